// Mesh topology on the GPU: first-seen global edge numbering, orientation signs,
// the CSR pattern of the assembled matrix and the deterministic element->nnz
// gather lists.
//
// Reference rule (src/mesh_gmsh.cpp:104-146, include/edgefem/mesh.hpp:56-60): scan
// tets in order with local pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3), then tris with
// (0,1),(1,2),(2,0); an edge is the sorted pair of gmsh IDs; a new edge gets the next
// index; the orientation sign is +1 iff conn[a] < conn[b].
//
// GPU formulation: every (element, local edge) "slot" s (tets: s = 6t + l, tris:
// s = 6T + 3f + l) is bucketed by the node holding the smaller ID; each bucket is
// sorted by (larger ID, slot), so every group of equal keys is one edge whose first
// slot is its first-seen occurrence.  Marking first slots and taking an exclusive
// scan over slot order gives exactly the first-seen numbering.  No global sort and no
// order-dependent atomics: bucket fill order is arbitrary, the per-bucket sort
// removes it.
#include "efm_internal.h"

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRow = 192;  // max distinct columns in one matrix row

__device__ inline void slot_nodes(int64_t s, int64_t T, const int32_t *tet_nodes,
                                  const int32_t *tri_nodes, int32_t &u, int32_t &v) {
  if (s < 6 * T) {
    int64_t t = s / 6;
    int l = (int)(s - t * 6);
    u = tet_nodes[t * 4 + tet_pair_a(l)];
    v = tet_nodes[t * 4 + tet_pair_b(l)];
  } else {
    int64_t q = s - 6 * T;
    int64_t f = q / 3;
    int l = (int)(q - f * 3);
    u = tri_nodes[f * 3 + tri_pair_a(l)];
    v = tri_nodes[f * 3 + tri_pair_b(l)];
  }
}

// Every node index must address node_ids; a bad index would make the later passes
// read out of bounds, so it is rejected before any of them runs.
__global__ void k_validate(int64_t count, const int32_t *__restrict__ nodes, int64_t N,
                           int32_t *__restrict__ err) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count)
    return;
  int32_t u = nodes[k];
  if (u < 0 || (int64_t)u >= N)
    atomicOr(err, 2);
}

__global__ void k_slot_keys(int64_t S, int64_t T, const int32_t *__restrict__ tet_nodes,
                            const int32_t *__restrict__ tri_nodes,
                            const int64_t *__restrict__ node_ids, int32_t *__restrict__ lo,
                            int64_t *__restrict__ hk, int8_t *__restrict__ tet_sign,
                            int8_t *__restrict__ tri_sign, int32_t *__restrict__ cnt) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S)
    return;
  int32_t u, v;
  slot_nodes(s, T, tet_nodes, tri_nodes, u, v);
  int64_t a = node_ids[u], b = node_ids[v];
  int8_t sign = (a < b) ? 1 : -1;
  // The key of mesh.hpp:56-60 is (min << 32) ^ max of the sorted ID pair; grouping by
  // (min node, max ID) is the same equivalence for IDs below 2^32.
  int32_t lnode = (a < b) ? u : v;
  int64_t hi = (a < b) ? b : a;
  if (a == b) {
    lnode = u;
    hi = a;
  }
  lo[s] = lnode;
  hk[s] = hi;
  if (s < 6 * T)
    tet_sign[s] = sign;
  else
    tri_sign[s - 6 * T] = sign;
  atomicAdd(&cnt[lnode], 1);
}

__global__ void k_bucket_fill(int64_t S, const int32_t *__restrict__ lo,
                              const int64_t *__restrict__ hk, const int32_t *__restrict__ bptr,
                              int32_t *__restrict__ cur, int32_t *__restrict__ bslot,
                              int64_t *__restrict__ bkey) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S)
    return;
  int32_t u = lo[s];
  int32_t pos = bptr[u] + atomicAdd(&cur[u], 1);
  bslot[pos] = (int32_t)s;
  bkey[pos] = hk[s];
}

// Sort each bucket by (key, slot) and mark the first slot of every key group.
__global__ void k_bucket_sort_mark(int64_t N, const int32_t *__restrict__ bptr,
                                   int32_t *__restrict__ bslot, int64_t *__restrict__ bkey,
                                   int32_t *__restrict__ is_first) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N)
    return;
  int32_t b0 = bptr[u], b1 = bptr[u + 1];
  for (int32_t i = b0 + 1; i < b1; ++i) {
    int64_t k = bkey[i];
    int32_t s = bslot[i];
    int32_t j = i - 1;
    while (j >= b0 && (bkey[j] > k || (bkey[j] == k && bslot[j] > s))) {
      bkey[j + 1] = bkey[j];
      bslot[j + 1] = bslot[j];
      --j;
    }
    bkey[j + 1] = k;
    bslot[j + 1] = s;
  }
  for (int32_t i = b0; i < b1; ++i)
    if (i == b0 || bkey[i] != bkey[i - 1])
      is_first[bslot[i]] = 1;
}

__global__ void k_bucket_assign(int64_t N, int64_t T, const int32_t *__restrict__ bptr,
                                const int32_t *__restrict__ bslot,
                                const int64_t *__restrict__ bkey, const int32_t *__restrict__ eid,
                                const int64_t *__restrict__ node_ids,
                                const int32_t *__restrict__ tet_nodes,
                                const int32_t *__restrict__ tri_nodes, int32_t *__restrict__ tet_edges,
                                int32_t *__restrict__ tri_edges, int64_t *__restrict__ edges,
                                int32_t *__restrict__ edge_nidx, int32_t *__restrict__ ntet,
                                int32_t *__restrict__ ntri) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N)
    return;
  int32_t b0 = bptr[u], b1 = bptr[u + 1];
  int32_t g0 = b0;
  while (g0 < b1) {
    int32_t g1 = g0 + 1;
    while (g1 < b1 && bkey[g1] == bkey[g0])
      ++g1;
    int32_t e = eid[bslot[g0]];
    int32_t nt = 0, nf = 0;
    for (int32_t i = g0; i < g1; ++i) {
      int64_t s = bslot[i];
      if (s < 6 * T) {
        tet_edges[s] = e;
        ++nt;
      } else {
        tri_edges[s - 6 * T] = e;
        ++nf;
      }
    }
    int64_t lid = node_ids[u];
    int64_t hid = bkey[g0];
    edges[2 * (int64_t)e] = lid < hid ? lid : hid;
    edges[2 * (int64_t)e + 1] = lid < hid ? hid : lid;
    int32_t su, sv;
    slot_nodes(bslot[g0], T, tet_nodes, tri_nodes, su, sv);
    edge_nidx[2 * (int64_t)e] = (int32_t)u;
    edge_nidx[2 * (int64_t)e + 1] = (su == (int32_t)u) ? sv : su;
    ntet[e] = nt;
    ntri[e] = nf;
    g0 = g1;
  }
}

__global__ void k_bucket_fill_incidence(int64_t N, int64_t T, const int32_t *__restrict__ bptr,
                                        const int32_t *__restrict__ bslot,
                                        const int64_t *__restrict__ bkey,
                                        const int32_t *__restrict__ eid,
                                        const int32_t *__restrict__ e2t_ptr,
                                        const int32_t *__restrict__ e2f_ptr,
                                        int32_t *__restrict__ e2t, int32_t *__restrict__ e2f) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N)
    return;
  int32_t b0 = bptr[u], b1 = bptr[u + 1];
  int32_t g0 = b0;
  while (g0 < b1) {
    int32_t g1 = g0 + 1;
    while (g1 < b1 && bkey[g1] == bkey[g0])
      ++g1;
    int32_t e = eid[bslot[g0]];
    int32_t kt = e2t_ptr[e], kf = e2f_ptr[e];
    for (int32_t i = g0; i < g1; ++i) {  // slots ascending within a group
      int64_t s = bslot[i];
      if (s < 6 * T)
        e2t[kt++] = (int32_t)s;
      else
        e2f[kf++] = (int32_t)(s - 6 * T);
    }
    g0 = g1;
  }
}

// Node -> incident edges (each edge listed at both ends; degenerate edges once).
__global__ void k_n2e_count(int64_t n, const int32_t *__restrict__ edge_nidx,
                            int32_t *__restrict__ cnt) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  int32_t a = edge_nidx[2 * e], b = edge_nidx[2 * e + 1];
  atomicAdd(&cnt[a], 1);
  if (b != a)
    atomicAdd(&cnt[b], 1);
}

__global__ void k_n2e_fill(int64_t n, const int32_t *__restrict__ edge_nidx,
                           const int32_t *__restrict__ ptr, int32_t *__restrict__ cur,
                           int32_t *__restrict__ n2e) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  int32_t a = edge_nidx[2 * e], b = edge_nidx[2 * e + 1];
  n2e[ptr[a] + atomicAdd(&cur[a], 1)] = (int32_t)e;
  if (b != a)
    n2e[ptr[b] + atomicAdd(&cur[b], 1)] = (int32_t)e;
}

// Sort every node's list ascending (removes the fill-order nondeterminism).
__global__ void k_n2e_sort(int64_t N, const int32_t *__restrict__ ptr, int32_t *__restrict__ n2e) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= N)
    return;
  int32_t b0 = ptr[u], b1 = ptr[u + 1];
  for (int32_t i = b0 + 1; i < b1; ++i) {
    int32_t v = n2e[i];
    int32_t j = i - 1;
    while (j >= b0 && n2e[j] > v) {
      n2e[j + 1] = n2e[j];
      --j;
    }
    n2e[j + 1] = v;
  }
}

// Sorted distinct columns of row e: all edges of the tets that contain edge e.
__device__ inline int row_columns(int32_t e, const int32_t *__restrict__ e2t_ptr,
                                  const int32_t *__restrict__ e2t,
                                  const int32_t *__restrict__ tet_edges, int32_t *list) {
  int len = 0;
  for (int32_t k = e2t_ptr[e]; k < e2t_ptr[e + 1]; ++k) {
    int32_t t = e2t[k] / 6;
    for (int j = 0; j < 6; ++j) {
      int32_t c = tet_edges[(int64_t)t * 6 + j];
      int p = 0;
      while (p < len && list[p] < c)
        ++p;
      if (p < len && list[p] == c)
        continue;
      if (len >= kMaxRow)
        return -1;
      for (int q = len; q > p; --q)
        list[q] = list[q - 1];
      list[p] = c;
      ++len;
    }
  }
  return len;
}

__global__ void k_row_count(int64_t n, const int32_t *__restrict__ e2t_ptr,
                            const int32_t *__restrict__ e2t, const int32_t *__restrict__ tet_edges,
                            int32_t *__restrict__ rowlen, int32_t *__restrict__ err) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  int32_t list[kMaxRow];
  int len = row_columns((int32_t)e, e2t_ptr, e2t, tet_edges, list);
  if (len < 0) {
    atomicOr(err, 1);
    len = 0;
  }
  if (len == 0) {
    // An edge owned by no tet (a stray surface edge) still gets its diagonal, as the
    // reference's Dirichlet pass inserts one with coeffRef.
    len = 1;
  }
  rowlen[e] = len;
}

__global__ void k_row_fill(int64_t n, const int32_t *__restrict__ e2t_ptr,
                           const int32_t *__restrict__ e2t, const int32_t *__restrict__ tet_edges,
                           const int32_t *__restrict__ row_ptr, int32_t *__restrict__ col_idx,
                           int32_t *__restrict__ row_idx) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  int32_t list[kMaxRow];
  int len = row_columns((int32_t)e, e2t_ptr, e2t, tet_edges, list);
  int32_t base = row_ptr[e];
  if (len <= 0) {
    col_idx[base] = (int32_t)e;
    row_idx[base] = (int32_t)e;
    return;
  }
  for (int k = 0; k < len; ++k) {
    col_idx[base + k] = list[k];
    row_idx[base + k] = (int32_t)e;
  }
}

__device__ inline int32_t find_col(const int32_t *__restrict__ col_idx, int32_t lo, int32_t hi,
                                   int32_t c) {
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (col_idx[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Contributor lists: for nnz (r, c) every (t, i, j) with tet_edges[t][i] = r and
// tet_edges[t][j] = c, in ascending t -- the triplet insertion order that
// setFromTriplets sums in (assemble_maxwell.cpp:185-205).  Each thread owns one row,
// so the counts and cursors of its nnz are private.
__global__ void k_contrib(int64_t n, int pass, const int32_t *__restrict__ e2t_ptr,
                          const int32_t *__restrict__ e2t, const int32_t *__restrict__ tet_edges,
                          const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                          int32_t *__restrict__ ccnt, const int32_t *__restrict__ cptr,
                          int32_t *__restrict__ contrib) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  int32_t r0 = row_ptr[e], r1 = row_ptr[e + 1];
  for (int32_t k = e2t_ptr[e]; k < e2t_ptr[e + 1]; ++k) {
    int32_t s = e2t[k];
    int32_t t = s / 6;
    int32_t i = s - t * 6;
    for (int j = 0; j < 6; ++j) {
      int32_t c = tet_edges[(int64_t)t * 6 + j];
      int32_t pos = find_col(col_idx, r0, r1, c);
      if (pos >= r1 || col_idx[pos] != c)
        continue;  // cannot happen: c is in row e by construction
      if (pass == 0) {
        ccnt[pos] += 1;
      } else {
        int32_t at = cptr[pos] + ccnt[pos];
        ccnt[pos] += 1;
        contrib[at] = t * 36 + i * 6 + j;
      }
    }
  }
}

} // namespace

extern "C" int efm_topology_create(int64_t n_nodes, const int64_t *node_ids, int64_t n_tets,
                                   const int32_t *tet_nodes, int64_t n_tris,
                                   const int32_t *tri_nodes, void *stream, efm_topology **out) {
  EFM_REQUIRE(out != nullptr, "efm_topology_create: out is NULL");
  *out = nullptr;
  EFM_REQUIRE(n_nodes > 0 && node_ids != nullptr, "efm_topology_create: no nodes");
  EFM_REQUIRE(n_tets >= 0 && n_tris >= 0, "efm_topology_create: negative element count");
  EFM_REQUIRE(n_tets == 0 || tet_nodes != nullptr, "efm_topology_create: tet_nodes is NULL");
  EFM_REQUIRE(n_tris == 0 || tri_nodes != nullptr, "efm_topology_create: tri_nodes is NULL");
  EFM_REQUIRE(n_tets * 36 < (int64_t)INT32_MAX, "efm_topology_create: too many tets for int32 maps");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = n_nodes, T = n_tets, F = n_tris;
  const int64_t S = 6 * T + 3 * F;
  EFM_REQUIRE(S > 0, "efm_topology_create: mesh has no elements");

  auto *topo = new efm_topology();
  std::unique_ptr<efm_topology> guard(topo);
  topo->N = N;
  topo->T = T;
  topo->F = F;
  int rc;
#define ALLOC(buf, bytes)                                                      \
  if ((rc = (buf).alloc(bytes)) != EFM_OK)                                     \
    return rc;
  ALLOC(topo->node_ids, N * 8);
  ALLOC(topo->tet_nodes, T * 16);
  ALLOC(topo->tri_nodes, F * 12);
  EFM_HIP_CHECK(hipMemcpyAsync(topo->node_ids.p, node_ids, N * 8, hipMemcpyDeviceToDevice, st));
  if (T)
    EFM_HIP_CHECK(hipMemcpyAsync(topo->tet_nodes.p, tet_nodes, T * 16, hipMemcpyDeviceToDevice, st));
  if (F)
    EFM_HIP_CHECK(hipMemcpyAsync(topo->tri_nodes.p, tri_nodes, F * 12, hipMemcpyDeviceToDevice, st));
  ALLOC(topo->tet_edges, T * 24 + 4);
  ALLOC(topo->tet_sign, T * 6 + 4);
  ALLOC(topo->tri_edges, F * 12 + 4);
  ALLOC(topo->tri_sign, F * 3 + 4);

  {
    DevBuf verr;
    ALLOC(verr, 4);
    EFM_HIP_CHECK(hipMemsetAsync(verr.p, 0, 4, st));
    if (T)
      hipLaunchKernelGGL(k_validate, dim3(grid_for(T * 4, kBlock)), dim3(kBlock), 0, st, T * 4,
                         topo->tet_nodes.as<int32_t>(), N, verr.as<int32_t>());
    if (F)
      hipLaunchKernelGGL(k_validate, dim3(grid_for(F * 3, kBlock)), dim3(kBlock), 0, st, F * 3,
                         topo->tri_nodes.as<int32_t>(), N, verr.as<int32_t>());
    EFM_LAUNCH_CHECK();
    int32_t h = 0;
    EFM_HIP_CHECK(hipMemcpyAsync(&h, verr.p, 4, hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    EFM_REQUIRE(h == 0, "efm_topology_create: element node index out of range");
  }
  DevBuf lo, hk, cnt, bptr, bslot, bkey, is_first, eid, work;
  ALLOC(lo, S * 4);
  ALLOC(hk, S * 8);
  ALLOC(cnt, N * 4);
  ALLOC(bptr, (N + 1) * 4);
  ALLOC(bslot, S * 4);
  ALLOC(bkey, S * 8);
  ALLOC(is_first, S * 4);
  ALLOC(eid, (S + 1) * 4);
  ALLOC(work, scan_work_bytes(S > N ? S : N) + scan_work_bytes(36 * T + 1));
  EFM_HIP_CHECK(hipMemsetAsync(cnt.p, 0, N * 4, st));
  EFM_HIP_CHECK(hipMemsetAsync(is_first.p, 0, S * 4, st));

  const int32_t *tn = topo->tet_nodes.as<int32_t>();
  const int32_t *fn = topo->tri_nodes.as<int32_t>();
  const int64_t *ids = topo->node_ids.as<int64_t>();
  hipLaunchKernelGGL(k_slot_keys, dim3(grid_for(S, kBlock)), dim3(kBlock), 0, st, S, T, tn, fn, ids,
                     lo.as<int32_t>(), hk.as<int64_t>(), topo->tet_sign.as<int8_t>(),
                     topo->tri_sign.as<int8_t>(), cnt.as<int32_t>());
  EFM_LAUNCH_CHECK();
  int64_t total = 0;
  if ((rc = exclusive_scan_i32(cnt.as<int32_t>(), N, bptr.as<int32_t>(), work.p, st, &total)) != EFM_OK)
    return rc;
  EFM_HIP_CHECK(hipMemsetAsync(cnt.p, 0, N * 4, st));
  hipLaunchKernelGGL(k_bucket_fill, dim3(grid_for(S, kBlock)), dim3(kBlock), 0, st, S,
                     lo.as<int32_t>(), hk.as<int64_t>(), bptr.as<int32_t>(), cnt.as<int32_t>(),
                     bslot.as<int32_t>(), bkey.as<int64_t>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bucket_sort_mark, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, N,
                     bptr.as<int32_t>(), bslot.as<int32_t>(), bkey.as<int64_t>(),
                     is_first.as<int32_t>());
  EFM_LAUNCH_CHECK();
  int64_t n_edges = 0;
  if ((rc = exclusive_scan_i32(is_first.as<int32_t>(), S, eid.as<int32_t>(), work.p, st, &n_edges)) != EFM_OK)
    return rc;
  topo->n_edges = n_edges;
  const int64_t n = n_edges;

  ALLOC(topo->edges, n * 16);
  ALLOC(topo->edge_nidx, n * 8);
  DevBuf ntet, ntri;
  ALLOC(ntet, n * 4);
  ALLOC(ntri, n * 4);
  ALLOC(topo->e2t_ptr, (n + 1) * 4);
  ALLOC(topo->e2f_ptr, (n + 1) * 4);
  ALLOC(topo->e2t, T * 24 + 4);
  ALLOC(topo->e2f, F * 12 + 4);
  hipLaunchKernelGGL(k_bucket_assign, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, N, T,
                     bptr.as<int32_t>(), bslot.as<int32_t>(), bkey.as<int64_t>(), eid.as<int32_t>(),
                     ids, tn, fn, topo->tet_edges.as<int32_t>(), topo->tri_edges.as<int32_t>(),
                     topo->edges.as<int64_t>(), topo->edge_nidx.as<int32_t>(), ntet.as<int32_t>(),
                     ntri.as<int32_t>());
  EFM_LAUNCH_CHECK();
  if ((rc = exclusive_scan_i32(ntet.as<int32_t>(), n, topo->e2t_ptr.as<int32_t>(), work.p, st, nullptr)) != EFM_OK)
    return rc;
  if ((rc = exclusive_scan_i32(ntri.as<int32_t>(), n, topo->e2f_ptr.as<int32_t>(), work.p, st, nullptr)) != EFM_OK)
    return rc;
  hipLaunchKernelGGL(k_bucket_fill_incidence, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, N, T,
                     bptr.as<int32_t>(), bslot.as<int32_t>(), bkey.as<int64_t>(), eid.as<int32_t>(),
                     topo->e2t_ptr.as<int32_t>(), topo->e2f_ptr.as<int32_t>(),
                     topo->e2t.as<int32_t>(), topo->e2f.as<int32_t>());
  EFM_LAUNCH_CHECK();
  // release the bucket temporaries before the pattern passes
  lo.alloc(0);
  hk.alloc(0);
  bslot.alloc(0);
  bkey.alloc(0);
  is_first.alloc(0);
  eid.alloc(0);

  // node -> edge incidence
  {
    DevBuf ncnt;
    ALLOC(ncnt, N * 4);
    ALLOC(topo->n2e_ptr, (N + 1) * 4);
    ALLOC(topo->n2e, n * 8 + 4);
    EFM_HIP_CHECK(hipMemsetAsync(ncnt.p, 0, N * 4, st));
    hipLaunchKernelGGL(k_n2e_count, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n,
                       topo->edge_nidx.as<int32_t>(), ncnt.as<int32_t>());
    EFM_LAUNCH_CHECK();
    if ((rc = exclusive_scan_i32(ncnt.as<int32_t>(), N, topo->n2e_ptr.as<int32_t>(), work.p, st, nullptr)) != EFM_OK)
      return rc;
    EFM_HIP_CHECK(hipMemsetAsync(ncnt.p, 0, N * 4, st));
    hipLaunchKernelGGL(k_n2e_fill, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n,
                       topo->edge_nidx.as<int32_t>(), topo->n2e_ptr.as<int32_t>(), ncnt.as<int32_t>(),
                       topo->n2e.as<int32_t>());
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_n2e_sort, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, N,
                       topo->n2e_ptr.as<int32_t>(), topo->n2e.as<int32_t>());
    EFM_LAUNCH_CHECK();
    EFM_HIP_CHECK(hipStreamSynchronize(st));
  }

  // CSR pattern
  DevBuf rowlen, err;
  ALLOC(rowlen, n * 4);
  ALLOC(err, 4);
  EFM_HIP_CHECK(hipMemsetAsync(err.p, 0, 4, st));
  hipLaunchKernelGGL(k_row_count, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n,
                     topo->e2t_ptr.as<int32_t>(), topo->e2t.as<int32_t>(),
                     topo->tet_edges.as<int32_t>(), rowlen.as<int32_t>(), err.as<int32_t>());
  EFM_LAUNCH_CHECK();
  ALLOC(topo->row_ptr, (n + 1) * 4);
  int64_t nnz = 0;
  if ((rc = exclusive_scan_i32(rowlen.as<int32_t>(), n, topo->row_ptr.as<int32_t>(), work.p, st, &nnz)) != EFM_OK)
    return rc;
  int32_t h_err = 0;
  EFM_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, 4, hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  if (h_err) {
    set_error("efm_topology_create: a matrix row has more than 192 distinct columns");
    return EFM_ERR_MESH;
  }
  topo->nnz = nnz;
  ALLOC(topo->col_idx, nnz * 4);
  ALLOC(topo->row_idx, nnz * 4);
  hipLaunchKernelGGL(k_row_fill, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n,
                     topo->e2t_ptr.as<int32_t>(), topo->e2t.as<int32_t>(),
                     topo->tet_edges.as<int32_t>(), topo->row_ptr.as<int32_t>(),
                     topo->col_idx.as<int32_t>(), topo->row_idx.as<int32_t>());
  EFM_LAUNCH_CHECK();

  // contributor lists
  DevBuf ccnt;
  ALLOC(ccnt, nnz * 4);
  ALLOC(topo->cptr, (nnz + 1) * 4);
  ALLOC(topo->contrib, T * 36 * 4 + 4);
  EFM_HIP_CHECK(hipMemsetAsync(ccnt.p, 0, nnz * 4, st));
  hipLaunchKernelGGL(k_contrib, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, 0,
                     topo->e2t_ptr.as<int32_t>(), topo->e2t.as<int32_t>(),
                     topo->tet_edges.as<int32_t>(), topo->row_ptr.as<int32_t>(),
                     topo->col_idx.as<int32_t>(), ccnt.as<int32_t>(), nullptr, nullptr);
  EFM_LAUNCH_CHECK();
  int64_t ctot = 0;
  if ((rc = exclusive_scan_i32(ccnt.as<int32_t>(), nnz, topo->cptr.as<int32_t>(), work.p, st, &ctot)) != EFM_OK)
    return rc;
  if (ctot != 36 * T) {
    set_error("efm_topology_create: contributor count mismatch (internal error)");
    return EFM_ERR_MESH;
  }
  EFM_HIP_CHECK(hipMemsetAsync(ccnt.p, 0, nnz * 4, st));
  hipLaunchKernelGGL(k_contrib, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, n, 1,
                     topo->e2t_ptr.as<int32_t>(), topo->e2t.as<int32_t>(),
                     topo->tet_edges.as<int32_t>(), topo->row_ptr.as<int32_t>(),
                     topo->col_idx.as<int32_t>(), ccnt.as<int32_t>(), topo->cptr.as<int32_t>(),
                     topo->contrib.as<int32_t>());
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipStreamSynchronize(st));
#undef ALLOC
  *out = guard.release();
  return EFM_OK;
}

extern "C" void efm_topology_destroy(efm_topology *topo) { delete topo; }

extern "C" int efm_topology_export_nodes(const efm_topology *topo, int32_t *edge_nodes,
                                         int32_t *n2e_ptr, int32_t *n2e, void *stream) {
  EFM_REQUIRE(topo, "efm_topology_export_nodes: NULL topology");
  hipStream_t st = (hipStream_t)stream;
  if (edge_nodes)
    EFM_HIP_CHECK(hipMemcpyAsync(edge_nodes, topo->edge_nidx.p, topo->n_edges * 8,
                                 hipMemcpyDeviceToDevice, st));
  if (n2e_ptr)
    EFM_HIP_CHECK(hipMemcpyAsync(n2e_ptr, topo->n2e_ptr.p, (topo->N + 1) * 4,
                                 hipMemcpyDeviceToDevice, st));
  if (n2e)
    EFM_HIP_CHECK(hipMemcpyAsync(n2e, topo->n2e.p, topo->n_edges * 8, hipMemcpyDeviceToDevice, st));
  return EFM_OK;
}

extern "C" int efm_topology_get_sizes(const efm_topology *topo, efm_topology_sizes *h_out) {
  EFM_REQUIRE(topo && h_out, "efm_topology_get_sizes: NULL argument");
  h_out->n_nodes = topo->N;
  h_out->n_tets = topo->T;
  h_out->n_tris = topo->F;
  h_out->n_edges = topo->n_edges;
  h_out->nnz = topo->nnz;
  return EFM_OK;
}

extern "C" int efm_topology_export(const efm_topology *topo, int64_t *edges, int32_t *tet_edges,
                                   int8_t *tet_sign, int32_t *tri_edges, int8_t *tri_sign,
                                   int32_t *row_ptr, int32_t *col_idx, void *stream) {
  EFM_REQUIRE(topo, "efm_topology_export: NULL topology");
  hipStream_t st = (hipStream_t)stream;
  auto cp = [&](void *dst, const DevBuf &src, size_t bytes) -> hipError_t {
    if (!dst || bytes == 0)
      return hipSuccess;
    return hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToDevice, st);
  };
  EFM_HIP_CHECK(cp(edges, topo->edges, topo->n_edges * 16));
  EFM_HIP_CHECK(cp(tet_edges, topo->tet_edges, topo->T * 24));
  EFM_HIP_CHECK(cp(tet_sign, topo->tet_sign, topo->T * 6));
  EFM_HIP_CHECK(cp(tri_edges, topo->tri_edges, topo->F * 12));
  EFM_HIP_CHECK(cp(tri_sign, topo->tri_sign, topo->F * 3));
  EFM_HIP_CHECK(cp(row_ptr, topo->row_ptr, (topo->n_edges + 1) * 4));
  EFM_HIP_CHECK(cp(col_idx, topo->col_idx, topo->nnz * 4));
  return EFM_OK;
}
