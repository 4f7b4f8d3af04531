// Element matrices and deterministic assembly.
//
// Pass 1 (one thread per tet): gradients of the barycentric coordinates, the Whitney
// curl-curl block K_e and mass block M_e (upper triangle, 21 entries each), multiplied
// by the orientation signs s_i s_j, kept in registers and written once as (K, M)
// pairs to a T x 21 scratch array.
// Pass 2 (one thread per nnz): sums the contributor list of the nnz (ascending tet,
// the order setFromTriplets sums duplicates in), applies the per-tet material and
// the Dirichlet rows/columns in the same pass.  No atomics.
#include "efm_element.h"
#include "efm_internal.h"

#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;

// K_e (edge_basis.cpp:47-63) and M_e (edge_basis.cpp:65-86), signed, upper triangle.
__global__ void __launch_bounds__(kBlock)
    k_element_blocks(int64_t T, const int32_t *__restrict__ tet_nodes,
                     const double *__restrict__ xyz, const int8_t *__restrict__ tet_sign,
                     double2 *__restrict__ km) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T)
    return;
  int4 nd = *reinterpret_cast<const int4 *>(tet_nodes + t * 4);
  V3 v[4] = {vload(xyz, nd.x), vload(xyz, nd.y), vload(xyz, nd.z), vload(xyz, nd.w)};
  V3 g[4];
  double V;
  gradients_and_volume(v, g, V);
  double s[6];
#pragma unroll
  for (int l = 0; l < 6; ++l)
    s[l] = (double)tet_sign[t * 6 + l];
  V3 c[6];
#pragma unroll
  for (int l = 0; l < 6; ++l)
    c[l] = vscale(vcross(g[tet_pair_a(l)], g[tet_pair_b(l)]), 2.0);
  double G[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      G[a][b] = vdot(g[a], g[b]);
  double2 *out = km + t * 21;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int j = i; j < 6; ++j) {
      double kv = V * vdot(c[i], c[j]);
      double term = whitney_mass_entry(G, V, i, j);
      double sg = s[i] * s[j];
      out[sym6(i, j)] = make_double2(kv * sg, term * sg);
    }
  }
}

// Gather for the K/M split (sweep.cpp:90-172): K = sum (K_e/mu) s s, M = sum (eps M_e) s s;
// Dirichlet rows/cols zero, K diagonal 1, M diagonal 0.
__global__ void __launch_bounds__(kBlock)
    k_gather_km(int64_t nnz, const int32_t *__restrict__ row_idx,
                const int32_t *__restrict__ col_idx, const int32_t *__restrict__ cptr,
                const int32_t *__restrict__ contrib, const double2 *__restrict__ km,
                const double2 *__restrict__ eps, const double2 *__restrict__ inv_mu,
                const uint8_t *__restrict__ dir, double2 *__restrict__ Kout,
                double2 *__restrict__ Mout) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz)
    return;
  int32_t r = row_idx[k], c = col_idx[k];
  if (dir && (dir[r] | dir[c])) {
    Kout[k] = make_double2(r == c ? 1.0 : 0.0, 0.0);
    Mout[k] = make_double2(0.0, 0.0);
    return;
  }
  double2 sk = make_double2(0.0, 0.0), sm = make_double2(0.0, 0.0);
  for (int32_t q = cptr[k]; q < cptr[k + 1]; ++q) {
    int32_t code = contrib[q];
    int32_t t = code / 36;
    int32_t ij = code - t * 36;
    int32_t i = ij / 6, j = ij - (ij / 6) * 6;
    double2 e = km[(int64_t)t * 21 + sym6(i, j)];
    double2 im = inv_mu ? inv_mu[t] : make_double2(1.0, 0.0);
    double2 ep = eps[t];
    sk = cadd(sk, make_double2(e.x * im.x, e.x * im.y));
    sm = cadd(sm, make_double2(ep.x * e.y, ep.y * e.y));
  }
  Kout[k] = sk;
  Mout[k] = sm;
}

// Gather for the direct form (assemble_maxwell.cpp:185-207, 244-265, 322-347):
// A = sum (K_e/mu - k0^2 eps M_e) s s  (+ j k0 on ABC diagonals), Dirichlet rows/cols
// zero with unit diagonal.
__global__ void __launch_bounds__(kBlock)
    k_gather_a(int64_t nnz, const int32_t *__restrict__ row_idx,
               const int32_t *__restrict__ col_idx, const int32_t *__restrict__ cptr,
               const int32_t *__restrict__ contrib, const double2 *__restrict__ km,
               const double2 *__restrict__ eps, const double2 *__restrict__ inv_mu,
               const uint8_t *__restrict__ dir, const uint8_t *__restrict__ abc, double k0,
               double2 *__restrict__ Aout) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz)
    return;
  int32_t r = row_idx[k], c = col_idx[k];
  if (dir && (dir[r] | dir[c])) {
    Aout[k] = make_double2(r == c ? 1.0 : 0.0, 0.0);
    return;
  }
  const double k0sq = k0 * k0;
  double2 sa = make_double2(0.0, 0.0);
  for (int32_t q = cptr[k]; q < cptr[k + 1]; ++q) {
    int32_t code = contrib[q];
    int32_t t = code / 36;
    int32_t ij = code - t * 36;
    int32_t i = ij / 6, j = ij - (ij / 6) * 6;
    double2 e = km[(int64_t)t * 21 + sym6(i, j)];
    double2 im = inv_mu ? inv_mu[t] : make_double2(1.0, 0.0);
    double2 ep = eps[t];
    double2 kk = make_double2(e.x * im.x, e.x * im.y);
    double2 mm = make_double2(k0sq * ep.x * e.y, k0sq * ep.y * e.y);
    sa = cadd(sa, csub(kk, mm));
  }
  if (abc && r == c && abc[r])
    sa.y += k0;
  Aout[k] = sa;
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

// 3x3 surface Whitney mass (edge_basis.cpp:88-130).
__device__ inline void tri_mass(const V3 v[3], double M[3][3]) {
  V3 nrm = vcross(vsub(v[1], v[0]), vsub(v[2], v[0]));
  double area2 = sqrt(vdot(nrm, nrm));
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      M[i][j] = 0.0;
  if (area2 < 1e-30)
    return;
  V3 nh = vscale(nrm, 1.0 / area2);
  double area = area2 / 2.0;
  V3 g[3];
  g[0] = vscale(vcross(nh, vsub(v[2], v[1])), 1.0 / area2);
  g[1] = vscale(vcross(nh, vsub(v[0], v[2])), 1.0 / area2);
  g[2] = vscale(vcross(nh, vsub(v[1], v[0])), 1.0 / area2);
  const double Id = area / 6.0, Io = area / 12.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int a = tri_pair_a(i), b = tri_pair_b(i);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      int c = tri_pair_a(j), d = tri_pair_b(j);
      double term = 0.0;
      term += vdot(g[b], g[d]) * (a == c ? Id : Io);
      term -= vdot(g[b], g[c]) * (a == d ? Id : Io);
      term -= vdot(g[a], g[d]) * (b == c ? Id : Io);
      term += vdot(g[a], g[c]) * (b == d ? Id : Io);
      M[i][j] = term;
    }
  }
}

// Port surface mass, one thread per global edge (row): walks the triangles containing
// the edge in mesh order and adds s_i s_j M_tri(i,j) for selected, non-Dirichlet pairs
// (wave_port.cpp:548-585).  Each thread writes only its own row.
__global__ void __launch_bounds__(kBlock)
    k_port_mass(int64_t n, const int32_t *__restrict__ e2f_ptr, const int32_t *__restrict__ e2f,
                const uint8_t *__restrict__ sel, const int32_t *__restrict__ tri_nodes,
                const int32_t *__restrict__ tri_edges, const int8_t *__restrict__ tri_sign,
                const double *__restrict__ xyz, const uint8_t *__restrict__ dir,
                const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col_idx,
                double *__restrict__ Ms) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  if (dir && dir[r])
    return;
  int32_t r0 = row_ptr[r], r1 = row_ptr[r + 1];
  for (int32_t k = e2f_ptr[r]; k < e2f_ptr[r + 1]; ++k) {
    int32_t code = e2f[k];
    int32_t f = code / 3;
    int32_t i = code - f * 3;
    if (!sel[f])
      continue;
    V3 v[3] = {vload(xyz, tri_nodes[(int64_t)f * 3 + 0]), vload(xyz, tri_nodes[(int64_t)f * 3 + 1]),
               vload(xyz, tri_nodes[(int64_t)f * 3 + 2])};
    double M[3][3];
    tri_mass(v, M);
    double si = (double)tri_sign[(int64_t)f * 3 + i];
    for (int j = 0; j < 3; ++j) {
      int32_t c = tri_edges[(int64_t)f * 3 + j];
      if (dir && dir[c])
        continue;
      int32_t pos = find_col(col_idx, r0, r1, c);
      if (pos >= r1 || col_idx[pos] != c)
        continue;  // pair not in the volume pattern (cannot happen for tet faces)
      double sj = (double)tri_sign[(int64_t)f * 3 + j];
      Ms[pos] += si * sj * M[i][j];
    }
  }
}

__global__ void k_mark(int64_t n_sel, const int32_t *__restrict__ idx, int64_t F,
                       uint8_t *__restrict__ mask) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_sel) {
    int32_t f = idx[k];
    if (f >= 0 && (int64_t)f < F)
      mask[f] = 1;
  }
}

struct PortCoefs {
  double2 c[8];
  const double *ms[8];
};

template <typename OutT>
__global__ void __launch_bounds__(kBlock)
    k_combine(int64_t nnz, const double2 *__restrict__ K, const double2 *__restrict__ M,
              double k0sq, int np, PortCoefs pc, OutT *__restrict__ A) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz)
    return;
  double2 kv = K[k], mv = M[k];
  double2 a = make_double2(kv.x - k0sq * mv.x, kv.y - k0sq * mv.y);
  for (int p = 0; p < np; ++p) {
    double m = pc.ms[p][k];
    a.x += pc.c[p].x * m;
    a.y += pc.c[p].y * m;
  }
  if constexpr (sizeof(OutT) == 16) {
    A[k] = a;
  } else {
    A[k] = make_float2((float)a.x, (float)a.y);
  }
}

int element_blocks(const efm_topology *topo, const double *xyz, hipStream_t st, DevBuf &km) {
  int rc = km.alloc((size_t)topo->T * 21 * sizeof(double2) + 16);
  if (rc != EFM_OK)
    return rc;
  hipLaunchKernelGGL(k_element_blocks, dim3(grid_for(topo->T, kBlock)), dim3(kBlock), 0, st,
                     topo->T, topo->tet_nodes.as<int32_t>(), xyz, topo->tet_sign.as<int8_t>(),
                     km.as<double2>());
  EFM_LAUNCH_CHECK();
  return EFM_OK;
}

} // namespace

extern "C" int efm_assemble_km(const efm_topology *topo, const double *xyz, const double *eps_tet,
                               const double *inv_mu_tet, const uint8_t *dirichlet, double *K_val,
                               double *M_val, void *stream) {
  EFM_REQUIRE(topo && xyz && eps_tet && K_val && M_val, "efm_assemble_km: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  DevBuf km;
  int rc = element_blocks(topo, xyz, st, km);
  if (rc != EFM_OK)
    return rc;
  hipLaunchKernelGGL(k_gather_km, dim3(grid_for(topo->nnz, kBlock)), dim3(kBlock), 0, st,
                     topo->nnz, topo->row_idx.as<int32_t>(), topo->col_idx.as<int32_t>(),
                     topo->cptr.as<int32_t>(), topo->contrib.as<int32_t>(), km.as<double2>(),
                     (const double2 *)eps_tet, (const double2 *)inv_mu_tet, dirichlet,
                     (double2 *)K_val, (double2 *)M_val);
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipStreamSynchronize(st));  // scratch is freed on return
  return EFM_OK;
}

extern "C" int efm_assemble_maxwell(const efm_topology *topo, const double *xyz,
                                    const double *eps_tet, const double *inv_mu_tet,
                                    const uint8_t *dirichlet, const uint8_t *abc_mask,
                                    double omega, double *A_val, void *stream) {
  EFM_REQUIRE(topo && xyz && eps_tet && A_val, "efm_assemble_maxwell: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  DevBuf km;
  int rc = element_blocks(topo, xyz, st, km);
  if (rc != EFM_OK)
    return rc;
  const double k0 = omega / kC0;
  hipLaunchKernelGGL(k_gather_a, dim3(grid_for(topo->nnz, kBlock)), dim3(kBlock), 0, st, topo->nnz,
                     topo->row_idx.as<int32_t>(), topo->col_idx.as<int32_t>(),
                     topo->cptr.as<int32_t>(), topo->contrib.as<int32_t>(), km.as<double2>(),
                     (const double2 *)eps_tet, (const double2 *)inv_mu_tet, dirichlet, abc_mask,
                     k0, (double2 *)A_val);
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  return EFM_OK;
}

extern "C" int efm_port_surface_mass(const efm_topology *topo, const double *xyz, int64_t n_sel,
                                     const int32_t *tri_sel, const uint8_t *dirichlet,
                                     double *Ms_val, void *stream) {
  EFM_REQUIRE(topo && xyz && Ms_val, "efm_port_surface_mass: NULL argument");
  EFM_REQUIRE(n_sel >= 0 && (n_sel == 0 || tri_sel), "efm_port_surface_mass: bad selection");
  hipStream_t st = (hipStream_t)stream;
  EFM_HIP_CHECK(hipMemsetAsync(Ms_val, 0, topo->nnz * sizeof(double), st));
  if (n_sel == 0)
    return EFM_OK;
  DevBuf mask;
  int rc = mask.alloc(topo->F + 1);
  if (rc != EFM_OK)
    return rc;
  EFM_HIP_CHECK(hipMemsetAsync(mask.p, 0, topo->F + 1, st));
  hipLaunchKernelGGL(k_mark, dim3(grid_for(n_sel, kBlock)), dim3(kBlock), 0, st, n_sel, tri_sel,
                     topo->F, mask.as<uint8_t>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_port_mass, dim3(grid_for(topo->n_edges, kBlock)), dim3(kBlock), 0, st,
                     topo->n_edges, topo->e2f_ptr.as<int32_t>(), topo->e2f.as<int32_t>(),
                     mask.as<uint8_t>(), topo->tri_nodes.as<int32_t>(),
                     topo->tri_edges.as<int32_t>(), topo->tri_sign.as<int8_t>(), xyz, dirichlet,
                     topo->row_ptr.as<int32_t>(), topo->col_idx.as<int32_t>(), Ms_val);
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  return EFM_OK;
}

extern "C" int efm_combine(const efm_topology *topo, const double *K_val, const double *M_val,
                           double k0sq, int n_ports, const double *const *Ms_vals,
                           const double *h_coef, int out_type, void *A_out, void *stream) {
  EFM_REQUIRE(topo && K_val && M_val && A_out, "efm_combine: NULL argument");
  EFM_REQUIRE(n_ports >= 0 && n_ports <= 8, "efm_combine: at most 8 ports");
  EFM_REQUIRE(n_ports == 0 || (Ms_vals && h_coef), "efm_combine: port arrays missing");
  EFM_REQUIRE(out_type == EFM_C128 || out_type == EFM_C64, "efm_combine: bad out_type");
  hipStream_t st = (hipStream_t)stream;
  PortCoefs pc{};
  for (int p = 0; p < n_ports; ++p) {
    pc.c[p] = make_double2(h_coef[2 * p], h_coef[2 * p + 1]);
    pc.ms[p] = Ms_vals[p];
  }
  if (out_type == EFM_C128)
    hipLaunchKernelGGL((k_combine<double2>), dim3(grid_for(topo->nnz, kBlock)), dim3(kBlock), 0, st,
                       topo->nnz, (const double2 *)K_val, (const double2 *)M_val, k0sq, n_ports,
                       pc, (double2 *)A_out);
  else
    hipLaunchKernelGGL((k_combine<float2>), dim3(grid_for(topo->nnz, kBlock)), dim3(kBlock), 0, st,
                       topo->nnz, (const double2 *)K_val, (const double2 *)M_val, k0sq, n_ports,
                       pc, (float2 *)A_out);
  EFM_LAUNCH_CHECK();
  return EFM_OK;
}

namespace {
__global__ void k_add_diagonal(int64_t m, const int32_t *__restrict__ rows, const double2 *__restrict__ vals,
                               const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                               double2 *__restrict__ A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m)
    return;
  const int32_t r = rows[i];
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    if (col[k] == r)
      A[k] = cadd(A[k], vals[i]);
}
} // namespace

extern "C" int efm_add_diagonal(const efm_topology *topo, int64_t m, const int32_t *rows, const double *vals,
                                double *A_val, void *stream) {
  EFM_REQUIRE(topo && A_val, "efm_add_diagonal: NULL argument");
  EFM_REQUIRE(m >= 0 && (m == 0 || (rows && vals)), "efm_add_diagonal: rows/vals missing");
  if (m == 0)
    return EFM_OK;
  hipStream_t st = (hipStream_t)stream;
  std::vector<int32_t> h(m);
  EFM_HIP_CHECK(hipMemcpyAsync(h.data(), rows, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<uint8_t> seen(topo->n_edges, 0);
  for (int64_t i = 0; i < m; ++i) {
    EFM_REQUIRE(h[i] >= 0 && h[i] < topo->n_edges, "efm_add_diagonal: row out of range");
    EFM_REQUIRE(!seen[h[i]], "efm_add_diagonal: rows must be distinct");
    seen[h[i]] = 1;
  }
  hipLaunchKernelGGL(k_add_diagonal, dim3(grid_for(m, 256)), dim3(256), 0, st, m, rows, (const double2 *)vals,
                     topo->row_ptr.as<int32_t>(), topo->col_idx.as<int32_t>(), (double2 *)A_val);
  EFM_LAUNCH_CHECK();
  return EFM_OK;
}
