// Post-solve field evaluation: E at one point of every tet and the tet-constant curl E
// from the edge solution.  The reference evaluates one point at a time
// (evaluate_edge_field, src/edge_basis.cpp:162-192; barycentric coordinates 132-151) in
// a host loop over the tets of the VTK export (src/io/vtk_export.cpp:14-40, at the
// centroid) and of the Huygens surface (src/post/huygens_surface.cpp:95-125, E at the
// centroid and curl E from whitney_edge_curls, src/edge_basis.cpp:34-45).  Here one
// thread takes one tet: the barycentric gradients once, the six Whitney functions at the
// point and the six edge curls, then every system's six coefficients s_e x_e.
//   E    = sum_e s_e x_e (l_a grad l_b - l_b grad l_a)
//   curl = sum_e s_e x_e 2 grad l_a x grad l_b
// HBM-bound: per tet the node/edge indices and signs, the gathered coordinates and
// solution values (each read once under ideal caching) and R x 96 B of output.
#include "efm_element.h"
#include "efm_internal.h"

using namespace efm;

namespace {

constexpr int kMaxSys = 8;

// Threads per block for R systems: the block's outputs (R x 3 values per tet) are staged
// in LDS (24 KB at most) and written as contiguous runs, since each thread's own 48 R
// bytes would make every store instruction touch a cache line per lane.
template <int R> constexpr int fields_block() { return R <= 2 ? 256 : (R <= 4 ? 128 : 64); }

template <int R>
__global__ void __launch_bounds__(fields_block<R>())
    k_tet_fields(int64_t T, const int32_t *__restrict__ tet_nodes, const int32_t *__restrict__ tet_edges,
                 const int8_t *__restrict__ tet_sign, const double *__restrict__ xyz,
                 const double2 *__restrict__ x, const double *__restrict__ lam, double2 *__restrict__ E,
                 double2 *__restrict__ curl) {
  constexpr int NB = fields_block<R>();
  __shared__ double2 stage[NB * R * 3];
  // XCD-contiguous tet ranges (block b runs on XCD b % 8, gridDim.x a multiple of 8): the
  // tets of an XCD are neighbours, so the nodes and edges they share hit in its L2
  const int64_t per_xcd = gridDim.x >> 3;
  const int64_t t0 = ((int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3)) * NB;
  if (t0 >= T)
    return;
  const int64_t t = min<int64_t>(t0 + threadIdx.x, T - 1);  // threads past T redo the last tet
  const int nt = (int)min<int64_t>(NB, T - t0);
  const int4 nd = *reinterpret_cast<const int4 *>(tet_nodes + t * 4);
  const V3 v[4] = {vload(xyz, nd.x), vload(xyz, nd.y), vload(xyz, nd.z), vload(xyz, nd.w)};
  V3 g[4];
  double vol;
  gradients_and_volume(v, g, vol);
  double l[4] = {0.25, 0.25, 0.25, 0.25};
  if (lam) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      l[i] = lam[t * 4 + i];
  }
  double2 eacc[R][3], cacc[R][3];
#pragma unroll
  for (int s = 0; s < R; ++s)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      eacc[s][k] = make_double2(0, 0);
      cacc[s][k] = make_double2(0, 0);
    }
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int a = tet_pair_a(e), b = tet_pair_b(e);
    const V3 w = vsub(vscale(g[b], l[a]), vscale(g[a], l[b]));
    const V3 c = vscale(vcross(g[a], g[b]), 2.0);
    const int32_t edge = tet_edges[t * 6 + e];
    const double sg = (double)tet_sign[t * 6 + e];
    const double wk[3] = {w.x, w.y, w.z}, ck[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const double2 xv = x[(int64_t)edge * R + s];
      const double2 cf = make_double2(sg * xv.x, sg * xv.y);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        eacc[s][k].x += cf.x * wk[k];
        eacc[s][k].y += cf.y * wk[k];
        cacc[s][k].x += cf.x * ck[k];
        cacc[s][k].y += cf.y * ck[k];
      }
    }
  }
  // [T][R][3] outputs through LDS: each thread's R x 3 values, then the block's contiguous
  // run of nt x R x 3 values, once for E and once for curl
  auto write = [&](double2 (&acc)[R][3], double2 *out) {
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        stage[(threadIdx.x * R + s) * 3 + k] = acc[s][k];
    __syncthreads();
    for (int i = threadIdx.x; i < nt * R * 3; i += NB)
      out[t0 * R * 3 + i] = stage[i];
    __syncthreads();
  };
  write(eacc, E);
  if (curl)
    write(cacc, curl);
}

template <int R>
void launch(const efm_topology *topo, const double *xyz, const double *x, const double *lam, double *E,
            double *curl, hipStream_t st) {
  const unsigned nb = (grid_for(topo->T, fields_block<R>()) + 7u) & ~7u;  // a multiple of 8
  hipLaunchKernelGGL((k_tet_fields<R>), dim3(nb), dim3(fields_block<R>()), 0, st, topo->T,
                     topo->tet_nodes.as<int32_t>(), topo->tet_edges.as<int32_t>(), topo->tet_sign.as<int8_t>(),
                     xyz, (const double2 *)x, lam, (double2 *)E, (double2 *)curl);
}

} // namespace

extern "C" int efm_tet_fields(const efm_topology *topo, const double *xyz, int nrhs, const double *x,
                              const double *lam, double *E, double *curl, void *stream) {
  EFM_REQUIRE(topo && xyz && x && E, "efm_tet_fields: NULL argument");
  EFM_REQUIRE(nrhs >= 1 && nrhs <= kMaxSys && (nrhs <= 4 || nrhs == 8),
              "efm_tet_fields: nrhs must be 1, 2, 3, 4 or 8");
  if (topo->T == 0)
    return EFM_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (nrhs) {
  case 1: launch<1>(topo, xyz, x, lam, E, curl, st); break;
  case 2: launch<2>(topo, xyz, x, lam, E, curl, st); break;
  case 3: launch<3>(topo, xyz, x, lam, E, curl, st); break;
  case 4: launch<4>(topo, xyz, x, lam, E, curl, st); break;
  default: launch<8>(topo, xyz, x, lam, E, curl, st); break;
  }
  EFM_LAUNCH_CHECK();
  return EFM_OK;
}
