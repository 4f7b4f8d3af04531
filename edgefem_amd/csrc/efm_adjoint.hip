// Adjoint gradient of S-parameters with respect to the per-tet permittivity.
//
// With A x = b, S = g^H x - delta and dA/d eps_e = -k0^2 M_e (the element mass scattered
// with the edge signs), dS_ja/d eps_e = k0^2 lam_j^T M_e x_a where lam_j solves
// A^T lam = conj(g_j).  For a real loss L, dL/d eps_e = sum Re(conj(dL/dS*) dS/d eps_e).
// The caller folds k0^2, conj(G) and the scaling between lam and the solved systems into
// one complex weight per (lam column, x column) term; this kernel evaluates
//   grad[d] = sum_t Re(w_t * lam_{l_t,e}^T M_e x_{r_t,e})
// per design tet e = design[d], recomputing M_e from the coordinates instead of storing
// element blocks.  No atomics: one thread per design tet.
#include "efm_element.h"
#include "efm_internal.h"

#include <algorithm>
#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxTerms = 2048;

__global__ void __launch_bounds__(kBlock)
    k_eps_gradient(int64_t D, int64_t T, const int32_t *__restrict__ design, const int32_t *__restrict__ tet_nodes,
                   const double *__restrict__ xyz, const int32_t *__restrict__ tet_edges,
                   const int8_t *__restrict__ tet_sign, int nl, const double2 *__restrict__ L, int nx,
                   const double2 *__restrict__ X, int n_terms, const int2 *__restrict__ terms,
                   const double2 *__restrict__ w, double *__restrict__ grad) {
  extern __shared__ unsigned char smem[];
  int2 *sterm = reinterpret_cast<int2 *>(smem);
  double2 *sw = reinterpret_cast<double2 *>(smem + ((n_terms * sizeof(int2) + 15) / 16) * 16);
  for (int i = threadIdx.x; i < n_terms; i += blockDim.x) {
    sterm[i] = terms[i];
    sw[i] = w[i];
  }
  __syncthreads();
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D)
    return;
  const int64_t t = design ? (int64_t)design[d] : d;
  if (t < 0 || t >= T) {  // invalid design entry: NaN, no out-of-range access
    grad[d] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }
  const int4 nd = *reinterpret_cast<const int4 *>(tet_nodes + t * 4);
  V3 v[4] = {vload(xyz, nd.x), vload(xyz, nd.y), vload(xyz, nd.z), vload(xyz, nd.w)};
  V3 g[4];
  double V;
  gradients_and_volume(v, g, V);
  double G[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      G[a][b] = vdot(g[a], g[b]);
  int32_t e[6];
  double s[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    e[k] = tet_edges[t * 6 + k];
    s[k] = (double)tet_sign[t * 6 + k];
  }
  double Me[21];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j)
      Me[sym6(i, j)] = whitney_mass_entry(G, V, i, j) * s[i] * s[j];
  double acc = 0.0;
  for (int q = 0; q < n_terms; ++q) {
    const int2 lr = sterm[q];
    double2 xl[6], xr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      xl[k] = L[(int64_t)e[k] * nl + lr.x];
      xr[k] = X[(int64_t)e[k] * nx + lr.y];
    }
    double2 sval = make_double2(0, 0);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double2 u = make_double2(0, 0);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double m = Me[sym6(i, j)];
        u.x += m * xr[j].x;
        u.y += m * xr[j].y;
      }
      sval = cfma(xl[i], u, sval);
    }
    const double2 wq = sw[q];
    acc += wq.x * sval.x - wq.y * sval.y;
  }
  grad[d] = acc;
}

} // namespace

extern "C" int efm_eps_gradient(const efm_topology *topo, const double *xyz, int64_t n_design,
                                const int32_t *design, int nl, const double *Lam, int nx, const double *X,
                                int n_terms, const int32_t *h_term_l, const int32_t *h_term_r,
                                const double *h_w, double *grad, void *stream) {
  EFM_REQUIRE(topo && xyz && Lam && X && grad, "efm_eps_gradient: NULL argument");
  EFM_REQUIRE(n_design >= 0 && (design || n_design == topo->T), "efm_eps_gradient: design list");
  EFM_REQUIRE(nl >= 1 && nx >= 1, "efm_eps_gradient: column counts must be positive");
  EFM_REQUIRE(n_terms >= 0 && n_terms <= kMaxTerms, "efm_eps_gradient: 0..2048 terms");
  EFM_REQUIRE(n_terms == 0 || (h_term_l && h_term_r && h_w), "efm_eps_gradient: terms missing");
  for (int q = 0; q < n_terms; ++q)
    EFM_REQUIRE(h_term_l[q] >= 0 && h_term_l[q] < nl && h_term_r[q] >= 0 && h_term_r[q] < nx,
                "efm_eps_gradient: term column out of range");
  if (n_design == 0)
    return EFM_OK;
  hipStream_t st = (hipStream_t)stream;
  DevBuf terms, wbuf;
  int rc;
  if ((rc = terms.alloc(sizeof(int2) * std::max(1, n_terms))) || (rc = wbuf.alloc(sizeof(double2) * std::max(1, n_terms))))
    return rc;
  std::vector<int2> ht(n_terms);
  for (int q = 0; q < n_terms; ++q)
    ht[q] = make_int2(h_term_l[q], h_term_r[q]);
  if (n_terms) {
    EFM_HIP_CHECK(hipMemcpyAsync(terms.p, ht.data(), sizeof(int2) * n_terms, hipMemcpyHostToDevice, st));
    EFM_HIP_CHECK(hipMemcpyAsync(wbuf.p, h_w, sizeof(double2) * n_terms, hipMemcpyHostToDevice, st));
  }
  const size_t shmem = ((n_terms * sizeof(int2) + 15) / 16) * 16 + n_terms * sizeof(double2);
  hipLaunchKernelGGL(k_eps_gradient, dim3(grid_for(n_design, kBlock)), dim3(kBlock), shmem, st, n_design, topo->T, design,
                     topo->tet_nodes.as<int32_t>(), xyz, topo->tet_edges.as<int32_t>(), topo->tet_sign.as<int8_t>(),
                     nl, (const double2 *)Lam, nx, (const double2 *)X, n_terms, terms.as<int2>(),
                     wbuf.as<double2>(), grad);
  EFM_LAUNCH_CHECK();
  // the term buffers are freed on return: wait for the kernel that reads them
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  return EFM_OK;
}
