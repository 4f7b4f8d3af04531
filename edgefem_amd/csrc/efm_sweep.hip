// Batched frequency sweep: many (frequency, excitation) systems that share one sparsity
// pattern, solved together by preconditioned COCG with per-system convergence masks.
//
//   A_s = K - k0_s^2 M + sum_p c_{s,p} M_s,p        (c_{s,p} = j beta_p(f_s) scale)
//
// Replaces the per-frequency loop of RectWaveguideDesign.frequency_sweep
// (python/edgefem/designs/waveguide.py:394-434), i.e. calculate_sparams_eigenmode
// (src/assemble_maxwell.cpp:637-787) per frequency, and the K/M reuse idea of
// frequency_sweep (src/sweep.cpp:192-329).
//
// Layout: every vector is [row][system]; the 64 lanes of a wavefront are 64 systems,
// so a column gather of a wave is one contiguous 1 KB read, and the shared matrix
// entries of a row are uniform across the wave.  The iteration is the single-reduction
// (Chronopoulos-Gear) form of COCG, so the SpMV gathers only one vector.
// The preconditioner is the auxiliary-gradient one (edge Jacobi + nodal correction);
// its per-system diagonals are formed on the fly from frequency-independent parts:
//   D_s(e)  = K_ee - k0_s^2 M_ee + c_{s,port(e)} Ms_ee
//   Dn_s(n) = dK(n) - k0_s^2 dM(n) + sum_p c_{s,p} dMs_p(n),   d*(n) = diag(G^T * G)
// Partial sums are reduced per system (one block per system) in a fixed order, so the
// result of a system does not depend on which other systems share the launch.
#include "efm_internal.h"

#include <algorithm>
#include <type_traits>
#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kSys = 64;           // systems per wavefront (one lane each)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxPorts = 4;

struct SwState {
  double2 rho, alpha, beta;
  double2 pend;           // alpha of the last step whose x and r updates are still pending
  double bnorm2, rnorm2;  // right-hand side and recursive residual norms^2
  double bn_out, rn_out;  // reliable updates: |b|^2 and the last true residual |b - A x|^2
  double2 rho_prev;       // reliable updates: rho at the start of the last iteration
  double rmax2;           // reliable updates: max |r|^2 since the last replacement
  int active, iters, breakdown, total_iters;
  int out_active;         // mask of the replacement kernels (residual, accumulate)
  int want;               // reliable updates: a residual replacement is due
  int saved_active, pad;
};

__device__ inline double2 bil(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 port_coef(const double2 *__restrict__ coef, int s, int P, int port) {
  return (port >= 0) ? coef[s * P + port] : make_double2(0, 0);
}

// Per-block partial sums of V values per system: the kWaves waves of a block handle
// the same 64 systems on different rows; combine them in LDS, fixed order.
template <int V>
__device__ inline void write_sys_partials(double2 (&v)[V], double2 *lds, double2 *partial,
                                          int block_linear) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < V; ++i)
    lds[(w * V + i) * 64 + lane] = v[i];
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      double2 a = lds[i * 64 + lane];
      for (int ww = 1; ww < kWaves; ++ww)
        a = cadd(a, lds[(ww * V + i) * 64 + lane]);
      partial[((int64_t)block_linear * V + i) * 64 + lane] = a;
    }
  }
}

// Persistent XCD-aware walk over row chunks of kWaves rows: block bx runs on XCD bx % 8.
struct ChunkWalk {
  int64_t begin, end, stride;
  __device__ ChunkWalk(int64_t nchunks) {
    const int64_t span = (nchunks + 7) / 8;
    const int xcd = blockIdx.x & 7;
    begin = xcd * span + (blockIdx.x >> 3);
    end = min(nchunks, (int64_t)(xcd + 1) * span);
    stride = gridDim.x >> 3;
  }
};

// Vector storage: complex FP64 (double2) or complex FP32 (float2); arithmetic and
// reductions are FP64 either way.
__device__ inline double2 ldv(const double2 *p) { return *p; }
__device__ inline double2 ldv(const float2 *p) {
  const float2 v = *p;
  return make_double2((double)v.x, (double)v.y);
}
__device__ inline void stv(double2 *p, double2 v) { *p = v; }
__device__ inline void stv(float2 *p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

// Broadcast of lane l's value to the whole wavefront (l wave-uniform).
__device__ inline int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---------------------------------------------------------------------------
// Chronopoulos-Gear COCG (one reduction point per iteration; the SpMV reads only u):
//   r0 = b, u0 = P r0, w0 = A u0, g0 = u0^T r0, d0 = u0^T w0, a0 = g0/d0, b0 = 0
//   p_i = u_i + b_i p_{i-1};  s_i = w_i + b_i s_{i-1}     (s = A p)
//   x_{i+1} = x_i + a_i p_i;   r_{i+1} = r_i - a_i s_i
//   u_{i+1} = P r_{i+1};  w_{i+1} = A u_{i+1};  g = u^T r;  d = u^T w
//   b_{i+1} = g_{i+1}/g_i;  a_{i+1} = g_{i+1} / (d_{i+1} - b_{i+1} g_{i+1} / a_i)
// (bilinear forms: COCG for complex-symmetric A).  Kernels per iteration: aux_node
// (nodal part of P on r_{i+1}, formed on the fly), update (p, s, x, r, u and partials
// of g, |r|^2), spmv (w and partials of d), finalize (scalars, convergence).
// x and r are stored every other iteration only: after a step i whose stores are
// skipped, x_i and r_i in memory lag by pend = alpha_i, and the next step applies
//   x_{i+2} = x_i + pend p_i + alpha_{i+1} p_{i+1},   r_{i+2} = r_i - pend s_i - alpha_{i+1} s_{i+1}
// (p_i and s_i are always stored).  A system that stops with a pending step gets
// x += pend p once after the loop.
// ---------------------------------------------------------------------------
// x = p = s = 0; r = b (b == nullptr: r already holds the right-hand side)
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_init(int64_t n, int B, const double2 *__restrict__ b, VT *__restrict__ x, VT *__restrict__ r,
              VT *__restrict__ p, VT *__restrict__ sv) {
  const double2 z = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * B; i += (int64_t)gridDim.x * kBlock) {
    stv(x + i, z);
    if (b)
      stv(r + i, b[i]);
    stv(p + i, z);
    stv(sv + i, z);
  }
}

// y[n][s] = Dn_s(n)^{-1} (G^T v)[n][s], NV vectors gathered per edge:
//   NV = 1: v = r;  NV = 2: v = r - alpha w (standard COCG: the new residual, w = A p);
//   NV = 3: v = r - alpha w - (alpha beta + pend) s (Chronopoulos-Gear, r and s as stored:
//   this is r_{i+1}).
// One node per wavefront; the node's edge list and gradient signs are loaded across the
// lanes and broadcast, so the gathers of its edges are independent of each other.
template <typename VT, int NV>
__global__ void __launch_bounds__(kBlock)
    k_sw_aux_node(int64_t N, int B, int P, const int32_t *__restrict__ n2e_ptr,
                  const int32_t *__restrict__ n2e, const int8_t *__restrict__ gsign,
                  const double2 *__restrict__ dnK, const double2 *__restrict__ dnM,
                  const double *__restrict__ dnMs, const double *__restrict__ k0sq,
                  const double2 *__restrict__ coef, const VT *__restrict__ r,
                  const VT *__restrict__ w, const VT *__restrict__ sv,
                  VT *__restrict__ y, const SwState *__restrict__ st) {
  constexpr int U = 4;
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.y * kSys + lane;
  const int wv = threadIdx.x >> 6;
  const bool act = s < B && st[s].active != 0;
  if (__ballot(act) == 0)
    return;
  const double2 alpha = NV >= 2 ? st[s].alpha : make_double2(0, 0);
  // r_{i+1} = r - alpha w - (alpha beta + pend) s with r, s as stored
  const double2 abp = NV == 3 ? cadd(cmul(st[s].alpha, st[s].beta), st[s].pend) : make_double2(0, 0);
  const double k2 = k0sq[s];
  for (int64_t nd = (int64_t)blockIdx.x * kWaves + wv; nd < N; nd += (int64_t)gridDim.x * kWaves) {
    double2 acc = make_double2(0, 0);
    const int32_t beg = n2e_ptr[nd], end = n2e_ptr[nd + 1];
    for (int32_t base = beg; base < end; base += 64) {
      const int len = min(64, end - base);
      int ej = 0, gj = 0;
      if (lane < len) {
        ej = n2e[base + lane];
        gj = gsign[base + lane];
      }
      for (int j = 0; j < len; j += U) {
        // the U gathers of a step are unconditional (one group of loads): past the list's
        // end they repeat its last edge with weight zero
        double2 rv[U];
#pragma unroll
        for (int t = 0; t < U; ++t) {
          const int64_t i = (int64_t)bcast(ej, min(j + t, len - 1)) * B + s;
          rv[t] = ldv(r + i);
          if (NV >= 2)
            rv[t] = csub(rv[t], bil(alpha, ldv(w + i)));
          if (NV == 3)
            rv[t] = csub(rv[t], bil(abp, ldv(sv + i)));
        }
#pragma unroll
        for (int t = 0; t < U; ++t) {
          const double gs = j + t < len ? (double)bcast(gj, min(j + t, len - 1)) : 0.0;
          acc.x += gs * rv[t].x;
          acc.y += gs * rv[t].y;
        }
      }
    }
    if (act) {
      double2 d = make_double2(dnK[nd].x - k2 * dnM[nd].x, dnK[nd].y - k2 * dnM[nd].y);
      for (int p = 0; p < P; ++p) {
        const double m = dnMs[(int64_t)p * N + nd];
        const double2 c = coef[s * P + p];
        d.x += c.x * m;
        d.y += c.y * m;
      }
      stv(y + nd * B + s, (d.x == 0.0 && d.y == 0.0) ? make_double2(0, 0) : cdiv(acc, d));
    }
  }
}

// dninv[n][s] = Dn_s(n)^{-1} (0 where Dn_s(n) = 0) for the systems of one solve.
__global__ void __launch_bounds__(kBlock)
    k_sw_dninv(int64_t N, int B, int P, const double2 *__restrict__ dnK, const double2 *__restrict__ dnM,
               const double *__restrict__ dnMs, const double *__restrict__ k0sq,
               const double2 *__restrict__ coef, double2 *__restrict__ dninv) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N * B; i += (int64_t)gridDim.x * kBlock) {
    const int64_t nd = i / B;
    const int s = (int)(i - nd * B);
    const double k2 = k0sq[s];
    double2 d = make_double2(dnK[nd].x - k2 * dnM[nd].x, dnK[nd].y - k2 * dnM[nd].y);
    for (int p = 0; p < P; ++p) {
      const double m = dnMs[(int64_t)p * N + nd];
      const double2 c = coef[s * P + p];
      d.x += c.x * m;
      d.y += c.y * m;
    }
    dninv[i] = (d.x == 0.0 && d.y == 0.0) ? make_double2(0, 0) : cdiv(make_double2(1, 0), d);
  }
}

// The same map as k_sw_aux_node (NV gathered vectors), with the node walk of
// k_sw_spmv_pre.  Groups of 64 systems are walked in turn, the blocks of an XCD walk a
// contiguous node range (so an edge gathered by both of its end nodes is read from HBM
// once), each wave preloads the edge-list bounds of its next 64 nodes, and the gathers of
// a batch of NE edges (NV vectors each) are issued with no branch: lanes past the node's
// degree hold a valid edge with gradient sign 0.
template <typename VT, int NV>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NV > 1 ? 3 : 4, 8)))
    k_sw_aux_pre(int64_t N, int B, int groups, const int32_t *__restrict__ n2e_ptr,
                 const int32_t *__restrict__ n2e, const int8_t *__restrict__ gsign,
                 const double2 *__restrict__ dninv, const VT *__restrict__ r,
                 const VT *__restrict__ w, const VT *__restrict__ sv, VT *__restrict__ y,
                 const SwState *__restrict__ st) {
  constexpr int NE = NV == 3 ? 7 : NV == 2 ? 10 : 14;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t nchunks = (N + kWaves - 1) / kWaves;
  ChunkWalk cw(nchunks);
  const int64_t nr = (cw.begin < cw.end) ? (cw.end - cw.begin + cw.stride - 1) / cw.stride : 0;
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
    const int s = g * kSys + lane;
    const bool act = s < B && st[s].active != 0;
    if (__ballot(act) == 0)
      continue;
    const double2 alpha = NV >= 2 ? st[s].alpha : make_double2(0, 0);
    const double2 ab = NV == 3 ? cadd(cmul(st[s].alpha, st[s].beta), st[s].pend) : make_double2(0, 0);
    for (int64_t jb = 0; jb < nr; jb += 64) {
      const int cnt = (int)min<int64_t>(64, nr - jb);
      int lo = 0, hi = 0;
      if (lane < cnt) {
        const int64_t nd = (cw.begin + (jb + lane) * cw.stride) * kWaves + wv;
        if (nd < N) {
          lo = n2e_ptr[nd];
          hi = n2e_ptr[nd + 1];
        }
      }
      for (int j = 0; j < cnt; ++j) {
        const int64_t nd = (cw.begin + (jb + j) * cw.stride) * kWaves + wv;
        const int64_t ndc = min<int64_t>(nd, N - 1);
        const double2 dinv = dninv[ndc * B + s];
        const int32_t beg = bcast(lo, j), len = bcast(hi, j) - beg;
        double2 sr = make_double2(0, 0), sw = make_double2(0, 0), ss = make_double2(0, 0);
        for (int32_t b = 0; b < len; b += 64) {
          const int m = min(64, len - b);
          const bool in = lane < m;
          const int32_t ek = n2e[beg + b + (in ? lane : 0)];
          const int ej = ek;
          const int gj = in ? (int)gsign[beg + b + lane] : 0;
          for (int e0 = 0; e0 < m; e0 += NE) {
            double2 xr[NE], xw[NE], xs[NE];
#pragma unroll
            for (int t = 0; t < NE; ++t) {
              const int64_t i = (int64_t)bcast(ej, min(e0 + t, 63)) * B + s;
              xr[t] = ldv(r + i);
              if (NV >= 2)
                xw[t] = ldv(w + i);
              if (NV == 3)
                xs[t] = ldv(sv + i);
            }
#pragma unroll
            for (int t = 0; t < NE; ++t) {
              // lanes past the batch's last edge hold sign 0 (e0 + t < 64 keeps the lane valid)
              const double gs = (e0 + t < m) ? (double)bcast(gj, min(e0 + t, 63)) : 0.0;
              sr.x += gs * xr[t].x;
              sr.y += gs * xr[t].y;
              if (NV >= 2) {
                sw.x += gs * xw[t].x;
                sw.y += gs * xw[t].y;
              }
              if (NV == 3) {
                ss.x += gs * xs[t].x;
                ss.y += gs * xs[t].y;
              }
            }
          }
        }
        if (act && nd < N) {
          // sum G (r - alpha w - (alpha beta + pend) s), the terms NV asks for
          const double2 acc = NV == 3 ? csub(csub(sr, cmul(alpha, sw)), cmul(ab, ss))
                                      : NV == 2 ? csub(sr, cmul(alpha, sw)) : sr;
          stv(y + nd * B + s, cmul(dinv, acc));
        }
      }
    }
  }
}

__device__ inline double2 edge_diag_inv(int64_t e, int s, int P, double k2, const double2 *dK,
                                        const double2 *dM, const double *dMs,
                                        const int8_t *rport, const double2 *coef) {
  const double2 c = port_coef(coef, s, P, rport[e]);
  const double m = dMs[e];
  double2 d = make_double2(dK[e].x - k2 * dM[e].x + c.x * m, dK[e].y - k2 * dM[e].y + c.y * m);
  return (d.x == 0.0 && d.y == 0.0) ? make_double2(1, 0) : cdiv(make_double2(1, 0), d);
}

// INIT: u = P r; partials g = u^T r, |r|^2.
// else: p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s; u = P r; partials,
// with x and r read and stored only when APPLY (the pending step of the previous
// iteration, pend, is folded in first: x += pend p_old, r -= pend s_old).
// Persistent: one block slice (gridDim.y == 1) walks the groups of 64 systems in turn.
template <typename VT, bool INIT, bool APPLY>
__global__ void __launch_bounds__(kBlock)
    k_sw_update(int64_t n, int B, int groups, int P, const int2 *__restrict__ en_eff,
                const double2 *__restrict__ dK, const double2 *__restrict__ dM,
                const double *__restrict__ dMs, const int8_t *__restrict__ rport,
                const double *__restrict__ k0sq, const double2 *__restrict__ coef,
                const VT *__restrict__ y, const VT *__restrict__ w,
                VT *__restrict__ u, VT *__restrict__ p, VT *__restrict__ sv,
                VT *__restrict__ x, VT *__restrict__ r, double2 *__restrict__ partial,
                const SwState *__restrict__ st) {
  __shared__ double2 lds[kWaves * 2 * 64];
  const int wv = threadIdx.x >> 6;
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
  const int s = g * kSys + (threadIdx.x & 63);
  double2 v[2] = {make_double2(0, 0), make_double2(0, 0)};
  const bool act = s < B && st[s].active != 0;
  if (act) {
    const double k2 = k0sq[s];
    const double2 alpha = st[s].alpha, beta = st[s].beta, pend = st[s].pend;
    for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
      const int64_t i = e * B + s;
      double2 rv = ldv(r + i);
      if (!INIT) {
        const double2 po = ldv(p + i), so = ldv(sv + i);
        const double2 pv = cfma(beta, po, ldv(u + i));
        const double2 svv = cfma(beta, so, ldv(w + i));
        if (APPLY) {
          stv(x + i, cfma(alpha, pv, cfma(pend, po, ldv(x + i))));
          rv = csub(rv, bil(pend, so));
        }
        rv = csub(rv, bil(alpha, svv));
        stv(p + i, pv);
        stv(sv + i, svv);
        if (APPLY)
          stv(r + i, rv);
      }
      const double2 di = edge_diag_inv(e, s, P, k2, dK, dM, dMs, rport, coef);
      double2 z = bil(di, rv);
      const int2 ab = en_eff[e];
      {  // both node values loaded unconditionally (node 0 for an edge without nodes)
        const bool hn = ab.x >= 0;
        const double2 yd =
            csub(ldv(y + (int64_t)(hn ? ab.y : 0) * B + s), ldv(y + (int64_t)(hn ? ab.x : 0) * B + s));
        z = hn ? cadd(z, yd) : z;
      }
      stv(u + i, z);
      v[0] = cadd(v[0], bil(rv, z));
      v[1].x += cabs2(rv);
    }
  }
  write_sys_partials<2>(v, lds, partial, g * gridDim.x + blockIdx.x);
  __syncthreads();  // lds reuse by the next group
  }
}

// ---------------------------------------------------------------------------
// Standard COCG on the batch (the default iteration form): per iteration
//   p = u + beta p;  w = A p, partials p^T w;  alpha = rho / p^T w
//   y = Dn^{-1} G^T (r - alpha w);  x += alpha p;  r -= alpha w;  u = D^{-1} r + G y,
//   partials u^T r, |r|^2;  beta = u^T r / rho
// Two reduction points per iteration against one for Chronopoulos-Gear, whose s = w +
// beta s recurrence drifts from A p and costs iterations on these indefinite systems
// (394/462/584 against 394/490/623 on the WR-90 16x8x32 sweep at 8.2/10/12.4 GHz, 2,333
// against 3,659 on a permittivity-10 slab).
// ---------------------------------------------------------------------------
// p = u + beta p for the active systems
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_pform(int64_t n, int B, const VT *__restrict__ u, VT *__restrict__ p, const SwState *__restrict__ st) {
  const int s = blockIdx.y * kSys + (threadIdx.x & 63);
  const int wv = threadIdx.x >> 6;
  if (s >= B || st[s].active == 0)
    return;
  const double2 beta = st[s].beta;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
    const int64_t i = e * B + s;
    stv(p + i, cfma(beta, ldv(p + i), ldv(u + i)));
  }
}

// Iteration form 2: the same recurrence with u never stored.  The update forms u only for
// its partial u^T r, and the p pass rebuilds it from r and the nodal y (unchanged between
// the two): p = D^{-1} r + G y + beta p.  One vector pass fewer per iteration (the u
// store) for a gather of y (node-sized, read in 64-system rows) and the edge diagonal.
// XLAG: the previous iteration's x += alpha p is made here, where p is read anyway, instead
// of in the update (which then reads neither x nor p): one vector pass fewer again.  The
// product is the same cfma of the same operands, so x is bit-identical; the last
// iteration's update of every system is made by k_sw_xflush after the loop.
template <typename VT, bool XLAG>
__global__ void __launch_bounds__(kBlock)
    k_sw_pform_u(int64_t n, int B, int P, const int2 *__restrict__ en_eff, const double2 *__restrict__ dK,
                 const double2 *__restrict__ dM, const double *__restrict__ dMs, const int8_t *__restrict__ rport,
                 const double *__restrict__ k0sq, const double2 *__restrict__ coef, const VT *__restrict__ y,
                 const VT *__restrict__ r, VT *__restrict__ p, VT *__restrict__ x, const SwState *__restrict__ st) {
  const int s = blockIdx.y * kSys + (threadIdx.x & 63);
  const int wv = threadIdx.x >> 6;
  if (s >= B || st[s].active == 0)
    return;
  const double2 beta = st[s].beta;
  const double k2 = k0sq[s];
  const bool xu = XLAG && st[s].iters > 0;  // an iteration done: its x update is pending
  const double2 alpha = st[s].alpha;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
    const int64_t i = e * B + s;
    const int2 ab = en_eff[e];
    double2 z = bil(edge_diag_inv(e, s, P, k2, dK, dM, dMs, rport, coef), ldv(r + i));
    {  // both node values loaded unconditionally (node 0 for an edge without nodes)
      const bool hn = ab.x >= 0;
      const double2 yd = csub(ldv(y + (int64_t)(hn ? ab.y : 0) * B + s), ldv(y + (int64_t)(hn ? ab.x : 0) * B + s));
      z = hn ? cadd(z, yd) : z;
    }
    const double2 po = ldv(p + i);
    if (xu)
      stv(x + i, cfma(alpha, po, ldv(x + i)));
    stv(p + i, cfma(beta, po, z));
  }
}

// The pending x += alpha p of the last iteration of every system that made one (not after
// a breakdown in the alpha step, which stops a system before its update).
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_xflush(int64_t n, int B, const VT *__restrict__ p, VT *__restrict__ x, const SwState *__restrict__ st) {
  const int s = blockIdx.y * kSys + (threadIdx.x & 63);
  const int wv = threadIdx.x >> 6;
  if (s >= B || st[s].iters == 0 || st[s].breakdown != 0)
    return;
  const double2 alpha = st[s].alpha;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
    const int64_t i = e * B + s;
    stv(x + i, cfma(alpha, ldv(p + i), ldv(x + i)));
  }
}

// x += alpha p (not with XLAG: made by the next p pass); r -= alpha w; u = D^{-1} r + G y
// (stored when STORE_U); partials u^T r, |r|^2.  Persistent: one block slice walks the
// groups of 64 systems in turn.
template <typename VT, bool STORE_U, bool XLAG = false>
__global__ void __launch_bounds__(kBlock)
    k_sw_update_std(int64_t n, int B, int groups, int P, const int2 *__restrict__ en_eff,
                    const double2 *__restrict__ dK, const double2 *__restrict__ dM,
                    const double *__restrict__ dMs, const int8_t *__restrict__ rport,
                    const double *__restrict__ k0sq, const double2 *__restrict__ coef,
                    const VT *__restrict__ y, const VT *__restrict__ w, const VT *__restrict__ p,
                    VT *__restrict__ u, VT *__restrict__ x, VT *__restrict__ r, double2 *__restrict__ partial,
                    const SwState *__restrict__ st) {
  __shared__ double2 lds[kWaves * 2 * 64];
  const int wv = threadIdx.x >> 6;
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
    const int s = g * kSys + (threadIdx.x & 63);
    double2 v[2] = {make_double2(0, 0), make_double2(0, 0)};
    const bool act = s < B && st[s].active != 0;
    if (act) {
      const double k2 = k0sq[s];
      const double2 alpha = st[s].alpha;
      for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
        const int64_t i = e * B + s;
        if (!XLAG)
          stv(x + i, cfma(alpha, ldv(p + i), ldv(x + i)));
        const double2 rv = csub(ldv(r + i), bil(alpha, ldv(w + i)));
        stv(r + i, rv);
        const double2 di = edge_diag_inv(e, s, P, k2, dK, dM, dMs, rport, coef);
        double2 z = bil(di, rv);
        const int2 ab = en_eff[e];
        {  // both node values loaded unconditionally (node 0 for an edge without nodes)
          const bool hn = ab.x >= 0;
          const double2 yd =
              csub(ldv(y + (int64_t)(hn ? ab.y : 0) * B + s), ldv(y + (int64_t)(hn ? ab.x : 0) * B + s));
          z = hn ? cadd(z, yd) : z;
        }
        if (STORE_U)
          stv(u + i, z);
        v[0] = cadd(v[0], bil(rv, z));
        v[1].x += cabs2(rv);
      }
    }
    write_sys_partials<2>(v, lds, partial, g * gridDim.x + blockIdx.x);
    __syncthreads();  // lds reuse by the next group
  }
}

// One row of A_s applied to a vector of the layout [row][system], called by a whole
// wavefront (row wave-uniform, one system per lane).  The row's column indices and
// shared K, M, M_s values are loaded once across the lanes (coalesced) and broadcast
// with readlane, so the gathers of v carry no dependent index load and are issued
// several at a time.
template <typename VT>
__device__ inline double2 sw_row_apply(int64_t row, int s, int B, int P, double k2,
                                       const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                       const double2 *__restrict__ K, const double2 *__restrict__ M,
                                       const double *__restrict__ Ms, const int8_t *__restrict__ rport,
                                       const double2 *__restrict__ coef, const VT *__restrict__ v) {
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const double2 c = port_coef(coef, s, P, rport[row]);
  const int32_t beg = row_ptr[row], end = row_ptr[row + 1];
  double2 acc = make_double2(0, 0);
  for (int32_t base = beg; base < end; base += 64) {
    const int len = min(64, end - base);
    int cj = 0;
    double kx = 0, ky = 0, mx = 0, my = 0, ms = 0;
    if (lane < len) {
      cj = col[base + lane];
      const double2 kv = K[base + lane], mv = M[base + lane];
      kx = kv.x;
      ky = kv.y;
      mx = mv.x;
      my = mv.y;
      ms = Ms[base + lane];
    }
    for (int j = 0; j < len; j += U) {
      double2 xv[U];
#pragma unroll
      for (int t = 0; t < U; ++t)
        if (j + t < len)
          xv[t] = ldv(v + (int64_t)bcast(cj, j + t) * B + s);
#pragma unroll
      for (int t = 0; t < U; ++t)
        if (j + t < len) {
          const int l = j + t;
          const double m = bcast(ms, l);
          const double2 a = make_double2(bcast(kx, l) - k2 * bcast(mx, l) + c.x * m,
                                         bcast(ky, l) - k2 * bcast(my, l) + c.y * m);
          acc = cfma(a, xv[t], acc);
        }
    }
  }
  return acc;
}

// Lane-staged metadata of one row (up to 64 entries; longer rows take the general path).
struct RowStage {
  int32_t beg, len;  // wave-uniform
  int cj;
  double kx, ky, mx, my, ms;
  double2 c;  // port coefficient of this lane's system on this row
};

__device__ inline void stage_row(RowStage &q, int64_t row, int s, int P, const int32_t *__restrict__ row_ptr,
                                 const int32_t *__restrict__ col, const double2 *__restrict__ K,
                                 const double2 *__restrict__ M, const double *__restrict__ Ms,
                                 const int8_t *__restrict__ rport, const double2 *__restrict__ coef) {
  const int lane = threadIdx.x & 63;
  q.beg = row_ptr[row];
  q.len = row_ptr[row + 1] - q.beg;
  q.c = port_coef(coef, s, P, rport[row]);
  q.cj = 0;
  q.kx = q.ky = q.mx = q.my = q.ms = 0.0;
  if (lane < q.len && lane < 64) {
    const int32_t k = q.beg + lane;
    q.cj = col[k];
    const double2 kv = K[k], mv = M[k];
    q.kx = kv.x;
    q.ky = kv.y;
    q.mx = mv.x;
    q.my = mv.y;
    q.ms = Ms[k];
  }
}

// w = A_s u; partials d = u^T w.  Persistent, XCD-aware: the blocks of one XCD walk
// the XCD's contiguous share of the rows, so the gathered columns of u (64 systems x
// 16 B per column) stay in that XCD's L2; blockIdx.y strides over the system groups.
// Each wave keeps one row staged ahead: the next row's indices and values are loaded
// while the current row's gathers are in flight.
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_spmv(int64_t n, int B, int P, int groups, const int32_t *__restrict__ row_ptr,
              const int32_t *__restrict__ col, const double2 *__restrict__ K,
              const double2 *__restrict__ M, const double *__restrict__ Ms,
              const int8_t *__restrict__ rport, const double *__restrict__ k0sq,
              const double2 *__restrict__ coef, const VT *__restrict__ u,
              VT *__restrict__ w, double2 *__restrict__ partial,
              const SwState *__restrict__ st) {
  constexpr int U = 16;
  __shared__ double2 lds[kWaves * 64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t nchunks = (n + kWaves - 1) / kWaves;
  ChunkWalk cw(nchunks);
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
    const int s = g * kSys + lane;
    double2 v[1] = {make_double2(0, 0)};
    const bool act = s < B && st[s].active != 0;
    // the wave's rows are walked only if one of its 64 systems is still active; the
    // gathers of inactive lanes read valid memory and are discarded
    if (__ballot(act) != 0) {
      const double k2 = k0sq[s];
      int64_t chunk = cw.begin;
      int64_t row = chunk * kWaves + wv;
      RowStage cur;
      if (chunk < cw.end && row < n)
        stage_row(cur, row, s, P, row_ptr, col, K, M, Ms, rport, coef);
      while (chunk < cw.end && row < n) {
        const int64_t nchunk = chunk + cw.stride;
        const int64_t nrow = nchunk * kWaves + wv;
        double2 acc = make_double2(0, 0);
        if (cur.len <= 64) {
          double2 xv[U];
#pragma unroll
          for (int t = 0; t < U; ++t)
            if (t < cur.len)
              xv[t] = ldv(u + (int64_t)bcast(cur.cj, t) * B + s);
          RowStage nxt;
          if (nchunk < cw.end && nrow < n)
            stage_row(nxt, nrow, s, P, row_ptr, col, K, M, Ms, rport, coef);
#pragma unroll
          for (int t = 0; t < U; ++t)
            if (t < cur.len) {
              const double m = bcast(cur.ms, t);
              const double2 a = make_double2(bcast(cur.kx, t) - k2 * bcast(cur.mx, t) + cur.c.x * m,
                                             bcast(cur.ky, t) - k2 * bcast(cur.my, t) + cur.c.y * m);
              acc = cfma(a, xv[t], acc);
            }
          for (int j = U; j < cur.len; j += 8) {
#pragma unroll
            for (int t = 0; t < 8; ++t)
              if (j + t < cur.len)
                xv[t] = ldv(u + (int64_t)bcast(cur.cj, j + t) * B + s);
#pragma unroll
            for (int t = 0; t < 8; ++t)
              if (j + t < cur.len) {
                const int l = j + t;
                const double m = bcast(cur.ms, l);
                const double2 a = make_double2(bcast(cur.kx, l) - k2 * bcast(cur.mx, l) + cur.c.x * m,
                                               bcast(cur.ky, l) - k2 * bcast(cur.my, l) + cur.c.y * m);
                acc = cfma(a, xv[t], acc);
              }
          }
          cur = nxt;
        } else {
          acc = sw_row_apply(row, s, B, P, k2, row_ptr, col, K, M, Ms, rport, coef, u);
          if (nchunk < cw.end && nrow < n)
            stage_row(cur, nrow, s, P, row_ptr, col, K, M, Ms, rport, coef);
        }
        if (act) {
          stv(w + row * B + s, acc);
          v[0] = cadd(v[0], bil(ldv(u + row * B + s), acc));
        }
        chunk = nchunk;
        row = nrow;
      }
    }
    write_sys_partials<1>(v, lds, partial, g * gridDim.x + blockIdx.x);
    __syncthreads();  // lds reuse by the next group
  }
}

// Row metadata for the pipelined SpMV: bounds and port of 64 upcoming rows are held one
// per lane, so staging a row's columns and values needs no dependent row_ptr load.
// Lanes past the row's length hold the row's own index as column and zero values, so
// a fixed number of gathers can be issued with no branch (exact wait counts).
struct RowStage2 {
  int32_t len;  // wave-uniform
  int port;     // wave-uniform
  int cj;
  double kx, ky, mx, my, ms;
  double2 c;
};

// One row with NU >= len gathers, all issued before the next row is staged.  KC / MC:
// K / M have nonzero imaginary parts (otherwise only their real parts are broadcast and
// multiplied).
template <int NU, bool KC, bool MC, typename VT, typename StageFn>
__device__ inline double2 sw_row_fixed(const RowStage2 &cur, RowStage2 &nxt, StageFn &&stage, int jn,
                                       int64_t rowc, int s, int B, double k2, const VT *__restrict__ u,
                                       double2 &ud) {
  ud = ldv(u + rowc * B + s);
  double2 xv[NU];
#pragma unroll
  for (int t = 0; t < NU; ++t)
    xv[t] = ldv(u + (int64_t)bcast(cur.cj, t) * B + s);
  stage(nxt, jn);
  // K x, M x and Ms x accumulated apart (each product takes one scalar operand), then
  // combined with this lane's k0^2 and port coefficient once per row
  double2 aK = make_double2(0, 0), aM = make_double2(0, 0), aS = make_double2(0, 0);
#pragma unroll
  for (int t = 0; t < NU; ++t) {
    const double kr = bcast(cur.kx, t), mr = bcast(cur.mx, t);
    if (KC) {
      aK = cfma(make_double2(kr, bcast(cur.ky, t)), xv[t], aK);
    } else {
      aK.x += kr * xv[t].x;
      aK.y += kr * xv[t].y;
    }
    if (MC) {
      aM = cfma(make_double2(mr, bcast(cur.my, t)), xv[t], aM);
    } else {
      aM.x += mr * xv[t].x;
      aM.y += mr * xv[t].y;
    }
  }
  if (cur.port >= 0) {
#pragma unroll
    for (int t = 0; t < NU; ++t) {
      const double m = bcast(cur.ms, t);
      aS.x += m * xv[t].x;
      aS.y += m * xv[t].y;
    }
  }
  return cfma(cur.c, aS, make_double2(aK.x - k2 * aM.x, aK.y - k2 * aM.y));
}

// w = A_s u; partials d = u^T w.  Groups of 64 systems are walked one after the other
// (gridDim.y == 1, so all resident blocks work on the same group and the XCD's share
// of u for that group, rows +- the matrix bandwidth, stays in its L2).  Each wave
// preloads the bounds of its next 64 rows across its lanes, stages the next row's
// columns and values while the current row's gathers are in flight, and issues a fixed
// number of gathers per row (13 or 20; longer rows take the general path).
template <typename VT, bool KC, bool MC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8)))
    k_sw_spmv_pre(int64_t n, int B, int P, int groups, const int32_t *__restrict__ row_ptr,
                  const int32_t *__restrict__ col, const double2 *__restrict__ K,
                  const double2 *__restrict__ M, const double *__restrict__ Ms,
                  const int8_t *__restrict__ rport, const double *__restrict__ k0sq,
                  const double2 *__restrict__ coef, const VT *__restrict__ u,
                  VT *__restrict__ w, double2 *__restrict__ partial,
                  const SwState *__restrict__ st) {
  __shared__ double2 lds[kWaves * 64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t nchunks = (n + kWaves - 1) / kWaves;
  ChunkWalk cw(nchunks);
  // rows of this wave: j = 0 .. nr-1, row_j = (begin + j stride) kWaves + wv (may be >= n
  // in the last chunk: such a row has length 0, reads row n-1 and writes nothing)
  const int64_t nr = (cw.begin < cw.end) ? (cw.end - cw.begin + cw.stride - 1) / cw.stride : 0;
  const double *Kd = reinterpret_cast<const double *>(K);
  const double *Md = reinterpret_cast<const double *>(M);
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
    const int s = g * kSys + lane;
    double2 v[1] = {make_double2(0, 0)};
    const bool act = s < B && st[s].active != 0;
    if (__ballot(act) != 0) {
      const double k2 = k0sq[s];
      for (int64_t jb = 0; jb < nr; jb += 64) {
        const int cnt = (int)min<int64_t>(64, nr - jb);
        int lo = 0, hi = 0, pt = -1;
        if (lane < cnt) {
          const int64_t row = (cw.begin + (jb + lane) * cw.stride) * kWaves + wv;
          if (row < n) {
            lo = row_ptr[row];
            hi = row_ptr[row + 1];
            pt = rport[row];
          }
        }
        auto row_of = [&](int j) { return min<int64_t>((cw.begin + (jb + j) * cw.stride) * kWaves + wv, n - 1); };
        // stage row j (j < cnt); every load is unconditional
        auto stage = [&](RowStage2 &q, int j) {
          const int32_t beg = bcast(lo, j);
          q.len = bcast(hi, j) - beg;
          q.port = bcast(pt, j);
          const double2 cv = coef[s * P + max(q.port, 0)];  // coef holds >= 1 entry per system
          q.c = (q.port >= 0) ? cv : make_double2(0, 0);
          const bool in = lane < q.len;
          const int32_t k = in ? beg + lane : 0;
          const int cl = col[k];
          double kx, ky = 0.0, mx, my = 0.0;
          if (KC) {
            const double2 kv = K[k];
            kx = kv.x;
            ky = kv.y;
          } else {
            kx = Kd[2 * (int64_t)k];
          }
          if (MC) {
            const double2 mv = M[k];
            mx = mv.x;
            my = mv.y;
          } else {
            mx = Md[2 * (int64_t)k];
          }
          const double mw = Ms[k];
          q.cj = in ? cl : (int)row_of(j);
          q.kx = in ? kx : 0.0;
          q.ky = in ? ky : 0.0;
          q.mx = in ? mx : 0.0;
          q.my = in ? my : 0.0;
          q.ms = in ? mw : 0.0;
        };
        RowStage2 cur, nxt;
        stage(cur, 0);
        for (int j = 0; j < cnt; ++j) {
          const int64_t row = (cw.begin + (jb + j) * cw.stride) * kWaves + wv;
          const int64_t rowc = min<int64_t>(row, n - 1);
          const int jn = min(j + 1, cnt - 1);
          double2 acc, ud;
          if (cur.len <= 13) {
            acc = sw_row_fixed<13, KC, MC>(cur, nxt, stage, jn, rowc, s, B, k2, u, ud);
            cur = nxt;
          } else if (cur.len <= 20) {
            acc = sw_row_fixed<20, KC, MC>(cur, nxt, stage, jn, rowc, s, B, k2, u, ud);
            cur = nxt;
          } else {
            acc = sw_row_apply(rowc, s, B, P, k2, row_ptr, col, K, M, Ms, rport, coef, u);
            ud = ldv(u + rowc * B + s);
            stage(cur, jn);
          }
          if (act && row < n) {
            stv(w + row * B + s, acc);
            v[0] = cadd(v[0], bil(ud, acc));
          }
        }
      }
    }
    write_sys_partials<1>(v, lds, partial, g * gridDim.x + blockIdx.x);
    __syncthreads();  // lds reuse by the next group
  }
}

// ---------------------------------------------------------------------------
// LDS-tiled SpMV.  Rows are cut into tiles of kTileRows consecutive rows.  The entries
// of a tile reference few distinct columns (on tetrahedral meshes ~56 for 8 rows of ~15
// entries), so the block loads each distinct column of u once per system group into LDS
// (one 64-lane load per column) and the rows read their columns from LDS: the vector
// memory work per row drops from one load per entry to ~7, which relieves the address
// unit that bounds the per-entry gathers.  Format, built once per sweep by
// k_sw_build_tiles:
//   ucol[t kTileCols + c], c < nu[t]: the distinct columns of tile t's entries and its
//     own row indices, sorted (slots nu[t] .. kTileCols-1 repeat the first one)
//   lidx[k]: slot of entry k's column (8 bits; lidx[nnz] = 0 for the padding entry)
//   dloc[r]: slot of row r's own index
//   nu[t] = -1: the tile does not fit (more than kTileCols distinct columns, more than
//     kTileKeys entries and rows, or a row longer than 64 entries); the tiled kernel is
//     then not used for the matrix.
// K, M and Ms carry one zero entry at index nnz, which lanes past a row's end read.
// ---------------------------------------------------------------------------
constexpr int kTileRows = 8;
constexpr int kTileCols = 72;
constexpr int kTileKeys = 512;
constexpr int kTileWaves = 8;  // blocks of 8 waves, one row of the tile each
constexpr int kTileBlock = kTileWaves * 64;
constexpr int kRowsPerWave = kTileRows / kTileWaves;
constexpr int kColSlots = kTileCols / kTileWaves;  // column slots of wave w: w + kTileWaves i

__global__ void __launch_bounds__(kBlock)
    k_sw_build_tiles(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                     int32_t *__restrict__ ucol, int32_t *__restrict__ nu, uint8_t *__restrict__ lidx,
                     uint8_t *__restrict__ dloc) {
  __shared__ int32_t key[kTileKeys];
  __shared__ int32_t scan[kBlock];
  __shared__ int too_long;
  const int tid = threadIdx.x;
  const int64_t t = blockIdx.x;
  const int64_t r0 = t * kTileRows;
  const int nr = (int)min<int64_t>(kTileRows, n - r0);
  const int32_t k0 = row_ptr[r0];
  const int ne = row_ptr[r0 + nr] - k0, m = ne + nr;
  if (tid == 0)
    too_long = 0;
  __syncthreads();
  if (tid < nr && row_ptr[r0 + tid + 1] - row_ptr[r0 + tid] > 64)
    too_long = 1;
  __syncthreads();
  if (too_long || m > kTileKeys) {
    if (tid == 0)
      nu[t] = -1;
    return;
  }
  for (int i = tid; i < kTileKeys; i += kBlock)
    key[i] = i < ne ? col[k0 + i] : (i < m ? (int32_t)(r0 + (i - ne)) : INT32_MAX);
  __syncthreads();
  for (int size = 2; size <= kTileKeys; size <<= 1) {  // bitonic sort
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < kTileKeys; i += kBlock) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const int32_t a = key[i], b = key[j];
          if ((a > b) == up) {
            key[i] = b;
            key[j] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  constexpr int PER = kTileKeys / kBlock;  // thread k flags keys PER k .. PER k + PER - 1
  int flag[PER], cnt = 0;
  int32_t mine[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = tid * PER + e;
    mine[e] = key[i];
    flag[e] = (mine[e] != INT32_MAX && (i == 0 || mine[e] != key[i - 1])) ? 1 : 0;
    cnt += flag[e];
  }
  scan[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {  // inclusive Hillis-Steele scan
    const int v = tid >= off ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  const int total = scan[kBlock - 1];
  if (total > kTileCols) {
    if (tid == 0)
      nu[t] = -1;
    return;
  }
  int pos = scan[tid] - cnt;
  const int32_t first = key[0];
  __syncthreads();  // every thread has read key[] before it is compacted
#pragma unroll
  for (int e = 0; e < PER; ++e)
    if (flag[e]) {
      key[pos] = mine[e];
      ucol[t * kTileCols + pos] = mine[e];
      ++pos;
    }
  for (int i = total + tid; i < kTileCols; i += kBlock)
    ucol[t * kTileCols + i] = first;
  if (tid == 0)
    nu[t] = total;
  __syncthreads();
  for (int i = tid; i < m; i += kBlock) {
    const int32_t c = i < ne ? col[k0 + i] : (int32_t)(r0 + (i - ne));
    int lo = 0, hi = total - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (key[mid] < c)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (i < ne)
      lidx[k0 + i] = (uint8_t)lo;
    else
      dloc[r0 + (i - ne)] = (uint8_t)lo;
  }
}

// Tile metadata of one step for one wave, loaded two steps ahead of its use.
struct TileMeta {
  int32_t uc;  // lane i < kColSlots: column of slot w + kTileWaves i
  int32_t rp;  // lane i <= kRowsPerWave: row_ptr of the wave's row i (row clamped to n)
  int32_t pt;  // lane i < kRowsPerWave: port of row i (row clamped to n - 1)
  int32_t dl;  // lane i < kRowsPerWave: slot of row i's own index
  int nu;      // distinct columns of the tile
};

// Row entries of one step, loaded one step ahead: lane e holds entry e of each of the
// wave's rows (the zero padding entry past the row's end).  Port rows (few) load their
// port masses and coefficient when they are computed.
struct TileRows {
  int li[kRowsPerWave];
  double kx[kRowsPerWave], ky[kRowsPerWave], mx[kRowsPerWave], my[kRowsPerWave];
  double k2;
  int lo[kRowsPerWave], len[kRowsPerWave], port[kRowsPerWave], dl[kRowsPerWave];
  int64_t row0;
  int g, nu;
};

// Accumulates K x and M x over a row's entries read from LDS (NU fixed entries, the
// lanes past the row's end holding zero coefficients; NU = 0: len entries).
template <int NU, bool KC, bool MC, typename VT>
__device__ inline void tile_row(const VT *su, int lane, int len, const TileRows &q, int r, double2 &aK,
                                double2 &aM) {
  auto entry = [&](int t) {
    const double2 x = ldv(su + bcast(q.li[r], t) * 64 + lane);
    const double kr = bcast(q.kx[r], t), mr = bcast(q.mx[r], t);
    if (KC) {
      aK = cfma(make_double2(kr, bcast(q.ky[r], t)), x, aK);
    } else {
      aK.x += kr * x.x;
      aK.y += kr * x.y;
    }
    if (MC) {
      aM = cfma(make_double2(mr, bcast(q.my[r], t)), x, aM);
    } else {
      aM.x += mr * x.x;
      aM.y += mr * x.y;
    }
  };
  if constexpr (NU > 0) {
#pragma unroll
    for (int t = 0; t < NU; ++t)
      entry(t);
  } else {
    for (int t = 0; t < len; ++t)
      entry(t);
  }
}

// w = A_s u; partials d = u^T w, on the tile format.  Each block walks its tiles (an
// XCD-contiguous range, as the other pipelined kernels) once per active system group,
// groups in turn, as one flat sequence of steps.  Step j: the columns of step j (in
// registers since step j-1) are written to LDS; the stores of step j-1 are issued; the
// metadata of step j+2 and the columns and entries of step j+1 are loaded; step j's rows
// are computed from LDS.  The stores precede the loads, so the wait for step j+1's loads
// at the top of the next step never includes stores issued after them.
template <typename VT, bool KC, bool MC>
__global__ void __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_sw_spmv_tile(int64_t n, int B, int P, int groups, const int32_t *__restrict__ row_ptr,
                   const int32_t *__restrict__ ucol, const int32_t *__restrict__ nu,
                   const uint8_t *__restrict__ lidx, const uint8_t *__restrict__ dloc,
                   const double2 *__restrict__ K, const double2 *__restrict__ M, const double *__restrict__ Ms,
                   const int8_t *__restrict__ rport, const double *__restrict__ k0sq,
                   const double2 *__restrict__ coef, const VT *__restrict__ u, VT *__restrict__ w,
                   double2 *__restrict__ partial, const SwState *__restrict__ st) {
  __shared__ VT su[kTileCols * 64];
  static_assert(kTileCols * 64 * sizeof(VT) >= kTileWaves * 64 * sizeof(double2), "partials reuse su");
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar branches)
  const int64_t ntl = (n + kTileRows - 1) / kTileRows;
  ChunkWalk cw(ntl);
  const int nt = (cw.begin < cw.end) ? (int)((cw.end - cw.begin + cw.stride - 1) / cw.stride) : 0;
  // groups with an active system (the same mask in every wave of the block) and this
  // lane's active flag per group
  uint32_t gmask = 0, mine = 0;
  for (int g = 0; g < groups; ++g) {
    const int s = g * kSys + lane;
    const bool a = s < B && st[s].active != 0;
    mine |= (a ? 1u : 0u) << g;
    if (__ballot(a) != 0)
      gmask |= 1u << g;
  }
  if (threadIdx.x < 64)
    for (int g = 0; g < groups; ++g)
      if (nt == 0 || !((gmask >> g) & 1))
        partial[((int64_t)g * gridDim.x + blockIdx.x) * 64 + lane] = make_double2(0, 0);
  const int nsteps = nt * __popc(gmask);
  if (nsteps == 0)
    return;
  auto group_of = [&](int j) {  // the (j / nt)-th active group
    uint32_t mk = gmask;
    for (int k = j / nt; k > 0; --k)
      mk &= mk - 1;
    return __ffs(mk) - 1;
  };
  auto tile_of = [&](int j) { return cw.begin + (int64_t)(j % nt) * cw.stride; };
  const double *Kd = reinterpret_cast<const double *>(K);
  const double *Md = reinterpret_cast<const double *>(M);
  const int32_t pad = (int32_t)row_ptr[n];  // the zero padding entry
  auto load_meta = [&](TileMeta &m, int j) {
    const int64_t t = tile_of(j);
    const int64_t row0 = t * kTileRows + wv * kRowsPerWave;
    m.uc = ucol[t * kTileCols + min(wv + kTileWaves * lane, kTileCols - 1)];
    m.rp = row_ptr[min<int64_t>(row0 + min(lane, kRowsPerWave), n)];
    const int64_t rr = min<int64_t>(row0 + min(lane, kRowsPerWave - 1), n - 1);
    m.pt = rport[rr];
    m.dl = dloc[rr];
    m.nu = nu[t];
  };
  // issues the loads of step j (no loaded value is used here).  Every column slot is
  // loaded (slots past nu hold a valid repeated column), so the loads are branch-free and
  // no register merge waits for them.
  using ST = typename std::conditional<std::is_same<VT, float2>::value, float, double>::type;
  auto load_data = [&](ST (&ux)[kColSlots], ST (&uy)[kColSlots], TileRows &q, const TileMeta &m, int j) {
    const int g = group_of(j);
    const int s = g * kSys + lane;
#pragma unroll
    for (int i = 0; i < kColSlots; ++i) {
      const VT t = u[(int64_t)bcast(m.uc, i) * B + s];
      ux[i] = t.x;
      uy[i] = t.y;
    }
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
      const int32_t lo = bcast(m.rp, r);
      const int len = bcast(m.rp, r + 1) - lo;
      const int32_t k = lane < len ? lo + lane : pad;
      q.li[r] = lidx[k];
      if (KC) {
        const double2 kv = K[k];
        q.kx[r] = kv.x;
        q.ky[r] = kv.y;
      } else {
        q.kx[r] = Kd[2 * (int64_t)k];
        q.ky[r] = 0.0;
      }
      if (MC) {
        const double2 mv = M[k];
        q.mx[r] = mv.x;
        q.my[r] = mv.y;
      } else {
        q.mx[r] = Md[2 * (int64_t)k];
        q.my[r] = 0.0;
      }
      q.lo[r] = lo;
      q.len[r] = len;
      q.port[r] = bcast(m.pt, r);
      q.dl[r] = bcast(m.dl, r);
    }
    q.k2 = k0sq[s];
    q.row0 = tile_of(j) * kTileRows + wv * kRowsPerWave;
    q.g = g;
    q.nu = m.nu;
  };
  ST ux[kColSlots], uy[kColSlots];
  double2 v = make_double2(0, 0);
  double2 pacc[kRowsPerWave];
  int64_t prow0 = -1;
  int ps = 0;
  bool pact = false;
  // step j: cur holds step j's entries, nxt receives step j+1's; mnext holds step j+1's
  // metadata, mfree receives step j+2's (the two register sets alternate, no copies)
  auto step = [&](int j, const TileRows &cur, TileRows &nxt, const TileMeta &mnext, TileMeta &mfree) {
    __syncthreads();  // the previous step's rows have read su
#pragma unroll
    for (int i = 0; i < kColSlots; ++i) {
      VT t;
      t.x = ux[i];
      t.y = uy[i];
      su[(wv + kTileWaves * i) * 64 + lane] = t;
    }
    __syncthreads();
    if (prow0 >= 0 && pact) {
#pragma unroll
      for (int r = 0; r < kRowsPerWave; ++r)
        if (prow0 + r < n)
          stv(w + (prow0 + r) * B + ps, pacc[r]);
    }
    load_data(ux, uy, nxt, mnext, min(j + 1, nsteps - 1));
    load_meta(mfree, min(j + 2, nsteps - 1));
    const int s = cur.g * kSys + lane;
    const bool act = (mine >> cur.g) & 1;
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
      double2 aK = make_double2(0, 0), aM = make_double2(0, 0);
      const int len = cur.len[r];
      if (len <= 13)
        tile_row<13, KC, MC>(su, lane, len, cur, r, aK, aM);
      else if (len <= 20)
        tile_row<20, KC, MC>(su, lane, len, cur, r, aK, aM);
      else
        tile_row<0, KC, MC>(su, lane, len, cur, r, aK, aM);
      double2 acc = make_double2(aK.x - cur.k2 * aM.x, aK.y - cur.k2 * aM.y);
      if (cur.port[r] >= 0) {  // port row: + c_p (Ms x)
        const double msv = Ms[lane < len ? cur.lo[r] + lane : pad];
        const double2 c = coef[s * P + cur.port[r]];
        double2 aS = make_double2(0, 0);
        for (int t = 0; t < len; ++t) {
          const double2 x = ldv(su + bcast(cur.li[r], t) * 64 + lane);
          const double m = bcast(msv, t);
          aS.x += m * x.x;
          aS.y += m * x.y;
        }
        acc = cfma(c, aS, acc);
      }
      pacc[r] = acc;
      if (act && cur.row0 + r < n)
        v = cadd(v, bil(ldv(su + cur.dl[r] * 64 + lane), acc));
    }
    prow0 = cur.row0;
    ps = s;
    pact = act;
    if (j + 1 == nsteps || group_of(j + 1) != cur.g) {
      // block partial of the group, fixed order; su is free once every wave is here (the
      // next step's first barrier orders these reads before su is rewritten)
      double2 *red = reinterpret_cast<double2 *>(su);
      __syncthreads();
      red[wv * 64 + lane] = v;
      __syncthreads();
      if (wv == 0) {
        double2 a = red[lane];
        for (int ww = 1; ww < kTileWaves; ++ww)
          a = cadd(a, red[ww * 64 + lane]);
        partial[((int64_t)cur.g * gridDim.x + blockIdx.x) * 64 + lane] = a;
      }
      v = make_double2(0, 0);
    }
  };
  TileRows ra, rb;
  TileMeta ma, mb;
  load_meta(ma, 0);
  load_meta(mb, min(1, nsteps - 1));
  load_data(ux, uy, ra, ma, 0);
  for (int j = 0; j < nsteps; j += 2) {
    step(j, ra, rb, mb, ma);
    if (j + 1 < nsteps)
      step(j + 1, rb, ra, ma, mb);
  }
  if (pact) {
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r)
      if (prow0 + r < n)
        stv(w + (prow0 + r) * B + ps, pacc[r]);
  }
}

// Bit 1: some K value has a nonzero imaginary part; bit 2: some M value has.
__global__ void k_sw_imag_flags(int64_t nnz, const double2 *__restrict__ K, const double2 *__restrict__ M,
                                int *__restrict__ flags) {
  int f = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    if (K[k].y != 0.0)
      f |= 1;
    if (M[k].y != 0.0)
      f |= 2;
  }
  if (f)
    atomicOr(flags, f);
}

// Per-system correction of frequency-dependent media (PML stretch, dispersive models),
//   A_s = K - k0_s^2 M + sum_p c_{s,p} M_s,p + C_s,
// with C_s on a fixed sub-pattern (correction row m is matrix row crow[m], entries
// cptr[m]..cptr[m+1] with columns ccol) and values cval[entry][nsys] per system.
// MODE 0 (after the SpMV, systems with active set): w += C_s u on the correction rows,
// partials of u^T C_s u per block (added to the SpMV's partials by the finalize).
// MODE 1 (before a residual, systems with out_active set): t[m] = C_s x for FP64 x.
// One wavefront per correction row and group of 64 systems (lane = system).
// Values: TERMS = false, cval[entry][ncols] with system s reading column colmap[s] (the
// systems of one frequency share a column); TERMS = true, C_s = sum_q alpha[s][q] C_q with
// kMaxTerms shared matrices cterm[entry][kMaxTerms] (unused terms zero): the terms of an
// entry are one wave-uniform 128-byte read, the per-system values are formed in registers.
constexpr int kMaxTerms = 8;
template <typename VT, typename IT, int MODE, bool TERMS, int NT = kMaxTerms>
__global__ void __launch_bounds__(kBlock)
    k_sw_corr(int64_t nrows, int B, int nsys, const int32_t *__restrict__ crow,
              const int32_t *__restrict__ cptr, const int32_t *__restrict__ ccol,
              const double2 *__restrict__ cval, int ncols, const int32_t *__restrict__ colmap,
              const double2 *__restrict__ cterm, const double2 *__restrict__ alpha,
              const IT *__restrict__ u, VT *__restrict__ w, double2 *__restrict__ t,
              double2 *__restrict__ partial, const SwState *__restrict__ st) {
  __shared__ double2 lds[kWaves * 64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = blockIdx.y;
  const int s = g * kSys + lane;
  double2 v[1] = {make_double2(0, 0)};
  const bool act = s < B && (MODE == 0 ? st[s].active : st[s].out_active) != 0;
  // the value column of this system; padding systems read column 0 and are masked out
  const int cs = (!TERMS && s < nsys) ? colmap[s] : 0;
  const double z = s < nsys ? 1.0 : 0.0;
  double2 al[TERMS ? NT : 1];
#pragma unroll
  for (int q = 0; q < (TERMS ? NT : 1); ++q)
    al[q] = (TERMS && s < nsys) ? alpha[(int64_t)s * kMaxTerms + q] : make_double2(0, 0);
  auto value = [&](int32_t k) -> double2 {
    if constexpr (TERMS) {  // the first NT terms (columns past the terms in use are zero)
      double2 a = make_double2(0, 0);
#pragma unroll
      for (int q = 0; q < NT; ++q)
        a = cfma(al[q], cterm[(int64_t)k * kMaxTerms + q], a);
      return a;
    } else {
      return cscale(cval[(int64_t)k * ncols + cs], z);
    }
  };
  if (__ballot(act) != 0) {
    for (int64_t m = (int64_t)blockIdx.x * kWaves + wv; m < nrows; m += (int64_t)gridDim.x * kWaves) {
      const int32_t k0 = cptr[m], k1 = cptr[m + 1];
      double2 acc = make_double2(0, 0);
      int32_t k = k0;
      for (; k + 4 <= k1; k += 4) {  // four independent value/vector loads in flight
        double2 a[4], x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] = value(k + q);
          x[q] = ldv(u + (int64_t)ccol[k + q] * B + s);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc = cfma(a[q], x[q], acc);
      }
      for (; k < k1; ++k)
        acc = cfma(value(k), ldv(u + (int64_t)ccol[k] * B + s), acc);
      if (act) {
        const int64_t i = (int64_t)crow[m] * B + s;
        if (MODE == 0) {
          stv(w + i, cadd(ldv(w + i), acc));
          v[0] = cadd(v[0], bil(ldv(u + i), acc));
        } else {
          t[m * B + s] = acc;
        }
      }
    }
  }
  if (MODE == 0)
    write_sys_partials<1>(v, lds, partial, g * gridDim.x + blockIdx.x);
}

// Residual replacement: r = b - A_s x in FP64 (x == nullptr: r = b), stored in the Krylov
// vector type; partials of |r|^2 (before rounding).  Systems without out_active are skipped.
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_residual(int64_t n, int B, int P, int groups, const int32_t *__restrict__ row_ptr,
                  const int32_t *__restrict__ col, const double2 *__restrict__ K,
                  const double2 *__restrict__ M, const double *__restrict__ Ms,
                  const int8_t *__restrict__ rport, const double *__restrict__ k0sq,
                  const double2 *__restrict__ coef, const double2 *__restrict__ b,
                  const double2 *__restrict__ x, VT *__restrict__ r, double2 *__restrict__ partial,
                  const SwState *__restrict__ st, const int32_t *__restrict__ cmap,
                  const double2 *__restrict__ tc) {
  __shared__ double2 lds[kWaves * 64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t nchunks = (n + kWaves - 1) / kWaves;
  ChunkWalk cw(nchunks);
  for (int g = blockIdx.y; g < groups; g += gridDim.y) {
    const int s = g * kSys + lane;
    double2 v[1] = {make_double2(0, 0)};
    const bool act = s < B && st[s].out_active != 0;
    if (__ballot(act) != 0) {
      const double k2 = k0sq[s];
      for (int64_t chunk = cw.begin; chunk < cw.end; chunk += cw.stride) {
        const int64_t row = chunk * kWaves + wv;
        if (row >= n)
          continue;
        double2 rv = b[row * B + s];
        if (x) {
          rv = csub(rv, sw_row_apply(row, s, B, P, k2, row_ptr, col, K, M, Ms, rport, coef, x));
          if (cmap) {  // the per-system correction's C_s x on its rows (k_sw_corr, MODE 1)
            const int32_t m = cmap[row];
            if (m >= 0)
              rv = csub(rv, tc[(int64_t)m * B + s]);
          }
        }
        if (act) {
          stv(r + row * B + s, rv);
          v[0].x += cabs2(rv);
        }
      }
    }
    write_sys_partials<1>(v, lds, partial, g * gridDim.x + blockIdx.x);
    __syncthreads();
  }
}

// Reliable update: x64 += d (the solution increment carried in the Krylov vector type
// since the last replacement), then d = 0, for the systems with out_active set.
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_accumulate(int64_t n, int B, VT *__restrict__ d, double2 *__restrict__ x,
                    const SwState *__restrict__ st) {
  const int s = blockIdx.y * kSys + (threadIdx.x & 63);
  const int wv = threadIdx.x >> 6;
  if (s >= B || !st[s].out_active)
    return;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
    const int64_t i = e * B + s;
    x[i] = cadd(x[i], ldv(d + i));
    stv(d + i, make_double2(0, 0));
  }
}

// Start of a residual replacement: the systems that asked for one become the active set
// of the replacement kernels (all = 1: every system, for the final accumulation); the
// others keep their state in saved_active.
__global__ void k_sw_ru_begin(int B, SwState *__restrict__ st, int all) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B)
    return;
  SwState &S = st[s];
  S.saved_active = S.active;
  const int sel = all ? 1 : S.want;
  S.active = sel;
  S.out_active = sel;
}

// After a solve: x += pend p for the systems that stopped with a pending step.
template <typename VT>
__global__ void __launch_bounds__(kBlock)
    k_sw_flush(int64_t n, int B, VT *__restrict__ x, const VT *__restrict__ p, const SwState *__restrict__ st) {
  const int s = blockIdx.y * kSys + (threadIdx.x & 63);
  const int wv = threadIdx.x >> 6;
  if (s >= B)
    return;
  const double2 pend = st[s].pend;
  if (pend.x == 0.0 && pend.y == 0.0)
    return;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + wv; e < n; e += (int64_t)gridDim.x * kWaves) {
    const int64_t i = e * B + s;
    stv(x + i, cfma(pend, ldv(p + i), ldv(x + i)));
  }
}

// Block-wide fixed-order sum of per-thread values.
template <int V>
__device__ inline void block_sum(double2 (&v)[V], double2 (*sh)[kBlock]) {
#pragma unroll
  for (int i = 0; i < V; ++i)
    sh[i][threadIdx.x] = v[i];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
#pragma unroll
      for (int i = 0; i < V; ++i)
        sh[i][threadIdx.x] = cadd(sh[i][threadIdx.x], sh[i][threadIdx.x + off]);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < V; ++i)
    v[i] = sh[i][0];
}

// Final true residual of every system: rn_out = |b - A x64|^2 from the residual kernel's
// partials (gxs blocks per group).
__global__ void __launch_bounds__(kBlock)
    k_sw_true_resid(int B, int gxs, const double2 *__restrict__ ps, SwState *__restrict__ st) {
  const int s = blockIdx.x;
  if (s >= B)
    return;
  const int gy = s / kSys, lane = s % kSys;
  __shared__ double2 sh[1][kBlock];
  double2 v[1] = {make_double2(0, 0)};
  for (int bx = threadIdx.x; bx < gxs; bx += kBlock)
    v[0] = cadd(v[0], ps[((int64_t)gy * gxs + bx) * 64 + lane]);
  block_sum<1>(v, sh);
  if (threadIdx.x == 0)
    st[s].rn_out = v[0].x;
}

// The system of a finalize block (one block per system, gridDim.x = the padded batch, a
// multiple of 64).  The partials of 8 consecutive systems share a 128-byte line, so each
// XCD (block b runs on XCD b % 8) takes a contiguous range of systems and reads every
// line into one L2; one system per block in launch order would fetch each line 8 times.
__device__ inline int fin_system() {
  const int per_xcd = gridDim.x >> 3;
  return (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
}

// One block per system: reduce its partials of g, |r|^2 (update grid, gxu blocks) and
// d (spmv grid, gxs blocks) in a fixed order and advance the scalars.
// skip: this iteration's update stored neither x nor r, so its alpha becomes pending.
template <bool INIT>
__global__ void __launch_bounds__(kBlock)
    k_sw_finalize(int B, int gxu, int gxs, const double2 *__restrict__ pu,
                  const double2 *__restrict__ ps, SwState *__restrict__ st,
                  int *__restrict__ active_flag, double tol2, int max_it, int max_inner, int skip) {
  const int s = fin_system();
  if (s >= B)
    return;
  const int gy = s / kSys, lane = s % kSys;
  double2 v[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int bx = threadIdx.x; bx < gxu; bx += kBlock) {
    const int64_t bl = (int64_t)gy * gxu + bx;
    v[0] = cadd(v[0], pu[(bl * 2 + 0) * 64 + lane]);
    v[1] = cadd(v[1], pu[(bl * 2 + 1) * 64 + lane]);
  }
  for (int bx = threadIdx.x; bx < gxs; bx += kBlock)
    v[2] = cadd(v[2], ps[((int64_t)gy * gxs + bx) * 64 + lane]);
  __shared__ double2 sh[3][kBlock];
  block_sum<3>(v, sh);
  if (threadIdx.x != 0)
    return;
  const double2 g = v[0], d = v[2];
  const double rn2 = v[1].x;
  SwState &S = st[s];
  if (INIT) {
    S.bnorm2 = rn2;
    S.rnorm2 = rn2;
    S.iters = 0;
    S.breakdown = 0;
    S.beta = make_double2(0, 0);
    S.pend = make_double2(0, 0);
    S.rho = g;
    S.active = (S.active && rn2 > 0.0) ? 1 : 0;
    if (S.active) {
      if (d.x == 0.0 && d.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.alpha = cdiv(g, d);
      }
    }
  } else if (S.active) {
    S.pend = skip ? S.alpha : make_double2(0, 0);
    S.iters += 1;
    S.total_iters += 1;
    S.rnorm2 = rn2;
    if (rn2 <= tol2 * S.bnorm2) {
      S.active = 0;
    } else if (S.total_iters >= max_it || S.iters >= max_inner || (S.rho.x == 0.0 && S.rho.y == 0.0)) {
      S.active = 0;
    } else {
      const double2 beta = cdiv(g, S.rho);
      const double2 den = csub(d, cdiv(cmul(beta, g), S.alpha));
      if (den.x == 0.0 && den.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.beta = beta;
        S.alpha = cdiv(g, den);
        S.rho = g;
      }
    }
  }
  active_flag[s] = S.active;
}

// Scalars of the standard form, one block per system.  STAGE 0 (after the initial u = P r):
// rho = u^T r, |b|^2; STAGE 1 (after the SpMV): alpha = rho / p^T w; STAGE 2 (after the
// update): convergence on |r|^2, beta = u^T r / rho.
// With reliable updates (ru_delta2 > 0) STAGE 2 also tracks max |r|^2 since the last
// replacement and asks for a replacement (want) once |r|^2 < ru_delta2 * max |r|^2; a
// system whose recursive residual meets the tolerance pauses (active = 0) until its true
// residual has been formed.  active_flag = active | want << 1.
// STAGE 3 (after a replacement: the update's u = P r partials on the true residual):
// convergence on the true |r|^2, else beta = u^T r / rho_prev with p kept, and the
// systems that took no part get their saved activity back.
template <int STAGE>
__global__ void __launch_bounds__(kBlock)
    k_sw_fin_std(int B, int gx, const double2 *__restrict__ part, int gx2, const double2 *__restrict__ part2,
                 SwState *__restrict__ st, int *__restrict__ active_flag, double tol2, int max_it, int max_inner,
                 double ru_delta2) {
  const int s = fin_system();
  if (s >= B)
    return;
  const int gy = s / kSys, lane = s % kSys;
  constexpr int V = STAGE == 1 ? 1 : 2;
  double2 v[V];
#pragma unroll
  for (int k = 0; k < V; ++k)
    v[k] = make_double2(0, 0);
  // four blocks' partials loaded before any is added (their round trips overlap), added in
  // the order of the plain loop
  int bx = threadIdx.x;
  for (; bx + 3 * kBlock < gx; bx += 4 * kBlock) {
    double2 t[4][V];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < V; ++k)
        t[u][k] = part[(((int64_t)gy * gx + bx + u * kBlock) * V + k) * 64 + lane];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < V; ++k)
        v[k] = cadd(v[k], t[u][k]);
  }
  for (; bx < gx; bx += kBlock) {
    const int64_t bl = (int64_t)gy * gx + bx;
#pragma unroll
    for (int k = 0; k < V; ++k)
      v[k] = cadd(v[k], part[(bl * V + k) * 64 + lane]);
  }
  if (STAGE == 1)  // the per-system correction's partials of u^T C_s u (gx2 blocks per group)
    for (int bx = threadIdx.x; bx < gx2; bx += kBlock)
      v[0] = cadd(v[0], part2[((int64_t)gy * gx2 + bx) * 64 + lane]);
  __shared__ double2 sh[V][kBlock];
  block_sum<V>(v, sh);
  if (threadIdx.x != 0)
    return;
  SwState &S = st[s];
  const bool ru = ru_delta2 > 0.0;
  if (STAGE == 0) {
    const double rn2 = v[1].x;
    S.bnorm2 = rn2;
    S.rnorm2 = rn2;
    S.bn_out = rn2;
    S.rn_out = rn2;
    S.rmax2 = rn2;
    S.want = 0;
    S.iters = 0;
    S.breakdown = 0;
    S.beta = make_double2(0, 0);
    S.pend = make_double2(0, 0);
    S.rho = v[0];
    S.rho_prev = v[0];
    S.active = (S.active && rn2 > 0.0) ? 1 : 0;
  } else if (STAGE == 1) {
    if (S.active) {
      const double2 d = v[0];
      if (d.x == 0.0 && d.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.alpha = cdiv(S.rho, d);
      }
    }
  } else if (STAGE == 2) {
    if (S.active) {
      const double2 g = v[0];
      const double rn2 = v[1].x;
      S.iters += 1;
      S.total_iters += 1;
      S.rnorm2 = rn2;
      if (rn2 <= tol2 * S.bnorm2) {
        S.active = 0;
        if (ru) {  // converged in the recursion: confirm on the true residual
          S.want = 1;
          S.rho_prev = S.rho;
        }
      } else if (S.total_iters >= max_it || S.iters >= max_inner || (S.rho.x == 0.0 && S.rho.y == 0.0)) {
        S.active = 0;
      } else {
        S.rho_prev = S.rho;
        S.beta = cdiv(g, S.rho);
        S.rho = g;
        if (ru) {
          S.rmax2 = fmax(S.rmax2, rn2);
          if (rn2 < ru_delta2 * S.rmax2)
            S.want = 1;
        }
      }
    }
  } else {  // STAGE 3
    if (S.want) {
      const double2 g = v[0];
      const double rn2 = v[1].x;
      S.want = 0;
      S.rnorm2 = rn2;
      S.rn_out = rn2;
      S.rmax2 = rn2;
      if (rn2 <= tol2 * S.bn_out || S.total_iters >= max_it) {
        S.active = 0;
      } else if (S.rho_prev.x == 0.0 && S.rho_prev.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.beta = cdiv(g, S.rho_prev);
        S.rho = g;
        S.active = 1;
      }
    } else {
      S.active = S.saved_active;
    }
  }
  active_flag[s] = S.active | (S.want << 1);
}

// Frequency-independent diagonals: edge diagonals of K, M, Ms; nodal diag(G^T X G).
__global__ void k_sw_edge_diag(int64_t n, const int32_t *__restrict__ row_ptr,
                               const int32_t *__restrict__ col, const double2 *__restrict__ K,
                               const double2 *__restrict__ M, const double *__restrict__ Ms,
                               double2 *__restrict__ dK, double2 *__restrict__ dM,
                               double *__restrict__ dMs) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  double2 a = make_double2(0, 0), b = make_double2(0, 0);
  double c = 0.0;
  for (int32_t k = row_ptr[e]; k < row_ptr[e + 1]; ++k)
    if (col[k] == e) {
      a = K[k];
      b = M[k];
      c = Ms[k];
    }
  dK[e] = a;
  dM[e] = b;
  dMs[e] = c;
}

__device__ inline int gsign_of(const int32_t *__restrict__ en, int32_t e, int32_t node) {
  const int32_t a = en[2 * (int64_t)e], b = en[2 * (int64_t)e + 1];
  if (a == b)
    return 0;
  return (node == b) ? 1 : ((node == a) ? -1 : 0);
}

// one thread per node: diag(G^T X G) for X in {K, M, Ms_p} (Dirichlet edges dropped)
__global__ void k_sw_node_diag(int64_t N, int P, const int32_t *__restrict__ n2e_ptr,
                               const int32_t *__restrict__ n2e, const int32_t *__restrict__ en,
                               const uint8_t *__restrict__ dir, const int32_t *__restrict__ row_ptr,
                               const int32_t *__restrict__ col, const double2 *__restrict__ K,
                               const double2 *__restrict__ M, const double *const *__restrict__ Msp,
                               double2 *__restrict__ dnK, double2 *__restrict__ dnM,
                               double *__restrict__ dnMs) {
  int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  double2 sk = make_double2(0, 0), sm = make_double2(0, 0);
  double sp[kMaxPorts] = {0, 0, 0, 0};
  for (int32_t k = n2e_ptr[nd]; k < n2e_ptr[nd + 1]; ++k) {
    const int32_t e = n2e[k];
    if (dir && dir[e])
      continue;
    const int ge = gsign_of(en, e, (int32_t)nd);
    if (!ge)
      continue;
    for (int32_t q = row_ptr[e]; q < row_ptr[e + 1]; ++q) {
      const int32_t f = col[q];
      if (dir && dir[f])
        continue;
      const int gf = gsign_of(en, f, (int32_t)nd);
      if (!gf)
        continue;
      const double sg = (double)(ge * gf);
      sk.x += sg * K[q].x;
      sk.y += sg * K[q].y;
      sm.x += sg * M[q].x;
      sm.y += sg * M[q].y;
      for (int p = 0; p < P; ++p)
        sp[p] += sg * Msp[p][q];
    }
  }
  dnK[nd] = sk;
  dnM[nd] = sm;
  for (int p = 0; p < P; ++p)
    dnMs[(int64_t)p * N + nd] = sp[p];
}

__global__ void k_sw_gsign(int64_t N, const int32_t *__restrict__ n2e_ptr, const int32_t *__restrict__ n2e,
                           const int32_t *__restrict__ en, const uint8_t *__restrict__ dir,
                           int8_t *__restrict__ gsign) {
  int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  for (int32_t k = n2e_ptr[nd]; k < n2e_ptr[nd + 1]; ++k) {
    const int32_t e = n2e[k];
    gsign[k] = (dir && dir[e]) ? 0 : (int8_t)gsign_of(en, e, (int32_t)nd);
  }
}

__global__ void k_sw_edge_prep(int64_t n, const int32_t *__restrict__ en, const uint8_t *__restrict__ dir,
                               int2 *__restrict__ en_eff) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  const int32_t a = en[2 * e], b = en[2 * e + 1];
  en_eff[e] = ((dir && dir[e]) || a == b) ? make_int2(-1, -1) : make_int2(a, b);
}

// port index of every row (-1 none), and the combined port mass; error flag if two
// ports share a row
__global__ void k_sw_ports(int64_t n, int P, const int32_t *__restrict__ row_ptr,
                           const double *const *__restrict__ Msp, double *__restrict__ Ms,
                           int8_t *__restrict__ rport, int *__restrict__ err) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  int port = -1;
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
    double m = 0.0;
    for (int p = 0; p < P; ++p) {
      const double v = Msp[p][k];
      if (v != 0.0) {
        if (port >= 0 && port != p)
          atomicOr(err, 1);
        port = p;
        m = v;
      }
    }
    Ms[k] = m;
  }
  rport[r] = (int8_t)port;
}

} // namespace

struct efm_sweep {
  int64_t n = 0, nnz = 0, N = 0;
  int P = 0, Bmax = 0;
  const int32_t *row_ptr = nullptr, *col = nullptr;
  const int32_t *n2e_ptr = nullptr, *n2e = nullptr;
  int gx_vec = 8, gx_nodes = 8, gx_spmv = 8, gy_spmv = 1;
  // SpMV kernel: 0 = one block slice per system group, 1 = pipelined group-sequential,
  // 2 = LDS-tiled (when every row tile fits); -1 = chosen on the first solve (in-place
  // timing over the iterations before its first convergence check)
  int spmv_variant = -1;
  int gx_pre = 8, gx_tile = 8;
  bool tiled_ok = false;
  // nodal kernel: 0 = one block slice per system group, 1 = pipelined, XCD-local node
  // walk (groups in turn); -1 = chosen on the first solve (in-place timing, as for the SpMV)
  int aux_variant = -1;
  int gx_aux = 8;
  int cplx = 3;  // bit 1: K complex, bit 2: M complex (selects the SpMV instantiation)
  // iteration: 0 = standard COCG (p in its own pass, two reductions), 1 = Chronopoulos-Gear,
  // 2 (default) = standard COCG with u rebuilt in the p pass instead of stored
  int form = 2;
  // standard forms: replay each full convergence-check chunk of iterations from a HIP graph
  // captured once per solve (after the autotuner has chosen the kernels)
  int use_graph = 0;
  hipStream_t capture_stream = nullptr;  // the chunk is captured here (a legacy null stream cannot capture)
  int gx_std[2] = {8, 8};  // grid of the standard form's update [VT c128, c64]
  // update grids (one block slice, as many blocks as are resident): [VT c128, c64][INIT, skip, apply]
  int gx_upd[2][3] = {{8, 8, 8}, {8, 8, 8}};
  float tune_ms[5] = {0.f, 0.f, 0.f, 0.f, 0.f};  // SpMV 0, 1, 2; nodal 0, 1
  efm::DevBuf K, M, Ms, rport, dK, dM, dMs, dnK, dnM, dnMs, gsign, en_eff;
  efm::DevBuf x, r, p, sv, u, w, y, part_u, part_s, state, flags, k0sq, coef;
  efm::DevBuf x64;  // reliable updates: FP64 solution (allocated on first use)
  efm::DevBuf dninv;  // [N][Bmax] inverse nodal diagonals of the systems of one solve
  efm::DevBuf tucol, tnu, tlidx, tdloc;  // row-tile format of the LDS-tiled SpMV
  // per-system correction C_s of frequency-dependent media (efm_sweep_set_correction):
  // c_rows rows of a fixed sub-pattern, values [c_nnz][c_nsys] in caller memory
  int64_t c_rows = 0, c_nnz = 0;
  int c_nsys = 0, c_ncols = 0, gx_corr = 8, ncu = 256;
  int c_terms = 0;  // 1: c_vals holds shared terms [entry][8] and c_alpha [system][8] their weights
  int c_nterms = kMaxTerms;  // the shared terms in use (columns past them are zero)
  const double2 *c_vals = nullptr;
  efm::DevBuf c_row, c_ptr, c_col, c_map, part_c, tcorr, c_colmap, c_alpha;
  ~efm_sweep() {
    if (capture_stream)
      (void)hipStreamDestroy(capture_stream);
  }
};

namespace {
// Grid of a persistent kernel: as many blocks of kBlock threads as are resident on the
// device at the kernel's occupancy (a multiple of 8 for the XCD-aware walks), capped by
// the work (items handled one per wave).
int resident_grid(const void *kernel, int ncu, int64_t items, int block = kBlock) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1)
    nb = 1;
  const int64_t need = (items + kWaves - 1) / kWaves;
  const int64_t cap = std::max<int64_t>(8, ((int64_t)ncu * nb / 8) * 8);
  const int64_t g = std::min<int64_t>(cap, std::max<int64_t>(need, 8));
  return (int)(((g + 7) / 8) * 8);
}
}  // namespace

extern "C" int efm_sweep_create(const efm_topology *topo, const double *K_val, const double *M_val,
                                int n_ports, const double *const *h_Ms_vals,
                                const uint8_t *dirichlet, int max_systems, void *stream,
                                efm_sweep **out) {
  EFM_REQUIRE(out, "efm_sweep_create: out is NULL");
  *out = nullptr;
  EFM_REQUIRE(topo && K_val && M_val, "efm_sweep_create: NULL argument");
  EFM_REQUIRE(n_ports >= 0 && n_ports <= kMaxPorts, "efm_sweep_create: 0..4 ports");
  EFM_REQUIRE(n_ports == 0 || h_Ms_vals, "efm_sweep_create: port masses missing");
  EFM_REQUIRE(max_systems >= 1 && max_systems <= 1024, "efm_sweep_create: 1..1024 systems");
  hipStream_t st = (hipStream_t)stream;
  auto s = std::make_unique<efm_sweep>();
  const int64_t n = topo->n_edges, nnz = topo->nnz, N = topo->N;
  s->n = n;
  s->nnz = nnz;
  s->N = N;
  s->P = n_ports;
  s->Bmax = ((max_systems + kSys - 1) / kSys) * kSys;
  s->row_ptr = topo->row_ptr.as<int32_t>();
  s->col = topo->col_idx.as<int32_t>();
  s->n2e_ptr = topo->n2e_ptr.as<int32_t>();
  s->n2e = topo->n2e.as<int32_t>();
  int rc;
  const size_t vb = (size_t)n * s->Bmax * sizeof(double2);
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      ncu = prop.multiProcessorCount;
  }
  s->ncu = ncu;
  const int groups = s->Bmax / kSys;
  // blocks per system group: enough waves to fill the chip, a multiple of 8 (XCD walk)
  auto gx_for = [&](int64_t items) {
    int64_t want = std::max<int64_t>(8, (int64_t)ncu * 8 / groups);
    int64_t need = (items + kWaves - 1) / kWaves;
    int64_t g = std::min(want, std::max<int64_t>(need, 1));
    return (int)(((g + 7) / 8) * 8);
  };
  s->gx_vec = gx_for(n);
  s->gx_nodes = gx_for(N);
  // SpMV grid: gx_spmv blocks per group slice, gy_spmv slices striding over the groups
  {
    const int64_t need = (n + kWaves - 1) / kWaves;
    const int64_t want = (int64_t)ncu * 2;
    s->gy_spmv = groups;
    const int64_t g = std::min<int64_t>(want, std::max<int64_t>(need, 8));
    s->gx_spmv = (int)(((g + 7) / 8) * 8);
    // pipelined kernels and the update: one block slice (gridDim.y = 1), as many blocks
    // as are resident
    s->gx_pre = resident_grid(reinterpret_cast<const void *>(&k_sw_spmv_pre<double2, true, true>), ncu, n);
    s->gx_aux = resident_grid(reinterpret_cast<const void *>(&k_sw_aux_pre<double2, 3>), ncu, N);
    s->gx_tile = resident_grid(reinterpret_cast<const void *>(&k_sw_spmv_tile<double2, true, true>), ncu,
                               ((n + kTileRows - 1) / kTileRows) * kWaves, kTileBlock);
    const void *upd[2][3] = {
        {reinterpret_cast<const void *>(&k_sw_update<double2, true, false>),
         reinterpret_cast<const void *>(&k_sw_update<double2, false, false>),
         reinterpret_cast<const void *>(&k_sw_update<double2, false, true>)},
        {reinterpret_cast<const void *>(&k_sw_update<float2, true, false>),
         reinterpret_cast<const void *>(&k_sw_update<float2, false, false>),
         reinterpret_cast<const void *>(&k_sw_update<float2, false, true>)}};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b)
        s->gx_upd[a][b] = resident_grid(upd[a][b], ncu, n);
    s->gx_std[0] = resident_grid(reinterpret_cast<const void *>(&k_sw_update_std<double2, true>), ncu, n);
    s->gx_std[1] = resident_grid(reinterpret_cast<const void *>(&k_sw_update_std<float2, true>), ncu, n);
  }
  int gx_upd_max = 8;
  for (int a = 0; a < 2; ++a) {
    for (int b = 0; b < 3; ++b)
      gx_upd_max = std::max(gx_upd_max, s->gx_upd[a][b]);
    gx_upd_max = std::max(gx_upd_max, s->gx_std[a]);
  }
  const int64_t ntl = (n + kTileRows - 1) / kTileRows;
  // K, M, Ms and the tile entry slots carry one zero entry at index nnz
  if ((rc = s->K.alloc((nnz + 1) * 16)) || (rc = s->M.alloc((nnz + 1) * 16)) ||
      (rc = s->Ms.alloc((nnz + 1) * 8)) || (rc = s->tucol.alloc(ntl * kTileCols * sizeof(int32_t))) ||
      (rc = s->tnu.alloc(ntl * sizeof(int32_t))) || (rc = s->tlidx.alloc(nnz + 1)) ||
      (rc = s->tdloc.alloc(n)) ||
      (rc = s->rport.alloc(n + 16)) || (rc = s->dK.alloc(n * 16)) || (rc = s->dM.alloc(n * 16)) ||
      (rc = s->dMs.alloc(n * 8)) || (rc = s->dnK.alloc(N * 16)) || (rc = s->dnM.alloc(N * 16)) ||
      (rc = s->dnMs.alloc((size_t)N * 8 * std::max(1, n_ports))) || (rc = s->gsign.alloc(2 * n + 16)) ||
      (rc = s->en_eff.alloc(n * sizeof(int2))) || (rc = s->x.alloc(vb)) || (rc = s->r.alloc(vb)) ||
      (rc = s->p.alloc(vb)) || (rc = s->sv.alloc(vb)) || (rc = s->u.alloc(vb)) || (rc = s->w.alloc(vb)) ||
      (rc = s->y.alloc((size_t)N * s->Bmax * sizeof(double2))) ||
      (rc = s->dninv.alloc((size_t)N * s->Bmax * sizeof(double2))) ||
      (rc = s->part_u.alloc((size_t)groups * gx_upd_max * 2 * 64 * sizeof(double2))) ||
      (rc = s->part_s.alloc((size_t)groups * std::max(std::max(s->gx_spmv, s->gx_pre), s->gx_tile) * 64 *
                            sizeof(double2))) ||
      (rc = s->state.alloc(s->Bmax * sizeof(SwState))) || (rc = s->flags.alloc(s->Bmax * sizeof(int))) ||
      (rc = s->k0sq.alloc(s->Bmax * sizeof(double))) ||
      (rc = s->coef.alloc((size_t)s->Bmax * std::max(1, n_ports) * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(s->K.p, K_val, nnz * 16, hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->M.p, M_val, nnz * 16, hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemsetAsync(s->K.as<double2>() + nnz, 0, 16, st));
  EFM_HIP_CHECK(hipMemsetAsync(s->M.as<double2>() + nnz, 0, 16, st));
  EFM_HIP_CHECK(hipMemsetAsync(s->Ms.as<double>() + nnz, 0, 8, st));
  EFM_HIP_CHECK(hipMemsetAsync(s->tlidx.as<uint8_t>() + nnz, 0, 1, st));
  // port arrays: device pointer table
  DevBuf ptab, err;
  if ((rc = ptab.alloc(sizeof(double *) * kMaxPorts)) || (rc = err.alloc(sizeof(int))))
    return rc;
  const double *hp[kMaxPorts] = {nullptr, nullptr, nullptr, nullptr};
  for (int p = 0; p < n_ports; ++p)
    hp[p] = h_Ms_vals[p];
  EFM_HIP_CHECK(hipMemcpyAsync(ptab.p, hp, sizeof(hp), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_sw_ports, dim3(grid_for(n, 256)), dim3(256), 0, st, n, n_ports, s->row_ptr,
                     (const double *const *)ptab.p, s->Ms.as<double>(), s->rport.as<int8_t>(), err.as<int>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sw_edge_diag, dim3(grid_for(n, 256)), dim3(256), 0, st, n, s->row_ptr, s->col,
                     s->K.as<double2>(), s->M.as<double2>(), s->Ms.as<double>(), s->dK.as<double2>(),
                     s->dM.as<double2>(), s->dMs.as<double>());
  EFM_LAUNCH_CHECK();
  const int32_t *en = topo->edge_nidx.as<int32_t>();
  hipLaunchKernelGGL(k_sw_node_diag, dim3(grid_for(N, 256)), dim3(256), 0, st, N, n_ports, s->n2e_ptr,
                     s->n2e, en, dirichlet, s->row_ptr, s->col, s->K.as<double2>(), s->M.as<double2>(),
                     (const double *const *)ptab.p, s->dnK.as<double2>(), s->dnM.as<double2>(),
                     s->dnMs.as<double>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sw_gsign, dim3(grid_for(N, 256)), dim3(256), 0, st, N, s->n2e_ptr, s->n2e, en,
                     dirichlet, s->gsign.as<int8_t>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sw_edge_prep, dim3(grid_for(n, 256)), dim3(256), 0, st, n, en, dirichlet,
                     s->en_eff.as<int2>());
  EFM_LAUNCH_CHECK();
  DevBuf cflag;
  if ((rc = cflag.alloc(sizeof(int))))
    return rc;
  EFM_HIP_CHECK(hipMemsetAsync(cflag.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_sw_imag_flags, dim3(std::max<int64_t>(1, std::min<int64_t>(grid_for(nnz, 256), 4096))),
                     dim3(256), 0, st, nnz, s->K.as<double2>(), s->M.as<double2>(), cflag.as<int>());
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipMemcpyAsync(&s->cplx, cflag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  // row tiles of the LDS-tiled SpMV; usable when every tile fits
  hipLaunchKernelGGL(k_sw_build_tiles, dim3((unsigned)ntl), dim3(kBlock), 0, st, n, s->row_ptr, s->col,
                     s->tucol.as<int32_t>(), s->tnu.as<int32_t>(), s->tlidx.as<uint8_t>(), s->tdloc.as<uint8_t>());
  EFM_LAUNCH_CHECK();
  std::vector<int32_t> h_nu(ntl);
  EFM_HIP_CHECK(hipMemcpyAsync(h_nu.data(), s->tnu.p, ntl * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  int h_err = 0;
  EFM_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  EFM_REQUIRE(h_err == 0, "efm_sweep_create: two ports share an edge row (not supported)");
  s->tiled_ok = std::all_of(h_nu.begin(), h_nu.end(), [](int32_t v) { return v >= 0; });
  *out = s.release();
  return EFM_OK;
}

extern "C" void efm_sweep_destroy(efm_sweep *s) { delete s; }

extern "C" int efm_sweep_set_variants(efm_sweep *s, int spmv, int aux) {
  EFM_REQUIRE(s, "efm_sweep_set_variants: NULL sweep");
  EFM_REQUIRE(spmv >= -1 && spmv <= 2 && aux >= -1 && aux <= 1,
              "efm_sweep_set_variants: SpMV variant must be -1..2, nodal variant -1..1");
  EFM_REQUIRE(spmv != 2 || s->tiled_ok,
              "efm_sweep_set_variants: the LDS-tiled SpMV (2) does not fit this matrix's row tiles");
  s->spmv_variant = spmv;
  s->aux_variant = aux;
  return EFM_OK;
}

extern "C" int efm_sweep_set_graph(efm_sweep *s, int enable) {
  EFM_REQUIRE(s, "efm_sweep_set_graph: NULL sweep");
  s->use_graph = enable ? 1 : 0;
  return EFM_OK;
}

extern "C" int efm_sweep_set_correction(efm_sweep *s, int64_t n_rows, const int32_t *h_rows,
                                        const int32_t *h_ptr, const int32_t *h_cols, void *stream) {
  EFM_REQUIRE(s, "efm_sweep_set_correction: NULL sweep");
  EFM_REQUIRE(n_rows >= 0 && n_rows <= s->n, "efm_sweep_set_correction: n_rows out of range");
  s->c_rows = 0;
  s->c_nnz = 0;
  s->c_vals = nullptr;
  s->c_nsys = 0;
  if (n_rows == 0)
    return EFM_OK;
  EFM_REQUIRE(h_rows && h_ptr && h_cols, "efm_sweep_set_correction: NULL argument");
  EFM_REQUIRE(h_ptr[0] == 0, "efm_sweep_set_correction: ptr[0] must be 0");
  std::vector<int32_t> map(s->n, -1);
  for (int64_t m = 0; m < n_rows; ++m) {
    EFM_REQUIRE(h_ptr[m + 1] >= h_ptr[m], "efm_sweep_set_correction: ptr not monotone");
    EFM_REQUIRE(h_rows[m] >= 0 && h_rows[m] < s->n, "efm_sweep_set_correction: row out of range");
    EFM_REQUIRE(map[h_rows[m]] < 0, "efm_sweep_set_correction: a row appears twice");
    map[h_rows[m]] = (int32_t)m;
  }
  const int64_t nz = h_ptr[n_rows];
  for (int64_t k = 0; k < nz; ++k)
    EFM_REQUIRE(h_cols[k] >= 0 && h_cols[k] < s->n, "efm_sweep_set_correction: column out of range");
  hipStream_t st = (hipStream_t)stream;
  const int groups = s->Bmax / kSys;
  const int64_t need = (n_rows + kWaves - 1) / kWaves;
  const int64_t want = std::max<int64_t>(8, (int64_t)s->ncu * 8 / groups);
  const int gx = (int)std::min(want, std::max<int64_t>(need, 1));
  int rc;
  if ((rc = s->c_row.alloc(n_rows * sizeof(int32_t))) || (rc = s->c_ptr.alloc((n_rows + 1) * sizeof(int32_t))) ||
      (rc = s->c_col.alloc(std::max<int64_t>(nz, 1) * sizeof(int32_t))) ||
      (rc = s->c_map.alloc(s->n * sizeof(int32_t))) ||
      (rc = s->part_c.alloc((size_t)groups * gx * 64 * sizeof(double2))) ||
      (rc = s->tcorr.alloc((size_t)n_rows * s->Bmax * sizeof(double2))) ||
      (rc = s->c_colmap.alloc((size_t)s->Bmax * sizeof(int32_t))) ||
      (rc = s->c_alpha.alloc((size_t)s->Bmax * kMaxTerms * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(s->c_row.p, h_rows, n_rows * sizeof(int32_t), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->c_ptr.p, h_ptr, (n_rows + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (nz > 0)
    EFM_HIP_CHECK(hipMemcpyAsync(s->c_col.p, h_cols, nz * sizeof(int32_t), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->c_map.p, map.data(), s->n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  s->c_rows = n_rows;
  s->c_nnz = nz;
  s->gx_corr = gx;
  return EFM_OK;
}

extern "C" int efm_sweep_set_correction_values(efm_sweep *s, int nsys, int ncols, const double *d_vals,
                                               const int32_t *h_col_of_system, void *stream) {
  EFM_REQUIRE(s, "efm_sweep_set_correction_values: NULL sweep");
  EFM_REQUIRE(s->c_rows > 0, "efm_sweep_set_correction_values: no correction pattern set");
  EFM_REQUIRE(nsys >= 1 && nsys <= s->Bmax, "efm_sweep_set_correction_values: nsys out of range");
  EFM_REQUIRE(ncols >= 1, "efm_sweep_set_correction_values: ncols < 1");
  EFM_REQUIRE(d_vals || s->c_nnz == 0, "efm_sweep_set_correction_values: NULL values");
  EFM_REQUIRE(h_col_of_system, "efm_sweep_set_correction_values: NULL column map");
  for (int i = 0; i < nsys; ++i)
    EFM_REQUIRE(h_col_of_system[i] >= 0 && h_col_of_system[i] < ncols,
                "efm_sweep_set_correction_values: column of a system out of range");
  hipStream_t st = (hipStream_t)stream;
  EFM_HIP_CHECK(hipMemcpyAsync(s->c_colmap.p, h_col_of_system, nsys * sizeof(int32_t), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  s->c_vals = (const double2 *)d_vals;
  s->c_nsys = nsys;
  s->c_ncols = ncols;
  s->c_terms = 0;
  return EFM_OK;
}

extern "C" int efm_sweep_set_correction_terms(efm_sweep *s, int nterms, const double *d_terms, int nsys,
                                              const double *h_alpha, void *stream) {
  EFM_REQUIRE(s, "efm_sweep_set_correction_terms: NULL sweep");
  EFM_REQUIRE(s->c_rows > 0, "efm_sweep_set_correction_terms: no correction pattern set");
  EFM_REQUIRE(nterms >= 1 && nterms <= kMaxTerms, "efm_sweep_set_correction_terms: 1..8 terms");
  EFM_REQUIRE(nsys >= 1 && nsys <= s->Bmax, "efm_sweep_set_correction_terms: nsys out of range");
  EFM_REQUIRE((d_terms || s->c_nnz == 0) && h_alpha, "efm_sweep_set_correction_terms: NULL argument");
  std::vector<double2> al((size_t)nsys * kMaxTerms, make_double2(0, 0));
  for (int i = 0; i < nsys; ++i)
    for (int q = 0; q < nterms; ++q)
      al[(size_t)i * kMaxTerms + q] = make_double2(h_alpha[2 * ((size_t)i * nterms + q)],
                                                   h_alpha[2 * ((size_t)i * nterms + q) + 1]);
  hipStream_t st = (hipStream_t)stream;
  EFM_HIP_CHECK(hipMemcpyAsync(s->c_alpha.p, al.data(), al.size() * sizeof(double2), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  s->c_vals = (const double2 *)d_terms;
  s->c_nsys = nsys;
  s->c_ncols = kMaxTerms;
  s->c_terms = 1;
  s->c_nterms = nterms;
  return EFM_OK;
}

extern "C" int efm_sweep_set_iteration(efm_sweep *s, int form) {
  EFM_REQUIRE(s, "efm_sweep_set_iteration: NULL sweep");
  EFM_REQUIRE(form >= 0 && form <= 2,
              "efm_sweep_set_iteration: form must be 0 (standard), 1 (Chronopoulos-Gear) or 2 (standard, u rebuilt "
              "in the p pass)");
  s->form = form;
  return EFM_OK;
}

extern "C" int efm_sweep_get_variants(const efm_sweep *s, int *h_spmv, int *h_aux, double *h_tune_ms) {
  EFM_REQUIRE(s && h_spmv && h_aux, "efm_sweep_get_variants: NULL argument");
  *h_spmv = s->spmv_variant;
  *h_aux = s->aux_variant;
  if (h_tune_ms)
    for (int v = 0; v < 5; ++v)
      h_tune_ms[v] = s->tune_ms[v];
  return EFM_OK;
}

namespace {

// Event pairs around the launches of kernel variants; pick() keeps the variant with the
// smallest mean time (stream synchronized before) and writes the means (0 for a
// variant without launches).
struct TuneLog {
  std::vector<hipEvent_t> e0, e1;
  std::vector<int> var;
  void add(hipEvent_t a, hipEvent_t b, int v) {
    e0.push_back(a);
    e1.push_back(b);
    var.push_back(v);
  }
  bool empty() const { return var.empty(); }
  int pick(int *variant, float *ms_out, int nv) {  // variants 0 .. nv-1 (nv <= 4)
    float sum[4] = {0.f, 0.f, 0.f, 0.f};
    int cnt[4] = {0, 0, 0, 0};
    int rc = EFM_OK;
    for (size_t i = 0; i < var.size(); ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0[i], e1[i]) != hipSuccess) {
        rc = EFM_ERR_HIP;
        break;
      }
      sum[var[i]] += ms;
      cnt[var[i]] += 1;
    }
    if (rc == EFM_OK) {
      int best = -1;
      for (int v = 0; v < nv; ++v) {
        ms_out[v] = cnt[v] ? sum[v] / cnt[v] : 0.f;
        if (cnt[v] && (best < 0 || sum[v] / cnt[v] < sum[best] / cnt[best]))
          best = v;
      }
      if (best >= 0)
        *variant = best;
    }
    release();
    return rc;
  }
  void release() {
    for (size_t i = 0; i < var.size(); ++i) {
      (void)hipEventDestroy(e0[i]);
      (void)hipEventDestroy(e1[i]);
    }
    e0.clear();
    e1.clear();
    var.clear();
  }
  ~TuneLog() { release(); }
};

// A captured chunk of iterations: the graph and its executable, released together.
struct ChunkGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  ~ChunkGraph() {
    if (x)
      (void)hipGraphExecDestroy(x);
    if (g)
      (void)hipGraphDestroy(g);
  }
};

// Ends a stream capture that an error left open (the partial graph is dropped).
struct CaptureGuard {
  hipStream_t st = nullptr;
  bool on = false;
  ~CaptureGuard() {
    if (on) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g)
        (void)hipGraphDestroy(g);
    }
  }
};

// One launch sequence over the padded system block [Bp]: everything a solve needs.
struct SwRun {
  efm_sweep *s;
  hipStream_t st;
  int Bp, groups, P, check;
  double tol2;
  int max_it;
  SwState *S;
  int *flags;
  double2 *pu, *ps;
  std::vector<int> hflags;

  int n_active(int &count) {
    int want = 0;
    return n_active(count, want);
  }
  // flags: bit 0 active, bit 1 a residual replacement is due (reliable updates)
  int n_active(int &count, int &want) {
    EFM_HIP_CHECK(hipMemcpyAsync(hflags.data(), flags, Bp * sizeof(int), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    count = 0;
    want = 0;
    for (int v : hflags) {
      count += (v & 1) != 0;
      want += (v & 2) != 0;
    }
    return EFM_OK;
  }

  // w = A_s u with SpMV variant `variant` (partials of u^T w in part_s)
  template <typename VT> int spmv_launch(int variant, const VT *u, VT *w) {
    const int64_t n = s->n;
    if (variant == 2) {
      const dim3 grid(s->gx_tile, 1);
      const double2 *Kv = s->K.as<double2>(), *Mv = s->M.as<double2>(), *cf = s->coef.as<double2>();
      const double *Msv = s->Ms.as<double>(), *k2 = s->k0sq.as<double>();
      const int8_t *rp = s->rport.as<int8_t>();
      const int32_t *uc = s->tucol.as<int32_t>(), *tn = s->tnu.as<int32_t>();
      const uint8_t *li = s->tlidx.as<uint8_t>(), *dl = s->tdloc.as<uint8_t>();
      switch (s->cplx & 3) {
      case 0:
        hipLaunchKernelGGL((k_sw_spmv_tile<VT, false, false>), grid, dim3(kTileBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, uc, tn, li, dl, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      case 1:
        hipLaunchKernelGGL((k_sw_spmv_tile<VT, true, false>), grid, dim3(kTileBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, uc, tn, li, dl, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      case 2:
        hipLaunchKernelGGL((k_sw_spmv_tile<VT, false, true>), grid, dim3(kTileBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, uc, tn, li, dl, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      default:
        hipLaunchKernelGGL((k_sw_spmv_tile<VT, true, true>), grid, dim3(kTileBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, uc, tn, li, dl, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
      }
    } else if (variant == 1) {
      const dim3 grid(s->gx_pre, 1);
      const double2 *Kv = s->K.as<double2>(), *Mv = s->M.as<double2>(), *cf = s->coef.as<double2>();
      const double *Msv = s->Ms.as<double>(), *k2 = s->k0sq.as<double>();
      const int8_t *rp = s->rport.as<int8_t>();
      switch (s->cplx & 3) {
      case 0:
        hipLaunchKernelGGL((k_sw_spmv_pre<VT, false, false>), grid, dim3(kBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, s->col, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      case 1:
        hipLaunchKernelGGL((k_sw_spmv_pre<VT, true, false>), grid, dim3(kBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, s->col, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      case 2:
        hipLaunchKernelGGL((k_sw_spmv_pre<VT, false, true>), grid, dim3(kBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, s->col, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
        break;
      default:
        hipLaunchKernelGGL((k_sw_spmv_pre<VT, true, true>), grid, dim3(kBlock), 0, st, n, Bp, P, groups,
                           s->row_ptr, s->col, Kv, Mv, Msv, rp, k2, cf, u, w, ps, S);
      }
    } else
      hipLaunchKernelGGL((k_sw_spmv<VT>), dim3(s->gx_spmv, s->gy_spmv), dim3(kBlock), 0, st, n, Bp, P, groups,
                         s->row_ptr, s->col, s->K.as<double2>(), s->M.as<double2>(), s->Ms.as<double>(),
                         s->rport.as<int8_t>(), s->k0sq.as<double>(), s->coef.as<double2>(), u, w, ps, S);
    EFM_LAUNCH_CHECK();
    if (s->c_rows > 0) {
      auto kc = !s->c_terms           ? k_sw_corr<VT, VT, 0, false>
                : s->c_nterms <= 4 ? k_sw_corr<VT, VT, 0, true, 4>
                                   : k_sw_corr<VT, VT, 0, true>;
      hipLaunchKernelGGL(kc, dim3(s->gx_corr, groups), dim3(kBlock), 0, st, s->c_rows, Bp, s->c_nsys,
                         s->c_row.as<int32_t>(), s->c_ptr.as<int32_t>(), s->c_col.as<int32_t>(), s->c_vals,
                         s->c_ncols, s->c_colmap.as<int32_t>(), s->c_vals, s->c_alpha.as<double2>(), u, w,
                         (double2 *)nullptr, s->part_c.as<double2>(), S);
      EFM_LAUNCH_CHECK();
    }
    return EFM_OK;
  }

  // y = Dn^{-1} G^T v with nodal variant `variant`, NV gathered vectors (see k_sw_aux_node)
  template <typename VT, int NV> int aux_launch(int variant, const VT *r, const VT *w, const VT *sv, VT *y) {
    const int8_t *gs = s->gsign.as<int8_t>();
    if (variant == 1)
      hipLaunchKernelGGL((k_sw_aux_pre<VT, NV>), dim3(s->gx_aux, 1), dim3(kBlock), 0, st, s->N, Bp, groups,
                         s->n2e_ptr, s->n2e, gs, s->dninv.as<double2>(), r, w, sv, y, S);
    else
      hipLaunchKernelGGL((k_sw_aux_node<VT, NV>), dim3(s->gx_nodes, groups), dim3(kBlock), 0, st, s->N, Bp, P,
                         s->n2e_ptr, s->n2e, gs, s->dnK.as<double2>(), s->dnM.as<double2>(), s->dnMs.as<double>(),
                         s->k0sq.as<double>(), s->coef.as<double2>(), r, w, sv, y, S);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // Preconditioned COCG on the systems whose state is active: x = A^{-1} r0, where r
  // holds r0 on entry (b == nullptr) or is set from b.  Vectors of type VT.
  template <typename VT> int cocg(const double2 *b, double inner_tol2, int max_inner) {
    const int64_t n = s->n;
    VT *x = s->x.as<VT>(), *r = s->r.as<VT>(), *p = s->p.as<VT>();
    VT *sv = s->sv.as<VT>(), *u = s->u.as<VT>(), *w = s->w.as<VT>(), *y = s->y.as<VT>();
    const int2 *en = s->en_eff.as<int2>();
    hipLaunchKernelGGL((k_sw_init<VT>), dim3(grid_for(n * Bp, kBlock)), dim3(kBlock), 0, st, n, Bp, b, x, r, p,
                       sv);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sw_dninv, dim3(grid_for(s->N * Bp, kBlock)), dim3(kBlock), 0, st, s->N, Bp, P,
                       s->dnK.as<double2>(), s->dnM.as<double2>(), s->dnMs.as<double>(), s->k0sq.as<double>(),
                       s->coef.as<double2>(), s->dninv.as<double2>());
    EFM_LAUNCH_CHECK();
    auto aux_v = [&](bool fused, int variant) -> int {
      return fused ? aux_launch<VT, 3>(variant, r, w, sv, y)
                   : aux_launch<VT, 1>(variant, r, (const VT *)nullptr, (const VT *)nullptr, y);
    };
    constexpr int vti = std::is_same<VT, float2>::value ? 1 : 0;
    auto update = [&](bool first, bool apply) -> int {
      const double2 *dK = s->dK.as<double2>(), *dM = s->dM.as<double2>();
      const double *dMs = s->dMs.as<double>(), *k2 = s->k0sq.as<double>();
      const int8_t *rp = s->rport.as<int8_t>();
      const double2 *cf = s->coef.as<double2>();
      const int vi = first ? 0 : (apply ? 2 : 1);
      const dim3 grid(s->gx_upd[vti][vi], 1);
      if (first)
        hipLaunchKernelGGL((k_sw_update<VT, true, false>), grid, dim3(kBlock), 0, st, n, Bp, groups, P, en, dK, dM,
                           dMs, rp, k2, cf, y, w, u, p, sv, x, r, pu, S);
      else if (apply)
        hipLaunchKernelGGL((k_sw_update<VT, false, true>), grid, dim3(kBlock), 0, st, n, Bp, groups, P, en, dK, dM,
                           dMs, rp, k2, cf, y, w, u, p, sv, x, r, pu, S);
      else
        hipLaunchKernelGGL((k_sw_update<VT, false, false>), grid, dim3(kBlock), 0, st, n, Bp, groups, P, en, dK, dM,
                           dMs, rp, k2, cf, y, w, u, p, sv, x, r, pu, S);
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    };
    auto spmv_v = [&](int variant) -> int { return spmv_launch<VT>(variant, u, w); };
    auto finalize = [&](bool first, bool skip, int spmv_variant) -> int {
      // the grids that wrote the partials: the SpMV variant's and the update's
      const int gxs = spmv_variant == 2 ? s->gx_tile : spmv_variant == 1 ? s->gx_pre : s->gx_spmv;
      const int gxu = s->gx_upd[vti][first ? 0 : (skip ? 1 : 2)];
      if (first)
        hipLaunchKernelGGL((k_sw_finalize<true>), dim3(Bp), dim3(kBlock), 0, st, Bp, gxu, gxs, pu, ps, S, flags,
                           inner_tol2, max_it, max_inner, 0);
      else
        hipLaunchKernelGGL((k_sw_finalize<false>), dim3(Bp), dim3(kBlock), 0, st, Bp, gxu, gxs, pu, ps, S, flags,
                           inner_tol2, max_it, max_inner, skip ? 1 : 0);
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    };
    // Autotuning in place: while a kernel variant is unset (-1), the iterations of the
    // first convergence-check chunk cycle through its candidates (SpMV it % nvs, nodal
    // it % 2), each launch timed with events where it runs in the iteration (so with the
    // cache state the iteration leaves); at the chunk's check the fastest mean is kept and
    // the means are reported in tune_ms.  Every candidate computes the same w (or y), so
    // the iterations are those of the chosen kernel up to rounding.  Setup launches use
    // variant 1.
    const int nvs = s->tiled_ok ? 3 : 2;
    TuneLog log_s, log_a;
    auto vs_of = [&](int it) { return s->spmv_variant >= 0 ? s->spmv_variant : (it >= 0 ? it % nvs : 1); };
    auto va_of = [&](int it) { return s->aux_variant >= 0 ? s->aux_variant : (it >= 0 ? it % 2 : 1); };
    auto run = [&](TuneLog &lg, bool timed, int v, auto &&launch) -> int {
      if (!timed)
        return launch();
      hipEvent_t e0, e1;
      EFM_HIP_CHECK(hipEventCreate(&e0));
      if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return EFM_ERR_HIP;
      }
      lg.add(e0, e1, v);
      EFM_HIP_CHECK(hipEventRecord(e0, st));
      int rc2 = launch();
      if (rc2)
        return rc2;
      EFM_HIP_CHECK(hipEventRecord(e1, st));
      return EFM_OK;
    };
    int rc, active = 0;
    if ((rc = aux_v(false, va_of(-1))) || (rc = update(true, false)) || (rc = spmv_v(vs_of(-1))) ||
        (rc = finalize(true, false, vs_of(-1))) || (rc = n_active(active)))
      return rc;
    int it = 0;
    const int cap = std::min(max_it, max_inner);
    while (active > 0 && it < cap) {
      const int chunk = std::min(check, cap - it);
      for (int c = 0; c < chunk; ++c, ++it) {
        const bool apply = (it & 1) != 0;  // x and r are stored on odd iterations
        const int va = va_of(it), vs = vs_of(it);
        if ((rc = run(log_a, s->aux_variant < 0, va, [&]() { return aux_v(true, va); })) ||
            (rc = update(false, apply)) ||
            (rc = run(log_s, s->spmv_variant < 0, vs, [&]() { return spmv_v(vs); })) ||
            (rc = finalize(false, !apply, vs)))
          return rc;
      }
      if ((rc = n_active(active)))  // synchronizes the stream: the events are complete
        return rc;
      if (s->spmv_variant < 0 && !log_s.empty() && (rc = log_s.pick(&s->spmv_variant, s->tune_ms, nvs)))
        return rc;
      if (s->aux_variant < 0 && !log_a.empty() && (rc = log_a.pick(&s->aux_variant, s->tune_ms + 3, 2)))
        return rc;
    }
    hipLaunchKernelGGL((k_sw_flush<VT>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, x, (const VT *)p, S);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // Reliable update (residual replacement) of the systems with `want` set, or of every
  // system (all = 1: the final accumulation, then the true residual of every system into
  // rn_out): x64 += x and x = 0; r = b - A x64 in FP64, stored as VT; y and u = P r with the
  // partials of u^T r and |r|^2; STAGE 3 finalize (convergence on the true residual, else
  // beta = u^T r / rho_prev with the search direction p kept).
  template <typename VT> int replace(const double2 *b, int all, int va, double ru_delta2) {
    const int64_t n = s->n;
    constexpr int vti = std::is_same<VT, float2>::value ? 1 : 0;
    VT *x = s->x.as<VT>(), *r = s->r.as<VT>(), *p = s->p.as<VT>();
    VT *sv = s->sv.as<VT>(), *u = s->u.as<VT>(), *w = s->w.as<VT>(), *y = s->y.as<VT>();
    double2 *x64 = s->x64.as<double2>();
    hipLaunchKernelGGL(k_sw_ru_begin, dim3((Bp + kBlock - 1) / kBlock), dim3(kBlock), 0, st, Bp, S, all);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_sw_accumulate<VT>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, x, x64, S);
    EFM_LAUNCH_CHECK();
    const bool corr = s->c_rows > 0;
    if (corr) {
      auto kc = !s->c_terms           ? k_sw_corr<VT, double2, 1, false>
                : s->c_nterms <= 4 ? k_sw_corr<VT, double2, 1, true, 4>
                                   : k_sw_corr<VT, double2, 1, true>;
      hipLaunchKernelGGL(kc, dim3(s->gx_corr, groups), dim3(kBlock), 0, st, s->c_rows, Bp, s->c_nsys,
                         s->c_row.as<int32_t>(), s->c_ptr.as<int32_t>(), s->c_col.as<int32_t>(), s->c_vals,
                         s->c_ncols, s->c_colmap.as<int32_t>(), s->c_vals, s->c_alpha.as<double2>(),
                         (const double2 *)x64, (VT *)nullptr, s->tcorr.as<double2>(), (double2 *)nullptr, S);
      EFM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_sw_residual<VT>), dim3(s->gx_spmv, s->gy_spmv), dim3(kBlock), 0, st, n, Bp, P, groups,
                       s->row_ptr, s->col, s->K.as<double2>(), s->M.as<double2>(), s->Ms.as<double>(),
                       s->rport.as<int8_t>(), s->k0sq.as<double>(), s->coef.as<double2>(), b,
                       (const double2 *)x64, r, ps, S, corr ? s->c_map.as<int32_t>() : (const int32_t *)nullptr,
                       corr ? (const double2 *)s->tcorr.as<double2>() : (const double2 *)nullptr);
    EFM_LAUNCH_CHECK();
    if (all) {
      hipLaunchKernelGGL(k_sw_true_resid, dim3(Bp), dim3(kBlock), 0, st, Bp, s->gx_spmv, ps, S);
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    }
    int rc;
    if ((rc = aux_launch<VT, 1>(va, r, (const VT *)nullptr, (const VT *)nullptr, y)))
      return rc;
    hipLaunchKernelGGL((k_sw_update<VT, true, false>), dim3(s->gx_upd[vti][0], 1), dim3(kBlock), 0, st, n, Bp,
                       groups, P, s->en_eff.as<int2>(), s->dK.as<double2>(), s->dM.as<double2>(), s->dMs.as<double>(),
                       s->rport.as<int8_t>(), s->k0sq.as<double>(), s->coef.as<double2>(), y, w, u, p, sv, x, r,
                       pu, S);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_sw_fin_std<3>), dim3(Bp), dim3(kBlock), 0, st, Bp, s->gx_upd[vti][0], pu, 0,
                       (const double2 *)nullptr, S, flags, tol2, max_it, max_it, ru_delta2);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // Standard preconditioned COCG (see k_sw_pform): the same contract as cocg().  With
  // ru_delta2 > 0 (reliable updates; b is the FP64 right-hand side and x64 the FP64
  // solution, zero on entry): the Krylov vectors are of type VT, x carries the solution
  // increment since the last replacement, and r is replaced by the FP64 residual
  // b - A x64 whenever a system asks for it (k_sw_fin_std STAGE 2), at the next
  // convergence check; the search direction is kept across a replacement, so the Krylov
  // recurrence is not restarted.  A system converges only on its true residual.
  template <typename VT> int cocg_std(const double2 *b, double inner_tol2, int max_inner, double ru_delta2 = 0.0) {
    const int64_t n = s->n;
    // form 2 without reliable updates (which read x during the loop): x updated one
    // iteration late, in the p pass
    const bool xlag = s->form == 2 && !(ru_delta2 > 0.0);
    VT *x = s->x.as<VT>(), *r = s->r.as<VT>(), *p = s->p.as<VT>();
    VT *sv = s->sv.as<VT>(), *u = s->u.as<VT>(), *w = s->w.as<VT>(), *y = s->y.as<VT>();
    constexpr int vti = std::is_same<VT, float2>::value ? 1 : 0;
    const double2 *dK = s->dK.as<double2>(), *dM = s->dM.as<double2>(), *cf = s->coef.as<double2>();
    const double *dMs = s->dMs.as<double>(), *k2 = s->k0sq.as<double>();
    const int8_t *rp = s->rport.as<int8_t>();
    const int2 *en = s->en_eff.as<int2>();
    hipLaunchKernelGGL((k_sw_init<VT>), dim3(grid_for(n * Bp, kBlock)), dim3(kBlock), 0, st, n, Bp, b, x, r, p,
                       sv);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sw_dninv, dim3(grid_for(s->N * Bp, kBlock)), dim3(kBlock), 0, st, s->N, Bp, P,
                       s->dnK.as<double2>(), s->dnM.as<double2>(), s->dnMs.as<double>(), k2, cf,
                       s->dninv.as<double2>());
    EFM_LAUNCH_CHECK();
    auto fin = [&](auto stage, int gx, const double2 *part) -> int {
      constexpr int STAGE = decltype(stage)::value;
      const bool corr = STAGE == 1 && s->c_rows > 0;
      hipLaunchKernelGGL((k_sw_fin_std<STAGE>), dim3(Bp), dim3(kBlock), 0, st, Bp, gx, part, corr ? s->gx_corr : 0,
                         corr ? (const double2 *)s->part_c.as<double2>() : (const double2 *)nullptr, S, flags,
                         inner_tol2, max_it, max_inner, ru_delta2);
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    };
    auto gxs_of = [&](int v) { return v == 2 ? s->gx_tile : v == 1 ? s->gx_pre : s->gx_spmv; };
    // autotuning in place, as in cocg()
    const int nvs = s->tiled_ok ? 3 : 2;
    TuneLog log_s, log_a;
    auto vs_of = [&](int it) { return s->spmv_variant >= 0 ? s->spmv_variant : it % nvs; };
    auto va_of = [&](int it) { return s->aux_variant >= 0 ? s->aux_variant : (it >= 0 ? it % 2 : 1); };
    auto run = [&](TuneLog &lg, bool timed, int v, auto &&launch) -> int {
      if (!timed)
        return launch();
      hipEvent_t e0, e1;
      EFM_HIP_CHECK(hipEventCreate(&e0));
      if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return EFM_ERR_HIP;
      }
      lg.add(e0, e1, v);
      EFM_HIP_CHECK(hipEventRecord(e0, st));
      int rc2 = launch();
      if (rc2)
        return rc2;
      EFM_HIP_CHECK(hipEventRecord(e1, st));
      return EFM_OK;
    };
    // u = P r, rho = u^T r
    int rc, active = 0;
    if ((rc = aux_launch<VT, 1>(va_of(-1), r, (const VT *)nullptr, (const VT *)nullptr, y)))
      return rc;
    hipLaunchKernelGGL((k_sw_update<VT, true, false>), dim3(s->gx_upd[vti][0], 1), dim3(kBlock), 0, st, n, Bp,
                       groups, P, en, dK, dM, dMs, rp, k2, cf, y, w, u, p, sv, x, r, pu, S);
    EFM_LAUNCH_CHECK();
    int want = 0;
    if ((rc = fin(std::integral_constant<int, 0>{}, s->gx_upd[vti][0], pu)) || (rc = n_active(active, want)))
      return rc;
    int it = 0;
    const int cap = std::min(max_it, max_inner);
    ChunkGraph graph;
    while ((active > 0 || want > 0) && it < cap) {
      if (active == 0) {  // every running system waits for its replacement
        if ((rc = replace<VT>(b, 0, va_of(it), ru_delta2)) || (rc = n_active(active, want)))
          return rc;
        continue;
      }
      const int chunk = std::min(check, cap - it);
      // a full chunk with the kernels chosen: replay the captured graph (capture it first)
      const bool replay = s->use_graph && chunk == check && s->spmv_variant >= 0 && s->aux_variant >= 0;
      if (replay && graph.x) {
        EFM_HIP_CHECK(hipGraphLaunch(graph.x, st));
        it += chunk;
        if ((rc = n_active(active, want)))
          return rc;
        if (want > 0 && ((rc = replace<VT>(b, 0, va_of(it), ru_delta2)) || (rc = n_active(active, want))))
          return rc;
        continue;
      }
      // the chunk is recorded on the sweep's own stream and launched on st
      CaptureGuard capture;
      const hipStream_t run_stream = st;
      if (replay) {
        if (!s->capture_stream)
          EFM_HIP_CHECK(hipStreamCreateWithFlags(&s->capture_stream, hipStreamNonBlocking));
        EFM_HIP_CHECK(hipStreamBeginCapture(s->capture_stream, hipStreamCaptureModeThreadLocal));
        capture.st = s->capture_stream;
        capture.on = true;
        st = s->capture_stream;
      }
      for (int c = 0; c < chunk; ++c, ++it) {
        const int va = va_of(it), vs = vs_of(it);
        if (s->form == 2 && xlag)
          hipLaunchKernelGGL((k_sw_pform_u<VT, true>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, P, en,
                             dK, dM, dMs, rp, k2, cf, (const VT *)y, (const VT *)r, p, x, S);
        else if (s->form == 2)
          hipLaunchKernelGGL((k_sw_pform_u<VT, false>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, P, en,
                             dK, dM, dMs, rp, k2, cf, (const VT *)y, (const VT *)r, p, x, S);
        else
          hipLaunchKernelGGL((k_sw_pform<VT>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, (const VT *)u,
                             p, S);
        EFM_LAUNCH_CHECK();
        if ((rc = run(log_s, s->spmv_variant < 0, vs, [&]() { return spmv_launch<VT>(vs, p, w); })) ||
            (rc = fin(std::integral_constant<int, 1>{}, gxs_of(vs), ps)) ||
            (rc = run(log_a, s->aux_variant < 0, va,
                      [&]() { return aux_launch<VT, 2>(va, r, w, (const VT *)nullptr, y); })))
          return rc;
        auto upd = s->form != 2 ? k_sw_update_std<VT, true, false>
                                : xlag ? k_sw_update_std<VT, false, true> : k_sw_update_std<VT, false, false>;
        hipLaunchKernelGGL(upd, dim3(s->gx_std[vti], 1), dim3(kBlock), 0, st, n, Bp, groups, P, en, dK, dM, dMs,
                           rp, k2, cf, (const VT *)y, (const VT *)w, (const VT *)p, u, x, r, pu, S);
        EFM_LAUNCH_CHECK();
        if ((rc = fin(std::integral_constant<int, 2>{}, s->gx_std[vti], pu)))
          return rc;
      }
      if (replay) {  // the chunk was captured, not run: instantiate and launch it
        capture.on = false;
        st = run_stream;
        EFM_HIP_CHECK(hipStreamEndCapture(s->capture_stream, &graph.g));
        EFM_HIP_CHECK(hipGraphInstantiate(&graph.x, graph.g, nullptr, nullptr, 0));
        EFM_HIP_CHECK(hipGraphLaunch(graph.x, st));
      }
      if ((rc = n_active(active, want)))  // synchronizes the stream: the events are complete
        return rc;
      if (s->spmv_variant < 0 && !log_s.empty() && (rc = log_s.pick(&s->spmv_variant, s->tune_ms, nvs)))
        return rc;
      if (s->aux_variant < 0 && !log_a.empty() && (rc = log_a.pick(&s->aux_variant, s->tune_ms + 3, 2)))
        return rc;
      if (want > 0 && ((rc = replace<VT>(b, 0, va_of(it), ru_delta2)) || (rc = n_active(active, want))))
        return rc;
    }
    if (xlag) {  // the last iteration's x updates
      hipLaunchKernelGGL((k_sw_xflush<VT>), dim3(s->gx_vec, groups), dim3(kBlock), 0, st, n, Bp, (const VT *)p, x,
                         S);
      EFM_LAUNCH_CHECK();
    }
    if (ru_delta2 > 0.0)
      return replace<VT>(b, 1, va_of(it), ru_delta2);
    return EFM_OK;
  }

  template <typename VT> int solve_inner(const double2 *b, double inner_tol2, int max_inner) {
    return s->form == 1 ? cocg<VT>(b, inner_tol2, max_inner) : cocg_std<VT>(b, inner_tol2, max_inner);
  }

};

}  // namespace

extern "C" int efm_sweep_solve(efm_sweep *s, int nsys, const double *h_k0sq, const double *h_coef,
                               const double *B, double *X, const efm_solve_opts *h_opts,
                               const efm_refine_opts *h_refine, int32_t *h_iters, double *h_resid,
                               void *stream) {
  EFM_REQUIRE(s && h_k0sq && B && X && h_opts, "efm_sweep_solve: NULL argument");
  EFM_REQUIRE(nsys >= 1 && nsys <= s->Bmax, "efm_sweep_solve: nsys out of range");
  EFM_REQUIRE(s->P == 0 || h_coef, "efm_sweep_solve: port coefficients missing");
  EFM_REQUIRE(h_opts->method == EFM_METHOD_COCG, "efm_sweep_solve: the sweep uses COCG");
  EFM_REQUIRE(!h_refine || h_refine->inner_type == EFM_C128 || h_refine->inner_type == EFM_C64,
              "efm_sweep_solve: inner_type must be EFM_C128 or EFM_C64");
  EFM_REQUIRE(!h_refine || (h_refine->replace_factor > 0 && h_refine->replace_factor < 1),
              "efm_sweep_solve: replace_factor must lie in (0, 1)");
  EFM_REQUIRE(!h_refine || s->form != 1,
              "efm_sweep_solve: reliable updates need the standard COCG form (0 or 2)");
  EFM_REQUIRE(s->c_rows == 0 || s->form != 1,
              "efm_sweep_solve: a per-system correction needs the standard COCG form (0 or 2)");
  EFM_REQUIRE(s->c_rows == 0 || ((s->c_vals || s->c_nnz == 0) && s->c_nsys == nsys),
              "efm_sweep_solve: correction values missing or set for another number of systems");
  hipStream_t st = (hipStream_t)stream;
  const int Bp = ((nsys + kSys - 1) / kSys) * kSys;  // padded system count this call
  const int P = s->P;
  // per-system scalars (padding systems get zero right-hand sides and stay inactive)
  std::vector<double> k2(Bp, 0.0);
  std::vector<double2> cf((size_t)Bp * std::max(1, P), make_double2(0, 0));
  for (int i = 0; i < nsys; ++i) {
    k2[i] = h_k0sq[i];
    for (int p = 0; p < P; ++p)
      cf[(size_t)i * P + p] = make_double2(h_coef[2 * ((size_t)i * P + p)], h_coef[2 * ((size_t)i * P + p) + 1]);
  }
  EFM_HIP_CHECK(hipMemcpyAsync(s->k0sq.p, k2.data(), Bp * sizeof(double), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->coef.p, cf.data(), cf.size() * sizeof(double2), hipMemcpyHostToDevice, st));
  // right-hand side with padding columns set to zero: rows of B are nsys wide
  DevBuf Bpad;
  const double2 *Bd = (const double2 *)B;
  if (Bp != nsys) {
    int rc = Bpad.alloc((size_t)s->n * Bp * sizeof(double2));
    if (rc != EFM_OK)
      return rc;
    EFM_HIP_CHECK(hipMemsetAsync(Bpad.p, 0, (size_t)s->n * Bp * sizeof(double2), st));
    EFM_HIP_CHECK(hipMemcpy2DAsync(Bpad.p, Bp * sizeof(double2), B, nsys * sizeof(double2),
                                   nsys * sizeof(double2), s->n, hipMemcpyDeviceToDevice, st));
    Bd = Bpad.as<double2>();
  }
  SwRun run;
  run.s = s;
  run.st = st;
  run.Bp = Bp;
  run.groups = Bp / kSys;
  run.P = P;
  run.check = h_opts->check_every > 0 ? h_opts->check_every : 16;
  run.tol2 = h_opts->tolerance * h_opts->tolerance;
  run.max_it = h_opts->max_iterations;
  run.S = s->state.as<SwState>();
  run.flags = s->flags.as<int>();
  run.pu = s->part_u.as<double2>();
  run.ps = s->part_s.as<double2>();
  run.hflags.resize(Bp);
  // every system starts active; the initial residual norms deactivate zero right-hand sides
  {
    std::vector<SwState> init(Bp);
    for (auto &v : init) {
      v = SwState{};
      v.active = 1;
      v.out_active = 1;
    }
    EFM_HIP_CHECK(hipMemcpyAsync(run.S, init.data(), Bp * sizeof(SwState), hipMemcpyHostToDevice, st));
  }
  int rc;
  const double2 *xsrc;
  if (!h_refine) {
    if ((rc = run.solve_inner<double2>(Bd, run.tol2, run.max_it)))
      return rc;
    xsrc = s->x.as<double2>();
  } else {
    if (!s->x64.p && (rc = s->x64.alloc((size_t)s->n * s->Bmax * sizeof(double2))))
      return rc;
    EFM_HIP_CHECK(hipMemsetAsync(s->x64.p, 0, (size_t)s->n * Bp * sizeof(double2), st));
    const double d2 = h_refine->replace_factor * h_refine->replace_factor;
    rc = (h_refine->inner_type == EFM_C64) ? run.cocg_std<float2>(Bd, run.tol2, run.max_it, d2)
                                           : run.cocg_std<double2>(Bd, run.tol2, run.max_it, d2);
    if (rc)
      return rc;
    xsrc = s->x64.as<double2>();
  }
  // solution: the first nsys columns
  if (Bp == nsys) {
    EFM_HIP_CHECK(hipMemcpyAsync(X, xsrc, (size_t)s->n * Bp * sizeof(double2), hipMemcpyDeviceToDevice, st));
  } else {
    EFM_HIP_CHECK(hipMemcpy2DAsync(X, nsys * sizeof(double2), xsrc, Bp * sizeof(double2),
                                   nsys * sizeof(double2), s->n, hipMemcpyDeviceToDevice, st));
  }
  std::vector<SwState> hs(Bp);
  EFM_HIP_CHECK(hipMemcpyAsync(hs.data(), run.S, Bp * sizeof(SwState), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  for (int i = 0; i < nsys; ++i) {
    if (h_iters)
      h_iters[i] = hs[i].total_iters;
    if (h_resid) {
      if (h_refine)
        h_resid[i] = hs[i].bn_out > 0 ? sqrt(hs[i].rn_out / hs[i].bn_out) : 0.0;
      else
        h_resid[i] = hs[i].bnorm2 > 0 ? sqrt(hs[i].rnorm2 / hs[i].bnorm2) : 0.0;
    }
  }
  return EFM_OK;
}
