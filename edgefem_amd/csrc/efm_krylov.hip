// Jacobi-preconditioned Krylov solvers on a CSR matrix for a block of independent
// right-hand sides (replaces solve_linear, src/solver.cpp:35-193).
//
//   COCG      complex-symmetric A (A^T = A): the eigenmode/K-M/rank-1 systems.
//   BiCGSTAB  general A (Bloch-periodic systems lose symmetry).
//
// Vectors and all reductions are FP64; matrix values are stored complex FP64 or
// complex FP32 (promoted on load).  Each iteration of COCG is two fused kernels:
//
//   spmv_p  : p_new = z + beta p_old (on the fly for every gathered column),
//             q = A p_new, partial dots p_new^T q -> alpha (last block reduces)
//   update  : x += alpha p, r -= alpha q, z = D^{-1} r, partial dots r^T z, r^H r
//             -> beta, convergence flag (last block reduces)
//
// A fixed grid with grid-stride loops keeps the number of partial sums small; the
// last block to finish (atomic ticket) reduces them in a fixed order, so results are
// deterministic.  Converged systems are masked: their vectors stop changing and the
// host polls a single counter every `check_every` iterations.
#include "efm_internal.h"

#include <type_traits>
#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kLanesPerRow = 8;  // lanes cooperating on one CSR row
constexpr int kRowsPerBlock = kBlock / kLanesPerRow;
constexpr int kMaxRhs = 8;

struct SysState {
  double2 rho;      // r^T z (COCG) / r0^T r (BiCGSTAB)
  double2 alpha;
  double2 beta;
  double2 omega;    // BiCGSTAB
  double bnorm2;
  double rnorm2;
  int active;       // 1 while iterating
  int iters;
  int breakdown;
  int pad;
};

struct Control {
  unsigned int ticket;   // last-block election
  int n_active;          // systems still iterating
};

__device__ inline double2 bil(double2 a, double2 b) {  // a*b (no conjugate)
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Block-wide sum of up to kMaxRhs complex values per thread into lds, result in lds[0..R).
template <int R>
__device__ inline void block_sum(double2 (&v)[R], double2 *lds) {
  // wavefront reduction with shuffles, then across the 4 waves through LDS
#pragma unroll
  for (int s = 0; s < R; ++s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v[s].x += __shfl_xor(v[s].x, off, 64);
      v[s].y += __shfl_xor(v[s].y, off, 64);
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < R; ++s)
      lds[wave * R + s] = v[s];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      double2 a = lds[s];
      for (int w = 1; w < kBlock / 64; ++w)
        a = cadd(a, lds[w * R + s]);
      lds[s] = a;
    }
  }
  __syncthreads();
}

// Last-block election; returns true in every thread of the final block.
__device__ inline bool last_block(unsigned int *ticket) {
  __shared__ bool am_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int prev = atomicAdd(ticket, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (am_last)
    __threadfence();
  return am_last;
}

// Deterministic reduction of partial[b * stride + s] over blocks b, by the last block.
template <int R>
__device__ inline void reduce_partials(const double2 *partial, int nb, int stride, double2 *lds,
                                       double2 (&out)[R]) {
  double2 v[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    v[s] = make_double2(0.0, 0.0);
  const volatile double *pv = reinterpret_cast<const volatile double *>(partial);
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      int64_t at = ((int64_t)b * stride + s) * 2;
      v[s].x += pv[at];
      v[s].y += pv[at + 1];
    }
  }
  block_sum<R>(v, lds);
#pragma unroll
  for (int s = 0; s < R; ++s)
    out[s] = lds[s];
}

// ---------------------------------------------------------------------------
// Initialisation: x = 0, r = b, z = D^{-1} b, p = 0; rho = b^T z, ||b||^2.
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_init(int64_t n, const double2 *__restrict__ b, const double2 *__restrict__ dinv,
           double2 *__restrict__ x, double2 *__restrict__ r, double2 *__restrict__ z,
           double2 *__restrict__ p, double2 *__restrict__ partial, SysState *__restrict__ st,
           Control *__restrict__ ctl, int method, double tol2) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  double2 rz[R], bb[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    rz[s] = make_double2(0, 0);
    bb[s] = make_double2(0, 0);
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double2 di = dinv[i];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      int64_t k = i * R + s;
      double2 bv = b[k];
      double2 zv = bil(di, bv);
      x[k] = make_double2(0, 0);
      r[k] = bv;
      z[k] = zv;
      p[k] = make_double2(0, 0);
      rz[s] = cadd(rz[s], bil(bv, zv));
      bb[s].x += cabs2(bv);
    }
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    v[s] = rz[s];
    v[R + s] = bb[s];
  }
  block_sum<2 * R>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)
      partial[(int64_t)blockIdx.x * 2 * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[2 * R];
  reduce_partials<2 * R>(partial, gridDim.x, 2 * R, lds, tot);
  if (threadIdx.x == 0) {
    int na = 0;
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      S.bnorm2 = tot[R + s].x;
      S.rnorm2 = tot[R + s].x;
      S.rho = tot[s];
      S.alpha = make_double2(0, 0);
      S.beta = make_double2(0, 0);
      S.omega = make_double2(1, 0);
      S.iters = 0;
      S.breakdown = 0;
      // zero right-hand side: x = 0 is the exact solution
      S.active = (S.bnorm2 > 0.0) ? 1 : 0;
      na += S.active;
    }
    ctl->n_active = na;
    ctl->ticket = 0;
  }
}

// ---------------------------------------------------------------------------
// COCG step 1: p_new = z + beta p_old, q = A p_new, alpha = rho / (p_new^T q)
// ---------------------------------------------------------------------------
template <typename VT, int R>
__global__ void __launch_bounds__(kBlock)
    k_cocg_spmv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                const VT *__restrict__ val, const double2 *__restrict__ z,
                const double2 *__restrict__ p_old, double2 *__restrict__ p_new,
                double2 *__restrict__ q, double2 *__restrict__ partial, SysState *__restrict__ st,
                Control *__restrict__ ctl) {
  __shared__ double2 lds[(kBlock / 64) * R];
  bool act[R];
  double2 beta[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    beta[s] = st[s].beta;
  }
  const int g = threadIdx.x / kLanesPerRow;       // row slot in block
  const int lane = threadIdx.x % kLanesPerRow;
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  for (int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock; row0 < n;
       row0 += (int64_t)gridDim.x * kRowsPerBlock) {
    int64_t row = row0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += kLanesPerRow) {
        double2 a = to_d2(val[k]);
        int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s) {
          double2 pc = cfma(beta[s], p_old[c * R + s], z[c * R + s]);
          acc[s] = cfma(a, pc, acc[s]);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
#pragma unroll
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
        acc[s].x += __shfl_xor(acc[s].x, off, kLanesPerRow);
        acc[s].y += __shfl_xor(acc[s].y, off, kLanesPerRow);
      }
    }
    if (row < n && lane == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        int64_t k = row * R + s;
        double2 pn = cfma(beta[s], p_old[k], z[k]);
        p_new[k] = pn;
        q[k] = acc[s];
        pq[s] = cadd(pq[s], bil(pn, acc[s]));
      }
    }
  }
  block_sum<R>(pq, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < R; ++s)
      partial[(int64_t)blockIdx.x * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[R];
  reduce_partials<R>(partial, gridDim.x, R, lds, tot);
  if (threadIdx.x == 0) {
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      if (!S.active)
        continue;
      if (tot[s].x == 0.0 && tot[s].y == 0.0) {
        S.breakdown = 1;
        S.alpha = make_double2(0, 0);
      } else {
        S.alpha = cdiv(S.rho, tot[s]);
      }
    }
    ctl->ticket = 0;
  }
}

// ---------------------------------------------------------------------------
// COCG step 2: x += alpha p, r -= alpha q, z = D^{-1} r; beta = (r^T z)_new / rho
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocg_update(int64_t n, const double2 *__restrict__ dinv, const double2 *__restrict__ p,
                  const double2 *__restrict__ q, double2 *__restrict__ x, double2 *__restrict__ r,
                  double2 *__restrict__ z, double2 *__restrict__ partial, SysState *__restrict__ st,
                  Control *__restrict__ ctl, double tol2, int max_it) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
  }
  double2 rz[R], rr[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    rz[s] = make_double2(0, 0);
    rr[s] = make_double2(0, 0);
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double2 di = dinv[i];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      int64_t k = i * R + s;
      double2 xv = cfma(alpha[s], p[k], x[k]);
      double2 aq = bil(alpha[s], q[k]);
      double2 rv = csub(r[k], aq);
      double2 zv = bil(di, rv);
      x[k] = xv;
      r[k] = rv;
      z[k] = zv;
      rz[s] = cadd(rz[s], bil(rv, zv));
      rr[s].x += cabs2(rv);
    }
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    v[s] = rz[s];
    v[R + s] = rr[s];
  }
  block_sum<2 * R>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)
      partial[(int64_t)blockIdx.x * 2 * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[2 * R];
  reduce_partials<2 * R>(partial, gridDim.x, 2 * R, lds, tot);
  if (threadIdx.x == 0) {
    int na = 0;
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      if (!S.active)
        continue;
      S.iters += 1;
      S.rnorm2 = tot[R + s].x;
      double2 rho_new = tot[s];
      if (S.rnorm2 <= tol2 * S.bnorm2) {
        S.active = 0;
      } else if (S.breakdown || (S.rho.x == 0.0 && S.rho.y == 0.0) || S.iters >= max_it) {
        S.active = 0;
      } else {
        S.beta = cdiv(rho_new, S.rho);
        S.rho = rho_new;
        ++na;
      }
    }
    ctl->n_active = na;
    ctl->ticket = 0;
  }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (van der Vorst 1992), right-preconditioned with D^{-1}:
//   rho_i = r0^H r; beta = (rho_i/rho_{i-1})(alpha/omega); p = r + beta (p - omega v)
//   y = D^{-1} p; v = A y; alpha = rho_i / (r0^H v); s = r - alpha v
//   zz = D^{-1} s; t = A zz; omega = (t^H s)/(t^H t); x += alpha y + omega zz; r = s - omega t
// Kernels: bs_p (p update, y), spmv_v (v = A y, r0^T v), bs_s (s, zz), spmv_t (t = A zz,
// t^H s, t^H t), bs_x (x, r, r0^T r, r^H r).  Same masking and reductions as COCG.
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_init(int64_t n, const double2 *__restrict__ b, double2 *__restrict__ x,
              double2 *__restrict__ r, double2 *__restrict__ r0, double2 *__restrict__ p,
              double2 *__restrict__ v, double2 *__restrict__ partial, SysState *__restrict__ st,
              Control *__restrict__ ctl) {
  __shared__ double2 lds[(kBlock / 64) * R];
  double2 bb[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    bb[s] = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      int64_t k = i * R + s;
      double2 bv = b[k];
      x[k] = make_double2(0, 0);
      r[k] = bv;
      r0[k] = bv;
      p[k] = make_double2(0, 0);
      v[k] = make_double2(0, 0);
      bb[s].x += cabs2(bv);
    }
  }
  block_sum<R>(bb, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < R; ++s)
      partial[(int64_t)blockIdx.x * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[R];
  reduce_partials<R>(partial, gridDim.x, R, lds, tot);
  if (threadIdx.x == 0) {
    int na = 0;
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      S.bnorm2 = tot[s].x;
      S.rnorm2 = tot[s].x;
      // r0 = b, so r0^T r = b^T b (bilinear)
      S.rho = make_double2(1, 0);
      S.alpha = make_double2(1, 0);
      S.omega = make_double2(1, 0);
      S.beta = make_double2(0, 0);
      S.iters = 0;
      S.breakdown = 0;
      S.active = (S.bnorm2 > 0.0) ? 1 : 0;
      na += S.active;
    }
    ctl->n_active = na;
    ctl->ticket = 0;
  }
}

// rho_1 = r0^H r for the first BiCGSTAB step; later steps get rho_i from bs_x.
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_dot_r0r(int64_t n, const double2 *__restrict__ r0, const double2 *__restrict__ r,
                 double2 *__restrict__ partial, SysState *__restrict__ st, Control *__restrict__ ctl) {
  __shared__ double2 lds[(kBlock / 64) * R];
  double2 d[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    d[s] = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      int64_t k = i * R + s;
      d[s] = cadd(d[s], bil(cconj(r0[k]), r[k]));
    }
  }
  block_sum<R>(d, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < R; ++s)
      partial[(int64_t)blockIdx.x * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[R];
  reduce_partials<R>(partial, gridDim.x, R, lds, tot);
  if (threadIdx.x == 0) {
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      if (!S.active)
        continue;
      // stash rho_i in beta's slot for bs_p: beta = (rho_i/rho)(alpha/omega)
      double2 rho_i = tot[s];
      if ((rho_i.x == 0.0 && rho_i.y == 0.0) || (S.omega.x == 0.0 && S.omega.y == 0.0)) {
        S.breakdown = 1;
        S.active = 0;
        continue;
      }
      S.beta = cmul(cdiv(rho_i, S.rho), cdiv(S.alpha, S.omega));
      S.rho = rho_i;
    }
    ctl->ticket = 0;
  }
}

// p = r + beta (p - omega v); y = D^{-1} p
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_p(int64_t n, const double2 *__restrict__ dinv, const double2 *__restrict__ r,
           const double2 *__restrict__ v, double2 *__restrict__ p, double2 *__restrict__ y,
           const SysState *__restrict__ st) {
  bool act[R];
  double2 beta[R], omega[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    beta[s] = st[s].beta;
    omega[s] = st[s].omega;
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double2 di = dinv[i];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      int64_t k = i * R + s;
      double2 pv = cadd(r[k], bil(beta[s], csub(p[k], bil(omega[s], v[k]))));
      p[k] = pv;
      y[k] = bil(di, pv);
    }
  }
}

// out = A in, with partial dots: mode 0 -> (r0^H out); mode 1 -> (out^H s, out^H out)
template <typename VT, int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_spmv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
              const VT *__restrict__ val, const double2 *__restrict__ in, double2 *__restrict__ out,
              const double2 *__restrict__ other, int mode, double2 *__restrict__ partial,
              SysState *__restrict__ st, Control *__restrict__ ctl) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[s].active != 0;
  const int g = threadIdx.x / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  double2 d1[R], d2[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    d1[s] = make_double2(0, 0);
    d2[s] = make_double2(0, 0);
  }
  for (int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock; row0 < n;
       row0 += (int64_t)gridDim.x * kRowsPerBlock) {
    int64_t row = row0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += kLanesPerRow) {
        double2 a = to_d2(val[k]);
        int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, in[c * R + s], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
#pragma unroll
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
        acc[s].x += __shfl_xor(acc[s].x, off, kLanesPerRow);
        acc[s].y += __shfl_xor(acc[s].y, off, kLanesPerRow);
      }
    }
    if (row < n && lane == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        int64_t k = row * R + s;
        out[k] = acc[s];
        double2 o = other[k];
        if (mode == 0) {
          d1[s] = cadd(d1[s], bil(cconj(o), acc[s]));
        } else {
          d1[s] = cadd(d1[s], bil(cconj(acc[s]), o));
          d2[s].x += cabs2(acc[s]);
        }
      }
    }
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    v[s] = d1[s];
    v[R + s] = d2[s];
  }
  block_sum<2 * R>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)
      partial[(int64_t)blockIdx.x * 2 * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[2 * R];
  reduce_partials<2 * R>(partial, gridDim.x, 2 * R, lds, tot);
  if (threadIdx.x == 0) {
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      if (!S.active)
        continue;
      if (mode == 0) {
        if (tot[s].x == 0.0 && tot[s].y == 0.0) {
          S.breakdown = 1;
          S.active = 0;
          continue;
        }
        S.alpha = cdiv(S.rho, tot[s]);
      } else {
        double tt = tot[R + s].x;
        S.omega = (tt > 0.0) ? cscale(tot[s], 1.0 / tt) : make_double2(0, 0);
      }
    }
    ctl->ticket = 0;
  }
}

// s = r - alpha v (stored in r); zz = D^{-1} s
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_s(int64_t n, const double2 *__restrict__ dinv, const double2 *__restrict__ v,
           double2 *__restrict__ r, double2 *__restrict__ zz, const SysState *__restrict__ st) {
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double2 di = dinv[i];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      int64_t k = i * R + s;
      double2 sv = csub(r[k], bil(alpha[s], v[k]));
      r[k] = sv;
      zz[k] = bil(di, sv);
    }
  }
}

// x += alpha y + omega zz; r = s - omega t; rho_new = r0^H r; ||r||^2
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_x(int64_t n, const double2 *__restrict__ y, const double2 *__restrict__ zz,
           const double2 *__restrict__ t, const double2 *__restrict__ r0, double2 *__restrict__ x,
           double2 *__restrict__ r, double2 *__restrict__ partial, SysState *__restrict__ st,
           Control *__restrict__ ctl, double tol2, int max_it) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R], omega[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
    omega[s] = st[s].omega;
  }
  double2 d1[R], d2[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    d1[s] = make_double2(0, 0);
    d2[s] = make_double2(0, 0);
  }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      int64_t k = i * R + s;
      double2 xv = cadd(x[k], cadd(bil(alpha[s], y[k]), bil(omega[s], zz[k])));
      double2 rv = csub(r[k], bil(omega[s], t[k]));
      x[k] = xv;
      r[k] = rv;
      d1[s] = cadd(d1[s], bil(cconj(r0[k]), rv));
      d2[s].x += cabs2(rv);
    }
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    v[s] = d1[s];
    v[R + s] = d2[s];
  }
  block_sum<2 * R>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)
      partial[(int64_t)blockIdx.x * 2 * R + s] = lds[s];
  }
  if (!last_block(&ctl->ticket))
    return;
  double2 tot[2 * R];
  reduce_partials<2 * R>(partial, gridDim.x, 2 * R, lds, tot);
  if (threadIdx.x == 0) {
    int na = 0;
    for (int s = 0; s < R; ++s) {
      SysState &S = st[s];
      if (!S.active)
        continue;
      S.iters += 1;
      S.rnorm2 = tot[R + s].x;
      if (S.rnorm2 <= tol2 * S.bnorm2) {
        S.active = 0;
        continue;
      }
      if (S.iters >= max_it || (S.omega.x == 0.0 && S.omega.y == 0.0)) {
        S.active = 0;
        continue;
      }
      double2 rho_i = tot[s];
      if (rho_i.x == 0.0 && rho_i.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
        continue;
      }
      S.beta = cmul(cdiv(rho_i, S.rho), cdiv(S.alpha, S.omega));
      S.rho = rho_i;
      ++na;
    }
    ctl->n_active = na;
    ctl->ticket = 0;
  }
}

// Jacobi inverse diagonal (1/a_rr; 1 where the diagonal is missing or zero).
template <typename VT>
__global__ void k_dinv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                       const VT *__restrict__ val, double2 *__restrict__ dinv) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  double2 d = make_double2(0, 0);
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    if (col[k] == r)
      d = to_d2(val[k]);
  dinv[r] = (d.x == 0.0 && d.y == 0.0) ? make_double2(1, 0) : cdiv(make_double2(1, 0), d);
}

template <typename VT, int R>
__global__ void __launch_bounds__(kBlock)
    k_spmv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
           const VT *__restrict__ val, const double2 *__restrict__ x, double2 *__restrict__ y) {
  const int g = threadIdx.x / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  for (int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock; row0 < n;
       row0 += (int64_t)gridDim.x * kRowsPerBlock) {
    int64_t row = row0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += kLanesPerRow) {
        double2 a = to_d2(val[k]);
        int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, x[c * R + s], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
#pragma unroll
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
        acc[s].x += __shfl_xor(acc[s].x, off, kLanesPerRow);
        acc[s].y += __shfl_xor(acc[s].y, off, kLanesPerRow);
      }
    }
    if (row < n && lane == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s)
        y[row * R + s] = acc[s];
    }
  }
}

} // namespace

struct efm_solver {
  int64_t n = 0, nnz = 0;
  const int32_t *row_ptr = nullptr;
  const int32_t *col = nullptr;
  int R = 1;
  int vtype = EFM_C128;
  int grid = 1024;
  efm::DevBuf val, dinv, x, r, z, p0, p1, q, r0, v, y, zz, t, partial, state, control;
  bool profile = false;
  double spmv_ms = 0.0;
  int64_t spmv_launches = 0;
  std::vector<hipEvent_t> evs;
  ~efm_solver() {
    for (auto e : evs)
      (void)hipEventDestroy(e);
  }
};

namespace {

template <int R> struct Dispatch {
  static int init(efm_solver *s, const double2 *b, hipStream_t st, int method, double tol2) {
    if (method == EFM_METHOD_COCG)
      hipLaunchKernelGGL(k_init<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, b, s->dinv.as<double2>(),
                         s->x.as<double2>(), s->r.as<double2>(), s->z.as<double2>(),
                         s->p0.as<double2>(), s->partial.as<double2>(), s->state.as<SysState>(),
                         s->control.as<Control>(), method, tol2);
    else
      hipLaunchKernelGGL(k_bs_init<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, b,
                         s->x.as<double2>(), s->r.as<double2>(), s->r0.as<double2>(),
                         s->p0.as<double2>(), s->v.as<double2>(), s->partial.as<double2>(),
                         s->state.as<SysState>(), s->control.as<Control>());
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  template <typename VT>
  static int cocg_iter(efm_solver *s, int it, hipStream_t st, double tol2, int max_it,
                       hipEvent_t e0, hipEvent_t e1) {
    double2 *p_old = (it & 1) ? s->p1.as<double2>() : s->p0.as<double2>();
    double2 *p_new = (it & 1) ? s->p0.as<double2>() : s->p1.as<double2>();
    if (e0)
      EFM_HIP_CHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL((k_cocg_spmv<VT, R>), dim3(s->grid), dim3(kBlock), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<VT>(), s->z.as<double2>(), p_old, p_new, s->q.as<double2>(),
                       s->partial.as<double2>(), s->state.as<SysState>(), s->control.as<Control>());
    EFM_LAUNCH_CHECK();
    if (e1)
      EFM_HIP_CHECK(hipEventRecord(e1, st));
    hipLaunchKernelGGL(k_cocg_update<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n,
                       s->dinv.as<double2>(), p_new, s->q.as<double2>(), s->x.as<double2>(),
                       s->r.as<double2>(), s->z.as<double2>(), s->partial.as<double2>(),
                       s->state.as<SysState>(), s->control.as<Control>(), tol2, max_it);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  template <typename VT>
  static int bicgstab_iter(efm_solver *s, int it, hipStream_t st, double tol2, int max_it,
                           hipEvent_t e0, hipEvent_t e1) {
    SysState *S = s->state.as<SysState>();
    Control *C = s->control.as<Control>();
    double2 *part = s->partial.as<double2>();
    if (it == 0) {
      hipLaunchKernelGGL(k_bs_dot_r0r<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n,
                         s->r0.as<double2>(), s->r.as<double2>(), part, S, C);
      EFM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_bs_p<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->dinv.as<double2>(),
                       s->r.as<double2>(), s->v.as<double2>(), s->p0.as<double2>(),
                       s->y.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    if (e0)
      EFM_HIP_CHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL((k_bs_spmv<VT, R>), dim3(s->grid), dim3(kBlock), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<VT>(), s->y.as<double2>(), s->v.as<double2>(),
                       s->r0.as<double2>(), 0, part, S, C);
    EFM_LAUNCH_CHECK();
    if (e1)
      EFM_HIP_CHECK(hipEventRecord(e1, st));
    hipLaunchKernelGGL(k_bs_s<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->dinv.as<double2>(),
                       s->v.as<double2>(), s->r.as<double2>(), s->zz.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_bs_spmv<VT, R>), dim3(s->grid), dim3(kBlock), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<VT>(), s->zz.as<double2>(), s->t.as<double2>(),
                       s->r.as<double2>(), 1, part, S, C);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bs_x<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->y.as<double2>(),
                       s->zz.as<double2>(), s->t.as<double2>(), s->r0.as<double2>(),
                       s->x.as<double2>(), s->r.as<double2>(), part, S, C, tol2, max_it);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  template <typename VT> static int spmv(efm_solver *s, const double2 *x, double2 *y, hipStream_t st) {
    hipLaunchKernelGGL((k_spmv<VT, R>), dim3(s->grid), dim3(kBlock), 0, st, s->n, s->row_ptr, s->col,
                       s->val.as<VT>(), x, y);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }
};

template <typename F> int with_rhs(int R, F &&f) {
  switch (R) {
  case 1: return f(std::integral_constant<int, 1>{});
  case 2: return f(std::integral_constant<int, 2>{});
  case 3: return f(std::integral_constant<int, 3>{});
  case 4: return f(std::integral_constant<int, 4>{});
  case 8: return f(std::integral_constant<int, 8>{});
  default:
    set_error("nrhs must be 1, 2, 3, 4 or 8");
    return EFM_ERR_ARG;
  }
}

} // namespace

extern "C" int efm_solver_create(int64_t n, int64_t nnz, const int32_t *row_ptr,
                                 const int32_t *col_idx, int nrhs, int value_type, void *stream,
                                 efm_solver **out) {
  EFM_REQUIRE(out, "efm_solver_create: out is NULL");
  *out = nullptr;
  EFM_REQUIRE(n > 0 && nnz > 0 && row_ptr && col_idx, "efm_solver_create: bad CSR");
  EFM_REQUIRE(nrhs == 1 || nrhs == 2 || nrhs == 3 || nrhs == 4 || nrhs == 8,
              "efm_solver_create: nrhs must be 1, 2, 3, 4 or 8");
  EFM_REQUIRE(value_type == EFM_C128 || value_type == EFM_C64, "efm_solver_create: bad value_type");
  (void)stream;
  auto s = std::make_unique<efm_solver>();
  s->n = n;
  s->nnz = nnz;
  s->row_ptr = row_ptr;
  s->col = col_idx;
  s->R = nrhs;
  s->vtype = value_type;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      ncu = prop.multiProcessorCount;
  }
  int64_t need = (n + kRowsPerBlock - 1) / kRowsPerBlock;
  int64_t g = (int64_t)ncu * 8;
  s->grid = (int)(need < g ? (need < 1 ? 1 : need) : g);
  const size_t vb = (size_t)n * nrhs * sizeof(double2);
  int rc;
  const size_t vbytes = (size_t)nnz * (value_type == EFM_C128 ? 16 : 8);
  if ((rc = s->val.alloc(vbytes)) || (rc = s->dinv.alloc(n * sizeof(double2))) ||
      (rc = s->x.alloc(vb)) || (rc = s->r.alloc(vb)) || (rc = s->z.alloc(vb)) ||
      (rc = s->p0.alloc(vb)) || (rc = s->p1.alloc(vb)) || (rc = s->q.alloc(vb)) ||
      (rc = s->partial.alloc((size_t)s->grid * 2 * kMaxRhs * sizeof(double2))) ||
      (rc = s->state.alloc(kMaxRhs * sizeof(SysState))) || (rc = s->control.alloc(sizeof(Control))))
    return rc;
  EFM_HIP_CHECK(hipMemset(s->control.p, 0, sizeof(Control)));
  EFM_HIP_CHECK(hipMemset(s->state.p, 0, kMaxRhs * sizeof(SysState)));
  *out = s.release();
  return EFM_OK;
}

extern "C" void efm_solver_destroy(efm_solver *s) { delete s; }

extern "C" int efm_solver_set_matrix(efm_solver *s, const void *values, void *stream) {
  EFM_REQUIRE(s && values, "efm_solver_set_matrix: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  EFM_HIP_CHECK(hipMemcpyAsync(s->val.p, values, s->val.bytes, hipMemcpyDeviceToDevice, st));
  if (s->vtype == EFM_C128)
    hipLaunchKernelGGL(k_dinv<double2>, dim3(grid_for(s->n, 256)), dim3(256), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<double2>(), s->dinv.as<double2>());
  else
    hipLaunchKernelGGL(k_dinv<float2>, dim3(grid_for(s->n, 256)), dim3(256), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<float2>(), s->dinv.as<double2>());
  EFM_LAUNCH_CHECK();
  return EFM_OK;
}

extern "C" int efm_solver_spmv(efm_solver *s, const double *x, double *y, void *stream) {
  EFM_REQUIRE(s && x && y, "efm_solver_spmv: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  return with_rhs(s->R, [&](auto RC) -> int {
    constexpr int R = decltype(RC)::value;
    if (s->vtype == EFM_C128)
      return Dispatch<R>::template spmv<double2>(s, (const double2 *)x, (double2 *)y, st);
    return Dispatch<R>::template spmv<float2>(s, (const double2 *)x, (double2 *)y, st);
  });
}

extern "C" int efm_solver_profile(efm_solver *s, int enable) {
  EFM_REQUIRE(s, "efm_solver_profile: NULL solver");
  s->profile = enable != 0;
  s->spmv_ms = 0.0;
  s->spmv_launches = 0;
  return EFM_OK;
}

extern "C" int efm_solver_profile_read(efm_solver *s, double *h_ms, int64_t *h_launches) {
  EFM_REQUIRE(s, "efm_solver_profile_read: NULL solver");
  if (h_ms)
    *h_ms = s->spmv_ms;
  if (h_launches)
    *h_launches = s->spmv_launches;
  return EFM_OK;
}

extern "C" int efm_solver_solve(efm_solver *s, const double *B, double *X,
                                const efm_solve_opts *h_opts, efm_solve_result *h_result,
                                int32_t *h_iters, double *h_resid, void *stream) {
  EFM_REQUIRE(s && B && X && h_opts, "efm_solver_solve: NULL argument");
  EFM_REQUIRE(h_opts->method == EFM_METHOD_COCG || h_opts->method == EFM_METHOD_BICGSTAB,
              "efm_solver_solve: unknown method");
  EFM_REQUIRE(h_opts->max_iterations >= 1, "efm_solver_solve: max_iterations < 1");
  hipStream_t st = (hipStream_t)stream;
  const double tol2 = h_opts->tolerance * h_opts->tolerance;
  const int max_it = h_opts->max_iterations;
  const int check = h_opts->check_every > 0 ? h_opts->check_every : 32;
  const bool cocg = h_opts->method == EFM_METHOD_COCG;
  int rc;
  if (!cocg) {
    const size_t vb = (size_t)s->n * s->R * sizeof(double2);
    if (!s->r0.p && ((rc = s->r0.alloc(vb)) || (rc = s->v.alloc(vb)) || (rc = s->y.alloc(vb)) ||
                     (rc = s->zz.alloc(vb)) || (rc = s->t.alloc(vb))))
      return rc;
  }
  if (s->profile && s->evs.size() < (size_t)2 * check) {
    for (auto e : s->evs)
      (void)hipEventDestroy(e);
    s->evs.assign(2 * check, nullptr);
    for (auto &e : s->evs)
      EFM_HIP_CHECK(hipEventCreate(&e));
  }
  rc = with_rhs(s->R, [&](auto RC) -> int {
    constexpr int R = decltype(RC)::value;
    return Dispatch<R>::init(s, (const double2 *)B, st, h_opts->method, tol2);
  });
  if (rc != EFM_OK)
    return rc;
  Control hc{};
  EFM_HIP_CHECK(hipMemcpyAsync(&hc, s->control.p, sizeof(Control), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  int it = 0;
  while (hc.n_active > 0 && it < max_it) {
    int chunk = check < (max_it - it) ? check : (max_it - it);
    for (int c = 0; c < chunk; ++c, ++it) {
      hipEvent_t e0 = s->profile ? s->evs[2 * c] : nullptr;
      hipEvent_t e1 = s->profile ? s->evs[2 * c + 1] : nullptr;
      rc = with_rhs(s->R, [&](auto RC) -> int {
        constexpr int R = decltype(RC)::value;
        if (cocg) {
          if (s->vtype == EFM_C128)
            return Dispatch<R>::template cocg_iter<double2>(s, it, st, tol2, max_it, e0, e1);
          return Dispatch<R>::template cocg_iter<float2>(s, it, st, tol2, max_it, e0, e1);
        }
        if (s->vtype == EFM_C128)
          return Dispatch<R>::template bicgstab_iter<double2>(s, it, st, tol2, max_it, e0, e1);
        return Dispatch<R>::template bicgstab_iter<float2>(s, it, st, tol2, max_it, e0, e1);
      });
      if (rc != EFM_OK)
        return rc;
    }
    EFM_HIP_CHECK(hipMemcpyAsync(&hc, s->control.p, sizeof(Control), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    if (s->profile) {
      for (int c = 0; c < chunk; ++c) {
        float ms = 0.f;
        EFM_HIP_CHECK(hipEventElapsedTime(&ms, s->evs[2 * c], s->evs[2 * c + 1]));
        s->spmv_ms += ms;
        s->spmv_launches += 1;
      }
    }
  }
  // the solution of iteration-masked COCG lives in x for every system
  EFM_HIP_CHECK(hipMemcpyAsync(X, s->x.p, (size_t)s->n * s->R * sizeof(double2),
                               hipMemcpyDeviceToDevice, st));
  SysState hs[kMaxRhs];
  EFM_HIP_CHECK(hipMemcpyAsync(hs, s->state.p, sizeof(SysState) * s->R, hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  efm_solve_result res{};
  res.converged = 1;
  for (int k = 0; k < s->R; ++k) {
    double rel = hs[k].bnorm2 > 0 ? sqrt(hs[k].rnorm2 / hs[k].bnorm2) : 0.0;
    bool conv = hs[k].bnorm2 == 0.0 || hs[k].rnorm2 <= tol2 * hs[k].bnorm2;
    if (!conv)
      res.converged = 0;
    if (hs[k].iters > res.iterations)
      res.iterations = hs[k].iters;
    if (rel > res.max_rel_residual)
      res.max_rel_residual = rel;
    if (h_iters)
      h_iters[k] = hs[k].iters;
    if (h_resid)
      h_resid[k] = rel;
  }
  if (h_result)
    *h_result = res;
  return EFM_OK;
}
