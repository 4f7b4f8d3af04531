// Jacobi-preconditioned Krylov solvers on a CSR matrix for a block of independent
// right-hand sides (replaces solve_linear, src/solver.cpp:35-193).
//
//   COCG      complex-symmetric A (A^T = A): the eigenmode/K-M systems.
//   BiCGSTAB  general A (Bloch-periodic systems lose symmetry).
//
// Vectors and all reductions are FP64; matrix values are stored complex FP64 or
// complex FP32 (promoted on load).  One COCG iteration is four launches:
//
//   cocg_spmv   p_new = z + beta p_old (on the fly for every gathered column),
//               q = A p_new; per-block partial sums of p_new^T q
//   finalize    one block: alpha = rho / sum(partials)
//   cocg_update x += alpha p, r -= alpha q, z = D^{-1} r; partial sums of r^T z, r^H r
//   finalize    one block: convergence test, beta, active-system count
//
// Partial sums are combined by a separate one-block kernel in a fixed order, so the
// results are deterministic and no kernel needs a device-scope fence (on gfx950 such
// a fence writes back the XCD's L2 from every wave that issues it).  Converged systems
// are masked: their vectors stop changing, and the host polls one counter every
// `check_every` iterations.
#include "efm_internal.h"

#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kLanesPerRow = 8;  // lanes cooperating on one CSR row
constexpr int kRowsPerBlock = kBlock / kLanesPerRow;
constexpr int kMaxRhs = 8;

struct SysState {
  double2 rho;      // r^T z (COCG) / r0^H r (BiCGSTAB)
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
  int n_active;     // systems still iterating
  int pad;
};

enum Stage : int {
  kFinInitCocg = 0,   // partials: [rz(R), bb(R)]
  kFinInitBs = 1,     // partials: [bb(R)]
  kFinAlpha = 2,      // partials: [pq(R)]
  kFinBeta = 3,       // partials: [rz(R), rr(R)]
  kFinBsRho0 = 4,     // partials: [r0^H r (R)]
  kFinBsAlpha = 5,    // partials: [r0^H v (R)]
  kFinBsOmega = 6,    // partials: [t^H s (R), t^H t (R)]
  kFinBsX = 7,        // partials: [r0^H r (R), r^H r (R)]
};

__device__ inline double2 bil(double2 a, double2 b) {  // a*b (no conjugate)
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Block-wide sum of V complex values per thread; result in lds[0..V) (all threads).
template <int V>
__device__ inline void block_sum(double2 (&v)[V], double2 *lds) {
#pragma unroll
  for (int s = 0; s < V; ++s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v[s].x += __shfl_xor(v[s].x, off, 64);
      v[s].y += __shfl_xor(v[s].y, off, 64);
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < V; ++s)
      lds[wave * V + s] = v[s];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < V; ++s) {
      double2 a = lds[s];
      for (int w = 1; w < kBlock / 64; ++w)
        a = cadd(a, lds[w * V + s]);
      lds[s] = a;
    }
  }
  __syncthreads();
}

template <int V>
__device__ inline void write_partials(double2 (&v)[V], double2 *lds, double2 *partial) {
  block_sum<V>(v, lds);
  if (threadIdx.x < V)
    partial[(int64_t)blockIdx.x * V + threadIdx.x] = lds[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Finalisation: one block reduces the per-block partials in a fixed order and
// advances the per-system scalars.
// ---------------------------------------------------------------------------
template <int R, int V>
__device__ inline void reduce_all(const double2 *__restrict__ partial, int nb, double2 *lds,
                                  double2 (&tot)[V]) {
  double2 v[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    v[s] = make_double2(0.0, 0.0);
  // U blocks' partials are loaded before any is added, so their round trips overlap; the
  // additions keep the order b, b + kBlock, ... of the plain loop (same sums, bit for bit)
  constexpr int U = V >= 32 ? 1 : (32 / V > 8 ? 8 : 32 / V);
  int b = threadIdx.x;
  for (; b + (U - 1) * kBlock < nb; b += U * kBlock) {
    double2 t[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int s = 0; s < V; ++s)
        t[u][s] = partial[(int64_t)(b + u * kBlock) * V + s];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int s = 0; s < V; ++s)
        v[s] = cadd(v[s], t[u][s]);
  }
  for (; b < nb; b += kBlock) {
#pragma unroll
    for (int s = 0; s < V; ++s)
      v[s] = cadd(v[s], partial[(int64_t)b * V + s]);
  }
  block_sum<V>(v, lds);
#pragma unroll
  for (int s = 0; s < V; ++s)
    tot[s] = lds[s];
}

template <int R, int Stage_>
__global__ void __launch_bounds__(kBlock)
    k_finalize(const double2 *__restrict__ partial, int nb, SysState *__restrict__ st,
               Control *__restrict__ ctl, double tol2, int max_it) {
  constexpr int V = (Stage_ == kFinInitBs || Stage_ == kFinAlpha || Stage_ == kFinBsRho0 ||
                     Stage_ == kFinBsAlpha)
                        ? R
                        : 2 * R;
  __shared__ double2 lds[(kBlock / 64) * V];
  double2 tot[V];
  reduce_all<R, V>(partial, nb, lds, tot);
  if (threadIdx.x != 0)
    return;
  int na = 0;
  for (int s = 0; s < R; ++s) {
    SysState &S = st[s];
    if constexpr (Stage_ == kFinInitCocg || Stage_ == kFinInitBs) {
      const double bb = (Stage_ == kFinInitCocg) ? tot[R + s].x : tot[s].x;
      S.bnorm2 = bb;
      S.rnorm2 = bb;
      S.rho = (Stage_ == kFinInitCocg) ? tot[s] : make_double2(1, 0);
      S.alpha = (Stage_ == kFinInitCocg) ? make_double2(0, 0) : make_double2(1, 0);
      S.beta = make_double2(0, 0);
      S.omega = make_double2(1, 0);
      S.iters = 0;
      S.breakdown = 0;
      S.active = (bb > 0.0) ? 1 : 0;  // zero right-hand side: x = 0 is exact
      na += S.active;
      continue;
    }
    if (!S.active)
      continue;
    if constexpr (Stage_ == kFinAlpha) {
      if (tot[s].x == 0.0 && tot[s].y == 0.0) {
        S.breakdown = 1;
        S.alpha = make_double2(0, 0);
      } else {
        S.alpha = cdiv(S.rho, tot[s]);
      }
    } else if constexpr (Stage_ == kFinBeta) {
      S.iters += 1;
      S.rnorm2 = tot[R + s].x;
      const double2 rho_new = tot[s];
      if (S.rnorm2 <= tol2 * S.bnorm2) {
        S.active = 0;
      } else if (S.breakdown || (S.rho.x == 0.0 && S.rho.y == 0.0) || S.iters >= max_it) {
        S.active = 0;
      } else {
        S.beta = cdiv(rho_new, S.rho);
        S.rho = rho_new;
      }
    } else if constexpr (Stage_ == kFinBsRho0) {
      const double2 rho_i = tot[s];
      if ((rho_i.x == 0.0 && rho_i.y == 0.0)) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.beta = cmul(cdiv(rho_i, S.rho), cdiv(S.alpha, S.omega));
        S.rho = rho_i;
      }
    } else if constexpr (Stage_ == kFinBsAlpha) {
      if (tot[s].x == 0.0 && tot[s].y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.alpha = cdiv(S.rho, tot[s]);
      }
    } else if constexpr (Stage_ == kFinBsOmega) {
      const double tt = tot[R + s].x;
      S.omega = (tt > 0.0) ? cscale(tot[s], 1.0 / tt) : make_double2(0, 0);
    } else if constexpr (Stage_ == kFinBsX) {
      S.iters += 1;
      S.rnorm2 = tot[R + s].x;
      const double2 rho_i = tot[s];
      if (S.rnorm2 <= tol2 * S.bnorm2) {
        S.active = 0;
      } else if (S.iters >= max_it || (S.omega.x == 0.0 && S.omega.y == 0.0)) {
        S.active = 0;
      } else if (rho_i.x == 0.0 && rho_i.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.beta = cmul(cdiv(rho_i, S.rho), cdiv(S.alpha, S.omega));
        S.rho = rho_i;
      }
    }
    na += S.active;
  }
  ctl->n_active = na;
}

// ---------------------------------------------------------------------------
// COCG.  The search direction and the preconditioned residual of every row live in
// one record W[row] = {z_0..z_{R-1}, p_0..p_{R-1}} (2R complex values), so the SpMV
// gathers both with one cache line per column.  Two records alternate between
// iterations: the SpMV reads W_a and writes p_new into W_b; the update writes the new
// z into W_b; the next SpMV reads W_b.
// ---------------------------------------------------------------------------
// x = 0, r = b, W0.z = D^{-1} b, W0.p = 0; partials rho = b^T z, ||b||^2
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocg_init(int64_t n, const double2 *__restrict__ b, const double2 *__restrict__ dinv,
                double2 *__restrict__ x, double2 *__restrict__ r, double2 *__restrict__ W,
                double2 *__restrict__ partial) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    v[s] = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double2 di = dinv[i];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      int64_t k = i * R + s;
      double2 bv = b[k];
      double2 zv = bil(di, bv);
      x[k] = make_double2(0, 0);
      r[k] = bv;
      W[i * 2 * R + s] = zv;
      W[i * 2 * R + R + s] = make_double2(0, 0);
      v[s] = cadd(v[s], bil(bv, zv));
      v[R + s].x += cabs2(bv);
    }
  }
  write_partials<2 * R>(v, lds, partial);
}

// Fused COCG SpMV kernels.  All compute, for every row and active system s,
//   p_new = z + beta p_old  (at every gathered column, from the record W_a),
//   q = A p_new,  W_b.p = p_new,  partial sums of p_new^T q,
// and differ only in how rows map onto lanes.  The autotuner times each variant on the
// installed matrix and keeps the fastest (efm_solver_autotune).
//
// Row-group variants: LPR lanes share a row and stride over its entries; rows are
// taken in chunks of 256/LPR.  Blocks are persistent and XCD-aware: block b runs on XCD
// b % 8 and walks chunks of that XCD's contiguous share of the rows, so the column
// gathers of one XCD stay within its own L2.
template <typename VT, int R, int LPR>
__global__ void __launch_bounds__(kBlock)
    k_cocg_spmv_rows(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                     const VT *__restrict__ val, const double2 *__restrict__ Wa,
                     double2 *__restrict__ Wb, double2 *__restrict__ q, double2 *__restrict__ partial,
                     const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * R];
  constexpr int kRows = kBlock / LPR;
  bool act[R];
  double2 beta[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    beta[s] = st[s].beta;
  }
  const int g = threadIdx.x / LPR;
  const int lane = threadIdx.x % LPR;
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  const int64_t nchunks = (n + kRows - 1) / kRows;
  const int64_t span = (nchunks + 7) / 8;
  const int xcd = blockIdx.x & 7;
  const int per_xcd = gridDim.x >> 3;
  const int64_t c_begin = xcd * span;
  const int64_t c_end = min(nchunks, c_begin + span);
  for (int64_t chunk = c_begin + (blockIdx.x >> 3); chunk < c_end; chunk += per_xcd) {
    const int64_t row = chunk * kRows + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      const int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += LPR) {
        const double2 a = to_d2(val[k]);
        const double2 *w = Wa + (int64_t)col[k] * 2 * R;
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, cfma(beta[s], w[R + s], w[s]), acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<LPR>(acc[s]);
    if (row < n && lane == 0) {
      const double2 *w = Wa + row * 2 * R;
      double2 *wb = Wb + row * 2 * R;
      // the whole record is written (z = 0; the update overwrites it) so that no partial
      // 64-byte line reaches HBM, which would need a read-modify-write
      double2 rec[2 * R];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        rec[s] = make_double2(0, 0);
        rec[R + s] = cfma(beta[s], w[R + s], w[s]);
      }
#pragma unroll
      for (int s = 0; s < 2 * R; ++s)
        wb[s] = rec[s];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        q[row * R + s] = acc[s];
        pq[s] = cadd(pq[s], bil(rec[R + s], acc[s]));
      }
    }
  }
  write_partials<R>(pq, lds, partial);
}

// Tiled variants.  A tile is a run of whole rows with at most TNNZ nonzeros and
// kBlock/RL rows.  Phase 1: thread t takes entries k = kb + t + 256 j, loads (value,
// column) coalesced, gathers the column's record and stores a * (z + beta p) in LDS;
// all of these loads are independent.  Phase 2: RL lanes per row sum the row's
// products from LDS; lane 0 writes q and p_new.  Persistent, XCD-aware tile walk.
constexpr int kTileRows = 64;
// nonzeros per tile: the LDS product buffer (tile_nnz * R complex doubles) stays <= 32 KB
__host__ __device__ constexpr int tile_nnz(int R) { return R <= 2 ? 1024 : (R <= 4 ? 512 : 256); }
// the smaller tiles of the second tiled variant (at least one entry per thread)
__host__ __device__ constexpr int tile_nnz2(int R) { return tile_nnz(R) / 2 < 256 ? 256 : tile_nnz(R) / 2; }

template <typename VT, int R, int TNNZ, int RL>
__global__ void __launch_bounds__(kBlock)
    k_cocg_spmv_tiled(int ntiles, int xcd_span, const int32_t *__restrict__ tile_row,
                      const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                      const VT *__restrict__ val, const double2 *__restrict__ Wa,
                      double2 *__restrict__ Wb, double2 *__restrict__ q,
                      double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 prod[TNNZ * R];
  __shared__ double2 lds[(kBlock / 64) * R];
  bool act[R];
  double2 beta[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    beta[s] = st[s].beta;
  }
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  const int xcd = blockIdx.x & 7;
  const int per_xcd = gridDim.x >> 3;  // grid is a multiple of 8
  const int lb = blockIdx.x >> 3;
  const int t_begin = xcd * xcd_span;
  const int t_end = min(ntiles, t_begin + xcd_span);
  for (int tile = t_begin + lb; tile < t_end; tile += per_xcd) {
    const int32_t r0 = tile_row[tile], r1 = tile_row[tile + 1];
    const int32_t kb = row_ptr[r0], ke = row_ptr[r1];
    constexpr int J = TNNZ / kBlock;
    double2 a[J];
    int32_t c[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int32_t k = kb + threadIdx.x + j * kBlock;
      if (k < ke) {
        a[j] = to_d2(val[k]);
        c[j] = col[k];
      } else {
        a[j] = make_double2(0, 0);
        c[j] = -1;
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int32_t lk = threadIdx.x + j * kBlock;
      if (c[j] >= 0) {
        const double2 *w = Wa + (int64_t)c[j] * 2 * R;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const double2 pc = cfma(beta[s], w[R + s], w[s]);
          prod[lk * R + s] = bil(a[j], pc);
        }
      }
    }
    __syncthreads();
    const int g = threadIdx.x / RL, l = threadIdx.x % RL;
    const int32_t row = r0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < r1) {
      const int32_t k0 = row_ptr[row] - kb, k1 = row_ptr[row + 1] - kb;
      for (int32_t k = k0 + l; k < k1; k += RL) {
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cadd(acc[s], prod[k * R + s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<RL>(acc[s]);
    if (row < r1 && l == 0) {
      const double2 *w = Wa + (int64_t)row * 2 * R;
      double2 *wb = Wb + (int64_t)row * 2 * R;
      // the whole record is written (z = 0; the update overwrites it) so that no partial
      // 64-byte line reaches HBM, which would need a read-modify-write
      double2 rec[2 * R];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        rec[s] = make_double2(0, 0);
        rec[R + s] = cfma(beta[s], w[R + s], w[s]);
      }
#pragma unroll
      for (int s = 0; s < 2 * R; ++s)
        wb[s] = rec[s];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        q[(int64_t)row * R + s] = acc[s];
        pq[s] = cadd(pq[s], bil(rec[R + s], acc[s]));
      }
    }
    __syncthreads();  // prod is reused by the next tile
  }
  write_partials<R>(pq, lds, partial);
}

// The tiled SpMV with the next tile's values, columns and row bounds loaded while the
// current tile's column gathers are in flight.  Every load of a tile is issued with no
// branch (entries past the tile's end read its first entry and contribute zero; rows past
// its end read its last row and write nothing), so the wait before the products covers
// the gathers only and the next tile's loads stay in flight across it.  desc[t] = (first
// row, end row, first nnz, end nnz) is read with uniform (scalar) loads.
template <typename VT, int R, int TNNZ, int RL>
__global__ void __launch_bounds__(kBlock)
    k_cocg_spmv_pipe(int ntiles, int xcd_span, const int4 *__restrict__ desc,
                     const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                     const VT *__restrict__ val, const double2 *__restrict__ Wa,
                     double2 *__restrict__ Wb, double2 *__restrict__ q,
                     double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 prod[TNNZ * R];
  __shared__ double2 lds[(kBlock / 64) * R];
  constexpr int J = TNNZ / kBlock;
  bool act[R];
  double2 beta[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    beta[s] = st[s].beta;
  }
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  const int xcd = blockIdx.x & 7;
  const int per_xcd = gridDim.x >> 3;
  const int lb = blockIdx.x >> 3;
  const int t_begin = xcd * xcd_span;
  const int t_end = min(ntiles, t_begin + xcd_span);
  const int g = threadIdx.x / RL, l = threadIdx.x % RL;
  // staged state of a tile: entries, columns, this thread's row bounds (relative)
  struct Stage {
    int4 d;
    double2 a[J];
    int32_t c[J];
    int32_t k0, k1;
  };
  auto stage = [&](Stage &x, int tile) {
    x.d = desc[tile];
    const int32_t kz = x.d.z < x.d.w ? x.d.z : 0;  // a valid entry either way
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int32_t k = x.d.z + threadIdx.x + j * kBlock;
      const bool in = k < x.d.w;
      const int32_t kk = in ? k : kz;
      const double2 av = ld_stream(val + kk);
      x.c[j] = ld_stream(col + kk);
      x.a[j] = in ? av : make_double2(0, 0);
    }
    const int32_t row = x.d.x + g;
    const int32_t rc = row < x.d.y ? row : x.d.y - 1;
    const int32_t rb = row_ptr[rc], re = row_ptr[rc + 1];
    x.k0 = rb - x.d.z;
    x.k1 = row < x.d.y ? re - x.d.z : x.k0;
  };
  int tile = t_begin + lb;
  if (tile >= t_end) {
    write_partials<R>(pq, lds, partial);
    return;
  }
  Stage cur, nxt;
  stage(cur, tile);
  for (; tile < t_end; tile += per_xcd) {
    // gathers of this tile and the records of its rows (lanes of a row read the same one)
    double2 wz[J][R], wp[J][R];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double2 *w = Wa + (int64_t)cur.c[j] * 2 * R;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        wz[j][s] = w[s];
        wp[j][s] = w[R + s];
      }
    }
    const int32_t row = cur.d.x + g;
    const int32_t rc = row < cur.d.y ? row : cur.d.y - 1;
    double2 rz[R], rp[R];
    {
      const double2 *w = Wa + (int64_t)rc * 2 * R;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        rz[s] = w[s];
        rp[s] = w[R + s];
      }
    }
    // the next tile's loads, behind the gathers
    stage(nxt, min(tile + per_xcd, t_end - 1));
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int32_t lk = threadIdx.x + j * kBlock;
#pragma unroll
      for (int s = 0; s < R; ++s)
        prod[lk * R + s] = bil(cur.a[j], cfma(beta[s], wp[j][s], wz[j][s]));
    }
    __syncthreads();
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    for (int32_t k = cur.k0 + l; k < cur.k1; k += RL)
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cadd(acc[s], prod[k * R + s]);
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<RL>(acc[s]);
    if (row < cur.d.y && l == 0) {
      double2 *wb = Wb + (int64_t)row * 2 * R;
      // the whole record is written (z = 0; the update overwrites it) so that no partial
      // 64-byte line reaches HBM
      double2 rec[2 * R];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        rec[s] = make_double2(0, 0);
        rec[R + s] = cfma(beta[s], rp[s], rz[s]);
      }
#pragma unroll
      for (int s = 0; s < 2 * R; ++s)
        wb[s] = rec[s];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        q[(int64_t)row * R + s] = acc[s];
        pq[s] = cadd(pq[s], bil(rec[R + s], acc[s]));
      }
    }
    __syncthreads();  // prod is reused by the next tile
    cur = nxt;
  }
  write_partials<R>(pq, lds, partial);
}

// x += alpha p, r -= alpha q, W_b.z = D^{-1} r; partials r^T z, r^H r
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocg_update(int64_t n, const double2 *__restrict__ dinv, double2 *__restrict__ Wb,
                  const double2 *__restrict__ q, double2 *__restrict__ x, double2 *__restrict__ r,
                  double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    v[s] = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double2 di = dinv[i];
    double2 *w = Wb + i * 2 * R;
    double2 rec[2 * R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      rec[s] = make_double2(0, 0);
      rec[R + s] = w[R + s];
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      const int64_t k = i * R + s;
      const double2 xv = cfma(alpha[s], rec[R + s], x[k]);
      const double2 rv = csub(r[k], bil(alpha[s], q[k]));
      const double2 zv = bil(di, rv);
      x[k] = xv;
      r[k] = rv;
      rec[s] = zv;
      v[s] = cadd(v[s], bil(rv, zv));
      v[R + s].x += cabs2(rv);
    }
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)  // whole record: no partial line to merge in HBM
      w[s] = rec[s];
  }
  write_partials<2 * R>(v, lds, partial);
}

// ---------------------------------------------------------------------------
// BiCGSTAB (van der Vorst 1992), right-preconditioned with D^{-1}:
//   rho_i = r0^H r; beta = (rho_i/rho_{i-1})(alpha/omega); p = r + beta (p - omega v)
//   y = D^{-1} p; v = A y; alpha = rho_i / (r0^H v); s = r - alpha v
//   zz = D^{-1} s; t = A zz; omega = (t^H s)/(t^H t); x += alpha y + omega zz; r = s - omega t
// ---------------------------------------------------------------------------
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_init(int64_t n, const double2 *__restrict__ b, double2 *__restrict__ x,
              double2 *__restrict__ r, double2 *__restrict__ r0, double2 *__restrict__ p,
              double2 *__restrict__ v, double2 *__restrict__ partial) {
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
  write_partials<R>(bb, lds, partial);
}

template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_dot_r0r(int64_t n, const double2 *__restrict__ r0, const double2 *__restrict__ r,
                 double2 *__restrict__ partial) {
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
  write_partials<R>(d, lds, partial);
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

// out = A in, partial dots: mode 0 -> (r0^H out); mode 1 -> (out^H s, out^H out)
template <typename VT, int R, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_bs_spmv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
              const VT *__restrict__ val, const double2 *__restrict__ in, double2 *__restrict__ out,
              const double2 *__restrict__ other, double2 *__restrict__ partial,
              const SysState *__restrict__ st) {
  constexpr int V = MODE == 0 ? R : 2 * R;
  __shared__ double2 lds[(kBlock / 64) * V];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[s].active != 0;
  const int g = threadIdx.x / kLanesPerRow;
  const int lane = threadIdx.x % kLanesPerRow;
  double2 d[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    d[s] = make_double2(0, 0);
  for (int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock; row0 < n;
       row0 += (int64_t)gridDim.x * kRowsPerBlock) {
    const int64_t row = row0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      const int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += kLanesPerRow) {
        const double2 a = to_d2(val[k]);
        const int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, in[c * R + s], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<kLanesPerRow>(acc[s]);
    if (row < n && lane == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        const int64_t k = row * R + s;
        out[k] = acc[s];
        const double2 o = other[k];
        if constexpr (MODE == 0) {
          d[s] = cadd(d[s], bil(cconj(o), acc[s]));
        } else {
          d[s] = cadd(d[s], bil(cconj(acc[s]), o));
          d[R + s].x += cabs2(acc[s]);
        }
      }
    }
  }
  write_partials<V>(d, lds, partial);
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

// x += alpha y + omega zz; r = s - omega t; partials r0^H r, ||r||^2
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_bs_x(int64_t n, const double2 *__restrict__ y, const double2 *__restrict__ zz,
           const double2 *__restrict__ t, const double2 *__restrict__ r0, double2 *__restrict__ x,
           double2 *__restrict__ r, double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R], omega[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
    omega[s] = st[s].omega;
  }
  double2 d[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    d[s] = make_double2(0, 0);
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
      d[s] = cadd(d[s], bil(cconj(r0[k]), rv));
      d[R + s].x += cabs2(rv);
    }
  }
  write_partials<2 * R>(d, lds, partial);
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
    const int64_t row = row0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (row < n) {
      const int32_t k0 = row_ptr[row], k1 = row_ptr[row + 1];
      for (int32_t k = k0 + lane; k < k1; k += kLanesPerRow) {
        const double2 a = to_d2(val[k]);
        const int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, x[c * R + s], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<kLanesPerRow>(acc[s]);
    if (row < n && lane == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s)
        y[row * R + s] = acc[s];
    }
  }
}


// ---------------------------------------------------------------------------
// Auxiliary preconditioner with a complex incidence (periodic systems): row e of the
// gradient G has up to two entries, ga[e] at node en2[e].x and gb[e] at node en2[e].y
// (-1 = none).  Slave nodes of a Floquet pair are folded onto their master with the
// phase, so G maps Bloch-periodic node potentials to reduced edge fields.
//   P r = D^{-1} r + G Dn^{-1} G^T r,  Dn = diag(G^T A G)   (bilinear transposes)
// ---------------------------------------------------------------------------
__device__ inline double2 caux_coef(const int2 *__restrict__ en2, const double2 *__restrict__ ga,
                                    const double2 *__restrict__ gb, int32_t f, int32_t node) {
  const int2 nn = en2[f];
  double2 c = make_double2(0, 0);
  if (nn.x == node)
    c = cadd(c, ga[f]);
  if (nn.y == node)
    c = cadd(c, gb[f]);
  return c;
}

template <typename VT>
__global__ void k_caux_diag(int64_t N, const int32_t *__restrict__ gt_ptr, const int32_t *__restrict__ gt_edge,
                            const double2 *__restrict__ gt_coef, const int2 *__restrict__ en2,
                            const double2 *__restrict__ ga, const double2 *__restrict__ gb,
                            const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                            const VT *__restrict__ val, double2 *__restrict__ dninv) {
  const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  double2 sum = make_double2(0, 0);
  for (int32_t q = gt_ptr[nd]; q < gt_ptr[nd + 1]; ++q) {
    const int32_t e = gt_edge[q];
    double2 inner = make_double2(0, 0);
    for (int32_t k = row_ptr[e]; k < row_ptr[e + 1]; ++k) {
      const double2 c = caux_coef(en2, ga, gb, col[k], (int32_t)nd);
      if (c.x != 0.0 || c.y != 0.0)
        inner = cfma(to_d2(val[k]), c, inner);
    }
    sum = cfma(gt_coef[q], inner, sum);
  }
  dninv[nd] = (sum.x == 0.0 && sum.y == 0.0) ? make_double2(0, 0) : cdiv(make_double2(1, 0), sum);
}

// y = Dn^{-1} G^T in
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_caux_node(int64_t N, const int32_t *__restrict__ gt_ptr, const int32_t *__restrict__ gt_edge,
                const double2 *__restrict__ gt_coef, const double2 *__restrict__ dninv,
                const double2 *__restrict__ in, double2 *__restrict__ y, const SysState *__restrict__ st) {
  const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  double2 acc[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    acc[s] = make_double2(0, 0);
  for (int32_t q = gt_ptr[nd]; q < gt_ptr[nd + 1]; ++q) {
    const int64_t e = gt_edge[q];
    const double2 g = gt_coef[q];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = cfma(g, in[e * R + s], acc[s]);
  }
  const double2 di = dninv[nd];
#pragma unroll
  for (int s = 0; s < R; ++s)
    if (st == nullptr || st[s].active)
      y[nd * R + s] = cmul(di, acc[s]);
}

// out = D^{-1} in + G y
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_caux_edge(int64_t n, const int2 *__restrict__ en2, const double2 *__restrict__ ga,
                const double2 *__restrict__ gb, const double2 *__restrict__ dinv, const double2 *__restrict__ y,
                const double2 *__restrict__ in, double2 *__restrict__ out, const SysState *__restrict__ st) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  const int2 nn = en2[e];
  const double2 di = dinv[e];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    if (st != nullptr && !st[s].active)
      continue;
    double2 v = cmul(di, in[e * R + s]);
    if (nn.x >= 0)
      v = cfma(ga[e], y[(int64_t)nn.x * R + s], v);
    if (nn.y >= 0)
      v = cfma(gb[e], y[(int64_t)nn.y * R + s], v);
    out[e * R + s] = v;
  }
}

// ---------------------------------------------------------------------------
// Low-rank term of the operator: A_eff = A + sum_p c_p w_p w_p^H, with w_p given on a
// compact list of edges (the rank-1 port loading of src/assemble_maxwell.cpp:209-242,
// applied as an operator instead of dense blocks in the sparsity pattern).  The ports
// are disjoint (checked on the host), so every list entry owns its row.  One block:
//   t_{p,s} = w_p^H in_s;  out_s += c_p w_p t_{p,s} on the port rows,
// and the change of the dot products the preceding SpMV wrote into its partials:
//   MODE 0 (COCG):          sum in_s . dout_s           (bilinear)       -> R values
//   MODE 1 (BiCGSTAB v):    sum conj(other_s) dout_s                     -> R values
//   MODE 2 (BiCGSTAB t):    sum conj(dout_s) other_s, |out_s|^2 change   -> 2R values
//   MODE 3 (plain SpMV):    none
// written as one extra partial block, so the finalisation sums one block more.
// ---------------------------------------------------------------------------
constexpr int kMaxPorts = 32;  // low-rank terms

template <int R, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_lowrank(int P, int64_t len, const int32_t *__restrict__ idx, const double2 *__restrict__ wv,
              const double2 *__restrict__ coef, const double2 *__restrict__ in, int in_stride, int in_off,
              double2 *__restrict__ out, const double2 *__restrict__ other, double2 *__restrict__ slot,
              const SysState *__restrict__ st) {
  constexpr int V = MODE == 2 ? 2 * R : R;
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxRhs];
  __shared__ double2 tsh[kMaxPorts * kMaxRhs];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = (MODE == 3) || st[s].active != 0;
  for (int p = 0; p < P; ++p) {
    double2 v[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      v[s] = make_double2(0, 0);
    for (int64_t i = threadIdx.x; i < len; i += kBlock) {
      const int32_t r = idx[p * len + i];
      if (r < 0)
        continue;
      const double2 w = cconj(wv[p * len + i]);
#pragma unroll
      for (int s = 0; s < R; ++s)
        v[s] = cfma(w, in[(int64_t)r * in_stride + in_off + s], v[s]);
    }
    block_sum<R>(v, lds);
    if (threadIdx.x < R)
      tsh[p * R + threadIdx.x] = lds[threadIdx.x];
    __syncthreads();
  }
  double2 c[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    c[s] = make_double2(0, 0);
  for (int64_t i = threadIdx.x; i < (int64_t)P * len; i += kBlock) {
    const int32_t r = idx[i];
    if (r < 0)
      continue;
    const int p = (int)(i / len);
    const double2 cw = cmul(coef[p], wv[i]);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      const int64_t k = (int64_t)r * R + s;
      const double2 d = cmul(cw, tsh[p * R + s]);
      const double2 o = out[k];
      const double2 nw = cadd(o, d);
      out[k] = nw;
      if constexpr (MODE == 0) {
        c[s] = cfma(in[(int64_t)r * in_stride + in_off + s], d, c[s]);
      } else if constexpr (MODE == 1) {
        c[s] = cfma(cconj(other[k]), d, c[s]);
      } else if constexpr (MODE == 2) {
        c[s] = cfma(cconj(d), other[k], c[s]);
        c[R + s].x += cabs2(nw) - cabs2(o);
      }
    }
  }
  if constexpr (MODE != 3) {
    block_sum<V>(c, lds);
    if (threadIdx.x < V)
      slot[threadIdx.x] = lds[threadIdx.x];
  }
}

// Jacobi diagonal of A_eff on the port rows: a_rr + c_p |w_{p,r}|^2.
template <typename VT>
__global__ void k_lowrank_dinv(int64_t m, int64_t len, const int32_t *__restrict__ idx,
                               const double2 *__restrict__ wv, const double2 *__restrict__ coef,
                               const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                               const VT *__restrict__ val, double2 *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m)
    return;
  const int32_t r = idx[i];
  if (r < 0)
    return;
  double2 d = make_double2(0, 0);
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    if (col[k] == r)
      d = to_d2(val[k]);
  d = cadd(d, cscale(coef[i / len], cabs2(wv[i])));
  dinv[r] = (d.x == 0.0 && d.y == 0.0) ? make_double2(1.0, 0.0) : cdiv(make_double2(1.0, 0.0), d);
}

// ---------------------------------------------------------------------------
// Auxiliary-space (discrete gradient) preconditioner
//   P r = D^{-1} r + G Dn^{-1} G^T r,   Dn = diag(G^T A G)
// G is the edge-node incidence of the discrete gradient (-1 at the lower-ID end,
// +1 at the higher-ID end), with Dirichlet edges removed.  P is complex symmetric,
// so COCG stays applicable.  The curl-free error components, on which the edge
// Jacobi step acts with the wrong scale (A ~ -k0^2 M there, diag(A) ~ K), are
// corrected on the nodes.
// ---------------------------------------------------------------------------
__device__ inline int grad_sign(const int32_t *__restrict__ en, int32_t e, int32_t node) {
  const int32_t a = en[2 * (int64_t)e], b = en[2 * (int64_t)e + 1];
  if (a == b)
    return 0;
  return (node == b) ? 1 : ((node == a) ? -1 : 0);
}

template <typename VT>
__global__ void k_aux_diag(int64_t N, const int32_t *__restrict__ n2e_ptr,
                           const int32_t *__restrict__ n2e, const int32_t *__restrict__ en,
                           const uint8_t *__restrict__ dir, const int32_t *__restrict__ row_ptr,
                           const int32_t *__restrict__ col, const VT *__restrict__ val,
                           double2 *__restrict__ dninv) {
  int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  double2 d = make_double2(0, 0);
  for (int32_t k = n2e_ptr[nd]; k < n2e_ptr[nd + 1]; ++k) {
    const int32_t e = n2e[k];
    if (dir && dir[e])
      continue;
    const int ge = grad_sign(en, e, (int32_t)nd);
    if (!ge)
      continue;
    for (int32_t q = row_ptr[e]; q < row_ptr[e + 1]; ++q) {
      const int32_t f = col[q];
      if (dir && dir[f])
        continue;
      const int gf = grad_sign(en, f, (int32_t)nd);
      if (!gf)
        continue;
      const double2 a = to_d2(val[q]);
      const double sg = (double)(ge * gf);
      d.x += sg * a.x;
      d.y += sg * a.y;
    }
  }
  dninv[nd] = (d.x == 0.0 && d.y == 0.0) ? make_double2(0, 0) : cdiv(make_double2(1, 0), d);
}

// Per-incidence gradient entries G[e, n] in {-1, 0, +1} (0 for Dirichlet or degenerate
// edges) and per-edge endpoints with the same masking folded in (-1, -1).
__global__ void k_aux_prep_nodes(int64_t N, const int32_t *__restrict__ n2e_ptr,
                                 const int32_t *__restrict__ n2e, const int32_t *__restrict__ en,
                                 const uint8_t *__restrict__ dir, int8_t *__restrict__ gsign) {
  int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  for (int32_t k = n2e_ptr[nd]; k < n2e_ptr[nd + 1]; ++k) {
    const int32_t e = n2e[k];
    gsign[k] = (dir && dir[e]) ? 0 : (int8_t)grad_sign(en, e, (int32_t)nd);
  }
}

__global__ void k_aux_prep_edges(int64_t n, const int32_t *__restrict__ en,
                                 const uint8_t *__restrict__ dir, int2 *__restrict__ en_eff) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n)
    return;
  const int32_t a = en[2 * e], b = en[2 * e + 1];
  const bool off = (dir && dir[e]) || a == b;
  en_eff[e] = off ? make_int2(-1, -1) : make_int2(a, b);
}

// y[n] = Dn^{-1} sum_e G[e,n] (r[e] - alpha q[e])   (alpha, q only when FUSED).
// Eight lanes per node stride over its incidence list; fixed-order shuffle reduction.
constexpr int kLanesPerNode = 8;
template <int R, bool FUSED>
__global__ void __launch_bounds__(kBlock)
    k_aux_node(int64_t N, const int32_t *__restrict__ n2e_ptr, const int32_t *__restrict__ n2e,
               const int8_t *__restrict__ gsign, const double2 *__restrict__ dninv,
               const double2 *__restrict__ r, const double2 *__restrict__ q,
               double2 *__restrict__ y, const SysState *__restrict__ st) {
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st == nullptr || st[s].active != 0;
    alpha[s] = FUSED ? st[s].alpha : make_double2(0, 0);
  }
  const int g = threadIdx.x / kLanesPerNode, lane = threadIdx.x % kLanesPerNode;
  constexpr int kNodesPerBlock = kBlock / kLanesPerNode;
  // XCD-aware walk (grid a multiple of 8): block b runs on XCD b % 8 and takes chunks of
  // that XCD's contiguous node range, so an edge gathered by both of its end nodes is
  // usually found in the same L2 the second time
  const int64_t nchunks = (N + kNodesPerBlock - 1) / kNodesPerBlock;
  const int64_t span = (nchunks + 7) / 8;
  const int64_t c_end = min(nchunks, ((int64_t)(blockIdx.x & 7) + 1) * span);
  for (int64_t ch = (int64_t)(blockIdx.x & 7) * span + (blockIdx.x >> 3); ch < c_end; ch += gridDim.x >> 3) {
    const int64_t nd = ch * kNodesPerBlock + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (nd < N) {
      for (int32_t k = n2e_ptr[nd] + lane; k < n2e_ptr[nd + 1]; k += kLanesPerNode) {
        const int32_t e = n2e[k];
        const int gs = gsign[k];
        if (!gs)
          continue;
        const double sg = (double)gs;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int64_t i = (int64_t)e * R + s;
          double2 rv = r[i];
          if (FUSED)
            rv = csub(rv, bil(alpha[s], q[i]));
          acc[s].x += sg * rv.x;
          acc[s].y += sg * rv.y;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<kLanesPerNode>(acc[s]);
    if (nd < N && lane == 0) {
      const double2 di = dninv[nd];
#pragma unroll
      for (int s = 0; s < R; ++s)
        if (act[s])
          y[nd * R + s] = bil(di, acc[s]);
    }
  }
}

// COCG update with the auxiliary preconditioner:
//   r <- r - alpha q; x <- x + alpha p; W_b.z = D^{-1} r + (y[head] - y[tail]);
//   partials r^T z, r^H r
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_aux_cocg_update(int64_t n, const int2 *__restrict__ en_eff,
                      const double2 *__restrict__ dinv, const double2 *__restrict__ y,
                      double2 *__restrict__ Wb, const double2 *__restrict__ q,
                      double2 *__restrict__ x, double2 *__restrict__ r,
                      double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    v[s] = make_double2(0, 0);
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const double2 di = dinv[e];
    const int2 ab = en_eff[e];
    double2 *w = Wb + e * 2 * R;
    double2 rec[2 * R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      rec[s] = make_double2(0, 0);
      rec[R + s] = w[R + s];
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      const int64_t k = e * R + s;
      const double2 xv = cfma(alpha[s], rec[R + s], x[k]);
      const double2 rv = csub(r[k], bil(alpha[s], q[k]));
      double2 zv = bil(di, rv);
      if (ab.x >= 0)
        zv = cadd(zv, csub(y[(int64_t)ab.y * R + s], y[(int64_t)ab.x * R + s]));
      x[k] = xv;
      r[k] = rv;
      rec[s] = zv;
      v[s] = cadd(v[s], bil(rv, zv));
      v[R + s].x += cabs2(rv);
    }
#pragma unroll
    for (int s = 0; s < 2 * R; ++s)  // whole record: no partial line to merge in HBM
      w[s] = rec[s];
  }
  write_partials<2 * R>(v, lds, partial);
}

// out = D^{-1} in + G y (plain application, for the initial residual and BiCGSTAB);
// with PARTIALS also the partials in^T out and in^H in.
template <int R, bool PARTIALS>
__global__ void __launch_bounds__(kBlock)
    k_aux_apply_edge(int64_t n, const int2 *__restrict__ en_eff,
                     const double2 *__restrict__ dinv, const double2 *__restrict__ y,
                     const double2 *__restrict__ in, double2 *__restrict__ out, int ostride,
                     double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st == nullptr || st[s].active != 0;
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    v[s] = make_double2(0, 0);
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const double2 di = dinv[e];
    const int2 ab = en_eff[e];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      const int64_t k = e * R + s;
      const double2 iv = in[k];
      double2 o = bil(di, iv);
      if (ab.x >= 0)
        o = cadd(o, csub(y[(int64_t)ab.y * R + s], y[(int64_t)ab.x * R + s]));
      out[e * ostride + s] = o;
      if (PARTIALS) {
        v[s] = cadd(v[s], bil(iv, o));
        v[R + s].x += cabs2(iv);
      }
    }
  }
  if (PARTIALS)
    write_partials<2 * R>(v, lds, partial);
}

// ---------------------------------------------------------------------------
// COCG with the search direction formed in its own pass (variant 9): per iteration
//   p = z + beta p           (k_cocgz_pform: a stream over z and p)
//   q = A p, partials p^T q  (k_cocgz_spmv: gathers p, R values per column in an array of
//                             its own, and writes q only)
//   x += alpha p, r -= alpha q, z = P r, partials   (k_cocgz_update)
// The fused forms build p inside the SpMV and gather a {z, p} record (twice the gathered
// bytes per nonzero).  Carrying q by the recurrence q_new = A z + beta q_old instead
// (gathering z) saves the p pass but adds iterations on these indefinite systems: on the
// 1.59 M-edge guide 4,489 iterations with an exact q every 8th, 4,129 every 2nd, against
// 4,030 here, which is the fastest of the four per solve.
// ---------------------------------------------------------------------------
// after the fused-form init: zf = W0.z, pf = 0, q = 0
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_init(int64_t n, const double2 *__restrict__ W, double2 *__restrict__ zf, double2 *__restrict__ pf,
                 double2 *__restrict__ q) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * R; i += (int64_t)gridDim.x * kBlock) {
    const int64_t row = i / R, sys = i - row * R;
    zf[i] = W[row * 2 * R + sys];
    pf[i] = make_double2(0, 0);
    q[i] = make_double2(0, 0);
  }
}

// q = A p with p gathered from its own array (R values per column), tiled and pipelined
// (the tile walk, staging and branch-free loads of k_cocg_spmv_pipe); partials p^T q.
// SPLIT (variant 11): the values are read as a real part per entry (vre) and, for the
// tiles with an imaginary part anywhere (imoff[tile] >= 0), an imaginary part per entry of
// that tile (vim + imoff[tile]); a tile whose values are all real streams 8 B per value
// instead of 16.  val is unused then.
template <typename VT, int R, int TNNZ, int RL, bool SPLIT = false>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_spmv(int ntiles, int xcd_span, const int4 *__restrict__ desc, const int32_t *__restrict__ row_ptr,
                 const int32_t *__restrict__ col, const VT *__restrict__ val, const double2 *__restrict__ pf,
                 double2 *__restrict__ q, double2 *__restrict__ partial, const SysState *__restrict__ st,
                 const double *__restrict__ vre = nullptr, const double *__restrict__ vim = nullptr,
                 const int32_t *__restrict__ imoff = nullptr) {
  // Products in one plane per system: a wavefront's writes are contiguous 16 B runs (no
  // bank conflicts), and a row's lanes read consecutive 16 B slots.
  __shared__ double2 prod[R * TNNZ];
  __shared__ double2 lds[(kBlock / 64) * R];
  constexpr int J = TNNZ / kBlock;
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[s].active != 0;
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  const int xcd = blockIdx.x & 7;
  const int per_xcd = gridDim.x >> 3;
  const int lb = blockIdx.x >> 3;
  const int t_begin = xcd * xcd_span;
  const int t_end = min(ntiles, t_begin + xcd_span);
  const int g = threadIdx.x / RL, l = threadIdx.x % RL;
  struct Stage {
    int4 d;
    double2 a[J];
    int32_t c[J];
    int32_t k0, k1;
  };
  auto stage = [&](Stage &x, int tile) {
    x.d = desc[tile];
    const int32_t kz = x.d.z < x.d.w ? x.d.z : 0;  // a valid entry either way
    if constexpr (SPLIT) {
      const int32_t io = imoff[tile];  // uniform over the block
      if (io >= 0) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int32_t k = x.d.z + threadIdx.x + j * kBlock;
          const bool in = k < x.d.w;
          const int32_t kk = in ? k : kz;
          const double ar = ld_stream(vre + kk);
          const double ai = ld_stream(vim + io + (kk - kz));
          x.c[j] = ld_stream(col + kk);
          x.a[j].x = in ? ar : 0.0;
          x.a[j].y = in ? ai : 0.0;
        }
      } else {
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int32_t k = x.d.z + threadIdx.x + j * kBlock;
          const bool in = k < x.d.w;
          const int32_t kk = in ? k : kz;
          const double ar = ld_stream(vre + kk);
          x.c[j] = ld_stream(col + kk);
          x.a[j].x = in ? ar : 0.0;
          x.a[j].y = 0.0;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int32_t k = x.d.z + threadIdx.x + j * kBlock;
        const bool in = k < x.d.w;
        const int32_t kk = in ? k : kz;
        const double2 av = ld_stream(val + kk);
        x.c[j] = ld_stream(col + kk);
        x.a[j].x = in ? av.x : 0.0;
        x.a[j].y = in ? av.y : 0.0;
      }
    }
    const int32_t row = x.d.x + g;
    const int32_t rc = row < x.d.y ? row : x.d.y - 1;
    const int32_t rb = row_ptr[rc], re = row_ptr[rc + 1];
    x.k0 = rb - x.d.z;
    x.k1 = row < x.d.y ? re - x.d.z : x.k0;
  };
  int tile = t_begin + lb;
  if (tile >= t_end) {
    write_partials<R>(pq, lds, partial);
    return;
  }
  Stage cur, nxt;
  stage(cur, tile);
  for (; tile < t_end; tile += per_xcd) {
    double2 zg[J][R];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const double2 *pc = pf + (int64_t)cur.c[j] * R;
#pragma unroll
      for (int s = 0; s < R; ++s)
        zg[j][s] = pc[s];
    }
    const int32_t row = cur.d.x + g;
    const int32_t rc = row < cur.d.y ? row : cur.d.y - 1;
    double2 pr[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      pr[s] = pf[(int64_t)rc * R + s];
    stage(nxt, min(tile + per_xcd, t_end - 1));
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int32_t lk = threadIdx.x + j * kBlock;
#pragma unroll
      for (int s = 0; s < R; ++s)
        prod[s * TNNZ + lk] = bil(cur.a[j], zg[j][s]);
    }
    __syncthreads();
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    for (int32_t k = cur.k0 + l; k < cur.k1; k += RL)
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cadd(acc[s], prod[s * TNNZ + k]);
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<RL>(acc[s]);
    if (row < cur.d.y && l == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        q[(int64_t)row * R + s] = acc[s];
        pq[s] = cadd(pq[s], bil(pr[s], acc[s]));
      }
    }
    __syncthreads();  // prod is reused by the next tile
    cur = nxt;
  }
  write_partials<R>(pq, lds, partial);
}

// p = z + beta p.  XLAG: also the previous iteration's x += alpha p, made here where p is
// read anyway instead of in the update (which then reads neither x nor p): one vector
// pass fewer per iteration.  The same cfma of the same operands, so x is bit-identical;
// the last iteration's update is made by k_cocgz_xflush after the loop.
template <int R, bool XLAG = false>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_pform(int64_t n, const double2 *__restrict__ zf, double2 *__restrict__ pf,
                  const SysState *__restrict__ st, double2 *__restrict__ x = nullptr) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * R; i += (int64_t)gridDim.x * kBlock) {
    const int sys = (int)(i % R);
    if (st[sys].active) {
      const double2 po = pf[i];
      if (XLAG && st[sys].iters > 0)
        x[i] = cfma(st[sys].alpha, po, x[i]);
      pf[i] = cfma(st[sys].beta, po, zf[i]);
    }
  }
}

// The pending x += alpha p of the last iteration of every system that made one (after a
// breakdown in the alpha step alpha is zero).
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_xflush(int64_t n, const double2 *__restrict__ pf, double2 *__restrict__ x,
                   const SysState *__restrict__ st) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * R; i += (int64_t)gridDim.x * kBlock) {
    const int sys = (int)(i % R);
    if (st[sys].iters > 0)
      x[i] = cfma(st[sys].alpha, pf[i], x[i]);
  }
}

// x += alpha p (not with XLAG: made by the next p pass), r -= alpha q, z = D^{-1} r (+ G y
// with the auxiliary preconditioner); partials r^T z, r^H r.  One (row, system) element per
// thread, so every load and store of a wavefront is one contiguous run (no partial lines).
template <int R, bool AUX, bool XLAG = false>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_update(int64_t n, const int2 *__restrict__ en_eff, const double2 *__restrict__ dinv,
                   const double2 *__restrict__ y, const double2 *__restrict__ pf, const double2 *__restrict__ q,
                   double2 *__restrict__ zf, double2 *__restrict__ x, double2 *__restrict__ r,
                   double2 *__restrict__ partial, const SysState *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * R];
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    act[s] = st[s].active != 0;
    alpha[s] = st[s].alpha;
  }
  double2 v[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    v[s] = make_double2(0, 0);
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n * R; k += (int64_t)gridDim.x * kBlock) {
    const int64_t e = k / R;
    const int sys = (int)(k - e * R);
    double2 al = make_double2(0, 0);
    bool a = false;
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        al = alpha[s];
        a = act[s];
      }
    if (!a)
      continue;
    const double2 di = dinv[e];
    const double2 rv = csub(r[k], bil(al, q[k]));
    double2 zv = bil(di, rv);
    if (AUX) {
      const int2 ab = en_eff[e];
      // both node values loaded unconditionally (node 0 for an edge without nodes), so they
      // issue with the row's other loads; the sum is taken only for edges with nodes
      const bool hn = ab.x >= 0;
      const double2 yd = csub(y[(int64_t)(hn ? ab.y : 0) * R + sys], y[(int64_t)(hn ? ab.x : 0) * R + sys]);
      zv = hn ? cadd(zv, yd) : zv;
    }
    if (!XLAG)
      x[k] = cfma(al, pf[k], x[k]);
    r[k] = rv;
    zf[k] = zv;
    const double2 rz = bil(rv, zv);
    const double rr = cabs2(rv);
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        v[s] = cadd(v[s], rz);
        v[R + s].x += rr;
      }
  }
  write_partials<2 * R>(v, lds, partial);
}

// x = 0, r = b, W0.p = 0 (W0.z comes from the aux kernels)
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cocg_init_aux(int64_t n, const double2 *__restrict__ b, double2 *__restrict__ x,
                    double2 *__restrict__ r, double2 *__restrict__ W) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * R;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] = make_double2(0, 0);
    r[i] = b[i];
    const int64_t row = i / R, s = i - row * R;
    W[row * 2 * R + R + s] = make_double2(0, 0);
  }
}

// ---------------------------------------------------------------------------
// Sliced-ELL SpMV of the p-pass form (variant 10): q = A p, partials p^T q.
// Rows are grouped in slices of 64/RL rows; a row's entries are spread over RL adjacent
// lanes (entry e of a row at step e / RL, lane e % RL of the row's lane group), and the
// slice's entries are stored step-major, so every step is one contiguous 64-entry load of
// values and columns.  Rows are taken in order of decreasing length inside windows of
// 4096 rows, so a slice pads little.  Each lane group sums its row in registers and
// reduces it with shuffles: no LDS, no barrier, so occupancy is set by registers only.
// Padding slots point at column 0 with value 0.  Slices are walked XCD-contiguously.
// ---------------------------------------------------------------------------
template <typename VT, int R, int RL>
__global__ void __launch_bounds__(kBlock)
    k_cocgz_sell(int nslices, int xcd_span, const int32_t *__restrict__ soff, const int32_t *__restrict__ srow,
                 const int32_t *__restrict__ scol, const VT *__restrict__ sval, const double2 *__restrict__ pf,
                 double2 *__restrict__ q, double2 *__restrict__ partial, const SysState *__restrict__ st) {
  constexpr int SR = 64 / RL;
  __shared__ double2 lds[(kBlock / 64) * R];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[s].active != 0;
  double2 pq[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    pq[s] = make_double2(0, 0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int xcd = blockIdx.x & 7;
  const int sb = xcd * xcd_span, se = min(nslices, sb + xcd_span);
  const int wstride = (gridDim.x >> 3) * (kBlock / 64);
  for (int sl = sb + (blockIdx.x >> 3) * (kBlock / 64) + wv; sl < se; sl += wstride) {
    const int o0 = soff[sl], o1 = soff[sl + 1];
    const int32_t row = srow[sl * SR + lane / RL];
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    int o = o0;
    for (; o + 1 < o1; o += 2) {
      const int64_t i0 = (int64_t)o * 64 + lane;
      const int32_t c0 = ld_stream(scol + i0), c1 = ld_stream(scol + i0 + 64);
      const double2 a0 = ld_stream(sval + i0), a1 = ld_stream(sval + i0 + 64);
      double2 g0[R], g1[R];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        g0[s] = pf[(int64_t)c0 * R + s];
        g1[s] = pf[(int64_t)c1 * R + s];
      }
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cfma(a1, g1[s], cfma(a0, g0[s], acc[s]));
    }
    if (o < o1) {
      const int64_t i0 = (int64_t)o * 64 + lane;
      const int32_t c0 = ld_stream(scol + i0);
      const double2 a0 = ld_stream(sval + i0);
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cfma(a0, pf[(int64_t)c0 * R + s], acc[s]);
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<RL>(acc[s]);
    if (row >= 0 && (lane % RL) == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        q[(int64_t)row * R + s] = acc[s];
        pq[s] = cadd(pq[s], bil(pf[(int64_t)row * R + s], acc[s]));
      }
    }
  }
  write_partials<R>(pq, lds, partial);
}

// Sliced-ELL structure from the CSR pattern, one wave per slice: for every slot the column
// (0 for padding) and the CSR position of its value (-1 for padding).
template <int RL>
__global__ void __launch_bounds__(kBlock)
    k_sell_build(int nslices, const int32_t *__restrict__ soff, const int32_t *__restrict__ srow,
                 const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col, int32_t *__restrict__ scol,
                 int32_t *__restrict__ svidx) {
  constexpr int SR = 64 / RL;
  const int sl = (int)(((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (sl >= nslices)
    return;
  const int32_t row = srow[sl * SR + lane / RL];
  const int32_t k0 = row >= 0 ? row_ptr[row] : 0, k1 = row >= 0 ? row_ptr[row + 1] : 0;
  for (int o = soff[sl], j = 0; o < soff[sl + 1]; ++o, ++j) {
    const int32_t k = k0 + j * RL + lane % RL;
    const bool in = row >= 0 && k < k1;
    const int64_t i = (int64_t)o * 64 + lane;
    scol[i] = in ? col[k] : 0;
    svidx[i] = in ? k : -1;
  }
}

// Sliced-ELL values from the CSR values (run after every matrix install).
template <typename VT>
__global__ void k_sell_vals(int64_t slots, const int32_t *__restrict__ svidx, const VT *__restrict__ val,
                            VT *__restrict__ sval) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = svidx[i];
    VT v;
    if (k >= 0) {
      v = val[k];
    } else {
      v.x = 0;
      v.y = 0;
    }
    sval[i] = v;
  }
}

} // namespace

// Split storage of variant 11: flag[t] = 1 if an entry of tile t (first tiling) has a
// nonzero imaginary part; one block per tile.
__global__ void __launch_bounds__(kBlock)
    k_split_flags(int ntiles, const int4 *__restrict__ desc, const double2 *__restrict__ val,
                  int32_t *__restrict__ flag) {
  const int t = blockIdx.x;
  if (t >= ntiles)
    return;
  const int4 d = desc[t];
  int any = 0;
  for (int32_t k = d.z + threadIdx.x; k < d.w; k += kBlock)
    any |= val[k].y != 0.0;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0)
    flag[t] = any;
}

// vre[k] = Re val[k] for every entry; vim[imoff[t] + k - d.z] = Im val[k] for the tiles with
// imoff[t] >= 0.  One block per tile.
__global__ void __launch_bounds__(kBlock)
    k_split_fill(int ntiles, const int4 *__restrict__ desc, const int32_t *__restrict__ imoff,
                 const double2 *__restrict__ val, double *__restrict__ vre, double *__restrict__ vim) {
  const int t = blockIdx.x;
  if (t >= ntiles)
    return;
  const int4 d = desc[t];
  const int32_t io = imoff[t];
  for (int32_t k = d.z + threadIdx.x; k < d.w; k += kBlock) {
    const double2 v = val[k];
    vre[k] = v.x;
    if (io >= 0)
      vim[io + (k - d.z)] = v.y;
  }
}

constexpr int kNumSpmvVariants = 12;  // rows<4>, rows<8>, rows<16>, tiled<T,4>, tiled<T/2,8>, (5 removed), pipe<T,4>, (7, 8 removed), p-pass form, p-pass form with sliced ELL, p-pass form with split real/imaginary values
// Numbers 5 (cooperative record gather), 7 (one cache line per nonzero) and 8 (tile-local
// column format) belonged to variants that were slower on every matrix measured; they were
// removed and their numbers are not reused.
constexpr bool kRetired(int v) { return v == 5 || v == 7 || v == 8; }
constexpr int kZForm = 9;             // COCG with p formed in its own pass (its own SpMV and update)
constexpr int kSellForm = 10;         // the same iteration with the sliced-ELL SpMV
constexpr int kSplitForm = 11;        // the same iteration, values stored as real parts plus the imaginary parts of complex tiles
constexpr int kSellRL = 4;            // lanes per row of the sliced-ELL SpMV

struct efm_solver {
  int64_t n = 0, nnz = 0;
  const int32_t *row_ptr = nullptr;
  const int32_t *col = nullptr;
  int R = 1;
  int vtype = EFM_C128;
  int grid = 1024;       // streaming kernels (grid-stride)
  int grid_spmv = 1024;  // row-group kernels (grid-stride)
  efm::DevBuf val, dinv, x, r, q, W0, W1, p0, r0, v, y, zz, t, partial, state, control;
  // row tiles of the tiled SpMV variants (ntiles = 0: a row is longer than a tile)
  efm::DevBuf tile_row[2];
  efm::DevBuf tile_desc;  // int4 (row0, row1, nnz0, nnz1) per tile of the first tiling
  int ntiles[2] = {0, 0}, xcd_span[2] = {0, 0}, grid_tiles[2] = {8, 8};
  int grid_tiles_res[2] = {8, 8};  // resident-block grids of the tiled variants
  int grid_pipe_res = 8;
  // variant 9 (p formed in its own pass): z and p arrays [n][R]; resident grid of its SpMV
  efm::DevBuf zf, pf;
  int grid_z = 8;
  // variant 10 (sliced ELL, built on first use): slice offsets in 64-slot steps, the row of
  // every lane group, columns and CSR value positions per slot, values per slot
  efm::DevBuf sell_off, sell_row, sell_col, sell_vidx, sell_val;
  int sell_nslices = 0, sell_xcd_span = 0, grid_sell = 8;
  int64_t sell_slots = 0;
  bool sell_built = false, sell_vals_ok = false;
  // variant 11 (split values, built on first use and after every set_matrix): real parts of
  // every entry, imaginary parts of the tiles that have any, the offset of each tile's
  // imaginary parts (-1: a real tile)
  efm::DevBuf sp_re, sp_im, sp_off;
  int64_t sp_im_n = 0;
  bool sp_ok = false;
  int grid_rows[3] = {8, 8, 8};
  int spmv_variant = -1;  // -1: choose by timing at the first solve
  double tune_ms[kNumSpmvVariants] = {0};
  // auxiliary gradient preconditioner (non-owning mesh arrays)
  bool aux = false;
  int64_t n_nodes = 0;
  int grid_nodes = 1;
  const int32_t *edge_nodes = nullptr, *n2e_ptr = nullptr, *n2e = nullptr;
  const uint8_t *dir = nullptr;
  efm::DevBuf dninv, ynode, gsign, en_eff;
  int grid_node_groups = 1;
  // complex-incidence auxiliary preconditioner (periodic systems)
  bool caux = false;
  int64_t caux_N = 0;
  efm::DevBuf caux_en2, caux_ga, caux_gb, caux_gt_ptr, caux_gt_edge, caux_gt_coef, caux_dninv, caux_y;
  // low-rank term (ports): P lists of len entries
  int lr_P = 0;
  int64_t lr_len = 0;
  efm::DevBuf lr_idx, lr_val, lr_coef;
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
  template <int Stage_>
  static int fin(efm_solver *s, int nb, hipStream_t st, double tol2, int max_it) {
    hipLaunchKernelGGL((k_finalize<R, Stage_>), dim3(1), dim3(kBlock), 0, st, s->partial.as<double2>(),
                       nb, s->state.as<SysState>(), s->control.as<Control>(), tol2, max_it);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // y = Dn^{-1} G^T in (nodes), then out = D^{-1} in + G y; partials when asked
  template <bool PARTIALS>
  static int aux_apply(efm_solver *s, const double2 *in, double2 *out, int ostride, hipStream_t st,
                       const SysState *S) {
    if (s->caux) {
      hipLaunchKernelGGL((k_caux_node<R>), dim3(grid_for(s->caux_N, kBlock)), dim3(kBlock), 0, st, s->caux_N,
                         s->caux_gt_ptr.as<int32_t>(), s->caux_gt_edge.as<int32_t>(), s->caux_gt_coef.as<double2>(),
                         s->caux_dninv.as<double2>(), in, s->caux_y.as<double2>(), S);
      EFM_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_caux_edge<R>), dim3(grid_for(s->n, kBlock)), dim3(kBlock), 0, st, s->n,
                         s->caux_en2.as<int2>(), s->caux_ga.as<double2>(), s->caux_gb.as<double2>(),
                         s->dinv.as<double2>(), s->caux_y.as<double2>(), in, out, S);
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    }
    hipLaunchKernelGGL((k_aux_node<R, false>), dim3(s->grid_node_groups), dim3(kBlock), 0, st,
                       s->n_nodes, s->n2e_ptr, s->n2e, s->gsign.as<int8_t>(), s->dninv.as<double2>(),
                       in, (const double2 *)nullptr, s->ynode.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_aux_apply_edge<R, PARTIALS>), dim3(s->grid), dim3(kBlock), 0, st, s->n,
                       s->en_eff.as<int2>(), s->dinv.as<double2>(), s->ynode.as<double2>(), in, out,
                       ostride, s->partial.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // Low-rank term after an SpMV that wrote `out` (and nb partial blocks of V values);
  // returns the partial block count including the correction block.
  template <int MODE>
  static int lowrank(efm_solver *s, const double2 *in, int in_stride, int in_off, double2 *out,
                     const double2 *other, int nb, hipStream_t st, int *nb_out) {
    constexpr int V = MODE == 2 ? 2 * R : R;
    *nb_out = nb;
    if (s->lr_P == 0)
      return EFM_OK;
    hipLaunchKernelGGL((k_lowrank<R, MODE>), dim3(1), dim3(kBlock), 0, st, s->lr_P, s->lr_len,
                       s->lr_idx.as<int32_t>(), s->lr_val.as<double2>(), s->lr_coef.as<double2>(), in, in_stride,
                       in_off, out, other, s->partial.as<double2>() + (int64_t)nb * V, s->state.as<SysState>());
    EFM_LAUNCH_CHECK();
    *nb_out = MODE == 3 ? nb : nb + 1;
    return EFM_OK;
  }

  // the variant-9 arrays from the fused-form init (z = W0.z, p = q = 0)
  static int init_zform(efm_solver *s, hipStream_t st) {
    hipLaunchKernelGGL(k_cocgz_init<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->W0.as<double2>(),
                       s->zf.as<double2>(), s->pf.as<double2>(), s->q.as<double2>());
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  static int init(efm_solver *s, const double2 *b, hipStream_t st, int method, double tol2) {
    if (method == EFM_METHOD_COCG && s->aux) {
      hipLaunchKernelGGL(k_cocg_init_aux<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, b,
                         s->x.as<double2>(), s->r.as<double2>(), s->W0.as<double2>());
      EFM_LAUNCH_CHECK();
      int rc = aux_apply<true>(s, s->r.as<double2>(), s->W0.as<double2>(), 2 * R, st, nullptr);
      if (rc != EFM_OK || (rc = init_zform(s, st)) != EFM_OK)
        return rc;
      return fin<kFinInitCocg>(s, s->grid, st, tol2, 0);
    }
    if (method == EFM_METHOD_COCG) {
      hipLaunchKernelGGL(k_cocg_init<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, b,
                         s->dinv.as<double2>(), s->x.as<double2>(), s->r.as<double2>(),
                         s->W0.as<double2>(), s->partial.as<double2>());
      EFM_LAUNCH_CHECK();
      int rc = init_zform(s, st);
      if (rc != EFM_OK)
        return rc;
      return fin<kFinInitCocg>(s, s->grid, st, tol2, 0);
    }
    hipLaunchKernelGGL(k_bs_init<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, b, s->x.as<double2>(),
                       s->r.as<double2>(), s->r0.as<double2>(), s->p0.as<double2>(),
                       s->v.as<double2>(), s->partial.as<double2>());
    EFM_LAUNCH_CHECK();
    return fin<kFinInitBs>(s, s->grid, st, tol2, 0);
  }

  // The sliced-ELL structure of variant 10 from the pattern (host ordering, device fill).
  static int build_sell(efm_solver *s, hipStream_t st) {
    constexpr int SR = 64 / kSellRL;
    const int64_t n = s->n;
    std::vector<int32_t> rp(n + 1);
    EFM_HIP_CHECK(hipMemcpyAsync(rp.data(), s->row_ptr, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    // rows by decreasing length inside windows of 4096 rows (stable: ties keep row order)
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i)
      order[i] = (int32_t)i;
    const int64_t win = 4096;
    for (int64_t w0 = 0; w0 < n; w0 += win) {
      const int64_t w1 = std::min(n, w0 + win);
      std::stable_sort(order.begin() + w0, order.begin() + w1,
                       [&](int32_t a, int32_t b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
    }
    const int64_t ns = (n + SR - 1) / SR;
    if (ns >= INT32_MAX)
      return EFM_ERR_ARG;
    std::vector<int32_t> srow(ns * SR, -1), soff(ns + 1, 0);
    int64_t steps = 0;
    for (int64_t sl = 0; sl < ns; ++sl) {
      int32_t mx = 0;
      for (int r = 0; r < SR; ++r) {
        const int64_t k = sl * SR + r;
        if (k < n) {
          srow[k] = order[k];
          mx = std::max(mx, rp[order[k] + 1] - rp[order[k]]);
        }
      }
      steps += (mx + kSellRL - 1) / kSellRL;
      if (steps * 64 >= INT32_MAX)
        return EFM_ERR_ARG;
      soff[sl + 1] = (int32_t)steps;
    }
    const int64_t slots = steps * 64;
    const size_t vsz = s->vtype == EFM_C128 ? sizeof(double2) : sizeof(float2);
    int rc;
    if ((rc = s->sell_off.alloc((ns + 1) * sizeof(int32_t))) || (rc = s->sell_row.alloc(ns * SR * sizeof(int32_t))) ||
        (rc = s->sell_col.alloc(slots * sizeof(int32_t))) || (rc = s->sell_vidx.alloc(slots * sizeof(int32_t))) ||
        (rc = s->sell_val.alloc(slots * vsz)))
      return rc;
    EFM_HIP_CHECK(hipMemcpyAsync(s->sell_off.p, soff.data(), (ns + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    EFM_HIP_CHECK(hipMemcpyAsync(s->sell_row.p, srow.data(), ns * SR * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sell_build<kSellRL>, dim3(grid_for(ns * 64, kBlock)), dim3(kBlock), 0, st, (int)ns,
                       s->sell_off.as<int32_t>(), s->sell_row.as<int32_t>(), s->row_ptr, s->col,
                       s->sell_col.as<int32_t>(), s->sell_vidx.as<int32_t>());
    EFM_LAUNCH_CHECK();
    EFM_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors above go out of scope
    s->sell_nslices = (int)ns;
    s->sell_xcd_span = (int)((ns + 7) / 8);
    s->sell_slots = slots;
    s->sell_built = true;
    s->sell_vals_ok = false;
    return EFM_OK;
  }

  // The split values of variant 11 from the installed complex-double matrix.
  static int build_split(efm_solver *s, hipStream_t st) {
    const int nt = s->ntiles[0];
    int rc;
    DevBuf flag;
    if ((rc = flag.alloc((size_t)nt * sizeof(int32_t))) || (rc = s->sp_off.alloc((size_t)nt * sizeof(int32_t))) ||
        (!s->sp_re.p && (rc = s->sp_re.alloc((size_t)s->nnz * sizeof(double)))))
      return rc;
    hipLaunchKernelGGL(k_split_flags, dim3(nt), dim3(kBlock), 0, st, nt, s->tile_desc.as<int4>(),
                       s->val.as<double2>(), flag.as<int32_t>());
    EFM_LAUNCH_CHECK();
    std::vector<int32_t> hf(nt), off(nt);
    std::vector<int4> hd(nt);
    EFM_HIP_CHECK(hipMemcpyAsync(hf.data(), flag.p, nt * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipMemcpyAsync(hd.data(), s->tile_desc.p, nt * sizeof(int4), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    int64_t m = 0;
    for (int t = 0; t < nt; ++t) {
      off[t] = hf[t] ? (int32_t)m : -1;
      if (hf[t])
        m += hd[t].w - hd[t].z;
    }
    if (m >= INT32_MAX)
      return EFM_ERR_ARG;
    if (m > s->sp_im_n || !s->sp_im.p) {
      if ((rc = s->sp_im.alloc((size_t)std::max<int64_t>(m, 1) * sizeof(double))))
        return rc;
      s->sp_im_n = m;
    }
    EFM_HIP_CHECK(hipMemcpyAsync(s->sp_off.p, off.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_split_fill, dim3(nt), dim3(kBlock), 0, st, nt, s->tile_desc.as<int4>(),
                       s->sp_off.as<int32_t>(), s->val.as<double2>(), s->sp_re.as<double>(), s->sp_im.as<double>());
    EFM_LAUNCH_CHECK();
    EFM_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors above go out of scope
    s->sp_ok = true;
    return EFM_OK;
  }

  // One fused SpMV of the given variant; *nb = number of partial sums written.
  template <typename VT>
  static int launch_spmv(efm_solver *s, int variant, const double2 *Wa, double2 *Wb, hipStream_t st,
                         int *nb) {
    SysState *S = s->state.as<SysState>();
    constexpr int TN = tile_nnz(R);
    switch (variant) {
    case 0:
      *nb = s->grid_rows[0];
      hipLaunchKernelGGL((k_cocg_spmv_rows<VT, R, 4>), dim3(*nb), dim3(kBlock), 0, st, s->n, s->row_ptr,
                         s->col, s->val.as<VT>(), Wa, Wb, s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case 1:
      *nb = s->grid_rows[1];
      hipLaunchKernelGGL((k_cocg_spmv_rows<VT, R, 8>), dim3(*nb), dim3(kBlock), 0, st, s->n, s->row_ptr,
                         s->col, s->val.as<VT>(), Wa, Wb, s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case 2:
      *nb = s->grid_rows[2];
      hipLaunchKernelGGL((k_cocg_spmv_rows<VT, R, 16>), dim3(*nb), dim3(kBlock), 0, st, s->n, s->row_ptr,
                         s->col, s->val.as<VT>(), Wa, Wb, s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case 3:
      *nb = s->grid_tiles_res[0];
      hipLaunchKernelGGL((k_cocg_spmv_tiled<VT, R, TN, 4>), dim3(*nb), dim3(kBlock), 0, st, s->ntiles[0],
                         s->xcd_span[0], s->tile_row[0].as<int32_t>(), s->row_ptr, s->col, s->val.as<VT>(),
                         Wa, Wb, s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case 4:
      *nb = s->grid_tiles_res[1];
      hipLaunchKernelGGL((k_cocg_spmv_tiled<VT, R, tile_nnz2(R), 8>), dim3(*nb), dim3(kBlock), 0, st,
                         s->ntiles[1], s->xcd_span[1], s->tile_row[1].as<int32_t>(), s->row_ptr, s->col,
                         s->val.as<VT>(), Wa, Wb, s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case 6:
      *nb = s->grid_pipe_res;
      hipLaunchKernelGGL((k_cocg_spmv_pipe<VT, R, TN, 4>), dim3(*nb), dim3(kBlock), 0, st, s->ntiles[0],
                         s->xcd_span[0], s->tile_desc.as<int4>(), s->row_ptr, s->col, s->val.as<VT>(), Wa, Wb,
                         s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case kSellForm: {
      int rc;
      if (!s->sell_built && (rc = build_sell(s, st)) != EFM_OK)
        return rc;
      if (!s->sell_vals_ok) {
        hipLaunchKernelGGL(k_sell_vals<VT>, dim3(grid_for(s->sell_slots, 256)), dim3(256), 0, st, s->sell_slots,
                           s->sell_vidx.as<int32_t>(), s->val.as<VT>(), s->sell_val.as<VT>());
        EFM_LAUNCH_CHECK();
        s->sell_vals_ok = true;
      }
      *nb = s->grid_sell;
      hipLaunchKernelGGL((k_cocgz_sell<VT, R, kSellRL>), dim3(*nb), dim3(kBlock), 0, st, s->sell_nslices,
                         s->sell_xcd_span, s->sell_off.as<int32_t>(), s->sell_row.as<int32_t>(),
                         s->sell_col.as<int32_t>(), s->sell_val.as<VT>(), s->pf.as<double2>(), s->q.as<double2>(),
                         s->partial.as<double2>(), S);
      break;
    }
    case kZForm:
      *nb = s->grid_z;
      hipLaunchKernelGGL((k_cocgz_spmv<VT, R, TN, 4>), dim3(*nb), dim3(kBlock), 0, st, s->ntiles[0],
                         s->xcd_span[0], s->tile_desc.as<int4>(), s->row_ptr, s->col, s->val.as<VT>(),
                         s->pf.as<double2>(), s->q.as<double2>(), s->partial.as<double2>(), S);
      break;
    case kSplitForm:
      if constexpr (std::is_same<VT, double2>::value) {
        int rc;
        if (!s->sp_ok && (rc = build_split(s, st)) != EFM_OK)
          return rc;
        *nb = s->grid_z;
        hipLaunchKernelGGL((k_cocgz_spmv<VT, R, TN, 4, true>), dim3(*nb), dim3(kBlock), 0, st, s->ntiles[0],
                           s->xcd_span[0], s->tile_desc.as<int4>(), s->row_ptr, s->col, s->val.as<VT>(),
                           s->pf.as<double2>(), s->q.as<double2>(), s->partial.as<double2>(), S,
                           (const double *)s->sp_re.as<double>(), (const double *)s->sp_im.as<double>(),
                           (const int32_t *)s->sp_off.as<int32_t>());
        break;
      }
      set_error("SpMV variant 11 needs complex-double matrix values");
      return EFM_ERR_ARG;
    default:
      set_error("unknown SpMV variant");
      return EFM_ERR_ARG;
    }
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  }

  // Time every available variant on the installed matrix (state and W0 must be valid:
  // called right after init; the timed launches only write W1.p, q and the partials,
  // which the first real iteration overwrites) and keep the fastest.
  template <typename VT> static int autotune(efm_solver *s, hipStream_t st) {
    hipEvent_t e0, e1;
    EFM_HIP_CHECK(hipEventCreate(&e0));
    EFM_HIP_CHECK(hipEventCreate(&e1));
    int best = 1;
    double best_ms = 1e30;
    for (int v = 0; v < kNumSpmvVariants; ++v) {
      s->tune_ms[v] = -1.0;
      if (v >= 3 && v <= 4 && s->ntiles[v - 3] == 0)
        continue;
      if (kRetired(v))
        continue;
      if (v == 6 && s->ntiles[0] == 0)
        continue;
      if (v == kZForm && (s->ntiles[0] == 0 || s->lr_P > 0))
        continue;
      if (v == kSellForm && s->lr_P > 0)
        continue;
      if (v == kSplitForm && (s->ntiles[0] == 0 || s->lr_P > 0 || !std::is_same<VT, double2>::value))
        continue;
      int nb;
      // variant 9 is timed with its p pass (run on W1 as scratch: the real p must stay as
      // init left it)
      auto launch = [&]() -> int {
        if (v == kZForm || v == kSellForm || v == kSplitForm) {
          hipLaunchKernelGGL(k_cocgz_pform<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->zf.as<double2>(),
                             s->W1.as<double2>(), s->state.as<SysState>());
          EFM_LAUNCH_CHECK();
        }
        return launch_spmv<VT>(s, v, s->W0.as<double2>(), s->W1.as<double2>(), st, &nb);
      };
      int rc = launch();  // warm-up
      if (rc != EFM_OK)
        return rc;
      EFM_HIP_CHECK(hipEventRecord(e0, st));
      const int reps = 5;
      for (int k = 0; k < reps; ++k)
        if ((rc = launch()) != EFM_OK)
          return rc;
      EFM_HIP_CHECK(hipEventRecord(e1, st));
      EFM_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      EFM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      s->tune_ms[v] = ms / reps;
      if (ms / reps < best_ms) {
        best_ms = ms / reps;
        best = v;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->spmv_variant = best;
    return EFM_OK;
  }

  template <typename VT>
  static int cocg_iter(efm_solver *s, int it, hipStream_t st, double tol2, int max_it,
                       hipEvent_t e0, hipEvent_t e1) {
    double2 *Wa = (it & 1) ? s->W1.as<double2>() : s->W0.as<double2>();
    double2 *Wb = (it & 1) ? s->W0.as<double2>() : s->W1.as<double2>();
    SysState *S = s->state.as<SysState>();
    if (s->spmv_variant == kZForm || s->spmv_variant == kSellForm || s->spmv_variant == kSplitForm) {
      int nb_spmv = 0;
      int rc;
      // x += alpha p (the previous iteration's), p = z + beta p, then q = A p
      hipLaunchKernelGGL((k_cocgz_pform<R, true>), dim3(s->grid), dim3(kBlock), 0, st, s->n, s->zf.as<double2>(),
                         s->pf.as<double2>(), S, s->x.as<double2>());
      EFM_LAUNCH_CHECK();
      if (e0)
        EFM_HIP_CHECK(hipEventRecord(e0, st));
      if ((rc = launch_spmv<VT>(s, s->spmv_variant, Wa, Wb, st, &nb_spmv)) != EFM_OK)
        return rc;
      if (e1)
        EFM_HIP_CHECK(hipEventRecord(e1, st));
      if ((rc = fin<kFinAlpha>(s, nb_spmv, st, tol2, max_it)) != EFM_OK)
        return rc;
      if (s->aux) {
        hipLaunchKernelGGL((k_aux_node<R, true>), dim3(s->grid_node_groups), dim3(kBlock), 0, st,
                           s->n_nodes, s->n2e_ptr, s->n2e, s->gsign.as<int8_t>(), s->dninv.as<double2>(),
                           s->r.as<double2>(), s->q.as<double2>(), s->ynode.as<double2>(), S);
        EFM_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_cocgz_update<R, true, true>), dim3(s->grid), dim3(kBlock), 0, st, s->n,
                           s->en_eff.as<int2>(), s->dinv.as<double2>(), s->ynode.as<double2>(),
                           s->pf.as<double2>(), s->q.as<double2>(), s->zf.as<double2>(), s->x.as<double2>(),
                           s->r.as<double2>(), s->partial.as<double2>(), S);
      } else {
        hipLaunchKernelGGL((k_cocgz_update<R, false, true>), dim3(s->grid), dim3(kBlock), 0, st, s->n,
                           (const int2 *)nullptr, s->dinv.as<double2>(), (const double2 *)nullptr,
                           s->pf.as<double2>(), s->q.as<double2>(), s->zf.as<double2>(), s->x.as<double2>(),
                           s->r.as<double2>(), s->partial.as<double2>(), S);
      }
      EFM_LAUNCH_CHECK();
      return fin<kFinBeta>(s, s->grid, st, tol2, max_it);
    }
    if (e0)
      EFM_HIP_CHECK(hipEventRecord(e0, st));
    int nb_spmv = 0;
    int rc = launch_spmv<VT>(s, s->spmv_variant, Wa, Wb, st, &nb_spmv);
    if (rc != EFM_OK)
      return rc;
    if (e1)
      EFM_HIP_CHECK(hipEventRecord(e1, st));
    if ((rc = lowrank<0>(s, Wb, 2 * R, R, s->q.as<double2>(), nullptr, nb_spmv, st, &nb_spmv)) != EFM_OK)
      return rc;
    rc = fin<kFinAlpha>(s, nb_spmv, st, tol2, max_it);
    if (rc != EFM_OK)
      return rc;
    if (s->aux) {
      hipLaunchKernelGGL((k_aux_node<R, true>), dim3(s->grid_node_groups), dim3(kBlock), 0, st,
                         s->n_nodes, s->n2e_ptr, s->n2e, s->gsign.as<int8_t>(), s->dninv.as<double2>(),
                         s->r.as<double2>(), s->q.as<double2>(), s->ynode.as<double2>(), S);
      EFM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_aux_cocg_update<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n,
                         s->en_eff.as<int2>(), s->dinv.as<double2>(), s->ynode.as<double2>(), Wb,
                         s->q.as<double2>(), s->x.as<double2>(), s->r.as<double2>(),
                         s->partial.as<double2>(), S);
      EFM_LAUNCH_CHECK();
      return fin<kFinBeta>(s, s->grid, st, tol2, max_it);
    }
    hipLaunchKernelGGL(k_cocg_update<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n,
                       s->dinv.as<double2>(), Wb, s->q.as<double2>(), s->x.as<double2>(),
                       s->r.as<double2>(), s->partial.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    return fin<kFinBeta>(s, s->grid, st, tol2, max_it);
  }

  template <typename VT>
  static int bicgstab_iter(efm_solver *s, int it, hipStream_t st, double tol2, int max_it,
                           hipEvent_t e0, hipEvent_t e1) {
    SysState *S = s->state.as<SysState>();
    double2 *part = s->partial.as<double2>();
    int rc;
    if (it == 0) {
      hipLaunchKernelGGL(k_bs_dot_r0r<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n,
                         s->r0.as<double2>(), s->r.as<double2>(), part);
      EFM_LAUNCH_CHECK();
      if ((rc = fin<kFinBsRho0>(s, s->grid, st, tol2, max_it)) != EFM_OK)
        return rc;
    }
    hipLaunchKernelGGL(k_bs_p<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->dinv.as<double2>(),
                       s->r.as<double2>(), s->v.as<double2>(), s->p0.as<double2>(),
                       s->y.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    if ((s->aux || s->caux) &&
        (rc = aux_apply<false>(s, s->p0.as<double2>(), s->y.as<double2>(), R, st, S)) != EFM_OK)
      return rc;
    if (e0)
      EFM_HIP_CHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL((k_bs_spmv<VT, R, 0>), dim3(s->grid_spmv), dim3(kBlock), 0, st, s->n,
                       s->row_ptr, s->col, s->val.as<VT>(), s->y.as<double2>(), s->v.as<double2>(),
                       s->r0.as<double2>(), part, S);
    EFM_LAUNCH_CHECK();
    if (e1)
      EFM_HIP_CHECK(hipEventRecord(e1, st));
    int nbv = 0;
    if ((rc = lowrank<1>(s, s->y.as<double2>(), R, 0, s->v.as<double2>(), s->r0.as<double2>(), s->grid_spmv, st,
                         &nbv)) != EFM_OK)
      return rc;
    if ((rc = fin<kFinBsAlpha>(s, nbv, st, tol2, max_it)) != EFM_OK)
      return rc;
    hipLaunchKernelGGL(k_bs_s<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->dinv.as<double2>(),
                       s->v.as<double2>(), s->r.as<double2>(), s->zz.as<double2>(), S);
    EFM_LAUNCH_CHECK();
    if ((s->aux || s->caux) &&
        (rc = aux_apply<false>(s, s->r.as<double2>(), s->zz.as<double2>(), R, st, S)) != EFM_OK)
      return rc;
    hipLaunchKernelGGL((k_bs_spmv<VT, R, 1>), dim3(s->grid_spmv), dim3(kBlock), 0, st, s->n,
                       s->row_ptr, s->col, s->val.as<VT>(), s->zz.as<double2>(), s->t.as<double2>(),
                       s->r.as<double2>(), part, S);
    EFM_LAUNCH_CHECK();
    int nbt = 0;
    if ((rc = lowrank<2>(s, s->zz.as<double2>(), R, 0, s->t.as<double2>(), s->r.as<double2>(), s->grid_spmv, st,
                         &nbt)) != EFM_OK)
      return rc;
    if ((rc = fin<kFinBsOmega>(s, nbt, st, tol2, max_it)) != EFM_OK)
      return rc;
    hipLaunchKernelGGL(k_bs_x<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->y.as<double2>(),
                       s->zz.as<double2>(), s->t.as<double2>(), s->r0.as<double2>(),
                       s->x.as<double2>(), s->r.as<double2>(), part, S);
    EFM_LAUNCH_CHECK();
    return fin<kFinBsX>(s, s->grid, st, tol2, max_it);
  }

  template <typename VT> static int spmv(efm_solver *s, const double2 *x, double2 *y, hipStream_t st) {
    hipLaunchKernelGGL((k_spmv<VT, R>), dim3(s->grid_spmv), dim3(kBlock), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<VT>(), x, y);
    EFM_LAUNCH_CHECK();
    int nb = 0;
    return lowrank<3>(s, x, R, 0, y, nullptr, 0, st, &nb);
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
  const int64_t cap = (int64_t)ncu * 8;
  int64_t need_spmv = (n + kRowsPerBlock - 1) / kRowsPerBlock;
  int64_t need_vec = (n + kBlock - 1) / kBlock;
  s->grid_spmv = (int)std::max<int64_t>(1, std::min(need_spmv, cap));
  s->grid = (int)std::max<int64_t>(1, std::min(need_vec, cap));
  const size_t vb = (size_t)n * nrhs * sizeof(double2);
  int rc;
  const size_t vbytes = (size_t)nnz * (value_type == EFM_C128 ? 16 : 8);
  // row tiles for the tiled SpMV variants: greedy runs of whole rows, <= TNNZ nonzeros
  // and <= 256 / RL rows per tile; blocks = 8 XCDs x min(tiles per XCD, ncu)
  std::vector<int32_t> h_rp(n + 1);
  EFM_HIP_CHECK(hipMemcpy(h_rp.data(), row_ptr, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int tv = 0; tv < 2; ++tv) {
    const int tcap = tv == 0 ? tile_nnz(nrhs) : tile_nnz2(nrhs);
    const int rcap = tv == 0 ? kBlock / 4 : kBlock / 8;
    std::vector<int32_t> tiles;
    tiles.push_back(0);
    bool tileable = true;
    for (int64_t r = 0; r < n;) {
      int64_t r1 = r;
      while (r1 < n && (r1 - r) < rcap && (h_rp[r1 + 1] - h_rp[r]) <= tcap)
        ++r1;
      if (r1 == r) {  // a single row longer than a tile
        tileable = false;
        break;
      }
      tiles.push_back((int32_t)r1);
      r = r1;
    }
    if (!tileable)
      continue;
    s->ntiles[tv] = (int)tiles.size() - 1;
    s->xcd_span[tv] = (s->ntiles[tv] + 7) / 8;
    const int64_t per_xcd = std::min<int64_t>(s->xcd_span[tv], (int64_t)ncu);
    s->grid_tiles[tv] = (int)(8 * std::max<int64_t>(1, per_xcd));
    if ((rc = s->tile_row[tv].alloc(tiles.size() * sizeof(int32_t))))
      return rc;
    EFM_HIP_CHECK(hipMemcpy(s->tile_row[tv].p, tiles.data(), tiles.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
    if (tv == 0) {
      std::vector<int4> dsc(s->ntiles[0]);
      for (int t = 0; t < s->ntiles[0]; ++t)
        dsc[t] = make_int4(tiles[t], tiles[t + 1], h_rp[tiles[t]], h_rp[tiles[t + 1]]);
      if ((rc = s->tile_desc.alloc(dsc.size() * sizeof(int4))))
        return rc;
      EFM_HIP_CHECK(hipMemcpy(s->tile_desc.p, dsc.data(), dsc.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
  }
  const int lprs[3] = {4, 8, 16};
  for (int v = 0; v < 3; ++v) {
    const int64_t rows = kBlock / lprs[v];
    const int64_t nchunks = (n + rows - 1) / rows;
    const int64_t span = (nchunks + 7) / 8;
    s->grid_rows[v] = (int)(8 * std::max<int64_t>(1, std::min<int64_t>(span, (int64_t)ncu)));
  }
  // tiled variants: as many blocks as are resident (a multiple of 8), so that each block's
  // tile walk runs once instead of in several rounds of blocks
  {
    const bool c64 = value_type == EFM_C64;
    auto occ = [&](const void *k) {
      int nb = 0;
      if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kBlock, 0) != hipSuccess || nb < 1)
        nb = 1;
      return nb;
    };
    auto tiled_k = [&](int tv) -> const void * {
#define EFM_TILED_K(RR)                                                                                         \
  (tv == 0 ? (c64 ? reinterpret_cast<const void *>(&k_cocg_spmv_tiled<float2, RR, tile_nnz(RR), 4>)             \
                  : reinterpret_cast<const void *>(&k_cocg_spmv_tiled<double2, RR, tile_nnz(RR), 4>))           \
           : (c64 ? reinterpret_cast<const void *>(&k_cocg_spmv_tiled<float2, RR, tile_nnz2(RR), 8>)            \
                  : reinterpret_cast<const void *>(&k_cocg_spmv_tiled<double2, RR, tile_nnz2(RR), 8>)))
      switch (nrhs) {
      case 1: return EFM_TILED_K(1);
      case 2: return EFM_TILED_K(2);
      case 3: return EFM_TILED_K(3);
      case 4: return EFM_TILED_K(4);
      default: return EFM_TILED_K(8);
      }
#undef EFM_TILED_K
    };
    for (int tv = 0; tv < 2; ++tv)
      if (s->ntiles[tv] > 0) {
        const int64_t per_xcd =
            std::min<int64_t>(s->xcd_span[tv], std::max<int64_t>(1, (int64_t)ncu * occ(tiled_k(tv)) / 8));
        s->grid_tiles_res[tv] = (int)(8 * per_xcd);
      }
    if (s->ntiles[0] > 0) {
#define EFM_PIPE_K(RR)                                                                                    \
  (c64 ? reinterpret_cast<const void *>(&k_cocg_spmv_pipe<float2, RR, tile_nnz(RR), 4>)                   \
       : reinterpret_cast<const void *>(&k_cocg_spmv_pipe<double2, RR, tile_nnz(RR), 4>))
      const void *kp = nrhs == 1 ? EFM_PIPE_K(1)
                     : nrhs == 2 ? EFM_PIPE_K(2)
                     : nrhs == 3 ? EFM_PIPE_K(3)
                     : nrhs == 4 ? EFM_PIPE_K(4)
                                 : EFM_PIPE_K(8);
#undef EFM_PIPE_K
      const int64_t per_xcd =
          std::min<int64_t>(s->xcd_span[0], std::max<int64_t>(1, (int64_t)ncu * occ(kp) / 8));
      s->grid_pipe_res = (int)(8 * per_xcd);
#define EFM_Z_K(RR)                                                                                       \
  (c64 ? reinterpret_cast<const void *>(&k_cocgz_spmv<float2, RR, tile_nnz(RR), 4>)                \
       : reinterpret_cast<const void *>(&k_cocgz_spmv<double2, RR, tile_nnz(RR), 4>))
      const void *kz = nrhs == 1 ? EFM_Z_K(1)
                     : nrhs == 2 ? EFM_Z_K(2)
                     : nrhs == 3 ? EFM_Z_K(3)
                     : nrhs == 4 ? EFM_Z_K(4)
                                 : EFM_Z_K(8);
#undef EFM_Z_K
      const int64_t per_xcd_z =
          std::min<int64_t>(s->xcd_span[0], std::max<int64_t>(1, (int64_t)ncu * occ(kz) / 8));
      s->grid_z = (int)(8 * per_xcd_z);
    }
  }
  {
    // sliced-ELL SpMV: as many blocks as are resident, a multiple of 8
    const bool c64 = value_type == EFM_C64;
    const void *ks = nullptr;
#define EFM_SELL_K(RR)                                                                                    \
  (c64 ? reinterpret_cast<const void *>(&k_cocgz_sell<float2, RR, kSellRL>)                              \
       : reinterpret_cast<const void *>(&k_cocgz_sell<double2, RR, kSellRL>))
    ks = nrhs == 1 ? EFM_SELL_K(1) : nrhs == 2 ? EFM_SELL_K(2) : nrhs == 3 ? EFM_SELL_K(3)
       : nrhs == 4 ? EFM_SELL_K(4) : EFM_SELL_K(8);
#undef EFM_SELL_K
    int nbs = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbs, ks, kBlock, 0) != hipSuccess || nbs < 1)
      nbs = 1;
    s->grid_sell = (int)std::max<int64_t>(8, ((int64_t)ncu * nbs / 8) * 8);
  }
  int gmax = std::max(s->grid, s->grid_spmv);
  gmax = std::max(gmax, s->grid_sell);
  for (int v = 0; v < 3; ++v)
    gmax = std::max(gmax, s->grid_rows[v]);
  for (int v = 0; v < 2; ++v)
    gmax = std::max(gmax, s->grid_tiles[v]);
  gmax = std::max(gmax, std::max(s->grid_tiles_res[0], s->grid_tiles_res[1]));
  gmax = std::max(gmax, s->grid_pipe_res);
  gmax = std::max(gmax, s->grid_z);
  if ((rc = s->val.alloc(vbytes)) || (rc = s->dinv.alloc(n * sizeof(double2))) ||
      (rc = s->x.alloc(vb)) || (rc = s->r.alloc(vb)) || (rc = s->q.alloc(vb)) ||
      (rc = s->W0.alloc(2 * vb)) || (rc = s->W1.alloc(2 * vb)) || (rc = s->zf.alloc(vb)) ||
      (rc = s->pf.alloc(vb)) ||
      (rc = s->partial.alloc((size_t)(gmax + 1) * 2 * kMaxRhs * sizeof(double2))) ||
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
  s->sell_vals_ok = false;  // the sliced-ELL values are refilled at the next SpMV of variant 10
  s->sp_ok = false;         // and the split values of variant 11 rebuilt
  if (s->vtype == EFM_C128)
    hipLaunchKernelGGL(k_dinv<double2>, dim3(grid_for(s->n, 256)), dim3(256), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<double2>(), s->dinv.as<double2>());
  else
    hipLaunchKernelGGL(k_dinv<float2>, dim3(grid_for(s->n, 256)), dim3(256), 0, st, s->n, s->row_ptr,
                       s->col, s->val.as<float2>(), s->dinv.as<double2>());
  EFM_LAUNCH_CHECK();
  if (s->lr_P > 0) {
    const int64_t m = s->lr_P * s->lr_len;
    if (s->vtype == EFM_C128)
      hipLaunchKernelGGL(k_lowrank_dinv<double2>, dim3(grid_for(m, 256)), dim3(256), 0, st, m, s->lr_len,
                         s->lr_idx.as<int32_t>(), s->lr_val.as<double2>(), s->lr_coef.as<double2>(), s->row_ptr,
                         s->col, s->val.as<double2>(), s->dinv.as<double2>());
    else
      hipLaunchKernelGGL(k_lowrank_dinv<float2>, dim3(grid_for(m, 256)), dim3(256), 0, st, m, s->lr_len,
                         s->lr_idx.as<int32_t>(), s->lr_val.as<double2>(), s->lr_coef.as<double2>(), s->row_ptr,
                         s->col, s->val.as<float2>(), s->dinv.as<double2>());
    EFM_LAUNCH_CHECK();
  }
  if (s->caux) {
    if (s->vtype == EFM_C128)
      hipLaunchKernelGGL(k_caux_diag<double2>, dim3(grid_for(s->caux_N, 256)), dim3(256), 0, st, s->caux_N,
                         s->caux_gt_ptr.as<int32_t>(), s->caux_gt_edge.as<int32_t>(), s->caux_gt_coef.as<double2>(),
                         s->caux_en2.as<int2>(), s->caux_ga.as<double2>(), s->caux_gb.as<double2>(), s->row_ptr,
                         s->col, s->val.as<double2>(), s->caux_dninv.as<double2>());
    else
      hipLaunchKernelGGL(k_caux_diag<float2>, dim3(grid_for(s->caux_N, 256)), dim3(256), 0, st, s->caux_N,
                         s->caux_gt_ptr.as<int32_t>(), s->caux_gt_edge.as<int32_t>(), s->caux_gt_coef.as<double2>(),
                         s->caux_en2.as<int2>(), s->caux_ga.as<double2>(), s->caux_gb.as<double2>(), s->row_ptr,
                         s->col, s->val.as<float2>(), s->caux_dninv.as<double2>());
    EFM_LAUNCH_CHECK();
  }
  if (s->aux) {
    if (s->vtype == EFM_C128)
      hipLaunchKernelGGL(k_aux_diag<double2>, dim3(grid_for(s->n_nodes, 256)), dim3(256), 0, st,
                         s->n_nodes, s->n2e_ptr, s->n2e, s->edge_nodes, s->dir, s->row_ptr, s->col,
                         s->val.as<double2>(), s->dninv.as<double2>());
    else
      hipLaunchKernelGGL(k_aux_diag<float2>, dim3(grid_for(s->n_nodes, 256)), dim3(256), 0, st,
                         s->n_nodes, s->n2e_ptr, s->n2e, s->edge_nodes, s->dir, s->row_ptr, s->col,
                         s->val.as<float2>(), s->dninv.as<double2>());
    EFM_LAUNCH_CHECK();
  }
  return EFM_OK;
}

extern "C" int efm_solver_set_lowrank(efm_solver *s, int n_terms, int64_t len, const int32_t *gidx,
                                      const double *gval, const double *h_coef, void *stream) {
  EFM_REQUIRE(s, "efm_solver_set_lowrank: NULL solver");
  EFM_REQUIRE(n_terms >= 0 && n_terms <= kMaxPorts, "efm_solver_set_lowrank: 0..32 terms");
  if (n_terms == 0 || len <= 0) {
    s->lr_P = 0;
    s->lr_len = 0;
    return EFM_OK;
  }
  EFM_REQUIRE(gidx && gval && h_coef, "efm_solver_set_lowrank: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t m = (int64_t)n_terms * len;
  std::vector<int32_t> h(m);
  EFM_HIP_CHECK(hipMemcpyAsync(h.data(), gidx, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<uint8_t> seen(s->n, 0);
  for (int64_t i = 0; i < m; ++i) {
    if (h[i] < 0)
      continue;
    EFM_REQUIRE(h[i] < s->n, "efm_solver_set_lowrank: edge index out of range");
    EFM_REQUIRE(!seen[h[i]], "efm_solver_set_lowrank: an edge appears twice (ports must be disjoint)");
    seen[h[i]] = 1;
  }
  int rc;
  if ((rc = s->lr_idx.alloc(m * sizeof(int32_t))) || (rc = s->lr_val.alloc(m * sizeof(double2))) ||
      (rc = s->lr_coef.alloc(kMaxPorts * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(s->lr_idx.p, gidx, m * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->lr_val.p, gval, m * sizeof(double2), hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->lr_coef.p, h_coef, n_terms * sizeof(double2), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  s->lr_P = n_terms;
  s->lr_len = len;
  return EFM_OK;
}

extern "C" int efm_solver_set_gradient_complex(efm_solver *s, int64_t n_nodes, const int32_t *h_edge_nodes,
                                               const double *h_edge_coefs, void *stream) {
  EFM_REQUIRE(s, "efm_solver_set_gradient_complex: NULL solver");
  if (n_nodes <= 0) {
    s->caux = false;
    return EFM_OK;
  }
  EFM_REQUIRE(h_edge_nodes && h_edge_coefs, "efm_solver_set_gradient_complex: NULL argument");
  EFM_REQUIRE(n_nodes < INT32_MAX, "efm_solver_set_gradient_complex: too many nodes");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  std::vector<int2> en2(n);
  std::vector<double2> ga(n), gb(n);
  std::vector<int32_t> cnt(n_nodes + 1, 0);
  for (int64_t e = 0; e < n; ++e) {
    const int32_t a = h_edge_nodes[2 * e], b = h_edge_nodes[2 * e + 1];
    EFM_REQUIRE(a >= -1 && a < n_nodes && b >= -1 && b < n_nodes, "efm_solver_set_gradient_complex: node index");
    en2[e] = make_int2(a, b);
    ga[e] = make_double2(h_edge_coefs[4 * e], h_edge_coefs[4 * e + 1]);
    gb[e] = make_double2(h_edge_coefs[4 * e + 2], h_edge_coefs[4 * e + 3]);
    if (a >= 0)
      ++cnt[a + 1];
    if (b >= 0 && b != a)
      ++cnt[b + 1];
  }
  for (int64_t k = 0; k < n_nodes; ++k)
    cnt[k + 1] += cnt[k];
  std::vector<int32_t> gt_edge(std::max<int32_t>(cnt[n_nodes], 1));
  std::vector<double2> gt_coef(std::max<int32_t>(cnt[n_nodes], 1));
  std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < n; ++e) {  // ascending edge order per node
    const int2 nn = en2[e];
    if (nn.x >= 0) {
      const double2 c = (nn.y == nn.x) ? cadd(ga[e], gb[e]) : ga[e];
      gt_edge[fill[nn.x]] = (int32_t)e;
      gt_coef[fill[nn.x]++] = c;
    }
    if (nn.y >= 0 && nn.y != nn.x) {
      gt_edge[fill[nn.y]] = (int32_t)e;
      gt_coef[fill[nn.y]++] = gb[e];
    }
  }
  int rc;
  if ((rc = s->caux_en2.alloc(n * sizeof(int2))) || (rc = s->caux_ga.alloc(n * sizeof(double2))) ||
      (rc = s->caux_gb.alloc(n * sizeof(double2))) || (rc = s->caux_gt_ptr.alloc((n_nodes + 1) * 4)) ||
      (rc = s->caux_gt_edge.alloc(gt_edge.size() * 4)) || (rc = s->caux_gt_coef.alloc(gt_coef.size() * 16)) ||
      (rc = s->caux_dninv.alloc(n_nodes * sizeof(double2))) ||
      (rc = s->caux_y.alloc((size_t)n_nodes * s->R * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_en2.p, en2.data(), n * sizeof(int2), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_ga.p, ga.data(), n * sizeof(double2), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_gb.p, gb.data(), n * sizeof(double2), hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_gt_ptr.p, cnt.data(), (n_nodes + 1) * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_gt_edge.p, gt_edge.data(), gt_edge.size() * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->caux_gt_coef.p, gt_coef.data(), gt_coef.size() * 16, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  s->caux_N = n_nodes;
  s->caux = true;
  s->aux = false;
  return EFM_OK;
}

extern "C" int efm_solver_set_gradient(efm_solver *s, int64_t n_nodes, const int32_t *edge_nodes,
                                       const int32_t *n2e_ptr, const int32_t *n2e,
                                       const uint8_t *dirichlet, void *stream) {
  EFM_REQUIRE(s, "efm_solver_set_gradient: NULL solver");
  if (n_nodes <= 0) {  // switch the auxiliary correction off
    s->aux = false;
    return EFM_OK;
  }
  EFM_REQUIRE(edge_nodes && n2e_ptr && n2e, "efm_solver_set_gradient: NULL mesh array");
  int rc;
  if ((rc = s->dninv.alloc((size_t)n_nodes * sizeof(double2))) ||
      (rc = s->ynode.alloc((size_t)n_nodes * s->R * sizeof(double2))))
    return rc;
  s->n_nodes = n_nodes;
  s->edge_nodes = edge_nodes;
  s->n2e_ptr = n2e_ptr;
  s->n2e = n2e;
  s->dir = dirichlet;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      ncu = prop.multiProcessorCount;
  }
  s->grid_nodes = (int)std::max<int64_t>(1, std::min<int64_t>((n_nodes + kBlock - 1) / kBlock, (int64_t)ncu * 8));
  const int64_t node_groups = (n_nodes + (kBlock / kLanesPerNode) - 1) / (kBlock / kLanesPerNode);
  // a multiple of 8 (the nodal kernel's XCD-aware walk)
  s->grid_node_groups = (int)(8 * std::max<int64_t>(1, std::min<int64_t>((node_groups + 7) / 8, (int64_t)ncu)));
  if ((rc = s->gsign.alloc((size_t)2 * s->n + 16)) || (rc = s->en_eff.alloc((size_t)s->n * sizeof(int2))))
    return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_aux_prep_nodes, dim3(grid_for(n_nodes, 256)), dim3(256), 0, st, n_nodes, n2e_ptr,
                     n2e, edge_nodes, dirichlet, s->gsign.as<int8_t>());
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_aux_prep_edges, dim3(grid_for(s->n, 256)), dim3(256), 0, st, s->n, edge_nodes,
                     dirichlet, s->en_eff.as<int2>());
  EFM_LAUNCH_CHECK();
  s->aux = true;
  s->caux = false;
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

extern "C" int efm_solver_set_spmv_variant(efm_solver *s, int variant) {
  EFM_REQUIRE(s, "efm_solver_set_spmv_variant: NULL solver");
  EFM_REQUIRE(variant >= -1 && variant < kNumSpmvVariants, "efm_solver_set_spmv_variant: bad variant");
  EFM_REQUIRE(variant < 0 || !kRetired(variant),
              "efm_solver_set_spmv_variant: variants 5, 7 and 8 were removed (slower on every matrix measured)");
  EFM_REQUIRE(variant < 3 || variant == kSellForm ||
                  s->ntiles[(variant == 6 || variant == kZForm || variant == kSplitForm) ? 0 : variant - 3] > 0,
              "efm_solver_set_spmv_variant: matrix has a row longer than a tile");
  EFM_REQUIRE(variant != kSplitForm || s->vtype == EFM_C128,
              "efm_solver_set_spmv_variant: variant 11 needs complex-double matrix values");
  s->spmv_variant = variant;
  return EFM_OK;
}

extern "C" int efm_solver_get_spmv_variant(const efm_solver *s, int *h_variant, double *h_tune_ms) {
  EFM_REQUIRE(s && h_variant, "efm_solver_get_spmv_variant: NULL argument");
  *h_variant = s->spmv_variant;
  if (h_tune_ms)
    for (int v = 0; v < kNumSpmvVariants; ++v)
      h_tune_ms[v] = s->tune_ms[v];
  return EFM_OK;
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
  EFM_REQUIRE(!(cocg && s->caux), "efm_solver_solve: the complex-incidence preconditioner needs BiCGSTAB");
  EFM_REQUIRE(!(cocg && (s->spmv_variant == kZForm || s->spmv_variant == kSellForm || s->spmv_variant == kSplitForm) &&
                s->lr_P > 0),
              "efm_solver_solve: SpMV variants 9-11 (p formed in its own pass) do not take low-rank port terms");
  int rc;
  if (!cocg) {
    const size_t vb = (size_t)s->n * s->R * sizeof(double2);
    if (!s->r0.p && ((rc = s->r0.alloc(vb)) || (rc = s->v.alloc(vb)) || (rc = s->y.alloc(vb)) ||
                     (rc = s->zz.alloc(vb)) || (rc = s->t.alloc(vb)) || (rc = s->p0.alloc(vb))))
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
  if (cocg && s->spmv_variant < 0) {
    rc = with_rhs(s->R, [&](auto RC) -> int {
      constexpr int R = decltype(RC)::value;
      if (s->vtype == EFM_C128)
        return Dispatch<R>::template autotune<double2>(s, st);
      return Dispatch<R>::template autotune<float2>(s, st);
    });
    if (rc != EFM_OK)
      return rc;
  }
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
  if (cocg && (s->spmv_variant == kZForm || s->spmv_variant == kSellForm || s->spmv_variant == kSplitForm)) {
    // the last iteration's x updates (the p pass makes them one iteration late)
    rc = with_rhs(s->R, [&](auto RC) -> int {
      constexpr int R = decltype(RC)::value;
      hipLaunchKernelGGL(k_cocgz_xflush<R>, dim3(s->grid), dim3(kBlock), 0, st, s->n, s->pf.as<double2>(),
                         s->x.as<double2>(), s->state.as<SysState>());
      EFM_LAUNCH_CHECK();
      return EFM_OK;
    });
    if (rc != EFM_OK)
      return rc;
  }
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
