// Batched BiCGSTAB over many independent systems of different sizes (periodic unit cells):
// the reduced matrices of all cells are concatenated into one block-diagonal CSR, and
// every cell carries R excitations.  Each (cell, excitation) pair is its own Krylov
// system with its own alpha, beta, omega, rho and convergence flag.
//
// Row tiling: every block owns a contiguous row range inside one cell, so the partial
// sums of a block belong to one cell and the finalisation of cell c reduces the
// contiguous block range of c (plus one slot per port term of c).  Nodes of the
// auxiliary preconditioner are tiled the same way.
//
// Iteration (van der Vorst 1992, right-preconditioned with P = D^{-1} + G Dn^{-1} G^T,
// the periodic auxiliary preconditioner with a complex incidence; the port loading
// sum_p c_p w_p w_p^H of every cell is applied after each SpMV):
//   rho_i = r0^H r; beta = (rho_i/rho_{i-1})(alpha/omega); p = r + beta (p - omega v)
//   y = P p; v = A y; alpha = rho_i/(r0^H v); s = r - alpha v; zz = P s; t = A zz;
//   omega = (t^H s)/(t^H t); x += alpha y + omega zz; r = s - omega t
// This is the many-cell form of calculate_sparams_periodic (src/assemble_maxwell.cpp:
// 576-635) called per unit cell by UnitCellDesign (python/edgefem/designs/
// unit_cell.py:700-777); the single-cell path is efm_solver_* with the same kernels'
// formulas.
#include "efm_internal.h"

#include <algorithm>
#include <vector>

using namespace efm;

namespace {

constexpr int kBlock = 256;
constexpr int kRowsPerTile = 1024;
constexpr int kLanes = 8;  // SpMV lanes per row
constexpr int kMaxR = 2;

struct CellSys {
  double2 rho, alpha, beta, omega;
  double bnorm2, rnorm2;
  int active, iters, breakdown, restart;  // restart: take r0 = r, p = r at the next p step
  int nrestart, pad[3];
};

constexpr int kMaxRestarts = 20;

// true if any of the R systems of `cell` is active (block-uniform)
template <int R> __device__ inline bool cell_active(const CellSys *st, int64_t cell) {
  bool a = false;
#pragma unroll
  for (int s = 0; s < R; ++s)
    a = a || st[cell * R + s].active != 0;
  return a;
}

struct Tile {
  int32_t cell;
  int32_t pad;
  int64_t r0, r1;
};

__device__ inline double2 bil(double2 a, double2 b) { return cmul(a, b); }

// block sum of V complex values; result in lds[0..V) for all threads
template <int V>
__device__ inline void bsum(double2 (&v)[V], double2 *lds) {
#pragma unroll
  for (int s = 0; s < V; ++s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v[s].x += __shfl_xor(v[s].x, off, 64);
      v[s].y += __shfl_xor(v[s].y, off, 64);
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int s = 0; s < V; ++s)
      lds[w * V + s] = v[s];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int s = 0; s < V; ++s) {
      double2 a = lds[s];
      for (int k = 1; k < kBlock / 64; ++k)
        a = cadd(a, lds[k * V + s]);
      lds[s] = a;
    }
  __syncthreads();
}

template <int V>
__device__ inline void put_partials(double2 (&v)[V], double2 *lds, double2 *partial, int64_t slot) {
  bsum<V>(v, lds);
  if (threadIdx.x < V)
    partial[slot * V + threadIdx.x] = lds[threadIdx.x];
}

enum : int { kInit = 0, kRho0 = 1, kAlpha = 2, kOmega = 3, kX = 4 };

// One block per cell: reduce the cell's partial slots (row tiles [b0, b1) and port slots
// [pb0, pb1)) and advance the R systems of the cell.
template <int R, int STAGE>
__global__ void __launch_bounds__(kBlock)
    k_cells_finalize(int64_t ncells, const int32_t *__restrict__ cell_b0, int npt, int64_t port_base,
                     const double2 *__restrict__ partial, CellSys *__restrict__ st, int *__restrict__ flags,
                     double tol2, int max_it) {
  constexpr int V = (STAGE == kOmega || STAGE == kX) ? 2 * R : R;
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const int64_t c = blockIdx.x;
  if (c >= ncells)
    return;
  double2 v[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    v[s] = make_double2(0, 0);
  for (int32_t b = cell_b0[c] + threadIdx.x; b < cell_b0[c + 1]; b += kBlock)
#pragma unroll
    for (int s = 0; s < V; ++s)
      v[s] = cadd(v[s], partial[(int64_t)b * V + s]);
  if (STAGE == kAlpha || STAGE == kOmega)
    for (int p = threadIdx.x; p < npt; p += kBlock)
#pragma unroll
      for (int s = 0; s < V; ++s)
        v[s] = cadd(v[s], partial[(port_base + c * npt + p) * V + s]);
  bsum<V>(v, lds);
  if (threadIdx.x != 0)
    return;
#pragma unroll
  for (int s = 0; s < R; ++s) {
    CellSys &S = st[c * R + s];
    const double2 tot = lds[s];
    if (STAGE == kInit) {
      S.bnorm2 = tot.x;
      S.rnorm2 = tot.x;
      S.rho = make_double2(1, 0);
      S.alpha = make_double2(1, 0);
      S.omega = make_double2(1, 0);
      S.beta = make_double2(0, 0);
      S.iters = 0;
      S.breakdown = 0;
      S.restart = 0;
      S.nrestart = 0;
      S.active = tot.x > 0.0 ? 1 : 0;
    } else if (!S.active) {
    } else if (STAGE == kRho0) {
      if (tot.x == 0.0 && tot.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.beta = cmul(cdiv(tot, S.rho), cdiv(S.alpha, S.omega));
        S.rho = tot;
      }
    } else if (STAGE == kAlpha) {
      S.restart = 0;
      if (tot.x == 0.0 && tot.y == 0.0) {
        S.breakdown = 1;
        S.active = 0;
      } else {
        S.alpha = cdiv(S.rho, tot);
      }
    } else if (STAGE == kOmega) {
      const double tt = lds[R + s].x;
      S.omega = (tt > 0.0) ? cscale(tot, 1.0 / tt) : make_double2(0, 0);
    } else if (STAGE == kX) {
      S.iters += 1;
      S.rnorm2 = lds[R + s].x;
      if (S.rnorm2 <= tol2 * S.bnorm2) {
        S.active = 0;
      } else if (S.iters >= max_it) {
        S.active = 0;
      } else if ((S.omega.x == 0.0 && S.omega.y == 0.0) || (tot.x == 0.0 && tot.y == 0.0) ||
                 cabs2(tot) < 1e-30 * S.rnorm2 * S.rnorm2) {
        // breakdown (omega = 0 or r0^H r ~ 0): restart from the current iterate with the
        // shadow residual r0 = r, so rho = |r|^2 and p = r
        if (S.nrestart >= kMaxRestarts) {
          S.breakdown = 1;
          S.active = 0;
        } else {
          S.nrestart += 1;
          S.restart = 1;
          S.rho = make_double2(S.rnorm2, 0.0);
          S.beta = make_double2(0, 0);
          S.omega = make_double2(1, 0);
          S.alpha = make_double2(1, 0);
        }
      } else {
        S.beta = cmul(cdiv(tot, S.rho), cdiv(S.alpha, S.omega));
        S.rho = tot;
      }
    }
    flags[c * R + s] = S.active;
  }
}

// x = 0, r = r0 = b, p = v = 0; partials |b|^2
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_init(const Tile *__restrict__ tiles, const double2 *__restrict__ b, double2 *__restrict__ x,
                 double2 *__restrict__ r, double2 *__restrict__ r0, double2 *__restrict__ p,
                 double2 *__restrict__ v, double2 *__restrict__ partial) {
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const Tile t = tiles[blockIdx.x];
  double2 d[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    d[s] = make_double2(0, 0);
  for (int64_t i = t.r0 + threadIdx.x; i < t.r1; i += kBlock)
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int64_t k = i * R + s;
      const double2 bv = b[k];
      x[k] = make_double2(0, 0);
      r[k] = bv;
      r0[k] = bv;
      p[k] = make_double2(0, 0);
      v[k] = make_double2(0, 0);
      d[s].x += cabs2(bv);
    }
  put_partials<R>(d, lds, partial, blockIdx.x);
}

template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_dot_r0r(const Tile *__restrict__ tiles, const double2 *__restrict__ r0, const double2 *__restrict__ r,
                    double2 *__restrict__ partial) {
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const Tile t = tiles[blockIdx.x];
  double2 d[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    d[s] = make_double2(0, 0);
  for (int64_t i = t.r0 + threadIdx.x; i < t.r1; i += kBlock)
#pragma unroll
    for (int s = 0; s < R; ++s)
      d[s] = cfma(cconj(r0[i * R + s]), r[i * R + s], d[s]);
  put_partials<R>(d, lds, partial, blockIdx.x);
}

// p = r + beta (p - omega v)
// (restarting systems: r0 = r, v = 0, p = r)
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_p(const Tile *__restrict__ tiles, const double2 *__restrict__ r, double2 *__restrict__ v,
              double2 *__restrict__ p, double2 *__restrict__ r0, const CellSys *__restrict__ st) {
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))
    return;
  bool act[R], rs[R];
  double2 beta[R], omega[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const CellSys &S = st[(int64_t)t.cell * R + s];
    act[s] = S.active != 0;
    rs[s] = S.restart != 0;
    beta[s] = S.beta;
    omega[s] = S.omega;
  }
  // one (row, system) element per thread: every access of a wavefront is contiguous
  for (int64_t k = t.r0 * R + threadIdx.x; k < t.r1 * R; k += kBlock) {
    const int sys = (int)(k % R);
    bool a = false, rsk = false;
    double2 bt = make_double2(0, 0), om = make_double2(0, 0);
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        a = act[s];
        rsk = rs[s];
        bt = beta[s];
        om = omega[s];
      }
    if (!a)
      continue;
    if (rsk) {
      r0[k] = r[k];
      v[k] = make_double2(0, 0);
      p[k] = r[k];
    } else {
      p[k] = cadd(r[k], bil(bt, csub(p[k], bil(om, v[k]))));
    }
  }
}

// nodal part of P: y = Dn^{-1} G^T in  (node tiles).  kNodeLanes lanes per node split the
// node's edge list, so its gathers are independent; a shuffle sums the lanes' parts.
constexpr int kNodeLanes = 8;

template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_aux_node(const Tile *__restrict__ ntiles, const int32_t *__restrict__ gt_ptr,
                     const int32_t *__restrict__ gt_edge, const double2 *__restrict__ gt_coef,
                     const double2 *__restrict__ dninv, const double2 *__restrict__ in, double2 *__restrict__ y,
                     const CellSys *__restrict__ st) {
  const Tile t = ntiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))
    return;
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[(int64_t)t.cell * R + s].active != 0;
  const int g = threadIdx.x / kNodeLanes, l = threadIdx.x % kNodeLanes;
  constexpr int kNodes = kBlock / kNodeLanes;
  for (int64_t nd0 = t.r0; nd0 < t.r1; nd0 += kNodes) {
    const int64_t nd = nd0 + g;
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    if (nd < t.r1) {
      // two of the lane's incident edges per step, both gathers issued together (past the
      // list's end the second repeats the last edge with a zero coefficient)
      const int32_t qe = gt_ptr[nd + 1];
      for (int32_t q = gt_ptr[nd] + l; q < qe; q += 2 * kNodeLanes) {
        const int32_t q2 = min(q + kNodeLanes, qe - 1);
        const int64_t e = gt_edge[q], e2 = gt_edge[q2];
        const double2 gc = gt_coef[q];
        const double2 gc2 = q + kNodeLanes < qe ? gt_coef[q2] : make_double2(0, 0);
        double2 v[R], v2[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
          v[s] = in[e * R + s];
          v2[s] = in[e2 * R + s];
        }
#pragma unroll
        for (int s = 0; s < R; ++s) {
          acc[s] = cfma(gc, v[s], acc[s]);
          acc[s] = cfma(gc2, v2[s], acc[s]);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
      for (int off = kNodeLanes / 2; off > 0; off >>= 1) {
        acc[s].x += __shfl_xor(acc[s].x, off, kNodeLanes);
        acc[s].y += __shfl_xor(acc[s].y, off, kNodeLanes);
      }
    // lane s of the node writes system s (the node's R values in one contiguous store)
    if (nd < t.r1 && l < R) {
      const double2 di = dninv[nd];
#pragma unroll
      for (int s = 0; s < R; ++s)
        if (s == l && act[s])
          y[nd * R + s] = cmul(di, acc[s]);
    }
  }
}

// out = D^{-1} in + G y  (row tiles)
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_aux_edge(const Tile *__restrict__ tiles, const int2 *__restrict__ en2, const double2 *__restrict__ ga,
                     const double2 *__restrict__ gb, const double2 *__restrict__ dinv, const double2 *__restrict__ y,
                     const double2 *__restrict__ in, double2 *__restrict__ out, const CellSys *__restrict__ st) {
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))
    return;
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[(int64_t)t.cell * R + s].active != 0;
  for (int64_t k = t.r0 * R + threadIdx.x; k < t.r1 * R; k += kBlock) {
    const int64_t e = k / R;
    const int sys = (int)(k - e * R);
    bool a = false;
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys)
        a = act[s];
    if (!a)
      continue;
    const int2 nn = en2[e];
    double2 v = cmul(dinv[e], in[k]);
    // both node values loaded unconditionally (node 0 where an edge has no node), so they
    // issue with the element's other loads; each is added only where the node exists
    const double2 ya = y[(int64_t)(nn.x >= 0 ? nn.x : 0) * R + sys];
    const double2 yb = y[(int64_t)(nn.y >= 0 ? nn.y : 0) * R + sys];
    v = nn.x >= 0 ? cfma(ga[e], ya, v) : v;
    v = nn.y >= 0 ? cfma(gb[e], yb, v) : v;
    out[k] = v;
  }
}

// out = A in (block-diagonal CSR, kLanes lanes per row); MODE 0: partials r0^H out;
// MODE 1: partials out^H other, |out|^2.  The tile's row pointers are staged in LDS once;
// each lane holds up to kEPL entries of its row (branch-free loads: lanes past the row's
// end read its first entry and contribute nothing), and the next step's entries are loaded
// while the current step's gathers are in flight.  Rows longer than kLanes * kEPL take an
// extra loop.
constexpr int kEPL = 3;

template <int R, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_cells_spmv(const Tile *__restrict__ tiles, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                 const double2 *__restrict__ val, const double2 *__restrict__ in, double2 *__restrict__ out,
                 const double2 *__restrict__ other, double2 *__restrict__ partial, const CellSys *__restrict__ st) {
  constexpr int V = MODE == 0 ? R : 2 * R;
  constexpr int kStepRows = kBlock / kLanes;
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  __shared__ int32_t srp[kRowsPerTile + 1];
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))  // the finalisation ignores the partials of inactive systems
    return;
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[(int64_t)t.cell * R + s].active != 0;
  const int nrows = (int)(t.r1 - t.r0);  // <= kRowsPerTile
  for (int i = threadIdx.x; i <= nrows; i += kBlock)
    srp[i] = row_ptr[t.r0 + i];
  __syncthreads();
  const int g = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
  double2 d[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    d[s] = make_double2(0, 0);
  struct Stage {
    double2 a[kEPL];
    int64_t c[kEPL];
  };
  auto stage = [&](Stage &x, int rl) {  // rl: the thread's row in the tile (may be >= nrows)
    const int rc = rl < nrows ? rl : nrows - 1;
    const int32_t kb = srp[rc], ke = rl < nrows ? srp[rc + 1] : kb;
#pragma unroll
    for (int j = 0; j < kEPL; ++j) {
      const int32_t k = kb + lane + j * kLanes;
      const bool inr = k < ke;
      const int32_t kk = inr ? k : kb;
      const double2 av = val[kk];
      x.c[j] = col[kk];
      x.a[j].x = inr ? av.x : 0.0;
      x.a[j].y = inr ? av.y : 0.0;
    }
  };
  Stage cur, nxt;
  stage(cur, g);
  for (int row0 = 0; row0 < nrows; row0 += kStepRows) {
    const int rl = row0 + g;
    double2 xv[kEPL][R];
#pragma unroll
    for (int j = 0; j < kEPL; ++j)
#pragma unroll
      for (int s = 0; s < R; ++s)
        xv[j][s] = in[cur.c[j] * R + s];
    stage(nxt, rl + kStepRows);
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
#pragma unroll
    for (int j = 0; j < kEPL; ++j)
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cfma(cur.a[j], xv[j][s], acc[s]);
    if (rl < nrows) {  // entries past kLanes * kEPL
      const int32_t ke = srp[rl + 1];
      for (int32_t k = srp[rl] + lane + kEPL * kLanes; k < ke; k += kLanes) {
        const double2 a = val[k];
        const int64_t c = col[k];
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[s] = cfma(a, in[c * R + s], acc[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
      for (int off = kLanes / 2; off > 0; off >>= 1) {
        acc[s].x += __shfl_xor(acc[s].x, off, kLanes);
        acc[s].y += __shfl_xor(acc[s].y, off, kLanes);
      }
    // lane s of the row writes system s (the row's R values in one contiguous store)
    if (rl < nrows && lane < R)
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (s != lane || !act[s])
          continue;
        const int64_t k = (t.r0 + rl) * R + s;
        out[k] = acc[s];
        const double2 o = other[k];
        if (MODE == 0) {
          d[s] = cfma(cconj(o), acc[s], d[s]);
        } else {
          d[s] = cfma(cconj(acc[s]), o, d[s]);
          d[R + s].x += cabs2(acc[s]);
        }
      }
    cur = nxt;
  }
  put_partials<V>(d, lds, partial, blockIdx.x);
}

// The same product on entry tiles (the single-system SpMV's scheme): the block of row tile t
// walks its sub-tiles sdesc[soff[t] .. soff[t+1]) of at most kSubNnz entries and kSubRows
// rows; each thread streams 4 entries of a sub-tile (branch-free loads, the next sub-tile's
// in flight during the gathers), the products go to LDS and kSubRL lanes per row sum them.
constexpr int kSubNnz = 1024;
constexpr int kSubRL = 4;
constexpr int kSubRows = kBlock / kSubRL;

template <int R, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_cells_spmv_t(const Tile *__restrict__ tiles, const int32_t *__restrict__ soff, const int4 *__restrict__ sdesc,
                   const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                   const double2 *__restrict__ val, const double2 *__restrict__ in, double2 *__restrict__ out,
                   const double2 *__restrict__ other, double2 *__restrict__ partial, const CellSys *__restrict__ st) {
  constexpr int V = MODE == 0 ? R : 2 * R;
  constexpr int J = kSubNnz / kBlock;
  __shared__ double2 prod[R * kSubNnz];
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))  // the finalisation ignores the partials of inactive systems
    return;
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[(int64_t)t.cell * R + s].active != 0;
  double2 d[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    d[s] = make_double2(0, 0);
  const int g = threadIdx.x / kSubRL, l = threadIdx.x % kSubRL;
  struct Stage {
    int4 dsc;
    double2 a[J];
    int32_t c[J];
    int32_t k0, k1;
  };
  auto stage = [&](Stage &x, int j) {
    x.dsc = sdesc[j];
    const int32_t kz = x.dsc.z < x.dsc.w ? x.dsc.z : 0;  // a valid entry either way
#pragma unroll
    for (int q = 0; q < J; ++q) {
      const int32_t k = x.dsc.z + threadIdx.x + q * kBlock;
      const bool inr = k < x.dsc.w;
      const int32_t kk = inr ? k : kz;
      const double2 av = ld_stream(val + kk);
      x.c[q] = ld_stream(col + kk);
      x.a[q].x = inr ? av.x : 0.0;
      x.a[q].y = inr ? av.y : 0.0;
    }
    const int32_t row = x.dsc.x + g;
    const int32_t rc = row < x.dsc.y ? row : x.dsc.y - 1;
    x.k0 = row_ptr[rc] - x.dsc.z;
    x.k1 = row < x.dsc.y ? row_ptr[rc + 1] - x.dsc.z : x.k0;
  };
  const int j0 = soff[blockIdx.x], j1 = soff[blockIdx.x + 1];
  Stage cur, nxt;
  if (j0 < j1)  // every tile holds a row, so at least one sub-tile
    stage(cur, j0);
  for (int j = j0; j < j1; ++j) {
    double2 zg[J][R];
#pragma unroll
    for (int q = 0; q < J; ++q)
#pragma unroll
      for (int s = 0; s < R; ++s)
        zg[q][s] = in[(int64_t)cur.c[q] * R + s];
    stage(nxt, min(j + 1, j1 - 1));
#pragma unroll
    for (int q = 0; q < J; ++q) {
      const int lk = threadIdx.x + q * kBlock;
#pragma unroll
      for (int s = 0; s < R; ++s)
        prod[s * kSubNnz + lk] = cmul(cur.a[q], zg[q][s]);
    }
    __syncthreads();
    double2 acc[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc[s] = make_double2(0, 0);
    for (int32_t k = cur.k0 + l; k < cur.k1; k += kSubRL)
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[s] = cadd(acc[s], prod[s * kSubNnz + k]);
#pragma unroll
    for (int s = 0; s < R; ++s)
      group_sum<kSubRL>(acc[s]);
    const int32_t row = cur.dsc.x + g;
    if (row < cur.dsc.y && l == 0) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (!act[s])
          continue;
        const int64_t k = (int64_t)row * R + s;
        out[k] = acc[s];
        const double2 o = other[k];
        if (MODE == 0) {
          d[s] = cfma(cconj(o), acc[s], d[s]);
        } else {
          d[s] = cfma(cconj(acc[s]), o, d[s]);
          d[R + s].x += cabs2(acc[s]);
        }
      }
    }
    __syncthreads();  // prod is reused by the next sub-tile
    cur = nxt;
  }
  put_partials<V>(d, lds, partial, blockIdx.x);
}

// Port terms, one block per term q (cell pcell[q], excitation-independent):
//   t_s = w_q^H in_s;  out_s += c_q w_q t_s on the term's edges; the change of the SpMV's
// dot products (MODE 0: r0^H dv; MODE 1: dt^H s and |t|^2) into the term's partial slot.
template <int R, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_cells_ports(const int64_t *__restrict__ poff, const int32_t *__restrict__ pidx, const double2 *__restrict__ pw,
                  const double2 *__restrict__ pcoef, const int32_t *__restrict__ pcell, const double2 *__restrict__ in,
                  double2 *__restrict__ out, const double2 *__restrict__ other, double2 *__restrict__ partial,
                  int64_t slot_base, const CellSys *__restrict__ st) {
  constexpr int V = MODE == 0 ? R : 2 * R;
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const int64_t q = blockIdx.x;
  const int64_t e0 = poff[q], e1 = poff[q + 1];
  const int64_t cell = pcell[q];
  bool act[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    act[s] = st[cell * R + s].active != 0;
  if (!cell_active<R>(st, cell))
    return;
  double2 tv[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    tv[s] = make_double2(0, 0);
  for (int64_t i = e0 + threadIdx.x; i < e1; i += kBlock) {
    const double2 w = cconj(pw[i]);
    const int64_t r = pidx[i];
#pragma unroll
    for (int s = 0; s < R; ++s)
      tv[s] = cfma(w, in[r * R + s], tv[s]);
  }
  bsum<R>(tv, lds);
  double2 t[R];
#pragma unroll
  for (int s = 0; s < R; ++s)
    t[s] = lds[s];
  __syncthreads();
  double2 c[V];
#pragma unroll
  for (int s = 0; s < V; ++s)
    c[s] = make_double2(0, 0);
  const double2 cq = pcoef[q];
  for (int64_t i = e0 + threadIdx.x; i < e1; i += kBlock) {
    const int64_t r = pidx[i];
    const double2 cw = cmul(cq, pw[i]);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (!act[s])
        continue;
      const int64_t k = r * R + s;
      const double2 d = cmul(cw, t[s]);
      const double2 o = out[k];
      const double2 nw = cadd(o, d);
      out[k] = nw;
      if (MODE == 0) {
        c[s] = cfma(cconj(other[k]), d, c[s]);
      } else {
        c[s] = cfma(cconj(d), other[k], c[s]);
        c[R + s].x += cabs2(nw) - cabs2(o);
      }
    }
  }
  put_partials<V>(c, lds, partial, slot_base + q);
}

// s = r - alpha v (in r)
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_s(const Tile *__restrict__ tiles, const double2 *__restrict__ v, double2 *__restrict__ r,
              const CellSys *__restrict__ st) {
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))
    return;
  bool act[R];
  double2 alpha[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const CellSys &S = st[(int64_t)t.cell * R + s];
    act[s] = S.active != 0;
    alpha[s] = S.alpha;
  }
  for (int64_t k = t.r0 * R + threadIdx.x; k < t.r1 * R; k += kBlock) {
    const int sys = (int)(k % R);
    bool a = false;
    double2 al = make_double2(0, 0);
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        a = act[s];
        al = alpha[s];
      }
    if (a)
      r[k] = csub(r[k], bil(al, v[k]));
  }
}

// x += alpha y + omega zz; r = s - omega t; partials r0^H r, |r|^2
template <int R>
__global__ void __launch_bounds__(kBlock)
    k_cells_x(const Tile *__restrict__ tiles, const double2 *__restrict__ y, const double2 *__restrict__ zz,
              const double2 *__restrict__ tvec, const double2 *__restrict__ r0, double2 *__restrict__ x,
              double2 *__restrict__ r, double2 *__restrict__ partial, const CellSys *__restrict__ st) {
  __shared__ double2 lds[(kBlock / 64) * 2 * kMaxR];
  const Tile t = tiles[blockIdx.x];
  if (!cell_active<R>(st, t.cell))  // the finalisation ignores the partials of inactive systems
    return;
  bool act[R];
  double2 alpha[R], omega[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const CellSys &S = st[(int64_t)t.cell * R + s];
    act[s] = S.active != 0;
    alpha[s] = S.alpha;
    omega[s] = S.omega;
  }
  double2 d[2 * R];
#pragma unroll
  for (int s = 0; s < 2 * R; ++s)
    d[s] = make_double2(0, 0);
  for (int64_t k = t.r0 * R + threadIdx.x; k < t.r1 * R; k += kBlock) {
    const int sys = (int)(k % R);
    bool a = false;
    double2 al = make_double2(0, 0), om = make_double2(0, 0);
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        a = act[s];
        al = alpha[s];
        om = omega[s];
      }
    if (!a)
      continue;
    x[k] = cadd(x[k], cadd(bil(al, y[k]), bil(om, zz[k])));
    const double2 rv = csub(r[k], bil(om, tvec[k]));
    r[k] = rv;
    const double2 dr = cfma(cconj(r0[k]), rv, make_double2(0, 0));
    const double rr = cabs2(rv);
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (s == sys) {
        d[s] = cadd(d[s], dr);
        d[R + s].x += rr;
      }
  }
  put_partials<2 * R>(d, lds, partial, blockIdx.x);
}

// Jacobi diagonal with the port terms (ports disjoint): a_rr + c_q |w|^2 on port rows
__global__ void k_cells_dinv(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                             const double2 *__restrict__ val, const double2 *__restrict__ extra,
                             double2 *__restrict__ dinv) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  double2 d = extra[r];
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    if (col[k] == r)
      d = cadd(d, val[k]);
  dinv[r] = (d.x == 0.0 && d.y == 0.0) ? make_double2(1, 0) : cdiv(make_double2(1, 0), d);
}

// diag(G^T A G) per node (the complex incidence of the periodic auxiliary space)
__global__ void k_cells_dn(int64_t N, const int32_t *__restrict__ gt_ptr, const int32_t *__restrict__ gt_edge,
                           const double2 *__restrict__ gt_coef, const int2 *__restrict__ en2,
                           const double2 *__restrict__ ga, const double2 *__restrict__ gb,
                           const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                           const double2 *__restrict__ val, double2 *__restrict__ dninv) {
  const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= N)
    return;
  double2 sum = make_double2(0, 0);
  for (int32_t q = gt_ptr[nd]; q < gt_ptr[nd + 1]; ++q) {
    const int32_t e = gt_edge[q];
    double2 inner = make_double2(0, 0);
    for (int32_t k = row_ptr[e]; k < row_ptr[e + 1]; ++k) {
      const int32_t f = col[k];
      const int2 nn = en2[f];
      double2 cf = make_double2(0, 0);
      if (nn.x == nd)
        cf = cadd(cf, ga[f]);
      if (nn.y == nd)
        cf = cadd(cf, gb[f]);
      if (cf.x != 0.0 || cf.y != 0.0)
        inner = cfma(val[k], cf, inner);
    }
    sum = cfma(gt_coef[q], inner, sum);
  }
  dninv[nd] = (sum.x == 0.0 && sum.y == 0.0) ? make_double2(0, 0) : cdiv(make_double2(1, 0), sum);
}

// Transposed gradient incidence, built on the device: per node its incident edges in
// ascending edge order (the order a sequential fill gives, so every sum over a node's
// list is reproducible), with the coefficient of that node's end (ga + gb when both ends
// of the edge fold onto the node).  Counts by integer atomics, offsets by a scan, a fill
// in any order, then a sort of each node's short list.
__global__ void k_gt_count(int64_t n, int64_t N, const int2 *__restrict__ en2, int32_t *__restrict__ cnt,
                           int *__restrict__ err) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int2 nn = en2[e];
    if (nn.x < -1 || nn.x >= N || nn.y < -1 || nn.y >= N) {
      *err = 1;
      continue;
    }
    if (nn.x >= 0)
      atomicAdd(cnt + nn.x, 1);
    if (nn.y >= 0 && nn.y != nn.x)
      atomicAdd(cnt + nn.y, 1);
  }
}

__global__ void k_gt_fill(int64_t n, const int2 *__restrict__ en2, const double2 *__restrict__ ga,
                          const double2 *__restrict__ gb, const int32_t *__restrict__ ptr,
                          int32_t *__restrict__ cursor, int32_t *__restrict__ gte, double2 *__restrict__ gtc) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int2 nn = en2[e];
    if (nn.x >= 0) {
      const int32_t q = ptr[nn.x] + atomicAdd(cursor + nn.x, 1);
      gte[q] = (int32_t)e;
      gtc[q] = (nn.y == nn.x) ? cadd(ga[e], gb[e]) : ga[e];
    }
    if (nn.y >= 0 && nn.y != nn.x) {
      const int32_t q = ptr[nn.y] + atomicAdd(cursor + nn.y, 1);
      gte[q] = (int32_t)e;
      gtc[q] = gb[e];
    }
  }
}

// insertion sort of each node's list by edge (an edge appears once per list)
__global__ void k_gt_sort(int64_t N, const int32_t *__restrict__ ptr, int32_t *__restrict__ gte,
                          double2 *__restrict__ gtc) {
  for (int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; nd < N; nd += (int64_t)gridDim.x * blockDim.x) {
    const int32_t lo = ptr[nd], hi = ptr[nd + 1];
    for (int32_t i = lo + 1; i < hi; ++i) {
      const int32_t key = gte[i];
      const double2 c = gtc[i];
      int32_t j = i - 1;
      while (j >= lo && gte[j] > key) {
        gte[j + 1] = gte[j];
        gtc[j + 1] = gtc[j];
        --j;
      }
      gte[j + 1] = key;
      gtc[j + 1] = c;
    }
  }
}

// ga[e], gb[e] from the interleaved per-edge coefficients (ga, gb)
__global__ void k_split_pairs(int64_t n, const double2 *__restrict__ ab, double2 *__restrict__ ga,
                              double2 *__restrict__ gb) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    ga[e] = ab[2 * e];
    gb[e] = ab[2 * e + 1];
  }
}

// Jacobi diagonal of the port loading: extra[r] = c_q |w_r|^2 on the rows of term q (one
// block per term; the terms' rows are disjoint, checked on the host)
__global__ void k_port_extra(const int64_t *__restrict__ poff, const int32_t *__restrict__ pidx,
                             const double2 *__restrict__ pw, const double2 *__restrict__ pcoef,
                             double2 *__restrict__ extra) {
  const int64_t q = blockIdx.x;
  const double2 c = pcoef[q];
  for (int64_t i = poff[q] + threadIdx.x; i < poff[q + 1]; i += blockDim.x) {
    const double2 w = pw[i];
    extra[pidx[i]] = cscale(c, w.x * w.x + w.y * w.y);
  }
}

// every column of a row must lie inside the row's cell; err = 1 otherwise
__global__ void k_cells_check(const Tile *__restrict__ tiles, const int64_t *__restrict__ row_off,
                              const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                              int *__restrict__ err) {
  const Tile t = tiles[blockIdx.x];
  const int64_t lo = row_off[t.cell], hi = row_off[t.cell + 1];
  for (int64_t r = t.r0 + threadIdx.x; r < t.r1; r += kBlock) {
    if (row_ptr[r + 1] <= row_ptr[r])  // the SpMV reads a row's first entry for its idle lanes
      err[0] = 1;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if (col[k] < lo || col[k] >= hi)
        err[0] = 1;
  }
}

} // namespace

struct efm_cells {
  int64_t ncells = 0, n = 0, nnz = 0, N = 0;
  int R = 2;
  int npt = 0;  // port terms per cell
  int64_t ntiles = 0, nntiles = 0, nterms = 0;
  efm::DevBuf row_ptr, col, val, dinv, tiles, ntiles_buf, cell_b0;
  efm::DevBuf en2, ga, gb, gt_ptr, gt_edge, gt_coef, dninv, ynode;
  efm::DevBuf poff, pidx, pw, pcoef, pcell;
  efm::DevBuf x, r, r0, p, v, y, zz, t, partial, state, flags;
  efm::DevBuf soff, sdesc;  // entry sub-tiles of the row tiles (nsub = 0: not built)
  int64_t nsub = 0;
  int spmv = 0;  // 1: k_cells_spmv_t on the sub-tiles, 0: row-parallel k_cells_spmv
};

namespace {
template <typename T> int upload(efm::DevBuf &d, const T *h, size_t count, hipStream_t st) {
  int rc = d.alloc(std::max<size_t>(count, 1) * sizeof(T));
  if (rc)
    return rc;
  if (count)
    EFM_HIP_CHECK(hipMemcpyAsync(d.p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  return EFM_OK;
}

std::vector<Tile> make_tiles(const int64_t *off, int64_t ncells, std::vector<int32_t> *first) {
  std::vector<Tile> t;
  if (first)
    first->assign(ncells + 1, 0);
  for (int64_t c = 0; c < ncells; ++c) {
    if (first)
      (*first)[c] = (int32_t)t.size();
    for (int64_t r = off[c]; r < off[c + 1]; r += kRowsPerTile)
      t.push_back(Tile{(int32_t)c, 0, r, std::min<int64_t>(off[c + 1], r + kRowsPerTile)});
  }
  if (first)
    (*first)[ncells] = (int32_t)t.size();
  return t;
}
} // namespace

extern "C" int efm_cells_create(int64_t ncells, const int64_t *h_row_off, const int32_t *row_ptr,
                                const int32_t *col, const double *val, int64_t nnz, const int64_t *h_node_off,
                                const int32_t *h_edge_nodes, const double *h_edge_coefs, int nrhs,
                                const int64_t *h_port_off, const int32_t *h_port_idx, const double *h_port_w,
                                const double *h_port_coef, void *stream, efm_cells **out) {
  EFM_REQUIRE(out, "efm_cells_create: out is NULL");
  *out = nullptr;
  EFM_REQUIRE(ncells >= 1 && h_row_off && row_ptr && col && val && h_node_off && h_edge_nodes &&
                  h_edge_coefs && h_port_off,
              "efm_cells_create: NULL argument");
  EFM_REQUIRE(nrhs == 1 || nrhs == 2, "efm_cells_create: 1 or 2 excitations per cell");
  hipStream_t st = (hipStream_t)stream;
  auto s = std::make_unique<efm_cells>();
  s->ncells = ncells;
  s->R = nrhs;
  s->npt = nrhs;
  s->n = h_row_off[ncells];
  s->nnz = nnz;
  s->N = h_node_off[ncells];
  const int64_t n = s->n, N = s->N;
  s->nterms = ncells * s->npt;
  for (int64_t c = 0; c < ncells; ++c)
    EFM_REQUIRE(h_row_off[c + 1] >= h_row_off[c] && h_node_off[c + 1] >= h_node_off[c],
                "efm_cells_create: offsets must be nondecreasing");
  // port terms: q = c * npt + p, entries [poff[q], poff[q+1]); rows inside the term's cell
  // and disjoint (checked on a sorted copy of the port rows)
  const int64_t npe = h_port_off[s->nterms];
  std::vector<int32_t> pcell(s->nterms);
  {
    std::vector<int32_t> rows(h_port_idx, h_port_idx + npe);
    for (int64_t q = 0; q < s->nterms; ++q) {
      pcell[q] = (int32_t)(q / s->npt);
      for (int64_t i = h_port_off[q]; i < h_port_off[q + 1]; ++i)
        EFM_REQUIRE(h_port_idx[i] >= h_row_off[pcell[q]] && h_port_idx[i] < h_row_off[pcell[q] + 1],
                    "efm_cells_create: port edge outside its cell");
    }
    std::sort(rows.begin(), rows.end());
    EFM_REQUIRE(std::adjacent_find(rows.begin(), rows.end()) == rows.end(),
                "efm_cells_create: port terms must be disjoint");
  }
  std::vector<int32_t> first;
  std::vector<Tile> tiles = make_tiles(h_row_off, ncells, &first);
  std::vector<Tile> ntl = make_tiles(h_node_off, ncells, nullptr);
  s->ntiles = (int64_t)tiles.size();
  s->nntiles = (int64_t)ntl.size();
  int rc;
  const size_t vb = (size_t)n * s->R * sizeof(double2);
  if ((rc = s->row_ptr.alloc((n + 1) * sizeof(int32_t))) || (rc = s->col.alloc(std::max<int64_t>(nnz, 1) * 4)) ||
      (rc = s->val.alloc(std::max<int64_t>(nnz, 1) * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(s->row_ptr.p, row_ptr, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->col.p, col, nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(s->val.p, val, nnz * sizeof(double2), hipMemcpyDeviceToDevice, st));
  {  // the row pointer must end at nnz and every column must stay inside its cell
    int32_t last = -1;
    EFM_HIP_CHECK(hipMemcpyAsync(&last, s->row_ptr.as<int32_t>() + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    EFM_REQUIRE(last == nnz, "efm_cells_create: row_ptr[n] != nnz");
  }
  {  // entry sub-tiles of every row tile for the SpMV: <= kSubNnz entries, <= kSubRows rows
    std::vector<int32_t> hrp(n + 1);
    EFM_HIP_CHECK(hipMemcpy(hrp.data(), s->row_ptr.p, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int4> sd;
    std::vector<int32_t> so(tiles.size() + 1, 0);
    bool ok = true;
    for (size_t ti = 0; ti < tiles.size() && ok; ++ti) {
      so[ti] = (int32_t)sd.size();
      for (int64_t r = tiles[ti].r0; r < tiles[ti].r1;) {
        int64_t r1 = r;
        while (r1 < tiles[ti].r1 && r1 - r < kSubRows && hrp[r1 + 1] - hrp[r] <= kSubNnz)
          ++r1;
        if (r1 == r) {  // a row longer than a sub-tile: keep the row-parallel kernel
          ok = false;
          break;
        }
        sd.push_back(make_int4((int32_t)r, (int32_t)r1, hrp[r], hrp[r1]));
        r = r1;
      }
    }
    so[tiles.size()] = (int32_t)sd.size();
    s->nsub = 0;
    if (ok && !sd.empty()) {
      if ((rc = upload(s->sdesc, sd.data(), sd.size(), st)) || (rc = upload(s->soff, so.data(), so.size(), st)))
        return rc;
      EFM_HIP_CHECK(hipStreamSynchronize(st));  // sd and so are freed at the end of this block
      s->nsub = (int64_t)sd.size();
      s->spmv = 1;
    }
  }
  // gradient incidence (global node indices) and its transpose, on the device
  {
    DevBuf ab, cnt, cursor, work, err;
    if ((rc = upload(s->en2, (const int2 *)h_edge_nodes, n, st)) ||
        (rc = upload(ab, (const double2 *)h_edge_coefs, 2 * n, st)) ||
        (rc = s->ga.alloc(std::max<int64_t>(n, 1) * sizeof(double2))) ||
        (rc = s->gb.alloc(std::max<int64_t>(n, 1) * sizeof(double2))) ||
        (rc = cnt.alloc((N + 1) * sizeof(int32_t))) || (rc = cursor.alloc((N + 1) * sizeof(int32_t))) ||
        (rc = s->gt_ptr.alloc((N + 1) * sizeof(int32_t))) || (rc = work.alloc(scan_work_bytes(N))) ||
        (rc = err.alloc(sizeof(int))))
      return rc;
    EFM_HIP_CHECK(hipMemsetAsync(cnt.p, 0, (N + 1) * sizeof(int32_t), st));
    EFM_HIP_CHECK(hipMemsetAsync(cursor.p, 0, (N + 1) * sizeof(int32_t), st));
    EFM_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), st));
    if (n > 0) {
      hipLaunchKernelGGL(k_split_pairs, dim3(grid_for(n, 256)), dim3(256), 0, st, n, ab.as<double2>(),
                         s->ga.as<double2>(), s->gb.as<double2>());
      EFM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gt_count, dim3(grid_for(n, 256)), dim3(256), 0, st, n, N, s->en2.as<int2>(),
                         cnt.as<int32_t>(), err.as<int>());
      EFM_LAUNCH_CHECK();
    }
    int64_t total = 0;
    if ((rc = exclusive_scan_i32(cnt.as<int32_t>(), N, s->gt_ptr.as<int32_t>(), work.p, st, &total)) != EFM_OK)
      return rc;
    int h_err = 0;
    EFM_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    EFM_REQUIRE(h_err == 0, "efm_cells_create: node index out of range");
    if ((rc = s->gt_edge.alloc(std::max<int64_t>(total, 1) * sizeof(int32_t))) ||
        (rc = s->gt_coef.alloc(std::max<int64_t>(total, 1) * sizeof(double2))))
      return rc;
    if (n > 0) {
      hipLaunchKernelGGL(k_gt_fill, dim3(grid_for(n, 256)), dim3(256), 0, st, n, s->en2.as<int2>(),
                         s->ga.as<double2>(), s->gb.as<double2>(), s->gt_ptr.as<int32_t>(), cursor.as<int32_t>(),
                         s->gt_edge.as<int32_t>(), s->gt_coef.as<double2>());
      EFM_LAUNCH_CHECK();
    }
    if (N > 0) {
      hipLaunchKernelGGL(k_gt_sort, dim3(grid_for(N, 256)), dim3(256), 0, st, N, s->gt_ptr.as<int32_t>(),
                         s->gt_edge.as<int32_t>(), s->gt_coef.as<double2>());
      EFM_LAUNCH_CHECK();
    }
    EFM_HIP_CHECK(hipStreamSynchronize(st));  // the temporaries are released here
  }
  if ((rc = upload(s->tiles, tiles.data(), tiles.size(), st)) ||
      (rc = upload(s->ntiles_buf, ntl.data(), ntl.size(), st)) ||
      (rc = upload(s->cell_b0, first.data(), first.size(), st)) ||
      (rc = upload(s->poff, h_port_off, s->nterms + 1, st)) || (rc = upload(s->pidx, h_port_idx, npe, st)) ||
      (rc = upload(s->pw, (const double2 *)h_port_w, npe, st)) ||
      (rc = upload(s->pcoef, (const double2 *)h_port_coef, s->nterms, st)) ||
      (rc = upload(s->pcell, pcell.data(), pcell.size(), st)))
    return rc;
  {
    DevBuf roff, err;
    if ((rc = upload(roff, h_row_off, ncells + 1, st)) || (rc = err.alloc(sizeof(int))))
      return rc;
    EFM_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), st));
    if (s->ntiles > 0) {
      hipLaunchKernelGGL(k_cells_check, dim3((int)s->ntiles), dim3(kBlock), 0, st, s->tiles.as<Tile>(),
                         roff.as<int64_t>(), s->row_ptr.as<int32_t>(), s->col.as<int32_t>(), err.as<int>());
      EFM_LAUNCH_CHECK();
    }
    int h_err = 0;
    EFM_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    EFM_REQUIRE(h_err == 0, "efm_cells_create: a column leaves its cell, or a row has no entry");
  }
  DevBuf ext;
  if ((rc = ext.alloc(std::max<int64_t>(n, 1) * sizeof(double2))) || (rc = s->dinv.alloc(n * sizeof(double2))) ||
      (rc = s->dninv.alloc(std::max<int64_t>(N, 1) * sizeof(double2))) ||
      (rc = s->ynode.alloc(std::max<int64_t>(N, 1) * s->R * sizeof(double2))) || (rc = s->x.alloc(vb)) ||
      (rc = s->r.alloc(vb)) || (rc = s->r0.alloc(vb)) || (rc = s->p.alloc(vb)) || (rc = s->v.alloc(vb)) ||
      (rc = s->y.alloc(vb)) || (rc = s->zz.alloc(vb)) || (rc = s->t.alloc(vb)) ||
      (rc = s->partial.alloc((size_t)(s->ntiles + s->nterms) * 2 * kMaxR * sizeof(double2))) ||
      (rc = s->state.alloc((size_t)ncells * s->R * sizeof(CellSys))) ||
      (rc = s->flags.alloc((size_t)ncells * s->R * sizeof(int))))
    return rc;
  EFM_HIP_CHECK(hipMemsetAsync(ext.p, 0, std::max<int64_t>(n, 1) * sizeof(double2), st));
  if (s->nterms > 0 && npe > 0) {
    hipLaunchKernelGGL(k_port_extra, dim3((unsigned)s->nterms), dim3(256), 0, st, s->poff.as<int64_t>(),
                       s->pidx.as<int32_t>(), s->pw.as<double2>(), s->pcoef.as<double2>(), ext.as<double2>());
    EFM_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_cells_dinv, dim3(grid_for(n, 256)), dim3(256), 0, st, n, s->row_ptr.as<int32_t>(),
                     s->col.as<int32_t>(), s->val.as<double2>(), ext.as<double2>(), s->dinv.as<double2>());
  EFM_LAUNCH_CHECK();
  if (N > 0) {
    hipLaunchKernelGGL(k_cells_dn, dim3(grid_for(N, 256)), dim3(256), 0, st, N, s->gt_ptr.as<int32_t>(),
                       s->gt_edge.as<int32_t>(), s->gt_coef.as<double2>(), s->en2.as<int2>(), s->ga.as<double2>(),
                       s->gb.as<double2>(), s->row_ptr.as<int32_t>(), s->col.as<int32_t>(), s->val.as<double2>(),
                       s->dninv.as<double2>());
    EFM_LAUNCH_CHECK();
  }
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  *out = s.release();
  return EFM_OK;
}

extern "C" void efm_cells_destroy(efm_cells *s) { delete s; }

extern "C" int efm_cells_set_spmv_variant(efm_cells *s, int variant) {
  EFM_REQUIRE(s, "efm_cells_set_spmv_variant: cells is NULL");
  EFM_REQUIRE(variant == 0 || variant == 1, "efm_cells_set_spmv_variant: variant must be 0 or 1");
  EFM_REQUIRE(variant == 0 || s->nsub > 0,
              "efm_cells_set_spmv_variant: no entry sub-tiles (a row holds more than 1024 entries)");
  s->spmv = variant;
  return EFM_OK;
}

namespace {
template <int R> int cells_solve(efm_cells *s, const double2 *B, double2 *X, const efm_solve_opts *o,
                                 int32_t *h_iters, double *h_resid, hipStream_t st) {
  const Tile *T = s->tiles.as<Tile>();
  const Tile *NT = s->ntiles_buf.as<Tile>();
  CellSys *S = s->state.as<CellSys>();
  double2 *part = s->partial.as<double2>();
  const int nt = (int)s->ntiles, nnt = (int)s->nntiles, nc = (int)s->ncells;
  const double tol2 = o->tolerance * o->tolerance;
  const int max_it = o->max_iterations;
  const int check = o->check_every > 0 ? o->check_every : 32;
  double2 *x = s->x.as<double2>(), *r = s->r.as<double2>(), *r0 = s->r0.as<double2>(), *p = s->p.as<double2>();
  double2 *v = s->v.as<double2>(), *y = s->y.as<double2>(), *zz = s->zz.as<double2>(), *t = s->t.as<double2>();
  auto fin = [&](auto stage_tag) -> int {
    constexpr int STG = decltype(stage_tag)::value;
    hipLaunchKernelGGL((k_cells_finalize<R, STG>), dim3(nc), dim3(kBlock), 0, st, (int64_t)nc,
                       s->cell_b0.as<int32_t>(), s->npt, (int64_t)s->ntiles, part, S, s->flags.as<int>(), tol2,
                       max_it);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  };
  auto precond = [&](const double2 *in, double2 *outv) -> int {
    if (nnt > 0) {
      hipLaunchKernelGGL((k_cells_aux_node<R>), dim3(nnt), dim3(kBlock), 0, st, NT, s->gt_ptr.as<int32_t>(),
                         s->gt_edge.as<int32_t>(), s->gt_coef.as<double2>(), s->dninv.as<double2>(), in,
                         s->ynode.as<double2>(), S);
      EFM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_cells_aux_edge<R>), dim3(nt), dim3(kBlock), 0, st, T, s->en2.as<int2>(),
                       s->ga.as<double2>(), s->gb.as<double2>(), s->dinv.as<double2>(), s->ynode.as<double2>(), in,
                       outv, S);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  };
  auto ports = [&](auto mode_tag, const double2 *in, double2 *outv, const double2 *other) -> int {
    constexpr int MODE = decltype(mode_tag)::value;
    if (s->nterms == 0)
      return EFM_OK;
    hipLaunchKernelGGL((k_cells_ports<R, MODE>), dim3((int)s->nterms), dim3(kBlock), 0, st, s->poff.as<int64_t>(),
                       s->pidx.as<int32_t>(), s->pw.as<double2>(), s->pcoef.as<double2>(), s->pcell.as<int32_t>(),
                       in, outv, other, part, (int64_t)s->ntiles, S);
    EFM_LAUNCH_CHECK();
    return EFM_OK;
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  int rc;
  hipLaunchKernelGGL((k_cells_init<R>), dim3(nt), dim3(kBlock), 0, st, T, B, x, r, r0, p, v, part);
  EFM_LAUNCH_CHECK();
  if ((rc = fin(std::integral_constant<int, kInit>{})))
    return rc;
  std::vector<int> hf((size_t)nc * R);
  auto n_active = [&](int &cnt) -> int {
    EFM_HIP_CHECK(hipMemcpyAsync(hf.data(), s->flags.p, hf.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    cnt = 0;
    for (int f : hf)
      cnt += f != 0;
    return EFM_OK;
  };
  int active = 0;
  if ((rc = n_active(active)))
    return rc;
  int it = 0;
  while (active > 0 && it < max_it) {
    const int chunk = std::min(check, max_it - it);
    for (int c = 0; c < chunk; ++c, ++it) {
      if (it == 0) {
        hipLaunchKernelGGL((k_cells_dot_r0r<R>), dim3(nt), dim3(kBlock), 0, st, T, r0, r, part);
        EFM_LAUNCH_CHECK();
        if ((rc = fin(std::integral_constant<int, kRho0>{})))
          return rc;
      }
      hipLaunchKernelGGL((k_cells_p<R>), dim3(nt), dim3(kBlock), 0, st, T, r, v, p, r0, S);
      EFM_LAUNCH_CHECK();
      if ((rc = precond(p, y)))
        return rc;
      if (s->spmv == 1)
        hipLaunchKernelGGL((k_cells_spmv_t<R, 0>), dim3(nt), dim3(kBlock), 0, st, T, s->soff.as<int32_t>(),
                           s->sdesc.as<int4>(), s->row_ptr.as<int32_t>(), s->col.as<int32_t>(),
                           s->val.as<double2>(), y, v, r0, part, S);
      else
        hipLaunchKernelGGL((k_cells_spmv<R, 0>), dim3(nt), dim3(kBlock), 0, st, T, s->row_ptr.as<int32_t>(),
                           s->col.as<int32_t>(), s->val.as<double2>(), y, v, r0, part, S);
      EFM_LAUNCH_CHECK();
      if ((rc = ports(I0{}, y, v, r0)) || (rc = fin(std::integral_constant<int, kAlpha>{})))
        return rc;
      hipLaunchKernelGGL((k_cells_s<R>), dim3(nt), dim3(kBlock), 0, st, T, v, r, S);
      EFM_LAUNCH_CHECK();
      if ((rc = precond(r, zz)))
        return rc;
      if (s->spmv == 1)
        hipLaunchKernelGGL((k_cells_spmv_t<R, 1>), dim3(nt), dim3(kBlock), 0, st, T, s->soff.as<int32_t>(),
                           s->sdesc.as<int4>(), s->row_ptr.as<int32_t>(), s->col.as<int32_t>(),
                           s->val.as<double2>(), zz, t, r, part, S);
      else
        hipLaunchKernelGGL((k_cells_spmv<R, 1>), dim3(nt), dim3(kBlock), 0, st, T, s->row_ptr.as<int32_t>(),
                           s->col.as<int32_t>(), s->val.as<double2>(), zz, t, r, part, S);
      EFM_LAUNCH_CHECK();
      if ((rc = ports(I1{}, zz, t, r)) || (rc = fin(std::integral_constant<int, kOmega>{})))
        return rc;
      hipLaunchKernelGGL((k_cells_x<R>), dim3(nt), dim3(kBlock), 0, st, T, y, zz, t, r0, x, r, part, S);
      EFM_LAUNCH_CHECK();
      if ((rc = fin(std::integral_constant<int, kX>{})))
        return rc;
    }
    if ((rc = n_active(active)))
      return rc;
  }
  EFM_HIP_CHECK(hipMemcpyAsync(X, x, (size_t)s->n * R * sizeof(double2), hipMemcpyDeviceToDevice, st));
  std::vector<CellSys> hs((size_t)nc * R);
  EFM_HIP_CHECK(hipMemcpyAsync(hs.data(), S, hs.size() * sizeof(CellSys), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  for (size_t k = 0; k < hs.size(); ++k) {
    if (h_iters)
      h_iters[k] = hs[k].iters;
    if (h_resid)
      h_resid[k] = hs[k].bnorm2 > 0 ? sqrt(hs[k].rnorm2 / hs[k].bnorm2) : 0.0;
  }
  return EFM_OK;
}
} // namespace

extern "C" int efm_cells_solve(efm_cells *s, const double *B, double *X, const efm_solve_opts *h_opts,
                               int32_t *h_iters, double *h_resid, void *stream) {
  EFM_REQUIRE(s && B && X && h_opts, "efm_cells_solve: NULL argument");
  EFM_REQUIRE(h_opts->method == EFM_METHOD_BICGSTAB, "efm_cells_solve: the cell batch uses BiCGSTAB");
  EFM_REQUIRE(h_opts->max_iterations >= 1, "efm_cells_solve: max_iterations < 1");
  hipStream_t st = (hipStream_t)stream;
  if (s->R == 1)
    return cells_solve<1>(s, (const double2 *)B, (double2 *)X, h_opts, h_iters, h_resid, st);
  return cells_solve<2>(s, (const double2 *)B, (double2 *)X, h_opts, h_iters, h_resid, st);
}
