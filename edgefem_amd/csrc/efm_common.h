// Shared helpers of the EdgeFEM MI355X library: error state, complex arithmetic on
// double2/float2, launch checks and the reference's fixed tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>

#include "../../include/edgefem_amd.h"

namespace efm {

// --------------------------------------------------------------------------
// Error state (thread-local message, integer status across the C ABI)
// --------------------------------------------------------------------------
void set_error(const std::string &msg);

#define EFM_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::efm::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));     \
      return EFM_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

#define EFM_REQUIRE(cond, msg)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::efm::set_error(msg);                                                   \
      return EFM_ERR_ARG;                                                      \
    }                                                                          \
  } while (0)

#define EFM_LAUNCH_CHECK() EFM_HIP_CHECK(hipGetLastError())

// --------------------------------------------------------------------------
// Complex arithmetic on HIP vector types
// --------------------------------------------------------------------------
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// a * b + c
__host__ __device__ inline double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm keeps the scaling safe for the magnitudes used here.
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + r * b.y;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + r * b.x;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Promote a stored matrix value (complex double or complex float) to double2.
__device__ inline double2 to_d2(double2 v) { return v; }
__device__ inline double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }

// Streaming (non-temporal) accesses for data used once per kernel, so that they do not
// evict reused data (gathered vectors) from the caches.
typedef double efm_v2d __attribute__((ext_vector_type(2)));
typedef float efm_v2f __attribute__((ext_vector_type(2)));
__device__ inline double2 ld_stream(const double2 *p) {
  const efm_v2d v = __builtin_nontemporal_load(reinterpret_cast<const efm_v2d *>(p));
  return make_double2(v.x, v.y);
}
__device__ inline double2 ld_stream(const float2 *p) {
  const efm_v2f v = __builtin_nontemporal_load(reinterpret_cast<const efm_v2f *>(p));
  return make_double2((double)v.x, (double)v.y);
}
__device__ inline int32_t ld_stream(const int32_t *p) { return __builtin_nontemporal_load(p); }
__device__ inline double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ inline void st_stream(double2 *p, double2 v) {
  efm_v2d w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<efm_v2d *>(p));
}

// A double moved between the lanes of each quad by a DPP quad permutation (a VALU move,
// no LDS traffic).  CTRL 0x4E reads lane ^ 2, 0xB1 reads lane ^ 1.  Every lane of the quad
// must be active.
template <int CTRL>
__device__ inline double dpp_quad(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Butterfly sum over aligned groups of W lanes (W a power of two <= 64), offsets W/2 down to
// 1 in that order, so every lane of a group ends with the same sum.  The last two steps
// (offsets 2 and 1) are DPP quad permutations; wider offsets go through __shfl_xor.  Every
// lane of a group must be active.
template <int W>
__device__ inline void group_sum(double2 &v) {
  static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "group width");
#pragma unroll
  for (int off = W / 2; off >= 4; off >>= 1) {
    v.x += __shfl_xor(v.x, off, W);
    v.y += __shfl_xor(v.y, off, W);
  }
  if constexpr (W >= 4) {
    v.x += dpp_quad<0x4E>(v.x);
    v.y += dpp_quad<0x4E>(v.y);
  }
  if constexpr (W >= 2) {
    v.x += dpp_quad<0xB1>(v.x);
    v.y += dpp_quad<0xB1>(v.y);
  }
}

// --------------------------------------------------------------------------
// Fixed tables of the reference
// --------------------------------------------------------------------------
// Local tet edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3): edge_basis.cpp:9-12, mesh_gmsh.cpp:106-109
__host__ __device__ inline int tet_pair_a(int l) { return (l < 3) ? 0 : (l < 5 ? 1 : 2); }
__host__ __device__ inline int tet_pair_b(int l) {
  return (l == 0) ? 1 : (l == 1) ? 2 : (l == 2) ? 3 : (l == 3) ? 2 : 3;
}
// Local tri edges (0,1),(1,2),(2,0): mesh_gmsh.cpp:110-112
__host__ __device__ inline int tri_pair_a(int l) { return l; }
__host__ __device__ inline int tri_pair_b(int l) { return (l + 1) % 3; }

// Upper-triangle index of (i,j) in a symmetric 6x6 block, i,j in [0,6).
__host__ __device__ inline int sym6(int i, int j) {
  if (i > j) {
    int t = i;
    i = j;
    j = t;
  }
  // rows: 0 -> 0..5, 1 -> 6..10, 2 -> 11..14, 3 -> 15..17, 4 -> 18..19, 5 -> 20
  return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

constexpr double kC0 = 299792458.0;

inline unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

} // namespace efm
