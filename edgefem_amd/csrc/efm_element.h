// Element geometry shared by the assembly and the gradient kernels: barycentric
// gradients and volume of a tet (src/edge_basis.cpp:14-26) and the Whitney mass entry
// (src/edge_basis.cpp:65-86).
#pragma once

#include "efm_common.h"

namespace efm {

struct V3 {
  double x, y, z;
};
__device__ inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 vscale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 vload(const double *xyz, int32_t node) {
  const double *p = xyz + (int64_t)node * 3;
  return {p[0], p[1], p[2]};
}

// src/edge_basis.cpp:14-26: B = [v0-v3, v1-v3, v2-v3]; grad l_i = column i of B^{-T}
// (row i of B^{-1} = cross product of the other two columns / det); grad l_3 = -sum;
// V = |det B| / 6.
__device__ inline void gradients_and_volume(const V3 v[4], V3 g[4], double &V) {
  V3 b0 = vsub(v[0], v[3]), b1 = vsub(v[1], v[3]), b2 = vsub(v[2], v[3]);
  V3 c12 = vcross(b1, b2), c20 = vcross(b2, b0), c01 = vcross(b0, b1);
  double det = vdot(b0, c12);
  double inv = 1.0 / det;
  g[0] = vscale(c12, inv);
  g[1] = vscale(c20, inv);
  g[2] = vscale(c01, inv);
  g[3] = {-g[0].x - g[1].x - g[2].x, -g[0].y - g[1].y - g[2].y, -g[0].z - g[1].z - g[2].z};
  V = fabs(det) / 6.0;
}

// M_e(i, j) = grad l_b . grad l_d I(a,c) - grad l_b . grad l_c I(a,d)
//           - grad l_a . grad l_d I(b,c) + grad l_a . grad l_c I(b,d),
// I = V/10 on equal indices, V/20 otherwise; edges i = (a,b), j = (c,d).
__device__ inline double whitney_mass_entry(const double G[4][4], double V, int i, int j) {
  const int a = tet_pair_a(i), b = tet_pair_b(i), c = tet_pair_a(j), d = tet_pair_b(j);
  const double Id = V / 10.0, Io = V / 20.0;
  double term = 0.0;
  term += G[b][d] * (a == c ? Id : Io);
  term -= G[b][c] * (a == d ? Id : Io);
  term -= G[a][d] * (b == c ? Id : Io);
  term += G[a][c] * (b == d ? Id : Io);
  return term;
}

} // namespace efm
