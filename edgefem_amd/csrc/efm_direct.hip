// Direct solve of a CSR system (plus rank-1 port terms) by a dense LU on the GPU
// (rocSOLVER getrf/getrs): the SparseLU path of solve_linear (src/solver.cpp:11-33), used
// when use_direct is set and as the fallback after a failed iterative solve
// (src/solver.cpp:104-115).  Dense storage limits it to small systems (n <= 40,000,
// 25.6 GB); the reference's sparse factor would not fit for the large ones either.
#include "efm_internal.h"

#include <algorithm>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

using namespace efm;

namespace {

constexpr int64_t kMaxDirectN = 40000;

// Column-major dense A from the CSR rows: A[i + j n] = val(i, j).
__global__ void k_csr_to_dense(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                               const double2 *__restrict__ val, double2 *__restrict__ A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
    A[i + (int64_t)col[k] * n] = val[k];
}

// A += c_p w_p w_p^H for P terms of len entries (index -1: padding); one thread per (p, a, b).
__global__ void k_dense_lowrank(int64_t n, int P, int64_t len, const int32_t *__restrict__ gidx,
                                const double2 *__restrict__ gval, const double2 *__restrict__ coef,
                                double2 *__restrict__ A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = len * len;
  if (t >= (int64_t)P * per)
    return;
  const int p = (int)(t / per);
  const int64_t a = (t - p * per) / len, b = (t - p * per) % len;
  const int32_t i = gidx[p * len + a], j = gidx[p * len + b];
  if (i < 0 || j < 0)
    return;
  const double2 wi = gval[p * len + a], wj = gval[p * len + b];
  const double2 v = cmul(coef[p], cmul(wi, make_double2(wj.x, -wj.y)));
  // distinct (i, j) per term: the port lists hold each edge once, so no two threads of one
  // term write the same entry; the terms are applied one launch each
  double2 &e = A[i + (int64_t)j * n];
  e = cadd(e, v);
}

// system-minor [i * nrhs + s] <-> column-major [i + s * n]
__global__ void k_to_colmajor(int64_t n, int nrhs, const double2 *__restrict__ in, double2 *__restrict__ out,
                              int forward) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nrhs)
    return;
  const int64_t i = t / nrhs;
  const int s = (int)(t - i * nrhs);
  if (forward)
    out[i + (int64_t)s * n] = in[t];
  else
    out[t] = in[i + (int64_t)s * n];
}

struct Handle {
  rocblas_handle h = nullptr;
  ~Handle() {
    if (h)
      (void)rocblas_destroy_handle(h);
  }
};

}  // namespace

extern "C" int efm_direct_solve(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *values,
                                int n_lowrank, int64_t lowrank_len, const int32_t *lr_idx, const double *lr_val,
                                const double *h_lr_coef, const double *B, double *X, int nrhs, int *h_info,
                                void *stream) {
  EFM_REQUIRE(n >= 1 && row_ptr && col_idx && values && B && X && h_info, "efm_direct_solve: NULL argument");
  EFM_REQUIRE(n <= kMaxDirectN, "efm_direct_solve: n above 40000 (the dense factor would not fit)");
  EFM_REQUIRE(nrhs >= 1, "efm_direct_solve: nrhs < 1");
  EFM_REQUIRE(n_lowrank == 0 || (lr_idx && lr_val && h_lr_coef && lowrank_len > 0),
              "efm_direct_solve: low-rank terms incomplete");
  hipStream_t st = (hipStream_t)stream;
  efm::DevBuf A, Bc, ipiv, info, coef;
  int rc;
  if ((rc = A.alloc((size_t)n * n * sizeof(double2))) || (rc = Bc.alloc((size_t)n * nrhs * sizeof(double2))) ||
      (rc = ipiv.alloc((size_t)n * sizeof(rocblas_int))) || (rc = info.alloc(sizeof(rocblas_int))) ||
      (rc = coef.alloc((size_t)std::max(1, n_lowrank) * sizeof(double2))))
    return rc;
  EFM_HIP_CHECK(hipMemsetAsync(A.p, 0, (size_t)n * n * sizeof(double2), st));
  hipLaunchKernelGGL(k_csr_to_dense, dim3(efm::grid_for(n, 256)), dim3(256), 0, st, n, row_ptr, col_idx,
                     (const double2 *)values, A.as<double2>());
  EFM_LAUNCH_CHECK();
  if (n_lowrank > 0) {
    EFM_HIP_CHECK(hipMemcpyAsync(coef.p, h_lr_coef, n_lowrank * sizeof(double2), hipMemcpyHostToDevice, st));
    for (int p = 0; p < n_lowrank; ++p) {
      hipLaunchKernelGGL(k_dense_lowrank, dim3(efm::grid_for(lowrank_len * lowrank_len, 256)), dim3(256), 0, st, n,
                         1, lowrank_len, lr_idx + (int64_t)p * lowrank_len,
                         (const double2 *)lr_val + (int64_t)p * lowrank_len, coef.as<double2>() + p, A.as<double2>());
      EFM_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_to_colmajor, dim3(efm::grid_for(n * nrhs, 256)), dim3(256), 0, st, n, nrhs,
                     (const double2 *)B, Bc.as<double2>(), 1);
  EFM_LAUNCH_CHECK();
  Handle hd;
  EFM_REQUIRE(rocblas_create_handle(&hd.h) == rocblas_status_success, "efm_direct_solve: rocblas handle");
  EFM_REQUIRE(rocblas_set_stream(hd.h, st) == rocblas_status_success, "efm_direct_solve: rocblas stream");
  const rocblas_int nn = (rocblas_int)n;
  EFM_REQUIRE(rocsolver_zgetrf(hd.h, nn, nn, (rocblas_double_complex *)A.p, nn, ipiv.as<rocblas_int>(),
                               info.as<rocblas_int>()) == rocblas_status_success,
              "efm_direct_solve: rocsolver_zgetrf failed");
  EFM_REQUIRE(rocsolver_zgetrs(hd.h, rocblas_operation_none, nn, (rocblas_int)nrhs, (rocblas_double_complex *)A.p,
                               nn, ipiv.as<rocblas_int>(), (rocblas_double_complex *)Bc.p,
                               nn) == rocblas_status_success,
              "efm_direct_solve: rocsolver_zgetrs failed");
  hipLaunchKernelGGL(k_to_colmajor, dim3(efm::grid_for(n * nrhs, 256)), dim3(256), 0, st, n, nrhs,
                     (const double2 *)Bc.p, (double2 *)X, 0);
  EFM_LAUNCH_CHECK();
  rocblas_int hinfo = 0;
  EFM_HIP_CHECK(hipMemcpyAsync(&hinfo, info.p, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  *h_info = (int)hinfo;  // > 0: U(info, info) is exactly zero (singular matrix)
  return EFM_OK;
}
