// C ABI housekeeping (error string, version) and the port-overlap S extraction.
#include "efm_internal.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace efm {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
} // namespace efm

using namespace efm;

extern "C" const char *efm_last_error(void) { return efm::g_last_error.c_str(); }

extern "C" int efm_version(void) { return 1; }

namespace {

constexpr int kBlock = 256;

// out[j * nrhs + a] = sum_k conj(g[j,k]) * X[idx[j,k] * nrhs + a]; one block per (j, a),
// fixed-order reduction.  Negative indices are padding.
__global__ void __launch_bounds__(kBlock)
    k_port_overlap(int64_t n, int64_t len, const int32_t *__restrict__ gidx,
                   const double2 *__restrict__ gval, const double2 *__restrict__ X, int nrhs,
                   double2 *__restrict__ out) {
  const int j = blockIdx.x / nrhs, a = blockIdx.x % nrhs;
  double2 acc = make_double2(0, 0);
  for (int64_t k = threadIdx.x; k < len; k += kBlock) {
    int32_t e = gidx[(int64_t)j * len + k];
    if (e < 0 || (int64_t)e >= n)
      continue;
    double2 g = gval[(int64_t)j * len + k];
    acc = cadd(acc, cmul(cconj(g), X[(int64_t)e * nrhs + a]));
  }
  __shared__ double2 lds[kBlock];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      lds[threadIdx.x] = cadd(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[blockIdx.x] = lds[0];
}

} // namespace

namespace {

// Staging buffer of efm_port_overlap: one per device and host thread, grown on demand and
// kept until the process ends (never freed, so no hipFree runs at thread or process exit).
constexpr int kMaxDevices = 64;
thread_local void *t_overlap_buf[kMaxDevices] = {};
thread_local size_t t_overlap_bytes[kMaxDevices] = {};

int overlap_staging(size_t bytes, double2 **out) {
  int dev = 0;
  EFM_HIP_CHECK(hipGetDevice(&dev));
  EFM_REQUIRE(dev >= 0 && dev < kMaxDevices, "efm_port_overlap: device ordinal out of range");
  if (t_overlap_bytes[dev] < bytes) {
    const size_t want = std::max<size_t>(bytes, 4096);
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
      set_error("hipMalloc of " + std::to_string(want) + " bytes failed");
      return EFM_ERR_NOMEM;
    }
    if (t_overlap_buf[dev]) {
      // the previous buffer may still be read by work queued on another stream
      EFM_HIP_CHECK(hipDeviceSynchronize());
      (void)hipFree(t_overlap_buf[dev]);
    }
    t_overlap_buf[dev] = p;
    t_overlap_bytes[dev] = want;
  }
  *out = static_cast<double2 *>(t_overlap_buf[dev]);
  return EFM_OK;
}

int overlap_run(int64_t n, int n_ports, int64_t len, const int32_t *gidx, const double *gval,
                const double *X, int nrhs, double *h_out, double2 *dev_out, hipStream_t st) {
  hipLaunchKernelGGL(k_port_overlap, dim3(n_ports * nrhs), dim3(kBlock), 0, st, n, len, gidx,
                     (const double2 *)gval, (const double2 *)X, nrhs, dev_out);
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipMemcpyAsync(h_out, dev_out, (size_t)n_ports * nrhs * sizeof(double2),
                               hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  return EFM_OK;
}

// Fixed staging array of efm_port_overlap_noalloc (see the header): every device that runs
// the library holds its own copy, 64 KiB, until the process ends.
constexpr int kOverlapScratch = 4096;  // complex entries = n_ports * nrhs at most
static __device__ __align__(256) double2 scratch[kOverlapScratch];

std::mutex g_scratch_mu;
double2 *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_address(double2 **out) {
  int dev = 0;
  EFM_HIP_CHECK(hipGetDevice(&dev));
  EFM_REQUIRE(dev >= 0 && dev < kMaxDevices, "efm_port_overlap_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    EFM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EFM_REQUIRE(p != nullptr, "efm_port_overlap_noalloc: scratch symbol has no device address");
    g_scratch_ptr[dev] = static_cast<double2 *>(p);
  }
  *out = g_scratch_ptr[dev];
  return EFM_OK;
}

} // namespace

extern "C" int efm_port_overlap(int64_t n, int n_ports, int64_t len, const int32_t *gidx,
                                const double *gval, const double *X, int nrhs, double *h_out,
                                void *stream) {
  EFM_REQUIRE(gidx && gval && X && h_out, "efm_port_overlap: NULL argument");
  EFM_REQUIRE(n > 0 && n_ports > 0 && len > 0 && nrhs > 0, "efm_port_overlap: bad sizes");
  double2 *buf = nullptr;
  int rc = overlap_staging((size_t)n_ports * nrhs * sizeof(double2), &buf);
  if (rc != EFM_OK)
    return rc;
  return overlap_run(n, n_ports, len, gidx, gval, X, nrhs, h_out, buf, (hipStream_t)stream);
}

extern "C" int efm_port_overlap_noalloc(int64_t n, int n_ports, int64_t len, const int32_t *gidx,
                                        const double *gval, const double *X, int nrhs,
                                        double *h_out, void *stream) {
  EFM_REQUIRE(gidx && gval && X && h_out, "efm_port_overlap_noalloc: NULL argument");
  EFM_REQUIRE(n > 0 && n_ports > 0 && len > 0 && nrhs > 0, "efm_port_overlap_noalloc: bad sizes");
  EFM_REQUIRE((int64_t)n_ports * nrhs <= kOverlapScratch,
              "efm_port_overlap_noalloc: n_ports * nrhs exceeds the 4096-entry scratch array");
  double2 *buf = nullptr;
  int rc = scratch_address(&buf);
  if (rc != EFM_OK)
    return rc;
  return overlap_run(n, n_ports, len, gidx, gval, X, nrhs, h_out, buf, (hipStream_t)stream);
}
