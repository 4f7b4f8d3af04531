// Exclusive prefix sums over int32 counts (int64 totals) for the setup passes.
// Three phases: per-block sums, one block scanning the block sums, per-block
// scan plus offset.  Deterministic; no atomics.
#include "efm_internal.h"

namespace efm {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;  // items per thread
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ inline long long block_exclusive_scan(long long v, long long *lds, long long *total) {
  // Hillis-Steele over the block in LDS; 256 threads.
  int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    long long add = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  long long incl = lds[t];
  if (total)
    *total = lds[kScanBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kScanBlock) k_tile_sums(const int32_t *in, int64_t n,
                                                           long long *sums) {
  __shared__ long long lds[kScanBlock];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    if (i < n)
      s += in[i];
  }
  long long tot;
  block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_sums(long long *sums, int64_t nt,
                                                           long long *grand) {
  __shared__ long long lds[kScanBlock];
  long long carry = 0;
  for (int64_t base = 0; base < nt; base += kScanBlock) {
    int64_t i = base + threadIdx.x;
    long long v = (i < nt) ? sums[i] : 0;
    long long tot;
    long long ex = block_exclusive_scan(v, lds, &tot);
    if (i < nt)
      sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0)
    *grand = carry;
}

template <typename OutT>
__global__ void __launch_bounds__(kScanBlock) k_tile_scan(const int32_t *in, int64_t n,
                                                           const long long *sums, OutT *out,
                                                           const long long *grand) {
  __shared__ long long lds[kScanBlock];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  // each thread owns kScanItems consecutive items
  long long loc[kScanItems];
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    long long v = (i < n) ? in[i] : 0;
    loc[k] = s;
    s += v;
  }
  long long ex = block_exclusive_scan(s, lds, nullptr);
  long long off = sums[blockIdx.x] + ex;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n)
      out[i] = (OutT)(off + loc[k]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    out[n] = (OutT)(*grand);
}
} // namespace

// out[0..n] = exclusive scan of in[0..n-1], out[n] = total.  Returns the total on the
// host (synchronises the stream).  `work` must hold scan_work_bytes(n).
size_t scan_work_bytes(int64_t n) {
  int64_t nt = (n + kScanTile - 1) / kScanTile;
  return (size_t)(nt + 1) * sizeof(long long) + 64;
}

template <typename OutT>
static int exclusive_scan_impl(const int32_t *in, int64_t n, OutT *out, void *work,
                               hipStream_t st, int64_t *h_total) {
  int64_t nt = (n + kScanTile - 1) / kScanTile;
  if (nt < 1)
    nt = 1;
  long long *sums = (long long *)work;
  long long *grand = sums + nt;
  hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nt), dim3(kScanBlock), 0, st, in, n, sums);
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, sums, nt, grand);
  EFM_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_tile_scan<OutT>), dim3((unsigned)nt), dim3(kScanBlock), 0, st, in, n,
                     sums, out, grand);
  EFM_LAUNCH_CHECK();
  if (h_total) {
    long long t = 0;
    EFM_HIP_CHECK(hipMemcpyAsync(&t, grand, sizeof(t), hipMemcpyDeviceToHost, st));
    EFM_HIP_CHECK(hipStreamSynchronize(st));
    *h_total = t;
  }
  return EFM_OK;
}

int exclusive_scan_i32(const int32_t *in, int64_t n, int32_t *out, void *work, hipStream_t st,
                       int64_t *h_total) {
  return exclusive_scan_impl<int32_t>(in, n, out, work, st, h_total);
}

} // namespace efm
