// Periodic (Floquet) elimination of slave edges on the sparse matrix.
//
// The reference copies A to a dense matrix and, per master/slave pair (m, s) with
// neither edge PEC, adds a * row s to row m and c * column s to column m, then zeroes
// the rows and columns of the non-PEC slaves with a unit diagonal
// (src/assemble_maxwell.cpp:496-574).  As a whole this is
//     A' = Z (E A F) Z + I_S,   E = I + sum a_k e_m e_s^T,   F = I + sum c_k e_s e_m^T,
// since masters and slaves are disjoint.  Entry-wise, for rows and columns outside S:
//     A'_rc = A_rc + [r = m_k] a_k A_{s_k c} + [c = m_j] c_j A_{r s_j}
//             + [r = m_k][c = m_j] a_k c_j A_{s_k s_j},
// so row r of A' is row r of A plus a_k times row s_k, with every slave column s_j
// folded onto its master m_j with factor c_j (slave columns whose master is PEC are
// dropped).  One thread per output row gathers those entries, sorts them by column and
// sums duplicates in a fixed order; a count pass sizes the output CSR.  Exact zeros
// stay in the pattern (the reference's sparseView drops them).
#include "efm_internal.h"

#include <vector>

using namespace efm;

namespace {

constexpr int kRowBlock = 64;
constexpr int kMaxRowEntries = 96;

__global__ void __launch_bounds__(kRowBlock)
    k_periodic_rows(int64_t n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                    const double2 *__restrict__ val, const int32_t *__restrict__ pair_of_master,
                    const int32_t *__restrict__ slave_state, const int32_t *__restrict__ pm,
                    const int32_t *__restrict__ ps, const double2 *__restrict__ ca,
                    const double2 *__restrict__ cc, int32_t *__restrict__ counts,
                    const int32_t *__restrict__ out_ptr, int32_t *__restrict__ out_col,
                    double2 *__restrict__ out_val, int *__restrict__ err) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  if (slave_state[r] != -2) {  // eliminated slave: unit row
    if (out_col) {
      out_col[out_ptr[r]] = (int32_t)r;
      out_val[out_ptr[r]] = make_double2(1.0, 0.0);
    } else {
      counts[r] = 1;
    }
    return;
  }
  int32_t ce[kMaxRowEntries];
  double2 ve[kMaxRowEntries];
  int m = 0;
  bool overflow = false;
  auto push_row = [&](int64_t src, double2 scale) {
    for (int32_t k = row_ptr[src]; k < row_ptr[src + 1]; ++k) {
      int32_t c = col[k];
      double2 v = cmul(scale, val[k]);
      const int32_t stc = slave_state[c];
      if (stc == -1)
        continue;  // slave column zeroed without a fold (its master is PEC)
      if (stc >= 0) {
        v = cmul(v, cc[stc]);
        c = pm[stc];
      }
      if (m >= kMaxRowEntries) {
        overflow = true;
        return;
      }
      // insertion keeps the columns sorted; equal columns keep their push order
      int j = m++;
      while (j > 0 && ce[j - 1] > c) {
        ce[j] = ce[j - 1];
        ve[j] = ve[j - 1];
        --j;
      }
      ce[j] = c;
      ve[j] = v;
    }
  };
  push_row(r, make_double2(1.0, 0.0));
  const int32_t k = pair_of_master[r];
  if (k >= 0)
    push_row(ps[k], ca[k]);
  if (overflow) {
    err[0] = 1;  // one writer value; any thread that overflows stores the same flag
    if (!out_col)
      counts[r] = 0;
    return;
  }
  int cnt = 0;
  for (int i = 0; i < m;) {
    const int32_t c = ce[i];
    double2 acc = ve[i];
    int j = i + 1;
    while (j < m && ce[j] == c)
      acc = cadd(acc, ve[j++]);
    if (out_col) {
      out_col[out_ptr[r] + cnt] = c;
      out_val[out_ptr[r] + cnt] = acc;
    }
    ++cnt;
    i = j;
  }
  if (!out_col)
    counts[r] = cnt;
}

} // namespace

extern "C" int efm_periodic_reduce(int64_t n, const int32_t *row_ptr, const int32_t *col, const double *val,
                                   int64_t n_pairs, const int32_t *h_master, const int32_t *h_slave,
                                   const double *h_coef_row, const double *h_coef_col,
                                   const uint8_t *h_dirichlet, int32_t *row_ptr_out, int32_t *col_out,
                                   double *val_out, int64_t cap, int64_t *h_nnz, void *stream) {
  EFM_REQUIRE(n > 0 && row_ptr && col && val && row_ptr_out && h_nnz, "efm_periodic_reduce: NULL argument");
  EFM_REQUIRE(n_pairs >= 0 && (n_pairs == 0 || (h_master && h_slave && h_coef_row && h_coef_col)),
              "efm_periodic_reduce: pair arrays missing");
  hipStream_t st = (hipStream_t)stream;
  // host tables: pair of each active master, state of each slave row/column
  std::vector<int32_t> pom(n, -1), sst(n, -2), pmv(std::max<int64_t>(n_pairs, 1)), psv(std::max<int64_t>(n_pairs, 1));
  std::vector<double2> cav(std::max<int64_t>(n_pairs, 1)), ccv(std::max<int64_t>(n_pairs, 1));
  for (int64_t k = 0; k < n_pairs; ++k) {
    const int32_t m = h_master[k], s = h_slave[k];
    EFM_REQUIRE(m >= 0 && m < n && s >= 0 && s < n && m != s, "efm_periodic_reduce: pair index out of range");
    pmv[k] = m;
    psv[k] = s;
    cav[k] = make_double2(h_coef_row[2 * k], h_coef_row[2 * k + 1]);
    ccv[k] = make_double2(h_coef_col[2 * k], h_coef_col[2 * k + 1]);
  }
  for (int64_t k = 0; k < n_pairs; ++k) {
    const int32_t m = pmv[k], s = psv[k];
    const bool mp = h_dirichlet && h_dirichlet[m], sp = h_dirichlet && h_dirichlet[s];
    EFM_REQUIRE(sst[m] == -2, "efm_periodic_reduce: a master edge is also a slave edge");
    if (sp)
      continue;  // PEC slave: already a unit row, not eliminated
    EFM_REQUIRE(sst[s] == -2, "efm_periodic_reduce: a slave edge appears in two pairs");
    sst[s] = mp ? -1 : (int32_t)k;
    if (!mp) {
      EFM_REQUIRE(pom[m] == -1, "efm_periodic_reduce: a master edge appears in two pairs");
      pom[m] = (int32_t)k;
    }
  }
  for (int64_t k = 0; k < n_pairs; ++k)
    EFM_REQUIRE(sst[pmv[k]] == -2, "efm_periodic_reduce: a master edge is also a slave edge");
  DevBuf dpom, dsst, dpm, dps, dca, dcc, cnt, work, err;
  int rc;
  const int64_t P = std::max<int64_t>(n_pairs, 1);
  if ((rc = dpom.alloc(n * 4)) || (rc = dsst.alloc(n * 4)) || (rc = dpm.alloc(P * 4)) || (rc = dps.alloc(P * 4)) ||
      (rc = dca.alloc(P * 16)) || (rc = dcc.alloc(P * 16)) || (rc = cnt.alloc(n * 4)) ||
      (rc = work.alloc(scan_work_bytes(n))) || (rc = err.alloc(sizeof(int))))
    return rc;
  EFM_HIP_CHECK(hipMemcpyAsync(dpom.p, pom.data(), n * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(dsst.p, sst.data(), n * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(dpm.p, pmv.data(), P * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(dps.p, psv.data(), P * 4, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(dca.p, cav.data(), P * 16, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemcpyAsync(dcc.p, ccv.data(), P * 16, hipMemcpyHostToDevice, st));
  EFM_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_periodic_rows, dim3(grid_for(n, kRowBlock)), dim3(kRowBlock), 0, st, n, row_ptr, col,
                     (const double2 *)val, dpom.as<int32_t>(), dsst.as<int32_t>(), dpm.as<int32_t>(),
                     dps.as<int32_t>(), dca.as<double2>(), dcc.as<double2>(), cnt.as<int32_t>(),
                     (const int32_t *)nullptr, (int32_t *)nullptr, (double2 *)nullptr, err.as<int>());
  EFM_LAUNCH_CHECK();
  int64_t total = 0;
  if ((rc = exclusive_scan_i32(cnt.as<int32_t>(), n, row_ptr_out, work.p, st, &total)) != EFM_OK)
    return rc;
  int h_err = 0;
  EFM_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  EFM_REQUIRE(h_err == 0, "efm_periodic_reduce: a reduced row has more than 96 entries");
  *h_nnz = total;
  if (!col_out || !val_out)
    return EFM_OK;
  EFM_REQUIRE(cap >= total, "efm_periodic_reduce: output capacity below the reduced nnz");
  hipLaunchKernelGGL(k_periodic_rows, dim3(grid_for(n, kRowBlock)), dim3(kRowBlock), 0, st, n, row_ptr, col,
                     (const double2 *)val, dpom.as<int32_t>(), dsst.as<int32_t>(), dpm.as<int32_t>(),
                     dps.as<int32_t>(), dca.as<double2>(), dcc.as<double2>(), cnt.as<int32_t>(), row_ptr_out,
                     col_out, (double2 *)val_out, err.as<int>());
  EFM_LAUNCH_CHECK();
  EFM_HIP_CHECK(hipStreamSynchronize(st));
  return EFM_OK;
}
