// Internal declarations shared by the library's translation units.
#pragma once

#include "efm_common.h"

namespace efm {

size_t scan_work_bytes(int64_t n);
int exclusive_scan_i32(const int32_t *in, int64_t n, int32_t *out, void *work, hipStream_t st,
                       int64_t *h_total);

// RAII device buffer used by the host-side drivers.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p)
      (void)hipFree(p);
  }
  int alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = b;
    if (b == 0)
      return EFM_OK;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of " + std::to_string(b) + " bytes failed");
      return EFM_ERR_NOMEM;
    }
    return EFM_OK;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

} // namespace efm

// Device-resident topology of one mesh.
struct efm_topology {
  int64_t N = 0, T = 0, F = 0, n_edges = 0, nnz = 0;
  efm::DevBuf node_ids;   // int64 [N]
  efm::DevBuf tet_nodes;  // int32 [T*4]
  efm::DevBuf tri_nodes;  // int32 [F*3]
  efm::DevBuf edges;      // int64 [n_edges*2]
  efm::DevBuf tet_edges;  // int32 [T*6]
  efm::DevBuf tet_sign;   // int8  [T*6]
  efm::DevBuf tri_edges;  // int32 [F*3]
  efm::DevBuf tri_sign;   // int8  [F*3]
  efm::DevBuf e2t_ptr;    // int32 [n_edges+1] edge -> tet slots (t*6+l), ascending
  efm::DevBuf e2t;        // int32 [6T]
  efm::DevBuf e2f_ptr;    // int32 [n_edges+1] edge -> tri slots (f*3+l), ascending
  efm::DevBuf e2f;        // int32 [3F]
  efm::DevBuf row_ptr;    // int32 [n_edges+1]
  efm::DevBuf col_idx;    // int32 [nnz]
  efm::DevBuf row_idx;    // int32 [nnz]
  efm::DevBuf cptr;       // int32 [nnz+1] nnz -> contributor codes
  efm::DevBuf contrib;    // int32 [36T]   codes t*36 + i*6 + j, ascending t per nnz
  efm::DevBuf edge_nidx;  // int32 [2*n_edges] node positions (lower-ID end, higher-ID end)
  efm::DevBuf n2e_ptr;    // int32 [N+1]   node -> incident edges, ascending
  efm::DevBuf n2e;        // int32 [2*n_edges]
};
