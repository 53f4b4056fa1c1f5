// Upload-time row ordering: every adjacency row is reordered so that
// neighbours of higher degree come first (stable within a class of equal
// floor(log2(degree))). BFS depths do not depend on the order of a row; a
// bottom-up scan stops at its first frontier neighbour, and hubs are the
// neighbours most likely to be in a large frontier (and, in the multi-source
// search, the ones whose source masks stay cached when every row starts with
// them). The reference scans rows in file order (expander.cuh:486-523).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <cstdint>

#include "ent_common.h"
#include "ent_internal.h"

namespace ent {

// vclass[v] = floor(log2(degree(v))) (0 for degree 0 or 1).
__global__ void k_vertex_class(const long long *off, long long n, unsigned char *vclass) {
  const long long nt = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += nt) {
    const long long d = off[v + 1] - off[v];
    vclass[v] = d > 1 ? (unsigned char)(63 - __clzll(d)) : 0;
  }
}

// key[i] = vclass[adj[i]].
__global__ void k_entry_class(const int *adj, long long m, const unsigned char *vclass, unsigned char *key) {
  const long long nt = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += nt) key[i] = vclass[adj[i]];
}

// Reorders g's device adjacency in place (through a second buffer). The CSR
// must not be in use by any search.
int order_rows_by_degree(ent_graph *g) {
  const int64_t n = g->n, m = g->m;
  if (m < 2) return ENT_OK;
  if (m > int64_t(UINT32_MAX) || n > int64_t(UINT32_MAX))
    return set_error(ENT_EINVAL, "row ordering supports m, n < 2^32 (m=%lld)", (long long)m);
  GraphStore &st = *g->store;
  hipStream_t s = g->stream;
  unsigned char *vclass = nullptr, *keys = nullptr, *keys_out = nullptr;
  int *adj_out = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = ENT_OK;
  auto cleanup = [&]() {
    ignore(hipFree(vclass));
    ignore(hipFree(keys));
    ignore(hipFree(keys_out));
    ignore(hipFree(adj_out));
    ignore(hipFree(tmp));
  };
#define ORD_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      rc = set_error(ENT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));  \
      cleanup();                                                                   \
      return rc;                                                                   \
    }                                                                              \
  } while (0)
  ORD_TRY(hipMalloc(&vclass, size_t(n)));
  ORD_TRY(hipMalloc(&keys, size_t(m)));
  ORD_TRY(hipMalloc(&keys_out, size_t(m)));
  ORD_TRY(hipMalloc(&adj_out, sizeof(int) * size_t(m)));
  const int grid = std::max(1, g->cus) * 8;
  hipLaunchKernelGGL(k_vertex_class, dim3(grid), dim3(256), 0, s, st.off, (long long)n, vclass);
  ORD_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_entry_class, dim3(grid), dim3(256), 0, s, st.adj, (long long)m, vclass, keys);
  ORD_TRY(hipGetLastError());
  const long long *beg = st.off, *end = st.off + 1;
  ORD_TRY(rocprim::segmented_radix_sort_pairs_desc(nullptr, tmp_bytes, keys, keys_out, st.adj, adj_out,
                                                   unsigned(m), unsigned(n), beg, end, 0, 6, s));
  ORD_TRY(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
  ORD_TRY(rocprim::segmented_radix_sort_pairs_desc(tmp, tmp_bytes, keys, keys_out, st.adj, adj_out, unsigned(m),
                                                   unsigned(n), beg, end, 0, 6, s));
  ORD_TRY(hipStreamSynchronize(s));
#undef ORD_TRY
  ignore(hipFree(st.adj));
  st.adj = adj_out;
  g->adj = adj_out;
  adj_out = nullptr;
  cleanup();
  return ENT_OK;
}

}  // namespace ent
