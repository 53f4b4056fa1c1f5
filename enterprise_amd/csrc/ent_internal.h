// Host-side definition of the opaque graph handle, shared by the
// single-source (bfs_plain.hip) and batched (bfs_batch.hip) translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "ent_bfs.h"
#include "ent_common.h"

struct ent_graph;
namespace ent {
struct Args;
struct Control;
struct BatchWorkspace;
template <class T>
inline void ignore(T) {}

// Resident CSR shared by every handle made with ent_graph_share; freed with
// the last handle.
struct GraphStore {
  int device = 0;
  long long *off = nullptr;              // int64[n + 1]
  int *adj = nullptr;                    // int32[m]
  unsigned long long *nzmask = nullptr;  // bit v: degree(v) > 0
  std::vector<int32_t> h_deg;            // host copy of degrees (for the static lists)
  ~GraphStore() {
    if (hipSetDevice(device) != hipSuccess) return;
    ignore(hipFree(off));
    ignore(hipFree(adj));
    ignore(hipFree(nzmask));
  }
};
// Reorders every adjacency row by descending floor(log2(neighbour degree)),
// stably (adj_order.hip).
int order_rows_by_degree(ent_graph *g);
// ent_graph_upload_ex with the adjacency either in host memory or already in
// device memory of `device` (adj_on_device; copied device to device).
int upload_graph(const int64_t *beg_pos, const int32_t *adj, bool adj_on_device, int64_t n, int64_t m, int device,
                 int32_t flags, ent_graph **out);
}  // namespace ent

struct ent_graph {
  int device = 0;
  int64_t n = 0, m = 0, n_active = 0, nwords = 0;
  hipStream_t stream = nullptr;
  std::shared_ptr<ent::GraphStore> store;  // resident CSR (shared)
  long long *off = nullptr;                 // = store->off
  int *adj = nullptr;                       // = store->adj
  unsigned long long *nzmask = nullptr;     // = store->nzmask
  int *hub_block = nullptr;
  int *piece_v = nullptr;
  int *piece_k = nullptr;
  int n_hub_block = 0, n_pieces = 0;
  int lists_t_block = -1, lists_t_grid = -1, lists_piece = -1;
  int *hub_slot = nullptr;  // LDS hub table source (hub_slots entries)
  int *hub_tag = nullptr;   // level at which hub_slot[s] joined the frontier
  int hub_slots = 0;        // entries built (0: none)
  int hub_min = 0;          // smallest degree in the table
  int lists_hub = -1, lists_hub_tw = -1;
  int *depth = nullptr;
  void *aux = nullptr;     // 256 bytes of device scratch for host-side calls (counters)
  unsigned long long *hashbuf = nullptr;  // kHashRows row checksums (ent_device_row_hash)
  int *parent = nullptr;   // parent array of ent_bfs_parents (allocated on first use)
  int trace_launches = 0;  // launches recorded by the last counting search
  void *stage[2] = {nullptr, nullptr};  // pinned staging for copies to pageable host memory
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  unsigned long long *visited = nullptr;
  unsigned long long *front[2] = {nullptr, nullptr};
  unsigned char *mark = nullptr;
  // Every mark is zero: the last search on the handle completed (a top-down
  // level marks only unvisited vertices, and the inspection that always
  // follows it clears every mark it finds). False until the first search.
  bool marks_clean = false;
  int *qbuf[2] = {nullptr, nullptr};  // striped frontier queues (kQStripes x qcap_s)
  int qcap_s = 0;
  ent::Control *ctl = nullptr;
  ent::Control *h_ctl = nullptr;           // pinned; the persistent form's state copy
  unsigned long long *h_mir = nullptr;     // host-coherent tagged state mirror, written by the kernels
  unsigned long long *h_mir_dev = nullptr; // the same memory as a device pointer
  unsigned search_seq = 0;                 // number of the latest search on this handle (never 0)
  ent_bfs_config cfg{};
  int cus = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  unsigned char graph_args[512] = {};  // launch arguments baked into the captured graph
  int graph_bs = 0, graph_grid = 0, graph_k = 0;
  size_t dev_bytes = 0;
  int last_grid = 0;
  int expect_launches = 12;
  ent::BatchWorkspace *batch = nullptr;  // lazily allocated multi-source workspace
  std::mutex mu;
};


#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ent::set_error(ENT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));  \
  } while (0)

namespace ent {

// Rows one ent_device_row_hash launch sums (its per-handle buffer's length).
constexpr int kHashRows = 512;

inline int dev_alloc(ent_graph *g, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    ignore(hipGetLastError());
    return set_error(ENT_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  g->dev_bytes += bytes;
  return ENT_OK;
}

inline void dev_free(ent_graph *g, void *p, size_t bytes) {
  if (!p) return;
  ignore(hipFree(p));
  g->dev_bytes -= std::min(g->dev_bytes, bytes);
}

// Frees the batched workspace (defined in bfs_batch.hip).
void batch_release(ent_graph *g);

// Device-to-host copy on g's stream, returning when it is complete (pinned
// or small destinations in one DMA, pageable ones through pinned staging).
// The caller holds g->mu.
int copy_to_host(ent_graph *g, void *dst, const void *src, size_t bytes);


}  // namespace ent
