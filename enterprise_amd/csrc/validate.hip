// All-vertex check of a BFS depth array on the device, behind
// ent_bfs_validate_device (ent_bfs.h). The reference checks VALIDATE_TIMES
// random reached vertices and their two-hop neighbourhoods on the host
// (validate.h:31-103); here every vertex and every edge is checked, in one
// pass over the CSR: 8 B per edge (adjacency entry and the neighbour's depth)
// plus 12 B per vertex.
//
// One lane per vertex scans a row shorter than kLaneRow; longer rows are
// scanned by the whole wavefront, one row at a time, 64 entries per step.
#include <hip/hip_runtime.h>

#include "ent_bfs.h"
#include "ent_common.h"
#include "ent_device.h"
#include "ent_internal.h"

namespace ent {
namespace {

constexpr int kLaneRow = 32;

// Edge rule between depths dv and du: both unreached, or both reached and at
// most one level apart.
__device__ __forceinline__ bool edge_ok(int dv, int du) {
  if ((dv < 0) != (du < 0)) return false;
  return dv < 0 || (du <= dv + 1 && du >= dv - 1);
}

__global__ __launch_bounds__(256) void k_validate(const long long *off, const int *adj, const int *depth, int n,
                                                  int root, unsigned long long *counts) {
  const int lane = lane_id();
  const long long nthreads = (long long)gridDim.x * blockDim.x;
  unsigned bad_value = 0, bad_edge = 0, no_parent = 0;
  // Wave-uniform trip count: every lane runs the same iterations (the
  // wavefront scans of long rows need all lanes).
  for (long long base = (long long)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n; base += nthreads) {
    const long long v = base + lane;
    const bool valid = v < n;
    const int dv = valid ? depth[v] : -1;
    long long s = 0, e = 0;
    bool edge_bad = false, parent = true;
    if (valid) {
      if (dv < -1 || dv >= n) ++bad_value;
      s = off[v];
      e = off[v + 1];
      parent = dv < 0 || v == root;
      if (e - s < kLaneRow) {
        for (long long i = s; i < e; ++i) {
          const int du = depth[adj[i]];
          edge_bad |= !edge_ok(dv, du);
          parent |= dv > 0 && du == dv - 1;
        }
      }
    }
    unsigned long long wm = __ballot(valid && e - s >= kLaneRow);
    while (wm) {
      const int src = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const long long ss = __shfl(s, src), ee = __shfl(e, src);
      const int d = __shfl(dv, src);
      bool eb = false, pa = false;
      for (long long q = ss + lane; q < ee; q += kWave) {
        const int du = depth[adj[q]];
        eb |= !edge_ok(d, du);
        pa |= d > 0 && du == d - 1;
      }
      const bool any_bad = __ballot(eb) != 0, any_parent = __ballot(pa) != 0;
      if (lane == src) {
        edge_bad = any_bad;
        parent = parent || any_parent;
      }
    }
    if (valid) {
      bad_edge += edge_bad ? 1u : 0u;
      no_parent += parent ? 0u : 1u;
    }
  }
  const int c1 = wave_sum(int(bad_value)), c2 = wave_sum(int(bad_edge)), c3 = wave_sum(int(no_parent));
  if (lane == 0) {
    if (c1) atomicAdd(counts + 1, (unsigned long long)c1);
    if (c2) atomicAdd(counts + 2, (unsigned long long)c2);
    if (c3) atomicAdd(counts + 3, (unsigned long long)c3);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) counts[0] = depth[root] != 0 ? 1ull : 0ull;
}

}  // namespace
}  // namespace ent

using namespace ent;

extern "C" int ent_bfs_validate_device(ent_graph *g, int32_t root, const int32_t *depth_dev, int64_t counts[4]) {
  clear_error();
  if (!g || !depth_dev || !counts) return set_error(ENT_EINVAL, "NULL argument");
  if (root < 0 || root >= g->n) return set_error(ENT_EINVAL, "root %d outside [0, %lld)", root, (long long)g->n);
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  unsigned long long *c = reinterpret_cast<unsigned long long *>(g->aux);
  HIP_TRY(hipMemsetAsync(c, 0, 4 * sizeof(unsigned long long), g->stream));
  const int grid = int(std::min<int64_t>(int64_t(g->cus) * 8, (g->n + 255) / 256));
  hipLaunchKernelGGL(k_validate, dim3(std::max(grid, 1)), dim3(256), 0, g->stream, g->off, g->adj, depth_dev,
                     int(g->n), int(root), c);
  HIP_TRY(hipGetLastError());
  unsigned long long h[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(h, c, sizeof h, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  for (int k = 0; k < 4; ++k) counts[k] = int64_t(h[k]);
  return ENT_OK;
}
