// Parent (BFS tree) output derived from a depth array, behind
// ent_bfs_parents[_device] (ent_bfs.h). The reference outputs depths only
// (bfs_gpu_opt.cuh:446); the tree it would have to pass is Graph500's:
// parent[root] = root, parent[v] a neighbour of v one level closer to the
// root, -1 for unreached vertices.
//
// One lane per vertex: it probes the first neighbours of its row four at a
// time (rows start with high-degree neighbours, which sit on the levels
// nearest the root most often) and stops at the first one with depth one
// less; rows not resolved within the probe go to the whole wavefront, one
// vertex at a time, with a ballot early exit. Every reached vertex other
// than the root has such a neighbour when the depths are a BFS from root;
// vertices that do not are counted in `bad` and get parent -2.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "ent_bfs.h"
#include "ent_common.h"
#include "ent_device.h"
#include "ent_internal.h"

namespace ent {

constexpr int kParentProbe = 16;

__global__ __launch_bounds__(256) void k_parents(const long long *off, const int *adj, const int *depth, int n,
                                                 int root, int *parent, unsigned int *bad) {
  const int lane = lane_id();
  const long long nthreads = (long long)gridDim.x * blockDim.x;
  unsigned int nbad = 0;
  // Wave-uniform trip count: every lane of a wavefront runs the same number
  // of iterations (the ballot loop below needs all lanes).
  for (long long base = (long long)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n; base += nthreads) {
    const long long v = base + lane;
    const bool valid = v < n;
    const int d = valid ? depth[v] : -1;
    int par = -1;
    long long s = 0, e = 0, p = 0;
    bool open = false;  // reached, not the root, parent not found yet
    if (valid && d > 0) {
      s = off[v];
      e = off[v + 1];
      open = true;
      const long long pe = min(e, s + kParentProbe);
      for (p = s; p < pe && open; p += 4) {
        int u[4], du[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = p + j < pe ? adj[p + j] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) du[j] = u[j] >= 0 ? depth[u[j]] : -2;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (open && du[j] == d - 1) {
            par = u[j];
            open = false;
          }
      }
      if (p > pe) p = pe;
    } else if (valid && d == 0) {
      par = int(v);
      open = v != root;  // depth 0 belongs to the root alone
    }
    unsigned long long lmask = __ballot(open && d > 0 && p < e);
    while (lmask) {
      const int src = __ffsll((long long)lmask) - 1;
      lmask &= lmask - 1;
      const long long pp = __shfl(p, src), ee = __shfl(e, src);
      const int dd = __shfl(d, src);
      int found = -1;
      for (long long q0 = pp; q0 < ee; q0 += kWave) {
        const long long q = q0 + lane;
        const int u = q < ee ? adj[q] : -1;
        const bool hit = u >= 0 && depth[u] == dd - 1;
        const unsigned long long hm = __ballot(hit);
        if (hm) {
          found = __shfl(u, __ffsll((long long)hm) - 1);
          break;
        }
      }
      if (lane == src) {
        par = found;
        open = found < 0;
      }
    }
    if (valid) {
      if (open) {  // no neighbour one level up: the depths are not a BFS from root
        par = -2;
        ++nbad;
      }
      parent[v] = par;
    }
  }
  if (nbad) atomicAdd(bad, nbad);
}

}  // namespace ent

using namespace ent;

extern "C" int ent_bfs_parents_device(ent_graph *g, int32_t root, const int32_t *depth_dev, int32_t *parent_dev) {
  clear_error();
  if (!g || !depth_dev || !parent_dev) return set_error(ENT_EINVAL, "NULL argument");
  if (root < 0 || root >= g->n) return set_error(ENT_EINVAL, "root %d outside [0, %lld)", root, (long long)g->n);
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  unsigned int *bad = reinterpret_cast<unsigned int *>(g->aux);
  HIP_TRY(hipMemsetAsync(bad, 0, sizeof(unsigned int), g->stream));
  const int grid = int(std::min<int64_t>(int64_t(g->cus) * 8, (g->n + 255) / 256));
  hipLaunchKernelGGL(k_parents, dim3(std::max(grid, 1)), dim3(256), 0, g->stream, g->off, g->adj, depth_dev,
                     int(g->n), int(root), parent_dev, bad);
  HIP_TRY(hipGetLastError());
  unsigned int h_bad = 0;
  HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof h_bad, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (h_bad)
    return set_error(ENT_EINVAL, "%u reached vertices have no neighbour one level closer to root %d", h_bad, root);
  return ENT_OK;
}

extern "C" int ent_bfs_parents(ent_graph *g, int32_t root, int32_t *depth_host, int32_t *parent_host,
                               ent_stats *st) {
  int rc = ent_bfs_device(g, root, nullptr, st);
  if (rc) return rc;
  if (!g->parent) {
    std::lock_guard<std::mutex> lk(g->mu);
    if (!g->parent && (rc = dev_alloc(g, reinterpret_cast<void **>(&g->parent), sizeof(int) * size_t(g->n))))
      return rc;
  }
  if ((rc = ent_bfs_parents_device(g, root, g->depth, g->parent))) return rc;
  if (depth_host && (rc = ent_copy_to_host(g, g->depth, depth_host, g->n))) return rc;
  if (parent_host && (rc = ent_copy_to_host(g, g->parent, parent_host, g->n))) return rc;
  return ENT_OK;
}
