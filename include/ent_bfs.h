/*
 * ent_bfs.h - C ABI of the MI355X (gfx950) direction-optimizing BFS.
 *
 * Drop-in for the reference's single-source GPU traversal
 *   template<vertex_t,index_t> int bfs_gpu_coalescing_mem(vertex_t* src_list,
 *       index_t* beg_pos, vertex_t* csr, index_t vert_count,
 *       index_t edge_count, index_t gpu_id)          (bfs_gpu_opt.cuh:293-300)
 * which allocates device memory and uploads the CSR (allocator.cuh:33-151),
 * then runs one level-synchronous BFS per root (bfs_tdbu_clfy_sort,
 * bfs_gpu_opt.cuh:40-288). Here those steps are separate entry points on an
 * opaque graph handle, so a graph is uploaded once and queried many times.
 *
 * Depth convention: int32, root = 0, unreached = -1. The reference keeps an
 * unsigned char depth with unreached = 255 (comm.h:69, :131); values 0..254
 * agree and 255 maps to -1.
 *
 * Errors: every function returns ENT_OK (0) or a negative ENT_E* code (see
 * ent_graph.h) and never exits the process; ent_last_error() gives the text.
 * Threading: any thread may call; calls on one handle are serialised by a
 * per-handle lock and run on the handle's own HIP stream.
 */
#ifndef ENT_BFS_H
#define ENT_BFS_H

#include <stdint.h>

#include "ent_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ent_graph ent_graph;

/* Direction policy. Bottom-up levels scan a vertex's own adjacency for a
 * parent, as the reference's do, so every policy but ENT_DIR_TOPDOWN needs
 * a symmetric adjacency (each edge stored both ways); ENT_DIR_TOPDOWN is
 * exact on any adjacency. */
#define ENT_DIR_AUTO 0      /* Beamer alpha/beta switch, both ways */
#define ENT_DIR_REFERENCE 1 /* reference rule: bottom-up once the frontier's
                               degree sum reaches 0.3 * m at a level >= 1,
                               never back (inspector.cuh:486-499) */
#define ENT_DIR_TOPDOWN 2   /* top-down only */
#define ENT_DIR_BOTTOMUP 3  /* bottom-up only */

/* Launch form. */
#define ENT_LAUNCH_PERSISTENT 0 /* one cooperative launch, grid barriers */
#define ENT_LAUNCH_LEVELS 1     /* one launch per phase, host sync per level
                                   (the reference's control flow) */
#define ENT_LAUNCH_GRAPH 2      /* per-phase launches captured in a hipGraph,
                                   replayed until the search ends */
#define ENT_LAUNCH_STREAM 3     /* per-phase launches queued without host sync,
                                   checked every graph_levels phases */

typedef struct ent_bfs_config {
  int32_t direction;       /* ENT_DIR_* */
  int32_t launch;          /* ENT_LAUNCH_* */
  int32_t alpha;           /* top-down -> bottom-up when m_f * alpha > m_u */
  int32_t beta;            /* bottom-up -> top-down when n_f * beta < n_active
                              and the frontier shrank */
  int32_t t_wave;          /* degree >= t_wave: one wavefront per vertex */
  int32_t t_block;         /* degree >= t_block: one block per vertex */
  int32_t t_grid;          /* degree >= t_grid: vertex split into pieces of
                              `piece` edges spread over the whole grid */
  int32_t piece;           /* edges per grid-class piece (<= t_grid) */
  int32_t block_size;      /* threads per block: 64..1024, multiple of 64 */
  int32_t blocks_per_cu;   /* blocks per compute unit (0 = occupancy max) */
  int32_t bu_probe;        /* neighbours one lane scans in bottom-up before
                              its wavefront takes the vertex over */
  int32_t bu_chunk_words;  /* 64-vertex bitmap words per wavefront chunk in
                              the bottom-up sweep: 1..64 */
  int32_t graph_levels;    /* phases per captured graph / queued batch */
  int32_t queue_small;     /* largest frontier expanded top-down from the
                              frontier queue in one launch (0 = auto,
                              negative = never) */
  int32_t batch_alpha;     /* multi-source search: bottom-up when the joint
                              frontier's degree sum * batch_alpha exceeds the
                              degree sum of vertices still missing sources */
  int32_t grid_blocks;     /* blocks per launch (0 = blocks_per_cu * CUs);
                              the persistent form never exceeds its
                              residency cap */
  int32_t hub_cache;       /* LDS hub table of 2^hub_cache entries (6..13;
                              0 = off): slot s holds the highest-degree vertex
                              v (degree >= t_wave) with v mod 2^hub_cache = s.
                              Bottom-up tests neighbours against the table's
                              frontier hubs before the frontier bitmap
                              (comm.h:172-177, expander.cuh:454-494) */
  int32_t hub_td;          /* with hub_cache: top-down flags hub neighbours in
                              LDS and marks each flagged hub once per block
                              (expander.cuh:96-109, :172-185) */
  int32_t queue_max_degree; /* a frontier goes through the queue only if all
                               its vertices have degree below this
                               (0 = t_block) */
  int32_t count_examined;   /* 1: run the kernel variant that counts the
                               work it examines (ent_stats.examined_*);
                               for measurement, not speed */
  int32_t batch_fused;      /* multi-source search with at most 256 sources
                               per workspace (ent_batch.h): a level that
                               pushes every source word and whose joint
                               frontier's degree sum is at most
                               n / batch_fused is pushed and settled in one
                               launch, with no sweep over every vertex
                               (0 = never) */
} ent_bfs_config;

typedef struct ent_stats {
  int32_t levels;           /* max depth + 1 */
  int32_t td_levels;        /* levels expanded top-down */
  int32_t bu_levels;        /* levels expanded bottom-up */
  int32_t phases;           /* kernel phases run (levels + inspections) */
  int64_t reached_vertices; /* N_c */
  int64_t reached_edges;    /* M_c = sum of degrees over reached vertices,
                               the reference's TEPS numerator
                               (validate.h:115-123) */
  int64_t alg_bytes;        /* 8 M_c + 12 N_c + 4 n */
  double device_ms;         /* device time of the search: constant 100 MHz clock from
                               k_init to the launch that found it finished */
  double host_ms;           /* wall time of the call */
  uint64_t bu_level_mask;   /* bit L set: level L ran bottom-up (L < 64) */
  int32_t grid_blocks;      /* blocks per level launch / persistent grid */
  int32_t block_size;       /* threads per block */
  int64_t examined_edges;   /* with count_examined: adjacency entries loaded
                               (every direction, the early exits included);
                               else -1 */
  int64_t examined_bytes;   /* with count_examined: 8 B per adjacency entry
                               loaded (entry + its status lookup), 8 B per
                               vertex whose offsets were loaded, and the bytes
                               of frontier / visited / mark / depth arrays
                               swept; else -1 */
} ent_stats;

/* Fills cfg with the library defaults. */
void ent_bfs_default_config(ent_bfs_config *cfg);

/* Uploads a CSR graph (host arrays, int64 offsets, int32 adjacency) to
 * `device` and allocates the search workspace. Replaces allocator.cuh:33-151.
 * n < 2^31; adjacency ids must lie in [0, n). */
int ent_graph_upload(const int64_t *beg_pos, const int32_t *adj, int64_t n,
                     int64_t m, int device, ent_graph **out);

/* Upload flags. By default the device copy of every adjacency row is
 * reordered by descending floor(log2(neighbour degree)) (stable within a
 * class): depths and counts do not depend on row order, and bottom-up scans
 * meet frontier hubs first. ENT_UPLOAD_KEEP_ORDER keeps the caller's order
 * (the reference's, expander.cuh:486-523). */
#define ENT_UPLOAD_KEEP_ORDER 1

/* ent_graph_upload with flags (ent_graph_upload = flags 0). */
int ent_graph_upload_ex(const int64_t *beg_pos, const int32_t *adj, int64_t n,
                        int64_t m, int device, int32_t flags, ent_graph **out);
int ent_graph_free(ent_graph *g);

/* Parents of a search from `root` whose depth array is depth_dev (device
 * pointers, n entries each): parent[root] = root; parent[v] = the first
 * neighbour u of v, in device row order, with depth[u] = depth[v] - 1; -1 for
 * unreached vertices. The result is a BFS tree that passes the Graph500
 * parent checks. Returns ENT_EINVAL (parents of the offending vertices set to
 * -2) when some reached vertex has no neighbour one level closer, i.e. the
 * depths are not a BFS from root. Needs a symmetric adjacency (a vertex's
 * neighbours are then its in-neighbours). No reference counterpart: the
 * reference outputs depths only (bfs_gpu_opt.cuh:446). */
int ent_bfs_parents_device(ent_graph *g, int32_t root, const int32_t *depth_dev,
                           int32_t *parent_dev);

/* Every-vertex check of a depth array on the device (depth_dev: n entries),
 * the invariants behind the reference's spot check (validate.h:31-103)
 * applied to all vertices instead of random samples:
 *   counts[0] = 1 if depth[root] != 0;
 *   counts[1] = vertices whose depth is below -1 or at least n;
 *   counts[2] = vertices with a neighbour whose depth differs from theirs by
 *               more than one, or with exactly one of the two unreached;
 *   counts[3] = reached vertices other than root with no neighbour one level
 *               closer.
 * All four are zero exactly when depth is the BFS from root (hop distances
 * are unique) on a symmetric adjacency. Returns ENT_OK whatever the counts. */
int ent_bfs_validate_device(ent_graph *g, int32_t root, const int32_t *depth_dev,
                            int64_t counts[4]);

/* One search from root, then its parents; depth_host and parent_host (n
 * entries, either may be NULL) receive both arrays. */
int ent_bfs_parents(ent_graph *g, int32_t root, int32_t *depth_host,
                    int32_t *parent_host, ent_stats *st);

/* Device time of the last search on g from HIP events on the handle's stream
 * (recorded before its first launch and behind its last queued launch),
 * waiting for those launches to drain. ent_bfs_device returns as soon as the
 * search is finished, and its stats.device_ms comes from the device's
 * constant clock; this is the event-timed figure of the same search. The
 * reference has no device timer: it times each root with host wall time
 * (bfs_gpu_opt.cuh:381, :411). */
int ent_last_device_ms(ent_graph *g, double *ms);

/* One launch of the last search run with count_examined = 1 (the analogue
 * of the reference's ENABLE_MONITORING per-level prints,
 * bfs_gpu_opt.cuh:66-290). phase: 0 top-down, 1 bottom-up, 2 inspection
 * after top-down, 3 top-down from the queue, 4 frontier rebuild. The
 * frontier counts are those of `level` when the launch starts. */
typedef struct ent_level_record {
  int32_t launch;
  int32_t level;
  int32_t phase;
  int32_t pad;
  int64_t frontier_vertices;
  int64_t frontier_edges;
} ent_level_record;

/* Copies up to max records of the last counting search into out and sets
 * *count to the number recorded: one per launch that ran a phase, i.e.
 * stats.phases - 1 (the last launch finds the frontier empty), at most 1024.
 * *count is 0 after a search without counting. */
int ent_level_trace(ent_graph *g, ent_level_record *out, int32_t max, int32_t *count);

/* Copies the device CSR back to host arrays (beg_pos: n + 1 entries, adj: m
 * entries; either may be NULL), in the order the device holds it. */
int ent_graph_download(ent_graph *g, int64_t *beg_pos, int32_t *adj);

/* A second handle on the same resident graph: it shares src's device CSR
 * (no copy; the CSR is freed with the last handle) and has its own stream
 * and search workspace, so searches on the two handles can run concurrently
 * on the device. No reference counterpart (the reference holds one graph and
 * one search per process, comm.h:137-158). */
int ent_graph_share(ent_graph *src, ent_graph **out);
int ent_graph_info(const ent_graph *g, int64_t *n, int64_t *m,
                   int64_t *n_active, int64_t *device_bytes);

int ent_set_config(ent_graph *g, const ent_bfs_config *cfg);
int ent_get_config(const ent_graph *g, ent_bfs_config *cfg);

/* One BFS from root (bfs_tdbu_clfy_sort, bfs_gpu_opt.cuh:40-288). The depth
 * array stays on the device; if depth_host != NULL it is also copied there
 * (n int32, outside the timed search). stats may be NULL. */
int ent_bfs(ent_graph *g, int32_t root, int32_t *depth_host, ent_stats *stats);

/* Same, writing the n-element int32 depth array to device memory depth_dev
 * (NULL = the handle's own buffer; 4-byte alignment suffices) and returning
 * once the search is done: the depths are final and may be read through any
 * stream, while launches that exit at once may still be queued on the
 * handle's stream (later calls on the handle are ordered behind them). */
int ent_bfs_device(ent_graph *g, int32_t root, int32_t *depth_dev,
                   ent_stats *stats);

/* Device pointer of the handle's depth buffer (valid until ent_graph_free). */
int32_t *ent_depth_device_ptr(ent_graph *g);

/* Copies count int32 values from device memory to host memory. */
int ent_copy_to_host(ent_graph *g, const int32_t *src_dev, int32_t *dst_host,
                     int64_t count);

/* Compatibility form of bfs_gpu_coalescing_mem (bfs_gpu_opt.cuh:293-447):
 * uploads the graph, runs one BFS per root in src_list, and prints per-root
 * and average TEPS as the reference does (roots reaching <= 2 levels are
 * skipped, bfs_gpu_opt.cuh:412). Returns 0 or a negative code. */
int ent_bfs_gpu_coalescing_mem(const int32_t *src_list, int32_t nsrc,
                               const int64_t *beg_pos, const int32_t *csr,
                               int64_t n, int64_t m, int32_t gpu_id);

/* Runs a copy kernel that reads and writes `bytes` bytes (reps + 1 times) and
 * reports the mean time per copy. Its hardware byte counters calibrate the
 * HBM traffic figures of the search kernels. No reference counterpart. */
int ent_calib_copy(int device, int64_t bytes, int32_t reps, double *ms_per_copy);

/* Device self-test of the wavefront primitives (DPP inclusive scan, ballot
 * compaction, k-th set bit). Returns 0 when all lanes agree with the host. */
int ent_selftest_wave(int device);

#ifdef __cplusplus
}
#endif

#endif /* ENT_BFS_H */
