/*
 * ent_batch.h - multi-source BFS on a resident graph (no counterpart in the
 * reference, which runs one root at a time, bfs_gpu_opt.cuh:370; the form
 * follows the iBFS idea the reference cites at readme.md:79).
 *
 * Up to 512 roots are searched together: each vertex carries W = ceil(k/64)
 * u64 masks (bit s = "source s"), so one adjacency read serves every source
 * whose frontier touches it. The result is k int32 depth rows, row-major
 * (depths[s * n + v]), each identical to what ent_bfs gives for roots[s].
 *
 * Memory: besides the graph, a workspace of (max_levels + 3) * (n * W * 8 +
 * ceil(n / 64) * 8) bytes holds the per-level frontier masks and joint
 * frontier bitmaps of the levels not yet written to the output (levels are
 * replayed in a pass that writes the output rows in full, coalesced lines).
 * The search itself has no depth limit: when it would need more levels than
 * are kept, a flush pass writes the kept levels to the rows and their slots
 * are reused (ent_batch_stats.flushes), so a high-diameter graph runs as one
 * batch in the same workspace.
 */
#ifndef ENT_BATCH_H
#define ENT_BATCH_H

#include <stdint.h>

#include "ent_bfs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ENT_BATCH_MAX_SOURCES 512

typedef struct ent_batch_stats {
  int32_t k;                /* sources */
  int32_t words;            /* u64 masks per vertex in this search: k / 64
                               rounded up to 1, 2, 4 or 8, whatever width the
                               workspace was reserved for */
  int32_t levels;           /* joint levels (max over sources of depth + 1) */
  int32_t phases;           /* kernel phases run */
  int64_t reached_edges;    /* sum over sources of M_c(s) */
  int64_t reached_vertices; /* sum over sources of N_c(s) */
  int64_t alg_bytes;        /* per 64-source group g: sum over levels l and
                               v in J_l(g) of (8 + 12 deg v), plus 4*64*n;
                               summed over groups */
  double device_ms;         /* HIP-event time of the batch */
  double host_ms;           /* wall time of the call */
  uint64_t bu_level_mask;   /* bit l: level l pulled (bottom-up) some of
                               its 64-source words */
  uint64_t td_level_mask;   /* bit l: level l pushed (top-down) some of its
                               words; both bits set: the level split its words
                               by frontier size */
  int32_t flushes;          /* passes that wrote the kept levels to the rows
                               before the search ended (one per max_levels + 1
                               levels beyond the first max_levels) */
  int32_t pad;
} ent_batch_stats;

/* Reserves the workspace for up to k sources, keeping max_levels joint
 * levels between flushes (1..250; 0 = keep the current reservation or the
 * default of 32 levels). A search with fewer sources runs in it at its own
 * mask width. */
int ent_batch_reserve(ent_graph *g, int32_t k, int32_t max_levels);

/* Frees the multi-source workspace (the graph stays resident). */
int ent_batch_release(ent_graph *g);

/* Device bytes the workspace would take for k sources and max_levels. */
int64_t ent_batch_workspace_bytes(const ent_graph *g, int32_t k,
                                  int32_t max_levels);

/* Multi-source BFS from roots[0..k-1] (1 <= k <= 512, roots may repeat),
 * writing k * n int32 depths to device memory depths_dev (4-byte aligned;
 * rows go out in 16-byte stores when it is 16-byte aligned and n % 4 == 0).
 * Row i is root i's depth array. Internally the sources are grouped into
 * 64-source words by root degree, so the time and the counts do not depend
 * on the order of roots. */
int ent_bfs_batch_device(ent_graph *g, const int32_t *roots, int32_t k,
                         int32_t *depths_dev, ent_batch_stats *stats);

/* Same, copying the k * n depths to host memory depths_host afterwards
 * (outside the timed search). The device rows are kept with the handle's
 * multi-source workspace and reused by later calls. */
int ent_bfs_batch(ent_graph *g, const int32_t *roots, int32_t k,
                  int32_t *depths_host, ent_batch_stats *stats);

/* Allocates / frees device memory through the library (for callers without
 * a device allocator of their own, e.g. ctypes). */
int ent_device_alloc(ent_graph *g, int64_t bytes, void **out);
int ent_device_free(ent_graph *g, void *p);

/* Position-keyed 64-bit checksum of each of `rows` rows of `cols` int32
 * values in device memory, computed on the device (for comparing large depth
 * arrays without copying them): sum over i, mod 2^64, of the splitmix64
 * finaliser of ((i << 32) ^ value_i) + 0x9E3779B97F4A7C15. Uses a buffer of
 * the handle (no allocation per call). */
int ent_device_row_hash(ent_graph *g, const int32_t *dev, int64_t rows,
                        int64_t cols, uint64_t *hashes_host);

/* Rows one ent_device_row_hash_noalloc call can sum. */
#define ENT_NOALLOC_HASH_ROWS 4096

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the checksums of ent_device_row_hash, for at most
 * ENT_NOALLOC_HASH_ROWS rows, with no graph handle and no allocation of any
 * kind. The per-row sums live in a fixed device array of
 * ENT_NOALLOC_HASH_ROWS * 8 bytes (32 KiB) that each GPU holds from the
 * first call on it until the process ends; every call on a GPU shares it,
 * so calls on one device from one process are serialized (each clears the
 * sums it uses). Runs on `device`'s null stream and returns once
 * hashes_host is written; the rows in `dev` must be complete (work queued
 * on other streams is not waited for). More rows than the maximum:
 * ENT_EINVAL. */
int ent_device_row_hash_noalloc(int32_t device, const int32_t *dev,
                                int64_t rows, int64_t cols,
                                uint64_t *hashes_host);

#ifdef __cplusplus
}
#endif

#endif /* ENT_BATCH_H */
