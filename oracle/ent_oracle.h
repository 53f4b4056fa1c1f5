/*
 * ent_oracle.h - CPU oracle for the Enterprise direction-optimizing BFS.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker the GPU library is
 * compared against; it is never linked into, loaded by, or called from the
 * product library. Only tests, the smoke check and the CPU-baseline leg of
 * the benchmark use it.
 *
 * Conventions (match the product library):
 *   CSR: int64 beg_pos[n+1], int32 adj[m]; depth: int32[n], unreached = -1.
 *   The reference stores depth as unsigned char with unreached = 255
 *   (comm.h:69, comm.h:131); build depth d in [0,254] <-> reference d, and
 *   build -1 <-> reference 255.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   ent_oracle_bfs            scipy.sparse.csgraph BFS on random graphs, closed
 *                             forms (path, cycle, star, complete graph,
 *                             hypercube, binary tree, grid, disjoint union) and
 *                             committed golden vectors
 *   ent_oracle_validate       accepts every oracle output above; rejects
 *                             planted mistakes (depth +-1, swapped vertices,
 *                             wrong root, unreached marked reached)
 *   ent_oracle_validate_spot  restates validate.h:31-103 verbatim in C;
 *                             checked on the same accept/reject cases
 *   ent_oracle_report         closed form: sum of degrees over reached vertices
 *   ent_oracle_bfs_parents /
 *   ent_oracle_validate_parents  parents checked against depth (Graph500
 *                             rules) and planted wrong parents rejected
 *   ent_oracle_bfs_do         equal to ent_oracle_bfs on every test graph
 *   ent_oracle_tuples_parse /
 *   ent_oracle_csr_build      the reference converter itself, compiled from
 *                             its source into oracle/_ref (tests/test_convert.py)
 *                             and its committed outputs (tests/golden/convert);
 *                             sorted and deduplicated rows against scipy's
 *                             COO to CSR conversion
 */
#ifndef ENT_ORACLE_H
#define ENT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Serial queue BFS; returns the number of levels (max depth + 1) or -1 on a
 * bad argument. */
int32_t ent_oracle_bfs(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                       int32_t root, int32_t *depth);

/* Serial queue BFS that also records parents (parent[root] = root, -1 when
 * unreached); the parent is the first discoverer in queue order. */
int32_t ent_oracle_bfs_parents(int64_t n, const int64_t *beg_pos,
                               const int32_t *adj, int32_t root,
                               int32_t *depth, int32_t *parent);

/* Direction-optimizing BFS on the host (Beamer's alpha/beta switch) with
 * `threads` OpenMP threads; the CPU baseline. */
int32_t ent_oracle_bfs_do(int64_t n, const int64_t *beg_pos,
                          const int32_t *adj, int32_t root, int32_t *depth,
                          int32_t threads);

/* All-vertex form of validate.h:64-99. Returns 0 when depth is a valid BFS
 * level assignment from root, else:
 *  -1 an edge joins depths differing by more than one, or joins a reached and
 *     an unreached vertex (validate.h:68-70 with INFTY = unreached)
 *  -5 depth[root] != 0
 *  -6 a reached vertex other than root has no neighbour one level up
 *  -7 a depth value below -1
 *  -8 bad argument */
int32_t ent_oracle_validate(int64_t n, const int64_t *beg_pos,
                            const int32_t *adj, int32_t root,
                            const int32_t *depth);

/* The reference's spot check restated for one chosen reached vertex base_v
 * (validate.h:57-99): returns 0 or -1..-4 exactly as the reference does, with
 * unreached treated as the largest depth value. */
int32_t ent_oracle_validate_spot(int64_t n, const int64_t *beg_pos,
                                 const int32_t *adj, const int32_t *depth,
                                 int32_t base_v);

/* Graph500 parent rules: parent[root] = root; for reached v != root, parent[v]
 * is a neighbour of v with depth[v] - 1; unreached v has parent -1.
 * Returns 0, or -9 (root), -10 (not a neighbour), -11 (wrong level),
 * -12 (unreached with a parent / reached without one). */
int32_t ent_oracle_validate_parents(int64_t n, const int64_t *beg_pos,
                                    const int32_t *adj, int32_t root,
                                    const int32_t *depth,
                                    const int32_t *parent);

/* validate.h:106-127: reached vertex count and the sum of their degrees (the
 * reference's TEPS numerator). */
void ent_oracle_report(int64_t n, const int64_t *beg_pos, const int32_t *depth,
                       int64_t *reached_edges, int64_t *reached_vertices);

/* The reference converter's text pass (tuple_text_to_bin.cpp:66-147): skips
 * skip_lines lines the way it does, then reads tokens separated by ' ',
 * '\t' and '\n' with atol. Returns the token count and stores up to cap
 * values in tokens (may be NULL) and the extremes in *vmin, *vmax; -1 when
 * the file cannot be read or has fewer lines than skip_lines. */
int64_t ent_oracle_tuples_parse(const char *path, int64_t skip_lines,
                                int64_t *tokens, int64_t cap, int64_t *vmin,
                                int64_t *vmax);

/* CSR of ne tuples (src[i], dst[i]) over n vertices, the reference
 * converter's construction (tuple_text_to_bin.cpp:190-248) plus options.
 * flags: 1 also store each tuple reversed, right after it; 2 drop (u, u);
 * 4 sort each row by neighbour (ties in tuple order); 8 drop repeated
 * neighbours of a row, keeping the first (implies 4). weight (NULL or ne
 * entries) follows its tuple into weight_out (NULL or m entries).
 * beg_pos: n + 1 entries; adj: capacity ne * 2. Returns m, or -1 when out
 * of memory. */
int64_t ent_oracle_csr_build(int64_t n, int64_t ne, const int32_t *src,
                             const int32_t *dst, const int32_t *weight,
                             int32_t flags, int64_t *beg_pos, int32_t *adj,
                             int32_t *weight_out);

#ifdef __cplusplus
}
#endif

#endif /* ENT_ORACLE_H */
