/*
 * ent_oracle.c - CPU oracle for the Enterprise BFS (test infrastructure only;
 * never part of the product library).
 *
 * What each function restates from the reference (iHeartGraph/Enterprise):
 *  - BFS semantics: depth[src] = 0 (expander.cuh:64); level L's frontier is
 *    {v : depth[v] == L} (inspector.cuh:84); every unvisited neighbour of the
 *    frontier gets L + 1 (expander.cuh:162-163, bottom-up form :519-523);
 *    the search ends when a level discovers nothing (bfs_gpu_opt.cuh:193-202).
 *    BFS hop distance is unique, so any correct BFS yields the same array.
 *  - validate.h:31-103 (spot check) and its all-vertex generalisation.
 *  - validate.h:106-127 (report: reached vertices and edges).
 */
#include "ent_oracle.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int32_t ent_oracle_bfs(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                       int32_t root, int32_t *depth) {
  return ent_oracle_bfs_parents(n, beg_pos, adj, root, depth, NULL);
}

int32_t ent_oracle_bfs_parents(int64_t n, const int64_t *beg_pos,
                               const int32_t *adj, int32_t root,
                               int32_t *depth, int32_t *parent) {
  if (n <= 0 || root < 0 || root >= n || !beg_pos || !depth) return -1;
  int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!queue) return -1;
  for (int64_t v = 0; v < n; ++v) depth[v] = -1;
  if (parent)
    for (int64_t v = 0; v < n; ++v) parent[v] = -1;
  int64_t head = 0, tail = 0;
  depth[root] = 0;
  if (parent) parent[root] = root;
  queue[tail++] = root;
  int32_t maxd = 0;
  while (head < tail) {
    int32_t u = queue[head++];
    int32_t du = depth[u];
    for (int64_t i = beg_pos[u]; i < beg_pos[u + 1]; ++i) {
      int32_t v = adj[i];
      if (depth[v] < 0) {
        depth[v] = du + 1;
        if (parent) parent[v] = u;
        if (du + 1 > maxd) maxd = du + 1;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
  return maxd + 1;
}

/* Direction-optimizing BFS (Beamer, Asanovic, Patterson, SC'12) on OpenMP
 * threads: the CPU baseline beside the GPU path, and a second CPU BFS the
 * tests compare with ent_oracle_bfs. Top-down levels expand a queue, each
 * thread claiming neighbours with a compare-and-swap on their depth and
 * collecting them in its own buffer (one shared reservation per thread and
 * level); bottom-up levels sweep the unvisited vertices against a frontier
 * bitmap (one bit per vertex), each thread writing whole 64-vertex words of
 * the next bitmap. The queue of a bottom-up level's output is rebuilt, in
 * parallel, only when the next level runs top-down. alpha = 15, beta = 18. */
int32_t ent_oracle_bfs_do(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                          int32_t root, int32_t *depth, int32_t threads) {
  if (n <= 0 || root < 0 || root >= n || !beg_pos || !depth) return -1;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#else
  threads = 1;
#endif
  const int64_t m = beg_pos[n];
  const int64_t nw = (n + 63) / 64;
  int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  uint64_t *fb = (uint64_t *)calloc((size_t)nw, sizeof(uint64_t)); /* frontier bits */
  uint64_t *nb = (uint64_t *)calloc((size_t)nw, sizeof(uint64_t)); /* next frontier bits */
  int64_t *tcount = (int64_t *)calloc((size_t)threads + 1, sizeof(int64_t));
  if (!cur || !nxt || !fb || !nb || !tcount) {
    free(cur);
    free(nxt);
    free(fb);
    free(nb);
    free(tcount);
    return -1;
  }
  int64_t i;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (i = 0; i < n; ++i) depth[i] = -1;
  depth[root] = 0;
  cur[0] = root;
  int64_t ncur = 1;
  int queue_ok = 1; /* cur holds the frontier of `level` */
  int bits_ok = 0;  /* fb holds it */
  int64_t mf = beg_pos[root + 1] - beg_pos[root];
  int64_t mu = m - mf;
  int bottom_up = 0;
  int32_t level = 0;
  const int64_t alpha = 15, beta = 18;
  const int64_t chunk = 64 * 64; /* bottom-up work unit: 64 whole bitmap words */
  while (ncur > 0) {
    if (!bottom_up && mf * alpha > mu) bottom_up = 1;
    else if (bottom_up && ncur * beta < n) bottom_up = 0;
    int64_t nnext = 0, mnext = 0;
    if (bottom_up) {
      if (!bits_ok) { /* frontier queue -> bitmap */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (i = 0; i < nw; ++i) fb[i] = 0;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (i = 0; i < ncur; ++i) __sync_fetch_and_or(&fb[cur[i] >> 6], 1ull << (cur[i] & 63));
      }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : nnext, mnext)
      for (i = 0; i < (n + chunk - 1) / chunk; ++i) {
        const int64_t v0 = i * chunk, v1 = v0 + chunk < n ? v0 + chunk : n;
        for (int64_t w = v0 >> 6; w <= (v1 - 1) >> 6; ++w) {
          uint64_t word = 0;
          const int64_t e = (w + 1) * 64 < v1 ? (w + 1) * 64 : v1;
          for (int64_t v = w * 64; v < e; ++v) {
            if (depth[v] >= 0) continue;
            for (int64_t k = beg_pos[v]; k < beg_pos[v + 1]; ++k) {
              const int32_t u = adj[k];
              if ((fb[u >> 6] >> (u & 63)) & 1ull) {
                depth[v] = level + 1;
                word |= 1ull << (v & 63);
                nnext += 1;
                mnext += beg_pos[v + 1] - beg_pos[v];
                break;
              }
            }
          }
          nb[w] = word;
        }
      }
      uint64_t *t = fb;
      fb = nb;
      nb = t;
      bits_ok = 1;
      queue_ok = 0;
    } else {
      if (!queue_ok) { /* bitmap -> queue: per-thread counts, prefix, fill */
#pragma omp parallel num_threads(threads)
        {
          int tid = 0, nt = 1;
#ifdef _OPENMP
          tid = omp_get_thread_num();
          nt = omp_get_num_threads();
#endif
          const int64_t w0 = nw * tid / nt, w1 = nw * (tid + 1) / nt;
          int64_t c = 0;
          for (int64_t w = w0; w < w1; ++w) c += __builtin_popcountll(fb[w]);
          tcount[tid + 1] = c;
#pragma omp barrier
#pragma omp single
          for (int k = 0; k < nt; ++k) tcount[k + 1] += tcount[k];
          int64_t o = tcount[tid];
          for (int64_t w = w0; w < w1; ++w)
            for (uint64_t x = fb[w]; x; x &= x - 1) cur[o++] = (int32_t)(w * 64 + __builtin_ctzll(x));
        }
        queue_ok = 1;
      }
      int64_t w = 0;
#pragma omp parallel num_threads(threads) reduction(+ : mnext)
      {
        int32_t buf[1024];
        int nbuf = 0;
#pragma omp for schedule(dynamic, 64) nowait
        for (i = 0; i < ncur; ++i) {
          int32_t u = cur[i];
          for (int64_t k = beg_pos[u]; k < beg_pos[u + 1]; ++k) {
            int32_t v = adj[k];
            if (depth[v] < 0 && __sync_bool_compare_and_swap(&depth[v], -1, level + 1)) {
              mnext += beg_pos[v + 1] - beg_pos[v];
              buf[nbuf++] = v;
              if (nbuf == 1024) {
                const int64_t slot = __sync_fetch_and_add(&w, (int64_t)nbuf);
                memcpy(nxt + slot, buf, sizeof(int32_t) * (size_t)nbuf);
                nbuf = 0;
              }
            }
          }
        }
        if (nbuf) {
          const int64_t slot = __sync_fetch_and_add(&w, (int64_t)nbuf);
          memcpy(nxt + slot, buf, sizeof(int32_t) * (size_t)nbuf);
        }
      }
      nnext = w;
      int32_t *t = cur;
      cur = nxt;
      nxt = t;
      queue_ok = 1;
      bits_ok = 0;
    }
    ncur = nnext;
    mu -= mnext;
    mf = mnext;
    ++level;
  }
  free(cur);
  free(nxt);
  free(fb);
  free(nb);
  free(tcount);
  return level;
}

int32_t ent_oracle_validate(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                            int32_t root, const int32_t *depth) {
  if (n <= 0 || root < 0 || root >= n || !beg_pos || !depth) return -8;
  if (depth[root] != 0) return -5;
  for (int64_t v = 0; v < n; ++v)
    if (depth[v] < -1) return -7;
  for (int64_t v = 0; v < n; ++v) {
    const int32_t dv = depth[v];
    int has_parent = (v == root);
    for (int64_t i = beg_pos[v]; i < beg_pos[v + 1]; ++i) {
      const int32_t du = depth[adj[i]];
      if ((dv < 0) != (du < 0)) return -1;
      if (dv < 0) continue;
      if (du > dv + 1 || du < dv - 1) return -1;
      if (du == dv - 1) has_parent = 1;
    }
    if (dv >= 0 && !has_parent) return -6;
  }
  return 0;
}

/* validate.h:57-99 for one base vertex. The reference's depth_t is unsigned
 * char, so INFTY (255) compares as the largest depth; map -1 to INT32_MAX. */
static int32_t spot_depth(const int32_t *depth, int32_t v) {
  return depth[v] < 0 ? 0x7fffffff : depth[v];
}

int32_t ent_oracle_validate_spot(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                                 const int32_t *depth, int32_t base_v) {
  if (n <= 0 || base_v < 0 || base_v >= n || depth[base_v] < 0) return -8;
  const int64_t base = depth[base_v];
  for (int64_t i = beg_pos[base_v]; i < beg_pos[base_v + 1]; ++i) {
    const int32_t c = adj[i];
    const int64_t cd = spot_depth(depth, c);
    if (cd > base + 1 || cd < base - 1) return -1;
    for (int64_t j = beg_pos[c]; j < beg_pos[c + 1]; ++j) {
      const int64_t x = spot_depth(depth, adj[j]);
      if (cd < base && x > base) return -2;
      if (cd > base && x < base) return -3;
      if (cd == base && (x > base + 1 || x < base - 1)) return -4;
    }
  }
  return 0;
}

static int is_neighbour(const int64_t *beg_pos, const int32_t *adj, int32_t v, int32_t u) {
  for (int64_t i = beg_pos[v]; i < beg_pos[v + 1]; ++i)
    if (adj[i] == u) return 1;
  return 0;
}

int32_t ent_oracle_validate_parents(int64_t n, const int64_t *beg_pos, const int32_t *adj,
                                    int32_t root, const int32_t *depth,
                                    const int32_t *parent) {
  if (n <= 0 || root < 0 || root >= n || !depth || !parent) return -8;
  if (parent[root] != root) return -9;
  for (int64_t v = 0; v < n; ++v) {
    if (v == root) continue;
    const int32_t p = parent[v];
    if (depth[v] < 0) {
      if (p != -1) return -12;
      continue;
    }
    if (p < 0 || p >= n) return -12;
    if (!is_neighbour(beg_pos, adj, (int32_t)v, p)) return -10;
    if (depth[p] != depth[v] - 1) return -11;
  }
  return 0;
}

void ent_oracle_report(int64_t n, const int64_t *beg_pos, const int32_t *depth,
                       int64_t *reached_edges, int64_t *reached_vertices) {
  int64_t e = 0, c = 0;
  for (int64_t v = 0; v < n; ++v)
    if (depth[v] >= 0) {
      ++c;
      e += beg_pos[v + 1] - beg_pos[v];
    }
  if (reached_edges) *reached_edges = e;
  if (reached_vertices) *reached_vertices = c;
}

/* ------------------------------------------------------------------------ */
/* Edge list to CSR (the reference's offline converter).                    */

int64_t ent_oracle_tuples_parse(const char *path, int64_t skip_lines, int64_t *tokens, int64_t cap,
                                int64_t *vmin, int64_t *vmax) {
  /* tuple_text_to_bin.cpp:66-147. The whole file is read into memory. */
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return -1;
  }
  const long size = ftell(f);
  rewind(f);
  /* Two NUL bytes past the end: the token scan stops there (the reference
   * reads past the end of a file that does not end in a separator). */
  char *buf = (char *)malloc((size_t)size + 2);
  if (!buf) {
    fclose(f);
    return -1;
  }
  if (size > 0 && fread(buf, 1, (size_t)size, f) != (size_t)size) {
    free(buf);
    fclose(f);
    return -1;
  }
  fclose(f);
  buf[size] = buf[size + 1] = 0;
  /* Line skipping as at :80-97: the byte at the current offset is never
   * examined before the offset advances, so a newline at offset 0 does not
   * end a line. */
  long head = 0;
  int64_t skipped = 0;
  while (skipped < skip_lines) {
    if (head >= size) {
      free(buf);
      return -1; /* :86-91: more lines to skip than the file has */
    }
    head++;
    if (head < size && buf[head] == '\n') {
      head++;
      skipped++;
    }
  }
  const char *ss = buf + head;
  const long len = size - head;
  long next = 0, curr = 0;
  int64_t count = 0, lo = INT64_MAX, hi = 0;
  /* :112-139: one atol per token; ' ', '\n', '\t' separate tokens. */
  while (next < len) {
    const int64_t a = (int64_t)atol(ss + curr);
    if (a > hi) hi = a;
    if (a < lo) lo = a;
    while (next < len && ss[next] != ' ' && ss[next] != '\n' && ss[next] != '\t') next++;
    while (next < len && (ss[next] == ' ' || ss[next] == '\n' || ss[next] == '\t')) next++;
    curr = next;
    if (tokens && count < cap) tokens[count] = a;
    count++;
  }
  free(buf);
  if (vmin) *vmin = lo;
  if (vmax) *vmax = hi;
  return count;
}

typedef struct {
  int32_t v;
  int64_t pos;
  int32_t w;
} row_entry;

static int row_entry_cmp(const void *x, const void *y) {
  const row_entry *a = (const row_entry *)x, *b = (const row_entry *)y;
  if (a->v != b->v) return a->v < b->v ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos ? 1 : 0);
}

int64_t ent_oracle_csr_build(int64_t n, int64_t ne, const int32_t *src, const int32_t *dst,
                             const int32_t *weight, int32_t flags, int64_t *beg_pos, int32_t *adj,
                             int32_t *weight_out) {
  /* tuple_text_to_bin.cpp:190-248: degrees per tuple (both ends when
   * reversed), row starts by prefix sum, then every tuple stored in file
   * order at its row's cursor, (u, v) in row u before (v, u) in row v. A
   * reversed self-loop is stored twice in its row (the reference writes
   * both copies to one slot and leaves the next slot unwritten). */
  const int rev = (flags & 1) != 0, no_loops = (flags & 2) != 0;
  const int dedup = (flags & 8) != 0, sorted = dedup || (flags & 4) != 0;
  int64_t *deg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  if (!deg) return -1;
  for (int64_t i = 0; i < ne; ++i) {
    if (no_loops && src[i] == dst[i]) continue;
    deg[src[i]]++;
    if (rev) deg[dst[i]]++;
  }
  beg_pos[0] = 0;
  for (int64_t v = 0; v < n; ++v) beg_pos[v + 1] = beg_pos[v] + deg[v];
  const int64_t m0 = beg_pos[n];
  row_entry *ent = (row_entry *)malloc(sizeof(row_entry) * (size_t)(m0 > 0 ? m0 : 1));
  if (!ent) {
    free(deg);
    return -1;
  }
  for (int64_t v = 0; v < n; ++v) deg[v] = beg_pos[v];
  int64_t pos = 0;
  for (int64_t i = 0; i < ne; ++i) {
    const int32_t u = src[i], v = dst[i];
    if (no_loops && u == v) continue;
    const int32_t w = weight ? weight[i] : 0;
    row_entry a = {v, pos++, w};
    ent[deg[u]++] = a;
    if (rev) {
      row_entry b = {u, pos++, w};
      ent[deg[v]++] = b;
    }
  }
  /* Options beyond the reference: rows sorted by neighbour (ties in file
   * order) and, with dedup, the first of each run of equal neighbours kept. */
  int64_t m = 0;
  for (int64_t v = 0; v < n; ++v) {
    const int64_t s = beg_pos[v], e = beg_pos[v + 1];
    if (sorted && e - s > 1) qsort(ent + s, (size_t)(e - s), sizeof(row_entry), row_entry_cmp);
    beg_pos[v] = m;
    for (int64_t p = s; p < e; ++p) {
      if (dedup && p > s && ent[p].v == ent[p - 1].v) continue;
      adj[m] = ent[p].v;
      if (weight_out) weight_out[m] = ent[p].w;
      m++;
    }
  }
  beg_pos[n] = m;
  free(ent);
  free(deg);
  return m;
}
