// Graph substrate: R-MAT and grid generators, CSR construction, the
// reference's binary CSR file format, and root pickers.
//
// The reference has no generator; it reads CSR files written by
// tuple_text_to_bin.cpp (int64 _beg_pos.bin / _csr.bin, :153-184) and
// narrows them to int32 (graph.hpp:60-66, :80-87). Its root picker is
// main.cu:55-78.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "ent_common.h"
#include "ent_convert.h"

namespace ent {

static thread_local std::string g_err;
static std::atomic<int> g_threads{0};

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void clear_error() { g_err.clear(); }

int host_threads() {
  int n = g_threads.load();
  if (n > 0) return n;
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  return hc > 16 ? 16 : int(hc);
}

namespace {

template <class T>
T *alloc_array(int64_t count) {
  if (count <= 0) count = 1;
  return static_cast<T *>(std::malloc(sizeof(T) * size_t(count)));
}

// Exclusive prefix sum of counts[0..n) into off[0..n], off[n] = total.
void exclusive_scan(const int64_t *counts, int64_t n, int64_t *off) {
  const int nt = host_threads();
  if (n < (1 << 20) || nt <= 1) {
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) {
      off[i] = s;
      s += counts[i];
    }
    off[n] = s;
    return;
  }
  const int64_t chunk = (n + nt - 1) / nt;
  std::vector<int64_t> part(nt + 1, 0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      int64_t b = t * chunk, e = std::min(n, b + chunk), s = 0;
      for (int64_t i = b; i < e; ++i) s += counts[i];
      part[t + 1] = s;
    });
  for (auto &th : pool) th.join();
  pool.clear();
  for (int t = 0; t < nt; ++t) part[t + 1] += part[t];
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      int64_t b = t * chunk, e = std::min(n, b + chunk), s = part[t];
      for (int64_t i = b; i < e; ++i) {
        off[i] = s;
        s += counts[i];
      }
    });
  for (auto &th : pool) th.join();
  off[n] = part[nt];
}

// Turns an unsorted adjacency bucket array (rows [start[v], start[v+1]) of
// `raw`) into a sorted, duplicate-free CSR in `out`.
int finish_csr(int64_t n, const int64_t *start, int32_t *raw, ent_csr *out) {
  int64_t *cnt = alloc_array<int64_t>(n);
  if (!cnt) return set_error(ENT_ENOMEM, "out of host memory (row counts)");
  parallel_for(n, 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int32_t *row = raw + start[v];
      int64_t len = start[v + 1] - start[v];
      std::sort(row, row + len);
      int64_t w = 0;
      for (int64_t i = 0; i < len; ++i)
        if (row[i] != v && (w == 0 || row[w - 1] != row[i])) row[w++] = row[i];
      cnt[v] = w;
    }
  });
  int64_t *off = alloc_array<int64_t>(n + 1);
  if (!off) {
    std::free(cnt);
    return set_error(ENT_ENOMEM, "out of host memory (offsets)");
  }
  exclusive_scan(cnt, n, off);
  const int64_t m = off[n];
  int32_t *adj = alloc_array<int32_t>(m);
  if (!adj) {
    std::free(cnt);
    std::free(off);
    return set_error(ENT_ENOMEM, "out of host memory (%lld edges)", (long long)m);
  }
  parallel_for(n, 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      std::memcpy(adj + off[v], raw + start[v], sizeof(int32_t) * size_t(cnt[v]));
  });
  std::free(cnt);
  out->n = n;
  out->m = m;
  out->beg_pos = off;
  out->adj = adj;
  return ENT_OK;
}

// R-MAT tuple i (Graph500 Kronecker recursion without noise) from a
// counter-based RNG, so tuples do not depend on the thread count; ids
// optionally permuted by a seeded Fisher-Yates shuffle.
struct Rmat {
  int32_t scale;
  uint64_t seed, ta, tb, tc;
  std::vector<int32_t> perm;
  Rmat(int32_t scale_, double a, double b, double c, uint64_t seed_, int32_t relabel) : scale(scale_), seed(seed_) {
    // Integer thresholds on a 32-bit uniform draw: quadrant 0 (a), 1 (b:
    // column bit set), 2 (c: row bit set), 3 (d: both bits set).
    ta = uint64_t(std::ldexp(a, 32));
    tb = uint64_t(std::ldexp(a + b, 32));
    tc = uint64_t(std::ldexp(a + b + c, 32));
    if (relabel) {
      const int64_t n = int64_t(1) << scale;
      perm.resize(size_t(n));
      std::iota(perm.begin(), perm.end(), 0);
      uint64_t st = mix64(seed ^ 0xA5A5A5A55A5A5A5Aull);
      for (int64_t i = n - 1; i > 0; --i) {
        uint64_t j = splitmix_next(st) % uint64_t(i + 1);
        std::swap(perm[size_t(i)], perm[size_t(j)]);
      }
    }
  }
  void tuple(int64_t i, int64_t &u, int64_t &v) const {
    uint64_t st = hash2(seed, uint64_t(i));
    int64_t r = 0, col = 0;
    for (int32_t l = 0; l < scale; ++l) {
      uint64_t x = splitmix_next(st) >> 32;
      int rb = 0, cb = 0;
      if (x < ta) {
      } else if (x < tb) {
        cb = 1;
      } else if (x < tc) {
        rb = 1;
      } else {
        rb = cb = 1;
      }
      r = (r << 1) | rb;
      col = (col << 1) | cb;
    }
    u = perm.empty() ? r : perm[size_t(r)];
    v = perm.empty() ? col : perm[size_t(col)];
  }
};

// Bucket fill shared by R-MAT and edge-list construction: edge(i, &u, &v)
// returns false for tuples to skip. Each kept tuple is stored both ways.
template <class EdgeFn>
int build_symmetric(int64_t n, int64_t ne, EdgeFn edge, ent_csr *out) {
  std::atomic<int64_t> *deg = static_cast<std::atomic<int64_t> *>(
      std::malloc(sizeof(std::atomic<int64_t>) * size_t(n + 1)));
  if (!deg) return set_error(ENT_ENOMEM, "out of host memory (degrees)");
  parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) deg[v].store(0, std::memory_order_relaxed);
  });
  parallel_for(ne, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int64_t u, v;
      if (!edge(i, u, v)) continue;
      deg[u].fetch_add(1, std::memory_order_relaxed);
      deg[v].fetch_add(1, std::memory_order_relaxed);
    }
  });
  int64_t *start = alloc_array<int64_t>(n + 1);
  if (!start) {
    std::free(deg);
    return set_error(ENT_ENOMEM, "out of host memory (bucket starts)");
  }
  {
    int64_t s = 0;
    for (int64_t v = 0; v < n; ++v) {
      start[v] = s;
      s += deg[v].load(std::memory_order_relaxed);
    }
    start[n] = s;
  }
  const int64_t total = start[n];
  int32_t *raw = alloc_array<int32_t>(total);
  if (!raw) {
    std::free(deg);
    std::free(start);
    return set_error(ENT_ENOMEM, "out of host memory (%lld raw edges)", (long long)total);
  }
  // Reuse deg as the per-row cursor.
  parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) deg[v].store(start[v], std::memory_order_relaxed);
  });
  parallel_for(ne, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int64_t u, v;
      if (!edge(i, u, v)) continue;
      raw[deg[u].fetch_add(1, std::memory_order_relaxed)] = int32_t(v);
      raw[deg[v].fetch_add(1, std::memory_order_relaxed)] = int32_t(u);
    }
  });
  std::free(deg);
  int rc = finish_csr(n, start, raw, out);
  std::free(start);
  std::free(raw);
  return rc;
}

}  // namespace
}  // namespace ent

using namespace ent;

extern "C" {

const char *ent_last_error(void) { return g_err.c_str(); }

void ent_set_host_threads(int32_t n) { g_threads.store(n > 0 ? n : 0); }

void ent_csr_free(ent_csr *g) {
  if (!g) return;
  std::free(g->beg_pos);
  std::free(g->adj);
  g->beg_pos = nullptr;
  g->adj = nullptr;
  g->n = g->m = 0;
}

int ent_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                 uint64_t seed, int32_t relabel, ent_csr *out) {
  clear_error();
  if (!out) return set_error(ENT_EINVAL, "out is NULL");
  if (scale < 1 || scale > 30) return set_error(ENT_EINVAL, "scale %d outside [1, 30]", scale);
  if (edgefactor < 1 || edgefactor > 1024)
    return set_error(ENT_EINVAL, "edgefactor %d outside [1, 1024]", edgefactor);
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return set_error(ENT_EINVAL, "bad R-MAT probabilities a=%g b=%g c=%g", a, b, c);
  const int64_t n = int64_t(1) << scale;
  const int64_t ne = n * edgefactor;
  const Rmat gen(scale, a, b, c, seed, relabel);
  auto edge = [&](int64_t i, int64_t &u, int64_t &v) -> bool {
    gen.tuple(i, u, v);
    return u != v;
  };
  return build_symmetric(n, ne, edge, out);
}

int ent_gen_rmat_tuples(int32_t scale, int32_t edgefactor, double a, double b, double c, uint64_t seed,
                        int32_t relabel, int32_t *src, int32_t *dst) {
  clear_error();
  if (!src || !dst) return set_error(ENT_EINVAL, "NULL output");
  if (scale < 1 || scale > 30) return set_error(ENT_EINVAL, "scale %d outside [1, 30]", scale);
  if (edgefactor < 1 || edgefactor > 1024)
    return set_error(ENT_EINVAL, "edgefactor %d outside [1, 1024]", edgefactor);
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return set_error(ENT_EINVAL, "bad R-MAT probabilities a=%g b=%g c=%g", a, b, c);
  const int64_t ne = (int64_t(1) << scale) * edgefactor;
  const Rmat gen(scale, a, b, c, seed, relabel);
  parallel_for(ne, 1 << 16, [&](int64_t bgn, int64_t end) {
    for (int64_t i = bgn; i < end; ++i) {
      int64_t u, v;
      gen.tuple(i, u, v);
      src[i] = int32_t(u);
      dst[i] = int32_t(v);
    }
  });
  return ENT_OK;
}

int ent_csr_from_edges(int64_t n, int64_t ne, const int32_t *src, const int32_t *dst,
                       ent_csr *out) {
  clear_error();
  if (!out || (ne > 0 && (!src || !dst))) return set_error(ENT_EINVAL, "NULL argument");
  if (n < 0 || n > INT32_MAX || ne < 0) return set_error(ENT_EINVAL, "bad sizes n=%lld ne=%lld", (long long)n, (long long)ne);
  for (int64_t i = 0; i < ne; ++i)
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n)
      return set_error(ENT_EINVAL, "edge %lld (%d,%d) out of range for n=%lld", (long long)i,
                       src[i], dst[i], (long long)n);
  auto edge = [&](int64_t i, int64_t &u, int64_t &v) -> bool {
    u = src[i];
    v = dst[i];
    return u != v;
  };
  return build_symmetric(n, ne, edge, out);
}

int ent_csr_from_arrays(int64_t n, int64_t m, const int64_t *beg_pos, const int32_t *adj,
                        ent_csr *out) {
  clear_error();
  if (!out || !beg_pos || (m > 0 && !adj)) return set_error(ENT_EINVAL, "NULL argument");
  if (n < 0 || n > INT32_MAX || m < 0) return set_error(ENT_EINVAL, "bad sizes");
  out->beg_pos = alloc_array<int64_t>(n + 1);
  out->adj = alloc_array<int32_t>(m);
  if (!out->beg_pos || !out->adj) {
    ent_csr_free(out);
    return set_error(ENT_ENOMEM, "out of host memory");
  }
  std::memcpy(out->beg_pos, beg_pos, sizeof(int64_t) * size_t(n + 1));
  if (m) std::memcpy(out->adj, adj, sizeof(int32_t) * size_t(m));
  out->n = n;
  out->m = m;
  return ENT_OK;
}

int ent_csr_check(const ent_csr *g, int32_t flags) {
  clear_error();
  if (!g || !g->beg_pos || (g->m > 0 && !g->adj)) return set_error(ENT_EINVAL, "NULL CSR");
  const int64_t n = g->n, m = g->m;
  if (n < 0 || n > INT32_MAX) return set_error(ENT_EINVAL, "n=%lld out of range", (long long)n);
  if (g->beg_pos[0] != 0) return set_error(ENT_EINVAL, "beg_pos[0]=%lld != 0", (long long)g->beg_pos[0]);
  if (g->beg_pos[n] != m)
    return set_error(ENT_EINVAL, "beg_pos[n]=%lld != m=%lld", (long long)g->beg_pos[n], (long long)m);
  for (int64_t v = 0; v < n; ++v)
    if (g->beg_pos[v + 1] < g->beg_pos[v])
      return set_error(ENT_EINVAL, "beg_pos decreases at vertex %lld", (long long)v);
  std::atomic<int64_t> bad{-1};
  std::atomic<int> why{0};
  parallel_for(n, 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e && bad.load(std::memory_order_relaxed) < 0; ++v) {
      for (int64_t i = g->beg_pos[v]; i < g->beg_pos[v + 1]; ++i) {
        int32_t u = g->adj[i];
        int w = 0;
        if (u < 0 || u >= n) w = 1;
        else if ((flags & 1) && i > g->beg_pos[v] && g->adj[i - 1] >= u) w = 2;
        else if ((flags & 2) && u == v) w = 3;
        else if (flags & 4) {
          const int32_t *rb = g->adj + g->beg_pos[u], *re = g->adj + g->beg_pos[u + 1];
          bool found = (flags & 1) ? std::binary_search(rb, re, int32_t(v))
                                   : std::find(rb, re, int32_t(v)) != re;
          if (!found) w = 4;
        }
        if (w) {
          int64_t expect = -1;
          if (bad.compare_exchange_strong(expect, v)) why.store(w);
          return;
        }
      }
    }
  });
  if (bad.load() >= 0) {
    static const char *what[] = {"", "neighbour id out of range", "row not sorted/duplicate",
                                 "self loop", "edge without its reverse"};
    return set_error(ENT_EINVAL, "CSR check failed at vertex %lld: %s", (long long)bad.load(),
                     what[why.load()]);
  }
  return ENT_OK;
}

int ent_gen_grid(int32_t width, int32_t height, int32_t delete_ppm, uint64_t seed,
                 int32_t keep_lcc, ent_csr *out) {
  clear_error();
  if (!out) return set_error(ENT_EINVAL, "out is NULL");
  if (width < 1 || height < 1 || int64_t(width) * height > INT32_MAX)
    return set_error(ENT_EINVAL, "bad grid %d x %d", width, height);
  if (delete_ppm < 0 || delete_ppm > 1000000) return set_error(ENT_EINVAL, "delete_ppm out of range");
  const int64_t W = width, H = height, n = W * H;
  const int64_t nh = (W - 1) * H;  // horizontal edge ids [0, nh), id = y*(W-1)+x
  auto kept = [&](int64_t eid) { return int64_t(hash2(seed, uint64_t(eid)) % 1000000u) >= delete_ppm; };
  auto right = [&](int64_t x, int64_t y) { return x + 1 < W && kept(y * (W - 1) + x); };
  auto down = [&](int64_t x, int64_t y) { return y + 1 < H && kept(nh + y * W + x); };
  // Keep-mask of vertices (all, or the largest component).
  std::vector<int32_t> newid(size_t(n), 0);
  int64_t nkeep = n;
  if (keep_lcc) {
    std::vector<int32_t> comp(size_t(n), -1);
    std::vector<int32_t> stack;
    int32_t ncomp = 0, best = -1;
    int64_t best_sz = 0;
    for (int64_t s = 0; s < n; ++s) {
      if (comp[size_t(s)] >= 0) continue;
      int64_t sz = 0;
      comp[size_t(s)] = ncomp;
      stack.push_back(int32_t(s));
      while (!stack.empty()) {
        int64_t v = stack.back();
        stack.pop_back();
        ++sz;
        int64_t x = v % W, y = v / W;
        int64_t nb[4];
        int k = 0;
        if (y > 0 && down(x, y - 1)) nb[k++] = v - W;
        if (x > 0 && right(x - 1, y)) nb[k++] = v - 1;
        if (right(x, y)) nb[k++] = v + 1;
        if (down(x, y)) nb[k++] = v + W;
        for (int i = 0; i < k; ++i)
          if (comp[size_t(nb[i])] < 0) {
            comp[size_t(nb[i])] = ncomp;
            stack.push_back(int32_t(nb[i]));
          }
      }
      if (sz > best_sz) {
        best_sz = sz;
        best = ncomp;
      }
      ++ncomp;
    }
    nkeep = 0;
    for (int64_t v = 0; v < n; ++v) newid[size_t(v)] = comp[size_t(v)] == best ? int32_t(nkeep++) : -1;
  } else {
    for (int64_t v = 0; v < n; ++v) newid[size_t(v)] = int32_t(v);
  }
  int64_t *off = alloc_array<int64_t>(nkeep + 1);
  if (!off) return set_error(ENT_ENOMEM, "out of host memory");
  // Row v's neighbours in increasing grid id: up, left, right, down. The
  // renumbering is monotone, so rows stay sorted.
  auto row = [&](int64_t v, int32_t *dst) -> int {
    int64_t x = v % W, y = v / W;
    int k = 0;
    if (y > 0 && down(x, y - 1)) { if (dst) dst[k] = newid[size_t(v - W)]; ++k; }
    if (x > 0 && right(x - 1, y)) { if (dst) dst[k] = newid[size_t(v - 1)]; ++k; }
    if (right(x, y)) { if (dst) dst[k] = newid[size_t(v + 1)]; ++k; }
    if (down(x, y)) { if (dst) dst[k] = newid[size_t(v + W)]; ++k; }
    return k;
  };
  std::vector<int32_t> kept_ids;
  kept_ids.reserve(size_t(nkeep));
  for (int64_t v = 0; v < n; ++v)
    if (newid[size_t(v)] >= 0) kept_ids.push_back(int32_t(v));
  std::vector<int64_t> cnt(static_cast<size_t>(nkeep));
  parallel_for(nkeep, 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) cnt[size_t(i)] = row(kept_ids[size_t(i)], nullptr);
  });
  exclusive_scan(cnt.data(), nkeep, off);
  int32_t *adj = alloc_array<int32_t>(off[nkeep]);
  if (!adj) {
    std::free(off);
    return set_error(ENT_ENOMEM, "out of host memory");
  }
  parallel_for(nkeep, 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) row(kept_ids[size_t(i)], adj + off[i]);
  });
  out->n = nkeep;
  out->m = off[nkeep];
  out->beg_pos = off;
  out->adj = adj;
  return ENT_OK;
}

int ent_csr_read_bin(const char *beg_file, const char *csr_file, ent_csr *out) {
  clear_error();
  if (!beg_file || !csr_file || !out) return set_error(ENT_EINVAL, "NULL argument");
  FILE *fb = std::fopen(beg_file, "rb");
  if (!fb) return set_error(ENT_EIO, "cannot open %s: %s", beg_file, std::strerror(errno));
  std::fseek(fb, 0, SEEK_END);
  long bsz = std::ftell(fb);
  std::fseek(fb, 0, SEEK_SET);
  if (bsz < 8 || bsz % 8) {
    std::fclose(fb);
    return set_error(ENT_EIO, "%s: size %ld is not a positive multiple of 8", beg_file, bsz);
  }
  const int64_t n = bsz / 8 - 1;
  if (n > INT32_MAX) {
    std::fclose(fb);
    return set_error(ENT_EINVAL, "%s: %lld vertices exceed int32 ids", beg_file, (long long)n);
  }
  int64_t *off = alloc_array<int64_t>(n + 1);
  if (!off) {
    std::fclose(fb);
    return set_error(ENT_ENOMEM, "out of host memory");
  }
  size_t got = std::fread(off, 8, size_t(n + 1), fb);
  std::fclose(fb);
  if (got != size_t(n + 1)) {
    std::free(off);
    return set_error(ENT_EIO, "%s: short read", beg_file);
  }
  FILE *fc = std::fopen(csr_file, "rb");
  if (!fc) {
    std::free(off);
    return set_error(ENT_EIO, "cannot open %s: %s", csr_file, std::strerror(errno));
  }
  std::fseek(fc, 0, SEEK_END);
  long csz = std::ftell(fc);
  std::fseek(fc, 0, SEEK_SET);
  if (csz < 0 || csz % 8) {
    std::fclose(fc);
    std::free(off);
    return set_error(ENT_EIO, "%s: size %ld is not a multiple of 8", csr_file, csz);
  }
  // The edge count is the file size, as in graph.hpp:46; beg_pos must agree.
  const int64_t m = csz / 8;
  if (off[0] != 0 || off[n] != m) {
    std::fclose(fc);
    std::free(off);
    return set_error(ENT_EINVAL, "offsets end at %lld but %s holds %lld ids", (long long)off[n],
                     csr_file, (long long)m);
  }
  int32_t *adj = alloc_array<int32_t>(m);
  if (!adj) {
    std::fclose(fc);
    std::free(off);
    return set_error(ENT_ENOMEM, "out of host memory");
  }
  const size_t blk = size_t(1) << 20;
  std::vector<int64_t> buf(blk);
  int64_t done = 0;
  int rc = ENT_OK;
  while (done < m) {
    size_t want = size_t(std::min<int64_t>(int64_t(blk), m - done));
    if (std::fread(buf.data(), 8, want, fc) != want) {
      rc = set_error(ENT_EIO, "%s: short read", csr_file);
      break;
    }
    for (size_t i = 0; i < want; ++i) {
      int64_t x = buf[i];
      if (x < 0 || x >= n) {
        rc = set_error(ENT_EINVAL, "%s: id %lld at %lld out of range", csr_file, (long long)x,
                       (long long)(done + int64_t(i)));
        break;
      }
      adj[done + int64_t(i)] = int32_t(x);
    }
    if (rc) break;
    done += int64_t(want);
  }
  std::fclose(fc);
  if (rc) {
    std::free(off);
    std::free(adj);
    return rc;
  }
  out->n = n;
  out->m = m;
  out->beg_pos = off;
  out->adj = adj;
  return ENT_OK;
}

int ent_csr_write_bin(const ent_csr *g, const char *beg_file, const char *csr_file) {
  clear_error();
  if (!g || !g->beg_pos || !beg_file || !csr_file) return set_error(ENT_EINVAL, "NULL argument");
  FILE *fb = std::fopen(beg_file, "wb");
  if (!fb) return set_error(ENT_EIO, "cannot create %s: %s", beg_file, std::strerror(errno));
  size_t w = std::fwrite(g->beg_pos, 8, size_t(g->n + 1), fb);
  std::fclose(fb);
  if (w != size_t(g->n + 1)) return set_error(ENT_EIO, "%s: short write", beg_file);
  FILE *fc = std::fopen(csr_file, "wb");
  if (!fc) return set_error(ENT_EIO, "cannot create %s: %s", csr_file, std::strerror(errno));
  const size_t blk = size_t(1) << 20;
  std::vector<int64_t> buf(blk);
  for (int64_t done = 0; done < g->m;) {
    size_t k = size_t(std::min<int64_t>(int64_t(blk), g->m - done));
    for (size_t i = 0; i < k; ++i) buf[i] = g->adj[done + int64_t(i)];
    if (std::fwrite(buf.data(), 8, k, fc) != k) {
      std::fclose(fc);
      return set_error(ENT_EIO, "%s: short write", csr_file);
    }
    done += int64_t(k);
  }
  std::fclose(fc);
  return ENT_OK;
}

static int count_nonisolated(const ent_csr *g) {
  int64_t c = 0;
  for (int64_t v = 0; v < g->n; ++v) c += g->beg_pos[v + 1] > g->beg_pos[v];
  return c > INT32_MAX ? INT32_MAX : int(c);
}

int ent_pick_roots(const ent_csr *g, int32_t k, uint64_t seed, int32_t *roots) {
  clear_error();
  if (!g || !roots || k < 0) return set_error(ENT_EINVAL, "bad argument");
  if (count_nonisolated(g) < k) return set_error(ENT_EINVAL, "graph has fewer than %d non-isolated vertices", k);
  uint64_t st = mix64(seed ^ 0x243F6A8885A308D3ull);
  for (int32_t i = 0; i < k;) {
    int32_t v = int32_t(splitmix_next(st) % uint64_t(g->n));
    if (g->beg_pos[v + 1] == g->beg_pos[v]) continue;
    bool fresh = true;
    for (int32_t j = 0; j < i; ++j)
      if (roots[j] == v) { fresh = false; break; }
    if (fresh) roots[i++] = v;
  }
  return ENT_OK;
}

int ent_pick_roots_glibc(const ent_csr *g, int32_t k, int32_t *roots) {
  clear_error();
  if (!g || !roots || k < 0) return set_error(ENT_EINVAL, "bad argument");
  if (count_nonisolated(g) < k) return set_error(ENT_EINVAL, "graph has fewer than %d non-isolated vertices", k);
  // glibc's rand() is random() over a 128-byte TYPE_3 state seeded with 1.
  char state[128];
  struct random_data rd;
  std::memset(&rd, 0, sizeof rd);
  if (initstate_r(1u, state, sizeof state, &rd) != 0) return set_error(ENT_EINTERNAL, "initstate_r failed");
  for (int32_t i = 0; i < k;) {
    int32_t r;
    random_r(&rd, &r);
    int32_t v = int32_t(int64_t(r) % g->n);
    if (g->beg_pos[v + 1] == g->beg_pos[v]) continue;
    bool fresh = true;
    for (int32_t j = 0; j < i; ++j)
      if (roots[j] == v) { fresh = false; break; }
    if (fresh) roots[i++] = v;
  }
  return ENT_OK;
}

int ent_pick_roots_mixed(const ent_csr *g, int32_t k, int32_t k_hub, uint64_t seed,
                         int32_t *roots) {
  clear_error();
  if (!g || !roots || k < 0 || k_hub < 0 || k_hub > k) return set_error(ENT_EINVAL, "bad argument");
  const int64_t n = g->n;
  std::vector<int32_t> ids(static_cast<size_t>(n));
  std::iota(ids.begin(), ids.end(), 0);
  auto deg = [&](int32_t v) { return g->beg_pos[v + 1] - g->beg_pos[v]; };
  int64_t top = std::max<int64_t>(1, (n + 99) / 100);
  std::nth_element(ids.begin(), ids.begin() + (top - 1), ids.end(), [&](int32_t x, int32_t y) {
    return deg(x) != deg(y) ? deg(x) > deg(y) : x < y;
  });
  std::vector<int32_t> hub(ids.begin(), ids.begin() + top);
  std::sort(hub.begin(), hub.end());
  hub.erase(std::remove_if(hub.begin(), hub.end(), [&](int32_t v) { return deg(v) == 0; }), hub.end());
  std::vector<int32_t> low;
  for (int64_t v = 0; v < n; ++v)
    if (deg(int32_t(v)) >= 1 && deg(int32_t(v)) <= 4 &&
        !std::binary_search(hub.begin(), hub.end(), int32_t(v)))
      low.push_back(int32_t(v));
  if (int64_t(hub.size()) < k_hub || int64_t(low.size()) < k - k_hub)
    return set_error(ENT_EINVAL, "not enough hub (%zu) or low-degree (%zu) vertices", hub.size(), low.size());
  uint64_t st = mix64(seed ^ 0x13198A2E03707344ull);
  // Partial Fisher-Yates draws without repeats.
  for (int32_t i = 0; i < k_hub; ++i) {
    size_t j = size_t(i) + size_t(splitmix_next(st) % uint64_t(hub.size() - size_t(i)));
    std::swap(hub[size_t(i)], hub[j]);
    roots[i] = hub[size_t(i)];
  }
  for (int32_t i = 0; i < k - k_hub; ++i) {
    size_t j = size_t(i) + size_t(splitmix_next(st) % uint64_t(low.size() - size_t(i)));
    std::swap(low[size_t(i)], low[j]);
    roots[k_hub + i] = low[size_t(i)];
  }
  return ENT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Text tuple files (the reference converter's input) and its output files.

namespace ent {
namespace {

inline bool tuple_sep(char ch) { return ch == ' ' || ch == '\n' || ch == '\t'; }

// atol on [p, end): leading white space, optional sign, decimal digits.
// Like atol it may read past the end of the token when the token itself
// starts with other white space ('\r', '\v', '\f').
int64_t text_atol(const char *p, const char *end) {
  while (p < end && (*p == ' ' || (*p >= '\t' && *p <= '\r'))) ++p;
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) neg = *p++ == '-';
  uint64_t x = 0;
  while (p < end && *p >= '0' && *p <= '9') x = x * 10 + uint64_t(*p++ - '0');
  return neg ? -int64_t(x) : int64_t(x);
}

// Token starts of the reference's scan (tuple_text_to_bin.cpp:112-139):
// offset 0, and every non-separator that follows a separator.
inline bool token_start(const char *d, int64_t p) { return p == 0 || (!tuple_sep(d[p]) && tuple_sep(d[p - 1])); }

struct MappedFile {
  const char *data = nullptr;
  size_t size = 0;
  int fd = -1;
  ~MappedFile() {
    if (data && size) munmap(const_cast<char *>(data), size);
    if (fd >= 0) close(fd);
  }
};

}  // namespace
}  // namespace ent

extern "C" {

void ent_tuples_free(ent_tuples *t) {
  if (!t) return;
  std::free(t->src);
  std::free(t->dst);
  t->src = t->dst = nullptr;
  t->n = t->ne = t->vmin = 0;
}

int ent_tuples_read(const char *path, int64_t skip_lines, ent_tuples *out) {
  clear_error();
  if (!path || !out) return set_error(ENT_EINVAL, "NULL argument");
  if (skip_lines < 0) return set_error(ENT_EINVAL, "skip_lines %lld < 0", (long long)skip_lines);
  std::memset(out, 0, sizeof *out);
  MappedFile f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) return set_error(ENT_EIO, "cannot open %s: %s", path, std::strerror(errno));
  struct stat sb;
  if (fstat(f.fd, &sb) != 0) return set_error(ENT_EIO, "cannot stat %s: %s", path, std::strerror(errno));
  f.size = size_t(sb.st_size);
  if (f.size) {
    void *p = mmap(nullptr, f.size, PROT_READ, MAP_PRIVATE, f.fd, 0);
    if (p == MAP_FAILED) return set_error(ENT_EIO, "cannot map %s: %s", path, std::strerror(errno));
    f.data = static_cast<const char *>(p);
  }
  const int64_t size = int64_t(f.size);
  // Line skipping as tuple_text_to_bin.cpp:80-97 does it: the offset advances
  // before the byte under it is examined.
  int64_t head = 0;
  for (int64_t skipped = 0; skipped < skip_lines;) {
    if (head >= size) return set_error(ENT_EINVAL, "%s has fewer than %lld lines to skip", path, (long long)skip_lines);
    ++head;
    if (head < size && f.data[head] == '\n') {
      ++head;
      ++skipped;
    }
  }
  const char *d = f.data + head;
  const int64_t len = size - head;
  const char *end = d + len;
  // Pass 1: token starts per chunk; pass 2: values at their global index.
  const int nt = host_threads();
  const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(int64_t(nt) * 8, len / 65536 + 1));
  const int64_t csz = (len + nchunk - 1) / nchunk;
  std::vector<int64_t> cnt(size_t(nchunk) + 1, 0);
  parallel_for(nchunk, 1, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      int64_t k = 0;
      for (int64_t p = c * csz, pe = std::min(len, (c + 1) * csz); p < pe; ++p) k += token_start(d, p);
      cnt[size_t(c) + 1] = k;
    }
  });
  for (int64_t c = 0; c < nchunk; ++c) cnt[size_t(c) + 1] += cnt[size_t(c)];
  const int64_t ntok = cnt[size_t(nchunk)];
  if (ntok == 0) return set_error(ENT_EINVAL, "%s holds no tuples", path);
  if (ntok & 1) return set_error(ENT_EINVAL, "%s holds an odd number of ids (%lld)", path, (long long)ntok);
  std::vector<int64_t> tok(static_cast<size_t>(ntok));
  std::vector<int64_t> lo(size_t(nchunk), INT64_MAX), hi(size_t(nchunk), INT64_MIN);
  parallel_for(nchunk, 1, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      int64_t k = cnt[size_t(c)], l = INT64_MAX, h = INT64_MIN;
      for (int64_t p = c * csz, pe = std::min(len, (c + 1) * csz); p < pe; ++p)
        if (token_start(d, p)) {
          const int64_t x = text_atol(d + p, end);
          tok[size_t(k++)] = x;
          l = std::min(l, x);
          h = std::max(h, x);
        }
      lo[size_t(c)] = l;
      hi[size_t(c)] = h;
    }
  });
  const int64_t vmin = *std::min_element(lo.begin(), lo.end());
  const int64_t vmax = *std::max_element(hi.begin(), hi.end());
  if (vmax - vmin >= int64_t(INT32_MAX))
    return set_error(ENT_EINVAL, "ids span [%lld, %lld]: 2^31 or more vertices", (long long)vmin, (long long)vmax);
  const int64_t ne = ntok / 2;
  out->src = alloc_array<int32_t>(ne);
  out->dst = alloc_array<int32_t>(ne);
  if (!out->src || !out->dst) {
    ent_tuples_free(out);
    return set_error(ENT_ENOMEM, "out of host memory (%lld tuples)", (long long)ne);
  }
  parallel_for(ne, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      out->src[i] = int32_t(tok[size_t(2 * i)] - vmin);
      out->dst[i] = int32_t(tok[size_t(2 * i + 1)] - vmin);
    }
  });
  out->n = vmax - vmin + 1;
  out->ne = ne;
  out->vmin = vmin;
  return ENT_OK;
}

int ent_csr_write_ref_files(const ent_csr *g, const int32_t *weight, const char *prefix) {
  clear_error();
  if (!g || !g->beg_pos || (g->m > 0 && !g->adj) || !prefix) return set_error(ENT_EINVAL, "NULL argument");
  const int64_t n = g->n, m = g->m;
  // Each file is written from int64 values produced block by block.
  auto write = [&](const char *suffix, int64_t count, auto value) -> int {
    const std::string name = std::string(prefix) + suffix;
    FILE *fp = std::fopen(name.c_str(), "wb");
    if (!fp) return set_error(ENT_EIO, "cannot create %s: %s", name.c_str(), std::strerror(errno));
    const int64_t blk = int64_t(1) << 20;
    std::vector<int64_t> buf(size_t(std::min(blk, std::max<int64_t>(count, 1))));
    for (int64_t done = 0; done < count;) {
      const int64_t k = std::min(blk, count - done);
      for (int64_t i = 0; i < k; ++i) buf[size_t(i)] = value(done + i);
      if (std::fwrite(buf.data(), 8, size_t(k), fp) != size_t(k)) {
        std::fclose(fp);
        return set_error(ENT_EIO, "%s: short write", name.c_str());
      }
      done += k;
    }
    if (std::fclose(fp) != 0) return set_error(ENT_EIO, "%s: %s", name.c_str(), std::strerror(errno));
    return ENT_OK;
  };
  int rc;
  if ((rc = write("_beg_pos.bin", n + 1, [&](int64_t i) { return g->beg_pos[i]; }))) return rc;
  if ((rc = write("_csr.bin", m, [&](int64_t i) { return int64_t(g->adj[i]); }))) return rc;
  if ((rc = write("_deg.bin", n, [&](int64_t i) { return g->beg_pos[i + 1] - g->beg_pos[i]; }))) return rc;
  // Row of entry i: advance a cursor over the rows (entries are visited in order).
  int64_t row = 0;
  if ((rc = write("_head.bin", m, [&](int64_t i) {
         while (g->beg_pos[row + 1] <= i) ++row;
         return row;
       })))
    return rc;
  if (weight && (rc = write("_weight.bin", m, [&](int64_t i) { return int64_t(weight[i]); }))) return rc;
  return ENT_OK;
}

}  // extern "C"
