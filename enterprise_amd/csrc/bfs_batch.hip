// Multi-source BFS (up to 512 roots per call) on a resident graph, behind
// include/ent_batch.h. No reference counterpart (the reference runs one root
// at a time, bfs_gpu_opt.cuh:370); each output row must equal the
// single-source result for its root, which the tests check bit for bit.
//
// Layout (vertex-major masks, W u64 words per vertex, bit s%64 of word s/64
// is source s):
//   seen  u64[n * W]              sources that reached v
//   lv    u64[R * n * W]          lv[L][v] = sources whose level-L frontier
//                                 holds v (new bits only; kept for the output);
//                                 defined only where jf[L] has v's bit
//   jf    u64[R * ceil(n/64)]     jf[L] = joint frontier of level L
//                                 (v in the level-L frontier of any source)
// Both are rings of R = lcap + 3 level slots (level L in slot L mod R). Before
// a level would reuse the slot of a level not yet written to the output rows,
// a flush pass writes the levels kept so far into the rows, so a search of
// any depth runs as one batch in a fixed workspace.
// Nothing reads lv[L][v] unless v is in jf[L], so lv is never cleared as a
// whole: a level clears lv[L + 2][v] only for the vertices that the next
// level's push can reach (unfinished, with edges), and the two readers of a
// row that may be stale (APPLY, and the pull's merge of pushed bits) read it
// only for such vertices.
// One launch index per phase, with the same launch-indexed state scheme as
// the single-source path (launch i reads the state and finished counters of
// launch i - 1; no fences, no contended atomics). The direction is chosen
// per 64-source word: words whose frontier is large are pulled, lagging ones
// pushed, and one level may do both:
//   TD     push: every joint-frontier vertex u ORs the push words of its
//          masks into lv[L+1][v] of its neighbours (minus seen) with
//          fire-and-forget atomics;
//   APPLY  after a TD-only level: seen |= lv[L+1], joint frontier bitmap,
//          counts;
//   BU     pull: every vertex with unseen bits of still-active pull-word
//          sources ORs the level-L masks of its neighbours, stopping once
//          every such source has been found (or the list ends); the bits
//          pushed to it at this level are merged in;
//   OUTPUT replays the kept levels per 64-vertex segment in LDS and writes
//          all k depth rows with 16-byte streaming stores (a flush: the same
//          for the levels kept so far, in the middle of the search).
// The pull runs in its own kernel, launched beside the others for the same
// launch index, so its mask registers do not cap their occupancy; with
// ENT_LAUNCH_PERSISTENT the whole search is one cooperative launch instead.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "ent_batch.h"
#include "ent_bfs.h"
#include "ent_common.h"
#include "ent_device.h"
#include "ent_internal.h"

namespace ent {

enum BPhase : int { BPH_TD = 0, BPH_BU = 1, BPH_APPLY = 2, BPH_OUTPUT = 3, BPH_TDF = 4 };
enum : int { BERR_TIMEOUT = 16 };

constexpr int kBStripes = 64;
constexpr int kBStripeU64 = 32;
// Counter words within a stripe: C_MFW + w is the degree sum of the vertices
// in the new frontier of the sources of word w.
enum : int {
  C_JWORDS = 0,
  C_MF = 1,
  C_ALG = 2,
  C_MUNF = 3,
  C_ACTIVE = 4,
  C_EDGES = 12,
  C_VERTS = 13,
  C_MFW = 14,
  C_WORK = 22,  // output: next 64-vertex segment to claim (not a count)
  C_COUNT = 23
};

struct BCounters {
  unsigned long long v[kBStripes * kBStripeU64];
};

struct BState {
  int level;
  int phase;
  int done;
  int error;
  int levels;
  int phases;
  unsigned int pull_words;  // words of this level expanded bottom-up (bit w)
  unsigned int push_words;  // words of this level expanded top-down
  unsigned long long bu_mask;  // levels that pulled some words
  unsigned long long td_mask;  // levels that pushed some words
  unsigned long long alg_bytes;
  unsigned long long total_edges;
  unsigned long long total_vertices;
  unsigned long long active[8];
  // Level ring (kept after the fields above: placed before pull_words, they
  // changed the pull kernel's code generation, 116 instead of 122 VGPRs and
  // 0.5 ms slower per R-MAT 23 batch).
  int base;       // first level not yet written to the output rows
  int flush;      // OUTPUT phase is a flush of levels base .. flush_end - 1
  int flush_end;
  int resume;     // phase to run after the flush
  int flushes;    // flushes so far
  int slot;       // ring slot of `level` in lv and jf
  // Degree sum of the vertices still missing an active source, as last
  // counted (a fused level does not sweep the vertices; the sum only falls,
  // so the last count bounds it from above).
  unsigned long long munf_last;
  int pulled;  // some level of this search pulled
  int pad;
};

struct BControl {
  BState st[2];
  GridSync sync;  // persistent form only
  BCounters cnt[3];
  int lphase[2];  // phase of launch i in lphase[i & 1] (-1 once done)
};

struct BArgs {
  const long long *off;
  const int *adj;
  const int *hub_block;
  const int *piece_v;
  const int *piece_k;
  unsigned long long *seen;
  unsigned long long *lv;
  unsigned long long *jf;  // ring of lcap + 3 joint-frontier bitmaps of nwords
  const unsigned long long *nz;  // bit v: degree(v) > 0
  int *out;
  BControl *ctl;
  long long m;
  long long nW;  // n * W
  int n;
  int nwords;
  int k;
  int lcap;
  int ring;  // lcap + 3 level slots in lv and jf
  int n_hub_block;
  int n_pieces;
  int t_wave, t_block, piece;
  int mode, alpha;
  int probe;
  int chunk_words;
  int out16;  // out is 16-byte aligned and n % 4 == 0: rows in 16-byte stores
  const int *rowmap;  // output row of the source at mask bit p (sources are placed by root degree)
  // Largest joint-frontier degree sum expanded as a fused level (0 = never;
  // W <= 4 only). When nonzero, every level also empties the ring slot two
  // levels ahead (see prep_word).
  long long fused;
};

template <int W>
struct alignas(8 * W) Mask {
  unsigned long long w[W];
};

template <int W>
__device__ __forceinline__ Mask<W> load_mask(const unsigned long long *base, long long v) {
  return *reinterpret_cast<const Mask<W> *>(base + v * W);
}
// Gather of v's mask for the pull: with H >= 0 (W = 8) only words 4H..4H+3
// (one 32-byte half), the rest zero; with H < 0 the whole mask.
template <int W, int H>
__device__ __forceinline__ Mask<W> gather_mask(const unsigned long long *base, long long v) {
  if constexpr (W == 8 && H >= 0) {
    typedef unsigned long long u64x4 __attribute__((ext_vector_type(4)));
    const u64x4 q = *reinterpret_cast<const u64x4 *>(base + v * W + 4 * H);
    Mask<W> m;
#pragma unroll
    for (int w = 0; w < W; ++w) m.w[w] = 0;
    m.w[4 * H + 0] = q.x;
    m.w[4 * H + 1] = q.y;
    m.w[4 * H + 2] = q.z;
    m.w[4 * H + 3] = q.w;
    return m;
  } else {
    return load_mask<W>(base, v);
  }
}
template <int W>
__device__ __forceinline__ void store_mask(unsigned long long *base, long long v, const Mask<W> &m) {
  *reinterpret_cast<Mask<W> *>(base + v * W) = m;
}

// Ring slot of level S.level + d, d = 0, 1 or 2 (the ring has at least four
// slots; the state carries the slot of its level, so no division is needed).
__device__ __forceinline__ int ring_slot(const BArgs &a, const BState &S, int d) {
  const int t = S.slot + d;
  return t >= a.ring ? t - a.ring : t;
}
__device__ __forceinline__ unsigned long long *level_ptr(const BArgs &a, int slot) {
  return a.lv + (long long)slot * a.nW;
}
__device__ __forceinline__ unsigned long long *jf_ptr(const BArgs &a, int slot) {
  return a.jf + (long long)slot * a.nwords;
}

// True if a push can reach v. On a symmetric graph (the only kind the pull
// searches exactly) a vertex without edges is no one's neighbour; a forced
// top-down search, exact on any adjacency, assumes every vertex can be.
__device__ __forceinline__ bool receives(const BArgs &a, long long deg) {
  return deg > 0 || a.mode == ENT_DIR_TOPDOWN;
}

template <int BS>
struct BSmem {
  static constexpr int NW = BS / kWave;
  union {
    struct {
      unsigned long long cand[NW][kWave];
      int excl[NW][kWave];
    } td;                              // top-down frontier staging
    int rows[ENT_BATCH_MAX_SOURCES];   // output: row of the source at each mask bit
  };
  unsigned char d8[NW][kWave * 68];  // output staging: 64 lanes x 64 sources (+4 pad)
  unsigned long long c[C_COUNT];
  long long seg;  // output: the segment the block claimed
  BState S;
};

__device__ __forceinline__ bool bitv(const unsigned long long *bm, int v) {
  return ((bm[v >> 6] >> (v & 63)) & 1ull) != 0;
}

// -- counters ----------------------------------------------------------------

template <int BS>
__device__ __forceinline__ void block_counts_reset(BSmem<BS> &sm) {
  if (threadIdx.x < C_COUNT) sm.c[threadIdx.x] = 0;
}

// Reduces per-lane values over the wavefront and adds them to LDS.
// Degree sums of the new frontier by source word, wave-uniform (scalar
// registers): each step adds the wave sum of the lanes' degrees, each capped
// at 2^25 so that 64 of them fit 32 bits (the sums only steer the direction
// choice).
template <int W>
struct WordSums {
  unsigned long long d[W];
};

template <int W>
__device__ __forceinline__ WordSums<W> word_sums_zero() {
  WordSums<W> z;
#pragma unroll
  for (int w = 0; w < W; ++w) z.d[w] = 0;
  return z;
}

// Adds the degree of this lane's vertex to the words where x has bits. All
// lanes of the wavefront must be active; invalid lanes pass x = 0.
template <int W>
__device__ __forceinline__ void word_sums_add(WordSums<W> &ws, const Mask<W> &x, long long deg) {
  const unsigned int dc = (unsigned int)min(deg, (long long)(1 << 25));
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const unsigned int part = x.w[w] ? dc : 0u;
    if (__ballot(part != 0)) ws.d[w] += (unsigned long long)(unsigned int)wave_sum(int(part));
  }
}

template <int BS, int W>
__device__ __forceinline__ void wave_counts(BSmem<BS> &sm, unsigned long long jw, unsigned long long mf,
                                            unsigned long long alg, unsigned long long munf, const Mask<W> &act,
                                            unsigned long long edges, unsigned long long verts,
                                            const WordSums<W> &mfw) {
  unsigned long long vals[6] = {jw, mf, alg, munf, edges, verts};
  Mask<W> ac = act;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int j = 0; j < 6; ++j) vals[j] += __shfl_xor(vals[j], o);
#pragma unroll
    for (int w = 0; w < W; ++w) ac.w[w] |= __shfl_xor(ac.w[w], o);
  }
  if (lane_id() == 0) {
    const int idx[6] = {C_JWORDS, C_MF, C_ALG, C_MUNF, C_EDGES, C_VERTS};
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (vals[j]) atomicAdd(&sm.c[idx[j]], vals[j]);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (ac.w[w]) atomicOr(&sm.c[C_ACTIVE + w], ac.w[w]);
      if (mfw.d[w]) atomicAdd(&sm.c[C_MFW + w], mfw.d[w]);
    }
  }
}

template <int BS>
__device__ __forceinline__ void block_counts_flush(const BArgs &a, BSmem<BS> &sm, int slot) {
  __syncthreads();
  if (threadIdx.x < C_COUNT) {
    const unsigned long long x = sm.c[threadIdx.x];
    if (x) {
      unsigned long long *dst = &a.ctl->cnt[slot].v[(blockIdx.x % kBStripes) * kBStripeU64 + threadIdx.x];
      if (threadIdx.x >= C_ACTIVE && threadIdx.x < C_ACTIVE + 8) atomicOr(dst, x);
      else atomicAdd(dst, x);
    }
  }
}

// -- state -------------------------------------------------------------------

__device__ BState next_bstate(const BArgs &a, const BState &p, const unsigned long long *tot, int W) {
  BState s = p;
  if (p.done) return s;
  s.phases = p.phases + 1;
  if (p.phase == BPH_TD) {  // pushed words settle in the pull phase, if any, else in APPLY
    s.phase = p.pull_words ? BPH_BU : BPH_APPLY;
    return s;
  }
  if (p.phase == BPH_OUTPUT) {
    if (p.flush) {  // levels base .. flush_end - 1 are in the rows: resume
      s.flush = 0;
      s.base = p.flush_end;
      s.phase = p.resume;
      return s;
    }
    s.total_edges = tot[C_EDGES];
    s.total_vertices = tot[C_VERTS];
    s.alg_bytes += 4ull * 64ull * (unsigned long long)W * (unsigned long long)a.n;
    s.done = 1;
    return s;
  }
  const int L = p.level;  // -1 before level 0
  if (L >= 0 && p.pull_words && L < 64) s.bu_mask |= 1ull << L;
  if (L >= 0 && p.push_words && L < 64) s.td_mask |= 1ull << L;
  s.alg_bytes += tot[C_ALG];
  if (tot[C_JWORDS] == 0) {
    s.levels = L + 1;
    s.phase = BPH_OUTPUT;
    return s;
  }
  for (int w = 0; w < 8; ++w) s.active[w] = w < W ? tot[C_ACTIVE + w] : 0ull;
  // A fused level counts no unfinished vertices: its successor steers with
  // the last count.
  const unsigned long long mf = tot[C_MF], munf = p.phase == BPH_TDF ? p.munf_last : tot[C_MUNF];
  s.munf_last = munf;
  // Direction per 64-source word: a word whose new frontier is large is
  // pulled, a word that lags is pushed, so a pull can stop once the large
  // words are covered instead of scanning every row for the lagging ones.
  unsigned pull = 0, live = 0;
  for (int w = 0; w < W; ++w) {
    if (!s.active[w]) continue;
    live |= 1u << w;
    bool bu;
    switch (a.mode) {
      case ENT_DIR_TOPDOWN: bu = false; break;
      case ENT_DIR_BOTTOMUP: bu = true; break;
      case ENT_DIR_REFERENCE:  // the reference's one-way switch, for every word at once
        bu = p.pull_words != 0 || (L + 1 >= 1 && 10ull * mf >= 3ull * (unsigned long long)a.m);
        break;
      default: bu = tot[C_MFW + w] * (unsigned long long)a.alpha > munf;
    }
    if (bu) pull |= 1u << w;
  }
  s.level = L + 1;
  s.slot = p.slot + 1 == a.ring ? 0 : p.slot + 1;
  s.pull_words = pull;
  s.push_words = live & ~pull;
  if (pull) s.pulled = 1;
  s.phase = s.push_words ? BPH_TD : BPH_BU;
  // A small all-push frontier is pushed and settled in one launch.
  if (pull == 0 && a.fused > 0 && mf <= (unsigned long long)a.fused) s.phase = BPH_TDF;
  // Level L' = s.level uses the slots of L' (its frontier), L' + 1 (its new
  // bits) and L' + 2 (cleared for the next push). The ring holds them beside
  // the levels base .. L' - 1 not yet in the rows while L' - base <= lcap;
  // beyond that, levels base .. L' (L' complete: it is the new frontier) are
  // flushed to the rows first.
  if (s.level - s.base > a.lcap) {
    s.resume = s.phase;
    s.phase = BPH_OUTPUT;
    s.flush = 1;
    s.flushes += 1;
    s.flush_end = s.level + 1;
  }
  return s;
}

template <int BS, int W>
__device__ __forceinline__ void begin_blaunch(const BArgs &a, BSmem<BS> &sm, long long i, bool coherent = false) {
  if (threadIdx.x < kWave) {
    BControl *c = a.ctl;
    const int lane = threadIdx.x;
    const BCounters &pc = c->cnt[(i + 2) % 3];
    // Totals of the 64 stripes: lane j sums (or ORs) counter j over the
    // stripes with independent loads, so no cross-lane steps are needed.
    __shared__ unsigned long long t[C_COUNT];
    if (lane < C_COUNT) {
      const bool is_or = lane >= C_ACTIVE && lane < C_ACTIVE + 8;
      unsigned long long acc = 0;
#pragma unroll 16
      for (int st = 0; st < kBStripes; ++st) {
        const unsigned long long *q = &pc.v[st * kBStripeU64 + lane];
        const unsigned long long x = coherent ? ld_agent(q) : *q;
        acc = is_or ? (acc | x) : (acc + x);
      }
      t[lane] = acc;
    }
    wave_sync();
    if (lane == 0) {
      BState prev;
      if (coherent) {
        const unsigned *s32 = reinterpret_cast<const unsigned *>(&c->st[(i + 1) & 1]);
        unsigned *d32 = reinterpret_cast<unsigned *>(&prev);
        for (int w = 0; w < int(sizeof(BState) / 4); ++w) d32[w] = ld_agent(s32 + w);
      } else {
        prev = c->st[(i + 1) & 1];
      }
      const BState S = next_bstate(a, prev, t, W);
      sm.S = S;
      if (blockIdx.x == 0) {
        c->st[i & 1] = S;
        c->lphase[i & 1] = S.done ? -1 : S.phase;
      }
    }
    if (blockIdx.x == 0 && lane < C_COUNT)
      for (int s = 0; s < kBStripes; ++s) c->cnt[(i + 1) % 3].v[s * kBStripeU64 + lane] = 0;
  }
  block_counts_reset(sm);
  __syncthreads();
}

// -- phases ------------------------------------------------------------------

// Keeps the words of f that this level pushes; true if any bit is left.
template <int W>
__device__ __forceinline__ bool push_mask(Mask<W> &f, unsigned words) {
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (!((words >> w) & 1u)) f.w[w] = 0;
    any |= f.w[w] != 0;
  }
  return any;
}

// Per-lane counts of a fused level's claims (the APPLY counts, taken where
// bits first land instead of in a sweep over every vertex).
template <int W>
struct Claims {
  unsigned long long jw, mf, alg;
  unsigned long long mfw[W];
  Mask<W> act;
};

// One edge of the push to u. Plain: OR the bits of f that u has not seen into
// lv[L + 1][u], fire and forget. CLAIM (a fused level): the bits also go into
// seen[u] at once, and the OR into lv[L + 1][u] returns the word's old value.
// That word was zero before the level began (prep_word), so the edge that
// finds it zero is the first to reach u in that word: it counts u's word, and
// the first word to claim u also sets u's bit in the new joint frontier jn.
template <int W, bool CLAIM>
__device__ __forceinline__ void push_edge(const BArgs &a, unsigned long long *dst, int u, const Mask<W> &f,
                                          unsigned long long *jn, Claims<W> &cl) {
  const Mask<W> sd = load_mask<W>(a.seen, u);
  if (!CLAIM) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const unsigned long long nb = f.w[w] & ~sd.w[w];
      if (nb) atomicOr(dst + (long long)u * W + w, nb);
    }
    return;
  }
  unsigned long long old[W];
  bool first = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const unsigned long long nb = f.w[w] & ~sd.w[w];
    old[w] = ~0ull;
    if (nb) {
      atomicOr(a.seen + (long long)u * W + w, nb);
      old[w] = atomicOr(dst + (long long)u * W + w, nb);
      cl.act.w[w] |= nb;
      first |= old[w] == 0;
    }
  }
  if (!first) return;
  const long long deg = a.off[u + 1] - a.off[u];
  const unsigned long long dc = (unsigned long long)min(deg, (long long)(1 << 25));
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (old[w] == 0) {
      cl.jw += 1;
      cl.alg += 8ull + 12ull * (unsigned long long)deg;
      cl.mfw[w] += dc;
    }
  }
  const unsigned long long bit = 1ull << (u & 63);
  if ((atomicOr(jn + (u >> 6), bit) & bit) == 0) cl.mf += (unsigned long long)deg;
}

template <int W, bool CLAIM>
__device__ __forceinline__ void push_range(const BArgs &a, unsigned long long *dst, long long s, long long e,
                                           int t, int stride, const Mask<W> &f, unsigned long long *jn,
                                           Claims<W> &cl) {
  for (long long i = s + t; i < e; i += stride) push_edge<W, CLAIM>(a, dst, a.adj[i], f, jn, cl);
}

// Empties word wd of a ring slot (masks lz, bitmap jz) for the level that
// will next use the slot: clears the mask of every vertex of the word in
// the slot's old joint frontier, then the word. Every nonzero mask of a slot
// belongs to a vertex in that slot's bitmap (a push reaches only vertices
// that join the new frontier, a pull stores only new bits), so once every
// word is emptied the slot is zero throughout, as a fused level's claims
// require. Every level empties the slot of level L + 2 this way (the fused
// level beside its sweep, APPLY and the pull in theirs).
template <int W>
__device__ __forceinline__ void prep_word(unsigned long long *lz, unsigned long long *jz, long long wd) {
  unsigned long long m = jz[wd];
  if (!m) return;
  jz[wd] = 0;
  for (; m; m &= m - 1) {
    const long long v = wd * 64 + (__ffsll((long long)m) - 1);
#pragma unroll
    for (int w = 0; w < W; ++w) lz[v * W + w] = 0;
  }
}

// Top-down level L. Plain (TD): pushes only; APPLY or the pull settles the
// new bits. CLAIM (TDF, a fused level): pushes with claims, so the new
// frontier's bitmap, seen and the counts are complete when the launch ends,
// and empties the slot of level L + 2 (prep_word) as APPLY and the pull do.
template <int BS, int W, bool CLAIM>
__device__ void btd_phase(const BArgs &a, BSmem<BS> &sm, int L) {
  const unsigned long long *fs = level_ptr(a, ring_slot(a, sm.S, 0));
  unsigned long long *dst = level_ptr(a, ring_slot(a, sm.S, 1));
  const unsigned long long *jf = jf_ptr(a, ring_slot(a, sm.S, 0));
  unsigned long long *jn = jf_ptr(a, ring_slot(a, sm.S, 1));
  Claims<W> cl;
  cl.jw = cl.mf = cl.alg = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) cl.mfw[w] = cl.act.w[w] = 0;
  const unsigned pw = sm.S.push_words;  // only these words are pushed at this level
  for (int p = blockIdx.x; p < a.n_pieces; p += gridDim.x) {
    const int v = a.piece_v[p];
    if (!bitv(jf, v)) continue;
    Mask<W> f = load_mask<W>(fs, v);
    if (!push_mask<W>(f, pw)) continue;
    const long long s = a.off[v] + (long long)a.piece_k[p] * a.piece;
    push_range<W, CLAIM>(a, dst, s, min(a.off[v + 1], s + a.piece), threadIdx.x, BS, f, jn, cl);
  }
  for (int h = blockIdx.x; h < a.n_hub_block; h += gridDim.x) {
    const int v = a.hub_block[h];
    if (!bitv(jf, v)) continue;
    Mask<W> f = load_mask<W>(fs, v);
    if (!push_mask<W>(f, pw)) continue;
    push_range<W, CLAIM>(a, dst, a.off[v], a.off[v + 1], threadIdx.x, BS, f, jn, cl);
  }
  const int wid = threadIdx.x / kWave, lane = lane_id();
  unsigned long long *wc = sm.td.cand[wid];
  int *we = sm.td.excl[wid];
  // A fused level's frontier is sparse: chunks of 64 bitmap words (4,096
  // vertices) keep a wavefront's lanes busy and its chunks per level few.
  const int CW = CLAIM ? kWave : a.chunk_words;
  const long long nchunks = ((long long)a.nwords + CW - 1) / CW;
  const long long nwv = (long long)gridDim.x * (BS / kWave);
  unsigned long long *lz = level_ptr(a, ring_slot(a, sm.S, 2));
  unsigned long long *jz = jf_ptr(a, ring_slot(a, sm.S, 2));
  for (long long ch = (long long)blockIdx.x * (BS / kWave) + wid; ch < nchunks; ch += nwv) {
    const long long w = ch * CW + lane;
    const unsigned long long word = (lane < CW && w < a.nwords) ? jf[w] : 0ull;
    // The slot of level L + 2, word w (a fused level's chunks cover every
    // word once, a word per lane), emptied beside the sweep.
    if (CLAIM && w < a.nwords) prep_word<W>(lz, jz, w);
    if (__ballot(word != 0) == 0) continue;
    const int cnt = __popcll(word);
    const int incl = wave_incl_scan(cnt);
    wc[lane] = word;
    we[lane] = incl - cnt;
    wave_sync();
    const int total = wave_read(incl, 63);
    for (int b0 = 0; b0 < total; b0 += kWave) {
      const int kk = b0 + lane;
      int v = 0;
      long long s = 0, e = 0;
      if (kk < total) {
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (we[mid] <= kk) lo = mid;
          else hi = mid - 1;
        }
        v = int((ch * CW + lo) * 64 + select_bit(wc[lo], kk - we[lo]));
        s = a.off[v];
        e = a.off[v + 1];
        if (e - s < a.t_wave) {
          Mask<W> f = load_mask<W>(fs, v);
          if (push_mask<W>(f, pw)) push_range<W, CLAIM>(a, dst, s, e, 0, 1, f, jn, cl);
        }
      }
      const long long d = e - s;
      unsigned long long wm = __ballot(kk < total && d >= a.t_wave && d < a.t_block);
      while (wm) {
        const int src = __ffsll((long long)wm) - 1;
        wm &= wm - 1;
        const int vv = __shfl(v, src);
        const long long ss = __shfl(s, src), ee = __shfl(e, src);
        Mask<W> f = load_mask<W>(fs, vv);
        if (push_mask<W>(f, pw)) push_range<W, CLAIM>(a, dst, ss, ee, lane, kWave, f, jn, cl);
      }
    }
    wave_sync();
  }
  if (CLAIM) {
    WordSums<W> mfw;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      unsigned long long x = cl.mfw[w];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
      mfw.d[w] = x;
    }
    wave_counts<BS, W>(sm, cl.jw, cl.mf, cl.alg, 0, cl.act, 0, 0, mfw);
  }
}

// prep_word inside a sweep with a lane per vertex: the masks of the word's
// old members (bits of oj, wave-uniform), then the word.
template <int W>
__device__ __forceinline__ void clear_old_members(unsigned long long *lz, unsigned long long *jz, long long wd,
                                                  unsigned long long oj, int lane) {
  if ((oj >> lane) & 1ull) {
    Mask<W> z;
#pragma unroll
    for (int w = 0; w < W; ++w) z.w[w] = 0;
    store_mask<W>(lz, wd * 64 + lane, z);
  }
  if (lane == 0) jz[wd] = 0;
}

// Tail of APPLY for the vertex of this lane: new bits x (the lane's vertex
// v < n when valid), seen update, counts, and the clearing of lv[L + 2][v]
// when the next level's push can still reach v.
template <int W>
__device__ __forceinline__ void settle_vertex(const BArgs &a, unsigned long long *clr, bool valid, long long v,
                                              const Mask<W> &sd,
                                              const Mask<W> &x, long long deg, const unsigned long long *active,
                                              unsigned long long &jw, unsigned long long &mf, unsigned long long &alg,
                                              unsigned long long &munf, Mask<W> &act, WordSums<W> &mfw,
                                              bool &any) {
  any = false;
  word_sums_add<W>(mfw, x, deg);  // x = 0 on invalid lanes
  if (!valid) return;
  int nzw = 0;
  Mask<W> ns;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    nzw += x.w[w] != 0;
    ns.w[w] = sd.w[w] | x.w[w];
    act.w[w] |= x.w[w];
  }
  any = nzw != 0;
  if (any) {
    store_mask<W>(a.seen, v, ns);
    jw += (unsigned long long)nzw;
    mf += (unsigned long long)deg;
    alg += (unsigned long long)nzw * (8ull + 12ull * (unsigned long long)deg);
  }
  bool unf = false;
#pragma unroll
  for (int w = 0; w < W; ++w) unf |= (active[w] & ~ns.w[w]) != 0;
  if (unf) munf += (unsigned long long)deg;
  // The next level pushes only bits of sources active now, and only to
  // vertices still missing one of them.
  if (unf && receives(a, deg)) {
    Mask<W> z;
#pragma unroll
    for (int w = 0; w < W; ++w) z.w[w] = 0;
    store_mask<W>(clr, v, z);
  }
}

template <int BS, int W>
__device__ void bapply_phase(const BArgs &a, BSmem<BS> &sm, int L) {
  unsigned long long *nx = level_ptr(a, ring_slot(a, sm.S, 1));
  unsigned long long *jn = jf_ptr(a, ring_slot(a, sm.S, 1));
  unsigned long long *clr = level_ptr(a, ring_slot(a, sm.S, 2));  // lv[L + 2]
  unsigned long long *pj = jf_ptr(a, ring_slot(a, sm.S, 2));
  const bool prep = W <= 4 && a.fused > 0;  // empty the slot of level L + 2 (prep_word)
  const int lane = lane_id();
  const long long nwv = (long long)gridDim.x * (BS / kWave);
  unsigned long long jw = 0, mf = 0, alg = 0, munf = 0;
  WordSums<W> mfw = word_sums_zero<W>();
  Mask<W> act;
#pragma unroll
  for (int w = 0; w < W; ++w) act.w[w] = 0;
  for (long long wd = (long long)blockIdx.x * (BS / kWave) + threadIdx.x / kWave; wd < a.nwords; wd += nwv) {
    const long long v = wd * 64 + lane;
    const bool valid = v < a.n;
    const unsigned long long oj = prep ? pj[wd] : 0ull;
    Mask<W> x, sd;
#pragma unroll
    for (int w = 0; w < W; ++w) x.w[w] = sd.w[w] = 0;
    long long deg = 0;
    if (valid) {
      x = load_mask<W>(nx, v);
      sd = load_mask<W>(a.seen, v);
      deg = a.off[v + 1] - a.off[v];
      // lv[L + 1][v] was cleared (then pushed to) iff v still missed an
      // active source and can receive pushes; otherwise it is stale and
      // nothing reached v. (Loaded with seen, not after it: the dependent
      // load costs the shared kernel a wavefront per SIMD in registers.)
      bool open = false;
#pragma unroll
      for (int w = 0; w < W; ++w) open |= (sm.S.active[w] & ~sd.w[w]) != 0;
      if (!(open && receives(a, deg))) {
#pragma unroll
        for (int w = 0; w < W; ++w) x.w[w] = 0;
      }
    }
    bool any;
    settle_vertex<W>(a, clr, valid, v, sd, x, deg, sm.S.active, jw, mf, alg, munf, act, mfw, any);
    const unsigned long long bits = __ballot(any);
    if (lane == 0) jn[wd] = bits;
    if (oj) clear_old_members<W>(clr, pj, wd, oj, lane);
  }
  wave_counts<BS, W>(sm, jw, mf, alg, munf, act, 0, 0, mfw);
}

// Bottom-up tail for the vertex of this lane. x = new bits: pulled bits of
// the pull words plus, when this level also pushed, the bits pushed to v.
// Both lie in ft = active & ~seen(v), whose popcount is ftc, so v stays
// unfinished for some active source iff x has fewer bits. The seen mask is
// read back only for vertices that gained bits. act is wave-uniform: it is
// merged with a wave OR only when some lane holds a source not yet in it.
template <int W>
__device__ __forceinline__ void settle_bu(const BArgs &a, unsigned long long *clr, bool valid, long long v, int ftc,
                                         const Mask<W> &x, long long deg, unsigned long long &jw,
                                         unsigned long long &mf, unsigned long long &alg, unsigned long long &munf,
                                         Mask<W> &act, WordSums<W> &mfw, bool &any) {
  int nzw = 0, xc = 0;
  bool fresh = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    nzw += x.w[w] != 0;
    xc += __popcll(x.w[w]);
    fresh |= (x.w[w] & ~act.w[w]) != 0;
  }
  any = nzw != 0;
  if (__ballot(fresh)) {
#pragma unroll
    for (int w = 0; w < W; ++w) act.w[w] |= wave_or64(x.w[w]);
  }
  word_sums_add<W>(mfw, x, deg);  // x = 0 on invalid lanes
  if (!valid) return;
  if (any) {
    Mask<W> ns = load_mask<W>(a.seen, v);
#pragma unroll
    for (int w = 0; w < W; ++w) ns.w[w] |= x.w[w];
    store_mask<W>(a.seen, v, ns);
    jw += (unsigned long long)nzw;
    mf += (unsigned long long)deg;
    alg += (unsigned long long)nzw * (8ull + 12ull * (unsigned long long)deg);
  }
  if (xc < ftc) {  // unfinished: the next level's push can still reach v
    munf += (unsigned long long)deg;
    if (receives(a, deg)) {
      Mask<W> z;
#pragma unroll
      for (int w = 0; w < W; ++w) z.w[w] = 0;
      store_mask<W>(clr, v, z);
    }
  }
}

// Bottom-up (pull) level: every vertex v gathers the level-L masks of its
// neighbours (only those in the joint frontier) until the OR covers
// tgt = active & ~seen(v) restricted to the pull words, or the list ends.
// Lanes probe their own first `probe` neighbours two at a time; longer
// remaining lists are scanned by the whole wavefront, one vertex at a time,
// with a wave OR per 64 neighbours. When this level also pushed some words
// (a top-down launch ran just before), the bits pushed to v are merged in.
template <int BS, int W, int H>
__device__ void bbu_phase_h(const BArgs &a, BSmem<BS> &sm, int L) {
  const unsigned long long *fs = level_ptr(a, ring_slot(a, sm.S, 0));
  const unsigned long long *jf = jf_ptr(a, ring_slot(a, sm.S, 0));  // joint frontier of level L
  unsigned long long *nx = level_ptr(a, ring_slot(a, sm.S, 1));
  unsigned long long *jn = jf_ptr(a, ring_slot(a, sm.S, 1));
  unsigned long long *clr = level_ptr(a, ring_slot(a, sm.S, 2));  // lv[L + 2]
  unsigned long long *pj = jf_ptr(a, ring_slot(a, sm.S, 2));
  const bool prep = W <= 4 && a.fused > 0;  // empty the slot of level L + 2 (prep_word)
  const int lane = lane_id();
  const long long nwv = (long long)gridDim.x * (BS / kWave);
  unsigned long long jw = 0, mf = 0, alg = 0, munf = 0;
  Mask<W> act;
  WordSums<W> mfw = word_sums_zero<W>();
#pragma unroll
  for (int w = 0; w < W; ++w) act.w[w] = 0;
  const unsigned pull = sm.S.pull_words;
  const bool merge = sm.S.push_words != 0;
  for (long long wd = (long long)blockIdx.x * (BS / kWave) + threadIdx.x / kWave; wd < a.nwords; wd += nwv) {
    const long long v = wd * 64 + lane;
    // A vertex without edges can neither pull nor be pushed to (the pull
    // runs only on a symmetric adjacency): nothing of it is loaded.
    const bool valid = v < a.n && ((a.nz[wd] >> lane) & 1ull) != 0;
    const unsigned long long oj = prep ? pj[wd] : 0ull;
    Mask<W> tgt, acc;
    bool need = false;
    int ftc = 0;
    long long s = 0, e = 0, p = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) tgt.w[w] = acc.w[w] = 0;
    if (valid) {
      const Mask<W> sd = load_mask<W>(a.seen, v);
      s = a.off[v];
      e = a.off[v + 1];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const unsigned long long ft = sm.S.active[w] & ~sd.w[w];
        ftc += __popcll(ft);
        tgt.w[w] = ((pull >> w) & 1u) ? ft : 0ull;
        need |= tgt.w[w] != 0;
      }
      need = need && e > s;
    }
    bool complete = !need;
    if (need) {
      const long long pe = min(e, s + a.probe);
      // Two neighbours per step: both joint-frontier tests, then both mask
      // gathers are in flight before the OR and the early-exit test.
      for (p = s; p < pe; p += 2) {
        const int u0 = a.adj[p];
        const bool two = p + 1 < pe;
        const int u1 = two ? a.adj[p + 1] : u0;
        const bool in0 = bitv(jf, u0), in1 = two && bitv(jf, u1);
        Mask<W> f0, f1;
#pragma unroll
        for (int w = 0; w < W; ++w) f0.w[w] = f1.w[w] = 0;
        if (in0) f0 = gather_mask<W, H>(fs, u0);
        if (in1) f1 = gather_mask<W, H>(fs, u1);
        bool all = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          acc.w[w] |= f0.w[w] | f1.w[w];
          all &= (acc.w[w] & tgt.w[w]) == tgt.w[w];
        }
        if (all) {
          complete = true;
          p += 2;
          break;
        }
      }
      if (p > pe) p = pe;
    }
    // Long lists: the whole wavefront scans one vertex at a time; the
    // vertex's target and running OR are wave-uniform while it does.
    unsigned long long lmask = __ballot(need && !complete && p < e);
    while (lmask) {
      const int src = __ffsll((long long)lmask) - 1;
      lmask &= lmask - 1;
      const long long pp = (long long)read_lane64((unsigned long long)p, src);
      const long long ee = (long long)read_lane64((unsigned long long)e, src);
      Mask<W> t2, a2;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        t2.w[w] = read_lane64(tgt.w[w], src);
        a2.w[w] = read_lane64(acc.w[w], src);
      }
      for (long long q0 = pp; q0 < ee; q0 += kWave) {
        const long long q = q0 + lane;
        Mask<W> f;
#pragma unroll
        for (int w = 0; w < W; ++w) f.w[w] = 0;
        if (q < ee) {
          const int u = a.adj[q];
          if (bitv(jf, u)) f = gather_mask<W, H>(fs, u);
        }
        bool all = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          a2.w[w] |= wave_or64(f.w[w]);
          all &= (a2.w[w] & t2.w[w]) == t2.w[w];
        }
        if (all) break;
      }
      if (lane == src) acc = a2;
    }
    Mask<W> x;
#pragma unroll
    for (int w = 0; w < W; ++w) x.w[w] = acc.w[w] & tgt.w[w];
    // Bits pushed to v by this level's top-down launch: lv[L + 1][v] was
    // cleared for v iff v missed an active source (ftc > 0) and has edges.
    if (valid && merge && ftc > 0 && receives(a, e - s)) {
      const Mask<W> xp = load_mask<W>(nx, v);
#pragma unroll
      for (int w = 0; w < W; ++w) x.w[w] |= xp.w[w];
    }
    // lv[L + 1] already holds zero, or the bits this level pushed to v:
    // store only when the pull found new bits.
    {
      bool pulled = false;
#pragma unroll
      for (int w = 0; w < W; ++w) pulled |= (acc.w[w] & tgt.w[w]) != 0;
      if (valid && pulled) store_mask<W>(nx, v, x);
    }
    bool any;
    settle_bu<W>(a, clr, valid, v, ftc, x, e - s, jw, mf, alg, munf, act, mfw, any);
    const unsigned long long bits = __ballot(any);
    if (lane == 0) jn[wd] = bits;
    if (oj) clear_old_members<W>(clr, pj, wd, oj, lane);
  }
  wave_counts<BS, W>(sm, jw, mf, alg, munf, act, 0, 0, mfw);
}

// Bottom-up level L, gathering one half of each neighbour's mask when every
// pull word lies in it (W = 8).
template <int BS, int W>
__device__ void bbu_phase(const BArgs &a, BSmem<BS> &sm, int L) {
  const unsigned pull = sm.S.pull_words;
  if (W == 8 && (pull & 0xF0u) == 0) bbu_phase_h<BS, W, 0>(a, sm, L);
  else if (W == 8 && (pull & 0x0Fu) == 0) bbu_phase_h<BS, W, 1>(a, sm, L);
  else bbu_phase_h<BS, W, -1>(a, sm, L);
}

// Transpose of a 64 x 64 bit matrix held one row per lane: returns, in lane
// c, the word whose bit r is bit c of lane r's word (six exchange steps of
// halving blocks).
__device__ __forceinline__ unsigned long long lane_transpose(unsigned long long x, int lane) {
  const unsigned long long m[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                   0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int j = 32 >> s;
    const unsigned long long o = __shfl_xor(x, j);
    x = (lane & j) ? (x & ~m[s]) | ((o >> j) & m[s]) : (x & m[s]) | ((o << j) & ~m[s]);
  }
  return x;
}

// 0xFF in byte j of the result for each set bit j of bits 4k .. 4k+3 of m.
__device__ __forceinline__ unsigned nibble_bytes(unsigned long long m, int k) {
  unsigned x = (unsigned)(m >> (4 * k)) & 0xFu;
  x = (x | (x << 7) | (x << 14) | (x << 21)) & 0x01010101u;
  return x * 0xFFu;
}

// Output replay with one 64-source group (W = 1) of levels L0 .. L0 + 63
// (fl: lane j holds the segment's bitmap word of level L0 + j). Each lane
// takes its own vertex's member levels from the transposed words and loads
// their masks two at a time; where every lane's masks are sparse (at most
// four sources each, as on a high-diameter graph) the table gets one byte
// per source bit, otherwise the nibble expansion.
template <class SlotFn>
__device__ __forceinline__ void replay_lanes(const BArgs &a, unsigned char *d8, unsigned long long fl, int L0, int lo,
                                             long long v, SlotFn slot) {
  const int lane = lane_id();
  unsigned long long mem = lane_transpose(fl, lane);  // bit j: v is in level L0 + j
  unsigned int *row = reinterpret_cast<unsigned int *>(d8 + lane * 68);
  while (__ballot(mem != 0)) {
    constexpr int kLv = 2;  // 4 costs the shared kernel a wavefront per SIMD
    unsigned long long mk[kLv];
    unsigned lvl[kLv];
#pragma unroll
    for (int t = 0; t < kLv; ++t) {
      mk[t] = 0;
      lvl[t] = 0;
      if (mem) {
        const int j = __ffsll((long long)mem) - 1;
        mem &= mem - 1;
        lvl[t] = (unsigned)(L0 - lo + j);
        mk[t] = a.lv[(long long)slot(L0 + j) * a.nW + v];
      }
    }
    bool dense = false;
#pragma unroll
    for (int t = 0; t < kLv; ++t) dense |= __popcll(mk[t]) > 4;
    if (__ballot(dense)) {
#pragma unroll 4
      for (int k = 0; k < 16; ++k) {
        unsigned w = row[k];
#pragma unroll
        for (int t = 0; t < kLv; ++t) {
          const unsigned msk = nibble_bytes(mk[t], k);
          w = (w & ~msk) | ((lvl[t] * 0x01010101u) & msk);
        }
        row[k] = w;
      }
    } else {
#pragma unroll
      for (int t = 0; t < kLv; ++t)
        for (unsigned long long b = mk[t]; b; b &= b - 1)
          d8[lane * 68 + (__ffsll((long long)b) - 1)] = (unsigned char)lvl[t];
    }
  }
}

// Output: per 64-vertex segment, replay the level masks of each 64-source
// group into an LDS table of levels (u8, 0xFF = unreached) and store the 64
// rows of the group with 256-byte coalesced writes. A vertex's mask line of
// level L is fetched only when the vertex is in that level's joint frontier
// (about 2 of 8 levels at R-MAT scale 23, none for a vertex without edges).
// With W > 1 the wavefronts of a block share one segment (wavefront w takes
// groups w, w + 4, ...), so each fetched line is read by every group while
// hot; with one group, the block claims four segments at once, one per
// wavefront.
// Blocks claim segments from a counter of the launch (work): the grid holds
// more blocks than are resident at once, and blocks that start late take
// fewer segments instead of an equal share.
// Levels lo .. hi - 1 are replayed; the table holds L - lo. With lo = 0 the
// rows are written whole (-1 where unreached); with lo > 0 earlier flushes
// wrote levels below lo, so only the depths found here are stored, one row
// per step with a lane per vertex. The final pass (not a flush) also counts
// the reached vertices and edges.
template <int BS, int W>
__device__ void boutput_phase(const BArgs &a, BSmem<BS> &sm, int lo, int hi, bool final,
                              unsigned long long *work) {
  constexpr int NWV = BS / kWave;
  const int wid = threadIdx.x / kWave, lane = lane_id();
  unsigned char *d8 = sm.d8[wid];
  unsigned long long edges = 0, verts = 0;
  // Ring slot of level L in lo .. hi - 1 (hi - lo < ring): level hi - 1 is
  // the state's level, so slot(lo) follows without a division, and one wrap
  // at most.
  int s0 = sm.S.slot - (hi - 1 - lo);
  if (s0 < 0) s0 += a.ring;
  auto slot = [&](int L) {
    const int t = s0 + (L - lo);
    return t >= a.ring ? t - a.ring : t;
  };
  // Output row of each mask bit (the first block barrier of the loop below
  // orders these stores before the reads).
  for (int p = threadIdx.x; p < 64 * W; p += BS) sm.rows[p] = p < a.k ? a.rowmap[p] : 0;
  // Segments per claim and this wavefront's segment and first group.
  constexpr int SPC = W == 1 ? NWV : 1;
  const int wseg = W == 1 ? wid : 0;
  const int g0 = W == 1 ? 0 : wid;
  for (;;) {
    if (threadIdx.x == 0) sm.seg = (long long)atomicAdd(work, (unsigned long long)SPC);
    __syncthreads();
    const long long wd = sm.seg + wseg;
    __syncthreads();
    if (wd - wseg >= a.nwords) break;
    if (wd >= a.nwords) continue;
    const long long v = wd * 64 + lane;
    const bool valid = v < a.n;
    if (final && valid && g0 == 0) {
      const Mask<W> sd = load_mask<W>(a.seen, v);
      int pc = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) pc += __popcll(sd.w[w]);
      verts += (unsigned long long)pc;
      edges += (unsigned long long)pc * (unsigned long long)(a.off[v + 1] - a.off[v]);
    }
    for (int g = g0; g < W; g += NWV) {
      unsigned int *row = reinterpret_cast<unsigned int *>(d8 + lane * 68);
#pragma unroll
      for (int j = 0; j < 16; ++j) row[j] = 0xFFFFFFFFu;
      for (int L0 = lo; L0 < hi; L0 += kWave) {
        // Lane j holds the segment's joint-frontier word of level L0 + j.
        const unsigned long long fl = L0 + lane < hi ? a.jf[(long long)slot(L0 + lane) * a.nwords + wd] : 0ull;
        unsigned long long lm = __ballot(fl != 0);
        if (W == 1) {
          if (lm) replay_lanes(a, d8, fl, L0, lo, v, slot);
          continue;
        }
        // Up to kLv levels per round trip: their mask loads are issued
        // together, then replayed (a segment's wavefront visits about six
        // levels at R-MAT scale 23, one lane's vertex about two).
        while (lm) {
          constexpr int kLv = 4;
          int jj[kLv];
          unsigned long long mk[kLv];
#pragma unroll
          for (int t = 0; t < kLv; ++t) {
            jj[t] = -1;
            mk[t] = 0;
            if (lm) {
              jj[t] = __ffsll((long long)lm) - 1;
              lm &= lm - 1;
            }
          }
#pragma unroll
          for (int t = 0; t < kLv; ++t) {
            if (jj[t] >= 0) {
              const unsigned long long fw = read_lane64(fl, jj[t]);
              if (((fw >> lane) & 1ull) != 0) mk[t] = a.lv[(long long)slot(L0 + jj[t]) * a.nW + v * W + g];
            }
          }
          // Each 4 bits of a mask become 4 table bytes at once (the bits
          // of one source are set at exactly one level).
          unsigned long long any = 0;
#pragma unroll
          for (int t = 0; t < kLv; ++t) any |= mk[t];
          if (any) {
            unsigned lrep[kLv];
#pragma unroll
            for (int t = 0; t < kLv; ++t) lrep[t] = (unsigned)(L0 - lo + jj[t]) * 0x01010101u;
#pragma unroll 4
            for (int k = 0; k < 16; ++k) {
              unsigned w = row[k];
#pragma unroll
              for (int t = 0; t < kLv; ++t) {
                const unsigned msk = nibble_bytes(mk[t], k);
                w = (w & ~msk) | (lrep[t] & msk);
              }
              row[k] = w;
            }
          }
        }
      }
      wave_sync();
      const int nb = min(64, a.k - 64 * g);
      if (a.out16 && lo == 0) {
        // 16-byte streaming stores: lane l writes vertices 4(l%16) .. +3 of
        // the segment for source 4j + l/16, j = 0..15 (every row starts on a
        // 16-byte boundary: out is 16-byte aligned and n a multiple of 4).
        const int q = lane & 15, sb = lane >> 4;
        const long long v0 = wd * 64 + 4 * q;
        if (v0 < a.n) {
          for (int j = 0; j < 16; ++j) {
            const int b = 4 * j + sb;
            if (b >= nb) break;
            typedef int v4i __attribute__((ext_vector_type(4)));
            const unsigned char d0 = d8[(4 * q + 0) * 68 + b], d1 = d8[(4 * q + 1) * 68 + b];
            const unsigned char d2 = d8[(4 * q + 2) * 68 + b], d3 = d8[(4 * q + 3) * 68 + b];
            int *dst = a.out + (long long)sm.rows[64 * g + b] * a.n + v0;
            v4i x;
            x.x = d0 == 0xFF ? -1 : int(d0);
            x.y = d1 == 0xFF ? -1 : int(d1);
            x.z = d2 == 0xFF ? -1 : int(d2);
            x.w = d3 == 0xFF ? -1 : int(d3);
            __builtin_nontemporal_store(x, reinterpret_cast<v4i *>(dst));
          }
        }
      } else if (valid) {
        for (int b = 0; b < nb; ++b) {
          const unsigned char d = d8[lane * 68 + b];
          int *dst = a.out + (long long)sm.rows[64 * g + b] * a.n + v;
          if (lo == 0) *dst = d == 0xFF ? -1 : int(d);
          else if (d != 0xFF) *dst = lo + int(d);
        }
      }
      wave_sync();
    }
  }
  Mask<W> none;
#pragma unroll
  for (int w = 0; w < W; ++w) none.w[w] = 0;
  wave_counts<BS, W>(sm, 0, 0, 0, 0, none, edges, verts, word_sums_zero<W>());
}

// Launch i runs as two kernels queued back to back: the bottom-up phase has
// its own kernel (its mask registers would otherwise cap the occupancy of
// every phase), the other phases share one. Both derive the same state for
// launch i; the one whose phase it is does the work. When the previous
// search on the workspace never pulled (W <= 4), the host queues k_blevel
// alone, with pulls = 1, and a pull runs there.
template <int BS, int W>
__global__ __launch_bounds__(BS) void k_blevel(BArgs a, long long i, int pulls) {
  __shared__ BSmem<BS> sm;
  begin_blaunch<BS, W>(a, sm, i);
  const BState &S = sm.S;
  if (S.done) return;
  const int slot = int(i % 3);
  if (S.phase == BPH_BU) {
    // With pulls set (W <= 4 only) the host queued no k_blevel_bu for this
    // launch: the pull runs here, at this kernel's occupancy.
    if (W <= 4 && pulls) {
      bbu_phase<BS, W>(a, sm, S.level);
      block_counts_flush(a, sm, slot);
    }
    return;
  }
  switch (S.phase) {
    case BPH_TD: btd_phase<BS, W, false>(a, sm, S.level); break;
    case BPH_TDF:
      if (W <= 4) btd_phase<BS, W, true>(a, sm, S.level);
      break;
    case BPH_APPLY: bapply_phase<BS, W>(a, sm, S.level); break;
    default:
      boutput_phase<BS, W>(a, sm, S.base, S.flush ? S.flush_end : S.levels, !S.flush, &a.ctl->cnt[slot].v[C_WORK]);
      break;
  }
  if (S.phase != BPH_TD) block_counts_flush(a, sm, slot);
}

template <int BS, int W>
__global__ __launch_bounds__(BS) void k_blevel_bu(BArgs a, long long i) {
  // k_blevel ran launch i just before on the same stream and recorded its
  // phase: unless it is a pull, leave without reading the state and counters.
  if (a.ctl->lphase[i & 1] != BPH_BU) return;
  __shared__ BSmem<BS> sm;
  begin_blaunch<BS, W>(a, sm, i);
  const BState &S = sm.S;
  if (S.done || S.phase != BPH_BU) return;
  bbu_phase<BS, W>(a, sm, S.level);
  block_counts_flush(a, sm, int(i % 3));
}

// The whole multi-source search in one cooperative launch: launches 0, 1,
// 2, ... of the per-phase form separated by grid barriers. Stops on
// completion, or when a barrier wait exceeds ~4 s in any block (the search is
// then marked failed). Exit decisions are taken from LDS after a block
// barrier, so a block leaves the loop as a whole.
// Wavefronts per SIMD the persistent kernel is compiled for (its register
// budget guarantees at least this many); the cooperative grid is sized from
// the same figure, so every block is resident.
constexpr int bpersist_waves_per_simd(int W) { return W >= 8 ? 3 : 4; }

template <int BS, int W>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(bpersist_waves_per_simd(W)))) void k_bpersist(
    BArgs a) {
  __shared__ BSmem<BS> sm;
  __shared__ int stop;
  for (long long i = 0;; ++i) {
    begin_blaunch<BS, W>(a, sm, i, true);
    const BState &S = sm.S;
    if (S.done || (S.error & BERR_TIMEOUT)) break;
    const int slot = int(i % 3);
    switch (S.phase) {
      case BPH_TD: btd_phase<BS, W, false>(a, sm, S.level); break;
      case BPH_TDF:
        if (W <= 4) btd_phase<BS, W, true>(a, sm, S.level);
        break;
      case BPH_APPLY: bapply_phase<BS, W>(a, sm, S.level); break;
      case BPH_BU: bbu_phase<BS, W>(a, sm, S.level); break;
      default:
        boutput_phase<BS, W>(a, sm, S.base, S.flush ? S.flush_end : S.levels, !S.flush,
                             &a.ctl->cnt[slot].v[C_WORK]);
        break;
    }
    if (S.phase != BPH_TD) block_counts_flush(a, sm, slot);
    const bool ok = grid_sync(&a.ctl->sync, 400000000ull);
    if (threadIdx.x == 0) {
      if (!ok) {
        atomicOr(&a.ctl->st[0].error, BERR_TIMEOUT);
        atomicOr(&a.ctl->st[1].error, BERR_TIMEOUT);
      }
      stop = !ok || (ld_agent(&a.ctl->st[0].error) & BERR_TIMEOUT) != 0;
    }
    __syncthreads();
    if (stop) break;
  }
}

// Zeroes seen, the level-0 joint frontier, lv[0] at the roots (the only
// vertices where it is read) and lv[1] where level 0 can push; clears the
// counters and the grid barrier; sets the pseudo-level -1 state. With fused
// levels, also empties the slots of levels 0 and 1 whole (prep_word; the
// level-1 masks are cleared at every vertex a push can reach).
__global__ void k_binit_zero(BArgs a, const int *roots) {
  const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nt = (long long)gridDim.x * blockDim.x;
  const int W = int(a.nW / a.n);
  const int sh = W == 1 ? 0 : W == 2 ? 1 : W == 4 ? 2 : 3;
  unsigned long long *l1 = a.lv + a.nW;
  for (long long i = gt; i < a.nW; i += nt) {
    a.seen[i] = 0;
    const long long v = i >> sh;
    if (receives(a, a.off[v + 1] - a.off[v])) l1[i] = 0;
  }
  for (long long i = gt; i < (long long)a.k * W; i += nt) a.lv[(long long)roots[i / W] * W + i % W] = 0;
  for (long long w = gt; w < a.nwords; w += nt) {
    if (a.fused > 0) {
      for (unsigned long long m = a.jf[w]; m; m &= m - 1) {
        const long long v = w * 64 + (__ffsll((long long)m) - 1);
        for (int j = 0; j < W; ++j) a.lv[v * W + j] = 0;
      }
      a.jf[a.nwords + w] = 0;
    }
    a.jf[w] = 0;
  }
  constexpr long long kC = (long long)kBStripes * kBStripeU64;
  for (long long i = gt; i < 3 * kC; i += nt) a.ctl->cnt[i / kC].v[i % kC] = 0;
  for (long long i = gt; i < kSyncStripes * kSyncStride; i += nt) a.ctl->sync.arrive[i] = 0;
  if (gt == 0) {
    a.ctl->sync.arrive_all = 0;
    a.ctl->sync.gen = 0;
  }
  if (gt == 0) {
    BState S{};
    S.level = -1;
    S.slot = a.ring - 1;
    S.phase = BPH_APPLY;
    a.ctl->st[0] = S;
    a.ctl->st[1] = S;
  }
}

// Level-0 counts computed on the host (frontier 0 = the roots).
struct BInit {
  unsigned long long jw, mf, alg, munf;
  unsigned long long mfw[8];  // degree sum of the distinct roots of word w
};

// Sets the source bits (level 0) and the level-0 counts.
__global__ void k_binit_src(BArgs a, const int *roots, BInit in) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < a.k) {
    const int v = roots[s];
    const int W = int(a.nW / a.n);
    const unsigned long long bit = 1ull << (s & 63);
    atomicOr(&a.lv[(long long)v * W + (s >> 6)], bit);
    atomicOr(&a.seen[(long long)v * W + (s >> 6)], bit);
    atomicOr(&a.jf[v >> 6], 1ull << (v & 63));
    atomicOr(&a.ctl->cnt[2].v[C_ACTIVE + (s >> 6)], bit);
  }
  if (s == 0) {
    unsigned long long *c = a.ctl->cnt[2].v;
    c[C_JWORDS] = in.jw;
    c[C_MF] = in.mf;
    c[C_ALG] = in.alg;
    c[C_MUNF] = in.munf;
    for (int w = 0; w < 8; ++w) c[C_MFW + w] = in.mfw[w];
  }
}

// Position-keyed checksum of each row (sum over i of mix(i, value), mod
// 2^64). Block (r, c) sums chunk c of row r and adds it to out[r].
__global__ void k_row_hash(const int *dev, long long cols, long long chunk, unsigned long long *out) {
  const long long r = blockIdx.y;
  const long long c0 = (long long)blockIdx.x * chunk;
  const long long c1 = min(cols, c0 + chunk);
  unsigned long long h = 0;
  for (long long i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
    unsigned long long z = ((unsigned long long)i << 32) ^ (unsigned int)dev[r * cols + i];
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h += z ^ (z >> 31);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  __shared__ unsigned long long red[4];
  if (lane_id() == 0) red[threadIdx.x / kWave] = h;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&out[r], red[0] + red[1] + red[2] + red[3]);
}

// Scratch of ent_device_row_hash_noalloc: one checksum accumulator per row,
// ENT_NOALLOC_HASH_ROWS rows. A file-scope device array: each GPU the library
// runs on holds its own copy until the process ends, at its own address
// (looked up once per device with hipGetSymbolAddress; host code never uses
// the name itself as a device address).
static __device__ __align__(256) unsigned long long scratch[ENT_NOALLOC_HASH_ROWS];

// Clears the first n words of p (a kernel, not hipMemsetAsync, so the
// entry point adds no memset node when a caller's stream is captured).
__global__ void k_zero_u64(unsigned long long *p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

struct BatchWorkspace {
  int words = 0;      // W (1, 2, 4 or 8)
  int lcap = 0;       // levels kept between flushes (ring of lcap + 3 slots)
  int64_t n = 0;
  unsigned long long *seen = nullptr;
  unsigned long long *lv = nullptr;
  unsigned long long *jf = nullptr;  // lcap + 3 bitmaps
  int *roots = nullptr;  // roots by mask bit, then the output row of each bit (2 x 512)
  BControl *ctl = nullptr;
  BControl *h_ctl = nullptr;
  size_t bytes = 0;
  int expect_launches = 16;
  int *rows = nullptr;     // device rows of ent_bfs_batch (grown on demand)
  size_t rows_bytes = 0;
  // The last search pulled at some level: queue k_blevel_bu beside every
  // launch (otherwise, with W <= 4, k_blevel runs a pull itself).
  bool pulled_last = true;
  // Every nonzero mask in the lv ring belongs to a vertex in its slot's
  // bitmap (what fused levels rely on). Holds after the rings are zeroed and
  // after a search that completed with fused levels allowed; a search that
  // did not, or failed, may leave masks the bitmaps do not list.
  bool rings_clean = false;
  int clean_words = 0;  // the mask width of the layout rings_clean describes
};

}  // namespace ent

using namespace ent;

namespace {

constexpr int kBatchBS = 256;
constexpr int kDefaultLevels = 32;

int words_for(int k) { return k <= 64 ? 1 : k <= 128 ? 2 : k <= 256 ? 4 : 8; }

size_t ws_bytes(int64_t n, int W, int lcap) {
  const size_t nW = size_t(n) * size_t(W);
  const size_t nwords = size_t((n + 63) / 64);
  return 8 * nW + (8 * nW + 8 * nwords) * size_t(lcap + 3) + sizeof(BControl) + 8 * ENT_BATCH_MAX_SOURCES;
}

int reserve(ent_graph *g, int k, int lcap) {
  if (k < 1 || k > ENT_BATCH_MAX_SOURCES) return set_error(ENT_EINVAL, "k=%d outside [1, 512]", k);
  if (lcap < 1 || lcap > 250) return set_error(ENT_EINVAL, "max_levels=%d outside [1, 250]", lcap);
  const int W = words_for(k);
  BatchWorkspace *b = g->batch;
  if (b && b->words >= W && b->lcap >= lcap) return ENT_OK;
  if (b) {
    batch_release(g);
    b = nullptr;
  }
  b = new (std::nothrow) BatchWorkspace();
  if (!b) return set_error(ENT_ENOMEM, "out of host memory");
  b->words = W;
  b->lcap = lcap;
  b->n = g->n;
  const size_t nW = size_t(g->n) * size_t(W);
  int rc;
  g->batch = b;
  if ((rc = dev_alloc(g, reinterpret_cast<void **>(&b->seen), 8 * nW)) ||
      (rc = dev_alloc(g, reinterpret_cast<void **>(&b->lv), 8 * nW * size_t(lcap + 3))) ||
      (rc = dev_alloc(g, reinterpret_cast<void **>(&b->jf), 8 * size_t(g->nwords) * size_t(lcap + 3))) ||
      (rc = dev_alloc(g, reinterpret_cast<void **>(&b->roots), 8 * ENT_BATCH_MAX_SOURCES)) ||
      (rc = dev_alloc(g, reinterpret_cast<void **>(&b->ctl), sizeof(BControl)))) {
    batch_release(g);
    return rc;
  }
  if (hipHostMalloc(reinterpret_cast<void **>(&b->h_ctl), sizeof(BControl), hipHostMallocDefault) != hipSuccess) {
    batch_release(g);
    return set_error(ENT_ENOMEM, "pinned host memory");
  }
  b->bytes = ws_bytes(g->n, W, lcap);
  return ENT_OK;
}

struct BKernels {
  const void *gen, *bu, *persist;
};

template <int W>
BKernels blevel_kernels() {
  return {reinterpret_cast<const void *>(&k_blevel<kBatchBS, W>),
          reinterpret_cast<const void *>(&k_blevel_bu<kBatchBS, W>),
          reinterpret_cast<const void *>(&k_bpersist<kBatchBS, W>)};
}

BKernels blevel_for(int W) {
  switch (W) {
    case 1: return blevel_kernels<1>();
    case 2: return blevel_kernels<2>();
    case 4: return blevel_kernels<4>();
    default: return blevel_kernels<8>();
  }
}

int run_batch(ent_graph *g, const int32_t *roots, int k, int *out, BState *fin) {
  BatchWorkspace *b = g->batch;
  // The search's own width: a workspace reserved for more sources holds the
  // narrower layout (every array is n * W words per level slot or less), and
  // a batch of 64 on a 512-source workspace costs what it costs on its own.
  const int W = words_for(k);
  BArgs a{};
  a.off = g->off;
  a.adj = g->adj;
  a.hub_block = g->hub_block;
  a.nz = g->nzmask;
  a.piece_v = g->piece_v;
  a.piece_k = g->piece_k;
  a.seen = b->seen;
  a.lv = b->lv;
  a.jf = b->jf;
  a.out = out;
  a.ctl = b->ctl;
  a.m = g->m;
  a.nW = g->n * W;
  a.n = int(g->n);
  a.nwords = int(g->nwords);
  a.k = k;
  a.lcap = b->lcap;
  a.ring = b->lcap + 3;
  a.n_hub_block = g->n_hub_block;
  a.n_pieces = g->n_pieces;
  const ent_bfs_config &c = g->cfg;
  a.t_wave = c.t_wave;
  a.t_block = c.t_block;
  a.piece = c.piece;
  a.mode = c.direction;
  a.alpha = c.batch_alpha > 0 ? c.batch_alpha : 4;
  a.probe = c.bu_probe;
  a.chunk_words = c.bu_chunk_words;
  a.out16 = ((reinterpret_cast<uintptr_t>(out) & 15) == 0 && (g->n & 3) == 0) ? 1 : 0;
  a.fused = (W <= 4 && c.batch_fused > 0) ? std::max<long long>(1, (long long)g->n / c.batch_fused) : 0;
  // Sources are placed at mask bits in descending order of root degree
  // (ties by vertex id), so each 64-source word holds roots whose frontiers
  // grow alike and the per-word direction can push the lagging words while it
  // pulls the others, whatever order the caller gave; the same set of roots
  // always gets the same words, and so the same per-word counts.
  // placed[0 .. k) are the roots by bit, placed[512 + p] the caller's index
  // of bit p (its output row).
  std::vector<int32_t> placed(2 * ENT_BATCH_MAX_SOURCES, 0);
  {
    std::vector<int> order(static_cast<size_t>(k));
    for (int s = 0; s < k; ++s) order[size_t(s)] = s;
    const int32_t *deg = g->store->h_deg.data();
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const int32_t dx = deg[size_t(roots[x])], dy = deg[size_t(roots[y])];
      return dx != dy ? dx > dy : roots[x] < roots[y];
    });
    for (int p = 0; p < k; ++p) {
      placed[size_t(p)] = roots[order[size_t(p)]];
      placed[size_t(ENT_BATCH_MAX_SOURCES + p)] = order[size_t(p)];
    }
  }
  a.rowmap = b->roots + ENT_BATCH_MAX_SOURCES;
  // Level-0 counts on the host: distinct (vertex, word) pairs of the roots.
  std::vector<std::pair<int32_t, int>> vw;
  std::vector<int32_t> vs;
  for (int p = 0; p < k; ++p) {
    vw.push_back({placed[size_t(p)], p >> 6});
    vs.push_back(placed[size_t(p)]);
  }
  std::sort(vw.begin(), vw.end());
  vw.erase(std::unique(vw.begin(), vw.end()), vw.end());
  std::sort(vs.begin(), vs.end());
  vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
  BInit init{};
  for (int32_t v : vs) init.mf += (unsigned long long)g->store->h_deg[size_t(v)];
  for (auto &p : vw) {
    const unsigned long long d = (unsigned long long)g->store->h_deg[size_t(p.first)];
    init.alg += 8ull + 12ull * d;
    init.mfw[p.second] += d;
  }
  init.jw = vw.size();
  init.munf = (unsigned long long)g->m;  // every vertex still needs sources
  HIP_TRY(hipMemcpyAsync(b->roots, placed.data(), sizeof(int32_t) * placed.size(), hipMemcpyHostToDevice, g->stream));
  const int grid = g->cus * 8;
  const bool was_clean = b->rings_clean && b->clean_words == W;  // same layout
  b->rings_clean = false;  // until this search completes
  b->clean_words = W;
  if (a.fused > 0 && !was_clean) {
    const long long nlv = a.nW * (long long)a.ring, njf = (long long)a.nwords * (long long)a.ring;
    hipLaunchKernelGGL(k_zero_u64, dim3(grid), dim3(256), 0, g->stream, b->lv, nlv);
    hipLaunchKernelGGL(k_zero_u64, dim3(grid), dim3(256), 0, g->stream, b->jf, njf);
    HIP_TRY(hipGetLastError());
  }
  {
    void *args[] = {&a, &b->roots};
    HIP_TRY(hipLaunchKernel(reinterpret_cast<const void *>(&k_binit_zero), dim3(grid), dim3(256), args, 0, g->stream));
  }
  {
    void *args[] = {&a, &b->roots, &init};
    HIP_TRY(hipLaunchKernel(reinterpret_cast<const void *>(&k_binit_src), dim3((k + 255) / 256), dim3(256), args, 0,
                            g->stream));
  }
  const BKernels kl = blevel_for(W);
  if (c.launch == ENT_LAUNCH_PERSISTENT) {
    // Every block must be resident: no more wavefronts per CU than the
    // kernel's register budget guarantees (4 SIMDs x bpersist_waves_per_simd
    // wavefronts: 3 blocks of 256 at W = 8, 4 below), at most 8 workgroups,
    // and within the occupancy the runtime reports.
    int occ = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kl.persist, kBatchBS, 0));
    const int wave_cap = 4 * bpersist_waves_per_simd(W) / (kBatchBS / kWave);
    int per_cu = std::min(std::min(occ, wave_cap), 8);
    if (c.blocks_per_cu > 0) per_cu = std::min(per_cu, c.blocks_per_cu);
    if (per_cu < 1) return set_error(ENT_EDEVICE, "persistent multi-source kernel cannot be resident");
    int pgrid = g->cus * per_cu;
    if (c.grid_blocks > 0) pgrid = std::min(pgrid, c.grid_blocks);
    void *args[] = {&a};
    HIP_TRY(hipLaunchCooperativeKernel(kl.persist, dim3(pgrid), dim3(kBatchBS), args, 0, g->stream));
    HIP_TRY(hipMemcpyAsync(b->h_ctl, b->ctl, sizeof(BState) * 2, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const BState &s0 = b->h_ctl->st[0], &s1 = b->h_ctl->st[1];
    *fin = s0.phases >= s1.phases ? s0 : s1;
    fin->error |= (s0.error | s1.error) & BERR_TIMEOUT;
    g->last_grid = pgrid;
    b->rings_clean = a.fused > 0 && fin->done && !fin->error;
    return ENT_OK;
  }
  const int lgrid = c.grid_blocks > 0 ? c.grid_blocks : g->cus * std::max(1, c.blocks_per_cu > 0 ? c.blocks_per_cu : 8);
  long long i = 0;
  int kk = std::max(2, b->expect_launches);
  // A search like the last one that never pulled queues one kernel per
  // launch; a pull it does meet runs in k_blevel (W <= 4). W = 8 always
  // queues the pull kernel, whose registers k_blevel lacks.
  const bool pull_kernel = W > 4 || b->pulled_last;
  int pulls = pull_kernel ? 0 : 1;
  // At most two launches per level (top-down, then pull or apply), one flush
  // per lcap + 1 levels and the output; levels <= n.
  const long long max_launches = 4LL * ((long long)g->n + 4);
  for (;;) {
    for (int j = 0; j < kk; ++j, ++i) {
      void *gargs[] = {&a, &i, &pulls};
      void *args[] = {&a, &i};
      HIP_TRY(hipLaunchKernel(kl.gen, dim3(lgrid), dim3(kBatchBS), gargs, 0, g->stream));
      if (pull_kernel) HIP_TRY(hipLaunchKernel(kl.bu, dim3(lgrid), dim3(kBatchBS), args, 0, g->stream));
    }
    HIP_TRY(hipMemcpyAsync(b->h_ctl, b->ctl, sizeof(BState) * 2, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *fin = b->h_ctl->st[(i - 1) & 1];
    if (fin->done) break;
    if (i > max_launches) return set_error(ENT_EINTERNAL, "batch did not finish in %lld launches", i);
    kk = std::min(2 * kk, 512);
  }
  b->expect_launches = std::min(fin->phases + 1, 512);
  b->pulled_last = fin->pulled != 0;
  g->last_grid = lgrid;
  b->rings_clean = a.fused > 0 && !fin->error;
  return ENT_OK;
}

}  // namespace

namespace ent {
void batch_release(ent_graph *g) {
  BatchWorkspace *b = g ? g->batch : nullptr;
  if (!b) return;
  const size_t nW = size_t(b->n) * size_t(b->words);
  dev_free(g, b->seen, 8 * nW);
  dev_free(g, b->lv, 8 * nW * size_t(b->lcap + 3));
  dev_free(g, b->jf, 8 * size_t(g->nwords) * size_t(b->lcap + 3));
  dev_free(g, b->roots, 8 * ENT_BATCH_MAX_SOURCES);
  dev_free(g, b->ctl, sizeof(BControl));
  dev_free(g, b->rows, b->rows_bytes);
  if (b->h_ctl) ignore(hipHostFree(b->h_ctl));
  delete b;
  g->batch = nullptr;
}
}  // namespace ent

namespace {

// Multi-source search with g->mu held. Any depth fits the reserved
// workspace: the level ring flushes kept levels to the rows as it goes
// (st->flushes).
int batch_locked(ent_graph *g, const int32_t *roots, int k, int32_t *depths_dev, ent_batch_stats *st,
                 std::chrono::steady_clock::time_point t0) {
  HIP_TRY(hipSetDevice(g->device));
  int rc = reserve(g, k, g->batch ? g->batch->lcap : kDefaultLevels);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  BState fin{};
  rc = run_batch(g, roots, k, depths_dev, &fin);
  if (rc) {
    ignore(hipStreamSynchronize(g->stream));
    return rc;
  }
  if (fin.error & BERR_TIMEOUT) {
    ignore(hipStreamSynchronize(g->stream));
    return set_error(ENT_EDEVICE, "persistent multi-source search: grid barrier wait timed out");
  }
  if (fin.error) {
    ignore(hipStreamSynchronize(g->stream));
    return set_error(ENT_EINTERNAL, "device reported error flags 0x%x", fin.error);
  }
  const unsigned long long edges = fin.total_edges, verts = fin.total_vertices, alg = fin.alg_bytes;
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventSynchronize(g->ev1));
  const auto t1 = std::chrono::steady_clock::now();
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->k = k;
    st->words = words_for(k);
    st->levels = fin.levels;
    st->phases = fin.phases;
    st->reached_edges = int64_t(edges);
    st->reached_vertices = int64_t(verts);
    st->alg_bytes = int64_t(alg);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->ev0, g->ev1) == hipSuccess) st->device_ms = ms;
    st->host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    st->bu_level_mask = fin.bu_mask;
    st->td_level_mask = fin.td_mask;
    st->flushes = fin.flushes;
  }
  return ENT_OK;
}

}  // namespace

extern "C" {

int64_t ent_batch_workspace_bytes(const ent_graph *g, int32_t k, int32_t max_levels) {
  if (!g || k < 1 || k > ENT_BATCH_MAX_SOURCES) return -1;
  return int64_t(ws_bytes(g->n, words_for(k), max_levels > 0 ? max_levels : kDefaultLevels));
}

int ent_batch_reserve(ent_graph *g, int32_t k, int32_t max_levels) {
  clear_error();
  if (!g) return set_error(ENT_EINVAL, "NULL graph");
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(hipStreamSynchronize(g->stream));
  const int lcap = max_levels > 0 ? max_levels : (g->batch ? g->batch->lcap : kDefaultLevels);
  return reserve(g, k, lcap);
}

int ent_batch_release(ent_graph *g) {
  clear_error();
  if (!g) return set_error(ENT_EINVAL, "NULL graph");
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(hipStreamSynchronize(g->stream));
  batch_release(g);
  return ENT_OK;
}

int ent_bfs_batch_device(ent_graph *g, const int32_t *roots, int32_t k, int32_t *depths_dev,
                         ent_batch_stats *st) {
  clear_error();
  if (!g || !roots || !depths_dev) return set_error(ENT_EINVAL, "NULL argument");
  if (k < 1 || k > ENT_BATCH_MAX_SOURCES) return set_error(ENT_EINVAL, "k=%d outside [1, 512]", k);
  if (reinterpret_cast<uintptr_t>(depths_dev) & 3) return set_error(ENT_EINVAL, "depths_dev not 4-byte aligned");
  for (int s = 0; s < k; ++s)
    if (roots[s] < 0 || roots[s] >= g->n) return set_error(ENT_EINVAL, "root %d = %d out of range", s, roots[s]);
  std::lock_guard<std::mutex> lk(g->mu);
  return batch_locked(g, roots, k, depths_dev, st, std::chrono::steady_clock::now());
}

int ent_bfs_batch(ent_graph *g, const int32_t *roots, int32_t k, int32_t *depths_host, ent_batch_stats *st) {
  clear_error();
  if (!g || !roots || !depths_host) return set_error(ENT_EINVAL, "NULL argument");
  if (k < 1 || k > ENT_BATCH_MAX_SOURCES) return set_error(ENT_EINVAL, "k=%d outside [1, 512]", k);
  for (int s = 0; s < k; ++s)
    if (roots[s] < 0 || roots[s] >= g->n) return set_error(ENT_EINVAL, "root %d = %d out of range", s, roots[s]);
  std::lock_guard<std::mutex> lk(g->mu);
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(g->device));
  int rc = reserve(g, k, g->batch ? g->batch->lcap : kDefaultLevels);
  if (rc) return rc;
  // Device rows kept with the workspace and reused by later calls.
  BatchWorkspace *b = g->batch;
  const size_t bytes = sizeof(int32_t) * size_t(k) * size_t(g->n);
  if (b->rows_bytes < bytes) {
    HIP_TRY(hipStreamSynchronize(g->stream));
    dev_free(g, b->rows, b->rows_bytes);
    b->rows = nullptr;
    b->rows_bytes = 0;
    if ((rc = dev_alloc(g, reinterpret_cast<void **>(&b->rows), bytes))) return rc;
    b->rows_bytes = bytes;
  }
  rc = batch_locked(g, roots, k, b->rows, st, t0);
  if (!rc) rc = copy_to_host(g, depths_host, b->rows, bytes);
  return rc;
}

int ent_device_alloc(ent_graph *g, int64_t bytes, void **out) {
  clear_error();
  if (!g || !out || bytes < 0) return set_error(ENT_EINVAL, "bad argument");
  HIP_TRY(hipSetDevice(g->device));
  *out = nullptr;
  hipError_t e = hipMalloc(out, size_t(std::max<int64_t>(bytes, 8)));
  if (e != hipSuccess) {
    ignore(hipGetLastError());
    return set_error(ENT_ENOMEM, "hipMalloc(%lld) failed: %s", (long long)bytes, hipGetErrorString(e));
  }
  return ENT_OK;
}

int ent_device_free(ent_graph *g, void *p) {
  // Keeps the previous error text: callers free buffers on their error paths.
  if (!g) return set_error(ENT_EINVAL, "NULL graph");
  if (!p) return ENT_OK;
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(hipFree(p));
  return ENT_OK;
}

int ent_device_row_hash(ent_graph *g, const int32_t *dev, int64_t rows, int64_t cols, uint64_t *hashes_host) {
  clear_error();
  if (!g || !dev || !hashes_host || rows < 1 || cols < 1) return set_error(ENT_EINVAL, "bad argument");
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  // The handle's own checksum buffer, kHashRows rows per launch: no
  // allocation, so calls on other handles' streams keep running meanwhile.
  unsigned long long *d = g->hashbuf;
  long long chunk = 1 << 16;
  const long long nchunks = (cols + chunk - 1) / chunk;
  for (int64_t r0 = 0; r0 < rows; r0 += kHashRows) {
    const int64_t nr = std::min<int64_t>(kHashRows, rows - r0);
    const int32_t *base = dev + r0 * cols;
    void *args[] = {const_cast<int32_t **>(&base), &cols, &chunk, &d};
    HIP_TRY(hipMemsetAsync(d, 0, sizeof(unsigned long long) * size_t(nr), g->stream));
    HIP_TRY(hipLaunchKernel(reinterpret_cast<const void *>(&k_row_hash), dim3(unsigned(nchunks), unsigned(nr)),
                            dim3(256), args, 0, g->stream));
    HIP_TRY(hipMemcpyAsync(hashes_host + r0, d, sizeof(unsigned long long) * size_t(nr), hipMemcpyDeviceToHost,
                           g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  return ENT_OK;
}

}  // extern "C"

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];         // one call at a time per device
unsigned long long *g_scratch[kMaxDevices];  // device address of `scratch` per device

// Device address of this device's `scratch`, looked up on the first call on
// the device (the caller has made it current and holds its lock).
int scratch_of(int device, unsigned long long **out) {
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ignore(hipGetLastError());
      return set_error(ENT_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_scratch[device] = static_cast<unsigned long long *>(p);
  }
  *out = g_scratch[device];
  return ENT_OK;
}

}  // namespace

extern "C" int ent_device_row_hash_noalloc(int32_t device, const int32_t *dev, int64_t rows, int64_t cols,
                                           uint64_t *hashes_host) {
  clear_error();
  if (!dev || !hashes_host || rows < 1 || cols < 1) return set_error(ENT_EINVAL, "bad argument");
  if (rows > ENT_NOALLOC_HASH_ROWS)
    return set_error(ENT_EINVAL, "rows=%lld exceeds the %d rows of the no-allocation scratch", (long long)rows,
                     ENT_NOALLOC_HASH_ROWS);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess) {
    ignore(hipGetLastError());
    return set_error(ENT_EDEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev || device >= kMaxDevices)
    return set_error(ENT_EINVAL, "device %d of %d", device, ndev);
  std::lock_guard<std::mutex> lk(g_scratch_mu[device]);
  HIP_TRY(hipSetDevice(device));
  unsigned long long *acc = nullptr;
  const int rc = scratch_of(device, &acc);
  if (rc) return rc;
  // The accumulators keep the previous call's sums: clear the ones this call
  // uses, then sum and copy on the null stream (the copy waits for both).
  const long long nr = rows;
  hipLaunchKernelGGL(k_zero_u64, dim3(unsigned((nr + 255) / 256)), dim3(256), 0, 0, acc, nr);
  HIP_TRY(hipGetLastError());
  long long chunk = 1 << 16;
  const long long nchunks = (cols + chunk - 1) / chunk;
  long long ncols = cols;
  void *args[] = {const_cast<int32_t **>(&dev), &ncols, &chunk, &acc};
  HIP_TRY(hipLaunchKernel(reinterpret_cast<const void *>(&k_row_hash), dim3(unsigned(nchunks), unsigned(nr)),
                          dim3(256), args, 0, nullptr));
  HIP_TRY(hipMemcpy(hashes_host, acc, sizeof(unsigned long long) * size_t(nr), hipMemcpyDeviceToHost));
  return ENT_OK;
}

