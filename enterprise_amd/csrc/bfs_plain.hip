// Single-source direction-optimizing BFS for gfx950 (wave64) behind the C ABI
// of ent_bfs.h.
//
// Reference behaviour restated (iHeartGraph/Enterprise):
//  - level loop, direction state and termination: bfs_tdbu_clfy_sort,
//    bfs_gpu_opt.cuh:40-288; termination when a level finds nothing
//    (:193-202);
//  - top-down expansion with degree classes: THD/WAP/CTA_expand_sort,
//    expander.cuh:77-423 (warp32 classes, thresholds 32 / 1024 from
//    bfs_gpu_opt.cuh:323-324) -> here thread / wavefront / block / grid
//    classes sized for 64-lane wavefronts;
//  - bottom-up expansion with early exit at the first frontier neighbour:
//    THD/WAP/CTA_bu_expand_sort, expander.cuh:434-819;
//  - status-array inspection and classification into class queues:
//    inspector.cuh:38-432 plus the bin-offset scans of scan.cuh:485-685 ->
//    here ballot/DPP ranks and LDS staging, no per-thread bins, no reaper;
//  - direction switch: inspector.cuh:486-499 (0.3 * m, one way), offered as
//    ENT_DIR_REFERENCE beside a Beamer alpha/beta rule.
//
// Device state (all in HBM, sized at upload):
//   off int64[n+1], adj int32[m], nzmask u64[nwords] (bit = degree > 0),
//   depth int32[n], visited u64[nwords], front[2] u64[nwords],
//   class queues q[2][3] int32[n] (thread / wave / block class) and grid
//   pieces qg[2] int64 ((v << 32) | piece index), Control (level state and
//   three rotating per-level counter slots).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ent_bfs.h"
#include "ent_common.h"
#include "ent_device.h"

namespace ent {

enum Phase : int { PH_TD = 0, PH_BU = 1, PH_INSPECT = 2 };
enum : int { ERR_NONE = 0, ERR_OVERFLOW = 1, ERR_LEVELS = 2, ERR_TIMEOUT = 4 };

struct LevelSlot {
  unsigned int qsize[4];  // class queue sizes of the frontier (class 3: pieces)
  unsigned int nf;        // frontier vertices
  unsigned int pad;
  unsigned long long mf;  // sum of frontier degrees
};

struct Control {
  int level;
  int phase;
  int done;
  int error;
  int bitmap_valid;  // frontier of `level` is stored in front[level & 1]
  int levels;
  int td_levels;
  int bu_levels;
  int phases;
  unsigned int arrive;
  unsigned int gen;
  int pad;
  unsigned long long total_nf;
  unsigned long long total_mf;
  unsigned long long prev_nf;
  unsigned long long bu_mask;
  LevelSlot slot[3];
};

struct Args {
  const long long *off;
  const int *adj;
  const unsigned long long *nzmask;
  int *depth;
  unsigned long long *visited;
  unsigned long long *front[2];
  int *q[2][3];
  long long *qg[2];
  Control *ctl;
  long long m;
  long long n_active;
  long long cap3;
  int n;
  int nwords;
  int t_wave, t_block, t_grid, piece;
  int mode, alpha, beta;
  int bu_probe;
  int bu_chunk_words;
};

template <int BS>
struct Smem {
  static constexpr int NW = BS / kWave;
  static constexpr int S0 = 2048, S1 = 512, S2 = 128;
  union {
    struct {
      int buf0[S0];
      int buf1[S1];
      int buf2[S2];
    } td;
    struct {
      unsigned long long cand[NW][kWave];
      unsigned long long found[NW][kWave];
      int excl[NW][kWave];
    } bu;
  } u;
  int cnt[3];
  unsigned int base[3];
  unsigned long long nf, mf;
  int L, ph, done, bm;
};

__device__ __forceinline__ bool in_front(const Args &a, const unsigned long long *fr, bool bm, int u,
                                         int L) {
  return bm ? ((fr[u >> 6] >> (u & 63)) & 1ull) != 0 : a.depth[u] == L;
}

// Appends v (degree e - s) to the class queues of parity `par`, counting it
// in `slot`. Classes 0..2 are staged in LDS and flushed per block; grid-class
// vertices are split into pieces written straight to the global piece queue.
template <int BS>
__device__ __forceinline__ void push_vertex(const Args &a, Smem<BS> &sm, int v, long long s, long long e,
                                            int par, LevelSlot *slot) {
  const long long d = e - s;
  atomicAdd(&sm.nf, 1ull);
  atomicAdd(&sm.mf, (unsigned long long)d);
  if (d >= a.t_grid) {
    const unsigned np = unsigned((d + a.piece - 1) / a.piece);
    const unsigned base = atomicAdd(&slot->qsize[3], np);
    // Every reserved index below cap3 is written, so a reader clamped to cap3
    // never sees an unwritten entry.
    for (unsigned k = 0; k < np && (long long)base + k < a.cap3; ++k)
      a.qg[par][base + k] = ((long long)v << 32) | k;
    if ((long long)base + np > a.cap3) atomicOr(&a.ctl->error, ERR_OVERFLOW);
    return;
  }
  const int c = d >= a.t_block ? 2 : (d >= a.t_wave ? 1 : 0);
  const int i = atomicAdd(&sm.cnt[c], 1);
  int *buf = c == 0 ? sm.u.td.buf0 : (c == 1 ? sm.u.td.buf1 : sm.u.td.buf2);
  const int cap = c == 0 ? Smem<BS>::S0 : (c == 1 ? Smem<BS>::S1 : Smem<BS>::S2);
  if (i < cap) {
    buf[i] = v;
  } else {
    const unsigned g = atomicAdd(&slot->qsize[c], 1u);
    if (g < unsigned(a.n)) a.q[par][c][g] = v;
    else atomicOr(&a.ctl->error, ERR_OVERFLOW);
  }
}

// Moves the block's LDS-staged queue entries to global memory with one
// reservation per class, and adds the block's vertex / degree counts.
template <int BS>
__device__ void flush_push(const Args &a, Smem<BS> &sm, int par, LevelSlot *slot, bool count) {
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    const int cap = c == 0 ? Smem<BS>::S0 : (c == 1 ? Smem<BS>::S1 : Smem<BS>::S2);
    const int k = min(sm.cnt[c], cap);
    sm.base[c] = k ? atomicAdd(&slot->qsize[c], unsigned(k)) : 0u;
  } else if (threadIdx.x == 3 && count && sm.nf) {
    atomicAdd(&slot->nf, unsigned(sm.nf));
    atomicAdd(&slot->mf, sm.mf);
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int cap = c == 0 ? Smem<BS>::S0 : (c == 1 ? Smem<BS>::S1 : Smem<BS>::S2);
    const int k = min(sm.cnt[c], cap);
    const int *buf = c == 0 ? sm.u.td.buf0 : (c == 1 ? sm.u.td.buf1 : sm.u.td.buf2);
    const unsigned base = sm.base[c];
    int *dst = a.q[par][c];
    for (int i = threadIdx.x; i < k; i += BS) {
      if (base + unsigned(i) < unsigned(a.n)) dst[base + i] = buf[i];
    }
    if (k > 0 && threadIdx.x == 0 && base + unsigned(k) > unsigned(a.n)) atomicOr(&a.ctl->error, ERR_OVERFLOW);
  }
  __syncthreads();
}

template <int BS>
__device__ __forceinline__ void td_visit(const Args &a, Smem<BS> &sm, int u, int L, int npar,
                                         LevelSlot *out) {
  const unsigned long long bit = 1ull << (u & 63);
  unsigned long long *w = a.visited + (u >> 6);
  if (*w & bit) return;  // may be stale; the atomic below decides
  const unsigned long long old = atomicOr(w, bit);
  if (old & bit) return;
  a.depth[u] = L + 1;
  push_vertex(a, sm, u, a.off[u], a.off[u + 1], npar, out);
}

// Top-down level L: expand the class queues of parity L & 1 (counts in slot
// L % 3) into those of parity (L + 1) & 1 (slot (L + 1) % 3).
template <int BS>
__device__ void td_phase(const Args &a, Smem<BS> &sm, int L) {
  LevelSlot *in = &a.ctl->slot[L % 3];
  LevelSlot *out = &a.ctl->slot[(L + 1) % 3];
  const int par = L & 1, npar = par ^ 1;
  const unsigned n0 = min(ld_agent(&in->qsize[0]), unsigned(a.n));
  const unsigned n1 = min(ld_agent(&in->qsize[1]), unsigned(a.n));
  const unsigned n2 = min(ld_agent(&in->qsize[2]), unsigned(a.n));
  const unsigned n3 = unsigned(min((long long)ld_agent(&in->qsize[3]), a.cap3));
  const int lane = lane_id();
  // Grid class: one block per piece of a huge adjacency list.
  for (unsigned i = blockIdx.x; i < n3; i += gridDim.x) {
    const long long x = a.qg[par][i];
    const int v = int(x >> 32);
    const long long k = x & 0xffffffffll;
    const long long s = a.off[v] + k * a.piece;
    const long long e = min(a.off[v + 1], s + a.piece);
    for (long long p = s + threadIdx.x; p < e; p += BS) td_visit(a, sm, a.adj[p], L, npar, out);
  }
  // Block class.
  for (unsigned i = blockIdx.x; i < n2; i += gridDim.x) {
    const int v = a.q[par][2][i];
    const long long s = a.off[v], e = a.off[v + 1];
    for (long long p = s + threadIdx.x; p < e; p += BS) td_visit(a, sm, a.adj[p], L, npar, out);
  }
  // Wavefront class: 64 lanes stride one adjacency list (coalesced).
  {
    const unsigned gw = blockIdx.x * (BS / kWave) + threadIdx.x / kWave;
    const unsigned nw = gridDim.x * (BS / kWave);
    for (unsigned i = gw; i < n1; i += nw) {
      const int v = a.q[par][1][i];
      const long long s = a.off[v], e = a.off[v + 1];
      for (long long p = s + lane; p < e; p += kWave) td_visit(a, sm, a.adj[p], L, npar, out);
    }
  }
  // Thread class: one lane per short adjacency list.
  {
    const unsigned gt = blockIdx.x * BS + threadIdx.x;
    const unsigned nt = gridDim.x * BS;
    for (unsigned i = gt; i < n0; i += nt) {
      const int v = a.q[par][0][i];
      const long long s = a.off[v], e = a.off[v + 1];
      for (long long p = s; p < e; ++p) td_visit(a, sm, a.adj[p], L, npar, out);
    }
  }
  flush_push(a, sm, npar, out, true);
}

// Bottom-up level L: every unvisited vertex with neighbours looks for a
// neighbour in frontier L and stops at the first one (expander.cuh:486-523).
// A wavefront owns a chunk of bitmap words: it compacts the chunk's candidate
// vertices (popcount + DPP scan), gives each lane one candidate for up to
// bu_probe neighbours, hands longer unresolved lists to the whole wavefront
// (ballot early exit), then writes its words of the next frontier and of the
// visited bitmap without atomics.
template <int BS>
__device__ void bu_phase(const Args &a, Smem<BS> &sm, int L, bool use_bitmap) {
  LevelSlot *out = &a.ctl->slot[(L + 1) % 3];
  const unsigned long long *fr = a.front[L & 1];
  unsigned long long *fn = a.front[(L + 1) & 1];
  const int wid = threadIdx.x / kWave, lane = lane_id();
  unsigned long long *wc = sm.u.bu.cand[wid];
  unsigned long long *wf = sm.u.bu.found[wid];
  int *we = sm.u.bu.excl[wid];
  const int CW = a.bu_chunk_words;
  const long long nchunks = ((long long)a.nwords + CW - 1) / CW;
  // XCD-aware numbering: blocks are dispatched round-robin over the 8 XCDs,
  // so give each XCD a contiguous range of chunks.
  long long eb = blockIdx.x;
  if ((gridDim.x & 7) == 0) eb = (long long)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const long long gw = eb * (BS / kWave) + wid;
  const long long nwv = (long long)gridDim.x * (BS / kWave);
  unsigned long long nf = 0, mf = 0;
  for (long long ch = gw; ch < nchunks; ch += nwv) {
    const long long w0 = ch * CW;
    const long long w = w0 + lane;
    const bool own = lane < CW && w < a.nwords;
    const unsigned long long vis = own ? a.visited[w] : ~0ull;
    const unsigned long long cand = ~vis;
    const int cnt = __popcll(cand);
    const int incl = wave_incl_scan(cnt);
    const int total = wave_read(incl, 63);
    wc[lane] = cand;
    we[lane] = incl - cnt;
    wf[lane] = 0;
    wave_sync();
    for (int b0 = 0; b0 < total; b0 += kWave) {
      const int k = b0 + lane;
      int v = 0, j = 0;
      long long s = 0, e = 0, p = 0;
      unsigned long long bitm = 0;
      bool found = false;
      if (k < total) {
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (we[mid] <= k) lo = mid;
          else hi = mid - 1;
        }
        j = lo;
        const int bp = select_bit(wc[j], k - we[j]);
        v = int((w0 + j) * 64 + bp);
        bitm = 1ull << bp;
        s = a.off[v];
        e = a.off[v + 1];
        const long long pe = min(e, s + a.bu_probe);
        for (p = s; p < pe; ++p) {
          if (in_front(a, fr, use_bitmap, a.adj[p], L)) {
            found = true;
            break;
          }
        }
      }
      if (found) {
        atomicOr(&wf[j], bitm);
        a.depth[v] = L + 1;
        nf += 1;
        mf += (unsigned long long)(e - s);
      }
      unsigned long long lmask = __ballot(k < total && !found && p < e);
      while (lmask) {
        const int src = __ffsll((long long)lmask) - 1;
        lmask &= lmask - 1;
        const int vv = __shfl(v, src);
        const int jj = __shfl(j, src);
        const long long ss = __shfl(s, src);
        const long long pp = __shfl(p, src);
        const long long ee = __shfl(e, src);
        const unsigned long long bb = __shfl(bitm, src);
        bool hit_any = false;
        for (long long q0 = pp; q0 < ee; q0 += kWave) {
          const long long q = q0 + lane;
          const bool hit = q < ee && in_front(a, fr, use_bitmap, a.adj[q], L);
          if (__ballot(hit)) {
            hit_any = true;
            break;
          }
        }
        if (hit_any && lane == 0) {
          atomicOr(&wf[jj], bb);
          a.depth[vv] = L + 1;
          nf += 1;
          mf += (unsigned long long)(ee - ss);
        }
      }
    }
    wave_sync();
    if (own) {
      const unsigned long long f = wf[lane];
      fn[w] = f;
      if (f) a.visited[w] = vis | f;
    }
    wave_sync();
  }
  if (nf) {
    atomicAdd(&sm.nf, nf);
    atomicAdd(&sm.mf, mf);
  }
  __syncthreads();
  if (threadIdx.x == 0 && sm.nf) {
    atomicAdd(&out->nf, unsigned(sm.nf));
    atomicAdd(&out->mf, sm.mf);
  }
  __syncthreads();
}

// Inspection after bottom-up: turns the frontier bitmap of level L into the
// class queues of parity L & 1 (slot L % 3) for a top-down expansion.
template <int BS>
__device__ void inspect_phase(const Args &a, Smem<BS> &sm, int L) {
  LevelSlot *slot = &a.ctl->slot[L % 3];
  const unsigned long long *fr = a.front[L & 1];
  const int par = L & 1;
  const long long gt = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nt = (long long)gridDim.x * BS;
  for (long long w = gt; w < a.nwords; w += nt) {
    unsigned long long x = fr[w];
    while (x) {
      const int bp = __ffsll((long long)x) - 1;
      x &= x - 1;
      const int v = int(w * 64 + bp);
      push_vertex(a, sm, v, a.off[v], a.off[v + 1], par, slot);
    }
  }
  flush_push(a, sm, par, slot, false);
}

// Chooses the next phase. Runs on one thread after every block finished the
// current phase.
__device__ void decide(const Args &a) {
  Control *c = a.ctl;
  const int L = c->level, ph = c->phase;
  c->phases += 1;
  if (ph == PH_INSPECT) {
    c->phase = PH_TD;
    return;
  }
  LevelSlot *nx = &c->slot[(L + 1) % 3];
  const unsigned long long nf = ld_agent(&nx->nf);
  const unsigned long long mf = ld_agent(&nx->mf);
  if (ph == PH_BU) {
    c->bu_levels += 1;
    if (L < 64) c->bu_mask |= 1ull << L;
  } else {
    c->td_levels += 1;
  }
  c->total_nf += nf;
  c->total_mf += mf;
  LevelSlot *z = &c->slot[(L + 2) % 3];
  for (int i = 0; i < 4; ++i) st_agent(&z->qsize[i], 0u);
  st_agent(&z->nf, 0u);
  st_agent(&z->mf, 0ull);
  if (nf == 0) {
    c->levels = L + 1;
    c->done = 1;
    return;
  }
  if (L + 1 >= a.n) {
    atomicOr(&c->error, ERR_LEVELS);
    c->done = 1;
    return;
  }
  const bool cur_bu = ph == PH_BU;
  bool next_bu;
  switch (a.mode) {
    case ENT_DIR_TOPDOWN: next_bu = false; break;
    case ENT_DIR_BOTTOMUP: next_bu = true; break;
    case ENT_DIR_REFERENCE: next_bu = cur_bu || 10ull * mf >= 3ull * (unsigned long long)a.m; break;
    default: {
      const unsigned long long mu = (unsigned long long)a.m - c->total_mf;
      if (!cur_bu) next_bu = mf * (unsigned long long)a.alpha > mu;
      else next_bu = !(nf * (unsigned long long)a.beta < (unsigned long long)a.n_active && nf < c->prev_nf);
    }
  }
  c->prev_nf = nf;
  c->level = L + 1;
  if (next_bu) {
    c->phase = PH_BU;
    c->bitmap_valid = cur_bu ? 1 : 0;
  } else {
    c->phase = cur_bu ? PH_INSPECT : PH_TD;
    c->bitmap_valid = 0;
  }
}

// depth = -1, visited = isolated vertices and padding bits, root at level 0,
// root queued in its class, counters cleared.
template <int BS>
__global__ __launch_bounds__(BS) void k_init(Args a, int root, int first_phase) {
  const long long gt = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nt = (long long)gridDim.x * BS;
  for (long long v = gt; v < a.n; v += nt) a.depth[v] = v == root ? 0 : -1;
  for (long long w = gt; w < a.nwords; w += nt) {
    unsigned long long vis = ~a.nzmask[w];
    if (w == (root >> 6)) vis |= 1ull << (root & 63);
    a.visited[w] = vis;
  }
  if (gt == 0) {
    Control *c = a.ctl;
    c->level = 0;
    c->phase = first_phase;
    c->done = 0;
    c->error = 0;
    c->bitmap_valid = 0;
    c->levels = 0;
    c->td_levels = c->bu_levels = c->phases = 0;
    c->arrive = 0;
    c->gen = 0;
    c->prev_nf = 0;
    c->bu_mask = 0;
    for (int s = 0; s < 3; ++s) {
      for (int i = 0; i < 4; ++i) c->slot[s].qsize[i] = 0;
      c->slot[s].nf = 0;
      c->slot[s].mf = 0;
    }
    const long long s0 = a.off[root], e0 = a.off[root + 1], d = e0 - s0;
    c->total_nf = 1;
    c->total_mf = (unsigned long long)d;
    c->slot[0].nf = 1;
    c->slot[0].mf = (unsigned long long)d;
    if (d >= a.t_grid) {
      const unsigned np = unsigned((d + a.piece - 1) / a.piece);
      for (unsigned k = 0; k < np && k < a.cap3; ++k) a.qg[0][k] = ((long long)root << 32) | k;
      c->slot[0].qsize[3] = np;
    } else {
      const int cls = d >= a.t_block ? 2 : (d >= a.t_wave ? 1 : 0);
      a.q[0][cls][0] = root;
      c->slot[0].qsize[cls] = 1;
    }
  }
}

template <int BS>
__device__ __forceinline__ void load_control(const Args &a, Smem<BS> &sm) {
  if (threadIdx.x == 0) {
    Control *c = a.ctl;
    sm.L = ld_agent(&c->level);
    sm.ph = ld_agent(&c->phase);
    sm.done = ld_agent(&c->done);
    sm.bm = ld_agent(&c->bitmap_valid);
    sm.cnt[0] = sm.cnt[1] = sm.cnt[2] = 0;
    sm.nf = sm.mf = 0;
  }
  __syncthreads();
}

template <int BS>
__device__ __forceinline__ void run_phase(const Args &a, Smem<BS> &sm) {
  if (sm.ph == PH_TD) td_phase<BS>(a, sm, sm.L);
  else if (sm.ph == PH_BU) bu_phase<BS>(a, sm, sm.L, sm.bm != 0);
  else inspect_phase<BS>(a, sm, sm.L);
}

// One phase per launch; the last block to finish picks the next phase.
template <int BS>
__global__ __launch_bounds__(BS) void k_level(Args a) {
  __shared__ Smem<BS> sm;
  load_control(a, sm);
  if (sm.done) return;
  run_phase<BS>(a, sm);
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned t = atomicAdd(&a.ctl->arrive, 1u);
    if (t == gridDim.x - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      decide(a);
      st_agent(&a.ctl->arrive, 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
  }
}

// Grid barrier whose last arrival runs decide(). Bounded: a wait longer than
// ~4 s marks the search failed and releases the waiter.
template <int BS>
__device__ __forceinline__ void grid_barrier_decide(const Args &a) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    Control *c = a.ctl;
    const unsigned g = ld_agent(&c->gen);
    const unsigned t = __hip_atomic_fetch_add(&c->arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (t == gridDim.x - 1) {
      decide(a);
      st_agent(&c->arrive, 0u);
      __hip_atomic_store(&c->gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint64_t t0 = rt_clock();
      while (__hip_atomic_load(&c->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(2);
        if (rt_clock() - t0 > 400000000ull) {
          atomicOr(&c->error, ERR_TIMEOUT);
          st_agent(&c->done, 1);
          break;
        }
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Whole search in one cooperative launch.
template <int BS>
__global__ __launch_bounds__(BS) void k_persist(Args a) {
  __shared__ Smem<BS> sm;
  for (;;) {
    load_control(a, sm);
    if (sm.done) break;
    run_phase<BS>(a, sm);
    grid_barrier_decide<BS>(a);
  }
}

// ---------------------------------------------------------------------------
// Wave primitive self-test.
__global__ void k_selftest(int *out) {
  const int lane = lane_id();
  const int x = (lane * 7 + 3) % 11;
  out[lane] = wave_incl_scan(x);
  const unsigned long long m = __ballot(x & 1);
  out[64 + lane] = (x & 1) ? mask_rank(m) : -1;
  const unsigned long long word = 0x8421084210842108ull ^ (unsigned long long)lane * 0x9E3779B97F4A7C15ull;
  const int pc = __popcll(word);
  out[128 + lane] = select_bit(word, lane % pc);
  out[192 + lane] = wave_sum(lane);
}

}  // namespace ent

// ===========================================================================
// Host side.
using namespace ent;

struct ent_graph {
  int device = 0;
  int64_t n = 0, m = 0, n_active = 0, nwords = 0;
  hipStream_t stream = nullptr;
  long long *off = nullptr;
  int *adj = nullptr;
  unsigned long long *nzmask = nullptr;
  int *depth = nullptr;
  unsigned long long *visited = nullptr;
  unsigned long long *front[2] = {nullptr, nullptr};
  int *q[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  long long *qg[2] = {nullptr, nullptr};
  int64_t cap3 = 0;
  Control *ctl = nullptr;
  Control *h_ctl = nullptr;
  ent_bfs_config cfg{};
  int cus = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  Args graph_args{};
  int graph_bs = 0, graph_grid = 0, graph_k = 0;
  size_t dev_bytes = 0;
  int last_grid = 0;
  std::mutex mu;
};

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_error(ENT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

namespace {

int dev_alloc(ent_graph *g, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(ENT_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  g->dev_bytes += bytes;
  return ENT_OK;
}

template <class T>
inline void ignore(T) {}

void release(ent_graph *g) {
  if (!g) return;
  ignore(hipSetDevice(g->device));
  if (g->gexec) ignore(hipGraphExecDestroy(g->gexec));
  if (g->graph) ignore(hipGraphDestroy(g->graph));
  ignore(hipFree(g->off));
  ignore(hipFree(g->adj));
  ignore(hipFree(g->nzmask));
  ignore(hipFree(g->depth));
  ignore(hipFree(g->visited));
  for (int p = 0; p < 2; ++p) {
    ignore(hipFree(g->front[p]));
    for (int c = 0; c < 3; ++c) ignore(hipFree(g->q[p][c]));
    ignore(hipFree(g->qg[p]));
  }
  ignore(hipFree(g->ctl));
  if (g->h_ctl) ignore(hipHostFree(g->h_ctl));
  if (g->ev0) ignore(hipEventDestroy(g->ev0));
  if (g->ev1) ignore(hipEventDestroy(g->ev1));
  if (g->stream) ignore(hipStreamDestroy(g->stream));
  delete g;
}

int64_t piece_capacity(int64_t n, int64_t m, int t_grid, int piece) {
  return std::min<int64_t>(n, m / std::max(1, t_grid) + 1) + m / std::max(1, piece) + 1;
}

int check_config(const ent_bfs_config *c) {
  if (c->direction < 0 || c->direction > 3) return set_error(ENT_EINVAL, "direction %d", c->direction);
  if (c->launch < 0 || c->launch > 3) return set_error(ENT_EINVAL, "launch %d", c->launch);
  if (c->alpha < 1 || c->beta < 1) return set_error(ENT_EINVAL, "alpha/beta must be >= 1");
  if (c->t_wave < 1 || c->t_block < c->t_wave || c->t_grid < c->t_block)
    return set_error(ENT_EINVAL, "thresholds must satisfy 1 <= t_wave <= t_block <= t_grid");
  if (c->piece < 64 || c->piece > c->t_grid) return set_error(ENT_EINVAL, "piece must be in [64, t_grid]");
  const int bs = c->block_size;
  if (bs != 64 && bs != 128 && bs != 256 && bs != 512 && bs != 1024)
    return set_error(ENT_EINVAL, "block_size %d not in {64,128,256,512,1024}", bs);
  if (c->blocks_per_cu < 0 || c->blocks_per_cu > 32) return set_error(ENT_EINVAL, "blocks_per_cu %d", c->blocks_per_cu);
  if (c->bu_probe < 0 || c->bu_probe > 1 << 20) return set_error(ENT_EINVAL, "bu_probe %d", c->bu_probe);
  if (c->bu_chunk_words < 1 || c->bu_chunk_words > 64) return set_error(ENT_EINVAL, "bu_chunk_words %d", c->bu_chunk_words);
  if (c->graph_levels < 1 || c->graph_levels > 256) return set_error(ENT_EINVAL, "graph_levels %d", c->graph_levels);
  return ENT_OK;
}

Args make_args(ent_graph *g, int *depth) {
  Args a{};
  a.off = g->off;
  a.adj = g->adj;
  a.nzmask = g->nzmask;
  a.depth = depth;
  a.visited = g->visited;
  a.front[0] = g->front[0];
  a.front[1] = g->front[1];
  for (int p = 0; p < 2; ++p) {
    for (int c = 0; c < 3; ++c) a.q[p][c] = g->q[p][c];
    a.qg[p] = g->qg[p];
  }
  a.ctl = g->ctl;
  a.m = g->m;
  a.n_active = g->n_active;
  a.cap3 = g->cap3;
  a.n = int(g->n);
  a.nwords = int(g->nwords);
  const ent_bfs_config &c = g->cfg;
  a.t_wave = c.t_wave;
  a.t_block = c.t_block;
  a.t_grid = c.t_grid;
  a.piece = c.piece;
  a.mode = c.direction;
  a.alpha = c.alpha;
  a.beta = c.beta;
  a.bu_probe = c.bu_probe;
  a.bu_chunk_words = c.bu_chunk_words;
  return a;
}

template <int BS>
struct Kernels {
  static const void *init() { return reinterpret_cast<const void *>(&k_init<BS>); }
  static const void *level() { return reinterpret_cast<const void *>(&k_level<BS>); }
  static const void *persist() { return reinterpret_cast<const void *>(&k_persist<BS>); }
};

struct KernelSet {
  const void *init, *level, *persist;
};

KernelSet kernels_for(int bs) {
  switch (bs) {
    case 64: return {Kernels<64>::init(), Kernels<64>::level(), Kernels<64>::persist()};
    case 128: return {Kernels<128>::init(), Kernels<128>::level(), Kernels<128>::persist()};
    case 512: return {Kernels<512>::init(), Kernels<512>::level(), Kernels<512>::persist()};
    case 1024: return {Kernels<1024>::init(), Kernels<1024>::level(), Kernels<1024>::persist()};
    default: return {Kernels<256>::init(), Kernels<256>::level(), Kernels<256>::persist()};
  }
}

int level_grid(ent_graph *g) {
  const int bpc = g->cfg.blocks_per_cu > 0 ? g->cfg.blocks_per_cu : std::max(1, 2048 / g->cfg.block_size);
  return g->cus * bpc;
}

hipError_t launch(const void *fn, int grid, int bs, void **args, hipStream_t s) {
  return hipLaunchKernel(fn, dim3(grid), dim3(bs), args, 0, s);
}

// Queues the per-phase kernel k times (used directly and under capture).
hipError_t queue_levels(ent_graph *g, const KernelSet &ks, Args *a, int k) {
  void *args[] = {a};
  const int grid = level_grid(g);
  g->last_grid = grid;
  for (int i = 0; i < k; ++i) {
    hipError_t e = launch(ks.level, grid, g->cfg.block_size, args, g->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t smem_bytes(int bs) {
  switch (bs) {
    case 64: return sizeof(Smem<64>);
    case 128: return sizeof(Smem<128>);
    case 512: return sizeof(Smem<512>);
    case 1024: return sizeof(Smem<1024>);
    default: return sizeof(Smem<256>);
  }
}

// Blocks per CU for the cooperative kernel: the occupancy query, capped by
// the LDS budget (160 KiB per CU, each block's LDS rounded up to a 2 KiB
// allocation granule), 32 wavefronts per CU and 16 workgroups per CU, and by
// blocks_per_cu when set. Every block must be resident at once.
int persistent_blocks_per_cu(ent_graph *g, const KernelSet &ks) {
  const int bs = g->cfg.block_size;
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ks.persist, bs, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const size_t lds_block = (smem_bytes(bs) + 2047) / 2048 * 2048;
  const int lds_limit = int((160 * 1024) / lds_block);
  const int wave_limit = 32 / (bs / kWave);
  int per_cu = std::min(std::min(occ, lds_limit), std::min(wave_limit, 16));
  if (g->cfg.blocks_per_cu > 0) per_cu = std::min(per_cu, g->cfg.blocks_per_cu);
  return per_cu;
}

int run_search(ent_graph *g, int root, int *depth_dev) {
  const ent_bfs_config &cfg = g->cfg;
  const KernelSet ks = kernels_for(cfg.block_size);
  Args a = make_args(g, depth_dev);
  const int bs = cfg.block_size;
  const int first_phase = cfg.direction == ENT_DIR_BOTTOMUP ? PH_BU : PH_TD;
  {
    void *args[] = {&a, &root, const_cast<int *>(&first_phase)};
    const int grid = int(std::min<int64_t>(int64_t(g->cus) * 8, (g->n + bs - 1) / bs + 1));
    HIP_TRY(launch(ks.init, std::max(1, grid), bs, args, g->stream));
  }
  const int64_t max_phases = 2 * g->n + 16;
  switch (cfg.launch) {
    case ENT_LAUNCH_PERSISTENT: {
      const int per_cu = persistent_blocks_per_cu(g, ks);
      if (per_cu < 1) return set_error(ENT_EDEVICE, "persistent kernel cannot be resident");
      void *args[] = {&a};
      g->last_grid = g->cus * per_cu;
      HIP_TRY(hipLaunchCooperativeKernel(ks.persist, dim3(g->last_grid), dim3(bs), args, 0, g->stream));
      break;
    }
    case ENT_LAUNCH_LEVELS:
    case ENT_LAUNCH_STREAM: {
      const int k = cfg.launch == ENT_LAUNCH_LEVELS ? 1 : cfg.graph_levels;
      for (int64_t ph = 0;; ph += k) {
        HIP_TRY(queue_levels(g, ks, &a, k));
        HIP_TRY(hipMemcpyAsync(g->h_ctl, g->ctl, 16, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (g->h_ctl->done) break;
        if (ph > max_phases) return set_error(ENT_EINTERNAL, "search did not finish in %lld phases", (long long)ph);
      }
      break;
    }
    case ENT_LAUNCH_GRAPH: {
      const int k = cfg.graph_levels;
      const int grid = level_grid(g);
      if (!g->gexec || g->graph_bs != bs || g->graph_grid != grid || g->graph_k != k ||
          std::memcmp(&g->graph_args, &a, sizeof(Args)) != 0) {
        if (g->gexec) ignore(hipGraphExecDestroy(g->gexec));
        if (g->graph) ignore(hipGraphDestroy(g->graph));
        g->gexec = nullptr;
        g->graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = queue_levels(g, ks, &a, k);
        hipGraph_t gr = nullptr;
        hipError_t e2 = hipStreamEndCapture(g->stream, &gr);
        if (e != hipSuccess || e2 != hipSuccess)
          return set_error(ENT_EDEVICE, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        g->graph = gr;
        HIP_TRY(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
        g->graph_args = a;
        g->graph_bs = bs;
        g->graph_grid = grid;
        g->graph_k = k;
      }
      for (int64_t ph = 0;; ph += k) {
        HIP_TRY(hipGraphLaunch(g->gexec, g->stream));
        HIP_TRY(hipMemcpyAsync(g->h_ctl, g->ctl, 16, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (g->h_ctl->done) break;
        if (ph > max_phases) return set_error(ENT_EINTERNAL, "search did not finish in %lld phases", (long long)ph);
      }
      break;
    }
  }
  return ENT_OK;
}

}  // namespace

extern "C" {

void ent_bfs_default_config(ent_bfs_config *c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->direction = ENT_DIR_AUTO;
  c->launch = ENT_LAUNCH_STREAM;
  c->alpha = 15;
  c->beta = 18;
  c->t_wave = 32;
  c->t_block = 2048;
  c->t_grid = 32768;
  c->piece = 8192;
  c->block_size = 256;
  c->blocks_per_cu = 0;
  c->bu_probe = 16;
  c->bu_chunk_words = 8;
  c->graph_levels = 8;
  c->hub_cache = 0;
}

int ent_graph_upload(const int64_t *beg_pos, const int32_t *adj, int64_t n, int64_t m, int device,
                     ent_graph **out) {
  clear_error();
  if (!out) return set_error(ENT_EINVAL, "out is NULL");
  *out = nullptr;
  if (!beg_pos || (m > 0 && !adj)) return set_error(ENT_EINVAL, "NULL CSR array");
  if (n < 1 || n >= (int64_t(1) << 31) - 64) return set_error(ENT_EINVAL, "n=%lld outside [1, 2^31-64)", (long long)n);
  if (m < 0 || beg_pos[0] != 0 || beg_pos[n] != m)
    return set_error(ENT_EINVAL, "offsets do not span [0, m=%lld]", (long long)m);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return set_error(ENT_EDEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev) return set_error(ENT_EINVAL, "device %d of %d", device, ndev);
  // Host-side structural checks: monotone offsets and in-range ids.
  {
    std::atomic<int64_t> bad{-1};
    parallel_for(n, 1 << 15, [&](int64_t b, int64_t e) {
      for (int64_t v = b; v < e; ++v) {
        if (beg_pos[v + 1] < beg_pos[v]) { bad.store(v); return; }
        for (int64_t i = beg_pos[v]; i < beg_pos[v + 1]; ++i)
          if (adj[i] < 0 || adj[i] >= n) { bad.store(v); return; }
      }
    });
    if (bad.load() >= 0) return set_error(ENT_EINVAL, "malformed CSR at vertex %lld", (long long)bad.load());
  }
  ent_graph *g = new (std::nothrow) ent_graph();
  if (!g) return set_error(ENT_ENOMEM, "out of host memory");
  g->device = device;
  g->n = n;
  g->m = m;
  g->nwords = (n + 63) / 64;
  ent_bfs_default_config(&g->cfg);
  int rc = ENT_OK;
#define UP_TRY(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      rc = set_error(ENT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
      release(g);                                                                      \
      return rc;                                                                       \
    }                                                                                  \
  } while (0)
#define UP_ALLOC(ptr, bytes)                                           \
  do {                                                                 \
    rc = dev_alloc(g, reinterpret_cast<void **>(&(ptr)), (bytes));     \
    if (rc) {                                                          \
      release(g);                                                      \
      return rc;                                                       \
    }                                                                  \
  } while (0)
  UP_TRY(hipSetDevice(device));
  UP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  UP_TRY(hipEventCreate(&g->ev0));
  UP_TRY(hipEventCreate(&g->ev1));
  UP_TRY(hipDeviceGetAttribute(&g->cus, hipDeviceAttributeMultiprocessorCount, device));
  std::vector<unsigned long long> nz(size_t(g->nwords), 0ull);
  int64_t active = 0;
  for (int64_t v = 0; v < n; ++v)
    if (beg_pos[v + 1] > beg_pos[v]) {
      nz[size_t(v >> 6)] |= 1ull << (v & 63);
      ++active;
    }
  g->n_active = active;
  g->cap3 = piece_capacity(n, m, g->cfg.t_grid, g->cfg.piece);
  UP_ALLOC(g->off, sizeof(long long) * size_t(n + 1));
  UP_ALLOC(g->adj, sizeof(int) * size_t(m));
  UP_ALLOC(g->nzmask, sizeof(unsigned long long) * size_t(g->nwords));
  UP_ALLOC(g->depth, sizeof(int) * size_t(n));
  UP_ALLOC(g->visited, sizeof(unsigned long long) * size_t(g->nwords));
  for (int p = 0; p < 2; ++p) {
    UP_ALLOC(g->front[p], sizeof(unsigned long long) * size_t(g->nwords));
    for (int c = 0; c < 3; ++c) UP_ALLOC(g->q[p][c], sizeof(int) * size_t(n));
    UP_ALLOC(g->qg[p], sizeof(long long) * size_t(g->cap3));
  }
  UP_ALLOC(g->ctl, sizeof(Control));
  UP_TRY(hipHostMalloc(reinterpret_cast<void **>(&g->h_ctl), sizeof(Control), hipHostMallocDefault));
  UP_TRY(hipMemcpy(g->off, beg_pos, sizeof(long long) * size_t(n + 1), hipMemcpyHostToDevice));
  if (m) UP_TRY(hipMemcpy(g->adj, adj, sizeof(int) * size_t(m), hipMemcpyHostToDevice));
  UP_TRY(hipMemcpy(g->nzmask, nz.data(), sizeof(unsigned long long) * size_t(g->nwords), hipMemcpyHostToDevice));
  UP_TRY(hipMemset(g->ctl, 0, sizeof(Control)));
  UP_TRY(hipDeviceSynchronize());
#undef UP_TRY
#undef UP_ALLOC
  *out = g;
  return ENT_OK;
}

int ent_graph_free(ent_graph *g) {
  clear_error();
  if (!g) return ENT_OK;
  {
    std::lock_guard<std::mutex> lk(g->mu);
    ignore(hipSetDevice(g->device));
    ignore(hipStreamSynchronize(g->stream));
  }
  release(g);
  return ENT_OK;
}

int ent_graph_info(const ent_graph *g, int64_t *n, int64_t *m, int64_t *n_active, int64_t *device_bytes) {
  clear_error();
  if (!g) return set_error(ENT_EINVAL, "NULL graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (n_active) *n_active = g->n_active;
  if (device_bytes) *device_bytes = int64_t(g->dev_bytes);
  return ENT_OK;
}

int ent_set_config(ent_graph *g, const ent_bfs_config *cfg) {
  clear_error();
  if (!g || !cfg) return set_error(ENT_EINVAL, "NULL argument");
  int rc = check_config(cfg);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g->mu);
  const int64_t need = piece_capacity(g->n, g->m, cfg->t_grid, cfg->piece);
  if (need > g->cap3) {
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipStreamSynchronize(g->stream));
    for (int p = 0; p < 2; ++p) {
      ignore(hipFree(g->qg[p]));
      g->qg[p] = nullptr;
      rc = dev_alloc(g, reinterpret_cast<void **>(&g->qg[p]), sizeof(long long) * size_t(need));
      if (rc) return rc;
    }
    g->cap3 = need;
  }
  g->cfg = *cfg;
  return ENT_OK;
}

int ent_get_config(const ent_graph *g, ent_bfs_config *cfg) {
  clear_error();
  if (!g || !cfg) return set_error(ENT_EINVAL, "NULL argument");
  *cfg = g->cfg;
  return ENT_OK;
}

int32_t *ent_depth_device_ptr(ent_graph *g) { return g ? g->depth : nullptr; }

int ent_bfs_device(ent_graph *g, int32_t root, int32_t *depth_dev, ent_stats *st) {
  clear_error();
  if (!g) return set_error(ENT_EINVAL, "NULL graph");
  if (root < 0 || root >= g->n) return set_error(ENT_EINVAL, "root %d outside [0, %lld)", root, (long long)g->n);
  std::lock_guard<std::mutex> lk(g->mu);
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(g->device));
  int *depth = depth_dev ? depth_dev : g->depth;
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  int rc = run_search(g, root, depth);
  if (rc) {
    ignore(hipStreamSynchronize(g->stream));
    return rc;
  }
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipMemcpyAsync(g->h_ctl, g->ctl, sizeof(Control), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  const auto t1 = std::chrono::steady_clock::now();
  const Control &c = *g->h_ctl;
  if (c.error)
    return set_error(ENT_EINTERNAL, "device reported error flags 0x%x (1 queue overflow, 2 level guard, 4 barrier timeout)",
                     c.error);
  if (!c.done) return set_error(ENT_EINTERNAL, "search ended without completion");
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->levels = c.levels;
    st->td_levels = c.td_levels;
    st->bu_levels = c.bu_levels;
    st->phases = c.phases;
    st->reached_vertices = int64_t(c.total_nf);
    st->reached_edges = int64_t(c.total_mf);
    st->alg_bytes = 8 * st->reached_edges + 12 * st->reached_vertices + 4 * g->n;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->ev0, g->ev1) == hipSuccess) st->device_ms = ms;
    st->host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    st->bu_level_mask = c.bu_mask;
    st->grid_blocks = g->last_grid;
    st->block_size = g->cfg.block_size;
  }
  return ENT_OK;
}

int ent_bfs(ent_graph *g, int32_t root, int32_t *depth_host, ent_stats *st) {
  int rc = ent_bfs_device(g, root, nullptr, st);
  if (rc || !depth_host) return rc;
  return ent_copy_to_host(g, g->depth, depth_host, g->n);
}

int ent_copy_to_host(ent_graph *g, const int32_t *src_dev, int32_t *dst_host, int64_t count) {
  clear_error();
  if (!g || !src_dev || !dst_host || count < 0) return set_error(ENT_EINVAL, "bad argument");
  std::lock_guard<std::mutex> lk(g->mu);
  HIP_TRY(hipSetDevice(g->device));
  HIP_TRY(hipMemcpyAsync(dst_host, src_dev, sizeof(int32_t) * size_t(count), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return ENT_OK;
}

int ent_bfs_gpu_coalescing_mem(const int32_t *src_list, int32_t nsrc, const int64_t *beg_pos,
                               const int32_t *csr, int64_t n, int64_t m, int32_t gpu_id) {
  if (!src_list || nsrc < 0) return set_error(ENT_EINVAL, "bad source list");
  ent_graph *g = nullptr;
  int rc = ent_graph_upload(beg_pos, csr, n, m, gpu_id, &g);
  if (rc) return rc;
  std::printf("In gpu bfs\n");
  double avg = 0.0;
  int counted = 0;
  for (int32_t i = 0; i < nsrc; ++i) {
    std::printf("Test %d\nStarted from: %d\n", i + 1, src_list[i]);
    ent_stats st;
    const auto t0 = std::chrono::steady_clock::now();
    rc = ent_bfs_device(g, src_list[i], nullptr, &st);
    const auto t1 = std::chrono::steady_clock::now();
    if (rc) break;
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    if (st.levels - 1 > 2) {
      ++counted;
      const double teps = double(st.reached_edges) / (1e9 * sec);
      avg = (teps + avg * (counted - 1)) / counted;
      std::printf("Traversed vertices: %lld\t\t\tTraversed edges: %lld\nTraversed time(s) :%g\t\t"
                  "Current TEPS (Billion): %g\nAverage TEPS (Billion): %g\n",
                  (long long)st.reached_vertices, (long long)st.reached_edges, sec, teps, avg);
    } else {
      std::printf("Traverse depth is %d\n", st.levels - 1);
    }
    std::printf("\n====================================\n");
  }
  if (!rc) std::printf("Final Average TEPS (Billion): %g\n", avg);
  ent_graph_free(g);
  return rc;
}

int ent_selftest_wave(int device) {
  clear_error();
  HIP_TRY(hipSetDevice(device));
  int *d = nullptr;
  HIP_TRY(hipMalloc(&d, 256 * sizeof(int)));
  hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, 0, d);
  std::vector<int> h(256);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, 256 * sizeof(int), hipMemcpyDeviceToHost);
  ignore(hipFree(d));
  if (e != hipSuccess) return set_error(ENT_EDEVICE, "selftest: %s", hipGetErrorString(e));
  int run = 0, odd_rank = 0;
  for (int lane = 0; lane < 64; ++lane) {
    const int x = (lane * 7 + 3) % 11;
    run += x;
    if (h[lane] != run) return set_error(ENT_EINTERNAL, "scan lane %d: %d != %d", lane, h[lane], run);
    if (x & 1) {
      if (h[64 + lane] != odd_rank) return set_error(ENT_EINTERNAL, "rank lane %d", lane);
      ++odd_rank;
    }
    const unsigned long long word = 0x8421084210842108ull ^ (unsigned long long)lane * 0x9E3779B97F4A7C15ull;
    const int r = lane % __builtin_popcountll(word);
    int pos = -1;
    for (int b = 0, c = 0; b < 64; ++b)
      if ((word >> b) & 1) {
        if (c == r) { pos = b; break; }
        ++c;
      }
    if (h[128 + lane] != pos) return set_error(ENT_EINTERNAL, "select lane %d: %d != %d", lane, h[128 + lane], pos);
    if (h[192 + lane] != 2016) return set_error(ENT_EINTERNAL, "wave_sum lane %d", lane);
  }
  return ENT_OK;
}

}  // extern "C"
