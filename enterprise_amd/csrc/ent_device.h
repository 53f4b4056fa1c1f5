// Wavefront (wave64) primitives for gfx950: DPP inclusive scan, ballot-based
// compaction ranks, k-th set bit, scoped atomics and a bounded grid barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ent {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Rank of this lane among the lanes whose bit is set in mask (mbcnt).
__device__ __forceinline__ int mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Inclusive prefix sum across the 64 lanes with DPP: row_shr 1, 2, 4, 8 build
// 16-lane row scans, then row_bcast:15 (rows 1 and 3) and row_bcast:31 (rows 2
// and 3) carry row totals forward. All lanes must be active.
__device__ __forceinline__ int wave_incl_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  return x;
}

// Bitwise OR over all 64 lanes with the same DPP pattern, returned as a
// wave-uniform value (read from lane 63). All lanes must be active.
__device__ __forceinline__ uint32_t wave_or32(uint32_t v) {
  int x = int(v);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  return uint32_t(__builtin_amdgcn_readlane(x, 63));
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
  return uint64_t(wave_or32(uint32_t(v))) | (uint64_t(wave_or32(uint32_t(v >> 32))) << 32);
}

// Value of a 64-bit register in lane src (src wave-uniform), as a scalar.
__device__ __forceinline__ uint64_t read_lane64(uint64_t v, int src) {
  return uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), src))) |
         (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), src))) << 32);
}

__device__ __forceinline__ int wave_read(int x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}

// Sum over all 64 lanes (every lane receives the total).
__device__ __forceinline__ int wave_sum(int x) { return wave_read(wave_incl_scan(x), 63); }

// Position of the r-th (0-based) set bit of x; x must hold more than r bits.
__device__ __forceinline__ int select_bit(uint64_t x, int r) {
  int pos = 0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    const int c = __popcll(x & ((1ull << s) - 1));
    if (r >= c) {
      r -= c;
      x >>= s;
      pos += s;
    }
  }
  return pos;
}

// Orders LDS traffic between lanes of one wavefront (in-order LDS, but keep
// the compiler from moving accesses across this point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 100 MHz constant clock.
__device__ __forceinline__ uint64_t rt_clock() { return wall_clock64(); }

// Grid barrier state for cooperative (all blocks resident) launches: 64
// striped arrival counters one 128-byte line apart, a global arrival counter
// and a generation number. All zero before the first use of a launch. The
// counters only grow: barrier b takes arrivals b * quota .. (b + 1) * quota - 1
// of a stripe (no block arrives at barrier b + 1 before every block has
// arrived at b), so no counter is ever reset and no reset has to be ordered
// before the next barrier's arrivals.
constexpr int kSyncStripes = 64;
constexpr int kSyncStride = 16;  // u64 words between stripes
struct GridSync {
  unsigned long long arrive_all;
  unsigned int gen;
  unsigned int pad[29];
  unsigned long long arrive[kSyncStripes * kSyncStride];
};

// Two-level arrival: blocks count on their stripe, the last of each stripe
// (per barrier) on arrive_all; returns true for the single last block.
// Thread 0 only, after the block's release fence. The last arriver of a
// stripe acquires the stripe's arrivals before it counts on arrive_all with a
// release-acquire increment, so the block that returns true has acquired
// every block's release and publishes it with its generation store.
__device__ __forceinline__ bool grid_arrive_last(GridSync *c) {
  const unsigned long long G = gridDim.x;
  const unsigned long long s = blockIdx.x % kSyncStripes;
  const unsigned long long quota = (G - s + kSyncStripes - 1) / kSyncStripes;
  const unsigned long long nstripes = G < kSyncStripes ? G : kSyncStripes;
  unsigned long long *sc = &c->arrive[s * kSyncStride];
  const unsigned long long t = __hip_atomic_fetch_add(sc, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((t + 1) % quota != 0) return false;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const unsigned long long t2 =
      __hip_atomic_fetch_add(&c->arrive_all, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  return (t2 + 1) % nstripes == 0;
}

// Grid barrier. Every wavefront first waits for its own stores to complete
// (the block barrier alone does not), so one agent-scope release per block
// publishes the whole block's writes. Waiters poll the generation with
// relaxed loads (an acquire load per poll would invalidate the caches on
// every iteration) and take one agent-scope acquire after the flip. Bounded:
// returns false (in every thread of the block) when the wait exceeded
// `limit` ticks of rt_clock.
__device__ __forceinline__ bool grid_sync(GridSync *c, uint64_t limit) {
  __shared__ int sync_ok;
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are done
  __syncthreads();
  if (threadIdx.x == 0) {
    sync_ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned g = ld_agent(&c->gen);
    if (grid_arrive_last(c)) {
      __hip_atomic_store(&c->gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint64_t t0 = rt_clock();
      while (ld_agent(&c->gen) == g) {
        __builtin_amdgcn_s_sleep(1);
        if (rt_clock() - t0 > limit) {
          sync_ok = 0;
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return sync_ok != 0;
}

}  // namespace ent
