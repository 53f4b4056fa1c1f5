// Edge list to CSR on the device, behind include/ent_convert.h. The
// reference builds its CSR on the host, tuple by tuple, in file order
// (tuple_text_to_binary_csr/tuple_text_to_bin.cpp:190-248); this file
// gives every stored entry one 64-bit record and radix-sorts the records:
//
//   record r (tuple i = r / 2 with ENT_CSR_REVERSE, else i = r; the odd
//   record of a reversed tuple is (dst, src)):
//     input order : key = row                      (stable sort keeps tuple order)
//     sorted rows : key = row << cb | col          (cb = bits of n)
//   a dropped self-loop gets row = n and sorts behind every row.
//
// Sorted rows without weights sort the keys alone; the other forms carry the
// record index (u32) to fetch the column and weight afterwards. Then, per
// tile of 4096 sorted records: the count of kept records (real row, and with
// dedup a key unlike its predecessor's), a scan of the tile counts, and one
// assembly pass that writes each kept record at its rank: its column and
// weight, and beg_pos at the row boundaries.
//
// HBM traffic per record: the radix sort's passes over 8-byte keys (and
// 4-byte indices), one pass per 8 key bits; the record and assembly kernels
// read the tuples once and write the CSR once.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ent_bfs.h"
#include "ent_common.h"
#include "ent_convert.h"
#include "ent_device.h"
#include "ent_internal.h"

namespace ent {
namespace {

constexpr int kBuildBlock = 256;

// Bits needed for the values 0..x.
int bits_for(uint64_t x) {
  int b = 1;
  while (b < 64 && (uint64_t(1) << b) <= x) ++b;
  return b;
}

struct BuildShape {
  long long ne, records;
  int n, rev, no_loops, sorted, dedup, cb;
};

__device__ __forceinline__ void record_entry(const BuildShape &s, const int *src, const int *dst, long long r,
                                             long long &tuple, unsigned long long &row, unsigned long long &col) {
  tuple = s.rev ? r >> 1 : r;
  const bool back = s.rev && (r & 1);
  const int u = src[tuple], v = dst[tuple];
  row = (unsigned long long)(back ? v : u);
  col = (unsigned long long)(back ? u : v);
  if (s.no_loops && u == v) row = (unsigned long long)s.n;
}

__global__ __launch_bounds__(kBuildBlock) void k_records(BuildShape s, const int *src, const int *dst,
                                                         unsigned long long *keys, unsigned *idx) {
  for (long long r = (long long)blockIdx.x * kBuildBlock + threadIdx.x; r < s.records;
       r += (long long)gridDim.x * kBuildBlock) {
    long long t;
    unsigned long long row, col;
    record_entry(s, src, dst, r, t, row, col);
    keys[r] = s.sorted ? (row << s.cb) | (row == (unsigned long long)s.n ? 0ull : col) : row;
    if (idx) idx[r] = unsigned(r);
  }
}

// A run of empty rows too long for one thread: beg[v0 .. v1) = val.
struct Gap {
  long long v0, v1, val;
};
constexpr int kGapCap = 1 << 16;
constexpr long long kGapSerial = 256;  // longest run a thread fills itself

// Sets beg[v] = val for v in [v0, v1): by this thread when the run is
// short, else listed for k_fill_gaps (or by this thread if the list is full).
__device__ __forceinline__ void fill_rows(long long *beg, long long v0, long long v1, long long val, Gap *gaps,
                                          unsigned long long *ngaps) {
  if (v1 - v0 > kGapSerial) {
    const unsigned long long g = atomicAdd(ngaps, 1ull);
    if (g < (unsigned long long)kGapCap) {
      gaps[g] = Gap{v0, v1, val};
      return;
    }
  }
  for (long long v = v0; v < v1; ++v) beg[v] = val;
}

// Tiles of kTile sorted records, one block each. Record i is kept when its
// row is a real row (< n) and, with dedup, its key differs from record
// i - 1's (records of a dropped loop sort last, behind every kept one).
constexpr int kTileItems = 16;
constexpr long long kTile = (long long)kBuildBlock * kTileItems;

__device__ __forceinline__ bool record_kept(const BuildShape &s, const unsigned long long *keys, long long i,
                                            int shift) {
  const unsigned long long k = keys[i];
  if ((k >> shift) >= (unsigned long long)s.n) return false;
  return !s.dedup || i == 0 || keys[i - 1] != k;
}

__global__ __launch_bounds__(kBuildBlock) void k_tile_count(BuildShape s, const unsigned long long *keys,
                                                            unsigned long long *tile_cnt) {
  const int shift = s.sorted ? s.cb : 0;
  const long long t0 = (long long)blockIdx.x * kTile;
  int c = 0;
#pragma unroll 4
  for (int j = 0; j < kTileItems; ++j) {
    const long long i = t0 + (long long)j * kBuildBlock + threadIdx.x;
    if (i < s.records) c += record_kept(s, keys, i, shift) ? 1 : 0;
  }
  __shared__ int part[kBuildBlock / kWave];
  const int w = wave_sum(c);
  if (lane_id() == 0) part[threadIdx.x / kWave] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < kBuildBlock / kWave; ++k) t += (unsigned long long)part[k];
    tile_cnt[blockIdx.x] = t;
  }
}

// Writes each kept record at its rank (tile_off[tile] + rank in the tile):
// its column and weight, and beg for the rows from the previous record's row
// (exclusive) to its own (inclusive) when its row starts there; the last kept
// record also fills beg for the rows after it, up to n. tile_off has
// ntiles + 1 entries (the last is m); beg is zero beforehand (for m = 0).
__global__ __launch_bounds__(kBuildBlock) void k_tile_assemble(BuildShape s, const unsigned long long *keys,
                                                               const unsigned *idx,
                                                               const unsigned long long *tile_off, const int *src,
                                                               const int *dst, const int *w_in, long long *beg,
                                                               int *adj, int *w_out, Gap *gaps,
                                                               unsigned long long *ngaps) {
  const int shift = s.sorted ? s.cb : 0;
  const unsigned long long colmask = (1ull << s.cb) - 1;
  const long long t0 = (long long)blockIdx.x * kTile;
  const long long m = (long long)tile_off[gridDim.x];
  long long base = (long long)tile_off[blockIdx.x];
  const int wid = threadIdx.x / kWave, lane = lane_id();
  __shared__ int wcnt[kBuildBlock / kWave];
  for (int j = 0; j < kTileItems; ++j) {
    const long long i = t0 + (long long)j * kBuildBlock + threadIdx.x;
    const bool kp = i < s.records && record_kept(s, keys, i, shift);
    const unsigned long long bal = __ballot(kp);
    if (lane == 0) wcnt[wid] = __popcll(bal);
    __syncthreads();
    long long pos = base + mask_rank(bal);
    long long step = 0;
    for (int k = 0; k < kBuildBlock / kWave; ++k) {
      if (k < wid) pos += wcnt[k];
      step += wcnt[k];
    }
    if (kp) {
      const unsigned long long key = keys[i];
      const long long row = (long long)(key >> shift);
      // Clamped: ids outside [0, n) (unchecked on the device form) cannot
      // send beg writes out of range.
      const long long prev = i == 0 ? -1 : (long long)min(keys[i - 1] >> shift, (unsigned long long)s.n);
      if (row != prev) fill_rows(beg, prev + 1, row + 1, pos, gaps, ngaps);
      if (pos == m - 1) fill_rows(beg, row + 1, (long long)s.n + 1, m, gaps, ngaps);
      if (idx) {
        long long t;
        unsigned long long r, col;
        record_entry(s, src, dst, (long long)idx[i], t, r, col);
        adj[pos] = int(col);
        if (w_out) w_out[pos] = w_in[t];
      } else {
        adj[pos] = int(key & colmask);
      }
    }
    base += step;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBuildBlock) void k_fill_gaps(const Gap *gaps, const unsigned long long *ngaps,
                                                           long long *beg) {
  const long long ng = (long long)min(*ngaps, (unsigned long long)kGapCap);
  for (long long g = blockIdx.x; g < ng; g += gridDim.x) {
    const Gap x = gaps[g];
    for (long long v = x.v0 + threadIdx.x; v < x.v1; v += kBuildBlock) beg[v] = x.val;
  }
}

int grid_for(long long items, int cus) {
  return int(std::max<long long>(1, std::min<long long>((long long)cus * 16, (items + kBuildBlock - 1) / kBuildBlock)));
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) ignore(hipFree(p));
  }
  int alloc(size_t bytes, const char *what) {
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      ignore(hipGetLastError());
      p = nullptr;
      return set_error(ENT_ENOMEM, "hipMalloc of %zu bytes (%s) failed", bytes, what);
    }
    return ENT_OK;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// The whole construction on `stream`; w_in/w_out may be null. Returns m.
int build_csr(int device, hipStream_t stream, long long n, long long ne, const int *src, const int *dst,
              const int *w_in, int flags, long long *beg, int *adj, long long adj_cap, int *w_out,
              long long *m_out, ent_build_stats *st) {
  if (n < 0 || n > INT32_MAX || ne < 0) return set_error(ENT_EINVAL, "bad sizes n=%lld ne=%lld", n, ne);
  if (flags & ~(ENT_CSR_REVERSE | ENT_CSR_NO_LOOPS | ENT_CSR_SORT | ENT_CSR_DEDUP))
    return set_error(ENT_EINVAL, "unknown flags 0x%x", flags);
  BuildShape s{};
  s.ne = ne;
  s.rev = (flags & ENT_CSR_REVERSE) != 0;
  s.records = ne * (s.rev ? 2 : 1);
  s.n = int(n);
  s.no_loops = (flags & ENT_CSR_NO_LOOPS) != 0;
  s.dedup = (flags & ENT_CSR_DEDUP) != 0;
  s.sorted = s.dedup || (flags & ENT_CSR_SORT) != 0;
  s.cb = bits_for(uint64_t(std::max<long long>(n - 1, 0)));
  if (s.records >= (1ll << 32)) return set_error(ENT_EINVAL, "%lld records: at most 2^32 - 1", s.records);
  if (adj_cap < s.records) return set_error(ENT_EINVAL, "adjacency capacity %lld < %lld records", adj_cap, s.records);
  const int kbits = s.sorted ? s.cb + bits_for(uint64_t(n)) : bits_for(uint64_t(n));
  if (kbits > 64) return set_error(ENT_EINVAL, "key of %d bits", kbits);
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const bool keys_only = s.sorted && !w_in;
  const size_t R = size_t(s.records);
  DevBuf k0, k1, i0, i1, cnt, tmp, gaps;
  int rc;
  if ((rc = k0.alloc(8 * R, "keys")) || (rc = k1.alloc(8 * R, "keys")) || (rc = cnt.alloc(16, "counts")) ||
      (rc = gaps.alloc(sizeof(Gap) * kGapCap, "row gaps")))
    return rc;
  if (!keys_only && ((rc = i0.alloc(4 * R, "record indices")) || (rc = i1.alloc(4 * R, "record indices")))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      ignore(hipEventDestroy(a));
      ignore(hipEventDestroy(b));
    }
  } evg{e0, e1};
  rocprim::double_buffer<unsigned long long> keys(k0.as<unsigned long long>(), k1.as<unsigned long long>());
  rocprim::double_buffer<unsigned> vals(i0.as<unsigned>(), i1.as<unsigned>());
  // Scratch for the sort and the scan of tile counts, sized once.
  const long long ntiles = std::max<long long>(1, (s.records + kTile - 1) / kTile);
  DevBuf tiles;
  if ((rc = tiles.alloc(8 * size_t(2 * ntiles + 1), "tile counts"))) return rc;
  unsigned long long *tile_cnt = tiles.as<unsigned long long>(), *tile_off = tile_cnt + ntiles;
  size_t tb_sort = 0, tb_scan = 0;
  if (keys_only)
    HIP_TRY(rocprim::radix_sort_keys(nullptr, tb_sort, keys, R, 0u, unsigned(kbits), stream));
  else
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb_sort, keys, vals, R, 0u, unsigned(kbits), stream));
  HIP_TRY(rocprim::inclusive_scan(nullptr, tb_scan, tile_cnt, tile_off + 1, size_t(ntiles),
                                  rocprim::plus<unsigned long long>(), stream));
  if ((rc = tmp.alloc(std::max(tb_sort, tb_scan), "sort scratch"))) return rc;
  HIP_TRY(hipEventRecord(e0, stream));
  if (R) {
    hipLaunchKernelGGL(k_records, dim3(grid_for(s.records, cus)), dim3(kBuildBlock), 0, stream, s, src, dst,
                       keys.current(), keys_only ? nullptr : vals.current());
    HIP_TRY(hipGetLastError());
    size_t tb = tb_sort;
    if (keys_only)
      HIP_TRY(rocprim::radix_sort_keys(tmp.p, tb, keys, R, 0u, unsigned(kbits), stream));
    else
      HIP_TRY(rocprim::radix_sort_pairs(tmp.p, tb, keys, vals, R, 0u, unsigned(kbits), stream));
  }
  // Kept records per tile, their running sum, then every kept record at its
  // rank. beg starts at zero: with no kept record nothing else writes it.
  unsigned long long *ngaps = cnt.as<unsigned long long>() + 1;
  HIP_TRY(hipMemsetAsync(ngaps, 0, 8, stream));
  HIP_TRY(hipMemsetAsync(tiles.p, 0, 8 * size_t(2 * ntiles + 1), stream));
  HIP_TRY(hipMemsetAsync(beg, 0, 8 * size_t(n + 1), stream));
  if (R) {
    hipLaunchKernelGGL(k_tile_count, dim3(unsigned(ntiles)), dim3(kBuildBlock), 0, stream, s, keys.current(),
                       tile_cnt);
    HIP_TRY(hipGetLastError());
    size_t tb = tb_scan;
    HIP_TRY(rocprim::inclusive_scan(tmp.p, tb, tile_cnt, tile_off + 1, size_t(ntiles),
                                    rocprim::plus<unsigned long long>(), stream));
    hipLaunchKernelGGL(k_tile_assemble, dim3(unsigned(ntiles)), dim3(kBuildBlock), 0, stream, s, keys.current(),
                       keys_only ? nullptr : vals.current(), tile_off, src, dst, w_in, beg, adj, w_out,
                       gaps.as<Gap>(), ngaps);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fill_gaps, dim3(cus * 4), dim3(kBuildBlock), 0, stream, gaps.as<Gap>(), ngaps, beg);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, stream));
  long long m = 0;
  HIP_TRY(hipMemcpyAsync(&m, beg + n, 8, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  *m_out = m;
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->records = s.records;
    st->m = m;
    st->key_bits = kbits;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) st->device_ms = ms;
  }
  return ENT_OK;
}

// Tuple ids are checked on the host: the kernels index rows by them.
int check_tuples(int64_t n, int64_t ne, const int32_t *src, const int32_t *dst) {
  if (n < 0 || n > INT32_MAX || ne < 0) return set_error(ENT_EINVAL, "bad sizes n=%lld ne=%lld", (long long)n, (long long)ne);
  std::atomic<int64_t> bad{-1};
  parallel_for(ne, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) {
        bad.store(i);
        return;
      }
  });
  const int64_t i = bad.load();
  if (i >= 0)
    return set_error(ENT_EINVAL, "tuple %lld (%d, %d) outside [0, %lld)", (long long)i, src[i], dst[i], (long long)n);
  return ENT_OK;
}

// Copies host tuples (and weights) to new device buffers on `stream`.
struct DevTuples {
  DevBuf s, d, w;
  int upload(hipStream_t stream, int64_t ne, const int32_t *src, const int32_t *dst, const int32_t *weight) {
    int rc;
    if ((rc = s.alloc(4 * size_t(ne), "tuples")) || (rc = d.alloc(4 * size_t(ne), "tuples"))) return rc;
    if (weight && (rc = w.alloc(4 * size_t(ne), "weights"))) return rc;
    if (ne) {
      HIP_TRY(hipMemcpyAsync(s.p, src, 4 * size_t(ne), hipMemcpyHostToDevice, stream));
      HIP_TRY(hipMemcpyAsync(d.p, dst, 4 * size_t(ne), hipMemcpyHostToDevice, stream));
      if (weight) HIP_TRY(hipMemcpyAsync(w.p, weight, 4 * size_t(ne), hipMemcpyHostToDevice, stream));
    }
    return ENT_OK;
  }
};

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) ignore(hipStreamDestroy(s));
  }
};

}  // namespace
}  // namespace ent

using namespace ent;

extern "C" {

int ent_csr_build_device(int32_t device, int64_t n, int64_t ne, const int32_t *src_dev, const int32_t *dst_dev,
                         int32_t flags, int64_t *beg_pos_dev, int32_t *adj_dev, int64_t adj_capacity,
                         int64_t *m_out, ent_build_stats *st) {
  clear_error();
  if (!beg_pos_dev || !m_out || (ne > 0 && (!src_dev || !dst_dev || !adj_dev)))
    return set_error(ENT_EINVAL, "NULL argument");
  HIP_TRY(hipSetDevice(device));
  StreamGuard sg;
  HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  long long m = 0;
  const int rc = build_csr(device, sg.s, n, ne, src_dev, dst_dev, nullptr, flags,
                           reinterpret_cast<long long *>(beg_pos_dev), adj_dev, adj_capacity, nullptr, &m, st);
  if (rc) return rc;
  *m_out = m;
  return ENT_OK;
}

int ent_csr_build(int32_t device, int64_t n, int64_t ne, const int32_t *src, const int32_t *dst,
                  const int32_t *weight, int32_t flags, ent_csr *out, int32_t **weight_out, ent_build_stats *st) {
  clear_error();
  if (!out || (ne > 0 && (!src || !dst))) return set_error(ENT_EINVAL, "NULL argument");
  int rc = check_tuples(n, ne, src, dst);
  if (rc) return rc;
  std::memset(out, 0, sizeof *out);
  if (weight_out) *weight_out = nullptr;
  HIP_TRY(hipSetDevice(device));
  StreamGuard sg;
  HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  const int64_t records = ne * ((flags & ENT_CSR_REVERSE) ? 2 : 1);
  DevTuples dt;
  DevBuf db, da, dwo;
  if ((rc = dt.upload(sg.s, ne, src, dst, weight)) || (rc = db.alloc(8 * size_t(n + 1), "offsets")) ||
      (rc = da.alloc(4 * size_t(records), "adjacency")) || (weight && (rc = dwo.alloc(4 * size_t(records), "weights"))))
    return rc;
  long long m = 0;
  if ((rc = build_csr(device, sg.s, n, ne, dt.s.as<int>(), dt.d.as<int>(), weight ? dt.w.as<int>() : nullptr, flags,
                      db.as<long long>(), da.as<int>(), records, weight ? dwo.as<int>() : nullptr, &m, st)))
    return rc;
  int64_t *beg = static_cast<int64_t *>(std::malloc(8 * size_t(n + 1)));
  int32_t *adj = static_cast<int32_t *>(std::malloc(4 * size_t(std::max<long long>(m, 1))));
  int32_t *wh = weight_out && weight ? static_cast<int32_t *>(std::malloc(4 * size_t(std::max<long long>(m, 1)))) : nullptr;
  auto fail = [&](int code) {
    std::free(beg);
    std::free(adj);
    std::free(wh);
    return code;
  };
  if (!beg || !adj || (weight_out && weight && !wh))
    return fail(set_error(ENT_ENOMEM, "out of host memory (%lld entries)", m));
  hipError_t e = hipMemcpyAsync(beg, db.p, 8 * size_t(n + 1), hipMemcpyDeviceToHost, sg.s);
  if (e == hipSuccess && m) e = hipMemcpyAsync(adj, da.p, 4 * size_t(m), hipMemcpyDeviceToHost, sg.s);
  if (e == hipSuccess && wh && m) e = hipMemcpyAsync(wh, dwo.p, 4 * size_t(m), hipMemcpyDeviceToHost, sg.s);
  if (e == hipSuccess) e = hipStreamSynchronize(sg.s);
  if (e != hipSuccess) return fail(set_error(ENT_EDEVICE, "copy of the CSR to the host failed: %s", hipGetErrorString(e)));
  out->n = n;
  out->m = m;
  out->beg_pos = beg;
  out->adj = adj;
  if (weight_out) *weight_out = wh;
  return ENT_OK;
}

int ent_graph_upload_edges(const int32_t *src, const int32_t *dst, int64_t n, int64_t ne, int32_t flags,
                           int32_t device, int32_t upload_flags, ent_graph **out) {
  clear_error();
  if (!out) return set_error(ENT_EINVAL, "out is NULL");
  *out = nullptr;
  if (ne > 0 && (!src || !dst)) return set_error(ENT_EINVAL, "NULL argument");
  if (n < 1) return set_error(ENT_EINVAL, "n=%lld < 1", (long long)n);
  int rc = check_tuples(n, ne, src, dst);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(device));
  StreamGuard sg;
  HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  const int64_t records = ne * ((flags & ENT_CSR_REVERSE) ? 2 : 1);
  DevTuples dt;
  DevBuf db, da;
  if ((rc = dt.upload(sg.s, ne, src, dst, nullptr)) || (rc = db.alloc(8 * size_t(n + 1), "offsets")) ||
      (rc = da.alloc(4 * size_t(records), "adjacency")))
    return rc;
  long long m = 0;
  if ((rc = build_csr(device, sg.s, n, ne, dt.s.as<int>(), dt.d.as<int>(), nullptr, flags, db.as<long long>(),
                      da.as<int>(), records, nullptr, &m, nullptr)))
    return rc;
  // The offsets come back for the host-side degree tables; the adjacency is
  // copied device to device into the handle's CSR.
  std::vector<int64_t> beg(size_t(n) + 1);
  HIP_TRY(hipMemcpyAsync(beg.data(), db.p, 8 * size_t(n + 1), hipMemcpyDeviceToHost, sg.s));
  HIP_TRY(hipStreamSynchronize(sg.s));
  return upload_graph(beg.data(), da.as<int32_t>(), true, n, m, device, upload_flags, out);
}

}  // extern "C"
