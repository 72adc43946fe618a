// sort_kernels.h -- deterministic stable counting sort of droplets by cell.
//
// The droplet update processes droplets tile by tile in cell order and sums
// each cell's source contributions in that order, which replaces the
// reference's unordered FP64 atomics (SprayParticles.cpp:426-441).  After
// each pass the droplets carry a new cell key; this sort produces, for the
// next pass, the permutation `perm` (sorted position -> storage position)
// and the CSR offsets `cstart[key]` of every cell.  The order within a cell
// is ascending storage position (a stable sort), which does not depend on
// the order in which the GPU threads run:
//   1. count:   cnt[key]++ (integer atomics), remembering each droplet's
//               arrival slot;
//   2. scan:    cstart = exclusive prefix sum of cnt (block scans, then each
//               block adds the sum of the block totals before it);
//   3. scatter: perm[cstart[key] + slot] = position;
//   4. order:   each cell's perm segment is sorted ascending: segments of up
//               to kSeqSortMax entries (a few droplets per cell) by one
//               thread (insertion sort, or a sorting network in registers);
//               longer ones (droplets
//               injected into the few cells of a nozzle, breakup children
//               of one cell) are listed and sorted by a workgroup each
//               (sort_long_groups_kernel), so no segment costs O(m^2) on
//               one thread.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace pelemp {

constexpr int kScanBlock = 1024;    // cells per block of the cell-count scan
constexpr int kScanThreads = 256;   // its threads: four cells each

__global__ void __launch_bounds__(256) count_keys_kernel(const uint32_t* __restrict__ key,
                                                          const int32_t* __restrict__ n_ptr,
                                                          int32_t* __restrict__ cnt,
                                                          int32_t* __restrict__ arrival) {
  const int32_t n = *n_ptr;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t k = key[s];
    if (k != kDeadKey) arrival[s] = atomicAdd(&cnt[k], 1);
  }
}

// Phase 1: per-block inclusive scan of the per-cell counts into out
// (exclusive form written) and block totals.  The count of a cell is c0, or
// c0 + c1 + c2 when c1 is given (the three groups of an incremental pass).
// With n_entries given, thread 0 also sets the entry count of the new buffer
// (n_src, or the previous sort's total out[n]) and logs the previous total
// (the droplets at the pass start) into log_slot when that is given.
__global__ void __launch_bounds__(kScanThreads) scan_block_kernel(const int32_t* __restrict__ c0,
                                                                   const int32_t* __restrict__ c1,
                                                                   const int32_t* __restrict__ c2, int32_t n,
                                                                   int32_t* __restrict__ out,
                                                                   int32_t* __restrict__ block_sum,
                                                                   const int32_t* __restrict__ n_src,
                                                                   int32_t* __restrict__ n_entries,
                                                                   int32_t* __restrict__ log_slot) {
  static_assert(kScanBlock == 4 * kScanThreads, "four cells per thread");
  __shared__ int32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int32_t i0 = blockIdx.x * kScanBlock + 4 * threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0 && n_entries) {
    const int32_t pt = out[n];
    *n_entries = n_src ? *n_src : pt;
    if (log_slot) *log_slot = pt;
  }
  int32_t v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int32_t i = i0 + q;
    v[q] = (i < n) ? (c1 ? c0[i] + c1[i] + c2[i] : c0[i]) : 0;
  }
  const int32_t tsum = v[0] + v[1] + v[2] + v[3];
  // inclusive scan of the thread sums over the wavefront, then over the
  // block's wavefronts
  int32_t x = tsum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int32_t wbase = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) wbase += (w < wave) ? wsum[w] : 0;
  int32_t r = wbase + x - tsum;  // exclusive prefix of this thread's first cell
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (i0 + q < n) out[i0 + q] = r;
    r += v[q];
  }
  if (threadIdx.x == kScanThreads - 1) block_sum[blockIdx.x] = r;
}

// Phase 2: each block adds the sum of the preceding blocks' totals; the last
// block writes the grand total to out[n] and *total.  Also clears n_long.
__global__ void __launch_bounds__(kScanThreads) scan_add_kernel(int32_t* __restrict__ out, int32_t n,
                                                                 const int32_t* __restrict__ block_sum,
                                                                 int32_t nblocks, int32_t* __restrict__ total,
                                                                 int32_t* __restrict__ n_long) {
  __shared__ int32_t red[kScanThreads / 64];
  int32_t part = 0;
  for (int32_t b = threadIdx.x; b < static_cast<int32_t>(blockIdx.x); b += kScanThreads) part += block_sum[b];
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  int32_t base = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) base += red[w];
  const int32_t i0 = blockIdx.x * kScanBlock + 4 * threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (i0 + q < n) out[i0 + q] += base;
  if (threadIdx.x == 0 && static_cast<int32_t>(blockIdx.x) == nblocks - 1) {
    const int32_t tot = base + block_sum[nblocks - 1];
    out[n] = tot;
    *total = tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_long = 0;
}

__global__ void __launch_bounds__(256) scatter_perm_kernel(const uint32_t* __restrict__ key,
                                                            const int32_t* __restrict__ n_ptr,
                                                            const int32_t* __restrict__ arrival,
                                                            const int32_t* __restrict__ cstart,
                                                            int32_t* __restrict__ perm,
                                                            int32_t* __restrict__ split_counters) {
  const int32_t n = *n_ptr;
  // the split lists are rebuilt after the sort
  if (blockIdx.x == 0 && threadIdx.x < 3) split_counters[threadIdx.x] = 0;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t k = key[s];
    if (k != kDeadKey) perm[cstart[k] + arrival[s]] = s;
  }
}

// ---------------------------------------------------------------------------
// Incremental re-sort after an update pass.  The update kernel has already
// classified every droplet at storage position s of the new buffer:
//   stayer (same cell as at the pass start): arrival = rank among the
//     stayers of its cell, in position order (computed by the cell owner);
//   mover into a cell with a larger key than its old cell ("from lower"):
//     arrival = (1 << 30) | atomic slot in mov_lo[key];
//   mover into a cell with a smaller key ("from higher"):
//     arrival = (2 << 30) | atomic slot in mov_hi[key].
// Storage positions are sorted by old key, so within a new cell the stable
// order is [movers from lower keys][stayers][movers from higher keys]; only
// the two mover groups (usually empty or one droplet) need ordering.
constexpr int32_t kArrRankMask = (1 << 30) - 1;

// One thread per tile after the sort: a tile of n > kSegDroplets droplets
// gets ceil(n / kSegDroplets) consecutive segment slots and its segments
// 1.. as work items (the order of the lists does not matter: every entry
// names its tile and segment).
__device__ inline void split_tile(const int32_t* __restrict__ cstart, int32_t t, const SplitDev& sp) {
  const int32_t n = cstart[(t + 1) * 64] - cstart[t * 64];
  const int32_t nseg = (n + kSegDroplets - 1) / kSegDroplets;
  if (nseg > 1) {
    const int32_t slot = atomicAdd(&sp.counters[2], nseg);
    const int32_t it = atomicAdd(&sp.counters[0], nseg - 1);
    const int32_t li = atomicAdd(&sp.counters[1], 1);
    sp.tile_slot[t] = slot;
    for (int32_t k = 1; k < nseg; ++k) sp.items[it + k - 1] = make_int2(t, k);
    sp.list[li] = make_int4(t, slot, nseg, 0);
  } else {
    sp.tile_slot[t] = -1;
  }
}

__global__ void __launch_bounds__(256) split_tiles_kernel(const int32_t* __restrict__ cstart, int32_t ntiles,
                                                           SplitDev sp) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) split_tile(cstart, t, sp);
}

__global__ void __launch_bounds__(256) scatter_incr_kernel(const uint32_t* __restrict__ key,
                                                            const int32_t* __restrict__ n_ptr,
                                                            const int32_t* __restrict__ arrival,
                                                            const int32_t* __restrict__ cstart,
                                                            const int32_t* __restrict__ mov_lo,
                                                            const int32_t* __restrict__ stay_cnt,
                                                            int32_t* __restrict__ perm, SplitDev sp) {
  const int32_t n = *n_ptr;
  // the split lists of the next pass are rebuilt after this kernel
  if (blockIdx.x == 0 && threadIdx.x < 3) sp.counters[threadIdx.x] = 0;
  // kU entries per thread and round, their loads issued together: each entry
  // is a chain of dependent loads (key and arrival, then the cell's offsets)
  constexpr int kU = 8;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t s0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s0 < n; s0 += kU * stride) {
    uint32_t k[kU];
    int32_t v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t s = s0 + u * stride;
      k[u] = (s < n) ? key[s] : kDeadKey;
      v[u] = (s < n) ? arrival[s] : 0;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (k[u] == kDeadKey) continue;
      const int cls = v[u] >> 30;
      const int32_t r = v[u] & kArrRankMask;
      int32_t p = cstart[k[u]];
      if (cls == 0) {
        // a stayer: its rank among the stayers of its segment of a heavy tile
        const int32_t seg = r >> kSegShift;
        p += mov_lo[k[u]] + (r & kSegRankMask);
        if (seg > 0) p += sp.seg_stay[(sp.tile_slot[k[u] / 64] + seg) * 64 + (k[u] % 64)];
      } else if (cls == 1) {
        p += r;
      } else {
        p += mov_lo[k[u]] + stay_cnt[k[u]] + r;
      }
      perm[p] = static_cast<int32_t>(s0 + u * stride);
    }
  }
}

__device__ inline void insertion_sort_global(int32_t* __restrict__ v, int32_t a, int32_t b) {
  for (int32_t i = a + 1; i < b; ++i) {
    const int32_t x = v[i];
    int32_t j = i - 1;
    while (j >= a && v[j] > x) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

// Segments of up to kSeqSortMax entries are sorted by their thread, up to
// kWaveSortMax by their wavefront; longer ones are listed and sorted by a
// workgroup each.  The list holds at most n / (kWaveSortMax + 1) entries.
constexpr int32_t kSeqSortMax = 16;
constexpr int32_t kWaveSortMax = 64;  // segments up to this length: sorted by the wavefront that finds them
constexpr int kLongSortThreads = 256;
constexpr int kLongSortChunk = 4096;  // most entries sorted in LDS at a time
constexpr int kLongSortGrid = 1024;   // workgroups working through the list

// perm[a, a + len) for len <= N sorted in registers: all loads issued at
// once, a bitonic network with compile-time indices, the stores
template <int N>
__device__ inline void sort_short(int32_t* __restrict__ perm, int32_t a, int32_t len) {
  int32_t v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = (q < len) ? perm[a + q] : INT32_MAX;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = v[i], y = v[l];
          const bool up = (i & k) == 0;
          v[i] = up ? min(x, y) : max(x, y);
          v[l] = up ? max(x, y) : min(x, y);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (q < len) perm[a + q] = v[q];
}

// perm[a, a + m) for m <= 64 sorted by the whole wavefront (uniform a, m):
// one entry per lane, a bitonic network over lanes
__device__ inline void wave_sort64(int32_t* __restrict__ perm, int32_t a, int32_t m, int lane) {
  int32_t v = (lane < m) ? perm[a + lane] : INT32_MAX;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int32_t w = __shfl_xor(v, j);
      const bool lower = (lane & j) == 0;
      const bool up = (lane & k) == 0;
      v = (lower == up) ? min(v, w) : max(v, w);
    }
  }
  if (lane < m) perm[a + lane] = v;
}

// Sort a segment in place, or list a long one: every lane of the wavefront
// calls this (len 0 for none).  Segments of kSeqSortMax < len <=
// kWaveSortMax are taken one after the other by the whole wavefront; one
// atomic per wavefront reserves the list entries of its long segments.
__device__ inline void sort_or_defer(int32_t* __restrict__ perm, int32_t a, int32_t len, int2* __restrict__ long_list,
                                     int32_t* __restrict__ n_long) {
  if (len > 1 && len <= 4) insertion_sort_global(perm, a, a + len);
  else if (len > 4 && len <= 8) sort_short<8>(perm, a, len);
  else if (len > 8 && len <= kSeqSortMax) sort_short<16>(perm, a, len);
  const int lane = threadIdx.x & 63;
  unsigned long long wb = __ballot(len > kSeqSortMax && len <= kWaveSortMax);
  while (wb != 0ull) {
    const int src = __ffsll(static_cast<long long>(wb)) - 1;
    wave_sort64(perm, __shfl(a, src), __shfl(len, src), lane);
    wb &= wb - 1ull;
  }
  const bool is_long = len > kWaveSortMax;
  const unsigned long long bal = __ballot(is_long);
  if (bal == 0ull) return;
  const int leader = __ffsll(static_cast<long long>(bal)) - 1;
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(n_long, __popcll(bal));
  base = __shfl(base, leader);
  if (is_long) {
    const unsigned long long lower = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
    long_list[base + __popcll(lower)] = make_int2(a, len);
  }
}

// Sort each listed segment perm[a, a + m) ascending (m > kWaveSortMax), one
// workgroup each: chunks
// of up to kLongSortChunk entries by a bitonic network in LDS sized to the
// chunk (the next power of two), then merges of sorted runs of doubling
// width through tmp[a, a + m) (each thread merges one slice of the output,
// found by a merge-path search).  Entries are distinct storage positions
// below INT32_MAX, which pads a network.
__global__ void __launch_bounds__(kLongSortThreads) sort_long_groups_kernel(const int2* __restrict__ long_list,
                                                                             const int32_t* __restrict__ n_long,
                                                                             int32_t* __restrict__ perm,
                                                                             int32_t* __restrict__ tmp) {
  __shared__ int32_t sh[kLongSortChunk];
  const int32_t nl = *n_long;
  const int tid = threadIdx.x;
  for (int32_t g = blockIdx.x; g < nl; g += gridDim.x) {
    const int32_t a = long_list[g].x;
    const int32_t m = long_list[g].y;
    for (int32_t c0 = 0; c0 < m; c0 += kLongSortChunk) {
      const int32_t len = min(kLongSortChunk, m - c0);
      int P = 128;
      while (P < len) P <<= 1;
      for (int e = tid; e < P; e += kLongSortThreads) sh[e] = (e < len) ? perm[a + c0 + e] : INT32_MAX;
      __syncthreads();
      for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int e = tid; e < P / 2; e += kLongSortThreads) {
            const int i = 2 * e - (e & (j - 1));  // e with a zero bit inserted at log2(j)
            const int l = i + j;
            const bool up = (i & k) == 0;
            const int32_t x = sh[i], y = sh[l];
            if ((x > y) == up) {
              sh[i] = y;
              sh[l] = x;
            }
          }
          __syncthreads();
        }
      }
      for (int e = tid; e < len; e += kLongSortThreads) perm[a + c0 + e] = sh[e];
      __syncthreads();
    }
    int32_t* src = perm + a;
    int32_t* dst = tmp + a;
    for (int32_t w = kLongSortChunk; w < m; w <<= 1) {
      for (int32_t lo = 0; lo < m; lo += 2 * w) {
        const int32_t mid = min(lo + w, m), hi = min(lo + 2 * w, m);
        const int32_t la = mid - lo, lb = hi - mid, tot = hi - lo;
        const int32_t per = (tot + kLongSortThreads - 1) / kLongSortThreads;
        const int32_t o0 = min(tid * per, tot), o1 = min(o0 + per, tot);
        if (o0 < o1) {
          const int32_t* A = src + lo;
          const int32_t* B = src + mid;
          // i = entries of A among the first o0 outputs
          int32_t ilo = max(0, o0 - lb), ihi = min(o0, la);
          while (ilo < ihi) {
            const int32_t i = (ilo + ihi) >> 1;
            if (A[i] < B[o0 - i - 1])
              ilo = i + 1;
            else
              ihi = i;
          }
          int32_t i = ilo, j = o0 - ilo;
          for (int32_t o = o0; o < o1; ++o) {
            if (j >= lb || (i < la && A[i] < B[j]))
              dst[lo + o] = A[i++];
            else
              dst[lo + o] = B[j++];
          }
        }
      }
      __syncthreads();
      int32_t* t = src;
      src = dst;
      dst = t;
    }
    if (src != perm + a)
      for (int32_t e = tid; e < m; e += kLongSortThreads) perm[a + e] = src[e];
    __syncthreads();
  }
}

// order the mover groups of each cell by storage position; the cell's
// counters are cleared for the next pass
__global__ void __launch_bounds__(256) order_movers_kernel(const int32_t* __restrict__ cstart,
                                                            int32_t* __restrict__ mov_lo,
                                                            int32_t* __restrict__ stay_cnt,
                                                            int32_t* __restrict__ mov_hi, int32_t ncells,
                                                            int32_t* __restrict__ perm, int2* __restrict__ long_list,
                                                            int32_t* __restrict__ n_long, int32_t ntiles,
                                                            SplitDev sp) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  // the next pass's heavy-tile split (split_tiles_kernel's work, in this
  // launch: cstart is final and scatter_incr has read the old lists)
  if (c < ntiles) split_tile(cstart, c, sp);
  int32_t lo = 0, hi = 0, a = 0, h0 = 0;
  if (c < ncells) {
    lo = mov_lo[c];
    hi = mov_hi[c];
    if (lo > 1 || hi > 1) {
      a = cstart[c];
      h0 = a + lo + stay_cnt[c];
    }
    // the counters start the next pass at zero
    mov_lo[c] = 0;
    stay_cnt[c] = 0;
    mov_hi[c] = 0;
  }
  sort_or_defer(perm, a, lo, long_list, n_long);
  sort_or_defer(perm, h0, hi, long_list, n_long);
}

// Sort each cell's segment of perm ascending (a segment holds the droplets
// of one cell); the cell's pass counters (mov_lo, stay_cnt, mov_hi, each
// ncells long from `counters`) are cleared for the next pass.
__global__ void __launch_bounds__(256) order_cells_kernel(const int32_t* __restrict__ cstart,
                                                           int32_t ncells,
                                                           int32_t* __restrict__ perm, int2* __restrict__ long_list,
                                                           int32_t* __restrict__ n_long,
                                                           int32_t* __restrict__ counters) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  int32_t a = 0, len = 0;
  if (c < ncells) {
    a = cstart[c];
    len = cstart[c + 1] - a;
    counters[c] = 0;
    counters[ncells + c] = 0;
    counters[2 * static_cast<size_t>(ncells) + c] = 0;
  }
  sort_or_defer(perm, a, len, long_list, n_long);
}

}  // namespace pelemp
