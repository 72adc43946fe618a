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
//   2. scan:    cstart = exclusive prefix sum of cnt (three-phase block scan);
//   3. scatter: perm[cstart[key] + slot] = position;
//   4. order:   each cell's perm segment is sorted ascending (segments hold
//               ~ppc entries; insertion sort, one thread per cell).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace pelemp {

constexpr int kScanBlock = 1024;

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

// Phase 1: per-block inclusive scan of cnt into out (exclusive form written)
// and block totals.
__global__ void __launch_bounds__(kScanBlock) scan_block_kernel(const int32_t* __restrict__ cnt,
                                                                 int32_t n, int32_t* __restrict__ out,
                                                                 int32_t* __restrict__ block_sum) {
  __shared__ int32_t sh[kScanBlock];
  const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  const int32_t v = (i < n) ? cnt[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int32_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;  // exclusive
  if (threadIdx.x == kScanBlock - 1) block_sum[blockIdx.x] = sh[threadIdx.x];
}

// Phase 2: exclusive scan of the block totals (single block, loops).
__global__ void __launch_bounds__(kScanBlock) scan_sums_kernel(int32_t* __restrict__ block_sum,
                                                                int32_t nblocks,
                                                                int32_t* __restrict__ total) {
  __shared__ int32_t sh[kScanBlock];
  int32_t carry = 0;
  for (int32_t base = 0; base < nblocks; base += kScanBlock) {
    const int32_t i = base + threadIdx.x;
    const int32_t v = (i < nblocks) ? block_sum[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int32_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblocks) block_sum[i] = carry + sh[threadIdx.x] - v;
    const int32_t tot = sh[kScanBlock - 1];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

// Phase 3: add block offsets; cstart[n] = total.
__global__ void __launch_bounds__(kScanBlock) scan_add_kernel(int32_t* __restrict__ out, int32_t n,
                                                               const int32_t* __restrict__ block_sum,
                                                               const int32_t* __restrict__ total) {
  const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n) out[i] += block_sum[blockIdx.x];
  if (i == 0) out[n] = *total;
}

__global__ void __launch_bounds__(256) scatter_perm_kernel(const uint32_t* __restrict__ key,
                                                            const int32_t* __restrict__ n_ptr,
                                                            const int32_t* __restrict__ arrival,
                                                            const int32_t* __restrict__ cstart,
                                                            int32_t* __restrict__ perm) {
  const int32_t n = *n_ptr;
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

// cnt = mov_lo + stay_cnt + mov_hi, per cell
__global__ void __launch_bounds__(256) sum3_kernel(const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                    const int32_t* __restrict__ c, int32_t n,
                                                    int32_t* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i] + c[i];
}

__global__ void __launch_bounds__(256) scatter_incr_kernel(const uint32_t* __restrict__ key,
                                                            const int32_t* __restrict__ n_ptr,
                                                            const int32_t* __restrict__ arrival,
                                                            const int32_t* __restrict__ cstart,
                                                            const int32_t* __restrict__ mov_lo,
                                                            const int32_t* __restrict__ stay_cnt,
                                                            int32_t* __restrict__ perm) {
  const int32_t n = *n_ptr;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const uint32_t k = key[s];
    if (k == kDeadKey) continue;
    const int32_t v = arrival[s];
    const int cls = v >> 30;
    const int32_t r = v & kArrRankMask;
    int32_t p = cstart[k];
    if (cls == 0) {
      p += mov_lo[k] + r;
    } else if (cls == 1) {
      p += r;
    } else {
      p += mov_lo[k] + stay_cnt[k] + r;
    }
    perm[p] = s;
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

// order the mover groups of each cell by storage position
__global__ void __launch_bounds__(256) order_movers_kernel(const int32_t* __restrict__ cstart,
                                                            const int32_t* __restrict__ mov_lo,
                                                            const int32_t* __restrict__ stay_cnt,
                                                            const int32_t* __restrict__ mov_hi, int32_t ncells,
                                                            int32_t* __restrict__ perm) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t lo = mov_lo[c];
  const int32_t hi = mov_hi[c];
  if (lo < 2 && hi < 2) return;
  const int32_t a = cstart[c];
  if (lo > 1) insertion_sort_global(perm, a, a + lo);
  if (hi > 1) {
    const int32_t h0 = a + lo + stay_cnt[c];
    insertion_sort_global(perm, h0, h0 + hi);
  }
}

// Sort each cell's segment of perm ascending (insertion sort in place; a
// segment holds the droplets of one cell, ~ppc entries).
__global__ void __launch_bounds__(256) order_cells_kernel(const int32_t* __restrict__ cstart,
                                                           int32_t ncells,
                                                           int32_t* __restrict__ perm) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t a = cstart[c];
  const int32_t b = cstart[c + 1];
  for (int32_t i = a + 1; i < b; ++i) {
    const int32_t x = perm[i];
    int32_t j = i - 1;
    while (j >= a && perm[j] > x) {
      perm[j + 1] = perm[j];
      --j;
    }
    perm[j + 1] = x;
  }
}

}  // namespace pelemp
