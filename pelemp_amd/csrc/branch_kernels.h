// branch_kernels.h -- the passes around the update kernel when films,
// breakup or splash are on (FP64 form):
//  * film_fab_kernel: film thickness of every cell, the sum over the film
//    droplets of the cell in cell-sorted order (fillFilmFab, WallFilm.H:9-26
//    and SprayParticles.cpp:275-288, which sums with atomics);
//  * sb_count_kernel / sb_create_kernel: the secondary droplets of the
//    records the update wrote (CreateSBDroplets, SpraySB.cpp:9-207).  The
//    children of the record at output position s are appended at
//    n + offset(s) + k (offset = exclusive scan of the child counts over s),
//    so the order, ids and slots of new droplets do not depend on thread
//    timing; each child registers as an arrival from a higher position in
//    its cell, which places it after every existing droplet of the cell.
#pragma once

#include <hip/hip_runtime.h>

#include "branches.h"
#include "layout.h"
#include "update_kernel.h"

namespace pelemp {

template <int NF>
__global__ void __launch_bounds__(256) film_fab_kernel(const int32_t* __restrict__ cstart,
                                                       const int32_t* __restrict__ perm,
                                                       const double* __restrict__ f, long long cap, int nx, int ny,
                                                       int nz, int ntx, int nty, double face_area,
                                                       double* __restrict__ wf) {
  using F = Fields<NF>;
  const int key = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncells = nx * ny * nz;
  if (key >= ncells) return;
  const int tile = key / kTileCells, lc = key % kTileCells;
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  const int cx = tx * kTile + lc % kTile, cy = ty * kTile + (lc / kTile) % kTile, cz = tz * kTile + lc / (kTile * kTile);
  double v = 0.;
  for (int32_t i = cstart[key]; i < cstart[key + 1]; ++i) {
    const int32_t o = perm[i];
    const double film = f[F::FILM * cap + o];
    if (film > 0.) v += film_fab_value(film, f[F::DIA * cap + o], face_area);
  }
  wf[(static_cast<long long>(cz) * ny + cy) * nx + cx] = v;
}

template <int NF>
__global__ void __launch_bounds__(256) sb_count_kernel(const int32_t* __restrict__ n_ptr,
                                                       const int32_t* __restrict__ sb_type,
                                                       const double* __restrict__ sb_data, long long cap,
                                                       double ppp_fact, int32_t* __restrict__ count) {
  const int32_t n = *n_ptr;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int ty = sb_type[s];
    int c = 0;
    if (ty >= kSBSplashDry) {
      c = 4;
    } else if (ty != kSBNone) {
      SBRec<NF> r;
      r.type = ty;
      r.num_dens = sb_data[11 * cap + s];
      c = sb_child_count<NF>(r, ppp_fact);
    }
    count[s] = c;
  }
}

// Exclusive scan out = scan(in) over [0, n) in 64-bit (chunks of 1024,
// block sums, adds): child counts can be large, so no 32-bit offsets.
__global__ void __launch_bounds__(kScanBlock) sb_scan_block_kernel(const int32_t* __restrict__ in,
                                                                    int64_t* __restrict__ v,
                                                                    const int32_t* __restrict__ n_ptr,
                                                                    int64_t* __restrict__ block_sum) {
  __shared__ int64_t sh[kScanBlock];
  const int32_t n = *n_ptr;
  const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  if (static_cast<int64_t>(blockIdx.x) * kScanBlock >= n) {
    if (threadIdx.x == 0) block_sum[blockIdx.x] = 0;
    return;
  }
  const int64_t x = (i < n) ? in[i] : 0;
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int64_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) v[i] = sh[threadIdx.x] - x;
  if (threadIdx.x == kScanBlock - 1) block_sum[blockIdx.x] = sh[threadIdx.x];
}

// exclusive scan of the block sums of the n_ptr entries (one workgroup,
// loops over chunks)
__global__ void __launch_bounds__(kScanBlock) sb_scan_sums_kernel(int64_t* __restrict__ block_sum,
                                                                  const int32_t* __restrict__ n_ptr) {
  __shared__ int64_t sh[kScanBlock];
  const int32_t nblocks = (*n_ptr + kScanBlock - 1) / kScanBlock;
  int64_t carry = 0;
  for (int32_t base = 0; base < nblocks; base += kScanBlock) {
    const int32_t i = base + threadIdx.x;
    const int64_t x = (i < nblocks) ? block_sum[i] : 0;
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int64_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblocks) block_sum[i] = carry + sh[threadIdx.x] - x;
    const int64_t tot = sh[kScanBlock - 1];
    __syncthreads();
    carry += tot;
  }
}

__global__ void __launch_bounds__(kScanBlock) sb_scan_add_kernel(int64_t* __restrict__ v,
                                                                  const int32_t* __restrict__ n_ptr,
                                                                  const int64_t* __restrict__ block_sum) {
  const int32_t n = *n_ptr;
  const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n) v[i] += block_sum[blockIdx.x];
}

// level geometry as the child pass needs it (passed by value)
struct SBGeom {
  double plo[3], phi[3], dxi[3];
  int bndry_lo[3], ncell[3], ntile[3];
};

// Children pass, one thread per child.  ids[0] next id, ids[1] next slot,
// ids[2] children of this pass.  sb_total_kernel sets ids[2] and the entry
// count; the parent of child r is the last record position s with
// offset[s] <= r (a record with no children has the offset of its
// successor), found by marking each record's first child and an inclusive
// max-scan over the children.
__global__ void sb_total_kernel(const int32_t* __restrict__ n_ptr, const int64_t* __restrict__ offset,
                                const int32_t* __restrict__ count, int64_t* __restrict__ ids,
                                int32_t* __restrict__ n_new, long long cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t n = *n_ptr;
  const long long total = (n > 0) ? offset[n - 1] + count[n - 1] : 0;
  const long long m = static_cast<long long>(n) + total;
  ids[2] = total;
  *n_new = static_cast<int32_t>(m < cap ? m : cap);
}

__global__ void __launch_bounds__(256) sb_parent_init_kernel(const int64_t* __restrict__ ids, long long cap,
                                                             int32_t* __restrict__ parent) {
  const long long total = ids[2] < cap ? ids[2] : cap;
  for (long long r = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; r < total;
       r += static_cast<long long>(gridDim.x) * blockDim.x)
    parent[r] = -1;
}

__global__ void __launch_bounds__(256) sb_parent_mark_kernel(const int32_t* __restrict__ n_ptr,
                                                             const int32_t* __restrict__ count,
                                                             const int64_t* __restrict__ offset, long long cap,
                                                             int32_t* __restrict__ parent) {
  const int32_t n = *n_ptr;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
    if (count[s] > 0 && offset[s] < cap) parent[offset[s]] = s;
}

// inclusive max-scan of parent[0, total) in chunks of kScanBlock
__global__ void __launch_bounds__(kScanBlock) sb_maxscan_block_kernel(int32_t* __restrict__ v,
                                                                      const int64_t* __restrict__ ids,
                                                                      long long cap,
                                                                      int32_t* __restrict__ block_max) {
  __shared__ int32_t sh[kScanBlock];
  const long long total = ids[2] < cap ? ids[2] : cap;
  const long long i = static_cast<long long>(blockIdx.x) * kScanBlock + threadIdx.x;
  if (static_cast<long long>(blockIdx.x) * kScanBlock >= total) {
    if (threadIdx.x == 0) block_max[blockIdx.x] = -1;
    return;
  }
  sh[threadIdx.x] = (i < total) ? v[i] : -1;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int32_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : -1;
    __syncthreads();
    sh[threadIdx.x] = sh[threadIdx.x] > t ? sh[threadIdx.x] : t;
    __syncthreads();
  }
  if (i < total) v[i] = sh[threadIdx.x];
  if (threadIdx.x == kScanBlock - 1) block_max[blockIdx.x] = sh[threadIdx.x];
}

// exclusive max-scan of the block maxima of the children (one workgroup)
__global__ void __launch_bounds__(kScanBlock) sb_maxscan_sums_kernel(int32_t* __restrict__ block_max,
                                                                     const int64_t* __restrict__ ids, long long cap) {
  __shared__ int32_t sh[kScanBlock];
  const long long total = ids[2] < cap ? ids[2] : cap;
  const int32_t nblocks = static_cast<int32_t>((total + kScanBlock - 1) / kScanBlock);
  int32_t carry = -1;
  for (int32_t base = 0; base < nblocks; base += kScanBlock) {
    const int32_t i = base + threadIdx.x;
    const int32_t x = (i < nblocks) ? block_max[i] : -1;
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int32_t t = (threadIdx.x >= static_cast<unsigned>(off)) ? sh[threadIdx.x - off] : -1;
      __syncthreads();
      sh[threadIdx.x] = sh[threadIdx.x] > t ? sh[threadIdx.x] : t;
      __syncthreads();
    }
    const int32_t excl = (threadIdx.x == 0) ? -1 : sh[threadIdx.x - 1];
    if (i < nblocks) block_max[i] = carry > excl ? carry : excl;
    const int32_t tot = sh[kScanBlock - 1];
    __syncthreads();
    carry = carry > tot ? carry : tot;
  }
}

__global__ void __launch_bounds__(kScanBlock) sb_maxscan_add_kernel(int32_t* __restrict__ v,
                                                                    const int64_t* __restrict__ ids, long long cap,
                                                                    const int32_t* __restrict__ block_max) {
  const long long total = ids[2] < cap ? ids[2] : cap;
  const long long i = static_cast<long long>(blockIdx.x) * kScanBlock + threadIdx.x;
  if (i < total) {
    const int32_t b = block_max[blockIdx.x];
    v[i] = v[i] > b ? v[i] : b;
  }
}

// child r of the pass: record s = parent[r], k = r - offset[s].  The waves
// step through r together (lane = r mod 64) so that the arrival counters of
// a cell are advanced once per run of consecutive children in that cell
// (the children of one record mostly share a cell): one atomic per run.
template <int NF>
__global__ void __launch_bounds__(256) sb_create_kernel(
    const int32_t* __restrict__ n_ptr, const int32_t* __restrict__ sb_type, const double* __restrict__ sb_data,
    const int64_t* __restrict__ offset, const int32_t* __restrict__ parent, const int64_t* __restrict__ ids,
    FuelConst<double, NF> fc, double sub_dt, double ppp_fact, uint64_t seed, double* __restrict__ f, long long cap,
    int32_t* __restrict__ slot, int64_t* __restrict__ cold_id, long long slot_cap, uint32_t* __restrict__ key,
    int32_t* __restrict__ arrival, int32_t* __restrict__ mov_hi, SBGeom gm, int* err) {
  using F = Fields<NF>;
  const int32_t n = *n_ptr;
  const long long total = ids[2];
  const long long id0 = ids[0], slot0 = ids[1];
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long base = blockIdx.x * static_cast<long long>(blockDim.x) + (threadIdx.x & ~63u); base < total;
       base += stride) {
    const long long r = base + lane;
    const long long q = static_cast<long long>(n) + r;
    const long long sl = slot0 + r;
    bool act = r < total;
    if (act && (q >= cap || sl >= slot_cap)) {
      set_err(err, PELEMP_E_NOMEM);
      act = false;
    }
    uint32_t kk = 0xffffffffu;
    if (act) {
      const int32_t s = parent[r];
      const int k = static_cast<int>(r - offset[s]);
      SBRec<NF> rc;
      rc.type = sb_type[s];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        rc.norm[d] = sb_data[(0 + d) * cap + s];
        rc.vel[d] = sb_data[(3 + d) * cap + s];
        rc.loc[d] = sb_data[(6 + d) * cap + s];
      }
      rc.T0 = sb_data[9 * cap + s];
      rc.ref_dia = sb_data[10 * cap + s];
      rc.num_dens = sb_data[11 * cap + s];
      rc.phi1 = sb_data[12 * cap + s];
      rc.phi2 = sb_data[13 * cap + s];
      rc.phi3 = sb_data[14 * cap + s];
#pragma unroll
      for (int spf = 0; spf < NF; ++spf) rc.Y0[spf] = sb_data[(kSBFields + spf) * cap + s];
      const int64_t parent_id = cold_id[slot[s]];
      BDrop<NF> c;
      sb_child<NF>(rc, k, parent_id, fc, sub_dt, ppp_fact, seed, c);
      // no valid position (SpraySB.cpp:187-190): a KH-stripped wall film has
      // no velocity, so its children's directions are 0/0
      if (!(isfinite(c.pos[0]) && isfinite(c.pos[1]) && isfinite(c.pos[2]))) set_err(err, PELEMP_E_BAD_CHILD);
      // periodic wrap of the new position (the Redistribute stand-in)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (gm.bndry_lo[d] == 0) c.pos[d] = wrap_periodic(c.pos[d], gm.plo[d], gm.phi[d]);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        f[(F::POS + d) * cap + q] = c.pos[d];
        f[(F::VEL + d) * cap + q] = c.vel[d];
      }
      f[F::T * cap + q] = c.T;
      f[F::DIA * cap + q] = c.dia;
#pragma unroll
      for (int spf = 0; spf < NF; ++spf) f[(F::Y + spf) * cap + q] = c.Y[spf];
      f[F::ND * cap + q] = c.numDens;
      f[F::FILM * cap + q] = c.film;
      f[F::BM1 * cap + q] = c.bm1;
      f[F::BM2 * cap + q] = c.bm2;
      f[F::N0 * cap + q] = c.N0;
      slot[q] = static_cast<int32_t>(sl);
      cold_id[sl] = id0 + r;
      int cc[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int ic = static_cast<int>(floor((c.pos[d] - gm.plo[d]) * gm.dxi[d]));
        ic = ic < 0 ? 0 : (ic >= gm.ncell[d] ? gm.ncell[d] - 1 : ic);
        cc[d] = ic;
      }
      kk = tile_key(cc[0], cc[1], cc[2], gm.ntile);
      key[q] = kk;
    }
    // runs of equal keys in consecutive lanes: the first lane of a run takes
    // the run's arrival indices with one atomic and hands them out
    const uint32_t kprev = __shfl_up(kk, 1);
    const bool same = act && lane > 0 && kprev == kk;
    const unsigned long long same_mask = __ballot(same);
    const bool lead = act && !same;
    int32_t first = 0;
    if (lead) {
      const int run = 1 + ((lane == 63) ? 0 : __builtin_ctzll(~(same_mask >> (lane + 1))));
      first = atomicAdd(&mov_hi[kk], run);
    }
    const unsigned long long lead_mask = __ballot(lead);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long mine = lead_mask & upto;
    const int leader = mine ? 63 - __builtin_clzll(mine) : lane;
    const int32_t b = __shfl(first, leader);
    if (act) arrival[q] = (b + (lane - leader)) | (2 << 30);
  }
}

// after sb_create: advance the id and slot counters by the children made
__global__ void sb_advance_kernel(int64_t* __restrict__ ids) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ids[0] += ids[2];
  ids[1] += ids[2];
}

}  // namespace pelemp
