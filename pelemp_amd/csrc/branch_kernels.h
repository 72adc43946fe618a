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

// exclusive scan of the block sums (one workgroup, loops over chunks)
__global__ void __launch_bounds__(kScanBlock) sb_scan_sums_kernel(int64_t* __restrict__ block_sum, int32_t nblocks) {
  __shared__ int64_t sh[kScanBlock];
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

// children of every record; ids[0] next id, ids[1] next slot, ids[2] the
// number of children (set by sb_total_kernel)
template <int NF>
__global__ void __launch_bounds__(256) sb_create_kernel(
    const int32_t* __restrict__ n_ptr, const int32_t* __restrict__ sb_type, const double* __restrict__ sb_data,
    const int64_t* __restrict__ offset, const int64_t* __restrict__ ids, FuelConst<double, NF> fc, double sub_dt,
    double ppp_fact, uint64_t seed, double* __restrict__ f, long long cap, int32_t* __restrict__ slot,
    int64_t* __restrict__ cold_id, long long slot_cap, uint32_t* __restrict__ key,
    int32_t* __restrict__ arrival, int32_t* __restrict__ mov_hi, SBGeom gm, int* err) {
  using F = Fields<NF>;
  const double* plo = gm.plo;
  const double* phi = gm.phi;
  const double* dxi = gm.dxi;
  const int* bndry_lo = gm.bndry_lo;
  const int* ncell = gm.ncell;
  const int* ntile = gm.ntile;
  const int32_t n = *n_ptr;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int ty = sb_type[s];
    if (ty == kSBNone) continue;
    SBRec<NF> r;
    r.type = ty;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      r.norm[q] = sb_data[(0 + q) * cap + s];
      r.vel[q] = sb_data[(3 + q) * cap + s];
      r.loc[q] = sb_data[(6 + q) * cap + s];
    }
    r.T0 = sb_data[9 * cap + s];
    r.ref_dia = sb_data[10 * cap + s];
    r.num_dens = sb_data[11 * cap + s];
    r.phi1 = sb_data[12 * cap + s];
    r.phi2 = sb_data[13 * cap + s];
    r.phi3 = sb_data[14 * cap + s];
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) r.Y0[spf] = sb_data[(kSBFields + spf) * cap + s];
    const int nc = sb_child_count<NF>(r, ppp_fact);
    const int64_t parent_id = cold_id[slot[s]];
    for (int k = 0; k < nc; ++k) {
      const long long rank = offset[s] + k;
      const long long q = static_cast<long long>(n) + rank;
      const long long sl = ids[1] + rank;
      if (q >= cap || sl >= slot_cap) {
        set_err(err, PELEMP_E_NOMEM);
        break;
      }
      BDrop<NF> c;
      sb_child<NF>(r, k, parent_id, fc, sub_dt, ppp_fact, seed, c);
      // no valid position (SpraySB.cpp:187-190): a KH-stripped wall film has
      // no velocity, so its children's directions are 0/0
      if (!(isfinite(c.pos[0]) && isfinite(c.pos[1]) && isfinite(c.pos[2]))) set_err(err, PELEMP_E_BAD_CHILD);
      // periodic wrap of the new position (the Redistribute stand-in)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (bndry_lo[d] == 0) c.pos[d] = wrap_periodic(c.pos[d], plo[d], phi[d]);
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
      cold_id[sl] = ids[0] + rank;
      int cc[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int ic = static_cast<int>(floor((c.pos[d] - plo[d]) * dxi[d]));
        ic = ic < 0 ? 0 : (ic >= ncell[d] ? ncell[d] - 1 : ic);
        cc[d] = ic;
      }
      const uint32_t kk = tile_key(cc[0], cc[1], cc[2], ntile);
      key[q] = kk;
      arrival[q] = atomicAdd(&mov_hi[kk], 1) | (2 << 30);
    }
  }
}

// after sb_create: advance the id / slot counters and the entry count
__global__ void sb_total_kernel(const int32_t* __restrict__ n_ptr, const int64_t* __restrict__ offset,
                                const int32_t* __restrict__ count, int64_t* __restrict__ ids,
                                int32_t* __restrict__ n_new, long long cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int32_t n = *n_ptr;
  const long long total = (n > 0) ? offset[n - 1] + count[n - 1] : 0;
  const long long m = static_cast<long long>(n) + total;
  ids[0] += total;
  ids[1] += total;
  ids[2] = total;
  *n_new = static_cast<int32_t>(m < cap ? m : cap);
}

}  // namespace pelemp
