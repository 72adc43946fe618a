// inject_kernels.h -- device side of the injection and initialisation
// routines (SprayInjection.H:10-348) and of appending droplets.
//
// The droplet loop of sprayInjection (SprayInjection.H:78-158) draws
// candidates until the injected mass reaches the target; the mass is summed
// candidate by candidate, so the stopping candidate and the injected mass
// are the reference loop's own only if that FP64 sum keeps its order.  The
// loop runs in batches of candidates, three launches each:
//   inj_mass_kernel   every candidate's parcel mass (0 when rejected) and
//                     the accepted count of each workgroup, chip-wide;
//   inj_scan_kernel   one wavefront: the ordered sum (uniform loads, every
//                     lane adding the same values), the stopping candidate,
//                     and the exclusive scan of the workgroup counts;
//   inj_write_kernel  the accepted candidates up to the stop recomputed in
//                     full and written at their output index (workgroup
//                     offset + rank from a block ballot), chip-wide.
// Adding 0 for a rejected candidate leaves the sum unchanged and cannot
// stop the loop, which matches the reference skipping it.  Parcel masses are
// not negative (the host rejects a non-positive liquid density), so the
// running sum does not decrease: a group of kInjSumGroup candidates whose
// sum stays below the target holds no stop, and only the group that reaches
// it (or turns the sum NaN) is summed again candidate by candidate.
#pragma once

#include <hip/hip_runtime.h>

#include "inject.h"

namespace pelemp {

constexpr int kInjBlock = 256;
constexpr int kInjSumGroup = 32;

// result words of one injection
struct InjResult {
  double cur_mass;       // injected mass (carried from batch to batch)
  long long stop;        // batch index of the candidate that reached the mass, -1 if none
  long long n_out;       // droplets kept up to the stop (with a stop)
  long long batch_kept;  // accepted candidates of the batch
  int outside;           // a kept droplet outside the domain (Where() failed)
};

template <int NF>
__global__ void __launch_bounds__(kInjBlock) inj_mass_kernel(InjJet<NF> j, long long k0,
                                                             double* __restrict__ mass,
                                                             int* __restrict__ blk_cnt) {
  __shared__ int s_wave[kInjBlock / 64];
  const long long i = static_cast<long long>(blockIdx.x) * kInjBlock + threadIdx.x;
  const double dia = inj_dia(j, k0 + i);
  const bool acc = inj_accept(j, dia);
  mass[i] = acc ? inj_parcel_mass(j, dia) : 0.;
  const unsigned long long bal = __ballot(acc);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kInjBlock / 64; ++w) t += s_wave[w];
    blk_cnt[blockIdx.x] = t;
  }
}

// One wavefront of 64 lanes; nb is a multiple of kInjSumGroup.  blk_cnt becomes the
// exclusive prefix of the workgroup counts.
__global__ void __launch_bounds__(64) inj_scan_kernel(const double* __restrict__ mass, long long nb,
                                                      int* __restrict__ blk_cnt, int nblk, double inject_mass,
                                                      InjResult* __restrict__ res) {
  constexpr int G = kInjSumGroup;
  const int lane = threadIdx.x;
  double cur = res->cur_mass;
  long long stop = -1;
  double m[G], mn[G];
#pragma unroll
  for (int q = 0; q < G; ++q) m[q] = mass[q];
  for (long long base = 0; base < nb; base += G) {
    // the next group's loads in flight while this group is summed
    const long long nx = (base + G < nb) ? base + G : base;
#pragma unroll
    for (int q = 0; q < G; ++q) mn[q] = mass[nx + q];
    double c = cur;
#pragma unroll
    for (int q = 0; q < G; ++q) c = c + m[q];
    if (c < inject_mass) {
      cur = c;
#pragma unroll
      for (int q = 0; q < G; ++q) m[q] = mn[q];
      continue;
    }
    for (int q = 0; q < G; ++q) {
      cur = cur + m[q];
      if (!(cur < inject_mass)) {
        stop = base + q;
        break;
      }
    }
    break;
  }
  int carry = 0;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    const int v = (b < nblk) ? blk_cnt[b] : 0;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(x, off);
      if (lane >= off) x += t;
    }
    if (b < nblk) blk_cnt[b] = carry + x - v;
    carry += __shfl(x, 63);
  }
  if (lane == 0) {
    res->cur_mass = cur;
    res->stop = stop;
    res->batch_kept = carry;
  }
}

// Staging in the AoS meaning (pos[3], rdata[10 + NF], id).
template <int NF>
__global__ void __launch_bounds__(kInjBlock) inj_write_kernel(InjJet<NF> j, long long k0,
                                                              const int* __restrict__ blk_off, long long out_base,
                                                              long long out_cap, long long next_id,
                                                              double* __restrict__ pos, double* __restrict__ rdata,
                                                              long long* __restrict__ id,
                                                              InjResult* __restrict__ res) {
  constexpr int NSR = 10 + NF;
  __shared__ int s_wave[kInjBlock / 64];
  const long long stop = res->stop;
  const long long blk0 = static_cast<long long>(blockIdx.x) * kInjBlock;
  if (stop >= 0 && blk0 > stop) return;  // the whole workgroup is past the stop
  const long long i = blk0 + threadIdx.x;
  InjDrop o;
  inj_candidate<NF>(j, k0 + i, o);
  const bool keep = o.accept && (stop < 0 || i <= stop);
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const unsigned long long bal = __ballot(o.accept);
  const unsigned long long lower = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
  if (lane == 0) s_wave[wv] = __popcll(bal);
  __syncthreads();
  int off = 0;
  for (int w = 0; w < wv; ++w) off += s_wave[w];
  if (!keep) return;
  const long long oi = out_base + blk_off[blockIdx.x] + off + __popcll(lower);
  if (!inj_inside(o.pos, j.plo, j.dxi, j.ncell)) atomicOr(&res->outside, 1);
  if (i == stop) res->n_out = oi + 1;
  if (oi < out_cap) {
    double* r = rdata + oi * NSR;
    for (int d = 0; d < 3; ++d) {
      pos[3 * oi + d] = o.pos[d];
      r[d] = o.vel[d];
    }
    r[3] = j.T;
    r[4] = o.dia;
    for (int s = 0; s < NF; ++s) r[5 + s] = NF > 1 ? j.Y[s] : 1.;
    r[5 + NF] = j.num_ppp;  // numDens
    r[6 + NF] = j.num_ppp;  // N0
    r[7 + NF] = 0.;         // BM1
    r[8 + NF] = j.initial_bm2;
    r[9 + NF] = 0.;         // film height
    id[oi] = next_id + oi;
  }
}

// uniformSprayInit (SprayInjection.H:198-348) on one rank: lattice droplet
// idx (x fastest) at ((i + 0.5) * dx_part) from the origin, written at
// staging position n - 1 - idx (the reference moves them to the tiles from
// the back of its list).
template <int NF>
__global__ void __launch_bounds__(256) lattice_kernel(long long n, int np0, int np1, double dxp0, double dxp1,
                                                      double dxp2, double v0, double v1, double v2, double dia,
                                                      double T, const double* __restrict__ Yp, double num_ppp,
                                                      double initial_bm2, long long next_id, double plo0,
                                                      double plo1, double plo2, double dxi0, double dxi1,
                                                      double dxi2, int nc0, int nc1, int nc2,
                                                      double* __restrict__ pos, double* __restrict__ rdata,
                                                      long long* __restrict__ id, int* __restrict__ outside) {
  constexpr int NSR = 10 + NF;
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const long long d1 = np0, d2 = np1;
  const long long i2 = idx / (d1 * d2);
  const long long rem = idx - i2 * d1 * d2;
  const long long i1 = rem / d1;
  const long long i0 = rem % d1;
  const double p[3] = {(static_cast<double>(i0) + 0.5) * dxp0, (static_cast<double>(i1) + 0.5) * dxp1,
                       (static_cast<double>(i2) + 0.5) * dxp2};
  const double plo[3] = {plo0, plo1, plo2}, dxi[3] = {dxi0, dxi1, dxi2};
  const int nc[3] = {nc0, nc1, nc2};
  if (!inj_inside(p, plo, dxi, nc)) atomicOr(outside, 1);
  const long long o = n - 1 - idx;
  double* r = rdata + o * NSR;
  for (int d = 0; d < 3; ++d) pos[3 * o + d] = p[d];
  r[0] = v0;
  r[1] = v1;
  r[2] = v2;
  r[3] = T;
  r[4] = dia;
  for (int s = 0; s < NF; ++s) r[5 + s] = Yp[s];
  r[5 + NF] = num_ppp;
  r[6 + NF] = num_ppp;
  r[7 + NF] = 0.;
  r[8 + NF] = initial_bm2;
  r[9 + NF] = 0.;
  id[o] = next_id + idx;
}

// Renumber the cold slots of the live droplets by sorted position, so the
// slot arrays can take new droplets (slots of deleted droplets are never
// reused otherwise): tmp_* receive the cold data in sorted order.
__global__ void __launch_bounds__(256) slot_compact_kernel(const int32_t* __restrict__ total,
                                                           const int32_t* __restrict__ perm,
                                                           int32_t* __restrict__ slot,
                                                           const int64_t* __restrict__ cold_id,
                                                           const double* __restrict__ cold_n0,
                                                           const double* __restrict__ cold_bm1,
                                                           const double* __restrict__ cold_bm2, int with_bm,
                                                           int64_t* __restrict__ tmp_id, double* __restrict__ tmp_n0,
                                                           double* __restrict__ tmp_bm1,
                                                           double* __restrict__ tmp_bm2) {
  const int32_t n = *total;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int32_t o = perm[s];
    const int32_t sl = slot[o];
    tmp_id[s] = cold_id[sl];
    if (with_bm) {
      tmp_n0[s] = cold_n0[sl];
      tmp_bm1[s] = cold_bm1[sl];
      tmp_bm2[s] = cold_bm2[sl];
    }
    slot[o] = s;
  }
}

// Storage compaction: row q of the live droplets, in sorted order, from
// buffer src (storage order) to dst (position = sorted position).
template <class T>
__global__ void __launch_bounds__(256) gather_rows_kernel(const int32_t* __restrict__ total,
                                                          const int32_t* __restrict__ perm, const T* __restrict__ src,
                                                          T* __restrict__ dst, int nrows, int64_t cap) {
  const int32_t n = *total;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int32_t o = perm[s];
    for (int q = 0; q < nrows; ++q) dst[q * cap + s] = src[q * cap + o];
  }
}

__global__ void __launch_bounds__(256) identity_perm_kernel(const int32_t* __restrict__ total,
                                                            int32_t* __restrict__ perm,
                                                            int32_t* __restrict__ n_entries) {
  const int32_t n = *total;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) perm[s] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_entries = n;
}

}  // namespace pelemp
