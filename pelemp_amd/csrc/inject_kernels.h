// inject_kernels.h -- device side of the injection and initialisation
// routines (SprayInjection.H:10-348) and of appending droplets.
//
// inject_kernel runs the droplet loop of sprayInjection (SprayInjection.H:
// 78-158): candidates are drawn until the injected mass reaches the target.
// One workgroup of kInjBlock threads draws kInjBlock candidates at a time in
// parallel (each candidate has its own counter-based random numbers); one
// thread then adds their masses in candidate order, the order of the
// reference's loop, so the stopping candidate and the injected mass are the
// reference loop's own; the kept droplets of the chunk get their output
// index from a block prefix count.  Injection counts are small (tens to
// thousands of droplets per step), so one workgroup is the right size: the
// launch is latency-bound either way.
#pragma once

#include <hip/hip_runtime.h>

#include "inject.h"

namespace pelemp {

constexpr int kInjBlock = 256;

// result words of one injection launch
struct InjResult {
  double cur_mass;   // injected mass
  long long n_out;   // droplets kept (may exceed the staging capacity)
  long long n_cand;  // candidates drawn
  int outside;       // a kept droplet outside the domain (Where() failed)
  int unfinished;    // the candidate limit was reached before the target
};

// Staging in the AoS meaning (pos[3], rdata[10 + NF], id).
template <int NF>
__global__ void __launch_bounds__(kInjBlock) inject_kernel(InjJet<NF> j, double inject_mass, long long max_cand,
                                                           long long out_cap, long long next_id,
                                                           double* __restrict__ pos, double* __restrict__ rdata,
                                                           long long* __restrict__ id,
                                                           InjResult* __restrict__ res) {
  constexpr int NSR = 10 + NF;
  __shared__ double s_mass[kInjBlock];
  __shared__ int s_acc[kInjBlock];
  __shared__ int s_wave[kInjBlock / 64];
  __shared__ int s_stop;
  __shared__ double s_cur;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  double cur = 0.;  // thread 0's running mass
  long long out_base = 0;
  long long base = 0;
  int outside = 0;
  bool done = !(cur < inject_mass);
  while (!done) {
    const long long k = base + tid;
    InjDrop o;
    inj_candidate<NF>(j, k, o);
    s_mass[tid] = o.mass;
    s_acc[tid] = o.accept ? 1 : 0;
    __syncthreads();
    if (tid == 0) {
      // the reference's loop condition, candidate by candidate
      int stop = -1;
      for (int q0 = 0; q0 < kInjBlock && stop < 0; q0 += 8) {
        double m[8];
        int a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          m[q] = s_mass[q0 + q];
          a[q] = s_acc[q0 + q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (stop < 0 && a[q]) {
            cur = cur + m[q];
            if (!(cur < inject_mass)) stop = q0 + q;
          }
        }
      }
      s_stop = stop;
      s_cur = cur;
    }
    __syncthreads();
    const int stop = s_stop;
    const bool keep = o.accept && (stop < 0 || tid <= stop);
    // output index: kept candidates before this one in the chunk
    const unsigned long long bal = __ballot(keep);
    const unsigned long long lower = (lane == 0) ? 0ull : (bal & ((~0ull) >> (64 - lane)));
    const int rank_w = __popcll(lower);
    if (lane == 0) s_wave[wv] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < kInjBlock / 64; ++w) {
      off += (w < wv) ? s_wave[w] : 0;
      tot += s_wave[w];
    }
    if (keep) {
      const long long oi = out_base + off + rank_w;
      if (!inj_inside(o.pos, j.plo, j.dxi, j.ncell)) outside = 1;
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
    out_base += tot;
    base += kInjBlock;
    done = (stop >= 0) || (base >= max_cand);
    __syncthreads();
  }
  // every thread that saw a droplet outside reports it
  if (outside) atomicOr(&res->outside, 1);
  if (tid == 0) {
    res->cur_mass = cur;
    res->n_out = out_base;
    res->n_cand = base;
    res->unfinished = (cur < inject_mass) ? 1 : 0;
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
