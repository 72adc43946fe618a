// transfer_kernels.h -- the step after the update: transferSource and
// addSpraySrc (SprayParticles.H:250-319) for one box covering the domain.
//
// transfer_add_kernel: one thread per cell of the domain grown by nghost, i
// fastest.  For each component pair (tmp comp -> act comp) the thread adds
// tmp's value to act; a valid cell first adds its periodic images among
// tmp's ghost cells (SumBoundary), in the fixed shift order (sz, sy, sx) in
// {-1,0,1}^3, z slowest.  Each act value receives one add per pair in pair
// order, as addSpraySrc adds per cell (momentum, density, energy, species).
// zero_fill_kernel then clears tmp (setVal(0)); it is a kernel rather than a
// memset so that the pass can be captured into a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pelemp {

constexpr int kXferMaxPairs = 64;

struct XferFab {
  double* data;
  int lo[3];
  int n[3];
  long long cstride;  // cells per component
};

struct XferMap {
  int npair;
  int from[kXferMaxPairs];
  int to[kXferMaxPairs];
};

struct XferGeom {
  int dlo[3], N[3];  // domain lower corner and size
  int periodic[3];
  int nghost;
};

__device__ inline bool xfer_has(const XferFab& f, int i, int j, int k) {
  return i >= f.lo[0] && i < f.lo[0] + f.n[0] && j >= f.lo[1] && j < f.lo[1] + f.n[1] && k >= f.lo[2] &&
         k < f.lo[2] + f.n[2];
}

__device__ inline long long xfer_cell(const XferFab& f, int i, int j, int k) {
  return (static_cast<long long>(k - f.lo[2]) * f.n[1] + (j - f.lo[1])) * f.n[0] + (i - f.lo[0]);
}

__global__ void __launch_bounds__(256) transfer_add_kernel(XferFab tmp, XferFab act, XferMap map, XferGeom g) {
  const int ng = g.nghost;
  const int R0 = g.N[0] + 2 * ng, R1 = g.N[1] + 2 * ng, R2 = g.N[2] + 2 * ng;
  const long long ncell = static_cast<long long>(R0) * R1 * R2;
  for (long long m = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; m < ncell;
       m += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int i = g.dlo[0] - ng + static_cast<int>(m % R0);
    const int j = g.dlo[1] - ng + static_cast<int>((m / R0) % R1);
    const int k = g.dlo[2] - ng + static_cast<int>(m / (static_cast<long long>(R0) * R1));
    const bool valid = i >= g.dlo[0] && i < g.dlo[0] + g.N[0] && j >= g.dlo[1] && j < g.dlo[1] + g.N[1] &&
                       k >= g.dlo[2] && k < g.dlo[2] + g.N[2];
    const long long ct = xfer_cell(tmp, i, j, k);
    const long long ca = xfer_cell(act, i, j, k);
    for (int q = 0; q < map.npair; ++q) {
      const double* T = tmp.data + map.from[q] * tmp.cstride;
      double v = T[ct];
      if (valid) {
        for (int sz = -1; sz <= 1; ++sz) {
          if (sz != 0 && !g.periodic[2]) continue;
          for (int sy = -1; sy <= 1; ++sy) {
            if (sy != 0 && !g.periodic[1]) continue;
            for (int sx = -1; sx <= 1; ++sx) {
              if (sx != 0 && !g.periodic[0]) continue;
              if (sx == 0 && sy == 0 && sz == 0) continue;
              const int gi = i + sx * g.N[0], gj = j + sy * g.N[1], gk = k + sz * g.N[2];
              if (xfer_has(tmp, gi, gj, gk)) v += T[xfer_cell(tmp, gi, gj, gk)];
            }
          }
        }
      }
      double* A = act.data + map.to[q] * act.cstride;
      A[ca] = A[ca] + v;
    }
  }
}

__global__ void __launch_bounds__(256) zero_fill_kernel(double* __restrict__ p, long long n) {
  for (long long m = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; m < n;
       m += static_cast<long long>(gridDim.x) * blockDim.x)
    p[m] = 0.;
}

}  // namespace pelemp
