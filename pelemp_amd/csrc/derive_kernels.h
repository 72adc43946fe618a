// derive_kernels.h -- SprayParticleContainer::computeDerivedVars
// (SprayDerive.cpp:9-155): per-cell spray diagnostics on a mesh FAB.
//
// One thread per cell of the FAB box.  A cell inside the domain sums its
// droplets in cell-sorted order (the reference adds with atomics) and adds the
// sums to the FAB; then every cell of the box gets the protected divides
// (d10 /= num, d32 /= surface, temperature and velocity /= mass, only where
// the denominator is non-zero).  The cell's variables are loaded once into
// registers, updated there and stored once (the FAB pointer with a run-time
// stride would otherwise serialise every load behind the previous store).  Variables from start_indx, in the order of
// SpraySetup.cpp:213-239: mass, density, num, volume, surface, volume
// fraction, d10, d32, wall-film height, wall-film mass, temperature,
// parcels, velocity(3), and for several fuels the liquid mass of each.
#pragma once

#include <hip/hip_runtime.h>

#include "droplet_physics.h"
#include "layout.h"
#include "transfer_kernels.h"

namespace pelemp {

constexpr int kDeriveBase = 15;

struct DeriveGeom {
  int dlo[3], ncell[3], ntile[3];
  double cell_vol, face_area;
  int pelec;                                    // CGS output (PeleC flavour)
  double scale[kDeriveBase + PELEMP_MAX_FUEL];  // SI -> CGS per variable
};

// R: storage type of the droplet fields (double: FP64 form, float: FP32
// form, whose liquid composition comes from the FP64 species masses m)
template <int NF, class R>
__global__ void __launch_bounds__(256) derive_kernel(const int32_t* __restrict__ cstart,
                                                     const int32_t* __restrict__ perm, const R* __restrict__ f,
                                                     const double* __restrict__ m, long long cap,
                                                     FuelConst<double, NF> fc, DeriveGeom g, XferFab mf,
                                                     int start_indx, int nvar) {
  const int ncells = g.ncell[0] * g.ncell[1] * g.ncell[2];
  const int32_t ntot = cstart[ncells];
  const long long nbox = mf.cstride;
  constexpr int mass_i = 0, dens_i = 1, num_i = 2, vol_i = 3, surf_i = 4, volf_i = 5, d10_i = 6, d32_i = 7,
                wfh_i = 8, wfm_i = 9, temp_i = 10, nump_i = 11, vel_i = 12, spec_i = 15;
  for (long long cell = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; cell < nbox;
       cell += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int i = mf.lo[0] + static_cast<int>(cell % mf.n[0]);
    const int j = mf.lo[1] + static_cast<int>((cell / mf.n[0]) % mf.n[1]);
    const int k = mf.lo[2] + static_cast<int>(cell / (static_cast<long long>(mf.n[0]) * mf.n[1]));
    double* V = mf.data + start_indx * mf.cstride + cell;
    const int ci = i - g.dlo[0], cj = j - g.dlo[1], ck = k - g.dlo[2];
    int32_t s0 = 0, s1 = 0;
    if (ci >= 0 && ci < g.ncell[0] && cj >= 0 && cj < g.ncell[1] && ck >= 0 && ck < g.ncell[2]) {
      const uint32_t key = tile_key(ci, cj, ck, g.ntile);
      s0 = cstart[key];
      s1 = cstart[key + 1];
    }
    const bool has = s1 > s0;
    if (!has && !(ntot > 0)) continue;
    // the cell's variables: loads issued before the droplet loop, so they
    // overlap with it
    double W[kDeriveBase + NF];
#pragma unroll
    for (int q = 0; q < kDeriveBase + NF; ++q) W[q] = (q < nvar) ? V[q * mf.cstride] : 0.;
    {
      if (has) {
        double acc[kDeriveBase + NF];
#pragma unroll
        for (int q = 0; q < kDeriveBase + NF; ++q) acc[q] = 0.;
        int32_t o_next = perm[s0];
        for (int32_t s = s0; s < s1; ++s) {
          const int32_t o = o_next;
          if (s + 1 < s1) o_next = perm[s + 1];
          double T_part, dia_part, num_ppp, film_hght, vel[3], Y[NF];
          if constexpr (sizeof(R) == 8) {
            using F = Fields<NF>;
            T_part = f[F::T * cap + o];
            dia_part = f[F::DIA * cap + o];
            num_ppp = f[F::ND * cap + o];
            film_hght = f[F::FILM * cap + o];
#pragma unroll
            for (int d = 0; d < 3; ++d) vel[d] = f[(F::VEL + d) * cap + o];
#pragma unroll
            for (int spf = 0; spf < NF; ++spf) Y[spf] = f[(F::Y + spf) * cap + o];
          } else {
            using G = Fields32;
            T_part = static_cast<double>(f[G::T * cap + o]);
            dia_part = static_cast<double>(f[G::DIA * cap + o]);
            num_ppp = static_cast<double>(f[G::ND * cap + o]);
            film_hght = 0.;  // the FP32 form has no wall films
#pragma unroll
            for (int d = 0; d < 3; ++d) vel[d] = static_cast<double>(f[(G::VEL + d) * cap + o]);
            double mt = 0.;
#pragma unroll
            for (int spf = 0; spf < NF; ++spf) mt += m[spf * cap + o];
#pragma unroll
            for (int spf = 0; spf < NF; ++spf) Y[spf] = m[spf * cap + o] / mt;
          }
          double rho_part = 0.;
#pragma unroll
          for (int spf = 0; spf < NF; ++spf) rho_part += Y[spf] / rhoL<double, NF>(fc, T_part, spf);
          rho_part = 1. / rho_part;
          const double surf = kPIc * dia_part * dia_part;
          const double vol = kPIc / 6. * rm::pow(dia_part, 3.);
          const double pmass = vol * rho_part;
          if (film_hght == 0.) {
            acc[mass_i] += num_ppp * pmass;
            acc[dens_i] += num_ppp * pmass / g.cell_vol;
            acc[num_i] += num_ppp;
            acc[vol_i] += num_ppp * vol;
            acc[surf_i] += num_ppp * surf;
            acc[volf_i] += num_ppp * vol / g.cell_vol;
            acc[d10_i] += num_ppp * dia_part;
            acc[d32_i] += num_ppp * vol * 6.;
            acc[temp_i] += num_ppp * pmass * T_part;
            acc[nump_i] += 1.;
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[vel_i + d] += num_ppp * pmass * vel[d];
#pragma unroll
            for (int spf = 0; spf < NF; ++spf) acc[spec_i + spf] += Y[spf] * pmass;
          } else {
            const double rad2 = rm::pow(0.5 * dia_part, 2.);
            const double cur_vol = kPIc / 6. * film_hght * (3. * rad2 + film_hght * film_hght);
            acc[wfh_i] += cur_vol / g.face_area;
            acc[wfm_i] += rho_part * cur_vol;
          }
        }
#pragma unroll
        for (int q = 0; q < kDeriveBase + NF; ++q)
          if (q < nvar) W[q] = W[q] + (g.pelec ? acc[q] * g.scale[q] : acc[q]);
      }
    }
    if (ntot > 0) {
      if (W[num_i] != 0.) W[d10_i] = W[d10_i] / W[num_i];
      if (W[surf_i] != 0.) W[d32_i] = W[d32_i] / W[surf_i];
      const double mass = W[mass_i];
      if (mass != 0.) {
        W[temp_i] = W[temp_i] / mass;
#pragma unroll
        for (int d = 0; d < 3; ++d) W[vel_i + d] = W[vel_i + d] / mass;
      }
    }
#pragma unroll
    for (int q = 0; q < kDeriveBase + NF; ++q)
      if (q < nvar) V[q * mf.cstride] = W[q];
  }
}

}  // namespace pelemp
