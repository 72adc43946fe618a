// spray_oracle_transfer.h -- CPU restatement of the steps on either side of
// the droplet update: the transfer of the spray source to the flow solver's
// source and the particle-CFL time step.
//
// TEST INFRASTRUCTURE (see spray_oracle.h).  Restates:
//   Source/PP_Spray/SprayParticles.H:250-295  transferSource
//     (tmpSource.SumBoundary(periodicity), MultiFab::Add or addSpraySrc over
//      the valid box grown by min(actSource ghosts, source_ghosts),
//      tmpSource.setVal(0))
//   Source/PP_Spray/SprayParticles.H:299-319  addSpraySrc
//   Source/PP_Spray/SprayParticles.cpp:89-140 estTimestep
//   Source/PP_Spray/SprayDerive.cpp:9-155     computeDerivedVars (list of
//     SpraySetup.cpp:213-239)
// for one level held in one box that covers the domain (the layout of this
// build's container).
//
// Deviations (documented): computeDerivedVars sums each cell's droplets in
// cell-sorted order (the reference adds with atomics) and uses the droplet's
// containing cell clamped into the domain (the reference indexes the
// unclamped cell).  SumBoundary adds the ghost cells of tmpSource to
// the valid cells they are periodic images of; a valid cell near a corner has
// several images and AMReX leaves their order to its communication tags.
// Here they are added in a fixed order: shifts (sz, sy, sx) in {-1,0,1}^3,
// z slowest, each image added to the running value of the valid cell.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../include/pelemp_spray.h"

namespace oracle {

// valid-box index space of the FAB: cell (i, j, k) in domain coordinates
inline int64_t fab_off(const pelemp_fab_view& f, int i, int j, int k, int comp) {
  const int64_t nx = f.hi[0] - f.lo[0] + 1, ny = f.hi[1] - f.lo[1] + 1, nz = f.hi[2] - f.lo[2] + 1;
  return ((static_cast<int64_t>(comp) * nz + (k - f.lo[2])) * ny + (j - f.lo[1])) * nx + (i - f.lo[0]);
}

inline bool fab_has(const pelemp_fab_view& f, int i, int j, int k) {
  return i >= f.lo[0] && i <= f.hi[0] && j >= f.lo[1] && j <= f.hi[1] && k >= f.lo[2] && k <= f.hi[2];
}

// transferSource; comps/dep_indx describe the component map of addSpraySrc.
// Returns PELEMP_OK or PELEMP_E_ARG.
inline int transfer_source(const pelemp_geom& g, const pelemp_spray_comps& SPI, const int* dep_indx, int nfuel,
                           int source_ghosts, pelemp_fab_view& tmp, pelemp_fab_view& act) {
  int ng_tmp = 1 << 30, ng_act = 1 << 30;
  for (int d = 0; d < 3; ++d) {
    ng_tmp = std::min(ng_tmp, std::min(g.domain_lo[d] - tmp.lo[d], tmp.hi[d] - g.domain_hi[d]));
    ng_act = std::min(ng_act, std::min(g.domain_lo[d] - act.lo[d], act.hi[d] - g.domain_hi[d]));
  }
  if (ng_tmp < 0 || ng_act < 0 || source_ghosts < 0) return PELEMP_E_ARG;
  const int nghost = std::min(std::min(ng_act, source_ghosts), ng_tmp);
  const int ncomp = tmp.ncomp;
  // component pairs (tmp comp -> act comp)
  int nmap = 0;
  int from[64], to[64];
  if (tmp.ncomp == act.ncomp) {
    if (ncomp > 64) return PELEMP_E_ARG;
    for (int c = 0; c < ncomp; ++c) {
      from[nmap] = c;
      to[nmap++] = c;
    }
  } else {
    for (int dir = 0; dir < 3; ++dir) {
      from[nmap] = SPI.momSrcIndx + dir;
      to[nmap++] = SPI.momIndx + dir;
    }
    from[nmap] = SPI.rhoSrcIndx;
    to[nmap++] = SPI.rhoIndx;
    from[nmap] = SPI.engSrcIndx;
    to[nmap++] = SPI.engIndx;
    for (int sp = 0; sp < nfuel; ++sp) {
      from[nmap] = SPI.specSrcIndx + sp;
      to[nmap++] = SPI.specIndx + dep_indx[sp];
    }
    for (int q = 0; q < nmap; ++q)
      if (from[q] < 0 || from[q] >= tmp.ncomp || to[q] < 0 || to[q] >= act.ncomp) return PELEMP_E_ARG;
  }
  int N[3];
  for (int d = 0; d < 3; ++d) N[d] = g.domain_hi[d] - g.domain_lo[d] + 1;
  const int lo[3] = {g.domain_lo[0] - nghost, g.domain_lo[1] - nghost, g.domain_lo[2] - nghost};
  const int hi[3] = {g.domain_hi[0] + nghost, g.domain_hi[1] + nghost, g.domain_hi[2] + nghost};
  for (int q = 0; q < nmap; ++q) {
    for (int k = lo[2]; k <= hi[2]; ++k) {
      for (int j = lo[1]; j <= hi[1]; ++j) {
        for (int i = lo[0]; i <= hi[0]; ++i) {
          double v = tmp.data[fab_off(tmp, i, j, k, from[q])];
          const bool valid = i >= g.domain_lo[0] && i <= g.domain_hi[0] && j >= g.domain_lo[1] &&
                             j <= g.domain_hi[1] && k >= g.domain_lo[2] && k <= g.domain_hi[2];
          if (valid) {
            // SumBoundary: the periodic images of this cell among tmp's ghosts
            for (int sz = -1; sz <= 1; ++sz) {
              for (int sy = -1; sy <= 1; ++sy) {
                for (int sx = -1; sx <= 1; ++sx) {
                  if (sx == 0 && sy == 0 && sz == 0) continue;
                  const int s3[3] = {sx, sy, sz};
                  bool ok = true;
                  for (int d = 0; d < 3; ++d)
                    if (s3[d] != 0 && g.bndry_lo[d] != 0) ok = false;  // not periodic
                  if (!ok) continue;
                  const int gi = i + sx * N[0], gj = j + sy * N[1], gk = k + sz * N[2];
                  if (fab_has(tmp, gi, gj, gk)) v += tmp.data[fab_off(tmp, gi, gj, gk, from[q])];
                }
              }
            }
          }
          double* dst = &act.data[fab_off(act, i, j, k, to[q])];
          *dst = *dst + v;
        }
      }
    }
  }
  // tmpSource.setVal(0.)
  const int64_t ntot = static_cast<int64_t>(tmp.hi[0] - tmp.lo[0] + 1) * (tmp.hi[1] - tmp.lo[1] + 1) *
                       (tmp.hi[2] - tmp.lo[2] + 1) * tmp.ncomp;
  for (int64_t m = 0; m < ntot; ++m) tmp.data[m] = 0.;
  return PELEMP_OK;
}

// estTimestep for particles vel[3*n] (active ones only), cfl = spray_cfl;
// max_jet_vel > 0 adds the injection bound cfl dx_0 / max_jet_vel.
inline double est_timestep(const pelemp_geom& g, bool fixed_parts, int64_t n, const double* vel, double cfl,
                           double max_jet_vel) {
  if (fixed_parts) return -1.;
  double dt = std::numeric_limits<double>::max();
  const double dxi[3] = {1. / g.dx[0], 1. / g.dx[1], 1. / g.dx[2]};
  for (int64_t i = 0; i < n; ++i) {
    double m = std::fabs(vel[3 * i]) * dxi[0];
    m = std::max(m, std::fabs(vel[3 * i + 1]) * dxi[1]);
    m = std::max(m, std::fabs(vel[3 * i + 2]) * dxi[2]);
    const double dt_part = (m > 0.) ? (cfl / m) : 1.E50;
    dt = std::min(dt, dt_part);
  }
  if (max_jet_vel > 0.) dt = std::min(dt, cfl * g.dx[0] / max_jet_vel);
  return dt;
}

// computeDerivedVars (SprayDerive.cpp:9-155) for particles in any order:
// pos[3n], rdata[n*NSR] (SprayComps layout), sorted here by containing cell
// like the update.  mf_var[ncomp][nz][ny][nx] on box [lo, hi] (covering the
// domain); the spray variables start at start_indx: mass, density, num,
// volume, surface, volume fraction, d10, d32, wall-film height, wall-film
// mass, temperature, parcels, velocity(3), then, when with_species, the
// liquid mass of each fuel.
template <int NF>
inline int compute_derived(const pelemp_geom& g, const pelemp_spray_data& fdat, int64_t n, const double* pos,
                           const double* rdata, pelemp_fab_view& mf, int start_indx, bool with_species,
                           const double* scale = nullptr) {
  using C = Comps<NF>;
  constexpr int NSR = C::pstateNum;
  // the per-fuel masses are listed only for several fuels (SpraySetup.cpp:236-240)
  const int nvar = 15 + ((with_species && NF > 1) ? NF : 0);
  if (start_indx < 0 || start_indx + nvar > mf.ncomp) return PELEMP_E_ARG;
  for (int d = 0; d < 3; ++d)
    if (mf.lo[d] > g.domain_lo[d] || mf.hi[d] < g.domain_hi[d]) return PELEMP_E_ARG;
  const double dxi[3] = {1. / g.dx[0], 1. / g.dx[1], 1. / g.dx[2]};
  const double cell_vol = g.dx[0] * g.dx[1] * g.dx[2];
  const double face_area = 1. * g.dx[0] * g.dx[0];
  std::vector<int64_t> key(n), order(n);
  for (int64_t i = 0; i < n; ++i) {
    int c[3];
    containing_cell(&pos[3 * i], g, dxi, c);
    key[i] = cell_key(c, g);
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  const int mass_i = 0, dens_i = 1, num_i = 2, vol_i = 3, surf_i = 4, volf_i = 5, d10_i = 6, d32_i = 7,
            wfh_i = 8, wfm_i = 9, temp_i = 10, nump_i = 11, vel_i = 12, spec_i = 15;
  int64_t a = 0;
  while (a < n) {
    int64_t b = a;
    while (b < n && key[order[b]] == key[order[a]]) ++b;
    double acc[16 + NF];
    for (int q = 0; q < 16 + NF; ++q) acc[q] = 0.;
    for (int64_t r = a; r < b; ++r) {
      const double* rd = &rdata[order[r] * NSR];
      const double T_part = rd[C::pstateT];
      const double dia_part = rd[C::pstateDia];
      double rho_part = 0.;
      for (int spf = 0; spf < NF; ++spf) rho_part += rd[C::pstateY + spf] / rhoL(fdat, T_part, spf);
      rho_part = 1. / rho_part;
      const double surf = kPI * dia_part * dia_part;
      const double vol = kPI / 6. * pm::pow(dia_part, 3.);
      const double pmass = vol * rho_part;
      const double num_ppp = rd[C::pstateNumDens];
      const double film_hght = rd[C::pstateFilmHght];
      if (film_hght == 0.) {
        acc[mass_i] += num_ppp * pmass;
        acc[dens_i] += num_ppp * pmass / cell_vol;
        acc[num_i] += num_ppp;
        acc[vol_i] += num_ppp * vol;
        acc[surf_i] += num_ppp * surf;
        acc[volf_i] += num_ppp * vol / cell_vol;
        acc[d10_i] += num_ppp * dia_part;
        acc[d32_i] += num_ppp * vol * 6.;
        acc[temp_i] += num_ppp * pmass * T_part;
        acc[nump_i] += 1.;
        for (int dir = 0; dir < 3; ++dir) acc[vel_i + dir] += num_ppp * pmass * rd[C::pstateVel + dir];
        // the fuel masses carry no num_ppp factor (SprayDerive.cpp:117-123)
        for (int spf = 0; spf < NF; ++spf) acc[spec_i + spf] += rd[C::pstateY + spf] * pmass;
      } else {
        const double rad2 = pm::pow(0.5 * dia_part, 2.);
        const double cur_vol = kPI / 6. * film_hght * (3. * rad2 + film_hght * film_hght);
        acc[wfh_i] += cur_vol / face_area;
        acc[wfm_i] += rho_part * cur_vol;
      }
    }
    int c[3];
    containing_cell(&pos[3 * order[a]], g, dxi, c);
    for (int q = 0; q < nvar; ++q) {
      double* v = &mf.data[fab_off(mf, c[0], c[1], c[2], start_indx + q)];
      *v = scale ? *v + acc[q] * scale[q] : *v + acc[q];
    }
    a = b;
  }
  if (n == 0) return PELEMP_OK;  // the reference visits no particle tile
  // protected_divide over the whole FAB: only where the denominator is non-zero
  for (int k = mf.lo[2]; k <= mf.hi[2]; ++k) {
    for (int j = mf.lo[1]; j <= mf.hi[1]; ++j) {
      for (int i = mf.lo[0]; i <= mf.hi[0]; ++i) {
        auto at = [&](int q) -> double& { return mf.data[fab_off(mf, i, j, k, start_indx + q)]; };
        if (at(num_i) != 0.) at(d10_i) = at(d10_i) / at(num_i);
        if (at(surf_i) != 0.) at(d32_i) = at(d32_i) / at(surf_i);
        if (at(mass_i) != 0.) {
          at(temp_i) = at(temp_i) / at(mass_i);
          for (int dir = 0; dir < 3; ++dir) at(vel_i + dir) = at(vel_i + dir) / at(mass_i);
        }
      }
    }
  }
  return PELEMP_OK;
}

}  // namespace oracle
