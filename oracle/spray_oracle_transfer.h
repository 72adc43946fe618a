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
// for one level held in one box that covers the domain (the layout of this
// build's container).
//
// Deviation (documented): SumBoundary adds the ghost cells of tmpSource to
// the valid cells they are periodic images of; a valid cell near a corner has
// several images and AMReX leaves their order to its communication tags.
// Here they are added in a fixed order: shifts (sz, sy, sx) in {-1,0,1}^3,
// z slowest, each image added to the running value of the valid cell.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

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

}  // namespace oracle
