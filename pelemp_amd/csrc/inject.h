// inject.h -- droplet injection of one spray jet: the candidate droplet of
// SprayParticleContainer::sprayInjection (SprayInjection.H:78-158) with
// SprayJet::get_new_particle / transform_loc_vel (SprayJet.cpp:148-184,
// SprayJet.H:165-221) and the diameter distributions
// (Distribution/Distributions.cpp:1-174).  FP64, the project's elementary
// functions, so that device and host give the same bits.
//
// amrex::Random() is replaced by a counter-based uniform number keyed by
// (jet seed, injection call, candidate index, draw slot): every candidate
// has its own fixed slots, so the candidates of one call can be drawn in
// parallel and the result does not depend on the launch shape.  Draw slots:
// 0 radial fraction, 1 azimuth, 2 hollow-cone spread, 3-4 the diameter
// (Box-Muller pair for the normal and log-normal distributions), 5 the
// fraction of dt the droplet has already travelled.
#pragma once

#include <cstdint>

#include "gas_props.h"
#include "pelemp_spray.h"

namespace pelemp {

constexpr int kChiBins = 100;
constexpr double kInjPI = 3.14159265358979323846;

PM_HDI uint64_t inj_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

PM_HDI double inj_uniform(uint64_t seed, int64_t call, int64_t cand, int draw) {
  const uint64_t a = inj_mix64(seed ^ inj_mix64(static_cast<uint64_t>(call)));
  const uint64_t r = inj_mix64(a ^ inj_mix64(static_cast<uint64_t>(cand) * 8ull + static_cast<uint64_t>(draw)));
  return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
}

// standard normal from two uniforms in [0, 1) (Box-Muller, cosine branch)
PM_HDI double inj_std_normal(double u1, double u2) {
  return std::sqrt(-2. * pm::log(1. - u1)) * pm::cos(2. * kInjPI * u2);
}

// A distribution as sampled: p0, p1 per type
//   UNIFORM: diameter;  NORMAL: mean, std;  LOGNORMAL: log mean, log std;
//   WEIBULL: scale, k;  CHISQUARED: d32 with the CDF table chi[]
struct InjDist {
  int type;
  double p0, p1;
  double chi[kChiBins];
};

PM_HDI double inj_sample_dia(const InjDist& d, double u1, double u2) {
  switch (d.type) {
    case PELEMP_DIST_UNIFORM: return d.p0;
    case PELEMP_DIST_NORMAL: return d.p0 + d.p1 * inj_std_normal(u1, u2);
    case PELEMP_DIST_LOGNORMAL: return pm::exp(d.p0 + d.p1 * inj_std_normal(u1, u2));
    case PELEMP_DIST_WEIBULL: {
      const double fact = -pm::log(1. - u1);
      return d.p0 * pm::pow(fact, 1. / d.p1);
    }
    default: {  // CHISQUARED: first table bin whose CDF reaches u1, left edge
      const double dmean = d.p0 / 3.;
      const double dxi = 12. / 100.;
      int curn = 0;
      double curr = d.chi[0];
      double curxi = 0.;
      while (u1 > curr && curn < kChiBins - 1) {
        curn++;
        curr = d.chi[curn];
        curxi += dxi;
      }
      return curxi * dmean;
    }
  }
}

// Everything a candidate needs, fixed for one call (host-computed).
template <int NF>
struct InjJet {
  double cent[3];
  double sp1, cp1, st1, ct1;  // jet direction angles (transform_loc_vel)
  double spread, swirl, hollow_spread, jet_dia;
  int hollow;
  double umag;      // jet_vel(time)
  double T, Y[NF];  // droplet temperature and composition
  double rho_part;  // liquid density at (T, Y), as sprayInjection forms it
  double min_dia;
  double num_ppp;
  double dt;        // sim_dt plus held-over time
  double initial_bm2;
  uint64_t seed;
  int64_t call;
  InjDist dist;
  // domain for the Where() check
  double plo[3], dxi[3];
  int ncell[3];
};

struct InjDrop {
  bool accept;
  double pos[3], vel[3], dia, mass;  // mass = num_ppp * pmass
};

// diameter of candidate k (draw slots 3-4), its acceptance and parcel mass:
// the part of a candidate the injected-mass sum needs
template <int NF>
PM_HDI double inj_dia(const InjJet<NF>& j, int64_t k) {
  return inj_sample_dia(j.dist, inj_uniform(j.seed, j.call, k, 3), inj_uniform(j.seed, j.call, k, 4));
}

template <int NF>
PM_HDI bool inj_accept(const InjJet<NF>& j, double dia) {
  return dia > j.min_dia && j.umag > 0.;
}

template <int NF>
PM_HDI double inj_parcel_mass(const InjJet<NF>& j, double dia) {
  const double pmass = kInjPI / 6. * j.rho_part * pm::pow(dia, 3.);
  return j.num_ppp * pmass;
}

template <int NF>
PM_HDI void inj_candidate(const InjJet<NF>& j, int64_t k, InjDrop& o) {
  double radp = inj_uniform(j.seed, j.call, k, 0);
  if (j.hollow) radp = 1.;
  const double phi_radial = inj_uniform(j.seed, j.call, k, 1) * 2. * kInjPI;
  const double cur_rad = radp * j.jet_dia / 2.;
  // get_new_particle (SprayJet.cpp:148-184)
  const double radp2 = 2. * cur_rad / j.jet_dia;
  double theta_spread = radp2 * j.spread / 2.;
  if (j.hollow) {
    const double r = inj_uniform(j.seed, j.call, k, 2) - 0.5;
    theta_spread += j.hollow_spread * r;
  }
  const double phi_swirl = j.swirl;
  const double dia = inj_dia(j, k);
  o.dia = dia;
  o.accept = inj_accept(j, dia);
  if (!o.accept) {
    o.mass = 0.;
    return;
  }
  // transform_loc_vel (SprayJet.H:165-221, 3D)
  double sp2 = pm::sin(phi_radial);
  double cp2 = pm::cos(phi_radial);
  const double st2 = pm::sin(theta_spread);
  const double ct2 = pm::cos(theta_spread);
  const double dp[3] = {j.cp1 * cp2 * j.ct1 - j.sp1 * sp2, j.sp1 * cp2 * j.ct1 + j.cp1 * sp2, -j.st1 * cp2};
  const double phivel = phi_radial + phi_swirl;
  sp2 = pm::sin(phivel);
  cp2 = pm::cos(phivel);
  const double v1 = j.st1 * ct2 + st2 * cp2 * j.ct1;
  double vel[3] = {j.cp1 * v1 - j.sp1 * sp2 * st2, j.sp1 * v1 + sp2 * st2 * j.cp1, j.ct1 * ct2 - j.st1 * st2 * cp2};
  double loc[3];
  for (int d = 0; d < 3; ++d) {
    loc[d] = j.cent[d] + cur_rad * dp[d];
    vel[d] *= j.umag;
    o.vel[d] = vel[d];
  }
  const double pmov = inj_uniform(j.seed, j.call, k, 5);
  for (int d = 0; d < 3; ++d) o.pos[d] = loc[d] + pmov * j.dt * vel[d];
  o.mass = inj_parcel_mass(j, dia);
}

// AMReX Where() on one box covering the domain: the containing cell
// floor((x - plo) / dx) must be a domain cell
PM_HDI bool inj_inside(const double* pos, const double* plo, const double* dxi, const int* ncell) {
  for (int d = 0; d < 3; ++d) {
    const double c = std::floor((pos[d] - plo[d]) * dxi[d]);
    if (!(c >= 0. && c < static_cast<double>(ncell[d]))) return false;
  }
  return true;
}

}  // namespace pelemp
