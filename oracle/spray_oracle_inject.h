// spray_oracle_inject.h -- CPU restatement of droplet injection and
// initialisation.
//
// TEST INFRASTRUCTURE (see spray_oracle.h).  Restates, as one sequential
// loop in the reference's order:
//   Source/PP_Spray/SprayInjection.H:10-180   sprayInjection
//   Source/PP_Spray/SprayInjection.H:198-348  uniformSprayInit (one rank)
//   Source/PP_Spray/SprayJet.H:71-86,165-239  jet_active, transform_loc_vel,
//                                              interpolateROI
//   Source/PP_Spray/SprayJet.cpp:148-184      get_new_particle
//   Source/PP_Spray/Distribution/Distributions.cpp:1-174
//
// Deviations (documented): amrex::Random() / RandomNormal() are replaced by
// counter-based uniforms keyed by (jet seed, injection call, candidate,
// draw slot) and a Box-Muller normal; every candidate has the fixed slots
// 0 radial fraction, 1 azimuth, 2 hollow spread, 3-4 diameter, 5 travelled
// fraction of dt.  The ChiSquared table walk stops at its last bin (the
// reference reads past it when the uniform exceeds the last CDF value); the
// ROI bracket index is kept inside the table; uniformSprayInit sets the
// droplet mass fractions from Y_part (the reference leaves them
// uninitialised).  PeleC jets (CGS) are converted to SI with the library's
// unit conversions (pelemp_amd/csrc/units.h), rates interpolated in CGS.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "../include/pelemp_spray.h"
#include "../pelemp_amd/csrc/pm_math.h"
#include "../pelemp_amd/csrc/units.h"

namespace oracle {

namespace inj {

constexpr double PI = 3.14159265358979323846;
constexpr double MIN_MASS = 1.0e-16;  // SprayUnits::min_mass, kg

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// amrex::Random() for draw `slot` of candidate `cand` of injection `call`
inline double random01(uint64_t seed, int64_t call, int64_t cand, int slot) {
  const uint64_t a = mix64(seed ^ mix64(static_cast<uint64_t>(call)));
  const uint64_t r = mix64(a ^ mix64(static_cast<uint64_t>(cand) * 8ull + static_cast<uint64_t>(slot)));
  return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
}

// amrex::RandomNormal(mean, std) from the two diameter draws
inline double random_normal(double mean, double sd, double u1, double u2) {
  const double z = std::sqrt(-2. * pm::log(1. - u1)) * pm::cos(2. * PI * u2);
  return mean + sd * z;
}

// DistBase and its five registered distributions
struct Dist {
  int type = PELEMP_DIST_UNIFORM;
  double m_diam = 0., m_mean = 0., m_std = 0., m_log_mean = 0., m_log_std = 0., m_k = 0., m_d32 = 0.;
  double rvals[100] = {};

  void init(int t, const double* p) {
    type = t;
    switch (t) {
      case PELEMP_DIST_UNIFORM: m_diam = p[0]; break;
      case PELEMP_DIST_NORMAL:
        m_mean = p[0];
        m_std = p[1];
        break;
      case PELEMP_DIST_LOGNORMAL: {  // Distributions.cpp:67-75
        const double mean = p[0], sd = p[1];
        m_mean = mean;
        m_log_mean = 2. * pm::log(mean) - 0.5 * pm::log(sd * sd + mean * mean);
        m_log_std = std::sqrt(std::max(-2. * pm::log(mean) + pm::log(sd * sd + mean * mean), 0.));
        break;
      }
      case PELEMP_DIST_WEIBULL:
        m_mean = p[0];
        m_k = p[1];
        break;
      default: {  // ChiSquared::init, Distributions.cpp:134-149
        m_d32 = p[0];
        const double xiend = 12.;
        const double dxi = xiend / 100.;
        const double rend = 1. - pm::exp(-xiend) * (1. + xiend * (1. + xiend * (0.5 + xiend / 6.)));
        for (int i = 0; i < 100; i++) {
          const double xi = dxi * (i + 1);
          const double rval = 1. - pm::exp(-xi) * (1. + xi * (1. + xi * (0.5 + xi / 6.)));
          rvals[i] = rval / rend;
        }
      }
    }
  }

  double chi_walk(double fact) const {
    const double dmean = m_d32 / 3.;
    const double dxi = 12. / 100.;
    int curn = 0;
    double curr = rvals[0];
    double curxi = 0.;
    while (fact > curr && curn < 99) {
      curn++;
      curr = rvals[curn];
      curxi += dxi;
    }
    return curxi * dmean;
  }

  double get_dia(double u1, double u2) const {
    switch (type) {
      case PELEMP_DIST_UNIFORM: return m_diam;
      case PELEMP_DIST_NORMAL: return random_normal(m_mean, m_std, u1, u2);
      case PELEMP_DIST_LOGNORMAL: return pm::exp(random_normal(m_log_mean, m_log_std, u1, u2));
      case PELEMP_DIST_WEIBULL: {
        const double fact = -pm::log(1. - u1);
        return m_mean * pm::pow(fact, 1. / m_k);
      }
      default: return chi_walk(u1);
    }
  }

  double get_avg_dia() const {
    switch (type) {
      case PELEMP_DIST_UNIFORM: return m_diam;
      case PELEMP_DIST_CHISQUARED: return chi_walk(0.5);
      default: return m_mean;
    }
  }
};

// interpolateROI (SprayJet.H:223-239)
inline double interpolate_roi(const pelemp_spray_jet& j, double time, const double* vals) {
  int i = 0;
  double ctime = j.roi_time[0];
  while (ctime < time && i < j.n_roi - 1) {
    ++i;
    ctime = j.roi_time[i];
  }
  if (i == 0) i = 1;
  const double time1 = j.roi_time[i - 1], time2 = j.roi_time[i];
  const double invt = (time - time1) / (time2 - time1);
  return vals[i - 1] + (vals[i] - vals[i - 1]) * invt;
}
inline double jet_vel(const pelemp_spray_jet& j, double time) {
  return j.n_roi > 0 ? interpolate_roi(j, time, j.roi_vel) : j.jet_vel;
}
inline double mass_flow_rate(const pelemp_spray_jet& j, double time) {
  return j.n_roi > 0 ? interpolate_roi(j, time, j.roi_mdot) : j.mass_flow;
}

// AMReX Where() for one box covering the domain
inline bool where(const pelemp_geom& g, const double* p) {
  for (int d = 0; d < 3; ++d) {
    const double c = std::floor((p[d] - g.plo[d]) * (1. / g.dx[d]));
    if (!(c >= 0. && c < static_cast<double>(g.domain_hi[d] - g.domain_lo[d] + 1))) return false;
  }
  return true;
}

}  // namespace inj

// sprayInjection for one jet.  sd and g are SI; the jet is in the flavour's
// units (CGS when pelec).  Appends the new droplets (AoS) to pos/rd/ids with
// ids next_id, next_id+1, ...; updates the jet's state.  Returns PELEMP_OK,
// PELEMP_E_BAD_INJECTION.
inline int spray_injection(const pelemp_spray_data& sd, const pelemp_geom& g, pelemp_spray_jet& jet_in, double time,
                           double sim_dt, int64_t next_id, bool pelec, std::vector<double>& pos,
                           std::vector<double>& rd, std::vector<int64_t>& ids) {
  using namespace inj;
  const int nf = sd.nfuel;
  const int nsr = 10 + nf;
  // jet_active (SprayJet.H:71-86) and SprayInjection.H:17-22
  if (time < jet_in.start_time || time > jet_in.end_time) return PELEMP_OK;
  if (sim_dt <= 0. || jet_vel(jet_in, time) < 0.) return PELEMP_OK;
  pelemp_spray_jet jet = jet_in;
  double mdot = mass_flow_rate(jet_in, time);
  double umag = jet_vel(jet_in, time);
  if (pelec) {
    pelemp::cgs::jet_to_si(jet);
    mdot *= pelemp::cgs::kMass;
    umag *= pelemp::cgs::kLen;
  }
  Dist dist;
  dist.init(jet.dist_type, jet.dist_param);
  // SprayInjection.H:28-43
  double rho_avg = 0.;
  for (int spf = 0; spf < nf; ++spf) rho_avg += (nf > 1 ? jet.jet_Y[spf] : 1.) / rhoL(sd, jet.jet_T, spf);
  rho_avg = 1. / rho_avg;
  const double Pi_six = PI / 6.;
  const double avg_dia = dist.get_avg_dia();
  const double avg_mass = Pi_six * rho_avg * pm::pow(avg_dia, 3.);
  const double min_dia = pm::cbrt(6. * MIN_MASS / (PI * rho_avg));
  if (avg_dia < min_dia || avg_mass < MIN_MASS) return PELEMP_E_BAD_INJECTION;
  const double initial_bm2 = sd.do_breakup == 2 ? -1. : 0.;
  // SprayInjection.H:49-73
  const double min_inj_parcel = jet.min_parcel;
  double dt = sim_dt;
  const double inject_mass0 = mdot * sim_dt;
  double inject_mass = inject_mass0;
  if (jet.sum_inj_mass > 0.) {
    dt += jet.sum_inj_time;
    inject_mass += jet.sum_inj_mass;
  }
  double num_ppp = jet.num_ppp;
  if (num_ppp < 0.) {
    num_ppp = inject_mass / avg_mass;
    jet.num_ppp = num_ppp;
  }
  if (inject_mass / (num_ppp * avg_mass) < min_inj_parcel) {
    jet.sum_inj_mass = inject_mass;
    jet.sum_inj_time = dt;
  } else {
    // transform_loc_vel's jet angles (SprayJet.H:175-181)
    const double norm_mag =
        std::sqrt(jet.norm[0] * jet.norm[0] + jet.norm[1] * jet.norm[1] + jet.norm[2] * jet.norm[2]);
    const double theta_jet = pm::acos(jet.norm[2] / norm_mag);
    const double phi_jet = pm::atan2(jet.norm[1] / norm_mag, jet.norm[0] / norm_mag);
    const double sp1 = pm::sin(phi_jet), cp1 = pm::cos(phi_jet);
    const double st1 = pm::sin(theta_jet), ct1 = pm::cos(theta_jet);
    const int64_t call = jet.n_injections;
    double cur_mass = 0.;
    int64_t cand = 0;
    const int64_t max_cand = 67108864;
    while (cur_mass < inject_mass) {
      if (cand >= max_cand) return PELEMP_E_BAD_INJECTION;
      const int64_t k = cand++;
      double radp = random01(jet.seed, call, k, 0);
      if (jet.hollow_spray) radp = 1.;
      const double phi_radial = random01(jet.seed, call, k, 1) * 2. * PI;
      const double cur_rad = radp * jet.jet_dia / 2.;
      // get_new_particle (SprayJet.cpp:148-184)
      const double radp2 = 2. * cur_rad / jet.jet_dia;
      double theta_spread = radp2 * jet.spread_angle / 2.;
      if (jet.hollow_spray) theta_spread += jet.hollow_spread * (random01(jet.seed, call, k, 2) - 0.5);
      const double phi_swirl = jet.swirl_angle;
      const double dia_part = dist.get_dia(random01(jet.seed, call, k, 3), random01(jet.seed, call, k, 4));
      const double T_part = jet.jet_T;
      if (!(dia_part > min_dia && umag > 0.)) continue;
      // transform_loc_vel (SprayJet.H:165-221)
      double sp2 = pm::sin(phi_radial), cp2 = pm::cos(phi_radial);
      const double st2 = pm::sin(theta_spread), ct2 = pm::cos(theta_spread);
      const double dp[3] = {cp1 * cp2 * ct1 - sp1 * sp2, sp1 * cp2 * ct1 + cp1 * sp2, -pm::sin(theta_jet) * cp2};
      const double phivel = phi_radial + phi_swirl;
      sp2 = pm::sin(phivel);
      cp2 = pm::cos(phivel);
      const double v1 = st1 * ct2 + st2 * cp2 * ct1;
      double vel[3] = {cp1 * v1 - sp1 * sp2 * st2, sp1 * v1 + sp2 * st2 * cp1, ct1 * ct2 - st1 * st2 * cp2};
      double loc[3];
      for (int d = 0; d < 3; ++d) {
        loc[d] = jet.cent[d] + cur_rad * dp[d];
        vel[d] *= umag;
      }
      // SprayInjection.H:117-155
      std::vector<double> r(nsr, 0.);
      for (int d = 0; d < 3; ++d) r[d] = vel[d];
      r[3] = T_part;
      r[4] = dia_part;
      double rho_part = 0.;
      if (nf > 1) {
        for (int spf = 0; spf < nf; ++spf) {
          r[5 + spf] = jet.jet_Y[spf];
          rho_part += jet.jet_Y[spf] / rhoL(sd, T_part, spf);
        }
        rho_part = 1. / rho_part;
      } else {
        rho_part = rhoL(sd, T_part, 0);
        r[5] = 1.;
      }
      const double pmov = random01(jet.seed, call, k, 5);
      double p[3];
      for (int d = 0; d < 3; ++d) p[d] = loc[d] + pmov * dt * r[d];
      const double pmass = Pi_six * rho_part * pm::pow(dia_part, 3.);
      r[7 + nf] = 0.;
      r[8 + nf] = initial_bm2;
      r[9 + nf] = 0.;
      r[6 + nf] = num_ppp;
      r[5 + nf] = num_ppp;
      const double new_mass = cur_mass + num_ppp * pmass;
      if (!where(g, p)) return PELEMP_E_BAD_INJECTION;
      pos.insert(pos.end(), p, p + 3);
      rd.insert(rd.end(), r.begin(), r.end());
      ids.push_back(next_id++);
      cur_mass = new_mass;
    }
    jet.n_injections += 1;
    // SprayInjection.H:160-180
    const double est_mdot = cur_mass / dt;
    if (est_mdot - mdot > 0.5 * mdot) jet.min_parcel += 1.;
    jet.total_inj_mass += cur_mass;
    jet.total_inj_time += dt;
    jet.sum_inj_mass = -1.;
    jet.sum_inj_time = -1.;
  }
  if (pelec) {
    pelemp::cgs::jet_state_to_cgs(jet, jet_in);
  } else {
    jet_in.num_ppp = jet.num_ppp;
    jet_in.sum_inj_mass = jet.sum_inj_mass;
    jet_in.sum_inj_time = jet.sum_inj_time;
    jet_in.min_parcel = jet.min_parcel;
    jet_in.total_inj_mass = jet.total_inj_mass;
    jet_in.total_inj_time = jet.total_inj_time;
    jet_in.n_injections = jet.n_injections;
  }
  return PELEMP_OK;
}

// uniformSprayInit on one rank (SprayInjection.H:198-348): lattice droplets
// in the order they reach the tiles (the reference pops its list from the
// back), ids in lattice order from next_id.  g is SI; vel and dia in the
// flavour's units.
inline int uniform_spray_init(const pelemp_spray_data& sd, const pelemp_geom& g, const int* num_part,
                              const double* vel_part, double dia_part, double T_part, const double* Y_part,
                              double num_ppp, int64_t next_id, bool pelec, std::vector<double>& pos,
                              std::vector<double>& rd, std::vector<int64_t>& ids) {
  const int nf = sd.nfuel;
  const int nsr = 10 + nf;
  const double f = pelec ? pelemp::cgs::kLen : 1.;
  std::vector<double> part_vals(nsr, 0.);
  for (int d = 0; d < 3; ++d) part_vals[d] = vel_part[d] * f;
  part_vals[3] = T_part;
  part_vals[4] = dia_part * f;
  for (int s = 0; s < nf; ++s) part_vals[5 + s] = Y_part[s];
  part_vals[5 + nf] = num_ppp;
  part_vals[6 + nf] = num_ppp;
  part_vals[7 + nf] = 0.;
  part_vals[8 + nf] = sd.do_breakup == 2 ? -1. : 0.;
  part_vals[9 + nf] = 0.;
  double dx_part[3];
  for (int d = 0; d < 3; ++d) dx_part[d] = (g.phi[d] - g.plo[d]) / static_cast<double>(num_part[d]);
  const int64_t total = static_cast<int64_t>(num_part[0]) * num_part[1] * num_part[2];
  std::vector<double> lp(3 * total);
  for (int64_t cur = 0; cur < total; ++cur) {
    // unflatten_particles (SprayInjection.H:182-196)
    const int64_t d1 = num_part[0], d2 = num_part[1];
    int64_t cidx = cur;
    const int64_t i2 = cidx / (d1 * d2);
    cidx -= i2 * d1 * d2;
    const int64_t idx[3] = {cidx % d1, cidx / d1, i2};
    for (int d = 0; d < 3; ++d) lp[3 * cur + d] = (static_cast<double>(idx[d]) + 0.5) * dx_part[d];
    if (!inj::where(g, &lp[3 * cur])) return PELEMP_E_BAD_INJECTION;
  }
  for (int64_t cur = total - 1; cur >= 0; --cur) {
    pos.insert(pos.end(), &lp[3 * cur], &lp[3 * cur] + 3);
    rd.insert(rd.end(), part_vals.begin(), part_vals.end());
    ids.push_back(next_id + cur);
  }
  return PELEMP_OK;
}

}  // namespace oracle
