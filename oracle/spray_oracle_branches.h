// spray_oracle_branches.h -- CPU restatement of the divergent branches of
// the spray update: wall films, TAB breakup, wall reflection and the
// Ahamed splash model, and the creation of the secondary droplets.
//
// TEST INFRASTRUCTURE (see spray_oracle.h).  Restates, in the reference's
// structure and operation order:
//   Source/PP_Spray/BreakupSplash/WallFilm.H:9-196     fillFilmFab,
//                                                      calculateFilmSource
//   Source/PP_Spray/BreakupSplash/TABBreakup.H:12-225  updateBreakupTAB,
//                                                      splitDropletTAB
//   Source/PP_Spray/BreakupSplash/AhamedSplash.H:22-339 find_tangents,
//                     get_splash_vels, get_ms_theta, splash_criteria,
//                     droplet_splashing
//   Source/PP_Spray/WallFunctions.H:89-234             check_wall, impose_wall
//   Source/PP_Spray/BreakupSplash/ReitzKHRT.H:7-178    updateBreakupKHRT
//   Source/PP_Spray/SpraySB.cpp:9-207                  CreateSBDroplets
// (Cartesian walls only; the EB branches are not part of this build.)
//
// Deviations (documented):
//  * The azimuth of each TAB breakup child comes from a counter-based
//    uniform number keyed by (seed, parent id, child index) instead of
//    amrex::Random() (SpraySB.cpp:153), so CPU and GPU draw the same values.
//  * New droplet ids are max_id + 1, + 2, ... in parent order (the parents in
//    cell-sorted order) instead of ParticleType::NextID().
//  * The film thickness of a cell (fillFilmFab) is summed over the film
//    particles in cell-sorted order (the reference uses atomics).
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

// splash_breakup (SprayFuelData.H:31-39)
enum SBType : int {
  SB_NONE = 0,
  SB_KH = 1,
  SB_RT = 2,
  SB_TAB = 3,
  SB_SPLASH_DRY = 4,
  SB_SPLASH_THERMAL = 5,
  SB_SPLASH_WET = 6
};

// splash_type (AhamedSplash.H:12-19)
enum SplashType : int { REBOUND = 0, DEPOSIT, SPLASH_DRY, THERMAL_BREAKUP, SPREAD, SPLASH_WET };

// One SBPtrs entry (SBData.H:7-19): what CreateSBDroplets needs of a parent.
template <int NF>
struct SBRecord {
  int type = SB_NONE;
  double norm[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, loc[3] = {0, 0, 0};
  double T0 = 0., ref_dia = 0., num_dens = 0.;
  double Y0[NF] = {};
  double phi1 = 0., phi2 = 0., phi3 = 0.;
};

// counter-based uniform number in [0, 1) for breakup child azimuths
inline uint64_t sb_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double sb_uniform(uint64_t seed, int64_t parent_id, int child) {
  const uint64_t r = sb_mix64(seed ^ sb_mix64(static_cast<uint64_t>(parent_id) * 0x10000ull +
                                              static_cast<uint64_t>(child)));
  return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
}

inline double vlen(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
inline double vdot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void vcross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------------------
// Wall film (WallFilm.H)

// Film volume per face area of one film particle (fillFilmFab, WallFilm.H:9-26)
template <int NF>
double film_fab_contribution(const Particle<NF>& p, double face_area) {
  using C = Comps<NF>;
  const double film_hght = p.rdata[C::pstateFilmHght];
  const double film_dia = p.rdata[C::pstateDia];
  const double film_vol = kPI * film_hght * (0.75 * film_dia * film_dia + film_hght * film_hght);
  return film_vol / face_area;
}

// calculateFilmSource (WallFilm.H:28-196), PELELM_USE_SPRAY flavour
template <class Mech, int NF>
void calculateFilmSource(double flow_dt, GasPhaseVals<Mech::NS, NF>& gpv, const pelemp_spray_data& fdat,
                         Particle<NF>& p, const double* cBoilT) {
  constexpr int NS = Mech::NS;
  using C = Comps<NF>;
  const double dt = flow_dt * fdat.dtmod;
  const double rule = 1. / 2.;
  const double C_eps = 1.E-15;
  const double min_height = 1.E-7;  // 1e-5 * len_conv
  double Y_skin[NS], h_film[NS], cp_n[NS], Ddiag[NS];
  double L_fuel[NF] = {0.0}, mi_dot[NF] = {0.0}, Y_film[NF], X_vapor[NF] = {0.0};
  for (int n = 0; n < NS; ++n) Ddiag[n] = 0.;
  double T_film = p.rdata[C::pstateT];
  double rho_film = 0., Tcrit = 0., cp_film = 0., lambda_film = 0., mw_film = 0.;
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = std::min(T_film, cBoilT[spf]);
    Y_film[spf] = p.rdata[C::pstateY + spf];
    rho_film += Y_film[spf] / rhoL(fdat, minT, spf);
    cp_film += Y_film[spf] * fdat.cp[spf];
    lambda_film += Y_film[spf] * lambdaL(fdat, minT, spf);
    Tcrit += Y_film[spf] * fdat.critT[spf];
    mw_film += Y_film[spf] / gpv.mw[fdat.indx[spf]];
  }
  mw_film = 1. / mw_film;
  rho_film = 1. / rho_film;
  T_film = std::min(0.999 * Tcrit, T_film);
  double film_height = p.rdata[C::pstateFilmHght];
  const double film_dia = p.rdata[C::pstateDia];
  const double rad2 = (0.5 * film_dia) * (0.5 * film_dia);
  const double film_area = kPI * rad2;
  double film_mass = rho_film * film_height * film_area;
  const double start_mass = film_mass;
  const double delT = std::max(gpv.T_fluid - T_film, 0.);
  const double T_skin = T_film + rule * delT;
  GasProps<Mech>::T2Cpi(T_skin, cp_n);
  GasProps<Mech>::T2Hi(T_film, h_film);
  for (int n = 0; n < NS; ++n) Y_skin[n] = 0.;
  double cp_skin = 0., mw_skin = 0., B_M = 0., sumXVap = 0.;
  calcVaporState<NS, NF>(fdat, gpv, rule, T_film, C_eps, mw_film, Y_film, h_film, cp_n, cBoilT, Y_skin, X_vapor,
                         L_fuel, B_M, sumXVap, cp_skin, mw_skin);
  double lambda_skin = 0., mu_skin = 0.;
  const double rho_skin = mw_skin * gpv.p_fluid / (RU * T_skin);
  GasProps<Mech>::transport(false, true, true, T_skin, rho_skin, Y_skin, Ddiag, mu_skin, lambda_skin);
  if (sumXVap > 0.) {
    const double h_heat = lambda_skin / film_height;
    const double q_conv = film_area * h_heat * (gpv.T_fluid - T_film);
    const double q_cond = -film_area * lambda_film * 2. * (T_film - fdat.wall_T) / film_height;
    const double h_mass = film_area * h_heat * pm::cbrt(lambda_skin / cp_skin) * pm::log1p(B_M) / lambda_skin;
    double sumL = 0., m_dot = 0.;
    for (int spf = 0; spf < NF; ++spf) {
      const int fspec = fdat.indx[spf];
      Ddiag[fspec] *= mw_skin / gpv.mw[fspec];
      if (X_vapor[spf] > 0.) {
        const double cur_rhoD = X_vapor[spf] * Ddiag[fspec] / sumXVap;
        mi_dot[spf] = -std::max(h_mass * pm::pow(cur_rhoD, 2. / 3.), 0.);
        m_dot += mi_dot[spf];
        sumL += mi_dot[spf] * L_fuel[spf];
      }
    }
    const double part_temp_src = (q_conv + q_cond + sumL) / (film_mass * cp_film);
    gpv.fluid_eng_src += q_conv;
    T_film += dt * part_temp_src;
    double new_mass = film_mass + dt * m_dot;
    film_height = new_mass / (rho_film * film_area);
    if (film_height > min_height) {
      if (NF > 1) {
        rho_film = 0.;
        for (int spf = 0; spf < NF; ++spf) {
          Y_film[spf] = std::min(1., std::max((Y_film[spf] * film_mass + mi_dot[spf] * dt) / new_mass, 0.));
          if (Y_film[spf] < 1.E-12) Y_film[spf] = 0.;
          rho_film += Y_film[spf] / rhoL(fdat, T_film, spf);
        }
        rho_film = 1. / rho_film;
      } else {
        rho_film = rhoL(fdat, T_film, 0);
      }
    } else {
      film_height = 0.;
      new_mass = 0.;
      p.id = -1;
    }
    film_mass = new_mass;
  }
  p.rdata[C::pstateT] = T_film;
  const double mdot_total = (film_mass - start_mass) / dt;
  gpv.fluid_mass_src = mdot_total;
  for (int spf = 0; spf < NF; ++spf) {
    const int fdspec = fdat.dep_indx[spf];
    const double oldY = p.rdata[C::pstateY + spf];
    const double newY = Y_film[spf];
    const double midot = (newY * film_mass - oldY * start_mass) / dt;
    gpv.fluid_Y_dot[spf] = midot;
    p.rdata[C::pstateY + spf] = newY;
    gpv.fluid_eng_src += midot * h_film[fdspec];
  }
  p.rdata[C::pstateDia] = film_dia;
  p.rdata[C::pstateFilmHght] = film_height;
}

// ---------------------------------------------------------------------------
// TAB breakup (TABBreakup.H)

inline double tab_drag_coef(double Reyn) {
  if (Reyn > 1000.) return 0.424;
  if (Reyn > 1.) return 24. / Reyn * (1. + pm::cbrt(Reyn * Reyn) / 6.);
  if (Reyn > 0.) return 24. / Reyn;
  return 0.;
}

// updateBreakupTAB (TABBreakup.H:12-184); returns the accumulated
// tangential velocity of the breakups over dt.
template <int NS, int NF>
double updateBreakupTAB(double Reyn_d, double dt, const double* cBoilT, const GasPhaseVals<NS, NF>& gpv,
                        const pelemp_spray_data& fdat, Particle<NF>& p) {
  using C = Comps<NF>;
  const double C_k = 8.;
  const double C_d = 10.;
  const double C_b = 0.5;
  const double C_F = 1. / 3.;
  const double Wet = 80.;
  const double k2 = 2. / 9.;
  const double vel_part[3] = {p.rdata[C::pstateVel], p.rdata[C::pstateVel + 1], p.rdata[C::pstateVel + 2]};
  const double T_part = p.rdata[C::pstateT];
  const double dia_part = p.rdata[C::pstateDia];
  double num_dens = p.rdata[C::pstateNumDens];
  double rho_part = 0., mu_part = 0.;
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = std::min(T_part, cBoilT[spf]);
    const double Y_part = p.rdata[C::pstateY + spf];
    rho_part += Y_part / rhoL(fdat, minT, spf);
    mu_part += Y_part * muL(fdat, minT, spf);
  }
  rho_part = 1. / rho_part;
  double Utan_total = 0.;
  const double sigma = fdat.sigma;
  double rad_part = 0.5 * dia_part;
  const double min_rad = 4. * pm::cbrt(kMinMass * 3. / (4. * kPI * rho_part));
  if (rad_part < min_rad) return 0.;
  double diff_vel[3];
  for (int d = 0; d < 3; ++d) diff_vel[d] = gpv.vel_fluid[d] - vel_part[d];
  const double dv2 = diff_vel[0] * diff_vel[0] + diff_vel[1] * diff_vel[1] + diff_vel[2] * diff_vel[2];
  const double dvl = std::sqrt(dv2);
  const double We_div_r = gpv.rho_fluid * dv2 / sigma;
  const double We_crit = C_k * C_b / C_F;
  double denom = rho_part * rad_part * rad_part;
  double td = 2. * denom / (C_d * mu_part);
  double omega2 = C_k * sigma / (denom * rad_part) - 1. / (td * td);
  const double tbconst = 2. * std::sqrt(3. * rho_part / gpv.rho_fluid) / dvl;
  if (omega2 <= 0.) {
    p.rdata[C::pstateBM1] = 0.;
    p.rdata[C::pstateBM2] = 0.;
    return 0.;
  }
  double omega = std::sqrt(omega2);
  double yn = p.rdata[C::pstateBM1];
  double ydotn = p.rdata[C::pstateBM2];
  double Reyn = Reyn_d;
  double C_D = tab_drag_coef(Reyn);
  const double tb_estconst = std::sqrt(3. * rho_part / gpv.rho_fluid) / dvl;
  double tb_est = 0.1 * rad_part * tb_estconst;
  double subdt = std::min(dt, tb_est);
  double curt = 0.;
  while (curt < dt) {
    const double We = We_div_r * rad_part;
    const double Wer = We / We_crit;
    const double A2 = (yn - Wer) * (yn - Wer) + ydotn * ydotn / omega2;
    const double A = std::sqrt(A2);
    const double ex = pm::exp(-subdt / td);
    const double cs = pm::cos(omega * subdt), sn = pm::sin(omega * subdt);
    const double ynp = Wer + ex * ((yn - Wer) * cs + 1. / omega * (ydotn + (yn - Wer) / td) * sn);
    const double ydotnp =
        (Wer - ynp) / td + omega * ex * (1. / omega * (ydotn + (yn - Wer) / td) * cs - (yn - Wer) * sn);
    if (Wer + A <= 1.) {
      yn = ynp;
      ydotn = ydotnp;
    } else {
      double tbv = 0.;
      if (std::fabs(yn) < 1.) {
        const double pv1 = (yn - Wer);
        const double pv2 = -ydotn / omega;
        double phi = pm::atan2(pv2, pv1);
        if (phi < 0.) phi = 2. * kPI + phi;
        double theta = pm::acos((1. - Wer) / A);
        if (theta < phi) {
          if (2. * kPI - theta >= phi) theta = -theta;
          theta += 2. * kPI;
        }
        tbv = (theta - phi) / omega;
      }
      if (tbv < subdt) {
        const double k1 = k2 * ((std::sqrt(Wet) - 1.) * pm::pow(We / Wet, 4.) + 1.);
        double Kbr = k1 * omega;
        if (We >= Wet) Kbr = k2 * omega * std::sqrt(We);
        const double etheta = std::max(-1., std::min(1., 1. - 1. / Wer));
        const double etb = pm::acos(etheta) / omega;
        const double rchild = rad_part * pm::exp(-Kbr * etb);
        num_dens *= pm::pow(rad_part / rchild, 3.);
        const double rsmr = std::sqrt(pm::pow(rad_part, 3.) / (rchild * rchild));
        const double AE2 = 3. * (1. - rad_part / rsmr + 5. * C_D * We / 72.) * omega2;
        const double AE = std::sqrt(AE2);
        const double Utan = AE * C_b * rad_part;
        Utan_total += Utan;
        Reyn = Reyn * rchild / rad_part;
        C_D = tab_drag_coef(Reyn);
        rad_part = rchild;
        yn = 0.;
        ydotn = 0.;
        tb_est = rad_part * tbconst;
        subdt = std::min(dt, 0.1 * tb_est);
        denom = rho_part * rad_part * rad_part;
        td = 2. * denom / (C_d * mu_part);
        omega2 = C_k * sigma / (denom * rad_part) - 1. / (td * td);
        if (omega2 <= 0.) {
          p.rdata[C::pstateDia] = 2. * rad_part;
          p.rdata[C::pstateNumDens] = num_dens;
          p.rdata[C::pstateBM1] = 0.;
          p.rdata[C::pstateBM2] = 0.;
          return Utan_total;
        }
        omega = std::sqrt(omega2);
      } else {
        yn = ynp;
        ydotn = ydotnp;
      }
    }
    if (curt + subdt > dt) subdt = dt - curt;
    curt += subdt;
  }
  p.rdata[C::pstateDia] = 2. * rad_part;
  p.rdata[C::pstateNumDens] = num_dens;
  p.rdata[C::pstateBM1] = yn;
  p.rdata[C::pstateBM2] = ydotn;
  return Utan_total;
}

// splitDropletTAB (TABBreakup.H:186-225)
template <int NF>
void splitDropletTAB(Particle<NF>& p, double max_num_ppp, double Utan_total, SBRecord<NF>& rf) {
  using C = Comps<NF>;
  if (p.id <= 0) return;
  const double num_dens = p.rdata[C::pstateNumDens];
  if (num_dens > max_num_ppp) {
    rf.type = SB_TAB;
    const double vel_part[3] = {p.rdata[C::pstateVel], p.rdata[C::pstateVel + 1], p.rdata[C::pstateVel + 2]};
    const double velMag = vlen(vel_part);
    for (int d = 0; d < 3; ++d) {
      rf.loc[d] = p.pos[d];
      rf.vel[d] = vel_part[d];
      rf.norm[d] = vel_part[d] / velMag;
    }
    rf.num_dens = p.rdata[C::pstateNumDens];
    rf.ref_dia = p.rdata[C::pstateDia];
    rf.phi1 = Utan_total;
    rf.phi2 = p.rdata[C::pstateBM1];
    rf.phi3 = p.rdata[C::pstateBM2];
    rf.T0 = p.rdata[C::pstateT];
    for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = p.rdata[C::pstateY + spf];
    p.id = -1;
  }
}

// ---------------------------------------------------------------------------
// KH-RT wave breakup (ReitzKHRT.H:7-178).  BM1 holds the shed mass (in
// units of N0 r^3), BM2 the RT clock (negative: RT off until the first KH
// event).  dt is dtmod * flow_dt (SprayParticles.cpp:407-410).
template <int NS, int NF>
void updateBreakupKHRT(Particle<NF>& p, double Reyn_d, double dt, const double* cBoilT, double avg_inject_mass,
                       double B0, double B1, double C3, const GasPhaseVals<NS, NF>& gpv,
                       const pelemp_spray_data& fdat, SBRecord<NF>& rf, bool make_new_drops) {
  using C = Comps<NF>;
  const double mthresh_jet = 0.03;
  const double mthresh_drop = 0.9;
  const double We_crit = 6.;
  const double A1 = 0.188;
  const double vel_part[3] = {p.rdata[C::pstateVel], p.rdata[C::pstateVel + 1], p.rdata[C::pstateVel + 2]};
  const double T_part = p.rdata[C::pstateT];
  const double dia_part = p.rdata[C::pstateDia];
  double num_dens = p.rdata[C::pstateNumDens];
  double rho_part = 0., mu_part = 0.;
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = std::min(T_part, cBoilT[spf]);
    const double Y_part = p.rdata[C::pstateY + spf];
    rho_part += Y_part / rhoL(fdat, minT, spf);
    mu_part += Y_part * muL(fdat, minT, spf);
  }
  rho_part = 1. / rho_part;
  const double min_rad = 4. * pm::cbrt(kMinMass * 3. / (4. * kPI * rho_part));
  double shed_mass = p.rdata[C::pstateBM1];
  double N0 = p.rdata[C::pstateN0];
  double rt_time = p.rdata[C::pstateBM2];
  const double sigma = fdat.sigma;
  double rad_part = 0.5 * dia_part;
  if (rad_part < min_rad) return;
  const double rad_part3 = pm::pow(rad_part, 3.);
  double diff_vel[3];
  for (int d = 0; d < 3; ++d) diff_vel[d] = gpv.vel_fluid[d] - vel_part[d];
  const double vel_mag = vlen(diff_vel);
  const double vel_sq = vel_mag * vel_mag;
  const double C_D = tab_drag_coef(Reyn_d);  // same piecewise law, ReitzKHRT.H:72-79
  const double We_g = gpv.rho_fluid * vel_sq * rad_part / sigma;
  const double Z = mu_part / std::sqrt(rho_part * rad_part * sigma);
  const double T = Z * std::sqrt(We_g);
  bool breakupRT = false;
  if (rt_time >= 0.) {
    const double a = 3. / 8. * C_D * gpv.rho_fluid * vel_sq / (rho_part * rad_part);
    const double momdiff = a * (rho_part - gpv.rho_fluid);
    const double omega_RT = std::sqrt(2. / (3. * std::sqrt(3. * sigma)) * momdiff * std::sqrt(momdiff) /
                                      (rho_part + gpv.rho_fluid));
    const double KRT = std::sqrt(momdiff / (3. * sigma));
    const double lam_RT = C3 * kPI / KRT;
    if (rad_part < lam_RT) {
      rt_time = 0.;
    } else {
      rt_time += dt;
    }
    const double tau_RT = 1. / omega_RT;
    if (rt_time > tau_RT && lam_RT < rad_part) {
      rt_time = 0.;
      num_dens *= pm::pow(rad_part / lam_RT, 3.);
      rad_part = lam_RT;
      breakupRT = true;
    }
  }
  if (!breakupRT) {
    const double lam_KH = rad_part * 9.02 * (1. + 0.45 * std::sqrt(Z)) * (1. + 0.4 * pm::pow(T, 0.7)) /
                          pm::pow(1. + 0.865 * pm::pow(We_g, 1.67), 0.6);
    const double omega_KH = (0.34 + 0.385 * We_g * std::sqrt(We_g)) / ((1. + Z) * (1. + 1.4 * pm::pow(T, 0.6))) *
                            std::sqrt(sigma / (rho_part * rad_part3));
    const double tau_KH = 3.726 * B1 * rad_part / (lam_KH * omega_KH);
    double rs = B0 * lam_KH;
    if (rs < rad_part && We_g > We_crit) {
      const double rp = (dt / tau_KH * rs + rad_part) / (1. + dt / tau_KH);
      const double rp3 = rp * rp * rp;
      double rs3 = rs * rs * rs;
      double np = num_dens * rad_part3 / rp3;
      rt_time = 0.;
      shed_mass += N0 * (rad_part3 - rp3);
      double ns = shed_mass / rs3;
      bool create_child_drops = false;
      if (make_new_drops) {
        const double rel_shed_drop = shed_mass / (num_dens * rad_part3);
        const double rel_shed_jet = 8. * shed_mass / avg_inject_mass;
        if (rel_shed_drop > mthresh_drop) {
          create_child_drops = true;
          shed_mass = num_dens * rad_part3;
          rs = std::max(rs, min_rad);
          rs3 = rs * rs * rs;
          ns = shed_mass / rs3;
          p.id = -1;
        } else if (rel_shed_jet > mthresh_jet && rs > min_rad && ns > num_dens) {
          create_child_drops = true;
          np = (np * rp3 - ns * rs3) / rp3;
          N0 = np;
          shed_mass = 0.;
        }
      }
      if (create_child_drops) {
        const double Utan = A1 * omega_KH * lam_KH;
        rf.type = SB_KH;
        for (int d = 0; d < 3; ++d) {
          rf.loc[d] = p.pos[d];
          rf.vel[d] = vel_part[d];
          rf.norm[d] = vel_part[d] / vel_mag;  // droplet velocity over the relative speed, as written
        }
        rf.num_dens = ns;
        rf.ref_dia = 2. * rs;
        rf.T0 = T_part;
        rf.phi1 = Utan;
        rf.phi2 = 0.;
        rf.phi3 = 0.;
        for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = p.rdata[C::pstateY + spf];
      }
      rad_part = rp;
      num_dens = np;
    }
  }
  p.rdata[C::pstateDia] = 2. * rad_part;
  p.rdata[C::pstateNumDens] = num_dens;
  p.rdata[C::pstateN0] = N0;
  p.rdata[C::pstateBM1] = shed_mass;
  p.rdata[C::pstateBM2] = rt_time;
}

// ---------------------------------------------------------------------------
// Splash model (AhamedSplash.H) and walls (WallFunctions.H)

// find_tangents (AhamedSplash.H:22-44), 3D
inline void find_tangents(const double pvel[3], double tanPsi[3], const double norm[3], double tanBeta[3]) {
  double testvec[3] = {-pvel[0], -pvel[1], -pvel[2]};
  double cr[3];
  vcross(testvec, norm, cr);
  if (vlen(cr) < 1.E-5) {
    testvec[0] += 1.;
    testvec[1] += 1.;
    testvec[2] += 1.;
  }
  vcross(testvec, norm, tanPsi);
  const double l1 = vlen(tanPsi);
  for (int d = 0; d < 3; ++d) tanPsi[d] /= l1;
  vcross(tanPsi, norm, tanBeta);
  const double l2 = vlen(tanBeta);
  for (int d = 0; d < 3; ++d) tanBeta[d] /= l2;
}

// get_splash_vels (AhamedSplash.H:47-88)
inline void get_splash_vels(double U0norm, double U0tan, double Kv, double del_film, double& uBeta_0,
                            double& uBeta_half, double& uBeta_pi, double& uPsi_coeff, double& usNorm) {
  const double Un_mag = std::fabs(U0norm);
  if (del_film < 0.01) {
    const double Ut_splash = 0.0065 * Kv * pm::exp(-0.004 * Kv) * Un_mag;
    uBeta_0 = U0tan + Ut_splash;
    uBeta_half = pm::cos(kPI / 6.) * U0tan;
    uBeta_pi = pm::cos(kPI / 3.) * U0tan - Ut_splash;
    uPsi_coeff = 0.2 * U0tan + Ut_splash;
    usNorm = 0.0028 * Kv * pm::exp(-0.0062 * Kv) * Un_mag;
  } else if (del_film < 0.1) {
    double Ut_splash;
    if (del_film > 0.05) {
      Ut_splash = (5.E-5 * Kv + 0.0203) * Un_mag;
      usNorm = 0.03 * U0tan + (1.32E-4 * Kv + 0.0158) * Un_mag;
    } else {
      Ut_splash = (3.45E-5 * Kv + 0.006) * Un_mag;
      usNorm = 0.03 * U0tan + (3.57E-5 * Kv - 0.0057) * Un_mag;
    }
    uBeta_0 = 0.3 * U0tan + Ut_splash;
    uBeta_half = 0.1 * U0tan;
    uBeta_pi = 0.033 * U0tan - Ut_splash;
    uPsi_coeff = 0.06 * U0tan + Ut_splash;
  } else {
    uBeta_0 = 0.1 * U0tan;
    uBeta_half = 0.05 * U0tan;
    uBeta_pi = 0.01 * U0tan;
    uPsi_coeff = 0.02 * U0tan;
    usNorm = 0.03 * U0tan;
  }
}

// get_ms_theta (AhamedSplash.H:90-122): splashed mass in the four directions
inline void get_ms_theta(double alpha, double ms, double del_film, double* ms_thetas) {
  if (del_film > 0.) {
    const double alpha_d = alpha * 180. / kPI;
    double ms_psi_0, ms_psi_pih;
    if (del_film > 0.1) {
      ms_psi_0 = (1.9E-4 * alpha_d * alpha_d + 0.0202 * alpha_d - 0.0107) * ms;
      ms_psi_pih = (6.4E-5 * alpha_d * alpha_d - 7.6E-3 * alpha_d + 0.4176) * ms;
    } else {
      ms_psi_0 = (-2.3E-4 * alpha_d * alpha_d + 0.0236 * alpha_d) * ms;
      ms_psi_pih = (6.E-5 * alpha_d * alpha_d - 6.8E-3 * alpha_d + 0.3619) * ms;
    }
    ms_thetas[0] = ms_psi_pih;
    ms_thetas[1] = ms_psi_0;
    ms_thetas[2] = ms_psi_pih;
    ms_thetas[3] = ms - ms_psi_0 - 2. * ms_psi_pih;
  } else {
    const double tanterm = 0.4 * pm::tan(kPI / 2. - alpha);
    ms_thetas[0] = 0.25 * ms;
    ms_thetas[1] = 0.25 * ms * (1. + tanterm);
    ms_thetas[2] = 0.25 * ms;
    ms_thetas[3] = 0.25 * ms * (1. - tanterm);
  }
}

// splash_criteria (AhamedSplash.H:125-197)
inline int splash_criteria(double Kv, double Tstar, double alpha, double Kcrit, double del_film, int& N_SB,
                           double& ms, double& dm) {
  int crit = THERMAL_BREAKUP;
  if (del_film < 0.01) {
    if (Tstar < 1.1) {
      if (Kv < Kcrit) {
        crit = DEPOSIT;
        ms = 0.;
        dm = 1.;
      } else {
        crit = SPLASH_DRY;
        ms = 0.75 * (1. - pm::exp(-3.E-3 * (Kv - Kcrit)));
        dm = 1. - ms;
        N_SB = SB_SPLASH_DRY;
      }
    } else {
      if (Kv < Kcrit) {
        crit = REBOUND;
      } else {
        crit = THERMAL_BREAKUP;
        ms = 0.75 * (1. - pm::exp(-0.0175 * (Kv - Kcrit)));
        dm = 0.;
        N_SB = SB_SPLASH_THERMAL;
      }
    }
  } else {
    const double Kcrit_wet = Kcrit + 226. * pm::pow(del_film, 0.72);
    if (Kv > Kcrit_wet) {
      crit = SPLASH_WET;
      double A = 4.5, B = 0.2;
      if (del_film < 0.5) {
        A = 10.42 * del_film + 0.32;
        B = 8.75 * del_film - 0.107;
      }
      const double ms0 = A * (1. - pm::exp(-0.003 * (Kv - Kcrit_wet))) + B;
      const double alpha_d = alpha * 180. / kPI;
      if (del_film > 0.1) {
        ms = (alpha_d > 45.) ? ms0 * (-0.0305 * alpha_d + 3.7496) : ms0 * 0.0525 * alpha_d;
      } else {
        ms = (alpha_d > 45.) ? ms0 * (-7.2E-3 * alpha_d + 1.6389)
                             : ms0 * (-4.6E-4 * alpha_d * alpha_d + 0.05 * alpha_d);
      }
      ms = std::min(ms, 1.);
    } else {
      crit = SPREAD;
      ms = 0.;
      dm = 1.;
    }
  }
  return crit;
}

// droplet_splashing (AhamedSplash.H:199-339), PELELM_USE_SPRAY flavour
// (pressure in Pa -> bar with 1e-5).
template <int NF>
void droplet_splashing(Particle<NF>& p, const pelemp_spray_data& fdat, double p_gas, const double vel_part[3],
                       double par_dot, const double normal[3], const double* cBoilT, SBRecord<NF>& rf,
                       double film_h) {
  using C = Comps<NF>;
  const double U0mag = vlen(vel_part);
  const double U0norm = vdot(normal, vel_part);
  const double U0tan = std::sqrt(U0mag * U0mag - U0norm * U0norm);
  const double T_part = p.rdata[C::pstateT];
  const double num_dens = p.rdata[C::pstateNumDens];
  double mu_part = 0., rho_part = 0., Tboil = 0.;
  for (int spf = 0; spf < NF; ++spf) {
    const int ysp = C::pstateY + spf;
    const double minT = std::min(T_part, cBoilT[spf]);
    mu_part += p.rdata[ysp] * muL(fdat, minT, spf);
    rho_part += p.rdata[ysp] / rhoL(fdat, minT, spf);
    Tboil += p.rdata[ysp] * cBoilT[spf];
  }
  rho_part = 1. / rho_part;
  const double Tstar = fdat.wall_T / Tboil;
  const double dia_part = p.rdata[C::pstateDia];
  const double pmass = kPI / 6. * rho_part * pm::pow(dia_part, 3.);
  const double We = rho_part * dia_part * U0norm * U0norm / fdat.sigma;
  const double Re_L = std::fabs(U0norm) * dia_part * rho_part / mu_part;
  const double Kv = std::sqrt(We * std::sqrt(Re_L));
  const double alpha = std::max(kPI / 6., pm::asin(std::fabs(U0norm) / U0mag));
  const double p_bar = p_gas * 1.E-5;
  const double Kcrit = (26.788 * pm::log(p_bar) + 33.546) * pm::exp(-6.E-3 * (fdat.wall_T - 260.)) + 36.;
  const double del_film = film_h / dia_part;
  double ms = 0., dm = 0.;
  double impact_loc[3];
  for (int d = 0; d < 3; ++d) impact_loc[d] = p.pos[d] - par_dot * vel_part[d] / U0norm;
  int nsb = rf.type;
  const int splash_flag = splash_criteria(Kv, Tstar, alpha, Kcrit, del_film, nsb, ms, dm);
  rf.type = nsb;
  if (splash_flag == REBOUND) {
    const double rbound = 0.823 * pm::exp(-1.9835E-3 * pm::pow(Kv, 1.6));
    for (int d = 0; d < 3; ++d) {
      p.pos[d] -= (1. + rbound) * par_dot * normal[d];
      p.rdata[C::pstateVel + d] -= U0norm * (1. + rbound) * normal[d];
    }
  } else if (ms > 0.) {
    for (int d = 0; d < 3; ++d) {
      rf.loc[d] = impact_loc[d];
      rf.norm[d] = normal[d];
      rf.vel[d] = vel_part[d];
    }
    rf.ref_dia = p.rdata[C::pstateDia];
    rf.phi1 = Kv;
    rf.phi2 = pmass * ms;
    rf.phi3 = del_film;
    rf.T0 = T_part;
    rf.num_dens = num_dens;
    for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = p.rdata[C::pstateY + spf];
    if (splash_flag == THERMAL_BREAKUP) {
      const double rem_mass = pmass * (1. - ms);
      const double new_dia = pm::cbrt(6. * rem_mass / (kPI * rho_part));
      double tanBeta[3], tanPsi[3];
      find_tangents(vel_part, tanPsi, normal, tanBeta);
      const double Un_splash = 0.0028 * Kv * pm::exp(-0.0062 * Kv) * std::fabs(U0mag);
      const double Ut_splash = 1.5 * std::sqrt(3.) * U0tan / kPI;
      for (int d = 0; d < 3; ++d) {
        p.pos[d] = impact_loc[d];
        p.rdata[C::pstateVel + d] = Ut_splash * tanBeta[d] + Un_splash * normal[d];
      }
      p.rdata[C::pstateDia] = new_dia;
      p.rdata[C::pstateBM1] = 0.;
      p.rdata[C::pstateBM2] = 0.;
    }
  }
  if (dm > 0.) {
    const double film_vol = num_dens * pmass * dm / rho_part;
    const double comb_dia = pm::cbrt(film_vol * 6. / kPI);
    const double comb_We = We * comb_dia / dia_part;
    const double comb_Re = Re_L * comb_dia / dia_part;
    const double theta_c = fdat.theta_c;
    const double st = pm::sin(theta_c);
    const double film_dia =
        comb_dia * std::sqrt((12. + comb_We) /
                             (6. * (1. - pm::cos(theta_c)) / (st * st) + 4. * comb_We / std::sqrt(comb_Re)));
    p.rdata[C::pstateDia] = film_dia;
    const double film_hght = film_vol / (0.25 * kPI * film_dia * film_dia);
    p.rdata[C::pstateFilmHght] = film_hght;
    p.rdata[C::pstateN0] = 1.;
    p.rdata[C::pstateNumDens] = 1.;
    for (int d = 0; d < 3; ++d) {
      p.pos[d] = impact_loc[d];
      p.rdata[C::pstateVel + d] = 0.;
    }
  }
}

// check_wall (WallFunctions.H:89-171), Cartesian faces: pos normalised by
// dx relative to plo; returns whether the particle is behind the wall.
inline bool check_wall(const double pos[3], const int bflags[3], const double dx[3], const int ijkc[3],
                       double& par_dot, double normal[3], const EB* eb = nullptr, const pelemp_geom* g = nullptr,
                       const int* ijkc_prev = nullptr) {
  bool wall_check = false;
  double bcentv[3];
  for (int dir = 0; dir < 3; ++dir) {
    if ((bflags[dir] == -1 || bflags[dir] == 1) && !wall_check) {
      normal[dir] = -static_cast<double>(bflags[dir]);
      bcentv[dir] = static_cast<double>(ijkc[dir]) + 0.5 * static_cast<double>(1 - bflags[dir]);
      wall_check = true;
    } else {
      normal[dir] = 0.;
      bcentv[dir] = 0.;
    }
  }
  // EB walls (WallFunctions.H:124-157)
  if (eb && !wall_check) wall_check = eb_wall(*eb, *g, ijkc, ijkc_prev, normal, bcentv);
  bool refl_check = false;
  if (wall_check) {
    const double tolerance = 2.220446049250313e-16;
    par_dot = (pos[0] - bcentv[0]) * dx[0] * normal[0] + (pos[1] - bcentv[1]) * dx[1] * normal[1] +
              (pos[2] - bcentv[2]) * dx[2] * normal[2];
    if (par_dot < tolerance) refl_check = true;
  }
  return refl_check;
}

// impose_wall (WallFunctions.H:173-234), Cartesian walls
template <int NF>
void impose_wall(bool do_splash, Particle<NF>& p, const pelemp_spray_data& fdat, const pelemp_geom& g,
                 const int bflags[3], const double* cBoilT, double p_fluid, SBRecord<NF>& rf, double film_h,
                 const EB* eb = nullptr, const int* ijkc_prev = nullptr) {
  using C = Comps<NF>;
  double par_dot = 2.;
  double normal[3];
  const double vel_part[3] = {p.rdata[C::pstateVel], p.rdata[C::pstateVel + 1], p.rdata[C::pstateVel + 2]};
  double normpos[3];
  int ijkc[3];
  for (int d = 0; d < 3; ++d) {
    normpos[d] = (p.pos[d] - g.plo[d]) / g.dx[d];
    ijkc[d] = static_cast<int>(std::floor(normpos[d]));
  }
  const bool do_refl = check_wall(normpos, bflags, g.dx, ijkc, par_dot, normal, eb, &g, ijkc_prev);
  if (do_refl) {
    const double Nw_Vp = vdot(normal, vel_part);
    if (do_splash) {
      droplet_splashing<NF>(p, fdat, p_fluid, vel_part, par_dot, normal, cBoilT, rf, film_h);
    } else {
      for (int dir = 0; dir < 3; ++dir) {
        p.rdata[C::pstateVel + dir] -= 2. * Nw_Vp * normal[dir];
        p.pos[dir] -= 2. * par_dot * normal[dir];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Secondary droplets (SpraySB.cpp:9-207).  Appends the children of one
// record to `out`; next_id is advanced per child.
template <int NF>
void create_sb_droplets(const SBRecord<NF>& rf, int64_t parent_id, const pelemp_spray_data& fdat, double sub_dt,
                        double breakup_ppp_fact, uint64_t seed, int64_t& next_id,
                        std::vector<Particle<NF>>& out) {
  using C = Comps<NF>;
  const double T0 = rf.T0;
  double Y0[NF];
  double rho_part = 0.;
  if (NF > 1) {
    for (int spf = 0; spf < NF; ++spf) {
      Y0[spf] = rf.Y0[spf];
      rho_part += Y0[spf] / rhoL(fdat, T0, spf);
    }
    rho_part = 1. / rho_part;
  } else {
    rho_part = rhoL(fdat, T0, 0);
    Y0[0] = 1.;
  }
  const double U0mag = vlen(rf.vel);
  if (rf.type >= SB_SPLASH_DRY) {
    double tanPsi[3], tanBeta[3];
    find_tangents(rf.vel, tanPsi, rf.norm, tanBeta);
    const double Kv = rf.phi1, ms = rf.phi2, del_film = rf.phi3;
    const double U0norm = vdot(rf.norm, rf.vel);
    const double alpha = std::max(kPI / 6., pm::asin(std::fabs(U0norm) / U0mag));
    const double U0tan = std::sqrt(U0mag * U0mag - U0norm * U0norm);
    double uBeta_0, uBeta_half, uBeta_pi, uPsi_coeff, usNorm;
    get_splash_vels(U0norm, U0tan, Kv, del_film, uBeta_0, uBeta_half, uBeta_pi, uPsi_coeff, usNorm);
    double ms_thetas[4];
    get_ms_theta(alpha, ms, del_film, ms_thetas);
    for (int k = 0; k < 4; ++k) {
      const double psi = 0.5 * kPI * (static_cast<double>(k) - 1.);
      Particle<NF> c{};
      c.id = next_id++;
      const double new_mass = ms_thetas[k];
      const double dia = pm::cbrt(6. * new_mass / (kPI * rho_part));
      c.rdata[C::pstateDia] = dia;
      double utBeta = uBeta_half;
      if (k == 1) {
        utBeta = uBeta_0;
      } else if (k == 3) {
        utBeta = uBeta_pi;
      }
      const double utPsi = uPsi_coeff * pm::sin(psi);
      for (int d = 0; d < 3; ++d) {
        const double pvel = usNorm * rf.norm[d] + utBeta * tanBeta[d] + utPsi * tanPsi[d];
        c.pos[d] = rf.loc[d] + dia * rf.norm[d];
        c.rdata[C::pstateVel + d] = pvel;
      }
      for (int spf = 0; spf < NF; ++spf) c.rdata[C::pstateY + spf] = Y0[spf];
      c.rdata[C::pstateBM1] = 0.;
      c.rdata[C::pstateBM2] = 0.;
      c.rdata[C::pstateFilmHght] = 0.;
      c.rdata[C::pstateN0] = rf.num_dens;
      c.rdata[C::pstateNumDens] = rf.num_dens;
      c.rdata[C::pstateT] = T0;
      out.push_back(c);
    }
  } else {
    const double Utan = rf.phi1;
    const double dmean = rf.ref_dia;
    double N_s = pm::pow(rf.num_dens, breakup_ppp_fact);
    const int N_d = std::max(1, static_cast<int>(rf.num_dens / N_s));
    N_s = rf.num_dens / static_cast<double>(N_d);
    double testvec[3] = {1., 0., 0.};
    double cr[3];
    vcross(testvec, rf.norm, cr);
    if (vlen(cr) < 1.E-5) {
      testvec[0] = 0.;
      testvec[1] = 1.;
      testvec[2] = 0.;
    }
    double tanPsi[3], tanBeta[3];
    vcross(testvec, rf.norm, tanPsi);
    const double l1 = vlen(tanPsi);
    for (int d = 0; d < 3; ++d) tanPsi[d] /= l1;
    vcross(tanPsi, rf.norm, tanBeta);
    const double l2 = vlen(tanBeta);
    for (int d = 0; d < 3; ++d) tanBeta[d] /= l2;
    for (int k = 0; k < N_d; ++k) {
      const double rnd = sb_uniform(seed, parent_id, k);
      Particle<NF> c{};
      c.id = next_id++;
      c.rdata[C::pstateDia] = dmean;
      c.rdata[C::pstateT] = T0;
      for (int spf = 0; spf < NF; ++spf) c.rdata[C::pstateY + spf] = Y0[spf];
      c.rdata[C::pstateBM1] = rf.phi2;
      c.rdata[C::pstateBM2] = rf.phi3;
      c.rdata[C::pstateFilmHght] = 0.;
      c.rdata[C::pstateN0] = N_s;
      c.rdata[C::pstateNumDens] = N_s;
      const double psi = rnd * 2. * kPI;
      const double sp = pm::sin(psi), cp = pm::cos(psi);
      for (int d = 0; d < 3; ++d) {
        const double pvel = rf.vel[d] + Utan * (sp * tanPsi[d] + cp * tanBeta[d]);
        c.pos[d] = rf.loc[d] + sub_dt * pvel;
        c.rdata[C::pstateVel + d] = pvel;
      }
      out.push_back(c);
    }
  }
}

}  // namespace oracle
