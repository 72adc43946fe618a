// branches.h -- the divergent branches of the droplet update (FP64 form):
// wall films, TAB breakup, reflective walls with the Ahamed splash model,
// and the kinematics of the secondary droplets.
//
// Follows, in operation order (every expression keeps the reference's
// association so the FP64 kernel reproduces the CPU restatement bit for bit):
//   Source/PP_Spray/BreakupSplash/WallFilm.H:9-196      film volume, film source
//   Source/PP_Spray/BreakupSplash/TABBreakup.H:12-225   TAB oscillator, split
//   Source/PP_Spray/BreakupSplash/AhamedSplash.H:22-339 splash regimes
//   Source/PP_Spray/WallFunctions.H:89-234              wall check / reflection
//   Source/PP_Spray/SpraySB.cpp:9-207                   secondary droplets
// Cartesian walls only.  A breakup child's azimuth comes from a
// counter-based uniform number keyed by (seed, parent id, child index).
#pragma once

#include "droplet_physics.h"
#include "eb.h"

namespace pelemp {

// splash_breakup (SprayFuelData.H:31-39)
constexpr int kSBNone = 0;
constexpr int kSBKH = 1;
constexpr int kSBTab = 3;
constexpr int kSBSplashDry = 4;
constexpr int kSBSplashThermal = 5;
constexpr int kSBSplashWet = 6;
// splash_type (AhamedSplash.H:12-19)
constexpr int kRebound = 0, kDeposit = 1, kSplashDry = 2, kThermalBreakup = 3, kSpread = 4, kSplashWet = 5;

// KH-RT model constants (SpraySetup.cpp:17-19), the mean injected parcel
// mass N d^3 of the jets (SprayParticles.cpp:230-239) and the model's time
// step dtmod * flow_dt (SprayParticles.cpp:407-410)
struct KhrtParams {
  double B0, B1, C3, avg_inject_mass, dt;
};

// Liquid properties not needed by the drag/evaporation path
// (SprayFuelData.H:155-173).
struct LiquidExtra {
  double lambda_coef[4 * PELEMP_MAX_FUEL];
  double mu_coef[4 * PELEMP_MAX_FUEL];
  double sigma, wall_T, theta_c;
};

PM_HDI double lambdaL(const LiquidExtra& x, double T, int spf) {
  const double a = x.lambda_coef[4 * spf], b = x.lambda_coef[4 * spf + 1];
  const double c = x.lambda_coef[4 * spf + 2], d = x.lambda_coef[4 * spf + 3];
  return a + T * (b + T * (c + T * d));
}
PM_HDI double muL(const LiquidExtra& x, double T, int spf) {
  const double a = x.mu_coef[4 * spf], b = x.mu_coef[4 * spf + 1];
  const double c = x.mu_coef[4 * spf + 2], d = x.mu_coef[4 * spf + 3];
  return a + ((d / T + c) / T + b) / T;
}

// What the secondary-droplet pass needs of a parent (SBData.H:7-19).
template <int NF>
struct SBRec {
  int type;
  double norm[3], vel[3], loc[3];
  double T0, ref_dia, num_dens, phi1, phi2, phi3;
  double Y0[NF];
};

// Droplet state seen by the branches (positions absolute, FP64).
template <int NF>
struct BDrop {
  double pos[3];
  double vel[3];
  double T, dia, Y[NF], numDens, N0, bm1, bm2, film;
  bool alive;
};

// Periodic shift by a whole number of periods L into [lo, hi): a wall film
// placed at the impact point of a grazing droplet can lie several periods
// away (AhamedSplash.H:238); within one period it is a single +-L.
PM_HDI double wrap_periodic(double x, double lo, double hi) {
  if (x >= hi || x < lo) {
    const double L = hi - lo;
    double n = floor((x - lo) / L);
    if (x >= hi && n < 1.) n = 1.;
    if (x < lo && n > -1.) n = -1.;
    x -= n * L;
  }
  return x;
}

PM_HDI double v3len(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
PM_HDI double v3dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
PM_HDI void v3cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------------------
// Wall film: film volume per face area (fillFilmFab, WallFilm.H:9-26)
PM_HDI double film_fab_value(double film_hght, double film_dia, double face_area) {
  const double film_vol = kPIc * film_hght * (0.75 * film_dia * film_dia + film_hght * film_hght);
  return film_vol / face_area;
}

// calculateFilmSource (WallFilm.H:28-196): updates the film droplet and
// returns its sources (not scaled by the number density).
template <class M, int NF>
PM_HDI void film_source(double flow_dt, const GasAt<double, M::NS>& g, const FuelConst<double, NF>& fd,
                        const LiquidExtra& lx, BDrop<NF>& d, const double* cBoilT, DropSrc<double, NF>& s) {
  constexpr int NS = M::NS;
  const double dt = flow_dt * fd.dtmod;
  const double rule = 1. / 2.;
  const double C_eps = 1.E-15;
  const double min_height = 1.E-7;  // 1e-5 cm (WallFilm.H:46)
  double Y_skin[NS], h_film[NS], cp_n[NS], Ddiag[NS];
  double L_fuel[NF], mi_dot[NF], Y_film[NF], X_vapor[NF];
#pragma unroll
  for (int n = 0; n < NS; ++n) Ddiag[n] = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    L_fuel[spf] = 0.;
    mi_dot[spf] = 0.;
    X_vapor[spf] = 0.;
  }
  s.mom[0] = s.mom[1] = s.mom[2] = 0.;
  s.mass = 0.;
  s.eng = 0.;
  double T_film = d.T;
  double rho_film = 0., Tcrit = 0., cp_film = 0., lambda_film = 0., mw_film = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = rmin(T_film, cBoilT[spf]);
    Y_film[spf] = d.Y[spf];
    rho_film += Y_film[spf] / rhoL<double, NF>(fd, minT, spf);
    cp_film += Y_film[spf] * fd.cp[spf];
    lambda_film += Y_film[spf] * lambdaL(lx, minT, spf);
    Tcrit += Y_film[spf] * fd.critT[spf];
    mw_film += Y_film[spf] / fd.mw_fuel[spf];
  }
  mw_film = 1. / mw_film;
  rho_film = 1. / rho_film;
  T_film = rmin(0.999 * Tcrit, T_film);
  double film_height = d.film;
  const double film_dia = d.dia;
  const double rad2 = (0.5 * film_dia) * (0.5 * film_dia);
  const double film_area = kPIc * rad2;
  double film_mass = rho_film * film_height * film_area;
  const double start_mass = film_mass;
  const double delT = rmax(g.T - T_film, 0.);
  const double T_skin = T_film + rule * delT;
  GasProps<M>::T2Cpi(T_skin, cp_n);
  GasProps<M>::T2Hi(T_film, h_film);
#pragma unroll
  for (int n = 0; n < NS; ++n) Y_skin[n] = 0.;
  double cp_skin = 0., mw_skin = 0., B_M = 0., sumXVap = 0.;
  calc_vapor_state<M, NF, double>(fd, g, rule, T_film, C_eps, mw_film, Y_film, h_film, cp_n, cBoilT, Y_skin, X_vapor,
                                  L_fuel, B_M, sumXVap, cp_skin, mw_skin);
  double lambda_skin = 0., mu_skin = 0.;
  const double rho_skin = mw_skin * g.p / (kRU * T_skin);
  // only the fuels' diffusivities are used (WallFilm.H:148-152)
  GasProps<M>::template transport<double, NS - NF>(false, true, true, T_skin, rho_skin, Y_skin, Ddiag, mu_skin,
                                                   lambda_skin);
  if (sumXVap > 0.) {
    const double h_heat = lambda_skin / film_height;
    const double q_conv = film_area * h_heat * (g.T - T_film);
    const double q_cond = -film_area * lambda_film * 2. * (T_film - lx.wall_T) / film_height;
    const double h_mass = film_area * h_heat * rm::cbrt(lambda_skin / cp_skin) * rm::log1p(B_M) / lambda_skin;
    double sumL = 0., m_dot = 0.;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      const int fspec = fuel_species<NS, NF>(spf);
      const double dd = Ddiag[fspec] * (mw_skin / fd.mw_fuel[spf]);
      if (X_vapor[spf] > 0.) {
        const double cur_rhoD = X_vapor[spf] * dd / sumXVap;
        mi_dot[spf] = -rmax(h_mass * rm::pow(cur_rhoD, 2. / 3.), 0.);
        m_dot += mi_dot[spf];
        sumL += mi_dot[spf] * L_fuel[spf];
      }
    }
    const double part_temp_src = (q_conv + q_cond + sumL) / (film_mass * cp_film);
    s.eng += q_conv;
    T_film += dt * part_temp_src;
    double new_mass = film_mass + dt * m_dot;
    film_height = new_mass / (rho_film * film_area);
    if (film_height > min_height) {
      if (NF > 1) {
        rho_film = 0.;
#pragma unroll
        for (int spf = 0; spf < NF; ++spf) {
          Y_film[spf] = rmin(1., rmax((Y_film[spf] * film_mass + mi_dot[spf] * dt) / new_mass, 0.));
          if (Y_film[spf] < 1.E-12) Y_film[spf] = 0.;
          rho_film += Y_film[spf] / rhoL<double, NF>(fd, T_film, spf);
        }
        rho_film = 1. / rho_film;
      } else {
        rho_film = rhoL<double, NF>(fd, T_film, 0);
      }
    } else {
      film_height = 0.;
      new_mass = 0.;
      d.alive = false;
    }
    film_mass = new_mass;
  }
  d.T = T_film;
  const double mdot_total = (film_mass - start_mass) / dt;
  s.mass = mdot_total;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const double oldY = d.Y[spf];
    const double newY = Y_film[spf];
    const double midot = (newY * film_mass - oldY * start_mass) / dt;
    s.Ydot[spf] = midot;
    d.Y[spf] = newY;
    s.eng += midot * dep_value<NS, NF, true>(fd, h_film, spf);
  }
  d.dia = film_dia;
  d.film = film_height;
}

// ---------------------------------------------------------------------------
// TAB breakup (TABBreakup.H:12-184)
PM_HDI double tab_cd(double Reyn) {
  if (Reyn > 1000.) return 0.424;
  if (Reyn > 1.) return 24. / Reyn * (1. + rm::cbrt(Reyn * Reyn) / 6.);
  if (Reyn > 0.) return 24. / Reyn;
  return 0.;
}

// updateBreakupTAB: of the gas state only the density rho_g and velocity
// vel_g at the droplet enter; d.dia, d.numDens, d.bm1 (y) and d.bm2 (dy/dt)
// are updated, the other fields read.
template <int NF>
PM_HDI double tab_update(double Reyn_d, double dt, const double* cBoilT, double rho_g, const double* vel_g,
                         const FuelConst<double, NF>& fd, const LiquidExtra& lx, BDrop<NF>& d) {
  const double C_k = 8.;
  const double C_d = 10.;
  const double C_b = 0.5;
  const double C_F = 1. / 3.;
  const double Wet = 80.;
  const double k2 = 2. / 9.;
  double num_dens = d.numDens;
  double rho_part = 0., mu_part = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = rmin(d.T, cBoilT[spf]);
    rho_part += d.Y[spf] / rhoL<double, NF>(fd, minT, spf);
    mu_part += d.Y[spf] * muL(lx, minT, spf);
  }
  rho_part = 1. / rho_part;
  double Utan_total = 0.;
  const double sigma = lx.sigma;
  double rad_part = 0.5 * d.dia;
  const double min_rad = 4. * rm::cbrt(kMinMass * 3. / (4. * kPIc * rho_part));
  if (rad_part < min_rad) return 0.;
  double dv[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) dv[q] = vel_g[q] - d.vel[q];
  const double dv2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
  const double dvl = std::sqrt(dv2);
  const double We_div_r = rho_g * dv2 / sigma;
  const double We_crit = C_k * C_b / C_F;
  double denom = rho_part * rad_part * rad_part;
  double td = 2. * denom / (C_d * mu_part);
  double omega2 = C_k * sigma / (denom * rad_part) - 1. / (td * td);
  const double tbconst = 2. * std::sqrt(3. * rho_part / rho_g) / dvl;
  if (omega2 <= 0.) {
    d.bm1 = 0.;
    d.bm2 = 0.;
    return 0.;
  }
  double omega = std::sqrt(omega2);
  double yn = d.bm1;
  double ydotn = d.bm2;
  double Reyn = Reyn_d;
  double C_D = tab_cd(Reyn);
  const double tb_estconst = std::sqrt(3. * rho_part / rho_g) / dvl;
  double tb_est = 0.1 * rad_part * tb_estconst;
  double subdt = rmin(dt, tb_est);
  double curt = 0.;
  // exp(-subdt/td), cos and sin(omega subdt) and 1/omega change only when a
  // breakup event or the last, shortened interval changes subdt, omega or
  // td: kept from the previous interval otherwise (the same values)
  double k_subdt = __builtin_nan(""), k_omega = 0., k_td = 0.;
  double ex = 0., cs = 0., sn = 0., inv_omega = 0.;
  while (curt < dt) {
    const double We = We_div_r * rad_part;
    const double Wer = We / We_crit;
    const double A2 = (yn - Wer) * (yn - Wer) + ydotn * ydotn / omega2;
    const double A = std::sqrt(A2);
    if (!(subdt == k_subdt && omega == k_omega && td == k_td)) {
      ex = rm::exp(-subdt / td);
      pm::sincos(omega * subdt, sn, cs);
      inv_omega = 1. / omega;
      k_subdt = subdt;
      k_omega = omega;
      k_td = td;
    }
    const double ynp = Wer + ex * ((yn - Wer) * cs + inv_omega * (ydotn + (yn - Wer) / td) * sn);
    const double ydotnp =
        (Wer - ynp) / td + omega * ex * (inv_omega * (ydotn + (yn - Wer) / td) * cs - (yn - Wer) * sn);
    if (Wer + A <= 1.) {
      yn = ynp;
      ydotn = ydotnp;
    } else {
      double tbv = 0.;
      if (std::fabs(yn) < 1.) {
        const double pv1 = (yn - Wer);
        const double pv2 = -ydotn / omega;
        double phi = rm::atan2(pv2, pv1);
        if (phi < 0.) phi = 2. * kPIc + phi;
        double theta = rm::acos((1. - Wer) / A);
        if (theta < phi) {
          if (2. * kPIc - theta >= phi) theta = -theta;
          theta += 2. * kPIc;
        }
        tbv = (theta - phi) / omega;
      }
      if (tbv < subdt) {
        const double k1 = k2 * ((std::sqrt(Wet) - 1.) * rm::pow(We / Wet, 4.) + 1.);
        double Kbr = k1 * omega;
        if (We >= Wet) Kbr = k2 * omega * std::sqrt(We);
        const double etheta = rmax(-1., rmin(1., 1. - 1. / Wer));
        const double etb = rm::acos(etheta) / omega;
        const double rchild = rad_part * rm::exp(-Kbr * etb);
        num_dens *= rm::pow(rad_part / rchild, 3.);
        const double rsmr = std::sqrt(rm::pow(rad_part, 3.) / (rchild * rchild));
        const double AE2 = 3. * (1. - rad_part / rsmr + 5. * C_D * We / 72.) * omega2;
        const double AE = std::sqrt(AE2);
        Utan_total += AE * C_b * rad_part;
        Reyn = Reyn * rchild / rad_part;
        C_D = tab_cd(Reyn);
        rad_part = rchild;
        yn = 0.;
        ydotn = 0.;
        tb_est = rad_part * tbconst;
        subdt = rmin(dt, 0.1 * tb_est);
        denom = rho_part * rad_part * rad_part;
        td = 2. * denom / (C_d * mu_part);
        omega2 = C_k * sigma / (denom * rad_part) - 1. / (td * td);
        if (omega2 <= 0.) {
          d.dia = 2. * rad_part;
          d.numDens = num_dens;
          d.bm1 = 0.;
          d.bm2 = 0.;
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
  d.dia = 2. * rad_part;
  d.numDens = num_dens;
  d.bm1 = yn;
  d.bm2 = ydotn;
  return Utan_total;
}

// splitDropletTAB (TABBreakup.H:186-225)
template <int NF>
PM_HDI void tab_split(BDrop<NF>& d, double max_num_ppp, double Utan_total, SBRec<NF>& rf) {
  if (!d.alive) return;
  if (d.numDens > max_num_ppp) {
    rf.type = kSBTab;
    const double velMag = v3len(d.vel);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      rf.loc[q] = d.pos[q];
      rf.vel[q] = d.vel[q];
      rf.norm[q] = d.vel[q] / velMag;
    }
    rf.num_dens = d.numDens;
    rf.ref_dia = d.dia;
    rf.phi1 = Utan_total;
    rf.phi2 = d.bm1;
    rf.phi3 = d.bm2;
    rf.T0 = d.T;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = d.Y[spf];
    d.alive = false;
  }
}

// ---------------------------------------------------------------------------
// KH-RT wave breakup (ReitzKHRT.H:7-178): BM1 is the shed mass (N0 r^3
// units), BM2 the RT clock (negative: RT off until the first KH event)
template <int NS, int NF>
PM_HDI void khrt_update(double Reyn_d, const double* cBoilT, const GasAt<double, NS>& g,
                        const FuelConst<double, NF>& fd, const LiquidExtra& lx, const KhrtParams& kp,
                        bool make_new_drops, BDrop<NF>& d, SBRec<NF>& rf) {
  const double mthresh_jet = 0.03;
  const double mthresh_drop = 0.9;
  const double We_crit = 6.;
  const double A1 = 0.188;
  const double dt = kp.dt;
  double num_dens = d.numDens;
  double rho_part = 0., mu_part = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = rmin(d.T, cBoilT[spf]);
    rho_part += d.Y[spf] / rhoL<double, NF>(fd, minT, spf);
    mu_part += d.Y[spf] * muL(lx, minT, spf);
  }
  rho_part = 1. / rho_part;
  const double min_rad = 4. * rm::cbrt(kMinMass * 3. / (4. * kPIc * rho_part));
  double shed_mass = d.bm1;
  double N0 = d.N0;
  double rt_time = d.bm2;
  const double sigma = lx.sigma;
  double rad_part = 0.5 * d.dia;
  if (rad_part < min_rad) return;
  const double rad_part3 = rm::pow(rad_part, 3.);
  double dv[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) dv[q] = g.vel[q] - d.vel[q];
  const double vel_mag = v3len(dv);
  const double vel_sq = vel_mag * vel_mag;
  const double C_D = tab_cd(Reyn_d);
  const double We_g = g.rho * vel_sq * rad_part / sigma;
  const double Z = mu_part / std::sqrt(rho_part * rad_part * sigma);
  const double T = Z * std::sqrt(We_g);
  bool breakupRT = false;
  if (rt_time >= 0.) {
    const double a = 3. / 8. * C_D * g.rho * vel_sq / (rho_part * rad_part);
    const double momdiff = a * (rho_part - g.rho);
    const double omega_RT =
        std::sqrt(2. / (3. * std::sqrt(3. * sigma)) * momdiff * std::sqrt(momdiff) / (rho_part + g.rho));
    const double KRT = std::sqrt(momdiff / (3. * sigma));
    const double lam_RT = kp.C3 * kPIc / KRT;
    if (rad_part < lam_RT) {
      rt_time = 0.;
    } else {
      rt_time += dt;
    }
    const double tau_RT = 1. / omega_RT;
    if (rt_time > tau_RT && lam_RT < rad_part) {
      rt_time = 0.;
      num_dens *= rm::pow(rad_part / lam_RT, 3.);
      rad_part = lam_RT;
      breakupRT = true;
    }
  }
  if (!breakupRT) {
    const double lam_KH = rad_part * 9.02 * (1. + 0.45 * std::sqrt(Z)) * (1. + 0.4 * rm::pow(T, 0.7)) /
                          rm::pow(1. + 0.865 * rm::pow(We_g, 1.67), 0.6);
    const double omega_KH = (0.34 + 0.385 * We_g * std::sqrt(We_g)) / ((1. + Z) * (1. + 1.4 * rm::pow(T, 0.6))) *
                            std::sqrt(sigma / (rho_part * rad_part3));
    const double tau_KH = 3.726 * kp.B1 * rad_part / (lam_KH * omega_KH);
    double rs = kp.B0 * lam_KH;
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
        const double rel_shed_jet = 8. * shed_mass / kp.avg_inject_mass;
        if (rel_shed_drop > mthresh_drop) {
          create_child_drops = true;
          shed_mass = num_dens * rad_part3;
          rs = rmax(rs, min_rad);
          rs3 = rs * rs * rs;
          ns = shed_mass / rs3;
          d.alive = false;
        } else if (rel_shed_jet > mthresh_jet && rs > min_rad && ns > num_dens) {
          create_child_drops = true;
          np = (np * rp3 - ns * rs3) / rp3;
          N0 = np;
          shed_mass = 0.;
        }
      }
      if (create_child_drops) {
        rf.type = kSBKH;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          rf.loc[q] = d.pos[q];
          rf.vel[q] = d.vel[q];
          rf.norm[q] = d.vel[q] / vel_mag;  // droplet velocity over the relative speed, as written
        }
        rf.num_dens = ns;
        rf.ref_dia = 2. * rs;
        rf.T0 = d.T;
        rf.phi1 = A1 * omega_KH * lam_KH;
        rf.phi2 = 0.;
        rf.phi3 = 0.;
#pragma unroll
        for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = d.Y[spf];
      }
      rad_part = rp;
      num_dens = np;
    }
  }
  d.dia = 2. * rad_part;
  d.numDens = num_dens;
  d.N0 = N0;
  d.bm1 = shed_mass;
  d.bm2 = rt_time;
}

// ---------------------------------------------------------------------------
// Splash (AhamedSplash.H) and walls (WallFunctions.H)

// find_tangents (AhamedSplash.H:22-44)
PM_HDI void find_tangents(const double* pvel, double* tanPsi, const double* norm, double* tanBeta) {
  double testvec[3] = {-pvel[0], -pvel[1], -pvel[2]};
  double cr[3];
  v3cross(testvec, norm, cr);
  if (v3len(cr) < 1.E-5) {
    testvec[0] += 1.;
    testvec[1] += 1.;
    testvec[2] += 1.;
  }
  v3cross(testvec, norm, tanPsi);
  const double l1 = v3len(tanPsi);
#pragma unroll
  for (int q = 0; q < 3; ++q) tanPsi[q] /= l1;
  v3cross(tanPsi, norm, tanBeta);
  const double l2 = v3len(tanBeta);
#pragma unroll
  for (int q = 0; q < 3; ++q) tanBeta[q] /= l2;
}

// splash_criteria (AhamedSplash.H:125-197)
PM_HDI int splash_criteria(double Kv, double Tstar, double alpha, double Kcrit, double del_film, int& N_SB,
                           double& ms, double& dm) {
  int crit = kThermalBreakup;
  if (del_film < 0.01) {
    if (Tstar < 1.1) {
      if (Kv < Kcrit) {
        crit = kDeposit;
        ms = 0.;
        dm = 1.;
      } else {
        crit = kSplashDry;
        ms = 0.75 * (1. - rm::exp(-3.E-3 * (Kv - Kcrit)));
        dm = 1. - ms;
        N_SB = kSBSplashDry;
      }
    } else {
      if (Kv < Kcrit) {
        crit = kRebound;
      } else {
        crit = kThermalBreakup;
        ms = 0.75 * (1. - rm::exp(-0.0175 * (Kv - Kcrit)));
        dm = 0.;
        N_SB = kSBSplashThermal;
      }
    }
  } else {
    const double Kcrit_wet = Kcrit + 226. * rm::pow(del_film, 0.72);
    if (Kv > Kcrit_wet) {
      crit = kSplashWet;
      double A = 4.5, B = 0.2;
      if (del_film < 0.5) {
        A = 10.42 * del_film + 0.32;
        B = 8.75 * del_film - 0.107;
      }
      const double ms0 = A * (1. - rm::exp(-0.003 * (Kv - Kcrit_wet))) + B;
      const double alpha_d = alpha * 180. / kPIc;
      if (del_film > 0.1) {
        ms = (alpha_d > 45.) ? ms0 * (-0.0305 * alpha_d + 3.7496) : ms0 * 0.0525 * alpha_d;
      } else {
        ms = (alpha_d > 45.) ? ms0 * (-7.2E-3 * alpha_d + 1.6389)
                             : ms0 * (-4.6E-4 * alpha_d * alpha_d + 0.05 * alpha_d);
      }
      ms = rmin(ms, 1.);
    } else {
      crit = kSpread;
      ms = 0.;
      dm = 1.;
    }
  }
  return crit;
}

// droplet_splashing (AhamedSplash.H:199-339), pressure in Pa
template <int NF>
PM_HDI void droplet_splashing(BDrop<NF>& d, const FuelConst<double, NF>& fd, const LiquidExtra& lx, double p_gas,
                              const double* vel_part, double par_dot, const double* normal, const double* cBoilT,
                              SBRec<NF>& rf, double film_h) {
  const double U0mag = v3len(vel_part);
  const double U0norm = v3dot(normal, vel_part);
  const double U0tan = std::sqrt(U0mag * U0mag - U0norm * U0norm);
  const double T_part = d.T;
  const double num_dens = d.numDens;
  double mu_part = 0., rho_part = 0., Tboil = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const double minT = rmin(T_part, cBoilT[spf]);
    mu_part += d.Y[spf] * muL(lx, minT, spf);
    rho_part += d.Y[spf] / rhoL<double, NF>(fd, minT, spf);
    Tboil += d.Y[spf] * cBoilT[spf];
  }
  rho_part = 1. / rho_part;
  const double Tstar = lx.wall_T / Tboil;
  const double dia_part = d.dia;
  const double pmass = kPIc / 6. * rho_part * rm::pow(dia_part, 3.);
  const double We = rho_part * dia_part * U0norm * U0norm / lx.sigma;
  const double Re_L = std::fabs(U0norm) * dia_part * rho_part / mu_part;
  const double Kv = std::sqrt(We * std::sqrt(Re_L));
  const double alpha = rmax(kPIc / 6., rm::asin(std::fabs(U0norm) / U0mag));
  const double p_bar = p_gas * 1.E-5;
  const double Kcrit = (26.788 * rm::log(p_bar) + 33.546) * rm::exp(-6.E-3 * (lx.wall_T - 260.)) + 36.;
  const double del_film = film_h / dia_part;
  double ms = 0., dm = 0.;
  double impact_loc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) impact_loc[q] = d.pos[q] - par_dot * vel_part[q] / U0norm;
  int nsb = rf.type;
  const int splash_flag = splash_criteria(Kv, Tstar, alpha, Kcrit, del_film, nsb, ms, dm);
  rf.type = nsb;
  if (splash_flag == kRebound) {
    const double rbound = 0.823 * rm::exp(-1.9835E-3 * rm::pow(Kv, 1.6));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      d.pos[q] -= (1. + rbound) * par_dot * normal[q];
      d.vel[q] -= U0norm * (1. + rbound) * normal[q];
    }
  } else if (ms > 0.) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      rf.loc[q] = impact_loc[q];
      rf.norm[q] = normal[q];
      rf.vel[q] = vel_part[q];
    }
    rf.ref_dia = d.dia;
    rf.phi1 = Kv;
    rf.phi2 = pmass * ms;
    rf.phi3 = del_film;
    rf.T0 = T_part;
    rf.num_dens = num_dens;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) rf.Y0[spf] = d.Y[spf];
    if (splash_flag == kThermalBreakup) {
      const double rem_mass = pmass * (1. - ms);
      const double new_dia = rm::cbrt(6. * rem_mass / (kPIc * rho_part));
      double tanBeta[3], tanPsi[3];
      find_tangents(vel_part, tanPsi, normal, tanBeta);
      const double Un_splash = 0.0028 * Kv * rm::exp(-0.0062 * Kv) * std::fabs(U0mag);
      const double Ut_splash = 1.5 * std::sqrt(3.) * U0tan / kPIc;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        d.pos[q] = impact_loc[q];
        d.vel[q] = Ut_splash * tanBeta[q] + Un_splash * normal[q];
      }
      d.dia = new_dia;
      d.bm1 = 0.;
      d.bm2 = 0.;
    }
  }
  if (dm > 0.) {
    const double film_vol = num_dens * pmass * dm / rho_part;
    const double comb_dia = rm::cbrt(film_vol * 6. / kPIc);
    const double comb_We = We * comb_dia / dia_part;
    const double comb_Re = Re_L * comb_dia / dia_part;
    const double theta_c = lx.theta_c;
    const double st = rm::sin(theta_c);
    const double film_dia =
        comb_dia * std::sqrt((12. + comb_We) /
                             (6. * (1. - rm::cos(theta_c)) / (st * st) + 4. * comb_We / std::sqrt(comb_Re)));
    d.dia = film_dia;
    d.film = film_vol / (0.25 * kPIc * film_dia * film_dia);
    d.N0 = 1.;
    d.numDens = 1.;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      d.pos[q] = impact_loc[q];
      d.vel[q] = 0.;
    }
  }
}

// check_wall (WallFunctions.H:89-171), Cartesian faces; pos normalised by dx
PM_HDI bool check_wall(const double* pos, const int* bflags, const double* dx, const int* ijkc, double& par_dot,
                       double* normal, const EbDev* eb = nullptr, const int* ijkc_prev = nullptr) {
  bool wall_check = false;
  double bcentv[3];
#pragma unroll
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
  if (eb && !wall_check) wall_check = eb_wall(*eb, ijkc, ijkc_prev, normal, bcentv);
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
PM_HDI void impose_wall(bool do_splash, BDrop<NF>& d, const FuelConst<double, NF>& fd, const LiquidExtra& lx,
                        const double* plo, const double* dx, const int* bflags, const double* cBoilT, double p_fluid,
                        SBRec<NF>& rf, double film_h, const EbDev* eb = nullptr, const int* ijkc_prev = nullptr) {
  double par_dot = 2.;
  double normal[3];
  const double vel_part[3] = {d.vel[0], d.vel[1], d.vel[2]};
  double normpos[3];
  int ijkc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    normpos[q] = (d.pos[q] - plo[q]) / dx[q];
    ijkc[q] = static_cast<int>(floor(normpos[q]));
  }
  if (check_wall(normpos, bflags, dx, ijkc, par_dot, normal, eb, ijkc_prev)) {
    const double Nw_Vp = v3dot(normal, vel_part);
    if (do_splash) {
      droplet_splashing<NF>(d, fd, lx, p_fluid, vel_part, par_dot, normal, cBoilT, rf, film_h);
    } else {
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
        d.vel[dir] -= 2. * Nw_Vp * normal[dir];
        d.pos[dir] -= 2. * par_dot * normal[dir];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Secondary droplets (SpraySB.cpp:9-207)

PM_HDI uint64_t sb_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
PM_HDI double sb_uniform(uint64_t seed, int64_t parent_id, int child) {
  const uint64_t r = sb_mix64(seed ^ sb_mix64(static_cast<uint64_t>(parent_id) * 0x10000ull +
                                              static_cast<uint64_t>(child)));
  return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
}

// number of children of one record
template <int NF>
PM_HDI int sb_child_count(const SBRec<NF>& rf, double ppp_fact) {
  if (rf.type == kSBNone) return 0;
  if (rf.type >= kSBSplashDry) return 4;
  const double N_s = rm::pow(rf.num_dens, ppp_fact);
  const int N_d = static_cast<int>(rf.num_dens / N_s);
  return N_d > 1 ? N_d : 1;
}

// get_splash_vels (AhamedSplash.H:47-88)
PM_HDI void get_splash_vels(double U0norm, double U0tan, double Kv, double del_film, double& uBeta_0,
                            double& uBeta_half, double& uBeta_pi, double& uPsi_coeff, double& usNorm) {
  const double Un_mag = std::fabs(U0norm);
  if (del_film < 0.01) {
    const double Ut_splash = 0.0065 * Kv * rm::exp(-0.004 * Kv) * Un_mag;
    uBeta_0 = U0tan + Ut_splash;
    uBeta_half = rm::cos(kPIc / 6.) * U0tan;
    uBeta_pi = rm::cos(kPIc / 3.) * U0tan - Ut_splash;
    uPsi_coeff = 0.2 * U0tan + Ut_splash;
    usNorm = 0.0028 * Kv * rm::exp(-0.0062 * Kv) * Un_mag;
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

// get_ms_theta (AhamedSplash.H:90-122)
PM_HDI void get_ms_theta(double alpha, double ms, double del_film, double* ms_thetas) {
  if (del_film > 0.) {
    const double alpha_d = alpha * 180. / kPIc;
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
    const double tanterm = 0.4 * rm::tan(kPIc / 2. - alpha);
    ms_thetas[0] = 0.25 * ms;
    ms_thetas[1] = 0.25 * ms * (1. + tanterm);
    ms_thetas[2] = 0.25 * ms;
    ms_thetas[3] = 0.25 * ms * (1. - tanterm);
  }
}

// Child k of a record (SpraySB.cpp:57-199): fills the droplet state of the
// child; N0 and the TAB state are returned in c as well.
template <int NF>
PM_HDI void sb_child(const SBRec<NF>& rf, int k, int64_t parent_id, const FuelConst<double, NF>& fd,
                     double sub_dt, double ppp_fact, uint64_t seed, BDrop<NF>& c) {
  const double T0 = rf.T0;
  double Y0[NF];
  double rho_part = 0.;
  if (NF > 1) {
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      Y0[spf] = rf.Y0[spf];
      rho_part += Y0[spf] / rhoL<double, NF>(fd, T0, spf);
    }
    rho_part = 1. / rho_part;
  } else {
    rho_part = rhoL<double, NF>(fd, T0, 0);
    Y0[0] = 1.;
  }
  const double U0mag = v3len(rf.vel);
  c.alive = true;
  c.T = T0;
  c.film = 0.;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) c.Y[spf] = Y0[spf];
  if (rf.type >= kSBSplashDry) {
    double tanPsi[3], tanBeta[3];
    find_tangents(rf.vel, tanPsi, rf.norm, tanBeta);
    const double Kv = rf.phi1, ms = rf.phi2, del_film = rf.phi3;
    const double U0norm = v3dot(rf.norm, rf.vel);
    const double alpha = rmax(kPIc / 6., rm::asin(std::fabs(U0norm) / U0mag));
    const double U0tan = std::sqrt(U0mag * U0mag - U0norm * U0norm);
    double uBeta_0, uBeta_half, uBeta_pi, uPsi_coeff, usNorm;
    get_splash_vels(U0norm, U0tan, Kv, del_film, uBeta_0, uBeta_half, uBeta_pi, uPsi_coeff, usNorm);
    double ms_thetas[4];
    get_ms_theta(alpha, ms, del_film, ms_thetas);
    const double psi = 0.5 * kPIc * (static_cast<double>(k) - 1.);
    const double new_mass = (k == 0) ? ms_thetas[0] : (k == 1) ? ms_thetas[1] : (k == 2) ? ms_thetas[2] : ms_thetas[3];
    const double dia = rm::cbrt(6. * new_mass / (kPIc * rho_part));
    c.dia = dia;
    double utBeta = uBeta_half;
    if (k == 1) {
      utBeta = uBeta_0;
    } else if (k == 3) {
      utBeta = uBeta_pi;
    }
    const double utPsi = uPsi_coeff * rm::sin(psi);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      c.vel[q] = usNorm * rf.norm[q] + utBeta * tanBeta[q] + utPsi * tanPsi[q];
      c.pos[q] = rf.loc[q] + dia * rf.norm[q];
    }
    c.bm1 = 0.;
    c.bm2 = 0.;
    c.N0 = rf.num_dens;
    c.numDens = rf.num_dens;
  } else {
    const double Utan = rf.phi1;
    double N_s = rm::pow(rf.num_dens, ppp_fact);
    int N_d = static_cast<int>(rf.num_dens / N_s);
    N_d = N_d > 1 ? N_d : 1;
    N_s = rf.num_dens / static_cast<double>(N_d);
    double testvec[3] = {1., 0., 0.};
    double cr[3];
    v3cross(testvec, rf.norm, cr);
    if (v3len(cr) < 1.E-5) {
      testvec[0] = 0.;
      testvec[1] = 1.;
      testvec[2] = 0.;
    }
    double tanPsi[3], tanBeta[3];
    v3cross(testvec, rf.norm, tanPsi);
    const double l1 = v3len(tanPsi);
#pragma unroll
    for (int q = 0; q < 3; ++q) tanPsi[q] /= l1;
    v3cross(tanPsi, rf.norm, tanBeta);
    const double l2 = v3len(tanBeta);
#pragma unroll
    for (int q = 0; q < 3; ++q) tanBeta[q] /= l2;
    const double rnd = sb_uniform(seed, parent_id, k);
    c.dia = rf.ref_dia;
    c.bm1 = rf.phi2;
    c.bm2 = rf.phi3;
    c.N0 = N_s;
    c.numDens = N_s;
    const double psi = rnd * 2. * kPIc;
    const double sp = rm::sin(psi), cp = rm::cos(psi);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double pvel = rf.vel[q] + Utan * (sp * tanPsi[q] + cp * tanBeta[q]);
      c.pos[q] = rf.loc[q] + sub_dt * pvel;
      c.vel[q] = pvel;
    }
  }
}

}  // namespace pelemp
