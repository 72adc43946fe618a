// droplet_physics.h -- per-droplet drag / evaporation update for the HIP
// kernels, templated on the arithmetic type R (double or float).
//
// Follows the reference method of Source/PP_Spray/Drag.H:4-441 (drag +
// Abramzon-Sirignano evaporation with adaptive substeps) and the SprayData
// fits of SprayFuelData.H:146-211, PeleLMeX flavour (SI units, velocity and
// temperature read directly from the state, no kinetic-energy terms).
//
// For R = double every expression keeps the reference's association order
// and uses the pm:: elementary functions, so a kernel compiled with
// -ffp-contract=off reproduces the FP64 method bit for bit.  Droplet-
// independent subexpressions (Watson-corrected latent heat at boiling times
// the fuel molecular weight, 1/boilT, fuel molecular weights) are hoisted
// into FuelConst by the host; hoisting a whole subexpression does not change
// any rounding.
//
// Species indices of the fuels are compile-time (fuel_species below), so
// the per-species register arrays are never indexed dynamically.
//
// For R = float the same sequence runs in single precision with hardware
// elementary functions; the caller keeps the source sums in double.
#pragma once

#include "gas_props.h"
#include "pelemp_spray.h"

namespace pelemp {

constexpr double kPIc = 3.14159265358979323846;
constexpr double kMinMass = 1.0e-16;  // SprayUnits::min_mass (kg)

// Saturating float->int for the substep count (the reference's
// static_cast<int> is undefined beyond INT_MAX, Drag.H:290,355).
template <class R>
PM_HD int sat_int(R x) {
  if (!(x == x)) return 0;
  if (x > R(1.0e9)) return 1000000000;
  if (x < R(-1.0e9)) return -1000000000;
  return static_cast<int>(x);
}

template <class R>
PM_HD R rmin(R a, R b) {
  return (b < a) ? b : a;  // std::min semantics
}
template <class R>
PM_HD R rmax(R a, R b) {
  return (a < b) ? b : a;  // std::max semantics
}

// Gas species index of liquid fuel spf.  Every compiled mechanism lists its
// NF fuel species last, which the context checks at creation; the index is
// then a compile-time constant and the per-species arrays stay in registers.
template <int NS, int NF>
PM_HD constexpr int fuel_species(int spf) {
  return NS - NF + spf;
}

// v[k] of a per-species register array at a run-time species index k, as a
// chain of selects so that the array stays in registers.  Each element goes
// through an empty asm statement: without it the compiler folds the selects
// of array elements into one load at a computed address, which moves the
// array (and everything it shares a stack slot with) to scratch memory.
template <int NS, class R>
PM_HDI R species_at(const R* v, int k) {
  R r = v[0];
#pragma unroll
  for (int n = 1; n < NS; ++n) {
    R vn = v[n];
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(vn));
#endif
    r = (k == n) ? vn : r;
  }
  return r;
}

// Fuel data hoisted per launch (droplet independent).
template <class R, int NF>
struct FuelConst {
  R mw_fuel[NF];   // molecular weight of fuel species indx[spf]
  R hboil_mw[NF];  // Hboil_ref * mw_fuel (SprayFuelData.H:190-196)
  R inv_boilT[NF];  // 1 / boilT
  R critT[NF];
  R boilT[NF];
  R cp[NF];
  R latent[NF];
  R rho_coef[4 * NF];
  R psat_coef[4 * NF];
  int indx[NF];
  int dep_indx[NF];  // gas species that receives the vapour of fuel spf (dep_fuel_species)
  bool dep_own;      // every fuel's vapour goes to its own species (dep_indx == indx)
  bool cc_psat[NF];  // psat_coef[4*spf+3] == 0: Clausius-Clapeyron
  R ref_T;
  R ln10;  // pm::log(10) (host and device give the same bits)
  R dtmod;
  R body_force[3];
  R y_tab_coef;  // unused placeholder for alignment
  bool mass_trans, mom_trans, fixed_parts;
  int do_breakup;
  bool pelec;  // PeleC flavour: drag work and kinetic energy in the energy source
};

template <class R, int NF>
inline FuelConst<R, NF> make_fuel_const(const pelemp_spray_data& f, const double* mw_all) {
  FuelConst<R, NF> c{};
  for (int spf = 0; spf < NF; ++spf) {
    const double mw = mw_all[f.indx[spf]];
    const double Hboil_ref =
        f.ref_latent[spf] * pm::pow((f.critT[spf] - f.ref_T) / (f.critT[spf] - f.boilT[spf]), -0.38);
    c.mw_fuel[spf] = static_cast<R>(mw);
    c.hboil_mw[spf] = static_cast<R>(Hboil_ref * mw);
    c.inv_boilT[spf] = static_cast<R>(1. / f.boilT[spf]);
    c.critT[spf] = static_cast<R>(f.critT[spf]);
    c.boilT[spf] = static_cast<R>(f.boilT[spf]);
    c.cp[spf] = static_cast<R>(f.cp[spf]);
    c.latent[spf] = static_cast<R>(f.latent[spf]);
    for (int q = 0; q < 4; ++q) {
      c.rho_coef[4 * spf + q] = static_cast<R>(f.rho_coef[4 * spf + q]);
      c.psat_coef[4 * spf + q] = static_cast<R>(f.psat_coef[4 * spf + q]);
    }
    c.indx[spf] = f.indx[spf];
    c.dep_indx[spf] = f.dep_indx[spf];
    c.cc_psat[spf] = (f.psat_coef[4 * spf + 3] == 0.);
  }
  c.dep_own = true;
  for (int spf = 0; spf < NF; ++spf) c.dep_own = c.dep_own && (f.dep_indx[spf] == f.indx[spf]);
  c.ref_T = static_cast<R>(f.ref_T);
  c.ln10 = static_cast<R>(pm::log(10.0));
  c.dtmod = static_cast<R>(f.dtmod);
  for (int d = 0; d < 3; ++d) c.body_force[d] = static_cast<R>(f.body_force[d]);
  c.mass_trans = f.mass_trans != 0;
  c.mom_trans = f.mom_trans != 0;
  c.fixed_parts = f.fixed_parts != 0;
  c.do_breakup = f.do_breakup;
  return c;
}

// Of the deposition species of fuel spf (dep_indx, Drag.H:76,112,426): the
// gas mass fraction and a per-species value such as the enthalpy.  DEP:
// the deposition species may differ from the fuel's own (checked at run
// time); kernels compiled without it assume dep_indx == indx.
template <int NS, int NF, bool DEP, class R>
PM_HDI R dep_value(const FuelConst<R, NF>& fd, const R* v, int spf) {
  if constexpr (!DEP) return v[fuel_species<NS, NF>(spf)];
  return fd.dep_own ? v[fuel_species<NS, NF>(spf)] : species_at<NS>(v, fd.dep_indx[spf]);
}

// Interpolated gas state at the droplet (GasPhaseVals, SprayFuelData.H:71-112)
template <class R, int NS>
struct GasAt {
  R vel[3];
  R T;
  R rho;
  R p;
  R Y[NS];
  R mw_mix;
};

// Droplet sources (GasPhaseVals::fluid_*), before the -1/V sub_dt/flow_dt factor
template <class R, int NF>
struct DropSrc {
  R mom[3];
  R mass;
  R Ydot[NF];
  R eng;
};

// Droplet state the update reads and writes.
template <class R, int NF>
struct Drop {
  R vel[3];
  R T;
  R dia;
  R Y[NF];
  R numDens;
  R bm1;  // TAB distortion y (used only when do_breakup == 1)
  bool alive;
  // mass bookkeeping form only: droplet mass on entry and the mass change of
  // each liquid species over the call (kept separately so that changes far
  // below the precision of R are not lost in mass - startmass)
  R mass;
  R dmk[NF];
};

// SprayData::rhoL (SprayFuelData.H:146-153)
template <class R, int NF>
PM_HDI R rhoL(const FuelConst<R, NF>& f, R T, int spf) {
  const R a = f.rho_coef[4 * spf], b = f.rho_coef[4 * spf + 1];
  const R c = f.rho_coef[4 * spf + 2], d = f.rho_coef[4 * spf + 3];
  return a + T * (b + T * (c + T * d));
}

// SprayData::psat (SprayFuelData.H:204-211)
template <class R, int NF>
PM_HDI R psat(const FuelConst<R, NF>& f, R T, int spf) {
  const R a = f.psat_coef[4 * spf], b = f.psat_coef[4 * spf + 1];
  const R c = f.psat_coef[4 * spf + 2], d = f.psat_coef[4 * spf + 3];
  if constexpr (sizeof(R) == 8) {
    // pm::pow(10, z) is pm::exp(z pm::log(10)) for every z (z = 0 included)
    return d * pm::exp((a - b / (T + c)) * f.ln10);
  } else {
    return d * rm::pow(R(10.), a - b / (T + c));
  }
}

// pm::log1p(b) and pm::pow(1 + b, 0.7) from one logarithm of u = 1 + b.
// pm::pow(u, 0.7) is pm::exp(0.7 pm::log(u)) for every u (u = 0, 1, inf and
// NaN included), and pm::log1p(b) is pm::log(u) (b / (u - 1)), or b itself
// for u == 1, NaN or infinite b; both results are bit for bit those of the
// separate calls.  lu returns pm::log(u).
// (FP64 only; the template lets single-precision instantiations of the
// callers compile the branch they never take.)
template <class R>
PM_HDI void log1p_pow07(R b, R& l1p, R& p07, R& lu) {
  static_assert(sizeof(R) == 8 || sizeof(R) == 4, "");
  const double bd = static_cast<double>(b);
  const double u = 1.0 + bd;
  const double l = pm::log(u);
  lu = static_cast<R>(l);
  l1p = static_cast<R>((!(bd == bd) || bd == pm::kInf || u == 1.0) ? bd : l * (bd / (u - 1.0)));
  p07 = static_cast<R>(pm::exp(0.7 * l));
}

// GasPhaseVals::define (SprayFuelData.H:101-111): mixture MW and pressure
template <class M, class R>
PM_HDI void gas_define(GasAt<R, M::NS>& g) {
  R s = R(0);
#pragma unroll
  for (int n = 0; n < M::NS; ++n) s += g.Y[n] / static_cast<R>(M::W[n]);
  g.p = g.rho * static_cast<R>(kRU) * s * g.T;
  g.mw_mix = R(1) / s;
}

// SprayData::calcBoilT (SprayFuelData.H:180-200)
template <int NF, class R>
PM_HDI void calc_boil_t(const FuelConst<R, NF>& f, R p, R* cBoilT) {
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const R cb = R(1) / (rm::log(static_cast<R>(kPATM) / p) * static_cast<R>(kRU) / f.hboil_mw[spf] +
                         f.inv_boilT[spf]);
    cBoilT[spf] = rmin(f.critT[spf], cb);
  }
}

// calcHeatCoeff (Drag.H:8-45).  lBM1: log(1 + B_M), which the caller has
// already taken; pm::pow(1 + B_M, phi) is pm::exp(phi lBM1) for every phi
// when B_M > C_eps (1 + B_M > 1, so lBM1 > 0).
template <class R>
PM_HDI R calc_heat_coeff(R ratio, R B_M, R B_eps, R C_eps, R Nu_0, R lBM1) {
  if (B_M <= C_eps) return R(0);
  if constexpr (sizeof(R) == 4) {
    // single precision: (1+B_M)^phi = exp(phi ln(1+B_M)) with the logarithm
    // taken once, and (1+B_T)^0.7 = exp(0.7 log1p(B_T)) sharing log1p(B_T)
    const int maxIter = 100;
    const R NU2 = Nu_0 - R(2);
    const R lnBM1 = lBM1;
    R phi = ratio / Nu_0;
    R B_T_old = rm::exp(phi * lnBM1) - R(1);
    R logB = rm::log1p(B_T_old);
    R invFT = B_T_old / (logB * rm::exp(R(0.7) * logB));
    R Nu_num = R(2) + NU2 * invFT;
    phi = ratio / Nu_num;
    R B_T = rm::exp(phi * lnBM1) - R(1);
    R error = rm::fabs(B_T - B_T_old);
    int k = 0;
    while (k < maxIter && error > B_eps) {
      B_T_old = B_T;
      logB = rm::log1p(B_T);
      invFT = B_T / (logB * rm::exp(R(0.7) * logB));
      Nu_num = R(2) + NU2 * invFT;
      phi = ratio / Nu_num;
      B_T = rm::exp(phi * lnBM1) - R(1);
      error = rm::fabs(B_T - B_T_old);
      k++;
    }
    logB = rm::log1p(B_T);
    invFT = B_T / (logB * rm::exp(R(0.7) * logB));
    Nu_num = R(2) + NU2 * invFT;
    return Nu_num * logB / B_T;
  }
  // the reference's sequence (pow(1+B_M, phi), log1p(B_T), pow(1+B_T, 0.7)
  // per step) with log(1 + B_M) taken once and log(1 + B_T) shared by
  // log1p and the 0.7 power: the same bits with two logarithms fewer per step
  const int maxIter = 100;
  const R NU2 = Nu_0 - R(2);
  R phi = ratio / Nu_0;
  R B_T_old = pm::exp(phi * lBM1) - R(1);
  R logB, p07, lu;
  log1p_pow07(B_T_old, logB, p07, lu);
  R invFT = B_T_old / (logB * p07);
  R Nu_num = R(2) + NU2 * invFT;
  phi = ratio / Nu_num;
  R B_T = pm::exp(phi * lBM1) - R(1);
  R error = rm::fabs(B_T - B_T_old);
  int k = 0;
  while (k < maxIter && error > B_eps) {  // B_eps: see spray_source
    B_T_old = B_T;
    log1p_pow07(B_T, logB, p07, lu);
    invFT = B_T / (logB * p07);
    Nu_num = R(2) + NU2 * invFT;
    phi = ratio / Nu_num;
    B_T = pm::exp(phi * lBM1) - R(1);
    error = rm::fabs(B_T - B_T_old);
    k++;
  }
  log1p_pow07(B_T, logB, p07, lu);
  invFT = B_T / (logB * p07);
  Nu_num = R(2) + NU2 * invFT;
  return Nu_num * logB / B_T;
}

// calcVaporState (Drag.H:47-135)
template <class M, int NF, class R, bool DEP = true>
PM_HDI void calc_vapor_state(const FuelConst<R, NF>& fd, const GasAt<R, M::NS>& g, R rule, R T_in,
                            R C_eps, R mw_part, const R* Y_l, const R* h_part, const R* cp_n,
                            const R* cBoilT, R* Y_skin, R* X_vapor, R* L_fuel, R& B_M, R& sumXVap,
                            R& cp_skin, R& mw_skin, R* sumYVap_out = nullptr) {
  constexpr int NS = M::NS;
  R mbar_vap = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const int fspec = fuel_species<NS, NF>(spf);
    const R boilT_ref = fd.boilT[spf];
    const R mw_fuel = fd.mw_fuel[spf];
    const R T_part = rmin(T_in, R(0.99) * cBoilT[spf]);
    const R part_latent = h_part[fspec] + fd.latent[spf] - fd.cp[spf] * (T_part - fd.ref_T);
    L_fuel[spf] = part_latent;
    R pres_sat = R(0);
    if (fd.cc_psat[spf]) {
      pres_sat = static_cast<R>(kPATM) *
                 rm::exp(part_latent * mw_fuel / static_cast<R>(kRU) * (R(1) / boilT_ref - R(1) / T_part));
    } else {
      pres_sat = psat<R, NF>(fd, T_part, spf);
    }
    const R X_fluid = dep_value<NS, NF, DEP>(fd, g.Y, spf) * g.mw_mix / mw_fuel;
    if (pres_sat > X_fluid * g.p) {
      const R Xl = Y_l[spf] * mw_part / mw_fuel;
      const R Xv = Xl * pres_sat / g.p;
      mbar_vap += Xv * mw_fuel;
      sumXVap += Xv;
      X_vapor[spf] = Xv;
    } else {
      X_vapor[spf] = R(0);
    }
  }
  const R totalmwx = mbar_vap + (R(1) - sumXVap) * g.mw_mix;
  R sumYSkin = R(0), sumYfFluid = R(0), sumYVap = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    if (X_vapor[spf] > R(0)) {
      const int fspec = fuel_species<NS, NF>(spf);
      const R mw_fuel = fd.mw_fuel[spf];
      const R Yfv = X_vapor[spf] * mw_fuel / totalmwx;
      sumYVap += Yfv;
      const R gasY = dep_value<NS, NF, DEP>(fd, g.Y, spf);
      sumYfFluid += gasY;
      const R Ysk = Yfv + rule * (gasY - Yfv);
      Y_skin[fspec] = Ysk;
      sumYSkin += Ysk;
    }
  }
  if (sumYVap_out) *sumYVap_out = sumYVap;
  const R renorm = (R(1) - sumYSkin) / (R(1) - sumYfFluid);
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    if (Y_skin[n] == R(0)) Y_skin[n] = g.Y[n] * renorm;
    cp_skin += Y_skin[n] * cp_n[n];
    mw_skin += Y_skin[n] / static_cast<R>(M::W[n]);
  }
  mw_skin = R(1) / mw_skin;
  B_M = (sumYVap - sumYfFluid) / rmax(C_eps, (R(1) - sumYVap));
  B_M = rmin(R(20), B_M);
}

// Rates of one substep of calculateSpraySource (Drag.H:204-347): drag,
// evaporation and heat transfer of a droplet in the gas state g.  This is the
// plain method's "rate evaluation" (the part the learned surrogate replaces).
template <class M, int NF, class R>
struct SubRates {
  R Reyn;
  R drag_force;    // 0.125 rho_s C_D pi d^2 |du|  (force / |du| ... per droplet)
  R drag_force_p;  // drag acceleration per unit relative velocity [1/s]
  R diff_vel[3];
  R m_dot;       // total mass rate of the droplet (<= 0) [kg/s]
  R mi_dot[NF];  // per liquid species
  R sumL;        // sum_k mi_dot_k L_k [W]
  R conv_src;    // convective heat into the droplet [W]
  R part_temp_src;  // dT_d/dt [K/s]
  R cp_part, Tboil, delT, inv_pmass;
  R h_part[M::NS];  // gas enthalpies at T_d (used by the species energy term)
  // film properties and transfer groups (what the learned surrogate sees):
  R ShRhoD;  // Sh* x sum_k rho D_k            [kg/(m s)]   (mass transfer)
  R lamNu;   // lambda_skin x Nu_eff           [W/(m K)]    (heat transfer)
  R rhoD;    // sum_k rho D_k of the vapour    [kg/(m s)]
  R mu_skin;  // film viscosity                [Pa s]
  R sumYVap;  // fuel vapour mass fraction at the surface
  R B_M;      // Spalding mass transfer number
};

template <class M, int NF, class R, bool DEP = true>
PM_HDI void substep_rates(const GasAt<R, M::NS>& g, const FuelConst<R, NF>& fd, R T_part, R dia_part,
                          R pmass, R rho_part, const R* Y_part, const R* vel_part, R y_tab,
                          const R* cBoilT, SubRates<M, NF, R>& o) {
  constexpr int NS = M::NS;
  const R rule = R(1.) / R(3.);
  const R C_eps = R(1.E-15);
  // The fixed-point tolerance of calcHeatCoeff (1e-7, Drag.H:155) is below
  // single-precision resolution for B_T of order one; the single-precision
  // form uses 1e-6 (an absolute tolerance, as in the reference).
  const R B_eps = (sizeof(R) == 4) ? R(1.E-6) : R(1.E-7);
  const bool mass_trans = fd.mass_trans;
  R Y_skin[NS], cp_n[NS], Ddiag[NS];
  R L_fuel[NF], X_vapor[NF];
#pragma unroll
  for (int n = 0; n < NS; ++n) Ddiag[n] = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    L_fuel[spf] = R(0);
    X_vapor[spf] = R(0);
    o.mi_dot[spf] = R(0);
  }
  R cp_part = R(0), Tboil = R(0), mw_part = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    Tboil += Y_part[spf] * cBoilT[spf];
    cp_part += Y_part[spf] * fd.cp[spf];
    mw_part += Y_part[spf] / fd.mw_fuel[spf];
  }
  mw_part = R(1) / mw_part;
  const R delT = rmax(g.T - T_part, R(0));
  R T_skin = g.T;
  if (mass_trans) T_skin = T_part + rule * delT;
  GasProps<M>::T2Cpi(T_skin, cp_n);
  GasProps<M>::T2Hi(T_part, o.h_part);
#pragma unroll
  for (int n = 0; n < NS; ++n) Y_skin[n] = R(0);
  R cp_skin = R(0), mw_skin = R(0), B_M = R(0), sumXVap = R(0);
  o.sumYVap = R(0);
  o.ShRhoD = R(0);
  o.rhoD = R(0);
  if (mass_trans) {
    calc_vapor_state<M, NF, R, DEP>(fd, g, rule, T_part, C_eps, mw_part, Y_part, o.h_part, cp_n, cBoilT, Y_skin,
                               X_vapor, L_fuel, B_M, sumXVap, cp_skin, mw_skin, &o.sumYVap);
  } else {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      Y_skin[n] = g.Y[n];
      cp_skin += g.Y[n] * cp_n[n];
    }
    mw_skin = g.mw_mix;
  }
  R lambda_skin = R(0), mu_skin = R(0);
  R rho_skin = g.rho;
  if (mass_trans) rho_skin = mw_skin * g.p / (static_cast<R>(kRU) * T_skin);
  // only the fuels' diffusivities are used (Drag.H:312-318)
  GasProps<M>::template transport<R, NS - NF>(true, mass_trans, mass_trans, T_skin, rho_skin, Y_skin, Ddiag, mu_skin,
                                             lambda_skin);
#pragma unroll
  for (int q = 0; q < 3; ++q) o.diff_vel[q] = g.vel[q] - vel_part[q];
  const R diff_vel_mag = rm::sqrt(o.diff_vel[0] * o.diff_vel[0] + o.diff_vel[1] * o.diff_vel[1] +
                                  o.diff_vel[2] * o.diff_vel[2]);
  const R Reyn = rho_skin * diff_vel_mag * dia_part / mu_skin;
  const R inv_pmass = R(1) / pmass;
  R drag_coef = R(0);
  if (Reyn > R(1000.)) {
    drag_coef = R(0.424);
  } else if (Reyn > R(1.)) {
    drag_coef = R(24.) / Reyn * (R(1.) + rm::cbrt(Reyn * Reyn) / R(6.));
  } else if (Reyn > R(0.)) {
    drag_coef = R(24.) / Reyn;
  }
  drag_coef *= (R(1.) + R(2.6432) * y_tab);
  o.drag_force = R(0.125) * rho_skin * drag_coef * static_cast<R>(kPIc) * dia_part * dia_part * diff_vel_mag;
  o.drag_force_p = R(0.75) * rho_skin * drag_coef * diff_vel_mag / (dia_part * rho_part);
  R part_temp_src = R(0);
  R m_dot = R(0);
  R sumL = R(0);
  R conv_src = R(0);
  if (mass_trans) {
    const R Pr_skin = mu_skin * cp_skin / lambda_skin;
    const R powR = rmax(rm::pow(rmin(R(400.), Reyn), R(0.077)), R(1.));
    R Nu_num = R(0);
    const R Nu_0 = R(1.) + powR * rm::cbrt(R(1.) + Reyn * Pr_skin);
    if (sumXVap > R(0)) {
      R rhoDtotal = R(0), cp_fuel = R(0), sumYSkin = R(0);
      R Dfuel[NF];
#pragma unroll
      for (int spf = 0; spf < NF; ++spf) {
        const int fspec = fuel_species<NS, NF>(spf);
        R dd = Ddiag[fspec];
        dd *= mw_skin / fd.mw_fuel[spf];
        const R X_norm = X_vapor[spf] / sumXVap;
        dd *= X_norm;
        Ddiag[fspec] = dd;
        Dfuel[spf] = dd;
        rhoDtotal += dd;
        const R ysk = Y_skin[fspec];
        cp_fuel += ysk * cp_n[fspec];
        sumYSkin += ysk;
      }
      cp_fuel /= sumYSkin;
      const R Sc_skin = mu_skin / rhoDtotal;
      R logB, invFM, lBM1;
      if constexpr (sizeof(R) == 4) {
        logB = rm::log1p(B_M);
        invFM = B_M / (logB * rm::exp(R(0.7) * logB));  // (1+B_M)^0.7 from log1p(B_M)
        lBM1 = logB;
      } else {
        // log1p(B_M), (1 + B_M)^0.7 and log(1 + B_M) from one logarithm
        R p07;
        log1p_pow07(B_M, logB, p07, lBM1);
        invFM = B_M / (logB * p07);
      }
      const R Sh_0 = R(1.) + powR * rm::cbrt(R(1.) + Reyn * Sc_skin);
      const R Sh_num = R(2.) + (Sh_0 - R(2.)) * invFM;
      const R mdotcoeff = static_cast<R>(kPIc) * dia_part * Sh_num * logB;
      const R ratio = cp_fuel * Sh_num * rhoDtotal / lambda_skin;
      Nu_num = calc_heat_coeff<R>(ratio, B_M, B_eps, C_eps, Nu_0, lBM1);
      o.ShRhoD = Sh_num * rhoDtotal;
      o.rhoD = rhoDtotal;
#pragma unroll
      for (int spf = 0; spf < NF; ++spf) {
        if (X_vapor[spf] > R(0)) {
          // Dfuel[spf] is Ddiag[indx[spf]] after the updates above (fuel
          // species are distinct)
          o.mi_dot[spf] = -rmax(Dfuel[spf] * mdotcoeff, R(0.));
          m_dot += o.mi_dot[spf];
          sumL += o.mi_dot[spf] * L_fuel[spf];
        }
      }
    } else {
      Nu_num = Nu_0;
    }
    conv_src = static_cast<R>(kPIc) * lambda_skin * dia_part * delT * Nu_num;
    part_temp_src = (sumL + conv_src) * inv_pmass / cp_part;
    o.lamNu = lambda_skin * Nu_num;
  } else {
    o.lamNu = R(0);
  }
  o.B_M = B_M;
  o.mu_skin = mu_skin;
  o.Reyn = Reyn;
  o.m_dot = m_dot;
  o.sumL = sumL;
  o.conv_src = conv_src;
  o.part_temp_src = part_temp_src;
  o.cp_part = cp_part;
  o.Tboil = Tboil;
  o.delT = delT;
  o.inv_pmass = inv_pmass;
}

// calculateSpraySource (Drag.H:137-441), PELELM_USE_SPRAY flavour.
// Updates d (vel, T, dia, Y, alive) and returns the sources in s and the
// Reynolds number of the last substep.
//
// MASSB (mass bookkeeping form, used with R = float): the droplet mass comes
// in as d.mass, the per-species mass changes sum_substeps mi_dot part_dt are
// returned in d.dmk, and the mass/species sources are built from those sums
// instead of from (pmass - startmass); in exact arithmetic the two agree for
// the explicit Euler update of Drag.H:371-400.
//
// pre (surrogate form): rates of the first substep predicted by the learned
// surrogate; when given they replace substep_rates for that substep.
template <class M, int NF, class R, bool MASSB = false, bool DEP = true>
PM_HDI R spray_source(R flow_dt, const GasAt<R, M::NS>& g, const FuelConst<R, NF>& fd,
                     Drop<R, NF>& d, const R* cBoilT, DropSrc<R, NF>& s,
                     const SubRates<M, NF, R>* pre = nullptr) {
  constexpr int NS = M::NS;
  const R C_eps = R(1.E-15);
  const R min_mass = R(kMinMass);
  const int nSubMax = 100;
  const bool mass_trans = fd.mass_trans;
  R Y_part[NF];
  SubRates<M, NF, R> o;
  s.mom[0] = s.mom[1] = s.mom[2] = R(0);
  s.mass = R(0);
  s.eng = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) s.Ydot[spf] = R(0);
  const R num_ppp = d.numDens;
  R vel_part[3] = {d.vel[0], d.vel[1], d.vel[2]};
  R y_tab = R(0);
  if (fd.do_breakup == 1) y_tab = d.bm1;
  if (fd.fixed_parts) vel_part[0] = vel_part[1] = vel_part[2] = R(0);
  R T_part = d.T;
  R dia_part = d.dia;
  R rho_part = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    Y_part[spf] = d.Y[spf];
    rho_part += Y_part[spf] / rhoL<R, NF>(fd, rmin(T_part, cBoilT[spf]), spf);
  }
  rho_part = R(1) / rho_part;
  R dt = flow_dt;
  int isub = 1;
  int nsub = 1;
  R pmass;
  if constexpr (MASSB) {
    pmass = d.mass;
    dia_part = rm::cbrt(R(6.) * pmass / (static_cast<R>(kPIc) * rho_part));
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) d.dmk[spf] = R(0);
  } else {
    pmass = static_cast<R>(kPIc / 6.) * rho_part * (dia_part * dia_part * dia_part);
  }
  const R startmass = pmass;
  R Reyn = R(0);
  R part_vel_src[3];
  while (isub <= nsub) {
    if (pre != nullptr && isub == 1) {
      o = *pre;
    } else {
      substep_rates<M, NF, R, DEP>(g, fd, T_part, dia_part, pmass, rho_part, Y_part, vel_part, y_tab, cBoilT, o);
    }
    Reyn = o.Reyn;
    const R Tboil = o.Tboil;
    const R* mi_dot = o.mi_dot;
#pragma unroll
    for (int q = 0; q < 3; ++q) part_vel_src[q] = o.drag_force_p * o.diff_vel[q] + fd.body_force[q];
    if (fd.mom_trans) {
      const R fm = num_ppp * o.drag_force;
#pragma unroll
      for (int q = 0; q < 3; ++q) s.mom[q] += fm * o.diff_vel[q];
      if (fd.pelec) {
        // drag work s_d,mu . u_d (Drag.H:282-286, PeleC flavour)
        const R S_dmu_dot_u =
            o.diff_vel[0] * vel_part[0] + o.diff_vel[1] * vel_part[1] + o.diff_vel[2] * vel_part[2];
        s.eng += num_ppp * o.drag_force * S_dmu_dot_u;
      }
      if (isub == 1) {
        const R inv_tau_var = o.drag_force_p;
        const int c = sat_int(flow_dt * inv_tau_var) + 1;
        nsub = (c < nsub) ? nsub : c;
        nsub = (nSubMax < nsub) ? nSubMax : nsub;
      }
    }
    const R part_temp_src = o.part_temp_src;
    const R m_dot = o.m_dot;
    if (mass_trans) {
      s.eng += num_ppp * o.conv_src;
      if (isub == 1 && o.delT > C_eps) {
        const R inv_tau_d = -m_dot / (R(0.2) * pmass);
        const R inv_tau_T = o.conv_src * o.inv_pmass / (o.cp_part * o.delT);
        const int a = sat_int(flow_dt * inv_tau_d);
        const int b = sat_int(flow_dt * inv_tau_T);
        const int c = ((a < b) ? b : a) + 1;
        nsub = (c < nsub) ? nsub : c;
        nsub = (nSubMax < nsub) ? nSubMax : nsub;
      }
    }
    if (isub == 1) dt = flow_dt / static_cast<R>(nsub);
    const R part_dt = fd.dtmod * dt;
    if (!fd.fixed_parts) {
#pragma unroll
      for (int q = 0; q < 3; ++q) vel_part[q] += part_dt * part_vel_src[q];
    }
    if (mass_trans) {
      const R newT = T_part + part_dt * part_temp_src;
      const R new_mass = pmass + m_dot * part_dt;
      if (new_mass > min_mass && newT < Tboil) {
        T_part = newT;
        if constexpr (MASSB) {
#pragma unroll
          for (int spf = 0; spf < NF; ++spf) d.dmk[spf] += mi_dot[spf] * part_dt;
        }
        if (NF > 1) {
          rho_part = R(0);
          R sumY = R(0);
#pragma unroll
          for (int spf = 0; spf < NF; ++spf) {
            Y_part[spf] = rmin(R(1.), rmax(R(0.), (Y_part[spf] * pmass + mi_dot[spf] * part_dt) / new_mass));
            if (Y_part[spf] <= R(1.E-12)) Y_part[spf] = R(0.);
            sumY += Y_part[spf];
          }
#pragma unroll
          for (int spf = 0; spf < NF; ++spf) {
            Y_part[spf] /= sumY;
            rho_part += Y_part[spf] / rhoL<R, NF>(fd, rmin(T_part, cBoilT[spf]), spf);
          }
          rho_part = R(1.) / rho_part;
        } else {
          rho_part = rhoL<R, NF>(fd, rmin(T_part, cBoilT[0]), 0);
        }
        pmass = new_mass;
        dia_part = rm::cbrt(R(6.) * pmass / (static_cast<R>(kPIc) * rho_part));
      } else {
        pmass = R(0.);
        d.alive = false;
        if constexpr (MASSB) {
#pragma unroll
          for (int spf = 0; spf < NF; ++spf) d.dmk[spf] = -d.Y[spf] * startmass;
        }
        nsub = isub;
        isub = nsub + 1;
      }
    }
    ++isub;
  }
  if (nsub > 1) {
    s.eng /= static_cast<R>(nsub);
#pragma unroll
    for (int q = 0; q < 3; ++q) s.mom[q] /= static_cast<R>(nsub);
  }
  if constexpr (MASSB) {
    R dm = R(0);
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) dm += d.dmk[spf];
    const R inv_dt = R(1) / (fd.dtmod * flow_dt);
    const R mdot_total = dm * inv_dt;
    s.mass = num_ppp * mdot_total;
    if (fd.mom_trans) {
#pragma unroll
      for (int q = 0; q < 3; ++q) s.mom[q] += num_ppp * vel_part[q] * mdot_total;
    }
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      const R midot = d.dmk[spf] * inv_dt;
      s.Ydot[spf] = num_ppp * midot;
      s.eng += num_ppp * midot * dep_value<NS, NF, DEP>(fd, o.h_part, spf);
      d.Y[spf] = Y_part[spf];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) d.vel[q] = vel_part[q];
    d.T = T_part;
    d.dia = dia_part;
    return Reyn;
  }
  const R mdot_total = (pmass - startmass) / (fd.dtmod * flow_dt);
  s.mass = num_ppp * mdot_total;
  if (fd.pelec) {
    // kinetic energy of the evaporated mass (Drag.H:418-421, PeleC flavour)
    const R part_ke = R(0.5) * (vel_part[0] * vel_part[0] + vel_part[1] * vel_part[1] + vel_part[2] * vel_part[2]);
    s.eng += num_ppp * part_ke * mdot_total;
  }
  if (fd.mom_trans) {
#pragma unroll
    for (int q = 0; q < 3; ++q) s.mom[q] += num_ppp * vel_part[q] * mdot_total;
  }
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const R oldY = d.Y[spf];
    const R newY = Y_part[spf];
    const R midot = (newY * pmass - oldY * startmass) / (fd.dtmod * flow_dt);
    s.Ydot[spf] = num_ppp * midot;
    s.eng += num_ppp * midot * dep_value<NS, NF, DEP>(fd, o.h_part, spf);
    d.Y[spf] = Y_part[spf];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) d.vel[q] = vel_part[q];
  d.T = T_part;
  d.dia = dia_part;
  return Reyn;
}

}  // namespace pelemp
