// surrogate.h -- physics side of the learned surrogate of the substep rates.
//
// The plain substep (substep_rates in droplet_physics.h, after Drag.H:213-347)
// evaluates the film ("skin") state of the droplet: species heat capacities
// at T_skin, mixture-averaged viscosity, conductivity and diffusivities, and
// the fixed-point iteration for the heat-transfer Spalding number
// (calcHeatCoeff, Drag.H:4-45).  The network replaces exactly that block by
// three quantities:
//   z0 = ln(sum_k rho D_k)      film diffusivity of the vapour (Drag.H:305-318)
//   z1 = ln(lambda_skin Nu*)    heat-transfer group, including the
//                               fixed point for B_T (Drag.H:328-345)
//   z2 = ln(mu_skin)            film viscosity (Drag.H:252-262)
// Everything else keeps the plain path's closed forms, computed here:
// the surface vapour state (Raoult, saturation pressure, latent heats;
// Drag.H:47-135), the Spalding mass transfer number B_M, the film density,
// the Reynolds and Schmidt numbers, the piecewise drag coefficient (which
// jumps by 7/6 at Re = 1, Drag.H:266-274, where a smooth network could not
// follow it) and the Abramzon-Sirignano Sherwood number:
//   Re     = rho_skin |du| d / mu,  Sc = mu / rho D
//   drag_p = 0.75 rho_skin C_D(Re) |du| / (d rho_l)          (Drag.H:275-279)
//   Sh*    = 2 + (Sh_0(Re, Sc) - 2) B_M / (ln(1+B_M)(1+B_M)^0.7)
//   m_dot  = -pi d Sh* rho D ln(1 + B_M)                      (Drag.H:332-340)
//   Q      =  pi d (T_g - T_d)^+ lambda Nu*                   (Drag.H:343)
// The split of m_dot among the liquid species follows the surface vapour
// mole fractions weighted by the binary diffusivities of each fuel in N2
// (the mixture-averaged rho D_k of the plain path, with the air-dominated
// film approximated by N2).
#pragma once

#include "droplet_physics.h"
#include "pelemp_spray.h"

namespace pelemp {

constexpr int kSurrIn = PELEMP_SURR_NIN;
constexpr int kSurrHid = 64;
constexpr int kSurrOut = PELEMP_SURR_NOUT;
constexpr int kSurrLin = 5;  // linear layers
// the relative-velocity input is ln(|u_rel| + kSurrU0)
constexpr double kSurrU0 = 0.01;

PM_HD constexpr int surr_fan_in(int L) { return L == 0 ? kSurrIn : kSurrHid; }
PM_HD constexpr int surr_fan_out(int L) { return L == kSurrLin - 1 ? kSurrOut : kSurrHid; }
PM_HD constexpr int surr_w_off(int L) {
  int o = 0;
  for (int i = 0; i < L; ++i) o += surr_fan_in(i) * surr_fan_out(i) + surr_fan_out(i);
  return o;
}
PM_HD constexpr int surr_b_off(int L) { return surr_w_off(L) + surr_fan_in(L) * surr_fan_out(L); }
constexpr int kSurrParams = surr_w_off(kSurrLin);
static_assert(kSurrParams == PELEMP_SURR_NPARAM, "parameter count");

// Exact part of the surface state (calcVaporState, Drag.H:47-135, with the
// skin mixture reduced to its molecular weight).
template <class R, int NF>
struct VaporLite {
  R B_M;
  R Ysk_f;      // fuel mass fraction of the film gas (1/3 rule)
  R Yfg;        // fuel vapour mass fraction of the gas
  R mw_skin;    // molecular weight of the film gas
  R X_vapor[NF];
  R L_fuel[NF];
  R h_fuel[NF];  // gas enthalpy of each fuel at T_d
  bool ok;      // every liquid species present evaporates (X_vapor > 0)
};

template <class M, int NF, class R>
PM_HDI void vapor_lite(const FuelConst<R, NF>& fd, const GasAt<R, M::NS>& g, R T_in, R mw_part, const R* Y_l,
                       const R* cBoilT, VaporLite<R, NF>& v) {
  constexpr int NS = M::NS;
  const R rule = R(1.) / R(3.);
  const R C_eps = R(1.E-15);
  R mbar_vap = R(0), sumXVap = R(0);
  v.ok = true;
  v.Yfg = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const int fspec = fuel_species<NS, NF>(spf);
    // h_k(T_d) of the fuel species only (the plain path evaluates all NS)
    const R hk = GasProps<M>::template h_one<R, -1>(T_in, fspec);
    v.h_fuel[spf] = hk;
    const R mw_fuel = fd.mw_fuel[spf];
    const R T_part = rmin(T_in, R(0.99) * cBoilT[spf]);
    const R part_latent = hk + fd.latent[spf] - fd.cp[spf] * (T_part - fd.ref_T);
    v.L_fuel[spf] = part_latent;
    R pres_sat;
    if (fd.cc_psat[spf]) {
      pres_sat = static_cast<R>(kPATM) *
                 rm::exp(part_latent * mw_fuel / static_cast<R>(kRU) * (R(1) / fd.boilT[spf] - R(1) / T_part));
    } else {
      pres_sat = psat<R, NF>(fd, T_part, spf);
    }
    const R gY = g.Y[fspec];
    v.Yfg += gY;
    const R X_fluid = gY * g.mw_mix / mw_fuel;
    if (pres_sat > X_fluid * g.p) {
      const R Xl = Y_l[spf] * mw_part / mw_fuel;
      const R Xv = Xl * pres_sat / g.p;
      mbar_vap += Xv * mw_fuel;
      sumXVap += Xv;
      v.X_vapor[spf] = Xv;
    } else {
      v.X_vapor[spf] = R(0);
      if (Y_l[spf] > R(0)) v.ok = false;
    }
  }
  const R totalmwx = mbar_vap + (R(1) - sumXVap) * g.mw_mix;
  R sumYVap = R(0), sumYfFluid = R(0), sumYSkin = R(0);
  R Ysk[NF];
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    Ysk[spf] = R(0);
    if (v.X_vapor[spf] > R(0)) {
      const int fspec = fuel_species<NS, NF>(spf);
      const R Yfv = v.X_vapor[spf] * fd.mw_fuel[spf] / totalmwx;
      sumYVap += Yfv;
      const R gasY = g.Y[fspec];
      sumYfFluid += gasY;
      Ysk[spf] = Yfv + rule * (gasY - Yfv);
      sumYSkin += Ysk[spf];
    }
  }
  // film molecular weight: vapour species from the 1/3 rule, every other
  // species renormalised from the gas (Drag.H:124-131)
  const R renorm = (R(1) - sumYSkin) / (R(1) - sumYfFluid);
  R inv_mw = R(0);
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    R y = g.Y[n] * renorm;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf)
      if (n == fuel_species<NS, NF>(spf) && Ysk[spf] != R(0)) y = Ysk[spf];
    inv_mw += y / static_cast<R>(M::W[n]);
  }
  v.mw_skin = R(1) / inv_mw;
  v.Ysk_f = sumYSkin;
  R B_M = (sumYVap - sumYfFluid) / rmax(C_eps, (R(1) - sumYVap));
  v.B_M = rmin(R(20), B_M);
  if (!(sumXVap > R(0))) v.ok = false;
}

// Raw network inputs (before normalisation).
template <class R>
PM_HDI void surr_features(R dia, R T_d, R urel, R T_g, R p, R Ysk_f, R Yl0, R Yl1, float* x) {
  x[0] = static_cast<float>(rm::log(dia));
  x[1] = static_cast<float>(T_d);
  x[2] = static_cast<float>(rm::log(urel + static_cast<R>(kSurrU0)));
  x[3] = static_cast<float>(T_g);
  x[4] = static_cast<float>(rm::log(p));
  x[5] = static_cast<float>(Ysk_f);
  x[6] = static_cast<float>(Yl0);
  x[7] = static_cast<float>(Yl1);
}

// Hull of the training distribution (see the norm struct).
template <class R>
PM_HDI bool surr_in_hull(const pelemp_surrogate_norm& n, R dia, R T_d, R Tboil_mix, R urel, R T_g, R p, R Yfg,
                         R B_M) {
  return dia >= R(n.d_min) && dia <= R(n.d_max) && T_g >= R(n.Tg_min) && T_g <= R(n.Tg_max) &&
         p >= R(n.p_min) && p <= R(n.p_max) && urel <= R(n.urel_max) && Yfg <= R(n.Yfg_max) &&
         T_d >= R(n.Td_min) && T_d <= R(n.Td_frac_max) * Tboil_mix && B_M >= R(n.BM_min);
}

// Binary diffusivity fit of fuel spf in N2 at 1 atm (ln D = cubic in ln T),
// used for the species split of m_dot.
template <class M, int NF, class R>
PM_HDI R fuel_n2_diff(R lt, int spf) {
  constexpr int NS = M::NS;
  const int fspec = fuel_species<NS, NF>(spf);
  // pair (1 = N2, fspec) in the row-major i<j pair order of fitdbin
  const int p = 1 * NS - 1 * 2 / 2 + (fspec - 1 - 1);
  const R f = static_cast<R>(M::fitdbin[4 * p]) +
              lt * (static_cast<R>(M::fitdbin[4 * p + 1]) +
                    lt * (static_cast<R>(M::fitdbin[4 * p + 2]) + lt * static_cast<R>(M::fitdbin[4 * p + 3])));
  return rm::exp(f);
}

// Substep rates from the de-normalised network outputs z (the SubRates the
// plain substep_rates would produce).
template <class M, int NF, class R>
PM_HDI void surr_rates(const R* z, const VaporLite<R, NF>& v, const GasAt<R, M::NS>& g, const FuelConst<R, NF>& fd,
                       R T_part, R dia, R pmass, R rho_part, const R* Y_part, const R* vel_part, R y_tab,
                       const R* cBoilT, SubRates<M, NF, R>& o) {
  constexpr int NS = M::NS;
  const R kpi = static_cast<R>(kPIc);
  R cp_part = R(0), Tboil = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    Tboil += Y_part[spf] * cBoilT[spf];
    cp_part += Y_part[spf] * fd.cp[spf];
  }
  const R delT = rmax(g.T - T_part, R(0));
  const R T_skin = T_part + R(1.) / R(3.) * delT;
#pragma unroll
  for (int q = 0; q < 3; ++q) o.diff_vel[q] = g.vel[q] - vel_part[q];
  const R umag = rm::sqrt(o.diff_vel[0] * o.diff_vel[0] + o.diff_vel[1] * o.diff_vel[1] + o.diff_vel[2] * o.diff_vel[2]);
  const R inv_pmass = R(1) / pmass;
  const R rhoD = rm::exp(z[0]);
  const R lamNu = rm::exp(z[1]);
  const R mu = rm::exp(z[2]);
  // drag (Drag.H:262-279), exact given the film viscosity
  const R rho_skin = v.mw_skin * g.p / (static_cast<R>(kRU) * T_skin);
  const R Reyn = rho_skin * umag * dia / mu;
  R drag_coef = R(0);
  if (Reyn > R(1000.)) {
    drag_coef = R(0.424);
  } else if (Reyn > R(1.)) {
    drag_coef = R(24.) / Reyn * (R(1.) + rm::cbrt(Reyn * Reyn) / R(6.));
  } else if (Reyn > R(0.)) {
    drag_coef = R(24.) / Reyn;
  }
  drag_coef *= (R(1.) + R(2.6432) * y_tab);
  o.drag_force = R(0.125) * rho_skin * drag_coef * kpi * dia * dia * umag;
  o.drag_force_p = R(0.75) * rho_skin * drag_coef * umag / (dia * rho_part);
  // mass transfer (Drag.H:299-340), exact given rho D and mu
  const R Sc = mu / rhoD;
  const R powR = rmax(rm::pow(rmin(R(400.), Reyn), R(0.077)), R(1.));
  const R logB = rm::log1p(v.B_M);
  const R invFM = v.B_M / (logB * rm::pow(R(1.) + v.B_M, R(0.7)));
  const R Sh_0 = R(1.) + powR * rm::cbrt(R(1.) + Reyn * Sc);
  const R Sh_num = R(2.) + (Sh_0 - R(2.)) * invFM;
  const R mdotcoeff = kpi * dia * Sh_num * logB;
  // species split: X_v,k D_k,N2 / sum
  const R lt = rm::log(T_skin);
  R wk[NF], wsum = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    wk[spf] = (v.X_vapor[spf] > R(0)) ? v.X_vapor[spf] * fuel_n2_diff<M, NF, R>(lt, spf) : R(0);
    wsum += wk[spf];
  }
  R m_dot = R(0), sumL = R(0);
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) {
    const R mi = -rmax(rhoD * (wk[spf] / wsum) * mdotcoeff, R(0.));
    o.mi_dot[spf] = mi;
    m_dot += mi;
    sumL += mi * v.L_fuel[spf];
  }
  const R conv_src = kpi * lamNu * dia * delT;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) o.h_part[fuel_species<NS, NF>(spf)] = v.h_fuel[spf];
  o.Reyn = Reyn;
  o.m_dot = m_dot;
  o.sumL = sumL;
  o.conv_src = conv_src;
  o.part_temp_src = (sumL + conv_src) * inv_pmass / cp_part;
  o.cp_part = cp_part;
  o.Tboil = Tboil;
  o.delT = delT;
  o.inv_pmass = inv_pmass;
  o.ShRhoD = Sh_num * rhoD;
  o.lamNu = lamNu;
  o.rhoD = rhoD;
  o.mu_skin = mu;
  o.B_M = v.B_M;
  o.sumYVap = R(0);
}

// Error indicator: the largest flow_dt / tau of the substep-count rule
// (Drag.H:287-292,348-359) evaluated with the predicted rates.  Below 1 the
// plain method takes a single substep; the surrogate is used only when it
// stays below the configured margin, so a prediction error cannot change
// the substep count and fast-changing droplets go to the plain path.
template <class M, int NF, class R>
PM_HDI R surr_indicator(const SubRates<M, NF, R>& o, R flow_dt, R pmass, bool mom_trans) {
  R ind = R(0);
  if (mom_trans) ind = flow_dt * o.drag_force_p;
  const R inv_tau_d = -o.m_dot / (R(0.2) * pmass);
  const R inv_tau_T = o.conv_src * o.inv_pmass / (o.cp_part * o.delT);
  ind = rmax(ind, flow_dt * inv_tau_d);
  ind = rmax(ind, flow_dt * inv_tau_T);
  return ind;
}

}  // namespace pelemp
