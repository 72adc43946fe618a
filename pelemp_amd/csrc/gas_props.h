// gas_props.h -- gas-phase species properties used by the droplet update.
//
// These stand in for the PelePhysics calls the reference makes from the
// droplet kernel (absent from the reference tree):
//   eos.molecular_weight   SprayParticles.cpp:332
//   eos.T2Cpi, eos.T2Hi    Drag.H:220-221, SpraySetup.cpp:286
//   trans.transport(...)   Drag.H:252-256  ("Simple" mixture-averaged model)
// All quantities are SI (J/kg/K, J/kg, Pa s, W/m/K, kg/m/s for rho*D), so the
// SprayUnits conversions of the LMeX flavour are identities here.
//
// Models (tables in mech_data.h, generated by mech/gen_mech.py):
//   cp_k/R  = a0 + a1 T + a2 T^2 + a3 T^3 + a4 T^4               (NASA-7)
//   h_k/RT  = a0 + a1 T/2 + a2 T^2/3 + a3 T^3/4 + a4 T^4/5 + a5/T
//   ln mu_k, ln lambda_k, ln D_ij(1 atm) = cubic in ln T
//   mu      = (sum_k X_k mu_k^6)^(1/6)
//   lambda  = (sum_k X_k lambda_k^(1/4))^4
//   Ddiag_k = rho D_k W_k / W_mix,
//   D_k     = (sum_{j!=k} Y_j) / (sum_{j!=k} X_j / D_kj(p)),
//             D_kj(p) = D_kj(1 atm) p_atm / p,  p = rho R T / W_mix
// Ddiag carries the factor W_k/W_mix (diffusion flux written with mole-
// fraction gradients); the droplet model removes it again when it converts
// Ddiag of the fuel with mw_skin / mw_fuel (Drag.H:312-314).
// The functions are templated on the arithmetic type: double evaluates with
// the project's pm:: elementary functions (bitwise identical on host and
// device), float with the gfx950 hardware approximations.
#pragma once

#include "mech_data.h"
#include "pm_math.h"

namespace pelemp {

constexpr double kRU = 8.31446261815324;  // J/mol/K
constexpr double kPATM = 101325.0;        // Pa

// Arithmetic-type dispatch for the elementary functions.
namespace rm {
PM_HD double exp(double x) { return pm::exp(x); }
PM_HD double log(double x) { return pm::log(x); }
PM_HD double log1p(double x) { return pm::log1p(x); }
PM_HD double pow(double x, double y) { return pm::pow(x, y); }
PM_HD double cbrt(double x) { return pm::cbrt(x); }
PM_HD double sqrt(double x) { return std::sqrt(x); }
PM_HD double fabs(double x) { return std::fabs(x); }
PM_HD double sin(double x) { return pm::sin(x); }
PM_HD double cos(double x) { return pm::cos(x); }
PM_HD double tan(double x) { return pm::tan(x); }
PM_HD double asin(double x) { return pm::asin(x); }
PM_HD double acos(double x) { return pm::acos(x); }
PM_HD double atan2(double y, double x) { return pm::atan2(y, x); }

#if defined(__HIP_DEVICE_COMPILE__)
// Single precision: the hardware base-2 exponential and logarithm
// (v_exp_f32, v_log_f32, about 1 ulp) without the rescaling steps for
// denormal arguments and results that the library forms add: the arguments
// here are positive physical quantities (temperatures, pressures, mixture
// sums, transfer numbers) far from the denormal range.
__device__ inline float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ inline float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
// log(1 + x) for x >= 0 with the rounding of 1 + x corrected (Goldberg):
// log(u) * x / (u - 1), u = fl(1 + x)
__device__ inline float log1p(float x) {
  const float u = 1.f + x;
  return (u == 1.f) ? x : log(u) * (x / (u - 1.f));
}
__device__ inline float pow(float x, float y) {
  return (x == 0.0f) ? ((y > 0.0f) ? 0.0f : __builtin_huge_valf()) : exp(y * log(x));
}
// x >= 0 on every call site (masses, Reynolds and Peclet groups, mixture
// sums): the hardware exp2/log2 pair instead of the software cbrtf
__device__ inline float cbrt(float x) { return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.f / 3.f)); }
__device__ inline float sqrt(float x) { return __builtin_sqrtf(x); }
__device__ inline float fabs(float x) { return __builtin_fabsf(x); }
#else
inline float exp(float x) { return std::exp(x); }
inline float log(float x) { return std::log(x); }
inline float log1p(float x) { return std::log1p(x); }
inline float pow(float x, float y) { return std::pow(x, y); }
inline float cbrt(float x) { return std::cbrt(x); }
inline float sqrt(float x) { return std::sqrt(x); }
inline float fabs(float x) { return std::fabs(x); }
#endif
}  // namespace rm

template <class M>
struct GasProps {
  static constexpr int NS = M::NS;

  // Molecular weights [kg/mol] (eos.molecular_weight * mass_conv).
  template <class R>
  static PM_HDI void molecular_weight(R* mw) {
#pragma unroll
    for (int n = 0; n < NS; ++n) mw[n] = static_cast<R>(M::W[n]);
  }

  // NASA-7 coefficient a_q of species n for the range of T (lo: T < Tmid).
  static PM_HD constexpr double coef(bool lo, int n, int q) {
    return lo ? M::nasa_lo[7 * n + q] : M::nasa_hi[7 * n + q];
  }
  static PM_HD constexpr double tmid_min() {
    double t = M::Tmid[0];
    for (int n = 1; n < NS; ++n) t = M::Tmid[n] < t ? M::Tmid[n] : t;
    return t;
  }
  static PM_HD constexpr double tmid_max() {
    double t = M::Tmid[0];
    for (int n = 1; n < NS; ++n) t = M::Tmid[n] > t ? M::Tmid[n] : t;
    return t;
  }

  // cp_n of species n with the coefficients of range lo (compile-time) or of
  // the range T falls in (a select per coefficient)
  template <class R, int RANGE>  // RANGE: 1 lower, 0 upper, -1 by T
  static PM_HDI R cp_one(R T, int n) {
    const bool lo = RANGE == 1 ? true : (RANGE == 0 ? false : T < static_cast<R>(M::Tmid[n]));
    const R a0 = static_cast<R>(coef(lo, n, 0));
    const R a1 = static_cast<R>(coef(lo, n, 1));
    const R a2 = static_cast<R>(coef(lo, n, 2));
    const R a3 = static_cast<R>(coef(lo, n, 3));
    const R a4 = static_cast<R>(coef(lo, n, 4));
    const R rw = static_cast<R>(kRU / M::W[n]);
    return rw * (a0 + T * (a1 + T * (a2 + T * (a3 + T * a4))));
  }
  template <class R, int RANGE>
  static PM_HDI R h_one(R T, int n) {
    const bool lo = RANGE == 1 ? true : (RANGE == 0 ? false : T < static_cast<R>(M::Tmid[n]));
    // coefficient selects between two compile-time constants
    const R b0 = static_cast<R>(coef(lo, n, 0));
    const R b1 = static_cast<R>(coef(lo, n, 1) * 0.5);
    const R b2 = static_cast<R>(coef(lo, n, 2) * (1.0 / 3.0));
    const R b3 = static_cast<R>(coef(lo, n, 3) * 0.25);
    const R b4 = static_cast<R>(coef(lo, n, 4) * 0.2);
    const R b5 = static_cast<R>(coef(lo, n, 5));
    const R rw = static_cast<R>(kRU / M::W[n]);
    return rw * (T * (b0 + T * (b1 + T * (b2 + T * (b3 + T * b4)))) + b5);
  }

  // Range of T for the whole wavefront: 1 when every lane is below every
  // species' Tmid, 0 when every lane is at or above it, -1 otherwise.  The
  // per-lane values are the same either way; a uniform range only drops the
  // per-coefficient selects.
  template <class R>
  static PM_HDI int wave_range(R T) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (__all(T < static_cast<R>(tmid_min()))) return 1;
    if (__all(!(T < static_cast<R>(tmid_max())))) return 0;
#else
    (void)T;
#endif
    return -1;
  }

  // Species heat capacities at constant pressure [J/kg/K].
  template <class R>
  static PM_HDI void T2Cpi(R T, R* cp) {
    const int range = wave_range(T);
    if (range == 1) {
#pragma unroll
      for (int n = 0; n < NS; ++n) cp[n] = cp_one<R, 1>(T, n);
    } else if (range == 0) {
#pragma unroll
      for (int n = 0; n < NS; ++n) cp[n] = cp_one<R, 0>(T, n);
    } else {
#pragma unroll
      for (int n = 0; n < NS; ++n) cp[n] = cp_one<R, -1>(T, n);
    }
  }

  // Species specific enthalpies [J/kg] (incl. heat of formation).
  template <class R>
  static PM_HDI void T2Hi(R T, R* h) {
    const int range = wave_range(T);
    if (range == 1) {
#pragma unroll
      for (int n = 0; n < NS; ++n) h[n] = h_one<R, 1>(T, n);
    } else if (range == 0) {
#pragma unroll
      for (int n = 0; n < NS; ++n) h[n] = h_one<R, 0>(T, n);
    } else {
#pragma unroll
      for (int n = 0; n < NS; ++n) h[n] = h_one<R, -1>(T, n);
    }
  }

  // Temperature from the mixture's specific internal energy e [J/kg] and
  // mass fractions (the eos.EY2T the PeleC flavour calls in the gather,
  // SprayInterpolation.H:143-146): Newton on e(T) = sum_k Y_k (h_k - R T / W_k)
  // with c_v = sum_k Y_k (c_p,k - R / W_k), from the guess T (the state's
  // temperature), at most 50 steps, until |dT| <= 1e-10 T.
  static PM_HDI double EY2T(double e, const double* Y, double T) {
    double h[NS], cp[NS];
    for (int it = 0; it < 50; ++it) {
      T2Hi<double>(T, h);
      T2Cpi<double>(T, cp);
      double ev = 0., cv = 0.;
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const double rw = kRU / M::W[n];
        ev += Y[n] * (h[n] - rw * T);
        cv += Y[n] * (cp[n] - rw);
      }
      const double dT = (e - ev) / cv;
      T = T + dT;
      if (std::fabs(dT) <= 1e-10 * T) break;
    }
    return T;
  }

  // Mixture-averaged transport at (T, rho, Y).  Ddiag[k] = rho D_k,mix for
  // k >= D_FIRST (the others are left as they are; each Ddiag[k] is computed
  // on its own, so the ones computed do not depend on D_FIRST).
  template <class R, int D_FIRST = 0>
  static PM_HDI void transport(bool get_mu, bool get_lam, bool get_D, R T,
                              R rho, const R* Y, R* Ddiag, R& mu, R& lam) {
    const R lt = rm::log(T);
    const R lt2 = lt * lt;
    const R lt3 = lt2 * lt;
    R X[NS];
    R sumYW = R(0);
#pragma unroll
    for (int n = 0; n < NS; ++n) sumYW += Y[n] / static_cast<R>(M::W[n]);
    const R wbar = R(1) / sumYW;
#pragma unroll
    for (int n = 0; n < NS; ++n) X[n] = Y[n] / static_cast<R>(M::W[n]) * wbar;
    if (get_mu) {
      R s = R(0);
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const R f = static_cast<R>(M::fitmu[4 * n]) + static_cast<R>(M::fitmu[4 * n + 1]) * lt +
                    static_cast<R>(M::fitmu[4 * n + 2]) * lt2 +
                    static_cast<R>(M::fitmu[4 * n + 3]) * lt3;
        if constexpr (sizeof(R) == 4) {
          s += X[n] * rm::exp(R(6) * f);  // mu_k^6 with one exponential
        } else {
          const R m = rm::exp(f);
          const R m3 = m * m * m;
          s += X[n] * (m3 * m3);
        }
      }
      if constexpr (sizeof(R) == 4) {
        mu = rm::exp(rm::log(s) * R(1. / 6.));  // s^(1/6)
      } else {
        mu = rm::cbrt(rm::sqrt(s));
      }
    }
    if (get_lam) {
      R s = R(0);
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const R f = static_cast<R>(M::fitlam[4 * n]) + static_cast<R>(M::fitlam[4 * n + 1]) * lt +
                    static_cast<R>(M::fitlam[4 * n + 2]) * lt2 +
                    static_cast<R>(M::fitlam[4 * n + 3]) * lt3;
        if constexpr (sizeof(R) == 4) {
          s += X[n] * rm::exp(R(0.25) * f);  // lambda_k^(1/4) with one exponential
        } else {
          const R l = rm::exp(f);
          s += X[n] * rm::sqrt(rm::sqrt(l));
        }
      }
      const R s2 = s * s;
      lam = s2 * s2;
    }
    if (get_D) {
      // inverse binary diffusivities at 1 atm, pairs i<j row-major
      constexpr int NP = NS * (NS - 1) / 2;
      R invD[NP > 0 ? NP : 1];
#pragma unroll
      for (int a = 0; a < NS; ++a) {
#pragma unroll
        for (int b = a + 1; b < NS; ++b) {
          const int p = a * NS - a * (a + 1) / 2 + (b - a - 1);
          if (b < D_FIRST) {  // the pair of two species whose Ddiag is not wanted
            invD[p] = R(0);
            continue;
          }
          const R f = static_cast<R>(M::fitdbin[4 * p]) + static_cast<R>(M::fitdbin[4 * p + 1]) * lt +
                      static_cast<R>(M::fitdbin[4 * p + 2]) * lt2 + static_cast<R>(M::fitdbin[4 * p + 3]) * lt3;
          invD[p] = rm::exp(-f);
        }
      }
      // p / p_atm from the ideal gas law
      const R pratio = rho * static_cast<R>(kRU) * T / wbar / static_cast<R>(kPATM);
#pragma unroll
      for (int i = D_FIRST; i < NS; ++i) {
        R term1 = R(0);
        R term2 = R(0);
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          if (j == i) continue;
          const int a = i < j ? i : j;
          const int b = i < j ? j : i;
          const int pidx = a * NS - a * (a + 1) / 2 + (b - a - 1);
          term1 += Y[j];
          term2 += X[j] * invD[pidx];
        }
        Ddiag[i] = rho * term1 / (term2 * pratio) * (static_cast<R>(M::W[i]) / wbar);
      }
    }
  }
};

}  // namespace pelemp
