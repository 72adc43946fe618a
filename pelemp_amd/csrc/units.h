// units.h -- the PeleC flavour's CGS inputs and outputs and the SI units the
// droplet model computes in.
//
// PeleC runs the spray in CGS with SprayUnits factors of 1
// (SprayFuelData.H:55-64); this build's gas properties are SI, so a PeleC
// context converts once at its boundary: SprayData and the geometry at
// creation, droplets at set/get, the gas state per cell in the gather and the
// sources per cell at the add (the factors below).  The thresholds the
// reference writes in CGS (min_mass 1e-13 g, film height 1e-5 cm) are the same
// physical values as the SI ones the model uses.  Host-only, shared by the
// library and the CPU oracle so both convert with the same arithmetic.
#pragma once

#include "pelemp_spray.h"

namespace pelemp {
namespace cgs {

constexpr double kLen = 1.e-2;     // cm -> m (also cm/s -> m/s, cm/s^2 -> m/s^2)
constexpr double kRho = 1.e3;      // g/cm^3 -> kg/m^3
constexpr double kSpecE = 1.e-4;   // erg/g -> J/kg
constexpr double kPres = 1.e-1;    // dyn/cm^2 -> Pa
constexpr double kSigma = 1.e-3;   // dyn/cm -> N/m
constexpr double kVisc = 1.e-1;    // poise -> Pa s
constexpr double kCond = 1.e-5;    // erg/(cm s K) -> W/(m K)
constexpr double kVol = 1.e-6;     // cm^3 -> m^3
// SI source -> CGS source per cell: momentum kg/(m^2 s^2) -> g/(cm^2 s^2),
// mass and species kg/(m^3 s) -> g/(cm^3 s), energy W/m^3 -> erg/(cm^3 s)
constexpr double kSrcMom = 1.e-1;
constexpr double kSrcMass = 1.e-3;
constexpr double kSrcEng = 1.e1;

inline void spray_data_to_si(pelemp_spray_data& sd) {
  sd.sigma *= kSigma;
  for (int d = 0; d < 3; ++d) sd.body_force[d] *= kLen;
  for (int f = 0; f < PELEMP_MAX_FUEL; ++f) {
    sd.cp[f] *= kSpecE;
    sd.latent[f] *= kSpecE;
    sd.ref_latent[f] *= kSpecE;
    sd.psat_coef[4 * f + 3] *= kPres;  // the Antoine factor d; a, b, c are in K
    for (int q = 0; q < 4; ++q) {
      sd.rho_coef[4 * f + q] *= kRho;
      sd.lambda_coef[4 * f + q] *= kCond;
      sd.mu_coef[4 * f + q] *= kVisc;
    }
  }
}

inline void geom_to_si(pelemp_geom& g) {
  for (int d = 0; d < 3; ++d) {
    g.plo[d] *= kLen;
    g.phi[d] *= kLen;
    g.dx[d] *= kLen;
  }
}

// one droplet in the reference's AoS meaning (pos[3], rdata[10 + nf]); BM1
// is a shed volume N0 r^3 for KH-RT (do_breakup 2), dimensionless for TAB
inline void droplet_to_si(double* pos, double* rd, int nf, int do_breakup) {
  for (int d = 0; d < 3; ++d) {
    pos[d] *= kLen;
    rd[d] *= kLen;
  }
  rd[4] *= kLen;           // diameter
  rd[9 + nf] *= kLen;      // film height
  if (do_breakup == 2) rd[7 + nf] *= kVol;
}

inline void droplet_to_cgs(double* pos, double* rd, int nf, int do_breakup) {
  for (int d = 0; d < 3; ++d) {
    pos[d] /= kLen;
    rd[d] /= kLen;
  }
  rd[4] /= kLen;
  rd[9 + nf] /= kLen;
  if (do_breakup == 2) rd[7 + nf] /= kVol;
}

constexpr double kMass = 1.e-3;    // g -> kg (also g/s -> kg/s)

// a spray jet's lengths, velocities, diameters and masses to SI; the
// held-over mass keeps its "none" value (-1)
inline void jet_to_si(pelemp_spray_jet& j) {
  for (int d = 0; d < 3; ++d) j.cent[d] *= kLen;
  j.jet_dia *= kLen;
  j.jet_vel *= kLen;
  j.max_jet_vel *= kLen;
  j.mass_flow *= kMass;
  j.dist_param[0] *= kLen;
  if (j.dist_type == PELEMP_DIST_NORMAL || j.dist_type == PELEMP_DIST_LOGNORMAL) j.dist_param[1] *= kLen;
  if (j.sum_inj_mass > 0.) j.sum_inj_mass *= kMass;
  j.total_inj_mass *= kMass;
}

// the injection state of an SI jet back into the CGS jet it came from
inline void jet_state_to_cgs(const pelemp_spray_jet& si, pelemp_spray_jet& j) {
  j.num_ppp = si.num_ppp;
  j.sum_inj_mass = si.sum_inj_mass > 0. ? si.sum_inj_mass / kMass : si.sum_inj_mass;
  j.sum_inj_time = si.sum_inj_time;
  j.min_parcel = si.min_parcel;
  j.total_inj_mass = si.total_inj_mass / kMass;
  j.total_inj_time = si.total_inj_time;
  j.n_injections = si.n_injections;
}

// SI -> CGS factor of each source component in deposition order:
// momentum(3), mass, species(nf), energy
inline void source_scale(int nf, double* s) {
  for (int d = 0; d < 3; ++d) s[d] = kSrcMom;
  s[3] = kSrcMass;
  for (int f = 0; f < nf; ++f) s[4 + f] = kSrcMass;
  s[4 + nf] = kSrcEng;
}

// SI -> CGS factor of each accumulated derived variable before its
// protected divide (computeDerivedVars order): mass, density, num, volume,
// surface, volume fraction, sum N d, sum 6 N V, film height, film mass,
// sum N m T, parcels, sum N m u (3), fuel masses
inline void derived_scale(int nf, double* s) {
  const double base[15] = {1.e3, 1.e-3, 1., 1.e6, 1.e4, 1., 1.e2, 1.e6, 1.e2, 1.e3, 1.e3, 1., 1.e5, 1.e5, 1.e5};
  for (int q = 0; q < 15; ++q) s[q] = base[q];
  for (int f = 0; f < nf; ++f) s[15 + f] = 1.e3;
}

}  // namespace cgs
}  // namespace pelemp
