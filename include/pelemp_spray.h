/* pelemp_spray.h -- C ABI of the MI355X spray droplet update.
 *
 * This library replaces the per-step Lagrangian spray update of PeleMP,
 * SprayParticleContainer::updateParticles (Source/PP_Spray/SprayParticles.H:
 * 229-243, SprayParticles.cpp:142-502), together with the setup it reads
 * (SprayData / SprayComps, Source/PP_Spray/SprayFuelData.H:8-212;
 * SpraySetup.cpp:255-297) and the particle-CFL time-step estimate
 * (SprayParticleContainer::estTimestep, SprayParticles.cpp:89-140).
 *
 * All entry points are plain C: pointers, sizes and PODs, no C++ or torch
 * types.  Every call returns a status code (PELEMP_OK == 0).  Where the
 * reference calls amrex::Abort, the library returns the matching error code
 * instead; the device-side checks record a sticky error word that the next
 * synchronising call reports.
 *
 * Layout conventions.
 *  - Mesh data (gas state, spray source) is an AMReX-style FAB: one box of
 *    cells [lo, hi] (inclusive, ghost cells included), components slowest,
 *    x fastest: data[((n*nz + (k-lo2))*ny + (j-lo1))*nx + (i-lo0)].
 *  - Particles cross the boundary in the reference's AoS meaning: position
 *    (3 doubles), real components rdata[0..NSR) with NSR = 10 + nfuel
 *    ordered as SprayComps (vel 0..2, T 3, dia 4, Y 5..4+nf, numDens 5+nf,
 *    N0 6+nf, BM1 7+nf, BM2 8+nf, FilmHght 9+nf), and a 64-bit id (> 0 is
 *    active).  Inside the library they live in device SoA tiles sorted by
 *    cell.
 */
#ifndef PELEMP_SPRAY_H
#define PELEMP_SPRAY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PELEMP_MAX_FUEL 8

/* Droplets per workgroup segment of a 4x4x4-cell tile.  A tile holding more
 * droplets than this is updated by several workgroups, one per segment of
 * its cell-sorted droplets; each cell's source sum is then the sum, in
 * segment order, of the per-segment sums (each in cell-sorted droplet
 * order).  This fixes the association order of the deposition, so results
 * do not depend on thread timing. */
#define PELEMP_SEGMENT_DROPLETS 4096

/* Status codes (reference: amrex::Abort sites listed per code). */
#define PELEMP_OK 0
#define PELEMP_E_STATE_BOX 1   /* SprayInterpolation.H:121-124 */
#define PELEMP_E_SRC_BOX 2     /* SprayParticles.cpp:420-425 */
#define PELEMP_E_LEFT_DOMAIN 3 /* SprayParticles.cpp:347-351 */
#define PELEMP_E_BAD_CONFIG 4  /* SpraySetup.cpp:92-100,152-183 */
#define PELEMP_E_HIP 5         /* HIP runtime error */
#define PELEMP_E_ARG 6         /* invalid argument / size */
#define PELEMP_E_NOMEM 7       /* device allocation failed */
#define PELEMP_E_BAD_CHILD 8   /* SpraySB.cpp:187-190: a secondary droplet
                                  without a valid position */
#define PELEMP_E_BAD_INJECTION 9 /* SprayInjection.H:39-43 (mean droplet too
                                    small), :149-152 (injected droplet
                                    outside the domain), :286-289 (lattice
                                    droplet outside), and the jet checks of
                                    SprayJet.cpp:28-133, SprayJet.H:71-138 */
#define PELEMP_E_EB_STENCIL 10 /* SprayInterpolation.H:562: no uncovered
                                  cell of large enough volume fraction in the
                                  EB stencil (the reference then indexes
                                  indx_array[-1]) */

/* Mirror of SprayData (SprayFuelData.H:115-212), fixed at PELEMP_MAX_FUEL
 * fuel slots; the first nfuel are used. */
typedef struct pelemp_spray_data {
  int32_t mass_trans;  /* evaporation on */
  int32_t mom_trans;   /* momentum coupling on */
  int32_t fixed_parts; /* particles fixed in place */
  int32_t do_splash;
  int32_t do_breakup; /* 0 none, 1 TAB, 2 KHRT */
  int32_t nfuel;
  double min_eb_vfrac;
  double ref_T;
  double sigma;   /* surface tension */
  double wall_T;
  double theta_c; /* contact angle [rad] */
  double dtmod;   /* fraction of dt the particle update uses */
  double body_force[3];
  double critT[PELEMP_MAX_FUEL];
  double boilT[PELEMP_MAX_FUEL];
  double cp[PELEMP_MAX_FUEL];
  double latent[PELEMP_MAX_FUEL];     /* ref_latent - h_gas(ref_T) */
  double ref_latent[PELEMP_MAX_FUEL];
  double psat_coef[4 * PELEMP_MAX_FUEL];
  double rho_coef[4 * PELEMP_MAX_FUEL];
  double lambda_coef[4 * PELEMP_MAX_FUEL];
  double mu_coef[4 * PELEMP_MAX_FUEL];
  int32_t indx[PELEMP_MAX_FUEL];     /* gas species index of each fuel */
  int32_t dep_indx[PELEMP_MAX_FUEL]; /* gas species receiving the vapour */
} pelemp_spray_data;

/* Mirror of the runtime part of SprayComps (SprayFuelData.H:20-28). */
typedef struct pelemp_spray_comps {
  int32_t rhoIndx, momIndx, engIndx, utempIndx, specIndx;
  int32_t rhoSrcIndx, momSrcIndx, engSrcIndx, specSrcIndx;
} pelemp_spray_comps;

/* One FAB: box [lo,hi] incl. ghosts, ncomp components, component-major. */
typedef struct pelemp_fab_view {
  double* data;
  int32_t lo[3];
  int32_t hi[3];
  int32_t ncomp;
} pelemp_fab_view;

/* The same FAB in single precision (a gas state kept in FP32). */
typedef struct pelemp_fab_view_f32 {
  float* data;
  int32_t lo[3];
  int32_t hi[3];
  int32_t ncomp;
} pelemp_fab_view_f32;

/* Level geometry: physical box, cell size, valid cell-index domain and
 * boundary kinds per face (0 periodic, 1 reflective wall, -1 outflow),
 * as built at SprayParticles.cpp:183-218. */
typedef struct pelemp_geom {
  double plo[3];
  double phi[3];
  double dx[3];
  int32_t domain_lo[3];
  int32_t domain_hi[3];
  int32_t bndry_lo[3];
  int32_t bndry_hi[3];
} pelemp_geom;

/* Gas mechanisms compiled into the library (species order of the state). */
#define PELEMP_MECH_DODECANE_AIR 0        /* O2 N2 NC12H26 */
#define PELEMP_MECH_DECANE_DODECANE_AIR 1 /* O2 N2 NC10H22 NC12H26 */
#define PELEMP_MECH_ALKANE6_AIR 2         /* O2 N2 NC8..NC16 (6 fuels) */
#define PELEMP_MECH_DECANE_AIR 3          /* O2 N2 NC10H22 (the decks' decane_3sp species) */

/* Arithmetic form of the per-droplet update. */
#define PELEMP_FORM_FP64 0      /* reference method, FP64 end to end */
#define PELEMP_FORM_FP32 1      /* FP32 droplet state + ODE, FP64 sums */
#define PELEMP_FORM_SURROGATE 2 /* FP32 + MLP for the evaporation rates */

typedef struct pelemp_spray_ctx pelemp_spray_ctx;

/* Create a context for one AMR level on the current HIP device.
 * Reference: readSprayParams + spraySetup (SpraySetup.cpp:73-297); the
 * latent-heat shift latent = ref_latent - h_gas(ref_T) is applied here when
 * apply_latent_shift != 0.  max_particles sizes the device tiles. */
int pelemp_spray_create(pelemp_spray_ctx** ctx, int32_t mech, int32_t form,
                        const pelemp_spray_data* sd,
                        const pelemp_spray_comps* comps,
                        const pelemp_geom* geom, int64_t max_particles,
                        int32_t apply_latent_shift);
/* The PeleC flavour of the update (built without PELELM_USE_SPRAY): the
 * state holds rho, rho*u, rho*E, rho*Y_k and T (T only as the guess of
 * eos.EY2T, SprayInterpolation.H:126-146), the energy source carries the
 * drag work and the kinetic energy of the evaporated mass (Drag.H:282-286,
 * 418-421), and everything is CGS (SprayData with the deck's unshifted
 * reference latent heat, geometry, droplets, state, sources, derived
 * variables, avg_inject_mass, jet velocity), as PeleC's SprayUnits factors of
 * 1 imply (SprayFuelData.H:55-64).  FP64 form only. */
int pelemp_spray_create_pelec(pelemp_spray_ctx** ctx, int32_t mech, int32_t form,
                              const pelemp_spray_data* sd_cgs,
                              const pelemp_spray_comps* comps,
                              const pelemp_geom* geom_cgs, int64_t max_particles);
int pelemp_spray_destroy(pelemp_spray_ctx* ctx);

/* The SprayData actually in use (after the latent shift). */
int pelemp_spray_get_data(const pelemp_spray_ctx* ctx, pelemp_spray_data* out);

/* Container-level parameters of the breakup and splash models, static
 * members of SprayParticleContainer in the reference (SprayParticles.H:
 * 388-405, defaults SpraySetup.cpp:15-19, read at :144-160).  rng_seed keys
 * the counter-based numbers that replace amrex::Random() for the azimuth of
 * breakup children (SpraySB.cpp:153). */
typedef struct pelemp_spray_params {
  double max_parcel_size;       /* max_parcel_size, default 100 */
  double breakup_parcel_factor; /* breakup_parcel_factor, default 0.5 */
  double khrt_B0, khrt_B1, khrt_C3; /* KHRT_B0/B1/C3, defaults 0.61, 7, 1 */
  uint64_t rng_seed;            /* default 0 */
  /* KH-RT: mean injected parcel mass sum_jets N d^3 / n_jets, which the
   * reference takes from its spray jets (SprayParticles.cpp:230-239);
   * default 0, as with no jets */
  double avg_inject_mass;
} pelemp_spray_params;
int pelemp_spray_set_params(pelemp_spray_ctx* ctx, const pelemp_spray_params* p);
int pelemp_spray_get_params(const pelemp_spray_ctx* ctx, pelemp_spray_params* out);

/* Particle transfer, host AoS meaning (see header comment):
 * pos[n*3], rdata[n*NSR], id[n].  Inactive ids (<= 0) are dropped. */
int pelemp_spray_set_particles(pelemp_spray_ctx* ctx, int64_t n,
                               const double* pos, const double* rdata,
                               const int64_t* id);
/* Copies back the active particles in the library's current (cell-sorted)
 * order; *n is in: capacity, out: count. */
int pelemp_spray_get_particles(pelemp_spray_ctx* ctx, int64_t* n, double* pos,
                               double* rdata, int64_t* id);
int pelemp_spray_num_particles(pelemp_spray_ctx* ctx, int64_t* n);
/* Appends particles (same AoS meaning) to those already in the context, as
 * the reference's injection routines append host particles to the tiles
 * (SprayInjection.H:164-179, :297-310); for injectors that run on the host,
 * e.g. a SprayJet subclass with its own get_new_particle.  Inside each cell
 * the new droplets follow the existing ones, in input order.  Ids must be
 * > 0; later secondary droplets take ids past the largest id in the
 * context. */
int pelemp_spray_add_particles(pelemp_spray_ctx* ctx, int64_t n,
                               const double* pos, const double* rdata,
                               const int64_t* id);

/* ---- Embedded boundaries (SprayInterpolation.H:201-634,
 *      WallFunctions.H:124-157, SprayParticles.cpp:192-199,271-291,413-418) */

/* The EB data of the level for the state box, as AMReX's
 * EBFArrayBoxFactory holds it (getEBCellFlagFab, getVolFrac, getCentroid,
 * getBndryCent, getBndryNormal), device pointers owned by the caller, each a
 * FAB over the box [lo, hi] (component-major, x fastest).  flags: AMReX
 * EBCellFlag words, cell type in bits 0-1 (0 regular, 1 single-valued,
 * 2 multi-valued, 3 covered), neighbour (i, j, k) in {-1,0,1}^3 connected
 * when bit PELEMP_EB_NGBR_BIT(i, j, k) is set.  Centroids are relative to
 * the cell centre in cell units; the boundary normal points out of the
 * fluid. */
#define PELEMP_EB_NGBR_BIT(i, j, k) (5 + ((i) + 1) + 3 * ((j) + 1) + 9 * ((k) + 1))
typedef struct pelemp_eb_view {
  const uint32_t* flags;
  const double* vfrac;
  const double* ccent; /* 3 components */
  const double* bcent; /* 3 components */
  const double* bnorm; /* 3 components */
  int32_t lo[3];
  int32_t hi[3];
} pelemp_eb_view;
/* Attach (eb != NULL) or detach the EB of the level.  With an EB that has
 * cut or covered cells in the box, the update interpolates with
 * eb_interp, scales deposits in cut cells by 1/vfrac, reflects droplets
 * (or splashes them) off the EB, and deposits a droplet into the closest
 * acceptable cell where eb_interp moves it; FP64 form.  The box must cover
 * the state box used by the updates. */
int pelemp_spray_set_eb(pelemp_spray_ctx* ctx, const pelemp_eb_view* eb);

/* ---- Injection (SprayInjection.H, SprayJet.{H,cpp}, Distributions.cpp) ---- */

/* Droplet-diameter distributions (Distribution/Distributions.cpp:1-174);
 * dist_param meaning per type. */
#define PELEMP_DIST_UNIFORM 0    /* [diameter] */
#define PELEMP_DIST_NORMAL 1     /* [mean_dia, std_dev] */
#define PELEMP_DIST_LOGNORMAL 2  /* [mean_dia, std_dev] of the diameter */
#define PELEMP_DIST_WEIBULL 3    /* [mean_dia, k] */
#define PELEMP_DIST_CHISQUARED 4 /* [d32] */

/* The members of SprayJet (SprayJet.H:262-293) that the injection reads and
 * updates; angles in radians, as stored by the SprayJet constructors
 * (SprayJet.cpp:30-37,116-121).  Units are the context's (SI, or CGS for a
 * PeleC context).  The rate-of-injection tables of readROI are caller-owned
 * arrays of n_roi points (n_roi = 0: constant jet_vel and mass_flow).
 * amrex::Random() draws are replaced by counter-based numbers keyed by
 * (seed, n_injections, candidate droplet, draw); n_injections counts the
 * calls that reached the droplet loop. */
typedef struct pelemp_spray_jet {
  double cent[3];
  double norm[3];       /* normalised by pelemp_spray_jet_check */
  double spread_angle;  /* m_spreadAngle: full cone angle */
  double swirl_angle;   /* m_swirlAngle */
  double hollow_spread; /* m_hollowSpread */
  double jet_dia;
  double jet_vel;       /* m_jetVel (constant jet) */
  double max_jet_vel;   /* m_maxJetVel (estTimestep's jet bound) */
  double mass_flow;     /* m_massFlow (constant jet) */
  double jet_T;
  double jet_Y[PELEMP_MAX_FUEL];
  double start_time, end_time;
  int32_t hollow_spray;
  int32_t dist_type;
  double dist_param[2];
  int32_t n_roi;
  int32_t reserved;
  const double* roi_time;
  const double* roi_mdot;
  const double* roi_vel;
  /* injection state, public in the reference (SprayJet.H:258-266) */
  double num_ppp;        /* m_numPPP; < 0: set by the first injection */
  double sum_inj_mass;   /* m_sumInjMass: mass held over, -1 = none */
  double sum_inj_time;   /* m_sumInjTime */
  double min_parcel;     /* m_minParcel, initially 1 */
  double total_inj_mass; /* m_totalInjMass */
  double total_inj_time; /* m_totalInjTime */
  uint64_t seed;
  int64_t n_injections;
} pelemp_spray_jet;

/* The constructor checks of SprayJet (SprayJet.cpp:28-36,55-57,76-85,
 * 123-134; SprayJet.H:71-79,121-139): angles in range, a jet centre inside
 * the domain, mass fractions summing to 1, non-negative rates; normalises
 * norm.  *avg_dia receives get_avg_dia() of the distribution. */
int pelemp_spray_jet_check(const pelemp_spray_ctx* ctx, pelemp_spray_jet* jet,
                           double* avg_dia);

/* SprayJet::readROI (SprayJet.cpp:194-256): parses a rate-of-injection file
 * (a header line, then "time[ms];mdot" lines) into n points of time [s],
 * mdot and velocity mdot / (rho_part * area * cd); mass_conv is 1e-3 for
 * the PeleLMeX flavour (g/s -> kg/s) and 1 for PeleC.  *n is in: capacity,
 * out: points.  Also returns the start / end time and the largest jet
 * velocity. */
int pelemp_spray_read_roi(const char* path, double rho_part, double cd,
                          double jet_dia, double mass_conv, int32_t* n,
                          double* time, double* mdot, double* vel,
                          double* start_time, double* end_time,
                          double* max_jet_vel);

/* SprayParticleContainer::sprayInjection (SprayInjection.H:10-180): injects
 * the droplets of one jet for a step of sim_dt at `time`, on the device,
 * and appends them to the context (ids continue past the largest id).  The
 * jet's state (num_ppp, held-over mass and time, min_parcel, totals) is
 * updated as the reference updates its SprayJet.  The droplet loop (draw a
 * candidate, keep it if its diameter exceeds the minimum, until the
 * injected mass reaches the target) runs with the candidates drawn in
 * parallel and the mass summed in candidate order.  *n_injected: droplets
 * added.  PELEMP_E_BAD_INJECTION for a droplet outside the domain (the
 * reference's "Bad injection particle" abort), a candidate limit reached
 * before the mass, or a liquid density at the jet temperature that is not
 * positive (the reference loop would not end).  Waits for `stream`. */
int pelemp_spray_inject(pelemp_spray_ctx* ctx, pelemp_spray_jet* jet,
                        double time, double sim_dt, int64_t* n_injected,
                        void* stream);

/* SprayParticleContainer::uniformSprayInit (SprayInjection.H:198-348), one
 * rank: num_part[0]*num_part[1]*num_part[2] droplets on a lattice of the
 * domain, (i + 0.5) * ProbLength / num_part, with the given velocity,
 * diameter, temperature, mass fractions and parcel size, generated on the
 * device and appended.  Ids continue past the largest id in lattice order
 * (x fastest); the reference appends them in reverse order, and so does
 * this call. */
int pelemp_spray_uniform_init(pelemp_spray_ctx* ctx, const int32_t num_part[3],
                              const double vel[3], double dia, double T,
                              const double* Y, double num_ppp);

/* ---- On-disk formats (SprayIO.cpp) ---- */

/* injection_data.log of a checkpoint (SprayIO.cpp:49-97): the number of
 * jets, then per jet its name, num_ppp, held-over mass and time,
 * min_parcel, total mass and time (15 significant digits). */
int pelemp_spray_write_injection_log(const char* path, int32_t njets,
                                     const char* const* names,
                                     const pelemp_spray_jet* jets);
/* PostInitRestart (SprayIO.cpp:100-160): reads such a file and sets the
 * state of every jet whose name matches; jets not in the file keep their
 * state.  Duplicate or empty names are an error. */
int pelemp_spray_read_injection_log(const char* path, int32_t njets,
                                    const char* const* names,
                                    pelemp_spray_jet* jets);

/* The ASCII particle file written with plot files (SprayIO.cpp:35-47,
 * AMReX ParticleContainer::WriteAsciiFile): the particle count, then one
 * line per particle: position, id, cpu, the real components in SprayComps
 * order; 15 significant digits. */
int pelemp_spray_write_ascii(pelemp_spray_ctx* ctx, const char* path);

/* Particle ids.  AMReX names a particle by (id, cpu): id < 2^39 and the
 * rank that made it, cpu < 2^24.  This library holds one 64-bit id per
 * droplet, id + cpu * 2^39, which is the plain id for droplets made on
 * rank 0; the on-disk formats below split it back into (id, cpu). */

/* Checkpoint of the droplets in the context, SprayParticleIO
 * (SprayIO.cpp:6-34), which writes them with AMReX's
 * ParticleContainer::Checkpoint(dir, "particles", real_comp_names) in
 * AMReX's particle checkpoint layout: dir/particles/Header
 * ("Version_Two_Dot_One_double", the SprayIO.cpp:9-33 component names,
 * particle count, next id, grid list), dir/particles/Level_0/Particle_H
 * (the box array: here the level's domain box) and
 * dir/particles/Level_0/DATA_00000 (per grid: the packed idcpu words as
 * int32 pairs, then position and components as doubles).  fuel_names name
 * the spray_mf_<fuel> components (a missing name is the fuel's index). */
int pelemp_spray_checkpoint(pelemp_spray_ctx* ctx, const char* dir,
                            int32_t nfuel_names, const char* const* fuel_names);
/* Restart(dir, "particles") (SpraySetup.cpp:320-325): the level-0
 * particles of such a checkpoint, from every grid and DATA file the header
 * lists, written by this library or by AMReX (Version_Two_Dot_One or
 * _Two_Dot_Zero, _double or _single); the next id becomes the larger of the
 * header's and one past the largest id.  PELEMP_E_BAD_CONFIG when the
 * component counts differ from the context's (10 + nf reals, no ints). */
int pelemp_spray_restart(pelemp_spray_ctx* ctx, const char* dir);

/* The same layout on host arrays, no device involved (a converter, or a
 * host that holds its particles itself).  pos is [n][3], rdata [n][nreal],
 * ids as above.  nboxes boxes (box_lo/box_hi [nboxes][3] cell indices) form
 * the level's box array; with more than one, each particle goes to the
 * first box that holds its cell, floor((pos - plo) / dx), and a particle in
 * none is PELEMP_E_ARG.  plo and dx may be NULL for one box. */
int pelemp_particles_write_checkpoint(const char* dir, const char* name,
                                      int64_t n, const double* pos,
                                      const double* rdata, const int64_t* ids,
                                      int32_t nreal,
                                      const char* const* real_names,
                                      int64_t next_id, int32_t nboxes,
                                      const int32_t* box_lo,
                                      const int32_t* box_hi,
                                      const double plo[3], const double dx[3]);
/* Particle count of one level, real-component count and next id of a
 * checkpoint at dir/name. */
int pelemp_particles_checkpoint_info(const char* dir, const char* name,
                                     int32_t level, int64_t* n,
                                     int32_t* nreal, int64_t* next_id);
/* Reads one level's particles (grid by grid, in file order; particles with
 * a negative AMReX id are skipped) into arrays of `capacity` particles. */
int pelemp_particles_read_checkpoint(const char* dir, const char* name,
                                     int32_t level, int64_t capacity,
                                     int64_t* n, double* pos, double* rdata,
                                     int64_t* ids);

/* The droplet update for one level, SprayParticleContainer::updateParticles
 * (SprayParticles.cpp:142-502) with the same argument meanings; `state` and
 * `source` are device FABs owned by the caller.  The spray sources are
 * accumulated into `source` (+=), as the reference's atomics do; each
 * cell's contributions are summed in a fixed order (cell-sorted droplet
 * order), so results are reproducible run to run.  Asynchronous on
 * `stream` (a hipStream_t, NULL = default stream); device-side check
 * failures are reported by pelemp_spray_sync. */
int pelemp_spray_update(pelemp_spray_ctx* ctx, const pelemp_fab_view* state,
                        pelemp_fab_view* source, double flow_dt, double time,
                        int32_t state_ghosts, int32_t source_ghosts,
                        int32_t isVirt, int32_t isGhost, int32_t do_move,
                        double spray_cfl_lev, void* stream);

/* pelemp_spray_update with the gas state in single precision: the state
 * MultiFab of a solver that keeps its fields in FP32 (updateParticles reads
 * the state as Real; here Real = float for the state only).  The state is
 * widened to double exactly where the droplet model computes in FP64, so the
 * FP64 form gives what pelemp_spray_update gives for the same values stored
 * as doubles.  The sources stay FP64.  PeleLMeX flavour. */
int pelemp_spray_update_f32gas(pelemp_spray_ctx* ctx,
                               const pelemp_fab_view_f32* state,
                               pelemp_fab_view* source, double flow_dt,
                               double time, int32_t state_ghosts,
                               int32_t source_ghosts, int32_t isVirt,
                               int32_t isGhost, int32_t do_move,
                               double spray_cfl_lev, void* stream);

/* How pelemp_spray_update writes the spray source.  No counterpart in the
 * reference, whose callers zero the source MultiFab (tmpSource.setVal(0.),
 * SprayParticles.H:294) and let the update add into it:
 *   PELEMP_SOURCE_ACCUMULATE (default): the update adds (+=) as the
 *     reference does;
 *   PELEMP_SOURCE_OVERWRITE: the first particle sub-cycle of a call writes
 *     every domain cell of the source (components the coupling flags switch
 *     on) with 0 + its sum, which is what adding into a zeroed FAB gives, so
 *     the caller's zero fill (and the update's read of the old values) is
 *     not needed.  Ghost cells of the source are left as they are (deposits
 *     redirected to them by an EB still add). */
#define PELEMP_SOURCE_ACCUMULATE 0
#define PELEMP_SOURCE_OVERWRITE 1
int pelemp_spray_set_source_mode(pelemp_spray_ctx* ctx, int32_t mode);

/* Waits for `stream` and returns the sticky device-side error recorded by
 * earlier updates (PELEMP_E_STATE_BOX, _SRC_BOX, _LEFT_DOMAIN, ...), then
 * clears it.  pelemp_spray_update itself does not wait for the device. */
int pelemp_spray_sync(pelemp_spray_ctx* ctx, void* stream);

/* moveKickDrift / moveKick wrappers (SprayParticles.cpp:37-87). */
int pelemp_spray_move_kick_drift(pelemp_spray_ctx* ctx,
                                 const pelemp_fab_view* state,
                                 pelemp_fab_view* source, double dt,
                                 double time, int32_t isVirt, int32_t isGhost,
                                 int32_t state_ghosts, int32_t source_ghosts,
                                 int32_t do_move, double spray_cfl_lev,
                                 void* stream);

/* Particle-CFL time step, SprayParticleContainer::estTimestep
 * (SprayParticles.cpp:89-140): the minimum over the droplets of
 * cfl / max_d(|u_d| / dx_d) (1e50 for a droplet at rest), then, when
 * max_jet_vel > 0, the minimum with cfl dx_0 / max_jet_vel (the bound the
 * reference takes from its spray jets, :125-131; max_jet_vel is the largest
 * jet velocity).  *dt is -1 when fixed_parts is set and the largest double
 * when there are no droplets and no jet bound.  Waits for `stream`. */
int pelemp_spray_est_timestep(pelemp_spray_ctx* ctx, double cfl,
                              double max_jet_vel, double* dt, void* stream);

/* SprayParticleContainer::transferSource (SprayParticles.H:250-295) for one
 * box covering the domain: the ghost cells of tmp_source are added to the
 * valid cells they are periodic images of (SumBoundary), tmp_source is added
 * to act_source over the domain grown by min(act_source ghosts,
 * source_ghosts) -- every component when both have the same number of
 * components (MultiFab::Add), otherwise through the spray component map of
 * addSpraySrc (SprayParticles.H:299-319: momentum, density, energy, fuel
 * species at specIndx + dep_indx) -- and tmp_source is zeroed.  Device
 * FABs owned by the caller; asynchronous on `stream`. */
int pelemp_spray_transfer_source(pelemp_spray_ctx* ctx, int32_t source_ghosts,
                                 pelemp_fab_view* tmp_source,
                                 pelemp_fab_view* act_source, void* stream);

/* SprayParticleContainer::computeDerivedVars (SprayDerive.cpp:9-155): adds
 * the per-cell spray diagnostics to the FAB mf_var (covering the domain)
 * from component start_indx on, in the order of SpraySetup.cpp:213-239:
 * spray_mass, spray_density, spray_num, spray_vol, spray_surf_area,
 * spray_vol_frac, d10, d32, wall_film_hght, wall_film_mass, spray_temp,
 * num_parcels, spray_x/y/z_vel, and with derive_plot_species == 1 and
 * several fuels spray_mass_<fuel> per fuel; then the protected divides
 * (d10, d32, temperature, velocity).  Each cell's droplets are summed in
 * cell-sorted order.  pelemp_spray_derive_count gives the number of
 * variables (NumDeriveVars, SprayParticles.H:366-369).  Asynchronous on
 * `stream`. */
int pelemp_spray_derive_count(const pelemp_spray_ctx* ctx,
                              int32_t derive_plot_species, int32_t* n);
int pelemp_spray_compute_derived(pelemp_spray_ctx* ctx, pelemp_fab_view* mf_var,
                                 int32_t start_indx, int32_t derive_plot_species,
                                 void* stream);

/* Active droplets at the start of each update pass since the previous call
 * (at most 4096 passes are logged); waits for the device.  The sum is the
 * droplet-step count of those passes. */
int pelemp_spray_pass_counts(pelemp_spray_ctx* ctx, int64_t* out, int32_t max,
                             int32_t* n_out);

/* Kernel timing: when enabled, the update records HIP events around its
 * kernels; the accumulated milliseconds and launch counts are returned per
 * kernel slot (0 update, 1 sort, 2 film thickness pass, 3 secondary-droplet
 * pass). */
int pelemp_spray_set_timing(pelemp_spray_ctx* ctx, int32_t enable);
int pelemp_spray_get_timing(pelemp_spray_ctx* ctx, double* ms, int64_t* launches,
                            int32_t nslots);

/* Bytes of device memory held by the context. */
int pelemp_spray_device_bytes(const pelemp_spray_ctx* ctx, int64_t* bytes);

/* ---- Learned surrogate of the substep rate evaluation -------------------
 * No counterpart in the reference: the surrogate replaces the film-property
 * / Spalding-number evaluation inside the substep of calculateSpraySource
 * (Drag.H:213-347) for PELEMP_FORM_SURROGATE.  Network: 8 inputs -> 4 hidden
 * layers of 64 (hard swish) -> 3 outputs; FP32 master parameters, 13251 floats,
 * per linear layer L = 0..4 the weights W_L[out][in] (row-major) followed by
 * the biases b_L[out]; fan-in 8,64,64,64,64, fan-out 64,64,64,64,3.
 *
 * Inputs (before normalisation): ln d [m], T_d [K], ln(|u_gas - u_d| +
 * 0.01 m/s), T_gas [K], ln p [Pa], fuel mass fraction of the film gas (1/3
 * rule between the surface vapour and the gas), liquid mass fractions Y_l0,
 * Y_l1.  Outputs (after de-normalisation):
 *   z0 = ln(sum_k rho D_k)   film diffusivity of the vapour  -> dm/dt
 *   z1 = ln(lambda Nu*)      heat-transfer group incl. the B_T fixed point
 *                            -> Q = pi d (T_gas - T_d) e^z1, dT_d/dt
 *   z2 = ln(mu_skin)         film viscosity -> Re, drag correction C_D Re/24
 * The surface vapour state, B_M, latent heats, film density, Re, Sc, the
 * piecewise drag coefficient and the Sherwood number stay exact closed
 * forms, so dm/dt = -pi d Sh*(Re, Sc, B_M) e^z0 ln(1 + B_M). */
#define PELEMP_SURR_NPARAM 13251
#define PELEMP_SURR_NIN 8
#define PELEMP_SURR_NOUT 3

typedef struct pelemp_surrogate_norm {
  float in_center[8];
  float in_scale[8];   /* x_n = (x - in_center) * in_scale */
  float out_mean[3];
  float out_std[3];    /* z = out_mean + out_std * y_n */
  /* training hull; droplets outside it take the plain FP32 path */
  float d_min, d_max;         /* m */
  float Tg_min, Tg_max;       /* K */
  float p_min, p_max;         /* Pa */
  float urel_max;             /* m/s */
  float Yfg_max;              /* fuel vapour mass fraction of the gas */
  float Td_min;               /* K */
  float Td_frac_max;          /* T_d <= Td_frac_max * sum_k Y_l,k T_boil,k */
  float BM_min;               /* Spalding number floor */
  float indicator_max;        /* error indicator: flow_dt / tau_{drag,mass,T}
                                 of the predicted rates must stay below it */
} pelemp_surrogate_norm;

/* Training samples generated by the plain FP64 rate evaluation (the same
 * device code as the FP64 update) at n counter-based random points of the
 * hull: x[n*8] raw (un-normalised) inputs, t[n*3] targets z0..z2, w[n]
 * 1 for a usable sample, 0 otherwise (surface vapour below the gas, B_M
 * under the floor).  Device pointers; asynchronous on `stream`. */
int pelemp_surrogate_sample(int32_t mech, const pelemp_spray_data* sd,
                            int32_t apply_latent_shift,
                            const pelemp_surrogate_norm* hull, int64_t n,
                            uint64_t seed, float* x, float* t, float* w,
                            void* stream);

/* FP32 forward pass of the network on n raw inputs (normalised inside with
 * `norm`): y[n*3] normalised outputs.  quant != 0 rounds the weights and each
 * layer's input to bfloat16 (round to nearest even), as the MFMA inference
 * does.  Device pointers. */
int pelemp_mlp_forward(const float* params, const pelemp_surrogate_norm* norm,
                       int32_t quant, int64_t n, const float* x, float* y,
                       void* stream);

/* Loss L = (1/n) sum_i w_i |y_i - (t_i - out_mean)/out_std|^2 over one batch
 * of n <= 65536 samples and its gradient with respect to the 13251
 * parameters (straight-through for the bfloat16 rounding when quant != 0).
 * grad: device, loss: host.  Waits for the device. */
int pelemp_mlp_loss_grad(const float* params, const pelemp_surrogate_norm* norm,
                         int32_t quant, int64_t n, const float* x,
                         const float* t, const float* w, float* grad,
                         double* loss, void* stream);

/* pelemp_mlp_loss_grad without any device allocation.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training or serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and waiting behaviour as
 * pelemp_mlp_loss_grad, but the batch workspace (layer inputs, deltas,
 * per-sample losses, per-chunk gradient partials, loss accumulator) is one
 * statically sized device array of PELEMP_MLP_NOALLOC_SCRATCH_BYTES per
 * device, held until the process ends, so n may not exceed
 * PELEMP_MLP_NOALLOC_MAX_BATCH (PELEMP_E_ARG otherwise).  Every call on a
 * device shares that device's workspace: calls that overlap on one device
 * (several host threads or streams) race and must be serialised by the
 * caller. */
#define PELEMP_MLP_NOALLOC_MAX_BATCH 4096
#define PELEMP_MLP_NOALLOC_SCRATCH_BYTES 10298112
int pelemp_mlp_loss_grad_noalloc(const float* params,
                                 const pelemp_surrogate_norm* norm,
                                 int32_t quant, int64_t n, const float* x,
                                 const float* t, const float* w, float* grad,
                                 double* loss, void* stream);

typedef struct pelemp_mlp_train_cfg {
  int32_t epochs;
  int32_t batch;        /* <= 65536 */
  int32_t quant;        /* train through the bfloat16 rounding */
  int32_t verbose;      /* print the loss of every epoch to stderr */
  double lr;            /* Adam step size, cosine-decayed to lr_final */
  double lr_final;
  double beta1, beta2, eps;
  uint64_t seed;        /* batch composition per epoch */
} pelemp_mlp_train_cfg;

/* Adam training of `params` (device, in/out) on n samples, with the
 * project's forward / backward / gradient-reduction kernels.  epoch_loss
 * (host, cfg->epochs entries) receives the mean training loss per epoch. */
int pelemp_mlp_train(float* params, const pelemp_surrogate_norm* norm,
                     int64_t n, const float* x, const float* t, const float* w,
                     const pelemp_mlp_train_cfg* cfg, double* epoch_loss,
                     void* stream);

/* Inference with the bfloat16 MFMA kernel (the code the surrogate update
 * runs): params on the host, x/y on the device as in pelemp_mlp_forward. */
int pelemp_mlp_infer_bf16(const float* params_host,
                          const pelemp_surrogate_norm* norm, int64_t n,
                          const float* x, float* y, void* stream);

/* Install a trained network (host parameters) in a context created with
 * PELEMP_FORM_SURROGATE. */
int pelemp_spray_set_surrogate(pelemp_spray_ctx* ctx, const float* params_host,
                               const pelemp_surrogate_norm* norm);

/* Droplet-steps taken by the surrogate and by the plain fallback since the
 * previous call (waits for the device). */
int pelemp_spray_surrogate_counts(pelemp_spray_ctx* ctx, int64_t* n_surrogate,
                                  int64_t* n_fallback);

/* Self-test hook: D = A (32x16) * B (16x32) with one bfloat16 MFMA
 * instruction; A, B row-major bfloat16 bit patterns, D row-major float.
 * Shows the operand layout the surrogate kernel assumes. */
int pelemp_mfma_probe(const uint16_t* a_host, const uint16_t* b_host,
                      float* d_host);

/* Self-test hook: evaluate the library's FP64 elementary functions on the
 * device (fn: 0 exp, 1 log, 2 log1p, 3 pow, 4 cbrt, 5 sin, 6 cos, 7 atan2,
 * 8 acos, 9 sqrt, 10 div) over n host values; used to show the device and
 * host implementations agree bit for bit.  Not part of the reference. */
int pelemp_math_eval(int32_t fn, int64_t n, const double* x, const double* y,
                     double* out);

/* Device-memory, stream and timer plumbing for callers that have no HIP
 * runtime of their own (the Python harness).  No reference counterpart. */
int pelemp_device_count(int32_t* n);
int pelemp_device_malloc(void** ptr, int64_t bytes);
int pelemp_device_free(void* ptr);
int pelemp_memcpy_htod(void* dst, const void* src, int64_t bytes);
int pelemp_memcpy_dtoh(void* dst, const void* src, int64_t bytes);
int pelemp_memcpy_dtod(void* dst, const void* src, int64_t bytes);
int pelemp_memset_device(void* ptr, int32_t value, int64_t bytes);
/* FP64 words [ptr, ptr + n_words) set to 0 by a kernel on `stream`, in
 * stream order (what a caller's MultiFab::setVal(0.) of the spray source
 * does between steps); a kernel, not a memset, so it may sit in a captured
 * graph. */
int pelemp_fab_set_zero(double* ptr, int64_t n_words, void* stream);
int pelemp_device_synchronize(void);
int pelemp_stream_create(void** stream);
int pelemp_stream_destroy(void* stream);
int pelemp_stream_synchronize(void* stream);
/* Event pair timer: record start (which = 0) / stop (which = 1) on a stream,
 * then read the elapsed milliseconds (waits for the stop event). */
int pelemp_timer_create(void** timer);
int pelemp_timer_record(void* timer, int32_t which, void* stream);
int pelemp_timer_elapsed(void* timer, double* ms);
int pelemp_timer_destroy(void* timer);

/* Last error message of the calling thread. */
const char* pelemp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PELEMP_SPRAY_H */
