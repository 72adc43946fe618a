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

/* Status codes (reference: amrex::Abort sites listed per code). */
#define PELEMP_OK 0
#define PELEMP_E_STATE_BOX 1   /* SprayInterpolation.H:121-124 */
#define PELEMP_E_SRC_BOX 2     /* SprayParticles.cpp:420-425 */
#define PELEMP_E_LEFT_DOMAIN 3 /* SprayParticles.cpp:347-351 */
#define PELEMP_E_BAD_CONFIG 4  /* SpraySetup.cpp:92-100,152-183 */
#define PELEMP_E_HIP 5         /* HIP runtime error */
#define PELEMP_E_ARG 6         /* invalid argument / size */
#define PELEMP_E_NOMEM 7       /* device allocation failed */

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
int pelemp_spray_destroy(pelemp_spray_ctx* ctx);

/* The SprayData actually in use (after the latent shift). */
int pelemp_spray_get_data(const pelemp_spray_ctx* ctx, pelemp_spray_data* out);

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
 * (SprayParticles.cpp:89-140, without the jet-velocity bound).  Returns -1
 * in *dt when fixed_parts is set or there are no particles. */
int pelemp_spray_est_timestep(pelemp_spray_ctx* ctx, double cfl, double* dt,
                              void* stream);

/* Active droplets at the start of each update pass since the previous call
 * (at most 4096 passes are logged); waits for the device.  The sum is the
 * droplet-step count of those passes. */
int pelemp_spray_pass_counts(pelemp_spray_ctx* ctx, int64_t* out, int32_t max,
                             int32_t* n_out);

/* Kernel timing: when enabled, the update records HIP events around its
 * kernels; the accumulated milliseconds and launch counts are returned per
 * kernel slot (0 update, 1 sort, 2 reorder). */
int pelemp_spray_set_timing(pelemp_spray_ctx* ctx, int32_t enable);
int pelemp_spray_get_timing(pelemp_spray_ctx* ctx, double* ms, int64_t* launches,
                            int32_t nslots);

/* Bytes of device memory held by the context. */
int pelemp_spray_device_bytes(const pelemp_spray_ctx* ctx, int64_t* bytes);

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
