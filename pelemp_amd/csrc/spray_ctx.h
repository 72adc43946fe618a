// spray_ctx.h -- device-resident state of one spray level and the launcher
// interface shared by the host translation unit and the per-form kernel
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "layout.h"
#include "pelemp_spray.h"

namespace pelemp {

// Per-update parameters computed on the host (SprayParticles.cpp:183-229).
struct PassParams {
  double flow_dt;
  double sub_dt;
  double cur_coef;  // -(1/V) sub_dt / flow_dt
  int do_move;
  int src_overwrite;  // write the domain cells of the source (first sub-cycle of an overwrite-mode call)
  int is_ghost;
  int state_ghosts;   // state box = domain grown by this many cells (negative: the state FAB's box)
  int source_ghosts;  // source box = domain grown by this many cells (negative: the source FAB's box)
  int at_bounds;
  int branches;        // film / breakup / splash kernel variant
  int make_new_drops;  // SprayParticles.cpp:298-299
  int do_splash_box;   // splash on (walls present)
  const pelemp_fab_view* state;  // box and components of the state (data is null for an FP32 state)
  const float* state32;          // FP32 state data, or null
  pelemp_fab_view* source;
};

struct Ctx;

// Launch one update pass of the given form (implemented in the per-form
// translation units).  Reads buffer ctx.cur via ctx.perm, writes buffer
// 1-ctx.cur and ctx.key.
int launch_update_fp64(Ctx& c, const PassParams& p, hipStream_t s);
int launch_update_fp32(Ctx& c, const PassParams& p, hipStream_t s);
int launch_update_surr(Ctx& c, const PassParams& p, hipStream_t s);
// PeleC flavour: the state box converted to SI (rho, Y, vel, T) into
// ctx.pelec_conv, once per update call, before the passes.
int launch_pelec_convert(Ctx& c, const pelemp_fab_view& st, hipStream_t s);

// Mechanism tables (spray_host.hip).
int mech_ns(int mech);
void mech_mw(int mech, std::vector<double>& mw);
void mech_h(int mech, double T, double* h);

struct Ctx {
  int mech = 0;
  int form = 0;
  int nf = 1;
  int ns = 3;
  pelemp_spray_data sd{};
  pelemp_spray_comps comps{};
  pelemp_geom geom{};
  std::vector<double> mw;  // species molecular weights

  int64_t cap = 0;
  int ncell[3] = {0, 0, 0};
  int ntile[3] = {0, 0, 0};
  int ncells = 0;
  int ntiles = 0;
  bool hot_brk = false;  // breakup state (BM1, BM2, N0) carried in the hot SoA
  int nhot = 0;          // hot fields per droplet
  size_t wsize = 8;      // bytes per hot field value

  void* fbuf[2] = {nullptr, nullptr};
  double* mbuf[2] = {nullptr, nullptr};  // FP32 form: liquid species masses
  int32_t* slot[2] = {nullptr, nullptr};
  int cur = 0;
  int32_t* perm = nullptr;
  int32_t* cstart = nullptr;  // ncells + 1
  int32_t* cnt = nullptr;
  int32_t* arrival = nullptr;
  int2* long_list = nullptr;  // (start, length) of perm segments too long for one thread
  int32_t* n_long = nullptr;  // entries of long_list (in scalars)
  bool counters_dirty = true;  // mov_lo / stay_cnt / mov_hi not cleared since the last update pass
  int32_t* mov_lo = nullptr;    // per cell: droplets arriving from lower keys
  int32_t* mov_hi = nullptr;    // per cell: droplets arriving from higher keys
  int32_t* stay_cnt = nullptr;  // per cell: droplets that stayed
  uint32_t* key = nullptr;  // key of each storage position of buffer cur
  int32_t* n_entries = nullptr;  // entries in buffer cur (device scalar)
  int32_t* block_sum = nullptr;
  int32_t* total = nullptr;
  int32_t* err = nullptr;
  int32_t* scalars = nullptr;  // n_entries, total, err, n_long, split counters[3] in one allocation
  // heavy tiles (more than kSegDroplets droplets): split lists of the next
  // pass and the per-segment partials; capacities from the droplet capacity
  SplitDev sp{};
  int32_t split_items_cap = 0;
  int32_t split_list_cap = 0;
  int64_t split_slots_cap = 0;
  double* dscratch = nullptr;  // reduction scratch
  // cold per-slot data (indexed by slot)
  int64_t* cold_id = nullptr;
  double* cold_n0 = nullptr;
  double* cold_bm1 = nullptr;
  double* cold_bm2 = nullptr;
  int64_t nslots = 0;  // slots handed out
  int64_t next_id = 1; // id of the next new droplet (without the branches;
                       // with them the device counter sb_ids[0] holds it)
  int64_t host_count = 0;  // active droplets after the last sort (host copy, when known)
  // active droplets at the start of every pass since the last read
  static constexpr int kCountLog = 4096;
  int32_t* count_log = nullptr;
  int64_t npass = 0;

  // branches (FP64 form with breakup / splash / films): container
  // parameters, film thickness per cell, secondary-droplet records and
  // the counters of new droplets (device scalars: next id, next slot)
  pelemp_spray_params params{};
  int src_mode = PELEMP_SOURCE_ACCUMULATE;
  bool branches = false;
  // PeleC flavour: CGS state (rho, rho*u, rho*E, rho*Y, T), sources, droplets
  // and geometry at the boundary; SI inside (src_scale: SI -> CGS sources)
  bool pelec = false;
  double src_scale[5 + PELEMP_MAX_FUEL] = {};
  double* pelec_conv = nullptr;  // the state box in SI (5 + NS fields)
  size_t pelec_conv_bytes = 0;
  double* wf = nullptr;
  int32_t* sb_type = nullptr;
  double* sb_data = nullptr;
  int32_t* sb_count = nullptr;   // children per position
  int64_t* sb_offset = nullptr;  // their exclusive scan
  int64_t* sb_block = nullptr;   // scan block sums over positions
  int32_t* sb_parent = nullptr;  // record position of each child of the pass
  int32_t* sb_pblock = nullptr;  // its max-scan block maxima
  int64_t* sb_ids = nullptr;    // [next id, next slot, children of the last pass]
  int32_t* n_new = nullptr;     // entries after the children pass
  int64_t slot_cap = 0;         // capacity of the cold (per-slot) arrays

  // embedded boundary (caller-owned device arrays): the view, whether the
  // box has cut or covered cells, the domain cells next to them (linear,
  // x fastest) and the per-droplet redirected deposits of a pass
  bool eb_set = false;
  bool eb_on = false;
  pelemp_eb_view eb{};
  int32_t* eb_cells = nullptr;
  int32_t n_eb_cells = 0;
  uint32_t* eb_rkey = nullptr;
  double* eb_rval = nullptr;

  // surrogate form: packed network (SurrWeights), its normalisation, and the
  // droplet-step counters [surrogate, fallback]
  void* surr_w = nullptr;
  bool has_surr = false;
  pelemp_surrogate_norm surr_norm{};
  unsigned long long* surr_cnt = nullptr;

  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double t_ms[4] = {0, 0, 0, 0};
  int64_t t_n[4] = {0, 0, 0, 0};

  size_t bytes = 0;
};

hipEvent_t ctx_event(Ctx& c);
void ctx_time_begin(Ctx& c, int slot, hipStream_t s, hipEvent_t* ev);
void ctx_time_end(Ctx& c, int slot, hipStream_t s, hipEvent_t ev0);
void set_last_error(const std::string& msg);

}  // namespace pelemp
