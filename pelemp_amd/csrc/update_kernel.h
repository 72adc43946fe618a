// update_kernel.h -- fused per-tile droplet update for gfx950.
//
// Replaces the per-particle body of SprayParticleContainer::updateParticles
// (SprayParticles.cpp:324-482) for the drag/evaporation path: gather of the
// gas state (trilinear, SprayInterpolation.H:99-199), boiling temperature,
// calculateSpraySource (Drag.H:137-441), source deposition into the
// containing cell (SprayParticles.cpp:419-441) and the drift with bounds
// check and periodic wrap (SprayParticles.cpp:443-479).
//
// Work decomposition.  Cells are grouped in 4x4x4 tiles; droplets are kept
// sorted by (tile, cell) so the droplets of one tile are contiguous.  One
// 256-thread workgroup owns one tile:
//   1. it stages the gas state of the tile plus a one-cell halo (6^3 cells)
//      in LDS, converting rho*Y_k to Y_k once per cell;
//   2. it walks the tile's droplets in chunks of 256 (one droplet per
//      thread, storage position via perm), runs the droplet physics in
//      registers and writes the droplet back at its sorted position;
//   3. after each chunk, the per-droplet source contributions are reduced
//      per cell in droplet order by the threads that own (cell, component)
//      accumulators, so every cell's sum has a fixed association order;
//   4. finally it adds each cell's sums to the source FAB (owner computes,
//      no atomics).
// Workgroups are mapped XCD-contiguously: consecutive tiles (which share
// halo cells) run on the same XCD and hit the same L2.
//
// Surrogate form: the first substep's rates come from the learned network
// (surrogate.h, surrogate_mlp.h) for droplets inside its hull and below its
// error indicator; every wavefront evaluates the network for its 64
// droplets with MFMA between the gather and the update, and the other
// droplets take the plain FP32 substep.  The packed network is staged in
// LDS once per workgroup, and the workgroups loop over tiles (grid sized to
// the machine) so the staging is amortised.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "branches.h"
#include "droplet_physics.h"
#include "eb.h"
#include "layout.h"
#include "surrogate_mlp.h"

namespace pelemp {

constexpr int kHaloW = kTile + 2;
constexpr int kHaloCells = kHaloW * kHaloW * kHaloW;

// The staged gas state in LDS is cell-major: the 5 + NS values of halo cell
// h (rho, Y[NS], vel[3], T) at gas[h * kGasStride + f], the stride padded to
// 16 bytes so that a droplet's trilinear gather reads each corner cell with
// a few wide LDS loads instead of one load per value.
template <class R, int NS>
constexpr int gas_stride() {
  constexpr int per16 = 16 / static_cast<int>(sizeof(R));
  return ((5 + NS + per16 - 1) / per16) * per16;
}
constexpr int kUpdThreads = 256;

// Error codes written to the sticky error word (first failure wins).
__device__ inline void set_err(int* err, int code) { atomicCAS(err, 0, code); }

struct FabDev {
  double* data;
  const float* data32;  // a gas state kept in FP32 (data is then unused)
  int lo[3];
  int n[3];
  long long cstride;  // cells per component
};

// The gas state of cell c of a state FAB with elements of type T (the
// PeleLMeX layout): v = rho, Y[NS] = rho Y / rho, vel[3], T, in double (an
// FP32 state is widened exactly).
template <int NS, class T>
__device__ __attribute__((always_inline)) inline void state_cell(const T* __restrict__ S, long long cs, long long c,
                                                                 int c_rho, int c_spec, int c_vel, int c_T,
                                                                 double* v) {
  const double rho = static_cast<double>(S[c_rho * cs + c]);
  const double inv_rho = 1. / rho;
  v[0] = rho;
#pragma unroll
  for (int n = 0; n < NS; ++n) v[1 + n] = static_cast<double>(S[(c_spec + n) * cs + c]) * inv_rho;
#pragma unroll
  for (int q = 0; q < 3; ++q) v[1 + NS + q] = static_cast<double>(S[(c_vel + q) * cs + c]);
  v[4 + NS] = static_cast<double>(S[c_T * cs + c]);
}

// The same for a FAB in double or in single precision (one uniform branch
// per cell).
template <int NS>
__device__ __attribute__((always_inline)) inline void state_cell(const FabDev& f, long long c, int c_rho, int c_spec,
                                                                 int c_vel, int c_T, double* v) {
  if (f.data32)
    state_cell<NS>(f.data32, f.cstride, c, c_rho, c_spec, c_vel, c_T, v);
  else
    state_cell<NS>(f.data, f.cstride, c, c_rho, c_spec, c_vel, c_T, v);
}

// Normalisation of the surrogate inputs / outputs as used in the kernel.
struct SurrKNorm {
  float c[kSurrIn], s[kSurrIn], mean[kSurrOut], sd[kSurrOut];
};

template <class R, int NF>
struct UpdateArgs {
  const R* __restrict__ fin;
  R* __restrict__ fout;
  const double* __restrict__ min;  // FP32 form: species masses (nf x cap)
  double* __restrict__ mout;
  long long cap;
  const int32_t* __restrict__ perm;
  const int32_t* __restrict__ cstart;
  const int32_t* __restrict__ slot_in;
  int32_t* __restrict__ slot_out;
  uint32_t* __restrict__ key_out;
  // incremental re-sort bookkeeping (see sort_kernels.h)
  int32_t* __restrict__ arrival_out;
  int32_t* __restrict__ mov_lo;
  int32_t* __restrict__ mov_hi;
  int32_t* __restrict__ stay_cnt;
  FabDev st;
  int c_rho, c_vel, c_spec, c_T;
  FabDev src;
  int c_rsrc, c_msrc, c_ssrc, c_esrc;
  double plo[3], phi[3], dx[3], dxi[3];
  int dom_lo[3];
  int ncell[3];
  int ntile[3];
  int bndry_lo[3], bndry_hi[3];
  FuelConst<R, NF> fc;
  R flow_dt;
  R sub_dt;
  double cur_coef;  // -(1/V) sub_dt / flow_dt
  R sub_dt_dxi[3];  // FP32 form: sub_dt / dx per direction
  int do_move;
  int src_overwrite;  // 1: the domain cells of the source get this pass's sums (written, not added)
  int is_ghost;
  int st_g;   // state ghost width: the gather may read cells up to st_g outside the domain (negative: the FAB)
  int src_g;  // source ghost width: a ghost droplet is deleted when its cell leaves the domain grown by src_g
  int* err;
  // heavy tiles: the split lists of this pass and the per-segment partials
  SplitDev sp;
  // surrogate form
  const SurrWeights* __restrict__ surr_w;
  SurrKNorm sn;
  pelemp_surrogate_norm hull;
  unsigned long long* surr_cnt;  // [surrogate, fallback] droplet-steps
  // branches (FP64 form): films, TAB, walls/splash, secondary-droplet records
  LiquidExtra lx;
  const double* __restrict__ wf;  // film thickness per cell (splash)
  int32_t* __restrict__ sb_type;  // record type per output position
  double* __restrict__ sb_data;   // record fields [kSBFields + NF][cap]
  int do_breakup;                 // 0 none, 1 TAB, 2 KH-RT
  int make_new_drops;
  int do_splash_box;
  double max_ppp;
  KhrtParams kh;  // KH-RT constants, mean injected parcel mass, dtmod * flow_dt
  // PeleC flavour (FP64): CGS state rho, rho*u, rho*E, rho*Y, T and CGS sources
  int pelec;
  int c_eng;
  const double* __restrict__ conv;  // PeleC: the state box in SI (pelec_convert_kernel)
  double src_scale[5 + PELEMP_MAX_FUEL];  // SI -> CGS factor per source component
  // embedded boundary (FP64 branch kernel, cut or covered cells in the box)
  int eb_on;
  EbDev eb;
  double min_eb_vfrac;
  double inv_vol;
  uint32_t* __restrict__ eb_rkey;  // per sorted position: key of the cell the deposit goes to, or kDeadKey
  double* __restrict__ eb_rval;    // that deposit [5 + NF][cap]
};


// One cell of the PeleC state (SprayInterpolation.H:126-146): Y, u, kinetic
// and internal energy formed in CGS, then rho, u, e to SI and T from EY2T
// with the state's T as the guess.  S points at the cell's first component.
template <class M>
__device__ __attribute__((always_inline)) inline void pelec_state_cell(const double* S, long long cs, int c_rho,
                                                                      int c_spec, int c_vel, int c_eng, int c_T,
                                                                      double& rho, double* Y, double* vel,
                                                                      double& T) {
  constexpr int NS = M::NS;
  const double cur_rho = S[c_rho * cs];
  const double inv_rho = 1. / cur_rho;
#pragma unroll
  for (int n = 0; n < NS; ++n) Y[n] = S[(c_spec + n) * cs] * inv_rho;
  double ke = 0.;
  double vc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    vc[q] = S[(c_vel + q) * cs] * inv_rho;
    ke += vc[q] * vc[q] / 2.;
  }
  const double intEng = S[c_eng * cs] * inv_rho - ke;
  rho = cur_rho * 1.e3;
#pragma unroll
  for (int q = 0; q < 3; ++q) vel[q] = vc[q] * 1.e-2;
  T = GasProps<M>::EY2T(intEng * 1.e-4, Y, S[c_T * cs]);
}

// The PeleC state of the whole state box in SI, once per update (one thread
// per cell): conv[(q * ncells) + cell] for q = rho, Y[NS], vel[3], T, so the
// update kernel stages PeleC cells without the EY2T solve.
template <class M>
__global__ void __launch_bounds__(256) pelec_convert_kernel(const double* __restrict__ S, long long ncells,
                                                            int c_rho, int c_spec, int c_vel, int c_eng, int c_T,
                                                            double* __restrict__ conv) {
  constexpr int NS = M::NS;
  for (long long c = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; c < ncells;
       c += static_cast<long long>(gridDim.x) * blockDim.x) {
    double rho, Y[NS], vel[3], T;
    pelec_state_cell<M>(S + c, ncells, c_rho, c_spec, c_vel, c_eng, c_T, rho, Y, vel, T);
    conv[c] = rho;
#pragma unroll
    for (int n = 0; n < NS; ++n) conv[(1 + n) * ncells + c] = Y[n];
#pragma unroll
    for (int q = 0; q < 3; ++q) conv[(1 + NS + q) * ncells + c] = vel[q];
    conv[(4 + NS) * ncells + c] = T;
  }
}

// fields of a secondary-droplet record (SBData.H:7-19): norm(3) vel(3)
// loc(3) T0 ref_dia num_dens phi1 phi2 phi3, then Y0[NF]
constexpr int kSBFields = 15;

// Cell (rx, ry, rz), relative to the domain's lower corner, inside the
// domain grown by g cells (a negative g: no limit).
__device__ __attribute__((always_inline)) inline bool in_grown_box(int rx, int ry, int rz, const int* ncell, int g) {
  if (g < 0) return true;
  return rx >= -g && rx < ncell[0] + g && ry >= -g && ry < ncell[1] + g && rz >= -g && rz < ncell[2] + g;
}

__device__ inline int xcd_tile(int b, int ntiles) {
  // gfx950 dispatches workgroup b to XCD b % 8; give each XCD a contiguous
  // range of tiles.
  if ((ntiles & 7) != 0) return b;
  const int per = ntiles >> 3;
  return (b & 7) * per + (b >> 3);
}

// check_bounds (SprayInterpolation.H:64-97)
__device__ __attribute__((always_inline)) inline bool dev_check_bounds(const double* pos, const double* plo, const double* phi,
                                        const double* dx, const int* blo, const int* bhi,
                                        int* bflags) {
  for (int hilo = 0; hilo < 2; ++hilo) {
    const int fact = (hilo == 0) ? 1 : -1;
    for (int dir = 0; dir < 3; ++dir) {
      const double dom_loc = (hilo == 0) ? plo[dir] : phi[dir];
      const int bflag = (hilo == 0) ? blo[dir] : bhi[dir];
      if (bflag != 0) {
        const double diff = static_cast<double>(fact) * (pos[dir] - dom_loc);
        if (diff < 0.) {
          if (bflag == 1) {
            bflags[dir] = -fact;
          } else {
            return true;
          }
        } else if (diff < 0.5 * dx[dir]) {
          bflags[dir] = -2 * fact;
        }
      }
    }
  }
  return false;
}

// Occupancy targets (workgroups of 4 waves per CU = waves per SIMD), with
// the units compiled without machine-level loop-invariant code motion (which
// kept hoisted FP64 constants live and spilled scalar registers): the plain
// FP64 kernel with up to two liquid species runs at 3 waves per SIMD (168
// registers, about 100 bytes per lane spilled; measured 18 % faster than 2
// waves without spills), the six-fuel one and the branch kernels at 2, the
// TAB branch kernel too (193 registers; it needed ~390 with the hoisting).
// The FP32 form with one liquid species runs at 5 waves per SIMD (96
// registers, 8 bytes per lane spilled; 10 % faster than 4 waves), with two
// at 4 (120 registers; at 5 waves it spilled 88 bytes per lane and ran 19 %
// slower), with six at 3 (168 registers, about 190 bytes per lane spilled;
// 15 % faster than 2 waves without spills).
template <class R, int NF, bool TAB = false, bool BR = false>
struct UpdOccupancy {
  static constexpr int kMinBlocks =
      TAB ? 2 : ((sizeof(R) == 8) ? ((!BR && NF <= 2) ? 3 : 2) : (NF == 1 ? 5 : (NF <= 2 ? 4 : 3)));
};

// Word of the source FAB for component q (mom 3, mass, Y nf, energy) of cell
// c of the tile at (otx, oty, otz): null when the component is switched off
// (SprayParticles.cpp:426-441) or the cell lies outside the FAB (an error
// unless the droplets are ghosts, :420-425).
template <class R, int NF>
__device__ __attribute__((always_inline)) inline double* src_word(const UpdateArgs<R, NF>& a, int otx, int oty,
                                                                   int otz, int c, int q) {
  if (q < 3 && !a.fc.mom_trans) return nullptr;
  if (q >= 3 && q < 4 + NF && !a.fc.mass_trans) return nullptr;
  const int comp = (q < 3) ? a.c_msrc + q : (q == 3) ? a.c_rsrc : (q < 4 + NF) ? a.c_ssrc + (q - 4) : a.c_esrc;
  const int gx = a.dom_lo[0] + otx + (c % kTile) - a.src.lo[0];
  const int gy = a.dom_lo[1] + oty + ((c / kTile) % kTile) - a.src.lo[1];
  const int gz = a.dom_lo[2] + otz + c / (kTile * kTile) - a.src.lo[2];
  if (gx < 0 || gx >= a.src.n[0] || gy < 0 || gy >= a.src.n[1] || gz < 0 || gz >= a.src.n[2]) {
    if (!a.is_ghost) set_err(a.err, PELEMP_E_SRC_BOX);
    return nullptr;
  }
  return &a.src.data[comp * a.src.cstride + (static_cast<long long>(gz) * a.src.n[1] + gy) * a.src.n[0] + gx];
}

// Type of the per-droplet source contributions staged in LDS for the ordered
// per-cell reduction: the plain kernels stage the droplet's source in the
// form's precision and the reduction multiplies it by the cell coefficient
// in double; the branch kernels stage coefficient x source in double (the
// coefficient varies per droplet in cut cells).
template <class R, bool BR>
using ContribT = std::conditional_t<BR, double, R>;

// Per-droplet values carried from the gather to the update.
template <class M, int NF, class R>
struct DropLocal {
  int32_t o;     // storage position in the input buffer
  int32_t slot;  // its cold-data slot
  int lc, cx, cy, cz;
  Drop<R, NF> d;
  R pos[3];
  R film;
  R bm2, n0;      // breakup state (FP64 form with breakup / splash)
  double mk[NF];  // FP32 form: liquid species masses
  GasAt<R, M::NS> g;
  R cBoilT[NF];
  int bflags[3];
  int ijkc[3];  // branch kernel: the gather's cell (unclamped, or the EB's closest cell)
  bool fe;      // branch kernel: EB weights were used
};

// Load droplet s of the tile: its storage position, slot and fields (the
// loads only; drop_gather does the rest).  Issued ahead of the work that
// needs them, so their latency overlaps with the halo staging or with the
// previous chunk's reductions.
template <class M, int NF, class R, bool POS64, bool BR = false>
__device__ __attribute__((always_inline)) inline void drop_load(const UpdateArgs<R, NF>& a, int32_t s,
                                                                DropLocal<M, NF, R>& L) {
  using F = Fields<NF>;
  const long long cap = a.cap;
  const int32_t o = a.perm[s];
  L.o = o;
  L.slot = a.slot_in[o];
  Drop<R, NF>& d = L.d;
  R* pos = L.pos;
  if constexpr (POS64) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      pos[q] = a.fin[(F::POS + q) * cap + o];
      d.vel[q] = a.fin[(F::VEL + q) * cap + o];
    }
    d.T = a.fin[F::T * cap + o];
    d.dia = a.fin[F::DIA * cap + o];
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) d.Y[spf] = a.fin[(F::Y + spf) * cap + o];
    d.numDens = a.fin[F::ND * cap + o];
    L.film = a.fin[F::FILM * cap + o];
    if constexpr (BR) {
      d.bm1 = a.fin[F::BM1 * cap + o];
      L.bm2 = a.fin[F::BM2 * cap + o];
      L.n0 = a.fin[F::N0 * cap + o];
    }
  } else {
    using G = Fields32;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      pos[q] = a.fin[(G::POS + q) * cap + o];
      d.vel[q] = a.fin[(G::VEL + q) * cap + o];
    }
    d.T = a.fin[G::T * cap + o];
    d.numDens = a.fin[G::ND * cap + o];
    L.film = R(0);  // no wall films in this form
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) L.mk[spf] = a.min[spf * cap + o];
  }
}

// Gather the gas state of droplet s of the tile (loaded by drop_load).
template <class M, int NF, class R, bool POS64, bool AT_BOUNDS, bool BR = false, bool EB = false>
__device__ __attribute__((always_inline)) inline void drop_gather(const UpdateArgs<R, NF>& a, const R* gas,
                                                                  const unsigned char* gvalid, const int32_t* cs,
                                                                  int otx, int oty, int otz, int32_t s,
                                                                  DropLocal<M, NF, R>& L) {
  constexpr int NS = M::NS;
  // local cell of sorted position s: largest c with cs[c] <= s
  int lc = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1)
    if (lc + step < kTileCells && cs[lc + step] <= s) lc += step;
  L.lc = lc;
  L.cx = otx + (lc % kTile);
  L.cy = oty + ((lc / kTile) % kTile);
  L.cz = otz + lc / (kTile * kTile);

  Drop<R, NF>& d = L.d;
  R* pos = L.pos;
  if constexpr (!POS64) {
    // the ODE's mass and mass fractions in single precision (the FP64
    // species masses stay the bookkeeping)
    double m = 0.;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) m += L.mk[spf];
    const R mf = static_cast<R>(m);
    const R im = R(1) / mf;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) d.Y[spf] = static_cast<R>(L.mk[spf]) * im;
    d.mass = mf;
    d.dia = R(0);
  }
  if constexpr (!BR) d.bm1 = R(0);
  d.alive = true;

  // trilinear corner cell and weights (SprayInterpolation.H:158-199)
  int ijk[3];
  R sx_hi[3];
  int* bflags = L.bflags;
  bflags[0] = bflags[1] = bflags[2] = 0;
  if constexpr (POS64) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double lx = (static_cast<double>(pos[q]) - a.plo[q]) * a.dxi[q] + 0.5;
      ijk[q] = static_cast<int>(floor(lx));
      sx_hi[q] = static_cast<R>(lx - static_cast<double>(ijk[q]));
    }
    if constexpr (AT_BOUNDS) {
      double pd[3] = {static_cast<double>(pos[0]), static_cast<double>(pos[1]), static_cast<double>(pos[2])};
      if (dev_check_bounds(pd, a.plo, a.phi, a.dx, a.bndry_lo, a.bndry_hi, bflags))
        set_err(a.err, PELEMP_E_LEFT_DOMAIN);
    }
  } else {
    const int c3[3] = {L.cx, L.cy, L.cz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const R l = pos[q] + R(0.5);
      const R fl = floor(l);
      ijk[q] = c3[q] + static_cast<int>(fl);
      sx_hi[q] = l - fl;
    }
  }
  R ssv[2][3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const R sx_lo = R(1.) - sx_hi[q];
    if (bflags[q] > 1) {
      ssv[0][q] = R(1.);
      ssv[1][q] = R(0.);
    } else if (bflags[q] < -1) {
      ssv[0][q] = R(0.);
      ssv[1][q] = R(1.);
    } else {
      ssv[0][q] = sx_lo;
      ssv[1][q] = sx_hi[q];
    }
  }
  // InterpolateGasPhase (SprayInterpolation.H:99-150)
  GasAt<R, NS>& g = L.g;
  g.rho = R(0);
  g.T = R(0);
#pragma unroll
  for (int q = 0; q < 3; ++q) g.vel[q] = R(0);
#pragma unroll
  for (int n = 0; n < NS; ++n) g.Y[n] = R(0);
  // add cell (rx, ry, rz) (domain-relative) with weight cw
  auto add_cell = [&](int rx, int ry, int rz, R cw) {
    int hx = rx - otx + 1, hy = ry - oty + 1, hz = rz - otz + 1;
    const bool inr = hx >= 0 && hx < kHaloW && hy >= 0 && hy < kHaloW && hz >= 0 && hz < kHaloW;
    if constexpr (AT_BOUNDS) {
      // a droplet more than half a cell behind a reflective wall (kept in the
      // boundary cell) reaches past the halo: read the state's ghost cells
      // directly, with the staging's arithmetic
      if (!inr) {
        const int lx = a.dom_lo[0] + rx - a.st.lo[0];
        const int ly = a.dom_lo[1] + ry - a.st.lo[1];
        const int lz = a.dom_lo[2] + rz - a.st.lo[2];
        if (!(lx >= 0 && lx < a.st.n[0] && ly >= 0 && ly < a.st.n[1] && lz >= 0 && lz < a.st.n[2]) ||
            !in_grown_box(rx, ry, rz, a.ncell, a.st_g)) {
          set_err(a.err, PELEMP_E_STATE_BOX);
          return;
        }
        const long long c = (static_cast<long long>(lz) * a.st.n[1] + ly) * a.st.n[0] + lx;
        if constexpr (sizeof(R) == 8) {
          if (a.pelec) {
            const double* V = a.conv + c;
            const long long cs = a.st.cstride;
            g.rho += cw * V[0];
#pragma unroll
            for (int n = 0; n < NS; ++n) g.Y[n] += cw * V[(1 + n) * cs];
#pragma unroll
            for (int q = 0; q < 3; ++q) g.vel[q] += cw * V[(1 + NS + q) * cs];
            g.T += cw * V[(4 + NS) * cs];
            return;
          }
        }
        double v[5 + NS];
        state_cell<NS>(a.st, c, a.c_rho, a.c_spec, a.c_vel, a.c_T, v);
        g.rho += cw * static_cast<R>(v[0]);
#pragma unroll
        for (int n = 0; n < NS; ++n) g.Y[n] += cw * static_cast<R>(v[1 + n]);
#pragma unroll
        for (int q = 0; q < 3; ++q) g.vel[q] += cw * static_cast<R>(v[1 + NS + q]);
        g.T += cw * static_cast<R>(v[4 + NS]);
        return;
      }
    }
    hx = inr ? hx : 0;
    hy = inr ? hy : 0;
    hz = inr ? hz : 0;
    const int h = (hz * kHaloW + hy) * kHaloW + hx;
    if (!inr || !gvalid[h]) set_err(a.err, PELEMP_E_STATE_BOX);
    const R* gc = gas + h * gas_stride<R, NS>();
    g.rho += cw * gc[0];
#pragma unroll
    for (int n = 0; n < NS; ++n) g.Y[n] += cw * gc[1 + n];
#pragma unroll
    for (int q = 0; q < 3; ++q) g.vel[q] += cw * gc[1 + NS + q];
    g.T += cw * gc[4 + NS];
  };
  bool eb_weights = false;
  if constexpr (BR) {
    L.fe = false;
    if constexpr (EB) {
      // the reference's ijkc: the unclamped containing cell
#pragma unroll
      for (int q = 0; q < 3; ++q)
        L.ijkc[q] = static_cast<int>(floor((static_cast<double>(pos[q]) - a.plo[q]) * a.dxi[q]));
      // eb_interp (SprayInterpolation.H:395-632) when the stencil has cut or
      // covered cells; it may move the droplet and its deposit cell
      EbStencil es;
      const int r = eb_interp(a.eb, L.pos, L.ijkc, ijk, a.plo, a.dx, a.dxi, a.min_eb_vfrac, es);
      if (r < 0) {
        set_err(a.err, PELEMP_E_EB_STENCIL);
      } else if (r > 0) {
        L.fe = true;
        eb_weights = true;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (es.w[q] > 0.) add_cell(es.idx[q][0], es.idx[q][1], es.idx[q][2], static_cast<R>(es.w[q]));
      }
    }
  }
  if (!eb_weights) {
    if constexpr (!AT_BOUNDS) {
      // the eight corners ijk-1 .. ijk of a droplet away from the walls lie
      // in the staged halo: one bounds check for the block, the halo cell
      // of corner (-1,-1,-1) and constant offsets for the others; the
      // per-corner state-box flags are read only when the staging found a
      // halo cell outside the state box (gvalid[kHaloCells] == 0)
      const int hx = ijk[0] - otx, hy = ijk[1] - oty, hz = ijk[2] - otz;
      const bool inr = hx >= 0 && hx < kHaloW - 1 && hy >= 0 && hy < kHaloW - 1 && hz >= 0 && hz < kHaloW - 1;
      if (!inr) set_err(a.err, PELEMP_E_STATE_BOX);
      const int h0 = inr ? (hz * kHaloW + hy) * kHaloW + hx : 0;
      const bool chk = gvalid[kHaloCells] == 0;
      const R* gb = gas + h0 * gas_stride<R, NS>();
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
#pragma unroll
          for (int ii = 0; ii < 2; ++ii) {
            const R cw = ssv[ii][0] * ssv[jj][1] * ssv[kk][2];
            if (cw > R(0.)) {
              constexpr int kOff[2][2][2] = {{{0, 1}, {kHaloW, kHaloW + 1}},
                                             {{kHaloW * kHaloW, kHaloW * kHaloW + 1},
                                              {kHaloW * kHaloW + kHaloW, kHaloW * kHaloW + kHaloW + 1}}};
              const int off = kOff[kk][jj][ii];
              if (chk && !gvalid[h0 + off]) set_err(a.err, PELEMP_E_STATE_BOX);
              const R* gc = gb + off * gas_stride<R, NS>();
              g.rho += cw * gc[0];
#pragma unroll
              for (int n = 0; n < NS; ++n) g.Y[n] += cw * gc[1 + n];
#pragma unroll
              for (int q = 0; q < 3; ++q) g.vel[q] += cw * gc[1 + NS + q];
              g.T += cw * gc[4 + NS];
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int kk = -1; kk < 1; ++kk) {
#pragma unroll
        for (int jj = -1; jj < 1; ++jj) {
#pragma unroll
          for (int ii = -1; ii < 1; ++ii) {
            const R cw = ssv[ii + 1][0] * ssv[jj + 1][1] * ssv[kk + 1][2];
            if (cw > R(0.)) add_cell(ijk[0] + ii, ijk[1] + jj, ijk[2] + kk, cw);
          }
        }
      }
    }
  }
  gas_define<M, R>(g);
  calc_boil_t<NF, R>(a.fc, g.p, L.cBoilT);
}

// Droplet physics, source contribution, drift, store and re-sort
// classification of droplet s (storage position s of the output buffer).
template <class M, int NF, class R, bool POS64, bool AT_BOUNDS, bool DEP = true>
__device__ __attribute__((always_inline)) inline void drop_update(const UpdateArgs<R, NF>& a, int tile, int t,
                                                                  int32_t s, DropLocal<M, NF, R>& L,
                                                                  const SubRates<M, NF, R>* pre, R* contrib,
                                                                  unsigned char* stayed) {
  constexpr int NSRC = 5 + NF;
  using F = Fields<NF>;
  const long long cap = a.cap;
  Drop<R, NF>& d = L.d;
  R* pos = L.pos;
  const R film = L.film;
  int* bflags = L.bflags;
  DropSrc<R, NF> sr;
  if (film > R(0.)) {
    set_err(a.err, PELEMP_E_BAD_CONFIG);  // wall-film droplets use the film pass
    sr.mom[0] = sr.mom[1] = sr.mom[2] = R(0);
    sr.mass = R(0);
    sr.eng = R(0);
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) sr.Ydot[spf] = R(0);
  } else {
    spray_source<M, NF, R, !POS64, DEP>(a.sub_dt, L.g, a.fc, d, L.cBoilT, sr, pre);
  }
  // the droplet's source, in the form's precision; the reduction multiplies
  // each by cur_coef in double (SprayParticles.cpp:419-441)
  R* cb = &contrib[t * NSRC];
#pragma unroll
  for (int q = 0; q < 3; ++q) cb[q] = sr.mom[q];
  cb[3] = sr.mass;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) cb[4 + spf] = sr.Ydot[spf];
  cb[4 + NF] = sr.eng;

  // drift (SprayParticles.cpp:443-479) and new cell key
  uint32_t key = kDeadKey;
  if constexpr (POS64) {
    double p3[3] = {static_cast<double>(pos[0]), static_cast<double>(pos[1]), static_cast<double>(pos[2])};
    if (a.do_move && !a.fc.fixed_parts && d.alive && !(film > R(0.))) {
#pragma unroll
      for (int q = 0; q < 3; ++q) p3[q] += static_cast<double>(a.sub_dt) * static_cast<double>(d.vel[q]);
      if constexpr (AT_BOUNDS) {
        if (dev_check_bounds(p3, a.plo, a.phi, a.dx, a.bndry_lo, a.bndry_hi, bflags)) d.alive = false;
      }
    }
    // a ghost droplet whose (unclamped, unwrapped) cell left the source box
    // is deleted (SprayParticles.cpp:477-479)
    if (a.is_ghost) {
      int gc[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) gc[q] = static_cast<int>(floor((p3[q] - a.plo[q]) * a.dxi[q]));
      if (!in_grown_box(gc[0], gc[1], gc[2], a.ncell, a.src_g)) d.alive = false;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (a.bndry_lo[q] == 0) {
        const double Lq = a.phi[q] - a.plo[q];
        if (p3[q] >= a.phi[q]) {
          p3[q] -= Lq;
        } else if (p3[q] < a.plo[q]) {
          p3[q] += Lq;
        }
      }
      pos[q] = static_cast<R>(p3[q]);
    }
    if (d.alive) {
      int c[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        int ic = static_cast<int>(floor((p3[q] - a.plo[q]) * a.dxi[q]));
        ic = ic < 0 ? 0 : (ic >= a.ncell[q] ? a.ncell[q] - 1 : ic);
        c[q] = ic;
      }
      key = tile_key(c[0], c[1], c[2], a.ntile);
    }
  } else {
    int c[3] = {L.cx, L.cy, L.cz};
    if (a.do_move && !a.fc.fixed_parts && d.alive && !(film > R(0.))) {
      bool left_src = false;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        R xi = pos[q] + a.sub_dt_dxi[q] * d.vel[q];
        const R sh = floor(xi);
        xi -= sh;
        if (xi >= R(1.)) xi = R(0.);  // rounding of a tiny negative xi
        int ic = c[q] + static_cast<int>(sh);
        const int n = a.ncell[q];
        // a ghost droplet whose unwrapped cell left the source box is
        // deleted (SprayParticles.cpp:477-479)
        if (a.is_ghost && a.src_g >= 0 && (ic < -a.src_g || ic >= n + a.src_g)) left_src = true;
        // periodic: ((ic % n) + n) % n, with the integer division only for a
        // move of more than one period
        if (ic < 0) {
          ic += n;
          if (ic < 0) ic = ((ic % n) + n) % n;
        } else if (ic >= n) {
          ic -= n;
          if (ic >= n) ic %= n;
        }
        c[q] = ic;
        pos[q] = xi;
      }
      if (a.is_ghost && left_src) d.alive = false;
    }
    if (d.alive) key = tile_key(c[0], c[1], c[2], a.ntile);
  }

  R* fo = a.fout;
  if constexpr (POS64) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      fo[(F::POS + q) * cap + s] = pos[q];
      fo[(F::VEL + q) * cap + s] = d.vel[q];
    }
    fo[F::T * cap + s] = d.T;
    fo[F::DIA * cap + s] = d.dia;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) fo[(F::Y + spf) * cap + s] = d.Y[spf];
    fo[F::ND * cap + s] = d.numDens;
    fo[F::FILM * cap + s] = film;
  } else {
    using G = Fields32;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      fo[(G::POS + q) * cap + s] = pos[q];
      fo[(G::VEL + q) * cap + s] = d.vel[q];
    }
    fo[G::T * cap + s] = d.T;
    fo[G::DIA * cap + s] = d.dia;
    fo[G::ND * cap + s] = d.numDens;
    // FP64 mass bookkeeping: accumulate the single-precision increments
    double m = 0.;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      double v = L.mk[spf] + static_cast<double>(d.dmk[spf]);
      v = v > 0. ? v : 0.;
      L.mk[spf] = v;
      m += v;
    }
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      // Y <= 1e-12 is set to zero (Drag.H:385-387)
      const double v = (L.mk[spf] <= 1.e-12 * m) ? 0. : L.mk[spf];
      a.mout[spf * cap + s] = v;
    }
  }
  a.slot_out[s] = L.slot;
  a.key_out[s] = key;
  // classify for the incremental re-sort
  const uint32_t k_old = static_cast<uint32_t>(tile) * kTileCells + L.lc;
  stayed[t] = (key == k_old) ? 1 : 0;
  if (key != kDeadKey && key != k_old) {
    // relative to the new cell `key`, the droplet comes from a cell with a
    // lower key (it precedes the stayers) or a higher key (it follows them)
    const int32_t arr = (k_old < key) ? (atomicAdd(&a.mov_lo[key], 1) | (1 << 30))
                                      : (atomicAdd(&a.mov_hi[key], 1) | (2 << 30));
    a.arrival_out[s] = arr;
  }
}

// FP64 form with the branches (SprayParticles.cpp:354-479), in three steps
// around the workgroup's TAB stage: br_source (film source or droplet source
// and the deposit contribution), tab_stage (the TAB update of the chunk's
// droplets, reordered for wavefront coherence) and br_finish (TAB split or
// KH-RT, drift, wall reflection / splash, store of the breakup state and of
// the secondary-droplet record).  The deposit contribution depends only on
// the source, so computing it before the breakup update changes nothing.
template <class M, int NF, bool EB = false>
__device__ __attribute__((always_inline)) inline void br_source(const UpdateArgs<double, NF>& a, int t,
                                                                int32_t s, DropLocal<M, NF, double>& L,
                                                                BDrop<NF>& b, double& Reyn_d, double* contrib,
                                                                unsigned char* redir) {
  constexpr int NSRC = 5 + NF;
  const long long cap = a.cap;
  Drop<double, NF>& d = L.d;
  const bool is_film = L.film > 0.;
  b.alive = true;
  b.film = L.film;
  b.numDens = d.numDens;
  b.N0 = L.n0;
  b.bm2 = L.bm2;
  DropSrc<double, NF> sr;
  Reyn_d = 0.;
  if (is_film) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      b.pos[q] = L.pos[q];
      b.vel[q] = d.vel[q];
    }
    b.T = d.T;
    b.dia = d.dia;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) b.Y[spf] = d.Y[spf];
    b.bm1 = d.bm1;
    film_source<M, NF>(a.sub_dt, L.g, a.fc, a.lx, b, L.cBoilT, sr);
  } else {
    Reyn_d = spray_source<M, NF, double, false>(a.sub_dt, L.g, a.fc, d, L.cBoilT, sr);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      b.pos[q] = L.pos[q];
      b.vel[q] = d.vel[q];
    }
    b.T = d.T;
    b.dia = d.dia;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) b.Y[spf] = d.Y[spf];
    b.bm1 = d.bm1;
    b.alive = d.alive;
  }
  // contributions cur_coef * source (SprayParticles.cpp:413-441); with an EB
  // 1/vfrac in a cut deposit cell, and eb_interp's closest cell may be a
  // neighbour: that deposit is kept aside for the redirect pass
  double coef = a.cur_coef;
  bool to_other = false;
  int dep[3] = {L.cx, L.cy, L.cz};
  if constexpr (EB) {
    // with EB weights, eb_interp's cell: the droplet's own, or the closest
    // acceptable cell of the stencil (a neighbour, possibly a ghost cell)
    if (L.fe) {
#pragma unroll
      for (int q = 0; q < 3; ++q) dep[q] = L.ijkc[q];
    }
    if (a.eb.single(dep[0], dep[1], dep[2])) {
      double cvol = a.inv_vol;
      cvol *= 1. / (a.eb.vf(dep[0], dep[1], dep[2]));
      coef = -cvol * a.sub_dt / a.flow_dt;
    }
    to_other = dep[0] != L.cx || dep[1] != L.cy || dep[2] != L.cz;
  }
  double* cb = &contrib[t * NSRC];
#pragma unroll
  for (int q = 0; q < 3; ++q) cb[q] = coef * sr.mom[q];
  cb[3] = coef * sr.mass;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) cb[4 + spf] = coef * sr.Ydot[spf];
  cb[4 + NF] = coef * sr.eng;
  if constexpr (EB) {
    redir[t] = to_other ? 1 : 0;
    a.eb_rkey[s] = to_other ? eb_cell_code(dep[0], dep[1], dep[2], a.ncell) : kDeadKey;
    if (to_other) {
#pragma unroll
      for (int q = 0; q < NSRC; ++q) a.eb_rval[q * cap + s] = cb[q];
    }
  }
}

// TAB stage of one chunk (TABBreakup.H:12-184 per droplet).  The number of
// TAB sub-intervals differs from droplet to droplet (0.1 r sqrt(3 rho_l /
// rho_g) / |u_g - u_d| against dt, shorter after a breakup event), and a
// wavefront runs as long as its slowest lane: the chunk's droplets are
// binned by an estimate of that number (rho_l ~ 700 kg/m^3: an ordering key
// only), their TAB inputs written to LDS in bin order, each lane updates the
// droplet of its position in that order and the results go back to the
// droplets' own lanes.  A droplet's result depends on its inputs only, so
// the order changes no value.  Every thread of the workgroup takes part.
template <int NF>
struct TabRec {
  static constexpr int kReyn = 0, kBoil = 1, kRho = 1 + NF, kGvel = 2 + NF, kT = 5 + NF, kDia = 6 + NF,
                       kND = 7 + NF, kBM1 = 8 + NF, kBM2 = 9 + NF, kVel = 10 + NF, kY = 13 + NF,
                       kUtan = 13 + 2 * NF, kN = 14 + 2 * NF;
};
constexpr int kTabBins = 16;

template <class M, int NF>
__device__ __attribute__((always_inline)) inline void tab_stage(const UpdateArgs<double, NF>& a, bool need, int t,
                                                                double Reyn_d, const DropLocal<M, NF, double>& L,
                                                                BDrop<NF>& b, double& Utan, double* rec,
                                                                int32_t* cnt) {
  using T = TabRec<NF>;
  constexpr int W = kUpdThreads;
  if (t < kTabBins) cnt[t] = 0;
  __syncthreads();
  int bin = 0, pos = 0;
  if (need) {
    double dv2 = 0.;
#pragma unroll
    for (int q = 0; q < 3; ++q) dv2 += (L.g.vel[q] - b.vel[q]) * (L.g.vel[q] - b.vel[q]);
    const double est = a.sub_dt * std::sqrt(dv2 * L.g.rho) / (0.05 * b.dia * 45.8257569495584);
    const double e = (est >= 0.) ? (est < kTabBins ? est : kTabBins) : 0.;  // NaN: bin 0
    bin = static_cast<int>(e);
    bin = bin < kTabBins ? bin : kTabBins - 1;
    pos = atomicAdd(&cnt[bin], 1);
  }
  __syncthreads();
  if (t == 0) {
    int32_t sum = 0;
    for (int k = 0; k < kTabBins; ++k) {
      const int32_t c = cnt[k];
      cnt[k] = sum;
      sum += c;
    }
    cnt[kTabBins] = sum;
  }
  __syncthreads();
  const int r = need ? cnt[bin] + pos : 0;
  if (need) {
    rec[T::kReyn * W + r] = Reyn_d;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      rec[(T::kBoil + spf) * W + r] = L.cBoilT[spf];
      rec[(T::kY + spf) * W + r] = b.Y[spf];
    }
    rec[T::kRho * W + r] = L.g.rho;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      rec[(T::kGvel + q) * W + r] = L.g.vel[q];
      rec[(T::kVel + q) * W + r] = b.vel[q];
    }
    rec[T::kT * W + r] = b.T;
    rec[T::kDia * W + r] = b.dia;
    rec[T::kND * W + r] = b.numDens;
    rec[T::kBM1 * W + r] = b.bm1;
    rec[T::kBM2 * W + r] = b.bm2;
  }
  __syncthreads();
  if (t < cnt[kTabBins]) {
    BDrop<NF> d;
    double cb[NF], vg[3];
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      cb[spf] = rec[(T::kBoil + spf) * W + t];
      d.Y[spf] = rec[(T::kY + spf) * W + t];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      vg[q] = rec[(T::kGvel + q) * W + t];
      d.vel[q] = rec[(T::kVel + q) * W + t];
    }
    d.T = rec[T::kT * W + t];
    d.dia = rec[T::kDia * W + t];
    d.numDens = rec[T::kND * W + t];
    d.bm1 = rec[T::kBM1 * W + t];
    d.bm2 = rec[T::kBM2 * W + t];
    const double U = tab_update<NF>(rec[T::kReyn * W + t], a.sub_dt, cb, rec[T::kRho * W + t], vg, a.fc, a.lx, d);
    rec[T::kDia * W + t] = d.dia;
    rec[T::kND * W + t] = d.numDens;
    rec[T::kBM1 * W + t] = d.bm1;
    rec[T::kBM2 * W + t] = d.bm2;
    rec[T::kUtan * W + t] = U;
  }
  __syncthreads();
  if (need) {
    b.dia = rec[T::kDia * W + r];
    b.numDens = rec[T::kND * W + r];
    b.bm1 = rec[T::kBM1 * W + r];
    b.bm2 = rec[T::kBM2 * W + r];
    Utan = rec[T::kUtan * W + r];
  }
}

template <class M, int NF, bool AT_BOUNDS, bool EB = false>
__device__ __attribute__((always_inline)) inline void br_finish(const UpdateArgs<double, NF>& a, int tile, int t,
                                                                int32_t s, DropLocal<M, NF, double>& L,
                                                                BDrop<NF>& b, double Reyn_d, double Utan,
                                                                unsigned char* stayed) {
  using F = Fields<NF>;
  const long long cap = a.cap;
  const bool is_film = L.film > 0.;
  SBRec<NF> rec;
  rec.type = kSBNone;
  // breakup (SprayParticles.cpp:398-412): the TAB split, or the KH-RT update
  // (which also runs for the TAB model when no new droplets can be made);
  // the TAB update itself ran in the tab stage
  if (b.alive && a.do_breakup > 0) {
    if (a.do_breakup == 1 && a.make_new_drops) {
      tab_split<NF>(b, a.max_ppp, Utan, rec);
    } else {
      khrt_update<M::NS, NF>(Reyn_d, L.cBoilT, L.g, a.fc, a.lx, a.kh, a.make_new_drops != 0, b, rec);
    }
  }

  // drift, bounds and walls (SprayParticles.cpp:443-479), periodic wrap
  const bool moves = a.do_move && !a.fc.fixed_parts && b.alive && !is_film;
  if (moves) {
#pragma unroll
    for (int q = 0; q < 3; ++q) b.pos[q] += a.sub_dt * b.vel[q];
    if (AT_BOUNDS || (EB && L.fe)) {
      if (dev_check_bounds(b.pos, a.plo, a.phi, a.dx, a.bndry_lo, a.bndry_hi, L.bflags)) {
        b.alive = false;
      } else {
        // film thickness of the gather's cell (the EB's closest cell when
        // eb_interp moved it); a droplet that started behind a wall sees
        // none (the reference reads its ghost cell, which holds no film)
        int fc[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if constexpr (EB) {
            fc[q] = L.ijkc[q];
          } else {
            fc[q] = static_cast<int>(floor((L.pos[q] - a.plo[q]) * a.dxi[q]));
          }
        }
        bool inside = true;
#pragma unroll
        for (int q = 0; q < 3; ++q) inside = inside && fc[q] >= 0 && fc[q] < a.ncell[q];
        const double film_h =
            (a.do_splash_box && inside)
                ? a.wf[(static_cast<long long>(fc[2]) * a.ncell[1] + fc[1]) * a.ncell[0] + fc[0]]
                : 0.;
        impose_wall<NF>(a.do_splash_box != 0, b, a.fc, a.lx, a.plo, a.dx, L.bflags, L.cBoilT, L.g.p, rec, film_h,
                        EB ? &a.eb : nullptr, EB ? L.ijkc : nullptr);
      }
    }
  }
  // a ghost droplet whose cell left the source box is deleted
  // (SprayParticles.cpp:477-479): the cell of the new position (before the
  // periodic shift), or the gather's cell when it did not move
  if (a.is_ghost) {
    int gc[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      gc[q] = static_cast<int>(floor((b.pos[q] - a.plo[q]) * a.dxi[q]));
      if constexpr (EB) {
        if (!moves && L.fe) gc[q] = L.ijkc[q];
      }
    }
    if (!in_grown_box(gc[0], gc[1], gc[2], a.ncell, a.src_g)) b.alive = false;
  }
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (a.bndry_lo[q] == 0) b.pos[q] = wrap_periodic(b.pos[q], a.plo[q], a.phi[q]);
  uint32_t key = kDeadKey;
  if (b.alive) {
    int c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int ic = static_cast<int>(floor((b.pos[q] - a.plo[q]) * a.dxi[q]));
      ic = ic < 0 ? 0 : (ic >= a.ncell[q] ? a.ncell[q] - 1 : ic);
      c[q] = ic;
    }
    key = tile_key(c[0], c[1], c[2], a.ntile);
  }
  double* fo = a.fout;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    fo[(F::POS + q) * cap + s] = b.pos[q];
    fo[(F::VEL + q) * cap + s] = b.vel[q];
  }
  fo[F::T * cap + s] = b.T;
  fo[F::DIA * cap + s] = b.dia;
#pragma unroll
  for (int spf = 0; spf < NF; ++spf) fo[(F::Y + spf) * cap + s] = b.Y[spf];
  fo[F::ND * cap + s] = b.numDens;
  fo[F::FILM * cap + s] = b.film;
  fo[F::BM1 * cap + s] = b.bm1;
  fo[F::BM2 * cap + s] = b.bm2;
  fo[F::N0 * cap + s] = b.N0;
  // secondary-droplet record of this position
  a.sb_type[s] = rec.type;
  if (rec.type != kSBNone) {
    double* r = a.sb_data;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      r[(0 + q) * cap + s] = rec.norm[q];
      r[(3 + q) * cap + s] = rec.vel[q];
      r[(6 + q) * cap + s] = rec.loc[q];
    }
    r[9 * cap + s] = rec.T0;
    r[10 * cap + s] = rec.ref_dia;
    r[11 * cap + s] = rec.num_dens;
    r[12 * cap + s] = rec.phi1;
    r[13 * cap + s] = rec.phi2;
    r[14 * cap + s] = rec.phi3;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) r[(kSBFields + spf) * cap + s] = rec.Y0[spf];
  }
  a.slot_out[s] = L.slot;
  a.key_out[s] = key;
  // classify for the incremental re-sort
  const uint32_t k_old = static_cast<uint32_t>(tile) * kTileCells + L.lc;
  stayed[t] = (key == k_old) ? 1 : 0;
  if (key != kDeadKey && key != k_old) {
    const int32_t arr = (k_old < key) ? (atomicAdd(&a.mov_lo[key], 1) | (1 << 30))
                                      : (atomicAdd(&a.mov_hi[key], 1) | (2 << 30));
    a.arrival_out[s] = arr;
  }
}

// Surrogate stage of one chunk: hull test, network (whole wavefront) and
// predicted rates.  Returns whether this lane's droplet uses them.
template <class M, int NF>
__device__ __attribute__((always_inline)) inline bool surr_stage_drop(const UpdateArgs<float, NF>& a,
                                                                      const SurrWeights& W, bool active,
                                                                      DropLocal<M, NF, float>& L,
                                                                      SubRates<M, NF, float>& o) {
  bool use = active && a.fc.mass_trans && !(L.film > 0.f);
  float xn[kSurrIn];
  VaporLite<float, NF> v;
  float rho_part = 1.f, pmass = 1.f, dia = 1.f;
  float vel_part[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < kSurrIn; ++q) xn[q] = 0.f;
  if (use) {
    const Drop<float, NF>& d = L.d;
    float rinv = 0.f, mwinv = 0.f, Tb = 0.f;
#pragma unroll
    for (int spf = 0; spf < NF; ++spf) {
      rinv += d.Y[spf] / rhoL<float, NF>(a.fc, rmin(d.T, L.cBoilT[spf]), spf);
      mwinv += d.Y[spf] / a.fc.mw_fuel[spf];
      Tb += d.Y[spf] * L.cBoilT[spf];
    }
    rho_part = 1.f / rinv;
    pmass = d.mass;
    dia = rm::cbrt(6.f * pmass / (static_cast<float>(kPIc) * rho_part));
    if (!a.fc.fixed_parts) {
#pragma unroll
      for (int q = 0; q < 3; ++q) vel_part[q] = d.vel[q];
    }
    vapor_lite<M, NF, float>(a.fc, L.g, d.T, 1.f / mwinv, d.Y, L.cBoilT, v);
    float du2 = 0.f;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float dv = L.g.vel[q] - vel_part[q];
      du2 += dv * dv;
    }
    const float urel = rm::sqrt(du2);
    use = v.ok && surr_in_hull<float>(a.hull, dia, d.T, Tb, urel, L.g.T, L.g.p, v.Yfg, v.B_M);
    if (use) {
      float x[kSurrIn];
      surr_features<float>(dia, d.T, urel, L.g.T, L.g.p, v.Ysk_f, d.Y[0], (NF > 1) ? d.Y[NF - 1] : 0.f, x);
#pragma unroll
      for (int q = 0; q < kSurrIn; ++q) xn[q] = (x[q] - a.sn.c[q]) * a.sn.s[q];
    }
  }
  float yn[kSurrOut];
  surr_mlp(W, xn, yn);  // every lane of the wavefront
  if (use) {
    float z[kSurrOut];
#pragma unroll
    for (int q = 0; q < kSurrOut; ++q) z[q] = a.sn.mean[q] + a.sn.sd[q] * yn[q];
    surr_rates<M, NF, float>(z, v, L.g, a.fc, L.d.T, dia, pmass, rho_part, L.d.Y, vel_part, 0.f, L.cBoilT, o);
    const float ind = surr_indicator<M, NF, float>(o, a.sub_dt, pmass, a.fc.mom_trans);
    use = (ind < a.hull.indicator_max) && (o.part_temp_src == o.part_temp_src);
  }
  return use;
}

// One segment of a tile: the tile's droplets seg * kSegDroplets .. +
// kSegDroplets (all of them when the tile holds no more).  slot >= 0 marks
// a heavy tile: the segment's per-cell sums and stayer counts go to the
// partials of segment slot slot + seg, which split_combine_kernel adds to
// the source in segment order.
template <class M, int NF, class R, bool POS64, bool AT_BOUNDS, bool SURR, bool BR = false, bool EB = false,
          bool TAB = false, bool DEP = true>
__device__ __attribute__((always_inline)) inline void process_tile(const UpdateArgs<R, NF>& a, int tile, int seg,
                                                                   int slot, R* gas, unsigned char* gvalid,
                                                                   int32_t* cs, ContribT<R, BR>* contrib,
                                                                   unsigned char* stayed, const SurrWeights* sw,
                                                                   unsigned& n_surr, unsigned& n_fb,
                                                                   unsigned char* redir = nullptr,
                                                                   double* tabrec = nullptr,
                                                                   int32_t* tabcnt = nullptr) {
  constexpr int NS = M::NS;
  constexpr int NSRC = 5 + NF;   // mom[3], mass, Ydot[NF], eng
  constexpr int NITEM = (kTileCells * NSRC + kUpdThreads - 1) / kUpdThreads;

  const int t = threadIdx.x;
  const int txy = a.ntile[0] * a.ntile[1];
  const int otz = (tile / txy) * kTile;
  const int oty = ((tile % txy) / a.ntile[0]) * kTile;
  const int otx = (tile % a.ntile[0]) * kTile;

  if (t <= kTileCells) cs[t] = a.cstart[tile * kTileCells + t];
  if (t == 0) gvalid[kHaloCells] = 1;  // cleared below if a halo cell is outside the state box
  __syncthreads();
  const int32_t d0 = cs[0] + seg * kSegDroplets;
  const int32_t d1 = (cs[kTileCells] - d0 < kSegDroplets) ? cs[kTileCells] : d0 + kSegDroplets;
  if (d1 <= d0) {  // uniform: empty tile
    if (a.src_overwrite) {
      // its cells get no spray source this pass
      for (int item = t; item < kTileCells * NSRC; item += kUpdThreads) {
        double* w = src_word(a, otx, oty, otz, item / NSRC, item % NSRC);
        if (w) *w = 0.;
      }
    }
    if constexpr (SURR) __syncthreads();  // cs is rewritten by the next tile
    return;
  }
  const int32_t seg_tag = seg << kSegShift;

  // the plain and surrogate kernels load each chunk's droplets ahead: the
  // first chunk's before the halo staging, the next chunk's before the
  // chunk's stayer ranks and reductions (LDS work)
  constexpr bool kAhead = !BR;
  DropLocal<M, NF, R> Lp;
  if constexpr (kAhead) {
    if (d0 + t < d1) drop_load<M, NF, R, POS64>(a, d0 + t, Lp);
  }

  // ---- stage gas halo ----
  for (int h = t; h < kHaloCells; h += kUpdThreads) {
    const int hx = h % kHaloW, hy = (h / kHaloW) % kHaloW, hz = h / (kHaloW * kHaloW);
    const int gx = a.dom_lo[0] + otx + hx - 1;
    const int gy = a.dom_lo[1] + oty + hy - 1;
    const int gz = a.dom_lo[2] + otz + hz - 1;
    const int lx = gx - a.st.lo[0], ly = gy - a.st.lo[1], lz = gz - a.st.lo[2];
    // the state box: inside the FAB and inside the domain grown by st_g
    const bool ok = lx >= 0 && lx < a.st.n[0] && ly >= 0 && ly < a.st.n[1] && lz >= 0 && lz < a.st.n[2] &&
                    in_grown_box(otx + hx - 1, oty + hy - 1, otz + hz - 1, a.ncell, a.st_g);
    gvalid[h] = ok ? 1 : 0;
    if (!ok) gvalid[kHaloCells] = 0;
    if constexpr (sizeof(R) == 8) {
      if (ok && a.pelec) {
        // PeleC flavour: the cell in SI, converted before the update
        const long long c = (static_cast<long long>(lz) * a.st.n[1] + ly) * a.st.n[0] + lx;
        const double* V = a.conv + c;
        const long long cs = a.st.cstride;
#pragma unroll
        for (int f = 0; f < 5 + NS; ++f) gas[h * gas_stride<R, NS>() + f] = V[f * cs];
        continue;
      }
    }
    if (ok) {
      const long long c = (static_cast<long long>(lz) * a.st.n[1] + ly) * a.st.n[0] + lx;
      double v[5 + NS];
      state_cell<NS>(a.st, c, a.c_rho, a.c_spec, a.c_vel, a.c_T, v);
#pragma unroll
      for (int f = 0; f < 5 + NS; ++f) gas[h * gas_stride<R, NS>() + f] = static_cast<R>(v[f]);
    } else {
      // outside the state box: a droplet reaching it is an error; the
      // cleared cell keeps what it reads defined
#pragma unroll
      for (int f = 0; f < 5 + NS; ++f) gas[h * gas_stride<R, NS>() + f] = R(0);
    }
  }
  __syncthreads();

  double acc[NITEM];
#pragma unroll
  for (int r = 0; r < NITEM; ++r) acc[r] = 0.;
  int32_t srank = 0;  // threads t < 64: stayers of cell t seen so far

  for (int32_t base = d0; base < d1; base += kUpdThreads) {
    const int32_t s = base + t;
    if constexpr (BR) {
      const bool act = s < d1;
      DropLocal<M, NF, R> L;
      BDrop<NF> b;
      double Reyn_d = 0., Utan = 0.;
      if (act) {
        drop_load<M, NF, R, true, true>(a, s, L);
        drop_gather<M, NF, R, true, AT_BOUNDS, true, EB>(a, gas, gvalid, cs, otx, oty, otz, s, L);
        br_source<M, NF, EB>(a, t, s, L, b, Reyn_d, contrib, redir);
      }
      if constexpr (TAB) tab_stage<M, NF>(a, act && b.alive, t, Reyn_d, L, b, Utan, tabrec, tabcnt);
      if (act) br_finish<M, NF, AT_BOUNDS, EB>(a, tile, t, s, L, b, Reyn_d, Utan, stayed);
    } else if constexpr (!SURR) {
      if (s < d1) {
        drop_gather<M, NF, R, POS64, AT_BOUNDS>(a, gas, gvalid, cs, otx, oty, otz, s, Lp);
        drop_update<M, NF, R, POS64, AT_BOUNDS, DEP>(a, tile, t, s, Lp, nullptr, contrib, stayed);
      }
    } else {
      const bool active = s < d1;
      if (active) drop_gather<M, NF, R, POS64, AT_BOUNDS>(a, gas, gvalid, cs, otx, oty, otz, s, Lp);
      SubRates<M, NF, R> o;
      const bool use = surr_stage_drop<M, NF>(a, *sw, active, Lp, o);
      if (active) {
        n_surr += use ? 1u : 0u;
        n_fb += use ? 0u : 1u;
        drop_update<M, NF, R, POS64, AT_BOUNDS, false>(a, tile, t, s, Lp, use ? &o : nullptr, contrib, stayed);
      }
    }
    if constexpr (kAhead) {
      if (s + kUpdThreads < d1) drop_load<M, NF, R, POS64>(a, s + kUpdThreads, Lp);
    }
    __syncthreads();
    // stayers: rank within their cell in position order (cell owner threads)
    if (t < kTileCells) {
      const int32_t cend = (base + kUpdThreads < d1) ? base + kUpdThreads : d1;
      const int32_t lo = cs[t] > base ? cs[t] : base;
      const int32_t hi = cs[t + 1] < cend ? cs[t + 1] : cend;
      for (int32_t i = lo; i < hi; ++i)
        if (stayed[i - base]) a.arrival_out[i] = seg_tag | srank++;
    }
    // ordered per-cell reduction of this chunk
    const int32_t chunk_end = (base + kUpdThreads < d1) ? base + kUpdThreads : d1;
#pragma unroll
    for (int r = 0; r < NITEM; ++r) {
      const int item = t + r * kUpdThreads;
      if (item < kTileCells * NSRC) {
        const int c = item / NSRC, q = item % NSRC;
        const int32_t lo = cs[c] > base ? cs[c] : base;
        const int32_t hi = cs[c + 1] < chunk_end ? cs[c + 1] : chunk_end;
        double v = acc[r];
        if constexpr (EB) {
          // deposits sent to another cell by eb_interp are left out here
          for (int32_t i = lo; i < hi; ++i)
            if (!redir[i - base]) v = v + contrib[(i - base) * NSRC + q];
        } else if constexpr (BR) {
          for (int32_t i = lo; i < hi; ++i) v = v + contrib[(i - base) * NSRC + q];
        } else {
          const double cc = a.cur_coef;
          for (int32_t i = lo; i < hi; ++i) v = v + cc * static_cast<double>(contrib[(i - base) * NSRC + q]);
        }
        acc[r] = v;
      }
    }
    __syncthreads();
  }

  if (slot >= 0) {
    // heavy tile: this segment's sums and stayer counts, combined later
    double* P = a.sp.part + static_cast<long long>(slot + seg) * (kTileCells * NSRC);
#pragma unroll
    for (int r = 0; r < NITEM; ++r) {
      const int item = t + r * kUpdThreads;
      if (item < kTileCells * NSRC) P[item] = acc[r];
    }
    if (t < kTileCells) a.sp.seg_stay[(slot + seg) * kTileCells + t] = srank;
    if constexpr (SURR) __syncthreads();  // cs is rewritten by the next tile
    return;
  }
  if (t < kTileCells) a.stay_cnt[tile * kTileCells + t] = srank;

  // add each cell's sums to the source FAB: a thread's items are distinct
  // (cell, component) words (the host checks that the source components do
  // not overlap), so all their loads go out before the stores.  In the
  // overwrite mode every cell gets 0 + its sum (the value an add into a
  // zeroed FAB gives) and the FAB is not read.
  double* dst[NITEM];
  double old[NITEM];
#pragma unroll
  for (int r = 0; r < NITEM; ++r) {
    dst[r] = nullptr;
    const int item = t + r * kUpdThreads;
    if (item < kTileCells * NSRC) {
      const int c = item / NSRC, q = item % NSRC;
      if (a.src_overwrite || cs[c + 1] > cs[c]) dst[r] = src_word(a, otx, oty, otz, c, q);
    }
    old[r] = (dst[r] && !a.src_overwrite) ? *dst[r] : 0.;
  }
#pragma unroll
  for (int r = 0; r < NITEM; ++r) {
    if (dst[r]) {
      const int q = (t + r * kUpdThreads) % NSRC;
      *dst[r] = a.pelec ? old[r] + acc[r] * a.src_scale[q] : old[r] + acc[r];
    }
  }
  if constexpr (SURR) __syncthreads();  // cs is rewritten by the next tile
}

// TAB: the branch kernel for the TAB breakup model (do_breakup == 1), with
// the workgroup TAB stage.  DEP: the deposition species are read at run time
// (dep_fuel_species); the plain kernels are also compiled without it for
// the usual case of every fuel depositing into its own species, which
// leaves the run-time selects out.
template <class M, int NF, class R, bool POS64, bool AT_BOUNDS, bool BR = false, bool EB = false, bool TAB = false,
          bool DEP = true>
__global__ void __launch_bounds__(kUpdThreads, (UpdOccupancy<R, NF, TAB, BR>::kMinBlocks))
    spray_update_kernel(UpdateArgs<R, NF> a) {
  constexpr int NG = gas_stride<R, M::NS>();  // staged gas fields: rho, Y[NS], vel[3], T (padded)
  constexpr int NSRC = 5 + NF;
  __shared__ __attribute__((aligned(16))) R gas[NG * kHaloCells];
  __shared__ unsigned char gvalid[kHaloCells + 1];  // + the all-valid flag
  __shared__ int32_t cs[kTileCells + 1];
  __shared__ ContribT<R, BR> contrib[kUpdThreads * NSRC];
  __shared__ unsigned char stayed[kUpdThreads];
  __shared__ unsigned char redir[EB ? kUpdThreads : 1];
  // the TAB stage's records and bin offsets (branch kernels)
  __shared__ double tabrec[TAB ? TabRec<NF>::kN * kUpdThreads : 1];
  __shared__ int32_t tabcnt[TAB ? kTabBins + 1 : 1];
  const int ntiles = a.ntile[0] * a.ntile[1] * a.ntile[2];
  unsigned n_surr = 0, n_fb = 0;
  // workgroups 0 .. ntiles-1: the first segment of every tile; the others:
  // one further segment of a heavy tile each (those beyond the pass's work
  // items return at once)
  int tile, seg = 0;
  if (static_cast<int>(blockIdx.x) < ntiles) {
    tile = xcd_tile(blockIdx.x, ntiles);
  } else {
    const int j = blockIdx.x - ntiles;
    if (j >= a.sp.counters[0]) return;
    const int2 it = a.sp.items[j];
    tile = it.x;
    seg = it.y;
  }
  process_tile<M, NF, R, POS64, AT_BOUNDS, false, BR, EB, TAB, DEP>(a, tile, seg, a.sp.tile_slot[tile], gas, gvalid, cs,
                                                          contrib, stayed, nullptr, n_surr, n_fb, redir, tabrec,
                                                          tabcnt);
}

// The sums of the heavy tiles (SprayParticles.cpp:419-441): per cell and
// component, the segments' partial sums added in segment order, then added
// to the source FAB by the cell's owner; per cell, the stayer counts of the
// segments turned into offsets (for the re-sort) and their total.
template <class R, int NF>
__global__ void __launch_bounds__(kUpdThreads) split_combine_kernel(UpdateArgs<R, NF> a) {
  constexpr int NSRC = 5 + NF;
  const int t = threadIdx.x;
  const int nsplit = a.sp.counters[1];
  const int txy = a.ntile[0] * a.ntile[1];
  for (int i = blockIdx.x; i < nsplit; i += gridDim.x) {
    const int4 e = a.sp.list[i];
    const int tile = e.x, slot = e.y, nseg = e.z;
    const int otz = (tile / txy) * kTile;
    const int oty = ((tile % txy) / a.ntile[0]) * kTile;
    const int otx = (tile % a.ntile[0]) * kTile;
    for (int item = t; item < kTileCells * NSRC; item += kUpdThreads) {
      const int c = item / NSRC, q = item % NSRC;
      const bool has = a.cstart[tile * kTileCells + c + 1] > a.cstart[tile * kTileCells + c];
      if (!has && !a.src_overwrite) continue;
      double* dst = src_word(a, otx, oty, otz, c, q);
      if (!dst) continue;
      double total = 0.;
      if (has)
        for (int k = 0; k < nseg; ++k)
          total = total + a.sp.part[static_cast<long long>(slot + k) * (kTileCells * NSRC) + item];
      const double old = a.src_overwrite ? 0. : *dst;
      *dst = a.pelec ? old + total * a.src_scale[q] : old + total;
    }
    if (t < kTileCells) {
      int32_t off = 0;
      for (int k = 0; k < nseg; ++k) {
        int32_t* cnt = &a.sp.seg_stay[(slot + k) * kTileCells + t];
        const int32_t v = *cnt;
        *cnt = off;
        off += v;
      }
      a.stay_cnt[tile * kTileCells + t] = off;
    }
  }
}

// Surrogate form: gridDim.x workgroups (a multiple of 8) loop over the
// tiles; the network is staged once per workgroup.
template <class M, int NF>
__global__ void __launch_bounds__(kUpdThreads, 2) spray_update_surr_kernel(UpdateArgs<float, NF> a) {
  constexpr int NG = gas_stride<float, M::NS>();
  constexpr int NSRC = 5 + NF;
  __shared__ __attribute__((aligned(16))) float gas[NG * kHaloCells];
  __shared__ unsigned char gvalid[kHaloCells + 1];  // + the all-valid flag
  __shared__ int32_t cs[kTileCells + 1];
  __shared__ float contrib[kUpdThreads * NSRC];
  __shared__ unsigned char stayed[kUpdThreads];
  __shared__ SurrWeights sw;
  surr_stage(sw, a.surr_w, threadIdx.x, kUpdThreads);
  __syncthreads();
  const int ntiles = a.ntile[0] * a.ntile[1] * a.ntile[2];
  unsigned n_surr = 0, n_fb = 0;
  // the first segment of every tile, then the further segments of the
  // heavy tiles
  const int nwork = ntiles + a.sp.counters[0];
  for (int b = blockIdx.x; b < nwork; b += gridDim.x) {
    int tile, seg = 0;
    if (b < ntiles) {
      tile = xcd_tile(b, ntiles);
    } else {
      const int2 it = a.sp.items[b - ntiles];
      tile = it.x;
      seg = it.y;
    }
    process_tile<M, NF, float, false, false, true>(a, tile, seg, a.sp.tile_slot[tile], gas, gvalid, cs, contrib,
                                                   stayed, &sw, n_surr, n_fb);
  }
  // one pair of counter updates per wavefront
  unsigned long long s1 = n_surr, s2 = n_fb;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    s1 += __shfl_xor(s1, off);
    s2 += __shfl_xor(s2, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&a.surr_cnt[0], s1);
    atomicAdd(&a.surr_cnt[1], s2);
  }
}

}  // namespace pelemp
