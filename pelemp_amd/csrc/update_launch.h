// update_launch.h -- host-side launch of the fused update kernel for one
// arithmetic form; included by spray_fp64.hip and spray_fp32.hip, which are
// compiled with different floating-point contraction settings.
#pragma once

#include "spray_ctx.h"
#include "update_kernel.h"

namespace pelemp {

inline FabDev make_fab(const pelemp_fab_view& v) {
  FabDev f;
  f.data = v.data;
  f.data32 = nullptr;
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = v.lo[d];
    f.n[d] = v.hi[d] - v.lo[d] + 1;
  }
  f.cstride = static_cast<long long>(f.n[0]) * f.n[1] * f.n[2];
  return f;
}

template <class M, int NF, class R, bool POS64, bool SURR = false>
int launch_update_t(Ctx& c, const PassParams& p, hipStream_t s) {
  UpdateArgs<R, NF> a{};
  a.fin = static_cast<const R*>(c.fbuf[c.cur]);
  a.fout = static_cast<R*>(c.fbuf[1 - c.cur]);
  a.min = c.mbuf[c.cur];
  a.mout = c.mbuf[1 - c.cur];
  a.cap = c.cap;
  a.perm = c.perm;
  a.cstart = c.cstart;
  a.slot_in = c.slot[c.cur];
  a.slot_out = c.slot[1 - c.cur];
  a.key_out = c.key;
  a.arrival_out = c.arrival;
  a.mov_lo = c.mov_lo;
  a.mov_hi = c.mov_hi;
  a.stay_cnt = c.stay_cnt;
  a.st = make_fab(*p.state);
  a.st.data32 = p.state32;
  a.c_rho = c.comps.rhoIndx;
  a.c_vel = c.comps.momIndx;
  a.c_spec = c.comps.specIndx;
  a.c_T = c.comps.utempIndx;
  a.src = make_fab(*p.source);
  a.c_rsrc = c.comps.rhoSrcIndx;
  a.c_msrc = c.comps.momSrcIndx;
  a.c_ssrc = c.comps.specSrcIndx;
  a.c_esrc = c.comps.engSrcIndx;
  for (int d = 0; d < 3; ++d) {
    a.plo[d] = c.geom.plo[d];
    a.phi[d] = c.geom.phi[d];
    a.dx[d] = c.geom.dx[d];
    a.dxi[d] = 1. / c.geom.dx[d];
    a.dom_lo[d] = c.geom.domain_lo[d];
    a.ncell[d] = c.ncell[d];
    a.ntile[d] = c.ntile[d];
    a.bndry_lo[d] = c.geom.bndry_lo[d];
    a.bndry_hi[d] = c.geom.bndry_hi[d];
    a.sub_dt_dxi[d] = static_cast<R>(p.sub_dt * a.dxi[d]);
  }
  a.fc = make_fuel_const<R, NF>(c.sd, c.mw.data());
  a.fc.pelec = c.pelec;
  a.pelec = c.pelec ? 1 : 0;
  a.c_eng = c.comps.engIndx;
  a.conv = c.pelec_conv;
  for (int q = 0; q < 5 + PELEMP_MAX_FUEL; ++q) a.src_scale[q] = c.src_scale[q];
  a.flow_dt = static_cast<R>(p.flow_dt);
  a.sub_dt = static_cast<R>(p.sub_dt);
  a.cur_coef = p.cur_coef;
  a.do_move = p.do_move;
  a.src_overwrite = p.src_overwrite;
  a.is_ghost = p.is_ghost;
  a.st_g = p.state_ghosts;
  a.src_g = p.source_ghosts;
  a.err = c.err;
  a.sp = c.sp;
  // one workgroup per tile, then one per further segment of a heavy tile
  // (the workgroups beyond the pass's work items return at once)
  const dim3 grid(c.ntiles + c.split_items_cap), block(kUpdThreads);
  if constexpr (SURR) {
    const pelemp_surrogate_norm& n = c.surr_norm;
    for (int q = 0; q < kSurrIn; ++q) {
      a.sn.c[q] = n.in_center[q];
      a.sn.s[q] = n.in_scale[q];
    }
    for (int q = 0; q < kSurrOut; ++q) {
      a.sn.mean[q] = n.out_mean[q];
      a.sn.sd[q] = n.out_std[q];
    }
    a.hull = n;
    a.surr_w = static_cast<const SurrWeights*>(c.surr_w);
    a.surr_cnt = c.surr_cnt;
    if (p.at_bounds) {
      set_last_error("walls are supported by the FP64 form only");
      return PELEMP_E_BAD_CONFIG;
    }
    // two workgroups per CU on 256 CUs, each looping over tiles
    const int nwg = c.ntiles < 512 ? c.ntiles : 512;
    hipLaunchKernelGGL((spray_update_surr_kernel<M, NF>), dim3(nwg), block, 0, s, a);
  } else if (p.branches) {
    if constexpr (POS64) {
      for (int q = 0; q < 4 * PELEMP_MAX_FUEL; ++q) {
        a.lx.lambda_coef[q] = c.sd.lambda_coef[q];
        a.lx.mu_coef[q] = c.sd.mu_coef[q];
      }
      a.lx.sigma = c.sd.sigma;
      a.lx.wall_T = c.sd.wall_T;
      a.lx.theta_c = c.sd.theta_c;
      a.wf = c.wf;
      a.sb_type = c.sb_type;
      a.sb_data = c.sb_data;
      a.do_breakup = c.sd.do_breakup;
      a.kh.B0 = c.params.khrt_B0;
      a.kh.B1 = c.params.khrt_B1;
      a.kh.C3 = c.params.khrt_C3;
      a.kh.avg_inject_mass = c.params.avg_inject_mass;
      a.kh.dt = c.sd.dtmod * p.flow_dt;
      a.make_new_drops = p.make_new_drops;
      a.do_splash_box = p.do_splash_box;
      a.max_ppp = c.params.max_parcel_size;
      a.inv_vol = 1. / (c.geom.dx[0] * c.geom.dx[1] * c.geom.dx[2]);
      a.eb_on = c.eb_on ? 1 : 0;
      if (c.eb_on) {
        a.eb.flags = c.eb.flags;
        a.eb.vfrac = c.eb.vfrac;
        a.eb.ccent = c.eb.ccent;
        a.eb.bcent = c.eb.bcent;
        a.eb.bnorm = c.eb.bnorm;
        for (int d = 0; d < 3; ++d) {
          a.eb.lo[d] = c.eb.lo[d] - c.geom.domain_lo[d];
          a.eb.n[d] = c.eb.hi[d] - c.eb.lo[d] + 1;
        }
        a.eb.cstride = static_cast<long long>(a.eb.n[0]) * a.eb.n[1] * a.eb.n[2];
        a.min_eb_vfrac = c.sd.min_eb_vfrac;
        a.eb_rkey = c.eb_rkey;
        a.eb_rval = c.eb_rval;
      }
      // the TAB model runs the kernel with the workgroup TAB stage
      const bool tab = c.sd.do_breakup == 1;
      if (c.eb_on) {
        if (p.at_bounds) {
          if (tab) {
            hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, true, true, true, true>), grid, block, 0, s, a);
          } else {
            hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, true, true, true>), grid, block, 0, s, a);
          }
        } else if (tab) {
          hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, false, true, true, true>), grid, block, 0, s, a);
        } else {
          hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, false, true, true>), grid, block, 0, s, a);
        }
      } else if (p.at_bounds) {
        if (tab) {
          hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, true, true, false, true>), grid, block, 0, s, a);
        } else {
          hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, true, true>), grid, block, 0, s, a);
        }
      } else if (tab) {
        hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, false, true, false, true>), grid, block, 0, s, a);
      } else {
        hipLaunchKernelGGL((spray_update_kernel<M, NF, double, true, false, true>), grid, block, 0, s, a);
      }
    } else {
      set_last_error("breakup, splash and wall films are supported by the FP64 form only");
      return PELEMP_E_BAD_CONFIG;
    }
  } else if (p.at_bounds) {
    if constexpr (POS64) {
      hipLaunchKernelGGL((spray_update_kernel<M, NF, R, POS64, true>), grid, block, 0, s, a);
    } else {
      set_last_error("walls are supported by the FP64 form only");
      return PELEMP_E_BAD_CONFIG;
    }
  } else if (a.fc.dep_own) {
    hipLaunchKernelGGL((spray_update_kernel<M, NF, R, POS64, false, false, false, false, false>), grid, block, 0, s,
                       a);
  } else {
    hipLaunchKernelGGL((spray_update_kernel<M, NF, R, POS64, false>), grid, block, 0, s, a);
  }
  // the heavy tiles' segment sums, added to the source in segment order
  const int cgrid = c.split_list_cap < 256 ? c.split_list_cap : 256;
  hipLaunchKernelGGL((split_combine_kernel<R, NF>), dim3(cgrid), block, 0, s, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_last_error(std::string("update launch: ") + hipGetErrorString(e));
    return PELEMP_E_HIP;
  }
  return PELEMP_OK;
}

template <class M>
int launch_pelec_convert_t(Ctx& c, const pelemp_fab_view& st, hipStream_t s) {
  const long long ncells = static_cast<long long>(st.hi[0] - st.lo[0] + 1) * (st.hi[1] - st.lo[1] + 1) *
                           (st.hi[2] - st.lo[2] + 1);
  const size_t bytes = 8 * static_cast<size_t>(5 + M::NS) * static_cast<size_t>(ncells);
  if (bytes > c.pelec_conv_bytes) {
    if (c.pelec_conv) (void)hipFree(c.pelec_conv);
    c.pelec_conv = nullptr;
    c.pelec_conv_bytes = 0;
    if (hipMalloc(&c.pelec_conv, bytes) != hipSuccess) {
      set_last_error("hipMalloc failed (PeleC state conversion)");
      return PELEMP_E_NOMEM;
    }
    c.pelec_conv_bytes = bytes;
  }
  const unsigned nb = static_cast<unsigned>(std::min<long long>((ncells + 255) / 256, 65535));
  hipLaunchKernelGGL(pelec_convert_kernel<M>, dim3(nb), dim3(256), 0, s, st.data, ncells, c.comps.rhoIndx,
                     c.comps.specIndx, c.comps.momIndx, c.comps.engIndx, c.comps.utempIndx, c.pelec_conv);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_last_error(std::string("PeleC state conversion: ") + hipGetErrorString(e));
    return PELEMP_E_HIP;
  }
  return PELEMP_OK;
}

template <class R, bool POS64, bool SURR = false>
int dispatch_update(Ctx& c, const PassParams& p, hipStream_t s) {
  if constexpr (SURR) {
    if (c.mech == PELEMP_MECH_DODECANE_AIR && c.nf == 1)
      return launch_update_t<Mech_dodecane_air, 1, float, false, true>(c, p, s);
    if (c.mech == PELEMP_MECH_DECANE_DODECANE_AIR && c.nf == 2)
      return launch_update_t<Mech_decane_dodecane_air, 2, float, false, true>(c, p, s);
    set_last_error("the surrogate form supports mechanisms 0 (one fuel) and 1 (two fuels)");
    return PELEMP_E_BAD_CONFIG;
  }
  if (c.mech == PELEMP_MECH_DODECANE_AIR && c.nf == 1)
    return launch_update_t<Mech_dodecane_air, 1, R, POS64>(c, p, s);
  if (c.mech == PELEMP_MECH_DECANE_DODECANE_AIR && c.nf == 2)
    return launch_update_t<Mech_decane_dodecane_air, 2, R, POS64>(c, p, s);
  if (c.mech == PELEMP_MECH_ALKANE6_AIR && c.nf == 6)
    return launch_update_t<Mech_alkane6_air, 6, R, POS64>(c, p, s);
  if (c.mech == PELEMP_MECH_DECANE_AIR && c.nf == 1)
    return launch_update_t<Mech_decane_air, 1, R, POS64>(c, p, s);
  set_last_error("unsupported mechanism / fuel count combination");
  return PELEMP_E_BAD_CONFIG;
}

}  // namespace pelemp
