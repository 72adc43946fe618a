"""Host-side mirror of PeleMP's SprayParticleContainer for the droplet update.

Method names and argument meanings follow the reference
(Source/PP_Spray/SprayParticles.H:38-415):

    updateParticles(level, state, source, flow_dt, time, state_ghosts,
                    source_ghosts, isVirt, isGhost, do_move, ltransparm,
                    spray_cfl_lev)                       SprayParticles.H:229-243
    moveKickDrift(state, source, level, dt, time, isVirtualPart, isGhostPart,
                  state_ghosts, source_ghosts, do_move, ltransparm,
                  spray_cfl_lev)                         SprayParticles.H:212-226
    moveKick(...)                                        SprayParticles.H:196-209
    estTimestep(level)                                   SprayParticles.H:110
    transferSource(source_ghosts, level, tmpSource,
                   actSource)                            SprayParticles.H:250-295
    getSprayData() / getSprayComps()                     SprayParticles.H:378-384
    sprayInjection(time, spray_jet, sim_dt, level)       SprayInjection.H:10-180
    uniformSprayInit(num_part, vel_part, dia_part, T_part, Y_part, level,
                     num_redist, num_ppp)                SprayInjection.H:198-348
    SprayParticleIO(level, is_checkpoint, dir)           SprayIO.cpp:6-98
    PostInitRestart(dir)                                 SprayIO.cpp:100-160

SprayJet mirrors Source/PP_Spray/SprayJet.{H,cpp} (constant or rate-shaped
jets with the five droplet-size distributions of Distribution/).

Every call goes through the C ABI of libpelemp_spray.so (HIP kernels for
gfx950).  `state` and `source` are device FABs (DeviceFab), the MultiFab of
one box.  Errors that the reference turns into amrex::Abort raise
PelempError with the matching code.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _capi as C


class DeviceFab:
    """One FAB in device memory: box [lo, hi] (incl. ghosts), ncomp components,
    component-major with x fastest (the MultiFab data of one box).  dtype
    float32 gives a single-precision FAB (a gas state kept in FP32, for
    updateParticles)."""

    def __init__(self, lo, hi, ncomp: int, data: Optional[np.ndarray] = None, dtype=np.float64):
        self.lo = tuple(int(v) for v in lo)
        self.hi = tuple(int(v) for v in hi)
        self.ncomp = int(ncomp)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError("a FAB holds float64 or float32 data")
        self.shape = (self.ncomp, self.hi[2] - self.lo[2] + 1, self.hi[1] - self.lo[1] + 1,
                      self.hi[0] - self.lo[0] + 1)
        self.buf = C.DeviceBuffer(self.dtype.itemsize * int(np.prod(self.shape)))
        if data is not None:
            if tuple(data.shape) != self.shape:
                raise ValueError(f"FAB data shape {data.shape} != {self.shape}")
            self.buf.upload(np.ascontiguousarray(data, dtype=self.dtype))
        else:
            self.buf.zero()
        self._view = C.fab_view(self.buf.ptr, self.lo, self.hi, self.ncomp, f32=self.is_f32)

    @property
    def is_f32(self) -> bool:
        return self.dtype == np.dtype(np.float32)

    def view(self):
        return self._view

    def numpy(self) -> np.ndarray:
        return self.buf.download(np.empty(self.shape, dtype=self.dtype))

    def setVal(self, v: float) -> None:
        if v != 0.0:
            self.buf.upload(np.full(self.shape, v, dtype=self.dtype))
        else:
            self.buf.zero()


class SprayJet:
    """A spray jet (SprayJet.{H,cpp}), built like the reference's direct
    constructor (SprayJet.cpp:95-141): angles in degrees, the distribution by
    name with its parameters (Uniform: diameter; Normal / LogNormal:
    mean_dia, std_dev; Weibull: mean_dia, k; ChiSquared: d32).  Units are
    the container's flavour units.  `seed` keys the counter-based random
    numbers of its injections."""

    def __init__(self, jet_name: str, spc: "SprayParticleContainer", jet_cent, jet_norm, spread_angle: float,
                 jet_dia: float, jet_vel: float, mass_flow: float, jet_temp: float, jet_Y=None,
                 dist_type: str = "Uniform", dist_params=(0.0, 0.0), start_time: float = 0.0,
                 end_time: float = 1.0e8, phi_swirl: float = 0.0, hollow_spray: bool = False,
                 hollow_spread: float = 0.0, inject_ppp: float = -1.0, seed: int = 0):
        self.m_jetName = jet_name
        self._spc = spc
        j = C.jet_c(jet_cent, jet_norm, spread_angle, jet_dia, jet_vel, mass_flow, jet_temp, jet_Y, dist_type,
                    dist_params, start_time, end_time, phi_swirl, hollow_spray, hollow_spread, inject_ppp, seed)
        self._j = j
        self._roi = None
        avg = ctypes.c_double()
        C.check(spc._lib.pelemp_spray_jet_check(spc._h, ctypes.byref(j), ctypes.byref(avg)), "SprayJet")
        self.m_avgDia = float(avg.value)

    def readROI(self, roi_file: str, rho_part: float, cd: float) -> None:
        """SprayJet::readROI (SprayJet.cpp:194-256): rate of injection from a
        file of 'time [ms];mdot' lines; sets start/end time and max_jet_vel."""
        mconv = 1.0e-3 if self._spc.flavour == "pelelmex" else 1.0
        n = ctypes.c_int32(0)
        lib = self._spc._lib
        st = lib.pelemp_spray_read_roi(roi_file.encode(), float(rho_part), float(cd), self._j.jet_dia, mconv,
                                       ctypes.byref(n), None, None, None, None, None, None)
        if st not in (C.OK, C.E_ARG) or (st == C.E_ARG and n.value == 0):
            C.check(st, "SprayJet.readROI")
        t, m, v = (np.zeros(max(n.value, 1)) for _ in range(3))
        t0, t1, vmax = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        C.check(lib.pelemp_spray_read_roi(roi_file.encode(), float(rho_part), float(cd), self._j.jet_dia, mconv,
                                          ctypes.byref(n), t.ctypes.data, m.ctypes.data, v.ctypes.data,
                                          ctypes.byref(t0), ctypes.byref(t1), ctypes.byref(vmax)), "SprayJet.readROI")
        self._roi = (t[:n.value].copy(), m[:n.value].copy(), v[:n.value].copy())
        j = self._j
        j.n_roi = n.value
        j.roi_time, j.roi_mdot, j.roi_vel = (a.ctypes.data for a in self._roi)
        j.start_time, j.end_time, j.max_jet_vel = t0.value, t1.value, vmax.value

    # accessors of SprayJet.H
    def jet_name(self) -> str:
        return self.m_jetName

    def num_ppp(self) -> float:
        return float(self._j.num_ppp)

    def set_num_ppp(self, v: float) -> None:
        self._j.num_ppp = float(v)

    def max_jet_vel(self) -> float:
        return float(self._j.max_jet_vel)

    def get_avg_dia(self) -> float:
        return self.m_avgDia

    def jet_active(self, time: float) -> bool:
        return not (time < self._j.start_time or time > self._j.end_time)

    def state(self) -> dict:
        """The public injection members (SprayJet.H:258-266)."""
        j = self._j
        return dict(num_ppp=j.num_ppp, m_sumInjMass=j.sum_inj_mass, m_sumInjTime=j.sum_inj_time,
                    m_minParcel=j.min_parcel, m_totalInjMass=j.total_inj_mass, m_totalInjTime=j.total_inj_time,
                    n_injections=j.n_injections)

    def reset_sum(self) -> None:
        self._j.sum_inj_mass = -1.0
        self._j.sum_inj_time = -1.0


def _stream_handle(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.handle)


class SprayParticleContainer:
    """Droplets of one AMR level, resident on the GPU."""

    def __init__(self, mech: int, spray_data: dict, comps: dict, geom: dict, max_particles: int,
                 form: int = C.FORM_FP64, apply_latent_shift: bool = True, flavour: str = "pelelmex"):
        """flavour "pelelmex" (SI, velocity and T in the state) or "pelec" (CGS
        throughout, rho*u and rho*E in the state, the SprayData latent heat
        unshifted)."""
        self._lib = C.lib()
        self.mech = mech
        self.form = form
        self.nf = int(spray_data["nfuel"])
        from .problem import mech_info

        species = mech_info(mech)["species"]
        self.fuel_names = [species[int(i)] for i in list(spray_data["indx"])[:self.nf]]
        self.geom = dict(geom)
        self.comps = dict(comps)
        self._sd = C.spray_data_c(spray_data)
        self._comps = C.comps_c(comps)
        self._geom = C.geom_c(geom)
        h = ctypes.c_void_p()
        self.flavour = flavour
        if flavour == "pelec":
            C.check(self._lib.pelemp_spray_create_pelec(ctypes.byref(h), mech, form, ctypes.byref(self._sd),
                                                        ctypes.byref(self._comps), ctypes.byref(self._geom),
                                                        int(max_particles)), "pelemp_spray_create_pelec")
        elif flavour == "pelelmex":
            C.check(self._lib.pelemp_spray_create(ctypes.byref(h), mech, form, ctypes.byref(self._sd),
                                                  ctypes.byref(self._comps), ctypes.byref(self._geom),
                                                  int(max_particles), 1 if apply_latent_shift else 0),
                    "pelemp_spray_create")
        else:
            raise ValueError(f"unknown flavour {flavour!r}")
        self._h = h
        self.m_sprayJets: list = []

    # -- lifetime ----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pelemp_spray_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data --------------------------------------------------------------
    @property
    def NSR(self) -> int:
        return 10 + self.nf

    def getSprayData(self) -> dict:
        out = C.SprayDataC()
        C.check(self._lib.pelemp_spray_get_data(self._h, ctypes.byref(out)), "pelemp_spray_get_data")
        return C.spray_data_dict(out)

    def getSprayComps(self) -> dict:
        return dict(self.comps)

    def set_params(self, max_parcel_size: float = 100.0, breakup_parcel_factor: float = 0.5,
                   rng_seed: int = 0, khrt_B0: float = 0.61, khrt_B1: float = 7.0, khrt_C3: float = 1.0,
                   avg_inject_mass: float = 0.0) -> None:
        """Container parameters of the breakup / splash models (max_parcel_size,
        breakup_parcel_factor, KHRT_B0/B1/C3; SpraySetup.cpp:15-19,138-160) and
        the mean injected parcel mass N d^3 the KH-RT model compares the shed
        mass with (SprayParticles.cpp:230-239)."""
        p = C.SprayParamsC()
        C.check(self._lib.pelemp_spray_get_params(self._h, ctypes.byref(p)), "pelemp_spray_get_params")
        p.max_parcel_size = float(max_parcel_size)
        p.breakup_parcel_factor = float(breakup_parcel_factor)
        p.rng_seed = int(rng_seed)
        p.khrt_B0 = float(khrt_B0)
        p.khrt_B1 = float(khrt_B1)
        p.khrt_C3 = float(khrt_C3)
        p.avg_inject_mass = float(avg_inject_mass)
        C.check(self._lib.pelemp_spray_set_params(self._h, ctypes.byref(p)), "pelemp_spray_set_params")

    def set_particles(self, pos: np.ndarray, rdata: np.ndarray, ids: np.ndarray) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rdata = np.ascontiguousarray(rdata, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        n = ids.shape[0]
        if pos.shape != (n, 3) or rdata.shape != (n, self.NSR):
            raise ValueError("particle arrays have the wrong shape")
        C.check(self._lib.pelemp_spray_set_particles(self._h, n, pos.ctypes.data, rdata.ctypes.data,
                                                     ids.ctypes.data), "pelemp_spray_set_particles")

    def num_particles(self) -> int:
        n = ctypes.c_int64()
        C.check(self._lib.pelemp_spray_num_particles(self._h, ctypes.byref(n)), "pelemp_spray_num_particles")
        return int(n.value)

    def get_particles(self):
        n = self.num_particles()
        pos = np.empty((n, 3))
        rdata = np.empty((n, self.NSR))
        ids = np.empty(n, dtype=np.int64)
        nn = ctypes.c_int64(n)
        C.check(self._lib.pelemp_spray_get_particles(self._h, ctypes.byref(nn), pos.ctypes.data,
                                                     rdata.ctypes.data, ids.ctypes.data),
                "pelemp_spray_get_particles")
        return pos, rdata, ids

    # -- the update ----------------------------------------------------------
    def updateParticles(self, level: int, state: DeviceFab, source: DeviceFab, flow_dt: float, time: float,
                        state_ghosts: int, source_ghosts: int, isVirt: bool, isGhost: bool, do_move: bool,
                        ltransparm=None, spray_cfl_lev: float = -1.0, stream=None, sync: bool = True) -> None:
        """SprayParticleContainer::updateParticles (SprayParticles.cpp:142-502).

        `ltransparm` is accepted for signature parity; transport data are
        compiled into the library.  With sync=False the call returns as soon
        as the kernels are queued; `sync()` then reports device-side errors.
        """
        del level, ltransparm
        if self.m_sprayJets and int(self._sd.do_breakup) == 2:
            # the KH-RT parcel-mass threshold from the jets (SprayParticles.cpp:230-239)
            p = C.SprayParamsC()
            C.check(self._lib.pelemp_spray_get_params(self._h, ctypes.byref(p)), "pelemp_spray_get_params")
            p.avg_inject_mass = self.avg_inject_mass()
            C.check(self._lib.pelemp_spray_set_params(self._h, ctypes.byref(p)), "pelemp_spray_set_params")
        sv, dv = state.view(), source.view()
        h = _stream_handle(stream)
        # a single-precision state goes through the FP32-state entry point
        fn = self._lib.pelemp_spray_update_f32gas if getattr(state, "is_f32", False) else self._lib.pelemp_spray_update
        C.check(fn(self._h, ctypes.byref(sv), ctypes.byref(dv), float(flow_dt), float(time), int(state_ghosts),
                   int(source_ghosts), int(isVirt), int(isGhost), int(do_move), float(spray_cfl_lev), h),
                "pelemp_spray_update")
        if sync:
            self.sync(stream)

    def moveKickDrift(self, state: DeviceFab, source: DeviceFab, level: int, dt: float, time: float,
                      isVirtualPart: bool, isGhostPart: bool, state_ghosts: int, source_ghosts: int,
                      do_move: bool = True, ltransparm=None, spray_cfl_lev: float = -1.0, stream=None,
                      sync: bool = True) -> None:
        """SprayParticleContainer::moveKickDrift (SprayParticles.cpp:59-87)."""
        self.updateParticles(level, state, source, dt, time, state_ghosts, source_ghosts, isVirtualPart,
                             isGhostPart, do_move, ltransparm, spray_cfl_lev, stream, sync)

    def moveKick(self, state: DeviceFab, source: DeviceFab, level: int, dt: float, time: float,
                 isVirtualPart: bool, isGhostPart: bool, state_ghosts: int, source_ghosts: int,
                 ltransparm=None, spray_cfl_lev: float = -1.0, stream=None, sync: bool = True) -> None:
        """SprayParticleContainer::moveKick (SprayParticles.cpp:37-57): update without moving."""
        self.updateParticles(level, state, source, dt, time, state_ghosts, source_ghosts, isVirtualPart,
                             isGhostPart, False, ltransparm, spray_cfl_lev, stream, sync)

    def set_source_mode(self, mode: int) -> None:
        """C.SOURCE_ACCUMULATE (+=, the reference's) or C.SOURCE_OVERWRITE (the
        update writes the source's domain cells; no zero fill needed)."""
        C.check(self._lib.pelemp_spray_set_source_mode(self._h, int(mode)), "pelemp_spray_set_source_mode")

    def sync(self, stream=None) -> None:
        C.check(self._lib.pelemp_spray_sync(self._h, _stream_handle(stream)), "device check")

    def estTimestep(self, level: int = 0, cfl: float = 0.5, max_jet_vel: float = 0.0, stream=None) -> float:
        """SprayParticleContainer::estTimestep (SprayParticles.cpp:89-140):
        cfl = spray_cfl; max_jet_vel is the largest velocity of the spray jets
        (the reference reads it from its jets; 0 for none)."""
        del level
        dt = ctypes.c_double()
        C.check(self._lib.pelemp_spray_est_timestep(self._h, float(cfl), float(max_jet_vel), ctypes.byref(dt),
                                                    _stream_handle(stream)), "pelemp_spray_est_timestep")
        return float(dt.value)

    DERIVE_NAMES = ("spray_mass", "spray_density", "spray_num", "spray_vol", "spray_surf_area", "spray_vol_frac",
                    "d10", "d32", "wall_film_hght", "wall_film_mass", "spray_temp", "num_parcels", "spray_x_vel",
                    "spray_y_vel", "spray_z_vel")

    def DeriveVarNames(self, derive_plot_species: int = 1) -> list:
        """Names of the derived spray variables (SpraySetup.cpp:213-239)."""
        names = list(self.DERIVE_NAMES)
        if derive_plot_species == 1 and len(self.fuel_names) > 1:
            names += ["spray_mass_" + f for f in self.fuel_names]
        return names

    def NumDeriveVars(self, derive_plot_species: int = 1) -> int:
        n = ctypes.c_int32()
        C.check(self._lib.pelemp_spray_derive_count(self._h, int(derive_plot_species), ctypes.byref(n)),
                "pelemp_spray_derive_count")
        return int(n.value)

    def computeDerivedVars(self, mf_var: "DeviceFab", level: int = 0, start_indx: int = 0,
                           derive_plot_species: int = 1, stream=None, sync: bool = True) -> None:
        """SprayParticleContainer::computeDerivedVars (SprayDerive.cpp:9-155):
        adds the per-cell spray diagnostics to mf_var from start_indx on."""
        del level
        v = mf_var.view()
        C.check(self._lib.pelemp_spray_compute_derived(self._h, ctypes.byref(v), int(start_indx),
                                                       int(derive_plot_species), _stream_handle(stream)),
                "pelemp_spray_compute_derived")
        if sync:
            C.device_synchronize()

    def transferSource(self, source_ghosts: int, level: int, tmpSource: "DeviceFab", actSource: "DeviceFab",
                       stream=None, sync: bool = True) -> None:
        """SprayParticleContainer::transferSource (SprayParticles.H:250-295):
        periodic SumBoundary of tmpSource, add into actSource (every component,
        or the addSpraySrc map when the component counts differ), then
        tmpSource.setVal(0)."""
        del level
        tv, av = tmpSource.view(), actSource.view()
        C.check(self._lib.pelemp_spray_transfer_source(self._h, int(source_ghosts), ctypes.byref(tv), ctypes.byref(av),
                                                       _stream_handle(stream)), "pelemp_spray_transfer_source")
        if sync:
            C.device_synchronize()

    # -- embedded boundary -----------------------------------------------------
    def set_eb(self, eb: Optional[dict]) -> None:
        """Attach the level's EB (dict of numpy arrays flags (uint32), vfrac,
        ccent[3], bcent[3], bnorm[3] over the box lo..hi, component-major), the
        data AMReX's EBFArrayBoxFactory holds; None detaches it.  The arrays
        are copied to the device and kept with the container."""
        if eb is None:
            C.check(self._lib.pelemp_spray_set_eb(self._h, None), "pelemp_spray_set_eb")
            self._eb_bufs = None
            return
        bufs = {}
        for k, dt in (("flags", np.uint32), ("vfrac", np.float64), ("ccent", np.float64), ("bcent", np.float64),
                      ("bnorm", np.float64)):
            a = np.ascontiguousarray(eb[k], dtype=dt)
            b = C.DeviceBuffer(a.nbytes)
            b.upload(a)
            bufs[k] = b
        v = C.EBViewC()
        v.flags, v.vfrac, v.ccent = bufs["flags"].ptr, bufs["vfrac"].ptr, bufs["ccent"].ptr
        v.bcent, v.bnorm = bufs["bcent"].ptr, bufs["bnorm"].ptr
        for d in range(3):
            v.lo[d] = int(eb["lo"][d])
            v.hi[d] = int(eb["hi"][d])
        C.check(self._lib.pelemp_spray_set_eb(self._h, ctypes.byref(v)), "pelemp_spray_set_eb")
        self._eb_bufs = bufs

    # -- injection and initialisation ----------------------------------------
    def add_particles(self, pos: np.ndarray, rdata: np.ndarray, ids: np.ndarray) -> None:
        """Append droplets (the tiles' host-particle copy of SprayInjection.H:
        164-179), e.g. from a host-side injector."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rdata = np.ascontiguousarray(rdata, dtype=np.float64)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        n = ids.shape[0]
        if pos.shape != (n, 3) or rdata.shape != (n, self.NSR):
            raise ValueError("particle arrays have the wrong shape")
        C.check(self._lib.pelemp_spray_add_particles(self._h, n, pos.ctypes.data, rdata.ctypes.data,
                                                     ids.ctypes.data), "pelemp_spray_add_particles")

    def sprayInjection(self, time: float, spray_jet: SprayJet, sim_dt: float, level: int = 0,
                       stream=None) -> int:
        """SprayParticleContainer::sprayInjection (SprayInjection.H:10-180) on
        the device; returns the number of droplets injected."""
        del level
        n = ctypes.c_int64()
        C.check(self._lib.pelemp_spray_inject(self._h, ctypes.byref(spray_jet._j), float(time), float(sim_dt),
                                              ctypes.byref(n), _stream_handle(stream)), "sprayInjection")
        if spray_jet not in self.m_sprayJets:
            self.m_sprayJets.append(spray_jet)
        return int(n.value)

    def uniformSprayInit(self, num_part, vel_part, dia_part: float, T_part: float, Y_part, level: int = 0,
                         num_redist: int = -1, num_ppp: float = 1.0) -> None:
        """SprayParticleContainer::uniformSprayInit (SprayInjection.H:198-348),
        one rank; num_redist only schedules the reference's rank-by-rank copy."""
        del level, num_redist
        npart = (ctypes.c_int32 * 3)(*[int(x) for x in num_part])
        vel = (ctypes.c_double * 3)(*[float(x) for x in vel_part])
        Y = np.ascontiguousarray(np.atleast_1d(np.asarray(Y_part, dtype=np.float64)))
        if Y.shape[0] != self.nf:
            raise ValueError("Y_part needs one value per fuel")
        C.check(self._lib.pelemp_spray_uniform_init(self._h, npart, vel, float(dia_part), float(T_part),
                                                    Y.ctypes.data, float(num_ppp)), "uniformSprayInit")

    def avg_inject_mass(self) -> float:
        """sum over jets of N d_avg^3 / n_jets (SprayParticles.cpp:230-239)."""
        if not self.m_sprayJets:
            return 0.0
        nj = float(len(self.m_sprayJets))
        return sum(js.num_ppp() * js.get_avg_dia() ** 3 / nj for js in self.m_sprayJets)

    # -- on-disk formats ------------------------------------------------------
    write_ascii_files = False

    def SprayParticleIO(self, level: int, is_checkpoint: bool, dir: str, fuel_names=None) -> None:
        """SprayParticleContainer::SprayParticleIO (SprayIO.cpp:6-98): the
        particle checkpoint under dir/particles, the spray<N>.p3d ASCII file
        next to dir when write_ascii_files, and injection_data.log of the jets
        in a checkpoint."""
        import os
        import re

        names = list(fuel_names) if fuel_names is not None else list(self.fuel_names)
        arr = (ctypes.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        C.check(self._lib.pelemp_spray_checkpoint(self._h, dir.encode(), len(names), arr), "SprayParticleIO")
        if level == 0 and self.write_ascii_files:
            m = re.search(r"(\d+)[^\d]*$", dir)
            num = m.group(1) if m else ""
            fname = os.path.join(os.path.dirname(dir), "spray" + num + ".p3d")
            C.check(self._lib.pelemp_spray_write_ascii(self._h, fname.encode()), "WriteAsciiFile")
        if is_checkpoint and self.m_sprayJets:
            os.makedirs(os.path.join(dir, "particles"), exist_ok=True)
            self.write_injection_log(os.path.join(dir, "particles", "injection_data.log"))

    def write_injection_log(self, path: str) -> None:
        jets = (C.SprayJetC * len(self.m_sprayJets))(*[js._j for js in self.m_sprayJets])
        nm = (ctypes.c_char_p * len(self.m_sprayJets))(*[js.m_jetName.encode() for js in self.m_sprayJets])
        C.check(self._lib.pelemp_spray_write_injection_log(path.encode(), len(self.m_sprayJets), nm, jets),
                "write injection log")

    def Restart(self, dir: str) -> None:
        """Restart(dir, "particles") (SpraySetup.cpp:320-325): the level-0
        particles of an AMReX particle checkpoint under dir/particles, written
        by SprayParticleIO or by PeleMP itself."""
        C.check(self._lib.pelemp_spray_restart(self._h, dir.encode()), "Restart")

    def PostInitRestart(self, dir: str = "") -> None:
        """SprayParticleContainer::PostInitRestart (SprayIO.cpp:100-160): the
        jets' injection state from dir/particles/injection_data.log."""
        import os

        if not self.m_sprayJets:
            return
        n = len(self.m_sprayJets)
        jets = (C.SprayJetC * n)(*[js._j for js in self.m_sprayJets])
        nm = (ctypes.c_char_p * n)(*[js.m_jetName.encode() for js in self.m_sprayJets])
        path = os.path.join(dir, "particles", "injection_data.log")
        C.check(self._lib.pelemp_spray_read_injection_log(path.encode(), n, nm, jets), "PostInitRestart")
        for js, jc in zip(self.m_sprayJets, jets):
            for f in ("num_ppp", "sum_inj_mass", "sum_inj_time", "min_parcel", "total_inj_mass", "total_inj_time"):
                setattr(js._j, f, getattr(jc, f))

    # -- timing ---------------------------------------------------------------
    def pass_counts(self) -> list:
        """Active droplets at the start of each update pass since the last call."""
        buf = (ctypes.c_int64 * 4096)()
        n = ctypes.c_int32()
        C.check(self._lib.pelemp_spray_pass_counts(self._h, buf, 4096, ctypes.byref(n)), "pelemp_spray_pass_counts")
        return [int(buf[i]) for i in range(n.value)]

    def set_timing(self, enable: bool) -> None:
        C.check(self._lib.pelemp_spray_set_timing(self._h, 1 if enable else 0), "pelemp_spray_set_timing")

    def get_timing(self):
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_int64 * 4)()
        C.check(self._lib.pelemp_spray_get_timing(self._h, ms, n, 4), "pelemp_spray_get_timing")
        return list(ms), list(n)

    def device_bytes(self) -> int:
        b = ctypes.c_int64()
        C.check(self._lib.pelemp_spray_device_bytes(self._h, ctypes.byref(b)), "pelemp_spray_device_bytes")
        return int(b.value)
