"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from pelemp_amd import _capi as C

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class BranchParamsC(ctypes.Structure):
    """oracle_branch_params (oracle_capi.cpp)."""
    _fields_ = [("max_ppp", ctypes.c_double), ("ppp_fact", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("next_id", ctypes.c_int64), ("cap", ctypes.c_int64), ("khrt_B0", ctypes.c_double),
                ("khrt_B1", ctypes.c_double), ("khrt_C3", ctypes.c_double), ("avg_inject_mass", ctypes.c_double),
                ("pelec", ctypes.c_int64), ("state_ghosts", ctypes.c_int64), ("source_ghosts", ctypes.c_int64)]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oracle_update.restype = ctypes.c_int
        L.oracle_update.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(C.SprayDataC),
                                    ctypes.POINTER(C.SprayCompsC), ctypes.POINTER(C.GeomC), ctypes.c_int64, vp, vp,
                                    vp, ctypes.POINTER(C.FabViewC), ctypes.POINTER(C.FabViewC), ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(BranchParamsC), ctypes.POINTER(ctypes.c_int64)]
        L.oracle_pm_eval.restype = None
        L.oracle_compiler.restype = ctypes.c_char_p
        L.oracle_compiler.argtypes = []
        L.oracle_pm_eval.argtypes = [ctypes.c_int, ctypes.c_int64, vp, vp, vp]
        L.oracle_gas_props.restype = ctypes.c_int
        L.oracle_gas_props.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, vp, vp,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_calc_heat_coeff.restype = ctypes.c_double
        L.oracle_calc_heat_coeff.argtypes = [ctypes.c_double] * 5
        for nm in ("oracle_psat", "oracle_rhoL", "oracle_muL", "oracle_lambdaL"):
            getattr(L, nm).restype = ctypes.c_double
            getattr(L, nm).argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.c_double, ctypes.c_int]
        L.oracle_calc_boil_t.restype = ctypes.c_double
        L.oracle_calc_boil_t.argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int]
        L.oracle_trilinear.restype = None
        L.oracle_trilinear.argtypes = [vp, vp, vp, vp, vp]
        L.oracle_check_bounds.restype = ctypes.c_int
        L.oracle_check_bounds.argtypes = [vp, ctypes.POINTER(C.GeomC), vp]
        L.oracle_spray_source_one.restype = ctypes.c_double
        L.oracle_spray_source_one.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(C.SprayDataC),
                                              ctypes.c_double, vp, vp, vp, vp, vp]
        L.oracle_film_source_one.restype = ctypes.c_int
        L.oracle_film_source_one.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(C.SprayDataC),
                                             ctypes.c_double, vp, vp, vp, vp, vp]
        L.oracle_find_tangents.restype = None
        L.oracle_find_tangents.argtypes = [vp, vp, vp, vp]
        L.oracle_create_children.restype = ctypes.c_int
        L.oracle_create_children.argtypes = [ctypes.c_int, ctypes.POINTER(C.SprayDataC), ctypes.c_int, vp,
                                             ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                             ctypes.c_int64, ctypes.c_int, vp, vp, vp]
        L.oracle_tab_update.restype = ctypes.c_double
        L.oracle_tab_update.argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.c_double, ctypes.c_double, vp, vp, vp]
        L.oracle_khrt_update.restype = ctypes.c_int
        L.oracle_khrt_update.argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.c_double, ctypes.c_double, vp, vp,
                                         ctypes.c_int, vp, vp, vp, vp]
        L.oracle_transfer_source.restype = ctypes.c_int
        L.oracle_transfer_source.argtypes = [ctypes.POINTER(C.GeomC), ctypes.POINTER(C.SprayCompsC), vp, ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(C.FabViewC), ctypes.POINTER(C.FabViewC)]
        L.oracle_compute_derived.restype = ctypes.c_int
        L.oracle_compute_derived.argtypes = [ctypes.c_int, ctypes.POINTER(C.GeomC), ctypes.POINTER(C.SprayDataC),
                                             ctypes.c_int64, vp, vp, ctypes.POINTER(C.FabViewC), ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int]
        L.oracle_ey2t.restype = ctypes.c_double
        L.oracle_ey2t.argtypes = [ctypes.c_int, ctypes.c_double, vp, ctypes.c_double]
        L.oracle_est_timestep.restype = ctypes.c_double
        L.oracle_est_timestep.argtypes = [ctypes.POINTER(C.GeomC), ctypes.c_int, ctypes.c_int64, vp, ctypes.c_double,
                                          ctypes.c_double]
        L.oracle_splash_criteria.restype = ctypes.c_int
        L.oracle_splash_criteria.argtypes = [ctypes.c_double] * 5 + [vp]
        L.oracle_num_threads.restype = ctypes.c_int
        for nm in ("oracle_inject", "oracle_uniform_init"):
            getattr(L, nm).restype = ctypes.c_int
        L.oracle_inject.argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.POINTER(C.GeomC), ctypes.POINTER(C.SprayJetC),
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int64), vp, vp, vp]
        L.oracle_uniform_init.argtypes = [ctypes.POINTER(C.SprayDataC), ctypes.POINTER(C.GeomC), vp, vp,
                                          ctypes.c_double, ctypes.c_double, vp, ctypes.c_double, ctypes.c_int64,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_int64), vp, vp, vp]
        L.oracle_dist_sample.restype = ctypes.c_double
        L.oracle_dist_sample.argtypes = [ctypes.c_int, vp, ctypes.c_double, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_double)]
        L.oracle_set_eb.restype = None
        L.oracle_set_eb.argtypes = [ctypes.POINTER(C.EBViewC)]
        L.oracle_eb_interp.restype = ctypes.c_int
        L.oracle_eb_interp.argtypes = [ctypes.POINTER(C.GeomC), ctypes.c_double, vp, vp, vp, vp]
        L.oracle_last_redirects.restype = ctypes.c_int64
        L.oracle_inj_random.restype = ctypes.c_double
        L.oracle_inj_random.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        _lib = L
    return _lib


def compiler() -> str:
    """The compiler and version that built the oracle library."""
    return lib().oracle_compiler().decode()


def pm_eval(fn: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    lib().oracle_pm_eval(fn, x.size, x.ctypes.data, None if yy is None else yy.ctypes.data, out.ctypes.data)
    return out


def gas_props(mech: int, T: float, rho: float, Y):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    ns = Y.size
    cp, h, rhoD = np.empty(ns), np.empty(ns), np.empty(ns)
    mu, lam = ctypes.c_double(), ctypes.c_double()
    st = lib().oracle_gas_props(mech, T, rho, Y.ctypes.data, cp.ctypes.data, h.ctypes.data, rhoD.ctypes.data,
                                ctypes.byref(mu), ctypes.byref(lam))
    assert st == 0
    return dict(cp=cp, h=h, rhoD=rhoD, mu=mu.value, lam=lam.value)


def latent_shifted(sdata: dict, mech: int, ns: int) -> dict:
    """SprayData after spraySetup's latent shift (SpraySetup.cpp:284-290)."""
    d = dict(sdata)
    h = gas_props(mech, sdata["ref_T"], 1.0, np.full(ns, 1.0 / ns))["h"]
    d["latent"] = [sdata["ref_latent"][i] - h[sdata["indx"][i]] for i in range(sdata["nfuel"])]
    return d


def update(mech: int, sdata: dict, comps: dict, geom: dict, pos, rdata, ids, state: np.ndarray, state_lo,
           state_hi, source: np.ndarray, src_lo, src_hi, flow_dt: float, do_move: bool = True,
           spray_cfl: float = -1.0, is_virt: bool = False, is_ghost: bool = False, nthreads: int = 0,
           branch: dict | None = None, state_ghosts: int | None = None, source_ghosts: int | None = None):
    """One updateParticles call on host arrays; `source` is updated in place.
    Returns (status, pos, rdata, ids) with the surviving droplets in cell order
    (then the secondary droplets created by breakup / splash).  `branch`
    (max_ppp, ppp_fact, seed, next_id, cap, khrt_B0, khrt_B1, khrt_C3,
    avg_inject_mass, pelec) sets the breakup / splash parameters and the
    flavour (pelec=1: CGS SprayData without the latent shift, CGS geometry,
    droplets, state and sources); its next_id is updated.  state_ghosts /
    source_ghosts bound the gather and the deposits to the domain grown by
    that many cells (None: the FABs' boxes only)."""
    nf = int(sdata["nfuel"])
    n = np.asarray(ids).shape[0]
    bp = BranchParamsC(max_ppp=100.0, ppp_fact=0.5, seed=0, next_id=0, cap=n, khrt_B0=0.61, khrt_B1=7.0,
                       khrt_C3=1.0, avg_inject_mass=0.0,
                       state_ghosts=-1 if state_ghosts is None else int(state_ghosts),
                       source_ghosts=-1 if source_ghosts is None else int(source_ghosts))
    if branch is not None:
        for k in ("max_ppp", "ppp_fact", "seed", "next_id", "cap", "khrt_B0", "khrt_B1", "khrt_C3",
                  "avg_inject_mass", "pelec"):
            if k in branch:
                setattr(bp, k, branch[k])
    cap = max(int(bp.cap), n)
    bp.cap = cap
    pos_out = np.zeros((cap, 3))
    pos_out[:n] = np.asarray(pos, dtype=np.float64)
    rd_out = np.zeros((cap, np.asarray(rdata).shape[1]))
    rd_out[:n] = np.asarray(rdata, dtype=np.float64)
    id_out = np.zeros(cap, dtype=np.int64)
    id_out[:n] = np.asarray(ids, dtype=np.int64)
    pos, rdata, ids = pos_out, rd_out, id_out
    sd = C.spray_data_c(sdata)
    cc = C.comps_c(comps)
    gg = C.geom_c(geom)
    assert state.flags["C_CONTIGUOUS"] and source.flags["C_CONTIGUOUS"]
    sv = C.fab_view(state.ctypes.data, state_lo, state_hi, state.shape[0])
    dv = C.fab_view(source.ctypes.data, src_lo, src_hi, source.shape[0])
    nout = ctypes.c_int64()
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    st = lib().oracle_update(mech, nf, ctypes.byref(sd), ctypes.byref(cc), ctypes.byref(gg), n, pos.ctypes.data,
                             rdata.ctypes.data, ids.ctypes.data, ctypes.byref(sv), ctypes.byref(dv), flow_dt,
                             int(is_virt), int(is_ghost), int(do_move), float(spray_cfl), int(nthreads),
                             ctypes.byref(bp), ctypes.byref(nout))
    if branch is not None:
        branch["next_id"] = int(bp.next_id)
    m = nout.value
    return st, pos[:m].copy(), rdata[:m].copy(), ids[:m].copy()


def calc_heat_coeff(ratio, B_M, B_eps=1e-7, C_eps=1e-15, Nu_0=2.0) -> float:
    return lib().oracle_calc_heat_coeff(ratio, B_M, B_eps, C_eps, Nu_0)


def spray_source_one(mech: int, sdata: dict, flow_dt: float, gas, rdata, pid: int = 1):
    """calculateSpraySource for one droplet in the gas state
    gas = [u, v, w, T, rho, Y_0..Y_NS-1].  Returns (Re, rdata_out, id_out,
    sources [mom3, mass, Ydot(nf), eng], cBoilT)."""
    nf = int(sdata["nfuel"])
    sd = C.spray_data_c(sdata)
    gas = np.ascontiguousarray(gas, dtype=np.float64)
    rd = np.array(rdata, dtype=np.float64, copy=True)
    pid_c = np.array([pid], dtype=np.int64)
    src = np.zeros(5 + nf)
    cb = np.zeros(nf)
    Re = lib().oracle_spray_source_one(mech, nf, ctypes.byref(sd), flow_dt, gas.ctypes.data, rd.ctypes.data,
                                       pid_c.ctypes.data, src.ctypes.data, cb.ctypes.data)
    return Re, rd, int(pid_c[0]), src, cb


def film_source_one(mech: int, sdata: dict, flow_dt: float, gas, rdata, pid: int = 1):
    """calculateFilmSource for one wall-film particle in the gas state
    gas = [u, v, w, T, rho, Y_0..Y_NS-1].  Returns (rdata_out, id_out,
    sources [mom3, mass, Ydot(nf), eng], cBoilT)."""
    nf = int(sdata["nfuel"])
    sd = C.spray_data_c(sdata)
    gas = np.ascontiguousarray(gas, dtype=np.float64)
    rd = np.array(rdata, dtype=np.float64, copy=True)
    pid_c = np.array([pid], dtype=np.int64)
    src = np.zeros(5 + nf)
    cb = np.zeros(nf)
    st = lib().oracle_film_source_one(mech, nf, ctypes.byref(sd), flow_dt, gas.ctypes.data, rd.ctypes.data,
                                      pid_c.ctypes.data, src.ctypes.data, cb.ctypes.data)
    assert st == 0, st
    return rd, int(pid_c[0]), src, cb


def find_tangents(vel, norm):
    """find_tangents (AhamedSplash.H:22-48): (tanPsi, tanBeta)."""
    v = np.ascontiguousarray(vel, dtype=np.float64)
    n = np.ascontiguousarray(norm, dtype=np.float64)
    tp, tb = np.zeros(3), np.zeros(3)
    lib().oracle_find_tangents(v.ctypes.data, n.ctypes.data, tp.ctypes.data, tb.ctypes.data)
    return tp, tb


def create_children(sdata: dict, rec_type: int, norm, vel, loc, T0: float, ref_dia: float, num_dens: float,
                    phi1: float, phi2: float, phi3: float, Y0, parent_id: int = 1, sub_dt: float = 1e-6,
                    ppp_fact: float = 0.5, seed: int = 0, first_id: int = 1000, max_out: int = 4096):
    """CreateSBDroplets (SpraySB.cpp:9-207) for one secondary-droplet record;
    returns (pos, rdata, ids) of the children."""
    nf = int(sdata["nfuel"])
    sd = C.spray_data_c(sdata)
    rec = np.concatenate([norm, vel, loc, [T0, ref_dia, num_dens, phi1, phi2, phi3], Y0]).astype(np.float64)
    assert rec.size == 15 + nf
    pos = np.zeros((max_out, 3))
    rd = np.zeros((max_out, 10 + nf))
    ids = np.zeros(max_out, dtype=np.int64)
    m = lib().oracle_create_children(nf, ctypes.byref(sd), int(rec_type), rec.ctypes.data, parent_id, sub_dt,
                                     ppp_fact, seed, first_id, max_out, pos.ctypes.data, rd.ctypes.data,
                                     ids.ctypes.data)
    assert m >= 0, m
    return pos[:m].copy(), rd[:m].copy(), ids[:m].copy()


def tab_update(sdata: dict, Reyn: float, dt: float, gas, rdata):
    """updateBreakupTAB for one n-dodecane droplet (mechanism 0) in the gas
    state [u, v, w, T, rho, Y...]; returns (Utan_total, rdata_out, cBoilT)."""
    sd = C.spray_data_c(sdata)
    gas = np.ascontiguousarray(gas, dtype=np.float64)
    rd = np.array(rdata, dtype=np.float64, copy=True)
    cb = np.zeros(1)
    U = lib().oracle_tab_update(ctypes.byref(sd), Reyn, dt, gas.ctypes.data, rd.ctypes.data, cb.ctypes.data)
    return U, rd, float(cb[0])


def khrt_update(sdata: dict, Reyn: float, dt: float, gas, rdata, pos=(0.0, 0.0, 0.0), B0=0.61, B1=7.0, C3=1.0,
                avg_inject_mass=0.0, make_new_drops=True):
    """updateBreakupKHRT for one n-dodecane droplet (mechanism 0) in the gas
    state [u, v, w, T, rho, Y...]; returns (id, rdata_out, record, cBoilT) with
    record = [type, loc 3, vel 3, norm 3, num_dens, ref_dia, T0, Utan]."""
    sd = C.spray_data_c(sdata)
    gas = np.ascontiguousarray(gas, dtype=np.float64)
    rd = np.array(rdata, dtype=np.float64, copy=True)
    kp = np.array([B0, B1, C3, avg_inject_mass], dtype=np.float64)
    ps = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(14)
    cb = np.zeros(1)
    pid = lib().oracle_khrt_update(ctypes.byref(sd), Reyn, dt, gas.ctypes.data, kp.ctypes.data, int(make_new_drops),
                                   ps.ctypes.data, rd.ctypes.data, out.ctypes.data, cb.ctypes.data)
    return pid, rd, out, float(cb[0])


def transfer_source(geom: dict, comps: dict, dep_indx, source_ghosts: int, tmp: np.ndarray, tmp_lo, tmp_hi,
                    act: np.ndarray, act_lo, act_hi) -> int:
    """transferSource on host FABs [ncomp, nz, ny, nx] (updated in place):
    periodic SumBoundary of tmp, add into act (all components when the counts
    match, else the addSpraySrc map), zero tmp.  Returns the status."""
    assert tmp.flags["C_CONTIGUOUS"] and act.flags["C_CONTIGUOUS"]
    dep = np.ascontiguousarray(dep_indx, dtype=np.int32)
    tv = C.fab_view(tmp.ctypes.data, tmp_lo, tmp_hi, tmp.shape[0])
    av = C.fab_view(act.ctypes.data, act_lo, act_hi, act.shape[0])
    return lib().oracle_transfer_source(ctypes.byref(C.geom_c(geom)), ctypes.byref(C.comps_c(comps)),
                                        dep.ctypes.data, int(dep.size), int(source_ghosts), ctypes.byref(tv),
                                        ctypes.byref(av))


def compute_derived(sdata: dict, geom: dict, pos, rdata, mf: np.ndarray, mf_lo, mf_hi, start_indx: int = 0,
                    with_species: bool = True, pelec: bool = False) -> int:
    """computeDerivedVars into the host FAB mf[ncomp, nz, ny, nx] (added to,
    then the protected divides).  Returns the status."""
    nf = int(sdata["nfuel"])
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rdata = np.ascontiguousarray(rdata, dtype=np.float64)
    assert mf.flags["C_CONTIGUOUS"]
    mv = C.fab_view(mf.ctypes.data, mf_lo, mf_hi, mf.shape[0])
    return lib().oracle_compute_derived(nf, ctypes.byref(C.geom_c(geom)), ctypes.byref(C.spray_data_c(sdata)),
                                        pos.shape[0], pos.ctypes.data, rdata.ctypes.data, ctypes.byref(mv),
                                        int(start_indx), int(with_species), int(pelec))


def ey2t(mech: int, e: float, Y, T0: float) -> float:
    """Temperature from the specific internal energy e [J/kg] (EY2T)."""
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    return lib().oracle_ey2t(mech, float(e), Y.ctypes.data, float(T0))


def est_timestep(geom: dict, vel, cfl: float = 0.5, max_jet_vel: float = 0.0, fixed_parts: bool = False) -> float:
    """estTimestep over the droplet velocities vel[n, 3]."""
    v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    return lib().oracle_est_timestep(ctypes.byref(C.geom_c(geom)), int(fixed_parts), v.shape[0], v.ctypes.data,
                                     float(cfl), float(max_jet_vel))


def splash_criteria(Kv: float, Tstar: float, alpha: float, Kcrit: float, del_film: float):
    """splash_criteria: (splash_type, N_SB, ms, dm)."""
    out = np.zeros(3)
    t = lib().oracle_splash_criteria(Kv, Tstar, alpha, Kcrit, del_film, out.ctypes.data)
    return t, int(out[0]), float(out[1]), float(out[2])


def _particles_out(nf: int, run):
    """Call run(n_ptr, pos, rd, ids) with growing output arrays."""
    cap = 1024
    while True:
        n = ctypes.c_int64(cap)
        pos = np.zeros((cap, 3))
        rd = np.zeros((cap, 10 + nf))
        ids = np.zeros(cap, dtype=np.int64)
        st = run(ctypes.byref(n), pos.ctypes.data, rd.ctypes.data, ids.ctypes.data)
        if st == C.E_ARG and n.value > cap:
            cap = n.value
            continue
        m = n.value if st == 0 else 0
        return st, pos[:m].copy(), rd[:m].copy(), ids[:m].copy()


def inject(sdata: dict, geom: dict, jet: "C.SprayJetC", time: float, sim_dt: float, next_id: int,
           pelec: bool = False):
    """sprayInjection of one jet (the jet's state is updated in place):
    (status, pos, rdata, ids) of the new droplets."""
    sd, g = C.spray_data_c(sdata), C.geom_c(geom)
    return _particles_out(int(sdata["nfuel"]), lambda n, p, r, i: lib().oracle_inject(
        ctypes.byref(sd), ctypes.byref(g), ctypes.byref(jet), float(time), float(sim_dt), int(next_id),
        int(pelec), n, p, r, i))


def uniform_init(sdata: dict, geom: dict, num_part, vel, dia: float, T: float, Y, num_ppp: float, next_id: int,
                 pelec: bool = False):
    """uniformSprayInit on one rank: (status, pos, rdata, ids)."""
    sd, g = C.spray_data_c(sdata), C.geom_c(geom)
    npart = np.ascontiguousarray(num_part, dtype=np.int32)
    v = np.ascontiguousarray(vel, dtype=np.float64)
    Yv = np.ascontiguousarray(np.atleast_1d(Y), dtype=np.float64)
    return _particles_out(int(sdata["nfuel"]), lambda n, p, r, i: lib().oracle_uniform_init(
        ctypes.byref(sd), ctypes.byref(g), npart.ctypes.data, v.ctypes.data, float(dia), float(T), Yv.ctypes.data,
        float(num_ppp), int(next_id), int(pelec), n, p, r, i))


def dist_sample(dist_type: int, params, u1: float, u2: float = 0.5):
    """One diameter of a distribution from the uniforms (u1, u2), and its
    get_avg_dia."""
    p = np.ascontiguousarray(list(params) + [0.0] * (2 - len(params)), dtype=np.float64)
    avg = ctypes.c_double()
    d = lib().oracle_dist_sample(int(dist_type), p.ctypes.data, float(u1), float(u2), ctypes.byref(avg))
    return d, avg.value


def inj_random(seed: int, call: int, cand: int, slot: int) -> float:
    return lib().oracle_inj_random(int(seed), int(call), int(cand), int(slot))


_eb_keep = None


def set_eb(eb) -> None:
    """The EB (dict of host arrays flags, vfrac, ccent, bcent, bnorm, lo, hi)
    for the following update calls; None detaches it."""
    global _eb_keep
    if eb is None:
        lib().oracle_set_eb(None)
        _eb_keep = None
        return
    arrs = {k: np.ascontiguousarray(eb[k], dtype=(np.uint32 if k == "flags" else np.float64))
            for k in ("flags", "vfrac", "ccent", "bcent", "bnorm")}
    v = C.EBViewC()
    v.flags, v.vfrac, v.ccent = arrs["flags"].ctypes.data, arrs["vfrac"].ctypes.data, arrs["ccent"].ctypes.data
    v.bcent, v.bnorm = arrs["bcent"].ctypes.data, arrs["bnorm"].ctypes.data
    for d in range(3):
        v.lo[d] = int(eb["lo"][d])
        v.hi[d] = int(eb["hi"][d])
    _eb_keep = (arrs, v)
    lib().oracle_set_eb(ctypes.byref(v))


def eb_interp(geom: dict, pos, min_eb_vfrac: float = 0.0):
    """eb_interp for one droplet with the current EB: (do_fe_interp,
    cells[8, 3] (domain-relative), weights[8], deposit cell, position)."""
    g = C.geom_c(geom)
    p = np.ascontiguousarray(pos, dtype=np.float64).copy()
    ijkc = np.floor((p - np.array(geom["plo"])) / np.array(geom["dx"])).astype(np.int32)
    idx = np.zeros((8, 3), dtype=np.int32)
    w = np.zeros(8)
    r = lib().oracle_eb_interp(ctypes.byref(g), float(min_eb_vfrac), p.ctypes.data, ijkc.ctypes.data,
                               idx.ctypes.data, w.ctypes.data)
    return r, idx, w, ijkc, p


def last_redirects() -> int:
    return int(lib().oracle_last_redirects())
