"""Seeded synthetic spray problems (gas state FAB + droplets) for the
benchmark configurations and the tests.

The gas recipe follows the configuration text: a periodic box of air
(Y_O2 = 0.233, Y_N2 = 0.767, Exec/SprayTests/PeleC/abramzon_test/
prob_parm.H:13-14), temperature T0 (1 + 0.05 sin(2 pi x/L) sin(2 pi y/L)
sin(2 pi z/L)), velocity (U0, 0, 0) + 0.5 (sin 2 pi y/L, sin 2 pi z/L,
sin 2 pi x/L) m/s, density from the ideal gas law at pressure p.  Ghost cells
are periodic copies of interior cells.  Droplets: uniform positions (config
1) or Poisson counts per cell (config 2), lognormal diameters, T_d = 300 K,
velocity = local gas velocity + N(0, 1) m/s per component.

State layout (PeleLMeX flavour): rho, velocity(3), rho Y_k (NS), rho h
(unused by the spray), T.  Spray source layout: rho, momentum(3), rho Y_f
(nf), energy.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi as C
from .fuels import spray_data

RU = 8.31446261815324
PATM = 101325.0

_MECH_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mech", "mech_data.json")
MECH_NAMES = {C.MECH_DODECANE_AIR: "dodecane_air", C.MECH_DECANE_DODECANE_AIR: "decane_dodecane_air",
              C.MECH_ALKANE6_AIR: "alkane6_air", C.MECH_DECANE_AIR: "decane_air"}


def mech_info(mech: int) -> dict:
    with open(_MECH_JSON) as f:
        return json.load(f)[MECH_NAMES[mech]]


@dataclass
class SprayProblem:
    name: str
    mech: int
    species: list
    W: np.ndarray
    nf: int
    geom: dict
    comps: dict
    sdata: dict
    state: np.ndarray  # (ncomp, nz, ny, nx) incl. ghosts
    state_lo: tuple
    state_hi: tuple
    src_shape: tuple
    src_lo: tuple
    src_hi: tuple
    pos: np.ndarray  # (n, 3)
    rdata: np.ndarray  # (n, 10 + nf)
    ids: np.ndarray  # (n,)
    flow_dt: float
    nsteps: int
    state_ghosts: int = 2
    source_ghosts: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def nsr(self) -> int:
        return 10 + self.nf

    def new_source(self) -> np.ndarray:
        return np.zeros(self.src_shape, dtype=np.float64)

    @property
    def n(self) -> int:
        return int(self.ids.shape[0])


def state_comps(ns: int, nf: int) -> dict:
    return dict(rhoIndx=0, momIndx=1, specIndx=4, engIndx=4 + ns, utempIndx=5 + ns,
                rhoSrcIndx=0, momSrcIndx=1, specSrcIndx=4, engSrcIndx=4 + nf)


def _periodic_fill(a: np.ndarray, g: int) -> np.ndarray:
    """Pad the last three axes of a by g cells, periodically."""
    return np.pad(a, [(0, 0)] * (a.ndim - 3) + [(g, g)] * 3, mode="wrap")


def gas_velocity(x, y, z, L, U0):
    tw = 2.0 * math.pi
    return (U0 + 0.5 * np.sin(tw * y / L[1]), 0.5 * np.sin(tw * z / L[2]), 0.5 * np.sin(tw * x / L[0]))


def build_problem(name: str, *, ncell=(32, 32, 32), dx=1.0e-4, p=PATM, T0=1000.0, U0=5.0,
                  mech=C.MECH_DODECANE_AIR, fuels=("NC12H26",), n_drops=100000, poisson=False,
                  d_median=20e-6, sigma_ln=0.4, Y_range=None, flow_dt=1e-6, nsteps=100, seed=1234,
                  state_ghosts=2, source_ghosts=1, sdata_kw=None) -> SprayProblem:
    info = mech_info(mech)
    species = info["species"]
    W = np.array(info["W"])
    ns = len(species)
    nf = len(fuels)
    rng = np.random.default_rng(seed)
    nx, ny, nz = ncell
    L = (nx * dx, ny * dx, nz * dx)
    plo = (0.0, 0.0, 0.0)
    # interior cell centres
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    tw = 2.0 * math.pi
    T = T0 * (1.0 + 0.05 * np.sin(tw * X / L[0]) * np.sin(tw * Y / L[1]) * np.sin(tw * Z / L[2]))
    u, v, w = gas_velocity(X, Y, Z, L, U0)
    Ygas = np.zeros(ns)
    Ygas[species.index("O2")] = 0.233
    Ygas[species.index("N2")] = 0.767
    wbar = 1.0 / np.sum(Ygas / W)
    rho = p * wbar / (RU * T)
    comps = state_comps(ns, nf)
    ncomp = 6 + ns
    st = np.zeros((ncomp, nz, ny, nx))
    st[comps["rhoIndx"]] = rho
    st[comps["momIndx"]] = u
    st[comps["momIndx"] + 1] = v
    st[comps["momIndx"] + 2] = w
    for k in range(ns):
        st[comps["specIndx"] + k] = rho * Ygas[k]
    st[comps["utempIndx"]] = T
    g = state_ghosts
    state = np.ascontiguousarray(_periodic_fill(st, g))
    state_lo = (-g, -g, -g)
    state_hi = (nx - 1 + g, ny - 1 + g, nz - 1 + g)
    sg = source_ghosts
    src_shape = (5 + nf, nz + 2 * sg, ny + 2 * sg, nx + 2 * sg)
    src_lo = (-sg, -sg, -sg)
    src_hi = (nx - 1 + sg, ny - 1 + sg, nz - 1 + sg)
    # droplets
    if poisson:
        lam = n_drops / (nx * ny * nz)
        counts = rng.poisson(lam, size=nx * ny * nz)
        n = int(counts.sum())
        cell = np.repeat(np.arange(nx * ny * nz), counts)
        ci = cell % nx
        cj = (cell // nx) % ny
        ck = cell // (nx * ny)
        frac = rng.random((n, 3))
        pos = np.stack([(ci + frac[:, 0]) * dx, (cj + frac[:, 1]) * dx, (ck + frac[:, 2]) * dx], axis=1)
    else:
        n = int(n_drops)
        pos = rng.random((n, 3)) * np.array(L)
    pos = np.minimum(pos, np.nextafter(np.array(L), 0.0))
    dia = d_median * np.exp(sigma_ln * rng.standard_normal(n))
    ug = np.stack(gas_velocity(pos[:, 0], pos[:, 1], pos[:, 2], L, U0), axis=1)
    vel = ug + rng.standard_normal((n, 3))
    nsr = 10 + nf
    rdata = np.zeros((n, nsr))
    rdata[:, 0:3] = vel
    rdata[:, 3] = 300.0
    rdata[:, 4] = dia
    if nf == 1:
        rdata[:, 5] = 1.0
    else:
        lo, hi = Y_range if Y_range is not None else (0.3, 0.7)
        y1 = rng.uniform(lo, hi, n)
        rdata[:, 5] = y1
        rdata[:, 6] = 1.0 - y1
        if nf > 2 and n > 0:
            raise ValueError("nf > 2: liquid compositions come from config5's Dirichlet draw")
    rdata[:, 5 + nf] = 1.0  # numDens
    rdata[:, 6 + nf] = 1.0  # N0
    ids = np.arange(1, n + 1, dtype=np.int64)
    geom = dict(plo=list(plo), phi=list(L), dx=[dx, dx, dx], domain_lo=[0, 0, 0],
                domain_hi=[nx - 1, ny - 1, nz - 1], bndry_lo=[0, 0, 0], bndry_hi=[0, 0, 0])
    kw = dict(sdata_kw or {})
    sdata = spray_data(list(fuels), species, **kw)
    return SprayProblem(name=name, mech=mech, species=species, W=W, nf=nf, geom=geom, comps=comps,
                        sdata=sdata, state=state, state_lo=state_lo, state_hi=state_hi, src_shape=src_shape,
                        src_lo=src_lo, src_hi=src_hi, pos=np.ascontiguousarray(pos),
                        rdata=np.ascontiguousarray(rdata), ids=ids, flow_dt=flow_dt, nsteps=nsteps,
                        state_ghosts=state_ghosts, source_ghosts=source_ghosts,
                        meta=dict(p=p, T0=T0, U0=U0, seed=seed))


def config3(ncell=(128, 128, 256), drops_per_cell: float = 3.0e7 / (128 * 128 * 256), z_frac: float = 1.0,
            T0: float = 600.0, p: float = 5.0 * PATM, wall_T: float = 450.0, theta_c_deg: float = 30.0,
            num_dens: float = 50.0, du_max: float = 150.0, seed: int = 777, flow_dt: float = 5e-7,
            nsteps: int = 100) -> SprayProblem:
    """Config 3: slab with a reflective wall at z = 0 (isothermal, wall_T =
    450 K), x/y periodic, outflow at z_hi; gas at temperature T0 (600 K) and
    pressure p (5 atm; both unspecified by the configuration and recorded
    here) moving at -20 m/s toward the wall; n-dodecane droplets with
    lognormal diameters (median 40 um, sigma_ln 0.6) whose relative velocity
    to the gas gives Weber numbers log-uniform in [1, 300], capped at du_max
    (150 m/s, so a droplet crosses less than one cell per step: the wall
    check of the method assumes it, WallFunctions.H:110-120), directed toward
    the wall (uniform on the lower hemisphere).  TAB breakup,
    splash/deposition and wall films on; contact angle theta_c_deg (30 deg,
    recorded); parcels of num_dens droplets.  `z_frac` < 1 places the
    droplets in the lowest fraction of the slab."""
    mech = C.MECH_DODECANE_AIR
    info = mech_info(mech)
    species = info["species"]
    W = np.array(info["W"])
    ns, nf = len(species), 1
    rng = np.random.default_rng(seed)
    nx, ny, nz = ncell
    dx = 1.0e-4
    L = (nx * dx, ny * dx, nz * dx)
    Ygas = np.zeros(ns)
    Ygas[species.index("O2")] = 0.233
    Ygas[species.index("N2")] = 0.767
    wbar = 1.0 / np.sum(Ygas / W)
    rho = p * wbar / (RU * T0)
    ug = np.array([0.0, 0.0, -20.0])
    comps = state_comps(ns, nf)
    g = 2
    shape = (6 + ns, nz + 2 * g, ny + 2 * g, nx + 2 * g)
    state = np.zeros(shape)
    state[comps["rhoIndx"]] = rho
    for d in range(3):
        state[comps["momIndx"] + d] = ug[d]
    for k in range(ns):
        state[comps["specIndx"] + k] = rho * Ygas[k]
    state[comps["utempIndx"]] = T0
    sg = 1
    n = int(round(drops_per_cell * nx * ny * nz * z_frac))
    pos = rng.random((n, 3)) * np.array([L[0], L[1], L[2] * z_frac])
    pos = np.minimum(pos, np.nextafter(np.array(L), 0.0))
    dia = 40e-6 * np.exp(0.6 * rng.standard_normal(n))
    sdata = spray_data(["NC12H26"], species, do_breakup=1, do_splash=True, wall_T=wall_T, theta_c_deg=theta_c_deg)
    We = np.exp(rng.uniform(0.0, math.log(300.0), n))
    du = np.minimum(np.sqrt(We * sdata["sigma"] / (rho * dia)), du_max)
    # direction uniform on the hemisphere pointing to the wall
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    v[:, 2] = -np.abs(v[:, 2])
    vel = ug[None, :] + du[:, None] * v
    rdata = np.zeros((n, 10 + nf))
    rdata[:, 0:3] = vel
    rdata[:, 3] = 300.0
    rdata[:, 4] = dia
    rdata[:, 5] = 1.0
    rdata[:, 5 + nf] = num_dens
    rdata[:, 6 + nf] = num_dens
    geom = dict(plo=[0.0, 0.0, 0.0], phi=list(L), dx=[dx, dx, dx], domain_lo=[0, 0, 0],
                domain_hi=[nx - 1, ny - 1, nz - 1], bndry_lo=[0, 0, 1], bndry_hi=[0, 0, -1])
    return SprayProblem(name="config3", mech=mech, species=species, W=W, nf=nf, geom=geom, comps=comps,
                        sdata=sdata, state=np.ascontiguousarray(state), state_lo=(-g, -g, -g),
                        state_hi=(nx - 1 + g, ny - 1 + g, nz - 1 + g), src_shape=(5 + nf, nz + 2 * sg, ny + 2 * sg,
                                                                                   nx + 2 * sg),
                        src_lo=(-sg, -sg, -sg), src_hi=(nx - 1 + sg, ny - 1 + sg, nz - 1 + sg),
                        pos=np.ascontiguousarray(pos), rdata=np.ascontiguousarray(rdata),
                        ids=np.arange(1, n + 1, dtype=np.int64), flow_dt=flow_dt, nsteps=nsteps, state_ghosts=g,
                        source_ghosts=sg, meta=dict(p=p, T0=T0, wall_T=wall_T, theta_c_deg=theta_c_deg,
                                                    num_dens=num_dens, seed=seed))


def hollow_cone_droplets(n: int, rng, L, nozzle, *, half_angle_deg: float = 10.0, jitter_deg: float = 3.0,
                         speed: float = 200.0, speed_jitter: float = 0.1, s_range=(0.01, 0.86), s_exp: float = 1.0,
                         d_median: float = 10e-6, sigma_ln: float = 0.5, nf: int = 6, alpha: float = 2.0,
                         T_d: float = 300.0):
    """Droplets of a hollow-cone jet along +z from `nozzle`: direction at
    polar angle half_angle_deg + U(-jitter, jitter) (degrees) and a uniform
    azimuth (the hollow-cone draw of SprayJet::get_new_particle,
    SprayJet.cpp:148-184, with the cone's rim angle), speed speed x (1 +
    U(-speed_jitter, speed_jitter)) along that direction, placed at a travelled
    distance s with density ~ s^s_exp on [s_range] x the box height (s_exp =
    1: droplets per cell fall off as 1/s along the cone, so the cells near
    the nozzle hold thousands), lognormal diameters (median, sigma_ln),
    liquid mass fractions from a Dirichlet(alpha) draw over the nf species,
    temperature T_d, one droplet per parcel.  Returns (pos, rdata)."""
    theta = np.radians(half_angle_deg + rng.uniform(-jitter_deg, jitter_deg, n))
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=1)
    a = s_exp + 1.0
    s0, s1 = s_range[0] * L[2], s_range[1] * L[2]
    sdist = (s0 ** a + rng.random(n) * (s1 ** a - s0 ** a)) ** (1.0 / a)
    pos = np.asarray(nozzle)[None, :] + sdist[:, None] * d
    pos = np.mod(pos, np.asarray(L))
    pos = np.minimum(pos, np.nextafter(np.asarray(L), 0.0))
    umag = speed * (1.0 + rng.uniform(-speed_jitter, speed_jitter, n))
    rdata = np.zeros((n, 10 + nf))
    rdata[:, 0:3] = umag[:, None] * d
    rdata[:, 3] = T_d
    rdata[:, 4] = d_median * np.exp(sigma_ln * rng.standard_normal(n))
    rdata[:, 5:5 + nf] = rng.dirichlet(np.full(nf, alpha), n) if nf > 1 else 1.0
    rdata[:, 5 + nf] = 1.0
    rdata[:, 6 + nf] = 1.0
    return np.ascontiguousarray(pos), np.ascontiguousarray(rdata)


def config5(ncell: int = 256, n_drops: float = 2.0e8, seed: int = 5555, droplets: bool = True,
            **jet_kw) -> SprayProblem:
    """Config 5 (dense uneven jet): ncell^3 periodic box, dx = 1e-4 m, gas at
    60 atm / 800 K with the sinusoidal recipe (no mean flow; the state is to
    be handed to the update in FP32), n_drops droplets of a hollow-cone jet
    (half-angle 10 deg +- 3 deg, 200 m/s +- 10 %, nozzle on the axis at 4 %
    of the box height, see hollow_cone_droplets) -- 0 droplets in most cells
    and hundreds to thousands in the cells of the cone -- of a six-component
    C8-C16 n-alkane liquid (Dirichlet(2) mass fractions), median diameter
    10 um (sigma_ln 0.5, recorded here), 20 steps of 2e-7 s.  With
    droplets=False only the gas and geometry are built (jet_droplets then
    draws the droplets)."""
    from .fuels import ALKANE6

    prob = build_problem("config5", ncell=(ncell, ncell, ncell), p=60.0 * PATM, T0=800.0, U0=0.0,
                         mech=C.MECH_ALKANE6_AIR, fuels=ALKANE6, n_drops=0, flow_dt=2e-7, nsteps=20, seed=seed)
    L = prob.geom["phi"]
    nozzle = np.array([0.5 * L[0], 0.5 * L[1], 0.04 * L[2]])
    prob.meta.update(nozzle=nozzle.tolist(), n_target=int(n_drops), jet=dict(jet_kw))
    if droplets and n_drops > 0:
        prob.pos, prob.rdata, prob.ids = jet_droplets(prob, int(n_drops))
    return prob


def jet_droplets(prob: SprayProblem, n_drops: int, threads: int | None = None):
    """The n_drops droplets of a config5 problem as (pos, rdata, ids), ids
    1..n_drops.  Blocks of 2^20 droplets have their own seeded generators
    and are drawn by a thread pool (numpy's generators release the GIL), so
    the set does not depend on the thread count."""
    from concurrent.futures import ThreadPoolExecutor

    L = prob.geom["phi"]
    nozzle = np.array(prob.meta["nozzle"])
    block = 1 << 20
    nblk = (n_drops + block - 1) // block
    pos = np.empty((n_drops, 3))
    rdata = np.empty((n_drops, 10 + prob.nf))

    def fill(b):
        lo, hi = b * block, min(n_drops, (b + 1) * block)
        rng = np.random.default_rng([prob.meta["seed"], b])
        pos[lo:hi], rdata[lo:hi] = hollow_cone_droplets(hi - lo, rng, L, nozzle, nf=prob.nf, **prob.meta["jet"])

    nt = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=nt) as ex:
        list(ex.map(fill, range(nblk)))
    return pos, rdata, np.arange(1, n_drops + 1, dtype=np.int64)


def config1(scale: float = 1.0, seed: int = 1234, **kw) -> SprayProblem:
    """Config 1: 32^3 periodic box, dx = 1e-4 m, air at 1 atm, T0 = 1000 K,
    1e5 n-dodecane droplets (median 20 um, sigma_ln 0.4), 100 steps of 1e-6 s.
    `scale` < 1 shrinks the droplet count (box unchanged)."""
    return build_problem("config1", ncell=(32, 32, 32), p=PATM, T0=1000.0, mech=C.MECH_DODECANE_AIR,
                         fuels=("NC12H26",), n_drops=int(round(1e5 * scale)), d_median=20e-6, sigma_ln=0.4,
                         flow_dt=1e-6, nsteps=100, seed=seed, **kw)


def decane_case(ncell: int = 32, drops_per_cell: float = 3.0, p: float = 10.0 * PATM, T0: float = 1500.0,
                seed: int = 777, **kw) -> SprayProblem:
    """An n-decane spray in the gas mechanism of the reference's decane decks
    (decane_3sp: O2, N2, NC10H22; Exec/SprayTests/PeleLMeX/abramzon_test):
    a periodic ncell^3 box, dx = 1e-4 m, air at the Abramzon case's 10 bar
    and 1500 K with the config-1 sinusoidal recipe, Poisson(drops_per_cell)
    droplets per cell (median 20 um, sigma_ln 0.4, T_d = 300 K), 1e-6 s
    steps."""
    n = drops_per_cell * ncell ** 3
    return build_problem("decane", ncell=(ncell, ncell, ncell), p=p, T0=T0, mech=C.MECH_DECANE_AIR,
                         fuels=("NC10H22",), n_drops=n, poisson=True, d_median=20e-6, sigma_ln=0.4, flow_dt=1e-6,
                         nsteps=100, seed=seed, **kw)


def config4(ncell: int = 128, drops_per_cell: float = 1.6e7 / 128 ** 3, seed: int = 2468, **kw) -> SprayProblem:
    """Config 4: the learned surrogate on the config-1 physics scaled to
    1.6e7 droplets: a 128^3 periodic box, dx = 1e-4 m, air at 1 atm and
    T0 = 1000 K (the config-1 gas recipe), Poisson(~7.63) n-dodecane droplets
    per cell (median 20 um, sigma_ln 0.4, T_d = 300 K), 1e-6 s steps.
    `ncell` < 128 gives the same recipe on a smaller box (for tests)."""
    n = drops_per_cell * ncell ** 3
    return build_problem("config4", ncell=(ncell, ncell, ncell), p=PATM, T0=1000.0, mech=C.MECH_DODECANE_AIR,
                         fuels=("NC12H26",), n_drops=n, poisson=True, d_median=20e-6, sigma_ln=0.4, flow_dt=1e-6,
                         nsteps=100, seed=seed, **kw)


def config2(ncell: int = 128, drops_per_cell: float = 1.6e7 / 128 ** 3, seed: int = 4321, **kw) -> SprayProblem:
    """Config 2: 128^3 periodic box at 30 atm, T0 = 900 K, Poisson(~7.63)
    droplets per cell of an n-decane/n-dodecane liquid (Y_decane ~ U[0.3,0.7]),
    median 15 um, sigma_ln 0.5, 50 steps of 5e-7 s.  `ncell` < 128 gives the
    same recipe on a smaller box (for tests)."""
    n = drops_per_cell * ncell ** 3
    return build_problem("config2", ncell=(ncell, ncell, ncell), p=30.0 * PATM, T0=900.0,
                         mech=C.MECH_DECANE_DODECANE_AIR, fuels=("NC10H22", "NC12H26"), n_drops=n, poisson=True,
                         d_median=15e-6, sigma_ln=0.5, Y_range=(0.3, 0.7), flow_dt=5e-7, nsteps=50, seed=seed, **kw)


def to_pelec(prob: SprayProblem, T_guess_factor: float = 1.0) -> SprayProblem:
    """The same problem in the PeleC flavour's CGS form: SprayData in CGS
    (the deck's reference latent heat, unshifted), geometry and droplets in
    cm, the state as rho, rho*u, rho*E, rho*Y and T (g/cm^3, g/(cm^2 s),
    erg/cm^3; E = e + |u|^2/2 with e the mixture's specific internal energy
    from the NASA-7 enthalpies), sources in CGS.  T_guess_factor scales the
    stored temperature, which the PeleC gather only uses as the EY2T guess."""
    import copy

    p = copy.deepcopy(prob)
    L, RHO, E = 1.0e-2, 1.0e3, 1.0e-4
    sd = dict(prob.sdata)
    sd["sigma"] = sd["sigma"] / 1.0e-3
    sd["body_force"] = [v / L for v in sd["body_force"]]
    for k, f in (("cp", E), ("ref_latent", E), ("latent", E)):
        sd[k] = [v / f for v in sd[k]]
    psat = list(sd["psat_coef"])
    for i in range(len(psat) // 4):
        psat[4 * i + 3] = psat[4 * i + 3] / 1.0e-1
    sd["psat_coef"] = psat
    sd["rho_coef"] = [v / RHO for v in sd["rho_coef"]]
    sd["lambda_coef"] = [v / 1.0e-5 for v in sd["lambda_coef"]]
    sd["mu_coef"] = [v / 1.0e-1 for v in sd["mu_coef"]]
    p.sdata = sd
    g = dict(prob.geom)
    for k in ("plo", "phi", "dx"):
        g[k] = [v / L for v in g[k]]
    p.geom = g
    p.pos = prob.pos / L
    rd = prob.rdata.copy()
    rd[:, 0:3] /= L
    rd[:, 4] /= L
    rd[:, 9 + prob.nf] /= L
    if sd["do_breakup"] == 2:
        rd[:, 7 + prob.nf] /= 1.0e-6
    p.rdata = rd
    cm = prob.comps
    st = prob.state
    ns = len(prob.species)
    rho = st[cm["rhoIndx"]]
    T = st[cm["utempIndx"]]
    Y = np.stack([st[cm["specIndx"] + k] / rho for k in range(ns)])
    # e(T) = sum_k Y_k (h_k - R T / W_k), SI, NASA-7 h_k/(R T / W_k)
    info = mech_info(prob.mech)
    e = np.zeros_like(T)
    for k in range(ns):
        lo, hi = np.array(info["low"][k]), np.array(info["high"][k])
        a = np.where((T < info["Tmid"][k])[None], lo[:, None, None, None], hi[:, None, None, None])
        hRT = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T
        e += Y[k] * (RU * T / prob.W[k]) * (hRT - 1.0)
    rho_c = rho / RHO
    u_c = np.stack([st[cm["momIndx"] + d] / L for d in range(3)])
    new = np.zeros_like(st)
    new[cm["rhoIndx"]] = rho_c
    for d in range(3):
        new[cm["momIndx"] + d] = rho_c * u_c[d]
    new[cm["engIndx"]] = rho_c * (e / E + 0.5 * np.sum(u_c * u_c, axis=0))
    for k in range(ns):
        new[cm["specIndx"] + k] = rho_c * Y[k]
    new[cm["utempIndx"]] = T * T_guess_factor
    p.state = new
    p.name = prob.name + "_pelec"
    p.meta = dict(prob.meta, flavour="pelec")
    return p


def eb_plane(prob: SprayProblem, point, normal, nsub: int = 6) -> dict:
    """A planar embedded boundary over the state box of `prob`, in the layout
    of AMReX's EB data (EBCellFlag words, volume fraction, cell centroid,
    boundary centroid and boundary normal, each relative to the cell centre
    in cell units).  The solid is the side {x : (x - point) . normal > 0};
    `normal` (made unit) is the boundary normal, pointing out of the fluid.
    Volume fractions and centroids come from nsub^3 sample points per cell;
    the boundary centroid is the point of the plane nearest the cell centre;
    a cell is connected to each neighbour that is not covered."""
    g = prob.geom
    lo, hi = np.array(prob.state_lo), np.array(prob.state_hi)
    n = hi - lo + 1
    dx = np.array(g["dx"])
    plo = np.array(g["plo"])
    dl = np.array(g["domain_lo"])
    nrm = np.asarray(normal, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    x0 = np.asarray(point, dtype=np.float64)
    # cell centres (physical), arrays indexed [k, j, i]
    ii = [np.arange(n[d]) + lo[d] for d in range(3)]
    cz, cy, cx = np.meshgrid(ii[2], ii[1], ii[0], indexing="ij")
    cent = np.stack([(cx - dl[0] + 0.5) * dx[0] + plo[0], (cy - dl[1] + 0.5) * dx[1] + plo[1],
                     (cz - dl[2] + 0.5) * dx[2] + plo[2]])
    sub = (np.arange(nsub) + 0.5) / nsub - 0.5
    vol = np.zeros(n[::-1])
    csum = np.zeros((3,) + tuple(n[::-1]))
    for a in sub:
        for b in sub:
            for c in sub:
                off = np.array([a, b, c])
                pt = cent + (off * dx)[:, None, None, None]
                fluid = np.tensordot(nrm, pt - x0[:, None, None, None], axes=1) <= 0.0
                vol += fluid
                csum += fluid[None] * off[:, None, None, None]
    vfrac = vol / nsub ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        ccent = np.where(vol[None] > 0, csum / np.maximum(vol, 1)[None], 0.0)
    covered = vol == 0
    cut = (vol > 0) & (vol < nsub ** 3)
    # boundary centroid: nearest point of the plane, in cell units
    dist = np.tensordot(nrm, cent - x0[:, None, None, None], axes=1)
    bc = (-dist[None] * nrm[:, None, None, None]) / dx[:, None, None, None]
    bcent = np.where(cut[None], bc, 0.0)
    bnorm = np.where(cut[None], np.broadcast_to(nrm[:, None, None, None], bc.shape), 0.0)
    ccent = np.where(cut[None], ccent, 0.0)
    vfrac = np.where(cut, vfrac, np.where(covered, 0.0, 1.0))
    flags = np.zeros(n[::-1], dtype=np.uint32)
    flags[cut] = 1
    flags[covered] = 3
    # connectivity to the 26 neighbours and self for cells that are not covered
    padc = np.pad(covered, 1, mode="edge")
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                nb = padc[1 + dk:1 + dk + n[2], 1 + dj:1 + dj + n[1], 1 + di:1 + di + n[0]]
                bit = np.uint32(1) << np.uint32(5 + (di + 1) + 3 * (dj + 1) + 9 * (dk + 1))
                flags |= np.where(~covered & ~nb, bit, np.uint32(0)).astype(np.uint32)
    return dict(flags=flags, vfrac=vfrac, ccent=ccent, bcent=bcent, bnorm=bnorm, lo=tuple(int(v) for v in lo),
                hi=tuple(int(v) for v in hi), point=x0, normal=nrm)
