"""GPU injection, lattice initialisation, appending and the particle files
against the oracle (SprayInjection.H:10-348, SprayJet.{H,cpp},
Distributions.cpp, SprayIO.cpp): injected droplets and the jet state bitwise
equal in the FP64 form (every distribution, swirl, hollow cone, a rate
profile, KH-RT's initial state, the PeleC flavour), then the update over the
grown droplet set bitwise equal (the new droplets follow the old ones in
each cell); storage and slot compaction when droplets have left; checkpoint
/ restart bit for bit."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from pelemp_amd import _capi as C
from pelemp_amd import problem as P
from tests.helpers import by_id

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _state(prob):
    from pelemp_amd.spray import DeviceFab

    return DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)


def _copy(j):
    return C.SprayJetC.from_buffer_copy(j)


def _jet_state(j):
    return (j.num_ppp, j.sum_inj_mass, j.sum_inj_time, j.min_parcel, j.total_inj_mass, j.total_inj_time,
            j.n_injections)


JETS = {
    "normal_swirl": dict(dist_type="Normal", dist_params=(20e-6, 6e-6), phi_swirl=15.0, inject_ppp=5.0),
    "lognormal_hollow": dict(dist_type="LogNormal", dist_params=(25e-6, 8e-6), hollow_spray=True,
                             hollow_spread=8.0, inject_ppp=3.0),
    "weibull": dict(dist_type="Weibull", dist_params=(22e-6, 3.0), inject_ppp=4.0),
    "chisquared": dict(dist_type="ChiSquared", dist_params=(30e-6,), inject_ppp=2.0),
    "uniform_first_ppp": dict(dist_type="Uniform", dist_params=(18e-6,), inject_ppp=-1.0),
}


def _make_jet(pc, prob, scale=1.0, mass_scale=1.0, **kw):
    from pelemp_amd.spray import SprayJet

    g = prob.geom
    cent = [0.5 * (a + b) for a, b in zip(g["plo"], g["phi"])]
    args = dict(jet_cent=cent, jet_norm=[0.3, -0.2, 1.0], spread_angle=20.0, jet_dia=1e-4 * scale,
                jet_vel=50.0 * scale, mass_flow=2e-3 * mass_scale, jet_temp=300.0,
                jet_Y=[0.4, 0.6][:prob.nf] if prob.nf > 1 else None, seed=3)
    args.update(kw)
    if "dist_params" in args and scale != 1.0:
        dp = list(args["dist_params"])
        if args.get("dist_type") != "Weibull":
            dp = [v * scale for v in dp]
        else:
            dp[0] *= scale
        args["dist_params"] = tuple(dp)
    return SprayJet("jet1", pc, **args)


def _inject_steps(pc, prob, jet, nsteps, dt, pelec=False, first_id=None):
    """GPU and oracle injections step by step; returns the oracle's new droplets."""
    oj = _copy(jet._j)
    next_id = first_id if first_id is not None else int(prob.ids.max()) + 1
    allp, allr, alli = [], [], []
    for step in range(nsteps):
        n_g = pc.sprayInjection(step * dt, jet, dt)
        st, p, r, i = O.inject(prob.sdata, prob.geom, oj, step * dt, dt, next_id, pelec=pelec)
        assert st == 0
        assert n_g == len(i), f"step {step}: {n_g} vs {len(i)} droplets"
        assert _jet_state(jet._j) == _jet_state(oj), f"jet state differs at step {step}"
        next_id += len(i)
        allp.append(p)
        allr.append(r)
        alli.append(i)
    return np.concatenate(allp), np.concatenate(allr), np.concatenate(alli)


@pytest.mark.parametrize("case", list(JETS))
def test_injection_bitwise(case):
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=16)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, **JETS[case])
    p, r, i = _inject_steps(pc, prob, jet, 4, 1e-6)
    # inject_ppp < 0 sizes the parcel to one step's mass: about one per step
    assert len(i) >= (4 if case == "uniform_first_ppp" else 10)
    gp, gr, gi = by_id(*pc.get_particles())
    new = gi >= i.min()
    assert np.array_equal(gi[new], i)
    assert np.array_equal(gp[new], p)
    assert np.array_equal(gr[new], r)
    # the droplets present before are untouched
    op, orr, oi = by_id(prob.pos, prob.rdata, prob.ids)
    assert np.array_equal(gi[~new], oi) and np.array_equal(gp[~new], op) and np.array_equal(gr[~new], orr)
    pc.close()


def test_injection_needing_a_second_batch_bitwise():
    # a narrow distribution just above the smallest diameter (0.65 um): about
    # 30% of the candidates are rejected, so one call draws ~86k candidates,
    # more than its first batch (1.05 x the ~69k expected + 4096), and goes on
    # in a second batch with the running mass and output offset carried over
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=16)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 80000)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, mass_scale=4e-3, dist_type="Normal", dist_params=(0.68e-6, 0.06e-6), inject_ppp=1.0)
    p, r, i = _inject_steps(pc, prob, jet, 1, 1e-6)
    assert len(i) > 50000
    gp, gr, gi = by_id(*pc.get_particles())
    new = gi >= i.min()
    assert np.array_equal(gi[new], i)
    assert np.array_equal(gp[new], p)
    assert np.array_equal(gr[new], r)
    pc.close()


def test_injection_then_update_bitwise():
    # inject, update, inject, update: sources and droplets equal the oracle's
    # bit for bit, which needs the new droplets after the old ones in every
    # cell on both sides
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    prob = P.config2(ncell=16)
    sd = O.latent_shifted(prob.sdata, prob.mech, len(prob.species))
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, **JETS["normal_swirl"])
    oj = _copy(jet._j)
    state = _state(prob)
    pos, rd, ids = prob.pos, prob.rdata, prob.ids
    next_id = int(ids.max()) + 1
    dt = prob.flow_dt
    for step in range(3):
        pc.sprayInjection(step * dt, jet, dt)
        st, p, r, i = O.inject(prob.sdata, prob.geom, oj, step * dt, dt, next_id)
        assert st == 0 and len(i) > 0
        next_id += len(i)
        pos, rd, ids = np.concatenate([pos, p]), np.concatenate([rd, r]), np.concatenate([ids, i])
        src_o = prob.new_source()
        st, pos, rd, ids = O.update(prob.mech, sd, prob.comps, prob.geom, pos, rd, ids, prob.state, prob.state_lo,
                                    prob.state_hi, src_o, prob.src_lo, prob.src_hi, dt)
        assert st == 0
        src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        pc.updateParticles(0, state, src_g, dt, step * dt, prob.state_ghosts, prob.source_ghosts, False, False, True)
        assert np.array_equal(src_g.numpy(), src_o), f"sources differ at step {step}"
    g = by_id(*pc.get_particles())
    o = by_id(pos, rd, ids)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    pc.close()


def test_injection_rate_profile_and_khrt_state():
    # a rate-of-injection profile (the SprayA file, time axis stretched to
    # the test's steps) into a KH-RT context: BM2 starts at -1 (the RT clock)
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config3(ncell=(16, 16, 32), drops_per_cell=1.0)
    sdata = dict(prob.sdata, do_breakup=2, do_splash=0)
    pc = SprayParticleContainer(prob.mech, sdata, prob.comps, prob.geom, prob.n + 20000)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, dist_type="Uniform", dist_params=(20e-6,), inject_ppp=2.0, jet_norm=[0, 0, 1])
    jet.readROI(os.path.join(GOLDEN, "roi_spraya_wbreakup.dat"), 700.0, 0.89)
    assert jet._j.n_roi > 10 and jet._j.start_time > 0
    prob.sdata = sdata
    oj = _copy(jet._j)
    next_id = int(prob.ids.max()) + 1
    t0 = jet._j.start_time
    for step in range(3):
        t = t0 + step * 3e-7
        n_g = pc.sprayInjection(t, jet, 1e-7)
        st, p, r, i = O.inject(sdata, prob.geom, oj, t, 1e-7, next_id)
        assert st == 0 and n_g == len(i)
        assert _jet_state(jet._j) == _jet_state(oj)
        next_id += len(i)
    gp, gr, gi = by_id(*pc.get_particles())
    assert np.all(gr[gi >= int(prob.ids.max()) + 1, 8 + prob.nf] == -1.0)
    pc.close()


def test_pelec_injection_bitwise():
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.to_pelec(P.config2(ncell=16), T_guess_factor=1.0)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000, flavour="pelec")
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    # CGS: cm, cm/s, g/s
    jet = _make_jet(pc, prob, scale=100.0, mass_scale=1.0e3, **JETS["normal_swirl"])
    p, r, i = _inject_steps(pc, prob, jet, 3, 1e-6, pelec=True)
    gp, gr, gi = by_id(*pc.get_particles())
    new = gi >= i.min()
    assert np.array_equal(gi[new], i) and np.array_equal(gp[new], p) and np.array_equal(gr[new], r)
    pc.close()


def test_uniform_init_then_update_bitwise():
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    prob = P.config2(ncell=16)
    sd = O.latent_shifted(prob.sdata, prob.mech, len(prob.species))
    npart = (20, 24, 28)
    n = int(np.prod(npart))
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, n)
    pc.uniformSprayInit(npart, (1.0, -2.0, 0.5), 20e-6, 310.0, (0.3, 0.7), num_ppp=4.0)
    st, pos, rd, ids = O.uniform_init(prob.sdata, prob.geom, npart, (1.0, -2.0, 0.5), 20e-6, 310.0, (0.3, 0.7),
                                      4.0, 1)
    assert st == 0
    g = by_id(*pc.get_particles())
    o = by_id(pos, rd, ids)
    for a, b in zip(g, o):
        assert np.array_equal(a, b)
    src_o = prob.new_source()
    assert O.update(prob.mech, sd, prob.comps, prob.geom, pos, rd, ids, prob.state, prob.state_lo, prob.state_hi,
                    src_o, prob.src_lo, prob.src_hi, prob.flow_dt)[0] == 0
    src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    pc.updateParticles(0, _state(prob), src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False,
                       False, True)
    assert np.array_equal(src_g.numpy(), src_o)
    pc.close()


def test_add_particles_after_droplets_left():
    # outflow faces: droplets leave during the update; adding then needs the
    # storage of the dead entries and their slots back (capacity 1.5x)
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    prob = P.config2(ncell=16)
    g = dict(prob.geom, bndry_lo=[-1, -1, -1], bndry_hi=[-1, -1, -1])
    prob.geom = g
    n0 = prob.n
    rd = prob.rdata.copy()
    rng = np.random.default_rng(2)
    # fast droplets near the faces leave within one step
    L = g["phi"][0]
    near = (prob.pos[:, 0] > 0.9 * L)
    rd[near, 0] = 0.2 * L / prob.flow_dt
    prob.rdata = rd
    sd = O.latent_shifted(prob.sdata, prob.mech, len(prob.species))
    cap = int(1.5 * n0)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, g, cap)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    src_o = prob.new_source()
    st, pos, rd2, ids = O.update(prob.mech, sd, prob.comps, g, prob.pos, prob.rdata, prob.ids, prob.state,
                                 prob.state_lo, prob.state_hi, src_o, prob.src_lo, prob.src_hi, prob.flow_dt)
    assert st == 0 and len(ids) < n0
    state = _state(prob)
    src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False, True)
    assert np.array_equal(src_g.numpy(), src_o)
    # more than the free storage after the pass, within the live capacity
    m = cap - len(ids)
    assert len(ids) + m <= cap < n0 + m
    new_pos = rng.uniform(0.1, 0.8, size=(m, 3)) * L
    new_rd = np.tile(prob.rdata[0], (m, 1))
    new_rd[:, 0:3] = rng.normal(0.0, 1.0, size=(m, 3))
    new_ids = np.arange(m, dtype=np.int64) + int(prob.ids.max()) + 1
    pc.add_particles(new_pos, new_rd, new_ids)
    pos = np.concatenate([pos, new_pos])
    rd2 = np.concatenate([rd2, new_rd])
    ids = np.concatenate([ids, new_ids])
    gg = by_id(*pc.get_particles())
    oo = by_id(pos, rd2, ids)
    for a, b in zip(gg, oo):
        assert np.array_equal(a, b)
    # and the update over the compacted set
    src_o = prob.new_source()
    assert O.update(prob.mech, sd, prob.comps, g, pos, rd2, ids, prob.state, prob.state_lo, prob.state_hi, src_o,
                    prob.src_lo, prob.src_hi, prob.flow_dt)[0] == 0
    src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False, True)
    assert np.array_equal(src_g.numpy(), src_o)
    pc.close()


def test_checkpoint_restart_and_ascii(tmp_path):
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=16)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, **JETS["normal_swirl"])
    pc.sprayInjection(0.0, jet, 1e-6)
    pc.write_ascii_files = True
    d = str(tmp_path / "chk00042")
    pc.SprayParticleIO(0, True, d)
    before = by_id(*pc.get_particles())
    # AMReX's particle checkpoint layout, read by the independent numpy reader
    from tests import amrex_particles as A

    h = A.read(d + "/particles")
    assert h["version"] == "Version_Two_Dot_One_double" and h["nparticles"] == len(before[2])
    assert h["real_names"][:5] == ["xvel", "yvel", "zvel", "temperature", "diam"]
    assert h["real_names"][-1] == "wall_film_height" and len(h["real_names"]) == before[1].shape[1]
    gid, gcpu, gpos, grd = h["levels"][0][0]
    assert not gcpu.any()
    o = np.argsort(gid)
    assert np.array_equal(gid[o], before[2]) and np.array_equal(gpos[o], before[0])
    assert np.array_equal(grd[o], before[1])
    # the ASCII file next to the checkpoint: count, then x y z id cpu rdata
    txt = open(str(tmp_path / "spray00042.p3d")).read().split("\n")
    assert int(txt[0]) == len(before[2])
    rows = np.array([[float(v) for v in ln.split()] for ln in txt[1:] if ln.strip()])
    o = np.argsort(rows[:, 3])
    np.testing.assert_allclose(rows[o, 0:3], before[0], rtol=1e-14)
    np.testing.assert_array_equal(rows[o, 3].astype(np.int64), before[2])
    np.testing.assert_allclose(rows[o, 5:], before[1], rtol=1e-14, atol=1e-300)
    # the jet log
    state0 = _jet_state(jet._j)
    pc2 = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000)
    jet2 = _make_jet(pc2, prob, **JETS["normal_swirl"])
    pc2.m_sprayJets.append(jet2)
    pc2.Restart(d)
    pc2.PostInitRestart(d)
    after = by_id(*pc2.get_particles())
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    # the log holds 15 significant digits (SprayIO.cpp:80)
    np.testing.assert_allclose(_jet_state(jet2._j)[:6], state0[:6], rtol=1e-14)
    # the next injection continues the id sequence past the restart's ids
    pc2.sprayInjection(1e-6, jet2, 1e-6)
    ids2 = pc2.get_particles()[2]
    assert ids2.max() > before[2].max() and len(np.unique(ids2)) == len(ids2)
    pc.close()
    pc2.close()


def test_fp32_injection_matches_within_fp32():
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=16)
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 20000, form=C.FORM_FP32)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    jet = _make_jet(pc, prob, **JETS["weibull"])
    p, r, i = _inject_steps(pc, prob, jet, 2, 1e-6)
    gp, gr, gi = by_id(*pc.get_particles())
    new = gi >= i.min()
    assert np.array_equal(gi[new], i)
    # positions are kept inside their cell in FP32 (dx = 1e-4 m)
    np.testing.assert_allclose(gp[new], p, rtol=0, atol=2e-11)
    np.testing.assert_allclose(gr[new][:, 0:5], r[:, 0:5], rtol=2e-7)
    pc.close()
