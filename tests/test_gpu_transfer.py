"""GPU transferSource / addSpraySrc, estTimestep and computeDerivedVars
against the oracle (SprayParticles.H:250-319, SprayParticles.cpp:89-140,
SprayDerive.cpp:9-155): bitwise equal in the FP64 form, for both component
maps, with a wall and with wall films."""
import numpy as np
import pytest

from oracle import oracle_py as O
from pelemp_amd import _capi as C
from pelemp_amd import problem as P
from tests.helpers import run_gpu

pytestmark = pytest.mark.gpu


def _container(prob, form=0):
    from pelemp_amd.spray import SprayParticleContainer

    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, max(prob.n, 1), form=form)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    return pc


@pytest.mark.parametrize("walls", [False, True])
@pytest.mark.parametrize("same_ncomp", [False, True])
def test_transfer_source_matches_oracle(walls, same_ncomp):
    from pelemp_amd.spray import DeviceFab

    prob = P.config3(ncell=(16, 16, 32), drops_per_cell=1.0) if walls else P.config2(ncell=16)
    pc = _container(prob)
    g = prob.geom
    n = [g["domain_hi"][d] - g["domain_lo"][d] + 1 for d in range(3)]
    rng = np.random.default_rng(11)
    nsrc = prob.src_shape[0]
    ng_t, ng_a = 2, 1
    t_lo, t_hi = (-ng_t,) * 3, tuple(n[d] - 1 + ng_t for d in range(3))
    a_lo, a_hi = (-ng_a,) * 3, tuple(n[d] - 1 + ng_a for d in range(3))
    nact = nsrc if same_ncomp else 5 + len(prob.species) + 1
    tmp = rng.normal(size=(nsrc, n[2] + 2 * ng_t, n[1] + 2 * ng_t, n[0] + 2 * ng_t))
    act = rng.normal(size=(nact, n[2] + 2 * ng_a, n[1] + 2 * ng_a, n[0] + 2 * ng_a))
    tmp_d = DeviceFab(t_lo, t_hi, nsrc, tmp)
    act_d = DeviceFab(a_lo, a_hi, nact, act)
    pc.transferSource(2, 0, tmp_d, act_d)
    dep = list(prob.sdata["dep_indx"][:prob.nf])
    assert O.transfer_source(g, prob.comps, dep, 2, tmp, t_lo, t_hi, act, a_lo, a_hi) == 0
    assert np.array_equal(act_d.numpy(), act)
    assert not tmp_d.numpy().any()
    pc.close()


def test_update_then_transfer_moves_the_source():
    # one FP64 step of config 2 (small), then transferSource into the state
    # layout: density, momentum, energy and fuel species receive the spray
    # sources; the temporary source is cleared
    from pelemp_amd.spray import DeviceFab

    prob = P.config2(ncell=16)
    pc, _, _, _, srcs = run_gpu(prob, 1)
    tmp_d = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0], srcs[0])
    nact = 5 + len(prob.species) + 1
    act_d = DeviceFab(prob.src_lo, prob.src_hi, nact)
    pc.transferSource(prob.source_ghosts, 0, tmp_d, act_d)
    a = act_d.numpy()
    s = srcs[0]
    cm = prob.comps
    # the update deposits into valid cells only, so the ghost images add zeros
    assert s[cm["rhoSrcIndx"]].any()
    assert np.array_equal(a[cm["rhoIndx"]], s[cm["rhoSrcIndx"]])
    assert np.array_equal(a[cm["engIndx"]], s[cm["engSrcIndx"]])
    for d in range(3):
        assert np.array_equal(a[cm["momIndx"] + d], s[cm["momSrcIndx"] + d])
    for sp in range(prob.nf):
        assert np.array_equal(a[cm["specIndx"] + prob.sdata["dep_indx"][sp]], s[cm["specSrcIndx"] + sp])
    assert not tmp_d.numpy().any()
    pc.close()


@pytest.mark.parametrize("form", [C.FORM_FP64, C.FORM_FP32])
def test_est_timestep_matches_oracle(form):
    prob = P.config2(ncell=16)
    pc = _container(prob, form)
    _, rd, ids = pc.get_particles()
    vel = rd[:, 0:3]
    for jet in (0.0, 1e4):
        dt = pc.estTimestep(0, 0.5, jet)
        assert dt == O.est_timestep(prob.geom, vel, 0.5, jet)
    pc.close()


@pytest.mark.parametrize("case", ["config2", "config3_films"])
def test_derived_vars_match_oracle(case):
    from pelemp_amd.spray import DeviceFab

    if case == "config2":
        prob = P.config2(ncell=16)
        nsteps = 2
    else:
        prob = P.config3(ncell=(16, 16, 32), drops_per_cell=3.0, z_frac=0.25)
        prob.sdata = dict(prob.sdata, do_breakup=0)
        nsteps = 4
    pc, pos, rd, ids, _ = run_gpu(prob, nsteps, cap=4 * prob.n)
    if case == "config3_films":
        assert (rd[:, 9 + prob.nf] > 0).any(), "no wall film formed"
    g = prob.geom
    n = [g["domain_hi"][d] - g["domain_lo"][d] + 1 for d in range(3)]
    nvar = pc.NumDeriveVars(1)
    assert nvar == len(pc.DeriveVarNames(1))
    start = 2
    lo, hi = (-1,) * 3, tuple(n[d] for d in range(3))
    rng = np.random.default_rng(4)
    init = rng.normal(size=(start + nvar, n[2] + 2, n[1] + 2, n[0] + 2))
    mf_d = DeviceFab(lo, hi, start + nvar, init)
    pc.computeDerivedVars(mf_d, 0, start, 1)
    ref = init.copy()
    assert O.compute_derived(prob.sdata, g, pos, rd, ref, lo, hi, start, True) == 0
    got = mf_d.numpy()
    assert np.array_equal(got, ref)
    pc.close()


def test_derived_vars_fp32_form_totals():
    from pelemp_amd.spray import DeviceFab

    prob = P.config2(ncell=16)
    pc, pos, rd, ids, _ = run_gpu(prob, 2, form=C.FORM_FP32)
    nvar = pc.NumDeriveVars(1)
    n = 16
    mf_d = DeviceFab((0, 0, 0), (n - 1,) * 3, nvar)
    pc.computeDerivedVars(mf_d, 0, 0, 1)
    ref = np.zeros((nvar, n, n, n))
    assert O.compute_derived(prob.sdata, prob.geom, pos, rd, ref, (0, 0, 0), (n - 1,) * 3, 0, True) == 0
    got = mf_d.numpy()
    # additive totals (mass, number, volume, surface, parcels, fuel masses)
    for q in (0, 2, 3, 4, 11, 15, 16):
        assert got[q].sum() == pytest.approx(ref[q].sum(), rel=1e-12)
    pc.close()


def test_fab_set_zero_in_stream_order():
    # the source reset between steps: a kernel on the caller's stream that
    # clears exactly the words asked for
    from pelemp_amd import _capi as C
    from pelemp_amd.spray import DeviceFab

    rng = np.random.default_rng(5)
    data = rng.normal(size=(3, 6, 5, 4)) + 10.0
    fab = DeviceFab((0, 0, 0), (3, 4, 5), 3, data)
    s = C.Stream()
    n = 2 * 6 * 5 * 4 + 7  # two components and a few words of the third
    C.check(C.lib().pelemp_fab_set_zero(fab.buf.ptr, n, s.handle), "zero")
    s.synchronize()
    got = fab.numpy().reshape(-1)
    assert not got[:n].any()
    assert np.array_equal(got[n:], data.reshape(-1)[n:])
