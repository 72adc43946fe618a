"""GPU update vs the CPU oracle for the arguments that bound a call
(SprayParticles.cpp:142-502): ghost droplets whose cell leaves the box grown
by source_ghosts are deleted (:477-479; FP64 plain and branch kernels,
bitwise against the oracle; FP32 against its own active call), and the
device-side checks report "state box too small" (state_ghosts smaller than
the stencil needs, SprayInterpolation.H:121-124), "source box too small"
(:420-425; dropped without error for ghost droplets) and "left the domain"
(:347-351; found by the host when the droplets are handed in) with the
oracle's status.  Every kernel returns normally on these paths; the host
reads the sticky error word."""
import numpy as np
import pytest

from pelemp_amd import problem as P
from tests.helpers import by_id

pytestmark = pytest.mark.gpu


def _lib():
    from pelemp_amd import _capi as C

    if C.device_count() < 1:
        pytest.skip("no HIP device")
    return C


def _oracle(prob, nsteps, dt, branch=None, src_box=None, **kw):
    from oracle import oracle_py as O

    sd = O.latent_shifted(prob.sdata, prob.mech, len(prob.species))
    pos, rd, ids = kw.pop("pos", prob.pos), prob.rdata, prob.ids
    geom = kw.pop("geom", prob.geom)
    srcs = []
    st = 0
    for _ in range(nsteps):
        if src_box is None:
            src, lo, hi = prob.new_source(), prob.src_lo, prob.src_hi
        else:
            lo, hi = src_box
            src = np.zeros((prob.src_shape[0],) + tuple(h - l + 1 for l, h in zip(lo[::-1], hi[::-1])))
        st, pos, rd, ids = O.update(prob.mech, sd, prob.comps, geom, pos, rd, ids, prob.state, prob.state_lo,
                                    prob.state_hi, src, lo, hi, dt, branch=branch, **kw)
        srcs.append(src)
        if st != 0:
            break
    return st, (pos, rd, ids), srcs


def _gpu(prob, nsteps, dt, form=0, params=None, src_box=None, state_ghosts=None, source_ghosts=None,
         is_ghost=False, pos=None, geom=None, cap=None):
    C = _lib()
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, geom or prob.geom, cap or prob.n, form=form)
    if params:
        pc.set_params(*params)
    try:
        pc.set_particles(prob.pos if pos is None else pos, prob.rdata, prob.ids)
    except C.PelempError as e:
        # the host checks the droplets it is handed (a droplet behind an
        # outflow face is reported here, before any update)
        pc.close()
        return e.code, None, []
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    sg = prob.state_ghosts if state_ghosts is None else state_ghosts
    og = prob.source_ghosts if source_ghosts is None else source_ghosts
    srcs = []
    code = 0
    for _ in range(nsteps):
        lo, hi = (prob.src_lo, prob.src_hi) if src_box is None else src_box
        src = DeviceFab(lo, hi, prob.src_shape[0])
        try:
            pc.updateParticles(0, state, src, dt, 0.0, sg, og, False, is_ghost, True, None, -1.0)
        except C.PelempError as e:
            code = e.code
            break
        srcs.append(src.numpy())
    out = by_id(*pc.get_particles()) if code == 0 else None
    pc.close()
    return code, out, srcs


def _same(g, o, gs, os_):
    gpos, grd, gids = g
    opos, ord_, oids = by_id(*o)
    assert np.array_equal(gids, oids), f"droplet sets differ ({gids.size} vs {oids.size})"
    assert np.array_equal(grd, ord_) and np.array_equal(gpos, opos)
    assert len(gs) == len(os_) and all(np.array_equal(a, b) for a, b in zip(gs, os_))


def test_ghost_droplets_deleted_fp64_plain():
    prob = P.config2(ncell=16)
    dt = 1e-5  # moves of up to ~0.9 cells: droplets cross the periodic faces
    print(f"ghost fp64 plain: {prob.n} droplets", flush=True)
    st, o, os_ = _oracle(prob, 2, dt, is_ghost=True, state_ghosts=2, source_ghosts=0)
    assert st == 0
    code, g, gs = _gpu(prob, 2, dt, is_ghost=True, source_ghosts=0)
    assert code == 0
    assert prob.n - o[2].size > 50  # ghost droplets did leave
    _same(g, o, gs, os_)


def test_ghost_droplets_deleted_branch_kernel():
    prob = P.config3(ncell=(16, 16, 32), drops_per_cell=3.0, z_frac=0.25)
    sd = dict(prob.sdata)
    sd["do_splash"] = 0
    prob.sdata = sd
    dt = prob.flow_dt  # moves below one cell toward the wall (WallFunctions.H:110-120)
    print(f"ghost fp64 branches: {prob.n} droplets", flush=True)
    branch = dict(max_ppp=100.0, ppp_fact=0.5, seed=3, next_id=0, cap=4 * prob.n)
    st, o, os_ = _oracle(prob, 2, dt, branch=branch, is_ghost=True, state_ghosts=prob.state_ghosts, source_ghosts=0)
    assert st == 0
    code, g, gs = _gpu(prob, 2, dt, params=(100.0, 0.5, 3), is_ghost=True, source_ghosts=0, cap=4 * prob.n)
    assert code == 0
    assert prob.n - o[2].size > 10
    _same(g, o, gs, os_)


def test_ghost_droplets_deleted_fp32():
    C = _lib()
    prob = P.config2(ncell=16)
    dt = 1e-5
    code, act, a_src = _gpu(prob, 1, dt, form=C.FORM_FP32, source_ghosts=0)
    assert code == 0
    code, gh, g_src = _gpu(prob, 1, dt, form=C.FORM_FP32, source_ghosts=0, is_ghost=True)
    assert code == 0
    # the deposits come first: the same sources
    assert np.array_equal(a_src[0], g_src[0])
    # the ghost call keeps the active call's droplets that did not cross a face
    L = np.array(prob.geom["phi"])
    o = np.argsort(prob.ids)
    p0 = prob.pos[o][np.searchsorted(prob.ids[o], act[2])]
    crossed = np.any(np.abs(act[0] - p0) > 0.5 * L, axis=1)
    assert crossed.sum() > 50
    assert np.array_equal(gh[2], act[2][~crossed])
    assert np.array_equal(gh[1], act[1][~crossed]) and np.array_equal(gh[0], act[0][~crossed])


def test_state_box_too_small_is_reported():
    C = _lib()
    prob = P.config1(scale=0.05)
    st, _, _ = _oracle(prob, 1, prob.flow_dt, state_ghosts=0)
    assert st == C.E_STATE_BOX
    code, _, _ = _gpu(prob, 1, prob.flow_dt, state_ghosts=0)
    assert code == C.E_STATE_BOX
    code, _, _ = _gpu(prob, 1, prob.flow_dt, state_ghosts=1)
    assert code == 0


def test_source_box_too_small_is_reported_and_ghosts_drop_it():
    C = _lib()
    prob = P.config1(scale=0.05)
    n = prob.geom["domain_hi"]
    box = ((0, 0, 0), (n[0], n[1], n[2] - 1))  # the top z layer left out
    st, _, _ = _oracle(prob, 1, prob.flow_dt, src_box=box)
    assert st == C.E_SRC_BOX
    code, _, _ = _gpu(prob, 1, prob.flow_dt, src_box=box)
    assert code == C.E_SRC_BOX
    st, o, os_ = _oracle(prob, 2, prob.flow_dt, src_box=box, is_ghost=True)
    assert st == 0
    code, g, gs = _gpu(prob, 2, prob.flow_dt, src_box=box, is_ghost=True, source_ghosts=-1)
    assert code == 0
    _same(g, o, gs, os_)


def test_left_domain_is_reported():
    C = _lib()
    prob = P.config1(scale=0.01)
    geom = dict(prob.geom, bndry_lo=[0, 0, -1], bndry_hi=[0, 0, -1])
    pos = prob.pos.copy()
    pos[7, 2] = -0.3 * prob.geom["dx"][2]
    st, _, _ = _oracle(prob, 1, prob.flow_dt, geom=geom, pos=pos)
    assert st == C.E_LEFT_DOMAIN
    code, _, _ = _gpu(prob, 1, prob.flow_dt, geom=geom, pos=pos)
    assert code == C.E_LEFT_DOMAIN
