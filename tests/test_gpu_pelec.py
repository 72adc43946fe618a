"""GPU PeleC flavour (CGS state rho, rho*u, rho*E, rho*Y, T; EY2T in the
gather; drag work and kinetic energy in the energy source) against the
oracle's PeleC flavour on identical CGS inputs, with and without the wall /
splash / film / TAB branches, bitwise in the FP64 form.  Both sides convert
the CGS droplets they are handed to SI with the same arithmetic; the oracle
is handed CGS droplets at every call, so over several steps the library is
handed the same droplets before each step (a library that kept its SI
droplets would differ by one conversion rounding per step, enough to move
the heat-transfer fixed point's stopping step)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from pelemp_amd import problem as P
from tests.helpers import by_id

pytestmark = pytest.mark.gpu


def _run_both(prob, nsteps, cap_factor=4):
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    cap = cap_factor * prob.n
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, cap, flavour="pelec")
    pc.set_params(100.0, 0.5, 7)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    branch = dict(max_ppp=100.0, ppp_fact=0.5, seed=7, next_id=0, cap=cap, pelec=1)
    pos, rd, ids = prob.pos, prob.rdata, prob.ids
    for step in range(nsteps):
        if step > 0:
            pc.set_particles(pos, rd, ids)
        src_o = prob.new_source()
        st, pos, rd, ids = O.update(prob.mech, dict(prob.sdata), prob.comps, prob.geom, pos, rd, ids, prob.state,
                                    prob.state_lo, prob.state_hi, src_o, prob.src_lo, prob.src_hi, prob.flow_dt,
                                    branch=branch)
        assert st == 0, f"oracle status {st} at step {step}"
        src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True, None, -1.0)
        gs = src_g.numpy()
        assert np.array_equal(gs, src_o), f"sources differ at step {step}"
    gpos, grd, gids = pc.get_particles()
    return pc, by_id(gpos, grd, gids), by_id(pos, rd, ids)


@pytest.mark.parametrize("case", ["config2", "config3_branches"])
def test_pelec_flavour_bitwise(case):
    # with the branches one step: a breakup or splash decision must not hinge
    # on the per-step conversion rounding of the oracle's input
    if case == "config2":
        base = P.config2(ncell=16)
        nsteps = 3
    else:
        base = P.config3(ncell=(16, 16, 32), drops_per_cell=3.0, z_frac=0.25)
        nsteps = 1
    prob = P.to_pelec(base, T_guess_factor=1.02)
    pc, g, o = _run_both(prob, nsteps)
    assert np.array_equal(g[2], o[2])
    assert np.array_equal(g[0], o[0])
    assert np.array_equal(g[1], o[1])
    if case == "config3_branches":
        assert o[2].max() > prob.n  # secondary droplets
    # the time step in the flavour's units (cm, cm/s) and the jet bound
    dt = pc.estTimestep(0, 0.5, 2.0e4)
    vel = g[1][:, 0:3]
    dxi = 1.0 / np.array(prob.geom["dx"])
    ref = min(0.5 / np.max(np.abs(vel) * dxi, axis=1).max(), 0.5 * prob.geom["dx"][0] / 2.0e4)
    assert dt == pytest.approx(ref, rel=1e-12)
    pc.close()


def test_pelec_derived_vars_in_cgs():
    from pelemp_amd.spray import DeviceFab

    prob = P.to_pelec(P.config2(ncell=16), T_guess_factor=1.02)
    pc, g, _ = _run_both(prob, 1)
    nvar = pc.NumDeriveVars(1)
    n = 16
    mf_d = DeviceFab((0, 0, 0), (n - 1,) * 3, nvar)
    pc.computeDerivedVars(mf_d, 0, 0, 1)
    ref = np.zeros((nvar, n, n, n))
    gpos, grd, _ = g
    assert O.compute_derived(prob.sdata, prob.geom, gpos, grd, ref, (0, 0, 0), (n - 1,) * 3, 0, True, True) == 0
    got = mf_d.numpy()
    # the oracle converts the CGS droplets handed out by the library back to
    # SI, one rounding away from the library's own SI values
    scale = np.maximum(np.abs(ref).max(axis=(1, 2, 3), keepdims=True), 1e-300)
    assert np.max(np.abs(got - ref) / scale) < 1e-12
    # mass in g: the total liquid mass matches pi/6 rho_L d^3 N in CGS
    assert got[0].sum() > 0
    pc.close()
