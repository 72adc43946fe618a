"""GPU FP64 branch kernels vs the CPU oracle on identical seeded inputs: wall
films (calculateFilmSource, fillFilmFab), TAB breakup, specular walls,
splash and the secondary-droplet pass (SprayParticles.cpp:275-497,
WallFunctions.H, AhamedSplash.H, TABBreakup.H, SpraySB.cpp).

The problem is a box with solid walls at z = lo and z = hi, periodic in x and
y, and droplets moving towards the lower wall (10 % of the parcels start in
the lowest quarter).  Bar: every droplet (original and secondary, matched by
id) and every source component within 1e-10 relative; the kernels use the
oracle's elementary functions and association order, so the comparison is
also checked bit for bit."""
import numpy as np
import pytest

from pelemp_amd import problem as P
from tests.helpers import by_id, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _run_both(prob, sdata, nsteps, seed=11, cap_factor=20, do_move=True, avg_inject_mass=0.0, error_at=None):
    from oracle import oracle_py as O
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    cap = cap_factor * prob.n
    sd_o = O.latent_shifted(sdata, prob.mech, len(prob.species))
    branch = dict(max_ppp=100.0, ppp_fact=0.5, seed=seed, next_id=0, cap=cap, avg_inject_mass=avg_inject_mass)
    pc = SprayParticleContainer(prob.mech, sdata, prob.comps, prob.geom, cap, form=0)
    pc.set_params(100.0, 0.5, seed, avg_inject_mass=avg_inject_mass)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    pos, rd, ids = prob.pos, prob.rdata, prob.ids
    for step in range(nsteps):
        src_o = prob.new_source()
        st, pos, rd, ids = O.update(prob.mech, sd_o, prob.comps, prob.geom, pos, rd, ids, prob.state, prob.state_lo,
                                    prob.state_hi, src_o, prob.src_lo, prob.src_hi, prob.flow_dt, do_move=do_move,
                                    branch=branch)
        src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        if step == error_at:
            # the same failure on both sides, at the same step
            from pelemp_amd import _capi as C

            assert st != 0, f"oracle status 0 at step {step}"
            with pytest.raises(C.PelempError) as e:
                pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False,
                                   False, do_move, None, -1.0)
            assert e.value.code == st
            return None, st
        assert st == 0, f"oracle status {st} at step {step}"
        pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           do_move, None, -1.0)
        gs = src_g.numpy()
        scale = np.max(np.abs(src_o), axis=(1, 2, 3), keepdims=True)
        err = np.max(np.abs(gs - src_o) / np.maximum(scale, 1e-300))
        assert err <= TOL, f"source differs at step {step}: {err}"
        assert pc.num_particles() == ids.size, f"droplet count differs at step {step}"
    gpos, grd, gids = pc.get_particles()
    return by_id(gpos, grd, gids), by_id(pos, rd, ids)


def _compare(g, o, nf):
    gpos, grd, gids = g
    opos, ord_, oids = o
    assert np.array_equal(gids, oids), "surviving droplet sets differ"
    assert rel_err(gpos, opos, floor=1e-3) <= TOL
    assert rel_err(grd[:, 0:3], ord_[:, 0:3], floor=1e-3) <= TOL  # velocity
    assert rel_err(grd[:, 3:5], ord_[:, 3:5]) <= TOL  # T, diameter
    assert rel_err(grd[:, 5:5 + nf], ord_[:, 5:5 + nf], floor=1e-6) <= TOL  # mass fractions
    # number density, N0, TAB y / dy/dt, film height
    assert rel_err(grd[:, 5 + nf:], ord_[:, 5 + nf:], floor=1e-12) <= TOL
    return np.array_equal(grd, ord_) and np.array_equal(gpos, opos)


def _problem(**kw):
    return P.config3(ncell=(16, 16, 32), drops_per_cell=3.0, z_frac=0.25, **kw)


@pytest.mark.parametrize("case", ["walls", "splash_film", "tab", "all"])
def test_config3_small_branch_parity(case):
    prob = _problem()
    sd = dict(prob.sdata)
    if case in ("walls", "tab"):
        sd["do_splash"] = 0
    if case in ("walls", "splash_film"):
        sd["do_breakup"] = 0
    g, o = _run_both(prob, sd, 5)
    n0 = prob.n
    if case in ("splash_film", "all"):
        assert (o[1][:, 5 + prob.nf + 4] > 0).sum() > 0, "no wall film formed"
    if case != "walls":
        assert o[2].max() > n0, "no secondary droplets created"
    bitwise = _compare(g, o, prob.nf)
    assert bitwise, "agreement within 1e-10 but not bitwise"


@pytest.mark.parametrize("case", ["khrt", "khrt_splash", "tab_no_move"])
def test_config3_small_khrt_parity(case):
    # KH-RT stripping / RT breakup with its children (relative speeds up to
    # 30 m/s, mean injected parcel 50 x (40 um)^3), alone and with splash; and
    # the TAB model without a move, where the reference runs the KH-RT update
    # as well (SprayParticles.cpp:401-411)
    prob = _problem(du_max=30.0)
    sd = dict(prob.sdata)
    sd["do_breakup"] = 1 if case == "tab_no_move" else 2
    if case == "khrt":
        sd["do_splash"] = 0
    # with splash, the run stops at step 3 (see the next test)
    nsteps = 3 if case == "khrt_splash" else 5
    g, o = _run_both(prob, sd, nsteps, do_move=(case != "tab_no_move"), avg_inject_mass=50 * 40e-6 ** 3)
    if case != "tab_no_move":
        assert o[2].max() > prob.n, "no secondary droplets created"
    else:
        # stripped parcels: diameters changed with no droplet moved
        assert np.any(o[1][:, 4] != prob.rdata[np.argsort(prob.ids)][:, 4])
    assert _compare(g, o, prob.nf), "agreement within 1e-10 but not bitwise"


def test_khrt_stripped_wall_film_is_an_error():
    # the KH-RT update also runs on wall films (SprayParticles.cpp:398-411
    # has no film check); a film has no velocity, so the direction of its
    # children is 0/0 and the reference aborts ("Bad breakup particle",
    # SpraySB.cpp:187-190).  Both sides report PELEMP_E_BAD_CHILD at the step
    # where the first film is stripped.
    from pelemp_amd import _capi as C

    prob = _problem(du_max=30.0)
    sd = dict(prob.sdata)
    sd["do_breakup"] = 2
    _, st = _run_both(prob, sd, 4, avg_inject_mass=50 * 40e-6 ** 3, error_at=3)
    assert st == C.E_BAD_CHILD


def test_config3_children_ids_and_run_to_run():
    # a second run with the same seed gives the same droplets; another seed
    # moves the splashed children (their angles are drawn per child)
    prob = _problem()
    g1, _ = _run_both(prob, dict(prob.sdata), 3, seed=3)
    g2, o2 = _run_both(prob, dict(prob.sdata), 3, seed=3)
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))
    ids = g1[2]
    assert np.unique(ids).size == ids.size


def test_restart_from_state_with_film_droplets(tmp_path):
    """A state the update makes itself: droplets placed on the reflective
    wall (film droplets among them) sit a rounding error behind it
    (z < plo).  Handed back
    through an AMReX particle checkpoint (Restart), the context accepts them
    (check_bounds keeps droplets behind a reflective face,
    SprayInterpolation.H:65-100) and the next update equals the oracle's
    from the same particles, bit for bit."""
    from oracle import oracle_py as O
    from pelemp_amd import _capi as C
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    prob = P.config3(ncell=(16, 16, 32), drops_per_cell=3.0)
    sdata = prob.sdata
    sd_o = O.latent_shifted(sdata, prob.mech, len(prob.species))
    cap = 20 * prob.n
    branch = dict(max_ppp=100.0, ppp_fact=0.5, seed=5, next_id=0, cap=cap)
    pos, rd, ids = prob.pos, prob.rdata, prob.ids
    for _ in range(2):
        st, pos, rd, ids = O.update(prob.mech, sd_o, prob.comps, prob.geom, pos, rd, ids, prob.state,
                                    prob.state_lo, prob.state_hi, prob.new_source(), prob.src_lo, prob.src_hi,
                                    prob.flow_dt, branch=branch)
        assert st == 0
    behind = pos[:, 2] < prob.geom["plo"][2]
    print(f"restart: {ids.size} droplets, {int(behind.sum())} behind the wall", flush=True)
    assert behind.sum() > 0 and np.any(rd[behind, -1] > 0.0)
    names = [f"c{q}" for q in range(rd.shape[1])]
    C.write_particle_checkpoint(str(tmp_path), "particles", pos, rd, ids, names, branch["next_id"])
    pc = SprayParticleContainer(prob.mech, sdata, prob.comps, prob.geom, cap, form=0)
    pc.set_params(100.0, 0.5, 5)
    pc.Restart(str(tmp_path))
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    src_g = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
    pc.updateParticles(0, state, src_g, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                       True, None, -1.0)
    src_o = prob.new_source()
    st, pos, rd, ids = O.update(prob.mech, sd_o, prob.comps, prob.geom, pos, rd, ids, prob.state, prob.state_lo,
                                prob.state_hi, src_o, prob.src_lo, prob.src_hi, prob.flow_dt, branch=branch)
    assert st == 0
    g = by_id(*pc.get_particles())
    o = by_id(pos, rd, ids)
    pc.close()
    assert np.array_equal(g[2], o[2])
    assert np.array_equal(g[1], o[1]) and np.array_equal(g[0], o[0])
    assert np.array_equal(src_g.numpy(), src_o)
