"""The source overwrite mode (pelemp_spray_set_source_mode): an update in
PELEMP_SOURCE_OVERWRITE writes every domain cell of the source FAB with the
value an update in the reference's accumulate mode adds into a zeroed FAB,
bit for bit, whatever the FAB held before; ghost cells keep their values.
Covered: tiles split across workgroups (a crowded cell), empty tiles (the
dense jet), two particle sub-cycles (the second adds), the FP64, FP32 and
surrogate forms."""
import numpy as np
import pytest

from pelemp_amd import problem as P

pytestmark = pytest.mark.gpu

FILL = 7.25


def _lib():
    from pelemp_amd import _capi as C

    if C.device_count() < 1:
        pytest.skip("no HIP device")
    return C


def _run(prob, form, mode, nsteps, spray_cfl=-1.0, state_dtype=np.float64, net=False):
    C = _lib()
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n, form=form)
    if net:
        from pelemp_amd import surrogate as S

        S.Surrogate.load(S.default_path(prob.mech)).install(pc)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    pc.set_source_mode(mode)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state.astype(state_dtype),
                      dtype=state_dtype)
    out = []
    for _ in range(nsteps):
        src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        if mode == C.SOURCE_OVERWRITE:
            src.setVal(FILL)
        pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True, None, spray_cfl)
        out.append(src.numpy())
    pc.close()
    return out


def _check(prob, acc, ovw):
    g = prob.source_ghosts
    for a, b in zip(acc, ovw):
        inner = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
        assert np.array_equal(a[inner], b[inner])
        ghost = np.ones(b.shape[1:], bool)
        ghost[g:-g, g:-g, g:-g] = False
        assert np.all(b[:, ghost] == FILL)


def _crowded(prob, m=9000):
    rng = np.random.default_rng(3)
    sel = rng.choice(prob.n, size=m, replace=False)
    dx = np.array(prob.geom["dx"])
    prob.pos[sel] = (np.array([6, 9, 4]) + rng.uniform(0.05, 0.95, (m, 3))) * dx
    return prob


@pytest.mark.parametrize("form", [0, 1, 2])
def test_overwrite_equals_accumulate_into_zero(form):
    C = _lib()
    prob = _crowded(P.config2(ncell=16))
    print(f"source mode: form {form}", flush=True)
    net = form == C.FORM_SURROGATE
    acc = _run(prob, form, C.SOURCE_ACCUMULATE, 2, net=net)
    ovw = _run(prob, form, C.SOURCE_OVERWRITE, 2, net=net)
    _check(prob, acc, ovw)


def test_overwrite_with_subcycles_and_empty_tiles():
    C = _lib()
    prob = P.config5(ncell=32, n_drops=5e4)
    print("source mode: jet, 2 sub-cycles", flush=True)
    acc = _run(prob, C.FORM_FP32, C.SOURCE_ACCUMULATE, 2, spray_cfl=0.9, state_dtype=np.float32)
    ovw = _run(prob, C.FORM_FP32, C.SOURCE_OVERWRITE, 2, spray_cfl=0.9, state_dtype=np.float32)
    _check(prob, acc, ovw)
    acc = _run(prob, C.FORM_FP64, C.SOURCE_ACCUMULATE, 1, spray_cfl=0.9)
    ovw = _run(prob, C.FORM_FP64, C.SOURCE_OVERWRITE, 1, spray_cfl=0.9)
    _check(prob, acc, ovw)
