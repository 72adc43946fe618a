"""FP32 form (FP32 droplet state and ODE, FP64 source sums) against the FP64
CPU oracle on identical seeded inputs.

Tolerances (stated, not bitwise): after 5 steps of config 2 (small box),
per-droplet diameter and temperature within 1e-4 relative, velocity within
1e-4 of max(|u|, 1 m/s), position within 1e-9 m; total evaporated mass
within 0.5 % of the oracle's; every source component within 1e-3 of its
largest magnitude.
"""
import numpy as np
import pytest

from pelemp_amd import _capi as C
from pelemp_amd import problem as P
from tests.helpers import by_id, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _liquid_mass(prob, rd, sdata):
    nf = prob.nf
    T = rd[:, 3]
    inv = np.zeros_like(T)
    for s in range(nf):
        a, b, c, d = sdata["rho_coef"][4 * s:4 * s + 4]
        inv += rd[:, 5 + s] / (a + T * (b + T * (c + T * d)))
    return float(np.sum(np.pi / 6.0 / inv * rd[:, 4] ** 3 * rd[:, 5 + nf]))


def test_fp32_form_vs_fp64_oracle():
    prob = P.config2(ncell=16)
    nsteps = 5
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps, form=C.FORM_FP32)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    gpos, grd, gids = by_id(gpos, grd, gids)
    opos, ord_, oids = by_id(opos, ord_, oids)
    assert np.array_equal(gids, oids)
    assert np.max(np.abs(grd[:, 4] / ord_[:, 4] - 1)) < 1e-4
    assert np.max(np.abs(grd[:, 3] / ord_[:, 3] - 1)) < 1e-4
    uscale = np.maximum(np.abs(ord_[:, 0:3]), 1.0)
    assert np.max(np.abs(grd[:, 0:3] - ord_[:, 0:3]) / uscale) < 1e-4
    L = np.array(prob.geom["phi"])
    dpos = np.abs(gpos - opos)
    dpos = np.minimum(dpos, L - dpos)  # periodic
    assert np.max(dpos) < 1e-9
    # evaporated mass = vapour deposited into the gas: sum of the mass source
    # times cell volume times dt over all steps and cells
    vol = np.prod(prob.geom["dx"])
    rc = prob.comps["rhoSrcIndx"]
    evap_o = sum(float(np.sum(s[rc])) for s in osrc) * vol * prob.flow_dt
    evap_g = sum(float(np.sum(s[rc])) for s in gsrc) * vol * prob.flow_dt
    assert evap_o > 0
    assert abs(evap_g / evap_o - 1) < 5e-3
    for gs, os_ in zip(gsrc, osrc):
        scale = np.max(np.abs(os_), axis=(1, 2, 3), keepdims=True)
        assert np.max(np.abs(gs - os_) / np.maximum(scale, 1e-300)) < 1e-3


def test_fp32_form_with_swapped_deposition_species():
    # dep_fuel_species in the FP32 form: the decane vapour goes to NC12H26
    # and the dodecane vapour to NC10H22, in a gas holding some of both;
    # the same tolerances against the FP64 oracle
    from tests.test_gpu_parity import _with_fuel_vapour

    prob = _with_fuel_vapour(P.config2(ncell=16))
    prob.sdata = dict(prob.sdata, dep_indx=[prob.sdata["indx"][1], prob.sdata["indx"][0]])
    nsteps = 3
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps, form=C.FORM_FP32)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    gpos, grd, gids = by_id(gpos, grd, gids)
    opos, ord_, oids = by_id(opos, ord_, oids)
    assert np.array_equal(gids, oids)
    assert np.max(np.abs(grd[:, 4] / ord_[:, 4] - 1)) < 1e-4
    assert np.max(np.abs(grd[:, 3] / ord_[:, 3] - 1)) < 1e-4
    cm = prob.comps
    for gs, os_ in zip(gsrc, osrc):
        scale = np.max(np.abs(os_), axis=(1, 2, 3), keepdims=True)
        # a saturated species has a round-off source in the FP64 method (its
        # mass fraction renormalised) and an exact zero in the FP32 form's
        # mass bookkeeping: species are measured against the mass source
        m = scale[cm["rhoSrcIndx"]]
        for k in range(prob.nf):
            scale[cm["specSrcIndx"] + k] = np.maximum(scale[cm["specSrcIndx"] + k], m)
        assert np.max(np.abs(gs - os_) / np.maximum(scale, 1e-300)) < 1e-3


def test_fp32_form_refuses_wall_films():
    # the FP32 form has no walls and so no film-height field
    from pelemp_amd._capi import PelempError
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=8)
    rd = prob.rdata.copy()
    rd[0, 9 + prob.nf] = 1e-6  # a film height
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n, form=C.FORM_FP32)
    with pytest.raises(PelempError):
        pc.set_particles(prob.pos, rd, prob.ids)
    pc.close()


def test_fp32_form_decane_mechanism():
    prob = P.decane_case(ncell=16)
    nsteps = 4
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps, form=C.FORM_FP32)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    gpos, grd, gids = by_id(gpos, grd, gids)
    opos, ord_, oids = by_id(opos, ord_, oids)
    assert np.array_equal(gids, oids)
    assert np.max(np.abs(grd[:, 4] / ord_[:, 4] - 1)) < 1e-4
    assert np.max(np.abs(grd[:, 3] / ord_[:, 3] - 1)) < 1e-4
    for gs, os_ in zip(gsrc, osrc):
        scale = np.max(np.abs(os_), axis=(1, 2, 3), keepdims=True)
        assert np.max(np.abs(gs - os_) / np.maximum(scale, 1e-300)) < 1e-3
