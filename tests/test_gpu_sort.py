"""Long per-cell groups in the droplet sort (sort_kernels.h): thousands of
droplets packed into one cell, which then cross into the next cells together,
take the workgroup sort of long segments in the full sort (set_particles)
and in the incremental one (the movers into a cell).  The order within a
cell is the stable storage order the deposits are summed in, so sources and
droplets stay bitwise equal to the oracle's."""
import numpy as np
import pytest

from pelemp_amd import problem as P
from tests.helpers import by_id, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


# 40: one listed segment sorted in LDS; 5000 and 9000: several LDS chunks
# merged through the scratch buffer (one and two merge passes)
@pytest.mark.parametrize("m", [40, 5000, 9000])
def test_crowded_cell_stays_bitwise(m):
    prob = P.config2(ncell=16)
    g = prob.geom
    dx = np.array(g["dx"])
    plo = np.array(g["plo"])
    rng = np.random.default_rng(m)
    assert prob.n > 2 * m
    sel = rng.choice(prob.n, size=m, replace=False)
    cell = np.array([5, 6, 7])
    frac = np.stack([rng.uniform(0.75, 0.98, m), rng.uniform(0.02, 0.98, m), rng.uniform(0.02, 0.98, m)], axis=1)
    prob.pos[sel] = plo + (cell + frac) * dx
    # 0.6 of a cell per step along +x: the group moves cell to cell together
    prob.rdata[sel, 0] = 0.6 * dx[0] / prob.flow_dt
    prob.rdata[sel, 1] = 0.0
    prob.rdata[sel, 2] = 0.0
    nsteps = 3
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    for step, (a, b) in enumerate(zip(gsrc, osrc)):
        assert np.array_equal(a, b), f"sources differ at step {step}"
    gp, gr, gi = by_id(gpos, grd, gids)
    op, or_, oi = by_id(opos, ord_, oids)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gp, op)
    assert np.array_equal(gr, or_)
    # the group really crossed: most of it sits two cells on after the steps
    moved = np.floor((gp[np.isin(gi, prob.ids[sel])] - plo) / dx).astype(int)
    assert np.count_nonzero(moved[:, 0] >= cell[0] + 1) > 0.9 * m


def _heavy_cell(prob, m, seed, drift_cells):
    """m droplets packed into one cell near its +x face; drift_cells of a
    cell per step along +x (a part of the group crosses, the rest stays)."""
    g = prob.geom
    dx = np.array(g["dx"])
    plo = np.array(g["plo"])
    rng = np.random.default_rng(seed)
    sel = rng.choice(prob.n, size=m, replace=False)
    cell = np.array([9, 3, 12])
    frac = np.stack([rng.uniform(0.0, 0.98, m), rng.uniform(0.02, 0.98, m), rng.uniform(0.02, 0.98, m)], axis=1)
    prob.pos[sel] = plo + (cell + frac) * dx
    prob.rdata[sel, 0] = drift_cells * dx[0] / prob.flow_dt * rng.uniform(0.2, 1.0, m)
    prob.rdata[sel, 1] = 0.0
    prob.rdata[sel, 2] = 0.0
    return sel


# a tile of more than PELEMP_SEGMENT_DROPLETS (4096) droplets is updated by
# several workgroups; the cell's stayers then sit in several segments, whose
# partial sums are added in segment order (the oracle sums the same way)
@pytest.mark.parametrize("m", [4097, 12000])
def test_heavy_tile_segments_bitwise(m):
    prob = P.config2(ncell=16)
    _heavy_cell(prob, m, m + 1, 0.5)
    nsteps = 3
    print(f"heavy tile fp64: {m} droplets in one cell", flush=True)
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    for step, (a, b) in enumerate(zip(gsrc, osrc)):
        assert np.array_equal(a, b), f"sources differ at step {step}"
    gp, gr, gi = by_id(gpos, grd, gids)
    op, or_, oi = by_id(opos, ord_, oids)
    assert np.array_equal(gi, oi) and np.array_equal(gp, op) and np.array_equal(gr, or_)


def test_heavy_tile_segments_fp32():
    from pelemp_amd import _capi as C

    prob = P.config2(ncell=16)
    sel = _heavy_cell(prob, 12000, 5, 0.5)
    nsteps = 3
    print("heavy tile fp32: 12000 droplets in one cell", flush=True)
    _, gpos, grd, gids, gsrc = run_gpu(prob, nsteps, form=C.FORM_FP32)
    opos, ord_, oids, osrc = run_oracle(prob, nsteps)
    gp, gr, gi = by_id(gpos, grd, gids)
    op, or_, oi = by_id(opos, ord_, oids)
    assert np.array_equal(gi, oi)
    assert np.max(np.abs(gr[:, 4] - or_[:, 4]) / or_[:, 4]) < 1e-4  # diameter
    assert np.max(np.abs(gr[:, 3] - or_[:, 3]) / or_[:, 3]) < 1e-4  # temperature
    assert np.max(np.abs(gp - op)) < 1e-3 * prob.geom["dx"][0]
    # the sources of the crowded cell: FP64 sums of FP32 contributions
    for a, b in zip(gsrc, osrc):
        scale = np.max(np.abs(b), axis=(1, 2, 3), keepdims=True)
        assert np.max(np.abs(a - b) / np.maximum(scale, 1e-300)) < 1e-3
    # the packed droplets are all still there, in the cell and its neighbour
    crowd = np.isin(gi, prob.ids[sel])
    assert crowd.sum() == 12000
