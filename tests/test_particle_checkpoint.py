"""AMReX's particle checkpoint layout (the format of SprayParticleIO,
SprayIO.cpp:6-34, and of Restart, SpraySetup.cpp:320-325) through the
host-only entry points pelemp_particles_write_checkpoint / _info /
_read_checkpoint, checked with the independent numpy reader and writer in
tests/amrex_particles.py:
 * known answers of the idcpu packing, derived by hand from the
   sign-magnitude layout (id 5 on rank 0 -> 0x80000000 0x05000000);
 * the library's files read back by the numpy reader: header lines, the box
   array text of Particle_H, grid assignment, every value bit for bit;
 * files written by the numpy writer (several grids and DATA files, empty
   grids, an invalid particle, two levels, Version_Two_Dot_Zero and _single
   data) read by the library;
 * errors: a particle in no box, a bad version string, truncated data, a
   level not in the file.
No GPU: these entry points do not touch the device."""
import os

import numpy as np
import pytest

from pelemp_amd import _capi as C
from tests import amrex_particles as A

NAMES = ["xvel", "yvel", "zvel", "temperature", "diam", "spray_mf_NC12H26", "number_density", "num_dens0",
         "unused1", "unused2", "wall_film_height"]


def _lib():
    if not os.path.exists(C.LIB_PATH):
        pytest.skip("library not built")
    return C


def _particles(n, seed=0, cpu_frac=0.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 1.0, (n, 3))
    rd = rng.standard_normal((n, len(NAMES))) * 10.0 ** rng.integers(-12, 4, (n, len(NAMES)))
    ids = rng.choice(np.arange(1, 50 * n + 2), size=n, replace=False).astype(np.int64)
    cpus = np.where(rng.uniform(size=n) < cpu_frac, rng.integers(1, 1 << 24, n), 0)
    return pos, rd, ids, cpus


def test_idcpu_known_answers():
    pairs = A.pack_idcpu([5, 300, 1, (1 << 39) - 1], [0, 7, (1 << 24) - 1, 0])
    u = pairs.view(np.uint32)
    assert list(u[0]) == [0x80000000, 0x05000000]
    assert list(u[1]) == [0x80000001, 0x2C000007]  # 300 = 0x12C
    assert list(u[2]) == [0x80000000, 0x01FFFFFF]
    assert list(u[3]) == [0xFFFFFFFF, 0xFF000000]
    ids, cpus = A.unpack_idcpu(pairs)
    assert list(ids) == [5, 300, 1, (1 << 39) - 1] and list(cpus) == [0, 7, (1 << 24) - 1, 0]
    # a negative id clears the sign bit: AMReX's invalid particles
    assert A.unpack_idcpu(A.pack_idcpu([-9], [3]))[0][0] == -9


def test_library_checkpoint_read_by_numpy(tmp_path):
    C = _lib()
    n = 2000
    pos, rd, ids, cpus = _particles(n, seed=1, cpu_frac=0.3)
    held = ids + (cpus << A.ID_BITS)
    boxes = [((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 7, 15)), ((8, 8, 0), (15, 15, 15)), ((16, 0, 0), (17, 1, 1))]
    plo, dx = (0.0, 0.0, 0.0), (1.0 / 16, 1.0 / 16, 1.0 / 16)
    C.write_particle_checkpoint(str(tmp_path), "particles", pos, rd, held, NAMES, 123456, boxes, plo, dx)
    pdir = str(tmp_path / "particles")
    h = A.read(pdir)
    assert h["version"] == "Version_Two_Dot_One_double" and h["dim"] == 3
    assert h["real_names"] == NAMES and h["int_names"] == []
    assert h["is_checkpoint"] == 1 and h["nparticles"] == n and h["next_id"] == 123456
    assert h["finest"] == 0 and h["ngrids"] == [len(boxes)]
    txt = open(os.path.join(pdir, "Level_0", "Particle_H")).read()
    assert txt == ("(4 0\n((0,0,0) (7,15,15) (0,0,0))\n((8,0,0) (15,7,15) (0,0,0))\n"
                   "((8,8,0) (15,15,15) (0,0,0))\n((16,0,0) (17,1,1) (0,0,0))\n)\n")
    cell = np.floor(pos / np.array(dx)).astype(int)
    grid = np.where(cell[:, 0] < 8, 0, np.where(cell[:, 1] < 8, 1, 2))
    size = os.path.getsize(os.path.join(pdir, "Level_0", "DATA_00000"))
    assert size == n * (2 * 4 + (3 + len(NAMES)) * 8)
    for g in range(len(boxes)):
        which, count, where = h["grids"][0][g]
        sel = np.flatnonzero(grid == g)
        assert which == 0 and count == sel.size
        if count == 0:
            assert h["levels"][0][g] is None
            continue
        gid, gcpu, gpos, grd = h["levels"][0][g]
        assert np.array_equal(gid, ids[sel]) and np.array_equal(gcpu, cpus[sel])
        assert np.array_equal(gpos, pos[sel]) and np.array_equal(grd, rd[sel])
    # and back through the library: same particles, grid by grid
    p2, r2, i2, nxt = C.read_particle_checkpoint(str(tmp_path), "particles")
    order = np.concatenate([np.flatnonzero(grid == g) for g in range(len(boxes))])
    assert nxt == 123456
    assert np.array_equal(i2, held[order]) and np.array_equal(p2, pos[order]) and np.array_equal(r2, rd[order])


@pytest.mark.parametrize("version", ["Version_Two_Dot_One_double", "Version_Two_Dot_Zero_double",
                                     "Version_Two_Dot_One_single"])
def test_numpy_checkpoint_read_by_library(tmp_path, version):
    C = _lib()
    grids0 = []
    want = []
    for g, m in enumerate([300, 0, 17, 1, 250]):
        if m == 0:
            grids0.append(None)
            continue
        pos, rd, ids, cpus = _particles(m, seed=10 + g, cpu_frac=0.0 if "Zero" in version else 0.5)
        if "single" in version:
            pos, rd = pos.astype(np.float32).astype(np.float64), rd.astype(np.float32).astype(np.float64)
        ids = ids + 100000 * g
        if g == 2:
            ids[4] = -ids[4]  # invalid: skipped
        grids0.append((ids, cpus, pos, rd))
        keep = ids >= 0
        want.append((ids[keep] + (cpus[keep] << A.ID_BITS), pos[keep], rd[keep]))
    pos1, rd1, ids1, cpus1 = _particles(40, seed=99)
    grids1 = [(ids1 + 10 ** 7, cpus1, pos1, rd1)]
    if "single" in version:
        grids1 = [(ids1 + 10 ** 7, cpus1, pos1.astype(np.float32).astype(np.float64),
                   rd1.astype(np.float32).astype(np.float64))]
    A.write(str(tmp_path / "particles"), NAMES, [grids0, grids1], 777, version=version, files_per_level=3)
    p, r, i, nxt = C.read_particle_checkpoint(str(tmp_path), "particles", 0)
    # grids are read in header order
    assert nxt == 777
    assert np.array_equal(i, np.concatenate([w[0] for w in want]))
    assert np.array_equal(p, np.concatenate([w[1] for w in want]))
    assert np.array_equal(r, np.concatenate([w[2] for w in want]))
    p, r, i, _ = C.read_particle_checkpoint(str(tmp_path), "particles", 1)
    assert np.array_equal(i, grids1[0][0] + (grids1[0][1] << A.ID_BITS))
    assert np.array_equal(p, grids1[0][2]) and np.array_equal(r, grids1[0][3])


def test_checkpoint_errors(tmp_path):
    C = _lib()
    pos, rd, ids, _ = _particles(10, seed=3)
    with pytest.raises(C.PelempError):  # a particle in no box
        C.write_particle_checkpoint(str(tmp_path / "a"), "particles", pos, rd, ids, NAMES, 11,
                                    [((0, 0, 0), (0, 3, 3)), ((4, 0, 0), (7, 3, 3))], (0, 0, 0), (0.5, 0.5, 0.5))
    C.write_particle_checkpoint(str(tmp_path / "b"), "particles", pos, rd, ids, NAMES, 11)
    pdir = tmp_path / "b" / "particles"
    with pytest.raises(C.PelempError):  # level not in the file
        C.read_particle_checkpoint(str(tmp_path / "b"), "particles", 1)
    data = pdir / "Level_0" / "DATA_00000"
    raw = data.read_bytes()
    data.write_bytes(raw[:-8])
    with pytest.raises(C.PelempError):  # truncated
        C.read_particle_checkpoint(str(tmp_path / "b"), "particles")
    data.write_bytes(raw)
    hdr = pdir / "Header"
    txt = hdr.read_text()
    hdr.write_text(txt.replace("Version_Two_Dot_One_double", "Version_One_Dot_One_triple"))
    with pytest.raises(C.PelempError):
        C.read_particle_checkpoint(str(tmp_path / "b"), "particles")
    hdr.write_text(txt)
    assert np.array_equal(C.read_particle_checkpoint(str(tmp_path / "b"), "particles")[2], ids)
    with pytest.raises(C.PelempError):  # no checkpoint there
        C.read_particle_checkpoint(str(tmp_path / "none"), "particles")
