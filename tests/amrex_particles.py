"""A numpy reader and writer of AMReX's particle checkpoint layout, written
from the published format (AMReX_WriteBinaryParticleData.H,
AMReX_ParticleIO.H), independent of the library's C++ code: the checks of
pelemp_spray_checkpoint / pelemp_particles_* read the library's files with
it and hand the library files made with it.

Header: version string, space dimension, real-component count and names,
int-component count and names, the legacy checkpoint flag, particle count,
next id, finest level, grid counts per level, then per level and grid
"which count where".  Level_L/DATA_nnnnn: per grid, count x (2 + ni) int32
(idcpu split high/low for Version_Two_Dot_One, plain id and cpu for
Version_Two_Dot_Zero), then count x (3 + nr) reals."""
import os

import numpy as np

ID_BITS = 39


def pack_idcpu(ids, cpus):
    """int32 pairs of AMReX's idcpu word: bit 63 set for id >= 0, |id| in
    bits 24-62, cpu in bits 0-23 (AMReX_Particle.H, ParticleIDWrapper /
    ParticleCPUWrapper)."""
    ids = np.asarray(ids, dtype=np.int64)
    w = np.where(ids >= 0, np.uint64(1) << np.uint64(63), np.uint64(0))
    w = w | (np.abs(ids).astype(np.uint64) << np.uint64(24)) | np.asarray(cpus, dtype=np.uint64)
    hi = (w >> np.uint64(32)).astype(np.uint32).view(np.int32)
    lo = (w & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    return np.stack([hi, lo], axis=1)


def unpack_idcpu(pairs):
    pairs = np.asarray(pairs, dtype=np.int32)
    w = (pairs[:, 0].view(np.uint32).astype(np.uint64) << np.uint64(32)) | pairs[:, 1].view(np.uint32).astype(
        np.uint64)
    sign = (w >> np.uint64(63)).astype(bool)
    mag = ((w >> np.uint64(24)) & np.uint64((1 << ID_BITS) - 1)).astype(np.int64)
    cpu = (w & np.uint64(0xFFFFFF)).astype(np.int64)
    return np.where(sign, mag, -mag), cpu


def read(pdir):
    """dict with the header fields and, per level, a list of grids
    (ids, cpus, pos, rdata)."""
    tok = open(os.path.join(pdir, "Header")).read().split()
    it = iter(tok)
    h = {"version": next(it), "dim": int(next(it))}
    nr = int(next(it))
    h["real_names"] = [next(it) for _ in range(nr)]
    ni = int(next(it))
    h["int_names"] = [next(it) for _ in range(ni)]
    h["is_checkpoint"] = int(next(it))
    h["nparticles"] = int(next(it))
    h["next_id"] = int(next(it))
    h["finest"] = int(next(it))
    h["ngrids"] = [int(next(it)) for _ in range(h["finest"] + 1)]
    h["grids"] = [[(int(next(it)), int(next(it)), int(next(it))) for _ in range(ng)] for ng in h["ngrids"]]
    assert next(it, None) is None
    single = h["version"].endswith("_single")
    packed = h["version"].startswith("Version_Two_Dot_One")
    rt = np.float32 if single else np.float64
    h["levels"] = []
    for lev, grids in enumerate(h["grids"]):
        out = []
        for which, count, where in grids:
            if count == 0:
                out.append(None)
                continue
            fn = os.path.join(pdir, f"Level_{lev}", f"DATA_{which:05d}")
            with open(fn, "rb") as f:
                f.seek(where)
                iv = np.fromfile(f, dtype=np.int32, count=count * (2 + ni)).reshape(count, 2 + ni)
                rv = np.fromfile(f, dtype=rt, count=count * (3 + nr)).reshape(count, 3 + nr)
            if packed:
                ids, cpus = unpack_idcpu(iv[:, :2])
            else:
                ids, cpus = iv[:, 0].astype(np.int64), iv[:, 1].astype(np.int64)
            out.append((ids, cpus, rv[:, :3].astype(np.float64), rv[:, 3:].astype(np.float64)))
        h["levels"].append(out)
    return h


def write(pdir, real_names, levels, next_id, version="Version_Two_Dot_One_double", files_per_level=1):
    """Writes a checkpoint from levels = [[grid, ...], ...], each grid None
    or (ids, cpus, pos, rdata); grids go round-robin over files_per_level
    DATA files, as AMReX's writers spread grids over ranks' files."""
    single = version.endswith("_single")
    packed = version.startswith("Version_Two_Dot_One")
    rt = np.float32 if single else np.float64
    os.makedirs(pdir, exist_ok=True)
    entries = []
    n = 0
    for lev, grids in enumerate(levels):
        ldir = os.path.join(pdir, f"Level_{lev}")
        os.makedirs(ldir, exist_ok=True)
        files = [open(os.path.join(ldir, f"DATA_{k:05d}"), "wb") for k in range(files_per_level)]
        ent = []
        for g, grid in enumerate(grids):
            k = g % files_per_level
            f = files[k]
            where = f.tell()
            if grid is None:
                ent.append((k, 0, where))
                continue
            ids, cpus, pos, rd = grid
            iv = pack_idcpu(ids, cpus) if packed else np.stack([ids, cpus], axis=1).astype(np.int32)
            iv.astype(np.int32).tofile(f)
            np.concatenate([pos, rd], axis=1).astype(rt).tofile(f)
            ent.append((k, len(ids), where))
            n += int(np.count_nonzero(np.asarray(ids) >= 0))
        for f in files:
            f.close()
        entries.append(ent)
    with open(os.path.join(pdir, "Header"), "w") as h:
        h.write(f"{version}\n3\n{len(real_names)}\n")
        for nm in real_names:
            h.write(nm + "\n")
        h.write(f"0\n1\n{n}\n{next_id}\n{len(levels) - 1}\n")
        for ent in entries:
            h.write(f"{len(ent)}\n")
        for ent in entries:
            for which, count, where in ent:
                h.write(f"{which} {count} {where}\n")
