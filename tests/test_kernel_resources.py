"""Register, scratch and occupancy budgets of the update kernels, read from
the compiler's resource report that the build writes
(pelemp_amd/lib/kernel_resources.json).  A change that pushes the per-droplet
arrays of a plain update kernel to scratch memory (an indexed access into a
register array does) or lowers its occupancy fails here, on the CPU, before
it shows as a slower GPU run."""
import json
import os
import re

import pytest

PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pelemp_amd", "lib",
                    "kernel_resources.json")
# spray_update_kernel<Mech, NF, R, POS64, AT_BOUNDS, BR, EB, TAB, DEP>
_UPD = re.compile(r"spray_update_kernelINS_\d+(Mech_\w+?)ELi(\d+)E([df])((?:Lb[01]E)+)")


def _kernels():
    if not os.path.exists(PATH):
        pytest.skip("no resource report: build the library first")
    with open(PATH) as f:
        return json.load(f)


def _updates():
    out = []
    for name, r in _kernels().items():
        m = _UPD.search(name)
        if m:
            flags = [int(b) for b in re.findall(r"Lb([01])E", m.group(4))]
            out.append((m.group(1), int(m.group(2)), m.group(3), flags, r))
    return out


def test_every_update_variant_is_reported():
    ups = _updates()
    assert ups, "no update kernels in the resource report"
    mechs = {u[0] for u in ups}
    assert {"Mech_dodecane_air", "Mech_decane_dodecane_air", "Mech_alkane6_air", "Mech_decane_air"} <= mechs


# (FP32?, fuels <= 2?) -> (least waves per SIMD, most scratch bytes per lane)
# of the plain kernels that config 2, 4 and 5 run (no walls, own-species
# deposition)
BUDGET = {(True, True): (4, 16), (True, False): (3, 256), (False, True): (3, 128), (False, False): (2, 640)}


def test_plain_update_kernels_keep_their_registers():
    checked = 0
    for mech, nf, r, flags, res in _updates():
        pos64, at_bounds, br, eb, tab, dep = (flags + [1] * 6)[:6]
        if br:
            continue
        waves, scratch = BUDGET[(r == "f", nf <= 2)]
        if r == "f" and nf == 1:
            waves = 5
        if at_bounds:
            waves = min(waves, 2)
        assert res["occupancy"] >= waves, (mech, nf, flags, res)
        if not at_bounds and not dep:
            checked += 1
            assert res.get("scratchsize", 0) <= scratch, (mech, nf, flags, res)
        else:
            # the wall and run-time deposition variants: a few more spills
            assert res.get("scratchsize", 0) <= 1024, (mech, nf, flags, res)
    assert checked == 8


def test_surrogate_kernel_keeps_its_registers():
    n = 0
    for name, res in _kernels().items():
        if "spray_update_surr_kernel" in name:
            n += 1
            assert res["occupancy"] >= 2 and res.get("scratchsize", 0) <= 64, (name, res)
    assert n == 2
