"""Build the HIP library libpelemp_spray.so for gfx950 (in-tree).

    python -m pelemp_amd.build

Compiles the three translation units with hipcc and links them into
pelemp_amd/lib/libpelemp_spray.so.  The FP64 units use -ffp-contract=off
(bitwise agreement with the CPU oracle); the FP32 unit allows contraction,
reciprocal multiplication for division and the hardware approximations of
the elementary functions (its results are held to stated tolerances against
the FP64 oracle, not to bitwise agreement).  Packed-FP32 instructions are
disabled for every unit.  The compiler's per-kernel register, scratch,
occupancy and LDS figures are written to pelemp_amd/lib/kernel_resources.json
(read by tests/test_kernel_resources.py).
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(os.path.dirname(HERE), "include")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libpelemp_spray.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", INC, "-Wno-unused-result",
]
# The update kernels' units are compiled without machine-level loop-invariant
# code motion: hoisting the materialised FP64 constants and uniform values
# out of the droplet and substep loops kept them live in scalar registers,
# which then spilled into vector-register lanes (v_writelane/v_readlane,
# VALU instructions); without it the FP64 update kernel runs 2 % faster,
# the FP32 one 2 %, with the same results.
NO_MLICM = ["-mllvm", "-disable-machine-licm"]
UNITS = {
    "spray_host.hip": ["-ffp-contract=off"],
    "spray_fp64.hip": ["-ffp-contract=off", *NO_MLICM],
    "spray_fp32.hip": ["-ffp-contract=fast", "-freciprocal-math", "-fapprox-func", *NO_MLICM],
    "spray_surrogate.hip": ["-ffp-contract=off"],
}


RESOURCES = os.path.join(LIBDIR, "kernel_resources.json")
_REMARK = re.compile(r"remark:\s+(Function Name|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]|"
                     r"LDS Size \[bytes/block\]): (\S+)")


def _resources(stderr: str) -> dict:
    """Per-kernel register, scratch, occupancy and LDS figures from hipcc's
    kernel-resource-usage remarks (device functions included)."""
    out, cur = {}, None
    for key, val in _REMARK.findall(stderr):
        if key == "Function Name":
            cur = out.setdefault(val, {})
        elif cur is not None:
            cur[key.split(" ")[0].lower()] = int(val)
    return out


def _compile(unit: str, extra: list[str]):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJDIR, unit.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, *extra, "-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {unit}")
    return obj, _resources(r.stderr)


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        res = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    objs = [o for o, _ in res]
    resources = {}
    for _, r in res:
        resources.update(r)
    with open(RESOURCES, "w") as f:
        json.dump(resources, f, indent=1, sort_keys=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
