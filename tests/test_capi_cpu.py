"""The C-ABI library loads on a host without a GPU and exports every entry
point declared in include/pelemp_spray.h (no compute calls are made)."""
import ctypes
import os
import re

from pelemp_amd import _capi as C

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pelemp_spray.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pelemp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 25
    lib = ctypes.CDLL(C.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) == set(C.exported_symbols())


def test_struct_layouts_match_header():
    # sizes of the PODs as the C compiler lays them out (8-byte alignment of doubles)
    assert ctypes.sizeof(C.SprayCompsC) == 9 * 4
    assert ctypes.sizeof(C.FabViewC) == 8 + 7 * 4 + 4  # pointer, lo[3], hi[3], ncomp, padding
    assert ctypes.sizeof(C.GeomC) == 9 * 8 + 12 * 4
    n_dbl = 6 + 3 + 5 * 8 + 4 * 32
    assert ctypes.sizeof(C.SprayDataC) == 6 * 4 + n_dbl * 8 + 2 * 8 * 4


def test_device_count_without_gpu_is_clean():
    # on a host without a GPU this returns 0 devices rather than failing
    assert C.device_count() >= 0
