"""ctypes binding of libpelemp_spray.so (include/pelemp_spray.h).

The structures mirror the C ABI one to one.  Loading fails loudly when the
library has not been built: there is no CPU fallback for the droplet update.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

MAX_FUEL = 8

OK = 0
E_STATE_BOX = 1
E_SRC_BOX = 2
E_LEFT_DOMAIN = 3
E_BAD_CONFIG = 4
E_HIP = 5
E_ARG = 6
E_NOMEM = 7
E_BAD_CHILD = 8
E_BAD_INJECTION = 9
E_EB_STENCIL = 10
ERROR_NAMES = {
    E_STATE_BOX: "state box too small",
    E_SRC_BOX: "source box too small",
    E_LEFT_DOMAIN: "particle left the domain",
    E_BAD_CONFIG: "bad configuration",
    E_HIP: "HIP error",
    E_ARG: "invalid argument",
    E_NOMEM: "out of device memory",
    E_BAD_CHILD: "secondary droplet without a valid position",
    E_BAD_INJECTION: "bad injection",
    E_EB_STENCIL: "EB stencil without an acceptable cell",
}

DIST_UNIFORM = 0
DIST_NORMAL = 1
DIST_LOGNORMAL = 2
DIST_WEIBULL = 3
DIST_CHISQUARED = 4
DIST_TYPES = {"Uniform": DIST_UNIFORM, "Normal": DIST_NORMAL, "LogNormal": DIST_LOGNORMAL,
              "Weibull": DIST_WEIBULL, "ChiSquared": DIST_CHISQUARED}

MECH_DODECANE_AIR = 0
MECH_DECANE_DODECANE_AIR = 1
MECH_ALKANE6_AIR = 2
MECH_DECANE_AIR = 3

FORM_FP64 = 0
FORM_FP32 = 1
FORM_SURROGATE = 2

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpelemp_spray.so")

_c_dbl = ctypes.c_double
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64


class SprayDataC(ctypes.Structure):
    _fields_ = [
        ("mass_trans", _c_i32), ("mom_trans", _c_i32), ("fixed_parts", _c_i32),
        ("do_splash", _c_i32), ("do_breakup", _c_i32), ("nfuel", _c_i32),
        ("min_eb_vfrac", _c_dbl), ("ref_T", _c_dbl), ("sigma", _c_dbl), ("wall_T", _c_dbl),
        ("theta_c", _c_dbl), ("dtmod", _c_dbl), ("body_force", _c_dbl * 3),
        ("critT", _c_dbl * MAX_FUEL), ("boilT", _c_dbl * MAX_FUEL), ("cp", _c_dbl * MAX_FUEL),
        ("latent", _c_dbl * MAX_FUEL), ("ref_latent", _c_dbl * MAX_FUEL),
        ("psat_coef", _c_dbl * (4 * MAX_FUEL)), ("rho_coef", _c_dbl * (4 * MAX_FUEL)),
        ("lambda_coef", _c_dbl * (4 * MAX_FUEL)), ("mu_coef", _c_dbl * (4 * MAX_FUEL)),
        ("indx", _c_i32 * MAX_FUEL), ("dep_indx", _c_i32 * MAX_FUEL),
    ]


class SprayCompsC(ctypes.Structure):
    _fields_ = [(n, _c_i32) for n in (
        "rhoIndx", "momIndx", "engIndx", "utempIndx", "specIndx",
        "rhoSrcIndx", "momSrcIndx", "engSrcIndx", "specSrcIndx")]


class FabViewC(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("lo", _c_i32 * 3), ("hi", _c_i32 * 3), ("ncomp", _c_i32)]


class FabViewF32C(ctypes.Structure):
    """pelemp_fab_view_f32: the same FAB with float data."""
    _fields_ = [("data", ctypes.c_void_p), ("lo", _c_i32 * 3), ("hi", _c_i32 * 3), ("ncomp", _c_i32)]


class GeomC(ctypes.Structure):
    _fields_ = [
        ("plo", _c_dbl * 3), ("phi", _c_dbl * 3), ("dx", _c_dbl * 3),
        ("domain_lo", _c_i32 * 3), ("domain_hi", _c_i32 * 3),
        ("bndry_lo", _c_i32 * 3), ("bndry_hi", _c_i32 * 3),
    ]


SURR_NPARAM = 13251
SURR_NIN = 8
SURR_NOUT = 3
MLP_NOALLOC_MAX_BATCH = 4096  # PELEMP_MLP_NOALLOC_MAX_BATCH
SEGMENT_DROPLETS = 4096  # PELEMP_SEGMENT_DROPLETS
SOURCE_ACCUMULATE = 0
SOURCE_OVERWRITE = 1
_c_flt = ctypes.c_float


class SurrogateNormC(ctypes.Structure):
    _fields_ = [
        ("in_center", _c_flt * SURR_NIN), ("in_scale", _c_flt * SURR_NIN),
        ("out_mean", _c_flt * SURR_NOUT), ("out_std", _c_flt * SURR_NOUT),
        ("d_min", _c_flt), ("d_max", _c_flt), ("Tg_min", _c_flt), ("Tg_max", _c_flt),
        ("p_min", _c_flt), ("p_max", _c_flt), ("urel_max", _c_flt), ("Yfg_max", _c_flt),
        ("Td_min", _c_flt), ("Td_frac_max", _c_flt), ("BM_min", _c_flt), ("indicator_max", _c_flt),
    ]


class MlpTrainCfgC(ctypes.Structure):
    _fields_ = [
        ("epochs", _c_i32), ("batch", _c_i32), ("quant", _c_i32), ("verbose", _c_i32),
        ("lr", _c_dbl), ("lr_final", _c_dbl), ("beta1", _c_dbl), ("beta2", _c_dbl), ("eps", _c_dbl),
        ("seed", ctypes.c_uint64),
    ]


class SprayParamsC(ctypes.Structure):
    _fields_ = [("max_parcel_size", _c_dbl), ("breakup_parcel_factor", _c_dbl), ("khrt_B0", _c_dbl),
                ("khrt_B1", _c_dbl), ("khrt_C3", _c_dbl), ("rng_seed", ctypes.c_uint64),
                ("avg_inject_mass", _c_dbl)]


class EBViewC(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_void_p), ("vfrac", ctypes.c_void_p), ("ccent", ctypes.c_void_p),
                ("bcent", ctypes.c_void_p), ("bnorm", ctypes.c_void_p), ("lo", _c_i32 * 3), ("hi", _c_i32 * 3)]


class SprayJetC(ctypes.Structure):
    _fields_ = [("cent", _c_dbl * 3), ("norm", _c_dbl * 3), ("spread_angle", _c_dbl), ("swirl_angle", _c_dbl),
                ("hollow_spread", _c_dbl), ("jet_dia", _c_dbl), ("jet_vel", _c_dbl), ("max_jet_vel", _c_dbl),
                ("mass_flow", _c_dbl), ("jet_T", _c_dbl), ("jet_Y", _c_dbl * MAX_FUEL), ("start_time", _c_dbl),
                ("end_time", _c_dbl), ("hollow_spray", _c_i32), ("dist_type", _c_i32), ("dist_param", _c_dbl * 2),
                ("n_roi", _c_i32), ("reserved", _c_i32), ("roi_time", ctypes.c_void_p),
                ("roi_mdot", ctypes.c_void_p), ("roi_vel", ctypes.c_void_p), ("num_ppp", _c_dbl),
                ("sum_inj_mass", _c_dbl), ("sum_inj_time", _c_dbl), ("min_parcel", _c_dbl),
                ("total_inj_mass", _c_dbl), ("total_inj_time", _c_dbl), ("seed", ctypes.c_uint64),
                ("n_injections", _c_i64)]


_vp = ctypes.c_void_p
_PNORM = ctypes.POINTER(SurrogateNormC)
_PROTOS = {
    "pelemp_spray_set_params": (_c_i32, [_vp, ctypes.POINTER(SprayParamsC)]),
    "pelemp_spray_get_params": (_c_i32, [_vp, ctypes.POINTER(SprayParamsC)]),
    "pelemp_spray_create": (_c_i32, [ctypes.POINTER(_vp), _c_i32, _c_i32, ctypes.POINTER(SprayDataC),
                                     ctypes.POINTER(SprayCompsC), ctypes.POINTER(GeomC), _c_i64, _c_i32]),
    "pelemp_spray_create_pelec": (_c_i32, [ctypes.POINTER(_vp), _c_i32, _c_i32, ctypes.POINTER(SprayDataC),
                                           ctypes.POINTER(SprayCompsC), ctypes.POINTER(GeomC), _c_i64]),
    "pelemp_spray_destroy": (_c_i32, [_vp]),
    "pelemp_spray_get_data": (_c_i32, [_vp, ctypes.POINTER(SprayDataC)]),
    "pelemp_spray_set_particles": (_c_i32, [_vp, _c_i64, _vp, _vp, _vp]),
    "pelemp_spray_get_particles": (_c_i32, [_vp, ctypes.POINTER(_c_i64), _vp, _vp, _vp]),
    "pelemp_spray_num_particles": (_c_i32, [_vp, ctypes.POINTER(_c_i64)]),
    "pelemp_spray_add_particles": (_c_i32, [_vp, _c_i64, _vp, _vp, _vp]),
    "pelemp_spray_set_eb": (_c_i32, [_vp, ctypes.POINTER(EBViewC)]),
    "pelemp_spray_jet_check": (_c_i32, [_vp, ctypes.POINTER(SprayJetC), ctypes.POINTER(_c_dbl)]),
    "pelemp_spray_read_roi": (_c_i32, [ctypes.c_char_p, _c_dbl, _c_dbl, _c_dbl, _c_dbl, ctypes.POINTER(_c_i32),
                                       _vp, _vp, _vp, ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl),
                                       ctypes.POINTER(_c_dbl)]),
    "pelemp_spray_inject": (_c_i32, [_vp, ctypes.POINTER(SprayJetC), _c_dbl, _c_dbl, ctypes.POINTER(_c_i64), _vp]),
    "pelemp_spray_uniform_init": (_c_i32, [_vp, _vp, _vp, _c_dbl, _c_dbl, _vp, _c_dbl]),
    "pelemp_spray_write_injection_log": (_c_i32, [ctypes.c_char_p, _c_i32, _vp, ctypes.POINTER(SprayJetC)]),
    "pelemp_spray_read_injection_log": (_c_i32, [ctypes.c_char_p, _c_i32, _vp, ctypes.POINTER(SprayJetC)]),
    "pelemp_spray_write_ascii": (_c_i32, [_vp, ctypes.c_char_p]),
    "pelemp_spray_checkpoint": (_c_i32, [_vp, ctypes.c_char_p, _c_i32, _vp]),
    "pelemp_spray_restart": (_c_i32, [_vp, ctypes.c_char_p]),
    "pelemp_particles_write_checkpoint": (_c_i32, [ctypes.c_char_p, ctypes.c_char_p, _c_i64, _vp, _vp, _vp, _c_i32,
                                                  _vp, _c_i64, _c_i32, _vp, _vp, _vp, _vp]),
    "pelemp_particles_checkpoint_info": (_c_i32, [ctypes.c_char_p, ctypes.c_char_p, _c_i32, _vp, _vp, _vp]),
    "pelemp_particles_read_checkpoint": (_c_i32, [ctypes.c_char_p, ctypes.c_char_p, _c_i32, _c_i64, _vp, _vp, _vp,
                                                 _vp]),
    "pelemp_spray_update": (_c_i32, [_vp, ctypes.POINTER(FabViewC), ctypes.POINTER(FabViewC), _c_dbl, _c_dbl,
                                     _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_dbl, _vp]),
    "pelemp_spray_update_f32gas": (_c_i32, [_vp, ctypes.POINTER(FabViewF32C), ctypes.POINTER(FabViewC), _c_dbl,
                                            _c_dbl, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_dbl, _vp]),
    "pelemp_spray_sync": (_c_i32, [_vp, _vp]),
    "pelemp_spray_set_source_mode": (_c_i32, [_vp, _c_i32]),
    "pelemp_spray_move_kick_drift": (_c_i32, [_vp, ctypes.POINTER(FabViewC), ctypes.POINTER(FabViewC), _c_dbl,
                                              _c_dbl, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_dbl, _vp]),
    "pelemp_spray_est_timestep": (_c_i32, [_vp, _c_dbl, _c_dbl, ctypes.POINTER(_c_dbl), _vp]),
    "pelemp_spray_derive_count": (_c_i32, [_vp, _c_i32, ctypes.POINTER(_c_i32)]),
    "pelemp_spray_compute_derived": (_c_i32, [_vp, ctypes.POINTER(FabViewC), _c_i32, _c_i32, _vp]),
    "pelemp_spray_transfer_source": (_c_i32, [_vp, _c_i32, ctypes.POINTER(FabViewC), ctypes.POINTER(FabViewC), _vp]),
    "pelemp_spray_pass_counts": (_c_i32, [_vp, ctypes.POINTER(_c_i64), _c_i32, ctypes.POINTER(_c_i32)]),
    "pelemp_spray_set_timing": (_c_i32, [_vp, _c_i32]),
    "pelemp_spray_get_timing": (_c_i32, [_vp, ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_i64), _c_i32]),
    "pelemp_spray_device_bytes": (_c_i32, [_vp, ctypes.POINTER(_c_i64)]),
    "pelemp_surrogate_sample": (_c_i32, [_c_i32, ctypes.POINTER(SprayDataC), _c_i32, _PNORM, _c_i64,
                                         ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "pelemp_mlp_forward": (_c_i32, [_vp, _PNORM, _c_i32, _c_i64, _vp, _vp, _vp]),
    "pelemp_mlp_loss_grad": (_c_i32, [_vp, _PNORM, _c_i32, _c_i64, _vp, _vp, _vp, _vp,
                                      ctypes.POINTER(_c_dbl), _vp]),
    "pelemp_mlp_loss_grad_noalloc": (_c_i32, [_vp, _PNORM, _c_i32, _c_i64, _vp, _vp, _vp, _vp,
                                              ctypes.POINTER(_c_dbl), _vp]),
    "pelemp_mlp_train": (_c_i32, [_vp, _PNORM, _c_i64, _vp, _vp, _vp, ctypes.POINTER(MlpTrainCfgC), _vp, _vp]),
    "pelemp_mlp_infer_bf16": (_c_i32, [_vp, _PNORM, _c_i64, _vp, _vp, _vp]),
    "pelemp_spray_set_surrogate": (_c_i32, [_vp, _vp, _PNORM]),
    "pelemp_spray_surrogate_counts": (_c_i32, [_vp, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    "pelemp_mfma_probe": (_c_i32, [_vp, _vp, _vp]),
    "pelemp_math_eval": (_c_i32, [_c_i32, _c_i64, _vp, _vp, _vp]),
    "pelemp_device_count": (_c_i32, [ctypes.POINTER(_c_i32)]),
    "pelemp_device_malloc": (_c_i32, [ctypes.POINTER(_vp), _c_i64]),
    "pelemp_device_free": (_c_i32, [_vp]),
    "pelemp_memcpy_htod": (_c_i32, [_vp, _vp, _c_i64]),
    "pelemp_memcpy_dtoh": (_c_i32, [_vp, _vp, _c_i64]),
    "pelemp_memcpy_dtod": (_c_i32, [_vp, _vp, _c_i64]),
    "pelemp_memset_device": (_c_i32, [_vp, _c_i32, _c_i64]),
    "pelemp_fab_set_zero": (_c_i32, [_vp, _c_i64, _vp]),
    "pelemp_device_synchronize": (_c_i32, []),
    "pelemp_stream_create": (_c_i32, [ctypes.POINTER(_vp)]),
    "pelemp_stream_destroy": (_c_i32, [_vp]),
    "pelemp_stream_synchronize": (_c_i32, [_vp]),
    "pelemp_timer_create": (_c_i32, [ctypes.POINTER(_vp)]),
    "pelemp_timer_record": (_c_i32, [_vp, _c_i32, _vp]),
    "pelemp_timer_elapsed": (_c_i32, [_vp, ctypes.POINTER(_c_dbl)]),
    "pelemp_timer_destroy": (_c_i32, [_vp]),
    "pelemp_last_error": (ctypes.c_char_p, []),
}

_lib = None


class PelempError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)} ({msg})")


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m pelemp_amd.build` "
                "(the spray update has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_PROTOS)


def check(code: int, where: str) -> None:
    if code != OK:
        msg = lib().pelemp_last_error()
        raise PelempError(code, where, msg.decode() if msg else "")


def fill_array(dst, values):
    for i, v in enumerate(values):
        dst[i] = v


def spray_data_c(d: dict) -> SprayDataC:
    """Build a SprayDataC from a dict with SprayData field names."""
    s = SprayDataC()
    for k in ("mass_trans", "mom_trans", "fixed_parts", "do_splash", "do_breakup", "nfuel"):
        setattr(s, k, int(d.get(k, 0)))
    for k in ("min_eb_vfrac", "ref_T", "sigma", "wall_T", "theta_c", "dtmod"):
        setattr(s, k, float(d.get(k, {"min_eb_vfrac": 0.05, "sigma": -1., "wall_T": -1., "theta_c": -1.,
                                       "dtmod": 0.5}.get(k, 0.))))
    fill_array(s.body_force, d.get("body_force", [0., 0., 0.]))
    for k in ("critT", "boilT", "cp", "latent", "ref_latent", "psat_coef", "rho_coef", "lambda_coef",
              "mu_coef"):
        fill_array(getattr(s, k), d.get(k, []))
    fill_array(s.indx, d.get("indx", []))
    fill_array(s.dep_indx, d.get("dep_indx", d.get("indx", [])))
    return s


def spray_data_dict(s: SprayDataC) -> dict:
    out = {}
    for name, _t in SprayDataC._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def jet_c(jet_cent, jet_norm, spread_angle: float, jet_dia: float, jet_vel: float, mass_flow: float,
          jet_temp: float, jet_Y=None, dist_type: str = "Uniform", dist_params=(0.0, 0.0), start_time: float = 0.0,
          end_time: float = 1.0e8, phi_swirl: float = 0.0, hollow_spray: bool = False, hollow_spread: float = 0.0,
          inject_ppp: float = -1.0, seed: int = 0) -> SprayJetC:
    """A SprayJetC as the reference's direct constructor fills SprayJet
    (SprayJet.cpp:95-141): angles given in degrees, stored in radians; the
    normal as given (pelemp_spray_jet_check normalises it)."""
    if dist_type not in DIST_TYPES:
        raise ValueError(f"unknown distribution {dist_type!r}")
    j = SprayJetC()
    for d in range(3):
        j.cent[d] = float(jet_cent[d])
        j.norm[d] = float(jet_norm[d])
    j.spread_angle = float(spread_angle) * np.pi / 180.0
    j.swirl_angle = float(phi_swirl) * np.pi / 180.0
    j.hollow_spread = float(hollow_spread) * np.pi / 180.0
    j.jet_dia = float(jet_dia)
    j.jet_vel = float(jet_vel)
    j.max_jet_vel = float(jet_vel)
    j.mass_flow = float(mass_flow)
    j.jet_T = float(jet_temp)
    for i, y in enumerate([1.0] if jet_Y is None else list(jet_Y)):
        j.jet_Y[i] = float(y)
    j.start_time = float(start_time)
    j.end_time = float(end_time)
    j.hollow_spray = 1 if hollow_spray else 0
    j.dist_type = DIST_TYPES[dist_type]
    j.dist_param[0] = float(dist_params[0])
    j.dist_param[1] = float(dist_params[1]) if len(dist_params) > 1 else 0.0
    j.num_ppp = float(inject_ppp)
    j.sum_inj_mass = -1.0
    j.sum_inj_time = -1.0
    j.min_parcel = 1.0
    j.seed = int(seed)
    return j


def fab_view(ptr: int, lo, hi, ncomp: int, f32: bool = False):
    v = FabViewF32C() if f32 else FabViewC()
    v.data = ptr
    fill_array(v.lo, lo)
    fill_array(v.hi, hi)
    v.ncomp = ncomp
    return v


def geom_c(g: dict) -> GeomC:
    c = GeomC()
    for k in ("plo", "phi", "dx", "domain_lo", "domain_hi", "bndry_lo", "bndry_hi"):
        fill_array(getattr(c, k), g[k])
    return c


def comps_c(d: dict) -> SprayCompsC:
    c = SprayCompsC()
    for name, _t in SprayCompsC._fields_:
        setattr(c, name, int(d[name]))
    return c


def math_eval(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    """Evaluate the library's device FP64 elementary function `fn` on x (, y)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    check(lib().pelemp_math_eval(fn, x.size, x.ctypes.data, None if yy is None else yy.ctypes.data,
                                 out.ctypes.data), "pelemp_math_eval")
    return out


class DeviceBuffer:
    """Device allocation owned by the library's HIP runtime."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        check(lib().pelemp_device_malloc(ctypes.byref(p), int(nbytes)), "pelemp_device_malloc")
        self.ptr = int(p.value)
        self.nbytes = int(nbytes)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().pelemp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().pelemp_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "pelemp_memcpy_htod")

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        check(lib().pelemp_memcpy_dtoh(a.ctypes.data, self.ptr, a.nbytes), "pelemp_memcpy_dtoh")
        return a

    def zero(self) -> None:
        check(lib().pelemp_memset_device(self.ptr, 0, self.nbytes), "pelemp_memset_device")


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().pelemp_stream_create(ctypes.byref(s)), "pelemp_stream_create")
        self.handle = int(s.value)

    def synchronize(self) -> None:
        check(lib().pelemp_stream_synchronize(self.handle), "pelemp_stream_synchronize")

    def __del__(self):
        try:
            lib().pelemp_stream_destroy(self.handle)
        except Exception:
            pass


class Timer:
    """HIP event pair recorded on a stream."""

    def __init__(self):
        t = ctypes.c_void_p()
        check(lib().pelemp_timer_create(ctypes.byref(t)), "pelemp_timer_create")
        self.handle = int(t.value)

    def start(self, stream=None) -> None:
        check(lib().pelemp_timer_record(self.handle, 0, stream), "pelemp_timer_record")

    def stop(self, stream=None) -> None:
        check(lib().pelemp_timer_record(self.handle, 1, stream), "pelemp_timer_record")

    def elapsed_ms(self) -> float:
        ms = ctypes.c_double()
        check(lib().pelemp_timer_elapsed(self.handle, ctypes.byref(ms)), "pelemp_timer_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            lib().pelemp_timer_destroy(self.handle)
        except Exception:
            pass


def device_count() -> int:
    n = ctypes.c_int32()
    check(lib().pelemp_device_count(ctypes.byref(n)), "pelemp_device_count")
    return int(n.value)


def device_synchronize() -> None:
    check(lib().pelemp_device_synchronize(), "pelemp_device_synchronize")


def write_particle_checkpoint(dir: str, name: str, pos, rdata, ids, real_names, next_id: int, boxes=None,
                              plo=None, dx=None) -> None:
    """AMReX particle checkpoint of host arrays (pelemp_particles_write_checkpoint).
    boxes: list of ((lo), (hi)) cell-index boxes, one box by default (then plo
    and dx are not needed)."""
    import numpy as np

    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    nreal = len(real_names)
    rdata = np.ascontiguousarray(rdata, dtype=np.float64).reshape(n, nreal)
    ids = np.ascontiguousarray(ids, dtype=np.int64).reshape(n)
    if boxes is None:
        boxes = [((0, 0, 0), (0, 0, 0))]
    lo = np.ascontiguousarray([b[0] for b in boxes], dtype=np.int32)
    hi = np.ascontiguousarray([b[1] for b in boxes], dtype=np.int32)
    nm = (ctypes.c_char_p * max(nreal, 1))(*[s.encode() for s in real_names])
    p3 = None if plo is None else (ctypes.c_double * 3)(*plo)
    d3 = None if dx is None else (ctypes.c_double * 3)(*dx)
    check(lib().pelemp_particles_write_checkpoint(
        dir.encode(), name.encode(), n, pos.ctypes.data, rdata.ctypes.data, ids.ctypes.data, nreal,
        ctypes.cast(nm, ctypes.c_void_p), int(next_id), len(boxes), lo.ctypes.data, hi.ctypes.data,
        ctypes.cast(p3, ctypes.c_void_p) if p3 is not None else None,
        ctypes.cast(d3, ctypes.c_void_p) if d3 is not None else None), "write_particle_checkpoint")


def read_particle_checkpoint(dir: str, name: str, level: int = 0):
    """(pos [n,3], rdata [n,nreal], ids [n], next_id) of one level of a
    particle checkpoint (pelemp_particles_read_checkpoint)."""
    import numpy as np

    n = ctypes.c_int64()
    nr = ctypes.c_int32()
    nxt = ctypes.c_int64()
    check(lib().pelemp_particles_checkpoint_info(dir.encode(), name.encode(), level, ctypes.byref(n),
                                                 ctypes.byref(nr), ctypes.byref(nxt)), "particle checkpoint info")
    cap = int(n.value)
    pos = np.zeros((max(cap, 1), 3))
    rd = np.zeros((max(cap, 1), max(int(nr.value), 1)))
    ids = np.zeros(max(cap, 1), dtype=np.int64)
    m = ctypes.c_int64()
    check(lib().pelemp_particles_read_checkpoint(dir.encode(), name.encode(), level, cap, ctypes.byref(m),
                                                 pos.ctypes.data, rd.ctypes.data, ids.ctypes.data),
          "read_particle_checkpoint")
    m = int(m.value)
    return pos[:m], rd[:m, :int(nr.value)], ids[:m], int(nxt.value)
