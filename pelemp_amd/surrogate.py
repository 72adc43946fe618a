"""Learned surrogate of the substep rates: training data, training on the
GPU and installation in a PELEMP_FORM_SURROGATE container.

Everything runs through the C ABI (include/pelemp_spray.h):
  pelemp_surrogate_sample   samples of the plain FP64 rate evaluation
  pelemp_mlp_train          Adam on the project's forward/backward kernels
  pelemp_mlp_forward        FP32 forward pass (validation)
  pelemp_mlp_infer_bf16     the bfloat16 MFMA inference of the update kernel
  pelemp_spray_set_surrogate

    python -m pelemp_amd.surrogate --mech 1 --samples 10000000 --epochs 24 \\
        --out pelemp_amd/surrogate_data/decane_dodecane_air.npz

The network (8 -> 4 x 64 hard swish -> 3) predicts the film diffusivity
ln(sum rho D_k), the heat-transfer group ln(lambda Nu*) and the film
viscosity ln(mu) from ln d, T_d, ln(|u_rel| + 0.01), T_gas, ln p, the film
fuel mass fraction and two liquid mass fractions; dm/dt, dT_d/dt and the
drag follow from them with the plain method's closed forms
(pelemp_amd/csrc/surrogate.h).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import _capi as C
from .fuels import spray_data
from .problem import PATM, mech_info

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_data")

# Training hull: the configuration's ranges (T_gas 600-1500 K, p 1-60 atm,
# d 1-100 um) plus the droplet-state ranges the workloads visit.
DEFAULT_HULL = dict(d_min=1.0e-6, d_max=1.0e-4, Tg_min=600.0, Tg_max=1500.0, p_min=1.0 * PATM, p_max=60.0 * PATM,
                    urel_max=30.0, Yfg_max=0.05, Td_min=280.0, Td_frac_max=0.98, BM_min=1.0e-6,
                    indicator_max=0.5)

FUELS_OF_MECH = {C.MECH_DODECANE_AIR: ("NC12H26",), C.MECH_DECANE_DODECANE_AIR: ("NC10H22", "NC12H26")}
OUTPUT_NAMES = ("ln(rho D)  [dm/dt]", "ln(lambda Nu*)  [heat, dT_d/dt]", "ln(mu_skin)  [drag, Re]")
SURR_U0 = 0.01  # the relative-velocity input is ln(|u_rel| + SURR_U0) (kSurrU0)


def input_ranges(hull: dict) -> list:
    """(lo, hi) of each raw input; the normalisation maps them to [-1, 1]."""
    return [(math.log(hull["d_min"]), math.log(hull["d_max"])), (hull["Td_min"], 700.0),
            (math.log(SURR_U0), math.log(hull["urel_max"] + SURR_U0)),
            (hull["Tg_min"], hull["Tg_max"]), (math.log(hull["p_min"]), math.log(hull["p_max"])), (0.0, 1.0),
            (0.0, 1.0), (0.0, 1.0)]


def make_norm(hull: dict, out_mean, out_std) -> C.SurrogateNormC:
    n = C.SurrogateNormC()
    for q, (lo, hi) in enumerate(input_ranges(hull)):
        n.in_center[q] = 0.5 * (lo + hi)
        n.in_scale[q] = 2.0 / (hi - lo)
    for q in range(C.SURR_NOUT):
        n.out_mean[q] = float(out_mean[q])
        n.out_std[q] = float(out_std[q])
    for k, v in hull.items():
        setattr(n, k, float(v))
    return n


def norm_dict(n: C.SurrogateNormC) -> dict:
    out = {}
    for name, typ in n._fields_:
        v = getattr(n, name)
        out[name] = [float(x) for x in v] if hasattr(v, "__len__") else float(v)
    return out


def norm_from_dict(d: dict) -> C.SurrogateNormC:
    n = C.SurrogateNormC()
    for name, _t in n._fields_:
        v = d[name]
        if isinstance(v, list):
            arr = getattr(n, name)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(n, name, v)
    return n


def layer_shapes():
    fan_in = [C.SURR_NIN, 64, 64, 64, 64]
    fan_out = [64, 64, 64, 64, C.SURR_NOUT]
    return list(zip(fan_out, fan_in))


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (N(0, 1/fan_in) weights, zero biases)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in layer_shapes():
        parts.append(rng.normal(0.0, 1.0 / math.sqrt(i), size=(o, i)).astype(np.float32).ravel())
        parts.append(np.zeros(o, dtype=np.float32))
    p = np.concatenate(parts)
    assert p.size == C.SURR_NPARAM
    return p


def mech_spray_data(mech: int) -> dict:
    return spray_data(FUELS_OF_MECH[mech], mech_info(mech)["species"])


class Samples:
    """Training samples resident on the device: x (n x 8 raw inputs),
    t (n x 3 targets), w (n usable flags)."""

    def __init__(self, mech: int, n: int, seed: int, hull: dict | None = None, stream=None):
        self.mech, self.n, self.seed = mech, int(n), int(seed)
        self.hull = dict(DEFAULT_HULL if hull is None else hull)
        self.x = C.DeviceBuffer(4 * self.n * C.SURR_NIN)
        self.t = C.DeviceBuffer(4 * self.n * C.SURR_NOUT)
        self.w = C.DeviceBuffer(4 * self.n)
        sd = C.spray_data_c(mech_spray_data(mech))
        hn = make_norm(self.hull, [0, 0, 0], [1, 1, 1])
        C.check(C.lib().pelemp_surrogate_sample(mech, ctypes.byref(sd), 1, ctypes.byref(hn), self.n, self.seed,
                                                self.x.ptr, self.t.ptr, self.w.ptr, stream),
                "pelemp_surrogate_sample")

    def host(self, m: int | None = None):
        m = self.n if m is None else min(m, self.n)
        x = self.x.download(np.empty((m, C.SURR_NIN), np.float32)) if m == self.n else \
            _download_prefix(self.x, (m, C.SURR_NIN))
        t = _download_prefix(self.t, (m, C.SURR_NOUT))
        w = _download_prefix(self.w, (m,))
        return x, t, w


def _download_prefix(buf: C.DeviceBuffer, shape) -> np.ndarray:
    a = np.empty(shape, np.float32)
    C.check(C.lib().pelemp_memcpy_dtoh(a.ctypes.data, buf.ptr, a.nbytes), "pelemp_memcpy_dtoh")
    return a


def forward(params: np.ndarray, norm: C.SurrogateNormC, x_dev: C.DeviceBuffer, n: int, quant: bool = False,
            stream=None) -> np.ndarray:
    """FP32 forward pass on the device (normalised outputs)."""
    pd = C.DeviceBuffer(4 * C.SURR_NPARAM)
    pd.upload(np.ascontiguousarray(params, np.float32))
    y = C.DeviceBuffer(4 * n * C.SURR_NOUT)
    C.check(C.lib().pelemp_mlp_forward(pd.ptr, ctypes.byref(norm), int(quant), n, x_dev.ptr, y.ptr, stream),
            "pelemp_mlp_forward")
    C.device_synchronize()
    return y.download(np.empty((n, C.SURR_NOUT), np.float32))


def infer_bf16(params: np.ndarray, norm: C.SurrogateNormC, x_dev: C.DeviceBuffer, n: int, stream=None) -> np.ndarray:
    """The bfloat16 MFMA inference (normalised outputs)."""
    p = np.ascontiguousarray(params, np.float32)
    y = C.DeviceBuffer(4 * n * C.SURR_NOUT)
    C.check(C.lib().pelemp_mlp_infer_bf16(p.ctypes.data, ctypes.byref(norm), n, x_dev.ptr, y.ptr, stream),
            "pelemp_mlp_infer_bf16")
    return y.download(np.empty((n, C.SURR_NOUT), np.float32))


def loss_grad(params: np.ndarray, norm: C.SurrogateNormC, x_dev, t_dev, w_dev, n: int, quant: bool = False,
              stream=None):
    pd = C.DeviceBuffer(4 * C.SURR_NPARAM)
    pd.upload(np.ascontiguousarray(params, np.float32))
    g = C.DeviceBuffer(4 * C.SURR_NPARAM)
    loss = ctypes.c_double()
    C.check(C.lib().pelemp_mlp_loss_grad(pd.ptr, ctypes.byref(norm), int(quant), n, x_dev.ptr, t_dev.ptr,
                                         None if w_dev is None else w_dev.ptr, g.ptr, ctypes.byref(loss), stream),
            "pelemp_mlp_loss_grad")
    return float(loss.value), g.download(np.empty(C.SURR_NPARAM, np.float32))


def loss_grad_noalloc(params_dev, norm: C.SurrogateNormC, x_dev, t_dev, w_dev, n: int, grad_dev,
                      quant: bool = False, stream=None) -> float:
    """pelemp_mlp_loss_grad_noalloc on caller-owned device buffers (params,
    inputs, targets, weights, gradient); returns the loss."""
    loss = ctypes.c_double()
    C.check(C.lib().pelemp_mlp_loss_grad_noalloc(params_dev.ptr, ctypes.byref(norm), int(quant), n, x_dev.ptr,
                                                 t_dev.ptr, None if w_dev is None else w_dev.ptr, grad_dev.ptr,
                                                 ctypes.byref(loss), stream),
            "pelemp_mlp_loss_grad_noalloc")
    return float(loss.value)


def train(params: np.ndarray, norm: C.SurrogateNormC, samples: Samples, *, epochs: int, batch: int = 4096,
          lr: float = 2e-3, lr_final: float = 2e-5, quant: bool = True, seed: int = 7, verbose: bool = False,
          stream=None):
    """Adam training on the device; returns (params, per-epoch loss)."""
    pd = C.DeviceBuffer(4 * C.SURR_NPARAM)
    pd.upload(np.ascontiguousarray(params, np.float32))
    cfg = C.MlpTrainCfgC(epochs=epochs, batch=batch, quant=int(quant), verbose=int(verbose), lr=lr, lr_final=lr_final,
                         beta1=0.9, beta2=0.999, eps=1e-8, seed=seed)
    hist = (ctypes.c_double * epochs)()
    C.check(C.lib().pelemp_mlp_train(pd.ptr, ctypes.byref(norm), samples.n, samples.x.ptr, samples.t.ptr,
                                     samples.w.ptr, ctypes.byref(cfg), hist, stream), "pelemp_mlp_train")
    return pd.download(np.empty(C.SURR_NPARAM, np.float32)), [float(h) for h in hist]


def output_errors(pred_n: np.ndarray, t: np.ndarray, w: np.ndarray, norm: C.SurrogateNormC) -> dict:
    """Per-output error of normalised predictions against targets: the
    outputs are logarithms, so the error is the relative error of the
    transfer group (and of dm/dt, the heat flux and the drag force)."""
    m = w > 0
    mean = np.array(norm.out_mean[:], np.float64)
    sd = np.array(norm.out_std[:], np.float64)
    z = mean + sd * pred_n[m].astype(np.float64)
    e = np.abs(np.expm1(z - t[m].astype(np.float64)))
    out = {}
    for q, name in enumerate(OUTPUT_NAMES):
        out[name] = dict(mean=float(e[:, q].mean()), p99=float(np.quantile(e[:, q], 0.99)),
                         max=float(e[:, q].max()))
    return out


class Surrogate:
    """A trained network: FP32 master parameters + normalisation."""

    def __init__(self, params: np.ndarray, norm: C.SurrogateNormC, mech: int, meta: dict | None = None):
        self.params = np.ascontiguousarray(params, np.float32)
        self.norm = norm
        self.mech = mech
        self.meta = dict(meta or {})

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params, norm=json.dumps(norm_dict(self.norm)), mech=self.mech,
                 meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path: str) -> "Surrogate":
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], norm_from_dict(json.loads(str(f["norm"]))), int(f["mech"]),
                       json.loads(str(f["meta"])))

    def install(self, pc) -> None:
        """Install in a SprayParticleContainer created with FORM_SURROGATE."""
        C.check(C.lib().pelemp_spray_set_surrogate(pc._h, self.params.ctypes.data, ctypes.byref(self.norm)),
                "pelemp_spray_set_surrogate")


def default_path(mech: int) -> str:
    return os.path.join(DATA_DIR, {C.MECH_DODECANE_AIR: "dodecane_air.npz",
                                   C.MECH_DECANE_DODECANE_AIR: "decane_dodecane_air.npz"}[mech])


def fit(mech: int, n_samples: int, epochs: int, *, batch: int = 4096, seed: int = 2024, lr: float = 2e-3,
        lr_final: float = 2e-5, n_val: int = 200_000, hull: dict | None = None, log=print) -> Surrogate:
    """Generate samples, train, validate on fresh samples; returns the network."""
    t0 = time.perf_counter()
    data = Samples(mech, n_samples, seed, hull)
    _, t, w = data.host(min(n_samples, 2_000_000))
    m = w > 0
    out_mean, out_std = t[m].mean(axis=0), t[m].std(axis=0)
    norm = make_norm(data.hull, out_mean, out_std)
    log(f"# samples: {n_samples} ({m.mean() * 100:.1f}% usable), {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    params, hist = train(init_params(seed), norm, data, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final,
                         verbose=True)
    log(f"# trained {epochs} epochs in {time.perf_counter() - t1:.1f} s; loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    val = Samples(mech, n_val, seed + 1, data.hull)
    vx, vt, vw = val.host()
    err_fp32 = output_errors(forward(params, norm, val.x, n_val), vt, vw, norm)
    err_bf16 = output_errors(infer_bf16(params, norm, val.x, n_val), vt, vw, norm)
    for name in OUTPUT_NAMES:
        log(f"#   {name}: mean {err_bf16[name]['mean']:.2e}  p99 {err_bf16[name]['p99']:.2e}  "
            f"max {err_bf16[name]['max']:.2e} (bf16 MFMA)")
    meta = dict(samples=n_samples, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed,
                loss_history=hist, val_samples=n_val, val_error_fp32=err_fp32, val_error_bf16=err_bf16,
                usable_fraction=float(m.mean()), train_seconds=time.perf_counter() - t1)
    return Surrogate(params, norm, mech, meta)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--mech", type=int, default=C.MECH_DECANE_DODECANE_AIR)
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    s = fit(a.mech, a.samples, a.epochs, batch=a.batch, seed=a.seed, lr=a.lr,
            log=lambda m: print(m, file=sys.stderr, flush=True))
    out = a.out or default_path(a.mech)
    s.save(out)
    print(json.dumps(dict(out=out, **{k: v for k, v in s.meta.items() if k != "loss_history"})))
    return 0


if __name__ == "__main__":
    sys.exit(main())
