"""GPU tests of the learned surrogate: MFMA operand layout, the FP32
forward pass and the gradients against the CPU reference network
(oracle/mlp_ref.py), the bfloat16 MFMA inference, the training samples
against the CPU oracle's droplet source, training, and the surrogate form of
the update against the FP64 oracle."""
import ctypes
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def _lib():
    from pelemp_amd import _capi as C

    if C.device_count() < 1:
        pytest.skip("no HIP device")
    return C


def _norm(C, seed=0):
    from pelemp_amd import surrogate as S

    rng = np.random.default_rng(seed)
    return S.make_norm(S.DEFAULT_HULL, rng.normal(0, 1, 3), rng.uniform(0.5, 2.0, 3))


def _raw_inputs(n, seed=1):
    """Raw inputs spread over the normalisation box."""
    from pelemp_amd import surrogate as S

    rng = np.random.default_rng(seed)
    rr = S.input_ranges(S.DEFAULT_HULL)
    return np.stack([rng.uniform(lo, hi, n) for lo, hi in rr], axis=1).astype(np.float32)


def _params(seed=2):
    from pelemp_amd import surrogate as S

    p = S.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    return (p + rng.normal(0, 0.05, p.size)).astype(np.float32)  # non-zero biases too


def _bf16_bits(a):
    from oracle import mlp_ref as R

    return (R.bf16_round(a).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_mfma_probe_layout():
    C = _lib()
    rng = np.random.default_rng(0)
    A = rng.normal(size=(32, 16)).astype(np.float32)
    B = rng.normal(size=(16, 32)).astype(np.float32)
    ab, bb = _bf16_bits(A), _bf16_bits(B)
    D = np.empty((32, 32), np.float32)
    C.check(C.lib().pelemp_mfma_probe(ab.ctypes.data, bb.ctypes.data, D.ctypes.data), "probe")
    Aq = (ab.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    Bq = (bb.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    ref = Aq @ Bq
    assert np.max(np.abs(D - ref)) < 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("quant", [False, True])
def test_fp32_forward_matches_cpu_network(quant):
    C = _lib()
    from oracle import mlp_ref as R
    from pelemp_amd import surrogate as S

    n = 3000
    x = _raw_inputs(n)
    norm = _norm(C)
    p = _params()
    xd = C.DeviceBuffer(x.nbytes)
    xd.upload(x)
    y = S.forward(p, norm, xd, n, quant=quant)
    xn = R.normalise_inputs(x, norm.in_center[:], norm.in_scale[:])
    yr, _ = R.forward(p.astype(np.float64), xn, quant=quant)
    err = np.abs(y - yr)
    # quantised: an FP32 rounding difference can move a bfloat16 activation by one step
    tol = 2e-3 if quant else 2e-5
    assert np.quantile(err, 0.999) < tol * max(1.0, np.abs(yr).max()), err.max()


@pytest.mark.parametrize("quant", [False, True])
def test_gradients_match_cpu_network(quant):
    C = _lib()
    from oracle import mlp_ref as R
    from pelemp_amd import surrogate as S

    n = 1500
    x = _raw_inputs(n, seed=4)
    rng = np.random.default_rng(5)
    t = rng.normal(0, 1, (n, 3)).astype(np.float32)
    w = (rng.uniform(size=n) > 0.1).astype(np.float32)
    norm = _norm(C, seed=6)
    p = _params(seed=7)
    bufs = []
    for a in (x, t, w):
        b = C.DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    loss, g = S.loss_grad(p, norm, bufs[0], bufs[1], bufs[2], n, quant=quant)
    xn = R.normalise_inputs(x, norm.in_center[:], norm.in_scale[:])
    lr, gr = R.loss_grad(p.astype(np.float64), xn, t, w, np.array(norm.out_mean[:]), np.array(norm.out_std[:]),
                         quant=quant)
    assert loss == pytest.approx(lr, rel=1e-4)
    # per-layer relative error of the gradient
    o = 0
    for i, m in zip(R.FAN_IN, R.FAN_OUT):
        k = i * m + m
        gg, rr = g[o:o + k], gr[o:o + k]
        rel = np.linalg.norm(gg - rr) / np.linalg.norm(rr)
        assert rel < (5e-3 if quant else 1e-4), (o, rel)
        o += k


def test_bf16_mfma_inference_matches_quantised_forward():
    C = _lib()
    from oracle import mlp_ref as R
    from pelemp_amd import surrogate as S

    n = 5000  # not a multiple of 64: the tail wavefront has idle lanes
    x = _raw_inputs(n, seed=8)
    norm = _norm(C, seed=9)
    p = _params(seed=10)
    xd = C.DeviceBuffer(x.nbytes)
    xd.upload(x)
    yb = S.infer_bf16(p, norm, xd, n)
    xn = R.normalise_inputs(x, norm.in_center[:], norm.in_scale[:])
    yq, _ = R.forward(p.astype(np.float64), xn, quant=True)
    y0, _ = R.forward(p.astype(np.float64), xn, quant=False)
    err = np.abs(yb - yq)
    scale = max(1.0, np.abs(yq).max())
    assert np.median(err) < 1e-4 * scale
    assert np.quantile(err, 0.999) < 3e-3 * scale
    # and the quantised network is close to the FP32 one
    assert np.quantile(np.abs(yb - y0), 0.99) < 0.05 * scale


# ---- training samples against the CPU oracle ------------------------------
def _splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _u(seed, i, j):
    r = _splitmix64(seed ^ _splitmix64((i * 16 + j) & MASK))
    return ((r >> 11) + 0.5) / 9007199254740992.0


def _f32(v):
    return float(np.float32(v))


def _sample_point(seed, i, hull, W, sd, fuel_idx, O):
    """The sampler's inputs of sample i, recomputed in FP64 on the host."""
    H = {k: _f32(v) for k, v in hull.items()}
    u = [_u(seed, i, j) for j in range(7)]
    dia = math.exp(math.log(H["d_min"]) + u[0] * (math.log(H["d_max"]) - math.log(H["d_min"])))
    Tg = H["Tg_min"] + u[1] * (H["Tg_max"] - H["Tg_min"])
    p = math.exp(math.log(H["p_min"]) + u[2] * (math.log(H["p_max"]) - math.log(H["p_min"])))
    urel = H["urel_max"] * u[3] * u[3]
    Yfg = H["Yfg_max"] * u[4] * u[4] * u[4]
    nf = len(fuel_idx)
    Yl = [1.0] if nf == 1 else [u[5], 1.0 - u[5]]
    Y = np.zeros(len(W))
    Y[0], Y[1] = 0.233 * (1 - Yfg), 0.767 * (1 - Yfg)
    for k, ix in enumerate(fuel_idx):
        Y[ix] = Yfg * Yl[k]
    s = float(np.sum(Y / W))
    rho = p / (8.31446261815324 * s * Tg)
    cB = [O.lib().oracle_calc_boil_t(ctypes.byref(sd), p, W[ix], k) for k, ix in enumerate(fuel_idx)]
    Tb = sum(Yl[k] * cB[k] for k in range(nf))
    Td = H["Td_min"] + u[6] * (H["Td_frac_max"] * Tb - H["Td_min"])
    return dia, Tg, p, urel, Y, rho, Yl, cB, Td


def _skin_state(sd, O, W, fuel_idx, Y, p, Td, Yl, cB):
    """calcVaporState (Drag.H:47-135): B_M, surface vapour mole fractions,
    film mass fractions (1/3 rule, other species renormalised) and the film
    molecular weight."""
    mw_mix = 1.0 / float(np.sum(Y / W))
    mw_part = 1.0 / sum(Yl[k] / W[ix] for k, ix in enumerate(fuel_idx))
    Xv = []
    for k, ix in enumerate(fuel_idx):
        Tp = min(Td, 0.99 * cB[k])
        ps = O.lib().oracle_psat(ctypes.byref(sd), Tp, k)
        Xv.append(Yl[k] * mw_part / W[ix] * ps / p if ps > Y[ix] * mw_mix / W[ix] * p else 0.0)
    tot = sum(x * W[ix] for x, ix in zip(Xv, fuel_idx)) + (1 - sum(Xv)) * mw_mix
    Ysk = np.zeros(len(W))
    yv = yf = 0.0
    for x, ix in zip(Xv, fuel_idx):
        if x > 0:
            yfv = x * W[ix] / tot
            yv += yfv
            yf += Y[ix]
            Ysk[ix] = yfv + (Y[ix] - yfv) / 3.0
    renorm = (1 - Ysk.sum()) / (1 - yf)
    Ysk = np.where(Ysk == 0, Y * renorm, Ysk)
    BM = min(20.0, (yv - yf) / max(1e-15, 1 - yv))
    return BM, Xv, Ysk, 1.0 / float(np.sum(Ysk / W))


def test_training_samples_match_oracle_droplet_source():
    C = _lib()
    from oracle import oracle_py as O
    from pelemp_amd import surrogate as S
    from pelemp_amd.problem import mech_info

    mech = C.MECH_DECANE_DODECANE_AIR
    n, seed = 256, 99
    smp = S.Samples(mech, n, seed)
    x, t, w = smp.host()
    assert w.mean() > 0.5
    info = mech_info(mech)
    W = np.array(info["W"], dtype=np.float64)
    sdata = O.latent_shifted(S.mech_spray_data(mech), mech, len(W))
    sd = C.spray_data_c(sdata)
    fuel_idx = sdata["indx"]
    checked = 0
    for i in range(n):
        if w[i] == 0:
            continue
        dia, Tg, p, urel, Y, rho, Yl, cB, Td = _sample_point(seed, i, S.DEFAULT_HULL, W, sd, fuel_idx, O)
        assert x[i, 0] == pytest.approx(math.log(dia), rel=1e-6)
        assert x[i, 1] == pytest.approx(Td, rel=1e-6)
        nf = len(fuel_idx)
        rd = np.zeros(10 + nf)
        rd[3], rd[4] = Td, dia
        rd[5:5 + nf] = Yl
        rd[5 + nf] = 1.0
        gas = np.concatenate([[urel, 0.0, 0.0, Tg, rho], Y])
        dt = 1e-9
        _, rout, pid, src, _ = O.spray_source_one(mech, sdata, dt, gas, rd)
        if pid <= 0:
            continue
        BM, Xv, Ysk, mw_skin = _skin_state(sd, O, W, fuel_idx, Y, p, Td, Yl, cB)
        T_skin = Td + (Tg - Td) / 3.0
        rho_skin = mw_skin * p / (8.31446261815324 * T_skin)
        tr = O.gas_props(mech, T_skin, rho_skin, Ysk)
        # film properties: the targets are the oracle's transport at the film state
        mu = tr["mu"]
        rhoD = sum(tr["rhoD"][ix] * mw_skin / W[ix] * x / sum(Xv) for x, ix in zip(Xv, fuel_idx))
        assert math.log(rhoD) == pytest.approx(float(t[i, 0]), abs=2e-5)
        assert math.log(mu) == pytest.approx(float(t[i, 2]), abs=2e-5)
        # heat: eng = Q + sum_k mdot_k h_k(T_d), Q = pi d (T_g - T_d) lambda Nu*
        h = O.gas_props(mech, Td, 1.0, np.full(len(W), 1.0 / len(W)))["h"]
        Q = src[4 + nf] - sum(src[4 + k] * h[ix] for k, ix in enumerate(fuel_idx))
        lamnu = Q / (math.pi * dia * (Tg - Td))
        assert math.log(lamnu) == pytest.approx(float(t[i, 1]), abs=2e-5)
        # the closed forms the surrogate keeps reproduce the oracle's source
        Re = rho_skin * urel * dia / mu
        CD = 0.424 if Re > 1000 else (24 / Re * (1 + (Re * Re) ** (1 / 3) / 6) if Re > 1 else 24 / Re)
        F = 0.125 * rho_skin * CD * math.pi * dia * dia * urel
        rinv = sum(Yl[k] / O.lib().oracle_rhoL(ctypes.byref(sd), min(Td, cB[k]), k) for k in range(nf))
        m0 = math.pi / 6 / rinv * dia ** 3
        mdot = src[3]
        assert src[0] == pytest.approx(F * urel * (1 + dt * mdot / m0), rel=1e-6)
        F2 = max(min(400.0, Re) ** 0.077, 1.0)
        Sh0 = 1 + F2 * (1 + Re * mu / rhoD) ** (1 / 3)
        Sh = 2 + (Sh0 - 2) * BM / (math.log1p(BM) * (1 + BM) ** 0.7)
        assert mdot == pytest.approx(-math.pi * dia * Sh * rhoD * math.log1p(BM), rel=1e-6)
        checked += 1
    assert checked > 100


def test_training_reduces_loss():
    C = _lib()
    from pelemp_amd import surrogate as S

    mech = C.MECH_DECANE_DODECANE_AIR
    smp = S.Samples(mech, 1 << 17, 5)
    _, t, w = smp.host()
    m = w > 0
    norm = S.make_norm(smp.hull, t[m].mean(0), t[m].std(0))
    p, hist = S.train(S.init_params(1), norm, smp, epochs=4, batch=2048, lr=3e-3, lr_final=3e-4)
    assert hist[-1] < 0.2 * hist[0], hist
    assert np.all(np.isfinite(p))


# ---- the surrogate form of the update ---------------------------------------
def _trained(mech_name="MECH_DECANE_DODECANE_AIR"):
    from pelemp_amd import _capi as C
    from pelemp_amd import surrogate as S

    path = S.default_path(getattr(C, mech_name))
    assert os.path.exists(path), f"trained network missing: {path}"
    return S.Surrogate.load(path)


@pytest.mark.parametrize("mech_name", ["MECH_DECANE_DODECANE_AIR", "MECH_DODECANE_AIR"])
def test_trained_network_error_per_output(mech_name):
    C = _lib()
    from pelemp_amd import surrogate as S

    net = _trained(mech_name)
    val = S.Samples(net.mech, 100_000, 777)
    _, t, w = val.host()
    err = S.output_errors(S.infer_bf16(net.params, net.norm, val.x, val.n), t, w, net.norm)
    for name in S.OUTPUT_NAMES:
        assert err[name]["mean"] < 5e-3, (name, err[name])
        assert err[name]["p99"] < 3e-2, (name, err[name])


def test_surrogate_form_evaporated_mass_vs_fp64_oracle():
    """Config-2 recipe on a 16^3 box, 10 steps: the surrogate form's
    evaporated mass (sum of the mass source) is within 0.5% of the FP64
    oracle's; most droplet-steps use the network."""
    C = _lib()
    from pelemp_amd import problem as P
    from pelemp_amd.spray import DeviceFab, SprayParticleContainer

    from tests.helpers import run_oracle

    prob = P.config2(ncell=16)
    nsteps = 10
    _, _, _, srcs_o = run_oracle(prob, nsteps)
    net = _trained()
    pc = SprayParticleContainer(prob.mech, prob.sdata, prob.comps, prob.geom, prob.n + 1024, form=C.FORM_SURROGATE)
    net.install(pc)
    pc.set_particles(prob.pos, prob.rdata, prob.ids)
    state = DeviceFab(prob.state_lo, prob.state_hi, prob.state.shape[0], prob.state)
    evap_s = 0.0
    evap_o = 0.0
    V = float(np.prod(prob.geom["dx"]))
    for k in range(nsteps):
        src = DeviceFab(prob.src_lo, prob.src_hi, prob.src_shape[0])
        pc.updateParticles(0, state, src, prob.flow_dt, 0.0, prob.state_ghosts, prob.source_ghosts, False, False,
                           True)
        evap_s += float(src.numpy()[0].sum()) * V * prob.flow_dt
        evap_o += float(srcs_o[k][0].sum()) * V * prob.flow_dt
    ns, nfb = ctypes.c_int64(), ctypes.c_int64()
    C.check(C.lib().pelemp_spray_surrogate_counts(pc._h, ctypes.byref(ns), ctypes.byref(nfb)), "counts")
    frac = ns.value / max(1, ns.value + nfb.value)
    rel = abs(evap_s - evap_o) / abs(evap_o)
    print(f"evaporated mass: surrogate {evap_s:.6e}, FP64 oracle {evap_o:.6e}, rel {rel:.2e}; "
          f"surrogate share {frac:.3f}")
    assert evap_o > 0
    assert rel < 5e-3
    assert frac > 0.9


def test_surrogate_form_refuses_other_deposition_species():
    # the network sees the fuel's own gas-phase mass fraction, so the
    # surrogate form needs each fuel's vapour in its own species
    from pelemp_amd import _capi as C
    from pelemp_amd import problem as P
    from pelemp_amd._capi import PelempError
    from pelemp_amd.spray import SprayParticleContainer

    prob = P.config2(ncell=8)
    sd = dict(prob.sdata, dep_indx=[prob.sdata["indx"][1], prob.sdata["indx"][0]])
    with pytest.raises(PelempError):
        SprayParticleContainer(prob.mech, sd, prob.comps, prob.geom, prob.n, form=C.FORM_SURROGATE)
