"""What the trained network gains over the method without it.

Without the network, a droplet model that skips the film evaluation
(reference path: calcVaporState's 1/3-rule film state, the transport call at
that state and calcHeatCoeff's B_T fixed point, Drag.H:204-347) falls back to
the classical frozen-property closure: viscosity, conductivity and vapour
diffusivity of the far-field gas, and Nu* = Nu_0 without the blowing
correction.  The committed network (pelemp_amd/surrogate_data/
decane_dodecane_air.npz, evaluated here in float64 with its bf16-rounded
weights and activations, as the MFMA kernel does) replaces exactly that
evaluation.  On hull points drawn like the training samples, the test
computes the three film quantities -- ln(sum rho D_k), ln(lambda Nu*),
ln(mu) -- with the CPU oracle (the plain FP64 method), the network and the
closure, and checks that the network's error is a small fraction of the
closure's: the accuracy it buys at the closure's cost class (one small MLP
instead of the transport call and fixed point).  Both committed networks:
the two-fuel one (config 2) and the n-dodecane one (configs 1 and 4)."""
import ctypes
import math

import numpy as np
import pytest

from tests.test_gpu_surrogate import _sample_point, _skin_state

RU = 8.31446261815324


def _film_quantities(O, mech, sd, W, fuel_idx, dia, Tg, p, urel, Y, rho, Yl, cB, Td):
    """Oracle film state: (z0, z1, z2) = ln rhoD, ln lambda Nu*, ln mu, or None
    outside the evaporating regime the network is trained on."""
    BM, Xv, Ysk, mw_skin = _skin_state(sd, O, W, fuel_idx, Y, p, Td, Yl, cB)
    if not (BM > 1e-6 and min(Xv) > 0):
        return None, None
    T_skin = Td + (Tg - Td) / 3.0
    rho_skin = mw_skin * p / (RU * T_skin)
    tr = O.gas_props(mech, T_skin, rho_skin, Ysk)
    mu, lam, cp = tr["mu"], tr["lam"], tr["cp"]
    rhoD = sum(tr["rhoD"][ix] * mw_skin / W[ix] * x / sum(Xv) for x, ix in zip(Xv, fuel_idx))
    Re = rho_skin * urel * dia / mu
    cp_skin = float(np.dot(Ysk, cp))
    Pr = mu * cp_skin / lam
    F2 = max(min(400.0, Re) ** 0.077, 1.0)
    Nu0 = 1.0 + F2 * (1.0 + Re * Pr) ** (1.0 / 3.0)
    Sh0 = 1.0 + F2 * (1.0 + Re * mu / rhoD) ** (1.0 / 3.0)
    Sh = 2.0 + (Sh0 - 2.0) * BM / (math.log1p(BM) * (1.0 + BM) ** 0.7)
    ysk_f = sum(Ysk[ix] for ix in fuel_idx)
    cp_fuel = sum(Ysk[ix] * cp[ix] for ix in fuel_idx) / ysk_f
    ratio = cp_fuel * Sh * rhoD / lam
    Nu = O.calc_heat_coeff(ratio, BM, 1e-7, 1e-15, Nu0)
    # the second liquid feature is 0 for a one-fuel liquid (surr_features)
    x = [math.log(dia), Td, math.log(urel + 0.01), Tg, math.log(p), ysk_f, Yl[0], Yl[-1] if len(Yl) > 1 else 0.0]
    return (math.log(rhoD), math.log(lam * Nu), math.log(mu)), x


def _closure(O, mech, W, fuel_idx, dia, Tg, p, urel, Y, rho):
    """Frozen far-field properties, no blowing correction of the heat transfer."""
    tr = O.gas_props(mech, Tg, rho, Y)
    mu, lam, cp = tr["mu"], tr["lam"], tr["cp"]
    mw = 1.0 / float(np.sum(Y / W))
    rhoD = float(np.mean([tr["rhoD"][ix] * mw / W[ix] for ix in fuel_idx]))
    Re = rho * urel * dia / mu
    Pr = mu * float(np.dot(Y, cp)) / lam
    F2 = max(min(400.0, Re) ** 0.077, 1.0)
    Nu0 = 1.0 + F2 * (1.0 + Re * Pr) ** (1.0 / 3.0)
    return math.log(rhoD), math.log(lam * Nu0), math.log(mu)


@pytest.mark.parametrize("mech_name", ["MECH_DECANE_DODECANE_AIR", "MECH_DODECANE_AIR"])
def test_network_beats_the_frozen_property_closure(mech_name):
    from oracle import mlp_ref as R
    from oracle import oracle_py as O
    from pelemp_amd import _capi as C
    from pelemp_amd import surrogate as S
    from pelemp_amd.problem import mech_info

    mech = getattr(C, mech_name)
    net = S.Surrogate.load(S.default_path(mech))
    n = S.norm_dict(net.norm)
    W = np.array(mech_info(mech)["W"], dtype=np.float64)
    sdata = O.latent_shifted(S.mech_spray_data(mech), mech, len(W))
    sd = C.spray_data_c(sdata)
    fuel_idx = sdata["indx"]
    exact, feats, clos = [], [], []
    for i in range(400):
        dia, Tg, p, urel, Y, rho, Yl, cB, Td = _sample_point(77, i, S.DEFAULT_HULL, W, sd, fuel_idx, O)
        z, x = _film_quantities(O, mech, sd, W, fuel_idx, dia, Tg, p, urel, Y, rho, Yl, cB, Td)
        if z is None:
            continue
        exact.append(z)
        feats.append(x)
        clos.append(_closure(O, mech, W, fuel_idx, dia, Tg, p, urel, Y, rho))
    exact, feats, clos = np.array(exact), np.array(feats), np.array(clos)
    assert exact.shape[0] > 250
    xn = R.normalise_inputs(feats, n["in_center"], n["in_scale"])
    yn, _ = R.forward(net.params.astype(np.float64), xn, quant=True)
    pred = np.asarray(n["out_mean"]) + np.asarray(n["out_std"]) * yn
    err_net = np.abs(pred - exact)
    err_clo = np.abs(clos - exact)
    for q, name in enumerate(S.OUTPUT_NAMES):
        # the network's mean error is a small fraction of the closure's, and
        # its largest error is below the closure's typical one
        assert err_net[:, q].mean() < 0.1 * err_clo[:, q].mean(), name
        assert np.quantile(err_net[:, q], 0.99) < np.median(err_clo[:, q]), name
    # in rate terms: dm/dt scales with rho D, the heat flux with lambda Nu*
    assert np.mean(np.abs(np.expm1(err_net[:, 0]))) < 0.01
    assert np.mean(np.abs(np.expm1(err_clo[:, 0]))) > 0.1
