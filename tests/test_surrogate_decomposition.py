"""CPU pin of the surrogate's decomposition of the substep: with the film
viscosity and diffusivity of the oracle's transport at the film state, the
closed forms the surrogate keeps (drag coefficient, Sherwood number,
m_dot = -pi d Sh* rho D ln(1 + B_M)) reproduce the CPU oracle's droplet
source (calculateSpraySource, Drag.H:137-441) on points of the training hull."""
import ctypes
import math

import numpy as np
import pytest

from tests.test_gpu_surrogate import _sample_point, _skin_state


def test_closed_forms_reproduce_oracle_source():
    from oracle import oracle_py as O
    from pelemp_amd import _capi as C
    from pelemp_amd import surrogate as S
    from pelemp_amd.problem import mech_info

    mech = C.MECH_DECANE_DODECANE_AIR
    W = np.array(mech_info(mech)["W"], dtype=np.float64)
    sdata = O.latent_shifted(S.mech_spray_data(mech), mech, len(W))
    sd = C.spray_data_c(sdata)
    fuel_idx = sdata["indx"]
    nf = len(fuel_idx)
    checked = 0
    for i in range(200):
        dia, Tg, p, urel, Y, rho, Yl, cB, Td = _sample_point(31, i, S.DEFAULT_HULL, W, sd, fuel_idx, O)
        BM, Xv, Ysk, mw_skin = _skin_state(sd, O, W, fuel_idx, Y, p, Td, Yl, cB)
        if not (BM > 1e-6 and min(Xv) > 0):
            continue
        rd = np.zeros(10 + nf)
        rd[3], rd[4] = Td, dia
        rd[5:5 + nf] = Yl
        rd[5 + nf] = 1.0
        dt = 1e-9
        _, _, pid, src, _ = O.spray_source_one(mech, sdata, dt, np.concatenate([[urel, 0.0, 0.0, Tg, rho], Y]), rd)
        assert pid > 0
        T_skin = Td + (Tg - Td) / 3.0
        rho_skin = mw_skin * p / (8.31446261815324 * T_skin)
        tr = O.gas_props(mech, T_skin, rho_skin, Ysk)
        mu = tr["mu"]
        rhoD = sum(tr["rhoD"][ix] * mw_skin / W[ix] * x / sum(Xv) for x, ix in zip(Xv, fuel_idx))
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


def test_planted_sherwood_error_is_caught():
    # a Sherwood number without the Spalding correction F(B_M) misses the
    # oracle's m_dot on the same points: the check above is not vacuous
    from oracle import oracle_py as O
    from pelemp_amd import _capi as C
    from pelemp_amd import surrogate as S
    from pelemp_amd.problem import mech_info

    mech = C.MECH_DECANE_DODECANE_AIR
    W = np.array(mech_info(mech)["W"], dtype=np.float64)
    sdata = O.latent_shifted(S.mech_spray_data(mech), mech, len(W))
    sd = C.spray_data_c(sdata)
    fuel_idx = sdata["indx"]
    nf = len(fuel_idx)
    misses = 0
    for i in range(60):
        dia, Tg, p, urel, Y, rho, Yl, cB, Td = _sample_point(32, i, S.DEFAULT_HULL, W, sd, fuel_idx, O)
        BM, Xv, Ysk, mw_skin = _skin_state(sd, O, W, fuel_idx, Y, p, Td, Yl, cB)
        if not (BM > 1e-3 and min(Xv) > 0):
            continue
        rd = np.zeros(10 + nf)
        rd[3], rd[4] = Td, dia
        rd[5:5 + nf] = Yl
        rd[5 + nf] = 1.0
        _, _, _, src, _ = O.spray_source_one(mech, sdata, 1e-9, np.concatenate([[urel, 0.0, 0.0, Tg, rho], Y]), rd)
        T_skin = Td + (Tg - Td) / 3.0
        rho_skin = mw_skin * p / (8.31446261815324 * T_skin)
        tr = O.gas_props(mech, T_skin, rho_skin, Ysk)
        rhoD = sum(tr["rhoD"][ix] * mw_skin / W[ix] * x / sum(Xv) for x, ix in zip(Xv, fuel_idx))
        Re = rho_skin * urel * dia / tr["mu"]
        Sh0 = 1 + max(min(400.0, Re) ** 0.077, 1.0) * (1 + Re * tr["mu"] / rhoD) ** (1 / 3)
        wrong = -math.pi * dia * Sh0 * rhoD * math.log1p(BM)
        if abs(wrong - src[3]) > 1e-4 * abs(src[3]):
            misses += 1
    assert misses > 10
