"""Gas-phase thermo/transport tables (the build's stand-in for PelePhysics)
against independent values: JANAF/NIST heat capacities and enthalpies,
literature viscosities, conductivities and diffusivities, direct
Chapman-Enskog evaluation (pins the ln-T fits), and limiting cases of the
mixing rules."""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from pelemp_amd import _capi as C
from pelemp_amd.mech import gen_mech as G
from pelemp_amd.problem import mech_info

RU = 8.31446261815324
MECH = C.MECH_DECANE_DODECANE_AIR  # O2 N2 NC10H22 NC12H26
SP = mech_info(MECH)["species"]
W = np.array(mech_info(MECH)["W"])


def pure(name):
    Y = np.zeros(len(SP))
    Y[SP.index(name)] = 1.0
    return Y


def props(T, Y, p=101325.0):
    Y = np.asarray(Y, dtype=float)
    wbar = 1.0 / np.sum(Y / W)
    rho = p * wbar / (RU * T)
    return O.gas_props(MECH, T, rho, Y), rho


@pytest.mark.parametrize("name,T,cp_molar,tol", [
    ("N2", 300.0, 29.12, 0.005),   # JANAF
    ("N2", 1000.0, 32.70, 0.005),
    ("O2", 300.0, 29.39, 0.005),
    ("O2", 1000.0, 34.87, 0.005),
    ("NC12H26", 298.15, 280.0, 0.02),  # NIST ideal gas, ~280 J/mol/K
    ("NC10H22", 298.15, 233.1, 0.02),
])
def test_cp_literature(name, T, cp_molar, tol):
    pr, _ = props(T, pure(name))
    k = SP.index(name)
    assert abs(pr["cp"][k] * W[k] / cp_molar - 1) < tol


def test_enthalpy_reference_states():
    pr, _ = props(298.15, pure("N2"))
    # elements in their standard state: h(298.15) ~ 0 (within 10 J/mol)
    assert abs(pr["h"][SP.index("N2")] * W[SP.index("N2")]) < 10.0
    assert abs(pr["h"][SP.index("O2")] * W[SP.index("O2")]) < 10.0
    # heats of formation (NIST gas phase): n-dodecane -290.9, n-decane -249.5 kJ/mol
    assert abs(pr["h"][SP.index("NC12H26")] * W[SP.index("NC12H26")] / -290.9e3 - 1) < 0.01
    assert abs(pr["h"][SP.index("NC10H22")] * W[SP.index("NC10H22")] / -249.5e3 - 1) < 0.01


def test_enthalpy_is_integral_of_cp():
    # dh/dT = cp, checked by central differences at several T
    for T in (350.0, 700.0, 1200.0):
        Y = np.full(len(SP), 0.25)
        hp = props(T + 0.5, Y)[0]["h"]
        hm = props(T - 0.5, Y)[0]["h"]
        cp = props(T, Y)[0]["cp"]
        assert np.max(np.abs((hp - hm) / 1.0 / cp - 1)) < 1e-6


@pytest.mark.parametrize("name,T,mu,lam,tol_mu,tol_lam", [
    ("N2", 300.0, 1.78e-5, 0.0259, 0.03, 0.05),
    ("O2", 300.0, 2.07e-5, 0.0266, 0.03, 0.05),
    ("N2", 1000.0, 4.15e-5, 0.0647, 0.04, 0.10),
])
def test_pure_transport_literature(name, T, mu, lam, tol_mu, tol_lam):
    pr, _ = props(T, pure(name))
    # a pure gas: the mixing rules must return the pure-species values
    assert abs(pr["mu"] / mu - 1) < tol_mu
    assert abs(pr["lam"] / lam - 1) < tol_lam


def test_fits_match_chapman_enskog():
    info = mech_info(MECH)
    data = [G.species_data(s) for s in SP]
    for T in (300.0, 700.0, 1500.0, 2500.0):
        for k, s in enumerate(SP):
            pr, _ = props(T, pure(s))
            assert abs(pr["mu"] / G.mu_ce(data[k], T) - 1) < 0.015
            assert abs(pr["lam"] / G.lam_eucken(data[k], T) - 1) < 0.03
    assert len(info["fitdbin"]) == len(SP) * (len(SP) - 1) // 2


def test_binary_diffusion_o2_n2():
    # D(O2,N2) at 300 K, 1 atm ~ 2.0e-5 m^2/s (Marrero & Mason)
    Y = np.zeros(len(SP))
    Y[SP.index("O2")] = 1e-6
    Y[SP.index("N2")] = 1.0 - 1e-6
    pr, rho = props(300.0, Y)
    k, b = SP.index("O2"), SP.index("N2")
    D = pr["rhoD"][k] / rho * W[b] / W[k]
    assert abs(D / 2.0e-5 - 1) < 0.08


def test_mixture_diffusion_pressure_scaling():
    # rho*D is independent of pressure at fixed T and composition
    Y = np.array([0.2, 0.7, 0.05, 0.05])
    a, _ = props(800.0, Y, 101325.0)
    b, _ = props(800.0, Y, 30 * 101325.0)
    assert np.max(np.abs(a["rhoD"] / b["rhoD"] - 1)) < 1e-12
    # trace species in a binary bath: Ddiag_k -> rho D_k,bath W_k / W_bath
    # (Ddiag carries W_k/W_mix; Drag.H:312-314 divides it out for the fuel)
    Y = np.zeros(len(SP))
    Y[SP.index("NC12H26")] = 1e-9
    Y[SP.index("N2")] = 1 - 1e-9
    pr, rho = props(500.0, Y)
    Dkb = G.dbin_ce(G.species_data("NC12H26"), G.species_data("N2"), 500.0)
    k, b = SP.index("NC12H26"), SP.index("N2")
    assert abs(pr["rhoD"][k] / rho / (Dkb * W[k] / W[b]) - 1) < 0.02
