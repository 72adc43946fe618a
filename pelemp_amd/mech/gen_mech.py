"""Generate the gas-phase thermo/transport tables used by the spray kernels.

The reference obtains species molecular weights, cp_k(T), h_k(T) and the
mixture-averaged transport properties from PelePhysics (EOS "Fuego" =
NASA-7 polynomials, transport "Simple" = log-log cubic fits of Chapman-Enskog
properties; call sites Source/PP_Spray/Drag.H:220-221,252-256,
SprayParticles.cpp:332).  PelePhysics is not part of the reference tree, so
this build supplies its own tables, generated here:

* N2, O2 NASA-7 coefficients and Lennard-Jones data are taken from the data
  files shipped with the reference
  (Util/MechanismPAH/PAHmechfiles/PAH.chthermo:3-6,39-42 and
  PAH.chtrans:2,8).
* n-alkane (CnH2n+2) ideal-gas heat capacity and enthalpy of formation come
  from Benson group additivity (groups C-(C)(H)3 and C-(C)2(H)2; S. W.
  Benson, Thermochemical Kinetics, 2nd ed., Wiley 1976, table A.1), fitted
  here by least squares into NASA-7 form over 250-1500 K (one polynomial
  used for both NASA ranges).
* n-alkane Lennard-Jones parameters come from the corresponding-states
  correlation eps/k = 0.77 Tc, sigma = 2.44 (Tc/Pc)^(1/3) [A, K, atm]
  (Bird, Stewart & Lightfoot, Transport Phenomena, eq. 1.4-11,12), with the
  critical constants listed below.
* Pure-species viscosity and binary diffusion coefficients follow
  Chapman-Enskog theory with the Neufeld-Janzen-Aziz (J. Chem. Phys. 57,
  1100 (1972)) collision-integral correlations; thermal conductivity uses
  the modified Eucken relation lambda = mu (cp + 5/4 R/W).  ln(mu), ln(lambda)
  and ln(D_ij at 1 atm) are fitted as cubics in ln T over 250-3000 K, the
  form the kernels evaluate.

Run: python -m pelemp_amd.mech.gen_mech   (rewrites mech_data.h / .json)
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

RU = 8.31446261815324  # J/mol/K
KB = 1.380649e-23
NA = 6.02214076e23
PATM = 101325.0
W_C = 12.011e-3
W_H = 1.008e-3

# NASA-7 data from PAH.chthermo (low range first, 300-1000-5000 K)
NASA_DATA = {
    "O2": dict(
        W=2 * 15.999e-3,
        Tmid=1000.0,
        low=[3.78245636e00, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09,
             3.24372837e-12, -1.06394356e03, 3.65767573e00],
        high=[3.28253784e00, 1.48308754e-03, -7.57966669e-07, 2.09470555e-10,
              -2.16717794e-14, -1.08845772e03, 5.45323129e00],
        eps=107.400, sigma=3.458,
    ),
    "N2": dict(
        W=2 * 14.007e-3,
        Tmid=1000.0,
        low=[3.29867700e00, 1.40824040e-03, -3.96322200e-06, 5.64151500e-09,
             -2.44485400e-12, -1.02089990e03, 3.95037200e00],
        high=[2.92664000e00, 1.48797680e-03, -5.68476000e-07, 1.00970380e-10,
              -6.75335100e-15, -9.22797700e02, 5.98052800e00],
        eps=97.530, sigma=3.621,
    ),
}

# Benson groups, Cp in cal/mol/K at these temperatures, dHf298 in kcal/mol
BENSON_T = [300.0, 400.0, 500.0, 600.0, 800.0, 1000.0, 1500.0]
BENSON_CH3 = dict(cp=[6.19, 7.84, 9.40, 10.79, 13.02, 14.77, 17.58], hf=-10.20)
BENSON_CH2 = dict(cp=[5.50, 6.95, 8.25, 9.35, 11.07, 12.34, 14.25], hf=-4.93)
CAL = 4.184

# critical constants (Tc [K], Pc [bar]) for the corresponding-states LJ data
ALKANE_CRIT = {
    8: (568.7, 24.90),
    10: (617.7, 21.10),
    12: (658.0, 18.20),
    13: (675.0, 16.80),
    14: (693.0, 15.70),
    16: (723.0, 14.00),
}


def alkane_nasa(n: int) -> dict:
    ncp = [(2 * a + (n - 2) * b) * CAL / RU for a, b in zip(BENSON_CH3["cp"], BENSON_CH2["cp"])]
    T = np.array(BENSON_T)
    # least squares quartic in T for cp/R (columns scaled for conditioning)
    A = np.vander(T / 1000.0, 5, increasing=True)
    c, *_ = np.linalg.lstsq(A, np.array(ncp), rcond=None)
    a = [c[i] / 1000.0 ** i for i in range(5)]
    hf = (2 * BENSON_CH3["hf"] + (n - 2) * BENSON_CH2["hf"]) * CAL * 1000.0  # J/mol
    T0 = 298.15
    h_over_R_poly = a[0] * T0 + a[1] * T0 ** 2 / 2 + a[2] * T0 ** 3 / 3 + a[3] * T0 ** 4 / 4 + a[4] * T0 ** 5 / 5
    a6 = hf / RU - h_over_R_poly
    a7 = 0.0  # entropy is not used by the spray update
    coef = [float(x) for x in a] + [float(a6), a7]
    Tc, Pc_bar = ALKANE_CRIT[n]
    Pc_atm = Pc_bar * 1e5 / PATM
    return dict(
        W=n * W_C + (2 * n + 2) * W_H,
        Tmid=1000.0,
        low=coef,
        high=coef,
        eps=0.77 * Tc,
        sigma=2.44 * (Tc / Pc_atm) ** (1.0 / 3.0),
    )


def species_data(name: str) -> dict:
    if name in NASA_DATA:
        return dict(NASA_DATA[name])
    assert name.startswith("NC") and "H" in name
    n = int(name[2:name.index("H")])
    return alkane_nasa(n)


def nasa_cp_R(d, T):
    a = d["low"] if T < d["Tmid"] else d["high"]
    return a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))


def omega22(Ts):
    return 1.16145 * Ts ** -0.14874 + 0.52487 * math.exp(-0.77320 * Ts) + 2.16178 * math.exp(-2.43787 * Ts)


def omega11(Ts):
    return (1.06036 * Ts ** -0.15610 + 0.19300 * math.exp(-0.47635 * Ts)
            + 1.03587 * math.exp(-1.52996 * Ts) + 1.76474 * math.exp(-3.89411 * Ts))


def mu_ce(d, T):
    """Chapman-Enskog pure-species viscosity [Pa s]."""
    m = d["W"] / NA
    sig = d["sigma"] * 1e-10
    return 5.0 / 16.0 * math.sqrt(math.pi * m * KB * T) / (math.pi * sig ** 2 * omega22(T / d["eps"]))


def lam_eucken(d, T):
    cp = nasa_cp_R(d, T) * RU / d["W"]
    return mu_ce(d, T) * (cp + 1.25 * RU / d["W"])


def dbin_ce(di, dj, T, p=PATM):
    """Chapman-Enskog binary diffusion coefficient [m^2/s] at pressure p."""
    mij = (di["W"] * dj["W"]) / (di["W"] + dj["W"]) / NA
    sig = 0.5 * (di["sigma"] + dj["sigma"]) * 1e-10
    eps = math.sqrt(di["eps"] * dj["eps"])
    return 3.0 / 16.0 * math.sqrt(2.0 * math.pi * (KB * T) ** 3 / mij) / (p * math.pi * sig ** 2 * omega11(T / eps))


FIT_T = np.geomspace(250.0, 3000.0, 200)


def logfit(fn):
    lt = np.log(FIT_T)
    y = np.log([fn(T) for T in FIT_T])
    c = np.polyfit(lt, y, 3)[::-1]  # c0 + c1 lnT + c2 lnT^2 + c3 lnT^3
    return [float(x) for x in c]


MECHS = {
    # name: species list (order = index in the gas state)
    "dodecane_air": ["O2", "N2", "NC12H26"],
    "decane_dodecane_air": ["O2", "N2", "NC10H22", "NC12H26"],
    "alkane6_air": ["O2", "N2", "NC8H18", "NC10H22", "NC12H26", "NC13H28", "NC14H30", "NC16H34"],
    # the species of the reference decks' decane_3sp chemistry (n-decane in air)
    "decane_air": ["O2", "N2", "NC10H22"],
}


def build_mech(species):
    data = [species_data(s) for s in species]
    out = dict(species=species, W=[d["W"] for d in data], Tmid=[d["Tmid"] for d in data],
               low=[d["low"] for d in data], high=[d["high"] for d in data],
               eps=[d["eps"] for d in data], sigma=[d["sigma"] for d in data])
    out["fitmu"] = [logfit(lambda T, d=d: mu_ce(d, T)) for d in data]
    out["fitlam"] = [logfit(lambda T, d=d: lam_eucken(d, T)) for d in data]
    pairs = []
    for i in range(len(data)):
        for j in range(i + 1, len(data)):
            pairs.append(logfit(lambda T, a=data[i], b=data[j]: dbin_ce(a, b, T)))
    out["fitdbin"] = pairs
    return out


def cxx_array(name, vals, per_line=4):
    s = [f"  static constexpr double {name}[{len(vals)}] = {{"]
    for i in range(0, len(vals), per_line):
        s.append("      " + ", ".join(float(v).hex() for v in vals[i:i + per_line]) + ",")
    s.append("  };")
    return "\n".join(s)


def struct_name(mname):
    return "Mech_" + mname


def emit_header(mechs) -> str:
    lines = [
        "// mech_data.h -- generated by pelemp_amd/mech/gen_mech.py; do not edit.",
        "// Gas-phase species tables (SI units): molecular weight [kg/mol], NASA-7",
        "// coefficients (cp/R, h/RT), and cubic fits in ln T of ln mu [Pa s],",
        "// ln lambda [W/m/K] and ln D_ij(1 atm) [m^2/s] (pairs i<j, row-major).",
        "#pragma once",
        "",
        "namespace pelemp {",
        "",
    ]
    for mname, m in mechs.items():
        ns = len(m["species"])
        lines.append(f"struct {struct_name(mname)} {{")
        lines.append(f"  static constexpr int NS = {ns};")
        lines.append(f"  // species: {', '.join(m['species'])}")
        lines.append(cxx_array("W", m["W"]))
        lines.append(cxx_array("Tmid", m["Tmid"]))
        lines.append(cxx_array("nasa_lo", [c for a in m["low"] for c in a], 7))
        lines.append(cxx_array("nasa_hi", [c for a in m["high"] for c in a], 7))
        lines.append(cxx_array("fitmu", [c for a in m["fitmu"] for c in a]))
        lines.append(cxx_array("fitlam", [c for a in m["fitlam"] for c in a]))
        lines.append(cxx_array("fitdbin", [c for a in m["fitdbin"] for c in a]))
        lines.append("};")
        lines.append("")
    lines.append("}  // namespace pelemp")
    lines.append("")
    return "\n".join(lines)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    mechs = {k: build_mech(v) for k, v in MECHS.items()}
    with open(os.path.join(here, "..", "csrc", "mech_data.h"), "w") as f:
        f.write(emit_header(mechs))
    with open(os.path.join(here, "mech_data.json"), "w") as f:
        json.dump(mechs, f, indent=1)


if __name__ == "__main__":
    main()
