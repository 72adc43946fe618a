"""Liquid fuel property sets (SprayData contents) used by the configurations.

Values are taken from the reference's run decks and converted to SI where
the deck is in CGS; the coefficient forms are those of SprayData
(SprayFuelData.H:146-211): rho and lambda cubic in T, mu cubic in 1/T,
psat = d 10^(a - b/(T + c)) (Antoine), one value = constant
(SpraySetup.cpp:22-58).
"""
from __future__ import annotations

import math

# CGS -> SI factors (erg/g -> J/kg, g/cm^3 -> kg/m^3, P -> Pa s,
# erg/(s cm K) -> W/(m K), dyn/cm^2 -> Pa, dyn/cm -> N/m)
_ERG_G = 1.0e-4
_G_CM3 = 1.0e3
_POISE = 0.1
_LAM = 1.0e-5
_DYNCM2 = 0.1
_DYNCM = 1.0e-3

FUELS = {
    # Exec/SprayTests/PeleC/SprayA_wbreakup/spraya-input:64-77 (CGS)
    "NC12H26": dict(
        critT=658.0,
        boilT=489.0,
        latent=3.59411e9 * _ERG_G,
        cp=2.217e7 * _ERG_G,
        lambda_coef=[c * _LAM for c in (2059.0484837579133, -2.4391762758823385,
                                        9.633688833837596e-06, 6.93350414269127e-07)],
        mu_coef=[c * _POISE for c in (-0.028611188933370675, 39.40765272136905,
                                      -18125.486152106896, 3021587.187559424)],
        rho_coef=[c * _G_CM3 for c in (1.7799145710513213, -0.006848424883513244,
                                       1.5016559178741312e-05, -1.2149370117669905e-08)],
        psat_coef=[4.265416493004907, 1741.1453093807672, -80.90351868881419, 1.0e6 * _DYNCM2],
        sigma=15.0 * _DYNCM,
        ref_T=298.15,
    ),
    # Exec/SprayTests/PeleLMeX/abramzon_test/inputs_2d:63-77 (MKS)
    "NC10H22": dict(
        critT=617.8,
        boilT=447.27,
        latent=3.5899e5,
        cp=2192.1,
        lambda_coef=[0.0, 0.0, 0.0, 0.0],
        mu_coef=[0.0, 0.0, 0.0, 0.0],
        rho_coef=[640.0, 0.0, 0.0, 0.0],
        psat_coef=[4.07857, 1501.268, -78.67, 1.0e5],
        sigma=None,
        ref_T=298.15,
    ),
}

# n-alkanes of the six-component liquid (C8-C16).  The reference tree holds no
# property set for these species (its decks cover n-dodecane, n-decane and
# n-heptane), so these are literature values (NIST Chemistry WebBook and Yaws,
# Thermophysical Properties of Chemicals and Hydrocarbons): critical and
# normal boiling temperatures, heat of vaporisation at 298.15 K, liquid heat
# capacity and density at 298 K.  The vapour pressure uses the reference's
# Clausius-Clapeyron branch (psat_coef[3] == 0, Drag.H:78-84) from the boiling
# point and latent heat; lambda and mu are not used by the droplet model
# without films or breakup and are left zero.
def _alkane(critT, boilT, dh_kj_mol, cp_j_molk, rho, mw):
    return dict(critT=critT, boilT=boilT, latent=dh_kj_mol * 1.0e3 / mw, cp=cp_j_molk / mw,
                lambda_coef=[0.0, 0.0, 0.0, 0.0], mu_coef=[0.0, 0.0, 0.0, 0.0], rho_coef=[rho, 0.0, 0.0, 0.0],
                psat_coef=[0.0, 0.0, 0.0, 0.0], sigma=None, ref_T=298.15)


FUELS["NC8H18"] = _alkane(568.7, 398.8, 41.5, 254.6, 699.0, 0.114232)
FUELS["NC13H28"] = _alkane(675.0, 508.6, 66.7, 406.7, 753.0, 0.184367)
FUELS["NC14H30"] = _alkane(693.0, 526.7, 71.3, 438.0, 759.0, 0.198394)
FUELS["NC16H34"] = _alkane(723.0, 560.0, 81.4, 501.6, 770.0, 0.226448)

# the six liquid species of the alkane mechanism, in its species order
ALKANE6 = ("NC8H18", "NC10H22", "NC12H26", "NC13H28", "NC14H30", "NC16H34")


def spray_data(fuels, species, *, mass_trans=True, mom_trans=True, fixed_parts=False, dtmod=1.0,
               body_force=(0.0, 0.0, 0.0), do_breakup=0, do_splash=False, sigma=None, wall_T=-1.0,
               theta_c_deg=-1.0, dep_species=None) -> dict:
    """SprayData dict for liquid `fuels` in a gas mechanism with `species`.

    `latent` is set to the reference latent heat; the library subtracts the
    gas enthalpy at ref_T when the context is created (SpraySetup.cpp:284-290).
    `dep_species`: the gas species that receive each fuel's vapour (the
    reference's `dep_fuel_species` input, SpraySetup.cpp:103-125); by default
    each fuel's own species.
    """
    nf = len(fuels)
    d = dict(mass_trans=int(mass_trans), mom_trans=int(mom_trans), fixed_parts=int(fixed_parts),
             do_splash=int(do_splash), do_breakup=int(do_breakup), nfuel=nf, dtmod=float(dtmod),
             body_force=list(body_force), min_eb_vfrac=0.05, wall_T=float(wall_T),
             theta_c=(theta_c_deg * math.pi / 180.0) if theta_c_deg >= 0 else -1.0)
    ref_T = {FUELS[f]["ref_T"] for f in fuels}
    if len(ref_T) != 1:
        raise ValueError("all fuels must share one reference temperature")
    d["ref_T"] = ref_T.pop()
    for k in ("critT", "boilT", "cp"):
        d[k] = [FUELS[f][k] for f in fuels]
    d["ref_latent"] = [FUELS[f]["latent"] for f in fuels]
    d["latent"] = list(d["ref_latent"])
    for k in ("psat_coef", "rho_coef", "lambda_coef", "mu_coef"):
        d[k] = [c for f in fuels for c in FUELS[f][k]]
    d["indx"] = [species.index(f) for f in fuels]
    if dep_species is None:
        d["dep_indx"] = list(d["indx"])
    else:
        if len(dep_species) != nf:
            raise ValueError("one deposition species per fuel")
        d["dep_indx"] = [species.index(f) for f in dep_species]
    sig = sigma if sigma is not None else FUELS[fuels[0]].get("sigma")
    d["sigma"] = float(sig) if sig is not None else -1.0
    return d
