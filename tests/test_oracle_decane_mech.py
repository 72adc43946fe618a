"""Pins of the decane_air mechanism (O2, N2, NC10H22: the species of the
reference's decane_3sp decks, Exec/SprayTests/PeleLMeX/abramzon_test): its
tables are the corresponding entries of the decane/dodecane mechanism (the
same generator and species data), and a droplet in it behaves as in the
four-species mechanism with no dodecane present, where every extra term is
an exact zero (Drag.H:47-441 sums over the species)."""
import numpy as np

from oracle import oracle_py as O
from pelemp_amd import _capi as C
from pelemp_amd import problem as P
from pelemp_amd.fuels import spray_data

RU = 8.31446261815324
PATM = 101325.0


def test_tables_are_the_four_species_mechanism_entries():
    a = P.mech_info(C.MECH_DECANE_AIR)
    b = P.mech_info(C.MECH_DECANE_DODECANE_AIR)
    assert a["species"] == ["O2", "N2", "NC10H22"]
    for k in ("W", "Tmid", "low", "high", "fitmu", "fitlam"):
        assert a[k] == b[k][:3], k
    # binary pairs (i<j, row-major): (O2,N2) (O2,NC10H22) (N2,NC10H22) are
    # pairs 0, 1, 3 of the four-species list
    assert a["fitdbin"] == [b["fitdbin"][0], b["fitdbin"][1], b["fitdbin"][3]]


def test_droplet_source_equals_the_dodecane_free_four_species_case():
    T, p = 1500.0, 10 * PATM
    Y3 = np.array([0.233, 0.767, 0.0])
    Y4 = np.array([0.233, 0.767, 0.0, 0.0])
    W3 = np.array(P.mech_info(C.MECH_DECANE_AIR)["W"])
    rho = p / (RU * T * np.sum(Y3 / W3))
    g3 = np.array([15.0, 0.0, 0.0, T, rho, *Y3])
    g4 = np.array([15.0, 0.0, 0.0, T, rho, *Y4])
    sd3 = O.latent_shifted(spray_data(["NC10H22"], P.mech_info(C.MECH_DECANE_AIR)["species"]), C.MECH_DECANE_AIR, 3)
    sd4 = O.latent_shifted(spray_data(["NC10H22"], P.mech_info(C.MECH_DECANE_DODECANE_AIR)["species"]),
                           C.MECH_DECANE_DODECANE_AIR, 4)
    assert sd3["latent"] == sd4["latent"]
    r = np.zeros(11)
    r[3], r[4], r[5], r[6] = 300.0, 100e-6, 1.0, 1.0
    for dt in (1e-6, 1e-4):
        Re3, out3, _, src3, cb3 = O.spray_source_one(C.MECH_DECANE_AIR, sd3, dt, g3, r)
        Re4, out4, _, src4, cb4 = O.spray_source_one(C.MECH_DECANE_DODECANE_AIR, sd4, dt, g4, r)
        assert Re3 == Re4
        assert np.array_equal(out3, out4) and np.array_equal(src3, src4) and np.array_equal(cb3, cb4)
        assert src3[3] < 0.0  # the droplet evaporates (the droplet-side mass rate)
