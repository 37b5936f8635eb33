"""Shared helpers for the tests: load golden decks, build small synthetic decks."""
import os

import numpy as np

from rismo2d_amd import deck as dk

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))


def golden_deck(g):
    """Rebuild the Deck stored in a golden fixture (exact arrays, not regenerated)."""
    return dk.Deck(x=g["x"], y=g["y"], z=g["z"], elems=g["elems"], nnd=g["nnd"],
                   manning=g["manning"], vt=g["vt"],
                   lines={dk.INLET_LINE: g["inlet"], dk.OUTLET_LINE: g["outlet"]},
                   U0=g["U0"], V0=g["V0"], S0=g["S0"], Q_inlet=float(g["Q"]),
                   S_inlet=float(g["S_in"]), S_outlet=float(g["S_out"]),
                   veg_dp=g.get("veg_dp"), veg_sp=g.get("veg_sp"),
                   turb=int(g["turb"]) if "turb" in g else 1, turb_st=g.get("turb_st"),
                   turb_kdest=g.get("turb_kdest"), hydrograph=g.get("hydrograph"),
                   turb_estx=g.get("turb_estx"), turb_esty=g.get("turb_esty"))
