"""A deck in the reference's input files run end to end on the GPU (rismo2d_amd.run):
the written golden decks reach the reference program's final state, stationary and
through the unsteady time loop with an inlet hydrograph."""
import os

import numpy as np
import pytest

from helpers import golden_deck, load_golden
from rismo2d_amd import deck as dk
from rismo2d_amd import rmx
from rismo2d_amd import run as rr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rmx.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["tri20x4", "elder20x4", "hydro20x4"])
def test_reference_deck_runs_to_reference_state(ctx, name, tmp_path):
    print(f"[gpu] reference deck run: {name}", flush=True)
    g = load_golden(name)
    kw = {}
    if "stationary" in g:
        kw = dict(tm_steps=int(g["tm_steps"]), dt=float(g["dt"]), theta=float(g["theta"]),
                  stationary=int(g["stationary"]))
    dk.write_reference_deck(golden_deck(g), str(tmp_path), **kw)
    d, run, counts, (U, V, S, z) = rr.run_reference_deck(os.path.join(str(tmp_path), "input.dat"),
                                                         ctx=ctx, solve=rr.PLAIN)
    print(f"    Newton steps per time step {counts}")
    assert all(k < (run["maxiter"] or 12) for k in counts)
    us = np.abs(g["final_U"]).max()
    assert np.abs(U - g["final_U"]).max() <= 1e-8 * us
    assert np.abs(V - g["final_V"]).max() <= 1e-8 * us
    Hr = g["final_S"] - g["final_z"]
    assert np.abs((S - z) - Hr).max() <= 1e-8 * np.abs(Hr).max()
    if run["stationary"] == 0:
        # the reference's Newton steps per time step, from its records
        ts = [int(g[f"s{k}_tstep"]) for k in range(int(g["nsteps"]))]
        assert counts == [ts.count(t) for t in sorted(set(ts))]
