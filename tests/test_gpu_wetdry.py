"""Drying and rewetting on the GPU (method 2, rmx_dry_rewet) against the reference program's
records, through the C ABI.

The decks: ``wetdry20x6`` (a bank rising out of the water: the far bank dries at the first
call, the strip beside it becomes marsh, the wet/dry line a slip boundary) and
``rewet20x6`` (the water rises from below the outlet level: elements of the bank get wet at
a later call and the equations grow from 937 to 1115) and ``marshmid20x6`` (marsh corners
lowered by half the dry limit: 61 midside nodes between dryLimit / 10 and dryLimit below
their original bed take the marsh flag).  The runner's time loop calls the
drying and rewetting on the reference's schedule (EqsUVS2D.cpp:200-222); every Newton step
applies the reference's own correction (rmx_set_solution), so that the device's state
follows the reference's trajectory and each drying pass sees the reference's input.

Bars: dry and marsh flags, boundary kinds, numbering and CSR pattern bit-exact; nodal state,
lowered bed, friction, inlet discharge and rotations within 1e-13 of their scale (the
update's arithmetic is the reference's, the rotations come from normals summed in another
order); the FP64 assembly within 1e-12 of the row maximum of the oracle's FP64 assembly of
the same state (the oracle runs in lockstep; its own records equal the reference's bit for
bit), and within 16 half-ulps of FP32 of the sum of the entry's |element contributions| of
the reference's FP32-stored matrix (its float accumulation); the right-hand side, an FP64
sum of element forces that cancels more as the Newton steps converge, within 1e-12 of the
first step's maximum (the scale of the forces) of the oracle's and the reference's.
"""
import numpy as np
import pytest

from helpers import golden_deck, load_golden
from oracle import pyoracle as po
from rismo2d_amd import rmx
from rismo2d_amd import run as rr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rmx.Context(0)
    yield c
    c.close()


def _rel(a, b, scale):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(scale, 1e-300))


@pytest.mark.parametrize("name", ["wetdry20x6", "rewet20x6", "marshmid20x6"])
def test_device_drying_follows_reference_records(ctx, name):
    print(f"[gpu] drying and rewetting on the device, {name}", flush=True)
    g = load_golden(name)
    method, freq, dry_limit, rewet_limit, _passes, count_down = g["dryrew"]
    ctx.upload_mesh(golden_deck(g))
    ctx.set_params(dry_method=int(method), dry_limit=float(dry_limit),
                   rewet_limit=float(rewet_limit), count_down=int(count_down))
    om = po.OracleModel(golden_deck(g))
    om.set_params(dry_method=int(method), dry_limit=float(dry_limit))
    om.set_dryrew(float(rewet_limit), int(count_down))
    calls = []

    class Cycle:                       # the device behind the runner's time loop
        k = 0

        def time_step(self, dt, theta, first):
            ctx.time_step(dt, theta, first)

        def slip_flow(self):
            ctx.slip_flow()

        def dry_rewet(self):
            calls.append((self.k,) + ctx.dry_rewet())
            assert om.dry_rewet() == calls[-1][1:]
            return calls[-1][1:]

        def relax_times(self):
            return ctx.relax_times()

    cyc = Cycle()

    def newton():
        k = cyc.k
        p = f"s{k}_"
        n = ctx.prepare()
        nf, ef = ctx.wetdry()
        np.testing.assert_array_equal(nf, g[p + "flag"].astype(np.int64) & (64 | 256),
                                      err_msg=f"dry/marsh flags at step {k}")
        eq = ctx.numbering()
        np.testing.assert_array_equal(eq, np.stack([g[p + "eqU"], g[p + "eqV"], g[p + "eqS"]], 1),
                                      err_msg=f"numbering at step {k}")
        assert n == len(g[p + "xsol"])
        rp, cols = ctx.pattern()
        np.testing.assert_array_equal(np.diff(rp), g[p + "width"])
        np.testing.assert_array_equal(cols, g[p + "cols"])
        U, V, S, z = ctx.state()
        us = np.abs(g[p + "U"]).max() + np.abs(g[p + "V"]).max()
        hs = np.abs(g[p + "S"]).max()
        errs = dict(U=_rel(U, g[p + "U"], us), V=_rel(V, g[p + "V"], us),
                    S=_rel(S, g[p + "S"], hs), z=_rel(z, g[p + "z"], hs))
        nod = ctx.nodal()
        cf, bcU, rotm = nod["cf"], nod["bcU"], nod["rotm"]
        np.testing.assert_array_equal(nod["kind"], g[p + "kind"].astype(np.int32))
        errs["cf"] = _rel(cf, g[p + "cf"], np.abs(g[p + "cf"]).max())
        errs["bcU"] = _rel(bcU, g[p + "bcU"], max(np.abs(g[p + "bcU"]).max(), 1e-12))
        ref_rot = np.stack([g[p + "r00"], g[p + "r01"], g[p + "r10"], g[p + "r11"]], 1)
        errs["rot"] = _rel(rotm, ref_rot, 1.0)
        ctx.assemble(rmx.F64)
        A, B = ctx.matrix(rmx.F64)
        assert om.prepare() == n
        Ao, Bo = om.assemble(64)
        rows = np.repeat(np.arange(n), np.diff(rp))
        rmax = np.zeros(n)
        np.maximum.at(rmax, rows, np.abs(Ao))
        errs["A"] = float((np.abs(A - Ao) / rmax[rows]).max())
        # the reference's FP32 matrix: float sums of float-rounded element contributions
        tol = np.maximum(16 * 2.0 ** -24 * om.assemble_abs(), 1e-12 * rmax[rows])
        errs["A32"] = float((np.abs(A - g[p + "aval"]) / tol).max())
        bs = np.abs(g["s0_bvec"]).max()
        errs["B"] = max(_rel(B, g[p + "bvec"], bs), _rel(B, Bo, bs))
        print(f"    step {k}: {n} equations, dry {int((ef == 1).sum())}, marsh "
              f"{int((ef == 2).sum())} elements; " +
              ", ".join(f"{key} {v:.1e}" for key, v in errs.items()), flush=True)
        for key in ("U", "V", "S", "z", "cf", "bcU", "rot"):
            assert errs[key] <= 1e-13, (k, key, errs[key])
        assert errs["A"] <= 1e-12 and errs["B"] <= 1e-12, (k, errs["A"], errs["B"])
        assert errs["A32"] <= 1.0, (k, errs["A32"])
        Xo, _ = om.newton_step(prec=32, solver=1, max_iter=1000, eps=1e-10)
        np.testing.assert_array_equal(Xo, g[p + "xsol"])
        ctx.set_solution(g[p + "xsol"])
        st = ctx.newton_update()
        cyc.k += 1
        return st

    maxit = int(g["maxiter"]) if "maxiter" in g else 12
    counts = rr.time_loop(cyc, 1, 0.0, 1.0, maxit, newton, stationary=True, dry_freq=int(freq))
    print(f"    calls (step, dried, wetted): {calls}")
    assert counts == [int(g["nsteps"])]
    U, V, S, z = ctx.state()
    us = np.abs(g["final_U"]).max() + np.abs(g["final_V"]).max()
    assert _rel(U, g["final_U"], us) <= 1e-13 and _rel(V, g["final_V"], us) <= 1e-13
    assert _rel(S, g["final_S"], np.abs(g["final_S"]).max()) <= 1e-13
    nf, ef = ctx.wetdry()
    np.testing.assert_array_equal(nf, g["final_flag"].astype(np.int64) & (64 | 256))
    assert calls[0][0] == 0 and calls[0][1] > 0 and (ef == 2).any()
    if name == "rewet20x6":
        assert any(k > 0 and w > 0 for k, _, w in calls)
    else:
        assert (ef == 1).sum() == calls[0][1]


@pytest.mark.parametrize("name", ["wetdryu20x6", "shelfu20x6"])
def test_device_unsteady_drying_follows_reference_records(ctx, name):
    """Unsteady flow (4 x 30 s, theta 0.7, drying every 3 Newton steps of each cycle),
    driven by the reference's corrections: at every Newton step the dry / marsh flags and
    those of the previous time step bit-exact, the state bit-exact to 1e-13 of its scale,
    the time derivatives and the previous state within 1e-12 of theirs, the numbering
    bit-exact.  ``wetdryu20x6``: the bank under a rising hydrograph.  ``shelfu20x6``: a
    shelf that gets wet at the second time step, so the time derivatives of nodes dry or
    marsh at the previous time step are zeroed (EqsUVS2D.cpp:1193-1202)."""
    print(f"[gpu] unsteady drying and rewetting on the device, {name}", flush=True)
    g = load_golden(name)
    method, freq, dry_limit, rewet_limit, _passes, count_down = g["dryrew"]
    dt, th = float(g["dt"]), float(g["theta"])
    ctx.upload_mesh(golden_deck(g))
    ctx.set_params(dry_method=int(method), dry_limit=float(dry_limit),
                   rewet_limit=float(rewet_limit), count_down=int(count_down))

    class Cycle:
        k = 0

        def time_step(self, dt, theta, first):
            ctx.time_step(dt, theta, first)

        def time_commit(self):
            ctx.time_commit()

        def set_inflow(self, Q, S):
            ctx.set_inflow(Q, S)

        def slip_flow(self):
            ctx.slip_flow()

        def dry_rewet(self):
            return ctx.dry_rewet()

        def relax_times(self):
            return ctx.relax_times()

    cyc = Cycle()
    worst = {}

    def newton():
        k = cyc.k
        p = f"s{k}_"
        ctx.prepare()
        nf, _ = ctx.wetdry()
        np.testing.assert_array_equal(nf, g[p + "flag"].astype(np.int64) & (64 | 128 | 256 | 512),
                                      err_msg=f"flags at step {k}")
        eq = ctx.numbering()
        np.testing.assert_array_equal(eq, np.stack([g[p + "eqU"], g[p + "eqV"], g[p + "eqS"]], 1))
        U, V, S, z = ctx.state()
        us = np.abs(g[p + "U"]).max() + np.abs(g[p + "V"]).max()
        errs = dict(U=_rel(U, g[p + "U"], us), V=_rel(V, g[p + "V"], us),
                    S=_rel(S, g[p + "S"], np.abs(g[p + "S"]).max()),
                    z=_rel(z, g[p + "z"], np.abs(g[p + "S"]).max()))
        ts = ctx.time_state()
        for key in ("dUdt", "dVdt", "dSdt", "Uo", "Vo", "So", "dUdto", "dVdto", "dSdto"):
            errs[key] = _rel(ts[key], g[p + key], max(np.abs(g[p + key]).max(), 1e-12))
        for key, v in errs.items():
            worst[key] = max(worst.get(key, 0.0), v)
        for key in ("U", "V", "S", "z"):
            assert errs[key] <= 1e-13, (k, key, errs[key])
        for key in ("dUdt", "dVdt", "dSdt", "Uo", "Vo", "So", "dUdto", "dVdto", "dSdto"):
            assert errs[key] <= 1e-12, (k, key, errs[key])
        ctx.set_solution(g[p + "xsol"])
        st = ctx.newton_update()
        cyc.k += 1
        return st

    counts = rr.time_loop(cyc, int(g["tm_steps"]), dt, th, int(g["maxiter"]), newton,
                          hydrograph=g["hydrograph"] if "hydrograph" in g else None,
                          dry_freq=int(freq))
    print("    worst over the steps: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
    tsteps = [int(g[f"s{k}_tstep"]) for k in range(int(g["nsteps"]))]
    assert counts == [tsteps.count(t) for t in range(1, int(g["tm_steps"]) + 1)]
    U, V, S, z = ctx.state()
    us = np.abs(g["final_U"]).max() + np.abs(g["final_V"]).max()
    assert _rel(U, g["final_U"], us) <= 1e-13 and _rel(S, g["final_S"], np.abs(g["final_S"]).max()) <= 1e-13


def test_device_flood_with_drying_follows_reference(ctx):
    """The 2 km flood channel (``flooddry200x6``: hydrograph, Elder's viscosity recomputed
    every Newton step, 5 x 30 s, theta 0.7) with a bank rising out of the water along one
    side and drying every 3 Newton steps, driven by the reference's corrections: the
    reference's Newton steps per time step, its final state within 1e-13 of its scale and its
    final dry / marsh flags bit-exact."""
    print("[gpu] flood with drying on the device, flooddry200x6", flush=True)
    g = load_golden("flooddry200x6")
    method, freq, dry_limit, rewet_limit, _passes, count_down = g["dryrew"]
    dt, th = float(g["dt"]), float(g["theta"])
    ctx.upload_mesh(golden_deck(g))
    ctx.set_params(dry_method=int(method), dry_limit=float(dry_limit),
                   rewet_limit=float(rewet_limit), count_down=int(count_down))

    class Cycle:
        k = 0

        def time_step(self, dt, theta, first):
            ctx.time_step(dt, theta, first)

        def time_commit(self):
            ctx.time_commit()

        def set_inflow(self, Q, S):
            ctx.set_inflow(Q, S)

        def slip_flow(self):
            ctx.slip_flow()

        def dry_rewet(self):
            return ctx.dry_rewet()

        def relax_times(self):
            return ctx.relax_times()

    cyc = Cycle()

    def newton():
        n = ctx.prepare()
        X = g[f"s{cyc.k}_xsol"]
        assert n == len(X), cyc.k
        ctx.set_solution(X)
        st = ctx.newton_update()
        cyc.k += 1
        return st

    counts = rr.time_loop(cyc, int(g["tm_steps"]), dt, th, 12, newton,
                          hydrograph=g["hydrograph"], dry_freq=int(freq))
    tsteps = [int(g[f"s{k}_tstep"]) for k in range(int(g["nsteps"]))]
    assert counts == [tsteps.count(t) for t in range(1, int(g["tm_steps"]) + 1)]
    U, V, S, z = ctx.state()
    us = np.abs(g["final_U"]).max() + np.abs(g["final_V"]).max()
    eu, ev = _rel(U, g["final_U"], us), _rel(V, g["final_V"], us)
    es = _rel(S, g["final_S"], np.abs(g["final_S"]).max())
    ez = _rel(z, g["final_z"], np.abs(g["final_S"]).max())
    print(f"    {sum(counts)} Newton steps; final state vs reference: u {eu:.1e}, v {ev:.1e}, "
          f"S {es:.1e}, z {ez:.1e}")
    assert max(eu, ev, es, ez) <= 1e-13
    nf, _ = ctx.wetdry()
    np.testing.assert_array_equal(nf & (64 | 256), g["final_flag"].astype(np.int64) & (64 | 256))
    np.testing.assert_array_equal(ctx.numbering(),
                                  np.stack([g["final_eqU"], g["final_eqV"], g["final_eqS"]], 1))
