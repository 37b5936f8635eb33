"""Generate the golden vectors of tests/golden/*.npz from the reference program.

Runs only in the oracle container (the reference sources must be present under
/root/reference/sources).  It builds oracle/_ref/rismo_harness (oracle/Makefile: the
reference's own objects, unmodified, plus oracle/ref_harness.cpp), writes each deck in
the reference's input format, runs the reference and stores what the harness recorded:

* ``tri20x4.npz``   6-node triangles, 20 x 4 cells: every Newton step's node state,
                    numbering, CSR pattern, assembled FP32 matrix and RHS, scale and
                    correction; step 0's element matrices, scaled system, ILU(0) factors,
                    ILU apply and MulVec of the scaled RHS; the final state.
* ``quad10x3.npz``  8-node quads, 10 x 3 cells, same records.
* ``veg20x4.npz``   the 20 x 4 triangle channel with emergent vegetation (Lindner) on
                    60 % of the elements in the outer halves of the width, same records.
* ``turb20x4.npz``  the 20 x 4 triangle channel with the algebraic eddy viscosity model
                    ($TM_TURBULENCE 2, st uniform in [0.5, 1.5] per element, KDest 0.6),
                    same records plus the nodal vt, uu, uv, vv of every step.
* ``turbq10x3.npz`` the 10 x 3 quad channel with the algebraic model and the element's vt
                    0.04 as lower bound ($TM_TURBULENCE 102), KDest in [0.4, 0.8].
* ``unst20x4.npz``  the 20 x 4 triangle channel as an unsteady flow: 3 time steps of
                    30 s with theta 0.7 from the uniform initial state (a transient);
                    every Newton step's records with the time derivatives and previous
                    time step's state, element matrices of Newton step 1 (non-zero time
                    derivatives).
* ``unstq10x3.npz`` the 10 x 3 quad channel, unsteady with the algebraic eddy viscosity
                    (3 steps of 20 s, theta 0.6; element matrices of Newton step 2).
* ``hydro20x4.npz`` the 20 x 4 triangle channel, unsteady (4 steps of 30 s, theta 0.7)
                    under an inlet hydrograph (QTINLET: Q 10 -> 16 -> 8 m3/s over 120 s);
                    element matrices of Newton step 5 (second time step).
* ``elder20x4.npz`` the 20 x 4 triangle channel with the anisotropic (Elder) eddy viscosity
                    on the constant model, recomputed at every Newton step so that the
                    tensor turns with the developing flow ($TM_TURBULENCE 211; estx in
                    [0.4, 1.2], esty in [0.05, 0.3] per element, KDest 0.6), same records.
* ``elderq10x3.npz`` the 10 x 3 quad channel, Elder on the algebraic model with the
                    element's vt as lower bound ($TM_TURBULENCE 112).
* ``channel200x20.npz``  the straight-channel workload (200 x 20 cells, 8,000 triangles,
                    bed noise sigma 0.05 m / 20 m, seed 0): numbering, per-step scale and
                    correction, BiCGStab iteration counts from the report, final state,
                    and a SHA-256 of the step-0 pattern.
* ``vegchannel200x20.npz``  the same channel with emergent vegetation (Lindner) on 30 % of
                    the elements of the outer 30 % of the width on either bank (stem
                    diameter U[0.01, 0.05] m, density U[0.5, 5] per m2), same records.
* ``meander40x64.npz``  a 40 x 64-cell reach of the meandering-channel generator (5,120
                    triangles, 125 m of the 10 km valley, the full 900 m section): numbering,
                    pattern SHA-256, the first two Newton steps' scale and correction and the
                    state after them (without drying and rewetting the floodplains'
                    steady flow does not converge, so the golden stops at the two steps the
                    reference's BiCGStab solves to its tolerance).
* ``flood200x6.npz``  a 2 km channel (200 x 6 cells) under the flood hydrograph 100 ->
                    1500 m3/s over 6,000 s, its first 10 implicit time steps of 30 s (theta
                    0.7), anisotropic (Elder) eddy viscosity recomputed every Newton step: the
                    Newton steps per time step and the state at the end of every time step.
* ``relax3_20x4.npz`` / ``relax3u20x4.npz``  the 20 x 4 triangle channel with relaxation
  method 3 (time-relaxed Newton steps), stationary (time step 100 s, relaxed time step from
  5 s) and unsteady (3 x 30 s, theta 0.7, relaxed time step from 2 s)
* ``wetdryu20x6.npz`` the bank of ``wetdry20x6`` as an unsteady flow under a rising inlet
                    hydrograph (4 steps of 30 s, theta 0.7), drying and rewetting every 3
                    Newton steps of every time step: every Newton step's records with the
                    previous time step's dry and marsh flags, time derivatives and previous
                    state.
* ``flooddry200x6.npz`` ``flood200x6`` with a bank rising 2.6 m across the outer 40 m of
                    the width, drying and rewetting every 3 Newton steps of every time step
                    (5 time steps; the Newton limit of 12 ends each, the marsh strip keeps
                    the iteration from converging): every Newton step's correction, the
                    final state and the final dry / marsh flags.
* ``cap20x4.npz``   the 20 x 4 triangle channel with a binding relaxation cap ($RELAX 2
                    with maxDeltaUV 0.02, maxDeltaS 0.005): every Newton step's records.
* ``marshmid20x6.npz`` a flat channel (20 x 6 cells, no bed noise) whose corner row at
                    y = 20 m lies half the dry limit (0.05 m) below the still water level and
                    whose rows above are dry land: the first drying call lowers that row's
                    marsh corners by half the dry limit, so the midside nodes beside them
                    sit a quarter or half of the dry limit below their original bed, inside
                    the band (dryLimit / 10, dryLimit) where the midside marsh flag is set
                    only by the reference's dryLimit / 10 test: every Newton step's records.
* ``wetdryq10x3.npz`` the bank of ``wetdry20x6`` on 10 x 3 eight-node quadrilaterals (the
                    bank rising 4.5 m from y = 10 m): the far third dry, the middle third
                    marsh, drying every 3 Newton steps: every Newton step's records.
* ``shelfu20x6.npz`` the flat channel with a shelf 0.1 m below the outlet level (dry land
                    above it), filling from 1.86 m as an unsteady flow (4 x 30 s, theta 0.7,
                    drying every 3 Newton steps of every time step): the shelf, dry at the
                    end of the first time step, gets wet during the second, where Timegrad
                    zeroes the time derivatives of nodes dry or marsh at the previous time
                    step: every Newton step's records.

The reference ships no test vectors of its own; these are the only golden data.
Usage:  python tests/golden/make_golden.py [deck names]
"""
from __future__ import annotations

import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from rismo2d_amd import deck as dk  # noqa: E402
from refdump import read_dump  # noqa: E402

REF = "/root/reference/sources"

DECKS = {
    "tri20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                    Q=10.0),
    "quad10x3": dict(nx=10, ny=3, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                     Q=10.0, quads=True),
    "veg20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                    Q=10.0, veg_frac=0.6, veg_band=0.5),
    "channel200x20": dict(nx=200, ny=20),
    "turb20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                     Q=10.0, turb=2, turb_st=(0.5, 1.5), turb_kdest=0.6),
    "turbq10x3": dict(nx=10, ny=3, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                      Q=10.0, quads=True, vt=0.04, turb=102, turb_st=(0.5, 1.5),
                      turb_kdest=(0.4, 0.8)),
    "unst20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                     Q=10.0),
    "unstq10x3": dict(nx=10, ny=3, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                      Q=10.0, quads=True, turb=2, turb_st=(0.5, 1.5)),
    "hydro20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                      Q=10.0),
    "elder20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                      Q=10.0, turb=211, turb_estx=(0.4, 1.2), turb_esty=(0.05, 0.3)),
    "elderq10x3": dict(nx=10, ny=3, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                       Q=10.0, quads=True, vt=0.04, turb=112, turb_st=(0.5, 1.5),
                       turb_kdest=(0.4, 0.8), turb_estx=(0.4, 1.2), turb_esty=(0.05, 0.3)),
    "wetdry20x6": dict(nx=20, ny=6, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                       Q=10.0),
    "vegchannel200x20": dict(nx=200, ny=20, veg_frac=0.3, veg_band=0.3),
    "meander40x64": dict(gen="meander", n_along=40, n_across=64, reach=125.0),
    "flooddry200x6": dict(nx=200, ny=6, length=2000.0, turb=211, turb_estx=(0.4, 1.2),
                          turb_esty=(0.05, 0.3)),
    "flood200x6": dict(nx=200, ny=6, length=2000.0, turb=211, turb_estx=(0.4, 1.2),
                       turb_esty=(0.05, 0.3)),
    "cap20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                    Q=10.0),
    "rewet20x6": dict(nx=20, ny=6, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                      Q=10.0),
    "wetdryu20x6": dict(nx=20, ny=6, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                        Q=10.0),
    "relax3_20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                        Q=10.0),
    "relax3u20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                        Q=10.0),
    "restart20x4": dict(nx=20, ny=4, length=100.0, width=20.0, noise_sigma=0.05, noise_corr=5.0,
                        Q=10.0),
    "marshmid20x6": dict(nx=20, ny=6, length=100.0, width=30.0, noise_sigma=0.0, Q=10.0),
    "shelfu20x6": dict(nx=20, ny=6, length=100.0, width=30.0, noise_sigma=0.0, Q=10.0),
    "wetdryq10x3": dict(nx=10, ny=3, length=100.0, width=30.0, noise_sigma=0.05, noise_corr=5.0,
                        Q=10.0, quads=True, vt=0.04),
}

# corner rows set to a bed level relative to the still water level S0: (y of the row, bed
# above S0); rows beyond the last listed one keep the last offset, rows before the first
# keep the generator's bed.  Midside nodes are the mean of their corners.
SHELF = {
    "marshmid20x6": [(20.0, -0.025), (25.0, 0.5)],
    "shelfu20x6": [(20.0, -0.1), (25.0, 0.5)],
}

# a bank that rises out of the water: the bed goes up linearly across the width from y0 to
# rise m at y1 (lattice lines, so that midside nodes stay the mean of their corners); with
# drying and rewetting every few Newton steps the far bank becomes dry, the strip beside it
# marsh, and the wet/dry line a slip boundary
BANKS = {
    "wetdry20x6": (15.0, 30.0, 4.5),
    "rewet20x6": (15.0, 30.0, 2.4),
    "wetdryu20x6": (15.0, 30.0, 4.5),
    "flooddry200x6": (60.0, 100.0, 2.6),
    "wetdryq10x3": (10.0, 30.0, 4.5),
}

# initial water level below the outlet's: the level rises over the Newton steps and the
# marsh strip and the dry bank above it get wet
INIT_S = {
    "rewet20x6": 1.5,
    "shelfu20x6": 1.86,
}

# inlet hydrographs (QTINLET), rows (t s, Q m3/s, S m)
HYDRO = {
    "hydro20x4": [[0.0, 10.0, 2.10], [60.0, 16.0, 2.16], [120.0, 8.0, 2.04]],
    "flood200x6": [[0.0, 100.0, 2.2], [6000.0, 1500.0, 2.2]],
    "flooddry200x6": [[0.0, 100.0, 2.2], [6000.0, 1500.0, 2.2]],
    "wetdryu20x6": [[0.0, 10.0, 2.10], [120.0, 30.0, 2.40]],
}

# decks kept to compact records (no per-step matrices or element records); "steps": the
# per-step corrections, "timesteps": the committed state of every time step instead
COMPACT = {"channel200x20": "steps", "vegchannel200x20": "steps", "meander40x64": "steps",
           "flood200x6": "timesteps", "flooddry200x6": "steps", "relax3_20x4": "steps", "relax3u20x4": "steps",
           "restart20x4": "steps"}

# time integration of the decks that are not one stationary flow step, and the Newton step
# whose element matrices are recorded (one with non-zero time derivatives)
RUNS = {
    "unst20x4": dict(tm_steps=3, dt=30.0, theta=0.7, stationary=0, elem_step=1),
    "unstq10x3": dict(tm_steps=3, dt=20.0, theta=0.6, stationary=0, elem_step=2),
    "hydro20x4": dict(tm_steps=4, dt=30.0, theta=0.7, stationary=0, elem_step=5),
    "wetdry20x6": dict(dryrew="2 3 0.001 0.005 100 5", elem_step=1),
    "rewet20x6": dict(dryrew="2 2 0.001 0.005 100 1", maxiter=5, elem_step=3),
    # the bank of wetdry20x6 under a rising hydrograph as an unsteady flow: drying and
    # rewetting inside every time step's cycle, nodes dry or marsh at the previous time
    # step (Predict, Timegrad)
    "wetdryu20x6": dict(dryrew="2 3 0.001 0.005 100 1", maxiter=8, tm_steps=4, dt=30.0,
                        theta=0.7, stationary=0, elem_step=6),
    "meander40x64": dict(maxiter=2, solver="1 6 1 0 60 2000 1.0e-10"),
    "flood200x6": dict(tm_steps=10, dt=30.0, theta=0.7, stationary=0),
    # the flood with a bank rising out of the water along one side: drying and rewetting
    # every 3 Newton steps of every time step, with the hydrograph and Elder's viscosity
    "flooddry200x6": dict(tm_steps=5, dt=30.0, theta=0.7, stationary=0,
                          dryrew="2 3 0.001 0.005 100 5"),
    "cap20x4": dict(relax="2 0.001 0.1 0.02 0.005 0.1", maxiter=40),
    # relaxation method 3 (the reference's default): time-relaxed Newton steps whose relaxed
    # time step grows from relaxTimeFlow to the time step; stationary, and unsteady
    "relax3_20x4": dict(relax="3 0.001 0.1 0.05 0.02 0.1", maxiter=40, dt=100.0, relax_time=5.0),
    "relax3u20x4": dict(relax="3 0.001 0.1 0.02 0.01 0.1", maxiter=60, tm_steps=3, dt=30.0,
                        theta=0.7, stationary=0, relax_time=2.0, elem_step=1),
    # a Newton limit the relaxed time step cannot reach the time step in: the cycle restarts
    # with the relaxed time step as the step's length, the rest of the interval follows
    "restart20x4": dict(relax="3 0.001 0.1 0.005 0.002 0.1", maxiter=4, tm_steps=2, dt=30.0,
                        theta=0.7, stationary=0, relax_time=0.5),
    "marshmid20x6": dict(dryrew="2 3 0.05 0.1 100 5", elem_step=1),
    "shelfu20x6": dict(dryrew="2 3 0.05 0.1 100 1", maxiter=8, tm_steps=4, dt=30.0, theta=0.7,
                       stationary=0, elem_step=6),
    "wetdryq10x3": dict(dryrew="2 3 0.001 0.005 100 5", elem_step=1),
}


def run_reference(deck, workdir, run):
    run = dict(run)
    elem_step = run.pop("elem_step", 0)
    dk.write_reference_deck(deck, workdir, **run)
    env = dict(os.environ, RMX_DUMP=os.path.join(workdir, "dump.bin"),
               RMX_DUMP_ELEM_STEP=str(elem_step), RMX_DUMP_ILU_STEP="0")
    exe = os.path.join(ROOT, "oracle", "_ref", "rismo_harness")
    subprocess.run([exe, "input.dat"], cwd=workdir, env=env, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
    report = open(os.path.join(workdir, "report.txt")).read()
    iters = [int(v) for v in re.findall(r"P_BCGSTABD::Iterate\)\s+(\d+)\. Iteration", report)]
    return read_dump(os.path.join(workdir, "dump.bin")), iters


def deck_arrays(d):
    out = dict(x=d.x, y=d.y, z=d.z, elems=d.elems, nnd=d.nnd, manning=d.manning, vt=d.vt,
               inlet=d.lines[dk.INLET_LINE], outlet=d.lines[dk.OUTLET_LINE], U0=d.U0, V0=d.V0,
               S0=d.S0, Q=d.Q_inlet, S_in=d.S_inlet, S_out=d.S_outlet)
    if d.veg_dp is not None:
        out["veg_dp"] = d.veg_dp
        out["veg_sp"] = d.veg_sp
    if d.hydrograph is not None:
        out["hydrograph"] = d.hydrograph
    if d.turb != 1:
        out["turb"] = d.turb
        for key in ("turb_st", "turb_kdest", "turb_estx", "turb_esty"):
            if getattr(d, key) is not None:
                out[key] = getattr(d, key)
    return out


def state_arrays(prefix, st):
    out = {}
    for k in ("U", "V", "S", "z", "cf", "kind", "rot", "bcU", "bcS", "r00", "r01", "r10", "r11",
              "eqU", "eqV", "eqS", "nx", "ny", "niox", "nioy", "vt", "uu", "uv", "vv",
              "dUdt", "dVdt", "dSdt", "Uo", "Vo", "So", "dUdto", "dVdto", "dSdto"):
        if k in st:
            out[f"{prefix}{k}"] = st[k]
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF}"],
                   check=True)
    names = sys.argv[1:] or list(DECKS)
    for name in names:
        kw = dict(DECKS[name])
        run = RUNS.get(name, {})
        d = dk.meander(**kw) if kw.pop("gen", "channel") == "meander" else dk.channel(**kw)
        if name in HYDRO:
            d.hydrograph = np.asarray(HYDRO[name], np.float64)
        if name in BANKS:
            y0, y1, rise = BANKS[name]
            z = d.z + rise * np.clip((d.y - y0) / (y1 - y0), 0.0, 1.0)
            # midside nodes: the mean of their corners, as the region reader sets them
            for e in range(d.n_elems):
                nnd = int(d.nnd[e])
                ncn = nnd // 2
                for i in range(ncn, nnd):
                    left, right = d.elems[e, i - ncn], d.elems[e, (i - ncn + 1) % ncn]
                    z[d.elems[e, i]] = 0.5 * (z[left] + z[right])
            d.z = z
        if name in SHELF:
            z = d.z.copy()
            rows = SHELF[name]
            for k, (y_row, dz) in enumerate(rows):
                y_next = rows[k + 1][0] if k + 1 < len(rows) else np.inf
                sel = (d.y >= y_row - 1e-9) & (d.y < y_next - 1e-9)
                z[sel] = d.S0[sel] + dz
            for e in range(d.n_elems):
                nnd = int(d.nnd[e])
                ncn = nnd // 2
                for i in range(ncn, nnd):
                    left, right = d.elems[e, i - ncn], d.elems[e, (i - ncn + 1) % ncn]
                    z[d.elems[e, i]] = 0.5 * (z[left] + z[right])
            d.z = z
        if name in INIT_S:
            d.S0 = np.full(d.n_nodes, INIT_S[name])
        with tempfile.TemporaryDirectory() as tmp:
            (steps, final), iters = run_reference(d, tmp, run)
        out = deck_arrays(d)
        out["nsteps"] = len(steps)
        for key in ("tm_steps", "dt", "theta", "stationary", "elem_step", "maxiter", "relax_time"):
            if key in run:
                out[key] = run[key]
        if "relax" in run:
            # $RELAX method relaxMin relaxMax maxDeltaUV maxDeltaS maxDeltaKD
            out["relax"] = np.asarray([float(v) for v in run["relax"].split()])
        if "dryrew" in run:
            # $DRYREW method freq dryLimit rewetLimit rewetPasses countDown
            out["dryrew"] = np.asarray([float(v) for v in run["dryrew"].split()])
        out["bicgstab_iters"] = np.asarray(iters, np.int64)
        full = name not in COMPACT
        for k, st in enumerate(steps):
            p = f"s{k}_"
            out[p + "scale"] = st["scale"]
            if COMPACT.get(name) != "timesteps":
                out[p + "xsol"] = st["xsol"]
            out[p + "tstep"] = st["tstep"]
            if COMPACT.get(name) == "timesteps" and k > 0 and st["tstep"] != steps[k - 1]["tstep"]:
                # the previous time step's committed state, as the next one starts
                t = int(steps[k - 1]["tstep"])
                for v in ("U", "V", "S"):
                    out[f"t{t}_{v}"] = st["state"][v + "o"]
            if full:
                out.update(state_arrays(p, st["state"]))
                if "dryrew" in run:
                    out[p + "flag"] = st["state"]["flag"]
                out[p + "width"] = st["width"]
                out[p + "cols"] = st["cols"]
                out[p + "aval"] = st["aval"]
                out[p + "bvec"] = st["bvec"]
            elif k == 0:
                out.update(state_arrays(p, st["state"]))
                h = hashlib.sha256(st["width"].tobytes() + st["cols"].tobytes()).hexdigest()
                out["s0_pattern_sha256"] = h
                out["s0_nnz"] = int(st["nnz"])
                out["s0_bvec"] = st["bvec"]
        s0 = steps[0]
        se = steps[run.get("elem_step", 0)]
        if full:
            ne = len(se["elems"])
            out["el_names"] = np.stack([np.pad(e["names"], (0, 8 - e["nnd"]), constant_values=0)
                                        for e in se["elems"]])
            out["el_nnd"] = np.asarray([e["nnd"] for e in se["elems"]])
            out["el_bound"] = np.asarray([e["bound"] for e in se["elems"]])
            K = np.zeros((ne, 24, 24))
            F = np.zeros((ne, 24))
            for i, e in enumerate(se["elems"]):
                nd = 3 * e["nnd"]
                K[i, :nd, :nd] = e["estifm"]
                F[i, :nd] = e["force"]
            out["el_estifm"] = K
            out["el_force"] = F
            for k in ("ascaled", "bscaled", "ilu", "iluapply", "mulvec"):
                out["s0_" + k] = s0[k]
        out.update(state_arrays("final_", final))
        if "dryrew" in run:
            out["final_flag"] = final["flag"]
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {len(steps)} Newton steps, BiCGStab iterations {iters}, "
              f"{os.path.getsize(path) / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
