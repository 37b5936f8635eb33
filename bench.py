"""Benchmark: time per Newton step to FP64 relative residual 1e-8 on the 2M-triangle
meandering channel (one MI355X).

One step = one Newton step of the UVS2D system from the deck's initial state (the state is
uploaded again before every step, outside the timed region, so every step solves the same
system): nodal friction and inlet
set-up, FP64 assembly, L2 scaling, FP32 copy of the matrix, ILU(0) factorisation, the
Krylov solve in FP32 inside FP64 iterative refinement to ||b - J x|| / ||b|| <= 1e-8 (FP64
J), Newton update.  The inputs (mesh, pattern, schedules) are resident on the device
before the timed region.

Printed (one JSON line): the step time of the FP32/FP64 path (FGMRES(15) with the
two-level preconditioner: coarse correction + strip ILU(0) sweeps); the plain FP64
ILU0-BiCGStab path on the same system, run for a bounded time: its elapsed time and the
residual it reached give a lower bound on the speedup (its full solve at this size takes
far longer than a benchmark run); the roofline of the dominant kernel (HIP events on the
library's stream); and a CPU baseline: the oracle's FP64 ILU0-BiCGStab Newton step on a
bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "time to FP64 residual 1e-8 per Newton step: learned+FP32 vs plain FP64"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-along", type=int, default=3200)
    ap.add_argument("--n-across", type=int, default=320)
    ap.add_argument("--strips", type=int, default=0)
    ap.add_argument("--segments", type=int, default=32)
    ap.add_argument("--coarse", type=int, default=8, help="coarse segments per strip (0: none)")
    ap.add_argument("--max-iter", type=int, default=1000, help="Krylov iterations per refinement sweep")
    ap.add_argument("--mkyrl", type=int, default=15, help="FGMRES restart length")
    ap.add_argument("--eps", type=float, default=1e-2, help="inner relative tolerance per sweep")
    ap.add_argument("--refine-max", type=int, default=8)
    ap.add_argument("--plain-iters", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling run: build and factor the system, time the FP32 SpMV and "
                         "strip sweeps, run a plain device copy of known size, print their "
                         "times and exit (no Newton step, no JSON benchmark line)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sweep_bytes(entries, rows, vbytes):
    """Algorithmic bytes of one triangular sweep: every kept factor entry (value + column)
    once, per row its entry offset, diagonal (upper sweep), right-hand side and result."""
    return entries * (vbytes + 4) + rows * (4 + 8 + 8)


def spmv_bytes(nnz, n, vbytes):
    """Algorithmic bytes of one SpMV: values + columns once, row pointers, x once, y."""
    return nnz * (vbytes + 4) + 4 * (n + 1) + 8 * n + 8 * n


def cpu_baseline():
    """The reference program itself when its build (oracle/_ref/rismo_harness, compiled from
    the reference's sources) is present, else the oracle; both on one core."""
    exe = os.path.join(HERE, "oracle", "_ref", "rismo_harness")
    if os.path.exists(exe):
        try:
            return cpu_baseline_reference(exe)
        except Exception as e:      # noqa: BLE001 - a failed baseline falls back, never stops
            log(f"reference baseline failed ({e}); timing the oracle instead")
    return cpu_baseline_port()


def _cpu_seconds(field):
    """'h:mm:ss:cc' of the reference's report (REPORT::PrintTime) in seconds."""
    h, m, sec, cc = (int(v) for v in field.split(":"))
    return 3600 * h + 60 * m + sec + cc / 100.0


def cpu_baseline_reference(exe):
    """One Newton step of the reference program on a 60-section reach of the meander
    generator: its own numbering, assembly, ILU(0) and BiCGStab to 1e-8 (at most 2000
    iterations), written as its input files; the value is the CPU time its report gives for
    the time step (the report's last CPU line)."""
    import re
    import subprocess
    import tempfile
    from rismo2d_amd import deck as dk
    na = 60
    d = dk.meander(n_along=na, n_across=320, reach=10000.0 * na / 3200)
    with tempfile.TemporaryDirectory() as wd:
        dk.write_reference_deck(d, wd, solver="1 6 1 0 60 2000 1.0e-8", maxiter=1)
        subprocess.run([exe, "input.dat"], cwd=wd, check=True, timeout=120,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        rep = open(os.path.join(wd, "report.txt")).read()
    cpu = re.findall(r"\bCPU:\s+(\d+:\d+:\d+:\d+)", rep)
    it = re.findall(r"P_BCGSTABD::Iterate\)\s+(\d+)\. Iteration \| accuracy = (\S+)", rep)
    if not cpu or not it:
        raise RuntimeError("no timing or iteration record in the reference's report")
    its, acc = int(it[-1][0]), float(it[-1][1])
    return dict(value=_cpu_seconds(cpu[-1]), unit="s per Newton step", cores=1,
                kind="reference",
                sample=(f"meander {na}x320 reach of the same generator ({d.n_elems} tri), one "
                        f"Newton step of the reference program built from its sources: "
                        f"numbering, assembly, ILU(0), BiCGStab {its} it to {acc:.1e}; CPU time "
                        f"of the step from its report, 1 core"),
                iters=its)


def cpu_baseline_port():
    """The oracle (C restatement of the reference's path, one core) on a bounded sample of
    the same workload: one Newton step, FP64 assembly + ILU(0) + BiCGStab to 1e-8, on a
    100-section reach of the meander generator."""
    from rismo2d_amd import deck as dk
    from oracle import pyoracle as po
    na = 100
    d = dk.meander(n_along=na, n_across=320, reach=10000.0 * na / 3200)
    m = po.OracleModel(d)
    m.set_params()
    t0 = time.perf_counter()
    n = m.prepare()
    rp, cols = m.pattern()
    A, B = m.assemble(64)
    po.scale(rp, A, B)
    L, _ = po.ilu0_factor(rp, cols, A)
    X, conv, its, acc = po.bicgstab(rp, cols, A, L, B, max_iter=2000, eps=1e-8)
    t = time.perf_counter() - t0
    return dict(value=t, unit="s per Newton step", cores=1, kind="port",
                sample=(f"meander {na}x320 reach of the same generator ({d.n_elems} tri, {n} eq), "
                        f"one Newton step with the oracle: FP64 assembly, ILU(0), BiCGStab "
                        f"{its} it to {acc:.1e} (converged {bool(conv)}), 1 core"),
                n_eq=n, iters=its)


def kernels_only(c, n, a):
    """The system of the first Newton step (FP64 assembly, L2 scaling, FP32 copy), its strip
    ILU(0) in FP32, five timed SpMV / sweep pairs, then three plain copies of an n-vector
    (the identity preconditioner's device copy: 8n bytes read, 8n written), the reference
    size for calibrating HBM byte counters."""
    from rismo2d_amd import rmx as R
    c.assemble(R.F64)
    c.scale_l2()
    c.convert(R.F32)
    c.silu0_factor(R.F32)
    kt = c.profile_kernels(R.F32, 5)
    x = np.random.default_rng(0).standard_normal(n)
    for _ in range(3):
        c.precond_apply(R.PREC_NONE, x, R.F32)
    lo_e, up_e = c.sweep_entries()
    print(json.dumps({"n": n, "nnz": c.nnz, "spmv_ms": kt.spmv_s * 1e3, "spmv_bytes": kt.spmv_bytes,
                      "sweep_fwd_ms": kt.silu_fwd_s * 1e3, "sweep_bwd_ms": kt.silu_bwd_s * 1e3,
                      "sweep_fwd_bytes": sweep_bytes(lo_e, n, 4),
                      "sweep_bwd_bytes": sweep_bytes(up_e, n, 4) + 4 * n,
                      "copy_bytes_each_way": 8 * n}), flush=True)
    c.close()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (the multi-GPU "
              "path is not built)", file=sys.stderr)
        sys.exit(2)
    from rismo2d_amd import deck as dk

    # the CPU baseline first: it starts the reference program as a child process, which is
    # done before this process loads the HIP library or touches the GPU
    cpu = None
    if not a.no_cpu and not a.kernels_only:
        try:
            cpu = cpu_baseline()
            log(f"cpu baseline: {cpu['value']:.1f}s ({cpu['sample']})")
        except Exception as e:       # the oracle library is test infrastructure: report its absence
            log(f"cpu baseline unavailable: {e}")

    from rismo2d_amd import rmx as R

    t = time.perf_counter()
    d = dk.meander(n_along=a.n_along, n_across=a.n_across, reach=10000.0 * a.n_along / 3200)
    log(f"deck {a.n_along}x{a.n_across}: {d.n_elems} triangles, {d.n_nodes} nodes "
        f"({time.perf_counter() - t:.1f}s)")
    c = R.Context(0)
    c.upload_mesh(d)
    c.set_params()
    t = time.perf_counter()
    n = c.prepare()
    c.set_coarse(a.coarse)
    c.set_strips(a.strips, a.segments)
    log(f"prepare: {n} equations, {c.nnz} non-zeros ({time.perf_counter() - t:.1f}s)")

    if a.kernels_only:
        kernels_only(c, n, a)
        return
    pk = R.PREC_SILU0C if a.coarse > 0 else R.PREC_SILU0
    opts = dict(method=R.FGMRES, prec=R.F32, precond=pk, max_iter=a.max_iter,
                mkyrl=a.mkyrl, eps=a.eps, refine=True, refine_tol=1e-8, refine_max=a.refine_max)

    def reset():
        # the deck's initial state again (host upload, outside the timed region), so that
        # every step solves the same system
        c.state_upload(d.U0, d.V0, d.S0)

    for w in range(a.warmup):
        reset()
        st = c.newton_step(**opts)
        log(f"warmup {w}: {st.solve.iters} it, residual {st.solve.true_resid:.2e}, "
            f"{st.t_prepare + st.t_assemble + st.t_factor + st.t_solve + st.t_update:.2f}s")
    times, stats = [], []
    for k in range(a.steps):
        reset()
        t0 = time.perf_counter()
        st = c.newton_step(**opts)
        times.append(time.perf_counter() - t0)
        stats.append(st)
        log(f"step {k}: {times[-1]:.2f}s (assemble {st.t_assemble:.3f}, factor {st.t_factor:.3f}, "
            f"solve {st.t_solve:.2f}: {st.solve.iters} it, {st.solve.sweeps} sweeps, residual "
            f"{st.solve.true_resid:.2e})")
    T = sum(times) / len(times)
    last = stats[-1]

    # kernel timing on the last step's system (HIP events on the library stream)
    kt = c.profile_kernels(R.F32, 5)
    lo_e, up_e = c.sweep_entries()
    fwd_bytes = sweep_bytes(lo_e, n, 4)
    bwd_bytes = sweep_bytes(up_e, n, 4) + 4 * n
    # the library reports the algorithmic bytes of the layout the FP32 SpMV ran on
    sp_bytes = kt.spmv_bytes if kt.spmv_bytes > 0 else spmv_bytes(c.nnz, n, 4)
    its = max(last.solve.iters, 1)
    spmv_per_it = 1                               # FGMRES: one product per iteration
    share = {"silu_sweep_fwd": kt.silu_fwd_s * its, "silu_sweep_bwd": kt.silu_bwd_s * its,
             "spmv_f32": kt.spmv_s * (spmv_per_it * its + last.solve.sweeps)}
    dom = max(share, key=share.get)
    dur, byt = {"silu_sweep_fwd": (kt.silu_fwd_s, fwd_bytes),
                "silu_sweep_bwd": (kt.silu_bwd_s, bwd_bytes),
                "spmv_f32": (kt.spmv_s, sp_bytes)}[dom]
    achieved = byt / dur
    # HBM bytes per launch of the dominant kernel from the committed counter summary of the
    # same workload (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, calibrated on a device copy of
    # known size), used when the system and the kernel's algorithmic bytes match
    traffic = None
    tf = os.path.join(HERE, "profiles", "meander2M_hbm_traffic.json")
    if os.path.exists(tf):
        rec = json.load(open(tf))
        k = rec.get("kernels", {}).get(dom)
        if k and rec.get("n") == n and rec.get("nnz") == c.nnz and k.get("algorithmic_bytes") == byt:
            traffic = k["traffic_bytes"]
    log(f"kernels: spmv {kt.spmv_s * 1e3:.3f} ms ({sp_bytes / kt.spmv_s / 1e9:.0f} GB/s), strip "
        f"sweeps {kt.silu_fwd_s * 1e3:.3f} / {kt.silu_bwd_s * 1e3:.3f} ms")

    # plain FP64 path on the same system (state reset, friction and inlet recomputed):
    # reference-order ILU(0), BiCGStab, bounded
    c.state_upload(d.U0, d.V0, d.S0)
    c.prepare()
    c.assemble(R.F64)
    c.scale_l2()
    t0 = time.perf_counter()
    c.ilu0_factor(R.F64)
    t_fac64 = time.perf_counter() - t0
    out = c.solve(R.BICGSTAB, R.F64, R.PREC_ILU0, max_iter=a.plain_iters, eps=1e-8)
    plain_it = out.seconds / max(out.iters, 1)
    plain_elapsed = t_fac64 + out.seconds
    kt64 = c.profile_kernels(R.F64, 2)
    log(f"plain FP64: factor {t_fac64:.2f}s, {out.iters} BiCGStab it in {out.seconds:.2f}s "
        f"({plain_it:.3f} s/it), residual {out.accuracy:.2e}; sweeps {kt64.ilu_fwd_s * 1e3:.1f} / "
        f"{kt64.ilu_bwd_s * 1e3:.1f} ms")
    c.close()

    line = {
        "metric": METRIC,
        "value": T,
        "unit": "s per Newton step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": T * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"meander channel {a.n_along}x{a.n_across} cells, "
                               f"{d.n_elems} Taylor-Hood triangles, {n} equations, {c.nnz} nnz; "
                               "Newton step 1 from the initial state",
                   "solver": f"FGMRES({a.mkyrl}), FP32 matrix, coarse correction "
                             f"({a.coarse} segments/strip x U,V,H aggregates) + strip ILU(0) "
                             f"({a.strips or '2 per CU'} strips x {a.segments} segments), "
                             f"FP64 iterative refinement to 1e-8 (inner tolerance {a.eps:g})",
                   "parallelism": "1 GPU"},
        "converged": bool(last.solve.converged),
        "true_resid": last.solve.true_resid,
        "krylov_iters": last.solve.iters,
        "refine_sweeps": last.solve.sweeps,
        "t_factor_s": last.t_factor,
        "t_solve_s": last.t_solve,
        "plain_fp64": {"factor_s": t_fac64, "s_per_bicgstab_iter": plain_it,
                       "iters_run": out.iters, "residual_reached": out.accuracy,
                       "elapsed_s": plain_elapsed, "converged": bool(out.converged),
                       "sweep_fwd_ms": kt64.ilu_fwd_s * 1e3, "sweep_bwd_ms": kt64.ilu_bwd_s * 1e3},
        # the plain path had not reached 1e-8 when stopped: its time to 1e-8 exceeds
        # elapsed_s, so elapsed_s / value bounds the speedup from below
        "speedup_vs_plain_lower_bound": (plain_elapsed / T) if (last.solve.converged and
                                                                 not out.converged) else None,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "algorithmic_bytes": byt, "avg_launch_ms": dur * 1e3},
        "cpu_baseline": cpu and {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
