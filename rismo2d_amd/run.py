"""Run a deck given in the reference's input files through the solve path.

The time loop is that of the reference's flow cycle (Compute.cpp:196-796 with
EQS_UVS2D::Execute, EqsUVS2D.cpp:71-1096). For each time step of ``$TM_STEPS``:
- unsteady (``$TM_STATIONARY 0``): Predict with dt and θ; QTINLET takes the inflow at the
  step's end (EqsUVS2D.cpp:318-320);
- Newton steps until the reference's convergence test holds (``nrconv``) or
  ``$TM_MAXITER`` steps, then the slip condition at rotated nodes;
- with ``$DRYREW`` calling it inside the cycle (its frequency below ``$TM_MAXITER``), the
  drying and rewetting before the first Newton step and then every ``dryRewFreq`` steps or
  after a converged one (EqsUVS2D.cpp:200-222): a call that wets elements, or one made while
  the relaxed time steps of methods 3 and 4 are shorter than the time step, holds the
  cycle's convergence until the next call; the call whose next one would fall beyond the
  Newton limit is the last;
- the cycle ends at ``$TM_MAXITER`` Newton steps whether or not they converged: the test
  that ends it (EqsUVS2D.cpp:893) holds at it == actualCycit - 1 before the restart with a
  shorter time step of relaxation methods 3 and 4 (:931-962) is examined, so that restart
  never runs (a golden deck whose relaxed time step stays far below the time step shows the
  reference stopping there);
- the step is committed as the old state of the next one.

A stationary deck solves its flow once.
"""
from __future__ import annotations

import numpy as np

from . import deck as dk
from . import rmx

# the solve of each Newton step: FGMRES(15) on the FP32 matrix with the two-level strip
# preconditioner inside FP64 refinement to 1e-8, or the reference's own plain path
FAST = dict(method=rmx.FGMRES, prec=rmx.F32, precond=rmx.PREC_SILU0C, mkyrl=15, eps=1e-2,
            refine=True, refine_tol=1e-8, refine_max=8, max_iter=1000)
PLAIN = dict(method=rmx.BICGSTAB, prec=rmx.F64, precond=rmx.PREC_ILU0, eps=1e-12,
             max_iter=1000)


def time_loop(m, steps: int, dt: float, theta: float, maxiter: int, newton, hydrograph=None,
              stationary: bool = False, on_step=None, dry_freq: int | None = None):
    """The flow cycle's time loop on a model ``m`` with the methods ``time_step(dt, theta,
    first)``, ``time_commit()``, ``set_inflow(Q, S)`` and ``slip_flow()``; ``newton()`` runs
    one Newton step and returns its statistics (``nrconv``).  ``dry_freq``: the drying and
    rewetting frequency (None: no drying); below ``maxiter`` ``m`` needs ``dry_rewet()``
    (elements got dry, got wet) and ``relax_times()`` (dt_UV, dt_H, dt).  Returns the
    Newton steps of every time step."""
    counts = []
    in_cycle = dry_freq is not None and dry_freq < maxiter
    for t in range(1, steps + 1):
        if not stationary:
            if t > 1:
                m.time_commit()
            m.time_step(dt, theta, first=(t == 1))
        if hydrograph is not None and not stationary:
            m.set_inflow(*rmx.hydrograph(hydrograph, t * dt))
        k = 0
        nrconv, dwconv, next_dw, maxit = True, True, 0, maxiter
        while True:
            if in_cycle and next_dw >= 0 and (k == next_dw or nrconv):
                next_dw = k + dry_freq
                _, wet = m.dry_rewet()
                dt_uv, dt_h, tdt = m.relax_times()
                dwconv = not (wet or dt_h < tdt or dt_uv < tdt)
                if next_dw > maxit:          # no further call in this cycle
                    maxit, next_dw, dwconv = next_dw, -1, True
            st = newton()
            k += 1
            nrconv = bool(st.nrconv)
            if on_step:
                on_step(t, k, st)
            if (nrconv and dwconv) or k >= maxiter:
                break
        # EQS_UVS2D::Execute ends with the slip condition (EqsUVS2D.cpp:1078)
        m.slip_flow()
        counts.append(k)
    return counts


def run_reference_deck(input_file: str, ctx: rmx.Context | None = None, solve: dict | None = None,
                       on_step=None):
    """Read ``input_file`` (read_reference_deck) and run it. ``solve``: the options of each
    Newton step (FAST by default, PLAIN the reference's BiCGStab / ILU(0) in FP64).
    ``on_step(t, k, stats)`` is called after every Newton step. Returns (deck, run settings,
    per time step the Newton step count, final (U, V, S, z))."""
    d, run = dk.read_reference_deck(input_file)
    own = ctx is None
    c = ctx or rmx.Context(0)
    try:
        c.upload_mesh(d)
        c.set_params(**run["params"])
        if solve is None:
            solve = FAST
        unsteady = run["stationary"] == 0
        dw = run["dryrew"]
        counts = time_loop(c, run["steps"] if unsteady else 1, run["dt"], run["theta"],
                           run["maxiter"], lambda: c.newton_step(**solve),
                           hydrograph=d.hydrograph, stationary=not unsteady, on_step=on_step,
                           dry_freq=dw["freq"] if dw["method"] else None)
        U, V, S, z = c.state()
        return d, run, counts, (np.asarray(U), np.asarray(V), np.asarray(S), np.asarray(z))
    finally:
        if own:
            c.close()
