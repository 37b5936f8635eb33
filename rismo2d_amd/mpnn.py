"""Host side of the learned preconditioner: weight initialisation, harvest of training
samples from the solver, and the training driver.  The network is evaluated and trained by
the device kernels of librmx.so (csrc/mpnn.hip); this module only prepares inputs and
calls the C ABI.

A training sample comes from one Newton system (FP64 matrix A, scaled right-hand side b):
the residual the two-level preconditioner M2 leaves on a vector v, r = v - A M2^-1 v, for
the vectors FGMRES feeds its preconditioner: the Arnoldi vectors v_1 = b / ||b||, ...
of the M2-preconditioned operator.  Training minimises ||r - A n(r)||^2 / ||r||^2.  Samples from several bathymetry seeds of one mesh
generator share the mesh graph, so one context holds them all.
"""
from __future__ import annotations

import numpy as np

from . import rmx

C, L, F, O = 64, 6, 8, 18


def shapes():
    """Parameter blocks in the order of the flat weight vector."""
    s = [("W_in", (F, C)), ("b_in", (C,))]
    for layer in range(L):
        s += [(f"W{layer}", (2 * C, C)), (f"b{layer}", (C,))]
    return s + [("W_out", (C, O)), ("b_out", (O,))]


def init_weights(seed: int = 0) -> np.ndarray:
    """Uniform(-1, 1) * sqrt(3 / fan_in) weights (layer weights halved for the residual
    stack), zero biases and a zero output layer, so training starts from "no correction";
    float32."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, sh in shapes():
        m = int(np.prod(sh))
        if name.startswith("b") or name == "W_out":
            parts.append(np.zeros(m))
            continue
        scale = np.sqrt(3.0 / sh[0])
        if name[1:].isdigit():
            scale *= 0.5
        parts.append(rng.uniform(-1.0, 1.0, m) * scale)
    theta = np.concatenate(parts).astype(np.float32)
    assert len(theta) == rmx.Context.mpnn_n_params()
    return theta


def harvest(ctx: "rmx.Context", deck, newton_steps: int = 2, per_system: int = 3,
            stride: int = 4, capacity: int = 0, dry: bool = False) -> int:
    """Run `newton_steps` Newton steps of `deck` on `ctx` (FGMRES with the two-level
    preconditioner in FP64 refinement) and store `per_system` samples of each Newton
    system, from every `stride`-th of its first Arnoldi vectors.  capacity > 0 clears the
    context's training set and sizes it first.  With `dry`, the drying and rewetting runs
    once before the first Newton step, as the flow cycle's first call does (the wet set then
    depends on the deck's bathymetry, and a context's training set holds one wet set: pass
    a capacity for every such deck).  A Newton step whose solve does not reach the FP64
    residual 1e-8 ends the harvest of the deck after its system's samples.  Returns the
    number of stored samples."""
    ctx.upload_mesh(deck)
    ctx.set_params()
    ctx.prepare()
    if dry:
        ctx.dry_rewet()
    count = 0
    for k in range(newton_steps):
        if k:
            ctx.prepare()
        ctx.assemble(rmx.F64)
        ctx.scale_l2()
        ctx.convert(rmx.F32)
        ctx.silu0_factor(rmx.F32)
        ctx.coarse_factor(rmx.F32)
        _, b = ctx.matrix(rmx.F64)
        basis = [b / np.linalg.norm(b)]
        for j in range(per_system * stride):
            v = basis[-1]
            z2 = ctx.precond_apply(rmx.PREC_SILU0C, v, rmx.F32)
            w = ctx.spmv(z2, rmx.F64)
            if j % stride == 0:
                count = ctx.mpnn_add_sample(v - w, capacity)
                capacity = 0
            for u in basis:                      # modified Gram-Schmidt
                w = w - (u @ w) * u
            basis.append(w / np.linalg.norm(w))
        st = ctx.newton_step(method=rmx.FGMRES, prec=rmx.F32, precond=rmx.PREC_SILU0C, eps=1e-3,
                             refine=True, refine_tol=1e-8, refine_max=20)
        if not st.solve.converged:
            # the next state would not be a Newton iterate: keep what is stored and stop
            break
    return count


def train(ctx: "rmx.Context", epochs: int = 50, lr: float = 1e-3) -> float:
    """Adam over the stored systems (one step per system on the mean loss of its samples);
    returns the mean sample loss of the last epoch."""
    return ctx.mpnn_train(epochs, lr)


def meander_deck(seed: int, n_along: int, plain_rise: float, n_across: int = 320):
    """A reach of the meander generator at its full resolution (3.125 m along, 320 cells
    across): `n_along` cells, seeded bathymetry, jitter and friction."""
    from . import deck as dk
    return dk.meander(n_along=n_along, n_across=n_across, reach=10000.0 * n_along / 3200,
                      seed=seed, plain_rise=plain_rise)


def train_meander(seeds, n_along: int = 320, plain_rise: float = 0.012, newton_steps: int = 3,
                  per_system: int = 16, epochs: int = 16, lr: float = 1e-3, theta=None,
                  log=print):
    """Train on reaches of the meander generator, one seed after another: per seed, harvest
    `per_system` samples from each of `newton_steps` Newton systems (after the first drying
    and rewetting call, whose wet set is the seed's own), then `epochs` passes of Adam over
    that seed's systems.  The weights and Adam's moments carry over from seed to seed.
    Returns (weights, per-seed records)."""
    import time
    ctx = rmx.Context(0)
    ctx.mpnn_set_weights(init_weights(0) if theta is None else theta)
    recs = []
    try:
        for s in seeds:
            t0 = time.perf_counter()
            d = meander_deck(s, n_along, plain_rise)
            try:
                k = harvest(ctx, d, newton_steps=newton_steps, per_system=per_system, stride=1,
                            capacity=newton_steps * per_system, dry=plain_rise > 0.0)
            except rmx.RmxError as e:
                log(f"seed {s}: skipped ({e})")
                recs.append({"seed": int(s), "skipped": str(e)})
                continue
            t1 = time.perf_counter()
            l0 = float(np.mean([ctx.mpnn_loss_grad(j)[0] for j in range(0, k, per_system)]))
            loss = ctx.mpnn_train(epochs, lr)
            t2 = time.perf_counter()
            recs.append({"seed": int(s), "elements": int(d.n_elems), "equations": int(ctx.n),
                         "samples": int(k), "loss_first_samples_before": l0,
                         "loss_last_epoch": float(loss), "harvest_s": t1 - t0,
                         "train_s": t2 - t1})
            log(f"seed {s}: {d.n_elems} elements, {ctx.n} equations, {k} samples; loss of the "
                f"systems' first samples {l0:.4f} before, last epoch {loss:.4f} "
                f"(harvest {t1 - t0:.1f}s, train {t2 - t1:.1f}s)")
        return ctx.mpnn_get_weights(), recs
    finally:
        ctx.close()


def main(argv=None):
    """python -m rismo2d_amd.mpnn --out W.npy [--seeds A B] [--n-along N] ...: train on the
    meander reaches of seeds A..B-1 and write the weights (float32, numpy format)."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--seeds", type=int, nargs=2, default=(1, 65))
    ap.add_argument("--n-along", type=int, default=320)
    ap.add_argument("--plain-rise", type=float, default=0.012)
    ap.add_argument("--newton-steps", type=int, default=3)
    ap.add_argument("--per-system", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="weights to start from (default: the "
                                                  "initialisation of seed 0)")
    ap.add_argument("--record", default=None, help="per-seed records (JSON)")
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[mpnn] {msg}", flush=True)

    theta, recs = train_meander(range(*a.seeds), n_along=a.n_along, plain_rise=a.plain_rise,
                                newton_steps=a.newton_steps, per_system=a.per_system,
                                epochs=a.epochs, lr=a.lr, log=log,
                                theta=None if a.init is None else np.load(a.init))
    np.save(a.out, theta)
    if a.record:
        with open(a.record, "w") as f:
            json.dump({"args": vars(a), "seeds": recs}, f, indent=1)
    log(f"weights written to {a.out}")


if __name__ == "__main__":
    main()
