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
            stride: int = 4, capacity: int = 0) -> int:
    """Run `newton_steps` Newton steps of `deck` on `ctx` (FGMRES with the two-level
    preconditioner in FP64 refinement) and store `per_system` samples of each Newton
    system, from every `stride`-th of its first Arnoldi vectors.  capacity > 0 clears the
    context's training set and sizes it first.  Returns the number of stored samples."""
    ctx.upload_mesh(deck)
    ctx.set_params()
    ctx.prepare()
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
            raise rmx.RmxError(-1, f"harvest: Newton step {k} did not converge")
    return count


def train(ctx: "rmx.Context", epochs: int = 50, lr: float = 1e-3) -> float:
    """Adam over the stored pairs; returns the mean loss of the last epoch."""
    return ctx.mpnn_train(epochs, lr)
