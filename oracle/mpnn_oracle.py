"""CPU restatement of the learned preconditioner network (test infrastructure only).

Nothing in the GPU library calls this module.  It restates, in float64 numpy, the network
that rismo2d_amd/csrc/mpnn.hip evaluates and trains on the GPU, so that the GPU forward
pass and gradients can be checked against it.  Its own gradient is pinned by central
finite differences (tests/test_mpnn_oracle.py).  The network has no counterpart in the
reference; the architecture follows the learned-preconditioner specification of the
project (6 message-passing layers, 64 channels).

Graph: the mesh nodes, edges between nodes that share an element (both directions, no
self loops), given as CSR (adj_ptr, adj).

    h0 = relu(f W_in + b_in)                      f: [np, 8] node features
    for l in 1..6:
        a  = mean_{j in N(i)} h_j                 (0 for an isolated node)
        p  = [h, a] W_l + b_l                     W_l: [128, 64]
        h  = h + relu(p)
    out = h W_out + b_out                          [np, 18]: 3x3 blocks D_i, E_i

The node features describe the state only; the correction of a residual r is linear in r:
q = r / max_j |a_ij| (row-scaled), q_i its (u, v, h) components at node i (0 where a
variable has no equation), qbar_i their mean over the neighbours, and
    n(r)_i = D_i q_i + E_i qbar_i        (at the equations of node i).

With bf16=True, h (after every layer), a, the layer weights W_l and the layer gradients
g_p are rounded to bfloat16 (round to nearest even) where the GPU kernels store them in
bfloat16; sums stay in float64 here (float32 on the GPU).
"""
from __future__ import annotations

import numpy as np

C, L, F, O = 64, 6, 8, 18


def shapes():
    s = [("W_in", (F, C)), ("b_in", (C,))]
    for layer in range(L):
        s += [(f"W{layer}", (2 * C, C)), (f"b{layer}", (C,))]
    s += [("W_out", (C, O)), ("b_out", (O,))]
    return s


def n_params():
    return sum(int(np.prod(sh)) for _, sh in shapes())


def unpack(theta):
    out, k = {}, 0
    for name, sh in shapes():
        m = int(np.prod(sh))
        out[name] = theta[k:k + m].reshape(sh)
        k += m
    return out


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def aggregate(h, adj_ptr, adj):
    """a_i = mean of h over the neighbours of i (0 without neighbours)."""
    n = len(adj_ptr) - 1
    deg = np.diff(adj_ptr)
    rows = np.repeat(np.arange(n), deg)
    a = np.zeros_like(h)
    np.add.at(a, rows, h[adj])
    return a / np.maximum(deg, 1)[:, None]


def aggregate_t(g, adj_ptr, adj):
    """Transpose of aggregate: out_j = sum over i with j in N(i) of g_i / deg_i."""
    n = len(adj_ptr) - 1
    deg = np.diff(adj_ptr)
    rows = np.repeat(np.arange(n), deg)
    w = (g / np.maximum(deg, 1)[:, None])[rows]
    out = np.zeros_like(g)
    np.add.at(out, adj, w)
    return out


def forward(theta, feat, adj_ptr, adj, bf=False, keep=False):
    P = unpack(theta)
    rnd = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    q_in = feat @ P["W_in"] + P["b_in"]
    h = rnd(np.maximum(q_in, 0.0))
    cache = {"q_in": q_in, "h": [h], "a": [], "p": []}
    for layer in range(L):
        a = rnd(aggregate(h, adj_ptr, adj))
        W = rnd(P[f"W{layer}"])
        p = h @ W[:C] + a @ W[C:] + P[f"b{layer}"]
        h = rnd(h + np.maximum(p, 0.0))
        cache["a"].append(a)
        cache["p"].append(p)
        cache["h"].append(h)
    out = h @ P["W_out"] + P["b_out"]
    return (out, cache) if keep else out


def backward(theta, feat, cache, g_out, adj_ptr, adj, bf=False):
    """Gradient of a loss with d loss / d out = g_out, through the cached forward pass."""
    P = unpack(theta)
    rnd = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    G = {}
    hL = cache["h"][-1]
    G["W_out"] = hL.T @ g_out
    G["b_out"] = g_out.sum(0)
    g_h = g_out @ P["W_out"].T
    for layer in reversed(range(L)):
        h_in, a, p = cache["h"][layer], cache["a"][layer], cache["p"][layer]
        W = rnd(P[f"W{layer}"])
        g_p = rnd(g_h * (p > 0.0))
        G[f"W{layer}"] = np.concatenate([h_in, a], 1).T @ g_p
        G[f"b{layer}"] = g_p.sum(0)
        g_x = g_p @ W.T
        g_h = g_h + g_x[:, :C] + aggregate_t(g_x[:, C:], adj_ptr, adj)
    g_q = g_h * (cache["q_in"] > 0.0)
    G["W_in"] = feat.T @ g_q
    G["b_in"] = g_q.sum(0)
    return np.concatenate([G[name].reshape(-1) for name, _ in shapes()])


def loss_grad(theta, feat, target, mask, adj_ptr, adj, bf=False):
    """loss = sum(mask * (out - target)^2) / sum(mask) and its gradient."""
    out, cc = forward(theta, feat, adj_ptr, adj, bf=bf, keep=True)
    cnt = max(float(mask.sum()), 1.0)
    d = (out - target) * mask
    loss = float((d * d).sum() / cnt)
    return loss, backward(theta, feat, cc, 2.0 * d / cnt, adj_ptr, adj, bf)


def row_inv(rowptr, vals):
    """1 / max_j |a_ij| per CSR row (0 for an empty or zero row)."""
    n = len(rowptr) - 1
    m = np.zeros(n)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    np.maximum.at(m, rows, np.abs(vals))
    return np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 0.0)


def node_components(x, eq):
    """Equation vector -> [np, 3] (0 where a node variable has no equation)."""
    out = np.zeros(eq.shape)
    m = eq >= 0
    out[m] = x[eq[m]]
    return out


def lin_apply(out, q, eq, adj_ptr, adj):
    """n = D q + E qbar as an equation vector."""
    qn = node_components(q, eq)
    qb = aggregate(qn, adj_ptr, adj)
    D = out[:, :9].reshape(-1, 3, 3)
    E = out[:, 9:].reshape(-1, 3, 3)
    nn = np.einsum("ivw,iw->iv", D, qn) + np.einsum("ivw,iw->iv", E, qb)
    nv = np.zeros(len(q))
    m = eq >= 0
    nv[eq[m]] = nn[m]
    return nv


def resid_loss_grad(theta, feat, r, rowptr, cols, vals, eq, adj_ptr, adj, bf=False):
    """loss = ||r - A n(r)||^2 / ||r||^2 (A in CSR) and its gradient."""
    import scipy.sparse as sp
    n = len(r)
    A = sp.csr_matrix((vals, cols, rowptr), shape=(n, n))
    out, cc = forward(theta, feat, adj_ptr, adj, bf=bf, keep=True)
    q = r * row_inv(rowptr, vals)
    nv = lin_apply(out, q, eq, adj_ptr, adj)
    res = r - A @ nv
    rr = float(r @ r)
    loss = float(res @ res) / rr
    g_n = node_components(-2.0 * (A.T @ res) / rr, eq)
    qn = node_components(q, eq)
    qb = aggregate(qn, adj_ptr, adj)
    g_out = np.concatenate([np.einsum("iv,iw->ivw", g_n, qn).reshape(-1, 9),
                            np.einsum("iv,iw->ivw", g_n, qb).reshape(-1, 9)], 1)
    return loss, backward(theta, feat, cc, g_out, adj_ptr, adj, bf)
