"""The learned preconditioner's CPU oracle (oracle/mpnn_oracle.py) is pinned by finite
differences and by the adjoint identity of its aggregation; no reference counterpart."""
import numpy as np

from oracle import mpnn_oracle as mo


def _graph(n, seed):
    """Random symmetric graph without self loops, CSR with sorted neighbours; node n-1
    isolated."""
    rng = np.random.default_rng(seed)
    M = rng.random((n, n)) < 0.3
    M = M | M.T
    np.fill_diagonal(M, False)
    M[n - 1, :] = False
    M[:, n - 1] = False
    ptr = np.concatenate([[0], np.cumsum(M.sum(1))]).astype(np.int32)
    adj = np.concatenate([np.nonzero(M[i])[0] for i in range(n)]).astype(np.int32)
    return ptr, adj


def test_aggregate_transpose_is_adjoint():
    ptr, adj = _graph(15, 1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((15, 4))
    y = rng.standard_normal((15, 4))
    lhs = (mo.aggregate(x, ptr, adj) * y).sum()
    rhs = (x * mo.aggregate_t(y, ptr, adj)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_aggregate_is_neighbour_mean():
    ptr, adj = _graph(9, 3)
    x = np.arange(9.0)[:, None] * np.ones((1, 2))
    a = mo.aggregate(x, ptr, adj)
    for i in range(9):
        nb = adj[ptr[i]:ptr[i + 1]]
        want = x[nb].mean(0) if len(nb) else np.zeros(2)
        np.testing.assert_allclose(a[i], want, rtol=1e-15)


def test_gradient_matches_central_differences():
    n = 14
    ptr, adj = _graph(n, 4)
    rng = np.random.default_rng(5)
    theta = rng.standard_normal(mo.n_params()) * 0.2
    feat = rng.standard_normal((n, mo.F))
    target = rng.standard_normal((n, mo.O))
    mask = (rng.random((n, mo.O)) < 0.8).astype(float)
    loss, grad = mo.loss_grad(theta, feat, target, mask, ptr, adj)
    idx = rng.choice(len(theta), 60, replace=False)
    # include every parameter block
    starts = np.cumsum([0] + [int(np.prod(sh)) for _, sh in mo.shapes()])[:-1]
    idx = np.unique(np.concatenate([idx, starts]))
    h = 1e-6
    for k in idx:
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (mo.loss_grad(tp, feat, target, mask, ptr, adj)[0]
              - mo.loss_grad(tm, feat, target, mask, ptr, adj)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38])
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2.0 ** -7              # above the tie rounds up
    assert r[3] == -2.5
    assert np.isfinite(r[4])


def test_residual_loss_gradient_matches_central_differences():
    n_nodes = 12
    ptr, adj = _graph(n_nodes, 6)
    rng = np.random.default_rng(7)
    eq = -np.ones((n_nodes, 3), np.int64)
    m = rng.random((n_nodes, 3)) < 0.85
    eq[m] = np.arange(int(m.sum()))
    n = int(m.sum())
    dense = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.3) + 4 * np.eye(n)
    rowptr = np.concatenate([[0], np.cumsum((dense != 0).sum(1))])
    cols = np.concatenate([np.nonzero(dense[i])[0] for i in range(n)])
    vals = dense[dense != 0]
    theta = rng.standard_normal(mo.n_params()) * 0.2
    feat = rng.standard_normal((n_nodes, mo.F))
    r = rng.standard_normal(n)
    loss, grad = mo.resid_loss_grad(theta, feat, r, rowptr, cols, vals, eq, ptr, adj)
    idx = np.unique(np.concatenate([rng.choice(len(theta), 40, replace=False),
                                    np.cumsum([0] + [int(np.prod(sh)) for _, sh in mo.shapes()])[:-1]]))
    h = 1e-4          # the loss is ~5e4 here: smaller steps lose digits to cancellation
    for k in idx:
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (mo.resid_loss_grad(tp, feat, r, rowptr, cols, vals, eq, ptr, adj)[0]
              - mo.resid_loss_grad(tm, feat, r, rowptr, cols, vals, eq, ptr, adj)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, grad[k])
