"""GPU parity of the HIP path (through the C ABI) against the oracle and the reference's
golden vectors.

Tolerances (floating point):
* integer work (numbering, CSR pattern, colouring-independent results): bit-exact;
* FP32 ILU(0) factors computed from the reference's own scaled FP32 matrix: bit-exact
  (the factor kernel keeps the reference's per-entry update order, FMA contraction off);
* element matrices, assembled FP64 entries, RHS: |gpu - oracle| <= 1e-12 * max|row| (the
  GPU contracts to FMA and sums element contributions in colour order);
* nodal friction and inlet discharge: 1e-13 relative (1e-10 at nodes with Lindner's
  vegetation iteration, which amplifies the ulp differences of device pow / acos / cos);
* nodal eddy viscosity and Reynolds stress of the algebraic model: 1e-13 relative;
* SpMV, ILU apply (FP64 accumulation in another order): 1e-12 relative to max|result|;
* FP32 SpMV on the row-group layout: 1e-13 of sum |a| |x| per row;
* Newton-converged u, v, h: within 1e-8 relative of the reference / oracle.
"""
import hashlib

import numpy as np
import pytest

from helpers import golden_deck, load_golden
from oracle import pyoracle as po
from rismo2d_amd import deck as dk
from rismo2d_amd import rmx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rmx.Context(0)
    yield c
    c.close()


def _rowmax(rp, A):
    w = np.diff(rp)
    rows = np.repeat(np.arange(len(w)), w)
    m = np.zeros(len(w))
    np.maximum.at(m, rows, np.abs(A))
    return m[rows], rows


@pytest.mark.parametrize("name", ["tri20x4", "quad10x3", "veg20x4", "turb20x4", "turbq10x3",
                                  "elder20x4", "elderq10x3"])
def test_setup_and_element_kernel_match_oracle(ctx, name):
    print(f"[gpu] setup/element parity on {name}", flush=True)
    g = load_golden(name)
    deck = golden_deck(g)
    ctx.upload_mesh(deck)
    ctx.set_params()
    n = ctx.prepare()
    om = po.OracleModel(deck)
    assert om.prepare() == n
    np.testing.assert_array_equal(ctx.numbering(), om.numbering())
    rp, cols = ctx.pattern()
    orp, ocols = om.pattern()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(cols, ocols)
    nod, onod = ctx.nodal(), om.nodal()
    np.testing.assert_array_equal(nod["kind"], onod["kind"])
    # Manning: 1e-13.  Nodes touching vegetated elements: 1e-10 (Lindner's drag iteration
    # amplifies the ulp differences of device pow / acos / cos against the host libm; a
    # flipped convergence decision would show as ~1e-3)
    veg = np.zeros(deck.n_nodes, bool)
    if deck.veg_dp is not None:
        for e in np.nonzero(deck.veg_dp > 0)[0]:
            veg[deck.elems[e, :deck.nnd[e]]] = True
    np.testing.assert_allclose(nod["cf"][~veg], onod["cf"][~veg], rtol=1e-13, atol=0)
    np.testing.assert_allclose(nod["cf"][veg], onod["cf"][veg], rtol=1e-10, atol=0)
    np.testing.assert_allclose(nod["bcU"], onod["bcU"], rtol=1e-13, atol=0)
    np.testing.assert_array_equal(nod["rotm"], onod["rotm"])
    if deck.turb % 10 == 2:
        # algebraic model: the oracle's nodal vt, uu are bit-exact with the reference's
        vt, uu = ctx.turbulence()
        ovt, ouu = om.turbulence_state()
        np.testing.assert_array_equal(ovt, g["s0_vt"])
        np.testing.assert_allclose(vt, ovt, rtol=1e-13, atol=0)
        np.testing.assert_allclose(uu, ouu, rtol=1e-13, atol=0)
    ne = ctx.n_model_elems()
    K, F = ctx.element_coefs(0, ne)
    for k in range(ne):
        Ko, Fo = om.coefs(k)
        sk = max(np.abs(Ko).max(), 1e-300)
        assert np.abs(K[k] - Ko).max() <= 1e-12 * sk, k
        sf = max(np.abs(Fo).max(), np.abs(Ko).max() * 1e-3)
        assert np.abs(F[k] - Fo).max() <= 1e-12 * sf, k
    # assembled FP64 system
    ctx.assemble(rmx.F64)
    A, B = ctx.matrix(rmx.F64)
    Ao, Bo = om.assemble(64)
    rm, _ = _rowmax(rp, Ao)
    print(f"    FP64 assembly: max |gpu - oracle| / max|row| = {(np.abs(A - Ao) / rm).max():.2e}")
    assert (np.abs(A - Ao) <= 1e-12 * rm).all()
    assert np.abs(B - Bo).max() <= 1e-12 * np.abs(Bo).max()
    # FP32 stock semantics against the reference's own assembled values
    ctx.assemble(rmx.F32)
    A32, B32 = ctx.matrix(rmx.F32)
    ref = g["s0_aval"]
    # each entry is a float sum of float-rounded element contributions; the GPU adds them
    # in colour order: allow 16 float half-ulps of the sum of |contributions|
    # (or 1e-12 of the row maximum where a contribution is itself a cancelled sum)
    sabs = om.assemble_abs()
    tol = np.maximum(16 * 2.0 ** -24 * sabs, 1e-12 * rm)
    ratio = (np.abs(A32.astype(np.float64) - ref) / tol).max()
    print(f"    FP32 assembly: max |gpu - ref| / tolerance = {ratio:.2e}")
    assert ratio <= 1.0
    np.testing.assert_allclose(B32, g["s0_bvec"], rtol=0, atol=1e-12 * np.abs(g["s0_bvec"]).max())


@pytest.mark.parametrize("name", ["tri20x4", "quad10x3"])
def test_ilu_spmv_match_reference_records(ctx, name):
    """Upload the reference's own scaled FP32 system, factor it on the GPU and compare
    with the reference's ILU(0) factors, ILU apply and MulVec."""
    print(f"[gpu] ILU/SpMV parity on {name}", flush=True)
    g = load_golden(name)
    rp = np.concatenate([[0], np.cumsum(g["s0_width"])]).astype(np.int32)
    cols = g["s0_cols"].astype(np.int32)
    ctx.csr_upload(rp, cols, g["s0_ascaled"], g["s0_bscaled"])
    ctx.convert(rmx.F32)
    ctx.ilu0_factor(rmx.F32)
    L = ctx.factors(rmx.F32)
    np.testing.assert_array_equal(L.astype(np.float64), g["s0_ilu"])
    y = ctx.ilu0_apply(g["s0_bscaled"], rmx.F32)
    ref = g["s0_iluapply"]
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    r = ctx.spmv(g["s0_bscaled"], rmx.F32)
    ref = g["s0_mulvec"]
    assert np.abs(r - ref).max() <= 1e-12 * np.abs(ref).max()
    # FP64 factorisation of the same matrix against the oracle's FP64 ILU(0)
    ctx.ilu0_factor(rmx.F64)
    L64 = ctx.factors(rmx.F64)
    Lo, _ = po.ilu0_factor(rp, cols, g["s0_ascaled"].astype(np.float64))
    rm, _ = _rowmax(rp, Lo)
    assert (np.abs(L64 - Lo) <= 1e-12 * rm).all()


def _relerr(a, b, scale):
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", ["tri20x4", "quad10x3", "veg20x4", "turb20x4", "turbq10x3",
                                  "elder20x4", "elderq10x3", "channel200x20"])
def test_newton_converges_to_reference_solution(ctx, name):
    """Full Newton iteration on the GPU (FP64 assembly, FP64 ILU0-BiCGStab) reaches the
    reference program's converged u, v, h within 1e-8 relative (the algebraic model's
    eddy viscosity frozen at the first step, as the reference's flow cycle keeps it)."""
    print(f"[gpu] Newton FP64 on {name}", flush=True)
    g = load_golden(name)
    deck = golden_deck(g)
    ctx.upload_mesh(deck)
    ctx.set_params()
    n = ctx.prepare()
    if name == "channel200x20":
        rp, cols = ctx.pattern()
        h = hashlib.sha256(np.diff(rp).astype(np.int32).tobytes() + cols.tobytes()).hexdigest()
        assert h == str(g["s0_pattern_sha256"])
        eq = ctx.numbering()
        np.testing.assert_array_equal(eq, np.stack([g["s0_eqU"], g["s0_eqV"], g["s0_eqS"]], 1))
    for k in range(int(g["nsteps"]) + 1):
        st = ctx.newton_step(method=rmx.BICGSTAB, prec=rmx.F64, eps=1e-12)
        assert st.solve.converged, (k, st.solve.accuracy)
        # BiCGStab's recursive residual reaches 1e-12; the true FP64 residual of these
        # small, ill-conditioned systems stays within 1e-9 (the gap depends on the
        # summation order of the sweeps), 1e-8 on the quad deck with the anisotropic
        # algebraic model, the worst conditioned (the reference's own first solve takes
        # 327 BiCGStab iterations)
        assert st.solve.true_resid < (1e-8 if name == "elderq10x3" else 1e-9)
    U, V, S, z = ctx.state()
    Hr = g["final_S"] - g["final_z"]
    us = np.abs(g["final_U"]).max()
    assert _relerr(U, g["final_U"], us) <= 1e-8
    assert _relerr(V, g["final_V"], us) <= 1e-8
    assert _relerr(S - z, Hr, np.abs(Hr).max()) <= 1e-8
    assert n == len(g["s0_xsol"])


def test_first_newton_correction_matches_reference(ctx):
    """Stock semantics (FP32 assembly, FP32 ILU, BiCGStab to 1e-10): the first Newton
    correction agrees with the reference's.  Both solve an FP32 matrix whose entries may
    differ by float rounding (different summation order), so the corrections agree to
    the FP32 perturbation of the system, 1e-6 relative, not to the solver tolerance."""
    g = load_golden("tri20x4")
    ctx.upload_mesh(golden_deck(g))
    ctx.prepare()
    ctx.assemble(rmx.F32)
    s = ctx.scale_l2()
    assert abs(s - float(g["s0_scale"])) <= 1e-13 * s
    ctx.ilu0_factor(rmx.F32)
    out = ctx.solve(rmx.BICGSTAB, rmx.F32, rmx.PREC_ILU0, eps=1e-10)
    assert out.converged and abs(out.iters - int(g["bicgstab_iters"][0])) <= 2
    X = ctx.solution()
    ref = g["s0_xsol"]
    assert np.abs(X - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("method", [rmx.BICGSTAB, rmx.FGMRES])
def test_mixed_precision_refinement_reaches_fp64_residual(ctx, method):
    """FP32 matrix + FP32 ILU(0) inside FP64 iterative refinement reaches true FP64
    relative residual 1e-8 and the FP64 solution."""
    print(f"[gpu] mixed precision, method {method}", flush=True)
    d = dk.channel(nx=60, ny=8, length=300.0, width=40.0, noise_corr=10.0, Q=40.0)
    ctx.upload_mesh(d)
    ctx.prepare()
    ctx.assemble(rmx.F64)
    ctx.scale_l2()
    ctx.ilu0_factor(rmx.F64)
    ref = ctx.solve(method, rmx.F64, rmx.PREC_ILU0, eps=1e-12)
    X64 = ctx.solution()
    assert ref.converged
    ctx.convert(rmx.F32)
    ctx.ilu0_factor(rmx.F32)
    out = ctx.solve(method, rmx.F32, rmx.PREC_ILU0, eps=1e-5, refine=True, refine_tol=1e-8)
    assert out.converged and out.true_resid <= 1e-8
    X = ctx.solution()
    assert np.abs(X - X64).max() <= 1e-6 * np.abs(X64).max()


def _restrict(rp, cols, vals, block):
    """Entries of a CSR matrix whose row and column lie in the same block, row order and
    entry order kept."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = block[cols] == block[rows]
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int32)
    return rp2, np.ascontiguousarray(cols[keep]), np.ascontiguousarray(vals[keep])


def _check_blocks(block, n, strips, segments, group):
    """Blocks: `strips` contiguous row ranges cut at n*b/strips (moved forward to the
    next node), each split into at most `segments` parts; a node's rows share a block."""
    strip = block // segments
    assert (np.diff(strip) >= 0).all() and strip[0] == 0 and strip[-1] == strips - 1
    starts = np.concatenate([[0], np.nonzero(np.diff(strip))[0] + 1])
    for b, a in enumerate(starts):
        nominal = n * b // strips
        assert a >= nominal and (a == 0 or group[a] != group[a - 1])
        assert (group[nominal:a] == group[max(nominal - 1, 0)]).all()
    same = group[1:] == group[:-1]
    assert (block[1:][same] == block[:-1][same]).all()


@pytest.mark.parametrize("sweep", ["block", "staged"])
@pytest.mark.parametrize("strips,segments", [(1, 1), (4, 1), (17, 1), (4, 8), (1, 16), (17, 4),
                                             (4, 32)])
def test_strip_ilu_matches_oracle(ctx, strips, segments, sweep, monkeypatch):
    """Strip ILU(0) = the oracle's ILU(0) factors of the whole matrix, swept with the
    entries that couple two blocks left out: FP64 and FP32 apply agree with the oracle's
    triangular solves on that restricted pattern to 1e-12 of max|result|.  Blocks of at
    most 2047 rows run the block sweeps (one wave per block, both triangles in one launch)
    unless RMX_SWEEP=1 selects the staged strip sweeps; larger blocks run the staged ones."""
    print(f"[gpu] strip ILU, {strips} strips x {segments} segments, {sweep} sweeps", flush=True)
    if sweep == "staged":
        monkeypatch.setenv("RMX_SWEEP", "1")
    else:
        monkeypatch.delenv("RMX_SWEEP", raising=False)
    g = load_golden("channel200x20")
    ctx.upload_mesh(golden_deck(g))
    n = ctx.prepare()
    ctx.set_strips(strips, segments)
    ctx.assemble(rmx.F64)
    ctx.scale_l2()
    rp, cols = ctx.pattern()
    A, B = ctx.matrix(rmx.F64)
    eq = ctx.numbering().reshape(-1)
    group = np.zeros(n, np.int64)
    group[eq[eq >= 0]] = np.nonzero(eq >= 0)[0] // 3
    nb = max(1, min(strips, n // 2048))
    block = ctx.blocks()
    _check_blocks(block, n, nb, segments, group)
    assert len(np.unique(block)) == nb * segments
    for prec, dt in ((rmx.F64, np.float64), (rmx.F32, np.float32)):
        if prec == rmx.F32:
            ctx.convert(rmx.F32)
        ctx.silu0_factor(prec)
        L, r = po.ilu0_factor(rp, cols, A.astype(dt))
        assert r == 0
        rp2, c2, L2 = _restrict(rp, cols, L, block)
        if nb * segments > 1:
            assert len(L2) < len(L)
        ref = po.ilu0_solve(rp2, c2, L2, B)
        ctx.kernel_stats(True)
        y = ctx.silu0_apply(B, prec)
        ran = ctx.kernel_stats(False)
        blocks = sweep == "block" and max(np.bincount(block)) <= 2047
        assert (ran["sweeps_block"][2] == 1) == blocks and (ran["sweep_lower"][2] == 1) != blocks
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max(), prec
    # one block is the reference-order ILU(0)
    if nb * segments == 1:
        ctx.ilu0_factor(rmx.F32)
        np.testing.assert_array_equal(ctx.ilu0_apply(B, rmx.F32), y)


@pytest.mark.parametrize("rows,pairs", [(3, 64), (7, 64), (32, 5), (3, 1)])
def test_block_sweeps_split_levels(ctx, rows, pairs, monkeypatch):
    """Block sweeps with levels split to at most `rows` rows and `pairs` entry pairs (a node
    needing more sits alone in its level): FP64 and FP32 apply within 1e-12 of the oracle's
    triangular solves of the block-restricted factors."""
    print(f"[gpu] block sweeps, levels of at most {rows} rows / {pairs} entry pairs", flush=True)
    monkeypatch.delenv("RMX_SWEEP", raising=False)
    monkeypatch.setenv("RMX_BLOCK_LEVEL_ROWS", str(rows))
    monkeypatch.setenv("RMX_BLOCK_LEVEL_PAIRS", str(pairs))
    g = load_golden("channel200x20")
    ctx.upload_mesh(golden_deck(g))
    ctx.prepare()
    ctx.set_strips(17, 4)
    ctx.assemble(rmx.F64)
    ctx.scale_l2()
    rp, cols = ctx.pattern()
    A, B = ctx.matrix(rmx.F64)
    block = ctx.blocks()
    for prec, dt in ((rmx.F64, np.float64), (rmx.F32, np.float32)):
        if prec == rmx.F32:
            ctx.convert(rmx.F32)
        ctx.silu0_factor(prec)
        L, _ = po.ilu0_factor(rp, cols, A.astype(dt))
        rp2, c2, L2 = _restrict(rp, cols, L, block)
        ref = po.ilu0_solve(rp2, c2, L2, B)
        ctx.kernel_stats(True)
        y = ctx.silu0_apply(B, prec)
        assert ctx.kernel_stats(False)["sweeps_block"][2] == 1
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max(), prec


@pytest.mark.parametrize("coarse_solve", ["device", "device-host-factor", "host"])
@pytest.mark.parametrize("strips,segments,cseg", [(4, 8, 2), (17, 4, 4), (1, 1, 1), (4, 8, 24)])
def test_two_level_apply_matches_dense_reference(ctx, strips, segments, cseg, coarse_solve,
                                                 monkeypatch):
    """Two-level apply z = z1 + M^-1 (r - A z1), z1 = P (R A P)^-1 R r, against the same
    formula in numpy: R A P summed and solved densely, M^-1 by the oracle's triangular
    solves of the block-restricted full ILU(0) factors.  FP64 and FP32 (matrix, factors and
    coarse matrix from the FP32 values): 1e-9 of max|z|.  The coarse system is solved by
    block cyclic reduction on the device (R A P is block tridiagonal in strips here; its
    blocks inverted and reduced on the device, or on the host with RMX_COARSE_HOST_FACTOR)
    or, with RMX_COARSE_HOST set when the coarse space is built, by the host band LU."""
    import scipy.sparse as sp
    print(f"[gpu] two-level apply, {strips} strips x {segments} segments, {cseg} coarse, "
          f"{coarse_solve} coarse solve", flush=True)
    monkeypatch.delenv("RMX_COARSE_HOST", raising=False)
    monkeypatch.delenv("RMX_COARSE_HOST_FACTOR", raising=False)
    if coarse_solve == "host":
        monkeypatch.setenv("RMX_COARSE_HOST", "1")
    elif coarse_solve == "device-host-factor":
        monkeypatch.setenv("RMX_COARSE_HOST_FACTOR", "1")
    g = load_golden("channel200x20")
    ctx.upload_mesh(golden_deck(g))
    n = ctx.prepare()
    ctx.set_coarse(cseg)
    ctx.set_strips(strips, segments)
    ctx.assemble(rmx.F64)
    ctx.scale_l2()
    rp, cols = ctx.pattern()
    A, B = ctx.matrix(rmx.F64)
    block = ctx.blocks()
    agg, nc = ctx.aggregates()
    nb = max(1, min(strips, n // 2048))
    # aggregates: numbered strip by strip, never spanning two strips, at most 3 per segment
    strip = block // segments
    assert (np.diff(agg[np.argsort(strip, kind="stable")]) >= -3 * cseg).all()
    first = np.zeros(nc, np.int64) - 1
    first[agg] = strip
    assert (first[agg] == strip).all() and (np.diff(first) >= 0).all()
    assert nc <= nb * cseg * 3 and len(np.unique(agg)) == nc
    P = sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))
    rng = np.random.default_rng(3)
    r = rng.standard_normal(n)
    for prec, dt in ((rmx.F64, np.float64), (rmx.F32, np.float32)):
        if prec == rmx.F32:
            ctx.convert(rmx.F32)
        ctx.silu0_factor(prec)
        ctx.coarse_factor(prec)
        Ap = A.astype(dt)
        M = sp.csr_matrix((Ap.astype(np.float64), cols, rp), shape=(n, n))
        Ac = (P.T @ M @ P).toarray()
        z1 = P @ np.linalg.solve(Ac, P.T @ r)
        r1 = r - M @ z1
        L, _ = po.ilu0_factor(rp, cols, Ap)
        rp2, c2, L2 = _restrict(rp, cols, L, block)
        ref = z1 + po.ilu0_solve(rp2, c2, L2, r1)
        y = ctx.precond_apply(rmx.PREC_SILU0C, r, prec)
        err = np.abs(y - ref).max() / np.abs(ref).max()
        print(f"    prec {prec}: coarse {nc} unknowns, max rel err {err:.2e}")
        assert err <= 1e-9, prec


@pytest.mark.parametrize("precond", [rmx.PREC_SILU0, rmx.PREC_SILU0C])
def test_strip_ilu_refinement_newton_reaches_reference(ctx, precond):
    """Newton with FP32 strip-ILU(0) BiCGStab (alone or after the coarse correction) inside
    FP64 refinement (true FP64 residual 1e-10 per step) reaches the reference program's
    converged state within 1e-8."""
    print(f"[gpu] Newton, strip ILU FP32 + FP64 refinement, precond {precond}", flush=True)
    g = load_golden("channel200x20")
    ctx.upload_mesh(golden_deck(g))
    ctx.prepare()
    ctx.set_coarse(4)
    ctx.set_strips(0, 16)
    for k in range(int(g["nsteps"]) + 1):
        st = ctx.newton_step(method=rmx.BICGSTAB, prec=rmx.F32, precond=precond,
                             eps=1e-4, refine=True, refine_tol=1e-10, refine_max=20)
        assert st.solve.converged and st.solve.true_resid <= 1e-10, (k, st.solve.true_resid)
    U, V, S, z = ctx.state()
    Hr = g["final_S"] - g["final_z"]
    us = np.abs(g["final_U"]).max()
    assert _relerr(U, g["final_U"], us) <= 1e-8
    assert _relerr(V, g["final_V"], us) <= 1e-8
    assert _relerr(S - z, Hr, np.abs(Hr).max()) <= 1e-8


def _random_csr(n, seed, density):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) * (rng.random((n, n)) < density)
    M = M + M.T * 0.3
    np.fill_diagonal(M, 0.0)
    M += np.diag(np.abs(M).sum(1) + 1.0)
    rp, cols, vals = [0], [], []
    for i in range(n):
        js = [i] + [j for j in rng.permutation(n) if j != i and M[i, j] != 0.0]
        cols += js
        vals += [M[i, j] for j in js]
        rp.append(len(cols))
    return M, np.asarray(rp, np.int32), np.asarray(cols, np.int32), np.asarray(vals)


def test_csr_kernels_against_dense(ctx):
    M, rp, cols, vals = _random_csr(300, 7, 0.03)
    b = np.random.default_rng(8).standard_normal(300)
    ctx.csr_upload(rp, cols, vals, b)
    np.testing.assert_allclose(ctx.spmv(b), M @ b, rtol=1e-12, atol=1e-12)
    ctx.ilu0_factor(rmx.F64)
    Lo, _ = po.ilu0_factor(rp, cols, vals)
    np.testing.assert_allclose(ctx.ilu0_apply(b), po.ilu0_solve(rp, cols, Lo, b), rtol=1e-11, atol=1e-12)
    for method in (rmx.BICGSTAB, rmx.FGMRES):
        out = ctx.solve(method, rmx.F64, rmx.PREC_ILU0, eps=1e-13)
        assert out.converged
        np.testing.assert_allclose(ctx.solution(), np.linalg.solve(M, b), rtol=1e-9, atol=1e-10)


def _grouped_csr(n_nodes, seed, far=None, nbr=5):
    """CSR whose rows come in runs of 1..6 consecutive rows with one column set (the rows
    of a node, coupled to up to `nbr` other nodes), columns in random order after the
    diagonal; `far` adds a coupling of the first node to the last row (a column span of the
    whole matrix)."""
    rng = np.random.default_rng(seed)
    sizes = 1 + np.arange(n_nodes) % 6
    first = np.concatenate([[0], np.cumsum(sizes)])
    n = int(first[-1])
    rows_of = [np.arange(first[k], first[k + 1]) for k in range(n_nodes)]
    rp, cols, vals = [0], [], []
    for k in range(n_nodes):
        nb = {k} | set(np.clip(k + rng.integers(-40, 41, size=nbr), 0, n_nodes - 1).tolist())
        cset = np.concatenate([rows_of[j] for j in sorted(nb)])
        if far and k == 0:
            cset = np.concatenate([cset, [n - 1]])
        for i in rows_of[k]:
            others = rng.permutation(cset[cset != i])
            js = np.concatenate([[i], others])
            v = rng.standard_normal(len(js))
            v[0] = np.abs(v[1:]).sum() + 1.0
            cols += js.tolist()
            vals += v.tolist()
            rp.append(len(cols))
    return n, np.asarray(rp, np.int32), np.asarray(cols, np.int32), np.asarray(vals)


def _csr_mul(rp, cols, vals, x):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, weights=vals * x[cols], minlength=len(rp) - 1)


def test_f32_spmv_row_groups(ctx):
    """FP32 SpMV on the row-group layout (rows sharing a column set, up to 4 per group,
    runs of 5 and 6 split), against a float64 product of the FP32 values; again after the
    FP32 matrix changes (the layout's values follow A32); with narrow rows (4 lanes per
    group), and on a matrix whose column span exceeds 16-bit offsets (CSR kernel)."""
    for n_nodes, far, nbr in ((1500, False, 5), (1500, False, 1), (22000, True, 5)):
        n, rp, cols, vals = _grouped_csr(n_nodes, 11, far, nbr)
        print(f"[gpu] FP32 SpMV row groups: n {n}, nnz {rp[-1]}, far coupling {far}", flush=True)
        b = np.random.default_rng(12).standard_normal(n)
        ctx.csr_upload(rp, cols, vals, b)
        ctx.convert(rmx.F32)
        for rnd in range(2):
            A32, _ = ctx.matrix(rmx.F32)
            a = A32.astype(np.float64)
            ref = _csr_mul(rp, cols, a, b)
            scale = _csr_mul(rp, cols, np.abs(a), np.abs(b))
            y = ctx.spmv(b, rmx.F32)
            assert (np.abs(y - ref) <= 1e-13 * scale).all(), rnd
            ctx.scale_l2()                       # new FP64 values, then a new FP32 copy
            ctx.convert(rmx.F32)
        # the FP32 products inside FGMRES (restart residual and basis) on the same layout
        ctx.ilu0_factor(rmx.F32)
        out = ctx.solve(rmx.FGMRES, rmx.F32, rmx.PREC_ILU0, eps=1e-5, mkyrl=20, max_iter=400,
                        refine=True, refine_tol=1e-10, refine_max=20)
        assert out.converged and out.true_resid <= 1e-10


def test_solve_proceed_chain_follows_the_reference(ctx):
    """EQS::Solve's fallback after a solver that has not converged (Solve.cpp:325-362), on
    a random diagonally dominant system: proceed > 0 continues with the next solver from X
    (when its tolerance is below the accuracy reached) and reaches the dense solution;
    proceed = -1 zeroes X and reports the solve interrupted; proceed = 0 keeps the
    unconverged X; a next solver whose tolerance is not below the accuracy reached is not
    run."""
    print("[gpu] proceed chain of the solve", flush=True)
    M, rp, cols, vals = _random_csr(300, 9, 0.03)
    b = np.random.default_rng(10).standard_normal(300)
    ref = np.linalg.solve(M, b)
    ctx.csr_upload(rp, cols, vals, b)
    ctx.ilu0_factor(rmx.F64)
    weak = dict(method=rmx.BICGSTAB, prec=rmx.F64, precond=rmx.PREC_NONE, max_iter=2, eps=1e-12)
    strong = dict(method=rmx.FGMRES, prec=rmx.F64, precond=rmx.PREC_ILU0, mkyrl=20, max_iter=500,
                  eps=1e-13)
    out = ctx.solve(**weak, proceed=1, next=strong)
    assert out.converged and out.solvers == 2 and out.iters > 2
    np.testing.assert_allclose(ctx.solution(), ref, rtol=1e-9, atol=1e-10)
    out = ctx.solve(**weak, proceed=-1)
    assert not out.converged and out.interrupted == 1 and out.solvers == 1
    assert not ctx.solution().any()
    out = ctx.solve(**weak, proceed=0)
    X = ctx.solution()
    assert not out.converged and out.interrupted == 0 and X.any()
    acc = out.accuracy
    out = ctx.solve(**weak, proceed=1, next=dict(strong, eps=1.5 * acc))
    assert not out.converged and out.solvers == 1
    # in refinement the chain continues from X as well (FP64 true residual reached)
    out = ctx.solve(**dict(weak, refine=True, refine_tol=1e-12, refine_max=1), proceed=1,
                    next=dict(strong, refine=True, refine_tol=1e-12, refine_max=5))
    assert out.converged and out.solvers == 2 and out.true_resid <= 1e-12
    np.testing.assert_allclose(ctx.solution(), ref, rtol=1e-9, atol=1e-10)


def test_full_pattern_ilu_is_exact_lu(ctx):
    M, rp, cols, vals = _random_csr(40, 3, 1.0)
    b = np.random.default_rng(4).standard_normal(40)
    ctx.csr_upload(rp, cols, vals, b)
    ctx.ilu0_factor(rmx.F64)
    np.testing.assert_allclose(ctx.ilu0_apply(b), np.linalg.solve(M, b), rtol=1e-10)


def _band_csr(n, half, seed):
    """Full band |i - j| <= half (row width up to 2 half + 1 <= 128), diagonally dominant,
    columns after the diagonal in random order."""
    rng = np.random.default_rng(seed)
    M = np.zeros((n, n))
    for i in range(n):
        lo, hi = max(0, i - half), min(n, i + half + 1)
        M[i, lo:hi] = rng.standard_normal(hi - lo)
        M[i, i] = np.abs(M[i, lo:hi]).sum() + 1.0
    rp, cols, vals = [0], [], []
    for i in range(n):
        js = [i] + [j for j in rng.permutation(n) if j != i and abs(i - j) <= half]
        cols += js
        vals += [M[i, j] for j in js]
        rp.append(len(cols))
    return M, np.asarray(rp, np.int32), np.asarray(cols, np.int32), np.asarray(vals)


def test_factorisation_of_rows_with_long_pivot_rows(ctx):
    """Rows whose pivot rows hold more entries after their diagonals than the
    factorisation stages in LDS (2048: 63 pivots x 63 entries in the band's middle) take
    the branch that reads them from memory; the band's first rows stay staged.  ILU(0) on
    a full band is its exact LU: the FP64 apply equals a dense solve, the FP64 factors the
    oracle's, and the FP32 factors are bit-identical to the oracle's FP32 factorisation."""
    print("[gpu] ILU(0) with pivot rows beyond the LDS staging (band 127)", flush=True)
    M, rp, cols, vals = _band_csr(400, 63, 11)
    b = np.random.default_rng(12).standard_normal(400)
    ctx.csr_upload(rp, cols, vals, b)
    ctx.ilu0_factor(rmx.F64)
    np.testing.assert_allclose(ctx.ilu0_apply(b), np.linalg.solve(M, b), rtol=1e-10, atol=1e-12)
    L64, _ = po.ilu0_factor(rp, cols, vals)
    np.testing.assert_allclose(ctx.factors(rmx.F64), L64, rtol=1e-12, atol=1e-14)
    ctx.convert(rmx.F32)
    ctx.ilu0_factor(rmx.F32)
    L32, _ = po.ilu0_factor(rp, cols, vals.astype(np.float32))
    np.testing.assert_array_equal(ctx.factors(rmx.F32), L32)


def test_pivot_repair_matches_oracle(ctx):
    """Pivots below 1e-20 are repaired as PRECO_ILU0::Factor does (Preco_ilu0.cpp:95-135):
    the hand-worked 3 x 3 case of the oracle's test bit for bit, and a 300-row system whose
    diagonal is zeroed in 6 rows (the repair runs in the order of the rows, each pass of the
    dataflow factorisation stopping at the first; their lower parts are zeroed as well, so the
    pivots stay exactly zero): FP32 factors bit-identical to the oracle's,
    FP64 within 1e-12, the FP64 apply a solve with those factors."""
    print("[gpu] ILU(0) pivot repair", flush=True)
    rp = np.array([0, 3, 6, 9], np.int32)
    cols = np.array([0, 1, 2, 1, 0, 2, 2, 0, 1], np.int32)
    A = np.array([0, 1, 0, 1, 1, 1, 2, 0, 1], np.float64)
    ctx.csr_upload(rp, cols, A, np.ones(3))
    ctx.ilu0_factor(rmx.F64)
    want = np.array([1, 2, 1, -1, -1, 0, 2, -0.0, 1])
    np.testing.assert_array_equal(ctx.factors(rmx.F64), want)
    ctx.convert(rmx.F32)
    ctx.ilu0_factor(rmx.F32)
    np.testing.assert_array_equal(ctx.factors(rmx.F32), want.astype(np.float32))
    M, rp, cols, vals = _random_csr(300, 21, 0.03)
    z = vals.copy()
    for r in (5, 50, 51, 120, 200, 298):        # zero diagonal and lower part: a zero pivot
        q = np.arange(rp[r], rp[r + 1])
        z[q[cols[q] <= r]] = 0.0
    b = np.random.default_rng(22).standard_normal(300)
    L64, rep64 = po.ilu0_factor(rp, cols, z)
    L32, rep32 = po.ilu0_factor(rp, cols, z.astype(np.float32))
    assert rep64 >= 6 and rep32 >= 6
    ctx.csr_upload(rp, cols, z, b)
    ctx.ilu0_factor(rmx.F64)
    np.testing.assert_allclose(ctx.factors(rmx.F64), L64, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ctx.ilu0_apply(b), po.ilu0_solve(rp, cols, L64, b), rtol=1e-10,
                               atol=1e-12)
    ctx.convert(rmx.F32)
    ctx.ilu0_factor(rmx.F32)
    np.testing.assert_array_equal(ctx.factors(rmx.F32), L32)
    print(f"    {rep32} pivots repaired", flush=True)


def test_error_paths(ctx):
    M, rp, cols, vals = _random_csr(20, 5, 0.2)
    bad = cols.copy()
    bad[rp[3]] = 4                      # diagonal not first in row 3
    with pytest.raises(rmx.RmxError):
        ctx.csr_upload(rp, bad, vals, np.ones(20))
    ctx.csr_upload(rp, cols, vals, np.ones(20))
    with pytest.raises(rmx.RmxError):
        ctx.solve(rmx.BICGSTAB, rmx.F32, rmx.PREC_ILU0)   # no FP32 matrix / factors
    with pytest.raises(rmx.RmxError):
        ctx.solve(rmx.FGMRES, rmx.F64, rmx.PREC_NONE, mkyrl=64)
    # a zero pivot with no later row to repair it from (PRECO_ILU0::Factor - 1)
    ctx.csr_upload(np.array([0, 1], np.int32), np.array([0], np.int32), np.zeros(1), np.ones(1))
    with pytest.raises(rmx.RmxError) as e:
        ctx.ilu0_factor(rmx.F64)
    assert e.value.code == -3
    # turbulence model codes outside the supported set are refused (tens digit 1 is the
    # anisotropic model, 2 is no model); the nodal quantities need the algebraic or the
    # anisotropic model (code 2 or + 10, with + 100 / + 200)
    for code in (3, 22, 402, -2):
        with pytest.raises(rmx.RmxError):
            ctx.set_params(turb=code)
    # relaxation methods the update does not implement, and a dry limit that leaves no
    # positive minimum depth, are refused at the boundary
    for kw in (dict(relax_method=1), dict(relax_method=5), dict(dry_limit=0.0),
               dict(dry_limit=-1e-3)):
        with pytest.raises(rmx.RmxError, match="params"):
            ctx.set_params(**kw)
    ctx.upload_mesh(golden_deck(load_golden("tri20x4")))
    with pytest.raises(rmx.RmxError):
        ctx.turbulence()
    ctx.set_params(turb=302)
    vt, uu = ctx.turbulence()
    assert (vt > 0).all() and (uu < 0).all()
    ctx.set_params(turb=1)


UNSTEADY = ["unst20x4", "unstq10x3", "hydro20x4"]


@pytest.mark.parametrize("name", UNSTEADY)
def test_unsteady_element_kernel_matches_oracle(ctx, name):
    """Theta-scheme time terms: at the golden state of the Newton step whose element
    matrices the reference recorded (non-zero time derivatives, previous time step's
    state), the GPU element kernel and assembly agree with the oracle (itself bit-exact
    with those records)."""
    print(f"[gpu] unsteady element parity on {name}", flush=True)
    g = load_golden(name)
    deck = golden_deck(g)
    p = f"s{int(g['elem_step'])}_"
    dt, th = float(g["dt"]), float(g["theta"])
    tkw = {k: g[p + k] for k in rmx.Context.TIME_KEYS}
    ctx.upload_mesh(deck)
    ctx.state_upload(g[p + "U"], g[p + "V"], g[p + "S"])
    ctx.time_step(dt, th, first=False)
    ctx.set_time_state(**tkw)
    inflow = None
    if "hydrograph" in g:
        inflow = rmx.hydrograph(g["hydrograph"], int(g[p + "tstep"]) * dt)
        ctx.set_inflow(*inflow)
    n = ctx.prepare()
    ts = ctx.time_state()
    for k in rmx.Context.TIME_KEYS:
        np.testing.assert_array_equal(ts[k], g[p + k])
    om = po.OracleModel(deck)
    om.set_state(g[p + "U"], g[p + "V"], g[p + "S"])
    om.time_step(dt, th)
    om.set_time_state(**tkw)
    if inflow is not None:
        om.set_inflow(*inflow)
    assert om.prepare() == n
    ne = ctx.n_model_elems()
    K, F = ctx.element_coefs(0, ne)
    for k in range(ne):
        Ko, Fo = om.coefs(k)
        sk = max(np.abs(Ko).max(), 1e-300)
        assert np.abs(K[k] - Ko).max() <= 1e-12 * sk, k
        sf = max(np.abs(Fo).max(), np.abs(Ko).max() * 1e-3)
        assert np.abs(F[k] - Fo).max() <= 1e-12 * sf, k
    ctx.assemble(rmx.F64)
    A, B = ctx.matrix(rmx.F64)
    Ao, Bo = om.assemble(64)
    rp, _ = ctx.pattern()
    rm, _ = _rowmax(rp, Ao)
    print(f"    FP64 assembly: max |gpu - oracle| / max|row| = {(np.abs(A - Ao) / rm).max():.2e}")
    assert (np.abs(A - Ao) <= 1e-12 * rm).all()
    # near a converged time step the residual is a small difference of large terms: scale
    # as for the element forces
    sb = max(np.abs(Bo).max(), 1e-3 * np.abs(Ao).max())
    assert np.abs(B - Bo).max() <= 1e-12 * sb
    ctx.time_step(0.0, 1.0)              # back to stationary flow for the tests that follow


@pytest.mark.parametrize("name", UNSTEADY)
def test_unsteady_newton_follows_reference(ctx, name):
    """The reference's time steps on the GPU: per time step Predict, as many Newton steps
    (FP64 assembly, FP64 ILU0-BiCGStab) as the reference took, Timegrad after every
    update, commit at the end.  At each later time step's start the state and the
    previous state agree with the reference's within 1e-8 of their scale, the previous
    time derivatives within 1e-9 of theirs; the final state within 1e-8."""
    print(f"[gpu] unsteady Newton on {name}", flush=True)
    g = load_golden(name)
    deck = golden_deck(g)
    dt, th = float(g["dt"]), float(g["theta"])
    ctx.upload_mesh(deck)
    ts = [int(g[f"s{k}_tstep"]) for k in range(int(g["nsteps"]))]
    first = True
    for t in sorted(set(ts)):
        ks = [k for k in range(len(ts)) if ts[k] == t]
        p = f"s{ks[0]}_"
        if not first:
            ctx.time_commit()
            U, V, S, z = ctx.state()
            us = np.abs(g[p + "U"]).max()
            assert _relerr(U, g[p + "U"], us) <= 1e-8, t
            assert _relerr(V, g[p + "V"], us) <= 1e-8, t
            assert _relerr(S - z, g[p + "S"] - g[p + "z"], np.abs(g[p + "S"] - g[p + "z"]).max()) <= 1e-8
            tsd = ctx.time_state()
            for k in ("dUdto", "dVdto", "dSdto"):
                sc = max(np.abs(g[p + k]).max(), 1e-12)
                err = _relerr(tsd[k], g[p + k], sc)
                print(f"    time step {t}: {k} rel err {err:.2e}")
                assert err <= 1e-9, (t, k)
        ctx.time_step(dt, th, first=first)
        if "hydrograph" in g:
            # QTINLET: the inflow at the time step's end (EqsUVS2D.cpp:318-320)
            ctx.set_inflow(*rmx.hydrograph(g["hydrograph"], t * dt))
        first = False
        for k in ks:
            st = ctx.newton_step(method=rmx.BICGSTAB, prec=rmx.F64, eps=1e-12)
            assert st.solve.converged and st.solve.true_resid < 1e-9, (t, k)
        assert st.nrconv, t
    U, V, S, z = ctx.state()
    us = np.abs(g["final_U"]).max()
    assert _relerr(U, g["final_U"], us) <= 1e-8
    assert _relerr(V, g["final_V"], us) <= 1e-8
    Hr = g["final_S"] - g["final_z"]
    assert _relerr(S - z, Hr, np.abs(Hr).max()) <= 1e-8
    ctx.time_step(0.0, 1.0)
