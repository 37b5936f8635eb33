"""CPU-side checks of the HIP library: it loads, exports every entry point that
include/rmx.h declares, and reports the missing device cleanly (no compute on CPU)."""
import ctypes

import numpy as np
import pytest

from rismo2d_amd import deck as dk
from rismo2d_amd import rmx


def test_library_exports_every_declared_symbol():
    names = rmx.header_functions()
    assert len(names) >= 25
    L = rmx.lib()
    for n in names:
        assert hasattr(L, n), n


def test_create_without_gpu_reports_no_device():
    # no torch here: torch brings its own HIP and HSA runtimes, and a process that maps
    # them next to the ones librmx.so links has two runtimes on one device
    h = ctypes.c_void_p()
    rc = rmx.lib().rmx_create(0, ctypes.byref(h))
    if rc == 0:
        rmx.lib().rmx_destroy(h)
        pytest.skip("a gfx950 GPU is present")
    assert rc == rmx.ERR_NODEVICE
    assert b"device" in rmx.lib().rmx_last_error()


def test_null_context_is_an_argument_error():
    rc = rmx.lib().rmx_assemble(None, 64)
    assert rc == -1
    assert b"null context" in rmx.lib().rmx_last_error()


def test_deck_generators_shapes():
    d = dk.channel(nx=200, ny=20)
    assert d.n_elems == 8000 and d.n_nodes == 16441
    q = dk.channel(nx=10, ny=3, quads=True)
    assert q.n_elems == 30 and (q.nnd == 8).all()
    m = dk.meander(n_along=40, n_across=8)
    assert m.n_elems == 2 * 40 * 8
    # every triangle is counter-clockwise
    e = m.elems[:, :3]
    x, y = m.x[e], m.y[e]
    area2 = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
    assert (area2 > 0).all()
    # midside nodes sit on their edges' midpoints
    for i, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
        mid = m.elems[:, 3 + i]
        np.testing.assert_allclose(m.x[mid], 0.5 * (m.x[m.elems[:, a]] + m.x[m.elems[:, b]]), atol=1e-9)


def test_hydrograph_lookup_matches_oracle():
    """The library's QTINLET lookup (host code, no device needed) against the oracle's,
    bit for bit, inside, at and beyond the hydrograph's points."""
    from oracle import pyoracle as po
    h = np.array([[0.0, 10.0, 2.10], [60.0, 16.0, 2.16], [120.0, 8.0, 2.04], [120.0, 9.0, 2.0],
                  [300.0, 12.0, 2.2]])
    for t in np.concatenate([np.linspace(-10.0, 320.0, 331), [60.0, 120.0, 120.06, 300.2]]):
        assert rmx.hydrograph(h, t) == po.hydrograph(h, t), t


def test_shipped_learned_weights_fit_the_network():
    """The weights shipped with the package (rismo2d_amd/mpnn_meander.npy) load without
    pickle, are float32 with the library's parameter count in the layout of mpnn.shapes(),
    finite, with a trained (non-zero) output layer; the record beside them lists the 64
    training seeds and the samples harvested."""
    import json
    import os
    from rismo2d_amd import mpnn
    here = os.path.dirname(mpnn.__file__)
    theta = np.load(os.path.join(here, "mpnn_meander.npy"))
    n = rmx.Context.mpnn_n_params()
    assert theta.dtype == np.float32 and theta.shape == (n,)
    assert sum(int(np.prod(sh)) for _, sh in mpnn.shapes()) == n
    assert np.isfinite(theta).all()
    k0 = sum(int(np.prod(sh)) for name, sh in mpnn.shapes() if name not in ("W_out", "b_out"))
    assert np.abs(theta[k0:]).max() > 0.0          # the initialisation's output layer is zero
    rec = json.load(open(os.path.join(here, "mpnn_meander.json")))
    seeds = [r["seed"] for r in rec["seeds"]]
    assert seeds == list(range(1, 65))
    assert sum(r.get("samples", 0) for r in rec["seeds"]) == 2640
