"""Reading decks in the reference's keyword input format (rismo2d_amd.deck.read_reference_deck).

* Round trips: the files write_reference_deck writes (the same files the reference program
  read to produce the golden vectors) read back to the same deck, bit for bit.
* Anchored on the reference: a golden deck written and read back gives the oracle the
  reference's own numbering, CSR pattern, node kinds and nodal state (the s0 records of the
  golden fixtures, which the reference program produced from those files).
* The readers' rules, on small hand-written files, each from the reference code cited:
  comment lines, element shapes the region file skips, nodes listed out of order, midside
  bed elevations, initial files with named or fixed columns, material defaults, repeated
  and accumulated line conditions.
* What the solve path does not support is rejected, not misread.
"""
import os
import re

import numpy as np
import pytest

from helpers import golden_deck, load_golden
from oracle import pyoracle as po
from rismo2d_amd import deck as dk

FIELDS = ("x", "y", "z", "elems", "nnd", "manning", "vt", "U0", "V0", "S0")


def _roundtrip(deck, tmp_path, **kw):
    dk.write_reference_deck(deck, str(tmp_path), **kw)
    return dk.read_reference_deck(os.path.join(str(tmp_path), "input.dat"))


def _same(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    for k in (dk.INLET_LINE, dk.OUTLET_LINE):
        np.testing.assert_array_equal(a.lines[k], b.lines[k])
    assert a.S_outlet == b.S_outlet and a.turb == b.turb
    ne = a.n_elems
    for f, dflt in (("veg_dp", 0.0), ("veg_sp", 0.0), ("turb_st", 1.0), ("turb_kdest", 0.6),
                    ("turb_estx", 0.0), ("turb_esty", 0.0)):
        va = getattr(a, f) if getattr(a, f) is not None else np.full(ne, dflt)
        vb = getattr(b, f) if getattr(b, f) is not None else np.full(ne, dflt)
        np.testing.assert_array_equal(va, vb, err_msg=f)
    if a.hydrograph is not None:
        np.testing.assert_array_equal(np.asarray(a.hydrograph, np.float64), b.hydrograph)
    else:
        assert b.hydrograph is None
        assert a.Q_inlet == b.Q_inlet and a.S_inlet == b.S_inlet


@pytest.mark.parametrize("make", [
    lambda: dk.channel(nx=20, ny=4),
    lambda: dk.channel(nx=10, ny=3, quads=True),
    lambda: dk.channel(nx=20, ny=4, veg_frac=0.5),
    lambda: dk.channel(nx=20, ny=4, turb=102, turb_st=(0.5, 2.0), turb_kdest=(0.4, 0.8)),
    lambda: dk.channel(nx=20, ny=4, turb=112, turb_estx=(0.4, 1.2), turb_esty=0.2),
], ids=["tri", "quad", "vegetation", "algebraic-vt", "elder"])
def test_round_trip_is_exact(make, tmp_path):
    d = make()
    r, run = _roundtrip(d, tmp_path, tm_steps=3, dt=600.0, theta=0.7, stationary=0)
    _same(d, r)
    assert (run["steps"], run["dt"], run["theta"], run["stationary"]) == (3, 600.0, 0.7, 0)
    assert "$GPDEGREE 5" in run["project"]


def test_round_trip_hydrograph(tmp_path):
    d = dk.channel(nx=20, ny=4)
    d.hydrograph = np.array([[0.0, 80.0, 2.05], [3600.0, 120.0, 2.15], [7200.0, 90.0, 2.1]])
    r, _ = _roundtrip(d, tmp_path)
    _same(d, r)
    assert (r.Q_inlet, r.S_inlet) == (80.0, 2.05)


@pytest.mark.parametrize("name", ["tri20x4", "quad10x3", "veg20x4", "turb20x4", "hydro20x4",
                                  "elderq10x3"])
def test_read_deck_gives_reference_numbering_and_pattern(name, tmp_path):
    """The reference program read these very files (written by write_reference_deck) to make
    the golden vectors: the deck read back must reproduce its numbering, pattern, node kinds
    and nodal state at the first Newton step."""
    g = load_golden(name)
    kw = {}
    if "stationary" in g:
        kw = dict(tm_steps=int(g["tm_steps"]), dt=float(g["dt"]), theta=float(g["theta"]),
                  stationary=int(g["stationary"]))
    r, run = _roundtrip(golden_deck(g), tmp_path, **kw)
    om = po.OracleModel(r)
    om.set_params()
    if r.turb % 10 == 2:
        om.set_turbulence(r.turb)
    if run["stationary"] == 0:
        # the first step of the unsteady cycle, from the settings read back; QTINLET takes
        # the inflow at the step's end
        om.time_step(run["dt"], run["theta"], first=True)
        if r.hydrograph is not None:
            om.set_inflow(*po.hydrograph(r.hydrograph, int(g["s0_tstep"]) * run["dt"]))
    n = om.prepare()
    eq = om.numbering()
    np.testing.assert_array_equal(eq, np.stack([g["s0_eqU"], g["s0_eqV"], g["s0_eqS"]], 1))
    rp, cols = om.pattern()
    np.testing.assert_array_equal(np.diff(rp), g["s0_width"])
    np.testing.assert_array_equal(cols, g["s0_cols"])
    assert n == len(g["s0_width"])
    np.testing.assert_array_equal(om.nodal()["kind"], g["s0_kind"].astype(np.int32))
    U, V, S, z = om.state()
    for a, k in ((U, "U"), (V, "V"), (S, "S"), (z, "z")):
        np.testing.assert_array_equal(a, g["s0_" + k], err_msg=k)


# ---------------------------------------------------------------------------------------
# the readers' rules on hand-written files

def _write(tmp_path, files):
    for name, text in files.items():
        (tmp_path / name).write_text(text)
    return str(tmp_path / "input.dat")


# two triangles on the unit square (corners 1..4, midsides 5..9); node lines out of order,
# a comment of each kind, a 'line' element in the region file (skipped by the region reader);
# node 9 (midside of the diagonal) carries a bed elevation that is not its corners' mean
REGION = """* region
9 3 0
# nodes
4 0 1 0.4
1 0 0 0.1
2 1 0 0.2
3 1 1 0.3
5 0.5 0 9.0
6 1 0.5 9.0
7 0.5 0.5 9.0
8 0 0.5 9.0
9 0.5 1 9.0
1 7 tri 1 2 3 5 6 7
3 5 line 1 2 5
2 7 tri 1 3 4 7 9 8
"""
CONTROL = """2 3
1 0 0 0
2 0 0 0
* control elements
1 4 line 1 4 8
2 9 line 2 3 6
3 4 tri 1 2 3
"""
MAT = """$RISMO2D 40000
* Manning, no turbulence line (class defaults)
   $ROUGHNESS 7 5 0.035 0 0.0 0.0 0.0
$ROUGHNESS 0 5 0.5 0 0.0 0.0 0.0
$TURBULENCE 0 9.0 9.0 0 0 9.0 0 0
"""
TM = """$RISMO2D 40100
$TM_STEPS 1 1 0
$TM_INTERVAL 1.0e6 0 0
$TM_STEP_NO 1
$TM_CYCLE 1
$TM_TURBULENCE 1
$TM_MAXITER 6
$TM_LINE QINLET 4 50.0 1.5
$TM_LINE QINLET 4 60.0 1.6
$TM_LINE OUTLET 9 1.4 0 0 0 0.0
"""
PROJECT = """$RISMO2D 40700
$REGIONFILE region.dat
$CONTROLFILE control.dat
$ASC_INITFILE init.dat
$MATERIALFILE mat.dat
$TIMESTEPFILE tm.dat
$RELAX 2
"""


def _files(init):
    return {"region.dat": REGION, "control.dat": CONTROL, "mat.dat": MAT, "tm.dat": TM,
            "input.dat": PROJECT, "init.dat": init}


def test_reader_rules(tmp_path):
    init = ("# 0.0 x 40000 S U dUdt V\n"
            "9 0\n" + "".join(f"{i} {1.0 + i} {0.1 * i} 7.0 {0.01 * i}\n" for i in range(1, 10)))
    d, run = dk.read_reference_deck(_write(tmp_path, _files(init)))
    # nodes by name; the line element of the region file is not an element
    np.testing.assert_array_equal(d.x[:4], [0, 1, 1, 0])
    assert d.n_elems == 2 and list(d.nnd) == [6, 6]
    np.testing.assert_array_equal(d.elems[1, :6], [0, 2, 3, 6, 8, 7])
    # midside bed elevations are their corners' means (Grid.cpp:354-375)
    zc = [0.1, 0.2, 0.3, 0.4]
    np.testing.assert_array_equal(d.z[4:], [0.5 * (zc[0] + zc[1]), 0.5 * (zc[1] + zc[2]),
                                            0.5 * (zc[0] + zc[2]), 0.5 * (zc[3] + zc[0]),
                                            0.5 * (zc[2] + zc[3])])
    # named initial columns in file order S U dUdt V
    np.testing.assert_array_equal(d.S0, 1.0 + np.arange(1, 10))
    np.testing.assert_array_equal(d.U0, 0.1 * np.arange(1, 10))
    np.testing.assert_array_equal(d.V0, 0.01 * np.arange(1, 10))
    # control lines by number (inlet 4 -> line 1, outlet 9 -> line 2), tri skipped
    np.testing.assert_array_equal(d.lines[dk.INLET_LINE], [[0, 3, 7]])
    np.testing.assert_array_equal(d.lines[dk.OUTLET_LINE], [[1, 2, 5]])
    # the last QINLET line counts; material defaults (no 0 lines come too late to matter)
    assert (d.Q_inlet, d.S_inlet, d.S_outlet) == (60.0, 1.6, 1.4)
    np.testing.assert_array_equal(d.manning, [0.035, 0.035])
    np.testing.assert_array_equal(d.vt, [1.0e-4, 1.0e-4])
    np.testing.assert_array_equal(d.turb_st, [1.0, 1.0])
    np.testing.assert_array_equal(d.turb_kdest, [0.6, 0.6])
    assert d.veg_dp is None and run["steps"] == 1


def test_initial_file_fixed_columns_and_defaults(tmp_path):
    # release 300 layout: name U V S K D C dUdt dVdt dSdt dummy; S defaults to the bed
    init = "# 0.0 x 300\n9 0\n" + "".join(
        f"{i} 0.5 0.25 {2.0 + i} 0 0 0 0 0 0 0\n" for i in range(1, 10))
    d, _ = dk.read_reference_deck(_write(tmp_path, _files(init)))
    np.testing.assert_array_equal(d.U0, np.full(9, 0.5))
    np.testing.assert_array_equal(d.S0, 2.0 + np.arange(1, 10))
    init = "# 0.0 x 40000 U\n9 0\n" + "".join(f"{i} 0.5\n" for i in range(1, 10))
    d, _ = dk.read_reference_deck(_write(tmp_path, _files(init)))
    np.testing.assert_array_equal(d.S0, d.z)
    np.testing.assert_array_equal(d.V0, np.zeros(9))


@pytest.mark.parametrize("edit,msg", [
    (("mat.dat", "$ROUGHNESS 7 5 0.035", "$ROUGHNESS 7 3 0.035"), "Manning"),
    (("mat.dat", "$ROUGHNESS 7 5", "$ROUGHNESS 6 5"), "not defined"),
    (("tm.dat", "$TM_LINE OUTLET 9 1.4 0 0 0 0.0\n",
      "$TM_LINE OUTLET 9 1.4 0 0 0 0.0\n$TM_LINE OUTLET 9 1.5 1 1 0 0.0\n"), "profile"),
    (("tm.dat", "$TM_LINE OUTLET 9", "$TM_LINE SLIP 9"), "not supported"),
    (("init.dat", "9 0\n", "8 0\n"), "wrong number of nodes"),
    (("input.dat", "$ASC_INITFILE", "$BIN_INITFILE"), "only ASCII"),
    (("input.dat", "$RELAX 2", "$RELAX 1"), "relaxation method 1"),
    (("input.dat", "$RELAX 2", "$RELAX 5"), "relaxation method 5"),
    (("input.dat", "$RELAX 2\n", "$RELAX 2\n$DRYREW 1 5\n"), "drying and rewetting method 1"),
    (("input.dat", "$RELAX 2\n", "$RELAX 2\n$DRYREW 3 5\n"), "drying and rewetting method 3"),
    # the time-step file: one boundary set with the flow cycle alone, and its Newton limit
    (("tm.dat", "$TM_CYCLE 1\n", "$TM_CYCLE 90 1\n"), "only the flow cycle"),
    (("tm.dat", "$TM_CYCLE 1\n", ""), "no \\$TM_CYCLE"),
    (("tm.dat", "$TM_MAXITER 6\n", ""), "no \\$TM_MAXITER"),
    (("tm.dat", "$TM_LINE OUTLET 9 1.4 0 0 0 0.0\n",
      "$TM_LINE OUTLET 9 1.4 0 0 0 0.0\n$TM_STEP_NO 5\n$TM_CYCLE 1\n"), "more than one boundary set"),
    (("tm.dat", "$TM_MAXITER 6\n", "$TM_MAXITER 6\n$TM_NODE QT 3 1.0 2.0\n"), "TM_NODE is not"),
    (("tm.dat", "$TM_MAXITER 6\n", "$TM_MAXITER 6\n$TM_BOUND_LINE 4 8\n"), "TM_BOUND_LINE is not"),
    (("tm.dat", "$TM_MAXITER 6\n", "$TM_MAXITER 6\n$TM_DISPERSION 1\n"), "DISPERSION 1"),
])
def test_reader_rejects_what_the_path_does_not_support(tmp_path, edit, msg):
    init = "# 0.0 x 40000 U V S\n9 0\n" + "".join(f"{i} 0 0 1.0\n" for i in range(1, 10))
    files = _files(init)
    name, old, new = edit
    assert old in files[name]
    files[name] = files[name].replace(old, new)
    with pytest.raises(dk.DeckFormatError, match=msg):
        dk.read_reference_deck(_write(tmp_path, files))


def test_drying_method_check_takes_the_flow_cycles_newton_limit(tmp_path):
    """Drying and rewetting by a method other than 2 is refused only where the flow cycle
    would call it (its frequency below the flow cycle's $TM_MAXITER entry, EqsUVS2D.cpp:201);
    method 2 inside the cycle is read with its parameters."""
    init = "# 0.0 x 40000 U V S\n9 0\n" + "".join(f"{i} 0 0 1.0\n" for i in range(1, 10))
    for dry, maxiter, ok in (("3 10", "6", True), ("3 10", "20 1", False), ("2 3", "20 1", True)):
        files = _files(init)
        files["input.dat"] = files["input.dat"].replace("$RELAX 2\n", f"$RELAX 2\n$DRYREW {dry}\n")
        files["tm.dat"] = files["tm.dat"].replace("$TM_MAXITER 6\n", f"$TM_MAXITER {maxiter}\n")
        path = _write(tmp_path, files)
        if ok:
            _, run = dk.read_reference_deck(path)
            assert run["dryrew"]["method"] == int(dry[0])
            assert run["params"]["dry_method"] == int(dry[0])
        else:
            with pytest.raises(dk.DeckFormatError, match="method 3"):
                dk.read_reference_deck(path)


def test_project_keywords_as_the_reference_scans_them(tmp_path):
    """$CONVERGENCE, $RELAX, $DRYREW and $MINMAX (Project.cpp:1780-1852): fields in order
    up to the first that does not convert, the reference's defaults for the rest, the
    rewetting limit raised to the drying limit; drying and rewetting switched off
    (method 0) may name any frequency."""
    init = "# 0.0 x 40000 U V S\n9 0\n" + "".join(f"{i} 0 0 1.0\n" for i in range(1, 10))
    files = _files(init)
    files["input.dat"] = files["input.dat"].replace(
        "$RELAX 2\n", "$RELAX 2 0.001 0.1 5.0 x 7.0\n$CONVERGENCE 1e-6\n"
                       "$DRYREW 0 3 0.01 0.002\n$MINMAX 0 0 0 50\n")
    _, run = dk.read_reference_deck(_write(tmp_path, files))
    assert run["params"] == dict(relax_method=2, max_delta_uv=5.0, max_delta_s=0.1,
                                 conv_uv=1e-6, conv_s=1e-4, dry_method=0, dry_limit=0.01,
                                 max_us=50.0, rewet_limit=0.01, count_down=5)
    assert run["dryrew"] == dict(method=0, freq=3, dry_limit=0.01, rewet_limit=0.01,
                                 rewet_passes=100, count_down=5)
    # $MAX_Us sets maxUs as $MINMAX's fourth field does; the later line wins
    # (Project.cpp:1786, 1830)
    for extra, want in (("$MINMAX 0 0 0 50\n$MAX_Us 5.0\n", 5.0),
                        ("$MAX_Us 5.0\n$MINMAX 0 0 0 50\n", 50.0),
                        ("$MAX_Us 5.0\n$MINMAX 0 0\n", 5.0)):
        files = _files(init)
        files["input.dat"] = files["input.dat"].replace("$RELAX 2\n", "$RELAX 2\n" + extra)
        _, run = dk.read_reference_deck(_write(tmp_path, files))
        assert run["params"]["max_us"] == want, extra


def test_relaxation_method_3_takes_relax_time_from_the_time_interval(tmp_path):
    """Without $RELAX the reference's default method 3 (Project.cpp:86) runs time-relaxed
    Newton steps; relaxTimeFlow is $TM_INTERVAL's second value, and the time interval when
    that is not positive or exceeds it (Timeint.cpp:697-719)."""
    init = "# 0.0 x 40000 U V S\n9 0\n" + "".join(f"{i} 0 0 1.0\n" for i in range(1, 10))
    files = _files(init)
    assert "$TM_INTERVAL" in files["tm.dat"]
    for tmi, want in (("100.0 5.0 0.0", 5.0), ("100.0 0.0 0.0", 100.0), ("100.0 500 0", 100.0),
                      ("100.0", 100.0)):
        f = dict(files)
        f["input.dat"] = f["input.dat"].replace("$RELAX 2\n", "")
        f["tm.dat"] = re.sub(r"\$TM_INTERVAL[^\n]*", "$TM_INTERVAL " + tmi, f["tm.dat"])
        _, run = dk.read_reference_deck(_write(tmp_path, f))
        p = run["params"]
        assert (p["relax_method"], p["relax_time"], p["interval"]) == (3, want, 100.0), tmi
        assert (p["max_delta_uv"], p["max_delta_s"]) == (0.1, 0.1)     # the defaults


def test_time_step_lists_are_read_per_cycle(tmp_path):
    """$TM_MAXITER / $TM_TURBULENCE / $TM_STATIONARY hold one value per cycle of the set
    (Timeint.cpp:2021-2135): the flow cycle's is the first token, whatever follows."""
    init = "# 0.0 x 40000 U V S\n9 0\n" + "".join(f"{i} 0 0 1.0\n" for i in range(1, 10))
    files = _files(init)
    files["tm.dat"] = files["tm.dat"].replace("$TM_MAXITER 6\n", "$TM_MAXITER 4 9\n").replace(
        "$TM_TURBULENCE 1\n", "$TM_TURBULENCE 2 1\n$TM_STATIONARY 1 0\n")
    _, run = dk.read_reference_deck(_write(tmp_path, files))
    assert (run["maxiter"], run["turb"], run["stationary"]) == (4, 2, 1)


def test_golden_deck_project_keywords_are_the_solve_defaults(tmp_path):
    """The golden decks' $RELAX / $CONVERGENCE / $DRYREW give the parameters the oracle and
    the device solve use by default."""
    g = load_golden("tri20x4")
    dk.write_reference_deck(golden_deck(g), str(tmp_path))
    _, run = dk.read_reference_deck(str(tmp_path / "input.dat"))
    assert run["params"] == dict(relax_method=2, max_delta_uv=10.0, max_delta_s=10.0,
                                 conv_uv=1e-7, conv_s=1e-7, dry_method=2, dry_limit=0.001,
                                 max_us=100.0, rewet_limit=0.005, count_down=5)


@pytest.mark.parametrize("exact", [True, False])
def test_restart_file_reads_back_as_initial_state(tmp_path, exact):
    """A state written as the reference's ASCII restart (header with the comma-separated
    column list, eleven node values, element lines) is an initial file: read back exactly
    with %.17g, to the reference's printed 7 digits otherwise."""
    d = dk.channel(nx=10, ny=3)
    rng = np.random.default_rng(5)
    U, V = rng.standard_normal(d.n_nodes), rng.standard_normal(d.n_nodes)
    S = d.z + 1.0 + 0.1 * rng.random(d.n_nodes)
    dk.write_reference_deck(d, str(tmp_path))
    dk.write_reference_restart(str(tmp_path / "init.dat"), d, U, V, S, time="3600.0",
                               exact=exact)
    header = (tmp_path / "init.dat").read_text().splitlines()[0]
    assert header.split()[1:4] == ["3600.0", "Release", "40700"]
    r, _ = dk.read_reference_deck(str(tmp_path / "input.dat"))
    if exact:
        for a, b in ((r.U0, U), (r.V0, V), (r.S0, S)):
            np.testing.assert_array_equal(a, b)
    else:
        for a, b in ((r.U0, U), (r.V0, V), (r.S0, S)):
            np.testing.assert_allclose(a, b, rtol=5e-7, atol=0)
