"""Reader for the record stream written by oracle/ref_harness.cpp.

Each record is ``char tag[8]`` + ``int64 nbytes`` + payload.  ``read_dump`` groups the
records by Newton step (a ``STEP`` record opens a step; ``FINAL`` is the state after the
last step).  Used only by tests/golden/make_golden.py in the oracle container.
"""
from __future__ import annotations

import numpy as np

STATE_FIELDS = ["name", "flag", "kind", "rot", "eqU", "eqV", "eqS", "x", "y", "z", "zor",
                "U", "V", "S", "cf", "vt", "nx", "ny", "niox", "nioy", "bcU", "bcS",
                "r00", "r01", "r11", "r10", "uu", "uv", "vv", "dUdt", "dVdt", "dSdt",
                "Uo", "Vo", "So", "dUdto", "dVdto", "dSdto"]


def _records(path):
    data = open(path, "rb").read()
    p = 0
    while p < len(data):
        tag = data[p:p + 8].rstrip(b"\0").decode()
        n = int(np.frombuffer(data[p + 8:p + 16], np.int64)[0])
        yield tag, data[p + 16:p + 16 + n]
        p += 16 + n


def read_dump(path):
    steps = []
    final = None
    cur = None
    hdr = None
    for tag, buf in _records(path):
        if tag == "STEP":
            cur = {"elems": []}
            steps.append(cur)
        elif tag == "TSTEP":
            cur["tstep"] = int(np.frombuffer(buf, np.int64)[0])
        elif tag == "STATEHDR":
            hdr = np.frombuffer(buf, np.int64)
        elif tag in ("STATE", "FINAL"):
            a = np.frombuffer(buf, np.float64).reshape(int(hdr[0]), int(hdr[1]))
            st = {k: a[:, i].copy() for i, k in enumerate(STATE_FIELDS[:a.shape[1]])}
            if tag == "STATE":
                cur["state"] = st
            else:
                final = st
        elif tag == "ELEM":
            a = np.frombuffer(buf, np.float64)
            nnd = int(a[1])
            ndf = 3 * nnd
            cur["elems"].append(dict(ret=int(a[0]), nnd=nnd, bound=bool(a[2]),
                                     names=a[3:3 + nnd].astype(np.int64),
                                     force=a[3 + nnd:3 + nnd + ndf].copy(),
                                     estifm=a[3 + nnd + ndf:].reshape(ndf, ndf).copy()))
        elif tag == "PATHDR":
            h = np.frombuffer(buf, np.int64)
            cur["n"], cur["nnz"], cur["realpr"] = int(h[0]), int(h[1]), int(h[2])
        elif tag == "WIDTH":
            cur["width"] = np.frombuffer(buf, np.int32).copy()
        elif tag == "COLS":
            cur["cols"] = np.frombuffer(buf, np.int32).copy()
        elif tag in ("AVAL", "ASCALED", "ILU"):
            cur[tag.lower()] = np.frombuffer(buf, np.float64).copy()
        elif tag in ("BVEC", "BSCALED", "ILUAPPLY", "MULVEC", "XSOL"):
            cur[tag.lower()] = np.frombuffer(buf, np.float64).copy()
        elif tag == "SCALE":
            cur["scale"] = float(np.frombuffer(buf, np.float64)[0])
        elif tag == "XHDR":
            h = np.frombuffer(buf, np.int64)
            cur["neq"], cur["err"] = int(h[0]), int(h[1])
    return steps, final
