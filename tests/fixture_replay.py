"""Replays the reference's chained BLAS golden vectors against a backend.

The fixtures (tests/golden/blas_fixtures.json) were extracted from
test/linalg/blas/{gemm,trsm,syrk,gemv,trsv}.zig. A backend is any object with
gemm/trsm/syrk/gemv/trsv taking the CBLAS argument list with numpy buffers
(float64 unless `dtype` says otherwise).
"""
import json
import os

import numpy as np

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blas_fixtures.json")

ENUMS = {
    ".row_major": 101, ".col_major": 102,
    ".no_trans": 111, ".trans": 112, ".conj_trans": 113, ".conj_no_trans": 114,
    ".upper": 121, ".lower": 122, ".non_unit": 131, ".unit": 132,
    ".left": 141, ".right": 142,
}
FLOAT_POS = {"gemm": (6, 11), "trsm": (7,), "syrk": (5, 8), "gemv": (4, 9), "trsv": ()}


def load():
    with open(FIXTURES) as f:
        return json.load(f)


def _resolve(fn, args, scalars, bufs):
    out = []
    for pos, tok in enumerate(args):
        if tok in ENUMS:
            out.append(ENUMS[tok])
        elif tok.endswith(".ptr"):
            out.append(bufs[tok[:-4]])
        elif tok in scalars:
            v = scalars[tok]
            out.append(float(v) if pos in FLOAT_POS[fn] else int(v))
        else:
            out.append(float(tok) if pos in FLOAT_POS[fn] else int(tok))
    return out


def replay(name, backend, dtype=np.float64, rel_floor=0.0):
    """Run every call of fixture `name`; return a list of failure strings.
    rel_floor raises the relative tolerance of the approximate expectations
    (for a float32 backend); exact expectations stay exact."""
    fx = load()[name]
    bufs = {}
    failures = []
    for ev in fx["events"]:
        if ev["op"] == "set":
            bufs[ev["name"]] = np.array(ev["values"], dtype=dtype)
        elif ev["op"] == "call":
            args = _resolve(ev["fn"], ev["args"], fx["scalars"], bufs)
            getattr(backend, ev["fn"])(*args)
        else:
            got = float(bufs[ev["name"]][ev["index"]])
            exp = ev["value"]
            if ev["rel"] == 0.0:
                ok = got == exp
            else:
                # std.testing.expectApproxEqRel: |x - y| <= tol * max(|x|, |y|)
                tol = max(ev["rel"], rel_floor)
                ok = abs(got - exp) <= tol * max(abs(got), abs(exp)) or got == exp
            if not ok:
                failures.append(f"{fx['source']}:{ev['line']} {ev['name']}[{ev['index']}] = {got!r}, expected {exp!r}")
    return failures
