"""Replays the reference's chained BLAS golden vectors against a backend.

The fixtures (tests/golden/blas_fixtures.json) were extracted from
test/linalg/blas/{gemm,trsm,syrk}.zig. A backend is any object with
gemm/trsm/syrk taking the CBLAS argument list with numpy float64 buffers.
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
FLOAT_POS = {"gemm": (6, 11), "trsm": (7,), "syrk": (5, 8)}


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


def replay(name, backend):
    """Run every call of fixture `name`; return a list of failure strings."""
    fx = load()[name]
    bufs = {}
    failures = []
    for ev in fx["events"]:
        if ev["op"] == "set":
            bufs[ev["name"]] = np.array(ev["values"], dtype=np.float64)
        elif ev["op"] == "call":
            args = _resolve(ev["fn"], ev["args"], fx["scalars"], bufs)
            getattr(backend, ev["fn"])(*args)
        else:
            got = bufs[ev["name"]][ev["index"]]
            exp = ev["value"]
            if ev["rel"] == 0.0:
                ok = got == exp
            else:
                # std.testing.expectApproxEqRel: |x - y| <= tol * max(|x|, |y|)
                ok = abs(got - exp) <= ev["rel"] * max(abs(got), abs(exp)) or got == exp
            if not ok:
                failures.append(f"{fx['source']}:{ev['line']} {ev['name']}[{ev['index']}] = {got!r}, expected {exp!r}")
    return failures
