"""Write the reference's BLAS golden vectors as JSON fixtures.

Run in a container that has the reference checkout (it reads the Zig test
files as text; nothing is compiled, imported or executed):

    python tests/golden/extract_blas_fixtures.py /root/reference

It reads the real-valued (f64) part of

    test/linalg/blas/gemm.zig:6-260   (8 chained calls, exact)
    test/linalg/blas/trsm.zig:6-892   (32 chained calls, rel 1e-7)
    test/linalg/blas/syrk.zig:6-287   (8 chained calls, exact)
    test/linalg/blas/gemv.zig:6-320   (8 calls, +-2 strides, exact)
    test/linalg/blas/trsv.zig:6-420   (16 calls, +-2 strides, rel 1e-7)

and records, in order, every buffer initialisation, every call with its
argument tokens and every expected element. The loop-generated vectors of
iamax.zig:6-28, scal.zig:6-40 and swap.zig:6-48 are recorded as the
formulas those tests state (written out below with their line numbers).
Output: tests/golden/blas_fixtures.json.
"""
from __future__ import annotations

import json
import os
import re
import sys

CALL_RE = re.compile(r"^\s*(gemm|trsm|syrk|gemv|trsv)\((.*)\)\s*catch unreachable;")
MEMCPY_RE = re.compile(r"^\s*@memcpy\((\w+)\.ptr,\s*&\[_\]f64\{")
EXPECT_RE = re.compile(
    r"expect(Equal|ApproxEqRel)\(([-0-9.eE+]+),\s*(\w+)\[(\d+)\](?:,\s*([-0-9.eE+]+))?\)")
SCALAR_RE = re.compile(r"^\s*const (\w+)(?::\s*f64)?\s*=\s*([-0-9.]+);")


def parse_file(path: str):
    events = []
    scalars = {}
    lines = open(path).read().splitlines()
    i = 0
    start = None
    while i < len(lines):
        line = lines[i]
        if start is None:
            if line.startswith("test "):
                start = i + 1
            i += 1
            continue
        # the complex-valued half of each test begins at the first cf64 value
        if "cf64.init" in line or "alloc(cf64" in line:
            break
        m = SCALAR_RE.match(line)
        if m:
            scalars[m.group(1)] = float(m.group(2))
            i += 1
            continue
        m = MEMCPY_RE.match(line)
        if m:
            name = m.group(1)
            vals = []
            i += 1
            while not lines[i].strip().startswith("});"):
                tok = lines[i].strip().rstrip(",")
                if tok:
                    vals.append(float(tok))
                i += 1
            events.append({"op": "set", "name": name, "values": vals, "line": i + 1})
            i += 1
            continue
        m = CALL_RE.match(line)
        if m:
            args = [a.strip() for a in m.group(2).split(",")]
            args = [a for a in args if a and a != ".{}"]
            events.append({"op": "call", "fn": m.group(1), "args": args, "line": i + 1})
            i += 1
            continue
        m = EXPECT_RE.search(line)
        if m:
            events.append({
                "op": "expect", "name": m.group(3), "index": int(m.group(4)),
                "value": float(m.group(2)),
                "rel": float(m.group(5)) if m.group(5) else 0.0, "line": i + 1})
        i += 1
    return {"source": os.path.relpath(path, os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(path))))),
            "scalars": scalars, "events": events}


def vector_fixtures():
    n = 1000
    return {
        # iamax.zig:9-27: zeros with ones at 127 and 456; first index wins.
        "iamax": {
            "source": "test/linalg/blas/iamax.zig:6-28",
            "n": n, "ones_at": [127, 456],
            "cases": [{"n": n, "incx": 1, "expect": 127}, {"n": n // 2, "incx": 2, "expect": 228}],
        },
        # scal.zig:19-40: x1 = i+1, scal(n, 2, x1, 1) -> 2(i+1);
        # x3 = i+1, scal(n/2, 2, x3, 2) -> even i doubled, odd i unchanged.
        "scal": {
            "source": "test/linalg/blas/scal.zig:6-40",
            "n": n, "alpha": 2.0,
            "cases": [{"n": n, "incx": 1}, {"n": n // 2, "incx": 2}],
        },
        # swap.zig:20-48: x1 = i+1, x2 = i+2, x3 = n-i+1, x4 = i+2.
        "swap": {
            "source": "test/linalg/blas/swap.zig:6-48",
            "n": n,
            "cases": [
                {"n": n, "incx": 1, "incy": 1, "x": "i+1", "y": "i+2",
                 "x_after": "i+2", "y_after": "i+1"},
                {"n": n, "incx": 1, "incy": -1, "x": "i+1", "y": "n-i+1",
                 "x_after": "i+2", "y_after": "n-i"},
                {"n": n // 2, "incx": 2, "incy": 2, "x": "i+1", "y": "i+2",
                 "x_after": "i%2==0 ? i+2 : i+1", "y_after": "i%2==0 ? i+1 : i+2"},
            ],
        },
    }


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    base = os.path.join(ref, "test", "linalg", "blas")
    names = ("gemm", "trsm", "syrk", "gemv", "trsv")
    out = {name: parse_file(os.path.join(base, name + ".zig")) for name in names}
    for name, fx in out.items():
        fx["source"] = f"test/linalg/blas/{name}.zig"
    out.update(vector_fixtures())
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blas_fixtures.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    for name in names:
        ev = out[name]["events"]
        print(name, "calls", sum(e["op"] == "call" for e in ev), "expects", sum(e["op"] == "expect" for e in ev))


if __name__ == "__main__":
    main()
