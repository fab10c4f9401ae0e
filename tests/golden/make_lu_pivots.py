"""Golden pivot vectors of zml's getrf for the large synthetic LU inputs.

The oracle's threaded closed form (bitwise equal to zml's blocked getrf, as
tests/test_oracle.py checks at small n) is run on the seeded matrices below
and the resulting 1-based pivot vector is stored with a few factor
checksums, so that the GPU tests can compare pivots at sizes the scalar
restatement would take hours for.

    python tests/golden/make_lu_pivots.py 16384

Matrix: column-major n x n, a_ij = -1 + 2 u(seed=42, i*n + j) (the oracle's
splitmix64 stream), the same matrix in either layout.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402


def main(n: int, seed: int = 42, threads: int = 0):
    a = O.fill_uniform(O.COL, n, n, seed)
    ip = np.zeros(n, np.int32)
    t0 = time.time()
    info = O.getrf_fast(n, n, a, n, ip, threads)
    dt = time.time() - t0
    lu = a.reshape(n, n).T  # row i, col j
    diag = np.diagonal(lu).copy()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"lu_pivots_n{n}_seed{seed}.npz")
    np.savez_compressed(out, ipiv=ip, diag=diag, info=np.int32(info))
    meta = {"n": n, "seed": seed, "info": int(info), "seconds": dt,
            "min_abs_u_diag": float(np.abs(diag).min()), "max_abs_u_diag": float(np.abs(diag).max())}
    print(json.dumps(meta))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 16384)
