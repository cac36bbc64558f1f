"""Writes tests/golden/golden_ref.json: SHA-256 digests of the reference
builder's output (every GLOBAL_BLOCK array and the payload streams) and its
scalar statistics, for the committed fixture matrices and several
compile-time configurations, plus digests of the reference reader's output
for the .mtx fixtures.

Run in a checkout where the reference is present and oracle/_ref is built
(`make -C oracle`): it calls the reference's own code through oracle/_ref.
The JSON it writes is data (digests and numbers) and is what the tests use;
the tests do not need the reference.

  python3 tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "cb-sparse-dcu_amd"))

from oracle import pyoracle  # noqa: E402

REF_DATA = "/root/reference/cb-spmv-dcu/data"
CONFIGS = [  # (block size, waves per workgroup, row owned, force COO64)
    (16, 4, True, False),
    (16, 4, False, False),
    (16, 8, True, False),
    (32, 4, True, False),
    (32, 8, False, False),
    (8, 4, True, False),
    (64, 4, True, False),
    (16, 4, True, True),
]


def digest(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def fixture_matrices():
    """Name -> (m, n, row, col, val) for every golden matrix."""
    from cbspmv import gen

    mats = {}
    m, n, r, c, v = pyoracle.ref_read_mtx(os.path.join(HERE, "1138_bus.mtx"))
    mats["1138_bus"] = (m, n, r, c, v)
    z = np.load(os.path.join(HERE, "heart2_pattern.npz"))
    ptr = z["row_ptr"].astype(np.int64)
    rows = np.repeat(np.arange(int(z["m"]), dtype=np.int32), np.diff(ptr))
    mats["heart2"] = (int(z["m"]), int(z["n"]), rows, z["col"].astype(np.int32), np.ones(rows.size))
    d = gen.poisson_4096()
    mats["poisson_4096"] = (d["m"], d["n"], d["row"], d["col"], d["val"])
    r, c, v = gen.random_coo(700, 900, 12, seed=101, dup=True)
    mats["random_dup"] = (700, 900, r, c, v)
    r, c, v = gen.random_coo(5000, 5000, 30, seed=102)
    mats["random_40k"] = (5000, 5000, r, c, v)
    return mats


def main():
    out = {"readers": {}, "builds": []}
    for name in ("1138_bus", "heart2"):
        src = os.path.join(REF_DATA, f"{name}.mtx")
        if os.path.exists(src):
            m, n, r, c, v = pyoracle.ref_read_mtx(src)
            out["readers"][name] = {"m": m, "n": n, "nnz": int(r.size), "row": digest(r.astype(np.int32)),
                                    "col": digest(c.astype(np.int32)), "val": digest(v.astype(np.float64))}
    for name, (m, n, r, c, v) in fixture_matrices().items():
        for bs, w, ro, f in CONFIGS:
            for comp in (False, True):
                if comp and (bs, w, f) != (16, 4, False):
                    continue
                b = pyoracle.RefBuild(m, n, r, c, v, bs=bs, waves=w, row_owned=ro, force_coo64=f, comp=comp)
                rec = {"matrix": name, "bs": bs, "waves": w, "row_owned": ro, "force_coo64": f, "comp": comp,
                       "arrays": {a: digest(b.array(a)) for a in pyoracle.ARRAY_NAMES},
                       "sizes": {a: int(b.array(a).size) for a in pyoracle.ARRAY_NAMES},
                       "scalars": {s: b.scalar(s) for s in pyoracle.SCALAR_NAMES}}
                out["builds"].append(rec)
                b.close()
    with open(os.path.join(HERE, "golden_ref.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"wrote {len(out['builds'])} builds, {len(out['readers'])} readers")


if __name__ == "__main__":
    main()
