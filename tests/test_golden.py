"""Builder and reader parity against the committed golden digests.

tests/golden/golden_ref.json holds SHA-256 digests of every GLOBAL_BLOCK
array and payload stream that the reference builder (coo2block_gather /
coo2block_comp, cb-spmv-dcu/src/coo2block.h:628 / :605) produced for the
fixture matrices under eight compile-time configurations, plus its scalar
statistics, and digests of the reference reader's output
(read_matrix_coo_data, cb-spmv-dcu/src/io/mmio_highlevel.h:313) for
1138_bus.mtx and heart2.mtx.  tests/golden/make_golden.py wrote it.  These
tests need neither the reference nor a GPU: both the product builder and the
oracle must reproduce every digest bit for bit.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from cbspmv import Plan, gen, read_mtx
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
with open(os.path.join(GOLD, "golden_ref.json")) as _fh:
    GOLDEN = json.load(_fh)

# load_stddev_* come from a -ffast-math build of the reference; the last bit
# of a sum of squares may differ, so those two compare to 1e-12 relative.
INEXACT = {"load_stddev_before", "load_stddev_after"}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


_MATS = {}


def matrix(name):
    if name not in _MATS:
        if name == "1138_bus":
            m, n, r, c, v = read_mtx(os.path.join(GOLD, "1138_bus.mtx"))
        elif name == "heart2":
            z = np.load(os.path.join(GOLD, "heart2_pattern.npz"))
            ptr = z["row_ptr"].astype(np.int64)
            m, n = int(z["m"]), int(z["n"])
            r = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
            c, v = z["col"].astype(np.int32), np.ones(r.size)
        elif name == "poisson_4096":
            d = gen.poisson_4096()
            m, n, r, c, v = d["m"], d["n"], d["row"], d["col"], d["val"]
        elif name == "random_dup":
            m, n = 700, 900
            r, c, v = gen.random_coo(m, n, 12, seed=101, dup=True)
        elif name == "random_40k":
            m, n = 5000, 5000
            r, c, v = gen.random_coo(m, n, 30, seed=102)
        else:
            raise KeyError(name)
        _MATS[name] = (m, n, r, c, v)
    return _MATS[name]


def _id(rec):
    return (f"{rec['matrix']}-bs{rec['bs']}-w{rec['waves']}-ro{int(rec['row_owned'])}"
            f"-f{int(rec['force_coo64'])}{'-comp' if rec['comp'] else ''}")


BUILDS = GOLDEN["builds"]


def _check(rec, array_fn, scalar_fn):
    bad = []
    for name, want in rec["arrays"].items():
        a = array_fn(name)
        if a.size != rec["sizes"][name] or digest(a) != want:
            bad.append(f"{name} (size {a.size} vs {rec['sizes'][name]})")
    assert not bad, "arrays differ from the reference builder: " + ", ".join(bad)
    for name, want in rec["scalars"].items():
        got = scalar_fn(name)
        if name in INEXACT:
            assert got == pytest.approx(want, rel=1e-12, abs=1e-300), name
        else:
            assert got == want, name


@pytest.mark.parametrize("rec", BUILDS, ids=_id)
def test_product_builder_matches_reference(rec):
    m, n, r, c, v = matrix(rec["matrix"])
    p = Plan(m, n, r, c, v, block_size=rec["bs"], waves=rec["waves"], row_owned=rec["row_owned"],
             force_coo64=rec["force_coo64"], builder="comp" if rec["comp"] else "gather")
    st = p.stats()
    try:
        _check(rec, p.export, lambda s: float(st[s]))
    finally:
        p.close()


@pytest.mark.parametrize("rec", BUILDS, ids=_id)
def test_oracle_builder_matches_reference(rec):
    m, n, r, c, v = matrix(rec["matrix"])
    o = pyoracle.OracleBuild(m, n, r, c, v, bs=rec["bs"], waves=rec["waves"], row_owned=rec["row_owned"],
                             force_coo64=rec["force_coo64"], comp=rec["comp"])
    try:
        _check(rec, o.array, o.scalar)
    finally:
        o.close()


def test_reader_matches_reference_reader():
    want = GOLDEN["readers"]["1138_bus"]
    m, n, r, c, v = read_mtx(os.path.join(GOLD, "1138_bus.mtx"))
    assert (m, n, r.size) == (want["m"], want["n"], want["nnz"])
    assert digest(r.astype(np.int32)) == want["row"]
    assert digest(c.astype(np.int32)) == want["col"]
    assert digest(v.astype(np.float64)) == want["val"]


def test_heart2_fixture_is_reference_reader_output():
    """heart2_pattern.npz stores the reference reader's heart2.mtx output as
    CSR pattern (values are forced to 1 by that reader)."""
    want = GOLDEN["readers"]["heart2"]
    m, n, r, c, v = matrix("heart2")
    assert (m, n, r.size) == (want["m"], want["n"], want["nnz"])
    assert digest(r) == want["row"]
    assert digest(c) == want["col"]
    assert digest(v) == want["val"]


def test_known_answers_1138_bus():
    """Statistics the reference prints for 1138_bus (bs 16, 4 waves)."""
    m, n, r, c, v = matrix("1138_bus")
    p = Plan(m, n, r, c, v)
    st = p.stats()
    assert (m, n, r.size) == (1138, 1138, 4054)
    assert st["super_sparse_blocks"] == 589 and st["analyzed_blocks"] == 692
    assert st["gather_flag"] == 1 and st["direct_coo_columns"] == 1 and st["compact_coo_indices"] == 1
    assert st["gb_nnzb"] == 196
    assert p.export("segment_rows").size == 72
    assert p.export("pattern_row_ptr").size - 1 == 1130
    pol = p.export("policy").tolist()
    assert pol == [1, 0, 0, 1, 4, 72, 0, 1130]
    p.close()


def test_known_answers_heart2():
    m, n, r, c, v = matrix("heart2")
    p = Plan(m, n, r, c, v)
    st = p.stats()
    assert r.size == 682797
    assert st["super_sparse_blocks"] == 222 and st["analyzed_blocks"] == 5127
    types = p.export("type_per_blk")
    base = types & 0x7F  # COO 0, CSR 1, DENSE 3; bit 7 marks uniform rows
    assert [int((base == t).sum()) for t in (0, 1, 3)] == [1131, 2764, 1232]
    assert int((types >= 0x80).sum()) == 2143
    assert p.export("segment_rows").size == 974
    assert p.export("pattern_row_ptr").size - 1 == 146
    assert p.export("policy").tolist() == [1, 1, 1, 0, 9, 17, 1, 146]
    p.close()
