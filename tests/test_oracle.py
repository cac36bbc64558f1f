"""Pins the oracle against what mathematics and independent libraries fix,
so that a dropped term, a wrong index or a transposed operand fails here
rather than being copied into both the oracle and the product:

  * SpMV (utils.h:98-111): scipy.sparse CSR product, a dense numpy product
    on tiny inputs, e_j column probes (exact), linearity, and the reference's
    own check_results_consistency (utils.h:85) where oracle/_ref is built;
  * builder: block ids by floor division, first-appearance block order,
    per-block entry counts, the gathered-column restore map as numpy.unique
    per block row, gb_row_ptr as a prefix sum (coo2block.h:7-146);
  * dispatch policy and slab layout: the product's exports equal the
    oracle's, regions are aligned, ordered and disjoint
    (cb-spmv-dcu.hpp:1100-1348, 1428-1450).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from cbspmv import Plan, gen, read_mtx
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _rand(m, n, per_row, seed, **kw):
    r, c, v = gen.random_coo(m, n, per_row, seed=seed, **kw)
    return m, n, r, c, v


CASES = [
    _rand(300, 400, 5, 1),
    _rand(1000, 1000, 12, 2, dup=True),
    _rand(777, 5000, 3, 3, empty_rows=0.3),
    _rand(64, 64, 40, 4, unit=True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_spmv_seq_matches_scipy(case):
    m, n, r, c, v = CASES[case]
    x = np.random.default_rng(case).uniform(-1, 1, n)
    y = pyoracle.spmv_seq(m, r, c, v, x)
    A = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsr()
    y_ref = A @ x
    assert pyoracle.rel_l2(y, y_ref) < 1e-14


def test_spmv_seq_dense_tiny():
    rng = np.random.default_rng(7)
    for m, n in ((1, 1), (3, 5), (17, 9)):
        D = rng.uniform(-1, 1, (m, n)) * (rng.random((m, n)) < 0.5)
        r, c = np.nonzero(D)
        x = rng.uniform(-1, 1, n)
        y = pyoracle.spmv_seq(m, r, c, D[r, c], x)
        np.testing.assert_allclose(y, D @ x, rtol=1e-14, atol=1e-15)


def test_spmv_seq_column_probes_exact():
    """A e_j is column j of A (duplicates summed), exactly."""
    m, n, r, c, v = CASES[1]
    A = sp.coo_matrix((v, (r, c)), shape=(m, n)).tocsc()
    for j in (0, 1, n // 2, n - 1):
        e = np.zeros(n)
        e[j] = 1.0
        y = pyoracle.spmv_seq(m, r, c, v, e)
        col = np.zeros(m)
        sl = slice(A.indptr[j], A.indptr[j + 1])
        # summing duplicates in input order is what both do for a single column
        for i, val in zip(A.indices[sl], A.data[sl]):
            col[i] = val
        mask = np.zeros(m, bool)
        mask[r[c == j]] = True
        assert np.all(y[~mask] == 0.0)
        np.testing.assert_allclose(y, col, rtol=1e-15, atol=0)


def test_spmv_seq_linearity():
    m, n, r, c, v = CASES[0]
    rng = np.random.default_rng(9)
    x1, x2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    a, b = 0.75, -2.5
    lhs = pyoracle.spmv_seq(m, r, c, v, a * x1 + b * x2)
    rhs = a * pyoracle.spmv_seq(m, r, c, v, x1) + b * pyoracle.spmv_seq(m, r, c, v, x2)
    assert pyoracle.rel_l2(lhs, rhs) < 1e-14


def test_csr_omp_equals_seq_bitwise():
    for m, n, r, c, v in CASES:
        x = np.random.default_rng(m).uniform(-1, 1, n)
        a = pyoracle.spmv_seq(m, r, c, v, x)
        b = pyoracle.spmv_csr_omp(m, r, c, v, x, 4)
        assert np.array_equal(a, b)


@pytest.mark.skipif(not pyoracle.ref_available(), reason="oracle/_ref not built (reference checkout absent)")
def test_reference_check_accepts_oracle_and_rejects_perturbation():
    m, n, r, c, v = CASES[1]
    x = np.random.default_rng(3).uniform(-1, 1, n)
    y = pyoracle.spmv_seq(m, r, c, v, x)
    assert pyoracle.ref_check(m, n, r, c, v, x, y)
    bad = y.copy()
    bad[m // 2] += 1e-3 * (abs(bad[m // 2]) + 1)
    assert not pyoracle.ref_check(m, n, r, c, v, x, bad)


# ---- builder closed forms


@pytest.mark.parametrize("bs", [8, 16, 32])
def test_block_ids_no_gather(bs):
    """Without gathering, block (bi, bj) = (row // bs, col // bs), numbered
    by first appearance, with the entry count of the block."""
    m = n = 600
    rng = np.random.default_rng(bs)
    # blocks of >= 9 entries so gathering stays off
    r, c = [], []
    for bi in range(m // bs):
        for bj in rng.choice(n // bs, 3, replace=False):
            cells = rng.choice(bs * bs, 12, replace=False)
            r += list(bi * bs + cells // bs)
            c += list(bj * bs + cells % bs)
    r, c = np.array(r), np.array(c)
    o = np.lexsort((rng.random(r.size), r))
    r, c = r[o].astype(np.int32), c[o].astype(np.int32)
    v = rng.uniform(-1, 1, r.size)
    ob = pyoracle.OracleBuild(m, n, r, c, v, bs=bs)
    assert ob.scalar("gather_flag") == 0
    key = (r // bs).astype(np.int64) * (n // bs + 1) + c // bs
    uniq, first, counts = np.unique(key, return_index=True, return_counts=True)
    order = np.argsort(first, kind="stable")
    want_r = (uniq[order] // (n // bs + 1)).astype(np.int32)
    want_c = (uniq[order] % (n // bs + 1)).astype(np.int32)
    got_r, got_c = ob.array("gb_row_idx"), ob.array("gb_col_idx")
    assert np.array_equal(got_r, want_r) and np.array_equal(got_c, want_c)
    # 12 distinct cells each: entry counts survive (no duplicates to merge)
    assert np.array_equal(ob.array("nnz_per_blk"), counts[order].astype(np.int32))
    ptr = np.concatenate([[0], np.cumsum(np.bincount(want_r, minlength=(m + bs - 1) // bs))]).astype(np.int32)
    assert np.array_equal(ob.array("gb_row_ptr"), ptr)


def test_gather_restore_map_is_unique_columns_per_block_row():
    m, n, r, c, v = _rand(2000, 3000, 4, 11)
    ob = pyoracle.OracleBuild(m, n, r, c, v)
    assert ob.scalar("gather_flag") == 1
    bc, off = ob.array("block_cols"), ob.array("block_cols_offset")
    nbr = (m + 15) // 16
    assert off.size == nbr + 1 and off[0] == 0
    for bi in range(nbr):
        sel = (r // 16) == bi
        assert np.array_equal(bc[off[bi]:off[bi + 1]], np.unique(c[sel]))
    # every block's gathered column span points inside its block row's list
    gr, gc = ob.array("gb_row_idx"), ob.array("gb_col_idx")
    for bi, bj in zip(gr, gc):
        assert bj * 16 < off[bi + 1] - off[bi]


def test_gather_ratio_value():
    m, n, r, c, v = _rand(2000, 3000, 4, 12)
    ob = pyoracle.OracleBuild(m, n, r, c, v)
    # the ratio is computed on the ungathered blocks
    key = np.unique((r // 16).astype(np.int64) * 1000 + c // 16, return_counts=True)[1]
    assert ob.scalar("analyzed_blocks") == key.size
    assert ob.scalar("super_sparse_blocks") == int((key <= 8).sum())
    assert ob.scalar("super_sparse_ratio") == (key <= 8).sum() / key.size


def test_payload_sums_to_matrix():
    """Summing the payload values per (row, col) gives back the matrix."""
    for m, n, r, c, v in CASES:
        ob = pyoracle.OracleBuild(m, n, r, c, v)
        assert np.isclose(ob.array("pl_val").sum(), v.sum(), rtol=1e-12, atol=1e-9)


# ---- policy and layout


def _policy_cases():
    yield "1138_bus", read_mtx(os.path.join(HERE, "golden", "1138_bus.mtx"))
    d = gen.poisson_4096()
    yield "poisson_4096", (d["m"], d["n"], d["row"], d["col"], d["val"])
    for i, (m, n, r, c, v) in enumerate(CASES):
        yield f"case{i}", (m, n, r, c, v)


@pytest.mark.parametrize("name,mat", list(_policy_cases()), ids=lambda x: x if isinstance(x, str) else "")
@pytest.mark.parametrize("row_owned", [True, False])
def test_policy_and_layout_match_oracle(name, mat, row_owned):
    m, n, r, c, v = mat
    p = Plan(m, n, r, c, v, row_owned=row_owned)
    ob = pyoracle.OracleBuild(m, n, r, c, v, row_owned=row_owned)
    pol = ob.policy()
    assert p.export("policy").tolist() == [pol[k] for k in
                                           ("unit_values", "use_pattern", "use_unrolled", "use_small_direct",
                                            "launch_waves", "grid", "aggregate_coo", "pattern_count")]
    regions = ["gb_row_idx", "gb_col_idx", "segment_rows", "nnz_per_blk", "type_per_blk", "pattern_row_ptr",
               "block_cols", "block_addr", "x", "y", "output_epochs", "total"]
    offs = []
    for reg in regions:
        a, b = p.layout_offset(reg), ob.layout(reg)
        assert a == b, reg
        if a >= 0:
            offs.append((reg, a))
    assert [o for _, o in offs] == sorted(o for _, o in offs)
    for reg, o in offs:
        if reg in ("block_addr", "x", "y"):
            assert o % 8 == 0
        elif reg != "type_per_blk" and reg != "total":
            assert o % 4 == 0
    assert np.array_equal(p.export("image"), ob.array("image"))


def test_reference_x_matches_std_mt19937(tmp_path):
    """gen.reference_x reproduces the reference driver's x
    (std::mt19937(42), uniform_real_distribution<double>(0, 1),
    main.hip.cpp:58-66): compared with a C++ program using the standard
    library directly."""
    import subprocess

    from cbspmv.gen import reference_x

    src = tmp_path / "x.cpp"
    src.write_text('#include <random>\n#include <cstdio>\nint main(){std::mt19937 e(42);'
                   'std::uniform_real_distribution<double> u(0.0, 1.0);'
                   'for(int i=0;i<1000;i++) std::printf("%.17g\\n", u(e));}\n')
    exe = tmp_path / "x"
    subprocess.run(["g++", "-O1", str(src), "-o", str(exe)], check=True)
    want = np.array([float(t) for t in subprocess.run([str(exe)], capture_output=True, text=True,
                                                      check=True).stdout.split()])
    assert np.array_equal(reference_x(1000), want)
