"""Builder parity on crafted edge cases: the product builder, the oracle and
(where oracle/_ref is built) the reference builder itself must agree on every
GLOBAL_BLOCK array and payload stream, and the format decisions must land on
the thresholds the reference states (coo2block.h:154-380):

  * a block with <= 32 entries is COO; 33..64 is COO only when COO64 is on
    (forced, or fewer than half of the 33..64-entry blocks are row-balanced,
    max - min nonempty-row count <= 1); >= bs*bs*15/16 is DENSE; otherwise
    CSR;
  * column gathering when > 15% of the analysed blocks hold <= 8 entries;
    direct COO columns when gathering and n <= 65535; compact (12-bit)
    indices when also bs == 16, n <= 4095 and nnz < 10000;
  * the long-row flag when a COO block row holds more than 8 entries.
"""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle

NAMES = pyoracle.ARRAY_NAMES
COO, CSR, DENSE = 0, 1, 3


def blocks_matrix(m, n, specs, bs=16, seed=0, shuffle_rows=True):
    """Row-major COO with `k` distinct cells in block (bi, bj) for each
    (bi, bj, k) in specs; within a row the column order is shuffled."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for bi, bj, k in specs:
        h = min(bs, m - bi * bs)
        w = min(bs, n - bj * bs)
        cells = rng.choice(h * w, size=k, replace=False)
        rows.append(bi * bs + cells // w)
        cols.append(bj * bs + cells % w)
    r = np.concatenate(rows).astype(np.int64)
    c = np.concatenate(cols).astype(np.int64)
    key = rng.random(r.size) if shuffle_rows else c.astype(float)
    o = np.lexsort((key, r))
    r, c = r[o].astype(np.int32), c[o].astype(np.int32)
    v = rng.uniform(-1, 1, r.size)
    return r, c, v


def check_all(m, n, r, c, v, bs=16, waves=4, row_owned=True, force_coo64=False, comp=False):
    """Builds with product, oracle and (if present) reference; asserts all
    arrays and scalars agree and returns the product plan."""
    p = Plan(m, n, r, c, v, block_size=bs, waves=waves, row_owned=row_owned, force_coo64=force_coo64,
             builder="comp" if comp else "gather")
    o = pyoracle.OracleBuild(m, n, r, c, v, bs=bs, waves=waves, row_owned=row_owned, force_coo64=force_coo64,
                             comp=comp)
    builds = [("oracle", o.array, o.scalar)]
    ref = None
    if pyoracle.ref_available(bs, waves, row_owned, force_coo64):
        ref = pyoracle.RefBuild(m, n, r, c, v, bs=bs, waves=waves, row_owned=row_owned, force_coo64=force_coo64,
                                comp=comp)
        builds.append(("reference", ref.array, ref.scalar))
    st = p.stats()
    for who, arr, sca in builds:
        for name in NAMES:
            a, b = p.export(name), arr(name)
            assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b), f"{name} vs {who}"
        for s in pyoracle.SCALAR_NAMES:
            if s.startswith("load_stddev"):
                assert float(st[s]) == pytest.approx(sca(s), rel=1e-12, abs=1e-300), f"{s} vs {who}"
            else:
                assert float(st[s]) == sca(s), f"{s} vs {who}"
    o.close()
    if ref is not None:
        ref.close()
    return p


def types_by_block(p):
    """{(brow, bcol): type without the uniform bit} for the built plan."""
    rows = p.export("gb_row_idx")
    cols = p.export("gb_col_idx")
    t = p.export("type_per_blk") & 0x7F
    return {(int(a), int(b)): int(x) for a, b, x in zip(rows, cols, t)}


@pytest.mark.parametrize("k,want", [(1, COO), (8, COO), (32, COO), (33, COO), (64, COO), (65, CSR),
                                    (239, CSR), (240, DENSE), (256, DENSE)])
def test_format_threshold_no_gather(k, want):
    # gathering stays off: at most 1 of 12 blocks holds <= 8 entries
    specs = [(0, 0, k), (0, 1, 9), (1, 1, 12), (2, 0, 20)] + [(3, j, 10 + j) for j in range(4)] + \
        [(2, j, 15) for j in range(1, 4)] + [(1, 3, 9)]
    r, c, v = blocks_matrix(64, 64, specs, seed=k)
    p = check_all(64, 64, r, c, v)
    assert p.stats()["gather_flag"] == 0
    t = types_by_block(p)
    assert t[(0, 0)] == want  # 33..64: the only candidate, unbalanced -> COO64
    assert t[(0, 1)] == COO and t[(1, 1)] == COO


@pytest.mark.parametrize("balanced", [True, False])
def test_coo64_extension(balanced):
    """33..64-entry blocks stay CSR when at least half of them spread their
    entries evenly over rows (max - min row count <= 1), and become COO
    otherwise or when COO64 is forced."""
    rng = np.random.default_rng(5)
    rows, cols = [], []
    for b in range(6):
        if balanced:  # 3 entries in each of 16 rows = 48
            for rr in range(16):
                cc = rng.choice(16, 3, replace=False)
                rows += [rr] * 3
                cols += list(b * 16 + np.sort(cc))
        else:  # 40 entries piled into a few rows
            for rr, k in ((0, 16), (1, 16), (2, 8)):
                cc = rng.choice(16, k, replace=False)
                rows += [rr] * cc.size
                cols += list(b * 16 + np.sort(cc))
    r, c = np.array(rows), np.array(cols)
    o = np.lexsort((c, r))
    r, c = r[o].astype(np.int32), c[o].astype(np.int32)
    v = rng.uniform(-1, 1, r.size)
    p = check_all(16, 96, r, c, v)
    t = set(types_by_block(p).values())
    assert t == ({CSR} if balanced else {COO})
    p2 = check_all(16, 96, r, c, v, force_coo64=True)
    assert set(types_by_block(p2).values()) == {COO}


def _sparse_rows(m, n, per_row, seed, extra=()):
    r, c, v = gen.random_coo(m, n, per_row, seed=seed)
    if extra:
        er, ec = zip(*extra)
        r = np.concatenate([r, np.array(er, np.int32)])
        c = np.concatenate([c, np.array(ec, np.int32)])
        v = np.concatenate([v, np.ones(len(extra))])
        o = np.lexsort((c, r))
        r, c, v = r[o], c[o], v[o]
    return r, c, v


@pytest.mark.parametrize("n,nnz_target,compact", [(4095, 3000, True), (4096, 3000, False),
                                                   (4095, 12000, False)])
def test_compact_indices_thresholds(n, nnz_target, compact):
    m = 2000
    r, c, v = gen.random_coo(m, n, max(1, nnz_target // m), seed=n + nnz_target)
    p = check_all(m, n, r, c, v)
    st = p.stats()
    assert st["gather_flag"] == 1 and st["direct_coo_columns"] == 1
    assert st["compact_coo_indices"] == int(compact)


def test_nnz_9999_10000_boundary():
    m = n = 3000
    base_r, base_c, base_v = gen.random_coo(m, n, 3, seed=11)
    for total, compact in ((9999, 1), (10000, 0)):
        k = total - base_r.size
        assert k > 0
        # add distinct entries in the last rows to reach the exact count
        rng = np.random.default_rng(total)
        have = set(zip(base_r.tolist(), base_c.tolist()))
        extra = []
        while len(extra) < k:
            rr, cc = int(rng.integers(0, m)), int(rng.integers(0, n))
            if (rr, cc) not in have:
                have.add((rr, cc))
                extra.append((rr, cc))
        r = np.concatenate([base_r, np.array([e[0] for e in extra], np.int32)])
        c = np.concatenate([base_c, np.array([e[1] for e in extra], np.int32)])
        v = np.concatenate([base_v, rng.uniform(-1, 1, k)])
        o = np.lexsort((c, r))
        r, c, v = r[o], c[o], v[o]
        assert r.size == total
        p = check_all(m, n, r, c, v)
        assert p.stats()["compact_coo_indices"] == compact


@pytest.mark.parametrize("n,direct", [(65535, 1), (65536, 0)])
def test_direct_columns_n_boundary(n, direct):
    m = 500
    r, c, v = gen.random_coo(m, n, 4, seed=n)
    p = check_all(m, n, r, c, v)
    st = p.stats()
    assert st["gather_flag"] == 1
    assert st["direct_coo_columns"] == direct


def test_gather_ratio_threshold():
    """Gathering turns on strictly above a 15% share of <= 8-entry blocks."""
    for n_small, want in ((3, 0), (4, 1)):
        # 20 blocks of 20 entries plus n_small blocks of 2 entries: 3/23 =
        # 13% stays off, 4/24 = 16.7% turns on
        specs = [(i // 5, i % 5, 20) for i in range(20)] + [(4, j, 2) for j in range(n_small)]
        r, c, v = blocks_matrix(80, 80, specs, seed=n_small)
        p = check_all(80, 80, r, c, v)
        assert p.stats()["gather_flag"] == want, n_small


def test_long_row_flag():
    """A COO block whose row holds > 8 entries sets the long-row flag on all of
    the block's packed indices (position bits active: direct, non-compact)."""
    m, n = 600, 5000
    rows = [0] * 12 + [1] * 2 + list(range(2, 600))
    cols = list(range(12)) + [100, 200] + [(7 * i) % n for i in range(2, 600)]
    r, c = np.array(rows, np.int32), np.array(cols, np.int32)
    o = np.lexsort((c, r))
    r, c = r[o], c[o]
    # pad nnz past 10000 so indices are not compact (position bits only apply
    # to direct non-compact blocks at bs 16)
    r2, c2, _ = gen.random_coo(m, n, 20, seed=3)
    keep = r2 >= 2
    r = np.concatenate([r, r2[keep]])
    c = np.concatenate([c, c2[keep]])
    key = r.astype(np.int64) * n + c
    _, u = np.unique(key, return_index=True)
    r, c = r[u], c[u]
    v = np.ones(r.size)
    p = check_all(m, n, r, c, v)
    st = p.stats()
    assert st["direct_coo_columns"] == 1 and st["compact_coo_indices"] == 0
    # direct non-compact words: row << 16 | column, position in row at bit
    # 20 (5 bits), long-row flag at bit 25
    k0 = int(p.export("nnz_per_blk")[0])
    assert (p.export("type_per_blk")[0] & 0x7F) == COO
    w = p.export("pl_idx")[:k0].astype(np.int64)
    assert np.all(w & (1 << 25))
    row0 = w[((w >> 16) & 15) == 0]
    assert row0.size == 12
    assert sorted(((row0 >> 20) & 31).tolist()) == list(range(12))
    assert np.array_equal(np.sort(row0 & 0xFFFF), np.arange(12))


@pytest.mark.parametrize("bs", [8, 16, 32, 64])
@pytest.mark.parametrize("row_owned", [True, False])
def test_random_block_sizes(bs, row_owned):
    if not pyoracle.ref_available(bs, 4, row_owned, False) and bs != 16:
        pass  # product vs oracle still checked
    m, n = 1500, 1300
    r, c, v = gen.random_coo(m, n, 9, seed=bs, dup=True, empty_rows=0.1)
    check_all(m, n, r, c, v, bs=bs, row_owned=row_owned)


def test_balance_on_large_irregular():
    """Segment balancing (nnz >= 10000, not row owned) on a power-law matrix
    with long rows."""
    m = n = 20000
    rng = np.random.default_rng(17)
    lens = np.minimum(np.floor(rng.pareto(1.5, m) + 1).astype(int), 3000)
    r = np.repeat(np.arange(m), lens).astype(np.int32)
    c = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    v = rng.uniform(-1, 1, r.size)
    for w in (4, 8):
        p = check_all(m, n, r, c, v, waves=w, row_owned=False)
        assert p.stats()["load_balanced"] == 1


def test_unit_values_uniform_rows():
    """All-ones blocks whose rows repeat the first row are marked uniform."""
    rows, cols = [], []
    for rr in range(32):
        for cc in (1, 5, 9, 17, 20, 30):
            rows.append(rr)
            cols.append(cc)
    r, c = np.array(rows, np.int32), np.array(cols, np.int32)
    v = np.ones(r.size)
    p = check_all(32, 32, r, c, v)
    t = p.export("type_per_blk")
    assert np.all(t & 0x80)


def test_duplicates_are_summed():
    m = n = 40
    r = np.array([0, 0, 0, 3, 3, 17, 17, 17], np.int32)
    c = np.array([5, 5, 2, 7, 7, 30, 30, 30], np.int32)
    v = np.array([1.0, 2.0, 4.0, 0.5, 0.25, 1.0, 1.0, 1.0])
    p = check_all(m, n, r, c, v)
    x = np.arange(n, dtype=np.float64) + 1
    y = pyoracle.spmv_seq(m, r, c, v, x)
    assert y[0] == 3 * 6 + 4 * 3 and y[3] == 0.75 * 8 and y[17] == 3 * 31
    p.close()


def test_rejects_unsorted_and_out_of_range():
    from cbspmv import CbspmvError

    with pytest.raises(CbspmvError):
        Plan(4, 4, np.array([1, 0]), np.array([0, 0]), np.ones(2))
    with pytest.raises(CbspmvError):
        Plan(4, 4, np.array([0, 5]), np.array([0, 0]), np.ones(2))
    with pytest.raises(CbspmvError):
        Plan(4, 4, np.array([0, 1]), np.array([0, -1]), np.ones(2))


def test_empty_matrix():
    p = Plan(10, 10, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    st = p.stats()
    assert st["gb_nnzb"] == 0
    p.close()
