"""Scattered-entry split, host side (no GPU): which entries the split takes,
and the layout the sweep kernel reads.

The rule is restated here in numpy, independently of split.cpp: the
row-major entries are cut into groups of about nnz / tiles entries at row
boundaries, and an entry is scattered when at most three entries of its group
read its x line (128 bytes of x).  The layout checks are the kernel's
preconditions: every exported entry decodes to its input entry, each input
entry appears once, a row belongs to one wave, each 64-entry chunk of a wave
is in row order and the chunks of a wave in column order, rows fit the
row tile and keys their bit fields."""
import numpy as np
import pytest

from cbspmv import Plan


def scattered_matrix(m, n, seed, far_per_row=(0, 4), near=200, long_row=None):
    """Rows of near-diagonal entries plus a few uniformly scattered ones;
    optionally one row with `long_row` scattered entries."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(m):
        k = int(rng.integers(1, 12))
        c = np.clip(r + rng.integers(-near, near, size=k), 0, n - 1)
        f = rng.integers(0, n, size=int(rng.integers(*far_per_row)))
        if long_row is not None and r == m // 3:
            f = rng.choice(n, size=long_row, replace=False)
        c = np.unique(np.concatenate([c, f]))
        rows.append(np.full(c.size, r))
        cols.append(c)
    row = np.concatenate(rows).astype(np.int32)
    col = np.concatenate(cols).astype(np.int32)
    return m, n, row, col, rng.uniform(-1, 1, row.size)


def scattered_rule(row, col, tiles, line):
    nnz = row.size
    target = max(1, nnz // tiles)
    cut, e = [0], target
    while e < nnz:
        while e < nnz and row[e] == row[e - 1]:
            e += 1
        if e >= nnz:
            break
        cut.append(e)
        e += target
    cut.append(nnz)
    far = np.zeros(nnz, dtype=bool)
    for a, b in zip(cut[:-1], cut[1:]):
        _, inv, cnt = np.unique(col[a:b] // line, return_inverse=True, return_counts=True)
        far[a:b] = cnt[inv] <= 3
    return far


def decode(p):
    key, src, wo, br, bs, sh = (p.export(k) for k in ("split_key", "split_src", "split_wave_off", "split_blk_row",
                                                      "split_blk_span", "split_shape"))
    cb, W, R = (int(v) for v in sh)
    cnt = np.diff(wo.astype(np.int64))
    blk = np.repeat(np.arange(br.size), cnt.reshape(-1, W).sum(1))
    wave = np.repeat(np.arange(cnt.size), cnt)
    rloc = (key >> cb).astype(np.int64)
    return dict(key=key, src=src.astype(np.int64), wo=wo.astype(np.int64), br=br, bs=bs, cb=cb, W=W, R=R,
                blk=blk, wave=wave, rloc=rloc, row=br[blk] + rloc, col=(key & ((1 << cb) - 1)).astype(np.int64),
                val=p.export("split_val"))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_split_takes_the_scattered_entries(dtype):
    m, n, row, col, val = scattered_matrix(60000, 60000, seed=1)
    p = Plan(m, n, row, col, val, dtype=dtype, split="on", tiles=256)
    d = decode(p)
    want = np.flatnonzero(scattered_rule(row, col, 256, 16 if dtype == "f64" else 32))
    assert want.size > 0
    assert np.array_equal(np.sort(d["src"]), want)
    assert p.stats()["split_entries"] == want.size
    assert np.array_equal(d["row"], row[d["src"]]) and np.array_equal(d["col"], col[d["src"]])
    assert np.array_equal(d["val"], val[d["src"]])
    p.close()


@pytest.mark.parametrize("seed,n,long_row", [(2, 70000, None), (3, 1500000, 9000)])
def test_split_layout_is_what_the_kernel_reads(seed, n, long_row):
    # the long row's 9000 scattered entries over 1.5M columns mostly read
    # lines of their own: its block holds it alone
    m, n, row, col, val = scattered_matrix(40000, n, seed=seed, far_per_row=(0, 6), long_row=long_row)
    p = Plan(m, n, row, col, val, split="on", tiles=64)
    d = decode(p)
    cb, W, R = d["cb"], d["W"], d["R"]
    assert (1 << cb) >= n and R <= 2048 and (R << cb) <= 1 << 32
    nb = d["br"].size
    assert p.stats()["split_blocks"] == nb and d["wo"].size == nb * W + 1 and d["wo"][0] == 0
    assert np.all(np.diff(d["wo"]) >= 0) and d["wo"][-1] == d["key"].size
    assert np.all(d["bs"] >= 1) and np.all(d["bs"] <= R) and np.all(d["rloc"] < d["bs"][d["blk"]])
    assert np.all(np.diff(d["br"]) > 0) and np.all(d["br"][:-1] + d["bs"][:-1] <= d["br"][1:])
    per_block = np.bincount(d["blk"], minlength=nb)
    single_row = np.array([np.unique(d["row"][d["blk"] == b]).size == 1 for b in range(nb)])
    assert np.all((per_block <= 8000) | single_row)
    if long_row is not None:
        assert single_row.any()
    # a row belongs to one wave; waves of a block hold ascending row ranges
    pairs = np.unique(np.stack([d["row"], d["wave"]]), axis=1)
    assert np.unique(pairs[0]).size == pairs.shape[1]
    for w in range(d["wo"].size - 1):
        a, b = d["wo"][w], d["wo"][w + 1]
        r, c = d["row"][a:b], d["col"][a:b]
        chunks = [(q, min(b - a, q + 64)) for q in range(0, b - a, 64)]
        for q, z in chunks:
            assert np.all(np.diff(r[q:z]) >= 0)                 # rows ascend inside a chunk
        for (q, z), (q2, z2) in zip(chunks, chunks[1:]):
            assert c[q:z].max() <= c[q2:z2].min()               # chunks in column order
    assert np.unique(d["src"]).size == d["src"].size
    p.close()


def test_split_policy():
    # x of 60000 FP64 entries (480 KB) is below the automatic threshold
    m, n, row, col, val = scattered_matrix(60000, 60000, seed=4)
    for split, expect in (("auto", False), ("off", False), ("on", True)):
        p = Plan(m, n, row, col, val, split=split)
        assert (p.stats()["split_entries"] > 0) == expect, split
        p.close()
    # BF16 plans split (their sweep accumulates in FP32), compensated and
    # column-panel plans do not
    p = Plan(m, n, row, col, val, split="on", dtype="bf16")
    assert p.stats()["split_entries"] > 0
    p.close()
    for opts in (dict(dtype="f32", compensated=True), dict(dtype="bf16", compensated=True), dict(panels=2, tile_waves=8)):
        p = Plan(m, n, row, col, val, split="on", **opts)
        assert p.stats()["split_entries"] == 0, opts
        p.close()
    with pytest.raises(KeyError):
        Plan(m, n, row, col, val, split="sometimes")
    # the C ABI rejects an option value outside the enum
    import ctypes
    from cbspmv import _lib
    from cbspmv.plan import make_opts
    o = make_opts()
    o.stream_split = 3
    h = ctypes.c_void_p()
    rc = _lib.lib().cbspmv_build(m, n, row.size, row.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 col.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(o), ctypes.byref(h))
    assert rc == _lib.ERR_ARG and not h.value


def test_split_is_not_saved(tmp_path):
    m, n, row, col, val = scattered_matrix(30000, 30000, seed=5)
    p = Plan(m, n, row, col, val, split="on")
    assert p.stats()["split_entries"] > 0
    path = tmp_path / "plan.cbp"
    p.save(path)
    q = Plan.load(path)
    assert q.stats()["split_entries"] == 0 and q.stats()["nnz"] == row.size
    assert np.array_equal(q.export("stream_key"), p.export("stream_key"))
    p.close()
    q.close()
