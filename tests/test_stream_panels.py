"""CPU checks of the column-panel layout of the stream schedule: entries a
tile cannot serve from LDS move to per-panel item streams; decoding tiles
plus panel items gives back the input matrix, every panel item's columns lie
in its panel and no item cuts a row of its panel, the per-row panel masks
match the items, and every row with panel entries keeps an entry in the
tiles (so the tile stores its y before the panel sums are added)."""
import numpy as np
import pytest

from cbspmv import Plan, gen
from test_stream_layout import decode_stream


def _mixed(m, n, per_row_far, near, seed):
    """Banded rows (entries within +-near of the diagonal) plus per_row_far
    uniformly scattered columns per row, rows in order."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    k_near = rng.integers(1, 12, size=m)
    k_far = rng.poisson(per_row_far, size=m)
    r_near = np.repeat(np.arange(m), k_near)
    c_near = np.clip(r_near + rng.integers(-near, near + 1, size=r_near.size), 0, n - 1)
    r_far = np.repeat(np.arange(m), k_far)
    c_far = rng.integers(0, n, size=r_far.size)
    r = np.concatenate([r_near, r_far])
    c = np.concatenate([c_near, c_far])
    o = np.lexsort((rng.random(r.size), r))
    r, c = r[o].astype(np.int32), c[o].astype(np.int32)
    return m, n, r, c, rng.uniform(-1, 1, r.size)


def _check_panels(plan, m, n, r, c, v, dtype, expect_panels):
    P, n_tile_items = (int(t) for t in plan.export("stream_panels"))
    assert P == expect_panels
    row, col, val, pad, items, tiles, lines = decode_stream(plan, dtype)
    keep = ~pad
    # same matrix (sums per (row, col))
    key_in = r.astype(np.int64) * n + c
    key_st = row[keep] * n + col[keep]
    u_in, inv_in = np.unique(key_in, return_inverse=True)
    u_st, inv_st = np.unique(key_st, return_inverse=True)
    assert np.array_equal(u_in, u_st)
    tol = 1e-12 if dtype == "f64" else 1e-5
    np.testing.assert_allclose(np.bincount(inv_st, val[keep]), np.bincount(inv_in, v), rtol=tol, atol=tol)
    if P == 1:
        assert n_tile_items == items.shape[0]
        return
    ptr = plan.export("stream_panel_items").astype(np.int64)
    mask = plan.export("stream_panel_mask")
    assert ptr.size == P + 1 and ptr[0] == n_tile_items and ptr[-1] == items.shape[0]
    assert np.all(np.diff(ptr) >= 0)
    width = -(-(-(-n // P)) // 16) * 16
    item_of = np.repeat(np.arange(items.shape[0]), items[:, 1] - items[:, 0])
    in_panel = item_of >= n_tile_items
    # panel items hold no staged keys (only columns and padding)
    cached = (plan.export("stream_key") & np.uint32(0x80000000)) != 0
    assert np.all(pad[in_panel] == cached[in_panel])
    panel_of_item = np.searchsorted(ptr, np.arange(items.shape[0]), side="right") - 1
    sel = in_panel & keep
    prow, pcol = row[sel], col[sel]
    assert np.all(pcol // width == panel_of_item[item_of[sel]])
    # panel items never cut a row: each (row, panel) lies in one item
    pair = prow * P + pcol // width
    first_item = {}
    for pr, it in zip(pair.tolist(), item_of[sel].tolist()):
        assert first_item.setdefault(pr, it) == it
    # row masks
    want = np.zeros(m, np.uint8)
    np.bitwise_or.at(want, prow, (1 << (pcol // width)).astype(np.uint8))
    assert np.array_equal(mask, want)
    # rows with panel entries keep a tile entry (a zero entry if need be)
    tile_rows = np.zeros(m, bool)
    tile_rows[row[~in_panel]] = True
    assert np.all(tile_rows[mask != 0])
    # item row offsets fit the key field
    cb = int(plan.export("stream_col_bits")[0])
    span = np.zeros(items.shape[0], np.int64)
    np.maximum.at(span, item_of, row - items[item_of, 2])
    assert np.all(span <= (1 << (31 - cb)) - 1)


@pytest.mark.parametrize("panels", [2, 4, 8])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_forced_panels_layout(panels, dtype):
    m, n, r, c, v = _mixed(60000, 60000, 3, 300, seed=panels)
    plan = Plan(m, n, r, c, v, dtype=dtype, panels=panels)
    _check_panels(plan, m, n, r, c, v, dtype, panels)


def test_auto_panels():
    # panels only on request for now; the same matrix with 8 MiB of x
    m, n, r, c, v = _mixed(40000, 1 << 20, 4, 200, seed=5)
    _check_panels(Plan(m, n, r, c, v), m, n, r, c, v, "f64", 1)
    _check_panels(Plan(m, n, r, c, v, panels=4), m, n, r, c, v, "f64", 4)
    _check_panels(Plan(m, n, r, c, v, dtype="f32", panels=2), m, n, r, c, v, "f32", 2)
    # compensated plans keep x whole
    p = Plan(m, n, r, c, v, dtype="f32", compensated=True, panels=0)
    assert int(p.export("stream_panels")[0]) == 1


def test_panels_keep_cut_rows_in_tiles():
    # a row long enough to be cut into pieces keeps its scattered entries in the tiles
    m, n, r, c, v = _mixed(3000, 400000, 2, 50, seed=7)
    rng = np.random.default_rng(7)
    keep = r != 1500
    lc = np.sort(rng.choice(n, 60000, replace=False)).astype(np.int32)
    at = np.searchsorted(r[keep], 1500)
    r = np.concatenate([r[keep][:at], np.full(lc.size, 1500, np.int32), r[keep][at:]])
    c = np.concatenate([c[keep][:at], lc, c[keep][at:]])
    v = rng.uniform(-1, 1, r.size)
    plan = Plan(m, n, r, c, v, item_nnz=2048, panels=4)
    _check_panels(plan, m, n, r, c, v, "f64", 4)
    row, col, val, pad, items, tiles, lines = decode_stream(plan)
    n_tile_items = int(plan.export("stream_panels")[1])
    item_of = np.repeat(np.arange(items.shape[0]), items[:, 1] - items[:, 0])
    assert not np.any((row == 1500) & (item_of >= n_tile_items))


def test_bad_panel_options_rejected():
    m, n, r, c, v = _mixed(100, 100, 1, 5, seed=1)
    with pytest.raises(Exception, match="stream_panels"):
        Plan(m, n, r, c, v, panels=3)
    with pytest.raises(Exception, match="stream_panels"):
        Plan(m, n, r, c, v, panels=2, tile_waves=12)


@pytest.mark.parametrize("panels", [2, 8])
def test_small_cache_panels_layout(panels):
    # a 4 KB LDS cache leaves most entries to the panels
    r, c, v = gen.random_coo(3000, 20000, 7, seed=panels)
    plan = Plan(3000, 20000, r, c, v, panels=panels, tile_kib=4)
    _check_panels(plan, 3000, 20000, r, c, v, "f64", panels)
    n_tile_items = int(plan.export("stream_panels")[1])
    assert plan.export("stream_items").reshape(-1, 4).shape[0] > n_tile_items
