"""CPU checks of the transcoded stream slab: decoding it with numpy gives back
the same matrix as the input COO (sum of duplicates), and its work items,
tiles and LDS slots satisfy the invariants the stream kernel relies on."""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle

CACHED = np.uint32(0x80000000)
LINE = 16
def tile_shape(plan):
    waves, nbytes = plan.export("stream_tile_shape")
    return int(waves), int(nbytes)


def tile_slots(plan, dtype):
    return tile_shape(plan)[1] // (8 if dtype == "f64" else 4)


def decode_stream(plan, dtype="f64"):
    """(row, col, val, pad) of every stream entry, decoded like the kernel
    does: key = cached flag | row offset from the item's base row | column or
    LDS slot; padding entries read the always-zero LDS slot."""
    key = plan.export("stream_key")
    val = plan.export("stream_val")
    cb = int(plan.export("stream_col_bits")[0])
    items = plan.export("stream_items").reshape(-1, 4).astype(np.int64)
    tiles = plan.export("stream_tiles").reshape(-1, 2).astype(np.int64)
    lines = plan.export("stream_tile_lines").astype(np.int64)
    omask = (1 << (31 - cb)) - 1
    off = ((key >> np.uint32(cb)) & np.uint32(omask)).astype(np.int64)
    low = (key & np.uint32((1 << cb) - 1)).astype(np.int64)
    cached = (key & CACHED) != 0
    item_of = np.repeat(np.arange(items.shape[0]), items[:, 1] - items[:, 0])
    assert item_of.size == key.size
    row = items[item_of, 2] + off
    tile_of = item_of // tile_shape(plan)[0]
    pad = cached & (low == tile_slots(plan, dtype) - 1)
    col = low.copy()
    cidx = np.nonzero(cached & ~pad)[0]
    slot = low[cidx]
    assert np.all(slot < tiles[tile_of[cidx], 1] * LINE), "LDS slot past the tile's staged lines"
    col[cidx] = lines[tiles[tile_of[cidx], 0] + slot // LINE] * LINE + slot % LINE
    assert np.all(val[pad] == 0.0)
    return row, col, val, pad, items, tiles, lines


def _matrices():
    yield "poisson_4096", gen.poisson_4096()
    for seed, (m, n, pr, dup, unit) in enumerate([(500, 700, 3, True, False), (3000, 3000, 40, False, False),
                                                  (300, 300, 80, True, True), (1001, 777, 6, False, False)]):
        r, c, v = gen.random_coo(m, n, pr, seed=seed, dup=dup, unit=unit, empty_rows=0.2 if seed == 3 else 0.0)
        yield f"rand{seed}", dict(m=m, n=n, row=r, col=c, val=v)
    # one long row (cut into pieces)
    r, c, v = gen.random_coo(200, 50000, 4, seed=9)
    rng = np.random.default_rng(9)
    lc = np.sort(rng.choice(50000, 30000, replace=False)).astype(np.int32)
    at = np.searchsorted(r, 77)
    r = np.concatenate([r[:at], np.full(lc.size, 77, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc, c[at:]])
    v = np.concatenate([v[:at], rng.uniform(-1, 1, lc.size), v[at:]])
    yield "long_row", dict(m=200, n=50000, row=r, col=c, val=v)


@pytest.mark.parametrize("bs", [4, 16, 32])
@pytest.mark.parametrize("item_nnz", [0, 256])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_stream_decodes_to_input_matrix(bs, item_nnz, dtype):
    epl = 4 if dtype == "f64" else 8
    for name, d in _matrices():
        plan = Plan(d["m"], d["n"], d["row"], d["col"], d["val"], block_size=bs, item_nnz=item_nnz, dtype=dtype)
        row, col, val, pad, items, tiles, lines = decode_stream(plan, dtype)
        keep = ~pad
        row, col, val = row[keep], col[keep], val[keep]
        assert np.all(row < d["m"]) and np.all(col < d["n"]), name
        # same matrix: compare sums per (row, col)
        key_in = d["row"].astype(np.int64) * d["n"] + d["col"]
        key_st = row * d["n"] + col
        u_in, inv_in = np.unique(key_in, return_inverse=True)
        s_in = np.zeros(u_in.size)
        np.add.at(s_in, inv_in, d["val"])
        u_st, inv_st = np.unique(key_st, return_inverse=True)
        s_st = np.zeros(u_st.size)
        np.add.at(s_st, inv_st, val)
        nz_in, nz_st = u_in[s_in != 0], u_st[s_st != 0]
        assert np.array_equal(nz_in, nz_st), name
        np.testing.assert_allclose(s_st[s_st != 0], s_in[s_in != 0], rtol=1e-14, atol=1e-15)
        # rows are contiguous runs: along the stream, rows never decrease
        assert np.all(np.diff(row) >= 0), name
        # items tile the stream in lane-width multiples; cuts fall between
        # rows unless the item says it continues a row
        assert items[0, 0] == 0 and np.all(items[1:, 0] == items[:-1, 1])
        assert np.all(items[:, 0] % epl == 0) and np.all(items[:, 1] % epl == 0)
        first_partial = items[:, 3] & 0xFF
        last_partial = (items[:, 3] >> 8) & 0xFF
        row_all, *_ = decode_stream(plan, dtype)
        for i in range(1, items.shape[0]):
            same_row = row_all[items[i, 0]] == row_all[items[i - 1, 1] - 1]
            assert bool(same_row) == bool(first_partial[i]) == bool(last_partial[i - 1]), (name, i)
        # staged lines: sorted, distinct, at most 511 per tile
        assert np.all(tiles[:, 1] <= tile_slots(plan, dtype) // LINE - 1)
        for t in range(tiles.shape[0]):
            ls = lines[tiles[t, 0]: tiles[t, 0] + tiles[t, 1]]
            assert np.all(np.diff(ls) > 0)
        plan.close()


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_slab_planes_hold_the_entries(dtype):
    """The device slab stores each chunk of 64 lanes x EPL entries as 16-byte
    lane planes (keys, then values); reading it back plane by plane gives the
    logical entry stream."""
    d = gen.poisson_4096()
    plan = Plan(d["m"], d["n"], d["row"], d["col"], d["val"], dtype=dtype, item_nnz=1000)
    key = plan.export("stream_key")
    val = plan.export("stream_val")
    slab = plan.export("stream_slab")
    items = plan.export("stream_items").reshape(-1, 4).astype(np.int64)
    epl = 4 if dtype == "f64" else 8
    vb = {"f64": 8, "f32": 4, "bf16": 2}[dtype]
    nk = epl * 4 // 16
    assert slab.size == key.size * (4 + vb)
    k_back = np.zeros(key.size, np.uint32)
    v_back = np.zeros(key.size, np.float64)
    for b, e, _, _ in items:
        for e0 in range(b, e, 64 * epl):
            G = min(64, (e - e0) // epl)
            base = e0 * (4 + vb)
            planes = slab[base: base + G * 16 * (nk + epl * vb // 16)].reshape(-1, G, 16)
            kk = planes[:nk].transpose(1, 0, 2).reshape(G, -1).view(np.uint32)
            k_back[e0: e0 + G * epl] = kk.reshape(-1)
            vraw = planes[nk:].transpose(1, 0, 2).reshape(G, -1)
            if vb == 8:
                vv = vraw.view(np.float64)
            elif vb == 4:
                vv = vraw.view(np.float32).astype(np.float64)
            else:
                vv = (vraw.view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            v_back[e0: e0 + G * epl] = vv.reshape(-1)
    assert np.array_equal(k_back, key)
    if vb == 8:
        assert np.array_equal(v_back, val)
    elif vb == 4:
        assert np.array_equal(v_back, val.astype(np.float32).astype(np.float64))
    else:
        assert np.allclose(v_back, val, rtol=2 ** -8, atol=0)
    plan.close()


def test_stream_spmv_emulated_matches_oracle():
    """y from the decoded stream (numpy) equals the oracle to 1e-12."""
    for name, d in _matrices():
        plan = Plan(d["m"], d["n"], d["row"], d["col"], d["val"])
        row, col, val, pad, *_ = decode_stream(plan)
        x = np.random.default_rng(5).uniform(-1, 1, d["n"])
        xx = np.where(pad, 0.0, x[np.where(pad, 0, col)])
        y = np.zeros(d["m"])
        np.add.at(y, row, val * xx)
        y_ref = pyoracle.spmv_seq(d["m"], d["row"], d["col"], d["val"], x)
        assert pyoracle.rel_l2(y, y_ref) <= 1e-12, name
        empty = plan.export("stream_empty_rows")
        has = np.zeros(d["m"], bool)
        has[row[~pad]] = True
        assert np.array_equal(np.nonzero(~has)[0], np.sort(empty)), name
        plan.close()


def test_default_tile_waves():
    """Waves per tile left open: 16 without column panels, 8 for panel plans
    and for the transposed plan of a backward pass; an explicit count is
    kept everywhere."""
    m, n = 500, 700
    r, c, v = gen.random_coo(m, n, 6, seed=21)
    for kw, want, want_t in ((dict(dtype="f64"), 16, 8), (dict(dtype="f32"), 16, 8), (dict(dtype="bf16"), 16, 8),
                             (dict(dtype="f32", compensated=True), 16, 8), (dict(dtype="f32", panels=2), 8, 8),
                             (dict(dtype="f32", tile_waves=5), 5, 5), (dict(dtype="f64", tile_waves=12), 12, 12)):
        plan = Plan(m, n, r, c, v, **kw)
        assert tile_shape(plan)[0] == want, kw
        plan.build_transpose()
        assert int(plan.export("transpose.stream_tile_shape")[0]) == want_t, kw
        plan.close()


def test_items_fill_one_round_of_tiles():
    """Items are cut as small as one round of tiles (one per CU, 256) allows:
    never more items than 256 tiles hold, and for a matrix of many short
    rows hardly fewer, so few CUs are left without a tile.  (The matrix is
    large enough that items stay above their 1024-entry minimum.)"""
    m = n = 400_000
    r, c, v = gen.random_coo(m, n, 12, seed=31)
    for dtype in ("f64", "f32"):
        plan = Plan(m, n, r, c, v, dtype=dtype)
        waves = tile_shape(plan)[0]
        items = plan.export("stream_items").reshape(-1, 4).shape[0]
        assert items <= 256 * waves, dtype
        assert items >= 0.99 * 256 * waves, (dtype, items)
        plan.close()
