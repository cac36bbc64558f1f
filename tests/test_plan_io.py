"""Plan persistence (cbspmv_save / cbspmv_load): a plan read back from its
file exports the same arrays, images and statistics as the plan that wrote
it, runs the same SpMV on the GPU, and damaged files are rejected."""
import os

import numpy as np
import pytest

from cbspmv import CbspmvError, Plan, gen
from oracle import pyoracle

NAMES = pyoracle.ARRAY_NAMES + ["image", "stream_slab", "stream_key", "stream_val", "stream_items", "stream_tiles",
                                "stream_tile_lines", "stream_empty_rows", "policy", "stream_col_bits", "stream_panels",
                                "stream_panel_items", "stream_panel_mask"]


def _plans():
    d = gen.poisson_4096()
    yield "poisson_4096", (d["m"], d["n"], d["row"], d["col"], d["val"]), dict()
    r, c, v = gen.random_coo(900, 700, 15, seed=3, dup=True)
    yield "dup_lpt_bs32", (900, 700, r, c, v), dict(block_size=32, row_owned=False)
    yield "f32", (900, 700, r, c, v), dict(dtype="f32", compensated=True)
    yield "bf16", (900, 700, r, c, v), dict(dtype="bf16")
    r2, c2, v2 = gen.random_coo(3000, 20000, 7, seed=5)
    yield "panels", (3000, 20000, r2, c2, v2), dict(panels=4, tile_kib=4)


@pytest.mark.parametrize("name,mat,opts", list(_plans()), ids=lambda v: v if isinstance(v, str) else "")
def test_round_trip(tmp_path, name, mat, opts):
    p = Plan(*mat, **opts)
    path = tmp_path / f"{name}.cbplan"
    p.save(path)
    q = Plan.load(path)
    assert q.dtype == p.dtype and (q.m, q.n, q.nnz) == (p.m, p.n, p.nnz)
    sp, sq = p.stats(), q.stats()
    assert sp == sq
    for a in NAMES:
        assert np.array_equal(p.export(a), q.export(a)), a
    for reg in ("gb_row_idx", "block_addr", "x", "y", "total"):
        assert p.layout_offset(reg) == q.layout_offset(reg)
    q.close()
    p.close()


def test_rejects_damaged_files(tmp_path):
    d = gen.poisson_4096()
    p = Plan(d["m"], d["n"], d["row"], d["col"], d["val"])
    path = tmp_path / "a.cbplan"
    p.save(path)
    raw = path.read_bytes()
    (tmp_path / "trunc.cbplan").write_bytes(raw[: len(raw) // 2])
    (tmp_path / "magic.cbplan").write_bytes(b"X" + raw[1:])
    (tmp_path / "tail.cbplan").write_bytes(raw + b"\0")
    for bad in ("trunc", "magic", "tail", "missing"):
        with pytest.raises(CbspmvError):
            Plan.load(tmp_path / f"{bad}.cbplan")
    p.close()


@pytest.mark.gpu
def test_loaded_plan_spmv(tmp_path, gpu):
    import torch

    d = gen.poisson_4096()
    p = Plan(d["m"], d["n"], d["row"], d["col"], d["val"])
    path = tmp_path / "c1.cbplan"
    p.save(path)
    q = Plan.load(path)
    x = torch.tensor(d["x"], device=gpu)
    y_ref = pyoracle.spmv_seq(d["m"], d["row"], d["col"], d["val"], d["x"])
    for sched in ("stream", "blocked"):
        q.set_schedule(sched)
        assert pyoracle.rel_l2(q.spmv(x).cpu().numpy(), y_ref) <= 1e-12
    p.set_schedule("stream")
    q.set_schedule("stream")
    assert torch.equal(p.spmv(x), q.spmv(x))   # same slab, same bits
    q.close()
    p.close()


def _corrupt(src, dst, find, patch_at, patch):
    """Copies plan file src to dst with `patch` written over the bytes that
    start `patch_at` bytes into the first occurrence of `find`."""
    raw = bytearray(open(src, "rb").read())
    at = bytes(raw).find(find)
    assert at >= 0, "array not found in the file"
    raw[at + patch_at:at + patch_at + len(patch)] = patch
    open(dst, "wb").write(bytes(raw))


@pytest.mark.parametrize("what", ["item_end", "item_row", "key_column", "key_row", "tile_lines", "split_piece"])
def test_damaged_index_arrays_are_refused(tmp_path, what):
    """A file whose lengths are consistent but whose indices point outside
    the arrays the kernels index is refused on load (CBSPMV_ERR_ARG), not
    uploaded: an item past the slab or past m, a key naming a column past
    n or a row past m, a tile with more lines than its table or its LDS
    cache, a cut row's pieces past the partial buffer."""
    m, n = 500, 800
    r, c, v = gen.random_coo(m, n, 9, seed=8)
    # row 250 takes every column: longer than four items, so it is cut into pieces
    keep = r != 250
    at = int(np.searchsorted(r, 250))
    r = np.concatenate([r[:at], np.full(n, 250, np.int32), r[keep][at:]])
    c = np.concatenate([c[:at], np.arange(n, dtype=np.int32), c[keep][at:]])
    v = np.ones(r.size)
    p = Plan(m, n, r, c, v, item_nnz=64, tile_waves=4)
    good = tmp_path / "good.cbplan"
    p.save(good)
    Plan.load(good).close()
    items = p.export("stream_items").astype(np.uint32)
    slab = p.export("stream_slab")
    bad = tmp_path / "bad.cbplan"
    cb = int(p.export("stream_col_bits")[0])
    if what == "item_end":
        _corrupt(good, bad, items.tobytes(), 4, np.uint32(items[1] + (1 << 20)).tobytes())
    elif what == "item_row":
        _corrupt(good, bad, items.tobytes(), 8, np.uint32(m + 7).tobytes())
    elif what in ("key_column", "key_row"):
        key = np.frombuffer(slab[:4].tobytes(), np.uint32)[0]
        if what == "key_column":
            new = np.uint32(n + 3)                                   # uncached, column past n
        else:
            new = np.uint32((key & ((1 << cb) - 1)) | ((((1 << (31 - cb)) - 1) << cb) & 0x7FFFFFFF))
        _corrupt(good, bad, slab[:4096].tobytes(), 0, new.tobytes())
    elif what == "tile_lines":
        tiles = p.export("stream_tiles").astype(np.uint32)
        _corrupt(good, bad, tiles.tobytes(), 4, np.uint32(1 << 20).tobytes())
    else:
        assert p.stats()["nnz"] and items.size
        splits_items = items.reshape(-1, 4)[:, 3]
        assert np.any(splits_items != 0), "fixture must cut a row into pieces"
        # splits are (row, first_item, pieces, partial_base) u32 records: find the row-250 record
        raw = open(good, "rb").read()
        want = np.uint32(250).tobytes()
        # the split record starts with the row; pieces two words later
        pos = [i for i in range(0, len(raw) - 16, 4) if raw[i:i + 4] == want
               and np.frombuffer(raw[i + 8:i + 12], np.uint32)[0] >= 2
               and np.frombuffer(raw[i + 8:i + 12], np.uint32)[0] < 64]
        assert pos
        b = bytearray(raw)
        b[pos[0] + 12:pos[0] + 16] = np.uint32(1 << 24).tobytes()   # partial_base far past the buffer
        open(bad, "wb").write(bytes(b))
    with pytest.raises(CbspmvError, match="invalid plan file"):
        Plan.load(bad)
    p.close()
