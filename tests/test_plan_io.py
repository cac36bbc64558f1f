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
    d = gen.config1()
    yield "config1", (d["m"], d["n"], d["row"], d["col"], d["val"]), dict()
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
    d = gen.config1()
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

    d = gen.config1()
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
