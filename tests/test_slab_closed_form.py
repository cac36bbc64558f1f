"""Closed-form pin of the canonical slab (cb_spmv's device layout,
cb-spmv-dcu/src/cb-spmv-dcu.hpp:1100-1348).

`closed_form_image` below recomputes the packed image with numpy alone, from
the builder arrays and the layout rules the reference states:

  * reserve_bytes (cb-spmv-dcu.hpp:1133-1139): each region starts at the
    running size rounded up to its own alignment (align_up, :16-19);
  * region order (:1141-1180): gb_row_idx, gb_col_idx, [row-owned:
    segment_rows/begin/end (int), segment_direct/uniform (u8)], nnz_per_blk
    (int), type_per_blk (u8), [pattern reuse: pattern_row_ptr/col_idx/
    output_ptr/output_rows], [gathered: block_cols, block_cols_offset],
    block addresses (8 bytes each);
  * per-block payloads in final block order (:1182-1226): COO idx (u32, or
    u16 when compact) then, unless values are unit, values at
    align_up(index bytes, 8); CSR row_ptr[bs+1] and col[k] as IndexTypeOrig
    (u8 up to bs 16, u16 above) then values at align_up(., 8); DENSE
    bs*bs values, or one row mask per row (u8 up to bs 8, u16/u32/u64 at
    16/32/64) for unit values; payload alignment 8 for valued payloads, the
    index (mask) type's size for unit ones (4 for COO, compact included);
  * then x (gb_n*bs doubles), y (gb_m*bs doubles), row-owned output epochs.
  * unit_values (:1105-1127): every COO/CSR value is 1 and every DENSE
    value is 0 or 1.

The reference packer cannot run here (it allocates device memory before it
computes addresses), so this is the independent check of the layout: it
shares no code with the product packer (packer.cpp) or the oracle's
(cb_oracle.cpp).  Where oracle/_ref is built the builder arrays come from
the reference builder itself; everywhere the product's arrays, which
test_golden.py pins to the reference's digests, feed it too.  The product
stores byte offsets from the slab base where the reference stores device
pointers (0 = no payload) and leaves the x region zero in its exported image.
"""
import os

import numpy as np
import pytest

from cbspmv import Plan, gen, read_mtx
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
COO, CSR, DENSE = 0, 1, 3


def _align(v, a):
    return (v + a - 1) // a * a


def _orig_bytes(bs):
    return 1 if bs <= 16 else 2


def _mask_bytes(bs):
    return {2: 1, 4: 1, 8: 1, 16: 2, 32: 4, 64: 8}[bs]


def closed_form_image(a, bs, row_owned, pattern_wave, gather, compact, mutate=None):
    """(offsets, image bytes up to the y region, unit_values) from builder
    arrays `a` (name -> numpy array) by the reference's layout rules.
    `mutate` plants one packer mistake (tests of this check only):
    "no_pad" puts values right after the indices, "swap_rowcol" swaps
    the block row and column regions, "seg_late" puts the segment arrays
    after type_per_blk, "drop_epochs" omits the epoch region."""
    pad = (lambda v, al: v) if mutate == "no_pad" else _align
    val_al = 8
    nnzpb = a["nnz_per_blk"].astype(np.int64)
    types = a["type_per_blk"].astype(np.int64) & 0x7F
    nb = nnzpb.size
    pl_idx, pl_val = a["pl_idx"], a["pl_val"]

    # walk the payload streams once to find each block's slice
    blocks = []
    pi = pv = 0
    for g in range(nb):
        k = int(nnzpb[g])
        if k <= 0:
            blocks.append(None)
            continue
        t = int(types[g])
        if t == COO:
            blocks.append((t, k, pl_idx[pi:pi + k], pl_val[pv:pv + k]))
            pi += k
            pv += k
        elif t == CSR:
            blocks.append((t, k, pl_idx[pi:pi + bs + 1 + k], pl_val[pv:pv + k]))
            pi += bs + 1 + k
            pv += k
        elif t == DENSE:
            blocks.append((t, k, None, pl_val[pv:pv + bs * bs]))
            pv += bs * bs
        else:
            raise AssertionError(f"block {g}: type {t}")
    assert pi == pl_idx.size and pv == pl_val.size, "payload streams do not match the block list"

    unit = True
    for b in blocks:
        if b is None:
            continue
        t, _, _, vals = b
        unit = unit and (bool(np.all((vals == 0.0) | (vals == 1.0))) if t == DENSE else bool(np.all(vals == 1.0)))

    off = {}
    cur = 0

    def take(name, nbytes, al):
        nonlocal cur
        cur = _align(cur, al)
        off[name] = cur
        cur += nbytes
        return off[name]

    S = a["segment_rows"].size
    def segments():
        take("segment_rows", 4 * S, 4)
        take("segment_begin", 4 * S, 4)
        take("segment_end", 4 * S, 4)
        take("segment_direct", S, 1)
        take("segment_uniform", S, 1)

    take("gb_col_idx" if mutate == "swap_rowcol" else "gb_row_idx", 4 * nb, 4)
    take("gb_row_idx" if mutate == "swap_rowcol" else "gb_col_idx", 4 * nb, 4)
    if row_owned and mutate != "seg_late":
        segments()
    take("nnz_per_blk", 4 * nb, 4)
    take("type_per_blk", nb, 1)
    if row_owned and mutate == "seg_late":
        segments()
    if pattern_wave:
        for name in ("pattern_row_ptr", "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows"):
            take(name, 4 * a[name].size, 4)
    if gather:
        take("block_cols", 4 * a["block_cols"].size, 4)
        take("block_cols_offset", 4 * a["block_cols_offset"].size, 4)
    take("block_addr", 8 * nb, 8)

    ob, mb = _orig_bytes(bs), _mask_bytes(bs)
    addr = np.zeros(nb, dtype=np.uint64)
    placed = []
    for g, b in enumerate(blocks):
        if b is None:
            continue
        t, k, idx, vals = b
        if t == COO:
            ib = k * (2 if compact else 4)
            size, al = (ib, 4) if unit else (pad(ib, 8) + 8 * k, val_al)
        elif t == CSR:
            ib = (bs + 1 + k) * ob
            size, al = (ib, ob) if unit else (pad(ib, 8) + 8 * k, val_al)
        else:
            size, al = (bs * mb, mb) if unit else (8 * bs * bs, val_al)
        at = _align(cur, al)
        cur = at + size
        addr[g] = at
        placed.append((at, b))
    take("x", 8 * int(a["gb_n"]) * bs, 8)
    take("y", 8 * int(a["gb_m"]) * bs, 8)
    if row_owned and mutate != "drop_epochs":
        take("output_epochs", 4 * int(a["gb_m"]), 4)
    off["total"] = cur

    img = np.zeros(off["y"], dtype=np.uint8)

    def put(at, arr):
        raw = np.ascontiguousarray(arr).view(np.uint8)
        img[at:at + raw.size] = raw

    put(off["gb_row_idx"], a["gb_row_idx"].astype(np.int32))
    put(off["gb_col_idx"], a["gb_col_idx"].astype(np.int32))
    if row_owned:
        for name in ("segment_rows", "segment_begin", "segment_end"):
            put(off[name], a[name].astype(np.int32))
        put(off["segment_direct"], a["segment_direct"].astype(np.uint8))
        put(off["segment_uniform"], a["segment_uniform"].astype(np.uint8))
    put(off["nnz_per_blk"], nnzpb.astype(np.int32))
    put(off["type_per_blk"], a["type_per_blk"].astype(np.uint8))
    if pattern_wave:
        for name in ("pattern_row_ptr", "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows"):
            put(off[name], a[name].astype(np.int32))
    if gather:
        put(off["block_cols"], a["block_cols"].astype(np.int32))
        put(off["block_cols_offset"], a["block_cols_offset"].astype(np.int32))
    put(off["block_addr"], addr)
    orig_t = np.uint8 if ob == 1 else np.uint16
    mask_t = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[mb]
    for at, (t, k, idx, vals) in placed:
        if t == COO:
            w = idx.astype(np.uint32)
            ib = k * (2 if compact else 4)
            put(at, w.astype(np.uint16) if compact else w)
            if not unit:
                put(at + pad(ib, 8), vals.astype(np.float64))
        elif t == CSR:
            put(at, idx.astype(np.int64).astype(orig_t))
            if not unit:
                put(at + pad((bs + 1 + k) * ob, 8), vals.astype(np.float64))
        else:
            if unit:
                bits = (vals.reshape(bs, bs) != 0).astype(np.uint64)
                masks = (bits << np.arange(bs, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
                put(at, masks.astype(mask_t))
            else:
                put(at, vals.astype(np.float64))
    return off, img, unit


# ---- cases ------------------------------------------------------------------

def _heart2():
    z = np.load(os.path.join(HERE, "golden", "heart2_pattern.npz"))
    ptr = z["row_ptr"].astype(np.int64)
    m, n = int(z["m"]), int(z["n"])
    r = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
    return m, n, r, z["col"].astype(np.int32), np.ones(r.size)


def _matrix(name):
    if name == "1138_bus":
        return read_mtx(os.path.join(HERE, "golden", "1138_bus.mtx"))
    if name == "heart2":
        return _heart2()
    if name == "poisson_4096":
        d = gen.poisson_4096()
        return d["m"], d["n"], d["row"], d["col"], d["val"]
    if name == "dup_dense":
        # duplicates, dense blocks with a zero, CSR blocks, valued
        rng = np.random.default_rng(11)
        m = n = 96
        D = rng.uniform(-1, 1, (m, n)) * (rng.random((m, n)) < 0.6)
        D[:16, :16] = rng.uniform(-1, 1, (16, 16))
        D[3, 5] = 0.0
        r, c = np.nonzero(D)
        v = D[r, c]
        r, c, v = np.concatenate([r, r[:50]]), np.concatenate([c, c[:50]]), np.concatenate([v, v[:50]])
        o = np.argsort(r, kind="stable")
        return m, n, r[o].astype(np.int32), c[o].astype(np.int32), v[o]
    if name == "unit_dense":
        # unit values with DENSE blocks: the row-mask payload
        rng = np.random.default_rng(12)
        m = n = 128
        D = (rng.random((m, n)) < 0.3).astype(float)
        D[32:64, 32:64] = 1.0
        D[40, 40] = 0.0
        r, c = np.nonzero(D)
        return m, n, r.astype(np.int32), c.astype(np.int32), np.ones(r.size)
    if name == "random":
        r, c, v = gen.random_coo(1500, 2500, 9, seed=5, empty_rows=0.2)
        return 1500, 2500, r, c, v
    raise KeyError(name)


OPTION_SETS = [
    dict(bs=16, row_owned=True, pattern_wave=True),
    dict(bs=16, row_owned=False, pattern_wave=True),
    dict(bs=16, row_owned=True, pattern_wave=False),
    dict(bs=32, row_owned=True, pattern_wave=True),
    dict(bs=8, row_owned=False, pattern_wave=False),
    dict(bs=64, row_owned=True, pattern_wave=True),
]
MATRICES = ["1138_bus", "heart2", "poisson_4096", "dup_dense", "unit_dense", "random"]

_ARRAYS = ["gb_row_idx", "gb_col_idx", "nnz_per_blk", "type_per_blk", "segment_rows", "segment_begin",
           "segment_end", "segment_direct", "segment_uniform", "pattern_row_ptr", "pattern_col_idx",
           "pattern_output_ptr", "pattern_output_rows", "block_cols", "block_cols_offset", "pl_idx", "pl_val"]


def _product(mat, o):
    m, n, r, c, v = mat
    p = Plan(m, n, r, c, v, block_size=o["bs"], row_owned=o["row_owned"], pattern_wave=o["pattern_wave"])
    return p


def _arrays_of(getter, st):
    a = {k: getter(k) for k in _ARRAYS}
    a["gb_m"], a["gb_n"] = st["gb_m"], st["gb_n"]
    return a


def _region_names(o, gather):
    names = ["gb_row_idx", "gb_col_idx", "nnz_per_blk", "type_per_blk", "block_addr", "x", "y", "total"]
    if o["row_owned"]:
        names += ["segment_rows", "segment_begin", "segment_end", "segment_direct", "segment_uniform",
                  "output_epochs"]
    if o["pattern_wave"]:
        names += ["pattern_row_ptr", "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows"]
    if gather:
        names += ["block_cols", "block_cols_offset"]
    return names


@pytest.mark.parametrize("o", OPTION_SETS, ids=lambda o: f"bs{o['bs']}_ro{int(o['row_owned'])}_pw{int(o['pattern_wave'])}")
@pytest.mark.parametrize("name", MATRICES)
def test_product_image_equals_closed_form(name, o):
    mat = _matrix(name)
    p = _product(mat, o)
    st = p.stats()
    a = _arrays_of(p.export, st)
    off, img, unit = closed_form_image(a, o["bs"], o["row_owned"], o["pattern_wave"],
                                       bool(st["gather_flag"]), bool(st["compact_coo_indices"]))
    assert bool(st["unit_values"]) == unit
    for reg in _region_names(o, bool(st["gather_flag"])):
        assert p.layout_offset(reg) == off[reg], reg
    got = p.export("image")
    assert got.size == img.size
    if not np.array_equal(got, img):
        bad = np.nonzero(got != img)[0]
        inv = sorted(off.items(), key=lambda kv: kv[1])
        region = [k for k, v in inv if v <= bad[0]][-1]
        raise AssertionError(f"{bad.size} bytes differ, first at {bad[0]} (region {region})")
    assert int(st["slab_bytes"]) == off["total"]
    p.close()


@pytest.mark.parametrize("o", [o for o in OPTION_SETS if o["bs"] in (16, 32)],
                         ids=lambda o: f"bs{o['bs']}_ro{int(o['row_owned'])}_pw{int(o['pattern_wave'])}")
@pytest.mark.parametrize("name", MATRICES)
def test_reference_arrays_give_product_image(name, o):
    """The same closed form fed with the reference builder's own arrays
    (oracle/_ref) reproduces the product image byte for byte."""
    if not pyoracle.ref_available(bs=o["bs"], waves=4, row_owned=o["row_owned"]):
        pytest.skip("oracle/_ref not built (the reference checkout is absent)")
    m, n, r, c, v = _matrix(name)
    rb = pyoracle.RefBuild(m, n, r, c, v, bs=o["bs"], waves=4, row_owned=o["row_owned"])
    a = {k: rb.array(k) for k in _ARRAYS}
    a["gb_m"], a["gb_n"] = int(rb.scalar("gb_m")), int(rb.scalar("gb_n"))
    off, img, _ = closed_form_image(a, o["bs"], o["row_owned"], o["pattern_wave"],
                                    bool(rb.scalar("gather_flag")), bool(rb.scalar("compact_coo_indices")))
    p = _product((m, n, r, c, v), o)
    assert np.array_equal(p.export("image"), img)
    assert p.layout_offset("total") == off["total"]
    p.close()
    rb.close()


@pytest.mark.parametrize("mutate", ["no_pad", "swap_rowcol", "seg_late", "drop_epochs"])
@pytest.mark.parametrize("name", ["poisson_4096", "dup_dense"])
def test_closed_form_catches_planted_mistakes(name, mutate):
    """Each planted packer mistake makes the closed form disagree with the
    product's image or its offsets, so the comparison above would catch the
    same mistake in either packer."""
    o = dict(bs=16, row_owned=True, pattern_wave=True)
    p = _product(_matrix(name), o)
    st = p.stats()
    a = _arrays_of(p.export, st)
    args = (16, True, True, bool(st["gather_flag"]), bool(st["compact_coo_indices"]))
    off, good, _ = closed_form_image(a, *args)
    assert np.array_equal(p.export("image"), good)
    off2, bad, _ = closed_form_image(a, *args, mutate=mutate)
    differs = bad.size != good.size or not np.array_equal(bad, good) or off2 != off
    assert differs, mutate
    p.close()
