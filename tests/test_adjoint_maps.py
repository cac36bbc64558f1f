"""CPU checks of the adjoint's host structures (no GPU):

  * the transposed plan is the CB build of A^T's COO (the input reordered
    column-major, stable within a column): its GLOBAL_BLOCK arrays equal the
    oracle builder's on that COO, bit for bit;
  * slot_of[k] / grad_slot[k] point at the forward / transposed stream entry
    that holds input entry k (decoded from the stream keys);
  * emulating the kernels with those maps gives the textbook results: the
    per-entry value sums reproduce the stream values (what cbspmv_set_values
    writes), and the stored-order SDDMM permuted by grad_slot equals
    g[row_k] * x[col_k] for every input entry.
"""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle
from test_stream_layout import decode_stream


class _Transposed:
    """Export view of a plan's transposed plan."""

    def __init__(self, plan):
        self.plan = plan

    def export(self, name):
        return self.plan.export("transpose." + name)


def _block_dense(nb=12, bs=16, seed=3):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for b in range(nb):
        rr, cc = np.meshgrid(np.arange(bs), np.arange(bs), indexing="ij")
        keep = rng.random(bs * bs) < 0.97
        rows.append((b * bs + rr.ravel())[keep])
        cols.append((b * bs + cc.ravel())[keep])
        rows.append(b * bs + rng.integers(0, bs, 6))
        cols.append(rng.integers(0, nb * bs, 6))
    r, c = np.concatenate(rows), np.concatenate(cols)
    o = np.lexsort((c, r))
    r, c = r[o].astype(np.int32), c[o].astype(np.int32)
    return nb * bs, nb * bs, r, c, rng.uniform(-1, 1, r.size)


def _cases():
    yield "dup", (700, 900) + gen.random_coo(700, 900, 12, seed=101, dup=True)
    yield "empty_rows", (1001, 777) + gen.random_coo(1001, 777, 6, seed=8, empty_rows=0.3)
    yield "dense_blocks", _block_dense()
    r, c, v = gen.random_coo(300, 40000, 4, seed=6)
    lc = np.sort(np.random.default_rng(6).choice(40000, 20000, replace=False)).astype(np.int32)
    at = np.searchsorted(r, 137)
    r = np.concatenate([r[:at], np.full(lc.size, 137, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc, c[at:]])
    v = np.concatenate([v[:at], np.linspace(-1, 1, lc.size), v[at:]])
    yield "long_row", (300, 40000, r, c, v)
    d = gen.poisson_4096()
    yield "poisson_4096", (d["m"], d["n"], d["row"], d["col"], d["val"])


CASES = dict(_cases())


def transposed_coo(m, n, r, c, v):
    o = np.argsort(c, kind="stable")
    return n, m, c[o], r[o], v[o]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("bs", [16, 32])
def test_transposed_plan_is_cb_build_of_transpose(name, bs):
    m, n, r, c, v = CASES[name]
    p = Plan(m, n, r, c, v, block_size=bs)
    p.build_transpose()
    tm, tn, tr, tc, tv = transposed_coo(m, n, r, c, v)
    ob = pyoracle.OracleBuild(tm, tn, tr, tc, tv, bs=bs)
    for a in pyoracle.ARRAY_NAMES:
        assert np.array_equal(p.export("transpose." + a), ob.array(a)), a
    ob.close()
    p.close()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_entry_maps_and_kernel_emulation(name, dtype):
    m, n, r, c, v = CASES[name]
    p = Plan(m, n, r, c, v, dtype=dtype)
    p.build_transpose()
    slot = p.export("slot_of").astype(np.int64)
    gslot = p.export("grad_slot").astype(np.int64)
    assert slot.size == gslot.size == r.size

    frow, fcol, fval, fpad, *_ = decode_stream(p, dtype)
    assert not fpad[slot].any()
    assert np.array_equal(frow[slot], r) and np.array_equal(fcol[slot], c)

    t = _Transposed(p)
    trow, tcol, tval, tpad, *_ = decode_stream(t, dtype)
    assert trow.size == p.stored_entries()
    assert not tpad[gslot].any()
    assert np.array_equal(trow[gslot], c) and np.array_equal(tcol[gslot], r)

    # value groups: the stream value of an entry is the sum of its inputs
    for s_map, sval in ((slot, fval), (gslot, tval)):
        sums = np.zeros(sval.size)
        np.add.at(sums, s_map, v)
        held = np.zeros(sval.size, bool)
        held[s_map] = True
        np.testing.assert_allclose(sval[held], sums[held], rtol=1e-14, atol=1e-14)
        assert np.all(sval[~held] == 0.0)

    # stored-order SDDMM (g gathered by column of A^T = row of A, x by row of
    # A^T) permuted to input order
    rng = np.random.default_rng(1)
    g, x = rng.uniform(-1, 1, m), rng.uniform(-1, 1, n)
    stored = np.where(tpad, 0.0, g[np.minimum(tcol, m - 1)] * x[np.minimum(trow, n - 1)])
    assert np.array_equal(stored[gslot], g[r] * x[c])
    p.close()


def test_transpose_requires_build():
    from cbspmv import CbspmvError

    m, n, r, c, v = CASES["dup"]
    p = Plan(m, n, r, c, v)
    assert p.stored_entries() == -1
    with pytest.raises(CbspmvError):
        p.export("transpose.gb_row_idx")
    p.close()


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_value_positions_address_the_slab(dtype):
    """The byte offsets cbspmv_set_values writes to (stream_value_offset)
    hold each entry's value in the device slab."""
    m, n, r, c, v = CASES["dense_blocks"]
    p = Plan(m, n, r, c, v, dtype=dtype, item_nnz=700)
    pos = p.export("stream_value_pos")
    slab = p.export("stream_slab")
    val = p.export("stream_val")
    vb = {"f64": 8, "f32": 4, "bf16": 2}[dtype]
    assert pos.size == val.size and np.all(pos % vb == 0) and pos.max() + vb <= slab.size
    assert np.unique(pos).size == pos.size
    raw = slab[(pos[:, None] + np.arange(vb)).ravel()].reshape(-1, vb).copy()
    if dtype == "f64":
        got = raw.view(np.float64).ravel()
        assert np.array_equal(got, val)
    elif dtype == "f32":
        got = raw.view(np.float32).ravel()
        assert np.array_equal(got, val.astype(np.float32))
    else:
        h = raw.view(np.uint16).ravel().astype(np.uint32) << 16
        got = h.view(np.float32)
        f = val.astype(np.float32).view(np.uint32).astype(np.uint64)
        want = (((f + 0x7FFF + ((f >> 16) & 1)) >> 16).astype(np.uint32) << 16).view(np.float32)
        assert np.array_equal(got, want)
    p.close()
