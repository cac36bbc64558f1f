"""Seeded synthetic workloads (row-major COO + x) for the five benchmark
matrices.  Every generator uses numpy's PCG64 (`numpy.random.default_rng`)
with the stated seed and only vectorised numpy calls, so a given numpy version
reproduces the same matrices on every host.

  poisson_4096
      4096^2, per-row nnz ~ Poisson(10) floored at 1, distinct uniform
      columns, values U[-1,1], x U[-1,1], seed 0.  Within a row, columns keep
      their draw order.
  stencil_powerlaw_1m
      1,048,576^2: rows [0, 2^19) are a clipped 9-point stencil on a 1024 x
      1024 grid (row r = grid point (r // 1024, r % 1024), neighbours in
      (di, dj) lexicographic order); rows [2^19, 2^20) have power-law lengths
      (alpha = 2, max 8192) with 70% of columns uniform within +-4096 of the
      diagonal (reflected at the matrix edge) and 30% uniform over all
      columns, no repeats in a row.  Calibration: the literal minimum length
      1 gives ~9.2M nnz, far from the ~30M the workload names, so lengths are
      floor(x) of a Pareto(alpha) truncated to [x_min, 8192] with x_min = 7,
      clipped to >= 1, which gives ~30.5M nnz.  Values U[-1,1], x U[-1,1].
  clustered_2m
      2,097,152^2, ~50M nnz: random distinct 64-aligned 64 x 64 tiles, each
      cell present with probability 0.05, plus 4 distinct uniform columns per
      row; values N(0,1) * 10^U(-3,3) with one scale per row; entries sorted
      by (row, col).
  batch_256 (one matrix: batch_matrix)
      256 square matrices, n log-uniform in [1e3, 2e5], per-row nnz ~
      Poisson(U[5,50]) floored at 1, banded (+-max(32, 4 lambda)) or uniform
      columns chosen per matrix, 1-8 FP32 right-hand sides U[-1,1].
  powerlaw_512k
      524,288^2, power-law lengths (alpha = 2.2, max 4096, x_min = 9
      calibrated to ~20M nnz), 70% of columns within +-2048 of the diagonal
      and 30% uniform, no repeats; values U[-1,1], x U[-1,1].
"""
import numpy as np

MASTER_SEED = 20251203


def _dedupe_rows(rng, rows, cols, n, redraw):
    """Redraws entries whose (row, col) repeats an earlier entry of the row
    (earlier in draw order) until every row has distinct columns."""
    for _ in range(64):
        key = rows.astype(np.int64) * n + cols
        order = np.argsort(key, kind="stable")
        sk = key[order]
        dup = np.zeros(key.size, dtype=bool)
        dup[order[1:][sk[1:] == sk[:-1]]] = True
        nd = int(dup.sum())
        if nd == 0:
            return cols
        idx = np.nonzero(dup)[0]
        cols[idx] = redraw(rng, rows[idx])
    raise RuntimeError("could not make row columns distinct")


def _row_lengths_expand(lengths, first_row):
    return np.repeat(np.arange(first_row, first_row + lengths.size, dtype=np.int64), lengths)


def _power_law_block(rng, first_row, nrows, n, alpha, x_min, l_max, near, near_frac):
    """Rows [first_row, first_row + nrows) with truncated-Pareto lengths."""
    u = rng.random(nrows)
    # inverse CDF of the Pareto(alpha) restricted to [x_min, l_max]
    a1 = alpha - 1.0
    lo, hi = x_min ** (-a1), l_max ** (-a1)
    x = (lo - u * (lo - hi)) ** (-1.0 / a1)
    lengths = np.clip(np.floor(x), 1, min(l_max, n)).astype(np.int64)
    rows = _row_lengths_expand(lengths, first_row)

    def draw(rng, r):
        r = r.astype(np.int64)
        pick_near = rng.random(r.size) < near_frac
        off = rng.integers(-near, near + 1, size=r.size)
        c_near = r + off
        c_near = np.where((c_near < 0) | (c_near >= n), r - off, c_near)
        c_far = rng.integers(0, n, size=r.size)
        return np.where(pick_near, c_near, c_far).astype(np.int64)

    cols = draw(rng, rows)
    cols = _dedupe_rows(rng, rows, cols, n, draw)
    return rows, cols


def poisson_4096(seed=0):
    rng = np.random.default_rng(seed)
    m = n = 4096
    k = np.maximum(1, rng.poisson(10, size=m))
    rows = np.repeat(np.arange(m, dtype=np.int64), k)
    cols = np.concatenate([rng.choice(n, size=int(kk), replace=False) for kk in k])
    val = rng.uniform(-1.0, 1.0, size=rows.size)
    x = rng.uniform(-1.0, 1.0, size=n)
    return dict(m=m, n=n, row=rows.astype(np.int32), col=cols.astype(np.int32), val=val, x=x,
                name="poisson10_4096")


def stencil9(first_grid_rows, grid=1024):
    """9-point stencil rows for grid points [0, first_grid_rows * grid)."""
    nrows = first_grid_rows * grid
    r = np.arange(nrows, dtype=np.int64)
    i, j = r // grid, r % grid
    rows, cols = [], []
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            ii, jj = i + di, j + dj
            ok = (ii >= 0) & (ii < grid) & (jj >= 0) & (jj < grid)
            rows.append(np.where(ok, r, -1))
            cols.append(ii * grid + jj)
    R = np.stack(rows, axis=1).reshape(-1)
    C = np.stack(cols, axis=1).reshape(-1)
    keep = R >= 0
    return R[keep], C[keep]


def stencil_powerlaw_1m(seed=MASTER_SEED + 2, x_min=7.0):
    rng = np.random.default_rng(seed)
    m = n = 1 << 20
    half = m // 2
    r1, c1 = stencil9(half // 1024)
    r2, c2 = _power_law_block(rng, half, m - half, n, alpha=2.0, x_min=x_min, l_max=8192,
                              near=4096, near_frac=0.7)
    rows = np.concatenate([r1, r2]).astype(np.int32)
    cols = np.concatenate([c1, c2]).astype(np.int32)
    val = rng.uniform(-1.0, 1.0, size=rows.size)
    x = rng.uniform(-1.0, 1.0, size=n)
    return dict(m=m, n=n, row=rows, col=cols, val=val, x=x, name="stencil_powerlaw_1M")


def clustered_2m(seed=MASTER_SEED + 3, scale=1):
    """scale > 1 shrinks m, n and the cluster count by that factor (tests)."""
    rng = np.random.default_rng(seed)
    m = n = (1 << 21) // scale
    tiles = m // 64
    n_clusters = 203_000 // scale
    tid = rng.choice(tiles * tiles, size=n_clusters, replace=False)
    ti, tj = tid // tiles, tid % tiles
    counts = rng.binomial(4096, 0.05, size=n_clusters)
    cl = np.repeat(np.arange(n_clusters), counts)
    cell = rng.integers(0, 4096, size=cl.size)

    def redraw_cell(rng, c):
        return rng.integers(0, 4096, size=c.size)

    cell = _dedupe_rows(rng, cl, cell, 4096, redraw_cell)
    rows_c = ti[cl] * 64 + cell // 64
    cols_c = tj[cl] * 64 + cell % 64
    rows_b = np.repeat(np.arange(m, dtype=np.int64), 4)
    cols_b = rng.integers(0, n, size=rows_b.size)
    rows = np.concatenate([rows_c, rows_b])
    cols = np.concatenate([cols_c, cols_b])

    def redraw_col(rng, r):
        return rng.integers(0, n, size=r.size)

    # background entries come last, so duplicates are always redrawn there
    cols = _dedupe_rows(rng, rows, cols, n, redraw_col)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order].astype(np.int32), cols[order].astype(np.int32)
    scale_row = 10.0 ** rng.uniform(-3.0, 3.0, size=m)
    val = rng.standard_normal(rows.size) * scale_row[rows]
    x = rng.uniform(-1.0, 1.0, size=n)
    return dict(m=m, n=n, row=rows, col=cols, val=val, x=x, name="block_clustered_2M")


def batch_matrix(i, master_seed=MASTER_SEED + 4):
    rng = np.random.default_rng([master_seed, i])
    n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
    lam = rng.uniform(5.0, 50.0)
    banded = bool(rng.random() < 0.5)
    k = np.minimum(np.maximum(1, rng.poisson(lam, size=n)), n)
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    bw = max(32, int(4 * lam))

    def draw(rng, r):
        if banded:
            c = r + rng.integers(-bw, bw + 1, size=r.size)
            return np.clip(c, 0, n - 1)
        return rng.integers(0, n, size=r.size)

    cols = _dedupe_rows(rng, rows, draw(rng, rows), n, draw) if n > 2 * bw or not banded else None
    if cols is None:   # tiny banded matrix: sample distinct columns per row directly
        cols = np.concatenate([rng.choice(n, size=int(kk), replace=False) for kk in k])
    val = rng.uniform(-1.0, 1.0, size=rows.size)
    nrhs = int(rng.integers(1, 9))
    X = rng.uniform(-1.0, 1.0, size=(nrhs, n)).astype(np.float32)
    return dict(m=n, n=n, row=rows.astype(np.int32), col=cols.astype(np.int32), val=val, X=X,
                nrhs=nrhs, banded=banded, lam=lam)


def batch_256(count=256, master_seed=MASTER_SEED + 4, workers=None):
    """The matrices are independent (one seed each), so they are drawn by
    `workers` threads (default: up to 16; numpy releases the GIL in the
    bulk draws and sorts); the result does not depend on the worker count."""
    import os

    workers = min(16, os.cpu_count() or 1) if workers is None else int(workers)
    if workers <= 1 or count < 4:
        return [batch_matrix(i, master_seed) for i in range(count)]
    import concurrent.futures as cf

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda i: batch_matrix(i, master_seed), range(count)))


def powerlaw_512k(seed=MASTER_SEED + 5, m=1 << 19, x_min=9.0):
    rng = np.random.default_rng(seed)
    n = m
    rows, cols = _power_law_block(rng, 0, m, n, alpha=2.2, x_min=x_min, l_max=4096,
                                  near=2048, near_frac=0.7)
    val = rng.uniform(-1.0, 1.0, size=rows.size)
    x = rng.uniform(-1.0, 1.0, size=n)
    return dict(m=m, n=n, row=rows.astype(np.int32), col=cols.astype(np.int32), val=val, x=x,
                name="powerlaw_512k_fp32")


def random_coo(m, n, per_row, seed=0, dup=False, unit=False, empty_rows=0.0):
    """Small random row-major matrices for parity tests."""
    rng = np.random.default_rng(seed)
    k = np.maximum(0 if empty_rows > 0 else 1, rng.poisson(per_row, size=m))
    if empty_rows > 0:
        k[rng.random(m) < empty_rows] = 0
    k = np.minimum(k, n if not dup else 4 * n)
    rows = np.repeat(np.arange(m, dtype=np.int64), k)
    if dup:
        cols = rng.integers(0, n, size=rows.size)
    else:
        cols = np.concatenate([rng.choice(n, size=int(kk), replace=False) for kk in k]) if rows.size else np.zeros(0, np.int64)
    val = np.ones(rows.size) if unit else rng.uniform(-1.0, 1.0, size=rows.size)
    return rows.astype(np.int32), cols.astype(np.int32), val


def reference_x(n, seed=42):
    """x of the reference's benchmark driver (cb-spmv-dcu/src/main.hip.cpp:58-66):
    std::mt19937(42) with std::uniform_real_distribution<double>(0, 1).  The
    legacy numpy RandomState seeds MT19937 the same way, and libstdc++'s
    generate_canonical<double, 53> combines two 32-bit outputs g1, g2 as
    (g1 + g2 * 2^32) / 2^64 in double arithmetic (clamped below 1)."""
    g = np.random.RandomState(seed).randint(0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    x = (g[0::2] + g[1::2] * 2.0 ** 32) / 2.0 ** 64
    x[x >= 1.0] = np.nextafter(1.0, 0.0)
    return x
