"""CB-SpMV benchmark on one MI355X: prints one JSON line.

A step is one SpMV y = A x over the resident plan (the reference protocol,
cb-spmv-dcu/src/cb-spmv-dcu.hpp:1631-1660: W untimed warm-up launches, then K
timed launches between events).  Default workload: stencil_powerlaw_1m (1,048,576^2,
~30.2M nnz: 9-point stencil half + power-law half, FP64, 16x16 blocks), the
largest single-GPU configuration the metric is quoted on.

value      = algorithmic (CSR-equivalent) bytes per SpMV / time per SpMV, GB/s:
             B_alg = nnz (s_v + 4) + 4 (m + 1) + s_x n + s_y m
roofline   = the same bytes over the average launch duration (HIP events on
             the launch stream) against 8 TB/s; a plan with a scattered-entry
             split launches two kernels per SpMV, timed together (`kernels`);
             `traffic` is read from the committed rocprofv3 PMC summary for
             this workload when present
cpu_baseline = the reference's own CPU path (check_results_consistency, its
             sequential FP64 SpMV, compiled from the reference sources into
             oracle/_ref), timed on the host on the same matrix; the oracle's
             restatement of the same loop when oracle/_ref is absent
extensions = the precision, autograd and batched workloads below, each run
             as a child process after the default line's own measurement
             (their metric, value and results each); --no-extensions skips
             them

Other workloads (--workload), one JSON line each:
  precision  the 2,097,152^2 block-clustered matrix (~50M nnz, values scaled
             by 10^U(-3,3) per row) with values stored FP64 / FP32 / BF16,
             FP32 native and compensated accumulation: GB/s of each, FP32/FP64
             throughput ratio, relative L2 error of each against the FP64
             oracle
  autograd   the 524,288^2 power-law matrix (~20M nnz, FP32) through
             torch.autograd (CBSparse with stored-order values,
             cb_spmv_autograd + backward): the line's value is fwd+bwd / fwd
             with both timed inside the training step; beside it the
             inference forward, the whole step eager and as a captured graph,
             the kernels alone, the value push after an optimiser step, the
             input-COO-order backward variants, gradient errors against FP64
  batched    256 matrices (n log-uniform in [1e3, 2e5], 1-8 FP32 right-hand
             sides each): autotuned plans (16/32 blocks x 4/9 waves), one
             persistent launch, eager and hipGraph-replayed; B_alg summed
             over matrices with x and y counted per right-hand side
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cb-sparse-dcu_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EXTENSION_TIMEOUT_S = 600   # each extension workload of the default line (they take 20-80 s)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(m, n, nnz, dtype):
    s_v = {"f64": 8, "f32": 4, "bf16": 2}[dtype]
    s_x = 8 if dtype == "f64" else 4
    return nnz * (s_v + 4) + 4 * (m + 1) + s_x * n + s_x * m


MATRICES = ("poisson_4096", "stencil_powerlaw_1m", "clustered_2m", "powerlaw_512k")


def load_workload(name):
    from cbspmv import gen

    if name not in MATRICES:
        raise SystemExit(f"unknown matrix {name}")
    return getattr(gen, name)()


def cpu_baseline(d, budget_s):
    """Times the reference CPU path on a bounded number of repetitions."""
    import numpy as np

    from oracle import pyoracle

    m, n, r, c, v, x = d["m"], d["n"], d["row"], d["col"], d["val"], d["x"]
    y_ref = pyoracle.spmv_seq(m, r, c, v, x)
    if pyoracle.ref_available():
        import ctypes

        L = pyoracle.ref_lib()
        ri, ci = np.ascontiguousarray(r, np.int32), np.ascontiguousarray(c, np.int32)
        vv, xx, yy = (np.ascontiguousarray(a, np.float64) for a in (v, x, y_ref))
        P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
        args = (int(ri.size), P(ri, ctypes.c_int), P(ci, ctypes.c_int), P(vv, ctypes.c_double), int(m), int(n),
                P(xx, ctypes.c_double), P(yy, ctypes.c_double))
        t1 = L.ref_time_check(*args, 1)
        reps = max(3, min(200, int(budget_s / max(t1, 1e-6))))
        t = L.ref_time_check(*args, reps)
        kind, what = "reference", "check_results_consistency (utils.h:85-136) from oracle/_ref"
    else:
        t0 = time.perf_counter()
        pyoracle.spmv_seq(m, r, c, v, x)
        t1 = time.perf_counter() - t0
        reps = max(3, min(200, int(budget_s / max(t1, 1e-6))))
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pyoracle.spmv_seq(m, r, c, v, x)
            ts.append(time.perf_counter() - t0)
        t = sorted(ts)[len(ts) // 2]
        kind, what = "port", "oracle restatement of utils.h:98-111"
    return y_ref, t, reps, kind, what


VENDOR_ALGS = {"csr_adaptive": 2, "csr_rowsplit": 3, "csr_lrb": 7, "csr_nnzsplit": 8}
BSR_BLOCK_DIMS = (2, 4, 8, 16)   # the reference's hipSPARSE BSR baseline sizes (benchmark/hipSparse-bsr/macros.h:9)


def vendor_baseline(d, dtype, reps, ours_ms):
    """rocSPARSE CSR SpMV (each rocsparse_spmv algorithm) and BSR SpMV
    (rocsparse_bsrmv at block dims 2, 4, 8, 16 after rocsparse_csr2bsr) from
    libvendor_csr.so on the same matrix and GPU, timed like ours; the
    comparators the reference's hipSPARSE baseline programs stand for
    (benchmark/hipSparse-csr/main.hip.cpp:104-181,
    benchmark/hipSparse-bsr/main.hip.cpp:99-187).  BSR launches move whole
    blocks (up to 256 values each), so they run fewer timed launches."""
    import ctypes

    import numpy as np

    from oracle import pyoracle

    path = os.path.join(ROOT, "cb-sparse-dcu_amd", "libvendor_csr.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.vendor_csr_spmv.restype = ctypes.c_int
    L.vendor_csr_spmv.argtypes = [ctypes.c_int] * 3 + [ip, ip, dp, dp, dp] + [ctypes.c_int] * 4 + [dp]
    L.vendor_bsr_spmv.restype = ctypes.c_int
    L.vendor_bsr_spmv.argtypes = ([ctypes.c_int] * 3 + [ip, ip, dp, dp, dp] + [ctypes.c_int] * 4 +
                                  [dp, ctypes.POINTER(ctypes.c_longlong)])
    m, n, r, c, v, x = d["m"], d["n"], d["row"], d["col"], d["val"], d["x"]
    ptr = np.zeros(m + 1, np.int32)
    np.cumsum(np.bincount(r, minlength=m), out=ptr[1:])
    order = np.lexsort((c, r))   # rocSPARSE's CSR: columns ascending within a row
    col = np.ascontiguousarray(c[order], np.int32)
    val = np.ascontiguousarray(v[order], np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    y = np.zeros(m)
    y_ref = pyoracle.spmv_csr_omp(m, r, c, v, x, min(16, os.cpu_count() or 1))
    b_alg = algorithmic_bytes(m, n, int(r.size), dtype)
    out = {}
    for name, alg in VENDOR_ALGS.items():
        ms = ctypes.c_double()
        rc = L.vendor_csr_spmv(m, n, int(r.size), ptr.ctypes.data_as(ip), col.ctypes.data_as(ip),
                               val.ctypes.data_as(dp), xx.ctypes.data_as(dp), y.ctypes.data_as(dp), alg,
                               int(dtype != "f64"), 20, reps, ctypes.byref(ms))
        if rc != 0:
            out[name] = {"status": rc}
            continue
        out[name] = {"us": round(ms.value * 1e3, 2), "GB_s": round(b_alg / (ms.value * 1e-3) / 1e9, 1),
                     "rel_l2_vs_fp64": pyoracle.rel_l2(y, y_ref)}
        log(f"rocSPARSE {name}: {ms.value * 1e3:.1f} us")
    for bd in BSR_BLOCK_DIMS:
        name = f"bsr_{bd}"
        ms, nnzb = ctypes.c_double(), ctypes.c_longlong()
        y[:] = 0.0
        rc = L.vendor_bsr_spmv(m, n, int(r.size), ptr.ctypes.data_as(ip), col.ctypes.data_as(ip),
                               val.ctypes.data_as(dp), xx.ctypes.data_as(dp), y.ctypes.data_as(dp), bd,
                               int(dtype != "f64"), 3, max(5, reps // 10), ctypes.byref(ms), ctypes.byref(nnzb))
        if rc != 0:
            out[name] = {"status": rc}
            continue
        out[name] = {"us": round(ms.value * 1e3, 2), "GB_s": round(b_alg / (ms.value * 1e-3) / 1e9, 1),
                     "blocks": int(nnzb.value), "rel_l2_vs_fp64": pyoracle.rel_l2(y, y_ref)}
        log(f"rocSPARSE {name}: {ms.value * 1e3:.1f} us ({nnzb.value} blocks)")
    ok = {k: e for k, e in out.items() if "us" in e}
    best = min(ok, key=lambda k: ok[k]["us"]) if ok else None
    return {"library": "rocSPARSE (ROCm): CSR rocsparse_spmv and BSR rocsparse_bsrmv, same matrix and GPU",
            "algorithms": out,
            "best": best, "best_us": ok[best]["us"] if best else None,
            "cbspmv_speedup_vs_best": round(ok[best]["us"] / (ours_ms * 1e3), 3) if best else None}


def cpu_parallel(d, threads, reps=10):
    """The oracle's row-parallel CSR form of the same sum (OpenMP over
    `threads` host threads), with the CSR arrays built once outside the
    timer: median seconds per SpMV over `reps` runs after one warm-up,
    timed in C++ around the OpenMP loop alone."""
    from oracle import pyoracle

    csr = pyoracle.CSR(d["m"], d["row"], d["col"], d["val"])
    return csr.time_spmv(d["x"], threads, reps)


def timed(fn, steps, warmup, stream):
    """Mean ms per call of fn over `steps` calls after `warmup` calls,
    between events on `stream` (the stream fn launches on)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_precision(args):
    """FP64 / FP32 / BF16 storage on the block-clustered matrix, then FP64
    and FP32 on the 30M-entry matrix of the default line."""
    import numpy as np
    import torch

    import cbspmv
    from cbspmv import gen
    from oracle import pyoracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    d = gen.clustered_2m()
    m, n, nnz = d["m"], d["n"], int(d["row"].size)
    log(f"precision matrix {m}x{n}, nnz {nnz}, generated in {time.perf_counter() - t0:.1f}s")
    threads = min(16, os.cpu_count() or 1)
    y64 = pyoracle.spmv_csr_omp(m, d["row"], d["col"], d["val"], d["x"], threads)
    stream = torch.cuda.current_stream()
    res = {}
    for dtype, comp in (("f64", False), ("f32", False), ("f32", True), ("bf16", False), ("bf16", True)):
        plan = cbspmv.Plan(m, n, d["row"], d["col"], d["val"], dtype=dtype, compensated=comp)
        vdt = plan.vector_dtype
        x = torch.tensor(d["x"], dtype=vdt, device=dev)
        y = torch.empty(m, dtype=vdt, device=dev)
        plan.upload(stream)
        ms = timed(lambda: plan.spmv(x, y, stream), args.steps, args.warmup, stream)
        rel = pyoracle.rel_l2(y.double().cpu().numpy(), y64)
        b_alg = algorithmic_bytes(m, n, nnz, dtype)
        key = dtype + ("_comp" if comp else "")
        res[key] = {"us": round(ms * 1e3, 2), "GB_s": round(b_alg / (ms * 1e-3) / 1e9, 1),
                    "GFLOP_s": round(2 * nnz / (ms * 1e-3) / 1e9, 1), "rel_l2_vs_fp64": rel,
                    "bytes_alg": b_alg, "stream_slab_bytes": plan.stats()["stream_bytes"]}
        log(f"{key}: {ms * 1e3:.1f} us, rel {rel:.2e}")
        plan.close()
    speed = res["f64"]["us"] / res["f32"]["us"]
    # the same comparison on the 30M-entry matrix of the default line, whose
    # FP32 x (4 MB) fits one XCD's L2 where the FP64 x (8 MB) does not
    d2 = gen.stencil_powerlaw_1m()
    m2, n2, nnz2 = d2["m"], d2["n"], int(d2["row"].size)
    cfg2 = {}
    for dtype in ("f64", "f32"):
        plan = cbspmv.Plan(m2, n2, d2["row"], d2["col"], d2["val"], dtype=dtype)
        vdt = plan.vector_dtype
        x = torch.tensor(d2["x"], dtype=vdt, device=dev)
        y = torch.empty(m2, dtype=vdt, device=dev)
        plan.upload(stream)
        ms = timed(lambda: plan.spmv(x, y, stream), args.steps, args.warmup, stream)
        cfg2[dtype] = {"us": round(ms * 1e3, 2),
                       "GB_s": round(algorithmic_bytes(m2, n2, nnz2, dtype) / (ms * 1e-3) / 1e9, 1)}
        plan.close()
    cfg2["f32_over_f64"] = round(cfg2["f64"]["us"] / cfg2["f32"]["us"], 3)
    res["stencil_powerlaw_1m"] = cfg2
    log(f"stencil_powerlaw_1m: f64 {cfg2['f64']['us']} us, f32 {cfg2['f32']['us']} us")
    return {"metric": "FP32-storage SpMV throughput / FP64 on the mixed-precision matrix", "value": round(speed, 3),
            "unit": "x", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res["f32"]["us"] / 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator, cbspmv/gen.py clustered_2m)",
            "config": {"workload": "clustered_2M_50M_nnz_mixed_precision", "m": m, "n": n, "nnz": nnz},
            "results": res}


def bench_decompose(args):
    """Times sub-matrices of stencil_powerlaw_1m on their own, each as its own plan run
    by the same kernel: the stencil half, the power-law half's entries near
    the diagonal (|col - row| <= 4096, the columns the generator draws near
    the row) and its uniformly scattered entries, and everything but the
    scattered entries.  Their times against the whole matrix's show where
    the launch spends its time and how far the parts overlap.

    Then the whole matrix again with only the scattered entries' columns
    moved (rows, values and every other entry unchanged; each a valid
    matrix whose product the kernel computes in full): folded into the
    first 131,072 or 524,288 columns (1 or 4 MB of FP64 x, gathered from
    L2), or next to the diagonal (row + 4097 + col mod 2048, reflected at
    the edge: columns the tile stages in LDS).  The three times bound what
    the scattered gathers cost, and what any placement of them could gain."""
    import numpy as np
    import torch

    import cbspmv
    from cbspmv import gen

    dev = torch.device("cuda:0")
    d = gen.stencil_powerlaw_1m()
    m, n, r, c, v = d["m"], d["n"], d["row"], d["col"], d["val"]
    half = m // 2
    far = np.abs(c.astype(np.int64) - r.astype(np.int64)) > 4096
    parts = {"whole": None, "stencil_half": r < half, "power_law_near": (r >= half) & ~far,
             "power_law_scattered": (r >= half) & far, "all_but_scattered": ~far}
    r64, c64 = r.astype(np.int64), c.astype(np.int64)
    near = r64 + 4097 + c64 % 2048
    near = np.where(near < n, near, r64 - 4097 - c64 % 2048)
    moved = {"scattered_cols_in_1MB": np.where(far, c64 % 131072, c64),
             "scattered_cols_in_4MB": np.where(far, c64 % 524288, c64),
             "scattered_cols_near_diagonal": np.where(far, near, c64)}
    stream = torch.cuda.current_stream()
    res = {}
    for name, sel in list(parts.items()) + [(k, None) for k in moved]:
        rs, cs, vs = (r, c, v) if sel is None else (r[sel], c[sel], v[sel])
        if name in moved:
            cs = moved[name].astype(np.int32)
        plan = cbspmv.Plan(m, n, rs, cs, vs, dtype=args.dtype, split="off")   # the one-stream kernel on each part
        vdt = plan.vector_dtype
        x = torch.tensor(d["x"], dtype=vdt, device=dev)
        y = torch.empty(m, dtype=vdt, device=dev)
        plan.upload(stream)
        ms = timed(lambda: plan.spmv(x, y, stream), args.steps, args.warmup, stream)
        st = plan.stats()
        res[name] = {"nnz": int(rs.size), "us": round(ms * 1e3, 2), "stream_slab_bytes": st["stream_bytes"],
                     "slab_TB_s": round(st["stream_bytes"] / (ms * 1e-3) / 1e12, 2)}
        log(f"{name}: {rs.size} entries, {ms * 1e3:.2f} us")
        plan.close()
        del x, y
    res["overlap_us"] = round(res["all_but_scattered"]["us"] + res["power_law_scattered"]["us"]
                              - res["whole"]["us"], 2)
    return {"decomposition": "stencil_powerlaw_1m", "dtype": args.dtype, "steps": args.steps, "warmup": args.warmup,
            "parts": res}


def bench_autograd(args):
    """fwd+bwd / fwd of the differentiable SpMV through its torch.autograd
    wrapper: y = cb_spmv(A, x) with A.values (stored order) and x both
    requiring gradients, then y.backward(g) (dx = A^T g and the gradient of
    every stored value, one fused pass over A^T).  fwd is the plain forward
    launch.  `value` times both inside the training step (events before,
    between and after them, so each runs after the other as in training);
    also reported: the step as a captured graph replay and eagerly, the
    forward of an inference loop (each launch after another forward, the
    matrix partly cached), the kernels alone, the value push after an
    optimizer step, and the input-order alternatives."""
    import ctypes

    import numpy as np
    import torch

    import cbspmv
    from cbspmv import gen
    from oracle import pyoracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    d = gen.powerlaw_512k()
    m, n, r, c, v = d["m"], d["n"], d["row"], d["col"], d["val"]
    nnz = int(r.size)
    log(f"autograd matrix {m}x{n}, nnz {nnz}, generated in {time.perf_counter() - t0:.1f}s")
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(7)
    g64 = rng.uniform(-1, 1, m)
    x = torch.tensor(d["x"], dtype=torch.float32, device=dev, requires_grad=True)
    g = torch.tensor(g64, dtype=torch.float32, device=dev)
    threads = min(16, os.cpu_count() or 1)
    xd, gd = d["x"].astype(np.float32).astype(np.float64), g.double().cpu().numpy()
    o = np.argsort(c, kind="stable")
    dx64 = pyoracle.CSR(n, c[o], r[o], v[o]).spmv(gd, threads)
    dv64 = gd[r] * xd[c]

    t1 = time.perf_counter()
    A = cbspmv.CBSparse(m, n, r, c, v, dtype="f32", requires_grad=True)
    log(f"CBSparse (A and A^T) built in {time.perf_counter() - t1:.1f}s, {A.stored} stored entries")
    y = torch.empty(m, dtype=torch.float32, device=dev)
    xs = x.detach()
    fwd = timed(lambda: A.plan.spmv(xs, y, stream), args.steps, args.warmup, stream)

    def step():
        A.values.grad = None
        x.grad = None
        cbspmv.cb_spmv_autograd(A, x).backward(g)

    eager = timed(step, args.steps, args.warmup, stream)
    err_dx = pyoracle.rel_l2(x.grad.double().cpu().numpy(), dx64)
    err_dv = pyoracle.rel_l2(A.to_input(A.values.grad).double().cpu().numpy(), dv64)
    graph_ms = None
    try:
        side = torch.cuda.Stream()
        side.wait_stream(stream)
        with torch.cuda.stream(side):
            for _ in range(3):
                step()
        stream.wait_stream(side)
        torch.cuda.synchronize()
        cg = torch.cuda.CUDAGraph()
        A.values.grad = None
        x.grad = None
        with torch.cuda.graph(cg):
            cbspmv.cb_spmv_autograd(A, x).backward(g)
        graph_ms = timed(cg.replay, args.steps, args.warmup, stream)
        err_graph = pyoracle.rel_l2(A.to_input(A.values.grad).double().cpu().numpy(), dv64)
    except Exception as e:   # capture is a measurement aid; the eager step stays the reference
        log(f"graph capture of the autograd step failed: {e}")
        err_graph = None

    # the forward and the backward inside the training step: events before,
    # between and after, on the launch stream (each kernel then runs where a
    # training loop runs it, after the other one)
    marks = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    for e in marks:
        A.values.grad = None
        x.grad = None
        e[0].record(stream)
        yy = cbspmv.cb_spmv_autograd(A, x)
        e[1].record(stream)
        yy.backward(g)
        e[2].record(stream)
        del yy   # frees the step's autograd graph
    torch.cuda.synchronize()
    fwd_step = sum(e[0].elapsed_time(e[1]) for e in marks) / len(marks)
    bwd_step = sum(e[1].elapsed_time(e[2]) for e in marks) / len(marks)
    # the kernels alone
    lib = cbspmv.lib()
    h, sh = A.plan.handle, ctypes.c_void_p(stream.cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    dx = torch.empty(n, dtype=torch.float32, device=dev)
    dw = torch.empty(A.stored, dtype=torch.float32, device=dev)
    bwd = timed(lambda: lib.cbspmv_backward_stored(h, P(g), P(xs), P(dx), P(dw), sh), args.steps, args.warmup,
                stream)
    # pushing updated values (after an optimizer step) to A and A^T
    w = A.values.detach().clone()
    push = timed(lambda: A.plan.set_values_stored(w, stream), max(5, args.steps // 4), 2, stream)
    # input-order alternatives (kernel level)
    vin = torch.tensor(v.astype(np.float32), device=dev)
    dvin = torch.empty(nnz, dtype=torch.float32, device=dev)
    alt = {}
    for mode in ("direct", "fused"):
        plan = cbspmv.Plan(m, n, r, c, v, dtype="f32", backward=mode)
        plan.build_transpose()
        plan.upload(stream)
        hh = plan.handle
        t = timed(lambda: lib.cbspmv_backward(hh, P(g), P(xs), P(dx), P(dvin), sh), args.steps, args.warmup, stream)
        alt[f"input_order_{mode}_bwd_us"] = round(t * 1e3, 2)
        plan.close()
    push_in = timed(lambda: A.plan.set_values(vin, stream), max(5, args.steps // 4), 2, stream)
    A.close()
    ratio = (fwd_step + bwd_step) / fwd_step
    res = {"fwd_us": round(fwd * 1e3, 2),
           "fwd_in_training_step_us": round(fwd_step * 1e3, 2), "bwd_in_training_step_us": round(bwd_step * 1e3, 2),
           "ratio_in_training_step": round(ratio, 3),
           "wrapper_step_graph_us": round(graph_ms * 1e3, 2) if graph_ms is not None else None,
           "wrapper_step_eager_us": round(eager * 1e3, 2),
           "ratio_graph_over_inference_fwd": round(graph_ms / fwd, 3) if graph_ms is not None else None,
           "ratio_eager_with_host_gaps": round(eager / fwd, 3),
           "bwd_kernel_us": round(bwd * 1e3, 2), "ratio_kernels": round((fwd + bwd) / fwd, 3),
           "value_push_stored_us": round(push * 1e3, 2), "value_push_input_order_us": round(push_in * 1e3, 2),
           "rel_l2_dx": err_dx, "rel_l2_dvals": err_dv, "rel_l2_dvals_graph": err_graph, **alt}
    log(f"autograd: {res}")
    return {"metric": "fwd+bwd / fwd time of the differentiable SpMV through torch.autograd "
                      "(dx = A^T g and dvals of every stored value), both timed inside the training step",
            "value": round(ratio, 3), "unit": "x", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": fwd_step + bwd_step,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded generator, cbspmv/gen.py powerlaw_512k)",
            "config": {"workload": "powerlaw_512k_20M_nnz_autograd", "m": m, "n": n, "nnz": nnz}, "results": res}


def bench_batched(args):
    import numpy as np
    import torch

    import cbspmv
    from cbspmv import gen
    from oracle import pyoracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    mats = gen.batch_256(count=args.batch_count)
    log(f"{len(mats)} matrices, {sum(int(d['row'].size) for d in mats)} nnz, generated in "
        f"{time.perf_counter() - t0:.1f}s")
    stream = torch.cuda.current_stream()
    plans, picks, X, Y = [], {}, [], []
    t0 = time.perf_counter()
    for d in mats:
        if args.no_autotune:
            p = cbspmv.Plan(d["m"], d["n"], d["row"], d["col"], d["val"], dtype="f32")
        else:
            p, _ = cbspmv.autotune(d["m"], d["n"], d["row"], d["col"], d["val"], nrhs=d["nrhs"], dtype="f32")
            w = int(p.export("stream_tile_shape")[0])
            bs = (d["m"] + p.stats()["gb_m"] - 1) // max(1, p.stats()["gb_m"])
            bs = 16 if bs <= 16 else 32
            picks[f"bs{bs}_w{w}"] = picks.get(f"bs{bs}_w{w}", 0) + 1
        plans.append(p)
        X.append(torch.tensor(d["X"], dtype=torch.float32, device=dev))
        Y.append(torch.empty((d["nrhs"], d["m"]), dtype=torch.float32, device=dev))
    log(f"plans ready in {time.perf_counter() - t0:.1f}s; autotuner picks {picks}")
    b = cbspmv.Batch(plans, stream)
    batched = timed(lambda: b.spmv(X, Y, stream), args.steps, args.warmup, stream)
    # correctness: every right-hand side against the FP64 oracle
    worst = 0.0
    for d, y in zip(mats, Y):
        ys = y.double().cpu().numpy()
        for k in range(d["nrhs"]):
            ref = pyoracle.spmv_seq(d["m"], d["row"], d["col"], d["val"], d["X"][k].astype(np.float64))
            worst = max(worst, pyoracle.rel_l2(ys[k], ref))
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        b.capture(X, Y, gs)
    gs.synchronize()
    graph = timed(lambda: b.replay(gs), args.steps, args.warmup, gs)

    def eager_loop():
        for p, x, y in zip(plans, X, Y):
            for k in range(x.shape[0]):
                p.spmv(x[k], y[k], stream)
    loop = timed(eager_loop, max(3, args.steps // 20), 2, stream)
    b_alg = 0
    for d in mats:
        nnz = int(d["row"].size)
        b_alg += nnz * 8 + 4 * (d["m"] + 1) + d["nrhs"] * 4 * (d["n"] + d["m"])
    gbs = b_alg / (graph * 1e-3) / 1e9
    b.close()
    for p in plans:
        p.close()
    return {"metric": "batched SpMV GB/s (sum of B_alg, x and y per right-hand side), hipGraph replay",
            "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": graph, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded generator, cbspmv/gen.py batch_256)",
            "config": {"workload": f"batched_{len(mats)}_matrices_1to8_rhs", "matrices": len(mats),
                       "nnz": sum(int(d["row"].size) for d in mats), "rhs": sum(d["nrhs"] for d in mats)},
            "results": {"graph_replay_ms": round(graph, 4), "eager_batched_ms": round(batched, 4),
                        "per_matrix_launch_loop_ms": round(loop, 4), "bytes_alg": b_alg,
                        "GB_s_eager": round(b_alg / (batched * 1e-3) / 1e9, 1),
                        "worst_rel_l2_vs_fp64": worst, "autotune_picks": picks}}


def bench_mtx(args):
    """The reference driver's protocol on one .mtx file: its reader (values
    forced to 1), x from mt19937(42) U[0,1), GLOBAL_BLOCK + coo2block_gather,
    cb_spmv (validated first launch, 100 warm-up, 1000 timed launches), the
    report of print_benchmark_report on stderr; validation here is the
    oracle's FP64 SpMV at 1e-12 relative L2."""
    import cbspmv
    from cbspmv import gen
    from oracle import pyoracle

    m, n, r, c, v = cbspmv.read_mtx(args.mtx)
    nnz = int(r.size)
    x = gen.reference_x(n)
    gb = cbspmv.GLOBAL_BLOCK(m, n, nnz, block_size=args.block_size)
    cbspmv.coo2block_gather(r, c, v, gb)
    import numpy as np

    y = np.zeros(m)
    met = cbspmv.cb_spmv(gb, x, y)
    err = pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v, x))
    name = os.path.basename(args.mtx)
    cbspmv.print_benchmark_report(name, m, n, nnz, gb, met, err <= 1e-12, file=sys.stderr)
    b_alg = algorithmic_bytes(m, n, nnz, "f64")
    ms = met.runtime
    return {"metric": "SpMV GB/s (CSR-equivalent algorithmic bytes) on one MI355X", "value": round(b_alg / (ms * 1e-3) / 1e9, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": 1000, "warmup": 100, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": f"Matrix Market file {name}",
            "config": {"workload": f"mtx_{name}", "m": m, "n": n, "nnz": nnz, "block_size": args.block_size},
            "gflops": round(met.gflops, 2), "rel_l2_vs_fp64_oracle": err}


def run_extensions(args):
    """The three extension workloads, measured after the default line in the
    same run: the precision ratio on the block-clustered matrix (and on the
    default line's matrix), fwd+bwd / fwd through torch.autograd on the
    512K^2 matrix, and the 256-matrix batched graph replay.  Each runs as a
    child process (`bench.py --workload W` with this run's steps and
    warm-up, at least 200 and 20: a workload's launches take at most a few
    milliseconds each), one after the other, so that a workload that fails, even by
    a crash, cannot take the default line with it; each entry keeps the
    workload's metric, value and results, or its exit status and the end of
    its error output."""
    import subprocess

    keep = ("metric", "value", "unit", "ms_per_step", "dtype", "config", "results")
    ext = {}
    for name in ("precision", "autograd", "batched"):
        t0 = time.perf_counter()
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", name, "--steps", str(max(args.steps, 200)),
               "--warmup", str(max(args.warmup, 20))]
        try:
            proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                                  timeout=EXTENSION_TIMEOUT_S)
        except subprocess.TimeoutExpired as e:   # reported in the line, never hidden
            err = e.stderr.decode(errors="replace") if isinstance(e.stderr, bytes) else (e.stderr or "")
            ext[name] = {"error": f"no result within {EXTENSION_TIMEOUT_S} s", "stderr_tail": err[-2000:]}
        else:
            sys.stderr.write(proc.stderr)
            lines = [ln for ln in proc.stdout.splitlines() if ln.startswith("{")]
            if proc.returncode == 0 and lines:
                r = json.loads(lines[-1])
                ext[name] = {k: r[k] for k in keep if k in r}
            else:   # reported in the line, never hidden
                ext[name] = {"error": f"exit status {proc.returncode}", "stderr_tail": proc.stderr[-2000:]}
        ext[name]["seconds"] = round(time.perf_counter() - t0, 1)
        log(f"extension {name}: {ext[name].get('value')} in {ext[name]['seconds']} s")
    return ext


def read_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload_key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--matrix", default="stencil_powerlaw_1m", choices=MATRICES)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32", "bf16"])
    ap.add_argument("--split", default="auto", choices=["auto", "off", "on"],
                    help="the plan's scattered-entry split (auto: the library's choice)")
    ap.add_argument("--compensated", action="store_true")
    ap.add_argument("--block-size", type=int, default=16)
    ap.add_argument("--schedule", default="auto", choices=["auto", "stream", "blocked"])
    ap.add_argument("--item-nnz", type=int, default=0)
    ap.add_argument("--tile-waves", type=int, default=0, help="stream schedule: waves per tile (0 = auto)")
    ap.add_argument("--panels", type=int, default=0, help="stream schedule: x column panels 1, 2, 4 or 8 (0 = none)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE comparator")
    ap.add_argument("--no-extensions", action="store_true",
                    help="default line only: skip the precision, autograd and batched workloads it also reports")
    ap.add_argument("--workload", default="spmv", choices=["spmv", "precision", "autograd", "batched"])
    ap.add_argument("--mtx", default=None,
                    help="run the reference driver's protocol on a Matrix Market file (main.hip.cpp:46-75)")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (for scaling PMC byte counters)")
    ap.add_argument("--decompose", action="store_true",
                    help="time sub-matrices of stencil_powerlaw_1m (stencil half, near and scattered power-law entries) "
                         "each with the same kernel, at --dtype")
    ap.add_argument("--calibrate-gather", action="store_true",
                    help="only run 32M random 8-byte gathers from an 8 MiB FP64 vector (index_select), the "
                         "request shape of the scattered x reads, for reading PMC counters per gather")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU runs are out of scope", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        nbytes = 256 << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 5}), flush=True)
        return

    if args.calibrate_gather:
        n, k = 1 << 20, 1 << 25
        g = torch.Generator(device="cpu").manual_seed(0)
        src = torch.rand(n, dtype=torch.float64, generator=g).to("cuda:0")
        idx = torch.randint(0, n, (k,), generator=g, dtype=torch.int64).to("cuda:0")
        out = torch.empty(k, dtype=torch.float64, device="cuda:0")
        for _ in range(5):
            torch.index_select(src, 0, idx, out=out)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_gather_count": k, "x_bytes": n * 8, "index_bytes": k * 8, "out_bytes": k * 8,
                          "launches": 5}), flush=True)
        return

    if args.decompose:
        print(json.dumps(bench_decompose(args)), flush=True)
        return
    if args.mtx:
        print(json.dumps(bench_mtx(args)), flush=True)
        return
    if args.workload != "spmv":
        fn = {"precision": bench_precision, "autograd": bench_autograd, "batched": bench_batched}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    import cbspmv
    from oracle import pyoracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    d = load_workload(args.matrix)
    m, n, nnz = d["m"], d["n"], int(d["row"].size)
    log(f"{args.matrix}: {m}x{n}, nnz {nnz}, generated in {time.perf_counter() - t0:.1f}s")
    plan = cbspmv.Plan(m, n, d["row"], d["col"], d["val"], block_size=args.block_size, dtype=args.dtype,
                       compensated=args.compensated, schedule=args.schedule, item_nnz=args.item_nnz,
                       tile_waves=args.tile_waves, panels=args.panels, split=args.split)
    st = plan.stats()
    log(f"plan built in {st['build_seconds']:.2f}s: {st['gb_nnzb']} blocks (COO {st['coo_blocks']}, CSR "
        f"{st['csr_blocks']}, DENSE {st['dense_blocks']}), gather {st['gather_flag']}, schedule {plan.schedule()}")
    vdt = plan.vector_dtype
    x = torch.tensor(d["x"], dtype=vdt, device=dev)
    y = torch.empty(m, dtype=vdt, device=dev)
    stream = torch.cuda.current_stream()
    plan.upload(stream)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        plan.spmv(x, y, stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        plan.spmv(x, y, stream)
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms = total_ms / args.steps
    log(f"timed {args.steps} SpMVs: {ms * 1e3:.2f} us each")

    y_host = y.double().cpu().numpy()
    no_cpu = args.no_cpu
    if no_cpu:
        y_ref = pyoracle.spmv_seq(m, d["row"], d["col"], d["val"], d["x"])
        cpu = None
    else:
        y_ref, t_cpu, reps, kind, what = cpu_baseline(d, args.cpu_budget)
    rel = pyoracle.rel_l2(y_host, y_ref)
    log(f"rel_l2 vs FP64 oracle: {rel:.3e}")

    b_alg = algorithmic_bytes(m, n, nnz, args.dtype)
    gbs = b_alg / (ms * 1e-3) / 1e9
    gflops = 2.0 * nnz / (ms * 1e-3) / 1e9
    workload = (f"{args.matrix}_{args.dtype}{'_comp' if args.compensated else ''}_bs{args.block_size}"
                f"{'' if args.split == 'auto' else '_split_' + args.split}")
    if not no_cpu:
        cpu_gbs = b_alg / t_cpu / 1e9
        host = pyoracle.host_threads()
        # the host cores that go with one GPU: at most 16, and no more than this process may use
        threads = max(1, min(16, host["affinity_cpus"] if host["affinity_cpus"] > 0 else host["omp_max_threads"]))
        t_omp = cpu_parallel(d, threads, reps=20)
        cpu = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": 1, "kind": kind,
               "sample": f"full workload ({nnz} nnz), median of {reps} runs of {what}; "
                         f"{t_cpu * 1e3:.1f} ms per SpMV, {2 * nnz / t_cpu / 1e9:.3f} GFLOP/s",
               "host_threads_available": os.cpu_count(),
               "omp_max_threads": host["omp_max_threads"],
               "hardware_concurrency": host["hardware_concurrency"],
               "affinity_cpus": host["affinity_cpus"],
               "openmp_csr": {"value": round(b_alg / t_omp / 1e9, 3), "unit": "GB/s", "cores": threads,
                              "kind": "port", "ms_per_spmv": round(t_omp * 1e3, 3),
                              "sample": "oracle row-parallel CSR SpMV (same per-row order), CSR built once, "
                                        "OpenMP loop alone timed in C++, median of 20 after 1 warm-up"}}
    out = {
        "metric": "SpMV GB/s (CSR-equivalent algorithmic bytes) on one MI355X",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded generator, cbspmv/gen.py)",
        "config": {"workload": workload, "m": m, "n": n, "nnz": nnz, "block_size": args.block_size,
                   "schedule": plan.schedule(), "blocks": st["gb_nnzb"], "gather": bool(st["gather_flag"]),
                   "split_entries": st["split_entries"], "build_seconds": round(st["build_seconds"], 3)},
        "gflops": round(gflops, 2),
        "pct_hbm_peak": round(100.0 * gbs / HBM_PEAK_GBS, 2),
        "rel_l2_vs_fp64_oracle": rel,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": read_traffic(workload),
                     "kernels": ("stream_spmv (the entries the stream serves) then sweep_spmv (the "
                                 f"{st['split_entries']} scattered entries, {st['split_blocks']} row blocks), "
                                 "timed together" if st["split_entries"] else "stream_spmv"),
                     "bytes_per_launch_algorithmic": b_alg,
                     "stream_slab_bytes": st["stream_bytes"], "canonical_slab_bytes": st["slab_bytes"]},
        "cpu_baseline": cpu,
        "vendor_baseline": None if args.no_vendor else vendor_baseline(d, args.dtype, 200, ms),
    }
    plan.close()
    del x, y
    if not args.no_extensions and args.matrix == "stencil_powerlaw_1m" and args.dtype == "f64":
        torch.cuda.empty_cache()
        out["extensions"] = run_extensions(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
