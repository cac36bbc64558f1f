"""CB-SpMV benchmark on one MI355X: prints one JSON line.

A step is one SpMV y = A x over the resident plan (the reference protocol,
cb-spmv-dcu/src/cb-spmv-dcu.hpp:1631-1660: W untimed warm-up launches, then K
timed launches between events).  Default workload: config 2 (1,048,576^2,
~30.2M nnz: 9-point stencil half + power-law half, FP64, 16x16 blocks), the
largest single-GPU configuration the metric is quoted on.

value      = algorithmic (CSR-equivalent) bytes per SpMV / time per SpMV, GB/s:
             B_alg = nnz (s_v + 4) + 4 (m + 1) + s_x n + s_y m
roofline   = the same bytes over the average launch duration (HIP events on
             the launch stream) against 8 TB/s; `traffic` is read from the
             committed rocprofv3 PMC summary for this workload when present
cpu_baseline = the reference's own CPU path (check_results_consistency, its
             sequential FP64 SpMV, compiled from the reference sources into
             oracle/_ref), timed on the host on the same matrix; the oracle's
             restatement of the same loop when oracle/_ref is absent
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(m, n, nnz, dtype):
    s_v = {"f64": 8, "f32": 4, "bf16": 2}[dtype]
    s_x = 8 if dtype == "f64" else 4
    return nnz * (s_v + 4) + 4 * (m + 1) + s_x * n + s_x * m


def load_workload(config):
    from cbspmv import gen

    if config == 1:
        return gen.config1()
    if config == 2:
        return gen.config2()
    if config == 3:
        return gen.config3()
    if config == 5:
        return gen.config5()
    raise SystemExit(f"unknown config {config}")


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
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32", "bf16"])
    ap.add_argument("--compensated", action="store_true")
    ap.add_argument("--block-size", type=int, default=16)
    ap.add_argument("--schedule", default="auto", choices=["auto", "stream", "blocked"])
    ap.add_argument("--item-nnz", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (for scaling PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
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

    import cbspmv
    from oracle import pyoracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    d = load_workload(args.config)
    m, n, nnz = d["m"], d["n"], int(d["row"].size)
    log(f"config {args.config}: {m}x{n}, nnz {nnz}, generated in {time.perf_counter() - t0:.1f}s")
    plan = cbspmv.Plan(m, n, d["row"], d["col"], d["val"], block_size=args.block_size, dtype=args.dtype,
                       compensated=args.compensated, schedule=args.schedule, item_nnz=args.item_nnz)
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
    workload = (f"config{args.config}_{args.dtype}{'_comp' if args.compensated else ''}_bs{args.block_size}")
    if not no_cpu:
        cpu_gbs = b_alg / t_cpu / 1e9
        cpu = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": 1, "kind": kind,
               "sample": f"full workload ({nnz} nnz), median of {reps} runs of {what}; "
                         f"{t_cpu * 1e3:.1f} ms per SpMV, {2 * nnz / t_cpu / 1e9:.3f} GFLOP/s",
               "host_threads_available": os.cpu_count()}
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
                   "schedule": plan.schedule(), "blocks": st["gb_nnzb"], "gather": bool(st["gather_flag"])},
        "gflops": round(gflops, 2),
        "pct_hbm_peak": round(100.0 * gbs / HBM_PEAK_GBS, 2),
        "rel_l2_vs_fp64_oracle": rel,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": read_traffic(workload),
                     "bytes_per_launch_algorithmic": b_alg,
                     "stream_slab_bytes": st["stream_bytes"], "canonical_slab_bytes": st["slab_bytes"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
