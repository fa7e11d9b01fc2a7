#!/usr/bin/env python3
"""Benchmark of the MI355X SpGEMM engine (one JSON line on stdout).

Default workload: one rank's share of the large uniform product --
A and B 1,048,576 x 1,048,576 with 16 uniform random columns per row, fp64
values U[-1,1) (seeds 42 and 43); C = A.B through the plain drop-in path with
operands already resident in HBM.  A step is one complete product (analyze,
symbolic, scan, numeric; output left in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]

value     = 2*P / step time (GFLOP/s), P = scalar products (the reference's flop basis).
roofline  = HBM-bound: algorithmic bytes of the dominant kernel (the single-pass
            kernel k_fused: 32 B per row + 20 B per A entry + 12 B per scalar
            product + 12 B per C entry) / its average duration, measured with HIP
            events on the engine's stream around that launch.
cpu_baseline = the reference's own spgemm() (oracle/_ref/refdrv, -np 1, OpenMP
            guided/64) on a bounded row sample, or the C oracle when that binary
            is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "spgemm-hpc-thesis_amd", "py"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (rows, cols, nnz per row, seed A, seed B)
    "uniform_2p20x16": (1 << 20, 1 << 20, 16, 42, 43),
    "uniform_16384x8": (16384, 16384, 8, 42, 43),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_operands(name):
    import workloads as W
    rows, cols, k, sa, sb = WORKLOADS[name]
    a = W.uniform_random(rows, cols, k, seed=sa)
    b = W.uniform_random(cols, cols, k, seed=sb)
    return a, b


def traffic_from_profiles(workload):
    """HBM bytes per numeric launch measured with rocprofv3 PMC counters
    (profiles/traffic_<workload>.json, written from the FETCH_SIZE/WRITE_SIZE runs)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(a, b, products_total, target_s=15.0):
    """The reference's spgemm() on the host cores over a bounded row sample."""
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    rows_total = a.rows
    # sample: leading rows of A with ~(target) seconds of work at ~50 Mproducts/s/thread
    per_row = products_total / max(rows_total, 1)
    budget_products = 25e6 * threads * target_s / 8.0
    sample_rows = int(max(1024, min(rows_total, budget_products / max(per_row, 1))))
    sub_p = int(a.row_ptr[sample_rows])
    asub = O.Csr.make(sample_rows, a.cols, a.row_ptr[: sample_rows + 1], a.col[:sub_p], a.val[:sub_p])
    bo = O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val)
    sample_products = O.scalar_products(asub, bo)
    if O.ref_available():
        with tempfile.TemporaryDirectory() as tmp:
            pa, pb = os.path.join(tmp, "a.csr"), os.path.join(tmp, "b.csr")
            O.write_csr(pa, asub)
            O.write_csr(pb, bo)
            out = O.ref_run("bench", pa, pb, threads, 1, 5, 1, timeout=600)
        r = json.loads(out.strip().splitlines()[-1])
        p90 = r["compute_p90_s"]
        kind = "reference"
    else:
        times = []
        O.product(asub, bo, threads=threads)
        for _ in range(5):
            t0 = time.perf_counter()
            O.product(asub, bo, threads=threads)
            times.append(time.perf_counter() - t0)
        p90 = O.percentile90(times)
        kind = "port"
    return {"value": round(2.0 * sample_products / p90 / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
            "kind": kind,
            "sample": f"first {sample_rows} of {rows_total} rows of A times full B ({sample_products} scalar "
                      f"products), warmup 1, 5 timed repeats, compute P90; OpenMP guided/64"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uniform_2p20x16", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this engine computes one rank's local product; "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import spgemm_amd as S
    import workloads as W

    log(f"bench: generating {args.workload}")
    a, b = make_operands(args.workload)
    P = W.products(a, b)
    eng = S.Engine(0, timing=True)
    da, db = eng.upload(a), eng.upload(b)  # resident in HBM before timing
    eng.synchronize()

    log("bench: warmup")
    nnz = 0
    for _ in range(args.warmup):
        nnz = eng.multiply_device(da, db)
    eng.synchronize()

    log(f"bench: timing {args.steps} steps")
    step_ms, numeric_ms, symbolic_ms = [], [], []
    eng.synchronize()  # one GPU, one process: no barrier beyond the device sync
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        nnz = eng.multiply_device(da, db)  # returns after the product is complete on the device
        step_ms.append((time.perf_counter() - ts) * 1e3)
        st = eng.stats()
        numeric_ms.append(st["ms_numeric"])
        symbolic_ms.append(st["ms_symbolic"])
    eng.synchronize()
    total_s = time.perf_counter() - t0
    ms_per_step = total_s * 1e3 / args.steps
    st = eng.stats()

    # parity spot check on sampled rows (bit-exact against the oracle)
    import oracle as O
    c = eng.result_host()
    rng = np.random.Generator(np.random.PCG64(11))
    rows = np.sort(rng.choice(a.rows, size=min(1024, a.rows), replace=False))
    rp = [0]
    cols, vals = [], []
    for r in rows:
        s, e = a.row_ptr[r], a.row_ptr[r + 1]
        cols.append(a.col[s:e]); vals.append(a.val[s:e]); rp.append(rp[-1] + int(e - s))
    sub = O.product(O.Csr.make(len(rows), a.cols, rp, np.concatenate(cols), np.concatenate(vals)),
                    O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val))
    exact = all(
        np.array_equal(c.col[c.row_ptr[r]:c.row_ptr[r + 1]], sub.col[sub.row_ptr[i]:sub.row_ptr[i + 1]]) and
        np.array_equal(c.val[c.row_ptr[r]:c.row_ptr[r + 1]].view(np.uint64),
                       sub.val[sub.row_ptr[i]:sub.row_ptr[i + 1]].view(np.uint64))
        for i, r in enumerate(rows))

    model_bytes = W.streaming_bytes(a, b.rows, P, nnz)
    numeric_avg = statistics.mean(numeric_ms)
    single = bool(st["single_pass"])
    if single:
        kernel = "k_fused<small B>"
        kernel_bytes = 32 * a.rows + 20 * a.nnz + 12 * P + 12 * nnz
    else:
        kernel = "k_numeric_esc<64,256>"
        kernel_bytes = 40 * a.rows + 20 * a.nnz + 12 * P + 12 * nnz
    achieved = kernel_bytes / (numeric_avg * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.workload)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        try:
            cpu = cpu_baseline(a, b, P)
        except Exception as e:  # reported, never fatal
            cpu = {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference", "sample": f"failed: {e}"}

    rows_, cols_, k, sa, sb = WORKLOADS[args.workload]
    out = {
        "metric": "SpGEMM GFLOP/s (2 x scalar products / product time), single product, plain drop-in",
        "value": round(2.0 * P / (ms_per_step * 1e-3) / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform random CSR, generated on the host, resident in HBM before timing)",
        "config": {"workload": args.workload, "rows": rows_, "cols": cols_, "nnz_per_row": k,
                   "seeds": [sa, sb], "products": P, "c_nnz": nnz, "path": "plain drop-in (spg_multiply_begin)",
                   "p50_ms": round(statistics.median(step_ms), 4),
                   "p99_ms": round(float(np.percentile(step_ms, 99)), 4),
                   "pipeline_model_gbs": round(model_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                   "single_pass": single,
                   "analyze_symbolic_ms_avg": round(statistics.mean(symbolic_ms), 4),
                   "dominant_kernel_ms_avg": round(numeric_avg, 4),
                   "rows_per_bin": st["rows_per_bin"],
                   "parity": f"bit-exact on {len(rows)} sampled rows" if exact else "MISMATCH"},
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": kernel_bytes,
                     "pipeline_bytes_model": "streaming Gustavson: A once, B row bounds + entries per A entry, "
                                             "C once",
                     "pipeline_bytes": model_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    da.free()
    db.free()
    eng.close()


if __name__ == "__main__":
    main()
