#!/usr/bin/env python3
"""Benchmark of the MI355X SpGEMM engine (one JSON line on stdout).

Default run: the single product of one rank's share -- A and B 1,048,576 x
1,048,576 with 16 uniform random columns per row, fp64 values U[-1,1) (seeds
42 and 43); C = A.B through the plain drop-in path with operands already
resident in HBM -- as the headline `value`, followed by two sections of the
same line measured in the same run:
  config.batched  4096 independent jobs (rows log-uniform in [256, 65536], 4..32
                  uniform columns per row, every third a Galerkin triple R.A.P on
                  a 27-point stencil with 2x2x2 aggregation) as one batched
                  pipeline replayed from a HIP graph, against sequential plain
                  launches of the same jobs
  config.served   a Poisson request stream (2000/s for 3 s) of the same jobs'
                  products plus 5% products of the 2^20 x16 shape over registered,
                  device-resident operands, served with continuous batching under
                  a 200 GB budget, against one product per launch; p50/p99 latency

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-sections] [--no-cpu-baseline]

Timing follows the reference's protocol (src/baselines/two_sided_spgemm.cpp:
126-192): W untimed warmup products, then K timed products in trials of 10
repeats, each product timed on its own; the reported time is the nearest-rank
P90 of the samples (percentile90, src/common/spgemm_common.cpp:916-924), and
value = 2 P / P90 with P the scalar products (scalarMultiplicationCount).  The
mean (ms_per_step) and the median are reported beside it.  Defaults: warmup 2,
50 steps (5 trials x 10 repeats), the reference CLI's defaults.

Other workloads (--workload, printed as their own line): uniform_16384x8;
rmat_s20_AA (R-MAT scale 20, edge factor 32, (.57,.19,.19,.05), A.A, whose
~2.3e10-entry output exceeds device memory: a step is one sweep over row chunks);
rmat_s20_rect (the same R-MAT times B 2^20 x 65536, 24 per row); batched_4096;
served_poisson.

roofline  = HBM-bound: algorithmic bytes of the dominant kernel / its average
            duration from HIP events on the engine's stream.  For the single-pass
            kernel k_fused: 32 B per row + 16 B per A entry (B row start and
            length, 4 B each when B has fewer than 2^29 entries, else 12 B; A
            value 8 B) + 12 B per scalar product + 12 B per C entry.  `traffic`
            is the calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per launch
            recorded in profiles/traffic_<workload>.md.
cpu_baseline = the reference's own spgemm() (oracle/_ref/refdrv, -np 1, OpenMP
            guided/64, every host core) under the same protocol on a bounded row
            sample; the C oracle when that binary is absent.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "spgemm-hpc-thesis_amd", "py"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
REPEATS = 10          # repeats per trial (the reference CLI's default)
CPU_WARMUP, CPU_REPEATS, CPU_TRIALS = 2, 10, 5
CHUNK_PRODUCTS = 3 << 30  # scalar products per row chunk of a sweep (bounds C per chunk)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _uniform(rows, cols, k, seed):
    import workloads as W
    return W.uniform_random(rows, cols, k, seed=seed)


def _rmat():
    import workloads as W
    return W.rmat(20, 32, seed=42)


WORKLOADS = {
    "uniform_2p20x16": dict(kind="single", desc={"rows": 1 << 20, "cols": 1 << 20, "nnz_per_row": 16, "seeds": [42, 43]},
                            make=lambda: (_uniform(1 << 20, 1 << 20, 16, 42), _uniform(1 << 20, 1 << 20, 16, 43))),
    "uniform_16384x8": dict(kind="single", desc={"rows": 16384, "cols": 16384, "nnz_per_row": 8, "seeds": [42, 43]},
                            make=lambda: (_uniform(16384, 16384, 8, 42), _uniform(16384, 16384, 8, 43))),
    "rmat_s20_AA": dict(kind="chunked", desc={"rmat_scale": 20, "edge_factor": 32, "abcd": [0.57, 0.19, 0.19, 0.05],
                                              "seed": 42, "product": "A.A"},
                        make=lambda: (lambda a: (a, a))(_rmat())),
    "rmat_s20_rect": dict(kind="single", desc={"rmat_scale": 20, "edge_factor": 32, "seed": 42,
                                               "B": "2^20 x 65536, 24 uniform columns per row, seed 43"},
                          make=lambda: (_rmat(), _uniform(1 << 20, 65536, 24, 43))),
    "batched_4096": dict(kind="batched", desc={"jobs": 4096, "seed": 42, "rows": "log-uniform [256, 65536]",
                                               "nnz_per_row": "uniform 4..32", "galerkin_fraction": "1/3 (n in 16..48)"}),
    "served_poisson": dict(kind="served", desc={"rate_per_s": 2000, "duration_s": 3.0, "big_fraction": 0.05,
                                                "big_shape": "2^20 x 2^20, 16/row", "pool_jobs": 384, "seed": 42,
                                                "budget_bytes": 200 << 30}),
}


# ---- timing protocol -------------------------------------------------------

def p90_nearest_rank(samples):
    """Nearest-rank 90th percentile: the ceil(0.9 n)-th smallest sample
    (percentile90, src/common/spgemm_common.cpp:916-924)."""
    s = sorted(samples)
    return s[max(1, math.ceil(0.90 * len(s))) - 1]


def run_protocol(step, warmup, steps):
    """warmup untimed steps, then `steps` steps each timed on its own, in
    trials of REPEATS (two_sided_spgemm.cpp:126-167).  step() returns after
    its work is complete on the device; returns the samples in seconds."""
    for _ in range(warmup):
        step()
    samples = []
    for _ in range(steps):
        t0 = time.perf_counter()
        step()
        samples.append(time.perf_counter() - t0)
    return samples


def timing_fields(samples):
    n = len(samples)
    return {"samples": n, "trials": math.ceil(n / REPEATS), "repeats": min(REPEATS, n),
            "p90_ms": round(p90_nearest_rank(samples) * 1e3, 4),
            "p50_ms": round(statistics.median(samples) * 1e3, 4),
            "mean_ms": round(statistics.mean(samples) * 1e3, 4)}


def host_cores():
    """(logical CPUs this process may run on, physical cores of the host)."""
    try:
        logical = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        logical = os.cpu_count() or 1
    cores = set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    phys = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif not line.strip() and core is not None:
                    cores.add((phys, core))
                    phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return logical, (len(cores) or logical)


def row_structure(row_ptr):
    """Row-count statistics of a CSR (population min/max/mean/stddev/cv and
    empty rows: appendRowMetrics, src/common/benchmark_metrics.cpp:21-47)."""
    n = np.diff(np.asarray(row_ptr, dtype=np.int64))
    if len(n) == 0:
        return {"row_nnz_min": 0, "row_nnz_max": 0, "mean_nnz_per_row": 0.0, "row_nnz_stddev": 0.0,
                "row_nnz_cv": 0.0, "empty_rows": 0}
    mean = float(n.mean())
    sd = float(n.std())
    return {"row_nnz_min": int(n.min()), "row_nnz_max": int(n.max()), "mean_nnz_per_row": round(mean, 6),
            "row_nnz_stddev": round(sd, 6), "row_nnz_cv": round(sd / mean, 6) if mean > 0 else 0.0,
            "empty_rows": int((n == 0).sum())}


def traffic_from_profiles(workload):
    """HBM bytes per launch of the dominant kernel, from the calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/traffic_<workload>.md (its line `hbm_bytes_per_launch: N`)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return int(line.split(":", 1)[1].strip().replace(",", ""))
    return None


# ---- CPU baselines (the reference's own code where it is built) -----------

def cpu_baseline(a, b, products_total, target_s=20.0):
    """The reference's spgemm() on every host core over a bounded row sample,
    warmup 2, 5 trials x 10 repeats, compute P90."""
    import oracle as O
    threads, physical = host_cores()
    runs = 1 + CPU_WARMUP + CPU_REPEATS * CPU_TRIALS
    # leading rows of A holding about target_s / runs seconds of the
    # reference's work per run (~8e6 products per second per thread), by the
    # rows' cumulative products (skewed matrices put their heaviest rows first)
    budget_products = 8e6 * threads * target_s / runs
    cum = np.concatenate([[0], np.cumsum(np.diff(b.row_ptr)[a.col], dtype=np.int64)])[a.row_ptr]
    sample_rows = int(min(a.rows, max(1, np.searchsorted(cum, budget_products, side="left"))))
    sub_p = int(a.row_ptr[sample_rows])
    asub = O.Csr.make(sample_rows, a.cols, a.row_ptr[: sample_rows + 1], a.col[:sub_p], a.val[:sub_p])
    bo = O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val)
    sample_products = O.scalar_products(asub, bo)
    if O.ref_available():
        with tempfile.TemporaryDirectory() as tmp:
            pa, pb = os.path.join(tmp, "a.csr"), os.path.join(tmp, "b.csr")
            O.write_csr(pa, asub)
            O.write_csr(pb, bo)
            out = O.ref_run("bench", pa, pb, threads, CPU_WARMUP, CPU_REPEATS, CPU_TRIALS, timeout=900)
        r = json.loads(out.strip().splitlines()[-1])
        p90, p50 = r["compute_p90_s"], r["compute_p50_s"]
        kind = "reference"
    else:
        O.product(asub, bo, threads=threads)
        samples = run_protocol(lambda: O.product(asub, bo, threads=threads), CPU_WARMUP, CPU_REPEATS * CPU_TRIALS)
        p90, p50 = p90_nearest_rank(samples), statistics.median(samples)
        kind = "port"
    return {"value": round(2.0 * sample_products / p90 / 1e9, 4), "unit": "GFLOP/s", "cores": threads,
            "kind": kind, "host_logical_cpus": threads, "host_physical_cores": physical,
            "p90_ms": round(p90 * 1e3, 3), "p50_ms": round(p50 * 1e3, 3),
            "sample": f"first {sample_rows} of {a.rows} rows of A times full B ({sample_products} scalar products); "
                      f"warmup {CPU_WARMUP}, {CPU_TRIALS} trials x {CPU_REPEATS} repeats, compute P90; "
                      f"OpenMP guided/64 on {threads} threads"}


def cpu_baseline_jobs(jobs, njobs):
    """The reference's spgemm() over the first njobs jobs in sequence (a
    Galerkin triple as (R.A).P), one sample per pass, warmup 2, 5 trials x 10
    repeats, P90; jobs/s = njobs / P90."""
    import oracle as O
    threads, physical = host_cores()
    if not O.ref_available():
        return None
    with tempfile.TemporaryDirectory() as tmp:
        lines = []
        for i in range(njobs):
            j = jobs[i]
            paths = []
            for k, m in enumerate(j[1:]):
                pth = os.path.join(tmp, f"j{i}_{k}.csr")
                O.write_csr(pth, _o(m))
                paths.append(pth)
            lines.append(" ".join([j[0]] + paths))
        man = os.path.join(tmp, "jobs.txt")
        with open(man, "w") as f:
            f.write("\n".join(lines) + "\n")
        out = O.ref_run("benchjobs", man, threads, CPU_WARMUP, CPU_REPEATS, CPU_TRIALS, timeout=900)
    r = json.loads(out.strip().splitlines()[-1])
    return {"value": round(r["jobs_per_s_p90"], 3), "unit": "jobs/s", "cores": threads, "kind": "reference",
            "host_logical_cpus": threads, "host_physical_cores": physical,
            "gflops": round(r["gflops_p90"], 4), "p90_ms": round(r["p90_s"] * 1e3, 3),
            "sample": f"first {njobs} jobs in sequence per sample ({r['products']} scalar products), "
                      f"warmup {CPU_WARMUP}, {CPU_TRIALS} trials x {CPU_REPEATS} repeats, P90; "
                      f"OpenMP guided/64 on {threads} threads"}


# ---- parity helpers (the serial-order oracle on sampled rows) -------------

def oracle_rows(a, b, rows):
    """Serial-order oracle on selected rows of A (bit-exact reference values)."""
    import oracle as O
    rp = [0]
    cols, vals = [], []
    for r in rows:
        s, e = a.row_ptr[r], a.row_ptr[r + 1]
        cols.append(a.col[s:e])
        vals.append(a.val[s:e])
        rp.append(rp[-1] + int(e - s))
    return O.product(O.Csr.make(len(rows), a.cols, rp, np.concatenate(cols), np.concatenate(vals)),
                     O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val), threads=host_cores()[0])


def rows_equal(c, r_c, ref, i):
    return (np.array_equal(c.col[c.row_ptr[r_c]:c.row_ptr[r_c + 1]], ref.col[ref.row_ptr[i]:ref.row_ptr[i + 1]]) and
            np.array_equal(c.val[c.row_ptr[r_c]:c.row_ptr[r_c + 1]].view(np.uint64),
                           ref.val[ref.row_ptr[i]:ref.row_ptr[i + 1]].view(np.uint64)))


class RelErrors:
    """Max-entry and Frobenius relative error of sampled rows against the
    serial-order oracle (the fields the reference's Full report prints,
    src/common/spgemm_common.cpp:590-851); defined where the column arrays
    agree, else the rows are a mismatch already."""

    def __init__(self):
        self.max_rel, self.diff2, self.ref2, self.ok, self.finite = 0.0, 0.0, 0.0, True, True

    def add(self, c, r_c, ref, i):
        cc = c.col[c.row_ptr[r_c]:c.row_ptr[r_c + 1]]
        rc = ref.col[ref.row_ptr[i]:ref.row_ptr[i + 1]]
        if not np.array_equal(cc, rc):
            self.ok = False
            return
        cv = c.val[c.row_ptr[r_c]:c.row_ptr[r_c + 1]]
        rv = ref.val[ref.row_ptr[i]:ref.row_ptr[i + 1]]
        self.finite &= bool(np.all(np.isfinite(cv)))
        d = np.abs(cv - rv)
        nz = rv != 0
        if np.any(nz):
            self.max_rel = max(self.max_rel, float(np.max(d[nz] / np.abs(rv[nz]))))
        self.diff2 += float(np.sum(d * d))
        self.ref2 += float(np.sum(rv * rv))

    def report(self, rows):
        if not self.ok:
            return {"rows": rows, "columns_match": False}
        fro = (self.diff2 ** 0.5) / (self.ref2 ** 0.5) if self.ref2 > 0 else 0.0
        return {"rows": rows, "columns_match": True, "max_entry_rel_error": self.max_rel,
                "frobenius_rel_error": fro, "values_finite": self.finite}


def chunk_bounds(rowp, budget):
    """Row chunks [r0, r1) whose scalar products stay within budget (a row
    above the budget gets a chunk of its own)."""
    cs = np.concatenate([[0], np.cumsum(rowp)])
    bounds, r0, n = [], 0, len(rowp)
    while r0 < n:
        r1 = int(np.searchsorted(cs, cs[r0] + budget, side="right")) - 1
        r1 = min(max(r1, r0 + 1), n)
        bounds.append((r0, r1))
        r0 = r1
    return bounds


def _o(m):
    import oracle as O
    return O.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val)


def streaming_bytes_counts(rows, a_nnz, p, c_nnz):
    """Streaming-Gustavson bytes of one product: A once, B row bounds and
    entries once per referencing A entry, C once."""
    return (8 * (rows + 1) + 12 * a_nnz) + (16 * a_nnz + 12 * p) + (8 * (rows + 1) + 12 * c_nnz)


# ---- the single product (headline) ------------------------------------------

def run_single(args, name, wl):
    import spgemm_amd as S
    import workloads as W

    chunked = wl["kind"] == "chunked"
    steps = args.steps if args.steps is not None else (1 if chunked else CPU_REPEATS * CPU_TRIALS)
    warmup = args.warmup if args.warmup is not None else (0 if chunked else CPU_WARMUP)
    log(f"bench: generating {name}")
    a, b = wl["make"]()
    if args.row_range:
        # rows [r0, r1) of A only: that block of C, as one product
        r0, r1 = (int(x) for x in args.row_range.split(":"))
        s0, s1 = int(a.row_ptr[r0]), int(a.row_ptr[r1])
        a = S.Csr.make(r1 - r0, a.cols, a.row_ptr[r0:r1 + 1] - s0, a.col[s0:s1], a.val[s0:s1])
        chunked = False
        name = f"{name}[rows {r0}:{r1}]"
    rowp = W.row_products(a, b)
    P = int(rowp.sum())
    eng = S.Engine(0, timing=True)
    da, db = eng.upload(a), eng.upload(b)  # resident in HBM before timing
    eng.synchronize()
    bounds = chunk_bounds(rowp, CHUNK_PRODUCTS) if chunked else [(0, a.rows)]
    log(f"bench: P={P}, {len(bounds)} row chunk(s)")
    kernel_ms, sym_ms, nnz_box = [], [], [0]

    def one_step():
        nnz_total = 0
        kms = 0.0
        for r0, r1 in bounds:
            nnz_total += eng.multiply_device(da, db, (r0, r1) if chunked else None)
            kms += eng.stats()["ms_numeric"]
            if chunked:
                log(f"bench: rows [{r0}, {r1}) done")
        kernel_ms.append(kms)
        sym_ms.append(eng.stats()["ms_symbolic"])
        nnz_box[0] = nnz_total

    log(f"bench: warmup {warmup}, timing {steps} steps")
    eng.synchronize()  # one GPU, one process: no barrier beyond the device sync
    samples = run_protocol(one_step, warmup, steps)
    eng.synchronize()
    kernel_ms, sym_ms = kernel_ms[warmup:], sym_ms[warmup:]
    nnz = nnz_box[0]
    st = eng.stats()
    tf = timing_fields(samples)
    structure = None
    if not chunked:
        structure = {"a": row_structure(a.row_ptr), "b": row_structure(b.row_ptr),
                     "c": row_structure(eng.result_row_ptr()), "csr_offset_bits": 64}

    # parity spot check: sampled rows (and the heaviest rows of a sweep)
    # bit-exact against the serial-order oracle
    rng = np.random.Generator(np.random.PCG64(11))
    sample = set(rng.choice(a.rows, size=min(512 if not chunked else 128, a.rows), replace=False).tolist())
    if chunked:
        sample.update(int(h) for h in np.argsort(rowp)[-2:])
    sample = np.array(sorted(sample))
    log(f"bench: parity check on {len(sample)} rows")
    ref = oracle_rows(a, b, sample)
    exact = True
    rel = RelErrors()
    pos = {int(r): i for i, r in enumerate(sample)}
    for r0, r1 in bounds:
        mine = [int(r) for r in sample if r0 <= r < r1]
        if not mine:
            continue
        if chunked:
            eng.multiply_device(da, db, (r0, r1))
        for r in mine:
            c = eng.result_rows(r - r0, r - r0 + 1)
            exact &= rows_equal(c, 0, ref, pos[r])
            rel.add(c, 0, ref, pos[r])

    single = bool(st["single_pass"])
    entry_bytes = (8 if not st["wide_format"] else 12) + 8  # B row start + length, A value
    big_rows = int(sum(st["rows_per_bin"][3:]))  # bins W3, W4, Block, Large run after the single-pass kernel
    if st["direct"]:
        # direct single pass (no analysis): per row A's and C's row offsets,
        # per A entry its column, value and B's two 32-bit row offsets
        kernel = f"k_fused<{'false' if st['wide_format'] else 'true'}, 2, true>"
        kernel_bytes = 16 * a.rows + 20 * a.nnz + 12 * P + 12 * nnz
    elif single and not chunked and big_rows == 0:
        # the timed values phase is the single-pass kernel alone (named as rocprofv3 names it)
        kernel = f"k_fused<{'false' if st['wide_format'] else 'true'}, 2>"
        kernel_bytes = 32 * a.rows + entry_bytes * a.nnz + 12 * P + 12 * nnz
    else:
        kernel = "numeric phase (single-pass and big-row kernels)" if single else "numeric phase"
        kernel_bytes = 40 * a.rows + entry_bytes * a.nnz + 12 * P + 12 * nnz
    kernel_avg = statistics.mean(kernel_ms)
    achieved = kernel_bytes / (kernel_avg * 1e-3) / 1e9
    model_bytes = W.streaming_bytes(a, b.rows, P, nnz)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        try:
            cpu = cpu_baseline(a, b, P)
        except Exception as e:  # reported, never fatal
            cpu = {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "reference", "sample": f"failed: {e}"}

    p90_s = tf["p90_ms"] * 1e-3
    cfg = {"workload": name, **wl["desc"], "products": P, "c_nnz": nnz, "a_nnz": a.nnz,
           "path": "plain drop-in (spg_multiply_begin)" if not chunked else
                   f"row chunks of <= {CHUNK_PRODUCTS} products (spg_multiply_rows), {len(bounds)} chunks",
           "timing": {"protocol": "reference: warmup, trials x repeats, each product timed, nearest-rank P90 "
                                  "(value = 2P / P90)", **tf},
           "gflops_mean": round(2.0 * P / (tf["mean_ms"] * 1e-3) / 1e9, 3),
           "pipeline_model_gbs_p90": round(model_bytes / p90_s / 1e9, 1),
           "single_pass": single,
           "direct_single_pass": bool(st["direct"]),
           "wide_format": bool(st["wide_format"]),
           "analyze_symbolic_ms_mean": round(statistics.mean(sym_ms), 4),
           "dominant_kernel_ms_mean": round(kernel_avg, 4),
           "rows_per_bin": st["rows_per_bin"],
           "max_row_products": int(rowp.max()) if len(rowp) else 0,
           "parity": f"bit-exact on {len(sample)} sampled rows" if exact else "MISMATCH",
           "full_mode_errors": rel.report(len(sample))}
    if structure:
        cfg["row_structure"] = structure
    out = {
        "metric": "SpGEMM GFLOP/s (2 x scalar products / P90 product time), single product, plain drop-in",
        "value": round(2.0 * P / p90_s / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": tf["mean_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded CSR generated on the host, resident in HBM before timing)",
        "config": cfg,
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(name),
                     "bytes_per_launch": kernel_bytes,
                     "pipeline_bytes_model": "streaming Gustavson: A once, B row bounds + entries per A entry, "
                                             "C once",
                     "pipeline_bytes": model_bytes},
        "cpu_baseline": cpu,
    }
    da.free()
    db.free()
    eng.close()
    return out


# ---- batched jobs -------------------------------------------------------------

def run_batched(args, wl, steps=10, warmup=2, cpu=True):
    """Batched jobs/s against sequential plain launches (same resident operands,
    identical bit-exact results)."""
    import oracle as O
    import spgemm_amd as S
    import workloads as W

    njobs = args.jobs if args.jobs else wl["desc"]["jobs"]
    procs = max(1, min(16, os.cpu_count() or 1))
    log(f"bench: generating {njobs} jobs on {procs} processes")
    jobs = W.batch_jobs(njobs, seed=wl["desc"]["seed"], processes=procs)
    ab = [i for i, j in enumerate(jobs) if j[0] == "ab"]
    rap = [i for i, j in enumerate(jobs) if j[0] == "rap"]
    a1 = [jobs[i][1] for i in ab] + [jobs[i][1] for i in rap]
    b1 = [jobs[i][2] for i in ab] + [jobs[i][2] for i in rap]
    p_list = [jobs[i][3] for i in rap]
    assert all(np.all(np.diff(p.row_ptr) == 1) for p in p_list)  # aggregation: one entry per row
    eng = S.Engine(0, timing=True)
    log("bench: building batch arenas")
    w1 = S.Batch(eng, a1, b1)
    w1.run()
    w2 = S.Batch.chained(eng, w1, [len(ab) + k for k in range(len(rap))], p_list)
    w2.run()
    log("bench: autotune")
    c1 = w1.autotune(2)
    c2 = w2.autotune(2)
    w1.set_graph(not args.no_graph)
    w2.set_graph(not args.no_graph)
    gpu_ms, nnz_box = [], [0]

    def batch_step():
        n1 = w1.run()
        t1 = w1.stats()["ms_total"]
        w2.refresh()
        n2 = w2.run()
        gpu_ms.append(t1 + w2.stats()["ms_total"])
        nnz_box[0] = n1 + n2

    eng.synchronize()
    samples = run_protocol(batch_step, warmup, steps)
    gpu_ms = gpu_ms[warmup:]
    tf = timing_fields(samples)
    nnz = nnz_box[0]

    # products and algorithmic bytes of both waves (wave 2: (R.A).P with one
    # P entry per row, so its products equal nnz(R.A))
    n1 = [w1.product_nnz(i) for i in range(len(a1))]
    n2 = [w2.product_nnz(k) for k in range(len(rap))]
    p_wave1 = [W.products(a, b) for a, b in zip(a1, b1)]
    P_total = sum(p_wave1) + sum(n1[len(ab):])
    model = sum(streaming_bytes_counts(a.rows, a.nnz, p, c) for a, p, c in zip(a1, p_wave1, n1))
    model += sum(streaming_bytes_counts(jobs[i][1].rows, n1[len(ab) + k], n1[len(ab) + k], n2[k])
                 for k, i in enumerate(rap))
    gpu_mean = statistics.mean(gpu_ms) if gpu_ms else tf["mean_ms"]

    if args.batch_only:  # counter profiles: the batch's dispatches are each kernel's last
        for o in (w2, w1, eng):
            o.close()
        p90_s = tf["p90_ms"] * 1e-3
        return {"metric": "batched pipeline only (profiling run: no sequential leg, parity or CPU baseline)",
                "value": round(njobs / p90_s, 2), "unit": "jobs/s", "steps": steps, "warmup": warmup,
                "ms_per_step": tf["mean_ms"],
                "config": {"workload": "batched_4096", **wl["desc"], "jobs": njobs, "products": int(P_total),
                           "c_nnz": int(nnz), "gpu_ms_mean": round(gpu_mean, 3)},
                "roofline": None, "cpu_baseline": None}

    # sequential plain launches: one drop-in multiply per product on resident
    # operands; the triples' second product takes the first one's device result
    log("bench: sequential plain launches")
    seq_eng = S.Engine(0)
    ra_eng = S.Engine(0)
    dev = [(seq_eng.upload(jobs[i][1]), seq_eng.upload(jobs[i][2]), seq_eng.upload(jobs[i][3]) if jobs[i][0] == "rap"
            else None) for i in range(njobs)]
    seq_eng.synchronize()

    def seq_all():
        for i in range(njobs):
            da, db, dp = dev[i]
            if dp is None:
                seq_eng.multiply_device(da, db)
            else:
                ra_eng.multiply_device(da, db)
                seq_eng.begin(ra_eng.result_view(), dp.view())
    seq_samples = run_protocol(seq_all, 1, 2)
    seq_s = p90_nearest_rank(seq_samples)

    # bit-exact: sampled jobs from the batch equal their standalone products
    rng = np.random.Generator(np.random.PCG64(5))
    check = sorted(rng.choice(njobs, size=min(64, njobs), replace=False).tolist())
    exact = True
    pos1 = {j: k for k, j in enumerate(ab + rap)}
    pos2 = {j: k for k, j in enumerate(rap)}
    for i in check:
        da, db, dp = dev[i]
        if dp is None:
            seq_eng.multiply_device(da, db)
            exact &= O.csr_digest(_o(seq_eng.result_host())) == O.csr_digest(_o(w1.result(pos1[i])))
        else:
            ra_eng.multiply_device(da, db)
            seq_eng.begin(ra_eng.result_view(), dp.view())
            exact &= O.csr_digest(_o(seq_eng.result_host())) == O.csr_digest(_o(w2.result(pos2[i])))
    # and a few jobs against the serial-order oracle
    for i in check[:8]:
        kind = jobs[i][0]
        a, b = jobs[i][1], jobs[i][2]
        ref = O.product(_o(a), _o(b))
        if kind == "ab":
            exact &= O.csr_digest(ref) == O.csr_digest(_o(w1.result(pos1[i])))
        else:
            ref = O.product(ref, _o(jobs[i][3]))
            exact &= O.csr_digest(ref) == O.csr_digest(_o(w2.result(pos2[i])))
    wide = w1.stats()["wide_format"]
    for tup in dev:
        for m in tup:
            if m is not None:
                m.free()
    for o in (w2, w1, seq_eng, ra_eng, eng):
        o.close()

    cpu_rec = None
    if cpu and not args.no_cpu_baseline:
        log("bench: cpu baseline (reference spgemm over the first 24 jobs)")
        try:
            cpu_rec = cpu_baseline_jobs(jobs, min(24, njobs))
        except Exception as e:  # reported, never fatal
            cpu_rec = {"value": None, "unit": "jobs/s", "cores": 0, "kind": "reference", "sample": f"failed: {e}"}
    p90_s = tf["p90_ms"] * 1e-3
    achieved = model / (gpu_mean * 1e-3) / 1e9
    return {
        "metric": f"batched SpGEMM jobs/s ({njobs} independent jobs, 1/3 Galerkin R.A.P), one batched "
                  + ("pipeline launched eagerly" if args.no_graph else "pipeline replayed from a HIP graph")
                  + " (jobs / P90 batch time)",
        "value": round(njobs / p90_s, 2),
        "unit": "jobs/s",
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": tf["mean_ms"],
        "config": {"workload": "batched_4096", **wl["desc"], "jobs": njobs, "products": int(P_total),
                   "c_nnz": int(nnz), "timing": {"protocol": "reference (nearest-rank P90 of timed batches)", **tf},
                   "gpu_ms_mean": round(gpu_mean, 3),
                   "sequential_jobs_per_s": round(njobs / seq_s, 2),
                   "speedup_vs_sequential": round(seq_s / p90_s, 3),
                   "gflops": round(2.0 * P_total / p90_s / 1e9, 2),
                   "autotune_choice": {"wave1": c1, "wave2": c2}, "wide_format_wave1": bool(wide),
                   "parity": ("bit-exact: 64 sampled jobs equal their standalone products, 8 equal the oracle"
                              if exact else "MISMATCH")},
        "roofline": {"bound": "hbm", "kernel": "whole batched pipeline (both waves, HIP events around each graph)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(model),
                     "pipeline_bytes_model": "streaming Gustavson summed over every product of both waves"},
        "cpu_baseline": cpu_rec,
    }


# ---- served requests ---------------------------------------------------------

def run_served(args, wl):
    """Poisson stream of products over resident operands: continuous batching
    against one product per launch, same stream, same results."""
    import threading
    import oracle as O
    import spgemm_amd as S
    import workloads as W

    d = wl["desc"]
    procs = max(1, min(16, os.cpu_count() or 1))
    log("bench: generating the operand pool")
    jobs = W.batch_jobs(d["pool_jobs"], seed=d["seed"], processes=procs)
    big_a = W.uniform_random(1 << 20, 1 << 20, 16, seed=42)
    big_b = W.uniform_random(1 << 20, 1 << 20, 16, seed=43)
    eng = S.Engine(0)
    pairs = []  # (A, B) host operands of every pool product
    for j in jobs:
        if j[0] == "ab":
            pairs.append((j[1], j[2]))
        else:
            ra = eng.multiply(j[1], j[2])  # the R.A operand of the second product of a triple
            pairs.append((j[1], j[2]))
            pairs.append((ra, j[3]))
    rng = np.random.Generator(np.random.PCG64(d["seed"]))
    n_req = int(d["rate_per_s"] * d["duration_s"])
    arrivals = np.cumsum(rng.exponential(1.0 / d["rate_per_s"], size=n_req))
    is_big = rng.random(n_req) < d["big_fraction"]
    pick = rng.integers(0, len(pairs), size=n_req)

    def serve(max_batch, saturated=False):
        """The request stream against a server; saturated: every request
        submitted at once (throughput bound, admission under pressure)."""
        srv = S.Server(0, budget_bytes=d["budget_bytes"], max_batch=max_batch)
        handles = []
        for a, b in pairs:
            handles.append((srv.register(a), srv.register(b)))
        big_h = (srv.register(big_a), srv.register(big_b))
        if saturated:
            # one untimed pass: the server's pooled batches grow to their
            # steady sizes (growing a batch that is already tens of GB
            # reallocates it, which stalls the device for seconds once)
            warm = [srv.submit_resident(*(big_h if is_big[k] else handles[pick[k]])) for k in range(n_req)]
            for r in warm:
                srv.wait(r)
                srv.release(r)
        st0 = srv.stats()
        rids = [None] * n_req
        lat = np.zeros(n_req)
        nnzs = np.zeros(n_req, np.int64)
        done_t = np.zeros(n_req)
        ready = threading.Semaphore(0)

        def waiter(k0):
            for k in range(k0, n_req, 4):
                ready.acquire()
                while rids[k] is None:
                    time.sleep(0.0001)
                nnzs[k], lat[k] = srv.wait(rids[k])
                done_t[k] = time.perf_counter()
                srv.release(rids[k])

        ws = [threading.Thread(target=waiter, args=(k,)) for k in range(4)]
        for w in ws:
            w.start()
        t0 = time.perf_counter()
        for k in range(n_req):
            delay = t0 + arrivals[k] - time.perf_counter()
            if delay > 0 and not saturated:
                time.sleep(delay)
            h = big_h if is_big[k] else handles[pick[k]]
            rids[k] = srv.submit_resident(*h)
            ready.release()
        for w in ws:
            w.join()
        wall = float(done_t.max() - t0)
        st = srv.stats()
        # the same products alone (engine on the same GPU) for a sample of requests
        exact = True
        for k in rng.choice(n_req, size=0 if saturated else min(24, n_req), replace=False):
            a, b = (big_a, big_b) if is_big[k] else pairs[pick[k]]
            rid = srv.submit_resident(*(big_h if is_big[k] else handles[pick[k]]))
            got = srv.result(rid)
            srv.release(rid)
            alone = eng.multiply(a, b)
            exact &= O.csr_digest(_o(got)) == O.csr_digest(_o(alone))
        srv.close()
        return {"throughput": n_req / wall, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "wall_s": wall,
                "batches": st["batches"] - st0["batches"], "busy_s": st["busy_s"] - st0["busy_s"], "nnz": nnzs,
                "exact": exact}

    log("bench: served with continuous batching")
    cb = serve(4096)
    log("bench: served one product per launch")
    seq = serve(1)
    log("bench: saturated stream (every request at once), both modes")
    sat_cb = serve(4096, saturated=True)
    sat_seq = serve(1, saturated=True)
    eng.close()
    # algorithmic bytes of the served stream (every request's product once)
    p_pair = {}
    total_bytes = 0
    for k in range(n_req):
        key = -1 if is_big[k] else int(pick[k])
        if key not in p_pair:
            a, b = (big_a, big_b) if key < 0 else pairs[key]
            p_pair[key] = (a.rows, a.nnz, W.products(a, b))
        rows, annz, p = p_pair[key]
        total_bytes += streaming_bytes_counts(rows, annz, p, int(cb["nnz"][k]))
    achieved = total_bytes / cb["busy_s"] / 1e9 if cb["busy_s"] > 0 else 0.0
    return {
        "metric": f"served SpGEMM products/s at a Poisson offered rate of {d['rate_per_s']}/s (p50/p99 latency in "
                  "config)",
        "value": round(cb["throughput"], 1),
        "unit": "products/s",
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(cb["wall_s"] * 1e3, 3),
        "config": {"workload": "served_poisson", **{k: v for k, v in d.items() if k != "budget_bytes"},
                   "budget_gb": d["budget_bytes"] / 2 ** 30, "requests": n_req,
                   "batches": cb["batches"], "p50_ms": round(cb["p50_ms"], 3), "p99_ms": round(cb["p99_ms"], 3),
                   "device_busy_s": round(cb["busy_s"], 4), "wall_s": round(cb["wall_s"], 4),
                   "one_per_launch": {"products_per_s": round(seq["throughput"], 1),
                                      "p50_ms": round(seq["p50_ms"], 3), "p99_ms": round(seq["p99_ms"], 3)},
                   "saturated": {"requests": n_req,
                                 "products_per_s": round(sat_cb["throughput"], 1),
                                 "one_per_launch_products_per_s": round(sat_seq["throughput"], 1),
                                 "speedup_vs_one_per_launch": round(sat_cb["throughput"] / sat_seq["throughput"], 3),
                                 "batches": sat_cb["batches"], "p99_ms": round(sat_cb["p99_ms"], 3),
                                 "warmup": "one untimed pass of the same requests on the same server first"},
                   "p50_speedup_vs_one_per_launch": round(seq["p50_ms"] / cb["p50_ms"], 3),
                   "p99_speedup_vs_one_per_launch": round(seq["p99_ms"] / cb["p99_ms"], 3),
                   "parity": ("bit-exact: 24 sampled requests per mode equal the product computed alone"
                              if cb["exact"] and seq["exact"] else "MISMATCH"),
                   "config_reading": "configs are read 0-based: config-3-style = the batched jobs' products, "
                                     "config-1 shape = the 2^20 x16 product"},
        "roofline": {"bound": "hbm", "kernel": "served batches (device busy time of the continuous-batching runs)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": None,
                     "pipeline_bytes_model": f"streaming Gustavson summed over the {n_req} served products "
                                             "(offered-rate bound: the device idles between arrivals)",
                     "pipeline_bytes": int(total_bytes)},
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="uniform_2p20x16", choices=sorted(WORKLOADS))
    ap.add_argument("--no-sections", action="store_true",
                    help="default workload: skip the batched and served sections (for profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--row-range", default=None, metavar="R0:R1",
                    help="single-product workloads: only rows [R0, R1) of A (that block of C, as one product; "
                         "for profiling a region such as R-MAT's heaviest rows)")
    ap.add_argument("--jobs", type=int, default=0, help="batched workload: number of jobs (default 4096)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched workload: launch the pipeline eagerly instead of replaying a HIP graph "
                         "(for kernel traces)")
    ap.add_argument("--batch-only", action="store_true",
                    help="batched workload: time the batch pipeline only, so that each kernel's last dispatch "
                         "is the batch's (for counter profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this engine computes one rank's local product, and only a "
              "single-GPU figure is measured here; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    if wl["kind"] == "batched":
        steps = args.steps if args.steps is not None else 10
        warmup = args.warmup if args.warmup is not None else 2
        out = run_batched(args, wl, steps, warmup)
    elif wl["kind"] == "served":
        out = run_served(args, wl)
    else:
        out = run_single(args, args.workload, wl)
        if args.workload == "uniform_2p20x16" and not args.no_sections and not args.row_range:
            sec_args = argparse.Namespace(**{**vars(args), "steps": None, "warmup": None})
            b = run_batched(sec_args, WORKLOADS["batched_4096"])
            out["config"]["batched"] = {k: v for k, v in b.items() if k != "metric"} | {"metric": b["metric"]}
            s = run_served(sec_args, WORKLOADS["served_poisson"])
            out["config"]["served"] = {k: v for k, v in s.items() if k != "metric"} | {"metric": s["metric"]}
    for k in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"):
        out.setdefault(k, {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                           "dtype": "f64", "data": "synthetic (seeded operands generated on the host, resident "
                                                   "in HBM before timing)"}[k])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
