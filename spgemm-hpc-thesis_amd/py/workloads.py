"""Synthetic CSR operands for the benchmark workloads and tests.

Every generator is deterministic in its seed (numpy PCG64) and returns the
engine's Csr (int64 row offsets, int32 columns, float64 values).

  uniform_random  square or rectangular, exactly k distinct uniform columns per
                  row (stored ascending), values 2u-1 with u in [0, 1) at 53 bits
                  -- the 16384^2 x 8 and 2^20^2 x 16 workloads.
  banded          the reference CLI's synthetic matrix, makeBandedMatrix
                  (src/common/spgemm_common.cpp:348-367): column (17 row + e) % cols,
                  value 1 / (1 + (row + e) % 13), unsorted when columns wrap.
  rmat            R-MAT graph (a, b, c, d), duplicate edges merged, weights 1-u in (0, 1].
  stencil27       27-point operator on an n^3 grid, 26 on the diagonal, -1 off it.
  aggregation     piecewise-constant 2x2x2 aggregation P (n^3 x ceil(n/2)^3, one 1.0 per row).
  transpose       CSR transpose (R = P^T).
"""
from __future__ import annotations

import numpy as np

from spgemm_amd import Csr


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uniform_random(rows: int, cols: int, per_row: int, seed: int = 42) -> Csr:
    if per_row > cols:
        raise ValueError("per_row exceeds cols")
    rng = _rng(seed)
    if per_row * 2 > cols:
        c = np.stack([np.sort(rng.permutation(cols)[:per_row]) for _ in range(rows)]) if rows else \
            np.zeros((0, per_row), np.int64)
    else:
        c = rng.integers(0, cols, size=(rows, per_row), dtype=np.int64)
        c.sort(axis=1)
        while per_row > 1:
            bad = np.nonzero((np.diff(c, axis=1) == 0).any(axis=1))[0]
            if len(bad) == 0:
                break
            redo = rng.integers(0, cols, size=(len(bad), per_row), dtype=np.int64)
            redo.sort(axis=1)
            c[bad] = redo
    vals = 2.0 * rng.random(rows * per_row) - 1.0
    rp = np.arange(rows + 1, dtype=np.int64) * per_row
    return Csr.make(rows, cols, rp, c.reshape(-1).astype(np.int32), vals)


def banded(rows: int, cols: int, per_row: int) -> Csr:
    """makeBandedMatrix (spgemm_common.cpp:348-367)."""
    r = np.repeat(np.arange(rows, dtype=np.int64), per_row)
    e = np.tile(np.arange(per_row, dtype=np.int64), rows)
    col = (r * 17 + e) % cols
    val = 1.0 / (1.0 + ((r + e) % 13).astype(np.float64))
    rp = np.arange(rows + 1, dtype=np.int64) * per_row
    return Csr.make(rows, cols, rp, col.astype(np.int32), val)


def from_coo(rows: int, cols: int, r: np.ndarray, c: np.ndarray, v: np.ndarray) -> Csr:
    """Stable CSR assembly keeping the given entry order inside each row."""
    order = np.argsort(r, kind="stable")
    counts = np.bincount(r, minlength=rows).astype(np.int64)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    return Csr.make(rows, cols, rp, c[order].astype(np.int32), v[order])


def rmat(scale: int, edge_factor: int, a=0.57, b=0.19, c=0.19, d=0.05, seed: int = 42) -> Csr:
    n = 1 << scale
    m = edge_factor * n
    rng = _rng(seed)
    row = np.zeros(m, dtype=np.int64)
    col = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    for _ in range(scale):
        u = rng.random(m)
        row <<= 1
        col <<= 1
        row |= (u >= ab)
        col |= ((u >= a) & (u < ab)) | (u >= abc)
    key = np.unique(row * n + col)  # merge duplicates, sorted by (row, col)
    r = (key // n).astype(np.int64)
    cc = (key % n).astype(np.int32)
    w = 1.0 - rng.random(len(key))  # (0, 1]
    counts = np.bincount(r, minlength=n).astype(np.int64)
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rp[1:])
    return Csr.make(n, n, rp, cc, w)


def stencil27(n: int) -> Csr:
    idx = np.arange(n ** 3, dtype=np.int64)
    i, j, k = idx // (n * n), (idx // n) % n, idx % n
    rows, cols, vals = [], [], []
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                ii, jj, kk = i + di, j + dj, k + dk
                ok = (ii >= 0) & (ii < n) & (jj >= 0) & (jj < n) & (kk >= 0) & (kk < n)
                rows.append(idx[ok])
                cols.append((ii * n * n + jj * n + kk)[ok])
                vals.append(np.full(int(ok.sum()), 26.0 if (di, dj, dk) == (0, 0, 0) else -1.0))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    return from_coo(n ** 3, n ** 3, r[order], c[order], v[order])


def aggregation(n: int) -> Csr:
    m = (n + 1) // 2
    idx = np.arange(n ** 3, dtype=np.int64)
    i, j, k = idx // (n * n), (idx // n) % n, idx % n
    agg = (i // 2) * m * m + (j // 2) * m + (k // 2)
    return Csr.make(n ** 3, m ** 3, np.arange(n ** 3 + 1, dtype=np.int64), agg.astype(np.int32),
                    np.ones(n ** 3))


def transpose(a: Csr) -> Csr:
    r = np.repeat(np.arange(a.rows, dtype=np.int64), np.diff(a.row_ptr))
    return from_coo(a.cols, a.rows, a.col.astype(np.int64), r, a.val)


def products(a: Csr, b: Csr) -> int:
    """P = sum over A entries of nnz(B row) (the flop basis / 2)."""
    blen = np.diff(b.row_ptr)
    return int(blen[a.col].sum()) if a.nnz else 0


def row_products(a: Csr, b: Csr) -> np.ndarray:
    blen = np.diff(b.row_ptr)
    per_entry = blen[a.col] if a.nnz else np.zeros(0, np.int64)
    cs = np.zeros(a.nnz + 1, dtype=np.int64)
    np.cumsum(per_entry, out=cs[1:])
    return cs[a.row_ptr[1:]] - cs[a.row_ptr[:-1]]


def streaming_bytes(a: Csr, b_rows: int, p: int, c_nnz: int) -> int:
    """Streaming-Gustavson traffic model: A once, B row bounds and entries once
    per referencing A entry, C once."""
    m = a.rows
    return (8 * (m + 1) + 12 * a.nnz) + (16 * a.nnz + 12 * p) + (8 * (m + 1) + 12 * c_nnz)


def compulsory_bytes(a: Csr, b: Csr, c_nnz: int) -> int:
    """A, B and C each moved once."""
    return (8 * (a.rows + 1) + 12 * a.nnz) + (8 * (b.rows + 1) + 12 * b.nnz) + (8 * (a.rows + 1) + 12 * c_nnz)


def variable_uniform(rows: int, cols: int, kmin: int, kmax: int, rng: np.random.Generator) -> Csr:
    """Rows with k ~ U{kmin..kmax} distinct uniform columns each (ascending),
    values U[-1, 1): a uniform kmax-subset per row, then a uniform k-subset of it."""
    K = min(kmax, cols)
    c = rng.integers(0, cols, size=(rows, K), dtype=np.int64)
    c.sort(axis=1)
    while K > 1:
        bad = np.nonzero((np.diff(c, axis=1) == 0).any(axis=1))[0]
        if len(bad) == 0:
            break
        redo = rng.integers(0, cols, size=(len(bad), K), dtype=np.int64)
        redo.sort(axis=1)
        c[bad] = redo
    k = np.minimum(rng.integers(kmin, kmax + 1, size=rows), K)
    order = np.argsort(rng.random((rows, K)), axis=1)
    rank = np.empty_like(order)
    np.put_along_axis(rank, order, np.arange(K)[None, :].repeat(rows, axis=0), axis=1)
    keep = rank < k[:, None]
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(k, out=rp[1:])
    return Csr.make(rows, cols, rp, c[keep].astype(np.int32), 2.0 * rng.random(int(rp[-1])) - 1.0)


def batch_job(i: int, seed: int = 42, min_rows: int = 256, max_rows: int = 65536, n_lo: int = 16, n_hi: int = 48):
    """Job i of the batched workload (its own random stream, so jobs can be
    generated in parallel or alone).  Job i with i % 3 == 2 is a Galerkin
    triple ("rap", R, A, P): A the 27-point stencil on an m^3 grid (m uniform
    in n_lo..n_hi), P the 2x2x2 aggregation, R = P^T.  Every other job is a
    plain product ("ab", A, B) with rows = cols log-uniform in
    [min_rows, max_rows] and U{4..32} uniform columns per row."""
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, i])))
    if i % 3 == 2:
        m = int(rng.integers(n_lo, n_hi + 1))
        p = aggregation(m)
        return ("rap", transpose(p), stencil27(m), p)
    rows = int(np.exp(rng.uniform(np.log(min_rows), np.log(max_rows))))
    return ("ab", variable_uniform(rows, rows, 4, 32, rng), variable_uniform(rows, rows, 4, 32, rng))


def batch_jobs(n: int = 4096, seed: int = 42, processes: int = 1, **kw):
    """Jobs 0..n-1 of the batched workload (see batch_job)."""
    if processes <= 1:
        return [batch_job(i, seed, **kw) for i in range(n)]
    import functools
    import multiprocessing as mp
    with mp.get_context("fork").Pool(processes) as pool:
        return pool.map(functools.partial(batch_job, seed=seed, **kw), range(n), chunksize=16)
