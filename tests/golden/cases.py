"""Deterministic inputs of the golden cases (shared by make_golden.py and the tests).

Each case is (name, A, B) with A and B oracle.Csr matrices.  Inputs come from:
  * the reference's own data files (tests/golden/data: tests/data/*.mtx,
    tests/data/hv15r_roundoff/*, matrices/cage8*.mtx), read with the oracle reader;
  * the input formulas of the reference's unit tests:
      smoke_<seed>   tests/spgemm_serial_smoke.cpp:10-21,33-36 (unsorted columns,
                     duplicate coordinates, explicit zeros, empty rows)
      dyadic_<i>     tests/spgemm_full_streaming.cpp:26-37,42-45 (std::mt19937(42))
      banded         makeBandedMatrix, src/common/spgemm_common.cpp:348-367
  * seeded random matrices with real fp64 roundoff (uniform values, duplicate
    and unsorted columns, heavy rows that need the large-row path).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import Csr, mm_read  # noqa: E402

DATA = os.path.join(HERE, "data")

MM_PAIRS = [
    ("rectangular", "rectangular_a.mtx", "rectangular_b.mtx"),
    ("cancellation", "cancellation.mtx", "cancellation.mtx"),
    ("roundoff_stages", "roundoff_stages.mtx", "roundoff_stages.mtx"),
    ("tiny_symmetric", "tiny_symmetric.mtx", "tiny_symmetric.mtx"),
    ("tiny_rectangular_t", "tiny_rectangular.mtx", "tiny_symmetric.mtx"),
    ("empty", "empty.mtx", "empty.mtx"),
    ("overflow_product", "overflow_product.mtx", "overflow_product.mtx"),
    ("cage8_AA", "cage8.mtx", "cage8.mtx"),
    ("cage8_AR", "cage8.mtx", "cage8_restriction.mtx"),
    ("hv15r_zero", "hv15r_roundoff/zero-cancellation-witness-A.mtx",
     "hv15r_roundoff/zero-cancellation-witness-B.mtx"),
    ("hv15r_absolute", "hv15r_roundoff/absolute-threshold-witness-A.mtx",
     "hv15r_roundoff/absolute-threshold-witness-B.mtx"),
]


def data_path(name: str) -> str:
    return os.path.join(DATA, name)


class MT19937:
    """std::mt19937 (32-bit Mersenne Twister, init_genrand seeding)."""

    def __init__(self, seed: int):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def __call__(self) -> int:
        if self.i >= 624:
            mt = self.mt
            for k in range(624):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def _csr(rows, cols, rp, col, val) -> Csr:
    return Csr.make(rows, cols, np.asarray(rp, np.int64), np.asarray(col, np.int32), np.asarray(val, np.float64))


def smoke_input(rows: int, cols: int, salt: int) -> Csr:
    """tests/spgemm_serial_smoke.cpp:10-21."""
    rp, col, val = [0], [], []
    for row in range(rows):
        entry = 0
        while cols and (row + salt) % 4 and entry < 2 * cols + 1:
            col.append((entry * 7 + row + salt) % cols)
            val.append((entry * 3 + row + salt) % 7 - 3)
            entry += 1
        rp.append(len(val))
    return _csr(rows, cols, rp, col, val)


def smoke_case(seed: int):
    """tests/spgemm_serial_smoke.cpp:33-36."""
    a = smoke_input(seed % 8, (seed * 3) % 7, seed)
    b = smoke_input(a.cols, (seed * 5) % 9, seed + 1)
    return a, b


def dyadic_cases():
    """tests/spgemm_full_streaming.cpp:26-45 (32 cases from one mt19937(42))."""
    rnd = MT19937(42)

    def inp(rows, cols):
        rp, col, val = [0], [], []
        for _ in range(rows):
            for _ in range(2 * cols):
                col.append(rnd() % cols)
                val.append(np.ldexp(float(int(rnd() % 9) - 4), -20))
            rp.append(len(val))
        return _csr(rows, cols, rp, col, val)

    out = []
    for i in range(32):
        a = inp(i % 7 + 1, 5)
        b = inp(5, i % 4 + 1)
        out.append((a, b))
    return out


def banded(rows: int, cols: int, per_row: int) -> Csr:
    """makeBandedMatrix (src/common/spgemm_common.cpp:348-367)."""
    r = np.repeat(np.arange(rows, dtype=np.int64), per_row)
    e = np.tile(np.arange(per_row, dtype=np.int64), rows)
    return _csr(rows, cols, np.arange(rows + 1, dtype=np.int64) * per_row, (r * 17 + e) % cols,
                1.0 / (1.0 + ((r + e) % 13).astype(np.float64)))


def random_csr(rows: int, cols: int, lens, seed: int, *, sort_cols=True, distinct=True, zero_frac=0.0,
               int_values=False) -> Csr:
    """Random CSR with the given per-row lengths (array or int)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = np.broadcast_to(np.asarray(lens, dtype=np.int64), (rows,)).copy()
    if distinct:
        lens = np.minimum(lens, cols)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.empty(int(rp[-1]), dtype=np.int64)
    for r in range(rows):
        n = int(lens[r])
        if n == 0:
            continue
        if distinct:
            c = rng.choice(cols, size=n, replace=False)
        else:
            c = rng.integers(0, cols, size=n)
        if sort_cols:
            c = np.sort(c)
        col[rp[r]:rp[r + 1]] = c
    if int_values:
        val = rng.integers(-4, 5, size=len(col)).astype(np.float64)
    else:
        val = 2.0 * rng.random(len(col)) - 1.0
    if zero_frac > 0:
        val[rng.random(len(col)) < zero_frac] = 0.0
    return _csr(rows, cols, rp, col, val)


def rmat(scale: int, edge_factor: int, seed: int, a=0.57, b=0.19, c=0.19) -> Csr:
    n = 1 << scale
    m = edge_factor * n
    rng = np.random.Generator(np.random.PCG64(seed))
    row = np.zeros(m, dtype=np.int64)
    col = np.zeros(m, dtype=np.int64)
    for _ in range(scale):
        u = rng.random(m)
        row = (row << 1) | (u >= a + b)
        col = (col << 1) | (((u >= a) & (u < a + b)) | (u >= a + b + c))
    key = np.unique(row * n + col)
    r = key // n
    w = 1.0 - rng.random(len(key))
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    return _csr(n, n, rp, key % n, w)


def synthetic_cases():
    """Named synthetic (A, B) pairs; small enough for the oracle in seconds."""
    cases = []
    cases.append(("uniform_2048x8", random_csr(2048, 2048, 8, 1), random_csr(2048, 2048, 8, 2)))
    cases.append(("uniform_rect", random_csr(777, 1500, 12, 3), random_csr(1500, 3001, 9, 4)))
    rng = np.random.Generator(np.random.PCG64(5))
    lens = rng.integers(0, 40, size=1200)
    cases.append(("ragged_dups_unsorted", random_csr(1200, 900, lens, 6, sort_cols=False, distinct=False),
                  random_csr(900, 700, rng.integers(0, 30, size=900), 7, sort_cols=False, distinct=False)))
    cases.append(("int_cancel_zeros", random_csr(1000, 64, 20, 8, distinct=False, int_values=True, zero_frac=0.1),
                  random_csr(64, 50, 10, 9, distinct=False, int_values=True, zero_frac=0.1)))
    heavy_lens = np.where(np.arange(300) % 37 == 0, 3000, 6)
    cases.append(("heavy_rows", random_csr(300, 4000, heavy_lens, 10, sort_cols=False),
                  random_csr(4000, 5000, np.random.Generator(np.random.PCG64(11)).integers(1, 60, size=4000), 12)))
    cases.append(("wide_dense_rows", random_csr(40, 200, 200, 13), random_csr(200, 70000, 150, 14)))
    cases.append(("banded_16384x8", banded(16384, 16384, 8), banded(16384, 16384, 8)))
    r10 = rmat(10, 32, 15)
    cases.append(("rmat_s10_AA", r10, r10))
    return cases


def mm_cases():
    return [(name, mm_read(data_path(a)), mm_read(data_path(b))) for name, a, b in MM_PAIRS]


def split_inner(a: Csr, b: Csr, bounds):
    """Stage operands (A[:, k0:k1] with columns shifted, B[k0:k1, :]) for each
    [k0, k1) in bounds, keeping stored order (a Trident-style C-stationary split)."""
    stages = []
    for k0, k1 in bounds:
        rows_a, cols_a, vals_a, rp_a = [], [], [], [0]
        for r in range(a.rows):
            s, e = a.row_ptr[r], a.row_ptr[r + 1]
            c = a.col[s:e]
            keep = (c >= k0) & (c < k1)
            cols_a.append(c[keep] - k0)
            vals_a.append(a.val[s:e][keep])
            rp_a.append(rp_a[-1] + int(keep.sum()))
        As = _csr(a.rows, k1 - k0, rp_a, np.concatenate(cols_a) if cols_a else [], np.concatenate(vals_a) if vals_a else [])
        s0, s1 = b.row_ptr[k0], b.row_ptr[k1]
        Bs = _csr(k1 - k0, b.cols, b.row_ptr[k0:k1 + 1] - s0, b.col[s0:s1], b.val[s0:s1])
        stages.append((As, Bs))
    return stages


def accumulate_cases():
    """Named staged products [(A_s, B_s), ...] with the final shape."""
    out = []
    cage = mm_read(data_path("cage8.mtx"))
    out.append(("cage8_3stages", cage.rows, cage.cols, split_inner(cage, cage, [(0, 300), (300, 700), (700, 1015)])))
    ro = mm_read(data_path("roundoff_stages.mtx"))
    out.append(("roundoff_2stages", ro.rows, ro.cols, split_inner(ro, ro, [(4, 8), (0, 4)])))
    a = random_csr(500, 400, 15, 21, distinct=False, int_values=True, zero_frac=0.2)
    b = random_csr(400, 300, 9, 22, distinct=False, int_values=True)
    out.append(("int_4stages", 500, 300, split_inner(a, b, [(0, 100), (100, 200), (200, 300), (300, 400)])))
    return out
