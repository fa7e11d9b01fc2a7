"""The allocation-free product (spg_multiply_noalloc): called through the C
ABI into device buffers the test owns, it equals spg_multiply_begin's result
bit for bit, on consecutive calls whose second input is smaller (the shared
scratch array keeps the last call's contents), on every single-pass tiling
and both entry-stream formats, including big rows whose sums cancel to
exactly zero; sizes beyond its fixed scratch and capacities below C's
symbolic bound are rejected."""
import ctypes

import numpy as np
import pytest

import oracle as O
import spgemm_amd as S
import workloads as W
from parity_util import device_to_host, digest

pytestmark = pytest.mark.gpu

SPG_OK, SPG_EINVAL, SPG_ENOMEM = 0, 1, 3


class DeviceBuffers:
    """C's three arrays in device memory owned by the test (hipMalloc)."""

    def __init__(self, rows, capacity=0):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []
        self.rows = rows
        self.row_ptr = self._alloc((rows + 1) * 8)
        self.col = self.val = 0
        self.capacity = 0
        if capacity:
            self.reserve(capacity)

    def reserve(self, capacity):
        self.col = self._alloc(max(capacity, 1) * 4)
        self.val = self._alloc(max(capacity, 1) * 8)
        self.capacity = capacity

    def _alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p.value

    def host(self, cols, nnz):
        rp = device_to_host(self.row_ptr, self.rows + 1, np.int64)
        return S.Csr(self.rows, cols, rp, device_to_host(self.col, nnz, np.int32),
                     device_to_host(self.val, nnz, np.float64))

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc_product(engine, a, b):
    """spg_multiply_noalloc with buffers sized by its own capacity query."""
    da, db = engine.upload(a), engine.upload(b)
    bufs = DeviceBuffers(a.rows)
    try:
        # capacity query: row offsets only; SPG_ENOMEM with C's symbolic bound
        # unless C is empty (then the product is complete already)
        s, bound = engine.multiply_noalloc(da, db, bufs.row_ptr, 0, 0, 0)
        if W.products(a, b):
            assert s == SPG_ENOMEM and bound > 0, (s, engine.last_error())
        else:
            assert s == SPG_OK and bound == 0, (s, engine.last_error())
        bufs.reserve(bound)
        s, nnz = engine.multiply_noalloc(da, db, bufs.row_ptr, bufs.col, bufs.val, bound)
        assert s == SPG_OK, engine.last_error()
        assert nnz <= bound
        st = engine.stats()
        return bufs.host(b.cols, nnz), st
    finally:
        bufs.free()
        da.free()
        db.free()


def pm1(rng, rows, cols, per_row):
    rp = np.arange(rows + 1, dtype=np.int64) * per_row
    return S.Csr.make(rows, cols, rp, rng.integers(0, cols, size=rows * per_row),
                      rng.choice(np.array([-1.0, 1.0]), size=rows * per_row))


def cases():
    rng = np.random.Generator(np.random.PCG64(71))
    rmat = W.rmat(13, 32, seed=42)
    # 40 rows of 24,000 +-1 products into 64 columns: running sums cross zero
    # and some end at exactly zero in the large-row path
    cancel = (pm1(rng, 40, 4000, 3000), pm1(rng, 4000, 64, 8))
    return [("rmat_s13_AA", rmat, rmat),                    # every bin, heavy rows
            ("large_exact_zeros", *cancel),
            ("uniform_3000x9", W.uniform_random(3000, 3000, 9, seed=5), W.uniform_random(3000, 3000, 9, seed=6)),
            ("empty_rows", S.Csr.make(3, 4, [0, 0, 0, 0], [], []), W.uniform_random(4, 5, 2, seed=1))]


@pytest.mark.parametrize("path,fmt", [(S.PATH_AUTO, S.FORMAT_AUTO), (S.PATH_SINGLE_PASS_PACKED, S.FORMAT_WIDE),
                                      (S.PATH_SINGLE_PASS_DEALT, S.FORMAT_AUTO)],
                         ids=["auto-narrow", "packed-wide", "dealt-narrow"])
def test_noalloc_equals_multiply_begin(engine, path, fmt):
    # Consecutive calls, each input smaller than the one before.
    engine.set_path(path)
    engine.set_format(fmt)
    try:
        for name, a, b in cases():
            print("case", name, flush=True)
            want = engine.multiply(a, b)  # spg_multiply_begin + spg_multiply_finish
            want_st = engine.stats()
            got, st = noalloc_product(engine, a, b)
            assert got.nnz == want.nnz, name
            assert digest(got) == digest(want), name
            assert st["products"] == want_st["products"] and st["nnz"] == want.nnz, name
            assert st["wide_format"] == (fmt == S.FORMAT_WIDE), name
            if name == "large_exact_zeros":
                assert st["rows_per_bin"][6] == 40 and want_st["compacted"]
                ref = O.product(O.Csr.make(a.rows, a.cols, a.row_ptr, a.col, a.val),
                                O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val))
                assert digest(got) == O.csr_digest(ref)
    finally:
        engine.set_path(S.PATH_AUTO)
        engine.set_format(S.FORMAT_AUTO)


def test_noalloc_limits_and_errors(engine):
    a = W.uniform_random(2000, 2000, 8, seed=2)
    da = engine.upload(a)
    bufs = DeviceBuffers(a.rows, 100)
    try:
        # capacity below the symbolic bound: rejected, the bound reported
        s, bound = engine.multiply_noalloc(da, da, bufs.row_ptr, bufs.col, bufs.val, 100)
        assert s == SPG_ENOMEM and bound == W.products(a, a)  # every row single-pass: bound = P
        assert "symbolic bound" in engine.last_error()
        # host views are refused
        L = S.lib()
        hv = S.host_view(a)
        n = ctypes.c_int64()
        assert L.spg_multiply_noalloc(engine._h, ctypes.byref(hv), ctypes.byref(hv), bufs.row_ptr, bufs.col,
                                      bufs.val, 100, ctypes.byref(n)) == SPG_EINVAL
    finally:
        bufs.free()
        da.free()
    # more rows than the fixed scratch holds: rejected before any kernel runs
    rows = 1 << 22
    big = S.Csr.make(rows, 1, np.arange(rows + 1, dtype=np.int64), np.zeros(rows, np.int32), np.ones(rows))
    one = S.Csr.make(1, 1, [0, 1], [0], [2.0])
    db, dbig = engine.upload(one), engine.upload(big)
    bufs = DeviceBuffers(rows, rows)
    try:
        s, _ = engine.multiply_noalloc(dbig, db, bufs.row_ptr, bufs.col, bufs.val, rows)
        assert s == SPG_ENOMEM and "maximum is 536870912 bytes" in engine.last_error()
    finally:
        bufs.free()
        dbig.free()
        db.free()
    # the engine and the entry point stay usable
    a2, b2 = W.uniform_random(500, 700, 5, seed=8), W.uniform_random(700, 900, 6, seed=9)
    got, _ = noalloc_product(engine, a2, b2)
    assert digest(got) == digest(engine.multiply(a2, b2))
