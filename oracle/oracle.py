"""ctypes bindings for oracle/liboracle.so (the C restatement in oracle.c).

Test infrastructure only: the checker, never the thing measured or shipped.
See oracle.c for the reference file:line each function restates and for the
checks that pin it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "refdrv")


@dataclass
class Csr:
    rows: int
    cols: int
    row_ptr: np.ndarray  # int64[rows+1]
    col: np.ndarray      # int32[nnz]
    val: np.ndarray      # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if len(self.row_ptr) else 0

    @staticmethod
    def make(rows, cols, row_ptr, col, val) -> "Csr":
        return Csr(int(rows), int(cols), np.ascontiguousarray(row_ptr, dtype=np.int64),
                   np.ascontiguousarray(col, dtype=np.int32), np.ascontiguousarray(val, dtype=np.float64))


class _OrcCsr(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double))]


class _FullReport(C.Structure):
    _fields_ = [("passed", C.c_int32), ("environment_ok", C.c_int32), ("matrices_ok", C.c_int32),
                ("worst_row", C.c_int32), ("worst_column", C.c_int32), ("reference_digits", C.c_int32),
                ("rejected", C.c_int64), ("row_nnz_mismatches", C.c_int64),
                ("support_mismatches", C.c_int64), ("reference_nnz", C.c_int64),
                ("max_abs_error", C.c_double), ("max_reference_error", C.c_double),
                ("max_error_ratio", C.c_double), ("worst_actual", C.c_double),
                ("worst_reference", C.c_double), ("worst_bound", C.c_double),
                ("max_entry_rel_error", C.c_double), ("frobenius_rel_error", C.c_double)]


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"oracle error {code}: {message}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_last_error.restype = C.c_char_p
        L.orc_product.argtypes = [P(_OrcCsr), P(_OrcCsr), P(_OrcCsr), C.c_int]
        L.orc_scalar_products.argtypes = [P(_OrcCsr), P(_OrcCsr), P(C.c_int64)]
        L.orc_partial_counts.argtypes = [P(_OrcCsr), P(_OrcCsr), P(C.c_int64), P(C.c_int32)]
        L.orc_partial_compare.argtypes = [C.c_int32, C.c_int32, C.c_int32, P(C.c_int64), P(_OrcCsr),
                                          P(C.c_int64)]
        L.orc_full_validate.argtypes = [P(_OrcCsr), P(_OrcCsr), P(_OrcCsr), P(_FullReport)]
        L.orc_full_environment.restype = C.c_int
        L.orc_percentile90.argtypes = [P(C.c_double), C.c_int64]
        L.orc_percentile90.restype = C.c_double
        L.orc_mm_read.argtypes = [C.c_char_p, P(_OrcCsr)]
        L.orc_free.argtypes = [P(_OrcCsr)]
        L.orc_acc_create.argtypes = [C.c_int32, C.c_int32]
        L.orc_acc_create.restype = C.c_void_p
        L.orc_acc_destroy.argtypes = [C.c_void_p]
        L.orc_acc_add.argtypes = [C.c_void_p, P(_OrcCsr), P(_OrcCsr)]
        L.orc_acc_finish.argtypes = [C.c_void_p, P(_OrcCsr)]
        _lib = L
    return _lib


def _view(m: Csr) -> _OrcCsr:
    s = _OrcCsr()
    s.rows, s.cols = m.rows, m.cols
    s.nnz = int(m.row_ptr[-1]) if len(m.row_ptr) else 0
    s.row_ptr = m.row_ptr.ctypes.data_as(C.POINTER(C.c_int64))
    s.col = m.col.ctypes.data_as(C.POINTER(C.c_int32))
    s.val = m.val.ctypes.data_as(C.POINTER(C.c_double))
    return s


def _take(s: _OrcCsr) -> Csr:
    n = int(s.nnz)
    rp = np.ctypeslib.as_array(s.row_ptr, shape=(s.rows + 1,)).copy()
    col = np.ctypeslib.as_array(s.col, shape=(max(n, 1),))[:n].copy()
    val = np.ctypeslib.as_array(s.val, shape=(max(n, 1),))[:n].copy()
    lib().orc_free(C.byref(s))
    return Csr(int(s.rows), int(s.cols), rp, col, val)


def _check(code: int):
    if code != 0:
        raise OracleError(code, lib().orc_last_error().decode(errors="replace"))


def product(a: Csr, b: Csr, threads: int = 0) -> Csr:
    """Serial-order product (spgemm_common.cpp:131-178; spgemm_exchange.cpp:274-324)."""
    out = _OrcCsr()
    _check(lib().orc_product(C.byref(_view(a)), C.byref(_view(b)), C.byref(out), threads))
    return _take(out)


def scalar_products(a: Csr, b: Csr) -> int:
    """scalarMultiplicationCount (spgemm_common.cpp:474-487)."""
    n = C.c_int64()
    _check(lib().orc_scalar_products(C.byref(_view(a)), C.byref(_view(b)), C.byref(n)))
    return int(n.value)


def partial_counts(a: Csr, b: Csr):
    """buildPartialReference (spgemm_common.cpp:541-554) -> (row_nnz int64[rows], finite)."""
    counts = np.zeros(max(a.rows, 1), dtype=np.int64)
    fin = C.c_int32()
    _check(lib().orc_partial_counts(C.byref(_view(a)), C.byref(_view(b)),
                                    counts.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(fin)))
    return counts[: a.rows], bool(fin.value)


def partial_validate(a: Csr, b: Csr, c: Csr):
    """Partial mode (spgemm_common.cpp:556-572) -> (passed, row_nnz_mismatches)."""
    counts, fin = partial_counts(a, b)
    mism = C.c_int64()
    cnt = np.ascontiguousarray(counts if len(counts) else np.zeros(1, dtype=np.int64))
    ok = lib().orc_partial_compare(a.rows, b.cols, int(fin), cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                   C.byref(_view(c)), C.byref(mism))
    return bool(ok), int(mism.value)


def full_validate(a: Csr, b: Csr, c: Csr) -> dict:
    """Full mode, componentwise_dot_bound_v1 (spgemm_common.cpp:590-851)."""
    rep = _FullReport()
    _check(lib().orc_full_validate(C.byref(_view(a)), C.byref(_view(b)), C.byref(_view(c)), C.byref(rep)))
    d = {name: getattr(rep, name) for name, _ in _FullReport._fields_}
    d["passed"] = bool(d["passed"]) and bool(d["environment_ok"]) and bool(d["matrices_ok"])
    d["criterion"] = "componentwise_dot_bound_v1"
    return d


def full_environment() -> bool:
    return bool(lib().orc_full_environment())


def percentile90(samples) -> float:
    s = np.ascontiguousarray(samples, dtype=np.float64)
    return float(lib().orc_percentile90(s.ctypes.data_as(C.POINTER(C.c_double)), len(s)))


def mm_read(path: str) -> Csr:
    """readMatrixMarket restated (matrix_market.cpp:28-144)."""
    out = _OrcCsr()
    _check(lib().orc_mm_read(path.encode(), C.byref(out)))
    return _take(out)


def accumulate(rows: int, cols: int, stages) -> Csr:
    """trident::LocalAccumulator (matrix_blocks.cpp:152-200) over [(A, B), ...]."""
    L = lib()
    h = L.orc_acc_create(rows, cols)
    try:
        for a, b in stages:
            _check(L.orc_acc_add(h, C.byref(_view(a)), C.byref(_view(b))))
        out = _OrcCsr()
        _check(L.orc_acc_finish(h, C.byref(out)))
        return _take(out)
    finally:
        L.orc_acc_destroy(h)


# ---------------------------------------------------------------------------
# Binary CSR files shared with oracle/_ref/refdrv
# ---------------------------------------------------------------------------
MAGIC = b"SPGCSR01"


def write_csr(path: str, m: Csr):
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([m.rows, m.cols], dtype=np.int32).tobytes())
        f.write(np.array([m.nnz], dtype=np.int64).tobytes())
        f.write(m.row_ptr.astype("<i8").tobytes())
        f.write(m.col.astype("<i4").tobytes())
        f.write(m.val.astype("<f8").tobytes())


def read_csr(path: str) -> Csr:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC:
        raise ValueError("bad csr magic: " + path)
    rows, cols = np.frombuffer(data, dtype="<i4", count=2, offset=8)
    nnz = int(np.frombuffer(data, dtype="<i8", count=1, offset=16)[0])
    off = 24
    rp = np.frombuffer(data, dtype="<i8", count=int(rows) + 1, offset=off).copy()
    off += 8 * (int(rows) + 1)
    col = np.frombuffer(data, dtype="<i4", count=nnz, offset=off).copy()
    off += 4 * nnz
    val = np.frombuffer(data, dtype="<f8", count=nnz, offset=off).copy()
    return Csr(int(rows), int(cols), rp, col, val)


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def ref_run(*args, timeout=600) -> str:
    """Run oracle/_ref/refdrv (the reference's own functions) and return stdout."""
    r = subprocess.run([REF_BIN, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise OracleError(r.returncode, r.stderr.strip())
    return r.stdout


def csr_digest(m: Csr) -> str:
    """SHA-256 over (rows, cols, row_ptr int64 LE, col int32 LE, value bits LE)."""
    import hashlib
    h = hashlib.sha256()
    h.update(np.array([m.rows, m.cols], dtype="<i4").tobytes())
    h.update(m.row_ptr.astype("<i8").tobytes())
    h.update(m.col.astype("<i4").tobytes())
    h.update(m.val.astype("<f8").view("<u8").tobytes())
    return h.hexdigest()
