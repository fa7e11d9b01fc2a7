"""Python bindings (ctypes) for libspgemm_amd.so, the MI355X SpGEMM engine.

This is a thin layer over the C ABI in include/spgemm_amd.h, used by the tests
and the benchmark.  Importing it never falls back to anything: if the HIP
library is missing the import of a context raises.

    eng = Engine()                      # spg_create on device 0
    C = eng.multiply(A, B)              # drop-in product, host CSR in/out
    dA = eng.upload(A)                  # operand resident in HBM
    eng.multiply_device(dA, dB)         # product of resident operands
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libspgemm_amd.so")

SPG_HOST, SPG_DEVICE = 0, 1

STATUS = {0: "SPG_OK", 1: "SPG_EINVAL", 2: "SPG_EOVERFLOW", 3: "SPG_ENOMEM", 4: "SPG_EMISSING_ROW",
          5: "SPG_ERANGE", 6: "SPG_EDEVICE", 7: "SPG_ENODEVICE", 8: "SPG_ESTATE", 9: "SPG_EPARSE",
          10: "SPG_EBUSY"}

EXPORTED = [
    "spg_create", "spg_destroy", "spg_status_string", "spg_last_error", "spg_set_stream", "spg_set_timing",
    "spg_last_stats", "spg_multiply_begin", "spg_multiply_finish", "spg_result_device",
    "spg_accumulator_create", "spg_accumulator_add", "spg_accumulator_finish", "spg_accumulator_copy",
    "spg_accumulator_destroy", "spg_matrix_read_mm", "spg_matrix_view", "spg_matrix_free",
    "spg_matrix_last_error", "spg_matrix_upload", "spg_device_matrix_view", "spg_device_matrix_free",
    "spg_synchronize", "spg_set_path", "spg_set_format", "spg_set_large", "spg_multiply_rows", "spg_result_rows",
    "spg_multiply_noalloc",
    "spg_batch_create", "spg_batch_create_chained", "spg_batch_refresh", "spg_batch_run", "spg_batch_count",
    "spg_batch_set_graph",
    "spg_batch_autotune", "spg_batch_product_nnz", "spg_batch_copy", "spg_batch_info", "spg_batch_last_stats",
    "spg_batch_destroy", "spg_batch_rebuild", "spg_batch_set_stream", "spg_batch_device_bytes", "spg_batch_trim",
    "spg_batch_estimate", "spg_batch_last_error", "spg_context_stream", "spg_context_bind",
    "spg_server_create", "spg_server_submit", "spg_server_register", "spg_server_unregister",
    "spg_server_submit_resident", "spg_server_wait", "spg_server_result", "spg_server_release", "spg_server_stats",
    "spg_server_destroy",
    "spg_matrix_write_csr", "spg_matrix_read_csr",
    "spg_halo_upload", "spg_halo_pack", "spg_device_halo_view", "spg_device_halo_free",
    "spg_validation_last_error", "spg_validate_full", "spg_full_report_text", "spg_partial_counts",
    "spg_validate_partial", "spg_spref_write", "spg_spref_read", "spg_sfull_write", "spg_sfull_open",
    "spg_sfull_next_row", "spg_sfull_row_cells", "spg_sfull_finish", "spg_sfull_close", "spg_sha256_hex",
]

PATH_AUTO, PATH_SINGLE_PASS, PATH_TWO_PHASE, PATH_SINGLE_PASS_DEALT, PATH_SINGLE_PASS_PACKED = 0, 1, 2, 3, 4
FORMAT_AUTO, FORMAT_WIDE = 0, 1
LARGE_AUTO, LARGE_WINDOWS = 0, 1


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


class CsrView(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("values", C.c_void_p),
                ("location", C.c_int32)]


class HaloView(C.Structure):
    _fields_ = [("n_rows", C.c_int32), ("global_rows", C.c_void_p), ("row_ptr", C.c_void_p),
                ("col_idx", C.c_void_p), ("values", C.c_void_p), ("location", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("products", C.c_int64), ("structural_nnz", C.c_int64), ("nnz", C.c_int64),
                ("rows_per_bin", C.c_int64 * 8), ("compacted", C.c_int32), ("bins", C.c_int32),
                ("single_pass", C.c_int32), ("wide_format", C.c_int32),
                ("ms_total", C.c_double), ("ms_symbolic", C.c_double), ("ms_numeric", C.c_double),
                ("direct", C.c_int32), ("partitioned", C.c_int32)]

    def as_dict(self) -> dict:
        return {"products": self.products, "structural_nnz": self.structural_nnz, "nnz": self.nnz,
                "rows_per_bin": list(self.rows_per_bin)[: self.bins], "compacted": bool(self.compacted),
                "single_pass": bool(self.single_pass), "wide_format": bool(self.wide_format),
                "ms_total": self.ms_total, "ms_symbolic": self.ms_symbolic, "ms_numeric": self.ms_numeric,
                "direct": bool(self.direct), "partitioned": bool(self.partitioned)}


class FullExample(C.Structure):
    _fields_ = [("row", C.c_int32), ("column", C.c_int32), ("accepted", C.c_int32),
                ("support_mismatch", C.c_int32), ("contributions", C.c_int64), ("actual", C.c_double),
                ("serial_expected", C.c_double), ("expected", C.c_longdouble), ("abs_error", C.c_longdouble),
                ("error_bound", C.c_longdouble), ("error_ratio", C.c_longdouble)]


class FullReport(C.Structure):
    _fields_ = [("environment_ok", C.c_int32), ("invalid_matrix", C.c_int32), ("passed", C.c_int32),
                ("reference_digits", C.c_int32), ("worst_row", C.c_int32), ("worst_column", C.c_int32),
                ("n_examples", C.c_int32), ("reserved_", C.c_int32),
                ("rows", C.c_int64), ("reference_nnz", C.c_int64), ("row_nnz_mismatches", C.c_int64),
                ("support_mismatches", C.c_int64), ("rejected_coordinates", C.c_int64),
                ("max_abs_error", C.c_double), ("max_reference_error", C.c_longdouble),
                ("max_error_ratio", C.c_longdouble), ("worst_actual", C.c_longdouble),
                ("worst_reference", C.c_longdouble), ("worst_bound", C.c_longdouble),
                ("examples", FullExample * 8)]


class SpgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.name = STATUS.get(status, str(status))


_lib = None


def lib():
    """Load libspgemm_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libspgemm_amd.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, P = C.c_void_p, C.POINTER
        L.spg_create.argtypes = [P(vp), C.c_int]
        L.spg_destroy.argtypes = [vp]
        L.spg_status_string.argtypes = [C.c_int]
        L.spg_status_string.restype = C.c_char_p
        L.spg_last_error.argtypes = [vp]
        L.spg_last_error.restype = C.c_char_p
        L.spg_set_stream.argtypes = [vp, vp]
        L.spg_set_timing.argtypes = [vp, C.c_int]
        L.spg_last_stats.argtypes = [vp, P(Stats)]
        L.spg_multiply_begin.argtypes = [vp, P(CsrView), P(CsrView), C.c_int32, P(HaloView), P(C.c_int64)]
        L.spg_multiply_finish.argtypes = [vp, vp, vp, vp]
        L.spg_result_device.argtypes = [vp, P(CsrView)]
        L.spg_accumulator_create.argtypes = [vp, C.c_int32, C.c_int32, P(vp)]
        L.spg_accumulator_add.argtypes = [vp, P(CsrView), P(CsrView)]
        L.spg_accumulator_finish.argtypes = [vp, P(C.c_int64)]
        L.spg_accumulator_copy.argtypes = [vp, vp, vp, vp]
        L.spg_accumulator_destroy.argtypes = [vp]
        L.spg_matrix_read_mm.argtypes = [C.c_char_p, P(vp)]
        L.spg_matrix_view.argtypes = [vp, P(CsrView)]
        L.spg_matrix_free.argtypes = [vp]
        L.spg_matrix_last_error.restype = C.c_char_p
        L.spg_matrix_upload.argtypes = [vp, P(CsrView), P(vp)]
        L.spg_device_matrix_view.argtypes = [vp, P(CsrView)]
        L.spg_device_matrix_free.argtypes = [vp]
        L.spg_synchronize.argtypes = [vp]
        L.spg_set_path.argtypes = [vp, C.c_int]
        L.spg_set_format.argtypes = [vp, C.c_int]
        L.spg_set_large.argtypes = [vp, C.c_int]
        L.spg_multiply_noalloc.argtypes = [vp, P(CsrView), P(CsrView), vp, vp, vp, C.c_int64, P(C.c_int64)]
        L.spg_result_rows.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, C.c_int64, P(C.c_int64)]
        L.spg_batch_create.argtypes = [vp, C.c_int32, P(CsrView), P(CsrView), P(vp)]
        L.spg_batch_create_chained.argtypes = [vp, vp, C.c_int32, P(C.c_int32), P(CsrView), P(vp)]
        L.spg_batch_refresh.argtypes = [vp]
        L.spg_batch_run.argtypes = [vp, P(C.c_int64)]
        L.spg_batch_count.argtypes = [vp, P(C.c_uint64)]
        L.spg_batch_rebuild.argtypes = [vp, C.c_int32, P(CsrView), P(CsrView)]
        L.spg_batch_device_bytes.argtypes = [vp, P(C.c_uint64)]
        L.spg_batch_trim.argtypes = [vp]
        L.spg_batch_estimate.argtypes = [vp, vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int64,
                                         P(C.c_uint64)]
        L.spg_batch_set_graph.argtypes = [vp, C.c_int]
        L.spg_batch_autotune.argtypes = [vp, C.c_int, P(C.c_int)]
        L.spg_batch_product_nnz.argtypes = [vp, C.c_int32, P(C.c_int64)]
        L.spg_batch_copy.argtypes = [vp, C.c_int32, vp, vp, vp]
        L.spg_batch_info.argtypes = [vp, C.c_int32, P(C.c_int32), P(C.c_int32)]
        L.spg_batch_last_stats.argtypes = [vp, P(Stats)]
        L.spg_batch_destroy.argtypes = [vp]
        L.spg_batch_last_error.argtypes = [vp]
        L.spg_batch_last_error.restype = C.c_char_p
        L.spg_server_create.argtypes = [C.c_int, C.c_uint64, C.c_int, P(vp)]
        L.spg_server_submit.argtypes = [vp, P(CsrView), P(CsrView), P(C.c_int64)]
        L.spg_server_register.argtypes = [vp, P(CsrView), P(C.c_int64)]
        L.spg_server_unregister.argtypes = [vp, C.c_int64]
        L.spg_server_submit_resident.argtypes = [vp, C.c_int64, C.c_int64, P(C.c_int64)]
        L.spg_server_wait.argtypes = [vp, C.c_int64, P(C.c_int64), P(C.c_double)]
        L.spg_server_result.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.spg_server_release.argtypes = [vp, C.c_int64]
        L.spg_server_stats.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_uint64)]
        L.spg_server_destroy.argtypes = [vp]
        L.spg_multiply_rows.argtypes = [vp, P(CsrView), P(CsrView), C.c_int32, P(HaloView), C.c_int32, C.c_int32,
                                        P(C.c_int64)]
        L.spg_matrix_write_csr.argtypes = [P(CsrView), C.c_char_p]
        L.spg_matrix_read_csr.argtypes = [C.c_char_p, P(vp)]
        L.spg_halo_upload.argtypes = [vp, P(HaloView), P(vp)]
        L.spg_halo_pack.argtypes = [vp, P(CsrView), C.c_int32, vp, C.c_int32, C.c_int32, P(vp)]
        L.spg_device_halo_view.argtypes = [vp, P(HaloView)]
        L.spg_device_halo_free.argtypes = [vp]
        cp = C.c_char_p
        L.spg_validation_last_error.restype = cp
        L.spg_validate_full.argtypes = [P(CsrView), P(CsrView), P(CsrView), cp, cp, C.c_int, P(FullReport)]
        L.spg_full_report_text.argtypes = [P(FullReport), vp, C.c_size_t, P(C.c_size_t)]
        L.spg_partial_counts.argtypes = [P(CsrView), P(CsrView), C.c_int, vp, P(C.c_int32)]
        L.spg_validate_partial.argtypes = [P(CsrView), P(CsrView), P(CsrView), cp, cp, C.c_int, P(C.c_int32),
                                           P(C.c_int64)]
        L.spg_spref_write.argtypes = [cp, cp, C.c_int32, C.c_int32, vp, C.c_int32]
        L.spg_spref_read.argtypes = [cp, cp, C.c_int32, C.c_int32, vp, P(C.c_int32)]
        L.spg_sfull_write.argtypes = [P(CsrView), P(CsrView), cp, cp, C.c_int]
        L.spg_sfull_open.argtypes = [cp, cp, C.c_int32, C.c_int32, P(vp)]
        L.spg_sfull_next_row.argtypes = [vp, P(C.c_int64)]
        L.spg_sfull_row_cells.argtypes = [vp, vp, vp, vp, vp, vp]
        L.spg_sfull_finish.argtypes = [vp]
        L.spg_sfull_close.argtypes = [vp]
        L.spg_sha256_hex.argtypes = [vp, C.c_size_t, C.c_char_p]
        _lib = L
    return _lib


def host_view(m: Csr) -> CsrView:
    """Host view; nnz is the length of the entry arrays (the engine checks it
    against row_ptr[rows])."""
    if len(m.col) != len(m.val):
        raise ValueError("column and value arrays differ in length")
    v = CsrView()
    v.rows, v.cols, v.nnz = m.rows, m.cols, len(m.col)
    v.row_ptr = m.row_ptr.ctypes.data
    v.col_idx = m.col.ctypes.data if len(m.col) else None
    v.values = m.val.ctypes.data if len(m.val) else None
    v.location = SPG_HOST
    return v


def read_mm(path: str) -> Csr:
    """readMatrixMarket through the engine's host reader (parallel parse)."""
    return _read_host_matrix(lib().spg_matrix_read_mm, path)


def read_csr(path: str) -> Csr:
    """A binary CSR file written by write_csr."""
    return _read_host_matrix(lib().spg_matrix_read_csr, path)


def write_csr(m: Csr, path: str):
    L = lib()
    v = host_view(m)
    s = L.spg_matrix_write_csr(C.byref(v), path.encode())
    if s != 0:
        raise SpgError(s, L.spg_matrix_last_error().decode())


def _read_host_matrix(fn, path: str) -> Csr:
    L = lib()
    h = C.c_void_p()
    s = fn(path.encode(), C.byref(h))
    if s != 0:
        raise SpgError(s, L.spg_matrix_last_error().decode())
    try:
        v = CsrView()
        L.spg_matrix_view(h, C.byref(v))
        n = int(v.nnz)
        rp = np.ctypeslib.as_array(C.cast(v.row_ptr, C.POINTER(C.c_int64)), shape=(v.rows + 1,)).copy()
        col = (np.ctypeslib.as_array(C.cast(v.col_idx, C.POINTER(C.c_int32)), shape=(n,)).copy()
               if n else np.zeros(0, np.int32))
        val = (np.ctypeslib.as_array(C.cast(v.values, C.POINTER(C.c_double)), shape=(n,)).copy()
               if n else np.zeros(0, np.float64))
        return Csr(int(v.rows), int(v.cols), rp, col, val)
    finally:
        L.spg_matrix_free(h)


class DeviceMatrix:
    """A CSR matrix resident in HBM, owned by the library (spg_matrix_upload)."""

    def __init__(self, engine: "Engine", m: Csr):
        L = lib()
        h = C.c_void_p()
        engine._check(L.spg_matrix_upload(engine._h, C.byref(host_view(m)), C.byref(h)))
        self._h, self._L = h, L
        self.rows, self.cols, self.nnz = m.rows, m.cols, len(m.col)

    def view(self) -> CsrView:
        v = CsrView()
        self._L.spg_device_matrix_view(self._h, C.byref(v))
        return v

    def free(self):
        if getattr(self, "_h", None):
            self._L.spg_device_matrix_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One spg_context (one device, one stream)."""

    def __init__(self, device: int = 0, timing: bool = False):
        L = lib()
        h = C.c_void_p()
        s = L.spg_create(C.byref(h), device)
        if s != 0:
            raise SpgError(s, "spg_create failed")
        self._h = h
        self._L = L
        self._path = PATH_AUTO
        self._format = FORMAT_AUTO
        if timing:
            L.spg_set_timing(h, 1)

    def close(self):
        if getattr(self, "_h", None):
            self._L.spg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, s: int):
        if s != 0:
            raise SpgError(s, self._L.spg_last_error(self._h).decode(errors="replace"))

    def set_stream(self, stream_handle: int | None):
        self._check(self._L.spg_set_stream(self._h, stream_handle))

    def set_timing(self, on: bool):
        self._check(self._L.spg_set_timing(self._h, 1 if on else 0))

    def set_path(self, path: int):
        """PATH_AUTO, PATH_SINGLE_PASS, PATH_TWO_PHASE or PATH_SINGLE_PASS_DEALT (results are identical)."""
        self._check(self._L.spg_set_path(self._h, int(path)))
        self._path = int(path)

    def set_format(self, fmt: int):
        """FORMAT_AUTO or FORMAT_WIDE (64-bit entry stream and B addressing
        whatever B's size; results are identical)."""
        self._check(self._L.spg_set_format(self._h, int(fmt)))
        self._format = int(fmt)

    def set_large(self, mode: int):
        """LARGE_AUTO or LARGE_WINDOWS (rows above 2048 products never
        partitioned by column window; results are identical)."""
        self._check(self._L.spg_set_large(self._h, int(mode)))

    def stats(self) -> dict:
        st = Stats()
        self._check(self._L.spg_last_stats(self._h, C.byref(st)))
        return st.as_dict()

    def begin(self, a_view: CsrView, b_view: CsrView, b_first_row: int = 0, halo: HaloView | None = None) -> int:
        n = C.c_int64()
        self._check(self._L.spg_multiply_begin(self._h, C.byref(a_view), C.byref(b_view), b_first_row,
                                               C.byref(halo) if halo is not None else None, C.byref(n)))
        return int(n.value)

    def finish(self, rows: int, cols: int, nnz: int) -> Csr:
        rp = np.zeros(rows + 1, dtype=np.int64)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        self._check(self._L.spg_multiply_finish(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return Csr(rows, cols, rp, col[:nnz], val[:nnz])

    def multiply(self, a: Csr, b: Csr, b_first_row: int = 0, halo=None) -> Csr:
        """Drop-in local multiply on host CSR matrices (spgemm_exchange.cpp:258-325).

        halo: None or (global_rows int32[n], row_ptr int64[n+1], col int32, val float64).
        """
        hv = None
        keep = []
        if halo is not None:
            g, rp, col, val = (np.ascontiguousarray(halo[0], np.int32), np.ascontiguousarray(halo[1], np.int64),
                               np.ascontiguousarray(halo[2], np.int32), np.ascontiguousarray(halo[3], np.float64))
            keep = [g, rp, col, val]
            hv = HaloView()
            hv.n_rows = len(g)
            hv.global_rows = g.ctypes.data if len(g) else None
            hv.row_ptr = rp.ctypes.data
            hv.col_idx = col.ctypes.data if len(col) else None
            hv.values = val.ctypes.data if len(val) else None
            hv.location = SPG_HOST
        nnz = self.begin(host_view(a), host_view(b), b_first_row, hv)
        del keep
        return self.finish(a.rows, b.cols, nnz)

    def upload(self, m: Csr) -> DeviceMatrix:
        return DeviceMatrix(self, m)

    def multiply_halo_device(self, a: DeviceMatrix, b_local: DeviceMatrix, b_first_row: int,
                             halo: "DeviceHalo | None") -> int:
        """Product with a device-resident local B block and device halo rows
        (directory built on the device); the result stays on the device."""
        hv = halo.view() if halo is not None else None
        return self.begin(a.view(), b_local.view(), b_first_row, hv)

    def synchronize(self):
        self._check(self._L.spg_synchronize(self._h))

    def multiply_device(self, a: DeviceMatrix, b: DeviceMatrix, rows: tuple[int, int] | None = None) -> int:
        """Product of HBM-resident operands (or C rows [r0, r1) of it); the
        result stays on the device."""
        if rows is None:
            return self.begin(a.view(), b.view())
        n = C.c_int64()
        av, bv = a.view(), b.view()
        self._check(self._L.spg_multiply_rows(self._h, C.byref(av), C.byref(bv), 0, None, int(rows[0]),
                                              int(rows[1]), C.byref(n)))
        return int(n.value)

    def multiply_noalloc(self, a: DeviceMatrix, b: DeviceMatrix, row_ptr: int, col: int, val: int,
                         capacity: int) -> tuple[int, int]:
        """spg_multiply_noalloc into caller-owned device buffers (addresses);
        returns (status, nnz) without raising, so that the capacity query
        (SPG_ENOMEM with the symbolic bound) can be read."""
        n = C.c_int64()
        av, bv = a.view(), b.view()
        s = self._L.spg_multiply_noalloc(self._h, C.byref(av), C.byref(bv), row_ptr, col, val, int(capacity),
                                         C.byref(n))
        return int(s), int(n.value)

    def last_error(self) -> str:
        return self._L.spg_last_error(self._h).decode(errors="replace")

    def result_view(self) -> CsrView:
        v = CsrView()
        self._check(self._L.spg_result_device(self._h, C.byref(v)))
        return v

    def result_rows(self, r0: int, r1: int) -> Csr:
        """Rows [r0, r1) of the last result (row_ptr starting at 0)."""
        rp = np.zeros(r1 - r0 + 1, dtype=np.int64)
        n = C.c_int64()
        s = self._L.spg_result_rows(self._h, r0, r1, rp.ctypes.data, None, None, 0, C.byref(n))
        cnt = int(n.value)
        if s != 0 and cnt == 0:
            self._check(s)
        col = np.zeros(max(cnt, 1), dtype=np.int32)
        val = np.zeros(max(cnt, 1), dtype=np.float64)
        self._check(self._L.spg_result_rows(self._h, r0, r1, rp.ctypes.data, col.ctypes.data, val.ctypes.data, cnt,
                                            C.byref(n)))
        v = self.result_view()
        return Csr(r1 - r0, v.cols, rp, col[:cnt], val[:cnt])

    def result_row_ptr(self) -> np.ndarray:
        """Row offsets of the last result (rows + 1 int64), without its entries."""
        v = self.result_view()
        out = np.empty(v.rows + 1, np.int64)
        hip = C.CDLL("libamdhip64.so")
        self.synchronize()
        rc = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(v.row_ptr), C.c_size_t(out.nbytes), 2)
        if rc != 0:
            raise SpgError(6, f"hipMemcpy of the result's row offsets failed ({rc})")
        return out

    def result_host(self) -> Csr:
        v = self.result_view()
        return self.finish(v.rows, v.cols, v.nnz)


class DeviceHalo:
    """Halo rows in HBM owned by the library: uploaded host rows
    (spg_halo_upload) or requested rows packed from a device B block
    (spg_halo_pack)."""

    def __init__(self, engine: "Engine", handle):
        self._e, self._L, self._h = engine, engine._L, handle

    @classmethod
    def upload(cls, engine: "Engine", global_rows, row_ptr, col, val) -> "DeviceHalo":
        g, rp = np.ascontiguousarray(global_rows, np.int32), np.ascontiguousarray(row_ptr, np.int64)
        cl, vl = np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64)
        hv = HaloView()
        hv.n_rows = len(g)
        hv.global_rows = g.ctypes.data if len(g) else None
        hv.row_ptr = rp.ctypes.data
        hv.col_idx = cl.ctypes.data if len(cl) else None
        hv.values = vl.ctypes.data if len(vl) else None
        hv.location = SPG_HOST
        h = C.c_void_p()
        engine._check(engine._L.spg_halo_upload(engine._h, C.byref(hv), C.byref(h)))
        return cls(engine, h)

    @classmethod
    def pack(cls, engine: "Engine", b_local: "DeviceMatrix", first_row: int, requests) -> "DeviceHalo":
        req = np.ascontiguousarray(requests, np.int32)
        h = C.c_void_p()
        bv = b_local.view()
        engine._check(engine._L.spg_halo_pack(engine._h, C.byref(bv), int(first_row),
                                              req.ctypes.data if len(req) else None, len(req), SPG_HOST,
                                              C.byref(h)))
        return cls(engine, h)

    def view(self) -> HaloView:
        v = HaloView()
        self._L.spg_device_halo_view(self._h, C.byref(v))
        return v

    def host(self):
        """(global_rows, row_ptr, col, val) copied back to the host."""
        import ctypes
        v = self.view()
        n = int(v.n_rows)
        hip = ctypes.CDLL("libamdhip64.so")
        def get(ptr, count, dtype):
            out = np.zeros(max(count, 1), dtype)
            if count:
                hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(count * out.itemsize), 2)
            return out[:count]
        rp = get(v.row_ptr, n + 1, np.int64)
        nnz = int(rp[-1]) if n else 0
        return get(v.global_rows, n, np.int32), rp, get(v.col_idx, nnz, np.int32), get(v.values, nnz, np.float64)

    def free(self):
        if self._h:
            self._L.spg_device_halo_free(self._h)
            self._h = None


class Batch:
    """Many independent products computed by one device pipeline
    (spg_batch_*).  Each result is bit-identical to the standalone product."""

    def __init__(self, engine: "Engine", a_list=None, b_list=None, _handle=None, _keep=None):
        self._e = engine
        self._L = engine._L
        if _handle is not None:
            self._h = _handle
            self._keep = _keep
            return
        n = len(a_list)
        av = (CsrView * max(n, 1))()
        bv = (CsrView * max(n, 1))()
        for i, (a, b) in enumerate(zip(a_list, b_list)):
            av[i] = host_view(a)
            bv[i] = host_view(b)
        h = C.c_void_p()
        engine._check(self._L.spg_batch_create(engine._h, n, av, bv, C.byref(h)))
        self._h = h
        self._keep = None
        self.n = n

    @classmethod
    def chained(cls, engine: "Engine", src: "Batch", src_index, b_list) -> "Batch":
        """A_i = result src_index[i] of batch src (which must have run)."""
        n = len(b_list)
        idx = (C.c_int32 * max(n, 1))(*[int(x) for x in src_index])
        bv = (CsrView * max(n, 1))()
        for i, b in enumerate(b_list):
            bv[i] = host_view(b)
        h = C.c_void_p()
        engine._check(engine._L.spg_batch_create_chained(engine._h, src._h, n, idx, bv, C.byref(h)))
        out = cls(engine, _handle=h, _keep=[src])
        out.n = n
        return out

    def _bcheck(self, s: int):
        if s != 0:
            raise SpgError(s, self._L.spg_batch_last_error(self._h).decode(errors="replace"))

    def refresh(self):
        self._e._check(self._L.spg_batch_refresh(self._h))

    def run(self) -> int:
        n = C.c_int64()
        self._bcheck(self._L.spg_batch_run(self._h, C.byref(n)))
        return int(n.value)

    def count(self) -> int:
        """Run the count phase only; returns the device bytes the value phase
        will allocate (the next run() continues from these counts)."""
        n = C.c_uint64()
        self._bcheck(self._L.spg_batch_count(self._h, C.byref(n)))
        return int(n.value)

    def rebuild(self, a_list, b_list):
        """Replace the products, reusing the batch's device buffers."""
        n = len(a_list)
        av = (CsrView * max(n, 1))()
        bv = (CsrView * max(n, 1))()
        for i, (a, b) in enumerate(zip(a_list, b_list)):
            av[i] = host_view(a)
            bv[i] = host_view(b)
        self._bcheck(self._L.spg_batch_rebuild(self._h, n, av, bv))
        self.n = n

    def device_bytes(self) -> int:
        n = C.c_uint64()
        self._bcheck(self._L.spg_batch_device_bytes(self._h, C.byref(n)))
        return int(n.value)

    def estimate(self, rows: int, a_nnz: int, b_rows: int, b_nnz: int, cols: int, c_entries: int) -> int:
        """Device bytes rebuilding this batch with products of these total
        sizes and running it with C's capacity at c_entries would add."""
        n = C.c_uint64()
        self._e._check(self._L.spg_batch_estimate(self._e._h, self._h, rows, a_nnz, b_rows, b_nnz, cols, c_entries,
                                                  C.byref(n)))
        return int(n.value)

    def trim(self):
        """Free the batch's device buffers (it holds no products afterwards)."""
        self._bcheck(self._L.spg_batch_trim(self._h))
        self.n = 0

    def set_graph(self, on: bool):
        self._e._check(self._L.spg_batch_set_graph(self._h, 1 if on else 0))

    def autotune(self, repeats: int = 2) -> int:
        """Time the pipeline variants and keep the fastest; returns its index:
        0 single pass (2 rows per wave), 1 single pass (1 row per wave),
        2 two-phase, 3 dealt 32-row tiles, 4 packed tiles of small rows."""
        c = C.c_int()
        self._bcheck(self._L.spg_batch_autotune(self._h, repeats, C.byref(c)))
        return int(c.value)

    def product_nnz(self, i: int) -> int:
        n = C.c_int64()
        self._bcheck(self._L.spg_batch_product_nnz(self._h, i, C.byref(n)))
        return int(n.value)

    def result(self, i: int) -> Csr:
        r, c = C.c_int32(), C.c_int32()
        self._e._check(self._L.spg_batch_info(self._h, i, C.byref(r), C.byref(c)))
        nnz = self.product_nnz(i)
        rp = np.zeros(r.value + 1, dtype=np.int64)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        self._bcheck(self._L.spg_batch_copy(self._h, i, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return Csr(int(r.value), int(c.value), rp, col[:nnz], val[:nnz])

    def stats(self) -> dict:
        st = Stats()
        self._e._check(self._L.spg_batch_last_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self._L.spg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Served products with continuous batching (spg_server_*).  submit() and
    submit_resident() return request ids; wait() blocks until the product is
    computed; result() copies it out; release() frees it."""

    def __init__(self, device: int = 0, budget_bytes: int = 200 << 30, max_batch: int = 4096):
        self._L = lib()
        h = C.c_void_p()
        s = self._L.spg_server_create(device, budget_bytes, max_batch, C.byref(h))
        if s != 0:
            raise SpgError(s, "spg_server_create failed")
        self._h = h
        self._shape = {}
        self._rshape = {}

    def _check(self, s: int, what: str):
        if s != 0:
            raise SpgError(s, what)

    def submit(self, a: Csr, b: Csr) -> int:
        rid = C.c_int64()
        av, bv = host_view(a), host_view(b)
        self._check(self._L.spg_server_submit(self._h, C.byref(av), C.byref(bv), C.byref(rid)), "submit")
        self._shape[rid.value] = (a.rows, b.cols)
        return int(rid.value)

    def register(self, m: Csr) -> int:
        h = C.c_int64()
        v = host_view(m)
        self._check(self._L.spg_server_register(self._h, C.byref(v), C.byref(h)), "register")
        self._rshape[h.value] = (m.rows, m.cols)
        return int(h.value)

    def unregister(self, handle: int):
        self._check(self._L.spg_server_unregister(self._h, handle), "unregister")

    def submit_resident(self, ha: int, hb: int) -> int:
        rid = C.c_int64()
        self._check(self._L.spg_server_submit_resident(self._h, ha, hb, C.byref(rid)), "submit_resident")
        self._shape[rid.value] = (self._rshape[ha][0], self._rshape[hb][1])
        return int(rid.value)

    def wait(self, rid: int):
        """(nnz, latency seconds) of a completed request."""
        n, lat = C.c_int64(), C.c_double()
        self._check(self._L.spg_server_wait(self._h, rid, C.byref(n), C.byref(lat)), f"request {rid}")
        return int(n.value), float(lat.value)

    def result(self, rid: int) -> Csr:
        nnz, _ = self.wait(rid)
        rows, cols = self._shape[rid]
        rp = np.zeros(rows + 1, dtype=np.int64)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        self._check(self._L.spg_server_result(self._h, rid, rp.ctypes.data, col.ctypes.data, val.ctypes.data),
                    "result")
        return Csr(rows, cols, rp, col[:nnz], val[:nnz])

    def release(self, rid: int):
        self._check(self._L.spg_server_release(self._h, rid), "release")
        self._shape.pop(rid, None)

    def stats(self) -> dict:
        b, c, busy, ch = C.c_int64(), C.c_int64(), C.c_double(), C.c_uint64()
        self._check(self._L.spg_server_stats(self._h, C.byref(b), C.byref(c), C.byref(busy), C.byref(ch)), "stats")
        return {"batches": b.value, "completed": c.value, "busy_s": busy.value, "charged_bytes": ch.value}

    def close(self):
        if getattr(self, "_h", None):
            self._L.spg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accumulator:
    """Staged accumulation on the device (trident::LocalAccumulator,
    src/trident/common/matrix_blocks.cpp:152-200): add_product(a, b) adds a*b
    into per-row running sums in call order; finish() drops exact zeros."""

    def __init__(self, engine: "Engine", rows: int, cols: int):
        self._e = engine
        self._L = engine._L
        h = C.c_void_p()
        engine._check(self._L.spg_accumulator_create(engine._h, rows, cols, C.byref(h)))
        self._h = h
        self.rows, self.cols = rows, cols

    def add_product(self, a: Csr, b: Csr):
        av, bv = host_view(a), host_view(b)
        self._e._check(self._L.spg_accumulator_add(self._h, C.byref(av), C.byref(bv)))

    def finish(self) -> Csr:
        n = C.c_int64()
        self._e._check(self._L.spg_accumulator_finish(self._h, C.byref(n)))
        nnz = int(n.value)
        rp = np.zeros(self.rows + 1, dtype=np.int64)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        self._e._check(self._L.spg_accumulator_copy(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return Csr(self.rows, self.cols, rp, col[:nnz], val[:nnz])

    def close(self):
        if getattr(self, "_h", None):
            self._L.spg_accumulator_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- validation (host code; no device needed) ------------------------------

def _vcheck(s: int):
    if s != 0:
        raise SpgError(s, lib().spg_validation_last_error().decode(errors="replace"))


def _enc(x):
    return None if x is None else str(x).encode()


def sha256_hex(data: bytes) -> str:
    out = C.create_string_buffer(65)
    buf = C.create_string_buffer(bytes(data), len(data))
    _vcheck(lib().spg_sha256_hex(buf, len(data), out))
    return out.value.decode()


def validate_full(a: Csr, b: Csr, c: Csr, sfull: str | None = None, key: str | None = None,
                  threads: int = 0) -> dict:
    """Full validation (componentwise_dot_bound_v1); expected rows from `sfull`
    (a .sfull file named by `key`) or rebuilt from a and b."""
    rep = FullReport()
    va, vb, vc = host_view(a), host_view(b), host_view(c)
    _vcheck(lib().spg_validate_full(C.byref(va), C.byref(vb), C.byref(vc), _enc(sfull), _enc(key), threads,
                                    C.byref(rep)))
    need = C.c_size_t()
    lib().spg_full_report_text(C.byref(rep), None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    lib().spg_full_report_text(C.byref(rep), buf, need.value, None)
    out = {f: getattr(rep, f) for f, _ in FullReport._fields_ if f not in ("examples", "reserved_")}
    out["examples"] = [{f: getattr(rep.examples[i], f) for f, _ in FullExample._fields_}
                       for i in range(rep.n_examples)]
    out["text"] = buf.value.decode()
    return out


def partial_counts(a: Csr, b: Csr, threads: int = 0):
    counts = np.zeros(a.rows, np.int64)
    fin = C.c_int32()
    va, vb = host_view(a), host_view(b)
    _vcheck(lib().spg_partial_counts(C.byref(va), C.byref(vb), threads, counts.ctypes.data, C.byref(fin)))
    return counts, bool(fin.value)


def validate_partial(a: Csr, b: Csr, c: Csr, spref: str | None = None, key: str | None = None,
                     threads: int = 0):
    passed, mism = C.c_int32(), C.c_int64()
    va, vb, vc = host_view(a), host_view(b), host_view(c)
    _vcheck(lib().spg_validate_partial(C.byref(va), C.byref(vb), C.byref(vc), _enc(spref), _enc(key), threads,
                                       C.byref(passed), C.byref(mism)))
    return bool(passed.value), int(mism.value)


def spref_write(path: str, key: str, rows: int, cols: int, counts, finite: bool):
    counts = np.ascontiguousarray(counts, dtype=np.int64)
    _vcheck(lib().spg_spref_write(path.encode(), key.encode(), rows, cols, counts.ctypes.data if rows else None,
                                  1 if finite else 0))


def spref_read(path: str, key: str, rows: int, cols: int):
    counts = np.zeros(rows, np.int64)
    fin = C.c_int32()
    _vcheck(lib().spg_spref_read(path.encode(), key.encode(), rows, cols, counts.ctypes.data if rows else None,
                                 C.byref(fin)))
    return counts, bool(fin.value)


def sfull_write(a: Csr, b: Csr, path: str, key: str, threads: int = 0):
    va, vb = host_view(a), host_view(b)
    _vcheck(lib().spg_sfull_write(C.byref(va), C.byref(vb), path.encode(), key.encode(), threads))


class SfullReader:
    """Row-by-row reader of a .sfull file; finish() checks the SHA-256 trailer."""

    def __init__(self, path: str, key: str, rows: int, cols: int):
        self._h = C.c_void_p()
        _vcheck(lib().spg_sfull_open(path.encode(), key.encode(), rows, cols, C.byref(self._h)))

    def row(self) -> dict:
        n = C.c_int64()
        _vcheck(lib().spg_sfull_next_row(self._h, C.byref(n)))
        k = int(n.value)
        col = np.zeros(k, np.int32)
        serial = np.zeros(k, np.float64)
        ref = np.zeros(k, np.longdouble)
        bound = np.zeros(k, np.longdouble)
        terms = np.zeros(k, np.int64)
        _vcheck(lib().spg_sfull_row_cells(self._h, col.ctypes.data, serial.ctypes.data, ref.ctypes.data,
                                          bound.ctypes.data, terms.ctypes.data))
        return {"column": col, "serial": serial, "reference": ref, "bound": bound, "terms": terms}

    def finish(self):
        _vcheck(lib().spg_sfull_finish(self._h))

    def close(self):
        if self._h:
            lib().spg_sfull_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
