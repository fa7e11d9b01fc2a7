"""Helpers shared by the GPU parity tests: conversions between the engine's
and the oracle's CSR types, digests, and row-sampled comparison with the
serial-order oracle for products too large to rebuild whole on the host."""
import ctypes

import numpy as np

import oracle as O
import spgemm_amd as S


def to_s(m) -> S.Csr:
    return S.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val)


def to_o(m) -> O.Csr:
    return O.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val)


def digest(m) -> str:
    return O.csr_digest(to_o(m))


def sorted_rows(c) -> bool:
    """Columns strictly increasing inside every row of a CSR."""
    d = np.diff(c.col.astype(np.int64))
    starts = c.row_ptr[1:-1]
    ok = d > 0
    ok[starts[(starts > 0) & (starts < len(c.col))] - 1] = True
    return bool(np.all(ok))


def sample_rows(a, rows):
    """The CSR made of the given rows of a (in the given order)."""
    rp = [0]
    cols, vals = [], []
    for r in rows:
        s, e = a.row_ptr[r], a.row_ptr[r + 1]
        cols.append(a.col[s:e])
        vals.append(a.val[s:e])
        rp.append(rp[-1] + int(e - s))
    return O.Csr.make(len(rows), a.cols, rp, np.concatenate(cols) if cols else np.zeros(0, np.int32),
                      np.concatenate(vals) if vals else np.zeros(0))


def oracle_rows(a, b, rows, threads=8):
    """Serial-order oracle product of the given rows of a with b."""
    return O.product(sample_rows(to_o(a), rows), to_o(b), threads=threads)


def rows_match(got_col, got_val, ref, i) -> bool:
    """Row i of the oracle CSR ref equals the given columns and value bits."""
    s, e = ref.row_ptr[i], ref.row_ptr[i + 1]
    return (np.array_equal(got_col, ref.col[s:e]) and
            np.array_equal(np.asarray(got_val).view(np.uint64), ref.val[s:e].view(np.uint64)))


def check_sampled_rows(a, b, c, nsample, seed):
    """nsample random rows of the host CSR c equal the oracle's, bit for bit."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rows = np.sort(rng.choice(a.rows, size=min(nsample, a.rows), replace=False))
    sub = oracle_rows(a, b, rows)
    for i, r in enumerate(rows):
        s, e = c.row_ptr[r], c.row_ptr[r + 1]
        assert rows_match(c.col[s:e], c.val[s:e], sub, i), r


def check_device_rows(engine, a, b, rows):
    """Rows of the engine's last (device) result equal the oracle's, fetched
    one row at a time through spg_result_rows."""
    rows = np.array(sorted(set(int(r) for r in rows)))
    sub = oracle_rows(a, b, rows)
    for i, r in enumerate(rows):
        got = engine.result_rows(int(r), int(r) + 1)
        assert rows_match(got.col, got.val, sub, i), r
    return len(rows)


def device_to_host(ptr, count, dtype):
    """A host copy of `count` elements at a device address (hipMemcpy)."""
    out = np.empty(count, dtype)
    if count:
        hip = ctypes.CDLL("libamdhip64.so")
        rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                           2)  # hipMemcpyDeviceToHost
        assert rc == 0, f"hipMemcpy failed ({rc})"
    return out
