"""Parity at the shapes the large workloads run: row-chunked execution
(spg_multiply_rows + spg_result_rows, the route for outputs larger than
device memory), the rectangular R-MAT product, and an output with more than
2^31 entries (int64 C offsets past the int32 range).

Bar: bit-exact.  Chunked results are compared with the whole product and the
serial-order oracle; full-size products, which the oracle cannot rebuild in
seconds, with size-independent properties over every row plus the oracle on
sampled rows (and on the rows where C's offsets cross 2^31)."""
import numpy as np
import pytest

import oracle as O
import spgemm_amd as S
import workloads as W
from parity_util import check_device_rows, device_to_host, digest, sorted_rows, to_o

pytestmark = pytest.mark.gpu


def concat(parts, cols):
    """Row-wise concatenation of host CSR chunks."""
    rp = [np.zeros(1, np.int64)]
    off = 0
    for p in parts:
        rp.append(p.row_ptr[1:] + off)
        off += p.nnz
    return S.Csr.make(sum(p.rows for p in parts), cols, np.concatenate(rp),
                      np.concatenate([p.col for p in parts]), np.concatenate([p.val for p in parts]))


def chunk_bounds_rmat13(rowp):
    """Chunk boundaries for R-MAT scale 13 A.A: one chunk of the heaviest row
    alone, boundaries inside single-pass tiles and analyze blocks (not
    multiples of 16 or 256), and one on each side of a large row whose
    predecessor is not large."""
    n = len(rowp)
    large = rowp > 2048
    h = int(next(r for r in range(n // 2, n) if large[r] and not large[r - 1]))
    cuts = sorted({0, 1, 777, 2305, h, h + 1, 5003, n})
    return list(zip(cuts[:-1], cuts[1:])), h


CHUNK_PATHS = [("auto-narrow", S.PATH_AUTO, S.FORMAT_AUTO),
               ("packed-wide", S.PATH_SINGLE_PASS_PACKED, S.FORMAT_WIDE),
               ("dealt-narrow", S.PATH_SINGLE_PASS_DEALT, S.FORMAT_AUTO),
               ("two_phase-wide", S.PATH_TWO_PHASE, S.FORMAT_WIDE)]


@pytest.mark.parametrize("path,fmt", [p[1:] for p in CHUNK_PATHS], ids=[p[0] for p in CHUNK_PATHS])
def test_row_chunks_equal_whole_product(engine, path, fmt):
    a = W.rmat(13, 32, seed=42)
    rowp = W.row_products(a, a)
    bounds, h = chunk_bounds_rmat13(rowp)
    assert len(bounds) >= 6 and rowp[h] > 2048 and rowp[h - 1] <= 2048
    want = O.csr_digest(O.product(to_o(a), to_o(a), threads=8))
    engine.set_path(path)
    engine.set_format(fmt)
    try:
        da = engine.upload(a)
        whole_nnz = engine.multiply_device(da, da)
        whole = engine.result_host()
        assert whole.nnz == whole_nnz and digest(whole) == want
        parts = []
        for r0, r1 in bounds:
            print("chunk", r0, r1, flush=True)
            n = engine.multiply_device(da, da, (r0, r1))
            st = engine.stats()
            assert st["products"] == int(rowp[r0:r1].sum())
            assert st["wide_format"] == (fmt == S.FORMAT_WIDE)
            # the chunk, fetched in two row ranges split inside it
            mid = (r0 + r1 + 1) // 2
            lo, hi = engine.result_rows(0, mid - r0), engine.result_rows(mid - r0, r1 - r0)
            assert lo.nnz + hi.nnz == n
            parts += [lo, hi]
        got = concat(parts, a.cols)
        assert got.rows == a.rows and got.nnz == whole_nnz
        assert digest(got) == want
        da.free()
    finally:
        engine.set_path(S.PATH_AUTO)
        engine.set_format(S.FORMAT_AUTO)


def test_row_chunks_of_host_views_and_errors(engine):
    # Host operands through spg_multiply_rows: chunk results equal the rows
    # of the whole product; empty and inverted ranges.
    a = W.uniform_random(3000, 2000, 9, seed=3)
    b = W.uniform_random(2000, 5000, 11, seed=4)
    want = O.product(to_o(a), to_o(b))
    da, db = engine.upload(a), engine.upload(b)
    parts = []
    for r0, r1 in ((0, 1000), (1000, 1000), (1000, 2999), (2999, 3000)):
        engine.multiply_device(da, db, (r0, r1))
        parts.append(engine.result_rows(0, r1 - r0))
    assert O.csr_digest(to_o(concat(parts, b.cols))) == O.csr_digest(want)
    for r0, r1 in ((5, 4), (-1, 3), (0, 3001)):
        with pytest.raises(S.SpgError) as e:
            engine.multiply_device(da, db, (r0, r1))
        assert e.value.name == "SPG_EINVAL"
    da.free()
    db.free()


def test_rmat_rect_s14_whole(engine):
    # The rectangular R-MAT case at oracle size: R-MAT scale 14, edge factor
    # 32 (0.57, 0.19, 0.19, 0.05) times B 2^14 x 1024 with 24 uniform columns
    # per row; heavy rows fill whole 1024-column rows of C.
    a = W.rmat(14, 32, seed=42)
    b = W.uniform_random(1 << 14, 1024, 24, seed=43)
    want = O.product(to_o(a), to_o(b), threads=8)
    for path, fmt in ((S.PATH_AUTO, S.FORMAT_AUTO), (S.PATH_SINGLE_PASS_PACKED, S.FORMAT_WIDE)):
        engine.set_path(path)
        engine.set_format(fmt)
        try:
            c = engine.multiply(a, b)
        finally:
            engine.set_path(S.PATH_AUTO)
            engine.set_format(S.FORMAT_AUTO)
        st = engine.stats()
        assert st["rows_per_bin"][6] > 0 and st["products"] == W.products(a, b)
        assert digest(c) == O.csr_digest(want)


def check_rows_by_chunks(engine, nrows, cols, span=1 << 16):
    """Properties of every row of the engine's last result, fetched in row
    chunks: offsets nondecreasing from 0 to nnz, columns strictly increasing
    inside rows and within [0, cols), values finite and nonzero."""
    total = 0
    for r0 in range(0, nrows, span):
        r1 = min(nrows, r0 + span)
        part = engine.result_rows(r0, r1)
        assert part.row_ptr[0] == 0 and np.all(np.diff(part.row_ptr) >= 0)
        if part.nnz:
            assert sorted_rows(part)
            assert part.col.min() >= 0 and part.col.max() < cols
            assert np.all(np.isfinite(part.val)) and np.all(part.val != 0)
        total += part.nnz
    return total


def test_rmat_rect_s20_properties(engine):
    # Full size: R-MAT scale 20 times B 2^20 x 65536 with 24 uniform columns
    # per row (seeds 42, 43).  Every row is checked for the CSR properties;
    # 512 sampled rows and the 4 heaviest are bit-exact against the oracle.
    a = W.rmat(20, 32, seed=42)
    b = W.uniform_random(1 << 20, 65536, 24, seed=43)
    rowp = W.row_products(a, b)
    da, db = engine.upload(a), engine.upload(b)
    nnz = engine.multiply_device(da, db)
    st = engine.stats()
    assert st["products"] == int(rowp.sum())
    assert st["rows_per_bin"][6] > 0 and nnz <= st["structural_nnz"] <= st["products"]
    assert check_rows_by_chunks(engine, a.rows, b.cols) == nnz
    rng = np.random.Generator(np.random.PCG64(17))
    rows = set(rng.choice(a.rows, size=512, replace=False).tolist())
    rows.update(int(r) for r in np.argsort(rowp)[-4:])
    print("checking", len(rows), "rows against the oracle", flush=True)
    assert check_device_rows(engine, a, b, rows) == len(rows)
    da.free()
    db.free()


def test_output_beyond_2p31_entries(engine):
    # 9,000,000 rows with 16 uniform columns each times B 2^20 x 2^24 with 16
    # per row: 2.304e9 scalar products and about as many output entries, so
    # C's row offsets pass 2^31 and its entry indices need 64 bits.  Every
    # row offset is checked, and the rows around the 2^31 crossing, the last
    # rows and 256 sampled rows are bit-exact against the oracle.
    rows = 9_000_000
    a = W.uniform_random(rows, 1 << 20, 16, seed=61)
    b = W.uniform_random(1 << 20, 1 << 24, 16, seed=62)
    da, db = engine.upload(a), engine.upload(b)
    nnz = engine.multiply_device(da, db)
    st = engine.stats()
    assert st["products"] == rows * 256
    assert nnz > (1 << 31) and nnz == st["nnz"] <= st["structural_nnz"] <= st["products"]
    v = engine.result_view()
    assert v.nnz == nnz and v.rows == rows
    offs = device_to_host(v.row_ptr, rows + 1, np.int64)  # every row offset
    assert offs[0] == 0 and offs[-1] == nnz
    assert np.all(np.diff(offs) >= 16) and np.all(np.diff(offs) <= 256)
    cross = int(np.searchsorted(offs, 1 << 31, side="right")) - 1
    assert offs[cross] <= (1 << 31) < offs[cross + 1]
    rng = np.random.Generator(np.random.PCG64(29))
    check = set(range(cross - 3, cross + 4)) | set(range(rows - 3, rows))
    check |= set(rng.choice(rows, size=256, replace=False).tolist())
    print("crossing row", cross, "checking", len(check), "rows", flush=True)
    assert check_device_rows(engine, a, b, check) == len(check)
    da.free()
    db.free()
