"""Parity of the HIP engine with the reference's serial order, through the C ABI.

Bar: bit-exact.  Every product's (row_ptr, col, value bits) digest must equal
the digest minted from the reference binary (golden.json), and products the
oracle computes in seconds are compared with the oracle on the same seeded
inputs.  At full workload sizes the checks are size-independent properties
plus oracle comparison on sampled rows.
"""
import math
import os

import numpy as np
import pytest

import cases
import oracle as O
import spgemm_amd as S
import workloads as W
from parity_util import check_sampled_rows, digest, sorted_rows, to_o, to_s

pytestmark = pytest.mark.gpu


def all_cases():
    out = list(cases.mm_cases()) + cases.synthetic_cases()
    out += [(f"smoke_{s}",) + cases.smoke_case(s) for s in range(100)]
    out += [(f"dyadic_{i}", a, b) for i, (a, b) in enumerate(cases.dyadic_cases())]
    return out


PATHS = [("auto", S.PATH_AUTO), ("single_pass", S.PATH_SINGLE_PASS), ("two_phase", S.PATH_TWO_PHASE),
         ("single_pass_dealt", S.PATH_SINGLE_PASS_DEALT), ("single_pass_packed", S.PATH_SINGLE_PASS_PACKED)]
# Every pipeline in both encodings of the entry stream: the 32-bit one that
# small B operands take, and the 64-bit one of B with 2^29 entries or more
# (big products and batches), forced here on small inputs.
FORMATS = [("narrow", S.FORMAT_AUTO), ("wide", S.FORMAT_WIDE)]
PATH_FORMATS = [(f"{p}-{f}", pv, fv) for p, pv in PATHS for f, fv in FORMATS]


def pin(engine, path, fmt):
    engine.set_path(path)
    engine.set_format(fmt)


def unpin(engine):
    engine.set_path(S.PATH_AUTO)
    engine.set_format(S.FORMAT_AUTO)


@pytest.fixture(params=PATH_FORMATS, ids=[p[0] for p in PATH_FORMATS])
def path_engine(request, engine):
    """The session engine pinned to one pipeline and one entry-stream format
    (results must not depend on either)."""
    pin(engine, request.param[1], request.param[2])
    yield engine
    unpin(engine)


def test_golden_products_bit_exact(path_engine, golden):
    engine = path_engine
    for name, a, b in all_cases():
        print("case", name, flush=True)
        c = engine.multiply(to_s(a), to_s(b))
        g = golden["products"][name]
        st = engine.stats()
        assert st["products"] == g["products"], name
        assert c.nnz == g["c_nnz"], name
        assert digest(c) == g["digest"], name
        assert st["wide_format"] == (engine._format == S.FORMAT_WIDE), name
        # AUTO skips the per-row analysis exactly when the longest rows of A
        # and B prove every row single-pass (and A has at most 2^20 entries)
        ma = int(np.diff(a.row_ptr).max()) if a.rows else 0
        mb = int(np.diff(b.row_ptr).max()) if b.rows else 0
        want_direct = (engine._path == S.PATH_AUTO and a.rows > 0 and b.rows > 0 and ma <= 256
                       and ma * mb <= 256 and a.nnz <= (1 << 20))
        assert st["direct"] == want_direct, name
        if a.nnz:
            # small products take the single pass unless two-phase is pinned
            assert st["single_pass"] == (engine._path != S.PATH_TWO_PHASE), name


def block_edge_case(rows, small_pattern, seed):
    # Rows of A whose product is small (<= 256 products: single pass) or
    # bigger (counted ahead) by small_pattern(r); some rows empty.
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = np.array([0 if r % 23 == 5 else (int(rng.integers(1, 9)) if small_pattern(r) else int(rng.integers(30, 60)))
                     for r in range(rows)], dtype=np.int64)
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    a = O.Csr.make(rows, 700, rp, rng.integers(0, 700, size=int(rp[-1])), 2.0 * rng.random(int(rp[-1])) - 1.0)
    b = O.Csr.make(700, 900, np.arange(701, dtype=np.int64) * 20, rng.integers(0, 900, size=14000),
                   2.0 * rng.random(14000) - 1.0)
    return a, b


EDGE_SHAPES = [
    ("one_row", 1, lambda r: True),
    ("block_all_small", 256, lambda r: True),
    ("block_plus_one", 257, lambda r: r % 3 != 0),
    ("sixteen_per_block", 513, lambda r: r % 16 == 0 or r % 16 == 1),  # 32 small rows per block
    ("seventeen_small", 300, lambda r: r < 17 or r > 290),
    ("no_small_rows", 260, lambda r: False),
    ("alternating", 1000, lambda r: (r // 7) % 2 == 0),
]


@pytest.mark.parametrize("path,fmt", [p[1:] for p in PATH_FORMATS], ids=[p[0] for p in PATH_FORMATS])
def test_single_pass_tile_edges(engine, path, fmt):
    # Block and tile boundaries of every single-pass tiling (8-row tiles,
    # dealt 32-row tiles, packed tiles per 256-row block) against the oracle.
    pin(engine, path, fmt)
    try:
        for i, (name, rows, pat) in enumerate(EDGE_SHAPES):
            print("case", name, flush=True)
            a, b = block_edge_case(rows, pat, 40 + i)
            c = engine.multiply(to_s(a), to_s(b))
            assert digest(c) == O.csr_digest(O.product(a, b)), name
    finally:
        unpin(engine)



BIN_CAPS = [(128, 1), (256, 2), (512, 3), (1024, 4), (2048, 5)]  # (products and A entries, bin)


def expected_bin(p, na):
    if p == 0:
        return 0
    for cap, b in BIN_CAPS:
        if p <= cap and na <= cap:
            return b
    return 6


def bin_edge_case(seed):
    # Rows of exactly 256/257, 512/513, 1024/1025, 2048/2049 products (and
    # neighbours), and rows whose A-entry count alone decides the bin: B rows
    # of 64, 16 and 1 entries make any product count reachable.
    rng = np.random.Generator(np.random.PCG64(seed))
    lens_b = np.array([64] * 100 + [16] * 100 + [1] * 100, dtype=np.int64)
    bp = np.zeros(301, dtype=np.int64)
    np.cumsum(lens_b, out=bp[1:])
    ncols = 3000
    b = O.Csr.make(300, ncols, bp, rng.integers(0, ncols, size=int(bp[-1])), 2.0 * rng.random(int(bp[-1])) - 1.0)
    rows = []
    for target in (255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 2047, 2048, 2049):
        for _ in range(3):
            q, r = divmod(target, 64)
            ent = list(rng.integers(0, 100, size=q)) + list(rng.integers(100, 200, size=r // 16)) + \
                list(rng.integers(200, 300, size=r % 16))
            rng.shuffle(ent)
            rows.append(ent)
    for na in (300, 600, 1100, 2100):  # products = A entries (B rows of one entry)
        rows.append(list(rng.integers(200, 300, size=na)))
    rows.insert(5, [])
    rows.insert(17, list(rng.integers(100, 200, size=3)))
    rp = np.zeros(len(rows) + 1, dtype=np.int64)
    np.cumsum([len(e) for e in rows], out=rp[1:])
    col = np.array([c for e in rows for c in e], dtype=np.int64)
    a = O.Csr.make(len(rows), 300, rp, col, 2.0 * rng.random(len(col)) - 1.0)
    p = np.array([int(lens_b[e].sum()) if e else 0 for e in rows])
    want = np.bincount([expected_bin(int(pi), len(e)) for pi, e in zip(p, rows)], minlength=7)
    return a, b, want


@pytest.mark.parametrize("path,fmt", [p[1:] for p in PATH_FORMATS], ids=[p[0] for p in PATH_FORMATS])
def test_bin_edges(engine, path, fmt):
    # Every bin boundary (by products and by A entries) on every pipeline and
    # entry-stream format, bit for bit against the oracle.
    pin(engine, path, fmt)
    try:
        a, b, want = bin_edge_case(77)
        c = engine.multiply(to_s(a), to_s(b))
        st = engine.stats()
        assert list(st["rows_per_bin"]) == list(want)
        assert all(x > 0 for x in want[1:])  # every bin is exercised
        assert st["wide_format"] == (fmt == S.FORMAT_WIDE)
        assert digest(c) == O.csr_digest(O.product(a, b))
    finally:
        unpin(engine)

def test_validators_accept_engine_output(engine):
    for name, a, b in list(cases.mm_cases()) + cases.synthetic_cases():
        if name == "overflow_product":
            continue
        print("case", name, flush=True)
        c = to_o(engine.multiply(to_s(a), to_s(b)))
        ok, mism = O.partial_validate(a, b, c)
        assert ok and mism == 0, name
        rep = O.full_validate(a, b, c)
        assert rep["passed"] and rep["rejected"] == 0 and rep["max_abs_error"] == 0.0, name
        assert rep["max_entry_rel_error"] == 0.0 and rep["frobenius_rel_error"] == 0.0


def test_overflow_product_keeps_infinities(engine, golden):
    # tests/check_partial_validation.cmake: the overflow case fails validation
    a = O.mm_read(cases.data_path("overflow_product.mtx"))
    c = to_o(engine.multiply(to_s(a), to_s(a)))
    assert np.all(np.isinf(c.val)) and c.nnz == 2
    assert not O.partial_validate(a, a, c)[0]
    assert not O.full_validate(a, a, c)["passed"]


def csr(rows, cols, rp, col, val):
    return S.Csr.make(rows, cols, rp, col, val)


def test_special_values(path_engine):
    engine = path_engine
    unit = csr(1, 1, [0, 1], [0], [1.0])
    # 1e-200^2 underflows to zero and is dropped (spgemm_validation_smoke.cpp:74-75)
    t = csr(1, 1, [0, 1], [0], [1e-200])
    c = engine.multiply(t, t)
    assert c.nnz == 0 and c.row_ptr.tolist() == [0, 0]
    st = engine.stats()
    assert st["compacted"] == (not st["single_pass"])  # two-phase counted the zero
    # subnormal operands survive (spgemm_full_streaming.cpp:108-110)
    sub = csr(1, 1, [0, 1], [0], [5e-324])
    c = engine.multiply(sub, unit)
    assert c.nnz == 1 and c.val.view(np.uint64)[0] == 1
    # NaN and Inf propagate; the validators reject them, the multiply does not
    for v in (math.nan, math.inf, -math.inf):
        c = engine.multiply(csr(1, 1, [0, 1], [0], [v]), unit)
        assert c.nnz == 1 and not math.isfinite(c.val[0])
    # order-dependent cancellation: {1e16,-1e16,1} keeps 1, {1e16,1,-1e16} drops all
    oa = csr(1, 1, [0, 1], [0], [1.0])
    assert engine.multiply(oa, csr(1, 1, [0, 3], [0, 0, 0], [1e16, -1e16, 1])).val.tolist() == [1.0]
    assert engine.multiply(oa, csr(1, 1, [0, 3], [0, 0, 0], [1e16, 1, -1e16])).nnz == 0
    # 1 + 2^-55 - 1 == 0 in serial order (spgemm_full_streaming.cpp:58-61)
    a = csr(1, 3, [0, 3], [0, 1, 2], [1, 1, 1])
    b = csr(3, 1, [0, 1, 2, 3], [0, 0, 0], [1, math.ldexp(1, -55), -1])
    assert engine.multiply(a, b).nnz == 0


def test_empty_and_degenerate_shapes(engine):
    z = csr(0, 5, [0], [], [])
    b = csr(5, 4, [0, 1, 1, 2, 2, 3], [0, 3, 1], [1.0, 2.0, 3.0])
    c = engine.multiply(z, b)
    assert c.rows == 0 and c.cols == 4 and c.nnz == 0 and c.row_ptr.tolist() == [0]
    ea = csr(3, 5, [0, 0, 0, 0], [], [])
    c = engine.multiply(ea, b)
    assert c.nnz == 0 and c.row_ptr.tolist() == [0, 0, 0, 0]
    nb = csr(5, 0, [0] * 6, [], [])
    a = csr(2, 5, [0, 2, 3], [0, 4, 2], [1.0, 1.0, 1.0])
    c = engine.multiply(a, nb)
    assert c.cols == 0 and c.nnz == 0
    # A entries that reference empty B rows only
    c = engine.multiply(csr(2, 5, [0, 2, 3], [1, 3, 3], [1.0, 2.0, 3.0]), b)
    assert c.nnz == 0 and engine.stats()["rows_per_bin"][0] == 2


def test_errors(engine):
    # The reference's spgemm() compares no shapes (only serialSpGEMM does,
    # spgemm_common.cpp:106-110): A.cols != B.rows multiplies when every A
    # column names a row of B ...
    a = csr(2, 3, [0, 1, 2], [0, 2], [1.0, 1.0])
    b = csr(4, 2, [0, 1, 1, 1, 2], [0, 1], [1.0, 1.0])
    c = engine.multiply(a, b)
    assert c.row_ptr.tolist() == [0, 1, 1] and c.col.tolist() == [0] and c.val.tolist() == [1.0]
    a5 = csr(1, 5, [0, 2], [1, 3], [2.0, 3.0])
    c = engine.multiply(a5, b)
    assert c.row_ptr.tolist() == [0, 1] and c.col.tolist() == [1] and c.val.tolist() == [3.0]
    # ... and a column that names no local row (and no halo slot) is the
    # reference's fail("missing a remote sparse row required by A")
    # (bRowForColumn, spgemm_exchange.cpp:75-87), at either end of the range
    for col in (7, 3, -1):
        bad = csr(1, 8, [0, 2], [0, col], [1.0, 1.0])
        with pytest.raises(S.SpgError) as e:
            engine.multiply(bad, csr(3, 2, [0, 1, 1, 2], [0, 1], [1.0, 1.0]))
        assert e.value.name == "SPG_EMISSING_ROW", col
        assert "missing a remote sparse row required by A" in str(e.value)
    # a malformed CSR view
    with pytest.raises(S.SpgError) as e:
        engine.multiply(csr(1, 1, [0, 2], [0], [1.0]), csr(1, 1, [0, 1], [0], [1.0]))
    assert e.value.name == "SPG_EINVAL"
    # the engine stays usable after errors
    c = engine.multiply(csr(1, 1, [0, 1], [0], [2.0]), csr(1, 1, [0, 1], [0], [3.0]))
    assert c.val.tolist() == [6.0]


def split_halo(b: O.Csr, first: int, count: int, seed: int, drop=None):
    """Local block rows [first, first+count) and the other rows as a shuffled halo."""
    s0, s1 = b.row_ptr[first], b.row_ptr[first + count]
    local = S.Csr.make(count, b.cols, b.row_ptr[first:first + count + 1] - s0, b.col[s0:s1], b.val[s0:s1])
    rows = [r for r in range(b.rows) if not (first <= r < first + count) and r != drop]
    rng = np.random.Generator(np.random.PCG64(seed))
    rng.shuffle(rows)
    rp, col, val = [0], [], []
    for r in rows:
        col.append(b.col[b.row_ptr[r]:b.row_ptr[r + 1]])
        val.append(b.val[b.row_ptr[r]:b.row_ptr[r + 1]])
        rp.append(rp[-1] + int(b.row_ptr[r + 1] - b.row_ptr[r]))
    halo = (np.array(rows, np.int32), np.array(rp, np.int64),
            np.concatenate(col) if col else np.zeros(0, np.int32), np.concatenate(val) if val else np.zeros(0))
    return local, halo


def test_halo_rows_resolve_like_the_full_product(engine, golden):
    # bRowForColumn (spgemm_exchange.cpp:75-87): local block first, then halo slots
    for name in ("uniform_rect", "ragged_dups_unsorted", "heavy_rows"):
        _, a, b = next(x for x in cases.synthetic_cases() if x[0] == name)
        first, count = b.rows // 3, b.rows // 2
        local, halo = split_halo(b, first, count, seed=1)
        c = engine.multiply(to_s(a), local, b_first_row=first, halo=halo)
        assert digest(c) == golden["products"][name]["digest"], name
    _, a, b = next(x for x in cases.synthetic_cases() if x[0] == "uniform_rect")
    referenced = int(a.col[0])
    local, halo = split_halo(b, 0, 10, seed=2, drop=referenced if referenced >= 10 else None)
    if referenced >= 10:
        with pytest.raises(S.SpgError) as e:
            engine.multiply(to_s(a), local, b_first_row=0, halo=halo)
        assert e.value.name == "SPG_EMISSING_ROW"


def test_device_resident_operands(engine, golden):
    for name in ("cage8_AA", "heavy_rows", "int_cancel_zeros"):
        _, a, b = next(x for x in list(cases.mm_cases()) + cases.synthetic_cases() if x[0] == name)
        da, db = engine.upload(to_s(a)), engine.upload(to_s(b))
        n = engine.multiply_device(da, db)
        c = engine.result_host()
        da.free()
        db.free()
        assert c.nnz == n and digest(c) == golden["products"][name]["digest"], name


def test_buffer_reuse_across_sizes(engine):
    big = W.uniform_random(20000, 20000, 12, seed=5)
    small = W.uniform_random(300, 20000, 3, seed=6)
    want_big = O.csr_digest(O.product(to_o(big), to_o(big)))
    want_small = O.csr_digest(O.product(to_o(small), to_o(big)))
    for _ in range(2):
        assert digest(engine.multiply(big, big)) == want_big
        assert digest(engine.multiply(small, big)) == want_small


def test_uniform_16384_x8_full(engine):
    # 16384^2, exactly 8 uniform columns per row, values U[-1,1): full size
    a = W.uniform_random(16384, 16384, 8, seed=42)
    b = W.uniform_random(16384, 16384, 8, seed=43)
    c = to_o(engine.multiply(a, b))
    assert engine.stats()["products"] == 16384 * 64
    ref = O.product(to_o(a), to_o(b))
    assert O.csr_digest(c) == O.csr_digest(ref)
    assert O.partial_validate(to_o(a), to_o(b), c) == (True, 0)
    rep = O.full_validate(to_o(a), to_o(b), c)
    assert rep["passed"] and rep["max_abs_error"] == 0.0


def test_uniform_2p20_x16_properties(engine):
    # 2^20 x 2^20, 16 uniform columns per row: properties at full size plus
    # bit-exact comparison with the oracle on 4096 sampled rows.
    n = 1 << 20
    a = W.uniform_random(n, n, 16, seed=42)
    b = W.uniform_random(n, n, 16, seed=43)
    c = engine.multiply(a, b)
    st = engine.stats()
    assert st["products"] == n * 256
    assert c.nnz == st["structural_nnz"] == st["nnz"]
    assert c.row_ptr[0] == 0 and c.row_ptr[-1] == c.nnz and np.all(np.diff(c.row_ptr) >= 0)
    assert sorted_rows(c)
    assert np.all(c.val != 0) and np.all(np.isfinite(c.val))
    assert c.col.min() >= 0 and c.col.max() < n
    check_sampled_rows(a, b, c, 4096, seed=7)


def test_rmat_s13_AA(engine):
    # R-MAT (0.57, 0.19, 0.19, 0.05) scale 13, edge factor 32: heavy rows take
    # the large-row path; whole product against the oracle.
    a = W.rmat(13, 32, seed=42)
    c = engine.multiply(a, a)
    st = engine.stats()
    assert st["rows_per_bin"][4] > 0 and st["rows_per_bin"][5] > 0  # bins W4 and Block
    assert digest(c) == O.csr_digest(O.product(to_o(a), to_o(a)))


def pm1_csr(rng, rows, cols, per_row):
    rp = np.arange(rows + 1, dtype=np.int64) * per_row
    col = rng.integers(0, cols, size=rows * per_row)
    val = rng.choice(np.array([-1.0, 1.0]), size=rows * per_row)
    return O.Csr.make(rows, cols, rp, col, val)


@pytest.mark.parametrize("spread,cols", [(1, 64), (4099, 300000), (16411, 1200000)])
def test_large_rows_exact_zero_sums(engine, spread, cols):
    # 40 rows of 24,000 products into 64 distinct columns with values +-1:
    # running sums cross zero between chunks and some end at exactly zero, so
    # the large-row extraction must drop them; whole product bit-exact against
    # the oracle.  The 64 columns are spread over 64, 300,000 and 1,200,000
    # columns: the presence bitmaps are in LDS; in HBM for the values and in
    # LDS (wide window) for the count; in HBM for both.
    rng = np.random.Generator(np.random.PCG64(31))
    a, b = pm1_csr(rng, 40, 4000, 3000), pm1_csr(rng, 4000, 64, 8)
    if spread > 1:
        b = O.Csr.make(b.rows, cols, b.row_ptr, b.col.astype(np.int64) * spread, b.val)
    print("case spread", spread, "cols", b.cols, flush=True)
    c = engine.multiply(to_s(a), to_s(b))
    assert engine.stats()["rows_per_bin"][6] == 40
    ref = O.product(a, b)
    assert ref.nnz < 40 * 64  # some coordinates cancel
    assert digest(c) == O.csr_digest(ref)


@pytest.mark.parametrize("path", [S.PATH_AUTO, S.PATH_TWO_PHASE], ids=["auto", "two_phase"])
def test_large_rows_rank_passes(engine, path):
    # Rows of 60,000 +-1 products into 50,000 columns (about 26,600 distinct
    # each, more than one LDS accumulator pass holds) and 3 rows of 12,000
    # products into 24 columns: the compact large-row kernel takes several
    # passes over rank ranges, running sums cancel to exactly zero inside and
    # across passes, and zeros are dropped; bit-exact against the oracle.
    rng = np.random.Generator(np.random.PCG64(33))
    b = pm1_csr(rng, 3000, 50000, 20)
    wide = pm1_csr(rng, 12, 3000, 3000)
    narrow_b = O.Csr.make(3000, 50000, b.row_ptr, (b.col.astype(np.int64) % 24) * 2083, b.val)
    a = O.Csr.make(15, 3000, np.concatenate([wide.row_ptr, wide.row_ptr[-1] + np.array([600, 1200, 1800])]),
                   np.concatenate([wide.col, rng.integers(0, 3000, size=1800)]),
                   np.concatenate([wide.val, rng.choice(np.array([-1.0, 1.0]), size=1800)]))
    engine.set_path(path)
    try:
        for bb in (b, narrow_b):
            c = engine.multiply(to_s(a), to_s(bb))
            assert engine.stats()["rows_per_bin"][6] == 15
            ref = O.product(a, bb)
            assert digest(c) == O.csr_digest(ref)
        assert ref.nnz < 3 * 24 + 12 * 50000  # exact zeros were dropped
    finally:
        engine.set_path(S.PATH_AUTO)


def test_heavy_row_tiles_in_analyze(engine):
    # Rows 0..63 hold 400 entries each (25,600 entries: one tile walked by all
    # waves of its workgroup), the other rows 5; whole product bit-exact, and
    # a column outside B's rows in the last quarter of the heavy tile's
    # entries is reported.
    rng = np.random.Generator(np.random.PCG64(34))
    lens = np.where(np.arange(300) < 64, 400, 5)
    rp = np.zeros(301, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    col = rng.integers(0, 3000, size=int(rp[-1]))
    a = O.Csr.make(300, 3000, rp, col, 2.0 * rng.random(int(rp[-1])) - 1.0)
    b = O.Csr.make(3000, 5000, np.arange(3001, dtype=np.int64) * 7, rng.integers(0, 5000, size=21000),
                   2.0 * rng.random(21000) - 1.0)
    c = engine.multiply(to_s(a), to_s(b))
    assert digest(c) == O.csr_digest(O.product(a, b))
    bad = col.copy()
    bad[int(rp[64]) - 3] = 3000 + 11
    with pytest.raises(S.SpgError) as e:
        engine.multiply(S.Csr.make(300, 3000, rp, bad, a.val), to_s(b))
    assert e.value.name == "SPG_EMISSING_ROW"


def test_large_rows_heaviest_first_order(engine):
    # Large rows of very different work, heaviest last in row order: the
    # engine reorders them by work; every row is still the oracle's.
    rng = np.random.Generator(np.random.PCG64(32))
    lens = np.concatenate([np.full(60, 300), [4000, 9000, 20000]])
    rp = np.zeros(len(lens) + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    a = O.Csr.make(len(lens), 5000, rp, rng.integers(0, 5000, size=int(rp[-1])), 2.0 * rng.random(int(rp[-1])) - 1.0)
    b = O.Csr.make(5000, 30000, np.arange(5001, dtype=np.int64) * 12, rng.integers(0, 30000, size=60000),
                   2.0 * rng.random(60000) - 1.0)
    c = engine.multiply(to_s(a), to_s(b))
    assert engine.stats()["rows_per_bin"][6] == len(lens)
    assert digest(c) == O.csr_digest(O.product(a, b))


SHIM = os.path.join(os.path.dirname(S.LIB_PATH), "shim_test")


def test_cpp_shim_matches_golden(golden, tmp_path):
    """The C++17 shim (reference-shaped spgemm/readMatrixMarket over the C ABI)
    gives the golden product and the reference's exception types."""
    import subprocess
    assert os.path.exists(SHIM), "build the library first (shim_test is built beside it)"
    for name, pa, pb in [("cage8_AA", "cage8.mtx", "cage8.mtx"),
                         ("cage8_AR", "cage8.mtx", "cage8_restriction.mtx"),
                         ("rectangular", "rectangular_a.mtx", "rectangular_b.mtx")]:
        if name not in golden["products"]:
            continue
        out = tmp_path / f"{name}.csr"
        print("case", name, flush=True)
        r = subprocess.run([SHIM, cases.data_path(pa), cases.data_path(pb), str(out)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "shim ok" in r.stdout and "validate ok" in r.stdout
        if name == "cage8_AA":
            assert "accumulate ok" in r.stdout
        c = O.read_csr(str(out))
        assert O.csr_digest(c) == golden["products"][name]["digest"], name


def test_device_halo_resolves_like_the_full_product(engine, golden):
    # bRowForColumn with the local block and the halo rows in HBM: the
    # directory is built on the device; the sender's packed rows
    # (spgemm_exchange.cpp:217-235) are the same rows, and give the same product
    for name in ("uniform_rect", "ragged_dups_unsorted", "heavy_rows"):
        _, a, b = next(x for x in cases.synthetic_cases() if x[0] == name)
        print("case", name, flush=True)
        first, count = b.rows // 3, b.rows // 2
        local, halo = split_halo(b, first, count, seed=1)
        da, dl, dbf = engine.upload(to_s(a)), engine.upload(local), engine.upload(to_s(b))
        up = S.DeviceHalo.upload(engine, *halo)
        engine.multiply_halo_device(da, dl, first, up)
        assert digest(engine.result_host()) == golden["products"][name]["digest"], name
        packed = S.DeviceHalo.pack(engine, dbf, 0, halo[0])
        g, rp, col, val = packed.host()
        assert g.tolist() == halo[0].tolist() and rp.tolist() == halo[1].tolist()
        assert col.tolist() == halo[2].tolist() and val.view(np.uint64).tolist() == np.asarray(halo[3]).view(np.uint64).tolist()
        engine.multiply_halo_device(da, dl, first, packed)
        assert digest(engine.result_host()) == golden["products"][name]["digest"], name
        for h in (up, packed):
            h.free()
        for m in (da, dl, dbf):
            m.free()


def test_device_halo_errors(engine):
    _, a, b = next(x for x in cases.synthetic_cases() if x[0] == "uniform_rect")
    print("case device halo errors", flush=True)
    first, count = 0, 10
    local, halo = split_halo(b, first, count, seed=3)
    da, dl, dbf = engine.upload(to_s(a)), engine.upload(local), engine.upload(to_s(b))
    g, rp, col, val = halo
    # duplicate halo row, row outside A's columns
    dup = np.array(g), rp
    dup[0][1] = dup[0][0]
    for rows, code in ((dup[0], "SPG_EINVAL"), (np.where(np.arange(len(g)) == 0, a.cols + 5, g), "SPG_EINVAL")):
        h = S.DeviceHalo.upload(engine, rows, rp, col, val)
        with pytest.raises(S.SpgError) as e:
            engine.multiply_halo_device(da, dl, first, h)
        assert e.value.name == code
        h.free()
    # a referenced row in neither the block nor the halo
    referenced = int(a.col[np.argmax(a.col >= count)])
    keep = [i for i in range(len(g)) if g[i] != referenced]
    sub_rp = [0]
    sub_col, sub_val = [], []
    for i in keep:
        sub_col.append(col[rp[i]:rp[i + 1]]); sub_val.append(val[rp[i]:rp[i + 1]])
        sub_rp.append(sub_rp[-1] + int(rp[i + 1] - rp[i]))
    h = S.DeviceHalo.upload(engine, g[keep], sub_rp, np.concatenate(sub_col), np.concatenate(sub_val))
    with pytest.raises(S.SpgError) as e:
        engine.multiply_halo_device(da, dl, first, h)
    assert e.value.name == "SPG_EMISSING_ROW"
    h.free()
    # a request for a row the sender does not own
    with pytest.raises(S.SpgError) as e:
        S.DeviceHalo.pack(engine, dl, first, [0, count + 3])
    assert e.value.name == "SPG_EMISSING_ROW"
    # the engine stays usable
    engine.multiply_halo_device(da, dbf, 0, None)
    for m in (da, dl, dbf):
        m.free()
