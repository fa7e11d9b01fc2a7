"""The CPU oracle pinned against the reference's golden vectors and unit tests.

Checks, per function of oracle/oracle.c:
  * products, P and Partial counts equal the reference binary's (golden.json,
    minted by oracle/_ref/refdrv from src/baselines/spgemm_exchange.cpp and
    src/common/spgemm_common.cpp) on every case, bit for bit;
  * the product equals an independent dense brute-force product on the
    small-integer cases of tests/spgemm_serial_smoke.cpp;
  * known answers of tests/spgemm_serial_smoke.cpp, tests/partial_reference_test.cpp,
    tests/spgemm_validation_smoke.cpp and tests/spgemm_full_streaming.cpp;
  * the HV15R witnesses against exact rational arithmetic
    (tests/full_roundoff_hv15r_test.py:27-34);
  * the MatrixMarket reader and the staged accumulator against the reference binary.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import cases
import oracle as O

TINY = math.ldexp(1.0, -55)


def csr(rows, cols, rp, col, val):
    return O.Csr.make(rows, cols, rp, col, val)


def all_cases():
    out = list(cases.mm_cases()) + cases.synthetic_cases()
    out += [(f"smoke_{s}",) + cases.smoke_case(s) for s in range(100)]
    out += [(f"dyadic_{i}", a, b) for i, (a, b) in enumerate(cases.dyadic_cases())]
    return out


def counts_digest(c):
    import hashlib
    return hashlib.sha256(np.asarray(c, dtype="<i8").tobytes()).hexdigest()


def test_products_match_reference_binary(golden):
    seen = 0
    for name, a, b in all_cases():
        g = golden["products"][name]
        c = O.product(a, b)
        assert O.csr_digest(c) == g["digest"], name
        assert c.nnz == g["c_nnz"], name
        assert O.scalar_products(a, b) == g["products"], name
        counts, fin = O.partial_counts(a, b)
        assert counts_digest(counts) == g["partial_digest"], name
        assert int(fin) == g["partial_finite"], name
        if "val_bits" in g:
            assert c.row_ptr.tolist() == g["row_ptr"]
            assert c.col.tolist() == g["col"]
            assert [f"{x:016x}" for x in c.val.view("<u8").tolist()] == g["val_bits"]
        seen += 1
    assert seen == len(golden["products"])


def test_product_independent_of_thread_count():
    _, a, b = next(x for x in cases.synthetic_cases() if x[0] == "ragged_dups_unsorted")
    d1 = O.csr_digest(O.product(a, b, threads=1))
    for t in (2, 3, 8):
        assert O.csr_digest(O.product(a, b, threads=t)) == d1


def dense(m):
    d = np.zeros((m.rows, m.cols))
    for r in range(m.rows):
        for p in range(m.row_ptr[r], m.row_ptr[r + 1]):
            d[r, m.col[p]] += m.val[p]
    return d


def test_dense_bruteforce_small_integers():
    # tests/spgemm_serial_smoke.cpp:33-55: exact small-integer products, output
    # columns strictly increasing and no stored zeros.
    for seed in range(100):
        a, b = cases.smoke_case(seed)
        c = O.product(a, b)
        assert c.rows == a.rows and c.cols == b.cols
        da, db = dense(a), dense(b)
        expected = np.zeros((a.rows, b.cols))
        for i in range(a.rows):
            for j in range(b.cols):
                s = 0.0
                for k in range(a.cols):
                    s += da[i, k] * db[k, j]
                expected[i, j] = s
        assert np.array_equal(dense(c), expected), seed
        for r in range(c.rows):
            cols = c.col[c.row_ptr[r]:c.row_ptr[r + 1]]
            assert np.all(np.diff(cols) > 0)
        assert np.all(c.val != 0)
        ok, mism = O.partial_validate(a, b, c)
        assert ok and mism == 0


def test_serial_smoke_known_answer():
    # tests/spgemm_serial_smoke.cpp:65-80
    a = csr(2, 3, [0, 2, 3], [0, 2, 1], [1.0, 2.0, 3.0])
    b = csr(3, 2, [0, 1, 2, 3], [1, 0, 1], [4.0, 5.0, 6.0])
    c = O.product(a, b)
    assert c.row_ptr.tolist() == [0, 1, 2] and c.col.tolist() == [1, 0] and c.val.tolist() == [16.0, 15.0]
    assert O.scalar_products(a, b) == 3


def test_partial_known_answers():
    # tests/partial_reference_test.cpp:58-62 exact cancellation -> {2,0,2}
    a = csr(3, 2, [0, 2, 2, 3], [0, 1, 1], [1, 1, 2])
    b = csr(2, 4, [0, 2, 4], [0, 1, 0, 2], [1, 2, -1, 3])
    assert O.partial_counts(a, b)[0].tolist() == [2, 0, 2]
    # :82-89 non-associative sums keep input order
    oa = csr(1, 1, [0, 1], [0], [1.0])
    assert O.partial_counts(oa, csr(1, 1, [0, 3], [0, 0, 0], [1e16, -1e16, 1]))[0].tolist() == [1]
    assert O.partial_counts(oa, csr(1, 1, [0, 3], [0, 0, 0], [1e16, 1, -1e16]))[0].tolist() == [0]


def test_partial_validation_smoke():
    # tests/spgemm_validation_smoke.cpp:44-75
    a = csr(3, 2, [0, 2, 2, 3], [0, 1, 1], [1, 1, 2])
    b = csr(2, 3, [0, 2, 4], [0, 1, 0, 2], [1, 2, -1, 3])
    product = csr(3, 3, [0, 2, 2, 4], [1, 2, 0, 2], [2, 3, -2, 6])
    assert O.partial_validate(a, b, product) == (True, 0)
    altered = csr(3, 3, [0, 2, 2, 4], [0, 2, 0, 2], [99, 3, -2, 6])
    assert O.partial_validate(a, b, altered) == (True, 0)  # values/columns not compared
    assert not O.full_validate(a, b, altered)["passed"]
    assert O.partial_validate(a, b, csr(3, 3, [0, 1, 2, 4], [1, 2, 0, 2], [2, 3, -2, 6])) == (False, 2)
    assert O.partial_validate(a, b, csr(3, 4, [0, 2, 2, 4], [1, 2, 0, 2], [2, 3, -2, 6])) == (False, 0)
    empty_a = csr(3, 2, [0, 0, 0, 0], [], [])
    empty_p = csr(3, 3, [0, 0, 0, 0], [], [])
    assert O.partial_validate(empty_a, b, empty_p) == (True, 0)
    assert O.partial_validate(a, b, empty_p) == (False, 2)
    assert O.partial_validate(empty_a, b, product) == (False, 2)
    huge = csr(1, 1, [0, 1], [0], [1e200])
    assert O.partial_validate(huge, huge, csr(1, 1, [0, 1], [0], [1.0]))[0] is False
    tiny = csr(1, 1, [0, 1], [0], [1e-200])
    assert O.partial_validate(tiny, tiny, csr(1, 1, [0, 0], [], [])) == (True, 0)
    for v in (math.nan, math.inf, -math.inf):
        bad = csr(3, 3, [0, 2, 2, 4], [1, 2, 0, 2], [v, 3, -2, 6])
        assert O.partial_validate(a, b, bad) == (False, 0)
        nb = csr(2, 3, [0, 2, 4], [0, 1, 0, 2], [1, v, -1, 3])
        assert O.partial_validate(a, nb, product)[0] is False


def max_abs_difference(l, r):
    """maxAbsoluteDifference (spgemm_common.cpp:489-528), as a Python check."""
    if l.rows != r.rows or l.cols != r.cols:
        return math.inf
    if not (np.all(np.isfinite(l.val)) and np.all(np.isfinite(r.val))):
        return math.inf
    m = 0.0
    for row in range(l.rows):
        i, j = int(l.row_ptr[row]), int(r.row_ptr[row])
        ie, je = int(l.row_ptr[row + 1]), int(r.row_ptr[row + 1])
        while i < ie or j < je:
            if j == je or (i < ie and l.col[i] < r.col[j]):
                d = abs(l.val[i]); i += 1
            elif i == ie or r.col[j] < l.col[i]:
                d = abs(r.val[j]); j += 1
            else:
                d = abs(l.val[i] - r.val[j]); i += 1; j += 1
            if not math.isfinite(d):
                return math.inf
            m = max(m, d)
    return m


def equivalent(a, b, c):
    # spgemm_full_streaming.cpp:20-25: streaming max_abs_error == materialized one
    rep = O.full_validate(a, b, c)
    if rep["environment_ok"] and rep["matrices_ok"]:
        assert rep["max_abs_error"] == max_abs_difference(c, O.product(a, b))
    return rep


def test_full_streaming_known_answers():
    # tests/spgemm_full_streaming.cpp:42-57
    for a, b in cases.dyadic_cases():
        c = O.product(a, b)
        assert equivalent(a, b, c)["passed"]
        if c.nnz:
            c2 = O.Csr(c.rows, c.cols, c.row_ptr, c.col, c.val.copy())
            c2.val[0] += 1
            assert not equivalent(a, b, c2)["passed"]
        missing = csr(a.rows, b.cols, np.zeros(a.rows + 1, np.int64), [], [])
        equivalent(a, b, missing)
    # :58-68 serial 1 + 2^-55 - 1 is zero; a reordered residual passes Full, fails Partial
    a = csr(1, 3, [0, 3], [0, 1, 2], [1, 1, 1])
    b = csr(3, 1, [0, 1, 2, 3], [0, 0, 0], [1, TINY, -1])
    reordered = csr(1, 1, [0, 1], [0], [TINY])
    assert O.partial_validate(a, b, reordered)[0] is False
    rep = equivalent(a, b, reordered)
    assert rep["passed"] and rep["max_abs_error"] == TINY and rep["row_nnz_mismatches"] == 1
    for v in (0.5e-10, 1e-10, 1.0, math.inf, math.nan):
        assert not O.full_validate(a, b, csr(1, 1, [0, 1], [0], [v]))["passed"]
    # :79-83 wrong column
    col_a = csr(1, 1, [0, 1], [0], [1])
    col_b = csr(1, 2, [0, 1], [0], [1])
    assert not equivalent(col_a, col_b, csr(1, 2, [0, 1], [1], [1]))["passed"]
    unit = csr(1, 1, [0, 1], [0], [1])
    for v in (1e200, 1e-200, math.inf, math.nan):
        large = csr(1, 1, [0, 1], [0], [v])
        equivalent(large, large, unit)
    # :92-97 malformed results
    for rp, cl, vl in (([0, 2], [0, 0], [0.5, 0.5]), ([0, 1], [1], [1])):
        assert not O.full_validate(unit, unit, csr(1, 1, rp, cl, vl))["passed"]
    # :98-104 scale matters
    small = csr(1, 1, [0, 1], [0], [1e-150])
    assert not O.full_validate(small, small, csr(1, 1, [0, 1], [0], [1e-300 + 1e-310]))["passed"]
    assert O.full_validate(small, small, O.product(small, small))["passed"]
    # :105-110 gradual underflow and subnormals
    und = csr(1, 1, [0, 1], [0], [1e-200])
    assert O.full_validate(und, und, csr(1, 1, [0, 0], [], []))["passed"]
    sub = csr(1, 1, [0, 1], [0], [5e-324])
    assert O.full_validate(sub, unit, sub)["passed"]
    # :111-114 spurious column without contributing terms
    assert not O.full_validate(col_a, col_b, csr(1, 2, [0, 2], [0, 1], [1, 1e-300]))["passed"]
    # :115-123 invalid inputs
    for rp, cl, vl in (([0, 1], [-1], [1]), ([0, 1], [1], [1]), ([0, 1], [0], [math.nan])):
        bad = csr(1, 1, rp, cl, vl)
        assert not O.full_validate(bad, unit, unit)["passed"]
        assert not O.full_validate(unit, bad, unit)["passed"]


def test_full_hv15r_witnesses():
    # tests/spgemm_full_streaming.cpp:139-166 (pinned historical max_abs_error)
    for name, row, column, err in (("zero-cancellation-witness", 426610, 448836, 1.8135240096338719e-20),
                                   ("absolute-threshold-witness", 45207, 1397793, 1.0186340659856796e-10)):
        a = O.mm_read(cases.data_path(f"hv15r_roundoff/{name}-A.mtx"))
        b = O.mm_read(cases.data_path(f"hv15r_roundoff/{name}-B.mtx"))
        zero = name.startswith("zero")
        rp = np.zeros(a.rows + 1, np.int64)
        cols, vals = [], []
        if not zero:
            cols, vals = [column], [73368.519419447723]
            rp[row + 1:] = 1
        cand = csr(a.rows, b.cols, rp, cols, vals)
        rep = O.full_validate(a, b, cand)
        assert rep["passed"] and rep["max_abs_error"] == err
        if zero:
            rp = np.zeros(a.rows + 1, np.int64)
            rp[row + 1:] = 1
            cols, vals = [column], [0.0]
        cand = csr(a.rows, b.cols, rp, cols, [vals[0] + 0.001])
        assert not O.full_validate(a, b, cand)["passed"]


def test_hv15r_witness_exact_rationals():
    # tests/full_roundoff_hv15r_test.py:27-34: the witness data against exact arithmetic
    for name in ("zero-cancellation-witness", "absolute-threshold-witness"):
        w = json.load(open(cases.data_path(f"hv15r_roundoff/{name}.json")))
        exact = sum((Fraction(t["a"]) * Fraction(t["b"]) for t in w["terms_in_serial_order"]), Fraction())
        assert float(exact) == w["exact_dot_as_float"]
        for label in ("serial", "trident"):
            assert float(abs(Fraction(w[label]) - exact)) == w[label + "_error_vs_exact"]
        # the oracle's serial value for that coordinate is the witness's serial value
        a = O.mm_read(cases.data_path(f"hv15r_roundoff/{name}-A.mtx"))
        b = O.mm_read(cases.data_path(f"hv15r_roundoff/{name}-B.mtx"))
        c = O.product(a, b)
        r = w["row"]
        got = c.val[c.row_ptr[r]:c.row_ptr[r + 1]][c.col[c.row_ptr[r]:c.row_ptr[r + 1]] == w["col"]]
        assert (got.tolist() or [0.0]) == [w["serial"]] or (w["serial"] != 0 and got.tolist() == [w["serial"]])


def test_full_report_relative_errors():
    a = csr(1, 2, [0, 2], [0, 1], [1.0, 2.0])
    b = csr(2, 2, [0, 1, 2], [0, 1], [3.0, 4.0])
    c = O.product(a, b)
    rep = O.full_validate(a, b, c)
    assert rep["max_entry_rel_error"] == 0 and rep["frobenius_rel_error"] == 0
    off = csr(1, 2, [0, 2], [0, 1], [3.0, 8.0 * (1 + 2 ** -40)])
    rep = O.full_validate(a, b, off)
    assert rep["max_entry_rel_error"] == pytest.approx(2 ** -40, rel=1e-12)
    assert not rep["passed"]


def test_mm_reader_matches_reference(golden):
    for rel, g in golden["mm_read"].items():
        path = cases.data_path(rel)
        if g["ok"]:
            m = O.mm_read(path)
            assert O.csr_digest(m) == g["digest"], rel
        else:
            with pytest.raises(O.OracleError) as e:
                O.mm_read(path)
            assert g["error"] in str(e.value), rel


def test_accumulate_matches_reference(golden):
    for name, rows, cols, stages in cases.accumulate_cases():
        c = O.accumulate(rows, cols, stages)
        g = golden["accumulate"][name]
        assert c.nnz == g["c_nnz"] and O.csr_digest(c) == g["digest"], name
    # roundoff_stages.mtx: staged order keeps 2^-55 that the serial order cancels
    assert golden["accumulate"]["roundoff_2stages"]["c_nnz"] == 1
    assert golden["products"]["roundoff_stages"]["c_nnz"] == 0


def test_percentile90_nearest_rank():
    # spgemm_common.cpp:916-924: index ceil(0.9 n) - 1 of the sorted samples
    assert O.percentile90([5.0]) == 5.0
    assert O.percentile90([2.0, 1.0]) == 2.0
    assert O.percentile90(list(range(10, 0, -1))) == 9.0
    assert O.percentile90([float(x) for x in range(1, 21)]) == 18.0


def test_reference_binary_agrees_when_present(tmp_path):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built here")
    a, b = cases.smoke_case(7)
    O.write_csr(str(tmp_path / "a.csr"), a)
    O.write_csr(str(tmp_path / "b.csr"), b)
    O.ref_run("spgemm", tmp_path / "a.csr", tmp_path / "b.csr", tmp_path / "c.csr")
    assert O.csr_digest(O.read_csr(str(tmp_path / "c.csr"))) == O.csr_digest(O.product(a, b))
