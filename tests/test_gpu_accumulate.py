"""Staged accumulation on the device against the reference's
trident::LocalAccumulator (golden digests minted from the reference binary)
and against the oracle's restatement."""
import numpy as np
import pytest

import cases
import oracle as O
import spgemm_amd as S

pytestmark = pytest.mark.gpu


def to_s(m) -> S.Csr:
    return S.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val)


@pytest.mark.parametrize("path", [S.PATH_AUTO, S.PATH_TWO_PHASE])
def test_accumulate_matches_reference(engine, golden, path):
    engine.set_path(path)
    try:
        for name, rows, cols, stages in cases.accumulate_cases():
            print("case", name, flush=True)
            acc = S.Accumulator(engine, rows, cols)
            for a, b in stages:
                acc.add_product(to_s(a), to_s(b))
            c = acc.finish()
            acc.close()
            g = golden["accumulate"][name]
            assert c.nnz == g["c_nnz"], name
            assert O.csr_digest(O.Csr.make(c.rows, c.cols, c.row_ptr, c.col, c.val)) == g["digest"], name
    finally:
        engine.set_path(S.PATH_AUTO)


def test_accumulate_finish_twice_and_more_stages(engine):
    a = S.Csr.make(2, 2, [0, 2, 3], [0, 1, 1], [1.0, 2.0, 3.0])
    b = S.Csr.make(2, 3, [0, 1, 3], [2, 0, 2], [1.0, -1.0, 4.0])
    acc = S.Accumulator(engine, 2, 3)
    acc.add_product(a, b)
    c1 = acc.finish()
    c2 = acc.finish()
    assert np.array_equal(c1.val, c2.val) and np.array_equal(c1.col, c2.col)
    # the same product again doubles every coordinate; minus it twice cancels
    acc.add_product(a, b)
    c3 = acc.finish()
    assert np.array_equal(c3.val, 2 * c1.val)
    neg = S.Csr.make(2, 2, [0, 2, 3], [0, 1, 1], [-1.0, -2.0, -3.0])
    acc.add_product(neg, b)
    acc.add_product(neg, b)
    assert acc.finish().nnz == 0
    with pytest.raises(S.SpgError) as e:
        acc.add_product(S.Csr.make(3, 2, [0, 0, 0, 0], [], []), b)  # wrong row count
    assert e.value.name == "SPG_EINVAL"
    acc.close()


def test_accumulate_large_rows_with_cancellation(engine):
    # Two stages of rows with 24,000 products each into 64 columns, values
    # +-1 (the large-row path with exact-zero running sums), against the
    # oracle's restatement of LocalAccumulator.
    rng = np.random.Generator(np.random.PCG64(33))

    def pm1(rows, cols, per_row):
        rp = np.arange(rows + 1, dtype=np.int64) * per_row
        return O.Csr.make(rows, cols, rp, rng.integers(0, cols, size=rows * per_row),
                          rng.choice(np.array([-1.0, 1.0]), size=rows * per_row))

    stages = [(pm1(30, 4000, 3000), pm1(4000, 64, 8)), (pm1(30, 2000, 3000), pm1(2000, 64, 8))]
    acc = S.Accumulator(engine, 30, 64)
    for a, b in stages:
        acc.add_product(to_s(a), to_s(b))
    c = acc.finish()
    acc.close()
    ref = O.accumulate(30, 64, stages)
    assert c.nnz == ref.nnz
    assert O.csr_digest(O.Csr.make(c.rows, c.cols, c.row_ptr, c.col, c.val)) == O.csr_digest(ref)
