"""Batched products (spg_batch_*): every C_i of a batch, chained batch, graph
replay and autotuned pipeline is bit-identical to the standalone product of
the same operands, and the standalone products match the serial-order oracle."""
import numpy as np
import pytest

import oracle as O
import spgemm_amd as S
import workloads as W

pytestmark = pytest.mark.gpu


def digest(m) -> str:
    return O.csr_digest(O.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val))


def small_jobs(n=24, seed=7):
    return W.batch_jobs(n, seed=seed, max_rows=2048, n_lo=6, n_hi=12)


def waves(jobs):
    ab = [(j[1], j[2]) for j in jobs if j[0] == "ab"]
    rap = [j[1:] for j in jobs if j[0] == "rap"]
    a1 = [a for a, _ in ab] + [r for r, _, _ in rap]
    b1 = [b for _, b in ab] + [a for _, a, _ in rap]
    return ab, rap, a1, b1


def test_batch_results_equal_standalone(engine):
    jobs = small_jobs()
    ab, rap, a1, b1 = waves(jobs)
    batch = S.Batch(engine, a1, b1)
    total = batch.run()
    got = [batch.result(i) for i in range(len(a1))]
    assert total == sum(g.nnz for g in got)
    for i, (a, b) in enumerate(zip(a1, b1)):
        print("product", i, a.rows, a.nnz, flush=True)
        alone = engine.multiply(a, b)
        assert digest(got[i]) == digest(alone), i
        if i % 5 == 0:
            ref = O.product(O.Csr.make(a.rows, a.cols, a.row_ptr, a.col, a.val),
                            O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val))
            assert O.csr_digest(ref) == digest(got[i]), i
    # second product of each Galerkin triple, chained on the device
    k0 = len(ab)
    chain = S.Batch.chained(engine, batch, [k0 + k for k in range(len(rap))], [p for _, _, p in rap])
    chain.run()
    for k, (r, a, p) in enumerate(rap):
        ra = engine.multiply(r, a)
        rap_alone = engine.multiply(ra, p)
        c = chain.result(k)
        assert digest(c) == digest(rap_alone), k
        # closed form: a coarse row of R.A.P sums the row sums of A over its
        # 2x2x2 fine cells (integer stencil values: exact)
        rs_a = np.add.reduceat(a.val, a.row_ptr[:-1]) if a.nnz else np.zeros(a.rows)
        rs_a[np.diff(a.row_ptr) == 0] = 0
        want = np.zeros(p.cols)
        np.add.at(want, p.col, rs_a)
        have = np.zeros(c.rows)
        for i in range(c.rows):
            have[i] = c.val[c.row_ptr[i]:c.row_ptr[i + 1]].sum()
        assert np.array_equal(want, have), k
    batch.close()
    chain.close()


def test_batch_graph_replay_and_autotune(engine):
    jobs = small_jobs(18, seed=11)
    _, _, a1, b1 = waves(jobs)
    batch = S.Batch(engine, a1, b1)
    batch.run()
    ref = [digest(batch.result(i)) for i in range(len(a1))]
    batch.set_graph(True)
    for rep in range(3):  # first graph run captures, later ones replay
        print("graph run", rep, flush=True)
        batch.run()
        assert [digest(batch.result(i)) for i in range(len(a1))] == ref
    choice = batch.autotune(1)
    assert choice in (0, 1, 2, 3, 4)
    for rep in range(2):
        batch.run()
        assert [digest(batch.result(i)) for i in range(len(a1))] == ref
    batch.close()


@pytest.mark.parametrize("path", [S.PATH_SINGLE_PASS_DEALT, S.PATH_SINGLE_PASS_PACKED])
def test_batch_dealt_tiles_equal_standalone(engine, path):
    # Single pass over dealt 32-row tiles or packed tiles of small rows (the
    # batch inherits the engine's path), eager and replayed from a graph,
    # against each product alone on the default path.
    jobs = small_jobs(20, seed=13)
    _, _, a1, b1 = waves(jobs)
    want = [digest(engine.multiply(a, b)) for a, b in zip(a1, b1)]
    engine.set_path(path)
    try:
        batch = S.Batch(engine, a1, b1)
        batch.run()
        assert [digest(batch.result(i)) for i in range(len(a1))] == want
        batch.set_graph(True)
        for rep in range(3):
            print("dealt graph run", rep, flush=True)
            batch.run()
            assert [digest(batch.result(i)) for i in range(len(a1))] == want
        batch.close()
    finally:
        engine.set_path(S.PATH_AUTO)


def test_batch_errors(engine):
    a = S.Csr.make(2, 3, [0, 1, 2], [0, 2], [1.0, 1.0])
    b = S.Csr.make(4, 2, [0, 1, 1, 1, 2], [0, 1], [1.0, 1.0])
    with pytest.raises(S.SpgError) as e:
        S.Batch(engine, [a], [b])
    assert e.value.name == "SPG_EINVAL"
    bad = S.Csr.make(1, 3, [0, 1], [7], [1.0])  # column outside B_i's rows
    ok_b = S.Csr.make(3, 2, [0, 1, 1, 2], [0, 1], [1.0, 1.0])
    batch = S.Batch(engine, [bad], [ok_b])
    with pytest.raises(S.SpgError) as e:
        batch.run()
    assert e.value.name == "SPG_ERANGE"
    batch.close()


def big_bin_jobs(seed=23):
    """Plain products whose rows reach bins W4 (<= 1024 products), Block
    (<= 2048) and Large: 24..60 uniform columns per row in A and B, plus a few
    of the small jobs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for rows in (300, 1100, 2500, 700):
        out.append((W.variable_uniform(rows, rows, 24, 60, rng), W.variable_uniform(rows, rows, 24, 60, rng)))
    out += [(j[1], j[2]) for j in small_jobs(9, seed=seed) if j[0] == "ab"]
    return out


@pytest.mark.parametrize("path,fmt", [(S.PATH_AUTO, S.FORMAT_WIDE), (S.PATH_SINGLE_PASS_PACKED, S.FORMAT_WIDE),
                                      (S.PATH_TWO_PHASE, S.FORMAT_AUTO), (S.PATH_SINGLE_PASS_DEALT, S.FORMAT_WIDE)],
                         ids=["auto-wide", "packed-wide", "two_phase-narrow", "dealt-wide"])
def test_batch_big_bins_and_wide_format(engine, path, fmt):
    # A batch whose rows fill bins W4, Block and Large, in the 64-bit format
    # of the big batches (B arenas of 2^29 entries or more) and the 32-bit
    # one, eager and replayed from a graph: every C_i equals its standalone
    # product on the default path, and every third the oracle.
    pairs = big_bin_jobs()
    a1, b1 = [a for a, _ in pairs], [b for _, b in pairs]
    want = [digest(engine.multiply(a, b)) for a, b in zip(a1, b1)]
    for i in range(0, len(pairs), 3):
        a, b = pairs[i]
        ref = O.product(O.Csr.make(a.rows, a.cols, a.row_ptr, a.col, a.val),
                        O.Csr.make(b.rows, b.cols, b.row_ptr, b.col, b.val))
        assert O.csr_digest(ref) == want[i], i
    engine.set_path(path)
    engine.set_format(fmt)
    try:
        batch = S.Batch(engine, a1, b1)
        batch.run()
        st = batch.stats()
        assert all(st["rows_per_bin"][k] > 0 for k in (4, 5, 6)), st["rows_per_bin"]
        assert st["wide_format"] == (fmt == S.FORMAT_WIDE)
        assert [digest(batch.result(i)) for i in range(len(a1))] == want
        batch.set_graph(True)
        for rep in range(2):
            print("graph run", rep, flush=True)
            batch.run()
            assert [digest(batch.result(i)) for i in range(len(a1))] == want
        batch.close()
    finally:
        engine.set_path(S.PATH_AUTO)
        engine.set_format(S.FORMAT_AUTO)


def test_batch_count_then_run_and_trim(engine):
    """The count phase alone (spg_batch_count) then the run gives the same
    products as one run; the reported value bytes are the buffer growth the
    run then shows; a trimmed batch holds no device bytes and rebuilds."""
    jobs = small_jobs(12, seed=23)
    _, _, a1, b1 = waves(jobs)
    ref = S.Batch(engine, a1, b1)
    ref.run()
    want = [digest(ref.result(i)) for i in range(len(a1))]
    batch = S.Batch(engine, a1, b1)
    before = batch.device_bytes()
    grow = batch.count()
    counted = batch.device_bytes()
    assert counted >= before and grow > 0
    total = batch.run()
    after = batch.device_bytes()
    # the value phase grows the buffers by the reported bytes, plus the
    # compaction buffers when a product has exact zeros (not in the bound)
    assert counted + grow <= after <= counted + grow + 12 * total + 8 * (sum(a.rows for a in a1) + 1) + 768
    assert [digest(batch.result(i)) for i in range(len(a1))] == want
    # counted again on warm buffers: nothing more to grow
    assert batch.count() == 0
    batch.run()
    assert [digest(batch.result(i)) for i in range(len(a1))] == want
    batch.trim()
    assert batch.device_bytes() == 0
    # the estimate of a rebuild's arena and count workspace is what they take
    a5, b5 = a1[:5], b1[:5]
    est = batch.estimate(sum(a.rows for a in a5), sum(a.nnz for a in a5), sum(b.rows for b in b5),
                         sum(b.nnz for b in b5), max(b.cols for b in b5), 0)
    batch.rebuild(a5, b5)
    batch.count()
    assert batch.device_bytes() == est
    batch.run()
    assert [digest(batch.result(i)) for i in range(5)] == want[:5]
