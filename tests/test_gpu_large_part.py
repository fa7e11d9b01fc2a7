"""Rows above 2048 scalar products partitioned by windows of 4096 columns
(SPG_LARGE_AUTO) and folded window by window in LDS: bit-exact against the
serial-order oracle and identical to the HBM-window kernels
(SPG_LARGE_WINDOWS), for rows whose windows cancel to exact zeros, rows cut
into several fold tasks, hub columns taking most of a window's products,
windows of many chunks, windows of 8 to 2,056 products (every per-lane item
count of a fold chunk), and C with more columns than the compact kernel
takes."""
import numpy as np
import pytest

import oracle as O
import spgemm_amd as S
from parity_util import digest, to_s

pytestmark = pytest.mark.gpu


def _csr(rows, cols, lens, col, val):
    rp = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    return O.Csr.make(rows, cols, rp, np.asarray(col, dtype=np.int64), np.asarray(val, dtype=np.float64))


def _both_modes(engine, a, b, paths=(S.PATH_AUTO, S.PATH_TWO_PHASE)):
    ref = O.csr_digest(O.product(a, b, threads=8))
    for path in paths:
        engine.set_path(path)
        try:
            for mode in (S.LARGE_AUTO, S.LARGE_WINDOWS):
                print("path", path, "large mode", mode, flush=True)
                engine.set_large(mode)
                c = engine.multiply(to_s(a), to_s(b))
                st = engine.stats()
                assert digest(c) == ref, (path, mode)
                if mode == S.LARGE_WINDOWS:
                    assert not st["partitioned"]
                yield mode, st, c
        finally:
            engine.set_large(S.LARGE_AUTO)
            engine.set_path(S.PATH_AUTO)


def test_partition_cancellation_tasks_hubs(engine):
    rng = np.random.Generator(np.random.PCG64(71))
    cols = 200_000  # 49 windows; above the compact kernel's 65,536 columns
    nb = 4000
    # B: rows of 40 random columns; rows 2k and 2k+1 are equal, so A entries
    # (2k, x) and (2k+1, -x) cancel to exact zeros in whatever windows they hit
    blen = np.full(nb, 40)
    bcol = rng.integers(0, cols, size=nb * 40)
    bval = rng.uniform(-1, 1, size=nb * 40)
    bcol = bcol.reshape(nb, 40)
    bval = bval.reshape(nb, 40)
    bcol[1::2] = bcol[0::2]
    bval[1::2] = bval[0::2]
    b = _csr(nb, cols, blen, bcol.ravel(), bval.ravel())
    # the same with every B row's first entry in column 7: a hub column that
    # takes a large share of window 0's products
    hub = O.Csr.make(nb, cols, b.row_ptr, np.where(np.arange(nb * 40) % 40 == 0, 7, b.col), b.val)
    rows = []
    # 0-7: cancelling pairs (some rows fully, some partly)
    for r in range(8):
        k = rng.choice(nb // 2, size=80, replace=False) * 2
        x = rng.uniform(-1, 1, size=80)
        cancel = 80 if r < 3 else 40
        ac = np.concatenate([k, k[:cancel] + 1])
        av = np.concatenate([x, -x[:cancel]])
        rows.append((ac, av))
    # 8-9: 2,000 entries each (80,000 products: several fold tasks and
    # windows of several chunks)
    for r in range(2):
        ac = rng.integers(0, nb, size=2000)
        rows.append((ac, rng.uniform(-1, 1, size=2000)))
    # 10-19: rows of 60-120 entries (2,400-4,800 products)
    for r in range(10):
        n = int(rng.integers(60, 121))
        rows.append((rng.integers(0, nb, size=n), rng.uniform(-1, 1, size=n)))
    lens = [len(c) for c, _ in rows]
    a = _csr(len(rows), nb, lens, np.concatenate([c for c, _ in rows]), np.concatenate([v for _, v in rows]))
    for bb in (b, hub):
        seen = {}
        for mode, st, c in _both_modes(engine, a, bb):
            seen[mode] = st["partitioned"]
            assert st["rows_per_bin"][6] == len(rows)
        assert seen[S.LARGE_AUTO]  # the partition ran
    ref = O.product(a, b)
    # rows 0-2 cancel but for columns two pairs share (the pairs' partial sums
    # need not cancel exactly then): most of their ~3,200 columns drop
    assert all(int(ref.row_ptr[r + 1] - ref.row_ptr[r]) < 300 for r in range(3))


def test_partition_beside_compact_rows(engine):
    # C of 40,000 columns: rows whose distinct columns fit the compact
    # kernel's accumulator fold there, the wider ones are partitioned; both
    # kinds in one product, through the batch pipeline too.
    rng = np.random.Generator(np.random.PCG64(72))
    cols, nb = 40_000, 3000
    b = _csr(nb, cols, np.full(nb, 30), rng.integers(0, cols, size=nb * 30), rng.uniform(-1, 1, size=nb * 30))
    lens = np.array([80] * 6 + [1500] * 3 + [120] * 6)  # 2,400 .. 45,000 products per row
    a = _csr(len(lens), nb, lens, rng.integers(0, nb, size=int(lens.sum())), rng.uniform(-1, 1, size=int(lens.sum())))
    seen = {}
    for mode, st, c in _both_modes(engine, a, b):
        seen[mode] = st["partitioned"]
    assert seen[S.LARGE_AUTO]
    want = O.csr_digest(O.product(a, b))
    batch = S.Batch(engine, [to_s(a), to_s(a)], [to_s(b), to_s(b)])
    batch.run()
    assert digest(batch.result(0)) == want and digest(batch.result(1)) == want
    assert batch.stats()["partitioned"]


@pytest.mark.parametrize("fmt", [S.FORMAT_AUTO, S.FORMAT_WIDE])
def test_partition_keep_zeros_accumulator(engine, fmt):
    # The staged accumulator keeps structural zeros (keep_zeros): windows with
    # exact zeros must keep them; equal under both large-row modes.
    rng = np.random.Generator(np.random.PCG64(73))
    cols, nb = 150_000, 2000
    b = _csr(nb, cols, np.full(nb, 30), rng.integers(0, cols, size=nb * 30), rng.choice([-1.0, 1.0], size=nb * 30))
    lens = np.full(6, 120)
    a = _csr(6, nb, lens, rng.integers(0, nb, size=720), rng.choice([-1.0, 1.0], size=720))
    out = {}
    engine.set_format(fmt)
    try:
        for mode in (S.LARGE_AUTO, S.LARGE_WINDOWS):
            engine.set_large(mode)
            acc = S.Accumulator(engine, 6, cols)
            acc.add_product(to_s(a), to_s(b))
            acc.add_product(to_s(a), to_s(b))
            out[mode] = digest(acc.finish())
            acc.close()
    finally:
        engine.set_large(S.LARGE_AUTO)
        engine.set_format(S.FORMAT_AUTO)
    assert out[S.LARGE_AUTO] == out[S.LARGE_WINDOWS]
    assert out[S.LARGE_AUTO] == O.csr_digest(O.accumulate(6, cols, [(a, b), (a, b)]))


def test_heavy_rows_analyze_pieces_and_count_launch(engine):
    # A first row tile of about 70,000 A entries (cut into analyze pieces),
    # rows of 480 K and 960 K products (counted by the heavy-row launch: C of
    # 65,536 columns, a narrow count bitmap), hub columns, rows the compact
    # kernel takes (at most 4096 distinct columns), an empty row inside the
    # heavy tile, and B rows of 24 entries (entry-major count batches) as well
    # as of up to 90 entries (group lookups).
    rng = np.random.Generator(np.random.PCG64(74))
    cols, nb = 65_536, 65_536
    for blen_hi in (24, 90):
        # (B rows of 0 entries: A entries whose products are none, several
        # entries starting at one product)
        blen = np.full(nb, 24) if blen_hi == 24 else rng.integers(0, blen_hi + 1, size=nb)
        bnnz = int(blen.sum())
        bcol = rng.integers(0, cols, size=bnnz)
        bcol[::7] = 5  # a hub column
        b = _csr(nb, cols, blen, bcol, rng.uniform(-1, 1, size=bnnz))
        lens = np.full(200, 8)
        lens[0], lens[1], lens[3] = 20_000, 40_000, 9_000
        lens[2] = 0
        lens[100:110] = 160  # ~3,800 products each: compact rows
        acol = rng.integers(0, nb, size=int(lens.sum()))
        a = _csr(len(lens), nb, lens, acol, rng.uniform(-1, 1, size=int(lens.sum())))
        want = O.scalar_products(a, b)
        assert want > 2 * (1 << 18)  # rows 0 and 1 take the heavy-row count launch
        seen = {}
        for mode, st, c in _both_modes(engine, a, b, paths=(S.PATH_AUTO,)):
            assert st["products"] == want
            assert st["rows_per_bin"][6] >= 13  # rows 0, 1, 3 and the compact rows are large
            seen[mode] = st["partitioned"]
        assert seen[S.LARGE_AUTO]
        # the last row tile partial (rows 64-69) and heavy: its pieces are
        # at most one per row
        lens2 = np.full(70, 4)
        lens2[64:70] = 10_000
        a2 = _csr(70, nb, lens2, rng.integers(0, nb, size=int(lens2.sum())), rng.uniform(-1, 1, size=int(lens2.sum())))
        for mode, st, c in _both_modes(engine, a2, b, paths=(S.PATH_AUTO,)):
            assert st["products"] == O.scalar_products(a2, b)


@pytest.mark.parametrize("hub", [False, True], ids=["spread", "hub"])
def test_partition_window_sizes(engine, hub):
    # Rows whose windows (4096 columns each) get 8 to 2,056 products: a fold
    # chunk of n products gives each lane ceil(n / 256) items, so every item
    # count from 1 to 4 and partial last chunks of several-chunk windows occur.
    # With `hub`, half of each B row's entries fall in one column of their
    # window (the sorted path); bit-exact against the oracle.
    rng = np.random.Generator(np.random.PCG64(72))
    cols, per = 200_000, 8
    nwin = (cols + 4095) // 4096
    sizes = [1, 8, 32, 33, 64, 65, 96, 97, 128, 129, 256, 257]  # B rows per window in a row: 8x products
    rows_b = []
    for w in range(nwin):
        lo, hi = w * 4096, min(cols, (w + 1) * 4096)
        for _ in range(260):
            c = rng.integers(lo, hi, size=per)
            if hub:
                c[: per // 2] = lo + 7
            rows_b.append(c)
    nb = len(rows_b)
    b = _csr(nb, cols, np.full(nb, per), np.concatenate(rows_b), rng.uniform(-1, 1, size=nb * per))
    arows = []
    for r in range(6):
        ws = rng.choice(nwin, size=len(sizes), replace=False)
        ks = []
        for w, m in zip(ws, np.roll(sizes, r)):
            ks.append(w * 260 + rng.choice(260, size=int(m), replace=False))
        k = np.concatenate(ks)
        rng.shuffle(k)
        arows.append(k)
    a = _csr(len(arows), nb, [len(k) for k in arows], np.concatenate(arows),
             rng.uniform(-1, 1, size=sum(len(k) for k in arows)))
    for mode, st, c in _both_modes(engine, a, b, paths=(S.PATH_AUTO,)):
        assert st["rows_per_bin"][6] == len(arows)
        if mode == S.LARGE_AUTO:
            assert st["partitioned"]
