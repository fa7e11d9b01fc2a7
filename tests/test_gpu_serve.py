"""Served products (spg_server_*): requests submitted concurrently from
several client threads, batched continuously on the device, each give the
result of computing the same product alone; admission respects the budget."""
import threading
import time

import numpy as np
import pytest

import oracle as O
import spgemm_amd as S
import workloads as W

pytestmark = pytest.mark.gpu


def digest(m) -> str:
    return O.csr_digest(O.Csr.make(m.rows, m.cols, m.row_ptr, m.col, m.val))


def test_served_mix_equals_alone(engine):
    jobs = W.batch_jobs(30, seed=19, max_rows=2048, n_lo=6, n_hi=10)
    host_pairs = [(j[1], j[2]) for j in jobs]  # plain products and the R.A of triples
    big = (W.uniform_random(1 << 15, 1 << 15, 16, seed=5), W.uniform_random(1 << 15, 1 << 15, 16, seed=6))
    srv = S.Server(0, budget_bytes=8 << 30, max_batch=64)
    try:
        pool = [W.uniform_random(512, 512, 12, seed=100 + k) for k in range(4)]
        handles = [srv.register(m) for m in pool]
        res_pairs = [(i % 4, (i * 3 + 1) % 4) for i in range(12)]
        ids = {}
        lock = threading.Lock()
        rng = np.random.Generator(np.random.PCG64(3))
        delays = rng.exponential(0.002, size=len(host_pairs) + len(res_pairs) + 1)

        def client(kind, items):
            for key, item in items:
                time.sleep(float(delays[key % len(delays)]))
                rid = srv.submit(*item) if kind == "host" else srv.submit_resident(handles[item[0]],
                                                                                     handles[item[1]])
                with lock:
                    ids[(kind, key)] = rid

        threads = [threading.Thread(target=client, args=("host", list(enumerate(host_pairs))[t::3]))
                   for t in range(3)]
        threads.append(threading.Thread(target=client, args=("res", list(enumerate(res_pairs)))))
        for t in threads:
            t.start()
        big_id = srv.submit(*big)
        for t in threads:
            t.join()
        for (kind, key), rid in sorted(ids.items()):
            print("request", kind, key, flush=True)
            if kind == "host":
                a, b = host_pairs[key]
            else:
                a, b = pool[res_pairs[key][0]], pool[res_pairs[key][1]]
            got = srv.result(rid)
            assert digest(got) == digest(engine.multiply(a, b)), (kind, key)
            srv.release(rid)
        got = srv.result(big_id)
        assert digest(got) == digest(engine.multiply(*big))
        srv.release(big_id)
        st = srv.stats()
        assert st["completed"] == len(ids) + 1 and st["batches"] >= 1
        for h in handles:
            srv.unregister(h)
        assert srv.stats()["charged_bytes"] == 0
    finally:
        srv.close()


def test_served_admission_and_errors():
    srv = S.Server(0, budget_bytes=1 << 20, max_batch=8)
    try:
        big = W.uniform_random(4096, 4096, 16, seed=1)  # estimate far above 1 MiB
        with pytest.raises(S.SpgError) as e:
            srv.submit(big, big)
        assert e.value.name == "SPG_ENOMEM"
        bad = S.Csr.make(1, 3, [0, 1], [7], [1.0])
        ok_b = S.Csr.make(3, 2, [0, 1, 1, 2], [0, 1], [1.0, 1.0])
        with pytest.raises(S.SpgError) as e:
            srv.submit(bad, ok_b)
        assert e.value.name == "SPG_ERANGE"
        small = S.Csr.make(2, 2, [0, 1, 2], [0, 1], [2.0, 3.0])
        rid = srv.submit(small, small)
        c = srv.result(rid)
        assert c.val.tolist() == [4.0, 9.0]
        srv.release(rid)
    finally:
        srv.close()


def _dup_heavy_pair(seed):
    """A (256 x 512, 24 distinct columns per row) times B (512 x 1024, every
    row holding columns 0..15): 384 scalar products per row (fewer than C's
    1024 columns, so no structural bound is tighter than the products) but
    only 16 entries per row of C, so C is 24x smaller than its bound."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rows, inner, cols, per, dense = 256, 512, 1024, 24, 16
    acol = np.concatenate([np.sort(rng.choice(inner, per, replace=False)) for _ in range(rows)])
    a = S.Csr.make(rows, inner, np.arange(rows + 1) * per, acol, rng.uniform(-1, 1, rows * per))
    bcol = np.tile(np.arange(dense), inner)
    b = S.Csr.make(inner, cols, np.arange(inner + 1) * dense, bcol, rng.uniform(-1, 1, inner * dense))
    return a, b


def _base_bytes(a, b):
    """An upper estimate of a product's arena and count workspace: 12 B per
    entry + 8 B per row of each operand, 12 B per A entry (entry stream),
    160 B per A row (row bins, counts, C's row offsets), 4 B per B row."""
    def operand(m):
        return 12 * int(m.row_ptr[-1]) + 8 * (m.rows + 1)
    return operand(a) + operand(b) + 12 * int(a.row_ptr[-1]) + 160 * (a.rows + 1) + 4 * (b.rows + 1)


SLACK = 256 << 10  # large-row bitmaps, small buffers and allocation rounding


def test_served_admission_by_counts(engine):
    a, b = _dup_heavy_pair(11)
    products = int(sum(b.row_ptr[c + 1] - b.row_ptr[c] for c in a.col))
    alone = engine.multiply(a, b)
    nnz = int(alone.row_ptr[-1])
    assert nnz == a.rows * 16 and products == 384 * a.rows
    base = _base_bytes(a, b)
    budget = base + 12 * nnz + SLACK
    assert base + 12 * products > budget  # the work bound would not admit it
    srv = S.Server(0, budget_bytes=budget, max_batch=8)
    try:
        rid = srv.submit(a, b)
        got = srv.result(rid)
        assert digest(got) == digest(alone)
        # held: the batch's measured buffers, C at its counted size
        held = srv.stats()["charged_bytes"]
        print("held bytes", held, "budget", budget, "work bound", 12 * products, flush=True)
        assert 12 * nnz <= held <= budget
        srv.release(rid)
        assert srv.stats()["charged_bytes"] == 0
    finally:
        srv.close()
    # in a mix: two such requests and small ones (single-pass rows, charged by
    # their work bound).  Every operand fits the budget at once but C of only
    # one of the two big requests does: the builder splits the batch or waits
    # for the first big request's release.
    smalls = [(W.uniform_random(40, 40, 3, seed=70 + k), W.uniform_random(40, 40, 3, seed=80 + k)) for k in range(4)]
    a2, b2 = _dup_heavy_pair(12)
    small_products = sum(int(sum(y.row_ptr[c + 1] - y.row_ptr[c] for c in x.col)) for x, y in smalls)
    budget = 2 * base + 12 * nnz + sum(_base_bytes(x, y) for x, y in smalls) + 12 * small_products + SLACK
    assert base + 12 * products > budget
    srv = S.Server(0, budget_bytes=budget, max_batch=8)
    try:
        ids = [srv.submit(a, b)] + [srv.submit(x, y) for x, y in smalls] + [srv.submit(a2, b2)]
        expect = [alone] + [engine.multiply(x, y) for x, y in smalls] + [engine.multiply(a2, b2)]
        for k, (rid, want) in enumerate(zip(ids, expect)):
            print("mixed request", k, flush=True)
            assert digest(srv.result(rid)) == digest(want), k
            srv.release(rid)
        assert srv.stats()["charged_bytes"] == 0
        assert srv.stats()["completed"] == len(ids)
    finally:
        srv.close()


def test_served_result_too_large_for_budget():
    # one A column of ones times one dense B row: 64 x 4096 entries of C (3 MiB)
    # from operands of a few KiB; the count phase finds C does not fit 1 MiB
    a = S.Csr.make(64, 1, np.arange(65), np.zeros(64), np.ones(64))
    b = S.Csr.make(1, 4096, [0, 4096], np.arange(4096), np.ones(4096))
    srv = S.Server(0, budget_bytes=1 << 20, max_batch=8)
    try:
        rid = srv.submit(a, b)
        with pytest.raises(S.SpgError) as e:
            srv.wait(rid)
        assert e.value.name == "SPG_ENOMEM"
        srv.release(rid)
        assert srv.stats()["charged_bytes"] == 0
        small = S.Csr.make(2, 2, [0, 1, 2], [0, 1], [2.0, 3.0])
        rid = srv.submit(small, small)
        assert srv.result(rid).val.tolist() == [4.0, 9.0]
        srv.release(rid)
    finally:
        srv.close()


def _resident_pool(srv, n, seed):
    mats = [W.uniform_random(4096, 4096, 12, seed=seed + k) for k in range(n)]
    return mats, [srv.register(m) for m in mats]


def test_served_saturated_tight_budget(engine):
    # 120 requests at once against a budget that holds a few batches of about
    # 10 MB per request: batches are bounded by the per-batch share of the
    # budget, and a batch the budget cuts short waits for a pooled batch
    # while the device is busy; every result equals the product alone.
    srv = S.Server(0, budget_bytes=128 << 20, max_batch=4096)
    try:
        mats, handles = _resident_pool(srv, 6, 40)
        pairs = [(k % 6, (k * 5 + 1) % 6) for k in range(120)]
        rids = [srv.submit_resident(handles[i], handles[j]) for i, j in pairs]
        want = {}
        for k, (rid, (i, j)) in enumerate(zip(rids, pairs)):
            got = srv.result(rid)
            if (i, j) not in want:
                want[(i, j)] = digest(engine.multiply(mats[i], mats[j]))
            assert digest(got) == want[(i, j)], k
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == 120
        assert st["batches"] >= 4  # the share and the budget bounded the batches
        for h in handles:
            srv.unregister(h)
        assert srv.stats()["charged_bytes"] == 0
    finally:
        srv.close()


def test_served_client_holds_every_result(engine):
    # A client that releases nothing until every request is done: admission
    # waits for pooled batches only while the device has a batch to run, so
    # all 60 requests complete when their results fit the budget together.
    srv = S.Server(0, budget_bytes=1 << 30, max_batch=4096)
    try:
        mats, handles = _resident_pool(srv, 4, 60)
        pairs = [(k % 4, (k + 1) % 4) for k in range(60)]
        rids = [srv.submit_resident(handles[i], handles[j]) for i, j in pairs]
        for rid in rids:
            srv.wait(rid)
        for k, (rid, (i, j)) in enumerate(zip(rids, pairs)):
            if k % 7 == 0:
                assert digest(srv.result(rid)) == digest(engine.multiply(mats[i], mats[j])), k
        for rid in rids:
            srv.release(rid)
        assert srv.stats()["completed"] == 60
    finally:
        srv.close()
