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
