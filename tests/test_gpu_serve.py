"""GPU: the serving loop (g4s_server_*).

Continuous batching: requests submitted while others run are admitted at
chunk boundaries, share the device launches, and each result equals the
same network run alone (det: the oracle's raster bit for bit); requests wait
for memory instead of exceeding the budget.
"""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_raster(oracle, n, nsyn, seed, scale, steps):
    net = oracle.Network(n, nsyn, seed, regime=1, thal_scale=scale)
    r, c, _ = net.run(steps)
    return r, c


REQS = [(1200, 300, 101, 1.0, 420), (3000, 1000, 102, 1.3, 250), (500, 499, 103, 0.9, 333),
        (8000, 100, 104, 1.5, 180), (2500, 1000, 105, 1.0, 510), (64, 20, 106, 1.2, 77)]


def test_alone_equals_mixed_det(g4s, oracle):
    """Staggered arrivals (some join while others are mid-run) -- every
    result is the oracle's raster of that network alone."""
    with g4s.Server("det", chunk=50) as srv:
        ids = []
        for k, (n, ns, sd, sc, st) in enumerate(REQS):
            ids.append(srv.submit(n, ns, sd, st, thal_scale=sc))
            time.sleep(0.01 * (k % 3))
        srv.drain(120)
        st = srv.stats()
        assert st["completed"] == len(REQS) and st["failed"] == 0
        assert st["max_active"] >= 2, st  # the requests did share launches
        for rid, (n, ns, sd, sc, steps) in zip(ids, REQS):
            info = srv.info(rid)
            assert info["state"] == "done" and info["t"] == steps
            assert info["submit_s"] <= info["admit_s"] <= info["done_s"]
            oraster, ocounts = oracle_raster(oracle, n, ns, sd, sc, steps)
            assert np.array_equal(g4s.raster_to_dense(srv.raster(rid), n), oraster), (n, ns, sd)
            assert np.array_equal(srv.counts(rid), ocounts)


def test_submit_from_threads_while_running(g4s, oracle):
    """Submissions from several threads during execution."""
    with g4s.Server("det", chunk=25) as srv:
        first = srv.submit(4000, 400, 201, 600)
        got = {}

        def worker(k):
            time.sleep(0.005 * k)
            got[k] = srv.submit(900 + 37 * k, 50 + k, 210 + k, 60 + 10 * k, thal_scale=1.1)

        th = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        srv.drain(120)
        for k, rid in got.items():
            n, ns, sd, steps = 900 + 37 * k, 50 + k, 210 + k, 60 + 10 * k
            oraster, _ = oracle_raster(oracle, n, ns, sd, 1.1, steps)
            assert np.array_equal(g4s.raster_to_dense(srv.raster(rid), n), oraster), k
        oraster, _ = oracle_raster(oracle, 4000, 400, 201, 1.0, 600)
        assert np.array_equal(g4s.raster_to_dense(srv.raster(first), 4000), oraster)


def test_budget_queues_requests(g4s, oracle):
    fp = g4s.footprint(3000, 500, chunk=50)
    with g4s.Server("det", chunk=50, budget=2 * fp + fp // 2) as srv:
        ids = [srv.submit(3000, 500, 300 + k, 120) for k in range(5)]
        srv.drain(120)
        st = srv.stats()
        assert st["completed"] == 5 and st["max_bytes"] <= 2 * fp + fp // 2 and st["max_active"] <= 2
        infos = [srv.info(r) for r in ids]
        assert max(i["admit_s"] for i in infos) > min(i["done_s"] for i in infos)  # some waited for memory
        for k, rid in enumerate(ids):
            oraster, _ = oracle_raster(oracle, 3000, 500, 300 + k, 1.0, 120)
            assert np.array_equal(g4s.raster_to_dense(srv.raster(rid), 3000), oraster)
        with pytest.raises(g4s.G4SError) as e:
            srv.submit(200000, 1000, 1, 10)  # can never fit
        assert e.value.code == -7


def test_atomic_requests_first_steps(g4s, oracle):
    with g4s.Server("atomic", chunk=100) as srv:
        ids = [srv.submit(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in REQS[:4]]
        srv.drain(120)
        for rid, (n, ns, sd, sc, st) in zip(ids, REQS[:4]):
            oraster, _ = oracle_raster(oracle, n, ns, sd, sc, st)
            dense = g4s.raster_to_dense(srv.raster(rid), n)
            assert np.array_equal(dense[:100], oraster[:100]), (n, sd)


def test_release_and_states(g4s):
    with g4s.Server("det", chunk=50) as srv:
        rid = srv.submit(1000, 100, 7, 150, keep_raster=False)
        srv.wait(rid, 60)
        assert srv.info(rid)["state"] == "done"
        assert srv.counts(rid).shape == (150,)
        with pytest.raises(g4s.G4SError):
            srv.raster(rid)
        srv.release(rid)
        with pytest.raises(g4s.G4SError):
            srv.info(rid)
        with pytest.raises(g4s.G4SError):
            srv.submit(10, 10, 1, 5)


def test_busy_server_det(g4s, oracle):
    """More requests resident than the light-load limit (192): later
    admissions take longer slices, and a second wave reuses the device
    blocks of the first; every sampled result is the oracle's."""
    rng = np.random.default_rng(7)
    reqs = [(int(rng.integers(1000, 4000)), 100, 500 + k, float(rng.uniform(0.9, 1.3)), int(rng.integers(300, 400)))
            for k in range(260)]
    wave2 = [(int(rng.integers(1000, 4000)), 100, 900 + k, 1.1, 150) for k in range(12)]
    with g4s.Server("det", chunk=50) as srv:
        ids = [srv.submit(n, ns, sd, st, thal_scale=sc, keep_raster=False) for (n, ns, sd, sc, st) in reqs]
        srv.drain(300)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["max_active"] >= 192, st
        ids2 = [srv.submit(n, ns, sd, st2, thal_scale=sc, keep_raster=False) for (n, ns, sd, sc, st2) in wave2]
        srv.drain(120)
        assert srv.stats()["completed"] == len(reqs) + len(wave2)
        sample = [(ids[k], reqs[k]) for k in (0, 1, 97, 198, 259)] + list(zip(ids2, wave2))
        for rid, (n, ns, sd, sc, steps) in sample:
            info = srv.info(rid)
            assert info["state"] == "done" and info["t"] == steps
            _, ocounts = oracle_raster(oracle, n, ns, sd, sc, steps)
            assert np.array_equal(srv.counts(rid), ocounts), (n, sd)
