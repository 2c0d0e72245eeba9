"""GPU tests of the serving loop (continuous batching over a structure pool).

Each request's result must equal the same structure evaluated alone by the pool bit for bit,
whatever batch the server put it in (a structure's kernels do not depend on the rest of the
launch, and every sum has a fixed order or is exact); requests for one structure are evaluated
in submission order.
"""
import sys
import threading

import numpy as np
import pytest

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu



def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def structures():
    return W.batch(seed=W.SEED0 + 14, count=12, nmin=64, nmax=2000)


def make_pool(ws, precision):
    p = ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=6, precision=precision)
    for w in ws:
        p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"])
    return p


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_alone_equal_results_in_the_mix(structures, precision):
    ws = structures
    rng = np.random.default_rng(3)
    rounds = 3
    dqs = [[None] * len(ws) for _ in range(rounds)]
    for i, w in enumerate(ws):
        q = w["charges"]
        for r in range(rounds):
            q = W.scc_update(q, rng)
            dqs[r][i] = q
    say(f"server over {len(ws)} structures x {rounds} SCC updates, {precision}")
    with make_pool(ws, precision) as pool:
        ids = sorted(pool.sizes)
        alone = [[pool.shift([i], [dqs[r][i]]) for i in ids] for r in range(rounds)]
        with ew.EwaldServer(pool, max_batch=5, max_inflight=2) as srv:
            tickets = [[srv.submit(i, dqs[r][i]) for i in ids] for r in range(rounds)]
            for r in range(rounds):
                for k, i in enumerate(ids):
                    shift, energy = srv.result(tickets[r][k])
                    s0, e0 = alone[r][k]
                    assert np.array_equal(shift, s0[0])
                    assert energy == float(e0[0])
            st = srv.stats()
            assert st["completed"] == rounds * len(ids) and st["failed"] == 0
            # requests that arrived together were batched (at most 5 per launch)
            assert 1.0 < st["mean_batch"] <= 5.0
            assert st["batches"] >= rounds * len(ids) / 5
            assert 0.0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]


def test_two_submitting_threads(structures):
    ws = structures[:8]
    say("server: two host threads submitting concurrently")
    rng = np.random.default_rng(5)
    dq = [W.scc_update(w["charges"], rng) for w in ws]
    with make_pool(ws, "fp64") as pool:
        ids = sorted(pool.sizes)
        alone = {i: pool.shift([i], [dq[i]]) for i in ids}
        results = {}
        with ew.EwaldServer(pool, max_batch=8, max_inflight=2) as srv:
            def client(mine):
                for _ in range(4):
                    t = [(i, srv.submit(i, dq[i])) for i in mine]
                    for i, tk in t:
                        results.setdefault(i, []).append(srv.result(tk))
            th = [threading.Thread(target=client, args=(ids[0::2],)),
                  threading.Thread(target=client, args=(ids[1::2],))]
            for t in th:
                t.start()
            for t in th:
                t.join()
        for i in ids:
            assert len(results[i]) == 4
            for shift, energy in results[i]:
                assert np.array_equal(shift, alone[i][0][0])


def test_close_finishes_submitted_requests(structures):
    ws = structures[:6]
    say("server: destroy completes what was submitted")
    with make_pool(ws, "mixed") as pool:
        ids = sorted(pool.sizes)
        srv = ew.EwaldServer(pool, max_batch=2, max_inflight=1)
        tickets = [srv.submit(i, ws[i]["charges"]) for i in ids for _ in range(2)]
        outs = [srv._out[t] for t in tickets]
        srv.close()
        for shift, energy in outs:
            assert np.all(np.isfinite(shift)) and np.isfinite(energy[0])


def test_server_errors(structures):
    w = structures[0]
    say("server error paths")
    with make_pool([w], "fp64") as pool:
        with ew.EwaldServer(pool) as srv:
            with pytest.raises(ew.EwaldError) as exc:
                srv.submit(3, w["charges"])
            assert exc.value.code == ew.EW_ERR_ARGUMENT
            with pytest.raises(ew.EwaldError) as exc:
                srv.submit(0, w["charges"] + 0.5)
            assert exc.value.code == ew.EW_ERR_NOT_NEUTRAL
            t = srv.submit(0, w["charges"])
            assert srv.result(t) is not None
