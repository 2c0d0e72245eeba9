"""The batched (256 structures) and serving (1,000 resident structures) workloads at their full
size: every structure of the batch against the CPU SPME, mixed against FP64, results alone
against results in the mix, and two rounds of SCC requests for all 1,000 served structures
against the same structures evaluated alone.

Workloads (fftewald_amd.workloads): batch() -- 256 triclinic structures, N log-uniform in
[64, 8192], angles U[70, 110] deg, density 0.1 A^-3, ~0.8 A grids, order 6; serving() -- 1,000
structures, N log-uniform in [100, 20000], built like the 98,304-atom cell, ~0.777 A grids.

Tolerances (max_i |x_i - ref_i| / max_i |ref_i|): FP64 against the CPU SPME at identical
parameters 1e-11; a structure alone and in any batch or served request: bit for bit (its kernels
do not depend on the rest of the launch, and every sum has a fixed order or is exact); mixed
against FP64 1e-6 (the mixed mode's stated accuracy).  Mixed energies are compared to 1e-6 of the self
energy alpha/sqrt(pi) sum q^2: an energy is a difference of terms of that size and can be much
smaller than they are (a relative error of the energy itself then says little).
"""
import os
import sys

import numpy as np
import pytest
from conftest import rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch():
    say("generating the batch workload (256 structures)")
    return W.batch()


@pytest.fixture(scope="module")
def serving():
    say("generating the serving workload (1,000 structures)")
    return W.serving(workers=min(16, os.cpu_count() or 1))


def make_pool(ws, precision):
    p = ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=ws[0]["order"], precision=precision)
    ids = [p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"]) for w in ws]
    return p, ids


def test_batch_every_structure_matches_cpu_spme(batch, oracle_mod):
    ws = batch
    assert len(ws) == 256
    say(f"batch FP64: {len(ws)} structures, {sum(len(w['charges']) for w in ws)} atoms, "
        "one grouped evaluation against the CPU SPME")
    pool, ids = make_pool(ws, "fp64")
    with pool:
        shifts, energies = pool.shift(ids, [w["charges"] for w in ws])
        prms = [pool.parameters(i) for i in ids]
    worst = 0.0
    for w, s, e, prm in zip(ws, shifts, energies, prms):
        want = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                     prm["order"], prm["alpha"], prm["r_cut"])
        err = rel_max(s, want)
        worst = max(worst, err)
        assert err < 1e-11, (len(w["charges"]), prm["grid"], err)
        assert e == pytest.approx(0.5 * float(np.dot(w["charges"], want)), rel=1e-11)
    say(f"batch FP64 worst relative shift error against the CPU SPME: {worst:.2e}")


def test_batch_mixed_against_fp64_and_alone_against_the_mix(batch):
    ws = batch
    rng = np.random.default_rng(20260404)
    dqs = [W.scc_update(w["charges"], rng) for w in ws]
    say("batch: mixed against FP64 for all 256 structures; random subsets against the full batch")
    p64, ids64 = make_pool(ws, "fp64")
    p32, ids32 = make_pool(ws, "mixed")
    with p64, p32:
        s64, e64 = p64.shift(ids64, dqs)
        s32, e32 = p32.shift(ids32, dqs)
        e_self = [p32.parameters(i)["alpha"] / np.sqrt(np.pi) * float(np.dot(d, d))
                  for i, d in zip(ids32, dqs)]
        for i in range(len(ws)):
            assert abs(e32[i] - e64[i]) <= 1e-6 * e_self[i]
        worst = 0.0
        for a, b in zip(s32, s64):
            worst = max(worst, rel_max(a, b))
        assert worst < 1e-6
        say(f"batch mixed against FP64: worst relative shift error {worst:.2e}")
        for size in (1, 17, 100):
            sub = sorted(rng.choice(len(ws), size=size, replace=False).tolist())
            for pool, full, efull in ((p64, s64, e64), (p32, s32, e32)):
                s, e = pool.shift(sub, [dqs[i] for i in sub])
                for k, i in enumerate(sub):
                    assert np.array_equal(s[k], full[i]), (size, i)
                    assert e[k] == efull[i], (size, i)


def test_serving_two_rounds_for_every_structure(serving):
    ws = serving
    assert len(ws) == 1000
    rng = np.random.default_rng(20260505)
    rounds = 2
    dqs = []
    for w in ws:
        q, per = w["charges"], []
        for _ in range(rounds):
            q = W.scc_update(q, rng)
            per.append(q)
        dqs.append(per)
    say(f"serving mixed: {len(ws)} resident structures, "
        f"{sum(len(w['charges']) for w in ws)} atoms, {rounds} rounds of SCC requests each")
    pool, ids = make_pool(ws, "mixed")
    with pool:
        e_self = [pool.parameters(i)["alpha"] / np.sqrt(np.pi) *
                  float(np.dot(w["charges"], w["charges"])) for i, w in zip(ids, ws)]
        alone = [[pool.shift([i], [dqs[i][r]]) for i in ids] for r in range(rounds)]
        with ew.EwaldServer(pool, max_batch=64, max_inflight=2) as srv:
            # every request goes in at once: the server batches them as slots free up, a
            # structure's second request after its first
            tickets = [[srv.submit(i, dqs[i][r]) for i in ids] for r in range(rounds)]
            worst = 0.0
            for r in range(rounds):
                for k, i in enumerate(ids):
                    shift, energy = srv.result(tickets[r][k])
                    s0, e0 = alone[r][k]
                    err = rel_max(shift, s0[0])
                    worst = max(worst, err)
                    assert np.array_equal(shift, s0[0]), (i, r, err)
                    assert energy == float(e0[0]), (i, r)
            st = srv.stats()
            assert st["completed"] == rounds * len(ids) and st["failed"] == 0
            assert st["mean_batch"] > 1.0
    say(f"serving: worst relative shift error in the mix against alone {worst:.2e}, "
        f"mean batch {st['mean_batch']:.1f}")


def test_serving_sample_matches_cpu_spme(serving, oracle_mod):
    """A seeded sample of the served structures, the three largest among them, served in FP64
    and checked against the CPU SPME."""
    ws = serving
    order = np.argsort([len(w["charges"]) for w in ws])
    rng = np.random.default_rng(77)
    pick = sorted(set(order[-3:].tolist()) | set(rng.choice(len(ws), 5, replace=False).tolist()))
    sample = [ws[i] for i in pick]
    say(f"serving FP64 sample of {len(sample)} structures (N {[len(w['charges']) for w in sample]})"
        " against the CPU SPME")
    pool, ids = make_pool(sample, "fp64")
    with pool:
        with ew.EwaldServer(pool, max_batch=4, max_inflight=2) as srv:
            tickets = [srv.submit(i, w["charges"]) for i, w in zip(ids, sample)]
            results = [srv.result(t) for t in tickets]
        prms = [pool.parameters(i) for i in ids]
    for w, (shift, energy), prm in zip(sample, results, prms):
        want = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                     prm["order"], prm["alpha"], prm["r_cut"])
        assert rel_max(shift, want) < 1e-11
        assert energy == pytest.approx(0.5 * float(np.dot(w["charges"], want)), rel=1e-11)
