"""GPU tests of the structure pool (grouped launches over many resident structures).

A pool evaluates each structure with the single-structure engine's kernel bodies and
variants, so every structure's shift and energy must equal the single engine's to round-off
(the pool gathers the mesh potential and adds the pair sum in one kernel, the single engine in
two: tolerance 1e-13 in FP64, 1e-6 in mixed precision, its stated accuracy), the CPU SPME oracle
at identical parameters to 1e-11 (FP64), and itself bitwise alone and in any batch: a
structure's kernels do not depend on what else is in the launch, and every sum, the spreading's
included, has a fixed order or is exact.
"""
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
def structures():
    # uneven sizes and triclinic cells as in the batched workload, kept small for the test
    return W.batch(count=10, nmin=64, nmax=3000)


TOL = {"fp64": 1e-13, "mixed": 1e-6}


def single(w, precision, dq=None):
    with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"],
                        precision=precision) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        return e.shift(w["charges"] if dq is None else dq) + (e.parameters(),)


def make_pool(ws, precision, capacity=0):
    p = ew.EwaldPool(capacity_bytes=capacity, tol=1e-10, r_cut=ws[0]["r_cut"],
                     order=ws[0]["order"], precision=precision)
    ids = [p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"]) for w in ws]
    return p, ids


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pool_matches_single_engine(structures, precision):
    ws = structures
    say(f"pool of {len(ws)} structures, N {[len(w['charges']) for w in ws]}, {precision}")
    with make_pool(ws, precision)[0] as p:
        ids = sorted(p.sizes)
        shifts, energies = p.shift(ids, [w["charges"] for w in ws])
        modes = set()
        for i, w in zip(ids, ws):
            s1, e1, prm = single(w, precision)
            modes.add(prm["pair_mode"])
            assert rel_max(shifts[i], s1) < TOL[precision]
            assert energies[i] == pytest.approx(e1, rel=TOL[precision])
        # both pair-list forms (linked tiles, image sums for small cells) are in the batch
        assert modes == {0, 1}


@pytest.mark.parametrize("order,prec", [(4, "fp64"), (9, "fp64"), (9, "mixed")])
def test_pool_other_orders(structures, order, prec):
    """The grouped kernels at a compiled order other than 6 (4) and at an order they take at run
    time (9), against the single engine at the same order and mesh.  The parity tests pin the
    single engine to the CPU SPME at order 4 and, for the run-time-order kernels, at order 12."""
    ws = structures[:5]
    say(f"pool of {len(ws)} structures at order {order}, {prec}")
    with ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=order, precision=prec) as p:
        ids = [p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"])
               for w in ws]
        shifts, energies = p.shift(ids, [w["charges"] for w in ws])
    for k, w in enumerate(ws):
        with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=order,
                            precision=prec) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            s1, e1 = e.shift(w["charges"])
        assert rel_max(shifts[k], s1) < TOL[prec], (order, k)
        assert energies[k] == pytest.approx(e1, rel=TOL[prec]), (order, k)


def test_pool_symmetric_pair_storage(structures):
    ws = structures[:5]
    say("pool with symmetric pair storage against single engines with the row list")
    with ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=6, pair_storage=1) as p:
        ids = [p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"])
               for w in ws]
        assert {p.parameters(i)["pair_mode"] for i in ids} <= {1, 2}
        shifts, energies = p.shift(ids, [w["charges"] for w in ws])
        for i, w in zip(ids, ws):
            s1, e1, _ = single(w, "fp64")
            assert rel_max(shifts[i], s1) < TOL["fp64"]
            assert energies[i] == pytest.approx(e1, rel=TOL["fp64"])


def test_pool_fp64_matches_cpu_spme(structures, oracle_mod):
    ws = structures[:4]
    say("pool FP64 against the CPU SPME oracle, 4 structures")
    with make_pool(ws, "fp64")[0] as p:
        ids = sorted(p.sizes)
        shifts, energies = p.shift(ids, [w["charges"] for w in ws])
        for i, w in zip(ids, ws):
            prm = p.parameters(i)
            want = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                         prm["order"], prm["alpha"], prm["r_cut"])
            assert rel_max(shifts[i], want) < 1e-11
            assert energies[i] == pytest.approx(0.5 * float(np.dot(w["charges"], want)),
                                                rel=1e-11)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_alone_and_in_the_mix(structures, precision):
    ws = structures
    say(f"subsets of the pool give bitwise the results of the whole batch ({precision})")
    rng = np.random.default_rng(7)
    dqs = [W.scc_update(w["charges"], rng) for w in ws]
    with make_pool(ws, precision)[0] as p:
        ids = sorted(p.sizes)
        full, efull = p.shift(ids, dqs)
        for sub in ([3], [7, 1], [9, 0, 5, 2]):
            s, e = p.shift(sub, [dqs[i] for i in sub])
            for k, i in enumerate(sub):
                assert np.array_equal(s[k], full[i]), (sub, i)
                assert e[k] == efull[i], (sub, i)


def test_device_buffers_and_graph_replays(structures):
    import torch
    ws = structures[:6]
    say("pool shift_device: graph replays against the host entry point")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    with make_pool(ws, "fp64")[0] as p:
        ids = sorted(p.sizes)
        dq_t = [torch.zeros(len(w["charges"]), dtype=torch.float64, device=dev) for w in ws]
        sh_t = [torch.zeros_like(d) for d in dq_t]
        en_t = torch.zeros(len(ws), dtype=torch.float64, device=dev)
        for rep in range(4):
            dqs = [W.scc_update(w["charges"], rng) for w in ws]
            for t, d in zip(dq_t, dqs):
                t.copy_(torch.from_numpy(d))
            torch.cuda.synchronize()
            p.shift_device(ids, [t.data_ptr() for t in dq_t], [t.data_ptr() for t in sh_t],
                           en_t.data_ptr())
            p.sync()
            want, ewant = p.shift(ids, dqs)
            for k in range(len(ws)):
                assert rel_max(sh_t[k].cpu().numpy(), want[k]) < TOL["fp64"]
            assert np.allclose(en_t.cpu().numpy(), ewant, rtol=TOL["fp64"], atol=0.0)


def test_admission_by_footprint_and_removal(structures):
    ws = structures[:3]
    say("pool capacity: admission refused, then accepted after a removal")
    with make_pool(ws[:1], "fp64")[0] as probe:
        need = probe.usage()[0]
    cap = int(need * 1.5)
    with ew.EwaldPool(capacity_bytes=cap, tol=1e-10, r_cut=ws[0]["r_cut"], order=6) as p:
        a = p.add(ws[0]["coords"], ws[0]["lattice"], charges=ws[0]["charges"], grid=ws[0]["grid"])
        used, capacity, count = p.usage()
        assert used == need and capacity == cap and count == 1
        with pytest.raises(ew.EwaldError) as exc:
            p.add(ws[0]["coords"], ws[0]["lattice"], charges=ws[0]["charges"],
                  grid=ws[0]["grid"])
        assert exc.value.code == ew.EW_ERR_MEMORY
        p.remove(a)
        b = p.add(ws[0]["coords"], ws[0]["lattice"], charges=ws[0]["charges"], grid=ws[0]["grid"])
        assert b == a and p.usage()[2] == 1
        s, e = p.shift([b], [ws[0]["charges"]])
        s1, e1, _ = single(ws[0], "fp64")
        assert rel_max(s[0], s1) < TOL["fp64"]


def test_pool_errors(structures):
    w = structures[0]
    say("pool error paths")
    with make_pool([w], "fp64")[0] as p:
        with pytest.raises(ew.EwaldError) as exc:
            p.shift([0, 0], [w["charges"], w["charges"]])
        assert exc.value.code == ew.EW_ERR_ARGUMENT
        with pytest.raises(ew.EwaldError) as exc:
            p.shift([5], [w["charges"]])
        assert exc.value.code == ew.EW_ERR_ARGUMENT
        with pytest.raises(ew.EwaldError) as exc:
            p.shift([0], [w["charges"] + 1.0])
        assert exc.value.code == ew.EW_ERR_NOT_NEUTRAL
        # vectors that do not match their structures are refused before the library reads them
        with pytest.raises(ValueError):
            p.shift([0], [w["charges"][:-1]])
        with pytest.raises(ValueError):
            p.shift([0], [w["charges"], w["charges"]])
        with pytest.raises(ew.EwaldError) as exc:
            p.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"],
                  species=np.zeros(len(w["charges"]), dtype=np.int32),
                  hubbard_u=np.array([-1.0]))
        assert exc.value.code == ew.EW_ERR_ARGUMENT
        assert p.usage()[2] == 1
