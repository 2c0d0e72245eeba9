"""GPU tests of the multi-vector evaluation (several charge vectors through one geometry in one
grouped launch, ew_shift_multi / ew_shift_multi_device).

Each vector's shift and energy must equal ew_shift of that vector alone to round-off (the
grouped launch gathers the mesh potential and adds the pair sum in one kernel, the
single-vector evaluation in two: FP64 1e-13, mixed 1e-6 relative); one vector is also
checked against the CPU SPME oracle at the engine's parameters (1e-11, FP64).
"""
import sys

import numpy as np
import pytest
from conftest import load_golden, rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-13, "mixed": 1e-6}


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def vectors(q0, nvec, seed):
    rng = np.random.default_rng(seed)
    out, q = [], q0
    for _ in range(nvec):
        q = W.scc_update(q, rng)
        out.append(q)
    return np.array(out)


def check_against_single(e, dqs, precision):
    shifts, energies = e.shift_multi(dqs)
    assert shifts.shape == dqs.shape and energies.shape == (len(dqs),)
    for v, q in enumerate(dqs):
        s1, e1 = e.shift(q)
        assert rel_max(shifts[v], s1) < TOL[precision], f"vector {v}"
        assert energies[v] == pytest.approx(e1, rel=TOL[precision]), f"vector {v}"
    return shifts, energies


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("storage", [0, 1])
def test_multi_equals_single_water(precision, storage):
    w = W.water(n_side=8)
    say(f"water 1536 atoms, 5 vectors, {precision}, pair storage {storage}")
    with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], order=6, precision=precision,
                        pair_storage=storage) as e:
        e.set_geometry(w["coords"], w["lattice"], species=w["species"],
                       hubbard_u=w["hubbard_u"], charges=w["charges"])
        assert e.parameters()["pair_mode"] == (2 if storage else 0)
        check_against_single(e, vectors(w["charges"], 5, 11), precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_equals_single_image_sum(precision):
    g = load_golden("ref_small_cell.npz")
    say(f"40 atoms, cell smaller than the cutoff sphere (image-summed pairs), {precision}")
    with ew.EwaldEngine(tol=1e-12, order=12, precision=precision) as e:
        e.set_geometry(g["coords"], g["lattice"], charges=g["charges"])
        assert e.parameters()["pair_mode"] == 1
        check_against_single(e, vectors(g["charges"], 8, 12), precision)


def test_multi_charged_cell():
    w = W.water(n_side=6)
    say("water 648 atoms, charged vectors (neutralising background), 3 vectors")
    rng = np.random.default_rng(13)
    dqs = vectors(w["charges"], 3, 14) + rng.uniform(-0.2, 0.2, (3, 1))
    with ew.EwaldEngine(tol=1e-10, charged=True) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        check_against_single(e, dqs, "fp64")
    with ew.EwaldEngine(tol=1e-10) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        with pytest.raises(ew.EwaldError) as exc:
            e.shift_multi(dqs)
        assert exc.value.code == ew.EW_ERR_NOT_NEUTRAL


def test_multi_matches_cpu_spme(oracle_mod):
    w = W.water()
    say("water 12288 atoms, 3 vectors, FP64, vector 1 against the CPU SPME oracle")
    dqs = vectors(w["charges"], 3, 15)
    with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"]) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        prm = e.parameters()
        shifts, energies = e.shift_multi(dqs)
    want = oracle_mod.spme_ewald(w["coords"], dqs[1], w["lattice"], prm["grid"], prm["order"],
                                 prm["alpha"], prm["r_cut"])
    assert rel_max(shifts[1], want) < 1e-11
    assert energies[1] == pytest.approx(0.5 * float(np.dot(dqs[1], want)), rel=1e-11)


def test_multi_vector_counts_and_replays():
    w = W.water(n_side=6)
    say("water 648 atoms: vector counts 2, 6, 3, 8, 1 and graph replays")
    dqs = vectors(w["charges"], 8, 16)
    with ew.EwaldEngine(tol=1e-10) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        ref = np.array([e.shift(q)[0] for q in dqs])
        for nv in (2, 6, 3, 8, 1, 6, 6):
            s, _ = e.shift_multi(dqs[:nv])
            assert rel_max(s, ref[:nv]) < TOL["fp64"], f"nvec {nv}"
        with pytest.raises(ew.EwaldError) as exc:
            e.shift_multi(np.tile(dqs[:1], (ew.MAX_VECTORS + 1, 1)))
        assert exc.value.code == ew.EW_ERR_ARGUMENT
        # a new geometry drops the per-vector buffers; the next call rebuilds them
        e.set_geometry(w["coords"] * 1.001, w["lattice"] * 1.001, charges=w["charges"])
        ref2 = np.array([e.shift(q)[0] for q in dqs[:4]])
        s, _ = e.shift_multi(dqs[:4])
        assert rel_max(s, ref2) < TOL["fp64"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_device_pointers(precision):
    import torch
    w = W.water(n_side=8)
    say(f"water 1536 atoms: device-pointer multi entry point, 4 vectors, {precision}")
    dqs = vectors(w["charges"], 4, 17)
    with ew.EwaldEngine(tol=1e-10, precision=precision) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        s_host, e_host = e.shift_multi(dqs)
        dq = torch.tensor(dqs, dtype=torch.float64, device="cuda")
        out = torch.empty_like(dq)
        en = torch.empty(4, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(3):
            e.shift_multi_device(4, dq.data_ptr(), out.data_ptr(), en.data_ptr())
            torch.cuda.synchronize()
            assert rel_max(out.cpu().numpy(), s_host) < TOL[precision]
            assert np.allclose(en.cpu().numpy(), e_host, rtol=TOL[precision], atol=0)
