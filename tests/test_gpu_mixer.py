"""GPU tests of the device charge mixer (ew_mixer_*), through the C ABI.

Tolerances:
  * the reference's own regression vectors (test/src/dftbp/unit/mixer/broyden.F90:25-64):
    atol = 100 eps, the reference test's own bound;
  * against the CPU restatement (oracle/broyden.py) on long vectors: 1e-9 relative per iterate
    and 1e-13 once converged.  The device sums its inner products in another order; two CPU
    restatements that differ only in that order (pairwise vs exactly rounded sums) already
    differ by up to 5e-11 in the iterations just before convergence, where the small Broyden
    system is ill-conditioned, and agree to 1e-15 again after it;
  * a device-resident SCC loop (ew_shift_device -> model response -> ew_mixer_mix_device)
    against the same loop on the CPU (SPME oracle + Broyden oracle): 1e-9 relative per iterate.
"""
import json
import os
import sys

import numpy as np
import pytest
from conftest import GOLDEN, rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W
from oracle.broyden import BroydenMixer, SimpleMixer

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def golden():
    with open(os.path.join(GOLDEN, "ref_broyden_h2o.json")) as f:
        return json.load(f)


def test_reference_regression_vectors_on_device():
    g = golden()
    p = g["params"]
    say("Broyden mixer: reference water regression vectors (4 iterations)")
    with ew.EwaldMixer(mix_param=p["mixParam"], max_iter=p["maxSccIter"], omega0=p["omega0"],
                       min_weight=p["minWeight"], max_weight=p["maxWeight"],
                       weight_fac=p["weightFac"]) as m:
        m.reset(3)
        q = np.array(g["expected_charges"][0])
        for i, d in enumerate(g["difference"]):
            q = m.mix(q, np.array(d))
            want = np.array(g["expected_charges"][i + 1])
            assert np.abs(q - want).max() <= g["atol_eps"] * EPS, f"iteration {i + 1}"
        assert m.iteration == 4


def _linear_problem(n, seed):
    rng = np.random.default_rng(seed)
    k = 12
    v = np.linalg.qr(rng.normal(size=(n, k)))[0]       # a few slow modes
    lam = np.linspace(0.5, 0.97, k)
    b = rng.normal(size=n)
    return lambda q: (v * lam) @ (v.T @ q) - 0.05 * q + b - q


@pytest.mark.parametrize("n", [1000, 100_003])
def test_broyden_matches_oracle_long_vectors(n):
    say(f"Broyden mixer against the CPU restatement, n = {n}, 25 iterations")
    resid = _linear_problem(n, 7 + n)
    cpu = BroydenMixer(0.2, 40)
    cpu.reset(n)
    with ew.EwaldMixer(mix_param=0.2, max_iter=40) as m:
        m.reset(n)
        q_gpu = np.zeros(n)
        q_cpu = np.zeros(n)
        for it in range(25):
            d = resid(q_cpu)
            q_cpu = cpu.mix(q_cpu, d)
            q_gpu = m.mix(q_gpu, resid(q_gpu))
            assert rel_max(q_gpu, q_cpu) < 1e-9, f"iteration {it + 1}"
        assert rel_max(q_gpu, q_cpu) < 1e-13
        assert np.abs(resid(q_gpu)).max() < 1e-12 * np.abs(q_gpu).max()


def test_simple_mixer_and_device_pointers():
    import torch
    n = 5000
    say("simple mixer and the device-pointer entry point")
    rng = np.random.default_rng(11)
    q = rng.normal(size=n)
    d = rng.normal(size=n)
    with ew.EwaldMixer(kind="simple", mix_param=0.3) as m:
        m.reset(n)
        out = m.mix(q, d)
        # the device fuses the multiply-add (one rounding instead of two)
        bound = 2 * EPS * (np.abs(q) + 0.3 * np.abs(d))
        assert np.all(np.abs(out - SimpleMixer(0.3).mix(q, d)) <= bound)
    with ew.EwaldMixer(mix_param=0.2, max_iter=10) as m_dev, \
            ew.EwaldMixer(mix_param=0.2, max_iter=10) as m_host:
        m_dev.reset(n)
        m_host.reset(n)
        tq = torch.tensor(q, dtype=torch.float64, device="cuda")
        td = torch.empty_like(tq)
        qh = q.copy()
        for it in range(6):
            dd = np.sin(qh * (it + 1)) * 0.1
            td.copy_(torch.from_numpy(dd))
            torch.cuda.synchronize()
            m_dev.mix_device(tq.data_ptr(), td.data_ptr(), torch.cuda.current_stream().cuda_stream)
            qh = m_host.mix(qh, dd)
            torch.cuda.synchronize()
            # same kernels, same order: the two entry points agree bit for bit
            assert np.array_equal(tq.cpu().numpy(), qh), f"iteration {it + 1}"


def test_mixer_errors():
    say("mixer error paths")
    with ew.EwaldMixer(max_iter=2) as m:
        with pytest.raises(ew.EwaldError) as e:
            m.mix(np.zeros(3), np.zeros(3))
        m.reset(3)
        q = np.array([1.0, 2.0, 3.0])
        q = m.mix(q, np.array([0.1, 0.0, -0.1]))
        q = m.mix(q, np.array([0.05, 0.01, -0.06]))
        with pytest.raises(ew.EwaldError) as e:
            m.mix(q, np.array([0.01, 0.0, -0.01]))
        assert e.value.code == ew.EW_ERR_ARGUMENT and "Maximal" in str(e.value)
        m.reset(5)  # a new cycle with another length
        assert m.iteration == 0
        m.mix(np.ones(5), np.full(5, 0.1))
        with pytest.raises(ew.EwaldError) as e:
            m.mix(np.array([np.nan, 0, 0, 0, 0]), np.zeros(5))
        assert e.value.code == ew.EW_ERR_NOT_FINITE
    with pytest.raises(ew.EwaldError):
        ew.EwaldMixer(max_iter=5000)
    with pytest.raises(ew.EwaldError):
        ew.EwaldMixer(mix_param=-1.0)


def test_device_scc_loop_matches_cpu_loop(oracle_mod):
    """An SCC cycle that never leaves the device: shift(dq) on the resident engine, a model
    charge response dq_out = q0 - chi (shift - mean(shift)), and the Broyden mix of
    (dq, dq_out - dq); the same loop on the CPU with the SPME oracle and the Broyden oracle."""
    import torch
    w = W.water(n_side=6)
    n = len(w["charges"])
    chi = 0.08
    iters = 14
    say(f"device SCC loop: water {n} atoms, Broyden, {iters} iterations, against the CPU loop")
    q0 = w["charges"]
    with ew.EwaldEngine(tol=1e-10) as e, ew.EwaldMixer(mix_param=0.2, max_iter=40) as m:
        e.set_geometry(w["coords"], w["lattice"], charges=q0)
        prm = e.parameters()
        m.reset(n)
        dev = torch.device("cuda")
        tq0 = torch.tensor(q0, dtype=torch.float64, device=dev)
        tq = tq0.clone()
        tshift = torch.empty_like(tq)
        tdiff = torch.empty_like(tq)
        ten = torch.empty(1, dtype=torch.float64, device=dev)
        sp = torch.cuda.current_stream().cuda_stream
        gpu_iterates = []
        for _ in range(iters):
            e.shift_device(tq.data_ptr(), tshift.data_ptr(), ten.data_ptr(), sp)
            tdiff.copy_(tq0 - chi * (tshift - tshift.mean()) - tq)
            m.mix_device(tq.data_ptr(), tdiff.data_ptr(), sp)
            gpu_iterates.append(tq.cpu().numpy())
    cpu = BroydenMixer(0.2, 40)
    cpu.reset(n)
    q = q0.copy()
    cpu_iterates = []
    for _ in range(iters):
        shift = oracle_mod.spme_ewald(w["coords"], q, w["lattice"], prm["grid"], prm["order"],
                                      prm["alpha"], prm["r_cut"])
        q = cpu.mix(q, q0 - chi * (shift - shift.mean()) - q)
        cpu_iterates.append(q)
    for it, (a, b) in enumerate(zip(gpu_iterates, cpu_iterates)):
        assert rel_max(a, b) < 1e-9, f"iteration {it + 1}"
    # the cycle converged: the last charges are a fixed point of the response
    last = cpu_iterates[-1]
    shift = oracle_mod.spme_ewald(w["coords"], last, w["lattice"], prm["grid"], prm["order"],
                                  prm["alpha"], prm["r_cut"])
    resid = q0 - chi * (shift - shift.mean()) - last
    assert np.abs(resid).max() < 1e-8 * np.abs(last).max()
