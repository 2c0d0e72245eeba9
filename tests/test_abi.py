"""C-ABI surface and host logic, without a GPU: the library loads, exports every symbol the
header declares, and its parameter chain agrees with the reference's (through the oracle)."""
import math
import os
import re

import numpy as np
import pytest
from conftest import ROOT

import fftewald_amd as ew
from fftewald_amd import workloads as W

HEADER = os.path.join(ROOT, "include", "fftewald_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ew_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert set(names) == set(ew.EXPORTED)
    L = ew.lib()
    for name in names:
        assert hasattr(L, name), name


def test_parameter_chain_matches_reference(oracle_mod):
    w = W.triclinic(n=3000, side=31.0)
    for tol in (1e-6, 1e-9, 1e-10):
        p = ew.ewald_parameters(w["charges"], w["lattice"], tol=tol)
        ref = oracle_mod.choose_parameters(w["lattice"], w["charges"], tol)
        assert p["r_cut"] == pytest.approx(ref["r_cut"], rel=1e-14)
        assert p["alpha"] == pytest.approx(ref["alpha"], rel=1e-13)
        assert all(k % 2 == 0 or k % 3 == 0 or k % 5 == 0 for k in p["grid"])
    # pinned cutoff: alpha follows from it (fe/fft/ewald_fft_3d_parameters.f90:114-121)
    p = ew.ewald_parameters(w["charges"], w["lattice"], tol=1e-10, r_cut=18.0)
    ref = oracle_mod.choose_parameters(w["lattice"], w["charges"], 1e-10, r_cut=18.0)
    assert p["alpha"] == pytest.approx(ref["alpha"], rel=1e-13)
    # pinned alpha: r_cut = sqrt(-ln tol)/alpha
    p = ew.ewald_parameters(w["charges"], w["lattice"], tol=1e-10, alpha=0.3)
    assert p["r_cut"] == pytest.approx(math.sqrt(-math.log(1e-10)) / 0.3, rel=1e-15)


def test_parameter_chain_matches_reference_fixture():
    """The library's chain (csrc/params.cpp) against the chain composed from the compiled
    reference functions (tests/golden/ref_truncation.npz): r_cut rule, alpha from the real-space
    budget, and the DFTB+ balanced alpha (alpha_rule "dftb").  The charges are rebuilt from
    sum q^2, which moves alpha by round-off only."""
    from conftest import load_golden
    g = load_golden("ref_truncation.npz")
    for r, lat in zip(g["chain"], g["lattice"]):
        n, q2, tol = int(r[0]), r[1], r[2]
        q = np.full(n, math.sqrt(q2 / n))
        p = ew.ewald_parameters(q, lat, tol=tol)
        assert p["r_cut"] == pytest.approx(r[4], rel=1e-15)
        assert p["alpha"] == pytest.approx(r[5], rel=1e-13)
        p = ew.ewald_parameters(q, lat, tol=tol, alpha_rule="dftb")
        assert p["alpha"] == pytest.approx(r[9], rel=1e-8)


def test_dftb_alpha_rule_matches_getOptimalAlphaEwald(oracle_mod):
    w = W.nacl()
    p = ew.ewald_parameters(w["charges"], w["lattice"], tol=1e-10, alpha_rule="dftb",
                            grid=16, order=8)
    a = oracle_mod.optimal_alpha_dftb(w["lattice"], 1e-10)
    assert p["alpha"] == pytest.approx(a, rel=1e-14)
    assert p["r_cut"] == pytest.approx(math.sqrt(-math.log(1e-10)) / a, rel=1e-14)


def test_grid_memory_grows_with_size_and_precision():
    small = W.triclinic(n=2000, side=27.0)
    big = W.triclinic(n=16000, side=54.0)
    a = ew.ewald_grid_memory(small["charges"], small["lattice"], tol=1e-8)
    b = ew.ewald_grid_memory(big["charges"], big["lattice"], tol=1e-8)
    c = ew.ewald_grid_memory(big["charges"], big["lattice"], tol=1e-8, precision="mixed")
    assert 0 < a < b and c < b


def test_parameter_errors():
    lat = np.eye(3) * 20.0
    q = np.array([1.0, -1.0])
    with pytest.raises(ew.EwaldError) as e:
        ew.ewald_parameters(q, lat, tol=1.5)
    assert e.value.code == ew.EW_ERR_TOLERANCE
    with pytest.raises(ew.EwaldError) as e:
        ew.ewald_parameters(q, np.zeros((3, 3)))
    assert e.value.code == ew.EW_ERR_CELL
    with pytest.raises(ew.EwaldError) as e:
        ew.ewald_parameters(q, lat, grid=(7, 7, 7))  # 7 is not 2^a 3^b 5^c
    assert e.value.code == ew.EW_ERR_ARGUMENT


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ew.EwaldError) as e:
        ew.EwaldEngine()
    assert e.value.code == ew.EW_ERR_DEVICE


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with a message instead of running."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_k_cut_override_sets_the_mesh(oracle_mod):
    """k_cut_in (fe/fft/ewald_fft_3d.f90:36-76, parameters.f90:127-130): the given Fourier cutoff
    replaces the budget's and the mesh follows from it; the chain's own cutoff reproduces the
    automatic mesh, a doubled one roughly doubles it, an explicit grid still wins, and a
    non-finite value is refused."""
    w = W.triclinic(n=3000, side=31.0)
    q, lat = w["charges"], w["lattice"]
    for tol in (1e-6, 1e-9, 1e-10):
        auto = ew.ewald_parameters(q, lat, tol=tol)
        m = oracle_mod.cell_metrics(lat)
        kc = oracle_mod.cutoff_from_fourier_budget(float(np.sum(q * q)), len(q),
                                                   m["volume"] ** (1.0 / 3.0), auto["alpha"], tol)
        same = ew.ewald_parameters(q, lat, tol=tol, k_cut=kc)
        assert same["grid"] == auto["grid"] and same["alpha"] == auto["alpha"]
        big = ew.ewald_parameters(q, lat, tol=tol, k_cut=2.0 * kc)
        for a in range(3):
            assert 1.8 * auto["grid"][a] <= big["grid"][a] <= 2.6 * auto["grid"][a]
        # the mesh resolves the requested cutoff along every axis
        assert big["k_cut"] >= 2.0 * kc
        pinned = ew.ewald_parameters(q, lat, tol=tol, k_cut=2.0 * kc, grid=(40, 40, 40))
        assert pinned["grid"] == (40, 40, 40)
    with pytest.raises(ew.EwaldError) as e:
        ew.ewald_parameters(q, lat, tol=1e-8, k_cut=float("nan"))
    assert e.value.code == ew.EW_ERR_ARGUMENT


def test_fortran_k_cut_in_reaches_the_mesh(tmp_path):
    """The Fortran ewald_parameters passes k_cut_in through (it used to drop it)."""
    import subprocess
    from conftest import PKG
    binp = os.path.join(ROOT, "tests", "fortran", "bin", "fortran_check")
    if not os.path.exists(binp):
        pytest.fail(f"{binp} not built (run __graft_entry__.build())")
    w = W.triclinic(n=3000, side=31.0)
    inp, out = tmp_path / "p_in.bin", tmp_path / "p_out.bin"
    auto = ew.ewald_parameters(w["charges"], w["lattice"], tol=1e-8)
    with open(inp, "wb") as f:
        np.array([len(w["charges"])], dtype=np.int32).tofile(f)
        w["charges"].astype(np.float64).tofile(f)
        w["lattice"].astype(np.float64).tofile(f)
        np.array([1e-8, 2.0 * auto["k_cut"]], dtype=np.float64).tofile(f)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(PKG, "lib") + ":" + env.get("LD_LIBRARY_PATH", "")
    r = subprocess.run([binp, str(inp), str(out), "params"], env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.fromfile(out, dtype=np.int32, count=6)
    kc = np.fromfile(out, dtype=np.float64, offset=24)
    assert tuple(g[:3]) == auto["grid"]
    want = ew.ewald_parameters(w["charges"], w["lattice"], tol=1e-8, k_cut=2.0 * auto["k_cut"])
    assert tuple(g[3:]) == want["grid"] and tuple(g[3:]) != tuple(g[:3])
    assert kc[1] == pytest.approx(want["k_cut"], rel=1e-15)
