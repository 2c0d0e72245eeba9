"""The Fortran host side (dftb_fast-ewald_3d_amd/fortran/fftewald_amd.f90) drives the GPU engine
through iso_c_binding, with the reference's own ewald_potential_force signature and with
DFTB+'s native arrays for the resident engine (tests/fortran/fortran_check.f90)."""
import os
import subprocess

import numpy as np
import pytest
from conftest import PKG, ROOT, load_golden, rel_max

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "fortran", "bin", "fortran_check")


def test_fortran_host_matches_reference(tmp_path, oracle_mod):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    g = load_golden("ref_nacl_c1.npz")
    n = len(g["charges"])
    inp, out = tmp_path / "input.bin", tmp_path / "output.bin"
    with open(inp, "wb") as f:
        np.array([n], dtype=np.int32).tofile(f)
        g["coords"].astype(np.float64).tofile(f)          # C [n][3] == Fortran coords(3,n)
        g["charges"].astype(np.float64).tofile(f)
        g["lattice"].astype(np.float64).tofile(f)         # C rows a_i == Fortran columns a_i
        np.array([1e-10, float(g["alpha"])], dtype=np.float64).tofile(f)
        np.array([16, 8, len(g["hubbard_u"])], dtype=np.int32).tofile(f)
        g["species"].astype(np.int32).tofile(f)
        g["hubbard_u"].astype(np.float64).tofile(f)
    print("\n[gpu-case] fortran_check nacl 512 atoms", flush=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(PKG, "lib") + ":" + env.get("LD_LIBRARY_PATH", "")
    r = subprocess.run([BIN, str(inp), str(out)], env=env, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    data = np.fromfile(out, dtype=np.float64)
    pot, shift, energy = data[:n], data[n:2 * n], data[2 * n]
    assert rel_max(pot, g["pot"]) < 1e-10
    sg = oracle_mod.short_gamma_shift(g["coords"], g["charges"], g["lattice"], g["species"],
                                      g["hubbard_u"])
    want = g["pot"] + sg
    assert rel_max(shift, want) < 1e-10
    assert energy == pytest.approx(0.5 * float(np.dot(g["charges"], want)), rel=1e-10)


def test_fortran_broyden_mixer_reference_vectors(tmp_path):
    """DFTB+'s own Broyden regression vectors (test/src/dftbp/unit/mixer/broyden.F90:25-64)
    through the Fortran host side and the device mixer, within that test's atol = 100 eps."""
    import json
    from conftest import GOLDEN
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    with open(os.path.join(GOLDEN, "ref_broyden_h2o.json")) as f:
        g = json.load(f)
    p = g["params"]
    n, niter = 3, len(g["difference"])
    inp, out = tmp_path / "mixer_in.bin", tmp_path / "mixer_out.bin"
    with open(inp, "wb") as f:
        np.array([n, niter, p["maxSccIter"]], dtype=np.int32).tofile(f)
        np.array([p["mixParam"], p["omega0"], p["minWeight"], p["maxWeight"], p["weightFac"]],
                 dtype=np.float64).tofile(f)
        np.array(g["expected_charges"][0], dtype=np.float64).tofile(f)
        np.array(g["difference"], dtype=np.float64).tofile(f)   # C [iter][n] == Fortran (n, iter)
    print("\n[gpu-case] fortran_check mixer: reference Broyden vectors", flush=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(PKG, "lib") + ":" + env.get("LD_LIBRARY_PATH", "")
    r = subprocess.run([BIN, str(inp), str(out), "mixer"], env=env, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    hist = np.fromfile(out, dtype=np.float64).reshape(niter, n)
    want = np.array(g["expected_charges"][1:])
    assert np.abs(hist - want).max() <= g["atol_eps"] * np.finfo(np.float64).eps


def test_fortran_external_charges(tmp_path, oracle_mod):
    """TCoulomb%getPotential / %getPotentialGradient / %addExternalPotGrad as DFTB+ would call
    them from Fortran (chargeCoords(3, m), deriv(3, N)), against the CPU restatement of the
    reference's ewald() and its derivatives (oracle/external.py)."""
    from oracle.external import external_gradients, external_potential
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")
    g = load_golden("ref_triclinic.npz")
    x, q, lat = g["coords"], g["charges"], g["lattice"]
    n = len(q)
    rng = np.random.default_rng(61)
    m = 6
    y = rng.random((m, 3)) @ lat
    qe = rng.normal(size=m)
    blur = rng.uniform(0.4, 1.0, m)
    eps = 0.3
    inp, out = tmp_path / "ext_in.bin", tmp_path / "ext_out.bin"
    with open(inp, "wb") as f:
        np.array([n], dtype=np.int32).tofile(f)
        x.astype(np.float64).tofile(f)
        q.astype(np.float64).tofile(f)
        lat.astype(np.float64).tofile(f)
        np.array([1e-12], dtype=np.float64).tofile(f)
        np.array([m], dtype=np.int32).tofile(f)
        y.tofile(f)
        qe.tofile(f)
        blur.tofile(f)
        np.array([eps], dtype=np.float64).tofile(f)
    print("\n[gpu-case] fortran_check external: 300 atoms, 6 point charges", flush=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(PKG, "lib") + ":" + env.get("LD_LIBRARY_PATH", "")
    r = subprocess.run([BIN, str(inp), str(out), "external"], env=env, capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    data = np.fromfile(out, dtype=np.float64)
    pot, potb = data[:n], data[n:2 * n]
    agrad = data[2 * n:5 * n].reshape(n, 3)
    egrad = data[5 * n:5 * n + 3 * m].reshape(m, 3)
    pgrad = data[5 * n + 3 * m:].reshape(n, 3)
    a_bal = np.sqrt(np.pi) / abs(np.linalg.det(lat)) ** (1.0 / 3.0)
    assert rel_max(pot, external_potential(x, y, qe, lat, alpha=a_bal)) < 1e-10
    d0, d1 = external_gradients(x, y, qe, lat, a_bal, target_charges=q)
    h0, _ = external_gradients(x, y, qe, lat, a_bal)
    assert rel_max(agrad, d0) < 1e-9
    assert rel_max(egrad, d1) < 1e-9
    assert rel_max(pgrad, h0) < 1e-9
    # the softened potential depends on alpha: the restatement at the engine's own alpha
    import fftewald_amd as ew
    with ew.EwaldEngine(tol=1e-12) as e:
        e.set_geometry(x, lat, charges=q)
        alpha = e.parameters()["alpha"]
    want = external_potential(x[:60], y, qe, lat, alpha=alpha, blur=blur, eps_soften=eps)
    assert rel_max(potb[:60], want) < 1e-10
