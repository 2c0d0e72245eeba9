"""GPU parity of the HIP engine, through its C ABI.

Tolerances (max_i |shift_i - ref_i| / max_i |ref_i| for shifts; relative for energies):
  * against the reference's own direct Ewald (golden fixtures)           1e-10
  * against the CPU SPME oracle at identical parameters (FP64)          1e-11
  * mixed precision against the FP64 engine                              1e-6
"""
import math
import sys

import numpy as np
import pytest
from conftest import load_golden, rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def engine_shift(w, species=None, hub=None, dq=None, **opts):
    with ew.EwaldEngine(**opts) as e:
        e.set_geometry(w["coords"], w["lattice"], species=species, hubbard_u=hub,
                       charges=w["charges"])
        s, en = e.shift(w["charges"] if dq is None else dq)
        return s, en, e.parameters()


def test_c1_nacl_matches_reference_direct_ewald():
    g = load_golden("ref_nacl_c1.npz")
    say("nacl 512 atoms, alpha=getOptimalAlphaEwald(1e-10), grid 16^3 order 8, FP64")
    s, en, prm = engine_shift(g, tol=1e-10, alpha=float(g["alpha"]), grid=16, order=8)
    assert prm["pair_mode"] == 1  # r_cut = 76 Bohr > cell: image-summed pairs
    e_ref = 0.5 * float(np.dot(g["charges"], g["pot"]))
    assert rel_max(s, g["pot"]) < 1e-10
    assert en == pytest.approx(e_ref, rel=1e-10)


def test_c1_alpha_rule_dftb_matches_fixture_alpha():
    g = load_golden("ref_nacl_c1.npz")
    say("nacl 512 atoms, alpha_rule=dftb")
    s, en, prm = engine_shift(g, tol=1e-10, alpha_rule="dftb", grid=16, order=8)
    assert prm["alpha"] == pytest.approx(float(g["alpha"]), rel=1e-14)
    assert rel_max(s, g["pot"]) < 1e-10


def test_madelung_on_gpu():
    g = load_golden("ref_madelung.npz")
    say("ideal rocksalt, tol 1e-12, automatic parameters")
    s, en, prm = engine_shift(g, tol=1e-12)
    m = -en / (len(g["charges"]) / 2) * 2.82 * W.AA
    assert m == pytest.approx(1.747564594633182, abs=1e-10)
    assert rel_max(s, g["pot"]) < 1e-10


def test_triclinic_reference_own_alpha():
    g = load_golden("ref_triclinic.npz")
    say("300 atoms sheared cell, reference balanced alpha, tol 1e-12")
    s, en, prm = engine_shift(g, tol=1e-12, alpha=float(g["alpha"]), order=12)
    assert rel_max(s, g["pot"]) < 1e-10
    assert en == pytest.approx(0.5 * float(np.dot(g["charges"], g["pot"])), rel=1e-10)


def test_small_cell_image_sum():
    g = load_golden("ref_small_cell.npz")
    say("40 atoms, cell smaller than the cutoff sphere")
    s, en, prm = engine_shift(g, tol=1e-12, order=12)
    assert prm["pair_mode"] == 1
    assert rel_max(s, g["pot"]) < 1e-10


def test_c1_short_gamma_scc_shift(oracle_mod):
    g = load_golden("ref_nacl_c1.npz")
    say("nacl 512 atoms + short gamma (U Na 0.2, Cl 0.3)")
    s, en, prm = engine_shift(g, species=g["species"], hub=g["hubbard_u"], tol=1e-10,
                              alpha=float(g["alpha"]), grid=16, order=8)
    sg = oracle_mod.short_gamma_shift(g["coords"], g["charges"], g["lattice"], g["species"],
                                      g["hubbard_u"])
    want = g["pot"] + sg
    assert rel_max(s, want) < 1e-10
    assert en == pytest.approx(0.5 * float(np.dot(g["charges"], want)), rel=1e-10)


def _spme_reference(oracle_mod, w, prm, species=None, hub=None):
    pot = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                prm["order"], prm["alpha"], prm["r_cut"])
    if species is not None:
        pot = pot + oracle_mod.short_gamma_shift(w["coords"], w["charges"], w["lattice"],
                                                 species, hub)
    return pot


def test_c2_water_fp64_matches_cpu_spme(oracle_mod):
    w = W.water()
    say("water 12288 atoms, grid 64^3, order 6, r_cut 10 A, U O 0.4 H 0.3, FP64")
    s, en, prm = engine_shift(w, species=w["species"], hub=w["hubbard_u"], tol=1e-10,
                              r_cut=w["r_cut"], grid=w["grid"], order=w["order"])
    assert prm["pair_mode"] == 0
    want = _spme_reference(oracle_mod, w, prm, w["species"], w["hubbard_u"])
    assert rel_max(s, want) < 1e-11
    assert en == pytest.approx(0.5 * float(np.dot(w["charges"], want)), rel=1e-11)


def test_c2_water_against_reference_direct_ewald():
    """The 12,288-atom water box against the reference's own direct Ewald (tol 1e-10,
    direct/ewald_direct_3d.f90:37-182; tests/golden/ref_water_c2.npz), Coulomb only.

    * the workload's production settings (64^3 mesh, order 6, r_cut 10 A, tol 1e-10): the SPME
      discretisation error is 8.6e-6 in the shifts and 2.7e-6 in the energy (measured with the
      CPU SPME at the same parameters); asserted below 2e-5 / 1e-5;
    * tightened settings (128^3, order 12, r_cut 10 A, tol 1e-12): <= 1e-10 in both, as on
      the 512-atom cell (the remaining 1.1e-11 is the direct sum's own truncation)."""
    g = load_golden("ref_water_c2.npz")
    w = {"coords": g["coords"], "lattice": g["lattice"], "charges": g["charges"]}
    e_ref = 0.5 * float(np.dot(g["charges"], g["pot"]))
    say("c2 water 12,288 atoms, 64^3 order 6 (production) against the reference direct Ewald")
    s, en, _ = engine_shift(w, tol=1e-10, grid=64, order=6, r_cut=10.0 * W.AA)
    err, err_e = rel_max(s, g["pot"]), abs(en - e_ref) / abs(e_ref)
    say(f"production settings: shift {err:.2e}, energy {err_e:.2e}")
    assert err < 2e-5 and err_e < 1e-5
    say("c2 water 12,288 atoms, 128^3 order 12 tol 1e-12 against the reference direct Ewald")
    s, en, _ = engine_shift(w, tol=1e-12, grid=128, order=12, r_cut=10.0 * W.AA)
    err, err_e = rel_max(s, g["pot"]), abs(en - e_ref) / abs(e_ref)
    say(f"tightened settings: shift {err:.2e}, energy {err_e:.2e}")
    assert err < 1e-10 and err_e < 1e-10


@pytest.mark.parametrize("K,order", [((45, 27, 50), 5), ((24, 30, 20), 4), ((36, 36, 36), 10)])
def test_odd_grids_and_orders(oracle_mod, K, order):
    rng = np.random.default_rng(sum(K) + order)
    lat = W.lattice_from_parameters(22.0 * W.AA, 19.0 * W.AA, 24.0 * W.AA, 84.0, 99.0, 71.0)
    n = 900
    w = {"coords": W.poisson_disc(n, lat, 1.0 * W.AA, rng), "lattice": lat,
         "charges": W.neutralise(rng.uniform(-1, 1, n))}
    say(f"900 atoms triclinic, grid {K}, order {order}")
    s, en, prm = engine_shift(w, tol=1e-8, grid=K, order=order)
    want = _spme_reference(oracle_mod, w, prm)
    assert rel_max(s, want) < 1e-11


def test_c3_fp64_and_mixed(oracle_mod):
    w = W.triclinic()
    say("triclinic 98304 atoms, grid 128^3, order 6, r_cut 10 A, FP64 then mixed")
    s64, e64, prm = engine_shift(w, tol=1e-10, r_cut=w["r_cut"], grid=w["grid"],
                                 order=w["order"])
    want = _spme_reference(oracle_mod, w, prm)
    assert rel_max(s64, want) < 1e-11
    assert e64 == pytest.approx(0.5 * float(np.dot(w["charges"], want)), rel=1e-11)
    say("triclinic 98304 atoms, mixed precision")
    s32, e32, prm32 = engine_shift(w, tol=1e-10, r_cut=w["r_cut"], grid=w["grid"],
                                   order=w["order"], precision="mixed")
    # default FFT form on the 128^3 mesh: the fused x-y plane kernels (register transforms) in
    # both precisions
    assert prm32["fft_planes"] == 1 and prm["fft_planes"] == 1
    err = rel_max(s32, s64)
    print(f"mixed vs FP64: shift {err:.3e}  energy {abs(e32 - e64) / abs(e64):.3e}")
    assert err < 1e-6
    # E = -219 Ha is what is left of E_self = -4437 Ha against E_real + E_rec: the FP32 mesh
    # error of E_rec (~1e-7 of it) shows up ~20x larger relative to the total.
    assert e32 == pytest.approx(e64, rel=1e-5)


def test_shift_device_and_repeated_calls():
    import torch
    w = W.water(n_side=8)
    say("water 1536 atoms: device-pointer entry point, repeated SCC iterations")
    rng = np.random.default_rng(7)
    with ew.EwaldEngine(tol=1e-8) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        q = w["charges"]
        for it in range(3):
            q = W.scc_update(q, rng)
            s_host, e_host = e.shift(q)
            dq = torch.tensor(q, dtype=torch.float64, device="cuda")
            out = torch.empty_like(dq)
            en = torch.empty(1, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            e.shift_device(dq.data_ptr(), out.data_ptr(), en.data_ptr())
            torch.cuda.synchronize()
            assert rel_max(out.cpu().numpy(), s_host) < 1e-14
            assert float(en.item()) == pytest.approx(e_host, rel=1e-14)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_repeated_evaluations_are_bitwise_identical(precision):
    """Every step of an evaluation sums in a fixed order, and the spreading accumulates each mesh
    point in fixed point (integer adds commute), so the same charges give bitwise the same shifts
    and energy however the lanes interleave: on a 64^3 mesh (owned spreading blocks, register
    plane transforms) and on small cells (per-tile spreading, Stockham passes)."""
    rng = np.random.default_rng(11)
    cases = [("water 1536, 64^3", W.water(n_side=8), dict(grid=64, order=6))]
    cases += [(f"small cell {i}", w, dict(grid=w["grid"], order=w["order"]))
              for i, w in enumerate(W.batch(count=3, nmin=64, nmax=300))]
    for name, w, kw in cases:
        say(f"{name} {precision}: the same charges twice, other charges between")
        with ew.EwaldEngine(tol=1e-10, r_cut=w.get("r_cut"), precision=precision, **kw) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            q1 = W.scc_update(w["charges"], rng)
            q2 = W.scc_update(w["charges"], rng)
            s1, e1 = e.shift(q1)
            for _ in range(3):
                e.shift(q2)
            s1b, e1b = e.shift(q1)
            assert np.array_equal(s1, s1b), name
            assert e1 == e1b, name


def test_error_paths():
    w = W.water(n_side=4)
    say("error paths: no geometry, non-neutral, coincident, non-finite")
    with ew.EwaldEngine() as e:
        with pytest.raises(ew.EwaldError) as err:
            e.shift(w["charges"])
        assert err.value.code == ew.EW_ERR_NO_GEOMETRY
        e.set_geometry(w["coords"], w["lattice"])
        with pytest.raises(ew.EwaldError) as err:
            e.shift(w["charges"] + 0.01)
        assert err.value.code == ew.EW_ERR_NOT_NEUTRAL
        q = w["charges"].copy()
        q[3] = np.nan
        with pytest.raises(ew.EwaldError) as err:
            e.shift(q)
        assert err.value.code == ew.EW_ERR_NOT_FINITE
        # charges beyond the fixed-point spreading's range are refused, not wrapped
        q = w["charges"].copy()
        q[3], q[4] = 2e4, -2e4
        with pytest.raises(ew.EwaldError) as err:
            e.shift(q)
        assert err.value.code == ew.EW_ERR_ARGUMENT
        # arrays that do not match the geometry are refused before the C side reads them
        with pytest.raises(ValueError):
            e.shift(w["charges"][:-1])
        with pytest.raises(ValueError):
            e.gradients(w["charges"], grad=np.zeros((len(w["charges"]) - 1, 3)))
        with pytest.raises(ValueError):
            e.stress(w["charges"][:-1])
        # a Fortran-ordered gradient array is accumulated in the right layout
        gf = np.asfortranarray(np.arange(3.0 * len(w["charges"])).reshape(-1, 3))
        diff = e.gradients(w["charges"], grad=gf) - e.gradients(w["charges"])
        assert np.abs(diff - gf).max() < 1e-6
        bad = w["coords"].copy()
        bad[5] = bad[9]
        with pytest.raises(ew.EwaldError) as err:
            e.set_geometry(bad, w["lattice"])
        assert err.value.code == ew.EW_ERR_COINCIDENT


def test_one_shot_potential_matches_engine():
    g = load_golden("ref_triclinic.npz")
    say("one-shot ewald_potential_force")
    pot, _ = ew.ewald_potential_force(g["coords"], g["charges"], g["lattice"], tol=1e-12,
                                      alpha=float(g["alpha"]), order=12)
    assert rel_max(pot, g["pot"]) < 1e-10


@pytest.mark.parametrize("name,kw", [
    ("ref_triclinic.npz", dict(tol=1e-12, order=12)),
    ("ref_small_cell.npz", dict(tol=1e-12, order=12)),
    ("ref_nacl_c1.npz", dict(tol=1e-10, grid=24, order=10)),
])
def test_forces_match_reference_direct(name, kw):
    g = load_golden(name)
    if "alpha" in g:
        kw = dict(kw, alpha=float(g["alpha"]))
    say(f"forces: {name} {kw}")
    pot, force = ew.ewald_potential_force(g["coords"], g["charges"], g["lattice"], force=True, **kw)
    assert rel_max(pot, g["pot"]) < 1e-10
    assert rel_max(force, g["force"]) < 1e-9


def test_engine_gradients_and_c2_forces(oracle_mod):
    w = W.water(n_side=10)
    say("water 3000 atoms: ew_gradients (FP64, mixed) against the CPU SPME force")
    opts = dict(tol=1e-10, r_cut=w["r_cut"], grid=(40, 40, 40), order=6)
    with ew.EwaldEngine(**opts) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        prm = e.parameters()
        g64 = e.gradients(w["charges"], grad=np.ones((len(w["charges"]), 3)))
    _, f = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                 prm["order"], prm["alpha"], prm["r_cut"], want_force=True)
    assert rel_max(1.0 - g64, f) < 1e-10
    with ew.EwaldEngine(precision="mixed", **opts) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        g32 = e.gradients(w["charges"])
    assert rel_max(-g32, f) < 1e-5


def test_graph_replays_match_eager():
    import torch
    w = W.water(n_side=8)
    say("water 1536 atoms: captured-graph replays against eager launches")
    rng = np.random.default_rng(11)
    dev = torch.device("cuda:0")
    with ew.EwaldEngine(tol=1e-8) as g, ew.EwaldEngine(tol=1e-8, eager=True) as e:
        g.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        dq = torch.empty(len(w["charges"]), dtype=torch.float64, device=dev)
        out = torch.empty_like(dq)
        en = torch.empty(1, dtype=torch.float64, device=dev)
        q = w["charges"]
        for it in range(5):
            q = W.scc_update(q, rng)
            dq.copy_(torch.from_numpy(q))
            torch.cuda.synchronize()
            g.shift_device(dq.data_ptr(), out.data_ptr(), en.data_ptr())  # same buffers: replay
            torch.cuda.synchronize()
            s_eager, e_eager = e.shift(q)                                # eager work in between
            assert rel_max(out.cpu().numpy(), s_eager) < 1e-14
            assert float(en.item()) == pytest.approx(e_eager, rel=1e-14)


# ---- charged cells: the neutralising background of DFTB+'s stock Ewald ----
# (src/dftbp/dftb/coulomb.F90:1292-1296).  Pins: the simple-cubic Wigner constant
# xi = -2.837297479480620 (potential at a point charge from its images and the uniform
# background, in units of q/L; literature value, independent of the engine's formulas) and the
# independence of the potential from the splitting parameter alpha.

XI_SC = -2.837297479480620


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_charged_point_charge_simple_cubic(precision, tol):
    L, q = 20.0, 0.7
    say(f"charged cell: one charge in a cubic box, {precision}")
    with ew.EwaldEngine(tol=1e-12, charged=True, precision=precision) as e:
        e.set_geometry(np.array([[1.3, 2.1, 0.4]]), np.eye(3) * L, charges=np.array([q]))
        s, en = e.shift(np.array([q]))
    assert s[0] == pytest.approx(q * XI_SC / L, rel=tol)
    assert en == pytest.approx(0.5 * q * q * XI_SC / L, rel=tol)


def test_charged_cell_alpha_independent_and_rejected_by_default():
    w = W.batch(seed=W.SEED0 + 21, count=1, nmin=300, nmax=300)[0]
    rng = np.random.default_rng(4)
    q = w["charges"] + 0.05 + 0.01 * rng.random(len(w["charges"]))  # net charge ~ +16
    say("charged cell: alpha independence, 300 atoms, net charge %.2f" % q.sum())
    out = []
    for scale in (1.0, 1.4):
        with ew.EwaldEngine(tol=1e-12, charged=True) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=q)
            a = e.parameters()["alpha"]
        with ew.EwaldEngine(tol=1e-12, charged=True, alpha=a * scale) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=q)
            out.append(e.shift(q))
    assert rel_max(out[1][0], out[0][0]) < 1e-9
    assert out[1][1] == pytest.approx(out[0][1], rel=1e-9)
    with ew.EwaldEngine(tol=1e-12) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=q)
        with pytest.raises(ew.EwaldError) as exc:
            e.shift(q)
        assert exc.value.code == ew.EW_ERR_NOT_NEUTRAL


def test_charged_cell_matches_cpu_spme_plus_background(oracle_mod):
    w = W.water(n_side=6)
    q = w["charges"].copy()
    q[::7] += 0.3
    say("charged water box against CPU SPME + background, FP64")
    with ew.EwaldEngine(tol=1e-10, charged=True, r_cut=w["r_cut"], grid=24, order=6) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=q)
        s, en = e.shift(q)
        prm = e.parameters()
    pot = oracle_mod.spme_ewald(w["coords"], q, w["lattice"], prm["grid"], prm["order"],
                                prm["alpha"], prm["r_cut"])
    vol = abs(np.linalg.det(w["lattice"]))
    want = pot - math.pi / (vol * prm["alpha"] ** 2) * q.sum()
    assert rel_max(s, want) < 1e-11
    assert en == pytest.approx(0.5 * float(np.dot(q, want)), rel=1e-11)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_symmetric_pair_storage_matches_row_list(precision, tol):
    w = W.water()
    say(f"water: symmetric pair storage against the full row list, {precision}")
    out = {}
    for storage in (1, 0):
        with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"],
                            precision=precision, pair_storage=storage) as e:
            e.set_geometry(w["coords"], w["lattice"], species=w["species"],
                           hubbard_u=w["hubbard_u"], charges=w["charges"])
            prm = e.parameters()
            out[storage] = e.shift(w["charges"]) + (prm,)
    assert out[1][2]["pair_mode"] == 2 and out[0][2]["pair_mode"] == 0
    # each unordered pair once: half the entries of the row list
    assert 2 * out[1][2]["n_pairs"] == out[0][2]["n_pairs"]
    assert out[1][2]["device_bytes"] < out[0][2]["device_bytes"]
    assert rel_max(out[1][0], out[0][0]) < tol
    assert out[1][1] == pytest.approx(out[0][1], rel=tol)


# ---- stress (TCoulomb%addStress convention: sigma = W / V, W = -dE/d(strain)) ----
# Pins: central differences of the reference's own direct Ewald energy (the oracle's restatement,
# an independent method) under symmetric strains, and for a charged cell the closed form
# trace(sigma) = E / V of a single charge in a cubic cell (E = q^2 xi / 2L scales as 1/L).

def _strained(coords, lat, a, b, h):
    eps = np.zeros((3, 3))
    eps[a, b] += 0.5 * h
    eps[b, a] += 0.5 * h
    m = np.eye(3) + eps
    return coords @ m.T, lat @ m.T


def test_stress_matches_strain_derivative_of_direct_energy(oracle_mod):
    g = load_golden("ref_triclinic.npz")
    q = g["charges"]
    say("stress: 300 charges, sheared cell, against strain derivatives of the direct energy")
    with ew.EwaldEngine(tol=1e-12) as e:
        e.set_geometry(g["coords"], g["lattice"], charges=q)
        st = e.stress(q)
    vol = abs(np.linalg.det(g["lattice"]))
    h = 1e-5
    want = np.zeros((3, 3))
    for a in range(3):
        for b in range(a, 3):
            en = []
            for sgn in (1, -1):
                c, l = _strained(g["coords"], g["lattice"], a, b, sgn * h)
                pot, _, _ = oracle_mod.direct_ewald(c, q, l, tol=1e-13)
                en.append(0.5 * float(np.dot(q, pot)))
            want[a, b] = want[b, a] = -(en[0] - en[1]) / (2 * h) / vol
    assert np.allclose(st, st.T, rtol=0, atol=1e-14)
    assert np.abs(st - want).max() / np.abs(want).max() < 1e-6


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-4)])
def test_stress_is_strain_derivative_of_engine_energy_linked(precision, tol):
    w = W.water(n_side=6)
    say(f"stress: water 648 atoms (linked pairs), strain derivatives of the SPME energy, {precision}")
    opts = dict(tol=1e-12, alpha=0.3, r_cut=w["r_cut"], grid=(36, 36, 36), order=8,
                precision=precision)
    with ew.EwaldEngine(**opts) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        assert e.parameters()["pair_mode"] == 0
        st = e.stress(w["charges"])
    vol = abs(np.linalg.det(w["lattice"]))
    h = 1e-5 if precision == "fp64" else 1e-3
    want = np.zeros((3, 3))
    for a in range(3):
        for b in range(a, 3):
            en = []
            for sgn in (1, -1):
                c, l = _strained(w["coords"], w["lattice"], a, b, sgn * h)
                with ew.EwaldEngine(**opts) as e:
                    e.set_geometry(c, l, charges=w["charges"])
                    en.append(e.shift(w["charges"])[1])
            want[a, b] = want[b, a] = -(en[0] - en[1]) / (2 * h) / vol
    assert np.abs(st - want).max() / np.abs(want).max() < tol


def test_stress_charged_cell_closed_form():
    L, q = 20.0, 0.7
    say("stress: one charge in a cubic cell with the neutralising background")
    with ew.EwaldEngine(tol=1e-12, charged=True) as e:
        e.set_geometry(np.array([[1.3, 2.1, 0.4]]), np.eye(3) * L, charges=np.array([q]))
        st = e.stress(np.array([q]))
    energy = 0.5 * q * q * XI_SC / L
    assert np.trace(st) == pytest.approx(energy / L ** 3, rel=1e-9)
    assert np.allclose(np.diag(st), np.trace(st) / 3, rtol=1e-9, atol=0)
    assert np.abs(st - np.diag(np.diag(st))).max() < 1e-12 * abs(np.trace(st))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_plane_and_separate_fft_passes_agree(precision, tol):
    w = W.water()
    say(f"water: fused x-y plane FFT (fft_mode 2) against separate passes (1), {precision}")
    out = {}
    for mode in (2, 1, 0):
        with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"],
                            precision=precision, fft_mode=mode) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            out[mode] = e.shift(w["charges"]) + (e.parameters()["fft_planes"],)
    assert out[2][2] == 1 and out[1][2] == 0
    # the default on this 64^3 mesh: the register x-y plane kernels (a compiled plane side)
    assert out[0][2] == 1
    assert rel_max(out[0][0], out[1][0]) < tol
    assert rel_max(out[2][0], out[1][0]) < tol
    assert out[2][1] == pytest.approx(out[1][1], rel=tol)
    with pytest.raises(ew.EwaldError) as exc:
        ew.EwaldEngine(fft_mode=3)
    assert exc.value.code == ew.EW_ERR_ARGUMENT


# ---- ragged and degenerate inputs ----

def _clustered(n=1500, seed=21):
    """Atoms crowded into one corner slab of a large triclinic cell: most SPME tiles are empty,
    the occupied ones hold many atoms, and pair-list rows are long and uneven."""
    rng = np.random.default_rng(seed)
    lat = W.lattice_from_parameters(40.0 * W.AA, 36.0 * W.AA, 44.0 * W.AA, 80.0, 95.0, 70.0)
    frac = rng.uniform(0.0, 1.0, (n, 3)) * np.array([0.35, 0.3, 1.0])
    coords = frac @ lat
    # keep a minimum separation so no two charges coincide
    keep = [0]
    for i in range(1, n):
        d = coords[keep] - coords[i]
        f = d @ np.linalg.inv(lat)
        f -= np.rint(f)
        if np.min(np.linalg.norm(f @ lat, axis=1)) > 0.9 * W.AA:
            keep.append(i)
    coords = coords[keep]
    q = W.neutralise(rng.uniform(-1.0, 1.0, len(coords)))
    return {"coords": coords, "lattice": lat, "charges": q}


@pytest.mark.parametrize("storage", [0, 1])
def test_clustered_atoms_empty_tiles_match_cpu_spme(oracle_mod, storage):
    w = _clustered()
    say(f"{len(w['charges'])} atoms clustered in a corner slab (empty tiles), storage {storage}")
    s, en, prm = engine_shift(w, tol=1e-10, r_cut=10.0 * W.AA, grid=(48, 45, 50), order=6,
                              pair_storage=storage)
    assert prm["pair_mode"] == (2 if storage else 0)
    want = _spme_reference(oracle_mod, w, prm)
    assert rel_max(s, want) < 1e-11
    assert en == pytest.approx(0.5 * float(np.dot(w["charges"], want)), rel=1e-11)


def test_degenerate_charges_and_two_atoms(oracle_mod):
    say("all-zero charges, linearity, and a two-atom cell (image-summed pairs)")
    w = W.water(n_side=4)
    with ew.EwaldEngine(tol=1e-10) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        s, en = e.shift(np.zeros(len(w["charges"])))
        assert np.all(s == 0.0) and en == 0.0
        # linearity: shift(2 q) = 2 shift(q) up to round-off
        s1, e1 = e.shift(w["charges"])
        s2, e2 = e.shift(2.0 * w["charges"])
        assert rel_max(s2, 2.0 * s1) < 1e-13
        assert e2 == pytest.approx(4.0 * e1, rel=1e-13)
    lat = np.eye(3) * 12.0
    two = {"coords": np.array([[1.0, 2.0, 3.0], [5.0, 4.0, 9.0]]), "lattice": lat,
           "charges": np.array([0.7, -0.7])}
    s, en, prm = engine_shift(two, tol=1e-12, grid=24, order=8)
    want = _spme_reference(oracle_mod, two, prm)
    assert rel_max(s, want) < 1e-11


def test_crowded_tiles_use_the_per_atom_pair_build(oracle_mod):
    """More than 1024 atoms in a sort tile (a coarse 16^3 mesh over 5000 atoms): the pair list
    is built one wave per atom instead of through the tile-staged kernel; same rows."""
    rng = np.random.default_rng(31)
    lat = np.eye(3) * 45.0 * W.AA
    n = 5000
    w = {"coords": W.poisson_disc(n, lat, 1.0 * W.AA, rng), "lattice": lat,
         "charges": W.neutralise(rng.uniform(-1, 1, n))}
    say("5000 atoms, 16^3 mesh: > 1024 atoms per sort tile (per-atom pair build)")
    s, en, prm = engine_shift(w, tol=1e-8, r_cut=9.0 * W.AA, grid=16, order=6)
    assert prm["pair_mode"] == 0
    want = _spme_reference(oracle_mod, w, prm)
    assert rel_max(s, want) < 1e-11


def test_int32_partners_when_offsets_exceed_16_bits(oracle_mod):
    """A cell with four sort-tile layers along z holds ~35,000 atoms per layer, so partners in
    the next layer lie more than 32,767 rows away: the fill falls back to int32 partners.  The
    shift must still match the CPU SPME at the same parameters, and the mixed engine FP64."""
    rng = np.random.default_rng(77)
    side = 112.0 * W.AA
    lat = np.eye(3) * side
    n = 140_000
    w = {"coords": rng.random((n, 3)) @ lat, "lattice": lat,
         "charges": W.neutralise(rng.uniform(-1.0, 1.0, n))}
    say(f"{n} random atoms, grid (128, 128, 32): int32 partner fallback")
    opts = dict(tol=1e-8, r_cut=8.0 * W.AA, grid=(128, 128, 32), order=6)
    s_rows, _, prm = engine_shift(w, **opts)
    assert prm["pair_mode"] == 0 and prm["index_bytes"] == 4
    assert rel_max(s_rows, _spme_reference(oracle_mod, w, prm)) < 1e-11
    s_mixed, _, prm_m = engine_shift(w, precision="mixed", **opts)
    assert prm_m["index_bytes"] == 4
    assert rel_max(s_mixed, s_rows) < 1e-6
