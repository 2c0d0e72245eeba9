"""Pins the CPU oracle (oracle/ewald_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed operand) fails it:
  * golden vectors written by the reference's own Fortran modules (tests/golden/),
  * closed forms (Madelung constants, the cubic B-spline, the Slater-density gamma integral),
  * invariances (alpha independence, translation, supercell replication, linearity, reciprocity),
  * brute force on tiny inputs (structure factors, DFT).
"""
import math

import numpy as np
import pytest
from conftest import load_golden, rel_max

from fftewald_amd import workloads as W

# ----------------------------------------------------------------------------- golden vectors


def test_direct_ewald_matches_reference_nacl(oracle_mod):
    g = load_golden("ref_nacl_c1.npz")
    pot, force, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"],
                                            tol=float(g["tol"]), alpha=float(g["alpha"]),
                                            want_force=True)
    assert rel_max(pot, g["pot"]) < 1e-13
    assert rel_max(force, g["force"]) < 1e-12


def test_direct_ewald_matches_reference_triclinic_own_alpha(oracle_mod):
    g = load_golden("ref_triclinic.npz")
    pot, force, alpha = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"],
                                                tol=float(g["tol"]), want_force=True)
    assert alpha == pytest.approx(float(g["alpha"]), rel=1e-15)
    assert rel_max(pot, g["pot"]) < 1e-13
    assert rel_max(force, g["force"]) < 1e-12


def test_direct_ewald_matches_reference_small_cell(oracle_mod):
    g = load_golden("ref_small_cell.npz")
    pot, force, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"],
                                            tol=float(g["tol"]), want_force=True)
    assert rel_max(pot, g["pot"]) < 1e-13
    assert rel_max(force, g["force"]) < 1e-12


def test_balanced_alpha_matches_reference_bitwise(oracle_mod):
    t = load_golden("ref_alpha.npz")["table"]
    for minG, minR, V, tol, a_ref, st_ref in t:
        a, st = oracle_mod.balanced_alpha(minG, minR, V, tol)
        assert st == int(st_ref)
        assert a == a_ref


def test_lambert_w0_matches_reference_bitwise(oracle_mod):
    g = load_golden("ref_truncation.npz")
    for x, w in zip(g["lw_x"], g["lw_w"]):
        assert oracle_mod.lambert_w0(float(x)) == w, x
    # and the defining identity W e^W = x on the same table (x > 0, no overflow; exp
    # amplifies the last-bit error of W by W)
    for x, w in zip(g["lw_x"], g["lw_w"]):
        if 1e-300 < x < 1e150:
            assert w * math.exp(w) == pytest.approx(x, rel=2e-15 * (2.0 + w))


def _chain_rows():
    g = load_golden("ref_truncation.npz")
    keys = ("n", "q2", "tol", "V", "r_cut", "alpha", "k_cut", "minG", "minR", "a_bal", "st")
    return g, [dict(zip(keys, r), lattice=lat) for r, lat in zip(g["chain"], g["lattice"])]


def test_budget_inversions_match_reference_bitwise(oracle_mod):
    """splitting_from_real_budget / cutoff_from_fourier_budget (fe/common/ewald_truncation.f90:
    53-104) against the compiled reference on the workload cells, tol 1e-6 ... 1e-12."""
    g, rows = _chain_rows()
    frac = float(g["budget_fraction"])
    for r in rows:
        n = int(r["n"])
        a = oracle_mod.splitting_from_real_budget(r["r_cut"], r["q2"], n, r["V"], frac * r["tol"])
        assert a == r["alpha"]
        k = oracle_mod.cutoff_from_fourier_budget(r["q2"], n, r["V"] ** (1.0 / 3.0), r["alpha"],
                                                 r["tol"])
        assert k == r["k_cut"]


def test_parameter_chain_matches_reference(oracle_mod):
    """The oracle's whole chain (r_cut rule and linked-cell cap, alpha from 0.2 tol, k_cut;
    fe/fft/ewald_fft_3d_parameters.f90:96-131) against the chain composed from the compiled
    reference functions with the rule's constants read from the reference source."""
    _, rows = _chain_rows()
    for r in rows:
        m = oracle_mod.cell_metrics(r["lattice"])
        assert m["volume"] == pytest.approx(r["V"], rel=1e-15)
        q = np.full(int(r["n"]), math.sqrt(r["q2"] / r["n"]))
        p = oracle_mod.choose_parameters(r["lattice"], q, r["tol"])
        assert p["r_cut"] == pytest.approx(r["r_cut"], rel=1e-15)
        assert p["alpha"] == pytest.approx(r["alpha"], rel=1e-14)
        assert p["k_cut"] == pytest.approx(r["k_cut"], rel=1e-14)


def test_dftb_optimal_alpha_matches_reference_balance(oracle_mod):
    """getOptimalAlphaEwald (src/dftbp/dftb/coulomb.F90:2097-2179, 2686-2748) bisects the same
    balance as the compiled balanced_alpha (fe/common/ewald_alpha_search.f90:83-181) with
    minG = |shortest reciprocal vector incl. 2 pi|, minR = |shortest lattice vector|; only the
    rounding order of the Fourier term and the cap (30 against 100 steps) differ.  Agreement to
    the last bisection bracket (relative 2^-30 of the starting bracket, here < 1e-8) whenever the
    reference converged; exact agreement on most rows."""
    _, rows = _chain_rows()
    exact = 0
    for r in rows:
        assert int(r["st"]) == 0
        a = oracle_mod.optimal_alpha_dftb(r["lattice"], r["tol"])
        assert a == pytest.approx(r["a_bal"], rel=1e-8)
        exact += a == r["a_bal"]
    assert exact >= len(rows) // 2
    # the 512-atom NaCl workload's alpha (written into ref_nacl_c1.npz) is the reference's
    g = load_golden("ref_nacl_c1.npz")
    nacl = [r for r in rows if int(r["n"]) == 512 and r["tol"] == 1e-10][0]
    assert float(g["alpha"]) == pytest.approx(nacl["a_bal"], rel=1e-8)


def test_planted_parameter_mistakes_miss_the_reference(oracle_mod):
    """Plausible slips in the chain fail the fixtures above: minG without the 2 pi, the whole
    tolerance as the real-space budget instead of its fraction, the cube edge replaced by the
    first lattice length."""
    g, rows = _chain_rows()
    frac = float(g["budget_fraction"])
    assert frac == 0.2 and float(g["prefactor"]) == 3.0
    for r in rows:
        n = int(r["n"])
        a, st = oracle_mod.balanced_alpha(r["minG"] / (2.0 * math.pi), r["minR"], r["V"],
                                          r["tol"])
        assert st != 0 or abs(a / r["a_bal"] - 1.0) > 1e-3
        a = oracle_mod.splitting_from_real_budget(r["r_cut"], r["q2"], n, r["V"], r["tol"])
        assert abs(a / r["alpha"] - 1.0) > 1e-3
        edge = float(np.linalg.norm(r["lattice"][0]))
        k = oracle_mod.cutoff_from_fourier_budget(r["q2"], n, edge, r["alpha"], r["tol"])
        if abs(edge / r["V"] ** (1.0 / 3.0) - 1.0) > 1e-3:
            assert abs(k / r["k_cut"] - 1.0) > 1e-6


def test_spme_reaches_reference_direct_ewald_on_water(oracle_mod):
    """The CPU SPME against the reference's direct Ewald on the 12,288-atom water box
    (tests/golden/ref_water_c2.npz): 1e-10 at 128^3 / order 12 / tol 1e-12; the production
    64^3 / order 6 setting's discretisation error (8.6e-6) is what the GPU test states."""
    import fftewald_amd as ew
    g = load_golden("ref_water_c2.npz")
    w = W.water()
    assert np.array_equal(w["coords"], g["coords"]) and np.array_equal(w["charges"], g["charges"])
    prm = ew.ewald_parameters(g["charges"], g["lattice"], tol=1e-12, grid=128, order=12,
                              r_cut=10.0 * W.AA)
    s = oracle_mod.spme_ewald(g["coords"], g["charges"], g["lattice"], prm["grid"], prm["order"],
                              prm["alpha"], prm["r_cut"])
    assert rel_max(s, g["pot"]) < 1e-10
    prm = ew.ewald_parameters(g["charges"], g["lattice"], tol=1e-10, grid=64, order=6,
                              r_cut=10.0 * W.AA)
    s = oracle_mod.spme_ewald(g["coords"], g["charges"], g["lattice"], prm["grid"], prm["order"],
                              prm["alpha"], prm["r_cut"])
    assert 1e-6 < rel_max(s, g["pot"]) < 2e-5


def test_linked_cell_real_space_matches_reference(oracle_mod):
    g = load_golden("ref_real_space.npz")
    pot, force = oracle_mod.real_space(g["coords"], g["charges"], g["lattice"], float(g["alpha"]),
                                       float(g["r_cut"]), method="cell_list", want_force=True)
    assert rel_max(pot, g["pot"]) < 1e-13
    assert rel_max(force, g["force"]) < 1e-12
    # and the image sum gives the same terms
    pot2 = oracle_mod.real_space(g["coords"], g["charges"], g["lattice"], float(g["alpha"]),
                                 float(g["r_cut"]), method="images")
    assert rel_max(pot2, g["pot"]) < 1e-13


# ----------------------------------------------------------------------------- closed forms

MADELUNG_NACL = 1.747564594633182
MADELUNG_CSCL = 1.76267477307099
MADELUNG_ZNS = 1.638055053388790


def _madelung(oracle_mod, coords, q, lat, r0):
    pot, _, _ = oracle_mod.direct_ewald(coords, q, lat, tol=1e-14, phase_tables=True)
    return -0.5 * float(np.dot(q, pot)) / (len(q) / 2) * r0


def test_madelung_reference_fixture():
    g = load_golden("ref_madelung.npz")
    r0 = 2.82 * W.AA
    m = -0.5 * float(np.dot(g["charges"], g["pot"])) / (len(g["charges"]) / 2) * r0
    assert m == pytest.approx(MADELUNG_NACL, abs=5e-12)


def test_madelung_nacl_cscl_zns(oracle_mod):
    w = W.nacl_ideal(reps=2)
    assert _madelung(oracle_mod, w["coords"], w["charges"], w["lattice"],
                     2.82 * W.AA) == pytest.approx(MADELUNG_NACL, abs=1e-12)
    a = 4.11
    cscl = np.array([[0, 0, 0], [0.5 * a, 0.5 * a, 0.5 * a]])
    m = _madelung(oracle_mod, cscl, np.array([1.0, -1.0]), np.eye(3) * a, math.sqrt(3) * a / 2)
    assert m == pytest.approx(MADELUNG_CSCL, abs=1e-12)
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]) * a
    zns = np.concatenate([fcc, fcc + 0.25 * a])
    q = np.array([1.0] * 4 + [-1.0] * 4)
    m = _madelung(oracle_mod, zns, q, np.eye(3) * a, math.sqrt(3) * a / 4)
    assert m == pytest.approx(MADELUNG_ZNS, abs=1e-12)


def test_cubic_bspline_closed_form(oracle_mod):
    # M_4 on [0,4]: M(x) = x^3/6 on [0,1], (-3x^3 + 12x^2 - 12x + 4)/6 on [1,2], symmetric
    def m4(x):
        x = min(x, 4 - x)
        if x <= 0:
            return 0.0
        if x <= 1:
            return x ** 3 / 6
        return (-3 * x ** 3 + 12 * x ** 2 - 12 * x + 4) / 6

    for u in (0.0, 0.25, 3.7, 10.999):
        th, dth = oracle_mod.bspline_weights(u, 4)
        w = u - math.floor(u)
        for o in range(4):
            assert th[o] == pytest.approx(m4(w + 3 - o), abs=1e-15)
        assert sum(th) == pytest.approx(1.0, abs=1e-15)
        assert sum(dth) == pytest.approx(0.0, abs=1e-15)
    for p in (5, 6, 10, 16):
        th, dth = oracle_mod.bspline_weights(0.37, p)
        assert sum(th) == pytest.approx(1.0, abs=1e-14)
        eps = 1e-6
        tp, _ = oracle_mod.bspline_weights(0.37 + eps, p)
        tm, _ = oracle_mod.bspline_weights(0.37 - eps, p)
        assert np.allclose((tp - tm) / (2 * eps), dth, atol=1e-8)


def test_bspline_moduli(oracle_mod):
    for K, p in ((16, 4), (27, 6), (64, 8)):
        b = oracle_mod.bspline_moduli(K, p)
        assert b[0] == pytest.approx(1.0, abs=1e-14)
        th, _ = oracle_mod.bspline_weights(0.0, p)
        mk = th[::-1][1:]  # M_p(k+1), k = 0..p-2
        for m in (1, K // 3, K - 1):
            s = sum(mk[k] * np.exp(2j * np.pi * m * k / K) for k in range(p - 1))
            assert b[m] == pytest.approx(1.0 / abs(s) ** 2, rel=1e-12)
        assert np.allclose(b[1:], b[1:][::-1], rtol=1e-12)


def _slater_gamma(R, ta, tb):
    """1/R - gamma_short by quadrature: Coulomb energy of two unit exponential densities
    rho(r) = t^3/(8 pi) exp(-t r) at distance R (independent of shortgammafuncs.F90)."""
    from scipy import integrate

    def va(s):  # potential of density a at distance s
        return 1.0 / s - math.exp(-ta * s) * (1.0 / s + ta / 2.0)

    def shell(r):  # angular average of va(|R + r|) over directions of r
        lo, hi = abs(R - r), R + r
        f = integrate.quad(lambda s: va(s) * s, lo, hi, epsabs=1e-14, epsrel=1e-13, limit=200)[0]
        return f / (2.0 * R * r)

    rho = lambda r: tb ** 3 / (8 * math.pi) * math.exp(-tb * r)  # noqa: E731
    g = integrate.quad(lambda r: 4 * math.pi * r * r * rho(r) * shell(r), 0, 60.0 / tb,
                       epsabs=1e-13, epsrel=1e-12, limit=400, points=[R])[0]
    return 1.0 / R - g


@pytest.mark.parametrize("ua,ub", [(0.3, 0.3), (0.2, 0.3), (0.4, 0.25)])
def test_short_gamma_is_the_slater_integral(oracle_mod, ua, ub):
    for R in (1.5, 3.0, 6.0):
        s = oracle_mod.exp_gamma(R, ua, ub)
        assert s == pytest.approx(_slater_gamma(R, 3.2 * ua, 3.2 * ub), rel=1e-8, abs=1e-12)


def test_short_gamma_limits(oracle_mod):
    assert oracle_mod.exp_gamma(0.0, 0.3, 0.3) == pytest.approx(-0.3)
    # symmetric in the two U, continuous as Ub -> Ua
    assert oracle_mod.exp_gamma(2.0, 0.2, 0.4) == pytest.approx(oracle_mod.exp_gamma(2.0, 0.4, 0.2),
                                                                rel=1e-14)
    assert oracle_mod.exp_gamma(2.0, 0.3, 0.3 + 1e-4) == pytest.approx(
        oracle_mod.exp_gamma(2.0, 0.3, 0.3), rel=1e-3)
    # short range: decays faster than 1/r; the cutoff sits where it falls to ~1e-10
    for u2, u1 in ((0.2, 0.2), (0.3, 0.2), (0.4, 0.4)):
        c = oracle_mod.exp_gamma_cutoff(u2, u1)
        assert oracle_mod.exp_gamma(c, u2, u1) <= 1e-10
        assert oracle_mod.exp_gamma(0.9 * c, u2, u1) > 1e-10
    assert oracle_mod.exp_gamma_cutoff(0.2, 0.2) == pytest.approx(39.9, abs=0.2)
    assert oracle_mod.exp_gamma_cutoff(0.4, 0.4) == pytest.approx(20.5, abs=0.2)


def test_short_gamma_shift_brute_force(oracle_mod):
    rng = np.random.default_rng(3)
    lat = W.lattice_from_parameters(9.0, 10.0, 11.0, 80, 95, 105)  # Bohr: small, many images
    pos = W.poisson_disc(6, lat, 1.5, rng)
    q = W.neutralise(rng.uniform(-1, 1, 6))
    sp = np.array([0, 1, 0, 1, 1, 0], dtype=np.int32)
    U = np.array([0.35, 0.5])
    got = oracle_mod.short_gamma_shift(pos, q, lat, sp, U)
    cut = {(a, b): oracle_mod.exp_gamma_cutoff(U[max(a, b)], U[min(a, b)]) for a in (0, 1)
           for b in (0, 1)}
    want = np.zeros(6)
    R = 8
    for i in range(6):
        want[i] = U[sp[i]] * q[i]
        for j in range(6):
            for t in np.ndindex(2 * R + 1, 2 * R + 1, 2 * R + 1):
                T = np.array(t) - R
                if i == j and not T.any():
                    continue
                r = np.linalg.norm(pos[j] - pos[i] + T @ lat)
                if r <= cut[(sp[i], sp[j])]:
                    want[i] -= oracle_mod.exp_gamma(r, U[sp[j]], U[sp[i]]) * q[j]
    assert rel_max(got, want) < 1e-13


# ----------------------------------------------------------------------------- invariances


def test_direct_ewald_alpha_independent(oracle_mod):
    g = load_golden("ref_triclinic.npz")
    a0 = float(g["alpha"])
    base, _, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"], tol=1e-15,
                                         alpha=a0, phase_tables=True)
    for f in (0.7, 1.5):
        pot, _, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"], tol=1e-15,
                                            alpha=f * a0, phase_tables=True)
        assert rel_max(pot, base) < 1e-12


def test_translation_permutation_supercell_linearity(oracle_mod):
    rng = np.random.default_rng(11)
    lat = W.lattice_from_parameters(12.0, 13.0, 14.0, 85, 95, 80)
    n = 24
    pos = W.poisson_disc(n, lat, 1.5, rng)
    q = W.neutralise(rng.uniform(-1, 1, n))
    q2 = W.neutralise(rng.uniform(-1, 1, n))
    f = lambda p, c, L=lat: oracle_mod.direct_ewald(p, c, L, tol=1e-14, phase_tables=True)[0]  # noqa
    base = f(pos, q)
    assert rel_max(f(pos + np.array([3.1, -2.0, 0.7]), q), base) < 1e-12
    perm = rng.permutation(n)
    assert rel_max(f(pos[perm], q[perm]), base[perm]) < 1e-12
    assert rel_max(f(pos, 2.0 * q - 0.5 * q2), 2.0 * base - 0.5 * f(pos, q2)) < 1e-12
    # reciprocity: sum a phi[b] = sum b phi[a]
    assert float(np.dot(q, f(pos, q2))) == pytest.approx(float(np.dot(q2, base)), rel=1e-11)
    # 2x1x1 supercell: per-atom potentials are unchanged
    lat2 = lat.copy()
    lat2[0] *= 2
    pos2 = np.concatenate([pos, pos + lat[0]])
    sup = f(pos2, np.concatenate([q, q]), lat2)
    assert rel_max(sup[:n], base) < 1e-11
    # a change of lattice basis (a3 -> a3 + a1) describes the same crystal
    lat3 = lat.copy()
    lat3[2] = lat[2] + lat[0]
    assert rel_max(f(pos, q, lat3), base) < 1e-11


def test_phase_tables_equal_cos_sin(oracle_mod):
    g = load_golden("ref_nacl_c1.npz")
    a, _, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"], tol=1e-10,
                                      alpha=float(g["alpha"]))
    b, _, _ = oracle_mod.direct_ewald(g["coords"], g["charges"], g["lattice"], tol=1e-10,
                                      alpha=float(g["alpha"]), phase_tables=True)
    assert rel_max(b, a) < 1e-13


# ----------------------------------------------------------------------------- FFT / NFFT / SPME


@pytest.mark.parametrize("shape", [(8, 6, 10), (7, 9, 5), (14, 1, 49), (60, 4, 3)])
def test_fft3d_matches_numpy(oracle_mod, shape):
    rng = np.random.default_rng(1)
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    fwd = oracle_mod.fft3d(g, -1)
    assert np.allclose(fwd, np.fft.fftn(g), atol=1e-11)
    bwd = oracle_mod.fft3d(g, +1)
    assert np.allclose(bwd, np.fft.ifftn(g) * g.size, atol=1e-11)


def test_next_fast_length_and_window(oracle_mod):
    assert [oracle_mod.next_fast_length(n) for n in (97, 99, 121, 1, 11)] == [98, 100, 126, 2, 12]
    assert oracle_mod.window_cutoff_from_budget(1e-9) == 11
    assert oracle_mod.window_cutoff_from_budget(1e-10) == 12
    assert oracle_mod.window_cutoff_from_budget(1e-4) == 6
    assert oracle_mod.window_cutoff_from_budget(1e-12) == 14
    w = oracle_mod.lambert_w0(3.7)
    assert w * math.exp(w) == pytest.approx(3.7, rel=1e-15)


def test_nfft_adjoint_matches_direct_structure_factors(oracle_mod):
    rng = np.random.default_rng(2)
    t = rng.uniform(-0.3, 1.4, (11, 3))
    q = rng.normal(size=11)
    M = (6, 8, 4)
    s = oracle_mod.nfft_adjoint(t, q, M, 12)
    for a3 in range(M[2]):
        for a2 in range(M[1]):
            for a1 in range(M[0]):
                k = [a if a < m // 2 else a - m for a, m in zip((a1, a2, a3), M)]
                want = np.sum(q * np.exp(2j * np.pi * (t @ np.array(k))))
                assert abs(s[a3, a2, a1] - want) < 1e-9 * np.abs(q).sum()
    # the forward transform is its transpose
    fh = rng.normal(size=(M[2], M[1], M[0])) + 1j * rng.normal(size=(M[2], M[1], M[0]))
    f = oracle_mod.nfft_forward(t, fh, 12)
    lhs = np.vdot(q, f)  # <q, A f>
    rhs = np.sum(np.conj(oracle_mod.nfft_adjoint(t, q, M, 12)) * fh)  # <A^H q, f>
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_reference_fast_path_matches_direct(oracle_mod):
    g = load_golden("ref_triclinic.npz")
    pot, force, prm = oracle_mod.fft_ewald(g["coords"], g["charges"], g["lattice"], tol=1e-9,
                                           want_force=True)
    assert prm["window_cutoff"] == 11
    assert rel_max(pot, g["pot"]) < 1e-7
    assert rel_max(force, g["force"]) < 1e-6


def test_spme_converges_to_direct(oracle_mod):
    g = load_golden("ref_nacl_c1.npz")
    a = float(g["alpha"])
    rc = math.sqrt(-math.log(1e-10)) / a
    pot = oracle_mod.spme_ewald(g["coords"], g["charges"], g["lattice"], (16, 16, 16), 8, a, rc)
    assert rel_max(pot, g["pot"]) < 1e-10
    g = load_golden("ref_triclinic.npz")
    a = float(g["alpha"])
    rc = math.sqrt(-math.log(1e-12)) / a
    errs = []
    for K, p in ((24, 6), (40, 8), (64, 12)):
        pot = oracle_mod.spme_ewald(g["coords"], g["charges"], g["lattice"], (K, K, K), p, a, rc)
        errs.append(rel_max(pot, g["pot"]))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-10


def test_spme_forces_match_finite_differences(oracle_mod):
    g = load_golden("ref_small_cell.npz")
    pos, q, lat = g["coords"], g["charges"], g["lattice"]
    a, K, p = 0.6, (20, 20, 20), 8
    _, f = oracle_mod.spme_reciprocal(pos, q, lat, K, p, a, want_force=True)
    eps = 1e-5
    for i in (0, 7):
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += eps
            pm[i, d] -= eps
            ep = 0.5 * np.dot(q, oracle_mod.spme_reciprocal(pp, q, lat, K, p, a))
            em = 0.5 * np.dot(q, oracle_mod.spme_reciprocal(pm, q, lat, K, p, a))
            assert f[i, d] == pytest.approx(-(ep - em) / (2 * eps), rel=1e-6, abs=1e-9)


def test_check_configuration(oracle_mod):
    assert oracle_mod.check_configuration([1.0, -1.0]) == 0
    assert oracle_mod.check_configuration([1.0, -0.9]) == 2
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.direct_ewald(np.zeros((2, 3)), [1.0, -1.0], np.eye(3) * 5.0)


def test_workload_generators():
    w = W.nacl()
    assert w["coords"].shape == (512, 3)
    assert abs(w["charges"].sum()) < 1e-12
    w = W.water(n_side=4)
    assert w["coords"].shape == (192, 3)
    d = np.linalg.norm(w["coords"][1::3] - w["coords"][0::3], axis=1)
    assert np.allclose(d, 0.96 * W.AA)
    lat = W.lattice_from_parameters(10, 10, 10, 90, 90, 75)
    assert np.degrees(np.arccos(lat[0] @ lat[1] / 100)) == pytest.approx(75)
    assert W.smooth235(97) == 100 and W.smooth235(121) == 125
