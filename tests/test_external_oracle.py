"""Pins the CPU restatement of the external-charge potential (oracle/external.py) before it is
used as the checker of ew_external_potential.

  * sources on the atoms: the potential is the Ewald potential of the atoms themselves, checked
    against the reference's own direct-Ewald output (tests/golden/ref_triclinic.npz,
    ref_small_cell.npz);
  * alpha independence with non-neutral sources (fails without the background term);
  * closed forms of the two variants: a blurred source differs from a point source by
    -q erfc(d/w)/d summed over its images, a softened one by q erfc(alpha d)(1/sqrt(d^2+eps^2)-1/d);
  * translation of sources and targets together by a lattice vector changes nothing;
  * the gradients are central differences of the energy sum_t q_t potential_t.
"""
import numpy as np
import pytest
from conftest import load_golden, rel_max
from scipy.special import erfc

from oracle.external import external_gradients, external_potential


@pytest.mark.parametrize("name", ["ref_triclinic.npz", "ref_small_cell.npz"])
def test_sources_on_atoms_give_the_reference_ewald_potential(name):
    g = load_golden(name)
    x, q, lat = g["coords"], g["charges"], g["lattice"]
    n = min(len(q), 24)
    # the potential of all atoms at the first n of them
    alpha = 3.5 / abs(np.linalg.det(lat)) ** (1.0 / 3.0)   # any alpha: the sum is exact
    pot = external_potential(x[:n], x, q, lat, alpha=alpha)
    assert rel_max(pot, g["pot"][:n]) < 1e-11


def _cell(seed):
    rng = np.random.default_rng(seed)
    lat = np.array([[11.0, 0.0, 0.0], [2.5, 10.0, 0.0], [-1.0, 1.5, 12.0]])
    targets = rng.random((6, 3)) @ lat
    sources = rng.random((5, 3)) @ lat
    q = rng.normal(size=5)          # not neutral
    return lat, targets, sources, q


def test_alpha_independence_needs_the_background():
    lat, x, y, q = _cell(3)
    assert abs(q.sum()) > 0.1
    p1 = external_potential(x, y, q, lat, alpha=0.35)
    p2 = external_potential(x, y, q, lat, alpha=0.55)
    assert rel_max(p2, p1) < 1e-12
    # the same check fails once the background term is left out
    vol = abs(np.linalg.det(lat))
    bg = lambda a: -np.pi / (vol * a * a) * q.sum()
    assert rel_max(p2 - bg(0.55), p1 - bg(0.35)) > 1e-3


def test_blur_closed_form():
    lat, x, y, q = _cell(5)
    w = np.array([0.4, 0.7, 1.1, 0.3, 0.9])
    alpha = 0.45
    point = external_potential(x, y, q, lat, alpha=alpha)
    blurred = external_potential(x, y, q, lat, alpha=alpha, blur=w)
    # erf(d/w)/d = 1/d - erfc(d/w)/d: the difference is -q erfc(d/w)/d over images with d < 10 w
    want = np.zeros(len(x))
    n = np.arange(-3, 4)
    L = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3) @ lat
    for t in range(len(x)):
        for s in range(len(y)):
            d = np.linalg.norm(x[t] - y[s] + L, axis=1)
            d = d[d < 10 * w[s]]
            want[t] -= q[s] * np.sum(erfc(d / w[s]) / d)
    assert np.abs((blurred - point) - want).max() < 1e-12 * np.abs(point).max()
    # a vanishing width is a point charge
    tiny = external_potential(x, y, q, lat, alpha=alpha, blur=np.full(5, 1e-4))
    assert rel_max(tiny, point) < 1e-13


def test_softening_closed_form():
    lat = np.eye(3) * 30.0
    x = np.array([[1.0, 2.0, 3.0]])
    y = np.array([[3.0, 2.5, 4.0]])
    q = np.array([0.7])
    alpha, eps = 0.5, 0.8
    d = np.linalg.norm(x[0] - y[0])
    p0 = external_potential(x, y, q, lat, alpha=alpha)
    pe = external_potential(x, y, q, lat, alpha=alpha, eps_soften=eps)
    want = q[0] * erfc(alpha * d) * (1.0 / np.sqrt(d * d + eps * eps) - 1.0 / d)
    assert abs((pe - p0)[0] - want) < 1e-14


def test_lattice_translation_invariance():
    lat, x, y, q = _cell(9)
    p = external_potential(x, y, q, lat, alpha=0.4)
    shifted = external_potential(x + lat[0] - 2 * lat[2], y - lat[1], q, lat, alpha=0.4)
    assert rel_max(shifted, p) < 1e-12


@pytest.mark.parametrize("blurred", [False, True])
def test_gradients_are_energy_derivatives(blurred):
    lat, x, y, q = _cell(13)
    qt = np.random.default_rng(14).normal(size=len(x))
    w = np.array([0.5, 0.8, 0.3, 1.0, 0.6]) if blurred else None
    alpha = 0.4
    d0, d1 = external_gradients(x, y, q, lat, alpha, target_charges=qt, blur=w)
    h = 1e-5
    energy = lambda xx, yy: float(qt @ external_potential(xx, yy, q, lat, alpha=alpha, blur=w))
    for t, c in [(0, 0), (2, 1), (5, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[t, c] += h
        xm[t, c] -= h
        assert (energy(xp, y) - energy(xm, y)) / (2 * h) == pytest.approx(d0[t, c], abs=1e-8)
    for s, c in [(0, 2), (3, 0), (4, 1)]:
        yp, ym = y.copy(), y.copy()
        yp[s, c] += h
        ym[s, c] -= h
        assert (energy(x, yp) - energy(x, ym)) / (2 * h) == pytest.approx(d1[s, c], abs=1e-8)
    # the Hamiltonian derivative is the gradient of the potential itself
    g0, none = external_gradients(x, y, q, lat, alpha, blur=w)
    assert none is None
    assert np.allclose(g0 * qt[:, None], d0, rtol=1e-12, atol=1e-15)
