"""CPU restatement of DFTB+'s periodic potential of external point charges (TCoulomb%getPotential).

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of ew_external_potential; the product
library never imports or calls it.

Follows src/dftbp/dftb/coulomb.F90: getPotential (3121-3160, periodic branch) ->
sumInvRPeriodicAsymm (1340-1418): potential_t = sum_s q_s ewald(x_t - y_s) with
  ewald           (2311-2345): reciprocal + real - pi/(V alpha^2), and -2 alpha/sqrt(pi) when
                               |x_t - y_s|^2 < tolSameDist^2 (the raw difference, as written);
  ewaldReciprocal (2348-2386): 2 * sum over half of the G != 0 lattice of
                               exp(-G^2/(4 alpha^2))/G^2 cos(G.r), times 4 pi / V;
  ewaldReal       (2523-2575): sum over lattice vectors L of erfc(alpha d)/sqrt(d^2 + eps^2),
                               d = |r + L|, skipping d < tolSameDist; with a blur width w:
                               [erf(d/w) (1 for d >= 10 w) - erf(alpha d)]/sqrt(d^2 + eps^2).
tolSameDist = 1e-5 (src/dftbp/common/accuracy.F90:69), erfArgLimit_ = 10 (coulomb.F90:214).
The reference takes its G and L lists from TCoulomb's setup; here they are every lattice point
within cutoffs chosen so that the dropped terms are below `tol` (the sums are then exact to
rounding, and independent of alpha -- except with eps_soften > 0, which softens only the
real-space denominator, so that the result depends on alpha as the reference's does).

Gradients follow addInvRPrimePeriodicAsymm (1934-2092) with derivEwaldReciprocal (2388-2425)
and derivEwaldReal (2577-2620): deriv0(:, t) += q_t q_s grad ewald(x_t - y_s), deriv1(:, s) -=
the same (q_t = 1 and no deriv1 for the Hamiltonian derivative); derivEwaldReal skips images
closer than tolSameDist^2 in distance, as written.

Pinned in tests/test_external_oracle.py: with the sources on the atoms the sum is the Ewald
potential of the direct-Ewald restatement (itself pinned by Madelung constants); the total is
alpha-independent for a non-neutral set of sources only with the background term; a blurred
source differs from a point source by the closed form -q erfc(d/w)/d over its images; the
softened form differs by q erfc(alpha d) (1/sqrt(d^2+eps^2) - 1/d); the gradients are the
central differences of sum_t q_t potential_t with respect to the atoms and the sources.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfc

TOL_SAME_DIST = 1.0e-5
ERF_ARG_LIMIT = 10.0


def _lattice_points(basis: np.ndarray, rmax: float) -> np.ndarray:
    """All n1 b1 + n2 b2 + n3 b3 (rows of basis) with |.| <= rmax."""
    inv = np.linalg.inv(basis)
    # |n_a| <= rmax * |row a of inv^T| bounds every point of the sphere
    nmax = np.ceil(rmax * np.linalg.norm(inv, axis=0)).astype(int)
    r = [np.arange(-k, k + 1) for k in nmax]
    n = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
    pts = n @ basis
    return pts[np.einsum("ij,ij->i", pts, pts) <= rmax * rmax]


def external_potential(targets, sources, charges, lat, alpha, blur=None, eps_soften=0.0,
                       tol=1.0e-17):
    """Potential (n_targets,) of the point charges at `sources` (m, 3) on `targets` (n, 3);
    lattice rows a_i (3, 3), atomic units."""
    x = np.asarray(targets, dtype=np.float64)
    y = np.asarray(sources, dtype=np.float64)
    q = np.asarray(charges, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    w = None if blur is None else np.asarray(blur, dtype=np.float64)
    vol = abs(np.linalg.det(lat))
    rec = 2.0 * np.pi * np.linalg.inv(lat).T          # rows G_i, G_i . a_j = 2 pi delta_ij
    lnt = np.sqrt(-np.log(tol))
    gmax = 2.0 * alpha * lnt
    rmax = lnt / alpha
    if w is not None:
        rmax = max(rmax, w.max() * lnt)
    g = _lattice_points(rec, gmax)
    g = g[np.einsum("ij,ij->i", g, g) > 0.0]
    # half space: the first non-zero component positive (no origin, no inversion partners)
    key = np.where(g[:, 0] != 0, g[:, 0], np.where(g[:, 1] != 0, g[:, 1], g[:, 2]))
    g = g[key > 0]
    g2 = np.einsum("ij,ij->i", g, g)
    gw = np.exp(-g2 / (4.0 * alpha * alpha)) / g2
    diff = x[:, None, :] - y[None, :, :]               # rr = coord0 - coord1, (n, m, 3)
    span = np.abs(diff).max() * np.sqrt(3.0)
    lvec = _lattice_points(lat, rmax + span)
    pot = np.zeros(len(x))
    eps2 = eps_soften * eps_soften
    for t in range(len(x)):
        for s in range(len(y)):
            rr = diff[t, s]
            rec_sum = 2.0 * np.sum(gw * np.cos(g @ rr)) * 4.0 * np.pi / vol   # 2348-2386
            d2 = np.einsum("ij,ij->i", rr + lvec, rr + lvec)                  # 2523-2575
            keep = d2 >= TOL_SAME_DIST * TOL_SAME_DIST
            d = np.sqrt(d2[keep])
            den = np.sqrt(d2[keep] + eps2)
            if w is None:
                real = np.sum(erfc(alpha * d) / den)
            else:
                near = d < ERF_ARG_LIMIT * w[s]
                # erf(d/w) - erf(alpha d) = erfc(alpha d) - erfc(d/w)
                real = np.sum((erfc(alpha * d) - np.where(near, erfc(d / w[s]), 0.0)) / den)
            val = rec_sum + real - np.pi / (vol * alpha * alpha)              # 2311-2345
            if rr @ rr < TOL_SAME_DIST * TOL_SAME_DIST:
                val -= 2.0 * alpha / np.sqrt(np.pi)
            pot[t] += val * q[s]
    return pot


def _lists(lat, alpha, tol, wmax, span):
    rec = 2.0 * np.pi * np.linalg.inv(lat).T
    lnt = np.sqrt(-np.log(tol))
    g = _lattice_points(rec, 2.0 * alpha * lnt)
    g = g[np.einsum("ij,ij->i", g, g) > 0.0]
    key = np.where(g[:, 0] != 0, g[:, 0], np.where(g[:, 1] != 0, g[:, 1], g[:, 2]))
    g = g[key > 0]
    rmax = max(lnt / alpha, wmax * lnt)
    return g, _lattice_points(lat, rmax + span)


def external_gradients(targets, sources, charges, lat, alpha, target_charges=None, blur=None,
                       tol=1.0e-17):
    """(deriv0 (n, 3), deriv1 (m, 3)) of addInvRPrimePeriodicAsymm; target_charges None is the
    Hamiltonian derivative (deriv0 = sum_s q_s grad ewald(x_t - y_s), deriv1 None)."""
    x = np.asarray(targets, dtype=np.float64)
    y = np.asarray(sources, dtype=np.float64)
    q = np.asarray(charges, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    w = None if blur is None else np.asarray(blur, dtype=np.float64)
    qt = np.ones(len(x)) if target_charges is None else np.asarray(target_charges, np.float64)
    vol = abs(np.linalg.det(lat))
    diff = x[:, None, :] - y[None, :, :]
    g, lvec = _lists(lat, alpha, tol, 0.0 if w is None else w.max(),
                     np.abs(diff).max() * np.sqrt(3.0))
    g2 = np.einsum("ij,ij->i", g, g)
    gw = np.exp(-g2 / (4.0 * alpha * alpha)) / g2
    d0 = np.zeros((len(x), 3))
    d1 = np.zeros((len(y), 3))
    for t in range(len(x)):
        for s in range(len(y)):
            rr = diff[t, s]
            # derivEwaldReciprocal (2388-2425)
            rec = 2.0 * ((-np.sin(g @ rr) * gw) @ g) * 4.0 * np.pi / vol
            # derivEwaldReal (2577-2620)
            v = rr + lvec
            d = np.sqrt(np.einsum("ij,ij->i", v, v))
            keep = d >= TOL_SAME_DIST * TOL_SAME_DIST
            v, d = v[keep], d[keep]
            f = alpha * d
            c = (-2.0 / np.sqrt(np.pi) * np.exp(-f * f) * f - erfc(f)) / d ** 3
            if w is not None:
                near = d < ERF_ARG_LIMIT * w[s]
                fb = d / w[s]
                c = c + np.where(near, (2.0 / np.sqrt(np.pi) * np.exp(-fb * fb) * fb + erfc(fb))
                                 / d ** 3, 0.0)
            real = c @ v
            ft = (rec + real) * q[s] * qt[t]
            d0[t] += ft
            d1[s] -= ft
    return d0, (None if target_charges is None else d1)
