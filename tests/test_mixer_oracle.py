"""The CPU restatement of DFTB+'s charge mixers (oracle/broyden.py), pinned before it is used as
the checker of the device mixer:
  * the reference's own regression vectors (test/src/dftbp/unit/mixer/broyden.F90:25-64, a water
    molecule over four iterations) within the reference test's tolerance, atol = 100 eps;
  * the first Broyden step is simple mixing, and a mixer that drops the stored history, or the
    omega0 regularisation, misses the reference vectors (the vectors pin those terms);
  * on a linear fixed-point problem Broyden converges where simple mixing at the same
    mixing parameter has not.
"""
import json
import os
import warnings

import numpy as np
import pytest
from conftest import GOLDEN

from oracle.broyden import BroydenMixer, SimpleMixer

EPS = np.finfo(np.float64).eps


def golden():
    with open(os.path.join(GOLDEN, "ref_broyden_h2o.json")) as f:
        return json.load(f)


def make(p, cls=BroydenMixer):
    return cls(p["mixParam"], p["maxSccIter"], p["omega0"], p["minWeight"], p["maxWeight"],
               p["weightFac"])


def run_reference_sequence(m, g):
    m.reset(3)
    q = np.array(g["expected_charges"][0])
    errs = []
    for i, d in enumerate(g["difference"]):
        q = m.mix(q, np.array(d))
        errs.append(float(np.abs(q - np.array(g["expected_charges"][i + 1])).max()))
    return errs


def test_broyden_matches_reference_regression_vectors():
    g = golden()
    errs = run_reference_sequence(make(g["params"]), g)
    assert max(errs) <= g["atol_eps"] * EPS


def test_first_step_is_simple_mixing():
    g = golden()
    p = g["params"]
    m = make(p)
    m.reset(3)
    q0 = np.array(g["expected_charges"][0])
    d0 = np.array(g["difference"][0])
    assert np.array_equal(m.mix(q0, d0), q0 + p["mixParam"] * d0)
    s = SimpleMixer(p["mixParam"])
    s.reset(3)
    assert np.array_equal(s.mix(q0, d0), q0 + p["mixParam"] * d0)


class _NoHistory(BroydenMixer):
    """Planted mistake: the stored |u> vectors of earlier iterations are never subtracted."""

    def mix(self, q_inp, q_diff):
        saved = self.uu.copy()
        self.uu[:] = 0.0
        out = super().mix(q_inp, q_diff)
        keep = self.i_iter - 1
        self.uu[:keep - 1] = saved[:keep - 1]
        return out


class _NoOmega(BroydenMixer):
    """Planted mistake: the omega0^2 regularisation of beta is dropped."""

    def __init__(self, *a):
        super().__init__(*a)
        self.omega0 = 1e-300


def test_planted_mistakes_miss_the_reference_vectors():
    g = golden()
    tol = g["atol_eps"] * EPS
    assert max(run_reference_sequence(make(g["params"], _NoHistory), g)) > tol
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # without omega0 the 3-vector system goes singular
        assert not max(run_reference_sequence(make(g["params"], _NoOmega), g)) <= tol


def test_broyden_solves_a_linear_fixed_point():
    # a symmetric response with a few slow modes (eigenvalues 0.5..0.97) over weak ones
    rng = np.random.default_rng(3)
    n = 200
    v, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.concatenate([np.linspace(0.5, 0.97, 8), rng.uniform(-0.1, 0.1, n - 8)])
    jac = (v * lam) @ v.T
    b = rng.normal(size=n)
    exact = np.linalg.solve(np.eye(n) - jac, b)

    def residual_after(mixer, steps):
        mixer.reset(n)
        q = np.zeros(n)
        for _ in range(steps):
            q = mixer.mix(q, jac @ q + b - q)
        return np.abs(q - exact).max() / np.abs(exact).max()

    assert residual_after(BroydenMixer(0.2, 100), 60) < 1e-10
    assert residual_after(SimpleMixer(0.2), 60) > 1e-3
