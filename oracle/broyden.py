"""CPU restatement of DFTB+'s charge mixers (simple and modified Broyden) for the device mixer.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the HIP mixer; the product library
never imports or calls it.

Follows src/dftbp/mixer/broydenmixer.F90 (TBroydenMixerReal: init 114-140, reset 144-169, mix
180-200, modifiedBroydenMixing 204-342; Johnson, PRB 38, 12807 (1988)) and
src/dftbp/mixer/simplemixer.F90 (mix 77-92).  The small linear system is solved with LAPACK's
dgetrf / dgetrs(trans='T') through scipy, as the reference's getrf / getrs wrappers do
(src/dftbp/math/lapackroutines.F90).

Pinned by the reference's own regression vectors (test/src/dftbp/unit/mixer/broyden.F90:25-64,
a water molecule's charges over four iterations, atol 100 eps), committed as
tests/golden/ref_broyden_h2o.json, and by the closed form of the first (simple-mixing) step.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

EPS = np.finfo(np.float64).eps


class SimpleMixer:
    """qInpResult <- qInpResult + mixParam * qDiff (simplemixer.F90:77-92)."""

    def __init__(self, mix_param: float):
        self.alpha = float(mix_param)

    def reset(self, n_elem: int) -> None:
        assert n_elem > 0

    def mix(self, q_inp: np.ndarray, q_diff: np.ndarray) -> np.ndarray:
        return q_inp + self.alpha * q_diff


class BroydenMixer:
    """Modified Broyden mixer with the reference's parameters and state."""

    def __init__(self, mix_param: float, max_iter: int, omega0: float = 0.01,
                 min_weight: float = 1.0, max_weight: float = 1.0e5, weight_fac: float = 1.0e-2):
        # broydenmixer.F90:114-140
        assert max_iter > 0 and mix_param > 0.0 and omega0 > 0.0
        self.m_iter = int(max_iter)
        self.alpha = float(mix_param)
        self.omega0 = float(omega0)
        self.min_weight = float(min_weight)
        self.max_weight = float(max_weight)
        self.weight_fac = float(weight_fac)
        self.n_elem = 0
        self.i_iter = 0

    def reset(self, n_elem: int) -> None:
        # broydenmixer.F90:144-169
        assert n_elem > 0
        m = self.m_iter - 1
        self.n_elem = int(n_elem)
        self.q_inp_last = np.zeros(n_elem)
        self.q_diff_last = np.zeros(n_elem)
        self.dF = np.zeros((max(m, 0), n_elem))
        self.uu = np.zeros((max(m, 0), n_elem))
        self.ww = np.zeros(max(m, 0))
        self.aa = np.zeros((max(m, 0), max(m, 0)))
        self.i_iter = 0

    def mix(self, q_inp: np.ndarray, q_diff: np.ndarray) -> np.ndarray:
        """Returns the mixed charges (the reference's qInpResult on exit)."""
        # broydenmixer.F90:180-200
        self.i_iter += 1
        if self.i_iter > self.m_iter:
            raise RuntimeError("Broyden mixer: Maximal nr. of steps exceeded")
        q = np.array(q_inp, dtype=np.float64, copy=True)
        qd = np.asarray(q_diff, dtype=np.float64)
        nn = self.i_iter
        nn_1 = nn - 1
        # first iteration: simple mix (272-277)
        if nn == 1:
            self.q_inp_last[:] = q
            self.q_diff_last[:] = qd
            return q + self.alpha * qd
        # weight of this iteration (285-293)
        w = np.sqrt(np.dot(qd, qd))
        w = self.weight_fac / w if w > self.weight_fac / self.max_weight else self.max_weight
        w = max(w, self.min_weight)
        self.ww[nn_1 - 1] = w
        # normalised |dF(m-1)> (296-300)
        df_uu = qd - self.q_diff_last
        inv_norm = max(np.sqrt(np.dot(df_uu, df_uu)), EPS)
        inv_norm = 1.0 / inv_norm
        df_uu = inv_norm * df_uu
        # a, c and beta (304-315); Fortran index ii = 1..nn-2 is row ii-1 here
        cc = np.zeros(nn_1)
        for ii in range(nn - 2):
            self.aa[ii, nn_1 - 1] = np.dot(self.dF[ii], df_uu)
            self.aa[nn_1 - 1, ii] = self.aa[ii, nn_1 - 1]
            cc[ii] = self.ww[ii] * np.dot(self.dF[ii], qd)
        self.aa[nn_1 - 1, nn_1 - 1] = 1.0
        cc[nn_1 - 1] = self.ww[nn_1 - 1] * np.dot(df_uu, qd)
        wv = self.ww[:nn_1]
        beta = np.empty((nn_1, nn_1))
        for ii in range(nn_1):
            beta[:, ii] = wv * wv[ii] * self.aa[:nn_1, ii]
            beta[ii, ii] += self.omega0 ** 2
        # LU decomposition and the transposed solve (318-320)
        lu, piv = sla.lu_factor(beta)
        cc = sla.lu_solve((lu, piv), cc, trans=1)
        # store |dF(m-1)>, build |u(m-1)> (323-330)
        self.dF[nn_1 - 1] = df_uu
        u = self.alpha * df_uu + inv_norm * (q - self.q_inp_last)
        self.q_inp_last[:] = q
        self.q_diff_last[:] = qd
        # new vector (333-340)
        q = q + self.alpha * qd
        for ii in range(nn - 2):
            q = q - self.ww[ii] * cc[ii] * self.uu[ii]
        q = q - self.ww[nn_1 - 1] * cc[nn_1 - 1] * u
        self.uu[nn_1 - 1] = u
        return q
