"""ctypes bindings of the test oracle (``oracle/lib/libewald_oracle.so``) and of the
reference's own Fortran modules compiled from source (``oracle/_ref/libewald_ref.so``).

TEST INFRASTRUCTURE ONLY.  The tests, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` use this package as the checker; the product library never imports,
links or calls it.

Layouts follow the C restatement: positions ``(n, 3)``, lattice ``(3, 3)`` with row i the
lattice vector a_i, charges ``(n,)``; everything float64, atomic units.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libewald_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libewald_ref.so")
REFERENCE_SRC = "/root/reference/src/dftbp/extlibs/fftewald"

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build(with_reference: bool | None = None) -> None:
    """Compile the oracle (and, where the reference sources exist, ``_ref``)."""
    subprocess.run(["make", "-C", HERE, "-s", "all"], check=True)
    if with_reference is None:
        with_reference = os.path.isdir(REFERENCE_SRC)
    if with_reference:
        subprocess.run(["make", "-C", HERE, "-s", "ref"], check=True)


def _ptr(a: np.ndarray | None, ty=_D):
    if a is None:
        return None
    return a.ctypes.data_as(ty)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle status {code}")
        self.code = code


class _CellMetrics(C.Structure):
    _fields_ = [("lat", C.c_double * 9), ("rec", C.c_double * 9), ("volume", C.c_double),
                ("latlen", C.c_double * 3), ("reclen", C.c_double * 3)]


class FftParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("r_cut", C.c_double), ("k_cut", C.c_double),
                ("window_cutoff", C.c_int), ("nmodes", C.c_int * 3)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_reference=False)
        L = C.CDLL(LIB_PATH)
        L.orc_cell_metrics.argtypes = [_D, C.POINTER(_CellMetrics)]
        L.orc_check_configuration.argtypes = [C.c_int, _D]
        L.orc_self_potential.restype = C.c_double
        L.orc_self_potential.argtypes = [C.c_double, C.c_double]
        L.orc_direct_cutoffs.argtypes = [C.c_double, C.c_double, _D, _D]
        L.orc_lambert_w0.restype = C.c_double
        L.orc_lambert_w0.argtypes = [C.c_double]
        L.orc_splitting_from_real_budget.restype = C.c_double
        L.orc_splitting_from_real_budget.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double,
                                                     C.c_double]
        L.orc_cutoff_from_fourier_budget.restype = C.c_double
        L.orc_cutoff_from_fourier_budget.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double,
                                                     C.c_double]
        L.orc_balanced_alpha.argtypes = [C.c_double] * 4 + [_D]
        L.orc_balanced_splitting.restype = C.c_double
        L.orc_balanced_splitting.argtypes = [C.POINTER(_CellMetrics), C.c_double]
        L.orc_optimal_alpha_dftb.argtypes = [_D, C.c_double, _D]
        for fn in (L.orc_real_space_images, L.orc_real_space_cell_list):
            fn.argtypes = [C.c_int, _D, _D, C.POINTER(_CellMetrics), C.c_double, C.c_double, _D, _D]
        L.orc_direct_ewald.argtypes = [C.c_int, _D, _D, _D, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, C.c_int, _D, _D, _D]
        L.orc_next_fast_length.argtypes = [C.c_int]
        L.orc_window_cutoff_from_budget.argtypes = [C.c_double]
        L.orc_choose_parameters.argtypes = [C.POINTER(_CellMetrics), C.c_double, C.c_int,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_int, C.c_int, C.POINTER(FftParams)]
        L.orc_grid_memory_gigabytes.restype = C.c_double
        L.orc_grid_memory_gigabytes.argtypes = [C.POINTER(_CellMetrics), C.c_double, C.c_int,
                                                C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.c_int, C.c_int]
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _D, C.c_int]
        L.orc_nfft_adjoint_3d.argtypes = [C.c_int, _D, _D, _D, _D, C.c_int, C.c_int, C.c_int,
                                          C.c_int, _D]
        L.orc_nfft_forward_3d.argtypes = [C.c_int, _D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                          _D, _D]
        L.orc_fft_ewald.argtypes = [C.c_int, _D, _D, _D, C.c_double, _D, _D, C.POINTER(FftParams)]
        L.orc_spme_reciprocal.argtypes = [C.c_int, _D, _D, _D, _I, C.c_int, C.c_double, _D, _D]
        L.orc_spme_ewald.argtypes = [C.c_int, _D, _D, _D, _I, C.c_int, C.c_double, C.c_double,
                                     _D, _D]
        L.orc_bspline_moduli.argtypes = [C.c_int, C.c_int, _D]
        L.orc_bspline_weights.argtypes = [C.c_double, C.c_int, _D, _D]
        L.orc_exp_gamma.restype = C.c_double
        L.orc_exp_gamma.argtypes = [C.c_double] * 3
        L.orc_exp_gamma_cutoff.restype = C.c_double
        L.orc_exp_gamma_cutoff.argtypes = [C.c_double] * 3
        L.orc_short_gamma_shift.argtypes = [C.c_int, _D, _D, _D, _I, _D, C.c_int, _D]
        _lib = L
    return _lib


def have_reference() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's Fortran modules (built from /root/reference by ``make ref``)."""
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.ref_direct_ewald.argtypes = [C.c_int, _D, _D, _D, C.c_double, C.c_double, C.c_int, _D, _D]
        L.ref_short_range.argtypes = [C.c_int, _D, _D, _D, C.c_double, C.c_double, _D, _D]
        L.ref_cell_metrics.argtypes = [_D, _D, _D, _D, _D]
        L.ref_balanced_alpha.argtypes = [C.c_double] * 4 + [_D, _I]
        L.ref_balanced_splitting.restype = C.c_double
        L.ref_balanced_splitting.argtypes = [_D, C.c_double]
        L.ref_direct_cutoffs.argtypes = [C.c_double, C.c_double, _D, _D]
        L.ref_lambert_w0.restype = C.c_double
        L.ref_lambert_w0.argtypes = [C.c_double]
        L.ref_splitting_from_real_budget.restype = C.c_double
        L.ref_splitting_from_real_budget.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double,
                                                     C.c_double]
        L.ref_cutoff_from_fourier_budget.restype = C.c_double
        L.ref_cutoff_from_fourier_budget.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double,
                                                     C.c_double]
        _ref = L
    return _ref


# --------------------------------------------------------------------------------------------
# oracle wrappers
# --------------------------------------------------------------------------------------------

def cell_metrics(lat) -> dict:
    lat = _f64(lat)
    m = _CellMetrics()
    st = lib().orc_cell_metrics(_ptr(lat), C.byref(m))
    if st:
        raise OracleError(st, "cell_metrics")
    return {"rec": np.array(m.rec[:]).reshape(3, 3), "volume": m.volume,
            "latlen": np.array(m.latlen[:]), "reclen": np.array(m.reclen[:]), "_c": m}


def check_configuration(q) -> int:
    q = _f64(q)
    return lib().orc_check_configuration(len(q), _ptr(q))


def direct_cutoffs(alpha: float, tol: float):
    r, k = C.c_double(), C.c_double()
    lib().orc_direct_cutoffs(alpha, tol, C.byref(r), C.byref(k))
    return r.value, k.value


def lambert_w0(x: float) -> float:
    return lib().orc_lambert_w0(x)


def splitting_from_real_budget(r_cut: float, q2sum: float, n: int, pseudo_volume: float,
                               budget: float) -> float:
    return lib().orc_splitting_from_real_budget(r_cut, q2sum, n, pseudo_volume, budget)


def cutoff_from_fourier_budget(q2sum: float, n: int, box_edge: float, alpha: float,
                               budget: float) -> float:
    return lib().orc_cutoff_from_fourier_budget(q2sum, n, box_edge, alpha, budget)


def balanced_alpha(minG, minR, volume, tol):
    a = C.c_double()
    st = lib().orc_balanced_alpha(minG, minR, volume, tol, C.byref(a))
    return a.value, st


def balanced_splitting(lat, tol: float) -> float:
    m = cell_metrics(lat)
    return lib().orc_balanced_splitting(C.byref(m["_c"]), tol)


def optimal_alpha_dftb(lat, tol: float) -> float:
    lat = _f64(lat)
    a = C.c_double()
    st = lib().orc_optimal_alpha_dftb(_ptr(lat), tol, C.byref(a))
    if st:
        raise OracleError(st, "optimal_alpha_dftb")
    return a.value


def real_space(pos, q, lat, alpha, r_cut, method="cell_list", want_force=False):
    pos, q = _f64(pos), _f64(q)
    n = len(q)
    m = cell_metrics(lat)
    pot = np.zeros(n)
    force = np.zeros((n, 3)) if want_force else None
    fn = lib().orc_real_space_cell_list if method == "cell_list" else lib().orc_real_space_images
    st = fn(n, _ptr(pos), _ptr(q), C.byref(m["_c"]), alpha, r_cut, _ptr(pot), _ptr(force))
    if st:
        raise OracleError(st, "real_space")
    return (pot, force) if want_force else pot


def direct_ewald(pos, q, lat, tol=1e-12, alpha=None, r_cut=None, k_cut=None, want_force=False,
                 real_space="images", phase_tables=False):
    """Direct O(N^2) Ewald; returns (pot, force or None, alpha)."""
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    pot = np.zeros(n)
    force = np.zeros((n, 3)) if want_force else None
    a = C.c_double()
    st = lib().orc_direct_ewald(n, _ptr(pos), _ptr(q), _ptr(lat), tol, alpha or 0.0,
                                r_cut or 0.0, k_cut or 0.0,
                                1 if real_space == "cell_list" else 0, 1 if phase_tables else 0,
                                _ptr(pot), _ptr(force), C.byref(a))
    if st:
        raise OracleError(st, "direct_ewald")
    return pot, force, a.value


def next_fast_length(n: int) -> int:
    return lib().orc_next_fast_length(n)


def window_cutoff_from_budget(b: float) -> int:
    return lib().orc_window_cutoff_from_budget(b)


def choose_parameters(lat, q, tol, alpha=0.0, r_cut=0.0, k_cut=0.0, nmodes=0, window=0):
    q = _f64(q)
    m = cell_metrics(lat)
    p = FftParams()
    st = lib().orc_choose_parameters(C.byref(m["_c"]), float(np.sum(q * q)), len(q), tol, alpha,
                                     r_cut, k_cut, nmodes, window, C.byref(p))
    if st:
        raise OracleError(st, "choose_parameters")
    return {"alpha": p.alpha, "r_cut": p.r_cut, "k_cut": p.k_cut,
            "window_cutoff": p.window_cutoff, "nmodes": tuple(p.nmodes[:])}


def grid_memory_gigabytes(lat, q, tol) -> float:
    q = _f64(q)
    m = cell_metrics(lat)
    return lib().orc_grid_memory_gigabytes(C.byref(m["_c"]), float(np.sum(q * q)), len(q), tol,
                                           0.0, 0.0, 0.0, 0, 0)


def fft3d(grid: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised DFT of a complex array indexed [i3, i2, i1] (i1 fastest)."""
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    n3, n2, n1 = g.shape
    st = lib().orc_fft3d(n1, n2, n3, g.view(np.float64).ctypes.data_as(_D), sign)
    if st:
        raise OracleError(st, "fft3d")
    return g


def nfft_adjoint(t, q, M, m):
    t = _f64(t)
    q = _f64(q)
    out = np.zeros((M[2], M[1], M[0]), dtype=np.complex128)
    st = lib().orc_nfft_adjoint_3d(len(q), _ptr(np.ascontiguousarray(t[:, 0])),
                                   _ptr(np.ascontiguousarray(t[:, 1])),
                                   _ptr(np.ascontiguousarray(t[:, 2])), _ptr(q), M[0], M[1],
                                   M[2], m, out.view(np.float64).ctypes.data_as(_D))
    if st:
        raise OracleError(st, "nfft_adjoint")
    return out


def nfft_forward(t, fhat, m):
    t = _f64(t)
    fhat = np.ascontiguousarray(fhat, dtype=np.complex128)
    M3, M2, M1 = fhat.shape
    out = np.zeros(len(t), dtype=np.complex128)
    st = lib().orc_nfft_forward_3d(len(t), _ptr(np.ascontiguousarray(t[:, 0])),
                                   _ptr(np.ascontiguousarray(t[:, 1])),
                                   _ptr(np.ascontiguousarray(t[:, 2])), M1, M2, M3, m,
                                   fhat.view(np.float64).ctypes.data_as(_D),
                                   out.view(np.float64).ctypes.data_as(_D))
    if st:
        raise OracleError(st, "nfft_forward")
    return out


def fft_ewald(pos, q, lat, tol=1e-10, want_force=False):
    """The reference fast path restated (linked cells + Gaussian NFFT)."""
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    pot = np.zeros(n)
    force = np.zeros((n, 3)) if want_force else None
    p = FftParams()
    st = lib().orc_fft_ewald(n, _ptr(pos), _ptr(q), _ptr(lat), tol, _ptr(pot), _ptr(force),
                             C.byref(p))
    if st:
        raise OracleError(st, "fft_ewald")
    params = {"alpha": p.alpha, "r_cut": p.r_cut, "k_cut": p.k_cut,
              "window_cutoff": p.window_cutoff, "nmodes": tuple(p.nmodes[:])}
    return pot, force, params


def spme_reciprocal(pos, q, lat, K, order, alpha, want_force=False):
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    Karr = np.ascontiguousarray(K, dtype=np.int32)
    pot = np.zeros(n)
    force = np.zeros((n, 3)) if want_force else None
    st = lib().orc_spme_reciprocal(n, _ptr(pos), _ptr(q), _ptr(lat), _ptr(Karr, _I), order, alpha,
                                   _ptr(pot), _ptr(force))
    if st:
        raise OracleError(st, "spme_reciprocal")
    return (pot, force) if want_force else pot


def spme_ewald(pos, q, lat, K, order, alpha, r_cut, want_force=False):
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    Karr = np.ascontiguousarray(K, dtype=np.int32)
    pot = np.zeros(n)
    force = np.zeros((n, 3)) if want_force else None
    st = lib().orc_spme_ewald(n, _ptr(pos), _ptr(q), _ptr(lat), _ptr(Karr, _I), order, alpha,
                              r_cut, _ptr(pot), _ptr(force))
    if st:
        raise OracleError(st, "spme_ewald")
    return (pot, force) if want_force else pot


def bspline_moduli(K: int, order: int) -> np.ndarray:
    out = np.zeros(K)
    lib().orc_bspline_moduli(K, order, _ptr(out))
    return out


def bspline_weights(u: float, order: int):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().orc_bspline_weights(u, order, _ptr(th), _ptr(dth))
    return th, dth


def exp_gamma(r: float, ua: float, ub: float) -> float:
    return lib().orc_exp_gamma(r, ua, ub)


def exp_gamma_cutoff(u2: float, u1: float, min_value: float = 0.0) -> float:
    return lib().orc_exp_gamma_cutoff(u2, u1, min_value)


def short_gamma_shift(pos, q, lat, species, hubbard_u):
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    u = _f64(hubbard_u)
    out = np.zeros(len(q))
    st = lib().orc_short_gamma_shift(len(q), _ptr(pos), _ptr(q), _ptr(lat), _ptr(sp, _I), _ptr(u),
                                     len(u), _ptr(out))
    if st:
        raise OracleError(st, "short_gamma_shift")
    return out


# --------------------------------------------------------------------------------------------
# reference (compiled Fortran) wrappers
# --------------------------------------------------------------------------------------------

def ref_direct_ewald(pos, q, lat, tol, alpha=None, want_force=False):
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    pot = np.zeros(n)
    force = np.zeros((n, 3))
    ref().ref_direct_ewald(n, _ptr(pos), _ptr(q), _ptr(lat), tol, alpha or 0.0,
                           1 if want_force else 0, _ptr(pot), _ptr(force))
    return (pot, force) if want_force else pot


def ref_short_range(pos, q, lat, alpha, r_cut):
    pos, q, lat = _f64(pos), _f64(q), _f64(lat)
    n = len(q)
    pot = np.zeros(n)
    force = np.zeros((n, 3))
    ref().ref_short_range(n, _ptr(pos), _ptr(q), _ptr(lat), alpha, r_cut, _ptr(pot), _ptr(force))
    return pot, force


def ref_balanced_alpha(minG, minR, volume, tol):
    a = C.c_double()
    st = C.c_int()
    ref().ref_balanced_alpha(minG, minR, volume, tol, C.byref(a), C.byref(st))
    return a.value, st.value


def ref_lambert_w0(x: float) -> float:
    return ref().ref_lambert_w0(x)


def ref_splitting_from_real_budget(r_cut, q2sum, n, pseudo_volume, budget) -> float:
    return ref().ref_splitting_from_real_budget(r_cut, q2sum, n, pseudo_volume, budget)


def ref_cutoff_from_fourier_budget(q2sum, n, box_edge, alpha, budget) -> float:
    return ref().ref_cutoff_from_fourier_budget(q2sum, n, box_edge, alpha, budget)


def ref_balanced_splitting(lat, tol):
    lat = _f64(lat)
    return ref().ref_balanced_splitting(_ptr(lat), tol)


def ref_cell_metrics(lat):
    lat = _f64(lat)
    rec = np.zeros((3, 3))
    vol = C.c_double()
    ll = np.zeros(3)
    rl = np.zeros(3)
    ref().ref_cell_metrics(_ptr(lat), _ptr(rec), C.byref(vol), _ptr(ll), _ptr(rl))
    return {"rec": rec, "volume": vol.value, "latlen": ll, "reclen": rl}
