"""ctypes binding of the MI355X periodic Coulomb engine (``lib/libfftewald_amd.so``).

Mirrors the reference library's public surface (src/dftbp/extlibs/fftewald/fft/ewald_fft_3d.f90:
``ewald_potential_force``, ``ewald_parameters``, ``ewald_grid_memory``) and the resident
per-SCC-iteration use TCoulomb makes of it (src/dftbp/dftb/coulomb.F90:336-357, 786-880), in the
style of the DFTB+ ctypes wrapper (tools/pythonapi/src/dftbplus/dftbplus.py:72-75, 384-399):
``np.ctypeslib.ndpointer`` arguments and one checking wrapper per call.

There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libfftewald_amd.so")

EW_OK = 0
EW_ERR_ARGUMENT = 1
EW_ERR_NOT_NEUTRAL = 2
EW_ERR_CELL = 3
EW_ERR_TOLERANCE = 4
EW_ERR_COINCIDENT = 5
EW_ERR_NO_GEOMETRY = 6
EW_ERR_DEVICE = 7
EW_ERR_MEMORY = 8
EW_ERR_NOT_FINITE = 9
EW_ERR_PENDING = 10

PRECISION_FP64 = 0
PRECISION_MIXED = 1

#: most charge vectors of one ew_shift_multi call
MAX_VECTORS = 8

#: every symbol the C ABI (include/fftewald_amd.h) exports
EXPORTED = ("ew_create", "ew_destroy", "ew_last_error", "ew_set_geometry", "ew_get_parameters",
            "ew_shift", "ew_shift_device", "ew_shift_device_noalloc", "ew_shift_multi", "ew_shift_multi_device",
            "ew_gradients", "ew_stress", "ew_external_potential", "ew_external_gradients",
            "ew_phase_times", "ew_trace_times", "ew_get_tuning",
            "ew_potential_force",
            "ew_parameters_for", "ew_grid_memory", "ew_pool_create", "ew_pool_destroy",
            "ew_pool_last_error", "ew_pool_add", "ew_pool_remove", "ew_pool_parameters",
            "ew_pool_usage", "ew_pool_shift", "ew_pool_shift_device", "ew_pool_sync",
            "ew_server_create", "ew_server_destroy", "ew_server_last_error", "ew_server_submit",
            "ew_server_wait", "ew_server_drain", "ew_server_stats", "ew_mixer_create",
            "ew_mixer_destroy", "ew_mixer_last_error", "ew_mixer_reset", "ew_mixer_mix",
            "ew_mixer_mix_device", "ew_mixer_iteration")

#: phases reported by ew_phase_times, in order
PHASES = ("gather_charges", "spread", "fft_x_r2c", "fft_y", "fft_z_influence", "fft_y_inverse",
          "fft_x_c2r", "pair_sum", "gather_combine")
#: spans reported by ew_trace_times (the evaluation's own launch shape), in order
TRACE = ("charges", "pair_sum", "spread_region", "spread_assemble", "fft_x_r2c", "fft_y",
         "fft_z_influence", "fft_y_inverse", "fft_x_c2r", "gather_mesh", "join_wait",
         "gather_combine", "evaluation")


class EwaldError(RuntimeError):
    """A non-zero status of the C ABI; ``code`` is the EW_ERR_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class Tuning(C.Structure):
    """Mirrors struct ew_tuning: launch shapes of one structure's evaluation."""
    _fields_ = [("spread_direct", C.c_int), ("spread_block", C.c_int * 3),
                ("fft_planes", C.c_int), ("fft_lines", C.c_int * 3), ("fft_radix", C.c_int * 3),
                ("pair_cosched", C.c_int), ("pair_spread_share", C.c_int),
                ("eval_us", C.c_float), ("fft_wave", C.c_int), ("fft_lanes", C.c_int * 3),
                ("pair_blocks", C.c_int), ("pair_fork", C.c_int)]

    def as_dict(self):
        return {"spread_direct": self.spread_direct, "spread_block": tuple(self.spread_block),
                "fft_planes": self.fft_planes, "fft_lines": tuple(self.fft_lines),
                "fft_radix": tuple(self.fft_radix), "pair_cosched": self.pair_cosched,
                "pair_spread_share": self.pair_spread_share, "eval_us": self.eval_us,
                "fft_wave": self.fft_wave, "fft_lanes": tuple(self.fft_lanes),
                "pair_blocks": self.pair_blocks, "pair_fork": self.pair_fork}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        t.spread_direct = int(d["spread_direct"])
        t.fft_planes = int(d["fft_planes"])
        t.pair_cosched = int(d.get("pair_cosched", 0))
        t.pair_spread_share = int(d.get("pair_spread_share", 0))
        t.fft_wave = int(d.get("fft_wave", 0))
        t.pair_blocks = int(d.get("pair_blocks", 0))
        t.pair_fork = int(d.get("pair_fork", 0))
        for a in range(3):
            t.fft_lanes[a] = int(d.get("fft_lanes", (0, 0, 0))[a])
            t.spread_block[a] = int(d["spread_block"][a])
            t.fft_lines[a] = int(d["fft_lines"][a])
            t.fft_radix[a] = int(d["fft_radix"][a])
        return t


TUNE = {"rules": 0, "auto": 1, "cached": 2, "given": 3}


class Options(C.Structure):
    _fields_ = [("tol", C.c_double), ("alpha", C.c_double), ("r_cut", C.c_double),
                ("grid", C.c_int * 3), ("order", C.c_int), ("precision", C.c_int),
                ("device", C.c_int), ("alpha_rule", C.c_int), ("profile", C.c_int),
                ("eager", C.c_int), ("charged", C.c_int), ("fft_mode", C.c_int),
                ("pair_storage", C.c_int), ("group_grid", C.c_int), ("k_cut", C.c_double),
                ("tune", C.c_int), ("tuning", Tuning)]


class ServerOptions(C.Structure):
    _fields_ = [("max_batch", C.c_int), ("max_inflight", C.c_int), ("reserved", C.c_int * 6)]


MIXER_BROYDEN = 0
MIXER_SIMPLE = 1


class MixerOptions(C.Structure):
    _fields_ = [("kind", C.c_int), ("max_iter", C.c_int), ("mix_param", C.c_double),
                ("omega0", C.c_double), ("min_weight", C.c_double), ("max_weight", C.c_double),
                ("weight_fac", C.c_double), ("device", C.c_int), ("reserved", C.c_int * 3)]


class ServerReport(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong),
                ("failed", C.c_longlong), ("batches", C.c_longlong), ("mean_batch", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Parameters(C.Structure):
    _fields_ = [("alpha", C.c_double), ("r_cut", C.c_double), ("k_cut", C.c_double),
                ("r_list", C.c_double), ("grid", C.c_int * 3), ("order", C.c_int),
                ("precision", C.c_int), ("pair_mode", C.c_int), ("index_bytes", C.c_int),
                ("fft_planes", C.c_int), ("gather_split", C.c_int), ("n_pairs", C.c_longlong),
                ("device_bytes", C.c_longlong)]

    def as_dict(self) -> dict:
        return {"alpha": self.alpha, "r_cut": self.r_cut, "k_cut": self.k_cut,
                "r_list": self.r_list, "grid": tuple(self.grid), "order": self.order,
                "precision": self.precision, "pair_mode": self.pair_mode,
                "index_bytes": self.index_bytes, "fft_planes": self.fft_planes,
                "gather_split": self.gather_split,
                "n_pairs": self.n_pairs, "device_bytes": self.device_bytes}


_D1 = np.ctypeslib.ndpointer(dtype=np.float64, ndim=1, flags="C_CONTIGUOUS")
_D2 = np.ctypeslib.ndpointer(dtype=np.float64, ndim=2, flags="C_CONTIGUOUS")
_I1 = np.ctypeslib.ndpointer(dtype=np.int32, ndim=1, flags="C_CONTIGUOUS")

_lib = None


def _optional(nd):
    """ndpointer that also accepts None (passed as NULL)."""

    class _Opt(nd):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return super().from_param(obj)

    return _Opt


def lib() -> C.CDLL:
    """Load the HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"fftewald_amd: {LIB_PATH} is missing -- build it first "
                              "(python -c 'import __graft_entry__ as g; g.build()')")
        try:
            # torch ships its own libamdhip64.so.7; loading it first makes the engine bind to the
            # same HIP runtime, so one process never holds two runtimes.
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        L.ew_create.argtypes = [C.POINTER(Options), C.POINTER(H)]
        L.ew_destroy.argtypes = [H]
        L.ew_last_error.restype = C.c_char_p
        L.ew_last_error.argtypes = [H]
        L.ew_set_geometry.argtypes = [H, C.c_int, _D2, _D2, _optional(_I1), _optional(_D1),
                                      C.c_int, _optional(_D1)]
        L.ew_get_parameters.argtypes = [H, C.POINTER(Parameters)]
        L.ew_shift.argtypes = [H, _D1, _D1, C.POINTER(C.c_double)]
        L.ew_shift_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ew_shift_device_noalloc.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ew_shift_multi.argtypes = [H, C.c_int, _D2, _D2, _D1]
        L.ew_shift_multi_device.argtypes = [H, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
        L.ew_gradients.argtypes = [H, _D1, _D2]
        L.ew_stress.argtypes = [H, _D1, _D2]
        L.ew_external_potential.argtypes = [H, C.c_int, _optional(_D2), _optional(_D1),
                                            _optional(_D1), C.c_double, _D1]
        L.ew_external_gradients.argtypes = [H, C.c_int, _optional(_D2), _optional(_D1),
                                            _optional(_D1), _optional(_D1), _D2, _optional(_D2)]
        L.ew_phase_times.argtypes = [H, np.ctypeslib.ndpointer(dtype=np.float64, ndim=1,
                                                                flags="C_CONTIGUOUS"),
                                     C.POINTER(C.c_longlong), C.c_int]
        L.ew_get_tuning.argtypes = [H, C.POINTER(Tuning)]
        L.ew_trace_times.argtypes = [H, np.ctypeslib.ndpointer(dtype=np.float64, ndim=1,
                                                                flags="C_CONTIGUOUS"),
                                     C.POINTER(C.c_longlong), C.c_int]
        L.ew_potential_force.argtypes = [C.c_int, _D2, _D1, _D2, C.POINTER(Options), _D1,
                                         _optional(_D2)]
        L.ew_parameters_for.argtypes = [C.c_int, _optional(_D1), _D2, C.POINTER(Options),
                                        C.POINTER(Parameters)]
        L.ew_grid_memory.argtypes = [C.c_int, _optional(_D1), _D2, C.POINTER(Options),
                                     C.POINTER(C.c_double)]
        L.ew_pool_create.argtypes = [C.POINTER(Options), C.c_longlong, C.POINTER(H)]
        L.ew_pool_destroy.argtypes = [H]
        L.ew_pool_last_error.restype = C.c_char_p
        L.ew_pool_last_error.argtypes = [H]
        L.ew_pool_add.argtypes = [H, C.c_int, _D2, _D2, _optional(_I1), _optional(_D1), C.c_int,
                                  _optional(_D1), _optional(_I1), C.POINTER(C.c_int)]
        L.ew_pool_remove.argtypes = [H, C.c_int]
        L.ew_pool_parameters.argtypes = [H, C.c_int, C.POINTER(Parameters)]
        L.ew_pool_usage.argtypes = [H, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_int)]
        L.ew_pool_shift.argtypes = [H, C.c_int, _I1, _D1, _D1, _D1]
        L.ew_pool_shift_device.argtypes = [H, C.c_int, _I1, C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p]
        L.ew_pool_sync.argtypes = [H]
        L.ew_server_create.argtypes = [H, C.POINTER(ServerOptions), C.POINTER(H)]
        L.ew_server_destroy.argtypes = [H]
        L.ew_server_last_error.restype = C.c_char_p
        L.ew_server_last_error.argtypes = [H]
        L.ew_server_submit.argtypes = [H, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_longlong)]
        L.ew_server_wait.argtypes = [H, C.c_longlong, C.c_double]
        L.ew_server_drain.argtypes = [H]
        L.ew_server_stats.argtypes = [H, C.POINTER(ServerReport), C.c_int]
        L.ew_mixer_create.argtypes = [C.POINTER(MixerOptions), C.POINTER(H)]
        L.ew_mixer_destroy.argtypes = [H]
        L.ew_mixer_last_error.restype = C.c_char_p
        L.ew_mixer_last_error.argtypes = [H]
        L.ew_mixer_reset.argtypes = [H, C.c_int]
        L.ew_mixer_mix.argtypes = [H, _D1, _D1]
        L.ew_mixer_mix_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ew_mixer_iteration.argtypes = [H, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def make_options(tol=None, alpha=None, r_cut=None, grid=None, order=None, precision="fp64",
                 device=0, alpha_rule="fftewald", profile=False, eager=False,
                 charged=False, fft_mode=0, pair_storage=0, k_cut=None, tune="rules",
                 tuning=None, group_grid=0) -> Options:
    o = Options()
    o.tol = tol or 0.0
    o.alpha = alpha or 0.0
    o.r_cut = r_cut or 0.0
    o.k_cut = k_cut or 0.0
    if grid is not None:
        g = (grid, grid, grid) if np.isscalar(grid) else tuple(grid)
        for a in range(3):
            o.grid[a] = int(g[a])
    o.order = order or 0
    o.precision = {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}[precision] \
        if isinstance(precision, str) else int(precision)
    o.device = device
    o.alpha_rule = {"fftewald": 0, "dftb": 1}[alpha_rule] if isinstance(alpha_rule, str) \
        else int(alpha_rule)
    o.profile = 2 if profile == "trace" else (1 if profile else 0)
    o.eager = 1 if eager else 0
    o.charged = 1 if charged else 0
    o.fft_mode = int(fft_mode)
    o.pair_storage = int(pair_storage)
    o.group_grid = int(group_grid)
    o.tune = TUNE[tune] if isinstance(tune, str) else int(tune)
    if tuning is not None:
        o.tuning = tuning if isinstance(tuning, Tuning) else Tuning.from_dict(tuning)
        if isinstance(tune, str) and tune == "rules":
            o.tune = TUNE["given"]
    return o


def _check(status: int, handle=None):
    if status != EW_OK:
        msg = lib().ew_last_error(handle).decode()
        raise EwaldError(status, msg)


def _f64(a, ndim) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != ndim:
        raise ValueError(f"expected a {ndim}-d array, got shape {a.shape}")
    return a


class EwaldEngine:
    """Resident engine for one structure: TCoulomb (+ TShortGamma) state on the device.

    ``set_geometry`` once per geometry step, then ``shift`` (host arrays) or ``shift_device``
    (device pointers, e.g. ``tensor.data_ptr()``) once per SCC iteration.
    """

    def __init__(self, **opts):
        self._h = C.c_void_p()
        self.options = make_options(**opts)
        _check(lib().ew_create(C.byref(self.options), C.byref(self._h)))
        self.n = 0

    def close(self):
        if self._h:
            lib().ew_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_geometry(self, coords, lattice, species=None, hubbard_u=None, charges=None):
        coords = _f64(coords, 2)
        lattice = _f64(lattice, 2)
        n = coords.shape[0]
        if coords.shape != (n, 3) or lattice.shape != (3, 3):
            raise ValueError("coords must be (n, 3) and lattice (3, 3)")
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        u = None if hubbard_u is None else _f64(hubbard_u, 1)
        q = None if charges is None else _f64(charges, 1)
        _check(lib().ew_set_geometry(self._h, n, coords, lattice, sp, u,
                                     0 if u is None else len(u), q), self._h)
        self.n = n

    def parameters(self) -> dict:
        p = Parameters()
        _check(lib().ew_get_parameters(self._h, C.byref(p)), self._h)
        return p.as_dict()

    def _per_atom(self, a, name):
        """a as a contiguous float64 vector with one entry per atom of the current geometry (the
        C side reads n entries; without a geometry it reports that itself)."""
        a = _f64(a, 1)
        if self.n and len(a) != self.n:
            raise ValueError(f"{name} must have {self.n} entries (one per atom), got {len(a)}")
        return a

    def shift(self, dq):
        """Returns (shift[n], energy) for charges dq (caller order)."""
        dq = self._per_atom(dq, "dq")
        out = np.empty(self.n)
        e = C.c_double()
        _check(lib().ew_shift(self._h, dq, out, C.byref(e)), self._h)
        return out, e.value

    def shift_device(self, dq_ptr: int, shift_ptr: int, energy_ptr: int = 0, stream: int = 0):
        _check(lib().ew_shift_device(self._h, C.c_void_p(dq_ptr), C.c_void_p(shift_ptr),
                                     C.c_void_p(energy_ptr or None), C.c_void_p(stream or None)),
               self._h)

    def shift_multi(self, dqs):
        """Returns (shift[nvec, n], energy[nvec]) for the charge vectors dqs[nvec, n], evaluated
        together in one grouped launch (nvec <= MAX_VECTORS)."""
        dqs = _f64(dqs, 2)
        nvec = dqs.shape[0]
        if dqs.shape[1] != self.n:
            raise ValueError(f"charge vectors must have {self.n} entries, got {dqs.shape[1]}")
        out = np.empty((nvec, self.n))
        e = np.empty(nvec)
        _check(lib().ew_shift_multi(self._h, nvec, dqs, out, e), self._h)
        return out, e

    def shift_multi_device(self, nvec: int, dq_ptr: int, shift_ptr: int, energy_ptr: int = 0,
                           stream: int = 0):
        _check(lib().ew_shift_multi_device(self._h, int(nvec), C.c_void_p(dq_ptr),
                                           C.c_void_p(shift_ptr), C.c_void_p(energy_ptr or None),
                                           C.c_void_p(stream or None)), self._h)

    def gradients(self, dq, grad=None):
        """grad + dU/dx (Ewald part) at charges dq; grad defaults to zeros (n, 3)."""
        dq = self._per_atom(dq, "dq")
        g = (np.zeros((self.n, 3)) if grad is None
             else np.array(grad, dtype=np.float64, order="C", copy=True))
        if g.shape != (self.n, 3):
            raise ValueError(f"grad must be ({self.n}, 3), got {g.shape}")
        _check(lib().ew_gradients(self._h, dq, g), self._h)
        return g

    def stress(self, dq, stress=None):
        """stress + the Ewald stress (3, 3) at charges dq (TCoulomb%addStress convention)."""
        dq = self._per_atom(dq, "dq")
        st = (np.zeros((3, 3)) if stress is None
              else np.array(stress, dtype=np.float64, order="C", copy=True))
        if st.shape != (3, 3):
            raise ValueError(f"stress must be (3, 3), got {st.shape}")
        _check(lib().ew_stress(self._h, dq, st), self._h)
        return st

    def external_potential(self, coords, charges, blur_widths=None, eps_soften=0.0):
        """Potential (n,) of external point charges at the resident atoms, caller order
        (TCoulomb%getPotential, periodic).  coords (m, 3) Bohr, charges (m,); blur_widths (m,)
        for Gaussian sources; eps_soften > 0 softens the real-space denominator."""
        q = _f64(charges, 1)
        m = len(q)
        xy = _f64(coords, 2) if m else None
        if m and xy.shape != (m, 3):
            raise ValueError("coords must be (m, 3) for m charges")
        w = None if blur_widths is None else _f64(blur_widths, 1)
        if w is not None and len(w) != m:
            raise ValueError("blur_widths must have one entry per charge")
        pot = np.empty(self.n)
        _check(lib().ew_external_potential(self._h, m, xy, q if m else None, w,
                                           float(eps_soften), pot), self._h)
        return pot

    def external_gradients(self, coords, charges, dq=None, blur_widths=None):
        """Derivatives of the atoms' interaction with external point charges.  With atom charges
        dq: (atom_grad (n, 3), ext_grad (m, 3)) = (dq_i sum_s q_s grad ewald(x_i - y_s), minus
        its sum over the atoms per source); without: (sum_s q_s grad ewald(x_i - y_s), None)."""
        q = _f64(charges, 1)
        m = len(q)
        xy = _f64(coords, 2) if m else None
        if m and xy.shape != (m, 3):
            raise ValueError("coords must be (m, 3) for m charges")
        w = None if blur_widths is None else _f64(blur_widths, 1)
        if w is not None and len(w) != m:
            raise ValueError("blur_widths must have one entry per charge")
        d = None if dq is None else self._per_atom(dq, "dq")
        ga = np.zeros((self.n, 3))
        gs = np.zeros((m, 3)) if d is not None else None
        _check(lib().ew_external_gradients(self._h, m, xy, q if m else None, w, d, ga,
                                           gs if (gs is not None and m) else None), self._h)
        return ga, gs

    def phase_times(self, reset: bool = False):
        """({phase: ms}, calls) accumulated since the last reset (needs profile=True)."""
        ms = np.zeros(len(PHASES))
        calls = C.c_longlong()
        _check(lib().ew_phase_times(self._h, ms, C.byref(calls), 1 if reset else 0), self._h)
        return dict(zip(PHASES, ms.tolist())), calls.value

    def tuning(self) -> dict:
        """Launch shapes of the current geometry (ew_get_tuning)."""
        t = Tuning()
        _check(lib().ew_get_tuning(self._h, C.byref(t)), self._h)
        return t.as_dict()

    def trace_times(self, reset: bool = False):
        """({span: ms}, calls) accumulated since the last reset, timed in the evaluation's own
        launch shape (needs profile="trace")."""
        ms = np.zeros(len(TRACE))
        calls = C.c_longlong()
        _check(lib().ew_trace_times(self._h, ms, C.byref(calls), 1 if reset else 0), self._h)
        return dict(zip(TRACE, ms.tolist())), calls.value


class EwaldPool:
    """Many resident structures evaluated together (grouped launches).

    ``add`` once per structure (geometry, pair list, weights and plans stay on the device),
    then ``shift`` (host arrays) or ``shift_device`` (device pointers) for any subset of them
    once per SCC iteration.  ``capacity_bytes`` caps the device memory of the structures; a
    structure that does not fit is refused with EW_ERR_MEMORY.
    """

    def __init__(self, capacity_bytes: int = 0, **opts):
        self._p = C.c_void_p()
        self.options = make_options(**opts)
        _check(lib().ew_pool_create(C.byref(self.options), int(capacity_bytes),
                                    C.byref(self._p)))
        self.sizes = {}

    def _check(self, status):
        if status != EW_OK:
            raise EwaldError(status, lib().ew_pool_last_error(self._p).decode())

    def close(self):
        if self._p:
            lib().ew_pool_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, coords, lattice, species=None, hubbard_u=None, charges=None, grid=None) -> int:
        coords = _f64(coords, 2)
        lattice = _f64(lattice, 2)
        n = coords.shape[0]
        if coords.shape != (n, 3) or lattice.shape != (3, 3):
            raise ValueError("coords must be (n, 3) and lattice (3, 3)")
        sp = None if species is None else np.ascontiguousarray(species, dtype=np.int32)
        u = None if hubbard_u is None else _f64(hubbard_u, 1)
        q = None if charges is None else _f64(charges, 1)
        g = None if grid is None else np.ascontiguousarray(grid, dtype=np.int32)
        sid = C.c_int()
        self._check(lib().ew_pool_add(self._p, n, coords, lattice, sp, u,
                                      0 if u is None else len(u), q, g, C.byref(sid)))
        self.sizes[sid.value] = n
        return sid.value

    def remove(self, sid: int):
        self._check(lib().ew_pool_remove(self._p, int(sid)))
        self.sizes.pop(sid, None)

    def parameters(self, sid: int) -> dict:
        p = Parameters()
        self._check(lib().ew_pool_parameters(self._p, int(sid), C.byref(p)))
        return p.as_dict()

    def usage(self):
        """(device bytes held, capacity, structure count)."""
        u, c, k = C.c_longlong(), C.c_longlong(), C.c_int()
        self._check(lib().ew_pool_usage(self._p, C.byref(u), C.byref(c), C.byref(k)))
        return u.value, c.value, k.value

    def shift(self, ids, dqs):
        """One grouped evaluation; dqs: one charge vector per id.  Returns (shifts, energies)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        if len(dqs) != len(ids):
            raise ValueError(f"{len(ids)} ids but {len(dqs)} charge vectors")
        vecs = [_f64(d, 1) for d in dqs]
        # the library walks the concatenation by each structure's atom count (an unknown id is
        # refused there before anything is read)
        for i, d in zip(ids, vecs):
            n = self.sizes.get(int(i))
            if n is not None and len(d) != n:
                raise ValueError(f"structure {int(i)} has {n} atoms, got {len(d)} charges")
        dq = np.concatenate(vecs) if vecs else np.zeros(0)
        out = np.empty_like(dq)
        en = np.empty(len(ids))
        self._check(lib().ew_pool_shift(self._p, len(ids), ids, dq, out, en))
        shifts, o = [], 0
        for i in ids:
            n = self.sizes[int(i)]
            shifts.append(out[o:o + n])
            o += n
        return shifts, en

    def shift_device(self, ids, dq_ptrs, shift_ptrs, energies_ptr: int, stream: int = 0):
        """Enqueues one grouped evaluation on device buffers (no synchronisation)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        k = len(ids)
        dp = (C.c_void_p * k)(*[C.c_void_p(int(x)) for x in dq_ptrs])
        sp = (C.c_void_p * k)(*[C.c_void_p(int(x)) for x in shift_ptrs])
        self._check(lib().ew_pool_shift_device(self._p, k, ids, dp, sp,
                                               C.c_void_p(int(energies_ptr)),
                                               C.c_void_p(stream or None)))

    def sync(self):
        self._check(lib().ew_pool_sync(self._p))


class EwaldServer:
    """Request loop with continuous batching over an EwaldPool.

    ``submit(id, dq)`` queues one SCC evaluation of a resident structure and returns a ticket;
    ``result(ticket)`` waits for it and returns (shift, energy).  While the server runs, only the
    server uses the pool.
    """

    def __init__(self, pool: "EwaldPool", max_batch: int = 64, max_inflight: int = 2):
        self._s = C.c_void_p()
        self.pool = pool
        o = ServerOptions()
        o.max_batch = max_batch
        o.max_inflight = max_inflight
        _check(lib().ew_server_create(pool._p, C.byref(o), C.byref(self._s)))
        self._out = {}  # ticket -> (shift array, energy array), alive until completion

    def _check(self, status):
        if status != EW_OK:
            raise EwaldError(status, lib().ew_server_last_error(self._s).decode())

    def close(self):
        if self._s:
            lib().ew_server_destroy(self._s)
            self._s = C.c_void_p()
            self._out.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, sid: int, dq) -> int:
        dq = _f64(dq, 1)
        n = self.pool.sizes.get(int(sid))
        if n is not None and len(dq) != n:
            raise ValueError(f"structure {sid} has {n} atoms, got {len(dq)} charges")
        # an unknown id is refused by the library before anything is written
        shift = np.empty(n if n is not None else 1)
        energy = np.empty(1)
        t = C.c_longlong()
        self._check(lib().ew_server_submit(self._s, int(sid), dq.ctypes.data, shift.ctypes.data,
                                           energy.ctypes.data, C.byref(t)))
        self._out[t.value] = (shift, energy)
        return t.value

    def submit_ptr(self, sid: int, dq_ptr: int, shift_ptr: int, energy_ptr: int) -> int:
        """submit() on raw host addresses (the caller keeps the buffers alive and sized: dq is
        copied at submission, shift and energy are written at completion); for load drivers."""
        t = C.c_longlong()
        self._check(lib().ew_server_submit(self._s, sid, dq_ptr, shift_ptr, energy_ptr,
                                           C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0) -> int:
        """Waits for a ticket from submit_ptr; returns its status (EW_ERR_PENDING at timeout)."""
        return lib().ew_server_wait(self._s, ticket, timeout)

    def result(self, ticket: int, timeout: float = -1.0):
        """Waits for a ticket; returns (shift, energy), or None at the timeout."""
        st = lib().ew_server_wait(self._s, int(ticket), float(timeout))
        if st == EW_ERR_PENDING:
            return None
        shift, energy = self._out.pop(ticket)
        self._check(st)
        return shift, float(energy[0])

    def drain(self):
        self._check(lib().ew_server_drain(self._s))

    def stats(self, reset: bool = False) -> dict:
        r = ServerReport()
        self._check(lib().ew_server_stats(self._s, C.byref(r), 1 if reset else 0))
        return r.as_dict()


def ewald_potential_force(positions, charges, lattice, force=False, **opts):
    """ewald_potential_force (fe/fft/ewald_fft_3d.f90:36-134): (pot, force or None)."""
    positions = _f64(positions, 2)
    charges = _f64(charges, 1)
    lattice = _f64(lattice, 2)
    n = len(charges)
    pot = np.empty(n)
    f = np.empty((n, 3)) if force else None
    o = make_options(**opts)
    _check(lib().ew_potential_force(n, positions, charges, lattice, C.byref(o), pot, f))
    return pot, f


def ewald_parameters(charges, lattice, **opts) -> dict:
    """ewald_parameters (fe/fft/ewald_fft_3d.f90:142-191), no device needed."""
    charges = _f64(charges, 1)
    lattice = _f64(lattice, 2)
    p = Parameters()
    o = make_options(**opts)
    _check(lib().ew_parameters_for(len(charges), charges, lattice, C.byref(o), C.byref(p)))
    return p.as_dict()


def ewald_grid_memory(charges, lattice, **opts) -> float:
    """ewald_grid_memory (fe/fft/ewald_fft_3d.f90:197-244) in gigabytes, no device needed."""
    charges = _f64(charges, 1)
    lattice = _f64(lattice, 2)
    g = C.c_double()
    o = make_options(**opts)
    _check(lib().ew_grid_memory(len(charges), charges, lattice, C.byref(o), C.byref(g)))
    return g.value


class EwaldMixer:
    """SCC charge mixer on the device: DFTB+'s modified Broyden (default) or simple mixer.

    ``reset(n)`` starts an SCC cycle with vectors of n entries; ``mix(q_inp, q_diff)`` returns
    the mixed charges (host arrays); ``mix_device(q_ptr, q_diff_ptr)`` mixes device arrays in
    place (q_inp -> mixed charges) without leaving the device.
    """

    def __init__(self, kind: str = "broyden", mix_param: float = 0.2, max_iter: int = 100,
                 omega0: float = 0.01, min_weight: float = 1.0, max_weight: float = 1.0e5,
                 weight_fac: float = 1.0e-2, device: int = 0):
        kinds = {"broyden": MIXER_BROYDEN, "simple": MIXER_SIMPLE}
        if kind not in kinds:
            raise ValueError(f"kind must be one of {sorted(kinds)}")
        o = MixerOptions()
        o.kind = kinds[kind]
        o.max_iter = max_iter
        o.mix_param = mix_param
        o.omega0 = omega0
        o.min_weight = min_weight
        o.max_weight = max_weight
        o.weight_fac = weight_fac
        o.device = device
        self._m = C.c_void_p()
        _check(lib().ew_mixer_create(C.byref(o), C.byref(self._m)))
        self.n = 0

    def _check(self, status):
        if status != EW_OK:
            raise EwaldError(status, lib().ew_mixer_last_error(self._m).decode())

    def close(self):
        if self._m:
            lib().ew_mixer_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self, n_elem: int):
        self._check(lib().ew_mixer_reset(self._m, int(n_elem)))
        self.n = int(n_elem)

    def mix(self, q_inp, q_diff) -> np.ndarray:
        q = np.array(q_inp, dtype=np.float64, copy=True)
        d = _f64(q_diff, 1)
        if self.n and (q.ndim != 1 or q.shape != d.shape or len(q) != self.n):
            raise ValueError(f"expected two vectors of {self.n} entries")
        self._check(lib().ew_mixer_mix(self._m, q, d))
        return q

    def mix_device(self, q_ptr: int, q_diff_ptr: int, stream: int = 0):
        self._check(lib().ew_mixer_mix_device(self._m, C.c_void_p(q_ptr), C.c_void_p(q_diff_ptr),
                                              C.c_void_p(stream or None)))

    @property
    def iteration(self) -> int:
        it = C.c_int()
        self._check(lib().ew_mixer_iteration(self._m, C.byref(it)))
        return it.value
