"""The allocation-free evaluation (ew_shift_device_noalloc) through the C ABI: its scratch lives
in the library's static device array, so it must give what ew_shift_device gives, call after
call, also when a call follows one on a larger structure (the array keeps the larger call's
contents), and refuse a geometry whose scratch exceeds the array.

Both entry points run the same kernels, and every sum of an evaluation (the spreading's
fixed-point accumulation included) has a fixed order or is exact, so the results agree bit for
bit.
"""
import ctypes as C
import sys

import numpy as np
import pytest
from conftest import rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _both(torch, e, q):
    """(shift, energy) of ew_shift_device and of ew_shift_device_noalloc for charges q."""
    dev = torch.device("cuda:0")
    n = len(q)
    dq = torch.from_numpy(q).to(dev)
    a = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    b = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    en = torch.full((2,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    e.shift_device(dq.data_ptr(), a.data_ptr(), en.data_ptr())
    st = ew.lib().ew_shift_device_noalloc(e._h, C.c_void_p(dq.data_ptr()),
                                          C.c_void_p(b.data_ptr()),
                                          C.c_void_p(en.data_ptr() + 8), None)
    assert st == ew.EW_OK, ew.lib().ew_last_error(e._h).decode()
    torch.cuda.synchronize()
    return (a.cpu().numpy(), float(en[0])), (b.cpu().numpy(), float(en[1]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_shift_device_large_then_smaller(precision):
    import torch
    # the 98,304-atom cell at 128^3 needs about 44 MB of scratch in mixed precision; FP64 takes
    # a 24,000-atom cell of the same density at 80^3
    big = W.triclinic(n=24000, side=62.1) if precision == "fp64" else W.triclinic()
    kbig = 80 if precision == "fp64" else 128
    small = W.water(n_side=8, box=8 * 3.1)
    rng = np.random.default_rng(3)
    for name, w, K in (("large", big, kbig), ("smaller", small, 32)):
        n = len(w["charges"])
        say(f"noalloc {precision}: {name}, {n} atoms, {K}^3")
        with ew.EwaldEngine(tol=1e-10, precision=precision, grid=K, order=6,
                            r_cut=10.0 * W.AA) as e:
            e.set_geometry(w["coords"], w["lattice"], species=w.get("species"),
                           hubbard_u=w.get("hubbard_u"), charges=w["charges"])
            for q in (w["charges"], W.scc_update(w["charges"], rng)):
                (sa, ea), (sb, eb) = _both(torch, e, q)
                assert np.isfinite(sb).all()
                assert np.array_equal(sb, sa)
                assert eb == ea


def test_noalloc_refuses_scratch_beyond_the_static_array():
    """The 98,304-atom cell on a 192^3 mesh needs over 110 MB of scratch in FP64 (mesh and
    spectrum), above 64 MiB: refused on the host before anything is launched; the handle still
    evaluates normally."""
    import torch
    w = W.triclinic()
    dev = torch.device("cuda:0")
    n = len(w["charges"])
    dq = torch.from_numpy(w["charges"]).to(dev)
    out = torch.zeros(n, dtype=torch.float64, device=dev)
    say("noalloc fp64 98,304 atoms at 192^3: expect EW_ERR_MEMORY")
    with ew.EwaldEngine(tol=1e-10, grid=192, order=6, r_cut=w["r_cut"]) as e:
        e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
        st = ew.lib().ew_shift_device_noalloc(e._h, C.c_void_p(dq.data_ptr()),
                                              C.c_void_p(out.data_ptr()), None, None)
        assert st == ew.EW_ERR_MEMORY
        assert "static limit" in ew.lib().ew_last_error(e._h).decode()
        e.shift_device(dq.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        assert np.isfinite(out.cpu().numpy()).all()
