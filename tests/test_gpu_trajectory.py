"""One engine driven through a trajectory: moved atoms, the short-range gamma dropped, a sheared
cell, a smaller and a larger system and back, with the same caller device buffers on every step.

A new geometry releases the previous one's buffers (reused when the sizes fit) and the captured
graphs, which are keyed by the caller's buffer addresses; every step must give what a fresh engine
gives at that geometry, through the host, device-pointer and multi-vector entry points and for
the gradients and the stress.

Tolerances (max_i |x_i - ref_i| / max_i |ref_i|): FP64 1e-12, mixed 1e-6 (shifts) and 1e-5
(energies, gradients, stress): the fresh engine may take other launch shapes (its buffers and
captured graphs are new), whose sums round differently.
"""
import sys

import numpy as np
import pytest
from conftest import rel_max

import fftewald_amd as ew
from fftewald_amd import workloads as W

pytestmark = pytest.mark.gpu

NVEC = 3


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _steps():
    rng = np.random.default_rng(31)
    a = W.water(n_side=8, box=8 * 3.1)
    moved = dict(a, coords=a["coords"] + rng.normal(scale=0.05, size=a["coords"].shape))
    sheared = dict(moved, lattice=moved["lattice"] + np.array([[0, 0, 0], [0.2, 0, 0], [0, 0, 0]])
                   * moved["lattice"][0, 0])
    small = W.water(seed=5, n_side=6, box=6 * 3.1)
    large = W.water(seed=6, n_side=10, box=10 * 3.1)
    bare = {k: v for k, v in moved.items() if k not in ("species", "hubbard_u")}
    return [("water 1536", a), ("same atoms moved", moved), ("no short-range gamma", bare),
            ("sheared cell", sheared), ("water 648 (smaller)", small),
            ("water 3000 (larger)", large), ("first again", a)]


def _set(e, w):
    e.set_geometry(w["coords"], w["lattice"], species=w.get("species"),
                   hubbard_u=w.get("hubbard_u"), charges=w["charges"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_trajectory_matches_fresh_engines(precision):
    import torch
    tol, tol_e = (1e-12, 1e-12) if precision == "fp64" else (1e-6, 1e-5)
    opts = dict(tol=1e-8, precision=precision)
    steps = _steps()
    nmax = max(len(w["charges"]) for _, w in steps)
    dev = torch.device("cuda:0")
    dq = torch.zeros(NVEC * nmax, dtype=torch.float64, device=dev)
    out = torch.zeros_like(dq)
    en = torch.zeros(NVEC, dtype=torch.float64, device=dev)
    rng = np.random.default_rng(17)
    first = None
    with ew.EwaldEngine(**opts) as e:
        for name, w in steps:
            say(f"trajectory step: {name}, {len(w['charges'])} atoms, {precision}")
            _set(e, w)
            n = len(w["charges"])
            qs = np.stack([w["charges"]] + [W.scc_update(w["charges"], rng)
                                            for _ in range(NVEC - 1)])
            with ew.EwaldEngine(**opts) as f:
                _set(f, w)
                assert f.parameters()["grid"] == e.parameters()["grid"]
                want = [f.shift(q) for q in qs]
                want_g = f.gradients(qs[1])
                want_st = f.stress(qs[1])
            # host arrays
            s, en_h = e.shift(qs[0])
            assert rel_max(s, want[0][0]) < tol
            assert en_h == pytest.approx(want[0][1], rel=tol_e)
            # device pointers: the same buffers on every step
            dq[:n] = torch.from_numpy(qs[1]).to(dev)
            torch.cuda.synchronize()
            e.shift_device(dq.data_ptr(), out.data_ptr(), en.data_ptr())
            torch.cuda.synchronize()
            assert rel_max(out[:n].cpu().numpy(), want[1][0]) < tol
            assert float(en[0].item()) == pytest.approx(want[1][1], rel=tol_e)
            # several vectors per call, again the same buffers
            dq[:NVEC * n] = torch.from_numpy(qs.reshape(-1)).to(dev)
            torch.cuda.synchronize()
            e.shift_multi_device(NVEC, dq.data_ptr(), out.data_ptr(), en.data_ptr())
            torch.cuda.synchronize()
            got = out[:NVEC * n].cpu().numpy().reshape(NVEC, n)
            for v in range(NVEC):
                assert rel_max(got[v], want[v][0]) < tol
                assert float(en[v].item()) == pytest.approx(want[v][1], rel=tol_e)
            assert rel_max(e.gradients(qs[1]), want_g) < tol_e
            st = e.stress(qs[1])
            assert np.abs(st - want_st).max() / np.abs(want_st).max() < tol_e
            if first is None:
                first = s
            elif name == "first again":
                assert rel_max(s, first) < tol


def test_caller_stream_evaluations_are_waited_for():
    """Evaluations enqueued on a caller's (non-default) stream and not synchronised: a new
    geometry, the gradients and a second evaluation on another stream must wait for them, since
    they reuse the same device buffers.  Every result equals a fresh engine's."""
    import torch
    dev = torch.device("cuda:0")
    a = W.water(n_side=8, box=8 * 3.1)
    b = W.water(seed=5, n_side=6, box=6 * 3.1)
    with ew.EwaldEngine(tol=1e-8) as f:
        _set(f, a)
        want_a = f.shift(a["charges"])
        want_ga = f.gradients(a["charges"])
        _set(f, b)
        want_b = f.shift(b["charges"])
    na, nb = len(a["charges"]), len(b["charges"])
    dq_a = torch.from_numpy(a["charges"]).to(dev)
    dq_b = torch.from_numpy(b["charges"]).to(dev)
    out_a = torch.zeros(na, dtype=torch.float64, device=dev)
    out_a2 = torch.zeros(na, dtype=torch.float64, device=dev)
    out_b = torch.zeros(nb, dtype=torch.float64, device=dev)
    en = torch.zeros(3, dtype=torch.float64, device=dev)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    say("caller streams: evaluation, gradients, second stream, new geometry, no sync between")
    with ew.EwaldEngine(tol=1e-8) as e:
        _set(e, a)
        e.shift_device(dq_a.data_ptr(), out_a.data_ptr(), en.data_ptr(), stream=s1.cuda_stream)
        g = e.gradients(a["charges"])  # host entry point on the handle's own stream
        e.shift_device(dq_a.data_ptr(), out_a2.data_ptr(), en.data_ptr() + 8,
                       stream=s2.cuda_stream)
        _set(e, b)  # releases a's buffers: must wait for both evaluations
        e.shift_device(dq_b.data_ptr(), out_b.data_ptr(), en.data_ptr() + 16,
                       stream=s1.cuda_stream)
        torch.cuda.synchronize()
    assert rel_max(out_a.cpu().numpy(), want_a[0]) < 1e-12
    assert rel_max(out_a2.cpu().numpy(), want_a[0]) < 1e-12
    assert rel_max(out_b.cpu().numpy(), want_b[0]) < 1e-12
    assert float(en[0]) == pytest.approx(want_a[1], rel=1e-12)
    assert float(en[1]) == pytest.approx(want_a[1], rel=1e-12)
    assert float(en[2]) == pytest.approx(want_b[1], rel=1e-12)
    assert rel_max(g, want_ga) < 1e-12


@pytest.mark.parametrize("multi", [False, True])
def test_more_buffer_sets_than_captured_graphs(multi):
    """Twelve distinct buffer sets (graphs are kept for eight) launched back to back on a
    caller's stream with no synchronisation: evicting a captured graph must wait for its
    launches.  Every result is checked."""
    import torch
    dev = torch.device("cuda:0")
    w = W.water(n_side=6, box=6 * 3.1)
    n = len(w["charges"])
    rng = np.random.default_rng(5)
    nv = 2 if multi else 1
    qs = [np.stack([W.scc_update(w["charges"], rng) for _ in range(nv)]) for _ in range(12)]
    with ew.EwaldEngine(tol=1e-8) as f:
        _set(f, w)
        want = [[f.shift(q[v]) for v in range(nv)] for q in qs]
    dqs = [torch.from_numpy(q.reshape(-1)).to(dev) for q in qs]
    outs = [torch.zeros(nv * n, dtype=torch.float64, device=dev) for _ in qs]
    ens = [torch.zeros(nv, dtype=torch.float64, device=dev) for _ in qs]
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    say(f"12 buffer sets on a caller stream, multi={multi}")
    with ew.EwaldEngine(tol=1e-8) as e:
        _set(e, w)
        for rep in range(2):
            for k in range(12):
                if multi:
                    e.shift_multi_device(nv, dqs[k].data_ptr(), outs[k].data_ptr(),
                                         ens[k].data_ptr(), stream=s.cuda_stream)
                else:
                    e.shift_device(dqs[k].data_ptr(), outs[k].data_ptr(), ens[k].data_ptr(),
                                   stream=s.cuda_stream)
            torch.cuda.synchronize()
            for k in range(12):
                got = outs[k].cpu().numpy().reshape(nv, n)
                for v in range(nv):
                    assert rel_max(got[v], want[k][v][0]) < 1e-12, (rep, k, v)
                    assert float(ens[k][v]) == pytest.approx(want[k][v][1], rel=1e-12)
