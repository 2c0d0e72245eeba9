"""Launch shapes (ew_tuning): every shape the autotuner may pick gives the same shifts and
energy as the fixed rules, to round-off (the shapes sum in different orders), and the
CPU SPME at identical parameters; the autotuner's choice is cached and reused.

Tolerances (max_i |x_i - ref_i| / max_i |ref_i|): FP64 1e-13 between shapes and 1e-11 against
the CPU SPME; mixed 1e-6 against the FP64 rules.
"""
import os
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


def _cases():
    w = W.water(n_side=8, box=8 * 3.1)  # 1536 atoms with short gamma, 32^3
    w["grid"] = (32, 32, 32)
    t = W.triclinic(n=3000, side=31.0)  # sheared, 45 x 27 x 50 mesh (odd radices)
    t["grid"] = (45, 27, 50)
    return [("water 1536, 32^3", w), ("triclinic 3000, 45x27x50", t)]


def _shapes(base):
    """A spread of shapes around the rules: write-out (copies, direct, owned blocks), blocks,
    register line transforms with each compiled lane count, the Stockham passes' planes, lines
    and radix orders, and the pair sum's placement (co-scheduling runs in Stockham launches)."""
    out = []
    for direct in (0, 1, 2):
        for blk in ((1, 1, 1), (2, 1, 1), (1, 2, 2), (2, 2, 2)):
            t = dict(base, spread_direct=direct, spread_block=blk)
            out.append(t)
    for lanes in ((4, 4, 4), (8, 8, 8), (4, 8, 16), (16, 16, 16)):
        out.append(dict(base, fft_wave=1, fft_lanes=lanes))
    for lanes in ((4, 4, 4), (4, 4, 8), (8, 8, 8)):
        out.append(dict(base, fft_wave=2, fft_lanes=lanes))
    for planes in (0, 1):
        for lines in ((8, 2, 4), (16, 8, 16), (32, 16, 2)):
            out.append(dict(base, fft_wave=0, fft_planes=planes, fft_lines=lines))
    for r in range(5):
        out.append(dict(base, fft_wave=0, fft_radix=(r, (r + 1) % 5, (r + 2) % 5)))
    for co, share in ((1, 0), (1, 35), (1, 100), (0, 0)):
        out.append(dict(base, fft_wave=0, pair_cosched=co, pair_spread_share=share))
    for nb in (64, 256, 1024):
        out.append(dict(base, pair_blocks=nb))
        out.append(dict(base, pair_blocks=nb, pair_fork=1))
    return out


@pytest.mark.parametrize("case", range(2))
def test_every_shape_matches_the_rules_and_the_oracle(case, oracle_mod):
    name, w = _cases()[case]
    opts = dict(tol=1e-10, grid=w["grid"], order=6, r_cut=8.0 * W.AA)
    geo = dict(species=w.get("species"), hubbard_u=w.get("hubbard_u"), charges=w["charges"])
    with ew.EwaldEngine(**opts) as e:
        e.set_geometry(w["coords"], w["lattice"], **geo)
        base = e.tuning()
        prm = e.parameters()
        ref, eref = e.shift(w["charges"])
    want = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                 prm["order"], prm["alpha"], prm["r_cut"])
    if w.get("species") is not None:
        want = want + oracle_mod.short_gamma_shift(w["coords"], w["charges"], w["lattice"],
                                                   w["species"], w["hubbard_u"])
    assert rel_max(ref, want) < 1e-11
    tried = 0
    for t in _shapes(base):
        say(f"shapes {name}: {t}")
        with ew.EwaldEngine(**opts, tuning=t) as e:
            try:
                e.set_geometry(w["coords"], w["lattice"], **geo)
            except ew.EwaldError as err:
                # a block larger than this mesh is refused, never run
                assert err.code == ew.EW_ERR_ARGUMENT and "spreading block" in str(err)
                continue
            got = e.tuning()
            s, en = e.shift(w["charges"])
            # the multi-vector grouped launch over the same shapes (its spreading blocks may
            # take more than 64 KiB of LDS, as the single-vector chain's may)
            sm, em = e.shift_multi(np.stack([w["charges"], 0.5 * w["charges"]]))
        assert rel_max(s, ref) < 1e-13, t
        assert rel_max(sm[0], ref) < 1e-13, t
        assert rel_max(sm[1], 0.5 * ref) < 1e-13, t
        assert em[0] == pytest.approx(eref, rel=1e-13)
        assert em[1] == pytest.approx(0.25 * eref, rel=1e-13)
        assert en == pytest.approx(eref, rel=1e-13)
        tried += 1
        assert got["fft_radix"] == tuple(t["fft_radix"])
    assert tried >= 12


def test_autotuned_shapes_agree_and_are_cached(tmp_path, monkeypatch):
    monkeypatch.setenv("EW_TUNE_CACHE", str(tmp_path / "tune.cache"))
    w = W.water(n_side=8, box=8 * 3.1)
    opts = dict(tol=1e-10, grid=(32, 32, 32), order=6, r_cut=8.0 * W.AA)
    geo = dict(species=w["species"], hubbard_u=w["hubbard_u"], charges=w["charges"])
    for prec, tol in (("fp64", 1e-13), ("mixed", 1e-6)):
        say(f"autotune water 1536 {prec}")
        with ew.EwaldEngine(**opts, precision="fp64") as e:
            e.set_geometry(w["coords"], w["lattice"], **geo)
            ref, eref = e.shift(w["charges"])
        with ew.EwaldEngine(**opts, precision=prec, tune="auto") as e:
            e.set_geometry(w["coords"], w["lattice"], **geo)
            tuned = e.tuning()
            s, en = e.shift(w["charges"])
        assert tuned["eval_us"] > 0.0
        assert rel_max(s, ref) < tol
        assert en == pytest.approx(eref, rel=10 * tol)
        # the choice is in the cache: a cache-only engine takes it without timing
        with ew.EwaldEngine(**opts, precision=prec, tune="cached") as e:
            e.set_geometry(w["coords"], w["lattice"], **geo)
            again = e.tuning()
        for k in ("spread_direct", "spread_block", "fft_planes", "fft_lines", "fft_radix",
                  "pair_cosched", "pair_spread_share", "fft_wave", "fft_lanes", "pair_blocks",
                  "pair_fork"):
            assert again[k] == tuned[k], k
    lines = open(tmp_path / "tune.cache").read().splitlines()
    assert len(lines) == 2 and all(l.startswith("v5:") for l in lines)
    assert all(len(l.split()) == 21 for l in lines)


# every compiled (axis length, lanes per line) of the register line transforms appears on
# some axis of these meshes
_WAVE_MESHES = [(32, 48, 64), (64, 32, 96), (96, 64, 32), (128, 96, 48), (48, 128, 128),
                (256, 32, 32), (32, 256, 64), (64, 64, 256), (32, 32, 48), (96, 96, 64),
                (128, 128, 96)]
_LANES = {32: (4, 8), 48: (4,), 64: (4, 8, 16), 96: (4, 8), 128: (4, 8, 16), 256: (8, 16)}
# square x-y planes with a compiled plane kernel
_PLANE_LANES = {32: (4,), 64: (8,), 96: (4,), 128: (8,)}


@pytest.mark.parametrize("mesh", _WAVE_MESHES)
def test_register_transforms_match_stockham(mesh, oracle_mod):
    """The register line transforms give the Stockham passes' shifts and energy (FP64 1e-13,
    mixed 2e-6 against mixed Stockham and 1e-6 against FP64) for every compiled lane count of
    every axis, in a sheared cell (gamma 75 deg); one FP64 case also against the CPU SPME
    (1e-11)."""
    n = 1500
    w = W.triclinic(n=n, side=26.0)
    opts = dict(tol=1e-10, grid=mesh, order=6, r_cut=7.0 * W.AA)
    geo = dict(charges=w["charges"])
    ref = {}
    for prec in ("fp64", "mixed"):
        with ew.EwaldEngine(**opts, precision=prec) as e:
            e.set_geometry(w["coords"], w["lattice"], **geo)
            base = e.tuning()
            stock = dict(base, fft_wave=0)
        with ew.EwaldEngine(**opts, precision=prec, tuning=stock) as e:
            e.set_geometry(w["coords"], w["lattice"], **geo)
            assert e.tuning()["fft_wave"] == 0
            ref[prec] = e.shift(w["charges"])
            prm = e.parameters()
        if prec == "fp64" and mesh == _WAVE_MESHES[0]:
            want = oracle_mod.spme_ewald(w["coords"], w["charges"], w["lattice"], prm["grid"],
                                         prm["order"], prm["alpha"], prm["r_cut"])
            assert rel_max(ref[prec][0], want) < 1e-11
        combos = max(len(_LANES[k]) for k in mesh)
        cases = [(1, tuple(_LANES[k][c % len(_LANES[k])] for k in mesh)) for c in range(combos)]
        if mesh[0] == mesh[1]:
            cases += [(2, (pl, pl, _LANES[mesh[2]][0])) for pl in _PLANE_LANES[mesh[0]]]
        for wave, lanes in cases:
            say(f"register transforms {mesh} wave {wave} lanes {lanes} {prec}")
            with ew.EwaldEngine(**opts, precision=prec,
                                tuning=dict(base, fft_wave=wave, fft_lanes=lanes)) as e:
                e.set_geometry(w["coords"], w["lattice"], **geo)
                got = e.tuning()
                assert got["fft_wave"] == wave and got["fft_lanes"] == lanes
                s, en = e.shift(w["charges"])
            r, er = ref[prec]
            tol = 1e-13 if prec == "fp64" else 2e-6
            assert rel_max(s, r) < tol, (mesh, lanes, prec)
            assert en == pytest.approx(er, rel=10 * tol)
            if prec == "mixed":
                assert rel_max(s, ref["fp64"][0]) < 1e-6


def test_owned_blocks_over_several_item_passes():
    """Owned spreading blocks whose atoms overflow one pass of the LDS item list (a dense cell
    and blocks of 2 x 2 x 2 sort tiles: over 2,000 candidate atoms per block, 1,024 per pass)
    give the block copies' shifts and energy, in FP64 and, in mixed precision, the FP64
    result to 1e-6; the same shapes in the multi-vector grouped launch."""
    w = W.triclinic(n=20000, side=55.0)
    opts = dict(tol=1e-10, grid=(64, 64, 64), order=6, r_cut=8.0 * W.AA)
    geo = dict(charges=w["charges"])
    with ew.EwaldEngine(**opts) as e:
        e.set_geometry(w["coords"], w["lattice"], **geo)
        base = e.tuning()
    ref = None
    for prec, tol in (("fp64", 1e-13), ("mixed", 1e-6)):
        for direct in (0, 2):
            t = dict(base, spread_direct=direct, spread_block=(2, 2, 2))
            say(f"owned passes {prec} direct {direct}")
            with ew.EwaldEngine(**opts, precision=prec, tuning=t) as e:
                e.set_geometry(w["coords"], w["lattice"], **geo)
                assert e.tuning()["spread_direct"] == direct
                s, en = e.shift(w["charges"])
                sm, em = e.shift_multi(np.stack([w["charges"], -w["charges"]]))
            if ref is None:
                ref, eref = s, en
            assert rel_max(s, ref) < tol
            assert en == pytest.approx(eref, rel=10 * tol)
            assert rel_max(sm[0], ref) < tol
            assert rel_max(sm[1], -ref) < tol
            assert em[1] == pytest.approx(eref, rel=10 * tol)
