"""Benchmark: SCC shift+energy evaluations per second of the SPME Ewald engine on one MI355X.

A step is one SCC-iteration evaluation (charges -> per-atom shift + energy) of the 98,304-atom
triclinic cell (a = b = c = 99.4 A, gamma = 75 deg, Poisson-disc positions, charges U[-1,1]),
128^3 mesh, B-spline order 6, r_cut 10 A, geometry resident on the device, charges already in
HBM.  The headline is the mixed-precision engine; the FP64 engine is timed in the same run and
the mixed shifts are checked against it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Per-kernel durations come from HIP events recorded by the engine on its
own stream (profile mode, a separate pass); the CPU baseline is the oracle's restatement of the
reference fast path (linked cells + Gaussian NFFT at the DFTB+ default tol 1e-9) on the host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "dftb_fast-ewald_3d_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def phase_bytes(n, K, order, nnz, rs, cs, ib, gamma, planes=False, inf_table=False):
    """Algorithmic (compulsory) bytes of one launch of each phase.

    rs: mesh real size, cs: pair coefficient size, ib: partner index size (2 or 4 bytes);
    planes: the x and y passes are one fused kernel per direction (the y phases move nothing);
    inf_table: the z pass reads the influence function from a per-geometry table (one real per
    half-spectrum point) instead of evaluating it."""
    G = K[0] * K[1] * K[2]
    Gh = (K[0] // 2 + 1) * K[1] * K[2]
    th = 3 * order * rs
    return {
        "gather_charges": 20 * n,
        "spread": (16 + th + 8) * n + rs * G,
        "fft_x_r2c": rs * G + 2 * rs * Gh,
        "fft_y": 0 if planes else 4 * rs * Gh,
        "fft_z_influence": (5 if inf_table else 4) * rs * Gh,
        "fft_y_inverse": 0 if planes else 4 * rs * Gh,
        "fft_x_c2r": 2 * rs * Gh + rs * G,
        "pair_sum": nnz * (ib + cs) + n * (8 + 8 + 8),
        "gather_combine": (16 + th + 8 + 8 + 4 + 8 + (8 if gamma else 0)) * n + rs * G,
    }


def trace_bytes(n, K, order, nnz, rs, cs, ib, gamma, planes=False, region=True, tuning=None,
                split_gather=False, inf_table=False):
    """Algorithmic bytes of each span of ew_trace_times (the evaluation's launch shape).

    The privatised spreading's region kernel reads the atoms' stencil origins, weights and
    charges (and the sort permutation when it reads the caller's charges); its assembly writes
    the mesh (the block copies between them are not algorithmic; with direct write-out the
    region kernel writes the mesh).  Without privatised blocks the whole spreading is the
    assembly span.  With the pair sum co-scheduled in the chain's launches its bytes go to
    those spans in the proportions of its row split."""
    pb = phase_bytes(n, K, order, nnz, rs, cs, ib, gamma, planes, inf_table)
    G = K[0] * K[1] * K[2]
    tuning = tuning or {}
    co = bool(tuning.get("pair_cosched", 0))
    direct = bool(tuning.get("spread_direct", 0))
    atoms = pb["spread"] - rs * G + (4 * n if region and not co else 0)
    out = {"charges": pb["gather_charges"], "pair_sum": pb["pair_sum"],
           "spread_region": atoms if region else 0,
           "spread_assemble": rs * G if region else pb["spread"],
           "fft_x_r2c": pb["fft_x_r2c"], "fft_y": pb["fft_y"],
           "fft_z_influence": pb["fft_z_influence"], "fft_y_inverse": pb["fft_y_inverse"],
           "fft_x_c2r": pb["fft_x_c2r"], "gather_mesh": 0, "join_wait": 0,
           "gather_combine": pb["gather_combine"]}
    if split_gather:
        # the mesh potential at every atom (stencil origins, weights, the mesh; one FP64 out)
        # before the join; after it the combine reads the mesh potential, pair sum, charge,
        # on-site term and caller index and writes the shift
        th = 3 * order * rs
        out["gather_mesh"] = (16 + th + 8) * n + rs * G
        out["gather_combine"] = (8 + 8 + 8 + 8 + 4 + 8) * n
    if region and direct:
        out["spread_region"] += rs * G
        out["spread_assemble"] = 0
    if co:
        share = tuning.get("pair_spread_share", 0) / 100.0 if region else 0.0
        ffts = ["fft_x_r2c", "fft_z_influence", "fft_x_c2r"] if planes else \
            ["fft_x_r2c", "fft_y", "fft_z_influence", "fft_y_inverse", "fft_x_c2r"]
        out["spread_region"] += share * pb["pair_sum"]
        for k in ffts:
            out[k] += (1.0 - share) * pb["pair_sum"] / len(ffts)
        out["pair_sum"] = 0
    return out


#: spans of ew_trace_times on the evaluation stream (the chain), in launch order
CHAIN = ("spread_region", "spread_assemble", "fft_x_r2c", "fft_y", "fft_z_influence",
         "fft_y_inverse", "fft_x_c2r", "gather_mesh", "join_wait", "gather_combine")


def trace_symbol(span, prec, index_bytes, order, planes=False, pair_mode=0, tuning=None,
                 grid=None, split_gather=False):
    """The kernel that carries a span of ew_trace_times, as rocprofv3 names it (template
    arguments included: the Stockham chain kernels carry the pair sum's co-scheduling flag, the
    register transforms their DFT length per lane and lanes per line)."""
    r = "float" if prec == "mixed" else "double"
    i = "short" if index_bytes == 2 else "int"
    tuning = tuning or {}
    wave = int(tuning.get("fft_wave", 0))
    gather = {"gather_mesh": f"void ewamd::k_gather_tile<{r}, {order}, true>" if split_gather
              else None,
              "gather_combine": "ewamd::k_shift_combine" if split_gather else
              f"void ewamd::k_gather_tile<{r}, {order}, false>"}
    if span in gather:
        return gather[span]
    if wave and grid is not None and span.startswith("fft"):
        ln = tuning.get("fft_lanes", (0, 0, 0))
        K = list(grid)
        if wave == 2:
            return {"fft_x_r2c": f"void ewamd::k_wplane_r2c<{r}, {K[0]}, {ln[0]}, false>",
                    "fft_z_influence": f"void ewamd::k_wfft_z<{r}, {K[2] // ln[2]}, {ln[2]}, false>",
                    "fft_x_c2r": f"void ewamd::k_wplane_c2r<{r}, {K[0]}, {ln[0]}, false>"}.get(span)
        return {"fft_x_r2c": f"void ewamd::k_wfft_x_r2c<{r}, {K[0] // ln[0]}, {ln[0]}, false>",
                "fft_y": f"void ewamd::k_wfft_lines<{r}, {K[1] // ln[1]}, {ln[1]}, -1, false>",
                "fft_z_influence": f"void ewamd::k_wfft_z<{r}, {K[2] // ln[2]}, {ln[2]}, false>",
                "fft_y_inverse": f"void ewamd::k_wfft_lines<{r}, {K[1] // ln[1]}, {ln[1]}, 1, false>",
                "fft_x_c2r": f"void ewamd::k_wfft_x_c2r<{r}, {K[0] // ln[0]}, {ln[0]}, false>"}.get(span)
    co = "true" if tuning.get("pair_cosched", 0) else "false"
    direct = int(tuning.get("spread_direct", 0))
    spread = {0: f"void ewamd::k_spread_region<{r}, {order}, {co}>",
              1: f"void ewamd::k_spread_region<{r}, {order}, {co}>",
              2: f"void ewamd::k_spread_owned<{r}, {order}, {co}>"}[direct]
    return {"charges": "ewamd::k_gather_charges",
            "pair_sum": f"void ewamd::k_pair_spmv_stream<{r}, {i}>" if pair_mode != 2 else
            f"void ewamd::k_pair_half<{r}>",
            "spread_region": spread,
            "spread_assemble": {0: f"void ewamd::k_spread_assemble<{r}>",
                                1: f"void ewamd::k_clear<{r}>", 2: None}[direct],
            "fft_x_r2c": f"void ewamd::k_fft_xy_r2c<{r}, {co}>" if planes else
            f"void ewamd::k_fft_x_r2c<{r}, {co}>",
            "fft_y": None if planes else f"void ewamd::k_fft_lines<{r}, 0, {co}>",
            "fft_z_influence": f"void ewamd::k_fft_lines<{r}, 2, {co}>",
            "fft_y_inverse": None if planes else f"void ewamd::k_fft_lines<{r}, 1, {co}>",
            "fft_x_c2r": f"void ewamd::k_fft_xy_c2r<{r}, {co}>" if planes else
            f"void ewamd::k_fft_x_c2r<{r}, {co}>"}.get(span)


def critical_path(trace, calls, tb, prec, prm, step_ms, tuning=None, split_gather=False):
    """The evaluation's critical path from its traced spans: the chain on the evaluation stream
    (spreading, FFTs, the wait for the pair sum, the gather) and, beside it, the charge gather
    and pair sum.  The pair-sum branch is on the critical path when its spans (charge gather and
    pair sum) are at least as long as the chain's spans before the join (the join wait itself
    holds a cross-queue latency of about 9 us whichever branch ends first, so it does not
    decide).  The dominant kernel is the longest span on the critical path."""
    us = {k: trace[k] / calls * 1e3 for k in trace}
    planes = fused_planes(prm, tuning)
    rows = []
    for k in CHAIN + ("charges", "pair_sum"):
        if k in ("fft_y", "fft_y_inverse") and planes:
            continue
        if k == "gather_mesh" and not split_gather:
            continue
        b = tb[k]
        rows.append({"span": k, "beside_chain": k in ("charges", "pair_sum"),
                     "kernel": trace_symbol(k, prec, prm["index_bytes"], prm["order"], planes,
                                            prm["pair_mode"], tuning, prm["grid"], split_gather),
                     "us": round(us[k], 2), "algorithmic_bytes": b,
                     "GBs": round(b / (us[k] * 1e-6) / 1e9, 1) if b and us[k] > 0 else None,
                     "frac": round(b / (us[k] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                     if b and us[k] > 0 else None})
    chain = [r for r in rows if not r["beside_chain"]]
    chain_us = sum(r["us"] for r in chain)
    chain_b = sum(r["algorithmic_bytes"] for r in chain)
    before_join = CHAIN[:CHAIN.index("join_wait")]
    chain_to_join = sum(us[k] for k in before_join if k in us)
    pair_on_path = us["charges"] + us["pair_sum"] >= chain_to_join
    on_path = [r for r in rows if r["kernel"] and (not r["beside_chain"] or pair_on_path)]
    dom = max(on_path, key=lambda r: r["us"])
    spread_fft = [r for r in chain if r["span"].startswith(("spread", "fft"))]
    sf_us = sum(r["us"] for r in spread_fft)
    sf_b = sum(r["algorithmic_bytes"] for r in spread_fft)
    # every kernel of one evaluation (the chain and the branch beside it) against the timed step
    step_b = sum(r["algorithmic_bytes"] for r in rows)
    return {
        "note": "HIP events in the timed launch shape (ew_trace_times, opts.profile = 2): the "
                "chain on the evaluation stream, the charge gather and pair sum beside it on "
                "the auxiliary stream; each traced evaluation waits for its last event",
        "spans": rows, "evaluation_us": round(us["evaluation"], 2),
        "timed_step_us": round(step_ms * 1e3, 2),
        "chain_us": round(chain_us, 2), "chain_bytes": chain_b,
        "chain_frac": round(chain_b / (chain_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "spread_fft_us": round(sf_us, 2), "spread_fft_bytes": sf_b,
        "spread_fft_frac": round(sf_b / (sf_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "pair_sum_on_critical_path": pair_on_path, "dominant": dom,
        "step_bytes": step_b,
        "step_frac": round(step_b / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }


def fused_planes(prm, tuning):
    """Whether the x and y passes of the launch shapes are one kernel per plane (the y phases
    then move nothing): the register plane kernels, or the Stockham plane kernels."""
    wave = int((tuning or {}).get("fft_wave", 0))
    return wave == 2 or (wave == 0 and bool(prm.get("fft_planes", 0)))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spme_triclinic98k_hbm_traffic.json")


def measured_traffic(symbol):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (already scaled by the copy calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["kernels"].get(symbol, {}).get("bytes")
    except (OSError, ValueError, KeyError):
        return None


def timed_run(e, torch, dqs, out, en, steps, warmup):
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for i in range(warmup):
        e.shift_device(dqs[i % len(dqs)].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for i in range(steps):
        e.shift_device(dqs[i % len(dqs)].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def timed_multi(e, torch, nvec, dq, out, en, steps, warmup):
    """ms per ew_shift_multi_device call of nvec vectors (device-resident [nvec][n] arrays)."""
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(warmup):
        e.shift_multi_device(nvec, dq.data_ptr(), out.data_ptr(), en.data_ptr(), sp)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        e.shift_multi_device(nvec, dq.data_ptr(), out.data_ptr(), en.data_ptr(), sp)
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def latency_run(e, torch, dqs, out, en, steps):
    """Per-evaluation latency with a synchronisation after each (p50/p99)."""
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    lat = []
    for i in range(steps):
        t = time.perf_counter()
        e.shift_device(dqs[i % len(dqs)].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
        stream.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def cpu_baseline(w, budget_s=10.0, max_reps=15):
    """The reference fast path (restated in the oracle) on the host cores.

    One untimed warm-up evaluation, then evaluations until budget_s seconds have passed (at most
    max_reps, the reference's own repetition count, coulomb.F90:392-423)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.fft_ewald(w["coords"], w["charges"], w["lattice"], tol=1e-9)
    times = []
    while len(times) < max_reps and sum(times) < budget_s:
        t = time.perf_counter()
        _, _, prm = oracle.fft_ewald(w["coords"], w["charges"], w["lattice"], tol=1e-9)
        times.append(time.perf_counter() - t)
    mean = float(np.mean(times))
    return {"value": round(1.0 / mean, 4), "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": (f"{len(times)} timed evaluations (+1 warm-up) of the 98,304-atom cell by the "
                       f"oracle's restatement of the reference fast path (linked-cell erfc + "
                       f"Gaussian NFFT, tol 1e-9, {prm['nmodes'][0]}^3 modes, 2x oversampled, "
                       f"window m={prm['window_cutoff']}, own mixed-radix FFT, OpenMP): mean "
                       f"{mean:.3f} s, min {min(times):.3f} s per evaluation")}


def copy_calibration(torch, mib=512, reps=5):
    """Plain device-to-device copies of a known size, for scaling HBM byte counters."""
    x = torch.ones(mib * 2**20 // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: {reps} copies of {mib} MiB (read {mib} MiB + write {mib} "
          f"MiB each)", file=sys.stderr, flush=True)
    del x, y


def batch_bytes(prm, n, rs):
    """Algorithmic bytes of one evaluation of one structure (all phases)."""
    pb = phase_bytes(n, prm["grid"], prm["order"], prm["n_pairs"], rs, rs, prm["index_bytes"],
                     False, bool(prm.get("fft_planes", 0)))
    return sum(pb.values())


def batch_leg(args, torch, ew, W, steps, warmup, sequential=True, tune="rules"):
    """Batched workload: 256 independent triclinic structures (N log-uniform in [64, 8192],
    angles U[70, 110] deg, density 0.1 A^-3, ~0.8 A grids, order 6) evaluated together by one
    grouped launch per kernel kind; FP64 and mixed; optionally the same structures through
    single-structure engines one after another.  tune: the members' launch shapes ("rules", or
    "auto": each member autotuned, or read from the tuning cache, as it is added).  Returns
    (results per precision, structures, atoms, worst mixed-vs-FP64 shift error)."""
    t = time.perf_counter()
    ws = W.batch()
    print(f"[bench] batch workload generated in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    dq_t = [torch.tensor(w["charges"], dtype=torch.float64, device=dev) for w in ws]
    sh_t = [torch.empty_like(d) for d in dq_t]
    en_t = torch.empty(len(ws), dtype=torch.float64, device=dev)
    natoms = sum(len(w["charges"]) for w in ws)
    res, shifts = {}, {}
    for prec in ("fp64", "mixed"):
        rs = 4 if prec == "mixed" else 8
        with ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=ws[0]["order"],
                          precision=prec, group_grid=args.group_grid, tune=tune) as pool:
            t = time.perf_counter()
            ids = [pool.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"])
                   for w in ws]
            torch.cuda.synchronize()
            geom_s = time.perf_counter() - t
            alg = sum(batch_bytes(pool.parameters(i), len(w["charges"]), rs)
                      for i, w in zip(ids, ws))
            dptr = [d.data_ptr() for d in dq_t]
            sptr = [d.data_ptr() for d in sh_t]
            for _ in range(warmup):
                pool.shift_device(ids, dptr, sptr, en_t.data_ptr(), sp)
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            for _ in range(steps):
                pool.shift_device(ids, dptr, sptr, en_t.data_ptr(), sp)
            t1.record(stream)
            torch.cuda.synchronize()
            ms = t0.elapsed_time(t1) / steps
            shifts[prec] = [s.cpu().numpy() for s in sh_t]
        res[prec] = {"ms_per_step": ms, "evals_per_s": len(ws) * 1e3 / ms,
                     "atoms_per_s": natoms * 1e3 / ms, "geometry_s": geom_s,
                     "algorithmic_bytes": alg, "GBs": alg / (ms * 1e-3) / 1e9}
        print(f"[bench] batch {prec}: {ms:.3f} ms/step ({len(ws)} structures, {natoms} atoms), "
              f"{len(ws) * 1e3 / ms:.0f} evals/s; geometry {geom_s:.1f} s", file=sys.stderr,
              flush=True)
        if not sequential:
            continue
        # the same structures one after another through single-structure engines
        engines = []
        for w in ws:
            e = ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"],
                               precision=prec)
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            engines.append(e)
        for e, d, o in zip(engines, dq_t, sh_t):
            e.shift_device(d.data_ptr(), o.data_ptr(), 0, sp)
        torch.cuda.synchronize()
        steps_seq = max(1, steps // 4)
        t0.record(stream)
        for _ in range(steps_seq):
            for e, d, o in zip(engines, dq_t, sh_t):
                e.shift_device(d.data_ptr(), o.data_ptr(), 0, sp)
        t1.record(stream)
        torch.cuda.synchronize()
        ms_seq = t0.elapsed_time(t1) / steps_seq
        for e in engines:
            e.close()
        res[prec]["sequential_ms_per_step"] = ms_seq
        res[prec]["grouped_over_sequential"] = ms_seq / ms
        print(f"[bench] batch {prec}: sequential engines {ms_seq:.3f} ms/step", file=sys.stderr,
              flush=True)
    err = max(float(np.abs(a - b).max() / np.abs(b).max())
              for a, b in zip(shifts["mixed"], shifts["fp64"]))
    return res, ws, natoms, err


def run_batch(args, torch, ew, W):
    """The batched workload as its own JSON line (see batch_leg)."""
    res, ws, natoms, err = batch_leg(args, torch, ew, W, args.steps, args.warmup,
                                     tune=args.tune)
    head = res[args.precision]
    line = {
        "metric": "SPME shift+energy evals/s (batch of 256 structures)",
        "value": round(head["evals_per_s"], 2), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 mesh/FFT/pair coefficients + fp64 accumulation" if args.precision ==
        "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": "256 triclinic structures, N log-uniform [64, 8192], angles "
                               "U[70,110] deg, density 0.1 A^-3, grids ~0.8 A (2^a3^b5^c), "
                               "order 6, r_cut 10 A, tol 1e-10; one grouped launch per kernel "
                               "kind, graph replay", "structures": len(ws), "atoms": natoms,
                   "precision": args.precision, "launch_shapes": args.tune},
        "fp64_evals_per_s": round(res["fp64"]["evals_per_s"], 2),
        "mixed_evals_per_s": round(res["mixed"]["evals_per_s"], 2),
        "atoms_per_s": {k: round(v["atoms_per_s"]) for k, v in res.items()},
        "grouped_over_sequential": {k: round(v["grouped_over_sequential"], 3)
                                    for k, v in res.items()},
        "mixed_vs_fp64_rel_err": {"shift_max": err},
        "geometry_setup_s": {k: round(v["geometry_s"], 2) for k, v in res.items()},
        "roofline": {"bound": "hbm", "kernel": "whole grouped evaluation",
                     "achieved": round(head["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": head["algorithmic_bytes"]},
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def poisson_load(srv, ids, bufs, qps, seconds, rng):
    """Offers SCC requests as a Poisson stream at `qps` for `seconds` (a structure gets a new
    request only after its previous one completed; bufs[sid] = (dq, shift, energy) host arrays
    reused per structure; an arrival that finds every structure busy is dropped).  Returns
    (offered, completed, elapsed, dropped)."""
    import queue
    import threading
    busy = set()
    lock = threading.Lock()
    q = queue.SimpleQueue()
    done = [0]
    ptrs = {sid: (b[0].ctypes.data, b[1].ctypes.data, b[2].ctypes.data) for sid, b in bufs.items()}

    def collector():
        while True:
            item = q.get()
            if item is None:
                return
            sid, t = item
            srv.wait(t)
            with lock:
                busy.discard(sid)
                done[0] += 1

    th = threading.Thread(target=collector)
    th.start()
    offered = dropped = 0
    n_ids = len(ids)
    gaps = rng.exponential(1.0 / qps, int(qps * seconds * 1.2) + 16)
    picks = rng.integers(n_ids, size=len(gaps) * 2)
    t0 = time.perf_counter()
    t_next = t0
    end = t0 + seconds
    k = 0
    for gap in gaps:
        t_next += gap
        if t_next >= end:
            break
        while True:
            g = t_next - time.perf_counter()
            if g <= 0:
                break
            if g > 3e-4:
                time.sleep(g - 2e-4)
        sid = None
        with lock:
            if len(busy) < n_ids:  # every structure has a request outstanding: arrival dropped
                for _ in range(64):
                    c = ids[picks[k % len(picks)]]
                    k += 1
                    if c not in busy:
                        sid = c
                        break
                if sid is None:
                    sid = next(c for c in ids if c not in busy)
                busy.add(sid)
        offered += 1
        if sid is None:
            dropped += 1
            continue
        p = ptrs[sid]
        q.put((sid, srv.submit_ptr(sid, p[0], p[1], p[2])))
    q.put(None)
    th.join()
    srv.drain()
    return offered, done[0], time.perf_counter() - t0, dropped


def serving_leg(args, ew, W, sweep, seconds, precisions=("mixed", "fp64"), tune="rules"):
    """Serving workload: 1,000 resident structures (N log-uniform in [100, 20000], built like
    the 98,304-atom cell); SCC charge updates arrive as a Poisson stream at the rates `sweep` and
    are served by the continuous-batching loop; throughput and p50/p99 latency per precision.
    Returns (results per precision, structures, atoms)."""
    t = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    ws = W.serving(workers=workers)
    print(f"[bench] serving workload generated in {time.perf_counter() - t:.1f} s "
          f"({workers} processes)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(2026)
    natoms = sum(len(w["charges"]) for w in ws)
    res = {}
    for prec in precisions:
        with ew.EwaldPool(tol=1e-10, r_cut=ws[0]["r_cut"], order=ws[0]["order"],
                          precision=prec, tune=tune) as pool:
            t = time.perf_counter()
            ids = [pool.add(w["coords"], w["lattice"], charges=w["charges"], grid=w["grid"])
                   for w in ws]
            setup = time.perf_counter() - t
            used = pool.usage()[0]
            bufs = {i: (W.scc_update(w["charges"], rng), np.empty(len(w["charges"])), np.empty(1))
                    for i, w in zip(ids, ws)}
            print(f"[bench] serving {prec}: {len(ids)} structures resident, {used / 2**30:.1f} "
                  f"GiB, set-up {setup:.1f} s", file=sys.stderr, flush=True)
            rows = []
            with ew.EwaldServer(pool, max_batch=args.max_batch, max_inflight=2) as srv:
                poisson_load(srv, ids, bufs, sweep[0], 0.5, rng)  # warm-up
                srv.stats(reset=True)
                for qps in sweep:
                    offered, completed, el, dropped = poisson_load(srv, ids, bufs, qps,
                                                                   seconds, rng)
                    st = srv.stats(reset=True)
                    row = {"offered_qps": qps, "achieved_qps": round(completed / el, 1),
                           "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                           "max_ms": round(st["max_ms"], 4),
                           "mean_batch": round(st["mean_batch"], 2), "requests": completed,
                           "dropped": dropped}
                    rows.append(row)
                    print(f"[bench] serving {prec} {qps:.0f} qps: achieved {row['achieved_qps']}"
                          f", p50 {row['p50_ms']} ms, p99 {row['p99_ms']} ms, batch "
                          f"{row['mean_batch']}, dropped {dropped}", file=sys.stderr, flush=True)
                    if completed < 0.9 * qps * el:  # past saturation: stop the sweep
                        break
            ok = [r for r in rows if r["p99_ms"] <= 2.0 and r["dropped"] == 0 and
                  r["achieved_qps"] >= 0.97 * r["offered_qps"]]
            best = max(ok, key=lambda r: r["achieved_qps"]) if ok else None
            res[prec] = {"sweep": rows, "sustained": best, "setup_s": round(setup, 1),
                         "resident_bytes": used}
    return res, ws, natoms


def run_serving(args, torch, ew, W):
    """The serving workload as its own JSON line (see serving_leg)."""
    res, ws, natoms = serving_leg(args, ew, W, [float(x) for x in args.qps.split(",")],
                                  args.seconds, tune=args.tune)
    head = res[args.precision]
    best = head["sustained"]
    line = {
        "metric": "serving: sustained SCC requests/s at p99 <= 2 ms (1,000 resident structures)",
        "value": best["achieved_qps"] if best else 0.0, "unit": "requests/s", "n_gpus": 1,
        "steps": None, "warmup": None, "ms_per_step": best["p50_ms"] if best else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 mesh/FFT/pair coefficients + fp64 accumulation" if args.precision ==
        "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": "1,000 resident triclinic structures (gamma 75 deg, density "
                               "0.1036 A^-3, Poisson-disc 1 A), N log-uniform [100, 20000], "
                               "grids ~0.777 A, order 6, r_cut 10 A, tol 1e-10; Poisson "
                               "arrivals of dq + N(0, 0.01) updates; continuous batching, "
                               f"max_batch {args.max_batch}, 2 batches in flight",
                   "structures": len(ws), "atoms": natoms, "precision": args.precision,
                   "seconds_per_rate": args.seconds, "launch_shapes": args.tune},
        "p50_ms": best["p50_ms"] if best else None, "p99_ms": best["p99_ms"] if best else None,
        "serving": res, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


def run_mesh(args, torch, ew, W):
    """Spreading and FFT bandwidth against mesh size: the 98,304-atom cell with its 128^3 mesh
    and finer meshes (same atoms, order 6, r_cut 10 A), every phase alone on the stream (HIP
    events, profiled evaluations), FP64 and mixed.  Reports each spread/FFT phase's algorithmic
    bytes over its time as a fraction of the 8 TB/s HBM peak."""
    w = W.triclinic()
    n = len(w["charges"])
    rng = np.random.default_rng(99)
    dq = torch.tensor(W.scc_update(w["charges"], rng), dtype=torch.float64, device="cuda:0")
    out = torch.empty(n, dtype=torch.float64, device="cuda:0")
    en = torch.empty(1, dtype=torch.float64, device="cuda:0")
    keys = ("spread", "fft_x_r2c", "fft_y", "fft_z_influence", "fft_y_inverse", "fft_x_c2r")
    rows = []
    for K in [int(k) for k in args.meshes.split(",")]:
        for prec in ("fp64", "mixed"):
            with ew.EwaldEngine(tol=1e-10, r_cut=w["r_cut"], grid=K, order=w["order"],
                                precision=prec, profile=True) as e:
                e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
                prm = e.parameters()
                tun = e.tuning()
                wave = int(tun.get("fft_wave", 0))
                for _ in range(3):
                    e.shift_device(dq.data_ptr(), out.data_ptr(), en.data_ptr())
                e.phase_times(reset=True)
                for _ in range(args.profile_steps):
                    e.shift_device(dq.data_ptr(), out.data_ptr(), en.data_ptr())
                ph, calls = e.phase_times()
            rs = 4 if prec == "mixed" else 8
            planes = fused_planes(prm, tun)
            pb = phase_bytes(n, prm["grid"], prm["order"], prm["n_pairs"], rs, rs,
                             prm["index_bytes"], False, planes, wave > 0)
            row = {"mesh": list(prm["grid"]), "precision": prec, "fft_planes": int(planes),
                   "fft_wave": wave}
            tot_b, tot_ms = 0, 0.0
            for k in keys:
                ms = ph[k] / calls
                if pb[k]:
                    row[k] = {"us": round(ms * 1e3, 1),
                              "frac": round(pb[k] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}
                tot_b += pb[k]
                tot_ms += ms
            row["spread_fft_frac"] = round(tot_b / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)
            rows.append(row)
            print(f"[bench] mesh {K}^3 {prec}: spread+FFT {tot_ms * 1e3:.1f} us, "
                  f"{row['spread_fft_frac']:.3f} of HBM peak", file=sys.stderr, flush=True)
    best = max(rows, key=lambda r: r["spread_fft_frac"])
    line = {
        "metric": "spread + FFT algorithmic bandwidth, fraction of HBM peak, against mesh size",
        "value": best["spread_fft_frac"], "unit": "fraction of 8 TB/s", "n_gpus": 1,
        "steps": args.profile_steps, "warmup": 3, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64 and fp32 meshes", "data": "synthetic",
        "config": {"workload": "triclinic 98,304 atoms (the headline cell), SPME meshes "
                               f"{args.meshes} per axis, order 6, r_cut 10 A; each phase alone "
                               "on the stream (HIP events, ~4 us of event overhead each)",
                   "atoms": n},
        "meshes": rows, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


def replicas_leg(args, torch, ew, W, w, qs, fp64_plain_ms):
    """Several geometries of the 98,304-atom cell (the cell and copies with every atom moved by
    N(0, 0.05 A): SCC cycles of neighbouring molecular-dynamics snapshots) resident in one pool,
    all evaluated in one grouped launch per kernel kind; evaluations per second against the
    plain FP64 engine (one structure, one charge vector per call)."""
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    rng = np.random.default_rng(31415)
    geos = [w["coords"]] + [w["coords"] + rng.normal(scale=0.05 * W.AA, size=w["coords"].shape)
                            for _ in range(args.replicas - 1)]
    dq_t = [torch.tensor(qs[k % len(qs)], dtype=torch.float64, device=dev)
            for k in range(args.replicas)]
    sh_t = [torch.empty_like(d) for d in dq_t]
    en_t = torch.empty(args.replicas, dtype=torch.float64, device=dev)
    out = {}
    for prec in ("fp64", "mixed"):
        with ew.EwaldPool(tol=1e-10, r_cut=w["r_cut"], order=w["order"], precision=prec,
                          group_grid=args.group_grid) as pool:
            ids = [pool.add(x, w["lattice"], charges=w["charges"], grid=w["grid"]) for x in geos]
            dptr = [d.data_ptr() for d in dq_t]
            sptr = [d.data_ptr() for d in sh_t]
            for _ in range(3):
                pool.shift_device(ids, dptr, sptr, en_t.data_ptr(), sp)
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            steps = max(5, args.steps // 4)
            t0.record(stream)
            for _ in range(steps):
                pool.shift_device(ids, dptr, sptr, en_t.data_ptr(), sp)
            t1.record(stream)
            torch.cuda.synchronize()
            ms = t0.elapsed_time(t1) / steps
        out[prec] = {"ms_per_launch": round(ms, 5),
                     "evals_per_s": round(args.replicas * 1e3 / ms, 1),
                     "over_plain_fp64": round(args.replicas * fp64_plain_ms / ms, 3)}
        print(f"[bench] {args.replicas} resident 98,304-atom geometries, {prec}: {ms:.3f} ms per "
              f"grouped evaluation, {args.replicas * 1e3 / ms:.0f} evals/s", file=sys.stderr,
              flush=True)
    return {"structures": args.replicas, "note": "one grouped launch per kernel kind over "
            "several resident geometries of the 98,304-atom cell (atoms moved by N(0, 0.05 A)); "
            "over_plain_fp64 divides by the plain FP64 engine's one-structure rate", **out}


def cpu_legs(w_c1, w_c2, threads):
    """Further CPU baselines on the host cores (reported only): the reference's direct Ewald
    (restated in the oracle) on the 512-atom NaCl cell at tol 1e-10, and the stock DFTB+ path's
    per-SCC-iteration cost, the product of the N x N Ewald matrix with the charges
    (coulomb.F90:1253-1308, 868; a symmetric random matrix of that size, whose cost does not
    depend on its entries) for the 512- and 12,288-atom cells."""
    import oracle
    legs = {}
    a = oracle.optimal_alpha_dftb(w_c1["lattice"], 1e-10)
    oracle.direct_ewald(w_c1["coords"], w_c1["charges"], w_c1["lattice"], tol=1e-10, alpha=a)
    t = time.perf_counter()
    reps = 0
    while reps < 3 or time.perf_counter() - t < 2.0:
        oracle.direct_ewald(w_c1["coords"], w_c1["charges"], w_c1["lattice"], tol=1e-10, alpha=a)
        reps += 1
        if reps >= 10:
            break
    legs["direct_ewald_nacl512"] = {"s_per_eval": round((time.perf_counter() - t) / reps, 4),
                                    "cores": threads, "kind": "port",
                                    "sample": f"{reps} evaluations, tol 1e-10, alpha {a:.4f}"}
    for name, w in (("nacl512", w_c1), ("water12288", w_c2)):
        n = len(w["charges"])
        rng = np.random.default_rng(5)
        m = rng.standard_normal((n, n))
        m = m + m.T
        q = np.ascontiguousarray(w["charges"])
        m @ q
        t = time.perf_counter()
        reps = 0
        while reps < 5 or time.perf_counter() - t < 1.0:
            m @ q
            reps += 1
            if reps >= 200:
                break
        legs[f"stock_hemv_{name}"] = {
            "s_per_scc_iteration": round((time.perf_counter() - t) / reps, 6), "cores": threads,
            "kind": "port", "matrix_bytes": 8 * n * n,
            "sample": f"{reps} products of the {n} x {n} FP64 matrix with the charges (numpy)"}
        del m
    print(f"[bench] cpu legs: {legs}", file=sys.stderr, flush=True)
    return legs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--latency-steps", type=int, default=500)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run plain 512 MiB copies first (for HBM counter scaling)")
    ap.add_argument("--workload", default="triclinic",
                    choices=["triclinic", "batch", "serving", "mesh"],
                    help="triclinic: the 98,304-atom cell (headline); batch: 256 structures; "
                         "serving: 1,000 resident structures under Poisson load; mesh: spread "
                         "and FFT bandwidth of the 98,304-atom cell against mesh size")
    ap.add_argument("--qps", default="5000,10000,20000,40000,60000,80000,100000,120000,140000",
                    help="serving: offered request rates")
    ap.add_argument("--seconds", type=float, default=2.0, help="serving: seconds per rate")
    ap.add_argument("--max-batch", type=int, default=64, help="serving: structures per launch")
    ap.add_argument("--vectors", default="2,4,8",
                    help="triclinic: charge vectors per multi-vector call (empty: skip)")
    ap.add_argument("--meshes", default="128,192,256,320",
                    help="mesh workload: mesh points per axis")
    ap.add_argument("--external", type=int, default=20000,
                    help="triclinic: external point charges timed once per precision (0: skip)")
    ap.add_argument("--group-grid", type=int, default=0, choices=[0, 1],
                    help="grouped launches (batch pools, the triclinic replicas): 0 one "
                         "workgroup per work item (default), 1 persistent grids")
    ap.add_argument("--tune", default=None, choices=["auto", "rules"],
                    help="launch shapes of the timed engines: the autotuner (the default for "
                         "the 98,304-atom cell) or the fixed rules (the default for the batch "
                         "and serving pools, where per-structure tuning is slower)")
    ap.add_argument("--replicas", type=int, default=4,
                    help="triclinic: resident geometries of the 98,304-atom cell evaluated in "
                         "one grouped launch (0: skip)")
    ap.add_argument("--no-workload-legs", action="store_true",
                    help="triclinic: skip the compact batch and serving legs")
    ap.add_argument("--headline-only", action="store_true",
                    help="triclinic: only the timed, latency and traced passes (every chain "
                         "kernel then runs in the timed launch shape, for kernel-trace "
                         "profiles)")
    args = ap.parse_args()
    if args.headline_only:
        args.vectors = ""
        args.external = 0
        args.no_cpu_baseline = True
        args.replicas = 0
        args.no_workload_legs = True
    args.vectors = [int(v) for v in args.vectors.split(",") if v.strip()]
    if args.tune is None:
        args.tune = "rules" if args.workload in ("batch", "serving") else "auto"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not supported "
              "(structures shard across GPUs by footprint)", file=sys.stderr)
        sys.exit(2)

    import torch

    import fftewald_amd as ew
    from fftewald_amd import workloads as W

    if args.copy_calibration:
        copy_calibration(torch)
    if args.workload == "batch":
        run_batch(args, torch, ew, W)
        return
    if args.workload == "serving":
        run_serving(args, torch, ew, W)
        return
    if args.workload == "mesh":
        run_mesh(args, torch, ew, W)
        return
    t = time.perf_counter()
    w = W.triclinic()
    print(f"[bench] workload generated in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    n = len(w["charges"])
    rng = np.random.default_rng(12345)
    qs = [w["charges"]]
    for _ in range(max(3, max(args.vectors, default=0) - 1)):
        qs.append(W.scc_update(qs[-1], rng))
    dev = torch.device("cuda:0")
    dqs = [torch.tensor(q, dtype=torch.float64, device=dev) for q in qs]
    out = torch.empty(n, dtype=torch.float64, device=dev)
    en = torch.empty(1, dtype=torch.float64, device=dev)
    opts = dict(tol=1e-10, r_cut=w["r_cut"], grid=w["grid"], order=w["order"])

    res = {}
    shifts = {}
    for prec in ("fp64", "mixed"):
        # the fixed launch-shape rules, for comparison with the autotuned engine below
        with ew.EwaldEngine(precision=prec, **opts) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            rules_ms = timed_run(e, torch, dqs, out, en, args.steps, args.warmup)
            rules = e.tuning()
        with ew.EwaldEngine(precision=prec, tune=args.tune, **opts) as e:
            t = time.perf_counter()
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            torch.cuda.synchronize()
            geom_s = time.perf_counter() - t
            prm = e.parameters()
            tuning = e.tuning()
            ms = timed_run(e, torch, dqs, out, en, args.steps, args.warmup)
            print(f"[bench] {prec}: rules {rules_ms * 1e3:.1f} us, {args.tune} {ms * 1e3:.1f} us "
                  f"per evaluation; shapes {tuning}", file=sys.stderr, flush=True)
            p50, p99 = latency_run(e, torch, dqs, out, en, args.latency_steps)
            s, energy = e.shift(qs[0])
            shifts[prec] = (s, energy)
            # the host-array entry point: charges over PCIe in, shifts back, synchronised
            t = time.perf_counter()
            for i in range(10):
                e.shift(qs[i % 4])
            host_ms = (time.perf_counter() - t) * 100.0
            self_energy = -prm["alpha"] / math.sqrt(math.pi) * float(np.dot(qs[0], qs[0]))
            # several charge vectors through the same geometry per call
            multi = {}
            for nv in args.vectors:
                dqm = torch.stack(dqs[:nv]).contiguous()
                outm = torch.empty_like(dqm)
                enm = torch.empty(nv, dtype=torch.float64, device=dev)
                msm = timed_multi(e, torch, nv, dqm, outm, enm, max(1, args.steps // 2),
                                  args.warmup)
                multi[nv] = {"ms_per_call": round(msm, 5), "evals_per_s": round(nv * 1e3 / msm, 1)}
                if prec == "mixed" and "fp64" in shifts:
                    s0 = shifts["fp64"][0]
                    multi[nv]["rel_err_vs_fp64"] = float(
                        np.abs(outm[0].cpu().numpy() - s0).max() / np.abs(s0).max())
                print(f"[bench] {prec} multi-vector x{nv}: {msm:.4f} ms/call, "
                      f"{nv * 1e3 / msm:.1f} evals/s", file=sys.stderr, flush=True)
                del dqm, outm, enm
            # the per-geometry work of a molecular-dynamics step: a new geometry (the atoms moved
            # a little), the Ewald gradients and the stress at the current charges
            md = None
            if not args.headline_only:
                rg = np.random.default_rng(4242)
                geo = []
                for k in range(4):
                    xk = w["coords"] + rg.normal(scale=0.01, size=w["coords"].shape)
                    t = time.perf_counter()
                    e.set_geometry(xk, w["lattice"], charges=w["charges"])
                    geo.append((time.perf_counter() - t) * 1e3)
                e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
                e.gradients(qs[0])
                t = time.perf_counter()
                for _ in range(3):
                    e.gradients(qs[0])
                grad_ms = (time.perf_counter() - t) / 3 * 1e3
                e.stress(qs[0])
                t = time.perf_counter()
                for _ in range(3):
                    e.stress(qs[0])
                stress_ms = (time.perf_counter() - t) / 3 * 1e3
                md = {"set_geometry_ms": round(float(np.mean(geo[1:])), 3),
                      "gradients_ms": round(grad_ms, 3), "stress_ms": round(stress_ms, 3)}
                print(f"[bench] {prec} geometry step {md['set_geometry_ms']:.2f} ms, gradients "
                      f"{grad_ms:.2f} ms, stress {stress_ms:.2f} ms", file=sys.stderr, flush=True)
            # external point charges (once per geometry in DFTB+): potential at the atoms
            # and both sides' gradients, host arrays, averaged over a few calls
            ext = None
            if args.external > 0:
                ry = np.random.default_rng(777)
                ey = ry.random((args.external, 3)) @ w["lattice"]
                eq = ry.normal(size=args.external)
                e.external_potential(ey, eq)
                t = time.perf_counter()
                for _ in range(3):
                    e.external_potential(ey, eq)
                tp = (time.perf_counter() - t) / 3
                e.external_gradients(ey, eq, dq=qs[0])
                t = time.perf_counter()
                for _ in range(3):
                    e.external_gradients(ey, eq, dq=qs[0])
                tg = (time.perf_counter() - t) / 3
                ext = {"charges": args.external, "potential_ms": round(tp * 1e3, 3),
                       "gradients_ms": round(tg * 1e3, 3)}
                print(f"[bench] {prec} external charges x{args.external}: potential "
                      f"{tp * 1e3:.2f} ms, gradients {tg * 1e3:.2f} ms", file=sys.stderr,
                      flush=True)
        sp = torch.cuda.current_stream().cuda_stream
        # the timed launch shape again, with HIP events between its kernels (ew_trace_times)
        with ew.EwaldEngine(precision=prec, profile="trace", tuning=tuning, **opts) as e:
            e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
            for i in range(3):
                e.shift_device(dqs[i % 4].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
            e.trace_times(reset=True)
            for i in range(args.profile_steps):
                e.shift_device(dqs[i % 4].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
            trace, tcalls = e.trace_times()
        rs = 4 if prec == "mixed" else 8
        planes = fused_planes(prm, tuning)
        split = bool(prm.get("gather_split", 0))
        tb = trace_bytes(n, prm["grid"], prm["order"], prm["n_pairs"], rs, rs,
                         prm["index_bytes"], False, planes, trace["spread_region"] > 0, tuning,
                         split, int(tuning.get("fft_wave", 0)) > 0)
        crit = critical_path(trace, tcalls, tb, prec, prm, ms, tuning, split)
        print(f"[bench] {prec} critical path: chain {crit['chain_us']:.1f} us "
              f"({crit['chain_frac']:.3f} of HBM peak), dominant {crit['dominant']['span']} "
              f"{crit['dominant']['us']:.1f} us", file=sys.stderr, flush=True)
        # every phase alone on the stream (ew_phase_times)
        phases, calls = {}, 0
        if not args.headline_only:
            with ew.EwaldEngine(precision=prec, profile=True, tuning=tuning, **opts) as e:
                e.set_geometry(w["coords"], w["lattice"], charges=w["charges"])
                for i in range(3):
                    e.shift_device(dqs[i % 4].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
                e.phase_times(reset=True)
                for i in range(args.profile_steps):
                    e.shift_device(dqs[i % 4].data_ptr(), out.data_ptr(), en.data_ptr(), sp)
                phases, calls = e.phase_times()
        pb = phase_bytes(n, prm["grid"], prm["order"], prm["n_pairs"], rs, rs,
                         prm["index_bytes"], False, planes, int(tuning.get("fft_wave", 0)) > 0)
        ph = {}
        for k, v in phases.items():
            avg_ms = v / calls
            ph[k] = {"ms": round(avg_ms, 5), "bytes": pb[k],
                     "GBs": round(pb[k] / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 and pb[k]
                     else None}
            if planes and k in ("fft_y", "fft_y_inverse"):
                ph[k]["note"] = "fused into the x-y plane kernels (event overhead only)"
            if k == "pair_sum" and avg_ms > 0:
                # in-cutoff pairs of the resident list summed per second (each linked pair twice)
                ph[k]["pairs_per_s"] = round(prm["n_pairs"] / (avg_ms * 1e-3), 1)
        res[prec] = {"ms_per_eval": ms, "evals_per_s": 1e3 / ms, "geometry_s": geom_s,
                     "multi": multi, "external": ext, "md_step": md, "critical_path": crit,
                     "tuning": tuning, "rules_ms": rules_ms, "rules": rules,
                     "self_energy": self_energy,
                     "p50_ms": p50, "p99_ms": p99, "params": prm, "phases": ph,
                     "host_api_ms": host_ms}
        print(f"[bench] {prec}: {ms:.4f} ms/eval, {1e3 / ms:.1f} evals/s, geometry {geom_s:.2f} s,"
              f" p50 {p50:.3f} ms p99 {p99:.3f} ms", file=sys.stderr, flush=True)

    replicas = None
    if args.replicas > 0:
        replicas = replicas_leg(args, torch, ew, W, w, qs, res["fp64"]["ms_per_eval"])
    workload_legs = None
    if not args.no_workload_legs:
        bres, bws, batoms, berr = batch_leg(args, torch, ew, W, max(5, args.steps // 5), 3,
                                            sequential=False)
        sres, sws, satoms = serving_leg(args, ew, W, [20000.0, 60000.0, 80000.0, 100000.0,
                                                      120000.0],
                                        1.0, precisions=("mixed",))
        sbest = sres["mixed"]["sustained"]
        workload_legs = {
            "batch_256": {"structures": len(bws), "atoms": batoms,
                          "fp64_evals_per_s": round(bres["fp64"]["evals_per_s"], 1),
                          "mixed_evals_per_s": round(bres["mixed"]["evals_per_s"], 1),
                          "ms_per_step": {k: round(v["ms_per_step"], 4) for k, v in bres.items()},
                          "mixed_vs_fp64_shift_max": berr,
                          "roofline_whole_evaluation": {
                              k: {"achieved_GBs": round(v["GBs"], 1),
                                  "frac": round(v["GBs"] / HBM_PEAK_GBS, 4),
                                  "algorithmic_bytes": v["algorithmic_bytes"]}
                              for k, v in bres.items()}},
            "serving_1000": {"structures": len(sws), "atoms": satoms, "precision": "mixed",
                             "sweep": sres["mixed"]["sweep"],
                             "sustained_qps_at_p99_2ms": sbest["achieved_qps"] if sbest else None,
                             "p50_ms": sbest["p50_ms"] if sbest else None,
                             "p99_ms": sbest["p99_ms"] if sbest else None,
                             "seconds_per_rate": 1.0}}
    s64, e64 = shifts["fp64"]
    s32, e32 = shifts["mixed"]
    err_shift = float(np.abs(s32 - s64).max() / np.abs(s64).max())
    err_energy = abs(e32 - e64) / abs(e64)
    head = res[args.precision]
    # the roofline kernel: the longest kernel on the timed step's critical path, timed by HIP
    # events in that launch shape (see critical_path)
    d = head["critical_path"]["dominant"]
    achieved = d["GBs"]
    roofline = {"bound": "hbm", "kernel": d["kernel"], "phase": d["span"], "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"],
                "traffic": measured_traffic(d["kernel"]),
                "algorithmic_bytes": d["algorithmic_bytes"], "avg_ms": round(d["us"] * 1e-3, 5),
                "launch_shape": "timed step (pair sum beside the chain, direct launches)",
                "whole_step": {"algorithmic_bytes": head["critical_path"]["step_bytes"],
                               "ms": round(head["ms_per_eval"], 5),
                               "achieved": round(head["critical_path"]["step_bytes"] /
                                                 (head["ms_per_eval"] * 1e-3) / 1e9, 1),
                               "frac": head["critical_path"]["step_frac"],
                               "note": "algorithmic bytes of every kernel of one evaluation "
                                       "(chain and pair branch) over the timed step"}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w)
        print(f"[bench] cpu baseline {cpu['value']:.4f} evals/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)
        cpu["other_legs"] = cpu_legs(W.nacl(), W.water(), cpu["cores"])
    line = {
        "metric": "SPME shift+energy evals/s (98k atoms), p99 latency; mixed vs FP64 @1e-6 rel",
        "value": round(head["evals_per_s"], 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_eval"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 mesh/FFT/pair coefficients + fp64 accumulation" if args.precision ==
        "mixed" else "fp64",
        "data": "synthetic",
        "config": {"workload": "triclinic 98,304 atoms, a=b=c=99.4 A, gamma=75 deg, "
                               "Poisson-disc d_min 1 A, q U[-1,1]; SPME 128^3 order 6, "
                               "r_cut 10 A, tol 1e-10 (alpha from the reference's real-space "
                               "budget rule); geometry resident, charges in HBM; launch "
                               "shapes from the autotuner (tuned at ew_set_geometry, outside "
                               "the timed region)",
                   "atoms": n, "grid": list(head["params"]["grid"]),
                   "order": head["params"]["order"], "alpha": head["params"]["alpha"],
                   "r_cut_bohr": head["params"]["r_cut"], "pairs": head["params"]["n_pairs"],
                   "precision": args.precision},
        "p50_ms": round(head["p50_ms"], 4),
        "p99_ms": round(head["p99_ms"], 4),
        "fp64_evals_per_s": round(res["fp64"]["evals_per_s"], 2),
        "mixed_evals_per_s": round(res["mixed"]["evals_per_s"], 2),
        "mixed_over_fp64": round(res["mixed"]["evals_per_s"] / res["fp64"]["evals_per_s"], 3),
        "mixed_vs_fp64_rel_err": {"shift_max": err_shift, "energy": err_energy},
        "mixed_gate": {"shift_rel_limit": 1e-6, "shift_rel_max": err_shift,
                       "shift_met": err_shift <= 1e-6,
                       "energy_rel": err_energy,
                       "energy_rel_to_self_energy": abs(e32 - e64) / abs(res["fp64"]["self_energy"]),
                       "energy_note": "the energy is a difference of large terms (self energy "
                                      f"{res['fp64']['self_energy']:.1f} Ha against E = "
                                      f"{e64:.1f} Ha): the FP32 mesh's ~1e-7 relative error of "
                                      "the reciprocal energy is amplified by that cancellation; "
                                      "the 1e-6 gate of the metric is on the shifts"},
        "critical_path": {k: v["critical_path"] for k, v in res.items()},
        "replicas_98k": replicas,
        "workload_legs": workload_legs,
        "launch_shapes": {k: {"tuned": v["tuning"], "rules": v["rules"],
                              "rules_ms_per_eval": round(v["rules_ms"], 5),
                              "tuned_ms_per_eval": round(v["ms_per_eval"], 5)}
                          for k, v in res.items()},
        "multi_vector": {k: v["multi"] for k, v in res.items()},
        "best_mixed_multi_over_plain_fp64": round(
            max(m["evals_per_s"] for m in res["mixed"]["multi"].values()) /
            res["fp64"]["evals_per_s"], 3) if res["mixed"]["multi"] else None,
        "geometry_setup_s": {k: round(v["geometry_s"], 3) for k, v in res.items()},
        "external_charges": {k: v["external"] for k, v in res.items()},
        "per_geometry_ms": {k: v["md_step"] for k, v in res.items()},
        "host_arrays_ms_per_eval": {k: round(v["host_api_ms"], 4) for k, v in res.items()},
        "phases": {k: v["phases"] for k, v in res.items()} if not args.headline_only else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
