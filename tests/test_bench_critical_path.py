"""bench.py's critical-path arithmetic on a synthetic trace (CPU only).

The bench reports the roofline of the longest kernel on the timed step's critical path and the
whole step's algorithmic bytes over its time; these checks pin which branch is on the path, the
dominant kernel's name and the byte sums, with numbers worked out by hand below.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test",
                                                  os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


PRM = {"index_bytes": 2, "order": 6, "pair_mode": 0, "grid": [128, 128, 128], "fft_planes": 1}
TUNING = {"fft_wave": 2, "fft_lanes": (8, 8, 4), "spread_direct": 2, "pair_cosched": 0}
# microseconds per span (the register plane transforms carry no separate y passes)
US = {"spread_region": 45.0, "spread_assemble": 4.0, "fft_x_r2c": 23.0, "fft_y": 0.0,
      "fft_z_influence": 14.0, "fft_y_inverse": 0.0, "fft_x_c2r": 18.0, "gather_mesh": 19.0,
      "join_wait": 9.0, "gather_combine": 7.5, "charges": 5.0, "pair_sum": 103.0,
      "evaluation": 182.0}
BYTES = {"spread_region": 18_000_000, "spread_assemble": 0, "fft_x_r2c": 17_000_000,
         "fft_y": 0, "fft_z_influence": 21_000_000, "fft_y_inverse": 0,
         "fft_x_c2r": 17_000_000, "gather_mesh": 18_000_000, "join_wait": 0,
         "gather_combine": 4_000_000, "charges": 2_000_000, "pair_sum": 258_000_000}


def _trace(us, calls):
    # ew_trace_times reports milliseconds summed over the traced evaluations
    return {k: v * calls * 1e-3 for k, v in us.items()}


def test_chain_on_the_path():
    b = _bench()
    calls = 10
    crit = b.critical_path(_trace(US, calls), calls, BYTES, "mixed", PRM, 0.14, TUNING,
                           split_gather=True)
    # chain before the join: 45 + 4 + 23 + 14 + 18 + 19 = 123 us > charges + pair sum = 108 us
    assert crit["pair_sum_on_critical_path"] is False
    assert crit["dominant"]["span"] == "spread_region"
    assert crit["dominant"]["kernel"] == "void ewamd::k_spread_owned<float, 6, false>"
    assert crit["dominant"]["us"] == pytest.approx(45.0)
    # 18 MB in 45 us = 400 GB/s = 0.05 of 8 TB/s
    assert crit["dominant"]["frac"] == pytest.approx(0.05)
    spans = {r["span"] for r in crit["spans"]}
    assert "fft_y" not in spans and "fft_y_inverse" not in spans  # fused into the planes
    # chain spans: 45 + 4 + 23 + 14 + 18 + 19 + 9 + 7.5 = 139.5 us, 95 MB
    assert crit["chain_us"] == pytest.approx(139.5)
    assert crit["chain_bytes"] == 95_000_000
    # every kernel of the evaluation: 95 + 2 + 258 = 355 MB over the 0.14 ms step
    assert crit["step_bytes"] == 355_000_000
    assert crit["step_frac"] == pytest.approx(355e6 / 0.14e-3 / 8e12, abs=1e-4)


def test_pair_branch_on_the_path():
    b = _bench()
    calls = 4
    us = dict(US, pair_sum=130.0)  # 5 + 130 = 135 us >= 123 us
    crit = b.critical_path(_trace(us, calls), calls, BYTES, "mixed", PRM, 0.15, TUNING,
                           split_gather=True)
    assert crit["pair_sum_on_critical_path"] is True
    assert crit["dominant"]["span"] == "pair_sum"
    assert crit["dominant"]["kernel"] == "void ewamd::k_pair_spmv_stream<float, short>"
    # spans without a kernel (the join wait, the owned blocks' empty assembly) never dominate
    assert all(r["kernel"] is not None or r["span"] != crit["dominant"]["span"]
               for r in crit["spans"])
