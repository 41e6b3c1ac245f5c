"""Host-side checks of bench.py's contract: one GPU only, the committed traffic summary parses,
and the workload names it offers are the synthetic generators'."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from motex_amd import synth  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_summary_parses():
    tr = bench.traffic_from_profile("hd_batched_kernel", "dna1200")
    assert tr is not None
    assert tr["bytes"] == tr["fetch_bytes"] + tr["write_bytes"] > 0
    assert tr["source"].startswith("profiles/")
    assert bench.traffic_from_profile("hd_batched_kernel", "no_such_workload") is None


def test_workload_choices_are_the_generators():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for name in synth.WORKLOADS:
        assert name in r.stdout


def test_roofline_fraction_from_issued_valu_is_a_fraction():
    vp = bench.valu_from_profile("hd_batched_kernel", "dna1200")
    assert vp is not None and vp["valu_insts"] > 0
    # the committed trace's kernel time: issued lane-ops over it must stay under the peak
    frac = vp["valu_insts"] * 64 / 8.5e-3 / 1e12 / bench.VALU_LANE_OPS_PEAK_TOPS
    assert 0 < frac <= 1
    assert bench.valu_from_profile("hd_batched_kernel", "no_such_workload") is None


def test_calibration_against_motexomp_is_recorded():
    for d in (0, 1):
        cal = bench.calibration(d, 16)
        assert cal is not None and cal["ratio_oracle_over_motexomp"] > 0
    assert bench.calibration(0, 1)["threads"] == 1
