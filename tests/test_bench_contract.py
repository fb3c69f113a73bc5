"""CPU checks of the benchmark's contract: the committed traffic summary it reads, its
algorithmic byte count and the refusal of multi-GPU runs (no GPU needed)."""

import os
import subprocess
import sys

import bench


def test_algorithmic_bytes_per_correlation():
    # one complex64 half spectrum (H x (W/2+1)) written and read back
    assert bench.algorithmic_bytes_per_correlation(4096, 4096) == 2 * 8 * 4096 * 2049 == 134_283_264


def test_committed_traffic_summary_is_readable_and_plausible():
    step = bench.read_traffic(bench.TRAFFIC_SUMMARY)
    colpass = bench.read_traffic(bench.TRAFFIC_SUMMARY, "colpass_bytes_per_launch")
    algorithmic = bench.algorithmic_bytes_per_correlation(4096, 4096) * 8 * 13
    assert step is not None and colpass is not None
    # measured HBM bytes cannot be below what the step must move, and the column pass is
    # part of the step
    assert algorithmic <= step < 3 * algorithmic
    assert colpass < step
    assert bench.read_traffic(bench.TRAFFIC_SUMMARY, "no_such_key") is None
    assert bench.read_traffic(os.path.join(os.path.dirname(bench.TRAFFIC_SUMMARY), "missing.md")) is None


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2
    assert "one GPU only" in res.stderr
    assert res.stdout == ""
