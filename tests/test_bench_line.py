"""The bench line's figures taken from committed rocprofv3 summaries, and its
one-GPU guard (CPU only: nothing here launches a kernel)."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _field(key):
    with open(bench.TRAFFIC_SUMMARY) as f:
        return float(re.search(rf"^{key}:\s*([0-9.eE+]+)", f.read(), re.M).group(1))


def test_traffic_scales_with_steps_per_launch():
    measured = _field("bytes_per_launch")
    steps = int(_field("steps_per_launch"))
    b, s = bench.traffic_from_summary(steps)
    assert s == steps and b == measured
    b2, _ = bench.traffic_from_summary(2 * steps)
    assert b2 == 2 * measured


def test_issue_figures_from_the_counter_summary():
    issue = bench.issue_from_summary()
    assert issue is not None
    assert issue["valu_per_wave_pair"] == _field("valu_per_wave_pair")
    assert 0.0 < issue["trans_share"] < 1.0
    assert 0.0 < issue["valu_issue_per_quad_cycle"] <= issue["issue_peak_per_quad_cycle"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
