"""bench.py's kernel-trace summary (--trace-summary) on a hand-made trace:
idle time, time with one and with two kernels running, busy union, the
window over the trace's last seconds, and the per-kernel totals."""
import csv
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for n, s, e in rows:
            w.writerow([n, s, e])


def test_trace_summary_counts_overlap_and_idle(tmp_path):
    import bench
    p = tmp_path / "t.csv"
    # a: 0-100 and 300-400 ns, b: 50-150 ns
    _write(p, [("void k_a<1>(Args)", 0, 100), ("void k_b<2>(X)", 50, 150), ("void k_a<1>(Args)", 300, 400)])
    s = bench.trace_summary(str(p))
    assert s["kernels"] == 3
    assert s["span_s"] == pytest.approx(400e-9)
    assert s["idle_s"] == pytest.approx(150e-9)
    assert s["busy_union_s"] == pytest.approx(250e-9)
    k = s["seconds_with_k_kernels_running"]
    assert k["1"] == pytest.approx(200e-9) and k["2"] == pytest.approx(50e-9)
    assert s["sum_of_kernel_durations_s"] == pytest.approx(300e-9)
    top = {t["kernel"]: t for t in s["top_kernels"]}
    assert top["void k_a<1>"]["calls"] == 2 and top["void k_a<1>"]["total_s"] == pytest.approx(200e-9)


def test_trace_summary_window_keeps_the_last_seconds(tmp_path):
    import bench
    p = tmp_path / "t.csv"
    sec = 1_000_000_000
    # an early kernel 10 s before the last ones: outside a 1 s window
    _write(p, [("void k_old(A)", 0, 100), ("void k_new(A)", 10 * sec, 10 * sec + 500),
               ("void k_new(A)", 10 * sec + 600, 10 * sec + 700)])
    s = bench.trace_summary(str(p), window_s=1.0)
    assert s["kernels"] == 2
    assert s["idle_s"] == pytest.approx(100e-9)
