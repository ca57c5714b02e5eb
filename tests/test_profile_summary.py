"""The window mode of the rocprof summary script, on a synthetic kernel trace:
a batch pass of 4 launch groups repeated (in sequence, then back to back),
followed by single-operator launches of one of the same kernels."""
import csv
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "topo-dda-forwin64_amd"))
import profile_summary as PS  # noqa: E402

FIELDS = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]
LENGTHS = (36, 40, 45, 50)


def write_trace(path):
    rows = []
    t = 0

    def rep(span):
        # four groups on four streams: all start together, the longest sets the span
        nonlocal t
        for k, L in enumerate(LENGTHS):
            rows.append({"Kind": "KERNEL_DISPATCH", "Kernel_Name": f"void tdda::k_pxr<float, {L}, 256>(int)",
                         "Start_Timestamp": t + k, "End_Timestamp": t + span - (3 - k)})
        t += span

    for _ in range(2):      # in sequence: 2 repetitions, 1000 ns each, other passes between
        rep(1000)
        t += 5000
    for _ in range(3):      # back to back: 3 repetitions of 800 ns
        rep(800)
    for k in range(5):      # one operator's launches: one kernel
        rows.append({"Kind": "KERNEL_DISPATCH", "Kernel_Name": "void tdda::k_pxr<float, 40, 512>(int)",
                     "Start_Timestamp": t, "End_Timestamp": t + 50})
        t += 60
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        w.writerows(rows)


def test_window_finds_the_batch_runs(tmp_path, capsys):
    p = tmp_path / "trace.csv"
    write_trace(p)
    spans = PS.window(str(p), "k_pxr<float", 3, 4, 0, distinct=True)
    assert spans == pytest.approx([0.8, 0.8, 0.8])
    spans = PS.window(str(p), "k_pxr<float", 2, 4, 12, distinct=True)  # skip the back-to-back runs
    assert spans == pytest.approx([1.0, 1.0])
    out = capsys.readouterr().out
    assert "first start to last end / runs 0.80 us" in out


def test_window_without_distinct_takes_the_last_dispatches(tmp_path):
    p = tmp_path / "trace.csv"
    write_trace(p)
    spans = PS.window(str(p), "k_pxr<float", 1, 5, 0)
    assert spans == pytest.approx([(4 * 60 + 50) / 1e3])
    with pytest.raises(SystemExit):
        PS.window(str(p), "k_pxr<float", 10, 4, 0, distinct=True)
