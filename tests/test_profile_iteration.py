"""The kernel-trace reader of the iteration profiler, on a synthetic trace
(two capped solves: k_mix, iterations of two applications, k_mix)."""
import csv
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "topo-dda-forwin64_amd"))
import profile_iteration as PI  # noqa: E402

FIELDS = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]


def write_trace(path):
    rows, t = [], 1000

    def k(name, dur, stream_overlap=0):
        nonlocal t
        start = t - stream_overlap
        rows.append({"Kind": "KERNEL_DISPATCH", "Kernel_Name": name, "Start_Timestamp": start,
                     "End_Timestamp": start + dur})
        t = max(t, start + dur)

    for solve in range(2):
        k("void tdda::k_mix<1>(int)", 500)
        t += 100  # idle
        for it in range(3):
            for mode in range(2):
                # two launch groups of k_pzy on two streams (overlapping), then x, then yz
                k("void tdda::k_pzy<float, 128>(int)", 900 if mode == 0 else 600)
                k("void tdda::k_pzy<float, 96>(int)", 400, stream_overlap=400)
                k("void tdda::k_pxr<float, 108, 256>(int)", 800)
                k("void tdda::k_pyz<float, 128>(int)", 700)
                k("void tdda::k_finish<float, 1>(int)", 50)
                t += 10  # launch gap
        k("void tdda::k_mix<2>(int)", 500)
        t += 5000
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        w.writerows(rows)


def test_spans_of_synthetic_trace(tmp_path, capsys):
    p = tmp_path / "trace.csv"
    write_trace(p)
    out = PI.spans(str(p))
    # loop: 3 iterations x 2 applications; each application 900|600 + 800 + 700 + 50 busy, then a gap of 10
    busy_ns = 3 * ((900 + 800 + 700 + 50) + (600 + 800 + 700 + 50))
    idle_ns = 5 * 10  # the loop runs from its first launch to its last: five of the six launch gaps
    assert out["busy_ms"] == pytest.approx(busy_ns / 1e6)
    assert out["idle_ms"] == pytest.approx(idle_ns / 1e6)
    assert out["loop_ms"] == pytest.approx((busy_ns + idle_ns) / 1e6)
    assert out["span_us"]["k_pzy"] == pytest.approx((0.9, 0.6))
    assert out["span_us"]["k_pxr"] == pytest.approx((0.8, 0.8))
    assert out["span_us"]["k_pyz"] == pytest.approx((0.7, 0.7))
    assert "k_pzy" in capsys.readouterr().out


def test_spans_needs_two_solves(tmp_path):
    p = tmp_path / "one.csv"
    with open(p, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        w.writerow({"Kind": "KERNEL_DISPATCH", "Kernel_Name": "void tdda::k_mix<1>(int)", "Start_Timestamp": 1,
                    "End_Timestamp": 2})
    with pytest.raises(SystemExit):
        PI.spans(str(p))
