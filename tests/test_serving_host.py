"""CPU tests of the serving driver's host-side summary (zeta_amd/serving.py):
latency percentiles over completed requests only, with cancelled and
never-admitted requests counted apart."""
import numpy as np

from zeta_amd import serving


def test_summary_counts_and_percentiles():
    out = np.zeros(5, dtype=[("score", "<i4"), ("nodes", "<u8"), ("depth", "<i4"), ("status", "<i4")])
    out["status"] = [0, 0, 0, -7, -100]  # done, done, done, cancelled, never admitted
    out["nodes"] = [10, 20, 30, 5, 0]
    depths = np.array([4, 4, 5, 6, 7])
    lat = np.array([100.0, 200.0, 300.0, np.nan, np.nan])
    s = serving.summarize(out, lat, 2.0, depths, 70)
    assert (s["requests"], s["completed"], s["cancelled"], s["not_admitted"]) == (5, 3, 1, 1)
    assert s["p50_ms"] == 200.0 and abs(s["p99_ms"] - 298.0) < 1e-9
    assert s["nodes"] == 70 and s["nodes_per_s"] == 35.0 and s["request_nodes"] == 65
    assert set(s["by_depth"]) == {4, 5}  # depths with no completed request are left out
    assert s["by_depth"][4]["n"] == 2 and s["by_depth"][4]["p50_ms"] == 150.0
    assert s["by_depth"][5] == {"n": 1, "p50_ms": 300.0, "p99_ms": 300.0}


def test_summary_without_completions():
    out = np.zeros(2, dtype=[("score", "<i4"), ("nodes", "<u8"), ("depth", "<i4"), ("status", "<i4")])
    out["status"] = [-7, -100]
    s = serving.summarize(out, np.array([np.nan, np.nan]), 1.0, np.array([8, 9]), 0)
    assert s["completed"] == 0 and s["p50_ms"] is None and s["p99_ms"] is None and s["by_depth"] == {}
