"""bench.py's readers of the committed profile summaries (host logic): every
traffic line and every BD-rate workload the bench can ask for parses, so the
bench line never dies on a summary it wrote itself."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

LINE = re.compile(r"^`search` workload=(\w+) nref=(\d+) config=(\d+),(\d+),(\d+): bytes_per_launch=(\d+)\s*$")


def test_every_traffic_line_parses():
    with open(bench.TRAFFIC_SUMMARY) as f:
        rows = [LINE.match(ln) for ln in f if ln.startswith("`search` ")]
    assert rows and all(rows), "a traffic line of the summary does not have the form the bench reads"
    for m in rows:
        wl, nref, wpg, wcu, reuse, want = m.groups()
        tune = {"waves_per_group": int(wpg), "waves_per_cu": int(wcu), "reuse_regions": int(reuse)}
        assert bench.measured_traffic(wl, tune, int(nref)) == int(want)


def test_traffic_of_an_unmeasured_configuration_is_none():
    tune = {"waves_per_group": 2, "waves_per_cu": 8, "reuse_regions": 1}
    assert bench.measured_traffic("1080p", tune, 1) is None


@pytest.mark.parametrize("workload,nref", [("1080p", 1), ("batch", 1), ("4k", 4)])
def test_bd_rate_of_every_workload(workload, nref):
    got = bench.measured_bd_rate(5, None, 0.0, workload, nref)
    assert got is not None and got["references"] == nref
    assert set(got["per_size"]) <= set(bench.WORKLOAD_SIZES[workload])
    assert got["max_pct"] == max(max(v["cubic_pct"] + v["pchip_pct"]) for v in got["per_size"].values())
    # another rule than the default is not in the summary
    assert bench.measured_bd_rate(8, None, 0.0, workload, nref) is None
