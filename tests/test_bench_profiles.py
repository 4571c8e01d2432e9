"""bench.py reads the committed rocprofv3 traffic summaries (profiles/pmc_traffic.json) for its
roofline objects: every entry must parse, be used for its own workload sizes, and be dropped for
other sizes (CPU only: no device call)."""
import json
import os

import pytest

import bench

TRAFFIC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles", "pmc_traffic.json")


def _entries():
    with open(TRAFFIC) as f:
        return {k: v for k, v in json.load(f).items() if isinstance(v, dict)}


@pytest.mark.parametrize("entry", sorted(_entries()))
def test_entry_has_workload_commit_and_bytes(entry):
    e = _entries()[entry]
    assert e["workload"]["particles"] > 0
    assert isinstance(e.get("commit"), str) and e["commit"]
    assert e["bytes"] > 0
    assert e.get("commands"), "the summary names the commands that made it"


@pytest.mark.parametrize("entry", sorted(k for k in _entries() if k != "build_lines"))
def test_kernel_traffic_used_only_for_its_workload(entry):
    wl = _entries()[entry]["workload"]
    b, src = bench.measured_traffic(entry, n=wl["particles"], P=wl.get("samples"))
    assert src["used"] and b == _entries()[entry]["bytes"]
    b2, src2 = bench.measured_traffic(entry, n=wl["particles"] + 1, P=wl.get("samples"))
    assert b2 is None and not src2["used"]


def test_build_roofline_sums_the_build_kernels():
    e = _entries()["build_lines"]
    n = e["workload"]["particles"]
    r = bench.build_roofline(n, 2, 5.0, "lines")
    per_kernel = sum(v["fetch_bytes"] + v["write_bytes"] for v in e["kernels"].values())
    assert r["traffic"] == per_kernel
    assert abs(per_kernel - e["bytes"]) <= 1e-6 * e["bytes"]
    assert r["traffic_source"]["commit"] == e["commit"]
    assert r["measured_GBs"] == pytest.approx(per_kernel / 5e-3 / 1e9)
    # another particle count: no measured traffic
    assert bench.build_roofline(n + 1, 2, 5.0, "lines")["traffic"] is None
