"""The committed rocprofv3 summaries the bench line reads: each names its
command, covers the kernel and workload the bench reports, and its traffic
block agrees with its own FETCH_SIZE/WRITE_SIZE figures and calibration."""
import importlib.util
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _blocks(name):
    text = open(os.path.join(ROOT, "profiles", name)).read()
    return text, [json.loads(b) for b in re.findall(r"```json\n(.*?)\n```", text, re.S)]


@pytest.mark.parametrize("name,kernel", [("k_wf_extend.md", "k_wf_extend"), ("k_wf_shade.md", "k_wf_shade"),
                                         ("k_render_plain.md", "k_render_plain")])
def test_traffic_summaries_cover_the_bench_kernels(bench, name, kernel):
    text, blocks = _blocks(name)
    assert "rocprofv3 --kernel-trace --stats" in text and "rocprofv3 --pmc FETCH_SIZE" in text
    assert "rocprofv3 --pmc WRITE_SIZE" in text
    b = [x for x in blocks if x.get("kernel") == kernel][0]
    assert b["workload"] == bench.WORKLOAD
    # traffic = calibrated fetch + write, the figures printed above the block
    assert b["traffic_bytes_per_launch"] == pytest.approx(b["fetch_bytes"] + b["write_bytes"], rel=1e-6)
    assert 1.5 < b["fetch_scale"] < 2.5 and 0.5 < b["write_scale"] < 1.5
    assert bench.profiled_traffic(kernel, bench.WORKLOAD) == b["traffic_bytes_per_launch"] > 0


def test_issue_counters_cover_the_closest_hit_and_shading_kernels(bench):
    for kernel in ("k_wf_extend", "k_wf_shade"):
        got = bench.profiled_issue(kernel)
        assert got is not None and got["kernel"].startswith(kernel + "<")
        for k in ("valu_issue_share", "active_lanes", "wait_share", "l2_hit"):
            assert 0.0 < got[k] < 1.0, (kernel, k, got[k])
    assert bench.profiled_issue("k_no_such_kernel") is None


def test_replay_summary_names_its_workload():
    text, blocks = _blocks("k_wfa_shade.md")
    assert "python bench.py --only views" in text
    b = [x for x in blocks if x.get("kernel") == "k_wfa_shade"][0]
    assert b["workload"].startswith("ring16-512x512") and b["traffic_bytes_per_launch"] > 0
