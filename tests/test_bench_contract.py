"""bench.py's host side, without a GPU: more than one GPU is refused with a message,
the committed rocprofv3 --pmc summary feeds `roofline.traffic` only for the traversal
variant it was measured with, and the per-launch byte model reproduces the figure the
profile summary quotes from the counts it quotes."""
import json
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

PMC_MD = os.path.join(ROOT, "profiles", "k_wf_trace_displaced_sphere_pmc.md")
PMC_JSON = os.path.join(ROOT, "profiles", "k_wf_trace_displaced_sphere_pmc.json")


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_only_for_the_profiled_variant():
    d = json.load(open(PMC_JSON))
    v = d["traversal_variant"]
    t = bench.measured_traffic("k_wf_trace_displaced_sphere", v)
    # the markdown summary states the scaled total in MB (written by hand beside the JSON)
    mb = float(re.search(r"\*\*([0-9.]+) MB of HBM traffic per launch\*\*", open(PMC_MD).read()).group(1))
    assert t == pytest.approx(mb * 1e6, rel=2e-3)
    for other in range(4):
        if other != v:
            assert bench.measured_traffic("k_wf_trace_displaced_sphere", other) is None
    assert bench.measured_traffic("k_wf_trace_displaced_sphere", None) is None
    assert bench.measured_traffic("no_such_kernel", v) is None


def test_trace_bytes_match_the_quoted_counts():
    # the counting frame of the 1080p / 64 spp workload (SAH-collapsed BVH): 2,291,441,269
    # node visits, 360,892,479 triangle tests, 132,710,400 camera and 40,813,025 extension
    # rays, 40,433,654 shadow rays, no instance entries (flattened), 248 iterations ->
    # 727,137,615 B per launch, the `roofline.algorithmic_bytes` those counts gave
    st = {"node_visits": 2291441269, "tri_tests": 360892479, "instance_visits": 0, "primary_rays": 132710400,
          "extension_rays": 40813025, "shadow_rays": 40433654}
    per_launch = bench.trace_bytes(st) / 248
    assert int(per_launch) == 727137615
    # by hand: 64 B per node, 48 B per triangle test, 84 B per camera or extension ray,
    # 44 B per shadow ray
    assert per_launch * 248 == (2291441269 * 64 + 360892479 * 48 + (132710400 + 40813025) * 84
                                + 40433654 * 44)
    # units: one node is 64 B, one triangle 48 B, an extension ray 84 B, a shadow ray 44 B
    assert bench.trace_bytes({"node_visits": 1, "tri_tests": 0, "instance_visits": 0, "primary_rays": 0,
                              "extension_rays": 0, "shadow_rays": 0}) == 64
    assert bench.trace_bytes({"node_visits": 0, "tri_tests": 1, "instance_visits": 0, "primary_rays": 1,
                              "extension_rays": 0, "shadow_rays": 1}) == 48 + 84 + 44
