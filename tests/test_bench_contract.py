"""The bench's contract pieces that need no GPU: a multi-GPU request is
refused with a message, and the committed rocprofv3 summaries the bench
reads (HBM traffic per launch, VALU issue, FP64 flops) hold entries for
every kernel its JSON line reports."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("kernel", ["query_dcd", "query_ccd", "hessian", "energy", "toi", "build_leaf_levels"])
def test_traffic_profile_covers_the_step_kernels(kernel):
    t = bench.measured_traffic(kernel)
    assert t is not None and t > 0


def test_stream_traffic_profile_covers_the_streams():
    assert bench.measured_traffic("hessian_pt", bench.STREAM_TRAFFIC_PROFILE, grid=1 << 22) > 0


@pytest.mark.parametrize("kernel", ["query_dcd", "query_ccd", "hessian", "toi"])
def test_issue_profile_covers_the_step_kernels(kernel):
    e = bench.measured_issue(kernel)
    assert e is not None
    assert 0.0 < e["valu_busy"] <= 1.0 and 0.0 < e["valu_utilization"] <= 1.0 and e["valu_insts_per_launch"] > 0


@pytest.mark.parametrize("kind", ["ee", "pt"])
def test_fp64_profile_covers_the_toi_streams(kind):
    assert bench.measured_flops64(kind) > 0
