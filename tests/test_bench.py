"""bench.py's contract on the CPU: one GPU only, and the committed rocprofv3
summaries it reads HBM traffic from parse to per-launch byte counts near
the algorithmic bytes of each kernel."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

from conftest import ROOT


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and p.stdout == ""


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


@pytest.mark.parametrize("key,summary,compulsory", [
    ("f64_k1", "pmc_traffic.txt", 2 * 8 * 512 ** 3),       # read t, write t+1
    ("f64_k8", "pmc_traffic.txt", 2 * 8 * 512 ** 3),       # read t, write t+8
    ("f32_k8", "pmc_traffic.txt", 2 * 4 * 512 ** 3),
    ("lap2d_8000", "filter_march_kernel_laplacian_stats.md", 16 * 8000 ** 2),
    ("lap3d_512", "filter_march_kernel_laplacian_stats.md", 16 * 512 ** 3),
])
def test_committed_traffic_parses(key, summary, compulsory):
    t = _bench().pmc_traffic(key, summary)
    assert t is not None, key
    # measured HBM bytes per launch: at least the compulsory bytes (less a
    # counter rounding margin), and within 30 % above them
    assert 0.99 * compulsory <= t <= 1.3 * compulsory, (key, t / compulsory)


def test_unknown_traffic_key_is_none():
    assert _bench().pmc_traffic("no_such_kernel") is None


@pytest.mark.parametrize("key,k", [("f64_k6", 6), ("f64_k8", 8), ("f32_k8", 8), ("bf16_k6", 6)])
def test_committed_valu_counts_parse(key, k):
    """SQ_INSTS_VALU per launch of the fused kernels: 512^3 cells x k
    levels, 9-12 VALU instructions per 64-lane cell-level once the tile
    halo (about 1.2-1.3x) is counted, so 10-20 per output cell-update."""
    v = _bench().pmc_traffic("valu_" + key)
    assert v is not None
    per = v * 64 / (512 ** 3 * k)
    assert 10 <= per <= 20, per
