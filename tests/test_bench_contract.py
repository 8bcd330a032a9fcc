"""bench.py's command-line contract and its roofline accounting (no GPU)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused(gpus):
    """--gpus N > 1 ends with exit code 2 and says why, before touching a GPU."""
    proc = _run("--gpus", gpus)
    assert proc.returncode == 2
    assert "one GPU only" in proc.stderr
    assert proc.stdout == ""


def test_workloads_are_listed():
    proc = _run("--help")
    assert proc.returncode == 0
    for name in ("exact_large", "matern", "loocv", "batched", "sparse", "predict",
                 "sparse_predict"):
        assert name in proc.stdout


def _lauum_tile_products(n, tile):
    """Tile products of X^T X over the lower output tiles (i >= j) with X lower
    triangular, so only k >= i contributes: counted one by one."""
    nt = -(-n // tile)
    return sum(1 for i in range(nt) for j in range(i + 1) for k in range(i, nt))


@pytest.mark.parametrize("n,tile", [(128, 128), (256, 128), (1000, 128), (640, 64),
                                    (7, 2)])
def test_lauum_flops_counts_the_lower_tiles(n, tile):
    """The roofline's algorithmic flops: 2 tile^3 per tile product."""
    expect = 2.0 * tile ** 3 * _lauum_tile_products(n, tile)
    assert bench.lauum_flops(n, tile) == pytest.approx(expect, rel=1e-15)


def test_lauum_flops_tends_to_n_cubed_over_three():
    n = 65536
    assert bench.lauum_flops(n) / (n ** 3 / 3.0) == pytest.approx(1.0, abs=0.01)


def test_profiled_traffic_reads_the_committed_summary():
    """The headline's HBM traffic comes from the committed rocprofv3 summary and
    is at least one read of the lower triangle of the float32 inverse factor."""
    n = 65536
    t = bench.profiled_traffic(n)
    assert t is not None
    assert t >= n * n / 2 * 4
    assert bench.profiled_traffic(16384) is None
