"""GPU parity of LD r^2 (the --print-r2 path on a .residuals matrix) through the
C ABI (tests marked `gpu`): against the oracle on the same inputs (the FP64
MFMA band Gram differs from the oracle's FP64 dots only in summation order:
1e-12 absolute on r^2) and against the reference CLI's .ld.gz (6 decimals)."""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def test_ld_r2_matches_reference_cli(gpu):
    z = load_golden(GOLDEN / "ld_r2_bp2000.npz")
    ws, we = pa.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    gpu.load_dense(z["resid"])
    r2 = gpu.ld_r2(ws, we)
    assert r2.shape == z["r2"].shape
    assert np.abs(r2 - z["r2"]).max() <= 5.01e-7
    assert np.abs(r2 - O.ld_r2(z["resid"], ws, we)).max() < 1e-12


@pytest.mark.parametrize("n,m,bp", [(2, 70, 50), (37, 129, 300), (300, 1000, 10 ** 6), (65, 200, 0)])
def test_ld_r2_matches_oracle(gpu, n, m, bp):
    """N = 2 (the smallest with N - 1 > 0), tiles cut by M % 64, one window
    spanning a whole chromosome, and ld_bp = 0 (single-SNP windows: no pairs
    unless positions repeat)."""
    rng = np.random.default_rng(n + m)
    X = (rng.standard_normal((m, n)) + rng.uniform(-2, 2, (m, 1))).astype(np.float32)  # uncentred rows
    pos = np.cumsum(rng.integers(0, 40, m)).astype(np.int32)  # repeated positions included
    chr_end = np.array(sorted(set([m // 3, m // 2, m - 1]))).astype(np.int32)
    ws, we = pa.ld_windows(pos, chr_end, bp)
    gpu.load_dense(X)
    r2 = gpu.ld_r2(ws, we)
    want = O.ld_r2(X, ws, we)
    assert r2.shape == want.shape
    if len(want):
        assert np.abs(r2 - want).max() < 1e-12


def test_ld_r2_errors(gpu):
    X = np.ones((10, 5), dtype=np.float32)
    gpu.load_dense(X)
    with pytest.raises(pa.GpuError):
        gpu.ld_r2(np.array([3, 2]), np.array([2, 2]))  # leads must increase
    with pytest.raises(pa.GpuError):
        gpu.ld_r2(np.array([8]), np.array([5]))  # window past the last SNP
