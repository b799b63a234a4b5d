"""The allocation-free LD entry point (pcaone_gpu_ld_r2_geno_noalloc) against
the allocating one (pcaone_gpu_ld_r2_geno) through the C ABI: the same pairs,
bit for bit, on two calls in a row (the second with fewer windows), its size
limit and its precondition."""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def test_noalloc_matches_allocating_entry_point(gpu):
    z = load_golden(GOLDEN / "ld_stats1_bp3000.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    ws, we = pa.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    gpu.load_bed(z["payload"], n, m)
    with pytest.raises(pa.GpuError):  # no LD layout yet
        gpu.ld_r2_geno_noalloc(ws, we)
    r2, sums, cfg = gpu.ld_r2_geno(ws, we, sums=True)
    # first call: every window
    r2n, sumsn = gpu.ld_r2_geno_noalloc(ws, we, sums=True, config=cfg)
    assert np.array_equal(sumsn, sums)
    assert np.array_equal(np.isnan(r2n), np.isnan(r2))
    ok = ~np.isnan(r2)
    assert np.array_equal(r2n[ok], r2[ok])
    # second call, smaller: windows 100 .. 399 only (their pairs of the full result)
    off = np.concatenate([[0], np.cumsum(we - 1)])
    w0, w1 = 100, 400
    r2s, sumss = gpu.ld_r2_geno_noalloc(ws[w0:w1], we[w0:w1], sums=True, config=cfg)
    sl = slice(off[w0], off[w1])
    assert np.array_equal(sumss, sums[sl])
    ok = ~np.isnan(r2[sl])
    assert np.array_equal(np.isnan(r2s), ~ok) and np.array_equal(r2s[ok], r2[sl][ok])
    # every tile shape gives the same sums without allocating
    for c in range(5):
        _, s_c = gpu.ld_r2_geno_noalloc(ws[w0:w1], we[w0:w1], sums=True, config=c)
        assert np.array_equal(s_c, sums[sl])


def test_noalloc_refuses_windows_beyond_its_scratch(gpu):
    n, m = 40, 6000
    pay = pa.sim_bed(3, n, m, 1, 0.1, 0.05, 0.5, 0.0)
    gpu.load_bed(pay, n, m)
    pos = np.arange(m, dtype=np.int32) * 10
    ws, we = pa.ld_windows(pos, np.array([m - 1], dtype=np.int32), 10 ** 6)  # 18M pairs: 144 MB of r^2
    small_ws, small_we = ws[:5], np.minimum(we[:5], 50)
    gpu.ld_r2_geno(small_ws, small_we)  # builds the layout
    with pytest.raises(pa.GpuError):
        gpu.ld_r2_geno_noalloc(ws, we)
    r2, _ = gpu.ld_r2_geno_noalloc(small_ws, small_we)
    assert r2.shape == (int((small_we - 1).sum()),)
