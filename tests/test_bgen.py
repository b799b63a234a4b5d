"""BGEN input (FileBgen, src/FileBgen.cpp:11-107, through the reference's
vendored bgen reader).

CPU: the oracle's reader against the reference CLI's own files.  Its
--ld --ld-stats 1 runs write the centred dosage matrix as float32
(.residuals, src/Data.cpp:196-245), which pins every dosage (the 8-bit
byte-pair path's vector loop and table, the any-depth path, layout 1, the
minor-allele swap) and F; the PCA fixtures pin the rest through the printed
eigenvalues and loadings.  Planted mistakes (one dosage rule for every
sample, no swap) must move those files.  The host-only container walk
(pcaone_bgen_info) against the files the simulator writes.

GPU: pcaone_gpu_load_bgen against the oracle, bit for bit where the sums are
exact (F, the minor allele, the FP32 matrix), over layouts, compressions, bit
depths, missing samples, a declared ploidy range, a sample count with more
than 256 residue classes in the minor-allele test, the MAF filter and
--pcangsd's probabilities; and against the reference's .residuals."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import parse_args
from conftest import GOLDEN, load_golden

LD_CASES = ["bgen_ld1_8", "bgen_ld1_16z", "bgen_ld1_5", "bgen_ld1_l1"]


def _write(tmp_path: Path, z, name="x.bgen") -> Path:
    p = tmp_path / name
    p.write_bytes(np.asarray(z["bgen"], dtype=np.uint8).tobytes())
    return p


def _residuals(z):
    raw = np.asarray(z["residuals"], dtype=np.uint8)
    m, n = np.frombuffer(raw[:8].tobytes(), dtype="<u4")
    return np.frombuffer(raw[8:].tobytes(), dtype="<f4").reshape(int(m), int(n))


def _centred_f32(X: np.ndarray) -> np.ndarray:
    """write_residuals's rows: each column minus its mean, cast to float32"""
    return (X - X.mean(axis=0)).T.astype(np.float32)


@pytest.mark.parametrize("case", LD_CASES)
def test_oracle_matrix_matches_reference_residuals(case, tmp_path):
    z = load_golden(GOLDEN / f"{case}.npz")
    R = _residuals(z)
    X, F, minor, keep = O.bgen_read(_write(tmp_path, z), standardize=False)
    assert len(keep) == R.shape[0] and X.shape[0] == R.shape[1]
    # every entry equal; the missing ones are 0 - (a column mean of ~1e-17)
    assert np.abs(_centred_f32(X).astype(np.float64) - R).max() < 1e-15
    assert minor.sum() > 0 and (minor == 0).sum() > 0  # both allele orders occur


def test_oracle_planted_mistakes_move_residuals(tmp_path):
    z = load_golden(GOLDEN / "bgen_ld1_8.npz")
    R = _residuals(z)
    path = _write(tmp_path, z)
    orig_dose, orig_minor = O.bgen_ref_dosage, O.bgen_minor_index

    def table_only(blk, layout, n):  # the 511-entry table for every sample
        d, miss = orig_dose(blk, layout, n)
        q, _, _, _, _, _, pr = O._bgen_header2(blk, n)
        return O.bgen_lut8()[2 * pr[0:2 * n:2].astype(np.uint32) + pr[1:2 * n:2]], miss

    try:
        O.bgen_ref_dosage = table_only
        X1, *_ = O.bgen_read(path, standardize=False)
        O.bgen_ref_dosage = orig_dose
        O.bgen_minor_index = lambda d: 0  # never swap to the minor allele
        X2, *_ = O.bgen_read(path, standardize=False)
    finally:
        O.bgen_ref_dosage, O.bgen_minor_index = orig_dose, orig_minor
    assert np.count_nonzero(np.abs(_centred_f32(X1).astype(np.float64) - R) > 1e-12) > 1000
    assert np.abs(_centred_f32(X2).astype(np.float64) - R).max() > 0.1


def test_lut8_closed_form():
    lut = O.bgen_lut8()
    assert lut.dtype == np.float32 and lut.size == 511
    assert lut[0] == 0 and lut[255] == 1 and lut[510] == 2
    # 7 decimals of i / 255: within half a unit of the 7th decimal, plus float rounding
    i = np.arange(511)
    assert np.abs(lut.astype(np.float64) - i / 255.0).max() <= 0.5e-7 + 1.2e-7


def test_minor_allele_rule_closed_forms():
    # N < 200: one class, freq = total / 200 whatever N is (the library's denominator)
    assert O.bgen_minor_index(np.full(150, 1.4, dtype=np.float32)) == 1  # 210 / 200 > 1: NaN interval
    assert O.bgen_minor_index(np.full(150, 0.2, dtype=np.float32)) == 0
    # N = 1000: ten classes; a clear minor allele stops after the first class
    d = np.zeros(1000, dtype=np.float32)
    d[::10] = 2.0  # class 0 holds every sample with dosage 2: freq 1 after it
    assert O.bgen_minor_index(d) == 1
    d = np.zeros(1000, dtype=np.float32)
    d[1::10] = 2.0  # class 0 is all 0: certain at freq 0
    assert O.bgen_minor_index(d) == 0


@pytest.mark.parametrize("layout,comp,bits", [(2, 0, 8), (2, 1, 3), (2, 2, 32), (1, 0, 16), (1, 1, 16)])
def test_container_walk(tmp_path, layout, comp, bits):
    p = tmp_path / "a.bgen"
    pa.sim_bgen(p, 5, 37, 23, 2, 0.1, miss=0.1, layout=layout, compression=comp, bits=bits,
                sample_ids=layout == 2)
    info = pa.bgen_info(p)
    assert info == {"nsamples": 37, "nvariants": 23, "layout": layout, "compression": comp}
    n, lay, blocks = O.bgen_blocks(p)
    assert (n, lay, len(blocks)) == (37, layout, 23)
    for K, blk in blocks:
        d, miss = O.bgen_ref_dosage(blk, layout, n)
        ok = ~miss
        assert K == 2 and np.all(d[ok] >= -1e-6) and np.all(d[ok] <= 2 + 1e-6)
        if layout == 2:  # missing samples store zero probabilities: dosage 0
            assert np.all(d[miss] == 0)


def test_container_errors(tmp_path):
    p = tmp_path / "a.bgen"
    pa.sim_bgen(p, 5, 40, 10, 2, 0.1)
    b = p.read_bytes()
    bad = tmp_path / "bad.bgen"
    bad.write_bytes(b[:16] + b"bxen" + b[20:])
    with pytest.raises(pa.GpuError):
        pa.bgen_info(bad)
    bad.write_bytes(b[:len(b) - 10])  # the last variant runs past the end
    with pytest.raises(pa.GpuError):
        pa.bgen_info(bad)
    with pytest.raises(pa.GpuError):
        pa.bgen_info(tmp_path / "missing.bgen")


# ---- GPU -------------------------------------------------------------------

FORMATS = [
    # (N, M, layout, compression, bits, declared minimum ploidy, missing rate)
    (230, 300, 2, 1, 8, 2, 0.02),
    (37, 120, 2, 0, 8, 2, 0.05),      # N < 200 and below the 32-sample vector loop
    (150, 200, 2, 2, 16, 1, 0.02),    # the any-depth path with per-sample ploidy bytes
    (205, 150, 2, 0, 5, 2, 0.0),
    (99, 100, 2, 1, 32, 2, 0.03),
    (180, 150, 1, 1, 16, 2, 0.02),
    (64, 90, 1, 0, 16, 2, 0.0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", FORMATS, ids=lambda f: "n{}_l{}_c{}_b{}_p{}".format(*f[:5], f[5]))
@pytest.mark.parametrize("standardize", [True, False])
def test_load_bgen_matches_oracle(gpu, tmp_path, fmt, standardize):
    n, m, layout, comp, bits, pmin, miss = fmt
    p = tmp_path / "a.bgen"
    pa.sim_bgen(p, 11 + n, n, m, 3, 0.1, miss=miss, layout=layout, compression=comp, bits=bits,
                pmin_declared=pmin)
    F, minor = gpu.load_bgen(p, standardize=standardize, nthreads=3)
    X, Fo, mo, keep = O.bgen_read(p, standardize=standardize)
    assert (gpu.N, gpu.M) == (n, len(keep))
    assert np.array_equal(minor, mo)
    got = gpu.dense_rows()
    want = X.T.astype(np.float32)
    if bits <= 16:
        # the dosages are multiples of 2^-39 or coarser and the sums stay below
        # 2^14: every double sum is exact in any order, so F is bit-equal
        assert np.array_equal(F, Fo)
        assert np.array_equal(got, want)
    else:
        # 32-bit values reach 2^-55: the sums round, in another order than the oracle's
        assert np.allclose(F, Fo, rtol=1e-15, atol=0)
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulp[want != 0].max() <= 1 and np.abs(got[want == 0]).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", LD_CASES)
def test_load_bgen_matches_reference_residuals(gpu, tmp_path, case):
    z = load_golden(GOLDEN / f"{case}.npz")
    R = _residuals(z)
    gpu.load_bgen(_write(tmp_path, z), standardize=False)
    got = gpu.dense_rows().astype(np.float64)
    assert got.shape == R.shape
    assert np.abs(got - R).max() < 1e-15


@pytest.mark.gpu
def test_load_bgen_many_residue_classes(gpu, tmp_path):
    """N = 30,011: 300 residue classes in the minor-allele test, more than one
    block's worth of classes, with variants whose test runs through them."""
    n, m = 30011, 24
    p = tmp_path / "big.bgen"
    pa.sim_bgen(p, 77, n, m, 2, 0.02, pmin=0.42, pmax=0.5, miss=0.01)
    F, minor = gpu.load_bgen(p)
    X, Fo, mo, keep = O.bgen_read(p)
    assert np.array_equal(minor, mo) and np.array_equal(F, Fo)
    assert np.array_equal(gpu.dense_rows(), X.T.astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("maf", [0.1, 0.9, 0.3])
def test_load_bgen_maf_filter(gpu, tmp_path, maf):
    n, m = 210, 400
    p = tmp_path / "a.bgen"
    pa.sim_bgen(p, 91, n, m, 3, 0.2, pmin=0.01, pmax=0.5, miss=0.02)
    F, minor = gpu.load_bgen(p, maf=maf)
    X, Fo, mo, keep = O.bgen_read(p, maf=maf)
    assert 0 < len(keep) < m
    assert gpu.M == len(keep) and np.array_equal(F, Fo)
    assert np.array_equal(gpu.dense_rows(), X.T.astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("bits,comp", [(8, 1), (12, 2), (16, 0)])
def test_load_bgen_pcangsd_probabilities(gpu, tmp_path, bits, comp):
    """--pcangsd reads the genotype probabilities as likelihoods; the EM on them
    matches the oracle's PCAngsd on the same two planes."""
    n, m = 160, 500
    p = tmp_path / "g.bgen"
    pa.sim_bgen(p, 123, n, m, 4, 0.1, miss=0.0, compression=comp, bits=bits)
    assert gpu.load_bgen(p, pcangsd=True) is None
    assert (gpu.N, gpu.M) == (n, m)
    P = O.bgen_probs(p)
    P0, P1 = (np.ascontiguousarray(P[:, :, i].T.astype(np.float64)) for i in (0, 1))  # (N, M)
    Fg, it = gpu.gl_em_maf(100, 1e-6)
    Fo, ito = O.em_maf_with_gl(P0, P1, 100, 1e-6)
    assert it == ito and np.allclose(Fg, Fo, rtol=0, atol=1e-12)
    gpu.rsvd_init(3, 7, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64)
    U, S, V, iters, diffs, _, _ = gpu.pcangsd(5, 0.0, 100, 1e-5, grm=False, dc=False)
    orc, _, ito2, dfo, _, _ = O.pcangsd_halko(P0, P1, 3, 7, 5, 0.0)
    assert iters == ito2
    # E is regenerated in FP32 on the device each EM iteration, in FP64 in the oracle
    assert np.max(np.abs(S - orc.S) / orc.S) < 1e-7
    assert O.principal_angles_max(U, orc.U) < 1e-5


@pytest.mark.gpu
def test_load_bgen_pcangsd_refuses_phased_or_mixed_ploidy(gpu, tmp_path):
    p = tmp_path / "g.bgen"
    pa.sim_bgen(p, 3, 50, 20, 2, 0.1, pmin_declared=1)
    with pytest.raises(pa.GpuError):
        gpu.load_bgen(p, pcangsd=True)


@pytest.mark.gpu
def test_iram_on_bgen_matches_reference_cli(gpu, tmp_path):
    z = load_golden(GOLDEN / "bgen_svd0_l1.npz")
    gpu.load_bgen(_write(tmp_path, z))
    gpu.gram_init(pa.PREC_FP64)
    k = z["eigvals"].size
    U, S, V, nconv, nops = gpu.iram(k, tol=1e-6)
    m = int(z["nsnps"])
    assert np.max(np.abs(S ** 2 / m * 2 - z["eigvals"]) / z["eigvals"]) < 5e-6  # 6 printed digits
    assert O.principal_angles_max(U[:, :2], z["eigvecs"][:, :2]) < 1e-4
