"""GPU parity of LD r^2 (the --print-r2 path on a .residuals matrix) through the
C ABI (tests marked `gpu`): against the oracle on the same inputs (the FP64
MFMA band Gram differs from the oracle's FP64 dots only in summation order:
1e-12 absolute on r^2) and against the reference CLI's .ld.gz (6 decimals);
the --USV path on PLINK input the same way (tests/golden/ld_usv_bp3000.npz)."""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import geno_sums_fast
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


def test_ld_usv_matches_reference_cli(gpu):
    """--USV on PLINK input: (I - U U^T) G in FP64 on the device, r^2 against
    the reference CLI's .ld.gz (6 decimals) and the oracle."""
    z = load_golden(GOLDEN / "ld_usv_bp3000.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    ws, we = pa.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    U = O.read_usv(z["U"])  # the matrix the reference applies (read_usv's column-major map)
    gpu.load_bed(z["payload"], n, m)
    gpu.ld_usv(U)
    r2 = gpu.ld_r2(ws, we)
    assert r2.shape == z["r2"].shape
    assert np.abs(r2 - z["r2"]).max() <= 5.01e-7
    R = O.ld_usv_matrix(O.decode_codes(z["payload"], n, m), U)
    assert np.abs(r2 - O.ld_r2(R, ws, we, centre=False)).max() < 1e-12


@pytest.mark.parametrize("n,m,k", [(3, 70, 1), (65, 130, 5), (257, 300, 33)])
def test_ld_usv_matches_oracle(gpu, n, m, k):
    """Sample and SNP counts across the 64-row and 256-thread tile edges, k on
    each register bound; a random (not orthonormal) U."""
    pay = pa.sim_bed(n + m, n, m, 2, 0.1, 0.05, 0.5, 0.05)
    rng = np.random.default_rng(k)
    U = rng.standard_normal((n, k)) / np.sqrt(n)
    pos = np.cumsum(rng.integers(0, 30, m)).astype(np.int32)
    chr_end = np.array([m // 2, m - 1], dtype=np.int32)
    ws, we = pa.ld_windows(pos, chr_end, 200)
    gpu.load_bed(pay, n, m)
    gpu.ld_usv(U)
    r2 = gpu.ld_r2(ws, we)
    with np.errstate(invalid="ignore", divide="ignore"):
        want = O.ld_r2(O.ld_usv_matrix(O.decode_codes(pay, n, m), U), ws, we, centre=False)
    assert r2.shape == want.shape
    # a monomorphic SNP (likely at N = 3) has a zero row: 0 * inf = NaN there, as
    # in the reference's 0 * (1 / 0)
    assert np.array_equal(np.isnan(r2), np.isnan(want))
    ok = ~np.isnan(want)
    if ok.any():
        assert np.abs(r2[ok] - want[ok]).max() < 1e-10
    # a new load drops the adjusted matrix: ld_r2 on PLINK genotypes alone is an error
    gpu.load_bed(pay, n, m)
    with pytest.raises(pa.GpuError):
        gpu.ld_r2(ws, we)


def test_ld_r2_errors(gpu):
    X = np.ones((10, 5), dtype=np.float32)
    gpu.load_dense(X)
    with pytest.raises(pa.GpuError):
        gpu.ld_r2(np.array([3, 2]), np.array([2, 2]))  # leads must increase
    with pytest.raises(pa.GpuError):
        gpu.ld_r2(np.array([8]), np.array([5]))  # window past the last SNP


# ---- LD straight from the packed genotypes (the --ld --ld-stats 1 route) ----
def _ld_geno_case():
    z = load_golden(GOLDEN / "ld_stats1_bp3000.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    codes = O.decode_codes(z["payload"], n, m)
    ws, we = pa.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    return z, n, m, codes, ws, we


@pytest.mark.parametrize("config", list(range(5)) + [-1])
def test_ld_geno_matches_reference_cli(gpu, config):
    """Every tile configuration (and the autotuner): integer sums bit-exact
    against the oracle's per-pair brute force, r^2 equal to the oracle's FP64
    closed form and within 1e-6 of the reference CLI's .ld.gz (6 printed
    decimals on float32 residuals), NaN where the reference prints nan."""
    z, n, m, codes, ws, we = _ld_geno_case()
    gpu.load_bed(z["payload"], n, m)
    r2, sums, used = gpu.ld_r2_geno(ws, we, sums=True, config=config)
    assert used >= 0 and (config < 0 or used == config)
    want = O.ld_geno_sums(codes, ws, we)
    assert sums.shape == want.shape and np.array_equal(sums.astype(np.int64), want)
    nan = np.isnan(z["r2"])
    assert np.array_equal(np.isnan(r2), nan)
    assert np.abs(r2[~nan] - z["r2"][~nan]).max() <= 1e-6
    ro = O.ld_geno_r2(codes, ws, we, want)
    assert np.abs(r2[~nan] - ro[~nan]).max() < 1e-12


def test_ld_geno_written_as_reference_text(gpu, tmp_path):
    """GPU r^2 through pcaone_write_ld_gz against the reference CLI's .ld.gz:
    every line's six SNP fields equal, NaN lines equal, the printed r^2 within
    one unit of the 6th decimal (the reference's r^2 comes from float32
    residuals, so a printed digit may round the other way)."""
    import gzip

    z, n, m, codes, ws, we = _ld_geno_case()
    gpu.load_bed(z["payload"], n, m)
    r2, _, _ = gpu.ld_r2_geno(ws, we)
    (tmp_path / "g.mbim").write_bytes(z["mbim"].tobytes())
    pa.write_ld_gz(tmp_path / "g.ld.gz", tmp_path / "g.mbim", ws, we, r2)
    got = gzip.decompress((tmp_path / "g.ld.gz").read_bytes()).splitlines()
    ref = gzip.decompress(z["ldgz"].tobytes()).splitlines()
    assert len(got) == len(ref) and got[0] == ref[0]
    same = 0
    for a, b in zip(got[1:], ref[1:]):
        fa, fb = a.split(b"\t"), b.split(b"\t")
        assert fa[:6] == fb[:6]
        if fb[6] == b"-nan":
            assert fa[6] == b"-nan"
        else:
            assert abs(float(fa[6]) - float(fb[6])) <= 1.0000001e-6
        same += a == b
    assert same >= 0.95 * (len(ref) - 1)


@pytest.mark.parametrize("n,m,bp,miss", [(2, 70, 300, 0.0), (3, 129, 200, 0.2), (67, 500, 1000, 0.05),
                                         (301, 1100, 20, 0.01), (1001, 700, 10 ** 6, 0.02)])
def test_ld_geno_sums_bit_exact(gpu, n, m, bp, miss):
    """N below one 64-sample step and N % 4 in {1, 2, 3}; SNP counts off the
    32 / 64 / 128 tile edges; windows of one pair up to whole chromosomes;
    missing rates 0 to 20 %; repeated positions."""
    pay = pa.sim_bed(n * 7 + m, n, m, 2, 0.1, 0.05, 0.5, miss)
    rng = np.random.default_rng(m)
    chr_end = np.array(sorted({m // 3, m - 1}), dtype=np.int32)
    gaps = rng.integers(0, 12, m)
    gaps[chr_end] = np.maximum(gaps[chr_end], 1)  # a chromosome's last position is its own (src/LD.cpp:130)
    pos = np.cumsum(gaps).astype(np.int32)
    ws, we = pa.ld_windows(pos, chr_end, bp)
    gpu.load_bed(pay, n, m)
    codes = O.decode_codes(pay, n, m)
    want = geno_sums_fast(codes, ws, we)
    for cfg in range(5):
        r2, sums, _ = gpu.ld_r2_geno(ws, we, sums=True, config=cfg)
        assert np.array_equal(sums.astype(np.int64), want), cfg
    ro = O.ld_geno_r2(codes, ws, we, want)
    ok = ~np.isnan(ro)
    assert np.array_equal(np.isnan(r2), ~ok)
    if ok.any():
        assert np.abs(r2[ok] - ro[ok]).max() < 1e-12


def test_ld_geno_windows_200_to_3000(gpu):
    """Uneven windows of about 200 to 2,500 SNPs (SNP density alternating 15x
    every 2,500 SNPs along two chromosomes; 7.8M pairs) in one launch: sums
    bit-exact, r^2 to the oracle."""
    n, m = 256, 10000
    pay = pa.sim_bed(77, n, m, 3, 0.1, 0.05, 0.5, 0.005)
    rng = np.random.default_rng(5)
    gaps = np.where((np.arange(m) // 2500) % 2 == 0, rng.integers(1, 20, m), rng.integers(60, 240, m))
    pos = np.cumsum(gaps).astype(np.int32)
    chr_end = np.array([m // 2 - 1, m - 1], dtype=np.int32)
    ws, we = pa.ld_windows(pos, chr_end, 30000)
    assert we.max() >= 2500 and np.percentile(we, 5) >= 190
    gpu.load_bed(pay, n, m)
    r2, sums, cfg = gpu.ld_r2_geno(ws, we, sums=True)
    codes = O.decode_codes(pay, n, m)
    want = geno_sums_fast(codes, ws, we)
    assert np.array_equal(sums.astype(np.int64), want)
    ro = O.ld_geno_r2(codes, ws, we, want)
    ok = ~np.isnan(ro)
    assert np.abs(r2[ok] - ro[ok]).max() < 1e-12
    # r^2 is a squared correlation
    assert np.all((r2[ok] >= 0) & (r2[ok] <= 1 + 1e-12))


def test_ld_geno_errors(gpu):
    X = np.ones((10, 5), dtype=np.float32)
    gpu.load_dense(X)
    with pytest.raises(pa.GpuError):
        gpu.ld_r2_geno(np.array([0]), np.array([3]))  # no genotypes loaded
    pay = pa.sim_bed(1, 20, 50, 1, 0.1, 0.05, 0.5, 0.0)
    gpu.load_bed(pay, 20, 50)
    with pytest.raises(pa.GpuError):
        gpu.ld_r2_geno(np.array([3, 2]), np.array([2, 2]))  # leads must increase
    with pytest.raises(pa.GpuError):
        gpu.ld_r2_geno(np.array([0]), np.array([3]), config=99)  # no such tile shape


# ---- producers and consumers around the LD path (ld_tools_bp3000) ----------
def _ld_tools():
    return load_golden(GOLDEN / "ld_tools_bp3000.npz")


def _read_residuals(path):
    raw = np.fromfile(path, dtype=np.uint32, count=2)
    return raw, np.fromfile(path, dtype=np.float32, offset=8).reshape(int(raw[0]), int(raw[1]))


def test_write_residuals_matches_reference_cli(gpu, tmp_path):
    """--ld after a PLINK PCA (ld-stats 0: G - U S V^T, re-centred) with the
    U, S, V the reference printed (6 significant digits), and ld-stats 1 (the
    centred genotypes) against the reference's file of the ld_stats1 fixture."""
    z = _ld_tools()
    n, m = int(z["nsamples"]), int(z["nsnps"])
    gpu.load_bed(z["payload"], n, m)
    gpu.write_residuals(tmp_path / "r0.residuals", z["U"], z["S"], z["V"], ld_stats=0)
    hdr, R = _read_residuals(tmp_path / "r0.residuals")
    assert list(hdr) == [m, n] and R.shape == z["resid0"].shape
    assert np.abs(R - z["resid0"]).max() < 2e-5  # the printed U, S, V carry 6 digits
    s = load_golden(GOLDEN / "ld_stats1_bp3000.npz")
    n1, m1 = int(s["nsamples"]), int(s["nsnps"])
    gpu.load_bed(s["payload"], n1, m1)
    gpu.write_residuals(tmp_path / "r1.residuals", ld_stats=1)
    hdr1, R1 = _read_residuals(tmp_path / "r1.residuals")
    assert list(hdr1) == [m1, n1]
    assert np.abs(R1 - s["resid"]).max() <= 2.0 ** -23  # float32 rounding of the same FP64 values


def test_write_mbim_matches_reference_cli(gpu, tmp_path):
    z = _ld_tools()
    n, m = int(z["nsamples"]), int(z["nsnps"])
    gpu.load_bed(z["payload"], n, m)
    F, _ = gpu.allele_freq()
    (tmp_path / "x.bim").write_text(str(z["bim"]))
    pa.write_mbim(tmp_path / "x.mbim", tmp_path / "x.bim", F)
    assert (tmp_path / "x.mbim").read_text() == str(z["mbim"])


@pytest.mark.parametrize("tag", ["pm", "pb", "pp"])
def test_ld_prune_matches_reference_cli(gpu, tmp_path, tag):
    """--ld-r2 0.2 on the reference's ld-stats 0 residuals with its .mbim (the
    lower-MAF SNP of a pair goes), with the plain .bim (the partner goes), and
    on the .bed with --USV of the same PCA (its .mbim): .ld.prune.in / .out equal."""
    z = _ld_tools()
    n, m = int(z["nsamples"]), int(z["nsnps"])
    (tmp_path / "x.bim").write_text(str(z["bim"]))
    (tmp_path / "x.mbim").write_text(str(z["mbim"]))
    if tag == "pp":  # --USV prefix: the SNP table is prefix.mbim (src/Cmd.cpp:162)
        gpu.load_bed(z["payload"], n, m)
        gpu.ld_usv(O.read_usv(z["U"]))
        bim = tmp_path / "x.mbim"
    else:
        gpu.load_dense(z["resid0"])
        bim = tmp_path / ("x.mbim" if tag == "pm" else "x.bim")
    gpu.ld_prune(bim, int(z["ld_bp"]), 0.2, tmp_path / "out")
    assert (tmp_path / "out.ld.prune.in").read_text() == str(z[f"prune_{tag}_in"])
    assert (tmp_path / "out.ld.prune.out").read_text() == str(z[f"prune_{tag}_out"])


def test_ld_clump_matches_reference_cli(gpu, tmp_path):
    """--clump (p1 1e-3, p2 0.05, r2 0.1, 2 kb) of an assoc-like file with
    distinct p-values and some SNPs absent, on the ld-stats 0 residuals."""
    z = _ld_tools()
    (tmp_path / "x.mbim").write_text(str(z["mbim"]))
    (tmp_path / "assoc.txt").write_text(str(z["assoc"]))
    gpu.load_dense(z["resid0"])
    gpu.ld_clump(tmp_path / "assoc.txt", tmp_path / "x.mbim", tmp_path / "x.clump", clump_bp=2000, clump_r2=0.1,
                 clump_p1=0.001, clump_p2=0.05)
    assert (tmp_path / "x.clump").read_text() == str(z["clump"])


def test_ld_tools_errors(gpu, tmp_path):
    z = _ld_tools()
    n, m = int(z["nsamples"]), int(z["nsnps"])
    (tmp_path / "short.bim").write_text("1\trs0\t0\t10\tA\tG\n")
    gpu.load_dense(z["resid0"])
    with pytest.raises(pa.GpuError):  # .bim length differs from the LD matrix
        gpu.ld_prune(tmp_path / "short.bim", 3000, 0.2, tmp_path / "o")
    (tmp_path / "a.txt").write_text("CHR\tBP\tPVAL\n1\t10\t0.1\n")
    (tmp_path / "x.mbim").write_text(str(z["mbim"]))
    with pytest.raises(pa.GpuError):  # no P column
        gpu.ld_clump(tmp_path / "a.txt", tmp_path / "x.mbim", tmp_path / "o.clump")
    with pytest.raises(pa.GpuError):  # residuals need PLINK genotypes
        gpu.write_residuals(tmp_path / "r.residuals", ld_stats=1)
    gpu.load_bed(z["payload"], n, m)
    with pytest.raises(pa.GpuError):  # ld-stats 0 needs U, S, V
        gpu._check(gpu._L.pcaone_gpu_write_residuals(gpu.h, str(tmp_path / "r").encode(), None, None, None, 0, 0))
