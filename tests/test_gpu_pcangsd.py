"""GPU parity of the PCAngsd path on genotype likelihoods (tests marked `gpu`):
emMAF_with_GL, the initial E, the fit_with_pi PCAngsd update and
pcangsd_standardize_E against the oracle on the same inputs, and the whole EM
(plus the GRM written to .cov) against the reference CLI's BEAGLE runs with
RSVD (tests/golden/gl_pcangsd_svd1.npz) and with IRAM
(tests/golden/gl_pcangsd_svd0.npz); --pcangsd on PLINK hard calls against the
reference CLI's in-core PLINK run (tests/golden/plink_pcangsd_svd1.npz).

Tolerances: F and Dc within FP64 rounding (sums in a different order); E is
stored FP32, so within 2^-23 of the largest entry; eigenvalues and the GRM as
for the RSVD arms (the reference prints 6 significant digits)."""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import eig_relerr, gl_planes, parse_args
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

ARMS = {"fp64": pa.PREC_FP64, "fp32": pa.PREC_FP32, "bf16x2": pa.PREC_SPLIT_BF16}
EIG_TOL = {"fp64": 5e-6, "fp32": 1e-5, "bf16x2": 1e-5}


def _gl(seed=11, n=77, m=513):
    return pa.sim_gl(seed, n, m, 3, 0.1, 0.05, 0.5, 2.0, 0.01)


@pytest.mark.parametrize("per_site", [3, 2])
def test_em_maf_matches_oracle(gpu, per_site):
    gl = _gl()
    gpu.load_gl(gl[:, :, :per_site])
    F, it = gpu.gl_em_maf(100, 1e-6)
    P0, P1 = gl_planes(gl)
    Fo, ito = O.em_maf_with_gl(P0, P1, 100, 1e-6)
    assert it == ito
    assert np.allclose(F, Fo, rtol=0, atol=1e-12)


@pytest.mark.parametrize("variant", [pa.RSVD_NORMAL, pa.RSVD_FANCY])
def test_pcangsd_update_matches_oracle(gpu, variant):
    gl = _gl(12, 70, 600)
    n, m, k = 70, 600, 3
    gpu.load_gl(gl)
    F, _ = gpu.gl_em_maf(100, 1e-6)
    gpu.rsvd_init(k, 7, variant=variant, bands=4, precision=pa.PREC_FP64)
    P0, P1 = gl_planes(gl)
    rng = np.random.default_rng(5)
    U = np.linalg.qr(rng.standard_normal((n, k)))[0]
    V = np.linalg.qr(rng.standard_normal((m, k)))[0]
    S = np.array([30.0, 20.0, 10.0])
    # S = 0: the prior falls back to F clamped to [1e-4, 1 - 1e-4], which is the
    # initial E wherever F lies inside that range
    E0, _ = gpu.pcangsd_update(U, np.zeros(k), V)
    want0 = O.pcangsd_E(P0, P1, F, U, np.zeros(k), V).T
    assert np.abs(E0 - want0).max() < 2.0 ** -22 * np.abs(want0).max()
    inside = (F >= 1e-4) & (F <= 1 - 1e-4)
    init = O.gl_initial_E(P0, P1, F).T
    assert np.abs(E0[inside] - init[inside]).max() < 2.0 ** -22 * np.abs(init).max()
    E, _ = gpu.pcangsd_update(U, S, V)
    want = O.pcangsd_E(P0, P1, F, U, S, V).T
    assert np.abs(E - want).max() < 2.0 ** -22 * np.abs(want).max()
    Es, Dc = gpu.pcangsd_update(U, S, V, standardize=True)
    wantE, wantDc = O.pcangsd_standardize(P0, P1, F, U, S, V)
    assert np.abs(Es - wantE.T).max() < 2.0 ** -22 * np.abs(wantE).max()
    assert np.allclose(Dc, wantDc, rtol=1e-12, atol=0)


@pytest.mark.parametrize("arm", list(ARMS))
def test_pcangsd_matches_reference_cli(gpu, arm):
    z = load_golden(GOLDEN / "gl_pcangsd_svd1.npz")
    d = parse_args(z["args"])
    m = int(z["nsnps"])
    gpu.load_gl(z["gl"])
    gpu.gl_em_maf(100, 1e-6)
    gpu.rsvd_init(d["k"], d["os"], variant=d["svd"], precision=ARMS[arm], seed=d["seed"], rand=d["rand"])
    U, S, V, it, diffs, Dc, Cm = gpu.pcangsd(d["p"], d["tol"], 100, 1e-5)
    ev = S ** 2 / m * 2
    assert eig_relerr(ev, z["eigvals"]) < EIG_TOL[arm]
    assert O.principal_angles_max(U, z["eigvecs"]) < 1e-4
    assert O.principal_angles_max(V, z["loadings"]) < 1e-4
    # the GRM of the reference's .cov, 6 significant digits
    assert np.max(np.abs(Cm - z["cov"])) < 2e-5 * np.max(np.abs(z["cov"]))
    # EM trajectory against the oracle
    P0, P1 = gl_planes(z["gl"])
    op, F, ito, dfo, Dco, Co = O.pcangsd_halko(P0, P1, d["k"], d["os"], d["p"], d["tol"])
    assert it == ito
    assert np.allclose(Dc, Dco, rtol=1e-6)
    if arm == "fp64":
        assert np.allclose(diffs, dfo, rtol=1e-6, atol=1e-12)
        assert eig_relerr(ev, op.eigvals()) < 1e-9


@pytest.mark.parametrize("arm", list(ARMS))
def test_grm_eigvecs2_matches_reference_cli(gpu, arm, tmp_path):
    """The GRM eigendecomposition on the device (pcaone_gpu_grm_eig) and the
    .eigvecs2 writer against the reference CLI's file: an orthonormal
    eigenbasis of the device's own C to FP64 rounding, |lambda| descending and
    equal to numpy's eigenvalues of that C, the leading well-separated vectors
    equal to the reference's up to sign, and the written file's header and
    sample names equal to the reference's."""
    z = load_golden(GOLDEN / "gl_pcangsd_svd1.npz")
    d = parse_args(z["args"])
    gpu.load_gl(z["gl"])
    gpu.gl_em_maf(100, 1e-6)
    gpu.rsvd_init(d["k"], d["os"], variant=d["svd"], precision=ARMS[arm], seed=d["seed"], rand=d["rand"])
    _, _, _, _, _, _, Cm = gpu.pcangsd(d["p"], d["tol"], 100, 1e-5)
    ev, U2 = gpu.grm_eig()
    n = Cm.shape[0]
    assert np.abs(U2.T @ U2 - np.eye(n)).max() < 1e-12
    lam = np.linalg.eigvalsh(Cm)
    want = np.sort(np.abs(lam))[::-1]
    assert np.abs(ev - want).max() < 1e-12 * want[0]
    R = Cm @ U2
    sgn = np.sign(np.einsum("ij,ij->j", U2, R))
    assert np.abs(R - U2 * (sgn * ev)).max() < 1e-12 * want[0]
    text = str(z["text_eigvecs2"])
    gold = np.array([[float(x) for x in ln.split("\t")[2].split()] for ln in text.splitlines()[1:]])
    for j in range(5):
        s = np.sign(U2[:, j] @ gold[:, j])
        assert np.abs(U2[:, j] - s * gold[:, j]).max() < 1e-4
    (tmp_path / "g.beagle.gz").write_bytes(z["beagle_gz"].tobytes())
    pa.write_eigvecs2(tmp_path / "g.eigvecs2", tmp_path / "g.beagle.gz", U2)
    got = (tmp_path / "g.eigvecs2").read_text().splitlines()
    ref = text.splitlines()
    assert got[0] == ref[0] and len(got) == len(ref)
    assert [ln.split("\t")[:2] for ln in got[1:]] == [ln.split("\t")[:2] for ln in ref[1:]]


def test_grm_eig_needs_a_grm(gpu):
    gpu.load_gl(_gl(3, 20, 100))
    with pytest.raises(pa.GpuError):
        gpu.grm_eig()


@pytest.mark.parametrize("arm", list(ARMS))
def test_pcangsd_plink_matches_reference_cli(gpu, arm):
    """--pcangsd on in-core PLINK input: hard-call likelihoods, genotype-count
    F, centred-genotype initial E, then the EM (no GRM step)."""
    z = load_golden(GOLDEN / "plink_pcangsd_svd1.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    gpu.load_bed(z["payload"], n, m)
    F = gpu.gl_from_bed()
    codes = O.decode_codes(z["payload"], n, m)
    Fo, _ = O.allele_freq(codes)
    assert np.array_equal(F, Fo)
    gpu.rsvd_init(d["k"], d["os"], variant=d["svd"], precision=ARMS[arm], seed=d["seed"], rand=d["rand"])
    U, S, V, it, diffs, Dc, Cm = gpu.pcangsd(d["p"], d["tol"], 100, 1e-5, grm=False, dc=False)
    assert Dc is None and Cm is None
    ev = S ** 2 / m * 2
    assert eig_relerr(ev, z["eigvals"]) < max(5e-6, EIG_TOL[arm])
    assert O.principal_angles_max(U, z["eigvecs"]) < 1e-4
    assert O.principal_angles_max(V, z["loadings"]) < 1e-4
    op, _, ito, dfo = O.pcangsd_plink(codes, d["k"], d["os"], d["p"], d["tol"])
    assert it == ito
    if arm == "fp64":
        # the device holds the missing entries' 0.333333 as FP32 (1.4e-8 off)
        # and E as FP32: the trajectory agrees to about 1e-8
        assert np.allclose(diffs, dfo, rtol=1e-4, atol=1e-8)
        assert np.allclose(diffs, z["em_diffs"], rtol=1e-4, atol=1e-8)  # the reference's log
        assert eig_relerr(ev, op.eigvals()) < 1e-7


def test_pcangsd_plink_initial_E_and_update(gpu):
    """The hard-call likelihoods through the update kernel: the initial E is the
    centred genotype (half-dosage units, missing 0), and the update equals the
    oracle's fit_with_pi on the hard calls."""
    pay = pa.sim_bed(17, 66, 301, 3, 0.1, 0.05, 0.5, 0.05)
    n, m, k = 66, 301, 3
    gpu.load_bed(pay, n, m)
    gpu.gl_from_bed()
    gpu.rsvd_init(k, 5, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64)
    codes = O.decode_codes(pay, n, m)
    F, _ = O.allele_freq(codes)
    P0, P1 = O.plink_hard_call_gl(codes)
    rng = np.random.default_rng(8)
    U = np.linalg.qr(rng.standard_normal((n, k)))[0]
    V = np.linalg.qr(rng.standard_normal((m, k)))[0]
    S = np.array([9.0, 5.0, 2.0])
    E, _ = gpu.pcangsd_update(U, S, V)
    want = O.pcangsd_E(P0, P1, F, U, S, V).T
    # P is FP32 on the device (0.333333 rounds), E is stored FP32
    assert np.abs(E - want).max() < 1e-6 * np.abs(want).max()
    # the first solve runs on the centred genotype: its top singular values match
    # the oracle's on BED2GENO - F
    gpu.rsvd_init(k, 5, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64)
    U0, S0, V0, _, _ = gpu.usv(7, 0.0)
    orc = O.RsvdOracle(O.centered_matrix(codes, F), k, 5, fancy=False)
    orc.compute_usv(7, 0.0)
    assert eig_relerr(S0, orc.S[:k]) < 1e-6  # E is FP32 on the device


def test_pcangsd_errors(gpu):
    gl = _gl(13, 40, 200)
    gpu.load_gl(gl)
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(2, 4, variant=pa.RSVD_NORMAL)  # allele frequencies not estimated yet
    with pytest.raises(pa.GpuError):
        gpu.load_gl(gl[:, :, :1])  # one likelihood per sample


@pytest.mark.parametrize("arm", ["fp64", "bf16x2"])
def test_pcangsd_iram_matches_reference(gpu, arm):
    """--svd 0 on likelihoods: IRAM in every EM iteration, flip_UV and the RMSE
    stop (src/Arnoldi.cpp:57-104), then pcangsd_standardize_E and the GRM,
    against the reference CLI's run with its defaults (ncv 20, itol 1e-6,
    maxiter 100, tol-em 1e-5): the RMSE it logs per iteration (6 significant
    digits), its eigenvalues, vectors and GRM."""
    z = load_golden(GOLDEN / "gl_pcangsd_svd0.npz")
    d = parse_args(z["args"])
    m = int(z["nsnps"])
    gpu.load_gl(z["gl"])
    gpu.gl_em_maf(100, 1e-6)
    gpu.gram_init(ARMS[arm], standardize=False)
    U, S, V, it, diffs, Dc, Cm = gpu.pcangsd_iram(d["k"])
    assert it == len(z["em_rmse"]) and diffs[-1] < 1e-5 and np.all(diffs[:-1] >= 1e-5)
    # the split-BF16 Gram products move the RMSEs by up to ~3e-7; the EM with
    # each solve's own vectors starts 4e-4 away (1.698e-3)
    assert np.allclose(diffs, z["em_rmse"], rtol=1e-4, atol=1e-9 if arm == "fp64" else 1e-6)
    assert eig_relerr(S ** 2 / m * 2, z["eigvals"]) < EIG_TOL[arm]
    assert O.principal_angles_max(U, z["eigvecs"]) < 1e-4
    assert O.principal_angles_max(V, z["loadings"]) < 1e-4
    # flip_UV: each column's entry of largest magnitude in U is positive
    assert np.all(U[np.abs(U).argmax(axis=0), np.arange(U.shape[1])] > 0)
    assert np.max(np.abs(Cm - z["cov"])) < 2e-5 * np.max(np.abs(z["cov"]))
    assert np.all(np.isfinite(Dc)) and np.allclose(np.diag(Cm), Dc / m)
