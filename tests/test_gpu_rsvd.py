"""GPU parity of the RSVD hot path through the C ABI (tests marked `gpu`).

Every result is compared with the CPU oracle on the same seeded inputs, and
the golden cases also with the reference CLI's own printed outputs.
Tolerances: FP64 arm within FP64 rounding of the oracle; FP32 and split-BF16
arms within the north-star bound, eigenvalue relative error <= 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import eig_relerr, oracle_matrix, oracle_run, parse_args
from conftest import golden_cases, load_golden

pytestmark = pytest.mark.gpu

ARMS = {"fp64": pa.PREC_FP64, "fp32": pa.PREC_FP32, "bf16x2": pa.PREC_SPLIT_BF16}
# relative (to max |entry|) tolerance of G / H after one computeGandH call
GH_TOL = {"fp64": 1e-11, "fp32": 3e-5, "bf16x2": 3e-5}
EIG_TOL = {"fp64": 1e-9, "fp32": 1e-5, "bf16x2": 1e-5}


def _payload(seed, n, m, npops=3, fst=0.1, miss=0.01):
    return pa.sim_bed(seed, n, m, npops, fst, 0.05, 0.5, miss)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 65, 257, 301])
def test_allele_freq_bit_exact(gpu, n):
    m = 517
    pay = _payload(100 + n, n, m, npops=1, fst=0.0, miss=0.05) if n > 1 else _payload(100, n, m, 1, 0.0, 0.0)
    gpu.load_bed(pay, n, m)
    F, cnt = gpu.allele_freq()
    codes = O.decode_codes(pay, n, m)
    Fo, co = O.allele_freq(codes)
    assert np.array_equal(cnt, co)
    assert np.array_equal(F, Fo)


def _edge_payload():
    """N % 4 == 2, a monomorphic SNP (sd = 0), an all-missing SNP and a
    missing-heavy SNP in an otherwise ordinary set."""
    n, m = 150, 1203
    pay = _payload(77, n, m, miss=0.01).reshape(m, -1).copy()
    b = pay.shape[1]
    pay[10, :] = 0x00           # all hom A1: F = 1, sd = 0
    pay[20, :] = 0x55           # all missing: F = 0
    rng = np.random.default_rng(1)
    row = O.decode_codes(pay[30], n, 1)[0]
    row[rng.random(n) < 0.6] = 1
    packed = np.zeros(b, dtype=np.uint8)
    for j, c in enumerate(row):
        packed[j >> 2] |= np.uint8(c << (2 * (j & 3)))
    pay[30] = packed
    return pay.reshape(-1), n, m


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("variant", [pa.RSVD_NORMAL, pa.RSVD_FANCY])
def test_gandh_matches_oracle(gpu, arm, variant):
    pay, n, m = _edge_payload()
    k, os_ = 4, 6
    gpu.load_bed(pay, n, m)
    gpu.rsvd_init(k, os_, variant=variant, bands=8, precision=ARMS[arm], perm=True)
    X = oracle_matrix(pay, n, m)
    orc = O.RsvdOracle(X, k, os_, fancy=variant == pa.RSVD_FANCY, bands=8)
    for pi in range(4):
        G, H = gpu.gandh(pi)
        Go, Ho = orc.compute_gandh(pi)
        eg = np.abs(G - Go).max() / np.abs(Go).max()
        eh = np.abs(H - Ho).max() / np.abs(Ho).max()
        assert eg < GH_TOL[arm] and eh < GH_TOL[arm], (pi, eg, eh)
    if variant == pa.RSVD_FANCY:
        assert np.array_equal(gpu.perm(), O.shuffle_indices(m))


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("case", golden_cases(), ids=lambda p: p.stem)
def test_usv_matches_reference_cli(gpu, case, arm):
    z = load_golden(case)
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    if d["svd"] == 0:
        pytest.skip("IRAM cases are covered by the Gram mat-vec tests")
    if "gl" in z or d["pcangsd"]:
        pytest.skip("PCAngsd cases are covered by the genotype-likelihood tests")
    if "dense" in z:
        gpu.load_dense(z["dense"])
    else:
        gpu.load_bed(z["payload"], n, m)
    gpu.rsvd_init(d["k"], d["os"], variant=d["svd"], bands=d["w"], precision=ARMS[arm], seed=d["seed"],
                  rand=d["rand"], perm=not d["S"], standardize=bool(d["scale"]))
    if d["emu"]:
        U, S, V, it, diffs = gpu.em(d["p"], d["tol"], 100, 1e-5)
        orc, _, _ = oracle_run(z)
        ev = S ** 2 / m * 2
        assert eig_relerr(ev, z["eigvals"]) < max(5e-6, EIG_TOL[arm])
        assert O.principal_angles_max(U[:, :2], z["eigvecs"][:, :2]) < 1e-4
        if arm == "fp64":
            assert eig_relerr(ev, orc.eigvals()) < EIG_TOL[arm]
        return
    U, S, V, ep, diffs = gpu.usv(d["p"], d["tol"])
    ev = S ** 2 / m * 2
    # the reference prints 6 significant digits
    assert eig_relerr(ev, z["eigvals"]) < max(5e-6, EIG_TOL[arm])
    npc = 3 if m == 1601 else 2  # PCs with a clear spectral gap (4 populations in d2)
    assert O.principal_angles_max(U[:, :npc], z["eigvecs"][:, :npc]) < 1e-4
    assert O.principal_angles_max(V[:, :npc], z["loadings"][:, :npc]) < 1e-4
    # and with the oracle on the same inputs, epoch for epoch
    orc, epo, _ = oracle_run(z)
    if arm == "fp64":
        assert ep == epo
        assert eig_relerr(ev, orc.eigvals()) < EIG_TOL[arm]
        assert O.principal_angles_max(U, orc.U) < 1e-6
        # 1 - mev per epoch (convergence decisions follow it)
        n_d = min(len(orc.diffs), d["p"])
        assert np.allclose(diffs[1:1 + n_d], orc.diffs[:n_d], rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("variant", [pa.RSVD_NORMAL, pa.RSVD_FANCY])
def test_dense_gandh_matches_oracle(gpu, arm, variant):
    """Dense FP32 input (FileBin layout): N = 333 (not a multiple of 64),
    M = 1003 (uneven last window), one constant feature (zero after centring)."""
    D = pa.sim_counts(5, 333, 1003, base=0.1)
    D[7] = 3.5
    gpu.load_dense(D)
    assert np.allclose(gpu.feature_means(), D.astype(np.float64).mean(axis=1), rtol=1e-13, atol=1e-13)
    gpu.rsvd_init(6, 8, variant=variant, bands=8, precision=ARMS[arm])
    X = O.dense_centered(D)
    orc = O.RsvdOracle(X, 6, 8, fancy=variant == pa.RSVD_FANCY, bands=8)
    # split arm: X itself carries a 16-bit (hi + lo BF16) representation error
    tol = {"fp64": 1e-11, "fp32": 3e-5, "bf16x2": 1e-4}[arm]
    for pi in range(3):
        G, H = gpu.gandh(pi)
        Go, Ho = orc.compute_gandh(pi)
        eg = np.abs(G - Go).max() / np.abs(Go).max()
        eh = np.abs(H - Ho).max() / np.abs(Ho).max()
        assert eg < tol and eh < tol, (pi, eg, eh)
    gpu.rsvd_init(6, 8, variant=variant, bands=8, precision=ARMS[arm])
    orc = O.RsvdOracle(X, 6, 8, fancy=variant == pa.RSVD_FANCY, bands=8)
    U, S, V, ep, _ = gpu.usv(8, 1e-5)
    orc.compute_usv(8, 1e-5)
    assert eig_relerr(S ** 2, orc.S ** 2) < EIG_TOL[arm]


@pytest.mark.parametrize("arm", list(ARMS))
def test_usv_edge_inputs(gpu, arm):
    pay, n, m = _edge_payload()
    gpu.load_bed(pay, n, m)
    X = oracle_matrix(pay, n, m)
    for variant, kw in ((pa.RSVD_NORMAL, dict(rand=0, seed=3)), (pa.RSVD_FANCY, dict(seed=5))):
        gpu.rsvd_init(3, 7, variant=variant, bands=4, precision=ARMS[arm], **kw)
        U, S, V, ep, _ = gpu.usv(6, 1e-6)
        orc = O.RsvdOracle(X, 3, 7, fancy=variant == pa.RSVD_FANCY, bands=4, **kw)
        epo = orc.compute_usv(6, 1e-6)
        assert eig_relerr(S, orc.S) < EIG_TOL[arm] / 2
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-10)
        if arm == "fp64":
            assert ep == epo


@pytest.mark.parametrize("arm", list(ARMS))
def test_emu_iterations_match_oracle(gpu, arm):
    """EMU EM-PCA: per-iteration 1 - mev and the final eigenvalues against the
    oracle's restatement of the EM loop (src/Halko.cpp:294-336)."""
    n, m = 240, 3000
    pay = _payload(55, n, m, miss=0.04)
    gpu.load_bed(pay, n, m)
    gpu.rsvd_init(3, 7, variant=pa.RSVD_NORMAL, precision=ARMS[arm])
    U, S, V, it, diffs = gpu.em(6, 0.0, 50, 1e-5)
    codes = O.decode_codes(pay, n, m)
    orc, ito, diffs_o = O.emu_halko(codes, 3, 7, 6, 0.0, maxiter=50, tolem=1e-5)
    assert it == ito
    assert np.allclose(diffs, diffs_o, rtol=1e-3 if arm != "fp64" else 1e-8, atol=1e-9)
    assert eig_relerr(S, orc.S) < EIG_TOL[arm]
    assert O.principal_angles_max(V, orc.V) < (1e-7 if arm == "fp64" else 1e-4)


def test_unstandardized_and_ploidy(gpu):
    pay = _payload(8, 400, 2400)
    gpu.load_bed(pay, 400, 2400)
    X = oracle_matrix(pay, 400, 2400, standardize=False)
    gpu.rsvd_init(4, 4, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64, standardize=False)
    U, S, V, ep, _ = gpu.usv(5, 0.0)
    orc = O.RsvdOracle(X, 4, 4, fancy=False)
    orc.compute_usv(5, 0.0)
    assert eig_relerr(S, orc.S) < 1e-10
    # haploid scaling divides by sqrt(1 F (1-F)) instead of sqrt(2 F (1-F))
    codes = O.decode_codes(pay, 400, 2400)
    F, _ = O.allele_freq(codes)
    Xh = O.standardize(O.centered_matrix(codes, F), F, ploidy=1)
    gpu.rsvd_init(4, 4, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64, ploidy=1)
    U, S, V, ep, _ = gpu.usv(5, 0.0)
    orc = O.RsvdOracle(Xh, 4, 4, fancy=False)
    orc.compute_usv(5, 0.0)
    assert eig_relerr(S, orc.S) < 1e-10


def test_explicit_omega_and_permutation(gpu):
    pay = _payload(21, 256, 2048)
    gpu.load_bed(pay, 256, 2048)
    rng = np.random.default_rng(4)
    om = rng.standard_normal((256, 10))
    perm = rng.permutation(2048).astype(np.int32)
    gpu.rsvd_init(5, 5, variant=pa.RSVD_FANCY, bands=16, precision=pa.PREC_FP64, omega=om, permutation=perm)
    assert np.array_equal(gpu.perm(), perm)
    U, S, V, ep, _ = gpu.usv(8, 1e-5)
    X = oracle_matrix(pay, 256, 2048)
    orc = O.RsvdOracle(X, 5, 5, fancy=True, bands=16, omega_override=om)
    orc.permutation = perm.astype(np.int64)
    orc.X = np.ascontiguousarray(X[:, perm])
    orc.compute_usv(8, 1e-5)
    assert eig_relerr(S, orc.S) < 1e-9
    assert O.principal_angles_max(V, orc.loadings_unpermuted()) < 1e-6


def test_errors_are_reported(gpu):
    pay = _payload(2, 40, 400)
    gpu.load_bed(pay, 40, 400)
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(20, 10, bands=4)       # 2 (k + os) > N
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(2, 2, bands=5)         # bands odd (src/Cmd.cpp:213)
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(2, 2, bands=2)         # bands below 4
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(2, 2, bands=512)       # more windows than SNPs
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(40, 40, bands=4)       # l > 64


@pytest.mark.parametrize("bands", [6, 12])
def test_even_bands_match_oracle(gpu, bands):
    """winSVD with an even window count that is not a power of two, which the
    reference accepts (src/Cmd.cpp:213): schedule, epochs and results equal
    the oracle's; the plain variant ignores -w."""
    n, m = 200, 2400
    pay = _payload(61, n, m)
    gpu.load_bed(pay, n, m)
    X = oracle_matrix(pay, n, m)
    gpu.rsvd_init(4, 6, variant=pa.RSVD_FANCY, bands=bands, precision=pa.PREC_FP64)
    U, S, V, ep, diffs = gpu.usv(10, 1e-4)
    orc = O.RsvdOracle(X, 4, 6, fancy=True, bands=bands)
    epo = orc.compute_usv(10, 1e-4)
    assert ep == epo
    assert eig_relerr(S, orc.S) < 1e-9
    assert O.principal_angles_max(V, orc.loadings_unpermuted()) < 1e-6
    gpu.rsvd_init(4, 6, variant=pa.RSVD_NORMAL, bands=bands, precision=pa.PREC_FP64)
    U1, S1, _, _, _ = gpu.usv(4, 0.0)
    orc = O.RsvdOracle(X, 4, 6, fancy=False)
    orc.compute_usv(4, 0.0)
    assert eig_relerr(S1, orc.S) < 1e-9


def test_split_bf16_vs_fp64_at_scale(gpu):
    """N = 4000, M = 120000, winSVD k = 10, l = 20: eigenvalues of the split
    arm against the FP64 arm on the same device (size-independent checks:
    orthonormal U, V; descending S)."""
    n, m = 4000, 120000
    pay = _payload(31, n, m, npops=5, fst=0.05, miss=0.005)
    gpu.load_bed(pay, n, m)
    res = {}
    for arm in ("fp64", "bf16x2"):
        gpu.rsvd_init(10, 10, variant=pa.RSVD_FANCY, bands=64, precision=ARMS[arm])
        res[arm] = gpu.usv(20, 1e-4)
    U, S, V, ep, _ = res["bf16x2"]
    U0, S0, V0, ep0, _ = res["fp64"]
    assert eig_relerr(S ** 2, S0 ** 2) < 1e-5
    assert O.principal_angles_max(U[:, :4], U0[:, :4]) < 1e-4
    assert np.allclose(U.T @ U, np.eye(10), atol=1e-9)
    assert np.allclose(V.T @ V, np.eye(10), atol=1e-9)
    assert np.all(np.diff(S) <= 0)
