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
from _cases import payload_of, csv_bytes, eig_relerr, oracle_matrix, oracle_run, parse_args
from conftest import GOLDEN, golden_cases, load_golden

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
def test_usv_matches_reference_cli(gpu, case, arm, tmp_path):
    z = load_golden(case)
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    if d["svd"] == 0:
        pytest.skip("IRAM cases are covered by the Gram mat-vec tests")
    if "gl" in z or d["pcangsd"]:
        pytest.skip("PCAngsd cases are covered by the genotype-likelihood tests")
    csv = "counts" in z
    if csv:  # the reference's --csv input, zstd-compressed, through the library's reader
        path = tmp_path / "x.csv.zst"
        path.write_bytes(csv_bytes(z["counts"]))
        gpu.load_csv(path, scale=d["scale"])
        assert (gpu.N, gpu.M) == (n, m)
    elif "dense" in z:
        gpu.load_dense(z["dense"])
    elif "bgen" in z:  # the reference's -g input through the library's BGEN reader
        path = tmp_path / "x.bgen"
        path.write_bytes(np.asarray(z["bgen"], dtype=np.uint8).tobytes())
        gpu.load_bgen(path, standardize=bool(d["scale"]) and not d["ld"])
        assert (gpu.N, gpu.M) == (n, m)
    else:
        gpu.load_bed(payload_of(z), n, m)
    ooc = {}
    if d["m"] > 0 and d["svd"] == 2:  # out-of-core winSVD: the -m block plan and permute_plink's order
        bs, _, factor = pa.ooc_plan(n, m, 2, d["k"], d["os"], d["w"], d["m"])
        ooc = {"blocksize": bs, "band_factor": factor}
        if not d["S"]:
            ooc["permutation"] = pa.plink_perm(n, m, 2, d["w"])
    gpu.rsvd_init(d["k"], d["os"], variant=d["svd"], bands=d["w"], precision=ARMS[arm], seed=d["seed"],
                  rand=d["rand"], perm=not d["S"], standardize=bool(d["scale"]), **ooc)
    if ooc:
        assert np.array_equal(gpu.perm(), ooc.get("permutation", np.arange(m)))
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
    if csv and d["scale"] == 2:
        npc = 1  # eigenvalues 2-4 of that case lie within 8 % of each other
    assert O.principal_angles_max(U[:, :npc], z["eigvecs"][:, :npc]) < 1e-4
    assert O.principal_angles_max(V[:, :npc], z["loadings"][:, :npc]) < 1e-4
    # and with the oracle on the same inputs, epoch for epoch
    orc, epo, _ = oracle_run(z)
    if arm == "fp64":
        # a CSV's transformed values and BGEN's centred dosages are held in FP32
        # on the device (as the reference's own .bin conversion of a CSV holds
        # them); the oracle's in FP64
        held32 = csv or "bgen" in z
        tol = 2e-7 if held32 else EIG_TOL[arm]
        assert ep == epo
        assert eig_relerr(ev, orc.eigvals()) < tol
        assert O.principal_angles_max(U, orc.U) < (1e-5 if held32 else 1e-6)
        # 1 - mev per epoch (convergence decisions follow it)
        n_d = min(len(orc.diffs), d["p"])
        assert np.allclose(diffs[1:1 + n_d], orc.diffs[:n_d], rtol=1e-5 if held32 else 1e-8, atol=1e-12)


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


# l = k + os across the contraction's tile counts: folded hi + lo epilogue
# (l % 16 in {0, 8}, l <= 64), the one-wave and whole-block l x l steps, and
# the general path above 64 columns (column-blocked contraction, wide
# linear algebra).  M = 6000 SNPs in 64 windows and N = 300 samples make the
# 64-window H products (N x 64 rows) and the long G segments run the streamed
# kernel, the short ones the 4-wave kernel.
L_SWEEP = [16, 32, 40, 48, 64, 100, 128]
_sweep_cache = {}


def _sweep_oracle(l):
    if l not in _sweep_cache:
        n, m = 300, 6000
        pay = _payload(90 + l, n, m, npops=3, fst=0.1, miss=0.01)
        X = oracle_matrix(pay, n, m)
        k = l // 2
        orc = O.RsvdOracle(X, k, l - k, fancy=True, bands=64)
        gh = [tuple(x.copy() for x in orc.compute_gandh(pi)) for pi in range(3)]
        orc2 = O.RsvdOracle(X, k, l - k, fancy=True, bands=64)
        epo = orc2.compute_usv(7, 1e-4)
        _sweep_cache[l] = (pay, n, m, k, gh, orc2, epo)
    return _sweep_cache[l]


# Entrywise G / H after Omega refreshes are compared in the FP64 arm only.
# Householder's R-diagonal sign (beta = -sign(pivot)) and flipOmg are
# discontinuous: a column whose pivot is near zero changes sign under any
# perturbation above FP64 rounding, and for noise columns (uncorrelated with
# the previous Omega) flipOmg does not undo it.  Emulating only the BF16 hi + lo
# rounding of the operands inside the FP64 oracle (N = 300, M = 6000, 64
# windows) already moves max |G| entries by 4e-4 at l = 64 and by O(0.1) at
# l >= 100 after one epoch, so the FP32 and split arms are held to the exact
# error bound of one refresh-free product (test_split_error_bound) and, end
# to end, to eigenvalues and subspaces (test_usv_l_sweep).


@pytest.mark.parametrize("l", L_SWEEP)
def test_gandh_l_sweep(gpu, l):
    pay, n, m, k, gh, _, _ = _sweep_oracle(l)
    gpu.load_bed(pay, n, m)
    gpu.rsvd_init(k, l - k, variant=pa.RSVD_FANCY, bands=64, precision=pa.PREC_FP64)
    for pi, (Go, Ho) in enumerate(gh):
        G, H = gpu.gandh(pi)
        eg = np.abs(G - Go).max() / np.abs(Go).max()
        eh = np.abs(H - Ho).max() / np.abs(Ho).max()
        assert eg < 1e-10 and eh < 1e-10, (l, pi, eg, eh)  # FP64 rounding through cond(H) <= 50 refreshes


U16, U24 = 2.0 ** -16, 2.0 ** -24


@pytest.mark.parametrize("arm", ["fp32", "bf16x2"])
@pytest.mark.parametrize("source", ["bed", "dense"])
@pytest.mark.parametrize("l", [20, 40, 64, 100])
def test_split_error_bound(gpu, l, source, arm):
    """G = X^T Omega and H = X G of the first plain-RSVD call against an
    entrywise error bound derived from the arithmetic, not fitted:

    * an FP64 value split into BF16 hi + lo keeps it to 2^-16 relative
      (each rounding to BF16 is within 2^-8 of its input); FP32 rounding is
      2^-24 relative;
    * packed genotypes enter as y' = 2 - 2 y (|y'| <= 2, exact), and
      G_i = d_i / 2 (colsum - sum_j y'_ij Omega_j / 2 - ...): the contraction
      error is d_i / 4 times that of sum_j y'_ij Omega_j;
    * dense values are split too and three of the four hi/lo products are
      kept: |ab - (ahbh + ahbl + albh)| <= 3 * 2^-16 |a| |b|;
    * FP32 accumulation over K terms adds at most K * 2^-24 of sum |terms|;
    * H = X G carries X's application to G's error plus its own operand and
      accumulation error on the operand d G.
    A dropped lo half (2^-8 per entry) breaks the bound by two orders of
    magnitude."""
    k = l // 2
    if source == "bed":
        n, m = 300, 5000
        pay = _payload(300 + l, n, m, miss=0.01)
        gpu.load_bed(pay, n, m)
        codes = O.decode_codes(pay, n, m)
        F, _ = O.allele_freq(codes)
        X = oracle_matrix(pay, n, m)
        sd = np.sqrt(2 * F * (1 - F))
        d = np.where(sd > 1e-9, 1.0 / sd, 1.0)
        A = np.full((n, m), 2.0) * d / 4          # G side: |y'| d / 4 <= 2 d / 4
        Ah = np.full((n, m), 0.5)                  # H side: |y'| / 4 on the operand d G
        opnd = U16 if arm == "bf16x2" else U24
        xop = 0.0                                  # the genotypes are exact
    else:
        n, m = 360, 3000
        D = pa.sim_counts(40 + l, n, m, base=0.1)
        gpu.load_dense(D)
        X = O.dense_centered(D)
        A = np.abs(D.T.astype(np.float64))
        Ah = A
        d = np.ones(m)
        opnd = U16 if arm == "bf16x2" else U24
        xop = 2 * U16 if arm == "bf16x2" else 0.0  # the split of D itself (FP32 D is exact in FP32)
    gpu.rsvd_init(k, l - k, variant=pa.RSVD_NORMAL, precision=ARMS[arm])
    G, H = gpu.gandh(0)
    orc = O.RsvdOracle(X, k, l - k, fancy=False)
    Go, Ho = orc.compute_gandh(0)
    Om = np.abs(orc.omg)
    eG = (opnd + xop + n * U24) * (A.T @ Om) * 1.01 + 1e-12 * np.abs(Go).max()
    assert np.all(np.abs(G - Go) <= eG), float(np.max(np.abs(G - Go) / eG))
    gp = np.abs(d[:, None] * G) if source == "bed" else np.abs(G)
    eH = (np.abs(X) @ eG + (opnd + xop + m * U24) * (Ah @ gp)) * 1.01 + 1e-12 * np.abs(Ho).max()
    assert np.all(np.abs(H - Ho) <= eH), float(np.max(np.abs(H - Ho) / eH))
    # the bound is not vacuous: it sits well below max |entry|
    assert eG.max() < 1e-2 * np.abs(Go).max() and eH.max() < 1e-2 * np.abs(Ho).max()


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("l", L_SWEEP)
def test_usv_l_sweep(gpu, l, arm):
    pay, n, m, k, _, orc, epo = _sweep_oracle(l)
    gpu.load_bed(pay, n, m)
    gpu.rsvd_init(k, l - k, variant=pa.RSVD_FANCY, bands=64, precision=ARMS[arm])
    U, S, V, ep, diffs = gpu.usv(7, 1e-4)
    # all k in the FP64 arm; in the others the structured PCs (npops - 1 = 2):
    # the unconverged estimates of the noise bulk after 7 epochs depend on the
    # sign decisions above (at l = 100: 9e-5 apart in the split arm, the two
    # structured eigenvalues 1e-8)
    top = k if arm == "fp64" else 2
    assert eig_relerr(S[:top] ** 2, orc.S[:top] ** 2) < EIG_TOL[arm]
    assert O.principal_angles_max(U[:, :2], orc.U[:, :2]) < (1e-7 if arm == "fp64" else 1e-4)
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-9)
    assert np.allclose(V.T @ V, np.eye(k), atol=1e-9)
    if arm == "fp64":
        assert ep == epo
        n_d = min(len(orc.diffs), 7)
        assert np.allclose(diffs[1:1 + n_d], orc.diffs[:n_d], rtol=1e-7, atol=1e-11)


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("variant", [pa.RSVD_NORMAL, pa.RSVD_FANCY])
def test_dense_l100_matches_oracle(gpu, arm, variant):
    """Dense input at the C4 operator width (k = 50, l = 100): G / H of the
    first epochs and the converged eigenvalues against the oracle."""
    D = pa.sim_counts(12, 420, 2600, base=0.1)
    gpu.load_dense(D)
    X = O.dense_centered(D)
    gpu.rsvd_init(50, 50, variant=variant, bands=16, precision=ARMS[arm])
    orc = O.RsvdOracle(X, 50, 50, fancy=variant == pa.RSVD_FANCY, bands=16)
    tol = {"fp64": 1e-10, "fp32": 3e-5, "bf16x2": 1e-4}[arm]
    # entrywise only where it is well posed (see test_gandh_l_sweep): every
    # epoch in the FP64 arm, before the first Omega refresh in the others
    npi = 3 if arm == "fp64" else (1 if variant == pa.RSVD_NORMAL else 0)
    for pi in range(npi):
        G, H = gpu.gandh(pi)
        Go, Ho = orc.compute_gandh(pi)
        eg = np.abs(G - Go).max() / np.abs(Go).max()
        eh = np.abs(H - Ho).max() / np.abs(Ho).max()
        assert eg < tol and eh < tol, (pi, eg, eh)
    gpu.rsvd_init(50, 50, variant=variant, bands=16, precision=ARMS[arm])
    orc = O.RsvdOracle(X, 50, 50, fancy=variant == pa.RSVD_FANCY, bands=16)
    U, S, V, ep, _ = gpu.usv(6, 1e-5)
    epo = orc.compute_usv(6, 1e-5)
    top = 50 if arm == "fp64" else 3
    assert eig_relerr(S[:top] ** 2, orc.S[:top] ** 2) < EIG_TOL[arm]
    if arm == "fp64":
        assert ep == epo


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
    gpu.load_bed(_payload(2, 400, 400), 400, 400)
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(60, 69, bands=4)       # l > 128
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(65, 0, bands=4)        # k > 64


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


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("source", ["bed", "dense"])
def test_rank_deficient_windows(gpu, arm, source):
    """winSVD windows of fewer SNPs than l: the first Omega refreshes see a
    partial H of rank below l, where CholeskyQR breaks down and the reference's
    Householder QR fills the trailing columns from rounding.  The solve
    restarts with the rank repair; its top eigenvalues match exact SVD and
    the oracle's Householder run."""
    if source == "bed":
        n, m, k = 300, 2048, 32               # 32 SNPs per window, l = 64
        pay = _payload(81, n, m, npops=3, fst=0.1, miss=0.01)
        gpu.load_bed(pay, n, m)
        X = oracle_matrix(pay, n, m)
    else:
        n, m, k = 3000, 640, 50               # 10 genes per window, l = 100
        D = pa.sim_counts(82, n, m, base=0.1)
        gpu.load_dense(D)
        X = O.dense_centered(D)
    gpu.rsvd_init(k, k, variant=pa.RSVD_FANCY, bands=64, precision=ARMS[arm])
    U, S, V, ep, _ = gpu.usv(20, 1e-7)
    s = np.linalg.svd(X, compute_uv=False)
    orc = O.RsvdOracle(X, k, k, fancy=True, bands=64)
    orc.compute_usv(20, 1e-7)
    top = 5
    # as close to exact SVD as the reference's own algorithm gets (the oracle's
    # Householder run), within the arm's arithmetic
    e_orc = eig_relerr(orc.S[:top] ** 2, s[:top] ** 2)
    assert e_orc < 1e-5
    assert eig_relerr(S[:top] ** 2, s[:top] ** 2) < max(10 * e_orc, 1e-9 if arm == "fp64" else 1e-5)
    assert np.allclose(U.T @ U, np.eye(k), atol=1e-8)


@pytest.mark.parametrize("arm", list(ARMS))
def test_matrix_rank_below_l(gpu, arm):
    """A matrix of rank below k + oversamples (20 distinct samples, each five
    times: rank <= 20, l = 40).  G and H are singular; the reference's
    Householder QR returns trailing columns set by rounding, CholeskyQR2 (with
    the rank repair when a pivot fails) an orthonormal completion of its own.
    The eigenvalues within the rank match exact SVD as closely as the
    oracle's Householder run does."""
    base = O.decode_codes(_payload(83, 20, 3000), 20, 3000)
    codes = np.repeat(base, 5, axis=1)        # 100 samples
    n, m = codes.shape[1], codes.shape[0]
    b = (n + 3) // 4
    pay = np.zeros((m, b), dtype=np.uint8)
    for j in range(n):
        pay[:, j >> 2] |= (codes[:, j] << (2 * (j & 3))).astype(np.uint8)
    pay = pay.reshape(-1)
    gpu.load_bed(pay, n, m)
    X = oracle_matrix(pay, n, m)
    s = np.linalg.svd(X, compute_uv=False)
    r = int(np.sum(s > 1e-9 * s[0]))
    assert r <= 20
    gpu.rsvd_init(20, 20, variant=pa.RSVD_FANCY, bands=8, precision=ARMS[arm])
    U, S, V, ep, _ = gpu.usv(10, 1e-6)
    orc = O.RsvdOracle(X, 20, 20, fancy=True, bands=8)
    orc.compute_usv(10, 1e-6)
    top = min(r, 10)
    e_orc = eig_relerr(orc.S[:top] ** 2, s[:top] ** 2)
    e_gpu = eig_relerr(S[:top] ** 2, s[:top] ** 2)
    assert e_gpu < max(10 * e_orc, 1e-8 if arm == "fp64" else 1e-5), (e_gpu, e_orc)
    assert np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    assert np.allclose(U[:, :top].T @ U[:, :top], np.eye(top), atol=1e-6)


def test_csv_reader_forms_and_errors(gpu, tmp_path):
    """The same counts as zstd and as plain text give the same matrix (the
    same solve, bit for bit); a last line without a newline is left out, as
    the reference counts lines; unaligned rows, non-numbers and --scale 2
    on non-integer counts are errors."""
    z = load_golden(GOLDEN / "csv_svd1_scale2.npz")
    counts = z["counts"]
    res = []
    for zst in (True, False):
        path = tmp_path / ("a.csv.zst" if zst else "a.csv")
        path.write_bytes(csv_bytes(counts, zstd=zst))
        gpu.load_csv(path, scale=2, nthreads=3)
        gpu.rsvd_init(4, 6, variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64)
        res.append(gpu.usv(5, 0.0)[1])
    assert np.array_equal(res[0], res[1])
    (tmp_path / "b.csv").write_bytes(csv_bytes(counts, zstd=False) + b"1,2,3")
    gpu.load_csv(tmp_path / "b.csv", scale=2)
    assert (gpu.N, gpu.M) == counts.shape[::-1]
    (tmp_path / "c.csv").write_text("1,2,3\n4,5\n")
    with pytest.raises(pa.GpuError):
        gpu.load_csv(tmp_path / "c.csv", scale=2)
    (tmp_path / "d.csv").write_text("1,2,3\n4,x,6\n")
    with pytest.raises(pa.GpuError):
        gpu.load_csv(tmp_path / "d.csv", scale=1)
    (tmp_path / "e.csv").write_text("1,2,3\n4,.5,6\n")
    with pytest.raises(pa.GpuError):
        gpu.load_csv(tmp_path / "e.csv", scale=2)
    gpu.load_csv(tmp_path / "e.csv", scale=1)  # std::stof reads .5; only --scale 2 needs integers

