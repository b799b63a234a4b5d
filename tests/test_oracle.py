"""Pins the CPU oracle (oracle/pcaone_oracle.py) before it is trusted as the checker.

* decode: brute force over every byte value and every N mod 4, following the
  reference's byte loop (src/FilePlink.cpp:41-52);
* allele frequency: bit-equal to the reference's sequential sum;
* Omega / shuffle: bit-equal to libstdc++ itself (a g++-built probe);
* RSVD: the reference binary's own printed outputs (tests/golden/), and the
  FP64 eigen-decomposition of X X^T for the converged leading eigenvalues;
* a planted mistake (wrong scaling / wrong seed) is caught by the same checks.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import eig_relerr, geno_sums_fast, oracle_matrix, oracle_run, parse_args
from conftest import ROOT, golden_cases, load_golden

BED2GENO_SPEC = {0b00: 1.0, 0b10: 0.5, 0b11: 0.0}  # A1 dosage / 2; 01 = missing


def _ref_byte_loop(payload, n, m):
    """Reference decode loop, one byte at a time (src/FilePlink.cpp:73-88)."""
    b = (n + 3) >> 2
    out = np.zeros((m, n), dtype=np.uint8)
    for i in range(m):
        j = 0
        for q in range(b):
            buf = int(payload[i * b + q])
            for _ in range(4):
                if j < n:
                    out[i, j] = buf & 3
                    buf >>= 2
                j += 1
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 13])
def test_decode_every_byte_value(n):
    b = (n + 3) >> 2
    payload = np.repeat(np.arange(256, dtype=np.uint8), b)  # row v = byte value v
    got = O.decode_codes(payload, n, 256)
    assert np.array_equal(got, _ref_byte_loop(payload, n, 256))
    for code, g in BED2GENO_SPEC.items():
        assert O.BED2GENO[code] == g
    assert O.BED2GENO[1] == -9.0


def test_allele_freq_matches_sequential_sum():
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 4, size=(300, 517)).astype(np.uint8)
    codes[5, :] = 1  # all missing -> F = 0
    F, cnt = O.allele_freq(codes)
    for i in range(codes.shape[0]):
        s, c = 0.0, 0
        for x in codes[i]:
            if x != 1:
                s += float(O.BED2GENO[x])
                c += 1
        want = 0.0 if c == 0 else s / c
        assert F[i] == want  # bit-equal
        assert cnt[i].sum() == codes.shape[1]


@pytest.fixture(scope="module")
def rng_probe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("probe") / "rng_probe"
    subprocess.run([gxx, "-O2", "-std=c++17", str(ROOT / "tests" / "helpers" / "rng_probe.cpp"), "-o", str(exe)],
                   check=True)
    return exe


def _probe(exe, mode, seed, n):
    out = subprocess.run([str(exe), mode, str(seed), str(n)], check=True, capture_output=True, text=True).stdout
    return out.split()


def test_normal_draws_match_libstdcxx(rng_probe):
    want = np.array([float(x) for x in _probe(rng_probe, "normal", 101, 400)])
    got = O.omega(80, 5, seed=101, rand=1).flatten(order="F")
    assert np.array_equal(got, want)


def test_uniform_draws_match_libstdcxx(rng_probe):
    want = np.array([float(x) for x in _probe(rng_probe, "uniform", 7, 300)])
    got = O.omega(100, 3, seed=7, rand=0).flatten(order="F")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 20001, 120000])
def test_shuffle_matches_libstdcxx(rng_probe, n):
    # n <= 46340 takes libstdc++'s two-positions-per-draw branch, larger n the plain one
    want = np.array([int(x) for x in _probe(rng_probe, "shuffle", 0, n)])
    assert np.array_equal(O.shuffle_indices(n), want)


@pytest.mark.parametrize("case", golden_cases(), ids=lambda p: p.stem)
def test_oracle_matches_reference_outputs(case):
    z = load_golden(case)
    orc, _, d = oracle_run(z)
    # reference prints 6 significant digits
    assert eig_relerr(orc.eigvals(), z["eigvals"]) < 5e-6
    assert np.max(np.abs(orc.S - z["sigvals"]) / z["sigvals"]) < 5e-6
    assert O.principal_angles_max(orc.U, z["eigvecs"]) < 3e-5
    assert O.principal_angles_max(orc.loadings_unpermuted(), z["loadings"]) < 3e-5


def test_planted_mistakes_are_caught():
    z = load_golden(ROOT / "tests" / "golden" / "svd1_p1.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    # (a) dropped standardisation
    X = oracle_matrix(z["payload"], n, m, standardize=False)
    orc = O.RsvdOracle(X, d["k"], d["os"], fancy=False, seed=d["seed"])
    orc.compute_usv(d["p"], d["tol"])
    assert eig_relerr(orc.eigvals(), z["eigvals"]) > 1e-2
    # (b) a different Omega draw (one power iteration leaves its trace)
    X = oracle_matrix(z["payload"], n, m)
    orc = O.RsvdOracle(X, d["k"], d["os"], fancy=False, seed=d["seed"] + 1)
    orc.compute_usv(d["p"], d["tol"])
    assert eig_relerr(orc.eigvals(), z["eigvals"]) > 1e-4
    # (c) winSVD without the shuffle differs from the shuffled reference run
    z2 = load_golden(ROOT / "tests" / "golden" / "svd2_w8.npz")
    d2 = parse_args(z2["args"])
    X = oracle_matrix(z2["payload"], int(z2["nsamples"]), int(z2["nsnps"]))
    orc = O.RsvdOracle(X, d2["k"], d2["os"], fancy=True, bands=d2["w"], perm=False)
    orc.compute_usv(d2["p"], d2["tol"])
    assert eig_relerr(orc.eigvals(), z2["eigvals"]) > 1e-5


def test_converged_rsvd_matches_eigh():
    z = load_golden(ROOT / "tests" / "golden" / "svd1_p7.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    X = oracle_matrix(z["payload"], n, m)
    orc = O.RsvdOracle(X, 5, 15, fancy=False)
    orc.compute_usv(20, 0.0)
    w = np.linalg.eigvalsh(X @ X.T)[::-1][:5] / m * 2
    assert eig_relerr(orc.eigvals()[:2], w[:2]) < 1e-10  # population PCs, large gap
    assert eig_relerr(orc.eigvals(), w) < 2e-3


def test_mev_and_flip_closed_forms():
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.standard_normal((50, 6)))
    assert abs(O.mev(q, q) - 1.0) < 1e-14
    assert abs(O.mev(q, -q) - 1.0) < 1e-14
    assert abs(O.mev(q[:, :3], q[:, 3:])) < 1e-14
    omg2 = rng.standard_normal((50, 3))
    new = omg2.copy()
    new[:, 1] = -omg2[:, 1] * 1.01  # much closer to -Omega2: flipped back
    o2, o = O.flip_omg(omg2, new)
    assert np.allclose(o[:, 1], omg2[:, 1] * 1.01) and np.allclose(o[:, 0], omg2[:, 0])
    assert np.array_equal(o2, o)


@pytest.mark.parametrize("bands", [4, 8, 64])
def test_fancy_schedule_closed_forms(bands):
    st = O.FancyState(bands)
    lb = int(np.log2(bands))
    for pi in range(lb + 3):
        reset, steps = O.fancy_epoch_plan(st, pi)
        upd = [b for b, _, u, _ in steps if u]
        assert reset == (2 ** pi >= bands)
        if pi == 0:
            assert upd == list(range(1, bands))  # every window after the first
        if pi >= lb:  # full band: one refresh, after the last window
            assert upd == [bands - 1]
            assert [a for _, a, _, _ in steps] == [1] * (bands // 2) + [2] * (bands // 2)
        if 0 < pi < lb:
            assert upd[0] == 2 ** (pi - 1) - 1  # the "adjacent" refresh


# --- genotype likelihoods (PCAngsd on BEAGLE input) --------------------------

def test_em_maf_with_gl_closed_forms():
    rng = np.random.default_rng(3)
    n, m = 40, 25
    # certain genotypes: the posterior is the genotype whatever the prior, so
    # one EM step lands on the dosage mean and the second step confirms it
    g = rng.integers(0, 3, size=(n, m))
    P0, P1 = (g == 0).astype(float), (g == 1).astype(float)
    F, it = O.em_maf_with_gl(P0, P1, 100, 1e-12)
    assert np.allclose(F, g.mean(axis=0) / 2.0, rtol=0, atol=1e-15) and it == 2
    # flat likelihoods carry no information: F stays at its start value 0.25
    flat = np.full((n, m), 1.0 / 3.0)
    F, it = O.em_maf_with_gl(flat, flat, 100, 1e-12)
    assert np.allclose(F, 0.25, rtol=0, atol=1e-15) and it == 1
    # the EM fixed point: F = mean posterior(F) / 2 at convergence
    P = rng.dirichlet([1.0, 1.0, 1.0], size=(n, m))
    F, _ = O.em_maf_with_gl(P[..., 0], P[..., 1], 500, 1e-14)
    post, *_ = O._posterior(P[..., 0], P[..., 1], F[None, :])
    assert np.allclose(post.mean(axis=0) / 2.0, F, rtol=0, atol=1e-12)
    # initial E is centred on 2F per site at that fixed point
    E = O.gl_initial_E(P[..., 0], P[..., 1], F)
    assert np.allclose(E.mean(axis=0), 0.0, atol=1e-12)


def test_pcangsd_update_closed_forms():
    rng = np.random.default_rng(4)
    n, m, k = 30, 20, 2
    P = rng.dirichlet([1.0, 1.0, 1.0], size=(n, m))
    F = rng.uniform(0.1, 0.4, m)
    U, S, V = rng.standard_normal((n, k)), np.abs(rng.standard_normal(k)), rng.standard_normal((m, k))
    # zero singular values: the individual frequency prior falls back to F
    E0 = O.pcangsd_E(P[..., 0], P[..., 1], F, U, np.zeros(k), V)
    assert np.allclose(E0, O.gl_initial_E(P[..., 0], P[..., 1], F), rtol=0, atol=1e-15)
    # huge loadings clamp the prior to [1e-4, 1 - 1e-4]: the posterior is the
    # brute-force one at the clamped prior
    E = O.pcangsd_E(P[..., 0], P[..., 1], F, U, S * 1e6, V)
    pt = np.where((U * S) @ V.T > 0, 1 - 1e-4, 1e-4)
    p = [P[..., 0] * (1 - pt) ** 2, P[..., 1] * 2 * pt * (1 - pt), P[..., 2] * pt ** 2]
    want = (p[1] + 2 * p[2]) / (p[0] + p[1] + p[2]) - 2 * F[None, :]
    assert np.allclose(E, want, rtol=1e-12, atol=1e-12)
    # the GRM diagonal term: E[(g - 2F)^2] / (2F(1-F)) under the posterior
    Es, Dc = O.pcangsd_standardize(P[..., 0], P[..., 1], F, U, S, V)
    ptc = np.clip(((U * S) @ V.T + 2 * F[None, :]) / 2, 1e-4, 1 - 1e-4)
    q = np.stack([P[..., 0] * (1 - ptc) ** 2, P[..., 1] * 2 * ptc * (1 - ptc), P[..., 2] * ptc ** 2])
    q /= q.sum(axis=0)
    var = sum(q[gi] * (gi - 2 * F[None, :]) ** 2 for gi in range(3))
    assert np.allclose(Dc, (var / (2 * F * (1 - F))[None, :]).sum(axis=1), rtol=1e-12)
    mean = q[1] + 2 * q[2]
    assert np.allclose(Es, (mean - 2 * F[None, :]) / np.sqrt(2 * F * (1 - F))[None, :], rtol=1e-12, atol=1e-12)


def test_pcangsd_grm_matches_reference():
    z = load_golden(ROOT / "tests" / "golden" / "gl_pcangsd_svd1.npz")
    from _cases import gl_planes
    P0, P1 = gl_planes(z["gl"])
    d = parse_args(z["args"])
    op, F, it, diffs, Dc, C = O.pcangsd_halko(P0, P1, d["k"], d["os"], d["p"], d["tol"])
    # the reference writes the GRM with 6 significant digits
    assert np.max(np.abs(C - z["cov"])) < 1e-5 * np.max(np.abs(z["cov"]))
    assert np.allclose(np.diag(C), Dc / P0.shape[1])
    # planted mistake: the GRM without the Dc diagonal differs visibly
    E, _ = O.pcangsd_standardize(P0, P1, F, op.U, op.S, op.V)
    assert np.max(np.abs(np.diag(E @ E.T / P0.shape[1]) - np.diag(z["cov"]))) > 1e-3


def test_pcangsd_iram_matches_reference():
    """The IRAM EM (--svd 0 on likelihoods) against the reference CLI's run: the
    RMSE it logs per iteration (6 significant digits), its eigenvalues, vectors
    and GRM.  Planted mistake: the EM with every solve's own singular vectors
    (the textbook update, not what the reference's solver object hands out)
    follows another trajectory and stops elsewhere."""
    z = load_golden(ROOT / "tests" / "golden" / "gl_pcangsd_svd0.npz")
    from _cases import gl_planes
    P0, P1 = gl_planes(z["gl"])
    m = P0.shape[1]
    k = parse_args(z["args"])["k"]
    U, S, V, F, it, diffs, Dc, C = O.pcangsd_iram(P0, P1, k)
    ref = z["em_rmse"]
    assert it == len(ref) == 17
    assert np.allclose(diffs, ref, rtol=1e-5, atol=0)
    assert eig_relerr(S ** 2 / m * 2, z["eigvals"]) < 5e-6
    assert O.principal_angles_max(U, z["eigvecs"]) < 1e-5
    assert O.principal_angles_max(V, z["loadings"]) < 1e-5
    assert np.all(U[np.abs(U).argmax(axis=0), np.arange(k)] > 0)
    assert np.max(np.abs(C - z["cov"])) < 1e-5 * np.max(np.abs(z["cov"]))
    *_, it2, diffs2, _, _ = O.pcangsd_iram(P0, P1, k, refresh_vectors=True)
    assert abs(diffs2[0] - ref[0]) > 0.1 * ref[0]
    U2, S2, *_ = O.pcangsd_iram(P0, P1, k, refresh_vectors=True)
    assert eig_relerr(S2 ** 2 / m * 2, z["eigvals"]) > 1e-3


def test_pcangsd_plink_planted_mistakes(monkeypatch):
    """The PLINK PCAngsd fixture pins the oracle: the EM trajectory the
    reference logs (6 significant digits) and its outputs.  An initial E in
    full-dosage units moves the first EM diff; missing hard calls read as AA
    (0, 0) instead of thirds move the eigenvalues."""
    z = load_golden(ROOT / "tests" / "golden" / "plink_pcangsd_svd1.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    codes = O.decode_codes(z["payload"], n, m)
    op, _, it, diffs = O.pcangsd_plink(codes, d["k"], d["os"], d["p"], d["tol"])
    assert eig_relerr(op.eigvals(), z["eigvals"]) < 5e-6
    assert it == len(z["em_diffs"])
    assert np.allclose(diffs, z["em_diffs"], rtol=1e-5, atol=0)
    centred = O.centered_matrix
    monkeypatch.setattr(O, "centered_matrix", lambda c, F: 2.0 * centred(c, F))
    _, _, _, bad = O.pcangsd_plink(codes, d["k"], d["os"], d["p"], d["tol"])
    assert abs(bad[0] / z["em_diffs"][0] - 1) > 1e-2
    monkeypatch.setattr(O, "centered_matrix", centred)
    hard = O.plink_hard_call_gl
    monkeypatch.setattr(O, "plink_hard_call_gl", lambda c: tuple(np.where((c == 1).T, 0.0, p) for p in hard(c)))
    bad, *_ = O.pcangsd_plink(codes, d["k"], d["os"], d["p"], d["tol"])
    assert eig_relerr(bad.eigvals(), z["eigvals"]) > 1e-4


# --- LD windows and r^2 ------------------------------------------------------

def _brute_windows(pos, chr_end, ld_bp):
    """Pairs (i, j), i < j in one chromosome with pos[j] - pos[i] <= ld_bp, for
    every i whose position differs from its chromosome's last one."""
    pairs, start = [], 0
    for e in chr_end:
        for i in range(start, e + 1):
            if pos[i] == pos[e]:
                continue
            pairs += [(i, j) for j in range(i + 1, e + 1) if pos[j] - pos[i] <= ld_bp]
        start = e + 1
    return pairs


def test_ld_oracle_matches_reference():
    z = load_golden(ROOT / "tests" / "golden" / "ld_r2_bp2000.npz")
    ws, we = O.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    ia = np.concatenate([np.full(e - 1, i) for i, e in zip(ws, we)])
    ib = np.concatenate([np.arange(i + 1, i + e) for i, e in zip(ws, we)])
    assert np.array_equal(ia, z["snp_a"]) and np.array_equal(ib, z["snp_b"])
    r2 = O.ld_r2(z["resid"], ws, we)
    assert np.abs(r2 - z["r2"]).max() <= 5.01e-7  # the reference prints 6 decimals
    # brute-force windows
    pairs = _brute_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    assert pairs == list(zip(ia.tolist(), ib.tolist()))
    # planted mistake: uncentred rows
    X = z["resid"] + 0.3
    Xc = X.astype(np.float64)
    sds = 1 / np.sqrt((Xc * Xc).sum(1) / (Xc.shape[1] - 1))
    wrong = np.concatenate([((Xc[np.arange(i + 1, i + e)] @ Xc[i]) * sds[i] * sds[np.arange(i + 1, i + e)]
                             / (Xc.shape[1] - 1)) ** 2 for i, e in zip(ws, we)])
    assert np.abs(wrong - z["r2"]).max() > 1e-3


def test_ld_usv_oracle_matches_reference():
    """--USV on PLINK input: the reference's PCA gives the .eigvecs table (6
    digits), its LD run the r^2 of (I - U U^T) G with U as read_usv maps that
    table."""
    z = load_golden(ROOT / "tests" / "golden" / "ld_usv_bp3000.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    codes = O.decode_codes(z["payload"], n, m)
    ws, we = O.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    ia = np.concatenate([np.full(e - 1, i) for i, e in zip(ws, we)])
    ib = np.concatenate([np.arange(i + 1, i + e) for i, e in zip(ws, we)])
    assert np.array_equal(ia, z["snp_a"]) and np.array_equal(ib, z["snp_b"])
    U = O.read_usv(z["U"])
    R = O.ld_usv_matrix(codes, U)
    r2 = O.ld_r2(R, ws, we, centre=False)
    assert np.abs(r2 - z["r2"]).max() <= 5.01e-7  # the reference prints 6 decimals
    # closed form: the product as two thin products
    G = O.centered_matrix(codes, O.allele_freq(codes)[0])
    assert np.allclose(R.T, G - U @ (U.T @ G), rtol=0, atol=1e-12)
    # planted mistakes: no adjustment, the table taken row-major (U as printed),
    # re-centred rows
    assert np.abs(O.ld_r2(G.T, ws, we, centre=False) - z["r2"]).max() > 1e-3
    assert np.abs(O.ld_r2(O.ld_usv_matrix(codes, z["U"]), ws, we, centre=False) - z["r2"]).max() > 1e-3
    assert np.abs(O.ld_r2(R, ws, we, centre=True) - z["r2"]).max() > 1e-5


def test_ld_r2_closed_forms():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(50)
    y = rng.standard_normal(50)
    y -= (y - y.mean()) @ (x - x.mean()) / ((x - x.mean()) @ (x - x.mean())) * (x - x.mean())  # orthogonal
    X = np.stack([x, 3 * x + 1, -x, y]).astype(np.float32)
    r2 = O.ld_r2(X, np.array([0]), np.array([4]))
    assert np.allclose(r2[:2], 1.0, atol=1e-6) and abs(r2[2]) < 1e-6
    # a correlation with a known value: y = x + noise of equal variance -> r^2 = 1/2 in expectation
    z = rng.standard_normal((2, 200000))
    r2 = O.ld_r2(np.stack([z[0], z[0] + z[1]]).astype(np.float32), np.array([0]), np.array([2]))
    assert abs(r2[0] - 0.5) < 0.01


def _pairs(ws, we):
    ia = np.concatenate([np.full(e - 1, i) for i, e in zip(ws, we) if e > 1])
    ib = np.concatenate([np.arange(i + 1, i + e) for i, e in zip(ws, we) if e > 1])
    return ia, ib


def test_ld_geno_oracle_matches_reference():
    """--ld --ld-stats 1 on PLINK input, then --print-r2 on its .residuals
    (tests/golden/ld_stats1_bp3000.npz): the window pairs, the residual matrix
    itself and the r^2 from the integer sums."""
    z = load_golden(ROOT / "tests" / "golden" / "ld_stats1_bp3000.npz")
    n, m = int(z["nsamples"]), int(z["nsnps"])
    codes = O.decode_codes(z["payload"], n, m)
    ws, we = O.ld_windows(z["pos"], z["chr_end"], int(z["ld_bp"]))
    ia, ib = _pairs(ws, we)
    assert np.array_equal(ia, z["snp_a"]) and np.array_equal(ib, z["snp_b"])
    R = O.ld_stats1_matrix(codes)
    assert np.abs(R - z["resid"].astype(np.float64)).max() < 1e-6  # the file holds float32
    assert np.abs(R.astype(np.float32) - z["resid"]).max() < 1e-14
    r2 = O.ld_geno_r2(codes, ws, we)
    nan = np.isnan(z["r2"])
    assert nan.any() and np.array_equal(np.isnan(r2), nan)  # the monomorphic SNP: sd = 0
    # 6 printed decimals (5e-7) plus the reference's float32 residuals (|dr^2| <= 5e-7)
    assert np.abs(r2[~nan] - z["r2"][~nan]).max() <= 1e-6
    # planted mistakes: missing read as hom A2 (w = 1), F from all samples, no re-centring term
    y = np.where(codes == 0, 2, np.where(codes == 2, 1, 0))
    wrong = np.where(codes == 1, 3, codes).astype(np.uint8)  # missing -> code 11 (y 0, counted)
    bad = O.ld_geno_r2(wrong, ws, we)
    assert np.nanmax(np.abs(bad - z["r2"])) > 1e-3
    Fall = y.sum(axis=1) / (2.0 * n)
    Xall = np.where(codes == 1, 0.0, y / 2.0 - Fall[:, None])
    assert np.nanmax(np.abs(O.ld_r2(Xall, ws, we, centre=False) - z["r2"])) > 1e-4


def test_ld_geno_sums_brute_force():
    """Integer sums two ways (per-pair dot products, and whole Gram matrices
    of y and the mask), on N % 4 in {1, 2, 3} and heavy missingness."""
    for n, m, miss in ((9, 40, 0.3), (34, 70, 0.05), (51, 33, 0.0)):
        pay = pa.sim_bed(n * m, n, m, 2, 0.1, 0.05, 0.5, miss)
        codes = O.decode_codes(pay, n, m)
        pos = np.arange(m) * 10
        ws, we = O.ld_windows(pos, np.array([m - 1]), 95)
        s = O.ld_geno_sums(codes, ws, we)
        y = np.where(codes == 0, 2, np.where(codes == 2, 1, 0)).astype(np.int64)
        w = (codes != 1).astype(np.int64)
        ia, ib = _pairs(ws, we)
        want = np.stack([(y @ y.T)[ia, ib], (y @ w.T)[ia, ib], (w @ y.T)[ia, ib], (w @ w.T)[ia, ib]], axis=1)
        assert np.array_equal(s, want)
        assert np.array_equal(geno_sums_fast(codes, ws, we), s)
        # tiny explicit loop for the first pair
        i, k = ia[0], ib[0]
        yy = sum(int(y[i, j]) * int(y[k, j]) for j in range(n))
        ww = sum(int(codes[i, j] != 1) * int(codes[k, j] != 1) for j in range(n))
        assert s[0, 0] == yy and s[0, 3] == ww
        # the closed-form r^2 equals the FP64 residual-matrix route
        r2 = O.ld_geno_r2(codes, ws, we, s)
        ref = O.ld_r2(O.ld_stats1_matrix(codes), ws, we)
        ok = ~np.isnan(ref)
        assert np.array_equal(np.isnan(r2), np.isnan(ref)) and np.abs(r2[ok] - ref[ok]).max() < 1e-12


def test_csv_planted_mistakes():
    """The CSV reader's transform: a float median, a missing median factor or
    no standardisation each move the reference's printed eigenvalues (the
    base of the log does not: standardisation removes it)."""
    z = load_golden(ROOT / "tests" / "golden" / "csv_svd1_scale2.npz")
    d = parse_args(z["args"])
    c = z["counts"].astype(np.float64)
    n = c.shape[1]
    lib = c.sum(axis=0)

    def run(X):
        orc = O.RsvdOracle(X, d["k"], d["os"], fancy=False, seed=d["seed"])
        orc.compute_usv(d["p"], d["tol"])
        return orc.eigvals()

    def std(X):
        X = X - X.mean(axis=1, keepdims=True)
        return (X / (np.sqrt((X * X).sum(axis=1)) / np.sqrt(n - 1.0))[:, None]).T.copy()

    assert eig_relerr(run(O.csv_matrix(z["counts"], 2)), z["eigvals"]) < 5e-6
    med_f = float(np.median(lib))  # the float midpoint instead of get_median's integer one
    assert eig_relerr(run(std(np.log10(c * med_f / lib + 1))), z["eigvals"]) > 1e-5
    med = float(np.sort(lib.astype(np.int64))[n // 2 - 1] + np.sort(lib.astype(np.int64))[n // 2]) // 2
    assert eig_relerr(run(std(np.log(c * med / lib + 1))), z["eigvals"]) < 5e-6
    assert eig_relerr(run(std(np.log10(c / lib + 1))), z["eigvals"]) > 1e-3
    X = np.log10(c * med / lib + 1)
    assert eig_relerr(run((X - X.mean(axis=1, keepdims=True)).T.copy()), z["eigvals"]) > 1e-3



# --- out of core (-m): block plan, permute_plink order, block schedule -------

OOC = ["svd2_ooc_w8_m5", "svd2_ooc_w4_d2", "svd2_ooc_w8_f1"]


@pytest.mark.parametrize("name", OOC)
def test_ooc_plan_and_perm_match_reference(name):
    z = load_golden(ROOT / "tests" / "golden" / f"{name}.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    bs, nb, factor = O.ooc_plan(n, m, d["k"], d["os"], 2, d["w"], d["m"])
    assert [bs, nb, factor] == [int(x) for x in z["ooc_plan"]]  # the reference's logged plan
    assert np.array_equal(O.plink_perm_indices(n, m, 2, d["w"]), z["ooc_perm"])  # its .perm.bim order


def test_ooc_plan_closed_forms():
    # IRAM: memory * 2^27 doubles per sample column, nothing else
    assert O.ooc_plan(1000, 10 ** 6, 10, 10, 0, 64, 1.0) == (134218, 8, 1)
    # Halko with room to spare: blocksize = (memory * 2^27 - 3nl - 2ml - 5m) / n
    n, m, l = 1000, 100000, 20
    bs, nb, f = O.ooc_plan(n, m, 10, 10, 1, 64, 0.5)
    assert bs == int(np.ceil((0.5 * 134217728 - 3 * n * l - 2 * m * l - 5 * m) / n)) and nb == -(-m // bs) and f == 1
    # winSVD: fewer blocks than bands -> one block per band; else bands * factor blocks
    assert O.ooc_plan(200, 2000, 5, 5, 2, 8, 0.001) == (250, 8, 1)
    bs, nb, f = O.ooc_plan(200, 2000, 5, 5, 2, 8, 0.0005)
    assert f == 5 and bs == -(-2000 // 40) and nb == 40
    with pytest.raises(ValueError):  # a single block is the reference's error
        O.ooc_plan(200, 2000, 5, 5, 1, 8, 1.0)


@pytest.mark.parametrize("m,nb", [(2000, 8), (1601, 4), (999, 8), (64, 4)])
def test_plink_perm_closed_forms(m, nb):
    p = O.plink_perm_indices(100, m, 2, nb)
    assert np.array_equal(np.sort(p), np.arange(m))
    # band b takes SNPs b, b + nb, b + 2 nb, ... in order from its first position
    bandsize = -(-m // nb)
    modr = m % nb
    starts = [b * bandsize if (modr == 0 or b < modr) else modr * bandsize + (bandsize - 1) * (b - modr)
              for b in range(nb)] + [m]
    for b in range(nb):
        seg = p[starts[b]:starts[b + 1]]
        assert np.array_equal(seg, b + nb * np.arange(len(seg)))


def test_ooc_schedule_closed_forms():
    # factor F: the first epoch refreshes after every F-th... band of 2F blocks;
    # the in-core plan is F = 1 with nblocks = bands
    for bands, F in [(8, 5), (4, 11), (8, 1)]:
        nblocks = bands * F
        st = O.FancyState(bands, nblocks, F)
        reset, steps = O.fancy_epoch_plan(st, 0)
        upd = [b for b, _, u, _ in steps if u]
        assert st.bandsize == 2 * F and not reset
        assert upd == [b for b in range(nblocks) if (b + 1) % F == 0 and b + 1 >= 2 * F] or F == 1
        lb = int(np.log2(bands))
        for pi in range(1, lb + 2):
            reset, steps = O.fancy_epoch_plan(st, pi)
            upd = [b for b, _, u, _ in steps if u]
            if pi >= lb:
                assert upd == [nblocks - 1] and reset
            else:
                assert upd[0] == 2 ** (pi - 1) * F - 1  # the "adjacent" refresh, counted in blocks
    st_in, st_f1 = O.FancyState(8), O.FancyState(8, 8, 1)
    for pi in range(5):
        assert O.fancy_epoch_plan(st_in, pi) == O.fancy_epoch_plan(st_f1, pi)


@pytest.mark.parametrize("name", ["svd2_ooc_w8_m5", "svd2_ooc_w4_d2"])
def test_ooc_planted_mistakes(name):
    # w8_m5 runs to convergence (the order of the SNPs still shows), w4_d2 stops
    # after 2 epochs (the block schedule shows too)
    z = load_golden(ROOT / "tests" / "golden" / f"{name}.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    X = oracle_matrix(z["payload"], n, m)
    bs, _, f = O.ooc_plan(n, m, d["k"], d["os"], 2, d["w"], d["m"])
    perm = O.plink_perm_indices(n, m, 2, d["w"])

    def run(**kw):
        orc = O.RsvdOracle(X, d["k"], d["os"], fancy=True, bands=d["w"], seed=d["seed"], **kw)
        orc.compute_usv(d["p"], d["tol"])
        return orc

    good = run(blocksize=bs, band_factor=f, perm_override=perm)
    assert eig_relerr(good.eigvals(), z["eigvals"]) < 5e-6
    # the in-core shuffle in place of permute_plink's interleave
    assert eig_relerr(run(blocksize=bs, band_factor=f).eigvals(), z["eigvals"]) > 1e-5
    if name == "svd2_ooc_w4_d2":
        # the in-core windows (bands, factor 1)
        assert eig_relerr(run(perm_override=perm).eigvals(), z["eigvals"]) > 1e-5
        # the blocks without the band factor in the schedule
        assert eig_relerr(run(blocksize=bs, band_factor=1, perm_override=perm).eigvals(), z["eigvals"]) > 1e-5
