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

import pcaone_oracle as O
from _cases import eig_relerr, oracle_matrix, oracle_run, parse_args
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
