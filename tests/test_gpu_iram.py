"""GPU parity of the IRAM boundary: the Spectra perform_op Gram mat-vec and the
restarted-Lanczos driver around it (tests marked `gpu`).

perform_op is checked against the oracle's X X^T x / X^T X x on the same
inputs; the IRAM solve against exact FP64 SVD (what IRAM converges to) and the
reference CLI's own --svd 0 output (tests/golden/bin_svd0.npz)."""
from __future__ import annotations

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import eig_relerr, oracle_matrix
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

ARMS = {"fp64": pa.PREC_FP64, "fp32": pa.PREC_FP32, "bf16x2": pa.PREC_SPLIT_BF16}
MV_TOL = {"fp64": 1e-12, "fp32": 3e-6, "bf16x2": 3e-5}


@pytest.mark.parametrize("arm", list(ARMS))
@pytest.mark.parametrize("source", ["bed", "dense"])
def test_gram_matvec_matches_oracle(gpu, arm, source):
    rng = np.random.default_rng(3)
    if source == "bed":
        n, m = 301, 2503
        pay = pa.sim_bed(17, n, m, 3, 0.1, 0.05, 0.5, 0.02)
        gpu.load_bed(pay, n, m)
        X = oracle_matrix(pay, n, m)
    else:
        D = pa.sim_counts(9, 301, 777, base=0.1)
        gpu.load_dense(D)
        X = O.dense_centered(D)
        n, m = X.shape
    gpu.gram_init(ARMS[arm], standardize=True)
    for side, dim in ((0, n), (1, m)):
        for _ in range(2):
            x = rng.standard_normal(dim)
            y = gpu.gram_matvec(side, x)
            want = X @ (X.T @ x) if side == 0 else X.T @ (X @ x)
            tol = MV_TOL[arm] * (3 if source == "dense" and arm == "bf16x2" else 1)
            assert np.abs(y - want).max() / np.abs(want).max() < tol, (side, arm)
    # linearity: A (a x + b z) = a A x + b A z
    x, z = rng.standard_normal(n), rng.standard_normal(n)
    lhs = gpu.gram_matvec(0, 2.0 * x - 3.0 * z)
    rhs = 2.0 * gpu.gram_matvec(0, x) - 3.0 * gpu.gram_matvec(0, z)
    assert np.abs(lhs - rhs).max() / np.abs(rhs).max() < 10 * MV_TOL[arm]


@pytest.mark.parametrize("arm", ["fp64", "bf16x2"])
def test_iram_matches_exact_svd(gpu, arm):
    n, m, k = 240, 3000, 6
    pay = pa.sim_bed(29, n, m, 4, 0.05, 0.05, 0.5, 0.01)
    gpu.load_bed(pay, n, m)
    gpu.gram_init(ARMS[arm], standardize=True)
    U, S, V, nconv, nops = gpu.iram(k, tol=1e-7)
    assert nconv == k and nops > 0
    X = oracle_matrix(pay, n, m)
    u, s, vt = np.linalg.svd(X, full_matrices=False)
    assert eig_relerr(S ** 2, s[:k] ** 2) < (1e-9 if arm == "fp64" else 1e-5)
    assert O.principal_angles_max(U[:, :3], u[:, :3]) < 1e-5
    assert np.allclose(V.T @ V, np.eye(k), atol=1e-5)


def test_iram_matches_reference_cli(gpu):
    z = load_golden(GOLDEN / "bin_svd0.npz")
    gpu.load_dense(z["dense"])
    gpu.gram_init(pa.PREC_FP64)
    k = z["eigvals"].size
    U, S, V, nconv, nops = gpu.iram(k, tol=1e-6)
    m = int(z["nsnps"])
    assert eig_relerr(S ** 2 / m * 2, z["eigvals"]) < 5e-6  # 6 printed digits
    assert O.principal_angles_max(U[:, :2], z["eigvecs"][:, :2]) < 1e-4


def test_iram_k50_ncv101_matches_exact_svd(gpu):
    """The C4 IRAM setting (top-50, ncv = 2 k + 1 = 101, src/Cmd.cpp:189) on a
    dense centred matrix: eigenvalues and the leading subspace against exact
    FP64 SVD."""
    D = pa.sim_counts(31, 1200, 700, base=0.2)
    gpu.load_dense(D)
    gpu.gram_init(pa.PREC_FP64)
    U, S, V, nconv, nops = gpu.iram(50, ncv=101, tol=1e-8)
    assert nconv == 50
    X = O.dense_centered(D)
    u, s, vt = np.linalg.svd(X, full_matrices=False)
    assert eig_relerr(S ** 2, s[:50] ** 2) < 1e-9
    assert O.principal_angles_max(U[:, :5], u[:, :5]) < 1e-6
    assert np.allclose(U.T @ U, np.eye(50), atol=1e-8)
