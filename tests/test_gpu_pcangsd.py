"""GPU parity of the PCAngsd path on genotype likelihoods (tests marked `gpu`):
emMAF_with_GL, the initial E, the fit_with_pi PCAngsd update and
pcangsd_standardize_E against the oracle on the same inputs, and the whole EM
(plus the GRM written to .cov) against the reference CLI's BEAGLE run
(tests/golden/gl_pcangsd_svd1.npz).

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


def test_pcangsd_errors(gpu):
    gl = _gl(13, 40, 200)
    gpu.load_gl(gl)
    with pytest.raises(pa.GpuError):
        gpu.rsvd_init(2, 4, variant=pa.RSVD_NORMAL)  # allele frequencies not estimated yet
    with pytest.raises(pa.GpuError):
        gpu.load_gl(gl[:, :, :1])  # one likelihood per sample
