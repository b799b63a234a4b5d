"""--project 1|2 with --USV (run_projection, src/Projection.cpp:147-208) on a
PLINK target: the oracle against the reference CLI's projected .eigvecs
(CPU), the device path against both (GPU).  The target leaves out SNPs, swaps
the alleles of others and moves a few, so matching, flips and missing
genotypes all act."""
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O

GOLD = Path(__file__).parent / "golden"


def _case():
    d = np.load(GOLD / "project_p12.npz")
    n, m = int(d["ntarget"]), int(d["ntarget_snps"])
    codes = O.decode_codes(d["tgt_payload"], n, m)
    return d, n, m, codes, str(d["tgt_bim"]).splitlines(), str(d["mbim"]).splitlines()


@pytest.mark.parametrize("method", [1, 2])
def test_oracle_matches_reference_cli(method):
    d, n, m, codes, bim, mbim = _case()
    U, W = O.project_plink(codes, bim, mbim, d["S"], d["V"], method)
    gold = d[f"U_m{method}"]
    assert U.shape == gold.shape
    assert np.abs(U - gold).max() <= 5.01e-6 * np.abs(gold).max()
    # eigenvalues as written: S^2 / (matched SNPs) * 2
    ev = np.array([float(x) for x in str(d[f"text_eigvals_m{method}"]).split()])
    assert np.allclose(ev, d["S"] ** 2 / W.shape[0] * 2, rtol=5e-6)


def test_oracle_planted_mistakes():
    """Ignoring the allele swap, or leaving missing genotypes in the fit as
    hom-A2, moves the reference's projection."""
    d, n, m, codes, bim, mbim = _case()
    gold = d["U_m2"]
    noswap = [ln.replace("\tG\tA", "\tA\tG") for ln in bim]
    U1, _ = O.project_plink(codes, noswap, mbim, d["S"], d["V"], 2)
    assert np.abs(U1 - gold).max() > 1e-3 * np.abs(gold).max()
    U2, _ = O.project_plink(np.where(codes == 1, 3, codes), bim, mbim, d["S"], d["V"], 2)
    assert np.abs(U2 - gold).max() > 1e-3 * np.abs(gold).max()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [1, 2])
def test_device_projection(gpu, tmp_path, method):
    d, n, m, codes, bim, mbim = _case()
    (tmp_path / "t.bim").write_text(str(d["tgt_bim"]))
    (tmp_path / "r.mbim").write_text(str(d["mbim"]))
    gpu.load_bed(d["tgt_payload"], n, m)
    U, W = gpu.project(tmp_path / "t.bim", tmp_path / "r.mbim", d["S"], d["V"], method=method)
    Uo, Wo = O.project_plink(codes, bim, mbim, d["S"], d["V"], method)
    assert W.shape == Wo.shape and np.abs(W - Wo).max() == 0.0
    assert np.abs(U - Uo).max() < 1e-10 * np.abs(Uo).max()
    gold = d[f"U_m{method}"]
    assert np.abs(U - gold).max() <= 5.01e-6 * np.abs(gold).max()


# ---- BEAGLE likelihoods (run_projection_beagle_gl, src/Projection.cpp:32-137)

def _gl_case():
    d = np.load(GOLD / "project_gl.npz")
    sites = [tuple(s.split("\t")) for s in d["sites"]]
    mbim = str(d["mbim"]).splitlines()
    gl = d["gl"]
    P0 = np.ascontiguousarray(gl[:, :, 0].T.astype(np.float64))
    P1 = np.ascontiguousarray(gl[:, :, 1].T.astype(np.float64))
    return d, sites, mbim, gl, P0, P1


def test_oracle_gl_projection_matches_reference_cli():
    d, sites, mbim, gl, P0, P1 = _gl_case()
    ri, fl, F = O.beagle_mbim_match(sites, mbim)
    nm = int((ri >= 0).sum())
    assert 0 < nm < len(sites) and 0 < fl.sum() < nm  # unmatched and swapped sites both occur
    U, VS, it, diffs = O.project_beagle_gl(P0, P1, ri, fl, F, d["S"], d["V"])
    assert it == len(d["diffs"])
    assert np.allclose(diffs, d["diffs"], rtol=5e-6, atol=1e-9)  # the log's 6 digits
    assert np.abs(U - d["U"]).max() <= 5.01e-6 * np.abs(d["U"]).max()
    # eigenvalues as written: S^2 / (matched sites) * 2
    ev = np.array([float(x) for x in str(d["text_eigvals"]).split()])
    assert np.allclose(ev, d["S"] ** 2 / nm * 2, rtol=5e-6)


def test_oracle_gl_projection_planted_mistakes():
    """Swaps ignored (F and the loading sign kept), or the posterior weights of
    the two homozygotes exchanged, move the reference's scores."""
    d, sites, mbim, gl, P0, P1 = _gl_case()
    ri, fl, F = O.beagle_mbim_match(sites, mbim)
    gold = d["U"]
    Fk = np.where(fl, 1.0 - F, F)
    U1, *_ = O.project_beagle_gl(P0, P1, ri, np.zeros_like(fl), Fk, d["S"], d["V"])
    assert np.abs(U1 - gold).max() > 1e-3 * np.abs(gold).max()
    U2, *_ = O.project_beagle_gl(1.0 - P0 - P1, P1, ri, fl, F, d["S"], d["V"])
    assert np.abs(U2 - gold).max() > 1e-3 * np.abs(gold).max()


@pytest.mark.gpu
@pytest.mark.parametrize("standardize", [True, False])
def test_device_gl_projection(gpu, tmp_path, standardize):
    d, sites, mbim, gl, P0, P1 = _gl_case()
    (tmp_path / "t.beagle.gz").write_bytes(np.asarray(d["beagle_gz"], dtype=np.uint8).tobytes())
    (tmp_path / "r.mbim").write_text(str(d["mbim"]))
    gpu.load_gl(gl)
    U, W, it, diffs = gpu.project_gl(tmp_path / "t.beagle.gz", tmp_path / "r.mbim", d["S"], d["V"],
                                     standardize=standardize)
    ri, fl, F = O.beagle_mbim_match(sites, mbim)
    Uo, VSo, ito, dfo = O.project_beagle_gl(P0, P1, ri, fl, F, d["S"], d["V"], standardize=standardize)
    assert W.shape == VSo.shape and np.array_equal(W, VSo)
    assert it == ito
    assert np.allclose(diffs, dfo, rtol=1e-8, atol=1e-13)
    assert np.abs(U - Uo).max() < 1e-10 * np.abs(Uo).max()
    if standardize:  # the reference CLI's own run
        assert np.abs(U - d["U"]).max() <= 5.01e-6 * np.abs(d["U"]).max()
    # the entry point that takes the matching as given
    U2, W2, it2, d2 = gpu.project_gl_matched(ri, fl, F, d["S"], d["V"], standardize=standardize)
    assert it2 == it and np.array_equal(U2, U) and np.array_equal(W2, W)
