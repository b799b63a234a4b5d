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
