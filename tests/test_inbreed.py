"""--inbreed 1 (per-site inbreeding coefficients and the HWE LRT under
population structure, src/Inbreeding.cpp): the oracle against the reference
CLI's .hwe on a PLINK set with --USV of the reference's own PCA (CPU), and the
device path against both (GPU)."""
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O

GOLD = Path(__file__).parent / "golden"


def _case():
    d = np.load(GOLD / "inbreed_d1.npz")
    n, m = int(d["nsamples"]), int(d["nsnps"])
    F = np.array([float(ln.split("\t")[6]) for ln in str(d["mbim"]).splitlines()])
    rows = [ln.split("\t") for ln in str(d["hwe"]).splitlines()[1:]]
    gold = np.array([[float(x) for x in r[1:]] for r in rows])  # HWE_P, LRT, coef
    return d, n, m, F, gold


def _close6(a, b):
    """Equal to the 6 significant digits the reference prints."""
    return np.all(np.abs(a - b) <= 5.01e-6 * np.maximum(np.abs(b), 1e-300) + 1e-300)


def test_oracle_matches_reference_cli():
    d, n, m, F, gold = _case()
    codes = O.decode_codes(d["payload"], n, m)
    coef, T, p, _ = O.inbreeding_em(codes, d["U"], d["S"], d["V"], F)
    assert _close6(coef, gold[:, 2]) and _close6(T, gold[:, 1]) and _close6(p, gold[:, 0])


def test_oracle_planted_mistakes():
    """No population structure (Pi = F) or missing genotypes read as hom A2
    move the reference's coefficients."""
    d, n, m, F, gold = _case()
    codes = O.decode_codes(d["payload"], n, m)
    coef0, *_ = O.inbreeding_em(codes, d["U"], d["S"] * 0.0, d["V"], F)
    assert np.abs(coef0 - gold[:, 2]).max() > 1e-2
    c2 = np.where(codes == 1, 3, codes)
    coef2, *_ = O.inbreeding_em(c2, d["U"], d["S"], d["V"], F)
    assert np.abs(coef2 - gold[:, 2]).max() > 1e-3


def test_hwe_writer_reproduces_reference_text(tmp_path):
    d, n, m, F, gold = _case()
    (tmp_path / "x.mbim").write_text(str(d["mbim"]))
    # the reference's own printed values go back out as its text
    pa.write_hwe(tmp_path / "x.hwe", tmp_path / "x.mbim", gold[:, 0], gold[:, 1], gold[:, 2])
    assert (tmp_path / "x.hwe").read_text() == str(d["hwe"])


@pytest.mark.gpu
def test_device_inbreeding_matches_oracle_and_reference(gpu, tmp_path):
    d, n, m, F, gold = _case()
    gpu.load_bed(d["payload"], n, m)
    coef, T, p, it = gpu.inbreeding(d["U"], d["S"], d["V"], F)
    codes = O.decode_codes(d["payload"], n, m)
    oc, oT, op, oit = O.inbreeding_em(codes, d["U"], d["S"], d["V"], F)
    assert it == oit
    # FP64 on both sides, sums in different orders; the p-value's slope
    # 1 / sqrt(2 pi T) near T = 0 magnifies the LRT's rounding
    assert np.abs(coef - oc).max() < 1e-12 and np.abs(T - oT).max() < 1e-9 and np.abs(p - op).max() < 1e-9
    assert _close6(coef, gold[:, 2]) and _close6(T, gold[:, 1]) and _close6(p, gold[:, 0])
    (tmp_path / "x.mbim").write_text(str(d["mbim"]))
    pa.write_hwe(tmp_path / "x.hwe", tmp_path / "x.mbim", p, T, coef)
    got = (tmp_path / "x.hwe").read_text().splitlines()
    ref = str(d["hwe"]).splitlines()
    assert len(got) == len(ref) and got[0] == ref[0]
    assert [g.split("\t")[0] for g in got] == [r.split("\t")[0] for r in ref]
    assert sum(g == r for g, r in zip(got, ref)) >= 0.95 * len(ref)
