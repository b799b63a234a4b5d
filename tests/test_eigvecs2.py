"""The GRM eigenvector output (.eigvecs2) of the PCAngsd branch
(src/Halko.cpp:317-331): the writer against the reference CLI's text, and the
fixture's vectors against the mathematics they are defined by (eigenvectors
of the .cov GRM ordered by |eigenvalue|, JacobiSVD's singular vectors of a
symmetric matrix)."""
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa

GOLD = Path(__file__).parent / "golden"


def _fixture():
    return np.load(GOLD / "gl_pcangsd_svd1.npz")


def _parse(text: str):
    lines = text.splitlines()
    names = [ln.split("\t")[0] for ln in lines[1:]]
    U = np.array([[float(x) for x in ln.split("\t")[2].split()] for ln in lines[1:]])
    return lines[0], names, U


def test_golden_vectors_are_grm_eigenvectors():
    d = _fixture()
    _, _, U = _parse(str(d["text_eigvecs2"]))
    lam, Q = np.linalg.eigh(d["cov"])
    o = np.argsort(-np.abs(lam), kind="stable")
    lam, Q = lam[o], Q[:, o]
    # the well separated leading vectors agree up to sign (the .cov text has 6 digits)
    for j in range(5):
        s = np.sign(U[:, j] @ Q[:, j])
        assert np.abs(U[:, j] - s * Q[:, j]).max() < 1e-5
    # the whole set is an orthonormal eigenbasis of the GRM, |lambda| descending
    assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < 1e-4
    R = d["cov"] @ U
    ev = np.einsum("ij,ij->j", U, R)
    assert np.all(np.diff(np.abs(ev)) <= 1e-4)
    assert np.abs(R - U * ev).max() < 1e-4


@pytest.mark.parametrize("gz", [True, False])
def test_writer_reproduces_reference_text(tmp_path, gz):
    import gzip

    d = _fixture()
    gold = str(d["text_eigvecs2"])
    _, names, U = _parse(gold)
    raw = d["beagle_gz"].tobytes()
    bg = tmp_path / ("g.beagle.gz" if gz else "g.beagle")
    bg.write_bytes(raw if gz else gzip.decompress(raw))
    pa.write_eigvecs2(tmp_path / "g.eigvecs2", bg, U)
    assert (tmp_path / "g.eigvecs2").read_text() == gold


def test_writer_rejects_bad_input(tmp_path):
    (tmp_path / "b.beagle").write_text("marker\tallele1\tallele2\tA\tA\tA\tB\tB\tB\nm1\t0\t1\t1\t0\t0\t1\t0\t0\n")
    with pytest.raises(pa.GpuError):  # 3 samples in U2, 2 in the header
        pa.write_eigvecs2(tmp_path / "b.eigvecs2", tmp_path / "b.beagle", np.eye(3))
    (tmp_path / "c.beagle").write_text("marker\tallele1\tallele2\tA\tA\n")
    with pytest.raises(pa.GpuError):  # columns not a multiple of 3
        pa.write_eigvecs2(tmp_path / "c.eigvecs2", tmp_path / "c.beagle", np.eye(1))
    pa.write_eigvecs2(tmp_path / "b.eigvecs2", tmp_path / "b.beagle", np.array([[1.0, -0.5], [2e-7, 3.0]]))
    assert (tmp_path / "b.eigvecs2").read_text() == "#FID\tIID\tPC1\tPC2\nA\tA\t   1 -0.5\nB\tB\t2e-07     3\n"
