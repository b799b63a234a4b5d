"""The .ld.gz writer (pcaone_write_ld_gz) against the text the reference CLI
wrote: ld_r2_big's lines (src/LD.cpp:419-442) for the ld_stats1_bp3000
fixture, whose SNP fields come from the reference's own .mbim."""
import gzip
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa

GOLD = Path(__file__).parent / "golden"


def _fixture():
    return np.load(GOLD / "ld_stats1_bp3000.npz")


def _windows(d):
    return pa.ld_windows(d["pos"], d["chr_end"], int(d["ld_bp"]))


def _golden_text(d):
    return gzip.decompress(d["ldgz"].tobytes())


def _r2_from_text(text: bytes):
    return np.array([float(line.split(b"\t")[6]) for line in text.splitlines()[1:]])


@pytest.mark.parametrize("level,threads", [(-1, 1), (1, 4), (6, 16)])
def test_writer_reproduces_reference_text(tmp_path, level, threads):
    d = _fixture()
    (tmp_path / "x.mbim").write_bytes(d["mbim"].tobytes())
    ws, we = _windows(d)
    gold = _golden_text(d)
    r2 = _r2_from_text(gold)  # the reference's own values, as printed
    assert np.isnan(r2).sum() == 27
    pa.write_ld_gz(tmp_path / "x.ld.gz", tmp_path / "x.mbim", ws, we, r2, level=level, nthreads=threads)
    got = gzip.decompress((tmp_path / "x.ld.gz").read_bytes())
    assert got == gold


def test_writer_chunks_are_gzip_members(tmp_path):
    """More pairs than one chunk (2^20): several members, read back in order."""
    m = 1500
    bim = "".join(f"{1 + i // 700} rs{i} 0 {10 * i + 5} A C\n" for i in range(m))
    (tmp_path / "b.bim").write_text(bim)
    ws = np.arange(0, m - 1, dtype=np.int32)
    we = np.minimum(m - ws, 1400).astype(np.int32)
    n = int((we - 1).sum())
    assert n > (1 << 20)
    r2 = np.random.default_rng(3).random(n)
    pa.write_ld_gz(tmp_path / "b.ld.gz", tmp_path / "b.bim", ws, we, r2, level=1, nthreads=3)
    lines = gzip.decompress((tmp_path / "b.ld.gz").read_bytes()).decode().splitlines()
    assert lines[0] == "CHR_A\tBP_A\tSNP_A\tCHR_B\tBP_B\tSNP_B\tR2"
    assert len(lines) == n + 1
    p = 0
    for w in (0, 700, m - 2):  # spot checks: first, a chromosome change, last window
        p = int((we[:w] - 1).sum())
        i = ws[w]
        for j in (1, we[w] - 1):
            k = i + j
            want = (f"{1 + i // 700}\t{10 * i + 5}\trs{i}\t{1 + k // 700}\t{10 * k + 5}\trs{k}\t"
                    f"{r2[p + j - 1]:f}")
            assert lines[1 + p + j - 1] == want


def test_writer_rejects_bad_input(tmp_path):
    (tmp_path / "s.bim").write_text("1 rs0 0 10 A C\n1 rs1 0 20 A C\n")
    ws = np.array([0], dtype=np.int32)
    we = np.array([2], dtype=np.int32)
    with pytest.raises(pa.GpuError):  # pair count differs from the windows'
        pa.write_ld_gz(tmp_path / "s.ld.gz", tmp_path / "s.bim", ws, we, np.zeros(2))
    with pytest.raises(pa.GpuError):  # window past the .bim
        pa.write_ld_gz(tmp_path / "s.ld.gz", tmp_path / "s.bim", ws, np.array([3], dtype=np.int32), np.zeros(2))
    with pytest.raises(pa.GpuError):  # missing .bim
        pa.write_ld_gz(tmp_path / "s.ld.gz", tmp_path / "none.bim", ws, we, np.zeros(1))
    (tmp_path / "t.bim").write_text("1 rs0 0\n")
    with pytest.raises(pa.GpuError):  # fewer than 4 fields
        pa.write_ld_gz(tmp_path / "t.ld.gz", tmp_path / "t.bim", ws[:0], we[:0], np.zeros(0))
