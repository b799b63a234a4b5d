"""The CSV reader (pcaone_read_csv, host only) that pcaone_gpu_load_csv runs
before its device transform: zstd and plain text against the counts, the
std::stoi library sizes, the reference's line counting and the errors."""
from pathlib import Path

import numpy as np
import pytest

import pcaone_amd as pa
from _cases import csv_bytes

GOLD = Path(__file__).parent / "golden"


@pytest.mark.parametrize("zst,threads", [(True, 1), (True, 7), (False, 3)])
def test_read_csv_matches_counts(tmp_path, zst, threads):
    counts = np.load(GOLD / "csv_svd1_scale2.npz")["counts"]
    p = tmp_path / "a.csv"
    p.write_bytes(csv_bytes(counts, zstd=zst))
    X, lib = pa.read_csv(p, threads)
    assert X.shape == counts.shape
    assert np.array_equal(X, counts.astype(np.float32))
    assert np.array_equal(lib, counts.sum(axis=0))


def test_read_csv_large_multi_chunk(tmp_path):
    """More text than one 4 MiB decoder output buffer and one 16 MiB parse batch."""
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 100000, (5000, 900))
    p = tmp_path / "big.csv.zst"
    p.write_bytes(csv_bytes(counts))
    X, lib = pa.read_csv(p, 8)
    assert np.array_equal(X, counts.astype(np.float32)) and np.array_equal(lib, counts.sum(axis=0))


def test_read_csv_text_rules(tmp_path):
    p = tmp_path / "b.csv"
    p.write_text("1,2.5,3\n4, 5,-6\n7,8,9")  # no final newline: the last line is not counted
    X, lib = pa.read_csv(p)
    assert X.tolist() == [[1, 2.5, 3], [4, 5, -6]]
    assert lib.tolist() == [5, 7, -3]  # std::stoi: "2.5" -> 2, " 5" -> 5
    for bad in ("1,2,3\n4,5\n", "1,2\n3,x\n", ""):
        p.write_text(bad)
        with pytest.raises(pa.GpuError):
            pa.read_csv(p)
    with pytest.raises(pa.GpuError):
        pa.read_csv(tmp_path / "missing.csv")
