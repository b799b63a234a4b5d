"""CPU-side checks of the C ABI library: it loads, exports exactly what
include/pcaone_gpu.h declares, and its host-only pieces (window schedule,
synthetic data, argument validation) behave.  No GPU compute is called."""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from conftest import ROOT


def header_symbols():
    text = (ROOT / "include" / "pcaone_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcaone_\w+)\s*\(", text)))


def test_library_exports_every_header_symbol(lib):
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/pcaone_gpu.h but not exported"
        assert s in pa.SIGNATURES, f"{s} missing from the ctypes signature table"
    assert set(pa.SIGNATURES) == set(syms)


def test_no_device_reports_error_instead_of_crashing(lib):
    if pa.device_count() > 0:
        pytest.skip("a HIP device is visible")
    h = C.c_void_p()
    assert lib.pcaone_gpu_create(0, C.byref(h)) < 0
    assert not h.value
    # null handles are rejected, not dereferenced
    assert lib.pcaone_gpu_load_bed(None, None, 1, 1) == -1
    assert lib.pcaone_gpu_last_error(None) == b"null handle"


def test_sim_bed_is_deterministic_across_threads():
    a = pa.sim_bed(5, 999, 3000, 3, 0.1, 0.05, 0.5, 0.01, nthreads=1)
    b = pa.sim_bed(5, 999, 3000, 3, 0.1, 0.05, 0.5, 0.01, nthreads=7)
    c = pa.sim_bed(6, 999, 3000, 3, 0.1, 0.05, 0.5, 0.01, nthreads=7)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_sim_bed_statistics():
    n, m, miss = 2000, 4000, 0.01
    pay, panc = pa.sim_bed(9, n, m, 3, 0.1, 0.05, 0.5, miss, want_p=True)
    codes = O.decode_codes(pay, n, m)
    F, cnt = O.allele_freq(codes)
    assert abs(cnt[:, 1].sum() / (n * m) - miss) < 1e-3
    assert np.all((F > 0) & (F < 1))  # monomorphic SNPs are redrawn
    assert np.all((panc >= 0.05) & (panc <= 0.5))
    assert abs(np.corrcoef(F, panc)[0, 1]) > 0.8
    # padding bits of the last byte stay zero (PLINK writes zeros there)
    b = (n + 3) >> 2
    assert n % 4 == 0 or np.all(pay.reshape(m, b)[:, -1] >> (2 * (n % 4)) == 0)


def test_sim_bed_rejects_bad_arguments():
    out = np.empty(10, dtype=np.uint8)
    f = pa.lib().pcaone_sim_bed
    P = C.POINTER(C.c_uint8)
    assert f(1, 0, 10, 1, 0.1, 0.05, 0.5, 0.0, out.ctypes.data_as(P), None, 1) == -1
    assert f(1, 4, 10, 5, 0.1, 0.05, 0.5, 0.0, out.ctypes.data_as(P), None, 1) == -1  # npops > N
    assert f(1, 4, 10, 1, 1.0, 0.05, 0.5, 0.0, out.ctypes.data_as(P), None, 1) == -1  # fst = 1


@pytest.mark.parametrize("bands", [4, 8, 16, 64, 256])
def test_host_window_schedule_matches_oracle(bands):
    st = O.FancyState(bands)
    bs = 1
    for pi in range(int(np.log2(bands)) + 4):
        reset, steps = O.fancy_epoch_plan(st, pi)
        acc, upd, zer, rst, bs = pa.fancy_plan(bands, pi, bs)
        assert rst == reset
        assert list(acc) == [s[1] for s in steps]
        assert list(upd) == [s[2] for s in steps]
        assert list(zer) == [s[3] for s in steps]
        assert bs == st.bandsize


@pytest.mark.parametrize("bands,factor", [(8, 5), (4, 11), (8, 1), (16, 3)])
def test_host_block_schedule_matches_oracle(bands, factor):
    nblocks = bands * factor - (1 if factor > 2 else 0)  # ceil rounding can leave fewer blocks
    st = O.FancyState(bands, nblocks, factor)
    bs = 1
    for pi in range(int(np.log2(bands)) + 4):
        reset, steps = O.fancy_epoch_plan(st, pi)
        acc, upd, zer, rst, bs = pa.fancy_plan(bands, pi, bs, nblocks, factor)
        assert rst == reset
        assert [list(acc), list(upd), list(zer)] == [[s[i] for s in steps] for i in (1, 2, 3)]
        assert bs == st.bandsize


def test_host_window_schedule_rejects_bad_bands():
    bs = C.c_uint64(1)
    a = np.zeros(64, dtype=np.intc)
    P = C.POINTER(C.c_int)
    for bands, nb, f in ((2, 0, 1), (5, 0, 1), (0, 0, 1), (8, 1, 1), (8, 40, 0)):
        assert pa.lib().pcaone_fancy_plan(bands, nb, f, 0, C.byref(bs), a.ctypes.data_as(P), a.ctypes.data_as(P),
                                          a.ctypes.data_as(P), None) == -1


def test_ooc_plan_and_plink_perm_match_oracle():
    from conftest import GOLDEN, load_golden
    from _cases import parse_args
    for name in ("svd2_ooc_w8_m5", "svd2_ooc_w4_d2", "svd2_ooc_w8_f1"):
        z = load_golden(GOLDEN / f"{name}.npz")
        d = parse_args(z["args"])
        n, m = int(z["nsamples"]), int(z["nsnps"])
        assert list(pa.ooc_plan(n, m, 2, d["k"], d["os"], d["w"], d["m"])) == [int(x) for x in z["ooc_plan"]]
        assert np.array_equal(pa.plink_perm(n, m, 2, d["w"]), z["ooc_perm"])
    for args in [(1000, 10 ** 6, 0, 10, 10, 64, 1.0), (1000, 100000, 1, 10, 10, 64, 0.5), (5000, 300000, 2, 20, 20,
                                                                                         64, 0.7)]:
        assert pa.ooc_plan(*args) == O.ooc_plan(args[0], args[1], args[3], args[4], args[2], args[5], args[6])
    for m, nb in [(1000, 8), (1601, 4), (999, 8), (64, 4), (123457, 64)]:
        assert np.array_equal(pa.plink_perm(100, m, 2, nb), O.plink_perm_indices(100, m, 2, nb))
    # a buffer of one GB over 2 MB rows: several buffers and a short last one
    assert np.array_equal(pa.plink_perm(8 * 10 ** 6, 1100, 1, 4), O.plink_perm_indices(8 * 10 ** 6, 1100, 1, 4))
    with pytest.raises(pa.GpuError):  # a single block
        pa.ooc_plan(200, 2000, 1, 5, 5, 8, 1.0)


def test_ld_windows_match_oracle():
    import pcaone_oracle as O
    from conftest import GOLDEN, load_golden
    z = load_golden(GOLDEN / "ld_r2_bp2000.npz")
    for bp in (0, 1, 150, int(z["ld_bp"]), 10 ** 7):
        ws, we = pa.ld_windows(z["pos"], z["chr_end"], bp)
        wo, eo = O.ld_windows(z["pos"], z["chr_end"], bp)
        assert np.array_equal(ws, wo) and np.array_equal(we, eo)
    with pytest.raises(pa.GpuError):
        pa.ld_windows(z["pos"], z["chr_end"][:-1], 100)  # the last chromosome must end at nsnps - 1


def test_write_eigs_reproduces_reference_files(tmp_path):
    """The four text files, byte for byte, from the values the reference wrote
    (already at 6 significant digits, so re-printing them is exact)."""
    from conftest import GOLDEN, load_golden
    z = load_golden(GOLDEN / "svd1_p7.npz")
    m = int(z["nsnps"])
    S = z["sigvals"]
    pa.write_eigs(tmp_path / "o", S, z["eigvecs"], z["loadings"], m, 2)
    for ext in ("sigvals", "eigvecs", "loadings"):
        assert (tmp_path / ("o." + ext)).read_text() == str(z["text_" + ext]), ext
    # eigvals are recomputed from S: equal to 6 significant digits
    ev = np.loadtxt(tmp_path / "o.eigvals")
    assert np.allclose(ev, z["eigvals"], rtol=1e-5)
    assert (tmp_path / "o.eigvals").read_text().count("\n") == len(S)
