"""Out-of-core genotypes (-m): the .bed file mapped in host memory and streamed
window by window through the device ring, against the same file held in HBM,
the oracle and the reference CLI's -m runs."""
import numpy as np
import pytest

import pcaone_amd as pa
import pcaone_oracle as O
from _cases import eig_relerr, oracle_run, parse_args
from conftest import GOLDEN, load_golden

pytestmark = pytest.mark.gpu


def _bed(tmp_path, payload, n, m, name="x"):
    pa.write_plink(tmp_path / name, payload, n, m)
    return tmp_path / (name + ".bed")


def _run(gpu, path, n, m, ooc, maxp, tol, **kw):
    gpu.load_bed_file(path, n, m, out_of_core=ooc)
    F = gpu.allele_freq()
    k, os_ = kw.pop("k"), kw.pop("os")
    gpu.rsvd_init(k, os_, **kw)
    G, H = gpu.gandh(0)
    U, S, V, ep, diffs = gpu.usv(maxp, tol)
    return F, G, H, U, S, V, ep


@pytest.mark.parametrize("variant,arm", [(pa.RSVD_FANCY, pa.PREC_FP64), (pa.RSVD_FANCY, pa.PREC_SPLIT_BF16),
                                         (pa.RSVD_NORMAL, pa.PREC_FP64), (pa.RSVD_NORMAL, pa.PREC_FP32)])
def test_streamed_matches_resident(gpu, tmp_path, variant, arm):
    # N % 4 == 3, a short last window, 3 % missing, and an -m block plan with
    # band factor 3 (24 windows over 8 bands)
    n, m = 1003, 6007
    pay = pa.sim_bed(31, n, m, 4, 0.08, 0.05, 0.5, 0.03)
    path = _bed(tmp_path, pay, n, m)
    bs = -(-m // 24)
    kw = dict(k=6, os=10, variant=variant, bands=8, precision=arm, blocksize=bs, band_factor=3)
    if variant == pa.RSVD_FANCY:
        kw["permutation"] = pa.plink_perm(n, m, 2, 8)
    print(f"streamed vs resident: variant {variant} arm {arm} n {n} m {m} blocksize {bs}", flush=True)
    rs = _run(gpu, path, n, m, True, 6, 0.0, **dict(kw))
    rr = _run(gpu, path, n, m, False, 6, 0.0, **dict(kw))
    for a, b in zip(rs[0], rr[0]):  # F and the code counts from the streamed pass, bit for bit
        assert np.array_equal(a, b)
    assert rs[6] == rr[6]
    tol = 1e-11 if arm == pa.PREC_FP64 else 1e-5
    for a, b in ((rs[1], rr[1]), (rs[2], rr[2])):  # G and H of pi = 0
        assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))
    assert eig_relerr(rs[4], rr[4]) < (1e-12 if arm == pa.PREC_FP64 else 1e-6)
    # streamed launches cover one window each, resident ones all windows: the
    # sums round in another order, and six unconverged epochs divide that by the
    # gap between PCs 3 and 4
    assert O.principal_angles_max(rs[3][:, :3], rr[3][:, :3]) < (1e-6 if arm == pa.PREC_FP64 else 1e-4)


@pytest.mark.parametrize("name", ["svd2_ooc_w8_m5", "svd2_ooc_w4_d2", "svd2_ooc_w8_f1"])
def test_streamed_winsvd_matches_reference_cli(gpu, tmp_path, name):
    """The reference CLI's out-of-core winSVD (-m): its block plan, permute_plink's
    order and its outputs, from a streamed file."""
    z = load_golden(GOLDEN / f"{name}.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    path = _bed(tmp_path, z["payload"], n, m)
    bs, nb, factor = pa.ooc_plan(n, m, 2, d["k"], d["os"], d["w"], d["m"])
    print(f"streamed winSVD {name}: blocksize {bs} nblocks {nb} factor {factor}", flush=True)
    gpu.load_bed_file(path, n, m, out_of_core=True)
    gpu.rsvd_init(d["k"], d["os"], variant=pa.RSVD_FANCY, bands=d["w"], precision=pa.PREC_FP64, seed=d["seed"],
                  permutation=pa.plink_perm(n, m, 2, d["w"]), blocksize=bs, band_factor=factor)
    U, S, V, ep, diffs = gpu.usv(d["p"], d["tol"])
    ev = S ** 2 / m * 2
    assert eig_relerr(ev, z["eigvals"]) < 5e-6  # 6 printed digits
    npc = 3 if m == 1601 else 2
    assert O.principal_angles_max(U[:, :npc], z["eigvecs"][:, :npc]) < 1e-4
    assert O.principal_angles_max(V[:, :npc], z["loadings"][:, :npc]) < 1e-4
    orc, epo, _ = oracle_run(z)
    assert ep == epo
    assert eig_relerr(ev, orc.eigvals()) < 1e-10
    n_d = min(len(orc.diffs), d["p"])
    assert np.allclose(diffs[1:1 + n_d], orc.diffs[:n_d], rtol=1e-8, atol=1e-12)


def test_streamed_emu_matches_reference_cli(gpu, tmp_path):
    """EMU through the reference's out-of-core reader (-m 0.0005, --svd 1)."""
    z = load_golden(GOLDEN / "emu_svd1.npz")
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    path = _bed(tmp_path, z["payload"], n, m)
    bs, nb, _ = pa.ooc_plan(n, m, 1, d["k"], d["os"], 64, d["m"])
    print(f"streamed EMU: blocksize {bs} nblocks {nb}", flush=True)
    gpu.load_bed_file(path, n, m, out_of_core=True)
    gpu.rsvd_init(d["k"], d["os"], variant=pa.RSVD_NORMAL, precision=pa.PREC_FP64, seed=d["seed"], blocksize=bs)
    U, S, V, it, diffs = gpu.em(d["p"], d["tol"], 100, 1e-5)
    assert eig_relerr(S ** 2 / m * 2, z["eigvals"]) < 5e-6
    assert O.principal_angles_max(U[:, :2], z["eigvecs"][:, :2]) < 1e-4


def test_streamed_gram_matvec_and_iram(gpu, tmp_path):
    n, m = 517, 3001
    pay = pa.sim_bed(37, n, m, 3, 0.1, 0.05, 0.5, 0.02)
    path = _bed(tmp_path, pay, n, m)
    rng = np.random.default_rng(5)
    xs, xm = rng.standard_normal(n), rng.standard_normal(m)
    out = {}
    for ooc in (True, False):
        gpu.load_bed_file(path, n, m, out_of_core=ooc)
        gpu.gram_init(pa.PREC_FP64)
        out[ooc] = (gpu.gram_matvec(0, xs), gpu.gram_matvec(1, xm), gpu.iram(4))
    for a, b in zip(out[True][:2], out[False][:2]):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    assert eig_relerr(out[True][2][1], out[False][2][1]) < 1e-10


def test_streamed_refusals_and_file_errors(gpu, tmp_path):
    n, m = 64, 200
    pay = pa.sim_bed(3, n, m, 2, 0.1, 0.05, 0.5, 0.0)
    path = _bed(tmp_path, pay, n, m)
    gpu.load_bed_file(path, n, m, out_of_core=True)
    with pytest.raises(pa.GpuError) as e:
        gpu.gl_from_bed()
    assert e.value.code == -3 and "resident" in str(e.value)
    with pytest.raises(pa.GpuError):
        gpu.ld_usv(np.ones((n, 1)))
    # a resident load after a streamed one works again
    gpu.load_bed_file(path, n, m, out_of_core=False)
    gpu.gl_from_bed()
    bad = tmp_path / "bad.bed"
    bad.write_bytes(bytes([0x6C, 0x1B, 0x00]) + bytes(m * 16))
    with pytest.raises(pa.GpuError, match="magic"):
        gpu.load_bed_file(bad, n, m, out_of_core=True)
    short = tmp_path / "short.bed"
    short.write_bytes(bytes([0x6C, 0x1B, 0x01]) + bytes(m * 16 - 1))
    with pytest.raises(pa.GpuError):
        gpu.load_bed_file(short, n, m, out_of_core=True)
    with pytest.raises(pa.GpuError):
        gpu.load_bed_file(tmp_path / "missing.bed", n, m, out_of_core=True)
