"""Writes the golden fixtures in tests/golden/ from the reference PCAone CLI.

Run in the build container only (it needs oracle/_ref/PCAone, built from the
reference sources by `make -f oracle/ref.mk`):

    python tests/golden/make_golden.py [case ...]

Inputs are seeded Balding-Nichols PLINK sets from pcaone_sim_bed (and dense,
likelihood and LD inputs); each case stores its input and the reference's
printed outputs (.eigvals, .sigvals, .eigvecs, .loadings with 6 significant
digits, src/Data.cpp:177-194; .cov; .ld.gz with 6 decimals) in one .npz.  The reference is run with -n 1 so its output does not depend on
OpenMP scheduling.
"""
from __future__ import annotations

import gzip
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import pcaone_amd as pa  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "PCAone"
OUT = Path(__file__).resolve().parent

DATASETS = {
    # name: (seed, N, M, npops, fst, miss)
    "d1": (11, 200, 2000, 3, 0.1, 0.01),
    "d2": (12, 203, 1601, 4, 0.05, 0.03),   # N % 4 == 3, uneven last window
    "d3": (13, 200, 2000, 3, 0.1, 0.05),    # 5 % missing for EMU
}

# larger sets regenerated from the simulator's seed (not stored): name -> (seed, N, M, npops, fst, miss)
BIG = {
    "c1": (2027, 2000, 20000, 3, 0.1, 0.01),   # the CPU-reference-parity workload
    "w64": (2028, 1000, 24000, 5, 0.05, 0.005),  # the full 64-window winSVD schedule
}
BIG_CASES = {
    "c1_svd1_p7": ("c1", "--svd 1 -k 10 --oversamples 10 -p 7 --tol-rsvd 0 --seed 101"),
    "w64_svd2": ("w64", "--svd 2 -k 10 --oversamples 10 -w 64"),
}

CASES = {
    # name: (dataset, CLI args)
    "svd1_p7": ("d1", "--svd 1 -k 5 --oversamples 5 -p 7 --tol-rsvd 0 --seed 101"),
    "svd1_p1": ("d1", "--svd 1 -k 5 --oversamples 5 -p 1 --tol-rsvd 0 --seed 101"),
    "svd1_p1_unif": ("d1", "--svd 1 -k 5 --oversamples 5 -p 1 --tol-rsvd 0 --seed 7 --rand 0"),
    "svd1_p3_nostd": ("d1", "--svd 1 -k 4 --oversamples 6 -p 3 --tol-rsvd 0 --scale 0"),
    "svd2_w8": ("d1", "--svd 2 -k 5 --oversamples 5 -w 8"),
    "svd2_w8_S": ("d1", "--svd 2 -k 5 --oversamples 5 -w 8 -S"),
    "svd2_w4_p2": ("d2", "--svd 2 -k 3 --oversamples 7 -w 4 -p 2 --tol-rsvd 0 --seed 3"),
    "svd1_d2": ("d2", "--svd 1 -k 6 --oversamples 6 -p 5 --tol-rsvd 0"),
    # EMU through the out-of-core reader (-m): the consistent EM path of the
    # reference (the in-core one indexes a permutation that --svd 1 never sets)
    "emu_svd1": ("d3", "--svd 1 -m 0.0005 --emu -k 3 --oversamples 7 -p 5 --tol-rsvd 0"),
    # winSVD out of core (-m GB): the block plan of src/Data.cpp:27-71 (blocksize,
    # nblocks, band factor 5 / 11 / 1) and permute_plink's SNP order
    "svd2_ooc_w8_m5": ("d1", "--svd 2 -k 5 --oversamples 5 -w 8 -m 0.0005"),
    "svd2_ooc_w4_d2": ("d2", "--svd 2 -k 3 --oversamples 7 -w 4 -p 2 --tol-rsvd 0 --seed 3 -m 0.0004"),
    "svd2_ooc_w8_f1": ("d1", "--svd 2 -k 5 --oversamples 5 -w 8 -m 0.001"),
    # PCAngsd on in-core PLINK hard calls (--svd 1: no shuffle, consistent indices)
    "plink_pcangsd_svd1": ("d3", "--svd 1 --pcangsd -k 3 --oversamples 7 -p 5 --tol-rsvd 0"),
}

# dense float32 .bin inputs (scRNA-like log1p CPM): name -> (seed, ncells, ngenes, base)
DENSE = {"b1": (21, 300, 410, 0.035)}
# genotype likelihoods (BEAGLE): name -> (seed, nsamples, nsites, npops, fst, depth, err)
GL = {"g1": (31, 120, 700, 4, 0.1, 2.0, 0.01)}
GL_CASES = {
    # PCAngsd EM; --svd 1 keeps the reference's in-core indices consistent
    "gl_pcangsd_svd1": ("g1", "--svd 1 -k 3 --oversamples 7 -p 5 --tol-rsvd 0"),
    # the same EM with IRAM (run_pca_with_arnoldi's in-core EM branch)
    "gl_pcangsd_svd0": ("g1", "--svd 0 -k 3"),
}
DENSE_CASES = {
    "bin_svd1_p5": ("b1", "--svd 1 -k 5 --oversamples 5 -p 5 --tol-rsvd 0"),
    "bin_svd2_w8": ("b1", "--svd 2 -k 5 --oversamples 5 -w 8"),
    "bin_svd0": ("b1", "--svd 0 -k 5"),
}


# LD r^2 on a dense residual matrix (-B .residuals --match-bim .mbim --print-r2):
# name -> (seed, nsamples, SNPs per chromosome, ld_bp)
LD = {"ld_r2_bp2000": (41, 150, (300, 250, 351), 2000)}
# LD with --USV on PLINK input (a PCA run for .eigvecs, then -b --USV --print-r2):
# name -> (seed, nsamples, SNPs per chromosome, ld_bp, k)
LD_USV = {"ld_usv_bp3000": (51, 160, (400, 350, 450), 3000, 3)}


def ld_inputs(seed, n, per_chr):
    """Residual-like float32 rows with LD decaying along the chromosome (an AR(1)
    chain of SNPs) and sorted positions with 1..400 bp gaps (one duplicated
    position inside chromosome 1)."""
    rng = np.random.default_rng(seed)
    m = int(sum(per_chr))
    X = np.empty((m, n))
    pos = np.empty(m, dtype=np.int64)
    chr_end = np.cumsum(per_chr) - 1
    start = 0
    for c, cnt in enumerate(per_chr):
        x = rng.standard_normal(n)
        p = int(rng.integers(1, 1000))
        for t in range(cnt):
            rho = rng.uniform(0.3, 0.95)
            x = rho * x + np.sqrt(1 - rho * rho) * rng.standard_normal(n)
            X[start + t] = x
            pos[start + t] = p
            p += int(rng.integers(1, 400))
        start += cnt
    pos[10] = pos[9]  # duplicated position inside chromosome 1
    X -= X.mean(axis=1, keepdims=True)
    return X.astype(np.float32), pos, chr_end


def run_ld(tmp: Path, name: str, X, pos, chr_end, ld_bp):
    pa.write_bin(tmp / (name + ".residuals"), X)
    chrom = np.searchsorted(chr_end, np.arange(X.shape[0])) + 1
    with open(tmp / (name + ".mbim"), "w") as f:
        for i in range(X.shape[0]):
            f.write(f"{chrom[i]}\trs{i}\t0\t{pos[i]}\tA\tG\t0.25\n")
    cmd = [str(REF), "-B", name + ".residuals", "--match-bim", name + ".mbim", "--print-r2", "--ld-bp",
           str(ld_bp), "-o", name, "-n", "1", "-v", "0"]
    subprocess.run(cmd, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    import gzip
    ia, ib, r2 = [], [], []
    with gzip.open(tmp / (name + ".ld.gz"), "rt") as f:
        f.readline()
        for line in f:
            t = line.rstrip("\n").split("\t")
            ia.append(int(t[2][2:]))
            ib.append(int(t[5][2:]))
            r2.append(float(t[6]))
    return np.array(ia), np.array(ib), np.array(r2)


def read_ld_gz(path: Path):
    import gzip
    ia, ib, r2 = [], [], []
    with gzip.open(path, "rt") as f:
        f.readline()
        for line in f:
            t = line.rstrip("\n").split("\t")
            ia.append(int(t[2][2:]))
            ib.append(int(t[5][2:]))
            r2.append(float(t[6]))
    return np.array(ia), np.array(ib), np.array(r2)


def run_ld_usv(tmp: Path, name: str, seed, n, per_chr, ld_bp, k):
    """Genotypes with 2 % missing on len(per_chr) chromosomes (sorted positions,
    1..400 bp gaps, one duplicated position), the reference's PCA for U, then
    its --USV LD."""
    m = int(sum(per_chr))
    pay = pa.sim_bed(seed, n, m, 3, 0.1, 0.05, 0.5, 0.02)
    rng = np.random.default_rng(seed)
    chrom = np.repeat(np.arange(1, len(per_chr) + 1), per_chr)
    pos = np.concatenate([np.cumsum(rng.integers(1, 400, c)) for c in per_chr]).astype(np.int64)
    pos[10] = pos[9]
    pa.write_plink(tmp / name, pay, n, m, chrom, pos)
    pca = [str(REF), "-b", name, "--svd", "1", "-k", str(k), "-p", "7", "--tol-rsvd", "0", "-V", "-o",
           name + "_pca", "-n", "1", "-v", "0"]  # -V also writes the .mbim that --USV reads
    subprocess.run(pca, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    ld = [str(REF), "-b", name, "--USV", name + "_pca", "--print-r2", "--ld-bp", str(ld_bp), "-o", name, "-n", "1",
          "-v", "0"]
    subprocess.run(ld, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    U = np.loadtxt(tmp / (name + "_pca.eigvecs"), ndmin=2)
    ia, ib, r2 = read_ld_gz(tmp / (name + ".ld.gz"))
    chr_end = np.cumsum(per_chr) - 1
    return dict(payload=pay, nsamples=n, nsnps=m, pos=pos, chr_end=chr_end, ld_bp=ld_bp, U=U, snp_a=ia, snp_b=ib,
                r2=r2)


# LD on PLINK input through the reference's two steps: a PCA run with --ld
# --ld-stats 1 writes the centred genotypes (.residuals, .mbim), then -B
# .residuals --print-r2.  name -> (seed, nsamples, SNPs per chromosome, ld_bp)
LD_STATS1 = {"ld_stats1_bp3000": (61, 203, (420, 380, 500), 3000)}


def run_ld_stats1(tmp: Path, name: str, seed, n, per_chr, ld_bp):
    """Genotypes with 2 % missing (N % 4 == 3), one monomorphic SNP (its r^2 is
    NaN: sd = 0), sorted positions with 1..400 bp gaps and one duplicated
    position on len(per_chr) chromosomes."""
    m = int(sum(per_chr))
    pay = pa.sim_bed(seed, n, m, 3, 0.1, 0.05, 0.5, 0.02).reshape(m, -1).copy()
    pay[25, :] = 0x00  # every sample hom A1
    pay = pay.reshape(-1)
    rng = np.random.default_rng(seed)
    chrom = np.repeat(np.arange(1, len(per_chr) + 1), per_chr)
    pos = np.concatenate([np.cumsum(rng.integers(1, 400, c)) for c in per_chr]).astype(np.int64)
    pos[40] = pos[39]
    pa.write_plink(tmp / name, pay, n, m, chrom, pos)
    res = [str(REF), "-b", name, "--svd", "1", "-k", "2", "-p", "3", "--ld", "--ld-stats", "1", "-o",
           name + "_res", "-n", "1", "-v", "0"]
    subprocess.run(res, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    ld = [str(REF), "-B", name + "_res.residuals", "--match-bim", name + "_res.mbim", "--print-r2", "--ld-bp",
          str(ld_bp), "-o", name, "-n", "1", "-v", "0"]
    subprocess.run(ld, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    raw = np.fromfile(tmp / (name + "_res.residuals"), dtype=np.uint32, count=2)
    resid = np.fromfile(tmp / (name + "_res.residuals"), dtype=np.float32, offset=8).reshape(int(raw[0]),
                                                                                           int(raw[1]))
    ia, ib, r2 = read_ld_gz(tmp / (name + ".ld.gz"))
    chr_end = np.cumsum(per_chr) - 1
    # the file bytes as well: the .ld.gz as the reference wrote it and the
    # .mbim its SNP fields come from (the text of both is the fixture)
    ldgz = np.frombuffer((tmp / (name + ".ld.gz")).read_bytes(), dtype=np.uint8)
    mbim = np.frombuffer((tmp / (name + "_res.mbim")).read_bytes(), dtype=np.uint8)
    return dict(payload=pay, nsamples=n, nsnps=m, pos=pos, chr_end=chr_end, ld_bp=ld_bp, resid=resid, snp_a=ia,
                snp_b=ib, r2=r2, ldgz=ldgz, mbim=mbim)


# CSV of counts (--csv, zstd-compressed): name -> (seed, ncells, ngenes)
CSV = {"c1": (41, 260, 380)}
CSV_CASES = {
    "csv_svd1_scale2": ("c1", "--svd 1 -k 4 --oversamples 6 -p 12 --tol-rsvd 0 --scale 2"),
    "csv_svd1_scale1": ("c1", "--svd 1 -k 4 --oversamples 6 -p 12 --tol-rsvd 0 --scale 1"),
}


def zstd_compress(data: bytes) -> bytes:
    """One zstd frame (the system's libzstd, level 3)."""
    import ctypes

    z = ctypes.CDLL("libzstd.so.1")
    z.ZSTD_compressBound.restype = ctypes.c_size_t
    z.ZSTD_compress.restype = ctypes.c_size_t
    z.ZSTD_isError.restype = ctypes.c_uint
    cap = z.ZSTD_compressBound(ctypes.c_size_t(len(data)))
    dst = ctypes.create_string_buffer(cap)
    n = z.ZSTD_compress(dst, ctypes.c_size_t(cap), data, ctypes.c_size_t(len(data)), 3)
    assert not z.ZSTD_isError(ctypes.c_size_t(n))
    return dst.raw[:n]


def csv_counts(seed, ncells, ngenes):
    """Integer counts, genes x cells: negative binomial (gamma-Poisson) around
    per-gene means, with 3 cell groups shifting a third of the genes."""
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.normal(0.0, 1.2, ngenes)) * 0.8
    grp = rng.integers(0, 3, ncells)
    shift = np.ones((ngenes, 3))
    shift[rng.random(ngenes) < 0.33, :] = np.exp(rng.normal(0, 0.8, (1, 3)))
    lib = np.exp(rng.normal(0.0, 0.4, ncells))
    lam = mu[:, None] * shift[:, grp] * lib[None, :]
    return rng.poisson(rng.gamma(2.0, lam / 2.0)).astype(np.int64)


# BGEN inputs (pcaone_sim_bgen): name -> (seed, N, M, npops, fst, miss, layout, compression, bits,
# declared minimum ploidy, sample identifier block)
BGEN = {
    "bg8": (81, 230, 1000, 3, 0.1, 0.02, 2, 1, 8, 2, True),      # zlib, the 8-bit byte-pair path
    "bg16z": (82, 150, 600, 3, 0.1, 0.02, 2, 2, 16, 1, False),   # zstd, 16 bits, ploidy range 1-2
    "bg5": (83, 205, 500, 4, 0.08, 0.0, 2, 0, 5, 2, False),      # uncompressed, 5 bits
    "bg1": (84, 180, 500, 3, 0.1, 0.02, 1, 1, 16, 2, False),     # layout 1 (v1.1), zlib
}
# No --pcangsd case: the reference's BGEN reader writes the probabilities into a
# P it never sizes (src/FileBgen.cpp:72-80) and the CLI ends with SIGSEGV.
BGEN_CASES = {
    "bgen_svd1_8": ("bg8", "--svd 1 -k 4 --oversamples 6 -p 5 --tol-rsvd 0"),
    "bgen_ld1_8": ("bg8", "--svd 1 -k 3 --oversamples 7 -p 1 --tol-rsvd 0 --ld --ld-stats 1"),
    "bgen_svd1_16z": ("bg16z", "--svd 1 -k 3 --oversamples 7 -p 6 --tol-rsvd 0"),
    "bgen_ld1_16z": ("bg16z", "--svd 1 -k 3 --oversamples 7 -p 1 --tol-rsvd 0 --ld --ld-stats 1"),
    "bgen_svd2_5": ("bg5", "--svd 2 -k 3 --oversamples 7 -w 8 -S"),
    "bgen_ld1_5": ("bg5", "--svd 1 -k 3 --oversamples 7 -p 1 --tol-rsvd 0 --ld --ld-stats 1"),
    "bgen_svd0_l1": ("bg1", "--svd 0 -k 4"),
    "bgen_ld1_l1": ("bg1", "--svd 1 -k 3 --oversamples 7 -p 1 --tol-rsvd 0 --ld --ld-stats 1"),
    "bgen_nostd_8": ("bg8", "--svd 1 -k 3 --oversamples 7 -p 4 --tol-rsvd 0 --scale 0"),
}

LD_TOOLS = {"ld_tools_bp3000": (71, 157, (230, 170), 3000)}
# --project 1/2 with --USV of a reference PCA: name -> (seed, nref, ntarget, nsnps)
PROJECT = {"project_p12": (61, 220, 90, 1500)}


def _pack_codes(codes: np.ndarray) -> np.ndarray:
    """codes (M x N, .bed 2-bit codes) -> .bed payload rows (missing padding)."""
    m, n = codes.shape
    nb = (n + 3) // 4
    c = np.full((m, nb * 4), 1, dtype=np.uint8)
    c[:, :n] = codes
    c = c.reshape(m, nb, 4)
    return (c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)).astype(np.uint8).reshape(-1)


def run_project(tmp: Path, name: str, seed, nref, ntgt, m):
    """One simulation split into a reference panel and target samples. The
    target .bim leaves out some SNPs, swaps the alleles of others (their codes
    flipped 00 <-> 11 to match) and moves a few to another position; the
    reference PCA is written with -V, then --USV of it projects the target with
    --project 1 and --project 2."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pcaone_oracle as O

    n = nref + ntgt
    pay = pa.sim_bed(seed, n, m, 3, 0.1, 0.05, 0.5, 0.02)
    codes = O.decode_codes(pay, n, m)
    rng = np.random.default_rng(seed)
    ref_codes, tgt_codes = codes[:, :nref], codes[:, nref:]
    pos = (np.arange(m) + 1) * 100
    pa.write_plink(tmp / (name + "_ref"), _pack_codes(ref_codes), nref, m, None, pos)
    keep = rng.random(m) > 0.05
    swap = (rng.random(m) < 0.04) & keep
    moved = (rng.random(m) < 0.01) & keep & ~swap
    t = tgt_codes[keep].copy()
    sw = swap[keep]
    t[sw] = np.where(t[sw] == 0, 3, np.where(t[sw] == 3, 0, t[sw]))
    idx = np.flatnonzero(keep)
    with open(tmp / (name + "_tgt.bim"), "w") as f:
        for i in idx:
            a1, a2 = ("G", "A") if swap[i] else ("A", "G")
            f.write(f"1\trs{i}\t0\t{pos[i] + (7 if moved[i] else 0)}\t{a1}\t{a2}\n")
    with open(tmp / (name + "_tgt.bed"), "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]))
        f.write(_pack_codes(t).tobytes())
    with open(tmp / (name + "_tgt.fam"), "w") as f:
        for j in range(ntgt):
            f.write(f"t{j} t{j} 0 0 0 -9\n")

    def run(*args):
        subprocess.run([str(REF), *args, "-n", "1", "-v", "0"], check=True, cwd=tmp, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)

    run("-b", name + "_ref", "--svd", "1", "-k", "3", "-p", "6", "--tol-rsvd", "0", "-V", "-o", name + "_pca")
    res = dict(tgt_payload=_pack_codes(t), ntarget=ntgt, ntarget_snps=int(keep.sum()),
               tgt_bim=(tmp / (name + "_tgt.bim")).read_text(), mbim=(tmp / (name + "_pca.mbim")).read_text(),
               S=np.loadtxt(tmp / (name + "_pca.sigvals"), ndmin=1, skiprows=1),
               V=np.loadtxt(tmp / (name + "_pca.loadings"), ndmin=2))
    for meth in (1, 2):
        run("-b", name + "_tgt", "--USV", name + "_pca", "--project", str(meth), "-o", f"{name}_m{meth}")
        res[f"U_m{meth}"] = np.loadtxt(tmp / f"{name}_m{meth}.eigvecs", ndmin=2)
        res[f"text_eigvals_m{meth}"] = (tmp / f"{name}_m{meth}.eigvals").read_text()
    return res


# BEAGLE likelihood projection onto a PLINK reference PCA: name -> (seed, nref, ntarget, nsnps)
PROJECT_GL = {"project_gl": (63, 220, 90, 1200)}


def run_project_gl(tmp: Path, name: str, seed, nref, ntgt, m):
    """A reference PCA (-V) on a PLINK panel, then the other samples of the same
    simulation as BEAGLE likelihoods (read depth ~ Poisson(4), 1% error) on
    most of its SNPs: markers by ID or as chromosome_position, alleles as
    letters or as 0-3 codes, some swapped (alleles and triples reversed), a
    few likelihood triples NaN (missing), some sites unmatched (an unknown
    marker, a third allele); then -G --USV --project 2, which
    runs the likelihood EM (run_projection_beagle_gl)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pcaone_oracle as O

    n = nref + ntgt
    pay = pa.sim_bed(seed, n, m, 3, 0.1, 0.05, 0.5, 0.02)
    codes = O.decode_codes(pay, n, m)
    rng = np.random.default_rng(seed)
    pos = (np.arange(m) + 1) * 100
    pa.write_plink(tmp / (name + "_ref"), _pack_codes(codes[:, :nref]), nref, m, None, pos)
    keep = rng.random(m) > 0.05
    swap = (rng.random(m) < 0.1) & keep
    by_pos = rng.random(m) < 0.2
    numeric = rng.random(m) < 0.5
    ncopy = np.select([codes == 0, codes == 2, codes == 3], [2, 1, 0], -1)  # copies of A (allele 1), -1 missing
    rows, sites = [], []
    for i in np.flatnonzero(keep):
        gl = np.empty((ntgt, 3))
        for j in range(ntgt):
            g = ncopy[i, nref + j]
            d = rng.poisson(4.0) if g >= 0 else 0
            lik = np.ones(3)
            for _ in range(d):
                b = rng.random() < g / 2.0  # the read carries A
                if rng.random() < 0.01:
                    b = not b
                for h, q in enumerate((1.0, 0.5, 0.0)):  # genotypes AA, AG, GG: fraction of A
                    lik[h] *= q * 0.99 + (1 - q) * 0.01 if b else (1 - q) * 0.99 + q * 0.01
            gl[j] = lik / lik.sum()
        a1, a2 = ("2", "0") if numeric[i] else ("G", "A")
        if not swap[i]:
            a1, a2 = a2, a1
        else:
            gl = gl[:, ::-1]
        gl = gl.astype(np.float32)
        gl[rng.random(ntgt) < 0.003] = np.nan
        marker = f"1_{pos[i]}" if by_pos[i] else f"rs{i}"
        if rng.random() < 0.02:  # unknown marker: unmatched
            marker = f"x{i}"
        if rng.random() < 0.03:  # another allele: unmatched
            a2 = "1" if numeric[i] else "C"
        rows.append(gl)
        sites.append((marker, a1, a2))
    glm = np.stack(rows)  # sites x samples x 3
    with gzip.open(tmp / (name + "_tgt.beagle.gz"), "wt") as f:
        f.write("marker\tallele1\tallele2\t" + "\t".join(f"t{j}\tt{j}\tt{j}" for j in range(ntgt)) + "\n")
        for (mk, a1, a2), gl in zip(sites, glm):
            f.write(f"{mk}\t{a1}\t{a2}\t" + "\t".join(repr(float(v)) for v in gl.reshape(-1)) + "\n")

    def run(*args):
        subprocess.run([str(REF), *args, "-n", "1", "-v", "0"], check=True, cwd=tmp, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)

    run("-b", name + "_ref", "--svd", "1", "-k", "3", "-p", "6", "--tol-rsvd", "0", "-V", "-o", name + "_pca")
    run("-G", name + "_tgt.beagle.gz", "--USV", name + "_pca", "--project", "2", "-o", name + "_out")
    log = (tmp / (name + "_out.log")).read_text()
    return dict(gl=glm, sites=np.array(["\t".join(s) for s in sites]),
                beagle_gz=np.frombuffer((tmp / (name + "_tgt.beagle.gz")).read_bytes(), dtype=np.uint8),
                mbim=(tmp / (name + "_pca.mbim")).read_text(),
                S=np.loadtxt(tmp / (name + "_pca.sigvals"), ndmin=1, skiprows=1),
                V=np.loadtxt(tmp / (name + "_pca.loadings"), ndmin=2),
                U=np.loadtxt(tmp / (name + "_out.eigvecs"), ndmin=2),
                text_eigvals=(tmp / (name + "_out.eigvals")).read_text(),
                diffs=np.array([float(x) for x in re.findall(r"GL projection iter =\s*\d+, diff =\s*([0-9.eE+-]+)",
                                                             log)]))


# --inbreed 1 with --USV of a PCA on the same PLINK set: name -> dataset
INBREED = {"inbreed_d1": "d1"}


def run_inbreed(tmp: Path, name: str, ds: str):
    """A PCA with -V (.eigvecs, .sigvals, .loadings, .mbim), then --USV of it
    with --inbreed 1 (.hwe)."""
    def run(*args):
        subprocess.run([str(REF), *args, "-n", "1", "-v", "0"], check=True, cwd=tmp, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)

    run("-b", ds, "--svd", "1", "-k", "3", "-p", "5", "--tol-rsvd", "0", "-V", "-o", name + "_pca")
    run("-b", ds, "--USV", name + "_pca", "--inbreed", "1", "-o", name)
    return dict(U=np.loadtxt(tmp / (name + "_pca.eigvecs"), ndmin=2),
                S=np.loadtxt(tmp / (name + "_pca.sigvals"), ndmin=1, skiprows=1),
                V=np.loadtxt(tmp / (name + "_pca.loadings"), ndmin=2),
                mbim=(tmp / (name + "_pca.mbim")).read_text(), hwe=(tmp / (name + ".hwe")).read_text())


def run_ld_tools(tmp: Path, name: str, seed, n, per_chr, ld_bp):
    """The LD consumers and producers on one PLINK set (3 % missing, one
    monomorphic SNP, 1..400 bp gaps, two chromosomes):
    - --ld (ld-stats 0) after a PCA: .residuals, .mbim and the U, S, V it used;
    - --ld-r2 pruning on those residuals with the .mbim (MAF choice), with a
      plain .bim (random choice), and on the .bed with --USV of that PCA;
    - --clump of an assoc-like file (CHR, SNP, BP, P; distinct p-values, a few
      SNPs left out) on the residuals."""
    m = int(sum(per_chr))
    pay = pa.sim_bed(seed, n, m, 3, 0.1, 0.05, 0.5, 0.03).reshape(m, -1).copy()
    rng = np.random.default_rng(seed)
    # LD: most SNPs copy their left neighbour's codes for ~85 % of the samples
    nb = pay.shape[1]
    codes = np.stack([(pay >> (2 * b)) & 3 for b in range(4)], axis=2).reshape(m, 4 * nb)
    for i in range(1, m):
        if rng.random() < 0.7:
            same = rng.random(4 * nb) < 0.85
            codes[i, same] = codes[i - 1, same]
    codes = codes.reshape(m, nb, 4)
    pay = np.zeros((m, nb), dtype=np.uint8)
    for b in range(4):
        pay |= (codes[:, :, b].astype(np.uint8) << (2 * b))
    pay[:, -1] &= np.uint8((1 << (2 * (n - 4 * (nb - 1)))) - 1) if n % 4 else np.uint8(0xFF)
    pay[:, -1] |= np.uint8(0x55 & ~((1 << (2 * (n - 4 * (nb - 1)))) - 1) & 0xFF) if n % 4 else np.uint8(0)
    pay[17, :] = 0xFF  # every sample hom A2
    pay = pay.reshape(-1)
    chrom = np.repeat(np.arange(1, len(per_chr) + 1), per_chr)
    pos = np.concatenate([np.cumsum(rng.integers(1, 400, c)) for c in per_chr]).astype(np.int64)
    pa.write_plink(tmp / name, pay, n, m, chrom, pos)

    def run(*args):
        subprocess.run([str(REF), *args, "-n", "1", "-v", "0"], check=True, cwd=tmp, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)

    run("-b", name, "--svd", "1", "-k", "3", "-p", "4", "--tol-rsvd", "0", "--ld", "-V", "-o", name + "_r0")
    raw = np.fromfile(tmp / (name + "_r0.residuals"), dtype=np.uint32, count=2)
    resid0 = np.fromfile(tmp / (name + "_r0.residuals"), dtype=np.float32, offset=8).reshape(int(raw[0]), int(raw[1]))
    res = dict(payload=pay, nsamples=n, nsnps=m, pos=pos, chrom=chrom, ld_bp=ld_bp, resid0=resid0,
               U=np.loadtxt(tmp / (name + "_r0.eigvecs"), ndmin=2),
               S=np.loadtxt(tmp / (name + "_r0.sigvals"), ndmin=1, skiprows=1),
               V=np.loadtxt(tmp / (name + "_r0.loadings"), ndmin=2),
               bim=(tmp / (name + ".bim")).read_text(), mbim=(tmp / (name + "_r0.mbim")).read_text())
    r0 = name + "_r0.residuals"
    run("-B", r0, "--match-bim", name + "_r0.mbim", "--ld-r2", "0.2", "--ld-bp", str(ld_bp), "-o", name + "_pm")
    run("-B", r0, "--match-bim", name + ".bim", "--ld-r2", "0.2", "--ld-bp", str(ld_bp), "-o", name + "_pb")
    run("-b", name, "--USV", name + "_r0", "--ld-r2", "0.2", "--ld-bp", str(ld_bp), "-o", name + "_pp")
    for tag in ("pm", "pb", "pp"):
        for kind in ("in", "out"):
            res[f"prune_{tag}_{kind}"] = (tmp / f"{name}_{tag}.ld.prune.{kind}").read_text()
    keep = rng.random(m) > 0.1
    pv = rng.permutation(m) + 1.0
    pv = pv / (m + 1.0) * 10 ** rng.uniform(-6, 0, m)  # distinct p-values over 6 decades
    lines = ["CHR\tSNP\tBP\tP"] + [f"{chrom[i]}\trs{i}\t{pos[i]}\t{pv[i]:.6g}" for i in range(m) if keep[i]]
    (tmp / "assoc.txt").write_text("\n".join(lines) + "\n")
    run("-B", r0, "--match-bim", name + "_r0.mbim", "--clump", "assoc.txt", "--clump-p1", "0.001", "--clump-p2",
        "0.05", "--clump-r2", "0.1", "--clump-bp", "2000", "-o", name + "_cl")
    res["assoc"] = (tmp / "assoc.txt").read_text()
    res["clump"] = (tmp / (name + "_cl.p0.clump")).read_text()
    return res


def run_case(tmp: Path, prefix: str, args: str, out: str, flag: str = "-b"):
    ooc_perm = "-m" in args.split() and "--svd 2" in args and "-S" not in args.split()
    # -v 3 keeps the permuted copy the out-of-core winSVD writes (src/Main.cpp:151-162)
    iram_em = flag == "-G" and "--svd 0" in args
    # -v 1 logs the IRAM EM's RMSE per iteration (src/Arnoldi.cpp:77-78)
    cmd = [str(REF), flag, prefix, *args.split(), "-o", out, "-V", "-n", "1", "-v",
           "3" if ooc_perm else "1" if iram_em else "0"]
    subprocess.run(cmd, check=True, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    res = {}
    res["eigvals"] = np.loadtxt(tmp / (out + ".eigvals"), ndmin=1)
    with open(tmp / (out + ".sigvals")) as f:
        f.readline()
        res["sigvals"] = np.loadtxt(f, ndmin=1)
    res["eigvecs"] = np.loadtxt(tmp / (out + ".eigvecs"), ndmin=2)
    res["loadings"] = np.loadtxt(tmp / (out + ".loadings"), ndmin=2)
    for ext in ("sigvals", "eigvals", "eigvecs", "loadings"):
        res["text_" + ext] = np.array((tmp / (out + "." + ext)).read_text())
    if "-m" in args.split() and "--svd 2" in args:
        # the block plan the reference logs and the SNP order of its permuted copy
        log = (tmp / (out + ".log")).read_text()
        plan = re.findall(r"after adjustment by PCAone: -w =\s*(\d+), blocksize =\s*(\d+), nblocks =\s*(\d+), "
                          r"factor =\s*(\d+)", log)
        res["ooc_plan"] = np.array([int(x) for x in plan[-1][1:]])
    if ooc_perm:
        ids = [ln.split()[1] for ln in (tmp / (prefix + ".bim")).read_text().splitlines()]
        pos = {v: i for i, v in enumerate(ids)}
        res["ooc_perm"] = np.array([pos[ln.split()[1]] for ln in
                                    (tmp / (out + ".perm.bim")).read_text().splitlines()], dtype=np.int64)
    if "--pcangsd" in args.split():
        # the EM trajectory the reference logs (6 decimals), src/Halko.cpp:309-310
        log = (tmp / (out + ".log")).read_text()
        res["em_diffs"] = np.array([float(x) for x in re.findall(
            r"individual allele frequencies estimated iter =\s*\d+, diff =\s*([0-9.eE+-]+)", log)])
    if iram_em:
        log = (tmp / (out + ".log")).read_text()
        res["em_rmse"] = np.array([float(x) for x in re.findall(r"RMSE =\s*([0-9.eE+-]+)", log)])

    return res


def main():
    if not REF.exists():
        raise SystemExit(f"{REF} missing: make -f oracle/ref.mk")
    only = set(sys.argv[1:])  # case names to (re)write; none: all of them

    def want(name):
        return not only or name in only

    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        payloads = {}
        for name, (seed, n, m, k, fst, miss) in DATASETS.items():
            pay = pa.sim_bed(seed, n, m, k, fst, 0.05, 0.5, miss)
            pa.write_plink(tmp / name, pay, n, m)
            payloads[name] = (pay, n, m)
        for case, (ds, args) in CASES.items():
            if not want(case):
                continue
            pay, n, m = payloads[ds]
            res = run_case(tmp, ds, args, case)
            np.savez_compressed(OUT / f"{case}.npz", payload=pay, nsamples=n, nsnps=m, args=args, **res)
            print(case, res["eigvals"])
        for name, (seed, n, m, npops, fst, depth, err) in GL.items():
            gl = pa.sim_gl(seed, n, m, npops, fst, 0.05, 0.5, depth, err)
            pa.write_beagle(tmp / (name + ".beagle.gz"), gl)
            for case, (ds, args) in GL_CASES.items():
                if ds != name or not want(case):
                    continue
                res = run_case(tmp, name + ".beagle.gz", args, case, flag="-G")
                res["cov"] = np.loadtxt(tmp / (case + ".cov"))
                # the GRM's JacobiSVD vectors, and the BEAGLE file their sample names come from
                res["text_eigvecs2"] = (tmp / (case + ".eigvecs2")).read_text()
                res["beagle_gz"] = np.frombuffer((tmp / (name + ".beagle.gz")).read_bytes(), dtype=np.uint8)
                np.savez_compressed(OUT / f"{case}.npz", gl=gl, nsamples=n, nsnps=m, args=args, **res)
                print(case, res["eigvals"])
        for name, (seed, n, per_chr, ld_bp) in LD.items():
            if not want(name):
                continue
            X, pos, chr_end = ld_inputs(seed, n, per_chr)
            ia, ib, r2 = run_ld(tmp, name, X, pos, chr_end, ld_bp)
            np.savez_compressed(OUT / f"{name}.npz", resid=X, pos=pos, chr_end=chr_end, ld_bp=ld_bp, snp_a=ia,
                                snp_b=ib, r2=r2)
            print(name, len(r2), "pairs")
        for name, (seed, n, per_chr, ld_bp, k) in LD_USV.items():
            if not want(name):
                continue
            res = run_ld_usv(tmp, name, seed, n, per_chr, ld_bp, k)
            np.savez_compressed(OUT / f"{name}.npz", **res)
            print(name, len(res["r2"]), "pairs")
        for name, (seed, nref, ntgt, m) in PROJECT.items():
            if not want(name):
                continue
            res = run_project(tmp, name, seed, nref, ntgt, m)
            np.savez_compressed(OUT / f"{name}.npz", **res)
            print(name, res["ntarget_snps"], "target SNPs", res["U_m2"][:2])
        for name, (seed, nref, ntgt, m) in PROJECT_GL.items():
            if not want(name):
                continue
            res = run_project_gl(tmp, name, seed, nref, ntgt, m)
            np.savez_compressed(OUT / f"{name}.npz", **res)
            print(name, len(res["sites"]), "target sites,", len(res["diffs"]), "EM iterations", res["U"][:1])
        for name, ds in INBREED.items():
            if not want(name):
                continue
            pay, n, m = payloads[ds]
            res = run_inbreed(tmp, name, ds)
            np.savez_compressed(OUT / f"{name}.npz", payload=pay, nsamples=n, nsnps=m, **res)
            print(name, res["hwe"].count("\n") - 1, "sites")
        for name, (seed, ncells, ngenes) in CSV.items():
            counts = csv_counts(seed, ncells, ngenes)
            text = "".join(",".join(str(int(v)) for v in row) + "\n" for row in counts)
            (tmp / (name + ".csv.zst")).write_bytes(zstd_compress(text.encode()))
            for case, (ds, args) in CSV_CASES.items():
                if ds != name or not want(case):
                    continue
                res = run_case(tmp, name + ".csv.zst", args, case, flag="--csv")
                np.savez_compressed(OUT / f"{case}.npz", counts=counts, nsamples=ncells, nsnps=ngenes, args=args,
                                    **res)
                print(case, res["eigvals"])
        for name, (seed, n, per_chr, ld_bp) in LD_TOOLS.items():
            if not want(name):
                continue
            res = run_ld_tools(tmp, name, seed, n, per_chr, ld_bp)
            np.savez_compressed(OUT / f"{name}.npz", **res)
            print(name, res["clump"].count("\n") - 1, "clumps,", res["prune_pm_out"].count("\n"), "pruned")
        for name, (seed, n, per_chr, ld_bp) in LD_STATS1.items():
            if not want(name):
                continue
            res = run_ld_stats1(tmp, name, seed, n, per_chr, ld_bp)
            np.savez_compressed(OUT / f"{name}.npz", **res)
            print(name, len(res["r2"]), "pairs", int(np.isnan(res["r2"]).sum()), "NaN")
        for name, (seed, n, m, npops, fst, miss) in BIG.items():
            pay = pa.sim_bed(seed, n, m, npops, fst, 0.05, 0.5, miss)
            pa.write_plink(tmp / name, pay, n, m)
            for case, (ds, args) in BIG_CASES.items():
                if ds != name or not want(case):
                    continue
                res = run_case(tmp, name, args, case)
                res.pop("text_loadings")  # the loadings are kept as float32 values
                res["loadings"] = res["loadings"].astype(np.float32)
                import hashlib
                np.savez_compressed(OUT / f"{case}.npz", sim=np.array([seed, n, m, npops]),
                                    sim_f=np.array([fst, 0.05, 0.5, miss]),
                                    payload_sha256=hashlib.sha256(pay.tobytes()).hexdigest(), nsamples=n, nsnps=m,
                                    args=args, **res)
                print(case, res["eigvals"])
        for name, (seed, n, m, npops, fst, miss, lay, comp, bits, pmin, ids) in BGEN.items():
            path = tmp / (name + ".bgen")
            pa.sim_bgen(path, seed, n, m, npops, fst, miss=miss, layout=lay, compression=comp, bits=bits,
                        pmin_declared=pmin, sample_ids=ids)
            raw = np.frombuffer(path.read_bytes(), dtype=np.uint8)
            for case, (ds, args) in BGEN_CASES.items():
                if ds != name or not want(case):
                    continue
                res = run_case(tmp, name + ".bgen", args, case, flag="-g")
                if "--ld" in args.split():
                    # the centred (unstandardised) matrix the reference writes as float32
                    res["residuals"] = np.fromfile(tmp / (case + ".residuals"), dtype=np.uint8)
                np.savez_compressed(OUT / f"{case}.npz", bgen=raw, nsamples=n, nsnps=m, args=args, **res)
                print(case, res["eigvals"])
        for name, (seed, n, m, base) in DENSE.items():
            D = pa.sim_counts(seed, n, m, base=base)
            pa.write_bin(tmp / (name + ".bin"), D)
            for case, (ds, args) in DENSE_CASES.items():
                if ds != name or not want(case):
                    continue
                res = run_case(tmp, name + ".bin", args, case, flag="-B")
                np.savez_compressed(OUT / f"{case}.npz", dense=D, nsamples=n, nsnps=m, args=args, **res)
                print(case, res["eigvals"])


if __name__ == "__main__":
    main()
