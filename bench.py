#!/usr/bin/env python3
"""Benchmark of the PCAone randomized-SVD hot path on one MI355X.

Headline workload (a "step" = one full winSVD solve, computeUSV from the first
computeGandH to U, S, V on the host): 10,000 samples x 500,000 SNPs, 5
Balding-Nichols populations (Fst 0.05, ancestral p ~ U(0.05, 0.5)), 0.5 %
missing, packed .bed resident in HBM, k = 20, l = 40, 64 windows, tol 1e-4,
at most 20 epochs.  Operator set-up (SNP shuffle, HBM layout, missing lists,
Omega draw) is the analogue of the reference's data loading; it is outside
the timed region and reported beside it (`init_s`).

Arms: split-BF16 MFMA (the measured value), plain FP32 (the drop-in it is
compared with) and FP64 (accuracy reference for both).

The other workloads of the same hot path are measured in the same run and
reported under "configs" (each with its own metric, roofline and CPU baseline):
  C5  LD r^2 straight from packed genotypes, 20,000 x 1,000,000 SNPs over 22
      chromosomes, 1 Mb windows (integer sums by int8 MFMA)
  C4  dense scRNA-like 200,000 cells x 20,000 genes FP32: RSVD k = 50 (l = 100)
      and IRAM top-50 (ncv = 101)
  C3  PCAngsd EM on genotype likelihoods, 5,000 x 200,000 sites, k = 4
  BGEN  the BGEN reader (8-bit zlib probabilities, 10,000 x 50,000) into the
      dense path
Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import gzip
import json
import os
import platform
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "PCA solve s + genotype GB/s: split-BF16 MFMA vs FP32 GPU at eig relerr≤1e-5"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
INT8_PEAK_TOPS = 5000.0
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch of k_packed_gemm_bf16 on the headline workload: rocprofv3
# FETCH_SIZE (x 2.0) + WRITE_SIZE (x 1.0) averaged over the 280 launches of one
# split-BF16 solve, the factors from a 512 MiB copy kernel
# (profiles/k_packed_gemm_bf16_c2_winsvd.md)
TRAFFIC = 291.3e6
# the same measure for the other workloads' dominant kernels (same calibration):
# k_ld_geno, the one full-call launch (profiles/k_ld_geno_c5_ld_r2.md)
TRAFFIC_C5 = 26.03e9
# k_dense_gemm_bf16x3, averaged over the 280 launches of a solve
# (profiles/k_dense_gemm_bf16x3_c4_dense_rsvd.md)
TRAFFIC_C4 = 3.130e9
# k_gl_E, one launch per EM iteration (profiles/k_gl_E_c3_pcangsd.md)
TRAFFIC_C3 = 12.49e9
REF = ROOT / "oracle" / "_ref" / "PCAone"
# worker threads of the CPU baselines: the GPU host's CPU share for one GPU
CPU_THREADS = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def principal_angle(A, B):
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1)))


def cpu_facts():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"threads_used": CPU_THREADS, "os_cpu_count": os.cpu_count(), "affinity": aff, "model": model}


def run_ref(args, cwd):
    """The reference CLI (built from its sources by oracle/ref.mk); returns
    (wall s, solve s = wall - logged reading time, log text) or None."""
    if not REF.exists():
        return None
    out = None
    for i, a in enumerate(args):
        if a == "-o":
            out = args[i + 1]
    t0 = time.perf_counter()
    r = subprocess.run([str(REF), *args], cwd=cwd, capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        log("reference CLI failed:", " ".join(args), r.stderr[-400:])
        return None
    text = (Path(cwd) / (out + ".log")).read_text() if out and (Path(cwd) / (out + ".log")).exists() else ""
    rd = re.search(r"total elapsed reading time:\s*([0-9.eE+-]+)", text)
    tot = re.search(r"total elapsed wall time:\s*([0-9.eE+-]+)", text)
    total_s = float(tot.group(1)) if tot else wall
    read_s = float(rd.group(1)) if rd else 0.0
    return wall, max(total_s - read_s, 1e-9), text


# ---------------------------------------------------------------------------
# C2: winSVD on packed genotypes (headline)
# ---------------------------------------------------------------------------
def c2_solve(g, pa, wl, prec):
    t0 = time.perf_counter()
    g.rsvd_init(wl["k"], wl["os"], variant=pa.RSVD_FANCY, bands=wl["bands"], precision=prec, seed=101)
    g.synchronize()
    t1 = time.perf_counter()
    U, S, V, ep, diffs = g.usv(wl["maxp"], wl["tol"])
    g.synchronize()
    return time.perf_counter() - t1, t1 - t0, U, S, V, ep


def c2_streamed(g, pa, wl, pay, n, m, prec, memory_gb=2.0):
    """Out of core (-m): the payload as a .bed file in host memory, every pass
    streamed window by window through the device ring, with the reference's
    -m block plan and permute_plink's SNP order (PCIe-bound by design)."""
    with tempfile.TemporaryDirectory() as td:
        pa.write_plink(Path(td) / "c2", pay, n, m)
        t0 = time.perf_counter()
        g.load_bed_file(Path(td) / "c2.bed", n, m, out_of_core=True)
        g.synchronize()
        load_s = time.perf_counter() - t0
        try:
            bs, nb, factor = pa.ooc_plan(n, m, 2, wl["k"], wl["os"], wl["bands"], memory_gb)
        except pa.GpuError:  # small shapes fit one block: stream the in-core windows
            memory_gb, bs, nb, factor = None, -(-m // wl["bands"]), wl["bands"], 1
        perm = pa.plink_perm(n, m, 2, wl["bands"])
        t0 = time.perf_counter()
        g.rsvd_init(wl["k"], wl["os"], variant=pa.RSVD_FANCY, bands=wl["bands"], precision=prec, seed=101,
                    permutation=perm, blocksize=bs, band_factor=factor)
        g.synchronize()
        t1 = time.perf_counter()
        U, S, V, ep, _ = g.usv(wl["maxp"], wl["tol"])
        g.synchronize()
        solve_s = time.perf_counter() - t1
    log(f"C2 streamed (-m {memory_gb}: {nb} blocks of {bs}, factor {factor}): load {load_s:.2f} s, "
        f"init {t1 - t0:.2f} s, solve {solve_s:.2f} s, {ep} epochs")
    return dict(plan={"memory_gb": memory_gb, "blocksize": bs, "nblocks": nb, "band_factor": factor},
                load_s=load_s, init_s=t1 - t0, solve_s=solve_s, S=S, U=U, epochs=ep)


def c2_cpu_baseline(g, pa, wl, msample, cores):
    """Reference CLI on the first `msample` SNPs of the same simulation, same
    winSVD settings; then the GPU's split-BF16 solve of that same sample
    against the reference's printed eigenvalues and PCs."""
    n = wl["N"]
    pay = pa.sim_bed(wl["seed"], n, msample, wl["npops"], wl["fst"], 0.05, 0.5, wl["miss"])
    log(f"C2: reference CLI on {msample} SNPs x {n} samples")
    with tempfile.TemporaryDirectory() as td:
        pa.write_plink(Path(td) / "sample", pay, n, msample)
        res = run_ref(["-b", "sample", "--svd", "2", "-k", str(wl["k"]), "--oversamples", str(wl["os"]), "-w",
                       str(wl["bands"]), "-p", str(wl["maxp"]), "--tol-rsvd", str(wl["tol"]), "-n", str(cores), "-o",
                       "out", "-v", "1"], td)
        if res is None:
            return None, None
        wall, solve_s, text = res
        ev_ref = np.loadtxt(Path(td) / "out.eigvals", ndmin=1)
        U_ref = np.loadtxt(Path(td) / "out.eigvecs", ndmin=2)
    ep = re.search(r"stops at epoch =\s*(\d+)", text)
    epochs = int(ep.group(1)) if ep else 7
    gbytes = epochs * msample * ((n + 3) >> 2) / 1e9
    cpu = {"value": round(gbytes / solve_s, 4), "unit": "GB/s", "cores": cores, "kind": "reference",
           "sample": f"reference PCAone CLI (-n {cores}) winSVD -k {wl['k']} --oversamples {wl['os']} "
                     f"-w {wl['bands']} on the first {msample} SNPs x {n} samples of the same simulation; "
                     f"solve {solve_s:.2f} s = wall {wall:.2f} s - reading; {epochs} epochs",
           "cpu": cpu_facts()}
    # the GPU on the same sample
    g.load_bed(pay, n, msample)
    g.rsvd_init(wl["k"], wl["os"], variant=pa.RSVD_FANCY, bands=wl["bands"], precision=pa.PREC_SPLIT_BF16, seed=101)
    U, S, V, ep_gpu, _ = g.usv(wl["maxp"], wl["tol"])
    ev = S ** 2 / msample * 2
    npc = wl["npops"] - 1
    acc = {"sample": f"first {msample} SNPs", "epochs_gpu": ep_gpu, "epochs_ref": epochs,
           "eig_relerr_max_vs_printed": float(np.max(np.abs(ev - ev_ref) / ev_ref)),
           f"max_angle_top{npc}_rad": principal_angle(U[:, :npc], U_ref[:, :npc]),
           "note": "the reference prints 6 significant digits (5e-6 relative)"}
    return cpu, acc


def run_c2(g, pa, args):
    wl = dict(N=10000, M=500000, npops=5, fst=0.05, miss=0.005, k=20, os=20, bands=64, maxp=20, tol=1e-4,
              seed=2026)
    if args.small:
        wl.update(N=2000, M=20000, npops=3, fst=0.1, miss=0.01, k=10, os=10)
    n, m, l = wl["N"], wl["M"], wl["k"] + wl["os"]
    t0 = time.perf_counter()
    pay = pa.sim_bed(wl["seed"], n, m, wl["npops"], wl["fst"], 0.05, 0.5, wl["miss"])
    log(f"C2: generated {m} x {n} genotypes in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    g.load_bed(pay, n, m)  # host -> HBM over PCIe, plus the allele-frequency count
    g.synchronize()
    load_s = time.perf_counter() - t0
    bytes_per_epoch = m * ((n + 3) >> 2)

    arms = {}
    want = set(args.arms.split(","))
    for name, prec, steps, warm in (("bf16x2", pa.PREC_SPLIT_BF16, args.steps, args.warmup),
                                    ("fp32", pa.PREC_FP32, max(1, args.steps), max(1, args.warmup)),
                                    ("fp64", pa.PREC_FP64, 1, 0)):
        if name not in want:
            continue
        for _ in range(warm):
            c2_solve(g, pa, wl, prec)
        ts, inits, res = [], [], None
        for _ in range(steps):
            t, ti, U, S, V, ep = c2_solve(g, pa, wl, prec)
            ts.append(t)
            inits.append(ti)
            res = (U, S, V, ep)
        arms[name] = dict(times=ts, init=inits, U=res[0], S=res[1], V=res[2], epochs=res[3])
        log(f"C2 {name}: {np.mean(ts) * 1e3:.1f} ms/solve over {steps} solves (init {np.mean(inits) * 1e3:.1f} ms), "
            f"{res[3]} epochs")

    roofline = None
    if "bf16x2" in want:
        # roofline of the dominant kernel, timed with HIP events on the handle's stream
        # (one extra split-BF16 solve with an event pair around every launch)
        g.profiling(True)
        g.reset_stats()
        _, _, _, _, _, ep_prof = c2_solve(g, pa, wl, pa.PREC_SPLIT_BF16)
        stats = g.kernel_stats()
        g.profiling(False)
        cnt = stats["gemm_G"][0] + stats["gemm_H"][0]
        ms = stats["gemm_G"][1] + stats["gemm_H"][1]
        flops = ep_prof * 4.0 * m * n * l  # algorithmic: two 2*M*N*l products per epoch
        achieved = flops / (ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": TRAFFIC,
                    "kernel": "k_packed_gemm_bf16 (gemm_G + gemm_H)",
                    "launches": cnt, "avg_launch_us": round(ms * 1e3 / cnt, 2),
                    "algorithmic_flops_per_launch": flops / cnt,
                    "kernel_ms_per_solve": {k: round(v[1], 3) for k, v in sorted(stats.items())},
                    "kernel_launches_per_solve": {k: v[0] for k, v in sorted(stats.items())}}
        if "fp32" in want:
            # the FP32 arm's contraction (k_packed_gemm_valu<float>: VALU FMA) against the FP32 vector peak
            g.profiling(True)
            g.reset_stats()
            _, _, _, _, _, ep32 = c2_solve(g, pa, wl, pa.PREC_FP32)
            st32 = g.kernel_stats()
            g.profiling(False)
            ms32 = st32["gemm_G"][1] + st32["gemm_H"][1]
            a32 = ep32 * 4.0 * m * n * l / (ms32 * 1e-3) / 1e12
            roofline["fp32_arm"] = {"kernel": "k_packed_gemm_valu<float> (gemm_G + gemm_H)", "bound": "valu",
                                    "achieved": round(a32, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                    "frac": round(a32 / FP32_PEAK_TFLOPS, 4), "contraction_ms_per_solve": round(ms32, 3),
                                    "launches": st32["gemm_G"][0] + st32["gemm_H"][0],
                                    "note": "peak: MI355X's FP32 figure; without packed FP32 (left off in this "
                                            "build) plain FMAs issue at half of it"}
        if want != {"bf16x2", "fp32", "fp64"}:
            log(f"C2 timing solve: contraction {cnt} launches, avg {roofline['avg_launch_us']} us, "
                f"{roofline['achieved']} TFLOP/s")
            log("C2 kernel ms per solve: " + json.dumps({k: [v[0], round(v[1], 3)] for k, v in
                                                         sorted(stats.items(), key=lambda x: -x[1][1])}))
    if want != {"bf16x2", "fp32", "fp64"}:
        return None
    ref = arms["fp64"]
    acc = {}
    npc = wl["npops"] - 1
    for name in ("bf16x2", "fp32"):
        a = arms[name]
        ev, ev0 = a["S"] ** 2 / m * 2, ref["S"] ** 2 / m * 2
        acc[name] = {"eig_relerr_max": float(np.max(np.abs(ev - ev0) / ev0)),
                     f"max_angle_top{npc}_rad": principal_angle(a["U"][:, :npc], ref["U"][:, :npc]),
                     "epochs": a["epochs"]}
    bf = arms["bf16x2"]
    t_bf = float(np.mean(bf["times"]))
    init_s = float(np.mean(bf["init"]))
    gbs = bf["epochs"] * bytes_per_epoch / t_bf / 1e9
    t_fp32 = float(np.mean(arms["fp32"]["times"]))
    st = c2_streamed(g, pa, wl, pay, n, m, pa.PREC_SPLIT_BF16)
    ev, ev0 = st["S"] ** 2 / m * 2, ref["S"] ** 2 / m * 2
    out_of_core = {
        "what": "the same payload as a .bed file in host memory, streamed per window through a two-slot device "
                "ring on every pass (the reference's -m block plan and permute_plink order); split-BF16",
        "plan": st["plan"], "load_s": round(st["load_s"], 4), "init_s": round(st["init_s"], 4),
        "solve_s": round(st["solve_s"], 4), "epochs": st["epochs"],
        "value_GBps": round(st["epochs"] * bytes_per_epoch / st["solve_s"] / 1e9, 3),
        # the top PCs; the trailing ones are the noise spectrum, where another SNP
        # order (permute_plink's, not the shuffle) settles differently within tol
        f"eig_relerr_top{npc}_vs_resident_fp64": float(np.max(np.abs(ev - ev0)[:npc] / ev0[:npc])),
        f"eig_relerr_all{wl['k']}_vs_resident_fp64": float(np.max(np.abs(ev - ev0) / ev0)),
        f"max_angle_top{npc}_vs_resident_fp64_rad": principal_angle(st["U"][:, :npc], ref["U"][:, :npc])}
    cpu, acc_ref = (None, None) if args.no_cpu else c2_cpu_baseline(g, pa, wl, min(args.cpu_snps, m), CPU_THREADS)
    if acc_ref:
        acc["bf16x2_vs_reference_cli_on_cpu_sample"] = acc_ref
    return {
        "metric": METRIC, "value": round(gbs, 3), "unit": "GB/s", "n_gpus": 1, "steps": len(bf["times"]),
        "warmup": args.warmup, "ms_per_step": round(t_bf * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"winSVD {n} samples x {m} SNPs packed .bed in HBM, {wl['npops']} Balding-Nichols "
                               f"pops Fst {wl['fst']}, {wl['miss'] * 100:g}% missing, k={wl['k']} l={l}, "
                               f"{wl['bands']} windows, tol {wl['tol']}",
                   "precision": "genotypes exact in BF16 x operand split into BF16 hi+lo, FP32 MFMA accumulation;"
                                " FP64 CholeskyQR2, Householder signs, l x l eigensolve",
                   "epochs": bf["epochs"], "solve_s": round(t_bf, 5)},
        "init_s": round(init_s, 5),
        "value_incl_init_GBps": round(bf["epochs"] * bytes_per_epoch / (t_bf + init_s) / 1e9, 3),
        "fp32_arm": {"solve_s": round(t_fp32, 5), "value_GBps": round(arms["fp32"]["epochs"] * bytes_per_epoch /
                                                                      t_fp32 / 1e9, 3)},
        "fp64_arm": {"solve_s": round(float(np.mean(ref["times"])), 5)},
        "speedup_vs_fp32": round(t_fp32 / t_bf, 3),
        "pcie_inclusive": {"load_bed_s": round(load_s, 4),
                           "value_GBps": round(bf["epochs"] * bytes_per_epoch / (t_bf + load_s) / 1e9, 3)},
        "accuracy_vs_fp64": acc,
        "out_of_core": out_of_core,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
# C5: LD r^2 straight from the packed genotypes
# ---------------------------------------------------------------------------
def c5_positions(M, nchr, seed):
    """Positions as a Poisson process whose mean spacing changes every 2,000
    SNPs (uniform in 333 .. 5,667 bp, mean 3 kb), so 1 Mb windows hold about
    200 to 3,000 SNPs; nchr chromosomes of equal SNP counts."""
    rng = np.random.default_rng(seed)
    nreg = M // 2000 + 1
    rate = rng.uniform(333, 5667, nreg)
    gaps = rng.exponential(np.repeat(rate, 2000)[:M]).astype(np.int64) + 1
    per = np.full(nchr, M // nchr)
    per[-1] += M - per.sum()
    pos = np.empty(M, np.int64)
    s = 0
    for c in range(nchr):
        pos[s:s + per[c]] = np.cumsum(gaps[s:s + per[c]])
        s += per[c]
    return pos.astype(np.int32), (np.cumsum(per) - 1).astype(np.int32)


def c5_cpu_baseline(pa, pay_rows, n, pos, chr_end, nsnp, cores):
    """The reference's own two steps on the first `nsnp` SNPs of chromosome 1:
    PCA with --ld --ld-stats 1 (writes the .residuals), then --print-r2 on them
    (ld_r2_big: single-threaded, text and gzip per pair)."""
    pay = np.ascontiguousarray(pay_rows[:nsnp]).reshape(-1)
    log(f"C5: reference CLI on {nsnp} SNPs x {n} samples")
    with tempfile.TemporaryDirectory() as td:
        pa.write_plink(Path(td) / "s", pay, n, nsnp, np.ones(nsnp, dtype=np.int64), pos[:nsnp].astype(np.int64))
        a = run_ref(["-b", "s", "--svd", "1", "-k", "2", "-p", "3", "--ld", "--ld-stats", "1", "-o", "res", "-n",
                     str(cores), "-v", "0"], td)
        b = run_ref(["-B", "res.residuals", "--match-bim", "res.mbim", "--print-r2", "--ld-bp", "1000000", "-o", "ld",
                     "-n", str(cores), "-v", "1"], td)
        if a is None or b is None:
            return None
        with gzip.open(Path(td) / "ld.ld.gz", "rt") as f:
            npairs = sum(1 for _ in f) - 1
    wall, solve_s, _ = b
    return {"value": round(npairs / solve_s, 1), "unit": "pairs/s", "cores": 1, "kind": "reference",
            "sample": f"reference CLI --print-r2 on its own --ld --ld-stats 1 .residuals of the first {nsnp} SNPs of "
                      f"chromosome 1 x {n} samples ({npairs} pairs, ld_r2_big single-threaded: {solve_s:.2f} s = wall "
                      f"- reading); per-pair rate, which scales linearly to the whole genome",
            "cpu": cpu_facts()}


def run_c5(g, pa, args):
    N, M = (20000, 1000000) if not args.small else (2000, 50000)
    t0 = time.perf_counter()
    pay = pa.sim_bed(2027, N, M, 5, 0.05, 0.05, 0.5, 0.005)
    pos, chr_end = c5_positions(M, 22, 7)
    ws, we = pa.ld_windows(pos, chr_end, 1000000)
    P = int((we - 1).sum())
    log(f"C5: generated {M} x {N} genotypes, {len(ws)} windows, {P} pairs in {time.perf_counter() - t0:.1f} s")
    g.load_bed(pay, N, M)
    g.synchronize()
    r2, _, cfg = g.ld_r2_geno(ws, we)  # builds the layout, autotunes the tile shape
    log(f"C5: tuned tile configuration {cfg}")
    g.profiling(True)
    walls, kms = [], []
    for _ in range(max(1, args.steps)):
        g.reset_stats()
        t0 = time.perf_counter()
        r2, _, _ = g.ld_r2_geno(ws, we, config=cfg)
        walls.append(time.perf_counter() - t0)
        kms.append(g.kernel_stats()["ld_geno"][1])
    g.profiling(False)
    ms = float(np.mean(kms))
    wall = float(np.mean(walls))
    ops = 4 * 2.0 * N * P  # the four integer Grams per pair
    achieved = ops / (ms * 1e-3) / 1e12
    ok = ~np.isnan(r2)
    out = {"metric": "LD r^2 pairs/s (int8 MFMA integer sums, FP64 r^2), r^2 resident in HBM",
           "value": round(P / (ms * 1e-3), 1), "unit": "pairs/s", "higher_is_better": True, "dtype": "int8",
           "config": {"workload": f"LD r^2, {N} samples x {M} SNPs packed .bed, 22 chromosomes, positions a "
                                  "Poisson process with mean spacing changing every 2,000 SNPs (333..5,667 bp), "
                                  "--ld-bp 1000000, 0.5% missing, --ld-stats 1 semantics",
                      "pairs": P, "windows": len(ws), "window_snps_p1_p50_p99": [int(x) for x in
                                                                             np.percentile(we, [1, 50, 99])],
                      "tile_config": cfg},
           "ms_per_step": round(ms, 3), "steps": len(kms),
           "pcie_inclusive": {"call_s": round(wall, 4), "pairs_per_s": round(P / wall, 1),
                              "note": "the whole C-ABI call, r^2 copied to host memory (8 B per pair)"},
           "int_ops_per_s": achieved * 1e12, "packed_GBps_read_once": round(M * ((N + 3) >> 2) / (ms * 1e-3) / 1e9, 2),
           "r2_range": [float(r2[ok].min()), float(r2[ok].max())],
           "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                        "frac": round(achieved / INT8_PEAK_TOPS, 4),
                        "traffic": None if args.small else TRAFFIC_C5,
                        "kernel": "k_ld_geno (one persistent launch)",
                        "algorithmic_ops_per_launch": ops, "avg_launch_us": round(ms * 1e3, 1)}}
    if not args.no_cpu:
        out["cpu_baseline"] = c5_cpu_baseline(pa, pay.reshape(M, -1), N, pos, chr_end, 1000, CPU_THREADS)
    return out


# ---------------------------------------------------------------------------
# C4: dense scRNA-like matrix, RSVD k = 50 and IRAM top-50
# ---------------------------------------------------------------------------
def run_c4(g, pa, args):
    ncells, ngenes = (200000, 20000) if not args.small else (20000, 2048)
    t0 = time.perf_counter()
    D = pa.sim_counts(44, ncells, ngenes, base=0.03, sdlog=1.5, phi=0.2, nthreads=CPU_THREADS)
    zeros = float((D[:200] == 0).mean())
    log(f"C4: generated {ngenes} x {ncells} counts ({zeros * 100:.1f}% zeros) in {time.perf_counter() - t0:.1f} s")
    g.load_dense(D)
    g.synchronize()
    m, n = ngenes, ncells
    res = {}
    for name, prec in (("bf16x2", pa.PREC_SPLIT_BF16), ("fp32", pa.PREC_FP32), ("fp64", pa.PREC_FP64)):
        steps = 1 if name == "fp64" else max(1, args.steps)
        g.rsvd_init(50, 50, variant=pa.RSVD_FANCY, bands=64, precision=prec)
        g.usv(20, 1e-4)  # warm-up
        ts = []
        for _ in range(steps):
            g.rsvd_init(50, 50, variant=pa.RSVD_FANCY, bands=64, precision=prec)
            g.synchronize()
            t0 = time.perf_counter()
            U, S, V, ep, _ = g.usv(20, 1e-4)
            ts.append(time.perf_counter() - t0)
        res[name] = (float(np.mean(ts)), U, S, ep)
        log(f"C4 RSVD {name}: {res[name][0] * 1e3:.1f} ms, {ep} epochs")
    # roofline of the split arm's dominant kernel (the dense bf16x3 contraction: HBM-bound)
    g.rsvd_init(50, 50, variant=pa.RSVD_FANCY, bands=64, precision=pa.PREC_SPLIT_BF16)
    g.profiling(True)
    g.reset_stats()
    _, _, _, ep, _ = g.usv(20, 1e-4)
    st = g.kernel_stats()
    g.profiling(False)
    cms = st["gemm_G"][1] + st["gemm_H"][1]
    cnt = st["gemm_G"][0] + st["gemm_H"][0]
    gbytes = ep * 2 * 4.0 * m * n / 1e9  # each epoch reads the FP32 matrix once per product
    t_bf, U, S, ep_bf = res["bf16x2"]
    S64, U64 = res["fp64"][2], res["fp64"][1]
    ev, ev0 = S ** 2, S64 ** 2
    out = {"metric": "dense PCA solve s (RSVD k=50 split-BF16 3-product MFMA; IRAM top-50)", "value": round(t_bf, 4),
           "unit": "s", "higher_is_better": False, "dtype": "bf16x3",
           "config": {"workload": f"{ncells} cells x {ngenes} genes log1p(CPM) FP32 in HBM, centred per gene, "
                                  f"NB(mean LogNormal(0,1.5) x 0.03, dispersion 0.2) counts, {zeros * 100:.1f}% zeros;"
                                  " winSVD k=50 l=100 64 windows tol 1e-4", "epochs": ep_bf},
           "fp32_arm_s": round(res["fp32"][0], 4), "fp64_arm_s": round(res["fp64"][0], 4),
           "speedup_vs_fp32": round(res["fp32"][0] / t_bf, 3),
           "accuracy_vs_fp64": {"eig_relerr_max_top10": float(np.max(np.abs(ev[:10] - ev0[:10]) / ev0[:10])),
                                "eig_relerr_max_top50": float(np.max(np.abs(ev - ev0) / ev0)),
                                "max_angle_top5_rad": principal_angle(U[:, :5], U64[:, :5])},
           "roofline": {"bound": "hbm", "achieved": round(gbytes / (cms * 1e-3), 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbytes / (cms * 1e-3) / HBM_PEAK_GBS, 4),
                        "traffic": None if args.small else TRAFFIC_C4,
                        "kernel": "k_dense_gemm_bf16x3 (gemm_G + gemm_H)", "launches": cnt,
                        "avg_launch_us": round(cms * 1e3 / cnt, 2),
                        "kernel_ms_per_solve": {k: round(v[1], 3) for k, v in sorted(st.items())}}}
    if not args.no_cpu:
        out["cpu_baseline"] = c4_cpu_baseline(pa, D, CPU_THREADS)
    if args.no_iram:
        return out
    # IRAM top-50 (Spectra op on the GPU, restarted Lanczos on the host): the
    # FP64 arm converges to itol 1e-6 and is the ground truth for every arm
    iram = {}
    for name, prec in (("bf16x2", pa.PREC_SPLIT_BF16), ("fp64", pa.PREC_FP64)):
        g.gram_init(prec)
        log(f"C4: IRAM top-50 {name}")
        t0 = time.perf_counter()
        _, Si, _, nconv, nops = g.iram(50, ncv=101, maxit=1000, tol=1e-6)
        iram[name] = (time.perf_counter() - t0, Si, nops)
        log(f"C4 IRAM {name}: {iram[name][0]:.2f} s ({nops} ops)")
    evi = iram["fp64"][1] ** 2

    def rel(a, b):
        return float(np.max(np.abs(a - b) / b))

    out["iram"] = {"bf16x2_s": round(iram["bf16x2"][0], 3), "bf16x2_nops": iram["bf16x2"][2],
                   "fp64_s": round(iram["fp64"][0], 3), "fp64_nops": iram["fp64"][2],
                   "perform_op_GBps_bf16x2": round(iram["bf16x2"][2] * 2 * 4.0 * m * n / iram["bf16x2"][0] / 1e9, 1),
                   "eig_relerr_top50_bf16x2_vs_fp64": rel(iram["bf16x2"][1] ** 2, evi)}
    out["accuracy_vs_iram_fp64"] = {
        "note": "converged top-50 of FP64 IRAM as the truth: RSVD k=50 at tol 1e-4 stops short of convergence on "
                "this flat spectrum in every arm, so the split and FP64 RSVD arms differ by their unconverged "
                "trajectories (Householder signs and flipOmg are discontinuous), not by their arithmetic",
        **{f"rsvd_{a}_top{t}": rel(res[a][2][:t] ** 2, evi[:t]) for a in ("bf16x2", "fp32", "fp64") for t in (5, 50)}}
    return out


def c4_cpu_baseline(pa, D, cores):
    """Reference CLI RSVD k=50 (winSVD, -w 64) on the first 10,000 cells of the
    same matrix (all genes), as .bin input: elements x epochs per second."""
    ncell = min(10000, D.shape[1])
    sub = np.ascontiguousarray(D[:, :ncell])
    log(f"C4: reference CLI on {sub.shape[0]} genes x {ncell} cells")
    with tempfile.TemporaryDirectory() as td:
        pa.write_bin(Path(td) / "d.bin", sub)
        r = run_ref(["-B", "d.bin", "--svd", "2", "-k", "50", "-w", "64", "-p", "20", "--tol-rsvd", "1e-4", "-n",
                     str(cores), "-o", "out", "-v", "1"], td)
    if r is None:
        return None
    wall, solve_s, text = r
    ep = re.search(r"stops at epoch =\s*(\d+)", text)
    epochs = int(ep.group(1)) if ep else 7
    gbytes = epochs * 2 * 4.0 * sub.size / 1e9
    return {"value": round(gbytes / solve_s, 3), "unit": "GB/s (FP32 matrix bytes per product)", "cores": cores,
            "kind": "reference",
            "sample": f"reference CLI (-n {cores}) winSVD -k 50 -w 64 on {sub.shape[0]} genes x {ncell} cells of the "
                      f"same matrix: solve {solve_s:.2f} s = wall {wall:.2f} s - reading, {epochs} epochs",
            "cpu": cpu_facts()}


# ---------------------------------------------------------------------------
# C3: PCAngsd EM on genotype likelihoods
# ---------------------------------------------------------------------------
def run_c3(g, pa, args):
    n, m = (5000, 200000) if not args.small else (1000, 20000)
    t0 = time.perf_counter()
    gl = pa.sim_gl(33, n, m, 4, 0.1, 0.05, 0.5, 2.0, 0.01, nthreads=CPU_THREADS)
    log(f"C3: generated {m} x {n} likelihoods in {time.perf_counter() - t0:.1f} s")
    g.load_gl(gl)
    out = {}
    res = {}
    for name, prec in (("bf16x2", pa.PREC_SPLIT_BF16), ("fp32", pa.PREC_FP32), ("fp64", pa.PREC_FP64)):
        g.profiling(name == "bf16x2")
        g.reset_stats()
        t0 = time.perf_counter()
        F, it_maf = g.gl_em_maf(100, 1e-6)
        g.rsvd_init(4, 10, variant=pa.RSVD_NORMAL, precision=prec)
        U, S, V, it, diffs, Dc, Cm = g.pcangsd(20, 1e-4, 100, 1e-5, grm=False, dc=True)
        g.synchronize()
        res[name] = (time.perf_counter() - t0, S, U, it)
        if name == "bf16x2":
            st = g.kernel_stats()
        log(f"C3 {name}: {res[name][0]:.2f} s, {it} EM iterations")
    g.profiling(False)
    t_bf, S, U, it = res["bf16x2"]
    S64, U64 = res["fp64"][1], res["fp64"][2]
    ev, ev0 = S ** 2, S64 ** 2
    em = st.get("gl_E", (0, 0.0))
    # the E update reads the two FP32 likelihoods and writes E (FP32, laid-out rows): 12 B per element
    ebytes = em[0] * 12.0 * n * m / 1e9 if em[0] else 0.0
    out = {"metric": "PCAngsd EM solve s (emMAF + EM iterations, split-BF16 3-product MFMA RSVD per iteration)",
           "value": round(t_bf, 3), "unit": "s", "higher_is_better": False, "dtype": "bf16x3",
           "config": {"workload": f"{n} samples x {m} sites, 4 pops Fst 0.1, depth Poisson(2), 1% error, FP32 GLs; "
                                  "--svd 1 k=4 l=14, tol-em 1e-5, tol-maf 1e-6", "em_iterations": it},
           "fp32_arm_s": round(res["fp32"][0], 3), "fp64_arm_s": round(res["fp64"][0], 3),
           "accuracy_vs_fp64": {"eig_relerr_max": float(np.max(np.abs(ev - ev0) / ev0)),
                                "max_angle_top3_rad": principal_angle(U[:, :3], U64[:, :3]),
                                "em_iterations_fp64": res["fp64"][3]},
           "kernel_ms": {k: round(v[1], 2) for k, v in sorted(st.items())}}
    if em[0]:
        out["roofline"] = {"bound": "hbm", "achieved": round(ebytes / (em[1] * 1e-3), 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(ebytes / (em[1] * 1e-3) / HBM_PEAK_GBS, 4),
                           "traffic": None if args.small else TRAFFIC_C3,
                           "kernel": "k_gl_E (E update: two FP32 likelihoods in, FP32 E out)",
                           "launches": em[0], "avg_launch_us": round(em[1] * 1e3 / em[0], 1)}
    # --project of the same likelihoods onto the FP64 arm's PCA (the last arm
    # run: its V; every site matched, F the emMAF frequencies): the EM of
    # run_projection_beagle_gl
    g.profiling(True)
    g.reset_stats()
    t0 = time.perf_counter()
    Up, _, pit, _ = g.project_gl_matched(np.arange(m), np.zeros(m, np.uint8), F, S64, V, maxiter=100, tolem=1e-5)
    tp = time.perf_counter() - t0
    pst = g.kernel_stats()
    g.profiling(False)
    es = pst["projgl_estep"]
    pbytes = es[0] * 8.0 * n * m  # two FP32 likelihoods per sample and site per E step
    out["projection"] = {
        "what": "--project of the same likelihoods onto this PCA (run_projection_beagle_gl: device E step, "
                "per-sample least squares, host sign alignment)", "value_s": round(tp, 4), "iterations": pit,
        # the projection's least squares and PCAngsd's standardised E are different estimators
        "angle_top3_vs_pcangsd_U_rad": principal_angle(Up[:, :3], U64[:, :3]),
        "roofline": {"bound": "hbm", "achieved": round(pbytes / (es[1] * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(pbytes / (es[1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "kernel": "k_projgl_estep", "launches": es[0], "avg_launch_us": round(es[1] * 1e3 / es[0], 1)},
        "kernel_ms": {k: round(v[1], 2) for k, v in sorted(pst.items())}}
    log(f"C3 projection: {tp:.3f} s, {pit} iterations")
    if not args.no_cpu:
        out["cpu_baseline"] = c3_cpu_baseline(pa, gl, CPU_THREADS)
    return out


def c3_cpu_baseline(pa, gl, cores):
    """Reference CLI PCAngsd (-G BEAGLE, --svd 1 -k 4) on the first 2,000 sites
    x 500 samples of the same likelihoods: site x sample x EM iterations per
    second (its time includes the GRM and the N x N JacobiSVD the reference
    runs for BEAGLE input, which grow as N^3: hence the few samples)."""
    ms, ns = min(2000, gl.shape[0]), min(500, gl.shape[1])
    sub = np.ascontiguousarray(gl[:ms, :ns])
    log(f"C3: reference CLI on {ms} sites x {ns} samples")
    with tempfile.TemporaryDirectory() as td:
        pa.write_beagle(Path(td) / "g.beagle.gz", sub)
        r = run_ref(["-G", "g.beagle.gz", "--svd", "1", "-k", "4", "-p", "20", "--tol-rsvd", "1e-4", "-n", str(cores),
                     "-o", "out", "-v", "1"], td)
    if r is None:
        return None
    wall, solve_s, text = r
    its = len(re.findall(r"individual allele frequencies estimated", text)) or 1
    rate = ms * ns * its / solve_s
    return {"value": round(rate, 1), "unit": "site-samples x EM iterations / s", "cores": cores, "kind": "reference",
            "sample": f"reference CLI (-n {cores}) -G --svd 1 -k 4 on the first {ms} sites x {ns} samples: "
                      f"solve {solve_s:.2f} s = wall {wall:.2f} s - reading, {its} EM iterations (GRM and its "
                      "JacobiSVD included)",
            "cpu": cpu_facts()}


# ---------------------------------------------------------------------------
# BGEN input: the reader feeding the dense path
# ---------------------------------------------------------------------------
def run_bgen(g, pa, args):
    """pcaone_gpu_load_bgen on a BGEN v1.2 file (layout 2, 8-bit, zlib) in host
    memory: container walk, decompression on the host's 16 worker threads,
    the chunk copied to HBM and decoded there (dosage, minor allele, F,
    centring, standardisation, FP32 matrix)."""
    n, m = (10000, 50000) if not args.small else (2000, 5000)
    with tempfile.TemporaryDirectory() as td:
        path = Path(td) / "x.bgen"
        t0 = time.perf_counter()
        pa.sim_bgen(path, 2031, n, m, 5, 0.05, miss=0.01, layout=2, compression=1, bits=8, nthreads=CPU_THREADS)
        fsize = path.stat().st_size
        log(f"BGEN: wrote {m} variants x {n} samples ({fsize / 1e6:.0f} MB) in {time.perf_counter() - t0:.1f} s")
        g.load_bgen(path, nthreads=CPU_THREADS)  # warm-up (page cache, pinned staging)
        walls, kms = [], []
        g.profiling(True)
        for _ in range(max(1, args.steps)):
            g.reset_stats()
            t0 = time.perf_counter()
            g.load_bgen(path, nthreads=CPU_THREADS)
            g.synchronize()
            walls.append(time.perf_counter() - t0)
            kms.append(g.kernel_stats()["bgen_dosage"])
        g.profiling(False)
        cpu = None if args.no_cpu else bgen_cpu_baseline(pa, n, min(2000, m), CPU_THREADS)
    wall = float(np.mean(walls))
    launches = kms[0][0]
    kms_ms = float(np.mean([k[1] for k in kms]))
    dbytes = m * (2.0 * n + n + 10)  # decompressed genotype blocks: header, ploidy bytes, two bytes per sample
    # the decode kernel's algorithmic bytes: the decompressed block in, the FP32 row out
    kbytes = m * (3.0 * n + 10 + 4.0 * n)
    achieved = kbytes / (kms_ms * 1e-3) / 1e9
    return {"metric": "BGEN read s (FileBgen::read_all: decompress on host threads, decode on the device)",
            "value": round(wall, 4), "unit": "s", "higher_is_better": False, "dtype": "u8 -> f32",
            "config": {"workload": f"BGEN v1.2 layout 2, 8-bit probabilities, zlib, {n} samples x {m} variants, 5 "
                                   "Balding-Nichols pops Fst 0.05, 1% missing; minor-allele dosages, centred and "
                                   "standardised FP32 matrix in HBM", "file_MB": round(fsize / 1e6, 1)},
            "variants_per_s": round(m / wall, 1), "decompressed_GBps": round(dbytes / wall / 1e9, 3),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "k_bgen_dosage (one workgroup per variant)", "launches": launches,
                         "kernel_ms_per_read": round(kms_ms, 3),
                         "note": "the read is bound by host decompression; the device decode is a small share"},
            "cpu_baseline": cpu}


def bgen_cpu_baseline(pa, n, msample, cores):
    """The reference CLI's read_all of the first `msample` variants of the same
    simulation (the simulator seeds each variant on its own), from its logged
    reading time."""
    with tempfile.TemporaryDirectory() as td:
        pa.sim_bgen(Path(td) / "s.bgen", 2031, n, msample, 5, 0.05, miss=0.01, layout=2, compression=1, bits=8,
                    nthreads=cores)
        log(f"BGEN: reference CLI on {msample} variants x {n} samples")
        r = run_ref(["-g", "s.bgen", "--svd", "1", "-k", "2", "-p", "1", "--tol-rsvd", "0", "-n", str(cores), "-o",
                     "out", "-v", "1"], td)
    if r is None:
        return None
    wall, _, text = r
    rd = re.search(r"total elapsed reading time:\s*([0-9.eE+-]+)", text)
    read_s = float(rd.group(1)) if rd else wall
    return {"value": round(msample / read_s, 1), "unit": "variants/s", "cores": cores, "kind": "reference",
            "sample": f"reference CLI (-n {cores}) -g on the first {msample} variants x {n} samples of the same "
                      f"simulation: logged reading time {read_s:.3f} s",
            "cpu": cpu_facts()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="small shapes for a quick check")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-snps", type=int, default=40000)
    ap.add_argument("--arms", default="bf16x2,fp32,fp64",
                    help="C2 arms to run (profiling runs use bf16x2 alone); the JSON line needs all three")
    ap.add_argument("--configs", default="c2,c5,c4,c3,bgen", help="workloads to run; the JSON line needs c2")
    ap.add_argument("--no-iram", action="store_true", help="C4 without the IRAM runs (short profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU sharding is a design "
              "in DESIGN.md, not part of this build.", file=sys.stderr)
        sys.exit(2)

    import pcaone_amd as pa

    if pa.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    g = pa.Gpu(0)
    todo = args.configs.split(",")
    head = run_c2(g, pa, args) if "c2" in todo else None
    configs = {}
    for name, fn in (("c5", run_c5), ("c4", run_c4), ("c3", run_c3), ("bgen", run_bgen)):
        if name in todo:
            configs[name.upper()] = fn(g, pa, args)
            log(f"{name.upper()}: " + json.dumps({k: v for k, v in configs[name.upper()].items()
                                                  if k in ("value", "unit", "roofline")}))
    g.close()
    if head is None:
        log("partial run (--configs / --arms): no JSON line; the configs:")
        log(json.dumps(configs))
        return
    head["configs"] = configs
    print(json.dumps(head), flush=True)


if __name__ == "__main__":
    try:
        main()
    finally:
        # release the handle and the page-locked buffers while the HIP runtime
        # is fully alive, never from finalizers during interpreter shutdown
        if "pcaone_amd" in sys.modules:
            sys.modules["pcaone_amd"].shutdown()
