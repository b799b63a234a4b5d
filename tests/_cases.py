"""Shared helpers for the tests: golden-case argument parsing and oracle runs."""
from __future__ import annotations

import shlex

import numpy as np

import pcaone_oracle as O


def parse_args(args) -> dict:
    """The subset of PCAone's CLI the golden cases use, with Param's defaults
    (src/Cmd.hpp:28-50) and oversamples = max(os, k) (src/Cmd.cpp:190)."""
    a = shlex.split(str(args))
    d = {"svd": 2, "k": 10, "os": 10, "p": 20, "tol": 1e-4, "seed": 101, "rand": 1, "w": 64, "S": False,
         "scale": 1, "m": 0.0, "emu": False, "pcangsd": False, "ld": False, "ld_stats": 0}
    i = 0
    flags = {"--svd": ("svd", int), "-k": ("k", int), "--oversamples": ("os", int), "-p": ("p", int),
             "--tol-rsvd": ("tol", float), "--seed": ("seed", int), "--rand": ("rand", int), "-w": ("w", int),
             "--scale": ("scale", int), "-m": ("m", float), "--ld-stats": ("ld_stats", int)}
    while i < len(a):
        if a[i] in ("-S", "--emu", "--pcangsd", "--ld"):
            d[{"-S": "S", "--emu": "emu", "--pcangsd": "pcangsd", "--ld": "ld"}[a[i]]] = True
            i += 1
        else:
            key, typ = flags[a[i]]
            d[key] = typ(a[i + 1])
            i += 2
    d["os"] = max(d["os"], d["k"])
    return d


def payload_of(z: dict):
    """The .bed payload of a fixture: stored, or regenerated from the simulator's
    seed (larger sets) and checked against the hash the fixture holds."""
    if "payload" in z:
        return z["payload"]
    import hashlib

    import pcaone_amd as pa

    seed, n, m, npops = (int(x) for x in z["sim"])
    fst, pmin, pmax, miss = (float(x) for x in z["sim_f"])
    pay = pa.sim_bed(seed, n, m, npops, fst, pmin, pmax, miss)
    assert hashlib.sha256(pay.tobytes()).hexdigest() == str(z["payload_sha256"]), "simulator output changed"
    return pay


def oracle_matrix(payload, n, m, standardize=True):
    codes = O.decode_codes(payload, n, m)
    F, _ = O.allele_freq(codes)
    X = O.centered_matrix(codes, F)
    return O.standardize(X, F) if standardize else X


def bgen_matrix(z: dict, d: dict):
    """FileBgen::read_all's matrix of a BGEN fixture; standardised unless
    --scale 0 or --ld (which runs the operator unstandardised,
    src/Halko.cpp:285-289)."""
    import os
    import tempfile

    fd, path = tempfile.mkstemp(suffix=".bgen")
    try:
        os.write(fd, np.asarray(z["bgen"], dtype=np.uint8).tobytes())
        os.close(fd)
        X, _, _, _ = O.bgen_read(path, standardize=bool(d["scale"]) and not d["ld"])
    finally:
        os.unlink(path)
    return X


def oracle_run(z: dict):
    d = parse_args(z["args"])
    n, m = int(z["nsamples"]), int(z["nsnps"])
    if "gl" in z:  # PCAngsd on BEAGLE likelihoods: the op of the last EM solve
        P0, P1 = gl_planes(z["gl"])
        if d["svd"] == 0:  # the IRAM EM: its final U, S, V
            U, S, V, *_ = O.pcangsd_iram(P0, P1, d["k"])
            return O.ExactSvd(U, S, V, m), None, d
        orc, *_ = O.pcangsd_halko(P0, P1, d["k"], d["os"], d["p"], d["tol"], seed=d["seed"], rand=d["rand"])
        return orc, None, d
    if "counts" in z:  # CSV of counts: the reader's --scale transform, no further centring
        X = O.csv_matrix(z["counts"], d["scale"])
        orc = O.RsvdOracle(X, d["k"], d["os"], fancy=d["svd"] == 2, bands=d["w"], seed=d["seed"], rand=d["rand"],
                           perm=not d["S"])
        return orc, orc.compute_usv(d["p"], d["tol"]), d
    if "dense" in z or "bgen" in z:
        X = bgen_matrix(z, d) if "bgen" in z else O.dense_centered(z["dense"])
        if d["svd"] == 0:  # IRAM: the converged top-k singular triplets
            u, s, vt = np.linalg.svd(X, full_matrices=False)
            return O.ExactSvd(u[:, :d["k"]], s[:d["k"]], vt[:d["k"]].T, m), None, d
        orc = O.RsvdOracle(X, d["k"], d["os"], fancy=d["svd"] == 2, bands=d["w"], seed=d["seed"], rand=d["rand"],
                           perm=not d["S"])
        return orc, orc.compute_usv(d["p"], d["tol"]), d
    if d["pcangsd"]:  # PCAngsd on PLINK hard calls: the op of the last EM solve
        codes = O.decode_codes(z["payload"], n, m)
        orc, *_ = O.pcangsd_plink(codes, d["k"], d["os"], d["p"], d["tol"], seed=d["seed"], rand=d["rand"])
        return orc, None, d
    if d["emu"]:
        codes = O.decode_codes(z["payload"], n, m)
        orc, _, _ = O.emu_halko(codes, d["k"], d["os"], d["p"], d["tol"], seed=d["seed"], rand=d["rand"])
        return orc, None, d
    X = oracle_matrix(payload_of(z), n, m, bool(d["scale"]))
    ooc = {}
    if d["m"] > 0 and d["svd"] == 2:  # out-of-core winSVD: the -m block plan and permute_plink's order
        bs, _, factor = O.ooc_plan(n, m, d["k"], d["os"], 2, d["w"], d["m"])
        ooc = {"blocksize": bs, "band_factor": factor}
        if not d["S"]:
            ooc["perm_override"] = O.plink_perm_indices(n, m, 2, d["w"])
    orc = O.RsvdOracle(X, d["k"], d["os"], fancy=d["svd"] == 2, bands=d["w"], seed=d["seed"], rand=d["rand"],
                       perm=not d["S"], **ooc)
    ep = orc.compute_usv(d["p"], d["tol"])
    return orc, ep, d


def gl_planes(gl):
    """(nsites, nsamples, 3) likelihoods -> the (nsamples, nsites) planes P0, P1
    the reference keeps (src/FileBeagle.cpp:99-100)."""
    return np.ascontiguousarray(gl[:, :, 0].T, dtype=np.float64), np.ascontiguousarray(gl[:, :, 1].T,
                                                                                        dtype=np.float64)


def eig_relerr(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.abs(b)))


def geno_sums_fast(codes, ws, we):
    """pcaone_oracle.ld_geno_sums by window matrix-vector products in FP64
    (integer sums below 2^53 are exact), for windows too many for its per-pair
    loop; checked equal to it in tests/test_oracle.py."""
    y = np.where(codes == 0, 2.0, np.where(codes == 2, 1.0, 0.0))
    w = (codes != 1).astype(np.float64)
    out = []
    for i, e in zip(ws, we):
        if e <= 1:
            continue
        k = slice(i + 1, i + e)
        out.append(np.stack([y[k] @ y[i], w[k] @ y[i], y[k] @ w[i], w[k] @ w[i]], axis=1))
    return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 4), dtype=np.int64)


def csv_bytes(counts: np.ndarray, zstd: bool = True) -> bytes:
    """The counts as the reference's CSV input: one feature per line, the
    samples comma-separated, zstd-compressed (the system's libzstd) or plain."""
    text = "".join(",".join(str(int(v)) for v in row) + "\n" for row in np.asarray(counts)).encode()
    if not zstd:
        return text
    import ctypes

    z = ctypes.CDLL("libzstd.so.1")
    z.ZSTD_compressBound.restype = ctypes.c_size_t
    z.ZSTD_compress.restype = ctypes.c_size_t
    z.ZSTD_isError.restype = ctypes.c_uint
    cap = z.ZSTD_compressBound(ctypes.c_size_t(len(text)))
    dst = ctypes.create_string_buffer(cap)
    n = z.ZSTD_compress(dst, ctypes.c_size_t(cap), text, ctypes.c_size_t(len(text)), 3)
    assert not z.ZSTD_isError(ctypes.c_size_t(n))
    return dst.raw[:n]

