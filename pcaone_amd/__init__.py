"""pcaone_amd -- MI355X (gfx950) implementation of PCAone's randomized-SVD hot path.

The product is the C ABI in ``include/pcaone_gpu.h`` implemented by
``pcaone_amd/lib/libpcaone_amd.so`` (hand-written HIP kernels + C++ host
engine).  This module is a thin ctypes view of that ABI for tests and the
benchmark; it never computes anything itself and raises if the library is
missing.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import weakref
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "lib" / "libpcaone_amd.so"

PREC_FP64, PREC_FP32, PREC_SPLIT_BF16 = 0, 1, 2
RSVD_NORMAL, RSVD_FANCY = 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x2": PREC_SPLIT_BF16}

_lib = None


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class RsvdParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("variant", "k", "oversamples", "bands", "precision", "seed", "rand",
                                       "perm", "standardize", "ploidy")] + [("blocksize", C.c_uint32),
                                                                          ("band_factor", C.c_int)]


_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)
_U64 = C.POINTER(C.c_uint64)
_INT = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol include/pcaone_gpu.h declares
SIGNATURES = {
    "pcaone_gpu_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pcaone_gpu_destroy": (None, [C.c_void_p]),
    "pcaone_gpu_last_error": (C.c_char_p, [C.c_void_p]),
    "pcaone_gpu_device_count": (C.c_int, []),
    "pcaone_gpu_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "pcaone_gpu_host_free": (C.c_int, [C.c_void_p]),
    "pcaone_gpu_load_bed": (C.c_int, [C.c_void_p, _U8, C.c_uint32, C.c_uint32]),
    "pcaone_gpu_load_bed_file": (C.c_int, [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_int]),
    "pcaone_gpu_allele_freq": (C.c_int, [C.c_void_p, _D, _I64]),
    "pcaone_gpu_load_dense_f32": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]),
    "pcaone_gpu_feature_means": (C.c_int, [C.c_void_p, _D]),
    "pcaone_ld_windows": (C.c_int, [_I32, C.c_uint32, _I32, C.c_uint32, C.c_int32, _I32, _I32,
                                    C.POINTER(C.c_uint32)]),
    "pcaone_gpu_ld_r2": (C.c_int, [C.c_void_p, _I32, _I32, C.c_uint32, _D, C.c_uint64, C.POINTER(C.c_uint64)]),
    "pcaone_gpu_ld_usv": (C.c_int, [C.c_void_p, _D, C.c_uint32]),
    "pcaone_gpu_ld_r2_geno": (C.c_int, [C.c_void_p, _I32, _I32, C.c_uint32, _D, C.c_uint64, C.POINTER(C.c_uint64),
                                        _I32, C.POINTER(C.c_int)]),
    "pcaone_gpu_ld_r2_geno_noalloc": (C.c_int, [C.c_void_p, _I32, _I32, C.c_uint32, _D, C.c_uint64,
                                                C.POINTER(C.c_uint64), _I32, C.c_int]),
    "pcaone_write_eigs": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_int, _D, _D, _D, C.c_int]),
    "pcaone_gpu_grm_eig": (C.c_int, [C.c_void_p, _D, _D]),
    "pcaone_gpu_load_csv": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_int,
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "pcaone_gpu_load_bgen": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _D, _U8]),
    "pcaone_gpu_project_gl": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, _D, _D, C.c_uint32, C.c_uint32,
                                        C.c_int, C.c_int, C.c_double, _D, _D, C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_int), _D]),
    "pcaone_gpu_project_gl_matched": (C.c_int, [C.c_void_p, _I32, _U8, _D, _D, _D, C.c_uint32, C.c_uint32,
                                                C.c_int, C.c_int, C.c_double, _D, _D, C.POINTER(C.c_uint32),
                                                C.POINTER(C.c_int), _D]),
    "pcaone_gpu_dense_rows": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "pcaone_bgen_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]),
    "pcaone_sim_bgen": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int]),
    "pcaone_read_csv": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]),
    "pcaone_gpu_project": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, _D, _D, C.c_uint32, C.c_uint32, C.c_int,
                                     C.c_int, _D, _D, C.POINTER(C.c_uint32)]),
    "pcaone_gpu_inbreeding": (C.c_int, [C.c_void_p, _D, _D, _D, C.c_uint32, _D, C.c_int, C.c_double, _D, _D, _D,
                                        C.POINTER(C.c_int)]),
    "pcaone_write_hwe": (C.c_int, [C.c_char_p, C.c_char_p, _D, _D, _D, C.c_uint32]),
    "pcaone_gpu_write_residuals": (C.c_int, [C.c_void_p, C.c_char_p, _D, _D, _D, C.c_uint32, C.c_int]),
    "pcaone_write_mbim": (C.c_int, [C.c_char_p, C.c_char_p, _D, C.c_uint32]),
    "pcaone_gpu_ld_prune": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_char_p]),
    "pcaone_gpu_ld_clump": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_char_p]),
    "pcaone_write_eigvecs2": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint32, _D]),
    "pcaone_write_ld_gz": (C.c_int, [C.c_char_p, C.c_char_p, _I32, _I32, C.c_uint32, _D, C.c_uint64, C.c_int,
                                     C.c_int]),
    "pcaone_gpu_load_gl": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_int]),
    "pcaone_gpu_gl_em_maf": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _D, C.POINTER(C.c_int)]),
    "pcaone_gpu_gl_from_bed": (C.c_int, [C.c_void_p]),
    "pcaone_gpu_pcangsd_update": (C.c_int, [C.c_void_p, _D, _D, _D, C.c_int, C.c_int, C.POINTER(C.c_float), _D]),
    "pcaone_gpu_pcangsd": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, _D, _D, _D,
                                     C.POINTER(C.c_int), _D, _D, _D]),
    "pcaone_gpu_pcangsd_iram": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                          _D, _D, _D, C.POINTER(C.c_int), _D, _D, _D]),
    "pcaone_gpu_rsvd_init": (C.c_int, [C.c_void_p, C.POINTER(RsvdParams), _D, _I32]),
    "pcaone_gpu_rsvd_dims": (C.c_int, [C.c_void_p, _I64, _I64, _I64]),
    "pcaone_gpu_rsvd_perm": (C.c_int, [C.c_void_p, _I32]),
    "pcaone_gpu_rsvd_gandh": (C.c_int, [C.c_void_p, C.c_int, _D, _D]),
    "pcaone_gpu_rsvd_usv": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _D, _D, _D, _INT, _D]),
    "pcaone_gpu_rsvd_set_flags": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "pcaone_gpu_rsvd_em": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, _D, _D, _D, _INT,
                                     _D]),
    "pcaone_gpu_gram_init": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "pcaone_gpu_gram_matvec": (C.c_int, [C.c_void_p, C.c_int, _D, _D]),
    "pcaone_gpu_iram": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, _D, _D, _D, _INT, _INT]),
    "pcaone_fancy_plan": (C.c_int, [C.c_int, C.c_uint32, C.c_int, C.c_int, _U64, _INT, _INT, _INT, _INT]),
    "pcaone_ooc_plan": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "pcaone_plink_perm": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _I32]),
    "pcaone_gpu_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "pcaone_gpu_reset_stats": (C.c_int, [C.c_void_p]),
    "pcaone_gpu_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t, _I64, _D]),
    "pcaone_gpu_synchronize": (C.c_int, [C.c_void_p]),
    "pcaone_sim_bed": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                 C.c_double, C.c_double, _U8, _D, C.c_int]),
    "pcaone_sim_counts": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                    C.POINTER(C.c_float), C.c_int]),
    "pcaone_sim_gl": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.POINTER(C.c_float), C.c_int]),
}


def lib() -> C.CDLL:
    """Load the in-tree shared library (fails loudly when it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pcaone_amd.build`")
        _lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a: np.ndarray | None, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def device_count() -> int:
    return lib().pcaone_gpu_device_count()


class _PinnedPool:
    """Page-locked host buffers for result arrays, reused by byte size: a
    buffer goes back to the pool when the last array viewing it is collected,
    so repeated solves neither re-pin memory nor fault pageable pages in.

    Nothing here calls HIP while the interpreter shuts down: the finalizers
    are not run at exit, and buffers still pooled or viewed then are left to
    the process exit.  ``shutdown()`` frees the pooled ones explicitly."""

    def __init__(self, keep: int = 4):
        self.free: dict[int, list[int]] = {}
        self.keep = keep
        self.closed = False

    def array(self, shape, order="C"):
        n = int(np.prod(shape))
        nbytes = max(n, 1) * 8
        stack = self.free.get(nbytes)
        if stack:
            ptr = stack.pop()
        else:
            out = C.c_void_p()
            if lib().pcaone_gpu_host_alloc(nbytes, C.byref(out)) != 0 or not out.value:
                return np.empty(shape, order=order)  # pageable: same results, slower copy-out
            ptr = out.value
        flat = np.ctypeslib.as_array((C.c_double * max(n, 1)).from_address(ptr))
        fin = weakref.finalize(flat, self._release, ptr, nbytes)
        fin.atexit = False
        a = flat[:n]
        if order == "F":
            return a.reshape(tuple(reversed(shape))).T
        return a.reshape(shape)

    def _release(self, ptr: int, nbytes: int):
        if self.closed or sys.is_finalizing():
            return
        stack = self.free.setdefault(nbytes, [])
        if len(stack) < self.keep:
            stack.append(ptr)
        elif _lib is not None:
            _lib.pcaone_gpu_host_free(ptr)

    def drain(self):
        """Free every pooled buffer now (arrays still alive keep theirs)."""
        self.closed = True
        if _lib is not None:
            for stack in self.free.values():
                for ptr in stack:
                    _lib.pcaone_gpu_host_free(ptr)
        self.free.clear()


_pinned = _PinnedPool()
_handles: "weakref.WeakSet[Gpu]" = weakref.WeakSet()


def shutdown():
    """Close every open handle and free the pooled page-locked buffers, while
    the HIP runtime is still fully alive.  Call it before the process ends
    (the benchmark and the test session do); without it the memory is simply
    returned by the process exit, and no HIP call runs during interpreter
    shutdown."""
    for g in list(_handles):
        g.close()
    _pinned.drain()


def sim_bed(seed: int, nsamples: int, nsnps: int, npops: int = 3, fst: float = 0.1, pmin: float = 0.05,
            pmax: float = 0.5, miss: float = 0.01, nthreads: int = 0, want_p: bool = False):
    """Balding-Nichols .bed payload (host only)."""
    nb = (nsamples + 3) >> 2
    out = np.empty(nsnps * nb, dtype=np.uint8)
    panc = np.empty(nsnps) if want_p else None
    rc = lib().pcaone_sim_bed(seed, nsamples, nsnps, npops, fst, pmin, pmax, miss, _p(out, C.c_uint8),
                              _p(panc, C.c_double), nthreads)
    if rc != 0:
        raise GpuError(rc, "pcaone_sim_bed failed")
    return (out, panc) if want_p else out


def sim_counts(seed: int, ncells: int, ngenes: int, base: float = 0.05, sdlog: float = 1.5, phi: float = 0.2,
               nthreads: int = 0) -> np.ndarray:
    """scRNA-like log1p(CPM) matrix, (ngenes, ncells) float32 (host only)."""
    out = np.empty((ngenes, ncells), dtype=np.float32)
    rc = lib().pcaone_sim_counts(seed, ncells, ngenes, base, sdlog, phi, _p(out, C.c_float), nthreads)
    if rc != 0:
        raise GpuError(rc, "pcaone_sim_counts failed")
    return out


def sim_gl(seed: int, nsamples: int, nsites: int, npops: int = 4, fst: float = 0.1, pmin: float = 0.05,
           pmax: float = 0.5, depth: float = 2.0, err: float = 0.01, nthreads: int = 0) -> np.ndarray:
    """Genotype likelihoods, (nsites, nsamples, 3) float32 (host only)."""
    out = np.empty((nsites, nsamples, 3), dtype=np.float32)
    rc = lib().pcaone_sim_gl(seed, nsamples, nsites, npops, fst, pmin, pmax, depth, err, _p(out, C.c_float),
                             nthreads)
    if rc != 0:
        raise GpuError(rc, "pcaone_sim_gl failed")
    return out


def sim_bgen(path: str | os.PathLike, seed: int, nsamples: int, nsnps: int, npops: int = 3, fst: float = 0.1,
             pmin: float = 0.05, pmax: float = 0.5, miss: float = 0.0, layout: int = 2, compression: int = 1,
             bits: int = 8, pmin_declared: int = 2, sample_ids: bool = False, nthreads: int = 0):
    """A BGEN file of Balding-Nichols genotype probabilities (host only)."""
    rc = lib().pcaone_sim_bgen(str(path).encode(), seed, nsamples, nsnps, npops, fst, pmin, pmax, miss, layout,
                               compression, bits, 1 if pmin_declared == 1 else 0, 1 if sample_ids else 0, nthreads)
    if rc != 0:
        raise GpuError(rc, "pcaone_sim_bgen failed")


def bgen_info(path: str | os.PathLike) -> dict:
    """Header of a BGEN file and a walk over its variant blocks (host only)."""
    n, v, lay, comp = C.c_uint32(0), C.c_uint32(0), C.c_int(0), C.c_int(0)
    rc = lib().pcaone_bgen_info(str(path).encode(), C.byref(n), C.byref(v), C.byref(lay), C.byref(comp))
    if rc != 0:
        raise GpuError(rc, f"pcaone_bgen_info failed on {path}")
    return {"nsamples": n.value, "nvariants": v.value, "layout": lay.value, "compression": comp.value}


def write_beagle(path: str | os.PathLike, gl: np.ndarray):
    """BEAGLE text (gzip) with exact float32 values: the reference parses each
    field with strtod (src/Utils.cpp:377-405)."""
    import gzip

    m, n, _ = gl.shape
    with gzip.open(path, "wt") as f:
        f.write("marker\tallele1\tallele2\t" + "\t".join(f"Ind{j}\tInd{j}\tInd{j}" for j in range(n)) + "\n")
        for s in range(m):
            vals = "\t".join(repr(float(v)) for v in gl[s].reshape(-1))
            f.write(f"1_{s + 1}\t0\t1\t{vals}\n")


def write_bin(path: str | os.PathLike, featmajor: np.ndarray):
    """FileBin layout: u32 nfeatures, u32 nsamples, then the float32 rows."""
    a = np.ascontiguousarray(featmajor, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(np.array(a.shape, dtype=np.uint32).tobytes())
        f.write(a.tobytes())


def ld_windows(pos: np.ndarray, chr_end: np.ndarray, ld_bp: int):
    """divide_pos_by_window -> (ws, we) int32 arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    ce = np.ascontiguousarray(chr_end, dtype=np.int32)
    ws = np.empty(len(pos), dtype=np.int32)
    we = np.empty(len(pos), dtype=np.int32)
    nw = C.c_uint32(0)
    rc = lib().pcaone_ld_windows(_p(pos, C.c_int32), len(pos), _p(ce, C.c_int32), len(ce), int(ld_bp),
                                 _p(ws, C.c_int32), _p(we, C.c_int32), C.byref(nw))
    if rc != 0:
        raise GpuError(rc, "pcaone_ld_windows failed")
    return ws[:nw.value].copy(), we[:nw.value].copy()


def write_eigs(prefix: str | os.PathLike, S: np.ndarray, U: np.ndarray, V: np.ndarray | None, nsnps: int,
               ploidy: int = 2):
    """The reference's .sigvals / .eigvals / .eigvecs (/ .loadings) files."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    U = np.asfortranarray(U, dtype=np.float64)
    Vf = None if V is None else np.asfortranarray(V, dtype=np.float64)
    rc = lib().pcaone_write_eigs(str(prefix).encode(), U.shape[0], nsnps, U.shape[1], _p(S, C.c_double),
                                 _p(U, C.c_double), _p(Vf, C.c_double), ploidy)
    if rc != 0:
        raise GpuError(rc, "pcaone_write_eigs failed")


def write_ld_gz(path: str | os.PathLike, bim: str | os.PathLike, ws: np.ndarray, we: np.ndarray, r2: np.ndarray,
                level: int = 1, nthreads: int = 16):
    """The reference's --print-r2 .ld.gz (header, then bim fields of both SNPs and
    r2 per window pair), formatted and compressed on nthreads host threads."""
    ws = np.ascontiguousarray(ws, dtype=np.int32)
    we = np.ascontiguousarray(we, dtype=np.int32)
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    rc = lib().pcaone_write_ld_gz(str(path).encode(), str(bim).encode(), _p(ws, C.c_int32), _p(we, C.c_int32),
                                  len(ws), _p(r2, C.c_double), len(r2), int(level), int(nthreads))
    if rc != 0:
        raise GpuError(rc, "pcaone_write_ld_gz failed")


def read_csv(path: str | os.PathLike, nthreads: int = 4):
    """The library's CSV reader alone: (counts (features, samples) float32, libsize int64)."""
    n, m = C.c_uint32(0), C.c_uint64(0)
    rc = lib().pcaone_read_csv(str(path).encode(), nthreads, None, 0, C.byref(n), C.byref(m), None)
    if rc != 0:
        raise GpuError(rc, "pcaone_read_csv failed")
    X = np.empty((m.value, n.value), dtype=np.float32)
    ls = np.empty(n.value, dtype=np.int64)
    rc = lib().pcaone_read_csv(str(path).encode(), nthreads, _p(X, C.c_float), X.size, C.byref(n), C.byref(m),
                               _p(ls, C.c_int64))
    if rc != 0:
        raise GpuError(rc, "pcaone_read_csv failed")
    return X, ls


def write_hwe(path: str | os.PathLike, bim: str | os.PathLike, hwe_p, lrt, coef):
    """The reference's .hwe table (ID, HWE_P, LRT, Inbreeding_coefficient)."""
    h = np.ascontiguousarray(hwe_p, dtype=np.float64)
    t = np.ascontiguousarray(lrt, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    rc = lib().pcaone_write_hwe(str(path).encode(), str(bim).encode(), _p(h, C.c_double), _p(t, C.c_double),
                                _p(c, C.c_double), len(c))
    if rc != 0:
        raise GpuError(rc, "pcaone_write_hwe failed")


def write_mbim(path: str | os.PathLike, bim: str | os.PathLike, F: np.ndarray):
    """The reference's .mbim: every .bim line with its allele frequency appended."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    rc = lib().pcaone_write_mbim(str(path).encode(), str(bim).encode(), _p(F, C.c_double), len(F))
    if rc != 0:
        raise GpuError(rc, "pcaone_write_mbim failed")


def write_eigvecs2(path: str | os.PathLike, beagle: str | os.PathLike, U2: np.ndarray):
    """The reference's .eigvecs2 (GRM eigenvectors, one row per BEAGLE sample)."""
    U2 = np.ascontiguousarray(U2, dtype=np.float64)
    n = U2.shape[0]
    if U2.shape != (n, n):
        raise ValueError("U2 must be square")
    rc = lib().pcaone_write_eigvecs2(str(path).encode(), str(beagle).encode(), n, _p(U2, C.c_double))
    if rc != 0:
        raise GpuError(rc, "pcaone_write_eigvecs2 failed")


def fancy_plan(bands: int, pi: int, bandsize: int, nblocks: int = 0, band_factor: int = 1):
    """Host schedule of one winSVD computeGandH call -> (acc, update, zero, reset, new_bandsize).
    nblocks > 0: the out-of-core schedule over nblocks blocks with the band factor."""
    bs = C.c_uint64(bandsize)
    nw = nblocks if nblocks else bands
    acc = np.zeros(nw, dtype=np.intc)
    upd = np.zeros(nw, dtype=np.intc)
    zer = np.zeros(nw, dtype=np.intc)
    rst = C.c_int(0)
    rc = lib().pcaone_fancy_plan(bands, nblocks, band_factor, pi, C.byref(bs), _p(acc, C.c_int), _p(upd, C.c_int),
                                 _p(zer, C.c_int), C.byref(rst))
    if rc != 0:
        raise GpuError(rc, "pcaone_fancy_plan failed")
    return acc, upd.astype(bool), zer, bool(rst.value), bs.value


def ooc_plan(nsamples: int, nsnps: int, svd: int, k: int, oversamples: int, bands: int, memory_gb: float):
    """The -m block plan (src/Data.cpp:27-71) -> (blocksize, nblocks, band_factor)."""
    bs, nb, f = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    rc = lib().pcaone_ooc_plan(nsamples, nsnps, svd, k, oversamples, bands, memory_gb, C.byref(bs), C.byref(nb),
                               C.byref(f))
    if rc != 0:
        raise GpuError(rc, "pcaone_ooc_plan failed (one block: remove -m)")
    return bs.value, nb.value, f.value


def plink_perm(nsamples: int, nsnps: int, buffer_gb: int, nbands: int) -> np.ndarray:
    """permute_plink's SNP order (src/FilePlink.cpp:298-405): perm[ib] = input SNP at ib."""
    out = np.empty(nsnps, dtype=np.int32)
    rc = lib().pcaone_plink_perm(nsamples, nsnps, buffer_gb, nbands, _p(out, C.c_int32))
    if rc != 0:
        raise GpuError(rc, "pcaone_plink_perm failed")
    return out


class Gpu:
    """One handle: device buffers + one HIP stream (not thread-safe)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.pcaone_gpu_create(device, C.byref(h))
        if rc != 0:
            raise GpuError(rc, f"pcaone_gpu_create({device}) failed")
        self.h = h
        self.N = self.M = 0
        _handles.add(self)

    def close(self):
        if getattr(self, "h", None):
            self._L.pcaone_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        # never tear the handle down from a finalizer during interpreter
        # shutdown: the process exit releases its device memory
        if not sys.is_finalizing():
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc < 0:
            raise GpuError(rc, self._L.pcaone_gpu_last_error(self.h).decode())
        return rc

    # ---- data ---------------------------------------------------------------
    def load_bed(self, payload: np.ndarray, nsamples: int, nsnps: int):
        payload = np.ascontiguousarray(payload, dtype=np.uint8)
        if payload.size != nsnps * ((nsamples + 3) >> 2):
            raise ValueError("payload size does not match nsnps * ceil(nsamples/4)")
        self._check(self._L.pcaone_gpu_load_bed(self.h, _p(payload, C.c_uint8), nsamples, nsnps))
        self.N, self.M = nsamples, nsnps

    def load_bed_file(self, filebed, nsamples: int, nsnps: int, out_of_core: bool = False):
        """A .bed file: into HBM, or (out_of_core, -m) mapped and streamed through a device ring."""
        self._check(self._L.pcaone_gpu_load_bed_file(self.h, str(filebed).encode(), nsamples, nsnps,
                                                     int(out_of_core)))
        self.N, self.M = nsamples, nsnps

    def load_dense(self, featmajor: np.ndarray):
        """featmajor: (nfeatures, nsamples) float32, the .bin body."""
        a = np.ascontiguousarray(featmajor, dtype=np.float32)
        m, n = a.shape
        self._check(self._L.pcaone_gpu_load_dense_f32(self.h, _p(a, C.c_float), n, m))
        self.N, self.M = n, m

    def load_gl(self, gl: np.ndarray):
        """gl: (nsites, nsamples, 3 or 2) float32 genotype likelihoods (BEAGLE order)."""
        a = np.ascontiguousarray(gl, dtype=np.float32)
        m, n, per = a.shape
        self._check(self._L.pcaone_gpu_load_gl(self.h, _p(a, C.c_float), n, m, per))
        self.N, self.M = n, m

    def gl_from_bed(self):
        """--pcangsd on the loaded PLINK set: its codes become hard-call
        likelihoods, F the genotype-count frequencies (returned), and
        rsvd_init then builds the centred-genotype initial E."""
        F = self.allele_freq()[0]
        self._check(self._L.pcaone_gpu_gl_from_bed(self.h))
        return F

    def gl_em_maf(self, maxiter: int = 100, tol: float = 1e-6):
        """emMAF_with_GL; returns (F, iterations)."""
        F = np.empty(self.M)
        it = C.c_int(0)
        self._check(self._L.pcaone_gpu_gl_em_maf(self.h, maxiter, tol, _p(F, C.c_double), C.byref(it)))
        return F, it.value

    def pcangsd_update(self, U, S, V, standardize: bool = False):
        """fit_with_pi (PCAngsd) or pcangsd_standardize_E from U (N, k), S (k), V (M, k);
        returns (E as (M, N) float32, Dc or None)."""
        U = np.asfortranarray(U, dtype=np.float64)
        V = np.asfortranarray(V, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64)
        k = U.shape[1]
        E = np.empty((self.M, self.N), dtype=np.float32)
        Dc = np.empty(self.N) if standardize else None
        self._check(self._L.pcaone_gpu_pcangsd_update(self.h, _p(U, C.c_double), _p(S, C.c_double),
                                                       _p(V, C.c_double), k, int(standardize), _p(E, C.c_float),
                                                       _p(Dc, C.c_double) if Dc is not None else None))
        return E, Dc

    def pcangsd(self, maxp: int, tol: float, maxiter: int = 100, tolem: float = 1e-5, grm: bool = True,
                dc: bool = True):
        """PCAngsd EM; returns U, S, V, iterations, diffs, Dc (or None), C (or None).
        BEAGLE input: dc and grm give pcangsd_standardize_E's Dc and the GRM;
        PLINK input (gl_from_bed) passes dc=False, grm=False as the reference
        stops after the EM there."""
        U = np.empty((self.N, self.k), order="F")
        S = np.empty(self.k)
        V = np.empty((self.M, self.k), order="F")
        it = C.c_int(0)
        diffs = np.zeros(max(maxiter, 1))
        Dc = np.empty(self.N) if dc else None
        Cm = np.empty((self.N, self.N)) if grm else None
        self._check(self._L.pcaone_gpu_pcangsd(self.h, maxp, tol, maxiter, tolem, _p(U, C.c_double),
                                                _p(S, C.c_double), _p(V, C.c_double), C.byref(it),
                                                _p(diffs, C.c_double), _p(Dc, C.c_double) if dc else None,
                                                _p(Cm, C.c_double) if grm else None))
        return U, S, V, it.value, diffs[:it.value], Dc, Cm

    def pcangsd_iram(self, k: int, ncv: int | None = None, maxit: int = 1000, tol: float = 1e-6,
                     maxiter: int = 100, tolem: float = 1e-5, grm: bool = True, dc: bool = True):
        """PCAngsd EM with IRAM (--svd 0 on likelihoods) after gram_init; returns
        U, S, V, iterations, RMSE per iteration, Dc (or None), C (or None)."""
        ncv = ncv if ncv is not None else max(20, 2 * k + 1)  # src/Cmd.cpp:189
        U = np.empty((self.N, k), order="F")
        S = np.empty(k)
        V = np.empty((self.M, k), order="F")
        it = C.c_int(0)
        diffs = np.zeros(max(maxiter, 1))
        Dc = np.empty(self.N) if dc else None
        Cm = np.empty((self.N, self.N)) if grm else None
        self._check(self._L.pcaone_gpu_pcangsd_iram(self.h, k, ncv, maxit, tol, maxiter, tolem, _p(U, C.c_double),
                                                     _p(S, C.c_double), _p(V, C.c_double), C.byref(it),
                                                     _p(diffs, C.c_double), _p(Dc, C.c_double) if dc else None,
                                                     _p(Cm, C.c_double) if grm else None))
        return U, S, V, it.value, diffs[:it.value], Dc, Cm

    def load_csv(self, path, scale: int = 2, scale_factor: float = 1.0, nthreads: int = 16):
        """A (zstd-compressed) CSV of counts, one feature per line, --scale on the device."""
        n, m = C.c_uint32(0), C.c_uint32(0)
        self._check(self._L.pcaone_gpu_load_csv(self.h, str(path).encode(), int(scale), float(scale_factor),
                                                 int(nthreads), C.byref(n), C.byref(m)))
        self.N, self.M = n.value, m.value

    def load_bgen(self, path, pcangsd: bool = False, maf: float = 0.0, standardize: bool = True, ploidy: int = 2,
                  nthreads: int = 16):
        """A BGEN file: minor-allele dosages into the dense path (returns F and
        the minor allele index of every variant), or with pcangsd the genotype
        probabilities as likelihoods (returns None)."""
        nv = bgen_info(path)["nvariants"]
        F = np.empty(nv) if not pcangsd else None
        minor = np.empty(nv, dtype=np.uint8) if not pcangsd else None
        n, m = C.c_uint32(0), C.c_uint32(0)
        self._check(self._L.pcaone_gpu_load_bgen(self.h, str(path).encode(), 1 if pcangsd else 0, float(maf),
                                                  1 if standardize else 0, int(ploidy), int(nthreads), C.byref(n),
                                                  C.byref(m), _p(F, C.c_double), _p(minor, C.c_uint8)))
        self.N, self.M = n.value, m.value
        return None if pcangsd else (F, minor)

    def project(self, bim, mbim, S, Vref, method: int = 2, standardize: bool = True):
        """--project on the loaded PLINK samples: (U (N, k), W (noverlap, k) the written loadings)."""
        S = np.ascontiguousarray(S, dtype=np.float64)
        Vref = np.asfortranarray(Vref, dtype=np.float64)
        k = len(S)
        U = np.empty((self.N, k), order="F")
        W = np.empty((self.M, k), order="F")
        nm = C.c_uint32(0)
        self._check(self._L.pcaone_gpu_project(self.h, str(bim).encode(), str(mbim).encode(), _p(S, C.c_double),
                                                _p(Vref, C.c_double), k, Vref.shape[0], int(method),
                                                int(standardize), _p(U, C.c_double), _p(W, C.c_double),
                                                C.byref(nm)))
        n = nm.value
        Wm = np.asfortranarray(W.reshape(-1, order="F")[: n * k].reshape((n, k), order="F"))
        return U, Wm

    def project_gl(self, beagle, mbim, S, Vref, standardize: bool = True, maxiter: int = 100,
                   tolem: float = 1e-5):
        """--project on the loaded BEAGLE likelihoods: (U (N, k), VS (noverlap, k),
        iterations, diffs)."""
        S = np.ascontiguousarray(S, dtype=np.float64)
        Vref = np.asfortranarray(Vref, dtype=np.float64)
        k = S.size
        U = np.empty((self.N, k), order="F")
        VS = np.empty(self.M * k)
        diffs = np.empty(max(maxiter, 1))
        nm, it = C.c_uint32(0), C.c_int(0)
        self._check(self._L.pcaone_gpu_project_gl(self.h, str(beagle).encode(), str(mbim).encode(), _p(S, C.c_double),
                                                   _p(Vref, C.c_double), k, Vref.shape[0], int(standardize),
                                                   int(maxiter), float(tolem), _p(U, C.c_double),
                                                   _p(VS, C.c_double), C.byref(nm), C.byref(it),
                                                   _p(diffs, C.c_double)))
        W = VS[:nm.value * k].reshape(k, nm.value).T.copy()
        return U, W, it.value, diffs[:it.value]

    def project_gl_matched(self, ref_idx, flip, F, S, Vref, standardize: bool = True, maxiter: int = 100,
                           tolem: float = 1e-5):
        """The likelihood projection with the site matching given (per loaded site)."""
        ref_idx = np.ascontiguousarray(ref_idx, dtype=np.int32)
        flip = np.ascontiguousarray(flip, dtype=np.uint8)
        F = np.ascontiguousarray(F, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64)
        Vref = np.asfortranarray(Vref, dtype=np.float64)
        k = S.size
        U = np.empty((self.N, k), order="F")
        VS = np.empty(self.M * k)
        diffs = np.empty(max(maxiter, 1))
        nm, it = C.c_uint32(0), C.c_int(0)
        self._check(self._L.pcaone_gpu_project_gl_matched(
            self.h, _p(ref_idx, C.c_int32), _p(flip, C.c_uint8), _p(F, C.c_double), _p(S, C.c_double),
            _p(Vref, C.c_double), k, Vref.shape[0], int(standardize), int(maxiter), float(tolem), _p(U, C.c_double),
            _p(VS, C.c_double), C.byref(nm), C.byref(it), _p(diffs, C.c_double)))
        W = VS[:nm.value * k].reshape(k, nm.value).T.copy()
        return U, W, it.value, diffs[:it.value]

    def inbreeding(self, U, S, V, F, maxiter: int = 100, tolem: float = 1e-5):
        """--inbreed 1 on the loaded PLINK hard calls: (coef, lrt, hwe_p, iterations)."""
        U = np.asfortranarray(U, dtype=np.float64)
        V = np.asfortranarray(V, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        m = V.shape[0]
        coef, lrt, hwe = np.empty(m), np.empty(m), np.empty(m)
        it = C.c_int(0)
        self._check(self._L.pcaone_gpu_inbreeding(self.h, _p(U, C.c_double), _p(S, C.c_double), _p(V, C.c_double),
                                                   U.shape[1], _p(F, C.c_double), int(maxiter), float(tolem),
                                                   _p(coef, C.c_double), _p(lrt, C.c_double), _p(hwe, C.c_double),
                                                   C.byref(it)))
        return coef, lrt, hwe, it.value

    def write_residuals(self, path, U=None, S=None, V=None, ld_stats: int = 0):
        """The reference's .residuals of the loaded PLINK genotypes (ld_stats 0:
        minus U S V^T of a PCA on them; 1: the centred genotypes)."""
        if ld_stats == 0:
            U = np.asfortranarray(U, dtype=np.float64)
            V = np.asfortranarray(V, dtype=np.float64)
            S = np.ascontiguousarray(S, dtype=np.float64)
            k = U.shape[1]
        else:
            k = 0
        self._check(self._L.pcaone_gpu_write_residuals(self.h, str(path).encode(), _p(U, C.c_double) if k else None,
                                                       _p(S, C.c_double) if k else None,
                                                       _p(V, C.c_double) if k else None, k, int(ld_stats)))

    def ld_prune(self, bim, ld_bp: int, r2_tol: float, prefix):
        """--ld-r2 pruning on the handle's LD matrix: prefix.ld.prune.in / .out."""
        self._check(self._L.pcaone_gpu_ld_prune(self.h, str(bim).encode(), int(ld_bp), float(r2_tol),
                                                 str(prefix).encode()))

    def ld_clump(self, assoc, bim, out, clump_bp: int = 250000, clump_r2: float = 0.5, clump_p1: float = 1e-4,
                 clump_p2: float = 0.01, colnames: str | None = None):
        """--clump on the handle's LD matrix: the .clump text to `out`."""
        self._check(self._L.pcaone_gpu_ld_clump(self.h, str(assoc).encode(), (colnames or "").encode(),
                                                 str(bim).encode(), int(clump_bp), float(clump_r2),
                                                 float(clump_p1), float(clump_p2), str(out).encode()))

    def grm_eig(self):
        """Eigenpairs of the GRM kept by the last pcangsd(grm=True): (|lambda|
        descending, U2 (N, N) with row i = sample i), as JacobiSVD(C) orders them."""
        ev = np.empty(self.N)
        U2 = np.empty((self.N, self.N))
        self._check(self._L.pcaone_gpu_grm_eig(self.h, _p(ev, C.c_double), _p(U2, C.c_double)))
        return ev, U2

    def ld_usv(self, U: np.ndarray):
        """--USV on the loaded PLINK set: the LD matrix becomes (I - U U^T) G,
        U the (N, k) matrix applied (the reference applies what its read_usv
        makes of the .eigvecs table)."""
        U = np.asfortranarray(U, dtype=np.float64)
        self._check(self._L.pcaone_gpu_ld_usv(self.h, _p(U, C.c_double), U.shape[1]))

    def ld_r2(self, ws: np.ndarray, we: np.ndarray) -> np.ndarray:
        """r^2 of the window pairs on the loaded dense matrix, window after window."""
        ws = np.ascontiguousarray(ws, dtype=np.int32)
        we = np.ascontiguousarray(we, dtype=np.int32)
        n = C.c_uint64(0)
        self._check(self._L.pcaone_gpu_ld_r2(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws), None, 0,
                                              C.byref(n)))
        out = np.empty(n.value)
        self._check(self._L.pcaone_gpu_ld_r2(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws),
                                              _p(out, C.c_double), n.value, C.byref(n)))
        return out

    def ld_r2_geno(self, ws: np.ndarray, we: np.ndarray, sums: bool = False, config: int = -1):
        """r^2 of the window pairs straight from the loaded PLINK genotypes
        (the --ld --ld-stats 1 route), window after window; returns (r2,
        integer sums (npairs, 4) or None, tile configuration used)."""
        ws = np.ascontiguousarray(ws, dtype=np.int32)
        we = np.ascontiguousarray(we, dtype=np.int32)
        n = C.c_uint64(0)
        self._check(self._L.pcaone_gpu_ld_r2_geno(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws), None, 0,
                                                   C.byref(n), None, None))
        out = np.empty(n.value)
        sm = np.empty((n.value, 4), dtype=np.int32) if sums else None
        cfg = C.c_int(config)
        self._check(self._L.pcaone_gpu_ld_r2_geno(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws),
                                                   _p(out, C.c_double), n.value, C.byref(n),
                                                   _p(sm, C.c_int32) if sums else None, C.byref(cfg)))
        return out, sm, cfg.value

    def ld_r2_geno_noalloc(self, ws: np.ndarray, we: np.ndarray, sums: bool = False, config: int = -1):
        """ld_r2_geno through the allocation-free entry point (the handle's LD
        layout must exist); returns (r2, integer sums or None)."""
        ws = np.ascontiguousarray(ws, dtype=np.int32)
        we = np.ascontiguousarray(we, dtype=np.int32)
        n = C.c_uint64(0)
        self._check(self._L.pcaone_gpu_ld_r2_geno_noalloc(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws),
                                                           None, 0, C.byref(n), None, config))
        out = np.empty(n.value)
        sm = np.empty((n.value, 4), dtype=np.int32) if sums else None
        self._check(self._L.pcaone_gpu_ld_r2_geno_noalloc(self.h, _p(ws, C.c_int32), _p(we, C.c_int32), len(ws),
                                                           _p(out, C.c_double), n.value, C.byref(n),
                                                           _p(sm, C.c_int32) if sums else None, config))
        return out, sm

    def dense_rows(self, r0: int = 0, nrows: int | None = None) -> np.ndarray:
        """Rows of the loaded dense matrix as held in HBM, (nrows, N) float32."""
        nrows = self.M - r0 if nrows is None else nrows
        out = np.empty((nrows, self.N), dtype=np.float32)
        self._check(self._L.pcaone_gpu_dense_rows(self.h, r0, nrows, _p(out, C.c_float)))
        return out

    def feature_means(self):
        mu = np.empty(self.M)
        self._check(self._L.pcaone_gpu_feature_means(self.h, _p(mu, C.c_double)))
        return mu

    def allele_freq(self):
        F = np.empty(self.M)
        cnt = np.empty((self.M, 4), dtype=np.int64)
        self._check(self._L.pcaone_gpu_allele_freq(self.h, _p(F, C.c_double), _p(cnt, C.c_int64)))
        return F, cnt

    # ---- RSVD operator --------------------------------------------------------
    def rsvd_init(self, k: int, oversamples: int, variant: int = RSVD_FANCY, bands: int = 64,
                  precision: int = PREC_SPLIT_BF16, seed: int = 101, rand: int = 1, perm: bool = True,
                  standardize: bool = True, ploidy: int = 2, omega: np.ndarray | None = None,
                  permutation: np.ndarray | None = None, blocksize: int = 0, band_factor: int = 1):
        p = RsvdParams(variant, k, oversamples, bands, precision, seed, rand, int(perm), int(standardize), ploidy,
                       blocksize, band_factor)
        om = None if omega is None else np.asfortranarray(omega, dtype=np.float64)
        pm = None if permutation is None else np.ascontiguousarray(permutation, dtype=np.int32)
        self._check(self._L.pcaone_gpu_rsvd_init(self.h, C.byref(p), _p(om, C.c_double), _p(pm, C.c_int32)))
        self.k, self.l = k, k + oversamples

    def perm(self) -> np.ndarray:
        out = np.empty(self.M, dtype=np.int32)
        self._check(self._L.pcaone_gpu_rsvd_perm(self.h, _p(out, C.c_int32)))
        return out

    def gandh(self, pi: int, want_g: bool = True, want_h: bool = True):
        G = np.empty((self.M, self.l), order="F") if want_g else None
        H = np.empty((self.N, self.l), order="F") if want_h else None
        self._check(self._L.pcaone_gpu_rsvd_gandh(self.h, pi, _p(G, C.c_double), _p(H, C.c_double)))
        return G, H

    def usv(self, maxp: int, tol: float):
        U = _pinned.array((self.N, self.k), order="F")
        S = np.empty(self.k)
        V = _pinned.array((self.M, self.k), order="F")
        ep = C.c_int(0)
        diffs = np.zeros(maxp + 1)
        self._check(self._L.pcaone_gpu_rsvd_usv(self.h, maxp, tol, _p(U, C.c_double), _p(S, C.c_double),
                                                 _p(V, C.c_double), C.byref(ep), _p(diffs, C.c_double)))
        return U, S, V, ep.value, diffs

    def set_flags(self, update: bool, standardize: bool):
        self._check(self._L.pcaone_gpu_rsvd_set_flags(self.h, int(update), int(standardize)))

    def em(self, maxp: int, tol: float, maxiter: int = 100, tolem: float = 1e-5):
        """EMU EM-PCA (--emu); returns U, S, V, iterations, per-iteration diffs."""
        U = _pinned.array((self.N, self.k), order="F")
        S = np.empty(self.k)
        V = _pinned.array((self.M, self.k), order="F")
        it = C.c_int(0)
        diffs = np.zeros(max(maxiter, 1))
        self._check(self._L.pcaone_gpu_rsvd_em(self.h, maxp, tol, maxiter, tolem, _p(U, C.c_double),
                                                _p(S, C.c_double), _p(V, C.c_double), C.byref(it),
                                                _p(diffs, C.c_double)))
        return U, S, V, it.value, diffs[:it.value]

    # ---- IRAM -------------------------------------------------------------------
    def gram_init(self, precision: int = PREC_FP64, standardize: bool = True, ploidy: int = 2):
        self._check(self._L.pcaone_gpu_gram_init(self.h, precision, int(standardize), ploidy))

    def gram_matvec(self, side: int, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.N if side == 0 else self.M)
        self._check(self._L.pcaone_gpu_gram_matvec(self.h, side, _p(x, C.c_double), _p(y, C.c_double)))
        return y

    def iram(self, k: int, ncv: int | None = None, maxit: int = 1000, tol: float = 1e-6):
        ncv = ncv if ncv is not None else max(20, 2 * k + 1)  # src/Cmd.cpp:189
        U = np.empty((self.N, k), order="F")
        S = np.empty(k)
        V = np.empty((self.M, k), order="F")
        nconv, nops = C.c_int(0), C.c_int(0)
        self._check(self._L.pcaone_gpu_iram(self.h, k, ncv, maxit, tol, _p(U, C.c_double), _p(S, C.c_double),
                                             _p(V, C.c_double), C.byref(nconv), C.byref(nops)))
        return U, S, V, nconv.value, nops.value

    # ---- timing ---------------------------------------------------------------
    def profiling(self, on: bool):
        self._check(self._L.pcaone_gpu_set_profiling(self.h, int(on)))

    def reset_stats(self):
        self._check(self._L.pcaone_gpu_reset_stats(self.h))

    def kernel_stats(self) -> dict:
        n = self._check(self._L.pcaone_gpu_kernel_stats(self.h, -1, None, 0, None, None))
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n):
            cnt = C.c_int64()
            ms = C.c_double()
            self._check(self._L.pcaone_gpu_kernel_stats(self.h, i, buf, 128, C.byref(cnt), C.byref(ms)))
            out[buf.value.decode()] = (cnt.value, ms.value)
        return out

    def synchronize(self):
        self._check(self._L.pcaone_gpu_synchronize(self.h))


def write_plink(prefix: str | os.PathLike, payload: np.ndarray, nsamples: int, nsnps: int,
                chrom: np.ndarray | None = None, pos: np.ndarray | None = None):
    """Write .bed/.bim/.fam for a payload (used to drive the reference CLI)."""
    prefix = str(prefix)
    with open(prefix + ".bed", "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]))
        f.write(np.ascontiguousarray(payload, dtype=np.uint8).tobytes())
    with open(prefix + ".fam", "w") as f:
        for j in range(nsamples):
            f.write(f"f{j} i{j} 0 0 0 -9\n")
    if chrom is None:
        chrom = np.ones(nsnps, dtype=np.int64)
    if pos is None:
        pos = np.arange(1, nsnps + 1, dtype=np.int64) * 100
    with open(prefix + ".bim", "w") as f:
        for i in range(nsnps):
            f.write(f"{chrom[i]}\trs{i}\t0\t{pos[i]}\tA\tG\n")
