"""CPU oracle for the PCAone randomized-SVD hot path (FP64, numpy).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.  It is a restatement, written from
the reference's behaviour, of:

* PLINK ``.bed`` decode and allele frequency  (src/FilePlink.cpp:29-131,
  src/Common.hpp:57-58)
* standardisation                             (src/Data.cpp:290-299, :7)
* the Omega draw through libstdc++ ``std::default_random_engine`` (minstd_rand0)
  and ``std::normal_distribution`` / ``std::uniform_real_distribution``
  (src/Halko.cpp:14-22, src/RSVD.hpp:25-54), filled column-major
* the SNP shuffle ``std::shuffle`` with a default-seeded engine
  (src/RSVD.hpp:63-79)
* ``flipOmg``                                  (src/RSVD.hpp:81-90)
* ``NormalRsvdOpData::computeGandH`` in-core   (src/Halko.cpp:97-126)
* ``FancyRsvdOpData::computeGandH`` in-core    (src/Halko.cpp:153-226)
* ``RsvdOpData::computeUSV``                   (src/Halko.cpp:45-95)
* ``mev``                                      (src/Utils.cpp:192-198)
* output scaling ``eigvals = S^2/M*ploidy``    (src/Halko.cpp:338-343,
  src/Data.cpp:177-194)
* EMU imputation and EM loop                   (src/Data.cpp:272-287,
  src/Halko.cpp:294-336)
* FileBin centring                             (src/FileBinary.cpp:21-30)
* emMAF_with_GL, initial E, PCAngsd update, pcangsd_standardize_E, GRM
                                               (src/Utils.cpp:501-531,
  src/FileBeagle.cpp:121-143, src/Data.cpp:249-343, src/Halko.cpp:317-331)
* PCAngsd on PLINK hard calls                  (src/FilePlink.cpp:89-104,
  src/Halko.cpp:294-316)
* the IRAM EM for PCAngsd, flip_UV, rmse       (src/Arnoldi.cpp:57-104,
  src/Utils.cpp:116-167; pinned by the reference CLI's --svd 0 run: the
  RMSE it logs at each iteration, its eigenvalues, vectors and GRM)
* LD windows and r^2, the --USV adjustment    (src/LD.cpp:20-23, 125-139,
  419-442, 460-466)
* the --ld-stats 1 residual matrix of PLINK input and its r^2 from integer
  sums                                         (src/Data.cpp:196-221,
  src/LD.cpp:419-442)

The third-party pieces restated here come from libstdc++ (GCC 11.4, the
toolchain the reference builds with): the Park-Miller minimal-standard LCG
(a=16807, m=2^31-1), ``generate_canonical<double,53>`` (two draws),
Marsaglia's polar method with one cached value, and the "downscaling"
``uniform_int_distribution`` together with the two-indices-per-draw
``std::shuffle``.  Householder QR signs follow LAPACK ``dlarfg`` (numpy), which is
the same convention as Eigen ``HouseholderQR`` (R_ii = -sign(alpha)*norm).

Pinning: see tests/test_oracle.py -- brute force over all 256 byte values for
decode, closed-form allele frequencies, libstdc++ outputs from a g++-built
probe for the RNG/shuffle, numpy ``eigh`` of X X^T for the converged RSVD, and
the reference binary's own outputs on seeded inputs (tests/golden/).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# src/Common.hpp:57-58 -- code 00 -> 1, 01 -> missing, 10 -> 0.5, 11 -> 0
BED2GENO = np.array([1.0, -9.0, 0.5, 0.0])
BED_MISSING_CODE = 1
VAR_TOL = 1e-9  # src/Data.hpp:7


# ----------------------------------------------------------------------------
# PLINK decode (src/FilePlink.cpp:36-62, 70-115)
# ----------------------------------------------------------------------------
def bed_bytes_per_snp(nsamples: int) -> int:
    """(N+3)>>2, src/FilePlink.hpp:18."""
    return (nsamples + 3) >> 2


def decode_codes(payload: np.ndarray, nsamples: int, nsnps: int) -> np.ndarray:
    """2-bit codes, shape (M, N), LSB-first within each byte; padding dropped."""
    b = bed_bytes_per_snp(nsamples)
    raw = np.asarray(payload, dtype=np.uint8).reshape(nsnps, b)
    codes = np.empty((nsnps, b * 4), dtype=np.uint8)
    for s in range(4):
        codes[:, s::4] = (raw >> (2 * s)) & 3
    return codes[:, :nsamples]


def allele_freq(codes: np.ndarray):
    """F_i = sum_nonmiss BED2GENO / n_nonmiss  (src/FilePlink.cpp:36-62).

    The partial sums of {1, 0.5, 0} are exact in double, so the reference's
    sequential sum equals ((n00 + 0.5*n10) / n) evaluated once.  Returns
    (F, counts) with counts[:, c] = number of samples with code c.
    """
    counts = np.stack([(codes == c).sum(axis=1) for c in range(4)], axis=1).astype(np.int64)
    n = counts[:, 0] + counts[:, 2] + counts[:, 3]
    s = counts[:, 0].astype(np.float64) + 0.5 * counts[:, 2].astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        F = np.where(n > 0, s / np.maximum(n, 1), 0.0)
    return F, counts


def centered_matrix(codes: np.ndarray, F: np.ndarray) -> np.ndarray:
    """X (N x M): BED2GENO[c] - F_i, missing -> 0 (src/FilePlink.cpp:106-114)."""
    g = BED2GENO[codes]  # (M, N)
    X = g - F[:, None]
    X[codes == BED_MISSING_CODE] = 0.0
    return np.ascontiguousarray(X.T)


def dense_centered(featmajor: np.ndarray) -> np.ndarray:
    """FileBin::read_all (src/FileBinary.cpp:21-30): X (N x M) holds the
    float32 values cast to double, each feature minus its mean over samples."""
    X = np.asarray(featmajor, dtype=np.float32).astype(np.float64).T.copy()
    X -= X.mean(axis=0)
    return X


def csv_matrix(counts: np.ndarray, scale: int, scale_factor: float = 1.0) -> np.ndarray:
    """FileCsv::read_all with parse_csvzstd (src/FileCsv.cpp:11-62, 82-142) on a
    counts matrix (features x samples, the CSV's lines x fields): library size
    libsize_j = sum over features of the integer counts, median the integer
    midpoint of the sorted sizes (get_median on ints, src/Utils.hpp:60-75);
    scale 2 log10(c * median / libsize + 1), 3 log1p(c / libsize * factor),
    4 c / libsize * factor, 0/1 c; scale > 0 then standardises each feature
    (standardize, src/Utils.cpp:34-43: minus its mean, over its norm / sqrt(N-1)
    when that exceeds 1e-10).  Returns X (N x M), samples by features.
    (For scales 3 and 4 the reference divides by a library size it never
    fills; this restatement uses the filled one, as the library does.)"""
    c = np.asarray(counts, dtype=np.float64)
    n = c.shape[1]
    lib = np.asarray(counts, dtype=np.int64).sum(axis=0)
    v = np.sort(lib)
    med = float((v[n // 2 - 1] + v[n // 2]) // 2 if n % 2 == 0 else v[n // 2])
    if scale == 2:
        X = np.log10(c * med / lib + 1.0)
    elif scale == 3:
        X = np.log1p(c / lib * scale_factor)
    elif scale == 4:
        X = c / lib * scale_factor
    else:
        X = c.copy()
    if scale > 0:
        X = X - X.mean(axis=1, keepdims=True)
        sd = np.sqrt((X * X).sum(axis=1)) / np.sqrt(n - 1.0)
        ok = sd > 1e-10
        X[ok] /= sd[ok, None]
    return X.T.copy()


def standardize(X: np.ndarray, F: np.ndarray, ploidy: int = 2) -> np.ndarray:
    """Divide column i by sqrt(ploidy*F_i*(1-F_i)) when above VAR_TOL
    (src/Data.cpp:290-299).  Evaluation order (ploidy*F)*(1-F) as in C++."""
    sd = np.sqrt((float(ploidy) * F) * (1.0 - F))
    out = X.copy()
    ok = sd > VAR_TOL
    out[:, ok] = out[:, ok] / sd[ok]
    return out


# ----------------------------------------------------------------------------
# libstdc++ random facilities
# ----------------------------------------------------------------------------
class MinstdRand0:
    """std::minstd_rand0 == std::default_random_engine in libstdc++."""

    A, M = 16807, 2147483647
    MIN, MAX = 1, 2147483646

    def __init__(self, seed: int = 1):
        self.seed(seed)

    def seed(self, s: int):
        s = int(s) % self.M
        self.x = 1 if s == 0 else s

    def __call__(self) -> int:
        self.x = (self.A * self.x) % self.M
        return self.x


def generate_canonical(g: MinstdRand0) -> float:
    """generate_canonical<double, 53>: k = 2 draws for a range of 2^31-2."""
    r = float(g.MAX - g.MIN + 1)
    s = float(g() - g.MIN)
    s += float(g() - g.MIN) * r
    ret = s / (r * r)
    if ret >= 1.0:
        ret = math.nextafter(1.0, 0.0)
    return ret


class NormalDistribution:
    """std::normal_distribution<double>: Marsaglia polar, caches the 2nd value."""

    def __init__(self, mean=0.0, stddev=1.0):
        self.mean, self.stddev = mean, stddev
        self.saved = None

    def __call__(self, g: MinstdRand0) -> float:
        if self.saved is not None:
            ret, self.saved = self.saved, None
        else:
            while True:
                x = 2.0 * generate_canonical(g) - 1.0
                y = 2.0 * generate_canonical(g) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2.0 * math.log(r2) / r2)
            self.saved = x * mult
            ret = y * mult
        return ret * self.stddev + self.mean


def omega(nrows: int, ncols: int, seed: int = 101, rand: int = 1) -> np.ndarray:
    """RsvdOpData::initOmg (src/Halko.cpp:14-22): column-major NullaryExpr fill."""
    g = MinstdRand0(seed)
    vals = np.empty(nrows * ncols)
    if rand:
        d = NormalDistribution(0.0, 1.0)
        for i in range(nrows * ncols):
            vals[i] = d(g)
    else:
        for i in range(nrows * ncols):
            vals[i] = generate_canonical(g) * 2.0 - 1.0
    return vals.reshape(ncols, nrows).T.copy()


def _uniform_int(g: MinstdRand0, a: int, b: int) -> int:
    """uniform_int_distribution<unsigned long>{a,b}(g): the downscaling branch
    (the engine range 2^31-3 is neither 2^32-1 nor 2^64-1)."""
    urngrange = g.MAX - g.MIN
    urange = b - a
    if urngrange > urange:
        uerange = urange + 1
        scaling = urngrange // uerange
        past = uerange * scaling
        while True:
            ret = g() - g.MIN
            if ret < past:
                break
        ret //= scaling
    elif urngrange == urange:
        ret = g() - g.MIN
    else:  # pragma: no cover - not reachable for the sizes used here
        raise NotImplementedError("upscaling branch")
    return ret + a


def shuffle_indices(n: int, seed: int = 1) -> np.ndarray:
    """std::shuffle of 0..n-1 with a default_random_engine{} (src/RSVD.hpp:69-71)."""
    idx = list(range(n))
    if n == 0:
        return np.array(idx, dtype=np.int64)
    g = MinstdRand0(seed)
    urngrange = g.MAX - g.MIN
    urange = n
    if urngrange // urange >= urange:
        i = 1
        if urange % 2 == 0:
            j = _uniform_int(g, 0, 1)
            idx[i], idx[j] = idx[j], idx[i]
            i += 1
        while i != n:
            swap_range = i + 1
            x = _uniform_int(g, 0, swap_range * (swap_range + 1) - 1)
            p1, p2 = x // (swap_range + 1), x % (swap_range + 1)
            idx[i], idx[p1] = idx[p1], idx[i]
            i += 1
            idx[i], idx[p2] = idx[p2], idx[i]
            i += 1
        return np.array(idx, dtype=np.int64)
    for i in range(1, n):
        j = _uniform_int(g, 0, i)
        idx[i], idx[j] = idx[j], idx[i]
    return np.array(idx, dtype=np.int64)


# ----------------------------------------------------------------------------
# RSVD (src/Halko.cpp, src/RSVD.hpp)
# ----------------------------------------------------------------------------
def flip_omg(omg2: np.ndarray, omg: np.ndarray):
    """flipOmg (src/RSVD.hpp:81-90); returns (new Omg2, new Omg)."""
    omg = omg.copy()
    for c in range(omg.shape[1]):
        if np.abs(omg2[:, c] - omg[:, c]).sum() > 2.0 * np.abs(omg2[:, c] + omg[:, c]).sum():
            omg[:, c] *= -1.0
    return omg.copy(), omg


def householder_q(H: np.ndarray) -> np.ndarray:
    """Thin Q of a Householder QR (LAPACK dgeqrf/dorgqr; same R-sign rule as Eigen)."""
    q, _ = np.linalg.qr(H, mode="reduced")
    return q


def mev(X: np.ndarray, Y: np.ndarray) -> float:
    """src/Utils.cpp:192-198."""
    return float(np.linalg.norm(X.T @ Y, axis=0).sum() / X.shape[1])


def fancy_windows(nsnps: int, bands: int):
    """Window bounds [start, stop) for in-core winSVD: blocksize = ceil(M/bands)
    (src/Halko.cpp:178, 190-191)."""
    blocksize = int(math.ceil(nsnps / bands))
    out = []
    for b in range(bands):
        start = b * blocksize
        stop = nsnps if (b + 1) * blocksize >= nsnps else (b + 1) * blocksize
        out.append((start, stop))
    return out


def ooc_plan(nsamples: int, nsnps: int, k: int, os: int, svd: int, bands: int, memory: float):
    """Data::prepare's out-of-core block plan (src/Data.cpp:27-71) for -m
    ``memory`` GB: returns (blocksize, nblocks, band_factor).  svd: 0 IRAM,
    1 Halko, 2 winSVD.  The products 3*n*l, 2*m*l and 5*m are 32-bit unsigned
    as in the reference; the one-block case is the reference's error."""
    u32 = lambda x: int(x) & 0xFFFFFFFF
    GB = 134217728  # doubles per GB
    if svd == 0:
        blocksize = u32(math.ceil(memory * GB / nsamples))
    else:
        l = u32(k + os)
        a, b, c = u32(3 * nsamples * l), u32(2 * nsnps * l), u32(5 * nsnps)
        m = float(u32(u32(a + b) + c)) / GB
        if memory > 1.1 * m:
            m = 0.0
        blocksize = u32(math.ceil(((m + memory) * GB - a - b - c) / nsamples))
    nblocks = u32(math.ceil(nsnps / blocksize))
    if nblocks == 1:
        raise ValueError("only one block exists. please remove -m option")
    factor = 1
    if svd == 2:
        if nblocks < bands:
            blocksize = u32(math.ceil(nsnps / bands))
        else:
            factor = u32(math.ceil(nblocks / bands))
            blocksize = u32(math.ceil(nsnps / (bands * factor)))
        nblocks = u32(math.ceil(nsnps / blocksize))
    return blocksize, nblocks, factor


def plink_perm_indices(nsamples: int, nsnps: int, gb: int, nbands: int) -> np.ndarray:
    """The SNP order permute_plink writes (src/FilePlink.cpp:298-405): perm[ib] =
    ia, SNP ia of the input at position ib of the permuted file.  Buffers of
    twoGB_snps SNPs (a multiple of nbands) are dealt round-robin into the bands,
    band b starting at bandidx[b]; the last buffer may be short."""
    bps = (nsamples + 3) >> 2
    two = int(math.floor(1073741824.0 * gb / bps))
    if two > nsnps:
        two = nsnps
    bufsize = two // nbands
    two = bufsize * nbands
    nblocks = (nsnps + two - 1) // two
    modr2 = nsnps % two
    modr = nsnps % nbands
    bandsize = (nsnps + nbands - 1) // nbands
    if modr == 0:
        bandidx = [i * bandsize for i in range(nbands)]
    else:
        bandidx = [i * bandsize if i < modr else modr * bandsize + (bandsize - 1) * (i - modr) for i in range(nbands)]
    bufidx = bufsize
    perm = np.full(nsnps, -1, dtype=np.int64)
    for i in range(nblocks):
        if i == nblocks - 1 and modr2 != 0:
            two2 = nsnps - (nblocks - 1) * two
            bufsize = (two2 + nbands - 1) // nbands
            modr2 = two2 % nbands
        for b in range(nbands):
            j = 0
            for j in range(bufsize - 1):
                perm[i * bufidx + bandidx[b] + j] = i * two + j * nbands + b
            j = bufsize - 1 if bufsize >= 1 else 0
            if i != nblocks - 1 or (i == nblocks - 1 and b < modr2) or modr2 == 0:
                perm[i * bufidx + bandidx[b] + j] = i * two + j * nbands + b
    if (perm < 0).any() or len(np.unique(perm)) != nsnps:
        raise ValueError("permute_plink's interleave is not a permutation for these sizes")
    return perm


def ooc_windows(nsnps: int, blocksize: int):
    """Block bounds [start, stop) of the out-of-core plan (src/Data.cpp:66-70)."""
    nb = int(math.ceil(nsnps / blocksize))
    return [(b * blocksize, min((b + 1) * blocksize, nsnps)) for b in range(nb)]


class FancyState:
    """The bandsize / H1 / H2 state that FancyRsvdOpData keeps between calls.
    In core nblocks = bands and band_factor = 1; out of core (-m) the bands
    are cut into nblocks blocks and the schedule counts blocks, starting at
    band_factor (src/Halko.cpp:227-269)."""

    def __init__(self, bands: int, nblocks: int | None = None, band_factor: int = 1):
        self.bands = bands
        self.nblocks = bands if nblocks is None else nblocks
        self.factor = band_factor
        self.bandsize = 1


def fancy_epoch_plan(state: FancyState, pi: int):
    """Advance ``state`` by one computeGandH call and return its plan:
    (reset_both, [(b, acc, update, zero), ...]) following src/Halko.cpp:159-223
    (in core) and src/Halko.cpp:227-269 (out of core)."""
    bands, nblocks, factor = state.bands, state.nblocks, state.factor
    reset = 2 ** pi >= bands
    if pi == 0:
        state.bandsize = factor
    state.bandsize = min(state.bandsize * 2, nblocks)
    bandsize = state.bandsize
    steps = []
    i = 1
    for b in range(nblocks):
        acc = 1 if i <= bandsize // 2 else 2
        adjacent = pi > 0 and (b + 1) == 2 ** (pi - 1) * factor and 2 ** pi < bands
        update, zero = False, 0
        if not ((b + 1) < bandsize and not adjacent):
            if (i == bandsize) or (i == bandsize // 2) or adjacent:
                update = True
                if i == bandsize:
                    zero = 1
                    i = 0
                else:
                    zero = 2
        steps.append((b, acc, update, zero))
        i += 1
    return reset, steps


@dataclass
class RsvdOracle:
    """FP64 restatement of RsvdOpData + Normal/Fancy computeGandH (in-core)."""

    X: np.ndarray                 # N x M, already centred / standardised
    k: int
    os: int
    fancy: bool = False
    bands: int = 64
    seed: int = 101
    rand: int = 1
    perm: bool = True             # winSVD shuffles SNP columns at pi == 0
    use_mev: bool = True
    omega_override: np.ndarray | None = None
    # out of core (-m): block length and band factor of ooc_plan, and the
    # permutation of permute_plink in place of the in-core shuffle
    blocksize: int = 0
    band_factor: int = 1
    perm_override: np.ndarray | None = None
    # state
    omg: np.ndarray = field(init=False, default=None)
    omg2: np.ndarray = field(init=False, default=None)
    permutation: np.ndarray | None = field(init=False, default=None)
    trace: list = field(init=False, default_factory=list)

    def __post_init__(self):
        self.N, self.M = self.X.shape
        self.l = self.k + self.os
        self.H1 = np.zeros((self.N, self.l))
        self.H2 = np.zeros((self.N, self.l))
        if self.blocksize:
            self.windows = ooc_windows(self.M, self.blocksize)
        else:
            self.windows = fancy_windows(self.M, self.bands)
        self.state = FancyState(self.bands, len(self.windows), self.band_factor)
        self.init_omg()

    def init_omg(self):
        if self.omega_override is not None:
            self.omg = self.omega_override.copy()
        else:
            self.omg = omega(self.N, self.l, self.seed, self.rand)
        self.omg2 = self.omg.copy()

    def _refresh(self, H):
        q = householder_q(H)
        self.omg2, self.omg = flip_omg(self.omg2, q)

    def compute_gandh(self, pi: int):
        if not self.fancy:
            if pi == 0:
                self.init_omg()
            else:
                self._refresh(self.H)
            G = self.X.T @ self.omg
            self.H = self.X @ G
            return G, self.H
        if pi == 0:
            self.init_omg()
        reset, steps = fancy_epoch_plan(self.state, pi)
        if reset:
            self.H1[:] = 0.0
            self.H2[:] = 0.0
        if pi == 0 and self.perm and self.permutation is None:
            self.permutation = (np.asarray(self.perm_override) if self.perm_override is not None
                                else shuffle_indices(self.M))
            self.X = np.ascontiguousarray(self.X[:, self.permutation])
        G = np.zeros((self.M, self.l))
        H = getattr(self, "H", np.zeros((self.N, self.l)))
        for (b, (s, e)), (_, acc, update, zero) in zip(enumerate(self.windows), steps):
            Xb = self.X[:, s:e]
            G[s:e] = Xb.T @ self.omg
            if acc == 1:
                self.H1 += Xb @ G[s:e]
            else:
                self.H2 += Xb @ G[s:e]
            if update:
                H = self.H1 + self.H2
                self._refresh(H)
                if zero == 1:
                    self.H1[:] = 0.0
                else:
                    self.H2[:] = 0.0
                self.trace.append((pi, b))
        self.H = H
        return G, H

    def compute_usv(self, p: int, tol: float):
        """RsvdOpData::computeUSV (src/Halko.cpp:45-95).  Returns epochs run."""
        k, l = self.k, self.l
        upre = None
        pi = 0
        epochs = 0
        self.diffs = []
        while pi <= p:
            G, H = self.compute_gandh(pi)
            epochs += 1
            q1, r1 = np.linalg.qr(G, mode="reduced")
            q2, r2 = np.linalg.qr(q1, mode="reduced")
            R = r2 @ r1
            B = np.linalg.solve(R.T, H.T)          # R^T B = H^T
            usvd, s, vt = np.linalg.svd(B, full_matrices=False)
            ucur = vt.T[:, :k]
            if pi > 0:
                diff = 1.0 - mev(ucur, upre) if self.use_mev else None
                self.diffs.append(diff)
                if diff < tol or pi == p:
                    if self.fancy and 2 ** pi < self.bands:
                        p = int(math.log2(self.bands))
                    else:
                        self.U = ucur
                        self.V = q2 @ usvd[:, :k]
                        self.S = s[:k]
                        break
                else:
                    upre = ucur
            else:
                upre = ucur
            pi += 1
        self.U = ucur
        return epochs

    def loadings_unpermuted(self) -> np.ndarray:
        """data->perm * V (src/Halko.cpp:339-340): row j of V belongs to SNP perm[j]."""
        if self.permutation is None:
            return self.V
        out = np.empty_like(self.V)
        out[self.permutation] = self.V
        return out

    def eigvals(self, ploidy: int = 2) -> np.ndarray:
        return self.S ** 2 / self.M * ploidy


@dataclass
class ExactSvd:
    """Top-k singular triplets of X (numpy LAPACK): what IRAM converges to
    (src/Arnoldi.cpp:57-68, PartialSVDSolver)."""

    U: np.ndarray
    S: np.ndarray
    V: np.ndarray
    M: int

    def eigvals(self, ploidy: int = 2) -> np.ndarray:
        return self.S ** 2 / self.M * ploidy

    def loadings_unpermuted(self) -> np.ndarray:
        return self.V


def emu_impute(X0: np.ndarray, miss: np.ndarray, F: np.ndarray, U, S, V) -> np.ndarray:
    """EMU update of the missing entries (src/FilePlink.cpp:253-261,
    src/Data.cpp:272-287): clamp(sum_k U_jk S_k V_ik, -F_i, 1 - F_i)."""
    pred = (U * S) @ V.T
    lo = np.broadcast_to(-F[None, :], pred.shape)
    hi = np.broadcast_to(1.0 - F[None, :], pred.shape)
    X = X0.copy()
    X[miss] = np.minimum(np.maximum(pred[miss], lo[miss]), hi[miss])
    return X


def emu_halko(codes: np.ndarray, k: int, os: int, maxp: int, tol: float, maxiter: int = 100,
              tolem: float = 1e-5, seed: int = 101, rand: int = 1, ploidy: int = 2):
    """run_pca_with_halko's EM branch for --emu with the plain RSVD
    (src/Halko.cpp:294-336, data read through FileBed::read_block_update,
    src/FilePlink.cpp:226-294).  Returns (oracle op after the final
    standardised solve, EM iterations, per-iteration diffs)."""
    F, _ = allele_freq(codes)
    X0 = centered_matrix(codes, F)
    miss = (codes == BED_MISSING_CODE).T
    op = RsvdOracle(X0.copy(), k, os, fancy=False, seed=seed, rand=rand)
    op.compute_usv(maxp, tol)
    diffs = []
    it = 0
    for it in range(1, maxiter + 1):
        vpre = op.V.copy()
        op.X = emu_impute(X0, miss, F, op.U, op.S, op.V)
        op.compute_usv(maxp, tol)
        diffs.append(1.0 - mev(op.V, vpre))
        if diffs[-1] < tolem:
            break
    op.X = standardize(emu_impute(X0, miss, F, op.U, op.S, op.V), F, ploidy)
    op.compute_usv(maxp, tol)
    return op, it, diffs


# ----------------------------------------------------------------------------
# genotype likelihoods: emMAF_with_GL and PCAngsd (BEAGLE input)
# ----------------------------------------------------------------------------
def _posterior(P0, P1, pt):
    """(p1 + 2 p2) / (p0 + p1 + p2) with the third GL = 1 - P0 - P1
    (src/Utils.cpp:511-515, src/Data.cpp:262-267)."""
    p0 = P0 * (1.0 - pt) * (1.0 - pt)
    p1 = P1 * 2.0 * pt * (1.0 - pt)
    p2 = (1.0 - P0 - P1) * pt * pt
    return (p1 + 2.0 * p2) / (p0 + p1 + p2), p0, p1, p2


def em_maf_with_gl(P0: np.ndarray, P1: np.ndarray, maxiter: int, tolmaf: float, f0: float = 0.25):
    """emMAF_with_GL (src/Utils.cpp:501-531) from F = 0.25 (src/FileBeagle.cpp:127).
    P0, P1: (N, M) -- the first two GLs of each sample and site.  Returns
    (F, iterations)."""
    n, m = P0.shape
    F = np.full(m, f0)
    it = 0
    for it in range(1, maxiter + 1):
        Ft = F.copy()
        post, *_ = _posterior(P0, P1, F[None, :])
        F = post.sum(axis=0) * (1.0 / (2.0 * n))
        if math.sqrt(((F - Ft) ** 2).sum() / m) < tolmaf:
            break
    return F, it


def gl_initial_E(P0, P1, F):
    """Initial E of FileBeagle::read_all (src/FileBeagle.cpp:133-143): posterior
    mean genotype at pt = F minus 2F; (N, M)."""
    post, *_ = _posterior(P0, P1, F[None, :])
    return post - 2.0 * F[None, :]


def pcangsd_E(P0, P1, F, U, S, V):
    """fit_with_pi, PCAngsd branch (src/Data.cpp:249-270)."""
    pt = ((U * S) @ V.T + 2.0 * F[None, :]) / 2.0
    pt = np.minimum(np.maximum(pt, 1e-4), 1.0 - 1e-4)
    post, *_ = _posterior(P0, P1, pt)
    return post - 2.0 * F[None, :]


def pcangsd_standardize(P0, P1, F, U, S, V):
    """pcangsd_standardize_E (src/Data.cpp:301-343): (E / sqrt(2F(1-F)), Dc)."""
    pt = ((U * S) @ V.T + 2.0 * F[None, :]) / 2.0
    pt = np.minimum(np.maximum(pt, 1e-4), 1.0 - 1e-4)
    post, p0, p1, p2 = _posterior(P0, P1, pt)
    E = post - 2.0 * F[None, :]
    norm = np.sqrt(2.0 * F * (1.0 - F))
    E = np.where(norm[None, :] > VAR_TOL, E / norm[None, :], E)
    ps = p0 + p1 + p2
    f2 = 2.0 * F[None, :]
    tmp = (0.0 - f2) ** 2 * (p0 / ps) + (1.0 - f2) ** 2 * (p1 / ps) + (2.0 - f2) ** 2 * (p2 / ps)
    Dc = (tmp / (2.0 * F * (1.0 - F))[None, :]).sum(axis=1)
    return E, Dc


def pcangsd_halko(P0, P1, k: int, os: int, maxp: int, tol: float, maxiter: int = 100, tolem: float = 1e-5,
                  tolmaf: float = 1e-6, seed: int = 101, rand: int = 1):
    """run_pca_with_halko's EM branch for PCAngsd on BEAGLE input with the plain
    RSVD (src/Halko.cpp:294-331): returns (op of the last solve, F, EM
    iterations, diffs, Dc, GRM)."""
    n, m = P0.shape
    F, _ = em_maf_with_gl(P0, P1, maxiter, tolmaf)
    op = RsvdOracle(gl_initial_E(P0, P1, F), k, os, fancy=False, seed=seed, rand=rand)
    op.compute_usv(maxp, tol)
    diffs, it = [], 0
    for it in range(1, maxiter + 1):
        vpre = op.V.copy()
        op.X = pcangsd_E(P0, P1, F, op.U, op.S, op.V)
        op.compute_usv(maxp, tol)
        diffs.append(1.0 - mev(op.V, vpre))
        if diffs[-1] < tolem:
            break
    E, Dc = pcangsd_standardize(P0, P1, F, op.U, op.S, op.V)
    C = E @ E.T / m
    C[np.diag_indices(n)] = Dc / m
    return op, F, it, diffs, Dc, C


def flip_uv(U: np.ndarray, V: np.ndarray):
    """flip_UV with ubase (src/Utils.cpp:116-131): a column pair changes sign when
    U's entry of largest magnitude (the first on ties) is negative."""
    U, V = U.copy(), V.copy()
    for i in range(U.shape[1]):
        x = int(np.argmax(np.abs(U[:, i])))
        if U[x, i] < 0:
            U[:, i] *= -1.0
            V[:, i] *= -1.0
    return U, V


def rmse_flip(X: np.ndarray, Y: np.ndarray) -> float:
    """rmse (src/Utils.cpp:154-167): Y's columns take the sign nearer X (flip_Y),
    then the root mean square of X - Y."""
    Z = Y.copy()
    for i in range(X.shape[1]):
        if np.abs(X[:, i] - Z[:, i]).sum() > 2.0 * np.abs(X[:, i] + Z[:, i]).sum():
            Z[:, i] *= -1.0
    return float(np.sqrt(((X - Z) ** 2).mean()))


def pcangsd_iram(P0, P1, k: int, maxiter: int = 100, tolem: float = 1e-5, tolmaf: float = 1e-6,
                 refresh_vectors: bool = False):
    """run_pca_with_arnoldi's in-core EM branch for PCAngsd (src/Arnoldi.cpp:57-104)
    with an exact eigendecomposition of the smaller Gram matrix in place of IRAM.
    The reference's PartialSVDSolver hands out the eigenvectors of its first
    solve on every later matrix_U / matrix_V call (it caches m_evecs,
    external/Spectra/contrib/PartialSVDSolver.h:177-206), so after the first
    solve only the singular values change: the smaller side's vectors stay, the
    other side is E times them over S.  refresh_vectors=True takes every solve's
    own vectors instead (a planted mistake for the tests).  Returns (U, S, V, F,
    iterations, RMSEs, Dc, GRM)."""
    n, m = P0.shape
    F, _ = em_maf_with_gl(P0, P1, 100, tolmaf)
    E = gl_initial_E(P0, P1, F)
    tall = n > m

    def eig_top(E):
        G = E.T @ E if tall else E @ E.T
        w, Q = np.linalg.eigh(G)
        o = np.argsort(w)[::-1][:k]
        return w[o], Q[:, o]

    def usv(E, lam, ev):
        s = np.sqrt(lam)
        return (E @ ev / s, s, ev.copy()) if tall else (ev.copy(), s, E.T @ ev / s)

    lam, ev0 = eig_top(E)
    U, S, V = usv(E, lam, ev0)
    U, V = flip_uv(U, V)
    diffs, it = [], 0
    for it in range(1, maxiter + 1):
        E = pcangsd_E(P0, P1, F, U, S, V)
        lam, evn = eig_top(E)
        U, S, V2 = usv(E, lam, evn if refresh_vectors else ev0)
        U, V2 = flip_uv(U, V2)
        diffs.append(rmse_flip(V2, V))
        V = V2
        if diffs[-1] < tolem:
            break
    Es, Dc = pcangsd_standardize(P0, P1, F, U, S, V)
    C = Es @ Es.T / m
    C[np.diag_indices(n)] = Dc / m
    return U, S, V, F, it, diffs, Dc, C


def plink_hard_call_gl(codes: np.ndarray):
    """The two GLs FileBed::read_all keeps per sample for --pcangsd on PLINK
    input (src/FilePlink.cpp:89-104): code 3 (RR) (1, 0), code 0 (AA) (0, 0),
    code 2 (RA) (0, 1), missing (0.333333, 0.333333).  Returns (P0, P1), (N, M)."""
    conds = [codes == 3, codes == 0, codes == 2]
    P0 = np.select(conds, [1.0, 0.0, 0.0], 0.333333)
    P1 = np.select(conds, [0.0, 0.0, 1.0], 0.333333)
    return np.ascontiguousarray(P0.T), np.ascontiguousarray(P1.T)


def pcangsd_plink(codes: np.ndarray, k: int, os: int, maxp: int, tol: float, maxiter: int = 100,
                  tolem: float = 1e-5, seed: int = 101, rand: int = 1):
    """run_pca_with_halko's EM branch for --pcangsd on in-core PLINK input with
    the plain RSVD (src/Halko.cpp:294-316): the first solve on the centred
    genotypes (src/FilePlink.cpp:106-114), then E <- fit_with_pi's PCAngsd
    update (src/Data.cpp:249-270) on the hard-call GLs with the PLINK allele
    frequencies, until 1 - mev(V, Vpre) < tolem.  No GRM: that step is for
    BEAGLE input only (src/Halko.cpp:317).  Returns (op, F, iterations, diffs)."""
    F, _ = allele_freq(codes)
    P0, P1 = plink_hard_call_gl(codes)
    op = RsvdOracle(centered_matrix(codes, F), k, os, fancy=False, seed=seed, rand=rand)
    op.compute_usv(maxp, tol)
    diffs, it = [], 0
    for it in range(1, maxiter + 1):
        vpre = op.V.copy()
        op.X = pcangsd_E(P0, P1, F, op.U, op.S, op.V)
        op.compute_usv(maxp, tol)
        diffs.append(1.0 - mev(op.V, vpre))
        if diffs[-1] < tolem:
            break
    return op, F, it, diffs


# ----------------------------------------------------------------------------
# LD r^2 on a dense (residual) matrix
# ----------------------------------------------------------------------------
def ld_windows(pos, chr_end, ld_bp: int):
    """divide_pos_by_window (src/LD.cpp:125-139), loop for loop: every SNP whose
    position differs from its chromosome's last position leads a window of
    the SNPs j >= i of that chromosome with pos[j] - pos[i] <= ld_bp.
    chr_end: 0-based index of each chromosome's last SNP.  Returns (ws, we),
    we counting the lead itself."""
    pos = [int(p) for p in pos]
    ws, we = [], []
    c = 0
    for i in range(len(pos)):
        if pos[i] == pos[chr_end[c]]:
            c += 1
            continue
        j = i
        while j <= chr_end[c]:
            if pos[j] - pos[i] > ld_bp:
                break
            j += 1
        ws.append(i)
        we.append(j - i)
    return np.array(ws, dtype=np.int64), np.array(we, dtype=np.int64)


def read_usv(table: np.ndarray) -> np.ndarray:
    """The matrix read_usv (src/Utils.cpp:255-285) returns for a text table
    (rows x k): it collects the values row after row and maps that buffer as a
    column-major rows x k matrix, so element (r, c) is the table's value at
    row-major position c * rows + r."""
    rows, k = table.shape
    return np.ascontiguousarray(np.asarray(table, dtype=np.float64).ravel().reshape(k, rows).T)


def ld_usv_matrix(codes: np.ndarray, U: np.ndarray) -> np.ndarray:
    """run_ld_stuff's --USV branch for PLINK input (src/LD.cpp:460-466):
    FileBed::read_all's centred genotypes G (N x M, missing 0) adjusted as
    (I - U U^T) G with U the matrix read_usv returns (N x k), not re-centred;
    returned feature-major (M x N)."""
    F, _ = allele_freq(codes)
    G = centered_matrix(codes, F)
    n = G.shape[0]
    return np.ascontiguousarray(((np.eye(n) - U @ U.T) @ G).T)


def ld_r2(featmajor: np.ndarray, ws, we, centre: bool = True) -> np.ndarray:
    """ld_r2_big (src/LD.cpp:419-442) on FileBin::read_all's matrix
    (src/FileBinary.cpp:21-30: float32 -> double, each feature centred by its
    mean; centre=False for the --USV matrix, used as it is): r = <x_i, x_k>
    (sds_i sds_k / (N - 1)), sds = 1 / sqrt(sum x^2 / (N - 1)) (calc_sds,
    :20-23).  r^2 for the pairs (ws[w], ws[w] + j), j = 1 .. we[w] - 1, window
    after window."""
    X = featmajor.astype(np.float64)
    if centre:
        X = X - X.mean(axis=1, keepdims=True)
    n = X.shape[1]
    df = 1.0 / (n - 1)
    sds = 1.0 / np.sqrt((X * X).sum(axis=1) * df)
    out = []
    for i, e in zip(ws, we):
        if e <= 1:
            continue
        k = np.arange(i + 1, i + e)
        r = (X[k] @ X[i]) * (sds[i] * sds[k] * df)
        out.append(r * r)
    return np.concatenate(out) if out else np.zeros(0)


def ld_stats1_matrix(codes: np.ndarray) -> np.ndarray:
    """The matrix Data::write_residuals writes with --ld-stats 1 for in-core
    PLINK input (src/Data.cpp:196-221): FileBed::read_all's G = BED2GENO - F
    with missing 0 (unstandardised: --ld sets setFlags(false, false),
    src/Halko.cpp:284-288), each SNP centred by its mean over samples;
    returned feature-major (M x N) in FP64 (the reference then casts to
    float32)."""
    F, _ = allele_freq(codes)
    G = centered_matrix(codes, F)
    G = G - G.mean(axis=0, keepdims=True)
    return np.ascontiguousarray(G.T)


def ld_geno_sums(codes: np.ndarray, ws, we) -> np.ndarray:
    """The integer sums behind the --ld-stats 1 r^2 of each window pair (i, k),
    window after window: with y the A1 dosage (0, 1, 2; missing 0) and w the
    non-missing mask, (sum y_i y_k, sum y_i w_k, sum w_i y_k, sum w_i w_k) over
    samples, as int64 (npairs x 4).  Brute force: one dot product per pair."""
    y = np.where(codes == 0, 2, np.where(codes == 2, 1, 0)).astype(np.int64)  # M x N
    w = (codes != BED_MISSING_CODE).astype(np.int64)
    out = []
    for i, e in zip(ws, we):
        for k in range(i + 1, i + e):
            out.append((int(y[i] @ y[k]), int(y[i] @ w[k]), int(w[i] @ y[k]), int(w[i] @ w[k])))
    return np.array(out, dtype=np.int64).reshape(-1, 4)


def ld_geno_r2(codes: np.ndarray, ws, we, sums: np.ndarray | None = None) -> np.ndarray:
    """r^2 of ld_r2_big (src/LD.cpp:419-442) on ld_stats1_matrix, from the
    integer sums in closed form.  With a = w (y / 2 - F) (the centred,
    missing-zero genotype), N mu = sum a = sum y / 2 - F n and
        sum a_i a_k = S_yy / 4 - F_k S_yw / 2 - F_i S_wy / 2 + F_i F_k S_ww,
        cov = sum a_i a_k - (N mu_i)(N mu_k) / N,
        ss_i = sum y_i^2 / 4 - F_i sum y_i + F_i^2 n_i - (N mu_i)^2 / N,
    r^2 = cov^2 / (ss_i ss_k) (the (N - 1) factors of calc_cor cancel)."""
    N = codes.shape[1]
    F, _ = allele_freq(codes)
    y = np.where(codes == 0, 2, np.where(codes == 2, 1, 0)).astype(np.int64)
    w = (codes != BED_MISSING_CODE).astype(np.int64)
    sy, syy, nn = y.sum(axis=1), (y * y).sum(axis=1), w.sum(axis=1)
    mun = 0.5 * sy - F * nn
    ss = 0.25 * syy - F * sy + F * F * nn - mun * mun / N
    if sums is None:
        sums = ld_geno_sums(codes, ws, we)
    ia = np.concatenate([np.full(e - 1, i) for i, e in zip(ws, we) if e > 1]) if len(sums) else np.zeros(0, int)
    ka = np.concatenate([np.arange(i + 1, i + e) for i, e in zip(ws, we) if e > 1]) if len(sums) else np.zeros(0, int)
    s = sums.astype(np.float64)
    cov = (0.25 * s[:, 0] - 0.5 * F[ka] * s[:, 1] - 0.5 * F[ia] * s[:, 2] + F[ia] * F[ka] * s[:, 3]
           - mun[ia] * mun[ka] / N)
    # a SNP without variation (monomorphic or all missing; ss is then exactly
    # 0) has an all-zero residual column and r = 0 * (1 / 0) = NaN in the
    # reference (calc_sds, ld_r2_big: src/LD.cpp:20-23, 427)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where((ss[ia] > 0) & (ss[ka] > 0), cov * cov / (ss[ia] * ss[ka]), np.nan)


def principal_angles_max(A: np.ndarray, B: np.ndarray) -> float:
    """Largest principal angle (radians) between span(A) and span(B), in FP64
    whatever the inputs' type (an FP32 QR leaves cos within 1e-7 of 1, where
    arccos resolves nothing below 4e-4)."""
    qa, _ = np.linalg.qr(np.asarray(A, dtype=np.float64))
    qb, _ = np.linalg.qr(np.asarray(B, dtype=np.float64))
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def inbreeding_em(codes: np.ndarray, U, S, V, F, maxiter: int = 100, tolem: float = 1e-5):
    """run_inbreeding_em with type 1 (src/Inbreeding.cpp:16-190) on hard calls
    (codes M x N as decode_codes returns them, 0..3 as .bed codes): Pi = clamp((U S V^T + 2F) / 2, 1e-4,
    1 - 1e-4) (src/FileUSV.cpp:11-22); per site the EM update of the
    inbreeding coefficient from observed / expected heterozygosity, with the
    SQUAREM step (alpha clamped to [1, 256], the coefficients to [-1, 1], the
    stabilising pass, RMSE < tolem) and AreClose(sv2, 0); then the LRT
    2 (log L_alt - log L_null) and its chi-square(1) p-value (kf_gammaq(1/2,
    x/2) = erfc(sqrt(x / 2)); NaN -> 1).  Returns (coef, lrt, p, iterations)."""
    from math import erfc, isnan, sqrt

    PI = np.clip((U @ np.diag(S) @ V.T + 2.0 * F[None, :]) * 0.5, 1e-4, 1.0 - 1e-4)
    c = np.asarray(codes).T  # samples x sites, as Pi

    def priors(fc, lrt):
        fadj = (1.0 - PI) * PI * fc[None, :]
        p0 = np.maximum(1e-4, (1.0 - PI) ** 2 + fadj)
        if lrt:
            p1 = np.maximum(1e-4, 2.0 * PI * (1.0 - PI) - 2.0 * fadj)
        else:
            p1 = np.maximum(1e-4, 2.0 * PI * (1.0 - PI) * (1.0 - fc[None, :]))
        p2 = np.maximum(1e-4, PI * PI + fadj)
        ps = 1.0 / (p0 + p1 + p2)
        return p0 * ps, p1 * ps, p2 * ps

    def coef_pass(fc):
        p0, p1, p2 = priors(fc, False)
        q0 = np.where(c == 3, p0, np.where(c == 1, 0.333333 * p0, 0.0))
        q1 = np.where(c == 2, p1, np.where(c == 1, 0.333333 * p1, 0.0))
        q2 = np.where(c == 0, p2, np.where(c == 1, 0.333333 * p2, 0.0))
        obs = (q1 / (q0 + q1 + q2)).sum(axis=0)
        exp = (2.0 * PI * (1.0 - PI)).sum(axis=0)
        f = np.clip(1.0 - obs / exp, -1.0, 1.0)
        return f, f - fc

    def close0(a):
        return a == 0.0

    Fc = np.zeros(V.shape[0])
    it = 0
    for it in range(maxiter):
        F0 = Fc.copy()
        Fc, D1 = coef_pass(Fc)
        sr2 = float((D1 * D1).sum())
        Fc, D2 = coef_pass(Fc)
        sv2 = float(((D2 - D1) ** 2).sum())
        if close0(sv2):
            break
        alpha = min(max(1.0, sqrt(sr2 / sv2)), 256.0)
        Fc = np.clip(F0 + 2 * alpha * D1 + alpha * alpha * (D2 - D1), -1.0, 1.0)
        Fc, _ = coef_pass(Fc)
        if sqrt(float(((F0 - Fc) ** 2).sum()) / len(Fc)) < tolem:
            break
    p0, p1, p2 = priors(Fc, True)
    la = np.where(c == 3, np.log(p0), np.where(c == 2, np.log(p1), np.where(c == 0, np.log(p2),
                                                                                    np.log(0.333333 * (p0 + p1 + p2)))))
    q = 1.0 - PI
    ln = np.where(c == 3, np.log(q * q), np.where(c == 2, np.log(2.0 * PI * q), np.where(
        c == 0, np.log(PI * PI), np.log(0.333333 * (q * q + 2.0 * PI * q + PI * PI)))))
    T = 2.0 * (la.sum(axis=0) - ln.sum(axis=0))
    p = np.array([1.0 if (t != t or t < 0) else erfc(sqrt(t / 2.0)) for t in T])
    p = np.array([1.0 if isnan(x) else x for x in p])
    return Fc, T, p, it + 1


def project_plink(codes: np.ndarray, bim_lines, mbim_lines, S, V, method: int, standardize: bool = True):
    """run_projection (src/Projection.cpp:147-208) on PLINK targets with the
    matching of FileBed's constructor (src/FilePlink.hpp:31-93): target .bim
    (split at tabs and spaces) against the .mbim (7 tab fields) by ID, then
    chromosome and position; alleles equal (F = freq) or swapped (F = 1 - freq,
    the loading row negated).  g = BED2GENO - F (missing 0, masked), divided
    by sqrt(2F(1-F)) when standardising and above 1e-9.  Method 1: U = G V S^-1;
    method 2: per sample the least-squares fit of V S on its observed SNPs
    (numpy lstsq: the reference uses ColPivHouseholderQR).  codes: M x N as
    decode_codes returns them.  Returns (U, W) with W the written loadings."""
    ref = {}
    for idx, line in enumerate(mbim_lines):
        t = line.split("\t")
        ref.setdefault(t[1], (idx, t[0], t[3], t[4], t[5], float(t[6])))
    rows, Fs, Ws = [], [], []
    S = np.asarray(S, dtype=np.float64)
    for j, line in enumerate(bim_lines):
        t = [x for x in __import__("re").split(r"[\t ]", line)]
        r = ref.get(t[1])
        if r is None or t[0] != r[1] or t[3] != r[2]:
            continue
        if t[4] == r[3] and t[5] == r[4]:
            f, sg = r[5], 1.0
        elif t[4] == r[4] and t[5] == r[3]:
            f, sg = 1.0 - r[5], -1.0
        else:
            continue
        rows.append(j)
        Fs.append(f)
        v = np.asarray(V[r[0]], dtype=np.float64) * sg
        Ws.append(v / S if method == 1 else v * S)
    c = np.asarray(codes)[rows]                     # matched SNPs x samples
    F = np.array(Fs)
    W = np.array(Ws)
    val = np.select([c == 0, c == 2, c == 3], [1.0, 0.5, 0.0], 0.0) - F[:, None]
    miss = c == 1
    val[miss] = 0.0
    if standardize:
        sd = np.sqrt(2.0 * F * (1.0 - F))
        ok = sd > 1e-9
        val[ok] /= sd[ok, None]
    n = c.shape[1]
    if method == 1:
        return val.T @ W, W
    U = np.empty((n, W.shape[1]))
    for i in range(n):
        o = ~miss[:, i]
        U[i] = np.linalg.lstsq(W[o], val[o, i], rcond=None)[0]
    return U, W



# ----------------------------------------------------------------------------
# BGEN input (FileBgen::read_all, src/FileBgen.cpp:11-107) through the
# reference's vendored bgen library: the container (external/bgen/header.cpp,
# variant.cpp), the genotype blocks (genotypes.cpp) and minor_certain
# (utils.cpp:46-51)
# ----------------------------------------------------------------------------
def bgen_lut8() -> np.ndarray:
    """genotypes.cpp's 511-entry table: i / 255 printed to 7 decimals, read as
    float (the literal of the table)."""
    return np.array([np.float32(float(f"{i / 255.0:.7f}")) for i in range(511)], dtype=np.float32)


def _zstd_decompress(data: bytes, size: int) -> bytes:
    import ctypes

    z = ctypes.CDLL("libzstd.so.1")
    z.ZSTD_decompress.restype = ctypes.c_size_t
    z.ZSTD_isError.restype = ctypes.c_uint
    dst = ctypes.create_string_buffer(max(size, 1))
    n = z.ZSTD_decompress(dst, ctypes.c_size_t(size), data, ctypes.c_size_t(len(data)))
    if z.ZSTD_isError(ctypes.c_size_t(n)) or n != size:
        raise ValueError("zstd decompression gave data of wrong length")
    return dst.raw[:size]


def bgen_blocks(path):
    """Header.cpp / Variant.cpp: (nsamples, layout, [(nalleles, decompressed
    genotype block)] in file order); the first variant sits at offset + 4."""
    import struct
    import zlib

    b = open(path, "rb").read()
    first, hlen, nvar, nsmp = struct.unpack_from("<IIII", b, 0)
    flags = struct.unpack_from("<I", b, max(20, hlen))[0]
    comp, layout = flags & 3, (flags >> 2) & 15
    off = first + 4
    out = []
    for _ in range(nvar):
        p = off
        if layout == 1:
            assert struct.unpack_from("<I", b, p)[0] == nsmp
            p += 4
        for _ in range(3):  # variant id, rsid, chromosome
            p += 2 + struct.unpack_from("<H", b, p)[0]
        p += 4
        K = 2
        if layout == 2:
            K = struct.unpack_from("<H", b, p)[0]
            p += 2
        for _ in range(K):
            p += 4 + struct.unpack_from("<I", b, p)[0]
        if layout == 1 and comp == 0:
            length = 6 * nsmp
        else:
            length = struct.unpack_from("<I", b, p)[0]
            p += 4
        blk = b[p:p + length]
        off = p + length
        if comp != 0 and layout == 2:
            dlen = struct.unpack_from("<I", blk, 0)[0]
            blk = blk[4:]
        else:
            dlen = 6 * nsmp if layout == 1 else len(blk)
        if comp == 1:
            blk = zlib.decompress(blk)
        elif comp == 2:
            blk = _zstd_decompress(blk, dlen)
        assert len(blk) == dlen
        out.append((K, blk))
    return nsmp, layout, out


def _bits(q: np.ndarray, start_bits: np.ndarray, B: int) -> np.ndarray:
    """the B-bit fields at the given bit offsets of the little-endian stream q"""
    qq = np.concatenate([q, np.zeros(8, dtype=np.uint8)]).astype(np.uint64)
    byte = (start_bits >> 3).astype(np.int64)
    w = np.zeros(len(start_bits), dtype=np.uint64)
    for i in range(5):
        w |= qq[byte + i] << np.uint64(8 * i)
    return ((w >> (start_bits & 7).astype(np.uint64)) & np.uint64((1 << B) - 1)).astype(np.uint32)


def _bgen_header2(blk: bytes, N: int):
    q = np.frombuffer(blk, dtype=np.uint8)
    pmin, pmax = int(q[6]), int(q[7])
    ploidy = q[8:8 + N]
    phased, B = int(q[8 + N]), int(q[9 + N])
    return q, pmin, pmax, ploidy, phased, B, q[10 + N:]


def bgen_ref_dosage(blk: bytes, layout: int, N: int):
    """Genotypes::get_allele_dosage's first step: the first allele's dosage
    (float32) of every sample and the missing mask.  ref_dosage_fast
    (constant diploid unphased 8-bit) takes its AVX2 loop, (2 hom + het) x
    (1/255f), for samples below 16 ceil((N - 32) / 16) and the lut8 table
    for the rest; ref_dosage_slow reads two B-bit values per sample,
    ((hom ploidy) + het (ploidy / 2)) x 1/(2^B - 1) in float (layout 1: three
    16-bit values, x 1/32768, missing when all three are 0)."""
    if layout == 1:
        v = np.frombuffer(blk, dtype="<u2").reshape(N, 3).astype(np.uint32)
        miss = (v[:, 0] == 0) & (v[:, 1] == 0) & (v[:, 2] == 0)
        d = (v[:, 0] * np.uint32(2) + v[:, 1]).astype(np.float32) * np.float32(1.0 / 32768)
        return d, miss
    q, pmin, pmax, ploidy, phased, B, pr = _bgen_header2(blk, N)
    miss = (ploidy & 0x80) != 0
    const = pmin == pmax
    if const and not phased and pmax == 2 and B == 8:
        i = 2 * pr[0:2 * N:2].astype(np.uint32) + pr[1:2 * N:2].astype(np.uint32)
        vec_end = 16 * (-(-(N - 32) // 16)) if N > 32 else 0
        d = bgen_lut8()[i]
        d[:vec_end] = i[:vec_end].astype(np.float32) * np.float32(np.float32(1.0) / np.float32(255.0))
        return d, miss
    cp = np.full(N, pmax, dtype=np.uint32) if const else (ploidy & 63).astype(np.uint32)
    n = np.arange(N, dtype=np.uint64)
    hom = _bits(pr, np.uint64(2 * B) * n, B)
    het = _bits(pr, np.uint64(2 * B) * n + np.uint64(B), B)
    maxval = np.uint32(int(2.0 ** B - 1.0))
    factor = np.float32(np.float32(1.0) / np.float32(maxval))
    with np.errstate(over="ignore"):
        d = (hom * cp + het * (cp // np.uint32(2))).astype(np.float32) * factor
    return d, miss


def bgen_minor_index(dose: np.ndarray) -> int:
    """Genotypes::find_minor_allele: residue classes n = idx2 (mod inc),
    inc = max(N / 100, 1), summed in order into a running double total;
    freq = total / (100 (idx2 + 1) 2) after each class (the denominator
    counts 100 samples per class whatever N is, so freq exceeds 1 for
    N < 200 at high dosage, where the interval is NaN and the test passes),
    stop at the first class where minor_certain(freq, 100 (idx2 + 1), 5)
    holds; minor index 0 when freq <= 0.5."""
    N = len(dose)
    inc = max(N // 100, 1)
    total, freq = 0.0, 0.0
    for idx2 in range(inc):
        for n in range(idx2, N, inc):
            total += float(dose[n])
        freq = total / (100 * (idx2 + 1) * 2)
        nchk = 100 * (idx2 + 1)
        v = (freq * (1 - freq)) / nchk
        delta = 5.0 * math.sqrt(v) if v >= 0 else math.nan  # std::sqrt of a negative: NaN
        if not ((freq - delta < 0.5) and (freq + delta > 0.5)):
            break
    return 0 if freq <= 0.5 else 1


def bgen_read(path, maf: float = 0.0, standardize: bool = True, ploidy: int = 2):
    """FileBgen::read_all without --pcangsd: per variant the minor allele's
    dosage (swap_allele_dosage: 2 - d in float; missing NaN), F = sum of d / 2
    over the non-missing samples / their count, kept when F > maf (maf above
    0.5 read as 1 - maf); X (N x kept) = d / 2 - F, missing 0, then
    Data::standardize_E when standardize.  Returns (X, F of every variant,
    minor index of every variant, kept indices).  The reference files the
    centred column of variant j under j instead of its kept index k
    (src/FileBgen.cpp:51); this restatement files it under k."""
    if maf > 0.5:
        maf = 1.0 - maf
    N, layout, blocks = bgen_blocks(path)
    cols, Fs, minors, keep = [], [], [], []
    for j, (_, blk) in enumerate(blocks):
        d, miss = bgen_ref_dosage(blk, layout, N)
        mi = bgen_minor_index(d)
        if mi != 0:
            d = np.float32(2.0) - d
        gs = float(np.sum(d[~miss].astype(np.float64) / 2.0))
        gc = int(np.count_nonzero(~miss))
        af = 0.0 if gc == 0 else gs / gc
        Fs.append(af)
        minors.append(mi)
        if af > maf:
            x = d.astype(np.float64) / 2.0 - af
            x[miss] = 0.0
            cols.append(x)
            keep.append(j)
    F = np.array(Fs)
    X = np.array(cols).T.copy()
    if standardize:
        X = standardize_cols(X, F[keep], ploidy)
    return X, F, np.array(minors, dtype=np.uint8), np.array(keep, dtype=np.int64)


def standardize_cols(X: np.ndarray, F: np.ndarray, ploidy: int = 2) -> np.ndarray:
    """Data::standardize_E on X (N x M) with F (M)."""
    return standardize(X, F, ploidy)


def bgen_probs(path) -> np.ndarray:
    """FileBgen::read_all with --pcangsd: P (M, N, 2) float32, the first two
    genotype probabilities of every sample (Genotypes::probabilities: lut8 of
    the two bytes for constant diploid unphased 8-bit data, else B-bit values x
    (float)(1 / ((float)2^B - 1)); layout 1: 16-bit values x 1/32768 with
    missing when all three are 0); missing samples NaN."""
    N, layout, blocks = bgen_blocks(path)
    P = np.empty((len(blocks), N, 2), dtype=np.float32)
    lut = bgen_lut8()
    for j, (_, blk) in enumerate(blocks):
        if layout == 1:
            v = np.frombuffer(blk, dtype="<u2").reshape(N, 3)
            p = v.astype(np.float32) * np.float32(1.0 / 32768)
            miss = (v == 0).all(axis=1)
            P[j] = p[:, :2]
        else:
            q, pmin, pmax, ploidy, phased, B, pr = _bgen_header2(blk, N)
            assert pmin == pmax == 2 and not phased
            miss = (ploidy & 0x80) != 0
            if B == 8:
                P[j, :, 0] = lut[pr[0:2 * N:2]]
                P[j, :, 1] = lut[pr[1:2 * N:2]]
            else:
                fac = np.float32(1.0 / float(np.float32(2.0 ** B) - np.float32(1)))
                n = np.arange(N, dtype=np.uint64)
                P[j, :, 0] = _bits(pr, np.uint64(2 * B) * n, B).astype(np.float32) * fac
                P[j, :, 1] = _bits(pr, np.uint64(2 * B) * n + np.uint64(B), B).astype(np.float32) * fac
        P[j, miss] = np.nan
    return P


# ----------------------------------------------------------------------------
# Projection of BEAGLE genotype likelihoods (run_projection_beagle_gl,
# src/Projection.cpp:32-137, with FileBeagle::read_all's matching,
# src/FileBeagle.cpp:12-118, and solve_projection_scores, :11-30)
# ----------------------------------------------------------------------------
_BEAGLE_ALLELE = {"0": "A", "1": "C", "2": "G", "3": "T"}


def beagle_mbim_match(sites, mbim_lines):
    """sites: (marker, allele1, allele2) per BEAGLE line, alleles 0-3 read as
    A, C, G, T.  A site finds its .mbim line by ID (column 2), else by
    chromosome_position (columns 1 and 4), the first line of a key winning;
    alleles equal give F = the .mbim frequency, swapped give F = 1 - it and a
    flip; anything else is unmatched.  Returns (ref_idx, -1 unmatched; flip;
    F) per site."""
    by_id, by_pos = {}, {}
    for idx, line in enumerate(mbim_lines):
        t = line.split("\t")
        assert len(t) == 7
        site = (idx, t[4], t[5], float(t[6]))
        by_id.setdefault(t[1], site)
        by_pos.setdefault(t[0] + "_" + t[3], site)
    ref_idx = np.full(len(sites), -1, dtype=np.int64)
    flip = np.zeros(len(sites), dtype=bool)
    F = np.zeros(len(sites))
    for j, (marker, a1, a2) in enumerate(sites):
        a1, a2 = _BEAGLE_ALLELE.get(a1, a1), _BEAGLE_ALLELE.get(a2, a2)
        r = by_id.get(marker)
        if r is None:
            r = by_pos.get(marker)
        if r is None:
            continue
        if a1 == r[1] and a2 == r[2]:
            ref_idx[j], F[j] = r[0], r[3]
        elif a1 == r[2] and a2 == r[1]:
            ref_idx[j], F[j], flip[j] = r[0], 1.0 - r[3], True
    return ref_idx, flip, F


def project_beagle_gl(P0, P1, ref_idx, flip, F, S, V, standardize: bool = True, maxiter: int = 100,
                      tolem: float = 1e-5):
    """P0, P1 (N x M): the first two likelihoods of every site; ref_idx, flip,
    F from beagle_mbim_match; S (k), V (reference SNPs x k).  On the matched
    sites L = +-V rows, VS = L S.  E step: pt = F (initial) or
    clamp((U_i S) L_j (times sqrt(2F(1-F)) when standardising) + F), both
    clamped to [1e-4, 1 - 1e-4]; posterior p11 = gl11 pt^2, p12 = gl12 2pt(1-pt),
    p22 = (1 - gl11 - gl12)(1-pt)^2; a non-finite or non-positive sum marks
    the entry missing (0); else g = (2 p11 + p12) / (2 sum) - F, divided by
    sqrt(2F(1-F)) when standardising and that exceeds 1e-9 (in the iterations
    such sites become 0, in the initial step they stay).  Each sample's scores
    are the least-squares fit of g on VS over its observed sites; after each
    iteration a column whose dot product with the previous is negative is
    negated, and the loop stops when |U - Uprev|_F / |Uprev|_F < tolem.
    Returns (U, VS, iterations, diffs)."""
    m_ = ref_idx >= 0
    L = np.asarray(V, dtype=np.float64)[ref_idx[m_]] * np.where(flip[m_], -1.0, 1.0)[:, None]
    S = np.asarray(S, dtype=np.float64)
    VS = L * S[None, :]
    Fm = np.asarray(F, dtype=np.float64)[m_]
    g11 = np.asarray(P0, dtype=np.float64)[:, m_]
    g12 = np.asarray(P1, dtype=np.float64)[:, m_]
    g22 = 1.0 - g11 - g12
    norm = np.sqrt(2.0 * Fm * (1.0 - Fm))
    n, k = g11.shape[0], S.size

    def estep(U, init):
        if init:
            pt = np.broadcast_to(np.minimum(np.maximum(Fm, 1e-4), 1.0 - 1e-4), g11.shape)
        else:
            z = (U * S[None, :]) @ L.T
            c = z * norm[None, :] if standardize else z
            pt = np.minimum(np.maximum(c + Fm[None, :], 1e-4), 1.0 - 1e-4)
        with np.errstate(invalid="ignore"):
            p11 = g11 * pt * pt
            p12 = g12 * 2.0 * pt * (1.0 - pt)
            p22 = g22 * (1.0 - pt) * (1.0 - pt)
            ps = p11 + p12 + p22
            miss = ~np.isfinite(ps) | (ps <= 0.0)
            G = (2.0 * p11 + p12) / (2.0 * ps) - Fm[None, :]
        G[miss] = 0.0
        if standardize:
            ok = norm > VAR_TOL
            G[:, ok] /= norm[None, ok]
            if not init:
                G[:, ~ok] = 0.0
            G[miss] = 0.0
        return G, miss

    def solve(G, miss):
        U = np.empty((n, k))
        for i in range(n):
            o = ~miss[i]
            U[i] = np.linalg.lstsq(VS[o], G[i, o], rcond=None)[0]
        return U

    G, miss = estep(None, True)
    U = solve(G, miss)
    diffs, it = [], 0
    for it in range(1, maxiter + 1):
        Up = U.copy()
        G, miss = estep(U, False)
        U = solve(G, miss)
        neg = np.sum(U * Up, axis=0) < 0.0
        U[:, neg] *= -1.0
        den = np.linalg.norm(Up)
        if den < 1e-12:
            den = 1.0
        diffs.append(float(np.linalg.norm(U - Up) / den))
        if diffs[-1] < tolem:
            break
    return U, VS, it, diffs
