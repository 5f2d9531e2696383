"""Host-side interface of the MI355X Jacquard fit, mirroring the reference's
batch boundary wrapper::calculate_relatedness_coefficients_gl (wrapper.rs:89-207):
a cohort of hard calls in, one Output record per pair out (coefficients.rs:143-151).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from ._lib import OUTPUT_DTYPE, Tune, check, ptr
from .synth import pack2

KINSHIP_VEC = np.array([1.0, 0.0, 0.5, 0.0, 0.5, 0.0, 0.5, 0.25, 0.0])  # cli.rs:150


def tune_default(iterations: int = 100) -> Tune:
    """sqp.rs:31-46 with sqp_max_iter = iterations and qp_max_iter = 10 (coefficients.rs:73)."""
    t = Tune()
    check(_lib.load().kestrel_tune_default(ctypes.byref(t), iterations))
    return t


def _maf_arg(maf: float | None) -> float:
    """None selects the reference's default 1/(2N) (KESTREL_MAF_DEFAULT);
    a given value must lie within 0 to 1 (wrapper.rs:101-106)."""
    if maf is None:
        return _lib.MAF_DEFAULT
    maf = float(maf)
    if not 0.0 <= maf <= 1.0:
        raise ValueError("Minimum allele frequency must be within 0 to 1")
    return maf


class Cohort:
    """Device-resident cohort of 2-bit hard calls (kestrel_cohort_create)."""

    def __init__(self, codes: np.ndarray | None = None, *, packed: np.ndarray | None = None,
                 n_sites: int | None = None, af: np.ndarray | None = None, maf: float | None = None,
                 af_iterations: int = 100, segment_sites: int | None = None, precision: str = "fp64"):
        """precision: "fp64" (the reference's arithmetic) or "mixed" (per-site
        terms in FP32, KESTREL_MIXED)."""
        L = _lib.load()
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if packed is None:
            codes = np.ascontiguousarray(codes, np.uint8)
            n_sites = codes.shape[1]
            packed = pack2(codes)
        packed = np.ascontiguousarray(packed, np.uint8)
        self.n_samples = packed.shape[0]
        self.n_sites = int(n_sites)
        afp = None
        if af is not None:
            afp = np.ascontiguousarray(af, np.float64).reshape(self.n_sites, 2)
        h = ctypes.c_void_p()
        check(L.kestrel_cohort_create(ctypes.byref(h), ptr(packed), self.n_samples, self.n_sites, packed.shape[1],
                                      ptr(afp), _maf_arg(maf), af_iterations,
                                      _lib.FP64 if precision == "fp64" else _lib.MIXED))
        self._h = h
        if segment_sites is not None:
            self.set_option("segment_sites", segment_sites)

    @classmethod
    def gt(cls, codes: np.ndarray | None = None, *, packed: np.ndarray | None = None, n_sites: int | None = None,
           af: np.ndarray | None = None) -> "Cohort":
        """Cohort for the CLS / Brier estimator of -f GT (kestrel_cohort_create_gt):
        sites with a missing call are dropped, frequencies counted unless given."""
        L = _lib.load()
        self = cls.__new__(cls)
        if packed is None:
            codes = np.ascontiguousarray(codes, np.uint8)
            n_sites = codes.shape[1]
            packed = pack2(codes)
        packed = np.ascontiguousarray(packed, np.uint8)
        self.n_samples = packed.shape[0]
        self.n_sites = int(n_sites)
        afp = None if af is None else np.ascontiguousarray(af, np.float64).reshape(self.n_sites, 2)
        h = ctypes.c_void_p()
        check(L.kestrel_cohort_create_gt(ctypes.byref(h), ptr(packed), self.n_samples, self.n_sites,
                                         packed.shape[1], ptr(afp)))
        self._h = h
        return self

    @classmethod
    def gt_multi(cls, gt2: np.ndarray | None = None, gt3: np.ndarray | None = None, gt4: np.ndarray | None = None,
                 *, af=None) -> "Cohort":
        """Cohort for the CLS / Brier estimator of -f GT with sites of 2, 3 and 4
        alleles (kestrel_cohort_create_gt_multi): gtA is (N, V_A) genotype
        indices in VCF order (j(j+1)/2 + i), 255 = a missing call (the site is
        dropped); af = (af2, af3, af4) with afA (V_A, A), or None (counted)."""
        L = _lib.load()
        self = cls.__new__(cls)
        arrs, counts, n = [], [], None
        for A, g in ((2, gt2), (3, gt3), (4, gt4)):
            if g is None:
                arrs.append(None)
                counts.append(0)
                continue
            g = np.ascontiguousarray(g, np.uint8)
            if g.ndim != 2:
                raise ValueError(f"gt{A} must have shape (samples, sites)")
            if n is not None and g.shape[0] != n:
                raise ValueError("buckets must have the same samples")
            n = g.shape[0]
            arrs.append(g)
            counts.append(g.shape[1])
        if n is None:
            raise ValueError("no genotypes given")
        packed = None
        if arrs[0] is not None and counts[0]:
            g2 = arrs[0]
            if ((g2 > 2) & (g2 != 255)).any():
                raise ValueError("gt2 holds genotype indices 0, 1, 2 or 255")
            packed = pack2(np.where(g2 == 255, 3, g2).astype(np.uint8))
        afs = [None, None, None]
        if af is not None:
            afs = [None if v == 0 or a is None else np.ascontiguousarray(a, np.float64).reshape(v, A)
                   for A, v, a in zip((2, 3, 4), counts, af)]
        self.n_samples, self.n_sites = int(n), int(sum(counts))
        self._bucket_sites = counts
        h = ctypes.c_void_p()
        check(L.kestrel_cohort_create_gt_multi(ctypes.byref(h), self.n_samples, ptr(packed), counts[0],
                                               0 if packed is None else packed.shape[1], ptr(arrs[1]), counts[1],
                                               ptr(arrs[2]), counts[2], ptr(afs[0]), ptr(afs[1]), ptr(afs[2])))
        self._h = h
        return self

    @classmethod
    def gl(cls, gl: np.ndarray, *, af: np.ndarray | None = None, maf: float | None = None,
           af_iterations: int = 100, segment_sites: int | None = None, precision: str = "fp64") -> "Cohort":
        """Cohort of soft genotype likelihoods (kestrel_cohort_create_gl): gl is
        (V, N, 3) site-major linear-scale likelihoods of 0/0, 0/1, 1/1, a NaN
        marking a missing call.  precision "mixed": option "precision"
        KESTREL_MIXED."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        L = _lib.load()
        self = cls.__new__(cls)
        g = np.ascontiguousarray(gl, np.float64)
        if g.ndim != 3 or g.shape[2] != 3:
            raise ValueError("gl must have shape (sites, samples, 3)")
        self.n_sites, self.n_samples = int(g.shape[0]), int(g.shape[1])
        afp = None if af is None else np.ascontiguousarray(af, np.float64).reshape(self.n_sites, 2)
        h = ctypes.c_void_p()
        check(L.kestrel_cohort_create_gl(ctypes.byref(h), ptr(g), self.n_samples, self.n_sites, ptr(afp),
                                         _maf_arg(maf), af_iterations))
        self._h = h
        if segment_sites is not None:
            self.set_option("segment_sites", segment_sites)
        if precision == "mixed":
            self.set_option("precision", _lib.MIXED)
        return self

    @classmethod
    def gl_multi(cls, gl2: np.ndarray | None = None, gl3: np.ndarray | None = None, gl4: np.ndarray | None = None,
                 *, maf: float | None = None, af_iterations: int = 100, segment_sites: int | None = None,
                 precision: str = "fp64") -> "Cohort":
        """Soft likelihoods of sites with 2, 3 and 4 alleles
        (kestrel_cohort_create_gl_multi): glA is (V_A, N, G) site-major,
        G = 3, 6, 10 in VCF genotype order.  precision "mixed": option
        "precision" KESTREL_MIXED for every part."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        L = _lib.load()
        self = cls.__new__(cls)
        arrs, counts, n = [], [], None
        for A, g in ((2, gl2), (3, gl3), (4, gl4)):
            G = A * (A + 1) // 2
            if g is None:
                arrs.append(None)
                counts.append(0)
                continue
            g = np.ascontiguousarray(g, np.float64)
            if g.ndim != 3 or g.shape[2] != G:
                raise ValueError(f"gl{A} must have shape (sites, samples, {G})")
            if n is not None and g.shape[1] != n:
                raise ValueError("buckets must have the same samples")
            n = g.shape[1]
            arrs.append(g)
            counts.append(g.shape[0])
        if n is None:
            raise ValueError("no likelihoods given")
        self.n_samples, self.n_sites = int(n), int(sum(counts))
        self._bucket_sites = counts
        h = ctypes.c_void_p()
        check(L.kestrel_cohort_create_gl_multi(ctypes.byref(h), self.n_samples, ptr(arrs[0]), counts[0],
                                               ptr(arrs[1]), counts[1], ptr(arrs[2]), counts[2],
                                               _maf_arg(maf), af_iterations))
        self._h = h
        if segment_sites is not None:
            self.set_option("segment_sites", segment_sites)
        if precision == "mixed":
            self.set_option("precision", _lib.MIXED)
        return self

    def multi_info(self):
        """(sites with 2, 3, 4 alleles after zero-frequency removal, the
        fitted frequencies of the input buckets {A: (V_A, A)}); for a gt_multi
        cohort the kept sites per bucket and the frequencies of the input
        sites (NaN rows for dropped sites)."""
        parts = np.zeros(3, np.uint32)
        afs = [np.zeros((v, A)) for A, v in zip((2, 3, 4), self._bucket_sites)]
        check(_lib.load().kestrel_cohort_multi_info(self._h, ptr(parts), ptr(afs[0]), ptr(afs[1]), ptr(afs[2])))
        return tuple(int(v) for v in parts), {2: afs[0], 3: afs[1], 4: afs[2]}

    def fit_pairs_gt(self, pairs: np.ndarray | None = None, n_select: int | None = None, iterations: int = 100):
        """CLS fit of pairs (coefficients.rs:166-278); returns (outputs, Q)."""
        if pairs is None:
            n_select = self.n_samples if n_select is None else n_select
            n = n_select * (n_select + 1) // 2
            pp = None
        else:
            pp = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
            n = pp.shape[0]
            n_select = 0
        out = np.zeros(n, OUTPUT_DTYPE)
        q = np.zeros((9, 9))
        check(_lib.load().kestrel_fit_pairs_gt(self._h, iterations, ptr(pp), n, n_select, ptr(out), ptr(q)))
        return out, q

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().kestrel_cohort_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, key: str, value: int):
        check(_lib.load().kestrel_set_option(self._h, key.encode(), int(value)))

    @property
    def n_kept(self) -> int:
        v = ctypes.c_uint32()
        check(_lib.load().kestrel_cohort_n_kept(self._h, ctypes.byref(v)))
        return v.value

    def allele_freqs(self):
        af = np.zeros((self.n_sites, 2))
        kept = np.zeros(self.n_sites, np.uint8)
        check(_lib.load().kestrel_cohort_allele_freqs(self._h, ptr(af), ptr(kept)))
        return af, kept.astype(bool)

    def fit_pairs(self, pairs: np.ndarray | None = None, n_select: int | None = None,
                  tune: Tune | None = None, iterations: int = 100) -> np.ndarray:
        """Pairs (M, 2) or None for the upper triangle incl. the diagonal of
        the first n_select samples (coefficients.rs:57, 87)."""
        t = tune or tune_default(iterations)
        if pairs is None:
            n_select = self.n_samples if n_select is None else n_select
            n = n_select * (n_select + 1) // 2
            pp = None
        else:
            pp = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
            n = pp.shape[0]
            n_select = 0
        out = np.zeros(n, OUTPUT_DTYPE)
        check(_lib.load().kestrel_fit_pairs(self._h, ctypes.byref(t), ptr(pp), n, n_select, ptr(out)))
        return out

    def autotune(self, pairs: np.ndarray | None = None, sample: int = 65536, seed: int = 0,
                 job_pairs: int | None = None, tune: Tune | None = None, iterations: int = 100) -> dict:
        """Launch-shape autotuner (kestrel_autotune): fits a sample of pairs
        (default: `sample` random off-diagonal pairs) with each candidate
        slots / streams / batch shape, predicts the time of a job of
        job_pairs pairs (default: every pair of the cohort) from each, keeps
        the fastest on this cohort and returns it with the predicted
        throughputs.  Results of later fits do not depend on the shape."""
        t = tune or tune_default(iterations)
        if pairs is None:
            rng = np.random.default_rng(seed)
            n = self.n_samples
            if n < 2:
                raise ValueError("tuning needs at least two samples")
            m = sample  # drawn with replacement: a timing sample
            x = rng.integers(0, n, 2 * m + 16)
            y = rng.integers(0, n, 2 * m + 16)
            keep = x != y
            pairs = np.stack([np.minimum(x, y), np.maximum(x, y)], 1)[keep][:m]
        pp = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        if job_pairs is None:
            job_pairs = self.n_samples * (self.n_samples + 1) // 2
        r = _lib.AutotuneResult()
        check(_lib.load().kestrel_autotune(self._h, ctypes.byref(t), ptr(pp), pp.shape[0], int(job_pairs),
                                           ctypes.byref(r)))
        return {"slots": r.slots, "streams": r.streams, "batch": r.batch, "pairs_per_s": r.pairs_per_s,
                "default_pairs_per_s": r.default_pairs_per_s, "trials": r.trials, "sample_pairs": int(pp.shape[0])}

    def evaluate(self, pairs: np.ndarray, points: np.ndarray, eps: float = 1e-8):
        """Objective, gradient and Hessian at points (M, 9) for pairs (M, 2)."""
        pp = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 9)
        n = pp.shape[0]
        f, g, h = np.zeros(n), np.zeros((n, 9)), np.zeros((n, 9, 9))
        ns = np.zeros(n, np.uint32)
        check(_lib.load().kestrel_evaluate(self._h, ptr(pp), ptr(pts), n, eps, ptr(f), ptr(g), ptr(h), ptr(ns)))
        return f, g, h, ns

    # ---- serving (kestrel_serve_start / kestrel_submit_queries / kestrel_wait)

    def serve_start(self, max_query_rows: int = 1024, max_pairs: int = 1 << 24, tune: Tune | None = None,
                    iterations: int = 100):
        """Start the serving engine: the cohort answers requests of query
        samples while others run (continuous batching)."""
        self._serve_tune = tune or tune_default(iterations)
        self._requests = {}
        check(_lib.load().kestrel_serve_start(self._h, ctypes.byref(self._serve_tune), int(max_query_rows),
                                              int(max_pairs)))

    def submit(self, query_codes: np.ndarray | None = None, subset: np.ndarray | None = None, *,
               query_packed: np.ndarray | None = None) -> int:
        """Queue query samples (codes (Q, n_sites) or 2-bit rows) against the
        cohort subset (indices) or the whole cohort; returns the request id.
        The result of wait(id) is (Q * len(subset)) records, x = query index,
        y = cohort index, in query-major order."""
        if query_packed is None:
            query_packed = pack2(np.ascontiguousarray(query_codes, np.uint8))
        qp = np.ascontiguousarray(query_packed, np.uint8)
        sub = None if subset is None else np.ascontiguousarray(subset, np.uint32)
        n_sub = self.n_samples if sub is None else sub.shape[0]
        out = np.zeros(qp.shape[0] * n_sub, OUTPUT_DTYPE)
        rid = ctypes.c_uint64()
        check(_lib.load().kestrel_submit_queries(self._h, ptr(qp), qp.shape[0], qp.shape[1], ptr(sub),
                                                 0 if sub is None else sub.shape[0], ptr(out), ctypes.byref(rid)))
        self._requests[rid.value] = out  # the engine writes into out until the request is done
        return rid.value

    def poll(self, request_id: int):
        """(done, latency_ms, outputs or None)."""
        done, lat = ctypes.c_int(), ctypes.c_double()
        try:
            check(_lib.load().kestrel_poll(self._h, request_id, ctypes.byref(done), ctypes.byref(lat)))
        except Exception:
            self._requests.pop(request_id, None)
            raise
        if not done.value:
            return False, None, None
        return True, lat.value, self._requests.pop(request_id)

    def wait(self, request_id: int):
        """Block until the request is done; returns (outputs, latency_ms)."""
        lat = ctypes.c_double()
        try:
            check(_lib.load().kestrel_wait(self._h, request_id, ctypes.byref(lat)))
        except Exception:
            self._requests.pop(request_id, None)
            raise
        return self._requests.pop(request_id), lat.value

    def serve_stats(self) -> dict:
        st = _lib.ServeStats()
        check(_lib.load().kestrel_serve_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "reserved"}

    def serve_stop(self):
        check(_lib.load().kestrel_serve_stop(self._h))

    def last_fit_stats(self):
        ms, sw, wall = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
        check(_lib.load().kestrel_last_fit_stats(self._h, ctypes.byref(ms), ctypes.byref(sw), ctypes.byref(wall)))
        busy, groups = ctypes.c_double(), ctypes.c_uint32()
        check(_lib.load().kestrel_last_fit_busy(self._h, ctypes.byref(busy), ctypes.byref(groups)))
        rms, rpairs = ctypes.c_double(), ctypes.c_uint64()
        check(_lib.load().kestrel_last_fit_recondition(self._h, ctypes.byref(rms), ctypes.byref(rpairs)))
        form = ctypes.c_int()
        check(_lib.load().kestrel_last_fit_form(self._h, ctypes.byref(form)))
        return {"sweep_ms": ms.value, "sweeps": sw.value, "wall_ms": wall.value, "busy_ms": busy.value,
                "groups": groups.value, "recondition_ms": rms.value, "recondition_pairs": rpairs.value,
                "form": {0: "table", 1: "class", 2: "lanes"}[form.value]}


def segment_bounds(n_kept: int, segment_sites: int = 4096, chunk: int = 32, pad: int = 256) -> np.ndarray:
    """Site segment bounds the device uses for a cohort with n_kept sites (the
    summation order of the sweep: sequential within a segment, segments folded
    left to right)."""
    v_pad = max(pad, (n_kept + pad - 1) // pad * pad)
    chunks = v_pad // chunk
    target = max(chunk, segment_sites) // chunk
    nseg = max(1, (chunks + target - 1) // target)
    per = (chunks + nseg - 1) // nseg
    b = [min(chunks, k * per) * chunk for k in range(nseg + 1)]
    b[-1] = v_pad
    return np.array(b, np.uint32)


def kinship(jacquard: np.ndarray) -> np.ndarray:
    """cli.rs:150,166: kinship = dot(delta, [1, 0, .5, 0, .5, 0, .5, .25, 0]) as an fma chain."""
    j = np.asarray(jacquard, np.float64)
    s = np.zeros(j.shape[:-1])
    for k in range(9):
        s = s + j[..., k] * KINSHIP_VEC[k]  # products with 0, .25, .5, 1 are exact
    return s


def build_arrays(outputs: np.ndarray, n: int):
    """lib.rs:120-161: (N, N, 9) Jacquard array with delta3<->delta5 and
    delta4<->delta6 swapped for (y, x), (N, N) kinship, (N,) inbreeding
    2*theta_xx - 1; entries without a fitted pair stay 0 as in the reference."""
    jac = np.zeros((n, n, 9))
    kin = np.zeros((n, n))
    inb = np.zeros(n)
    swap = [0, 1, 4, 5, 2, 3, 6, 7, 8]
    k = kinship(outputs["jacquard"])
    for o, kk in zip(outputs, k):
        x, y = int(o["x"]), int(o["y"])
        jac[x, y] = o["jacquard"]
        kin[x, y] = kk
        if x != y:
            jac[y, x] = o["jacquard"][swap]
            kin[y, x] = kk
        else:
            inb[x] = 2.0 * kk - 1.0
    return jac, kin, inb


class Coefficients:
    """lib.rs:36-44: the arrays returned by relatedness_coefficients_gt."""

    def __init__(self, jacquard, kinship, inbreeding):
        self.jacquard = jacquard
        self.kinship = kinship
        self.inbreeding = inbreeding


def _gt_indices(genotypes) -> np.ndarray:
    """(V, N, 2) allele indices -> (N, V) unphased genotype index
    (coefficients.rs:18-38)."""
    g = np.asarray(genotypes)
    if g.ndim != 3 or g.shape[2] != 2:
        raise ValueError("genotypes must have shape (sites, samples, 2)")
    g = g.astype(np.int64)
    lo, hi = np.minimum(g[..., 0], g[..., 1]), np.maximum(g[..., 0], g[..., 1])
    if hi.size and hi.max() > 3:
        raise ValueError("alleles beyond the fourth are not supported (vcf.rs:126-160 buckets 2, 3, 4 alleles)")
    return np.ascontiguousarray((hi * (hi + 1) // 2 + lo).T.astype(np.uint8))


def relatedness_coefficients_gt(genotypes, allele_frequencies=None, threads=None, iterations: int = 100,
                                alleles=None):
    """lib.rs:46-73 (a stub in the reference): the CLS / Brier estimator of
    wrapper.rs:22-87 over every pair of samples.  genotypes: (V, N, 2) allele
    indices; alleles: (V,) allele count of each site (2, 3 or 4; default: one
    more than the largest allele index seen, at least 2), which puts the site in
    the reference's 2-, 3- or 4-allele bucket; allele_frequencies: (V, A_max)
    with the first A columns of a site used, or None (counted).  threads is
    accepted for the reference's signature and unused (the fit runs on the
    GPU)."""
    if iterations < 1:
        raise ValueError("iterations must be non-zero")
    idx = _gt_indices(genotypes)
    g = np.asarray(genotypes)
    n_v = idx.shape[1]
    if alleles is None:
        alleles = np.maximum(2, g.max(axis=(1, 2)).astype(np.int64) + 1) if n_v else np.zeros(0, np.int64)
    alleles = np.asarray(alleles, np.int64).reshape(n_v)
    if n_v and (alleles.min() < 2 or alleles.max() > 4):
        raise ValueError("sites must have 2, 3 or 4 alleles")
    if n_v and (g.max(axis=(1, 2)) >= alleles).any():
        raise ValueError("an allele index exceeds its site's allele count")
    af = None if allele_frequencies is None else np.asarray(allele_frequencies, np.float64).reshape(n_v, -1)
    if (alleles == 2).all():
        co = Cohort.gt(idx, af=None if af is None else af[:, :2])
    else:
        sel = [np.flatnonzero(alleles == A) for A in (2, 3, 4)]
        co = Cohort.gt_multi(*(idx[:, s] for s in sel),
                             af=None if af is None else [af[s, :A] for A, s in zip((2, 3, 4), sel)])
    try:
        out, _ = co.fit_pairs_gt(iterations=iterations)
    finally:
        co.close()
    return Coefficients(*build_arrays(out, idx.shape[0]))


def relatedness_coefficients_gl(genotype_likelihoods, allele_frequencies=None, threads=None,
                                iterations: int = 100, minimum_allele_frequency=None):
    """lib.rs:75-109 (a stub in the reference): the maximum-likelihood fit of
    wrapper.rs:89-207 over every pair of samples.  genotype_likelihoods:
    (V, N, 3) linear-scale likelihoods of 0/0, 0/1, 1/1; a row with a NaN is a
    missing call.  Hard calls (one-hot rows) take the 2-bit path, other rows
    the likelihood path; both give the reference's records.  Given allele
    frequencies are rejected as in wrapper.rs:113-122.  Returns (jacquard,
    kinship, inbreeding)."""
    if allele_frequencies is not None:
        raise NotImplementedError("allele frequencies with genotype likelihoods (wrapper.rs:113-122)")
    if iterations < 1:
        raise ValueError("iterations must be non-zero")
    gl = np.asarray(genotype_likelihoods, np.float64)
    if gl.ndim != 3:
        raise ValueError("genotype_likelihoods must have shape (sites, samples, genotypes)")
    if gl.shape[2] in (6, 10):  # every site tri- or tetra-allelic
        co = Cohort.gl_multi(**{("gl3" if gl.shape[2] == 6 else "gl4"): gl}, maf=minimum_allele_frequency,
                             af_iterations=iterations)
        try:
            out = co.fit_pairs(iterations=iterations)
        finally:
            co.close()
        return build_arrays(out, gl.shape[1])
    if gl.shape[2] != 3:
        raise ValueError("genotype_likelihoods need 3, 6 or 10 genotypes (2, 3 or 4 alleles)")
    missing = np.isnan(gl).any(axis=2)
    one_hot = ((gl == 1.0).sum(axis=2) == 1) & ((gl == 0.0).sum(axis=2) == 2)
    n = gl.shape[1]
    if (missing | one_hot).all():
        codes = np.where(missing, 3, np.argmax(np.nan_to_num(gl, nan=0.0), axis=2)).astype(np.uint8)
        co = Cohort(np.ascontiguousarray(codes.T), maf=minimum_allele_frequency, af_iterations=iterations)
    else:
        co = Cohort.gl(gl, maf=minimum_allele_frequency, af_iterations=iterations)
    try:
        out = co.fit_pairs(iterations=iterations)
    finally:
        co.close()
    return build_arrays(out, n)


def run_cli(argv=None) -> None:
    """The kestrel command line from Python (lib.rs:111-117 `run_cli`, the
    module's console-script entry point, pyproject.toml:17-18): the arguments
    are `sys.argv` (argv[0] the program name) and the shipped `kestrel` binary
    runs them as a child process, its log lines passed through to stderr.
    As with clap's `Args::parse_from` in cli.rs:73, a usage error ends the
    process with status 2 and `--help` / `--version` with status 0
    (SystemExit); a failure of the run raises RuntimeError with the command's
    error message (the reference's `PyRuntimeError(format!("{:#}", e))`)."""
    import subprocess
    import sys

    args = [str(a) for a in (sys.argv if argv is None else argv)]
    exe = os.path.join(_lib.HERE, "kestrel")
    if not os.path.exists(exe):
        raise ImportError(f"{exe} is not built; run __graft_entry__.build() (make -C kestrel_amd)")
    err = []
    with subprocess.Popen([exe] + args[1:], stderr=subprocess.PIPE, text=True) as p:
        for line in p.stderr:
            sys.stderr.write(line)
            err.append(line)
    if p.returncode == 2:  # usage error, reported by the argument parser
        raise SystemExit(2)
    if p.returncode != 0:
        msg = [ln[len("Error: "):].rstrip("\n") for ln in err if ln.startswith("Error: ")]
        raise RuntimeError(msg[-1] if msg else f"kestrel exited with status {p.returncode}")
    if any(a in ("-h", "--help", "-V", "--version") for a in args[1:]):
        raise SystemExit(0)
