"""kestrel_evaluate_noalloc (device buffers in, device buffers out, scratch in a
fixed device array) against kestrel_evaluate, through the C ABI: bit-identical
f, g, H and n_xy on two calls in a row, the second with fewer points; points
beyond the scratch limit and pairs naming samples beyond the cohort are
rejected with KESTREL_EINVAL."""
import ctypes

import numpy as np
import pytest

from kestrel_amd import Cohort, _lib, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _hip():
    """The HIP runtime the library links (already loaded by it)."""
    _lib.load()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


H2D, D2H = 1, 2


class Dev:
    """A device copy of a host array (the caller's buffers)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert _hip().hipMalloc(ctypes.byref(self.p), max(1, self.a.nbytes)) == 0
        assert _hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert _hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        _hip().hipFree(self.p)


def noalloc(co, pairs, pts, eps=1e-8):
    n = len(pairs)
    dp = Dev(np.asarray(pairs, np.uint32))
    dx = Dev(np.asarray(pts, np.float64))
    f, g, h = Dev(np.full(n, np.nan)), Dev(np.full((n, 9), np.nan)), Dev(np.full((n, 81), np.nan))
    ns = Dev(np.zeros(n, np.uint32))
    rc = _lib.load().kestrel_evaluate_noalloc(co._h, dp.p, dx.p, n, eps, f.p, g.p, h.p, ns.p)
    if rc:
        return rc, None
    return rc, (f.get(), g.get(), h.get().reshape(n, 9, 9), ns.get())


def check_same(got, exp):
    f, g, h, ns = got
    fe, ge, he, nse = exp
    assert np.array_equal(ns, nse)
    for a, b in ((f, fe), (g, ge), (h, he)):
        assert ((a == b) | (np.isnan(a) & np.isnan(b))).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_evaluate(precision):
    say(f"case: kestrel_evaluate_noalloc vs kestrel_evaluate, 120 x 12000, {precision}")
    c = synth.generate(120, 12000, seed=91, n_family_pairs=8, missing=0.03)
    rng = np.random.default_rng(3)
    with Cohort(c.codes, precision=precision) as co:
        for n in (700, 90):  # two calls in a row, the second smaller
            pairs = rng.integers(0, 120, (n, 2))
            pts = rng.dirichlet(np.ones(9), n)
            rc, got = noalloc(co, pairs, pts)
            assert rc == 0, _lib.load().kestrel_last_error()
            exp = co.evaluate(pairs, pts)
            check_same(got, exp)
            if precision == "fp64":  # the table sweep (option class_sweep 0) as well
                co.set_option("class_sweep", 0)
                rc, got0 = noalloc(co, pairs, pts)
                assert rc == 0
                check_same(got0, co.evaluate(pairs, pts))
                co.set_option("class_sweep", 1)
        # points outside the sweeps' domain (eps = 0, negative coefficients)
        pairs = rng.integers(0, 120, (40, 2))
        pts = rng.dirichlet(np.ones(9), 40)
        pts[:10] = np.eye(9)[rng.integers(0, 9, 10)]
        pts[10:15, 2] = -0.3
        for eps in (0.0, 1e-8):
            rc, got = noalloc(co, pairs, pts, eps)
            assert rc == 0
            check_same(got, co.evaluate(pairs, pts, eps=eps))
        # a pair naming a sample beyond the cohort
        bad = np.array([[0, 5], [3, 120]])
        rc, _ = noalloc(co, bad, np.full((2, 9), 1 / 9))
        assert rc == _lib.EINVAL
        # more points than the scratch array holds (the table sweep's slots
        # are the smallest: beyond them every form rejects)
        big = 64 * 2**20 // (88 + 448 * 3) + 512
        rc, _ = noalloc(co, np.zeros((big, 2), np.uint32), np.full((big, 9), 1 / 9))
        assert rc == _lib.EINVAL
        assert b"limit" in _lib.load().kestrel_last_error()
        # and still correct after the rejected calls
        pairs = rng.integers(0, 120, (33, 2))
        pts = rng.dirichlet(np.ones(9), 33)
        rc, got = noalloc(co, pairs, pts)
        assert rc == 0
        check_same(got, co.evaluate(pairs, pts))


@pytest.mark.parametrize("rows", ["soft", "one-hot"])
def test_noalloc_multi_allelic_cohort(rows):
    say(f"case: kestrel_evaluate_noalloc on a 2/3/4-allele likelihood cohort, {rows} rows")
    if rows == "soft":
        gl, maf = synth.multi_likelihoods(40, (3000, 600, 200), seed=12), None
    else:  # hard calls: the 2-allele part in the class-sorted form
        gl, maf = synth.multi_onehot(synth.multi_genotypes(40, (3000, 600, 200), seed=12, missing=0.02)), 0.0
    with Cohort.gl_multi(*gl, maf=maf) as co:
        rng = np.random.default_rng(8)
        for n in (300, 17):
            pairs = rng.integers(0, 40, (n, 2))
            pts = rng.dirichlet(np.ones(9), n)
            rc, got = noalloc(co, pairs, pts)
            assert rc == 0, _lib.load().kestrel_last_error()
            check_same(got, co.evaluate(pairs, pts))
