"""Pins of the oracle's soft-likelihood paths (vcf.rs:338-410 rows through
allele/mod.rs:37-134, mixture.rs:96-236 and coefficients.rs:40-141)."""
import numpy as np
import pytest

from kestrel_amd import synth

E = 1e-8


@pytest.fixture(scope="module")
def soft(O):
    c = synth.generate(40, 3000, seed=21, n_family_pairs=8, missing=0.02)
    gl = synth.soft_likelihoods(c.codes, seed=22, mean_depth=3.0)  # (V, N, 3)
    g = np.ascontiguousarray(gl.transpose(1, 0, 2))  # oracle layout (N, V, 3)
    return c, g


def test_one_hot_rows_reduce_to_the_code_paths(O):
    """Hard calls as likelihood rows take exactly the 2-bit paths' arithmetic."""
    c = synth.generate(24, 1500, seed=23, missing=0.05)
    gl = np.eye(3)[np.minimum(c.codes, 2)]
    gl[c.codes == 3] = np.nan
    af = O.allele_freqs(c.codes)
    assert np.array_equal(O.allele_freqs_gl(gl), af)
    k = O.keep_sites(af)
    seg = np.array([0, 700, int(k.sum())], np.uint32)
    for order, sg in ((O.ORDER_REF, None), (O.ORDER_SEG, seg)):
        a = O.fit_pairs(c.codes[:, k], af[k], order=order, seg=sg)
        b = O.fit_pairs_gl(gl[:, k], af[k], order=order, seg=sg)
        for f in ("x", "y", "jacquard", "obj", "iters", "cond", "n_sites"):
            assert np.array_equal(a[f], b[f], equal_nan=True), f


def test_c_is_bilinear_in_the_likelihood_rows(O):
    """c(lx, ly) = sum_ij lx_i ly_j c(e_i, e_j): the nine blocks of
    mixture_row2 each carry one genotype pair (a wrong block index fails)."""
    rng = np.random.default_rng(3)
    eye = np.eye(3)
    for _ in range(100):
        p = rng.uniform(0.01, 0.99)
        af = np.array([p, 1 - p])
        lx, ly = rng.random(3), rng.random(3)
        c = O.mixture_row2(lx, ly, af)
        ref = sum(lx[i] * ly[j] * O.mixture_row2(eye[i], eye[j], af) for i in range(3) for j in range(3))
        np.testing.assert_allclose(c, ref, rtol=1e-14, atol=1e-300)


def test_soft_allele_freqs_are_stationary_points(O, soft):
    """f(p) = -(1/n) sum_i log(e + p^2 l0 + 2p(1-p) l1 + (1-p)^2 l2) over the
    called rows (allele/objective.rs:85-110); the SQP result is a root of f'."""
    from scipy.optimize import brentq

    c, g = soft
    af = O.allele_freqs_gl(g, maf=0.0)
    checked = 0
    for s in range(0, g.shape[1], 7):
        rows = g[:, s]
        rows = rows[~np.isnan(rows).any(1)]
        l0, l1, l2 = rows.T

        def d(p):
            den = E + p * p * l0 + 2 * p * (1 - p) * l1 + (1 - p) ** 2 * l2
            return ((2 * p * l0 + (2 - 4 * p) * l1 - 2 * (1 - p) * l2) / den).sum()

        p = af[s, 0]
        if p in (0.0, 1.0):
            continue
        lo, hi = max(1e-12, p - 1e-3), min(1 - 1e-12, p + 1e-3)
        if d(lo) * d(hi) > 0:
            continue
        assert abs(brentq(d, lo, hi, xtol=1e-15) - p) <= 1e-8, s
        checked += 1
    assert checked > 300


def test_soft_fit_satisfies_kkt_and_beats_random_points(O, soft):
    c, g = soft
    af = O.allele_freqs_gl(g)
    k = O.keep_sites(af)
    gk, afk = g[:, k], af[k]
    rng = np.random.default_rng(4)
    pairs = rng.integers(0, g.shape[0], (30, 2))
    out = O.fit_pairs_gl(gk, afk, pairs=pairs)
    for o in out:
        rows, valid, n = O.pair_rows_gl(gk[o["x"]], gk[o["y"]], afk)
        assert n == o["n_sites"]
        rows = rows[valid == 1]
        x = o["jacquard"]
        assert abs(x.sum() - 1.0) < 1e-12 and (x >= 0).all()
        L = rows @ x + E
        gr = -(rows / L[:, None]).sum(0) / n
        lam = x @ gr
        if o["iters"] < 100:
            for xi, gi in zip(x, gr):
                assert (gi >= lam - 1e-8 - 1e-10) if xi == 0.0 else abs(gi - lam) <= 1e-8 + 1e-10
        f = -np.log(L).mean()
        assert abs(f - o["obj"]) <= 1e-12 * abs(f)
        pts = rng.dirichlet(np.ones(9), 300)
        assert (-np.log(rows @ pts.T + E).mean(0) >= o["obj"] - 1e-12).all()


def test_partial_nan_row_is_missing(O):
    af = np.array([[0.3, 0.7]])
    lx = np.array([[0.2, np.nan, 1.0]])
    ly = np.array([[1.0, 0.5, 0.1]])
    rows, valid, n = O.pair_rows_gl(lx, ly, af)
    assert n == 0 and valid[0] == 0
