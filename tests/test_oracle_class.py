"""The oracle's class order (ORDER_CLASS, oracle/ko_class.c) pinned without the
device: its sums against a long-double evaluation of the reference's own
per-site terms (mixture_row2 rows, jacquard/objective.rs, grad_hess.rs), so a
wrong position, factor or power in any genotype-pair pattern fails; its fits
against the reference order; its expansion against the reference's rows on
single sites of every genotype pair."""
import numpy as np
import pytest

from kestrel_amd import synth


def _cohort(O, n, v, seed, missing):
    c = synth.generate(n, v, seed=seed, n_family_pairs=max(2, n // 8), missing=missing, inbred_fraction=0.05)
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    return np.ascontiguousarray(c.codes[:, kept]), af[kept]


def _longdouble_sums(O, gx, gy, af, pt, eps=1e-8):
    rows, valid, n = O.pair_rows_codes(gx, gy, af)
    r = rows[valid.astype(bool)].astype(np.longdouble)
    L = np.longdouble(eps) + r @ np.asarray(pt, np.longdouble)
    F = np.sum(np.log(L))
    G = (r / L[:, None]).sum(0)
    c = r / L[:, None]
    H = c.T @ c
    return F, G, H, n


@pytest.mark.parametrize("seed,missing", [(3, 0.05), (4, 0.3)])
def test_class_sums_against_long_double(O, seed, missing):
    codes, af = _cohort(O, 30, 9000, seed, missing)
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(9)
    for it in range(24):
        x, y = rng.integers(0, 30, 2)
        pt = rng.dirichlet(np.ones(9))
        if it % 3 == 0:
            pt[rng.integers(0, 9, 4)] = 0.0
            pt /= pt.sum()
        seg = np.array([0, 4096, codes.shape[1]], np.uint32) if it % 2 else None
        F, G, H, n = O.class_sums(codes[x], codes[y], af, pt, 1e-8, seg=seg)
        Fr, Gr, Hr, nr = _longdouble_sums(O, codes[x], codes[y], af, pt)
        assert n == nr
        assert abs(F - Fr) <= 1e-14 * abs(Fr)
        assert np.all(np.abs(G - Gr) <= 1e-14 * np.abs(Gr))
        scale = np.sqrt(np.outer(np.diag(Hr), np.diag(Hr)))[iu]
        assert np.all(np.abs(H - Hr[iu]) <= 1e-14 * scale)


def test_class_expansion_on_single_sites(O):
    """One site of each genotype pair: the expanded sums are the site's own
    c / L and c c^T / L^2 (the positions and factors of every pattern)."""
    rng = np.random.default_rng(0)
    for gx in range(3):
        for gy in range(3):
            p = rng.uniform(0.05, 0.95)
            af = np.array([[p, 1.0 - p]])
            pt = rng.dirichlet(np.ones(9))
            F, G, H, n = O.class_sums(np.array([gx]), np.array([gy]), af, pt)
            c, _, _ = O.pair_rows_codes(np.array([gx]), np.array([gy]), af)
            L = 1e-8 + c[0] @ pt
            assert n == 1
            np.testing.assert_allclose(G, c[0] / L, rtol=1e-14, atol=0)
            cc = np.outer(c[0], c[0]) / L ** 2
            np.testing.assert_allclose(H, cc[np.triu_indices(9)], rtol=1e-14, atol=0)
            assert abs(F - np.log(L)) <= 1e-15 * abs(np.log(L))
    # a missing call on either side contributes nothing
    F, G, H, n = O.class_sums(np.array([3, 0]), np.array([1, 3]), np.array([[0.3, 0.7], [0.4, 0.6]]), np.full(9, 1 / 9))
    assert n == 0 and F == 0.0 and not G.any() and not H.any()


def test_class_fits_against_reference_order(O):
    codes, af = _cohort(O, 24, 6000, 9, 0.02)
    a = O.fit_pairs(codes, af, order=O.ORDER_CLASS)
    b = O.fit_pairs(codes, af, order=O.ORDER_REF)
    lc = np.log10(b["cond"])
    well = np.isfinite(lc) & (lc < 12)
    # the fits stop at |g_i - lambda| <= 1e-8 (sqp.rs:95-109), so two summation
    # orders may stop 1e-8 apart: the stated bars (max |dDelta_k| <= 1e-6,
    # relative objective <= 1e-9)
    assert np.abs(a["jacquard"] - b["jacquard"]).max(1)[well].max() <= 1e-6
    assert (np.abs(a["obj"] - b["obj"]) / np.abs(b["obj"]))[well].max() <= 1e-9
    assert np.array_equal(a["n_sites"], b["n_sites"])
