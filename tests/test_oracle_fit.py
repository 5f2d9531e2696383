"""Pins of the allele-frequency fit (allele/mod.rs:37-134), the pair fit
(sqp.rs:49-93 via coefficients.rs:40-141) and the golden fixtures."""
import math
import os

import numpy as np
import pytest

from kestrel_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden", "c1_small.npz")


def test_allele_freqs_match_stationary_point(O):
    """For one-hot likelihoods the HWE objective (allele/objective.rs:85-110) is
    f(p) = -(1/n)[n0 log(p^2 + e) + n1 log(2p(1-p) + e) + n2 log((1-p)^2 + e)];
    its maximiser, found independently by a root of f'(p), is the allele count
    (2 n0 + n1) / (2 n) up to the e = 1e-8 regulariser.  The SQP lands on it."""
    from scipy.optimize import brentq

    c = synth.generate(150, 4000, seed=3, missing=0.05)
    af = O.allele_freqs(c.codes, maf=0.0)
    g = c.codes
    n0, n1, n2 = (g == 0).sum(0), (g == 1).sum(0), (g == 2).sum(0)
    e = 1e-8
    count = (2 * n0 + n1) / (2.0 * (n0 + n1 + n2))
    checked = 0
    for s in range(g.shape[1]):
        if count[s] in (0.0, 1.0):
            assert af[s, 0] == count[s]
            continue
        def d(p):
            return (n0[s] * 2 * p / (p * p + e) + n1[s] * (2 - 4 * p) / (2 * p * (1 - p) + e)
                    - n2[s] * 2 * (1 - p) / ((1 - p) ** 2 + e))
        lo, hi = max(1e-12, count[s] - 1e-3), min(1 - 1e-12, count[s] + 1e-3)
        p_star = brentq(d, lo, hi, xtol=1e-15)
        # the SQP stops at |g_i - lambda| <= 1e-8 (sqp.rs:95-109)
        assert abs(af[s, 0] - p_star) <= 1e-8, (s, af[s, 0], p_star)
        assert abs(p_star - count[s]) <= 2e-6
        checked += 1
    assert checked > 3000
    np.testing.assert_allclose(af[:, 0] + af[:, 1], 1.0, atol=1e-15)


def test_maf_zeroes_rare_alleles_and_drops_site(O):
    # 100 samples, site 0: one het (freq 1/200); site 1: 3 hets
    codes = np.zeros((100, 2), np.uint8)
    codes[0, 0] = 1
    codes[:3, 1] = 1
    af = O.allele_freqs(codes, maf=0.01)
    assert af[0, 1] == 0.0 and af[0, 0] == 1.0
    assert af[1, 1] > 0.0
    assert list(O.keep_sites(af)) == [False, True]


def test_site_without_called_sample_is_an_error(O):
    codes = np.zeros((10, 3), np.uint8)
    codes[:, 1] = 3
    with pytest.raises(ValueError):
        O.allele_freqs(codes)


def test_missing_calls_are_skipped(O):
    """NaN rows are dropped (lanevector.rs:43-77): adding missing calls to a
    site does not change its frequency."""
    g = np.array([[0], [1], [2], [1], [0]], np.uint8)
    g2 = np.vstack([g, np.full((7, 1), 3, np.uint8)])
    a = O.allele_freqs(g, maf=0.0)
    b = O.allele_freqs(g2, maf=0.0)
    assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def fit_c1(O):
    c = synth.generate(48, 20000, seed=5, n_family_pairs=12, missing=0.01, inbred_fraction=0.0)
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    codes, afk = c.codes[:, kept], af[kept]
    return c, codes, afk, O.fit_pairs(codes, afk, order=O.ORDER_REF)


def test_fit_satisfies_kkt_and_beats_random_points(O, fit_c1):
    c, codes, afk, out = fit_c1
    rng = np.random.default_rng(0)
    for i in rng.choice(len(out), 25, replace=False):
        o = out[i]
        rows, valid, n = O.pair_rows_codes(codes[o["x"]], codes[o["y"]], afk)
        rows = rows[valid == 1]
        x = o["jacquard"]
        assert abs(x.sum() - 1.0) < 1e-12 and (x >= 0).all()
        L = rows @ x + 1e-8
        g = -(rows / L[:, None]).sum(0) / n
        lam = x @ g
        tol = 1e-8 + 1e-10
        for xi, gi in zip(x, g):
            if xi == 0.0:
                assert gi >= lam - tol
            else:
                assert abs(gi - lam) <= tol
        f = -np.log(L).mean()
        assert abs(f - o["obj"]) <= 1e-12 * abs(f)
        pts = rng.dirichlet(np.ones(9), 300)
        fr = -np.log(rows @ pts.T + 1e-8).mean(0)
        assert (fr >= o["obj"] - 1e-12).all()


def test_relatives_recover_expected_coefficients(fit_c1):
    """Statistical sanity (old/pedigrees.py:197-239 kinships): PO 1/4 via
    delta8 = 1; FS (1/4, 1/2, 1/4) on delta7..9; HS delta8 = 1/2; FC delta8 = 1/4."""
    c, codes, afk, out = fit_c1
    idx = {(int(o["x"]), int(o["y"])): o for o in out}
    kin_vec = np.array([1, 0, .5, 0, .5, 0, .5, .25, 0])
    expect = {"PO": 0.25, "FS": 0.25, "HS": 0.125, "FC": 0.0625}
    for x, y, kind in c.pairs:
        o = idx[(min(x, y), max(x, y))]
        kin = o["jacquard"] @ kin_vec
        assert abs(kin - expect[kind]) < 0.03, (kind, kin)
        if kind == "PO":
            assert o["jacquard"][7] > 0.9
    unrelated = [o for o in out if o["x"] != o["y"] and (int(o["x"]), int(o["y"])) not in
                 {(min(a, b), max(a, b)) for a, b, _ in c.pairs} and o["x"] >= 2 * len(c.pairs)
                 and o["y"] >= 2 * len(c.pairs)]
    assert np.median([o["jacquard"][8] for o in unrelated]) > 0.9


def test_self_pairs_start_point_and_values(fit_c1):
    """coefficients.rs:88-92: a sample with itself starts at delta1 = delta7 = 1/2;
    an outbred sample fits delta7 = 1 (kinship 1/2)."""
    _, _, _, out = fit_c1
    selfp = out[out["x"] == out["y"]]
    kin = selfp["jacquard"] @ np.array([1, 0, .5, 0, .5, 0, .5, .25, 0])
    assert np.median(kin) > 0.49


def test_segmented_order_within_tolerance_of_reference_order(O, fit_c1):
    c, codes, afk, out = fit_c1
    seg = np.linspace(0, codes.shape[1], 6).astype(np.uint32)
    out2 = O.fit_pairs(codes, afk, order=O.ORDER_SEG, seg=seg)
    lc = np.log10(out["cond"])
    well = np.isfinite(lc) & (lc < 12)
    d = np.abs(out["jacquard"] - out2["jacquard"]).max(1)
    assert d[well].max() <= 1e-6
    rel = np.abs(out["obj"] - out2["obj"]) / np.abs(out["obj"])
    assert rel[well].max() <= 1e-9


def test_iteration_cap_reports_nan_condition(O, fit_c1):
    c, codes, afk, _ = fit_c1
    out = O.fit_pairs(codes, afk, pairs=np.array([[0, 1], [2, 3]]), iterations=1)
    for o in out:
        if o["iters"] == 1:
            assert math.isnan(o["cond"])


def test_samples_beyond_cohort_rejected(O, fit_c1):
    _, codes, afk, _ = fit_c1
    with pytest.raises(ValueError):
        O.fit_pairs(codes, afk, nsel=codes.shape[0] + 1)


def test_golden_fixture_reproduces(O):
    z = np.load(GOLD)
    af = O.allele_freqs(z["codes"])
    assert np.array_equal(af, z["af"])
    kept = O.keep_sites(af)
    assert np.array_equal(kept, z["kept"])
    codes, afk = z["codes"][:, kept], af[kept]
    ref = O.fit_pairs(codes, afk, order=O.ORDER_REF)
    seg = O.fit_pairs(codes, afk, order=O.ORDER_SEG, seg=z["seg"])
    for k in ("jacquard", "obj", "iters", "n_sites", "x", "y"):
        assert np.array_equal(ref[k], z["out_ref"][k]), k
        assert np.array_equal(seg[k], z["out_seg"][k]), k
