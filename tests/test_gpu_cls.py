"""GPU parity of the CLS / Brier estimator (CLI default -f GT) against the
oracle restatement of coefficients.rs:166-320.  The device sums c in the
reference's own sequential site order and Q in its 8-lane tile order, so the
records are compared bit for bit."""
import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, synth

pytestmark = pytest.mark.gpu


def same(a, b):
    return (a == b) | (np.isnan(a) & np.isnan(b))


def check(out, exp):
    for k in ("x", "y", "jacquard", "obj", "iters", "cond", "n_sites"):
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} pairs differ"


@pytest.mark.parametrize("n,v,seed", [(40, 3001, 1), (90, 12345, 2)])
def test_cls_bitexact(O, n, v, seed):
    print(f"case: cls {n} x {v}", flush=True)
    c = synth.generate(n, v, seed=seed, n_family_pairs=8, missing=0.0)
    with Cohort.gt(c.codes) as co:
        out, q = co.fit_pairs_gt()
        af, kept = co.allele_freqs()
    exp, qe = O.cls_fit_pairs(c.codes)
    assert np.array_equal(q, qe)
    assert np.array_equal(af, O.gt_allele_counts(c.codes))
    check(out, exp)


def test_cls_drops_sites_with_missing_calls_and_keeps_monomorphic(O):
    print("case: cls missing/monomorphic", flush=True)
    c = synth.generate(30, 2000, seed=3, missing=0.0)
    codes = c.codes.copy()
    codes[:, :20] = 0            # monomorphic sites stay on the GT path
    codes[4, 100:130] = 3        # sites with a missing call are dropped (vcf.rs:217-221)
    with Cohort.gt(codes) as co:
        out, q = co.fit_pairs_gt(iterations=50)
        _, kept = co.allele_freqs()
    assert kept[:20].all() and not kept[100:130].any()
    keep = (codes != 3).all(0)
    exp, qe = O.cls_fit_pairs(codes[:, keep], iterations=50)
    assert np.array_equal(q, qe)
    check(out, exp)


def test_cls_pairs_subset_and_errors(O):
    print("case: cls subset", flush=True)
    c = synth.generate(25, 1500, seed=4, missing=0.0)
    with Cohort.gt(c.codes) as co:
        pairs = np.array([[3, 9], [9, 3], [7, 7], [0, 24]])
        out, _ = co.fit_pairs_gt(pairs, iterations=7)
        exp, _ = O.cls_fit_pairs(c.codes, pairs=pairs, iterations=7)
        check(out, exp)
        sub, _ = co.fit_pairs_gt(n_select=6)
        assert len(sub) == 21
        with pytest.raises(KestrelError):
            co.fit_pairs_gt(n_select=26)
        with pytest.raises(KestrelError):
            co.fit_pairs()  # the ML fit needs an ML cohort


# ---------------------------------------------------------------- 3 and 4 alleles

def kept_buckets(idx):
    """The oracle's inputs: sites with a missing call (255) dropped per bucket."""
    keep = [~(b == 255).any(axis=0) for b in idx]
    return [b[:, k] for b, k in zip(idx, keep)], keep


@pytest.mark.parametrize("sites,seed", [((300, 90, 40), 11), ((0, 70, 0), 12), ((0, 0, 33), 13),
                                        ((1000, 3, 1), 14)])
def test_cls_multi_bitexact(O, sites, seed):
    print(f"case: cls multi {sites} seed {seed}", flush=True)
    idx = synth.multi_genotypes(48, sites, seed=seed, missing=0.003)
    kb, keep = kept_buckets(idx)
    with Cohort.gt_multi(*idx) as co:
        out, q = co.fit_pairs_gt()
        parts, afs = co.multi_info()
    exp, qe = O.cls_fit_pairs_multi(*kb)
    assert parts == tuple(int(k.sum()) for k in keep)
    for A, k, b in zip((2, 3, 4), keep, kb):
        assert np.isnan(afs[A][~k]).all()
        assert np.array_equal(afs[A][k], O.gt_allele_counts_generic(A, b))
    assert np.array_equal(q, qe)
    check(out, exp)


def test_cls_multi_given_frequencies_and_pairs(O):
    print("case: cls multi given af, pair list, iteration cap", flush=True)
    idx = synth.multi_genotypes(30, (200, 60, 25), seed=15)
    rng = np.random.default_rng(3)
    af = [rng.dirichlet(np.ones(A), v) for A, v in zip((2, 3, 4), (200, 60, 25))]
    pairs = np.array([[3, 9], [9, 3], [7, 7], [0, 29], [12, 13]])
    with Cohort.gt_multi(*idx, af=af) as co:
        out, q = co.fit_pairs_gt(pairs, iterations=5)
        sub, _ = co.fit_pairs_gt(n_select=7)
    exp, qe = O.cls_fit_pairs_multi(*idx, af=af, pairs=pairs, iterations=5)
    assert np.array_equal(q, qe)
    check(out, exp)
    exp, _ = O.cls_fit_pairs_multi(*idx, af=af, nsel=7)
    check(sub, exp)


def test_cls_multi_with_only_two_alleles_equals_biallelic_entry(O):
    print("case: cls gt_multi(2-allele only) vs gt", flush=True)
    c = synth.generate(26, 2100, seed=16, missing=0.0)
    with Cohort.gt(c.codes) as a, Cohort.gt_multi(c.codes) as b:
        oa, qa = a.fit_pairs_gt()
        ob, qb = b.fit_pairs_gt()
    assert np.array_equal(qa, qb)
    check(ob, oa)


def test_cls_multi_rejections():
    print("case: cls multi rejections", flush=True)
    idx = synth.multi_genotypes(10, (20, 8, 4), seed=17)
    bad = idx[1].copy()
    bad[2, 3] = 6  # a 3-allele site has genotype indices 0..5
    with pytest.raises(KestrelError):
        Cohort.gt_multi(idx[0], bad, idx[2])
    with pytest.raises(KestrelError):  # frequencies for some buckets only
        Cohort.gt_multi(*idx, af=[np.full((20, 2), 0.5), None, None])
    with pytest.raises(KestrelError):  # every site dropped for a missing call
        Cohort.gt_multi(None, np.full((10, 5), 255, np.uint8)).fit_pairs_gt()
