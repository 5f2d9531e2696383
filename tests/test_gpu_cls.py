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
