"""The oracle's reference-order fit in the reference's own AVX-512 loop shape
(ORDER_FAST: c_s rows built per 8-site block and pushed through the NaN
filter, the getexp/getmant log, the Hessian in three 15-accumulator passes per
32-block tile) against the oracle's 8-lane loops (ORDER_REF): the same bits.
This is the form the benchmark's CPU baseline times."""
import numpy as np
import pytest

from kestrel_amd import synth


@pytest.fixture(scope="module")
def fast(O):
    if not O.lib().ko_has_avx512():
        pytest.skip("the fast loops need an AVX-512 host")
    return O


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return ((a == b) | (np.isnan(a) & np.isnan(b))).all()


def assert_same_records(a, b):
    for k in ("x", "y", "jacquard", "obj", "iters", "n_sites", "cond"):
        assert same(a[k], b[k]), k


def test_log8_matches_scalar_log(fast):
    O = fast
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-320, 300, 3000), rng.uniform(0.5, 2.0, 2000),
                         [0.0, -0.0, -1.0, np.inf, -np.inf, np.nan, 5e-324, 2.2250738585072014e-308, 1.0, 0.75,
                          1.5, 1.4999999999999998]])
    for i, x in enumerate(xs):
        a = O.lib().ko_lane_log8_fast(float(x), i % 8)
        b = O.log(float(x))
        assert (a == b) or (np.isnan(a) and np.isnan(b)), (x, a, b)


def test_fast_fit_matches_reference_order_c1(fast):
    O = fast
    c = synth.config("c1")
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    assert_same_records(O.fit_pairs(kc, ka, order=O.ORDER_FAST), O.fit_pairs(kc, ka, order=O.ORDER_REF))


@pytest.mark.parametrize("v", [7, 8, 9, 255, 263, 1031])
def test_fast_fit_ragged_and_missing(fast, v):
    """Site counts that leave remainders of every length, heavy ragged
    missingness (partial blocks, whole missing blocks), self pairs, a pair
    without usable sites, and the tile boundaries (> 32 blocks)."""
    O = fast
    rng = np.random.default_rng(v)
    c = synth.generate(12, v, seed=v, n_family_pairs=2, missing_per_sample_max=0.6, maf=0.05)
    codes = c.codes.copy()
    codes[11, :] = 3  # a sample with no call
    af = O.allele_freqs(codes[:11])
    kept = O.keep_sites(af)
    kc, ka = np.ascontiguousarray(codes[:, kept]), af[kept]
    pairs = np.concatenate([rng.integers(0, 12, (60, 2)), [[3, 3], [0, 11], [11, 11]]])
    a = O.fit_pairs(kc, ka, pairs=pairs, order=O.ORDER_FAST)
    b = O.fit_pairs(kc, ka, pairs=pairs, order=O.ORDER_REF)
    assert_same_records(a, b)
    assert (a["n_sites"][-2:] == 0).all()
