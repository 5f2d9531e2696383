"""Soft genotype likelihoods on the device (kestrel_cohort_create_gl, sweep_gl)
against the oracle.  Bars as the hard-call parity: allele frequencies, masks,
n_xy, Delta, objective, iterations and condition bit-exact in the device's
summation order; the reference's order within the stated tolerance; one-hot
rows through the likelihood path equal the 2-bit path bit for bit."""
import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, build_arrays, relatedness_coefficients_gl, segment_bounds, synth

pytestmark = pytest.mark.gpu

TOL_DELTA = 1e-6


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(out, exp):
    for k in ("x", "y", "jacquard", "obj", "iters", "n_sites", "cond"):
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} pairs differ"


@pytest.fixture(scope="module")
def soft():
    c = synth.generate(64, 9000, seed=31, n_family_pairs=10, missing=0.02)
    gl = synth.soft_likelihoods(c.codes, seed=32, mean_depth=4.0)  # (V, N, 3)
    gl[5, 7, 1] = np.nan  # one value missing: the whole call is missing (lanevector.rs:100-104)
    co = Cohort.gl(gl)
    yield c, gl, co
    co.close()


def test_one_hot_rows_equal_the_2bit_path():
    """A cohort of one-hot likelihood rows is held as its 2-bit codes (it
    serves, which likelihood-row cohorts do not) with the codes cohort's
    records; and one-hot rows through the likelihood sweep (a cohort made
    soft by one extra sample, frequencies given) give the same records for
    the pairs among the one-hot samples."""
    say("case: one-hot likelihood rows vs 2-bit codes, 40 x 4000")
    c = synth.generate(40, 4000, seed=33, missing=0.03)
    gl = np.eye(3)[np.minimum(c.codes, 2)]
    gl[c.codes == 3] = np.nan
    g = np.ascontiguousarray(gl.transpose(1, 0, 2))
    with Cohort.gl(g) as a, Cohort(c.codes) as b:
        afa, ka = a.allele_freqs()
        afb, kb = b.allele_freqs()
        assert np.array_equal(afa, afb) and np.array_equal(ka, kb)
        base = b.fit_pairs()
        assert_bitexact(a.fit_pairs(), base)
        a.serve_start(max_query_rows=4, max_pairs=1000)  # held as 2-bit codes
        a.serve_stop()
    soft_sample = np.full((4000, 1, 3), 0.5)
    soft_sample[:, 0, 0] = 1.0
    with Cohort.gl(np.concatenate([g, soft_sample], 1), af=afb) as m, Cohort(c.codes, af=afb) as b2:
        b2.set_option("class_sweep", 0)  # the likelihood sweep keeps the table sweep's order
        with pytest.raises(KestrelError):
            m.serve_start(max_query_rows=4, max_pairs=1000)  # likelihood rows
        pairs = np.array([[i, j] for i in range(40) for j in range(i, 40)], np.uint32)
        assert_bitexact(m.fit_pairs(pairs=pairs), b2.fit_pairs(pairs=pairs))


def test_claim_order_tiles_keep_records(soft):
    """Pairs claimed in sample tiles (option pair_tile, automatic for
    likelihood rows) give the same records, in list order, as list order."""
    c, gl, co = soft
    say("case: soft fit, claim order in tiles of 64 / 5 samples vs list order")
    co.set_option("slots", 256)  # fewer slots than pairs: the claim order matters
    try:
        co.set_option("pair_tile", 0)
        base = co.fit_pairs()
        for tile in (-1, 5):
            co.set_option("pair_tile", tile)
            assert_bitexact(co.fit_pairs(), base)
        rng = np.random.default_rng(3)
        pairs = rng.integers(0, 64, (3000, 2))
        co.set_option("pair_tile", 0)
        lst = co.fit_pairs(pairs)
        co.set_option("pair_tile", 8)
        assert_bitexact(co.fit_pairs(pairs), lst)
    finally:
        co.set_option("pair_tile", -1)
        co.set_option("slots", 0)


def test_soft_allele_freqs_bitexact(O, soft):
    c, gl, co = soft
    say("case: soft allele frequencies 64 x 9000")
    af, kept = co.allele_freqs()
    exp = O.allele_freqs_gl(gl.transpose(1, 0, 2))
    assert np.array_equal(af, exp)
    assert np.array_equal(kept, O.keep_sites(exp))


def test_soft_fit_bitexact_and_tolerance(O, soft):
    c, gl, co = soft
    say("case: soft fit, all 2080 pairs")
    af, kept = co.allele_freqs()
    g = np.ascontiguousarray(gl.transpose(1, 0, 2)[:, kept])
    out = co.fit_pairs()
    seg = np.minimum(segment_bounds(co.n_kept), g.shape[1]).astype(np.uint32)
    assert_bitexact(out, O.fit_pairs_gl(g, af[kept], order=O.ORDER_SEG, seg=seg))
    ref = O.fit_pairs_gl(g, af[kept], order=O.ORDER_REF)
    lc = np.log10(ref["cond"])
    well = np.isfinite(lc) & (lc < 12)
    assert np.abs(out["jacquard"] - ref["jacquard"]).max(1)[well].max() <= TOL_DELTA
    assert (np.abs(out["obj"] - ref["obj"]) / np.abs(ref["obj"]))[well].max() <= 1e-9


def test_soft_mixed_precision_form(soft):
    """The mixed-precision form on likelihood rows (option "precision"): Delta
    and objective within 1e-6 of the FP64 form on the pairs both bring to
    convergence with log10 condition below 12; n_xy equal."""
    c, gl, co = soft
    say("case: soft fit, mixed-precision form vs FP64 form")
    b = co.fit_pairs()
    with Cohort.gl(gl, precision="mixed") as m:
        a = m.fit_pairs()
    assert np.array_equal(a["n_sites"], b["n_sites"])
    lb, la = np.log10(b["cond"]), np.log10(a["cond"])
    both = np.isfinite(lb) & np.isfinite(la) & (lb < 12)
    assert both.sum() > 0.9 * len(b)
    d = np.abs(a["jacquard"] - b["jacquard"]).max(1)[both]
    say(f"soft mixed vs fp64: max |dDelta| {d.max():.3g} over {both.sum()} pairs; iterations "
        f"{a['iters'].mean():.2f} vs {b['iters'].mean():.2f}")
    assert d.max() <= 1e-6
    assert (np.abs(a["obj"] - b["obj"]) / np.abs(b["obj"]))[both].max() <= 1e-6


def test_rows_without_an_exact_one(O, soft):
    """Rows scaled by their maximum are stored as the two entries beside the
    exact 1.0 (16 bytes); a cohort with a called row that has no entry equal
    to 1 keeps the three likelihoods (24 bytes).  Both match the oracle."""
    c, gl, _ = soft
    say("case: a called row without an exact 1.0, 64 x 9000")
    g2 = gl.copy()
    g2[100, 3] = (0.5, 0.25, 0.125)
    g2[200, 9] = (1.0, 1.0, 0.25)  # two exact ones: the first is the one left out
    with Cohort.gl(g2) as co:
        af, kept = co.allele_freqs()
        g = np.ascontiguousarray(g2.transpose(1, 0, 2)[:, kept])
        pairs = np.array([[3, 9], [3, 40], [9, 9], [0, 63], [12, 3]])
        out = co.fit_pairs(pairs)
        seg = np.minimum(segment_bounds(co.n_kept), g.shape[1]).astype(np.uint32)
        assert_bitexact(out, O.fit_pairs_gl(g, af[kept], pairs=pairs, order=O.ORDER_SEG, seg=seg))


def test_soft_evaluate_bitexact(O, soft):
    c, gl, co = soft
    say("case: soft evaluate, 100 points")
    af, kept = co.allele_freqs()
    g = gl.transpose(1, 0, 2)[:, kept]
    rng = np.random.default_rng(1)
    pairs = rng.integers(0, 64, (100, 2))
    pairs[0] = (7, 9)  # sample 7 has the partly missing call
    pts = rng.dirichlet(np.ones(9), 100)
    f, gr, h, ns = co.evaluate(pairs, pts)
    seg = np.minimum(segment_bounds(co.n_kept), g.shape[1]).astype(np.uint32)
    for i in range(100):
        rows, valid, n = O.pair_rows_gl(g[pairs[i, 0]], g[pairs[i, 1]], af[kept])
        assert ns[i] == n
        assert f[i] == O.obj_seg(rows, valid, seg, pts[i])
        ge, he = O.gh_seg(rows, valid, seg, pts[i])
        assert np.array_equal(gr[i], ge) and np.array_equal(h[i], he)


def test_relatedness_coefficients_gl_soft(O):
    say("case: relatedness_coefficients_gl on soft rows, 20 x 2500")
    c = synth.generate(20, 2500, seed=34, missing=0.01)
    gl = synth.soft_likelihoods(c.codes, seed=35)
    jac, kin, inb = relatedness_coefficients_gl(gl)
    g = gl.transpose(1, 0, 2)
    af = O.allele_freqs_gl(g)
    k = O.keep_sites(af)
    seg = np.minimum(segment_bounds(int(k.sum())), int(k.sum())).astype(np.uint32)
    ej, ek, ei = build_arrays(O.fit_pairs_gl(g[:, k], af[k], order=O.ORDER_SEG, seg=seg), 20)
    assert np.array_equal(jac, ej) and np.array_equal(kin, ek) and np.array_equal(inb, ei)


def test_serving_rejects_soft_cohort(soft):
    c, gl, co = soft
    with pytest.raises(Exception):
        co.serve_start(max_query_rows=4, max_pairs=100)
