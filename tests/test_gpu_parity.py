"""GPU parity: libkestrel_amd (gfx950 kernels, called through the C ABI)
against the oracle.

Bars:
  * allele frequencies, kept-site masks, n_xy, iterations: bit-exact;
  * objective / gradient / Hessian and the whole fit: bit-exact against the
    oracle evaluated in the device's summation order -- the class order of
    the class-sorted sweep (the default for FP64 fits of hard calls), the
    segmented order of the table sweep (option class_sweep 0);
  * against the reference's own 8-lane order: max |dDelta_k| <= 1e-6,
    |d kinship| <= 1e-6, relative objective <= 1e-9, log10 condition within
    0.1, on pairs with log10(condition) < 12 (the reference's order differs
    only in floating-point summation order).
"""
import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, build_arrays, kinship, segment_bounds, synth

pytestmark = pytest.mark.gpu

TOL_DELTA = 1e-6
TOL_KIN = 1e-6
TOL_OBJ_REL = 1e-9
TOL_LOG10_COND = 0.1


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def seg_for(n_kept, V, segment_sites=4096):
    return np.minimum(segment_bounds(n_kept, segment_sites), V).astype(np.uint32)


def class_fit(O, codes, af, **kw):
    """The oracle's records in the class-sorted sweep's order (the default
    form of FP64 fits of hard calls)."""
    return O.fit_pairs(np.ascontiguousarray(codes), af, order=O.ORDER_CLASS, **kw)


def class_eval(O, gx, gy, af, pt):
    F, G, H, n = O.class_sums(gx, gy, af, pt)
    h = np.zeros((9, 9))
    iu = np.triu_indices(9)
    h[iu] = H / n
    h[(iu[1], iu[0])] = H / n
    return -F / n, -G / n, h, n


def assert_bitexact(out, exp):
    for k in ("x", "y", "jacquard", "obj", "iters", "n_sites", "cond"):
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} pairs differ"


def assert_within_tolerance(out, ref):
    lc = np.log10(ref["cond"])
    well = np.isfinite(lc) & (lc < 12)
    d = np.abs(out["jacquard"] - ref["jacquard"]).max(1)
    assert d[well].max() <= TOL_DELTA
    assert np.abs(kinship(out["jacquard"]) - kinship(ref["jacquard"]))[well].max() <= TOL_KIN
    assert (np.abs(out["obj"] - ref["obj"]) / np.abs(ref["obj"]))[well].max() <= TOL_OBJ_REL
    lo = np.log10(out["cond"])
    assert np.abs(lo[well] - lc[well]).max() <= TOL_LOG10_COND


@pytest.fixture(scope="module")
def golden():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_small.npz"))
    return {k: z[k] for k in z.files}


def test_golden_allele_freqs_and_fit(O, golden):
    say("case: golden c1_small (40 x 3000)")
    with Cohort(golden["codes"], segment_sites=int(golden["segment_sites"])) as co:
        af, kept = co.allele_freqs()
        assert np.array_equal(af, golden["af"])
        assert np.array_equal(kept, golden["kept"])
        assert co.n_kept == int(golden["kept"].sum())
        out = co.fit_pairs()
        co.set_option("class_sweep", 0)  # the table sweep: the fixture's segmented order
        table = co.fit_pairs()
    assert_bitexact(table, golden["out_seg"])
    assert_within_tolerance(table, golden["out_ref"])
    assert_bitexact(out, class_fit(O, golden["codes"][:, golden["kept"]], golden["af"][golden["kept"]]))
    assert_within_tolerance(out, golden["out_ref"])


@pytest.fixture(scope="module")
def medium(O):
    c = synth.generate(96, 15000, seed=31, n_family_pairs=16, missing=0.02, inbred_fraction=0.05)
    co = Cohort(c.codes)
    af, kept = co.allele_freqs()
    codes, afk = c.codes[:, kept], af[kept]
    yield c, co, codes, afk
    co.close()


def test_allele_freqs_bitexact(O, medium):
    c, co, _, _ = medium
    say("case: allele freqs 96 x 15000")
    af, kept = co.allele_freqs()
    exp = O.allele_freqs(c.codes)
    assert np.array_equal(af, exp)
    assert np.array_equal(kept, O.keep_sites(exp))


def test_evaluate_matches_oracle(O, medium):
    """The operator boundary (objective + gradient + Hessian at given points)."""
    c, co, codes, afk = medium
    say("case: evaluate 300 points")
    rng = np.random.default_rng(0)
    pairs = rng.integers(0, codes.shape[0], (300, 2))
    pts = rng.dirichlet(np.ones(9), 300)
    pts[:20] = np.eye(9)[rng.integers(0, 9, 20)] * 0.5 + 0.5 / 9  # points with large 1/L spread
    f, g, h, ns = co.evaluate(pairs, pts)
    co.set_option("class_sweep", 0)
    ft, gt, ht, nt = co.evaluate(pairs, pts)
    co.set_option("class_sweep", 1)
    seg = seg_for(co.n_kept, codes.shape[1])
    for i in range(300):
        rows, valid, n = O.pair_rows_codes(codes[pairs[i, 0]], codes[pairs[i, 1]], afk)
        assert ns[i] == n and nt[i] == n
        fe, ge, he, _ = class_eval(O, codes[pairs[i, 0]], codes[pairs[i, 1]], afk, pts[i])
        assert f[i] == fe and np.array_equal(g[i], ge) and np.array_equal(h[i], he), i
        fe = O.obj_seg(rows, valid, seg, pts[i])
        ge, he = O.gh_seg(rows, valid, seg, pts[i])
        assert ft[i] == fe and np.array_equal(gt[i], ge) and np.array_equal(ht[i], he), i
        if i < 40:
            comp = rows[valid == 1]
            fr = O.obj_ref(comp, pts[i])
            gr, hr = O.gh_ref(comp, pts[i])
            assert abs(f[i] - fr) <= 1e-13 * abs(fr)
            np.testing.assert_allclose(g[i], gr, rtol=1e-12, atol=1e-13 * np.abs(gr).max())
            np.testing.assert_allclose(h[i], hr, rtol=1e-12, atol=1e-13 * np.abs(hr).max())


def test_fit_all_pairs_bitexact_and_tolerance(O, medium):
    c, co, codes, afk = medium
    say("case: fit 96 x 15000 all 4656 pairs")
    out = co.fit_pairs()
    co.set_option("class_sweep", 0)
    table = co.fit_pairs()
    co.set_option("class_sweep", 1)
    assert (out["status"] == 0).all() and (table["status"] == 0).all()
    ref = O.fit_pairs(codes, afk, order=O.ORDER_REF)
    assert_bitexact(out, class_fit(O, codes, afk))
    assert_within_tolerance(out, ref)
    seg = seg_for(co.n_kept, codes.shape[1])
    assert_bitexact(table, O.fit_pairs(codes, afk, order=O.ORDER_SEG, seg=seg))
    assert_within_tolerance(table, ref)


def test_schedule_independence(O, medium):
    """Results do not depend on how many pairs are in flight or their order."""
    c, co, codes, afk = medium
    say("case: schedule independence")
    rng = np.random.default_rng(1)
    pairs = rng.integers(0, codes.shape[0], (700, 2))
    a = co.fit_pairs(pairs)
    co.set_option("slots", 256)
    b = co.fit_pairs(pairs)
    perm = rng.permutation(700)
    co.set_option("slots", 512)
    cc = co.fit_pairs(pairs[perm])
    co.set_option("slots", 0)
    assert_bitexact(a, b)
    assert_bitexact(a[perm], cc)
    # two slot groups on two streams (used when there are >= 4 x slots pairs)
    co.set_option("slots", 256)
    many = rng.integers(0, codes.shape[0], (3000, 2))
    two = co.fit_pairs(many)
    assert co.last_fit_stats()["groups"] == 2
    co.set_option("streams", 1)
    one = co.fit_pairs(many)
    assert co.last_fit_stats()["groups"] == 1
    co.set_option("streams", 2)
    co.set_option("slots", 0)
    assert_bitexact(two, one)
    # iterations queued per host synchronisation
    for batch in (1, 3, 16):
        co.set_option("batch", batch)
        assert_bitexact(co.fit_pairs(many), two)
    # the batch captured once per fit into a HIP graph and relaunched: every
    # fit (many relaunches each) equals the eager one, with eager fits between
    co.set_option("graph", 1)
    for batch in (8, 3):
        co.set_option("batch", batch)
        assert_bitexact(co.fit_pairs(many), two)
        co.set_option("graph", 0)
        assert_bitexact(co.fit_pairs(many[:700]), two[:700])
        co.set_option("graph", 1)
    co.set_option("slots", 0)
    assert_bitexact(co.fit_pairs(many), two)
    co.set_option("graph", 0)
    co.set_option("batch", 8)


def test_autotune_sets_a_shape_that_keeps_results(medium):
    """kestrel_autotune tries launch shapes on a sample, keeps the fastest;
    the fit afterwards is bit-identical to the default shape's."""
    c, co, codes, afk = medium
    say("case: autotune")
    rng = np.random.default_rng(11)
    many = rng.integers(0, codes.shape[0], (4000, 2))
    before = co.fit_pairs(many)
    r = co.autotune(sample=30000)
    assert r["trials"] >= 6 and r["pairs_per_s"] > 0 and r["default_pairs_per_s"] > 0
    assert r["streams"] in (1, 2) and r["batch"] in (4, 8, 16)
    assert_bitexact(co.fit_pairs(many), before)
    co.set_option("slots", 0)
    co.set_option("streams", 2)
    co.set_option("batch", 8)
    with pytest.raises(KestrelError):
        co.autotune(pairs=np.zeros((10, 2), np.uint32))  # fewer than 256 sample pairs


def test_pair_order_and_subset(O, medium):
    """(y, x) fits the swapped problem (lib.rs:137-149 swaps delta3<->delta5,
    delta4<->delta6); n_select restricts to the first samples (cli.rs:41-45)."""
    c, co, codes, afk = medium
    say("case: pair order and subset")
    pairs = np.array([[3, 17], [17, 3], [5, 5]])
    out = co.fit_pairs(pairs)
    swap = [0, 1, 4, 5, 2, 3, 6, 7, 8]
    np.testing.assert_allclose(out[1]["jacquard"], out[0]["jacquard"][swap], atol=1e-6)
    sub = co.fit_pairs(n_select=10)
    assert len(sub) == 55 and sub["x"].max() < 10
    full = co.fit_pairs(n_select=12)
    idx = {(int(o["x"]), int(o["y"])): o for o in full}
    for o in sub:
        e = idx[(int(o["x"]), int(o["y"]))]
        assert np.array_equal(o["jacquard"], e["jacquard"]) and o["iters"] == e["iters"]
    jac, kin, inb = build_arrays(sub, 10)
    assert np.array_equal(jac[3, 7][swap], jac[7, 3]) and kin[3, 7] == kin[7, 3]
    with pytest.raises(KestrelError):
        co.fit_pairs(n_select=97)


def test_iteration_cap(O, medium):
    c, co, codes, afk = medium
    say("case: iteration cap 1 and 2")
    for it in (1, 2):
        out = co.fit_pairs(np.array([[0, 1], [2, 9], [4, 4]]), iterations=it)
        exp = class_fit(O, codes, afk, pairs=np.array([[0, 1], [2, 9], [4, 4]]), iterations=it)
        assert_bitexact(out, exp)
        assert (out["iters"] <= it).all()


def test_edge_cases(O):
    say("case: edge cases")
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 3, (12, 700)).astype(np.uint8)
    codes[:, :50] = 0          # monomorphic sites are dropped
    codes[5, :] = 3            # a sample with no call: its pairs have no usable site
    codes[6, ::2] = 3          # ragged missingness
    with Cohort(codes) as co:
        af, kept = co.allele_freqs()
        assert not kept[:50].any()
        out = co.fit_pairs()
        exp_af = O.allele_freqs(codes)
        assert np.array_equal(af, exp_af)
        bad = (out["x"] == 5) | (out["y"] == 5)
        assert (out["status"][bad] == 1).all() and (out["n_sites"][bad] == 0).all()
        assert np.isnan(out["jacquard"][bad]).all()
        assert (out["status"][~bad] == 0).all()
        good = out[~bad]
        kc, ka = codes[:, kept], af[kept]
        exp = class_fit(O, kc, ka, pairs=np.stack([good["x"], good["y"]], 1))
        assert_bitexact(good, exp)
    # all samples missing at one site: the reference panics; the library reports it
    codes2 = codes.copy()
    codes2[:, 60] = 3
    with pytest.raises(KestrelError):
        Cohort(codes2)
    # a cohort with no segregating site: every pair has no usable site
    with Cohort(np.zeros((4, 40), np.uint8)) as co:
        assert co.n_kept == 0
        out = co.fit_pairs()
        assert (out["status"] == 1).all()


def test_minimum_cohort(O):
    """Two samples, one segregating site."""
    say("case: 2 samples x 1 site")
    codes = np.array([[0], [2]], np.uint8)  # p = 1/2 (a site at exactly 1/(2N) is a threshold tie)
    with Cohort(codes) as co:
        af, kept = co.allele_freqs()
        assert co.n_kept == 1 and np.array_equal(kept, O.keep_sites(O.allele_freqs(codes)))
        out = co.fit_pairs()
    exp = class_fit(O, codes, af)
    assert len(out) == 3 and (out["n_sites"] == 1).all()
    assert_bitexact(out, exp)


@pytest.mark.parametrize("v", [255, 256, 257, 4095, 4096, 4097, 8191, 8192, 8193, 16385])
def test_padding_and_segment_boundaries(O, v):
    """Kept-site counts either side of the 256-site row padding, the table
    sweep's 4,096-site segments and the class-sorted sweep's 8,192-site
    blocks, in both forms."""
    say(f"case: 24 samples x {v} kept sites")
    c = synth.generate(24, v + v // 5 + 400, seed=40 + v, n_family_pairs=4, missing=0.03, maf=0.05)
    seg_sites = np.flatnonzero(O.keep_sites(O.allele_freqs(c.codes)))[:v]
    assert len(seg_sites) == v
    codes = np.ascontiguousarray(c.codes[:, seg_sites])
    with Cohort(codes) as co:
        af, kept = co.allele_freqs()
        assert kept.all() and co.n_kept == v
        out = co.fit_pairs()
        co.set_option("class_sweep", 0)
        table = co.fit_pairs()
    assert_bitexact(out, class_fit(O, codes, af))
    assert_bitexact(table, O.fit_pairs(codes, af, order=O.ORDER_SEG, seg=seg_for(v, v)))


def test_many_small_segments(O):
    say("case: segments of 32 sites over 1,000 sites")
    c = synth.generate(20, 1000, seed=44, n_family_pairs=4, missing=0.02)
    with Cohort(c.codes, segment_sites=32) as co:
        co.set_option("class_sweep", 0)  # segments are the table sweep's
        af, kept = co.allele_freqs()
        out = co.fit_pairs()
        n = co.n_kept
    kc = c.codes[:, kept]
    seg = np.minimum(segment_bounds(n, 32), kc.shape[1]).astype(np.uint32)
    assert len(seg) - 1 >= 20
    assert_bitexact(out, O.fit_pairs(kc, af[kept], order=O.ORDER_SEG, seg=seg))


def test_c3_full_size_properties(O):
    """The full C3 job (4,000 x 50,000, per-sample missingness U(0, 0.3), 8.0 M
    pairs): size-independent properties of every record, n_xy against counts
    of co-called sites, designated relatives, and a seeded sample of pairs
    bit-exact against the oracle."""
    say("case: C3 4000 x 50000 all 8,002,000 pairs")
    c = synth.config("c3")
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        out = co.fit_pairs()
    n = c.codes.shape[0]
    assert len(out) == n * (n + 1) // 2
    assert (out["status"] == 0).all()
    d = out["jacquard"]
    assert (d >= 0).all() and np.abs(d.sum(1) - 1).max() < 1e-12
    rng = np.random.default_rng(9)
    idx = rng.integers(0, len(out), 200)
    kc = c.codes[:, kept]
    called = kc != 3
    for i in idx[:40]:
        x, y = int(out["x"][i]), int(out["y"][i])
        assert out["n_sites"][i] == int((called[x] & called[y]).sum())
    pairs = np.stack([out["x"][idx], out["y"][idx]], 1)
    exp = class_fit(O, kc, af[kept], pairs=pairs, nthreads=16)
    assert_bitexact(out[idx], exp)
    expect = {"PO": 0.25, "FS": 0.25, "HS": 0.125, "FC": 0.0625}
    pos = lambda x, y: x * n - x * (x - 1) // 2 + (y - x)  # noqa: E731  row-major upper triangle
    for x, y, kind in c.pairs:
        a, b = min(x, y), max(x, y)
        k = kinship(out["jacquard"][pos(a, b):pos(a, b) + 1])[0]
        assert abs(k - expect[kind]) < 0.035, (kind, k)


def test_claims_in_runs_and_skipped_masked_sites(O):
    """Cohorts with many missing calls fit with claims in runs (a 256-slot
    workgroup's lanes claim consecutive pairs, mostly of one first sample) and
    sweeps that pass over the sites where that shared sample is missing.  The
    records are bit-identical to per-lane claims with full sweeps, for any
    number of pairs in flight and for a pair list whose runs mix samples, and
    to the oracle."""
    say("case: claims in runs, masked sites skipped (missing U(0, 0.5) per sample)")
    c = synth.generate(260, 5000, seed=77, missing_per_sample_max=0.5, n_family_pairs=8)
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, 260, (2500, 2))
    with Cohort(c.codes) as co:
        co.set_option("class_sweep", 0)  # claims in runs are the table sweep's
        af, kept = co.allele_freqs()
        auto = co.fit_pairs()  # missing calls >= 2 %: on
        co.set_option("skip_masked", 0)
        off = co.fit_pairs()
        off_list = co.fit_pairs(pairs)
        co.set_option("skip_masked", 1)
        on_list = co.fit_pairs(pairs)
        co.set_option("slots", 256)
        few = co.fit_pairs()
        co.set_option("streams", 1)
        one = co.fit_pairs()
    assert_bitexact(auto, off)
    assert_bitexact(few, off)
    assert_bitexact(one, off)
    assert_bitexact(on_list, off_list)
    kc = c.codes[:, kept]
    idx = rng.integers(0, len(auto), 300)
    sample = np.stack([auto["x"][idx], auto["y"][idx]], 1)
    exp = O.fit_pairs(kc, af[kept], pairs=sample, order=O.ORDER_SEG, seg=seg_for(int(kept.sum()), kc.shape[1]))
    assert_bitexact(auto[idx], exp)


def test_mixed_precision_form():
    """KESTREL_MIXED (FP32 table and Hessian terms; L, gradient, objective and
    solver FP64): the frequencies, kept sites and n_xy are the FP64 form's;
    Delta is within the form's bar (1e-6) of the FP64 form on the pairs both
    forms bring to convergence with log10 condition below 12, and so is the
    relative objective."""
    say("case: mixed-precision form vs FP64 form, 160 x 20000")
    c = synth.generate(160, 20000, seed=41, n_family_pairs=8)
    with Cohort(c.codes, precision="mixed") as a, Cohort(c.codes) as b:
        afa, ka = a.allele_freqs()
        afb, kb = b.allele_freqs()
        assert np.array_equal(afa, afb) and np.array_equal(ka, kb)
        oa, ob = a.fit_pairs(), b.fit_pairs()
    assert np.array_equal(oa["x"], ob["x"]) and np.array_equal(oa["n_sites"], ob["n_sites"])
    lb, la = np.log10(ob["cond"]), np.log10(oa["cond"])
    both = np.isfinite(lb) & np.isfinite(la) & (lb < 12)
    assert both.sum() > 0.9 * len(ob)
    d = np.abs(oa["jacquard"] - ob["jacquard"]).max(1)[both]
    say(f"mixed vs fp64: max |dDelta| {d.max():.3g} over {both.sum()} pairs; iterations "
        f"{oa['iters'].mean():.2f} vs {ob['iters'].mean():.2f}")
    assert d.max() <= 1e-6
    assert (np.abs(oa["obj"] - ob["obj"]) / np.abs(ob["obj"]))[both].max() <= 1e-6
    with pytest.raises(ValueError):
        Cohort(c.codes, precision="fp16")


def test_mixed_form_with_claims_in_runs_and_skipped_masked_sites():
    """The mixed-precision form on a cohort with many missing calls: claims in
    runs and the pass over the wavefront's masked sites (option skip_masked 1;
    off by default for this form) give the records of per-lane claims with
    full sweeps bit for bit, for the whole triangle and for a pair list."""
    say("case: mixed form, claims in runs, masked sites skipped (missing U(0, 0.5))")
    c = synth.generate(260, 5000, seed=78, missing_per_sample_max=0.5, n_family_pairs=8)
    rng = np.random.default_rng(6)
    pairs = rng.integers(0, 260, (2500, 2))
    with Cohort(c.codes, precision="mixed") as co:
        auto = co.fit_pairs()
        co.set_option("skip_masked", 1)
        on = co.fit_pairs()
        on_list = co.fit_pairs(pairs)
        co.set_option("skip_masked", 0)
        off = co.fit_pairs()
        off_list = co.fit_pairs(pairs)
    assert_bitexact(on, off)
    assert_bitexact(auto, off)
    assert_bitexact(on_list, off_list)
    assert np.isfinite(on["jacquard"]).all()


def test_given_allele_frequencies(O):
    say("case: given allele frequencies")
    rng = np.random.default_rng(4)
    codes = rng.integers(0, 3, (20, 900)).astype(np.uint8)
    p = rng.uniform(0.05, 0.95, 900)
    af = np.stack([p, 1 - p], 1)
    af[:7, 1] = 0.0
    af[:7, 0] = 1.0
    with Cohort(codes, af=af) as co:
        got, kept = co.allele_freqs()
        assert np.array_equal(got, af) and not kept[:7].any() and kept[7:].all()
        out = co.fit_pairs()
    kc, ka = codes[:, kept], af[kept]
    assert_bitexact(out, class_fit(O, kc, ka))


def test_c2_sample_of_pairs(O):
    """Full-size cohort (1,000 x 100,000): a seeded sample of pairs against the
    oracle, and size-independent properties of every designated relative."""
    say("case: C2 1000 x 100000 sampled pairs")
    c = synth.config("c2")
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        exp_af = O.allele_freqs(c.codes)
        assert np.array_equal(af, exp_af)
        rng = np.random.default_rng(5)
        pairs = np.concatenate([rng.integers(0, 1000, (120, 2)), np.array([(x, y) for x, y, _ in c.pairs])])
        out = co.fit_pairs(pairs)
    codes, afk = c.codes[:, kept], af[kept]
    assert_bitexact(out, class_fit(O, codes, afk, pairs=pairs, nthreads=16))
    assert_within_tolerance(out, O.fit_pairs(codes, afk, pairs=pairs, order=O.ORDER_REF, nthreads=16))
    s = out["jacquard"].sum(1)
    assert np.abs(s - 1).max() < 1e-12 and (out["jacquard"] >= 0).all()
    expect = {"PO": 0.25, "FS": 0.25, "HS": 0.125, "FC": 0.0625}
    kin = kinship(out["jacquard"][120:])
    for (x, y, kind), k in zip(c.pairs, kin):
        assert abs(k - expect[kind]) < 0.03, (kind, k)


def test_c2_whole_job_sample(O):
    """The whole C2 job (500,500 records in the default schedule: two slot
    groups, claims as slots free up): a seeded sample of 2,000 records
    bit-exact against the oracle, every record on the simplex."""
    say("case: C2 1000 x 100000, all 500,500 records, 2,000 checked against the oracle")
    c = synth.config("c2")
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        out = co.fit_pairs()
    n = c.codes.shape[0]
    assert len(out) == n * (n + 1) // 2 and (out["status"] == 0).all()
    s = out["jacquard"].sum(1)
    assert np.abs(s - 1).max() < 1e-12 and (out["jacquard"] >= 0).all()
    idx = np.random.default_rng(11).choice(len(out), 2000, replace=False)
    pairs = np.stack([out["x"][idx], out["y"][idx]], 1)
    codes, afk = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    assert_bitexact(out[idx], class_fit(O, codes, afk, pairs=pairs, nthreads=16))


def test_device_arithmetic_selftest():
    """recip_rn == IEEE 1/x and the positive-domain log == full log, bit for bit."""
    import ctypes
    from kestrel_amd import _lib
    say("case: arithmetic self-test 2^30 inputs")
    bad = ctypes.c_uint64()
    _lib.check(_lib.load().kestrel_selftest(1 << 30, 12345, ctypes.byref(bad)))
    assert bad.value == 0
