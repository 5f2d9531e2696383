"""Tri/tetra-allelic sites on the device (kestrel_cohort_create_gl_multi,
sweep_glm) against the oracle: input-bucket allele frequencies and the
re-bucketed part sizes bit-exact; pair records bit-exact in the device's
summation order (2-allele sites, then 3, then 4, each padded to 256 and
segmented); the reference's order within the stated tolerance."""
import numpy as np
import pytest

from kestrel_amd import Cohort, relatedness_coefficients_gl, segment_bounds, synth

pytestmark = pytest.mark.gpu

N = 20


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(out, exp):
    for k in ("x", "y", "jacquard", "obj", "iters", "n_sites", "cond"):
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} pairs differ"


def device_order(parts, segment_sites=4096):
    """Pads and segment bounds of the device's parts (each part's sites
    rounded up to 256, the 2-allele part at least 256; a 3- or 4-allele part
    cut into as many segments as the 2-allele part, in whole 32-site chunks,
    at most segment_sites each)."""
    pads, seg, base, main_nseg = {}, [0], 0, 1
    for A in (2, 3, 4):
        n = parts[A][0].shape[0]
        if A > 2 and n == 0:
            continue
        if A == 2:
            b = segment_bounds(n, segment_sites)
            main_nseg = len(b) - 1
        else:
            v_pad = (n + 255) // 256 * 256
            per = -(-v_pad // main_nseg)
            b = segment_bounds(n, min(segment_sites, max(32, -(-per // 32) * 32)))
        pads[A] = int(b[-1])
        seg += [int(v) + base for v in b[1:]]
        base += int(b[-1])
    return pads, np.array(seg, np.uint32)


@pytest.fixture(scope="module")
def multi(O):
    b2, b3, b4 = synth.multi_likelihoods(N, (900, 260, 120), seed=51)
    co = Cohort.gl_multi(b2, b3, b4)
    parts, af_in = O.multi_prepare({2: b2, 3: b3, 4: b4})
    yield co, parts, af_in, (b2, b3, b4)
    co.close()


def test_multi_frequencies_and_parts(multi):
    co, parts, af_in, _ = multi
    say("case: multi-allelic frequencies and re-bucketing")
    counts, af = co.multi_info()
    assert counts == tuple(parts[A][0].shape[0] for A in (2, 3, 4))
    assert counts[0] > 900 and counts[1] > 0 and counts[2] > 0  # 3/4-allele sites moved down
    for A in (2, 3, 4):
        assert np.array_equal(af[A], af_in[A]), A


def test_multi_fit_bitexact_and_tolerance(O, multi):
    co, parts, af_in, _ = multi
    say(f"case: multi-allelic fit, {N * (N + 1) // 2} pairs")
    out = co.fit_pairs()
    pairs = np.stack([out["x"], out["y"]], 1)
    pads, seg = device_order(parts)
    assert_bitexact(out, O.multi_fit_pairs(parts, pairs, order=O.ORDER_SEG, pads=pads, seg=seg))
    ref = O.multi_fit_pairs(parts, pairs, order=O.ORDER_REF)
    lc = np.log10(ref["cond"])
    well = np.isfinite(lc) & (lc < 12)
    assert well.sum() > 100
    assert np.abs(out["jacquard"] - ref["jacquard"]).max(1)[well].max() <= 1e-6
    assert (np.abs(out["obj"] - ref["obj"]) / np.abs(ref["obj"]))[well].max() <= 1e-9
    say("case: multi-allelic fit with the batch in a HIP graph")
    co.set_option("graph", 1)
    try:
        assert_bitexact(co.fit_pairs(), out)
    finally:
        co.set_option("graph", 0)


def test_multi_parts_cut_like_the_main_part(O, multi):
    """With several main-part segments the 3- and 4-allele parts are cut into
    as many (whole 32-site chunks), and the records follow that order."""
    co, parts, af_in, _ = multi
    say("case: multi-allelic fit with 256-site segments")
    co.set_option("segment_sites", 256)
    try:
        out = co.fit_pairs()
    finally:
        co.set_option("segment_sites", 4096)
    pads, seg = device_order(parts, 256)
    assert len(seg) - 1 >= 8
    pairs = np.stack([out["x"], out["y"]], 1)
    assert_bitexact(out, O.multi_fit_pairs(parts, pairs, order=O.ORDER_SEG, pads=pads, seg=seg))


@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_multi_biallelic_rows_compact_and_full(O, scale):
    """The 2-allele part is held as 16-byte rows when every called row has an
    exact 1.0 (scale 1: rows scaled by their maximum) and as 24-byte rows when
    some has none (scale 0.5); both give the oracle's records bit for bit."""
    say(f"case: multi-allelic fit, 2-allele rows scaled by {scale}")
    b2, b3, b4 = synth.multi_likelihoods(14, (300, 90, 40), seed=53, absent=0.0)
    b2 = b2 * scale
    with Cohort.gl_multi(b2, b3, b4) as co:
        out = co.fit_pairs()
    parts, _ = O.multi_prepare({2: b2, 3: b3, 4: b4})
    pairs = np.stack([out["x"], out["y"]], 1)
    pads, seg = device_order(parts)
    assert_bitexact(out, O.multi_fit_pairs(parts, pairs, order=O.ORDER_SEG, pads=pads, seg=seg))


def test_multi_evaluate_bitexact(O, multi):
    co, parts, af_in, _ = multi
    say("case: multi-allelic evaluate")
    rng = np.random.default_rng(2)
    pairs = rng.integers(0, N, (40, 2))
    pts = rng.dirichlet(np.ones(9), 40)
    f, g, h, ns = co.evaluate(pairs, pts)
    pads, seg = device_order(parts)
    for i in range(40):
        rows, valid = O.multi_pair_rows(parts, int(pairs[i, 0]), int(pairs[i, 1]), O.ORDER_SEG, pads)
        assert ns[i] == valid.sum()
        assert f[i] == O.obj_seg(rows, valid, seg, pts[i])
        ge, he = O.gh_seg(rows, valid, seg, pts[i])
        assert np.array_equal(g[i], ge) and np.array_equal(h[i], he)


def test_biallelic_bucket_alone_equals_create_gl(multi):
    co, parts, af_in, (b2, _, _) = multi
    say("case: gl_multi with only 2-allele sites vs create_gl")
    with Cohort.gl_multi(b2) as a, Cohort.gl(b2) as b:
        assert_bitexact(a.fit_pairs(), b.fit_pairs())


def test_relatedness_coefficients_gl_tetra_only(O):
    say("case: relatedness_coefficients_gl on 4-allele sites")
    _, _, b4 = synth.multi_likelihoods(12, (0, 0, 400), seed=52)
    jac, kin, inb = relatedness_coefficients_gl(b4)
    parts, _ = O.multi_prepare({4: b4})
    tri = [(x, y) for x in range(12) for y in range(x, 12)]
    pads, seg = device_order(parts)
    exp = O.multi_fit_pairs(parts, tri, order=O.ORDER_SEG, pads=pads, seg=seg)
    from kestrel_amd import build_arrays
    ej, ek, ei = build_arrays(exp, 12)
    assert np.array_equal(jac, ej) and np.array_equal(kin, ek) and np.array_equal(inb, ei)


def test_multi_mixed_precision_form(O):
    """The mixed-precision form on 2/3/4-allele likelihood rows (option
    "precision"; sweep_gl_mixed + sweep_glm_mixed with FP32 coefficient
    tables): Delta, kinship and objective within 1e-6 of the FP64 form on the
    pairs both bring to convergence with log10 condition below 12, n_xy equal,
    log10 condition within 0.1 below 100; the option switches back to
    bit-exact FP64 records."""
    from kestrel_amd import _lib, kinship
    say("case: multi-allelic fit, mixed-precision form vs FP64 form, 40 x (1500, 400, 160)")
    b2, b3, b4 = synth.multi_likelihoods(40, (1500, 400, 160), seed=57)
    with Cohort.gl_multi(b2, b3, b4) as co:
        b = co.fit_pairs()
        co.set_option("precision", _lib.MIXED)
        a = co.fit_pairs()
        co.set_option("precision", _lib.FP64)
        again = co.fit_pairs()
    assert_bitexact(again, b)
    assert np.array_equal(a["n_sites"], b["n_sites"])
    lb, la = np.log10(b["cond"]), np.log10(a["cond"])
    both = np.isfinite(lb) & np.isfinite(la) & (lb < 12)
    assert both.sum() > 0.9 * len(b)
    d = np.abs(a["jacquard"] - b["jacquard"]).max(1)[both]
    k = np.abs(kinship(a["jacquard"]) - kinship(b["jacquard"]))[both]
    r = (np.abs(a["obj"] - b["obj"]) / np.abs(b["obj"]))[both]
    say(f"multi mixed vs fp64: max |dDelta| {d.max():.3g}, kinship {k.max():.3g}, objective {r.max():.3g} over "
        f"{both.sum()} pairs; iterations {a['iters'].mean():.2f} vs {b['iters'].mean():.2f}")
    assert d.max() <= 1e-6 and k.max() <= 1e-6 and r.max() <= 1e-6
    c = both & (lb < 2.0)
    assert np.abs(la - lb)[c].max() <= 0.1


def test_multi_onehot_class_form(O):
    """Hard calls at 2-, 3- and 4-allele sites as one-hot rows (minimum allele
    frequency 0, so the re-bucketed 2-allele rows stay one-hot): the 2-allele
    part is held as 2-bit codes and swept in the class-sorted order, the 3- and
    4-allele parts in their segments after it.  Records and evaluations
    bit-exact against the oracle's class order with parts; the site-table
    form (option "class_sweep" 0) against the segmented order; both within
    the stated tolerance of the reference's order."""
    say("case: one-hot multi-allelic cohort in the class form, 48 x (1500, 300, 120)")
    b2, b3, b4 = synth.multi_onehot(synth.multi_genotypes(48, (1500, 300, 120), seed=71, missing=0.02))
    parts, _ = O.multi_prepare({2: b2, 3: b3, 4: b4}, maf=0.0)
    rng = np.random.default_rng(4)
    with Cohort.gl_multi(b2, b3, b4, maf=0.0) as co:
        out = co.fit_pairs()
        assert co.last_fit_stats()["form"] == "class"
        pairs = np.stack([out["x"], out["y"]], 1)
        pads, seg = device_order(parts)
        sel = np.unique(np.concatenate([rng.choice(len(out), 240, replace=False),
                                        np.nonzero(out["x"] == out["y"])[0][:8]]))
        assert_bitexact(out[sel], O.multi_class_fit_pairs(parts, pairs[sel], pads, seg))
        ref = O.multi_fit_pairs(parts, pairs[sel], order=O.ORDER_REF)
        lc = np.log10(ref["cond"])
        well = np.isfinite(lc) & (lc < 12)
        assert well.sum() > 150
        assert np.abs(out["jacquard"][sel] - ref["jacquard"]).max(1)[well].max() <= 1e-6
        assert (np.abs(out["obj"][sel] - ref["obj"]) / np.abs(ref["obj"]))[well].max() <= 1e-9
        # evaluations in the same order
        ep = rng.integers(0, 48, (24, 2))
        pts = rng.dirichlet(np.ones(9), 24)
        f, g, h, ns = co.evaluate(ep, pts)
        for i in range(24):
            fe, ge, he = O.multi_class_eval(parts, ep[i, 0], ep[i, 1], pads, seg, pts[i])
            assert f[i] == fe and np.array_equal(g[i], ge) and np.array_equal(h[i], he), i
        # the site-table form of the same cohort
        co.set_option("class_sweep", 0)
        t = co.fit_pairs()
        assert co.last_fit_stats()["form"] == "table"
        assert_bitexact(t[sel], O.multi_fit_pairs(parts, pairs[sel], order=O.ORDER_SEG, pads=pads, seg=seg))


def test_multi_onehot_block_shapes_agree(O):
    """The hard-call parts' sweep runs blocks of 512 pairs when they give
    every CU one, and blocks of 256 otherwise; both shapes give the same bits.
    140,000 evaluation points on one segment (274 blocks of 512, more than
    the 256 CUs of an MI355X) against 40 of them evaluated alone (one block of
    256), a few of which are checked against the oracle."""
    say("case: hard-call parts in blocks of 512 and of 256, 48 x (1500, 300, 120), 140000 points")
    b2, b3, b4 = synth.multi_onehot(synth.multi_genotypes(48, (1500, 300, 120), seed=71, missing=0.02))
    parts, _ = O.multi_prepare({2: b2, 3: b3, 4: b4}, maf=0.0)
    pads, seg = device_order(parts)
    rng = np.random.default_rng(6)
    n = 140_000
    ep = rng.integers(0, 48, (n, 2))
    pts = rng.dirichlet(np.ones(9), n)
    idx = rng.choice(n, 40, replace=False)
    with Cohort.gl_multi(b2, b3, b4, maf=0.0) as co:
        big = co.evaluate(ep, pts)
        small = co.evaluate(ep[idx], pts[idx])
    for a, b in zip(small, big):
        assert same(a, b[idx]).all()
    for i in range(4):
        fe, ge, he = O.multi_class_eval(parts, ep[idx[i], 0], ep[idx[i], 1], pads, seg, pts[idx[i]])
        assert small[0][i] == fe and np.array_equal(small[1][i], ge) and np.array_equal(small[2][i], he), i
