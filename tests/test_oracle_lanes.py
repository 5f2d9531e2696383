"""The oracle's wavefront order (ORDER_LANES: each segment summed by 64
strided lanes, then a butterfly) against its segmented order (ORDER_SEG) and
the reference's order (ORDER_REF).  It is the same function summed in another
order, so: where only one lane of a segment holds usable sites the gradient
and Hessian sums are the segmented order's bit for bit (the same terms in the
same order, plus exact zeros from the other lanes); elsewhere every output
stays within a few units of rounding of both other orders, and fits reach the
same records within the reference tolerances."""
import numpy as np

from kestrel_amd import segment_bounds, synth


def rows_of(O, c, x, y):
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    kc, ka = c.codes[:, kept], af[kept]
    rows, valid, _ = O.pair_rows_codes(kc[x], kc[y], ka)
    return rows, valid, kc, ka


def point(seed):
    p = np.random.default_rng(seed).dirichlet(np.ones(9))
    return p


def test_one_lane_segments_match_segmented_order_bitwise(O):
    c = synth.generate(8, 2000, seed=5, missing=0.0)
    rows, valid, _, _ = rows_of(O, c, 1, 4)
    seg = np.array([0, 256, 640, 1024, rows.shape[0]], np.uint32)
    for lane in (0, 5, 63):
        v = np.zeros_like(valid)
        for k in range(len(seg) - 1):
            s = np.arange(seg[k], seg[k + 1])
            v[s[(s - seg[k]) % 64 == lane]] = valid[s[(s - seg[k]) % 64 == lane]]
        assert v.sum() > 10
        x = point(lane)
        g1, h1 = O.gh_lanes(rows, v, seg, x)
        g2, h2 = O.gh_seg(rows, v, seg, x)
        assert np.array_equal(g1, g2) and np.array_equal(h1, h2), lane


def test_close_to_segmented_and_reference_orders(O):
    c = synth.generate(12, 5000, seed=9, missing=0.05, n_family_pairs=2)
    rows, valid, _, _ = rows_of(O, c, 2, 7)
    n = rows.shape[0]
    for seg in (segment_bounds(n), np.array([0, n], np.uint32), np.arange(0, n + 96, 96).clip(0, n).astype(np.uint32)):
        seg = np.unique(np.minimum(seg, n)).astype(np.uint32)
        for s in range(5):
            x = point(100 + s)
            fl, fs = O.obj_lanes(rows, valid, seg, x), O.obj_seg(rows, valid, seg, x)
            fr = O.obj_ref(rows[valid == 1], x)
            assert abs(fl - fs) <= 1e-13 * abs(fs) and abs(fl - fr) <= 1e-13 * abs(fr)
            (gl, hl), (gs, hs) = O.gh_lanes(rows, valid, seg, x), O.gh_seg(rows, valid, seg, x)
            assert np.allclose(gl, gs, rtol=1e-12, atol=0) and np.allclose(hl, hs, rtol=1e-12, atol=0)
            # a different order: not the same bits everywhere
            if s == 0 and len(seg) == 2:
                assert not (np.array_equal(gl, gs) and np.array_equal(hl, hs))


def test_fits_match_other_orders(O):
    c = synth.generate(16, 3000, seed=21, missing=0.03, n_family_pairs=3)
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    seg = np.minimum(segment_bounds(kc.shape[1], segment_sites=1024), kc.shape[1]).astype(np.uint32)
    a = O.fit_pairs(kc, ka, order=O.ORDER_LANES, seg=seg)
    b = O.fit_pairs(kc, ka, order=O.ORDER_SEG, seg=seg)
    r = O.fit_pairs(kc, ka, order=O.ORDER_REF)
    assert np.array_equal(a["n_sites"], b["n_sites"])
    for other in (b, r):
        assert np.abs(a["jacquard"] - other["jacquard"]).max() <= 1e-6
        assert (np.abs(a["obj"] - other["obj"]) <= 1e-9 * np.abs(other["obj"])).all()
    assert (a["iters"] == b["iters"]).mean() > 0.9
