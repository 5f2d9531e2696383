"""The operator boundary (kestrel_evaluate, the reference's Objective /
GradHess, sqp.rs:8-14) on the whole domain the reference accepts: any point,
any eps.  Points outside the sweeps' domain (eps = 0, negative or huge
coefficients, L = 0, negative or subnormal L) are evaluated with IEEE
division and per-site logs where a 16-site group's L are not all positive
normal numbers; f, g and H are bit-exact against the oracle in the device's
segmented order (special values included), and the objective's special values
are those of the reference's own order (L = 0 -> +inf, L < 0 -> NaN).  With
the class-sorted sweep (the default for 2-bit cohorts) points inside the
domain take its order and the others the same exact evaluation.  The pair fit
rejects an eps outside [1e-18, 1e14]."""
import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, segment_bounds, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return ((a == b) | (np.isnan(a) & np.isnan(b))).all()


def domain_points():
    """(point, eps) cases: each leaves the sweeps' domain in its own way."""
    e = np.eye(9)
    cases = []
    for k in range(9):
        cases.append((e[k], 0.0))            # eps = 0 at a vertex: L = 0 where c_k = 0
    cases.append((np.full(9, 1 / 9), 0.0))   # eps = 0, every L > 0
    cases.append((e[8], 1e-310))              # subnormal L where c_8 ... = 0
    cases.append((np.full(9, 1e200), 1e-8))   # products overflow: per-site logs
    cases.append((np.full(9, 1e-200), 1e-300))  # products underflow: per-site logs
    p = np.full(9, 1 / 9)
    p[0] = -0.5                                # negative coefficients: some L < 0
    cases.append((p, 1e-8))
    p = np.full(9, 1 / 9)
    p[8] = -2.0                                # every L < 0: pairs of negatives in every group
    cases.append((p, 1e-8))
    p = np.full(9, 1 / 9)
    p[3] = np.nan
    cases.append((p, 1e-8))
    cases.append((np.full(9, 1 / 9), 1e20))    # eps beyond the domain, finite L
    cases.append((np.full(9, 1 / 9), 1e-8))    # inside the domain (the sweeps)
    return cases


def check_against_oracle(O, co, rows_of, seg, pairs):
    cases = domain_points()
    for i, (pt, eps) in enumerate(cases):
        pp = np.repeat(np.asarray(pairs)[None, :], 1, 0)
        f, g, h, ns = co.evaluate(pp, pt[None, :], eps=eps)
        rows, valid = rows_of(pairs)
        fe = O.obj_seg(rows, valid, seg, pt, eps)
        ge, he = O.gh_seg(rows, valid, seg, pt, eps)
        assert same(f[0], fe), (i, eps, f[0], fe)
        assert same(g[0], ge), (i, eps)
        assert same(h[0], he), (i, eps)
        # the objective's meaning in the reference's own order
        fr = O.obj_ref(rows[valid == 1], pt, eps)
        if np.isfinite(fr):
            assert np.isfinite(f[0]) and abs(f[0] - fr) <= 1e-12 * abs(fr), (i, f[0], fr)
        else:
            assert same(f[0], fr), (i, f[0], fr)
    return len(cases)


def test_domain_2bit(O):
    say("case: evaluate outside the sweeps' domain, 2-bit cohort 30 x 5000")
    c = synth.generate(30, 5000, seed=61, n_family_pairs=4, missing=0.05)
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        codes, afk = np.ascontiguousarray(c.codes[:, kept]), af[kept]
        seg = np.minimum(segment_bounds(co.n_kept), codes.shape[1]).astype(np.uint32)

        def rows_of(p):
            rows, valid, _ = O.pair_rows_codes(codes[p[0]], codes[p[1]], afk)
            return rows, valid
        co.set_option("class_sweep", 0)  # the table sweep: every point in the segmented order
        n = check_against_oracle(O, co, rows_of, seg, (3, 17))
        check_against_oracle(O, co, rows_of, seg, (5, 5))
        table = [co.evaluate(np.array([[3, 17]]), pt[None, :], eps=eps) for pt, eps in domain_points()]
        co.set_option("class_sweep", 1)  # the class-sorted sweep for the points inside the domain
        for i, (pt, eps) in enumerate(domain_points()):
            f, g, h, _ = co.evaluate(np.array([[3, 17]]), pt[None, :], eps=eps)
            inside = i == len(domain_points()) - 1
            if inside:
                F, G, H, m = O.class_sums(codes[3], codes[17], afk, pt, eps)
                hh = np.zeros((9, 9))
                iu = np.triu_indices(9)
                hh[iu] = H / m
                hh[(iu[1], iu[0])] = H / m
                exp = (np.array([-F / m]), (-G / m)[None, :], hh[None, :, :])
            else:
                exp = table[i][:3]
            assert same(f, exp[0]) and same(g, exp[1]) and same(h, exp[2]), (i, eps)
        # a batch mixing both kinds of points gives each point's own result
        # (both forms)
        co.set_option("class_sweep", 0)
        pts = np.array([p for p, _ in domain_points()])
        for i, pt in enumerate(pts):
            ft, gt, ht, _ = co.evaluate(np.tile([3, 17], (len(pts), 1)), pts, eps=1e-8)
            fi, gi, hi, _ = co.evaluate(np.array([[3, 17]]), pt[None, :], eps=1e-8)
            assert same(ft[i], fi[0]) and same(gt[i], gi[0]) and same(ht[i], hi[0])
        co.set_option("class_sweep", 1)
        f, g, h, _ = co.evaluate(np.tile([3, 17], (len(pts), 1)), pts, eps=1e-8)
        for i, pt in enumerate(pts):
            fi, gi, hi, _ = co.evaluate(np.array([[3, 17]]), pt[None, :], eps=1e-8)
            assert same(f[i], fi[0]) and same(g[i], gi[0]) and same(h[i], hi[0])
        # the fit's points lie on the simplex: its eps must keep every L positive normal
        from kestrel_amd import tune_default
        for eps in (0.0, 1e-19, 1e15, float("nan")):
            t = tune_default(100)
            t.epsilon = eps
            with pytest.raises(KestrelError):
                co.fit_pairs(np.array([[0, 1]]), tune=t)
    say(f"{n} cases bit-exact against the oracle")


def test_domain_soft(O):
    say("case: evaluate outside the sweeps' domain, soft likelihoods 24 x 3000")
    c = synth.generate(24, 3000, seed=62, missing=0.02)
    gl = synth.soft_likelihoods(c.codes, seed=63, mean_depth=3.0)
    with Cohort.gl(gl) as co:
        af, kept = co.allele_freqs()
        g = gl.transpose(1, 0, 2)[:, kept]
        seg = np.minimum(segment_bounds(co.n_kept), g.shape[1]).astype(np.uint32)

        def rows_of(p):
            rows, valid, _ = O.pair_rows_gl(g[p[0]], g[p[1]], af[kept])
            return rows, valid
        check_against_oracle(O, co, rows_of, seg, (2, 11))


def test_domain_multi_allelic(O):
    say("case: evaluate outside the sweeps' domain, 2/3/4-allele likelihoods")
    from test_gpu_multi import device_order
    b2, b3, b4 = synth.multi_likelihoods(16, (700, 200, 90), seed=64)
    parts, _ = O.multi_prepare({2: b2, 3: b3, 4: b4})
    pads, seg = device_order(parts)
    with Cohort.gl_multi(b2, b3, b4) as co:
        def rows_of(p):
            return O.multi_pair_rows(parts, int(p[0]), int(p[1]), O.ORDER_SEG, pads)
        check_against_oracle(O, co, rows_of, seg, (1, 9))


def test_likelihoods_must_be_finite_and_non_negative():
    gl = np.full((10, 4, 3), 0.5)
    gl[3, 1, 2] = -0.1
    with pytest.raises(KestrelError):
        Cohort.gl(gl)
    gl[3, 1, 2] = np.inf
    with pytest.raises(KestrelError):
        Cohort.gl(gl)
