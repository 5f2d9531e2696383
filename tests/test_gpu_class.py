"""GPU parity of the class-sorted sweep (cohort option "class_sweep" 1):
evaluations and whole fits bit-exact against the oracle's class order, and
within the stated tolerance of the reference's own order (max |dDelta_k| <=
1e-6, kinship 1e-6, relative objective 1e-9, log10 condition 0.1 on pairs with
log10 condition < 12)."""
import numpy as np
import pytest

from kestrel_amd import Cohort, kinship, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(out, exp):
    for k in ("x", "y", "jacquard", "obj", "iters", "n_sites", "cond"):
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} pairs differ"


def assert_within_tolerance(out, ref):
    lc = np.log10(ref["cond"])
    well = np.isfinite(lc) & (lc < 12)
    assert np.abs(out["jacquard"] - ref["jacquard"]).max(1)[well].max() <= 1e-6
    assert np.abs(kinship(out["jacquard"]) - kinship(ref["jacquard"]))[well].max() <= 1e-6
    assert (np.abs(out["obj"] - ref["obj"]) / np.abs(ref["obj"]))[well].max() <= 1e-9
    assert np.abs(np.log10(out["cond"])[well] - lc[well]).max() <= 0.1


def class_cohort(n, v, seed, missing):
    c = synth.generate(n, v, seed=seed, n_family_pairs=max(2, n // 8), missing=missing, inbred_fraction=0.05)
    co = Cohort(c.codes)
    co.set_option("class_sweep", 1)
    af, kept = co.allele_freqs()
    return c, co, np.ascontiguousarray(c.codes[:, kept]), af[kept]


def expected_eval(O, codes, afk, pair, pt):
    F, G, H, n = O.class_sums(codes[pair[0]], codes[pair[1]], afk, pt)
    h = np.zeros((9, 9))
    iu = np.triu_indices(9)
    h[iu] = H / n
    h[(iu[1], iu[0])] = H / n
    return -F / n, -G / n, h, n


# site counts around the 8,192-site blocks of the lists (one block, a partial
# last block, several) and missing fractions that empty whole pattern lists
@pytest.mark.parametrize("n,v,seed,missing", [(40, 3000, 11, 0.01), (48, 9500, 12, 0.05), (24, 17500, 13, 0.3)])
def test_class_evaluate_bitexact(O, n, v, seed, missing):
    say(f"case: class-sorted evaluate {n} x {v} missing {missing}")
    c, co, codes, afk = class_cohort(n, v, seed, missing)
    with co:
        rng = np.random.default_rng(seed)
        m = 160
        pairs = rng.integers(0, n, (m, 2))
        pairs[:8, 1] = pairs[:8, 0]  # self pairs
        pts = rng.dirichlet(np.ones(9), m)
        pts[8:24] = 0.0
        pts[8:24, rng.integers(0, 9, 16)] = 1.0  # vertices: L = eps on sites whose c_k is 0
        pts[24:40] = np.where(rng.random((16, 9)) < 0.5, 0.0, pts[24:40])
        pts[24:40] /= np.maximum(pts[24:40].sum(1, keepdims=True), 1e-300)
        pts[24:40, 8] += (pts[24:40].sum(1) == 0)
        f, g, h, ns = co.evaluate(pairs, pts)
        for i in range(m):
            fe, ge, he, ne = expected_eval(O, codes, afk, pairs[i], pts[i])
            assert ns[i] == ne
            assert f[i] == fe and np.array_equal(g[i], ge) and np.array_equal(h[i], he), i
            if i % 8 == 0:  # the reference's own order
                rows, valid, _ = O.pair_rows_codes(codes[pairs[i, 0]], codes[pairs[i, 1]], afk)
                comp = rows[valid == 1]
                fr = O.obj_ref(comp, pts[i])
                gr, hr = O.gh_ref(comp, pts[i])
                assert abs(f[i] - fr) <= 1e-13 * abs(fr)
                np.testing.assert_allclose(g[i], gr, rtol=1e-12, atol=1e-13 * np.abs(gr).max())
                np.testing.assert_allclose(h[i], hr, rtol=1e-12, atol=1e-13 * np.abs(hr).max())


def test_class_fit_all_pairs_bitexact_and_tolerance(O):
    say("case: class-sorted fit 64 x 10000 all pairs")
    c, co, codes, afk = class_cohort(64, 10000, 21, 0.01)
    with co:
        out = co.fit_pairs()
        assert (out["status"] == 0).all()
        assert_bitexact(out, O.fit_pairs(codes, afk, order=O.ORDER_CLASS))
        assert_within_tolerance(out, O.fit_pairs(codes, afk, order=O.ORDER_REF))
        # the same records with a pair list in another order and few slots
        rng = np.random.default_rng(2)
        idx = rng.permutation(len(out))[:500]
        pairs = np.stack([out["x"][idx], out["y"][idx]], 1)
        co.set_option("slots", 256)
        sub = co.fit_pairs(pairs)
        assert_bitexact(sub, out[idx])


def test_class_fit_high_missingness(O):
    say("case: class-sorted fit 80 x 20000, missing U(0, 0.3) per sample")
    rng = np.random.default_rng(5)
    c = synth.generate(80, 20000, seed=22, n_family_pairs=10, missing=0.0)
    codes_all = c.codes.copy()
    rates = rng.uniform(0, 0.3, 80)
    codes_all[rng.random(codes_all.shape) < rates[:, None]] = 3
    with Cohort(codes_all) as co:
        co.set_option("class_sweep", 1)
        af, kept = co.allele_freqs()
        codes, afk = np.ascontiguousarray(codes_all[:, kept]), af[kept]
        pairs = rng.integers(0, 80, (600, 2))
        out = co.fit_pairs(pairs)
    assert_bitexact(out, O.fit_pairs(codes, afk, pairs=pairs, order=O.ORDER_CLASS))
    assert_within_tolerance(out, O.fit_pairs(codes, afk, pairs=pairs, order=O.ORDER_REF))


def test_class_fit_c2_sample(O):
    """1,000 x 100,000 (the C2 cohort): a seeded sample of pairs with the
    designated relatives, bit-exact against the oracle's class order."""
    c = synth.config("c2")
    say("case: class-sorted fit, C2 cohort, 200 pairs + the relatives")
    with Cohort(c.codes) as co:
        co.set_option("class_sweep", 1)
        af, kept = co.allele_freqs()
        codes, afk = np.ascontiguousarray(c.codes[:, kept]), af[kept]
        rng = np.random.default_rng(7)
        pairs = np.concatenate([rng.integers(0, 1000, (200, 2)), np.array([(x, y) for x, y, _ in c.pairs])])
        out = co.fit_pairs(pairs)
    assert_bitexact(out, O.fit_pairs(codes, afk, pairs=pairs, order=O.ORDER_CLASS, nthreads=16))
    assert_within_tolerance(out, O.fit_pairs(codes, afk, pairs=pairs, order=O.ORDER_REF, nthreads=16))


def test_class_list_limit_keeps_records():
    """Option "class_list_mib" caps the lists' device memory: evaluations run
    in chunks of the points that fit and fits with fewer slots, bit-identical
    to the defaults; a cap below 256 slots a group is an error."""
    say("case: class-sorted lists under a memory cap, 48 x 9500")
    from kestrel_amd import _lib
    c, co, codes, afk = class_cohort(48, 9500, 31, 0.05)
    with co:
        rng = np.random.default_rng(7)
        pairs = rng.integers(0, 48, (700, 2))
        pts = rng.dirichlet(np.ones(9), 700)
        e0 = co.evaluate(pairs, pts)
        f0 = co.fit_pairs()
        sweeps0 = co.last_fit_stats()["sweeps"]
        # 9,500 sites: two blocks, 37,928 bytes of lists a slot
        co.set_option("class_list_mib", 10)  # 256 points at a time: three chunks
        e1 = co.evaluate(pairs, pts)
        for a, b in zip(e1, e0):
            assert same(a, b).all()
        co.set_option("class_list_mib", 20)  # 256 slots a group
        f1 = co.fit_pairs()
        assert co.last_fit_stats()["sweeps"] > sweeps0  # 1,176 pairs through 2 x 256 slots instead of 1,280
        assert_bitexact(f1, f0)
        co.set_option("class_list_mib", 1)  # not even 256 slots
        with pytest.raises(_lib.KestrelError):
            co.fit_pairs()
        with pytest.raises(_lib.KestrelError):
            co.evaluate(pairs, pts)
        co.set_option("class_list_mib", 0)
        assert_bitexact(co.fit_pairs(), f0)


def test_last_fit_form_names_the_sweep():
    """kestrel_last_fit_form: the class-sorted form for FP64 fits of 2-bit
    rows, the site-table form with option "class_sweep" 0 or in the mixed
    form, the wavefront form with "pair_lanes" 64; records of one form do not
    depend on the others having run."""
    say("case: last fit form, 40 x 3000")
    from kestrel_amd import _lib
    c, co, codes, afk = class_cohort(40, 3000, 41, 0.02)
    with co:
        a = co.fit_pairs()
        assert co.last_fit_stats()["form"] == "class"
        co.set_option("class_sweep", 0)
        co.fit_pairs()
        assert co.last_fit_stats()["form"] == "table"
        co.set_option("class_sweep", 1)
        co.set_option("pair_lanes", 64)
        co.fit_pairs()
        assert co.last_fit_stats()["form"] == "lanes"
        co.set_option("pair_lanes", 1)
        co.set_option("precision", _lib.MIXED)
        co.fit_pairs()
        assert co.last_fit_stats()["form"] == "table"
        co.set_option("precision", _lib.FP64)
        assert_bitexact(co.fit_pairs(), a)
        assert co.last_fit_stats()["form"] == "class"
