"""Fits with a pair's sites across a wavefront (option "pair_lanes" 64,
sweep_lanes_fp64 and sweep_lanes_gl) through the C ABI: records bit-exact
against the oracle in the wavefront order (ORDER_LANES) -- whole cohorts with
missing calls and self pairs, segments that are not multiples of 64 sites, C1
in full, soft likelihood rows of both strides -- and within the reference
tolerances of the default one-pair-per-lane form.  The option is refused for
the forms it does not cover."""
import numpy as np
import pytest

from kestrel_amd import Cohort, KestrelError, kinship, segment_bounds, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return (a == b) | (np.isnan(a) & np.isnan(b))


def assert_bitexact(out, exp, keys=("x", "y", "jacquard", "obj", "iters", "n_sites", "cond")):
    for k in keys:
        ok = same(out[k], exp[k]).reshape(len(out), -1).all(1)
        assert ok.all(), f"{k}: {int((~ok).sum())} of {len(out)} records differ"


def close(out, ref):
    well = np.isfinite(ref["cond"]) & (ref["cond"] < 1e12)
    assert np.abs(out["jacquard"] - ref["jacquard"])[well].max() <= 1e-6
    assert np.abs(kinship(out["jacquard"]) - kinship(ref["jacquard"]))[well].max() <= 1e-6
    assert (np.abs(out["obj"] - ref["obj"]) <= 1e-9 * np.abs(ref["obj"]))[well].all()


@pytest.mark.parametrize("samples,sites,segment_sites", [(40, 5000, None), (24, 3000, 96), (20, 700, 32)])
def test_lanes_bitexact_against_oracle(O, samples, sites, segment_sites):
    say(f"case: pair_lanes 64, {samples} x {sites}, segments of {segment_sites or 'default'} sites")
    c = synth.generate(samples, sites, seed=samples + sites, n_family_pairs=3, missing=0.04)
    with Cohort(c.codes, segment_sites=segment_sites) as co:
        af, kept = co.allele_freqs()
        base = co.fit_pairs()
        co.set_option("pair_lanes", 64)
        out = co.fit_pairs()
        co.set_option("pair_lanes", 1)
        again = co.fit_pairs()
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    kw = {} if segment_sites is None else {"segment_sites": segment_sites}
    seg = np.minimum(segment_bounds(int(kept.sum()), **kw), kc.shape[1]).astype(np.uint32)
    assert_bitexact(out, O.fit_pairs(kc, ka, order=O.ORDER_LANES, seg=seg))
    assert_bitexact(again, base)  # the option leaves the default form as it was
    close(out, base)
    say(f"{len(out)} records bit-exact in the wavefront order")


def test_lanes_c1_and_pair_list(O):
    say("case: pair_lanes 64, C1 64 x 10000 all pairs, then a 7-pair list with repeats")
    c = synth.config("c1")
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        co.set_option("pair_lanes", 64)
        out = co.fit_pairs()
        pairs = np.array([[3, 9], [9, 3], [5, 5], [0, 63], [3, 9], [12, 40], [40, 12]], np.uint32)
        sub = co.fit_pairs(pairs=pairs)
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    seg = np.minimum(segment_bounds(int(kept.sum())), kc.shape[1]).astype(np.uint32)
    assert_bitexact(out, O.fit_pairs(kc, ka, order=O.ORDER_LANES, seg=seg))
    assert_bitexact(sub, O.fit_pairs(kc, ka, pairs=pairs, order=O.ORDER_LANES, seg=seg))
    close(out, O.fit_pairs(kc, ka, order=O.ORDER_REF))


def test_lanes_graph_two_streams_and_autotune(O):
    say("case: pair_lanes 64 with HIP-graph relaunch, two streams and small slot groups; autotune keeps the option")
    c = synth.generate(36, 4000, seed=91, missing=0.03, n_family_pairs=2)
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        co.set_option("pair_lanes", 64)
        eager = co.fit_pairs()
        for opts in ({"graph": 1}, {"graph": 0, "streams": 2, "slots": 256}, {"batch": 3, "slots": 512}):
            for k, v in opts.items():
                co.set_option(k, v)
            assert_bitexact(co.fit_pairs(), eager)
            assert_bitexact(co.fit_pairs(), eager)  # a graph replayed again
        co.set_option("graph", 0)
        co.autotune(sample=600)
        assert_bitexact(co.fit_pairs(), eager)  # still the wavefront form after tuning
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    seg = np.minimum(segment_bounds(int(kept.sum())), kc.shape[1]).astype(np.uint32)
    assert_bitexact(eager, O.fit_pairs(kc, ka, order=O.ORDER_LANES, seg=seg))


@pytest.mark.parametrize("stride", [2, 3])
def test_lanes_soft_bitexact_against_oracle(O, stride):
    say(f"case: pair_lanes 64, soft likelihoods 28 x 3000, rows of stride {stride}")
    c = synth.generate(28, 3000, seed=93, missing=0.03, n_family_pairs=2)
    gl = synth.soft_likelihoods(c.codes, seed=94, mean_depth=3.0)
    if stride == 3:
        gl = gl.copy()
        gl[100, 3] = (0.5, 0.25, 0.125)  # a called row without an exact 1.0: 24-byte rows
    with Cohort.gl(gl) as co:
        af, kept = co.allele_freqs()
        base = co.fit_pairs()
        co.set_option("pair_lanes", 64)
        out = co.fit_pairs()
    g = np.ascontiguousarray(gl.transpose(1, 0, 2)[:, kept])
    seg = np.minimum(segment_bounds(int(kept.sum())), g.shape[1]).astype(np.uint32)
    assert_bitexact(out, O.fit_pairs_gl(g, af[kept], order=O.ORDER_LANES, seg=seg))
    close(out, base)


def test_lanes_refused_where_not_covered():
    say("case: pair_lanes 64 on tri/tetra-allelic and mixed cohorts, bad values")
    c = synth.generate(12, 800, seed=3)
    b2, b3, b4 = synth.multi_likelihoods(12, (300, 80, 40), seed=5)
    with Cohort.gl_multi(b2, b3, b4) as co:
        co.set_option("pair_lanes", 64)
        with pytest.raises(KestrelError):
            co.fit_pairs()
    with Cohort(c.codes, precision="mixed") as co:
        co.set_option("pair_lanes", 64)
        with pytest.raises(KestrelError):
            co.fit_pairs()
        for v in (0, 2, 32, -2):
            with pytest.raises(KestrelError):
                co.set_option("pair_lanes", v)
        co.set_option("pair_lanes", -1)  # automatic: the mixed form keeps one pair a lane
        assert len(co.fit_pairs(pairs=np.array([[0, 1]], np.uint32))) == 1


def test_lanes_automatic_by_job_size(O):
    say("case: pair_lanes -1, 30 x 2000: a 435-pair job in the wavefront order, 4,097 pairs in the default form")
    c = synth.generate(30, 2000, seed=77, missing=0.02)
    with Cohort(c.codes) as co:
        af, kept = co.allele_freqs()
        co.set_option("pair_lanes", -1)
        small = co.fit_pairs(pairs=np.array([[i, j] for i in range(30) for j in range(i + 1, 30)], np.uint32))
        rng = np.random.default_rng(1)
        big_pairs = rng.integers(0, 30, (4097, 2)).astype(np.uint32)
        big = co.fit_pairs(pairs=big_pairs)
    kc, ka = np.ascontiguousarray(c.codes[:, kept]), af[kept]
    seg = np.minimum(segment_bounds(int(kept.sum())), kc.shape[1]).astype(np.uint32)
    sp = np.stack([small["x"], small["y"]], 1)
    assert_bitexact(small, O.fit_pairs(kc, ka, pairs=sp, order=O.ORDER_LANES, seg=seg))
    idx = np.arange(0, 4097, 37)
    assert_bitexact(big[idx], O.fit_pairs(kc, ka, pairs=big_pairs[idx], order=O.ORDER_CLASS))  # the default form
