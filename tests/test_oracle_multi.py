"""Pins of the oracle's tri/tetra-allelic paths: the generic mixture over
IIS_3/IIS_4 (mixture.rs:8-60), the 3/4-allele frequency fit (allele/*) and the
re-bucketing of remove_zero_frequencies (wrapper.rs:237-410)."""
import numpy as np
import pytest

E = 1e-8


def gl_index(a, b):
    i, j = min(a, b), max(a, b)
    return j * (j + 1) // 2 + i


def test_generic_a2_equals_the_biallelic_forms(O):
    rng = np.random.default_rng(0)
    for _ in range(100):
        p = rng.uniform(0.01, 0.99)
        af = np.array([p, 1 - p])
        lx, ly = rng.random(3), rng.random(3)
        assert np.array_equal(O.mixture_generic(2, lx, ly, af), O.mixture_generic2(lx, ly, af))
    gl = rng.random((45, 3))
    gl[7, 1] = np.nan
    a, ia = O.allele_freq_rows(gl, maf=0.0)
    b, ib = O.allele_freqA_rows(2, gl, maf=0.0)
    assert np.array_equal(a, b) and ia == ib


@pytest.mark.parametrize("A", [3, 4])
def test_state_columns_sum_to_one(O, A):
    """For every IBD state, the conditional genotype-pair probabilities of A
    alleles sum to one (one-hot rows select single (gx, gy) terms)."""
    G = A * (A + 1) // 2
    eye = np.eye(G)
    rng = np.random.default_rng(A)
    for _ in range(20):
        af = rng.dirichlet(np.ones(A))
        tot = sum(O.mixture_generic(A, eye[i], eye[j], af) for i in range(G) for j in range(G))
        np.testing.assert_allclose(tot, 1.0, rtol=0, atol=1e-14)


@pytest.mark.parametrize("A", [3, 4])
def test_generic_is_bilinear(O, A):
    G = A * (A + 1) // 2
    eye = np.eye(G)
    rng = np.random.default_rng(10 + A)
    for _ in range(20):
        af = rng.dirichlet(np.ones(A))
        lx, ly = rng.random(G), rng.random(G)
        c = O.mixture_generic(A, lx, ly, af)
        ref = sum(lx[i] * ly[j] * O.mixture_generic(A, eye[i], eye[j], af) for i in range(G) for j in range(G))
        np.testing.assert_allclose(c, ref, rtol=1e-13, atol=1e-300)


def test_absent_allele_reduces_to_biallelic(O):
    """A third allele of frequency 0 with likelihoods only on 00/01/11 gives
    the biallelic c_s (to rounding: the op orders differ)."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform(0.05, 0.95)
        lx3, ly3 = rng.random(3), rng.random(3)
        lx6, ly6 = np.zeros(6), np.zeros(6)
        lx6[:3], ly6[:3] = lx3, ly3
        c3 = O.mixture_generic(3, lx6, ly6, np.array([p, 1 - p, 0.0]))
        c2 = O.mixture_row2(lx3, ly3, np.array([p, 1 - p]))
        np.testing.assert_allclose(c3, c2, rtol=1e-14, atol=1e-300)


def _grad(gl, x):
    """-(1/n) sum over rows of d log(eps + sum_g w_g(x) l_g) / dx (allele/objective.rs)."""
    A = len(x)
    rows = gl[~np.isnan(gl).any(1)]
    d = np.full(len(rows), E)
    dd = np.zeros((len(rows), A))
    for j in range(A):
        for i in range(j + 1):
            lg = rows[:, gl_index(i, j)]
            if i == j:
                d += x[i] * x[i] * lg
                dd[:, i] += 2 * x[i] * lg
            else:
                d += 2 * x[i] * x[j] * lg
                dd[:, i] += 2 * x[j] * lg
                dd[:, j] += 2 * x[i] * lg
    return -(dd / d[:, None]).sum(0) / len(rows)


@pytest.mark.parametrize("A", [3, 4])
def test_multi_allele_frequencies_are_kkt_points(O, A):
    G = A * (A + 1) // 2
    rng = np.random.default_rng(20 + A)
    checked = 0
    for _ in range(30):
        true = rng.dirichlet(np.ones(A) * 2)
        n = 60
        geno = [sorted(rng.choice(A, 2, p=true)) for _ in range(n)]
        gl = np.full((n, G), 0.01) + rng.random((n, G)) * 0.05
        for r, (a, b) in enumerate(geno):
            gl[r, gl_index(a, b)] = 1.0
        gl[rng.integers(0, n)] = np.nan
        x, it = O.allele_freqA_rows(A, gl, maf=0.0)
        assert abs(x.sum() - 1) < 1e-12 and (x >= 0).all()
        if it >= 100:
            continue
        g = _grad(gl, x)
        lam = x @ g
        for xa, ga in zip(x, g):
            assert (ga >= lam - 1e-8 - 1e-10) if xa == 0.0 else abs(ga - lam) <= 1e-8 + 1e-10
        checked += 1
    assert checked >= 25


def test_one_hot_three_alleles_give_allele_counts(O):
    rng = np.random.default_rng(30)
    geno = [sorted(rng.choice(3, 2, p=[0.5, 0.3, 0.2])) for _ in range(400)]
    gl = np.zeros((400, 6))
    for r, (a, b) in enumerate(geno):
        gl[r, gl_index(a, b)] = 1.0
    x, _ = O.allele_freqA_rows(3, gl, maf=0.0)
    counts = np.bincount(np.array(geno).ravel(), minlength=3) / 800.0
    np.testing.assert_allclose(x, counts, atol=2e-6)


def test_zero_frequency_allele_moves_site_to_a_smaller_bucket(O):
    """A tri-allelic site whose allele 1 is never seen becomes biallelic on
    alleles (0, 2) with likelihood entries 00, 02, 22 (wrapper.rs:380-410)."""
    rng = np.random.default_rng(31)
    n = 50
    b3 = np.zeros((2, n, 6))
    for s in range(2):
        for r in range(n):
            a, b = sorted(rng.choice([0, 2], 2)) if s == 0 else sorted(rng.choice(3, 2))
            b3[s, r, gl_index(a, b)] = 1.0
    b2 = np.zeros((1, n, 3))
    b2[0, :, 1] = 1.0  # all heterozygous: both alleles at 1/2
    parts, af_in = O.multi_prepare({2: b2, 3: b3, 4: None})
    assert af_in[3][0, 1] == 0.0
    af2, gl2 = parts[2]
    af3, gl3 = parts[3]
    assert af2.shape == (2, 2) and af3.shape == (1, 3)
    np.testing.assert_array_equal(af2[1], af_in[3][0, [0, 2]])
    np.testing.assert_array_equal(gl2[:, 1], b3[0][:, [0, 3, 5]])
    np.testing.assert_array_equal(gl3[:, 0], b3[1])


def test_multi_fit_kkt(O):
    rng = np.random.default_rng(32)
    n = 16
    b2 = rng.random((300, n, 3))
    b3 = rng.random((40, n, 6))
    b4 = rng.random((12, n, 10))
    for b in (b2, b3, b4):
        b /= b.max(axis=2, keepdims=True)
    b2[5, 3] = np.nan
    parts, _ = O.multi_prepare({2: b2, 3: b3, 4: b4})
    pairs = [(0, 1), (2, 2), (3, 9), (7, 15)]
    out = O.multi_fit_pairs(parts, pairs)
    for (x, y), o in zip(pairs, out):
        rows, _ = O.multi_pair_rows(parts, x, y, O.ORDER_REF)
        assert o["n_sites"] == len(rows)
        L = rows @ o["jacquard"] + E
        g = -(rows / L[:, None]).sum(0) / len(rows)
        lam = o["jacquard"] @ g
        if o["iters"] < 100:
            for xi, gi in zip(o["jacquard"], g):
                assert (gi >= lam - 1e-8 - 1e-10) if xi == 0.0 else abs(gi - lam) <= 1e-8 + 1e-10
        assert abs(-np.log(L).mean() - o["obj"]) <= 1e-12 * abs(o["obj"])


def test_threaded_multi_fit_equals_the_pair_loop(O):
    """ko_fit_pairs_multi (the CPU baseline of the multi-allelic bench) gives
    the records of the per-pair loop, bit for bit, with the 2- and 4-allele
    buckets of ragged sizes (8-lane parts and remainders) and an empty one."""
    rng = np.random.default_rng(33)
    n = 12
    b2 = rng.random((203, n, 3))
    b3 = rng.random((29, n, 6))
    b4 = rng.random((11, n, 10))
    for b in (b2, b3, b4):
        b /= b.max(axis=2, keepdims=True)
    b2[7, 4] = np.nan
    b3[3, 0] = np.nan
    pairs = np.array([(0, 1), (2, 2), (3, 9), (7, 11), (4, 10), (5, 6)], np.uint32)
    for buckets in ({2: b2, 3: b3, 4: b4}, {2: b2, 3: None, 4: b4}):
        parts, _ = O.multi_prepare(buckets)
        want = O.multi_fit_pairs(parts, pairs)
        got = O.multi_fit_pairs_threads(parts, pairs, nthreads=3)
        for k in ("x", "y", "iters", "n_sites"):
            np.testing.assert_array_equal(got[k], want[k])
        for k in ("jacquard", "obj", "cond"):
            np.testing.assert_array_equal(got[k].view(np.uint64), want[k].view(np.uint64))
    with pytest.raises(ValueError):
        O.multi_fit_pairs_threads(parts, np.array([(0, n)], np.uint32))
