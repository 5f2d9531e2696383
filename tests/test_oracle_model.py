"""Pins of the identity-state model: c_s = P(gx, gy | Jacquard state k; p).

Reference: mixture.rs:96-236 (mixture_row2), iis.rs:26-462, conditional.rs:10-50.
The expected values come from a brute-force derivation in exact rational
arithmetic: each Jacquard state is a partition of the four allele slots
(x1, x2, y1, y2) into identical-by-descent classes, each class draws an
allele independently with its population frequency, and genotypes are
unordered.  This is independent of the reference's hand-unrolled products.
"""
import itertools
import json
import os
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(__file__)

# Jacquard states 1..9 as partitions of slots 0=x1, 1=x2, 2=y1, 3=y2
STATES = [
    [[0, 1, 2, 3]],
    [[0, 1], [2, 3]],
    [[0, 1, 2], [3]],
    [[0, 1], [2], [3]],
    [[2, 3, 0], [1]],
    [[2, 3], [0], [1]],
    [[0, 2], [1, 3]],
    [[0, 2], [1], [3]],
    [[0], [1], [2], [3]],
]


def brute(freqs, gx, gy):
    """P(unordered genotype gx of x, gy of y | state k) for k = 1..9."""
    out = []
    n_alleles = len(freqs)
    for part in STATES:
        total = Fraction(0)
        for assign in itertools.product(range(n_alleles), repeat=len(part)):
            slot = [None] * 4
            prob = Fraction(1)
            for cls, a in zip(part, assign):
                prob *= freqs[a]
                for s in cls:
                    slot[s] = a
            if tuple(sorted(slot[:2])) == gx and tuple(sorted(slot[2:])) == gy:
                total += prob
        out.append(total)
    return out


ONEHOT = np.eye(3)
GENO2 = [(0, 0), (0, 1), (1, 1)]  # VCF triangular order (conditional.rs:251-283)


@pytest.mark.parametrize("p", [Fraction(1, 2), Fraction(1, 100), Fraction(37, 64), Fraction(997, 1000),
                               Fraction(1, 3)])
def test_mixture_row2_onehot_equals_jacquard_probabilities(O, p):
    q = 1 - p
    af = np.array([float(p), float(q)])
    for gx in range(3):
        for gy in range(3):
            c = O.mixture_row2(ONEHOT[gx], ONEHOT[gy], af)
            exp = np.array([float(v) for v in brute([p, q], GENO2[gx], GENO2[gy])])
            np.testing.assert_allclose(c, exp, rtol=4e-16 * 8, atol=0)
            # zeros are exact zeros
            assert np.array_equal(c == 0, exp == 0)


def test_state_columns_sum_to_one(O):
    rng = np.random.default_rng(1)
    for p in rng.uniform(0.001, 0.999, 50):
        af = np.array([p, 1 - p])
        tot = np.zeros(9)
        for gx in range(3):
            for gy in range(3):
                tot += O.mixture_row2(ONEHOT[gx], ONEHOT[gy], af)
        np.testing.assert_allclose(tot, 1.0, rtol=0, atol=1e-15)


def test_mixture_row2_equals_generic_path(O):
    """mixture_row2 and accumulate_iis over IIS_2 are two encodings of one model
    (mixture.rs:8-60 vs 96-236); they agree to rounding for arbitrary likelihoods."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        p = rng.uniform(0.005, 0.995)
        af = np.array([p, 1 - p])
        lx, ly = rng.random(3), rng.random(3)
        a = O.mixture_row2(lx, ly, af)
        b = O.mixture_generic2(lx, ly, af)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-300)


def test_nan_likelihood_gives_nan_row(O):
    """A missing call (NaN likelihoods, vcf.rs:355,397) makes every entry NaN so
    the row is dropped (lanevector.rs:100-104)."""
    af = np.array([0.3, 0.7])
    c = O.mixture_row2(np.full(3, np.nan), ONEHOT[1], af)
    assert np.isnan(c).all()
    c = O.mixture_row2(ONEHOT[2], np.full(3, np.nan), af)
    assert np.isnan(c).all()


def _geno_index(i, j):
    return j * (j + 1) // 2 + i


def test_conditional_probabilities_all_modes_bruteforce(O):
    """iis.rs:279-462 for every genotype pair of four alleles (all 15 modes)."""
    freqs = [Fraction(1, 10), Fraction(2, 10), Fraction(3, 10), Fraction(4, 10)]
    fl = [float(f) for f in freqs]
    seen = set()
    genos = [(i, j) for j in range(4) for i in range(j + 1)]
    for (i, j) in genos:
        for (k, l) in genos:
            mode = O.iis_mode(i, j, k, l)
            seen.add(mode)
            c = O.conditional_probabilities(fl[i], fl[j], fl[k], fl[l], mode)
            exp = np.array([float(v) for v in brute(freqs, (i, j), (k, l))])
            np.testing.assert_allclose(c, exp, rtol=1e-14, atol=0)
    # mode 10 (ij-ji) needs k > l, which the triangular genotype order never
    # produces; iis.rs:211-215 defines it as mode 9
    assert seen == set(range(1, 16)) - {10}
    for _ in range(3):
        a = O.conditional_probabilities(0.1, 0.3, 0.3, 0.1, 10)
        b = O.conditional_probabilities(0.1, 0.3, 0.1, 0.3, 9)
        assert np.array_equal(a, b)


def test_iis_modes_match_ibd_table(O):
    """iis.rs:26-104 and ibd.py:4-69 use one equality-bit table; under the
    Jacquard state ibd.py assigns to a pattern, that identity pattern of four
    distinct-enough alleles has non-zero probability."""
    fx = json.load(open(os.path.join(HERE, "golden", "ibd_state_map.json")))["bits_to_delta"]
    freqs = [0.1, 0.2, 0.3, 0.4]
    for i, j, k, l in itertools.product(range(4), repeat=4):
        bits = "".join(str(int(b)) for b in (i == j, i == k, i == l, j == k, j == l, k == l))
        mode = O.iis_mode(i, j, k, l)
        assert (bits in fx) == (mode != 0)
        if mode:
            delta = fx[bits]
            c = O.conditional_probabilities(freqs[i], freqs[j], freqs[k], freqs[l], mode)
            assert c[delta - 1] > 0


def test_isa_variants_agree(O, tmp_path):
    """The three compiled ISA variants of the lane loops return identical bits."""
    import subprocess
    import sys
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); import oracle as O;"
        "rng=np.random.default_rng(5); rows=rng.random((1003,9))*0.3; x=rng.dirichlet(np.ones(9));"
        "g,h=O.gh_ref(rows,x); f=O.obj_ref(rows,x);"
        "print(O.isa(), repr(f), g.tobytes().hex()[:64], h.tobytes().hex()[-64:])"
    ) % os.path.join(os.path.dirname(HERE), "oracle")
    outs = set()
    for isa in ("base", "avx2", "avx512"):
        r = subprocess.run([sys.executable, "-c", code], env={**os.environ, "KO_ISA": isa}, capture_output=True,
                           text=True, check=True)
        outs.add(r.stdout.split(" ", 1)[1])
    assert len(outs) == 1
