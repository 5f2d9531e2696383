"""relatedness_coefficients_gt / _gl (lib.rs:46-109) on the device against the
oracle: the arrays are assembled from records that are bit-identical to the
oracle's."""
import numpy as np
import pytest

from kestrel_amd import build_arrays, relatedness_coefficients_gl, relatedness_coefficients_gt, segment_bounds, synth

pytestmark = pytest.mark.gpu


def test_relatedness_coefficients_gt(O):
    print("case: relatedness_coefficients_gt 30 x 2000", flush=True)
    c = synth.generate(30, 2000, seed=7, missing=0.0, n_family_pairs=6)
    codes = c.codes  # (N, V), no missing call
    alleles = np.stack([(codes >= 1), (codes == 2)], -1).astype(np.uint8)  # 0/0, 0/1, 1/1
    res = relatedness_coefficients_gt(alleles.transpose(1, 0, 2))
    exp, _ = O.cls_fit_pairs(codes)
    jac, kin, inb = build_arrays(exp, 30)
    assert np.array_equal(res.jacquard, jac)
    assert np.array_equal(res.kinship, kin)
    assert np.array_equal(res.inbreeding, inb)


def test_relatedness_coefficients_gt_multiallelic(O):
    """Sites with 3 and 4 alleles (allele count per site given, or inferred
    from the largest allele index) on the three-bucket CLS path; each bucket
    keeps the input's site order."""
    print("case: relatedness_coefficients_gt multi-allelic 20 x (150, 40, 20)", flush=True)
    idx = synth.multi_genotypes(20, (150, 40, 20), seed=9)
    genos = [(i, j) for j in range(4) for i in range(j + 1)]
    cols = [(A, b[:, s]) for A, b in zip((2, 3, 4), idx) for s in range(b.shape[1])]
    order = np.random.default_rng(2).permutation(len(cols))  # buckets interleaved in the input
    g = np.array([[genos[v] for v in cols[k][1]] for k in order], np.uint8)  # (V, N, 2)
    gidx = np.array([cols[k][1] for k in order])  # (V, N) genotype indices

    def expected(alleles):
        buckets = [np.ascontiguousarray(gidx[alleles == A].T) for A in (2, 3, 4)]
        exp, _ = O.cls_fit_pairs_multi(*buckets)
        return build_arrays(exp, 20)

    alleles = np.array([cols[k][0] for k in order])
    res = relatedness_coefficients_gt(g, alleles=alleles)
    jac, kin, inb = expected(alleles)
    assert np.array_equal(res.jacquard, jac) and np.array_equal(res.kinship, kin)
    assert np.array_equal(res.inbreeding, inb)
    # inferred counts put a site whose highest allele is absent in a smaller bucket
    res2 = relatedness_coefficients_gt(g)
    jac2, _, _ = expected(np.maximum(2, g.max(axis=(1, 2)) + 1))
    assert np.array_equal(res2.jacquard, jac2)


def test_relatedness_coefficients_gl(O):
    print("case: relatedness_coefficients_gl 24 x 3000", flush=True)
    c = synth.generate(24, 3000, seed=8, missing=0.02, n_family_pairs=6)
    gl = np.eye(3)[np.minimum(c.codes, 2)]  # (N, V, 3) one-hot
    gl[c.codes == 3] = np.nan
    jac, kin, inb = relatedness_coefficients_gl(gl.transpose(1, 0, 2))
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    kc = np.ascontiguousarray(c.codes[:, kept])
    exp = O.fit_pairs(kc, af[kept], order=O.ORDER_CLASS)  # one-hot rows: the 2-bit path's class-sorted sweep
    ej, ek, ei = build_arrays(exp, 24)
    assert np.array_equal(jac, ej) and np.array_equal(kin, ek) and np.array_equal(inb, ei)
