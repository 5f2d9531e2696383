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


def test_relatedness_coefficients_gl(O):
    print("case: relatedness_coefficients_gl 24 x 3000", flush=True)
    c = synth.generate(24, 3000, seed=8, missing=0.02, n_family_pairs=6)
    gl = np.eye(3)[np.minimum(c.codes, 2)]  # (N, V, 3) one-hot
    gl[c.codes == 3] = np.nan
    jac, kin, inb = relatedness_coefficients_gl(gl.transpose(1, 0, 2))
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    kc = c.codes[:, kept]
    seg = np.minimum(segment_bounds(int(kept.sum())), kc.shape[1]).astype(np.uint32)
    exp = O.fit_pairs(kc, af[kept], order=O.ORDER_SEG, seg=seg)
    ej, ek, ei = build_arrays(exp, 24)
    assert np.array_equal(jac, ej) and np.array_equal(kin, ek) and np.array_equal(inb, ei)
