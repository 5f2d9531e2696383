"""Write the golden fixtures of tests/golden from the oracle.

c1_small.npz: a seeded cohort of 40 samples x 3000 biallelic sites (4
designated relative pairs, 2% missing), its allele frequencies (allele/mod.rs
SQP, maf 1/(2N)), kept-site mask, and the Output records of every pair
(x <= y) in the reference's 8-lane order ("out_ref") and in the segmented
order of the device sweep ("out_seg", with its segment bounds "seg").

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from kestrel_amd import segment_bounds, synth  # noqa: E402


def main():
    c = synth.generate(40, 3000, seed=20240601, n_family_pairs=4, missing=0.02, inbred_fraction=0.1)
    af = O.allele_freqs(c.codes)
    kept = O.keep_sites(af)
    codes, afk = c.codes[:, kept], af[kept]
    seg = np.minimum(segment_bounds(int(kept.sum()), segment_sites=1024), codes.shape[1]).astype(np.uint32)
    out_ref = O.fit_pairs(codes, afk, order=O.ORDER_REF, nthreads=1)
    out_seg = O.fit_pairs(codes, afk, order=O.ORDER_SEG, seg=seg, nthreads=1)
    path = os.path.join(ROOT, "tests", "golden", "c1_small.npz")
    np.savez_compressed(path, codes=c.codes, af=af, kept=kept, seg=seg, segment_sites=np.array(1024),
                        out_ref=out_ref, out_seg=out_seg, pairs=np.array([(x, y) for x, y, _ in c.pairs]))
    print("wrote", path, len(out_ref), "pairs,", int(kept.sum()), "kept sites")


if __name__ == "__main__":
    main()
