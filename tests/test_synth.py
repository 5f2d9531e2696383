import numpy as np

from kestrel_amd import synth


def test_generator_is_seeded_and_shaped():
    a = synth.generate(30, 500, seed=1, n_family_pairs=4)
    b = synth.generate(30, 500, seed=1, n_family_pairs=4)
    assert np.array_equal(a.codes, b.codes) and a.pairs == b.pairs
    assert a.codes.shape == (30, 500) and a.codes.max() <= 3
    assert [k for _, _, k in a.pairs] == ["PO", "FS", "HS", "FC"]
    miss = (a.codes == 3).mean()
    assert 0.0 < miss < 0.03


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 4, (7, 1001)).astype(np.uint8)
    packed = synth.pack2(codes)
    assert packed.shape == (7, 1024 // 4)
    assert np.array_equal(synth.unpack2(packed, 1001), codes)
    # padding is code 3 (missing)
    assert (synth.unpack2(packed, 1024)[:, 1001:] == 3).all()
    # bit layout: site s in bits 2*(s%4) of byte s//4
    assert packed[0, 0] == codes[0, 0] | codes[0, 1] << 2 | codes[0, 2] << 4 | codes[0, 3] << 6


def test_per_sample_missingness_range():
    c = synth.generate(200, 2000, seed=2, missing_per_sample_max=0.3)
    rates = (c.codes == 3).mean(1)
    assert rates.min() < 0.05 and 0.2 < rates.max() < 0.33
