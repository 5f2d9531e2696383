"""Pins of the restated natural log (log/scalar.rs:17-70, log/avx512.rs:15-79)."""
import math
import struct

import numpy as np


def ulp_diff(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def test_log_within_one_ulp_of_libm(O):
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        np.exp(rng.uniform(-700, 700, 20000)),
        rng.uniform(0.5, 2.0, 20000),
        1.0 + rng.uniform(-1e-6, 1e-6, 2000),
        rng.uniform(1e-8, 1.0, 20000),
        np.array([1.0, 2.0, 0.5, 0.75, 1.5, 1e-8, 1e-8 + 1e-24, 1e308, 5e-324, 2.2250738585072014e-308,
                  2.2250738585072009e-308, 1e-310]),
    ])
    worst = 0
    for x in xs:
        worst = max(worst, ulp_diff(O.log(float(x)), math.log(float(x))))
    assert worst <= 1


def test_log_special_cases(O):
    assert O.log(0.0) == -math.inf
    assert O.log(-0.0) == -math.inf
    assert math.isnan(O.log(-1.0))
    assert math.isnan(O.log(math.nan))
    assert O.log(math.inf) == math.inf
    assert O.log(1.0) == 0.0


def test_getmant_form_matches_scalar_form(O):
    """The reference's AVX-512 lanes (getexp/getmant P75_1P5) and its scalar
    remainder (ilogb of d*4/3) split d = m 2^e identically on positive normals,
    so lanes and remainder rows compute the same log bits."""
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(1e-8, 2.0, 20000)])
    # mantissas at the 1.5 split
    for e in range(-30, 30):
        for m in (1.5, np.nextafter(1.5, 0), np.nextafter(1.5, 2), 1.0, np.nextafter(1.0, 0), 0.75):
            xs = np.append(xs, m * 2.0 ** e)
    for x in xs:
        assert O.log(float(x)) == O.log_getmant(float(x)), x
