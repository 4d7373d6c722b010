"""The random-number helpers divide a 16-bit or 8-bit integer by 65535 or 255
(Random.h:259-287, rnd16 / rnd8_4; Utils.h:126-140, decodeNormal).  The
kernels form the quotient without a divide -- the product with the rounded
reciprocal and one FMA correction step (unitQuotient in
raytracedgl1_amd/csrc/rg_device_math.h):

    q  = x * C                 C = float32(1 / D)
    r  = fma(-q, D, x)
    q' = fma(r, C, q)

This test evaluates that sequence with every operation correctly rounded to
float32 (exact rational arithmetic, then round-to-nearest-even), for every
x in [0, D], and requires q' to equal the correctly rounded float32 quotient
x / D.  The random numbers, and so every sampling decision, therefore stay
bit-exact with the oracle's plain divisions.  It also shows that the product
alone would not be enough.
"""
from fractions import Fraction

import numpy as np
import pytest


def _round_f32(v: Fraction) -> Fraction:
    """Round-to-nearest-even of an exact rational to float32 (normal range)."""
    if v == 0:
        return Fraction(0)
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = a.numerator.bit_length() - a.denominator.bit_length() - 24
    while a >= Fraction(2) ** (e + 24):
        e += 1
    while a < Fraction(2) ** (e + 23):
        e -= 1
    m = a / Fraction(2) ** e  # in [2^23, 2^24)
    n, rem = divmod(m.numerator, m.denominator)
    twice = 2 * rem
    if twice > m.denominator or (twice == m.denominator and n % 2 == 1):
        n += 1
    return sign * Fraction(n) * Fraction(2) ** e


def test_round_f32_matches_numpy():
    rng = np.random.default_rng(5)
    for v in rng.uniform(-1e3, 1e3, 2000):
        assert float(_round_f32(Fraction(float(v)))) == float(np.float32(v))


@pytest.mark.parametrize("d", [255, 65535])
def test_unit_quotient_is_exact(d):
    c = Fraction(float(np.float32(1.0) / np.float32(d)))
    xs = np.arange(d + 1, dtype=np.float32)
    want = xs / np.float32(d)  # IEEE float32 division, correctly rounded
    product_only_misses = 0
    for x in range(d + 1):
        fx = Fraction(x)
        q = _round_f32(fx * c)
        r = _round_f32(-q * d + fx)
        q1 = _round_f32(r * c + q)
        assert float(q1) == float(want[x]), (d, x)
        product_only_misses += float(q) != float(want[x])
    # the correction step is what makes it exact
    assert product_only_misses > 0
