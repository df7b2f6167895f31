"""div_by (oscars_amd/csrc/om_device.h): the quotient from a correctly rounded
reciprocal plus one FMA correction must be the IEEE-rounded quotient
(Markstein's theorem), since the amplitude's N = R/D and D/c feed the
bit-sensitive 1 - N.B and the phase.  Restated on the CPU with exactly rounded
FMAs (rational arithmetic) and compared with Python's IEEE division, on
amplitude-like operands and on random ones."""
from fractions import Fraction

import numpy as np


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div_by(q, b, inv):
    y = q * inv
    e = fma(-y, b, q)
    return fma(e, inv, y)


def test_div_by_is_the_ieee_quotient():
    rng = np.random.default_rng(11)
    c = 2.99792458e8
    cases = []
    # observation distances of tens of metres, components of R, and D / c
    d = rng.uniform(1.0, 60.0, 6000)
    comp = d * rng.uniform(-1.0, 1.0, 6000)
    cases += list(zip(comp, d)) + [(x, c) for x in d]
    # random magnitudes and mantissas
    a = rng.uniform(1.0, 2.0, 6000) * np.exp2(rng.integers(-40, 40, 6000))
    b = rng.uniform(1.0, 2.0, 6000) * np.exp2(rng.integers(-40, 40, 6000))
    cases += list(zip(a, b))
    bad = [(q, bb) for q, bb in cases if div_by(float(q), float(bb), 1.0 / float(bb)) != float(q) / float(bb)]
    assert not bad, bad[:5]


def test_the_correction_is_needed():
    # without the FMA correction, q * (1/b) misses the IEEE quotient in a fair
    # share of cases, so the check above can tell a dropped correction
    rng = np.random.default_rng(5)
    q = rng.uniform(1.0, 2.0, 4000)
    b = rng.uniform(1.0, 2.0, 4000)
    wrong = sum(float(x) * (1.0 / float(y)) != float(x) / float(y) for x, y in zip(q, b))
    assert wrong > 100
