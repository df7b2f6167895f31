"""The mixed path's sin/cos error bound (OM_SINCOS_FAST_ERR, used by the
borderline test in oscars_amd/csrc/lw_common.h) against the reduced-accuracy
polynomials of sincos_coeffs.h, evaluated on the CPU in the kernel's operation
order (sincos_fast64 in om_device.h).  numpy rounds each product and sum
separately where the kernel fuses them; that adds at most a few 1e-16, far
inside the bound's margin.  Huge phases go through the kernel's reduction
modulo pi, done here in exact rational arithmetic with the kernel's roundings,
against 50-digit mpmath sin/cos."""
import os
import re
from fractions import Fraction

import mpmath as mp
import numpy as np

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oscars_amd", "csrc")


def _header():
    return open(os.path.join(CSRC, "sincos_coeffs.h")).read()


def _coeffs():
    co = {m.group(1): float(m.group(2)) for m in re.finditer(r"#define (OM_\w+) \(([-0-9.e+]+)\)", _header())}
    return [co["OM_SF_%d" % i] for i in range(4)], [co["OM_QF_%d" % i] for i in range(5)]


def _bound():
    return float(re.search(r"#define OM_SINCOS_FAST_ERR \(([0-9.e+-]+)\)", _header()).group(1))


def poly(r, sf=None, qf=None):
    S, Q = _coeffs()
    S = sf or S
    Q = qf or Q
    u = r * r
    ps = (u * S[3] + S[2]) * u + S[1]
    ps = ps * u + S[0]
    sr = (r * u) * ps + r
    pq = (u * Q[4] + Q[3]) * u + Q[2]
    pq = pq * u + Q[1]
    pq = pq * u + Q[0]
    return sr, u * pq + 1.0


def test_reduced_sincos_within_bound():
    rng = np.random.default_rng(7)
    e = _bound()
    rmax = np.pi / 2 + 2e-5
    worst = 0.0
    for it in range(4):
        r = rng.uniform(-rmax, rmax, 2_000_000)
        if it == 0:
            r[:6] = [rmax, -rmax, 0.0, 1e-300, np.pi / 4, -1e-8]
        s, c = poly(r)
        worst = max(worst, float(np.max(np.abs(s - np.sin(r)))), float(np.max(np.abs(c - np.cos(r)))))
    print("reduced-accuracy sin/cos worst error %.3e, bound %.3e" % (worst, e))
    assert worst < e
    # the bound keeps a margin but is not loose enough to hide a wrong polynomial
    assert worst > 0.5 * e


def _fast64(phi):
    """sincos_fast64 of one double: exact reduction modulo pi as the kernel rounds it."""
    one_over_pi = 0.31830988618379067154
    pi_hi = 3.1415926535897931160
    pi_lo = 1.2246467991473532072e-16
    magic = 6755399441055744.0
    kk = float(Fraction(phi) * Fraction(one_over_pi) + Fraction(magic))  # fma, rounded once
    k = kk - magic
    r = float(Fraction(phi) - Fraction(k) * Fraction(pi_hi))
    r = float(Fraction(r) - Fraction(k) * Fraction(pi_lo))
    s, c = poly(np.array([r]))
    sg = -1.0 if int(k) % 2 else 1.0
    return sg * float(s[0]), sg * float(c[0])


def test_huge_phases_within_bound():
    mp.mp.dps = 50
    rng = np.random.default_rng(3)
    e = _bound()
    phis = np.concatenate([rng.uniform(-6e11, 6e11, 1500), rng.uniform(-1e4, 1e4, 300),
                           np.pi * np.arange(-3, 4) + 1e-9])
    worst = 0.0
    for phi in phis:
        s, c = _fast64(float(phi))
        x = mp.mpf(float(phi))
        worst = max(worst, abs(s - float(mp.sin(x))), abs(c - float(mp.cos(x))))
    print("huge phases: worst error %.3e, bound %.3e" % (worst, e))
    assert worst < e


def test_bound_catches_a_perturbed_polynomial():
    # a plausible mistake (one coefficient off by 1e-5 relative) exceeds the bound
    S, Q = _coeffs()
    r = np.linspace(-np.pi / 2, np.pi / 2, 200001)
    s, _ = poly(r, sf=[S[0] * (1 + 1e-5)] + S[1:])
    assert float(np.max(np.abs(s - np.sin(r)))) > _bound()
    _, c = poly(r, qf=Q[:2] + [Q[2] * (1 + 1e-3)] + Q[3:])
    assert float(np.max(np.abs(c - np.cos(r)))) > _bound()
