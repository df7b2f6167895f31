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


def _sh():
    co = {m.group(1): float(m.group(2)) for m in re.finditer(r"#define (OM_SH_\d) \(([-0-9.e+]+)\)", _header())}
    return [co["OM_SH_%d" % i] for i in range(8)]


def _sin64(x, sh=None):
    """sin64 (om_device.h) of one double: exact reduction modulo pi as the
    kernel rounds it, then the Estrin evaluation (numpy roundings)."""
    S = sh or _sh()
    one_over_pi, pi_hi, pi_lo, magic = 0.31830988618379067154, 3.1415926535897931160, 1.2246467991473532072e-16, \
        6755399441055744.0
    kk = float(Fraction(x) * Fraction(one_over_pi) + Fraction(magic))
    k = kk - magic
    r = float(Fraction(x) - Fraction(k) * Fraction(pi_hi))
    r = float(Fraction(r) - Fraction(k) * Fraction(pi_lo))
    u = r * r
    u2 = u * u
    u4 = u2 * u2
    q0 = (u2 * (u * S[3] + S[2])) + (u * S[1] + S[0])
    q1 = (u2 * (u * S[7] + S[6])) + (u * S[5] + S[4])
    s = (r * u) * (u4 * q1 + q0) + r
    return -s if int(k) % 2 else s


def test_field_sine_within_two_ulp():
    """The field models' sine: within two ulp of the exact value (half an ulp
    from the reduction's last rounding, the rest from the polynomial; numpy's
    separate roundings here stay inside the same two), near its zeros too, for
    arguments up to 1e5 rad."""
    mp.mp.dps = 50
    rng = np.random.default_rng(9)
    xs = np.concatenate([rng.uniform(-1e5, 1e5, 1500), rng.uniform(-10, 10, 1500),
                         np.pi * np.arange(-40, 41) + rng.uniform(-1e-6, 1e-6, 81), [1e-300, -1e-20, 0.0]])
    worst = 0.0
    for x in xs:
        ref = float(mp.sin(mp.mpf(float(x))))
        got = _sin64(float(x))
        ulp = np.spacing(abs(ref)) if ref != 0 else 5e-324
        worst = max(worst, abs(got - ref) / ulp)
    print("field sine: worst error %.2f ulp" % worst)
    assert worst <= 2.05
    # a coefficient off by 1e-12 relative is already visible
    S = _sh()
    bad = max(abs(_sin64(float(x), [S[0]] + [S[1] * (1 + 1e-12)] + S[2:]) - float(mp.sin(mp.mpf(float(x))))) /
              np.spacing(abs(float(mp.sin(mp.mpf(float(x)))))) for x in np.linspace(1.2, 1.5, 50))
    assert bad > 2.0
