"""The mixed power-density term (pd_term<true> in oscars_amd/csrc/lw.hip)
against the FP64 term (pd_term<false>, the reference's operations,
src/OSCARSSR.cc:2886-2921), evaluated on the CPU in the kernels' operation
order: IEEE square root and quotients for the FP64 term; fused multiply-adds
(exact product and sum, one rounding, done here in rational arithmetic), a
reciprocal square root and a reciprocal refined by two Newton steps for the
mixed term.  The hardware estimates the Newton steps start from are taken as
the exact values off by a random relative error of up to 2^-20.

The kernels' bound: |mixed - FP64| <= (Kx / x + K0) T + KQ Q with
Kx = 260 u, K0 = 64 u, KQ = 150 u (u = 2^-53), x = 1 - N.B,
T = |Num|^2 / (x^5 |R|^2) >= |term| and Q = x |Ac|^2 / (x^5 |R|^2).  It is
checked on electrons of 3 GeV (1 - beta = 1.45e-8) seen on axis, inside and
outside the radiation cone, with the acceleration perpendicular to the
velocity (a magnetic field: Num then cancels to about x |Ac|) and in any
direction, with and without a surface normal, down to the smallest x."""
import math
import random
from fractions import Fraction

U = 2.0 ** -53
KX, K0, KQ = 260 * U, 64 * U, 150 * U


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def estimate(v, rng):
    return v * (1.0 + rng.uniform(-2.0 ** -20, 2.0 ** -20))


def rsqrt_nr(x, rng):
    y = estimate(1.0 / math.sqrt(x), rng)
    h = 0.5 * x
    e = fma(-(h * y), y, 0.5)
    y = fma(y, e, y)
    e = fma(-(h * y), y, 0.5)
    return fma(y, e, y)


def recip_nr(x, rng):
    y = estimate(1.0 / x, rng)
    e = fma(-x, y, 1.0)
    y = fma(y, e, y)
    e = fma(-x, y, 1.0)
    return fma(y, e, y)


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def mag2_fma(a):
    return fma(a[2], a[2], fma(a[1], a[1], a[0] * a[0]))


def term64(obs, X, B, Ac, nrm):
    """The FP64 term (0 when a directional surface skips it)."""
    R = tuple(obs[k] - X[k] for k in range(3))
    r2 = dot(R, R)
    m = math.sqrt(r2)
    N1 = tuple(R[k] / m for k in range(3))
    ndn = dot(N1, nrm) if nrm else 1.0
    Num = cross(N1, cross(tuple(N1[k] - B[k] for k in range(3)), Ac))
    x = 1 - dot(B, N1)
    x2 = x * x
    den = x2 * x2 * x
    return 0.0 if (nrm and ndn <= 0) else dot(Num, Num) / den / r2 * ndn


def term_mixed(obs, X, B, Ac, nrm, rng):
    """Returns the term (0 when a directional surface skips it), T, Q and x."""
    R = tuple(obs[k] - X[k] for k in range(3))
    r2 = mag2_fma(R)
    im = rsqrt_nr(r2, rng)
    N1 = tuple(R[k] * im for k in range(3))
    ndn = fma(N1[2], nrm[2], fma(N1[1], nrm[1], N1[0] * nrm[0])) if nrm else 1.0
    NmB = tuple(N1[k] - B[k] for k in range(3))
    x = fma(-B[0], N1[0], fma(-B[1], N1[1], fma(-B[2], N1[2], 1.0)))
    nac = fma(N1[2], Ac[2], fma(N1[1], Ac[1], N1[0] * Ac[0]))
    Num = tuple(fma(NmB[k], nac, -(Ac[k] * x)) for k in range(3))
    x2 = x * x
    inv = recip_nr((x2 * x2) * (x * r2), rng)
    t = mag2_fma(Num) * inv
    term = 0.0 if (nrm and ndn <= 0) else t * ndn
    return term, t, x * mag2_fma(Ac) * inv, x, ndn


def _cases(rng, n):
    gamma = 3.0 / 0.51099895e-3
    beta = math.sqrt(1 - 1 / gamma ** 2)
    for i in range(n):
        kind = i % 4
        th = {0: 0.0, 1: rng.uniform(0, 3e-4), 2: rng.uniform(0, 3e-4), 3: rng.uniform(1e-3, 0.05)}[kind]
        ph = rng.uniform(0, 2 * math.pi)
        B = (beta * math.sin(th) * math.cos(ph), beta * math.sin(th) * math.sin(ph), beta * math.cos(th))
        X = (rng.uniform(-1e-4, 1e-4), rng.uniform(-1e-4, 1e-4), rng.uniform(-1.5, 1.5))
        if kind == 0:  # on axis of the electron's direction
            d = 20.0 - X[2]
            obs = (X[0] + B[0] / B[2] * d, X[1] + B[1] / B[2] * d, 20.0)
        else:
            obs = (rng.uniform(-0.04, 0.04), rng.uniform(-0.04, 0.04), 20.0)
        Ac = [rng.uniform(-1.0, 1.0) * 10 ** rng.uniform(9, 12) for _ in range(3)]
        if i % 5 != 4:  # perpendicular to the velocity, as in a magnetic field
            c = dot(Ac, B) / dot(B, B)
            Ac = [Ac[k] - c * B[k] for k in range(3)]
        Ac = tuple(Ac)
        if i % 3 == 0:
            nrm = None
        elif i % 3 == 1:
            nrm = (0.0, 0.0, 1.0)
        else:
            v = [rng.gauss(0, 1) for _ in range(3)]
            s = math.sqrt(sum(c * c for c in v))
            nrm = tuple(c / s for c in v)
        yield obs, X, B, Ac, nrm


def test_mixed_term_within_bound():
    rng = random.Random(11)
    worst = 0.0
    xs = []
    for obs, X, B, Ac, nrm in _cases(rng, 1600):
        t64 = term64(obs, X, B, Ac, nrm)
        tm, t, q, x, ndn = term_mixed(obs, X, B, Ac, nrm, rng)
        assert abs(tm) <= t * (1 + 1e-14)
        if nrm and ndn <= -16 * U:  # skipped by both paths: counts 0 in the kernels' bound
            assert tm == 0.0 and t64 == 0.0
            continue
        bound = (KX / x + K0) * t + KQ * q
        worst = max(worst, abs(tm - t64) / bound)
        xs.append(x)
    print("mixed power-density term: worst |mixed - FP64| / bound = %.3f, smallest x %.3e" % (worst, min(xs)))
    assert min(xs) < 2e-8
    assert worst < 1.0
    # the bound is not so loose that it would hide a term an order off
    assert worst > 0.02


def test_bound_catches_a_dropped_power():
    # a plausible mistake, x^4 in place of x^5 in the mixed denominator, breaks the bound
    rng = random.Random(5)
    hit = 0
    n = 0
    for obs, X, B, Ac, nrm in _cases(rng, 40):
        t64 = term64(obs, X, B, Ac, nrm)
        tm, t, q, x, ndn = term_mixed(obs, X, B, Ac, nrm, rng)
        if t64 == 0.0 or abs(ndn) < 1e-3:  # skipped, or grazing: far below the bound's scale T
            continue
            continue
        n += 1
        hit += abs(tm * x - t64) > (KX / x + K0) * t + KQ * q
    assert n > 15 and hit == n


def test_bound_catches_a_dropped_acceleration_term():
    # Num = (N-B)(N.Ac) without the -Ac (1 - N.B) part: outside the bound for
    # accelerations perpendicular to the velocity
    rng = random.Random(6)
    hit = n = 0
    for i, (obs, X, B, Ac, nrm) in enumerate(_cases(rng, 40)):
        if i % 5 == 4 or nrm:
            continue
        R = tuple(obs[k] - X[k] for k in range(3))
        r = math.sqrt(dot(R, R))
        N1 = tuple(c / r for c in R)
        x = 1 - dot(B, N1)
        nac = dot(N1, Ac)
        bad = sum(((N1[k] - B[k]) * nac) ** 2 for k in range(3)) / x ** 5 / r ** 2
        tm, t, q, x, _ = term_mixed(obs, X, B, Ac, nrm, rng)
        n += 1
        hit += abs(bad - term64(obs, X, B, Ac, nrm)) > (KX / x + K0) * t + KQ * q
    assert n > 5 and hit == n


def test_grazing_skip_decisions():
    """A directional surface whose normal is perpendicular to N up to rounding:
    the two paths skip different terms, and each such term lies within the
    bound the kernels count for it (N.normal > -16 u in the mixed path)."""
    rng = random.Random(21)
    gamma = 3.0 / 0.51099895e-3
    beta = math.sqrt(1 - 1 / gamma ** 2)
    flips = 0
    worst = 0.0
    for i in range(800):
        a, b = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        X = (0.0, 0.0, rng.uniform(-1.0, 1.0))
        obs = (a, 0.0, X[2] + b)
        r = math.sqrt(a * a + b * b)
        nrm = (b / r, rng.uniform(-1e-17, 1e-17), -a / r)
        th = rng.uniform(0, 1e-4)
        B = (beta * math.sin(th), 0.0, beta * math.cos(th))
        Ac = (0.0, rng.uniform(1e9, 1e12), 0.0)
        t64 = term64(obs, X, B, Ac, nrm)
        tm, t, q, x, ndn = term_mixed(obs, X, B, Ac, nrm, rng)
        flips += (tm == 0.0) != (t64 == 0.0)
        if ndn <= -16 * U:
            assert tm == 0.0 and t64 == 0.0
            continue
        worst = max(worst, abs(tm - t64) / ((KX / x + K0) * t + KQ * q))
    print("grazing normals: %d of 800 skip decisions differ, worst |mixed - FP64| / bound %.3f" % (flips, worst))
    assert flips > 10
    assert worst < 1.0
