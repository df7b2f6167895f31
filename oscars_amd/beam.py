"""Beam sampling for multi-particle sums (``nparticles=``).

Restated from (paths relative to the reference root):
  random stream           src/TRandomA.cc:26-66 (std::mt19937 seeded with an
                          int, std::normal_distribution<double>(0, 1),
                          std::uniform_real_distribution<double>(0, 1)), and
                          include/OSCARSPY.h:85-87 (set_seed)
  beam parameters         src/OSCARSSR_Python.cc:3873-4040 (horizontal
                          direction, distribution, energy spread, Twiss
                          combinations, lattice reference, dispersion)
  Twiss setters           src/TParticleBeam.cc:219-346, 385-391
  new random particle     src/TParticleBeam.cc:692-850 (filament, Gaussian, KV)

``RefRandom`` reproduces the C++ standard library's generators as libstdc++
implements them, so that a seeded run draws the same particles, bit for bit,
as the reference seeded with the same number: the Mersenne Twister is the one
the C++ standard fixes (its 10000th output from the default seed is part of
the standard); uniform doubles come from two 32-bit outputs as libstdc++'s
generate_canonical<double, 53> forms them; normals come from the Marsaglia
polar method with the second value of each pair kept for the next call, as
libstdc++'s normal_distribution does.  The stream is global, as the
reference's is: every ``sr`` object draws from it, and ``set_seed`` restarts it.
"""
import math
import os

import numpy as np

from . import geometry as G

PI = G.PI

_N, _M = 624, 397
_UPPER, _LOWER, _MATRIX = 0x80000000, 0x7FFFFFFF, 0x9908B0DF
_TWO32 = 4294967296.0
_TWO64 = 18446744073709551616.0
_BELOW_ONE = 1.0 - 2.0 ** -53


class RefRandom:
    """std::mt19937 + std::normal_distribution + std::uniform_real_distribution
    (libstdc++), with the methods ``sample_particles`` draws through."""

    def __init__(self, seed):
        s = int(seed) & 0xFFFFFFFF  # the engine keeps the seed modulo 2^32
        mt = [s]
        for i in range(1, _N):
            p = mt[-1]
            mt.append((1812433253 * (p ^ (p >> 30)) + i) & 0xFFFFFFFF)
        self._mt = np.array(mt, dtype=np.uint64)
        self._out = []
        self._pos = 0
        self._saved = None  # the normal distribution's second value of a pair

    def _twist(self):
        mt = self._mt
        up, lo, mat = np.uint64(_UPPER), np.uint64(_LOWER), np.uint64(_MATRIX)
        one = np.uint64(1)

        def step(k0, k1, src):
            y = (mt[k0:k1] & up) | (mt[k0 + 1:k1 + 1] & lo)
            mt[k0:k1] = src ^ (y >> one) ^ np.where((y & one) != 0, mat, np.uint64(0))

        # words are replaced in order, each from words already replaced where
        # the recurrence reaches back (k + 397 wraps to k - 227)
        step(0, _N - _M, mt[_M:_N].copy())
        step(_N - _M, 2 * (_N - _M), mt[0:_N - _M].copy())
        step(2 * (_N - _M), _N - 1, mt[_N - _M:_M - 1].copy())
        y = (int(mt[_N - 1]) & _UPPER) | (int(mt[0]) & _LOWER)
        mt[_N - 1] = int(mt[_M - 1]) ^ (y >> 1) ^ (_MATRIX if y & 1 else 0)
        y = mt.copy()
        y ^= y >> np.uint64(11)
        y ^= (y << np.uint64(7)) & np.uint64(0x9D2C5680)
        y ^= (y << np.uint64(15)) & np.uint64(0xEFC60000)
        y ^= y >> np.uint64(18)
        self._out = (y & np.uint64(0xFFFFFFFF)).tolist()
        self._pos = 0

    def u32(self):
        """Next 32-bit output of the engine."""
        if self._pos == len(self._out):
            self._twist()
        v = self._out[self._pos]
        self._pos += 1
        return v

    def random(self):
        """uniform_real_distribution<double>(0, 1)."""
        g1 = self.u32()
        g2 = self.u32()
        r = (float(g1) + float(g2) * _TWO32) / _TWO64
        return _BELOW_ONE if r >= 1.0 else r

    def standard_normal(self):
        """normal_distribution<double>(0, 1)."""
        if self._saved is not None:
            v, self._saved = self._saved, None
            return v
        while True:
            x = 2.0 * self.random() - 1.0
            y = 2.0 * self.random() - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2 * math.log(r2) / r2)
        self._saved = x * mult
        return y * mult


_stream = None


def set_seed(seed):
    """sr.set_seed: restart the global stream.  The reference converts the
    argument to a double and then to an int (truncation towards zero)."""
    global _stream
    _stream = RefRandom(int(float(seed)))


def stream():
    """The global stream; before any set_seed it is seeded from the operating
    system, as the reference seeds it from std::random_device."""
    global _stream
    if _stream is None:
        _stream = RefRandom(int.from_bytes(os.urandom(4), "little"))
    return _stream


# ---------------------------------------------------------------- beam parameters
def twiss_from(beta=None, alpha=None, gamma=None):
    """(beta, alpha, gamma) at the lattice reference point from the combination
    the user gave; None when none of the reference's combinations applies (a
    lone alpha or gamma leaves the Twiss parameters unset)."""
    has = (4 if beta is not None else 0) | (2 if alpha is not None else 0) | (1 if gamma is not None else 0)
    if has == 7:
        return [float(v) for v in beta], [float(v) for v in alpha], [float(v) for v in gamma]
    if has in (6, 4):
        b = [float(v) for v in beta]
        a = [float(v) for v in alpha] if has == 6 else [0.0, 0.0]
        if b[0] <= 0 or b[1] <= 0:
            raise ValueError("Beta cannot be <= 0")
        return b, a, [(1. + a[0] * a[0]) / b[0], (1. + a[1] * a[1]) / b[1]]
    if has == 5:
        b, g = [float(v) for v in beta], [float(v) for v in gamma]
        return b, [math.sqrt(g[0] * b[0] - 1), math.sqrt(g[1] * b[1] - 1)], g
    if has == 3:
        a, g = [float(v) for v in alpha], [float(v) for v in gamma]
        if g[0] == 0 or g[1] == 0:
            raise ValueError("Gamma cannot be <= 0")
        return [(1. + a[0] * a[0]) / g[0], (1. + a[1] * a[1]) / g[1]], a, g
    return None


def twiss_at_x0(beta, alpha, gamma, x0, u0, reference=(0.0, 0.0, 0.0)):
    """Twiss parameters drifted from the lattice reference point to x0."""
    d = [x0[k] - reference[k] for k in range(3)]
    D = math.sqrt(G.mag2(d))
    L = D if (d[0] * u0[0] + d[1] * u0[1] + d[2] * u0[2]) >= 0 else -1. * D
    bx0 = [beta[i] - 2 * L * alpha[i] + L * L * gamma[i] for i in range(2)]
    ax0 = [alpha[i] - L * gamma[i] for i in range(2)]
    return bx0, ax0, list(gamma)


def beam_directions(u0, horizontal):
    """Unit horizontal and vertical directions of the beam from the horizontal
    direction as the binding passes it (already a unit vector)."""
    h = G.unit(horizontal)
    c = G.cross(horizontal, u0)
    vd = G.unit(c)
    return h, [-vd[0], -vd[1], -vd[2]]


# src/TParticleBeam.cc:135-200: name aliases -> electron beam at 3 GeV, 0.5 A
_PREDEFINED = {
    "NSLSII": dict(beta=[1.5, 0.8], sigma_energy_GeV=0.0, distribution="filament"),
    "NSLSII-LONGSTRAIGHT": dict(beta=[18.0, 3.1], sigma_energy_GeV=3. * 0.00089, distribution="gaussian"),
    "NSLSII-SHORTSTRAIGHT": dict(beta=[1.5, 0.8], sigma_energy_GeV=3. * 0.00089, distribution="gaussian"),
}


def predefined_beam(name):
    """Parameters of a predefined beam (NSLSII, NSLSII-LONGSTRAIGHT,
    NSLSII-SHORTSTRAIGHT; "NSLS2" and "NSLS-II" are aliases of "NSLSII")."""
    u = name.upper()
    for alias in ("NSLS-II", "NSLS2"):
        if u == alias or u.startswith(alias + "-"):
            u = "NSLSII" + u[len(alias):]
            break
    if u not in _PREDEFINED:
        raise ValueError("Error in predefined beam name / definition")
    d = dict(_PREDEFINED[u])
    d.update(energy_GeV=3.0, current=0.5, emittance=[0.55e-9, 0.008e-9])
    return d


def distribution_of(emittance, distribution=""):
    d = (distribution or "").lower()
    if d == "":
        return "gaussian" if emittance is not None else "filament"
    if d not in ("filament", "gaussian", "kv"):
        raise ValueError("Unknown beam distribution name: " + distribution)
    return d


# ---------------------------------------------------------------- sampling
def _frame(beam, u0):
    hin = beam.get("horizontal_unit")
    if hin is None:
        hd = beam.get("horizontal_direction")
        hin = G.unit([float(v) for v in hd] if hd else [-v for v in G.orthogonal(beam["d0"])])
    return beam_directions(u0, hin)


def beam_frame(beam):
    """(horizontal, vertical) unit directions of a beam description, the
    reference's fHorizontalDirection / fVerticalDirection
    (src/TParticleBeam.cc:374-379)."""
    ideal = G.ideal_particle(beam.get("type", "electron"), beam["energy_GeV"], beam["x0"], beam["d0"],
                             beam.get("t0", 0.0), beam.get("current", 0.0), beam["q"], beam["m"])
    return _frame(beam, beam.get("u0") or ideal.u0)


def sample_particles(beam, n, seed=None, rng=None):
    """n particles (geometry.Particle) drawn from a beam description (the dict
    kept by sr.set_particle_beam).  Draws come from ``rng`` (anything with
    ``standard_normal()`` and ``random()``), else from a fresh RefRandom(seed)
    when a seed is given, else from the global stream."""
    if rng is None:
        rng = RefRandom(seed) if seed is not None else stream()
    q, m = beam["q"], beam["m"]
    sigE = beam.get("sigma_energy_GeV", 0.0) or 0.0
    x0 = beam["x0"]
    ideal = G.ideal_particle(beam.get("type", "electron"), beam["energy_GeV"], x0, beam["d0"], beam.get("t0", 0.0),
                             beam.get("current", 0.0), q, m)
    E0 = ideal.e0
    u0 = beam.get("u0") or ideal.u0
    hdir, vdir = _frame(beam, u0)
    eps = [float(v) for v in beam["emittance"]] if beam.get("emittance") is not None else [0.0, 0.0]
    dist = beam.get("distribution") or distribution_of(beam.get("emittance"))
    tw = twiss_from(beam.get("beta"), beam.get("alpha"), beam.get("gamma"))
    if tw is None:
        if dist != "filament" and (eps[0] * eps[0] + eps[1] * eps[1]) > 1e-99:
            # the reference samples with zero beta here and its particles are NaN
            raise ValueError("beam has an emittance but no Twiss parameters (beta, alpha, gamma)")
        tw = ([0.0, 0.0], [0.0, 0.0], [0.0, 0.0])
    bX0, aX0, gX0 = twiss_at_x0(tw[0], tw[1], tw[2], x0, u0, beam.get("lattice_reference") or (0.0, 0.0, 0.0))
    eta = [float(v) for v in beam["eta"]] if beam.get("eta") is not None else [0.0, 0.0]
    mc2 = G.kg_to_gev(m)
    t0 = beam.get("t0", 0.0)
    cur = beam.get("current", 0.0)
    out = []
    for _ in range(n):
        if dist == "filament":
            out.append(G.Particle(q, m, ideal.x0, ideal.b0, t0, cur))
            out[-1].energy_GeV = E0
            continue
        ENew = E0 + sigE * rng.standard_normal()
        if ENew < mc2:
            ENew = mc2
        GammaE = 1 if ENew / mc2 < 1 else ENew / mc2
        Beta = math.sqrt(1.0 - 1.0 / (GammaE * GammaE)) if GammaE != 1 else 0
        c = G.cross(hdir, vdir)
        BetaNew = [c[0] * Beta, c[1] * Beta, c[2] * Beta]
        XNew = list(x0)
        if dist == "gaussian":
            dEoverE = (ENew - E0) / E0
            if (eps[0] * eps[0] + eps[1] * eps[1]) > 1e-99:
                hX, hXP, vX, vXP = _gaussian_offsets(rng, eps, bX0, aX0, gX0)
                XNew = [XNew[k] + (hdir[k] * hX + eta[0] * dEoverE) for k in range(3)]
                XNew = [XNew[k] + (vdir[k] * vX + eta[1] * dEoverE) for k in range(3)]
                BetaNew = G.rot_axis(BetaNew, hXP, vdir)
                BetaNew = G.rot_axis(BetaNew, vXP, [-v for v in hdir])
        else:
            hX, hXP, vX, vXP = _kv_offsets(rng, eps, bX0, aX0)
            XNew = [XNew[k] + hdir[k] * hX for k in range(3)]
            XNew = [XNew[k] + vdir[k] * vX for k in range(3)]
            BetaNew = G.rot_axis(BetaNew, hXP, vdir)
            BetaNew = G.rot_axis(BetaNew, vXP, [-v for v in hdir])
        p = G.Particle(q, m, XNew, BetaNew, t0, cur)
        p.energy_GeV = ENew
        out.append(p)
    return out


def _gaussian_offsets(rng, eps, bX0, aX0, gX0):
    # src/TParticleBeam.cc:736-764; pow(r, 2) is r * r, as the compiler folds it,
    # and the sin and cos of one angle come from one sincos call
    hConv = -1 if aX0[0] > 0 else 1
    vConv = -1 if aX0[1] > 0 else 1
    hSigX = math.sqrt(eps[0] * bX0[0])
    hSigXP = math.sqrt(eps[0] * gX0[0])
    r = eps[0] / (hSigX * hSigXP)
    hRho = math.sqrt(1 - r * r)
    vSigX = math.sqrt(eps[1] * bX0[1])
    vSigXP = math.sqrt(eps[1] * gX0[1])
    r = eps[1] / (vSigX * vSigXP)
    vRho = math.sqrt(1 - r * r)
    hu, hv = rng.random(), rng.random()
    s, c = G.sincos(2 * PI * hv)
    hX = hSigX * math.sqrt(-2 * math.log(hu)) * (math.sqrt(1 - hRho * hRho) * c + hRho * s)
    hXP = hConv * hSigXP * math.sqrt(-2 * math.log(hu)) * s
    vu, vv = rng.random(), rng.random()
    s, c = G.sincos(2 * PI * vv)
    vX = vSigX * math.sqrt(-2 * math.log(vu)) * (math.sqrt(1 - vRho * vRho) * c + vRho * s)
    vXP = vConv * vSigXP * math.sqrt(-2 * math.log(vu)) * s
    return hX, hXP, vX, vXP


def _kv_offsets(rng, eps, bX0, aX0):
    # src/TParticleBeam.cc:789-826: uniform inside the phase-space ellipse
    hConv = -1 if aX0[0] > 0 else 1
    vConv = -1 if aX0[1] > 0 else 1
    hTheta = rng.random() * 2 * PI
    vTheta = rng.random() * 2 * PI
    hs, hc = G.sincos(hTheta)
    vs, vc = G.sincos(vTheta)
    hAX = math.sqrt(eps[0] * bX0[0]) * hc
    vAX = math.sqrt(eps[1] * bX0[1]) * vc
    hAXP = -math.sqrt(eps[0] / bX0[0]) * (aX0[0] * hc + hs)
    vAXP = -math.sqrt(eps[1] / bX0[1]) * (aX0[1] * vc + vs)
    hR = math.sqrt(hAX * hAX + hAXP * hAXP)
    vR = math.sqrt(vAX * vAX + vAXP * vAXP)
    hRealTheta = math.atan(hAXP / hAX)
    vRealTheta = math.atan(vAXP / vAX)
    if hAX < 0 and hAXP > 0:
        hRealTheta += PI
    elif hAX < 0 and hAXP < 0:
        hRealTheta -= PI
    if vAX < 0 and vAXP > 0:
        vRealTheta += PI
    elif vAX < 0 and vAXP < 0:
        vRealTheta -= PI
    hNewR = math.sqrt(rng.random()) * hR
    vNewR = math.sqrt(rng.random()) * vR
    hs, hc = G.sincos(hRealTheta)
    vs, vc = G.sincos(vRealTheta)
    return hNewR * hc, hNewR * hs * hConv, vNewR * vc, vNewR * vs * vConv
