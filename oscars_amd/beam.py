"""Gaussian beam sampling for multi-particle sums (``nparticles=``).

Restates TParticleBeam::GetNewParticle for the Gaussian distribution
(src/TParticleBeam.cc:692-773) and the Twiss transport to the start point
(src/TParticleBeam.cc:219-246, 312-328): the Twiss parameters given by the
user refer to the lattice reference point (default the origin) and are
drifted to x0 before sampling.  The random stream is numpy's PCG64 (seeded
with ``seed``), not the reference's std::mt19937, so two runs match
particle-for-particle only through explicit particle lists.
"""
import math

import numpy as np

from . import geometry as G

PI = G.PI


def beam_directions(d0, horizontal_direction=None):
    u0 = G.unit(d0)
    h = [float(v) for v in horizontal_direction] if horizontal_direction else [-v for v in G.orthogonal(d0)]
    h = G.unit(h)
    # src/TParticleBeam.cc:374-380: vertical = -(H x U0)
    c = G.cross(h, u0)
    vd = G.unit(c)
    return u0, h, [-vd[0], -vd[1], -vd[2]]


def twiss_at_x0(beta, alpha, x0, u0, reference=(0.0, 0.0, 0.0)):
    gamma = [(1. + alpha[0] * alpha[0]) / beta[0], (1. + alpha[1] * alpha[1]) / beta[1]]
    d = [x0[k] - reference[k] for k in range(3)]
    D = math.sqrt(G.mag2(d))
    L = D if (d[0] * u0[0] + d[1] * u0[1] + d[2] * u0[2]) >= 0 else -1. * D
    bx0 = [beta[i] - 2 * L * alpha[i] + L * L * gamma[i] for i in range(2)]
    ax0 = [alpha[i] - L * gamma[i] for i in range(2)]
    return bx0, ax0, gamma


def sample_particles(beam, n, seed=0, rng=None):
    """n particles (geometry.Particle) drawn from a Gaussian beam description
    (the dict kept by sr.set_particle_beam)."""
    rng = rng if rng is not None else np.random.Generator(np.random.PCG64(seed))
    q, m = beam["q"], beam["m"]
    E0 = beam["energy_GeV"]
    sigE = beam.get("sigma_energy_GeV", 0.0) or 0.0
    x0 = beam["x0"]
    u0, hdir, vdir = beam_directions(beam["d0"], beam.get("horizontal_direction"))
    eps = beam.get("emittance") or [0.0, 0.0]
    has_twiss = beam.get("beta") is not None
    if has_twiss:
        alpha = beam.get("alpha") or [0.0, 0.0]
        bX0, aX0, gX0 = twiss_at_x0(beam["beta"], alpha, x0, u0, beam.get("lattice_reference") or (0, 0, 0))
    eta = beam.get("eta") or [0.0, 0.0]
    out = []
    mc2 = G.kg_to_gev(m)
    for _ in range(n):
        ENew = E0 + sigE * rng.standard_normal()
        if ENew < mc2:
            ENew = mc2
        GammaE = 1 if ENew / mc2 < 1 else ENew / mc2
        Beta = math.sqrt(1.0 - 1.0 / (GammaE * GammaE)) if GammaE != 1 else 0
        dEoverE = (ENew - E0) / E0
        XNew = list(x0)
        c = G.cross(hdir, vdir)
        BetaNew = [c[0] * Beta, c[1] * Beta, c[2] * Beta]
        if (eps[0] * eps[0] + eps[1] * eps[1]) > 1e-99 and has_twiss:
            hConv = -1 if aX0[0] > 0 else 1
            vConv = -1 if aX0[1] > 0 else 1
            hSigX = math.sqrt(eps[0] * bX0[0])
            hSigXP = math.sqrt(eps[0] * gX0[0])
            hRho = math.sqrt(1 - (eps[0] / (hSigX * hSigXP)) ** 2)
            vSigX = math.sqrt(eps[1] * bX0[1])
            vSigXP = math.sqrt(eps[1] * gX0[1])
            vRho = math.sqrt(1 - (eps[1] / (vSigX * vSigXP)) ** 2)
            hu, hv = rng.random(), rng.random()
            hX = hSigX * math.sqrt(-2 * math.log(hu)) * (math.sqrt(1 - hRho * hRho) * math.cos(2 * PI * hv) +
                                                         hRho * math.sin(2 * PI * hv))
            hXP = hConv * hSigXP * math.sqrt(-2 * math.log(hu)) * math.sin(2 * PI * hv)
            vu, vv = rng.random(), rng.random()
            vX = vSigX * math.sqrt(-2 * math.log(vu)) * (math.sqrt(1 - vRho * vRho) * math.cos(2 * PI * vv) +
                                                         vRho * math.sin(2 * PI * vv))
            vXP = vConv * vSigXP * math.sqrt(-2 * math.log(vu)) * math.sin(2 * PI * vv)
            XNew = [XNew[k] + (hdir[k] * hX + eta[0] * dEoverE) for k in range(3)]
            XNew = [XNew[k] + (vdir[k] * vX + eta[1] * dEoverE) for k in range(3)]
            BetaNew = G.rot_axis(BetaNew, hXP, vdir)
            BetaNew = G.rot_axis(BetaNew, vXP, [-vdir_k for vdir_k in hdir])
        p = G.Particle(q, m, XNew, BetaNew, beam.get("t0", 0.0), beam.get("current", 0.0))
        p.energy_GeV = ENew
        out.append(p)
    return out
