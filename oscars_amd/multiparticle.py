"""Multi-particle (incoherent) sums on one GPU: beam sampling, uneven
trajectory windows and uneven observation requests, evaluated in one
persistent launch through the batched C ABI (om_batch_setup / om_batch_run).

The reference's equivalent is the nparticles loop of the calculate_* methods
(src/OSCARSSR.cc:4016-4063, 1778-1818), which samples, integrates and sums
one particle after another.
"""
import math

import numpy as np

from . import _lib
from . import beam as B
from . import geometry as G

C = G.C


def om_particles(particles, windows):
    """om_particle structs from geometry.Particle objects and (ctstart, ctstop, npoints) windows."""
    out = []
    for p, (a, b, n) in zip(particles, windows):
        o = _lib.om_particle()
        o.x0[:] = p.x0
        o.beta0[:] = p.b0
        o.t0 = p.t0
        o.q = p.q
        o.m = p.m
        o.gamma = p.gamma
        o.qomg = p.qomg
        o.current = p.current
        o.ctstart = a
        o.ctstop = b
        o.npoints = int(n)
        out.append(o)
    return out


def uneven_windows(particles, rng, npts_range=(8001, 20001), step_m=1e-4, z_start=(-1.0, -0.85)):
    """Random start z and trajectory lengths: particle i starts at z_start
    drawn uniformly, with T0 = 0 at its start, and runs for n_i - 1 steps of
    step_m (c*dt), n_i uniform in npts_range."""
    windows = []
    for p in particles:
        zs = rng.uniform(*z_start)
        p.x0 = [p.x0[0], p.x0[1], zs]
        n = int(rng.integers(npts_range[0], npts_range[1] + 1))
        windows.append((0.0, (n - 1) * step_m, n))
    return windows


def random_requests(n_req, rng, grid_range=(1, 41), e_range=(128, 1024), half_width=0.002, distance=30.0,
                    energy_span=(100.0, 3000.0)):
    """n_req requests: square grids n x n (n uniform in grid_range) on +-half_width at
    `distance`, each with m energies (m uniform in e_range) evenly spaced over energy_span."""
    reqs = []
    for _ in range(n_req):
        n = int(rng.integers(grid_range[0], grid_range[1] + 1))
        m = int(rng.integers(e_range[0], e_range[1] + 1))
        obs, nv, cont = G.rectangle("XY", n, n, 2 * half_width, 2 * half_width, [0, 0, 0], [0, 0, distance], 0, 3)
        e = np.array(G.energies(m, energy_span[0], energy_span[1]))
        reqs.append((np.array(obs), e))
    return reqs


def config_beam(current=0.5):
    """Gaussian beam of the batched benchmark: 3 GeV, emittances 0.9 nm / 8 pm,
    beta 1.5 m / 0.8 m, alpha 0, sigma_E/E 9e-4 (energy spread 2.7 MeV)."""
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, current)
    return dict(q=p0.q, m=p0.m, energy_GeV=3.0, x0=[0.0, 0.0, -1.0], d0=[0, 0, 1], t0=0.0, current=current,
                sigma_energy_GeV=3.0 * 9e-4, beta=[1.5, 0.8], alpha=[0.0, 0.0], emittance=[0.9e-9, 8e-12])


class BatchedFlux:
    """Resident particle set + requests on one device; run() is one step."""

    def __init__(self, device, fields, particles, windows, requests, req_of, weights, mode="mixed",
                 polarization=(0.0,) * 6, precision=0.01, max_level=-2, max_level_extended=0, level_cap=0,
                 waves_per_cu=0, efields=(), deep_level=0):
        self.dev = device
        self.n_particles = len(particles)
        self.requests = requests
        device.trajectories(fields, om_particles(particles, windows), efields=efields)
        conv = device.conv(precision, max_level, max_level_extended, 0, mode)
        device.batch_setup(requests, req_of, weights, list(polarization), conv, waves_per_cu, level_cap, deep_level)

    def tune(self, candidates=(2, 4, 8, 16), sample_items=20000):
        return self.dev.batch_tune(candidates, sample_items)

    def tune_deep(self, deep_levels=(0, 13, 15)):
        """Hand-over level (0: none) timed on the whole workload; the fastest is kept."""
        return self.dev.batch_tune_deep(deep_levels)

    def run(self, recompute_trajectories=True):
        return self.dev.batch_run(recompute_trajectories)


def assign_round_robin(n_particles, n_req):
    req_of = np.arange(n_particles) % n_req
    counts = np.bincount(req_of, minlength=n_req)
    weights = 1.0 / counts[req_of]
    return req_of, weights
