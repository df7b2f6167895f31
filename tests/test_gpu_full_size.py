"""Parity at the benchmark configurations' real sizes.

Field-map adjoint (the bench's adjoint workload): 4001-sample By(z) map over
+-1 m with 0.5 % per-pole amplitude errors and 20 um pole jitter (seed 5), 3 GeV
through a 20001-point RK4 trajectory, flux on a 51 x 51 grid (+-5 mm at 30 m)
at E1 - 1.5, E1, 3 E1 - 4.5, 3 E1 (E1 = 152.11 eV), levels frozen at the
highest level the error-free map converges at.  The forward flux is checked
against the FP64 oracle on a seeded subset of the grid (bar 1e-9); the gradient
of a loss over that subset, sum (F - F_ideal)^2, against 4th-order central
differences of oracle/fdref.c in long double (bar 1e-4, as in
tests/test_gpu_adjoint.py; the adjoint gets the subset's weights, zero
elsewhere).

Batched multi-particle flux with the benchmark's trajectory lengths (uneven
8001-20001-point windows) and request shapes (n x n points at 30 m x 128-300
energies): FP64 and mixed batches against the oracle's particle loop (bars 1e-9
and 1e-5, denominators floored at 1e-6 of each request's maximum)."""
import numpy as np
import pytest

from tests.common import UNDULATOR, rel_err

gpu = pytest.mark.gpu

N, Z0, DZ = 4001, -1.0, 5e-4
E1 = 152.11
ENERGIES = np.array([E1 - 1.5, E1, 3 * E1 - 4.5, 3 * E1])


def _grid():
    from oscars_amd import geometry as G
    obs, _, _ = G.rectangle("XY", 51, 51, 0.01, 0.01, [0, 0, 0], [0, 0, 30.0], 0, 2)
    return np.array(obs)


@pytest.fixture(scope="module")
def adjoint_case():
    from oscars_amd import _lib
    from oscars_amd import adjoint as A
    from oscars_amd import geometry as G
    from oscars_amd import multiparticle as M
    dev = _lib.Device(0)
    err = A.undulator_pole_map(N, Z0, DZ, sigma_amp=0.005, sigma_z=20e-6, seed=5)
    ideal = A.undulator_pole_map(N, Z0, DZ)
    obs = _grid()
    f, keep = A.grid_field([0, 0, Z0], [0, 0, DZ], [1, 1, N], ideal)
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
    dev.trajectories([f], M.om_particles([p0], [(0.0, 2.0, 20001)]))
    _, lev, _, _, _ = dev.flux(0, obs, ENERGIES, [0.0] * 6, dev.conv(mode="tiled"))
    level = int(lev.max())
    prob = A.FieldMapProblem(dev, [0, 0, Z0], [0, 0, DZ], [1, 1, N], 3.0, [0, 0, -1], 0.0, 2.0, 20001, obs,
                             ENERGIES, level)
    rng = np.random.default_rng(41)
    sub = np.unique(np.concatenate([[0, 1300, 2600], rng.choice(len(obs), 9, replace=False)]))
    return prob, err, ideal, obs, sub, level


@gpu
def test_adjoint_forward_full_size_matches_oracle(adjoint_case):
    from oracle.pyoracle import Oracle
    prob, err, ideal, obs, sub, level = adjoint_case
    F = prob.forward(err)
    o = Oracle(nthreads=8)
    o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_grid([0, 0, Z0], [0, 0, DZ], [1, 1, N], err)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(20001)
    worst = 0.0
    for k, e in enumerate(ENERGIES):
        v, _, conv, _ = o.flux_points(obs[sub], e, precision=0.0, max_level=level, carry_quirk=False)
        assert not conv.any()
        worst = max(worst, rel_err(F[sub, k], v, 1e-6))
    print("full-size adjoint forward (level %d, %d points x 4 energies) vs oracle: %.3e" % (level, len(sub), worst))
    assert worst < 1e-9


@gpu
def test_adjoint_gradient_full_size_matches_long_double_differences(adjoint_case):
    from oracle.pyoracle import fdref_flux
    prob, err, ideal, obs, sub, level = adjoint_case
    F_ideal = prob.forward(ideal)
    F = prob.forward(err)  # the reverse pass differentiates the last forward
    weights = np.zeros_like(F)
    weights[sub] = 2.0 * (F[sub] - F_ideal[sub])
    gf, ga, _ = prob.backward(weights)

    def flux(s, ang):
        from oscars_amd import adjoint as A
        from oscars_amd import geometry as G
        p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
        b0, _, _ = A.beta0_from_angles(p0.b0, ang[0], ang[1])
        return fdref_flux(s, Z0, DZ, [0, 0, -1], b0, 0.5, 0.0, 2.0, 20001, obs[sub], ENERGIES, level,
                          long_double=True)

    f_ideal = flux(ideal, (0.0, 0.0))

    def loss(s, ang):
        return float(np.sum((flux(s, ang) - f_ideal) ** 2))

    def fd4(fun, e):
        return (-fun(2 * e) + 8 * fun(e) - 8 * fun(-e) + fun(-2 * e)) / (12 * e)

    picks = [(2000, 1, 3e-4), (1234, 1, 3e-4), (2890, 1, 3e-4), (611, 1, 3e-4), (2471, 0, 3e-4)]
    for i, c, step in picks:
        def fun(d, i=i, c=c):
            s = err.copy()
            s[i, c] += d
            return loss(s, (0.0, 0.0))
        fd = fd4(fun, step)
        rel = abs(gf[i, c] - fd) / abs(fd)
        print("full size: sample %d comp %d adjoint %.10e fd %.10e rel %.2e" % (i, c, gf[i, c], fd, rel))
        assert rel < 1e-4
    for k in range(2):
        def fun(d, k=k):
            a = [0.0, 0.0]
            a[k] += d
            return loss(err, a)
        fd = fd4(fun, 3e-8)
        rel = abs(ga[k] - fd) / abs(fd)
        print("full size: angle %d adjoint %.10e fd %.10e rel %.2e" % (k, ga[k], fd, rel))
        assert rel < 1e-4


@gpu
@pytest.mark.parametrize("mode,tol", [("tiled", 1e-9), ("mixed", 1e-5)])
def test_batched_flux_real_trajectory_lengths(mode, tol):
    from oscars_amd import beam as B
    from oscars_amd import geometry as G
    from oscars_amd import multiparticle as M
    from oscars_amd.sr import sr
    from tests.test_gpu_batch import oracle_sums
    rng = np.random.Generator(np.random.PCG64(2024))
    parts = B.sample_particles(M.config_beam(), 3, rng=rng)
    windows = M.uneven_windows(parts, rng)  # 8001-20001 points
    obs3, _, _ = G.rectangle("XY", 3, 3, 0.004, 0.004, [0, 0, 0], [0, 0, 30.0], 0, 3)
    reqs = [(np.array(obs3), np.array(G.energies(128, 100.0, 3000.0))),
            (np.array([[0.0005, -0.0003, 30.0]]), np.array(G.energies(300, 100.0, 3000.0)))]
    req_of, weights = M.assign_round_robin(len(parts), len(reqs))
    print("trajectory points:", [w[2] for w in windows])
    ref = oracle_sums(parts, windows, reqs, req_of, weights)
    s = sr()
    s.add_bfield_undulator(**UNDULATOR)
    dev = s._device()
    bf = M.BatchedFlux(dev, s._fields, parts, windows, reqs, req_of, weights, mode=mode)
    outs, st = bf.run(recompute_trajectories=True)
    assert st["n_unconverged"] == 0
    for r in range(len(reqs)):
        e = rel_err(outs[r], ref[r], 1e-6)
        print("%s, real lengths, request %d: max rel err %.3e" % (mode, r, e))
        assert e < tol
