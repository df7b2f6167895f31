"""Reverse pass (HIP adjoint) against central finite differences.

Loss = sum over a 3 x 3 grid at 30 m (asymmetric in x and y, so that no
derivative vanishes by symmetry) and two energies (first and third harmonic) of
(F - F_ideal)^2, F the frozen-level flux (levels 0..10, precision 0) of a 3 GeV
electron through a 4001-sample By(z) map with per-pole amplitude and position
errors; F_ideal the same with the error-free map.

Two finite-difference references, both with the 4th-order central stencil:
  * oracle/fdref.c in long double (its double build equals the oracle to 1e-15):
    64-bit significands put the rounding noise of the 1e11-rad phase far below
    the derivative, and the quotients converge to ~1e-6 over steps
    1e-3 .. 3e-5 T and 1e-7 .. 3e-9 rad.  Bar: 1e-4 relative per derivative.
  * the FP64 oracle (the reference CPU path restated): its own rounding noise is
    ~5e-6 of the flux, which limits difference quotients to ~1e-2 here.
    Bar: 2e-2 on a directional derivative.
"""
import numpy as np
import pytest

gpu = pytest.mark.gpu

N_GRID = 4001
Z0, DZ = -1.0, 5e-4
LEVEL = 10
OBS = np.array([[x, y, 30.0] for x in (-0.002, 0.0, 0.0015) for y in (0.0, 0.0007, 0.0015)])
ENERGIES = np.array([150.0, 452.0])


def maps():
    from oscars_amd import adjoint as A
    err = A.undulator_pole_map(N_GRID, Z0, DZ, sigma_amp=0.005, sigma_z=20e-6, seed=7)
    ideal = A.undulator_pole_map(N_GRID, Z0, DZ)
    return err, ideal


def oracle_flux(samples, angles):
    from oracle.pyoracle import Oracle
    from oscars_amd import adjoint as A
    from oscars_amd import geometry as G
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
    b0, _, _ = A.beta0_from_angles(p0.b0, angles[0], angles[1])
    o = Oracle(nthreads=8)
    o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_grid([0, 0, Z0], [0, 0, DZ], [1, 1, N_GRID], samples)
    o.set_particle([0, 0, -1], list(b0), 0.0)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(4001)
    out = np.zeros((len(OBS), len(ENERGIES)))
    for k, e in enumerate(ENERGIES):
        v, lev, conv, ef = o.flux_points(OBS, e, precision=0.0, max_level=LEVEL, carry_quirk=False)
        assert not conv.any()  # precision 0: every point sums levels 0..LEVEL
        out[:, k] = v
    return out


def fdref_flux(samples, angles):
    from oracle.pyoracle import fdref_flux
    from oscars_amd import adjoint as A
    from oscars_amd import geometry as G
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
    b0, _, _ = A.beta0_from_angles(p0.b0, angles[0], angles[1])
    return fdref_flux(samples, Z0, DZ, [0, 0, -1], b0, 0.5, 0.0, 2.0, 4001, OBS, ENERGIES, LEVEL, long_double=True)


def fd4(fun, e):
    return (-fun(2 * e) + 8 * fun(e) - 8 * fun(-e) + fun(-2 * e)) / (12 * e)


@pytest.fixture(scope="module")
def problem():
    from oscars_amd import _lib
    from oscars_amd import adjoint as A
    dev = _lib.Device(0)
    return A.FieldMapProblem(dev, [0, 0, Z0], [0, 0, DZ], [1, 1, N_GRID], 3.0, [0, 0, -1], 0.0, 2.0, 4001, OBS,
                             ENERGIES, LEVEL)


@gpu
def test_adjoint_forward_matches_oracle(problem):
    err, ideal = maps()
    for s, ang in ((err, (0.0, 0.0)), (ideal, (2e-5, -1e-5))):
        f = problem.forward(s, ang)
        ref = oracle_flux(s, ang)
        rel = np.max(np.abs(f - ref) / np.abs(ref))
        print("forward vs oracle: %.3e" % rel)
        assert rel < 1e-9


@pytest.fixture(scope="module")
def gradients(problem):
    err, ideal = maps()
    F = problem.forward(err, (0.0, 0.0))
    f_ideal = problem.forward(ideal, (0.0, 0.0))
    F = problem.forward(err, (0.0, 0.0))
    gf, ga, gx = problem.backward(2.0 * (F - f_ideal))
    print("backward stats", problem.last_stats["backward"])
    return err, ideal, gf, ga


@gpu
def test_adjoint_matches_long_double_finite_differences(gradients):
    err, ideal, gf, ga = gradients
    f_ideal = fdref_flux(ideal, (0.0, 0.0))

    def loss(s, ang):
        return float(np.sum((fdref_flux(s, ang) - f_ideal) ** 2))

    # Bz couples to the motion only through the small vx: its derivatives are
    # ~1e-4 of the others, so its difference quotient uses a 30x larger step to
    # stay above the same absolute noise (truncation remains negligible).
    picks = [(2000, 1, 3e-4), (1500, 1, 3e-4), (2612, 1, 3e-4), (420, 1, 3e-4), (3560, 1, 3e-4), (2200, 0, 3e-4),
             (1777, 2, 1e-2)]
    for i, c, step in picks:
        def fun(d, i=i, c=c):
            s = err.copy()
            s[i, c] += d
            return loss(s, (0.0, 0.0))
        fd = fd4(fun, step)
        rel = abs(gf[i, c] - fd) / abs(fd)
        print("sample %d comp %d: adjoint %.10e fd %.10e rel %.2e" % (i, c, gf[i, c], fd, rel))
        assert rel < 1e-4
    for k in range(2):
        def fun(d, k=k):
            a = [0.0, 0.0]
            a[k] += d
            return loss(err, a)
        fd = fd4(fun, 3e-8)
        rel = abs(ga[k] - fd) / abs(fd)
        print("angle %d: adjoint %.10e fd %.10e rel %.2e" % (k, ga[k], fd, rel))
        assert rel < 1e-4


@gpu
def test_adjoint_matches_oracle_finite_differences(gradients):
    err, ideal, gf, ga = gradients
    f_ideal = oracle_flux(ideal, (0.0, 0.0))
    z = Z0 + DZ * np.arange(N_GRID)
    direction = np.zeros_like(err)
    direction[:, 1] = np.cos(2 * np.pi * z / 0.049) * np.exp(-z * z)
    fd = fd4(lambda d: float(np.sum((oracle_flux(err + d * direction, (0.0, 0.0)) - f_ideal) ** 2)), 3e-4)
    ad = float(np.sum(gf * direction))
    rel = abs(ad - fd) / abs(fd)
    print("directional vs FP64 oracle: adjoint %.8e fd %.8e rel %.2e" % (ad, fd, rel))
    assert rel < 2e-2


@gpu
def test_autograd_function_matches_problem(problem):
    import torch
    from oscars_amd import adjoint as A
    err, ideal = maps()
    f_ideal = torch.from_numpy(problem.forward(ideal, (0.0, 0.0)))
    s = torch.tensor(err, dtype=torch.float64, requires_grad=True)
    a = torch.zeros(2, dtype=torch.float64, requires_grad=True)
    F = A.FieldMapFlux.apply(s, a, problem)
    loss = ((F - f_ideal) ** 2).sum()
    loss.backward()
    F2 = problem.forward(err, (0.0, 0.0))
    gf, ga, _ = problem.backward(2.0 * (F2 - f_ideal.numpy()))
    assert np.allclose(s.grad.numpy(), gf, rtol=1e-12, atol=0)
    assert np.allclose(a.grad.numpy(), ga, rtol=1e-12, atol=0)


@gpu
def test_adjoint_gradient_is_bitwise_reproducible(problem):
    """The grid-sample gradient is summed per sample in a fixed record order
    (no floating-point atomics), so repeated reverse passes agree bit for bit."""
    err, ideal = maps()
    f_ideal = problem.forward(ideal, (0.0, 0.0))
    out = []
    for _ in range(3):
        F = problem.forward(err, (0.0, 0.0))
        gf, ga, gx = problem.backward(2.0 * (F - f_ideal))
        out.append((gf.copy(), np.array(ga), np.array(gx)))
    for gf, ga, gx in out[1:]:
        assert np.array_equal(gf, out[0][0]) and np.array_equal(ga, out[0][1]) and np.array_equal(gx, out[0][2])


@gpu
def test_device_resident_adjoint_equals_host_path(problem):
    """om_adjoint_forward_dev / om_adjoint_backward_dev (samples, flux and
    gradients as GPU tensors) give the host path's values bit for bit, and the
    autograd function on GPU tensors leaves its gradients on the GPU."""
    import torch
    from oscars_amd import adjoint as A
    err, ideal = maps()
    f_ideal = problem.forward(ideal, (0.0, 0.0))
    ang = (2e-6, -1e-6)
    F = problem.forward(err, ang)
    gf, ga, gx = problem.backward(2.0 * (F - f_ideal))
    s_t = torch.tensor(err, dtype=torch.float64, device="cuda")
    F_t = problem.forward_dev(s_t, ang)
    assert F_t.is_cuda and np.array_equal(F_t.cpu().numpy(), F)
    gf_t, ga_t, gx_t = problem.backward_dev(2.0 * (F_t - torch.as_tensor(f_ideal, device="cuda")))
    assert gf_t.is_cuda and ga_t.is_cuda and gx_t.is_cuda
    assert np.array_equal(gf_t.cpu().numpy(), gf)
    assert np.array_equal(gx_t.cpu().numpy(), gx)
    assert np.allclose(ga_t.cpu().numpy(), ga, rtol=1e-14, atol=0)
    # autograd on GPU tensors
    s = s_t.clone().requires_grad_(True)
    a = torch.tensor(ang, dtype=torch.float64, device="cuda", requires_grad=True)
    Fa = A.FieldMapFlux.apply(s, a, problem)
    assert Fa.is_cuda
    ((Fa - torch.as_tensor(f_ideal, device="cuda")) ** 2).sum().backward()
    assert s.grad.is_cuda and a.grad.is_cuda
    assert np.array_equal(s.grad.cpu().numpy(), gf)
    assert np.allclose(a.grad.cpu().numpy(), ga, rtol=1e-14, atol=0)
    print("device-resident adjoint: flux and grid gradients bit-identical to the host path")
