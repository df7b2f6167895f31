"""Pins of the oracle that do not come from the reference: closed forms,
invariants, an independent spline solver, and a high-precision brute force of
the Lienard-Wiechert sum.  Each check is chosen so that a plausible slip (a
dropped term, a wrong sign or index, a transposed operand, a wrong constant)
fails it."""
import math

import numpy as np
import pytest

from tests.common import planar, UNDULATOR

C = 2.99792458e8
QE = 1.602176462e-19
ME = 9.10938356e-31
HBAR = 1.054571596e-34


@pytest.fixture(scope="module")
def o_planar():
    from oracle.pyoracle import Oracle
    o = Oracle(nthreads=8)
    planar(o)
    o.calculate_trajectory()
    return o


def test_undulator_field_closed_form():
    from oracle.pyoracle import Oracle
    o = Oracle()
    o.add_bfield_undulator(**UNDULATOR)
    lam, n = 0.049, 31
    L = lam * (n + 2)
    for z in (0.0, 0.01, -0.3, 0.7):  # inside the full-field region
        assert o.get_bfield([0, 0, z])[1] == pytest.approx(math.sin(2 * math.pi * z / lam), abs=1e-15)
    z = -L / 2 + 0.2 * lam  # outer half period: quarter amplitude
    assert o.get_bfield([0, 0, z])[1] == pytest.approx(0.25 * math.sin(2 * math.pi * z / lam), abs=1e-15)
    z = L / 2 - 0.7 * lam  # next half period: three quarters
    assert o.get_bfield([0, 0, z])[1] == pytest.approx(0.75 * math.sin(2 * math.pi * z / lam), abs=1e-15)
    assert np.all(o.get_bfield([0, 0, L / 2 + 1e-6]) == 0)
    assert np.all(o.get_bfield([0.3, 0.2, -L / 2 - 1e-6]) == 0)


def test_grid_field_is_linear_interpolation():
    from oracle.pyoracle import Oracle
    o = Oracle()
    z = np.array([-1 + 0.01 * i for i in range(201)])
    s = np.stack([0.1 * z, 2 * z + 0.5, np.zeros_like(z)], axis=1)
    o.add_bfield_grid([0, 0, -1], [0, 0, 0.01], [1, 1, 201], s)
    for q in (-0.995, -0.2345, 0.0, 0.61718):
        b = o.get_bfield([0, 0, q])
        assert b[1] == pytest.approx(2 * q + 0.5, abs=1e-13)
        assert b[0] == pytest.approx(0.1 * q, abs=1e-13)
    # open interval: zero on and beyond the end samples
    assert np.all(o.get_bfield([0, 0, -1.0]) == 0) and np.all(o.get_bfield([0, 0, 1.0 + 1e-9]) == 0)


def test_uniform_field_circle():
    """Constant By from a grid map: the orbit is a circle of radius p/(eB)."""
    from oracle.pyoracle import Oracle
    o = Oracle()
    n = 101
    o.add_bfield_grid([0, 0, -2], [0, 0, 0.04], [1, 1, n], np.tile([0.0, 0.5, 0.0], (n, 1)))
    o.set_particle_beam(energy_GeV=1.0, x0=[0, 0, -1], d0=[0, 0, 1], current=0.1)
    o.set_ctstartstop(0, 1.5)
    o.set_npoints_trajectory(15001)
    T = o.calculate_trajectory()
    P = o.particle()
    gamma = P["gamma"]
    beta = math.sqrt(1 - 1 / gamma ** 2)
    rho = gamma * ME * beta * C / (QE * 0.5)
    x, z = T.p[:, 0], T.p[:, 2]
    # electron in +By bends towards +x: centre at (rho, -1)
    r = np.sqrt((x - rho) ** 2 + (z + 1) ** 2)
    assert np.max(np.abs(r - rho)) / rho < 1e-9
    # |beta| is conserved by the magnetic force
    b = np.linalg.norm(T.p[:, 3:6], axis=1)
    assert np.max(np.abs(b - beta)) < 1e-12


def test_undulator_deflection_amplitude(o_planar):
    """Peak x' inside the undulator equals K/gamma, K = e B lambda / (2 pi m c)."""
    T = o_planar.traj
    K = QE * 1.0 * 0.049 / (2 * math.pi * ME * C)
    gamma = o_planar.particle()["gamma"]
    z = T.p[:, 2]
    inner = np.abs(z) < 0.5
    xp = T.p[inner, 3] / T.p[inner, 5]
    assert np.max(np.abs(xp)) == pytest.approx(K / gamma, rel=2e-4)
    # Q1: the stored acceleration of knot i is the derivative at knot i-1
    qomg = o_planar.particle()["qomg"]
    i = 5000
    By = o_planar.get_bfield(T.p[i - 1, 0:3])[1]
    vz = T.p[i - 1, 5] * C
    assert T.p[i, 6] == pytest.approx(qomg * (-vz * By) / C, rel=1e-12)


def test_spline_matches_scipy_clamped(o_planar):
    """The spline is the clamped cubic spline with end slopes = end secants."""
    from scipy.interpolate import CubicSpline
    T = o_planar.traj
    sl = slice(8000, 8400)  # inside the undulator, where x oscillates
    t = T.t[sl].copy()
    y = T.p[sl, 0].copy()
    from oracle.pyoracle import Trajectory
    p9 = np.tile(y[:, None], (1, 9))
    Tr = Trajectory(t, p9)
    s0 = (y[1] - y[0]) / (t[1] - t[0])
    sN = (y[-1] - y[-2]) / (t[-1] - t[-2])
    cs = CubicSpline(t, y, bc_type=((1, s0), (1, sN)))
    assert np.max(np.abs(Tr.ypp[:, 0] - cs(t, 2))) < 1e-8 * np.max(np.abs(cs(t, 2)))
    tm = 0.5 * (t[:-1] + t[1:])
    vals = np.array([_spline_eval(Tr, x)[0] for x in tm[::7]])
    assert np.max(np.abs(vals - cs(tm[::7]))) < 1e-12 * np.max(np.abs(y)) + 1e-20


def _spline_eval(Tr, x):
    import ctypes as Cc
    from oracle.pyoracle import oracle_lib, _dp
    out = np.zeros(9)
    oracle_lib().orc_spline_value(Cc.byref(Tr.c()), float(x), out.ctypes.data_as(_dp))
    return out


def test_level_grid(o_planar):
    T = o_planar.traj
    ts, te = T.t[0], T.t[-1]
    for L in (0, 3, 9):
        lev = T.level(L)
        assert len(lev) == 2 ** L
        want = ts + (te - ts) / 2 ** (L + 1) + np.arange(2 ** L) * (te - ts) / 2 ** L
        assert np.allclose(lev[:, 0], want, rtol=0, atol=1e-22)
        assert T.dt_inclusive(L) == (te - ts) / (2 ** (L + 1) - 1)
    # levels 0..L together form the uniform midpoint grid of 2^(L+1)-1 points
    allt = np.sort(np.concatenate([T.level(L)[:, 0] for L in range(6)]))
    assert np.allclose(np.diff(allt), (te - ts) / 2 ** 6, rtol=1e-9)


def test_polarization_sum_rules(o_planar):
    """For an orthonormal (h, v) basis: |E.h|^2 + |E.v|^2 = |E_cl|^2 + |E_cr|^2 and
    all = lh + lv + |Ez|^2."""
    obs = [0.001, 0.0007, 30]
    # max_level=8 stops every polarization at the same level (convergence needs L > 8),
    # so the identities hold term by term
    kw = dict(npoints=15, energy_range_eV=[140, 600], carry_quirk=False, max_level=8)
    res = {}
    for pol in ("", "lh", "lv", "cl", "cr"):
        e, v, lev, conv, ef = o_planar.calculate_spectrum(obs, polarization=pol, **kw)
        res[pol] = (v, ef)
    lh, lv, cl, cr = res["lh"][0], res["lv"][0], res["cl"][0], res["cr"][0]
    assert np.allclose(lh + lv, cl + cr, rtol=1e-12)
    v_all, ef_all = res[""]
    C2 = v_all / np.sum(np.abs(ef_all) ** 2, axis=1)
    ez = C2 * np.abs(ef_all[:, 2]) ** 2
    assert np.allclose(v_all, lh + lv + ez, rtol=1e-12)


def test_scaling_with_current_and_mirror_symmetry():
    from oracle.pyoracle import Oracle
    vals = []
    for current in (0.5, 1.0):
        o = Oracle(nthreads=8)
        o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=current)
        o.add_bfield_undulator(**UNDULATOR)
        o.set_ctstartstop(0, 2)
        o.set_npoints_trajectory(8001)
        # points in +-y pairs, exactly mirrored
        obs = np.array([[x, y, 30.0] for x in (-0.004, 0.0, 0.003) for y in (-0.005, -0.002, 0.002, 0.005)])
        v, lev, conv, ef = o.flux_points(obs, 200.0)
        p, _, _ = o.power_density_points(obs, [0, 0, 1])
        vals.append((v, p))
    assert np.allclose(vals[1][0], 2 * vals[0][0], rtol=1e-14)
    assert np.allclose(vals[1][1], 2 * vals[0][1], rtol=1e-14)
    # y -> -y mirror: the planar orbit has y = 0, so the map is exactly symmetric in y
    v = vals[0][0].reshape(3, 4)
    assert np.array_equal(v, v[:, ::-1])
    p = vals[0][1].reshape(3, 4)
    assert np.array_equal(p, p[:, ::-1])


def test_on_axis_first_harmonic_kim(o_planar):
    """Peak on-axis flux near the first harmonic vs the far-field formula
    1.744e14 N^2 E^2[GeV] I F_1(K) photons/s/0.1%bw/mrad^2 (K.-J. Kim), at 30 m
    (1 mrad^2 = 900 mm^2).  The 1/4 and 3/4 end poles add about one period
    ((N+1)^2/N^2 = 1.066) and the finite distance a few per cent, so the ratio
    must lie in [1.0, 1.12]; a wrong constant or unit factor falls far outside."""
    from scipy.special import jv
    K = QE * 1.0 * 0.049 / (2 * math.pi * ME * C)
    gamma = o_planar.particle()["gamma"]
    E1 = 2 * gamma ** 2 * 2 * math.pi * HBAR * C / (0.049 * (1 + K ** 2 / 2)) / QE
    xi = K ** 2 / (4 * (1 + K ** 2 / 2))
    F1 = K ** 2 / (1 + K ** 2 / 2) ** 2 * (jv(0, xi) - jv(1, xi)) ** 2
    want = 1.744e14 * 31 ** 2 * 3.0 ** 2 * 0.5 * F1 / 900.0
    e = np.linspace(E1 * 0.97, E1 * 1.005, 36)
    v, lev, conv, ef = o_planar.spectrum([0, 0, 30], e, carry_quirk=False)
    assert E1 == pytest.approx(152.11, abs=0.02)
    assert 1.0 <= np.max(v) / want <= 1.12


def _lw_mpmath(T, obs, energy, L, current=0.5):
    """Level sums 0..L in 40-digit mpmath, except the quantities whose double
    rounding the FP64 formula cannot avoid and which are therefore part of its
    result: the phase -w (t + D/c) (~1e10 rad), and the three differences that
    cancel to ~1/gamma^2: 1 - n.b, n - b and 1 - |b|^2.  Those are taken in
    double precision, in the reference's operation order (with the double unit
    vector n); the products, cross products, denominators, constants and sums
    are exact."""
    import mpmath as mp
    mp.mp.dps = 40
    w_f = energy * 1.602176462e-19 / 1.054571596e-34
    E = [mp.mpc(0)] * 3
    Eabs = [mp.mpf(0)] * 3
    for l in range(L + 1):
        for row in T.level(l):
            X = [mp.mpf(x) for x in row[1:4]]
            B = [mp.mpf(x) for x in row[4:7]]
            A = [mp.mpf(x) for x in row[7:10]]
            R = [mp.mpf(obs[k]) - X[k] for k in range(3)]
            D = mp.sqrt(sum(r * r for r in R))
            N = [r / D for r in R]
            Rf = [obs[k] - row[1 + k] for k in range(3)]
            Df = math.sqrt(Rf[0] * Rf[0] + Rf[1] * Rf[1] + Rf[2] * Rf[2])
            Nf = [r / Df for r in Rf]
            omf = 1 - (Nf[0] * row[4] + Nf[1] * row[5] + Nf[2] * row[6])
            phif = -w_f * (row[0] + Df / 2.99792458e8)
            om = mp.mpf(omf)
            N = [mp.mpf(x) for x in Nf]
            NmB = [mp.mpf(Nf[k] - row[4 + k]) for k in range(3)]
            s1 = mp.mpf(1 - (row[4] * row[4] + row[5] * row[5] + row[6] * row[6]))
            cr1 = [NmB[1] * A[2] - NmB[2] * A[1], NmB[2] * A[0] - NmB[0] * A[2], NmB[0] * A[1] - NmB[1] * A[0]]
            cr2 = [N[1] * cr1[2] - N[2] * cr1[1], N[2] * cr1[0] - N[0] * cr1[2], N[0] * cr1[1] - N[1] * cr1[0]]
            ph = mp.expj(mp.mpf(phif))
            for k in range(3):
                term = (s1 * NmB[k] / (D * D * om * om) + cr2[k] / (D * om * om)) * ph
                E[k] += term
                Eabs[k] += abs(term)
    pi = mp.mpf("3.14159265358979323846")
    c = mp.mpf("2.99792458e8")
    eps0, mu0, h = mp.mpf("8.854187817e-12"), mp.mpf("1.2566370614e-6"), mp.mpf("6.62606876e-34")
    q = -mp.mpf("1.602176462e-19")
    C0 = q / (4 * pi * c * eps0 * mp.mpf("2.5066282746310002416123552393401"))
    C2 = 4 * pi * mp.mpf(current) / (h * abs(q) * mu0 * c) * mp.mpf("1e-6") * mp.mpf("0.001")
    dt = (mp.mpf(T.t[-1]) - mp.mpf(T.t[0])) / (2 ** (L + 1) - 1)
    flux = C2 * sum(abs(C0 * dt * e) ** 2 for e in E)
    cond = float(max(Eabs[k] / abs(E[k]) for k in range(3) if abs(E[k]) > 0))
    return float(flux), cond


@pytest.mark.parametrize("case", ["undulator", "straight"])
def test_lw_sum_brute_force_mpmath(case):
    """The oracle's level sums against _lw_mpmath.  'undulator' is dominated by
    the acceleration (far-field) term; 'straight' (no field, a = 0) has only
    the velocity (near-field) term, so both halves of the integrand are pinned."""
    from oracle.pyoracle import Oracle
    o = Oracle(nthreads=1)
    o.set_particle_beam(energy_GeV=3 if case == "undulator" else 0.05, x0=[0, 0, -0.1], d0=[0, 0, 1], current=0.5)
    if case == "undulator":
        o.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=2)
        obs, energy = [0.0004, -0.0002, 5.0], 150.0
    else:
        obs, energy = [0.01, 0.004, 0.3], 2.0
    o.set_ctstartstop(0, 0.2)
    o.set_npoints_trajectory(401)
    T = o.calculate_trajectory()
    L = 6
    v, lev, conv, ef = o.spectrum(obs, [energy], max_level=L, carry_quirk=False)
    flux, cond = _lw_mpmath(T, obs, energy, L)
    assert not conv[0]
    assert abs(flux - v[0]) / v[0] < 1e-14 * cond + 1e-13
