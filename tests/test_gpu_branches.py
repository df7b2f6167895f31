"""GPU parity of the device branches that the default calls reach but the
headline configurations do not: backward integration (T0 inside the
ctstartstop window, which is set_particle_beam's default window (-3, 3)),
extended levels above max_level, and sums capped below convergence with
their NotConverged flags and quantity='level' / 'precision' outputs.

The oracle is pinned to the reference build for each of these branches in
tests/test_oracle_vs_ref.py (test_trajectory_backward_and_forward,
test_extended_levels, test_nonconverged_carry).  The device reports levels
and NotConverged per point, as the reference's own GPU kernel does
(src/OSCARSSR_Cuda.cu:440-498), so the oracle runs with carry_quirk=False.

Bars: values 1e-9 relative per point (denominator floored at 1e-6 of the
largest value), levels and flags identical; trajectory knots 1e-12 of each
channel's largest magnitude (B fields) or 1e-8 (with E fields, where
sqrt(1 - beta^2) magnifies last-bit differences of the device sine).
"""
import numpy as np
import pytest

from tests.common import UNDULATOR, EPU, rel_err, with_efields

gpu = pytest.mark.gpu


def _knot_err(p, ref):
    scale = np.max(np.abs(ref), axis=0)
    scale[scale == 0] = 1
    return float(np.max(np.abs(p - ref) / scale))


def _same_levels(s, olev, oconv, cap=None):
    """The device's per-point level output is the last level summed: the
    converged level, or the cap for a point that did not converge (its
    NotConverged flag set; quantity='level' reports -1 for it, as the
    reference does).  The oracle reports -1 there."""
    lev = np.asarray(s._last_levels)
    conv = np.asarray(s._last_conv).astype(bool)
    assert np.array_equal(conv, oconv.astype(bool))
    assert np.array_equal(np.where(conv, lev, -1), olev)
    if cap is not None:
        assert np.all(lev[~conv] == cap)


def _pair(setup):
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    o, s = Oracle(nthreads=8), sr()
    setup(o)
    setup(s)
    return o, s


def _backward_undulator(x):
    # T0 = 0 at z = 0, the window ct in (-1, 1): 10000 knots backward, 10001 forward
    x.set_particle_beam(energy_GeV=3, x0=[0, 0, 0], d0=[0, 0, 1], current=0.5)
    x.add_bfield_undulator(**UNDULATOR)
    x.set_ctstartstop(-1, 1)
    x.set_npoints_trajectory(20001)


def _default_window(x):
    # set_particle_beam's default window ct in (-3, 3) at 10000 points per metre
    x.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    x.add_bfield_undulator(**UNDULATOR)


def _backward_epu(x):
    x.set_particle_beam(energy_GeV=3, x0=[0, 0, -0.3], d0=[0, 0, 1], current=0.5)
    x.add_bfield_epu(**EPU)
    x.set_ctstartstop(-1.2, 1.8)
    x.set_npoints_trajectory(30001)


def _backward_eb(x):
    with_efields(x)
    x.set_ctstartstop(-1, 1)
    x.set_npoints_trajectory(8001)


@gpu
@pytest.mark.parametrize("case", ["undulator", "default_window", "epu", "eb"])
def test_backward_trajectory_and_spectrum(case):
    """Knots, a spectrum and its levels with T0 inside the window."""
    setup = dict(undulator=_backward_undulator, default_window=_default_window, epu=_backward_epu,
                 eb=_backward_eb)[case]
    o, s = _pair(setup)
    T = o.calculate_trajectory()
    s.calculate_trajectory()
    t, p, ypp = s.get_trajectory_arrays()
    from oracle import pyoracle as P
    nb = P.C.c_long(0)
    nf = o.L.orc_npoints_forward_backward(o.ctstart, o.ctstop, o.npoints, o.P.t0, P.C.byref(nb))
    print("%s: %d backward + %d forward knots" % (case, nb.value, nf), flush=True)
    assert nb.value > 0 and len(t) == len(T.t) == nb.value + nf
    assert np.array_equal(t, T.t)
    ek = _knot_err(p, T.p)
    ey = _knot_err(ypp, T.ypp)
    # Y'' is the knots' second difference over h^2 (h = 3.3e-13 s): a knot
    # difference of 1e-15 of a channel becomes up to ~1e-9 of that channel's
    # largest Y''.  Its share of a level sample is Y'' h^2 / 6, which is back
    # at the knots' scale, so the spectrum below is the bar that matters.
    bar = 1e-8 if case == "eb" else 1e-12
    print("%s: knots %.3e, spline %.3e" % (case, ek, ey), flush=True)
    assert ek < bar and ey < max(1e-9, bar * 100)
    e = np.linspace(100.0, 900.0, 24)
    ov, olev, oconv, _ = o.spectrum([0, 0, 30], e, carry_quirk=False)
    for mode in ("plain", "tiled"):
        res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode=mode)
        v = np.array([r[1] for r in res])
        err = rel_err(v, ov, 1e-6)
        print("%s %s spectrum: %.3e" % (case, mode, err), flush=True)
        _same_levels(s, olev, oconv)
        assert err < 1e-9, (mode, err)


@gpu
def test_backward_flux_map_and_power_density():
    """A flux map and a power-density map on the backward-integrated
    trajectory (the default window), levels identical."""
    o, s = _pair(_default_window)
    kw = dict(npoints=[9, 7], width=[0.01, 0.008], translation=[0, 0, 30])
    res = s.calculate_flux_rectangle(energy_eV=152.11, **kw)
    v = np.array([r[1] for r in res])
    _, ov, olev, oconv, _, _ = o.calculate_flux_rectangle(152.11, carry_quirk=False, **kw)
    _same_levels(s, olev, oconv)
    ef = rel_err(v, ov, 1e-6)
    res = s.calculate_power_density_rectangle(**kw)
    v = np.array([r[1] for r in res])
    _, ov, olev, oconv, _ = o.calculate_power_density_rectangle(carry_quirk=False, **kw)
    _same_levels(s, olev, oconv)
    ep = rel_err(v, ov, 1e-6)
    print("default window: flux %.3e, power density %.3e" % (ef, ep))
    assert ef < 1e-9 and ep < 1e-9


@gpu
@pytest.mark.parametrize("field", ["undulator", "epu", "grid", "two_fields"])
def test_windows_of_their_own_in_one_launch(field):
    """Particles with windows of their own in one RK4 launch: forward only,
    backward only, and both ways with different lengths.  A particle's forward
    and backward integrations run side by side in one wave, each on two lanes
    that evaluate a step's fields two stages at a time (and for the EPU on two
    more, one per part), so lanes of a wave step in opposite directions and
    stop after different counts; each particle's knots must be the oracle's
    integration of its own window.  Fields: the planar undulator, the EPU, a
    By(z) map with pole errors and an undulator plus a Gaussian kick (the
    general variant)."""
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    windows = [(0.0, 0.6, 3001), (-0.5, 0.0, 2501), (-0.3, 0.7, 5001), (-0.8, 0.2, 4001), (-0.1, 0.9, 2001),
               (-0.6, 0.6, 6001), (0.0, 0.3, 1501)]
    samples = None
    if field == "grid":
        from oscars_amd import adjoint as A
        samples = A.undulator_pole_map(4001, -1.0, 5e-4, sigma_amp=0.005, sigma_z=20e-6, seed=3)

    def setup(x):
        x.set_particle_beam(energy_GeV=3, x0=[0, 0, 0], d0=[0, 0, 1], current=0.5)
        if field == "epu":
            x.add_bfield_epu(**EPU)
        elif field == "grid":
            x.add_bfield_grid([0, 0, -1.0], [0, 0, 5e-4], [1, 1, 4001], samples)
        else:
            x.add_bfield_undulator(**UNDULATOR)
            if field == "two_fields":
                x.add_bfield_gaussian(bfield=[0.05, 0, 0], sigma=[0, 0, 0.2], translation=[0, 0, 0.3])

    s = sr()
    setup(s)
    dev = s._device()
    dev.trajectories(s._fields, [s._om_particle(s._particle, ctstart=a, ctstop=b, npoints=n) for a, b, n in windows])
    for j, (a, b, n) in enumerate(windows):
        o = Oracle(nthreads=8)
        setup(o)
        o.set_ctstartstop(a, b)
        o.set_npoints_trajectory(n)
        T = o.calculate_trajectory()
        t, p, ypp = dev.trajectory(j)
        ek = _knot_err(p, T.p)
        print("%s window (%g, %g), %d knots: %.3e" % (field, a, b, len(t), ek), flush=True)
        assert np.array_equal(t, T.t) and ek < 1e-12


def _planar8001(x):
    x.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    x.add_bfield_undulator(**UNDULATOR)
    x.set_ctstartstop(0, 2)
    x.set_npoints_trajectory(8001)


@pytest.fixture(scope="module")
def planar8001():
    o, s = _pair(_planar8001)
    o.calculate_trajectory()
    return o, s


@gpu
@pytest.mark.parametrize("mode", ["plain", "tiled", "mixed"])
def test_extended_levels_spectrum(planar8001, mode):
    """Levels above max_level use the extended spacing T/(2^L - 1)
    (src/TParticleTrajectoryInterpolated.cc:242-251): max_level 9 and
    max_level_extended 12 on 1000-2000 eV, where points converge at 10-12."""
    o, s = planar8001
    e = np.linspace(1000.0, 2000.0, 33)
    kw = dict(max_level=9, max_level_extended=12)
    ov, olev, oconv, _ = o.spectrum([0, 0, 30], e, carry_quirk=False, **kw)
    res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode=mode, **kw)
    v = np.array([r[1] for r in res])
    err = rel_err(v, ov, 1e-6)
    print("extended levels %s: %.3e; levels %s" % (mode, err, np.unique(olev)))
    assert np.any(olev > 9)  # the extended levels are reached
    _same_levels(s, olev, oconv, cap=12)
    assert err < (1e-5 if mode == "mixed" else 1e-9)


@gpu
def test_extended_levels_flux_map(planar8001):
    o, s = planar8001
    kw = dict(npoints=[7, 7], width=[0.006, 0.006], translation=[0, 0, 30])
    cap = dict(max_level=8, max_level_extended=11)
    res = s.calculate_flux_rectangle(energy_eV=700.0, **kw, **cap)
    v = np.array([r[1] for r in res])
    _, ov, olev, oconv, _, _ = o.calculate_flux_rectangle(700.0, carry_quirk=False, **kw, **cap)
    err = rel_err(v, ov, 1e-6)
    print("extended flux map: %.3e, levels %s" % (err, np.unique(olev)))
    assert np.any(olev > 8)
    _same_levels(s, olev, oconv, cap=11)
    assert err < 1e-9


@gpu
@pytest.mark.parametrize("quantity", ["flux", "precision", "level"])
@pytest.mark.parametrize("mode", ["plain", "tiled"])
def test_capped_spectrum(planar8001, quantity, mode):
    """max_level 11 on 100-2000 eV, where points converge at levels 10-12:
    five stop unconverged at the cap.  Their value is the sum at the cap, their
    level -1, NotConverged set, and quantity='precision' gives the last
    level's relative change."""
    o, s = planar8001
    e = np.linspace(100.0, 2000.0, 40)
    kw = dict(max_level=11, quantity=quantity)
    ov, olev, oconv, _ = o.spectrum([0, 0, 30], e, carry_quirk=False, **kw)
    res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode=mode, **kw)
    v = np.array([r[1] for r in res])
    nconv = int(np.sum(oconv == 0))
    print("capped spectrum %s %s: %d of %d unconverged" % (mode, quantity, nconv, len(e)))
    assert 0 < nconv < len(e)
    _same_levels(s, olev, oconv, cap=11)
    if quantity == "level":
        assert np.array_equal(v, ov) and np.all(v[oconv == 0] == -1)
    else:
        assert rel_err(v, ov, 1e-6) < 1e-9


@gpu
@pytest.mark.parametrize("quantity", ["flux", "precision", "level"])
def test_capped_flux_map(planar8001, quantity):
    o, s = planar8001
    kw = dict(npoints=[9, 9], width=[0.01, 0.01], translation=[0, 0, 30])
    cap = dict(max_level=11, quantity=quantity)
    res = s.calculate_flux_rectangle(energy_eV=900.0, **kw, **cap)
    v = np.array([r[1] for r in res])
    _, ov, olev, oconv, _, _ = o.calculate_flux_rectangle(900.0, carry_quirk=False, **kw, **cap)
    nconv = int(np.sum(oconv == 0))
    print("capped flux map %s: %d of %d unconverged" % (quantity, nconv, len(ov)))
    assert 0 < nconv < len(ov)
    _same_levels(s, olev, oconv, cap=11)
    if quantity == "level":
        assert np.array_equal(v, ov)
    else:
        assert rel_err(v, ov, 1e-6) < 1e-9


@gpu
@pytest.mark.parametrize("quantity", ["power density", "precision", "level"])
def test_capped_power_density(quantity):
    """Power density on the EPU at 20 m (+-5 mm) capped at level 10, where
    points converge at levels 10-11: the level-11 points stop unconverged."""
    from tests.common import epu
    o, s = _pair(lambda x: epu(x, npoints=8001))
    kw = dict(npoints=[9, 9], width=[0.01, 0.01], translation=[0, 0, 20])
    cap = dict(max_level=10, quantity=quantity)
    res = s.calculate_power_density_rectangle(**kw, **cap)
    v = np.array([r[1] for r in res])
    _, ov, olev, oconv, _ = o.calculate_power_density_rectangle(carry_quirk=False, **kw, **cap)
    nconv = int(np.sum(oconv == 0))
    print("capped power density %s: %d of %d unconverged" % (quantity, nconv, len(ov)))
    assert 0 < nconv < len(ov)
    _same_levels(s, olev, oconv, cap=10)
    if quantity == "level":
        assert np.array_equal(v, ov)
    else:
        assert rel_err(v, ov, 1e-6) < 1e-9


@gpu
@pytest.mark.parametrize("electric", [False, True])
def test_time_dependent_fields(electric):
    """Fields with frequency / frequency_phase / time_offset on the device:
    the time factor is taken at every RK4 stage time (the oracle is pinned to
    the reference build in tests/test_oracle_vs_ref.py::test_time_dependent_fields).
    Field values at three times, knots (1e-12 of each channel; 1e-8 with the
    electric field, as test_electric_fields) and a spectrum at 1e-9, levels
    identical."""
    from tests.common import time_dependent
    o, s = _pair(lambda x: time_dependent(x, electric=electric))
    for t in (0.0, 1.3e-9, 4.1e-9):
        for x in ([0.0, 0.0, -0.7], [0.0004, 0.0, 0.02], [0.0, 0.001, 0.75], [0.001, 0.002, 0.9]):
            ref = o.get_bfield(x, t)
            got = np.array(s.get_bfield(x, t))
            assert np.allclose(got, ref, rtol=1e-14, atol=1e-15 * np.max(np.abs(ref))), (x, t, got, ref)
            refe = o.get_efield(x, t)
            assert np.allclose(np.array(s.get_efield(x, t)), refe, rtol=1e-14, atol=1e-15 * np.max(np.abs(refe) + 1))
    T = o.calculate_trajectory()
    s.calculate_trajectory()
    t, p, _ = s.get_trajectory_arrays()
    assert np.array_equal(t, T.t)
    ek = _knot_err(p, T.p)
    e = np.linspace(100.0, 900.0, 16)
    ov, olev, oconv, _ = o.spectrum([0, 0, 30], e, carry_quirk=False)
    v = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode="plain")])
    es = rel_err(v, ov, 1e-6)
    print("time-dependent fields (electric %s): knots %.3e, spectrum %.3e" % (electric, ek, es))
    assert ek < (1e-8 if electric else 1e-12) and es < 1e-9
    _same_levels(s, olev, oconv)


@gpu
@pytest.mark.parametrize("backward", [False, True])
def test_step_halving_uniform_box(backward):
    """Step halving in a uniform box, where no transcendental function enters
    the equations of motion: the device's sub-steps are the reference's
    operations in the same order, so the knots are bit-identical to the
    oracle's (pinned to the reference build)."""
    from tests.common import halving_box
    o, s = _pair(lambda x: halving_box(x, backward=backward))
    T = o.calculate_trajectory()
    n, depth = o.halvings()
    s.calculate_trajectory()
    t, p, ypp = s.get_trajectory_arrays()
    print("halving box (backward %s): %d splits, depth %d; knots %.3e" % (backward, n, depth, _knot_err(p, T.p)))
    assert n > 0 and np.array_equal(t, T.t) and np.array_equal(p, T.p)


@gpu
@pytest.mark.parametrize("backward", [False, True])
def test_step_halving(backward):
    """RK4 steps that meet |beta| >= 1 are halved on the device as the
    reference halves them (src/OSCARSSR.cc:1362-1378; the oracle's restatement
    is pinned in tests/test_oracle_vs_ref.py::test_step_halving): 8 steps of
    1 m through a 40 T bump, thousands of sub-steps, 14 levels deep.  Knots
    within 1e-9 of each channel's largest value: the bump turns the electron
    through several radians, so last-bit differences of the device's exp and
    the C library's grow along the trajectory."""
    from tests.common import halving_fields
    o, s = _pair(lambda x: halving_fields(x, backward=backward))
    T = o.calculate_trajectory()
    n, depth = o.halvings()
    s.calculate_trajectory()
    t, p, ypp = s.get_trajectory_arrays()
    assert np.array_equal(t, T.t)
    ek = _knot_err(p, T.p)
    print("step halving (backward %s): %d splits, depth %d; knots %.3e" % (backward, n, depth, ek))
    assert n > 0 and ek < 1e-9
