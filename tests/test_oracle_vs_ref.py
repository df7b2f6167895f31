"""The C oracle against the reference's compiled C++ core (oracle/_ref), where
it was built from the reference sources.  Every comparison is bit-exact,
including the thread-partition-dependent reporting of levels (carry_quirk)."""
import math

import numpy as np
import pytest

from tests.common import planar, epu

pytestmark = pytest.mark.skipif(not __import__("oracle.pyoracle", fromlist=["x"]).ref_available(),
                                reason="reference core not built (oracle/_ref/libref.so)")


def pair(setup, nthreads=3, **kw):
    from oracle.pyoracle import Oracle, RefSim
    o = Oracle(nthreads=nthreads)
    r = RefSim(nthreads=nthreads)
    setup(o, **kw)
    setup(r, **kw)
    return o, r


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_particle_and_fields():
    o, r = pair(planar)
    assert o.particle() == {k: (list(map(float, v)) if isinstance(v, list) else float(v))
                            for k, v in r.particle().items()}
    for z in np.linspace(-0.9, 0.9, 37):
        for x in (0.0, 1e-3):
            assert same(o.get_bfield([x, 0, z]), r.get_bfield([x, 0, z]))


@pytest.mark.parametrize("npts", [1001, 2000])
def test_trajectory_backward_and_forward(npts):
    """T0 inside the window: points both before and after T0."""
    def setup(s):
        s.set_particle_beam(energy_GeV=2.5, x0=[0.001, 0, -0.2], d0=[0.0005, 0, 1], current=0.3)
        s.add_bfield_undulator(bfield=[0, 0.8, 0.05], period=[0, 0, 0.05], nperiods=10, phase=0.3, taper=0.01)
        s.set_ctstartstop(-0.5, 0.6)
        s.set_npoints_trajectory(npts)
    o, r = pair(setup)
    T = o.calculate_trajectory()
    t, p, ypp = r.calculate_trajectory()
    assert same(T.t, t) and same(T.p, p) and same(T.ypp, ypp)


def test_epu_trajectory_and_spectrum():
    o, r = pair(epu, npoints=8001)
    T = o.calculate_trajectory()
    t, p, ypp = r.calculate_trajectory()
    assert same(T.t, t) and same(T.p, p)
    for pol in ("", "cl", "cr", "lv"):
        e, v, lev, conv, ef = o.calculate_spectrum([0.0005, 0.0002, 20], npoints=23, energy_range_eV=[200, 900],
                                                   polarization=pol)
        re, rv, rconv = r.calculate_spectrum([0.0005, 0.0002, 20], npoints=23, energy_range_eV=[200, 900],
                                             polarization=pol)
        assert same(v, rv), pol


@pytest.mark.parametrize("plane,rot", [("XY", (0, 0, 0)), ("YX", (0, 0, 0.2)), ("XZ", (math.pi / 2, 0, 0)),
                                       ("ZY", (0, 0.1, 0))])
def test_flux_planes_and_rotations(plane, rot):
    o, r = pair(planar, npoints=8001)
    kw = dict(plane=plane, rotations=rot, translation=[0.001, 0, 30])
    cont, v, lev, conv, ef, obs = o.calculate_flux_rectangle(300.0, [5, 4], [0.01, 0.008], **kw)
    xyz, rv, rconv = r.calculate_flux_rectangle(300.0, [5, 4], [0.01, 0.008], **kw)
    assert same(cont, xyz) and same(v, rv)


@pytest.mark.parametrize("angle", [0.0, 0.7])
def test_linear_polarization_angle(angle):
    o, r = pair(planar, npoints=8001)
    e, v, _, _, _ = o.calculate_spectrum([0, 0.0003, 30], npoints=11, energy_range_eV=[140, 170], angle=angle)
    re, rv, _ = r.calculate_spectrum([0, 0.0003, 30], npoints=11, energy_range_eV=[140, 170], angle=angle)
    assert same(v, rv)


@pytest.mark.parametrize("nthreads", [1, 3])
@pytest.mark.parametrize("quantity", ["flux", "level", "precision"])
def test_nonconverged_carry(nthreads, quantity):
    """max_level below the convergence threshold: the reported level/precision
    follows the reference's per-thread carry, which depends on nthreads."""
    o, r = pair(planar, nthreads=nthreads, npoints=8001)
    kw = dict(max_level=8, quantity=quantity)
    e, v, lev, conv, _ = o.calculate_spectrum([0, 0, 30], npoints=17, energy_range_eV=[100, 500], **kw)
    re, rv, rconv = r.calculate_spectrum([0, 0, 30], npoints=17, energy_range_eV=[100, 500], **kw)
    assert same(v, rv)
    assert bool(conv.all()) == rconv
    cont, v, lev, conv, _, _ = o.calculate_flux_rectangle(200.0, [4, 3], [0.01, 0.01], translation=[0, 0, 30],
                                                          max_level=9, **{"quantity": quantity})
    xyz, rv, rconv = r.calculate_flux_rectangle(200.0, [4, 3], [0.01, 0.01], translation=[0, 0, 30], max_level=9,
                                                quantity=quantity)
    assert same(v, rv)


def test_extended_levels():
    """Levels above max_level are 'extended' (spacing T/(2^L - 1))."""
    o, r = pair(planar, npoints=8001)
    kw = dict(max_level=9, max_level_extended=12)
    e, v, lev, conv, _ = o.calculate_spectrum([0, 0, 30], npoints=9, energy_range_eV=[1000, 2000], **kw)
    re, rv, _ = r.calculate_spectrum([0, 0, 30], npoints=9, energy_range_eV=[1000, 2000], **kw)
    assert same(v, rv)


@pytest.mark.parametrize("normal", [0, -1, 1])
def test_power_density_directional(normal):
    o, r = pair(epu, npoints=8001)
    kw = dict(translation=[0.01, 0, 20], rotations=(0.2, 0.3, 0), normal=normal)
    cont, v, lev, conv, obs = o.calculate_power_density_rectangle([6, 5], [0.08, 0.08], **kw)
    xyz, rv, rconv = r.calculate_power_density_rectangle([6, 5], [0.08, 0.08], **kw)
    assert same(cont, xyz) and same(v, rv)


def test_grid_file(tmp_path):
    path = str(tmp_path / "g.txt")
    z = np.linspace(-1, 1, 801)
    by = np.where(np.abs(z) < 0.8, np.sin(2 * np.pi * z / 0.05), 0.0)
    with open(path, "w") as f:
        f.write("# test map\n0\n0\n1\n0\n0\n1\n-1 # z0\n0.0025\n801\n")
        for b in by:
            f.write("0 %.17g 0\n" % b)

    def setup(s):
        s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
        s.add_bfield_file(path)
        s.set_ctstartstop(0, 2)
        s.set_npoints_trajectory(3001)
    o, r = pair(setup)
    T = o.calculate_trajectory()
    t, p, ypp = r.calculate_trajectory()
    assert same(T.p, p)
    e, v, _, _, _ = o.calculate_spectrum([0, 0, 30], npoints=9, energy_range_eV=[300, 900])
    re, rv, _ = r.calculate_spectrum([0, 0, 30], npoints=9, energy_range_eV=[300, 900])
    assert same(v, rv)


PD_POINTS = [[[x, y, 20.0], [0.1 * x, 0.2, 1.0]] for x in (-0.03, 0.0, 0.012) for y in (-0.01, 0.004)] + \
    [[[0.0, 0.0, 20.0], [0.0, 0.0, -1.0]]]


@pytest.mark.parametrize("normal", [0, 1, -1])
def test_power_density_points(normal):
    """calculate_power_density on points with normals: the product's point set-up
    (rotate, reverse, translate, unit normal) and the oracle's per-point-normal
    sum against the reference, bit for bit."""
    from oscars_amd.sr import surface_points
    o, r = pair(epu, npoints=8001)
    kw = dict(rotations=[0.01, -0.02, 0.3], translation=[0.001, 0, 0.5])
    xyz, rv, _ = r.calculate_power_density(PD_POINTS, normal=normal, **kw)
    obs, nrm = surface_points(PD_POINTS, normal, **kw)
    assert same(obs, xyz)
    ov, _, _ = o.power_density_surface(np.array(obs), np.array(nrm), directional=normal != 0)
    assert same(ov, rv)


@pytest.mark.parametrize("dim", [1, 3])
def test_power_density_line(dim):
    """A line of points has no normals: N1.normal is taken as 1."""
    o, r = pair(planar, npoints=4001)
    xyz, rv, _ = r.calculate_power_density_line([-0.01, 0.0, 30], [0.012, 0.002, 30], 12, dim=dim)
    obs = [[-0.01 + i * ((0.012 - -0.01) / 11), 0.0 + i * (0.002 / 11), 30.0 + i * ((30 - 30) / 11)]
           for i in range(12)]
    ov, _, _ = o.power_density_surface(np.array(obs), None)
    assert same(ov, rv)
    if dim == 3:
        assert same(xyz, obs)
    else:
        c = np.array(obs[0]) + (np.array(obs[-1]) - np.array(obs[0])) / 2
        d = [(-1. if i < 6 else 1.) * math.sqrt(float(np.sum((np.array(obs[i]) - c) ** 2))) for i in range(12)]
        assert np.allclose(xyz[:, 0], d, rtol=1e-15, atol=0) and not xyz[:, 1:].any()


@pytest.mark.parametrize("setup", [planar, epu])
def test_total_power(setup):
    o, r = pair(setup, npoints=4001)
    p, prec, lev = o.total_power()
    assert p == r.calculate_total_power()
    assert prec == r.calculate_total_power(quantity="precision")
    assert float(lev) == r.calculate_total_power(quantity="level")
    # coarse max_level: the reference raises, the oracle reports level -1
    assert o.total_power(max_level=5)[2] == -1
    with pytest.raises(ValueError, match="did not converge"):
        r.calculate_total_power(max_level=5)


@pytest.mark.parametrize("setup", [planar, epu])
def test_efield_vs_time(setup):
    """Time-domain electric field at one observer: one value per knot, time
    i dt + |R| / c with dt the integration step."""
    o, r = pair(setup, npoints=3001)
    o.calculate_trajectory()
    dt = ((o.ctstop - o.ctstart) / 2.99792458e8 / float(3001 - 1))
    for obs in ([0, 0, 30], [0.003, -0.001, 20]):
        rt, re = r.calculate_efield_vs_time(obs)
        ot, oe = o.efield_time(obs, dt)
        assert same(ot, rt) and same(oe, re)


def test_power_density_stl(tmp_path):
    """STL surfaces: the product's reader gives the reference's triangles and
    centres, the oracle's occluded power density equals the reference's bit
    for bit, the blocker shadows part of the screen, and the STL writer
    matches the reference's file after the 80-byte header (whose tail the
    reference leaves uninitialised)."""
    from oscars_amd import io as IO
    from tests.common import stl_scene
    scene = stl_scene(str(tmp_path))
    o, r = pair(planar, npoints=2001)
    for k, (path, scale, tr) in enumerate(scene):
        r.add_stl(path, scale=scale, translation=tr, name="s%d" % k)
    tri_r, v_r = r.calculate_power_density_stl(max_level=13, stlofile=str(tmp_path / "ref.stl"))
    tri = np.concatenate([IO.read_stl(p, s, t) for p, s, t in scene])
    assert same(tri[:, :9].reshape(-1, 3, 3), tri_r)
    o.calculate_trajectory()
    v_o, lev, _ = o.power_density_stl(tri, max_level=13)
    assert same(v_o, v_r)
    v_screen, _, _ = o.power_density_stl(tri[:36], max_level=13)
    shadowed = v_o[:36] < v_screen * (1 - 1e-6)
    assert 0 < shadowed.sum() < 36 and same(v_o[:36][~shadowed], v_screen[~shadowed])
    IO.write_stl(str(tmp_path / "mine.stl"), tri, [s for p, s, t in scene for _ in range(len(IO.read_stl(p)))])
    a, b = (tmp_path / "ref.stl").read_bytes(), (tmp_path / "mine.stl").read_bytes()
    assert a[80:] == b[80:] and a[:37] == b[:37]


@pytest.mark.parametrize("normal", [0, -1])
def test_three_corner_rectangle(normal):
    """x0x1x2 surfaces: the product's points, normal and stored coordinates
    against the reference's rectangle, and the oracle's power density on them."""
    from oscars_amd import geometry as G
    o, r = pair(epu, npoints=4001)
    corners = [[-0.03, -0.02, 20.0], [0.03, -0.01, 20.5], [-0.025, 0.02, 19.8]]
    kw = dict(rotations=[0.01, 0.2, -0.1], translation=[0.002, 0, 0.3])
    r.set_corners(corners)
    try:
        for dim in (2, 3):
            xyz, rv, _ = r.calculate_power_density_rectangle([5, 4], [1, 1], normal=normal, dim=dim, **kw)
            obs, nv, cont = G.parallelogram(5, 4, *corners, kw["rotations"], kw["translation"], normal, dim)
            assert same(cont, xyz)
            ov, _, _ = o.power_density_points(np.array(obs), nv, directional=normal != 0)
            assert same(ov, rv)
    finally:
        r.set_corners(None)


def test_gaussian_uniform_quadrupole_fields():
    """The oracle's Gaussian, uniform-box and quadrupole models: fields at
    points inside and outside each, the trajectory and a spectrum, bit for bit."""
    from tests.common import other_fields
    o, r = pair(other_fields)
    for x in ([0.001, 0.002, -0.3], [0.0003, -0.0001, 0.41], [0.002, 0.003, 0.1], [0.002, 0.003, 0.16],
              [0.0, 0.0, 0.9]):
        assert same(o.get_bfield(x), r.get_bfield(x))
    rt, rp, rypp = r.calculate_trajectory()
    T = o.calculate_trajectory()
    assert same(T.t, rt) and same(T.p, rp) and same(T.ypp, rypp)
    e, v, lev, conv, _ = o.calculate_spectrum([0, 0, 30], npoints=20, energy_range_eV=[10, 2000])
    re, rv, _ = r.calculate_spectrum([0, 0, 30], npoints=20, energy_range_eV=[10, 2000])
    assert same(v, rv)


@pytest.mark.parametrize("magnetic", [True, False])
def test_electric_fields(magnetic):
    """Electric fields in the equations of motion (with and without magnetic
    fields): trajectory and spectrum bit for bit."""
    from tests.common import with_efields
    o, r = pair(with_efields, magnetic=magnetic)
    rt, rp, rypp = r.calculate_trajectory()
    T = o.calculate_trajectory()
    assert same(T.t, rt) and same(T.p, rp) and same(T.ypp, rypp)
    e, v, lev, conv, _ = o.calculate_spectrum([0, 0, 30], npoints=8, energy_range_eV=[100, 800])
    re, rv, _ = r.calculate_spectrum([0, 0, 30], npoints=8, energy_range_eV=[100, 800])
    assert same(v, rv)


def test_removed_efields_keep_the_eb_equations():
    """The reference picks its equations of motion when a field is added or
    cleared: after its electric fields are removed an undulator is still
    integrated with the E+B equations, which is the oracle's E+B integration
    with one zero electric field."""
    from tests.common import UNDULATOR
    from oracle.pyoracle import RefSim
    r = RefSim()
    r.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    r.add_bfield_undulator(**UNDULATOR)
    r.add_efield_uniform(efield=[0, 5e5, 1e4], width=[0, 0, 0.3], translation=[0, 0, 0.5])
    r.remove_efield("")
    r.set_ctstartstop(0, 2)
    r.set_npoints_trajectory(4001)
    rt, rp, _ = r.calculate_trajectory()
    from oracle.pyoracle import Oracle
    o = Oracle(nthreads=1)
    o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_undulator(**UNDULATOR)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(4001)
    plain = o.calculate_trajectory().p.copy()
    o.add_efield_uniform(efield=[0, 0, 0])
    T = o.calculate_trajectory()
    assert same(T.t, rt) and same(T.p, rp)
    assert not np.array_equal(plain, rp)  # the B-only equations differ in the last bits


@pytest.mark.parametrize("electric", [False, True])
def test_time_dependent_fields(electric):
    """Fields with frequency / frequency_phase / time_offset
    (src/TField3D_IdealUndulator.cc:147-151, TField3D_Gaussian.cc:107-111,
    TField3D_UniformBox.cc:115-119, TField3D_Quadrupole.cc:78-82): the field
    at points and times, the trajectory (whose derivatives see the RK4 stage
    times) and a spectrum, bit for bit."""
    from tests.common import time_dependent
    o, r = pair(time_dependent, electric=electric)
    for t in (0.0, 1.3e-9, 4.1e-9):
        for x in ([0.0, 0.0, -0.7], [0.0004, 0.0, 0.02], [0.0, 0.001, 0.75], [0.001, 0.002, 0.9]):
            assert same(o.get_bfield(x, t), r.get_bfield(x, t))
            assert same(o.get_efield(x, t), r.get_efield(x, t))
    assert not same(o.get_bfield([0, 0, 0.02], 0.0), o.get_bfield([0, 0, 0.02], 1.3e-9))
    rt, rp, rypp = r.calculate_trajectory()
    T = o.calculate_trajectory()
    assert same(T.t, rt) and same(T.p, rp) and same(T.ypp, rypp)
    e, v, lev, conv, _ = o.calculate_spectrum([0, 0, 30], npoints=12, energy_range_eV=[100, 900])
    re, rv, _ = r.calculate_spectrum([0, 0, 30], npoints=12, energy_range_eV=[100, 900])
    assert same(v, rv)


def test_time_dependent_grid(tmp_path):
    """A grid map always carries cos(2 pi f (t + t0) + phase)
    (src/TField3D_Grid.cc:277), so frequency_phase alone scales it by
    cos(phase); with a frequency it oscillates."""
    from tests.common import undulator_like, write_oscars_grid
    path = str(tmp_path / "g.txt")
    write_oscars_grid(path, [0, 0, -0.5], [0, 0, 0.0025], [1, 1, 401], undulator_like)
    for f, ph, t0 in ((0.0, 0.4, 0.0), (9e7, 0.4, 1e-9)):
        from oracle.pyoracle import Oracle, RefSim
        o, r = Oracle(nthreads=3), RefSim(nthreads=3)
        for s in (o, r):
            s.set_time_dependence(f, ph, t0)
            s.add_bfield_file(path)
            s.set_time_dependence(0.0, 0.0, 0.0)
            s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
            s.set_ctstartstop(0, 2)
            s.set_npoints_trajectory(2001)
        for t in (0.0, 2e-9):
            for z in (-0.3, 0.01, 0.2):
                assert same(o.get_bfield([0, 0, z], t), r.get_bfield([0, 0, z], t))
        rt, rp, rypp = r.calculate_trajectory()
        T = o.calculate_trajectory()
        assert same(T.t, rt) and same(T.p, rp) and same(T.ypp, rypp)


@pytest.mark.parametrize("backward", [False, True])
@pytest.mark.parametrize("case", ["bump", "box"])
def test_step_halving(backward, case):
    """RK4 steps that meet |beta| >= 1 are halved recursively
    (src/OSCARSSR.cc:1362-1378), the stored acceleration being the last slope
    the halving computed: the trajectory bit for bit, and it differs from the
    oracle's count of split steps is checked to be nonzero."""
    from tests.common import halving_box, halving_fields
    o, r = pair(halving_fields if case == "bump" else halving_box, backward=backward)
    rt, rp, rypp = r.calculate_trajectory()
    T = o.calculate_trajectory()
    n, depth = o.halvings()
    assert n > 0 and depth < 40
    assert same(T.t, rt) and same(T.p, rp) and same(T.ypp, rypp)
    beta = np.sqrt(np.sum(T.p[:, 3:6] ** 2, axis=1))
    assert np.all(beta < 1)
