"""Total power and power density on surfaces of points on the GPU against the
oracle (pinned to the reference in tests/test_oracle_vs_ref.py).

Total power: the integrand carries 1/(1 - B.B)^3 with 1 - B.B ~ 1/gamma^2
~ 3e-8 at 3 GeV, so a last-bit difference in beta becomes ~1e-8 of a sample.
Against the oracle on its own trajectory the bar is therefore 1e-9 (the
device's RK4 differs from glibc's in the last bits of sin); against the oracle
summing the device's own trajectory it is 1e-13 (left: the tree order of the
device's level sums and x*x*x for pow(x, 3)).  The stopping level must agree.
Power density: 1e-9 relative with a floor at 1e-6 of the maximum, levels
identical."""
import numpy as np
import pytest

from tests.common import epu, planar, rel_err

gpu = pytest.mark.gpu


def _pair(setup, npoints):
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    o = Oracle(nthreads=8)
    setup(o, npoints=npoints)
    o.calculate_trajectory()
    s = sr()
    setup(s, npoints=npoints)
    return o, s


@gpu
@pytest.mark.parametrize("setup", [planar, epu])
def test_total_power(setup):
    o, s = _pair(setup, 4001)
    p, prec, lev = o.total_power()
    gp = s.calculate_total_power()
    gl = s.calculate_total_power(quantity="level")
    gr = s.calculate_total_power(quantity="precision")
    print("total power %s: %.12g W (oracle %.12g), level %d, precision %.3e (oracle %.3e)"
          % (setup.__name__, gp, p, gl, gr, prec))
    assert gl == lev
    assert abs(gp - p) / abs(p) < 1e-9
    assert abs(gr - prec) <= 1e-6 * abs(prec) + 1e-15
    from oracle.pyoracle import Trajectory
    t, pk, ypp = s._device().trajectory(0)
    o.traj = Trajectory(t, pk, ypp)
    p2, _, lev2 = o.total_power()
    print("  on the device's trajectory: oracle %.15g, rel diff %.3e" % (p2, abs(gp - p2) / abs(p2)))
    assert lev2 == gl and abs(gp - p2) / abs(p2) < 1e-13
    with pytest.raises(ValueError, match="did not converge"):
        s.calculate_total_power(max_level=5)
    with pytest.raises(ValueError, match="'quantity' must be"):
        s.calculate_total_power(quantity="flux")


POINTS = [[[x, y, 20.0], [0.1 * x, 0.2, 1.0]] for x in np.linspace(-0.03, 0.03, 7) for y in (-0.01, 0.0, 0.004)] + \
    [[[0.0, 0.0, 20.0], [0.0, 0.0, -1.0]]]


@gpu
@pytest.mark.parametrize("normal", [0, 1, -1])
def test_power_density_points(normal):
    from oscars_amd.sr import surface_points
    o, s = _pair(epu, 8001)
    kw = dict(rotations=[0.01, -0.02, 0.3], translation=[0.001, 0, 0.5])
    res = s.calculate_power_density(POINTS, normal=normal, **kw)
    obs, nrm = surface_points(POINTS, normal, **kw)
    assert [r[0] for r in res] == obs
    v = np.array([r[1] for r in res])
    ov, olev, _ = o.power_density_surface(np.array(obs), np.array(nrm), directional=normal != 0,
                                          carry_quirk=False)
    e = rel_err(v, ov, 1e-6) if np.any(ov) else float(np.max(np.abs(v)))
    print("power density on points, normal %d: %.3e" % (normal, e))
    assert e < 1e-9
    assert np.array_equal(s._last_levels, olev)


@gpu
@pytest.mark.parametrize("dim", [1, 3])
def test_power_density_line(dim):
    o, s = _pair(planar, 4001)
    res = s.calculate_power_density_line(x1=[-0.01, 0.0, 30], x2=[0.012, 0.002, 30], npoints=41, dim=dim)
    obs = [[-0.01 + i * ((0.012 - -0.01) / 40), 0.0 + i * (0.002 / 40), 30.0 + i * ((30 - 30) / 40)]
           for i in range(41)]
    ov, olev, _ = o.power_density_surface(np.array(obs), None, carry_quirk=False)
    v = np.array([r[1] for r in res])
    e = rel_err(v, ov, 1e-6)
    print("power density on a line (dim %d): %.3e" % (dim, e))
    assert e < 1e-9
    if dim == 3:
        assert [r[0] for r in res] == obs


@gpu
@pytest.mark.parametrize("setup", [planar, epu])
def test_efield_vs_time(setup):
    """Time-domain E-field, relative to the largest component: against the
    oracle on the device's own trajectory to 1e-13 (the oracle's pow(x, 3)
    against x*x*x, otherwise the same operations; measured 3e-16), and on the
    oracle's trajectory to 1e-9 (measured 1e-14); times to 1e-15."""
    from oracle.pyoracle import Trajectory
    o, s = _pair(setup, 3001)
    dt = ((s.get_ctstop() - s.get_ctstart()) / 2.99792458e8 / float(3001 - 1))
    for obs in ([0, 0, 30], [0.003, -0.001, 20]):
        res = s.calculate_efield_vs_time(obs)
        gt = np.array([r[0] for r in res])
        ge = np.array([r[1] for r in res])
        ot, oe = o.efield_time(obs, dt)
        scale = np.max(np.abs(oe))
        e1 = float(np.max(np.abs(ge - oe)) / scale)
        t, pk, ypp = s._device().trajectory(0)
        o2 = o.traj
        o.traj = Trajectory(t, pk, ypp)
        ot2, oe2 = o.efield_time(obs, dt)
        o.traj = o2
        e2 = float(np.max(np.abs(ge - oe2)) / scale)
        print("efield vs time %s %s: %.3e (oracle trajectory), %.3e (device trajectory)" % (setup.__name__, obs, e1,
                                                                                             e2))
        assert np.max(np.abs(gt - ot2) / ot2) < 1e-15
        assert e2 < 1e-13 and e1 < 1e-9


@gpu
def test_power_density_stl(tmp_path):
    """STL screen with a blocker: on the device's own trajectory the oracle and
    the device take the same occlusion decisions and agree to 1e-12 with the
    same levels; on the oracle's trajectory to 1e-9 (a ray grazing a triangle
    edge could decide differently there; none does in this scene)."""
    from oracle.pyoracle import Trajectory
    from tests.common import stl_scene
    scene = stl_scene(str(tmp_path))
    o, s = _pair(planar, 2001)
    for k, (path, scale, tr) in enumerate(scene):
        s.add_stl(path, scale=scale, translation=tr, name="s%d" % k)
    res = s.calculate_power_density_stl(max_level=13)
    v = np.array([r[1] for r in res])
    lev = s._last_levels.copy()
    from oscars_amd import io as IO
    tri = np.concatenate([IO.read_stl(p, sc, t) for p, sc, t in scene])
    ov, olev, _ = o.power_density_stl(tri, max_level=13, carry_quirk=False)
    e1 = rel_err(v, ov, 1e-6)
    t, pk, ypp = s._device().trajectory(0)
    o.traj = Trajectory(t, pk, ypp)
    ov2, olev2, _ = o.power_density_stl(tri, max_level=13, carry_quirk=False)
    e2 = rel_err(v, ov2, 1e-6)
    print("power density on STL (%d triangles): %.3e (oracle trajectory), %.3e (device trajectory), levels %s"
          % (len(v), e1, e2, sorted(set(lev.tolist()))))
    assert np.array_equal(lev, olev2) and e2 < 1e-12
    assert e1 < 1e-9
    with pytest.raises(ValueError, match="already exists"):
        s.add_stl(scene[0][0], name="s0")


@gpu
def test_three_corner_rectangles():
    """x0x1x2 surfaces for the power density and the flux: points from
    geometry.parallelogram (pinned to the reference), values against the oracle."""
    from oscars_amd import geometry as G
    o, s = _pair(epu, 8001)
    corners = [[-0.03, -0.02, 20.0], [0.03, -0.01, 20.5], [-0.025, 0.02, 19.8]]
    kw = dict(rotations=[0.01, 0.2, -0.1], translation=[0.002, 0, 0.3])
    res = s.calculate_power_density_rectangle(npoints=[7, 5], x0x1x2=corners, normal=-1, dim=3, **kw)
    obs, nv, cont = G.parallelogram(7, 5, *corners, kw["rotations"], kw["translation"], -1, 3)
    assert [r[0] for r in res] == cont
    ov, _, _ = o.power_density_points(np.array(obs), nv, directional=True, carry_quirk=False)
    e = rel_err(np.array([r[1] for r in res]), ov, 1e-6)
    small = [[-0.002, -0.002, 20.0], [0.002, -0.0015, 20.0], [-0.0018, 0.002, 20.0]]
    res = s.calculate_flux_rectangle(energy_eV=700.0, npoints=[6, 5], x0x1x2=small, mode="plain")
    obs, _, cont = G.parallelogram(6, 5, *small)
    assert [r[0] for r in res] == cont
    fv, _, _, _ = o.flux_points(np.array(obs), 700.0, carry_quirk=False)
    e2 = rel_err(np.array([r[1] for r in res]), fv, 1e-6)
    print("three-corner surfaces: power density %.3e, flux %.3e" % (e, e2))
    assert e < 1e-9 and e2 < 1e-9


@gpu
@pytest.mark.parametrize("scene", ["sphere", "grazing"])
def test_power_density_stl_hierarchy_equals_every_triangle(tmp_path, scene):
    """The bounding-volume hierarchy finds the same hits as testing every
    triangle for every ray (mode 'plain', the reference's loop): values and
    levels bit for bit.  'sphere': a 1280-triangle sphere of 10 mm at 5 m
    (the layout of the reference's CAD example), its far side shadowed by its
    near side; 'grazing': a plate in the beam's plane beside a screen, its
    triangles nearly parallel to the rays that pass them."""
    import time
    from oscars_amd.sr import sr
    from tests.common import stl_icosphere, stl_plate, write_stl_triangles
    s = sr()
    planar(s, npoints=4001)
    if scene == "sphere":
        path = str(tmp_path / "sphere.stl")
        write_stl_triangles(path, stl_icosphere(10.0, 3))
        s.add_stl(path, scale=1e-3, translation=[0, 0, 5])
    else:
        side = [(n, [a[0], a[2], a[1]], [b[0], b[2], b[1]], [c[0], c[2], c[1]])
                for n, a, b, c in stl_plate(-4, 4, 1000, 20000, 0.3, 8, 40)]  # the plane y = 0.3 mm, along z
        screen = stl_plate(-6, 6, -3, 3, 20000, 12, 6)
        path = str(tmp_path / "grazing.stl")
        write_stl_triangles(path, side + screen)
        s.add_stl(path, scale=1e-3)
    out = {}
    for mode in ("plain", "tiled"):
        t0 = time.perf_counter()
        res = s.calculate_power_density_stl(max_level=11, mode=mode)
        dt = time.perf_counter() - t0
        out[mode] = (np.array([r[1] for r in res]), s._last_levels.copy(), s.last_stats["kernel_ms"])
        print("%s, %d triangles, %s: %.1f ms (kernels %.1f ms)" % (scene, len(res), mode, 1e3 * dt,
                                                                    out[mode][2]))
    v0, l0, _ = out["plain"]
    v1, l1, _ = out["tiled"]
    shadowed = int(np.sum(v0 == 0))
    print("%s: %d of %d centres see no unblocked sample" % (scene, shadowed, len(v0)))
    assert np.array_equal(l0, l1)
    assert np.array_equal(v0, v1)
    assert 0 < np.count_nonzero(v0) < len(v0) or scene == "grazing"
