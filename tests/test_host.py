"""Host-side checks that need no GPU: the C ABI library loads and exports every
function include/oscars_mi.h declares; the Python drop-in's geometry is
bit-identical to the oracle's; argument errors raise the reference's
ValueError messages; compute calls fail loudly without a device."""
import math
import os
import re

import numpy as np
import pytest

from tests.common import planar

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "oscars_mi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(om_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    import ctypes
    from oscars_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
        assert isinstance(getattr(L, n), ctypes._CFuncPtr)
    assert set(_lib.EXPORTS) <= set(names)


def test_no_device_is_an_error():
    from oscars_amd import _lib
    if _lib.lib().om_device_count() > 0:
        pytest.skip("a device is visible")
    from oscars_amd.sr import sr
    s = sr()
    planar(s)
    with pytest.raises(ValueError, match="GPU"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200])


@pytest.mark.parametrize("plane", ["XY", "YX", "XZ", "ZX", "YZ", "ZY", "xy"])
@pytest.mark.parametrize("npts", [(7, 5), (1, 4), (3, 1), (1, 1)])
@pytest.mark.parametrize("normal,dim", [(0, 2), (-1, 3), (1, 2)])
def test_rectangle_bit_identical(plane, npts, normal, dim):
    from oracle.pyoracle import Oracle
    from oscars_amd import geometry as G
    rot = (0.1, -0.2, 0.3)
    tr = (0.001, -0.002, 30.0)
    obs, nv, cont = G.rectangle(plane, npts[0], npts[1], 0.02, 0.013, rot, tr, normal, dim)
    o = Oracle()
    oo, onv, oc = o.rectangle(plane, npts, (0.02, 0.013), rot, tr, normal, dim)
    assert np.array_equal(np.array(obs), oo)
    assert np.array_equal(np.array(cont), oc)
    assert np.array_equal(np.array(nv), onv)


def test_energies_polarization_particle_bit_identical():
    from oracle.pyoracle import Oracle, polarization_vector
    from oscars_amd import geometry as G
    o = Oracle()
    for n, a, b in ((2000, 100, 2000), (512, 100, 3000), (7, 1.5, 3.25), (1, 5.0, 9.0)):
        assert np.array_equal(np.array(G.energies(n, a, b)), o.energies(n, a, b))
    for pol, ang in (("lh", 0), ("lv", 0), ("cl", 0), ("cr", 0), ("linear", 0.37), ("circular-left", 0),
                     ("all", 0)):
        for h, p in (((1, 0, 0), (0, 0, 1)), ((0.6, 0.8, 0), (0.1, 0.2, 1.0))):
            assert np.array_equal(np.array(G.polarization_vector(pol, ang, h, p)), polarization_vector(pol, ang, h, p))
    for e, d0 in ((3.0, (0, 0, 1)), (2.5, (0.001, -0.002, 1)), (0.2, (1, 1, 1))):
        p = G.ideal_particle("electron", e, [0, 0.001, -1], d0, 0.0, 0.4)
        o.set_particle_beam(energy_GeV=e, x0=[0, 0.001, -1], d0=d0, current=0.4)
        q = o.particle()
        assert p.b0 == q["b0"] and p.gamma == q["gamma"] and p.qomg == q["qomg"]


def test_reference_error_messages():
    from oscars_amd.sr import sr
    s = sr()
    with pytest.raises(ValueError, match="No particle beam defined"):
        s.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], width=[0.01, 0.01], translation=[0, 0, 30])
    planar(s)
    with pytest.raises(ValueError, match="'quantity' must be"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200], quantity="bogus")
    with pytest.raises(ValueError, match="Polarization requested not recognized"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200], polarization="bogus")
    with pytest.raises(ValueError, match="cannot specify 'all' and 'angle'"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200], polarization="all", angle=0.3)
    with pytest.raises(ValueError, match="an entry in 'npoints' is <= 0"):
        s.calculate_flux_rectangle(energy_eV=100, npoints=[0, 3], width=[0.01, 0.01], translation=[0, 0, 30])
    with pytest.raises(ValueError, match="'dim' must be 2 or 3"):
        s.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], width=[0.01, 0.01], dim=4)
    with pytest.raises(ValueError, match="'normal' must be -1, 0, or 1"):
        s.calculate_power_density_rectangle(npoints=[3, 3], width=[0.01, 0.01], normal=2)
    with pytest.raises(ValueError, match="gpu is 1 and nthreads > 0"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200], gpu=1, nthreads=2)
    with pytest.raises(ValueError, match="'gpu' must be 0 or 1"):
        s.calculate_spectrum(obs=[0, 0, 30], npoints=5, energy_range_eV=[100, 200], gpu=3)
    with pytest.raises(ValueError, match="not a valid surface"):
        s.calculate_flux_rectangle(energy_eV=100, npoints=[3, 3], width=[0.01, 0.01], plane="XQ")


def test_oscars_field_file_reader(tmp_path):
    from oscars_amd.sr import read_oscars_field_file
    from oracle.pyoracle import read_oscars_grid
    path = os.path.join(ROOT, "tests", "golden", "grid_field.txt")
    a = read_oscars_field_file(path)
    b = read_oscars_grid(path)
    assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2]
    assert np.array_equal(a[3], b[3])
    bad = tmp_path / "short.txt"
    bad.write_text("# x\n0\n0\n1\n0\n0\n1\n0\n0.1\n5\n0 1 0\n")
    with pytest.raises(ValueError):
        read_oscars_field_file(str(bad))


def test_beam_sampling_statistics():
    """Gaussian beam sampling follows the reference's formulas: moments at x0
    after drifting the Twiss parameters from the lattice reference point."""
    from oscars_amd import beam as B
    from oscars_amd import geometry as G
    p0 = G.ideal_particle("electron", 3.0, [0, 0, 0], [0, 0, 1], 0.0, 0.5)
    b = dict(q=p0.q, m=p0.m, energy_GeV=3.0, x0=[0, 0, 0], d0=[0, 0, 1], t0=0.0, current=0.5,
             sigma_energy_GeV=0.0027, beta=[1.5, 0.8], alpha=[0.0, 0.0], emittance=[0.9e-9, 8e-12])
    ps = B.sample_particles(b, 4000, seed=7)
    x = np.array([p.x0[0] for p in ps])
    y = np.array([p.x0[1] for p in ps])
    xp = np.array([p.b0[0] / p.b0[2] for p in ps])
    e = np.array([p.energy_GeV for p in ps])
    assert np.std(x) == pytest.approx(math.sqrt(0.9e-9 * 1.5), rel=0.05)
    assert np.std(y) == pytest.approx(math.sqrt(8e-12 * 0.8), rel=0.05)
    assert np.std(xp) == pytest.approx(math.sqrt(0.9e-9 / 1.5), rel=0.05)
    assert np.std(e) == pytest.approx(0.0027, rel=0.05)
    assert abs(np.corrcoef(x, xp)[0, 1]) < 0.05


def test_mode_keywords_and_halbach():
    """The power-density extensions reject unknown modes before any GPU work;
    add_bfield_halbach (the reference's model returns zero everywhere,
    src/TField3D_Halbach.cc:58-65) adds nothing to the field list."""
    from oscars_amd.sr import sr
    s = sr()
    planar(s)
    n = len(s._fields)
    s.add_bfield_halbach(bfield=[0, 1, 0], period=[0, 0, 0.05], nperiods=10)
    assert len(s._fields) == n
    with pytest.raises(ValueError, match="mode must be"):
        s.calculate_power_density_rectangle(npoints=[3, 3], width=[0.01, 0.01], mode="fast")
    with pytest.raises(ValueError, match="mode must be"):
        s.calculate_power_density_line([0, 0, 20], [0.01, 0, 20], npoints=3, mode="fast")
    with pytest.raises(ValueError, match="mode must be"):
        s.calculate_power_density_stl(mode="fast")


def test_icosphere_is_closed_and_outward():
    """The synthetic STL sphere: 20 * 4^k triangles, every edge shared by two
    triangles, normals pointing away from the centre, rows as read_stl lays them out."""
    from oscars_amd import geometry as G
    tris = G.icosphere(2.0, 2, (1.0, 0.0, 5.0))
    assert len(tris) == 320
    edges = {}
    for _, a, b, c in tris:
        for p, q in ((a, b), (b, c), (c, a)):
            k = tuple(sorted((tuple(np.round(p, 12)), tuple(np.round(q, 12)))))
            edges[k] = edges.get(k, 0) + 1
    assert set(edges.values()) == {2}
    rows = G.stl_rows(tris)
    assert rows.shape == (320, 15)
    centre = np.array([1.0, 0.0, 5.0])
    assert np.all(np.einsum("ij,ij->i", rows[:, 9:12], rows[:, 12:15] - centre) > 0)
    assert np.allclose(np.linalg.norm(rows[:, 0:3] - centre, axis=1), 2.0)
    assert np.allclose(rows[:, 12:15], (rows[:, 0:3] + rows[:, 3:6] + rows[:, 6:9]) / 3)
