"""Field values at points (get_bfield / get_efield, summed on the device by
om_field_points) and the accumulated results of add_to_spectrum /
add_to_flux / add_to_power_density.

The oracle's field sums are pinned to the reference's GetB / GetE here and in
tests/test_field_files.py; the device sums are then held to the oracle."""
import ctypes
import os
import sys
import tempfile

import numpy as np
import pytest

from tests.common import EPU, UNDULATOR, other_fields, undulator_like, with_efields, write_oscars_grid


def _ref():
    from oracle import pyoracle as P
    if not P.ref_available():
        pytest.skip("reference build (oracle/_ref/libref.so) not present")
    return P


def _points(n, seed, zspan=1.2):
    rng = np.random.default_rng(seed)
    p = np.column_stack([rng.uniform(-0.004, 0.004, n), rng.uniform(-0.003, 0.003, n), rng.uniform(-zspan, zspan, n)])
    p[0] = [0, 0, 0]
    return p


def test_oracle_efield_matches_reference():
    P = _ref()
    from oracle.pyoracle import Oracle
    o, r = Oracle(nthreads=1), P.RefSim()
    with_efields(o)
    with_efields(r)
    for x in _points(200, 3):
        assert np.array_equal(o.get_efield(x), r.get_efield(x)), x
        assert np.array_equal(o.get_bfield(x), r.get_bfield(x)), x


# ---------------------------------------------------------------- accumulated results (host side)

def test_add_to_spectrum_accumulates_with_weights():
    from oscars_amd import io as IO
    from oscars_amd import sr as SR
    s = SR.sr()
    e = np.linspace(100.0, 900.0, 9)
    a, b = np.arange(1.0, 10.0), np.arange(10.0, 19.0) ** 2
    assert s.get_spectrum() == []
    s.add_to_spectrum([[x, y] for x, y in zip(e, a)], weight=0.5)
    s.add_to_spectrum([[x, y] for x, y in zip(e, b)])
    s.add_to_spectrum([[x, y] for x, y in zip(e + 1, b)], weight=-2.0)  # energies come from the first input
    got = s.get_spectrum()
    assert [g[0] for g in got] == list(e)
    assert np.array_equal([g[1] for g in got], 0.5 * a + b - 2.0 * b)
    with pytest.raises(ValueError, match="spectra dimensions do not match"):
        s.add_to_spectrum([[1.0, 2.0]])
    with pytest.raises(ValueError, match="No points in spectrum."):
        s.add_to_spectrum([])
    with tempfile.TemporaryDirectory() as d:
        s.write_spectrum(ofile=os.path.join(d, "a.txt"), bofile=os.path.join(d, "a.dat"))
        IO.write_spectrum_text(os.path.join(d, "b.txt"), [g[0] for g in got], [g[1] for g in got])
        IO.write_spectrum_binary(os.path.join(d, "b.dat"), [g[0] for g in got], [g[1] for g in got])
        for k in ("txt", "dat"):
            assert open(os.path.join(d, "a." + k), "rb").read() == open(os.path.join(d, "b." + k), "rb").read()


@pytest.mark.parametrize("which", ["flux", "power_density"])
def test_add_to_points_accumulates(which):
    from oscars_amd import sr as SR
    s = SR.sr()
    add, get = getattr(s, "add_to_" + which), getattr(s, "get_" + which)
    pts = [[0.001 * i, -0.002 * i, 30.0] for i in range(7)]
    add([[p, float(i)] for i, p in enumerate(pts)], weight=3.0)
    add([[p, 1.0] for p in pts], 0.25)
    got = get()
    assert [g[0] for g in got] == pts
    assert [g[1] for g in got] == [3.0 * i + 0.25 for i in range(7)]
    with pytest.raises(ValueError, match="spectra dimensions do not match"):
        add([[pts[0], 1.0]])
    with pytest.raises(ValueError, match="No points in flux."):
        add([])
    with pytest.raises(ValueError, match="number of elements not 3"):
        add([[[0.0, 1.0], 2.0]])


# ---------------------------------------------------------------- device field sums

def _fieldsets(tmp):
    path = os.path.join(tmp, "g.txt")
    write_oscars_grid(path, [-0.004, -0.003, -0.5], [0.002, 0.002, 0.0025], [5, 4, 401], undulator_like)

    def und(x):
        x.add_bfield_undulator(**UNDULATOR)

    def epu_(x):
        x.add_bfield_epu(**EPU)

    def grid(x):
        from oscars_amd import sr as SR
        kw = dict(rotations=[0.001, -0.002, 0.01], translation=[0.0001, 0, 0.002])
        if isinstance(x, SR.sr):
            x.add_bfield_file(path, iformat="OSCARS", **kw)
        else:
            x.add_bfield_grid(*SR.read_oscars_field_file(path), **kw)

    return {"undulator": und, "epu": epu_, "other": other_fields, "grid3d": grid}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["undulator", "epu", "other", "grid3d"])
def test_get_bfield_matches_oracle(case):
    """Device sums against the oracle's at 300 points in and around the
    magnets: within 1e-13 of the largest field."""
    from oracle.pyoracle import Oracle
    from oscars_amd import sr as SR
    with tempfile.TemporaryDirectory() as d:
        make = _fieldsets(d)[case]
        o, s = Oracle(nthreads=1), SR.sr()
        make(o)
        make(s)
        pts = _points(300, 11)
        want = np.array([o.get_bfield(x) for x in pts])
        got = np.array([s.get_bfield(list(x)) for x in pts])
    scale = max(float(np.max(np.abs(want))), 1e-300)
    assert np.max(np.abs(want)) > 0.1
    assert np.max(np.abs(got - want)) <= 1e-13 * scale, np.max(np.abs(got - want)) / scale


@pytest.mark.gpu
def test_get_efield_and_names():
    """Electric fields against the oracle; a name selects its fields only, an
    unknown name gives zeros, the named parts add up to the total."""
    from oracle.pyoracle import Oracle
    from oscars_amd import sr as SR
    o, s = Oracle(nthreads=1), SR.sr()
    with_efields(o)
    with_efields(s)
    pts = _points(200, 5)
    want = np.array([o.get_efield(x) for x in pts])
    got = np.array([s.get_efield(list(x)) for x in pts])
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    t = SR.sr()
    t.add_bfield_undulator(**UNDULATOR, name="u")
    t.add_bfield_gaussian(bfield=[0, 0.8, 0.05], sigma=[0, 0, 0.1], translation=[0, 0, 0.2], name="g")
    t.add_efield_uniform(efield=[1e5, 0, 0], width=[0, 0, 1], name="u")
    for x in pts[:20]:
        tot = np.array(t.get_bfield(list(x)))
        parts = np.array(t.get_bfield(list(x), name="u")) + np.array(t.get_bfield(list(x), name="g"))
        assert np.allclose(tot, parts, rtol=0, atol=1e-15 * max(1.0, np.max(np.abs(tot))))
        assert t.get_bfield(list(x), name="nothing") == [0.0, 0.0, 0.0]
        assert t.get_efield(list(x), name="g") == [0.0, 0.0, 0.0]
    assert t.get_efield([0, 0, 0], name="u") == [1e5, 0.0, 0.0]
    with pytest.raises(ValueError, match="Incorrect format in input"):
        t.get_bfield([0, 0])


@pytest.mark.gpu
def test_get_bfield_leaves_trajectory_fields_alone():
    """A field lookup between two spectra of a grid-field trajectory leaves
    the second spectrum identical to the first."""
    from oscars_amd import sr as SR
    with tempfile.TemporaryDirectory() as d:
        s = SR.sr()
        s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
        _fieldsets(d)["grid3d"](s)
        s.set_ctstartstop(0, 2)
        s.set_npoints_trajectory(4001)
        s.calculate_trajectory()
        kw = dict(obs=[0, 0, 30], energy_range_eV=[100, 1500], npoints=16)
        a = s.calculate_spectrum(**kw)
        other = SR.sr()
        _fieldsets(d)["grid3d"](other)
        s.get_bfield([0, 0, 0.01])
        other.get_bfield([0, 0, 0.01])
        b = s.calculate_spectrum(**kw)
    assert a == b


def test_remove_fields_by_name():
    """remove_bfield / remove_efield drop every field of that name and keep
    the rest; "" removes the unnamed fields (the reference's own removal of an
    unnamed field with "" is test_removed_efields_keep_the_eb_equations in
    tests/test_oracle_vs_ref.py); a removal that leaves magnetic fields only
    keeps the E+B equations of motion (a zero electric field here)."""
    from oscars_amd import sr as SR
    s = SR.sr()
    s.add_bfield_undulator(**UNDULATOR, name="u")
    s.add_bfield_gaussian(bfield=[0, 0.8, 0.05], sigma=[0, 0, 0.1], name="g")
    s.add_bfield_uniform(bfield=[0.1, 0, 0], width=[0, 0, 0.2], name="g")
    s.add_bfield_uniform(bfield=[0.2, 0, 0], width=[0, 0, 0.2])
    s.add_efield_uniform(efield=[1e5, 0, 0], width=[0, 0, 1], name="e")
    s.add_efield_uniform(efield=[2e5, 0, 0], width=[0, 0, 1])
    s.remove_bfield("")
    assert [f.name for f in s._fields] == ["u", "g", "g"]
    s.remove_efield("")
    assert [f.name for f in s._efields] == ["e"]
    s.remove_bfield("g")
    assert [f.name for f in s._fields] == ["u"]
    assert len(s._efields_for_derivs()) == 1
    s.remove_efield("e")
    assert s._efields == []
    eb = s._efields_for_derivs()
    assert len(eb) == 1 and list(eb[0].field) == [0.0, 0.0, 0.0] and list(eb[0].shape) == [0.0, 0.0, 0.0]
    s.add_bfield_gaussian(bfield=[0, 0.1, 0], sigma=[0, 0, 0.1])  # an addition chooses again
    assert s._efields_for_derivs() == []
    t = SR.sr()
    t.add_bfield_undulator(**UNDULATOR, name="u")
    t.remove_bfield("u")
    assert t._efields_for_derivs() == []


def test_grid_samples_outlive_clear_bfields(tmp_path):
    """An electric-field map keeps its samples when the magnetic fields are
    cleared: the field itself owns the array its grid pointer reads."""
    import gc
    from oscars_amd import sr as SR
    z = -1 + 0.02 * np.arange(101)
    path = tmp_path / "e.dat"
    write_oscars_grid(str(path), [0, 0, -1], [0, 0, 0.02], [1, 1, 101], lambda x, y, z: [1e5 * z, 2e5, -3e5 * z * z])
    s = SR.sr()
    s.add_efield_file(str(path), "OSCARS")
    s.add_bfield_undulator(**UNDULATOR)
    f = s._efields[-1]
    addr = ctypes.cast(f.grid_data, ctypes.c_void_p).value
    s.clear_bfields()
    gc.collect()
    assert f._keep.ctypes.data == addr and sys.getrefcount(f._keep) >= 2
    got = np.ctypeslib.as_array(f.grid_data, shape=(101 * 3,)).reshape(101, 3)
    assert np.array_equal(got[:, 0], 1e5 * z) and np.all(got[:, 1] == 2e5)


@pytest.mark.gpu
def test_removed_efields_keep_the_eb_equations_on_the_device():
    """After its electric field is removed an undulator trajectory is the
    oracle's E+B integration with a zero electric field (the reference's own,
    in tests/test_oracle_vs_ref.py), knots within 1e-13."""
    from oracle.pyoracle import Oracle
    from oscars_amd import sr as SR
    o, s = Oracle(nthreads=1), SR.sr()
    for x in (o, s):
        x.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
        x.add_bfield_undulator(**UNDULATOR)
        x.set_ctstartstop(0, 2)
        x.set_npoints_trajectory(4001)
    s.add_efield_uniform(efield=[0, 5e5, 1e4], width=[0, 0, 0.3], translation=[0, 0, 0.5], name="e")
    s.remove_efield("e")
    o.add_efield_uniform(efield=[0, 0, 0])
    T = o.calculate_trajectory()
    s.calculate_trajectory()
    t, p, _ = s._device().trajectory(0)
    err = float(np.max(np.abs(p - T.p) / np.maximum(np.max(np.abs(T.p), axis=0), 1e-300)))
    assert len(t) == len(T.t) and np.max(np.abs(t - T.t)) <= 1e-13 * np.max(np.abs(T.t))
    assert err <= 1e-13, err
