"""GPU parity: the HIP path (through the C ABI) against the C oracle.

Tolerances: plain FP64 and tiled FP64 paths 1e-9 relative per point
(denominator floored at 1e-6 of the map maximum, where the sum cancels to
near zero); mixed precision 1e-5 on the same measure.  Converged levels must
be identical.
"""
import numpy as np
import pytest

from tests.common import planar, epu, rel_err

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle_planar():
    from oracle.pyoracle import Oracle
    o = Oracle(nthreads=8)
    planar(o)
    o.calculate_trajectory()
    return o


@pytest.fixture(scope="module")
def gpu_planar():
    from oscars_amd.sr import sr
    g = sr()
    planar(g)
    return g


@gpu
def test_trajectory_knots(oracle_planar, gpu_planar):
    t, p, ypp = gpu_planar.get_trajectory_arrays()
    T = oracle_planar.traj
    assert len(t) == len(T.t) == 20001
    assert np.array_equal(t, T.t)
    # The device RK4 evaluates the field's sine with its own polynomial, which
    # can differ from the C library's in the last bit; differences are measured
    # per channel against that channel's largest magnitude.
    same = np.mean(np.all(p == T.p, axis=1))
    print("bit-identical knots: %.4f" % same)
    scale = np.max(np.abs(T.p), axis=0)
    scale[scale == 0] = 1
    err = np.max(np.abs(p - T.p) / scale)
    yscale = np.max(np.abs(T.ypp), axis=0)
    yscale[yscale == 0] = 1
    yerr = np.max(np.abs(ypp - T.ypp) / yscale)
    print("knots: max channel-relative difference %.3e, spline %.3e" % (err, yerr))
    assert err < 1e-12
    assert yerr < 1e-10


@gpu
def test_level_samples(oracle_planar, gpu_planar):
    dev = gpu_planar._device()
    for L in (0, 5, 11):
        a = dev.level(0, L)
        b = oracle_planar.traj.level(L)
        assert np.array_equal(a[:, 0], b[:, 0]), L
        scale = np.max(np.abs(oracle_planar.traj.p), axis=0)
        scale[scale == 0] = 1
        assert np.max(np.abs(a[:, 1:] - b[:, 1:]) / scale) < 1e-12, L


@gpu
@pytest.mark.parametrize("energy", [152.11, 260.0])
@pytest.mark.parametrize("mode", ["plain", "tiled"])
def test_flux_map_fp64(oracle_planar, gpu_planar, energy, mode):
    kw = dict(npoints=[21, 21], width=[0.02, 0.02], translation=[0, 0, 30])
    res = gpu_planar.calculate_flux_rectangle(energy_eV=energy, mode=mode, **kw)
    v = np.array([r[1] for r in res])
    xyz = np.array([r[0] for r in res])
    lev = gpu_planar._last_levels
    cont, ov, olev, oconv, _, _ = oracle_planar.calculate_flux_rectangle(energy, carry_quirk=False, **kw)
    assert np.array_equal(xyz, cont)
    assert np.array_equal(lev, olev)
    e = rel_err(v, ov, 1e-6)
    print("flux %s %.2f eV: max rel err %.3e" % (mode, energy, e))
    assert e < 1e-9


@gpu
@pytest.mark.parametrize("mode,tol", [("plain", 1e-9), ("tiled", 1e-9), ("mixed", 1e-5)])
def test_spectrum(oracle_planar, gpu_planar, mode, tol):
    res = gpu_planar.calculate_spectrum(obs=[0, 0, 30], npoints=150, energy_range_eV=[100, 2000], mode=mode)
    e = np.array([r[0] for r in res])
    v = np.array([r[1] for r in res])
    lev = gpu_planar._last_levels
    oe, ov, olev, oconv, _ = oracle_planar.calculate_spectrum([0, 0, 30], npoints=150, energy_range_eV=[100, 2000],
                                                              carry_quirk=False)
    assert np.array_equal(e, oe)
    err = rel_err(v, ov, 1e-6)
    print("spectrum %s: max rel err %.3e, level mismatches %d" % (mode, err, int(np.sum(lev != olev))))
    assert np.array_equal(lev, olev)
    assert err < tol


@gpu
@pytest.mark.parametrize("mode,ref", [("mixed", "tiled"), ("point_mixed", "plain")])
def test_mixed_levels_match_fp64(gpu_planar, mode, ref):
    """Every stop decision of the mixed paths matches the FP64 path: decisions
    within the sin/cos error bound are recomputed in FP64 (counted in
    n_rechecked)."""
    kw = dict(npoints=[41, 41], width=[0.02, 0.02], translation=[0, 0, 30])
    e = np.linspace(140.0, 460.0, 64) if mode == "mixed" else np.array([152.11])
    _, v64 = gpu_planar.calculate_flux_rectangle_energies(e, mode=ref, **kw)
    lev64 = gpu_planar._last_levels.copy()
    _, v32 = gpu_planar.calculate_flux_rectangle_energies(e, mode=mode, **kw)
    lev32 = gpu_planar._last_levels
    st = gpu_planar.last_stats
    err = rel_err(v32.ravel(), v64.ravel(), 1e-6)
    print("%s vs %s: %d points, level mismatches %d, rechecked %d, max rel err %.3e"
          % (mode, ref, lev32.size, int(np.sum(lev32 != lev64)), st["n_rechecked"], err))
    assert np.array_equal(lev32, lev64)
    assert err < 1e-5


@gpu
@pytest.mark.parametrize("pol", ["lh", "lv", "cl", "cr"])
def test_spectrum_polarization(oracle_planar, gpu_planar, pol):
    res = gpu_planar.calculate_spectrum(obs=[0.001, 0.0005, 30], npoints=40, energy_range_eV=[120, 800],
                                        polarization=pol)
    v = np.array([r[1] for r in res])
    _, ov, olev, _, _ = oracle_planar.calculate_spectrum([0.001, 0.0005, 30], npoints=40, energy_range_eV=[120, 800],
                                                         polarization=pol, carry_quirk=False)
    assert np.array_equal(gpu_planar._last_levels, olev)
    assert rel_err(v, ov, 1e-6) < 1e-9


@gpu
@pytest.mark.parametrize("quantity", ["precision", "level"])
def test_flux_quantities(oracle_planar, gpu_planar, quantity):
    kw = dict(npoints=[7, 5], width=[0.01, 0.01], translation=[0, 0, 30])
    res = gpu_planar.calculate_flux_rectangle(energy_eV=300.0, quantity=quantity, **kw)
    v = np.array([r[1] for r in res])
    _, ov, _, _, _, _ = oracle_planar.calculate_flux_rectangle(300.0, quantity=quantity, carry_quirk=False, **kw)
    if quantity == "level":
        assert np.array_equal(v, ov)
    else:
        assert rel_err(v, ov, 1e-12) < 1e-6


@gpu
@pytest.mark.parametrize("normal", [0, 1])
def test_power_density_epu(normal):
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    o = Oracle(nthreads=8)
    epu(o)
    g = sr()
    epu(g)
    kw = dict(npoints=[21, 21], width=[0.08, 0.08], translation=[0, 0, 20], normal=normal)
    res = g.calculate_power_density_rectangle(**kw)
    v = np.array([r[1] for r in res])
    cont, ov, olev, _, _ = o.calculate_power_density_rectangle(carry_quirk=False, **kw)
    assert np.array_equal(np.array([r[0] for r in res]), cont)
    assert np.array_equal(g._last_levels, olev)
    e = rel_err(v, ov, 1e-6)
    print("power density normal=%d: max rel err %.3e" % (normal, e))
    assert e < 1e-9


@gpu
def test_grid_field_trajectory():
    """1-D z grid map (4001 samples, staged in LDS by the RK4 kernel) against the oracle's knots."""
    from oracle.pyoracle import Oracle
    from oscars_amd import adjoint as A
    from oscars_amd.sr import sr
    samples = A.undulator_pole_map(4001, -1.0, 5e-4, sigma_amp=0.005, sigma_z=20e-6, seed=7)
    o = Oracle(nthreads=8)
    o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_grid([0, 0, -1.0], [0, 0, 5e-4], [1, 1, 4001], samples)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(4001)
    ref = o._traj()
    s = sr()
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_grid([0, 0, -1.0], [0, 0, 5e-4], [1, 1, 4001], samples)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(4001)
    t, p, ypp = s.get_trajectory_arrays()
    print("grid trajectory: zero knots %d of %d; x at end %.6e (oracle %.6e)" %
          (int(np.sum(np.all(p == 0, axis=1))), len(p), p[-1, 0], ref.p[-1, 0]))
    scale = np.max(np.abs(ref.p), axis=0)
    scale[scale == 0] = 1
    err = np.max(np.abs(p - ref.p) / scale)
    print("grid trajectory: max channel-relative difference %.3e" % err)
    assert np.array_equal(t, ref.t)
    assert err < 1e-12


@gpu
def test_device_resident_entry_points_match_host(gpu_planar):
    """om_flux_dev / om_spectrum_dev / om_power_density_dev on torch device
    tensors give the host entry points' values and levels bit for bit."""
    import torch
    from oscars_amd import geometry as G
    dev = gpu_planar._device()
    obs, _, _ = G.rectangle("XY", 9, 7, 0.01, 0.01, [0, 0, 0], [0, 0, 30], 0, 2)
    obs = np.array(obs)
    e = np.array([152.11, 300.0, 455.0])
    pol = [0.0] * 6
    for mode in ("plain", "mixed", "point_mixed"):
        conv = dev.conv(mode=mode)
        v, lev, cv, ef, st = dev.flux(0, obs, e, pol, conv, want_e=True)
        obs_t = torch.tensor(obs, dtype=torch.float64, device="cuda")
        e_t = torch.tensor(e, dtype=torch.float64, device="cuda")
        n = len(obs) * len(e)
        out_t = torch.zeros(n, dtype=torch.float64, device="cuda")
        lev_t = torch.zeros(n, dtype=torch.int32, device="cuda")
        cv_t = torch.zeros(n, dtype=torch.uint8, device="cuda")
        ef_t = torch.zeros(n, 6, dtype=torch.float64, device="cuda")
        st_d = dev.flux_dev(0, obs_t, e_t, pol, conv, out_t, lev_t, cv_t, ef_t)
        assert np.array_equal(out_t.cpu().numpy(), v), mode
        assert np.array_equal(lev_t.cpu().numpy(), lev), mode
        assert np.array_equal(cv_t.cpu().numpy(), cv), mode
        efd = ef_t.cpu().numpy()
        assert np.array_equal(efd[:, 0::2] + 1j * efd[:, 1::2], ef), mode
        assert st_d["evals"] == st["evals"] and st_d["max_level"] == st["max_level"]
    conv = dev.conv()
    v, lev, cv, _, _ = dev.spectrum(0, [0, 0, 30], e, pol, conv)
    out_t = torch.zeros(len(e), dtype=torch.float64, device="cuda")
    dev.spectrum_dev(0, torch.tensor([0.0, 0.0, 30.0], dtype=torch.float64, device="cuda"),
                     torch.tensor(e, dtype=torch.float64, device="cuda"), pol, conv, out_t)
    assert np.array_equal(out_t.cpu().numpy(), v)
    v, lev, cv, _ = dev.power_density(0, obs, [0, 0, 1], 0, conv)
    out_t = torch.zeros(len(obs), dtype=torch.float64, device="cuda")
    lev_t = torch.zeros(len(obs), dtype=torch.int32, device="cuda")
    dev.power_density_dev(0, torch.tensor(obs, dtype=torch.float64, device="cuda"), [0, 0, 1], 0, conv, out_t, lev_t)
    assert np.array_equal(out_t.cpu().numpy(), v)
    assert np.array_equal(lev_t.cpu().numpy(), lev)


@gpu
def test_output_files(gpu_planar, tmp_path):
    """ofile / bofile of the drop-in methods write the returned values in the
    reference's formats (the writers themselves are pinned in tests/test_io.py)."""
    from oscars_amd import io as IO
    res = gpu_planar.calculate_spectrum(obs=[0, 0, 30], npoints=20, energy_range_eV=[100, 500],
                                        ofile=str(tmp_path / "s.txt"), bofile=str(tmp_path / "s.dat"))
    IO.write_spectrum_text(str(tmp_path / "r.txt"), [r[0] for r in res], [r[1] for r in res])
    assert (tmp_path / "s.txt").read_bytes() == (tmp_path / "r.txt").read_bytes()
    assert len((tmp_path / "s.dat").read_bytes()) == 20 * 8
    res = gpu_planar.calculate_flux_rectangle(energy_eV=152.11, npoints=[5, 4], width=[0.01, 0.01],
                                              translation=[0, 0, 30], ofile=str(tmp_path / "f.txt"),
                                              bofile=str(tmp_path / "f.dat"))
    lines = (tmp_path / "f.txt").read_text().splitlines()
    assert len(lines) == 20 and len(lines[0].split()) == 3
    avg = gpu_planar.average_flux(ifiles=[str(tmp_path / "f.txt")] * 2)
    assert np.allclose([a[1] for a in avg], [float(l.split()[2]) for l in lines], rtol=0, atol=0)


@gpu
def test_gaussian_uniform_quadrupole_fields():
    """Gaussian, uniform-box and quadrupole fields on the device: the
    trajectory against the oracle (the device's exp and sin differ from
    glibc's in the last bits) and the spectrum to 1e-9, levels identical."""
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    from tests.common import other_fields
    o = Oracle(nthreads=8)
    other_fields(o)
    T = o.calculate_trajectory()
    s = sr()
    other_fields(s)
    s.calculate_trajectory()
    t, p, _ = s._device().trajectory(0)
    scale = np.maximum(np.max(np.abs(T.p), axis=0), 1e-300)
    et = float(np.max(np.abs(p - T.p) / scale))
    assert np.array_equal(t, T.t)
    e = np.linspace(10.0, 2000.0, 40)
    ov, olev, _, _ = o.spectrum([0, 0, 30], e, carry_quirk=False)
    res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode="plain")
    v = np.array([r[1] for r in res])
    es = rel_err(v, ov, 1e-6)
    print("other fields: trajectory %.3e, spectrum %.3e" % (et, es))
    assert et < 1e-11 and es < 1e-9
    assert np.array_equal(s._last_levels, olev)


@gpu
@pytest.mark.parametrize("dims", ["xyz", "xz"])
def test_multi_dimensional_grid_file(tmp_path, dims):
    """2-D and 3-D OSCARS field maps on the device against the oracle (whose
    grid is pinned to the reference's in tests/test_field_files.py)."""
    from oracle.pyoracle import Oracle
    from oscars_amd import sr as SR
    from tests.common import undulator_like, write_oscars_grid
    n = [5 if "x" in dims else 1, 4 if "y" in dims else 1, 401]
    path = str(tmp_path / "g.txt")
    write_oscars_grid(path, [-0.004, -0.003, -0.5], [0.002, 0.002, 0.0025], n, undulator_like)
    kw = dict(rotations=[0.001, -0.002, 0.01], translation=[0.0001, 0, 0.002])
    o, s = Oracle(nthreads=8), SR.sr()
    for x in (o, s):
        x.set_particle_beam(energy_GeV=3, x0=[0.0002, -0.0001, -1], d0=[0, 0, 1], current=0.5)
    st, sp, npts, data = SR.read_oscars_field_file(path)
    o.add_bfield_grid(st, sp, npts, data, **kw)
    s.add_bfield_file(path, iformat="OSCARS", **kw)
    for x in (o, s):
        x.set_ctstartstop(0, 2)
        x.set_npoints_trajectory(4001)
    T = o.calculate_trajectory()
    s.calculate_trajectory()
    t, p, _ = s._device().trajectory(0)
    et = float(np.max(np.abs(p - T.p) / np.maximum(np.max(np.abs(T.p), axis=0), 1e-300)))
    e = np.linspace(100.0, 600.0, 12)
    ov, olev, _, _ = o.spectrum([0, 0, 30], e, carry_quirk=False)
    v = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode="plain")])
    es = rel_err(v, ov, 1e-6)
    print("%s grid: trajectory %.3e, spectrum %.3e" % (dims, et, es))
    assert et < 1e-12 and es < 1e-9


@gpu
@pytest.mark.parametrize("magnetic", [True, False])
def test_electric_fields(magnetic):
    """Electric fields on the device (om_trajectories_eb): the trajectory
    against the oracle (pinned to the reference) and the spectrum to 1e-9.
    The equations of motion with E take sqrt(1 - v.v / c^2) at every step, a
    difference of numbers 3e-8 apart at 3 GeV, which magnifies last-bit
    differences of the device's sine: with fields of 300 MV/m the knots
    differed by 2e-7 of a channel's largest value; with these 3 MV/m fields
    by 1e-13.  The knots' bar is 1e-8."""
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    from tests.common import with_efields
    o = Oracle(nthreads=8)
    with_efields(o, magnetic=magnetic)
    T = o.calculate_trajectory()
    s = sr()
    with_efields(s, magnetic=magnetic)
    s.calculate_trajectory()
    t, p, _ = s._device().trajectory(0)
    et = float(np.max(np.abs(p - T.p) / np.maximum(np.max(np.abs(T.p), axis=0), 1e-300)))
    e = np.linspace(100.0, 800.0, 16)
    ov, olev, _, _ = o.spectrum([0, 0, 30], e, carry_quirk=False)
    v = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), mode="plain")])
    es = rel_err(v, ov, 1e-6)
    print("electric fields (magnetic %s): trajectory %.3e, spectrum %.3e" % (magnetic, et, es))
    assert np.array_equal(t, T.t) and et < 1e-8 and es < 1e-9
