"""Shared scenario set-up for the tests: the same calls drive the product
(oscars_amd.sr), the oracle and, where built, the reference core."""
import math

UNDULATOR = dict(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
EPU = dict(bfieldA=[0, 0.6, 0], bfieldB=[0.6, 0, 0], period=[0, 0, 0.057], nperiods=40, phase=math.pi / 2)


def planar(s, npoints=20001):
    """Planar undulator scenario (1.0 T, 49 mm, 31 periods, 3 GeV, 0.5 A)."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_undulator(**UNDULATOR)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(npoints)


def epu(s, npoints=40001):
    """Elliptically polarising undulator scenario (0.6 T x 0.6 T, 57 mm, 40 periods)."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    s.add_bfield_epu(**EPU)
    s.set_ctstartstop(0, 3)
    s.set_npoints_trajectory(npoints)


def rel_err(a, b, floor_frac=0.0):
    """max |a-b| / max(|b|, floor_frac * max|b|) elementwise."""
    import numpy as np
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    den = np.maximum(np.abs(b), floor_frac * np.max(np.abs(b)))
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(a - b) / den))


# Beams with spread for the sampling tests: set_particle_beam keyword sets
# (x0, d0 and the beam-spread arguments; unset arguments are left out).
BEAM_CASES = [
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5, sigma_energy_GeV=0.0027,
         emittance=[0.9e-9, 8e-12], beta=[1.5, 0.8]),
    dict(energy_GeV=3.0, x0=[0.001, -0.002, -1.5], d0=[0.01, 0.002, 1], current=0.5, sigma_energy_GeV=0.0027,
         emittance=[0.9e-9, 8e-12], beta=[1.5, 0.8], alpha=[0.3, -0.2], lattice_reference=[0, 0, 0.7],
         eta=[0.01, 0.002], horizontal_direction=[1, 0.1, 0]),
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], sigma_energy_GeV=0.001, emittance=[0.5e-9, 8e-12],
         beta=[2.0, 1.0], gamma=[0.7, 1.2]),
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], sigma_energy_GeV=0.001, emittance=[0.5e-9, 8e-12],
         alpha=[0.1, 0.2], gamma=[0.7, 1.2]),
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], sigma_energy_GeV=0.001, emittance=[0.5e-9, 8e-12],
         beta=[2.0, 1.0], alpha=[0.5, -0.5], gamma=[0.625, 1.25], distribution="KV"),
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], sigma_energy_GeV=0.001),
    dict(energy_GeV=3.0, x0=[0, 0, -1.5], d0=[0, 0, 1], sigma_energy_GeV=0.001, distribution="gaussian"),
]

# Predefined beams with and without overrides (set_particle_beam keywords).
PREDEFINED_CASES = [
    dict(beam="NSLSII-LONGSTRAIGHT", x0=[0, 0, -1]),
    dict(beam="NSLS2-SHORTSTRAIGHT", x0=[0, 0, -1], energy_GeV=2.5, sigma_energy_GeV=0.004,
         lattice_reference=[0, 0, 0.5]),
    dict(beam="nsls-ii", x0=[0, 0, -1]),
    dict(beam="NSLSII", x0=[0, 0, -1], distribution="gaussian"),
    dict(beam="NSLSII-LONGSTRAIGHT", x0=[0.0001, 0, -1], d0=[0.001, 0, 1], emittance=[1e-9, 1e-11], beta=[2, 1],
         alpha=[0.2, 0.1], eta=[0.02, 0]),
]


def stl_plate(x0, x1, y0, y1, z, nx, ny):
    """Triangles (normal, a, b, c) of a flat nx x ny rectangle at height z, two per cell."""
    import numpy as np
    xs, ys = np.linspace(x0, x1, nx + 1), np.linspace(y0, y1, ny + 1)
    out = []
    for i in range(nx):
        for j in range(ny):
            a, b = [xs[i], ys[j], z], [xs[i + 1], ys[j], z]
            c, d = [xs[i + 1], ys[j + 1], z], [xs[i], ys[j + 1], z]
            out += [([0, 0, -1], a, b, c), ([0, 0, -1], a, c, d)]
    return out


def write_stl_triangles(path, tris, header=b"test mesh"):
    """A binary STL file of (normal, a, b, c) triangles."""
    import numpy as np
    from oscars_amd import io as IO
    rec = np.zeros(len(tris), dtype=IO._STL_REC)
    for k, (n, a, b, c) in enumerate(tris):
        rec[k]["n"], rec[k]["a"], rec[k]["b"], rec[k]["c"] = n, a, b, c
    with open(path, "wb") as f:
        f.write(header.ljust(80, b" "))
        f.write(np.array([len(tris)], dtype="<u4").tobytes())
        f.write(rec.tobytes())


def stl_icosphere(radius, subdiv, centre=(0.0, 0.0, 0.0)):
    """Triangles (normal, a, b, c) of a subdivided icosahedron on a sphere."""
    from oscars_amd import geometry as G
    return G.icosphere(radius, subdiv, centre)


def stl_scene(dirname):
    """A screen of 36 triangles at 20 m with a 4-triangle blocker at 10 m in
    front of part of it (coordinates in mm, scale 1e-3).  Returns the two
    (path, scale, translation) entries."""
    import os
    screen, blocker = os.path.join(dirname, "screen.stl"), os.path.join(dirname, "blocker.stl")
    write_stl_triangles(screen, stl_plate(-10, 10, -5, 5, 20000, 6, 3))
    write_stl_triangles(blocker, stl_plate(-2, 3, -1, 2, 10000, 2, 1))
    return [(screen, 1e-3, [0.0, 0.0, 0.0]), (blocker, 1e-3, [0.0005, 0.0, 0.0])]


def other_fields(s, npoints=4001):
    """Gaussian, uniform-box and quadrupole fields together, rotated and
    translated (3 GeV electron from z = -1 m)."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_gaussian(bfield=[0, 0.8, 0.05], sigma=[0, 0, 0.1], rotations=[0.01, 0.02, 0.3],
                          translation=[0.0, 0.0, -0.3])
    s.add_bfield_uniform(bfield=[0.1, 0.3, 0], width=[0, 0, 0.4], rotations=[0, 0.05, 0], translation=[0, 0, 0.4])
    s.add_bfield_quadrupole(K=2.0, width=0.3, rotations=[0, 0, 0.1], translation=[0.001, 0, 0.0])
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(npoints)


def write_oscars_grid(path, start, step, n, func):
    """An OSCARS-format text field map: header, then z fastest, y, x."""
    with open(path, "w") as f:
        f.write("# test map\n")
        for a in range(3):
            f.write("%.17g\n%.17g\n%d\n" % (start[a], step[a], n[a]))
        for ix in range(n[0]):
            for iy in range(n[1]):
                for iz in range(n[2]):
                    x = [start[0] + ix * step[0], start[1] + iy * step[1], start[2] + iz * step[2]]
                    f.write("%.17g %.17g %.17g\n" % tuple(func(*x)))


def undulator_like(x, y, z):
    """A 3-D test field: a 50 mm period in z within |z| < 0.45 m, with x and y dependence."""
    k = 2 * _np.pi / 0.05
    env = 1.0 if abs(z) < 0.45 else 0.0
    return [0.02 * _np.cos(k * z) * x, env * 0.9 * _np.cos(k * z) * (1 + 3 * x * x), 0.1 * y]


import numpy as _np  # noqa: E402


def with_efields(s, npoints=4001, magnetic=True):
    """A planar undulator with electric fields added: a Gaussian kick, a
    uniform box and an undulator-like E field (V/m); without the magnet when
    magnetic is False."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    if magnetic:
        s.add_bfield_undulator(**UNDULATOR)
    s.add_efield_gaussian(efield=[2e6, -1e6, 3e5], sigma=[0, 0, 0.1], translation=[0, 0, -0.4])
    s.add_efield_uniform(efield=[0, 5e5, 1e4], width=[0, 0, 0.3], translation=[0, 0, 0.5])
    s.add_efield_undulator(efield=[0, 3e6, 0], period=[0, 0, 0.049], nperiods=31)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(npoints)


def _with_td(s, frequency, frequency_phase, time_offset, add, **kw):
    """add(**kw) with a time-dependent factor cos(2 pi f (t + t0) + phase): the
    product takes it as keyword arguments, the oracle and the reference build
    as a setting for the fields added next."""
    if hasattr(s, "set_time_dependence"):
        s.set_time_dependence(frequency, frequency_phase, time_offset)
        add(**kw)
        s.set_time_dependence(0.0, 0.0, 0.0)
    else:
        add(frequency=frequency, frequency_phase=frequency_phase, time_offset=time_offset, **kw)


def time_dependent(s, npoints=4001, electric=True):
    """Time-dependent fields over a pass of 6.7 ns (ct 0..2 m): the planar
    undulator at 150 MHz (one cycle over the pass, phase 0.3, offset 1 ns), a
    Gaussian kick at 80 MHz, a uniform box at 40 MHz and a quadrupole at
    200 MHz; with an electric field (uniform box at 120 MHz) when electric."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    _with_td(s, 1.5e8, 0.3, 1e-9, s.add_bfield_undulator, **UNDULATOR)
    _with_td(s, 8e7, -1.1, 0.0, s.add_bfield_gaussian, bfield=[0.02, 0.3, 0], sigma=[0, 0, 0.08],
             translation=[0, 0, -0.7])
    _with_td(s, 4e7, 0.0, 2e-9, s.add_bfield_uniform, bfield=[0.05, 0, 0.01], width=[0, 0, 0.2],
             translation=[0, 0, 0.75])
    _with_td(s, 2e8, 0.5, 0.0, s.add_bfield_quadrupole, K=1.5, width=0.2, translation=[0.0005, 0, 0.9])
    if electric:
        _with_td(s, 1.2e8, 0.2, 0.0, s.add_efield_uniform, efield=[2e5, -1e5, 0], width=[0, 0, 0.5],
                 translation=[0, 0, 0.1])
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(npoints)


def halving_fields(s, backward=False):
    """Steps that meet |beta| >= 1 and are halved (src/OSCARSSR.cc:1362-1378):
    a 40 T Gaussian bump (sigma 0.3 m) crossed with 1 m steps, where RK4's
    rotation per step (omega_c dt ~ 4 rad) grows |v| past c, and a 2 T uniform
    box around it; 3 GeV electrons.  With backward, T0 sits inside the window."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -0.5 if backward else -4], d0=[0, 0, 1], current=0.5)
    s.add_bfield_gaussian(bfield=[0, 40.0, 0], sigma=[0, 0, 0.3], translation=[0, 0, 0.0])
    s.add_bfield_uniform(bfield=[0.3, 2.0, 0], width=[0, 0, 3.0], translation=[0, 0, 0.2])
    if backward:
        s.set_ctstartstop(-3.5, 4.5)
        s.set_npoints_trajectory(9)
    else:
        s.set_ctstartstop(0, 8)
        s.set_npoints_trajectory(9)


def write_binary_grid(path, start, step, n, func, comment=b"", fmt=b"OSCARS", version=1):
    """A binary OSCARS field map (the layout src/TField3D_Grid.cc:875-1085
    reads): comment and format strings with int32 lengths, int32 version,
    float32 start / step and int32 count per axis, float32 Bx By Bz per point
    (z fastest, then y, then x)."""
    import numpy as np
    with open(path, "wb") as f:
        f.write(np.array([len(comment)], dtype="<i4").tobytes() + comment)
        f.write(np.array([version, len(fmt)], dtype="<i4").tobytes() + fmt)
        for a in range(3):
            f.write(np.array([start[a], step[a]], dtype="<f4").tobytes() + np.array([n[a]], dtype="<i4").tobytes())
        vals = []
        for ix in range(n[0]):
            for iy in range(n[1]):
                for iz in range(n[2]):
                    x = [float(np.float32(start[0])) + ix * float(np.float32(step[0])),
                         float(np.float32(start[1])) + iy * float(np.float32(step[1])),
                         float(np.float32(start[2])) + iz * float(np.float32(step[2]))]
                    vals.append(func(*x))
        f.write(np.array(vals, dtype="<f4").tobytes())


def halving_box(s, backward=False):
    """Step halving without transcendental functions: a uniform 10 T box
    crossed with 5 m steps.  Every RK4 stage then leaves |v| above c unless the
    rotation per sub-step is below ~3e-4 rad, so each step splits down to
    depth 14, as in the reference."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -2 if backward else -15], d0=[0, 0, 1], current=0.5)
    s.add_bfield_uniform(bfield=[0.5, 10.0, 0], width=[0, 0, 20.0], translation=[0, 0, 0])
    s.set_ctstartstop(-20, 20) if backward else s.set_ctstartstop(0, 40)
    s.set_npoints_trajectory(9)
