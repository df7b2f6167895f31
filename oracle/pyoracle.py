"""ctypes bindings for the CPU checkers (test infrastructure only).

* ``Oracle``   - this project's C restatement (oracle/liboracle.so).
* ``RefSim``   - the reference's own C++ core (oracle/_ref/libref.so), present
                 only where it was built from the reference sources.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module.  Both classes expose the same scenario-level API as the product's
``oscars_amd.sr.sr`` (set_particle_beam, add_bfield_*, set_ctstartstop,
set_npoints_trajectory, calculate_*) so one test drives all three the same way.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libref.so")
# the reference core with its GPU boundary served by liboscars_mi.so (oracle/Makefile, ref_mi)
LIB_REF_MI = os.path.join(HERE, "_ref", "libref_mi.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)
_lp = C.POINTER(C.c_long)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _v3(x):
    return (C.c_double * 3)(*[float(v) for v in x])


MAX_LEVEL = 26


class orc_field(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("field", C.c_double * 3), ("fieldB", C.c_double * 3), ("period", C.c_double * 3),
        ("center", C.c_double * 3),
        ("nperiods", C.c_int),
        ("phase", C.c_double), ("taper", C.c_double),
        ("period_length", C.c_double), ("unit", C.c_double * 3), ("und_length", C.c_double),
        ("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
        ("xstart", C.c_double), ("ystart", C.c_double), ("zstart", C.c_double),
        ("xstep", C.c_double), ("ystep", C.c_double), ("zstep", C.c_double),
        ("xstop", C.c_double), ("ystop", C.c_double), ("zstop", C.c_double),
        ("dimx", C.c_int),
        ("rot", C.c_double * 3), ("trans", C.c_double * 3),
        ("data", _dp),
        ("shape", C.c_double * 3), ("K", C.c_double), ("width", C.c_double),
        ("freq", C.c_double), ("fphase", C.c_double), ("toff", C.c_double),
    ]


class orc_fields(C.Structure):
    _fields_ = [("f", orc_field * 16), ("n", C.c_int)]


class orc_particle(C.Structure):
    _fields_ = [("x0", C.c_double * 3), ("b0", C.c_double * 3), ("t0", C.c_double),
                ("q", C.c_double), ("m", C.c_double), ("current", C.c_double),
                ("gamma", C.c_double), ("qomg", C.c_double)]


class orc_traj(C.Structure):
    _fields_ = [("n", C.c_long), ("t", _dp), ("p", _dp), ("ypp", _dp),
                ("tstart", C.c_double), ("tstop", C.c_double)]


class orc_opts(C.Structure):
    _fields_ = [("precision", C.c_double), ("max_level", C.c_int), ("max_level_extended", C.c_int),
                ("quantity", C.c_int), ("nthreads", C.c_int), ("carry_quirk", C.c_int),
                ("weight", C.c_double)]


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(LIB_ORACLE):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_ORACLE)
        L.orc_add_undulator.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp, C.c_double, C.c_double]
        L.orc_add_epu.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, _dp, C.c_double, C.c_double]
        L.orc_add_grid.argtypes = [C.c_void_p, _dp, _dp, _lp, _dp, _dp, _dp]
        L.orc_add_gaussian.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_trajectory_rk4_eb.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_long,
                                            _dp, _dp]
        L.orc_add_grid_exact.argtypes = [C.c_void_p, _dp, _dp, _dp, _lp, _dp, _dp, _dp]
        L.orc_add_uniform.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_add_quadrupole.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp]
        L.orc_get_b.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.orc_particle_electron.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double]
        L.orc_particle_set.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.orc_npoints_forward_backward.argtypes = [C.c_double, C.c_double, C.c_long, C.c_double, _lp]
        L.orc_npoints_forward_backward.restype = C.c_long
        L.orc_trajectory_rk4.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_long, _dp, _dp]
        L.orc_spline.argtypes = [C.c_long, _dp, _dp, _dp]
        L.orc_spline_value.argtypes = [C.c_void_p, C.c_double, _dp]
        L.orc_level_samples.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.orc_dt_inclusive.argtypes = [C.c_void_p, C.c_int]
        L.orc_dt_inclusive.restype = C.c_double
        L.orc_rectangle.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                    C.c_int, _dp, _dp, _dp]
        L.orc_energies.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_polarization_code.argtypes = [C.c_char_p]
        L.orc_polarization_vector.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_spectrum.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_int, _dp, _dp, C.c_void_p, _dp, _ip, _up, _dp]
        L.orc_flux.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_long, C.c_double, _dp, C.c_void_p, _dp, _ip, _up,
                               _dp]
        L.orc_power_density.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_long, C.c_int, C.c_void_p, _dp, _ip,
                                        _up]
        L.orc_power_density_surface.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_long, C.c_int, C.c_void_p,
                                                _dp, _ip, _up]
        L.orc_total_power.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _dp, _ip]
        L.orc_efield_time.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.orc_power_density_stl.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_long, C.c_void_p, _dp, _ip, _up]
        L.orc_efield_time.restype = None
        _oracle = L
    return _oracle


class Trajectory:
    """Knots (t, 9 channels) + spline second derivatives, as numpy arrays."""

    def __init__(self, t, p, ypp=None):
        self.t = np.ascontiguousarray(t, dtype=np.float64)
        self.p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 9)
        if ypp is None:
            ypp = np.zeros_like(self.p)
            oracle_lib().orc_spline(len(self.t), self.t.ctypes.data_as(_dp), self.p.ctypes.data_as(_dp),
                                    ypp.ctypes.data_as(_dp))
        self.ypp = np.ascontiguousarray(ypp, dtype=np.float64).reshape(-1, 9)

    def c(self):
        s = orc_traj()
        s.n = len(self.t)
        s.t = self.t.ctypes.data_as(_dp)
        s.p = self.p.ctypes.data_as(_dp)
        s.ypp = self.ypp.ctypes.data_as(_dp)
        s.tstart = float(self.t[0])
        s.tstop = float(self.t[-1])
        self._keep = s
        return s

    def level(self, L, extended=False):
        out = np.zeros((2 ** L, 10))
        oracle_lib().orc_level_samples(C.byref(self.c()), L, int(extended), out.ctypes.data_as(_dp))
        return out

    def dt_inclusive(self, L):
        return oracle_lib().orc_dt_inclusive(C.byref(self.c()), L)


def polarization_vector(polarization, angle, hdir, pdir):
    L = oracle_lib()
    code = L.orc_polarization_code(polarization.encode())
    if code < 0:
        raise ValueError("Polarization requested not recognized")
    out = np.zeros(6)
    L.orc_polarization_vector(code, float(angle), _v3(hdir), _v3(pdir), out.ctypes.data_as(_dp))
    return out


def _pol_string(polarization, angle):
    pol = "all"
    if polarization:
        pol = polarization
    elif angle is not None:
        pol = "linear"
    return pol


def read_oscars_grid(path):
    """Parse the OSCARS text field format (src/TField3D_Grid.cc:353-548)."""
    with open(path) as f:
        lines = f.read().splitlines()
    hv = [float(l.split()[0]) for l in lines[1:10]]
    start = [hv[0], hv[3], hv[6]]
    step = [hv[1], hv[4], hv[7]]
    npts = [int(hv[2]), int(hv[5]), int(hv[8])]
    n = npts[0] * npts[1] * npts[2]
    data = np.array([[float(v) for v in l.split()[:3]] for l in lines[10:10 + n]])
    return start, step, npts, data


class Oracle:
    """Scenario-level driver of the C restatement (same setup API as oscars.sr)."""

    def __init__(self, nthreads=1):
        self.L = oracle_lib()
        self.F = orc_fields()
        self.FE = orc_fields()  # electric fields
        self.P = orc_particle()
        self.ctstart, self.ctstop = 0.0, 0.0
        self.npoints = 0
        self.npoints_per_meter = 10000
        self.nthreads = nthreads
        self.traj = None
        self._grids = []
        self._td = (0.0, 0.0, 0.0)

    def set_time_dependence(self, frequency=0.0, frequency_phase=0.0, time_offset=0.0):
        """frequency, frequency_phase and time_offset of the fields added after
        this call (as RefSim.set_time_dependence)."""
        self._td = (float(frequency), float(frequency_phase), float(time_offset))

    def _added(self, F):
        f = F.f[F.n - 1]
        f.freq, f.fphase, f.toff = self._td

    # -- setup (src/OSCARSSR_Python.cc kwarg semantics) --
    def add_bfield_undulator(self, bfield, period, nperiods, phase=0.0, translation=(0, 0, 0), taper=0.0):
        self.L.orc_add_undulator(C.byref(self.F), _v3(bfield), _v3(period), int(nperiods), _v3(translation),
                                 float(phase), float(taper))
        self._added(self.F)
        self.traj = None

    def add_bfield_epu(self, bfieldA, bfieldB, period, nperiods, phase=0.0, translation=(0, 0, 0), taper=0.0):
        self.L.orc_add_epu(C.byref(self.F), _v3(bfieldA), _v3(bfieldB), _v3(period), int(nperiods),
                           _v3(translation), float(phase), float(taper))
        self._added(self.F)
        self.traj = None

    def add_bfield_gaussian(self, bfield, sigma, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self.L.orc_add_gaussian(C.byref(self.F), _v3(bfield), _v3(translation), _v3(sigma), _v3(rotations))
        self._added(self.F)
        self.traj = None

    def add_efield_gaussian(self, efield, sigma, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self.L.orc_add_gaussian(C.byref(self.FE), _v3(efield), _v3(translation), _v3(sigma), _v3(rotations))
        self._added(self.FE)
        self.traj = None

    def add_efield_uniform(self, efield, width=(0, 0, 0), rotations=(0, 0, 0), translation=(0, 0, 0)):
        self.L.orc_add_uniform(C.byref(self.FE), _v3(efield), _v3(width), _v3(translation), _v3(rotations))
        self._added(self.FE)
        self.traj = None

    def add_efield_undulator(self, efield, period, nperiods, translation=(0, 0, 0)):
        self.L.orc_add_undulator(C.byref(self.FE), _v3(efield), _v3(period), int(nperiods), _v3(translation), 0.0,
                                 0.0)
        self._added(self.FE)
        self.traj = None

    def add_bfield_uniform(self, bfield, width=(0, 0, 0), rotations=(0, 0, 0), translation=(0, 0, 0)):
        self.L.orc_add_uniform(C.byref(self.F), _v3(bfield), _v3(width), _v3(translation), _v3(rotations))
        self._added(self.F)
        self.traj = None

    def add_bfield_quadrupole(self, K, width, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self.L.orc_add_quadrupole(C.byref(self.F), float(K), float(width), _v3(rotations), _v3(translation))
        self._added(self.F)
        self.traj = None

    def add_bfield_grid(self, start, step, npts, samples, rotations=(0, 0, 0), translation=(0, 0, 0)):
        s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
        self._grids.append(s)
        rc = self.L.orc_add_grid(C.byref(self.F), _v3(start), _v3(step), (C.c_long * 3)(*npts),
                                 s.ctypes.data_as(_dp), _v3(rotations), _v3(translation))
        if rc:
            raise ValueError("unsupported grid (rc=%d)" % rc)
        self._added(self.F)
        self.traj = None

    def add_bfield_file(self, ifile, iformat="OSCARS", rotations=(0, 0, 0), translation=(0, 0, 0)):
        start, step, npts, data = read_oscars_grid(ifile)
        self.add_bfield_grid(start, step, npts, data, rotations, translation)

    def add_bfield_grid_exact(self, start, stop, step, npts, samples, rotations=(0, 0, 0), translation=(0, 0, 0)):
        """A grid whose start and stop are given as they are (a resampled table)."""
        s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
        self._grids.append(s)
        rc = self.L.orc_add_grid_exact(C.byref(self.F), _v3(start), _v3(stop), _v3(step), (C.c_long * 3)(*npts),
                                       s.ctypes.data_as(_dp), _v3(rotations), _v3(translation))
        if rc:
            raise ValueError("unsupported grid (rc=%d)" % rc)
        self._added(self.F)
        self.traj = None

    def get_bfield(self, x, t=0.0):
        out = np.zeros(3)
        self.L.orc_get_b(C.byref(self.F), _v3(x), float(t), out.ctypes.data_as(_dp))
        return out

    def get_efield(self, x, t=0.0):
        out = np.zeros(3)
        self.L.orc_get_b(C.byref(self.FE), _v3(x), float(t), out.ctypes.data_as(_dp))
        return out

    def set_particle_beam(self, energy_GeV, x0=(0, 0, 0), d0=(0, 0, 1), current=0.0, type="electron"):
        self.L.orc_particle_electron(C.byref(self.P), float(energy_GeV), _v3(x0), _v3(d0), float(current))
        self.set_ctstartstop(-3, 3)
        self.traj = None

    def set_particle(self, x0, beta0, t0=0.0):
        self.L.orc_particle_set(C.byref(self.P), _v3(x0), _v3(beta0), float(t0))
        self.traj = None

    def particle(self):
        P = self.P
        return dict(x0=list(P.x0), b0=list(P.b0), t0=P.t0, gamma=P.gamma, q=P.q, m=P.m, qomg=P.qomg,
                    current=P.current)

    def set_ctstartstop(self, a, b):
        self.npoints = int(self.npoints_per_meter * (b - a))
        self.ctstart, self.ctstop = float(a), float(b)
        self.traj = None

    def set_npoints_trajectory(self, n):
        self.npoints = int(n)
        self.traj = None

    # -- trajectory --
    def calculate_trajectory(self):
        nb = C.c_long(0)
        nf = self.L.orc_npoints_forward_backward(self.ctstart, self.ctstop, self.npoints, self.P.t0, C.byref(nb))
        n = nf + nb.value
        t = np.zeros(n)
        p = np.zeros((n, 9))
        self.L.orc_trajectory_rk4_eb(C.byref(self.F), C.byref(self.FE) if self.FE.n > 0 else None, C.byref(self.P),
                                     self.ctstart, self.ctstop, self.npoints, t.ctypes.data_as(_dp),
                                     p.ctypes.data_as(_dp))
        self.traj = Trajectory(t, p)
        return self.traj

    def halvings(self):
        """(steps the last trajectory integration split in two, the deepest split)."""
        self.L.orc_halving_count.restype = C.c_long
        return int(self.L.orc_halving_count()), int(self.L.orc_halving_depth())

    def _traj(self):
        if self.traj is None:
            self.calculate_trajectory()
        return self.traj

    def _opts(self, precision, max_level, max_level_extended, quantity, nthreads, carry_quirk, weight=1.0):
        o = orc_opts()
        o.precision = precision
        o.max_level = max_level
        o.max_level_extended = max_level_extended
        o.quantity = quantity
        o.nthreads = nthreads if nthreads else self.nthreads
        o.carry_quirk = int(carry_quirk)
        o.weight = weight
        return o

    @staticmethod
    def _quantity(q, pd=False):
        q = q.upper()
        if q == ("POWER DENSITY" if pd else "FLUX"):
            return 0
        if q == "PRECISION":
            return 1
        if q == "LEVEL":
            return 2
        raise ValueError("bad quantity")

    def spectrum(self, obs, energies_eV, polarization="", angle=None, horizontal_direction=(1, 0, 0),
                 propogation_direction=(0, 0, 1), precision=0.01, max_level=-2, max_level_extended=0,
                 quantity="flux", nthreads=0, carry_quirk=True, weight=1.0):
        """Returns (values, levels, converged, efield[n,3] complex)."""
        T = self._traj()
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        n = len(e)
        pol = polarization_vector(_pol_string(polarization, angle), angle or 0.0, horizontal_direction,
                                  propogation_direction)
        out = np.zeros(n)
        lev = np.zeros(n, dtype=np.int32)
        conv = np.zeros(n, dtype=np.uint8)
        ef = np.zeros((n, 6))
        o = self._opts(precision, max_level, max_level_extended, self._quantity(quantity), nthreads, carry_quirk,
                       weight)
        self.L.orc_spectrum(C.byref(T.c()), C.byref(self.P), _v3(obs), n, e.ctypes.data_as(_dp),
                            pol.ctypes.data_as(_dp), C.byref(o), out.ctypes.data_as(_dp),
                            lev.ctypes.data_as(_ip), conv.ctypes.data_as(_up), ef.ctypes.data_as(_dp))
        return out, lev, conv, ef[:, 0::2] + 1j * ef[:, 1::2]

    def energies(self, n, e0, e1):
        out = np.zeros(n)
        self.L.orc_energies(int(n), float(e0), float(e1), out.ctypes.data_as(_dp))
        return out

    def rectangle(self, plane, npoints, width, rotations=(0, 0, 0), translation=(0, 0, 0), normal=0, dim=2):
        n = npoints[0] * npoints[1]
        obs = np.zeros((n, 3))
        cont = np.zeros((n, 3))
        nv = np.zeros(3)
        rc = self.L.orc_rectangle(plane.encode(), int(npoints[0]), int(npoints[1]), float(width[0]),
                                  float(width[1]), _v3(rotations), _v3(translation), int(normal), int(dim),
                                  obs.ctypes.data_as(_dp), nv.ctypes.data_as(_dp), cont.ctypes.data_as(_dp))
        if rc:
            raise ValueError("not a valid surface")
        return obs, nv, cont

    def flux_points(self, obs_xyz, energy_eV, polarization="", angle=None, horizontal_direction=(1, 0, 0),
                    propogation_direction=(0, 0, 1), precision=0.01, max_level=-2, max_level_extended=0,
                    quantity="flux", nthreads=0, carry_quirk=True, weight=1.0):
        T = self._traj()
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        n = len(obs)
        pol = polarization_vector(_pol_string(polarization, angle), angle or 0.0, horizontal_direction,
                                  propogation_direction)
        out = np.zeros(n)
        lev = np.zeros(n, dtype=np.int32)
        conv = np.zeros(n, dtype=np.uint8)
        ef = np.zeros((n, 6))
        o = self._opts(precision, max_level, max_level_extended, self._quantity(quantity), nthreads, carry_quirk,
                       weight)
        self.L.orc_flux(C.byref(T.c()), C.byref(self.P), obs.ctypes.data_as(_dp), n, float(energy_eV),
                        pol.ctypes.data_as(_dp), C.byref(o), out.ctypes.data_as(_dp), lev.ctypes.data_as(_ip),
                        conv.ctypes.data_as(_up), ef.ctypes.data_as(_dp))
        return out, lev, conv, ef[:, 0::2] + 1j * ef[:, 1::2]

    def power_density_points(self, obs_xyz, normal, directional=False, precision=0.01, max_level=-2,
                             max_level_extended=0, quantity="power density", nthreads=0, carry_quirk=True,
                             weight=1.0):
        T = self._traj()
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        n = len(obs)
        out = np.zeros(n)
        lev = np.zeros(n, dtype=np.int32)
        conv = np.zeros(n, dtype=np.uint8)
        o = self._opts(precision, max_level, max_level_extended, self._quantity(quantity, pd=True), nthreads,
                       carry_quirk, weight)
        self.L.orc_power_density(C.byref(T.c()), C.byref(self.P), obs.ctypes.data_as(_dp), _v3(normal), n,
                                 int(directional), C.byref(o), out.ctypes.data_as(_dp), lev.ctypes.data_as(_ip),
                                 conv.ctypes.data_as(_up))
        return out, lev, conv

    def power_density_surface(self, obs_xyz, normals=None, directional=False, precision=0.01, max_level=-2,
                              max_level_extended=0, quantity="power density", nthreads=0, carry_quirk=True):
        """Power density at points with per-point unit normals[n, 3], or none."""
        T = self._traj()
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        n = len(obs)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(n, 3)
        out = np.zeros(n)
        lev = np.zeros(n, dtype=np.int32)
        conv = np.zeros(n, dtype=np.uint8)
        o = self._opts(precision, max_level, max_level_extended, self._quantity(quantity, pd=True), nthreads,
                       carry_quirk)
        self.L.orc_power_density_surface(C.byref(T.c()), C.byref(self.P), obs.ctypes.data_as(_dp),
                                         None if nrm is None else nrm.ctypes.data_as(_dp), n, int(directional),
                                         C.byref(o), out.ctypes.data_as(_dp), lev.ctypes.data_as(_ip),
                                         conv.ctypes.data_as(_up))
        return out, lev, conv

    def efield_time(self, obs, dt):
        """(t[n], E[n, 3]) at obs from the current trajectory's knots, step dt."""
        T = self._traj()
        n = len(T.t)
        t = np.zeros(n)
        e = np.zeros((n, 3))
        self.L.orc_efield_time(C.byref(T.c()), C.byref(self.P), _v3(obs), float(dt), t.ctypes.data_as(_dp),
                               e.ctypes.data_as(_dp))
        return t, e

    def power_density_stl(self, tri15, precision=0.01, max_level=-2, max_level_extended=0, quantity="power density",
                          carry_quirk=True):
        """tri15[n, 15] = A, B, C, normal, centre per triangle."""
        T = self._traj()
        t = np.ascontiguousarray(tri15, dtype=np.float64).reshape(-1, 15)
        n = len(t)
        out = np.zeros(n)
        lev = np.zeros(n, dtype=np.int32)
        conv = np.zeros(n, dtype=np.uint8)
        o = self._opts(precision, max_level, max_level_extended, self._quantity(quantity, pd=True), 1, carry_quirk)
        self.L.orc_power_density_stl(C.byref(T.c()), C.byref(self.P), t.ctypes.data_as(_dp), n, C.byref(o),
                                     out.ctypes.data_as(_dp), lev.ctypes.data_as(_ip), conv.ctypes.data_as(_up))
        return out, lev, conv

    def total_power(self, precision=0.01, max_level=-2, max_level_extended=0):
        """(power [W], precision, level); level -1 when not converged."""
        T = self._traj()
        o = self._opts(precision, max_level, max_level_extended, 0, 1, True)
        pw, pr, lv = C.c_double(0), C.c_double(0), C.c_int(0)
        self.L.orc_total_power(C.byref(T.c()), C.byref(self.P), C.byref(o), C.byref(pw), C.byref(pr), C.byref(lv))
        return pw.value, pr.value, lv.value

    # -- reference-shaped entry points --
    def calculate_spectrum(self, obs, npoints=0, energy_range_eV=None, energy_points_eV=None, **kw):
        if energy_points_eV is not None:
            e = np.asarray(energy_points_eV, dtype=np.float64)
        else:
            e = self.energies(npoints, energy_range_eV[0], energy_range_eV[1])
        v, lev, conv, ef = self.spectrum(obs, e, **kw)
        return e, v, lev, conv, ef

    def calculate_flux_rectangle(self, energy_eV, npoints, width, plane="XY", translation=(0, 0, 0),
                                 rotations=(0, 0, 0), normal=0, dim=2, **kw):
        obs, nv, cont = self.rectangle(plane, npoints, width, rotations, translation, normal, dim)
        v, lev, conv, ef = self.flux_points(obs, energy_eV, **kw)
        return cont, v, lev, conv, ef, obs

    def calculate_power_density_rectangle(self, npoints, width, plane="XY", translation=(0, 0, 0),
                                          rotations=(0, 0, 0), normal=0, dim=2, **kw):
        obs, nv, cont = self.rectangle(plane, npoints, width, rotations, translation, normal, dim)
        v, lev, conv = self.power_density_points(obs, nv, directional=(normal != 0), **kw)
        return cont, v, lev, conv, obs


_ref = None


def fdref_flux(samples, z0, dz, x0, beta0, current, ctstart, ctstop, npoints, obs, energies_eV, level,
               long_double=True):
    """Frozen-level flux (n_obs x n_e) through a 1-D z grid map from fdref.c, in
    long double (finite-difference reference) or double (checked against the oracle)."""
    L = oracle_lib()
    fn = L.fdref_flux_ld if long_double else L.fdref_flux_d
    fn.argtypes = [_dp, C.c_long, C.c_double, C.c_double, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_long,
                   _dp, C.c_long, _dp, C.c_long, C.c_int, _dp]
    s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
    obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    e = np.ascontiguousarray(energies_eV, dtype=np.float64).reshape(-1)
    out = np.zeros(len(obs) * len(e))
    rc = fn(s.ctypes.data_as(_dp), len(s), float(z0), float(dz), _v3(x0), _v3(beta0), float(current),
            float(ctstart), float(ctstop), int(npoints), obs.ctypes.data_as(_dp), len(obs), e.ctypes.data_as(_dp),
            len(e), int(level), out.ctypes.data_as(_dp))
    if rc:
        raise ValueError("fdref_flux: bad arguments")
    return out.reshape(len(obs), len(e))


def ref_available():
    return os.path.exists(LIB_REF)


_ref_mi = None


def ref_mi_available():
    return os.path.exists(LIB_REF_MI)


def ref_lib(with_gpu=False):
    """The reference core (libref.so), or with_gpu the same core compiled with
    its GPU boundary and linked against liboscars_mi.so (libref_mi.so; load
    torch first, as oscars_amd._lib does)."""
    global _ref, _ref_mi
    if with_gpu:
        if _ref_mi is None:
            _ref_mi = _bind_ref(C.CDLL(LIB_REF_MI))
            _ref_mi.ref_set_gpu.argtypes = [C.c_int]
            _ref_mi.ref_set_gpu.restype = None
        return _ref_mi
    if _ref is None:
        _ref = _bind_ref(C.CDLL(LIB_REF))
    return _ref


def _bind_ref(L):
    if True:
        L.ref_last_error.restype = C.c_char_p
        L.ref_new.restype = C.c_void_p
        L.ref_delete.argtypes = [C.c_void_p]
        L.ref_add_undulator.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp, C.c_double, C.c_double]
        L.ref_add_epu.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, _dp, C.c_double, C.c_double]
        L.ref_add_grid_file.argtypes = [C.c_void_p, C.c_char_p, _dp, _dp]
        L.ref_set_beam.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double]
        L.ref_add_gaussian.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.ref_add_field_file.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, _dp, _dp, _dp, C.c_int]
        L.ref_add_efield.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp]
        L.ref_add_field_interpolated.argtypes = [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_char_p), C.c_char_p,
                                                 C.c_double, _dp, _dp, _dp, C.c_int]
        L.ref_add_uniform.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.ref_add_quadrupole.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp]
        L.ref_set_particle.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.ref_get_particle.argtypes = [C.c_void_p, _dp]
        L.ref_set_ctstartstop.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.ref_set_npoints_trajectory.argtypes = [C.c_void_p, C.c_long]
        L.ref_get_bfield.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ref_get_efield.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ref_set_time_dependence.argtypes = [C.c_double, C.c_double, C.c_double]
        L.ref_set_time_dependence.restype = None
        L.ref_remove_efield.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_remove_efield.restype = None
        L.ref_set_beam_append.argtypes = [C.c_int, C.c_double]
        L.ref_set_beam_append.restype = None
        L.ref_calculate_trajectory.argtypes = [C.c_void_p]
        L.ref_trajectory_npoints.argtypes = [C.c_void_p]
        L.ref_trajectory_npoints.restype = C.c_long
        L.ref_trajectory.argtypes = [C.c_void_p, _dp]
        L.ref_spline_ypp.argtypes = [C.c_void_p, _dp]
        L.ref_level.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ref_spectrum.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, _dp, C.c_char_p,
                                   C.c_double, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _dp,
                                   _dp, _ip]
        L.ref_flux_rectangle.argtypes = [C.c_void_p, C.c_double, C.c_char_p, C.c_int, C.c_int, C.c_double,
                                         C.c_double, _dp, _dp, C.c_int, C.c_int, C.c_char_p, C.c_double, _dp,
                                         _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip]
        L.ref_power_density_rectangle.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_double,
                                                  C.c_double, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                  C.c_int, C.c_int, C.c_int, _dp, _dp, _ip]
        L.ref_set_beam_full.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                        C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_char_p, C.c_char_p]
        L.ref_set_seed.argtypes = [C.c_void_p, C.c_int]
        L.ref_set_nparticles.argtypes = [C.c_int]
        L.ref_set_corners.argtypes = [_dp]
        L.ref_set_trajectory_calculation.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.ref_write_trajectory.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p]
        L.ref_power_density_points.argtypes = [C.c_void_p, C.c_long, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double,
                                               C.c_int, C.c_int, C.c_int, _dp, _dp, _ip]
        L.ref_power_density_line.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_int, C.c_int, _dp, _dp, _ip]
        L.ref_total_power.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
        L.ref_efield_time.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_long), _dp, _dp]
        L.ref_add_stl.argtypes = [C.c_void_p, C.c_char_p, C.c_double, _dp, _dp, C.c_char_p]
        L.ref_clear_stl.argtypes = [C.c_void_p]
        L.ref_power_density_stl.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                            C.POINTER(C.c_long), _dp, _dp]
        L.ref_read_trajectory.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
        L.ref_set_trajectory_points.argtypes = [C.c_void_p, C.c_long, _dp, C.c_int, C.c_int]
        L.ref_random_normal.argtypes = [C.c_void_p]
        L.ref_random_normal.restype = C.c_double
        L.ref_random_uniform.argtypes = [C.c_void_p]
        L.ref_random_uniform.restype = C.c_double
        L.ref_new_particle.argtypes = [C.c_void_p]
        L.ref_write_spectrum.argtypes = [C.c_char_p, C.c_char_p, C.c_long, _dp, _dp]
        L.ref_write_points.argtypes = [C.c_char_p, C.c_char_p, C.c_long, _dp, _dp, C.c_int]
        L.ref_average_spectra.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_char_p, C.c_char_p,
                                          C.POINTER(C.c_long)]
        L.ref_average_points.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_char_p,
                                         C.POINTER(C.c_long), _dp, _dp, C.c_long]
    return L


def _ref_chk(L, rc):
    if rc:
        raise ValueError(L.ref_last_error().decode())


def ref_write_spectrum(text_path, bin_path, energies, values):
    """The reference's TSpectrumContainer writers on the given values."""
    L = ref_lib()
    e = np.ascontiguousarray(energies, dtype=np.float64)
    v = np.ascontiguousarray(values, dtype=np.float64)
    _ref_chk(L, L.ref_write_spectrum((text_path or "").encode(), (bin_path or "").encode(), len(e),
                                     e.ctypes.data_as(_dp), v.ctypes.data_as(_dp)))


def ref_write_points(text_path, bin_path, xyz, values, dim):
    """The reference's T3DScalarContainer writers on the given points and values."""
    L = ref_lib()
    p = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(values, dtype=np.float64)
    _ref_chk(L, L.ref_write_points((text_path or "").encode(), (bin_path or "").encode(), len(p),
                                   p.ctypes.data_as(_dp), v.ctypes.data_as(_dp), int(dim)))


def ref_average_spectra(files, binary, out_text="", out_bin=""):
    L = ref_lib()
    arr = (C.c_char_p * len(files))(*[f.encode() for f in files])
    n = C.c_long(0)
    _ref_chk(L, L.ref_average_spectra(int(binary), len(files), arr, out_text.encode(), out_bin.encode(),
                                      C.byref(n)))
    return n.value


def ref_average_points(files, binary, dim, out_text="", out_bin="", cap=1 << 20):
    L = ref_lib()
    arr = (C.c_char_p * len(files))(*[f.encode() for f in files])
    n = C.c_long(0)
    xyz = np.zeros((cap, 3))
    v = np.zeros(cap)
    _ref_chk(L, L.ref_average_points(int(binary), len(files), arr, int(dim), out_text.encode(), out_bin.encode(),
                                     C.byref(n), xyz.ctypes.data_as(_dp), v.ctypes.data_as(_dp), cap))
    return xyz[:n.value], v[:n.value]


_NOANGLE = 112188979912321.0


class RefSim:
    """Scenario-level driver of the reference's compiled C++ core."""

    def __init__(self, nthreads=1, gpu=False):
        """gpu: the reference core whose GPU boundary is this repository's
        liboscars_mi.so (libref_mi.so); its calculations then pass gpu=1."""
        self.L = ref_lib(with_gpu=gpu)
        self.h = self.L.ref_new()
        self.nthreads = nthreads
        self.L.ref_set_time_dependence(0.0, 0.0, 0.0)
        if gpu:
            self.L.ref_set_gpu(1)

    def __del__(self):
        try:
            self.L.ref_delete(self.h)
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise ValueError(self.L.ref_last_error().decode())

    def add_bfield_undulator(self, bfield, period, nperiods, phase=0.0, translation=(0, 0, 0), taper=0.0):
        self._chk(self.L.ref_add_undulator(self.h, _v3(bfield), _v3(period), int(nperiods), _v3(translation),
                                           float(phase), float(taper)))

    def add_bfield_epu(self, bfieldA, bfieldB, period, nperiods, phase=0.0, translation=(0, 0, 0), taper=0.0):
        self._chk(self.L.ref_add_epu(self.h, _v3(bfieldA), _v3(bfieldB), _v3(period), int(nperiods),
                                     _v3(translation), float(phase), float(taper)))

    def add_bfield_file_format(self, ifile, iformat="OSCARS", rotations=(0, 0, 0), translation=(0, 0, 0), scale=()):
        sc = np.ascontiguousarray(list(scale) or [0.0], dtype=np.float64)
        self._chk(self.L.ref_add_field_file(self.h, ifile.encode(), iformat.encode(), _v3(rotations), _v3(translation),
                                            sc.ctypes.data_as(_dp), len(scale)))

    def add_efield_gaussian(self, efield, sigma, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_efield(self.h, 3, _v3(efield), _v3(sigma), _v3((0, 0, 0)), 0, _v3(rotations),
                                        _v3(translation)))

    def add_efield_uniform(self, efield, width=(0, 0, 0), rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_efield(self.h, 4, _v3(efield), _v3(width), _v3((0, 0, 0)), 0, _v3(rotations),
                                        _v3(translation)))

    def add_efield_undulator(self, efield, period, nperiods, translation=(0, 0, 0)):
        self._chk(self.L.ref_add_efield(self.h, 0, _v3(efield), _v3((0, 0, 0)), _v3(period), int(nperiods),
                                        _v3((0, 0, 0)), _v3(translation)))

    def add_bfield_interpolated(self, mapping, iformat, parameter, rotations=(0, 0, 0), translation=(0, 0, 0),
                                scale=()):
        params = np.ascontiguousarray([float(m[0]) for m in mapping], dtype=np.float64)
        files = (C.c_char_p * len(mapping))(*[str(m[1]).encode() for m in mapping])
        sc = np.ascontiguousarray(list(scale) or [0.0], dtype=np.float64)
        self._chk(self.L.ref_add_field_interpolated(self.h, len(mapping), params.ctypes.data_as(_dp), files,
                                                    iformat.encode(), float(parameter), _v3(rotations),
                                                    _v3(translation), sc.ctypes.data_as(_dp), len(scale)))

    def add_bfield_gaussian(self, bfield, sigma, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_gaussian(self.h, _v3(bfield), _v3(translation), _v3(sigma), _v3(rotations)))

    def add_bfield_uniform(self, bfield, width=(0, 0, 0), rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_uniform(self.h, _v3(bfield), _v3(width), _v3(translation), _v3(rotations)))

    def add_bfield_quadrupole(self, K, width, rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_quadrupole(self.h, float(K), float(width), _v3(rotations), _v3(translation)))

    def add_bfield_file(self, ifile, iformat="OSCARS", rotations=(0, 0, 0), translation=(0, 0, 0)):
        self._chk(self.L.ref_add_grid_file(self.h, ifile.encode(), _v3(rotations), _v3(translation)))

    def get_bfield(self, x, t=0.0):
        out = np.zeros(3)
        self.L.ref_get_bfield(self.h, _v3(x), float(t), out.ctypes.data_as(_dp))
        return out

    def get_efield(self, x, t=0.0):
        out = np.zeros(3)
        self.L.ref_get_efield(self.h, _v3(x), float(t), out.ctypes.data_as(_dp))
        return out

    def set_time_dependence(self, frequency=0.0, frequency_phase=0.0, time_offset=0.0):
        """frequency, frequency_phase and time_offset of the fields added after
        this call (the binding's keyword arguments of every add_*field)."""
        self.L.ref_set_time_dependence(float(frequency), float(frequency_phase), float(time_offset))

    def remove_efield(self, name=""):
        self.L.ref_remove_efield(self.h, name.encode())

    def set_particle_beam(self, energy_GeV, x0=(0, 0, 0), d0=(0, 0, 1), current=0.0, type="electron"):
        self._chk(self.L.ref_set_beam(self.h, float(energy_GeV), _v3(x0), _v3(d0), float(current)))

    def set_particle_beam_full(self, energy_GeV, x0=(0, 0, 0), d0=(0, 0, 1), t0=0.0, current=0.0,
                               horizontal_direction=None, emittance=None, sigma_energy_GeV=0.0, beta=None,
                               alpha=None, gamma=None, lattice_reference=None, eta=None, distribution=None,
                               beam=None):
        """set_particle_beam with the beam-spread arguments; None = not given
        (with a predefined ``beam``, d0=None and energy_GeV < 0 too)."""
        def opt(v, n):
            if v is None:
                return None
            a = np.ascontiguousarray(v, dtype=np.float64).reshape(n)
            self._keep.append(a)
            return a.ctypes.data_as(_dp)
        self._keep = []
        self._chk(self.L.ref_set_beam_full(self.h, float(energy_GeV), _v3(x0), opt(d0, 3), float(t0), float(current),
                                           opt(horizontal_direction, 3), opt(emittance, 2), float(sigma_energy_GeV),
                                           opt(beta, 2), opt(alpha, 2), opt(gamma, 2), opt(lattice_reference, 3),
                                           opt(eta, 2), None if distribution is None else distribution.encode(),
                                           None if beam is None else beam.encode()))

    def add_particle_beam_full(self, weight=1.0, **kw):
        """add_particle_beam: set_particle_beam_full keeping the beams there."""
        self.L.ref_set_beam_append(1, float(weight))
        try:
            self.set_particle_beam_full(**kw)
        finally:
            self.L.ref_set_beam_append(0, 1.0)

    def set_seed(self, seed):
        self.L.ref_set_seed(self.h, int(seed))

    def random_normal(self):
        return self.L.ref_random_normal(self.h)

    def random_uniform(self):
        return self.L.ref_random_uniform(self.h)

    def new_particle(self):
        """Draw the next random particle of the beam and make it current."""
        self._chk(self.L.ref_new_particle(self.h))

    def set_particle(self, x0, beta0, t0=0.0):
        self._chk(self.L.ref_set_particle(self.h, _v3(x0), _v3(beta0), float(t0)))

    def particle(self):
        o = np.zeros(12)
        self.L.ref_get_particle(self.h, o.ctypes.data_as(_dp))
        return dict(x0=list(o[0:3]), b0=list(o[3:6]), t0=o[6], gamma=o[7], q=o[8], m=o[9], qomg=o[10],
                    current=o[11])

    def set_ctstartstop(self, a, b):
        self.L.ref_set_ctstartstop(self.h, float(a), float(b))

    def set_npoints_trajectory(self, n):
        self.L.ref_set_npoints_trajectory(self.h, int(n))

    def calculate_trajectory(self):
        self._chk(self.L.ref_calculate_trajectory(self.h))
        n = self.L.ref_trajectory_npoints(self.h)
        a = np.zeros((n, 10))
        self.L.ref_trajectory(self.h, a.ctypes.data_as(_dp))
        ypp = np.zeros((n, 9))
        self.L.ref_spline_ypp(self.h, ypp.ctypes.data_as(_dp))
        return a[:, 0].copy(), a[:, 1:].copy(), ypp

    def trajectory(self):
        """Knots of the current trajectory: (t[n], p9[n, 9], spline ypp[n, 9])."""
        n = self.L.ref_trajectory_npoints(self.h)
        a = np.zeros((n, 10))
        self.L.ref_trajectory(self.h, a.ctypes.data_as(_dp))
        ypp = np.zeros((n, 9))
        self.L.ref_spline_ypp(self.h, ypp.ctypes.data_as(_dp))
        return a[:, 0].copy(), a[:, 1:].copy(), ypp

    def calculate_power_density(self, points, normal=0, rotations=(0, 0, 0), translation=(0, 0, 0), precision=0.01,
                                max_level=-2, max_level_extended=0, quantity="power density", nthreads=0,
                                nparticles=0):
        q = {"POWER DENSITY": 0, "POWERDENSITY": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        self.L.ref_set_nparticles(int(nparticles))
        xn = np.ascontiguousarray([list(p[0]) + list(p[1]) for p in points], dtype=np.float64)
        n = len(xn)
        xyz = np.zeros((n, 3))
        v = np.zeros(n)
        conv = C.c_int(0)
        self._chk(self.L.ref_power_density_points(self.h, n, xn.ctypes.data_as(_dp), int(normal), _v3(rotations),
                                                  _v3(translation), nthreads or self.nthreads, float(precision),
                                                  int(max_level), int(max_level_extended), q,
                                                  xyz.ctypes.data_as(_dp), v.ctypes.data_as(_dp), C.byref(conv)))
        return xyz, v, bool(conv.value)

    def calculate_power_density_line(self, x1, x2, npoints, normal=0, dim=1, precision=0.01, max_level=-2,
                                     max_level_extended=0, nthreads=0, nparticles=0):
        self.L.ref_set_nparticles(int(nparticles))
        xyz = np.zeros((npoints, 3))
        v = np.zeros(npoints)
        conv = C.c_int(0)
        self._chk(self.L.ref_power_density_line(self.h, _v3(x1), _v3(x2), int(npoints), int(normal), int(dim),
                                                nthreads or self.nthreads, float(precision), int(max_level),
                                                int(max_level_extended), xyz.ctypes.data_as(_dp),
                                                v.ctypes.data_as(_dp), C.byref(conv)))
        return xyz, v, bool(conv.value)

    def add_stl(self, ifile, rotations=(0, 0, 0), translation=(0, 0, 0), scale=1.0, name=""):
        self._chk(self.L.ref_add_stl(self.h, ifile.encode(), float(scale), _v3(rotations), _v3(translation),
                                     name.encode()))

    def clear_stl(self):
        self.L.ref_clear_stl(self.h)

    def calculate_power_density_stl(self, precision=0.01, max_level=-2, max_level_extended=0,
                                    quantity="power density", nparticles=0, stlofile=None):
        """(triangles[n, 3, 3], values[n]) as the reference computes them."""
        q = {"POWER DENSITY": 0, "POWERDENSITY": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        self.L.ref_set_nparticles(int(nparticles))
        n = C.c_long(0)
        args = (float(precision), int(max_level), int(max_level_extended), q, 1,
                stlofile.encode() if stlofile else None)
        self._chk(self.L.ref_power_density_stl(self.h, *args, C.byref(n), None, None))
        tri = np.zeros((n.value, 3, 3))
        v = np.zeros(n.value)
        self.L.ref_set_nparticles(int(nparticles))
        self._chk(self.L.ref_power_density_stl(self.h, *args, C.byref(n), tri.ctypes.data_as(_dp),
                                               v.ctypes.data_as(_dp)))
        return tri, v

    def calculate_efield_vs_time(self, obs):
        """(t[n], E[n, 3]) as the reference computes them."""
        n = C.c_long(0)
        self._chk(self.L.ref_efield_time(self.h, _v3(obs), C.byref(n), None, None))
        t = np.zeros(n.value)
        e = np.zeros((n.value, 3))
        self._chk(self.L.ref_efield_time(self.h, _v3(obs), C.byref(n), t.ctypes.data_as(_dp), e.ctypes.data_as(_dp)))
        return t, e

    def calculate_total_power(self, precision=0.01, max_level=-2, max_level_extended=0, quantity="power"):
        q = {"POWER": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        out = C.c_double(0)
        self._chk(self.L.ref_total_power(self.h, float(precision), int(max_level), int(max_level_extended), q,
                                         C.byref(out)))
        return out.value

    def set_trajectory_calculation(self, method="RK4", precision=-1):
        self._chk(self.L.ref_set_trajectory_calculation(self.h, method.encode(), float(precision)))

    def set_corners(self, x0x1x2=None):
        """Rectangles from three corners (x0x1x2) instead of plane and widths; None: off."""
        if x0x1x2 is None:
            self.L.ref_set_corners(None)
        else:
            a = np.ascontiguousarray(x0x1x2, dtype=np.float64).reshape(9)
            self.L.ref_set_corners(a.ctypes.data_as(_dp))

    def write_trajectory(self, text_path=None, bin_path=None, fmt=""):
        self._chk(self.L.ref_write_trajectory(self.h, text_path.encode() if text_path else None,
                                              bin_path.encode() if bin_path else None, fmt.encode()))

    def read_trajectory(self, path, fmt="", binary=False):
        self._chk(self.L.ref_read_trajectory(self.h, path.encode(), fmt.encode(), int(binary)))

    def set_trajectory_points(self, t, p9, has_b=True, has_bp=True):
        rows = np.ascontiguousarray(np.column_stack([t, p9]), dtype=np.float64)
        self._chk(self.L.ref_set_trajectory_points(self.h, len(rows), rows.ctypes.data_as(_dp), int(has_b),
                                                   int(has_bp)))

    def level(self, L):
        out = np.zeros((2 ** L, 10))
        self._chk(self.L.ref_level(self.h, int(L), out.ctypes.data_as(_dp)))
        return out

    def calculate_spectrum(self, obs, npoints=0, energy_range_eV=None, energy_points_eV=None, polarization="",
                           angle=None, horizontal_direction=(1, 0, 0), propogation_direction=(0, 0, 1),
                           precision=0.01, max_level=-2, max_level_extended=0, quantity="flux", nthreads=0,
                           nparticles=0):
        q = {"FLUX": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        self.L.ref_set_nparticles(int(nparticles))
        if energy_points_eV is not None:
            e_in = np.ascontiguousarray(energy_points_eV, dtype=np.float64)
            n = len(e_in)
            ep = e_in.ctypes.data_as(_dp)
            e0 = e1 = 0.0
        else:
            n = npoints
            ep = None
            e0, e1 = energy_range_eV
        e = np.zeros(n)
        v = np.zeros(n)
        conv = C.c_int(0)
        self._chk(self.L.ref_spectrum(self.h, _v3(obs), n, float(e0), float(e1), ep, polarization.encode(),
                                      _NOANGLE if angle is None else float(angle), _v3(horizontal_direction),
                                      _v3(propogation_direction), nthreads or self.nthreads, float(precision),
                                      int(max_level), int(max_level_extended), q, e.ctypes.data_as(_dp),
                                      v.ctypes.data_as(_dp), C.byref(conv)))
        return e, v, bool(conv.value)

    def calculate_flux_rectangle(self, energy_eV, npoints, width, plane="XY", translation=(0, 0, 0),
                                 rotations=(0, 0, 0), normal=0, dim=2, polarization="", angle=None,
                                 horizontal_direction=(1, 0, 0), propogation_direction=(0, 0, 1), precision=0.01,
                                 max_level=-2, max_level_extended=0, quantity="flux", nthreads=0, nparticles=0):
        q = {"FLUX": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        self.L.ref_set_nparticles(int(nparticles))
        n = npoints[0] * npoints[1]
        xyz = np.zeros((n, 3))
        v = np.zeros(n)
        conv = C.c_int(0)
        self._chk(self.L.ref_flux_rectangle(self.h, float(energy_eV), plane.encode(), int(npoints[0]),
                                            int(npoints[1]), float(width[0]), float(width[1]), _v3(rotations),
                                            _v3(translation), int(normal), int(dim), polarization.encode(),
                                            _NOANGLE if angle is None else float(angle),
                                            _v3(horizontal_direction), _v3(propogation_direction),
                                            nthreads or self.nthreads, float(precision), int(max_level),
                                            int(max_level_extended), q, xyz.ctypes.data_as(_dp),
                                            v.ctypes.data_as(_dp), C.byref(conv)))
        return xyz, v, bool(conv.value)

    def calculate_power_density_rectangle(self, npoints, width, plane="XY", translation=(0, 0, 0),
                                          rotations=(0, 0, 0), normal=0, dim=2, precision=0.01, max_level=-2,
                                          max_level_extended=0, quantity="power density", nthreads=0,
                                          nparticles=0):
        q = {"POWER DENSITY": 0, "PRECISION": 1, "LEVEL": 2}[quantity.upper()]
        self.L.ref_set_nparticles(int(nparticles))
        n = npoints[0] * npoints[1]
        xyz = np.zeros((n, 3))
        v = np.zeros(n)
        conv = C.c_int(0)
        self._chk(self.L.ref_power_density_rectangle(self.h, plane.encode(), int(npoints[0]), int(npoints[1]),
                                                     float(width[0]), float(width[1]), _v3(rotations),
                                                     _v3(translation), int(normal), int(dim),
                                                     nthreads or self.nthreads, float(precision), int(max_level),
                                                     int(max_level_extended), q, xyz.ctypes.data_as(_dp),
                                                     v.ctypes.data_as(_dp), C.byref(conv)))
        return xyz, v, bool(conv.value)
