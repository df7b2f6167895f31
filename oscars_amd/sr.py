"""Drop-in ``oscars.sr``-style simulation object running on one MI355X.

Mirrors the reference Python API for the radiation hot path
(src/OSCARSSR_Python.cc: ``set_particle_beam`` :3561, ``add_bfield_undulator``
:1309, ``add_bfield_epu`` :1479, ``add_bfield_file`` :526, ``set_ctstartstop``
:345, ``set_npoints_trajectory`` :409, ``calculate_trajectory`` :5181,
``calculate_spectrum`` :5653, ``calculate_flux_rectangle`` :7977,
``calculate_power_density_rectangle`` :6471, ``set_trajectory`` :5374,
``set_seed``, ``set_new_particle`` :4589, ``average_spectra`` :8460): same
keyword names, defaults, return shapes and ValueError messages.  Every
calculation runs on one GPU through liboscars_mi.so; there is no CPU path, so
``gpu=0`` still uses the GPU and ``ngpu`` is accepted and ignored.

Extensions beyond the reference (keyword ``mode`` and the method
``calculate_flux_rectangle_energies``) select the kernel family and evaluate
many energies in one call.
"""
import math
import sys

import numpy as np

from . import _lib
from . import beam as B
from . import geometry as G
from . import io as IO

_NOANGLE = 112188979912321


def _v3(x, what):
    try:
        v = [float(a) for a in x]
    except TypeError:
        raise ValueError("Incorrect format in '%s'" % what)
    if len(v) != 3:
        raise ValueError("Incorrect format in '%s'" % what)
    return v


def _v2(x, what):
    try:
        v = [float(a) for a in x]
    except TypeError:
        raise ValueError("Incorrect format in '%s'" % what)
    if len(v) != 2:
        raise ValueError("Incorrect format in '%s'" % what)
    return v


def _warn(msg):
    sys.stderr.write(msg)


def _check_pd_mode(mode):
    if mode not in ("plain", "tiled", "mixed", "point_mixed"):
        raise ValueError("mode must be 'plain', 'tiled', 'mixed' or 'point_mixed'")


def _named(f, name):
    """Tag a field with its name (get_bfield / get_efield select by it)."""
    f.name = name
    return f


def surface_points(points, normal=0, rotations=None, translation=None):
    """Observation points and unit normals of calculate_power_density: each
    [[x, y, z], [nx, ny, nz]] rotated, the normal reversed for normal = -1,
    the point translated (src/OSCARSSR_Python.cc:6208-6240, src/TSurfacePoint.cc:121)."""
    rot = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
    tr = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
    obs, nrm = [], []
    for p in points:
        if len(p) != 2:
            raise ValueError("Incorrect format in 'points'")
        try:
            X, N = _v3(p[0], "points"), _v3(p[1], "points")
        except ValueError:
            raise ValueError("Incorrect format in 'points': Point or Normal does not have 3 elements")
        X, N = G.rot_xyz(X, rot), G.rot_xyz(N, rot)
        if normal == -1:
            N = [-N[0], -N[1], -N[2]]
        obs.append([X[k] + tr[k] for k in range(3)])
        nrm.append(G.unit(N))
    return obs, nrm


class sr:
    """Simulation object: fields, one particle beam, trajectory window, and the
    spectrum / flux / power-density calculations."""

    def __init__(self, nthreads=2, gpu=0, device=0):
        self._fields = []        # list of dicts kept for om_field construction
        self._efields = []       # electric fields (om_field)
        self._stale_eb = None    # E+B equations kept after a removal (None: follow the fields)
        self._beam = None        # the first beam's description
        self._beams = []         # (name, description, weight) per beam
        self._wsum = []          # running sums of the weights
        self._particle = None
        self._ctstart = 0.0
        self._ctstop = 0.0
        self._npoints = 0
        self._npoints_per_meter = 10000
        self._nthreads = nthreads
        self._gpu_global = gpu
        self._device_index = device
        self._dev = None
        self._traj_valid = False
        self.last_stats = None
        self._stl = []           # (name, tri[n, 15], scale) per STL object
        self._sum_spectrum = []  # add_to_spectrum: [[E, V], ...]
        self._sum_flux = []      # add_to_flux: [[[x, y, z], V], ...]
        self._sum_pd = []        # add_to_power_density: as add_to_flux

    # ------------------------------------------------------------ device
    def _device(self):
        if self._dev is None:
            self._dev = _lib.Device(self._device_index)
        return self._dev

    def print_gpu(self):
        print(self._device().info())

    def check_gpu(self):
        """Number of GPUs the library sees (src/OSCARSSR_Python.cc:181-190)."""
        return _lib.device_count()

    def set_gpu_global(self, gpu):
        """Kept for the reference's API: every calculation here runs on the GPU."""
        if gpu not in (0, 1):
            raise ValueError("'gpu' must be 0 or 1")
        self._gpu_global = int(gpu)

    def set_nthreads_global(self, nthreads):
        """Kept for the reference's API (its CPU thread count); the GPU path has
        no host threads to set."""
        if nthreads < 1:
            raise ValueError("'nthreads' must be >= 1")
        self._nthreads = int(nthreads)

    # ------------------------------------------------------------ fields
    @staticmethod
    def _check_name(name):
        if len(name) > 0 and name[0] == "_":
            raise ValueError("'name' cannot begin with '_'.  This is reserved for internal use.  Please pick a "
                             "different name")

    @staticmethod
    def _set_time(f, frequency, frequency_phase, time_offset):
        """The factor cos(2 pi frequency (t + time_offset) + frequency_phase) of
        every model (src/TField3D_IdealUndulator.cc:147-151, TField3D_IdealEPU.cc:165-169,
        TField3D_Gaussian.cc:107-111, TField3D_UniformBox.cc:115-119,
        TField3D_Quadrupole.cc:78-82, TField3D_Grid.cc:277)."""
        f.frequency = float(frequency)
        f.frequency_phase = float(frequency_phase)
        f.time_offset = float(time_offset)
        return f

    def _undulator(self, ftype, fieldA, fieldB, period, nperiods, phase, rotations, translation, taper, frequency,
                   frequency_phase, time_offset, name):
        self._check_name(name)
        rot = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        tr = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        f = _lib.om_field()
        f.type = ftype
        f.field[:] = G.rot_xyz(fieldA, rot)
        f.fieldB[:] = G.rot_xyz(fieldB, rot)
        f.period[:] = G.rot_xyz(_v3(period, "period"), rot)
        f.nperiods = int(nperiods)
        f.translation[:] = tr
        f.phase = float(phase)
        f.taper = float(taper)
        return _named(self._set_time(f, frequency, frequency_phase, time_offset), name)

    def add_bfield_undulator(self, bfield, period, nperiods, phase=0.0, rotations=None, translation=None, taper=0.0,
                             frequency=0.0, frequency_phase=0.0, time_offset=0.0, name=""):
        """Ideal undulator (src/OSCARSSR_Python.cc:1309-1478, src/TField3D_IdealUndulator.cc:63-152)."""
        self._push_b(self._undulator(_lib.OM_FIELD_UNDULATOR, _v3(bfield, "bfield"), [0.0, 0.0, 0.0], period,
                                     nperiods, phase, rotations, translation, taper, frequency, frequency_phase,
                                     time_offset, name))
        self._traj_valid = False

    def add_bfield_epu(self, period, nperiods, bfieldA=None, bfieldB=None, phase=0.0, rotations=None,
                       translation=None, taper=0.0, frequency=0.0, frequency_phase=0.0, time_offset=0.0, name=""):
        """Ideal elliptically polarising undulator: two undulator fields with
        phase shifts of +-phase/2 (src/OSCARSSR_Python.cc:1479-1617,
        src/TField3D_IdealEPU.cc:64-170)."""
        if not bfieldA and not bfieldB:
            raise ValueError("Must define at least one of 'bfieldA', 'bfieldB'")
        fa = _v3(bfieldA, "bfieldA") if bfieldA else [0.0, 0.0, 0.0]
        fb = _v3(bfieldB, "bfieldB") if bfieldB else [0.0, 0.0, 0.0]
        self._push_b(self._undulator(_lib.OM_FIELD_EPU, fa, fb, period, nperiods, phase, rotations, translation,
                                     taper, frequency, frequency_phase, time_offset, name))
        self._traj_valid = False

    def _simple_field(self, ftype, name, frequency, frequency_phase, time_offset, rotations, translation):
        self._check_name(name)
        f = _lib.om_field()
        f.type = ftype
        f.rotations[:] = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        f.translation[:] = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        return self._set_time(f, frequency, frequency_phase, time_offset)

    def add_bfield_gaussian(self, bfield, sigma, rotations=None, translation=None, frequency=0, frequency_phase=0,
                            time_offset=0, name=""):
        """Gaussian field bfield * prod exp(-(x_k / sigma_k)^2 / 2) about the
        translation; sigma_k <= 0 leaves axis k flat (src/OSCARSSR_Python.cc,
        add_bfield_gaussian; src/TField3D_Gaussian.cc:27-112)."""
        B, S = _v3(bfield, "bfield"), _v3(sigma, "sigma")
        f = self._simple_field(_lib.OM_FIELD_GAUSSIAN, name, frequency, frequency_phase, time_offset, rotations,
                               translation)
        f.field[:] = B
        f.shape[:] = S
        self._push_b(_named(f, name))
        self._traj_valid = False

    def add_bfield_uniform(self, bfield, width=None, rotations=None, translation=None, frequency=0,
                           frequency_phase=0, time_offset=0, name=""):
        """Uniform field inside a box of full widths `width` about the
        translation (a width <= 0 leaves that axis unbounded)
        (src/TField3D_UniformBox.cc:45-120)."""
        B = _v3(bfield, "bfield")
        W = _v3(width, "width") if width is not None and len(width) != 0 else [0.0, 0.0, 0.0]
        f = self._simple_field(_lib.OM_FIELD_UNIFORM, name, frequency, frequency_phase, time_offset, rotations,
                               translation)
        f.field[:] = B
        f.shape[:] = W
        self._push_b(_named(f, name))
        self._traj_valid = False

    def add_bfield_quadrupole(self, K, width, rotations=None, translation=None, frequency=0, frequency_phase=0,
                              time_offset=0, name=""):
        """Quadrupole (K y, K x, 0) [T] within |z| <= width / 2 of its frame
        (src/TField3D_Quadrupole.cc:20-83)."""
        f = self._simple_field(_lib.OM_FIELD_QUADRUPOLE, name, frequency, frequency_phase, time_offset, rotations,
                               translation)
        f.K = float(K)
        f.width = float(width)
        self._push_b(_named(f, name))
        self._traj_valid = False

    def add_bfield_halbach(self, bfield, period, nperiods, gap=0.0, magnet_height=0.0, magnet_width=0.0,
                           npoints_per_period=4, name=""):
        """The reference's Halbach model returns a zero field everywhere
        (src/TField3D_Halbach.cc:61-68) and its binding is not registered in
        the module's method table (src/OSCARSSR_Python.cc:9385, 9518); a zero
        contributes nothing to the field sum, so it is accepted and not sent to
        the device."""
        self._check_name(name)
        _v3(bfield, "bfield") if not isinstance(bfield, (int, float)) else None

    # ------------------------------------------------------------ electric fields
    def _add_efield(self, f):
        self._efields.append(f)
        self._stale_eb = None
        self._traj_valid = False

    def add_efield_gaussian(self, efield, sigma, rotations=None, translation=None, frequency=0, frequency_phase=0,
                            time_offset=0, name=""):
        """Gaussian electric field [V/m] (the Gaussian model of add_bfield_gaussian)."""
        E, S = _v3(efield, "efield"), _v3(sigma, "sigma")
        f = self._simple_field(_lib.OM_FIELD_GAUSSIAN, name, frequency, frequency_phase, time_offset, rotations,
                               translation)
        f.field[:] = E
        f.shape[:] = S
        self._add_efield(_named(f, name))

    def add_efield_uniform(self, efield, width=None, rotations=None, translation=None, frequency=0,
                           frequency_phase=0, time_offset=0, name=""):
        """Uniform electric field [V/m] inside a box (the model of add_bfield_uniform)."""
        E = _v3(efield, "efield")
        W = _v3(width, "width") if width is not None and len(width) != 0 else [0.0, 0.0, 0.0]
        f = self._simple_field(_lib.OM_FIELD_UNIFORM, name, frequency, frequency_phase, time_offset, rotations,
                               translation)
        f.field[:] = E
        f.shape[:] = W
        self._add_efield(_named(f, name))

    def add_efield_undulator(self, efield, period, nperiods, phase=0.0, rotations=None, translation=None, taper=0.0,
                             frequency=0, frequency_phase=0, time_offset=0, name=""):
        """Ideal-undulator electric field [V/m] (the model of add_bfield_undulator,
        src/OSCARSSR_Python.cc:2859-2990)."""
        self._add_efield(self._undulator(_lib.OM_FIELD_UNDULATOR, _v3(efield, "efield"), [0.0, 0.0, 0.0], period,
                                         nperiods, phase, rotations, translation, taper, frequency, frequency_phase,
                                         time_offset, name))

    def add_efield_file(self, ifile, iformat, rotations=None, translation=None, scale=None, frequency=0,
                        frequency_phase=0, time_offset=0, name=""):
        """Electric-field map [V/m] from a file, in any format add_bfield_file reads
        (src/OSCARSSR_Python.cc:2102-2252)."""
        n0 = len(self._fields)
        self.add_bfield_file(ifile, iformat=iformat, rotations=rotations, translation=translation, scale=scale,
                             frequency=frequency, frequency_phase=frequency_phase, time_offset=time_offset,
                             name=name)
        self._add_efield(self._fields.pop(n0))

    def add_efield_interpolated(self, mapping, iformat, parameter, rotations=None, translation=None, scale=None,
                                frequency=0, frequency_phase=0, time_offset=0, name=""):
        """Electric-field map interpolated between maps, as add_bfield_interpolated."""
        n0 = len(self._fields)
        self.add_bfield_interpolated(mapping, iformat, parameter, rotations, translation, scale, frequency,
                                     frequency_phase, time_offset, name)
        self._add_efield(self._fields.pop(n0))

    def _push_b(self, f):
        self._fields.append(f)
        self._stale_eb = None

    def _efields_for_derivs(self):
        # src/OSCARSSR.cc:1185-1200: the equations of motion are chosen when a
        # field is added or the fields are cleared, not when one is removed; a
        # removal that leaves magnetic fields only keeps the E+B equations,
        # which a zero electric field reproduces here
        if self._stale_eb and not self._efields and self._fields:
            f = _lib.om_field()
            f.type = _lib.OM_FIELD_UNIFORM
            return [f]
        return self._efields

    def _remove(self, fields, name):
        if self._stale_eb is None:
            self._stale_eb = len(self._efields) > 0 or len(self._fields) == 0
        # src/TFieldContainer.cc:52-70: every field whose name equals `name`
        # goes, so "" removes the unnamed ones
        return [f for f in fields if getattr(f, "name", "") != name]

    def remove_bfield(self, name):
        """Remove the magnetic fields of this name (src/TFieldContainer.cc:52-70);
        the current trajectory stays, as in the reference."""
        self._fields = self._remove(self._fields, name)

    def remove_efield(self, name):
        """Remove the electric fields of this name."""
        self._efields = self._remove(self._efields, name)

    def clear_efields(self):
        self._efields = []
        self._stale_eb = None

    def add_bfield_grid(self, start, step, npoints, samples, rotations=None, translation=None):
        """Regular 1-D field map from header values and a (n, 3) sample array
        (what add_bfield_file reads from an OSCARS-format file)."""
        s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
        f = _lib.om_field()
        f.type = _lib.OM_FIELD_GRID
        f.grid_start[:] = [float(v) for v in start]
        f.grid_step[:] = [float(v) for v in step]
        f.grid_n[:] = [int(v) for v in npoints]
        if int(np.prod(list(npoints))) != len(s):
            raise ValueError("error reading file.  Check format")
        f.grid_data = s.ctypes.data_as(_lib._dp)
        f.rotations[:] = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        f.translation[:] = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        f._keep = s  # grid_data points into s: the field owns its samples
        self._push_b(f)
        self._traj_valid = False

    def _add_grid_exact(self, start, stop, step, npoints, samples, rotations, translation):
        s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
        f = _lib.om_field()
        f.type = _lib.OM_FIELD_GRID
        f.grid_exact = 1
        f.grid_start[:] = [float(v) for v in start]
        f.grid_stop[:] = [float(v) for v in stop]
        f.grid_step[:] = [float(v) for v in step]
        f.grid_n[:] = [int(v) for v in npoints]
        f.grid_data = s.ctypes.data_as(_lib._dp)
        f.rotations[:] = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        f.translation[:] = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        f._keep = s  # grid_data points into s: the field owns its samples
        self._push_b(f)
        self._traj_valid = False

    def add_bfield_file(self, ifile="", bifile="", iformat="", rotations=None, translation=None, scale=None,
                        frequency=0.0, frequency_phase=0.0, time_offset=0.0, name=""):
        """Field map from a file (src/OSCARSSR_Python.cc:526-640,
        src/TField3D_Grid.cc:36-75, 353-1694): 'OSCARS' (regular 1-, 2- or 3-D
        grid; scale = [x, y, z step factors, Bx, By, Bz factors]),
        'OSCARS1D <columns>' (an irregular 1-D table resampled at 10000 points
        per metre), 'SRW', 'SPECTRA', or a binary OSCARS grid given as bifile."""
        if len(ifile) != 0 and len(bifile) != 0:
            raise ValueError("cannot specify both 'ifile' and 'bifile'")
        if len(ifile) != 0 and len(iformat) == 0:
            raise ValueError("'iformat' is blank")
        rot = _v3(rotations, "rotations") if rotations else None
        tr = _v3(translation, "translation") if translation else None
        sc = [float(v) for v in scale] if scale else []
        self._check_name(name)
        n0 = len(self._fields)
        try:
            if len(bifile) != 0:
                start, stop, step, npts, data = read_binary_field_file(bifile, sc)
                self._add_grid_exact(start, stop, step, npts, data, rot, tr)
            else:
                self._read_bfield_file(ifile, iformat, rot, tr, sc)
        except (OSError, ValueError, IndexError) as ex:
            raise ValueError("Could not import magnetic field.  Check 'ifile' and 'iformat' are correct (%s)" % ex)
        if len(self._fields) > n0:
            self._set_time(self._fields[-1], frequency, frequency_phase, time_offset)
            _named(self._fields[-1], name)

    def _read_bfield_file(self, ifile, iformat, rotations, translation, sc):
        fmt = iformat.upper()
        if fmt == "OSCARS":
            start, step, npts, data = read_oscars_field_file(ifile)
            if not sc:
                self.add_bfield_grid(start, step, npts, data, rotations, translation)
                return
            g = field_grid_scaled(start, step, npts, data, sc)
            self._add_grid_exact(g[0], g[1], g[2], npts, g[3], rotations, translation)
        elif fmt[:8] == "OSCARS1D":
            axis, first, last, step, n, data = read_oscars1d_field_file(ifile, iformat[8:], sc)
            start, stop, steps, npts = [0.0] * 3, [0.0] * 3, [0.0] * 3, [1, 1, 1]
            start[axis], stop[axis], steps[axis], npts[axis] = first, last, step, n
            self._add_grid_exact(start, stop, steps, npts, data, rotations, translation)
        elif fmt in ("SRW", "SPECTRA"):
            start, stop, step, npts, data = (read_srw_field_file if fmt == "SRW" else read_spectra_field_file)(ifile)
            self._add_grid_exact(start, stop, step, npts, data, rotations, translation)
        elif fmt == "BINARY":
            start, stop, step, npts, data = read_binary_field_file(ifile, sc)
            self._add_grid_exact(start, stop, step, npts, data, rotations, translation)
        else:
            raise ValueError("incorrect format given")

    def add_bfield_interpolated(self, mapping, iformat, parameter, rotations=None, translation=None, scale=None,
                                frequency=0, frequency_phase=0, time_offset=0, name="", ofile=""):
        """A field map interpolated between maps taken at several values of a
        parameter (e.g. an undulator gap): per sample and component, the
        reference's cubic spline over the parameter values, evaluated at
        `parameter` (src/TField3D_Grid.cc:1694-1940).  OSCARS-format maps."""
        if not iformat:
            raise ValueError("'iformat' is blank")
        if iformat.upper() != "OSCARS":
            raise ValueError("interpolation of %s maps is not supported by this implementation" % iformat)
        self._check_name(name)
        pairs = []
        for m in mapping:
            if len(m) != 2:
                raise ValueError("Incorrect format in 'mapping'")
            pairs.append((float(m[0]), str(m[1])))
        pairs.sort(key=lambda a: a[0])
        maps = [read_oscars_field_file(f) for _, f in pairs]
        for m in maps[1:]:
            if m[0] != maps[0][0] or m[1] != maps[0][1] or m[2] != maps[0][2]:
                raise ValueError("Header value is incorrect format")
        start, step, npts, _ = maps[0]
        sc = [float(v) for v in scale] if scale else []
        samples = []
        for m in maps:  # field factors apply to each map as it is read
            f = np.array(m[3], dtype=np.float64).reshape(-1, 3).copy()
            for k in range(3):
                if len(sc) > 3 + k and sc[3 + k] != 1:
                    f[:, k] *= sc[3 + k]
            samples.append(f)
        data = spline_at([a for a, _ in pairs], np.stack(samples), float(parameter))
        g = field_grid_scaled(start, step, npts, data, sc[:3])
        self._add_grid_exact(g[0], g[1], g[2], npts, g[3], rotations, translation)
        self._set_time(self._fields[-1], frequency, frequency_phase, time_offset)
        _named(self._fields[-1], name)

    def clear_bfields(self):
        # src/OSCARSSR.cc:186-196: the current trajectory stays until a field is added
        self._fields = []
        self._stale_eb = None

    # ------------------------------------------------------------ beam / trajectory window
    def set_particle_beam(self, *args, **kw):
        """Clear the beams, add this one and make its ideal particle current
        (src/OSCARSSR_Python.cc:3561-3580); arguments as add_particle_beam."""
        self.clear_particle_beams()
        self.add_particle_beam(*args, **kw)

    def clear_particle_beams(self):
        """Remove every beam; the current particle stays (src/OSCARSSR.cc:540-547)."""
        self._beams, self._wsum, self._beam = [], [], None

    def _beam_index(self):
        # src/TParticleBeamContainer.cc:135-166: one uniform draw picks a beam
        # by weight when there is more than one
        if len(self._beams) == 1:
            return 0
        r = B.stream().random() * self._wsum[-1]
        for i, w in enumerate(self._wsum):
            if r < w:
                return i
        raise ValueError("did not find a beam for this weight")

    def _beam_by_name(self, name):
        """A beam by name, or one drawn by weight for "" (src/TParticleBeamContainer.cc:107-121)."""
        if not self._beams:
            raise ValueError("No particle beam defined")
        if name == "":
            return self._beams[self._beam_index()][1]
        for n, b, _ in self._beams:
            if n == name:
                return b
        raise ValueError("beam name not in map")

    def add_particle_beam(self, type="electron", name="", energy_GeV=-1, d0=None, x0=None, beam="",
                          sigma_energy_GeV=0, t0=0, current=0, weight=1, rotations=None, translation=None,
                          horizontal_direction=None, beta=None, alpha=None, gamma=None, emittance=None, eta=None,
                          lattice_reference=None, mass=0, charge=0, ctstartstop=None, distribution=""):
        """Add a beam; random particles come from the beams in proportion to
        their weights (src/OSCARSSR_Python.cc:3699-4070)."""
        pre = None
        if beam:
            pre = B.predefined_beam(beam)
            type, current, t0 = "electron", pre["current"], 0.0
            if energy_GeV >= 0:
                pre["energy_GeV"] = energy_GeV
        if energy_GeV == -1:
            energy_GeV = 0
        pos = _v3(x0, "x0") if x0 is not None else [0.0, 0.0, 0.0]
        if energy_GeV < 0:
            raise ValueError("We do not currently support negative energy beams.  Please change 'energy_GeV' to >= 0")
        direction = _v3(d0, "d0") if d0 is not None else [0.0, 0.0, 1.0]
        rot = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        tr = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        # the horizontal direction comes from the direction as given, before
        # the rotations
        hdir = _v3(horizontal_direction, "horizontal_direction") if horizontal_direction is not None else \
            [-v for v in G.orthogonal(direction)]
        hdir = G.unit(hdir)
        pos = G.rot_xyz(pos, rot)
        direction = G.rot_xyz(direction, rot)
        pos = [pos[k] + tr[k] for k in range(3)]
        if pre is not None:
            # a predefined beam keeps its own direction unless d0 is given
            direction = direction if d0 is not None else [0.0, 0.0, 1.0]
            energy_GeV = pre["energy_GeV"]
        elif type == "custom" and (mass == 0 or charge == 0):
            raise ValueError("'mass' or 'charge' is zero")
        p = G.ideal_particle(type, energy_GeV, pos, direction, t0, current, charge, mass)
        if sigma_energy_GeV < 0:
            raise ValueError("'sigma_energy_GeV' cannot be less than zero")
        two = lambda v, name: None if v is None else _v2(v, name)  # noqa: E731
        beam_p = dict(emittance=two(emittance, "emittance"), beta=two(beta, "beta"), alpha=two(alpha, "alpha"),
                      gamma=two(gamma, "gamma"), eta=two(eta, "gamma"),
                      lattice_reference=None if lattice_reference is None else _v3(lattice_reference,
                                                                                   "lattice_reference"))
        user_twiss = B.twiss_from(beam_p["beta"], beam_p["alpha"], beam_p["gamma"])  # the reference's range checks
        dist = B.distribution_of(emittance, "")
        if pre is not None:
            if beam_p["emittance"] is None:
                beam_p["emittance"] = pre["emittance"]
                dist = pre["distribution"]
            if not sigma_energy_GeV > 0:
                sigma_energy_GeV = pre["sigma_energy_GeV"]
            if user_twiss is None:
                beam_p.update(beta=pre["beta"], alpha=[0.0, 0.0], gamma=None)
        if distribution:
            dist = B.distribution_of(emittance, distribution)
        my = name if name != "" else "_beam%d" % len(self._beams)
        if any(n == my for n, _, _ in self._beams):
            raise ValueError("Error in predefined beam name / definition" if pre is not None else
                             "invalid argument in adding particle beam.  possibly 'name' already exists")
        b = dict(type=type, name=my, energy_GeV=energy_GeV, x0=pos, d0=direction, t0=float(t0),
                 current=float(current), sigma_energy_GeV=float(sigma_energy_GeV),
                 horizontal_unit=hdir, distribution=dist, q=p.q, m=p.m, **beam_p)
        self._beams.append((my, b, float(weight)))
        self._wsum.append(float(weight) + (self._wsum[-1] if self._wsum else 0.0))
        if len(self._beams) == 1:
            # the first beam: its ideal particle is current (set_particle_beam's
            # SetNewParticle("", "ideal") draws nothing with one beam)
            self._beam = b
            self._particle = p
        if ctstartstop is not None:
            if len(ctstartstop) != 2:
                raise ValueError("Incorrect format in 'ctstartstop'")
            self.set_ctstartstop(ctstartstop[0], ctstartstop[1])
        else:
            self.set_ctstartstop(-3, 3)
        self._traj_valid = False

    def set_new_particle(self, beam="", particle=""):
        """A new current particle from the named beam, or one drawn by weight:
        ideal, or random for anything else (src/OSCARSSR.cc:506-520)."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        b = self._beam_by_name(beam)
        if particle.lower() == "ideal":
            self._particle = G.ideal_particle(b["type"], b["energy_GeV"], b["x0"], b["d0"], b["t0"], b["current"],
                                              b["q"], b["m"])
        else:
            # src/TParticleBeam.cc:667-689: anything but "ideal" draws a random particle
            self._particle = B.sample_particles(b, 1)[0]
        self._traj_valid = False

    def set_particle(self, x0, beta0, t0=0.0):
        """Explicit initial conditions for the current particle (extension)."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        b = self._beam
        self._particle = G.Particle(b["q"], b["m"], x0, beta0, t0, b["current"])
        self._traj_valid = False

    def set_trajectory_calculation(self, method="RK4", precision=-1):
        """The trajectory integrator (src/OSCARSSR.cc:784-800).  RK4 is the one
        here.  The reference's RKAS stores the start time twice and its
        trajectory spline is then NaN (and it stops with "stepsize underflow"
        on ideal undulators), so it is refused rather than imitated."""
        m = method.upper()
        if m == "RK4":
            if precision != -1:
                raise ValueError("RK4 Trajectory does not use the precision parameter and you are attempting to "
                                 "set it")
            return
        if m == "RKAS":
            raise ValueError("RKAS trajectories are not supported by this implementation (use RK4)")
        raise ValueError("Method requested is invalid.  Try again")

    def set_ctstartstop(self, start, stop):
        # src/OSCARSSR.cc:637-647.  As in the reference, a trajectory already
        # computed or set stays in use: the new window applies to the next one
        # (a new particle or field, or calculate_trajectory).
        self._npoints = int(self._npoints_per_meter * (stop - start))
        self._ctstart = float(start)
        self._ctstop = float(stop)

    def set_npoints_trajectory(self, npoints):
        if npoints < 1:
            raise ValueError("npoints must be >= 1")
        self._npoints = int(npoints)

    def set_npoints_per_meter_trajectory(self, npoints):
        self._npoints_per_meter = int(npoints)
        if self._ctstart != self._ctstop:
            self.set_ctstartstop(self._ctstart, self._ctstop)

    def get_ctstart(self):
        return self._ctstart

    def get_ctstop(self):
        return self._ctstop

    def get_npoints_trajectory(self):
        return self._npoints

    def _om_particle(self, p, ctstart=None, ctstop=None, npoints=None):
        o = _lib.om_particle()
        o.x0[:] = p.x0
        o.beta0[:] = p.b0
        o.t0 = p.t0
        o.q = p.q
        o.m = p.m
        o.gamma = p.gamma
        o.qomg = p.qomg
        o.current = p.current
        o.ctstart = self._ctstart if ctstart is None else ctstart
        o.ctstop = self._ctstop if ctstop is None else ctstop
        o.npoints = self._npoints if npoints is None else npoints
        return o

    def _ensure_trajectory(self):
        if self._particle is None:
            raise ValueError("No particle beam defined")
        if self._traj_valid:
            return
        if self._ctstart > self._particle.t0 or self._ctstop < self._particle.t0:
            raise ValueError("start time is greater than T0 or stop time is less than T0")
        if self._ctstart >= self._ctstop:
            raise ValueError("start time is greater than stop time.  check that set_ctstartstop is set")
        self._device().trajectories(self._fields, [self._om_particle(self._particle)], efields=self._efields_for_derivs())
        self._traj_valid = True

    def calculate_trajectory(self, ofile="", bofile="", oformat=""):
        """Integrate the current particle's trajectory (always anew, replacing
        one given by set_trajectory) and return it as get_trajectory does;
        ofile / bofile write it as text / binary in oformat
        (src/OSCARSSR_Python.cc:5181-5246)."""
        self._traj_valid = False
        self._ensure_trajectory()
        if ofile or bofile:
            t, p, _ = self._device().trajectory(0)
            if ofile:
                IO.write_trajectory_text(ofile, t, p, oformat)
            if bofile:
                IO.write_trajectory_binary(bofile, t, p, oformat)
        return self.get_trajectory()

    def get_trajectory(self):
        """[[t, [x, y, z], [bx, by, bz], [ax, ay, az]], ...] of the current
        trajectory; [] when none has been computed or set."""
        if not self._traj_valid:
            return []
        t, p, _ = self._device().trajectory(0)
        return [[float(t[i]), p[i, 0:3].tolist(), p[i, 3:6].tolist(), p[i, 6:9].tolist()] for i in range(len(t))]

    def set_trajectory(self, beam="", trajectory=None, ifile=None, bifile=None, iformat=""):
        """Use a given trajectory instead of integrating one: a list of
        [t, [x, y, z]] / [t, [x, y, z], [bx, by, bz]] / [t, x, b, [ax, ay, az]]
        points (t in s, a/c in 1/s; beta and a/c missing from the first point
        are rebuilt from spline slopes), or a text / binary file in the
        reference's trajectory formats (src/OSCARSSR_Python.cc:5374-5566).
        Times must increase strictly."""
        if self._beam is None:
            raise ValueError("Must previously have defined a particle beam to associate this trajectory with.  try: "
                             "add_particle_beam() or set_particle_beam() first")
        if ifile is not None and bifile is not None:
            raise ValueError("Cannot specify both text and binary input files")
        if trajectory is not None and (ifile is not None or bifile is not None):
            raise ValueError("Cannot specify both an input file and a trajectory list")
        if beam and beam != self._beam.get("name", ""):
            raise ValueError("no beam by that name found")
        if beam:
            self.set_new_particle(particle="ideal")
        if ifile is not None or bifile is not None:
            if ifile:
                t, p9, has_b, has_bp = IO.read_trajectory_text(ifile, iformat)
            elif bifile:
                t, p9, has_b, has_bp = IO.read_trajectory_binary(bifile, iformat)
            else:
                return
        elif trajectory is not None:
            if len(trajectory) < 1:
                raise ValueError("Input trajectory is not the correct format")
            has_b = len(trajectory[0]) >= 3
            has_bp = len(trajectory[0]) >= 4
            t = np.zeros(len(trajectory))
            p9 = np.zeros((len(trajectory), 9))
            for i, pt in enumerate(trajectory):
                try:
                    if len(pt) <= 1:
                        raise ValueError
                    t[i] = float(pt[0])
                    p9[i, 0:3] = _v3(pt[1], "trajectory")
                    if has_b:
                        p9[i, 3:6] = _v3(pt[2], "trajectory")
                    if has_bp:
                        p9[i, 6:9] = _v3(pt[3], "trajectory")
                except (ValueError, TypeError, IndexError):
                    raise ValueError("Incorrect format in 'trajectory' entry number %d" % i)
        else:
            raise ValueError("Need to specify where to get the trajectory from, ifile, trajectory, etc..")
        construct = (0 if has_b else 1) | (0 if has_bp else 2)
        self._device().trajectory_set(self._om_particle(self._particle), t, p9, construct)
        self._traj_valid = True

    def get_trajectory_arrays(self):
        """(t, knots[n, 9], spline second derivatives[n, 9]) as numpy arrays (extension)."""
        self._ensure_trajectory()
        return self._device().trajectory(0)

    # ------------------------------------------------------------ argument helpers
    @staticmethod
    def _quantity(q, pd=False):
        Q = q.upper()
        if Q in (("POWER DENSITY", "POWERDENSITY") if pd else ("FLUX",)):
            return 0
        if Q == "PRECISION":
            return 1
        if Q == "LEVEL":
            return 2
        if pd:
            raise ValueError("'quantity' must be: 'power density', 'precision', 'level', or blank")
        raise ValueError("'quantity' must be: 'flux', 'precision', 'level', or blank")

    @staticmethod
    def _check_common(nparticles, nthreads, gpu):
        if nparticles < 0:
            raise ValueError("'nparticles' must be >= 1 (sort of)")
        if nthreads < 0:
            raise ValueError("'nthreads' must be > 0")
        if gpu not in (0, 1, -1):
            raise ValueError("'gpu' must be 0 or 1")
        if nthreads > 0 and gpu == 1:
            raise ValueError("gpu is 1 and nthreads > 0.  Both are not currently allowed.")

    @staticmethod
    def _polarization(polarization, angle, hdir, pdir):
        if polarization == "all" and angle != _NOANGLE:
            raise ValueError("cannot specify 'all' and 'angle'.  This is to save you")
        pol = "all"
        if polarization:
            pol = polarization
        elif angle != _NOANGLE:
            pol = "linear"
        return G.polarization_vector(pol, angle if angle != _NOANGLE else 0.0, hdir, pdir)

    def set_seed(self, seed):
        """Restart the global random stream (include/OSCARSPY.h:85-87): the
        same seed gives the reference's particles, draw for draw."""
        B.set_seed(seed)

    def _particles_for(self, nparticles):
        """Particles to sum: the current one (nparticles=0) or nparticles drawn
        from the global stream in the reference's order (one new particle per
        iteration of src/OSCARSSR.cc:1782-1790)."""
        if nparticles == 0:
            return None
        if len(self._beams) == 1:
            return B.sample_particles(self._beam, nparticles)
        return [B.sample_particles(self._beam_by_name(""), 1)[0] for _ in range(nparticles)]

    def get_emittance(self, name=""):
        """[horizontal, vertical] emittance of a beam (src/OSCARSSR_Python.cc:4396-4410)."""
        e = self._beam_by_name(name).get("emittance")
        return [float(e[0]), float(e[1])] if e is not None else [0.0, 0.0]

    def get_beam_horizontal_direction(self, name=""):
        """Unit horizontal direction of a beam (src/OSCARSSR_Python.cc:4745-4760)."""
        return list(B.beam_frame(self._beam_by_name(name))[0])

    def get_beam_vertical_direction(self, name=""):
        """Unit vertical direction of a beam, -unit(horizontal x u0)."""
        return list(B.beam_frame(self._beam_by_name(name))[1])

    def set_twiss_parameters(self, beam="", beta=None, alpha=None, gamma=None, lattice_reference=None):
        """Twiss parameters of a beam, any two of beta, alpha, gamma or all three,
        or beta alone with alpha 0 (src/OSCARSSR_Python.cc:4269-4360)."""
        two = lambda v, n: None if v is None else _v2(v, n)  # noqa: E731
        tw = dict(beta=two(beta, "beta"), alpha=two(alpha, "alpha"), gamma=two(gamma, "gamma"))
        ref = None if lattice_reference is None else _v3(lattice_reference, "lattice_reference")
        if B.twiss_from(tw["beta"], tw["alpha"], tw["gamma"]) is None:
            return
        # the beam is looked up inside a matching branch only, as the
        # reference does, so a call without effect draws nothing from the stream
        b = self._beam_by_name(beam)
        b.update(tw, lattice_reference=ref)
        self._traj_valid = False

    def set_particle_beam_size(self, sigma=None, sigmap=None, beam="", lattice_reference=None, eta=None):
        """Checks its arguments and changes nothing, as the reference's binding
        does (src/OSCARSSR_Python.cc:4105-4160)."""
        for v, n in ((sigma, "sigma"), (sigmap, "sigmap")):
            if v is not None and len(v) != 0:
                _v2(v, n)
        if lattice_reference is not None:
            _v3(lattice_reference, "lattice_reference")

    def _batched(self, nparticles, obs, energies, pol, precision, max_level, max_level_extended, mode):
        """nparticles particles in one persistent launch (tiled / mixed kernels)."""
        from . import multiparticle as M
        plist = self._particles_for(nparticles)
        dev = self._device()
        windows = [(self._ctstart, self._ctstop, self._npoints)] * len(plist)
        w = [1.0 / len(plist)] * len(plist)
        bf = M.BatchedFlux(dev, self._fields, plist, windows, [(np.asarray(obs), np.asarray(energies))],
                           [0] * len(plist), w, mode=mode, polarization=pol, precision=precision,
                           max_level=max_level, max_level_extended=max_level_extended, efields=self._efields_for_derivs())
        self._traj_valid = False
        outs, st = bf.run(recompute_trajectories=False)
        self.last_stats = st
        self._last_conv = np.array([st["n_unconverged"] == 0])
        return outs[0].reshape(-1)

    def _run_particles(self, nparticles, fn):
        """Call fn(device, particle_index, weight, accumulate) for every particle."""
        plist = self._particles_for(nparticles)
        if plist is None:
            self._ensure_trajectory()
            return fn(self._device(), 0, 1.0, False)
        dev = self._device()
        dev.trajectories(self._fields, [self._om_particle(p) for p in plist], efields=self._efields_for_derivs())
        self._traj_valid = False
        w = 1.0 / float(len(plist))
        out = None
        for i in range(len(plist)):
            out = fn(dev, i, w, i > 0, out)
        return out

    # ------------------------------------------------------------ spectrum
    def calculate_spectrum(self, obs=None, npoints=0, energy_range_eV=None, energy_points_eV=None, points_eV=None,
                           polarization="", angle=_NOANGLE, horizontal_direction=None, propogation_direction=None,
                           precision=0.01, max_level=-2, max_level_extended=0, nparticles=0, nthreads=0, gpu=-1,
                           ngpu=None, quantity="flux", ofile="", bofile="", mode="tiled"):
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        if points_eV is not None and energy_points_eV is None:
            energy_points_eV = points_eV
        if energy_points_eV is not None:
            if len(energy_points_eV) == 0:
                raise ValueError("empty list of energy points")
            e = [float(v) for v in energy_points_eV]
        else:
            e0 = e1 = 0.0
            if energy_range_eV is not None:
                if len(energy_range_eV) != 2:
                    raise ValueError("'energy_range_eV' must be a list of length 2")
                e0, e1 = float(energy_range_eV[0]), float(energy_range_eV[1])
            if npoints < 1:
                npoints = abs(int(e1 - e0) + 1) if math.fabs(e1 - e0) + 1 > 100 else 100
            e = G.energies(npoints, e0, e1)
        if obs is None:
            raise ValueError("Incorrect format in 'obs'")
        o = _v3(obs, "obs")
        hdir = _v3(horizontal_direction, "horizontal_direction") if horizontal_direction else [1.0, 0.0, 0.0]
        pdir = _v3(propogation_direction, "propogation_direction") if propogation_direction else [0.0, 0.0, 1.0]
        q = self._quantity(quantity)
        pol = self._polarization(polarization, angle, hdir, pdir)
        e_arr = np.array(e)

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, _, st = dev.spectrum(ip, o, e_arr, pol, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        if nparticles > 0 and mode in ("tiled", "mixed") and q == 0:
            v = self._batched(nparticles, [o], e_arr, pol, precision, max_level, max_level_extended, mode)
        else:
            v = self._run_particles(nparticles, run)
        if not self._last_conv.all():
            _warn("Not all points converged to desired precision.  Can try increasing 'max_level_extended'\n")
        # src/OSCARSSR_Python.cc:5918-5924
        if ofile != "":
            IO.write_spectrum_text(ofile, e, v)
        if bofile != "":
            IO.write_spectrum_binary(bofile, e, v)
        return [[e[i], float(v[i])] for i in range(len(e))]

    # ------------------------------------------------------------ flux
    def _rect(self, npoints, plane, width, rotations, translation, normal, dim, x0x1x2):
        if dim not in (2, 3):
            raise ValueError("'dim' must be 2 or 3")
        if npoints is None or len(npoints) != 2:
            raise ValueError("'npoints' must be [int, int]")
        n1, n2 = int(npoints[0]), int(npoints[1])
        if n1 <= 0 or n2 <= 0:
            raise ValueError("an entry in 'npoints' is <= 0")
        if abs(normal) > 1:
            raise ValueError("'normal' must be -1, 0, or 1")
        rot = _v3(rotations, "rotations") if rotations else [0.0, 0.0, 0.0]
        tr = _v3(translation, "translation") if translation else [0.0, 0.0, 0.0]
        if x0x1x2 is not None:
            # src/OSCARSSR_Python.cc:6633-6655: three corners, rotated then translated
            if len(x0x1x2) != 3:
                raise ValueError("'x0x1x2' must have 3 XYZ points defined correctly")
            X = [_v3(x, "x0x1x2") for x in x0x1x2]
            if n1 < 2 or n2 < 2:
                raise ValueError("'npoints' must be >= 2 in each direction for 'x0x1x2'")
            return G.parallelogram(n1, n2, X[0], X[1], X[2], rot, tr, normal, dim)
        if width is None or len(width) != 2 or not (width[0] > 0 and width[1] > 0):
            raise ValueError("must specify 'width' (positive only)")
        return G.rectangle(plane, n1, n2, float(width[0]), float(width[1]), rot, tr, normal, dim)

    def calculate_flux_rectangle(self, energy_eV, npoints=None, plane="XY", normal=0, dim=2, width=None,
                                 rotations=None, translation=None, x0x1x2=None, polarization="", angle=_NOANGLE,
                                 horizontal_direction=None, propogation_direction=None, nparticles=0, nthreads=0,
                                 gpu=-1, ngpu=None, precision=0.01, max_level=-2, max_level_extended=0,
                                 quantity="flux", ofile="", bofile="", mode="plain"):
        if self._beam is None:
            raise ValueError("No particle beam defined")
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        obs, nv, cont = self._rect(npoints, plane, width, rotations, translation, normal, dim, x0x1x2)
        hdir = _v3(horizontal_direction, "translation") if horizontal_direction else [1.0, 0.0, 0.0]
        pdir = _v3(propogation_direction, "translation") if propogation_direction else [0.0, 0.0, 1.0]
        q = self._quantity(quantity)
        pol = self._polarization(polarization, angle, hdir, pdir)
        obs_a = np.array(obs)

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, _, st = dev.flux(ip, obs_a, [float(energy_eV)], pol, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        v = self._run_particles(nparticles, run)
        if not self._last_conv.all():
            _warn("Not all points converged to desired precision.  Can try increasing 'max_level_extended'\n")
        # src/OSCARSSR_Python.cc:8310-8318: the container's points (local x1, x2 for dim 2)
        if ofile != "":
            IO.write_points_text(ofile, cont, v, dim)
        if bofile != "":
            IO.write_points_binary(bofile, cont, v, dim)
        return [[list(cont[i]), float(v[i])] for i in range(len(cont))]

    def calculate_flux_rectangle_energies(self, energies_eV, npoints=None, plane="XY", normal=0, dim=2, width=None,
                                          rotations=None, translation=None, polarization="", angle=_NOANGLE,
                                          horizontal_direction=None, propogation_direction=None, nparticles=0,
                                          precision=0.01, max_level=-2, max_level_extended=0, quantity="flux",
                                          mode="tiled"):
        """Flux on a rectangle at many energies in one call (extension).
        Returns (coords[n_obs, 3], flux[n_obs, n_energies])."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        obs, nv, cont = self._rect(npoints, plane, width, rotations, translation, normal, dim, None)
        hdir = _v3(horizontal_direction, "translation") if horizontal_direction else [1.0, 0.0, 0.0]
        pdir = _v3(propogation_direction, "translation") if propogation_direction else [0.0, 0.0, 1.0]
        q = self._quantity(quantity)
        pol = self._polarization(polarization, angle, hdir, pdir)
        obs_a = np.array(obs)
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, _, st = dev.flux(ip, obs_a, e, pol, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        if nparticles > 0 and mode in ("tiled", "mixed") and q == 0:
            v = self._batched(nparticles, obs_a, e, pol, precision, max_level, max_level_extended, mode)
        else:
            v = self._run_particles(nparticles, run)
        return np.array(cont), v.reshape(len(obs), len(e))

    # ------------------------------------------------------------ power density
    def calculate_power_density_rectangle(self, npoints=None, plane="XY", normal=0, width=None, rotations=None,
                                          translation=None, x0x1x2=None, nparticles=0, gpu=-1, ngpu=None,
                                          nthreads=0, precision=0.01, max_level=-2, max_level_extended=0, dim=2,
                                          quantity="power density", ofile="", bofile="", mode="plain"):
        """Power density [W/mm^2] on a rectangle (src/OSCARSSR_Python.cc:6471-6797).
        Extension: mode "mixed" evaluates each term with reciprocal-based
        division and square root (within about 1e-6 of the FP64 terms) and
        recomputes in FP64 every point whose stop decision lies within that
        error's bound, so the levels are the FP64 path's."""
        _check_pd_mode(mode)
        if self._beam is None:
            raise ValueError("No particle beam defined")
        if npoints is None:
            npoints = [51, 51]
        if width is None:
            width = [0.010, 0.010]
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        obs, nv, cont = self._rect(npoints, plane, width, rotations, translation, normal, dim, x0x1x2)
        q = self._quantity(quantity, pd=True)
        directional = normal != 0
        obs_a = np.array(obs)

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, st = dev.power_density(ip, obs_a, nv, directional, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        v = self._run_particles(nparticles, run)
        if not self._last_conv.all():
            _warn("Not all points converged to desired precision.  Can try increasing 'max_level_extended'\n")
        # src/OSCARSSR_Python.cc:6755-6763
        if ofile != "":
            IO.write_points_text(ofile, cont, v, dim)
        if bofile != "":
            IO.write_points_binary(bofile, cont, v, dim)
        return [[list(cont[i]), float(v[i])] for i in range(len(cont))]

    def _pd_points(self, obs, normals, directional, nparticles, precision, max_level, max_level_extended, q,
                   mode="plain"):
        _check_pd_mode(mode)
        obs_a = np.asarray(obs, dtype=np.float64).reshape(-1, 3)

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, st = dev.power_density_surface(ip, obs_a, normals, directional, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        return self._run_particles(nparticles, run)

    def calculate_power_density(self, points, normal=0, rotations=None, translation=None, nparticles=0, gpu=-1,
                                ngpu=None, nthreads=0, precision=0.01, max_level=-2, max_level_extended=0,
                                quantity="power density", ofile="", bofile="", mode="plain"):
        """Power density [W/mm^2] at [[x, y, z], [nx, ny, nz]] points, rotated
        then translated (normals rotated only); returns [[[x, y, z], pd], ...]
        in absolute coordinates (src/OSCARSSR_Python.cc:6105-6380).  Extension:
        mode as in calculate_power_density_rectangle."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        obs, nrm = surface_points(points, normal, rotations, translation)
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        q = self._quantity(quantity, pd=True)
        v = self._pd_points(obs, np.array(nrm).reshape(-1, 3), normal != 0, nparticles, precision, max_level,
                            max_level_extended, q, mode)
        if not self._last_conv.all():
            _warn("Not all points converged to desired precision.  Can try increasing 'max_level_extended'\n")
        if ofile != "":
            IO.write_points_text(ofile, obs, v, 3)
        if bofile != "":
            IO.write_points_binary(bofile, obs, v, 3)
        return [[list(obs[i]), float(v[i])] for i in range(len(obs))]

    def calculate_power_density_line(self, x1, x2, npoints=0, ofile="", bofile="", normal=0, nparticles=0, gpu=-1,
                                     nthreads=0, precision=0.01, max_level=-2, max_level_extended=0, dim=1,
                                     mode="plain"):
        """Power density at npoints evenly spaced from x1 to x2 (a surface
        without normals); dim 1 returns the signed distance from the middle,
        dim 3 the points (src/OSCARSSR_Python.cc:7355-7560).  Extension: mode
        as in calculate_power_density_rectangle."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        if dim not in (1, 3):
            raise ValueError("'dim' must be 1 or 3")
        if npoints < 2:
            raise ValueError("'npoints' must be >= 2")
        a = _v3(x1, "x1") if len(x1) != 0 else [0.0, 0.0, 0.0]
        b = _v3(x2, "x2") if len(x2) != 0 else [0.0, 0.0, 0.0]
        step = [(b[k] - a[k]) / (npoints - 1) for k in range(3)]
        obs = [[a[k] + i * step[k] for k in range(3)] for i in range(npoints)]
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        v = self._pd_points(obs, None, normal != 0, nparticles, precision, max_level, max_level_extended, 0, mode)
        if dim == 3:
            pos = obs
        else:
            # src/OSCARSSR.cc:2712-2722: signed distance from the middle of the line
            c = [a[k] + (obs[-1][k] - a[k]) / 2 for k in range(3)]
            pos = []
            for i in range(npoints):
                d = [obs[i][k] - c[k] for k in range(3)]
                pos.append([(-1. if i < npoints // 2 else 1.) * math.sqrt(G.mag2(d)), 0.0, 0.0])
        if ofile != "":
            IO.write_points_text(ofile, pos, v, dim)
        if bofile != "":
            IO.write_points_binary(bofile, pos, v, dim)
        return [[list(pos[i]), float(v[i])] for i in range(npoints)]

    # ------------------------------------------------------------ STL surfaces
    def add_stl(self, ifile, rotations=None, translation=None, scale=1.0, name=""):
        """Add the triangles of a binary STL file, scaled then translated
        (src/OSCARSSR_Python.cc:6852-6927, src/TTriangle3DContainer.cc:110-228).
        As in the reference, ``rotations`` is checked but not applied."""
        if rotations is not None:
            _v3(rotations, "rotations")
        tr = _v3(translation, "translation") if translation is not None else [0.0, 0.0, 0.0]
        my = name if name != "" else "_stl" + str(len(self._stl))
        if any(n == my for n, _, _ in self._stl):
            raise ValueError("stl with this name already exists")
        self._stl.append((my, IO.read_stl(ifile, float(scale), tr), float(scale)))

    def clear_stl(self):
        self._stl = []

    def remove_stl(self, name):
        self._stl = [s for s in self._stl if s[0] != name]

    def calculate_power_density_stl(self, farfield_origin=None, ofile="", bofile="", stlofile="", nparticles=0,
                                    gpu=-1, ngpu=None, nthreads=0, precision=0.01, max_level=-2,
                                    max_level_extended=0, quantity="power density", mode="tiled"):
        """Power density [W/mm^2] at the centre of every STL triangle, with the
        samples whose path to the centre meets another triangle left out;
        returns [[[A, B, C], pd], ...] (src/OSCARSSR_Python.cc:6993-7213).
        ``farfield_origin`` is checked and unused, and ``ofile`` / ``bofile``
        write nothing, as in the reference; ``stlofile`` writes the triangles.
        Extension: mode "plain" tests every triangle for every ray, as the
        reference does; any other mode (default) walks a bounding-volume
        hierarchy that finds the same hits, so the results are the same."""
        _check_pd_mode(mode)
        if self._beam is None:
            raise ValueError("No particle beam defined")
        if farfield_origin is not None:
            _v3(farfield_origin, "farfield_origin")
        self._check_common(nparticles, nthreads, 0 if gpu == 0 else gpu)
        q = self._quantity(quantity, pd=True)
        if not self._stl:
            return []
        tri = np.concatenate([t for _, t, _ in self._stl])

        def run(dev, ip, w, acc, out=None):
            conv = dev.conv(precision, max_level, max_level_extended, q, mode, w, acc)
            v, lev, cv, st = dev.power_density_stl(ip, tri, conv, out=out)
            self.last_stats = st
            self._last_levels = lev
            self._last_conv = cv
            return v

        v = self._run_particles(nparticles, run)
        if stlofile != "":
            IO.write_stl(stlofile, tri, np.concatenate([[s] * len(t) for _, t, s in self._stl]))
        return [[[tri[i, 0:3].tolist(), tri[i, 3:6].tolist(), tri[i, 6:9].tolist()], float(v[i])]
                for i in range(len(tri))]

    def calculate_efield_vs_time(self, obs, ofile=""):
        """Electric field in the time domain at obs: [[t, [Ex, Ey, Ez]], ...],
        one entry per trajectory knot at t = i dt + |R| / c (uneven in the lab
        frame) (src/OSCARSSR_Python.cc:9048-9110).  ofile writes 'Ex Ey Ez t'
        lines, the layout the reference's container writer gives this data."""
        if self._beam is None:
            raise ValueError("No particle beam defined")
        o = _v3(obs, "obs")
        self._ensure_trajectory()
        t, e = self._device().efield_time(0, o)
        if ofile != "":
            IO.write_points_text(ofile, e, t, 3)
        return [[float(t[i]), e[i].tolist()] for i in range(len(t))]

    def calculate_total_power(self, precision=0.01, max_level=-2, max_level_extended=0, quantity="power"):
        """Total radiated power [W] of the current particle, or the precision or
        level reached (src/OSCARSSR_Python.cc:5974-6060)."""
        Q = quantity.upper()
        if Q not in ("POWER", "PRECISION", "LEVEL"):
            raise ValueError("'quantity' must be: 'power', 'precision', 'level', or blank")
        if self._beam is None:
            raise ValueError("No particle beam defined")
        self._ensure_trajectory()
        dev = self._device()
        pw, pr, lv, st = dev.total_power(0, dev.conv(precision, max_level, max_level_extended, 0, "plain"))
        self.last_stats = st
        if lv == -1:
            raise ValueError("total power calculation did not converge.  Try increasing max_level or decreasing "
                             "precision")
        return {"POWER": pw, "PRECISION": pr, "LEVEL": float(lv)}[Q]

    # ------------------------------------------------------------ field values
    def _field_at(self, fields, x, t, name):
        try:
            X = _v3(x, "x")
        except ValueError:
            raise ValueError("Incorrect format in input")
        sel = [f for f in fields if name == "" or getattr(f, "name", "") == name]
        return self._device().field_points(sel, [X], float(t))[0].tolist()

    def get_bfield(self, x, t=0.0, name=""):
        """Summed magnetic field [T] at x and time t [s], or of the fields named
        `name` only (src/OSCARSSR_Python.cc:1927-1975, src/TFieldContainer.cc:85-107)."""
        return self._field_at(self._fields, x, t, name)

    def get_efield(self, x, t=0.0, name=""):
        """Summed electric field [V/m] at x and time t [s], as get_bfield."""
        return self._field_at(self._efields, x, t, name)

    # ------------------------------------------------------------ accumulated results
    @staticmethod
    def _accumulate(store, new, weight):
        # src/OSCARSSR.cc:2465-2530: the first call sets the points, later
        # calls add weight * value to the values point by point
        if len(store) == 0:
            return [[p, v * weight] for p, v in new]
        if len(store) != len(new):
            raise ValueError("spectra dimensions do not match")
        return [[p, v + w * weight] for (p, v), (_, w) in zip(store, new)]

    def add_to_spectrum(self, spectrum, weight=1.0):
        """Add weight * spectrum ([[E, V], ...]) to the stored spectrum
        (src/OSCARSSR_Python.cc:8623-8656)."""
        if len(spectrum) < 1:
            raise ValueError("No points in spectrum.")
        new = []
        for p in spectrum:
            if len(p) != 2:
                raise ValueError("GetSpectrumFromList reporting not 2 points")
            new.append((float(p[0]), float(p[1])))
        self._sum_spectrum = self._accumulate(self._sum_spectrum, new, float(weight))

    def get_spectrum(self):
        """The stored spectrum [[E, V], ...] (src/OSCARSSR_Python.cc:8675-8680)."""
        return [[e, v] for e, v in self._sum_spectrum]

    def write_spectrum(self, ofile="", bofile=""):
        """Write the stored spectrum as text and/or binary
        (src/OSCARSSR_Python.cc:8381-8420)."""
        e = [p for p, _ in self._sum_spectrum]
        v = [w for _, w in self._sum_spectrum]
        if ofile != "":
            IO.write_spectrum_text(ofile, e, v)
        if bofile != "":
            IO.write_spectrum_binary(bofile, e, v)

    @staticmethod
    def _points_list(points, what):
        if len(points) < 1:
            raise ValueError("No points in %s." % what)
        new = []
        for p in points:
            if len(p) != 2:
                raise ValueError("GetT3DScalarContainerFromList reporting not 2 points")
            if len(p[0]) != 3:
                raise ValueError("number of elements not 3")
            new.append(([float(c) for c in p[0]], float(p[1])))
        return new

    def add_to_flux(self, flux, weight=1.0):
        """Add weight * flux ([[[x, y, z], V], ...]) to the stored flux
        (src/OSCARSSR_Python.cc:8877-8910)."""
        self._sum_flux = self._accumulate(self._sum_flux, self._points_list(flux, "flux"), float(weight))

    def get_flux(self):
        """The stored flux [[[x, y, z], V], ...] (src/OSCARSSR_Python.cc:8929-8934)."""
        return [[list(p), v] for p, v in self._sum_flux]

    def add_to_power_density(self, power_density, weight=1.0):
        """Add weight * power density ([[[x, y, z], V], ...]) to the stored one
        (src/OSCARSSR_Python.cc:8957-8990; its empty-input message says flux)."""
        self._sum_pd = self._accumulate(self._sum_pd, self._points_list(power_density, "flux"), float(weight))

    def get_power_density(self):
        """The stored power density (src/OSCARSSR_Python.cc:9009-9014)."""
        return [[list(p), v] for p, v in self._sum_pd]

    # ------------------------------------------------------------ averaging of saved results
    @staticmethod
    def average_spectra(ifiles=None, bifiles=None, spectra=None, ofile="", bofile="", weights=None):
        """Average spectra from text files, binary files or [[E, V], ...] lists
        (src/OSCARSSR_Python.cc:8460-8598)."""
        return IO.average_spectra(ifiles, bifiles, spectra, ofile, bofile, weights)

    @staticmethod
    def average_flux(ifiles=None, bifiles=None, ofile="", bofile="", dim=2):
        """Average flux files of the same points (src/OSCARSSR_Python.cc:8694-8790)."""
        return IO.average_points(ifiles, bifiles, ofile, bofile, dim)

    @staticmethod
    def average_power_density(ifiles=None, bifiles=None, ofile="", bofile="", dim=2):
        """Average power-density files of the same points (src/OSCARSSR_Python.cc:8694-8790)."""
        return IO.average_points(ifiles, bifiles, ofile, bofile, dim)


def spline_at(X, Y, x):
    """include/TOMATH.h:40-203 (TSpline1D3) through the points (X[i], Y[i]),
    Y[i] an array of any shape, evaluated at x; the operations run element by
    element in the reference's order."""
    X = [float(v) for v in X]
    Y = np.asarray(Y, dtype=np.float64)
    N = len(X)
    if N < 3:
        raise ValueError("TSpline1D3 does not see enough points for interpolation.  Please use >= 3 points")
    ypp = np.empty_like(Y)
    u = np.empty_like(Y)
    dydx0 = (Y[1] - Y[0]) / (X[1] - X[0])
    ypp[0] = -0.5
    u[0] = (3. / (X[1] - X[0])) * ((Y[1] - Y[0]) / (X[1] - X[0]) - dydx0)
    for i in range(1, N - 1):
        sig = (X[i] - X[i - 1]) / (X[i + 1] - X[i - 1])
        p = sig * ypp[i - 1] + 2.0
        ypp[i] = (sig - 1.) / p
        ui = (Y[i + 1] - Y[i]) / (X[i + 1] - X[i]) - (Y[i] - Y[i - 1]) / (X[i] - X[i - 1])
        u[i] = (6. * ui / (X[i + 1] - X[i - 1]) - sig * u[i - 1]) / p
    qn = 0.5
    dydxN = (Y[N - 1] - Y[N - 2]) / (X[N - 1] - X[N - 2])
    un = (3. / (X[N - 1] - X[N - 2])) * (dydxN - (Y[N - 1] - Y[N - 2]) / (X[N - 1] - X[N - 2]))
    ypp[N - 1] = (un - qn * u[N - 2]) / (qn * ypp[N - 2] + 1.)
    for k in range(N - 2, -1, -1):
        ypp[k] = ypp[k] * ypp[k + 1] + u[k]
    klo, khi = 0, N - 1
    while khi - klo > 1:
        k = (khi + klo) >> 1
        if X[k] > x:
            khi = k
        else:
            klo = k
    h = X[khi] - X[klo]
    if abs(h) < 1e-50:
        raise ValueError("TSpline1D3 sees that the stepsize h is zero")
    a = (X[khi] - x) / h
    b = (x - X[klo]) / h
    return a * Y[klo] + b * Y[khi] + ((a * a * a - a) * ypp[klo] + (b * b * b - b) * ypp[khi]) * (h * h) / 6.


def field_grid_scaled(start, step, npts, data, scaling):
    """Start, stop, step and samples of an OSCARS grid read with scaling
    factors (src/TField3D_Grid.cc:405-451, 520-530): steps scaled, the start
    re-centred about the unscaled middle, the field components scaled."""
    st, sp, sc = [], [], []
    for a in range(3):
        s_in, d_in, n = float(start[a]), float(step[a]), int(npts[a])
        d = d_in * scaling[a] if len(scaling) > a else d_in
        middle = s_in + d_in * (n - 1) / 2.
        s0 = middle - d * (n - 1) / 2.
        st.append(s0)
        sp.append(s0 + (n - 1) * d)
        sc.append(d)
    f = np.array(data, dtype=np.float64).reshape(-1, 3).copy()
    for k in range(3):
        if len(scaling) > 3 + k and scaling[3 + k] != 1:
            f[:, k] *= scaling[3 + k]
    return st, sp, sc, f


def read_oscars1d_field_file(path, columns, scaling=(), comment="#"):
    """An irregular 1-D field table, 'OSCARS1D' format with its column names
    (X, Y or Z and any of Bx By Bz; Ex/Fx... are the same columns), resampled
    on 10000 points per metre by linear interpolation as
    src/TField3D_Grid.cc:572-873 does.  Returns (axis, first, last, step, n,
    samples[n, 3])."""
    order, axis, has = [], None, [False, False, False]
    for c in columns.split():
        if len(order) > 3:
            raise ValueError("spatial or B-field dimensions are too large(index>3)")
        if c in ("X", "Y", "Z"):
            axis = "XYZ".index(c)
            order.append(0)
        elif c in ("Bx", "Ex", "Fx", "By", "Ey", "Fy", "Bz", "Ez", "Fz"):
            k = "xyz".index(c[1])
            has[k] = True
            order.append(1 + k)
        else:
            raise ValueError("only excepts X Y Z Bx By Bz")
    if sum(1 for o in order if o == 0) != 1:
        raise ValueError("spatial or B-field dimensions are too large")
    try:
        lines = open(path).read().split("\n")
    except OSError:
        raise ValueError("cannot open file for reading")
    rows = []
    for line in lines[1:]:
        t = line.strip(" \t")
        if t == "" or t[0] == comment:
            continue
        toks = line.replace(",", " ").split()
        v = [0.0] * 4
        for i, o in enumerate(order):
            try:
                v[o] = float(toks[i])
            except (IndexError, ValueError):
                raise ValueError("something is incorrect with data format or iformat string")
        for i in range(min(len(scaling), 4)):
            if i < len(order):
                v[order[i]] *= scaling[i]
        rows.append(v)
    rows.sort(key=lambda r: r[0])
    if len(rows) < 2:
        raise ValueError("not enough data points")
    first, last = rows[0][0], rows[-1][0]
    n = int((last - first) * 10000)
    step = (last - first) / float(n - 1)
    out = np.zeros((n, 3))
    min_bin = after = before = 0
    for i in range(n):
        z = i * step + first
        for j in range(min_bin + 1, len(rows)):
            if rows[j][0] > z:
                after, before, min_bin = j, j - 1, j - 1
                break
        a, b = rows[after], rows[before]
        for k in range(3):
            if has[k]:
                slope = (a[1 + k] - b[1 + k]) / (a[0] - b[0])
                out[i, k] = b[1 + k] + (z - b[0]) * slope
    return axis, first, last, step, n, out


def _grid_limits(start, step, n):
    return [start[a] + (n[a] - 1) * step[a] for a in range(3)]


def read_srw_field_file(path, comment="#"):
    """An SRW field map (src/TField3D_Grid.cc:1368-1541): a title line, nine
    header lines '#value' (start, step, n for x, y, z; the start is used as
    given), then 'Bx By Bz' lines with x fastest.  Returns start, stop, step,
    n and the samples with z fastest."""
    try:
        lines = open(path).read().split("\n")
    except OSError:
        raise ValueError("cannot open file for reading SRW format")
    hv = []
    for line in lines[1:10]:
        if not line or line[0] != comment:
            raise ValueError("something is wrong with the comment character, it was not seen")
        hv.append(float(line[1:].split()[0]))
    start, step = [hv[0], hv[3], hv[6]], [hv[1], hv[4], hv[7]]
    n = [int(hv[2]), int(hv[5]), int(hv[8])]
    if n[0] < 1 or n[1] < 1:
        raise ValueError("invalid dimensions")
    total = n[0] * n[1] * n[2]
    rows = lines[10:10 + total]
    try:
        vals = np.array([[float(v) for v in r.split()[:3]] for r in rows], dtype=np.float64)
    except ValueError:
        raise ValueError("input file stream failure")
    if vals.shape != (total, 3):
        raise ValueError("input file stream failure")
    data = vals.reshape(n[2], n[1], n[0], 3).transpose(2, 1, 0, 3).reshape(total, 3)
    return start, _grid_limits(start, step, n), step, n, data


def read_spectra_field_file(path):
    """A SPECTRA field map (src/TField3D_Grid.cc:1543-1692): a title line, then
    'dx dy dz nx ny nz' with steps in mm; the grid is centred on the origin;
    then 'Bx By Bz' lines with z fastest."""
    try:
        lines = open(path).read().split("\n")
    except OSError:
        raise ValueError("cannot open file")
    h = lines[1].split()
    step = [float(h[0]) / 1000., float(h[1]) / 1000., float(h[2]) / 1000.]
    n = [int(h[3]), int(h[4]), int(h[5])]
    if n[0] < 1 or n[1] < 1:
        raise ValueError("invalid number of points in at least one dimension")
    start = [-(step[a] * (n[a] - 1)) / 2 for a in range(3)]
    total = n[0] * n[1] * n[2]
    rows = lines[2:2 + total]
    try:
        data = np.array([[float(v) for v in r.split()[:3]] for r in rows], dtype=np.float64)
    except ValueError:
        raise ValueError("file stream failure")
    if data.shape != (total, 3):
        raise ValueError("file stream failure")
    return start, _grid_limits(start, step, n), step, n, data


def read_binary_field_file(path, scaling=()):
    """A binary OSCARS grid (src/TField3D_Grid.cc:875-1085, format version 1,
    format string 'OSCARS'): int32 comment length and comment, int32 version,
    int32 format length and format; float32 start and step and int32 count per
    axis (x, y, z); then float32 Bx By Bz per point, z fastest, then y, then x.
    The start is re-centred about the middle as the reader computes it
    (XStartIN + XStepIN * (NX - 1) / 2. with the product in float32), field
    scale factors multiply the float32 values.  Returns start, stop, step,
    npoints and the samples [n, 3]."""
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0

    def i32():
        nonlocal pos
        if pos + 4 > len(buf):
            raise ValueError("error reading file.  Check format")
        v = int(np.frombuffer(buf, dtype="<i4", count=1, offset=pos)[0])
        pos += 4
        return v

    def f32():
        nonlocal pos
        if pos + 4 > len(buf):
            raise ValueError("error reading file.  Check format")
        v = np.frombuffer(buf, dtype="<f4", count=1, offset=pos)[0]
        pos += 4
        return v

    ncomment = i32()
    pos += max(ncomment, 0)
    version = i32()
    nfmt = i32()
    fmt = buf[pos:pos + nfmt].decode("latin-1")
    pos += nfmt
    if version != 1:
        raise ValueError("File Version number incorrect")
    if fmt != "OSCARS":
        raise ValueError("binary field files of format '%s' are not supported by this implementation" % fmt)
    hdr = [(f32(), f32(), i32()) for _ in range(3)]
    npts = [h[2] for h in hdr]
    if min(npts) < 1:
        raise ValueError("invalid number of points in at least one dimension")
    start, stop, step = [], [], []
    for a, (s_in, d_in, n) in enumerate(hdr):
        d = float(d_in) * scaling[a] if len(scaling) > a else float(d_in)
        middle = float(s_in) + float(np.float32(d_in * np.float32(n - 1))) / 2.
        s0 = middle - d * (n - 1) / 2.
        start.append(s0)
        stop.append(s0 + (n - 1) * d)
        step.append(d)
    n = npts[0] * npts[1] * npts[2]
    if pos + 12 * n > len(buf):
        raise ValueError("error reading file.  Check format")
    raw = np.frombuffer(buf, dtype="<f4", count=3 * n, offset=pos).reshape(n, 3).copy()
    for k in range(3):
        if len(scaling) > 3 + k and scaling[3 + k] != 1:
            raw[:, k] = (raw[:, k].astype(np.float64) * scaling[3 + k]).astype(np.float32)
    return start, stop, step, npts, raw.astype(np.float64)


def read_oscars_field_file(path):
    """Header values and samples of an OSCARS-format text field map
    (src/TField3D_Grid.cc:353-548: a comment line, nine header lines whose
    first token is the value, then one 'Bx By Bz' line per grid point)."""
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < 10:
        raise ValueError("error reading file.  Check format")
    hv = [float(l.split()[0]) for l in lines[1:10]]
    start = [hv[0], hv[3], hv[6]]
    step = [hv[1], hv[4], hv[7]]
    npts = [int(hv[2]), int(hv[5]), int(hv[8])]
    if min(npts) < 1:
        raise ValueError("invalid number of points in at least one dimension")
    n = npts[0] * npts[1] * npts[2]
    rows = lines[10:10 + n]
    if len(rows) < n:
        raise ValueError("error reading file.  Check format")
    data = np.array([[float(v) for v in l.split()[:3]] for l in rows])
    return start, step, npts, data
