"""ctypes bindings of the C ABI in include/oscars_mi.h (liboscars_mi.so).

The library is built in-tree by ``oscars_amd/csrc/Makefile`` (called from
``__graft_entry__.build()``).  There is no CPU fallback: if the library is
missing, or no HIP device is visible, every compute call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboscars_mi.so")

OM_FIELD_UNDULATOR, OM_FIELD_EPU, OM_FIELD_GRID, OM_FIELD_GAUSSIAN, OM_FIELD_UNIFORM, OM_FIELD_QUADRUPOLE = range(6)
OM_MODE_PLAIN, OM_MODE_TILED, OM_MODE_MIXED, OM_MODE_POINT_MIXED = 0, 1, 2, 3
MODES = {"plain": OM_MODE_PLAIN, "tiled": OM_MODE_TILED, "mixed": OM_MODE_MIXED, "point_mixed": OM_MODE_POINT_MIXED}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)
_lp = C.POINTER(C.c_long)


class om_field(C.Structure):
    _fields_ = [("type", C.c_int), ("nperiods", C.c_int),
                ("field", C.c_double * 3), ("fieldB", C.c_double * 3), ("period", C.c_double * 3),
                ("translation", C.c_double * 3), ("phase", C.c_double), ("taper", C.c_double),
                ("grid_start", C.c_double * 3), ("grid_step", C.c_double * 3), ("grid_n", C.c_long * 3),
                ("grid_data", _dp), ("rotations", C.c_double * 3), ("shape", C.c_double * 3), ("K", C.c_double),
                ("width", C.c_double), ("grid_exact", C.c_int), ("grid_stop", C.c_double * 3),
                ("frequency", C.c_double), ("frequency_phase", C.c_double), ("time_offset", C.c_double)]


class om_particle(C.Structure):
    _fields_ = [("x0", C.c_double * 3), ("beta0", C.c_double * 3), ("t0", C.c_double), ("q", C.c_double),
                ("m", C.c_double), ("gamma", C.c_double), ("qomg", C.c_double), ("current", C.c_double),
                ("ctstart", C.c_double), ("ctstop", C.c_double), ("npoints", C.c_long)]


class om_conv(C.Structure):
    _fields_ = [("precision", C.c_double), ("max_level", C.c_int), ("max_level_extended", C.c_int),
                ("quantity", C.c_int), ("mode", C.c_int), ("weight", C.c_double), ("accumulate", C.c_int)]


class om_request(C.Structure):
    _fields_ = [("n_obs", C.c_long), ("obs_xyz", _dp), ("n_e", C.c_long), ("energies_eV", _dp)]


class om_batch_opts(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int), ("max_level", C.c_int), ("deep_level", C.c_int)]


class om_stats(C.Structure):
    _fields_ = [("evals", C.c_double), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("max_level", C.c_int), ("n_unconverged", C.c_long), ("n_rechecked", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = ["om_device_count", "om_device_info", "om_context_create", "om_context_destroy", "om_last_error",
           "om_trajectories", "om_trajectories_eb", "om_trajectory_set", "om_trajectory_get", "om_power_density_surface",
           "om_total_power", "om_efield_time", "om_field_points", "om_power_density_stl", "om_level_get",
           "om_trajectory_status", "om_flux",
           "om_spectrum", "om_power_density", "om_batch_setup", "om_batch_run", "om_batch_tune",
           "om_batch_profile",
           "om_adjoint_forward", "om_adjoint_backward", "om_flux_dev", "om_spectrum_dev", "om_power_density_dev",
           "om_flux_dev_noalloc", "om_adjoint_forward_dev", "om_adjoint_backward_dev", "om_batch_tune_deep"]

_lib = None


def lib():
    """Load liboscars_mi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oscars_amd: %s is missing - build it with __graft_entry__.build() "
                           "(there is no CPU fallback)" % LIB_PATH)
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as
    # the one this library links; the first loaded serves the whole process, and
    # a torch built against another ROCm release finds no GPU when it is not its
    # own.  Loading torch first keeps one runtime, torch's, for both (device
    # tensors then pass straight to the *_dev entry points).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.om_device_count.restype = C.c_int
    L.om_device_info.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
    L.om_context_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.om_context_destroy.argtypes = [C.c_void_p]
    L.om_last_error.argtypes = [C.c_void_p]
    L.om_last_error.restype = C.c_char_p
    L.om_trajectories.argtypes = [C.c_void_p, C.POINTER(om_field), C.c_int, C.POINTER(om_particle), C.c_int]
    L.om_trajectories_eb.argtypes = [C.c_void_p, C.POINTER(om_field), C.c_int, C.POINTER(om_field), C.c_int,
                                     C.POINTER(om_particle), C.c_int]
    L.om_trajectory_get.argtypes = [C.c_void_p, C.c_int, _lp, _dp, _dp, _dp]
    L.om_trajectory_set.argtypes = [C.c_void_p, C.POINTER(om_particle), C.c_long, _dp, _dp, C.c_int]
    L.om_level_get.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]
    L.om_trajectory_status.argtypes = [C.c_void_p, C.c_int]
    L.om_flux.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long, _dp, C.c_long, _dp, C.POINTER(om_conv), _dp, _ip,
                          _up, _dp, C.POINTER(om_stats)]
    L.om_spectrum.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_long, _dp, C.POINTER(om_conv), _dp, _ip, _up,
                              _dp, C.POINTER(om_stats)]
    L.om_power_density.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long, _dp, C.c_int, C.POINTER(om_conv), _dp,
                                   _ip, _up, C.POINTER(om_stats)]
    L.om_power_density_surface.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long, _dp, C.c_int, C.POINTER(om_conv),
                                           _dp, _ip, _up, C.POINTER(om_stats)]
    L.om_total_power.argtypes = [C.c_void_p, C.c_int, C.POINTER(om_conv), _dp, _dp, _ip, C.POINTER(om_stats)]
    L.om_efield_time.argtypes = [C.c_void_p, C.c_int, _dp, _lp, _dp, _dp]
    L.om_field_points.argtypes = [C.c_void_p, C.POINTER(om_field), C.c_int, _dp, C.c_long, C.c_double, _dp]
    L.om_power_density_stl.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long, C.POINTER(om_conv), _dp, _ip, _up,
                                       C.POINTER(om_stats)]
    _vp = C.c_void_p  # device addresses
    L.om_flux_dev.argtypes = [C.c_void_p, C.c_int, _vp, C.c_long, _vp, C.c_long, _dp, C.POINTER(om_conv), _vp, _vp,
                              _vp, _vp, C.POINTER(om_stats)]
    L.om_flux_dev_noalloc.argtypes = [C.c_void_p, C.c_int, _vp, C.c_long, _vp, C.c_long, _dp, C.POINTER(om_conv), _vp, _vp,
                              _vp, _vp, C.POINTER(om_stats)]
    L.om_spectrum_dev.argtypes = [C.c_void_p, C.c_int, _vp, _vp, C.c_long, _dp, C.POINTER(om_conv), _vp, _vp, _vp,
                                  _vp, C.POINTER(om_stats)]
    L.om_power_density_dev.argtypes = [C.c_void_p, C.c_int, _vp, C.c_long, _dp, C.c_int, C.POINTER(om_conv), _vp,
                                       _vp, _vp, C.POINTER(om_stats)]
    L.om_batch_setup.argtypes = [C.c_void_p, C.POINTER(om_request), C.c_int, _ip, _dp, _dp, C.POINTER(om_conv),
                                 C.POINTER(om_batch_opts)]
    L.om_batch_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(_dp), C.POINTER(om_stats)]
    L.om_batch_tune.argtypes = [C.c_void_p, _ip, C.c_int, C.c_long, _ip, _dp]
    L.om_batch_tune_deep.argtypes = [C.c_void_p, _ip, C.c_int, _ip, _dp]
    L.om_batch_profile.argtypes = [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_ulonglong), C.POINTER(om_stats)]
    L.om_adjoint_forward.argtypes = [C.c_void_p, C.POINTER(om_field), C.c_int, C.c_int, C.POINTER(om_particle),
                                     _dp, C.c_long, _dp, C.c_long, C.c_int, _dp, C.POINTER(om_stats)]
    L.om_adjoint_backward.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(om_stats)]
    L.om_adjoint_forward_dev.argtypes = [C.c_void_p, C.POINTER(om_field), C.c_int, C.c_int, _vp,
                                         C.POINTER(om_particle), _vp, C.c_long, _vp, C.c_long, C.c_int, _vp,
                                         C.POINTER(om_stats)]
    L.om_adjoint_backward_dev.argtypes = [C.c_void_p, _vp, _vp, _vp, _vp, C.POINTER(om_stats)]
    _lib = L
    return L


def device_count():
    """GPUs visible to the library (0 without one)."""
    return int(lib().om_device_count())


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


class Device:
    """One HIP device context (device buffers, stream, trajectory set)."""

    def __init__(self, device=0):
        L = lib()
        n = L.om_device_count()
        if n < 1:
            raise ValueError("You are requesting the GPU, but none were found")
        h = C.c_void_p()
        rc = L.om_context_create(int(device), C.byref(h))
        if rc:
            raise ValueError("could not create a HIP context on device %d (rc=%d)" % (device, rc))
        self.L = L
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.L.om_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise ValueError(self.L.om_last_error(self.h).decode())

    def info(self):
        buf = C.create_string_buffer(512)
        self.L.om_device_info(self.device, buf, 512)
        return buf.value.decode()

    def trajectories(self, fields, particles, check_status=True, efields=()):
        """RK4 + spline for `particles`.  check_status=False returns as soon as the
        work is queued; a |beta| >= 1 trajectory is then reported by the next
        call that uses it (every sum checks the status first)."""
        fa = (om_field * max(1, len(fields)))(*fields)
        pa = (om_particle * len(particles))(*particles)
        if efields:
            ea = (om_field * len(efields))(*efields)
            self._chk(self.L.om_trajectories_eb(self.h, fa, len(fields), ea, len(efields), pa, len(particles)))
        else:
            self._chk(self.L.om_trajectories(self.h, fa, len(fields), pa, len(particles)))
        if not check_status:
            return
        for i in range(len(particles)):
            if self.L.om_trajectory_status(self.h, i) != 0:
                raise ValueError("trajectory %d: |beta| >= 1 met during RK4 integration" % i)

    def trajectory_set(self, particle, t, p9, construct=0):
        """Make (t[n], p9[n, 9]) the one resident trajectory; construct bit 0 /
        bit 1 rebuild beta / a/c from the spline slopes on the device."""
        t = np.ascontiguousarray(t, dtype=np.float64)
        p9 = np.ascontiguousarray(p9, dtype=np.float64).reshape(-1, 9)
        if len(p9) != len(t):
            raise ValueError("t and p9 differ in length")
        self._chk(self.L.om_trajectory_set(self.h, C.byref(particle), len(t), _ptr(t), _ptr(p9), int(construct)))

    def trajectory(self, i=0):
        n = C.c_long(0)
        self._chk(self.L.om_trajectory_get(self.h, i, C.byref(n), None, None, None))
        t = np.zeros(n.value)
        p = np.zeros((n.value, 9))
        ypp = np.zeros((n.value, 9))
        self._chk(self.L.om_trajectory_get(self.h, i, C.byref(n), _ptr(t), _ptr(p), _ptr(ypp)))
        return t, p, ypp

    def level(self, i, L, extended=False):
        out = np.zeros((2 ** L, 10))
        self._chk(self.L.om_level_get(self.h, i, int(L), int(extended), _ptr(out)))
        return out

    @staticmethod
    def conv(precision=0.01, max_level=-2, max_level_extended=0, quantity=0, mode="plain", weight=1.0,
             accumulate=False):
        c = om_conv()
        c.precision = precision
        c.max_level = max_level
        c.max_level_extended = max_level_extended
        c.quantity = quantity
        c.mode = MODES[mode] if isinstance(mode, str) else int(mode)
        c.weight = weight
        c.accumulate = int(accumulate)
        return c

    def flux(self, particle, obs_xyz, energies_eV, pol6, conv, out=None, want_e=False):
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(np.atleast_1d(energies_eV), dtype=np.float64)
        n = len(obs) * len(e)
        out = np.zeros(n) if out is None else out
        lev = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.uint8)
        ef = np.zeros((n, 6)) if want_e else None
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_flux(self.h, int(particle), _ptr(obs), len(obs), _ptr(e), len(e), _ptr(pol),
                                 C.byref(conv), _ptr(out), _ptr(lev, _ip), _ptr(cv, _up), _ptr(ef), C.byref(st)))
        return out, lev, cv, (ef[:, 0::2] + 1j * ef[:, 1::2]) if want_e else None, st.as_dict()

    def spectrum(self, particle, obs, energies_eV, pol6, conv, out=None, want_e=False):
        o = np.ascontiguousarray(obs, dtype=np.float64)
        e = np.ascontiguousarray(energies_eV, dtype=np.float64)
        n = len(e)
        out = np.zeros(n) if out is None else out
        lev = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.uint8)
        ef = np.zeros((n, 6)) if want_e else None
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_spectrum(self.h, int(particle), _ptr(o), _ptr(e), n, _ptr(pol), C.byref(conv),
                                     _ptr(out), _ptr(lev, _ip), _ptr(cv, _up), _ptr(ef), C.byref(st)))
        return out, lev, cv, (ef[:, 0::2] + 1j * ef[:, 1::2]) if want_e else None, st.as_dict()

    def power_density(self, particle, obs_xyz, normal, directional, conv, out=None):
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        n = len(obs)
        out = np.zeros(n) if out is None else out
        lev = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.uint8)
        nv = np.ascontiguousarray(normal, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_power_density(self.h, int(particle), _ptr(obs), n, _ptr(nv), int(directional),
                                          C.byref(conv), _ptr(out), _ptr(lev, _ip), _ptr(cv, _up), C.byref(st)))
        return out, lev, cv, st.as_dict()

    def power_density_surface(self, particle, obs_xyz, normals, directional, conv, out=None):
        """Power density at points with a unit normal each (normals[n, 3]) or
        without normals (normals None)."""
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        n = len(obs)
        out = np.zeros(n) if out is None else out
        lev = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.uint8)
        nv = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(n, 3)
        st = om_stats()
        self._chk(self.L.om_power_density_surface(self.h, int(particle), _ptr(obs), n, _ptr(nv), int(directional),
                                                  C.byref(conv), _ptr(out), _ptr(lev, _ip), _ptr(cv, _up),
                                                  C.byref(st)))
        return out, lev, cv, st.as_dict()

    def power_density_stl(self, particle, tri15, conv, out=None):
        """Power density at STL triangle centres; tri15[n, 15] = A, B, C, normal, centre."""
        t = np.ascontiguousarray(tri15, dtype=np.float64).reshape(-1, 15)
        n = len(t)
        out = np.zeros(n) if out is None else out
        lev = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.uint8)
        st = om_stats()
        self._chk(self.L.om_power_density_stl(self.h, int(particle), _ptr(t), n, C.byref(conv), _ptr(out),
                                              _ptr(lev, _ip), _ptr(cv, _up), C.byref(st)))
        return out, lev, cv, st.as_dict()

    def efield_time(self, particle, obs):
        """(t[n], E[n, 3]): the time-domain electric field at obs, one value per knot."""
        o = np.ascontiguousarray(obs, dtype=np.float64).reshape(3)
        n = C.c_long(0)
        self._chk(self.L.om_efield_time(self.h, int(particle), _ptr(o), C.byref(n), None, None))
        t = np.zeros(n.value)
        e = np.zeros((n.value, 3))
        self._chk(self.L.om_efield_time(self.h, int(particle), _ptr(o), C.byref(n), _ptr(t), _ptr(e)))
        return t, e

    def field_points(self, fields, xyz, t=0.0):
        """Sum of `fields` at the rows of xyz[n, 3] and time t [s], evaluated on the device: [n, 3]."""
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(x)
        fa = (om_field * max(1, len(fields)))(*fields)
        self._chk(self.L.om_field_points(self.h, fa, len(fields), _ptr(x), len(x), float(t), _ptr(out)))
        return out

    def total_power(self, particle, conv):
        """(power [W], precision, level, stats); level -1: not converged."""
        pw, pr, lv = C.c_double(0), C.c_double(0), C.c_int(0)
        st = om_stats()
        self._chk(self.L.om_total_power(self.h, int(particle), C.byref(conv), C.byref(pw), C.byref(pr), C.byref(lv),
                                        C.byref(st)))
        return pw.value, pr.value, lv.value, st.as_dict()

    # ------------------------------------------------------------ device-resident variants
    @staticmethod
    def _dev_ptr(t, dtype, n, what):
        """Address of a contiguous GPU torch tensor of `n` elements of `dtype` (None: NULL)."""
        if t is None:
            return None
        if not (t.is_cuda and t.dtype == dtype and t.is_contiguous() and t.numel() == n):
            raise ValueError("%s: need a contiguous %s GPU tensor of %d elements" % (what, dtype, n))
        return t.data_ptr()

    def flux_dev(self, particle, obs_t, e_t, pol6, conv, out_t, lev_t=None, conv_t=None, efield_t=None):
        """om_flux on device tensors: obs_t (n_obs, 3) and e_t (n_e) float64, out_t (n_obs*n_e) float64,
        lev_t int32, conv_t uint8, efield_t (n_obs*n_e, 6) float64 (optional).  Returns the stats."""
        import torch
        n_obs, n_e = obs_t.shape[0], e_t.numel()
        n = n_obs * n_e
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_flux_dev(self.h, int(particle), self._dev_ptr(obs_t, torch.float64, 3 * n_obs, "obs"),
                                     n_obs, self._dev_ptr(e_t, torch.float64, n_e, "energies"), n_e, _ptr(pol),
                                     C.byref(conv), self._dev_ptr(out_t, torch.float64, n, "out"),
                                     self._dev_ptr(lev_t, torch.int32, n, "levels"),
                                     self._dev_ptr(conv_t, torch.uint8, n, "converged"),
                                     self._dev_ptr(efield_t, torch.float64, 6 * n, "efield"), C.byref(st)))
        return st.as_dict()

    def flux_dev_noalloc(self, particle, obs_t, e_t, pol6, conv, out_t, lev_t=None, conv_t=None, efield_t=None):
        """om_flux_dev_noalloc: flux_dev with the library's static device scratch,
        allocating nothing.  Returns the stats."""
        import torch
        n_obs, n_e = obs_t.shape[0], e_t.numel()
        n = n_obs * n_e
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_flux_dev_noalloc(self.h, int(particle), self._dev_ptr(obs_t, torch.float64, 3 * n_obs, "obs"),
                                             n_obs, self._dev_ptr(e_t, torch.float64, n_e, "energies"), n_e, _ptr(pol),
                                             C.byref(conv), self._dev_ptr(out_t, torch.float64, n, "out"),
                                             self._dev_ptr(lev_t, torch.int32, n, "levels"),
                                             self._dev_ptr(conv_t, torch.uint8, n, "converged"),
                                             self._dev_ptr(efield_t, torch.float64, 6 * n, "efield"), C.byref(st)))
        return st.as_dict()

    def spectrum_dev(self, particle, obs_t, e_t, pol6, conv, out_t, lev_t=None, conv_t=None, efield_t=None):
        """om_spectrum on device tensors (obs_t: 3 float64); see flux_dev."""
        import torch
        n = e_t.numel()
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_spectrum_dev(self.h, int(particle), self._dev_ptr(obs_t, torch.float64, 3, "obs"),
                                         self._dev_ptr(e_t, torch.float64, n, "energies"), n, _ptr(pol),
                                         C.byref(conv), self._dev_ptr(out_t, torch.float64, n, "out"),
                                         self._dev_ptr(lev_t, torch.int32, n, "levels"),
                                         self._dev_ptr(conv_t, torch.uint8, n, "converged"),
                                         self._dev_ptr(efield_t, torch.float64, 6 * n, "efield"), C.byref(st)))
        return st.as_dict()

    def power_density_dev(self, particle, obs_t, normal, directional, conv, out_t, lev_t=None, conv_t=None):
        """om_power_density on device tensors; see flux_dev."""
        import torch
        n = obs_t.shape[0]
        nv = np.ascontiguousarray(normal, dtype=np.float64)
        st = om_stats()
        self._chk(self.L.om_power_density_dev(self.h, int(particle), self._dev_ptr(obs_t, torch.float64, 3 * n, "obs"),
                                              n, _ptr(nv), int(directional), C.byref(conv),
                                              self._dev_ptr(out_t, torch.float64, n, "out"),
                                              self._dev_ptr(lev_t, torch.int32, n, "levels"),
                                              self._dev_ptr(conv_t, torch.uint8, n, "converged"), C.byref(st)))
        return st.as_dict()

    # ------------------------------------------------------------ batched multi-particle sums
    def batch_setup(self, requests, req_of, weight, pol6, conv, waves_per_cu=0, max_level=0, deep_level=0):
        """requests: list of (obs_xyz[n,3], energies[m]); particle i adds weight[i]*flux to requests[req_of[i]]."""
        self._batch_keep = []
        ra = (om_request * len(requests))()
        shapes = []
        for r, (obs, e) in enumerate(requests):
            o = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
            en = np.ascontiguousarray(e, dtype=np.float64)
            self._batch_keep += [o, en]
            ra[r].n_obs = len(o)
            ra[r].obs_xyz = _ptr(o)
            ra[r].n_e = len(en)
            ra[r].energies_eV = _ptr(en)
            shapes.append((len(o), len(en)))
        ro = np.ascontiguousarray(req_of, dtype=np.int32)
        w = np.ascontiguousarray(weight, dtype=np.float64)
        pol = np.ascontiguousarray(pol6, dtype=np.float64)
        opts = om_batch_opts(int(waves_per_cu), int(max_level), int(deep_level))
        self._chk(self.L.om_batch_setup(self.h, ra, len(requests), _ptr(ro, _ip), _ptr(w), _ptr(pol),
                                        C.byref(conv), C.byref(opts)))
        self._batch_shapes = shapes

    def batch_run(self, recompute_trajectories=False):
        outs = [np.zeros(n * m) for n, m in self._batch_shapes]
        arr = (_dp * len(outs))(*[_ptr(o) for o in outs])
        st = om_stats()
        self._chk(self.L.om_batch_run(self.h, int(bool(recompute_trajectories)), arr, C.byref(st)))
        return [o.reshape(n, m) for o, (n, m) in zip(outs, self._batch_shapes)], st.as_dict()

    def batch_profile(self, recompute_trajectories=True):
        """One batch step launched stage by stage with HIP events between the
        stages: ms of (trajectories, level rows, persistent kernel, deep-level
        kernel, FP64 recheck kernel) and the device counters."""
        ms = np.zeros(5)
        cnt = (C.c_ulonglong * 9)()
        st = om_stats()
        self._chk(self.L.om_batch_profile(self.h, int(bool(recompute_trajectories)), _ptr(ms), cnt, C.byref(st)))
        names = ("units", "unconverged", "evals", "fix_listed", "handed_over", "main_amplitudes", "main_evals",
                 "rechecked_in_unit", "main_slots")
        stages = ("trajectories", "levels", "main", "deep", "fix")
        return dict(zip(stages, ms.tolist())), {k: int(v) for k, v in zip(names, cnt)}, st.as_dict()

    def adjoint_forward(self, fields, grad_field, particle, obs_xyz, energies_eV, level):
        """Frozen-level flux (n_obs x n_e) of one particle; keeps what adjoint_backward needs."""
        fa = (om_field * max(1, len(fields)))(*fields)
        obs = np.ascontiguousarray(obs_xyz, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(energies_eV, dtype=np.float64).reshape(-1)
        out = np.zeros(len(obs) * len(e))
        st = om_stats()
        self._chk(self.L.om_adjoint_forward(self.h, fa, len(fields), int(grad_field), C.byref(particle), _ptr(obs),
                                            len(obs), _ptr(e), len(e), int(level), _ptr(out), C.byref(st)))
        self._adj_shape = (len(obs), len(e))
        return out.reshape(len(obs), len(e)), st.as_dict()

    def adjoint_backward(self, dloss_dflux, n_grid):
        g = np.ascontiguousarray(dloss_dflux, dtype=np.float64).reshape(-1)
        if g.size != self._adj_shape[0] * self._adj_shape[1]:
            raise ValueError("dloss_dflux has %d values, the forward produced %d" %
                             (g.size, self._adj_shape[0] * self._adj_shape[1]))
        gf = np.zeros((max(n_grid, 1), 3))
        gx = np.zeros(3)
        gb = np.zeros(3)
        st = om_stats()
        self._chk(self.L.om_adjoint_backward(self.h, _ptr(g), _ptr(gf) if n_grid > 0 else None, _ptr(gx), _ptr(gb),
                                             C.byref(st)))
        return gf[:n_grid], gx, gb, st.as_dict()

    def adjoint_forward_dev(self, fields, grad_field, samples_t, particle, obs_t, e_t, level, out_t):
        """om_adjoint_forward_dev on GPU tensors: the grid field's samples samples_t (n, 3) float64 (None: the
        om_field's host samples), obs_t (n_obs, 3), e_t (n_e), out_t (n_obs * n_e).  Returns the stats."""
        import torch
        fa = (om_field * max(1, len(fields)))(*fields)
        n_obs, n_e = obs_t.shape[0], e_t.numel()
        ns = samples_t.numel() if samples_t is not None else 0
        st = om_stats()
        self._chk(self.L.om_adjoint_forward_dev(
            self.h, fa, len(fields), int(grad_field), self._dev_ptr(samples_t, torch.float64, ns, "samples"),
            C.byref(particle), self._dev_ptr(obs_t, torch.float64, 3 * n_obs, "obs"), n_obs,
            self._dev_ptr(e_t, torch.float64, n_e, "energies"), n_e, int(level),
            self._dev_ptr(out_t, torch.float64, n_obs * n_e, "out"), C.byref(st)))
        self._adj_shape = (n_obs, n_e)
        return st.as_dict()

    def adjoint_backward_dev(self, dflux_t, gf_t, gx_t, gb_t):
        """om_adjoint_backward_dev on GPU tensors: dflux_t (n_obs * n_e), gf_t (n_grid, 3) or None, gx_t and gb_t
        (3) or None.  Returns the stats."""
        import torch
        n = self._adj_shape[0] * self._adj_shape[1]
        st = om_stats()
        self._chk(self.L.om_adjoint_backward_dev(
            self.h, self._dev_ptr(dflux_t, torch.float64, n, "dloss_dflux"),
            self._dev_ptr(gf_t, torch.float64, gf_t.numel() if gf_t is not None else 0, "grad_field"),
            self._dev_ptr(gx_t, torch.float64, 3, "grad_x0"), self._dev_ptr(gb_t, torch.float64, 3, "grad_beta0"),
            C.byref(st)))
        return st.as_dict()

    def batch_tune_deep(self, deep_levels=(0, 13, 15)):
        """om_batch_tune_deep: (best hand-over level, {level: ms})."""
        c = np.ascontiguousarray(deep_levels, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int(0)
        self._chk(self.L.om_batch_tune_deep(self.h, _ptr(c, _ip), len(c), C.byref(best), _ptr(ms)))
        return best.value, dict(zip([int(x) for x in c], ms.tolist()))

    def batch_tune(self, candidates=(2, 4, 8, 16, 32), sample_items=20000):
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = C.c_int(0)
        self._chk(self.L.om_batch_tune(self.h, _ptr(c, _ip), len(c), int(sample_items), C.byref(best), _ptr(ms)))
        return best.value, dict(zip([int(x) for x in c], ms.tolist()))
