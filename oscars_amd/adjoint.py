"""Gradients of flux maps with respect to a 1-D field map and the beam angles.

The reference computes flux only (OSCARSSR::CalculateFlux, src/OSCARSSR.cc:3905-4068);
this module differentiates that calculation at a frozen level (every output
point sums trajectory levels 0..level, i.e. precision 0 and max_level=level):

    flux = FieldMapFlux.apply(samples, angles, problem)

``samples`` is the (n, 3) table of a TField3D_Grid map (src/TField3D_Grid.cc:153-278),
``angles`` the electron's horizontal and vertical angles (x', y') applied to the
design direction as TParticleBeam::GetNewParticle applies them
(src/TParticleBeam.cc:752-773).  The backward pass runs the HIP adjoint kernels
(RK4 with sqrt(N) checkpoints, spline, LW sum) through om_adjoint_backward.
"""
import math

import numpy as np

from . import _lib
from . import geometry as G

C = G.C


# ---------------------------------------------------------------- field maps
def undulator_pole_map(n=4001, z0=-1.0, dz=5e-4, b0=1.0, period=0.049, nperiods=31, sigma_amp=0.0, sigma_z=0.0,
                       seed=0):
    """By(z) samples of the ideal planar undulator (src/TField3D_IdealUndulator.cc:113-152:
    length period*(nperiods+2), quarter amplitude in the outer half periods,
    three quarters in the next) with per-pole errors: pole k (half period k) is
    scaled by (1 + sigma_amp g_k) and shifted along z by sigma_z h_k, g and h
    standard normal from numpy PCG64(seed).  Returns an (n, 3) array."""
    rng = np.random.Generator(np.random.PCG64(seed))
    npoles = 2 * (nperiods + 2)
    g = rng.standard_normal(npoles)
    h = rng.standard_normal(npoles)
    L = period * (nperiods + 2)
    z = z0 + dz * np.arange(n)
    out = np.zeros((n, 3))
    for i, d in enumerate(z):
        if d > L / 2 or d < -L / 2:
            continue
        k = min(int((d + L / 2) / (period / 2)), npoles - 1)
        scale = 0.25 if k in (0, npoles - 1) else (0.75 if k in (1, npoles - 2) else 1.0)
        out[i, 1] = b0 * scale * (1 + sigma_amp * g[k]) * math.sin(2 * math.pi * (d - sigma_z * h[k]) / period)
    return out


def grid_field(start, step, npts, samples, translation=(0.0, 0.0, 0.0)):
    """om_field of a grid map (no rotation: the reverse pass does not support one)."""
    s = np.ascontiguousarray(samples, dtype=np.float64).reshape(-1, 3)
    f = _lib.om_field()
    f.type = _lib.OM_FIELD_GRID
    f.grid_start[:] = [float(v) for v in start]
    f.grid_step[:] = [float(v) for v in step]
    f.grid_n[:] = [int(v) for v in npts]
    f.grid_data = s.ctypes.data_as(_lib._dp)
    f.rotations[:] = [0.0, 0.0, 0.0]
    f.translation[:] = [float(v) for v in translation]
    return f, s


# ---------------------------------------------------------------- beam angles
def _rot_matrix(A, axis):
    V = G.unit(axis)
    ca, sa = math.cos(A), math.sin(A)
    return np.array([[ca + V[0] * V[0] * (1 - ca), V[0] * V[1] * (1 - ca) - V[2] * sa, V[0] * V[2] * (1 - ca) + V[1] * sa],
                     [V[1] * V[0] * (1 - ca) + V[2] * sa, ca + V[1] * V[1] * (1 - ca), V[1] * V[2] * (1 - ca) - V[0] * sa],
                     [V[2] * V[0] * (1 - ca) - V[1] * sa, V[2] * V[1] * (1 - ca) + V[0] * sa, ca + V[2] * V[2] * (1 - ca)]])


def _rot_matrix_dA(A, axis):
    V = G.unit(axis)
    ca, sa = math.cos(A), math.sin(A)
    return np.array([[-sa + V[0] * V[0] * sa, V[0] * V[1] * sa - V[2] * ca, V[0] * V[2] * sa + V[1] * ca],
                     [V[1] * V[0] * sa + V[2] * ca, -sa + V[1] * V[1] * sa, V[1] * V[2] * sa - V[0] * ca],
                     [V[2] * V[0] * sa - V[1] * ca, V[2] * V[1] * sa + V[0] * ca, -sa + V[2] * V[2] * sa]])


def beta0_from_angles(beta_ref, xp, yp, hdir=(1.0, 0.0, 0.0), vdir=(0.0, 1.0, 0.0)):
    """beta_ref rotated by x' about the vertical, then by y' about -horizontal
    (the order of src/TParticleBeam.cc:769-770).  Returns (beta0, d beta0/dx', d beta0/dy')."""
    b = np.asarray(beta_ref, dtype=np.float64)
    mh = [-v for v in hdir]
    Rx, Ry = _rot_matrix(xp, vdir), _rot_matrix(yp, mh)
    beta0 = Ry @ (Rx @ b)
    d_xp = Ry @ (_rot_matrix_dA(xp, vdir) @ b)
    d_yp = _rot_matrix_dA(yp, mh) @ (Rx @ b)
    return beta0, d_xp, d_yp


# ---------------------------------------------------------------- problem
class FieldMapProblem:
    """One electron through a differentiated grid map (plus optional fixed
    fields), flux on obs x energies at a frozen level, on one device."""

    def __init__(self, device, grid_start, grid_step, grid_n, energy_GeV, x0, ctstart, ctstop, npoints, obs,
                 energies_eV, level, current=0.5, t0=0.0, other_fields=(), d0=(0.0, 0.0, 1.0)):
        self.dev = device
        self.grid = (list(grid_start), list(grid_step), list(grid_n))
        self.n_grid = int(np.prod(grid_n))
        self.p0 = G.ideal_particle("electron", energy_GeV, list(x0), list(d0), t0, current)
        self.beta_ref = np.array(self.p0.b0)
        self.window = (float(ctstart), float(ctstop), int(npoints))
        self.obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
        self.energies = np.ascontiguousarray(energies_eV, dtype=np.float64).reshape(-1)
        self.level = int(level)
        self.other = list(other_fields)
        self._jac = None
        self._dev_inputs = None
        self.last_stats = {}

    def _particle(self, angles):
        beta0, dx, dy = beta0_from_angles(self.beta_ref, float(angles[0]), float(angles[1]))
        p = G.Particle(self.p0.q, self.p0.m, self.p0.x0, beta0.tolist(), self.p0.t0, self.p0.current)
        o = _lib.om_particle()
        o.x0[:] = p.x0
        o.beta0[:] = p.b0
        o.t0 = p.t0
        o.q = p.q
        o.m = p.m
        o.gamma = p.gamma
        o.qomg = p.qomg
        o.current = p.current
        o.ctstart, o.ctstop, o.npoints = self.window
        return o, np.stack([dx, dy], axis=1)

    def forward(self, samples, angles=(0.0, 0.0)):
        f, keep = grid_field(*self.grid, samples)
        part, jac = self._particle(angles)
        fields = [f] + self.other
        flux, st = self.dev.adjoint_forward(fields, 0, part, self.obs, self.energies, self.level)
        self._keep = keep
        self._jac = jac
        self.last_stats["forward"] = st
        return flux

    def backward(self, dflux):
        """(dLoss/dsamples (n, 3), dLoss/d(x', y'), dLoss/dx0) for the last forward."""
        gf, gx, gb, st = self.dev.adjoint_backward(dflux, self.n_grid)
        self.last_stats["backward"] = st
        return gf, gb @ self._jac, gx

    # device-resident variants: float64 torch tensors on the GPU in and out
    def forward_dev(self, samples_t, angles=(0.0, 0.0)):
        """forward() with the samples (n, 3) as a GPU tensor; returns the flux
        (n_obs, n_e) as a GPU tensor.  The two beam angles are host numbers
        (they set the particle the trajectory is launched with)."""
        import torch
        f, _ = grid_field(*self.grid, np.zeros((0, 3)))
        f.grid_data = None
        part, jac = self._particle(angles)
        out = torch.empty(len(self.obs) * len(self.energies), dtype=torch.float64, device=samples_t.device)
        if self._dev_inputs is None or self._dev_inputs[0].device != samples_t.device:
            self._dev_inputs = (torch.as_tensor(self.obs, device=samples_t.device),
                                torch.as_tensor(self.energies, device=samples_t.device))
        st = self.dev.adjoint_forward_dev([f] + self.other, 0, samples_t.contiguous(), part, self._dev_inputs[0],
                                          self._dev_inputs[1], self.level, out)
        self._jac = jac
        self.last_stats["forward"] = st
        return out.view(len(self.obs), len(self.energies))

    def backward_dev(self, dflux_t):
        """backward() on GPU tensors: (dLoss/dsamples (n, 3), dLoss/d(x', y') (2), dLoss/dx0 (3))."""
        import torch
        dev = dflux_t.device
        gf = torch.empty((self.n_grid, 3), dtype=torch.float64, device=dev)
        gx = torch.empty(3, dtype=torch.float64, device=dev)
        gb = torch.empty(3, dtype=torch.float64, device=dev)
        st = self.dev.adjoint_backward_dev(dflux_t.contiguous().view(-1), gf, gx, gb)
        self.last_stats["backward"] = st
        return gf, gb @ torch.as_tensor(self._jac, device=dev), gx


try:
    import torch

    class FieldMapFlux(torch.autograd.Function):
        """flux[n_obs, n_e] = F(samples[n, 3], angles[2]); float64 tensors on any device."""

        @staticmethod
        def forward(ctx, samples, angles, problem):
            ctx.problem = problem
            ctx.dev_samples = samples.device
            ctx.dev_angles = angles.device
            ctx.on_gpu = samples.is_cuda
            if samples.is_cuda:  # samples, flux and gradients stay on the device
                return problem.forward_dev(samples.detach().to(torch.float64),
                                           [float(a) for a in angles.detach().cpu().tolist()])
            flux = problem.forward(samples.detach().cpu().numpy(), angles.detach().cpu().numpy())
            return torch.from_numpy(flux).to(samples.device)

        @staticmethod
        def backward(ctx, gout):
            if ctx.on_gpu:
                gf, ga, _ = ctx.problem.backward_dev(gout.detach().to(torch.float64))
                return gf, ga.to(ctx.dev_angles), None
            gf, ga, _ = ctx.problem.backward(gout.detach().cpu().numpy())
            return (torch.from_numpy(np.ascontiguousarray(gf)).to(ctx.dev_samples),
                    torch.from_numpy(np.ascontiguousarray(ga)).to(ctx.dev_angles), None)
except ImportError:  # torch is optional for the numpy interface
    FieldMapFlux = None
