"""nparticles= on the GPU with the reference's random stream: after
sr.set_seed(s) the product draws the particles the reference draws after
set_seed(s) (pinned bit for bit in test_beam_sampling.py), so its
multi-particle spectrum, flux map and power density equal the oracle summed
over those particles with weights 1/N.  Bars: 1e-9 FP64, 1e-5 mixed, relative
with a floor at 1e-6 of the maximum."""
import numpy as np
import pytest

from oscars_amd import beam as B
from tests.common import BEAM_CASES, UNDULATOR, rel_err

gpu = pytest.mark.gpu
SEED = 5
NP = 4
BEAM = dict(BEAM_CASES[1], d0=[0, 0, 1], x0=[0.0005, -0.0002, -1.0])


def _pair():
    from oracle.pyoracle import Oracle
    from oscars_amd.sr import sr
    s = sr()
    s.set_particle_beam(**BEAM)
    s.add_bfield_undulator(**UNDULATOR)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(6001)
    o = Oracle(nthreads=8)
    o.set_particle_beam(3.0, x0=BEAM["x0"], d0=BEAM["d0"], current=0.5)
    o.add_bfield_undulator(**UNDULATOR)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(6001)
    return s, o


def _oracle_sum(s, o, fn):
    tot = None
    for p in B.sample_particles(s._beam, NP, seed=SEED):
        o.set_particle(p.x0, p.b0)
        o.calculate_trajectory()
        v = fn(o) * (1.0 / NP)
        tot = v if tot is None else tot + v
    return tot


@gpu
@pytest.mark.parametrize("mode,tol", [("plain", 1e-9), ("tiled", 1e-9), ("mixed", 1e-5)])
def test_nparticles_spectrum(mode, tol):
    s, o = _pair()
    e = np.linspace(140.0, 600.0, 64)
    ref = _oracle_sum(s, o, lambda o: o.spectrum([0, 0, 30], e, carry_quirk=False)[0])
    s.set_seed(SEED)
    res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(e), nparticles=NP, mode=mode)
    v = np.array([r[1] for r in res])
    err = rel_err(v, ref, 1e-6)
    print("nparticles spectrum %s: max rel err %.3e" % (mode, err))
    assert err < tol


@gpu
@pytest.mark.parametrize("mode,tol", [("plain", 1e-9), ("mixed", 1e-5)])
def test_nparticles_flux_rectangle(mode, tol):
    s, o = _pair()
    obs, _, _ = o.rectangle("XY", [11, 11], [0.004, 0.004], translation=[0, 0, 30])
    ref = _oracle_sum(s, o, lambda o: o.flux_points(obs, 440.0, carry_quirk=False)[0])
    s.set_seed(SEED)
    res = s.calculate_flux_rectangle(energy_eV=440.0, npoints=[11, 11], width=[0.004, 0.004],
                                     translation=[0, 0, 30], nparticles=NP, mode=mode)
    v = np.array([r[1] for r in res])
    err = rel_err(v, ref, 1e-6)
    print("nparticles flux %s: max rel err %.3e" % (mode, err))
    assert err < tol


@gpu
def test_nparticles_power_density():
    s, o = _pair()
    obs, nv, _ = o.rectangle("XY", [11, 11], [0.02, 0.02], translation=[0, 0, 30])
    ref = _oracle_sum(s, o, lambda o: o.power_density_points(obs, nv, directional=False, carry_quirk=False)[0])
    s.set_seed(SEED)
    res = s.calculate_power_density_rectangle(npoints=[11, 11], width=[0.02, 0.02], translation=[0, 0, 30],
                                              nparticles=NP)
    v = np.array([r[1] for r in res])
    err = rel_err(v, ref, 1e-6)
    print("nparticles power density: max rel err %.3e" % err)
    assert err < 1e-9
    s.set_seed(SEED)
    res = s.calculate_power_density_rectangle(npoints=[11, 11], width=[0.02, 0.02], translation=[0, 0, 30],
                                              nparticles=NP, mode="mixed")
    vm = np.array([r[1] for r in res])
    em = rel_err(vm, ref, 1e-6)
    print("nparticles power density, mixed: max rel err %.3e" % em)
    assert em < 1e-5
