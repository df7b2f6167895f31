"""The reference's own C++ calculations with gpu=1, served by this library.

oracle/_ref/libref_mi.so is the reference core compiled from its sources with
its GPU boundary enabled (-DCUDA) and linked with
oscars_amd/reference_shim/OSCARSSR_MI.cc over liboscars_mi.so (oracle/Makefile,
target ref_mi).  OSCARSSR::CalculateSpectrum / CalculateFlux /
CalculatePowerDensity then take their GPU branch (src/OSCARSSR.cc:1733-1747,
3974-3988, 2734-2744) into OSCARSSR_Cuda_Calculate*GPU, which hand the
reference's trajectory to om_trajectory_set and sum on the MI355X.  The
results are compared with the same calculations of the plain reference build
on the CPU: values within 1e-9 relative (denominator floored at 1e-6 of the
largest value), convergence flags equal.  Skipped where the reference build is
absent."""
import numpy as np
import pytest

from tests.common import EPU, UNDULATOR, rel_err

gpu = pytest.mark.gpu


def _sims(setup):
    import torch  # noqa: F401  (the HIP runtime torch ships must be loaded first)
    from oracle import pyoracle as P
    if not (P.ref_available() and P.ref_mi_available()):
        pytest.skip("reference builds (oracle/_ref/libref.so, libref_mi.so) not present")
    cpu, mi = P.RefSim(nthreads=8), P.RefSim(nthreads=8, gpu=True)
    setup(cpu)
    setup(mi)
    return cpu, mi


def _planar(s):
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_undulator(**UNDULATOR)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(8001)


@gpu
@pytest.mark.parametrize("polarization", ["", "linear-horizontal", "circular-left", "linear"])
def test_reference_spectrum_on_gpu(polarization):
    cpu, mi = _sims(_planar)
    kw = dict(npoints=40, energy_range_eV=[100, 1200], polarization=polarization,
              angle=0.3 if polarization == "linear" else None)
    e, v, conv = cpu.calculate_spectrum([0, 0.0002, 30], **kw)
    e2, v2, conv2 = mi.calculate_spectrum([0, 0.0002, 30], **kw)
    err = rel_err(v2, v, 1e-6)
    print("reference spectrum gpu=1 (%s): %.3e" % (polarization or "all", err))
    assert np.array_equal(e, e2) and conv == conv2 and err < 1e-9


@gpu
@pytest.mark.parametrize("quantity", ["flux", "level"])
def test_reference_flux_on_gpu(quantity):
    cpu, mi = _sims(_planar)
    kw = dict(npoints=[11, 9], width=[0.01, 0.008], translation=[0, 0, 30], quantity=quantity)
    x, v, conv = cpu.calculate_flux_rectangle(152.11, **kw)
    x2, v2, conv2 = mi.calculate_flux_rectangle(152.11, **kw)
    err = rel_err(v2, v, 1e-6)
    print("reference flux map gpu=1 (%s): %.3e" % (quantity, err))
    assert np.array_equal(x, x2) and conv == conv2
    assert np.array_equal(v, v2) if quantity == "level" else err < 1e-9


@gpu
def test_reference_power_density_on_gpu():
    def epu(s):
        s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
        s.add_bfield_epu(**EPU)
        s.set_ctstartstop(0, 3)
        s.set_npoints_trajectory(12001)

    cpu, mi = _sims(epu)
    kw = dict(npoints=[11, 11], width=[0.02, 0.02], translation=[0, 0, 20])
    x, v, conv = cpu.calculate_power_density_rectangle(**kw)
    x2, v2, conv2 = mi.calculate_power_density_rectangle(**kw)
    err = rel_err(v2, v, 1e-6)
    print("reference power density gpu=1: %.3e" % err)
    assert np.array_equal(x, x2) and conv == conv2 and err < 1e-9


@gpu
def test_reference_multi_particle_on_gpu():
    """nparticles=3 from a beam with spread: the reference draws each particle
    (SetNewParticle) and integrates its trajectory on the host in both builds,
    so the same seed gives the same particles; the sums run on the GPU in one."""
    def beam(s):
        s.set_particle_beam_full(energy_GeV=3.0, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5,
                                 sigma_energy_GeV=0.0027, emittance=[0.9e-9, 8e-12], beta=[1.5, 0.8])
        s.add_bfield_undulator(**UNDULATOR)
        s.set_ctstartstop(0, 2)
        s.set_npoints_trajectory(6001)
        s.set_seed(17)

    cpu, mi = _sims(beam)
    e, v, _ = cpu.calculate_spectrum([0, 0, 30], npoints=24, energy_range_eV=[100, 800], nparticles=3)
    e2, v2, _ = mi.calculate_spectrum([0, 0, 30], npoints=24, energy_range_eV=[100, 800], nparticles=3)
    err = rel_err(v2, v, 1e-6)
    print("reference nparticles=3 gpu=1: %.3e" % err)
    assert err < 1e-9
