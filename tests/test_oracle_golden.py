"""The C oracle against golden vectors written by the reference's own C++ core
(tests/golden/make_golden.py).  Bit-exact: same knots, same level samples,
same spectrum / flux / power-density doubles, same converged levels."""
import os

import numpy as np
import pytest

from tests.common import planar, epu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.fixture(scope="module")
def o_planar():
    from oracle.pyoracle import Oracle
    o = Oracle(nthreads=1)
    planar(o)
    o.calculate_trajectory()
    return o


def test_trajectory_and_levels(o_planar):
    g = load("planar_traj.npz")
    T = o_planar.traj
    assert len(T.t) == int(g["n"])
    idx = g["idx"]
    assert np.array_equal(T.t[idx], g["t"])
    assert np.array_equal(T.p[idx], g["p"])
    assert np.array_equal(T.ypp[idx], g["ypp"])
    lev = T.level(10)
    assert np.array_equal(lev[g["level10_idx"]], g["level10"])


@pytest.mark.parametrize("pol", ["", "lh", "cl"])
@pytest.mark.parametrize("quantity", ["flux", "level"])
def test_spectrum(o_planar, pol, quantity):
    g = load("planar_spectrum.npz")
    e, v, lev, conv, ef = o_planar.calculate_spectrum([0, 0, 30], npoints=37, energy_range_eV=[100, 2000],
                                                      polarization=pol, quantity=quantity, nthreads=1)
    assert np.array_equal(e, g["energies"])
    assert np.array_equal(v, g["%s_%s" % (pol or "all", quantity)])


@pytest.mark.parametrize("energy", [152.11, 2000.0])
def test_flux_map(o_planar, energy):
    g = load("planar_flux.npz")
    cont, v, lev, conv, ef, obs = o_planar.calculate_flux_rectangle(energy, [11, 9], [0.02, 0.02],
                                                                   translation=[0, 0, 30], nthreads=1)
    assert np.array_equal(cont, g["xyz"])
    assert np.array_equal(v, g["flux_%g" % energy])


@pytest.mark.parametrize("normal", [0, 1])
def test_power_density(normal):
    from oracle.pyoracle import Oracle
    g = load("epu_power.npz")
    o = Oracle(nthreads=1)
    epu(o)
    cont, v, lev, conv, obs = o.calculate_power_density_rectangle([11, 9], [0.08, 0.08], translation=[0, 0, 20],
                                                                  normal=normal, nthreads=1)
    assert np.array_equal(cont, g["xyz"])
    assert np.array_equal(v, g["pd_normal%d" % normal])


def test_grid_field():
    from oracle.pyoracle import Oracle
    g = load("grid_flux.npz")
    o = Oracle(nthreads=1)
    o.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_file(os.path.join(GOLD, "grid_field.txt"))
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(4001)
    T = o.calculate_trajectory()
    assert len(T.t) == int(g["n"])
    assert np.array_equal(T.t[::20], g["t"])
    assert np.array_equal(T.p[::20], g["p"])
    cont, v, lev, conv, ef, obs = o.calculate_flux_rectangle(150.0, [7, 5], [0.01, 0.01], translation=[0, 0, 30],
                                                            nthreads=1)
    assert np.array_equal(v, g["flux"])
