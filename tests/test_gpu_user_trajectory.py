"""set_trajectory on the GPU: a trajectory given as a file or a list is
splined on the device (and, where the file holds positions only, beta and
a/c are rebuilt from spline slopes first), then summed by the same LW kernels.

Reference values: tests/golden/user_trajectory.npz, written by the reference
build (tests/golden/make_traj_golden.py): the planar-undulator trajectory
thinned to uneven steps, read by the reference from a 'T X Y Z' text file.
Bars: knots and spline second derivatives within 1e-13 of the channel's
largest value (the same double operations in the same order); spectrum and
power density 1e-9 FP64, 1e-5 mixed."""
import os
import tempfile

import numpy as np
import pytest

from oscars_amd import io as IO
from tests.common import UNDULATOR, rel_err

gpu = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "user_trajectory.npz")


def _sr():
    from oscars_amd.sr import sr
    s = sr()
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    return s


def _chan_err(a, b):
    scale = np.maximum(np.max(np.abs(b), axis=0), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


@gpu
def test_set_trajectory_text_file_matches_reference():
    g = np.load(GOLD)
    s = _sr()
    p9 = np.zeros((len(g["t_in"]), 9))
    p9[:, 0:3] = g["x_in"]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "txyz.txt")
        IO.write_trajectory_text(path, g["t_in"], p9, "T X Y Z")
        s.set_trajectory(ifile=path)
    t, p, ypp = s._device().trajectory(0)
    k = g["check"]
    ep, ey = _chan_err(p[k], g["p9"]), _chan_err(ypp[k], g["ypp"])
    same = np.mean(p[k] == g["p9"])
    print("user trajectory: knots %.3e, ypp %.3e (knot values bit-identical: %.1f%%)" % (ep, ey, 100 * same))
    assert np.array_equal(t, g["t"])
    assert ep < 1e-13 and ey < 1e-13
    for mode, tol in (("plain", 1e-9), ("tiled", 1e-9), ("mixed", 1e-5)):
        res = s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=list(g["energies"]), mode=mode)
        e = rel_err(np.array([r[1] for r in res]), g["spectrum"], 1e-6)
        print("user trajectory spectrum %s: %.3e" % (mode, e))
        assert e < tol
    res = s.calculate_power_density_rectangle(npoints=[4, 3], width=[0.009, 0.003],
                                              translation=[-0.0015, -0.0005, 30])
    e = rel_err(np.array([r[1] for r in res]), g["pd"])
    print("user trajectory power density: %.3e" % e)
    assert e < 1e-9


@gpu
def test_trajectory_round_trips():
    """calculate_trajectory's own knots given back through a binary file and
    through a list give the same knots, and the spectrum of the integrated
    trajectory to 1e-12 (the given trajectory is splined by the sequential
    sweep, the integrated one by the windowed parallel form)."""
    s = _sr()
    s.add_bfield_undulator(**UNDULATOR)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(4001)
    e = list(np.linspace(150.0, 600.0, 40))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.dat")
        traj = s.calculate_trajectory(bofile=path)
        ref = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=e, mode="plain")])
        t0, p0, _ = s._device().trajectory(0)
        s.set_trajectory(bifile=path)
        t1, p1, _ = s._device().trajectory(0)
        assert np.array_equal(t0, t1) and np.array_equal(p0, p1)
        v = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=e, mode="plain")])
        assert rel_err(v, ref, 1e-6) < 1e-12
    s.set_trajectory(trajectory=traj)
    t2, p2, _ = s._device().trajectory(0)
    assert np.array_equal(t0, t2) and np.array_equal(p0, p2)
    v = np.array([r[1] for r in s.calculate_spectrum(obs=[0, 0, 30], energy_points_eV=e, mode="plain")])
    assert rel_err(v, ref, 1e-6) < 1e-12
    # the window set afterwards applies to the next trajectory only, as in the reference
    s.set_npoints_trajectory(2001)
    assert len(s.get_trajectory()) == len(traj)
    assert len(s.calculate_trajectory()) != len(traj)
