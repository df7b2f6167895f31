"""Writes tests/golden/user_trajectory.npz with the reference build
(oracle/_ref/libref.so, built by oracle/Makefile from the reference sources):
the planar-undulator trajectory of tests/common.py (2001 points) thinned to
uneven time steps (every knot with index % 3 != 1: steps of dt and 2 dt),
handed to the reference as a 'T X Y Z' text file, so that it rebuilds beta and
a/c from spline slopes and splines the result; then its knots, spline second
derivatives (every 5th knot), a spectrum and a power-density map from that
trajectory.
Run from the repository root: python tests/golden/make_traj_golden.py"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as P  # noqa: E402
from tests.common import planar  # noqa: E402

ENERGIES = np.linspace(150.0, 600.0, 24)
PD_OBS = [[x, y, 30.0] for x in (-0.006, -0.002, 0.0, 0.003) for y in (-0.002, 0.0, 0.001)]


def thinned_trajectory():
    s = P.RefSim()
    planar(s, npoints=2001)
    t, p9, _ = s.calculate_trajectory()
    keep = np.array([i for i in range(len(t)) if i % 3 != 1])
    return t[keep], p9[keep]


def main():
    t, p9 = thinned_trajectory()
    s = P.RefSim()
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.set_trajectory_points(t, p9)  # as given: the writer below keeps T X Y Z only
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "txyz.txt")
        s.write_trajectory(path, None, "T X Y Z")
        s.read_trajectory(path)
    tk, pk, ypp = s.trajectory()
    _, spec, _ = s.calculate_spectrum([0, 0, 30], energy_points_eV=ENERGIES)
    _, pd, _ = s.calculate_power_density_rectangle([4, 3], [0.009, 0.003], translation=[-0.0015, -0.0005, 30])
    every = np.arange(0, len(tk), 5)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "user_trajectory.npz"), t_in=t, x_in=p9[:, 0:3],
                        t=tk, check=every, p9=pk[every], ypp=ypp[every], energies=ENERGIES, spectrum=spec, pd=pd)


if __name__ == "__main__":
    main()
