"""Write the golden fixtures in this directory from the reference's own C++ core.

The reference library is oracle/_ref/libref.so, compiled from the reference
sources by oracle/Makefile (``make ref``); this script only runs it and stores
inputs and full-double outputs (in-memory values, never the reference's
6-digit text or float32 binary writers) as small .npz files:

  planar_traj.npz     trajectory knots / spline second derivatives (every 50th knot)
                      and level-10 samples (every 8th) for the planar undulator
  planar_spectrum.npz 37-point spectrum at (0,0,30), polarizations all/lh/cl, flux and level
  planar_flux.npz     11x9 flux map at 152.11 eV and 2000 eV, +-10 mm at 30 m
  epu_power.npz       11x9 power density of the EPU at 20 m, normal 0 and 1
  grid_flux.npz       1-D By(z) grid map (401 samples): trajectory knots and a 7x5 flux map

Run:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import RefSim  # noqa: E402
from tests.common import planar, epu, UNDULATOR  # noqa: E402


def grid_file(path):
    """A 1-D By(z) map from the ideal undulator field sampled on a coarse grid."""
    from oracle.pyoracle import Oracle
    o = Oracle()
    o.add_bfield_undulator(**UNDULATOR)
    n = 401
    z = np.array([-1 + 0.005 * i for i in range(n)])
    with open(path, "w") as f:
        f.write("# grid map for the golden fixture\n")
        f.write("0 # x start\n0 # x step\n1 # nx\n0 # y start\n0 # y step\n1 # ny\n")
        f.write("-1 # z start\n0.005 # z step\n%d # nz\n" % n)
        for zi in z:
            b = o.get_bfield([0, 0, zi])
            f.write("%.17g %.17g %.17g\n" % (b[0], b[1], b[2]))
    return n


def main():
    r = RefSim(nthreads=4)
    planar(r)
    t, p, ypp = r.calculate_trajectory()
    lev10 = r.level(10)
    np.savez_compressed(os.path.join(HERE, "planar_traj.npz"), t=t[::50], p=p[::50], ypp=ypp[::50],
                        idx=np.arange(0, len(t), 50), n=len(t), level10=lev10[::8], level10_idx=np.arange(0, 1024, 8))
    out = {}
    for pol in ("", "lh", "cl"):
        for q in ("flux", "level"):
            e, v, conv = r.calculate_spectrum([0, 0, 30], npoints=37, energy_range_eV=[100, 2000], polarization=pol,
                                              quantity=q, nthreads=1)
            out["%s_%s" % (pol or "all", q)] = v
            out["energies"] = e
    np.savez_compressed(os.path.join(HERE, "planar_spectrum.npz"), **out)
    out = {}
    for energy in (152.11, 2000.0):
        xyz, v, conv = r.calculate_flux_rectangle(energy, [11, 9], [0.02, 0.02], translation=[0, 0, 30], nthreads=1)
        out["flux_%g" % energy] = v
        out["xyz"] = xyz
    np.savez_compressed(os.path.join(HERE, "planar_flux.npz"), **out)

    r2 = RefSim(nthreads=4)
    epu(r2)
    out = {}
    for normal in (0, 1):
        xyz, v, conv = r2.calculate_power_density_rectangle([11, 9], [0.08, 0.08], translation=[0, 0, 20],
                                                            normal=normal, nthreads=1)
        out["pd_normal%d" % normal] = v
        out["xyz"] = xyz
    np.savez_compressed(os.path.join(HERE, "epu_power.npz"), **out)

    path = os.path.join(HERE, "grid_field.txt")
    grid_file(path)
    r3 = RefSim(nthreads=4)
    r3.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    r3.add_bfield_file(path)
    r3.set_ctstartstop(0, 2)
    r3.set_npoints_trajectory(4001)
    t, p, ypp = r3.calculate_trajectory()
    xyz, v, conv = r3.calculate_flux_rectangle(150.0, [7, 5], [0.01, 0.01], translation=[0, 0, 30], nthreads=1)
    np.savez_compressed(os.path.join(HERE, "grid_flux.npz"), t=t[::20], p=p[::20], n=len(t), flux=v, xyz=xyz)
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
