"""Writes the output-format fixtures under tests/golden/io/ with the reference's
own container code (oracle/_ref/libref.so, built by oracle/Makefile from the
reference sources): its spectrum and point-value writers (text and float32
binary) on fixed values, and its averages of fixed input files.  Run from the
repository root: python tests/golden/make_io_golden.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as P  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "io")


def values(n, seed):
    """Values that exercise the text format: signs, zeros, tiny and huge
    exponents, 3-digit exponents, numbers on a rounding boundary."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
    fixed = [0.0, -0.0, 1.0, -1.0, 1e-300, -2.5e300, 1.2345675e-5, 9.9999995e9, 123456.5, 0.1, 1 / 3.0]
    v[:len(fixed)] = fixed
    return v


def main():
    os.makedirs(OUT, exist_ok=True)
    e = np.linspace(100.0, 2000.0, 50)
    P.ref_write_spectrum(os.path.join(OUT, "spectrum.txt"), os.path.join(OUT, "spectrum.dat"), e, values(50, 1))
    xyz = np.column_stack([np.linspace(-0.01, 0.01, 40), np.linspace(0.02, -0.02, 40), np.full(40, 30.0)])
    for dim in (2, 3):
        P.ref_write_points(os.path.join(OUT, "points%d.txt" % dim), os.path.join(OUT, "points%d.dat" % dim), xyz,
                           values(40, 2 + dim), dim)
    # inputs for averaging: three spectra and three point sets (text and binary)
    for k in range(3):
        P.ref_write_spectrum(os.path.join(OUT, "avg_in%d.txt" % k), os.path.join(OUT, "avg_in%d.dat" % k), e,
                             np.abs(values(50, 10 + k)))
        P.ref_write_points(os.path.join(OUT, "avgp_in%d.txt" % k), os.path.join(OUT, "avgp_in%d.dat" % k), xyz,
                           np.abs(values(40, 20 + k)), 3)
    ins = [os.path.join(OUT, "avg_in%d" % k) for k in range(3)]
    P.ref_average_spectra([p + ".txt" for p in ins], 0, os.path.join(OUT, "avg_out_text.txt"),
                          os.path.join(OUT, "avg_out_text.dat"))
    P.ref_average_spectra([p + ".dat" for p in ins], 1, os.path.join(OUT, "avg_out_bin.txt"),
                          os.path.join(OUT, "avg_out_bin.dat"))
    insp = [os.path.join(OUT, "avgp_in%d" % k) for k in range(3)]
    P.ref_average_points([p + ".txt" for p in insp], 0, 3, os.path.join(OUT, "avgp_out_text.txt"),
                         os.path.join(OUT, "avgp_out_text.dat"))
    P.ref_average_points([p + ".dat" for p in insp], 1, 3, os.path.join(OUT, "avgp_out_bin.txt"),
                         os.path.join(OUT, "avgp_out_bin.dat"))
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
