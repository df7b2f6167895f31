"""Writes tests/golden/beam_sampling.npz with the reference's own random
stream and beam sampler (oracle/_ref/libref.so, built by oracle/Makefile from
the reference sources): raw normal/uniform draws after set_seed for a few
seeds, and the first particles drawn from each beam of tests/common.py's
BEAM_CASES and PREDEFINED_CASES after set_seed(42) (row 0 of each
predefined case is the ideal particle the beam starts with).  Run from the repository root:
python tests/golden/make_beam_golden.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as P  # noqa: E402
from tests.common import BEAM_CASES, PREDEFINED_CASES  # noqa: E402

SEEDS = [0, 1, 42, -7, 2147483647]
PATTERN = "NUUNNUNUUUNNNUN" * 8   # N = normal, U = uniform
N_PARTICLES = 100


def main():
    s = P.RefSim()
    draws = np.zeros((len(SEEDS), len(PATTERN)))
    for i, seed in enumerate(SEEDS):
        s.set_seed(seed)
        for j, c in enumerate(PATTERN):
            draws[i, j] = s.random_normal() if c == "N" else s.random_uniform()
    parts = np.zeros((len(BEAM_CASES), N_PARTICLES, 6))
    for i, kw in enumerate(BEAM_CASES):
        s.set_particle_beam_full(**kw)
        s.set_seed(42)
        for k in range(N_PARTICLES):
            s.new_particle()
            p = s.particle()
            parts[i, k, :3] = p["x0"]
            parts[i, k, 3:] = p["b0"]
    pre = np.zeros((len(PREDEFINED_CASES), N_PARTICLES + 1, 6))
    for i, kw in enumerate(PREDEFINED_CASES):
        s.set_particle_beam_full(**ref_args(kw))
        for k in range(N_PARTICLES + 1):
            if k:
                s.new_particle()
            p = s.particle()
            pre[i, k, :3] = p["x0"]
            pre[i, k, 3:] = p["b0"]
            if k == 0:
                s.set_seed(42)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "beam_sampling.npz"), seeds=np.array(SEEDS),
                        pattern=np.array(PATTERN), draws=draws, particles=parts, predefined=pre)


def ref_args(kw):
    """set_particle_beam keywords of a predefined case -> RefSim.set_particle_beam_full."""
    kw = dict(kw)
    return dict(kw, energy_GeV=kw.pop("energy_GeV", -1.0), d0=kw.get("d0"))


if __name__ == "__main__":
    main()
