"""Beam sampling for nparticles= against the reference's own random stream.

The reference draws particles from a global std::mt19937 through libstdc++'s
normal_distribution and uniform_real_distribution (src/TRandomA.cc), one new
particle per iteration of its multi-particle loops (src/OSCARSSR.cc:1782-1790).
oscars_amd.beam restates both; these tests pin it:
  - the engine against the C++ standard's own check value;
  - raw draws and particles against tests/golden/beam_sampling.npz, written by
    the reference build (tests/golden/make_beam_golden.py), bit for bit;
  - the same live, on more particles, where the reference build is present;
  - sr.set_seed / nparticles= / set_new_particle("random") draw the same
    particles as the reference's set_seed + SetNewParticle loop.
"""
import math
import os

import numpy as np
import pytest

from oscars_amd import beam as B
from oscars_amd import geometry as G
from oscars_amd.sr import sr
from tests.common import BEAM_CASES, PREDEFINED_CASES

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "beam_sampling.npz"))


def _sr_particles(kw, seed, n):
    s = sr()
    s.set_particle_beam(**kw)
    s.set_seed(seed)
    ps = s._particles_for(n)
    return np.array([p.x0 + p.b0 for p in ps])


def test_engine_check_value():
    """The C++ standard fixes mt19937's 10000th output from the default seed
    5489 at 4123659995; it also equals the classic genrand_int32 stream
    (numpy's legacy MT19937 seeding of an integer)."""
    r = B.RefRandom(5489)
    out = [r.u32() for _ in range(10000)]
    assert out[-1] == 4123659995
    r = B.RefRandom(12345)
    assert [r.u32() for _ in range(1500)] == \
        np.random.RandomState(12345).randint(0, 2 ** 32, size=1500, dtype=np.uint64).tolist()


def test_draws_match_reference_fixture():
    for i, seed in enumerate(GOLD["seeds"]):
        r = B.RefRandom(int(seed))
        got = [r.standard_normal() if c == "N" else r.random() for c in str(GOLD["pattern"])]
        assert np.array_equal(np.array(got), GOLD["draws"][i]), seed


def test_uniform_and_normal_forms():
    """What the definitions fix: a uniform is (g1 + g2 2^32) / 2^64 of two engine
    outputs, kept below 1; normals come in pairs from one accepted polar point
    (x, y): first y*m, then x*m, with m = sqrt(-2 ln r2 / r2)."""
    a, b = B.RefRandom(7), B.RefRandom(7)
    for _ in range(100):
        g1, g2 = b.u32(), b.u32()
        assert a.random() == (g1 + g2 * 2.0 ** 32) / 2.0 ** 64
    a, b = B.RefRandom(9), B.RefRandom(9)
    for _ in range(50):
        while True:
            x, y = 2.0 * b.random() - 1.0, 2.0 * b.random() - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        m = math.sqrt(-2 * math.log(r2) / r2)
        assert a.standard_normal() == y * m
        assert a.standard_normal() == x * m
    # a new seed drops the kept second value, as SetSeed recreates the distribution
    B.set_seed(3)
    first = B.stream().standard_normal()
    B.set_seed(3)
    assert B.stream().standard_normal() == first


@pytest.mark.parametrize("case", range(len(BEAM_CASES)))
def test_particles_match_reference_fixture(case):
    got = _sr_particles(BEAM_CASES[case], 42, GOLD["particles"].shape[1])
    assert np.array_equal(got, GOLD["particles"][case])


def _sr_predefined(kw, n):
    s = sr()
    s.set_particle_beam(**kw)
    rows = [s._particle.x0 + s._particle.b0]
    s.set_seed(42)
    rows += [p.x0 + p.b0 for p in s._particles_for(n)]
    return np.array(rows)


@pytest.mark.parametrize("case", range(len(PREDEFINED_CASES)))
def test_predefined_beams_match_reference_fixture(case):
    """Predefined beams (src/TParticleBeam.cc:135-200) with the binding's
    overrides: the ideal particle they start with, then 100 draws."""
    g = GOLD["predefined"][case]
    assert np.array_equal(_sr_predefined(PREDEFINED_CASES[case], g.shape[0] - 1), g)


def test_predefined_beam_errors():
    with pytest.raises(ValueError, match="Error in predefined beam name"):
        sr().set_particle_beam(beam="NSLS-III")


def test_sincos_is_the_library_pair():
    """geometry.sincos returns the C library's sincos, which agrees with
    separate sin and cos to within one unit in the last place."""
    for x in np.random.default_rng(1).uniform(-10, 10, 2000):
        s, c = G.sincos(x)
        assert abs(s - math.sin(x)) <= 2 * math.ulp(math.sin(x))
        assert abs(c - math.cos(x)) <= 2 * math.ulp(math.cos(x))


def test_stream_is_shared_and_continues():
    """One global stream, as the reference's: two sr objects draw in turn
    from it, and nparticles= continues where the last call stopped."""
    kw = BEAM_CASES[0]
    ref = _sr_particles(kw, 11, 30)
    a, b = sr(), sr()
    a.set_particle_beam(**kw)
    b.set_particle_beam(**kw)
    a.set_seed(11)
    got = [a._particles_for(10), b._particles_for(5)]
    b.set_new_particle(particle="random")
    got.append([b._particle])
    got.append(a._particles_for(14))
    got = np.array([p.x0 + p.b0 for g in got for p in g])
    assert np.array_equal(got, ref)
    a.set_new_particle(particle="ideal")
    assert a._particle.x0 == [0.0, 0.0, -1.5]


def test_sample_statistics():
    """Moments of a Gaussian beam at x0 (2e4 particles): sigma_x = sqrt(eps
    beta), sigma_x' = sqrt(eps gamma), energy spread, and the x-x' correlation
    -alpha/sqrt(beta gamma) of the drifted Twiss parameters."""
    kw = dict(energy_GeV=3.0, x0=[0, 0, 0], d0=[0, 0, 1], sigma_energy_GeV=0.0027, emittance=[0.9e-9, 8e-12],
              beta=[1.5, 0.8], alpha=[0.4, 0.0])
    s = sr()
    s.set_particle_beam(**kw)
    s.set_seed(1)
    ps = s._particles_for(20000)
    x = np.array([p.x0[0] for p in ps])
    xp = np.array([p.b0[0] / p.b0[2] for p in ps])
    y = np.array([p.x0[1] for p in ps])
    e = np.array([p.energy_GeV for p in ps])
    g = (1 + 0.4 ** 2) / 1.5
    assert np.std(x) == pytest.approx(math.sqrt(0.9e-9 * 1.5), rel=0.03)
    assert np.std(xp) == pytest.approx(math.sqrt(0.9e-9 * g), rel=0.03)
    assert np.std(y) == pytest.approx(math.sqrt(8e-12 * 0.8), rel=0.03)
    assert np.std(e) == pytest.approx(0.0027, rel=0.03)
    assert np.corrcoef(x, xp)[0, 1] == pytest.approx(-0.4 / math.sqrt(1.5 * g), abs=0.03)


def test_beam_argument_errors():
    s = sr()
    with pytest.raises(ValueError, match="Beta cannot be <= 0"):
        s.set_particle_beam(energy_GeV=3, emittance=[1e-9, 1e-11], beta=[0, 1])
    with pytest.raises(ValueError, match="Unknown beam distribution name"):
        s.set_particle_beam(energy_GeV=3, distribution="flat")
    with pytest.raises(ValueError, match="sigma_energy_GeV"):
        s.set_particle_beam(energy_GeV=3, sigma_energy_GeV=-1)
    s.set_particle_beam(energy_GeV=3, emittance=[1e-9, 1e-11])
    with pytest.raises(ValueError, match="no Twiss parameters"):
        s.set_new_particle(particle="random")


def _ref():
    from oracle import pyoracle as P
    if not P.ref_available():
        pytest.skip("reference build (oracle/_ref/libref.so) not present")
    return P


@pytest.mark.parametrize("case", range(len(BEAM_CASES)))
def test_particles_live_against_reference_build(case):
    P = _ref()
    s = P.RefSim()
    kw = BEAM_CASES[case]
    s.set_particle_beam_full(**kw)
    for seed in (5, 77):
        s.set_seed(seed)
        ref = []
        for _ in range(1000):
            s.new_particle()
            p = s.particle()
            ref.append(p["x0"] + p["b0"])
        assert np.array_equal(_sr_particles(kw, seed, 1000), np.array(ref)), (case, seed)


@pytest.mark.parametrize("case", range(len(PREDEFINED_CASES)))
def test_predefined_live_against_reference_build(case):
    P = _ref()
    from tests.golden.make_beam_golden import ref_args
    s = P.RefSim()
    kw = PREDEFINED_CASES[case]
    s.set_particle_beam_full(**ref_args(kw))
    rows = [s.particle()["x0"] + s.particle()["b0"]]
    s.set_seed(42)
    for _ in range(1000):
        s.new_particle()
        p = s.particle()
        rows.append(p["x0"] + p["b0"])
    assert np.array_equal(_sr_predefined(kw, 1000), np.array(rows))


def test_nparticles_spectrum_matches_reference_build():
    """The reference's own nparticles= spectrum (set_seed, then one new particle
    and trajectory per iteration) equals the oracle summed over the particles
    that sr draws after the same set_seed, with weights 1/N."""
    P = _ref()
    from oracle.pyoracle import Oracle
    from tests.common import UNDULATOR
    kw = BEAM_CASES[0]
    kw = dict(kw, x0=[0, 0, -1])
    r = P.RefSim(nthreads=1)
    r.set_particle_beam_full(**kw)
    r.add_bfield_undulator(**UNDULATOR)
    r.set_ctstartstop(0, 2)
    r.set_npoints_trajectory(6001)
    r.set_seed(5)
    e = np.linspace(140, 600, 40)
    _, vref, _ = r.calculate_spectrum([0, 0, 30], energy_points_eV=e, nparticles=3)
    s = sr()
    s.set_particle_beam(**kw)
    s.set_seed(5)
    o = Oracle(nthreads=1)
    o.set_particle_beam(3.0, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    o.add_bfield_undulator(**UNDULATOR)
    o.set_ctstartstop(0, 2)
    o.set_npoints_trajectory(6001)
    tot = np.zeros(len(e))
    for p in s._particles_for(3):
        o.set_particle(p.x0, p.b0)
        o.calculate_trajectory()
        tot += o.spectrum([0, 0, 30], e)[0] * (1.0 / 3)
    assert np.max(np.abs(tot - vref) / np.abs(vref)) < 1e-14


def test_several_beams_live_against_reference_build():
    """Two beams of weights 1 and 3: each particle first draws its beam by
    weight, then its own offsets, as the reference does draw for draw."""
    P = _ref()
    s = P.RefSim()
    a, b = BEAM_CASES[0], BEAM_CASES[2]
    s.set_particle_beam_full(**a)
    s.add_particle_beam_full(weight=3.0, **b)
    s.set_seed(11)
    ref = []
    for _ in range(600):
        s.new_particle()
        p = s.particle()
        ref.append(p["x0"] + p["b0"])
    o = sr()
    o.set_particle_beam(**a)
    o.add_particle_beam(weight=3.0, **b)
    o.set_seed(11)
    got = np.array([p.x0 + p.b0 for p in o._particles_for(600)])
    assert np.array_equal(got, np.array(ref))
    # the beams are picked about 1:3
    o.set_seed(12)
    picks = np.array([o._beam_index() for _ in range(4000)])
    assert abs(np.mean(picks == 1) - 0.75) < 0.03


def test_beam_accessors():
    o = sr()
    with pytest.raises(ValueError, match="No particle beam defined"):
        o.get_emittance()
    o.set_particle_beam(**BEAM_CASES[1])
    o.add_particle_beam(name="second", weight=2.0, **BEAM_CASES[2])
    assert o.get_emittance(name="_beam0") == BEAM_CASES[1]["emittance"]
    assert o.get_emittance(name="second") == BEAM_CASES[2]["emittance"]
    with pytest.raises(ValueError, match="beam name not in map"):
        o.get_emittance(name="third")
    with pytest.raises(ValueError, match="possibly 'name' already exists"):
        o.add_particle_beam(name="second", **BEAM_CASES[2])
    h = np.array(o.get_beam_horizontal_direction(name="_beam0"))
    v = np.array(o.get_beam_vertical_direction(name="_beam0"))
    u0 = np.array(BEAM_CASES[1]["d0"], dtype=float) / np.linalg.norm(BEAM_CASES[1]["d0"])
    assert abs(np.linalg.norm(h) - 1) < 1e-15 and abs(np.linalg.norm(v) - 1) < 1e-15
    assert abs(h @ v) < 1e-15 and abs(v @ u0) < 1e-15
    assert np.allclose(v, -np.cross(h, u0) / np.linalg.norm(np.cross(h, u0)), rtol=0, atol=1e-16)
    o.set_particle_beam_size(sigma=[1e-5, 1e-6], sigmap=[1e-6, 1e-7])  # accepted, no effect
    o.clear_particle_beams()
    with pytest.raises(ValueError, match="No particle beam defined"):
        o.set_new_particle()


def test_set_twiss_parameters_is_the_set_particle_beam_form():
    """Twiss parameters given afterwards sample the same particles as the same
    parameters given to set_particle_beam (pinned to the reference above)."""
    kw = dict(BEAM_CASES[1])
    tw = {k: kw.pop(k) for k in ("beta", "alpha", "lattice_reference")}
    o = sr()
    o.set_particle_beam(**kw, beta=[9.0, 9.0])
    o.set_twiss_parameters(**tw)
    o.set_seed(3)
    got = np.array([p.x0 + p.b0 for p in o._particles_for(200)])
    assert np.array_equal(got, _sr_particles(BEAM_CASES[1], 3, 200))


def test_set_particle_beam_positional_arguments():
    """set_particle_beam takes add_particle_beam's arguments by position too
    (the binding's 'type, name, energy_GeV, d0, x0, ...' order,
    src/OSCARSSR_Python.cc:3743-3766)."""
    from oscars_amd import sr as SR
    a, b = SR.sr(), SR.sr()
    a.set_particle_beam("electron", "b", 3.0, [0, 0, 1], [0, 0, -1], current=0.5)
    b.set_particle_beam(type="electron", name="b", energy_GeV=3.0, d0=[0, 0, 1], x0=[0, 0, -1], current=0.5)
    assert a._beams[0][0] == b._beams[0][0] == "b"
    assert a._beams[0][1] == b._beams[0][1]


def test_set_twiss_without_effect_draws_nothing():
    """With several beams, a set_twiss_parameters call that names no Twiss
    combination draws no uniform for the beam choice: the reference looks the
    beam up inside a matching branch only."""
    from oscars_amd import beam as Bm
    from oscars_amd import sr as SR
    s = SR.sr()
    s.add_particle_beam(energy_GeV=3.0, x0=[0, 0, -1], name="a")
    s.add_particle_beam(energy_GeV=3.0, x0=[0, 0, -1], name="b", weight=2)
    s.set_seed(7)
    s.set_twiss_parameters(alpha=[0.1, 0.2])  # alpha alone: no combination
    x = Bm.stream().random()
    s.set_seed(7)
    assert Bm.stream().random() == x
    s.set_seed(7)
    s.set_twiss_parameters(beta=[1.0, 2.0])  # a combination: the beam is drawn
    assert Bm.stream().random() != x
