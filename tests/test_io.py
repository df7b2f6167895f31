"""Output files and averaging (oscars_amd/io.py) against the reference's own
container code: byte-identical files and values.

Fixtures in tests/golden/io/ were written by the reference's TSpectrumContainer
and T3DScalarContainer (tests/golden/make_io_golden.py, through
oracle/_ref/libref.so).  Where the reference build is present the same
comparisons also run live on random values."""
import os
import tempfile

import numpy as np
import pytest

from oscars_amd import io as IO

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")


def _bytes(p):
    return open(p, "rb").read()


def _values(n, seed):
    # the generator of tests/golden/make_io_golden.py
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-30, 30, n)
    fixed = [0.0, -0.0, 1.0, -1.0, 1e-300, -2.5e300, 1.2345675e-5, 9.9999995e9, 123456.5, 0.1, 1 / 3.0]
    v[:len(fixed)] = fixed
    return v


XYZ = np.column_stack([np.linspace(-0.01, 0.01, 40), np.linspace(0.02, -0.02, 40), np.full(40, 30.0)])
E = np.linspace(100.0, 2000.0, 50)


def test_spectrum_writers_match_reference_files():
    with tempfile.TemporaryDirectory() as d:
        IO.write_spectrum_text(os.path.join(d, "s.txt"), E, _values(50, 1))
        IO.write_spectrum_binary(os.path.join(d, "s.dat"), E, _values(50, 1))
        assert _bytes(os.path.join(d, "s.txt")) == _bytes(os.path.join(G, "spectrum.txt"))
        assert _bytes(os.path.join(d, "s.dat")) == _bytes(os.path.join(G, "spectrum.dat"))


@pytest.mark.parametrize("dim", [2, 3])
def test_point_writers_match_reference_files(dim):
    v = _values(40, 2 + dim)
    with tempfile.TemporaryDirectory() as d:
        IO.write_points_text(os.path.join(d, "p.txt"), XYZ, v, dim)
        IO.write_points_binary(os.path.join(d, "p.dat"), XYZ, v, dim)
        assert _bytes(os.path.join(d, "p.txt")) == _bytes(os.path.join(G, "points%d.txt" % dim))
        assert _bytes(os.path.join(d, "p.dat")) == _bytes(os.path.join(G, "points%d.dat" % dim))


@pytest.mark.parametrize("kind", ["text", "bin"])
def test_average_spectra_matches_reference(kind):
    ins = [os.path.join(G, "avg_in%d.%s" % (k, "txt" if kind == "text" else "dat")) for k in range(3)]
    with tempfile.TemporaryDirectory() as d:
        o_t, o_b = os.path.join(d, "a.txt"), os.path.join(d, "a.dat")
        kw = dict(ifiles=ins) if kind == "text" else dict(bifiles=ins)
        res = IO.average_spectra(ofile=o_t, bofile=o_b, **kw)
        assert len(res) == 50
        assert _bytes(o_t) == _bytes(os.path.join(G, "avg_out_%s.txt" % kind))
        assert _bytes(o_b) == _bytes(os.path.join(G, "avg_out_%s.dat" % kind))


@pytest.mark.parametrize("kind", ["text", "bin"])
def test_average_points_matches_reference(kind):
    ins = [os.path.join(G, "avgp_in%d.%s" % (k, "txt" if kind == "text" else "dat")) for k in range(3)]
    with tempfile.TemporaryDirectory() as d:
        o_t, o_b = os.path.join(d, "a.txt"), os.path.join(d, "a.dat")
        kw = dict(ifiles=ins) if kind == "text" else dict(bifiles=ins)
        res = IO.average_points(ofile=o_t, bofile=o_b, dim=3, **kw)
        assert len(res) == 40 and len(res[0][0]) == 3
        assert _bytes(o_t) == _bytes(os.path.join(G, "avgp_out_%s.txt" % kind))
        assert _bytes(o_b) == _bytes(os.path.join(G, "avgp_out_%s.dat" % kind))


def test_average_invariants():
    """What the definition fixes: the average of N copies of one file is that
    file's values (N = 1, 2, 4: exact, the weights are powers of two); weights
    in the in-memory form scale linearly; a shorter later input is an error."""
    v = np.abs(np.random.default_rng(5).standard_normal(30)) * 10.0 ** np.arange(-15, 15)  # finite in float32
    e = np.linspace(1.0, 30.0, 30)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "s.dat")
        IO.write_spectrum_binary(p, e, v)
        f32 = v.astype(np.float32).astype(np.float64)
        for n in (1, 2, 4):
            res = IO.average_spectra(bifiles=[p] * n)
            assert np.array_equal(np.array([r[1] for r in res]), f32)
        res = IO.average_spectra(spectra=[[[a, b] for a, b in zip(e, v)]] * 2, weights=[0.25, 0.5])
        assert np.allclose(np.array([r[1] for r in res]), 0.75 * v, rtol=1e-15, atol=0)
        short = os.path.join(d, "short.dat")
        IO.write_spectrum_binary(short, e[:10], v[:10])
        with pytest.raises(ValueError, match="files are not the same length"):
            IO.average_spectra(bifiles=[p, short])
        assert len(IO.average_spectra(bifiles=[short, p])) == 10  # points come from the first input
    with pytest.raises(ValueError, match="No files/spectra given"):
        IO.average_spectra()
    with pytest.raises(ValueError, match="either text or binary files may be added, but not both"):
        IO.average_points(ifiles=["a"], bifiles=["b"])


def _ref():
    from oracle import pyoracle as P
    if not P.ref_available():
        pytest.skip("reference build (oracle/_ref/libref.so) not present")
    return P


def test_writers_live_against_reference_build():
    P = _ref()
    rng = np.random.default_rng(99)
    n = 500
    e = np.sort(rng.uniform(1, 1e4, n))
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
    xyz = rng.standard_normal((n, 3)) * 10.0 ** rng.integers(-5, 5, (n, 3))
    with tempfile.TemporaryDirectory() as d:
        j = lambda *a: os.path.join(d, *a)  # noqa: E731
        P.ref_write_spectrum(j("r.txt"), j("r.dat"), e, v)
        IO.write_spectrum_text(j("o.txt"), e, v)
        IO.write_spectrum_binary(j("o.dat"), e, v)
        assert _bytes(j("r.txt")) == _bytes(j("o.txt"))
        assert _bytes(j("r.dat")) == _bytes(j("o.dat"))
        for dim in (2, 3):
            P.ref_write_points(j("rp.txt"), j("rp.dat"), xyz, v, dim)
            IO.write_points_text(j("op.txt"), xyz, v, dim)
            IO.write_points_binary(j("op.dat"), xyz, v, dim)
            assert _bytes(j("rp.txt")) == _bytes(j("op.txt"))
            assert _bytes(j("rp.dat")) == _bytes(j("op.dat"))
        # averages of five random text files of 2-D points
        files = []
        for k in range(5):
            files.append(j("in%d.txt" % k))
            IO.write_points_text(files[-1], xyz, np.abs(rng.standard_normal(n)) * 1e13, 2)
        xr, vr = P.ref_average_points(files, 0, 2)
        res = IO.average_points(ifiles=files, dim=2)
        assert np.array_equal(np.array([r[1] for r in res]), vr)
        assert np.array_equal(np.array([r[0] for r in res]), xr)
