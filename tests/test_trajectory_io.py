"""Trajectory files (calculate_trajectory's ofile/bofile, set_trajectory's
ifile/bifile) against the reference's own TParticleTrajectoryPoints code:
writers byte for byte, readers value for value, on trajectories from the
reference build, and the reader's rules (format from a '#' header, '*'
columns, format override, binary header, |beta| >= 1).  The construction of
beta and a/c from spline slopes and the spline of a given trajectory run on
the GPU (tests/test_gpu_parity.py)."""
import os
import tempfile

import numpy as np
import pytest

from oscars_amd import io as IO
from tests.common import planar


def _ref():
    from oracle import pyoracle as P
    if not P.ref_available():
        pytest.skip("reference build (oracle/_ref/libref.so) not present")
    return P


def _bytes(p):
    return open(p, "rb").read()


@pytest.fixture(scope="module")
def ref_traj():
    P = _ref()
    s = P.RefSim()
    planar(s, npoints=1001)
    t, p9, _ = s.calculate_trajectory()
    return s, t, p9


@pytest.mark.parametrize("fmt", ["", "T X Y Z", "t z bx by bpx BPZ", "X T"])
def test_writers_match_reference(ref_traj, fmt):
    s, t, p9 = ref_traj
    with tempfile.TemporaryDirectory() as d:
        j = lambda n: os.path.join(d, n)  # noqa: E731
        s.write_trajectory(j("r.txt"), j("r.dat"), fmt)
        IO.write_trajectory_text(j("o.txt"), t, p9, fmt)
        IO.write_trajectory_binary(j("o.dat"), t, p9, fmt)
        assert _bytes(j("r.txt")) == _bytes(j("o.txt"))
        assert _bytes(j("r.dat")) == _bytes(j("o.dat"))


def test_writer_rejects_unknown_column(ref_traj):
    s, t, p9 = ref_traj
    with tempfile.TemporaryDirectory() as d:
        with pytest.raises(ValueError, match="format specifier not recognized"):
            IO.write_trajectory_text(os.path.join(d, "o.txt"), t, p9, "T Q")
        with pytest.raises(ValueError, match="at least one element"):
            IO.write_trajectory_binary(os.path.join(d, "o.dat"), t, p9, "   ")


@pytest.mark.parametrize("fmt", ["", "T X Y Z BX BY BZ BPX BPY BPZ", "t x y z bx by bz"])
def test_readers_match_reference(ref_traj, fmt):
    """Files written in each format, read back by both (the reference rebuilds
    missing columns; the columns the file holds must agree bit for bit)."""
    P = _ref()
    s, t, p9 = ref_traj
    cols = (fmt.upper() or IO.TRAJ_DEFAULT_FORMAT).split()
    with tempfile.TemporaryDirectory() as d:
        for binary in (False, True):
            path = os.path.join(d, "in.dat" if binary else "in.txt")
            (IO.write_trajectory_binary if binary else IO.write_trajectory_text)(path, t, p9, fmt)
            r = P.RefSim()
            r.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
            r.read_trajectory(path, "", binary)
            rt, rp, _ = r.trajectory()
            ot, op, has_b, has_bp = (IO.read_trajectory_binary if binary else IO.read_trajectory_text)(path)
            assert has_b == ("BX" in cols) and has_bp == ("BPX" in cols)
            assert np.array_equal(ot, rt)
            held = [IO._TRAJ_COLUMN[c] - 1 for c in cols if c != "T"]
            assert np.array_equal(op[:, held], rp[:, held])


def test_reader_rules():
    """'#' line without T and a position is a comment; '*' skips a column; an
    explicit format wins over the file's; a binary file read with a format
    has no header; beta >= 1 is refused."""
    P = _ref()
    with tempfile.TemporaryDirectory() as d:
        j = lambda n: os.path.join(d, n)  # noqa: E731
        rows = [[1e-9 * i, 0.001 * i, 99.0, 0.1 * i, 0.5] for i in range(6)]
        with open(j("a.txt"), "w") as f:
            f.write("# just a comment\n")
            for r in rows:
                f.write("%.17g %.17g %.17g %.17g %.17g\n" % tuple(r))
        ot, op, hb, hbp = IO.read_trajectory_text(j("a.txt"), "T X * Z BZ")
        assert np.array_equal(ot, [r[0] for r in rows])
        assert np.array_equal(op[:, 0], [r[1] for r in rows]) and np.array_equal(op[:, 2], [r[3] for r in rows])
        assert np.all(op[:, 1] == 0) and np.array_equal(op[:, 5], [0.5] * 6) and hb and not hbp
        r = P.RefSim()
        r.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
        r.read_trajectory(j("a.txt"), "T X * Z BZ")
        rt, rp, _ = r.trajectory()
        assert np.array_equal(rt, ot) and np.array_equal(rp[:, [0, 1, 2, 5]], op[:, [0, 1, 2, 5]])
        # header format and override
        with open(j("b.txt"), "w") as f:
            f.write("#T Z X\n")
            for r_ in rows:
                f.write("%.17g %.17g %.17g\n" % (r_[0], r_[1], r_[3]))
        ot, op, _, _ = IO.read_trajectory_text(j("b.txt"))
        assert np.array_equal(op[:, 2], [r_[1] for r_ in rows]) and np.array_equal(op[:, 0], [r_[3] for r_ in rows])
        ot, op, _, _ = IO.read_trajectory_text(j("b.txt"), "T X Z")
        assert np.array_equal(op[:, 0], [r_[1] for r_ in rows])
        # binary data read with an explicit format (no header)
        a = np.array([[r_[0], r_[1], r_[3]] for r_ in rows])
        open(j("c.dat"), "wb").write(a.tobytes() + b"\x00" * 7)
        ot, op, _, _ = IO.read_trajectory_binary(j("c.dat"), "T X Z")
        assert len(ot) == 6 and np.array_equal(op[:, 2], a[:, 2])
        with open(j("d.txt"), "w") as f:
            f.write("0 0 0 0 0.6 0 0.8\n")
        with pytest.raises(ValueError, match="beyond einstein"):
            IO.read_trajectory_text(j("d.txt"), "T X Y Z BX BY BZ")
        with pytest.raises(ValueError, match="time and position"):
            IO.read_trajectory_text(j("d.txt"), "T BX BY BZ")
        with pytest.raises(ValueError, match="does not match the data"):
            IO.read_trajectory_text(j("d.txt"), "T X Y Z BX BY BZ BPX")


def test_set_trajectory_argument_errors():
    from oscars_amd.sr import sr
    s = sr()
    with pytest.raises(ValueError, match="Must previously have defined a particle beam"):
        s.set_trajectory(trajectory=[[0, [0, 0, 0]]])
    s.set_particle_beam(energy_GeV=3)
    with pytest.raises(ValueError, match="Cannot specify both text and binary"):
        s.set_trajectory(ifile="a", bifile="b")
    with pytest.raises(ValueError, match="Cannot specify both an input file and a trajectory list"):
        s.set_trajectory(ifile="a", trajectory=[])
    with pytest.raises(ValueError, match="Need to specify where"):
        s.set_trajectory()
    with pytest.raises(ValueError, match="not the correct format"):
        s.set_trajectory(trajectory=[])
    with pytest.raises(ValueError, match="entry number 1"):
        s.set_trajectory(trajectory=[[0, [0, 0, 0], [0, 0, 0.5]], [1e-9, [0, 0, 0.3]]])
    assert s.set_trajectory(ifile="") is None  # a file argument that names nothing does nothing
