"""Output files and averaging of the reference's result containers.

Restated from (paths relative to the reference root):
  spectrum text / binary writers     src/TSpectrumContainer.cc:282-350
  spectrum averaging (text, binary,  src/TSpectrumContainer.cc:372-610 and the
  in-memory spectra)                 Python wrapper src/OSCARSSR_Python.cc:8460-8598
  point-value (flux, power density)  src/T3DScalarContainer.cc:396-470
  text / binary writers
  point-value averaging              src/T3DScalarContainer.cc:104-380 and
                                     src/OSCARSSR_Python.cc:8694-8790

Formats: text is C++ std::scientific at the default precision (6 digits after
the point), one point per line, fields separated by one space; binary is the
values cast to float32, native byte order, with no header.  Averages keep the
reference's order of operations, including its compensated running sum, so
averaged values equal the reference's to the last bit.
"""
import math

import numpy as np


def _fmt(x):
    # as a C++ stream prints: a NaN keeps its sign ("-nan" for the default NaN of
    # an invalid operation on x86)
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return "%.6e" % x


# ---------------------------------------------------------------- spectra
def write_spectrum_text(path, energies, values, header=""):
    """TSpectrumContainer::WriteToFileText: optional header line, then 'E V' per line."""
    try:
        f = open(path, "w")
    except OSError:
        raise ValueError("cannot open file for writing")
    with f:
        if header != "":
            f.write(header + "\n")
        for e, v in zip(energies, values):
            f.write(_fmt(e) + " " + _fmt(v) + "\n")


def write_spectrum_binary(path, energies, values):
    """TSpectrumContainer::WriteToFileBinary: (float32 E, float32 V) pairs."""
    a = np.empty((len(energies), 2), dtype=np.float32)
    with np.errstate(over="ignore"):  # beyond the float32 range: inf, as the C++ cast gives
        a[:, 0] = np.asarray(energies, dtype=np.float64)
        a[:, 1] = np.asarray(values, dtype=np.float64)
    try:
        f = open(path, "wb")
    except OSError:
        raise ValueError("cannot open file for binary write")
    with f:
        f.write(a.tobytes())


class _Compensated:
    """Values with the reference containers' compensated running sum
    (AddToFlux / AddToPoint: y = v - c; t = s + y; c = (t - s) - y; s = t)."""

    def __init__(self):
        self.x = []
        self.v = []
        self.c = []

    def add_point(self, x, v):
        self.x.append(x)
        self.v.append(v)
        self.c.append(0.0)

    def add_to(self, i, v):
        s = self.v[i]
        y = v - self.c[i]
        t = s + y
        self.c[i] = (t - s) - y
        self.v[i] = t


def _text_records(path, width):
    """Whitespace-separated numbers of a text file, `width` per record; reading
    stops at the first incomplete or unparsable record (as istream >> does)."""
    try:
        toks = open(path).read().split()
    except OSError:
        return None
    out = []
    for i in range(0, len(toks) - width + 1, width):
        try:
            out.append([float(t) for t in toks[i:i + width]])
        except ValueError:
            break
    return out


def _binary_records(path, width):
    data = open(path, "rb").read()
    n = len(data) // (4 * width)
    a = np.frombuffer(data[:n * 4 * width], dtype=np.float32).reshape(n, width)
    return [[float(x) for x in row] for row in a]


def _average(records, weights, err_len):
    """Point-by-point weighted sum over inputs: points come from input 0 and
    the loop ends when input 0 runs out; a later input that runs out first is
    an error (the reference's lock-step read)."""
    acc = _Compensated()
    n0 = len(records[0])
    for ip in range(n0):
        for i, rec in enumerate(records):
            if ip >= len(rec):
                raise ValueError(err_len(i))
            row = rec[ip]
            if i == 0:
                acc.add_point(row[:-1], row[-1] * weights[i])
            else:
                acc.add_to(ip, row[-1] * weights[i])
    return acc


def average_spectra(ifiles=None, bifiles=None, spectra=None, ofile="", bofile="", weights=None):
    """sr.average_spectra: average spectra from text files, binary files or
    in-memory [[E, V], ...] lists; returns [[E, V], ...] and optionally writes
    the result (src/OSCARSSR_Python.cc:8460-8598)."""
    ifiles = list(ifiles or [])
    bifiles = list(bifiles or [])
    spectra = list(spectra or [])
    weights = list(weights or [])
    nt, nb, ns, nw = len(ifiles), len(bifiles), len(spectra), len(weights)
    if nt != 0 and nb != 0 and ns != 0:
        raise ValueError("either text or binary or spectra may be added, but not a combination.")
    if nt + nb + ns < 1:
        raise ValueError("No files/spectra given.  You need at least one as input in a list.")
    if nw != 0 and nt + nb + ns != nw:
        raise ValueError("The number of weights does not match the number of inputs (nor is it zero).")
    if nt > 0 and nb > 0:
        raise ValueError("Currently adding mixed types of binary and text files is not supported.")
    if nt > 0 or nb > 0:
        files = ifiles if nt > 0 else bifiles
        # the wrapper passes no weights to the file readers: each file has 1/N
        w = [1.0 / float(len(files))] * len(files)
        recs = []
        for p in files:
            r = _text_records(p, 2) if nt > 0 else (_binary_records(p, 2) if _exists(p) else None)
            if r is None:
                raise ValueError("Cannot open one or more files of input")
            recs.append(r)
        acc = _average(recs, w, lambda i: "files are not the same length")
    else:
        n0 = len(spectra[0])
        for s in spectra:
            if len(s) != n0:
                raise ValueError("Incorrect size in one of the spectra")
        w = weights if nw != 0 else [1.0 / float(ns)] * ns
        acc = _Compensated()
        for k, s in enumerate(spectra):
            for ip in range(n0):
                x = float(s[ip][0])
                y = float(s[ip][1]) * w[k]
                if k == 0:
                    acc.add_point([x], y)
                else:
                    acc.add_to(ip, y)
    e = [x[0] for x in acc.x]
    if ofile != "":
        write_spectrum_text(ofile, e, acc.v)
    if bofile != "":
        write_spectrum_binary(bofile, e, acc.v)
    return [[e[i], acc.v[i]] for i in range(len(e))]


def _exists(p):
    try:
        open(p, "rb").close()
        return True
    except OSError:
        return False


# ---------------------------------------------------------------- point values (flux, power density)
def _xyz(points):
    out = []
    for p in points:
        p = list(p)
        out.append(p + [0.0] * (3 - len(p)))
    return out


def write_points_text(path, points, values, dim=2):
    """T3DScalarContainer::WriteToFileText: 'X Y V' (dim 2) or 'X Y Z V' (dim 3)
    per line; any other dim is an error once there is a point to write."""
    try:
        f = open(path, "w")
    except OSError:
        raise ValueError("cannot open output file")
    with f:
        for p, v in zip(_xyz(points), values):
            if dim not in (2, 3):
                raise ValueError("incorrect dimensions")
            if dim == 2:
                f.write(_fmt(p[0]) + " " + _fmt(p[1]) + " " + _fmt(v) + "\n")
            else:
                f.write(_fmt(p[0]) + " " + _fmt(p[1]) + " " + _fmt(p[2]) + " " + _fmt(v) + "\n")


def write_points_binary(path, points, values, dim=2):
    """T3DScalarContainer::WriteToFileBinary: float32 (X, Y, V) or (X, Y, Z, V) records."""
    if dim not in (2, 3):
        try:
            open(path, "wb").close()
        except OSError:
            raise ValueError("cannot open out file named: " + path)
        raise ValueError("incorrect dimensions")
    p = np.asarray(_xyz(points), dtype=np.float64).reshape(-1, 3)
    a = np.empty((len(p), dim + 1), dtype=np.float32)
    with np.errstate(over="ignore"):  # beyond the float32 range: inf, as the C++ cast gives
        a[:, 0] = p[:, 0]
        a[:, 1] = p[:, 1]
        if dim == 3:
            a[:, 2] = p[:, 2]
        a[:, dim] = np.asarray(values, dtype=np.float64)
    try:
        f = open(path, "wb")
    except OSError:
        raise ValueError("cannot open out file named: " + path)
    with f:
        f.write(a.tobytes())


def average_points(ifiles=None, bifiles=None, ofile="", bofile="", dim=2):
    """sr.average_flux / sr.average_power_density: average point-value files
    (same points, same order); returns [[[x, y, z], v], ...]
    (src/OSCARSSR_Python.cc:8694-8790, src/T3DScalarContainer.cc:104-380)."""
    ifiles = list(ifiles or [])
    bifiles = list(bifiles or [])
    if len(ifiles) != 0 and len(bifiles) != 0:
        raise ValueError("either text or binary files may be added, but not both.")
    if len(ifiles) + len(bifiles) < 1:
        raise ValueError("No files given.  You need at least one file as input in a list.")
    files = ifiles or bifiles
    n = float(len(files))
    if ifiles:
        if dim not in (2, 3):
            raise ValueError("'dim' must be 2 or 3")
        # text inputs are not checked for being open: an unreadable file reads as empty
        recs = [(_text_records(p, dim + 1) or []) for p in files]
        err = lambda i: "Inconsistent file format found in file " + files[i]  # noqa: E731
    else:
        for p in files:
            if not _exists(p):
                raise ValueError("Cannot open input file: " + p)
        if dim not in (2, 3):
            raise ValueError("T3DScalarContainer::AverageFromFilesBinary I do not know how to calculate in this "
                             "dimension")
        recs = [_binary_records(p, dim + 1) for p in files]
        err = lambda i: ("T3DScalarContainer::AverageFromFilesBinary index here must be 0%s  Please report this."  # noqa
                         % ("." if dim == 2 else ".."))
    acc = _Compensated()
    for ip in range(len(recs[0])):
        for i, rec in enumerate(recs):
            if ip >= len(rec):
                raise ValueError(err(i))
            row = rec[ip]
            v = row[-1] / n
            if i == 0:
                xyz = row[:-1] + ([0.0] if dim == 2 else [])
                acc.add_point(xyz, v)
            else:
                acc.add_to(ip, v)
    if ofile != "":
        write_points_text(ofile, acc.x, acc.v, dim)
    if bofile != "":
        write_points_binary(bofile, acc.x, acc.v, dim)
    return [[list(acc.x[i]), acc.v[i]] for i in range(len(acc.x))]



# ---------------------------------------------------------------- trajectories
# src/TParticleTrajectoryPoints.cc:259-818.  Columns: T [s], X Y Z [m], BX BY BZ
# (beta), BPX BPY BPZ (a/c [1/s]).
TRAJ_DEFAULT_FORMAT = "T X Y Z BX BY BZ BPX BPY BPZ"
_TRAJ_COLUMN = {"T": 0, "X": 1, "Y": 2, "Z": 3, "BX": 4, "BY": 5, "BZ": 6, "BPX": 7, "BPY": 8, "BPZ": 9}


def _fmt35(x):
    # std::scientific at precision 35
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return "%.35e" % x


def _traj_rows(t, p9):
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    p9 = np.asarray(p9, dtype=np.float64).reshape(-1, 9)
    return np.column_stack([t, p9])


def _format_words(fmt):
    words = fmt.upper().split()
    if len(words) < 1:
        raise ValueError("Format must contain at least one element")
    return words


def write_trajectory_text(path, t, p9, fmt=""):
    """TParticleTrajectoryPoints::WriteToFile: '# <format>' then one line per
    knot, values in std::scientific with 35 digits after the point."""
    try:
        f = open(path, "w")
    except OSError:
        raise ValueError("cannot open output file")
    rows = _traj_rows(t, p9)
    with f:
        if fmt.upper() == "":
            f.write("# " + TRAJ_DEFAULT_FORMAT + "\n")
            for r in rows:
                f.write(" ".join(_fmt35(v) for v in r) + "\n")
            return
        words = _format_words(fmt)
        f.write("# " + fmt + "\n")
        for r in rows:
            vals = []
            for w in words:
                if w not in _TRAJ_COLUMN:
                    raise ValueError("format specifier not recognized")
                vals.append(_fmt35(r[_TRAJ_COLUMN[w]]))
            f.write(" ".join(vals) + "\n")


def write_trajectory_binary(path, t, p9, fmt=""):
    """TParticleTrajectoryPoints::WriteToFileBinary: int32 length and the
    upper-case format string, then float64 values knot by knot."""
    try:
        f = open(path, "wb")
    except OSError:
        raise ValueError("cannot open output file")
    rows = _traj_rows(t, p9)
    with f:
        form = fmt.upper() or TRAJ_DEFAULT_FORMAT
        if fmt.upper() != "":
            words = _format_words(fmt)
        else:
            words = TRAJ_DEFAULT_FORMAT.split()
        f.write(np.array([len(form)], dtype=np.int32).tobytes())
        f.write(form.encode())
        if len(rows) == 0:
            return
        for w in words:
            if w not in _TRAJ_COLUMN:
                raise ValueError("format specifier not recognized")
        cols = [_TRAJ_COLUMN[w] for w in words]
        f.write(np.ascontiguousarray(rows[:, cols]).tobytes())


_CNUM = None


def _c_double(tok):
    """A whole token read as istream >> double reads it (no 'inf'/'nan',
    no underscores), else None."""
    global _CNUM
    if _CNUM is None:
        import re
        _CNUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")
    return float(tok) if _CNUM.match(tok) else None


def _traj_has(words):
    has = lambda *k: any(w in words for w in k)  # noqa: E731
    if not (has("T") and has("X", "Y", "Z")):
        raise ValueError("trajectory must contain time and position in 1d at a minimum")
    return has("BX", "BY", "BZ"), has("BPX", "BPY", "BPZ")


def _store(words, vals, row):
    for w, v in zip(words, vals):
        if w in _TRAJ_COLUMN:
            row[_TRAJ_COLUMN[w]] = v


def read_trajectory_text(path, fmt=""):
    """TParticleTrajectoryPoints::ReadFromFileFormat.  A first line that
    starts with '#' and names T and one of X/Y/Z gives the format (unless fmt
    is given); any other '#' line is a comment; columns named '*' are skipped.
    Returns (t[n], p9[n, 9], has_beta, has_beta_prime)."""
    try:
        lines = open(path).read().split("\n")
    except OSError:
        raise ValueError("cannot open input file specified")
    if lines and lines[-1] == "":
        lines.pop()
    found = ""
    head = lines[0].upper().split() if lines else []
    if head and head[0][0] != "#":
        start = 0
    else:
        start = 1
        if head:
            head = [head[0][1:]] + head[1:] if len(head[0]) != 1 else head[1:]
            if "T" in head and any(k in head for k in ("X", "Y", "Z")):
                found = "".join(" " + h for h in head)
    form = fmt.upper()
    if fmt == "" and found != "":
        form = found
    form = form.upper() or TRAJ_DEFAULT_FORMAT
    words = form.split()
    has_b, has_bp = _traj_has(words)
    rows = []
    for line in lines[start:]:
        toks = line.split()
        vals = []
        it = 0
        for w in words:
            if w == "*":
                it += 1
                vals.append(0.0)
                continue
            v = _c_double(toks[it]) if it < len(toks) else None
            if v is None:
                raise ValueError("When parsing the input trajectory file the format specified does not match the "
                                 "data in length")
            it += 1
            vals.append(v)
        row = [0.0] * 10
        _store(words, vals, row)
        if math.sqrt(row[4] * row[4] + row[5] * row[5] + row[6] * row[6]) >= 1.0:
            raise ValueError("Magnitude of beta >= 1 in input file.  Sorry, this is beyond einstein.  Line:\n" + line)
        rows.append(row)
    a = np.array(rows, dtype=np.float64).reshape(-1, 10)
    return a[:, 0].copy(), a[:, 1:].copy(), has_b, has_bp


def read_trajectory_binary(path, fmt=""):
    """TParticleTrajectoryPoints::ReadFromFileBinary.  Without fmt the file
    starts with an int32 length and the format string; with fmt the file is
    read as data from its first byte.  A last incomplete record is dropped."""
    try:
        data = open(path, "rb").read()
    except OSError:
        raise ValueError("cannot open input file: " + path)
    off = 0
    form = fmt
    if fmt == "":
        n = int(np.frombuffer(data[:4], dtype=np.int32)[0]) if len(data) >= 4 else 0
        if n < 1:
            raise ValueError("Trying to read binary file format from file header failed.  Incorrect format.")
        form = data[4:4 + n].decode("latin-1")
        off = 4 + n
    words = _format_words(form)
    has_b, has_bp = _traj_has(words)
    nrec = (len(data) - off) // (8 * len(words))
    vals = np.frombuffer(data[off:off + 8 * len(words) * nrec], dtype=np.float64).reshape(nrec, len(words))
    a = np.zeros((nrec, 10))
    for k, w in enumerate(words):
        if w in _TRAJ_COLUMN:
            a[:, _TRAJ_COLUMN[w]] = vals[:, k]
    for i in range(nrec):
        if math.sqrt(a[i, 4] * a[i, 4] + a[i, 5] * a[i, 5] + a[i, 6] * a[i, 6]) >= 1.0:
            raise ValueError("Magnitude of beta >= 1 in input file.  Sorry, this is beyond einstein.  ipoint: %d" % i)
    return a[:, 0].copy(), a[:, 1:].copy(), has_b, has_bp


# ---------------------------------------------------------------- STL surfaces
# Binary STL (80-byte header, uint32 count, then per triangle float32 normal,
# three float32 vertices and a 2-byte attribute), as the reference reads and
# writes it (src/TTriangle3DContainer.cc:110-228, src/TSTLContainer.cc:57-113).
_STL_REC = np.dtype([("n", "<f4", 3), ("a", "<f4", 3), ("b", "<f4", 3), ("c", "<f4", 3), ("s", "<i2")])


def read_stl(path, scale=1.0, translation=(0.0, 0.0, 0.0)):
    """Triangles of a binary STL file as the reference stores them:
    tri[n, 15] = A, B, C (vertices times scale, then translated), the normal
    as written in the file, and the centre ((A + B + C) / 3 before the
    translation, then translated).  The reference ignores rotations here."""
    try:
        data = open(path, "rb").read()
    except OSError:
        raise ValueError("input file cannot be opened")
    n = int(np.frombuffer(data[80:84], dtype="<u4")[0]) if len(data) >= 84 else 0
    rec = np.frombuffer(data[84:84 + 50 * n], dtype=_STL_REC, count=n)
    s = float(scale)
    a = rec["a"].astype(np.float64) * s
    b = rec["b"].astype(np.float64) * s
    c = rec["c"].astype(np.float64) * s
    centre = (a + b + c) / 3.
    t = np.asarray(translation, dtype=np.float64).reshape(1, 3)
    return np.concatenate([a + t, b + t, c + t, rec["n"].astype(np.float64), centre + t], axis=1)


def write_stl(path, tri, scales):
    """TSTLContainer::WriteSTLFile: the header, the count, then each triangle
    with its vertices multiplied by its file's scale (again) as float32 and
    a zero attribute.  scales[k] is the scale of triangle k's file."""
    tri = np.asarray(tri, dtype=np.float64).reshape(-1, 15)
    try:
        f = open(path, "wb")
    except OSError:
        raise ValueError("output file cannot be opened")
    rec = np.zeros(len(tri), dtype=_STL_REC)
    sc = np.asarray(scales, dtype=np.float64).reshape(-1, 1)
    with np.errstate(over="ignore"):
        rec["n"] = tri[:, 9:12]
        rec["a"] = tri[:, 0:3] * sc
        rec["b"] = tri[:, 3:6] * sc
        rec["c"] = tri[:, 6:9] * sc
    with f:
        head = b"OSCARS OSCARS OSCARS OSCARS OSCARS   "
        f.write(head + b"\x00" * (80 - len(head)))
        f.write(np.array([len(tri)], dtype="<u4").tobytes())
        f.write(rec.tobytes())
