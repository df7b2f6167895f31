"""Host-side geometry and particle set-up for the drop-in ``sr`` class.

Everything here is scalar IEEE-754 double arithmetic written in the
reference's operation order, with ``math`` (the C library's sin/cos/sqrt), so
observation points, energies, polarization vectors and initial particle
conditions are bit-identical to what the reference computes before it calls
its integrand.  Reference files restated (paths relative to the reference root):

  include/TOSCARS.h:20-76                    constants
  src/TParticleBeam.cc:97-130, 594-598, 667-689  ideal particle of a beam
  src/TParticleA.cc:454-485                  gamma, Q/M/gamma
  src/TSurfacePoints_Rectangle.cc:129-317    rectangle points and 2-D coordinates
  src/TSpectrumContainer.cc:57-91            energy grid
  src/OSCARSSR.cc:1887-1919                  polarization vectors
  src/TVector3D.cc:208-295                   rotations
"""
import ctypes
import ctypes.util
import math

PI = 3.14159265358979323846
C = 2.99792458e8
H = 6.62606876e-34
HBAR = 1.054571596e-34
QE = 1.602176462e-19
ME = 9.10938356e-31
EPSILON0 = 8.854187817E-12
MU0 = 1.2566370614E-6


def kg_to_gev(m):
    return 1e-9 * m * C * C / QE


def gev_to_kg(m):
    return QE * m / (1e-9 * C * C)


def ev_to_angular_frequency(e):
    return e * QE / HBAR


def mag2(v):
    return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]


def unit(v):
    m = math.sqrt(mag2(v))
    return [v[0] / m, v[1] / m, v[2] / m]


def cross(a, v):
    return [a[1] * v[2] - v[1] * a[2], a[2] * v[0] - v[2] * a[0], a[0] * v[1] - v[0] * a[1]]


def orthogonal(v):
    # include/TVector3D.h:132-143
    xx = -v[0] if v[0] < 0.0 else v[0]
    yy = -v[1] if v[1] < 0.0 else v[1]
    zz = -v[2] if v[2] < 0.0 else v[2]
    if xx < yy:
        return [0.0, v[2], -v[1]] if xx < zz else [v[1], -v[0], 0.0]
    return [-v[2], 0.0, v[0]] if yy < zz else [v[1], -v[0], 0.0]


def _libm_sincos():
    try:
        f = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6").sincos
    except (OSError, AttributeError):
        return None
    f.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f.restype = None
    return f


_SINCOS = _libm_sincos()


def sincos(x):
    """(sin x, cos x) as compiled C++ gets them where one function takes both
    of one argument: the compiler merges the two calls into the C library's
    sincos, whose last bit can differ from separate sin and cos calls."""
    if _SINCOS is None:
        return math.sin(x), math.cos(x)
    s, c = ctypes.c_double(), ctypes.c_double()
    _SINCOS(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def rot_xyz(v, r):
    """TVector3D::RotateSelfXYZ: about X, then Y, then Z."""
    x, y, z = float(v[0]), float(v[1]), float(v[2])
    s, c = sincos(r[0])
    y, z = c * y - s * z, s * y + c * z
    s, c = sincos(r[1])
    z, x = c * z - s * x, s * z + c * x
    s, c = sincos(r[2])
    x, y = c * x - s * y, s * x + c * y
    return [x, y, z]


def rot_axis(o, A, vin):
    """TVector3D::RotateSelf(angle, axis)."""
    V = unit(vin)
    sa, ca = sincos(A)
    M = [[ca + V[0] * V[0] * (1 - ca), V[0] * V[1] * (1 - ca) - V[2] * sa, V[0] * V[2] * (1 - ca) + V[1] * sa],
         [V[1] * V[0] * (1 - ca) + V[2] * sa, ca + V[1] * V[1] * (1 - ca), V[1] * V[2] * (1 - ca) - V[0] * sa],
         [V[2] * V[0] * (1 - ca) - V[1] * sa, V[2] * V[1] * (1 - ca) + V[0] * sa, ca + V[2] * V[2] * (1 - ca)]]
    return [M[i][0] * o[0] + M[i][1] * o[1] + M[i][2] * o[2] for i in range(3)]


class Particle:
    """Initial conditions of one particle (the reference's TParticleA)."""

    def __init__(self, q, m, x0, b0, t0=0.0, current=0.0):
        self.q = float(q)
        self.m = float(m)
        self.x0 = [float(v) for v in x0]
        self.t0 = float(t0)
        self.current = float(current)
        self.gamma = 1.0
        self.qomg = 0.0
        self.set_beta(b0)

    def set_beta(self, b0):
        # src/TParticleA.cc:454-485
        self.b0 = [float(v) for v in b0]
        beta2 = mag2(self.b0) if mag2(self.b0) > 0 else 0
        if beta2 == 1:
            return
        self.gamma = 1.0 / math.sqrt(1.0 - beta2) if beta2 != 0 else 1
        if self.m == 0 or self.gamma == 0:
            return
        self.qomg = self.q / self.m / self.gamma


PARTICLE_TYPES = {
    "electron": (-QE, ME), "anti-positron": (-QE, ME),
    "positron": (QE, ME), "anti-electron": (QE, ME),
    "muon": (-QE, gev_to_kg(0.1056583715)), "anti-muon": (QE, gev_to_kg(0.1056583715)),
}


def ideal_particle(ptype, energy_GeV, x0, d0, t0=0.0, current=0.0, charge=0.0, mass=0.0):
    """TParticleBeam constructor + GetNewParticle("ideal")."""
    if ptype == "custom":
        q, m = float(charge), float(mass)
    else:
        if ptype not in PARTICLE_TYPES:
            raise ValueError("particle type not recognized: %s" % ptype)
        q, m = PARTICLE_TYPES[ptype]
    u0 = unit(d0) if mag2(d0) > 0 else [0.0, 0.0, 0.0]
    e0 = m if energy_GeV < kg_to_gev(m) else float(energy_GeV)
    gamma = 1 if e0 / kg_to_gev(m) < 1 else e0 / kg_to_gev(m)
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma)) if gamma != 1 else 0
    p = Particle(q, m, x0, [u0[0] * beta, u0[1] * beta, u0[2] * beta], t0, current)
    p.e0 = e0
    p.u0 = u0
    return p


def rectangle(plane, n1, n2, w1, w2, rotations=(0, 0, 0), translation=(0, 0, 0), normal=0, dim=2):
    """Observation points, the normal and the coordinates the reference stores
    (TSurfacePoints_Rectangle::Init / GetXYZ / GetX1 / GetX2)."""
    P = plane.upper()
    s1 = w1 / (n1 - 1) if n1 > 1 else 0.0
    s2 = w2 / (n2 - 1) if n2 > 1 else 0.0
    st = [0.0, 0.0, 0.0]
    a1 = [0.0, 0.0, 0.0]
    a2 = [0.0, 0.0, 0.0]
    nv = [0.0, 0.0, 0.0]

    def start(i1, i2):
        # which coordinates receive -W1/2 and -W2/2
        if n1 > 1 and n2 > 1:
            st[i1], st[i2] = -w1 / 2., -w2 / 2.
        elif n1 <= 1 and n2 <= 1:
            pass
        elif n1 <= 1:
            st[i2] = -w2 / 2.
        else:
            st[i1] = -w1 / 2.

    axes = {"XY": (0, 1, 2, 1.0), "YX": (1, 0, 2, -1.0), "XZ": (0, 2, 1, -1.0), "ZX": (2, 0, 1, 1.0),
            "YZ": (1, 2, 0, 1.0), "ZY": (2, 1, 0, -1.0)}
    if P not in axes:
        raise ValueError("not a valid surface string: XY YX XZ ZX YZ ZY")
    i1, i2, inorm, sgn = axes[P]
    start(i1, i2)
    a1[i1] = s1
    a2[i2] = s2
    nv[inorm] = sgn
    st = rot_xyz(st, rotations)
    a1 = rot_xyz(a1, rotations)
    a2 = rot_xyz(a2, rotations)
    st = [st[k] + translation[k] for k in range(3)]
    nv = rot_xyz(nv, rotations)
    if normal == -1:
        nv = [nv[0] * -1, nv[1] * -1, nv[2] * -1]
    elif normal not in (0, 1):
        raise ValueError("normal must be -1, 0, or 1")
    m1 = math.sqrt(mag2(a1))
    m2 = math.sqrt(mag2(a2))
    obs = []
    cont = []
    for i in range(n1 * n2):
        ix1 = i // n2
        ix2 = i % n2
        o = [st[k] + a1[k] * float(ix1) + a2[k] * float(ix2) for k in range(3)]
        obs.append(o)
        if dim == 2:
            cont.append([s1 * ix1 - m1 * (n1 - 1) / 2., s2 * ix2 - m2 * (n2 - 1) / 2., 0.0])
        else:
            cont.append(o)
    return obs, nv, cont


def parallelogram(n1, n2, x0, x1, x2, rotations=(0, 0, 0), translation=(0, 0, 0), normal=0, dim=2):
    """Observation points, normal and stored coordinates of a surface given by
    three corners (TSurfacePoints_Rectangle::Init from X0, X1, X2,
    src/TSurfacePoints_Rectangle.cc:76-123; the binding rotates, then
    translates the corners, src/OSCARSSR_Python.cc:6633-6655)."""
    P = []
    for X in (x0, x1, x2):
        X = rot_xyz(X, rotations)
        P.append([X[k] + translation[k] for k in range(3)])
    X0, X1, X2 = P
    d1 = [X1[k] - X0[k] for k in range(3)]
    d2 = [X2[k] - X0[k] for k in range(3)]
    w1, w2 = math.sqrt(mag2(d1)), math.sqrt(mag2(d2))
    s1, s2 = w1 / (n1 - 1), w2 / (n2 - 1)
    a1 = [d1[k] / (n1 - 1) for k in range(3)]
    a2 = [d2[k] / (n2 - 1) for k in range(3)]
    nv = unit(cross(a1, a2))
    if normal == -1:
        nv = [nv[0] * -1, nv[1] * -1, nv[2] * -1]
    elif normal not in (0, 1):
        raise ValueError("fNormal is of the incorrect form.  Must be -1, 0, 1")
    nv = unit(nv)  # each surface point normalises its normal again (src/TSurfacePoint.cc:16)
    m1, m2 = math.sqrt(mag2(a1)), math.sqrt(mag2(a2))
    obs, cont = [], []
    for i in range(n1 * n2):
        ix1, ix2 = i // n2, i % n2
        o = [X0[k] + a1[k] * float(ix1) + a2[k] * float(ix2) for k in range(3)]
        obs.append(o)
        if dim == 2:
            cont.append([s1 * ix1 - m1 * (n1 - 1) / 2., s2 * ix2 - m2 * (n2 - 1) / 2., 0.0])
        else:
            cont.append(o)
    return obs, nv, cont


def energies(n, e0, e1):
    """TSpectrumContainer::Init(N, EFirst, ELast)."""
    if n < 1:
        raise ValueError("no points specified")
    if n == 1:
        return [float(e0)]
    return [e0 + (e1 - e0) / float(n - 1) * float(i) for i in range(n)]


def polarization_code(s):
    s = s.lower().replace(" ", "-")
    if "all" in s:
        return 0
    if "linear-horizontal" in s or "lh" in s:
        return 1
    if "linear-vertical" in s or "lv" in s:
        return 2
    if s == "linear":
        return 3
    if "circular-left" in s or "cl" in s:
        return 4
    if "circular-right" in s or "cr" in s:
        return 5
    raise ValueError("Polarization requested not recognized")


def polarization_vector(polarization, angle, hdir, pdir):
    """(Re, Im) x 3 of the reference's PolarizationVector (zeros for 'all')."""
    code = polarization_code(polarization)
    V = unit(cross(pdir, hdir))
    r2 = 1. / math.sqrt(2)
    out = [0.0] * 6
    if code == 1:
        for k in range(3):
            out[2 * k] = float(hdir[k])
    elif code == 2:
        for k in range(3):
            out[2 * k] = V[k]
    elif code == 3:
        A = rot_axis([float(h) for h in hdir], angle, pdir)
        for k in range(3):
            out[2 * k] = A[k]
    elif code in (4, 5):
        for k in range(3):
            vr, vi = 0. * V[k], 1. * V[k]
            re = hdir[k] + vr if code == 4 else hdir[k] - vr
            im = 0. + vi if code == 4 else 0. - vi
            out[2 * k] = re * r2
            out[2 * k + 1] = im * r2
    return out


def icosphere(radius, subdiv, centre=(0.0, 0.0, 0.0)):
    """Triangles (normal, a, b, c) of an icosahedron subdivided `subdiv` times
    and pushed out to a sphere of `radius` (20 * 4^subdiv triangles, outward
    unit normals): a synthetic closed mesh for calculate_power_density_stl."""
    import numpy as np
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    v = [np.array(p, dtype=float) / np.linalg.norm(p) for p in v]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    for _ in range(subdiv):
        cache = {}

        def middle(i, j):
            k = (min(i, j), max(i, j))
            if k not in cache:
                m = (v[i] + v[j]) / 2
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = middle(a, b), middle(b, c), middle(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    P = np.array(v) * radius + np.asarray(centre, dtype=float)
    out = []
    for a, b, c in f:
        n = np.cross(P[b] - P[a], P[c] - P[a])
        out.append((list(n / np.linalg.norm(n)), list(P[a]), list(P[b]), list(P[c])))
    return out


def stl_rows(triangles):
    """tri[n, 15] (A, B, C, normal, centre (A + B + C) / 3) of (normal, a, b, c)
    triangles, the layout read_stl gives, for Device.power_density_stl."""
    import numpy as np
    rows = []
    for n, a, b, c in triangles:
        a, b, c = np.asarray(a, float), np.asarray(b, float), np.asarray(c, float)
        rows.append(np.concatenate([a, b, c, np.asarray(n, float), (a + b + c) / 3.]))
    return np.array(rows).reshape(-1, 15)
