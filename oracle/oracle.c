/* CPU oracle: plain-C restatement of the reference OSCARS CPU path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reproduces the reference binary
 * bit for bit: every expression keeps the reference's association order, the
 * file is compiled with -ffp-contract=off, and the transcendental calls are the
 * same glibc ones (sin, cos, sqrt, pow, cexp) the reference makes.
 *
 * Reference files restated (paths relative to the reference root):
 *   include/TOSCARS.h:20-76                constants
 *   src/TVector3D.cc, include/TVector3D.h  vector op order
 *   src/TVector3DC.cc, include/TVector3DC.h complex vector op order
 *   src/TField3D_IdealUndulator.cc:63-152, src/TField3D_IdealEPU.cc:64-170,
 *   src/TField3D_Grid.cc:153-278, 353-548, src/TFieldContainer.cc:85-107
 *   src/TParticleBeam.cc:70-95, 667-689, src/TParticleA.cc:454-485
 *   src/OSCARSSR.cc:910-1009 (RK4 driver), 1247-1279 (DerivativesB), 1321-1386 (RK4)
 *   include/TOMATH.h:92-203 (spline), src/TParticleTrajectoryInterpolated.cc:124-270,
 *   src/TParticleTrajectoryInterpolatedPoints.cc:21-60 (levels)
 *   src/OSCARSSR.cc:1829-2046 (spectrum), 4075-4275 (flux), 2817-2976 (power density),
 *   2279-2387, 4281-4388, 2981-3083 (thread partitions)
 *   src/TSurfacePoints_Rectangle.cc:129-317, src/TSpectrumContainer.cc:57-91
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

/* ---------------------------------------------------------------- constants
 * include/TOSCARS.h:20-76 (1998-era values, kept verbatim) */
static const double K_PI = 3.14159265358979323846;
static double k_twopi(void) { return 2.0 * K_PI; }
static double k_fourpi(void) { return 4.0 * K_PI; }
static double k_pi2(void) { return K_PI * K_PI; }
static const double K_SQRT2PI = 2.5066282746310002416123552393401;
static const double K_C = 2.99792458e8;
static const double K_H = 6.62606876e-34;
static const double K_HBAR = 1.054571596e-34;
static const double K_QE = 1.602176462e-19;
static const double K_ME = 9.10938356e-31;
static const double K_EPS0 = 8.854187817E-12;
static const double K_MU0 = 1.2566370614E-6;
static double ev_to_omega(double e) { return e * K_QE / K_HBAR; }
static double kg_to_gev(double m) { return 1e-9 * m * K_C * K_C / K_QE; }

/* ---------------------------------------------------------------- vectors
 * src/TVector3D.cc: Mag2 = x*x + y*y + z*z, Dot likewise, Cross as written,
 * UnitVector = each component / Mag(). */
static double v_mag2(const double* a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static double v_dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v_cross(const double* a, const double* v, double* o) {
  double x = a[1] * v[2] - v[1] * a[2];
  double y = a[2] * v[0] - v[2] * a[0];
  double z = a[0] * v[1] - v[0] * a[1];
  o[0] = x; o[1] = y; o[2] = z;
}
static void v_unit(const double* a, double* o) {
  double m = sqrt(v_mag2(a));
  o[0] = a[0] / m; o[1] = a[1] / m; o[2] = a[2] / m;
}
/* src/TVector3D.cc RotateSelfX/Y/Z */
static void v_rot_x(double* v, double ang) {
  double s = sin(ang), c = cos(ang), yy = v[1];
  v[1] = c * yy - s * v[2];
  v[2] = s * yy + c * v[2];
}
static void v_rot_y(double* v, double ang) {
  double s = sin(ang), c = cos(ang), zz = v[2];
  v[2] = c * zz - s * v[0];
  v[0] = s * zz + c * v[0];
}
static void v_rot_z(double* v, double ang) {
  double s = sin(ang), c = cos(ang), xx = v[0];
  v[0] = c * xx - s * v[1];
  v[1] = s * xx + c * v[1];
}
static void v_rot_xyz(double* v, const double* r) {
  v_rot_x(v, r[0]);
  v_rot_y(v, r[1]);
  v_rot_z(v, r[2]);
}
/* src/TVector3D.cc:265-295 RotateSelf(angle, axis) */
static void v_rot_axis(double* o, double A, const double* vin) {
  double V[3];
  v_unit(vin, V);
  double M[3][3];
  M[0][0] = cos(A) + V[0] * V[0] * (1 - cos(A));
  M[0][1] = V[0] * V[1] * (1 - cos(A)) - V[2] * sin(A);
  M[0][2] = V[0] * V[2] * (1 - cos(A)) + V[1] * sin(A);
  M[1][0] = V[1] * V[0] * (1 - cos(A)) + V[2] * sin(A);
  M[1][1] = cos(A) + V[1] * V[1] * (1 - cos(A));
  M[1][2] = V[1] * V[2] * (1 - cos(A)) - V[0] * sin(A);
  M[2][0] = V[2] * V[0] * (1 - cos(A)) - V[1] * sin(A);
  M[2][1] = V[2] * V[1] * (1 - cos(A)) + V[0] * sin(A);
  M[2][2] = cos(A) + V[2] * V[2] * (1 - cos(A));
  double O[3] = {o[0], o[1], o[2]};
  o[0] = M[0][0] * O[0] + M[0][1] * O[1] + M[0][2] * O[2];
  o[1] = M[1][0] * O[0] + M[1][1] * O[1] + M[1][2] * O[2];
  o[2] = M[2][0] * O[0] + M[2][1] * O[1] + M[2][2] * O[2];
}

/* ---------------------------------------------------------------- fields */
void orc_fields_init(orc_fields* F) { memset(F, 0, sizeof(*F)); }

void orc_fields_free(orc_fields* F) {
  for (int i = 0; i < F->n; ++i) free(F->f[i].data);
  memset(F, 0, sizeof(*F));
}

static void undulator_geometry(orc_field* f) {
  /* src/TField3D_IdealUndulator.cc:101-106 */
  f->period_length = sqrt(v_mag2(f->period));
  v_unit(f->period, f->unit);
  f->und_length = sqrt(v_mag2(f->period)) * (double) (f->nperiods + 2);
}

static orc_field* new_field(orc_fields* F, int type) {
  if (F->n >= ORC_MAX_FIELDS) return NULL;
  orc_field* f = &F->f[F->n++];
  memset(f, 0, sizeof(*f));
  f->type = type;
  return f;
}

int orc_add_gaussian(orc_fields* F, const double* bfield, const double* center, const double* sigma,
                     const double* rotations) {
  orc_field* f = new_field(F, ORC_FIELD_GAUSSIAN);
  if (!f) return 1;
  for (int k = 0; k < 3; ++k) {
    f->field[k] = bfield[k];
    f->center[k] = center[k];
    f->shape[k] = sigma[k];
    f->rot[k] = rotations[k];
  }
  v_rot_xyz(f->field, f->rot);
  return 0;
}

int orc_add_uniform(orc_fields* F, const double* bfield, const double* width, const double* center,
                    const double* rotations) {
  orc_field* f = new_field(F, ORC_FIELD_UNIFORM);
  if (!f) return 1;
  for (int k = 0; k < 3; ++k) {
    f->field[k] = bfield[k];
    f->center[k] = center[k];
    f->shape[k] = width[k];
    f->rot[k] = rotations[k];
  }
  v_rot_xyz(f->field, f->rot);
  return 0;
}

int orc_add_quadrupole(orc_fields* F, double K, double width, const double* rotations, const double* translation) {
  orc_field* f = new_field(F, ORC_FIELD_QUADRUPOLE);
  if (!f) return 1;
  f->K = K;
  f->width = width;
  for (int k = 0; k < 3; ++k) {
    f->rot[k] = rotations[k];
    f->trans[k] = translation[k];
  }
  return 0;
}

int orc_add_undulator(orc_fields* F, const double* bfield, const double* period, int nperiods,
                      const double* translation, double phase, double taper) {
  if (F->n >= ORC_MAX_FIELDS) return 1;
  orc_field* f = &F->f[F->n++];
  memset(f, 0, sizeof(*f));
  f->type = ORC_FIELD_UNDULATOR;
  memcpy(f->field, bfield, 3 * sizeof(double));
  memcpy(f->period, period, 3 * sizeof(double));
  memcpy(f->center, translation, 3 * sizeof(double));
  f->nperiods = nperiods;
  f->phase = phase;
  f->taper = taper;
  undulator_geometry(f);
  return 0;
}

int orc_add_epu(orc_fields* F, const double* bfieldA, const double* bfieldB, const double* period, int nperiods,
                const double* translation, double phase, double taper) {
  if (F->n >= ORC_MAX_FIELDS) return 1;
  orc_field* f = &F->f[F->n++];
  memset(f, 0, sizeof(*f));
  f->type = ORC_FIELD_EPU;
  memcpy(f->field, bfieldA, 3 * sizeof(double));
  memcpy(f->fieldB, bfieldB, 3 * sizeof(double));
  memcpy(f->period, period, 3 * sizeof(double));
  memcpy(f->center, translation, 3 * sizeof(double));
  f->nperiods = nperiods;
  f->phase = phase;
  f->taper = taper;
  undulator_geometry(f);
  return 0;
}

/* src/TField3D_Grid.cc:353-548 (OSCARS text format, no scaling) */
/* A resampled OSCARS1D table (src/TField3D_Grid.cc:572-873): start and stop as given. */
int orc_add_grid_exact(orc_fields* F, const double* start, const double* stop, const double* step, const long* npts,
                       const double* samples, const double* rotations, const double* translation) {
  double zero[3] = {0, 0, 0};
  int rc = orc_add_grid(F, zero, step, npts, samples, rotations, translation);
  if (rc) return rc;
  orc_field* f = &F->f[F->n - 1];
  f->xstart = npts[0] > 1 ? start[0] : 0; f->xstop = npts[0] > 1 ? stop[0] : 0;
  f->ystart = npts[1] > 1 ? start[1] : 0; f->ystop = npts[1] > 1 ? stop[1] : 0;
  f->zstart = npts[2] > 1 ? start[2] : 0; f->zstop = npts[2] > 1 ? stop[2] : 0;
  return 0;
}

int orc_add_grid(orc_fields* F, const double* start_in, const double* step_in, const long* npts,
                 const double* samples, const double* rotations, const double* translation) {
  if (F->n >= ORC_MAX_FIELDS) return 1;
  int NX = (int) npts[0], NY = (int) npts[1], NZ = (int) npts[2];
  if (NX < 1 || NY < 1) return 2;
  int hasx = NX > 1, hasy = NY > 1, hasz = NZ > 1;
  if (hasx + hasy + hasz < 1) return 3;
  orc_field* f = &F->f[F->n++];
  memset(f, 0, sizeof(*f));
  f->type = ORC_FIELD_GRID;
  double XStep = step_in[0], YStep = step_in[1], ZStep = step_in[2];
  double MiddleX = start_in[0] + step_in[0] * (NX - 1) / 2.;
  double XStart = MiddleX - XStep * (NX - 1) / 2.;
  double MiddleY = start_in[1] + step_in[1] * (NY - 1) / 2.;
  double YStart = MiddleY - YStep * (NY - 1) / 2.;
  double MiddleZ = start_in[2] + step_in[2] * (NZ - 1) / 2.;
  double ZStart = MiddleZ - ZStep * (NZ - 1) / 2.;
  f->nx = NX; f->ny = NY; f->nz = NZ;
  f->xstart = XStart; f->ystart = YStart; f->zstart = ZStart;
  f->xstep = XStep; f->ystep = YStep; f->zstep = ZStep;
  f->xstop = f->xstart + (double) (size_t) (f->nx - 1) * f->xstep;
  f->ystop = f->ystart + (double) (size_t) (f->ny - 1) * f->ystep;
  f->zstop = f->zstart + (double) (size_t) (f->nz - 1) * f->zstep;
  /* 1 x, 2 y, 3 z for one axis; 12, 13, 23, 123 for several */
  if (hasx + hasy + hasz == 1) f->dimx = hasx ? 1 : (hasy ? 2 : 3);
  else f->dimx = hasx && hasy && hasz ? 123 : (hasx && hasy ? 12 : (hasx ? 13 : 23));
  long n = (long) NX * NY * NZ;
  f->data = (double*) malloc(3 * n * sizeof(double));
  for (long i = 0; i < n; ++i) {
    double v[3] = {samples[3 * i], samples[3 * i + 1], samples[3 * i + 2]};
    v_rot_xyz(v, rotations);
    memcpy(f->data + 3 * i, v, sizeof(v));
  }
  memcpy(f->rot, rotations, 3 * sizeof(double));
  memcpy(f->trans, translation, 3 * sizeof(double));
  return 0;
}

/* one ideal-undulator sinusoid with terminations; src/TField3D_IdealUndulator.cc:129-145 */
static int und_region(const orc_field* f, double D, double PhaseShift, double* amp_scale) {
  double L = f->und_length, PL = f->period_length;
  if (D > L / 2. + PhaseShift || D < -L / 2. + PhaseShift) return 0;
  if (D < -L / 2. + PhaseShift + PL || D > L / 2. + PhaseShift - PL) {
    if (D < -L / 2. + PhaseShift + PL / 2. || D > L / 2. + PhaseShift - PL / 2.) {
      *amp_scale = 0.25;
    } else {
      *amp_scale = 0.75;
    }
    return 2;
  }
  return 1;
}

/* cos(2 pi f (T + t0) + phase), src/TField3D_IdealUndulator.cc:151 */
static double td_factor(const orc_field* f, double T) {
  return cos(k_twopi() * f->freq * (T + f->toff) + f->fphase);
}

static void td_apply(const orc_field* f, double T, double* out) {
  double c = td_factor(f, T);
  out[0] = out[0] * c; out[1] = out[1] * c; out[2] = out[2] * c;
}

static void field_eval(const orc_field* f, const double* X, double T, double* out) {
  out[0] = out[1] = out[2] = 0;
  if (f->type == ORC_FIELD_UNDULATOR || f->type == ORC_FIELD_EPU) {
    double d[3] = {X[0] - f->center[0], X[1] - f->center[1], X[2] - f->center[2]};
    double D = v_dot(d, f->unit);
    double TaperCorrection = 1 + D * f->taper;
    for (int part = 0; part < (f->type == ORC_FIELD_EPU ? 2 : 1); ++part) {
      double PhaseShift;
      const double* amp;
      if (f->type == ORC_FIELD_UNDULATOR) {
        PhaseShift = f->phase * sqrt(v_mag2(f->period)) / k_twopi();
        amp = f->field;
      } else if (part == 0) {
        PhaseShift = f->phase / 2. * sqrt(v_mag2(f->period)) / k_twopi();
        amp = f->field;
      } else {
        PhaseShift = -f->phase / 2. * sqrt(v_mag2(f->period)) / k_twopi();
        amp = f->fieldB;
      }
      double scale = 1;
      int r = und_region(f, D, PhaseShift, &scale);
      /* outside the undulator the reference returns zero before the time factor */
      if (r == 0 && f->type == ORC_FIELD_UNDULATOR) return;
      if (r == 0) continue;
      double s = sin(k_twopi() * (D - PhaseShift) / f->period_length);
      double v[3];
      for (int k = 0; k < 3; ++k) {
        double a = (r == 2) ? amp[k] * scale : amp[k];
        v[k] = a * s * TaperCorrection;
      }
      if (f->type == ORC_FIELD_EPU && part == 1) {
        out[0] += v[0]; out[1] += v[1]; out[2] += v[2];
      } else {
        out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
      }
    }
    if (f->freq != 0) td_apply(f, T, out);
    return;
  }
  double x[3] = {X[0], X[1], X[2]};
  v_rot_xyz(x, f->rot);
  if (f->type == ORC_FIELD_GAUSSIAN || f->type == ORC_FIELD_UNIFORM) {
    /* src/TField3D_Gaussian.cc:83-112, src/TField3D_UniformBox.cc:98-120 */
    double RX[3] = {x[0] - f->center[0], x[1] - f->center[1], x[2] - f->center[2]};
    if (f->type == ORC_FIELD_UNIFORM) {
      if ((!(f->shape[0] <= 0) && fabs(RX[0]) > fabs(f->shape[0] / 2.)) ||
          (!(f->shape[1] <= 0) && fabs(RX[1]) > fabs(f->shape[1] / 2.)) ||
          (!(f->shape[2] <= 0) && fabs(RX[2]) > fabs(f->shape[2] / 2.)))
        return;
      out[0] = f->field[0]; out[1] = f->field[1]; out[2] = f->field[2];
      if (f->freq != 0) td_apply(f, T, out);
      return;
    }
    double Fraction = 1;
    for (int k = 0; k < 3; ++k)
      if (f->shape[k] > 0) Fraction *= exp(-pow((RX[k]) / f->shape[k], 2) / 2.);
    out[0] = Fraction * f->field[0]; out[1] = Fraction * f->field[1]; out[2] = Fraction * f->field[2];
    if (f->freq != 0) td_apply(f, T, out);
    return;
  }
  if (f->type == ORC_FIELD_QUADRUPOLE) {
    /* src/TField3D_Quadrupole.cc:62-83 */
    x[0] -= f->trans[0]; x[1] -= f->trans[1]; x[2] -= f->trans[2];
    if (fabs(x[2]) > f->width / 2) return;
    double Ret[3] = {f->K * x[1], f->K * x[0], 0};
    v_rot_xyz(Ret, f->rot);
    out[0] = Ret[0]; out[1] = Ret[1]; out[2] = Ret[2];
    if (f->freq != 0) td_apply(f, T, out);
    return;
  }
  /* grid, src/TField3D_Grid.cc:153-278 (1-D branches) */
  x[0] -= f->trans[0]; x[1] -= f->trans[1]; x[2] -= f->trans[2];
  if (f->nx > 1 && (x[0] <= f->xstart || x[0] >= f->xstop)) return;
  if (f->ny > 1 && (x[1] <= f->ystart || x[1] >= f->ystop)) return;
  if (f->nz > 1 && (x[2] <= f->zstart || x[2] >= f->zstop)) return;
  if (f->dimx > 3) {
    /* src/TField3D_Grid.cc:174-268 */
    size_t nx = f->nx > 1 ? (size_t) ((x[0] - f->xstart) / f->xstep) : 0;
    double dx = f->nx > 1 ? (x[0] - f->xstart) - nx * f->xstep : 0;
    size_t ny = f->ny > 1 ? (size_t) ((x[1] - f->ystart) / f->ystep) : 0;
    double dy = f->ny > 1 ? (x[1] - f->ystart) - ny * f->ystep : 0;
    size_t nz = f->nz > 1 ? (size_t) ((x[2] - f->zstart) / f->zstep) : 0;
    double dz = f->nz > 1 ? (x[2] - f->zstart) - nz * f->zstep : 0;
#define GI(ix, iy, iz) (f->data + 3 * ((ix) * f->ny * f->nz + (iy) * f->nz + (iz)))
#define LERP(r, lo, hi, d, st) for (int k = 0; k < 3; ++k) (r)[k] = (lo)[k] + (d) * ((hi)[k] - (lo)[k]) / (st)
    double v00[3], v10[3], v01[3], v11[3], v0[3], v1[3], F3[3];
    if (f->dimx == 123) {
      LERP(v00, GI(nx, ny, nz), GI(nx + 1, ny, nz), dx, f->xstep);
      LERP(v10, GI(nx, ny + 1, nz), GI(nx + 1, ny + 1, nz), dx, f->xstep);
      LERP(v01, GI(nx, ny, nz + 1), GI(nx + 1, ny, nz + 1), dx, f->xstep);
      LERP(v11, GI(nx, ny + 1, nz + 1), GI(nx + 1, ny + 1, nz + 1), dx, f->xstep);
      LERP(v0, v00, v10, dy, f->ystep);
      LERP(v1, v01, v11, dy, f->ystep);
      LERP(F3, v0, v1, dz, f->zstep);
    } else if (f->dimx == 12) {
      LERP(v0, GI(nx, ny, 0), GI(nx + 1, ny, 0), dx, f->xstep);
      LERP(v1, GI(nx, ny + 1, 0), GI(nx + 1, ny + 1, 0), dx, f->xstep);
      LERP(F3, v0, v1, dy, f->ystep);
    } else if (f->dimx == 13) {
      LERP(v0, GI(nx, 0, nz), GI(nx + 1, 0, nz), dx, f->xstep);
      LERP(v1, GI(nx, 0, nz + 1), GI(nx + 1, 0, nz + 1), dx, f->xstep);
      LERP(F3, v0, v1, dz, f->zstep);
    } else {
      LERP(v0, GI(0, ny, nz), GI(0, ny + 1, nz), dy, f->ystep);
      LERP(v1, GI(0, ny, nz + 1), GI(0, ny + 1, nz + 1), dy, f->ystep);
      LERP(F3, v0, v1, dz, f->zstep);
    }
#undef GI
#undef LERP
    double c = td_factor(f, T);
    for (int k = 0; k < 3; ++k) out[k] = F3[k] * c;
    return;
  }
  double coord, start, step;
  if (f->dimx == 1) { coord = x[0]; start = f->xstart; step = f->xstep; }
  else if (f->dimx == 2) { coord = x[1]; start = f->ystart; step = f->ystep; }
  else { coord = x[2]; start = f->zstart; step = f->zstep; }
  size_t n0 = (size_t) ((coord - start) / step);
  double dd = (coord - start) - (double) n0 * step;
  const double* a = f->data + 3 * n0;
  const double* b = f->data + 3 * (n0 + 1);
  for (int k = 0; k < 3; ++k) {
    double v = a[k] + (b[k] - a[k]) * dd / step;
    out[k] = v * td_factor(f, T);
  }
}

/* src/TFieldContainer.cc:85-107 */
void orc_get_b(const orc_fields* F, const double* x, double t, double* out) {
  double sum[3] = {0, 0, 0};
  for (int i = 0; i < F->n; ++i) {
    double v[3];
    field_eval(&F->f[i], x, t, v);
    sum[0] += v[0]; sum[1] += v[1]; sum[2] += v[2];
  }
  out[0] = sum[0]; out[1] = sum[1]; out[2] = sum[2];
}

/* ---------------------------------------------------------------- particle
 * electron beam + ideal particle: src/TParticleBeam.cc:97-130, 667-689,
 * src/TParticleA.cc:454-485 */
static void particle_set_gamma(orc_particle* P) {
  double Beta2 = v_mag2(P->b0) > 0 ? v_mag2(P->b0) : 0;
  if (Beta2 == 1) return;
  P->gamma = Beta2 != 0 ? 1.0 / sqrt(1.0 - Beta2) : 1;
  if (P->m == 0 || P->gamma == 0) return;
  P->qomg = P->q / P->m / P->gamma;
}

void orc_particle_electron(orc_particle* P, double energy_GeV, const double* x0, const double* d0, double current) {
  memset(P, 0, sizeof(*P));
  P->q = -K_QE;
  P->m = K_ME;
  P->current = current;
  double u0[3] = {0, 0, 0};
  if (v_mag2(d0) > 0) v_unit(d0, u0);
  double E0 = energy_GeV < kg_to_gev(P->m) ? P->m : energy_GeV;
  double Gamma = E0 / kg_to_gev(P->m) < 1 ? 1 : E0 / kg_to_gev(P->m);
  double Beta = Gamma != 1 ? sqrt(1.0 - 1.0 / (Gamma * Gamma)) : 0;
  memcpy(P->x0, x0, 3 * sizeof(double));
  P->b0[0] = u0[0] * Beta; P->b0[1] = u0[1] * Beta; P->b0[2] = u0[2] * Beta;
  P->t0 = 0;
  particle_set_gamma(P);
}

void orc_particle_set(orc_particle* P, const double* x0, const double* b0, double t0) {
  memcpy(P->x0, x0, 3 * sizeof(double));
  memcpy(P->b0, b0, 3 * sizeof(double));
  P->t0 = t0;
  particle_set_gamma(P);
}

/* ---------------------------------------------------------------- RK4 */
typedef struct {
  const orc_fields* F;
  const orc_particle* P;
  int err_gamma;
  const orc_fields* FE; /* electric fields, or NULL */
} rk_ctx;

/* src/OSCARSSR.cc:1247-1279 */
static void derivs_b(rk_ctx* c, double t, const double* x, double* dxdt) {
  if (c->FE && c->FE->n > 0) {
    /* src/OSCARSSR.cc:1283-1318 */
    double OneMinusEB = (1. - (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) / (K_C * K_C));
    if (OneMinusEB <= 0) c->err_gamma = 1;
    double X[3] = {x[0], x[2], x[4]};
    double B[3], E[3];
    orc_get_b(c->F, X, t, B);
    orc_get_b(c->FE, X, t, E);
    double Q = c->P->q / c->P->m * sqrt(1. - (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) / (K_C * K_C));
    double BetaDotE = (x[1] * E[0] + x[3] * E[1] + x[5] * E[2]) / K_C;
    dxdt[0] = x[1];
    dxdt[1] = Q * (E[0] - x[5] * B[1] + x[3] * B[2] - x[1] * BetaDotE / K_C);
    dxdt[2] = x[3];
    dxdt[3] = Q * (E[1] + x[5] * B[0] - x[1] * B[2] - x[3] * BetaDotE / K_C);
    dxdt[4] = x[5];
    dxdt[5] = Q * (E[2] + x[1] * B[1] - x[3] * B[0] - x[5] * BetaDotE / K_C);
    return;
  }
  double OneMinus = (1. - (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) / (K_C * K_C));
  if (OneMinus <= 0) {
    c->err_gamma = 1;
    return;
  }
  double X[3] = {x[0], x[2], x[4]};
  double B[3];
  orc_get_b(c->F, X, t, B);
  double q = c->P->qomg;
  dxdt[0] = x[1];
  dxdt[1] = q * (-x[5] * B[1] + x[3] * B[2]);
  dxdt[2] = x[3];
  dxdt[3] = q * (x[5] * B[0] - x[1] * B[2]);
  dxdt[4] = x[5];
  dxdt[5] = q * (x[1] * B[1] - x[3] * B[0]);
}

/* steps split by the last orc_trajectory_rk4_eb call (test instrumentation) */
static long g_halvings = 0;
static int g_halving_depth = 0;
long orc_halving_count(void) { return g_halvings; }
int orc_halving_depth(void) { return g_halving_depth; }

/* src/OSCARSSR.cc:1321-1386 (NR rk4 with gamma-error step halving) */
static void rk4_step(rk_ctx* c, double* y, double* dydx, double x, double h, double* yout, int depth) {
  double dym[6], dyt[6], yt[6], yo[6];
  double hh = h * 0.5;
  double h6 = h / 6.0;
  double xh = x + hh;
  int i;
  for (i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  derivs_b(c, xh, yt, dyt);
  for (i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  derivs_b(c, xh, yt, dym);
  for (i = 0; i < 6; ++i) {
    yt[i] = y[i] + h * dym[i];
    dym[i] += dyt[i];
  }
  derivs_b(c, x + h, yt, dyt);
  for (i = 0; i < 6; ++i) yo[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  int this_err = 1. - (yo[1] * yo[1] + yo[3] * yo[3] + yo[5] * yo[5]) / (K_C * K_C) <= 0 ? 1 : 0;
  if (c->err_gamma || this_err) {
    ++g_halvings;
    if (depth + 1 > g_halving_depth) g_halving_depth = depth + 1;
    c->err_gamma = 0;
    double yh[6];
    derivs_b(c, x, y, dydx);
    rk4_step(c, y, dydx, x, h / 2., yh, depth + 1);
    derivs_b(c, x + h / 2., yh, dydx);
    rk4_step(c, yh, dydx, x + h / 2., h / 2., yo, depth + 1);
  }
  for (i = 0; i < 6; ++i) yout[i] = yo[i];
}

long orc_npoints_forward_backward(double ctstart, double ctstop, long npoints, double t0, long* nbwd) {
  double DeltaT = ((ctstop - ctstart) / K_C / (double) (size_t) (npoints - 1));
  size_t nf = (size_t) (1 + (ctstop - t0) / K_C / DeltaT);
  size_t nb = (size_t) ((t0 - ctstart) / K_C / DeltaT);
  if (nbwd) *nbwd = (long) nb;
  return (long) nf;
}

static void reverse_rows(double* t, double* p, long n) {
  for (long i = 0, j = n - 1; i < j; ++i, --j) {
    double tt = t[i]; t[i] = t[j]; t[j] = tt;
    for (int k = 0; k < 9; ++k) {
      double v = p[9 * i + k]; p[9 * i + k] = p[9 * j + k]; p[9 * j + k] = v;
    }
  }
}

/* src/OSCARSSR.cc:910-1009.  Writes nf+nb points (caller sizes from
 * orc_npoints_forward_backward). */
int orc_trajectory_rk4(const orc_fields* F, const orc_particle* P, double ctstart, double ctstop, long npoints,
                       double* out_t, double* out_p) {
  return orc_trajectory_rk4_eb(F, NULL, P, ctstart, ctstop, npoints, out_t, out_p);
}

int orc_trajectory_rk4_eb(const orc_fields* F, const orc_fields* FE, const orc_particle* P, double ctstart,
                          double ctstop, long npoints, double* out_t, double* out_p) {
  rk_ctx c = {F, P, 0, FE};
  g_halvings = 0;
  g_halving_depth = 0;
  double DeltaT = ((ctstop - ctstart) / K_C / (double) (size_t) (npoints - 1));
  size_t NF = (size_t) (1 + (ctstop - P->t0) / K_C / DeltaT);
  size_t NB = (size_t) ((P->t0 - ctstart) / K_C / DeltaT);
  double x[6], dxdt[6];
  x[0] = P->x0[0]; x[1] = P->b0[0] * K_C;
  x[2] = P->x0[1]; x[3] = P->b0[1] * K_C;
  x[4] = P->x0[2]; x[5] = P->b0[2] * K_C;
  long n = 0;
  derivs_b(&c, P->t0 / K_C, x, dxdt);
  for (size_t i = 0; i != NF; ++i) {
    double t = P->t0 / K_C + DeltaT * (double) i;
    if (i == 0) derivs_b(&c, t, x, dxdt);
    double* o = out_p + 9 * n;
    o[0] = x[0]; o[1] = x[2]; o[2] = x[4];
    o[3] = x[1] / K_C; o[4] = x[3] / K_C; o[5] = x[5] / K_C;
    o[6] = dxdt[1] / K_C; o[7] = dxdt[3] / K_C; o[8] = dxdt[5] / K_C;
    out_t[n] = t;
    ++n;
    derivs_b(&c, t, x, dxdt);
    rk4_step(&c, x, dxdt, t, DeltaT, x, 0);
  }
  reverse_rows(out_t, out_p, n);
  x[0] = P->x0[0]; x[1] = P->b0[0] * K_C;
  x[2] = P->x0[1]; x[3] = P->b0[1] * K_C;
  x[4] = P->x0[2]; x[5] = P->b0[2] * K_C;
  double DeltaTReversed = -DeltaT;
  for (size_t i = 0; i != NB; ++i) {
    double t1 = P->t0 / K_C + DeltaTReversed * (double) (i);
    double t2 = P->t0 / K_C + DeltaTReversed * (double) (i + 1);
    derivs_b(&c, t1, x, dxdt);
    rk4_step(&c, x, dxdt, t1, DeltaTReversed, x, 0);
    double* o = out_p + 9 * n;
    o[0] = x[0]; o[1] = x[2]; o[2] = x[4];
    o[3] = x[1] / K_C; o[4] = x[3] / K_C; o[5] = x[5] / K_C;
    o[6] = dxdt[1] / K_C; o[7] = dxdt[3] / K_C; o[8] = dxdt[5] / K_C;
    out_t[n] = t2;
    ++n;
  }
  reverse_rows(out_t, out_p, n);
  return 0;
}

/* ---------------------------------------------------------------- spline
 * include/TOMATH.h:141-203, componentwise over the 9 channels.  The
 * tridiagonal factors depend only on t, so they are the same for every
 * channel; the arithmetic below is the per-channel arithmetic of the
 * reference's TParticleTrajectoryPoint operators. */
int orc_spline(long N, const double* X, const double* Y, double* YPP) {
  if (N < 3) return 1;
  double* u = (double*) malloc(9 * N * sizeof(double));
  for (int k = 0; k < 9; ++k) {
    YPP[k] = -0.5;
    double dydx0 = (Y[9 + k] - Y[k]) / (X[1] - X[0]);
    u[k] = ((Y[9 + k] - Y[k]) / (X[1] - X[0]) - dydx0) * (3. / (X[1] - X[0]));
  }
  for (long i = 1; i < N - 1; ++i) {
    double sig = (X[i] - X[i - 1]) / (X[i + 1] - X[i - 1]);
    for (int k = 0; k < 9; ++k) {
      double p = sig * YPP[9 * (i - 1) + k] + 2.0;
      YPP[9 * i + k] = (sig - 1.) / p;
      double ui = (Y[9 * (i + 1) + k] - Y[9 * i + k]) / (X[i + 1] - X[i]) -
                  (Y[9 * i + k] - Y[9 * (i - 1) + k]) / (X[i] - X[i - 1]);
      u[9 * i + k] = (ui * 6. / (X[i + 1] - X[i - 1]) - sig * u[9 * (i - 1) + k]) / p;
    }
  }
  double qn = 0.5;
  for (int k = 0; k < 9; ++k) {
    double dydxN = (Y[9 * (N - 1) + k] - Y[9 * (N - 2) + k]) / (X[N - 1] - X[N - 2]);
    double un = (dydxN - (Y[9 * (N - 1) + k] - Y[9 * (N - 2) + k]) / (X[N - 1] - X[N - 2])) *
                (3. / (X[N - 1] - X[N - 2]));
    YPP[9 * (N - 1) + k] = (un - u[9 * (N - 2) + k] * qn) / (YPP[9 * (N - 2) + k] * qn + 1.);
  }
  for (long j = N - 2; j >= 0; --j) {
    for (int k = 0; k < 9; ++k) YPP[9 * j + k] = YPP[9 * j + k] * YPP[9 * (j + 1) + k] + u[9 * j + k];
  }
  free(u);
  return 0;
}

/* include/TOMATH.h:92-126 */
void orc_spline_value(const orc_traj* T, double x, double* out) {
  int klo = 0, khi = (int) T->n - 1, k;
  while (khi - klo > 1) {
    k = (khi + klo) >> 1;
    if (T->t[k] > x) khi = k; else klo = k;
  }
  double h = T->t[khi] - T->t[klo];
  double a = (T->t[khi] - x) / h;
  double b = (x - T->t[klo]) / h;
  double ca = a * a * a - a;
  double cb = b * b * b - b;
  const double* Y0 = T->p + 9 * klo;
  const double* Y1 = T->p + 9 * khi;
  const double* P0 = T->ypp + 9 * klo;
  const double* P1 = T->ypp + 9 * khi;
  for (int c = 0; c < 9; ++c) {
    out[c] = Y0[c] * a + Y1[c] * b + (P0[c] * ca + P1[c] * cb) * (h * h) / 6.;
  }
}

long orc_level_npoints(int level) { return (long) pow(2, level); }

double orc_dt_inclusive(const orc_traj* T, int level) {
  return (T->tstop - T->tstart) / (pow(2., level + 1) - 1);
}

/* src/TParticleTrajectoryInterpolated.cc:124-161 (memory levels) and
 * src/TParticleTrajectoryInterpolatedPoints.cc:21-45 (extended levels) */
void orc_level_samples(const orc_traj* T, int level, int extended, double* out) {
  long np = (long) pow(2, level);
  double spacing = extended ? (T->tstop - T->tstart) / (pow(2., level) - 1)
                            : (T->tstop - T->tstart) / pow(2., level);
  double tstart = T->tstart + (T->tstop - T->tstart) / pow(2., level + 1);
  for (long i = 0; i < np; ++i) {
    double t = tstart + spacing * (double) i;
    out[10 * i] = t;
    orc_spline_value(T, t, out + 10 * i + 1);
  }
}

/* ---------------------------------------------------------------- geometry */
int orc_rectangle(const char* planein, int NX1, int NX2, double W1, double W2, const double* rot, const double* tr,
                  int normal, int dim, double* out_obs, double* out_normal, double* out_container) {
  char P[8] = {0};
  for (int i = 0; i < 7 && planein[i]; ++i) P[i] = (char) toupper((unsigned char) planein[i]);
  double s1 = NX1 > 1 ? W1 / (NX1 - 1) : 0;
  double s2 = NX2 > 1 ? W2 / (NX2 - 1) : 0;
  double st[3] = {0, 0, 0}, a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0}, nv[3] = {0, 0, 0};
  if (strcmp(P, "XY") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[0] = -W1 / 2.; st[1] = -W2 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[1] = -W2 / 2.; }
    else { st[0] = -W1 / 2.; }
    a1[0] = s1; a2[1] = s2; nv[2] = 1;
  } else if (strcmp(P, "YX") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[0] = -W2 / 2.; st[1] = -W1 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[0] = -W2 / 2.; }
    else { st[1] = -W1 / 2.; }
    a2[0] = s2; a1[1] = s1; nv[2] = -1;
  } else if (strcmp(P, "XZ") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[0] = -W1 / 2.; st[2] = -W2 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[2] = -W2 / 2.; }
    else { st[0] = -W1 / 2.; }
    a1[0] = s1; a2[2] = s2; nv[1] = -1;
  } else if (strcmp(P, "ZX") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[0] = -W2 / 2.; st[2] = -W1 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[0] = -W2 / 2.; }
    else { st[2] = -W1 / 2.; }
    a2[0] = s2; a1[2] = s1; nv[1] = 1;
  } else if (strcmp(P, "YZ") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[1] = -W1 / 2.; st[2] = -W2 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[2] = -W2 / 2.; }
    else { st[1] = -W1 / 2.; }
    a1[1] = s1; a2[2] = s2; nv[0] = 1;
  } else if (strcmp(P, "ZY") == 0) {
    if (NX1 > 1 && NX2 > 1) { st[1] = -W2 / 2.; st[2] = -W1 / 2.; }
    else if (NX1 <= 1 && NX2 <= 1) { }
    else if (NX1 <= 1) { st[1] = -W2 / 2.; }
    else { st[2] = -W1 / 2.; }
    a2[1] = s2; a1[2] = s1; nv[0] = -1;
  } else {
    return 1;
  }
  v_rot_xyz(st, rot);
  v_rot_xyz(a1, rot);
  v_rot_xyz(a2, rot);
  st[0] += tr[0]; st[1] += tr[1]; st[2] += tr[2];
  v_rot_xyz(nv, rot);
  if (normal == -1) { nv[0] *= -1; nv[1] *= -1; nv[2] *= -1; }
  else if (normal != 0 && normal != 1) return 2;
  if (out_normal) memcpy(out_normal, nv, sizeof(nv));
  long n = (long) NX1 * NX2;
  double m1 = sqrt(v_mag2(a1)), m2 = sqrt(v_mag2(a2));
  for (long i = 0; i < n; ++i) {
    int ix1 = (int) (i / NX2);
    int ix2 = (int) (i % NX2);
    double* o = out_obs + 3 * i;
    for (int k = 0; k < 3; ++k) o[k] = st[k] + a1[k] * (double) ix1 + a2[k] * (double) ix2;
    if (out_container) {
      double* c = out_container + 3 * i;
      if (dim == 2) {
        c[0] = s1 * ix1 - m1 * (NX1 - 1) / 2.;
        c[1] = s2 * ix2 - m2 * (NX2 - 1) / 2.;
        c[2] = 0;
      } else {
        c[0] = o[0]; c[1] = o[1]; c[2] = o[2];
      }
    }
  }
  return 0;
}

/* src/TSpectrumContainer.cc:57-91 */
void orc_energies(int N, double EFirst, double ELast, double* out) {
  if (N == 1) { out[0] = EFirst; return; }
  for (int i = 0; i < N; ++i) out[i] = EFirst + (ELast - EFirst) / (double) (size_t) (N - 1) * (double) (i);
}

/* ---------------------------------------------------------------- polarization */
int orc_polarization_code(const char* in) {
  char s[64] = {0};
  int n = 0;
  for (; in[n] && n < 63; ++n) s[n] = (char) (in[n] == ' ' ? '-' : tolower((unsigned char) in[n]));
  if (strstr(s, "all")) return 0;
  if (strstr(s, "linear-horizontal") || strstr(s, "lh")) return 1;
  if (strstr(s, "linear-vertical") || strstr(s, "lv")) return 2;
  if (strcmp(s, "linear") == 0) return 3;
  if (strstr(s, "circular-left") || strstr(s, "cl")) return 4;
  if (strstr(s, "circular-right") || strstr(s, "cr")) return 5;
  return -1;
}

/* src/OSCARSSR.cc:1887-1919 */
void orc_polarization_vector(int code, double angle, const double* H, const double* Pd, double* out) {
  double V[3], c[3];
  v_cross(Pd, H, c);
  v_unit(c, V);
  double r2 = 1. / sqrt(2);
  memset(out, 0, 6 * sizeof(double));
  if (code == 1) {
    for (int k = 0; k < 3; ++k) out[2 * k] = H[k];
  } else if (code == 2) {
    for (int k = 0; k < 3; ++k) out[2 * k] = V[k];
  } else if (code == 3) {
    double A[3] = {H[0], H[1], H[2]};
    v_rot_axis(A, angle, Pd);
    for (int k = 0; k < 3; ++k) out[2 * k] = A[k];
  } else if (code == 4 || code == 5) {
    for (int k = 0; k < 3; ++k) {
      /* (H + V*i) or (H - V*i), with V*i = (0*V, 1*V) as libstdc++ forms it */
      double vr = 0. * V[k], vi = 1. * V[k];
      double re = code == 4 ? H[k] + vr : H[k] - vr;
      double im = code == 4 ? 0. + vi : 0. - vi;
      out[2 * k] = re * r2;
      out[2 * k + 1] = im * r2;
    }
  }
}

/* ---------------------------------------------------------------- complex helpers
 * GCC's inline complex product: (ac - bd) + i(ad + bc). */
static void c_mul(double ar, double ai, double br, double bi, double* rr, double* ri) {
  double ac = ar * br, bd = ai * bi, ad = ar * bi, bc = ai * br;
  *rr = ac - bd;
  *ri = ad + bc;
}
static double pol_mag2(const double* P) {
  double s = 0;
  for (int k = 0; k < 3; ++k) {
    double r, i;
    c_mul(P[2 * k], P[2 * k + 1], P[2 * k], -P[2 * k + 1], &r, &i);
    s = k == 0 ? r : s + r;
  }
  return s;
}
/* E <- (E . P) P  (bilinear dot, src/TVector3DC.cc:123-127) */
static void pol_project(double* E, const double* P) {
  double sr = 0, si = 0;
  for (int k = 0; k < 3; ++k) {
    double r, i;
    c_mul(E[2 * k], E[2 * k + 1], P[2 * k], P[2 * k + 1], &r, &i);
    if (k == 0) { sr = r; si = i; } else { sr += r; si += i; }
  }
  for (int k = 0; k < 3; ++k) c_mul(P[2 * k], P[2 * k + 1], sr, si, &E[2 * k], &E[2 * k + 1]);
}
static double cvec_mag2(const double* E) {
  double s = 0;
  for (int k = 0; k < 3; ++k) {
    double r, i;
    c_mul(E[2 * k], E[2 * k + 1], E[2 * k], -E[2 * k + 1], &r, &i);
    s = k == 0 ? r : s + r;
  }
  return s;
}

/* ---------------------------------------------------------------- level cache */
typedef struct {
  const orc_traj* T;
  int stop_memory;
  double* lev[ORC_MAX_LEVEL + 64];
  pthread_mutex_t mu;
} level_cache;

static void lc_init(level_cache* c, const orc_traj* T, int stop_memory) {
  memset(c, 0, sizeof(*c));
  c->T = T;
  c->stop_memory = stop_memory;
  pthread_mutex_init(&c->mu, NULL);
}
static void lc_free(level_cache* c) {
  for (int i = 0; i < ORC_MAX_LEVEL + 64; ++i) free(c->lev[i]);
  pthread_mutex_destroy(&c->mu);
}
static const double* lc_get(level_cache* c, int L) {
  pthread_mutex_lock(&c->mu);
  if (!c->lev[L]) {
    long np = (long) pow(2, L);
    double* buf = (double*) malloc(10 * np * sizeof(double));
    orc_level_samples(c->T, L, L > c->stop_memory, buf);
    c->lev[L] = buf;
  }
  const double* r = c->lev[L];
  pthread_mutex_unlock(&c->mu);
  return r;
}

/* ---------------------------------------------------------------- LW sums */
typedef struct {
  /* inputs */
  const orc_traj* T;
  const orc_particle* P;
  level_cache* lc;
  const orc_opts* o;
  const double* pol6;
  int use_pol;
  int stop_memory, stop_ext;
  /* spectrum: one obs, omega per point; flux: obs per point, one omega */
  const double* obs_single;
  const double* obs_list;
  const double* omega_list;
  double omega_single;
  double C0, C2;
  /* power density */
  const double* normal;
  const double* normal_list; /* per-point normals (surfaces of points), or NULL */
  int has_normal;            /* 0: surface without normals, N1.Normal taken as 1 */
  int directional;
  double pd_const;
  /* outputs */
  double* out_value;
  int* out_level;
  unsigned char* out_conv;
  double* out_e;
  /* thread partition */
  long i0, i1, stride;
  int mode; /* 0 spectrum/flux, 1 power */
} lw_job;

/* One output point of the spectrum / flux sums; src/OSCARSSR.cc:1939-2038 and
 * 4167-4266.  Result_Precision / Result_Level live in *rp / *rl so the caller
 * decides whether they carry over between points (the reference carries them
 * within one thread). */
static void lw_point(lw_job* J, const double* obs, double Omega, long i, double* rp, int* rl) {
  const orc_traj* T = J->T;
  double SumE[6] = {0, 0, 0, 0, 0, 0};
  double ThisMag = -1, LastMag = -1, ThisPhase = -1, LastPhase = -1, MaxDPhase = 0;
  int LastLevel = 0;
  int converged_here = 0;
  (void) ThisPhase;
  for (int L = 0; L <= J->stop_ext; ++L) {
    LastLevel = L;
    const double* S = lc_get(J->lc, L);
    long NT = (long) pow(2, L);
    MaxDPhase = 0;
    for (long iT = 0; iT != NT; ++iT) {
      const double* s = S + 10 * iT;
      double Time = s[0];
      const double* X = s + 1;
      const double* B = s + 4;
      const double* A = s + 7;
      double R[3] = {obs[0] - X[0], obs[1] - X[1], obs[2] - X[2]};
      double N[3];
      v_unit(R, N);
      double D = sqrt(v_mag2(R));
      ThisPhase = -Omega * (Time + D / K_C);
      double PhaseTestValue = fabs(ThisPhase - LastPhase);
      if (iT != 0 && PhaseTestValue > MaxDPhase) MaxDPhase = PhaseTestValue;
      LastPhase = ThisPhase;
      double complex ex = cexp(CMPLX(0, ThisPhase));
      double cr = creal(ex), ci = cimag(ex);
      double s1 = 1 - v_mag2(B);
      double NmB[3] = {N[0] - B[0], N[1] - B[1], N[2] - B[2]};
      double om = 1 - v_dot(N, B);
      double om2 = om * om;
      double den1 = D * D * om2;
      double c1[3], c2[3];
      v_cross(NmB, A, c1);
      v_cross(N, c1, c2);
      double den2 = D * om2;
      for (int k = 0; k < 3; ++k) {
        double v = NmB[k] * s1 / den1 + c2[k] / den2;
        SumE[2 * k] += v * cr;
        SumE[2 * k + 1] += v * ci;
      }
    }
    double dt = orc_dt_inclusive(T, L);
    double E[6];
    for (int k = 0; k < 6; ++k) E[k] = SumE[k] * dt;
    if (J->use_pol) pol_project(E, J->pol6);
    ThisMag = cvec_mag2(E);
    *rp = fabs(ThisMag - LastMag) / LastMag;
    if ((L > 8 && *rp < J->o->precision && MaxDPhase < K_PI) ||
        (L > 8 && MaxDPhase < K_PI && ThisMag == LastMag)) {
      *rl = L;
      converged_here = 1;
      break;
    }
    LastMag = ThisMag;
  }
  double f = J->C0 * orc_dt_inclusive(T, LastLevel);
  for (int k = 0; k < 6; ++k) SumE[k] *= f;
  if (J->use_pol) pol_project(SumE, J->pol6);
  double v;
  if (J->o->quantity == 1) v = *rp * J->o->weight;
  else if (J->o->quantity == 2) v = ((double) *rl) * J->o->weight;
  else v = J->C2 * cvec_mag2(SumE) * J->o->weight;
  J->out_value[i] = v;
  if (J->out_level) J->out_level[i] = J->o->carry_quirk ? *rl : (converged_here ? LastLevel : -1);
  if (J->out_conv) J->out_conv[i] = J->o->carry_quirk ? (*rl != -1) : (unsigned char) converged_here;
  if (J->out_e) memcpy(J->out_e + 6 * i, SumE, sizeof(SumE));
}

/* src/OSCARSSR.cc:2843-2966 */
static void pd_point(lw_job* J, const double* Obs, const double* Normal, long i, double* rp, int* rl) {
  const orc_traj* T = J->T;
  double ThisSum = -1, LastSum = -1;
  int LastLevel = 0;
  double Sum = 0;
  int converged_here = 0;
  (void) ThisSum;
  for (int L = 0; L <= J->stop_ext; ++L) {
    LastLevel = L;
    double Last_Beta[3] = {0, 0, 0};
    double BetaDiffMax = -1;
    const double* S = lc_get(J->lc, L);
    long NT = (long) pow(2, L);
    for (long iT = 0; iT != NT; ++iT) {
      const double* s = S + 10 * iT;
      const double* X = s + 1;
      const double* B = s + 4;
      const double* A = s + 7;
      double dB[3] = {B[0] - Last_Beta[0], B[1] - Last_Beta[1], B[2] - Last_Beta[2]};
      double BetaDiff = sqrt(v_mag2(dB));
      if (iT > 0 && BetaDiff > BetaDiffMax) BetaDiffMax = BetaDiff;
      Last_Beta[0] = B[0]; Last_Beta[1] = B[1]; Last_Beta[2] = B[2];
      double R[3] = {Obs[0] - X[0], Obs[1] - X[1], Obs[2] - X[2]};
      double N1[3], N2[3], N3[3], o[3], c[3];
      v_unit(R, N1);
      {
        /* include/TVector3D.h:132-143 Orthogonal() */
        double xx = N1[0] < 0.0 ? -N1[0] : N1[0];
        double yy = N1[1] < 0.0 ? -N1[1] : N1[1];
        double zz = N1[2] < 0.0 ? -N1[2] : N1[2];
        if (xx < yy) {
          if (xx < zz) { o[0] = 0; o[1] = N1[2]; o[2] = -N1[1]; }
          else { o[0] = N1[1]; o[1] = -N1[0]; o[2] = 0; }
        } else {
          if (yy < zz) { o[0] = -N1[2]; o[1] = 0; o[2] = N1[0]; }
          else { o[0] = N1[1]; o[1] = -N1[0]; o[2] = 0; }
        }
      }
      v_unit(o, N2);
      v_cross(N1, N2, c);
      v_unit(c, N3);
      /* src/OSCARSSR.cc:2846, 2908: without normals N1.Normal is taken as 1 */
      double N1DotNormal = J->has_normal ? v_dot(N1, Normal) : 1;
      if (J->directional && N1DotNormal <= 0) continue;
      double NmB[3] = {N1[0] - B[0], N1[1] - B[1], N1[2] - B[2]};
      double c1[3], Num[3];
      v_cross(NmB, A, c1);
      v_cross(N1, c1, Num);
      double Denominator = pow(1 - v_dot(B, N1), 5);
      double d2 = v_dot(Num, N2), d3 = v_dot(Num, N3);
      double r2 = v_mag2(R);
      Sum += d2 * d2 / Denominator / r2 * N1DotNormal;
      Sum += d3 * d3 / Denominator / r2 * N1DotNormal;
    }
    double ThisSumL = Sum * orc_dt_inclusive(T, L);
    *rp = fabs(ThisSumL - LastSum) / LastSum;
    if ((L > 8 && *rp < J->o->precision && BetaDiffMax < 2. / (J->P->gamma)) ||
        (L > 8 && BetaDiffMax < 2. / (J->P->gamma) && ThisSumL == LastSum)) {
      *rl = L;
      LastSum = ThisSumL;
      converged_here = 1;
      break;
    } else if (L > 8 && ThisSumL == LastSum) {
      *rl = L;
      *rp = 0;
      converged_here = 1;
      break;
    }
    LastSum = ThisSumL;
  }
  Sum *= J->pd_const * orc_dt_inclusive(T, LastLevel);
  Sum /= 1e6;
  if (!J->directional) {
    if (Sum < 0) Sum *= -1;
  }
  double v;
  if (J->o->quantity == 1) v = *rp * J->o->weight;
  else if (J->o->quantity == 2) v = ((double) *rl) * J->o->weight;
  else v = Sum * J->o->weight;
  J->out_value[i] = v;
  if (J->out_level) J->out_level[i] = J->o->carry_quirk ? *rl : (converged_here ? LastLevel : -1);
  if (J->out_conv) J->out_conv[i] = J->o->carry_quirk ? (*rl != -1) : (unsigned char) converged_here;
}

static void* lw_thread(void* arg) {
  lw_job* J = (lw_job*) arg;
  double rp = -1;
  int rl = -1;
  for (long i = J->i0; i <= J->i1; i += J->stride) {
    if (!J->o->carry_quirk) { rp = -1; rl = -1; }
    if (J->mode == 1) {
      pd_point(J, J->obs_list + 3 * i, J->normal_list ? J->normal_list + 3 * i : J->normal, i, &rp, &rl);
    } else if (J->obs_list) {
      lw_point(J, J->obs_list + 3 * i, J->omega_single, i, &rp, &rl);
    } else {
      lw_point(J, J->obs_single, J->omega_list[i], i, &rp, &rl);
    }
  }
  return NULL;
}

static void level_stops(const orc_opts* o, int* stop_mem, int* stop_ext) {
  int m = o->max_level >= -1 && o->max_level <= ORC_MAX_LEVEL ? o->max_level : ORC_MAX_LEVEL;
  *stop_mem = m;
  *stop_ext = o->max_level_extended > m ? o->max_level_extended : m;
}

/* Runs J over n points with the reference's thread partition: strided for
 * the spectrum (src/OSCARSSR.cc:2318-2348), contiguous ranges otherwise
 * (src/OSCARSSR.cc:4320-4328). */
static int run_partitioned(lw_job* base, long npoints, int strided) {
  int nt = base->o->nthreads < 1 ? 1 : base->o->nthreads;
  long NTA = npoints > nt ? nt : npoints;
  if (NTA < 1) return 0;
  lw_job* jobs = (lw_job*) malloc(NTA * sizeof(lw_job));
  pthread_t* th = (pthread_t*) malloc(NTA * sizeof(pthread_t));
  long NPer = npoints / NTA, NRem = npoints % NTA;
  for (long it = 0; it < NTA; ++it) {
    jobs[it] = *base;
    if (strided) {
      jobs[it].i0 = it;
      jobs[it].i1 = npoints - 1;
      jobs[it].stride = NTA;
    } else {
      long iFirst = it < NRem ? NPer * it + it : NPer * it + NRem;
      long iLast = it < NRem ? iFirst + NPer : iFirst + NPer - 1;
      jobs[it].i0 = iFirst;
      jobs[it].i1 = iLast;
      jobs[it].stride = 1;
    }
  }
  if (NTA == 1) {
    lw_thread(&jobs[0]);
  } else {
    for (long it = 0; it < NTA; ++it) pthread_create(&th[it], NULL, lw_thread, &jobs[it]);
    for (long it = 0; it < NTA; ++it) pthread_join(th[it], NULL);
  }
  free(jobs);
  free(th);
  return 0;
}

static void job_common(lw_job* J, const orc_traj* T, const orc_particle* P, level_cache* lc, const orc_opts* o,
                       const double* pol6) {
  memset(J, 0, sizeof(*J));
  J->T = T;
  J->P = P;
  J->lc = lc;
  J->o = o;
  J->pol6 = pol6;
  J->use_pol = pol6 && pol_mag2(pol6) > 0.001;
  level_stops(o, &J->stop_memory, &J->stop_ext);
  J->C0 = P->q / (k_fourpi() * K_C * K_EPS0 * K_SQRT2PI);
}

int orc_spectrum(const orc_traj* T, const orc_particle* P, const double* obs, int n_omega, const double* energies_eV,
                 const double* pol6, const orc_opts* o, double* out_value, int* out_level, unsigned char* out_conv,
                 double* out_e) {
  lw_job J;
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  job_common(&J, T, P, &lc, o, pol6);
  J.C2 = k_fourpi() * fabs(P->current) / (K_H * fabs(P->q) * K_MU0 * K_C) * 1e-6 * 0.001;
  double* om = (double*) malloc(n_omega * sizeof(double));
  for (int i = 0; i < n_omega; ++i) om[i] = ev_to_omega(energies_eV[i]);
  J.obs_single = obs;
  J.omega_list = om;
  J.out_value = out_value; J.out_level = out_level; J.out_conv = out_conv; J.out_e = out_e;
  run_partitioned(&J, n_omega, 1);
  free(om);
  lc_free(&lc);
  return 0;
}

int orc_flux(const orc_traj* T, const orc_particle* P, const double* obs_xyz, long n_obs, double energy_eV,
             const double* pol6, const orc_opts* o, double* out_value, int* out_level, unsigned char* out_conv,
             double* out_e) {
  lw_job J;
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  job_common(&J, T, P, &lc, o, pol6);
  J.C2 = k_fourpi() * P->current / (K_H * fabs(P->q) * K_MU0 * K_C) * 1e-6 * 0.001;
  J.obs_list = obs_xyz;
  J.omega_single = ev_to_omega(energy_eV);
  J.out_value = out_value; J.out_level = out_level; J.out_conv = out_conv; J.out_e = out_e;
  run_partitioned(&J, n_obs, 0);
  lc_free(&lc);
  return 0;
}

int orc_power_density(const orc_traj* T, const orc_particle* P, const double* obs_xyz, const double* normal,
                      long n_obs, int directional, const orc_opts* o, double* out_value, int* out_level,
                      unsigned char* out_conv) {
  lw_job J;
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  job_common(&J, T, P, &lc, o, NULL);
  J.mode = 1;
  J.obs_list = obs_xyz;
  J.normal = normal;
  J.normal_list = NULL;
  J.has_normal = 1;
  J.directional = directional;
  J.pd_const = fabs(P->q * P->current) / (16 * k_pi2() * K_EPS0 * K_C);
  J.out_value = out_value; J.out_level = out_level; J.out_conv = out_conv;
  run_partitioned(&J, n_obs, 0);
  lc_free(&lc);
  return 0;
}

/* Power density on a surface of points (src/OSCARSSR_Python.cc:6105-6380 and
 * 7355-7560; src/OSCARSSR.cc:2817-2978): normals[3 n] are unit normals per
 * point, or NULL for a surface without normals (a line of points). */
int orc_power_density_surface(const orc_traj* T, const orc_particle* P, const double* obs_xyz,
                              const double* normals, long n_obs, int directional, const orc_opts* o,
                              double* out_value, int* out_level, unsigned char* out_conv) {
  lw_job J;
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  job_common(&J, T, P, &lc, o, NULL);
  J.mode = 1;
  J.obs_list = obs_xyz;
  J.normal = NULL;
  J.normal_list = normals;
  J.has_normal = normals != NULL;
  J.directional = directional;
  J.pd_const = fabs(P->q * P->current) / (16 * k_pi2() * K_EPS0 * K_C);
  J.out_value = out_value; J.out_level = out_level; J.out_conv = out_conv;
  run_partitioned(&J, n_obs, 0);
  lc_free(&lc);
  return 0;
}

/* Total radiated power of one particle (src/OSCARSSR.cc:3725-3843): one
 * running sum over all levels of (|A/c|^2 - |B x A/c|^2) / (1 - B.B)^3.
 * Returns 0, or 1 when no level met the precision (the reference throws). */
int orc_total_power(const orc_traj* T, const orc_particle* P, const orc_opts* o, double* out_power,
                    double* out_precision, int* out_level) {
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  double Result_Precision = -1;
  int Result_Level = -1;
  double LastSum = -1;
  double Sum = 0;
  for (int L = 0; L <= se; ++L) {
    double Last_Beta[3] = {0, 0, 0};
    double BetaDiffMax = -1;
    const double* S = lc_get(&lc, L);
    long NT = (long) pow(2, L);
    for (long iT = 0; iT != NT; ++iT) {
      const double* B = S + 10 * iT + 4;
      const double* A = S + 10 * iT + 7;
      double Gamma6 = 1. / pow(1 - v_dot(B, B), 3);
      double dB[3] = {B[0] - Last_Beta[0], B[1] - Last_Beta[1], B[2] - Last_Beta[2]};
      double BetaDiff = sqrt(v_mag2(dB));
      if (iT > 0 && BetaDiff > BetaDiffMax) BetaDiffMax = BetaDiff;
      Last_Beta[0] = B[0]; Last_Beta[1] = B[1]; Last_Beta[2] = B[2];
      double c[3];
      v_cross(B, A, c);
      Sum += (v_mag2(A) - v_mag2(c)) * Gamma6;
    }
    double ThisSum = Sum * orc_dt_inclusive(T, L);
    Result_Precision = fabs(ThisSum - LastSum) / LastSum;
    if (L > 8 && Result_Precision < o->precision && BetaDiffMax < 2. / (P->gamma)) {
      Result_Level = L;
      LastSum = ThisSum;
      break;
    } else if (L > 8 && ThisSum == LastSum) {
      Result_Level = L;
      Result_Precision = 0;
      break;
    }
    LastSum = ThisSum;
  }
  lc_free(&lc);
  *out_power = LastSum * fabs(P->q * P->current) / (6 * K_PI * K_EPS0 * K_C);
  if (out_precision) *out_precision = Result_Precision;
  if (out_level) *out_level = Result_Level;
  return Result_Level == -1 ? 1 : 0;
}

/* TTriangle3D::RayIntersectionDistance (src/TTriangle3D.cc:159-215), Moller and
 * Trumbore: distance along the (normalised) ray to triangle ABC, or -1. */
static double ray_triangle(const double* A, const double* B, const double* Cv, const double* O, const double* Din) {
  double D[3], e1[3], e2[3], p[3], tv[3], q[3];
  v_unit(Din, D);
  for (int k = 0; k < 3; ++k) {
    e1[k] = B[k] - A[k];
    e2[k] = Cv[k] - A[k];
    tv[k] = O[k] - A[k];
  }
  v_cross(D, e2, p);
  double det = v_dot(e1, p);
  if (det > -1e-12 && det < 1e-12) return -1;
  double inv_det = 1.0 / det;
  double u = v_dot(tv, p) * inv_det;
  if (u < 0.0 || u > 1.0) return -1;
  v_cross(tv, e1, q);
  double v = v_dot(D, q) * inv_det;
  if (v < 0.0 || u + v > 1.0) return -1;
  double t = v_dot(e2, q) * inv_det;
  if (t < 0) return -1;
  return t;
}

/* Power density at the centres of STL triangles (src/OSCARSSR.cc:3318-3527):
 * tri[15 n] = A, B, C, normal (as read), centre per triangle.  A trajectory
 * sample whose ray to the centre meets another triangle first is skipped. */
int orc_power_density_stl(const orc_traj* T, const orc_particle* P, const double* tri, long n, const orc_opts* o,
                          double* out_value, int* out_level, unsigned char* out_conv) {
  level_cache lc;
  int sm, se;
  level_stops(o, &sm, &se);
  lc_init(&lc, T, sm);
  double pd_const = fabs(P->q * P->current) / (16 * k_pi2() * K_EPS0 * K_C);
  double Result_Precision = -1;
  int Result_Level = -1;
  for (long i = 0; i < n; ++i) {
    const double* Obs = tri + 15 * i + 12;
    const double* Normal = tri + 15 * i + 9;
    double LastSum = -1;
    int LastLevel = 0, converged_here = 0;
    double Sum = 0;
    for (int L = 0; L <= se; ++L) {
      LastLevel = L;
      double Last_Beta[3] = {0, 0, 0};
      double BetaDiffMax = -1;
      const double* S = lc_get(&lc, L);
      long NT = (long) pow(2, L);
      for (long iT = 0; iT != NT; ++iT) {
        const double* X = S + 10 * iT + 1;
        const double* B = S + 10 * iT + 4;
        const double* A = S + 10 * iT + 7;
        double R[3] = {Obs[0] - X[0], Obs[1] - X[1], Obs[2] - X[2]};
        double dist = sqrt(v_mag2(R)), Ru[3];
        v_unit(R, Ru);
        int blocked = 0;
        for (long j = 0; j < n && !blocked; ++j) {
          if (j == i) continue;
          const double* tj = tri + 15 * j;
          double t = ray_triangle(tj, tj + 3, tj + 6, X, Ru);
          if (t > 0 && t < dist) blocked = 1;
        }
        if (blocked) continue;
        double dB[3] = {B[0] - Last_Beta[0], B[1] - Last_Beta[1], B[2] - Last_Beta[2]};
        double BetaDiff = sqrt(v_mag2(dB));
        if (iT > 0 && BetaDiff > BetaDiffMax) BetaDiffMax = BetaDiff;
        Last_Beta[0] = B[0]; Last_Beta[1] = B[1]; Last_Beta[2] = B[2];
        double N1[3], N2[3], N3[3], oo[3], c[3];
        v_unit(R, N1);
        {
          double xx = N1[0] < 0.0 ? -N1[0] : N1[0];
          double yy = N1[1] < 0.0 ? -N1[1] : N1[1];
          double zz = N1[2] < 0.0 ? -N1[2] : N1[2];
          if (xx < yy) {
            if (xx < zz) { oo[0] = 0; oo[1] = N1[2]; oo[2] = -N1[1]; }
            else { oo[0] = N1[1]; oo[1] = -N1[0]; oo[2] = 0; }
          } else {
            if (yy < zz) { oo[0] = -N1[2]; oo[1] = 0; oo[2] = N1[0]; }
            else { oo[0] = N1[1]; oo[1] = -N1[0]; oo[2] = 0; }
          }
        }
        v_unit(oo, N2);
        v_cross(N1, N2, c);
        v_unit(c, N3);
        double N1DotNormal = fabs(v_dot(N1, Normal));
        double NmB[3] = {N1[0] - B[0], N1[1] - B[1], N1[2] - B[2]};
        double c1[3], Num[3];
        v_cross(NmB, A, c1);
        v_cross(N1, c1, Num);
        double Denominator = pow(1 - v_dot(B, N1), 5);
        double d2 = v_dot(Num, N2), d3 = v_dot(Num, N3);
        double r2 = v_mag2(R);
        Sum += d2 * d2 / Denominator / r2 * N1DotNormal;
        Sum += d3 * d3 / Denominator / r2 * N1DotNormal;
      }
      double ThisSum = Sum * orc_dt_inclusive(T, L);
      Result_Precision = fabs(ThisSum - LastSum) / LastSum;
      if ((L > 8 && Result_Precision < o->precision && BetaDiffMax < 2. / (P->gamma)) ||
          (L > 8 && BetaDiffMax < 2. / (P->gamma) && ThisSum == LastSum)) {
        Result_Level = L;
        converged_here = 1;
        break;
      } else if (L > 8 && ThisSum == LastSum) {
        Result_Level = L;
        Result_Precision = 0;
        converged_here = 1;
        break;
      }
      LastSum = ThisSum;
    }
    Sum *= pd_const * orc_dt_inclusive(T, LastLevel);
    Sum /= 1e6;
    double v;
    if (o->quantity == 1) v = Result_Precision * o->weight;
    else if (o->quantity == 2) v = ((double) Result_Level) * o->weight;
    else v = Sum * o->weight;
    out_value[i] = v;
    if (out_level) out_level[i] = o->carry_quirk ? Result_Level : (converged_here ? LastLevel : -1);
    if (out_conv) out_conv[i] = (unsigned char) converged_here;
  }
  lc_free(&lc);
  return 0;
}

/* Electric field in the time domain at Obs, one value per knot
 * (src/OSCARSSR.cc:4471-4512); dt is the trajectory's step. */
void orc_efield_time(const orc_traj* T, const orc_particle* P, const double* Obs, double dt, double* out_t,
                     double* out_e) {
  double C0 = P->q / (k_fourpi() * K_EPS0);
  for (long iT = 0; iT != T->n; ++iT) {
    const double* X = T->p + 9 * iT;
    const double* B = X + 3;
    const double* A = X + 6;
    double R[3] = {Obs[0] - X[0], Obs[1] - X[1], Obs[2] - X[2]};
    double N[3];
    v_unit(R, N);
    double D = sqrt(v_mag2(R));
    double Mult = C0 * pow(1.0 / (1.0 - v_dot(N, B)), 3);
    double NmB[3] = {N[0] - B[0], N[1] - B[1], N[2] - B[2]};
    double r2 = v_mag2(R), b2 = v_mag2(B);
    double c1[3], c2[3];
    v_cross(NmB, A, c1);
    v_cross(N, c1, c2);
    for (int k = 0; k < 3; ++k) {
      double Near = ((1.0 - b2) * NmB[k]) / r2;
      double Far = ((1.0 / K_C) * c2[k]) / D;
      out_e[3 * iT + k] = Mult * (Near + Far);
    }
    out_t[iT] = (double) iT * dt + D / K_C;
  }
}

double orc_evals_from_levels(long n, const int* levels) {
  double s = 0;
  for (long i = 0; i < n; ++i) {
    int L = levels[i];
    if (L >= 0) s += pow(2., L + 1) - 1;
  }
  return s;
}
