// Field models, equations of motion, the RK4 step and the spline evaluation
// shared by the trajectory kernels (traj.hip) and the reverse pass (adjoint.hip).
//
// Reference behaviour restated (paths relative to the reference root):
//   derivatives src/OSCARSSR.cc:1247-1279, step src/OSCARSSR.cc:1321-1386
//   fields      src/TField3D_IdealUndulator.cc:113-152, src/TField3D_IdealEPU.cc:116-170,
//               src/TField3D_Grid.cc:153-278, sum src/TFieldContainer.cc:85-107
//   spline      include/TOMATH.h:92-126
#pragma once

#include "om_device.h"
#include "om_types.h"

namespace om {

// ---------------------------------------------------------------- fields
// Written without pointers into the field description or run-time array
// indices, so a single field held by value stays in registers.
__device__ __forceinline__ bool undulator_part(const FieldDev& f, double D, double PhaseShift, double a0, double a1,
                                               double a2, double taper_corr, double out[3], bool add) {
  double L = f.und_length, PL = f.period_length;
  if (D > L / 2. + PhaseShift || D < -L / 2. + PhaseShift) return false;
  double scale = 1;
  bool term = false;
  if (D < -L / 2. + PhaseShift + PL || D > L / 2. + PhaseShift - PL) {
    term = true;
    scale = (D < -L / 2. + PhaseShift + PL / 2. || D > L / 2. + PhaseShift - PL / 2.) ? 0.25 : 0.75;
  }
  double s = sin64(div_by(2.0 * kPi * (D - PhaseShift), PL, f.inv_period_length));
  double b0 = term ? a0 * scale : a0;
  double b1 = term ? a1 * scale : a1;
  double b2 = term ? a2 * scale : a2;
  double v0 = b0 * s * taper_corr, v1 = b1 * s * taper_corr, v2 = b2 * s * taper_corr;
  out[0] = add ? out[0] + v0 : v0;
  out[1] = add ? out[1] + v1 : v1;
  out[2] = add ? out[2] + v2 : v2;
  return true;
}

// cos(2 pi f (t + t0) + phase), src/TField3D_IdealUndulator.cc:151 (the same
// expression in every model)
__device__ __forceinline__ double td_factor(const FieldDev& f, double t) {
  return cos64(2.0 * kPi * f.freq * (t + f.toff) + f.fphase);
}

__device__ __forceinline__ void td_apply(const FieldDev& f, double t, double out[3]) {
  double c = td_factor(f, t);
  out[0] = out[0] * c;
  out[1] = out[1] * c;
  out[2] = out[2] * c;
}

__device__ __forceinline__ void rot_xyz(double v[3], const FieldDev& f) {
  // src/TVector3D.cc RotateSelfX, Y, Z with the angle's sin/cos precomputed on the host
  double s = f.rot_sin[0], c = f.rot_cos[0], yy = v[1];
  v[1] = c * yy - s * v[2];
  v[2] = s * yy + c * v[2];
  s = f.rot_sin[1]; c = f.rot_cos[1];
  double zz = v[2];
  v[2] = c * zz - s * v[0];
  v[0] = s * zz + c * v[0];
  s = f.rot_sin[2]; c = f.rot_cos[2];
  double xx = v[0];
  v[0] = c * xx - s * v[1];
  v[1] = s * xx + c * v[1];
}

// Cell index (size_t)((x - start) / step) of src/TField3D_Grid.cc:178-180 for a
// coordinate strictly inside the map: the quotient is positive and below the
// sample count, where truncation to int gives the same index in one
// conversion instead of the unsigned 64-bit sequence.
__device__ __forceinline__ int cell_of(double q) { return (int) q; }

// Two- and three-dimensional grids (src/TField3D_Grid.cc:174-268): the cell
// index per axis, then linear steps along x, y and z in the reference's order,
// each ((d * (hi - lo)) / step).  The index is kept inside the map.
__device__ __forceinline__ long grid_cell(const FieldDev& f, int a, double c, double& d) {
  if (f.n[a] <= 1) {
    d = 0;
    return 0;
  }
  int i = cell_of(div_by(c - f.start[a], f.step[a], f.inv_step[a]));
  if (i > f.n[a] - 2) i = (int) (f.n[a] - 2);
  d = (c - f.start[a]) - (double) i * f.step[a];
  return i;
}

__device__ __forceinline__ void lerp3(const double* lo, const double* hi, double d, double step, double istep,
                                      double* r) {
  for (int k = 0; k < 3; ++k) r[k] = lo[k] + div_by(d * (hi[k] - lo[k]), step, istep);
}

__device__ __noinline__ void grid_multi(const FieldDev& f, const double* X, double out[3]) {
  double dx, dy, dz;
  long nx = grid_cell(f, 0, X[0], dx), ny = grid_cell(f, 1, X[1], dy), nz = grid_cell(f, 2, X[2], dz);
  const long NY = f.n[1], NZ = f.n[2];
  auto at = [&](long ix, long iy, long iz) { return f.data + 3 * (ix * NY * NZ + iy * NZ + iz); };
  double v00[3], v10[3], v01[3], v11[3], v0[3], v1[3];
  if (f.dims == 7) {
    lerp3(at(nx, ny, nz), at(nx + 1, ny, nz), dx, f.step[0], f.inv_step[0], v00);
    lerp3(at(nx, ny + 1, nz), at(nx + 1, ny + 1, nz), dx, f.step[0], f.inv_step[0], v10);
    lerp3(at(nx, ny, nz + 1), at(nx + 1, ny, nz + 1), dx, f.step[0], f.inv_step[0], v01);
    lerp3(at(nx, ny + 1, nz + 1), at(nx + 1, ny + 1, nz + 1), dx, f.step[0], f.inv_step[0], v11);
    lerp3(v00, v10, dy, f.step[1], f.inv_step[1], v0);
    lerp3(v01, v11, dy, f.step[1], f.inv_step[1], v1);
    lerp3(v0, v1, dz, f.step[2], f.inv_step[2], out);
  } else if (f.dims == 3) {
    lerp3(at(nx, ny, 0), at(nx + 1, ny, 0), dx, f.step[0], f.inv_step[0], v0);
    lerp3(at(nx, ny + 1, 0), at(nx + 1, ny + 1, 0), dx, f.step[0], f.inv_step[0], v1);
    lerp3(v0, v1, dy, f.step[1], f.inv_step[1], out);
  } else if (f.dims == 5) {
    lerp3(at(nx, 0, nz), at(nx + 1, 0, nz), dx, f.step[0], f.inv_step[0], v0);
    lerp3(at(nx, 0, nz + 1), at(nx + 1, 0, nz + 1), dx, f.step[0], f.inv_step[0], v1);
    lerp3(v0, v1, dz, f.step[2], f.inv_step[2], out);
  } else {  // 6: y and z
    lerp3(at(0, ny, nz), at(0, ny + 1, nz), dy, f.step[1], f.inv_step[1], v0);
    lerp3(at(0, ny, nz + 1), at(0, ny + 1, nz + 1), dy, f.step[1], f.inv_step[1], v1);
    lerp3(v0, v1, dz, f.step[2], f.inv_step[2], out);
  }
}

// One model's value at (x, y, z) and time t [s].  The time factor is applied
// where the reference applies it: not to the zero an undulator, a uniform
// box, a quadrupole or a grid returns outside its region.
__device__ __forceinline__ void field_eval(const FieldDev& f, double x, double y, double z, double t,
                                           double out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (f.type == kUndulator || f.type == kEPU) {
    double D = (x - f.center[0]) * f.unit[0] + (y - f.center[1]) * f.unit[1] + (z - f.center[2]) * f.unit[2];
    double tc = 1 + D * f.taper;
    bool in = undulator_part(f, D, f.phase_shift_a, f.field[0], f.field[1], f.field[2], tc, out, false);
    if (f.type == kEPU) undulator_part(f, D, f.phase_shift_b, f.fieldB[0], f.fieldB[1], f.fieldB[2], tc, out, true);
    if (f.td && (in || f.type == kEPU)) td_apply(f, t, out);
    return;
  }
  double X[3] = {x, y, z};
  rot_xyz(X, f);
  if (f.type == kGaussian || f.type == kUniform) {
    // src/TField3D_Gaussian.cc:90-109, src/TField3D_UniformBox.cc:105-117: the
    // point rotated by the model's angles, relative to its centre
    double R0 = X[0] - f.center[0], R1 = X[1] - f.center[1], R2 = X[2] - f.center[2];
    if (f.type == kUniform) {
      if ((f.shape[0] > 0 && fabs(R0) > fabs(f.shape[0] / 2.)) || (f.shape[1] > 0 && fabs(R1) > fabs(f.shape[1] / 2.)) ||
          (f.shape[2] > 0 && fabs(R2) > fabs(f.shape[2] / 2.)))
        return;
      out[0] = f.field[0]; out[1] = f.field[1]; out[2] = f.field[2];
      if (f.td) td_apply(f, t, out);
      return;
    }
    double frac = 1;
    if (f.shape[0] > 0) { double u = R0 / f.shape[0]; frac *= exp(-(u * u) / 2.); }
    if (f.shape[1] > 0) { double u = R1 / f.shape[1]; frac *= exp(-(u * u) / 2.); }
    if (f.shape[2] > 0) { double u = R2 / f.shape[2]; frac *= exp(-(u * u) / 2.); }
    out[0] = frac * f.field[0]; out[1] = frac * f.field[1]; out[2] = frac * f.field[2];
    if (f.td) td_apply(f, t, out);
    return;
  }
  if (f.type == kQuadrupole) {
    // src/TField3D_Quadrupole.cc:66-80: (K y, K x, 0) inside the width, rotated back
    X[0] -= f.trans[0]; X[1] -= f.trans[1]; X[2] -= f.trans[2];
    if (fabs(X[2]) > f.width / 2) return;
    double V[3] = {f.K * X[1], f.K * X[0], 0};
    rot_xyz(V, f);
    out[0] = V[0]; out[1] = V[1]; out[2] = V[2];
    if (f.td) td_apply(f, t, out);
    return;
  }
  X[0] -= f.trans[0]; X[1] -= f.trans[1]; X[2] -= f.trans[2];
  if (f.n[0] > 1 && (X[0] <= f.start[0] || X[0] >= f.stop[0])) return;
  if (f.n[1] > 1 && (X[1] <= f.start[1] || X[1] >= f.stop[1])) return;
  if (f.n[2] > 1 && (X[2] <= f.start[2] || X[2] >= f.stop[2])) return;
  if (f.axis < 0) {
    grid_multi(f, X, out);
    if (f.td) td_apply(f, t, out);
    return;
  }
  int a = f.axis;
  double coord = a == 0 ? X[0] : (a == 1 ? X[1] : X[2]);
  double start = a == 0 ? f.start[0] : (a == 1 ? f.start[1] : f.start[2]);
  double step = a == 0 ? f.step[0] : (a == 1 ? f.step[1] : f.step[2]);
  double istep = a == 0 ? f.inv_step[0] : (a == 1 ? f.inv_step[1] : f.inv_step[2]);
  long na = a == 0 ? f.n[0] : (a == 1 ? f.n[1] : f.n[2]);
  int n0 = cell_of(div_by(coord - start, step, istep));
  if (n0 > na - 2) n0 = (int) (na - 2);  // keeps the n0+1 read inside the map
  double dd = (coord - start) - (double) n0 * step;
  const double* lo = f.data + 3 * n0;
  double l0 = lo[0], l1 = lo[1], l2 = lo[2], h0 = lo[3], h1 = lo[4], h2 = lo[5];
  // src/TField3D_Grid.cc:277: the factor cos(2 pi f (t + t0) + phase), 1 without time dependence
  const double c = f.td ? td_factor(f, t) : 1.0;
  out[0] = (l0 + div_by((h0 - l0) * dd, step, istep)) * c;
  out[1] = (l1 + div_by((h1 - l1) * dd, step, istep)) * c;
  out[2] = (l2 + div_by((h2 - l2) * dd, step, istep)) * c;
}

__device__ __forceinline__ void get_b(const FieldDev* __restrict__ F, int nf, double x, double y, double z, double t,
                                      double B[3]) {
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = 0; i < nf; ++i) {
    double v[3];
    field_eval(F[i], x, y, z, t, v);
    s0 += v[0]; s1 += v[1]; s2 += v[2];
  }
  B[0] = s0; B[1] = s1; B[2] = s2;
}

// ---------------------------------------------------------------- compact undulator / EPU
// The few numbers an ideal undulator or EPU needs, small enough to stay in
// registers for the whole integration.  The region tests become selects:
// outside the magnet the amplitude is multiplied by zero, which gives the same
// sum as the reference's early return.
struct UndParams {
  double c0, c1, c2, u0, u1, u2;  // centre, unit vector along the period
  double taper, half_len, PL, inv_PL, half_PL;
  double psA, psB;
  double a0, a1, a2, b0, b1, b2;  // field A, field B (zero for a planar undulator)
  int epu;
  // The region thresholds of the two parts (phase shifts psA, psB), formed on
  // the host in the reference's operation order (src/TField3D_IdealUndulator.cc:130-137):
  // -L/2+PS, L/2+PS, (-L/2+PS)+PL, (L/2+PS)-PL, (-L/2+PS)+PL/2, (L/2+PS)-PL/2
  double thA[6], thB[6];
};

// A copy of a wave-uniform double held in vector registers.  The integration
// is one lane's dependent chain with the vector registers mostly free, while
// the scalar registers cannot hold every constant of the field: kept there, the
// constants spill into vector lanes and each use costs two v_readlane.  The
// identity DPP move makes the compiler treat the value as per-lane.
__device__ __forceinline__ double in_vgpr(double x) {
  long long b = __builtin_bit_cast(long long, x);
  int lo = __builtin_amdgcn_update_dpp(0, (int) (b & 0xffffffff), 0xE4, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), 0xE4, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long) hi << 32) | (unsigned) lo);
}

// The value of x in the neighbouring lane (lanes 2k and 2k+1 exchange), by a
// DPP quad permutation of both halves: a move, exact.
__device__ __forceinline__ double lane_pair_other(double x) {
  long long b = __builtin_bit_cast(long long, x);
  int lo = __builtin_amdgcn_update_dpp(0, (int) (b & 0xffffffff), 0xB1, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long) hi << 32) | (unsigned) lo);
}

__device__ __forceinline__ UndParams und_in_vgpr(const UndParams& f) {
  UndParams v = f;
  v.c0 = in_vgpr(f.c0); v.c1 = in_vgpr(f.c1); v.c2 = in_vgpr(f.c2);
  v.u0 = in_vgpr(f.u0); v.u1 = in_vgpr(f.u1); v.u2 = in_vgpr(f.u2);
  v.taper = in_vgpr(f.taper); v.half_len = in_vgpr(f.half_len); v.PL = in_vgpr(f.PL);
  v.inv_PL = in_vgpr(f.inv_PL); v.half_PL = in_vgpr(f.half_PL);
  v.psA = in_vgpr(f.psA); v.psB = in_vgpr(f.psB);
  v.a0 = in_vgpr(f.a0); v.a1 = in_vgpr(f.a1); v.a2 = in_vgpr(f.a2);
  v.b0 = in_vgpr(f.b0); v.b1 = in_vgpr(f.b1); v.b2 = in_vgpr(f.b2);
  for (int k = 0; k < 6; ++k) {
    v.thA[k] = in_vgpr(f.thA[k]);
    v.thB[k] = in_vgpr(f.thB[k]);
  }
  return v;
}

// 0 outside the magnet, 1/4 and 3/4 on the terminating half periods, else 1.
__device__ __forceinline__ double und_scale(const double th[6], double D) {
  bool out = D > th[1] || D < th[0];
  bool term = D < th[2] || D > th[3];
  bool quarter = D < th[4] || D > th[5];
  double sc = term ? (quarter ? 0.25 : 0.75) : 1.0;
  return out ? 0.0 : sc;
}

__device__ __forceinline__ void und_eval(const UndParams& f, double x, double y, double z, double B[3]) {
  double D = (x - f.c0) * f.u0 + (y - f.c1) * f.u1 + (z - f.c2) * f.u2;
  double tc = 1 + D * f.taper;
  double scA = und_scale(f.thA, D);
  double sA = sin64(div_by(2.0 * kPi * (D - f.psA), f.PL, f.inv_PL));
  // terminated poles: (field*scale)*sin*taper; full poles: field*sin*taper (scale 1 is exact)
  double v0 = f.a0 * scA * sA * tc, v1 = f.a1 * scA * sA * tc, v2 = f.a2 * scA * sA * tc;
  if (f.epu) {
    double scB = und_scale(f.thB, D);
    double sB = sin64(div_by(2.0 * kPi * (D - f.psB), f.PL, f.inv_PL));
    v0 = v0 + f.b0 * scB * sB * tc;
    v1 = v1 + f.b1 * scB * sB * tc;
    v2 = v2 + f.b2 * scB * sB * tc;
  }
  B[0] = 0.0 + v0;
  B[1] = 0.0 + v1;
  B[2] = 0.0 + v2;
}

// ---------------------------------------------------------------- RK4
// A single 1-D grid along z without rotation, with its samples staged in LDS
// (kGridZ variant of the equations of motion).
struct GridZ {
  const double* data;  // 3 per sample (LDS when staged)
  long n;
  double start, stop, step, inv_step, trans;
};

struct RkCtx {
  const FieldDev* F;
  int nf;
  double qomg;
  int err;
  GridZ g;
  int nfe;     // electric fields, stored after the nf magnetic ones in F
  double qom;  // Q / M
  int halved;  // steps halved so far (src/OSCARSSR.cc:1362-1378)
  // kGridZ: the samples of the last cell read from LDS (the cell changes every
  // few steps, so most stages skip the LDS round trip on their serial chain)
  int gc_n0 = -1;
  double gc[6] = {0, 0, 0, 0, 0, 0};
};

// Field-model variants of the equations of motion:
//   M = kFieldsGeneric  fields read from F (any number, any type)
//   M = kOneUndulator   exactly one undulator/EPU, passed by value in f0 (registers)
//   M = kAxisUndX/Y/Z   one planar undulator with its period along z, no taper and
//                       a field along x, y or z only.  The general formula then
//                       reduces exactly: D = (z - c_z) (the other terms of the dot
//                       product are zeros), the taper factor is 1, and the two zero
//                       field components contribute zeros to v x B.  Only the sign
//                       of an exactly-zero result can differ, which never changes a
//                       nonzero value downstream.  The |beta| >= 1 test is left to
//                       the end of the step (rk4), where it is checked on the result.
//   M = kGridZ          one 1-D grid along z without rotation, samples in LDS: the
//                       grid branch of field_eval with the identity rotation and the
//                       x/y range tests (n = 1 there) removed, which changes no value.
//   M = kUndZ / kEpuZ   one undulator (kUndZ) or EPU (kEpuZ) with its period along
//                       z and no taper, fields in any direction: und_eval with
//                       D = z - c_z and the taper factor 1, both exact (the dot
//                       product's other terms are zeros, 1 + D 0 = 1), branch-free
//                       like the axis variants.  kEpuZ runs in lane pairs, each
//                       lane evaluating one of the two parts (rk4_kernel).
constexpr int kFieldsGeneric = 0, kOneUndulator = 1, kAxisUndX = 2, kAxisUndY = 3, kAxisUndZ = 4, kGridZ = 5,
              kFieldsEB = 6, kUndZ = 7, kEpuZ = 8;

__device__ __forceinline__ void grid_z_eval(const GridZ& g, int& cn0, double cs[6], double z, double B[3]) {
  double coord = z - g.trans;
  if (coord <= g.start || coord >= g.stop) {
    B[0] = B[1] = B[2] = 0;
    return;
  }
  int n0 = cell_of(div_by(coord - g.start, g.step, g.inv_step));
  if (n0 > g.n - 2) n0 = (int) (g.n - 2);
  if (n0 != cn0) {
    const double* lo = g.data + 3 * n0;
#pragma unroll
    for (int k = 0; k < 6; ++k) cs[k] = lo[k];
    cn0 = n0;
  }
  double dd = (coord - g.start) - (double) n0 * g.step;
#pragma unroll
  for (int k = 0; k < 3; ++k) B[k] = cs[k] + div_by((cs[3 + k] - cs[k]) * dd, g.step, g.inv_step);
}

// |beta| >= 1 test of src/OSCARSSR.cc:1260-1264 and 1368,
//   1. - (v1 v1 + v3 v3 + v5 v5) / (c c) <= 0,
// decided exactly without the division: with q the same sum of squares and
// b = RN(c c), RN(q / b) >= 1 exactly when q / b >= 1 - 2^-54 (the midpoint
// below 1 rounds to 1), i.e. when q - b >= -b 2^-54, and q - b is exact
// wherever it decides the outcome (q within a factor 2 of b).
constexpr double kC2 = kC * kC;
constexpr double kBetaOneMargin = -kC2 * 5.551115123125783e-17;  // -b 2^-54

__device__ __forceinline__ bool beta_at_least_one(double v1, double v3, double v5) {
  double q = v1 * v1 + v3 * v3 + v5 * v5;
  return q - kC2 >= kBetaOneMargin;
}

// Equations of motion at time t [s] (the field models read t only when they
// are time-dependent).  A state with |beta| >= 1 sets c.err and leaves d as it
// was (src/OSCARSSR.cc:1260-1264); with electric fields the reference sets the
// flag and still writes d.  Here every variant sets the flag and writes d: a
// flagged step is discarded whole (rk4 redoes it by halving from its start,
// with the slope recomputed there), so no kept value depends on d then.
template <int M>
__device__ __forceinline__ void derivs(RkCtx& c, const UndParams& f0, double t, const double* x, double* d) {
  if (M >= kAxisUndX && M <= kAxisUndZ) {
    // Branch-free, so the |beta| test and the magnet's extent are evaluated
    // beside the sine rather than before it on the serial chain.  Outside the
    // magnet the field is an exact zero (the reference returns zero there).
    constexpr int K = M - kAxisUndX;
    const bool bad = beta_at_least_one(x[1], x[3], x[5]);
    double D = x[4] - f0.c2;
    double sc = und_scale(f0.thA, D);
    double a = K == 0 ? f0.a0 : (K == 1 ? f0.a1 : f0.a2);
    double s = sin64(div_by(2.0 * kPi * (D - f0.psA), f0.PL, f0.inv_PL));
    double v = sc != 0.0 ? a * sc * s : 0.0;
    double n1, n3, n5;
    if (K == 0) {
      n1 = 0.0;
      n3 = c.qomg * (x[5] * v);
      n5 = c.qomg * (-(x[3] * v));
    } else if (K == 1) {
      n1 = c.qomg * (-(x[5] * v));
      n3 = 0.0;
      n5 = c.qomg * (x[1] * v);
    } else {
      n1 = c.qomg * (x[3] * v);
      n3 = c.qomg * (-(x[1] * v));
      n5 = 0.0;
    }
    // A state with |beta| >= 1 sets the flag, which discards the whole step
    // (rk4 redoes it by halving from its start), so d need not keep its old
    // values as the reference's derivatives do: no kept result depends on them.
    d[0] = x[1];
    d[1] = n1;
    d[2] = x[3];
    d[3] = n3;
    d[4] = x[5];
    d[5] = n5;
    c.err |= bad ? 1 : 0;
    return;
  }
  if (M == kUndZ || M == kEpuZ) {
    const bool bad = beta_at_least_one(x[1], x[3], x[5]);
    double D = x[4] - f0.c2;
    double scA = und_scale(f0.thA, D);
    double sA = sin64(div_by(2.0 * kPi * (D - f0.psA), f0.PL, f0.inv_PL));
    double v0 = f0.a0 * scA * sA, v1 = f0.a1 * scA * sA, v2 = f0.a2 * scA * sA;
    if (M == kEpuZ) {
      // Two lanes integrate the same state: one holds part A in f0's A slots,
      // the other part B there (rk4_kernel swaps them).  Each adds the other's
      // part to its own, A + B in one lane and B + A in the other, the same
      // double (IEEE addition commutes), which is the reference's A + B.
      v0 = v0 + lane_pair_other(v0);
      v1 = v1 + lane_pair_other(v1);
      v2 = v2 + lane_pair_other(v2);
    }
    double B0 = 0.0 + v0, B1 = 0.0 + v1, B2 = 0.0 + v2;
    double n1 = c.qomg * (-x[5] * B1 + x[3] * B2);
    double n3 = c.qomg * (x[5] * B0 - x[1] * B2);
    double n5 = c.qomg * (x[1] * B1 - x[3] * B0);
    // A state with |beta| >= 1 sets the flag, which discards the whole step
    // (rk4 redoes it by halving from its start), so d need not keep its old
    // values as the reference's derivatives do: no kept result depends on them.
    d[0] = x[1];
    d[1] = n1;
    d[2] = x[3];
    d[3] = n3;
    d[4] = x[5];
    d[5] = n5;
    c.err |= bad ? 1 : 0;
    return;
  }
  if (M == kFieldsEB) {
    // src/OSCARSSR.cc:1283-1318: magnetic and electric fields, the factor
    // Q/M sqrt(1 - beta^2) taken at each evaluation
    if (beta_at_least_one(x[1], x[3], x[5])) c.err = 1;
    double bb = (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) / (kC * kC);
    double B[3], E[3];
    get_b(c.F, c.nf, x[0], x[2], x[4], t, B);
    get_b(c.F + c.nf, c.nfe, x[0], x[2], x[4], t, E);
    double q = c.qom * sqrt(1. - bb);
    double be = (x[1] * E[0] + x[3] * E[1] + x[5] * E[2]) / kC;
    d[0] = x[1];
    d[1] = q * (E[0] - x[5] * B[1] + x[3] * B[2] - x[1] * be / kC);
    d[2] = x[3];
    d[3] = q * (E[1] + x[5] * B[0] - x[1] * B[2] - x[3] * be / kC);
    d[4] = x[5];
    d[5] = q * (E[2] + x[1] * B[1] - x[3] * B[0] - x[5] * be / kC);
    return;
  }
  // the flag discards the whole step (see above), so d is written regardless
  const bool bad = beta_at_least_one(x[1], x[3], x[5]);
  double B[3];
  if (M == kGridZ) {
    grid_z_eval(c.g, c.gc_n0, c.gc, x[4], B);
  } else if (M == kOneUndulator) {
    und_eval(f0, x[0], x[2], x[4], B);
  } else {
    get_b(c.F, c.nf, x[0], x[2], x[4], t, B);
  }
  d[0] = x[1];
  d[1] = c.qomg * (-x[5] * B[1] + x[3] * B[2]);
  d[2] = x[3];
  d[3] = c.qomg * (x[5] * B[0] - x[1] * B[2]);
  d[4] = x[5];
  d[5] = c.qomg * (x[1] * B[1] - x[3] * B[0]);
  c.err |= bad ? 1 : 0;
}

// One classic RK4 step from (t, y) with slope dydx into yo
// (src/OSCARSSR.cc:1321-1360); returns true when |beta| >= 1 was met in a
// stage or in the result (the step must then be halved).
template <int M>
__device__ __forceinline__ bool rk4_try(RkCtx& c, const UndParams& f0, const double* y, const double* dydx,
                                        double t, double h, double* yo) {
  double dym[6], dyt[6], yt[6];
  double hh = h * 0.5;
  double h6 = h / 6.0;
  double xh = t + hh;
  for (int i = 0; i < 6; ++i) dyt[i] = dydx[i];  // a stage with |beta| >= 1 keeps earlier values
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  derivs<M>(c, f0, xh, yt, dyt);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  for (int i = 0; i < 6; ++i) dym[i] = dyt[i];
  derivs<M>(c, f0, xh, yt, dym);
  for (int i = 0; i < 6; ++i) {
    yt[i] = y[i] + h * dym[i];
    dym[i] += dyt[i];
  }
  derivs<M>(c, f0, t + h, yt, dyt);
  for (int i = 0; i < 6; ++i) yo[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  return c.err || beta_at_least_one(yo[1], yo[3], yo[5]);
}

// Deepest halving of one step (h / 2^kMaxHalving); the reference recurses
// without limit, here the integration stops with an error status instead.
constexpr int kMaxHalving = 40;

// State and slope of one integration, passed by value to the halving walk so
// that the integration loop never takes their address (they stay in registers).
struct RkPair {
  double y[6], d[6];
  int err, halved, ok;
};

// src/OSCARSSR.cc:1362-1378: a step that met |beta| >= 1 is split in two
// halves, recursively.  The recursion is walked with an explicit stack of the
// sub-steps' start times and lengths: a failed sub-step recomputes the slope at
// its start and descends into its first half; a completed first half
// recomputes the slope at the midpoint and continues with the second half.
// The slope ends as the last one computed, as the reference's does.  ok = 0
// when the depth limit is reached.  Out of line: it runs only for steps that
// failed, and everything it needs arrives by value.
template <int M>
__device__ __noinline__ RkPair rk4_halving(RkCtx c, const UndParams f0, RkPair s, double t, double h) {
  double sx[kMaxHalving + 1], sh[kMaxHalving + 1];
  bool second[kMaxHalving + 1];
  sx[0] = t;
  sh[0] = h;
  second[0] = false;
  int d = 0;  // the failed sub-step
  s.ok = 1;
  while (true) {
    // split sub-step d: slope at its start, then its first half
    c.err = 0;
    derivs<M>(c, f0, sx[d], s.y, s.d);
    if (d == kMaxHalving) {
      s.ok = 0;
      break;
    }
    sx[d + 1] = sx[d];
    sh[d + 1] = sh[d] / 2.;
    second[d + 1] = false;
    ++d;
    bool done = false;
    while (true) {
      double yo[6];
      if (rk4_try<M>(c, f0, s.y, s.d, sx[d], sh[d], yo)) break;  // sub-step d fails: split it
      for (int i = 0; i < 6; ++i) s.y[i] = yo[i];
      while (d > 0 && second[d]) --d;  // a completed second half completes its parent
      if (d == 0) {
        done = true;
        break;
      }
      // sub-step d was a first half: its sibling starts at parent start + parent length / 2
      sx[d] = sx[d - 1] + sh[d - 1] / 2.;
      second[d] = true;
      derivs<M>(c, f0, sx[d], s.y, s.d);
    }
    if (done) break;
  }
  s.err = c.err;
  return s;
}

// RK4 step with the reference's halving; returns false only when the halving
// depth limit was reached.
template <int M>
__device__ __forceinline__ bool rk4(RkCtx& c, const UndParams& f0, double* y, double* dydx, double t, double h) {
  double yo[6];
  if (rk4_try<M>(c, f0, y, dydx, t, h, yo)) {
    ++c.halved;
    RkPair s;
    for (int i = 0; i < 6; ++i) {
      s.y[i] = y[i];
      s.d[i] = dydx[i];
    }
    s = rk4_halving<M>(c, f0, s, t, h);
    for (int i = 0; i < 6; ++i) {
      y[i] = s.y[i];
      dydx[i] = s.d[i];
    }
    c.err = s.err;
    return s.ok != 0;
  }
  for (int i = 0; i < 6; ++i) y[i] = yo[i];
  return true;
}

// Knot interval of x: a first guess from the (nearly) uniform knot grid, then
// moved to the reference bisection's answer (largest klo <= n-2 with
// t[klo] <= x), which is exact for increasing knots (include/TOMATH.h:92-126).
__device__ __forceinline__ long spline_interval(const double* __restrict__ X, long n, double x) {
  double h0 = (X[n - 1] - X[0]) / (double) (n - 1);
  long k = (long) ((x - X[0]) / h0);
  if (k < 0) k = 0;
  if (k > n - 2) k = n - 2;
  while (k < n - 2 && X[k + 1] <= x) ++k;
  while (k > 0 && X[k] > x) --k;
  return k;
}

// include/TOMATH.h:92-126 GetValue.  The interval is found from a guess on the
// uniform knot grid and then moved to the reference bisection's answer
// (largest klo <= n-2 with t[klo] <= x), which is exact for increasing knots.
__device__ __forceinline__ void spline_value(const double* __restrict__ X, const double* __restrict__ Y,
                                             const double* __restrict__ YPP, long n, double x, double* out) {
  long klo = spline_interval(X, n, x), khi = klo + 1;
  double h = X[khi] - X[klo];
  double a = (X[khi] - x) / h;
  double b = (x - X[klo]) / h;
  double ca = a * a * a - a;
  double cb = b * b * b - b;
  const double* Y0 = Y + 9 * klo;
  const double* Y1 = Y + 9 * khi;
  const double* P0 = YPP + 9 * klo;
  const double* P1 = YPP + 9 * khi;
#pragma unroll
  for (int c = 0; c < 9; ++c) out[c] = Y0[c] * a + Y1[c] * b + (P0[c] * ca + P1[c] * cb) * (h * h) / 6.;
}

}  // namespace om
