// Device-side helpers shared by the gfx950 kernels.
//
// All FP64 expressions that feed the retarded-time phase keep the reference's
// association order (src/TVector3D.cc, src/OSCARSSR.cc:1975-1993) and the
// library is compiled with -ffp-contract=off, so that the phase argument is
// bit-identical to the reference CPU path.  FMAs are written explicitly only
// where they cannot change that argument (sin/cos polynomials, exact
// Cody-Waite reduction, accumulations).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sincos_coeffs.h"

namespace om {

// include/TOSCARS.h:20-76 (values kept verbatim; they are physical facts)
constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 2.99792458e8;

// ---------------------------------------------------------------- vectors
struct V3 {
  double x, y, z;
};

__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double mag2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// src/TVector3D.cc Cross: (fY*Vz - Vy*fZ, fZ*Vx - Vz*fX, fX*Vy - Vx*fY)
__device__ __forceinline__ V3 cross(V3 a, V3 v) {
  return V3{a.y * v.z - v.y * a.z, a.z * v.x - v.z * a.x, a.x * v.y - v.x * a.y};
}

// ---------------------------------------------------------------- exact quotient
// q / b for a divisor whose correctly rounded reciprocal inv = RN(1/b) is
// known: y = RN(q*inv) is within one ulp of q/b, the remainder q - y*b is exact
// in one FMA, and one correction gives RN(q/b) (Markstein's theorem), i.e. the
// same double as the IEEE division, in three instructions.
__device__ __forceinline__ double div_by(double q, double b, double inv) {
  double y = q * inv;
  double e = __builtin_fma(-y, b, q);
  return __builtin_fma(e, inv, y);
}

// ---------------------------------------------------------------- sin/cos of huge phases
// The phase phi = -omega (t + D/c) reaches ~5e11 rad.  Reduction by pi/2 with
// a two-part constant: k = nearest integer to phi*2/pi (obtained with the
// 1.5*2^52 rounding constant, whose low mantissa bits then hold k mod 2^32),
// r = phi - k*P1 is exact (phi and k*P1 are both multiples of 2^-52 and
// |r| < 1), then r -= k*P2 rounds once.  |r| <= pi/4 + 2e-5.
constexpr double kTwoOverPi = 0.63661977236758134308;
constexpr double kPio2Hi = 1.5707963267948966192;      // nearest double to pi/2
constexpr double kPio2Lo = 6.123233995736766036e-17;   // pi/2 - kPio2Hi (rounded)
constexpr double kRoundMagic = 6755399441055744.0;     // 1.5 * 2^52
constexpr double kOneOverPi = 0.31830988618379067154;
constexpr double kPiHi = 3.1415926535897931160;      // nearest double to pi
constexpr double kPiLo = 1.2246467991473532072e-16;  // pi - kPiHi (rounded)

struct Reduced {
  double r;
  int q;
};

__device__ __forceinline__ Reduced reduce_pio2(double phi) {
  double kk = __builtin_fma(phi, kTwoOverPi, kRoundMagic);
  int q = (int) (uint32_t) __double_as_longlong(kk);
  double k = kk - kRoundMagic;
  double r = __builtin_fma(-k, kPio2Hi, phi);
  r = __builtin_fma(-k, kPio2Lo, r);
  return Reduced{r, q};
}

// FP64 polynomial sin/cos on the reduced argument, quadrant fix-up by q & 3.
__device__ __forceinline__ void sincos_reduced64(Reduced a, double* s, double* c) {
  double r = a.r;
  double u = r * r;
  double ps = __builtin_fma(u, OM_S64_5, OM_S64_4);
  ps = __builtin_fma(u, ps, OM_S64_3);
  ps = __builtin_fma(u, ps, OM_S64_2);
  ps = __builtin_fma(u, ps, OM_S64_1);
  ps = __builtin_fma(u, ps, OM_S64_0);
  double sr = __builtin_fma(r * u, ps, r);
  double pc = __builtin_fma(u, OM_C64_5, OM_C64_4);
  pc = __builtin_fma(u, pc, OM_C64_3);
  pc = __builtin_fma(u, pc, OM_C64_2);
  pc = __builtin_fma(u, pc, OM_C64_1);
  pc = __builtin_fma(u, pc, OM_C64_0);
  double cr = __builtin_fma(u * u, pc, __builtin_fma(-0.5, u, 1.0));
  int q = a.q;
  double ss = (q & 1) ? cr : sr;
  double cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ void sincos_huge64(double phi, double* s, double* c) {
  sincos_reduced64(reduce_pio2(phi), s, c);
}

// sin(x) for the field models (|x| up to ~1e5 rad), within two ulp.
// The argument is reduced exactly modulo pi (k = nearest integer to x/pi,
// r = x - k*P1 exact for |r| < 2, then r -= k*P2), sin(r) = r + r^3 SH(r^2)
// on |r| <= pi/2 + 2e-5 with SH of degree 7 evaluated by Estrin's scheme, and
// the sign flipped for odd k: one polynomial and no quadrant select, with a
// short dependency chain (the RK4 integration is one serial chain of these).
__device__ __forceinline__ double sin64(double x) {
  double kk = __builtin_fma(x, kOneOverPi, kRoundMagic);
  unsigned q = (unsigned) (uint32_t) __double_as_longlong(kk);
  double k = kk - kRoundMagic;
  double r = __builtin_fma(-k, kPiHi, x);
  r = __builtin_fma(-k, kPiLo, r);
  double u = r * r;
  double u2 = u * u;
  double u4 = u2 * u2;
  double p01 = __builtin_fma(u, OM_SH_1, OM_SH_0);
  double p23 = __builtin_fma(u, OM_SH_3, OM_SH_2);
  double p45 = __builtin_fma(u, OM_SH_5, OM_SH_4);
  double p67 = __builtin_fma(u, OM_SH_7, OM_SH_6);
  double q0 = __builtin_fma(u2, p23, p01);
  double q1 = __builtin_fma(u2, p67, p45);
  double sh = __builtin_fma(u4, q1, q0);
  double sr = __builtin_fma(r * u, sh, r);
  unsigned long long sg = (unsigned long long) (q << 31) << 32;  // bit 0 of k -> sign bit
  return __longlong_as_double(__double_as_longlong(sr) ^ (long long) sg);
}

// cos(x) for the time dependence of the field models: the reduction by pi/2
// and polynomial pair of sincos_huge64, within about one ulp.
__device__ __forceinline__ double cos64(double x) {
  double s, c;
  sincos_reduced64(reduce_pio2(x), &s, &c);
  return c;
}

// Reduced-accuracy sin/cos of a huge phase, in FP64 (mixed path).  The phase
// is reduced modulo pi with the same exact two-part scheme (k = nearest
// integer to phi/pi, r = phi - k*P1 exact for |r| < 2, then r -= k*P2),
// |r| <= pi/2 + 2e-5; sin(r) = r + r^3 SF(r^2) and cos(r) = 1 + r^2 QF(r^2)
// with the low-degree polynomials of sincos_coeffs.h, whose absolute error,
// FP64 rounding included, is at most OM_SINCOS_FAST_ERR; sin(phi) and cos(phi)
// are those of r with the sign flipped for odd k (one shift and two xors
// instead of a quadrant select).
__device__ __forceinline__ void sincos_fast64(double phi, double* s, double* c) {
  double kk = __builtin_fma(phi, kOneOverPi, kRoundMagic);
  unsigned q = (unsigned) (uint32_t) __double_as_longlong(kk);
  double k = kk - kRoundMagic;
  double r = __builtin_fma(-k, kPiHi, phi);
  r = __builtin_fma(-k, kPiLo, r);
  double u = r * r;
  double ps = __builtin_fma(u, OM_SF_3, OM_SF_2);
  ps = __builtin_fma(u, ps, OM_SF_1);
  ps = __builtin_fma(u, ps, OM_SF_0);
  double sr = __builtin_fma(r * u, ps, r);
  double pq = __builtin_fma(u, OM_QF_4, OM_QF_3);
  pq = __builtin_fma(u, pq, OM_QF_2);
  pq = __builtin_fma(u, pq, OM_QF_1);
  pq = __builtin_fma(u, pq, OM_QF_0);
  double cr = __builtin_fma(u, pq, 1.0);
  unsigned long long sg = (unsigned long long) (q << 31) << 32;  // bit 0 of k -> sign bit
  *s = __longlong_as_double(__double_as_longlong(sr) ^ (long long) sg);
  *c = __longlong_as_double(__double_as_longlong(cr) ^ (long long) sg);
}

}  // namespace om
