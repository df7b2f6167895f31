// Device helpers shared by the Lienard-Wiechert kernels (lw.hip, batch.hip).
//
// Reference behaviour restated (paths relative to the reference root):
//   integrand and level test  src/OSCARSSR.cc:1975-2012
//   polarization projection   src/OSCARSSR.cc:1999-2001, src/TVector3DC.cc:123-127
#pragma once

#include "om_device.h"
#include "launch_types.h"
#include "om_types.h"

namespace om {

__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double& rr, double& ri) {
  double ac = ar * br, bd = ai * bi, ad = ar * bi, bc = ai * br;
  rr = ac - bd;
  ri = ad + bc;
}

// E <- (E . P) P with the bilinear dot (src/TVector3DC.cc:123-127)
__device__ __forceinline__ void pol_project(double E[6], const double P[6]) {
  double sr = 0, si = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double r, i;
    cmul(E[2 * k], E[2 * k + 1], P[2 * k], P[2 * k + 1], r, i);
    if (k == 0) { sr = r; si = i; } else { sr += r; si += i; }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) cmul(P[2 * k], P[2 * k + 1], sr, si, E[2 * k], E[2 * k + 1]);
}

__device__ __forceinline__ double cmag2(const double E[6]) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double r = E[2 * k] * E[2 * k] - E[2 * k + 1] * (-E[2 * k + 1]);
    s = k == 0 ? r : s + r;
  }
  return s;
}

// Real amplitude vector of one sample, src/OSCARSSR.cc:1975-1990:
//   V = (1-|B|^2)(N-B)/(D^2 (1-N.B)^2) + N x ((N-B) x A)/(D (1-N.B)^2)
// and tau = t + D/c.  tau and N = R/D follow the reference's operations
// exactly (the phase, and 1 - N.B, which cancels to ~1/(2 gamma^2), are
// sensitive to the last bit): D is the IEEE square root, and R/D and D/c are
// the IEEE quotients, formed from the correctly rounded reciprocals 1/D (one
// division) and 1/c (a constant) with Markstein's correction (div_by).  The
// two final scalings of V use a Newton-refined hardware reciprocal of
// D (1-N.B)^2 and the reciprocal of D (V moves by a few 1e-16 relative).
struct Amp {
  double tau;
  V3 V;
};

constexpr double kInvC = 1.0 / kC;  // correctly rounded at compile time

// 1/x to within about one ulp: the hardware estimate and two Newton steps.
__device__ __forceinline__ double recip_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}

__device__ __forceinline__ Amp amplitude(V3 obs, const double* __restrict__ s) {
  V3 X = v3(s[1], s[2], s[3]);
  V3 B = v3(s[4], s[5], s[6]);
  V3 Ac = v3(s[7], s[8], s[9]);
  V3 R = sub(obs, X);
  double D2 = mag2(R);
  double D = sqrt(D2);
  double invD = 1.0 / D;
  V3 N = v3(div_by(R.x, D, invD), div_by(R.y, D, invD), div_by(R.z, D, invD));
  Amp a;
  a.tau = s[0] + div_by(D, kC, kInvC);
  double s1 = 1 - mag2(B);
  V3 NmB = sub(N, B);
  double om1 = 1 - dot(N, B);
  double om2 = om1 * om1;
  V3 c2 = cross(N, cross(NmB, Ac));
  double q2 = recip_nr(D * om2);
  double q1 = s1 * q2 * invD;
  a.V = v3(NmB.x * q1 + c2.x * q2, NmB.y * q1 + c2.y * q2, NmB.z * q1 + c2.z * q2);
  return a;
}

// Wave-scope barrier for LDS data exchanged only inside one wave (the waves of
// a workgroup walk slices of different lengths, so a workgroup barrier cannot
// be used inside the chunk loop).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void sincos_mode(bool mixed, double phi, double& sn, double& cs) {
  if (mixed) {
    sincos_fast64(phi, &sn, &cs);
  } else {
    sincos_huge64(phi, &sn, &cs);
  }
}

// Samples [i0, i1) of one level of one point's sum, the 64 lanes of a wave
// splitting them (lane j takes samples i0+j, i0+j+64, ...; amp_at(i) gives
// sample i's amplitude).  MIXED: reduced-accuracy sin/cos of the exactly reduced phase, and
// `an` returns the slice's sum of |V|_1 (for the error bound); else FP64.  The
// largest phase step between consecutive samples takes the previous sample's
// phase from the neighbouring lane (for sample i0 > 0, from amp_at(i0 - 1)).
// The partial sums are combined by a butterfly, so every lane ends with the
// same totals (pairs add the same two values).  Used for few points with many
// samples: whole waves per point keep their latency low.
template <bool MIXED, class AmpAt>
__device__ __forceinline__ void slice_level_sum(long i0, long i1, double negw, int lane, AmpAt amp_at,
                                                double tot[6], double& maxd, double& an) {
  double ps[6] = {0, 0, 0, 0, 0, 0};
  double md = 0, al = 0;
  double prev_last = i0 > 0 && i1 > i0 ? negw * amp_at(i0 - 1).tau : 0.0;
  for (long base = i0; base < i1; base += 64) {
    long i = base + lane;
    bool ok = i < i1;
    double phi = 0;
    if (ok) {
      Amp a = amp_at(i);
      phi = negw * a.tau;
      double sn, cs;
      sincos_mode(MIXED, phi, sn, cs);
      ps[0] = __builtin_fma(a.V.x, cs, ps[0]); ps[1] = __builtin_fma(a.V.x, sn, ps[1]);
      ps[2] = __builtin_fma(a.V.y, cs, ps[2]); ps[3] = __builtin_fma(a.V.y, sn, ps[3]);
      ps[4] = __builtin_fma(a.V.z, cs, ps[4]); ps[5] = __builtin_fma(a.V.z, sn, ps[5]);
      if (MIXED) al += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);
    }
    double prev = __shfl_up(phi, 1);
    if (lane == 0) prev = (base == 0) ? phi : prev_last;
    if (ok) md = fmax(md, fabs(phi - prev));
    prev_last = __shfl(phi, 63);
  }
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) ps[k] += __shfl_xor(ps[k], off);
    md = fmax(md, __shfl_xor(md, off));
    if (MIXED) al += __shfl_xor(al, off);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) tot[k] = ps[k];
  maxd = md;
  an = al;
}

// slice_level_sum over stored sample rows (10 doubles per sample), with the
// next chunk's rows loaded before the current chunk is computed, so the row
// latency overlaps the arithmetic.  Same operations and order as
// slice_level_sum with amp_at(i) = amplitude(obs, rows + 10 i).
template <bool MIXED>
__device__ __forceinline__ void slice_level_sum_rows(const double* __restrict__ rows, long i0, long i1, V3 obs,
                                                     double negw, int lane, double tot[6], double& maxd,
                                                     double& an) {
  double ps[6] = {0, 0, 0, 0, 0, 0};
  double md = 0, al = 0;
  double prev_last = i0 > 0 && i1 > i0 ? negw * amplitude(obs, rows + 10 * (i0 - 1)).tau : 0.0;
  double cur[10], nxt[10];
  const long valid0 = i0 + lane < i1 ? i0 + lane : i0;  // a row inside the slice (or row i0)
  {
#pragma unroll
    for (int q = 0; q < 10; ++q) nxt[q] = rows[10 * valid0 + q];
  }
  for (long base = i0; base < i1; base += 64) {
#pragma unroll
    for (int q = 0; q < 10; ++q) cur[q] = nxt[q];
    {
      long in = base + 64 + lane;
      in = in < i1 ? in : valid0;  // past the end: reload a valid row (unused)
#pragma unroll
      for (int q = 0; q < 10; ++q) nxt[q] = rows[10 * in + q];
    }
    long i = base + lane;
    bool ok = i < i1;
    double phi = 0;
    if (ok) {
      Amp a = amplitude(obs, cur);
      phi = negw * a.tau;
      double sn, cs;
      sincos_mode(MIXED, phi, sn, cs);
      ps[0] = __builtin_fma(a.V.x, cs, ps[0]); ps[1] = __builtin_fma(a.V.x, sn, ps[1]);
      ps[2] = __builtin_fma(a.V.y, cs, ps[2]); ps[3] = __builtin_fma(a.V.y, sn, ps[3]);
      ps[4] = __builtin_fma(a.V.z, cs, ps[4]); ps[5] = __builtin_fma(a.V.z, sn, ps[5]);
      if (MIXED) al += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);
    }
    double prev = __shfl_up(phi, 1);
    if (lane == 0) prev = (base == 0) ? phi : prev_last;
    if (ok) md = fmax(md, fabs(phi - prev));
    prev_last = __shfl(phi, 63);
  }
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) ps[k] += __shfl_xor(ps[k], off);
    md = fmax(md, __shfl_xor(md, off));
    if (MIXED) al += __shfl_xor(al, off);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) tot[k] = ps[k];
  maxd = md;
  an = al;
}

// Mixed path: bound of the error the reduced-accuracy sin/cos (sincos_fast64)
// puts into a level's stop decision.  Its cos/sin of the exactly reduced phase
// are within OM_SINCOS_FAST_ERR of the exact values (the polynomials' largest
// error on a dense grid, with a margin for FP64 rounding).  Each real
// component of the sum, Re/Im of sum_i V_k,i e^{i phi}, is then off by at most
// OM_SINCOS_FAST_ERR * sum_i |V_k,i|, so the complex vector E = sum * dt is off
// by d <= sqrt(2) OM_SINCOS_FAST_ERR an dt, an = sum over the samples so far of
// |V|_1 (the polarization projection does not increase the norm), and |E|^2 by
// dm <= 2 |E| d + 3 d^2 (|E| from the mixed sum, within d of the exact one).
// prec = |mag - lastmag| / lastmag then moves by at most
// (dm + dm_last (1 + prec)) / (lastmag - dm_last).  A decision whose margin
// |prec - precision| is within 1.25 times that bound is "borderline" (the FP64
// path might decide the other way); such points are recomputed in FP64.
constexpr double kSinCosMixedErr = OM_SINCOS_FAST_ERR;

__device__ __forceinline__ bool level_borderline(int L, double dt, double precision, double mag_prev, double mag,
                                                 double prec, double maxd, double an, double& dm_last) {
  double d = 1.4142135623730951 * kSinCosMixedErr * an * dt;
  double dm = 2.0 * sqrt(mag) * d + 3.0 * d * d;
  double den = mag_prev - dm_last;
  double b = 1.25 * (dm + dm_last * (1.0 + prec)) / den;
  dm_last = dm;
  return L > 8 && maxd < kPi && (den <= 0.0 || fabs(prec - precision) <= b);
}

}  // namespace om
