// Device helpers shared by the Lienard-Wiechert kernels (lw.hip, batch.hip).
//
// Reference behaviour restated (paths relative to the reference root):
//   integrand and level test  src/OSCARSSR.cc:1975-2012
//   polarization projection   src/OSCARSSR.cc:1999-2001, src/TVector3DC.cc:123-127
#pragma once

#include "om_device.h"
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
// and tau = t + D/c, both in the reference's operation order.
struct Amp {
  double tau;
  V3 V;
};

__device__ __forceinline__ Amp amplitude(V3 obs, const double* __restrict__ s) {
  V3 X = v3(s[1], s[2], s[3]);
  V3 B = v3(s[4], s[5], s[6]);
  V3 Ac = v3(s[7], s[8], s[9]);
  V3 R = sub(obs, X);
  double D2 = mag2(R);
  double D = sqrt(D2);
  V3 N = v3(R.x / D, R.y / D, R.z / D);
  Amp a;
  a.tau = s[0] + D / kC;
  double s1 = 1 - mag2(B);
  V3 NmB = sub(N, B);
  double om1 = 1 - dot(N, B);
  double om2 = om1 * om1;
  double den1 = D * D * om2;
  double den2 = D * om2;
  V3 c2 = cross(N, cross(NmB, Ac));
  a.V = v3(NmB.x * s1 / den1 + c2.x / den2, NmB.y * s1 / den1 + c2.y / den2, NmB.z * s1 / den1 + c2.z / den2);
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
    float sf, cf;
    sincos_reduced32(reduce_pio2(phi), &sf, &cf);
    sn = (double) sf;
    cs = (double) cf;
  } else {
    sincos_huge64(phi, &sn, &cs);
  }
}

}  // namespace om
