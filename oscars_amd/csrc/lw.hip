// Lienard-Wiechert frequency-domain sums for gfx950.
//
// Reference behaviour restated (paths relative to the reference root):
//   spectrum integrand     src/OSCARSSR.cc:1939-2038
//   flux-map integrand     src/OSCARSSR.cc:4167-4266
//   power-density integrand src/OSCARSSR.cc:2843-2966
//
// Output point = (observation point, photon energy) for spectrum / flux, or an
// observation point for power density.  Each point sums the trajectory level
// samples 0..L and stops at the first L > 8 meeting the reference's
// convergence test; the per-point state (complex E sum, last |E|^2, level,
// converged flag) lives in HBM so a launch can process a range of levels and
// the host can continue with more levels for the points that need them.
//
// Kernels:
//   lw_point_kernel   one lane = one output point, everything per sample in FP64
//                     (the plain FP64 path; the trajectory row is wave-uniform and
//                     is read through the scalar cache)
//   lw_tile_kernel    one wave = one observation point x 64 photon energies; the
//                     per-(obs, sample) retarded time tau = t + D/c and real
//                     amplitude vector V are computed once per sample by one lane
//                     and broadcast through LDS to the 64 energy lanes, which each
//                     evaluate phi = -omega*tau, sin/cos and the complex MAC.
//                     FP64 variant: bit-identical phase, FP64 sin/cos polynomial;
//                     mixed variant: the same FP64 phase, reduced exactly
//                     modulo pi, then low-degree FP64 sin/cos polynomials
//                     (error <= 4.8e-9).  Amplitudes and the complex
//                     accumulation stay FP64 in both: with packed FP32
//                     disabled, v_fma_f32 and v_fma_f64 issue at the same rate
//                     on gfx950, so FP32 would cost accuracy and save nothing.
//   pd_point_kernel   power density, one lane = one observation point, a level's
//                     samples split over the S waves of a workgroup.  FP64
//                     variant: the reference's operations; mixed variant:
//                     reciprocal square root and reciprocal refined by Newton
//                     steps instead of the square root and quotients, with an
//                     error bound per point that sends borderline stop
//                     decisions to the FP64 variant
//   lw_finalize_kernel  C0*dt scaling, polarization projection, flux / precision /
//                     level output, optional weighted Kahan accumulation over particles
#include <mutex>

#include "lw_common.h"

namespace om {




// Per-level convergence test, src/OSCARSSR.cc:1997-2012.
__device__ __forceinline__ bool level_converged(const LwArgs& A, int L, const double sum[6], double maxd,
                                                double& lastmag, double& prec) {
  double E[6];
  double dt = A.lv.dt_incl[L];
#pragma unroll
  for (int k = 0; k < 6; ++k) E[k] = sum[k] * dt;
  if (A.use_pol) pol_project(E, A.pol);
  double mag = cmag2(E);
  prec = fabs(mag - lastmag) / lastmag;
  bool c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == lastmag);
  lastmag = mag;
  return c;
}

// ---------------------------------------------------------------- level-slice combination
// A workgroup of S waves serves 64 output points; wave w sums the contiguous
// slice [ns*w/S, ns*(w+1)/S) of each level.  Each wave leaves its partial
// complex sum, its largest phase step inside the slice and the phases of its
// first and last sample in LDS; every wave then forms the same totals, in wave
// order, so all waves agree on which points have converged.
template <int S>
struct SliceLds {
  double ps[S][6][64];
  double maxd[S][64];
  double first[S][64];
  double last[S][64];
  double an[S][64];
};

template <int S>
__device__ __forceinline__ void slice_publish(SliceLds<S>& sh, int w, int lane, const double ps[6], double maxd,
                                              double first, double last) {
#pragma unroll
  for (int k = 0; k < 6; ++k) sh.ps[w][k][lane] = ps[k];
  sh.maxd[w][lane] = maxd;
  sh.first[w][lane] = first;
  sh.last[w][lane] = last;
}

template <int S>
__device__ __forceinline__ void slice_combine(SliceLds<S>& sh, int lane, long ns, double tot[6], double& maxd) {
  maxd = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) tot[k] = 0;
  bool have_prev = false;
  double prev_last = 0;
  for (int w = 0; w < S; ++w) {
    long i0 = ns * w / S, i1 = ns * (w + 1) / S;
    if (i1 <= i0) continue;
#pragma unroll
    for (int k = 0; k < 6; ++k) tot[k] += sh.ps[w][k][lane];
    maxd = fmax(maxd, sh.maxd[w][lane]);
    if (have_prev) maxd = fmax(maxd, fabs(sh.first[w][lane] - prev_last));
    prev_last = sh.last[w][lane];
    have_prev = true;
  }
}

// ---------------------------------------------------------------- one lane per point
// MIXED = false is the plain FP64 path; MIXED = true evaluates sin/cos with the
// reduced-accuracy polynomials on the exactly reduced FP64 phase.  Every lane reads the same
// trajectory row (scalar loads) and computes its own amplitude and phase.
template <bool MIXED, int S>
__global__ void __launch_bounds__(64 * S) lw_point_kernel(LwArgs A) {
  __shared__ SliceLds<S> sh;
  int lane = threadIdx.x & 63;
  int w = threadIdx.x >> 6;
  long g = (long) blockIdx.x * 64 + lane;
  bool valid;
  long p;
  if (A.list) {  // compacted: the points still active at this level
    valid = g < A.n_list;
    p = valid ? (long) A.list[g] : 0;
  } else {
    valid = g < A.npts;
    p = g;
  }
  long pc = valid ? p : 0;
  long io = pc / A.n_omega;
  long iw = pc - io * A.n_omega;
  V3 obs = v3(A.obs[3 * io], A.obs[3 * io + 1], A.obs[3 * io + 2]);
  double negw = -A.omega[iw];
  int lev = A.st.level[pc];
  bool active = valid && !A.st.conv[pc] && lev == A.La - 1;
  double sum[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sum[k] = A.st.sum[6 * pc + k];
  double lastmag = A.st.lastmag[pc];
  double prec = A.st.prec[pc];
  double an_run = 0, dm_last = 0;
  bool border = false;
  if (MIXED) {
    an_run = A.st.err[2 * pc];
    dm_last = A.st.err[2 * pc + 1];
  }
  for (int L = A.La; L <= A.Lb; ++L) {
    if (!__any(active)) break;
    const double* __restrict__ rows = A.lv.rows[L];
    long ns = 1L << L;
    long i0 = ns * w / S, i1 = ns * (w + 1) / S;
    double ps[6] = {0, 0, 0, 0, 0, 0};
    double maxd = 0, first = 0, lastphase = 0, an = 0;
    if (i1 > i0) {
      Amp a0 = amplitude(obs, rows + 10 * i0);
      first = negw * a0.tau;
      lastphase = first;
    }
    for (long i = i0; i < i1; ++i) {
      Amp a = amplitude(obs, rows + 10 * i);
      double phi = negw * a.tau;
      maxd = fmax(maxd, fabs(phi - lastphase));
      lastphase = phi;
      double sn, cs;
      sincos_mode(MIXED, phi, sn, cs);
      ps[0] = __builtin_fma(a.V.x, cs, ps[0]); ps[1] = __builtin_fma(a.V.x, sn, ps[1]);
      ps[2] = __builtin_fma(a.V.y, cs, ps[2]); ps[3] = __builtin_fma(a.V.y, sn, ps[3]);
      ps[4] = __builtin_fma(a.V.z, cs, ps[4]); ps[5] = __builtin_fma(a.V.z, sn, ps[5]);
      if (MIXED) an += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);
    }
    double tot[6];
    if (S > 1) {
      slice_publish<S>(sh, w, lane, ps, maxd, first, lastphase);
      if (MIXED) sh.an[w][lane] = an;
      __syncthreads();
      slice_combine<S>(sh, lane, ns, tot, maxd);
      if (MIXED) {
        an = 0;
        for (int v = 0; v < S; ++v) an += sh.an[v][lane];
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) tot[k] = ps[k];
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sum[k] += tot[k];
      lev = L;
      double mag_prev = lastmag;
      bool c = level_converged(A, L, sum, maxd, lastmag, prec);
      if (MIXED) {
        an_run += an;
        border = border ||
                 level_borderline(L, A.lv.dt_incl[L], A.precision, mag_prev, lastmag, prec, maxd, an_run, dm_last);
      }
      if (c) {
        A.st.conv[pc] = 1;
        active = false;
      }
    }
  }
  if (w == 0 && valid && lev >= A.La) {
#pragma unroll
    for (int k = 0; k < 6; ++k) A.st.sum[6 * pc + k] = sum[k];
    A.st.lastmag[pc] = lastmag;
    A.st.prec[pc] = prec;
    A.st.level[pc] = lev;
    if (MIXED) {
      A.st.err[2 * pc] = an_run;
      A.st.err[2 * pc + 1] = dm_last;
      if (border) A.st.border[pc] = 1;
    }
  }
}

// ---------------------------------------------------------------- tiled obs x 64 energies
// One workgroup of S waves per (observation point, 64-energy tile).  Within a
// wave, lane j computes tau and V of sample base+j of the wave's slice and
// leaves them in LDS; then all 64 lanes (energies) walk those 64 samples.
template <bool MIXED, int S>
__global__ void __launch_bounds__(64 * S) lw_tile_kernel(LwArgs A, long n_tiles_w) {
  __shared__ SliceLds<S> sh;
  __shared__ double s_tau[S][64];
  __shared__ double s_v[S][3][64];
  int lane = threadIdx.x & 63;
  int w = threadIdx.x >> 6;
  long tile = blockIdx.x;
  long io, iw;
  bool valid;
  if (A.list) {
    // compacted tile: active points of one observation point (the first entry is never a pad)
    int pt = A.list[64 * tile + lane];
    io = A.list[64 * tile] / A.n_omega;
    valid = pt >= 0;
    iw = valid ? pt - io * A.n_omega : 0;
  } else {
    io = tile / n_tiles_w;
    iw = (tile - io * n_tiles_w) * 64 + lane;
    valid = iw < A.n_omega;
  }
  long iwc = valid ? iw : A.n_omega - 1;
  long p = io * A.n_omega + iwc;
  V3 obs = v3(A.obs[3 * io], A.obs[3 * io + 1], A.obs[3 * io + 2]);
  double negw = -A.omega[iwc];
  int lev = A.st.level[p];
  bool active = valid && !A.st.conv[p] && lev == A.La - 1;
  double sum[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sum[k] = A.st.sum[6 * p + k];
  double lastmag = A.st.lastmag[p];
  double prec = A.st.prec[p];
  double an_run = 0, dm_last = 0;
  bool border = false;
  if (MIXED) {
    an_run = A.st.err[2 * p];
    dm_last = A.st.err[2 * p + 1];
  }
  double* tau_w = s_tau[w];
  double* vx_w = s_v[w][0];
  double* vy_w = s_v[w][1];
  double* vz_w = s_v[w][2];
  for (int L = A.La; L <= A.Lb; ++L) {
    if (!__any(active)) break;
    const double* __restrict__ rows = A.lv.rows[L];
    long ns = 1L << L;
    long i0 = ns * w / S, i1 = ns * (w + 1) / S;
    double ps[6] = {0, 0, 0, 0, 0, 0};
    double maxd = 0, first = 0, lastphase = 0, an = 0;
    for (long base = i0; base < i1; base += 64) {
      long si = base + lane;
      if (si < i1) {
        Amp a = amplitude(obs, rows + 10 * si);
        tau_w[lane] = a.tau;
        vx_w[lane] = a.V.x;
        vy_w[lane] = a.V.y;
        vz_w[lane] = a.V.z;
        if (MIXED) an += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);  // this lane's samples
      }
      wave_sync();  // the wave's own LDS writes are visible to all its lanes
      int cnt = (int) ((i1 - base) < 64 ? (i1 - base) : 64);
      if (base == i0) {
        first = negw * tau_w[0];
        lastphase = first;
      }
      int k = 0;
      for (; k + 4 <= cnt; k += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          double phi = negw * tau_w[k + u];
          maxd = fmax(maxd, fabs(phi - lastphase));
          lastphase = phi;
          double sn, cs;
          sincos_mode(MIXED, phi, sn, cs);
          double vx = vx_w[k + u], vy = vy_w[k + u], vz = vz_w[k + u];
          ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
          ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
          ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
        }
      }
      for (; k < cnt; ++k) {
        double phi = negw * tau_w[k];
        maxd = fmax(maxd, fabs(phi - lastphase));
        lastphase = phi;
        double sn, cs;
        sincos_mode(MIXED, phi, sn, cs);
        double vx = vx_w[k], vy = vy_w[k], vz = vz_w[k];
        ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
        ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
        ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
      }
      wave_sync();  // all lanes done reading before the next chunk is written
    }
    if (MIXED) {  // the samples' |V|_1 summed over the wave (the same for every energy lane)
      for (int off = 32; off > 0; off >>= 1) an += __shfl_xor(an, off);
    }
    double tot[6];
    if (S > 1) {
      slice_publish<S>(sh, w, lane, ps, maxd, first, lastphase);
      if (MIXED) sh.an[w][lane] = an;
      __syncthreads();
      slice_combine<S>(sh, lane, ns, tot, maxd);
      if (MIXED) {
        an = 0;
        for (int v = 0; v < S; ++v) an += sh.an[v][lane];
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) tot[k] = ps[k];
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sum[k] += tot[k];
      lev = L;
      double mag_prev = lastmag;
      bool c = level_converged(A, L, sum, maxd, lastmag, prec);
      if (MIXED) {
        an_run += an;
        border = border ||
                 level_borderline(L, A.lv.dt_incl[L], A.precision, mag_prev, lastmag, prec, maxd, an_run, dm_last);
      }
      if (c) {
        A.st.conv[p] = 1;
        active = false;
      }
    }
  }
  if (w == 0 && valid && lev >= A.La) {
#pragma unroll
    for (int k = 0; k < 6; ++k) A.st.sum[6 * p + k] = sum[k];
    A.st.lastmag[p] = lastmag;
    A.st.prec[p] = prec;
    A.st.level[p] = lev;
    if (MIXED) {
      A.st.err[2 * p] = an_run;
      A.st.err[2 * p + 1] = dm_last;
      if (border) A.st.border[p] = 1;
    }
  }
}

// ---------------------------------------------------------------- power density
// One workgroup of S waves per 64 observation points (lane = point): wave w
// sums the contiguous slice [ns*w/S, ns*(w+1)/S) of each level, the rows being
// wave-uniform (scalar loads); the slices' partial sums are added in wave order
// through LDS, so all waves hold the same totals and stop decisions.
//
// MIXED = false: the reference's operations for each term (IEEE square root
// and quotients).  MIXED = true: N = R / |R| from a Newton-refined reciprocal
// square root, one Newton-refined reciprocal of (1-N.B)^5 |R|^2 in place of
// the square root and the five quotients, and
// Num = N x ((N-B) x Ac) = (N-B)(N.Ac) - Ac (1-N.B)  (|N| = 1).
// With x = 1 - N.B, T = |Num|^2 / (x^5 |R|^2) >= |term| and
// Q = x |Ac|^2 / (x^5 |R|^2), a mixed term lies within
// (Kx / x + K0) T + KQ Q of the FP64 term (u = 2^-53):
//   - N differs by at most 7.5 u between the paths, x by 14 u, so x^5 by
//     70 u / x relative;
//   - Num, a function of N with slope at most 6 |Ac|, moves by 45 u |Ac|, and
//     the two paths' roundings and x add 46 u |Ac|, so
//     |d|Num|^2| <= 2 |Num| 91 u |Ac| <= 91 u (|Num|^2 / x + x |Ac|^2);
//   - the powers, reciprocal, quotients and N.normal (a directional surface
//     may skip a term the other path keeps when N.normal is within 14 u of 0)
//     add 41 u of T.
// Kx = 260 u, K0 = 64 u, KQ = 150 u keep a margin of 1.6 over those sums
// (tests/test_pd_mixed_bound.py checks the bound on the CPU, for electrons of
// 3 GeV, down to x = 1/(2 gamma^2)).  A level adds (Kx / min x + K0) sum T +
// KQ sum Q to the running bound e, the sums taken over the terms kept and over
// the skipped ones with N.normal > -16 u; both paths' summation roundings add at
// most 2 n u sum T after n samples.
constexpr double kPdU = 1.1102230246251565e-16;  // 2^-53
constexpr double kPdKx = 260 * kPdU;
constexpr double kPdK0 = 64 * kPdU;
constexpr double kPdKQ = 150 * kPdU;

// 1/sqrt(x) to within about two ulp: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  double e = __builtin_fma(-(h * y), y, 0.5);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-(h * y), y, 0.5);
  return __builtin_fma(y, e, y);
}

__device__ __forceinline__ double mag2_fma(V3 a) {
  return __builtin_fma(a.z, a.z, __builtin_fma(a.y, a.y, a.x * a.x));
}

// One term of the power-density sum, src/OSCARSSR.cc:2886-2921.  Adds it to
// ps unless the surface is directional and N.normal <= 0; the mixed path adds
// T and Q to tsum and qsum and keeps the smallest x.
template <bool MIXED>
__device__ __forceinline__ void pd_term(const PdArgs& A, V3 obs, V3 nrm, const double* __restrict__ s, double& ps,
                                        double& tsum, double& qsum, double& xmin) {
  V3 X = v3(s[1], s[2], s[3]);
  V3 B = v3(s[4], s[5], s[6]);
  V3 Ac = v3(s[7], s[8], s[9]);
  V3 R = sub(obs, X);
  if (!MIXED) {
    double r2 = mag2(R);
    double m = sqrt(r2);
    V3 N1 = v3(R.x / m, R.y / m, R.z / m);
    // src/OSCARSSR.cc:2908: a surface without normals takes N1.normal as 1
    double ndn = A.has_normal ? dot(N1, nrm) : 1.0;
    V3 Num = cross(N1, cross(sub(N1, B), Ac));
    double x = 1 - dot(B, N1);
    double x2 = x * x;
    double den = x2 * x2 * x;
    // (Num.N2)^2 + (Num.N3)^2 = |Num|^2 because Num is perpendicular to N1
    double term = mag2(Num) / den / r2 * ndn;
    if (!(A.directional && ndn <= 0)) ps += term;
  } else {
    double r2 = mag2_fma(R);
    double im = rsqrt_nr(r2);
    V3 N1 = v3(R.x * im, R.y * im, R.z * im);
    double ndn = A.has_normal ? __builtin_fma(N1.z, nrm.z, __builtin_fma(N1.y, nrm.y, N1.x * nrm.x)) : 1.0;
    V3 NmB = sub(N1, B);
    double x = __builtin_fma(-B.x, N1.x, __builtin_fma(-B.y, N1.y, __builtin_fma(-B.z, N1.z, 1.0)));
    double nac = __builtin_fma(N1.z, Ac.z, __builtin_fma(N1.y, Ac.y, N1.x * Ac.x));
    V3 Num = v3(__builtin_fma(NmB.x, nac, -(Ac.x * x)), __builtin_fma(NmB.y, nac, -(Ac.y * x)),
                __builtin_fma(NmB.z, nac, -(Ac.z * x)));
    double x2 = x * x;
    double inv = recip_nr((x2 * x2) * (x * r2));
    double t = mag2_fma(Num) * inv;
    bool keep = !(A.directional && ndn <= 0);
    if (keep) ps = __builtin_fma(t, ndn, ps);
    // a skipped term counts in the bound only where the FP64 path might keep it
    bool counted = keep || ndn > -16 * kPdU;
    tsum += counted ? t : 0.0;
    qsum = __builtin_fma(counted ? x * mag2_fma(Ac) : 0.0, inv, qsum);
    xmin = fmin(xmin, x);
  }
}

// Per-point state of the power-density level loop.
struct PdPoint {
  double sum, lastsum, prec, e_run, b_run;
  int lev;
  bool border;
};

template <bool MIXED>
__device__ __forceinline__ PdPoint pd_load(const PdArgs& A, long pc) {
  PdPoint q;
  q.sum = A.st.sum[6 * pc];
  q.lastsum = A.st.lastmag[pc];
  q.prec = A.st.prec[pc];
  q.lev = A.st.level[pc];
  q.e_run = MIXED ? A.st.err[2 * pc] : 0.0;
  q.b_run = MIXED ? A.st.err[2 * pc + 1] : 0.0;
  q.border = false;
  return q;
}

template <bool MIXED>
__device__ __forceinline__ void pd_store(const PdArgs& A, long pc, const PdPoint& q) {
  A.st.sum[6 * pc] = q.sum;
  A.st.lastmag[pc] = q.lastsum;
  A.st.prec[pc] = q.prec;
  A.st.level[pc] = q.lev;
  if (MIXED) {
    A.st.err[2 * pc] = q.e_run;
    A.st.err[2 * pc + 1] = q.b_run;
    if (q.border) A.st.border[pc] = 1;
  }
}

// End of level L for an active point: adds the level's partial sum ps (and, in
// the mixed path, its bound terms), then the reference's stop test,
// src/OSCARSSR.cc:2940-2953.  Returns true when the point stops (converged, or
// a mixed decision within the error bound, to be recomputed in FP64).
template <bool MIXED>
__device__ __forceinline__ bool pd_level_end(const PdArgs& A, int L, double ps, double tsum, double qsum, double xmin,
                                             PdPoint& q) {
  long ns = 1L << L;
  q.sum += ps;
  q.lev = L;
  double dt = A.lv.dt_incl[L];
  double thisSum = q.sum * dt;
  double prev = q.lastsum;
  q.prec = fabs(thisSum - q.lastsum) / q.lastsum;
  double bd = A.lv.beta_diff[L];
  bool c = (L > 8 && q.prec < A.precision && bd < A.two_over_gamma) ||
           (L > 8 && bd < A.two_over_gamma && thisSum == q.lastsum);
  bool c2 = !c && L > 8 && thisSum == q.lastsum;
  if (c2) q.prec = 0;
  q.lastsum = thisSum;
  if (MIXED) {
    // error bounds of this level's and the last level's sums (the last from
    // the running bounds as they stood before this level)
    double dlast = L > 0 ? (q.e_run + 2.0 * (double) ns * kPdU * q.b_run) * A.lv.dt_incl[L - 1] +
                               4 * kPdU * fabs(prev)
                         : 0.0;
    q.e_run += (kPdKx / xmin + kPdK0) * tsum + kPdKQ * qsum;
    q.b_run += tsum;
    double dthis = (q.e_run + 4.0 * (double) ns * kPdU * q.b_run) * dt + 4 * kPdU * fabs(thisSum);
    // no error at all (every term so far skipped far from N.normal = 0):
    // the sums are the FP64 path's, and so is the decision
    if (L > 8 && dthis + dlast > 0.0) {
      double tol = 1.25 * (dthis + dlast);
      double den = prev - dlast;
      bool amb_equal = fabs(thisSum - prev) <= tol;
      bool amb_prec = bd < A.two_over_gamma &&
                      (den <= 0.0 || fabs(q.prec - A.precision) <= 1.25 * (dthis + dlast * (1.0 + q.prec)) / den);
      q.border = q.border || amb_equal || amb_prec;
    }
  }
  return c || c2 || q.border;
}

__device__ __forceinline__ V3 pd_normal(const PdArgs& A, long pc) {
  return A.normals ? v3(A.normals[3 * pc], A.normals[3 * pc + 1], A.normals[3 * pc + 2])
                   : v3(A.normal[0], A.normal[1], A.normal[2]);
}

template <int S>
struct PdLds {
  double ps[S][64];
  double ts[S][64];
  double qs[S][64];
  double xm[S][64];
};

template <bool MIXED, int S>
__global__ void __launch_bounds__(64 * S) pd_point_kernel(PdArgs A) {
  __shared__ PdLds<S> sh;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  long g = (long) blockIdx.x * 64 + lane;
  bool valid;
  long p;
  if (A.list) {  // compacted: the points still active at this level
    valid = g < A.n_list;
    p = valid ? (long) A.list[g] : 0;
  } else {
    valid = g < A.npts;
    p = g;
  }
  long pc = valid ? p : 0;
  V3 obs = v3(A.obs[3 * pc], A.obs[3 * pc + 1], A.obs[3 * pc + 2]);
  V3 nrm = pd_normal(A, pc);
  PdPoint q = pd_load<MIXED>(A, pc);
  bool active = valid && !A.st.conv[pc] && q.lev == A.La - 1;
  for (int L = A.La; L <= A.Lb; ++L) {
    if (S > 1 ? !__syncthreads_or(active) : !__any(active)) break;
    const double* __restrict__ rows = A.lv.rows[L];
    long ns = 1L << L;
    long i0 = ns * w / S, i1 = ns * (w + 1) / S;
    double ps = 0, tsum = 0, qsum = 0, xmin = 2.0;
    for (long i = i0; i < i1; ++i) pd_term<MIXED>(A, obs, nrm, rows + 10 * i, ps, tsum, qsum, xmin);
    if (S > 1) {
      sh.ps[w][lane] = ps;
      if (MIXED) {
        sh.ts[w][lane] = tsum;
        sh.qs[w][lane] = qsum;
        sh.xm[w][lane] = xmin;
      }
      __syncthreads();
      ps = 0;
      tsum = 0;
      qsum = 0;
      xmin = 2.0;
      for (int v = 0; v < S; ++v) {
        long j0 = ns * v / S, j1 = ns * (v + 1) / S;
        if (j1 <= j0) continue;
        ps += sh.ps[v][lane];
        if (MIXED) {
          tsum += sh.ts[v][lane];
          qsum += sh.qs[v][lane];
          xmin = fmin(xmin, sh.xm[v][lane]);
        }
      }
      // the next level's __syncthreads_or keeps these reads ahead of its writes
    }
    if (active && pd_level_end<MIXED>(A, L, ps, tsum, qsum, xmin, q)) {
      A.st.conv[pc] = 1;
      active = false;
    }
  }
  if (w == 0 && valid && q.lev >= A.La) pd_store<MIXED>(A, pc, q);
}

// One wave per listed point (few points left, or the mixed path's FP64
// recheck): lane j sums samples j, j + 64, ... of each level, and the lanes'
// partial sums are combined by a butterfly in a fixed order, so every lane
// holds the same total and takes the same decision.
template <bool MIXED>
__global__ void __launch_bounds__(64) pd_wave_kernel(PdArgs A) {
  const int lane = threadIdx.x;
  long pc = A.list[blockIdx.x];
  V3 obs = v3(A.obs[3 * pc], A.obs[3 * pc + 1], A.obs[3 * pc + 2]);
  V3 nrm = pd_normal(A, pc);
  PdPoint q = pd_load<MIXED>(A, pc);
  if (A.st.conv[pc] || q.lev != A.La - 1) return;
  for (int L = A.La; L <= A.Lb; ++L) {
    const double* __restrict__ rows = A.lv.rows[L];
    long ns = 1L << L;
    double ps = 0, tsum = 0, qsum = 0, xmin = 2.0;
    for (long i = lane; i < ns; i += 64) pd_term<MIXED>(A, obs, nrm, rows + 10 * i, ps, tsum, qsum, xmin);
    for (int off = 32; off > 0; off >>= 1) {
      ps += __shfl_xor(ps, off);
      if (MIXED) {
        tsum += __shfl_xor(tsum, off);
        qsum += __shfl_xor(qsum, off);
        xmin = fmin(xmin, __shfl_xor(xmin, off));
      }
    }
    if (pd_level_end<MIXED>(A, L, ps, tsum, qsum, xmin, q)) {
      if (lane == 0) A.st.conv[pc] = 1;
      break;
    }
  }
  if (lane == 0 && q.lev >= A.La) pd_store<MIXED>(A, pc, q);
}

// ---------------------------------------------------------------- power density on STL triangles
// src/OSCARSSR.cc:3318-3527.  One wave per triangle centre; its lanes take 64
// consecutive samples of a level at a time.  Each lane asks whether any other
// triangle lies between its sample and the centre (src/TTriangle3D.cc:159-215,
// Moller and Trumbore, same operations), walking a bounding-volume hierarchy
// whose boxes are inflated by more than the test's rounding can move an
// accepted intersection, so it finds a hit exactly when the test of every
// triangle would.  Unblocked samples add the power density integrand with
// |N1 . normal| and move the beta-step maximum, the step taken from the last
// unblocked sample before (the lane's ballot gives it), as in the reference;
// the 64 terms are added by a butterfly in a fixed order.
constexpr int kStlWaves = 4;

__device__ __forceinline__ bool stl_blocks(const double* __restrict__ T, V3 O, V3 D, double dist) {
  // T: A, edge1, edge2 of one triangle; D: the unit ray direction, normalised again
  double m = sqrt(mag2(D));
  V3 R = v3(D.x / m, D.y / m, D.z / m);
  V3 A = v3(T[0], T[1], T[2]), e1 = v3(T[3], T[4], T[5]), e2 = v3(T[6], T[7], T[8]);
  V3 p = cross(R, e2);
  double det = dot(e1, p);
  if (det > -1e-12 && det < 1e-12) return false;
  double inv = 1.0 / det;
  V3 tv = sub(O, A);
  double u = dot(tv, p) * inv;
  if (u < 0.0 || u > 1.0) return false;
  V3 q = cross(tv, e1);
  double v = dot(R, q) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  double t = dot(e2, q) * inv;
  if (t < 0) return false;
  return t > 0 && t < dist;
}

// Does the segment X + t D, t in [t0, t1], meet the box?  An axis the
// direction does not move along is a plain containment test.
__device__ __forceinline__ bool stl_box_hit(const double* __restrict__ b, V3 X, V3 D, V3 inv, double t0, double t1) {
  double lo[3] = {b[0], b[1], b[2]}, hi[3] = {b[3], b[4], b[5]};
  double x[3] = {X.x, X.y, X.z}, d[3] = {D.x, D.y, D.z}, iv[3] = {inv.x, inv.y, inv.z};
  bool out = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double ta = (lo[a] - x[a]) * iv[a], tb = (hi[a] - x[a]) * iv[a];
    bool flat = d[a] == 0.0;
    out = out || (flat && (x[a] < lo[a] || x[a] > hi[a]));
    t0 = flat ? t0 : fmax(t0, fmin(ta, tb));
    t1 = flat ? t1 : fmin(t1, fmax(ta, tb));
  }
  return !out && t0 <= t1;
}

// Any triangle other than `self` between X and the point at distance dist
// along the unit direction N1.
__device__ __forceinline__ bool stl_any_hit(const StlArgs& A, V3 X, V3 N1, double dist, long self) {
  V3 inv = v3(1.0 / N1.x, 1.0 / N1.y, 1.0 / N1.z);
  int i = 0;
  while (i < A.bvh_nodes) {
    const int* m = A.bvh_meta + 3 * i;
    int next = m[0];
    if (!stl_box_hit(A.bvh_box + 6 * i, X, N1, inv, -A.pad_t, dist + A.pad_t)) {
      i = next;
      continue;
    }
    int cnt = m[2];
    if (cnt == 0) {  // inner node: its first child follows it
      ++i;
      continue;
    }
    for (int k = 0; k < cnt; ++k) {
      int j = A.bvh_order[m[1] + k];
      if (j != self && stl_blocks(A.tri + 12 * j, X, N1, dist)) return true;
    }
    i = next;
  }
  return false;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

__global__ void __launch_bounds__(64 * kStlWaves) pd_stl_kernel(StlArgs A) {
  const int lane = threadIdx.x & 63;
  const long p = (long) blockIdx.x * kStlWaves + (threadIdx.x >> 6);
  if (p >= A.ntri) return;
  int lev = A.st.level[p];
  if (A.st.conv[p] || lev != A.La - 1) return;
  const V3 obs = v3(A.centre[3 * p], A.centre[3 * p + 1], A.centre[3 * p + 2]);
  const V3 nrm = v3(A.tri[12 * p + 9], A.tri[12 * p + 10], A.tri[12 * p + 11]);
  double sum = A.st.sum[6 * p];
  double lastsum = A.st.lastmag[p];
  double prec = A.st.prec[p];
  bool done = false;
  for (int L = A.La; L <= A.Lb && !done; ++L) {
    const double* __restrict__ rows = A.lv.rows[L];
    const long ns = 1L << L;
    V3 lastB = v3(0, 0, 0);  // the last unblocked sample's beta (none yet: zero)
    double bdmax = -1;
    for (long base = 0; base < ns; base += 64) {
      const long i = base + lane;
      const bool valid = i < ns;
      const double* s = rows + 10 * (valid ? i : 0);
      V3 X = v3(s[1], s[2], s[3]);
      V3 R = sub(obs, X);
      double r2 = mag2(R);
      double dist = sqrt(r2);
      V3 N1 = v3(R.x / dist, R.y / dist, R.z / dist);
      const bool blocked = !valid || stl_any_hit(A, X, N1, dist, p);
      const unsigned long long open = __ballot(!blocked);
      V3 B = v3(s[4], s[5], s[6]);
      V3 Ac = v3(s[7], s[8], s[9]);
      // the last unblocked sample before this one: in this group, else carried
      const unsigned long long below = lane ? open & ((1ull << lane) - 1) : 0ull;
      const int src = below ? 63 - __clzll(below) : lane;
      V3 pB = v3(__shfl(B.x, src), __shfl(B.y, src), __shfl(B.z, src));
      V3 prevB = below ? pB : lastB;
      double term = 0, bd = -1;
      if (!blocked) {
        V3 dB = sub(B, prevB);
        bd = i > 0 ? sqrt(mag2(dB)) : -1.0;
        double ndn = fabs(dot(N1, nrm));
        V3 Num = cross(N1, cross(sub(N1, B), Ac));
        double x = 1 - dot(B, N1);
        double x2 = x * x;
        double den = x2 * x2 * x;
        // (Num.N2)^2 + (Num.N3)^2 = |Num|^2 because Num is perpendicular to N1
        term = mag2(Num) / den / r2 * ndn;
      }
      sum += wave_sum(term);
      bdmax = fmax(bdmax, wave_max(bd));
      if (open) {
        const int last = 63 - __clzll(open);
        lastB = v3(__shfl(B.x, last), __shfl(B.y, last), __shfl(B.z, last));
      }
    }
    lev = L;
    double thisSum = sum * A.lv.dt_incl[L];
    prec = fabs(thisSum - lastsum) / lastsum;
    bool c = (L > 8 && prec < A.precision && bdmax < A.two_over_gamma) ||
             (L > 8 && bdmax < A.two_over_gamma && thisSum == lastsum);
    bool c2 = !c && L > 8 && thisSum == lastsum;
    if (c2) prec = 0;
    if (c || c2) done = true;
    else lastsum = thisSum;
  }
  if (lane == 0) {
    A.st.sum[6 * p] = sum;
    A.st.lastmag[p] = lastsum;
    A.st.prec[p] = prec;
    A.st.level[p] = lev;
    if (done) A.st.conv[p] = 1;
  }
}

// ---------------------------------------------------------------- total power
// src/OSCARSSR.cc:3785-3802: the integrand (|A/c|^2 - |B x A/c|^2) / (1 - B.B)^3
// summed over one level's samples.  Each workgroup sums a contiguous slice
// (strided over its lanes, then a fixed tree), and tp_final_kernel adds the
// slices in order: the result does not depend on timing.
constexpr int kTpThreads = 256;

__global__ void __launch_bounds__(kTpThreads) tp_level_kernel(TpArgs A) {
  __shared__ double red[kTpThreads];
  long per = (A.npts + gridDim.x - 1) / gridDim.x;
  long i0 = (long) blockIdx.x * per, i1 = i0 + per < A.npts ? i0 + per : A.npts;
  double s = 0;
  for (long i = i0 + threadIdx.x; i < i1; i += kTpThreads) {
    const double* r = A.rows + 10 * i;
    V3 B = v3(r[4], r[5], r[6]);
    V3 Ac = v3(r[7], r[8], r[9]);
    double x = 1 - dot(B, B);
    double g6 = 1. / (x * x * x);
    s += (mag2(Ac) - mag2(cross(B, Ac))) * g6;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kTpThreads / 2; k > 0; k >>= 1) {
    if ((int) threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) A.partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kTpThreads) tp_final_kernel(const double* __restrict__ partial, int n,
                                                              double* __restrict__ out) {
  __shared__ double red[kTpThreads];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += kTpThreads) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kTpThreads / 2; k > 0; k >>= 1) {
    if ((int) threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------- finalize

__global__ void __launch_bounds__(256) lw_finalize_kernel(FinArgs F) {
  long p = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= F.npts) return;
  int L = F.st.level[p];
  int conv = F.st.conv[p];
  double v;
  if (F.kind == 0) {
    double E[6];
    double f = F.C0 * F.dt_incl[L];
#pragma unroll
    for (int k = 0; k < 6; ++k) E[k] = F.st.sum[6 * p + k] * f;
    if (F.use_pol) pol_project(E, F.pol);
    if (F.out_e) {
#pragma unroll
      for (int k = 0; k < 6; ++k) F.out_e[6 * p + k] = E[k];
    }
    v = F.C2 * cmag2(E);
  } else {
    double s = F.st.sum[6 * p] * (F.pd_const * F.dt_incl[L]);
    s /= 1e6;
    if (!F.directional && s < 0) s *= -1;
    v = s;
  }
  if (F.quantity == 1) v = F.st.prec[p];
  else if (F.quantity == 2) v = conv ? (double) L : -1.0;
  v *= F.weight;
  if (F.accumulate) {
    double sum = F.out[p];
    double y = v - F.out_comp[p];
    double t = sum + y;
    F.out_comp[p] = (t - sum) - y;
    F.out[p] = t;
  } else {
    F.out[p] = v;
    F.out_comp[p] = 0;  // a new sum starts here
  }
  if (F.out_level) F.out_level[p] = L;
  if (F.out_conv) F.out_conv[p] = (unsigned char) conv;
}

// Listed points, one wave per point, the lanes splitting each level's
// samples, levels La..Lb from the point's state.  Used where few points are
// active (the last levels of a map, and the FP64 recomputation of the mixed
// path's borderline points): a wave per point keeps their latency low.
// MIXED: reduced-accuracy sin/cos with the error-bound bookkeeping of the per-point kernel.
// Four waves (points) per workgroup.
template <bool MIXED>
__global__ void __launch_bounds__(256) lw_wave_kernel(LwArgs A) {
  int lane = threadIdx.x & 63;
  long g = (long) blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= A.n_list) return;  // whole waves leave; no workgroup barrier below
  long p = A.list[g];
  if (A.st.conv[p] || A.st.level[p] != A.La - 1) return;
  long io = p / A.n_omega;
  long iw = p - io * A.n_omega;
  V3 obs = v3(A.obs[3 * io], A.obs[3 * io + 1], A.obs[3 * io + 2]);
  double negw = -A.omega[iw];
  double sum[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sum[k] = A.st.sum[6 * p + k];
  double lastmag = A.st.lastmag[p];
  double prec = A.st.prec[p];
  double an_run = 0, dm_last = 0;
  bool border = false;
  if (MIXED) {
    an_run = A.st.err[2 * p];
    dm_last = A.st.err[2 * p + 1];
  }
  int lev = A.La - 1;
  bool c = false;
  for (int L = A.La; L <= A.Lb && !c; ++L) {
    const double* __restrict__ rows = A.lv.rows[L];
    double tot[6], maxd, an;
    slice_level_sum_rows<MIXED>(rows, 0, 1L << L, obs, negw, lane, tot, maxd, an);
#pragma unroll
    for (int k = 0; k < 6; ++k) sum[k] += tot[k];
    lev = L;
    double mag_prev = lastmag;
    c = level_converged(A, L, sum, maxd, lastmag, prec);
    if (MIXED) {
      an_run += an;
      border = border ||
               level_borderline(L, A.lv.dt_incl[L], A.precision, mag_prev, lastmag, prec, maxd, an_run, dm_last);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) A.st.sum[6 * p + k] = sum[k];
    A.st.lastmag[p] = lastmag;
    A.st.prec[p] = prec;
    A.st.level[p] = lev;
    if (c) A.st.conv[p] = 1;
    if (MIXED) {
      A.st.err[2 * p] = an_run;
      A.st.err[2 * p + 1] = dm_last;
      if (border) A.st.border[p] = 1;
    }
  }
}

__global__ void __launch_bounds__(256) state_reset_kernel(State st, long npts) {
  long p = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
#pragma unroll
  for (int k = 0; k < 6; ++k) st.sum[6 * p + k] = 0;
  if (st.err) {
    st.err[2 * p] = 0;
    st.err[2 * p + 1] = 0;
  }
  if (st.border) st.border[p] = 0;
  st.lastmag[p] = -1;
  st.prec[p] = -1;
  st.level[p] = -1;
  st.conv[p] = 0;
}

// Points flagged by the mixed path, listed (any order) and reset to the start
// of the level loop for the FP64 recomputation.
// Slots for the threads of a 256-thread workgroup with `want` set, taken from
// *counter with one atomic per workgroup (a counter shared by every workgroup
// of a large grid serialises its atomics): wave counts are combined in LDS, the
// base is broadcast, and each thread's slot follows its wave's and lane's order.
// Every thread of the workgroup must call it (it contains barriers).
__device__ __forceinline__ int block_reserve(bool want, int* counter, int* sh /* 5 ints of LDS */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m = __ballot(want);
  if (lane == 0) sh[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int i = 0; i < 4; ++i) {
      int c = sh[i];
      sh[i] = tot;
      tot += c;
    }
    sh[4] = tot > 0 ? atomicAdd(counter, tot) : 0;
  }
  __syncthreads();
  return sh[4] + sh[w] + __popcll(m & ((1ULL << lane) - 1ULL));
}

__global__ void __launch_bounds__(256) border_list_kernel(State st, long npts, int* __restrict__ list,
                                                          int* __restrict__ count) {
  __shared__ int sh[5];
  long p = (long) blockIdx.x * blockDim.x + threadIdx.x;
  bool a = p < npts && st.border[p];
  int slot = block_reserve(a, count, sh);
  if (!a) return;
  list[slot] = (int) p;
#pragma unroll
  for (int k = 0; k < 6; ++k) st.sum[6 * p + k] = 0;
  st.lastmag[p] = -1;
  st.prec[p] = -1;
  st.level[p] = -1;
  st.conv[p] = 0;
}

__global__ void __launch_bounds__(256) count_active_kernel(State st, long npts, int* out) {
  long p = (long) blockIdx.x * blockDim.x + threadIdx.x;
  int a = (p < npts && !st.conv[p]) ? 1 : 0;
  // wave-level reduce then one atomic per wave (vector memory atomic)
  unsigned long long m = __ballot(a);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(out, __popcll(m));
}

// Tiles of the next level, one workgroup per observation point: its points
// still active (unconverged, summed up to level La - 1) in increasing energy
// order, padded with -1 to a multiple of 64 and appended at a position taken
// from a tile counter.  Converged lanes then no longer ride along in partly
// finished tiles (the order of the tiles does not affect any result).
__global__ void __launch_bounds__(256) tile_list_kernel(State st, long n_omega, int La, int* __restrict__ list,
                                                        int* __restrict__ ntiles) {
  __shared__ int wcount[4];
  __shared__ int base_s, total_s;
  long io = blockIdx.x;
  int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long p0 = io * n_omega;
  int cnt = 0;
  for (long e = tid; e < n_omega; e += 256) cnt += (!st.conv[p0 + e] && st.level[p0 + e] == La - 1) ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off);
  if (lane == 0) wcount[wv] = cnt;
  __syncthreads();
  if (tid == 0) {
    int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    total_s = total;
    base_s = total > 0 ? atomicAdd(ntiles, (total + 63) / 64) : 0;
  }
  __syncthreads();
  int total = total_s;
  if (total == 0) return;
  int* out = list + 64L * base_s;
  int running = 0;
  for (long c = 0; c < n_omega; c += 256) {
    long e = c + tid;
    bool a = e < n_omega && !st.conv[p0 + e] && st.level[p0 + e] == La - 1;
    unsigned long long m = __ballot(a);
    int before = __popcll(m & ((1ULL << lane) - 1ULL));
    __syncthreads();  // wcount of the previous chunk has been read
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int k = 0; k < wv; ++k) woff += wcount[k];
    if (a) out[running + woff + before] = (int) (p0 + e);
    running += wcount[0] + wcount[1] + wcount[2] + wcount[3];
  }
  int padded = (total + 63) / 64 * 64;
  for (int i = total + tid; i < padded; i += 256) out[i] = -1;
}

// Points still active at level La (unconverged, summed up to La - 1), compacted
// for the point kernel.  Every lane takes part in the counter atomic (lane 0
// adds its wave's count) and the wave's base comes from readfirstlane.
__global__ void __launch_bounds__(256) point_list_kernel(State st, long npts, int La, int* __restrict__ list,
                                                         int* __restrict__ count) {
  __shared__ int sh[5];
  long p = (long) blockIdx.x * blockDim.x + threadIdx.x;
  bool a = p < npts && !st.conv[p] && st.level[p] == La - 1;
  int slot = block_reserve(a, count, sh);
  if (a) list[slot] = (int) p;
}

// Photon energy [eV] to angular frequency, e * qe / hbar (the host's operations).
__global__ void __launch_bounds__(256) omega_from_ev_kernel(const double* __restrict__ e, long n,
                                                            double* __restrict__ om, double qe, double hbar) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) om[i] = e[i] * qe / hbar;
}

// Sum over points of 2^(L+1) - 1 (integer), highest level, unconverged count,
// into out[0..2] (cleared by the launcher): each thread takes kStatsPer points,
// waves reduce with shuffles, and thread 0 of each workgroup does the atomics.
constexpr int kStatsPer = 8;

__global__ void __launch_bounds__(256) level_stats_kernel(const int* __restrict__ level,
                                                          const unsigned char* __restrict__ conv, long n,
                                                          unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sh[3][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long ev = 0, nu = 0;
  int mx = -1;
  for (int k = 0; k < kStatsPer; ++k) {
    long p = ((long) blockIdx.x * kStatsPer + k) * blockDim.x + threadIdx.x;
    if (p < n) {
      int L = level[p];
      ev += L >= 0 ? (unsigned long long) ((2L << L) - 1) : 0ULL;
      nu += conv[p] ? 0ULL : 1ULL;
      mx = L > mx ? L : mx;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    ev += __shfl_down(ev, off);
    nu += __shfl_down(nu, off);
    int o = __shfl_down(mx, off);
    mx = o > mx ? o : mx;
  }
  if (lane == 0) {
    sh[0][w] = ev;
    sh[1][w] = nu;
    sh[2][w] = (unsigned long long) (mx + 1);  // level + 1 (0: none)
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long e = 0, u = 0, m = 0;
    for (int i = 0; i < 4; ++i) {
      e += sh[0][i];
      u += sh[1][i];
      m = sh[2][i] > m ? sh[2][i] : m;
    }
    atomicAdd(&out[0], e);
    atomicAdd(&out[2], u);
    atomicMax(&out[1], m);
  }
}

__global__ void __launch_bounds__(64) u64_fill_kernel(unsigned long long* __restrict__ p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0ULL;
}

__global__ void __launch_bounds__(64) int_fill_kernel(int* __restrict__ p, int n, int v) {
  int i = threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace om

// ---------------------------------------------------------------- launchers (host)
// Waves per workgroup: enough to put ~4096 waves on the chip when there are
// few output points, one wave per 64 points otherwise.
static int split_for(long groups) {
  long target = 4096;
  int s = 1;
  while (s < 8 && groups * s < target) s *= 2;
  return s;
}

template <bool M>
static void launch_point(const om::LwArgs* A, int split, hipStream_t s) {
  long blocks = ((A->list ? A->n_list : A->npts) + 63) / 64;
  dim3 g((unsigned) blocks);
  switch (split) {
    case 1: hipLaunchKernelGGL((om::lw_point_kernel<M, 1>), g, dim3(64), 0, s, *A); break;
    case 2: hipLaunchKernelGGL((om::lw_point_kernel<M, 2>), g, dim3(128), 0, s, *A); break;
    case 4: hipLaunchKernelGGL((om::lw_point_kernel<M, 4>), g, dim3(256), 0, s, *A); break;
    default: hipLaunchKernelGGL((om::lw_point_kernel<M, 8>), g, dim3(512), 0, s, *A); break;
  }
}

template <bool M>
static void launch_tile(const om::LwArgs* A, long n_obs, int split, hipStream_t s) {
  long ntw = (A->n_omega + 63) / 64;
  dim3 g((unsigned) (n_obs * ntw));
  switch (split) {
    case 1: hipLaunchKernelGGL((om::lw_tile_kernel<M, 1>), g, dim3(64), 0, s, *A, ntw); break;
    case 2: hipLaunchKernelGGL((om::lw_tile_kernel<M, 2>), g, dim3(128), 0, s, *A, ntw); break;
    case 4: hipLaunchKernelGGL((om::lw_tile_kernel<M, 4>), g, dim3(256), 0, s, *A, ntw); break;
    default: hipLaunchKernelGGL((om::lw_tile_kernel<M, 8>), g, dim3(512), 0, s, *A, ntw); break;
  }
}

// split <= 0 picks the number of waves per workgroup from the grid size.
extern "C" hipError_t om_launch_lw_point(const om::LwArgs* A, int mixed, int split, hipStream_t s) {
  if (split <= 0) split = split_for(((A->list ? A->n_list : A->npts) + 63) / 64);
  if (mixed) launch_point<true>(A, split, s); else launch_point<false>(A, split, s);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_lw_tile(const om::LwArgs* A, long n_obs, int mixed, int split, hipStream_t s) {
  if (split <= 0) split = split_for(n_obs * ((A->n_omega + 63) / 64));
  if (mixed) launch_tile<true>(A, n_obs, split, s); else launch_tile<false>(A, n_obs, split, s);
  return hipGetLastError();
}

// Compacted tiles: A->list holds ntiles x 64 point indices (launch_tile's grid
// is n_obs x ceil(n_omega/64) tiles, so pass ntiles as n_obs with n_omega <= 64
// tiles per "observation": the kernel reads io from the list instead).
extern "C" hipError_t om_launch_lw_tile_list(const om::LwArgs* A, long ntiles, int mixed, int split, hipStream_t s) {
  if (split <= 0) split = split_for(ntiles);
  dim3 g((unsigned) ntiles);
  if (mixed) {
    switch (split) {
      case 1: hipLaunchKernelGGL((om::lw_tile_kernel<true, 1>), g, dim3(64), 0, s, *A, 1L); break;
      case 2: hipLaunchKernelGGL((om::lw_tile_kernel<true, 2>), g, dim3(128), 0, s, *A, 1L); break;
      case 4: hipLaunchKernelGGL((om::lw_tile_kernel<true, 4>), g, dim3(256), 0, s, *A, 1L); break;
      default: hipLaunchKernelGGL((om::lw_tile_kernel<true, 8>), g, dim3(512), 0, s, *A, 1L); break;
    }
  } else {
    switch (split) {
      case 1: hipLaunchKernelGGL((om::lw_tile_kernel<false, 1>), g, dim3(64), 0, s, *A, 1L); break;
      case 2: hipLaunchKernelGGL((om::lw_tile_kernel<false, 2>), g, dim3(128), 0, s, *A, 1L); break;
      case 4: hipLaunchKernelGGL((om::lw_tile_kernel<false, 4>), g, dim3(256), 0, s, *A, 1L); break;
      default: hipLaunchKernelGGL((om::lw_tile_kernel<false, 8>), g, dim3(512), 0, s, *A, 1L); break;
    }
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_lw_wave(const om::LwArgs* A, int mixed, hipStream_t s) {
  long blocks = (A->n_list + 3) / 4;
  if (blocks > 0) {
    if (mixed)
      hipLaunchKernelGGL(om::lw_wave_kernel<true>, dim3((unsigned) blocks), dim3(256), 0, s, *A);
    else
      hipLaunchKernelGGL(om::lw_wave_kernel<false>, dim3((unsigned) blocks), dim3(256), 0, s, *A);
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_omega_from_ev(const double* e, long n, double* om, double qe, double hbar,
                                               hipStream_t s) {
  hipLaunchKernelGGL(om::omega_from_ev_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, e, n, om, qe,
                     hbar);
  return hipGetLastError();
}

// out[0] = sum of 2^(L+1)-1, out[1] = highest level + 1, out[2] = unconverged points.
// ---------------------------------------------------------------- allocation-free scratch
// Device scratch of om_flux_dev_noalloc: per-point state, level sample rows,
// lists and counters of one call are carved from it instead of allocated.
// Each GPU the library runs on holds its own copy for the life of the process;
// calls that overlap on one GPU share it and race.
namespace {
constexpr size_t kScratchBytes = (size_t) 64 << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

// The device address of `scratch` on `device`, which must be current: looked
// up once per device with hipGetSymbolAddress (in host code the array's name
// is not a device address) and kept.
extern "C" hipError_t om_scratch_setup(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**) &p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  *bytes = kScratchBytes;
  return hipSuccess;
}

extern "C" hipError_t om_launch_level_stats(const int* level, const unsigned char* conv, long n,
                                            unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(om::u64_fill_kernel, dim3(1), dim3(64), 0, s, out, 3);
  long per_block = 256L * om::kStatsPer;
  hipLaunchKernelGGL(om::level_stats_kernel, dim3((unsigned) ((n + per_block - 1) / per_block)), dim3(256), 0, s,
                     level, conv, n, out);
  return hipGetLastError();
}

// Lists (and resets) the points the mixed path flagged as borderline.
extern "C" hipError_t om_launch_border_list(const om::State* st, long npts, int* list, int* count, hipStream_t s) {
  hipLaunchKernelGGL(om::int_fill_kernel, dim3(1), dim3(64), 0, s, count, 1, 0);
  hipLaunchKernelGGL(om::border_list_kernel, dim3((unsigned) ((npts + 255) / 256)), dim3(256), 0, s, *st, npts, list,
                     count);
  return hipGetLastError();
}

// Builds the compacted point list for level La (counter cleared first by a kernel).
extern "C" hipError_t om_launch_point_list(const om::State* st, long npts, int La, int* list, int* count,
                                           hipStream_t s) {
  hipLaunchKernelGGL(om::int_fill_kernel, dim3(1), dim3(64), 0, s, count, 1, 0);
  hipLaunchKernelGGL(om::point_list_kernel, dim3((unsigned) ((npts + 255) / 256)), dim3(256), 0, s, *st, npts, La,
                     list, count);
  return hipGetLastError();
}

// Builds the compacted tile list for level La (counter cleared first by a kernel).
extern "C" hipError_t om_launch_tile_list(const om::State* st, long n_obs, long n_omega, int La, int* list,
                                          int* ntiles, hipStream_t s) {
  hipLaunchKernelGGL(om::int_fill_kernel, dim3(1), dim3(64), 0, s, ntiles, 1, 0);
  hipLaunchKernelGGL(om::tile_list_kernel, dim3((unsigned) n_obs), dim3(256), 0, s, *st, n_omega, La, list, ntiles);
  return hipGetLastError();
}

// Power density: S waves per 64 points, chosen so that the grid holds about
// 8192 waves (at least 8 per SIMD) unless split > 0 names S; a list of at most
// kPdWaveListMax points runs one wave per point.
constexpr long kPdWaveListMax = 8192;
extern "C" hipError_t om_launch_pd_point(const om::PdArgs* A, int mixed, int split, hipStream_t s) {
  if (A->list && A->n_list <= kPdWaveListMax && split <= 0) {  // few points: one wave each
    if (mixed) hipLaunchKernelGGL(om::pd_wave_kernel<true>, dim3((unsigned) A->n_list), dim3(64), 0, s, *A);
    else hipLaunchKernelGGL(om::pd_wave_kernel<false>, dim3((unsigned) A->n_list), dim3(64), 0, s, *A);
    return hipGetLastError();
  }
  long groups = ((A->list ? A->n_list : A->npts) + 63) / 64;
  if (split <= 0) {
    split = 1;
    while (split < 8 && groups * split < 8192) split *= 2;
  }
  dim3 g((unsigned) groups);
  if (mixed) {
    switch (split) {
      case 1: hipLaunchKernelGGL((om::pd_point_kernel<true, 1>), g, dim3(64), 0, s, *A); break;
      case 2: hipLaunchKernelGGL((om::pd_point_kernel<true, 2>), g, dim3(128), 0, s, *A); break;
      case 4: hipLaunchKernelGGL((om::pd_point_kernel<true, 4>), g, dim3(256), 0, s, *A); break;
      default: hipLaunchKernelGGL((om::pd_point_kernel<true, 8>), g, dim3(512), 0, s, *A); break;
    }
  } else {
    switch (split) {
      case 1: hipLaunchKernelGGL((om::pd_point_kernel<false, 1>), g, dim3(64), 0, s, *A); break;
      case 2: hipLaunchKernelGGL((om::pd_point_kernel<false, 2>), g, dim3(128), 0, s, *A); break;
      case 4: hipLaunchKernelGGL((om::pd_point_kernel<false, 4>), g, dim3(256), 0, s, *A); break;
      default: hipLaunchKernelGGL((om::pd_point_kernel<false, 8>), g, dim3(512), 0, s, *A); break;
    }
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_pd_stl(const om::StlArgs* A, hipStream_t s) {
  long blocks = (A->ntri + om::kStlWaves - 1) / om::kStlWaves;
  hipLaunchKernelGGL(om::pd_stl_kernel, dim3((unsigned) blocks), dim3(64 * om::kStlWaves), 0, s, *A);
  return hipGetLastError();
}

// partial must hold om_tp_blocks(npts) doubles
extern "C" int om_tp_blocks(long npts) {
  long b = (npts + 4L * om::kTpThreads - 1) / (4L * om::kTpThreads);
  return (int) (b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

extern "C" hipError_t om_launch_total_power_level(const om::TpArgs* A, hipStream_t s) {
  int blocks = om_tp_blocks(A->npts);
  hipLaunchKernelGGL(om::tp_level_kernel, dim3((unsigned) blocks), dim3(om::kTpThreads), 0, s, *A);
  hipLaunchKernelGGL(om::tp_final_kernel, dim3(1), dim3(om::kTpThreads), 0, s, (const double*) A->partial, blocks,
                     A->out);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_finalize(const om::FinArgs* F, hipStream_t s) {
  long blocks = (F->npts + 255) / 256;
  hipLaunchKernelGGL(om::lw_finalize_kernel, dim3((unsigned) blocks), dim3(256), 0, s, *F);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_state_reset(const om::State* st, long npts, hipStream_t s) {
  long blocks = (npts + 255) / 256;
  hipLaunchKernelGGL(om::state_reset_kernel, dim3((unsigned) blocks), dim3(256), 0, s, *st, npts);
  return hipGetLastError();
}

// Active (unconverged) points into *out (cleared first by a kernel).
extern "C" hipError_t om_launch_count_active(const om::State* st, long npts, int* out, hipStream_t s) {
  hipLaunchKernelGGL(om::int_fill_kernel, dim3(1), dim3(64), 0, s, out, 1, 0);
  long blocks = (npts + 255) / 256;
  hipLaunchKernelGGL(om::count_active_kernel, dim3((unsigned) blocks), dim3(256), 0, s, *st, npts, out);
  return hipGetLastError();
}
