// Trajectory kernels for gfx950: batched RK4 in the field models, the
// clamped cubic spline over the RK4 knots, and the per-level sample grids.
//
// Reference behaviour restated (paths relative to the reference root):
//   RK4 driver  src/OSCARSSR.cc:910-1009  (acceleration stored one step late)
//   derivatives src/OSCARSSR.cc:1247-1279, step src/OSCARSSR.cc:1321-1386
//   fields      src/TField3D_IdealUndulator.cc:113-152, src/TField3D_IdealEPU.cc:116-170,
//               src/TField3D_Grid.cc:153-278, sum src/TFieldContainer.cc:85-107
//   spline      include/TOMATH.h:92-203
//   levels      src/TParticleTrajectoryInterpolated.cc:124-161, 228-270,
//               src/TParticleTrajectoryInterpolatedPoints.cc:21-60
//
// Layout in HBM: knots are packed over particles; t[k] and p[9k..9k+8]
// (x y z bx by bz ax ay az) and ypp[9k..].  Level samples are AoS rows of 10
// doubles (t x y z bx by bz ax ay az); level L of a particle starts at row
// 2^L - 1 of that particle's level block, so one 80-byte row is what every
// lane of a wave reads when it walks the trajectory.
#include <string.h>

#include "om_device.h"
#include "om_types.h"

namespace om {

// ---------------------------------------------------------------- fields
// Written without pointers into the field description or run-time array
// indices, so a single field held by value stays in registers.
__device__ __forceinline__ void undulator_part(const FieldDev& f, double D, double PhaseShift, double a0, double a1,
                                               double a2, double taper_corr, double out[3], bool add) {
  double L = f.und_length, PL = f.period_length;
  if (D > L / 2. + PhaseShift || D < -L / 2. + PhaseShift) return;
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

__device__ __forceinline__ void field_eval(const FieldDev& f, double x, double y, double z, double out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (f.type == kUndulator || f.type == kEPU) {
    double D = (x - f.center[0]) * f.unit[0] + (y - f.center[1]) * f.unit[1] + (z - f.center[2]) * f.unit[2];
    double tc = 1 + D * f.taper;
    undulator_part(f, D, f.phase_shift_a, f.field[0], f.field[1], f.field[2], tc, out, false);
    if (f.type == kEPU) undulator_part(f, D, f.phase_shift_b, f.fieldB[0], f.fieldB[1], f.fieldB[2], tc, out, true);
    return;
  }
  double X[3] = {x, y, z};
  rot_xyz(X, f);
  X[0] -= f.trans[0]; X[1] -= f.trans[1]; X[2] -= f.trans[2];
  if (f.n[0] > 1 && (X[0] <= f.start[0] || X[0] >= f.stop[0])) return;
  if (f.n[1] > 1 && (X[1] <= f.start[1] || X[1] >= f.stop[1])) return;
  if (f.n[2] > 1 && (X[2] <= f.start[2] || X[2] >= f.stop[2])) return;
  int a = f.axis;
  double coord = a == 0 ? X[0] : (a == 1 ? X[1] : X[2]);
  double start = a == 0 ? f.start[0] : (a == 1 ? f.start[1] : f.start[2]);
  double step = a == 0 ? f.step[0] : (a == 1 ? f.step[1] : f.step[2]);
  double istep = a == 0 ? f.inv_step[0] : (a == 1 ? f.inv_step[1] : f.inv_step[2]);
  long na = a == 0 ? f.n[0] : (a == 1 ? f.n[1] : f.n[2]);
  long n0 = (long) (size_t) div_by(coord - start, step, istep);
  if (n0 > na - 2) n0 = na - 2;  // keeps the n0+1 read inside the map
  double dd = (coord - start) - (double) n0 * step;
  const double* lo = f.data + 3 * n0;
  double l0 = lo[0], l1 = lo[1], l2 = lo[2], h0 = lo[3], h1 = lo[4], h2 = lo[5];
  out[0] = (l0 + div_by((h0 - l0) * dd, step, istep)) * 1.0;
  out[1] = (l1 + div_by((h1 - l1) * dd, step, istep)) * 1.0;
  out[2] = (l2 + div_by((h2 - l2) * dd, step, istep)) * 1.0;
}

__device__ __forceinline__ void get_b(const FieldDev* __restrict__ F, int nf, double x, double y, double z,
                                      double B[3]) {
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = 0; i < nf; ++i) {
    double v[3];
    field_eval(F[i], x, y, z, v);
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
};

__device__ __forceinline__ double und_scale(const UndParams& f, double D, double PS) {
  double L2 = f.half_len;
  bool out = D > L2 + PS || D < -L2 + PS;
  bool term = D < -L2 + PS + f.PL || D > L2 + PS - f.PL;
  bool quarter = D < -L2 + PS + f.half_PL || D > L2 + PS - f.half_PL;
  double sc = term ? (quarter ? 0.25 : 0.75) : 1.0;
  return out ? 0.0 : sc;
}

__device__ __forceinline__ void und_eval(const UndParams& f, double x, double y, double z, double B[3]) {
  double D = (x - f.c0) * f.u0 + (y - f.c1) * f.u1 + (z - f.c2) * f.u2;
  double tc = 1 + D * f.taper;
  double scA = und_scale(f, D, f.psA);
  double sA = sin64(div_by(2.0 * kPi * (D - f.psA), f.PL, f.inv_PL));
  // terminated poles: (field*scale)*sin*taper; full poles: field*sin*taper (scale 1 is exact)
  double v0 = f.a0 * scA * sA * tc, v1 = f.a1 * scA * sA * tc, v2 = f.a2 * scA * sA * tc;
  if (f.epu) {
    double scB = und_scale(f, D, f.psB);
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
struct RkCtx {
  const FieldDev* F;
  int nf;
  double qomg;
  int err;
};

// ONE: exactly one undulator/EPU field, passed by value in f0 so it lives in registers.
template <bool ONE>
__device__ __forceinline__ void derivs(RkCtx& c, const UndParams& f0, const double* x, double* d) {
  // |beta| >= 1 test of src/OSCARSSR.cc:1260-1264; only its sign is used, so
  // the product with 1/c^2 stands in for the division without changing it
  // except within one rounding of |beta| = 1.
  double OneMinus = 1. - (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) * (1.0 / (kC * kC));
  if (OneMinus <= 0) {
    c.err = 1;
    return;
  }
  double B[3];
  if (ONE) {
    und_eval(f0, x[0], x[2], x[4], B);
  } else {
    get_b(c.F, c.nf, x[0], x[2], x[4], B);
  }
  d[0] = x[1];
  d[1] = c.qomg * (-x[5] * B[1] + x[3] * B[2]);
  d[2] = x[3];
  d[3] = c.qomg * (x[5] * B[0] - x[1] * B[2]);
  d[4] = x[5];
  d[5] = c.qomg * (x[1] * B[1] - x[3] * B[0]);
}

// Classic RK4 step; returns false when |beta| >= 1 was met (the reference then
// halves the step recursively; that branch never triggers for physical beams,
// so the kernel reports it as an error instead of recursing).
template <bool ONE>
__device__ __forceinline__ bool rk4(RkCtx& c, const UndParams& f0, double* y, const double* dydx, double h) {
  double dym[6], dyt[6], yt[6];
  double hh = h * 0.5;
  double h6 = h / 6.0;
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  derivs<ONE>(c, f0, yt, dyt);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  derivs<ONE>(c, f0, yt, dym);
  for (int i = 0; i < 6; ++i) {
    yt[i] = y[i] + h * dym[i];
    dym[i] += dyt[i];
  }
  derivs<ONE>(c, f0, yt, dyt);
  double yo[6];
  for (int i = 0; i < 6; ++i) yo[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  bool bad = 1. - (yo[1] * yo[1] + yo[3] * yo[3] + yo[5] * yo[5]) * (1.0 / (kC * kC)) <= 0;
  if (c.err || bad) return false;
  for (int i = 0; i < 6; ++i) y[i] = yo[i];
  return true;
}

// The knot is stored with velocity and acceleration in SI units; knot_finish_kernel
// divides them by c afterwards (in parallel), which keeps six divisions off the
// sequential RK4 chain and gives the same doubles.
__device__ __forceinline__ void store_knot(double* __restrict__ t, double* __restrict__ p, long k, double tt,
                                           const double* x, const double* d) {
  t[k] = tt;
  double* o = p + 9 * k;
  o[0] = x[0]; o[1] = x[2]; o[2] = x[4];
  o[3] = x[1]; o[4] = x[3]; o[5] = x[5];
  o[6] = d[1]; o[7] = d[3]; o[8] = d[5];
}

__global__ void __launch_bounds__(256) knot_finish_kernel(double* __restrict__ p, long total) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total * 6) return;
  long k = i / 6;
  int c = 3 + (int) (i - k * 6);
  p[9 * k + c] = p[9 * k + c] / kC;
}

// One thread per particle.  Forward points go to knot nb+i, backward points to
// knot nb-1-i, which is the order the reference reaches after its two array
// reversals.
// ONE: a single undulator/EPU arrives by value in the kernel arguments (read
// through the scalar cache into registers); otherwise fields are read from F.
template <bool ONE>
__global__ void __launch_bounds__(64) rk4_kernel(const UndParams f0, const FieldDev* __restrict__ F, int nfields,
                                                 TrajDesc* __restrict__ desc, int np, double* __restrict__ t,
                                                 double* __restrict__ p) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  TrajDesc D = desc[ip];
  RkCtx c{F, nfields, D.qomg, 0};
  double* tk = t + D.knot_offset;
  double* pk = p + 9 * D.knot_offset;
  double x[6], dxdt[6];
  x[0] = D.x0[0]; x[1] = D.b0[0] * kC;
  x[2] = D.x0[1]; x[3] = D.b0[1] * kC;
  x[4] = D.x0[2]; x[5] = D.b0[2] * kC;
  int status = 0;
  derivs<ONE>(c, f0, x, dxdt);
  for (long i = 0; i < D.nf; ++i) {
    double tt = D.t0_over_c + D.dt * (double) i;
    if (i == 0) derivs<ONE>(c, f0, x, dxdt);
    store_knot(tk, pk, D.nb + i, tt, x, dxdt);
    derivs<ONE>(c, f0, x, dxdt);
    if (!rk4<ONE>(c, f0, x, dxdt, D.dt)) { status = 1; break; }
  }
  x[0] = D.x0[0]; x[1] = D.b0[0] * kC;
  x[2] = D.x0[1]; x[3] = D.b0[1] * kC;
  x[4] = D.x0[2]; x[5] = D.b0[2] * kC;
  double dtr = -D.dt;
  for (long i = 0; i < D.nb && status == 0; ++i) {
    double t2 = D.t0_over_c + dtr * (double) (i + 1);
    derivs<ONE>(c, f0, x, dxdt);
    if (!rk4<ONE>(c, f0, x, dxdt, dtr)) { status = 1; break; }
    store_knot(tk, pk, D.nb - 1 - i, t2, x, dxdt);
  }
  if (c.err) status = 1;
  desc[ip].status = status;
}

// ---------------------------------------------------------------- spline
// include/TOMATH.h:141-203.  spline_pre_kernel computes, in parallel, the
// parts of the tridiagonal sweep that do not depend on the previous knot:
// sig_i, sig_i - 1 and w_i = (secant difference) * 6 / (t[i+1] - t[i-1]) per
// channel.  spline_kernel then runs the sequential recurrences
//   p = sig*ypp[i-1] + 2,  ypp[i] = (sig-1)/p,  u[i] = (w_i - sig*u[i-1])/p
// and the back substitution, one lane per (particle, channel).
__global__ void __launch_bounds__(256) spline_pre_kernel(const TrajDesc* __restrict__ desc, int np,
                                                         const double* __restrict__ t, const double* __restrict__ p,
                                                         double* __restrict__ sig, double* __restrict__ sm1,
                                                         double* __restrict__ u, long total) {
  long g = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  // particle of knot g: last descriptor with knot_offset <= g
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (desc[mid].knot_offset <= g) lo = mid; else hi = mid;
  }
  int ip = lo;
  long o = desc[ip].knot_offset;
  long N = desc[ip].n;
  long i = g - o;
  if (i < 1 || i >= N - 1) return;
  const double* X = t + o;
  double s = (X[i] - X[i - 1]) / (X[i + 1] - X[i - 1]);
  sig[g] = s;
  sm1[g] = s - 1.;
  const double* Y = p + 9 * o;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    double ui = (Y[9 * (i + 1) + c] - Y[9 * i + c]) / (X[i + 1] - X[i]) -
                (Y[9 * i + c] - Y[9 * (i - 1) + c]) / (X[i] - X[i - 1]);
    u[9 * g + c] = ui * 6. / (X[i + 1] - X[i - 1]);
  }
}

__global__ void __launch_bounds__(64) spline_kernel(const TrajDesc* __restrict__ desc, int np,
                                                    const double* __restrict__ t, const double* __restrict__ p,
                                                    const double* __restrict__ sig, const double* __restrict__ sm1,
                                                    double* __restrict__ ypp, double* __restrict__ u) {
  int lane = threadIdx.x;
  int ip = blockIdx.x;
  if (ip >= np || lane >= 9) return;
  long N = desc[ip].n;
  long o = desc[ip].knot_offset;
  if (N < 3) return;
  const double* X = t + o;
  const double* Y = p + 9 * o + lane;
  const double* SG = sig + o;
  const double* SM = sm1 + o;
  double* P = ypp + 9 * o + lane;
  double* U = u + 9 * o + lane;
  P[0] = -0.5;
  double dydx0 = (Y[9] - Y[0]) / (X[1] - X[0]);
  U[0] = ((Y[9] - Y[0]) / (X[1] - X[0]) - dydx0) * (3. / (X[1] - X[0]));
  double pprev = -0.5, uprev = U[0];
  double s_n = SG[1], m_n = SM[1], w_n = U[9];
  for (long i = 1; i < N - 1; ++i) {
    double s = s_n, m = m_n, w = w_n;
    if (i + 1 < N - 1) {  // next knot's inputs, loaded ahead of the dependent chain
      s_n = SG[i + 1];
      m_n = SM[i + 1];
      w_n = U[9 * (i + 1)];
    }
    double pp = s * pprev + 2.0;
    double yi = m / pp;
    double ui = (w - s * uprev) / pp;
    P[9 * i] = yi;
    U[9 * i] = ui;
    pprev = yi;
    uprev = ui;
  }
  double qn = 0.5;
  double xa = X[N - 2], xb = X[N - 1], ya = Y[9 * (N - 2)], yb = Y[9 * (N - 1)];
  double dydxN = (yb - ya) / (xb - xa);
  double un = (dydxN - (yb - ya) / (xb - xa)) * (3. / (xb - xa));
  double last = (un - uprev * qn) / (pprev * qn + 1.);
  P[9 * (N - 1)] = last;
  // back substitution in blocks of 8 knots: the 16 loads of a block are issued
  // before its dependent chain of fused updates
  long k = N - 2;
  while (k >= 7) {
    double pk[8], uk[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      pk[j] = P[9 * (k - j)];
      uk[j] = U[9 * (k - j)];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      last = pk[j] * last + uk[j];
      P[9 * (k - j)] = last;
    }
    k -= 8;
  }
  for (; k >= 0; --k) {
    last = P[9 * k] * last + U[9 * k];
    P[9 * k] = last;
  }
}

// ---------------------------------------------------------------- level samples
struct LevelJob {
  const double* t;     // packed knot times
  const double* p;
  const double* ypp;
  long knot_offset;
  long n;
  double tstart, tstop;
  double* out;          // destination rows (10 doubles) for this level
  long npts;            // 2^L
  double t_first, spacing;
};

// include/TOMATH.h:92-126 GetValue.  The interval is found from a guess on the
// uniform knot grid and then moved to the reference bisection's answer
// (largest klo <= n-2 with t[klo] <= x), which is exact for increasing knots.
__device__ __forceinline__ void spline_value(const double* __restrict__ X, const double* __restrict__ Y,
                                             const double* __restrict__ YPP, long n, double x, double* out) {
  double h0 = (X[n - 1] - X[0]) / (double) (n - 1);
  long k = (long) ((x - X[0]) / h0);
  if (k < 0) k = 0;
  if (k > n - 2) k = n - 2;
  while (k < n - 2 && X[k + 1] <= x) ++k;
  while (k > 0 && X[k] > x) --k;
  long klo = k, khi = k + 1;
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

__global__ void __launch_bounds__(256) level_kernel(LevelJob J) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J.npts) return;
  double tt = J.t_first + J.spacing * (double) i;
  double v[9];
  spline_value(J.t + J.knot_offset, J.p + 9 * J.knot_offset, J.ypp + 9 * J.knot_offset, J.n, tt, v);
  double* o = J.out + 10 * i;
  o[0] = tt;
#pragma unroll
  for (int c = 0; c < 9; ++c) o[1 + c] = v[c];
}

// Largest |beta_i - beta_{i-1}| over one level (src/OSCARSSR.cc:2884-2890);
// it does not depend on the observation point, so it is computed once per level.
__global__ void __launch_bounds__(256) beta_diff_kernel(const double* __restrict__ rows, long npts,
                                                        double* __restrict__ out) {
  __shared__ double red[256];
  double m = -1;
  for (long i = 1 + threadIdx.x; i < npts; i += blockDim.x) {
    const double* a = rows + 10 * i + 4;
    const double* b = rows + 10 * (i - 1) + 4;
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    m = d > m ? d : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int) threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x + s] > red[threadIdx.x] ? red[threadIdx.x + s] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace om

// ---------------------------------------------------------------- launchers (host)
extern "C" hipError_t om_launch_rk4(const om::FieldDev* F, const om::FieldDev* F0_host, int nfields,
                                    om::TrajDesc* desc, int np, double* t, double* p, hipStream_t s) {
  int threads = 64;
  int blocks = (np + threads - 1) / threads;
  om::UndParams u;
  memset(&u, 0, sizeof(u));
  const om::FieldDev& f = *F0_host;
  if (nfields == 1 && (f.type == om::kUndulator || f.type == om::kEPU)) {
    u.c0 = f.center[0]; u.c1 = f.center[1]; u.c2 = f.center[2];
    u.u0 = f.unit[0]; u.u1 = f.unit[1]; u.u2 = f.unit[2];
    u.taper = f.taper;
    u.half_len = f.und_length / 2.;
    u.PL = f.period_length;
    u.inv_PL = f.inv_period_length;
    u.half_PL = f.period_length / 2.;
    u.psA = f.phase_shift_a;
    u.psB = f.phase_shift_b;
    u.a0 = f.field[0]; u.a1 = f.field[1]; u.a2 = f.field[2];
    u.b0 = f.fieldB[0]; u.b1 = f.fieldB[1]; u.b2 = f.fieldB[2];
    u.epu = f.type == om::kEPU;
    hipLaunchKernelGGL(om::rk4_kernel<true>, dim3(blocks), dim3(threads), 0, s, u, F, nfields, desc, np, t, p);
  } else {
    hipLaunchKernelGGL(om::rk4_kernel<false>, dim3(blocks), dim3(threads), 0, s, u, F, nfields, desc, np, t, p);
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_knot_finish(double* p, long total, hipStream_t s) {
  long blocks = (total * 6 + 255) / 256;
  hipLaunchKernelGGL(om::knot_finish_kernel, dim3((unsigned) blocks), dim3(256), 0, s, p, total);
  return hipGetLastError();
}

// sig and sm1 are scratch arrays of one double per knot
extern "C" hipError_t om_launch_spline(const om::TrajDesc* desc, int np, const double* t, const double* p,
                                       double* ypp, double* u, double* sig, double* sm1, long total,
                                       hipStream_t s) {
  long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(om::spline_pre_kernel, dim3((unsigned) blocks), dim3(256), 0, s, desc, np, t, p, sig, sm1, u,
                     total);
  hipLaunchKernelGGL(om::spline_kernel, dim3(np), dim3(64), 0, s, desc, np, t, p, sig, sm1, ypp, u);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_level(const om::LevelJob* J, hipStream_t s) {
  long blocks = (J->npts + 255) / 256;
  hipLaunchKernelGGL(om::level_kernel, dim3((unsigned) blocks), dim3(256), 0, s, *J);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_beta_diff(const double* rows, long npts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(om::beta_diff_kernel, dim3(1), dim3(256), 0, s, rows, npts, out);
  return hipGetLastError();
}
