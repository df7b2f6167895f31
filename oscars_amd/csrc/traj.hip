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
#include "launch_types.h"
#include "om_types.h"
#include "rk4_common.h"

namespace om {

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

// ---------------------------------------------------------------- paired stages
// The fields depend on position and time, never on the velocity.  In an RK4
// step from (t, y) with h2 = h/2 the four stages take their fields at the
// positions y_p, y_p + h2 k1_p, y_p + h2 k2_p and y_p + h k3_p (times t, t + h2,
// t + h2, t + h), and the position rows of a slope are the stage input's
// velocity: k1_p = y_v, k2_p = (y + h2 k1)_v, k3_p = (y + h2 k2)_v.  So the
// fields of stages 1 and 2 follow from y alone and those of stages 3 and 4 once
// k1 and k2 are known: the step's four field evaluations pair up, (1, 2) and
// then (3, 4), and two lanes evaluate a pair at once and exchange the results.
// Everything else runs in both lanes in the order of derivs and rk4_try, so the
// doubles are theirs.
// Scalar members (an array member would make the lane selects go through
// scratch memory).
struct StageField {
  double b0, b1, b2;  // kAxisUnd*: b0 is the field component, b1 = b2 = 0
  double e0, e1, e2;  // kFieldsEB: the electric field
};

// x in the lane of the other stage of a field pair: the neighbouring lane, or
// for kEpuZ (whose neighbouring lanes hold the EPU's two parts) the lane two
// away, by a DPP quad permutation of both halves: exact.
template <int M>
__device__ __forceinline__ double stage_other(double x) {
  constexpr int Q = M == kEpuZ ? 0x4E : 0xB1;
  long long b = __builtin_bit_cast(long long, x);
  int lo = __builtin_amdgcn_update_dpp(0, (int) (b & 0xffffffff), Q, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), Q, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long) hi << 32) | (unsigned) lo);
}

// grid_z_eval without the cached cell and without branches: the lanes of a
// pair evaluate different stages, so a cache miss in one lane would run the
// reload with the other masked off, and the test itself sits on the chain.
// The cell index is clamped into the map (the value is then replaced by the
// zero outside it), and the samples are read from LDS every time; the doubles
// are grid_z_eval's.
__device__ __forceinline__ void grid_z_direct(const GridZ& g, double z, double B[3]) {
  double coord = z - g.trans;
  bool out = coord <= g.start || coord >= g.stop;
  int n0 = cell_of(div_by(coord - g.start, g.step, g.inv_step));
  n0 = n0 > (int) (g.n - 2) ? (int) (g.n - 2) : n0;
  n0 = n0 < 0 ? 0 : n0;
  const double* lo = g.data + 3 * n0;
  double dd = (coord - g.start) - (double) n0 * g.step;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = lo[k] + div_by((lo[3 + k] - lo[k]) * dd, g.step, g.inv_step);
    B[k] = out ? 0.0 : v;
  }
}

// The field part of derivs<M> at (px, py, pz, t).
template <int M>
__device__ __forceinline__ StageField field_at(RkCtx& c, const UndParams& f0, double px, double py, double pz,
                                               double t) {
  StageField F;
  F.e0 = F.e1 = F.e2 = 0.0;
  double B[3] = {0.0, 0.0, 0.0}, E[3] = {0.0, 0.0, 0.0};
  if (M >= kAxisUndX && M <= kAxisUndZ) {
    constexpr int K = M - kAxisUndX;
    double D = pz - f0.c2;
    double sc = und_scale(f0.thA, D);
    double a = K == 0 ? f0.a0 : (K == 1 ? f0.a1 : f0.a2);
    double s = sin64(div_by(2.0 * kPi * (D - f0.psA), f0.PL, f0.inv_PL));
    F.b0 = sc != 0.0 ? a * sc * s : 0.0;
    F.b1 = F.b2 = 0.0;
    return F;
  } else if (M == kUndZ || M == kEpuZ) {
    double D = pz - f0.c2;
    double scA = und_scale(f0.thA, D);
    double sA = sin64(div_by(2.0 * kPi * (D - f0.psA), f0.PL, f0.inv_PL));
    double v0 = f0.a0 * scA * sA, v1 = f0.a1 * scA * sA, v2 = f0.a2 * scA * sA;
    if (M == kEpuZ) {
      // this lane's part (A or B, see rk4_kernel) plus the neighbour's, as in derivs
      v0 = v0 + lane_pair_other(v0);
      v1 = v1 + lane_pair_other(v1);
      v2 = v2 + lane_pair_other(v2);
    }
    F.b0 = 0.0 + v0;
    F.b1 = 0.0 + v1;
    F.b2 = 0.0 + v2;
    return F;
  } else if (M == kGridZ) {
    grid_z_direct(c.g, pz, B);
  } else if (M == kOneUndulator) {
    und_eval(f0, px, py, pz, B);
  } else if (M == kFieldsEB) {
    get_b(c.F, c.nf, px, py, pz, t, B);
    get_b(c.F + c.nf, c.nfe, px, py, pz, t, E);
  } else {
    get_b(c.F, c.nf, px, py, pz, t, B);
  }
  F.b0 = B[0];
  F.b1 = B[1];
  F.b2 = B[2];
  F.e0 = E[0];
  F.e1 = E[1];
  F.e2 = E[2];
  return F;
}

// The rest of derivs<M>, given the fields at the position of x.
template <int M>
__device__ __forceinline__ void derivs_with(RkCtx& c, const StageField& F, const double* x, double* d) {
  if (M == kFieldsEB) {
    if (beta_at_least_one(x[1], x[3], x[5])) c.err = 1;
    double bb = (x[1] * x[1] + x[3] * x[3] + x[5] * x[5]) / (kC * kC);
    double q = c.qom * sqrt(1. - bb);
    double be = (x[1] * F.e0 + x[3] * F.e1 + x[5] * F.e2) / kC;
    d[0] = x[1];
    d[1] = q * (F.e0 - x[5] * F.b1 + x[3] * F.b2 - x[1] * be / kC);
    d[2] = x[3];
    d[3] = q * (F.e1 + x[5] * F.b0 - x[1] * F.b2 - x[3] * be / kC);
    d[4] = x[5];
    d[5] = q * (F.e2 + x[1] * F.b1 - x[3] * F.b0 - x[5] * be / kC);
    return;
  }
  const bool bad = beta_at_least_one(x[1], x[3], x[5]);
  if (M >= kAxisUndX && M <= kAxisUndZ) {
    constexpr int K = M - kAxisUndX;
    double v = F.b0;
    if (K == 0) {
      d[1] = 0.0;
      d[3] = c.qomg * (x[5] * v);
      d[5] = c.qomg * (-(x[3] * v));
    } else if (K == 1) {
      d[1] = c.qomg * (-(x[5] * v));
      d[3] = 0.0;
      d[5] = c.qomg * (x[1] * v);
    } else {
      d[1] = c.qomg * (x[3] * v);
      d[3] = c.qomg * (-(x[1] * v));
      d[5] = 0.0;
    }
  } else {
    d[1] = c.qomg * (-x[5] * F.b1 + x[3] * F.b2);
    d[3] = c.qomg * (x[5] * F.b0 - x[1] * F.b2);
    d[5] = c.qomg * (x[1] * F.b1 - x[3] * F.b0);
  }
  d[0] = x[1];
  d[2] = x[3];
  d[4] = x[5];
  c.err |= bad ? 1 : 0;
}

// The fields at A (evaluated by the slot-0 lane) and B (slot 1), both
// returned in both lanes.
template <int M>
__device__ __forceinline__ void field_pair(RkCtx& c, const UndParams& f0, int slot, double xa, double ya, double za,
                                           double ta, double xb, double yb, double zb, double tb, StageField& FA,
                                           StageField& FB) {
  StageField m = field_at<M>(c, f0, slot ? xb : xa, slot ? yb : ya, slot ? zb : za, slot ? tb : ta);
  constexpr bool one = M >= kAxisUndX && M <= kAxisUndZ;
  StageField o;
  o.b0 = stage_other<M>(m.b0);
  o.b1 = one ? 0.0 : stage_other<M>(m.b1);
  o.b2 = one ? 0.0 : stage_other<M>(m.b2);
  o.e0 = M == kFieldsEB ? stage_other<M>(m.e0) : 0.0;
  o.e1 = M == kFieldsEB ? stage_other<M>(m.e1) : 0.0;
  o.e2 = M == kFieldsEB ? stage_other<M>(m.e2) : 0.0;
  FA.b0 = slot ? o.b0 : m.b0;
  FA.b1 = slot ? o.b1 : m.b1;
  FA.b2 = slot ? o.b2 : m.b2;
  FA.e0 = slot ? o.e0 : m.e0;
  FA.e1 = slot ? o.e1 : m.e1;
  FA.e2 = slot ? o.e2 : m.e2;
  FB.b0 = slot ? m.b0 : o.b0;
  FB.b1 = slot ? m.b1 : o.b1;
  FB.b2 = slot ? m.b2 : o.b2;
  FB.e0 = slot ? m.e0 : o.e0;
  FB.e1 = slot ? m.e1 : o.e1;
  FB.e2 = slot ? m.e2 : o.e2;
}

// derivs<M> at (t, y) into dydx followed by rk4<M>(y, dydx, t, h), with the
// step's fields evaluated in lane pairs.  A step that meets |beta| >= 1 is
// redone by the halving walk from y with the slope dydx, as in rk4.
template <int M>
__device__ __forceinline__ bool rk4_paired(RkCtx& c, const UndParams& f0, int slot, double* y, double* dydx, double t,
                                           double h) {
  double hh = h * 0.5;
  double h6 = h / 6.0;
  double xh = t + hh;
  double k2[6], k3[6], k4[6], yt[6], y3[6], yo[6];
  StageField F1, F2, F3, F4;
  field_pair<M>(c, f0, slot, y[0], y[2], y[4], t, y[0] + hh * y[1], y[2] + hh * y[3], y[4] + hh * y[5], xh, F1, F2);
  derivs_with<M>(c, F1, y, dydx);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];  // its positions are stage 2's above
  derivs_with<M>(c, F2, yt, k2);
  for (int i = 0; i < 6; ++i) y3[i] = y[i] + hh * k2[i];
  field_pair<M>(c, f0, slot, y3[0], y3[2], y3[4], xh, y[0] + h * y3[1], y[2] + h * y3[3], y[4] + h * y3[5], t + h,
                F3, F4);
  derivs_with<M>(c, F3, y3, k3);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + h * k3[i];  // its positions are stage 4's above
  derivs_with<M>(c, F4, yt, k4);
  for (int i = 0; i < 6; ++i) yo[i] = y[i] + h6 * (dydx[i] + k4[i] + 2.0 * (k3[i] + k2[i]));
  if (c.err || beta_at_least_one(yo[1], yo[3], yo[5])) {
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

// Lanes per particle.  The forward and the backward integration both start
// from the state at T0 and never meet (src/OSCARSSR.cc:943-1006), so they run
// in two halves of a particle's lanes: a window with T0 inside costs the longer
// of the two instead of their sum.  Each direction takes two lanes, one per
// stage of a field pair (bit 0: stage, bit 1: direction); kEpuZ two more, the
// EPU's parts A and B (bit 0: part, bit 1: stage, bit 2: direction).
template <int M>
constexpr int rk4_lanes() {
  return M == kEpuZ ? 8 : 4;
}

// Forward points go to knot nb+i, backward points to knot nb-1-i, which is the
// order the reference reaches after its two array reversals.  The loop is the
// same for both directions (step h = +-dt, time T0/c + h i): the forward lane
// stores the initial knot first and each later knot after the step that
// reaches it, like the reference's store before the next step; its last step
// is taken and not stored, as there (after the loop, so that every step of the
// loop stores).  The fErrorGamma flag (c.err) that the
// reference carries from the forward into the backward loop is always clear
// after a completed RK4 step, so it reaches the backward loop only when there
// is no forward step, which the backward lane reproduces.
// ONE: a single undulator/EPU arrives by value in the kernel arguments (read
// through the scalar cache into registers); otherwise fields are read from F.
template <int M>
__global__ void __launch_bounds__(64) rk4_kernel(const UndParams f_arg, const GridZ g, const FieldDev* __restrict__ F,
                                                 int nfields, TrajDesc* __restrict__ desc, int np,
                                                 double* __restrict__ t, double* __restrict__ p, int nefields) {
  extern __shared__ double lds_grid[];
  // Highest issue priority for these waves: the integration is one dependent
  // chain per particle, and when it runs beside other kernels (the previous
  // step's sums) it should not wait behind their waves on its SIMD.
  __builtin_amdgcn_s_setprio(3);
  constexpr int P = rk4_lanes<M>();
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int ip = gid / P;
  const int sub = gid % P;
  const int dir = sub / (P / 2);                   // 0 forward, 1 backward
  const int slot = M == kEpuZ ? (sub >> 1) & 1 : sub & 1;  // the stage of a field pair this lane evaluates
  RkCtx c{F, nfields, 0.0, 0, g, nefields, 0.0, 0};
  if (M == kGridZ) {
    // stage the map in LDS (all lanes), then integrate
    for (long i = threadIdx.x; i < 3 * g.n; i += blockDim.x) lds_grid[i] = g.data[i];
    __syncthreads();
    c.g.data = lds_grid;
  }
  if (ip >= np) return;  // all lanes of a particle together
  // the single undulator's constants, in vector registers (see in_vgpr)
  UndParams f0 = (M == kOneUndulator || (M >= kAxisUndX && M <= kAxisUndZ) || M == kUndZ || M == kEpuZ)
                     ? und_in_vgpr(f_arg) : f_arg;
  if (M == kEpuZ && (sub & 1)) {
    // this lane evaluates part B in the place of part A
    f0.psA = f0.psB;
    f0.a0 = f0.b0; f0.a1 = f0.b1; f0.a2 = f0.b2;
    for (int k = 0; k < 6; ++k) f0.thA[k] = f0.thB[k];
  }
  TrajDesc D = desc[ip];
  c.qomg = D.qomg;
  c.qom = D.qom;
  double* tk = t + D.knot_offset;
  double* pk = p + 9 * D.knot_offset;
  double x[6], dxdt[6];
  x[0] = D.x0[0]; x[1] = D.b0[0] * kC;
  x[2] = D.x0[1]; x[3] = D.b0[1] * kC;
  x[4] = D.x0[2]; x[5] = D.b0[2] * kC;
  const bool fwd = dir == 0 && D.nf > 0;
  // steps whose result is stored: nb backward, nf - 1 forward
  const long n = dir ? D.nb : (D.nf > 0 ? D.nf - 1 : 0);
  const double h = dir ? -D.dt : D.dt;
  int status = 0;
  if (dir == 0 || D.nf == 0) derivs<M>(c, f0, D.t0_over_c, x, dxdt);
  if (fwd) {
    derivs<M>(c, f0, D.t0_over_c + h * 0.0, x, dxdt);
    store_knot(tk, pk, D.nb, D.t0_over_c + h * 0.0, x, dxdt);
  }
  for (long i = 0; i < n; ++i) {
    double t1 = D.t0_over_c + h * (double) i;
    if (!rk4_paired<M>(c, f0, slot, x, dxdt, t1, h)) { status = 1; break; }
    store_knot(tk, pk, dir ? D.nb - 1 - i : D.nb + 1 + i, D.t0_over_c + h * (double) (i + 1), x, dxdt);
  }
  if (fwd && status == 0) {
    // the forward's last step: taken (its halving can fail) and not stored
    double t1 = D.t0_over_c + h * (double) (D.nf - 1);
    if (!rk4_paired<M>(c, f0, slot, x, dxdt, t1, h)) status = 1;
  }
  // the two directions' outcomes, combined in the lanes of the particle
  int halved = c.halved;
  status |= __shfl_xor(status, P / 2);
  halved += __shfl_xor(halved, P / 2);
  if (sub == 0) {
    desc[ip].status = status;
    desc[ip].halved = halved;
  }
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

// The three recurrences of the sweep (p_i = sig_i P_{i-1} + 2, P_i = (sig_i-1)/p_i;
// U_i = (w_i - sig_i U_{i-1})/p_i; Y''_k = P_k Y''_{k+1} + U_k) are contractions on
// the uniform RK4 knot grid (factors ~0.07, ~0.14 and ~0.27 per knot).  Every
// knot therefore evaluates its own value by running the same operations over
// the previous (or next) kWin = 64 knots, starting from an approximate value:
// the starting error shrinks below 1e-36 of the result, the last steps see the
// sequential sweep's inputs and produce its doubles.  Only entries that are
// ~1e-20 of their channel's scale (flat stretches where Y'' vanishes) can come
// out a few units different.  Exactly as in the sequential sweep, P_0 = -1/2,
// U_0 is the (zero) secant-slope term and Y''_{N-1} = (un - U_{N-2}/2)/(P_{N-2}/2 + 1).
constexpr long kSplineWin = 64;

__device__ __forceinline__ int particle_of_knot(const TrajDesc* __restrict__ desc, int np, long g) {
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (desc[mid].knot_offset <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// P_i (into pv) and p_i (into ppv) per knot
__global__ void __launch_bounds__(256) spline_p_kernel(const TrajDesc* __restrict__ desc, int np,
                                                       const double* __restrict__ sig, const double* __restrict__ sm1,
                                                       double* __restrict__ pv, double* __restrict__ ppv, long total) {
  long g = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int ip = particle_of_knot(desc, np, g);
  long o = desc[ip].knot_offset, N = desc[ip].n, i = g - o;
  if (N < 3 || i > N - 2) return;
  if (i == 0) {
    pv[g] = -0.5;
    ppv[g] = 1.0;
    return;
  }
  long j0 = i - kSplineWin > 1 ? i - kSplineWin : 1;
  double P = j0 == 1 ? -0.5 : -0.26794919243112270;  // the fixed point for sig = 1/2
  double pp = 1.0;
  for (long j = j0; j <= i; ++j) {
    pp = sig[o + j] * P + 2.0;
    P = sm1[o + j] / pp;
  }
  pv[g] = P;
  ppv[g] = pp;
}

// U_i per (knot, channel); w_i is in u (from spline_pre_kernel), U_i goes to u2
__global__ void __launch_bounds__(256) spline_u_kernel(const TrajDesc* __restrict__ desc, int np,
                                                       const double* __restrict__ t, const double* __restrict__ p,
                                                       const double* __restrict__ sig, const double* __restrict__ ppv,
                                                       const double* __restrict__ u, double* __restrict__ u2,
                                                       long total) {
  long gg = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (gg >= 9 * total) return;
  long g = gg / 9;
  int c = (int) (gg - 9 * g);
  int ip = particle_of_knot(desc, np, g);
  long o = desc[ip].knot_offset, N = desc[ip].n, i = g - o;
  if (N < 3 || i > N - 2) return;
  const double* X = t + o;
  const double* Y = p + 9 * o + c;
  double dydx0 = (Y[9] - Y[0]) / (X[1] - X[0]);
  double u0 = ((Y[9] - Y[0]) / (X[1] - X[0]) - dydx0) * (3. / (X[1] - X[0]));
  if (i == 0) {
    u2[9 * g + c] = u0;
    return;
  }
  long j0 = i - kSplineWin > 1 ? i - kSplineWin : 1;
  double U = j0 == 1 ? u0 : 0.0;
  for (long j = j0; j <= i; ++j) U = (u[9 * (o + j) + c] - sig[o + j] * U) / ppv[o + j];
  u2[9 * g + c] = U;
}

// Y''_k per (knot, channel), back substitution from the end (or from k + kWin)
__global__ void __launch_bounds__(256) spline_b_kernel(const TrajDesc* __restrict__ desc, int np,
                                                       const double* __restrict__ t, const double* __restrict__ p,
                                                       const double* __restrict__ pv, const double* __restrict__ u2,
                                                       double* __restrict__ ypp, long total) {
  long gg = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (gg >= 9 * total) return;
  long g = gg / 9;
  int c = (int) (gg - 9 * g);
  int ip = particle_of_knot(desc, np, g);
  long o = desc[ip].knot_offset, N = desc[ip].n, k = g - o;
  if (N < 3) return;
  const double* X = t + o;
  const double* Y = p + 9 * o + c;
  double qn = 0.5;
  double xa = X[N - 2], xb = X[N - 1], ya = Y[9 * (N - 2)], yb = Y[9 * (N - 1)];
  double dydxN = (yb - ya) / (xb - xa);
  double un = (dydxN - (yb - ya) / (xb - xa)) * (3. / (xb - xa));
  double last = (un - u2[9 * (o + N - 2) + c] * qn) / (pv[o + N - 2] * qn + 1.);
  if (k == N - 1) {
    ypp[9 * g + c] = last;
    return;
  }
  long j1 = k + kSplineWin < N - 1 ? k + kSplineWin : N - 1;
  double v = j1 == N - 1 ? last : 0.0;
  for (long j = j1 - 1; j >= k; --j) v = pv[o + j] * v + u2[9 * (o + j) + c];
  ypp[9 * g + c] = v;
}

// Trajectories given by the user (any increasing knot times): the sequential
// sweep of include/TOMATH.h:141-203 itself, one lane per channel c0..c0+nc-1
// of the nine knot channels; u is scratch of the knot layout.
__global__ void __launch_bounds__(64) spline_seq_kernel(const double* __restrict__ X, const double* __restrict__ p,
                                                        double* __restrict__ ypp, double* __restrict__ u, long n,
                                                        int c0, int nc) {
  if ((int) threadIdx.x >= nc) return;
  const int c = c0 + (int) threadIdx.x;
  const double* Y = p + c;
  double* P = ypp + c;
  double* U = u + c;
  const double dydx0 = (Y[9] - Y[0]) / (X[1] - X[0]);
  P[0] = -0.5;
  U[0] = (3. / (X[1] - X[0])) * ((Y[9] - Y[0]) / (X[1] - X[0]) - dydx0);
  for (long i = 1; i < n - 1; ++i) {
    const double sig = (X[i] - X[i - 1]) / (X[i + 1] - X[i - 1]);
    const double pp = sig * P[9 * (i - 1)] + 2.0;
    P[9 * i] = (sig - 1.) / pp;
    const double ui = (Y[9 * (i + 1)] - Y[9 * i]) / (X[i + 1] - X[i]) - (Y[9 * i] - Y[9 * (i - 1)]) / (X[i] - X[i - 1]);
    U[9 * i] = (6. * ui / (X[i + 1] - X[i - 1]) - sig * U[9 * (i - 1)]) / pp;
  }
  const double qn = 0.5;
  const double hN = X[n - 1] - X[n - 2];
  const double dydxN = (Y[9 * (n - 1)] - Y[9 * (n - 2)]) / hN;
  const double un = (3. / hN) * (dydxN - (Y[9 * (n - 1)] - Y[9 * (n - 2)]) / hN);
  P[9 * (n - 1)] = (un - qn * U[9 * (n - 2)]) / (qn * P[9 * (n - 2)] + 1.);
  for (long k = n - 2; k >= 0; --k) P[9 * k] = P[9 * k] * P[9 * (k + 1)] + U[9 * k];
}

// Spline slope at the knots (include/TOMATH.h:131-139 GetDerivative), divided
// by div, from channels c_src..c_src+2 into c_dst..c_dst+2: the reference's
// construction of beta from positions (div = c) and of a/c from beta (div = 1)
// (src/TParticleTrajectoryPoints.cc:820-862).
__global__ void __launch_bounds__(256) knot_slope_kernel(const double* __restrict__ X, double* __restrict__ p,
                                                         const double* __restrict__ ypp, long n, int c_src,
                                                         int c_dst, double div) {
  long g = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * n) return;
  long i = g / 3;
  int k = (int) (g - 3 * i);
  long j = i >= n - 1 ? n - 2 : i;
  const double h = X[j + 1] - X[j];
  const double y0 = p[9 * j + c_src + k], y1 = p[9 * (j + 1) + c_src + k];
  const double d = (y1 - y0) / h - ypp[9 * j + c_src + k] * h / 3 - ypp[9 * (j + 1) + c_src + k] * h / 6;
  p[9 * i + c_dst + k] = d / div;
}

// Electric field in the time domain at one observer, one knot per lane
// (src/OSCARSSR.cc:4471-4512): E = C0 / (1 - N.B)^3 (near + far field) and
// the arrival time i dt + |R| / c.  pow(x, 3) is x * x * x here.
__global__ void __launch_bounds__(256) efield_time_kernel(const double* __restrict__ p, long n, double ox, double oy,
                                                          double oz, double c0, double dt,
                                                          double* __restrict__ out_t, double* __restrict__ out_e) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* k = p + 9 * i;
  const double Rx = ox - k[0], Ry = oy - k[1], Rz = oz - k[2];
  const double r2 = Rx * Rx + Ry * Ry + Rz * Rz;
  const double D = sqrt(r2);
  const double Nx = Rx / D, Ny = Ry / D, Nz = Rz / D;
  const double Bx = k[3], By = k[4], Bz = k[5];
  const double Ax = k[6], Ay = k[7], Az = k[8];
  const double q = 1.0 / (1.0 - (Nx * Bx + Ny * By + Nz * Bz));
  const double mult = c0 * (q * q * q);
  const double b2 = Bx * Bx + By * By + Bz * Bz;
  const double Mx = Nx - Bx, My = Ny - By, Mz = Nz - Bz;
  // (N - B) x A, then N x that
  const double cx = My * Az - Ay * Mz, cy = Mz * Ax - Az * Mx, cz = Mx * Ay - Ax * My;
  const double fx = Ny * cz - cy * Nz, fy = Nz * cx - cz * Nx, fz = Nx * cy - cx * Ny;
  const double ic = 1.0 / kC;
  const double nfx = ((1.0 - b2) * Mx) / r2, nfy = ((1.0 - b2) * My) / r2, nfz = ((1.0 - b2) * Mz) / r2;
  const double ffx = (ic * fx) / D, ffy = (ic * fy) / D, ffz = (ic * fz) / D;
  out_e[3 * i + 0] = mult * (nfx + ffx);
  out_e[3 * i + 1] = mult * (nfy + ffy);
  out_e[3 * i + 2] = mult * (nfz + ffz);
  out_t[i] = (double) i * dt + D / kC;
}

// ---------------------------------------------------------------- field values at points
// Sum of the fields at n points (x, y, z rows): one point per lane.

__global__ void __launch_bounds__(256) field_points_kernel(const FieldDev* __restrict__ F, int nf,
                                                           const double* __restrict__ xyz, long n, double t,
                                                           double* __restrict__ out) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double B[3];
  get_b(F, nf, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], t, B);
  out[3 * i + 0] = B[0];
  out[3 * i + 1] = B[1];
  out[3 * i + 2] = B[2];
}

// ---------------------------------------------------------------- level samples

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
                                    om::TrajDesc* desc, int np, double* t, double* p, hipStream_t s,
                                    int nefields) {
  if (nefields > 0) {
    // electric fields (after the magnetic ones in F): the general equations of motion
    om::UndParams u0;
    memset(&u0, 0, sizeof(u0));
    om::GridZ g0;
    memset(&g0, 0, sizeof(g0));
    int threads0 = 64, blocks0 = (np * om::rk4_lanes<om::kFieldsEB>() + threads0 - 1) / threads0;
    hipLaunchKernelGGL(om::rk4_kernel<om::kFieldsEB>, dim3(blocks0), dim3(threads0), 0, s, u0, g0, F, nfields, desc,
                       np, t, p, nefields);
    return hipGetLastError();
  }
  int threads = 64;
  // rk4_lanes lanes per particle: 4 (two per direction), 8 for kEpuZ
  int blocks = (np * 4 + threads - 1) / threads;   // every variant but kEpuZ
  int blocks8 = (np * 8 + threads - 1) / threads;  // kEpuZ
  static_assert(om::rk4_lanes<om::kEpuZ>() == 8 && om::rk4_lanes<om::kAxisUndY>() == 4 &&
                    om::rk4_lanes<om::kOneUndulator>() == 4 && om::rk4_lanes<om::kFieldsGeneric>() == 4,
                "lanes per particle");
  om::UndParams u;
  memset(&u, 0, sizeof(u));
  om::GridZ g;
  memset(&g, 0, sizeof(g));
  const om::FieldDev& f = *F0_host;
  // one z grid without rotation whose samples fit in LDS: the staged-grid variant
  const size_t grid_bytes = (size_t) 3 * f.n[2] * sizeof(double);
  // the specialised variants hold time-independent fields only
  if (nfields == 1 && !f.td && f.type == om::kGrid && f.axis == 2 && f.rot_sin[0] == 0 && f.rot_sin[1] == 0 &&
      f.rot_sin[2] == 0 && f.rot_cos[0] == 1 && f.rot_cos[1] == 1 && f.rot_cos[2] == 1 &&
      grid_bytes <= 160 * 1024) {
    g.data = f.data;
    g.n = f.n[2];
    g.start = f.start[2];
    g.stop = f.stop[2];
    g.step = f.step[2];
    g.inv_step = f.inv_step[2];
    g.trans = f.trans[2];
    hipError_t e = hipFuncSetAttribute((const void*) om::rk4_kernel<om::kGridZ>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int) grid_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(om::rk4_kernel<om::kGridZ>, dim3(blocks), dim3(threads), grid_bytes, s, u, g, F, nfields, desc,
                       np, t, p, 0);
    return hipGetLastError();
  }
  if (nfields == 1 && !f.td && (f.type == om::kUndulator || f.type == om::kEPU)) {
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
    for (int k = 0; k < 2; ++k) {
      double ps = k == 0 ? u.psA : u.psB;
      double* th = k == 0 ? u.thA : u.thB;
      th[0] = -u.half_len + ps;
      th[1] = u.half_len + ps;
      th[2] = -u.half_len + ps + u.PL;
      th[3] = u.half_len + ps - u.PL;
      th[4] = -u.half_len + ps + u.half_PL;
      th[5] = u.half_len + ps - u.half_PL;
    }
    // planar undulator along z, untapered, field along one axis: the reduced equations
    int nz = (f.field[0] != 0) + (f.field[1] != 0) + (f.field[2] != 0);
    bool axis_z = f.unit[0] == 0 && f.unit[1] == 0 && f.unit[2] == 1;
    if (f.type == om::kUndulator && axis_z && f.taper == 0 && nz == 1) {
      if (f.field[0] != 0)
        hipLaunchKernelGGL(om::rk4_kernel<om::kAxisUndX>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc, np,
                           t, p, 0);
      else if (f.field[1] != 0)
        hipLaunchKernelGGL(om::rk4_kernel<om::kAxisUndY>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc, np,
                           t, p, 0);
      else
        hipLaunchKernelGGL(om::rk4_kernel<om::kAxisUndZ>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc, np,
                           t, p, 0);
    } else if (axis_z && f.taper == 0 && f.type == om::kEPU) {
      hipLaunchKernelGGL(om::rk4_kernel<om::kEpuZ>, dim3(blocks8), dim3(threads), 0, s, u, g, F, nfields, desc, np, t,
                         p, 0);
    } else if (axis_z && f.taper == 0) {
      hipLaunchKernelGGL(om::rk4_kernel<om::kUndZ>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc, np, t,
                         p, 0);
    } else {
      hipLaunchKernelGGL(om::rk4_kernel<om::kOneUndulator>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc,
                         np, t, p, 0);
    }
  } else {
    hipLaunchKernelGGL(om::rk4_kernel<om::kFieldsGeneric>, dim3(blocks), dim3(threads), 0, s, u, g, F, nfields, desc, np,
                       t, p, 0);
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_knot_finish(double* p, long total, hipStream_t s) {
  long blocks = (total * 6 + 255) / 256;
  hipLaunchKernelGGL(om::knot_finish_kernel, dim3((unsigned) blocks), dim3(256), 0, s, p, total);
  return hipGetLastError();
}

// scratch: sig, sm1, pv, ppv (one double per knot), u, u2 (nine per knot)
extern "C" hipError_t om_launch_spline(const om::TrajDesc* desc, int np, const double* t, const double* p,
                                       double* ypp, double* u, double* sig, double* sm1, double* pv, double* ppv,
                                       double* u2, long total, hipStream_t s) {
  unsigned b1 = (unsigned) ((total + 255) / 256), b9 = (unsigned) ((9 * total + 255) / 256);
  hipLaunchKernelGGL(om::spline_pre_kernel, dim3(b1), dim3(256), 0, s, desc, np, t, p, sig, sm1, u, total);
  hipLaunchKernelGGL(om::spline_p_kernel, dim3(b1), dim3(256), 0, s, desc, np, sig, sm1, pv, ppv, total);
  hipLaunchKernelGGL(om::spline_u_kernel, dim3(b9), dim3(256), 0, s, desc, np, t, p, sig, ppv, u, u2, total);
  hipLaunchKernelGGL(om::spline_b_kernel, dim3(b9), dim3(256), 0, s, desc, np, t, p, pv, u2, ypp, total);
  return hipGetLastError();
}

// Spline of a user trajectory (one particle's knots), with the optional
// construction of beta (construct & 1) and a/c (construct & 2) first.
extern "C" hipError_t om_launch_spline_user(const double* t, double* p, double* ypp, double* u, long n,
                                            int construct, hipStream_t s) {
  unsigned b3 = (unsigned) ((3 * n + 255) / 256);
  if (construct & 1) {
    hipLaunchKernelGGL(om::spline_seq_kernel, dim3(1), dim3(64), 0, s, t, p, ypp, u, n, 0, 3);
    hipLaunchKernelGGL(om::knot_slope_kernel, dim3(b3), dim3(256), 0, s, t, p, ypp, n, 0, 3, om::kC);
  }
  if (construct & 2) {
    hipLaunchKernelGGL(om::spline_seq_kernel, dim3(1), dim3(64), 0, s, t, p, ypp, u, n, 3, 3);
    hipLaunchKernelGGL(om::knot_slope_kernel, dim3(b3), dim3(256), 0, s, t, p, ypp, n, 3, 6, 1.0);
  }
  hipLaunchKernelGGL(om::spline_seq_kernel, dim3(1), dim3(64), 0, s, t, p, ypp, u, n, 0, 9);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_efield_time(const double* p, long n, const double* obs, double c0, double dt,
                                            double* out_t, double* out_e, hipStream_t s) {
  unsigned blocks = (unsigned) ((n + 255) / 256);
  hipLaunchKernelGGL(om::efield_time_kernel, dim3(blocks), dim3(256), 0, s, p, n, obs[0], obs[1], obs[2], c0, dt,
                     out_t, out_e);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_field_points(const om::FieldDev* F, int nf, const double* xyz, long n, double t,
                                             double* out, hipStream_t s) {
  unsigned blocks = (unsigned) ((n + 255) / 256);
  hipLaunchKernelGGL(om::field_points_kernel, dim3(blocks), dim3(256), 0, s, F, nf, xyz, n, t, out);
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
