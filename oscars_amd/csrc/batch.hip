// Batched multi-particle Lienard-Wiechert sums in one persistent launch.
//
// Every particle of the resident trajectory set contributes weight * flux to
// one request (an observation grid x an energy list); requests may differ in
// grid size and energy count, and particles in trajectory length.  The work
// unit is (particle, observation point, 64-energy tile); a persistent grid of
// single-wave workgroups pulls units from a device counter (a vector atomic on
// global memory), runs the reference's level loop for its 64 energy lanes and
// adds the weighted flux into the request's output with FP64 atomics
// (an incoherent sum over particles; the summation order, and so the last
// bits, depend on scheduling).
//
// Reference behaviour restated (paths relative to the reference root):
//   multi-particle loop with weight 1/N   src/OSCARSSR.cc:4016-4063, 1778-1818
//   level loop and convergence           src/OSCARSSR.cc:4179-4245
//   level sample times                   src/TParticleTrajectoryInterpolated.cc:124-161, 264-270
//
// Level samples of all particles live in one block: particle p, level L starts
// at row p*stride + 2^L - 1 (stride = 2^(Lcap+1) - 1 rows of 10 doubles).
// Deeper levels (rare: slowly converging points) are interpolated from the
// spline inside the sum kernel with the same expressions, so a batch follows
// the same level rule as the single-particle path up to max_level.
#include "lw_common.h"
#include "rk4_common.h"

namespace om {

struct LevelJobBatch {
  const double* t;
  const double* p;
  const double* ypp;
  const TrajDesc* desc;
  const double* tspan;  // per particle: tstart, tstop (2 doubles)
  double* rows;
  long stride;          // rows per particle
  int np;
  int Lcap;
  int stop_mem;         // levels above it use the extended spacing
};

__global__ void __launch_bounds__(256) level_batch_kernel(LevelJobBatch J) {
  long g = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long) J.np * J.stride) return;
  int ip = (int) (g / J.stride);
  long s = g - (long) ip * J.stride;
  int L = 63 - __clzll((unsigned long long) (s + 1));  // s in [2^L - 1, 2^(L+1) - 1)
  long i = s - ((1L << L) - 1);
  double ts = J.tspan[2 * ip], te = J.tspan[2 * ip + 1];
  // src/TParticleTrajectoryInterpolated.cc:138-147, 264-270 and
  // src/TParticleTrajectoryInterpolatedPoints.cc:33-37 (same operations as the host)
  double p2L = (double) (1L << L);
  double spacing = L > J.stop_mem ? (te - ts) / (p2L - 1) : (te - ts) / p2L;
  double tt = ts + (te - ts) / (2.0 * p2L) + spacing * (double) i;
  TrajDesc D = J.desc[ip];
  double v[9];
  spline_value(J.t + D.knot_offset, J.p + 9 * D.knot_offset, J.ypp + 9 * D.knot_offset, D.n, tt, v);
  double* o = J.rows + 10 * g;
  o[0] = tt;
#pragma unroll
  for (int c = 0; c < 9; ++c) o[1 + c] = v[c];
}

struct ReqDev {
  const double* obs;    // 3 per point
  const double* omega;  // per energy
  double* out;          // n_obs * n_e
  long n_obs, n_e, n_tiles_e;
};

struct BatchArgs {
  const double* rows;
  long stride;
  int Lb;               // highest level held in rows (materialised by level_batch_kernel)
  int Lmax;             // highest level summed; levels above Lb are interpolated in the kernel
  int stop_mem;         // levels above it use the extended spacing
  const double* t;      // knots, spline second derivatives and descriptors (levels above Lb)
  const double* p;
  const double* ypp;
  const TrajDesc* desc;
  const double* tspan;
  const int* req_of;
  const double* weight;
  const long* item_prefix;  // np + 1
  int np;
  const ReqDev* req;
  double C0, C2;
  double pol[6];
  int use_pol;
  double precision;
  unsigned long long* counter;
  long total_items;
  unsigned long long* n_unconverged;
  unsigned long long* evals;  // sum over lanes of 2^(L+1)-1 (as integer)
  // mixed path: units with lanes whose stop decision lay within the FP32 error
  // bound, as (item, lane mask) pairs (room for every unit), recomputed in FP64
  // by lw_batch_fix_kernel
  unsigned long long* fix_list;
  unsigned long long* n_fix;  // units listed
};

// Amplitude of a sample interpolated from the spline at time tt (levels above
// the stored rows; rare).
__device__ __forceinline__ Amp interpolated_amplitude(const BatchArgs& A, const TrajDesc& D, V3 obs,
                                                                double tt) {
  double srow[10];
  srow[0] = tt;
  spline_value(A.t + D.knot_offset, A.p + 9 * D.knot_offset, A.ypp + 9 * D.knot_offset, D.n, tt, srow + 1);
  return amplitude(obs, srow);
}

// Lane 0's value as a wave-uniform scalar (two 32-bit readfirstlane moves).
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  unsigned lo = (unsigned) __builtin_amdgcn_readfirstlane((int) (unsigned) v);
  unsigned hi = (unsigned) __builtin_amdgcn_readfirstlane((int) (unsigned) (v >> 32));
  return ((unsigned long long) hi << 32) | lo;
}

// One work unit (particle, observation point, 64-energy tile): the reference's
// level loop for the lanes of lane_mask whose energy exists, weighted flux
// added into the request's output.  Lane j computes the amplitude of sample
// base+j and leaves it in LDS; then every lane walks those 64 samples at its
// own energy.  MIXED: lanes whose stop decision lay within the FP32 error
// bound add nothing; the unit is listed with their mask for lw_batch_fix_kernel.
template <bool MIXED>
__device__ __forceinline__ void batch_unit(const BatchArgs& A, long item, unsigned long long lane_mask, int lane,
                                           double* s_tau, double (*s_v)[64]) {
  // particle of this work unit: last p with item_prefix[p] <= item
  int lo = 0, hi = A.np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (A.item_prefix[mid] <= item) lo = mid; else hi = mid;
  }
  int ip = lo;
  long local = item - A.item_prefix[ip];
  const ReqDev R = A.req[A.req_of[ip]];
  long io = local / R.n_tiles_e;
  long et = local - io * R.n_tiles_e;
  long ie = et * 64 + lane;
  bool valid = ie < R.n_e && ((lane_mask >> lane) & 1ULL);
  long iec = ie < R.n_e ? ie : R.n_e - 1;
  V3 obs = v3(R.obs[3 * io], R.obs[3 * io + 1], R.obs[3 * io + 2]);
  double negw = -R.omega[iec];
  double T = A.tspan[2 * ip + 1] - A.tspan[2 * ip];
  const double* prow = A.rows + 10 * (long) ip * A.stride;
  const double ts = A.tspan[2 * ip], te = A.tspan[2 * ip + 1];
  const TrajDesc D = A.desc[ip];
  double sum[6] = {0, 0, 0, 0, 0, 0};
  double lastmag = -1, prec = -1;
  bool active = valid;
  bool conv = false;
  int lev = 0;
  double an_run = 0, dm_last = 0;
  bool border = false;
  for (int L = 0; L <= A.Lmax; ++L) {
    if (!__any(active)) break;
    const double* __restrict__ rows = prow + 10 * ((1L << L) - 1);
    long ns = 1L << L;
    // sample times of level L (the expressions of level_batch_kernel)
    double p2L = (double) (1L << L);
    double spacing = L > A.stop_mem ? (te - ts) / (p2L - 1) : (te - ts) / p2L;
    double tfirst = ts + (te - ts) / (2.0 * p2L);
    bool stored = L <= A.Lb;
    double ps[6] = {0, 0, 0, 0, 0, 0};
    double maxd = 0, lastphase = 0;
    double anl = 0;  // mixed path: sum of |V|_1 over the samples this lane computed
    for (long base = 0; base < ns; base += 64) {
      long si = base + lane;
      if (si < ns) {
        Amp a;
        if (stored) {
          a = amplitude(obs, rows + 10 * si);
        } else {
          a = interpolated_amplitude(A, D, obs, tfirst + spacing * (double) si);
        }
        s_tau[lane] = a.tau;
        s_v[0][lane] = a.V.x;
        s_v[1][lane] = a.V.y;
        s_v[2][lane] = a.V.z;
        if (MIXED) anl += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);
      }
      wave_sync();
      int cnt = (int) ((ns - base) < 64 ? (ns - base) : 64);
      if (base == 0) lastphase = negw * s_tau[0];
      int k = 0;
      for (; k + 4 <= cnt; k += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          double phi = negw * s_tau[k + u];
          maxd = fmax(maxd, fabs(phi - lastphase));
          lastphase = phi;
          double sn, cs;
          sincos_mode(MIXED, phi, sn, cs);
          double vx = s_v[0][k + u], vy = s_v[1][k + u], vz = s_v[2][k + u];
          ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
          ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
          ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
        }
      }
      for (; k < cnt; ++k) {
        double phi = negw * s_tau[k];
        maxd = fmax(maxd, fabs(phi - lastphase));
        lastphase = phi;
        double sn, cs;
        sincos_mode(MIXED, phi, sn, cs);
        double vx = s_v[0][k], vy = s_v[1][k], vz = s_v[2][k];
        ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
        ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
        ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
      }
      wave_sync();
    }
    if (MIXED) {
      // every lane walked the same samples: the level's |V|_1 sum is wave-uniform
      for (int off = 32; off > 0; off >>= 1) anl += __shfl_xor(anl, off);
    }
    if (active) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sum[k] += ps[k];
      lev = L;
      // convergence test, src/OSCARSSR.cc:4227-4239
      double E[6];
      double dt = T / (2.0 * (double) (1L << L) - 1);
#pragma unroll
      for (int k = 0; k < 6; ++k) E[k] = sum[k] * dt;
      if (A.use_pol) pol_project(E, A.pol);
      double mag = cmag2(E);
      double mag_prev = lastmag;
      prec = fabs(mag - lastmag) / lastmag;
      bool c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == lastmag);
      lastmag = mag;
      if (MIXED) {
        an_run += anl;
        border = border || level_borderline(L, dt, A.precision, mag_prev, mag, prec, maxd, an_run, dm_last);
      }
      if (c) {
        conv = true;
        active = false;
      }
    }
  }
  // mixed path: the unit is listed once, with the mask of its borderline lanes
  // (all lanes take part in the atomic; lanes 0 and 1 store the entry)
  bool fixed = false;
  if (MIXED) {
    bool want = valid && border;
    unsigned long long m = __ballot(want);
    unsigned long long slot = atomicAdd(A.n_fix, (lane == 0 && m != 0ULL) ? 1ULL : 0ULL);
    unsigned long long j = uniform_u64(slot);
    if (m != 0ULL && lane < 2) A.fix_list[2 * j + lane] = lane == 0 ? (unsigned long long) item : m;
    fixed = want;
  }
  if (valid && !fixed) {
    double dt = T / (2.0 * (double) (1L << lev) - 1);
    double f = A.C0 * dt;
    double E[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) E[k] = sum[k] * f;
    if (A.use_pol) pol_project(E, A.pol);
    double flux = A.C2 * cmag2(E) * A.weight[ip];
    atomicAdd(&R.out[io * R.n_e + ie], flux);
  }
  // bookkeeping: lanes that stopped without converging, and evaluations done
  // (all lanes take part; only lane 0 adds a nonzero amount)
  bool counted = valid && !fixed;
  unsigned long long nc = __ballot(counted && !conv);
  unsigned long long ev = counted ? (unsigned long long) ((2L << lev) - 1) : 0ULL;
  for (int off = 32; off > 0; off >>= 1) ev += __shfl_down(ev, off);
  atomicAdd(A.n_unconverged, lane == 0 ? (unsigned long long) __popcll(nc) : 0ULL);
  atomicAdd(A.evals, lane == 0 ? ev : 0ULL);
}

// Four waves per SIMD: the register budget is set for the hot stored-row path;
// the rare interpolation path may spill.
template <bool MIXED>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) lw_batch_kernel(BatchArgs A) {
  __shared__ double s_tau[64];
  __shared__ double s_v[3][64];
  int lane = threadIdx.x;
  while (true) {
    // Next work unit.  Every lane executes the atomic (lane 0 adds 1, the others
    // 0) and the index is taken from lane 0 with readfirstlane, so it is a
    // scalar and the loop exit is a uniform branch.  No lane-conditional block
    // surrounds a cross-lane operation, which keeps the compiler from splitting
    // the wave between two copies of the loop.
    unsigned long long mine = atomicAdd(A.counter, lane == 0 ? 1ULL : 0ULL);
    long item = (long) uniform_u64(mine);
    if (item >= A.total_items) break;
    batch_unit<MIXED>(A, item, ~0ULL, lane, s_tau, s_v);
  }
}

// FP64 recomputation of the lanes the mixed batch kernel listed: for each
// listed unit, each lane of its mask (one energy) is summed by a whole wave,
// the 64 lanes splitting each level's samples (a few points per unit: a wave
// per point costs about 1/64 of re-walking the unit).  Levels, convergence
// test and weighting are the FP64 kernel's.  A grid-stride loop over the
// listed count (read from device memory, so the launch can sit in a captured
// graph); each lane adds the flux of its own energy at the end of the unit.
__global__ void __launch_bounds__(64) lw_batch_fix_kernel(BatchArgs A) {
  int lane = threadIdx.x;
  long n = (long) uniform_u64(*A.n_fix);
  for (long e = blockIdx.x; e < n; e += gridDim.x) {
    long item = (long) uniform_u64(A.fix_list[2 * e]);
    unsigned long long mask = uniform_u64(A.fix_list[2 * e + 1]);
    int lo = 0, hi = A.np;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (A.item_prefix[mid] <= item) lo = mid; else hi = mid;
    }
    int ip = lo;
    long local = item - A.item_prefix[ip];
    const ReqDev R = A.req[A.req_of[ip]];
    long io = local / R.n_tiles_e;
    long et = local - io * R.n_tiles_e;
    V3 obs = v3(R.obs[3 * io], R.obs[3 * io + 1], R.obs[3 * io + 2]);
    const double ts = A.tspan[2 * ip], te = A.tspan[2 * ip + 1];
    const double T = te - ts;
    const double* prow = A.rows + 10 * (long) ip * A.stride;
    const TrajDesc D = A.desc[ip];
    double my_flux = 0;
    unsigned long long n_conv = 0, ev = 0;
    for (unsigned long long m = mask; m != 0ULL; m &= m - 1ULL) {
      int b = __builtin_ctzll(m);
      long ie = et * 64 + b;
      double negw = -R.omega[ie];
      double sum[6] = {0, 0, 0, 0, 0, 0};
      double lastmag = -1, prec = -1;
      int lev = 0;
      bool c = false;
      for (int L = 0; L <= A.Lmax && !c; ++L) {
        long ns = 1L << L;
        double p2L = (double) (1L << L);
        double spacing = L > A.stop_mem ? (te - ts) / (p2L - 1) : (te - ts) / p2L;
        double tfirst = ts + (te - ts) / (2.0 * p2L);
        const double* __restrict__ rows = prow + 10 * ((1L << L) - 1);
        double tot[6], maxd;
        if (L <= A.Lb) {
          wave_level_sum(ns, negw, lane, [&](long i) { return amplitude(obs, rows + 10 * i); }, tot, maxd);
        } else {
          wave_level_sum(ns, negw, lane,
                         [&](long i) { return interpolated_amplitude(A, D, obs, tfirst + spacing * (double) i); },
                         tot, maxd);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) sum[k] += tot[k];
        lev = L;
        // convergence test, src/OSCARSSR.cc:4227-4239
        double E[6];
        double dt = T / (2.0 * (double) (1L << L) - 1);
#pragma unroll
        for (int k = 0; k < 6; ++k) E[k] = sum[k] * dt;
        if (A.use_pol) pol_project(E, A.pol);
        double mag = cmag2(E);
        prec = fabs(mag - lastmag) / lastmag;
        c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == lastmag);
        lastmag = mag;
      }
      double dt = T / (2.0 * (double) (1L << lev) - 1);
      double f = A.C0 * dt;
      double E[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) E[k] = sum[k] * f;
      if (A.use_pol) pol_project(E, A.pol);
      double flux = A.C2 * cmag2(E) * A.weight[ip];
      my_flux = lane == b ? flux : my_flux;
      n_conv += c ? 1ULL : 0ULL;
      ev += (unsigned long long) ((2L << lev) - 1);
    }
    if ((mask >> lane) & 1ULL) atomicAdd(&R.out[io * R.n_e + et * 64 + lane], my_flux);
    unsigned long long nu = (unsigned long long) __popcll(mask) - n_conv;
    atomicAdd(A.n_unconverged, lane == 0 ? nu : 0ULL);
    atomicAdd(A.evals, lane == 0 ? ev : 0ULL);
  }
}

// Clears n doubles (outputs) and the counters without a memset node, so the
// sequence can be captured in a HIP graph and replayed.
__global__ void __launch_bounds__(256) batch_clear_kernel(double* __restrict__ out, long n,
                                                          unsigned long long* __restrict__ counters, int nc) {
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
  if (i < nc) counters[i] = 0ULL;
}

}  // namespace om

extern "C" hipError_t om_launch_level_batch(const om::LevelJobBatch* J, hipStream_t s) {
  long total = (long) J->np * J->stride;
  long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(om::level_batch_kernel, dim3((unsigned) blocks), dim3(256), 0, s, *J);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_batch_clear(double* out, long n, unsigned long long* counters, int nc,
                                            hipStream_t s) {
  long m = n > nc ? n : nc;
  long blocks = (m + 255) / 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(om::batch_clear_kernel, dim3((unsigned) blocks), dim3(256), 0, s, out, n, counters, nc);
  return hipGetLastError();
}

extern "C" hipError_t om_launch_lw_batch(const om::BatchArgs* A, int mixed, int n_groups, hipStream_t s) {
  if (mixed) {
    hipLaunchKernelGGL(om::lw_batch_kernel<true>, dim3(n_groups), dim3(64), 0, s, *A);
    hipLaunchKernelGGL(om::lw_batch_fix_kernel, dim3(n_groups), dim3(64), 0, s, *A);
  } else {
    hipLaunchKernelGGL(om::lw_batch_kernel<false>, dim3(n_groups), dim3(64), 0, s, *A);
  }
  return hipGetLastError();
}
