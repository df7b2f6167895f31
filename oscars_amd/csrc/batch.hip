// Batched multi-particle Lienard-Wiechert sums in one persistent launch.
//
// Every particle of the resident trajectory set contributes weight * flux to
// one request (an observation grid x an energy list); requests may differ in
// grid size and energy count, and particles in trajectory length.  The work
// unit is (particle, observation point, block of up to kBatchEB energies); a
// persistent grid of single-wave workgroups pulls units from a device counter
// (a vector atomic on global memory), runs the reference's level loop for the
// block's energies, re-packing the still-active ones into full waves at every
// level, and adds the weighted flux into the request's output with FP64
// atomics (an incoherent sum over particles; the summation order, and so the
// last bits, depend on scheduling).
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

// LDS of one persistent wave: the amplitudes of the current super-block of
// kSuper samples (one double of padding after every 16 samples, so replicas
// that walk different parts of it read different banks) and the active list.
constexpr int kSuper = 256;
constexpr int kBlkPad = kSuper + kSuper / 16;
__device__ __forceinline__ int blk(int k) { return k + (k >> 4); }

struct BatchLds {
  double tau[kBlkPad];
  double vx[kBlkPad];
  double vy[kBlkPad];
  double vz[kBlkPad];
  unsigned short list[kBatchEB];
};

// One work unit (particle ip, observation point io, energies e0 .. e0+n_eb-1).
struct Unit {
  long item;
  int ip;
  long io, e0;
  int n_eb;
  V3 obs;
  double ts, te, T;
  const double* prow;
  TrajDesc D;
  ReqDev R;
};

__device__ __forceinline__ Unit unit_of(const BatchArgs& A, long item) {
  Unit U;
  U.item = item;
  // particle of this work unit: last p with item_prefix[p] <= item
  int lo = 0, hi = A.np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (A.item_prefix[mid] <= item) lo = mid; else hi = mid;
  }
  U.ip = lo;
  const long local = item - A.item_prefix[U.ip];
  U.R = A.req[A.req_of[U.ip]];
  U.io = local / U.R.n_blocks_e;
  U.e0 = (local - U.io * U.R.n_blocks_e) * kBatchEB;
  U.n_eb = (int) (U.R.n_e - U.e0 < kBatchEB ? U.R.n_e - U.e0 : kBatchEB);
  U.obs = v3(U.R.obs[3 * U.io], U.R.obs[3 * U.io + 1], U.R.obs[3 * U.io + 2]);
  U.ts = A.tspan[2 * U.ip];
  U.te = A.tspan[2 * U.ip + 1];
  U.T = U.te - U.ts;
  U.prow = A.rows + 10 * (long) U.ip * A.stride;
  U.D = A.desc[U.ip];
  return U;
}

// Flags of an energy in the wave's scratch.
constexpr int kFConverged = 1;
constexpr int kFBorderline = 2;  // mixed pass: a stop decision lay within the sin/cos error bound
constexpr int kFHanded = 4;      // handed to lw_batch_deep_kernel
constexpr int kFRecheck = 8;     // recomputed in FP64 inside the unit
constexpr int kFIdle = 16;       // not part of the FP64 recheck pass

// The reference's level loop for the energies of a unit whose flags carry none
// of the bits in `skip`.  SC32: reduced-accuracy sin/cos (sincos_fast64) with
// the borderline test, else full-accuracy sin/cos.  Level by level, the still
// active energies are packed into groups: 64 energies, one per lane, and for
// the last r < 64 of a level, m = 64 / r replicas of the r energies, replica q
// walking the q-th contiguous part of every super-block.  The level's samples
// go in super-blocks of kSuper: the lanes compute their amplitudes into LDS
// once, then every group walks the super-block at its energies (a group's
// partial sums wait in the wave's scratch while the others walk), so a
// sample's amplitude is computed once per level whatever the number of
// groups.  The replicas' partial sums are added in a fixed tree order.  The
// largest phase step takes the sample before each part from LDS (or from the
// previous super-block), so it is the same maximum over the same consecutive
// pairs as a sequential walk.  HANDOVER: energies still active at level Ldeep
// are handed to lw_batch_deep_kernel.
template <bool SC32, bool HANDOVER>
__device__ __forceinline__ void unit_levels(const BatchArgs& A, const Unit& U, int lane, double* __restrict__ ws,
                                            int* __restrict__ wsi, BatchLds& S, int skip, unsigned long long& n_amp,
                                            unsigned long long& n_ev, unsigned long long& n_slot) {
  double* w_sum = ws;                    // 6 x kBatchEB
  double* w_lastmag = ws + 6 * kBatchEB;
  double* w_prec = ws + 7 * kBatchEB;
  double* w_an = ws + 8 * kBatchEB;      // mixed: sum of |V|_1 over the samples so far
  double* w_dm = ws + 9 * kBatchEB;      // mixed: error bound of lastmag
  double* w_grp = ws + kBatchWsD * kBatchEB;  // partial sums of the groups between super-blocks
  int* w_level = wsi;
  int* w_flags = wsi + kBatchEB;
  const int n_eb = U.n_eb;
  for (int L = 0; L <= A.Lmax; ++L) {
    if (HANDOVER && L == A.Ldeep) {
      // hand the energies still active to the one-wave-per-point kernel
      for (int t = 0; t < n_eb; t += 64) {
        int e = t + lane;
        int ec = e < n_eb ? e : 0;
        bool a = e < n_eb && (w_flags[ec] & skip) == 0;
        unsigned long long m = __ballot(a);
        unsigned long long first = atomicAdd(A.n_deep, lane == 0 ? (unsigned long long) __popcll(m) : 0ULL);
        unsigned long long j = uniform_u64(first) + (unsigned long long) __popcll(m & ((1ULL << lane) - 1ULL));
        if (a && j < (unsigned long long) A.deep_cap) {
          A.deep_code[j] = (unsigned long long) U.item * kBatchEB + (unsigned long long) e;
#pragma unroll
          for (int k = 0; k < kBatchWsD; ++k) A.deep_state[k * A.deep_cap + j] = ws[k * kBatchEB + e];
          A.deep_level[j] = w_level[e];
          A.deep_flags[j] = w_flags[e];
          w_flags[e] |= kFHanded;
        }
      }
      __syncthreads();
    }
    // pack the energies still active
    int n_act = 0;
    for (int t = 0; t < n_eb; t += 64) {
      int e = t + lane;
      bool a = e < n_eb && (w_flags[e < n_eb ? e : 0] & skip) == 0;
      unsigned long long m = __ballot(a);
      if (a) S.list[n_act + __popcll(m & ((1ULL << lane) - 1ULL))] = e;
      n_act += __popcll(m);
    }
    __syncthreads();
    if (n_act == 0) break;
    const double* __restrict__ rows = U.prow + 10 * ((1L << L) - 1);
    const long ns = 1L << L;
    // sample times of level L (the expressions of level_batch_kernel)
    const double p2L = (double) (1L << L);
    const double spacing = L > A.stop_mem ? (U.te - U.ts) / (p2L - 1) : (U.te - U.ts) / p2L;
    const double tfirst = U.ts + (U.te - U.ts) / (2.0 * p2L);
    const bool stored = L <= A.Lb;
    const double dt = U.T / (2.0 * (double) (1L << L) - 1);
    const int n_groups = (n_act + 63) >> 6;  // 64 energies each, the last one r <= 64
    const long n_sb = (ns + kSuper - 1) / kSuper;
    n_amp += (unsigned long long) ns;
    double anl = 0;   // mixed: sum of |V|_1 over the samples this lane computed
    double tprev = 0; // tau of the last sample of the previous super-block
    double ps[6] = {0, 0, 0, 0, 0, 0};
    double maxd = 0;
    for (long sb = 0; sb < n_sb; ++sb) {
      const long base = sb * kSuper;
      const int cnt = (int) ((ns - base) < kSuper ? (ns - base) : kSuper);
#pragma unroll 1
      for (int h = 0; h < kSuper; h += 64) {
        const long si = base + h + lane;
        if (si < ns) {
          Amp a;
          if (stored) {
            a = amplitude(U.obs, rows + 10 * si);
          } else {
            a = interpolated_amplitude(A, U.D, U.obs, tfirst + spacing * (double) si);
          }
          S.tau[blk(h + lane)] = a.tau;
          S.vx[blk(h + lane)] = a.V.x;
          S.vy[blk(h + lane)] = a.V.y;
          S.vz[blk(h + lane)] = a.V.z;
          if (SC32) anl += fabs(a.V.x) + fabs(a.V.y) + fabs(a.V.z);
        }
      }
      wave_sync();
      const bool last_sb = sb == n_sb - 1;
      if (SC32 && last_sb) {
        // every sample of the level was computed by one lane: the |V|_1 sum is wave-uniform
        for (int off = 32; off > 0; off >>= 1) anl += __shfl_xor(anl, off);
      }
      for (int grp = 0; grp < n_groups; ++grp) {
        const int g = grp << 6;
        const int W = n_act - g < 64 ? n_act - g : 64;  // energies in this group
        const int m = 64 / W;                            // replicas of each
        const int q = lane / W;                          // this lane's replica (q >= m: idle)
        const int j = lane - q * W;
        const bool act = q < m;
        const int e = S.list[g + (act ? j : 0)];  // idle lanes shadow the group's first energy
        const double negw = -U.R.omega[U.e0 + e];
        if (sb == 0) {
          n_ev += (unsigned long long) ns * (unsigned long long) W;
          const long full = ns >> 6, part = ns & 63;
          n_slot += 64ULL * ((unsigned long long) full * (unsigned long long) ((64 + m - 1) / m) +
                             (unsigned long long) ((part + m - 1) / m));
        }
        if (n_groups > 1) {
          // this group's partial sums from the previous super-block (lane-private)
          double* gp = w_grp + grp * 7 * 64 + lane;
          if (sb == 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c) ps[c] = 0;
            maxd = 0;
          } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) ps[c] = gp[c * 64];
            maxd = gp[6 * 64];
          }
        }
        double lastphase;  // phase of the sample before the part being walked
        auto walk = [&](int k0, int k1) {
          int k = k0;
          for (; k + 4 <= k1; k += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int b = blk(k + u);
              double phi = negw * S.tau[b];
              maxd = fmax(maxd, fabs(phi - lastphase));
              lastphase = phi;
              double sn, cs;
              sincos_mode(SC32, phi, sn, cs);
              double vx = S.vx[b], vy = S.vy[b], vz = S.vz[b];
              ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
              ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
              ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
            }
          }
          for (; k < k1; ++k) {
            const int b = blk(k);
            double phi = negw * S.tau[b];
            maxd = fmax(maxd, fabs(phi - lastphase));
            lastphase = phi;
            double sn, cs;
            sincos_mode(SC32, phi, sn, cs);
            double vx = S.vx[b], vy = S.vy[b], vz = S.vz[b];
            ps[0] = __builtin_fma(vx, cs, ps[0]); ps[1] = __builtin_fma(vx, sn, ps[1]);
            ps[2] = __builtin_fma(vy, cs, ps[2]); ps[3] = __builtin_fma(vy, sn, ps[3]);
            ps[4] = __builtin_fma(vz, cs, ps[4]); ps[5] = __builtin_fma(vz, sn, ps[5]);
          }
        };
        if (m == 1) {  // one energy per lane: the whole super-block, wave-uniform bounds
          lastphase = negw * (base > 0 ? tprev : S.tau[0]);
          walk(0, cnt);
        } else {
          const int k0 = act ? (q * cnt) / m : 0;
          const int k1 = act ? ((q + 1) * cnt) / m : 0;
          lastphase = negw * (k0 > 0 ? S.tau[blk(k0 - 1)] : (base > 0 ? tprev : S.tau[0]));
          walk(k0, k1);
        }
        if (!last_sb) {
          if (n_groups > 1) {
            double* gp = w_grp + grp * 7 * 64 + lane;
#pragma unroll
            for (int c = 0; c < 6; ++c) gp[c * 64] = ps[c];
            gp[6 * 64] = maxd;
          }
          continue;
        }
        // the level is complete for this group
        if (m > 1) {
          // replicas q..q+2s-1 into replica q, for s = 1, 2, 4, ...: replica 0
          // ends with the totals, in the same order for every group shape
          for (int st = 1; st < m; st <<= 1) {
            const int src = lane + st * W < 64 ? lane + st * W : lane;
            const bool take = act && (q % (2 * st)) == 0 && q + st < m;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
              double o = __shfl(ps[c], src);
              ps[c] = take ? ps[c] + o : ps[c];
            }
            double om = __shfl(maxd, src);
            maxd = take ? fmax(maxd, om) : maxd;
          }
        }
        if (q == 0) {
          double sum[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) sum[c] = w_sum[c * kBatchEB + e] + ps[c];
          // convergence test, src/OSCARSSR.cc:4227-4239
          double E[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) E[c] = sum[c] * dt;
          if (A.use_pol) pol_project(E, A.pol);
          const double mag = cmag2(E);
          const double mag_prev = w_lastmag[e];
          const double prec = fabs(mag - mag_prev) / mag_prev;
          const bool c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == mag_prev);
          int flags = w_flags[e] | (c ? kFConverged : 0);
          if (SC32) {
            double an_run = w_an[e] + anl;
            double dm_last = w_dm[e];
            if (level_borderline(L, dt, A.precision, mag_prev, mag, prec, maxd, an_run, dm_last))
              flags |= kFBorderline;
            w_an[e] = an_run;
            w_dm[e] = dm_last;
          }
#pragma unroll
          for (int c = 0; c < 6; ++c) w_sum[c * kBatchEB + e] = sum[c];
          w_lastmag[e] = mag;
          w_prec[e] = prec;
          w_level[e] = L;
          w_flags[e] = flags;
        }
      }
      tprev = S.tau[blk(cnt - 1)];
      wave_sync();  // the next super-block overwrites the amplitudes
    }
    __syncthreads();  // this level's state is read by other lanes at the next one
  }
}

// One work unit: the level loop for the block's energies, weighted flux added
// into the request's output with FP64 atomics.  MIXED: a first pass with the
// reduced-accuracy sin/cos; the energies whose stop decision lay within its error bound are
// then recomputed from level 0 with FP64 sin/cos in a second pass of the same
// unit (their amplitudes shared, the rows still in cache), so every energy
// stops where the FP64 path stops.  The per-energy state lives in the wave's
// scratch (ws, wsi) between levels.
template <bool MIXED>
__device__ __forceinline__ void batch_unit(const BatchArgs& A, long item, int lane, double* __restrict__ ws,
                                           int* __restrict__ wsi, BatchLds& S) {
  const Unit U = unit_of(A, item);
  const int n_eb = U.n_eb;
  double* w_sum = ws;
  double* w_lastmag = ws + 6 * kBatchEB;
  double* w_prec = ws + 7 * kBatchEB;
  double* w_an = ws + 8 * kBatchEB;
  double* w_dm = ws + 9 * kBatchEB;
  int* w_level = wsi;
  int* w_flags = wsi + kBatchEB;
  for (int e = lane; e < n_eb; e += 64) {
#pragma unroll
    for (int k = 0; k < 6; ++k) w_sum[k * kBatchEB + e] = 0;
    w_lastmag[e] = -1;
    w_prec[e] = -1;
    w_an[e] = 0;
    w_dm[e] = 0;
    w_level[e] = -1;
    w_flags[e] = 0;
  }
  __syncthreads();  // the state written above is read by other lanes
  unsigned long long n_amp = 0, n_ev = 0, n_slot = 0;  // wave-uniform work counts of this unit
  unit_levels<MIXED, true>(A, U, lane, ws, wsi, S, kFConverged | kFHanded, n_amp, n_ev, n_slot);
  if (MIXED) {
    // FP64 recheck: reset the borderline energies, park the others
    int n_re = 0;
    for (int t = 0; t < n_eb; t += 64) {
      const int e = t + lane;
      const int ec = e < n_eb ? e : 0;
      const int flags = w_flags[ec];
      const bool want = e < n_eb && (flags & (kFBorderline | kFHanded)) == kFBorderline;
      n_re += __popcll(__ballot(want));
      if (want) {
#pragma unroll
        for (int k = 0; k < 6; ++k) w_sum[k * kBatchEB + e] = 0;
        w_lastmag[e] = -1;
        w_prec[e] = -1;
        w_level[e] = -1;
        w_flags[e] = kFRecheck;
      } else if (e < n_eb) {
        w_flags[e] = flags | kFIdle;
      }
    }
    __syncthreads();
    if (n_re > 0) {
      unit_levels<false, false>(A, U, lane, ws, wsi, S, kFConverged | kFHanded | kFIdle, n_amp, n_ev, n_slot);
    }
    atomicAdd(A.n_recheck, lane == 0 ? (unsigned long long) n_re : 0ULL);
  }
  // outputs and bookkeeping for the block's energies (all lanes take part in
  // the atomics; only lane 0 adds nonzero counts)
  for (int t = 0; t < n_eb; t += 64) {
    const int e = t + lane;
    const int ec = e < n_eb ? e : 0;
    const int lev = w_level[ec];
    const int flags = w_flags[ec];
    const bool v = e < n_eb && (flags & kFHanded) == 0;  // handed-over energies finish in the deep kernel
    if (v) {
      double f = A.C0 * (U.T / (2.0 * (double) (1L << lev) - 1));
      double E[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) E[k] = w_sum[k * kBatchEB + e] * f;
      if (A.use_pol) pol_project(E, A.pol);
      atomicAdd(&U.R.out[U.io * U.R.n_e + U.e0 + e], A.C2 * cmag2(E) * A.weight[U.ip]);
    }
    unsigned long long nc = __ballot(v && !(flags & kFConverged));
    unsigned long long ev = v ? (unsigned long long) ((2L << lev) - 1) : 0ULL;
    for (int off = 32; off > 0; off >>= 1) ev += __shfl_down(ev, off);
    atomicAdd(A.n_unconverged, lane == 0 ? (unsigned long long) __popcll(nc) : 0ULL);
    atomicAdd(A.evals, lane == 0 ? ev : 0ULL);
  }
  atomicAdd(A.main_amps, lane == 0 ? n_amp : 0ULL);
  atomicAdd(A.main_evals, lane == 0 ? n_ev : 0ULL);
  atomicAdd(A.main_slots, lane == 0 ? n_slot : 0ULL);
  __syncthreads();  // the next unit overwrites the scratch
}

// Four waves per SIMD: the register budget is set for the hot stored-row path;
// the rare interpolation path may spill.
template <bool MIXED>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) lw_batch_kernel(BatchArgs A) {
  __shared__ BatchLds S;
  int lane = threadIdx.x;
  double* ws = A.ws + (long) blockIdx.x * kBatchWsTotal;
  int* wsi = A.wsi + (long) blockIdx.x * kBatchWsI * kBatchEB;
  while (true) {
    // Next work unit.  Every lane executes the atomic (lane 0 adds 1, the others
    // 0) and the index is taken from lane 0 with readfirstlane, so it is a
    // scalar and the loop exit is a uniform branch.  No lane-conditional block
    // surrounds a cross-lane operation, which keeps the compiler from splitting
    // the wave between two copies of the loop.
    unsigned long long mine = atomicAdd(A.counter, lane == 0 ? 1ULL : 0ULL);
    long item = (long) uniform_u64(mine);
    if (item >= A.total_items) break;
    batch_unit<MIXED>(A, item, lane, ws, wsi, S);
  }
}
// Waves per point in the per-point kernels: each level's samples are split
// into kPointWaves contiguous slices, one per wave, combined in LDS.
constexpr int kPointWaves = 8;

// One point (code = item * kBatchEB + energy-in-block) summed by a workgroup
// of kPointWaves waves from level L0 with the given state; then its flux,
// counts and (MIXED, borderline) its fix-up entry.  Levels, convergence test
// and weighting are the batch kernel's.  Every thread returns with the same
// decisions; thread 0 does the atomics.
template <bool MIXED>
__device__ __forceinline__ void batch_point(const BatchArgs& A, unsigned long long code, int L0, double sum[6],
                                            double lastmag, double an_run, double dm_last, int flags, int lev,
                                            double (*sh)[kPointWaves], unsigned long long* sh_slot) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  long item = (long) (code / kBatchEB);
  long e_in = (long) (code - (unsigned long long) item * kBatchEB);
  int lo = 0, hi = A.np;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (A.item_prefix[mid] <= item) lo = mid; else hi = mid;
  }
  const int ip = lo;
  const long local = item - A.item_prefix[ip];
  const ReqDev R = A.req[A.req_of[ip]];
  const long io = local / R.n_blocks_e;
  const long ie = (local - io * R.n_blocks_e) * kBatchEB + e_in;
  const V3 obs = v3(R.obs[3 * io], R.obs[3 * io + 1], R.obs[3 * io + 2]);
  const double ts = A.tspan[2 * ip], te = A.tspan[2 * ip + 1];
  const double T = te - ts;
  const double* prow = A.rows + 10 * (long) ip * A.stride;
  const TrajDesc D = A.desc[ip];
  const double negw = -R.omega[ie];
  bool c = false;
  for (int L = L0; L <= A.Lmax && !c; ++L) {
    const long ns = 1L << L;
    const double p2L = (double) (1L << L);
    const double spacing = L > A.stop_mem ? (te - ts) / (p2L - 1) : (te - ts) / p2L;
    const double tfirst = ts + (te - ts) / (2.0 * p2L);
    const double* __restrict__ rows = prow + 10 * ((1L << L) - 1);
    const long i0 = ns * w / kPointWaves, i1 = ns * (w + 1) / kPointWaves;
    double t6[6], md, al;
    if (L <= A.Lb) {
      slice_level_sum_rows<MIXED>(rows, i0, i1, obs, negw, lane, t6, md, al);
    } else {
      slice_level_sum<MIXED>(
          i0, i1, negw, lane, [&](long i) { return interpolated_amplitude(A, D, obs, tfirst + spacing * (double) i); },
          t6, md, al);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sh[k][w] = t6[k];
      sh[6][w] = md;
      sh[7][w] = al;
    }
    __syncthreads();
    double tot[6] = {0, 0, 0, 0, 0, 0};
    double maxd = 0, anl = 0;
    for (int v = 0; v < kPointWaves; ++v) {  // in slice order: every thread forms the same totals
#pragma unroll
      for (int k = 0; k < 6; ++k) tot[k] += sh[k][v];
      maxd = fmax(maxd, sh[6][v]);
      anl += sh[7][v];
    }
    __syncthreads();  // sh is rewritten at the next level
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
    double prec = fabs(mag - lastmag) / lastmag;
    c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == lastmag);
    if (MIXED) {
      an_run += anl;
      if (level_borderline(L, dt, A.precision, lastmag, mag, prec, maxd, an_run, dm_last)) flags |= 2;
    }
    lastmag = mag;
  }
  c = c || (flags & 1);
  bool fixed = false;
  if (MIXED && (flags & 2)) {  // the same for every thread
    if (threadIdx.x == 0) *sh_slot = atomicAdd(A.n_fix, 1ULL);
    __syncthreads();
    unsigned long long j = *sh_slot;
    __syncthreads();
    if (j < (unsigned long long) A.fix_cap) {
      if (threadIdx.x == 0) A.fix_list[j] = code;
      fixed = true;
    }
  }
  if (threadIdx.x == 0 && !fixed) {
    double f = A.C0 * (T / (2.0 * (double) (1L << lev) - 1));
    double E[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) E[k] = sum[k] * f;
    if (A.use_pol) pol_project(E, A.pol);
    atomicAdd(&R.out[io * R.n_e + ie], A.C2 * cmag2(E) * A.weight[ip]);
    if (!c) atomicAdd(A.n_unconverged, 1ULL);
    atomicAdd(A.evals, (unsigned long long) ((2L << lev) - 1));
  }
}

// Points the batch kernel handed over at level Ldeep (slowly converging, with
// many samples per level): one workgroup per point from its stored state.  A
// grid-stride loop over the handed-over count (read from device memory, so the
// launch can sit in a captured graph).
template <bool MIXED>
__global__ void __launch_bounds__(64 * kPointWaves) lw_batch_deep_kernel(BatchArgs A) {
  __shared__ double sh[8][kPointWaves];
  __shared__ unsigned long long sh_slot;
  long n = (long) *A.n_deep;
  if (n > A.deep_cap) n = A.deep_cap;
  for (long k = blockIdx.x; k < n; k += gridDim.x) {
    unsigned long long code = A.deep_code[k];
    double st[kBatchWsD];
#pragma unroll
    for (int q = 0; q < kBatchWsD; ++q) st[q] = A.deep_state[q * A.deep_cap + k];
    int lev = A.deep_level[k];
    int flags = A.deep_flags[k];
    batch_point<MIXED>(A, code, lev + 1, st, st[6], st[8], st[9], flags, lev, sh, &sh_slot);
  }
}

// FP64 recomputation of the points the mixed kernels listed, from level 0, so
// their levels and values are the FP64 path's.  One wave per point, its lanes
// splitting each level's samples (a unit rarely holds more than a few such
// points, so sharing the amplitude across them would leave most lanes idle):
// no workgroup barrier, and a point's low levels occupy one wave, not eight.
// Four waves per workgroup; the waves stride over the list independently.
__global__ void __launch_bounds__(256) lw_batch_fix_kernel(BatchArgs A) {
  const int lane = threadIdx.x & 63;
  const long wave = (long) blockIdx.x * 4 + (threadIdx.x >> 6);
  const long n_waves = (long) gridDim.x * 4;
  long n = (long) *A.n_fix;
  if (n > A.fix_cap) n = A.fix_cap;
  for (long k = wave; k < n; k += n_waves) {
    const unsigned long long code = A.fix_list[k];
    const long item = (long) (code / kBatchEB);
    const long e_in = (long) (code - (unsigned long long) item * kBatchEB);
    int lo = 0, hi = A.np;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (A.item_prefix[mid] <= item) lo = mid; else hi = mid;
    }
    const int ip = lo;
    const long local = item - A.item_prefix[ip];
    const ReqDev R = A.req[A.req_of[ip]];
    const long io = local / R.n_blocks_e;
    const long ie = (local - io * R.n_blocks_e) * kBatchEB + e_in;
    const V3 obs = v3(R.obs[3 * io], R.obs[3 * io + 1], R.obs[3 * io + 2]);
    const double ts = A.tspan[2 * ip], te = A.tspan[2 * ip + 1];
    const double T = te - ts;
    const double* prow = A.rows + 10 * (long) ip * A.stride;
    const TrajDesc D = A.desc[ip];
    const double negw = -R.omega[ie];
    double sum[6] = {0, 0, 0, 0, 0, 0};
    double lastmag = -1;
    int lev = 0;
    bool c = false;
    for (int L = 0; L <= A.Lmax && !c; ++L) {
      const long ns = 1L << L;
      double tot[6], maxd, an;
      if (L <= A.Lb) {
        slice_level_sum_rows<false>(prow + 10 * ((1L << L) - 1), 0, ns, obs, negw, lane, tot, maxd, an);
      } else {
        const double p2L = (double) (1L << L);
        const double spacing = L > A.stop_mem ? (te - ts) / (p2L - 1) : (te - ts) / p2L;
        const double tfirst = ts + (te - ts) / (2.0 * p2L);
        slice_level_sum<false>(
            0, ns, negw, lane,
            [&](long i) { return interpolated_amplitude(A, D, obs, tfirst + spacing * (double) i); }, tot, maxd,
            an);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) sum[q] += tot[q];
      lev = L;
      // convergence test, src/OSCARSSR.cc:4227-4239 (the batch kernel's)
      double E[6];
      const double dt = T / (2.0 * (double) (1L << L) - 1);
#pragma unroll
      for (int q = 0; q < 6; ++q) E[q] = sum[q] * dt;
      if (A.use_pol) pol_project(E, A.pol);
      const double mag = cmag2(E);
      const double prec = fabs(mag - lastmag) / lastmag;
      c = (L > 8 && prec < A.precision && maxd < kPi) || (L > 8 && maxd < kPi && mag == lastmag);
      lastmag = mag;
    }
    // every lane holds the same totals (butterfly sums): lane 0 adds the point
    const double f = A.C0 * (T / (2.0 * (double) (1L << lev) - 1));
    double E[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) E[q] = sum[q] * f;
    if (A.use_pol) pol_project(E, A.pol);
    if (lane == 0) {
      atomicAdd(&R.out[io * R.n_e + ie], A.C2 * cmag2(E) * A.weight[ip]);
      if (!c) atomicAdd(A.n_unconverged, 1ULL);
      atomicAdd(A.evals, (unsigned long long) ((2L << lev) - 1));
    }
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

// The batch: the persistent kernel (stage bit 0), the per-point kernel for the
// points it handed over (bit 1) and, mixed, the FP64 recomputation of
// borderline points (bit 2); stages = 7 launches all three in order.
extern "C" hipError_t om_launch_lw_batch_stages(const om::BatchArgs* A, int mixed, int n_groups, int stages,
                                                hipStream_t s) {
  int pg = n_groups / om::kPointWaves > 0 ? n_groups / om::kPointWaves : 1;  // same wave count
  dim3 pb(64 * om::kPointWaves);
  if (mixed) {
    if (stages & 1) hipLaunchKernelGGL(om::lw_batch_kernel<true>, dim3(n_groups), dim3(64), 0, s, *A);
    if (stages & 2) hipLaunchKernelGGL(om::lw_batch_deep_kernel<true>, dim3(pg), pb, 0, s, *A);
    if (stages & 4) {
      int fb = n_groups / 4 > 0 ? n_groups / 4 : 1;  // four single-wave points per workgroup
      hipLaunchKernelGGL(om::lw_batch_fix_kernel, dim3(fb), dim3(256), 0, s, *A);
    }
  } else {
    if (stages & 1) hipLaunchKernelGGL(om::lw_batch_kernel<false>, dim3(n_groups), dim3(64), 0, s, *A);
    if (stages & 2) hipLaunchKernelGGL(om::lw_batch_deep_kernel<false>, dim3(pg), pb, 0, s, *A);
  }
  return hipGetLastError();
}

extern "C" hipError_t om_launch_lw_batch(const om::BatchArgs* A, int mixed, int n_groups, hipStream_t s) {
  return om_launch_lw_batch_stages(A, mixed, n_groups, 7, s);
}
