// Launch-argument structures shared by the host library (api.cpp) and the
// gfx950 kernels: plain data passed by value to the kernels.
#pragma once

#include "om_types.h"

namespace om {

// ---- single-particle Lienard-Wiechert sums (lw.hip)
struct Levels {
  const double* rows[kMaxLevel + 1];  // 2^L rows of 10 doubles per level (device)
  double dt_incl[kMaxLevel + 1];      // (tstop - tstart) / (2^(L+1) - 1)
  double beta_diff[kMaxLevel + 1];    // per level max |beta_i - beta_{i-1}| (power density)
};

struct State {
  double* sum;      // 6 doubles per point (Re/Im of Ex, Ey, Ez), power density: sum[0]
  double* err;      // mixed path (else null): 2 per point, sum over samples of |V|_1 and the
                    // error bound of lastmag
  double* lastmag;  // |E|^2 (or power sum) after the previous level
  double* prec;     // last relative change
  int* level;       // last level summed (-1 before any)
  int* conv;        // 1 once converged
  int* border;      // mixed path (else null): 1 if a level decision lay within the error bound
};

struct LwArgs {
  Levels lv;
  State st;
  int La, Lb;
  long npts;
  const double* obs;    // 3 per observation point
  const double* omega;  // angular frequency per energy
  long n_omega;         // point p -> (obs p / n_omega, energy p % n_omega)
  double pol[6];
  int use_pol;
  double precision;
  const int* list;      // null: dense grid.  Tile kernel: 64 point indices per tile (-1 pads);
                        // point kernel: n_list point indices, one per lane.
  long n_list;
};

struct PdArgs {
  Levels lv;
  State st;
  int La, Lb;
  long npts;
  const double* obs;
  double normal[3];         // one normal for every point (rectangles) ...
  const double* normals;    // ... or one per point (surfaces of points), when not null
  int has_normal;           // 0: a surface without normals (N1.normal taken as 1)
  int directional;
  double precision;
  double two_over_gamma;
  const int* list;          // null: every point; else n_list point indices, one per lane
  long n_list;
};

// Power density at the centres of STL triangles with occlusion by the others.
struct StlArgs {
  Levels lv;
  State st;
  int La, Lb;
  long ntri;
  const double* tri;  // 12 per triangle: A, B - A, C - A, normal as read
  const double* centre;  // 3 per triangle
  double precision;
  double two_over_gamma;
  // bounding-volume hierarchy in depth-first order (one leaf with every
  // triangle: the reference's test of every triangle for every ray)
  const double* bvh_box;  // 6 per node: low x y z, high x y z (inflated)
  const int* bvh_meta;    // 3 per node: next node when missed or done, first listed triangle, count (0: inner)
  const int* bvh_order;   // triangle indices in leaf order
  int bvh_nodes;
  double pad_t;           // slack of the ray segment at both ends [m]
};

// Total power: per-level sums of the integrand over the level's samples.
struct TpArgs {
  const double* rows;   // 2^L samples of (t, x, b, a/c)
  long npts;
  double* partial;      // one sum per workgroup
  double* out;          // the level's sum
};

struct FinArgs {
  State st;
  long npts;
  int kind;       // 0 spectrum/flux, 1 power density
  int quantity;   // 0 value, 1 precision, 2 level
  double C0, C2;  // spectrum/flux constants (src/OSCARSSR.cc:1882-1885)
  double pd_const;
  int directional;
  double dt_incl[kMaxLevel + 1];
  double pol[6];
  int use_pol;
  double weight;
  int accumulate;      // add weight*value into out with Kahan compensation (out_comp)
  double* out;
  double* out_comp;
  double* out_e;       // 6 per point (may be null)
  int* out_level;
  unsigned char* out_conv;
};

// ---- level samples (traj.hip, batch.hip)
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

// ---- batched multi-particle sums (batch.hip)
// A work unit is (particle, observation point, block of up to kBatchEB
// energies); the wave that takes it keeps the block's per-energy state in its
// own scratch: kBatchWsD doubles (sum[6], lastmag, prec, and for the mixed path
// the |V|_1 sum and the error bound of lastmag) and kBatchWsI ints (level,
// flags) per energy, field-major.
constexpr int kBatchEB = 512;
constexpr int kBatchWsD = 10;
constexpr int kBatchWsI = 2;
// plus, per wave, 7 partial sums x 64 lanes for each of the kBatchEB / 64
// energy groups of a level (kept between super-blocks of samples)
constexpr int kBatchWsTotal = kBatchWsD * kBatchEB + (kBatchEB / 64) * 7 * 64;

struct ReqDev {
  const double* obs;    // 3 per point
  const double* omega;  // per energy
  double* out;          // n_obs * n_e
  long n_obs, n_e, n_blocks_e;  // n_blocks_e = ceil(n_e / kBatchEB)
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
  const long* item_prefix;  // np + 1: first work unit of each particle
  int np;
  const ReqDev* req;
  double C0, C2;
  double pol[6];
  int use_pol;
  double precision;
  unsigned long long* counter;
  long total_items;
  unsigned long long* n_unconverged;
  unsigned long long* evals;  // sum over points of 2^(L+1)-1 (as integer)
  double* ws;                 // per persistent wave: kBatchWsTotal doubles
  int* wsi;                   // per persistent wave: kBatchWsI x kBatchEB ints
  // mixed path: points the batch kernel recomputed in FP64 inside their unit
  // (their stop decision lay within the reduced-accuracy sin/cos error bound)
  unsigned long long* n_recheck;
  // mixed path: handed-over points whose stop decision lay within the
  // reduced-accuracy sin/cos error bound, as item * kBatchEB + energy-in-block, recomputed in FP64 by
  // lw_batch_fix_kernel
  unsigned long long* fix_list;
  long fix_cap;
  unsigned long long* n_fix;  // points flagged (those beyond fix_cap keep the mixed result)
  // points still active at level Ldeep are handed to lw_batch_deep_kernel (one
  // wave per point) with their state: code as in fix_list, kBatchWsD doubles
  // (field-major, deep_cap per field), level and flags
  int Ldeep;
  unsigned long long* deep_code;
  double* deep_state;
  int* deep_level;
  int* deep_flags;
  long deep_cap;
  unsigned long long* n_deep;  // points handed over (those beyond deep_cap stay with their unit)
  // work done inside lw_batch_kernel itself (for its roofline): amplitude
  // evaluations, one per (unit, energy group, sample), integrand evaluations,
  // one per (active energy, sample), and lane slots of the integrand loop
  // (idle lanes included)
  unsigned long long* main_amps;
  unsigned long long* main_evals;
  unsigned long long* main_slots;
};

}  // namespace om
