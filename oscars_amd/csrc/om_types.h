// Plain-data structures shared by the host library and the gfx950 kernels.
#pragma once

#include <stdint.h>

namespace om {

constexpr int kMaxLevel = 26;  // TParticleA::kMaxTrajectoryLevel (include/TParticleA.h:72)
constexpr int kMaxFields = 16;

enum FieldType : int { kUndulator = 0, kEPU = 1, kGrid = 2, kGaussian = 3, kUniform = 4, kQuadrupole = 5 };

// One magnetic field model, evaluated on the device inside the RK4 kernel.
struct FieldDev {
  int type;
  int nperiods;
  int axis;  // grid: 0 x, 1 y, 2 z when one-dimensional, else -1
  int pad;
  double field[3], fieldB[3], period[3], center[3];
  double phase, taper;
  double period_length, unit[3], und_length;
  double inv_period_length;  // RN(1/period_length), for the exact Markstein quotient
  double phase_shift_a, phase_shift_b;
  // 1-D grid (src/TField3D_Grid.cc:153-278, open interval, linear interpolation)
  long n[3];
  double start[3], step[3], stop[3], inv_step[3];
  double rot[3], trans[3];
  double rot_sin[3], rot_cos[3];
  const double* data;  // device pointer, 3 per sample
  // Gaussian / uniform box / quadrupole: field (rotated on the host) in field[],
  // centre in center[], sigma or full width per axis (<= 0: axis ignored),
  // gradient and width along the axis of a quadrupole, its translation in trans[]
  double shape[3];
  double K, width;
  int dims;  // grid: axes with more than one sample, bit 0 x, bit 1 y, bit 2 z
  // time dependence: the model's value is multiplied by
  // cos(2 pi freq (t + toff) + fphase) where td is set (any model with a
  // frequency; a grid also with a phase alone, src/TField3D_Grid.cc:277)
  int td;
  double freq, fphase, toff;
};

// One particle's trajectory request (all per particle so that a batch can have
// uneven start/stop and lengths).
struct TrajDesc {
  double x0[3], b0[3];
  double t0_over_c;  // T0 / c in seconds (reference stores T0 in metres)
  double qomg;       // Q / M / gamma
  double qom;        // Q / M (equations of motion with electric fields)
  double dt;         // RK4 step [s]
  long nf, nb;       // forward / backward point counts
  long knot_offset;  // first knot in the packed knot arrays
  long n;            // nf + nb
  int status;        // written by the kernel: 0 ok, 1 the step halving met its depth limit
  int halved;        // written by the kernel: number of steps that were halved (|beta| >= 1 met)
};

}  // namespace om
