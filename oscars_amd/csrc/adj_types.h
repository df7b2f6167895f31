// Kernel argument blocks of the reverse pass, shared by adjoint.hip and the host library.
#pragma once

#include "om_types.h"

namespace om {

struct AdjLwArgs {
  const double* rows;   // ns samples x 10 (t, x, beta, a/c), levels 0..L concatenated
  long ns;
  const double* obs;    // n_obs x 3
  long n_obs;
  const double* omega;  // n_e angular frequencies
  long n_e;
  const double* g;      // n_obs x n_e x 6: dLoss/d(Re S_k), dLoss/d(Im S_k)
  long obs_per_chunk;
  double* partial;      // n_chunks x ns x 9
};

struct AdjRkArgs {
  const FieldDev* F;
  int nfields;
  int gfield;
  // grid-sample gradient contributions: 10 records per knot (5 field
  // evaluations, each adding to two neighbouring samples), record r = grid
  // sample grec_key[r] (ngrid: none) and the 3 values grec_val[3r..3r+2];
  // summed per sample in record order afterwards, so the gradient does not
  // depend on scheduling
  int* grec_key;       // null: no grid to differentiate
  double* grec_val;
  int ngrid;
  double qomg, dt;
  long nf, K, nseg;
  const double* ckpt;  // nseg x 6
  const double* kb;    // nf x 9: dLoss/d(knot x, beta, a/c)
  double* segmap;      // nseg x 42: M (6x6 row-major), b (6)
  double* lam;         // (nseg + 1) x 6: lambda at segment starts
  double* ybuf;        // nf x 6 recomputed states
  long lds_cap;        // doubles of dynamic LDS the segment kernels may stage the fields in (set by the launcher)
};

}  // namespace om
