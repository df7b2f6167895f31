/* Body of fdref.c, included once per floating type (REAL, FN, SIN, COS, SQRT). */

typedef struct {
  const REAL* data; /* 3 per sample */
  long n;
  REAL start, step, stop;
} FN(grid);

/* src/TField3D_Grid.cc:153-278, z-only branch: zero outside the open interval,
 * linear interpolation between samples n0 and n0+1 */
static void FN(grid_b)(const FN(grid) * g, REAL z, REAL B[3]) {
  B[0] = B[1] = B[2] = 0;
  if (z <= g->start || z >= g->stop) return;
  size_t n0 = (size_t) ((z - g->start) / g->step);
  if ((long) n0 > g->n - 2) n0 = (size_t) (g->n - 2);
  REAL dd = (z - g->start) - (REAL) n0 * g->step;
  const REAL* a = g->data + 3 * n0;
  for (int k = 0; k < 3; ++k) B[k] = a[k] + (a[3 + k] - a[k]) * dd / g->step;
}

/* src/OSCARSSR.cc:1247-1279 */
static void FN(derivs)(const FN(grid) * g, REAL qomg, const REAL* x, REAL* d) {
  REAL B[3];
  FN(grid_b)(g, x[4], B);
  d[0] = x[1];
  d[1] = qomg * (-x[5] * B[1] + x[3] * B[2]);
  d[2] = x[3];
  d[3] = qomg * (x[5] * B[0] - x[1] * B[2]);
  d[4] = x[5];
  d[5] = qomg * (x[1] * B[1] - x[3] * B[0]);
}

/* src/OSCARSSR.cc:1321-1386 (the |beta| >= 1 branch never triggers here) */
static void FN(rk4)(const FN(grid) * g, REAL qomg, REAL* y, const REAL* dydx, REAL h) {
  REAL dym[6], dyt[6], yt[6];
  REAL hh = h * (REAL) 0.5, h6 = h / (REAL) 6;
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
  FN(derivs)(g, qomg, yt, dyt);
  for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
  FN(derivs)(g, qomg, yt, dym);
  for (int i = 0; i < 6; ++i) {
    yt[i] = y[i] + h * dym[i];
    dym[i] += dyt[i];
  }
  FN(derivs)(g, qomg, yt, dyt);
  for (int i = 0; i < 6; ++i) y[i] = y[i] + h6 * (dydx[i] + dyt[i] + (REAL) 2 * dym[i]);
}

/* Frozen-level flux (levels 0..level, no level test) of an electron starting at
 * T0 = ctstart; out is n_obs x n_e, energy fastest.  Returns 0 on success. */
int FN(fdref_flux)(const double* samples, long n_grid, double z0, double dz, const double* x0, const double* beta0,
                   double current, double ctstart, double ctstop, long npoints, const double* obs, long n_obs,
                   const double* energies_eV, long n_e, int level, double* out) {
  const REAL kC = (REAL) 2.99792458e8, kQe = (REAL) 1.602176462e-19, kMe = (REAL) 9.10938356e-31;
  const REAL kH = (REAL) 6.62606876e-34, kHbar = (REAL) 1.054571596e-34;
  const REAL kEps0 = (REAL) 8.854187817E-12, kMu0 = (REAL) 1.2566370614E-6;
  const REAL kPi = (REAL) 3.14159265358979323846264338327950288L;
  if (n_grid < 2 || npoints < 3 || level < 0 || level > 24) return 1;
  REAL* data = (REAL*) malloc(3 * n_grid * sizeof(REAL));
  for (long i = 0; i < 3 * n_grid; ++i) data[i] = (REAL) samples[i];
  /* midpoint form of the z start (src/TField3D_Grid.cc:425-430) */
  REAL step = (REAL) dz;
  REAL mid = (REAL) z0 + step * (REAL) (n_grid - 1) / (REAL) 2;
  FN(grid) g = {data, n_grid, mid - step * (REAL) (n_grid - 1) / (REAL) 2, step, 0};
  g.stop = g.start + (REAL) (n_grid - 1) * g.step;
  REAL q = -kQe;
  REAL b2 = (REAL) beta0[0] * beta0[0] + (REAL) beta0[1] * beta0[1] + (REAL) beta0[2] * beta0[2];
  REAL gamma = (REAL) 1 / SQRT((REAL) 1 - b2);
  REAL qomg = q / kMe / gamma;
  /* RK4 knots, acceleration stored one step late (src/OSCARSSR.cc:910-1009), T0 = ctstart */
  REAL DeltaT = ((REAL) ctstop - (REAL) ctstart) / kC / (REAL) (npoints - 1);
  long N = npoints;
  REAL* X = (REAL*) malloc(N * sizeof(REAL));
  REAL* Y = (REAL*) malloc(9 * N * sizeof(REAL));
  REAL* P = (REAL*) malloc(9 * N * sizeof(REAL));
  REAL* U = (REAL*) malloc(9 * N * sizeof(REAL));
  REAL y[6] = {(REAL) x0[0], (REAL) beta0[0] * kC, (REAL) x0[1], (REAL) beta0[1] * kC, (REAL) x0[2],
               (REAL) beta0[2] * kC};
  REAL d[6];
  FN(derivs)(&g, qomg, y, d);
  for (long i = 0; i < N; ++i) {
    X[i] = (REAL) ctstart / kC + DeltaT * (REAL) i;
    REAL* o = Y + 9 * i;
    o[0] = y[0]; o[1] = y[2]; o[2] = y[4];
    o[3] = y[1] / kC; o[4] = y[3] / kC; o[5] = y[5] / kC;
    o[6] = d[1] / kC; o[7] = d[3] / kC; o[8] = d[5] / kC;
    FN(derivs)(&g, qomg, y, d);
    FN(rk4)(&g, qomg, y, d, DeltaT);
  }
  /* spline with end slopes equal to the end secants (include/TOMATH.h:141-203) */
  for (int k = 0; k < 9; ++k) {
    P[k] = (REAL) -0.5;
    U[k] = 0;
  }
  for (long i = 1; i < N - 1; ++i) {
    REAL sig = (X[i] - X[i - 1]) / (X[i + 1] - X[i - 1]);
    for (int k = 0; k < 9; ++k) {
      REAL p = sig * P[9 * (i - 1) + k] + 2;
      P[9 * i + k] = (sig - 1) / p;
      REAL ui = (Y[9 * (i + 1) + k] - Y[9 * i + k]) / (X[i + 1] - X[i]) -
                (Y[9 * i + k] - Y[9 * (i - 1) + k]) / (X[i] - X[i - 1]);
      U[9 * i + k] = (ui * 6 / (X[i + 1] - X[i - 1]) - sig * U[9 * (i - 1) + k]) / p;
    }
  }
  for (int k = 0; k < 9; ++k) P[9 * (N - 1) + k] = (0 - (REAL) 0.5 * U[9 * (N - 2) + k]) / ((REAL) 0.5 * P[9 * (N - 2) + k] + 1);
  for (long i = N - 2; i >= 0; --i)
    for (int k = 0; k < 9; ++k) P[9 * i + k] = P[9 * i + k] * P[9 * (i + 1) + k] + U[9 * i + k];
  /* level samples 0..level at memory spacing (src/TParticleTrajectoryInterpolated.cc:124-161, 264-270) */
  REAL ts = X[0], te = X[N - 1], T = te - ts;
  long ns = (2L << level) - 1;
  REAL* S = (REAL*) malloc(10 * ns * sizeof(REAL));
  long js = 0;
  for (int L = 0; L <= level; ++L) {
    REAL p2 = (REAL) (1L << L);
    for (long i = 0; i < (1L << L); ++i, ++js) {
      REAL x = ts + T / (2 * p2) + (T / p2) * (REAL) i;
      long k = (long) ((x - X[0]) / DeltaT);
      if (k < 0) k = 0;
      if (k > N - 2) k = N - 2;
      while (k < N - 2 && X[k + 1] <= x) ++k;
      while (k > 0 && X[k] > x) --k;
      REAL h = X[k + 1] - X[k], a = (X[k + 1] - x) / h, b = (x - X[k]) / h;
      REAL* o = S + 10 * js;
      o[0] = x;
      for (int c = 0; c < 9; ++c)
        o[1 + c] = a * Y[9 * k + c] + b * Y[9 * (k + 1) + c] +
                   ((a * a * a - a) * P[9 * k + c] + (b * b * b - b) * P[9 * (k + 1) + c]) * (h * h) / 6;
    }
  }
  /* LW sum and flux (src/OSCARSSR.cc:4123-4126, 4179-4245) */
  REAL C0 = q / (4 * kPi * kC * kEps0 * SQRT(2 * kPi));
  REAL C2 = 4 * kPi * (REAL) current / (kH * -q * kMu0 * kC) * (REAL) 1e-6 * (REAL) 0.001;
  REAL dtL = T / ((REAL) (2L << level) - 1);
  for (long io = 0; io < n_obs; ++io) {
    REAL ob[3] = {(REAL) obs[3 * io], (REAL) obs[3 * io + 1], (REAL) obs[3 * io + 2]};
    for (long ie = 0; ie < n_e; ++ie) {
      REAL w = (REAL) energies_eV[ie] * kQe / kHbar;
      REAL sr[3] = {0, 0, 0}, si[3] = {0, 0, 0};
      for (long j = 0; j < ns; ++j) {
        const REAL* s = S + 10 * j;
        REAL R[3] = {ob[0] - s[1], ob[1] - s[2], ob[2] - s[3]};
        REAL D = SQRT(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        REAL Nv[3] = {R[0] / D, R[1] / D, R[2] / D};
        const REAL* be = s + 4;
        const REAL* ac = s + 7;
        REAL bb = be[0] * be[0] + be[1] * be[1] + be[2] * be[2];
        REAL W[3] = {Nv[0] - be[0], Nv[1] - be[1], Nv[2] - be[2]};
        REAL kk = 1 - (Nv[0] * be[0] + Nv[1] * be[1] + Nv[2] * be[2]);
        REAL Yv[3] = {W[1] * ac[2] - ac[1] * W[2], W[2] * ac[0] - ac[2] * W[0], W[0] * ac[1] - ac[0] * W[1]};
        REAL NY[3] = {Nv[1] * Yv[2] - Yv[1] * Nv[2], Nv[2] * Yv[0] - Yv[2] * Nv[0], Nv[0] * Yv[1] - Yv[0] * Nv[1]};
        REAL d1 = D * D * kk * kk, d2 = D * kk * kk;
        REAL phi = -w * (s[0] + D / kC);
        REAL cs = COS(phi), sn = SIN(phi);
        for (int c = 0; c < 3; ++c) {
          REAL A = W[c] * (1 - bb) / d1 + NY[c] / d2;
          sr[c] += A * cs;
          si[c] += A * sn;
        }
      }
      REAL f = 0;
      for (int c = 0; c < 3; ++c) {
        REAL er = sr[c] * C0 * dtL, ei = si[c] * C0 * dtL;
        f += er * er + ei * ei;
      }
      out[io * n_e + ie] = (double) (C2 * f);
    }
  }
  free(S);
  free(U);
  free(P);
  free(Y);
  free(X);
  free(data);
  return 0;
}
