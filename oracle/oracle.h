/* CPU oracle: a plain-C restatement of the reference OSCARS CPU path for the
 * synchrotron-radiation hot path (RK4 trajectory -> cubic spline -> level
 * samples -> Lienard-Wiechert spectrum / flux / power-density sums).
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP product path and
 * the "port" CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (oscars_amd/) never calls it.
 *
 * Every function follows the reference operation by operation (same
 * association order, no FMA contraction) so that it reproduces the reference
 * binary bit for bit; each cites the reference file:line it restates.
 */
#ifndef OSCARS_ORACLE_H
#define OSCARS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_FIELDS 16
#define ORC_MAX_LEVEL 26 /* TParticleA::kMaxTrajectoryLevel, include/TParticleA.h:72 */

enum { ORC_FIELD_UNDULATOR = 0, ORC_FIELD_EPU = 1, ORC_FIELD_GRID = 2, ORC_FIELD_GAUSSIAN = 3, ORC_FIELD_UNIFORM = 4,
       ORC_FIELD_QUADRUPOLE = 5 };

typedef struct {
  int type;
  /* undulator / EPU (src/TField3D_IdealUndulator.cc:63-107, src/TField3D_IdealEPU.cc:64-112) */
  double field[3], fieldB[3], period[3], center[3];
  int nperiods;
  double phase, taper;
  double period_length, unit[3], und_length;
  /* grid (src/TField3D_Grid.cc:353-548): regular grid, 1-D along one axis */
  long nx, ny, nz;
  double xstart, ystart, zstart, xstep, ystep, zstep, xstop, ystop, zstop;
  int dimx; /* 1 = X, 2 = Y, 3 = Z */
  double rot[3], trans[3];
  double* data; /* 3 * n values, owned */
  /* Gaussian (sigma) / uniform box (width) per axis, quadrupole gradient and width */
  double shape[3];
  double K, width;
  /* time dependence: the field is multiplied by
   * cos(2 pi freq (T + toff) + fphase) (src/TField3D_IdealUndulator.cc:147-151
   * and the same in every model; the grid always, src/TField3D_Grid.cc:277) */
  double freq, fphase, toff;
} orc_field;

typedef struct {
  orc_field f[ORC_MAX_FIELDS];
  int n;
} orc_fields;

typedef struct {
  double x0[3], b0[3], t0; /* t0 in metres (c*t), like the reference */
  double q, m, current, gamma, qomg;
} orc_particle;

/* Trajectory + spline.  p is n rows of 9 (x y z bx by bz ax ay az), ypp the same. */
typedef struct {
  long n;
  double* t;
  double* p;
  double* ypp;
  double tstart, tstop;
} orc_traj;

/* ---- fields ---- */
void orc_fields_init(orc_fields* F);
void orc_fields_free(orc_fields* F);
/* src/TField3D_Gaussian.cc:27-60, TField3D_UniformBox.cc:45-86, TField3D_Quadrupole.cc:20-40 */
int orc_add_gaussian(orc_fields* F, const double* bfield, const double* center, const double* sigma,
                     const double* rotations);
int orc_add_uniform(orc_fields* F, const double* bfield, const double* width, const double* center,
                    const double* rotations);
int orc_add_quadrupole(orc_fields* F, double K, double width, const double* rotations, const double* translation);
int orc_add_undulator(orc_fields* F, const double* bfield, const double* period, int nperiods,
                      const double* translation, double phase, double taper);
int orc_add_epu(orc_fields* F, const double* bfieldA, const double* bfieldB, const double* period, int nperiods,
                const double* translation, double phase, double taper);
/* Grid from header values + samples as the OSCARS text reader would produce them. */
int orc_add_grid_exact(orc_fields* F, const double* start, const double* stop, const double* step, const long* npts,
                       const double* samples, const double* rotations, const double* translation);
int orc_add_grid(orc_fields* F, const double* start_in, const double* step_in, const long* npts,
                 const double* samples, const double* rotations, const double* translation);
void orc_get_b(const orc_fields* F, const double* x, double t, double* out);
/* RK4 steps (and sub-steps) the last trajectory integration split in two */
long orc_halving_count(void);
/* deepest split of that integration (1: a step halved once) */
int orc_halving_depth(void);

/* ---- particle ---- */
void orc_particle_electron(orc_particle* P, double energy_GeV, const double* x0, const double* d0, double current);
void orc_particle_set(orc_particle* P, const double* x0, const double* b0, double t0);

/* ---- trajectory ---- */
long orc_npoints_forward_backward(double ctstart, double ctstop, long npoints, double t0, long* nbwd);
/* with electric fields FE as well (NULL: none), src/OSCARSSR.cc:1283-1318 */
int orc_trajectory_rk4_eb(const orc_fields* F, const orc_fields* FE, const orc_particle* P, double ctstart,
                          double ctstop, long npoints, double* out_t, double* out_p);
int orc_trajectory_rk4(const orc_fields* F, const orc_particle* P, double ctstart, double ctstop, long npoints,
                       double* out_t, double* out_p);
int orc_spline(long n, const double* t, const double* p, double* ypp);
void orc_spline_value(const orc_traj* T, double x, double* out9);
long orc_level_npoints(int level);
/* memory levels: tstart + T/2^(L+1) + i T/2^L; extended: spacing T/(2^L - 1) */
void orc_level_samples(const orc_traj* T, int level, int extended, double* out /* 10 per sample */);
double orc_dt_inclusive(const orc_traj* T, int level);

/* ---- geometry ---- */
int orc_rectangle(const char* plane, int n1, int n2, double w1, double w2, const double* rotations,
                  const double* translation, int normal, int dim, double* out_obs /*3n*/, double* out_normal /*3*/,
                  double* out_container /*3n: the coordinates the reference stores*/);
void orc_energies(int n, double e0, double e1, double* out);

/* ---- polarization (src/OSCARSSR.cc:1887-1919) ---- */
/* pol: 0 all, 1 lh, 2 lv, 3 linear(angle), 4 cl, 5 cr; out = 6 doubles (re,im) x 3 */
int orc_polarization_code(const char* s);
void orc_polarization_vector(int code, double angle, const double* hdir, const double* pdir, double* out6);

/* ---- integrals ----
 * carry_quirk=1 reproduces the reference's per-thread carry of Result_Level /
 * Result_Precision across points (so quantity 1/2 and the not-converged flags
 * depend on nthreads); carry_quirk=0 reports each point on its own
 * (level -1 / not converged when it did not converge).
 * out_level: level actually summed to (LastLevel); out_conv: 1 if converged;
 * out_e: 6 doubles per point, C0*dt*SumE after polarization (re, im per axis).
 */
typedef struct {
  double precision;
  int max_level, max_level_extended;
  int quantity;
  int nthreads;
  int carry_quirk;
  double weight;
} orc_opts;

int orc_spectrum(const orc_traj* T, const orc_particle* P, const double* obs, int n_omega, const double* energies_eV,
                 const double* pol6, const orc_opts* o, double* out_value, int* out_level, unsigned char* out_conv,
                 double* out_e);
int orc_flux(const orc_traj* T, const orc_particle* P, const double* obs_xyz, long n_obs, double energy_eV,
             const double* pol6, const orc_opts* o, double* out_value, int* out_level, unsigned char* out_conv,
             double* out_e);
int orc_power_density(const orc_traj* T, const orc_particle* P, const double* obs_xyz, const double* normal,
                      long n_obs, int directional, const orc_opts* o, double* out_value, int* out_level,
                      unsigned char* out_conv);

int orc_power_density_surface(const orc_traj* T, const orc_particle* P, const double* obs_xyz,
                              const double* normals, long n_obs, int directional, const orc_opts* o,
                              double* out_value, int* out_level, unsigned char* out_conv);

int orc_total_power(const orc_traj* T, const orc_particle* P, const orc_opts* o, double* out_power,
                    double* out_precision, int* out_level);

int orc_power_density_stl(const orc_traj* T, const orc_particle* P, const double* tri, long n, const orc_opts* o,
                          double* out_value, int* out_level, unsigned char* out_conv);
void orc_efield_time(const orc_traj* T, const orc_particle* P, const double* Obs, double dt, double* out_t,
                     double* out_e);

/* Count of integrand evaluations sum_p (2^(L_p+1) - 1) given per-point levels. */
double orc_evals_from_levels(long n, const int* levels);

#ifdef __cplusplus
}
#endif
#endif
