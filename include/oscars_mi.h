/* oscars_mi.h - C ABI of the MI355X (gfx950) synchrotron-radiation hot path.
 *
 * Drop-in replacement for the reference's GPU boundary
 * (include/OSCARSSR_Cuda.h:28-66 of dhidas/OSCARS), redesigned as a true C ABI:
 * plain pointers and sizes, no C++ types, status codes instead of exceptions.
 * The caller (the C++/Python host, see oscars_amd/) samples particles and
 * describes fields; this library integrates trajectories on the device and
 * evaluates the Lienard-Wiechert sums.  Everything is synchronous: results
 * are in the caller's buffers when a call returns.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   om_device_count   <- OSCARSSR_Cuda_GetDeviceCount        include/OSCARSSR_Cuda.h:28
 *   om_device_info    <- OSCARSSR_Cuda_GetDeviceProperties   include/OSCARSSR_Cuda.h:30
 *   om_trajectories   <- the host-side OSR.CalculateTrajectory() call inside the
 *                        CUDA drivers (src/OSCARSSR_Cuda.cu:602-607, 1085-1088) and
 *                        OSCARSSR::CalculateTrajectoryRK4 (src/OSCARSSR.cc:910-1009)
 *   om_flux           <- OSCARSSR_Cuda_CalculateFluxGPU         include/OSCARSSR_Cuda.h:32-44
 *   om_spectrum       <- OSCARSSR_Cuda_CalculateSpectrumGPU     include/OSCARSSR_Cuda.h:46-57
 *   om_power_density  <- OSCARSSR_Cuda_CalculatePowerDensityGPU include/OSCARSSR_Cuda.h:59-66
 * All functions return 0 on success; om_last_error() gives the message.
 * A context is not thread-safe: use one per host thread.
 */
#ifndef OSCARS_MI_H
#define OSCARS_MI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct om_context om_context;

enum {
  OM_FIELD_UNDULATOR = 0,
  OM_FIELD_EPU = 1,
  OM_FIELD_GRID = 2,
  OM_FIELD_GAUSSIAN = 3,   /* src/TField3D_Gaussian.cc:27-112 */
  OM_FIELD_UNIFORM = 4,    /* src/TField3D_UniformBox.cc:45-120 */
  OM_FIELD_QUADRUPOLE = 5  /* src/TField3D_Quadrupole.cc:20-83 */
};
/* Kernel families: PLAIN one lane per output point, all FP64 (bit-identical
 * phase; the reference-faithful baseline); TILED one wave per observation
 * point x 64 energies sharing the per-sample amplitude, FP64; MIXED the tiled
 * kernel with reduced-accuracy sin/cos (FP64 polynomials within 4.8e-9) of the
 * exactly reduced FP64 phase, borderline stop decisions rechecked at full
 * accuracy; POINT_MIXED the per-point kernel with the same sin/cos
 * (single-energy maps).  Power density: PLAIN and TILED evaluate the
 * reference's FP64 term; MIXED and POINT_MIXED a term built on a Newton-refined
 * reciprocal square root and reciprocal (within about 1e-8 of the FP64 one),
 * borderline stop decisions recomputed with the FP64 term. */
enum { OM_MODE_PLAIN = 0, OM_MODE_TILED = 1, OM_MODE_MIXED = 2, OM_MODE_POINT_MIXED = 3 };
enum { OM_Q_VALUE = 0, OM_Q_PRECISION = 1, OM_Q_LEVEL = 2 };

/* One magnetic field model (src/TField3D_IdealUndulator.cc, TField3D_IdealEPU.cc,
 * TField3D_Grid.cc, TField3D_Gaussian.cc, TField3D_UniformBox.cc,
 * TField3D_Quadrupole.cc).  Undulator/EPU: field (fieldB for the EPU's second
 * component), period vector, nperiods, centre (translation), phase, taper.
 * Grid: the OSCARS text-file header values (start, step, n per axis; one,
 * two or three axes with n > 1, samples with z fastest, then y, then x), the
 * 3*n samples and the rotations/translation.
 * Gaussian: peak field, centre (translation), sigma per axis (shape; <= 0:
 * no fall-off along that axis), rotations.  Uniform: field, centre
 * (translation), full width per axis (shape; <= 0: unbounded), rotations.
 * Quadrupole: gradient K [T/m], width along the magnet axis, rotations,
 * translation.  Any model: frequency [Hz], frequency_phase and time_offset
 * [s] multiply it by cos(2 pi frequency (t + time_offset) + frequency_phase)
 * (src/TField3D_IdealUndulator.cc:147-151 and the same in every model; a grid
 * carries the factor also at zero frequency, src/TField3D_Grid.cc:277). */
typedef struct {
  int type;
  int nperiods;
  double field[3];
  double fieldB[3];
  double period[3];
  double translation[3];
  double phase;
  double taper;
  double grid_start[3];
  double grid_step[3];
  long grid_n[3];
  const double* grid_data;
  double rotations[3];
  double shape[3];
  double K;
  double width;
  /* grid: 0 = the OSCARS text header (start re-centred as the reader does),
   * 1 = start and grid_stop taken as given (a resampled OSCARS1D table) */
  int grid_exact;
  double grid_stop[3];
  double frequency;
  double frequency_phase;
  double time_offset;
} om_field;

/* One particle: initial position [m], beta, T0 [m] (c*t, as the reference),
 * charge [C], mass [kg], Q/M/gamma, current [A] and its time window. */
typedef struct {
  double x0[3];
  double beta0[3];
  double t0;
  double q;
  double m;
  double gamma;
  double qomg;
  double current;
  double ctstart;
  double ctstop;
  long npoints;
} om_particle;

/* Convergence and output controls (reference kwargs precision, max_level,
 * max_level_extended, quantity) plus the kernel family. */
typedef struct {
  double precision;
  int max_level;
  int max_level_extended;
  int quantity;
  int mode;
  double weight;
  int accumulate; /* 1: out_value += weight*value (Kahan-compensated across calls) */
} om_conv;

typedef struct {
  double evals;        /* sum over points of 2^(L+1)-1 */
  double kernel_ms;    /* LW kernels only, HIP events on the library stream */
  double total_ms;     /* whole call, host wall clock */
  int max_level;       /* highest level any point reached */
  long n_unconverged;
  long n_rechecked;    /* mixed modes: points whose stop decision lay within the sin/cos error
                          bound and were recomputed with full-accuracy sin/cos */
} om_stats;

int om_device_count(void);
int om_device_info(int device, char* buf, size_t len);

int om_context_create(int device, om_context** out);
void om_context_destroy(om_context* ctx);
const char* om_last_error(const om_context* ctx);

/* Integrate np trajectories on the device (RK4 + clamped cubic spline); they
 * stay resident in the context, indexed 0..np-1.  Replaces any previous set. */
int om_trajectories(om_context* ctx, const om_field* fields, int nfields, const om_particle* particles, int np);
/* A trajectory given by the caller instead of integrated: set_trajectory and
 * its file readers (src/OSCARSSR_Python.cc:5374-5566, src/OSCARSSR.cc:1140-1171).
 * n >= 3 knots at strictly increasing times t [s]; p9 per knot = x [m], beta,
 * a/c [1/s].  construct bit 0 builds beta from the positions' spline slope / c,
 * bit 1 builds a/c from beta's spline slope (src/TParticleTrajectoryPoints.cc:820-862);
 * those columns of p9 are then ignored.  The knots are splined as the reference
 * sets up its interpolated trajectory (src/TParticleA.cc:358-369).  Replaces any
 * previous set with this one trajectory (index 0); particle supplies Q/M/gamma
 * and the current.  Batched sums need integrated trajectories. */
int om_trajectory_set(om_context* ctx, const om_particle* particle, long n, const double* t, const double* p9,
                      int construct);
/* om_trajectories with electric fields as well (add_efield_*): the equations
 * of motion with E and B (src/OSCARSSR.cc:1186-1198, 1210-1245, 1283-1318). */
int om_trajectories_eb(om_context* ctx, const om_field* fields, int nfields, const om_field* efields, int nefields,
                       const om_particle* particles, int np);
/* Knots of trajectory i: *n in = capacity, out = count; any pointer may be NULL. */
int om_trajectory_get(om_context* ctx, int i, long* n, double* t, double* p9, double* ypp9);
/* Level samples of trajectory i (2^level rows of t,x,y,z,bx,by,bz,ax,ay,az). */
int om_level_get(om_context* ctx, int i, int level, int extended, double* rows);
/* Electric field in the time domain at obs[3] from trajectory i, one value per
 * knot: calculate_efield_vs_time (src/OSCARSSR_Python.cc:9048-9110,
 * src/OSCARSSR.cc:4471-4512).  out_t[k] = k dt + |obs - x_k| / c [s] (dt: the
 * integration step; 0 for a trajectory given by om_trajectory_set), out_e3[3 k]
 * = E [V/m].  *n in = capacity, out = knot count; nothing is written when the
 * capacity is short. */
int om_efield_time(om_context* ctx, int i, const double* obs, long* n, double* out_t, double* out_e3);
/* Sum of the given fields at n points (xyz: n rows of x, y, z [m]) into
 * out3 (n rows), evaluated on the device: get_bfield / get_efield
 * (src/OSCARSSR_Python.cc:1927-1975, src/OSCARSSR.cc:201-218, 381-398,
 * src/TFieldContainer.cc:85-107).  The fields are those of this call only; the
 * context's trajectories and their fields are left as they were.  No fields
 * gives zeros. */
int om_field_points(om_context* ctx, const om_field* fields, int nfields, const double* xyz, long n, double t,
                    double* out3);
/* Status of trajectory i: 0 ok (steps that met |beta| >= 1 were halved as the
 * reference halves them, src/OSCARSSR.cc:1362-1378), 1 a step still met
 * |beta| >= 1 after 40 halvings (the reference recurses without limit). */
int om_trajectory_status(om_context* ctx, int i);

/* Flux on n_obs observation points x n_e photon energies (point index
 * obs*n_e + e); n_e = 1 is the reference's flux map at one energy.
 * pol6: polarization vector (Re, Im) x 3, all zeros for 'all'.
 * out_level: the last level summed per point (its converged level, or the
 * level cap where out_conv is 0: the value is the sum at the cap, as in the
 * reference); with quantity OM_Q_LEVEL out_value holds the converged level
 * or -1, the reference's Result_Level (src/OSCARSSR_Cuda.cu:483-498).
 * out_conv: 1 converged, 0 NotConverged (decided per point, as the
 * reference's GPU kernel does).
 * out_e: complex E (6 doubles per point) after C0*dt and polarization; may be NULL. */
int om_flux(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV, long n_e,
            const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
            double* out_e, om_stats* stats);
/* Spectrum at one observation point (om_flux with n_obs = 1 and the
 * spectrum's |current| constant). */
int om_spectrum(om_context* ctx, int particle, const double* obs, const double* energies_eV, long n_e,
                const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
                double* out_e, om_stats* stats);
/* Power density [W/mm^2] on n_obs points with one surface normal. */
int om_power_density(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* normal,
                     int directional, const om_conv* conv, double* out_value, int* out_level,
                     unsigned char* out_conv, om_stats* stats);

/* Power density on a surface of points: calculate_power_density(points=...)
 * and calculate_power_density_line (src/OSCARSSR_Python.cc:6105-6380,
 * 7355-7560; integrand src/OSCARSSR.cc:2817-2978).  normals[3 n_obs] holds a
 * unit normal per point, or NULL for a surface without normals (the line),
 * where the reference takes N.normal as 1.  Host pointers. */
int om_power_density_surface(om_context* ctx, int particle, const double* obs_xyz, long n_obs,
                             const double* normals, int directional, const om_conv* conv, double* out_value,
                             int* out_level, unsigned char* out_conv, om_stats* stats);

/* Power density at the centres of STL triangles: calculate_power_density_stl
 * (src/OSCARSSR_Python.cc:6993-7213, src/OSCARSSR.cc:3318-3527).  tri15[15 n]
 * per triangle: vertices A, B, C [m], the normal as read from the file, the
 * centre.  A trajectory sample is skipped for a triangle when the segment to
 * its centre meets another triangle (src/TTriangle3D.cc:159-215).  Host
 * pointers. */
int om_power_density_stl(om_context* ctx, int particle, const double* tri15, long n, const om_conv* conv,
                         double* out_value, int* out_level, unsigned char* out_conv, om_stats* stats);

/* Total radiated power of trajectory `particle` [W]: calculate_total_power
 * (src/OSCARSSR_Python.cc:5974-6060, src/OSCARSSR.cc:3725-3843).  conv gives
 * precision and the level limits; *out_level is -1 when no level met the
 * precision (the reference raises "total power calculation did not
 * converge"), with *out_power the last level's value. */
int om_total_power(om_context* ctx, int particle, const om_conv* conv, double* out_power, double* out_precision,
                   int* out_level, om_stats* stats);

/* Device-resident variants: the same calculations with obs_xyz, energies_eV
 * and every output (out_value, out_level, out_conv, out_e) device pointers on
 * the context's device (e.g. torch tensors' data pointers); pol6, normal, conv
 * and stats stay host pointers.  No host transfer of inputs or outputs; the
 * call still returns once the outputs are written.  No counterpart in the
 * reference (its GPU entry points take host vectors); for callers that keep
 * their grids and results in HBM. */
int om_flux_dev(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV,
                long n_e, const double* pol6, const om_conv* conv, double* out_value, int* out_level,
                unsigned char* out_conv, double* out_e, om_stats* stats);
int om_spectrum_dev(om_context* ctx, int particle, const double* obs, const double* energies_eV, long n_e,
                    const double* pol6, const om_conv* conv, double* out_value, int* out_level,
                    unsigned char* out_conv, double* out_e, om_stats* stats);
int om_power_density_dev(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* normal,
                         int directional, const om_conv* conv, double* out_value, int* out_level,
                         unsigned char* out_conv, om_stats* stats);

/* om_flux_dev without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Inputs and outputs are the
 * caller's device buffers as in om_flux_dev; the per-point state, the level
 * sample rows (rebuilt at every call), lists and counters come from a static
 * device array of 64 MiB per GPU (one copy per device, held until the process
 * ends), so no hipMalloc, hipFree or pinned host allocation happens.  A
 * request needing more scratch than 64 MiB (about 56 bytes per output point,
 * 76 in the mixed modes, plus 160 bytes x 2^L rows for the deepest level L
 * summed) fails with an error.  conv->accumulate must be 0.  The trajectory is
 * the context's current one (om_trajectories before, which allocates).  Every
 * call on a GPU shares that GPU's scratch: calls that overlap on one GPU, from
 * two contexts or threads, race. */
int om_flux_dev_noalloc(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV,
                        long n_e, const double* pol6, const om_conv* conv, double* out_value, int* out_level,
                        unsigned char* out_conv, double* out_e, om_stats* stats);

/* ---- batched multi-particle sums (one persistent launch) ----
 * Replaces the reference's multi-particle loop (src/OSCARSSR.cc:4016-4063,
 * 1778-1818 and the CUDA host pipeline src/OSCARSSR_Cuda.cu:974-1147), which
 * recomputes each particle's trajectory and sums the particles one after the
 * other.  Particle i of the resident trajectory set (om_trajectories) adds
 * weight[i] * flux to the output grid of request req_of[i]; requests may have
 * different observation grids and energy lists, particles different
 * trajectory lengths.  Levels 0..opts->max_level of every particle are stored
 * as sample rows; deeper levels are interpolated in the kernel, so every point
 * follows the conv's level rule; points unconverged at its last level are
 * counted in stats->n_unconverged.  With opts->deep_level set, points still
 * active at that level are finished by a workgroup per point (their samples
 * split over its waves).  MIXED mode: points whose stop decision lay within
 * the reduced-accuracy sin/cos error bound are recomputed with full-accuracy
 * sin/cos in the same run (stats->n_rechecked counts them), so stopping levels
 * match the FP64 kernel. */
typedef struct {
  long n_obs;
  const double* obs_xyz;     /* 3 per point */
  long n_e;
  const double* energies_eV;
} om_request;

typedef struct {
  int waves_per_cu; /* persistent grid = CUs x this; 0: the tuned value (8 before any tuning) */
  int max_level;    /* highest level held as stored sample rows (0: up to 18 within 40% of the
                       device memory); deeper levels, up to the conv's max_level rule, are
                       interpolated */
  int deep_level;   /* points still active at this level are finished one workgroup per point
                       (0: none, every point finishes in its work unit) */
} om_batch_opts;

int om_batch_setup(om_context* ctx, const om_request* req, int n_req, const int* req_of, const double* weight,
                   const double* pol6, const om_conv* conv, const om_batch_opts* opts);
/* One step: [RK4 + spline when recompute_trajectories] + level samples + all
 * sums, replayed from a captured HIP graph; out[r] receives request r's grid
 * (n_obs x n_e, energy fastest). */
int om_batch_run(om_context* ctx, int recompute_trajectories, double* const* out, om_stats* stats);
/* One step of the batch launched stage by stage (no graph), with HIP events
 * between the stages, for profiling: stage_ms[5] = ms of the trajectories
 * (RK4 + spline; 0 without recompute), the level rows, the persistent kernel,
 * the deep-level kernel and the FP64 recheck kernel; counts[9] = the device
 * counters: units taken, unconverged points, evals, handed-over points listed
 * for the FP64 recheck kernel, points handed to the deep-level kernel,
 * amplitude evaluations and integrand evaluations inside the persistent
 * kernel, points rechecked in FP64 inside the persistent kernel, and its
 * integrand lane slots (idle lanes included).  No counterpart in the
 * reference. */
int om_batch_profile(om_context* ctx, int recompute_trajectories, double* stage_ms, unsigned long long* counts,
                     om_stats* stats);
/* Times the persistent kernel on the first max(sample_items, 16 x CUs x the
 * largest candidate) work units for each candidate waves-per-CU value and
 * keeps the fastest (returned in *best). */
int om_batch_tune(om_context* ctx, const int* candidates, int n, long sample_items, int* best, double* ms_out);
/* Times the whole batch launch (persistent kernel, deep-level and recheck
 * kernels) on the full workload for each candidate hand-over level (0: none;
 * else at least 9, as om_batch_opts.deep_level) and keeps the fastest
 * (returned in *best; the times in ms_out[n] when not NULL).  No counterpart
 * in the reference. */
int om_batch_tune_deep(om_context* ctx, const int* deep_levels, int n, int* best, double* ms_out);

/* ---- reverse pass (adjoint) ------------------------------------------------
 * The reference has no derivative path; these entry points differentiate its
 * flux calculation (OSCARSSR::CalculateFlux, src/OSCARSSR.cc:3905-4068 with the
 * point loop :4075-4275) at a frozen level: every output point sums levels
 * 0..level (precision 0, so the level test never stops early).
 *
 * om_adjoint_forward: RK4 + spline for one particle (ctstart must equal T0),
 * keeping the state every ceil(sqrt(N)) steps, then flux (n_obs x n_e, energy
 * fastest) with polarization 'all'.  Grid fields must carry no rotation.
 * grad_field: index into fields of the grid whose samples are differentiated
 * (-1: none). */
int om_adjoint_forward(om_context* ctx, const om_field* fields, int nfields, int grad_field,
                       const om_particle* particle, const double* obs_xyz, long n_obs, const double* energies_eV,
                       long n_e, int level, double* out_flux, om_stats* stats);
/* Given dLoss/dflux (n_obs x n_e) for the last om_adjoint_forward: dLoss/d(grid
 * samples) (3 per sample, x y z field components; NULL to skip), dLoss/dx0 [m]
 * and dLoss/dbeta0 (at fixed gamma).  stats->evals counts the (point, energy,
 * sample) terms differentiated. */
int om_adjoint_backward(om_context* ctx, const double* dloss_dflux, double* grad_field, double* grad_x0,
                        double* grad_beta0, om_stats* stats);

/* om_adjoint_forward / om_adjoint_backward on device memory (no counterpart in
 * the reference): obs_xyz, energies_eV, out_flux, dloss_dflux and every
 * gradient are device pointers, so a caller holding its tensors in HBM (an
 * autograd function) moves nothing over PCIe.  grad_samples: the samples of
 * fields[grad_field] on the device (3 per sample, unrotated), used in place of
 * that om_field's host samples (its grid_data may be NULL); NULL: the host
 * samples.  The calls return once the results are written. */
int om_adjoint_forward_dev(om_context* ctx, const om_field* fields, int nfields, int grad_field,
                           const double* grad_samples, const om_particle* particle, const double* obs_xyz, long n_obs,
                           const double* energies_eV, long n_e, int level, double* out_flux, om_stats* stats);
int om_adjoint_backward_dev(om_context* ctx, const double* dloss_dflux, double* grad_field, double* grad_x0,
                            double* grad_beta0, om_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
