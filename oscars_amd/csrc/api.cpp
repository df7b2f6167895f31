// Host side of the C ABI (include/oscars_mi.h): device context, trajectory
// set, level-sample cache and the level-by-level driver of the LW kernels.
//
// Reference behaviour restated (paths relative to the reference root):
//   level stop rules     src/OSCARSSR.cc:1866-1868, 4108-4110, 2836-2838
//   RK4 point counts     src/OSCARSSR.cc:914-919
//   spectrum/flux/power constants src/OSCARSSR.cc:1882-1885, 4123-4126, 2957
//   undulator geometry   src/TField3D_IdealUndulator.cc:101-106, 123-125
//   EPU phase shifts     src/TField3D_IdealEPU.cc:126-150
//   grid header handling src/TField3D_Grid.cc:413-437
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/oscars_mi.h"
#include "adj_types.h"
#include "launch_types.h"
#include "om_types.h"


extern "C" {
hipError_t om_launch_level_batch(const om::LevelJobBatch* J, hipStream_t s);
hipError_t om_launch_batch_clear(double* out, long n, unsigned long long* counters, int nc, hipStream_t s);
hipError_t om_launch_lw_batch(const om::BatchArgs* A, int mixed, int n_groups, hipStream_t s);
hipError_t om_scratch_setup(int device, void** ptr, size_t* bytes);
hipError_t om_launch_lw_batch_stages(const om::BatchArgs* A, int mixed, int n_groups, int stages, hipStream_t s);
hipError_t om_launch_rk4(const om::FieldDev* F, const om::FieldDev* F0_host, int nfields, om::TrajDesc* desc,
                         int np, double* t, double* p, hipStream_t s, int nefields);
hipError_t om_launch_knot_finish(double* p, long total, hipStream_t s);
hipError_t om_launch_field_points(const om::FieldDev* F, int nf, const double* xyz, long n, double t, double* out,
                                  hipStream_t s);
hipError_t om_launch_efield_time(const double* p, long n, const double* obs, double c0, double dt, double* out_t,
                                 double* out_e, hipStream_t s);
int om_tp_blocks(long npts);
hipError_t om_launch_pd_stl(const om::StlArgs* A, hipStream_t s);
hipError_t om_launch_total_power_level(const om::TpArgs* A, hipStream_t s);
hipError_t om_launch_spline_user(const double* t, double* p, double* ypp, double* u, long n, int construct,
                                 hipStream_t s);
hipError_t om_launch_spline(const om::TrajDesc* desc, int np, const double* t, const double* p, double* ypp,
                            double* u, double* sig, double* sm1, double* pv, double* ppv, double* u2, long total,
                            hipStream_t s);
hipError_t om_launch_level(const om::LevelJob* J, hipStream_t s);
hipError_t om_launch_beta_diff(const double* rows, long npts, double* out, hipStream_t s);
hipError_t om_launch_lw_point(const om::LwArgs* A, int mixed, int split, hipStream_t s);
hipError_t om_launch_lw_tile(const om::LwArgs* A, long n_obs, int mixed, int split, hipStream_t s);
hipError_t om_launch_lw_tile_list(const om::LwArgs* A, long ntiles, int mixed, int split, hipStream_t s);
hipError_t om_launch_tile_list(const om::State* st, long n_obs, long n_omega, int La, int* list, int* ntiles,
                               hipStream_t s);
hipError_t om_launch_point_list(const om::State* st, long npts, int La, int* list, int* count, hipStream_t s);
hipError_t om_launch_border_list(const om::State* st, long npts, int* list, int* count, hipStream_t s);
hipError_t om_launch_lw_wave(const om::LwArgs* A, int mixed, hipStream_t s);
hipError_t om_launch_omega_from_ev(const double* e, long n, double* om, double qe, double hbar, hipStream_t s);
hipError_t om_launch_level_stats(const int* level, const unsigned char* conv, long n, unsigned long long* out,
                                 hipStream_t s);
hipError_t om_launch_pd_point(const om::PdArgs* A, int mixed, int split, hipStream_t s);
hipError_t om_launch_finalize(const om::FinArgs* F, hipStream_t s);
hipError_t om_launch_state_reset(const om::State* st, long npts, hipStream_t s);
hipError_t om_launch_count_active(const om::State* st, long npts, int* out, hipStream_t s);
hipError_t om_launch_adj_ckpt(const double* p_raw, long K, long nseg, double* ckpt, hipStream_t s);
hipError_t om_launch_adj_lw(const om::AdjLwArgs* A, int nchunks, hipStream_t s);
hipError_t om_launch_adj_chunk_sum(const double* partial, long ns, int nchunks, double* out, hipStream_t s);
hipError_t om_launch_adj_spline(const double* rows, long ns, const double* sb, int L, const double* X, long n,
                                int* klo, double* yb, double* yppb, double* Pw, double* ppw, double* sgw, double* tb,
                                double* cc, hipStream_t s);
hipError_t om_launch_adj_rk4(const om::AdjRkArgs* A, hipStream_t s);
hipError_t om_launch_adj_fill(double* p, long n, hipStream_t s);
hipError_t om_launch_adj_key_fill(int* k, long n, int v, hipStream_t s);
hipError_t om_launch_adj_dloss(const double* dl, const double* E, double f, double* g, long npts, hipStream_t s);
hipError_t om_launch_adj_lam_out(const double* lam, double* gx0, double* gb0, double c, hipStream_t s);
hipError_t om_adj_grad_sort_bytes(long n, int end_bit, size_t* bytes);
hipError_t om_launch_adj_grad_gather(const int* key, int* skey, int* sidx, const double* val, long n, int ngrid,
                                     int end_bit, void* temp, size_t temp_bytes, double* gbar, hipStream_t s);
}

namespace {

// include/TOSCARS.h:20-76
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2Pi = 2.5066282746310002416123552393401;
constexpr double kC = 2.99792458e8;
constexpr double kH = 6.62606876e-34;
constexpr double kHbar = 1.054571596e-34;
constexpr double kQe = 1.602176462e-19;
constexpr double kEps0 = 8.854187817E-12;
constexpr double kMu0 = 1.2566370614E-6;
inline double two_pi() { return 2.0 * kPi; }
inline double four_pi() { return 4.0 * kPi; }
inline double ev_to_omega(double e) { return e * kQe / kHbar; }

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}


// Nonzero while the allocation-free entry point runs on this thread: any
// device or pinned allocation is then a programming error and throws.
thread_local int g_no_alloc = 0;

struct NoAllocScope {
  NoAllocScope() { ++g_no_alloc; }
  ~NoAllocScope() { --g_no_alloc; }
};

// Bump allocator over the static device scratch array that the
// allocation-free entry point uses in place of the context's buffers.
struct Arena {
  unsigned char* base = nullptr;
  size_t cap = 0, used = 0;
  template <class T>
  T* take(size_t count) {
    size_t bytes = (count * sizeof(T) + 255) & ~(size_t) 255;
    if (bytes > cap - used)
      throw std::length_error("the request needs more device scratch than the allocation-free entry point has (" +
                              std::to_string(cap >> 20) + " MiB)");
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (g_no_alloc) throw std::logic_error("device allocation inside the allocation-free entry point");
    if (p) check(hipFree(p), "hipFree");
    p = nullptr;
    n = 0;
    if (count == 0) return;
    check(hipMalloc((void**) &p, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void release() {
    if (p) (void) hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Pinned host staging for small uploads that must not block the host (a copy
// from pageable memory can): the buffer is rewritten only after the copy
// queued from it has run.
struct PinnedStage {
  void* p = nullptr;
  size_t n = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
  void* get(size_t bytes) {
    if (pending) check(hipEventSynchronize(done), "staging wait");
    pending = false;
    if (g_no_alloc && (!done || bytes > n))
      throw std::logic_error("host staging allocation inside the allocation-free entry point");
    if (!done) check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "event");
    if (bytes > n) {
      if (p) check(hipHostFree(p), "hipHostFree");
      p = nullptr;
      n = 0;
      check(hipHostMalloc(&p, bytes, 0), "hipHostMalloc");
      n = bytes;
    }
    return p;
  }
  void queued(hipStream_t s) {
    check(hipEventRecord(done, s), "event");
    pending = true;
  }
  void release() {
    if (pending && done) (void) hipEventSynchronize(done);
    if (p) (void) hipHostFree(p);
    if (done) (void) hipEventDestroy(done);
    p = nullptr;
    n = 0;
    done = nullptr;
    pending = false;
  }
};

// HIP event pairs around kernel launches, read once after the stream has been
// synchronised (no wait per launch).  Events are created on first use and kept.
struct LaunchTimer {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  void begin(hipStream_t s) {
    if (ev.size() < 2 * used + 2) {
      hipEvent_t a = nullptr, b = nullptr;
      check(hipEventCreate(&a), "event");
      check(hipEventCreate(&b), "event");
      ev.push_back(a);
      ev.push_back(b);
    }
    check(hipEventRecord(ev[2 * used], s), "event");
  }
  void end(hipStream_t s) {
    check(hipEventRecord(ev[2 * used + 1], s), "event");
    ++used;
  }
  // Sum of the recorded intervals [ms]; the stream must have been synchronised.
  double take() {
    double t = 0;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "elapsed");
      t += ms;
    }
    used = 0;
    return t;
  }
  void release() {
    for (auto e : ev) (void) hipEventDestroy(e);
    ev.clear();
    used = 0;
  }
};

struct LevelCache {
  double* rows[om::kMaxLevel + 1] = {};  // device buffers, kept across trajectory sets
  bool built[om::kMaxLevel + 1] = {};    // rows hold samples of the current trajectory
  int extended[om::kMaxLevel + 1] = {};
  double beta_diff[om::kMaxLevel + 1] = {};
  bool have_bd[om::kMaxLevel + 1] = {};
};

}  // namespace

struct om_context {
  int device = 0;
  void* scratch = nullptr;  // the device's static scratch (om_flux_dev_noalloc), looked up at creation
  size_t scratch_bytes = 0;
  int split = 0;  // waves per workgroup for the LW kernels (0: chosen from the grid size)
  // device samples of a grid field whose om_field carries no host samples (om_adjoint_forward_dev)
  const double* dev_grid_samples = nullptr;
  int batch_wpc = 0;  // persistent batch kernel: waves per CU chosen by om_batch_tune (0: default)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // trajectories
  int np = 0;
  std::vector<om::TrajDesc> desc;
  DevBuf<om::TrajDesc> d_desc;
  PinnedStage stage_desc, stage_fields, stage_rb;
  LaunchTimer timer;
  DevBuf<double> d_t, d_p, d_ypp, d_u, d_sig, d_sm1, d_pv, d_ppv, d_u2;
  std::vector<double> tstart, tstop;
  std::vector<om_particle> parts;
  std::vector<LevelCache> levels;
  DevBuf<om::FieldDev> d_fields;
  std::vector<DevBuf<double>> grid_bufs;
  // work buffers
  DevBuf<double> d_obs, d_omega, d_sum, d_err, d_lastmag, d_prec, d_out, d_outcomp, d_oute;
  DevBuf<int> d_level, d_conv, d_count, d_outlevel, d_list, d_border;
  DevBuf<unsigned char> d_outconv;
  DevBuf<unsigned long long> d_stats;
  DevBuf<double> d_bd;
  DevBuf<double> d_normals, d_tp, d_tp_partial, d_et, d_stl, d_fp, d_fp_out;
  DevBuf<double> d_bvh_box;
  DevBuf<int> d_bvh_meta, d_bvh_order;
  DevBuf<om::FieldDev> d_fp_fields;
  std::vector<om::FieldDev> fields_host;
  int nfields = 0;
  int nefields = 0;  // electric fields, after the magnetic ones in d_fields
  long total_knots = 0;
  bool desc_pending = false;  // trajectories launched, status not read back yet
  bool user_traj = false;     // slot 0 holds knots given by the caller (nothing to recompute)
  // batched multi-particle state
  struct Batch {
    bool ready = false;
    int np = 0, n_req = 0, Lb = 0, Lmax = 0, stop_mem = 0, mixed = 0, groups = 0;
    long stride = 0, total_items = 0, total_out = 0;
    std::vector<long> out_off, out_n;
    DevBuf<double> rows, tspan, obs, omega, out, weight;
    DevBuf<int> req_of;
    DevBuf<long> prefix;
    DevBuf<om::ReqDev> req;
    DevBuf<unsigned long long> counters, fix;
    DevBuf<double> ws;  // per persistent wave: the state of one unit's energies
    DevBuf<int> wsi;
    DevBuf<unsigned long long> deep_code;  // points handed to the one-wave-per-point kernel
    DevBuf<double> deep_state;
    DevBuf<int> deep_level, deep_flags;
    om::BatchArgs args;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // [without, with] trajectory recomputation
    void drop_graphs() {
      for (auto& e : exec)
        if (e) {
          (void) hipGraphExecDestroy(e);
          e = nullptr;
        }
    }
    void release() {
      drop_graphs();
      rows.release(); tspan.release(); obs.release(); omega.release(); out.release(); weight.release();
      req_of.release(); prefix.release(); req.release(); counters.release(); fix.release();
      ws.release(); wsi.release();
      deep_code.release(); deep_state.release(); deep_level.release(); deep_flags.release();
      ready = false;
    }
  } batch;

  // reverse pass: checkpoints and the forward results it needs
  struct Adj {
    bool armed = false;   // the next trajectory launch keeps checkpoints
    bool ready = false;   // a forward pass is stored for om_adjoint_backward
    long K = 0, nseg = 0, nf = 0, ngrid = 0;
    int gfield = -1, L = 0;
    long n_obs = 0, n_e = 0;
    double C0 = 0, C2 = 0, dt_L = 0;
    DevBuf<double> E;  // 6 per output: C0 dt_L S (no polarization projection), on the device
    DevBuf<double> dl;  // dLoss/dflux uploaded from the host
    DevBuf<double> ckpt, rows, obs, omega, g, partial, sb, yb, yppb, Pw, ppw, sgw, ub, cc, segmap, lam, ybuf, gbar;
    DevBuf<int> klo;
    DevBuf<int> grec_key, grec_skey, grec_sidx;  // grid-gradient records and their stable sort
    DevBuf<double> grec_val;
    DevBuf<unsigned char> sort_tmp;
    void release() {
      E.release(); dl.release();
      ckpt.release(); rows.release(); obs.release(); omega.release(); g.release(); partial.release();
      sb.release(); yb.release(); yppb.release(); Pw.release(); ppw.release(); sgw.release(); ub.release();
      cc.release();
      segmap.release(); lam.release(); ybuf.release(); gbar.release(); klo.release();
      grec_key.release(); grec_skey.release(); grec_sidx.release(); grec_val.release(); sort_tmp.release();
      ready = false;
    }
  } adj;

  void free_levels() {
    for (auto& lc : levels)
      for (int L = 0; L <= om::kMaxLevel; ++L)
        if (lc.rows[L]) {
          (void) hipFree(lc.rows[L]);
          lc.rows[L] = nullptr;
        }
    levels.clear();
  }

  ~om_context() {
    batch.release();
    adj.release();
    free_levels();
    for (auto& g : grid_bufs) g.release();
    d_desc.release(); d_t.release(); d_p.release(); d_ypp.release(); d_u.release(); d_fields.release();
    stage_desc.release(); stage_fields.release(); stage_rb.release();
    timer.release();
    d_sig.release(); d_sm1.release(); d_pv.release(); d_ppv.release(); d_u2.release();
    d_obs.release(); d_omega.release(); d_sum.release(); d_err.release(); d_lastmag.release();
    d_prec.release(); d_out.release(); d_outcomp.release(); d_oute.release(); d_level.release();
    d_conv.release(); d_count.release(); d_outlevel.release(); d_outconv.release(); d_bd.release();
    d_list.release(); d_border.release();
    d_normals.release(); d_tp.release(); d_tp_partial.release(); d_et.release(); d_stl.release(); d_fp.release(); d_fp_out.release(); d_fp_fields.release();
    d_bvh_box.release(); d_bvh_meta.release(); d_bvh_order.release();
    if (ev0) (void) hipEventDestroy(ev0);
    if (ev1) (void) hipEventDestroy(ev1);
    if (stream) (void) hipStreamDestroy(stream);
  }
};

namespace {

// Device-to-host read on the context's stream through pinned memory, then a wait
// for this stream only (a copy into pageable memory can also wait for work
// queued on other streams).
void read_back(om_context* ctx, void* dst, const void* dsrc, size_t bytes) {
  if (g_no_alloc) {  // straight into the caller's host memory: nothing is allocated
    check(hipMemcpyAsync(dst, dsrc, bytes, hipMemcpyDeviceToHost, ctx->stream), "read back");
    check(hipStreamSynchronize(ctx->stream), "sync");
    return;
  }
  void* ph = ctx->stage_rb.get(bytes);
  check(hipMemcpyAsync(ph, dsrc, bytes, hipMemcpyDeviceToHost, ctx->stream), "read back");
  check(hipStreamSynchronize(ctx->stream), "sync");
  std::memcpy(dst, ph, bytes);
}


template <class F>
int guarded(om_context* ctx, F&& f) {
  try {
    check(hipSetDevice(ctx->device), "hipSetDevice");
    f();
    return 0;
  } catch (std::exception const& e) {
    ctx->err = e.what();
    return 1;
  }
}

// src/TField3D_IdealUndulator.cc:101-106, 123-125; src/TField3D_IdealEPU.cc:126-150
om::FieldDev make_field(om_context* ctx, const om_field& f) {
  om::FieldDev d;
  std::memset(&d, 0, sizeof(d));
  d.type = f.type;
  d.nperiods = f.nperiods;
  for (int k = 0; k < 3; ++k) {
    d.field[k] = f.field[k];
    d.fieldB[k] = f.fieldB[k];
    d.period[k] = f.period[k];
    d.center[k] = f.translation[k];
  }
  d.phase = f.phase;
  d.taper = f.taper;
  // time dependence (src/TField3D_IdealUndulator.cc:147-151 and the other
  // models; a grid multiplies by the factor even at zero frequency,
  // src/TField3D_Grid.cc:277, which is 1 unless frequency_phase is set)
  d.freq = f.frequency;
  d.fphase = f.frequency_phase;
  d.toff = f.time_offset;
  d.td = f.frequency != 0 || (f.type == OM_FIELD_GRID && f.frequency_phase != 0) ? 1 : 0;
  double pm = std::sqrt(f.period[0] * f.period[0] + f.period[1] * f.period[1] + f.period[2] * f.period[2]);
  d.period_length = pm;
  d.inv_period_length = 1.0 / pm;
  for (int k = 0; k < 3; ++k) d.unit[k] = f.period[k] / pm;
  d.und_length = pm * (double) (f.nperiods + 2);
  if (f.type == OM_FIELD_UNDULATOR) {
    d.phase_shift_a = f.phase * pm / two_pi();
  } else if (f.type == OM_FIELD_EPU) {
    d.phase_shift_a = f.phase / 2. * pm / two_pi();
    d.phase_shift_b = -f.phase / 2. * pm / two_pi();
  } else if (f.type == OM_FIELD_GRID) {
    int nx = (int) f.grid_n[0], ny = (int) f.grid_n[1], nz = (int) f.grid_n[2];
    if (nx < 1 || ny < 1 || nz < 1) throw std::out_of_range("invalid number of points in at least one dimension");
    int naxes = (nx > 1) + (ny > 1) + (nz > 1);
    if (naxes < 1) throw std::out_of_range("invalid dimensions");
    int nn[3] = {nx, ny, nz};
    d.axis = -1;
    d.dims = 0;
    for (int a = 0; a < 3; ++a) {
      double start, stop;
      if (f.grid_exact) {
        // src/TField3D_Grid.cc:716-745: a resampled table keeps its first and last points
        start = nn[a] > 1 ? f.grid_start[a] : 0;
        stop = nn[a] > 1 ? f.grid_stop[a] : 0;
      } else {
        // src/TField3D_Grid.cc:413-451: header start re-centred, stop from the step
        double middle = f.grid_start[a] + f.grid_step[a] * (nn[a] - 1) / 2.;
        start = middle - f.grid_step[a] * (nn[a] - 1) / 2.;
        stop = start + (double) (nn[a] - 1) * f.grid_step[a];
      }
      d.n[a] = nn[a];
      d.start[a] = start;
      d.step[a] = f.grid_step[a];
      d.inv_step[a] = 1.0 / f.grid_step[a];
      d.stop[a] = stop;
      if (nn[a] > 1) d.dims |= 1 << a;
      if (nn[a] > 1 && naxes == 1) d.axis = a;
    }
    long n = (long) nx * ny * nz;
    if (n < 2) throw std::invalid_argument("field grid needs at least two samples");
    ctx->grid_bufs.emplace_back();
    DevBuf<double>& g = ctx->grid_bufs.back();
    g.ensure(3 * n);
    if (!f.grid_data) {
      // samples already on the device (om_adjoint_forward_dev), unrotated
      if (!ctx->dev_grid_samples) throw std::invalid_argument("grid field without samples");
      if (f.rotations[0] != 0 || f.rotations[1] != 0 || f.rotations[2] != 0)
        throw std::invalid_argument("device grid samples cannot be rotated");
      check(hipMemcpyAsync(g.p, ctx->dev_grid_samples, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
            "grid copy");
    } else {
      std::vector<double> data(3 * n);
      for (long i = 0; i < n; ++i) {
        double v[3] = {f.grid_data[3 * i], f.grid_data[3 * i + 1], f.grid_data[3 * i + 2]};
        // samples are rotated by the same angles as the point (src/TField3D_Grid.cc:536-538)
        double s = std::sin(f.rotations[0]), c = std::cos(f.rotations[0]), yy = v[1];
        v[1] = c * yy - s * v[2];
        v[2] = s * yy + c * v[2];
        s = std::sin(f.rotations[1]); c = std::cos(f.rotations[1]);
        double zz = v[2];
        v[2] = c * zz - s * v[0];
        v[0] = s * zz + c * v[0];
        s = std::sin(f.rotations[2]); c = std::cos(f.rotations[2]);
        double xx = v[0];
        v[0] = c * xx - s * v[1];
        v[1] = s * xx + c * v[1];
        data[3 * i] = v[0]; data[3 * i + 1] = v[1]; data[3 * i + 2] = v[2];
      }
      check(hipMemcpy(g.p, data.data(), data.size() * sizeof(double), hipMemcpyHostToDevice), "grid upload");
    }
    d.data = g.p;
    for (int k = 0; k < 3; ++k) {
      d.rot[k] = f.rotations[k];
      d.rot_sin[k] = std::sin(f.rotations[k]);
      d.rot_cos[k] = std::cos(f.rotations[k]);
      d.trans[k] = f.translation[k];
    }
  } else if (f.type == OM_FIELD_GAUSSIAN || f.type == OM_FIELD_UNIFORM || f.type == OM_FIELD_QUADRUPOLE) {
    double v[3] = {f.field[0], f.field[1], f.field[2]};
    for (int k = 0; k < 3; ++k) {
      d.rot[k] = f.rotations[k];
      d.rot_sin[k] = std::sin(f.rotations[k]);
      d.rot_cos[k] = std::cos(f.rotations[k]);
      d.trans[k] = f.translation[k];
      d.shape[k] = f.shape[k];
    }
    // the constructors rotate the field once (src/TField3D_Gaussian.cc:44, TField3D_UniformBox.cc:66)
    double yy = v[1];
    v[1] = d.rot_cos[0] * yy - d.rot_sin[0] * v[2];
    v[2] = d.rot_sin[0] * yy + d.rot_cos[0] * v[2];
    double zz = v[2];
    v[2] = d.rot_cos[1] * zz - d.rot_sin[1] * v[0];
    v[0] = d.rot_sin[1] * zz + d.rot_cos[1] * v[0];
    double xx = v[0];
    v[0] = d.rot_cos[2] * xx - d.rot_sin[2] * v[1];
    v[1] = d.rot_sin[2] * xx + d.rot_cos[2] * v[1];
    for (int k = 0; k < 3; ++k) d.field[k] = v[k];
    d.K = f.K;
    d.width = f.width;
  } else {
    throw std::invalid_argument("unknown field type");
  }
  return d;
}

// level stop rules, src/OSCARSSR.cc:1866-1868
void level_stops(const om_conv& c, int& stop_mem, int& stop_ext) {
  stop_mem = (c.max_level >= -1 && c.max_level <= om::kMaxLevel) ? c.max_level : om::kMaxLevel;
  stop_ext = c.max_level_extended > stop_mem ? c.max_level_extended : stop_mem;
}

double dt_inclusive(double ts, double te, int L) { return (te - ts) / (std::pow(2., L + 1) - 1); }

// Make sure levels [0, Lb] of trajectory ip exist on the device with the right
// (memory / extended) spacing.
void ensure_levels(om_context* ctx, int ip, int Lb, int stop_mem, LevelCache* lc_own = nullptr,
                   Arena* ar = nullptr) {
  LevelCache& lc = lc_own ? *lc_own : ctx->levels[ip];
  const om::TrajDesc& D = ctx->desc[ip];
  double ts = ctx->tstart[ip], te = ctx->tstop[ip];
  for (int L = 0; L <= Lb; ++L) {
    int ext = L > stop_mem ? 1 : 0;
    if (lc.rows[L] && lc.built[L] && lc.extended[L] == ext) continue;
    long np = 1L << L;
    if (!lc.rows[L] && ar) lc.rows[L] = ar->take<double>((size_t) np * 10);
    if (!lc.rows[L]) {
      if (g_no_alloc) throw std::logic_error("device allocation inside the allocation-free entry point");
      check(hipMalloc((void**) &lc.rows[L], (size_t) np * 10 * sizeof(double)), "hipMalloc level");
    }
    lc.built[L] = true;
    lc.extended[L] = ext;
    lc.have_bd[L] = false;
    om::LevelJob J;
    J.t = ctx->d_t.p;
    J.p = ctx->d_p.p;
    J.ypp = ctx->d_ypp.p;
    J.knot_offset = D.knot_offset;
    J.n = D.n;
    J.tstart = ts;
    J.tstop = te;
    J.out = lc.rows[L];
    J.npts = np;
    // src/TParticleTrajectoryInterpolated.cc:138-147, 264-270 and
    // src/TParticleTrajectoryInterpolatedPoints.cc:33-37
    J.spacing = ext ? (te - ts) / (std::pow(2., L) - 1) : (te - ts) / std::pow(2., L);
    J.t_first = ts + (te - ts) / std::pow(2., L + 1);
    check(om_launch_level(&J, ctx->stream), "level kernel");
  }
}

void ensure_beta_diff(om_context* ctx, int ip, int Lb) {
  LevelCache& lc = ctx->levels[ip];
  ctx->d_bd.ensure(om::kMaxLevel + 1);
  bool any = false;
  for (int L = 0; L <= Lb; ++L) {
    if (lc.have_bd[L]) continue;
    check(om_launch_beta_diff(lc.rows[L], 1L << L, ctx->d_bd.p + L, ctx->stream), "beta diff");
    any = true;
  }
  if (!any) return;
  double h[om::kMaxLevel + 1];
  read_back(ctx, h, ctx->d_bd.p, sizeof(h));
  for (int L = 0; L <= Lb; ++L) {
    if (!lc.have_bd[L]) {
      lc.beta_diff[L] = h[L];
      lc.have_bd[L] = true;
    }
  }
}

void fill_levels(om_context* ctx, int ip, int Lb, om::Levels& lv, LevelCache* lc_own = nullptr) {
  std::memset(&lv, 0, sizeof(lv));
  LevelCache& lc = lc_own ? *lc_own : ctx->levels[ip];
  for (int L = 0; L <= om::kMaxLevel; ++L) {
    lv.rows[L] = L <= Lb ? lc.rows[L] : nullptr;
    lv.dt_incl[L] = dt_inclusive(ctx->tstart[ip], ctx->tstop[ip], L);
    lv.beta_diff[L] = lc.beta_diff[L];
  }
}

// A device buffer of n elements: from the arena when one is given, else the
// context's cached buffer b (grown as needed).
template <class T>
T* scratch_buf(DevBuf<T>& b, size_t n, Arena* ar) {
  if (ar) return ar->take<T>(n);
  b.ensure(n);
  return b.p;
}

void state_for(om_context* ctx, long npts, bool mixed, om::State& st, Arena* ar = nullptr) {
  st.sum = scratch_buf(ctx->d_sum, 6 * npts, ar);
  st.lastmag = scratch_buf(ctx->d_lastmag, npts, ar);
  st.prec = scratch_buf(ctx->d_prec, npts, ar);
  st.level = scratch_buf(ctx->d_level, npts, ar);
  st.conv = scratch_buf(ctx->d_conv, npts, ar);
  st.err = mixed ? scratch_buf(ctx->d_err, 2 * npts, ar) : nullptr;
  st.border = mixed ? scratch_buf(ctx->d_border, npts, ar) : nullptr;
  check(om_launch_state_reset(&st, npts, ctx->stream), "state reset");
}

long count_active(om_context* ctx, const om::State& st, long npts) {
  ctx->d_count.ensure(1);
  check(om_launch_count_active(&st, npts, ctx->d_count.p, ctx->stream), "count active");
  int h = 0;
  read_back(ctx, &h, ctx->d_count.p, sizeof(int));
  return h;
}

// First launch covers levels up to this one; later launches add one level
// at a time for the points still running.
constexpr int kFirstLevels = 13;
// Per-point modes: levels with at most this many active points run one wave per
// point (the per-point kernel would fill only a fraction of the GPU).
constexpr long kWaveListMax = 16384;
constexpr int kListFrom = 9;  // first level that can stop a point (src/OSCARSSR.cc:2006: L > 8)

// Waits for the last om_trajectories and reads back the per-trajectory status.
void sync_desc(om_context* ctx) {
  if (!ctx->desc_pending) return;
  read_back(ctx, ctx->desc.data(), ctx->d_desc.p, ctx->desc.size() * sizeof(om::TrajDesc));
  ctx->desc_pending = false;
}

// status 1: a step still met |beta| >= 1 after halving it om::kMaxHalving
// times (the reference recurses without limit, src/OSCARSSR.cc:1362-1378)
const char* const kHalvingLimitMsg =
    "trajectory integration met |beta| >= 1 and halving the RK4 step 40 times did not resolve it";

void check_particle(om_context* ctx, int ip) {
  if (ip < 0 || ip >= ctx->np) throw std::out_of_range("particle index out of range (no trajectory computed?)");
  sync_desc(ctx);
  if (ctx->desc[ip].status != 0) throw std::out_of_range(kHalvingLimitMsg);
}

template <class T>
void upload(om_context* ctx, DevBuf<T>& b, const T* h, size_t n) {
  b.ensure(n);
  check(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
}

// upload through pinned staging: returns without waiting for the device
template <class T>
void upload_staged(om_context* ctx, PinnedStage& st, DevBuf<T>& b, const T* h, size_t n) {
  b.ensure(n);
  void* ph = st.get(n * sizeof(T));
  std::memcpy(ph, h, n * sizeof(T));
  check(hipMemcpyAsync(b.p, ph, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  st.queued(ctx->stream);
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Finalize kernel into the caller's buffers: host pointers (copied back) or,
// with dev_io, device pointers written in place.  Statistics are reduced on
// the device.
void finish_outputs(om_context* ctx, om::FinArgs& F, long npts, double* out_value, int* out_level,
                    unsigned char* out_conv, double* out_e, om_stats* stats, const om_conv& conv, bool dev_io,
                    Arena* ar = nullptr) {
  F.out_comp = scratch_buf(ctx->d_outcomp, npts, ar);
  F.weight = conv.weight;
  F.accumulate = conv.accumulate;
  if (dev_io) {
    F.out = out_value;  // a running sum (accumulate) is already there
    F.out_e = out_e;
    F.out_level = out_level ? out_level : scratch_buf(ctx->d_outlevel, npts, ar);
    F.out_conv = out_conv ? out_conv : scratch_buf(ctx->d_outconv, npts, ar);
  } else {
    ctx->d_out.ensure(npts);
    ctx->d_outlevel.ensure(npts);
    ctx->d_outconv.ensure(npts);
    if (out_e) ctx->d_oute.ensure(6 * npts);
    F.out = ctx->d_out.p;
    F.out_e = out_e ? ctx->d_oute.p : nullptr;
    F.out_level = ctx->d_outlevel.p;
    F.out_conv = ctx->d_outconv.p;
    if (conv.accumulate) {
      // The running sum lives in the caller's buffer; the compensation term is
      // kept on the device between calls of the same size.
      check(hipMemcpyAsync(ctx->d_out.p, out_value, npts * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
            "acc upload");
    }
  }
  check(om_launch_finalize(&F, ctx->stream), "finalize");
  unsigned long long st3[3] = {0, 0, 0};
  unsigned long long* d_stats = stats ? scratch_buf(ctx->d_stats, 3, ar) : nullptr;
  if (stats) check(om_launch_level_stats(F.out_level, F.out_conv, npts, d_stats, ctx->stream), "level stats");
  if (!dev_io) {
    check(hipMemcpyAsync(out_value, ctx->d_out.p, npts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
          "out");
    if (out_level)
      check(hipMemcpyAsync(out_level, ctx->d_outlevel.p, npts * sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
            "lev");
    if (out_conv) check(hipMemcpyAsync(out_conv, ctx->d_outconv.p, npts, hipMemcpyDeviceToHost, ctx->stream), "conv");
    if (out_e)
      check(hipMemcpyAsync(out_e, ctx->d_oute.p, 6 * npts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "efield");
  }
  if (stats) read_back(ctx, st3, d_stats, sizeof(st3));  // also waits for the copies above
  else check(hipStreamSynchronize(ctx->stream), "sync");
  if (stats) {
    stats->evals = (double) st3[0];
    stats->max_level = (int) st3[1] - 1;
    stats->n_unconverged = (long) st3[2];
  }
}

// RK4 + knot finish + spline for the resident descriptors (also used inside
// graph capture by the batched path).
void launch_trajectories(om_context* ctx, int np) {
  om::FieldDev none;
  std::memset(&none, 0, sizeof(none));
  check(om_launch_rk4(ctx->d_fields.p, ctx->nfields > 0 ? &ctx->fields_host[0] : &none, ctx->nfields,
                      ctx->d_desc.p, np, ctx->d_t.p, ctx->d_p.p, ctx->stream, ctx->nefields), "rk4");
  if (ctx->adj.armed) {
    // states every K steps of particle 0, taken from the raw knots before the division by c
    auto& A = ctx->adj;
    A.nf = ctx->desc[0].nf;
    A.K = (long) std::ceil(std::sqrt((double) A.nf));
    A.nseg = (A.nf + A.K - 1) / A.K;
    A.ckpt.ensure(6 * A.nseg);
    check(om_launch_adj_ckpt(ctx->d_p.p + 9 * ctx->desc[0].knot_offset, A.K, A.nseg, A.ckpt.p, ctx->stream),
          "checkpoints");
  }
  check(om_launch_knot_finish(ctx->d_p.p, ctx->total_knots, ctx->stream), "knot finish");
  check(om_launch_spline(ctx->d_desc.p, np, ctx->d_t.p, ctx->d_p.p, ctx->d_ypp.p, ctx->d_u.p, ctx->d_sig.p,
                         ctx->d_sm1.p, ctx->d_pv.p, ctx->d_ppv.p, ctx->d_u2.p, ctx->total_knots, ctx->stream),
        "spline");
}

// ar: the allocation-free entry point's scratch (device pointers only, dev_io);
// every buffer of the call then comes from it and nothing is allocated.
int lw_sums(om_context* ctx, int ip, const double* obs_xyz, long n_obs, const double* energies_eV, long n_e,
            const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
            double* out_e, om_stats* stats, bool spectrum, bool dev_io, Arena* ar = nullptr) {
  return guarded(ctx, [&] {
    Timer wall;
    ctx->timer.used = 0;
    if (ar && !dev_io) throw std::logic_error("the scratch arena takes device-resident inputs and outputs");
    check_particle(ctx, ip);
    if (!conv) throw std::invalid_argument("conv is NULL");
    if (n_obs < 1 || n_e < 1) throw std::length_error("no points specified");
    if (conv->mode < OM_MODE_PLAIN || conv->mode > OM_MODE_POINT_MIXED) throw std::invalid_argument("unknown mode");
    long npts = n_obs * n_e;
    int stop_mem, stop_ext;
    level_stops(*conv, stop_mem, stop_ext);
    if (stop_ext < 0) throw std::out_of_range("max_level leaves no level to sum");
    if (!obs_xyz || !energies_eV || !out_value) throw std::invalid_argument("missing inputs or outputs");
    const double* d_obs = nullptr;
    double* d_omega = nullptr;
    if (ar) {  // the caller's device grid is read in place
      d_obs = obs_xyz;
      d_omega = ar->take<double>(n_e);
      check(om_launch_omega_from_ev(energies_eV, n_e, d_omega, kQe, kHbar, ctx->stream), "omega");
    } else if (dev_io) {
      ctx->d_obs.ensure(3 * n_obs);
      ctx->d_omega.ensure(n_e);
      check(hipMemcpyAsync(ctx->d_obs.p, obs_xyz, 3 * n_obs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
            "obs");
      check(om_launch_omega_from_ev(energies_eV, n_e, ctx->d_omega.p, kQe, kHbar, ctx->stream), "omega");
    } else {
      std::vector<double> om(n_e);
      for (long i = 0; i < n_e; ++i) om[i] = ev_to_omega(energies_eV[i]);
      upload(ctx, ctx->d_obs, obs_xyz, 3 * n_obs);
      upload(ctx, ctx->d_omega, om.data(), n_e);
    }
    bool mixed = conv->mode == OM_MODE_MIXED;
    bool mixed_any = mixed || conv->mode == OM_MODE_POINT_MIXED;
    om::LwArgs A;
    std::memset(&A, 0, sizeof(A));
    state_for(ctx, npts, mixed_any, A.st, ar);
    A.npts = npts;
    A.obs = ar ? d_obs : ctx->d_obs.p;
    A.omega = ar ? d_omega : ctx->d_omega.p;
    // level samples: the context's cache, or rebuilt into the arena per call
    LevelCache arena_levels;
    LevelCache* lc = ar ? &arena_levels : nullptr;
    const size_t list_n = (size_t) n_obs * ((n_e + 63) / 64) * 64;
    int* d_list = nullptr;
    int* d_count = nullptr;
    if (ar) {
      d_list = ar->take<int>(list_n);
      d_count = ar->take<int>(1);
      check(hipEventRecord(ctx->ev0, ctx->stream), "event");
    }
    A.n_omega = n_e;
    double pm2 = 0;
    for (int k = 0; k < 3; ++k) {
      A.pol[2 * k] = pol6 ? pol6[2 * k] : 0;
      A.pol[2 * k + 1] = pol6 ? pol6[2 * k + 1] : 0;
      double r = A.pol[2 * k] * A.pol[2 * k] - A.pol[2 * k + 1] * (-A.pol[2 * k + 1]);
      pm2 = k == 0 ? r : pm2 + r;
    }
    A.use_pol = pm2 > 0.001;
    A.precision = conv->precision;
    float kms = 0;
    int La = 0;
    bool tiled = !(conv->mode == OM_MODE_PLAIN || conv->mode == OM_MODE_POINT_MIXED);
    // Levels 0..8 (no point can stop before level 9) in one dense launch, then
    // one launch per level over the points still active, compacted into full
    // waves (tiles of one observation point for the tiled kernels).
    int Lb = stop_ext < kListFrom - 1 ? stop_ext : kListFrom - 1;
    // precision <= 0: no point stops before the cap (prec < precision never
    // holds), so the dense launch takes every level and nothing is compacted
    if (conv->precision <= 0 && stop_ext == stop_mem) Lb = stop_ext;
    long ntiles = 0;
    long n_recheck = 0;
    int Ldone = 0;  // deepest level summed so far (its samples exist)
    while (true) {
      ensure_levels(ctx, ip, Lb, stop_mem, lc, ar);
      fill_levels(ctx, ip, Lb, A.lv, lc);
      A.La = La;
      A.Lb = Lb;
      if (!ar) ctx->timer.begin(ctx->stream);
      if (!tiled && A.list && A.n_list <= kWaveListMax) {
        // few points left: one wave per point
        check(om_launch_lw_wave(&A, conv->mode == OM_MODE_POINT_MIXED ? 1 : 0, ctx->stream), "lw_wave");
      } else if (!tiled) {
        check(om_launch_lw_point(&A, conv->mode == OM_MODE_POINT_MIXED ? 1 : 0, ctx->split, ctx->stream),
              "lw_point");
      } else if (A.list) {
        check(om_launch_lw_tile_list(&A, ntiles, mixed ? 1 : 0, ctx->split, ctx->stream), "lw_tile");
      } else {
        check(om_launch_lw_tile(&A, n_obs, mixed ? 1 : 0, ctx->split, ctx->stream), "lw_tile");
      }
      if (!ar) ctx->timer.end(ctx->stream);
      Ldone = Lb;
      if (Lb >= stop_ext) break;
      La = Lb + 1;
      Lb = La;
      if (!ar) {
        d_list = scratch_buf(ctx->d_list, list_n, nullptr);
        d_count = scratch_buf(ctx->d_count, 1, nullptr);
      }
      if (tiled)
        check(om_launch_tile_list(&A.st, n_obs, n_e, La, d_list, d_count, ctx->stream), "tile list");
      else
        check(om_launch_point_list(&A.st, npts, La, d_list, d_count, ctx->stream), "point list");
      int nt = 0;
      read_back(ctx, &nt, d_count, sizeof(int));
      if (nt == 0) break;
      ntiles = nt;
      A.list = d_list;
      A.n_list = nt;
    }
    // Mixed path: points whose stop decision lay within the sin/cos error bound are
    // recomputed from level 0 in FP64, one wave per point (their state was
    // reset when listed), so every point stops where the FP64 path stops: one
    // launch over the levels the main pass reached, then one per deeper level.
    if (mixed_any) {
      if (!ar) {
        d_list = scratch_buf(ctx->d_list, list_n, nullptr);
        d_count = scratch_buf(ctx->d_count, 1, nullptr);
      }
      check(om_launch_border_list(&A.st, npts, d_list, d_count, ctx->stream), "border list");
      int nb = 0;
      read_back(ctx, &nb, d_count, sizeof(int));
      n_recheck = nb;
      if (nb > 0) {
        om::State st64 = A.st;
        st64.err = nullptr;
        st64.border = nullptr;
        A.st = st64;
        A.list = d_list;
        A.n_list = nb;
        La = 0;
        Lb = Ldone;
        while (true) {
          ensure_levels(ctx, ip, Lb, stop_mem, lc, ar);
          fill_levels(ctx, ip, Lb, A.lv, lc);
          A.La = La;
          A.Lb = Lb;
          if (!ar) ctx->timer.begin(ctx->stream);
          check(om_launch_lw_wave(&A, 0, ctx->stream), "lw_wave recheck");
          if (!ar) ctx->timer.end(ctx->stream);
          if (Lb >= stop_ext) break;
          La = Lb + 1;
          Lb = La;
          check(om_launch_point_list(&A.st, npts, La, d_list, d_count, ctx->stream), "point list");
          int nt = 0;
          read_back(ctx, &nt, d_count, sizeof(int));
          if (nt == 0) break;
          A.n_list = nt;
        }
      }
    }
    const om_particle& P = ctx->parts[ip];
    om::FinArgs F;
    std::memset(&F, 0, sizeof(F));
    F.st = A.st;
    F.npts = npts;
    F.kind = 0;
    F.quantity = conv->quantity;
    // src/OSCARSSR.cc:1882-1885 (spectrum: |I|) and 4123-4126 (flux: I)
    F.C0 = P.q / (four_pi() * kC * kEps0 * kSqrt2Pi);
    F.C2 = spectrum ? four_pi() * std::fabs(P.current) / (kH * std::fabs(P.q) * kMu0 * kC) * 1e-6 * 0.001
                    : four_pi() * P.current / (kH * std::fabs(P.q) * kMu0 * kC) * 1e-6 * 0.001;
    for (int L = 0; L <= om::kMaxLevel; ++L) F.dt_incl[L] = A.lv.dt_incl[L];
    std::memcpy(F.pol, A.pol, sizeof(F.pol));
    F.use_pol = A.use_pol;
    if (ar) check(hipEventRecord(ctx->ev1, ctx->stream), "event");
    finish_outputs(ctx, F, npts, out_value, out_level, out_conv, out_e, stats, *conv, dev_io, ar);
    if (ar) {  // the whole call's device time (the stream is synchronised)
      check(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1), "elapsed");
    } else {
      kms = ctx->timer.take();
    }
    if (stats) {
      stats->kernel_ms = kms;
      stats->total_ms = wall.ms();
      stats->n_rechecked = n_recheck;
    }
  });
}

}  // namespace

extern "C" {

int om_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int om_device_info(int device, char* buf, size_t len) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 1;
  std::snprintf(buf, len, "%s (%s) CUs=%d clock=%d kHz mem=%.1f GiB L2=%d KiB", p.name, p.gcnArchName,
                p.multiProcessorCount, p.clockRate, (double) p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0),
                p.l2CacheSize / 1024);
  return 0;
}

int om_context_create(int device, om_context** out) {
  if (!out) return 1;
  *out = nullptr;
  int n = om_device_count();
  if (n < 1) return 2;
  if (device < 0 || device >= n) return 3;
  om_context* c = new om_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return 4;
  }
  if (om_scratch_setup(device, &c->scratch, &c->scratch_bytes) != hipSuccess) {
    delete c;
    return 5;
  }
  *out = c;
  return 0;
}

void om_context_destroy(om_context* ctx) {
  if (!ctx) return;
  (void) hipSetDevice(ctx->device);
  delete ctx;
}

const char* om_last_error(const om_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int om_trajectories(om_context* ctx, const om_field* fields, int nfields, const om_particle* particles, int np) {
  return om_trajectories_eb(ctx, fields, nfields, nullptr, 0, particles, np);
}

int om_trajectories_eb(om_context* ctx, const om_field* fields, int nfields, const om_field* efields, int nefields,
                       const om_particle* particles, int np) {
  return guarded(ctx, [&] {
    if (nfields < 0 || nefields < 0 || nfields + nefields > om::kMaxFields) throw std::out_of_range("too many fields");
    if (nefields > 0 && ctx->adj.armed) throw std::invalid_argument("the reverse pass supports magnetic fields only");
    if (np < 1) throw std::length_error("no particles");
    if ((int) ctx->levels.size() != np) ctx->free_levels();
    for (auto& lc : ctx->levels)
      for (int L = 0; L <= om::kMaxLevel; ++L) {
        lc.built[L] = false;
        lc.have_bd[L] = false;
      }
    for (auto& g : ctx->grid_bufs) g.release();
    ctx->grid_bufs.clear();
    ctx->grid_bufs.reserve(nfields);
    std::vector<om::FieldDev> F;
    for (int i = 0; i < nfields; ++i) F.push_back(make_field(ctx, fields[i]));
    for (int i = 0; i < nefields; ++i) F.push_back(make_field(ctx, efields[i]));
    ctx->batch.drop_graphs();
    ctx->batch.ready = false;
    ctx->adj.ready = false;
    ctx->fields_host = F;
    ctx->nfields = nfields;
    ctx->nefields = nefields;
    ctx->d_fields.ensure(F.size() > 0 ? F.size() : 1);
    if (!F.empty()) upload_staged(ctx, ctx->stage_fields, ctx->d_fields, F.data(), F.size());
    ctx->np = 0;
    ctx->user_traj = false;
    ctx->desc.assign(np, om::TrajDesc());
    ctx->parts.assign(particles, particles + np);
    ctx->tstart.assign(np, 0);
    ctx->tstop.assign(np, 0);
    long total = 0;
    for (int i = 0; i < np; ++i) {
      const om_particle& P = particles[i];
      om::TrajDesc& D = ctx->desc[i];
      std::memset(&D, 0, sizeof(D));
      // src/OSCARSSR.cc:871-884 argument checks, 914-919 point counts
      if (P.ctstart > P.t0 || P.ctstop < P.t0)
        throw std::out_of_range("start time is greater than T0 or stop time is less than T0");
      if (P.ctstart >= P.ctstop)
        throw std::out_of_range("start time is greater than stop time.  check that set_ctstartstop is set");
      if (P.npoints < 2) throw std::length_error("Not enough points in this trajectory");
      double DeltaT = ((P.ctstop - P.ctstart) / kC / (double) (size_t) (P.npoints - 1));
      D.nf = (long) (size_t) (1 + (P.ctstop - P.t0) / kC / DeltaT);
      D.nb = (long) (size_t) ((P.t0 - P.ctstart) / kC / DeltaT);
      D.n = D.nf + D.nb;
      if (D.n < 3) throw std::length_error("TSpline1D3 does not see enough points for interpolation.  Please use >= 3 points");
      D.dt = DeltaT;
      D.t0_over_c = P.t0 / kC;
      D.qomg = P.qomg;
      D.qom = P.q / P.m;
      for (int k = 0; k < 3; ++k) {
        D.x0[k] = P.x0[k];
        D.b0[k] = P.beta0[k];
      }
      D.knot_offset = total;
      total += D.n;
      // trajectory time window: first and last knot times
      ctx->tstart[i] = D.nb > 0 ? D.t0_over_c + (-DeltaT) * (double) D.nb : D.t0_over_c;
      ctx->tstop[i] = D.t0_over_c + DeltaT * (double) (D.nf - 1);
    }
    ctx->total_knots = total;
    ctx->d_t.ensure(total);
    ctx->d_p.ensure(9 * total);
    ctx->d_ypp.ensure(9 * total);
    ctx->d_u.ensure(9 * total);
    ctx->d_sig.ensure(total);
    ctx->d_sm1.ensure(total);
    ctx->d_pv.ensure(total);
    ctx->d_ppv.ensure(total);
    ctx->d_u2.ensure(9 * total);
    upload_staged(ctx, ctx->stage_desc, ctx->d_desc, ctx->desc.data(), np);
    launch_trajectories(ctx, np);
    // The integration runs on; its status is read back by the first call that
    // needs it (sync_desc), so the host can queue other work meanwhile.
    ctx->desc_pending = true;
    if ((int) ctx->levels.size() != np) ctx->levels.assign(np, LevelCache());
    ctx->np = np;
  });
}

int om_trajectory_set(om_context* ctx, const om_particle* particle, long n, const double* t, const double* p9,
                      int construct) {
  return guarded(ctx, [&] {
    if (!particle || !t || !p9) throw std::invalid_argument("missing trajectory inputs");
    if (construct < 0 || construct > 3) throw std::invalid_argument("construct must be 0..3");
    // src/TParticleA.cc:358-369 and include/TOMATH.h:141-160
    if (n < 2) throw std::length_error("Not enough points in this trajectory");
    if (n < 3) throw std::length_error("TSpline1D3 does not see enough points for interpolation.  Please use >= 3 points");
    for (long i = 1; i < n; ++i)
      if (!(t[i] > t[i - 1])) throw std::invalid_argument("trajectory times must increase strictly");
    sync_desc(ctx);
    if (ctx->levels.size() != 1) ctx->free_levels();
    for (auto& lc : ctx->levels)
      for (int L = 0; L <= om::kMaxLevel; ++L) {
        lc.built[L] = false;
        lc.have_bd[L] = false;
      }
    ctx->batch.drop_graphs();
    ctx->batch.ready = false;
    ctx->adj.ready = false;
    ctx->np = 0;
    ctx->user_traj = true;
    om::TrajDesc D;
    std::memset(&D, 0, sizeof(D));
    D.n = D.nf = n;
    D.t0_over_c = t[0];
    D.qomg = particle->qomg;
    for (int k = 0; k < 3; ++k) {
      D.x0[k] = p9[k];
      D.b0[k] = p9[3 + k];
    }
    ctx->desc.assign(1, D);
    ctx->parts.assign(particle, particle + 1);
    ctx->tstart.assign(1, t[0]);
    ctx->tstop.assign(1, t[n - 1]);
    ctx->total_knots = n;
    ctx->d_t.ensure(n);
    ctx->d_p.ensure(9 * n);
    ctx->d_ypp.ensure(9 * n);
    ctx->d_u.ensure(9 * n);
    upload(ctx, ctx->d_desc, ctx->desc.data(), 1);
    upload(ctx, ctx->d_t, t, n);
    upload(ctx, ctx->d_p, p9, 9 * n);
    check(om_launch_spline_user(ctx->d_t.p, ctx->d_p.p, ctx->d_ypp.p, ctx->d_u.p, n, construct, ctx->stream),
          "user spline");
    // the host buffers are the caller's: wait before returning
    check(hipStreamSynchronize(ctx->stream), "sync");
    if (ctx->levels.size() != 1) ctx->levels.assign(1, LevelCache());
    ctx->np = 1;
  });
}

int om_efield_time(om_context* ctx, int i, const double* obs, long* n, double* out_t, double* out_e3) {
  return guarded(ctx, [&] {
    check_particle(ctx, i);
    if (!obs || !n) throw std::invalid_argument("missing inputs");
    const om::TrajDesc& D = ctx->desc[i];
    long cap = *n;
    *n = D.n;
    if (cap < D.n || !out_t || !out_e3) return;
    ctx->d_et.ensure(4 * D.n);
    // src/OSCARSSR.cc:4487: C0 = Q / (4 pi eps0)
    double c0 = ctx->parts[i].q / (four_pi() * kEps0);
    check(om_launch_efield_time(ctx->d_p.p + 9 * D.knot_offset, D.n, obs, c0, D.dt, ctx->d_et.p, ctx->d_et.p + D.n,
                                ctx->stream),
          "efield time");
    read_back(ctx, out_t, ctx->d_et.p, D.n * sizeof(double));
    read_back(ctx, out_e3, ctx->d_et.p + D.n, 3 * D.n * sizeof(double));
  });
}

int om_field_points(om_context* ctx, const om_field* fields, int nfields, const double* xyz, long n, double t,
                    double* out3) {
  return guarded(ctx, [&] {
    if (nfields < 0 || nfields > om::kMaxFields) throw std::out_of_range("too many fields");
    if (n < 0 || (n > 0 && (!xyz || !out3))) throw std::invalid_argument("missing inputs");
    if (n == 0) return;
    if (nfields == 0) {
      std::memset(out3, 0, 3 * n * sizeof(double));
      return;
    }
    // Grid samples for these fields go up after the trajectory's own and come
    // down again before returning: the trajectory's fields stay as they are.
    const size_t keep = ctx->grid_bufs.size();
    try {
      std::vector<om::FieldDev> F;
      for (int i = 0; i < nfields; ++i) F.push_back(make_field(ctx, fields[i]));
      upload_staged(ctx, ctx->stage_fields, ctx->d_fp_fields, F.data(), F.size());
      upload_staged(ctx, ctx->stage_rb, ctx->d_fp, xyz, (size_t) (3 * n));
      ctx->d_fp_out.ensure(3 * n);
      check(om_launch_field_points(ctx->d_fp_fields.p, nfields, ctx->d_fp.p, n, t, ctx->d_fp_out.p, ctx->stream),
            "field points");
      read_back(ctx, out3, ctx->d_fp_out.p, 3 * n * sizeof(double));
    } catch (...) {
      check(hipStreamSynchronize(ctx->stream), "sync");
      for (size_t k = keep; k < ctx->grid_bufs.size(); ++k) ctx->grid_bufs[k].release();
      ctx->grid_bufs.resize(keep);
      throw;
    }
    for (size_t k = keep; k < ctx->grid_bufs.size(); ++k) ctx->grid_bufs[k].release();
    ctx->grid_bufs.resize(keep);
  });
}

int om_trajectory_status(om_context* ctx, int i) {
  if (!ctx || i < 0 || i >= ctx->np) return -1;
  if (guarded(ctx, [&] { sync_desc(ctx); })) return -1;
  return ctx->desc[i].status;
}

int om_trajectory_get(om_context* ctx, int i, long* n, double* t, double* p9, double* ypp9) {
  return guarded(ctx, [&] {
    if (i < 0 || i >= ctx->np) throw std::out_of_range("trajectory index out of range");
    const om::TrajDesc& D = ctx->desc[i];
    if (!n) throw std::invalid_argument("n is NULL");
    long cap = *n;
    *n = D.n;
    if (cap < D.n) return;
    if (t)
      check(hipMemcpy(t, ctx->d_t.p + D.knot_offset, D.n * sizeof(double), hipMemcpyDeviceToHost), "t");
    if (p9)
      check(hipMemcpy(p9, ctx->d_p.p + 9 * D.knot_offset, 9 * D.n * sizeof(double), hipMemcpyDeviceToHost), "p");
    if (ypp9)
      check(hipMemcpy(ypp9, ctx->d_ypp.p + 9 * D.knot_offset, 9 * D.n * sizeof(double), hipMemcpyDeviceToHost),
            "ypp");
  });
}

int om_level_get(om_context* ctx, int i, int level, int extended, double* rows) {
  return guarded(ctx, [&] {
    check_particle(ctx, i);
    if (level < 0 || level > om::kMaxLevel) throw std::out_of_range("level out of range");
    ensure_levels(ctx, i, level, extended ? level - 1 : om::kMaxLevel);
    check(hipStreamSynchronize(ctx->stream), "sync");
    check(hipMemcpy(rows, ctx->levels[i].rows[level], (1L << level) * 10 * sizeof(double), hipMemcpyDeviceToHost),
          "level copy");
  });
}

int om_flux(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV, long n_e,
            const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
            double* out_e, om_stats* stats) {
  return lw_sums(ctx, particle, obs_xyz, n_obs, energies_eV, n_e, pol6, conv, out_value, out_level, out_conv, out_e,
                 stats, false, false);
}

int om_spectrum(om_context* ctx, int particle, const double* obs, const double* energies_eV, long n_e,
                const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
                double* out_e, om_stats* stats) {
  return lw_sums(ctx, particle, obs, 1, energies_eV, n_e, pol6, conv, out_value, out_level, out_conv, out_e, stats,
                 true, false);
}

int om_flux_dev(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV, long n_e,
                const double* pol6, const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv,
                double* out_e, om_stats* stats) {
  return lw_sums(ctx, particle, obs_xyz, n_obs, energies_eV, n_e, pol6, conv, out_value, out_level, out_conv, out_e,
                 stats, false, true);
}

int om_flux_dev_noalloc(om_context* ctx, int particle, const double* obs_xyz, long n_obs, const double* energies_eV,
                        long n_e, const double* pol6, const om_conv* conv, double* out_value, int* out_level,
                        unsigned char* out_conv, double* out_e, om_stats* stats) {
  if (conv && conv->accumulate) {
    ctx->err = "accumulate is not supported by the allocation-free entry point";
    return 1;
  }
  if (!ctx->scratch) {
    ctx->err = "no device scratch for the allocation-free entry point";
    return 1;
  }
  NoAllocScope scope;
  Arena ar{static_cast<unsigned char*>(ctx->scratch), ctx->scratch_bytes, 0};
  return lw_sums(ctx, particle, obs_xyz, n_obs, energies_eV, n_e, pol6, conv, out_value, out_level, out_conv, out_e,
                 stats, false, true, &ar);
}

int om_spectrum_dev(om_context* ctx, int particle, const double* obs, const double* energies_eV, long n_e,
                    const double* pol6, const om_conv* conv, double* out_value, int* out_level,
                    unsigned char* out_conv, double* out_e, om_stats* stats) {
  return lw_sums(ctx, particle, obs, 1, energies_eV, n_e, pol6, conv, out_value, out_level, out_conv, out_e, stats,
                 true, true);
}

namespace {
// normal: one host normal for all points; normals: one per point (host, or
// device with dev_io), used when normal is null; neither: a surface without normals.
int power_density(om_context* ctx, int ip, const double* obs_xyz, long n_obs, const double* normal, int directional,
                  const om_conv* conv, double* out_value, int* out_level, unsigned char* out_conv, om_stats* stats,
                  bool dev_io, const double* normals = nullptr) {
  return guarded(ctx, [&] {
    Timer wall;
    ctx->timer.used = 0;
    check_particle(ctx, ip);
    if (!conv) throw std::invalid_argument("conv is NULL");
    if (n_obs < 1) throw std::length_error("no points specified");
    if (!obs_xyz || !out_value) throw std::invalid_argument("missing inputs or outputs");
    int stop_mem, stop_ext;
    level_stops(*conv, stop_mem, stop_ext);
    if (stop_ext < 0) throw std::out_of_range("max_level leaves no level to sum");
    if (dev_io) {
      ctx->d_obs.ensure(3 * n_obs);
      check(hipMemcpyAsync(ctx->d_obs.p, obs_xyz, 3 * n_obs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
            "obs");
    } else {
      upload(ctx, ctx->d_obs, obs_xyz, 3 * n_obs);
    }
    // MIXED / POINT_MIXED: the reciprocal-based terms, borderline stop
    // decisions recomputed from level 0 with the FP64 terms
    bool mixed = conv->mode == OM_MODE_MIXED || conv->mode == OM_MODE_POINT_MIXED;
    om::PdArgs A;
    std::memset(&A, 0, sizeof(A));
    state_for(ctx, n_obs, mixed, A.st);
    A.npts = n_obs;
    A.obs = ctx->d_obs.p;
    A.has_normal = normal || normals ? 1 : 0;
    if (normal) {
      for (int k = 0; k < 3; ++k) A.normal[k] = normal[k];
    } else if (normals) {
      ctx->d_normals.ensure(3 * n_obs);
      check(hipMemcpyAsync(ctx->d_normals.p, normals, 3 * n_obs * sizeof(double),
                           dev_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream),
            "normals");
      A.normals = ctx->d_normals.p;
    }
    A.directional = directional;
    A.precision = conv->precision;
    A.two_over_gamma = 2. / (ctx->parts[ip].gamma);
    float kms = 0;
    // Levels 0..13 in one launch over every point, then one launch per level
    // over the points still active (compacted list); the same for the mixed
    // path's rechecked points, in FP64.
    auto level_loop = [&](int La, int Lb, bool mix) {
      int Ldone = La;
      while (true) {
        ensure_levels(ctx, ip, Lb, stop_mem);
        ensure_beta_diff(ctx, ip, Lb);
        fill_levels(ctx, ip, Lb, A.lv);
        A.La = La;
        A.Lb = Lb;
        ctx->timer.begin(ctx->stream);
        check(om_launch_pd_point(&A, mix ? 1 : 0, ctx->split, ctx->stream), "pd_point");
        ctx->timer.end(ctx->stream);
        Ldone = Lb;
        if (Lb >= stop_ext) break;
        La = Lb + 1;
        Lb = La;
        int* d_list = scratch_buf(ctx->d_list, n_obs, nullptr);
        int* d_count = scratch_buf(ctx->d_count, 1, nullptr);
        check(om_launch_point_list(&A.st, n_obs, La, d_list, d_count, ctx->stream), "point list");
        int nt = 0;
        read_back(ctx, &nt, d_count, sizeof(int));
        if (nt == 0) break;
        A.list = d_list;
        A.n_list = nt;
      }
      return Ldone;
    };
    int Ldone = level_loop(0, stop_ext < kFirstLevels ? stop_ext : kFirstLevels, mixed);
    long n_recheck = 0;
    if (mixed) {
      int* d_list = scratch_buf(ctx->d_list, n_obs, nullptr);
      int* d_count = scratch_buf(ctx->d_count, 1, nullptr);
      check(om_launch_border_list(&A.st, n_obs, d_list, d_count, ctx->stream), "border list");
      int nb = 0;
      read_back(ctx, &nb, d_count, sizeof(int));
      n_recheck = nb;
      if (nb > 0) {
        A.st.err = nullptr;
        A.st.border = nullptr;
        A.list = d_list;
        A.n_list = nb;
        level_loop(0, Ldone, false);
      }
    }
    const om_particle& P = ctx->parts[ip];
    om::FinArgs F;
    std::memset(&F, 0, sizeof(F));
    F.st = A.st;
    F.npts = n_obs;
    F.kind = 1;
    F.quantity = conv->quantity;
    // src/OSCARSSR.cc:2957
    F.pd_const = std::fabs(P.q * P.current) / (16 * (kPi * kPi) * kEps0 * kC);
    F.directional = directional;
    for (int L = 0; L <= om::kMaxLevel; ++L) F.dt_incl[L] = A.lv.dt_incl[L];
    finish_outputs(ctx, F, n_obs, out_value, out_level, out_conv, nullptr, stats, *conv, dev_io);
    kms = ctx->timer.take();  // the stream is synchronised
    if (stats) {
      stats->kernel_ms = kms;
      stats->total_ms = wall.ms();
      stats->n_rechecked = n_recheck;
    }
  });
}
}  // namespace

int om_power_density(om_context* ctx, int ip, const double* obs_xyz, long n_obs, const double* normal,
                     int directional, const om_conv* conv, double* out_value, int* out_level,
                     unsigned char* out_conv, om_stats* stats) {
  if (!normal) {
    if (ctx) ctx->err = "normal is NULL";
    return 1;
  }
  return power_density(ctx, ip, obs_xyz, n_obs, normal, directional, conv, out_value, out_level, out_conv, stats,
                       false);
}

int om_power_density_surface(om_context* ctx, int ip, const double* obs_xyz, long n_obs, const double* normals,
                             int directional, const om_conv* conv, double* out_value, int* out_level,
                             unsigned char* out_conv, om_stats* stats) {
  return power_density(ctx, ip, obs_xyz, n_obs, nullptr, directional, conv, out_value, out_level, out_conv, stats,
                       false, normals);
}

namespace {
// Bounding-volume hierarchy over STL triangles for the occlusion test of
// om_power_density_stl, in depth-first order: an inner node's first child
// follows it, and meta[0] names the node to visit when a node is missed or
// done.  Leaves hold at most kBvhLeaf triangles (listed in `order`).
//
// The boxes are inflated so that the hierarchy never drops a triangle the
// reference's test (Moller and Trumbore, src/TTriangle3D.cc:159-215) would
// accept.  That test accepts |det| >= 1e-12 m^2; its barycentric u is
// tv.p / det with |tv| <= Rmax (the largest distance from a trajectory point
// to the mesh) and |p| <= |e2|, so rounding moves u by at most
// 3 u_r Rmax |e2| / 1e-12 (v likewise with |e1|), and the accepted point by at
// most 6.6e-4 Rmax |e1| |e2| per m^2.  Each box is inflated by 4 times that,
// plus 1e-9 of its coordinates, and the segment is extended at both ends by
// the largest inflation.  The inflation costs only extra tests.
constexpr int kBvhLeaf = 4;

struct StlBvh {
  std::vector<double> box;  // 6 per node
  std::vector<int> meta;    // 3 per node
  std::vector<int> order;
  double pad_t = 0;
};

void stl_bvh_build(const double* tri, long n, double rmax, bool one_leaf, StlBvh& h) {
  std::vector<double> lo(3 * n), hi(3 * n), mid(3 * n);
  double pad_max = 0;
  for (long i = 0; i < n; ++i) {
    const double* t = tri + 12 * i;
    double e1 = std::sqrt(t[3] * t[3] + t[4] * t[4] + t[5] * t[5]);
    double e2 = std::sqrt(t[6] * t[6] + t[7] * t[7] + t[8] * t[8]);
    double pad_rel = 4 * 6.6e-4 * rmax * e1 * e2;
    for (int k = 0; k < 3; ++k) {
      double a = t[k], b = t[k] + t[3 + k], c = t[k] + t[6 + k];
      double mn = std::min(a, std::min(b, c)), mx = std::max(a, std::max(b, c));
      double pad = pad_rel + 1e-9 * std::max(std::fabs(mn), std::fabs(mx)) + 1e-12;
      pad_max = std::max(pad_max, pad);
      lo[3 * i + k] = mn - pad;
      hi[3 * i + k] = mx + pad;
      mid[3 * i + k] = 0.5 * (mn + mx);
    }
  }
  h.pad_t = pad_max + 1e-9 * rmax;
  h.order.resize(n);
  for (long i = 0; i < n; ++i) h.order[i] = (int) i;
  h.box.clear();
  h.meta.clear();
  std::function<void(long, long)> build = [&](long first, long count) {
    long node = (long) h.meta.size() / 3;
    double bl[3] = {1e300, 1e300, 1e300}, bh[3] = {-1e300, -1e300, -1e300};
    double cl[3] = {1e300, 1e300, 1e300}, ch[3] = {-1e300, -1e300, -1e300};
    for (long k = first; k < first + count; ++k) {
      int j = h.order[k];
      for (int a = 0; a < 3; ++a) {
        bl[a] = std::min(bl[a], lo[3 * j + a]);
        bh[a] = std::max(bh[a], hi[3 * j + a]);
        cl[a] = std::min(cl[a], mid[3 * j + a]);
        ch[a] = std::max(ch[a], mid[3 * j + a]);
      }
    }
    for (int a = 0; a < 3; ++a) h.box.push_back(bl[a]);
    for (int a = 0; a < 3; ++a) h.box.push_back(bh[a]);
    h.meta.push_back(0);
    h.meta.push_back((int) first);
    h.meta.push_back(0);
    if (one_leaf || count <= kBvhLeaf) {
      h.meta[3 * node + 2] = (int) count;
    } else {
      int axis = 0;
      for (int a = 1; a < 3; ++a)
        if (ch[a] - cl[a] > ch[axis] - cl[axis]) axis = a;
      long half = count / 2;
      std::nth_element(h.order.begin() + first, h.order.begin() + first + half, h.order.begin() + first + count,
                       [&](int x, int y) {
                         double cx = mid[3 * x + axis], cy = mid[3 * y + axis];
                         return cx < cy || (cx == cy && x < y);
                       });
      build(first, half);
      build(first + half, count - half);
    }
    h.meta[3 * node] = (int) (h.meta.size() / 3);  // the node after this subtree
  };
  build(0, n);
}
}  // namespace

int om_power_density_stl(om_context* ctx, int ip, const double* tri15, long n, const om_conv* conv, double* out_value,
                         int* out_level, unsigned char* out_conv, om_stats* stats) {
  return guarded(ctx, [&] {
    Timer wall;
    ctx->timer.used = 0;
    check_particle(ctx, ip);
    if (!conv || !tri15 || !out_value) throw std::invalid_argument("missing inputs or outputs");
    if (n < 1) throw std::length_error("no points specified");
    int stop_mem, stop_ext;
    level_stops(*conv, stop_mem, stop_ext);
    if (stop_ext < 0) throw std::out_of_range("max_level leaves no level to sum");
    // A, B - A, C - A and the normal per triangle; the centres apart
    std::vector<double> tri(12 * n), centre(3 * n);
    for (long i = 0; i < n; ++i) {
      const double* t = tri15 + 15 * i;
      for (int k = 0; k < 3; ++k) {
        tri[12 * i + k] = t[k];
        tri[12 * i + 3 + k] = t[3 + k] - t[k];
        tri[12 * i + 6 + k] = t[6 + k] - t[k];
        tri[12 * i + 9 + k] = t[9 + k];
        centre[3 * i + k] = t[12 + k];
      }
    }
    upload(ctx, ctx->d_stl, tri.data(), tri.size());
    upload(ctx, ctx->d_obs, centre.data(), centre.size());
    // Rmax: a bound on the distance from any trajectory point (within
    // |ctstop - ctstart| of x0, the speed being below c) to any mesh point
    const om_particle& Pp = ctx->parts[ip];
    double mlo[3] = {1e300, 1e300, 1e300}, mhi[3] = {-1e300, -1e300, -1e300};
    for (long i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k)
        for (int v = 0; v < 3; ++v) {
          double c = tri15[15 * i + 3 * v + k];
          mlo[k] = std::min(mlo[k], c);
          mhi[k] = std::max(mhi[k], c);
        }
    double far2 = 0, diag2 = 0;
    for (int k = 0; k < 3; ++k) {
      double d = std::max(std::fabs(Pp.x0[k] - mlo[k]), std::fabs(Pp.x0[k] - mhi[k]));
      far2 += d * d;
      diag2 += (mhi[k] - mlo[k]) * (mhi[k] - mlo[k]);
    }
    double rmax = std::sqrt(far2) + std::fabs(Pp.ctstop - Pp.ctstart) + std::sqrt(diag2);
    // PLAIN: one leaf, every triangle tested for every ray as in the reference
    StlBvh bvh;
    stl_bvh_build(tri.data(), n, rmax, conv->mode == OM_MODE_PLAIN, bvh);
    upload(ctx, ctx->d_bvh_box, bvh.box.data(), bvh.box.size());
    upload(ctx, ctx->d_bvh_meta, bvh.meta.data(), bvh.meta.size());
    upload(ctx, ctx->d_bvh_order, bvh.order.data(), bvh.order.size());
    om::StlArgs A;
    std::memset(&A, 0, sizeof(A));
    state_for(ctx, n, false, A.st);
    A.ntri = n;
    A.tri = ctx->d_stl.p;
    A.centre = ctx->d_obs.p;
    A.bvh_box = ctx->d_bvh_box.p;
    A.bvh_meta = ctx->d_bvh_meta.p;
    A.bvh_order = ctx->d_bvh_order.p;
    A.bvh_nodes = (int) (bvh.meta.size() / 3);
    A.pad_t = bvh.pad_t;
    A.precision = conv->precision;
    A.two_over_gamma = 2. / (ctx->parts[ip].gamma);
    int La = 0;
    int Lb = stop_ext < kFirstLevels ? stop_ext : kFirstLevels;
    while (true) {
      ensure_levels(ctx, ip, Lb, stop_mem);
      fill_levels(ctx, ip, Lb, A.lv);
      A.La = La;
      A.Lb = Lb;
      ctx->timer.begin(ctx->stream);
      check(om_launch_pd_stl(&A, ctx->stream), "pd_stl");
      ctx->timer.end(ctx->stream);
      if (Lb >= stop_ext) break;
      if (count_active(ctx, A.st, n) == 0) break;
      La = Lb + 1;
      Lb = La;
    }
    const om_particle& P = ctx->parts[ip];
    om::FinArgs F;
    std::memset(&F, 0, sizeof(F));
    F.st = A.st;
    F.npts = n;
    F.kind = 1;
    F.quantity = conv->quantity;
    // src/OSCARSSR.cc:3506
    F.pd_const = std::fabs(P.q * P.current) / (16 * (kPi * kPi) * kEps0 * kC);
    F.directional = 1;  // the terms carry |N1 . normal|: no sign to fold
    for (int L = 0; L <= om::kMaxLevel; ++L) F.dt_incl[L] = A.lv.dt_incl[L];
    finish_outputs(ctx, F, n, out_value, out_level, out_conv, nullptr, stats, *conv, false);
    double kms = ctx->timer.take();
    if (stats) {
      stats->kernel_ms = kms;
      stats->total_ms = wall.ms();
    }
  });
}

int om_total_power(om_context* ctx, int ip, const om_conv* conv, double* out_power, double* out_precision,
                   int* out_level, om_stats* stats) {
  return guarded(ctx, [&] {
    Timer wall;
    ctx->timer.used = 0;
    check_particle(ctx, ip);
    if (!conv || !out_power) throw std::invalid_argument("missing inputs or outputs");
    int stop_mem, stop_ext;
    level_stops(*conv, stop_mem, stop_ext);
    if (stop_ext < 0) throw std::out_of_range("max_level leaves no level to sum");
    const om_particle& P = ctx->parts[ip];
    ctx->d_tp.ensure(om::kMaxLevel + 1);
    // src/OSCARSSR.cc:3757-3822: one running sum over levels; levels are summed
    // in chunks on the device and the stopping rule runs here on their sums
    double Sum = 0, LastSum = -1, Result_Precision = -1;
    int Result_Level = -1, La = 0;
    double evals = 0;
    while (La <= stop_ext && Result_Level == -1) {
      int Lb = La == 0 ? (stop_ext < kFirstLevels ? stop_ext : kFirstLevels) : La;
      ensure_levels(ctx, ip, Lb, stop_mem);
      ensure_beta_diff(ctx, ip, Lb);
      ctx->d_tp_partial.ensure(om_tp_blocks(1L << Lb));
      for (int L = La; L <= Lb; ++L) {
        om::TpArgs T;
        T.rows = ctx->levels[ip].rows[L];
        T.npts = 1L << L;
        T.partial = ctx->d_tp_partial.p;
        T.out = ctx->d_tp.p + L;
        ctx->timer.begin(ctx->stream);
        check(om_launch_total_power_level(&T, ctx->stream), "total power");
        ctx->timer.end(ctx->stream);
      }
      double h[om::kMaxLevel + 1];
      read_back(ctx, h + La, ctx->d_tp.p + La, (Lb - La + 1) * sizeof(double));
      for (int L = La; L <= Lb; ++L) {
        Sum += h[L];
        evals += (double) (1L << L);
        double ThisSum = Sum * dt_inclusive(ctx->tstart[ip], ctx->tstop[ip], L);
        Result_Precision = std::fabs(ThisSum - LastSum) / LastSum;
        if (L > 8 && Result_Precision < conv->precision && ctx->levels[ip].beta_diff[L] < 2. / P.gamma) {
          Result_Level = L;
          LastSum = ThisSum;
          break;
        } else if (L > 8 && ThisSum == LastSum) {
          Result_Level = L;
          Result_Precision = 0;
          break;
        }
        LastSum = ThisSum;
      }
      La = Lb + 1;
    }
    *out_power = LastSum * std::fabs(P.q * P.current) / (6 * kPi * kEps0 * kC);
    if (out_precision) *out_precision = Result_Precision;
    if (out_level) *out_level = Result_Level;
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->evals = evals;
      stats->kernel_ms = ctx->timer.take();
      stats->total_ms = wall.ms();
      stats->max_level = Result_Level;
      stats->n_unconverged = Result_Level == -1 ? 1 : 0;
    }
  });
}

int om_power_density_dev(om_context* ctx, int ip, const double* obs_xyz, long n_obs, const double* normal,
                         int directional, const om_conv* conv, double* out_value, int* out_level,
                         unsigned char* out_conv, om_stats* stats) {
  return power_density(ctx, ip, obs_xyz, n_obs, normal, directional, conv, out_value, out_level, out_conv, stats,
                       true);
}

// batch counters: [0] next work unit, [1] unconverged, [2] evals, [3] handed-over
// points listed for the FP64 recheck kernel, [4] handed over, [5] amplitudes,
// [6] evals and [8] integrand lane slots inside the persistent kernel, [7]
// points rechecked in FP64 inside their unit
constexpr int kBatchCounters = 9;
constexpr long kBatchFixCap = 1L << 27;
constexpr long kBatchDeepCap = 1L << 22;
constexpr int kBatchNoHandOver = 99;         // default hand-over level: none (above every level)
constexpr int kBatchRowsCap = 17;            // highest level stored as rows by default ...
constexpr double kBatchRowsFraction = 0.4;   // ... within this fraction of the device memory

namespace {
// Per-wave scratch of the persistent batch kernel for a grid of `groups` waves.
void batch_scratch(om_context* ctx, long groups, om::BatchArgs& A) {
  auto& B = ctx->batch;
  B.ws.ensure((size_t) groups * om::kBatchWsTotal);
  B.wsi.ensure((size_t) groups * om::kBatchWsI * om::kBatchEB);
  A.ws = B.ws.p;
  A.wsi = B.wsi.p;
}
}  // namespace

int om_batch_setup(om_context* ctx, const om_request* req, int n_req, const int* req_of, const double* weight,
                   const double* pol6, const om_conv* conv, const om_batch_opts* opts) {
  return guarded(ctx, [&] {
    auto& B = ctx->batch;
    B.release();
    int np = ctx->np;
    if (np < 1) throw std::out_of_range("no trajectories: call om_trajectories first");
    if (n_req < 1 || !req || !req_of || !weight || !conv) throw std::invalid_argument("missing batch inputs");
    if (conv->mode != OM_MODE_TILED && conv->mode != OM_MODE_MIXED)
      throw std::invalid_argument("batched sums use the tiled kernels: mode must be TILED or MIXED");
    if (ctx->user_traj) throw std::invalid_argument("batched sums need integrated trajectories (om_trajectories)");
    sync_desc(ctx);
    for (int i = 0; i < np; ++i) {
      if (ctx->desc[i].status != 0) throw std::out_of_range(kHalvingLimitMsg);
      if (req_of[i] < 0 || req_of[i] >= n_req) throw std::out_of_range("req_of out of range");
      if (ctx->parts[i].q != ctx->parts[0].q || ctx->parts[i].current != ctx->parts[0].current)
        throw std::invalid_argument("all particles of a batch must share charge and current");
    }
    int stop_mem, stop_ext;
    level_stops(*conv, stop_mem, stop_ext);
    int cap = (opts && opts->max_level > 0) ? opts->max_level : kBatchRowsCap;
    if (cap > 20) throw std::out_of_range("batch rows above level 20 are not supported");
    if (stop_ext < 0) throw std::out_of_range("max_level leaves no level to sum");
    if (!(opts && opts->max_level > 0)) {
      size_t free_b = 0, total_b = 0;
      check(hipMemGetInfo(&free_b, &total_b), "memory info");
      const double budget = kBatchRowsFraction * (double) total_b;
      while (cap > 10 && (double) np * (double) ((2L << cap) - 1) * 80.0 > budget) --cap;
    }
    B.Lb = stop_ext < cap ? stop_ext : cap;
    B.Lmax = stop_ext;
    B.stop_mem = stop_mem;
    B.stride = (2L << B.Lb) - 1;
    B.np = np;
    B.n_req = n_req;
    // requests: observation points, angular frequencies, output regions
    std::vector<double> obs, om;
    std::vector<om::ReqDev> rd(n_req);
    std::vector<long> obs_off(n_req), om_off(n_req);
    B.out_off.assign(n_req, 0);
    B.out_n.assign(n_req, 0);
    long total_out = 0;
    for (int r = 0; r < n_req; ++r) {
      if (req[r].n_obs < 1 || req[r].n_e < 1) throw std::length_error("no points specified");
      obs_off[r] = (long) obs.size();
      om_off[r] = (long) om.size();
      for (long k = 0; k < 3 * req[r].n_obs; ++k) obs.push_back(req[r].obs_xyz[k]);
      for (long k = 0; k < req[r].n_e; ++k) om.push_back(ev_to_omega(req[r].energies_eV[k]));
      B.out_off[r] = total_out;
      B.out_n[r] = req[r].n_obs * req[r].n_e;
      total_out += B.out_n[r];
    }
    B.total_out = total_out;
    upload(ctx, B.obs, obs.data(), obs.size());
    upload(ctx, B.omega, om.data(), om.size());
    B.out.ensure(total_out);
    for (int r = 0; r < n_req; ++r) {
      rd[r].obs = B.obs.p + obs_off[r];
      rd[r].omega = B.omega.p + om_off[r];
      rd[r].out = B.out.p + B.out_off[r];
      rd[r].n_obs = req[r].n_obs;
      rd[r].n_e = req[r].n_e;
      rd[r].n_blocks_e = (req[r].n_e + om::kBatchEB - 1) / om::kBatchEB;
    }
    upload(ctx, B.req, rd.data(), n_req);
    std::vector<long> prefix(np + 1, 0);
    long total_points = 0;
    for (int i = 0; i < np; ++i) {
      const om::ReqDev& R = rd[req_of[i]];
      prefix[i + 1] = prefix[i] + R.n_obs * R.n_blocks_e;
      total_points += R.n_obs * R.n_e;
    }
    B.total_items = prefix[np];
    upload(ctx, B.prefix, prefix.data(), prefix.size());
    upload(ctx, B.req_of, req_of, np);
    upload(ctx, B.weight, weight, np);
    std::vector<double> ts(2 * np);
    for (int i = 0; i < np; ++i) {
      ts[2 * i] = ctx->tstart[i];
      ts[2 * i + 1] = ctx->tstop[i];
    }
    upload(ctx, B.tspan, ts.data(), ts.size());
    B.rows.ensure((size_t) np * B.stride * 10);
    B.counters.ensure(kBatchCounters);
    const om_particle& P = ctx->parts[0];
    om::BatchArgs& A = B.args;
    std::memset(&A, 0, sizeof(A));
    A.rows = B.rows.p;
    A.stride = B.stride;
    A.Lb = B.Lb;
    A.Lmax = B.Lmax;
    A.stop_mem = B.stop_mem;
    A.t = ctx->d_t.p;
    A.p = ctx->d_p.p;
    A.ypp = ctx->d_ypp.p;
    A.desc = ctx->d_desc.p;
    A.tspan = B.tspan.p;
    A.req_of = B.req_of.p;
    A.weight = B.weight.p;
    A.item_prefix = B.prefix.p;
    A.np = np;
    A.req = B.req.p;
    // src/OSCARSSR.cc:4123-4126
    A.C0 = P.q / (four_pi() * kC * kEps0 * kSqrt2Pi);
    A.C2 = four_pi() * P.current / (kH * std::fabs(P.q) * kMu0 * kC) * 1e-6 * 0.001;
    double pm2 = 0;
    for (int k = 0; k < 3; ++k) {
      A.pol[2 * k] = pol6 ? pol6[2 * k] : 0;
      A.pol[2 * k + 1] = pol6 ? pol6[2 * k + 1] : 0;
      double r = A.pol[2 * k] * A.pol[2 * k] - A.pol[2 * k + 1] * (-A.pol[2 * k + 1]);
      pm2 = k == 0 ? r : pm2 + r;
    }
    A.use_pol = pm2 > 0.001;
    A.precision = conv->precision;
    A.counter = B.counters.p;
    A.n_unconverged = B.counters.p + 1;
    A.evals = B.counters.p + 2;
    A.n_fix = B.counters.p + 3;
    A.n_deep = B.counters.p + 4;
    A.main_amps = B.counters.p + 5;
    A.main_evals = B.counters.p + 6;
    A.n_recheck = B.counters.p + 7;
    A.main_slots = B.counters.p + 8;
    A.Ldeep = (opts && opts->deep_level > 0) ? opts->deep_level : kBatchNoHandOver;
    if (A.Ldeep < 9) throw std::out_of_range("deep_level must be at least 9");
    A.deep_cap = total_points < kBatchDeepCap ? total_points : kBatchDeepCap;
    B.deep_code.ensure(A.deep_cap);
    B.deep_state.ensure((size_t) om::kBatchWsD * A.deep_cap);
    B.deep_level.ensure(A.deep_cap);
    B.deep_flags.ensure(A.deep_cap);
    A.deep_code = B.deep_code.p;
    A.deep_state = B.deep_state.p;
    A.deep_level = B.deep_level.p;
    A.deep_flags = B.deep_flags.p;
    A.total_items = B.total_items;
    B.mixed = conv->mode == OM_MODE_MIXED ? 1 : 0;
    if (B.mixed) {
      // room for every point up to kBatchFixCap; points flagged beyond it keep the mixed result
      A.fix_cap = total_points < kBatchFixCap ? total_points : kBatchFixCap;
      B.fix.ensure(A.fix_cap);
      A.fix_list = B.fix.p;
    }
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "attribute");
    int wpc = (opts && opts->waves_per_cu > 0) ? opts->waves_per_cu : (ctx->batch_wpc > 0 ? ctx->batch_wpc : 8);
    if (wpc > 64) throw std::out_of_range("waves per CU out of range");
    B.groups = cus * wpc;
    batch_scratch(ctx, B.groups, A);
    B.ready = true;
  });
}

namespace {
void batch_levels(om_context* ctx) {
  auto& B = ctx->batch;
  om::LevelJobBatch J;
  J.t = ctx->d_t.p;
  J.p = ctx->d_p.p;
  J.ypp = ctx->d_ypp.p;
  J.desc = ctx->d_desc.p;
  J.tspan = B.tspan.p;
  J.rows = B.rows.p;
  J.stride = B.stride;
  J.np = B.np;
  J.Lcap = B.Lb;
  J.stop_mem = B.stop_mem;
  check(om_launch_level_batch(&J, ctx->stream), "level batch");
}
}  // namespace

int om_batch_run(om_context* ctx, int recompute_trajectories, double* const* out, om_stats* stats) {
  return guarded(ctx, [&] {
    Timer wall;
    auto& B = ctx->batch;
    if (!B.ready) throw std::logic_error("om_batch_setup has not been called for the current trajectories");
    int g = recompute_trajectories ? 1 : 0;
    if (!B.exec[g]) {
      hipGraph_t graph;
      check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      if (g) launch_trajectories(ctx, B.np);
      batch_levels(ctx);
      check(om_launch_batch_clear(B.out.p, B.total_out, B.counters.p, kBatchCounters, ctx->stream), "clear");
      check(om_launch_lw_batch(&B.args, B.mixed, B.groups, ctx->stream), "lw batch");
      check(hipStreamEndCapture(ctx->stream, &graph), "end capture");
      hipError_t e = hipGraphInstantiate(&B.exec[g], graph, nullptr, nullptr, 0);
      (void) hipGraphDestroy(graph);
      check(e, "graph instantiate");
    }
    check(hipEventRecord(ctx->ev0, ctx->stream), "event");
    check(hipGraphLaunch(B.exec[g], ctx->stream), "graph launch");
    check(hipEventRecord(ctx->ev1, ctx->stream), "event");
    check(hipEventSynchronize(ctx->ev1), "sync");
    float ms = 0;
    check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "elapsed");
    unsigned long long cnt[kBatchCounters];
    read_back(ctx, cnt, B.counters.p, sizeof(cnt));
    if (out) {
      for (int r = 0; r < B.n_req; ++r)
        if (out[r])
          check(hipMemcpy(out[r], B.out.p + B.out_off[r], B.out_n[r] * sizeof(double), hipMemcpyDeviceToHost),
                "out");
    }
    if (stats) {
      stats->evals = (double) cnt[2];
      stats->kernel_ms = ms;
      stats->total_ms = wall.ms();
      stats->max_level = B.Lmax;
      stats->n_unconverged = (long) cnt[1];
      stats->n_rechecked = (long) (cnt[3] + cnt[7]);
    }
  });
}

int om_batch_profile(om_context* ctx, int recompute_trajectories, double* stage_ms, unsigned long long* counts,
                     om_stats* stats) {
  return guarded(ctx, [&] {
    Timer wall;
    auto& B = ctx->batch;
    if (!B.ready) throw std::logic_error("om_batch_setup has not been called for the current trajectories");
    if (!stage_ms || !counts) throw std::invalid_argument("missing profile outputs");
    hipEvent_t ev[6];
    for (int i = 0; i < 6; ++i) check(hipEventCreate(&ev[i]), "event");
    struct Guard {
      hipEvent_t* e;
      ~Guard() {
        for (int i = 0; i < 6; ++i) (void) hipEventDestroy(e[i]);
      }
    } guard{ev};
    // the sequence om_batch_run replays, launched one stage at a time
    check(hipEventRecord(ev[0], ctx->stream), "event");
    if (recompute_trajectories) launch_trajectories(ctx, B.np);
    check(hipEventRecord(ev[1], ctx->stream), "event");
    batch_levels(ctx);
    check(om_launch_batch_clear(B.out.p, B.total_out, B.counters.p, kBatchCounters, ctx->stream), "clear");
    check(hipEventRecord(ev[2], ctx->stream), "event");
    check(om_launch_lw_batch_stages(&B.args, B.mixed, B.groups, 1, ctx->stream), "lw batch");
    check(hipEventRecord(ev[3], ctx->stream), "event");
    check(om_launch_lw_batch_stages(&B.args, B.mixed, B.groups, 2, ctx->stream), "lw batch deep");
    check(hipEventRecord(ev[4], ctx->stream), "event");
    if (B.mixed) check(om_launch_lw_batch_stages(&B.args, B.mixed, B.groups, 4, ctx->stream), "lw batch fix");
    check(hipEventRecord(ev[5], ctx->stream), "event");
    check(hipEventSynchronize(ev[5]), "sync");
    for (int i = 0; i < 5; ++i) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "elapsed");
      stage_ms[i] = ms;
    }
    unsigned long long cnt[kBatchCounters];
    read_back(ctx, cnt, B.counters.p, sizeof(cnt));
    for (int i = 0; i < kBatchCounters; ++i) counts[i] = cnt[i];
    if (stats) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, ev[0], ev[5]), "elapsed");
      stats->evals = (double) cnt[2];
      stats->kernel_ms = ms;
      stats->total_ms = wall.ms();
      stats->max_level = B.Lmax;
      stats->n_unconverged = (long) cnt[1];
      stats->n_rechecked = (long) (cnt[3] + cnt[7]);
    }
  });
}

int om_batch_tune(om_context* ctx, const int* candidates, int n, long sample_items, int* best, double* ms_out) {
  return guarded(ctx, [&] {
    auto& B = ctx->batch;
    if (!B.ready) throw std::logic_error("om_batch_setup has not been called");
    if (n < 1) throw std::invalid_argument("no candidates");
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "attribute");
    batch_levels(ctx);
    om::BatchArgs A = B.args;
    // every wave of the largest grid takes about 16 units, so the timing is not
    // dominated by the tail of a few deep units
    int cmax = 1;
    for (int i = 0; i < n; ++i) cmax = candidates[i] > cmax ? candidates[i] : cmax;
    long sample = sample_items > 16L * cus * cmax ? sample_items : 16L * cus * cmax;
    A.total_items = sample < B.total_items ? sample : B.total_items;
    if (cmax > 64) throw std::out_of_range("waves per CU out of range");
    batch_scratch(ctx, (long) cus * cmax, A);
    B.args.ws = A.ws;
    B.args.wsi = A.wsi;
    double best_ms = 1e300;
    int best_c = candidates[0];
    for (int i = 0; i < n; ++i) {
      if (candidates[i] < 1 || candidates[i] > 64) throw std::out_of_range("waves per CU out of range");
      check(om_launch_batch_clear(B.out.p, B.total_out, B.counters.p, kBatchCounters, ctx->stream), "clear");
      check(hipEventRecord(ctx->ev0, ctx->stream), "event");
      check(om_launch_lw_batch(&A, B.mixed, cus * candidates[i], ctx->stream), "lw batch");
      check(hipEventRecord(ctx->ev1, ctx->stream), "event");
      check(hipEventSynchronize(ctx->ev1), "sync");
      float ms = 0;
      check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "elapsed");
      if (ms_out) ms_out[i] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best_c = candidates[i];
      }
    }
    ctx->batch_wpc = best_c;
    B.groups = cus * best_c;
    B.drop_graphs();
    if (best) *best = best_c;
  });
}

int om_batch_tune_deep(om_context* ctx, const int* deep_levels, int n, int* best, double* ms_out) {
  return guarded(ctx, [&] {
    auto& B = ctx->batch;
    if (!B.ready) throw std::logic_error("om_batch_setup has not been called");
    if (n < 1 || !deep_levels) throw std::invalid_argument("no candidates");
    for (int i = 0; i < n; ++i)
      if (deep_levels[i] != 0 && deep_levels[i] < 9) throw std::out_of_range("deep_level must be 0 or at least 9");
    batch_levels(ctx);
    double best_ms = 1e300;
    int best_d = deep_levels[0];
    for (int i = 0; i < n; ++i) {
      B.args.Ldeep = deep_levels[i] > 0 ? deep_levels[i] : kBatchNoHandOver;
      check(om_launch_batch_clear(B.out.p, B.total_out, B.counters.p, kBatchCounters, ctx->stream), "clear");
      check(hipEventRecord(ctx->ev0, ctx->stream), "event");
      check(om_launch_lw_batch(&B.args, B.mixed, B.groups, ctx->stream), "lw batch");
      check(hipEventRecord(ctx->ev1, ctx->stream), "event");
      check(hipEventSynchronize(ctx->ev1), "sync");
      float ms = 0;
      check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "elapsed");
      if (ms_out) ms_out[i] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best_d = deep_levels[i];
      }
    }
    B.args.Ldeep = best_d > 0 ? best_d : kBatchNoHandOver;
    B.drop_graphs();
    if (best) *best = best_d;
  });
}


namespace {
// dev_io: obs_xyz, energies_eV and out_flux are device pointers, and the grid
// field's samples come from dev_samples (its om_field holds none).
int adjoint_forward(om_context* ctx, const om_field* fields, int nfields, int grad_field, const double* dev_samples,
                    const om_particle* particle, const double* obs_xyz, long n_obs, const double* energies_eV, long n_e,
                    int level, double* out_flux, om_stats* stats, bool dev_io) {
  int rc = guarded(ctx, [&] {
    ctx->adj.ready = false;
    if (!particle || !obs_xyz || !energies_eV || !out_flux) throw std::invalid_argument("missing adjoint inputs");
    if (n_obs < 1 || n_e < 1) throw std::length_error("no points specified");
    if (level < 0 || level > 20) throw std::out_of_range("adjoint level must be in 0..20");
    if (nfields < 0 || (nfields > 0 && !fields)) throw std::invalid_argument("fields");
    if (grad_field < -1 || grad_field >= nfields) throw std::out_of_range("grad_field out of range");
    for (int i = 0; i < nfields; ++i) {
      if (fields[i].type > OM_FIELD_GRID)
        throw std::invalid_argument("the reverse pass supports undulator, EPU and grid fields only");
      if (fields[i].type == OM_FIELD_GRID &&
          (fields[i].grid_n[0] > 1) + (fields[i].grid_n[1] > 1) + (fields[i].grid_n[2] > 1) != 1)
        throw std::invalid_argument("the reverse pass supports one-dimensional grids only");
    }
    for (int i = 0; i < nfields; ++i)
      if (fields[i].type == OM_FIELD_GRID &&
          (fields[i].rotations[0] != 0 || fields[i].rotations[1] != 0 || fields[i].rotations[2] != 0))
        throw std::invalid_argument("the reverse pass supports grid fields without rotations only");
    if (grad_field >= 0 && fields[grad_field].type != OM_FIELD_GRID)
      throw std::invalid_argument("grad_field must name a grid field");
    for (int i = 0; i < nfields; ++i)
      if (fields[i].frequency != 0 || (fields[i].type == OM_FIELD_GRID && fields[i].frequency_phase != 0))
        throw std::invalid_argument("the reverse pass supports time-independent fields only");
    if (dev_samples && grad_field < 0) throw std::invalid_argument("device samples need a grad_field");
    ctx->adj.armed = true;
  });
  if (rc) return rc;
  std::vector<om_field> fl(fields, fields + nfields);
  if (dev_samples) {
    fl[grad_field].grid_data = nullptr;
    ctx->dev_grid_samples = dev_samples;
  }
  rc = om_trajectories(ctx, fl.data(), nfields, particle, 1);
  ctx->dev_grid_samples = nullptr;
  ctx->adj.armed = false;
  if (rc) return rc;
  rc = guarded(ctx, [&] {
    sync_desc(ctx);
    if (ctx->desc[0].status != 0) throw std::out_of_range(kHalvingLimitMsg);
    if (ctx->desc[0].halved != 0)
      throw std::invalid_argument("the reverse pass does not follow halved RK4 steps (|beta| >= 1 was met)");
    if (ctx->desc[0].nb != 0)
      throw std::invalid_argument("the reverse pass needs ctstart == T0 (no backward integration)");
    ctx->adj.E.ensure(6 * n_obs * n_e);
  });
  if (rc) return rc;
  om_conv c;
  std::memset(&c, 0, sizeof(c));
  c.precision = 0;  // no level test: every point sums levels 0..level
  c.max_level = level;
  c.max_level_extended = 0;
  c.quantity = 0;
  // one lane per (point, energy) for a few energies: a tile's 64 energy lanes
  // would be mostly idle
  c.mode = n_e < 64 ? OM_MODE_PLAIN : OM_MODE_TILED;
  c.weight = 1;
  c.accumulate = 0;
  if (dev_io) {
    rc = lw_sums(ctx, 0, obs_xyz, n_obs, energies_eV, n_e, nullptr, &c, out_flux, nullptr, nullptr, ctx->adj.E.p,
                 stats, false, true);
  } else {
    std::vector<double> E(6 * n_obs * n_e);
    rc = lw_sums(ctx, 0, obs_xyz, n_obs, energies_eV, n_e, nullptr, &c, out_flux, nullptr, nullptr, E.data(), stats,
                 false, false);
    if (!rc) rc = guarded(ctx, [&] { upload(ctx, ctx->adj.E, E.data(), E.size()); });
  }
  if (rc) return rc;
  return guarded(ctx, [&] {
    auto& A = ctx->adj;
    const om_particle& P = ctx->parts[0];
    A.gfield = grad_field;
    A.ngrid = 0;
    if (grad_field >= 0) {
      const om::FieldDev& f = ctx->fields_host[grad_field];
      A.ngrid = f.n[0] * f.n[1] * f.n[2];
    }
    A.L = level;
    A.n_obs = n_obs;
    A.n_e = n_e;
    A.C0 = P.q / (four_pi() * kC * kEps0 * kSqrt2Pi);
    A.C2 = four_pi() * P.current / (kH * std::fabs(P.q) * kMu0 * kC) * 1e-6 * 0.001;
    A.dt_L = dt_inclusive(ctx->tstart[0], ctx->tstop[0], level);
    A.obs.ensure(3 * n_obs);
    A.omega.ensure(n_e);
    if (dev_io) {
      check(hipMemcpyAsync(A.obs.p, obs_xyz, 3 * n_obs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream), "obs");
      check(om_launch_omega_from_ev(energies_eV, n_e, A.omega.p, kQe, kHbar, ctx->stream), "omega");
    } else {
      std::vector<double> om(n_e);
      for (long i = 0; i < n_e; ++i) om[i] = ev_to_omega(energies_eV[i]);
      upload(ctx, A.obs, obs_xyz, 3 * n_obs);
      upload(ctx, A.omega, om.data(), n_e);
    }
    A.ready = true;
  });
}
}  // namespace

int om_adjoint_forward(om_context* ctx, const om_field* fields, int nfields, int grad_field,
                       const om_particle* particle, const double* obs_xyz, long n_obs, const double* energies_eV,
                       long n_e, int level, double* out_flux, om_stats* stats) {
  return adjoint_forward(ctx, fields, nfields, grad_field, nullptr, particle, obs_xyz, n_obs, energies_eV, n_e, level,
                         out_flux, stats, false);
}

int om_adjoint_forward_dev(om_context* ctx, const om_field* fields, int nfields, int grad_field,
                           const double* grad_samples, const om_particle* particle, const double* obs_xyz, long n_obs,
                           const double* energies_eV, long n_e, int level, double* out_flux, om_stats* stats) {
  return adjoint_forward(ctx, fields, nfields, grad_field, grad_samples, particle, obs_xyz, n_obs, energies_eV, n_e,
                         level, out_flux, stats, true);
}

namespace {
// dev_io: dloss_dflux and every output are device pointers.
int adjoint_backward(om_context* ctx, const double* dloss_dflux, double* grad_field, double* grad_x0,
                     double* grad_beta0, om_stats* stats, bool dev_io) {
  return guarded(ctx, [&] {
    Timer wall;
    auto& A = ctx->adj;
    if (!A.ready) throw std::logic_error("om_adjoint_forward has not been called for the current trajectory");
    if (!dloss_dflux) throw std::invalid_argument("dloss_dflux is NULL");
    const long npts = A.n_obs * A.n_e;
    const om::TrajDesc& D = ctx->desc[0];
    // dLoss/dS from dLoss/dF: F = C2 |E|^2, E = C0 dt_L S
    const double* dl = dloss_dflux;
    if (!dev_io) {
      upload(ctx, A.dl, dloss_dflux, npts);
      dl = A.dl.p;
    }
    A.g.ensure(6 * npts);
    check(om_launch_adj_dloss(dl, A.E.p, 2.0 * A.C2 * A.C0 * A.dt_L, A.g.p, npts, ctx->stream), "dloss");
    // level samples 0..L of particle 0, concatenated
    int stop_mem = A.L;
    ensure_levels(ctx, 0, A.L, stop_mem);
    long ns = (2L << A.L) - 1;
    A.rows.ensure(10 * ns);
    for (int L = 0; L <= A.L; ++L)
      check(hipMemcpyAsync(A.rows.p + 10 * ((1L << L) - 1), ctx->levels[0].rows[L], (1L << L) * 10 * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream), "rows");
    hipEvent_t e0 = ctx->ev0, e1 = ctx->ev1;
    check(hipEventRecord(e0, ctx->stream), "event");
    // 1. LW sum: dLoss/d(sample x, beta, a/c)
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "attribute");
    long target = (long) cus * 8 * 64;  // about 8 waves per CU
    long nchunks = (target + ns - 1) / ns;
    if (nchunks > A.n_obs) nchunks = A.n_obs;
    if (nchunks < 1) nchunks = 1;
    long per = (A.n_obs + nchunks - 1) / nchunks;
    nchunks = (A.n_obs + per - 1) / per;
    A.partial.ensure(9 * ns * nchunks);
    A.sb.ensure(9 * ns);
    om::AdjLwArgs L;
    L.rows = A.rows.p;
    L.ns = ns;
    L.obs = A.obs.p;
    L.n_obs = A.n_obs;
    L.omega = A.omega.p;
    L.n_e = A.n_e;
    L.g = A.g.p;
    L.obs_per_chunk = per;
    L.partial = A.partial.p;
    check(om_launch_adj_lw(&L, (int) nchunks, ctx->stream), "adjoint LW");
    check(om_launch_adj_chunk_sum(A.partial.p, ns, (int) nchunks, A.sb.p, ctx->stream), "adjoint sum");
    // 2. spline evaluation and spline build: dLoss/d(knots)
    long n = D.n;
    A.klo.ensure(ns);
    A.yb.ensure(9 * n);
    A.yppb.ensure(9 * n);
    A.Pw.ensure(n);
    A.ppw.ensure(n);
    A.sgw.ensure(n);
    A.ub.ensure(9 * n);
    A.cc.ensure(9 * n);
    const double* X = ctx->d_t.p + D.knot_offset;
    check(om_launch_adj_spline(A.rows.p, ns, A.sb.p, A.L, X, n, A.klo.p, A.yb.p, A.yppb.p, A.Pw.p, A.ppw.p, A.sgw.p,
                               A.ub.p, A.cc.p, ctx->stream), "adjoint spline");
    // 3. RK4 with checkpoints: dLoss/d(initial state), dLoss/d(grid samples)
    A.segmap.ensure(42 * A.nseg);
    A.lam.ensure(6 * (A.nseg + 1));
    A.ybuf.ensure(6 * A.nf);
    A.gbar.ensure(A.ngrid > 0 ? 3 * A.ngrid : 3);
    // grid-sample gradients: 10 records per knot, summed per sample in record order
    const long nrec = 10 * A.nf;
    int end_bit = 1;
    while ((1L << end_bit) <= A.ngrid) ++end_bit;
    if (A.ngrid > 0) {
      A.grec_key.ensure(nrec);
      A.grec_skey.ensure(nrec);
      A.grec_sidx.ensure(nrec);
      A.grec_val.ensure(3 * nrec);
      size_t tmp = 0;
      check(om_adj_grad_sort_bytes(nrec, end_bit, &tmp), "sort size");
      A.sort_tmp.ensure(tmp > 0 ? tmp : 1);
      check(om_launch_adj_key_fill(A.grec_key.p, nrec, (int) A.ngrid, ctx->stream), "clear");
    }
    om::AdjRkArgs R;
    R.F = ctx->d_fields.p;
    R.nfields = ctx->nfields;
    R.gfield = A.gfield;
    R.grec_key = A.ngrid > 0 ? A.grec_key.p : nullptr;
    R.grec_val = A.ngrid > 0 ? A.grec_val.p : nullptr;
    R.ngrid = (int) A.ngrid;
    R.qomg = D.qomg;
    R.dt = D.dt;
    R.nf = A.nf;
    R.K = A.K;
    R.nseg = A.nseg;
    R.ckpt = A.ckpt.p;
    R.kb = A.yb.p;
    R.segmap = A.segmap.p;
    R.lam = A.lam.p;
    R.ybuf = A.ybuf.p;
    check(om_launch_adj_rk4(&R, ctx->stream), "adjoint rk4");
    if (A.ngrid > 0)
      check(om_launch_adj_grad_gather(A.grec_key.p, A.grec_skey.p, A.grec_sidx.p, A.grec_val.p, nrec, (int) A.ngrid,
                                      end_bit, A.sort_tmp.p, A.sort_tmp.n, A.gbar.p, ctx->stream),
            "grid gradient");
    check(hipEventRecord(e1, ctx->stream), "event");
    if (dev_io) {
      if (grad_field && A.ngrid > 0)
        check(hipMemcpyAsync(grad_field, A.gbar.p, 3 * A.ngrid * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream),
              "grad field");
      check(om_launch_adj_lam_out(A.lam.p, grad_x0, grad_beta0, kC, ctx->stream), "initial-state gradients");
      check(hipStreamSynchronize(ctx->stream), "sync");
    } else {
      double lam0[6];
      check(hipMemcpyAsync(lam0, A.lam.p, sizeof(lam0), hipMemcpyDeviceToHost, ctx->stream), "lambda");
      if (grad_field && A.ngrid > 0)
        check(hipMemcpyAsync(grad_field, A.gbar.p, 3 * A.ngrid * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
              "grad field");
      check(hipStreamSynchronize(ctx->stream), "sync");
      if (grad_x0) {
        grad_x0[0] = lam0[0];
        grad_x0[1] = lam0[2];
        grad_x0[2] = lam0[4];
      }
      // state velocity v = beta c
      if (grad_beta0) {
        grad_beta0[0] = lam0[1] * kC;
        grad_beta0[1] = lam0[3] * kC;
        grad_beta0[2] = lam0[5] * kC;
      }
    }
    if (stats) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      stats->evals = (double) npts * (double) ns;
      stats->kernel_ms = ms;
      stats->total_ms = wall.ms();
      stats->max_level = A.L;
      stats->n_unconverged = 0;
    }
  });
}
}  // namespace

int om_adjoint_backward(om_context* ctx, const double* dloss_dflux, double* grad_field, double* grad_x0,
                        double* grad_beta0, om_stats* stats) {
  return adjoint_backward(ctx, dloss_dflux, grad_field, grad_x0, grad_beta0, stats, false);
}

int om_adjoint_backward_dev(om_context* ctx, const double* dloss_dflux, double* grad_field, double* grad_x0,
                            double* grad_beta0, om_stats* stats) {
  return adjoint_backward(ctx, dloss_dflux, grad_field, grad_x0, grad_beta0, stats, true);
}

}  // extern "C"
