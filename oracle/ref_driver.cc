// C-ABI driver around the reference OSCARS C++ core (compiled from the
// reference sources by oracle/Makefile into oracle/_ref/libref.so).
//
// Test infrastructure only: the tests, smoke() and bench.py's cpu_baseline
// use it as the checker and as the CPU baseline.  Nothing in the product
// (oscars_amd/) links or loads it.
//
// Every entry point mirrors what the reference Python binding does for the
// same call, so the numbers are the ones a user of `oscars.sr` would get:
//   set_particle_beam   src/OSCARSSR_Python.cc:3561-3575, 3699-4070
//   add_bfield_undulator src/OSCARSSR_Python.cc:1309-1400
//   add_bfield_epu       src/OSCARSSR_Python.cc:1479-1600
//   add_bfield_file      src/OSCARSSR_Python.cc:526-700
//   calculate_spectrum   src/OSCARSSR_Python.cc:5653-5928
//   calculate_flux_rectangle          src/OSCARSSR_Python.cc:7977-8350
//   calculate_power_density_rectangle src/OSCARSSR_Python.cc:6471-6797

#include "OSCARSSR.h"
#include "TField3D_IdealUndulator.h"
#include "TField3D_IdealEPU.h"
#include "TField3D_Grid.h"
#include "TField3D_Gaussian.h"
#include "TField3D_UniformBox.h"
#include "TField3D_Quadrupole.h"
#include "TSurfacePoints_Rectangle.h"
#include "TSurfacePoints_3D.h"
#include "TSpectrumContainer.h"
#include "T3DScalarContainer.h"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {
thread_local std::string g_err;
int g_nparticles = 0;  // NParticles of the next calculate_* calls (0: the current particle)

int fail(std::exception const& e) {
  g_err = e.what();
  return 1;
}

TVector3D v3(double const* p) { return TVector3D(p[0], p[1], p[2]); }

TParticleA& current_particle(OSCARSSR* s) {
  // GetCurrentParticle() returns a const reference to the (non-const) member
  return const_cast<TParticleA&>(s->GetCurrentParticle());
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// GPU argument of the calculations below (src/OSCARSSR.cc:1713-1729): 1 sends
// them through OSCARSSR::Calculate*GPU, i.e. through the GPU boundary the
// library was linked with (oscars_amd/reference_shim/OSCARSSR_MI.cc in
// _ref/libref_mi.so; the plain build has none and throws).
static int g_gpu = 0;
void ref_set_gpu(int gpu) { g_gpu = gpu; }

// Time dependence (frequency [Hz], frequency_phase, time_offset [s]) given to
// the field models added after this call (the binding's keyword arguments).
static double g_freq = 0, g_fphase = 0, g_toff = 0;
void ref_set_time_dependence(double frequency, double frequency_phase, double time_offset) {
  g_freq = frequency;
  g_fphase = frequency_phase;
  g_toff = time_offset;
}

void* ref_new() { return new OSCARSSR(); }
void ref_delete(void* h) { delete static_cast<OSCARSSR*>(h); }

int ref_add_undulator(void* h, double const* bfield, double const* period, int nperiods,
                      double const* translation, double phase, double taper) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->AddMagneticField((TField*) new TField3D_IdealUndulator(v3(bfield), v3(period), nperiods, v3(translation),
                                                              phase, taper, g_freq, g_fphase, g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_add_epu(void* h, double const* bfieldA, double const* bfieldB, double const* period, int nperiods,
                double const* translation, double phase, double taper) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->AddMagneticField((TField*) new TField3D_IdealEPU(v3(bfieldA), v3(bfieldB), v3(period), nperiods,
                                                        v3(translation), phase, taper, g_freq, g_fphase, g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_add_grid_file(void* h, const char* path, double const* rotations, double const* translation) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->AddMagneticField(path, "OSCARS", v3(rotations), v3(translation), std::vector<double>(), g_freq, g_fphase,
                        g_toff, "");
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// add_bfield_file(ifile, iformat, rotations, translation, scale) (src/OSCARSSR_Python.cc:526-700)
int ref_add_field_file(void* h, const char* path, const char* format, double const* rotations,
                       double const* translation, double const* scaling, int nscale) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    std::vector<double> sc(scaling, scaling + nscale);
    s->AddMagneticField(path, format, v3(rotations), v3(translation), sc, g_freq, g_fphase, g_toff, "");
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// add_efield_gaussian / add_efield_uniform / add_efield_undulator (src/OSCARSSR_Python.cc:2636, 2793, 2963)
int ref_add_efield(void* h, int type, double const* field, double const* shape, double const* period, int nperiods,
                   double const* rotations, double const* translation) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    if (type == 3)
      s->AddElectricField((TField*) new TField3D_Gaussian(v3(field), v3(translation), v3(shape), v3(rotations), g_freq,
                                                         g_fphase, g_toff, ""));
    else if (type == 4)
      s->AddElectricField((TField*) new TField3D_UniformBox(v3(field), v3(shape), v3(translation), v3(rotations),
                                                           g_freq, g_fphase, g_toff, ""));
    else
      s->AddElectricField((TField*) new TField3D_IdealUndulator(v3(field), v3(period), nperiods, v3(translation), 0, 0,
                                                               g_freq, g_fphase, g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// add_bfield_interpolated(mapping, iformat, parameter, rotations, translation, scale)
int ref_add_field_interpolated(void* h, int n, double const* params, const char* const* files, const char* format,
                               double parameter, double const* rotations, double const* translation,
                               double const* scaling, int nscale) {
  try {
    std::vector<std::pair<double, std::string> > M;
    for (int i = 0; i < n; ++i) M.push_back(std::make_pair(params[i], std::string(files[i])));
    std::vector<double> sc(scaling, scaling + nscale);
    static_cast<OSCARSSR*>(h)->AddMagneticFieldInterpolated(M, format, parameter, v3(rotations), v3(translation), sc,
                                                            g_freq, g_fphase, g_toff, "", "");
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// set_particle_beam(type='electron', energy_GeV=E, x0=..., d0=..., current=I)
// add_bfield_gaussian / add_bfield_uniform / add_bfield_quadrupole
// (src/OSCARSSR_Python.cc: the constructors as the binding calls them)
int ref_add_gaussian(void* h, double const* bfield, double const* translation, double const* sigma,
                     double const* rotations) {
  try {
    static_cast<OSCARSSR*>(h)->AddMagneticField(
        (TField*) new TField3D_Gaussian(v3(bfield), v3(translation), v3(sigma), v3(rotations), g_freq, g_fphase,
                                        g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_add_uniform(void* h, double const* bfield, double const* width, double const* translation,
                    double const* rotations) {
  try {
    static_cast<OSCARSSR*>(h)->AddMagneticField(
        (TField*) new TField3D_UniformBox(v3(bfield), v3(width), v3(translation), v3(rotations), g_freq, g_fphase,
                                          g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_add_quadrupole(void* h, double K, double width, double const* rotations, double const* translation) {
  try {
    static_cast<OSCARSSR*>(h)->AddMagneticField(
        (TField*) new TField3D_Quadrupole(K, width, v3(rotations), v3(translation), g_freq, g_fphase, g_toff, ""));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_set_beam(void* h, double energy_GeV, double const* x0, double const* d0, double current) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->ClearParticleBeams();
    TVector3D Direction = v3(d0);
    TVector3D Horizontal = -Direction.Orthogonal();
    Horizontal = Horizontal.UnitVector();
    TParticleBeam* b = &(s->AddParticleBeam("electron", "", v3(x0), Direction, energy_GeV, 0, current, 1));
    b->SetHorizontalDirection(Horizontal);
    s->SetCTStartStop(-3, 3);
    s->SetNewParticle("", "ideal");
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// Replace the current particle's initial conditions (explicit multi-particle lists).
int ref_set_particle(void* h, double const* x0, double const* beta0, double t0) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TParticleA& P = current_particle(s);
    P.SetInitialParticleConditions(v3(x0), v3(beta0), t0);
    P.ResetTrajectoryData();
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// sr.set_particle_beam with the beam-spread arguments, in the binding's order
// (src/OSCARSSR_Python.cc:3873-4040).  A null pointer is an argument not given.
// add_particle_beam in place of set_particle_beam for the next
// ref_set_beam_full calls: keep the beams already there, give the new one this
// weight and leave the current particle alone (src/OSCARSSR_Python.cc:3561-3580).
static int g_append_beam = 0;
static double g_beam_weight = 1;
void ref_set_beam_append(int append, double weight) {
  g_append_beam = append;
  g_beam_weight = weight;
}

int ref_set_beam_full(void* h, double energy_GeV, double const* x0, double const* d0, double t0, double current,
                      double const* horizontal, double const* emittance, double sigma_energy_GeV,
                      double const* beta, double const* alpha, double const* gamma,
                      double const* lattice_reference, double const* eta, char const* distribution,
                      char const* predefined) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    if (!g_append_beam) s->ClearParticleBeams();
    double const W = g_append_beam ? g_beam_weight : 1;
    TVector3D Direction = d0 ? v3(d0) : TVector3D(0, 0, 1);
    TVector3D Horizontal = horizontal ? v3(horizontal) : -Direction.Orthogonal();
    Horizontal = Horizontal.UnitVector();
    TParticleBeam* b = 0x0;
    if (predefined) {
      // a predefined beam (null d0: direction not given; energy < 0: not given)
      b = &(s->AddParticleBeam(std::string(predefined), "", W));
      if (energy_GeV >= 0) b->SetE0(energy_GeV);
      if (d0) b->SetU0(Direction);
      b->SetX0(v3(x0));
    } else {
      b = &(s->AddParticleBeam("electron", "", v3(x0), Direction, energy_GeV, t0, current, W));
    }
    b->SetHorizontalDirection(Horizontal);
    if (emittance) {
      b->SetEmittance(TVector2D(emittance[0], emittance[1]));
      b->SetBeamDistribution(TParticleBeam::kBeamDistribution_Gaussian);
    } else if (!predefined) {
      b->SetBeamDistribution(TParticleBeam::kBeamDistribution_Filament);
    }
    if (distribution) b->SetBeamDistribution(std::string(distribution));
    if (sigma_energy_GeV < 0) throw std::invalid_argument("'sigma_energy_GeV' cannot be less than zero");
    if (sigma_energy_GeV > 0) b->SetSigmaEnergyGeV(sigma_energy_GeV);
    TVector2D Beta = beta ? TVector2D(beta[0], beta[1]) : TVector2D(0, 0);
    TVector2D Alpha = alpha ? TVector2D(alpha[0], alpha[1]) : TVector2D(0, 0);
    TVector2D Gamma = gamma ? TVector2D(gamma[0], gamma[1]) : TVector2D(0, 0);
    int const has = (beta ? 4 : 0) | (alpha ? 2 : 0) | (gamma ? 1 : 0);
    bool const has_ref = lattice_reference != 0;
    TVector3D L = has_ref ? v3(lattice_reference) : TVector3D(0, 0, 0);
    b->SetTwissLatticeReference(L);
    switch (has) {
      case 7: b->SetTwissParameters(Beta, Alpha, Gamma, L, has_ref); break;
      case 6: b->SetTwissBetaAlpha(Beta, Alpha, L, has_ref); break;
      case 5: b->SetTwissBetaGamma(Beta, Gamma, L, has_ref); break;
      case 3: b->SetTwissAlphaGamma(Alpha, Gamma, L, has_ref); break;
      case 4: b->SetTwissBetaAlpha(Beta, TVector2D(0, 0), L, has_ref); break;
      default: break;
    }
    b->SetEta(eta ? TVector2D(eta[0], eta[1]) : TVector2D(0, 0));
    s->SetCTStartStop(-3, 3);
    if (!g_append_beam) s->SetNewParticle("", "ideal");
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// nparticles= of the calculate_* calls that follow.
void ref_set_nparticles(int n) { g_nparticles = n; }

// set_trajectory_calculation("RKAS" / "RK4", precision) (src/OSCARSSR.cc:784-800)
int ref_set_trajectory_calculation(void* h, const char* method, double precision) {
  try {
    static_cast<OSCARSSR*>(h)->SetTrajectoryCalculation(method, precision);
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// sr.set_seed (include/OSCARSPY.h:85-87) and the draws of the global stream.
void ref_set_seed(void* h, int seed) { static_cast<OSCARSSR*>(h)->SetSeed(seed); }
double ref_random_normal(void* h) { return static_cast<OSCARSSR*>(h)->GetRandomNormal(); }
double ref_random_uniform(void* h) { return static_cast<OSCARSSR*>(h)->GetRandomUniform(); }

// One random particle of the beam as the multi-particle loops take it
// (src/OSCARSSR.cc:497-505, called per particle at src/OSCARSSR.cc:1786).
int ref_new_particle(void* h) {
  try {
    static_cast<OSCARSSR*>(h)->SetNewParticle();
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_get_particle(void* h, double* out /* x0[3] b0[3] t0 gamma q m qovermgamma current */) {
  OSCARSSR* s = static_cast<OSCARSSR*>(h);
  TParticleA const& P = s->GetCurrentParticle();
  out[0] = P.GetX0().GetX(); out[1] = P.GetX0().GetY(); out[2] = P.GetX0().GetZ();
  out[3] = P.GetB0().GetX(); out[4] = P.GetB0().GetY(); out[5] = P.GetB0().GetZ();
  out[6] = P.GetT0();
  out[7] = P.GetGamma();
  out[8] = P.GetQ();
  out[9] = P.GetM();
  out[10] = P.GetQoverMGamma();
  out[11] = P.GetCurrent();
  return 0;
}

void ref_set_ctstartstop(void* h, double a, double b) { static_cast<OSCARSSR*>(h)->SetCTStartStop(a, b); }
void ref_set_npoints_trajectory(void* h, long n) { static_cast<OSCARSSR*>(h)->SetNPointsTrajectory((size_t) n); }

int ref_get_bfield(void* h, double const* x, double t, double* out) {
  TVector3D B = static_cast<OSCARSSR*>(h)->GetB(v3(x), t);
  out[0] = B.GetX(); out[1] = B.GetY(); out[2] = B.GetZ();
  return 0;
}

void ref_remove_efield(void* h, char const* name) { static_cast<OSCARSSR*>(h)->RemoveElectricField(name); }

int ref_get_efield(void* h, double const* x, double t, double* out) {
  TVector3D E = static_cast<OSCARSSR*>(h)->GetE(v3(x), t);
  out[0] = E.GetX(); out[1] = E.GetY(); out[2] = E.GetZ();
  return 0;
}

int ref_calculate_trajectory(void* h) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->CalculateTrajectory(current_particle(s));
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

long ref_trajectory_npoints(void* h) {
  OSCARSSR* s = static_cast<OSCARSSR*>(h);
  return (long) current_particle(s).GetTrajectory().GetNPoints();
}

// out: n rows of [t, x, y, z, bx, by, bz, ax, ay, az]
int ref_trajectory(void* h, double* out) {
  OSCARSSR* s = static_cast<OSCARSSR*>(h);
  TParticleTrajectoryPoints& T = current_particle(s).GetTrajectory();
  for (size_t i = 0; i != T.GetNPoints(); ++i) {
    TParticleTrajectoryPoint const& p = T.GetPoint(i);
    double* o = out + 10 * i;
    o[0] = T.GetT(i);
    o[1] = p.GetX().GetX(); o[2] = p.GetX().GetY(); o[3] = p.GetX().GetZ();
    o[4] = p.GetB().GetX(); o[5] = p.GetB().GetY(); o[6] = p.GetB().GetZ();
    o[7] = p.GetAoverC().GetX(); o[8] = p.GetAoverC().GetY(); o[9] = p.GetAoverC().GetZ();
  }
  return 0;
}

// calculate_trajectory's ofile/bofile writers on the current trajectory
// (src/OSCARSSR_Python.cc:5217-5229); a null path writes nothing.
int ref_write_trajectory(void* h, const char* text_path, const char* bin_path, const char* format) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    if (text_path) s->WriteTrajectory(text_path, format);
    if (bin_path) s->WriteTrajectoryBinary(bin_path, format);
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// set_trajectory(ifile=...) / set_trajectory(bifile=...) for the current particle
// (src/OSCARSSR.cc:1140-1171).
int ref_read_trajectory(void* h, const char* path, const char* format, int binary) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    if (binary) s->CurrentParticleReadTrajectoryBinary(path, format);
    else s->CurrentParticleReadTrajectory(path, format);
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// set_trajectory(trajectory=[...]): n points of [t, x(3), b(3), a/c(3)]; has_b /
// has_bp say which of b and a/c the points carry (src/OSCARSSR_Python.cc:5460-5506).
int ref_set_trajectory_points(void* h, long n, double const* rows, int has_b, int has_bp) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TParticleTrajectoryPoints& T = s->GetNewTrajectory();
    for (long i = 0; i < n; ++i) {
      double const* r = rows + 10 * i;
      T.AddPoint(v3(r + 1), has_b ? v3(r + 4) : TVector3D(0, 0, 0), has_b && has_bp ? v3(r + 7) : TVector3D(0, 0, 0),
                 r[0]);
    }
    if (!has_b) T.ConstructBetaAtPoints();
    if (!has_bp) T.ConstructAoverCAtPoints();
    s->SetupTrajectoryInterpolated();
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// Spline second derivatives: n rows of 9 values
int ref_spline_ypp(void* h, double* out) {
  OSCARSSR* s = static_cast<OSCARSSR*>(h);
  TOMATH::TSpline1D3<TParticleTrajectoryPoint> const& S = current_particle(s).GetTrajectoryInterpolated().GetSpline();
  for (size_t i = 0; i != S.GetNPoints(); ++i) {
    TParticleTrajectoryPoint const& p = S.GetYPP((int) i);
    double* o = out + 9 * i;
    o[0] = p.GetX().GetX(); o[1] = p.GetX().GetY(); o[2] = p.GetX().GetZ();
    o[3] = p.GetB().GetX(); o[4] = p.GetB().GetY(); o[5] = p.GetB().GetZ();
    o[6] = p.GetAoverC().GetX(); o[7] = p.GetAoverC().GetY(); o[8] = p.GetAoverC().GetZ();
  }
  return 0;
}

// Level-L samples (2^L rows of 10 values, same layout as ref_trajectory)
int ref_level(void* h, int level, double* out) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TParticleTrajectoryPoints const& T = current_particle(s).GetTrajectoryLevel(level);
    for (size_t i = 0; i != T.GetNPoints(); ++i) {
      TParticleTrajectoryPoint const& p = T.GetPoint(i);
      double* o = out + 10 * i;
      o[0] = T.GetT(i);
      o[1] = p.GetX().GetX(); o[2] = p.GetX().GetY(); o[3] = p.GetX().GetZ();
      o[4] = p.GetB().GetX(); o[5] = p.GetB().GetY(); o[6] = p.GetB().GetZ();
      o[7] = p.GetAoverC().GetX(); o[8] = p.GetAoverC().GetY(); o[9] = p.GetAoverC().GetZ();
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

static int make_pol(const char* pol, double angle, std::string& Polarization) {
  Polarization = "all";
  if (std::strlen(pol) != 0) {
    Polarization = pol;
  } else if (angle != 112188979912321) {
    Polarization = "linear";
  }
  return 0;
}

// calculate_spectrum(obs, energy_range_eV=[e0,e1], npoints=n | energy_points_eV=list)
int ref_spectrum(void* h, double const* obs, int npoints, double e0, double e1, double const* energies,
                 const char* pol, double angle, double const* hdir, double const* pdir,
                 int nthreads, double precision, int max_level, int max_level_extended, int quantity,
                 double* out_energy, double* out_value, int* out_all_converged) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TSpectrumContainer C;
    if (energies) {
      C.Init(std::vector<double>(energies, energies + npoints));
    } else {
      C.Init((size_t) npoints, e0, e1);
    }
    std::string P;
    make_pol(pol, angle, P);
    s->CalculateSpectrum(v3(obs), C, P, angle, v3(hdir), v3(pdir), g_nparticles, nthreads, g_gpu, -1, std::vector<int>(),
                         precision, max_level, max_level_extended, quantity);
    if (out_all_converged) *out_all_converged = C.AllConverged() ? 1 : 0;
    for (size_t i = 0; i != C.GetNPoints(); ++i) {
      out_energy[i] = C.GetEnergy(i);
      out_value[i] = C.GetFlux(i);
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// Corners x0, x1, x2 (9 values) for the next rectangles instead of plane and
// widths, or none (src/OSCARSSR_Python.cc:6633-6660): rotated, then translated.
static std::vector<double> g_corners;
void ref_set_corners(double const* x9) {
  g_corners.assign(x9 ? x9 : (double const*) 0, x9 ? x9 + 9 : (double const*) 0);
}

static void make_rect(TSurfacePoints_Rectangle& R, const char* plane, int n1, int n2, double w1, double w2,
                      double const* rotations, double const* translation, int normal) {
  if (g_corners.size() == 9) {
    TVector3D X[3];
    for (int i = 0; i < 3; ++i) {
      X[i] = v3(g_corners.data() + 3 * i);
      X[i].RotateSelfXYZ(v3(rotations));
      X[i] += v3(translation);
    }
    R.Init(n1, n2, X[0], X[1], X[2], normal);
    return;
  }
  R.Init(plane, n1, n2, w1, w2, v3(rotations), v3(translation), normal);
}

int ref_flux_rectangle(void* h, double energy_eV, const char* plane, int n1, int n2, double w1, double w2,
                       double const* rotations, double const* translation, int normal, int dim,
                       const char* pol, double angle, double const* hdir, double const* pdir,
                       int nthreads, double precision, int max_level, int max_level_extended, int quantity,
                       double* out_xyz, double* out_value, int* out_all_converged) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TSurfacePoints_Rectangle R;
    make_rect(R, plane, n1, n2, w1, w2, rotations, translation, normal);
    T3DScalarContainer C;
    std::string P;
    make_pol(pol, angle, P);
    s->CalculateFlux(R, energy_eV, C, P, angle, v3(hdir), v3(pdir), g_nparticles, nthreads, g_gpu, -1, std::vector<int>(),
                     precision, max_level, max_level_extended, dim, quantity);
    if (out_all_converged) *out_all_converged = C.AllConverged() ? 1 : 0;
    for (size_t i = 0; i != C.GetNPoints(); ++i) {
      T3DScalar const& p = C.GetPoint(i);
      out_xyz[3 * i + 0] = p.GetX().GetX();
      out_xyz[3 * i + 1] = p.GetX().GetY();
      out_xyz[3 * i + 2] = p.GetX().GetZ();
      out_value[i] = p.GetV();
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

int ref_power_density_rectangle(void* h, const char* plane, int n1, int n2, double w1, double w2,
                                double const* rotations, double const* translation, int normal, int dim,
                                int nthreads, double precision, int max_level, int max_level_extended,
                                int quantity, double* out_xyz, double* out_value, int* out_all_converged) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TSurfacePoints_Rectangle R;
    make_rect(R, plane, n1, n2, w1, w2, rotations, translation, normal);
    T3DScalarContainer C;
    bool const Directional = normal == 0 ? false : true;
    s->CalculatePowerDensity(R, C, dim, Directional, precision, max_level, max_level_extended, g_nparticles, nthreads,
                             g_gpu, -1, std::vector<int>(), quantity);
    if (out_all_converged) *out_all_converged = C.AllConverged() ? 1 : 0;
    for (size_t i = 0; i != C.GetNPoints(); ++i) {
      T3DScalar const& p = C.GetPoint(i);
      out_xyz[3 * i + 0] = p.GetX().GetX();
      out_xyz[3 * i + 1] = p.GetX().GetY();
      out_xyz[3 * i + 2] = p.GetX().GetZ();
      out_value[i] = p.GetV();
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

static int pd_surface(OSCARSSR* s, TSurfacePoints_3D const& S, int dim, bool directional, int nthreads,
                      double precision, int max_level, int max_level_extended, int quantity, double* out_xyz,
                      double* out_value, int* out_all_converged) {
  T3DScalarContainer C;
  s->CalculatePowerDensity(S, C, dim, directional, precision, max_level, max_level_extended, g_nparticles, nthreads, 0,
                           -1, std::vector<int>(), quantity);
  if (out_all_converged) *out_all_converged = C.AllConverged() ? 1 : 0;
  for (size_t i = 0; i != C.GetNPoints(); ++i) {
    T3DScalar const& p = C.GetPoint(i);
    out_xyz[3 * i + 0] = p.GetX().GetX();
    out_xyz[3 * i + 1] = p.GetX().GetY();
    out_xyz[3 * i + 2] = p.GetX().GetZ();
    out_value[i] = p.GetV();
  }
  return 0;
}

// calculate_power_density(points=[[x, n], ...], normal, rotations, translation)
// (src/OSCARSSR_Python.cc:6105-6380): points xn[6 n] (position, normal).
int ref_power_density_points(void* h, long n, double const* xn, int normal, double const* rotations,
                             double const* translation, int nthreads, double precision, int max_level,
                             int max_level_extended, int quantity, double* out_xyz, double* out_value,
                             int* out_all_converged) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TSurfacePoints_3D S;
    for (long i = 0; i < n; ++i) {
      TVector3D X = v3(xn + 6 * i), N = v3(xn + 6 * i + 3);
      X.RotateSelfXYZ(v3(rotations));
      N.RotateSelfXYZ(v3(rotations));
      if (normal == -1) N *= -1;
      X += v3(translation);
      S.AddPoint(X, N);
    }
    return pd_surface(s, S, 3, normal != 0, nthreads, precision, max_level, max_level_extended, quantity, out_xyz,
                      out_value, out_all_converged);
  } catch (std::exception const& e) { return fail(e); }
}

// calculate_power_density_line(x1, x2, npoints, normal, dim) (src/OSCARSSR_Python.cc:7355-7560)
int ref_power_density_line(void* h, double const* x1, double const* x2, int npoints, int normal, int dim,
                           int nthreads, double precision, int max_level, int max_level_extended, double* out_xyz,
                           double* out_value, int* out_all_converged) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    TVector3D const Step = (v3(x2) - v3(x1)) / (npoints - 1);
    TSurfacePoints_3D S;
    for (int i = 0; i != npoints; ++i) S.AddPoint(v3(x1) + i * Step);
    return pd_surface(s, S, dim, normal != 0, nthreads, precision, max_level, max_level_extended, 0, out_xyz,
                      out_value, out_all_converged);
  } catch (std::exception const& e) { return fail(e); }
}

// add_stl (src/OSCARSSR_Python.cc:6852-6927)
int ref_add_stl(void* h, const char* path, double scale, double const* rotations, double const* translation,
                const char* name) {
  try {
    static_cast<OSCARSSR*>(h)->AddSTLFile(path, scale, v3(rotations), v3(translation), name);
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

void ref_clear_stl(void* h) { static_cast<OSCARSSR*>(h)->ClearSTL(); }

// calculate_power_density_stl (src/OSCARSSR_Python.cc:6993-7213): *n in = capacity,
// out = triangle count; out_tri[9 k] = A, B, C; out_value[k].  stl_out: stlofile.
int ref_power_density_stl(void* h, double precision, int max_level, int max_level_extended, int quantity,
                          int nthreads, const char* stl_out, long* n, double* out_tri, double* out_value) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    s->CalculatePowerDensitySTL(TVector3D(0, 0, 0), precision, max_level, max_level_extended, g_nparticles, nthreads,
                                0, -1, std::vector<int>(), quantity);
    TSTLContainer const& S = s->GetSTLContainer();
    if (stl_out) S.WriteSTLFile(stl_out);
    long cap = *n;
    *n = (long) S.GetNPoints();
    if (cap < *n) return 0;
    for (long i = 0; i < *n; ++i) {
      TTriangle3D T = S.GetPoint(i);
      for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) out_tri[9 * i + 3 * v + k] = T[v][k];
      out_value[i] = T.GetValue();
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// calculate_efield_vs_time (src/OSCARSSR_Python.cc:9048-9110): *n in = capacity,
// out = count; out_t[k], out_e[3 k].
int ref_efield_time(void* h, double const* obs, long* n, double* out_t, double* out_e) {
  try {
    OSCARSSR* s = static_cast<OSCARSSR*>(h);
    T3DScalarContainer XYZT;
    s->CalculateElectricFieldTimeDomain(v3(obs), XYZT);
    long cap = *n;
    *n = (long) XYZT.GetNPoints();
    if (cap < *n) return 0;
    for (long i = 0; i < *n; ++i) {
      T3DScalar const& p = XYZT.GetPoint(i);
      out_t[i] = p.GetV();
      out_e[3 * i + 0] = p.GetX().GetX();
      out_e[3 * i + 1] = p.GetX().GetY();
      out_e[3 * i + 2] = p.GetX().GetZ();
    }
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// calculate_total_power (src/OSCARSSR_Python.cc:5974-6060)
int ref_total_power(void* h, double precision, int max_level, int max_level_extended, int quantity, double* out) {
  try {
    *out = static_cast<OSCARSSR*>(h)->CalculateTotalPower(precision, max_level, max_level_extended, quantity);
  } catch (std::exception const& e) { return fail(e); }
  return 0;
}

// ---- the reference's own output containers: writers and file averaging
// (src/TSpectrumContainer.cc, src/T3DScalarContainer.cc), driven from given
// values so their files can be compared byte for byte.

int ref_write_spectrum(const char* text_path, const char* bin_path, long n, double const* e, double const* v) {
  try {
    TSpectrumContainer S;
    for (long i = 0; i != n; ++i) S.AddPoint(e[i], v[i]);
    if (text_path && text_path[0]) S.WriteToFileText(text_path);
    if (bin_path && bin_path[0]) S.WriteToFileBinary(bin_path);
  } catch (std::exception const& ex) { return fail(ex); }
  return 0;
}

int ref_write_points(const char* text_path, const char* bin_path, long n, double const* xyz, double const* v,
                     int dim) {
  try {
    T3DScalarContainer C;
    for (long i = 0; i != n; ++i) C.AddPoint(v3(xyz + 3 * i), v[i]);
    if (text_path && text_path[0]) C.WriteToFileText(text_path, dim);
    if (bin_path && bin_path[0]) C.WriteToFileBinary(bin_path, dim);
  } catch (std::exception const& ex) { return fail(ex); }
  return 0;
}

// binary = 0: text files, 1: binary files.  The averaged container is written to
// out_text / out_bin (either may be empty) and its point count returned in *n.
int ref_average_spectra(int binary, int nfiles, const char* const* files, const char* out_text, const char* out_bin,
                        long* n) {
  try {
    std::vector<std::string> F(files, files + nfiles);
    TSpectrumContainer S;
    if (binary) S.AverageFromFilesBinary(F); else S.AverageFromFilesText(F);
    if (out_text && out_text[0]) S.WriteToFileText(out_text);
    if (out_bin && out_bin[0]) S.WriteToFileBinary(out_bin);
    *n = (long) S.GetNPoints();
  } catch (std::exception const& ex) { return fail(ex); }
  return 0;
}

int ref_average_points(int binary, int nfiles, const char* const* files, int dim, const char* out_text,
                       const char* out_bin, long* n, double* out_xyz, double* out_value, long cap) {
  try {
    std::vector<std::string> F(files, files + nfiles);
    T3DScalarContainer C;
    if (binary) C.AverageFromFilesBinary(F, dim); else C.AverageFromFilesText(F, dim);
    if (out_text && out_text[0]) C.WriteToFileText(out_text, dim);
    if (out_bin && out_bin[0]) C.WriteToFileBinary(out_bin, dim);
    *n = (long) C.GetNPoints();
    for (size_t i = 0; i != C.GetNPoints() && (long) i < cap; ++i) {
      T3DScalar const& p = C.GetPoint(i);
      out_xyz[3 * i + 0] = p.GetX().GetX();
      out_xyz[3 * i + 1] = p.GetX().GetY();
      out_xyz[3 * i + 2] = p.GetX().GetZ();
      out_value[i] = p.GetV();
    }
  } catch (std::exception const& ex) { return fail(ex); }
  return 0;
}

}  // extern "C"
