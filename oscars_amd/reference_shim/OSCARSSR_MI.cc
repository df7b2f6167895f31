// OSCARSSR_MI.cc - the reference's GPU boundary served by liboscars_mi.so.
//
// A maintainer of the reference compiles this file in place of
// src/OSCARSSR_Cuda.cu, with -DCUDA so that OSCARSSR::Calculate*GPU
// (src/OSCARSSR.cc:2405-2445, 3090-3136, 4395-4445) reach it, and links
// liboscars_mi.so.  It implements the declarations of include/OSCARSSR_Cuda.h:28-66
// over the plain C ABI of include/oscars_mi.h:
//
//   OSCARSSR_Cuda_GetDeviceCount          -> om_device_count
//   OSCARSSR_Cuda_GetDeviceProperties     -> om_device_info
//   OSCARSSR_Cuda_CalculateFluxGPU        -> om_trajectory_set + om_flux
//   OSCARSSR_Cuda_CalculateSpectrumGPU    -> om_trajectory_set + om_spectrum
//   OSCARSSR_Cuda_CalculatePowerDensityGPU-> om_trajectory_set + om_power_density_surface
//
// As the reference's own GPU host does (src/OSCARSSR_Cuda.cu:598-607), each
// particle comes from OSR (a new one per step when NParticles > 0) and its
// trajectory from OSR's integrator when it has none; the knots go to the
// device, which splines them and evaluates the sums.  The containers arrive
// with their points set (value 0) and receive AddToPoint / AddToFlux, the
// NotConverged flags and the 1/N weight, as there.  MaxLevel arrives as the
// reference's LevelStopWithExtended and every level up to it is summed with
// the stored-level spacing, as the reference's GPU kernels do.  Errors are
// thrown as std::invalid_argument / std::runtime_error with the library's
// message.
#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "OSCARSSR_Cuda.h"
#include "TVector3DC.h"
#include "oscars_mi.h"

namespace {

constexpr int kMaxDevices = 64;

// One library context per device, created on first use (each owns a stream
// and its cached buffers; the reference drives its GPUs from one host thread).
om_context* context_for(OSCARSSR& OSR, std::vector<int> const& gpus) {
  static om_context* ctx[kMaxDevices] = {};
  int const n = OSR.CheckGPU();
  if (n == 0) throw std::invalid_argument("No GPU found");
  int dev = -1;
  for (int g : gpus)
    if (g >= 0 && g < n) {
      dev = g;
      break;
    }
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("GPUs selected do not match hardware");
  if (!ctx[dev] && om_context_create(dev, &ctx[dev]) != 0) throw std::invalid_argument("No GPU found");
  return ctx[dev];
}

void ok(om_context* ctx, int rc) {
  if (rc != 0) throw std::runtime_error(om_last_error(ctx));
}

// Polarization vector (Re, Im) x 3 from the reference's polarization names
// (src/OSCARSSR.cc:1891-1919): zeros for "all"; "linear" turns the horizontal
// direction by Angle about the propagation direction.
void polarization(std::string name, double angle, TVector3D const& hdir, TVector3D const& pdir, double pol6[6]) {
  std::transform(name.begin(), name.end(), name.begin(), ::tolower);
  std::replace(name.begin(), name.end(), ' ', '-');
  auto has = [&](char const* a, char const* b) {
    return name.find(a) != std::string::npos || (b && name.find(b) != std::string::npos);
  };
  std::complex<double> const I(0, 1);
  TVector3D const vdir = pdir.Cross(hdir).UnitVector();
  TVector3DC P(0, 0, 0);
  if (has("all", nullptr)) {
  } else if (has("linear-horizontal", "lh")) {
    P = hdir;
  } else if (has("linear-vertical", "lv")) {
    P = vdir;
  } else if (name == "linear") {
    TVector3D a = hdir;
    a.RotateSelf(angle, pdir);
    P = a;
  } else if (has("circular-left", "cl")) {
    P = 1. / sqrt(2) * (TVector3DC(hdir) + vdir * I);
  } else if (has("circular-right", "cr")) {
    P = 1. / sqrt(2) * (TVector3DC(hdir) - vdir * I);
  } else {
    throw std::invalid_argument("Polarization requested not recognized");
  }
  std::complex<double> const c[3] = {P.GetX(), P.GetY(), P.GetZ()};
  for (int k = 0; k < 3; ++k) {
    pol6[2 * k] = c[k].real();
    pol6[2 * k + 1] = c[k].imag();
  }
}

// The current particle's trajectory onto the device (knot times [s], x, beta,
// a/c as the reference stores them), integrated by OSR first if it has none.
void trajectory_to_device(OSCARSSR& OSR, om_context* ctx) {
  if (OSR.GetTrajectory().GetNPoints() == 0) OSR.CalculateTrajectory();
  TParticleA const& P = OSR.GetCurrentParticle();
  TParticleTrajectoryPoints const& T = OSR.GetTrajectory();
  size_t const n = T.GetNPoints();
  std::vector<double> t(n), p9(9 * n);
  for (size_t i = 0; i < n; ++i) {
    t[i] = T.GetT(i);
    TVector3D const v[3] = {T.GetX(i), T.GetB(i), T.GetAoverC(i)};
    for (int k = 0; k < 3; ++k) {
      p9[9 * i + 3 * k + 0] = v[k].GetX();
      p9[9 * i + 3 * k + 1] = v[k].GetY();
      p9[9 * i + 3 * k + 2] = v[k].GetZ();
    }
  }
  om_particle op = {};
  for (int k = 0; k < 3; ++k) {
    op.x0[k] = k == 0 ? P.GetX0().GetX() : (k == 1 ? P.GetX0().GetY() : P.GetX0().GetZ());
    op.beta0[k] = k == 0 ? P.GetB0().GetX() : (k == 1 ? P.GetB0().GetY() : P.GetB0().GetZ());
  }
  op.t0 = P.GetT0();
  op.q = P.GetQ();
  op.m = P.GetM();
  op.gamma = P.GetGamma();
  op.qomg = P.GetQoverMGamma();
  op.current = P.GetCurrent();
  ok(ctx, om_trajectory_set(ctx, &op, (long) n, t.data(), p9.data(), 0));
}

// conv for the reference's GPU call: every level up to MaxLevel summed with
// the stored-level spacing (the kernels of src/OSCARSSR_Cuda.cu double the
// sample count per level without the extended spacing)
om_conv conv_for(double precision, int max_level, int quantity) {
  om_conv c = {};
  c.precision = precision;
  c.max_level = max_level;
  c.max_level_extended = 0;
  c.quantity = quantity;
  c.mode = OM_MODE_PLAIN;
  c.weight = 1.0;
  c.accumulate = 0;
  return c;
}

// Particle loop of the reference's GPU hosts: one particle (the current one)
// for NParticles == 0, else NParticles new ones; sums(i) adds particle i.
template <class Sums>
int particle_loop(OSCARSSR& OSR, om_context* ctx, int nparticles, Sums&& sums) {
  int const n = nparticles == 0 ? 1 : nparticles;
  for (int i = 0; i < n; ++i) {
    if (nparticles != 0) OSR.SetNewParticle();
    trajectory_to_device(OSR, ctx);
    sums();
  }
  return n;
}

}  // namespace

extern "C" int OSCARSSR_Cuda_GetDeviceCount() { return om_device_count(); }

std::string OSCARSSR_Cuda_GetDeviceProperties(int const i) {
  char buf[512];
  if (om_device_info(i, buf, sizeof buf) != 0) return "";
  return buf;
}

extern "C" void OSCARSSR_Cuda_CalculateFluxGPU(OSCARSSR& OSR, TSurfacePoints const& Surface, double const Energy_eV,
                                                T3DScalarContainer& FluxContainer, std::string const& PolarizationIn,
                                                double const Angle, TVector3D const& HorizontalDirection,
                                                TVector3D const& PropogationDirection, int const NParticles,
                                                std::vector<int> const& GPUVector, double const Precision,
                                                int const MaxLevel, int const ReturnQuantity) {
  om_context* ctx = context_for(OSR, GPUVector);
  double pol6[6];
  polarization(PolarizationIn, Angle, HorizontalDirection, PropogationDirection, pol6);
  size_t const ns = Surface.GetNPoints();
  std::vector<double> obs(3 * ns), v(ns);
  std::vector<unsigned char> cv(ns);
  for (size_t i = 0; i < ns; ++i) {
    TVector3D const X = Surface.GetPoint(i).GetPoint();
    obs[3 * i] = X.GetX();
    obs[3 * i + 1] = X.GetY();
    obs[3 * i + 2] = X.GetZ();
  }
  om_conv const c = conv_for(Precision, MaxLevel, ReturnQuantity);
  int const n = particle_loop(OSR, ctx, NParticles, [&] {
    ok(ctx, om_flux(ctx, 0, obs.data(), (long) ns, &Energy_eV, 1, pol6, &c, v.data(), nullptr, cv.data(), nullptr,
                    nullptr));
    for (size_t i = 0; i < ns; ++i) {
      FluxContainer.AddToPoint(i, v[i]);
      if (!cv[i]) FluxContainer.SetNotConverged(i);
    }
  });
  FluxContainer.WeightAll(1.0 / (double) n);
}

extern "C" void OSCARSSR_Cuda_CalculateSpectrumGPU(OSCARSSR& OSR, TVector3D const& ObservationPoint,
                                                    TSpectrumContainer& Spectrum, std::string const& PolarizationIn,
                                                    double const Angle, TVector3D const& HorizontalDirection,
                                                    TVector3D const& PropogationDirection, int const NParticles,
                                                    std::vector<int> const& GPUVector, double const Precision,
                                                    int const MaxLevel, int const ReturnQuantity) {
  om_context* ctx = context_for(OSR, GPUVector);
  double pol6[6];
  polarization(PolarizationIn, Angle, HorizontalDirection, PropogationDirection, pol6);
  size_t const ne = Spectrum.GetNPoints();
  std::vector<double> e(ne), v(ne);
  std::vector<unsigned char> cv(ne);
  for (size_t i = 0; i < ne; ++i) e[i] = Spectrum.GetEnergy(i);
  double const obs[3] = {ObservationPoint.GetX(), ObservationPoint.GetY(), ObservationPoint.GetZ()};
  om_conv const c = conv_for(Precision, MaxLevel, ReturnQuantity);
  int const n = particle_loop(OSR, ctx, NParticles, [&] {
    ok(ctx, om_spectrum(ctx, 0, obs, e.data(), (long) ne, pol6, &c, v.data(), nullptr, cv.data(), nullptr, nullptr));
    for (size_t i = 0; i < ne; ++i) {
      Spectrum.AddToFlux(i, v[i]);
      if (!cv[i]) Spectrum.SetNotConverged(i);
    }
  });
  Spectrum.Scale(1.0 / (double) n);
}

extern "C" void OSCARSSR_Cuda_CalculatePowerDensityGPU(OSCARSSR& OSR, TSurfacePoints const& Surface,
                                                        T3DScalarContainer& PowerDensityContainer, int const NParticles,
                                                        std::vector<int> const& GPUVector, double const Precision,
                                                        int const MaxLevel, int const ReturnQuantity) {
  om_context* ctx = context_for(OSR, GPUVector);
  size_t const ns = Surface.GetNPoints();
  bool const has_normal = Surface.HasNormal();
  std::vector<double> obs(3 * ns), nrm(has_normal ? 3 * ns : 0), v(ns);
  std::vector<unsigned char> cv(ns);
  for (size_t i = 0; i < ns; ++i) {
    TSurfacePoint const sp = Surface.GetPoint(i);
    obs[3 * i] = sp.GetPoint().GetX();
    obs[3 * i + 1] = sp.GetPoint().GetY();
    obs[3 * i + 2] = sp.GetPoint().GetZ();
    if (has_normal) {
      nrm[3 * i] = sp.GetNormalX();
      nrm[3 * i + 1] = sp.GetNormalY();
      nrm[3 * i + 2] = sp.GetNormalZ();
    }
  }
  om_conv const c = conv_for(Precision, MaxLevel, ReturnQuantity);
  int const n = particle_loop(OSR, ctx, NParticles, [&] {
    // The reference's GPU call carries no Directional flag; its kernel weighs
    // every sample by the signed N . normal (src/OSCARSSR_Cuda.cu:2519).  The
    // non-directional sum here weighs them the same way and returns the
    // absolute value: equal wherever the surface faces the source.
    ok(ctx, om_power_density_surface(ctx, 0, obs.data(), (long) ns, has_normal ? nrm.data() : nullptr, 0, &c,
                                     v.data(), nullptr, cv.data(), nullptr));
    for (size_t i = 0; i < ns; ++i) {
      PowerDensityContainer.AddToPoint(i, v[i]);
      if (!cv[i]) PowerDensityContainer.SetNotConverged(i);
    }
  });
  PowerDensityContainer.WeightAll(1.0 / (double) n);
}
