// Lorentz-sphere monostatic RCS, written against include/xfdtd/xfdtd.hpp.
//
// The scene and call sequence are those of the reference's
// test/test_3d_fdtd.cpp:158-211 (testLorentzSphereMonoStaticRCS): a sphere
// of radius 15 nm made of LorentzMedium(2.25, 1, 4e16, 0.28e16) at
// dl = radius / 30 in a 2.5-radius air cube, 12-cell CPML on every face, TFSF
// at index 22 with a Gaussian of tau = lambda_min / 6e8 (lambda_min = 20 dl)
// travelling along +z, broadband NF-FF box at index 15 observed back along
// -z, the default cfl; then the incident wave's FFT.  With the medium
// argument `drude` it runs testDrudeSphereMonostaticRCS (test_3d_fdtd.cpp:
// 213-270) instead: radius 3.75 mm and DrudeMedium(1, 1.8e11, 2e10), the
// rest alike; with `debye`, testDebySphereMonostaticRCS (test_3d_fdtd.cpp:
// 268-321): radius 0.25 m and DebyMedium(1.16, 1.01, 6.497e-10, 2.95e-4).
//
// usage: lorentz_sphere_monostatic_rcs [output_dir] [total_time_steps] [precision] [lorentz|drude|debye]
//                                      [domain edge in radii, default 2.5]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xfdtd/xfdtd.hpp"

using xfdtd::Boundary;
using xfdtd::Cube;
using xfdtd::GaussianWaveform;
using xfdtd::LorentzMedium;
using xfdtd::Material;
using xfdtd::NffftBroadBand;
using xfdtd::Object;
using xfdtd::PML;
using xfdtd::PointVector;
using xfdtd::Simulation;
using xfdtd::Sphere;
using xfdtd::TFSF3D;

int main(int argc, char **argv) {
  const std::string out_dir = argc > 1 ? argv[1] : "./lorentz_sphere";
  const size_t total_time_steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 1000;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;
  const std::string kind = argc > 4 ? argv[4] : "lorentz";
  if (kind != "lorentz" && kind != "drude" && kind != "debye") {
    std::fprintf(stderr, "unknown medium %s (lorentz, drude or debye)\n", kind.c_str());
    return 2;
  }

  const double radius = kind == "drude" ? 3.75e-3 : kind == "debye" ? 0.25 : 15e-9;
  const double dl = radius / 30;
  // the air cube's edge in radii (the reference's 2.5 puts the TFSF surface
  // at index 22 inside the sphere; 3 or more keep the sphere in the total
  // field region)
  const double radii = argc > 5 ? std::atof(argv[5]) : 2.5;
  const double domain_size = radius * radii;
  const double domain_origin = -domain_size / 2;
  std::vector<std::shared_ptr<Object>> objects;
  objects.emplace_back(std::make_shared<Object>(
      "domain",
      std::make_unique<Cube>(PointVector{domain_origin, domain_origin, domain_origin},
                             PointVector{domain_size, domain_size, domain_size}),
      Material{"air", 1, 1, 0, 0}));
  std::unique_ptr<Material> medium;
  if (kind == "drude")
    medium = std::make_unique<xfdtd::DrudeMedium>("random", 1, 1.8e11, 2e10);
  else if (kind == "debye")
    medium = std::make_unique<xfdtd::DebyMedium>("random", 1.16, 1.01, 6.497e-10, 2.95e-4);
  else
    medium = std::make_unique<LorentzMedium>("random", 2.25, 1, 4e16, 0.28e16);
  objects.emplace_back(std::make_shared<Object>("scatter", std::make_unique<Sphere>(PointVector{0, 0, 0}, radius),
                                                std::move(medium)));

  std::vector<std::shared_ptr<Boundary>> boundaries;
  for (auto o : {xfdtd::Orientation::XN, xfdtd::Orientation::XP, xfdtd::Orientation::YN,
                 xfdtd::Orientation::YP, xfdtd::Orientation::ZN, xfdtd::Orientation::ZP})
    boundaries.emplace_back(std::make_shared<PML>(o, 12));

  constexpr int n = 20;
  const double lambda_min = dl * n;
  const double tau = lambda_min / 6e8;
  const double t_0 = 4.5 * tau;
  constexpr double incident_theta = 0, incident_phi = 0, polarization = 0;
  constexpr size_t tfsf_index = 22;
  auto tfsf = std::make_unique<TFSF3D>(tfsf_index, tfsf_index, tfsf_index, 1.0, incident_theta, incident_phi,
                                       polarization, std::make_unique<GaussianWaveform>(1, tau, t_0));
  constexpr size_t output_index = 15;
  auto mono_rcs = std::make_unique<NffftBroadBand>(output_index, output_index, output_index,
                                                   incident_theta + M_PI, incident_phi + M_PI, out_dir);

  Simulation simulation(dl, objects, boundaries, std::move(tfsf), std::move(mono_rcs));
  simulation.setPrecision(precision);
  const auto t0 = std::chrono::steady_clock::now();
  try {
    simulation.run(total_time_steps);
    simulation.outputTFSFIncidentWaveFastFourierTransform(out_dir + "/incident_wave_fft");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: %.3f s; far field written to %s\n",
              (size_t)simulation.getNx(), (size_t)simulation.getNy(), (size_t)simulation.getNz(),
              total_time_steps, precision, s, out_dir.c_str());
  return 0;
}
