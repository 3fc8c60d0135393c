// PEC-sphere bistatic RCS, written against include/xfdtd/xfdtd.hpp.
//
// The scene and call sequence are those of the reference's
// examples/3d_pec_sphere_bistatic_rcs.cpp:34-103 (80-cell air domain at
// dl = 7.5 mm, 12-cell CPML on every face, PEC sphere of radius 0.1 m, TFSF
// at index 27, NF-FF box at index 17, 1 GHz, 180 theta values on [0, pi)).
// Two differences follow from the header: theta/phi/frequencies are
// std::vector<double> (no xtensor), and the optional movie monitor is left
// out, as it is commented out in the reference too.
//
// usage: pec_sphere_bistatic_rcs [output_dir] [total_time_steps] [precision]
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
using xfdtd::Material;
using xfdtd::NffftFd;
using xfdtd::Object;
using xfdtd::PML;
using xfdtd::PointVector;
using xfdtd::Simulation;
using xfdtd::Sphere;
using xfdtd::TFSF3D;

int main(int argc, char **argv) {
  const std::string out_dir = argc > 1 ? argv[1] : "./pec_sphere_bistatic_rcs";
  const size_t total_time_steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 500;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;

  constexpr double dl = 7.5e-3;
  std::vector<std::shared_ptr<Object>> objects;
  const double domain_origin = -40 * dl, domain_size = 80 * dl;
  objects.emplace_back(std::make_shared<Object>(
      "domain",
      std::make_unique<Cube>(PointVector{domain_origin, domain_origin, domain_origin},
                             PointVector{domain_size, domain_size, domain_size}),
      Material{"air", 1, 1, 0, 0}));
  constexpr double radius = 0.1;
  objects.emplace_back(std::make_shared<Object>(
      "scatter", std::make_unique<Sphere>(PointVector{0, 0, 0}, radius),
      Material{"pec", 1, 1, 1e10, 0}));

  std::vector<std::shared_ptr<Boundary>> boundaries;
  for (auto o : {xfdtd::Orientation::XN, xfdtd::Orientation::XP, xfdtd::Orientation::YN,
                 xfdtd::Orientation::YP, xfdtd::Orientation::ZN, xfdtd::Orientation::ZP})
    boundaries.emplace_back(std::make_shared<PML>(o, 12));

  constexpr int n = 20;
  constexpr double lambda_min = dl * n;
  constexpr double tau = lambda_min / 6e8;
  constexpr double t_0 = 4.5 * tau;
  constexpr size_t tfsf_index = 27;
  auto tfsf = std::make_unique<TFSF3D>(tfsf_index, tfsf_index, tfsf_index, 1.0, 0.0, 0.0, 0.0,
                                       std::make_unique<GaussianWaveform>(1, tau, t_0));

  constexpr size_t output_index = 17;
  std::vector<double> theta(180);
  for (size_t i = 0; i < theta.size(); ++i) theta[i] = M_PI * (double)i / 180.0;  // endpoint excluded
  auto nffft = std::make_unique<NffftFd>(output_index, output_index, output_index,
                                         std::vector<double>{1e9}, theta, std::vector<double>{0.0},
                                         out_dir);

  Simulation simulation(dl, objects, boundaries, std::move(tfsf), std::move(nffft), 1);
  simulation.setPrecision(precision);
  const auto t0 = std::chrono::steady_clock::now();
  try {
    simulation.run(total_time_steps);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: %.3f s; NF-FF written to %s\n",
              (size_t)simulation.getNx(), (size_t)simulation.getNy(), (size_t)simulation.getNz(),
              total_time_steps, precision, s, out_dir.c_str());
  return 0;
}
