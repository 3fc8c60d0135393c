// 2-D scattering by a PEC cylinder (TMz), written against include/xfdtd/xfdtd.hpp.
//
// The scene and call sequence are those of the reference's
// examples/2d_cylinder_scatter.cpp:26-106: a 350 x 300 cell vacuum plane at
// dl = lambda(20 GHz) / 20 with a 10-cell CPML on the four x and y faces, a
// z-axis PEC cylinder of radius 3 cm at the centre, a TFSF2D source 50 cells
// in travelling along phi = pi/4 with a cosine-modulated Gaussian
// (tau = 1.7 / 8 GHz, t_0 = 0.8 tau, 12 GHz), a broadband NF-FF box 40 cells
// in observed along (pi/2, 5 pi/4) (back towards the source), a movie of Ez
// every 100 steps, cfl 0.8; then the incident wave's FFT.
//
// usage: cylinder_2d_scatter [output_dir] [total_time_steps] [precision]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xfdtd/xfdtd.hpp"

using xfdtd::Boundary;
using xfdtd::CosineModulatedGaussianWaveform;
using xfdtd::Cube;
using xfdtd::Cylinder;
using xfdtd::Material;
using xfdtd::MovieMonitor;
using xfdtd::NffftBroadBand;
using xfdtd::Object;
using xfdtd::PML;
using xfdtd::PointVector;
using xfdtd::Simulation;
using xfdtd::TFSF2D;
using xfdtd::TimeDomainFieldMonitor;

int main(int argc, char **argv) {
  const std::string out_dir = argc > 1 ? argv[1] : "./2d_cylinder_scatter";
  const size_t total_time_steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 1400;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;

  constexpr size_t nx = 350, ny = 300, pml_size = 10;
  constexpr double center_frequency = 12e9, max_frequency = 20e9;
  constexpr double min_lambda = 3e8 / max_frequency;
  constexpr double dx = min_lambda / 20, dy = dx;
  constexpr double tau = 1.7 / (max_frequency - center_frequency);
  constexpr double t_0 = 0.8 * tau;
  constexpr double cylinder_x = nx * 0.5 * dx, cylinder_y = ny * 0.5 * dx, cylinder_radius = 0.03;

  std::vector<std::shared_ptr<Object>> objects;
  objects.emplace_back(std::make_shared<Object>(
      "free_space", std::make_unique<Cube>(PointVector{0, 0, 0}, PointVector{nx * dx, ny * dy, 0}),
      Material{"air", 1, 1, 0, 0}));
  objects.emplace_back(std::make_shared<Object>(
      "objectA",
      std::make_unique<Cylinder>(xfdtd::Axis::Z, PointVector{cylinder_x, cylinder_y, 0}, cylinder_radius, 0),
      Material{"pec", 1, 1, 1e24, 0}));

  std::vector<std::shared_ptr<Boundary>> boundaries;
  for (auto o : {xfdtd::Orientation::XN, xfdtd::Orientation::XP, xfdtd::Orientation::YN, xfdtd::Orientation::YP})
    boundaries.emplace_back(std::make_shared<PML>(o, (int)pml_size));

  auto movie = std::make_shared<MovieMonitor>(
      std::make_unique<TimeDomainFieldMonitor>(std::make_unique<Cube>(PointVector{0, 0, 0}, PointVector{nx * dx, ny * dy, 0}),
                                               xfdtd::PlaneType::XY, xfdtd::EMComponent::EZ,
                                               out_dir + "/movie_ez", ""),
      total_time_steps, 100);

  auto tfsf = std::make_unique<TFSF2D>(
      50, 50, M_PI * 0.25, 1,
      std::make_unique<CosineModulatedGaussianWaveform>(1, tau, t_0, center_frequency));
  auto nffft = std::make_unique<NffftBroadBand>(40, 40, 0, M_PI / 2, M_PI * 1.25, out_dir + "/nffft");

  Simulation simulation(dx, objects, boundaries, std::move(tfsf), std::move(nffft), 0.8);
  simulation.addMonitor(movie);
  simulation.setPrecision(precision);
  const auto t0 = std::chrono::steady_clock::now();
  try {
    simulation.run(total_time_steps);
    simulation.outputTFSFIncidentWaveFastFourierTransform(out_dir + "/tfsf_incident_wave_fft");
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: %.3f s; output in %s\n", (size_t)simulation.getNx(),
              (size_t)simulation.getNy(), (size_t)simulation.getNz(), total_time_steps, precision, s,
              out_dir.c_str());
  return 0;
}
