// Quarter-wave microstrip transformer S parameters, written against
// include/xfdtd/xfdtd.hpp.
//
// The scene and call sequence are those of the reference's
// examples/s_parameter_network.cpp:15-147: a 60 x 100 x 15 cell air box at
// 0.2 mm over a 4.6 eps_r substrate, 50 / 70.7 / 100 ohm microstrip lines as
// PEC ObjectPlanes, a 50 ohm Gaussian voltage source at port 1 and a 100 ohm
// resistor at port 2, CPML on the x and y faces, voltage and current
// monitors at both ports, and a two-port Network over 20 MHz .. 7.98 GHz.
//
// usage: quarter_wave_transformer [output_dir] [total_time_steps] [precision]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xfdtd/xfdtd.hpp"

using namespace xfdtd;

int main(int argc, char **argv) {
  const std::string out_dir = argc > 1 ? argv[1] : "./quarter_wave_transformer";
  const size_t total_time_steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 5000;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;

  const Material air{"air", 1, 1, 0, 0};
  const Material pec{"pec", 1, 1, 1e10, 0};
  constexpr double size = 2e-4;
  std::vector<std::shared_ptr<Object>> objects;
  objects.emplace_back(std::make_shared<Object>(
      "domain", std::make_unique<Cube>(PointVector{-5 * size, -5 * size, 0}, PointVector{60 * size, 100 * size, 15 * size}),
      air));
  constexpr double x = 50 * size, y = 90 * size, z = 5 * size;
  objects.emplace_back(std::make_shared<Object>(
      "substrate", std::make_unique<Cube>(PointVector{0, 0, 0}, PointVector{x, y, z}), Material{"substrate", 4.6, 1, 0, 0}));
  constexpr double w50 = 9 * size, l50 = 20 * size, w70 = 5 * size, l70 = 50 * size, w100 = 2 * size, l100 = 20 * size;
  objects.emplace_back(std::make_shared<ObjectPlane>(
      "msl_50", std::make_unique<Cube>(PointVector{4e-3, 0, z}, PointVector{w50, l50, 0}), pec));
  objects.emplace_back(std::make_shared<ObjectPlane>(
      "msl_70_7", std::make_unique<Cube>(PointVector{4.4e-3, l50, z}, PointVector{w70, l70, 0}), pec));
  objects.emplace_back(std::make_shared<ObjectPlane>(
      "msl_100", std::make_unique<Cube>(PointVector{4.8e-3, l50 + l70, z}, PointVector{w100, l100, 0}), pec));

  constexpr double l_min = size * 20;
  const double tau = l_min / 6e8, t_0 = 4.5 * tau;
  std::vector<std::shared_ptr<LumpedElement>> lumped_elements;
  lumped_elements.emplace_back(std::make_shared<VoltageSource>(
      std::make_unique<Cube>(PointVector{4e-3, 0, 0}, PointVector{w50, 2 * size, z}), Orientation::ZP, 50,
      std::make_unique<GaussianWaveform>(1, tau, t_0)));
  lumped_elements.emplace_back(std::make_shared<Resistor>(
      std::make_unique<Cube>(PointVector{4.8e-3, y - 2 * size, 0}, PointVector{w100, 2 * size, z}), Axis::Z, 100));

  std::vector<std::shared_ptr<Boundary>> boundaries;
  for (auto o : {Orientation::XN, Orientation::XP, Orientation::YN, Orientation::YP})
    boundaries.emplace_back(std::make_shared<PML>(o, 8));

  auto v1 = std::make_shared<VoltageMonitor>(std::make_unique<Cube>(PointVector{4e-3, 0, 0}, PointVector{w50, 2 * size, z}),
                                             Orientation::ZP, out_dir, "v1.npy");
  auto v2 = std::make_shared<VoltageMonitor>(
      std::make_unique<Cube>(PointVector{4.8e-3, y - 2 * size, 0}, PointVector{w100, 2 * size, z}), Orientation::ZP, out_dir,
      "v2.npy");
  auto c1 = std::make_shared<CurrentMonitor>(
      std::make_unique<Cube>(PointVector{4e-3, 0, 0.4e-3}, PointVector{9 * size, 2 * size, 1 * size}), Orientation::ZP,
      out_dir, "c1.npy");
  auto c2 = std::make_shared<CurrentMonitor>(
      std::make_unique<Cube>(PointVector{4.8e-3, y - 2 * size, 0.4e-3}, PointVector{2 * size, 2 * size, 1 * size}),
      Orientation::ZP, out_dir, "c2.npy");
  std::vector<std::shared_ptr<Monitor>> monitors{v1, v2, c1, c2};

  std::vector<std::unique_ptr<Port>> ports;
  ports.emplace_back(std::make_unique<Port>(1, 50, true, v1, c1));
  ports.emplace_back(std::make_unique<Port>(2, 100, false, v2, c2));
  std::vector<double> frequencies;  // xt::arange<double>(2e7, 8e9, 2e7)
  for (size_t q = 0; 2e7 + 2e7 * (double)q < 8e9; ++q) frequencies.push_back(2e7 + 2e7 * (double)q);
  auto network = std::make_unique<Network>(std::move(ports), frequencies, out_dir);

  Simulation simulation(size, objects, lumped_elements, boundaries, monitors, std::move(network));
  simulation.setPrecision(precision);
  const auto t0 = std::chrono::steady_clock::now();
  try {
    simulation.run(total_time_steps);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: %.3f s; S parameters written to %s\n",
              (size_t)simulation.getNx(), (size_t)simulation.getNy(), (size_t)simulation.getNz(), total_time_steps,
              precision, s, out_dir.c_str());
  return 0;
}
