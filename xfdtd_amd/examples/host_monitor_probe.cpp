// A monitor of one's own on the host (xfdtd::HostMonitor, the reference's
// Monitor::init / update / outputData virtuals of monitor/monitor.h:28-37):
// an Ez recorder that reads one cell through Simulation::getFieldBox after
// every step, beside the engine's own point probe at the same cell.  The
// scene is a 40-cell air cube at dl = 5 mm with 8-cell CPML, an eps_r = 4
// sphere of radius 40 mm and a TFSF Gaussian along +z.  Both series go to
// [output_dir]/host_ez.npy and probe_ez.npy; they are the same numbers.
//
// usage: host_monitor_probe [output_dir] [total_time_steps] [precision]
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xfdtd/xfdtd.hpp"

using namespace xfdtd;

class EzRecorder : public HostMonitor {
 public:
  EzRecorder(size_t i, size_t j, size_t k, std::string dir)
      : HostMonitor(std::move(dir), "host_ez.npy"), _i{i}, _j{j}, _k{k} {}
  void init(const Simulation &sim) override { _values.reserve(sim.totalTimeSteps()); }
  void update(const Simulation &sim, size_t) override {
    _values.push_back(sim.getFieldBox(EMComponent::EZ, {_i, _i + 1, _j, _j + 1, _k, _k + 1})[0]);
  }
  void outputData() override { dumpNpy(getOutputPath() + "/" + getOutputFileName(), _values); }

 private:
  size_t _i, _j, _k;
  std::vector<double> _values;
};

int main(int argc, char **argv) {
  const std::string out_dir = argc > 1 ? argv[1] : "./host_monitor_probe";
  const size_t steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 200;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;
  constexpr double dl = 5e-3;
  Simulation sim{dl, 0.9f};
  sim.setPrecision(precision);
  sim.addObject(std::make_shared<Object>(
      "domain", std::make_unique<Cube>(PointVector{-0.1, -0.1, -0.1}, PointVector{0.2, 0.2, 0.2}),
      Material::createAir()));
  sim.addObject(std::make_shared<Object>("ball", std::make_unique<Sphere>(PointVector{0, 0, 0}, 0.04),
                                         Material{"ball", 4, 1, 0, 0}));
  for (auto o : {Orientation::XN, Orientation::XP, Orientation::YN, Orientation::YP, Orientation::ZN,
                 Orientation::ZP})
    sim.addBoundary(std::make_shared<PML>(o, 8));
  const double tau = 20 * dl / 6e8;
  sim.addTFSFSource(std::make_shared<TFSF3D>(12, 12, 12, 1.0, 0.0, 0.0, 0.0,
                                             std::make_unique<GaussianWaveform>(tau, 4.5 * tau)));
  sim.addProbe(Probe{EMComponent::EZ, 24, 27, 30});
  auto rec = std::make_shared<EzRecorder>(24, 27, 30, out_dir);
  sim.addMonitor(rec);
  try {
    sim.run(steps);
    dumpNpy(out_dir + "/probe_ez.npy", sim.getProbeData());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: host monitor and probe written to %s\n",
              (size_t)sim.getNx(), (size_t)sim.getNy(), (size_t)sim.getNz(), steps, precision, out_dir.c_str());
  return 0;
}
