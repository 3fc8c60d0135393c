// Dielectric-sphere scattering, written against include/xfdtd/xfdtd.hpp.
//
// The scene and call sequence are those of the reference's
// examples/3d_diel_sphere_scatter.cpp:18-96: a Simulation{dl} built with
// addObject / addBoundary / addTFSFSource / addNFFFT (dl = 7.5 mm, an air
// domain of 2 r + 18 dl around an eps_r = 3, mu_r = 2 sphere of radius
// 0.1 m, 8-cell CPML on every face, TFSF at index 18 with a Gaussian of
// tau = 20 dl / 6e8 and t_0 = 4.5 tau at normal incidence, the NF-FF box at
// index 15 at 1 GHz over 181 theta on [0, pi] at phi = 0), run, then the
// incident wave's Fourier transform at 1 GHz written as incident_fft.npy
// beside the NF-FF output.  As in the other examples theta / phi /
// frequencies are std::vector<double> (no xtensor).
//
// usage: diel_sphere_scatter [output_dir] [total_time_steps] [precision]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "xfdtd/xfdtd.hpp"

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
  const std::string out_dir = argc > 1 ? argv[1] : "./3d_dielectric_sphere_scatter";
  const size_t total_time_steps = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 2000;
  const int precision = argc > 3 ? std::atoi(argv[3]) : 64;

  constexpr double dl{7.5e-3};
  Simulation simulation{dl};
  simulation.setPrecision(precision);

  constexpr double radius{0.1};
  constexpr double domain_size{2 * radius + 18 * dl};
  simulation.addObject(std::make_shared<Object>(
      "domain",
      std::make_unique<Cube>(PointVector{-domain_size / 2, -domain_size / 2, -domain_size / 2},
                             PointVector{domain_size, domain_size, domain_size}),
      Material::createAir()));
  simulation.addObject(std::make_shared<Object>(
      "sphere", std::make_unique<Sphere>(PointVector{0, 0, 0}, radius), Material{"sphere", 3, 2, 0, 0}));

  constexpr size_t pml_thickness{8};
  for (auto o : {xfdtd::Orientation::XN, xfdtd::Orientation::XP, xfdtd::Orientation::YN,
                 xfdtd::Orientation::YP, xfdtd::Orientation::ZN, xfdtd::Orientation::ZP})
    simulation.addBoundary(std::make_shared<PML>(o, pml_thickness));

  constexpr size_t n{20};
  constexpr double lambda_min{dl * n};
  constexpr double tau{lambda_min / 6e8};
  constexpr double t_0{4.5 * tau};
  constexpr size_t inject_index{pml_thickness + 10};
  auto tfsf = std::make_shared<TFSF3D>(inject_index, inject_index, inject_index, 1.0, 0.0, 0.0, 0.0,
                                       std::make_unique<GaussianWaveform>(tau, t_0));
  simulation.addTFSFSource(tfsf);

  std::vector<double> theta(181);
  for (size_t i = 0; i < theta.size(); ++i) theta[i] = M_PI * (double)i / 180.0;
  simulation.addNFFFT(std::make_shared<NffftFd>(inject_index - 3, inject_index - 3, inject_index - 3,
                                                std::vector<double>{1e9}, theta, std::vector<double>{0.0},
                                                out_dir));

  const auto t0 = std::chrono::steady_clock::now();
  try {
    simulation.run(total_time_steps);
    xfdtd::dumpNpy(out_dir + "/incident_fft.npy", tfsf->getIncidentWaveFourierTransform({1e9}));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "simulation failed: %s\n", e.what());
    return 1;
  }
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("grid %zu x %zu x %zu, %zu steps, FP%d: It costs %.3f seconds to run Dielectric Sphere Scatter.\n",
              (size_t)simulation.getNx(), (size_t)simulation.getNy(), (size_t)simulation.getNz(),
              total_time_steps, precision, s);
  return 0;
}
