// sim_capi.cpp -- include/xfdtd_sim.h over the C++ Simulation.
#include <cstring>
#include <string>

#include "xfdtd/xfdtd.hpp"
#include "xfdtd_engine.h"
#include "xfdtd_sim.h"

using namespace xfdtd;

struct xf_sim {
  Simulation sim;
  std::shared_ptr<NffftFd> nffft;
  std::vector<std::shared_ptr<Monitor>> line_monitors;  // voltage / current monitors, in order
  std::shared_ptr<Network> network;
  xf_sim(double dl, float cfl) : sim(dl, cfl) {}
};

static thread_local std::string g_sim_err;
extern "C" const char *xf_sim_last_error(void) { return g_sim_err.c_str(); }

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_sim_err = e.what();
  } catch (...) {
    g_sim_err = "unknown error";
  }
  return -1;
}

extern "C" xf_sim *xf_sim_create(double cell_size, float cfl) {
  try {
    return new xf_sim(cell_size, cfl);
  } catch (const std::exception &e) {
    g_sim_err = e.what();
    return nullptr;
  }
}

extern "C" void xf_sim_destroy(xf_sim *s) { delete s; }

static Material mat(float eps_r, float mu_r, double sigma_e, double sigma_m) {
  return Material("material", eps_r, mu_r, sigma_e, sigma_m);
}

extern "C" int xf_sim_add_cube(xf_sim *s, const double origin[3], const double size[3], float eps_r,
                               float mu_r, double sigma_e, double sigma_m) {
  return guarded([&] {
    s->sim.addObject(std::make_shared<Object>(
        "cube", std::make_unique<Cube>(PointVector{origin[0], origin[1], origin[2]},
                                       PointVector{size[0], size[1], size[2]}),
        mat(eps_r, mu_r, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_sphere(xf_sim *s, const double center[3], double radius, float eps_r,
                                 float mu_r, double sigma_e, double sigma_m) {
  return guarded([&] {
    s->sim.addObject(std::make_shared<Object>(
        "sphere", std::make_unique<Sphere>(PointVector{center[0], center[1], center[2]}, radius),
        mat(eps_r, mu_r, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_cylinder(xf_sim *s, int axis, const double center[3], double radius, double height,
                                   float eps_r, float mu_r, double sigma_e, double sigma_m) {
  return guarded([&] {
    if (axis < 0 || axis > 2) throw std::invalid_argument("bad cylinder axis");
    s->sim.addObject(std::make_shared<Object>(
        "cylinder",
        std::make_unique<Cylinder>(static_cast<Axis>(axis), PointVector{center[0], center[1], center[2]}, radius,
                                   height),
        mat(eps_r, mu_r, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_drude_cube(xf_sim *s, const double origin[3], const double size[3], double eps_inf,
                                     double omega_p, double nv, double sigma_e, double sigma_m) {
  return guarded([&] {
    s->sim.addObject(std::make_shared<Object>(
        "drude_cube", std::make_unique<Cube>(PointVector{origin[0], origin[1], origin[2]},
                                             PointVector{size[0], size[1], size[2]}),
        DrudeMedium("drude", eps_inf, omega_p, nv, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_drude_sphere(xf_sim *s, const double center[3], double radius, double eps_inf,
                                       double omega_p, double nv, double sigma_e, double sigma_m) {
  return guarded([&] {
    s->sim.addObject(std::make_shared<Object>(
        "drude_sphere", std::make_unique<Sphere>(PointVector{center[0], center[1], center[2]}, radius),
        DrudeMedium("drude", eps_inf, omega_p, nv, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_dispersive_object(xf_sim *s, int shape, const double p[6], int kind, const double mp[4],
                                            double sigma_e, double sigma_m) {
  return guarded([&] {
    std::unique_ptr<Shape> sh;
    if (shape == 0)
      sh = std::make_unique<Cube>(PointVector{p[0], p[1], p[2]}, PointVector{p[3], p[4], p[5]});
    else if (shape == 1)
      sh = std::make_unique<Sphere>(PointVector{p[0], p[1], p[2]}, p[3]);
    else if (shape == 2 && p[5] >= 0 && p[5] <= 2)
      sh = std::make_unique<Cylinder>(static_cast<Axis>((int)p[5]), PointVector{p[0], p[1], p[2]}, p[3], p[4]);
    else
      throw std::invalid_argument("bad shape");
    std::unique_ptr<Material> m;
    if (kind == 1)
      m = std::make_unique<DrudeMedium>("drude", mp[0], mp[1], mp[2], sigma_e, sigma_m);
    else if (kind == 2)
      m = std::make_unique<DebyMedium>("debye", mp[0], mp[1], mp[2], sigma_e, sigma_m);
    else if (kind == 3)
      m = std::make_unique<LorentzMedium>("lorentz", mp[0], mp[1], mp[2], mp[3], sigma_e, sigma_m);
    else
      throw std::invalid_argument("bad dispersive kind");
    s->sim.addObject(std::make_shared<Object>("dispersive", std::move(sh), std::move(m)));
  });
}

extern "C" int xf_sim_drude_box(xf_sim *s, int64_t box[6]) {
  return guarded([&] {
    xf_engine *e = s->sim.engine();
    if (!e) throw std::runtime_error("simulation not initialised");
    if (xf_engine_drude_box(e, box) != XF_OK) throw std::runtime_error(xf_last_error());
  });
}

extern "C" int xf_sim_set_drude_j(xf_sim *s, int comp, const double *in) {
  return guarded([&] {
    xf_engine *e = s->sim.engine();
    if (!e) throw std::runtime_error("simulation not initialised");
    if (xf_engine_set_drude_j(e, comp, in) != XF_OK) throw std::runtime_error(xf_last_error());
  });
}

extern "C" int xf_sim_get_drude_j(xf_sim *s, int comp, double *out) {
  return guarded([&] {
    xf_engine *e = s->sim.engine();
    if (!e) throw std::runtime_error("simulation not initialised");
    if (xf_engine_get_drude_j(e, comp, out) != XF_OK) throw std::runtime_error(xf_last_error());
  });
}

extern "C" int xf_sim_add_pml(xf_sim *s, int orientation, int thickness, int order, double sigma_ratio,
                              double alpha_min, double alpha_max, double kappa_max) {
  return guarded([&] {
    if (orientation < 0 || orientation > 5) throw std::invalid_argument("bad orientation");
    s->sim.addBoundary(std::make_shared<PML>(static_cast<Orientation>(orientation), thickness, order,
                                             sigma_ratio, alpha_min, alpha_max, kappa_max));
  });
}

extern "C" int xf_sim_set_tfsf3d(xf_sim *s, const size_t distance[3], double e0, double theta,
                                 double phi, double psi, double amplitude, double tau, double t0) {
  return guarded([&] {
    s->sim.addTFSFSource(std::make_shared<TFSF3D>(distance[0], distance[1], distance[2], e0, theta, phi,
                                                  psi, std::make_shared<GaussianWaveform>(amplitude, tau, t0)));
  });
}

extern "C" int xf_sim_set_nffft_fd(xf_sim *s, const size_t distance[3], const double *freqs, int nf,
                                   const double *theta, int ntheta, const double *phi, int nphi,
                                   const char *output_dir) {
  return guarded([&] {
    s->nffft = std::make_shared<NffftFd>(
        distance[0], distance[1], distance[2], std::vector<double>(freqs, freqs + nf),
        std::vector<double>(theta, theta + ntheta), std::vector<double>(phi, phi + nphi),
        output_dir ? std::string(output_dir) : std::string());
    s->sim.addNFFFT(s->nffft);
  });
}

extern "C" int xf_sim_set_nffft_broadband(xf_sim *s, const size_t distance[3], double theta, double phi,
                                          const char *output_dir) {
  return guarded([&] {
    s->sim.addNFFFT(std::make_shared<NffftBroadBand>(distance[0], distance[1], distance[2], theta, phi,
                                                     output_dir ? std::string(output_dir) : std::string()));
  });
}

extern "C" int xf_sim_add_current_source(xf_sim *s, const double origin[3], const double size[3],
                                         int orientation, double resistance, double amplitude,
                                         double tau, double t0) {
  return guarded([&] {
    if (orientation < 0 || orientation > 5) throw std::invalid_argument("bad orientation");
    s->sim.addLumpedElement(std::make_shared<CurrentSource>(
        std::make_unique<Cube>(PointVector{origin[0], origin[1], origin[2]},
                               PointVector{size[0], size[1], size[2]}),
        static_cast<Orientation>(orientation), resistance,
        std::make_unique<GaussianWaveform>(amplitude, tau, t0)));
  });
}

static std::unique_ptr<Cube> cube_of(const double origin[3], const double size[3]) {
  return std::make_unique<Cube>(PointVector{origin[0], origin[1], origin[2]}, PointVector{size[0], size[1], size[2]});
}

static std::unique_ptr<Waveform> waveform_of(int wkind, const double *wp);

extern "C" int xf_sim_set_tfsf2d(xf_sim *s, size_t distance_x, size_t distance_y, double phi_inc, double ez_i0,
                                 int wkind, const double wp[4]) {
  return guarded([&] {
    s->sim.addTFSFSource(
        std::make_shared<TFSF2D>(distance_x, distance_y, phi_inc, ez_i0, std::shared_ptr<Waveform>(waveform_of(wkind, wp))));
  });
}

static std::unique_ptr<Waveform> waveform_of(int wkind, const double *wp) {
  if (wkind == 1) return std::make_unique<SinusoidalWaveform>(wp[0], wp[1], wp[2]);
  if (wkind == 2) return std::make_unique<CosineModulatedGaussianWaveform>(wp[0], wp[1], wp[2], wp[3]);
  if (wkind != 0) throw std::invalid_argument("bad waveform");
  return std::make_unique<GaussianWaveform>(wp[0], wp[1], wp[2]);
}

extern "C" int xf_sim_add_lumped(xf_sim *s, int kind, const double origin[3], const double size[3], int dir,
                                 double value, int wkind, const double wp[4]) {
  return guarded([&] {
    if (kind <= 1 ? (dir < 0 || dir > 5) : (dir < 0 || dir > 2)) throw std::invalid_argument("bad orientation or axis");
    std::shared_ptr<LumpedElement> le;
    switch (kind) {
      case 0:
        le = std::make_shared<CurrentSource>(cube_of(origin, size), static_cast<Orientation>(dir), value,
                                             waveform_of(wkind, wp));
        break;
      case 1:
        le = std::make_shared<VoltageSource>(cube_of(origin, size), static_cast<Orientation>(dir), value,
                                             waveform_of(wkind, wp));
        break;
      case 2: le = std::make_shared<Resistor>(cube_of(origin, size), static_cast<Axis>(dir), value); break;
      case 3: le = std::make_shared<Capacitor>(cube_of(origin, size), static_cast<Axis>(dir), value); break;
      case 4: le = std::make_shared<Inductor>(cube_of(origin, size), static_cast<Axis>(dir), value); break;
      default: throw std::invalid_argument("bad lumped element kind");
    }
    s->sim.addLumpedElement(std::move(le));
  });
}

extern "C" int xf_sim_add_plane(xf_sim *s, const double origin[3], const double size[3], float eps_r, float mu_r,
                                double sigma_e, double sigma_m) {
  return guarded([&] {
    s->sim.addObject(std::make_shared<ObjectPlane>("plane", cube_of(origin, size), mat(eps_r, mu_r, sigma_e, sigma_m)));
  });
}

extern "C" int xf_sim_add_line_monitor(xf_sim *s, int kind, const double origin[3], const double size[3],
                                       int orientation, const char *dir, const char *file) {
  return guarded([&] {
    if (orientation < 0 || orientation > 5 || kind < 0 || kind > 1) throw std::invalid_argument("bad monitor");
    std::shared_ptr<Monitor> m;
    const std::string d = dir ? dir : "", f = file ? file : "";
    if (kind == 0)
      m = std::make_shared<VoltageMonitor>(cube_of(origin, size), static_cast<Orientation>(orientation), d, f);
    else
      m = std::make_shared<CurrentMonitor>(cube_of(origin, size), static_cast<Orientation>(orientation), d, f);
    s->line_monitors.push_back(m);
    s->sim.addMonitor(m);
  });
}

extern "C" int xf_sim_get_line_monitors(xf_sim *s, double *out) {
  return guarded([&] {
    auto v = s->sim.getLineMonitorData();
    size_t off = 0;
    for (const auto &m : v) {
      std::memcpy(out + off, m.data(), m.size() * sizeof(double));
      off += m.size();
    }
  });
}

extern "C" int xf_sim_set_network(xf_sim *s, int nports, const int *index, const double *impedance,
                                  const int *is_source, const int *vmon, const int *imon, const double *freqs,
                                  int nf, const char *output_dir) {
  return guarded([&] {
    std::vector<std::unique_ptr<Port>> ports;
    const int nm = (int)s->line_monitors.size();
    for (int q = 0; q < nports; ++q) {
      if (vmon[q] < 0 || vmon[q] >= nm || imon[q] < 0 || imon[q] >= nm) throw std::invalid_argument("bad port monitor");
      auto v = std::dynamic_pointer_cast<VoltageMonitor>(s->line_monitors[vmon[q]]);
      auto c = std::dynamic_pointer_cast<CurrentMonitor>(s->line_monitors[imon[q]]);
      if (!v || !c) throw std::invalid_argument("a port needs a voltage and a current monitor");
      ports.push_back(std::make_unique<Port>(index[q], std::complex<double>(impedance[2 * q], impedance[2 * q + 1]),
                                             is_source[q] != 0, v, c));
    }
    s->network = std::make_shared<Network>(std::move(ports), std::vector<double>(freqs, freqs + nf),
                                           output_dir ? std::string(output_dir) : std::string());
    s->sim.addNetwork(s->network);
  });
}

extern "C" int xf_sim_network_size(const xf_sim *s) {
  if (!s->network) return 0;
  int src = 0;
  for (const auto &p : s->network->ports()) src += p->isActivatedSource() ? 1 : 0;
  return src * (int)s->network->ports().size();
}

extern "C" int xf_sim_network_sparams(xf_sim *s, int *keys, double *out) {
  return guarded([&] {
    if (!s->network) throw std::runtime_error("no network");
    s->sim.getLineMonitorData();
    auto sp = s->network->calculateSParameters(s->sim.getDt());
    const size_t nf = s->network->frequencies().size();
    size_t q = 0;
    for (const auto &e : sp) {
      keys[q] = e.first;
      std::memcpy(out + 2 * nf * q, e.second.data(), nf * sizeof(std::complex<double>));
      ++q;
    }
  });
}

extern "C" int xf_sim_add_probe(xf_sim *s, int comp, size_t i, size_t j, size_t k) {
  return guarded([&] {
    if (comp < 0 || comp > 5) throw std::invalid_argument("bad component");
    s->sim.addProbe(Probe{static_cast<EMComponent>(comp), i, j, k});
  });
}

static std::unique_ptr<TimeDomainFieldMonitor> field_monitor(const double origin[3], const double size[3],
                                                              int plane_type, int comp, const char *dir,
                                                              const char *file) {
  if (comp < 0 || comp > 5) throw std::invalid_argument("bad component");
  if (plane_type < 0 || plane_type > 2) throw std::invalid_argument("bad plane type");
  return std::make_unique<TimeDomainFieldMonitor>(
      std::make_unique<Cube>(PointVector{origin[0], origin[1], origin[2]}, PointVector{size[0], size[1], size[2]}),
      static_cast<PlaneType>(plane_type), static_cast<EMComponent>(comp), dir ? dir : "", file ? file : "");
}

extern "C" int xf_sim_add_field_monitor(xf_sim *s, const double origin[3], const double size[3], int plane_type,
                                        int comp, const char *dir, const char *file) {
  return guarded([&] { s->sim.addMonitor(field_monitor(origin, size, plane_type, comp, dir, file)); });
}

extern "C" int xf_sim_add_movie_monitor(xf_sim *s, const double origin[3], const double size[3], int plane_type,
                                        int comp, const char *dir, size_t total_steps, size_t plot_step) {
  return guarded([&] {
    s->sim.addMonitor(std::make_shared<MovieMonitor>(field_monitor(origin, size, plane_type, comp, dir, ""),
                                                     total_steps, plot_step));
  });
}

extern "C" int xf_sim_plan(xf_sim *s, size_t total_steps) {
  return guarded([&] { s->sim.plan(total_steps); });
}

extern "C" int xf_sim_check_run(xf_sim *s, size_t total_steps) {
  return guarded([&] { s->sim.checkRun(total_steps); });
}

extern "C" int xf_sim_plan_coeff(xf_sim *s, int which, double *out) {
  return guarded([&] {
    auto v = s->sim.planCoefficient(which);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

extern "C" int xf_sim_init(xf_sim *s, size_t total_steps, int precision) {
  return guarded([&] {
    if (precision != 32 && precision != 64) throw std::invalid_argument("precision must be 32 or 64");
    s->sim.setPrecision(precision);
    s->sim.init(total_steps);
  });
}

extern "C" int xf_sim_step(xf_sim *s, size_t n) {
  return guarded([&] { s->sim.step(n); });
}

extern "C" int xf_sim_sync(xf_sim *s) {
  return guarded([&] { s->sim.sync(); });
}

extern "C" size_t xf_sim_current_step(const xf_sim *s) { return s->sim.currentTimeStep(); }

extern "C" int xf_sim_dims(const xf_sim *s, size_t out[3]) {
  out[0] = s->sim.getNx();
  out[1] = s->sim.getNy();
  out[2] = s->sim.getNz();
  return 0;
}

extern "C" double xf_sim_dt(const xf_sim *s) { return s->sim.getDt(); }

extern "C" int xf_sim_get_field(xf_sim *s, int comp, double *out) {
  return guarded([&] {
    if (comp < 0 || comp > 5) throw std::invalid_argument("bad component");
    auto v = s->sim.getField(static_cast<EMComponent>(comp));
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

extern "C" int xf_sim_set_field(xf_sim *s, int comp, const double *in) {
  return guarded([&] {
    if (comp < 0 || comp > 5) throw std::invalid_argument("bad component");
    const auto sh = s->sim.fieldShape(static_cast<EMComponent>(comp));
    std::vector<double> v(in, in + sh[0] * sh[1] * sh[2]);
    s->sim.setField(static_cast<EMComponent>(comp), v);
  });
}

extern "C" int64_t xf_sim_nffft_size(const xf_sim *s) {
  return s->sim.engine() ? xf_engine_nffft_size(s->sim.engine()) : 0;
}

extern "C" int xf_sim_get_nffft(xf_sim *s, double *out) {
  return guarded([&] {
    auto v = s->sim.getNffftAccumulators();
    std::memcpy(out, v.data(), v.size() * sizeof(std::complex<double>));
  });
}

extern "C" int xf_sim_farfield(xf_sim *s, double *out) {
  return guarded([&] {
    FarField r = s->sim.calculateFarField();
    const size_t M = r.e_theta.size();
    const std::vector<std::complex<double>> *c[8] = {&r.n_theta, &r.n_phi, &r.l_theta, &r.l_phi,
                                                     &r.e_theta, &r.e_phi, &r.h_theta, &r.h_phi};
    for (int q = 0; q < 8; ++q) std::memcpy(out + 2 * M * q, c[q]->data(), M * sizeof(std::complex<double>));
    std::memcpy(out + 16 * M, r.power_theta.data(), M * sizeof(double));
    std::memcpy(out + 17 * M, r.power_phi.data(), M * sizeof(double));
  });
}

extern "C" int64_t xf_sim_broadband_samples_size(const xf_sim *s) {
  return s->sim.engine() ? xf_engine_broadband_samples_size(s->sim.engine()) : 0;
}

extern "C" int xf_sim_get_broadband_samples(xf_sim *s, double *out) {
  return guarded([&] {
    auto v = s->sim.getBroadBandSamples();
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

extern "C" int xf_sim_broadband_farfield(xf_sim *s, const double *wave_number, double *out) {
  return guarded([&] {
    const size_t Ns = 2 * s->sim.totalTimeSteps() - 1;
    std::vector<double> k;
    if (wave_number) k.assign(wave_number, wave_number + Ns);
    BroadBandFarField r = s->sim.calculateBroadBandFarField(wave_number ? &k : nullptr);
    std::memcpy(out, r.frequencies.data(), Ns * sizeof(double));
    std::memcpy(out + Ns, r.wave_number.data(), Ns * sizeof(double));
    const std::vector<std::complex<double>> *c[14] = {&r.n_x,     &r.n_y,     &r.n_z,     &r.l_x,   &r.l_y,
                                                      &r.l_z,     &r.n_theta, &r.n_phi,   &r.l_theta,
                                                      &r.l_phi,   &r.e_theta, &r.e_phi,   &r.h_theta,
                                                      &r.h_phi};
    for (int q = 0; q < 14; ++q)
      std::memcpy(out + 2 * Ns + 2 * Ns * q, c[q]->data(), Ns * sizeof(std::complex<double>));
    std::memcpy(out + 30 * Ns, r.power_theta.data(), Ns * sizeof(double));
    std::memcpy(out + 31 * Ns, r.power_phi.data(), Ns * sizeof(double));
  });
}

extern "C" int xf_sim_broadband_ez(xf_sim *s, double *out) {
  return guarded([&] {
    BroadBandFarField r = s->sim.calculateBroadBandFarField(nullptr);
    if (r.ez.empty()) throw std::runtime_error("the far-zone Ez is defined on 2-D grids only");
    std::memcpy(out, r.ez.data(), r.ez.size() * sizeof(std::complex<double>));
  });
}

extern "C" int xf_sim_output_nffft(xf_sim *s) {
  return guarded([&] { s->sim.outputData(); });
}

extern "C" int xf_sim_get_probes(xf_sim *s, double *out) {
  return guarded([&] {
    auto v = s->sim.getProbeData();
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

extern "C" struct xf_engine *xf_sim_engine(xf_sim *s) { return s->sim.engine(); }

extern "C" int xf_sim_set_design(xf_sim *s, const size_t box[6], const float *eps_r) {
  return guarded([&] {
    std::array<size_t, 6> b;
    for (int q = 0; q < 6; ++q) b[q] = box[q];
    if (b[1] <= b[0] || b[3] <= b[2] || b[5] <= b[4]) throw std::invalid_argument("empty design box");
    const size_t nd = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]);
    s->sim.setDesign(b, std::vector<float>(eps_r, eps_r + nd));
  });
}

extern "C" int xf_sim_update_design(xf_sim *s, const float *eps_r) {
  return guarded([&] {
    const size_t nd = s->sim.designCells();
    s->sim.updateDesign(std::vector<float>(eps_r, eps_r + nd));
  });
}

extern "C" int xf_sim_set_dft_plane(xf_sim *s, int comp, int axis, size_t index, size_t lo1, size_t hi1,
                                    size_t lo2, size_t hi2, double freq) {
  return guarded([&] {
    if (comp < 0 || comp > 5) throw std::invalid_argument("bad component");
    s->sim.setDftPlane(static_cast<EMComponent>(comp), axis, index, lo1, hi1, lo2, hi2, freq);
  });
}

extern "C" int xf_sim_get_dft(xf_sim *s, double *out) {
  return guarded([&] {
    auto v = s->sim.getDft();
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  });
}

extern "C" int xf_sim_grad_eps(xf_sim *s, int checkpoint_every, double *grad, double *J) {
  return guarded([&] {
    std::vector<double> g;
    *J = s->sim.gradEps(g, checkpoint_every);
    std::memcpy(grad, g.data(), g.size() * sizeof(double));
  });
}

extern "C" int xf_sim_adjoint_run(xf_sim *s, size_t nsteps, double *fields, size_t n) {
  return guarded([&] {
    std::vector<double> v(fields, fields + n);
    s->sim.adjointRun(nsteps, v);
    std::memcpy(fields, v.data(), n * sizeof(double));
  });
}

extern "C" int xf_sim_checkpoint(xf_sim *s, int slot) {
  return guarded([&] { s->sim.checkpoint(slot); });
}

extern "C" int xf_sim_restore(xf_sim *s, int slot) {
  return guarded([&] { s->sim.restore(slot); });
}
