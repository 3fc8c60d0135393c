// host.cpp -- CPU set-up of the reference's Simulation, handed to the engine.
//
// Everything here runs once per simulation in FP64 with the reference's
// formulas and operand order (compiled with -ffp-contract=off), then the
// per-step work goes through include/xfdtd_engine.h.  Citations are to the
// reference's src/ tree.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "xfdtd/xfdtd.hpp"
#include "xfdtd_engine.h"

namespace xfdtd {

namespace {

bool isLessOrEqual(double a, double b) {
  // util/float_compare.h:14-17
  return (std::abs(a - b) < constant::TOLERABLE_EPSILON) ? true : (a < b);
}
bool isGreaterOrEqual(double a, double b) {
  return (std::abs(a - b) < constant::TOLERABLE_EPSILON) ? true : (a > b);
}

void check(xf_status s) {
  if (s != XF_OK) throw std::runtime_error(std::string("xfdtd engine: ") + xf_last_error());
}

int compIndex(EMComponent c) { return static_cast<int>(c); }

}  // namespace

// Host-side tables of one simulation, computed by Simulation::plan and
// handed to the engine by Simulation::upload.
struct PlanData {
  int nmat = 0;
  std::vector<uint8_t> midx;
  std::vector<double> ca, cb[12], cpsi[12];
  int dn[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
  std::vector<double> be[3], ae[3], bm[3], am[3];
  struct Src {
    int kind;  // 0 source term, 1 inductor
    int comp, is, ie, js, je, ks, ke;
    std::vector<double> coff, wave, c2;
  };
  std::vector<Src> srcs;
  struct Tfsf {
    bool present = false;
    xf_tfsf_desc desc{};
    std::vector<double> full[3], half[3], source;
  } tfsf;
  struct Nf {
    bool present = false;
    xf_nffft_desc desc{};
    std::vector<std::complex<double>> tj, tm;
  } nf;
  struct Bb {
    bool present = false;
    int64_t box[6] = {};  // li, ri, lj, rj, lk, rk
  } bb;
  struct Design {
    bool present = false;
    int64_t box[6] = {};
    std::vector<double> coef, eps_r;
  } design;
  struct Dft {
    bool present = false;
    int comp = 0, axis = 0;
    int64_t index = 0, lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0;
    std::vector<std::complex<double>> w;
  } dft;
  struct Drude {
    bool present = false;
    int64_t box[6] = {};
    std::vector<double> coef;  // [nmat][6]: ca, cb, 1 + k, k, beta, dl/2
  } drude;
};

// shape/cube.cpp:27-37
bool Cube::isPointInside(const PointVector &p) const {
  return isLessOrEqual(_point[0], p[0]) && isLessOrEqual(_point[1], p[1]) &&
         isLessOrEqual(_point[2], p[2]) && isGreaterOrEqual(_point[0] + _size[0], p[0]) &&
         isGreaterOrEqual(_point[1] + _size[1], p[1]) && isGreaterOrEqual(_point[2] + _size[2], p[2]);
}

// shape/sphere.cpp:27-36
bool Sphere::isPointInside(const PointVector &p) const {
  const double dx = p[0] - _center[0], dy = p[1] - _center[1], dz = p[2] - _center[2];
  const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  return isLessOrEqual(dist, _radius);
}
std::unique_ptr<Shape> Sphere::getWrappedBox() const {
  return std::make_unique<Cube>(
      PointVector{_center[0] - _radius, _center[1] - _radius, _center[2] - _radius},
      PointVector{2 * _radius, 2 * _radius, 2 * _radius});
}

// shape/cylinder.cpp:21-42: inside the wrapped box, then within the radius
// of the axis
bool Cylinder::isPointInside(const PointVector &p) const {
  if (!getWrappedBox()->isPointInside(p)) return false;
  const int u = _axis == Axis::X ? 1 : 0, v = _axis == Axis::Z ? 1 : 2;
  const double du = p[u] - _center[u], dv = p[v] - _center[v];
  return isLessOrEqual(std::sqrt(du * du + dv * dv), _radius);
}
// shape/cylinder.cpp:44-60
std::unique_ptr<Shape> Cylinder::getWrappedBox() const {
  PointVector lo, size;
  const int ax = static_cast<int>(_axis);
  for (int a = 0; a < 3; ++a) {
    lo[a] = _center[a] - (a == ax ? _height / 2 : _radius);
    size[a] = a == ax ? _height : 2 * _radius;
  }
  return std::make_unique<Cube>(lo, size);
}

// waveform/gaussian_waveform.cpp:18-24
double GaussianWaveform::getValueByTime(double time) const {
  return getAmplitude() * std::exp(-std::pow((time - _t_0) / _tau, 2));
}
double SinusoidalWaveform::getValueByTime(double time) const {
  return getAmplitude() * std::sin(2 * constant::PI * _frequency * time + _phase);
}
double CosineModulatedGaussianWaveform::getValueByTime(double time) const {
  return getAmplitude() * std::cos(2 * constant::PI * _f * (time - _t_0)) *
         std::exp(-std::pow((time - _t_0) / _tau, 2));
}

// util/dft.h:11-26
std::vector<std::complex<double>> dft(const std::vector<double> &x, double dt,
                                      const std::vector<double> &frequencies) {
  using namespace std::complex_literals;
  std::vector<std::complex<double>> res(frequencies.size());
  for (size_t i = 0; i < frequencies.size(); ++i) {
    std::complex<double> sum{0.0, 0.0};
    for (size_t j = 0; j < x.size(); ++j) {
      const double t = (double)j * dt;
      sum += x[j] * std::exp(-2.0 * 1i * constant::PI * frequencies[i] * t);
    }
    res[i] = sum * dt;
  }
  return res;
}

// network/port.cpp:19-29
void Port::calculateSParameter(const std::vector<double> &frequencies, double dt) {
  const std::complex<double> z = _impedance;
  const double k = std::sqrt(std::real(z));
  const auto v = dft(_voltage->getValueArray(), dt, frequencies);
  const auto i = dft(_current->getValueArray(), dt, frequencies);
  _a.resize(v.size());
  _b.resize(v.size());
  for (size_t q = 0; q < v.size(); ++q) {
    _a[q] = 0.5 * (v[q] + z * i[q]) / k;
    _b[q] = 0.5 * (v[q] - std::conj(z) * i[q]) / k;
  }
}

// network/network.cpp:27-45
std::map<int, std::vector<std::complex<double>>> Network::calculateSParameters(double dt) {
  for (auto &p : _ports) p->calculateSParameter(_frequencies, dt);
  std::map<int, std::vector<std::complex<double>>> s;
  for (size_t i = 0; i < _ports.size(); ++i) {
    if (!_ports[i]->isActivatedSource()) continue;
    for (size_t j = 0; j < _ports.size(); ++j) {
      const auto &b = _ports[j]->getB();
      const auto &a = _ports[i]->getA();
      std::vector<std::complex<double>> v(b.size());
      for (size_t q = 0; q < b.size(); ++q) v[q] = b[q] / a[q];
      s[_ports[j]->getPortIndex() * 10 + _ports[i]->getPortIndex()] = std::move(v);
    }
  }
  return s;
}

void Waveform::init(const std::vector<double> &time_array) {
  _value.resize(time_array.size());
  for (size_t q = 0; q < time_array.size(); ++q) _value[q] = getValueByTime(time_array[q]);
}

// perfect_match_layer.cpp:133-155 (P faces) and 302-325 (N faces)
PML::Profile PML::computeProfile(double dl, double dt) const {
  Profile p;
  const int d = _thickness;
  p.b_e.resize(d); p.a_e.resize(d); p.b_m.resize(d); p.a_m.resize(d);
  p.kappa_e.resize(d); p.kappa_m.resize(d);
  const double e2m{constant::MU_0 / constant::EPSILON_0};
  const double sigma_max = _sigma_ratio * (_order + 1) / (150.0 * constant::PI * dl);
  const bool pos = isOrientatingPositive();
  for (int i = 0; i < d; ++i) {
    double rho_e, rho_m;
    if (pos) {
      rho_e = (i + 1 - 0.75) / static_cast<double>(d);
      rho_m = (i + 1 - 0.25) / static_cast<double>(d);
    } else {
      rho_e = (d - i - 0.75) / static_cast<double>(d);
      rho_m = (d - i - 0.25) / static_cast<double>(d);
    }
    const double sigma_e = sigma_max * std::pow(rho_e, _order);
    double sigma_m = sigma_max * std::pow(rho_m, _order);
    sigma_m = sigma_m * e2m;
    const double kappa_e = 1 + (_kappa_max - 1) * std::pow(rho_e, _order);
    const double kappa_m = 1 + (_kappa_max - 1) * std::pow(rho_m, _order);
    const double alpha_e = _alpha_min + (_alpha_max - _alpha_min) * (1 - rho_e);
    double alpha_m = _alpha_min + (_alpha_max - _alpha_min) * (1 - rho_m);
    alpha_m = alpha_m * e2m;
    p.b_e[i] = std::exp((-dt / constant::EPSILON_0) * (sigma_e / kappa_e + alpha_e));
    p.a_e[i] = (1 / dl) * (p.b_e[i] - 1) * sigma_e / (kappa_e * (sigma_e + kappa_e * alpha_e));
    p.b_m[i] = std::exp((-dt / constant::MU_0) * (sigma_m / kappa_m + alpha_m));
    p.a_m[i] = (1 / dl) * (p.b_m[i] - 1) * sigma_m / (kappa_m * (sigma_m + kappa_m * alpha_m));
    p.kappa_e[i] = kappa_e;
    p.kappa_m[i] = kappa_m;
  }
  return p;
}

// tfsf/tfsf.cpp:15-33
TFSF::TFSF(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z, double e_0,
           double theta_inc, double phi_inc, double psi, std::shared_ptr<Waveform> waveform)
    : _distance_x{distance_x}, _distance_y{distance_y}, _distance_z{distance_z}, _e_0{e_0},
      _theta_inc{theta_inc}, _phi_inc{phi_inc}, _psi{psi}, _sin_theta_inc{std::sin(theta_inc)},
      _cos_theta_inc{std::cos(theta_inc)}, _sin_phi_inc{std::sin(phi_inc)},
      _cos_phi_inc{std::cos(phi_inc)}, _sin_psi{std::sin(psi)}, _cos_psi{std::cos(psi)},
      _k{PointVector{_sin_theta_inc * _cos_phi_inc, _sin_theta_inc * _sin_phi_inc, _cos_theta_inc}},
      _waveform{std::move(waveform)} {}

// util/dft.h:11-27
std::vector<std::complex<double>> TFSF::getIncidentWaveFourierTransform(
    const std::vector<double> &frequencies) const {
  using namespace std::complex_literals;
  const auto &v = _waveform->getValue();
  std::vector<std::complex<double>> res(frequencies.size());
  for (size_t i = 0; i < frequencies.size(); ++i) {
    std::complex<double> sum{0.0, 0.0};
    for (size_t j = 0; j < v.size(); ++j) {
      auto t{j * _dt};
      sum += v[j] * std::exp(-2.0 * 1i * constant::PI * frequencies[i] * t);
    }
    res[i] = sum * _dt;
  }
  return res;
}

std::pair<std::vector<double>, std::vector<std::complex<double>>> TFSF::getIncidentWaveFastFourierTransform()
    const {
  const auto &w = _waveform->getValue();
  const size_t T = w.size(), Ns = 2 * T - 1;
  std::pair<std::vector<double>, std::vector<std::complex<double>>> r;
  r.first.resize(Ns / 2 + 1);
  for (size_t q = 0; q < r.first.size(); ++q) r.first[q] = (double)q / ((double)Ns * _dt);
  r.second.resize(Ns);
  for (size_t q = 0; q < Ns; ++q) {
    // bin m = q - (T - 1) of the shifted transform, X(m) = sum_n w[n] e^{-2 pi i m n / Ns}
    const long long m = (long long)q - (long long)(T - 1);
    const size_t mm = (size_t)((m % (long long)Ns + (long long)Ns) % (long long)Ns);
    double re = 0, im = 0;
    for (size_t n = 0; n < T; ++n) {
      const double a = -2 * constant::PI * (double)((mm * n) % Ns) / (double)Ns;
      re += w[n] * std::cos(a);
      im += w[n] * std::sin(a);
    }
    r.second[q] = {re, im};
  }
  return r;
}

// tfsf/tfsf_3d.cpp:13-31
TFSF3D::TFSF3D(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z,
               double e_0, double theta_inc, double phi_inc, double psi,
               std::shared_ptr<Waveform> waveform)
    : TFSF(distance_x, distance_y, distance_z, e_0, theta_inc, phi_inc, psi, std::move(waveform)) {
  const double sp = _sin_phi_inc, cp = _cos_phi_inc, st = _sin_theta_inc, ct = _cos_theta_inc;
  const double u[3][3] = {{-sp, ct * cp, st * cp}, {cp, ct * sp, st * sp}, {0, -st, ct}};
  const double ke[3] = {_sin_psi, _cos_psi, 0};
  for (int r = 0; r < 3; ++r) _transform_e[r] = u[r][0] * ke[0] + u[r][1] * ke[1] + u[r][2] * ke[2];
  const PointVector &k = _k;
  const PointVector &e = _transform_e;
  _transform_h = {k[1] * e[2] - k[2] * e[1], k[2] * e[0] - k[0] * e[2], k[0] * e[1] - k[1] * e[0]};
}

// tfsf/tfsf_2d.cpp:10-24 (TFSF with theta = pi/2, psi = 0).  The
// polarisation (sin psi, cos psi) is taken as (sin psi, cos psi, 0): the
// reference builds it as a 3-vector from two values, whose intended third
// entry is 0.
TFSF2D::TFSF2D(SpatialIndex distance_x, SpatialIndex distance_y, double phi_inc, double ez_i0,
               std::shared_ptr<Waveform> waveform)
    : TFSF(distance_x, distance_y, 0, ez_i0, constant::PI / 2, phi_inc, 0, std::move(waveform)) {
  const double sp = _sin_phi_inc, cp = _cos_phi_inc, st = _sin_theta_inc, ct = _cos_theta_inc;
  const double u[3][3] = {{-sp, ct * cp, st * cp}, {cp, ct * sp, st * sp}, {0, -st, ct}};
  const double ke[3] = {_sin_psi, _cos_psi, 0};
  for (int r = 0; r < 3; ++r) _transform_e[r] = u[r][0] * ke[0] + u[r][1] * ke[1] + u[r][2] * ke[2];
  const PointVector &k = _k;
  const PointVector &e = _transform_e;
  _transform_h = {k[1] * e[2] - k[2] * e[1], k[2] * e[0] - k[0] * e[2], k[0] * e[1] - k[1] * e[0]};
}

// nffft/nffft_broadband.cpp:21-28
NffftBroadBand::NffftBroadBand(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z,
                               double theta, double phi, std::string output_dir_path)
    : NFFFT(distance_x, distance_y, distance_z, std::move(output_dir_path)), _theta{theta}, _phi{phi} {}

// include/nffft/nffft_fd.h constructor
NffftFd::NffftFd(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z,
                 std::vector<double> frequencies, std::vector<double> theta,
                 std::vector<double> phi, std::string output_dir_path)
    : NFFFT(distance_x, distance_y, distance_z, std::move(output_dir_path)),
      _frequencies{std::move(frequencies)}, _theta{std::move(theta)}, _phi{std::move(phi)} {
  _wave_number.resize(_frequencies.size());
  for (size_t q = 0; q < _frequencies.size(); ++q)
    _wave_number[q] = 2 * constant::PI * _frequencies[q] / constant::C_0;
}

// ------------------------------------------------------------- Simulation

Simulation::Simulation(double cell_size, float cfl) : _cell_size{cell_size}, _cfl{cfl} {
  // grid/grid_space.cpp:14-27
  for (int a = 0; a < 3; ++a) {
    _min[a] = std::numeric_limits<double>::max();
    _max[a] = std::numeric_limits<double>::min();
  }
}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<Boundary>> boundaries,
                       std::unique_ptr<TFSF> tfsf, float cfl)
    : Simulation(cell_size, cfl) {
  _objects = std::move(objects);
  _boundaries = std::move(boundaries);
  _tfsf = std::move(tfsf);
}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<Boundary>> boundaries,
                       std::unique_ptr<TFSF> tfsf, std::unique_ptr<NFFFT> nffft, float cfl)
    : Simulation(cell_size, std::move(objects), std::move(boundaries), std::move(tfsf), cfl) {
  _nffft = std::move(nffft);
}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
                       std::vector<std::shared_ptr<Boundary>> boundaries,
                       std::vector<std::shared_ptr<Monitor>> monitors, float cfl)
    : Simulation(cell_size, cfl) {
  _objects = std::move(objects);
  _lumped_elements = std::move(lumped_elements);
  _boundaries = std::move(boundaries);
  _monitors = std::move(monitors);
}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
                       std::vector<std::shared_ptr<Monitor>> monitors, float cfl)
    : Simulation(cell_size, std::move(objects), std::move(lumped_elements), {}, std::move(monitors), cfl) {}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<Monitor>> monitors, float cfl)
    : Simulation(cell_size, cfl) {
  _objects = std::move(objects);
  _monitors = std::move(monitors);
}
Simulation::Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
                       std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
                       std::vector<std::shared_ptr<Boundary>> boundaries,
                       std::vector<std::shared_ptr<Monitor>> monitors, std::unique_ptr<Network> network,
                       float cfl)
    : Simulation(cell_size, std::move(objects), std::move(lumped_elements), std::move(boundaries),
                 std::move(monitors), cfl) {
  _network = std::move(network);
}
Simulation::~Simulation() {
  if (_engine) xf_engine_destroy(_engine);
}

void Simulation::addNetwork(std::shared_ptr<Network> network) { _network = std::move(network); }

void Simulation::addObject(std::shared_ptr<Object> o) { _objects.push_back(std::move(o)); }
void Simulation::addBoundary(std::shared_ptr<Boundary> b) { _boundaries.push_back(std::move(b)); }
void Simulation::addLumpedElement(std::shared_ptr<LumpedElement> l) {
  _lumped_elements.push_back(std::move(l));
}
void Simulation::addTFSFSource(std::shared_ptr<TFSF> t) { _tfsf = std::move(t); }
void Simulation::addNFFFT(std::shared_ptr<NFFFT> n) { _nffft = std::move(n); }
void Simulation::addProbe(Probe p) { _probes.push_back(p); }
void Simulation::addMonitor(std::shared_ptr<Monitor> m) {
  if (_plan) throw std::runtime_error("Simulation already planned");
  _monitors.push_back(std::move(m));
}

// grid/grid_space.cpp:576-618
void Simulation::resizeSpaceSize(const double nmin[3], const double nmax[3]) {
  for (int a = 0; a < 3; ++a) {
    if (_min[a] > nmin[a]) _min[a] = nmin[a];
    if (_max[a] < nmax[a]) _max[a] = nmax[a];
  }
  for (int a = 0; a < 3; ++a) {
    size_t nn = static_cast<size_t>(std::round((_max[a] - _min[a]) / _cell_size));
    if (_max[a] < _min[a]) nn = 1;
    _n[a] = std::max(_n[a], nn);
  }
}

// simulation_init.cpp:105-146 with grid_space.cpp:364-549, 625-638, 737-757
void Simulation::initGridSpace() {
  if (_objects.empty()) throw std::runtime_error("Simulation has no objects");
  for (const auto &o : _objects) {
    auto box = o->getWrappedBox();
    auto *c = dynamic_cast<Cube *>(box.get());
    if (!c) throw std::runtime_error("Shape Type Error");
    const double mn[3] = {c->getXmin(), c->getYmin(), c->getZmin()};
    const double mx[3] = {c->getXmax(), c->getYmax(), c->getZmax()};
    resizeSpaceSize(mn, mx);
  }
  for (const auto &b : _boundaries) {
    const int ori = static_cast<int>(b->getOrientation());
    const int a = ori / 2;
    const double size = b->getSize() * _cell_size;
    if (_n[a] <= 1) throw std::runtime_error("GridSpace is not extended");
    double mn[3] = {_min[0], _min[1], _min[2]}, mx[3] = {_max[0], _max[1], _max[2]};
    if (ori % 2 == 0) mn[a] = _min[a] - size; else mx[a] = _max[a] + size;
    resizeSpaceSize(mn, mx);
  }
  // grid/grid_space.cpp:501-536: THREE_DIMENSION, or TWO_DIMENSION (the
  // TMz branch of simulation_update.cpp:157-168) with one z cell
  for (int a = 0; a < 3; ++a) _n[a] = std::max<size_t>(_n[a], 1);
  _two_d = _n[2] == 1;
  if (_n[0] <= 1 || _n[1] <= 1)
    throw std::runtime_error("only 3-D grids and 2-D grids with one z cell are supported");
  for (int a = 0; a < 3; ++a) {
    const size_t n = _n[a];
    _ec[a].resize(n + 1); _hc[a].resize(n); _es[a].resize(n); _hs[a].resize(n + 1);
    for (size_t i = 0; i <= n; ++i) _ec[a][i] = _min[a] + static_cast<double>(i) * _cell_size;
    for (size_t i = 0; i < n; ++i) _hc[a][i] = (_ec[a][i] + _ec[a][i + 1]) / 2;
    for (size_t i = 0; i < n; ++i) _es[a][i] = _ec[a][i + 1] - _ec[a][i];
    for (size_t i = 0; i <= n; ++i) {
      const double lo = (i == 0) ? _hc[a][0] - _cell_size : _hc[a][i - 1];
      const double hi = (i == n) ? _hc[a][n - 1] + _cell_size : _hc[a][i];
      _hs[a][i] = hi - lo;
    }
  }
}

// grid/grid_space.cpp:551-574
size_t Simulation::handleTransform(int a, double x) const {
  if (x < _min[a]) throw std::out_of_range("coordinate is out of range");
  size_t best = 0;
  double bv = std::abs(_ec[a][0] - x);
  for (size_t i = 1; i < _ec[a].size(); ++i) {
    const double v = std::abs(_ec[a][i] - x);
    if (v < bv) { bv = v; best = i; }
  }
  return best;
}

namespace {

struct Mat {
  double eps, mu, sigma, sigma_m;
};

// fdtd_basic_coff.cpp:77-91 for one (material, node size) pair
double coffSelfE(const Mat &m, double dt) {
  return (2 * m.eps - dt * m.sigma) / (2 * m.eps + dt * m.sigma);
}
double coffSelfH(const Mat &m, double dt) {
  return (2 * m.mu - dt * m.sigma_m) / (2 * m.mu + dt * m.sigma_m);
}
double coffEa(const Mat &m, double dt, double db) {  // cecha
  return -(2 * dt / db) / (2 * m.eps + m.sigma * dt);
}
double coffEb(const Mat &m, double dt, double da) {  // cechb
  return (2 * dt / da) / (2 * m.eps + m.sigma * dt);
}
double coffHa(const Mat &m, double dt, double db) {  // chcea
  return (2 * dt / db) / (2 * m.mu + m.sigma_m * dt);
}
double coffHb(const Mat &m, double dt, double da) {  // chceb
  return -(2 * dt / da) / (2 * m.mu + m.sigma_m * dt);
}

}  // namespace

void Simulation::plan(size_t total_time_steps) {
  if (_plan) throw std::runtime_error("Simulation already planned");
  _plan = std::make_shared<PlanData>();
  _total_time_steps = total_time_steps;
  // simulation_init.cpp:10-22
  initGridSpace();
  planMonitors();
  const int nx = (int)_n[0], ny = (int)_n[1], nz = (int)_n[2];
  const int N[3] = {nx, ny, nz};
  // fdtd_basic_coff.cpp:11-29, 342-346: cfl arrives as float
  const double cfl = static_cast<double>(_cfl);
  const double dx = _cell_size, dy = _cell_size, dz = _cell_size;
  if (_two_d)  // fdtd_basic_coff.cpp:13-19, 347-349
    _dt = cfl / (constant::C_0 * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy)));
  else
    _dt = cfl / (constant::C_0 * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)));
  const double dt = _dt;
  if (_two_d) {
    if (!_lumped_elements.empty() || _has_design || _has_dft || !_line_plans.empty() ||
        (_nffft && !dynamic_cast<NffftBroadBand *>(_nffft.get())))
      throw std::runtime_error(
          "a 2-D grid takes objects, PML on x and y, TFSF2D, NffftBroadBand, field monitors and probes only");
    if (_tfsf && !dynamic_cast<TFSF2D *>(_tfsf.get())) throw std::runtime_error("a 2-D grid needs TFSF2D");
  } else if (_tfsf && dynamic_cast<TFSF2D *>(_tfsf.get())) {
    throw std::runtime_error("TFSF2D needs a 2-D grid");
  }

  // ---- materials: 0 = default vacuum of the staggered arrays
  // (fdtd_basic_coff.cpp:313-339), 1..n = objects in insertion order
  std::vector<Mat> mats;
  mats.push_back({constant::EPSILON_0, constant::MU_0, 0.0, 0.0});
  bool any_drude = false;
  for (const auto &o : _objects) {
    const Material &m = o->getMaterial();
    any_drude = any_drude || m.isDispersion();
    mats.push_back({m.getPermittivityE(), m.getPermeabilityM(), m.getElectricalConductivity(),
                    m.getMagneticConductivity()});
  }
  // material index over (nx+1)(ny+1)(nz+1): object/object.cpp:53-73 evaluates
  // each shape at the cell centres (h-node coordinates) of the (nx,ny,nz)
  // cells; later objects overwrite earlier ones.  Only cells whose centre lies
  // in the shape's wrapped box (plus one cell) are tested: outside it every
  // shape here reports "outside".
  const size_t sj = (size_t)nz + 1, si = (size_t)(ny + 1) * sj;
  std::vector<uint8_t> midx((size_t)(nx + 1) * si, 0);
  // Drude cells: object/object.cpp:109-115 (the intended tagging that the
  // reference leaves commented out) marks the cells of a dispersive object's
  // mask with its index; later dispersive objects overwrite earlier ones and
  // other objects leave the mark alone.  dmask holds 1 + object index.
  std::vector<uint8_t> dmask(any_drude ? midx.size() : 0, 0);
  if (_objects.size() > 254) throw std::runtime_error("more than 254 objects");
  // ObjectPlane sigma_e per E component and cell (NaN: none); a later
  // object whose cell mask covers the cell resets it, as its arrays do
  std::map<size_t, std::array<double, 3>> psig;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  // z of a cell centre: 0 on a 2-D grid (grid/grid_space.cpp:241-247)
  auto czc = [&](size_t k) { return _two_d ? 0.0 : _hc[2][k]; };
  for (size_t q = 0; q < _objects.size(); ++q) {
    const Shape *sh = _objects[q]->getShape();
    auto box = sh->getWrappedBox();
    auto *c = dynamic_cast<Cube *>(box.get());
    const double lo[3] = {c->getXmin(), c->getYmin(), c->getZmin()};
    const double hi[3] = {c->getXmax(), c->getYmax(), c->getZmax()};
    if (dynamic_cast<const ObjectPlane *>(_objects[q].get())) {
      // object/object_plane.cpp:8-50: the two tangential components on the
      // plane, [s, e) along one and [s, e] along the other in-plane axis
      size_t st[3], en[3];
      for (int a = 0; a < 3; ++a) {
        st[a] = handleTransform(a, lo[a]);
        en[a] = handleTransform(a, hi[a]);
      }
      int n = -1;
      for (int a = 2; a >= 0; --a)
        if (st[a] == en[a]) n = a;
      if (n < 0) throw std::runtime_error("ObjectPlane " + _objects[q]->getName() + " is not a plane");
      const double sig = _objects[q]->getMaterial().getElectricalConductivity();
      for (int cc = 0; cc < 3; ++cc) {
        if (cc == n) continue;
        size_t r0[3], r1[3];
        for (int a = 0; a < 3; ++a) {
          r0[a] = st[a];
          r1[a] = a == n ? st[a] + 1 : (a == cc ? en[a] : en[a] + 1);
        }
        for (size_t i = r0[0]; i < r1[0]; ++i)
          for (size_t j = r0[1]; j < r1[1]; ++j)
            for (size_t k = r0[2]; k < r1[2]; ++k) {
              auto it = psig.emplace(i * si + j * sj + k, std::array<double, 3>{nan, nan, nan}).first;
              it->second[cc] = sig;
            }
      }
      continue;
    }
    for (auto it = psig.begin(); it != psig.end();) {
      const size_t cidx = it->first, i = cidx / si, j = (cidx % si) / sj, k = cidx % sj;
      if ((int)i < nx && (int)j < ny && (int)k < nz && sh->isPointInside({_hc[0][i], _hc[1][j], czc(k)}))
        it = psig.erase(it);
      else
        ++it;
    }
    int r0[3], r1[3];
    for (int a = 0; a < 3; ++a) {
      r0[a] = 0;
      r1[a] = N[a];
      while (r0[a] < N[a] && _hc[a][r0[a]] < lo[a] - _cell_size) ++r0[a];
      while (r1[a] > 0 && _hc[a][r1[a] - 1] > hi[a] + _cell_size) --r1[a];
    }
    const bool drude = _objects[q]->getMaterial().isDispersion();
    for (int i = r0[0]; i < r1[0]; ++i)
      for (int j = r0[1]; j < r1[1]; ++j)
        for (int k = r0[2]; k < r1[2]; ++k)
          if (sh->isPointInside({_hc[0][i], _hc[1][j], czc(k)})) {
            midx[(size_t)i * si + (size_t)j * sj + k] = (uint8_t)(q + 1);
            if (drude) dmask[(size_t)i * si + (size_t)j * sj + k] = (uint8_t)(q + 1);
          }
  }

  // ---- design box: the cells keep slot 0 (vacuum mu, no loss) for H and
  // carry per-cell E coefficients (set below)
  if (_has_design) {
    for (int a = 0; a < 3; ++a)
      if (_design_box[2 * a + 1] > (size_t)N[a] || _design_box[2 * a] >= _design_box[2 * a + 1])
        throw std::runtime_error("design box outside the grid");
    for (size_t i = _design_box[0]; i < _design_box[1]; ++i)
      for (size_t j = _design_box[2]; j < _design_box[3]; ++j)
        for (size_t k = _design_box[4]; k < _design_box[5]; ++k) {
          if (any_drude && dmask[i * si + j * sj + k])
            throw std::runtime_error("the design box overlaps Drude cells");
          midx[i * si + j * sj + k] = 0;
        }
  }

  // ---- Drude slots: a Drude cell whose material slot is its own Drude
  // object's uses that slot; one covered by a later plain object gets a slot
  // with that object's standard coefficients and the Drude object's ADE
  struct DrudeSlot {
    int m_base, d;
  };
  std::vector<DrudeSlot> dslots;
  std::vector<int> drude_of(mats.size(), 0);  // slot -> 1 + Drude object (0: none)
  int64_t dbox[6] = {N[0], 0, N[1], 0, N[2], 0};
  if (any_drude) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t c = (size_t)i * si + (size_t)j * sj + k;
          const int d = dmask[c];
          if (!d) continue;
          const int mb = midx[c];
          int slot = -1;
          if (mb == d) {
            slot = mb;
          } else {
            for (size_t q = 0; q < dslots.size(); ++q)
              if (dslots[q].m_base == mb && dslots[q].d == d) slot = (int)(mats.size() + q);
            if (slot < 0) {
              slot = (int)(mats.size() + dslots.size());
              dslots.push_back({mb, d});
            }
          }
          midx[c] = (uint8_t)slot;
          const int at[3] = {i, j, k};
          for (int a = 0; a < 3; ++a) {
            dbox[2 * a] = std::min<int64_t>(dbox[2 * a], at[a]);
            dbox[2 * a + 1] = std::max<int64_t>(dbox[2 * a + 1], at[a] + 1);
          }
        }
    for (size_t q = 1; q < mats.size(); ++q)
      if (_objects[q - 1]->getMaterial().isDispersion()) drude_of[q] = (int)q;
  }
  for (size_t q = 0; q < dslots.size(); ++q) {
    mats.push_back(mats[dslots[q].m_base]);
    drude_of.push_back(dslots[q].d);
  }

  // ---- lumped elements in insertion order (lumped_element/
  // {current_source,voltage_source,resistor,capacitor,inductor}.cpp).  A cell
  // whose E coefficients a lumped element or a plane changes gets a material
  // slot of its own kind: its base material with those components replaced.
  struct SrcInfo {
    int kind;  // 0: source term (current or voltage source), 1: inductor
    int comp, is, ie, js, je, ks, ke;
    std::vector<double> coff, wave, c2;
  };
  std::vector<SrcInfo> srcs;
  struct CellOv {
    int base = 0;
    uint8_t kind[3] = {0, 0, 0};  // per component: 0 none, 1 lumped ce/ca/cb, 2 plane sigma
    double v[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  };
  std::map<size_t, CellOv> ovs;
  for (const auto &e : psig) {
    CellOv &o = ovs[e.first];
    o.base = midx[e.first];
    for (int cc = 0; cc < 3; ++cc)
      if (!std::isnan(e.second[cc])) {
        o.kind[cc] = 2;
        o.v[cc][0] = e.second[cc];
      }
  }
  auto cell_sigma = [&](size_t cidx, int comp) {
    auto it = psig.find(cidx);
    if (it != psig.end() && !std::isnan(it->second[comp])) return it->second[comp];
    return mats[midx[cidx]].sigma;
  };
  for (const auto &le : _lumped_elements) {
    enum { CS, VS, RES, CAP, IND } kind;
    auto *cs = dynamic_cast<CurrentSource *>(le.get());
    auto *vs = dynamic_cast<VoltageSource *>(le.get());
    auto *rs = dynamic_cast<Resistor *>(le.get());
    auto *cp = dynamic_cast<Capacitor *>(le.get());
    auto *nd = dynamic_cast<Inductor *>(le.get());
    int a;
    double value;
    if (cs) { kind = CS; a = static_cast<int>(cs->getOrientation()) / 2; value = cs->getResistance(); }
    else if (vs) { kind = VS; a = static_cast<int>(vs->getOrientation()) / 2; value = vs->getResistance(); }
    else if (rs) { kind = RES; a = static_cast<int>(rs->getAxis()); value = rs->getResistance(); }
    else if (cp) { kind = CAP; a = static_cast<int>(cp->getAxis()); value = cp->getCapacitance(); }
    else if (nd) { kind = IND; a = static_cast<int>(nd->getAxis()); value = nd->getInductance(); }
    else throw std::runtime_error("unsupported lumped element");
    const Cube &cube = le->getCube();
    const double lo[3] = {cube.getXmin(), cube.getYmin(), cube.getZmin()};
    const double hi[3] = {cube.getXmax(), cube.getYmax(), cube.getZmax()};
    size_t st[3], en[3];
    for (int b = 0; b < 3; ++b) {
      st[b] = handleTransform(b, lo[b]);
      en[b] = handleTransform(b, hi[b]);
      if (en[b] < st[b]) throw std::range_error("end < start");
      if (en[b] == st[b]) throw std::runtime_error("lumped element box has no cells");
    }
    SrcInfo si_;
    si_.kind = kind == IND ? 1 : 0;
    si_.comp = a;  // XF_EX / XF_EY / XF_EZ
    si_.is = (int)st[0]; si_.ie = (int)en[0]; si_.js = (int)st[1]; si_.je = (int)en[1];
    si_.ks = (int)st[2]; si_.ke = (int)en[2];
    const size_t ni = en[0] - st[0], nj = en[1] - st[1], nk = en[2] - st[2];
    double rf = 0;
    if (kind == CS) {
      // current_source.cpp:34-116
      double af;
      if (a == 0) { rf = value * nk * nj / ni; af = cs->waveform().getAmplitude() / (nk * nj); }
      else if (a == 1) { rf = value * nk * ni / nj; af = cs->waveform().getAmplitude() / (nk * ni); }
      else { rf = value * ni * nj / nk; af = cs->waveform().getAmplitude() / (ni * nj); }
      if (static_cast<int>(cs->getOrientation()) % 2 == 0) af *= -1;
      cs->waveform().setAmplitude(af);
    } else if (kind == VS) {
      // voltage_source.cpp:40-107 (its YP branch tests YN twice and never
      // runs; taken as intended)
      double af;
      if (a == 0) { rf = value * nj * nk / ni; af = vs->waveform().getAmplitude() / ni; }
      else if (a == 1) { rf = value * ni * nk / nj; af = vs->waveform().getAmplitude() / nj; }
      else { rf = value * ni * nj / nk; af = vs->waveform().getAmplitude() / nk; }
      if (static_cast<int>(vs->getOrientation()) % 2 == 0) af *= -1;
      vs->waveform().setAmplitude(af);
    } else if (kind == RES || kind == IND) {
      // resistor.cpp:22-76, inductor.cpp:19-68
      if (a == 0) rf = value * nj * nk / ni;
      else if (a == 1) rf = value * ni * nk / nj;
      else rf = value * ni * nj / nk;
    } else {
      // capacitor.cpp:25-61
      if (a == 0) rf = value * ni / (nj * nk);
      else if (a == 1) rf = value * nj / (ni * nk);
      else rf = value * nk / (ni * nj);
    }
    if (kind == CS || kind == VS) {
      // waveform times xt::linspace(0.5 dt, (T - 0.5) dt, T)
      const size_t T = total_time_steps;
      std::vector<double> times(T);
      const double start = 0.5 * dt, stop = (T - 0.5) * dt;
      const double step = (stop - start) / std::fmax(1.0, (double)T - 1.0);
      for (size_t n = 0; n < T; ++n) times[n] = (n + 1 == T && T > 1) ? stop : start + step * (double)n;
      Waveform &w = kind == CS ? cs->waveform() : vs->waveform();
      w.init(times);
      si_.wave = w.getValue();
    }
    for (size_t i = st[0]; i < en[0]; ++i)
      for (size_t j = st[1]; j < en[1]; ++j)
        for (size_t k = st[2]; k < en[2]; ++k) {
          const size_t cidx = i * si + j * sj + k;
          const int mb = midx[cidx];
          if (drude_of[mb]) throw std::runtime_error("a lumped element on a Drude cell is not supported");
          if (_has_design && i >= _design_box[0] && i < _design_box[1] && j >= _design_box[2] &&
              j < _design_box[3] && k >= _design_box[4] && k < _design_box[5])
            throw std::runtime_error("a lumped element in the design box is not supported");
          const double eps = mats[mb].eps, sg = cell_sigma(cidx, a);
          double da, db, dc;
          // the reference's Y branches of Resistor, Capacitor and Inductor
          // assign _da twice and never _db; the intended (dz, dx) is taken
          if (a == 0) { da = _hs[1][j]; db = _hs[2][k]; dc = _es[0][i]; }
          else if (a == 1) { da = _hs[2][k]; db = _hs[0][i]; dc = _es[1][j]; }
          else { da = _hs[0][i]; db = _hs[1][j]; dc = _es[2][k]; }
          if (kind == IND) {
            si_.c2.push_back(dt * dc / (rf * da * db));
            si_.coff.push_back(-(2 * dt) / (2 * eps + dt * sg));
            continue;
          }
          double ce, ca_, cb_;
          if (kind == CAP) {
            const double beta = 2 * rf * dc / (da * db);
            ce = (2 * eps - dt * sg + beta) / (2 * eps + dt * sg + beta);
            ca_ = -(2 * dt) / ((2 * eps + dt * sg + beta) * db);
            cb_ = (2 * dt) / ((2 * eps + dt * sg + beta) * da);
          } else {
            double beta;
            if (kind == VS) {
              const double alpha = da * db * rf;
              beta = dt * dc / alpha;
              si_.coff.push_back(-2 * dt / ((2 * eps + dt * sg + beta) * alpha));
            } else {
              beta = (dt * dc) / (rf * da * db);
              if (kind == CS) si_.coff.push_back(-2 * dt / ((2 * eps + dt * sg + beta) * (db * da)));
            }
            ce = (2 * eps - dt * sg - beta) / (2 * eps + dt * sg + beta);
            ca_ = -2 * dt / ((2 * eps + dt * sg + beta) * db);
            cb_ = 2 * dt / ((2 * eps + dt * sg + beta) * da);
          }
          CellOv &o = ovs[cidx];
          o.base = mb;
          o.kind[a] = 1;
          o.v[a][0] = ce;
          o.v[a][1] = ca_;
          o.v[a][2] = cb_;
        }
    if (kind != RES && kind != CAP) srcs.push_back(std::move(si_));
  }
  for (const auto &e : ovs)
    if (drude_of[e.second.base]) throw std::runtime_error("an ObjectPlane on a Drude cell is not supported");
  // one slot per distinct (base, overrides): slots of equal overrides share
  // their table entries, so only cells with identical values share a slot
  std::map<std::vector<double>, int> slot_of;
  std::vector<std::pair<size_t, int>> cell_slot;
  std::vector<const CellOv *> slot_ov;
  const int nbase = (int)mats.size();
  for (const auto &e : ovs) {
    const CellOv &o = e.second;
    std::vector<double> key = {(double)o.base};
    for (int cc = 0; cc < 3; ++cc) {
      key.push_back(o.kind[cc]);
      for (int u = 0; u < 3; ++u) key.push_back(o.kind[cc] ? o.v[cc][u] : 0.0);
    }
    auto it = slot_of.find(key);
    int slot;
    if (it == slot_of.end()) {
      slot = nbase + (int)slot_ov.size();
      slot_of.emplace(std::move(key), slot);
      slot_ov.push_back(&o);
    } else {
      slot = it->second;
    }
    cell_slot.emplace_back(e.first, slot);
  }
  const int nmat = nbase + (int)slot_ov.size();
  if (nmat > 255) throw std::runtime_error("more than 255 material slots");

  // ---- coefficient tables (fdtd_basic_coff.cpp:93-121)
  auto axlen = [&](int t) { static const int ax[12] = {2, 1, 0, 2, 1, 0, 2, 1, 0, 2, 1, 0}; return N[ax[t]] + 1; };
  std::vector<double> ca(6 * (size_t)nmat);
  std::vector<std::vector<double>> cb(12), cpsi(12);
  for (int t = 0; t < 12; ++t) {
    cb[t].assign((size_t)nmat * axlen(t), 0.0);
  }
  auto fill_material = [&](int slot, const Mat &m) {
    for (int c = 0; c < 3; ++c) ca[c * nmat + slot] = coffSelfE(m, dt);
    for (int c = 3; c < 6; ++c) ca[c * nmat + slot] = coffSelfH(m, dt);
    for (int k = 0; k <= nz; ++k) {
      cb[XF_CB_EXHY][(size_t)slot * (nz + 1) + k] = coffEa(m, dt, _hs[2][k]);
      cb[XF_CB_EYHX][(size_t)slot * (nz + 1) + k] = coffEb(m, dt, _hs[2][k]);
      if (k < nz) {
        cb[XF_CB_HXEY][(size_t)slot * (nz + 1) + k] = coffHa(m, dt, _es[2][k]);
        cb[XF_CB_HYEX][(size_t)slot * (nz + 1) + k] = coffHb(m, dt, _es[2][k]);
      }
    }
    for (int j = 0; j <= ny; ++j) {
      cb[XF_CB_EXHZ][(size_t)slot * (ny + 1) + j] = coffEb(m, dt, _hs[1][j]);
      cb[XF_CB_EZHX][(size_t)slot * (ny + 1) + j] = coffEa(m, dt, _hs[1][j]);
      if (j < ny) {
        cb[XF_CB_HXEZ][(size_t)slot * (ny + 1) + j] = coffHb(m, dt, _es[1][j]);
        cb[XF_CB_HZEX][(size_t)slot * (ny + 1) + j] = coffHa(m, dt, _es[1][j]);
      }
    }
    for (int i = 0; i <= nx; ++i) {
      cb[XF_CB_EYHZ][(size_t)slot * (nx + 1) + i] = coffEa(m, dt, _hs[0][i]);
      cb[XF_CB_EZHY][(size_t)slot * (nx + 1) + i] = coffEb(m, dt, _hs[0][i]);
      if (i < nx) {
        cb[XF_CB_HYEZ][(size_t)slot * (nx + 1) + i] = coffHa(m, dt, _es[0][i]);
        cb[XF_CB_HZEY][(size_t)slot * (nx + 1) + i] = coffHb(m, dt, _es[0][i]);
      }
    }
  };
  for (int q = 0; q < nbase; ++q) fill_material(q, mats[q]);
  // the override slots: the base material, then each plane component's
  // coefficients from the plane's sigma_e (whole tables), then each lumped
  // component's ce / ca (db) / cb (da) at the cells that use the slot
  auto fill_e_comp = [&](int slot, int comp, const Mat &m) {
    ca[comp * nmat + slot] = coffSelfE(m, dt);
    if (comp == 0) {
      for (int k = 0; k <= nz; ++k) cb[XF_CB_EXHY][(size_t)slot * (nz + 1) + k] = coffEa(m, dt, _hs[2][k]);
      for (int j = 0; j <= ny; ++j) cb[XF_CB_EXHZ][(size_t)slot * (ny + 1) + j] = coffEb(m, dt, _hs[1][j]);
    } else if (comp == 1) {
      for (int i = 0; i <= nx; ++i) cb[XF_CB_EYHZ][(size_t)slot * (nx + 1) + i] = coffEa(m, dt, _hs[0][i]);
      for (int k = 0; k <= nz; ++k) cb[XF_CB_EYHX][(size_t)slot * (nz + 1) + k] = coffEb(m, dt, _hs[2][k]);
    } else {
      for (int j = 0; j <= ny; ++j) cb[XF_CB_EZHX][(size_t)slot * (ny + 1) + j] = coffEa(m, dt, _hs[1][j]);
      for (int i = 0; i <= nx; ++i) cb[XF_CB_EZHY][(size_t)slot * (nx + 1) + i] = coffEb(m, dt, _hs[0][i]);
    }
  };
  for (size_t q = 0; q < slot_ov.size(); ++q) {
    const CellOv &o = *slot_ov[q];
    const int slot = nbase + (int)q;
    fill_material(slot, mats[o.base]);
    for (int cc = 0; cc < 3; ++cc)
      if (o.kind[cc] == 2) {
        Mat m = mats[o.base];
        m.sigma = o.v[cc][0];
        fill_e_comp(slot, cc, m);
      }
  }
  for (const auto &cs_ : cell_slot) {
    const size_t cidx = cs_.first;
    const int slot = cs_.second;
    const CellOv &o = *slot_ov[slot - nbase];
    const int i = (int)(cidx / si), j = (int)((cidx % si) / sj), k = (int)(cidx % sj);
    for (int cc = 0; cc < 3; ++cc) {
      if (o.kind[cc] != 1) continue;
      ca[cc * nmat + slot] = o.v[cc][0];
      if (cc == 0) {
        cb[XF_CB_EXHY][(size_t)slot * (nz + 1) + k] = o.v[cc][1];
        cb[XF_CB_EXHZ][(size_t)slot * (ny + 1) + j] = o.v[cc][2];
      } else if (cc == 1) {
        cb[XF_CB_EYHZ][(size_t)slot * (nx + 1) + i] = o.v[cc][1];
        cb[XF_CB_EYHX][(size_t)slot * (nz + 1) + k] = o.v[cc][2];
      } else {
        cb[XF_CB_EZHX][(size_t)slot * (ny + 1) + j] = o.v[cc][1];
        cb[XF_CB_EZHY][(size_t)slot * (nx + 1) + i] = o.v[cc][2];
      }
    }
    midx[cidx] = (uint8_t)slot;
  }

  // ---- CPML (perfect_match_layer.cpp:49-96, 157-431): per axis, capture
  // c_psi = coefficient * dl, then divide the coefficient by kappa
  std::vector<double> be[3], ae[3], bm[3], am[3];
  int dn[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
  bool seen[6] = {false, false, false, false, false, false};
  for (int t = 0; t < 12; ++t) cpsi[t] = cb[t];
  for (auto &v : cpsi)
    for (double &x : v) x = x * _cell_size;
  // tables whose derivative axis is x / y / z: E pair then H pair
  static const int axis_tables[3][4] = {{XF_CB_EYHZ, XF_CB_EZHY, XF_CB_HYEZ, XF_CB_HZEY},
                                        {XF_CB_EXHZ, XF_CB_EZHX, XF_CB_HXEZ, XF_CB_HZEX},
                                        {XF_CB_EXHY, XF_CB_EYHX, XF_CB_HXEY, XF_CB_HYEX}};
  for (int a = 0; a < 3; ++a) {
    be[a].assign(N[a] + 1, 0.0); ae[a].assign(N[a] + 1, 0.0);
    bm[a].assign(N[a] + 1, 0.0); am[a].assign(N[a] + 1, 0.0);
  }
  std::vector<char> touched_e[3], touched_h[3];
  for (int a = 0; a < 3; ++a) { touched_e[a].assign(N[a] + 1, 0); touched_h[a].assign(N[a] + 1, 0); }
  for (const auto &b : _boundaries) {
    auto *pml = dynamic_cast<PML *>(b.get());
    if (!pml) throw std::runtime_error("only PML boundaries are supported");
    const int ori = static_cast<int>(pml->getOrientation());
    if (seen[ori]) throw std::runtime_error("two boundaries on the same face");
    seen[ori] = true;
    const int a = ori / 2;
    const bool pos = pml->isOrientatingPositive();
    const int d = (int)pml->getSize();
    if (d < 1 || d >= N[a]) throw std::runtime_error("PML thickness does not fit the grid");
    const PML::Profile pr = pml->computeProfile(_cell_size, dt);
    const int start = pos ? N[a] - d : 0;
    if (pos) dp[a] = d; else dn[a] = d;
    for (int l = 0; l < d; ++l) {
      const int ne = start + l + (pos ? 0 : 1);
      const int nh = start + l;
      if (touched_e[a][ne] || touched_h[a][nh]) throw std::runtime_error("CPML faces overlap");
      touched_e[a][ne] = touched_h[a][nh] = 1;
      be[a][ne] = pr.b_e[l]; ae[a][ne] = pr.a_e[l];
      bm[a][nh] = pr.b_m[l]; am[a][nh] = pr.a_m[l];
      for (int m = 0; m < nmat; ++m) {
        for (int q = 0; q < 2; ++q) {
          double &c = cb[axis_tables[a][q]][(size_t)m * (N[a] + 1) + ne];
          c = c / pr.kappa_e[l];
        }
        for (int q = 2; q < 4; ++q) {
          double &c = cb[axis_tables[a][q]][(size_t)m * (N[a] + 1) + nh];
          c = c / pr.kappa_m[l];
        }
      }
    }
  }
  // The engine applies the lumped and TFSF corrections before CPML, the
  // reference after; that only matters for cells in both, which are refused.
  auto in_slab_e = [&](int a, int n) { return (n >= 1 && n <= dn[a]) || (dp[a] > 0 && n >= N[a] - dp[a] && n <= N[a] - 1); };
  auto in_slab_h = [&](int a, int n) { return (n < dn[a]) || (dp[a] > 0 && n >= N[a] - dp[a] && n < N[a]); };
  for (const auto &s : srcs) {
    const int lo[3] = {s.is, s.js, s.ks}, hi[3] = {s.ie, s.je, s.ke};
    for (int a = 0; a < 3; ++a)
      for (int n = lo[a]; n < hi[a]; ++n)
        if (in_slab_e(a, n))
          throw std::runtime_error(s.kind == 1 ? "inductor inside a CPML slab is not supported"
                                               : "current source inside a CPML slab is not supported");
  }

  // ---- design box coefficients: fdtd_basic_coff.cpp:77-91 for each cell's
  // own permittivity (eps_r rounded to float as a Material's), lossless
  if (_has_design) {
    PlanData::Design &D = _plan->design;
    for (int a = 0; a < 3; ++a) {
      D.box[2 * a] = (int64_t)_design_box[2 * a];
      D.box[2 * a + 1] = (int64_t)_design_box[2 * a + 1];
      for (size_t n = _design_box[2 * a]; n < _design_box[2 * a + 1]; ++n)
        if (in_slab_e(a, (int)n) || n == 0 || n >= (size_t)N[a])
          throw std::runtime_error("design box must lie inside the grid and outside the CPML slabs");
    }
    designCoefficients(dt, D.coef, D.eps_r);
    D.present = true;
  }
  // ---- dispersive coefficients, XF_ADE_STRIDE per slot: DrudeMedium::init
  // (dispersive_material.cpp:127-144), DebyMedium::init (205-222),
  // LorentzMedium::init (17-36), with dl the grid's base size
  // (dispersive_material.h:47-48)
  if (any_drude && dbox[1] > dbox[0]) {
    PlanData::Drude &D = _plan->drude;
    for (int a = 0; a < 3; ++a)
      for (int64_t n = dbox[2 * a]; n < dbox[2 * a + 1]; ++n)
        if (in_slab_e(a, (int)n) || in_slab_h(a, (int)n))
          throw std::runtime_error("Drude cells inside a CPML slab are not supported");
    for (int q = 0; q < 6; ++q) D.box[q] = dbox[q];
    D.coef.assign(XF_ADE_STRIDE * (size_t)nmat, 0.0);
    const double dl = _cell_size, eps0 = constant::EPSILON_0;
    for (int slot = 0; slot < nbase; ++slot) {
      if (!drude_of[slot]) continue;
      const Material &m = _objects[drude_of[slot] - 1]->getMaterial();
      const double nv = m.drudeNv(), wp = m.drudeOmegaP(), einf = m.drudeEpsInf();
      const double sigma_e = m.getElectricalConductivity();
      double *c = &D.coef[XF_ADE_STRIDE * (size_t)slot];
      const int kind = m.dispersionKind();
      c[XF_ADE_KIND] = kind;
      c[5] = 0.5 * dl;
      if (kind == 2) {
        const double tau = m.dispersionTau(), de = m.dispersionEpsS() - einf;
        const double kk = (2 * tau - dt) / (2 * tau + dt);
        const double beta = (2 * de * eps0 * dt) / (2 * tau + dt);
        const double temp = 2 * eps0 * einf + beta + sigma_e * dt;
        c[0] = (2 * eps0 * einf + beta - sigma_e * dt) / temp;
        c[1] = (2 * dt) / (temp * dl);
        c[2] = 1 + kk;
        c[3] = kk;
        c[4] = beta;
        c[9] = dt;
      } else if (kind == 3) {
        const double de = m.dispersionEpsS() - einf;
        const double t0 = nv * dt + 1;
        const double alpha = (2 - wp * wp * dt * dt) / t0;
        const double xi = (nv * dt - 1) / t0;
        const double gamma = (eps0 * de * wp * wp * dt * dt) / t0;
        const double t1 = 4 * eps0 * einf + gamma + 2 * sigma_e * dt;
        c[0] = (4 * eps0 * einf - 2 * sigma_e * dt) / t1;
        c[1] = gamma / t1;
        c[2] = 1 + alpha;
        c[3] = alpha;
        c[4] = gamma;
        c[7] = xi;
        c[8] = (4 * dt) / (t1 * dl);
        c[10] = 2 * dt;
      } else {
        const double kk = (2 - nv * dt) / (2 + nv * dt);
        const double beta = (wp * wp * eps0 * dt) / (2 + nv * dt);
        const double temp = 2 * eps0 * einf + beta * dt + sigma_e * dt;
        c[0] = (2 * eps0 * einf - beta * dt - sigma_e * dt) / temp;
        c[1] = 2 * dt / (temp * dl);
        c[2] = 1 + kk;
        c[3] = kk;
        c[4] = beta;
      }
    }
    D.present = true;
  }
  // ---- DFT plane: weights dt exp(-i 2 pi f t dt) (nffft_fd.cpp:102-103)
  if (_has_dft) {
    PlanData::Dft &F = _plan->dft;
    F.comp = _dft_comp; F.axis = _dft_axis;
    F.index = (int64_t)_dft_plane[0]; F.lo1 = (int64_t)_dft_plane[1]; F.hi1 = (int64_t)_dft_plane[2];
    F.lo2 = (int64_t)_dft_plane[3]; F.hi2 = (int64_t)_dft_plane[4];
    F.w.resize(total_time_steps);
    for (size_t t = 0; t < total_time_steps; ++t) {
      const double ph = -2 * constant::PI * _dft_freq * ((double)t * dt);
      F.w[t] = std::complex<double>(dt * std::cos(ph), dt * std::sin(ph));
    }
    F.present = true;
  }

  // ---- the plan: every table the engine needs
  PlanData &P = *_plan;
  P.nmat = nmat;
  P.midx = std::move(midx);
  P.ca = std::move(ca);
  for (int t = 0; t < 12; ++t) { P.cb[t] = std::move(cb[t]); P.cpsi[t] = std::move(cpsi[t]); }
  for (int a = 0; a < 3; ++a) {
    P.dn[a] = dn[a]; P.dp[a] = dp[a];
    P.be[a] = std::move(be[a]); P.ae[a] = std::move(ae[a]);
    P.bm[a] = std::move(bm[a]); P.am[a] = std::move(am[a]);
  }
  for (auto &s_ : srcs) {
    PlanData::Src ps;
    ps.kind = s_.kind;
    ps.comp = s_.comp; ps.is = s_.is; ps.ie = s_.ie; ps.js = s_.js; ps.je = s_.je; ps.ks = s_.ks; ps.ke = s_.ke;
    ps.coff = std::move(s_.coff);
    ps.wave = std::move(s_.wave);
    ps.c2 = std::move(s_.c2);
    P.srcs.push_back(std::move(ps));
  }

  // ---- TFSF (tfsf/tfsf.cpp:35-53, tfsf_3d.cpp:33-151, tfsf_2d.cpp:26-97)
  if (_tfsf) {
    TFSF *t3 = _tfsf.get();
    auto *t2d = dynamic_cast<TFSF2D *>(t3);
    auto *t3d = dynamic_cast<TFSF3D *>(t3);
    if (!t2d && !t3d) throw std::runtime_error("only TFSF3D and TFSF2D sources are supported");
    const auto dist = t3->getDistance();
    for (int a = 0; a < (t2d ? 2 : 3); ++a)
      if (dist[a] < 1 || 2 * dist[a] >= (size_t)N[a]) throw std::runtime_error("TFSF box does not fit the grid");
    const int l[3] = {(int)dist[0], (int)dist[1], (int)dist[2]};
    const int r[3] = {nx - l[0], ny - l[1], nz - l[2]};
    for (int a = 0; a < (t2d ? 2 : 3); ++a) {
      // corrected E cells sit at l..r, H cells at l-1 and r
      if (in_slab_e(a, l[a]) || in_slab_e(a, r[a]) || in_slab_h(a, l[a] - 1) || in_slab_h(a, r[a]) ||
          dn[a] >= l[a] || (dp[a] > 0 && r[a] >= N[a] - dp[a]))
        throw std::runtime_error("TFSF box must lie outside the CPML slabs");
    }
    t3->_dt = dt;
    std::vector<double> time_arr(total_time_steps);
    for (size_t n = 0; n < total_time_steps; ++n) time_arr[n] = static_cast<double>(n) * dt;
    t3->_waveform->init(time_arr);
    PlanData::Tfsf &tp = P.tfsf;
    tp.source.resize(total_time_steps);
    for (size_t n = 0; n < total_time_steps; ++n) tp.source[n] = t3->_e_0 * (*t3->_waveform)(n);
    const double th = t3->_theta_inc, ph = t3->_phi_inc;
    const double ratio = 1 / std::sqrt(std::pow(std::sin(th), 4) * (std::pow(std::cos(ph), 4) + std::pow(std::sin(ph), 4)) +
                                       std::pow(std::cos(th), 4));
    const PointVector k = t3->_k;
    const size_t bn[3] = {(size_t)(r[0] - l[0]), (size_t)(r[1] - l[1]), (size_t)(r[2] - l[2])};
    // tfsf_2d.cpp:40-43 takes sqrt(nz^2) where tfsf_3d.cpp has nz^2
    const double diag =
        (t2d ? std::sqrt(std::pow(bn[0], 2) + std::pow(bn[1], 2) + std::sqrt(std::pow(bn[2], 2)))
             : std::sqrt(std::pow(bn[0], 2) + std::pow(bn[1], 2) + std::pow(bn[2], 2))) *
        ratio;
    const size_t L = static_cast<size_t>(std::ceil(diag)) + 4 + 1;
    double org[3];
    for (int a = 0; a < 3; ++a) org[a] = (k[a] >= 0) ? (double)l[a] : (double)r[a];
    if (t2d) org[2] = 0;  // tfsf_2d.cpp:60-77
    for (int a = 0; a < 3; ++a) org[a] -= 2 * k[a] / ratio;
    std::vector<double> *full = tp.full, *half = tp.half;
    for (int a = 0; a < 3; ++a) {
      full[a].resize(bn[a] + 1);
      half[a].resize(bn[a] + 2);
      for (size_t i = 0; i < bn[a] + 1; ++i) full[a][i] = ((double)(i + (size_t)l[a]) - org[a]) * k[a] * ratio;
      for (size_t i = 0; i < bn[a] + 2; ++i) half[a][i] = ((double)(i + (size_t)l[a]) - org[a] - 0.5) * k[a] * ratio;
    }
    // every interpolation index (tfsf_3d.cpp:418-422) must fall inside the
    // 1-D grid: bound the sum of the three projections from both sides
    double psum = 0, pmin = -0.5;
    for (int a = 0; a < (t2d ? 2 : 3); ++a) {
      double mx = full[a][0], mn = full[a][0];
      for (double v : half[a]) { mx = std::max(mx, v); mn = std::min(mn, v); }
      for (double v : full[a]) { mx = std::max(mx, v); mn = std::min(mn, v); }
      psum += mx;
      pmin += mn;
    }
    if (psum + 1 >= (double)L - 1 || pmin < 0)
      throw std::runtime_error("TFSF projection leaves the 1-D grid");
    const double scaled_dl = dx / ratio;
    xf_tfsf_desc &td = tp.desc;
    td.li = l[0]; td.ri = r[0]; td.lj = l[1]; td.rj = r[1]; td.lk = l[2]; td.rk = r[2];
    td.aux_size = (int64_t)L;
    for (int a = 0; a < 3; ++a) {
      td.proj_full[a] = full[a].data();
      td.proj_half[a] = half[a].data();
      td.transform_e[a] = t2d ? t2d->_transform_e[a] : t3d->_transform_e[a];
      td.transform_h[a] = t2d ? t2d->_transform_h[a] : t3d->_transform_h[a];
    }
    td.two_d = t2d ? 1 : 0;
    td.ceie = 1.0;
    td.chih = 1.0;
    td.ceihi = -(dt / (constant::EPSILON_0 * scaled_dl));
    td.chiei = -(dt / (constant::MU_0 * scaled_dl));
    td.abc_c1 = (constant::C_0 * dt - scaled_dl) / (constant::C_0 * dt + scaled_dl);
    td.abc_c2 = 2 * scaled_dl / (constant::C_0 * dt + scaled_dl);
    const double d3[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
      td.ca[a] = dt / (constant::EPSILON_0 * d3[a]);
      td.cb[a] = dt / (constant::MU_0 * d3[a]);
    }
    td.total_steps = (int64_t)total_time_steps;
    tp.present = true;
  }

  // ---- broadband NF-FF (nffft/nffft.cpp:22-34, nffft_broadband.cpp:30-91)
  if (auto *bb = dynamic_cast<NffftBroadBand *>(_nffft.get())) {
    const auto dist = bb->getDistance();
    for (int a = 0; a < 3; ++a)
      if (_two_d && a == 2 ? dist[a] != 0 : (dist[a] < 1 || 2 * dist[a] >= (size_t)N[a]))
        throw std::runtime_error("NF-FF box does not fit the grid");
    P.bb.present = true;
    const int64_t ext[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
      P.bb.box[2 * a] = (int64_t)dist[a];
      P.bb.box[2 * a + 1] = ext[a] - (int64_t)dist[a];
    }
    for (int q = 0; q < 6; ++q) _nffft_box[q] = (size_t)P.bb.box[q];
  }
  // ---- NF-FF (nffft/nffft.cpp:22-34, nffft_fd.cpp:17-106)
  if (_nffft && !P.bb.present) {
    auto *nf = dynamic_cast<NffftFd *>(_nffft.get());
    if (!nf) throw std::runtime_error("unsupported NFFFT type");
    const auto dist = nf->getDistance();
    for (int a = 0; a < 3; ++a)
      if (dist[a] < 1 || 2 * dist[a] >= (size_t)N[a]) throw std::runtime_error("NF-FF box does not fit the grid");
    PlanData::Nf &np_ = P.nf;
    xf_nffft_desc &nd = np_.desc;
    nd.li = (int64_t)dist[0]; nd.ri = nx - (int64_t)dist[0];
    nd.lj = (int64_t)dist[1]; nd.rj = ny - (int64_t)dist[1];
    nd.lk = (int64_t)dist[2]; nd.rk = nz - (int64_t)dist[2];
    for (int q = 0; q < 6; ++q) _nffft_box[q] = (size_t)(&nd.li)[q];
    nd.nf = (int)nf->frequencies().size();
    nd.total_steps = (int64_t)total_time_steps;
    const size_t T = total_time_steps;
    std::vector<std::complex<double>> &tj = np_.tj, &tm = np_.tm;
    tj.resize((size_t)nd.nf * T);
    tm.resize((size_t)nd.nf * T);
    using namespace std::complex_literals;
    for (int n = 0; n < nd.nf; ++n)
      for (size_t t = 0; t < T; ++t) {
        const double f = nf->frequencies()[n];
        tj[n * T + t] = dt * std::exp(-1i * 2.0 * constant::PI * (f * (t + 0.5) * dt));
        tm[n * T + t] = dt * std::exp(-1i * 2.0 * constant::PI * (f * t * dt));
      }
    np_.present = true;
  }

}

void Simulation::upload() {
  PlanData &P = *_plan;
  xf_grid_desc gd;
  gd.nx = (int64_t)_n[0]; gd.ny = (int64_t)_n[1]; gd.nz = (int64_t)_n[2];
  gd.precision = _precision;
  gd.n_materials = P.nmat;
  check(xf_engine_create(&gd, &_engine));
  check(xf_engine_set_horizon(_engine, (int64_t)_total_time_steps));
  const double *cbp[12], *cpp[12];
  for (int t = 0; t < 12; ++t) { cbp[t] = P.cb[t].data(); cpp[t] = P.cpsi[t].data(); }
  for (int a = 0; a < 3; ++a)
    if (P.dn[a] + P.dp[a] > 0)
      check(xf_engine_set_cpml(_engine, a, P.dn[a], P.dp[a], P.be[a].data(), P.ae[a].data(), P.bm[a].data(),
                               P.am[a].data()));
  check(xf_engine_set_materials(_engine, P.midx.data(), P.ca.data(), cbp, cpp));
  for (const auto &s : P.srcs) {
    if (s.kind == 1)
      check(xf_engine_add_inductor(_engine, s.comp, s.is, s.ie, s.js, s.je, s.ks, s.ke, s.coff.data(),
                                   s.c2.data()));
    else
      check(xf_engine_add_current_source(_engine, s.comp, s.is, s.ie, s.js, s.je, s.ks, s.ke, s.coff.data(),
                                         s.wave.data(), (int64_t)_total_time_steps));
  }
  for (const LineMonitorPlan &lp : _line_plans) {
    const size_t n = lp.terms.size();
    std::vector<int> comp(n), group(n);
    std::vector<int64_t> a(3 * n), b(3 * n);
    std::vector<double> w(n);
    for (size_t q = 0; q < n; ++q) {
      comp[q] = lp.terms[q].comp;
      group[q] = lp.terms[q].group;
      w[q] = lp.terms[q].w;
      for (int u = 0; u < 3; ++u) {
        a[3 * q + u] = lp.terms[q].a[u];
        b[3 * q + u] = lp.terms[q].b[u];
      }
    }
    int index = 0;
    check(xf_engine_add_line_monitor(_engine, (int)n, comp.data(), a.data(), b.data(), w.data(), group.data(),
                                     lp.sign, (int64_t)_total_time_steps, &index));
  }
  if (P.tfsf.present) {
    xf_tfsf_desc td = P.tfsf.desc;
    for (int a = 0; a < 3; ++a) {
      td.proj_full[a] = P.tfsf.full[a].data();
      td.proj_half[a] = P.tfsf.half[a].data();
    }
    td.source = P.tfsf.source.data();
    check(xf_engine_set_tfsf(_engine, &td));
  }
  if (P.bb.present) check(xf_engine_set_nffft_broadband(_engine, P.bb.box, (int64_t)_total_time_steps));
  if (P.nf.present) {
    xf_nffft_desc nd = P.nf.desc;
    nd.tj = reinterpret_cast<const double *>(P.nf.tj.data());
    nd.tm = reinterpret_cast<const double *>(P.nf.tm.data());
    check(xf_engine_set_nffft(_engine, &nd));
  }
  for (const Probe &p : _probes)
    check(xf_engine_add_probe(_engine, compIndex(p.component), (int64_t)p.i, (int64_t)p.j, (int64_t)p.k));
  if (P.design.present)
    check(xf_engine_set_design(_engine, P.design.box, P.design.coef.data(), P.design.eps_r.data()));
  if (P.drude.present) check(xf_engine_set_dispersive(_engine, P.drude.box, P.drude.coef.data()));
  if (P.dft.present)
    check(xf_engine_set_dft_plane(_engine, P.dft.comp, P.dft.axis, P.dft.index, P.dft.lo1, P.dft.hi1, P.dft.lo2,
                                  P.dft.hi2, reinterpret_cast<const double *>(P.dft.w.data()),
                                  (int64_t)P.dft.w.size()));
}

// fdtd_basic_coff.cpp:77-91 for each design cell's own permittivity (eps_r
// rounded to float as a Material's), lossless: [cell][9] E coefficients in
// the order of xf_engine_set_design, and the cells' eps_r
void Simulation::designCoefficients(double dt, std::vector<double> &coef, std::vector<double> &eps) const {
  const size_t nd = (_design_box[1] - _design_box[0]) * (_design_box[3] - _design_box[2]) *
                    (_design_box[5] - _design_box[4]);
  if (_design_eps.size() != nd) throw std::invalid_argument("design eps_r has the wrong size");
  coef.resize(9 * nd);
  eps.resize(nd);
  size_t q = 0;
  for (size_t i = _design_box[0]; i < _design_box[1]; ++i)
    for (size_t j = _design_box[2]; j < _design_box[3]; ++j)
      for (size_t k = _design_box[4]; k < _design_box[5]; ++k, ++q) {
        const Mat m{static_cast<double>(_design_eps[q]) * constant::EPSILON_0, constant::MU_0, 0.0, 0.0};
        double *c = &coef[9 * q];
        c[0] = coffSelfE(m, dt);
        c[1] = coffSelfE(m, dt);
        c[2] = coffSelfE(m, dt);
        c[3] = coffEa(m, dt, _hs[2][k]);  // cexhy
        c[4] = coffEb(m, dt, _hs[2][k]);  // ceyhx
        c[5] = coffEb(m, dt, _hs[1][j]);  // cexhz
        c[6] = coffEa(m, dt, _hs[1][j]);  // cezhx
        c[7] = coffEa(m, dt, _hs[0][i]);  // ceyhz
        c[8] = coffEb(m, dt, _hs[0][i]);  // cezhy
        eps[q] = static_cast<double>(_design_eps[q]);
      }
}

void Simulation::updateDesign(std::vector<float> eps_r) {
  if (!_engine || !_has_design) throw std::runtime_error("no initialised simulation with a design box");
  const size_t nd = (_design_box[1] - _design_box[0]) * (_design_box[3] - _design_box[2]) *
                    (_design_box[5] - _design_box[4]);
  if (eps_r.size() != nd) throw std::invalid_argument("design eps_r has the wrong size");
  _design_eps = std::move(eps_r);
  std::vector<double> coef, eps;
  designCoefficients(getDt(), coef, eps);
  check(xf_engine_update_design(_engine, coef.data(), eps.data()));
}

void Simulation::setDesign(const std::array<size_t, 6> &box, std::vector<float> eps_r) {
  if (_plan) throw std::runtime_error("Simulation already planned");
  _design_box = box;
  _design_eps = std::move(eps_r);
  _has_design = true;
}

void Simulation::setDftPlane(EMComponent component, int axis, size_t index, size_t lo1, size_t hi1, size_t lo2,
                             size_t hi2, double frequency) {
  if (_plan) throw std::runtime_error("Simulation already planned");
  if (axis < 0 || axis > 2) throw std::invalid_argument("axis 0..2");
  _dft_comp = compIndex(component);
  _dft_axis = axis;
  _dft_plane[0] = index; _dft_plane[1] = lo1; _dft_plane[2] = hi1; _dft_plane[3] = lo2; _dft_plane[4] = hi2;
  _dft_freq = frequency;
  _has_dft = true;
}

std::vector<std::complex<double>> Simulation::getDft() const {
  if (!_engine || !_has_dft) throw std::runtime_error("no DFT plane");
  std::vector<std::complex<double>> out((_dft_plane[2] - _dft_plane[1]) * (_dft_plane[4] - _dft_plane[3]));
  check(xf_engine_get_dft(_engine, reinterpret_cast<double *>(out.data())));
  return out;
}

double Simulation::gradEps(std::vector<double> &grad, int checkpoint_every) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  if (!_has_design) throw std::runtime_error("no design box");
  grad.assign(_design_eps.size(), 0.0);
  double J = 0;
  check(xf_engine_grad_eps(_engine, checkpoint_every, grad.data(), &J));
  return J;
}

void Simulation::adjointRun(size_t n, std::vector<double> &fields) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  size_t total = 0;
  for (int c = 0; c < 6; ++c) {
    auto sh = fieldShape(static_cast<EMComponent>(c));
    total += sh[0] * sh[1] * sh[2];
  }
  // with Drude cells the three J adjoints follow, in the E shapes
  if (_plan && _plan->drude.present)
    for (int c = 0; c < 3; ++c) {
      auto sh = fieldShape(static_cast<EMComponent>(c));
      total += sh[0] * sh[1] * sh[2];
    }
  if (fields.size() != total) throw std::invalid_argument("adjoint fields have the wrong size");
  check(xf_engine_adjoint_run(_engine, (int64_t)n, fields.data()));
}

void Simulation::checkpoint(int slot) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  check(xf_engine_checkpoint(_engine, slot));
}

void Simulation::restore(int slot) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  check(xf_engine_restore(_engine, slot));
}

void Simulation::init(size_t total_time_steps) {
  if (_engine) throw std::runtime_error("Simulation already initialised");
  if (!_plan) plan(total_time_steps);
  else if (_total_time_steps != total_time_steps) throw std::runtime_error("plan made for another step count");
  upload();
  for (const auto &mp : _monitors)
    if (auto *hm = dynamic_cast<HostMonitor *>(mp.get())) hm->init(*this);
}

// per-cell value of coefficient array `which` (cexe, cexhy, ..., chzey order of
// fdtd_basic_coff.h) implied by the plan, in the reference's array shape
std::vector<double> Simulation::planCoefficient(int which) const {
  if (!_plan) throw std::runtime_error("no plan");
  const PlanData &P = *_plan;
  static const int table_of[18] = {-1, XF_CB_EXHY, XF_CB_EXHZ, -1, XF_CB_EYHZ, XF_CB_EYHX,
                                   -1, XF_CB_EZHX, XF_CB_EZHY, -1, XF_CB_HXEY, XF_CB_HXEZ,
                                   -1, XF_CB_HYEZ, XF_CB_HYEX, -1, XF_CB_HZEX, XF_CB_HZEY};
  static const int axis_of_table[12] = {2, 1, 0, 2, 1, 0, 2, 1, 0, 2, 1, 0};
  if (which < 0 || which >= 18) throw std::invalid_argument("coefficient index 0..17");
  const int comp = which / 3;
  const auto sh = fieldShape(static_cast<EMComponent>(comp));
  const size_t nz1 = _n[2] + 1, ny1 = _n[1] + 1;
  std::vector<double> out(sh[0] * sh[1] * sh[2]);
  const int t = table_of[which];
  for (size_t i = 0; i < sh[0]; ++i)
    for (size_t j = 0; j < sh[1]; ++j)
      for (size_t k = 0; k < sh[2]; ++k) {
        const int m = P.midx[(i * ny1 + j) * nz1 + k];
        double v;
        if (t < 0) {
          v = P.ca[comp * P.nmat + m];
        } else {
          const int a = axis_of_table[t];
          const size_t n = a == 0 ? i : (a == 1 ? j : k);
          v = P.cb[t][(size_t)m * (_n[a] + 1) + n];
        }
        out[(i * sh[1] + j) * sh[2] + k] = v;
      }
  // design cells: their own E coefficients
  if (P.design.present && comp < 3) {
    // position of `which` in the per-cell record (cexe, ceye, ceze, cexhy,
    // ceyhx, cexhz, cezhx, ceyhz, cezhy)
    static const int slot_of[9] = {0, 3, 5, 1, 7, 4, 2, 6, 8};
    const int64_t *b = P.design.box;
    size_t q = 0;
    for (int64_t i = b[0]; i < b[1]; ++i)
      for (int64_t j = b[2]; j < b[3]; ++j)
        for (int64_t k = b[4]; k < b[5]; ++k, ++q)
          out[((size_t)i * sh[1] + (size_t)j) * sh[2] + (size_t)k] = P.design.coef[9 * q + slot_of[which]];
  }
  return out;
}

// Steps in chunks that end at the movie monitors' sample steps; after step t
// (monitor update slot, simulation_update.cpp:35) each due monitor writes
// its frame (movie_monitor.cpp:46-59).
void Simulation::step(size_t n) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  size_t cur = currentTimeStep();
  const size_t end = cur + n;
  std::vector<HostMonitor *> host;
  for (const auto &mp : _monitors)
    if (auto *hm = dynamic_cast<HostMonitor *>(mp.get())) host.push_back(hm);
  while (cur < end) {
    size_t next = host.empty() ? end : cur + 1;  // one past the next sample step
    for (const MonitorPlan &m : _monitor_plans) {
      if (m.plot_step == 0) continue;
      size_t t = ((cur + m.plot_step - 1) / m.plot_step) * m.plot_step;  // first multiple >= cur
      if (m.total > 0 && m.total - 1 >= cur && m.total - 1 < t) t = m.total - 1;
      if (t < end) next = std::min(next, t + 1);
    }
    check(xf_engine_step(_engine, (int64_t)(next - cur)));
    cur = next;
    const size_t t = cur - 1;
    for (const MonitorPlan &m : _monitor_plans) {
      if (m.plot_step == 0 || !(t % m.plot_step == 0 || t == m.total - 1)) continue;
      char name[32];
      std::snprintf(name, sizeof name, "%04zu.npy", t);
      writeMonitor(m, name);
    }
    if (!host.empty()) {
      sync();
      for (HostMonitor *hm : host) hm->update(*this, t);
    }
  }
}

void Simulation::sync() {
  if (_engine) check(xf_engine_sync(_engine));
}

size_t Simulation::currentTimeStep() const {
  return _engine ? (size_t)xf_engine_current_step(_engine) : 0;
}

void Simulation::checkRun(size_t total_time_steps) {
  std::cout << "Simulation Check:" << '\n';
  if (!_plan) plan(total_time_steps);
  std::ofstream ofs{"simulation_check", std::ios::out};
  ofs << "nx:" << _n[0] << " ny:" << _n[1] << " nz:" << _n[2] << '\n';
  ofs << "Total size:" << _n[0] * _n[1] * _n[2] << '\n';
  std::ofstream object_ofs{"simulation_object_check", std::ios::out};
  object_ofs << "Object:" << '\n';
  for (const auto &o : _objects) object_ofs << static_cast<std::string>(*o) << '\n';
}

void Simulation::run(size_t total_time_steps) {
  init(total_time_steps);
  step(total_time_steps);
  sync();
  outputData();
}

std::array<size_t, 3> Simulation::fieldShape(EMComponent c) const {
  const size_t nx = _n[0], ny = _n[1], nz = _n[2];
  switch (c) {
    case EMComponent::EX: return {nx, ny + 1, nz + 1};
    case EMComponent::EY: return {nx + 1, ny, nz + 1};
    case EMComponent::EZ: return {nx + 1, ny + 1, nz};
    case EMComponent::HX: return {nx + 1, ny, nz};
    case EMComponent::HY: return {nx, ny + 1, nz};
    default: return {nx, ny, nz + 1};
  }
}

std::vector<double> Simulation::getField(EMComponent c) const {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  const auto sh = fieldShape(c);
  std::vector<double> out(sh[0] * sh[1] * sh[2]);
  check(xf_engine_get_field(_engine, compIndex(c), out.data()));
  return out;
}

std::vector<double> Simulation::getFieldBox(EMComponent c, const std::array<size_t, 6> &box) const {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  const auto sh = fieldShape(c);
  for (int a = 0; a < 3; ++a)
    if (box[2 * a] > box[2 * a + 1] || box[2 * a + 1] > sh[a]) throw std::out_of_range("field box outside the component");
  const int64_t b[6] = {(int64_t)box[0], (int64_t)(box[1] - box[0]), (int64_t)box[2], (int64_t)(box[3] - box[2]),
                        (int64_t)box[4], (int64_t)(box[5] - box[4])};
  std::vector<double> out((size_t)b[1] * b[3] * b[5]);
  check(xf_engine_get_box(_engine, compIndex(c), b, out.data()));
  return out;
}

void Simulation::setField(EMComponent c, const std::vector<double> &v) {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  const auto sh = fieldShape(c);
  if (v.size() != sh[0] * sh[1] * sh[2]) throw std::invalid_argument("field size mismatch");
  check(xf_engine_set_field(_engine, compIndex(c), v.data()));
}

std::vector<std::complex<double>> Simulation::getNffftAccumulators() const {
  if (!_engine || !_nffft) throw std::runtime_error("no NF-FF surface");
  std::vector<std::complex<double>> out((size_t)xf_engine_nffft_size(_engine));
  check(xf_engine_get_nffft(_engine, reinterpret_cast<double *>(out.data())));
  return out;
}

// nffft_fd.cpp:379-421: the per-face sums come from the engine; the area
// factors, angular projections and far-zone fields follow the reference.
FarField Simulation::calculateFarField() const {
  auto *nf = dynamic_cast<NffftFd *>(_nffft.get());
  if (!_engine || !nf) throw std::runtime_error("no NF-FF surface");
  using C = std::complex<double>;
  using namespace std::complex_literals;
  const size_t NF = nf->frequencies().size(), NT = nf->theta().size(), NP = nf->phi().size();
  const size_t M = NF * NT * NP;
  std::vector<C> sums(M * 12);
  check(xf_engine_farfield_sums(_engine, _cell_size, nf->waveNumber().data(), nf->theta().data(), (int)NT,
                                nf->phi().data(), (int)NP, reinterpret_cast<double *>(sums.data())));
  FarField r;
  for (auto *v : {&r.n_theta, &r.n_phi, &r.l_theta, &r.l_phi, &r.e_theta, &r.e_phi, &r.h_theta, &r.h_phi})
    v->resize(M);
  r.power_theta.resize(M);
  r.power_phi.resize(M);
  const double dl = _cell_size;
  const double area = dl * dl;
  const double eta = constant::ETA_0;
  for (size_t n = 0; n < NF; ++n)
    for (size_t t = 0; t < NT; ++t)
      for (size_t p = 0; p < NP; ++p) {
        const size_t m = (n * NT + t) * NP + p;
        const C *s = &sums[m * 12];
        C a_x{0, 0}, a_y{0, 0}, a_z{0, 0}, f_x{0, 0}, f_y{0, 0}, f_z{0, 0};
        // calculateFarfieldX: (sum_jy, sum_jz, sum_my, sum_mz)
        a_y += s[0] * area; a_z += s[1] * area; f_y += s[2] * area; f_z += s[3] * area;
        // calculateFarfieldY: (sum_jz, sum_jx, sum_mz, sum_mx)
        a_z += s[4] * area; a_x += s[5] * area; f_z += s[6] * area; f_x += s[7] * area;
        // calculateFarfieldZ: (sum_jx, sum_jy, sum_mx, sum_my)
        a_x += s[8] * area; a_y += s[9] * area; f_x += s[10] * area; f_y += s[11] * area;
        const double cos_t = std::cos(nf->theta()[t]), sin_t = std::sin(nf->theta()[t]);
        const double cos_p = std::cos(nf->phi()[p]), sin_p = std::sin(nf->phi()[p]);
        const double cos_t_cos_p = cos_t * cos_p, cos_t_sin_p = cos_t * sin_p;
        const C a_theta = cos_t_cos_p * a_x + cos_t_sin_p * a_y - sin_t * a_z;
        const C a_phi = -sin_p * a_x + cos_p * a_y;
        const C f_theta = cos_t_cos_p * f_x + cos_t_sin_p * f_y - sin_t * f_z;
        const C f_phi = -sin_p * f_x + cos_p * f_y;
        const double k = nf->waveNumber()[n];
        const double coff = std::pow(k, 2) / (32 * constant::PI * constant::PI * eta);
        r.n_theta[m] = a_theta;
        r.n_phi[m] = a_phi;
        r.l_theta[m] = f_theta;
        r.l_phi[m] = f_phi;
        r.e_theta[m] = (-1i * k) * (eta * a_theta + f_phi) / (4 * constant::PI);
        r.e_phi[m] = (1i * k) * (-eta * a_phi + f_theta) / (4 * constant::PI);
        r.h_theta[m] = (1i * k) * (a_phi - f_theta / eta) / (4 * constant::PI);
        r.h_phi[m] = (-1i * k) * (a_theta + f_phi / eta) / (4 * constant::PI);
        r.power_theta[m] = coff * std::pow(std::abs(eta * a_theta + f_phi), 2);
        r.power_phi[m] = coff * std::pow(std::abs(-eta * a_phi + f_theta), 2);
      }
  return r;
}

std::vector<double> Simulation::getBroadBandSamples() const {
  if (!_engine || !dynamic_cast<NffftBroadBand *>(_nffft.get())) throw std::runtime_error("no broadband NF-FF");
  std::vector<double> out((size_t)xf_engine_broadband_samples_size(_engine));
  check(xf_engine_get_broadband_samples(_engine, out.data()));
  return out;
}

// nffft_broadband.cpp:316-345: bins f = linspace(-fs/2, fs/2, 2T-1) (xtensor's
// start + q step), k = 2 pi f / c0; the surface integrals come from the
// engine, the projections and far-zone fields follow the reference.
BroadBandFarField Simulation::calculateBroadBandFarField(const std::vector<double> *wave_number) const {
  auto *bb = dynamic_cast<NffftBroadBand *>(_nffft.get());
  if (!_engine || !bb) throw std::runtime_error("no broadband NF-FF surface");
  using C = std::complex<double>;
  using namespace std::complex_literals;
  const size_t T = _total_time_steps, Ns = 2 * T - 1;
  BroadBandFarField r;
  r.frequencies.resize(Ns);
  r.wave_number.resize(Ns);
  const double fs = 1 / _dt;
  const double start = -fs / 2, stop = fs / 2;
  const double step = (stop - start) / (double)(Ns > 1 ? Ns - 1 : 1);
  for (size_t q = 0; q < Ns; ++q) {
    r.frequencies[q] = start + step * (double)q;
    r.wave_number[q] = 2 * constant::PI * r.frequencies[q] / constant::C_0;
  }
  const std::vector<double> &k = wave_number ? *wave_number : r.wave_number;
  if (k.size() != Ns) throw std::invalid_argument("wave_number needs 2T-1 entries");
  std::vector<C> sums(6 * Ns);
  check(xf_engine_broadband_sums(_engine, _cell_size, k.data(), bb->theta(), bb->phi(),
                                 reinterpret_cast<double *>(sums.data())));
  auto part = [&](int u) { return std::vector<C>(sums.begin() + u * Ns, sums.begin() + (u + 1) * Ns); };
  r.n_x = part(0); r.n_y = part(1); r.n_z = part(2);
  r.l_x = part(3); r.l_y = part(4); r.l_z = part(5);
  for (auto *v : {&r.n_theta, &r.n_phi, &r.l_theta, &r.l_phi, &r.e_theta, &r.e_phi, &r.h_theta, &r.h_phi})
    v->resize(Ns);
  r.power_theta.resize(Ns);
  r.power_phi.resize(Ns);
  const double cos_t = std::cos(bb->theta()), sin_t = std::sin(bb->theta());
  const double cos_p = std::cos(bb->phi()), sin_p = std::sin(bb->phi());
  const double eta = constant::ETA_0;
  for (size_t q = 0; q < Ns; ++q) {
    const C n_theta = cos_t * cos_p * r.n_x[q] + cos_t * sin_p * r.n_y[q] - sin_t * r.n_z[q];
    const C n_phi = -sin_p * r.n_x[q] + cos_p * r.n_y[q];
    const C l_theta = cos_t * cos_p * r.l_x[q] + cos_t * sin_p * r.l_y[q] - sin_t * r.l_z[q];
    const C l_phi = -sin_p * r.l_x[q] + cos_p * r.l_y[q];
    const double kq = r.wave_number[q];
    r.n_theta[q] = n_theta;
    r.n_phi[q] = n_phi;
    r.l_theta[q] = l_theta;
    r.l_phi[q] = l_phi;
    r.e_theta[q] = (-1i * kq) * (eta * n_theta + l_phi) / (4 * constant::PI);
    r.e_phi[q] = (1i * kq) * (-eta * n_phi + l_theta) / (4 * constant::PI);
    r.h_theta[q] = (1i * kq) * (n_phi - l_theta / eta) / (4 * constant::PI);
    r.h_phi[q] = (-1i * kq) * (n_theta + l_phi / eta) / (4 * constant::PI);
    const double coff = std::pow(kq, 2) / (32 * constant::PI * constant::PI * eta);
    r.power_theta[q] = coff * std::pow(std::abs(eta * n_theta + l_phi), 2);
    r.power_phi[q] = coff * std::pow(std::abs(-eta * n_phi + l_theta), 2);
  }
  if (_two_d) {
    r.ez.resize(Ns);
    for (size_t q = 0; q < Ns; ++q) r.ez[q] = 0.5 * std::sqrt(1i * r.wave_number[q]) * (r.l_phi[q] + eta * r.n_theta[q]);
  }
  return r;
}

namespace {

// minimal .npy (format 1.0) writer, little-endian C order
void writeNpy(const std::string &path, const void *data, const char *descr, size_t elem,
              const std::vector<size_t> &shape) {
  std::string sh = "(";
  for (size_t q = 0; q < shape.size(); ++q) sh += std::to_string(shape[q]) + (shape.size() == 1 || q + 1 < shape.size() ? "," : "");
  sh += ")";
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr += '\n';
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const char magic[] = "\x93NUMPY";
  f.write(magic, 6);
  const char ver[2] = {1, 0};
  f.write(ver, 2);
  const uint16_t hl = (uint16_t)hdr.size();
  f.write(reinterpret_cast<const char *>(&hl), 2);
  f.write(hdr.data(), (std::streamsize)hdr.size());
  size_t n = elem;
  for (size_t s : shape) n *= s;
  f.write(static_cast<const char *>(data), (std::streamsize)n);
}

void makeDirs(const std::string &path) {
  std::string cur;
  for (size_t q = 0; q < path.size(); ++q) {
    cur += path[q];
    if (path[q] == '/' || q + 1 == path.size()) {
      if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
        throw std::runtime_error("cannot create directory " + cur);
    }
  }
}

}  // namespace

namespace {
std::vector<size_t> npyShape(size_t n, const std::vector<size_t> &shape) {
  if (shape.empty()) return {n};
  size_t m = 1;
  for (size_t s : shape) m *= s;
  if (m != n) throw std::invalid_argument("dumpNpy: shape holds " + std::to_string(m) + " values, data " + std::to_string(n));
  return shape;
}
}  // namespace

void dumpNpy(const std::string &path, const std::vector<double> &v, const std::vector<size_t> &shape) {
  const auto sh = npyShape(v.size(), shape);
  if (path.find('/') != std::string::npos) makeDirs(path.substr(0, path.rfind('/')));
  writeNpy(path, v.data(), "<f8", 8, sh);
}

void dumpNpy(const std::string &path, const std::vector<std::complex<double>> &v, const std::vector<size_t> &shape) {
  const auto sh = npyShape(v.size(), shape);
  if (path.find('/') != std::string::npos) makeDirs(path.substr(0, path.rfind('/')));
  writeNpy(path, v.data(), "<c16", 16, sh);
}

// grid boxes of the monitors (GridSpace::getGridBox, grid_space.cpp:306-329)
// and the plane check of TimeDomainFieldMonitor::init (field_monitor.cpp:30-87)
// VoltageMonitor / CurrentMonitor as term lists for the engine, in the
// reference's loop order (monitor/voltage_monitor.cpp:20-93,
// monitor/current_monitor.cpp:20-122)
void Simulation::planLineMonitors() {
  _line_plans.clear();
  const int64_t N[3] = {(int64_t)_n[0], (int64_t)_n[1], (int64_t)_n[2]};
  for (const auto &mp : _monitors) {
    auto *vm = dynamic_cast<VoltageMonitor *>(mp.get());
    auto *cm = dynamic_cast<CurrentMonitor *>(mp.get());
    if (!vm && !cm) continue;
    const Cube *c = mp->getShape();
    if (!c) throw std::runtime_error("Shape Type Error");
    const double lo[3] = {c->getXmin(), c->getYmin(), c->getZmin()};
    const double hi[3] = {c->getXmax(), c->getYmax(), c->getZmax()};
    int64_t st[3], en[3];
    for (int a = 0; a < 3; ++a) {
      st[a] = (int64_t)handleTransform(a, lo[a]);
      en[a] = (int64_t)handleTransform(a, hi[a]);
      if (en[a] < st[a]) throw std::range_error("end < start");
    }
    const int ori = static_cast<int>(vm ? vm->getOrientation() : cm->getOrientation());
    const int a = ori / 2;
    LineMonitorPlan lp;
    lp.monitor = mp.get();
    auto term = [&](int comp, int group, int64_t i, int64_t j, int64_t k, int64_t bi, int64_t bj, int64_t bk,
                    double w) {
      LineTermPlan t;
      t.comp = comp;
      t.group = group;
      t.a[0] = i; t.a[1] = j; t.a[2] = k;
      t.b[0] = bi; t.b[1] = bj; t.b[2] = bk;
      t.w = w;
      lp.terms.push_back(t);
    };
    const int64_t is = st[0], ie = en[0], js = st[1], je = en[1], ks = st[2], ke = en[2];
    if (vm) {
      // value += coff(c) E_a(i,j,k), coff = -dc / (cells across), negated for N
      const double across = (double)(a == 0 ? (je - js) * (ke - ks) : a == 1 ? (ie - is) * (ke - ks)
                                                                             : (ie - is) * (je - js));
      for (int64_t i = is; i < ie; ++i)
        for (int64_t j = js; j < je; ++j)
          for (int64_t k = ks; k < ke; ++k) {
            const int64_t idx = a == 0 ? i : (a == 1 ? j : k);
            double w = -_hs[a][idx] / across;
            if (ori % 2 == 0) w *= -1;
            term(a, 0, i, j, k, -1, -1, -1, w);
          }
      if (lp.terms.empty()) throw std::runtime_error("voltage monitor box has no cells");
      lp.sign = 1.0;
    } else {
      // the loop integral of H around the box at its start plane
      bool ok = st[a] < N[a];
      for (int b = 0; b < 3; ++b)
        if (b != a && (st[b] < 1 || en[b] + 1 > N[b])) ok = false;
      if (!ok) throw std::out_of_range("current monitor loop outside the grid");
      if (a == 0) {
        for (int64_t j = js; j <= je; ++j) term(4, 0, is, j, ks - 1, is, j, ke, _es[1][j]);
        for (int64_t k = ks; k <= ke; ++k) term(5, 1, is, je, k, is, js - 1, k, _es[2][k]);
      } else if (a == 1) {
        for (int64_t k = ks; k <= ke; ++k) term(5, 0, is - 1, js, k, ie, js, k, _es[2][k]);
        for (int64_t i = is; i <= ie; ++i) term(3, 1, i, js, ke, i, js, ks - 1, _es[0][i]);
      } else {
        for (int64_t i = is; i <= ie; ++i) term(3, 0, i, js - 1, ks, i, je, ks, _es[0][i]);
        for (int64_t j = js; j <= je; ++j) term(4, 1, ie, j, ks, is - 1, j, ks, _es[1][j]);
      }
      lp.sign = ori % 2 == 0 ? -1.0 : 1.0;
    }
    _line_plans.push_back(std::move(lp));
  }
}

void Simulation::planMonitors() {
  _monitor_plans.clear();
  planLineMonitors();
  for (const auto &mp : _monitors) {
    const Monitor *mon = mp.get();
    if (dynamic_cast<const VoltageMonitor *>(mon) || dynamic_cast<const CurrentMonitor *>(mon) ||
        dynamic_cast<const HostMonitor *>(mon))
      continue;
    size_t plot = 0, total = 0;
    if (auto *mv = dynamic_cast<const MovieMonitor *>(mon)) {
      if (mv->getPlotStep() == 0) throw std::invalid_argument("MovieMonitor plot_step must be positive");
      plot = mv->getPlotStep();
      total = mv->getTotalTimeSteps();
      mon = &mv->getMonitor();
    }
    auto *fm = dynamic_cast<const TimeDomainFieldMonitor *>(mon);
    if (!fm)
      throw std::runtime_error(
          "only TimeDomainFieldMonitor, MovieMonitor, VoltageMonitor, CurrentMonitor and HostMonitor are supported");
    const Cube *c = fm->getShape();
    if (!c) throw std::runtime_error("Shape Type Error");
    const double lo[3] = {c->getXmin(), c->getYmin(), c->getZmin()};
    const double hi[3] = {c->getXmax(), c->getYmax(), c->getZmax()};
    MonitorPlan m;
    m.monitor = fm;
    m.comp = compIndex(fm->getComponent());
    m.plot_step = plot;
    m.total = total;
    size_t num[3];
    for (int a = 0; a < 3; ++a) {
      const size_t s = handleTransform(a, lo[a]), e = handleTransform(a, hi[a]);
      if (e < s) throw std::range_error("end < start");
      num[a] = e - s;
      m.box[2 * a] = (int64_t)s;
      m.box[2 * a + 1] = (int64_t)(num[a] == 0 ? 1 : num[a]);
    }
    const PlaneType pt = fm->getPlaneType();
    if ((pt == PlaneType::XY && num[2] > 1) || (pt == PlaneType::YZ && num[0] > 1) ||
        (pt == PlaneType::ZX && num[1] > 1))
      throw std::runtime_error("Error: invalid grid box");
    const auto sh = fieldShape(fm->getComponent());
    for (int a = 0; a < 3; ++a)
      if ((size_t)(m.box[2 * a] + m.box[2 * a + 1]) > sh[a])
        throw std::out_of_range("monitor box outside the component array");
    _monitor_plans.push_back(m);
  }
}

void Simulation::writeMonitor(const MonitorPlan &m, const std::string &file) const {
  const std::string dir = m.monitor->getOutputPath();
  try {
    makeDirs(dir);
  } catch (const std::exception &) {
    std::fprintf(stderr, "Error: cannot create directory %s\n", dir.c_str());
    return;
  }
  std::vector<double> v((size_t)(m.box[1] * m.box[3] * m.box[5]));
  check(xf_engine_get_box(_engine, m.comp, m.box, v.data()));
  writeNpy(dir + "/" + file, v.data(), "<f8", 8, {(size_t)m.box[1], (size_t)m.box[3], (size_t)m.box[5]});
}

// nffft_fd.cpp:579-634, then the field monitors' final slices
// (field_monitor.cpp:91-121)
void Simulation::outputData() const {
  // simulation.cpp:144-156: NF-FF, then the monitors, then the network
  outputNffft();
  if (_engine) {
    for (const auto &mp : _monitors)
      if (auto *hm = dynamic_cast<HostMonitor *>(mp.get())) hm->outputData();
    for (const MonitorPlan &m : _monitor_plans)
      if (m.plot_step == 0) writeMonitor(m, m.monitor->getOutputFileName());
    if (!_line_plans.empty()) {
      getLineMonitorData();
      for (const LineMonitorPlan &lp : _line_plans) writeLineMonitor(lp);
    }
    if (_network) {
      if (_line_plans.empty()) getLineMonitorData();
      outputNetwork();
    }
  }
}

// VoltageMonitor::outputData (voltage_monitor.cpp:95-117): time
// linspace(0, T dt, T); CurrentMonitor::outputData (current_monitor.cpp:
// 124-138): linspace(dt/2, (T + 1/2) dt, T); rows [time, value]
void Simulation::writeLineMonitor(const LineMonitorPlan &lp) const {
  const Monitor *m = lp.monitor;
  const bool voltage = dynamic_cast<const VoltageMonitor *>(m) != nullptr;
  const std::vector<double> &v = voltage ? static_cast<const VoltageMonitor *>(m)->getValueArray()
                                         : static_cast<const CurrentMonitor *>(m)->getValueArray();
  const size_t T = _total_time_steps;
  std::vector<double> out(2 * T);
  const double start = voltage ? 0.0 : 0.5 * _dt;
  const double stop = voltage ? (double)T * _dt : ((double)T + 0.5) * _dt;
  const double step = (stop - start) / std::fmax(1.0, (double)T - 1.0);
  for (size_t n = 0; n < T; ++n) out[n] = (n + 1 == T && T > 1) ? stop : start + step * (double)n;
  for (size_t n = 0; n < T; ++n) out[T + n] = v[n];
  try {
    makeDirs(m->getOutputPath());
  } catch (const std::exception &) {
    std::fprintf(stderr, "Error: cannot create directory %s\n", m->getOutputPath().c_str());
    return;
  }
  writeNpy(m->getOutputPath() + "/" + m->getOutputFileName(), out.data(), "<f8", 8, {2, T});
}

// Network::outputData (network.cpp:27-60): s{10 j + i}.npy = [frequencies; S]
void Simulation::outputNetwork() const {
  auto s = _network->calculateSParameters(_dt);
  const std::string &dir = _network->getOutputPath();
  try {
    makeDirs(dir);
  } catch (const std::exception &) {
    std::fprintf(stderr, "Error: cannot create directory %s\n", dir.c_str());
    return;
  }
  const auto &f = _network->frequencies();
  for (const auto &e : s) {
    std::vector<std::complex<double>> data(2 * f.size());
    for (size_t q = 0; q < f.size(); ++q) {
      data[q] = f[q];
      data[f.size() + q] = e.second[q];
    }
    writeNpy(dir + "/s" + std::to_string(e.first) + ".npy", data.data(), "<c16", 16, {2, f.size()});
  }
}

std::vector<std::vector<double>> Simulation::getLineMonitorData() const {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  const size_t T = _total_time_steps;
  std::vector<double> buf(_line_plans.size() * T);
  if (!_line_plans.empty()) check(xf_engine_get_line_monitors(_engine, buf.data(), (int64_t)T));
  std::vector<std::vector<double>> out(_line_plans.size());
  for (size_t q = 0; q < _line_plans.size(); ++q) {
    out[q].assign(buf.begin() + q * T, buf.begin() + (q + 1) * T);
    Monitor *m = _line_plans[q].monitor;
    if (auto *vm = dynamic_cast<VoltageMonitor *>(m)) vm->_value = out[q];
    if (auto *cm = dynamic_cast<CurrentMonitor *>(m)) cm->_value = out[q];
  }
  return out;
}

void Simulation::outputNffft() const {
  if (auto *bb = dynamic_cast<NffftBroadBand *>(_nffft.get())) {
    if (_engine && !bb->getOutputDirPath().empty()) outputBroadBand(*bb);
    return;
  }
  auto *nf = dynamic_cast<NffftFd *>(_nffft.get());
  if (!_engine || !nf || nf->getOutputDirPath().empty()) return;
  const FarField r = calculateFarField();
  const std::string dir = nf->getOutputDirPath();
  try {
    makeDirs(dir);
  } catch (const std::exception &) {
    std::fprintf(stderr, "Error: cannot create directory %s\n", dir.c_str());
    return;
  }
  const std::vector<size_t> shp = {nf->frequencies().size(), nf->theta().size(), nf->phi().size()};
  auto p = [&](const char *name) { return dir + "/" + name; };
  writeNpy(p("frequencies.npy"), nf->frequencies().data(), "<f8", 8, {nf->frequencies().size()});
  writeNpy(p("theta.npy"), nf->theta().data(), "<f8", 8, {nf->theta().size()});
  writeNpy(p("phi.npy"), nf->phi().data(), "<f8", 8, {nf->phi().size()});
  writeNpy(p("e_theta.npy"), r.e_theta.data(), "<c16", 16, shp);
  writeNpy(p("e_phi.npy"), r.e_phi.data(), "<c16", 16, shp);
  writeNpy(p("h_theta.npy"), r.h_theta.data(), "<c16", 16, shp);
  writeNpy(p("h_phi.npy"), r.h_phi.data(), "<c16", 16, shp);
  writeNpy(p("n_theta.npy"), r.n_theta.data(), "<c16", 16, shp);
  writeNpy(p("n_phi.npy"), r.n_phi.data(), "<c16", 16, shp);
  writeNpy(p("l_theta.npy"), r.l_theta.data(), "<c16", 16, shp);
  writeNpy(p("l_phi.npy"), r.l_phi.data(), "<c16", 16, shp);
  writeNpy(p("power_theta.npy"), r.power_theta.data(), "<f8", 8, shp);
  writeNpy(p("power_phi.npy"), r.power_phi.data(), "<f8", 8, shp);
}

// NffftBroadBand::outputData and outputFarFieldParameters
// (nffft_broadband.cpp:156-200, 468-479)
void Simulation::outputBroadBand(const NffftBroadBand &bb) const {
  const BroadBandFarField r = calculateBroadBandFarField();
  const std::string dir = bb.getOutputDirPath();
  try {
    makeDirs(dir);
  } catch (const std::exception &) {
    std::fprintf(stderr, "Error: cannot create directory %s\n", dir.c_str());
    return;
  }
  {
    std::ofstream w(dir + "/far_field_parameter.dat");
    w << "Fs:\t" << 1 / _dt << '\n';
    w << "number of samples:\t" << r.frequencies.size() << '\n';
    w << "Theta:\n" << bb.theta() << '\n';
    w << "Phi:\n" << bb.phi() << '\n';
  }
  const std::vector<size_t> shp = {r.frequencies.size()};
  auto p = [&](const char *name) { return dir + "/" + name; };
  if (_two_d) {  // nffft_broadband.cpp:173-179: the 2-D far field only
    writeNpy(p("EZ.npy"), r.ez.data(), "<c16", 16, shp);
    return;
  }
  writeNpy(p("e_theta.npy"), r.e_theta.data(), "<c16", 16, shp);
  writeNpy(p("e_phi.npy"), r.e_phi.data(), "<c16", 16, shp);
  writeNpy(p("h_theta.npy"), r.h_theta.data(), "<c16", 16, shp);
  writeNpy(p("h_phi.npy"), r.h_phi.data(), "<c16", 16, shp);
  writeNpy(p("power_theta.npy"), r.power_theta.data(), "<f8", 8, shp);
  writeNpy(p("power_phi.npy"), r.power_phi.data(), "<f8", 8, shp);
}

void Simulation::outputTFSFIncidentWaveFastFourierTransform(const std::string &path) const {
  if (!_tfsf) return;
  try {
    makeDirs(path);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Error: %s\n", e.what());
    return;
  }
  const auto [f, v] = _tfsf->getIncidentWaveFastFourierTransform();
  writeNpy(path + "/frequencies.npy", f.data(), "<f8", 8, {f.size()});
  writeNpy(path + "/value.npy", v.data(), "<c16", 16, {v.size()});
}

std::vector<double> Simulation::getProbeData() const {
  if (!_engine) throw std::runtime_error("Simulation is not initialised");
  std::vector<double> out(_probes.size() * _total_time_steps);
  if (!_probes.empty()) check(xf_engine_get_probes(_engine, out.data(), (int64_t)_total_time_steps));
  return out;
}

}  // namespace xfdtd
