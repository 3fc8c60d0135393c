// xfdtd.hpp -- C++ host API of the MI355X FDTD engine.
//
// The classes keep the names, constructor arguments and call order of the
// reference (Mrwatermolen/XFDTD, src/include/...), so a scatter program written
// against the reference builds against this header with one change: xtensor
// arrays become std::vector<double> (xtensor is not a dependency here), and
// field/coefficient getters return host copies instead of references into
// host memory, because the fields live in HBM.
//
//   Simulation      include/simulation/simulation.h:23-170
//   Shape/Cube/Sphere include/shape/*.h
//   Material        include/material/material.h:14-78
//   Object          include/object/object.h
//   Boundary/PML    include/boundary/boundary.h, perfect_match_layer.h
//   Waveform/GaussianWaveform  include/waveform/*.h
//   TFSF/TFSF3D     include/tfsf/tfsf.h, tfsf_3d.h
//   NFFFT/NffftFd   include/nffft/nffft.h, nffft_fd.h
//   LumpedElement/CurrentSource include/lumped_element/*.h
//
// Set-up runs on the CPU in FP64 exactly as the reference computes it; each
// time step runs on the GPU through the C-ABI in include/xfdtd_engine.h.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <map>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

struct xf_engine;

namespace xfdtd {

struct PlanData;

namespace constant {
inline constexpr double PI{M_PI};
inline constexpr double EPSILON_0{8.854187817e-12};
inline constexpr double MU_0{4 * PI * 1e-7};
inline constexpr double SQUARED_C_0{1 / (EPSILON_0 * MU_0)};
inline const double ETA_0{std::sqrt(MU_0 / EPSILON_0)};
inline const double C_0{std::sqrt(SQUARED_C_0)};
inline constexpr double TOLERABLE_EPSILON{10 * std::numeric_limits<double>::epsilon()};
}  // namespace constant

using PointVector = std::array<double, 3>;
using SpatialIndex = size_t;
enum class Orientation { XN, XP, YN, YP, ZN, ZP };
enum class Axis { X, Y, Z };
enum class EMComponent { EX, EY, EZ, HX, HY, HZ };
enum class PlaneType { XY, YZ, ZX };  // util/type_define.h:30

// ------------------------------------------------------------------ shapes
class Shape {
 public:
  virtual ~Shape() = default;
  virtual bool isPointInside(const PointVector &point) const = 0;
  virtual std::unique_ptr<Shape> getWrappedBox() const = 0;
  virtual std::unique_ptr<Shape> clone() const = 0;
  // shape/shape.cpp:4, cube.cpp:16-21, sphere.cpp:17-21: the description
  // Simulation::checkRun writes
  virtual explicit operator std::string() const { return "Shape: "; }
};

class Cube : public Shape {
 public:
  Cube(PointVector point, PointVector size) : _point{point}, _size{size} {}
  bool isPointInside(const PointVector &point) const override;
  std::unique_ptr<Shape> getWrappedBox() const override { return std::make_unique<Cube>(*this); }
  std::unique_ptr<Shape> clone() const override { return std::make_unique<Cube>(*this); }
  double getXmin() const { return _point[0]; }
  double getYmin() const { return _point[1]; }
  double getZmin() const { return _point[2]; }
  double getXmax() const { return _point[0] + _size[0]; }
  double getYmax() const { return _point[1] + _size[1]; }
  double getZmax() const { return _point[2] + _size[2]; }
  PointVector getSize() const { return _size; }
  PointVector getPoint() const { return _point; }
  explicit operator std::string() const override {
    return "Cube: " + std::to_string(_point[0]) + " " + std::to_string(_point[1]) + " " +
           std::to_string(_point[2]) + " " + std::to_string(_size[0]) + " " + std::to_string(_size[1]) + " " +
           std::to_string(_size[2]);
  }

 private:
  PointVector _point, _size;
};

class Sphere : public Shape {
 public:
  Sphere(PointVector center, double radius) : _center{center}, _radius{radius} {}
  bool isPointInside(const PointVector &point) const override;
  std::unique_ptr<Shape> getWrappedBox() const override;
  std::unique_ptr<Shape> clone() const override { return std::make_unique<Sphere>(*this); }
  double getRadius() const { return _radius; }
  PointVector getCenter() const { return _center; }
  explicit operator std::string() const override {
    return "Sphere: " + std::to_string(_center[0]) + " " + std::to_string(_center[1]) + " " +
           std::to_string(_center[2]) + " " + std::to_string(_radius);
  }

 private:
  PointVector _center;
  double _radius;
};

// shape/cylinder.cpp: a cylinder along `axis`; its wrapped box spans the
// radius across and height/2 either side of the centre along the axis.  A
// 2-D scene's cylinder has height 0 in the z = 0 plane.
class Cylinder : public Shape {
 public:
  Cylinder(Axis axis, PointVector center, double radius, double height)
      : _axis{axis}, _center{center}, _radius{radius}, _height{height} {}
  bool isPointInside(const PointVector &point) const override;
  std::unique_ptr<Shape> getWrappedBox() const override;
  std::unique_ptr<Shape> clone() const override { return std::make_unique<Cylinder>(*this); }
  Axis getAxis() const { return _axis; }
  double getRadius() const { return _radius; }
  double getHeight() const { return _height; }
  PointVector getCenter() const { return _center; }

 private:
  Axis _axis;
  PointVector _center;
  double _radius, _height;
};

// --------------------------------------------------------------- material
class Material {
 public:
  Material(std::string_view name, float eps_r, float mu_r, double sigma_e, double sigma_m,
           bool is_dispersion = false)
      : _name{name}, _eps_r{eps_r}, _mu_r{mu_r}, _eps{_eps_r * constant::EPSILON_0},
        _mu{_mu_r * constant::MU_0}, _sigma_e{sigma_e}, _sigma_m{sigma_m},
        _is_dispersion{is_dispersion} {}
  static std::unique_ptr<Material> createAir(const std::string &name = "XFDTD_AIR") {
    return std::make_unique<Material>(name, 1, 1, 0, 0);
  }
  static std::unique_ptr<Material> createPEC() {
    return std::make_unique<Material>("XFDTD_PEC", 1, 1, 1e10, 0);
  }
  virtual ~Material() = default;
  double getPermittivityE() const { return _eps; }
  double getPermeabilityM() const { return _mu; }
  double getElectricalConductivity() const { return _sigma_e; }
  double getMagneticConductivity() const { return _sigma_m; }
  bool isDispersion() const { return _is_dispersion; }
  const std::string &name() const { return _name; }
  // material/material.cpp:18-27
  explicit operator std::string() const {
    return "Material: " + _name + "\neps_r: " + std::to_string(_eps_r) + "\nmu_r: " + std::to_string(_mu_r) +
           "\neps: " + std::to_string(_eps) + "\nmu: " + std::to_string(_mu) + "\nsigma_e: " +
           std::to_string(_sigma_e) + "\nsigma_m: " + std::to_string(_sigma_m) +
           "\nis_dispersion: " + std::to_string(static_cast<int>(_is_dispersion));
  }
  // Drude parameters of a DrudeMedium (eps_inf, omega_p, nv); a plain
  // Material holds 1, 0, 0 and is not dispersive
  double drudeEpsInf() const { return _drude_eps_inf; }
  double drudeOmegaP() const { return _drude_omega_p; }
  double drudeNv() const { return _drude_nv; }
  // the dispersive kind: 0 none, 1 Drude, 2 Debye, 3 Lorentz; the Debye and
  // Lorentz static permittivity and the Debye relaxation time
  int dispersionKind() const { return _disp_kind; }
  double dispersionEpsS() const { return _disp_eps_s; }
  double dispersionTau() const { return _disp_tau; }

 protected:
  double _drude_eps_inf = 1, _drude_omega_p = 0, _drude_nv = 0;
  int _disp_kind = 0;
  double _disp_eps_s = 1, _disp_tau = 0;

 private:
  std::string _name;
  float _eps_r, _mu_r;
  double _eps, _mu, _sigma_e, _sigma_m;
  bool _is_dispersion;
};

// material/dispersive_material.h:102-131: a LinearDispersiveMaterial
// (eps_r = mu_r = 1 in the standard arrays) whose cells take the Drude E
// update.  The parameters live in the Material base, so an Object built from
// a copy keeps them.
class DrudeMedium : public Material {
 public:
  DrudeMedium(std::string_view name, double eps_inf, double pole_omega, double nv, double sigma_e = 0,
              double sigma_m = 0)
      : Material(name, 1, 1, sigma_e, sigma_m, true) {
    _drude_eps_inf = eps_inf;
    _drude_omega_p = pole_omega;
    _drude_nv = nv;
    _disp_kind = 1;
  }
};

// material/dispersive_material.h:137-141, dispersive_material.cpp:195-275:
// eps(w) = eps_inf + (eps_s - eps_inf) / (1 + i w tau)
class DebyMedium : public Material {
 public:
  DebyMedium(std::string_view name, double eps_s, double eps_inf, double tau, double sigma_e = 0,
             double sigma_m = 0)
      : Material(name, 1, 1, sigma_e, sigma_m, true) {
    _drude_eps_inf = eps_inf;
    _disp_eps_s = eps_s;
    _disp_tau = tau;
    _disp_kind = 2;
  }
};

// material/dispersive_material.h:59-100, dispersive_material.cpp:7-117:
// eps(w) = eps_inf + (eps_s - eps_inf) omega_p^2 / (omega_p^2 + 2 i w nv - w^2)
class LorentzMedium : public Material {
 public:
  LorentzMedium(std::string_view name, double eps_s, double eps_inf, double omega_p, double nv,
                double sigma_e = 0, double sigma_m = 0)
      : Material(name, 1, 1, sigma_e, sigma_m, true) {
    _drude_eps_inf = eps_inf;
    _drude_omega_p = omega_p;
    _drude_nv = nv;
    _disp_eps_s = eps_s;
    _disp_kind = 3;
  }
};

class Object {
 public:
  Object(std::string_view name, std::unique_ptr<Shape> shape, Material material)
      : _name{name}, _shape{std::move(shape)}, _material{std::make_unique<Material>(material)} {}
  Object(std::string_view name, std::unique_ptr<Shape> shape, std::unique_ptr<Material> material)
      : _name{name}, _shape{std::move(shape)}, _material{std::move(material)} {}
  virtual ~Object() = default;
  std::unique_ptr<Shape> getWrappedBox() const { return _shape->getWrappedBox(); }
  const Shape *getShape() const { return _shape.get(); }
  const Material &getMaterial() const { return *_material; }
  const std::string &getName() const { return _name; }
  // object/object.cpp:43-47
  explicit operator std::string() const {
    return "Object: " + _name + "\n" + static_cast<std::string>(*_shape) + "\n" +
           static_cast<std::string>(*_material);
  }

 private:
  std::string _name;
  std::unique_ptr<Shape> _shape;
  std::unique_ptr<Material> _material;
};

// object/object_plane.cpp:4-50: a cube with zero size along one axis that
// sets sigma_e of the two tangential E components on that plane (and
// nothing else) in place of the cell-centre material of an Object
class ObjectPlane : public Object {
 public:
  ObjectPlane(std::string_view name, std::unique_ptr<Shape> shape, std::unique_ptr<Material> material)
      : Object(name, std::move(shape), std::move(material)) {}
  ObjectPlane(std::string_view name, std::unique_ptr<Shape> shape, Material material)
      : Object(name, std::move(shape), std::move(material)) {}
};

// --------------------------------------------------------------- waveform
class Waveform {
 public:
  explicit Waveform(double amplitude = 1.0) : _amplitude{amplitude} {}
  virtual ~Waveform() = default;
  virtual std::unique_ptr<Waveform> clone() const = 0;
  virtual double getValueByTime(double time) const = 0;
  virtual void init(const std::vector<double> &time_array);
  double operator()(size_t step) const { return _value.at(step); }
  const std::vector<double> &getValue() const { return _value; }
  double getAmplitude() const { return _amplitude; }
  void setAmplitude(double amplitude) { _amplitude = amplitude; }

 private:
  double _amplitude;
  std::vector<double> _value;
};

class GaussianWaveform : public Waveform {
 public:
  GaussianWaveform(double amplitude, double tau, double t_0)
      : Waveform{amplitude}, _tau{tau}, _t_0{t_0} {}
  GaussianWaveform(double tau, double t_0) : _tau{tau}, _t_0{t_0} {}
  std::unique_ptr<Waveform> clone() const override { return std::make_unique<GaussianWaveform>(*this); }
  double getValueByTime(double time) const override;

 private:
  double _tau, _t_0;
};

// waveform/sinusoidal_waveform.cpp: amplitude sin(2 pi f t + phase)
class SinusoidalWaveform : public Waveform {
 public:
  SinusoidalWaveform(double amplitude, double frequency, double phase)
      : Waveform{amplitude}, _frequency{frequency}, _phase{phase} {}
  SinusoidalWaveform(double frequency, double phase) : _frequency{frequency}, _phase{phase} {}
  std::unique_ptr<Waveform> clone() const override { return std::make_unique<SinusoidalWaveform>(*this); }
  double getValueByTime(double time) const override;

 private:
  double _frequency, _phase;
};

// waveform/cosine_modulated_gaussian_waveform.cpp:
// amplitude cos(2 pi f (t - t_0)) exp(-((t - t_0) / tau)^2)
class CosineModulatedGaussianWaveform : public Waveform {
 public:
  CosineModulatedGaussianWaveform(double amplitude, double tau, double t_0, double modulation_frequency)
      : Waveform{amplitude}, _tau{tau}, _t_0{t_0}, _f{modulation_frequency} {}
  CosineModulatedGaussianWaveform(double tau, double t_0, double modulation_frequency)
      : _tau{tau}, _t_0{t_0}, _f{modulation_frequency} {}
  std::unique_ptr<Waveform> clone() const override {
    return std::make_unique<CosineModulatedGaussianWaveform>(*this);
  }
  double getValueByTime(double time) const override;

 private:
  double _tau, _t_0, _f;
};

// --------------------------------------------------------------- boundary
class Boundary {
 public:
  virtual ~Boundary() = default;
  virtual SpatialIndex getSize() const = 0;
  virtual Orientation getOrientation() const = 0;
};

class PML : public Boundary {
 public:
  PML(Orientation orientation, int thickness, int order = 4, double sigma_ratio = 1,
      double alpha_min = 0, double alpha_max = 0.05, double kappa_max = 10)
      : _orientation{orientation}, _thickness{thickness}, _order{order}, _sigma_ratio{sigma_ratio},
        _alpha_min{alpha_min}, _alpha_max{alpha_max}, _kappa_max{kappa_max} {}
  SpatialIndex getSize() const override { return (SpatialIndex)_thickness; }
  Orientation getOrientation() const override { return _orientation; }
  bool isOrientatingPositive() const {
    return _orientation == Orientation::XP || _orientation == Orientation::YP ||
           _orientation == Orientation::ZP;
  }
  // profile tables of perfect_match_layer.cpp:133-155 / 302-325 (per layer)
  struct Profile {
    std::vector<double> b_e, a_e, b_m, a_m, kappa_e, kappa_m;
  };
  Profile computeProfile(double dl, double dt) const;

 private:
  Orientation _orientation;
  int _thickness, _order;
  double _sigma_ratio, _alpha_min, _alpha_max, _kappa_max;
};

// ------------------------------------------------------------------- TFSF
class TFSF {
 public:
  TFSF(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z, double e_0,
       double theta_inc, double phi_inc, double psi, std::shared_ptr<Waveform> waveform);
  virtual ~TFSF() = default;
  double getIncidentFieldAmplitude() const { return _e_0; }
  double getIncidentTheta() const { return _theta_inc; }
  double getIncidentPhi() const { return _phi_inc; }
  double getIncidentPsi() const { return _psi; }
  PointVector getKVector() const { return _k; }
  std::array<SpatialIndex, 3> getDistance() const { return {_distance_x, _distance_y, _distance_z}; }
  // tfsf/tfsf.cpp:63-66 with util/dft.h
  std::vector<std::complex<double>> getIncidentWaveFourierTransform(
      const std::vector<double> &frequencies) const;
  // tfsf/tfsf.cpp:68-78: the waveform (without e_0) zero-padded to 2T-1
  // points, fftshift(fft(w)), and rfftfreq(2T-1, dt) (T entries, as the
  // reference writes them)
  std::pair<std::vector<double>, std::vector<std::complex<double>>> getIncidentWaveFastFourierTransform() const;
  const std::vector<double> &getIncidentWaveValue() const { return _waveform->getValue(); }
  Waveform &waveform() { return *_waveform; }
  double dt() const { return _dt; }

 protected:
  SpatialIndex _distance_x, _distance_y, _distance_z;
  double _e_0, _theta_inc, _phi_inc, _psi;
  double _sin_theta_inc, _cos_theta_inc, _sin_phi_inc, _cos_phi_inc, _sin_psi, _cos_psi;
  PointVector _k;
  std::shared_ptr<Waveform> _waveform;
  double _dt = 0;
  friend class Simulation;
};

class TFSF3D : public TFSF {
 public:
  TFSF3D(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z, double e_0,
         double theta_inc, double phi_inc, double psi, std::shared_ptr<Waveform> waveform);
  PointVector getTransformE() const { return _transform_e; }
  PointVector getTransformH() const { return _transform_h; }

 private:
  PointVector _transform_e, _transform_h;
  friend class Simulation;
};

// tfsf/tfsf_2d.cpp:10-24: the TMz plane wave of a 2-D scene (a grid with
// one z cell): theta = pi/2, psi = 0, travelling along phi_inc; its Ez is
// -ez_i0 * waveform.  Only Ez, Hx and Hy take face corrections.
class TFSF2D : public TFSF {
 public:
  TFSF2D(SpatialIndex distance_x, SpatialIndex distance_y, double phi_inc, double ez_i0,
         std::shared_ptr<Waveform> waveform);
  PointVector getTransformE() const { return _transform_e; }
  PointVector getTransformH() const { return _transform_h; }

 private:
  PointVector _transform_e, _transform_h;
  friend class Simulation;
};

// ------------------------------------------------------------------ NF-FF
class NFFFT {
 public:
  NFFFT(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z,
        std::string output_dir_path)
      : _distance_x{distance_x}, _distance_y{distance_y}, _distance_z{distance_z},
        _output_dir_path{std::move(output_dir_path)} {}
  virtual ~NFFFT() = default;
  std::array<SpatialIndex, 3> getDistance() const { return {_distance_x, _distance_y, _distance_z}; }
  const std::string &getOutputDirPath() const { return _output_dir_path; }

 protected:
  SpatialIndex _distance_x, _distance_y, _distance_z;
  std::string _output_dir_path;
};

// far-field results of NffftFd::calculateFarfield (nffft_fd.cpp:379-421),
// each [nf][ntheta][nphi] C order
struct FarField {
  std::vector<std::complex<double>> n_theta, n_phi, l_theta, l_phi, e_theta, e_phi, h_theta, h_phi;
  std::vector<double> power_theta, power_phi;
};

class NffftFd : public NFFFT {
 public:
  NffftFd(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z,
          std::vector<double> frequencies, std::vector<double> theta, std::vector<double> phi,
          std::string output_dir_path);
  const std::vector<double> &frequencies() const { return _frequencies; }
  const std::vector<double> &theta() const { return _theta; }
  const std::vector<double> &phi() const { return _phi; }
  const std::vector<double> &waveNumber() const { return _wave_number; }

 private:
  std::vector<double> _frequencies, _wave_number, _theta, _phi;
};

// NffftBroadBand (nffft/nffft_broadband.cpp:21-28): the NF-FF surface of
// NffftFd observed in one direction (theta, phi) over every frequency bin of
// the run, the monostatic radar cross-section path.
class NffftBroadBand : public NFFFT {
 public:
  NffftBroadBand(SpatialIndex distance_x, SpatialIndex distance_y, SpatialIndex distance_z, double theta,
                 double phi, std::string output_dir_path);
  double theta() const { return _theta; }
  double phi() const { return _phi; }

 private:
  double _theta, _phi;
};

// NffftBroadBand::calculateFarField (nffft_broadband.cpp:316-345): every
// array has one entry per bin, 2T-1 bins from -fs/2 to fs/2 (fs = 1/dt)
struct BroadBandFarField {
  std::vector<double> frequencies, wave_number;
  std::vector<std::complex<double>> n_x, n_y, n_z, l_x, l_y, l_z;
  std::vector<std::complex<double>> n_theta, n_phi, l_theta, l_phi, e_theta, e_phi, h_theta, h_phi;
  std::vector<double> power_theta, power_phi;
  // 2-D grids: the far-zone Ez of nffft_broadband.cpp:173-179,
  // 0.5 sqrt(i k) (l_phi + eta0 n_theta) (empty on a 3-D grid)
  std::vector<std::complex<double>> ez;
};

// ---------------------------------------------------------- lumped element
class LumpedElement {
 public:
  explicit LumpedElement(std::unique_ptr<Cube> cube) : _cube{std::move(cube)} {}
  virtual ~LumpedElement() = default;
  const Cube &getCube() const { return *_cube; }

 protected:
  std::unique_ptr<Cube> _cube;
};

class CurrentSource : public LumpedElement {
 public:
  CurrentSource(std::unique_ptr<Cube> cube, Orientation orientation, double resistance,
                std::unique_ptr<Waveform> waveform)
      : LumpedElement{std::move(cube)}, _orientation{orientation},
        _resistance{resistance == 0 ? 1e-20 : resistance}, _waveform{std::move(waveform)} {}
  Orientation getOrientation() const { return _orientation; }
  double getResistance() const { return _resistance; }
  Waveform &waveform() { return *_waveform; }

 private:
  Orientation _orientation;
  double _resistance;
  std::unique_ptr<Waveform> _waveform;
  friend class Simulation;
};

// lumped_element/voltage_source.cpp:12-22: a voltage source of the given
// internal resistance (0 becomes 1e-20), spread over the cells along its axis
class VoltageSource : public LumpedElement {
 public:
  VoltageSource(std::unique_ptr<Cube> cube, Orientation orientation, double resistance,
                std::unique_ptr<Waveform> waveform)
      : LumpedElement{std::move(cube)}, _orientation{orientation},
        _resistance{resistance == 0 ? 1e-20 : resistance}, _waveform{std::move(waveform)} {}
  Orientation getOrientation() const { return _orientation; }
  double getResistance() const { return _resistance; }
  Waveform &waveform() { return *_waveform; }

 private:
  Orientation _orientation;
  double _resistance;
  std::unique_ptr<Waveform> _waveform;
};

// lumped_element/resistor.cpp, capacitor.cpp, inductor.cpp: a resistance,
// capacitance or inductance along one axis of the cube's cells
class Resistor : public LumpedElement {
 public:
  Resistor(std::unique_ptr<Cube> cube, Axis axis, double resistance)
      : LumpedElement{std::move(cube)}, _axis{axis}, _resistance{resistance} {}
  Axis getAxis() const { return _axis; }
  double getResistance() const { return _resistance; }

 private:
  Axis _axis;
  double _resistance;
};

class Capacitor : public LumpedElement {
 public:
  Capacitor(std::unique_ptr<Cube> cube, Axis axis, double capacitance)
      : LumpedElement{std::move(cube)}, _axis{axis}, _capacitance{capacitance} {}
  Axis getAxis() const { return _axis; }
  double getCapacitance() const { return _capacitance; }

 private:
  Axis _axis;
  double _capacitance;
};

class Inductor : public LumpedElement {
 public:
  Inductor(std::unique_ptr<Cube> cube, Axis axis, double inductance)
      : LumpedElement{std::move(cube)}, _axis{axis}, _inductance{inductance} {}
  Axis getAxis() const { return _axis; }
  double getInductance() const { return _inductance; }

 private:
  Axis _axis;
  double _inductance;
};

// ---------------------------------------------------------------- monitor
// Point sample of one component after every step (the monitor slot of the
// run loop, simulation_update.cpp:215-219).
struct Probe {
  EMComponent component;
  SpatialIndex i, j, k;
};

// monitor/monitor.h:11-67: an output location for field data
class Monitor {
 public:
  Monitor(std::unique_ptr<Cube> shape, std::string output_dir_path, std::string output_file_name)
      : _shape{std::move(shape)}, _dir{std::move(output_dir_path)}, _file{std::move(output_file_name)} {}
  virtual ~Monitor() = default;
  virtual const std::string &getOutputPath() const { return _dir; }
  virtual const std::string &getOutputFileName() const { return _file; }
  virtual void setOutputDirPath(const std::string &p) { _dir = p; }
  virtual void setOutputFileName(const std::string &f) { _file = f; }
  const Cube *getShape() const { return _shape.get(); }

 private:
  std::unique_ptr<Cube> _shape;
  std::string _dir, _file;
};

// monitor/field_monitor.h:14-38, field_monitor.cpp:15-121: one component
// over the grid box of `shape` (one cell thick across the plane), dumped as
// .npy (float64, C order) by outputData
class Simulation;

// A monitor of the caller's own: the reference's Monitor virtuals init /
// update / outputData (monitor/monitor.h:28-37) as host-side hooks.  init(sim)
// runs once the engine holds the scene, update(sim, t) after every time step
// t (the step's fields readable through sim.getField / sim.getFieldBox),
// outputData() from Simulation::outputData.  The engine waits for each step
// while a host monitor is present, so on large grids read boxes, not whole
// fields.  (Boundaries have no such hook: the CPML runs inside the curl
// kernels, and a host-side updateE / updateH would move the fields across
// PCIe twice a step; addBoundary takes PML only.)
class HostMonitor : public Monitor {
 public:
  explicit HostMonitor(std::string output_dir_path = "", std::string output_file_name = "")
      : Monitor(nullptr, std::move(output_dir_path), std::move(output_file_name)) {}
  virtual void init(const Simulation &) {}
  virtual void update(const Simulation &sim, size_t time_step) = 0;
  virtual void outputData() {}
};

class TimeDomainFieldMonitor : public Monitor {
 public:
  TimeDomainFieldMonitor(std::unique_ptr<Cube> shape, PlaneType plane_type, EMComponent component,
                         std::string output_dir_path, std::string output_file_name)
      : Monitor(std::move(shape), std::move(output_dir_path), std::move(output_file_name)),
        _plane_type{plane_type}, _component{component} {}
  PlaneType getPlaneType() const { return _plane_type; }
  EMComponent getComponent() const { return _component; }

 private:
  PlaneType _plane_type;
  EMComponent _component;
};

// monitor/movie_monitor.h, movie_monitor.cpp:46-59: the wrapped monitor's
// data after every step t with t % plot_step == 0 or t == total_time_steps - 1,
// as "%04d.npy" (t) in the wrapped monitor's directory
class MovieMonitor : public Monitor {
 public:
  MovieMonitor(std::unique_ptr<Monitor> monitor, size_t total_time_steps, size_t plot_step)
      : Monitor(nullptr, "", ""), _monitor{std::move(monitor)}, _total_time_steps{total_time_steps},
        _plot_step{plot_step} {}
  const std::string &getOutputPath() const override { return _monitor->getOutputPath(); }
  const std::string &getOutputFileName() const override { return _monitor->getOutputFileName(); }
  void setOutputDirPath(const std::string &p) override { _monitor->setOutputDirPath(p); }
  void setOutputFileName(const std::string &f) override { _monitor->setOutputFileName(f); }
  Monitor &getMonitor() const { return *_monitor; }
  size_t getTotalTimeSteps() const { return _total_time_steps; }
  size_t getPlotStep() const { return _plot_step; }

 private:
  std::unique_ptr<Monitor> _monitor;
  size_t _total_time_steps, _plot_step;
};

// monitor/voltage_monitor.cpp, monitor/current_monitor.cpp: the voltage
// along the orientation's axis over the cube's grid box, or the current
// through it (the loop integral of H around the box at its start plane),
// one value per step; outputData writes [time, value] (2 x T, float64)
class VoltageMonitor : public Monitor {
 public:
  VoltageMonitor(std::unique_ptr<Cube> shape, Orientation orientation, std::string output_dir_path,
                 std::string output_file_name)
      : Monitor(std::move(shape), std::move(output_dir_path), std::move(output_file_name)),
        _orientation{orientation} {}
  Orientation getOrientation() const { return _orientation; }
  double getValue(size_t i) const { return _value.at(i); }
  const std::vector<double> &getValueArray() const { return _value; }

 private:
  Orientation _orientation;
  std::vector<double> _value;
  friend class Simulation;
};

class CurrentMonitor : public Monitor {
 public:
  CurrentMonitor(std::unique_ptr<Cube> shape, Orientation orientation, std::string output_dir_path,
                 std::string output_file_name)
      : Monitor(std::move(shape), std::move(output_dir_path), std::move(output_file_name)),
        _orientation{orientation} {}
  Orientation getOrientation() const { return _orientation; }
  double getValue(size_t i) const { return _value.at(i); }
  const std::vector<double> &getValueArray() const { return _value; }

 private:
  Orientation _orientation;
  std::vector<double> _value;
  friend class Simulation;
};

// ---------------------------------------------------------------- network
// network/port.cpp: a port of reference impedance z measured by a voltage
// and a current monitor; a = (V + z I) / (2 sqrt(Re z)), b = (V - conj(z) I)
// / (2 sqrt(Re z)) of the monitors' DFTs (util/dft.h)
class Port {
 public:
  Port(int index, std::complex<double> impedance, bool is_activated_source, std::shared_ptr<VoltageMonitor> voltage,
       std::shared_ptr<CurrentMonitor> current)
      : _index{index}, _impedance{impedance}, _is_source{is_activated_source}, _voltage{std::move(voltage)},
        _current{std::move(current)} {}
  void calculateSParameter(const std::vector<double> &frequencies, double dt);
  int getPortIndex() const { return _index; }
  bool isActivatedSource() const { return _is_source; }
  const std::vector<std::complex<double>> &getA() const { return _a; }
  const std::vector<std::complex<double>> &getB() const { return _b; }
  const std::shared_ptr<VoltageMonitor> &voltageMonitor() const { return _voltage; }
  const std::shared_ptr<CurrentMonitor> &currentMonitor() const { return _current; }

 private:
  int _index;
  std::complex<double> _impedance;
  bool _is_source;
  std::shared_ptr<VoltageMonitor> _voltage;
  std::shared_ptr<CurrentMonitor> _current;
  std::vector<std::complex<double>> _a, _b;
};

// network/network.cpp: S_ji = b_j / a_i for every activated source port i,
// written as s{10 j + i}.npy = [frequencies; S] (2 x nf, complex128)
class Network {
 public:
  Network(std::vector<std::unique_ptr<Port>> ports, std::vector<double> frequencies, std::string output_path)
      : _ports{std::move(ports)}, _frequencies{std::move(frequencies)}, _output_path{std::move(output_path)} {}
  const std::vector<std::unique_ptr<Port>> &ports() const { return _ports; }
  const std::vector<double> &frequencies() const { return _frequencies; }
  const std::string &getOutputPath() const { return _output_path; }
  // S parameters keyed 10 j + i, from the ports' monitor data
  std::map<int, std::vector<std::complex<double>>> calculateSParameters(double dt);
  void outputData(double dt);

 private:
  std::vector<std::unique_ptr<Port>> _ports;
  std::vector<double> _frequencies;
  std::string _output_path;
};

// .npy files (xt::dump_npy in the reference): float64 or complex128, C
// order, shape {v.size()} when `shape` is empty; the file's directory is
// created when missing
void dumpNpy(const std::string &path, const std::vector<double> &v, const std::vector<size_t> &shape = {});
void dumpNpy(const std::string &path, const std::vector<std::complex<double>> &v,
             const std::vector<size_t> &shape = {});

// util/dft.h: sum_j x_j exp(-2 pi i f j dt) dt for every f
std::vector<std::complex<double>> dft(const std::vector<double> &x, double dt,
                                      const std::vector<double> &frequencies);

// ------------------------------------------------------------- simulation
class Simulation {
 public:
  inline static constexpr float DEFAULT_CFL{0.99};

  explicit Simulation(double cell_size, float cfl = DEFAULT_CFL);
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<Boundary>> boundaries, std::unique_ptr<TFSF> tfsf,
             float cfl = DEFAULT_CFL);
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<Boundary>> boundaries, std::unique_ptr<TFSF> tfsf,
             std::unique_ptr<NFFFT> nffft, float cfl = DEFAULT_CFL);
  // include/simulation/simulation.h:34-49: lumped-element circuits
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
             std::vector<std::shared_ptr<Boundary>> boundaries = {},
             std::vector<std::shared_ptr<Monitor>> monitors = {}, float cfl = DEFAULT_CFL);
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
             std::vector<std::shared_ptr<Monitor>> monitors, float cfl = DEFAULT_CFL);
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<Monitor>> monitors, float cfl = DEFAULT_CFL);
  Simulation(double cell_size, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<LumpedElement>> lumped_elements,
             std::vector<std::shared_ptr<Boundary>> boundaries, std::vector<std::shared_ptr<Monitor>> monitors,
             std::unique_ptr<Network> network, float cfl = DEFAULT_CFL);
  ~Simulation();
  Simulation(const Simulation &) = delete;
  Simulation &operator=(const Simulation &) = delete;

  void addObject(std::shared_ptr<Object> object);
  void addBoundary(std::shared_ptr<Boundary> boundary);
  void addLumpedElement(std::shared_ptr<LumpedElement> lumped_element);
  void addTFSFSource(std::shared_ptr<TFSF> tfsf);
  void addNFFFT(std::shared_ptr<NFFFT> nffft);
  void addProbe(Probe probe);
  void addMonitor(std::shared_ptr<Monitor> monitor);
  void addNetwork(std::shared_ptr<Network> network);
  // values of the voltage and current monitors (in the order added) after
  // the steps run so far; also stored in the monitor objects
  std::vector<std::vector<double>> getLineMonitorData() const;

  // ---- differentiation (no counterpart in the reference)
  // Per-cell relative permittivity of the cells [box[0],box[1]) x
  // [box[2],box[3]) x [box[4],box[5]) (grid indices), lossless, mu_r = 1,
  // applied after the objects as one-cell Objects would be; eps_r in C order.
  void setDesign(const std::array<size_t, 6> &box, std::vector<float> eps_r);
  // after init: new eps_r for the same design box (the engine keeps its
  // fields; restore a step-0 checkpoint before the next gradient run)
  void updateDesign(std::vector<float> eps_r);
  size_t designCells() const {
    return (_design_box[1] - _design_box[0]) * (_design_box[3] - _design_box[2]) * (_design_box[5] - _design_box[4]);
  }
  // running DFT of one component on a plane (axis fixed at `index`, the
  // other two axes in increasing order over [lo1,hi1) x [lo2,hi2)), sample
  // after step t weighted by dt exp(-i 2 pi f t dt); J = sum |F|^2
  void setDftPlane(EMComponent component, int axis, size_t index, size_t lo1, size_t hi1, size_t lo2,
                   size_t hi2, double frequency);
  std::vector<std::complex<double>> getDft() const;
  // J and dJ/d eps_r per design cell: forward run from step 0 with a state
  // checkpoint every `checkpoint_every` steps (0: automatic), then the
  // discrete adjoint back through each recomputed segment
  double gradEps(std::vector<double> &grad, int checkpoint_every = 0);
  // transpose of n steps of the field map on packed adjoint fields
  void adjointRun(size_t n, std::vector<double> &fields);
  void checkpoint(int slot);
  void restore(int slot);

  // Simulation::run (simulation_update.cpp:5-40): init, the time loop on the
  // GPU, then NF-FF output when an output directory is set.
  void run(size_t total_time_steps);

  // The same loop in pieces, for callers that interleave host work.
  // plan() does the CPU set-up only (no device needed); init() plans if
  // needed and uploads the tables to a new engine.
  void plan(size_t total_time_steps);
  void init(size_t total_time_steps);
  // Simulation::checkRun (simulation.cpp:98-113): the CPU set-up only, then
  // "simulation_check" (grid size) and "simulation_object_check" (every
  // object's description) in the working directory
  void checkRun(size_t total_time_steps);
  void step(size_t n);
  void sync();
  size_t currentTimeStep() const;
  size_t totalTimeSteps() const { return _total_time_steps; }

  // FP64 (default; parity with the reference) or FP32 fields.
  void setPrecision(int bits) { _precision = bits; }
  int precision() const { return _precision; }

  double getDt() const { return _dt; }
  SpatialIndex getNx() const { return _n[0]; }
  SpatialIndex getNy() const { return _n[1]; }
  SpatialIndex getNz() const { return _n[2]; }
  std::vector<double> getField(EMComponent c) const;
  // component c over the grid-index box [box[0],box[1]) x [box[2],box[3]) x
  // [box[4],box[5]), C order (only the box crosses to the host)
  std::vector<double> getFieldBox(EMComponent c, const std::array<size_t, 6> &box) const;
  void setField(EMComponent c, const std::vector<double> &v);
  std::array<size_t, 3> fieldShape(EMComponent c) const;

  // NF-FF accumulators and far field (NffftFd::calculateFarfield)
  std::vector<std::complex<double>> getNffftAccumulators() const;
  FarField calculateFarField() const;
  // NffftBroadBand: the stored face-current series ([T][8][Q], see
  // xf_engine_get_broadband_samples) and the far field of every bin; the
  // surface integrals use `wave_number` (2T-1 entries) when it is given,
  // else the bins' own
  std::vector<double> getBroadBandSamples() const;
  BroadBandFarField calculateBroadBandFarField(const std::vector<double> *wave_number = nullptr) const;
  // Simulation::outputData (simulation.cpp:144-156): NffftFd::outputData or
  // NffftBroadBand::outputData (nffft_broadband.cpp:156-200) when the NF-FF
  // has an output directory, then every field monitor's final
  // slice (movie monitors write their frames while stepping)
  void outputData() const;
  // simulation.cpp:170-188: frequencies.npy and value.npy of
  // TFSF::getIncidentWaveFastFourierTransform under `path`
  void outputTFSFIncidentWaveFastFourierTransform(const std::string &path) const;
  std::vector<double> getProbeData() const;  // [n_probes][total_steps]

  xf_engine *engine() const { return _engine; }
  // per-cell coefficient array `which` (0..17: cexe, cexhy, cexhz, ceye, ceyhz,
  // ceyhx, ceze, cezhx, cezhy, chxh, chxey, chxez, chyh, chyez, chyex, chzh,
  // chzex, chzey) implied by the plan's material map and tables
  std::vector<double> planCoefficient(int which) const;
  // node coordinates of the uniform grid (grid_space.cpp:625-638)
  const std::vector<double> &eNodeCoord(int axis) const { return _ec[axis]; }

 private:
  double _cell_size;
  float _cfl;
  int _precision = 64;
  std::vector<std::shared_ptr<Object>> _objects;
  std::vector<std::shared_ptr<Boundary>> _boundaries;
  std::vector<std::shared_ptr<LumpedElement>> _lumped_elements;
  std::shared_ptr<TFSF> _tfsf;
  std::shared_ptr<NFFFT> _nffft;
  std::vector<Probe> _probes;
  std::vector<std::shared_ptr<Monitor>> _monitors;
  struct MonitorPlan {
    const Monitor *monitor;   // the TimeDomainFieldMonitor
    int comp;
    int64_t box[6];           // i0, ni, j0, nj, k0, nk
    size_t plot_step, total;  // movie schedule (plot_step 0: final slice only)
  };
  std::vector<MonitorPlan> _monitor_plans;
  struct LineTermPlan {
    int comp, group;
    int64_t a[3], b[3];
    double w;
  };
  struct LineMonitorPlan {
    Monitor *monitor;  // a VoltageMonitor or CurrentMonitor
    std::vector<LineTermPlan> terms;
    double sign;
  };
  std::vector<LineMonitorPlan> _line_plans;
  std::shared_ptr<Network> _network;
  void planLineMonitors();
  void planMonitors();
  void writeMonitor(const MonitorPlan &m, const std::string &file) const;
  void outputBroadBand(const NffftBroadBand &bb) const;
  void outputNffft() const;
  void writeLineMonitor(const LineMonitorPlan &lp) const;
  void outputNetwork() const;
  bool _has_design = false;
  std::array<size_t, 6> _design_box{};
  std::vector<float> _design_eps;
  void designCoefficients(double dt, std::vector<double> &coef, std::vector<double> &eps) const;
  bool _has_dft = false;
  int _dft_comp = 0, _dft_axis = 0;
  size_t _dft_plane[5] = {};  // index, lo1, hi1, lo2, hi2
  double _dft_freq = 0;

  // GridSpace; _two_d: TWO_DIMENSION, one z cell (grid/grid_space.cpp:501-536)
  double _min[3], _max[3];
  size_t _n[3] = {0, 0, 0};
  bool _two_d = false;
  std::vector<double> _ec[3], _hc[3], _es[3], _hs[3];
  double _dt = 0;
  size_t _total_time_steps = 0;
  size_t _nffft_box[6] = {};
  xf_engine *_engine = nullptr;
  std::shared_ptr<PlanData> _plan;

  void initGridSpace();
  void upload();
  void resizeSpaceSize(const double nmin[3], const double nmax[3]);
  size_t handleTransform(int axis, double x) const;
};

}  // namespace xfdtd
