/*
 * xfdtd_sim.h -- C-ABI over the C++ Simulation of include/xfdtd/xfdtd.hpp.
 *
 * The reference is a C++ library with no foreign-language binding; these
 * entry points are the plain-C form of its construction and run calls
 * (include/simulation/simulation.h:27-66), for callers that cannot use C++
 * (ctypes, cgo, JNI, the test harness):
 *
 *   xf_sim_create                Simulation(double cell_size, float cfl)   simulation.cpp:20-23
 *   xf_sim_add_cube / _sphere    addObject(Object(name, Cube|Sphere, Material))  simulation.cpp:115-117
 *   xf_sim_add_drude_cube / _sphere  addObject(Object(name, shape, DrudeMedium(...)))  dispersive_material.h:112-113
 *   xf_sim_add_pml               addBoundary(PML(...))                      simulation.cpp:119-121
 *   xf_sim_set_tfsf3d            addTFSFSource(TFSF3D(..., GaussianWaveform))  simulation.cpp:128-130
 *   xf_sim_add_cylinder          addObject(Object(name, Cylinder, Material))   shape/cylinder.cpp:11-15
 *   xf_sim_set_tfsf2d            addTFSFSource(TFSF2D(dx, dy, phi, ez_i0, waveform))  tfsf/tfsf_2d.cpp:10-24
 *   xf_sim_set_nffft_fd          addNFFFT(NffftFd(...))                     simulation.cpp:132-134
 *   xf_sim_add_current_source    addLumpedElement(CurrentSource(...))       simulation.cpp:123-126
 *   xf_sim_add_probe             addMonitor (point sample)                  simulation.cpp:136-138
 *   xf_sim_add_field_monitor     addMonitor(TimeDomainFieldMonitor(...))    field_monitor.cpp:15-121
 *   xf_sim_add_movie_monitor     addMonitor(MovieMonitor(TimeDomainFieldMonitor(...), T, plot_step))
 *                                                                           movie_monitor.cpp:6-59
 *   xf_sim_init / xf_sim_step    run(T) split into init + time loop         simulation_update.cpp:5-40
 *   xf_sim_get_field             getEx()..getHz()                           simulation.h:92-97
 *   xf_sim_farfield              NffftFd::calculateFarfield                 nffft_fd.cpp:379-421
 *   xf_sim_set_nffft_broadband   addNFFFT(NffftBroadBand(...))              nffft_broadband.cpp:21-28
 *   xf_sim_broadband_farfield    NffftBroadBand::calculateFarField          nffft_broadband.cpp:316-466
 *   xf_sim_broadband_ez          the 2-D far-zone Ez of outputData          nffft_broadband.cpp:171-179
 *   xf_sim_output_nffft          Simulation::outputData: NffftFd::outputData and the field
 *                                monitors' final slices                     simulation.cpp:144-156
 *
 * Every call returns 0 on success and -1 on failure; xf_sim_last_error()
 * then holds the message of the C++ exception the reference would throw.
 */
#ifndef XFDTD_SIM_H
#define XFDTD_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xf_sim xf_sim;

const char *xf_sim_last_error(void);
xf_sim *xf_sim_create(double cell_size, float cfl);
void xf_sim_destroy(xf_sim *s);

int xf_sim_add_cube(xf_sim *s, const double origin[3], const double size[3], float eps_r,
                    float mu_r, double sigma_e, double sigma_m);
int xf_sim_add_sphere(xf_sim *s, const double center[3], double radius, float eps_r, float mu_r,
                      double sigma_e, double sigma_m);
/* addObject(Object(name, Cube|Sphere, DrudeMedium(name, eps_inf, pole_omega,
   nv, sigma_e, sigma_m)))  (material/dispersive_material.h:112-113) */
int xf_sim_add_drude_cube(xf_sim *s, const double origin[3], const double size[3], double eps_inf,
                          double omega_p, double nv, double sigma_e, double sigma_m);
int xf_sim_add_drude_sphere(xf_sim *s, const double center[3], double radius, double eps_inf,
                            double omega_p, double nv, double sigma_e, double sigma_m);
/* addObject(Object(name, shape, medium)) for any dispersive medium:
   shape 0 Cube (p = origin[3], size[3]), 1 Sphere (p = center[3], radius),
   2 Cylinder (p = center[3], radius, height, axis); kind 1
   DrudeMedium (mp = eps_inf, omega_p, nv), 2 DebyMedium (eps_s, eps_inf,
   tau), 3 LorentzMedium (eps_s, eps_inf, omega_p, nv)
   (material/dispersive_material.h:59-141) */
int xf_sim_add_dispersive_object(xf_sim *s, int shape, const double p[6], int kind, const double mp[4],
                                 double sigma_e, double sigma_m);
int xf_sim_add_pml(xf_sim *s, int orientation, int thickness, int order, double sigma_ratio,
                   double alpha_min, double alpha_max, double kappa_max);
int xf_sim_set_tfsf3d(xf_sim *s, const size_t distance[3], double e0, double theta, double phi,
                      double psi, double amplitude, double tau, double t0);
/* Object(name, Cylinder(axis 0 x / 1 y / 2 z, center, radius, height), Material) */
int xf_sim_add_cylinder(xf_sim *s, int axis, const double center[3], double radius, double height,
                        float eps_r, float mu_r, double sigma_e, double sigma_m);
/* TFSF2D(distance_x, distance_y, phi_inc, ez_i0, waveform) for a 2-D scene
   (objects with no z extent: one z cell); waveform kind as xf_sim_add_lumped */
int xf_sim_set_tfsf2d(xf_sim *s, size_t distance_x, size_t distance_y, double phi_inc, double ez_i0,
                      int wkind, const double wp[4]);
int xf_sim_set_nffft_fd(xf_sim *s, const size_t distance[3], const double *freqs, int nf,
                        const double *theta, int ntheta, const double *phi, int nphi,
                        const char *output_dir);
/* NffftBroadBand(dx, dy, dz, theta, phi, output_dir): monostatic RCS path;
   replaces any NffftFd (the reference's Simulation holds one NFFFT) */
int xf_sim_set_nffft_broadband(xf_sim *s, const size_t distance[3], double theta, double phi,
                               const char *output_dir);
int xf_sim_add_current_source(xf_sim *s, const double origin[3], const double size[3],
                              int orientation, double resistance, double amplitude, double tau,
                              double t0);
/* lumped element, in insertion order (lumped_element/{current_source,voltage_source,resistor,capacitor,inductor}.cpp): kind 0
   CurrentSource, 1 VoltageSource (dir = orientation 0..5, value =
   resistance), 2 Resistor, 3 Capacitor, 4 Inductor (dir = axis 0..2, value
   = R, C or L).  Sources take a waveform: wkind 0 Gaussian (amplitude, tau,
   t0), 1 sinusoid (amplitude, frequency, phase), 2 cosine-modulated
   Gaussian (amplitude, tau, t0, frequency) in wp[0..3]. */
int xf_sim_add_lumped(xf_sim *s, int kind, const double origin[3], const double size[3], int dir,
                      double value, int wkind, const double wp[4]);
/* ObjectPlane(Cube(origin, size), material) (object/object_plane.cpp:4-50) */
int xf_sim_add_plane(xf_sim *s, const double origin[3], const double size[3], float eps_r, float mu_r,
                     double sigma_e, double sigma_m);
/* VoltageMonitor (kind 0) / CurrentMonitor (kind 1) (Cube(origin, size),
   orientation, dir, file): one value per step; xf_sim_output_nffft writes
   dir/file = [time, value].  Monitors are numbered in the order added. */
int xf_sim_add_line_monitor(xf_sim *s, int kind, const double origin[3], const double size[3],
                            int orientation, const char *dir, const char *file);
/* values [monitor][total_steps] of the steps run so far */
int xf_sim_get_line_monitors(xf_sim *s, double *out);
/* Network(ports, frequencies, output_dir) of Port(index[q], impedance
   (re, im), is_source, voltage monitor vmon[q], current monitor imon[q])
   (network/port.cpp, network/network.cpp) */
int xf_sim_set_network(xf_sim *s, int nports, const int *index, const double *impedance,
                       const int *is_source, const int *vmon, const int *imon, const double *freqs, int nf,
                       const char *output_dir);
/* number of S parameters (source ports x ports); keys 10 j + i and the
   values [n][nf] interleaved complex, in increasing key order */
int xf_sim_network_size(const xf_sim *s);
int xf_sim_network_sparams(xf_sim *s, int *keys, double *out);
int xf_sim_add_probe(xf_sim *s, int comp, size_t i, size_t j, size_t k);
/* TimeDomainFieldMonitor(Cube(origin, size), plane_type (0 XY, 1 YZ, 2 ZX),
   component, dir, file): the component over the cube's grid box, written by
   xf_sim_output_nffft (the end of run) as dir/file (.npy, float64) */
int xf_sim_add_field_monitor(xf_sim *s, const double origin[3], const double size[3], int plane_type, int comp,
                             const char *dir, const char *file);
/* MovieMonitor of such a monitor: dir/%04d.npy after every step t with
   t % plot_step == 0 or t == total_steps - 1, written while stepping */
int xf_sim_add_movie_monitor(xf_sim *s, const double origin[3], const double size[3], int plane_type, int comp,
                             const char *dir, size_t total_steps, size_t plot_step);

/* CPU set-up only (Simulation::init without the device): grid, materials,
   coefficient tables, CPML, TFSF and NF-FF tables.  Needs no GPU. */
int xf_sim_plan(xf_sim *s, size_t total_steps);
/* Simulation::checkRun (simulation.cpp:98-113): plan if not planned, then
   write "simulation_check" and "simulation_object_check" in the working
   directory */
int xf_sim_check_run(xf_sim *s, size_t total_steps);
/* per-cell coefficient array `which` (0..17: cexe, cexhy, cexhz, ceye, ceyhz,
   ceyhx, ceze, cezhx, cezhy, chxh, chxey, chxez, chyh, chyez, chyex, chzh,
   chzex, chzey; fdtd_basic_coff.h) implied by the plan, in the reference's
   array shape: the check that material index + tables reproduce the
   reference's 18 arrays */
int xf_sim_plan_coeff(xf_sim *s, int which, double *out);
/* precision: 64 (FP64 fields, parity with the reference) or 32 */
int xf_sim_init(xf_sim *s, size_t total_steps, int precision);
int xf_sim_step(xf_sim *s, size_t n);
int xf_sim_sync(xf_sim *s);
size_t xf_sim_current_step(const xf_sim *s);

int xf_sim_dims(const xf_sim *s, size_t out[3]);
double xf_sim_dt(const xf_sim *s);
int xf_sim_get_field(xf_sim *s, int comp, double *out);
int xf_sim_set_field(xf_sim *s, int comp, const double *in);
/* accumulators as xf_engine_get_nffft; returns their complex count */
int64_t xf_sim_nffft_size(const xf_sim *s);
int xf_sim_get_nffft(xf_sim *s, double *out);
/* out: 8 complex arrays (n_theta, n_phi, l_theta, l_phi, e_theta, e_phi,
   h_theta, h_phi) then 2 real arrays (power_theta, power_phi), each
   [nf][ntheta][nphi] */
int xf_sim_farfield(xf_sim *s, double *out);
/* broadband NF-FF: face-current series as xf_engine_get_broadband_samples,
   and the far field of the Ns = 2T-1 bins.  wave_number[Ns] overrides the
   bins' wave numbers in the surface integrals (NULL: the bins' own).  out:
   frequencies[Ns], wave_number[Ns], then 14 complex arrays [Ns] (n_x, n_y,
   n_z, l_x, l_y, l_z, n_theta, n_phi, l_theta, l_phi, e_theta, e_phi,
   h_theta, h_phi), then power_theta[Ns], power_phi[Ns]: 32 Ns doubles */
int64_t xf_sim_broadband_samples_size(const xf_sim *s);
int xf_sim_get_broadband_samples(xf_sim *s, double *out);
int xf_sim_broadband_farfield(xf_sim *s, const double *wave_number, double *out);
/* 2-D grids: the far-zone Ez of every bin, 0.5 sqrt(i k) (l_phi + eta0
   n_theta) (nffft_broadband.cpp:171-179), interleaved complex [Ns] */
int xf_sim_broadband_ez(xf_sim *s, double *out);
/* Simulation::outputData: NF-FF files (NffftFd or NffftBroadBand) and the
   field monitors' final slices */
int xf_sim_output_nffft(xf_sim *s);
int xf_sim_get_probes(xf_sim *s, double *out);
/* the engine behind the simulation (valid after xf_sim_init) */
struct xf_engine *xf_sim_engine(xf_sim *s);

/* Differentiation (no counterpart in the reference; see xfdtd_engine.h).
   Before xf_sim_init: a design box of per-cell eps_r (lossless, applied after
   the objects) and a DFT monitor plane whose sum |F|^2 is the objective. */
int xf_sim_set_design(xf_sim *s, const size_t box[6], const float *eps_r);
/* after xf_sim_init: new eps_r for the same design box, coefficients
   recomputed on the host and uploaded (xf_engine_update_design) */
int xf_sim_update_design(xf_sim *s, const float *eps_r);
int xf_sim_set_dft_plane(xf_sim *s, int comp, int axis, size_t index, size_t lo1, size_t hi1, size_t lo2,
                         size_t hi2, double freq);
int xf_sim_get_dft(xf_sim *s, double *out);
/* J and dJ/d eps_r per design cell, from step 0 (xf_engine_grad_eps) */
int xf_sim_grad_eps(xf_sim *s, int checkpoint_every, double *grad, double *J);
/* transpose of nsteps steps on n packed adjoint field values, in place */
int xf_sim_adjoint_run(xf_sim *s, size_t nsteps, double *fields, size_t n);
int xf_sim_checkpoint(xf_sim *s, int slot);
/* the Drude box and the Drude J of component comp (0..2) over it
   (xf_engine_drude_box, xf_engine_get_drude_j) */
int xf_sim_drude_box(xf_sim *s, int64_t box[6]);
int xf_sim_get_drude_j(xf_sim *s, int comp, double *out);
int xf_sim_set_drude_j(xf_sim *s, int comp, const double *in);
int xf_sim_restore(xf_sim *s, int slot);

#ifdef __cplusplus
}
#endif
#endif
