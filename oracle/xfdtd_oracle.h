/*
 * xfdtd_oracle.h -- FP64 CPU restatement of the XFDTD reference's 3-D run loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * No product code links or calls it.
 *
 * The builder calls mirror the reference's Simulation API
 * (include/simulation/simulation.h:27-66): objects, boundaries (PML), one TFSF3D
 * source, one NffftFd, lumped current sources, then init + step.  Every
 * numerical routine in xfdtd_oracle.c cites the reference file:line it restates.
 */
#ifndef XFDTD_ORACLE_H
#define XFDTD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xo_sim xo_sim;

/* Orientation order of include/util/type_define.h:14 */
enum { XO_XN = 0, XO_XP, XO_YN, XO_YP, XO_ZN, XO_ZP };
/* field components */
enum { XO_EX = 0, XO_EY, XO_EZ, XO_HX, XO_HY, XO_HZ };

xo_sim *xo_sim_create(double cell_size, float cfl);
void xo_sim_destroy(xo_sim *s);
const char *xo_last_error(void);

/* Object(name, Cube(origin,size) | Sphere(center,radius), Material(eps_r,mu_r,sigma_e,sigma_m)) */
int xo_sim_add_cube(xo_sim *s, const double origin[3], const double size[3],
                    float eps_r, float mu_r, double sigma_e, double sigma_m);
int xo_sim_add_sphere(xo_sim *s, const double center[3], double radius,
                      float eps_r, float mu_r, double sigma_e, double sigma_m);
/* Object(name, Cylinder(axis 0 x / 1 y / 2 z, center, radius, height), Material) */
int xo_sim_add_cylinder(xo_sim *s, int axis, const double center[3], double radius, double height,
                        float eps_r, float mu_r, double sigma_e, double sigma_m);
/* make the last added object a DrudeMedium(eps_inf, omega_p, nv) with its
   sigma_e / sigma_m (its eps_r and mu_r become 1) */
int xo_sim_set_drude(xo_sim *s, double eps_inf, double omega_p, double nu);
/* make the last added object a DebyMedium(eps_s, eps_inf, tau) or a
   LorentzMedium(eps_s, eps_inf, omega_p, nv) with its sigma_e / sigma_m */
int xo_sim_set_debye(xo_sim *s, double eps_s, double eps_inf, double tau);
int xo_sim_set_lorentz(xo_sim *s, double eps_s, double eps_inf, double omega_p, double nu);
/* polarisation current J of component comp (0..2, E shapes) */
int xo_sim_get_drude_j(xo_sim *s, int comp, double *out);
int xo_sim_set_drude_j(xo_sim *s, int comp, const double *in);
/* PML(orientation, thickness, order, sigma_ratio, alpha_min, alpha_max, kappa_max) */
int xo_sim_add_pml(xo_sim *s, int orientation, int thickness, int order,
                   double sigma_ratio, double alpha_min, double alpha_max,
                   double kappa_max);
/* TFSF3D(dx,dy,dz,e0,theta,phi,psi, GaussianWaveform(amplitude,tau,t0)) */
int xo_sim_set_tfsf(xo_sim *s, const size_t distance[3], double e0,
                    double theta, double phi, double psi, double amplitude,
                    double tau, double t0);
/* TFSF2D(dx, dy, phi_inc, ez_i0, waveform) on a 2-D grid (nz = 1); waveform
   kind 0 Gaussian (amp, tau, t0), 1 sinusoid (amp, f, phase), 2
   cosine-modulated Gaussian (amp, tau, t0, f) in wp[4] */
int xo_sim_set_tfsf2d(xo_sim *s, size_t distance_x, size_t distance_y, double phi, double ez_i0, int wkind,
                      const double wp[4]);
/* NffftFd(dx,dy,dz, frequencies, theta, phi) */
int xo_sim_set_nffft(xo_sim *s, const size_t distance[3], const double *freqs,
                     int nf, const double *theta, int ntheta,
                     const double *phi, int nphi);
/* NffftBroadBand(dx,dy,dz, theta, phi): face-current series of every step */
int xo_sim_set_nffft_broadband(xo_sim *s, const size_t distance[3], double theta, double phi);
/* CurrentSource(Cube, orientation, resistance, GaussianWaveform(amp,tau,t0)) */
int xo_sim_add_current_source(xo_sim *s, const double origin[3],
                              const double size[3], int orientation,
                              double resistance, double amplitude, double tau,
                              double t0);
/* Lumped element (lumped_element/{current_source,voltage_source,resistor,capacitor,inductor}.cpp) in insertion order: kind 0
   CurrentSource, 1 VoltageSource (dir = orientation 0..5, value =
   resistance), 2 Resistor, 3 Capacitor, 4 Inductor (dir = axis 0..2, value =
   R, C or L).  Sources carry a waveform: wkind 0 Gaussian (amp, tau, t0),
   1 sinusoid (amp, frequency, phase), 2 cosine-modulated Gaussian (amp, tau,
   t0, frequency) in wp[0..3]. */
int xo_sim_add_lumped(xo_sim *s, int kind, const double origin[3], const double size[3], int dir,
                      double value, int wkind, const double wp[4]);
/* ObjectPlane(Cube(origin, size) with one zero size, material): sets the
   tangential sigma_e on the plane only (object/object_plane.cpp:8-50) */
int xo_sim_add_plane(xo_sim *s, const double origin[3], const double size[3], float eps_r, float mu_r,
                     double sigma_e, double sigma_m);
/* VoltageMonitor (kind 0) / CurrentMonitor (kind 1) over Cube(origin, size)
   with an orientation; returns the monitor's index */
int xo_sim_add_line_monitor(xo_sim *s, int kind, const double origin[3], const double size[3],
                            int orientation);
int xo_sim_get_line_monitors(const xo_sim *s, double *out /* nmon * total_steps */);
int xo_sim_get_inductor_current(const xo_sim *s, int q, double *out);
int xo_sim_get_lumped_wave(const xo_sim *s, int q, double *out /* total_steps */);
/* point probe sampled after every step (monitor update slot) */
int xo_sim_add_probe(xo_sim *s, int comp, size_t i, size_t j, size_t k);

int xo_sim_init(xo_sim *s, size_t total_steps);
/* run `n` steps of the time loop starting at the current step */
int xo_sim_step(xo_sim *s, size_t n);
/* run only the E half-step / H half-step of the current step (tests) */

int xo_sim_dims(const xo_sim *s, size_t out[3]);
double xo_sim_dt(const xo_sim *s);
size_t xo_sim_field_size(const xo_sim *s, int comp);
int xo_sim_get_field(const xo_sim *s, int comp, double *out);
int xo_sim_set_field(xo_sim *s, int comp, const double *in);
/* 18 coefficient arrays in the order cexe,cexhy,cexhz,ceye,ceyhz,ceyhx,
   ceze,cezhx,cezhy,chxh,chxey,chxez,chyh,chyez,chyex,chzh,chzex,chzey */
int xo_sim_get_coeff(const xo_sim *s, int which, double *out);
/* CPML profile of boundary `b` (insertion order): b_e,a_e,b_m,a_m,kappa_e,kappa_m */
int xo_sim_get_pml_profile(const xo_sim *s, int b, double *out);
/* TFSF: aux size and current e_inc/h_inc (sizes L and L-1) */
size_t xo_sim_tfsf_aux_size(const xo_sim *s);
int xo_sim_get_tfsf_aux(const xo_sim *s, double *e_inc, double *h_inc);
int xo_sim_get_waveform(const xo_sim *s, double *out /* total_steps */);
/* NF-FF accumulators, 24 currents in the order documented in xfdtd_oracle.c */
size_t xo_sim_nffft_size(const xo_sim *s);
int xo_sim_get_nffft(const xo_sim *s, double *out /* interleaved re,im */);
/* far field: out holds 8 complex arrays (n_theta, n_phi, l_theta, l_phi,
   e_theta, e_phi, h_theta, h_phi) + 2 real (power_theta, power_phi), each of
   nf*ntheta*nphi, packed as documented in xfdtd_oracle.c */
int xo_sim_farfield(xo_sim *s, double *out);
int xo_sim_get_probes(const xo_sim *s, double *out /* nprobe * total_steps */);
/* broadband NF-FF: the 2T-1 bin frequencies and wave numbers (either may be
   NULL); the six surface integrals n_x..l_z per bin (interleaved re,im; k
   NULL = the bins' own wave numbers); the face-current series [T][8][Q] */
int xo_sim_bb_frequencies(const xo_sim *s, double *freq, double *k);
int xo_sim_bb_sums(xo_sim *s, const double *k, double *out);
size_t xo_sim_bb_samples_size(const xo_sim *s);
int xo_sim_get_bb_samples(const xo_sim *s, double *out);
size_t xo_sim_current_step(const xo_sim *s);
int xo_set_threads(int n);

/* Adjoint (no reference counterpart; the checker of the GPU adjoint).
   Design box [i0,i1)x[j0,j1)x[k0,k1) with per-cell eps_r (float, C order),
   lossless, mu_r = 1, applied after the objects. */
int xo_sim_set_design(xo_sim *s, const size_t box[6], const float *eps_r);
/* DFT monitor of component comp on the plane axis = index, the two other
   axes in [lo1,hi1) x [lo2,hi2) (increasing axis order), sampled after every
   step with weight dt exp(-i 2 pi f t dt) */
int xo_sim_set_dft_plane(xo_sim *s, int comp, int axis, size_t index, size_t lo1, size_t hi1,
                         size_t lo2, size_t hi2, double freq);
int xo_sim_get_dft(const xo_sim *s, double *out /* interleaved complex */);
/* forward run from step 0 to total_steps, then the reverse run: J = sum |F|^2
   and dJ/d eps_r per design cell */
int xo_sim_grad(xo_sim *s, double *grad, double *J);
/* reverse-run nsteps of the field map on packed adjoint fields (in place) */
int xo_sim_adjoint_run(xo_sim *s, size_t nsteps, double *fields);

#ifdef __cplusplus
}
#endif
#endif
