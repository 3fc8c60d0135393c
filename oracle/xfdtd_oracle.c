/*
 * xfdtd_oracle.c -- FP64 CPU restatement of the XFDTD reference's 3-D run loop.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP engine.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The reference (Mrwatermolen/XFDTD) cannot be built here: its xtensor / xtl /
 * xtensor-blas / xtensor-fftw / FFTW3 / BLAS / LAPACK dependencies are absent.
 * This file restates, loop for loop, the reference code on the run-loop path.
 * Paths below are relative to the reference's src/ directory.
 *
 *   grid sizing / coordinates   grid/grid_space.cpp:364-464, 501-549, 576-638, 737-757
 *   shape masks                 grid/grid_space.cpp:274-304, shape/sphere.cpp:27-30,
 *                               shape/cube.cpp:27-37, util/float_compare.h:9-24
 *   materials                   material/material.cpp:7-16, object/object.cpp:53-73
 *   dt and coefficients         fdtd_basic_coff/fdtd_basic_coff.cpp:11-122, 313-354
 *   E / H curl updates          simulation/simulation_update.cpp:50-108, 130-197
 *   run-loop order              simulation/simulation_update.cpp:5-40, simulation_init.cpp:10-22
 *   CPML                        boundary/perfect_match_layer.cpp:12-718
 *   TFSF 3-D                    tfsf/tfsf.cpp:15-82, tfsf/tfsf_3d.cpp:13-483
 *   Gaussian waveform           waveform/gaussian_waveform.cpp:18-24
 *   NF-FF frequency domain      nffft/nffft.cpp:22-34, nffft/nffft_fd.cpp:17-203, 379-556
 *   face-centre interpolation   electromagnetic_field/linear_interpolation_emf.cpp:286-421
 *   current source              lumped_element/current_source.cpp:5-187
 *   Drude ADE                   material/dispersive_material.cpp:118-193,
 *                               simulation/simulation_update.cpp:221-280
 *
 * Resolved ambiguities (the reference leaves these to uninitialised memory or
 * to an xtensor behaviour that no reference test pins):
 *   - TFSF e_inc/h_inc and the XP/YP CPML psi arrays are never initialised by
 *     the reference (xarray::resize); they are zero here.
 *   - NffftFd::calculateFarfield broadcasts the (nf,) wave-number array against
 *     (nf,ntheta,nphi) arrays, which in xtensor aligns trailing axes; the
 *     evident intent (one wave number per frequency, axis 0) is used here.
 *   - CurrentSource's Y branch never assigns _db (current_source.cpp:85-88);
 *     the evident intent (da = hz[k], db = hx[i], dc = ey[j]) is used.
 *   - xt::linspace samples are start + step*i with the last sample = stop.
 *   - the 3-vector dot of TFSF3D's transform (BLAS dgemv) sums left to right.
 *   - Dispersive objects: Object::defaultInit throws before the cells are
 *     tagged (object/object.cpp:109-115), so the reference never runs its
 *     Drude update.  The commented-out intent is used: the cells of a Drude
 *     object's shape mask (the mask every material uses) take the Drude E
 *     update, later Drude objects overwriting earlier ones; other objects
 *     change the standard coefficient arrays only (they never set the index).
 */
#include "xfdtd_oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* util/constant.h */
#define XO_PI 3.14159265358979323846
static const double EPS0 = 8.854187817e-12;
static double MU0(void) { return 4 * XO_PI * 1e-7; }
static double C0(void) { return sqrt(1 / (EPS0 * MU0())); }
static double ETA0(void) { return sqrt(MU0() / EPS0); }
#define TOL (10 * DBL_EPSILON)

static __thread char g_err[512];
const char *xo_last_error(void) { return g_err; }
#define FAIL(...)                                  \
  do {                                             \
    snprintf(g_err, sizeof g_err, __VA_ARGS__);    \
    return -1;                                     \
  } while (0)

/* util/float_compare.h:14-24 */
static int is_less_or_equal(double a, double b) {
  return (fabs(a - b) < TOL) ? 1 : (a < b);
}
static int is_greater_or_equal(double a, double b) {
  return (fabs(a - b) < TOL) ? 1 : (a > b);
}

typedef struct {
  double *d;
  size_t n0, n1, n2;
} a3;
#define AT(a, i, j, k) ((a).d[((size_t)(i) * (a).n1 + (size_t)(j)) * (a).n2 + (size_t)(k)])

static int a3_alloc(a3 *a, size_t n0, size_t n1, size_t n2, double fill) {
  a->n0 = n0;
  a->n1 = n1;
  a->n2 = n2;
  size_t n = n0 * n1 * n2;
  a->d = (double *)malloc((n ? n : 1) * sizeof(double));
  if (!a->d) return -1;
  for (size_t q = 0; q < n; ++q) a->d[q] = fill;
  return 0;
}
static void a3_free(a3 *a) {
  free(a->d);
  a->d = NULL;
}

typedef struct {
  int kind; /* 0 cube, 1 sphere, 2 cylinder (p: centre, r: radius, h: height) */
  int plane; /* ObjectPlane: a zero-thickness cube that sets sigma_e only */
  int axis;  /* cylinder axis: 0 x, 1 y, 2 z */
  double p[3], s[3], r, h;
  float eps_r, mu_r;
  double sigma_e, sigma_m;
  /* dispersive media (material/dispersive_material.h): drude = 1
     DrudeMedium(eps_inf, omega_p, nv), 2 DebyMedium(eps_s, eps_inf, tau),
     3 LorentzMedium(eps_s, eps_inf, omega_p, nv) */
  int drude;
  double eps_inf, omega_p, nu, eps_s, tau;
} obj_t;

typedef struct {
  int ori, d, order;
  double sigma_ratio, alpha_min, alpha_max, kappa_max;
  double dl;
  size_t start;
  double *be, *ae, *bm, *am, *ke, *km;
  /* psi and c_psi for the four components (ea, eb, ha, hb), each laid out
     [layer][t1][t2] over the full transverse extent of the component */
  double *psi[4], *cpsi[4];
  size_t tn1[4], tn2[4];
} pml_t;

typedef struct {
  size_t dist[3];
  double e0, theta, phi, psi, amp, tau, t0;
  /* the waveform (kinds of lumped_wave: 0 Gaussian, 1 sinusoid, 2
     cosine-modulated Gaussian, whose frequency is wp3); dim2: TFSF2D */
  int wkind, dim2;
  double wp3;
  size_t li, ri, lj, rj, lk, rk;
  double te[3], th[3], ratio, scaled_dl, ceihi, chiei, ceie, chih, a, b;
  size_t L;
  double *e_inc, *h_inc, *wave;
  double *px_full, *py_full, *pz_full, *px_half, *py_half, *pz_half;
} tfsf_t;

typedef struct {
  size_t dist[3];
  int nf, nt, np;
  double *freqs, *theta, *phi, *k;
  size_t li, ri, lj, rj, lk, rk, nxb, nyb, nzb;
  double complex *acc; /* 24 current arrays, see nf_off() */
  double complex *tj, *tm;
} nffft_t;

/* NffftBroadBand (nffft/nffft_broadband.cpp:21-91): the same surface and
   face currents as NffftFd, kept as time series */
typedef struct {
  size_t dist[3];
  double theta, phi;
  size_t li, ri, lj, rj, lk, rk, nxb, nyb, nzb, Q;
  double *samples; /* [T][8][Q]: step, current c of nf_off(), surface cell */
} bb_t;

/* lumped elements in insertion order (lumped_element/{current_source,voltage_source,resistor,capacitor,inductor}.cpp) */
enum { LE_CURRENT = 0, LE_VOLTAGE = 1, LE_RESISTOR = 2, LE_CAPACITOR = 3, LE_INDUCTOR = 4 };
/* waveforms: 0 Gaussian (amp, tau, t0), 1 sinusoid (amp, f, phase),
   2 cosine-modulated Gaussian (amp, tau, t0, f) */
typedef struct {
  double origin[3], size[3];
  int kind;
  int ori;   /* orientation (sources) or 2 * axis (R, C, L) */
  double R;  /* resistance, capacitance or inductance */
  int wkind;
  double amp, tau, t0, wp3;
  double rf;
  size_t is, ie, js, je, ks, ke;
  double *wave, *coff_i;
  double *jc, *cecjc, *cjcec; /* inductor: current and its two coefficients */
} csrc_t;

/* voltage (kind 0) and current (kind 1) monitors (monitor/voltage_monitor.cpp,
   monitor/current_monitor.cpp) */
typedef struct {
  int kind, ori;
  double origin[3], size[3];
  size_t is, ie, js, je, ks, ke;
  double *value; /* [T] */
} lmon_t;

typedef struct {
  int comp;
  size_t i, j, k;
} probe_t;

struct xo_sim {
  double dl;
  float cfl;
  /* GridSpace */
  double minc[3], maxc[3];
  size_t n[3];
  double *ec[3], *hc[3], *es[3], *hs[3];
  /* TWO_DIMENSION grid (grid/grid_space.cpp:501-536): nz == 1, the TMz
     fields Ez, Hx, Hy carry the solution and Ex, Ey, Hz stay zero */
  int dim2;
  double dt;
  size_t T, cur;
  int inited;
  obj_t *objs;
  int nobj;
  pml_t *pml;
  int npml;
  int has_tfsf;
  tfsf_t tf;
  int has_nf;
  nffft_t nf;
  int has_bb;
  bb_t bb;
  csrc_t *src;
  int nsrc;
  lmon_t *lmon;
  int nlmon;
  probe_t *probes;
  int nprobe;
  double *probe_buf;
  a3 f[6];    /* Ex Ey Ez Hx Hy Hz */
  a3 c[18];   /* coefficient arrays */
  /* per-cell permittivity of a design box (lossless, mu_r = 1), applied
     after the objects as one-cell objects would be */
  int has_design;
  size_t dbox[6];        /* i0,i1,j0,j1,k0,k1 */
  float *design_eps;     /* eps_r per design cell, C order */
  /* DFT monitor of one component on a plane (the adjoint objective) */
  int has_dft;
  int dft_comp, dft_axis;
  size_t dft_index, dft_lo1, dft_hi1, dft_lo2, dft_hi2;
  double dft_freq;
  double complex *dft_acc;  /* [hi1-lo1][hi2-lo2] */
  /* Drude cells: dmask[(i*ny + j)*nz + k] = 1 + object index (0: none);
     per-object coefficients and the polarisation currents J (E shapes) */
  int has_drude;
  int *dmask;
  double *dk, *dbeta, *dca, *dcb;  /* [nobj]; Lorentz: alpha, gamma, ca, cb */
  double *dxi, *dcc;               /* [nobj]; Lorentz only: xi, cc */
  int has_lorentz;
  a3 J[3];
  a3 Jp[3], Ep[3];  /* Lorentz: J and E of the step before (E shapes) */
};

/* coefficient indices */
enum {
  CEXE, CEXHY, CEXHZ, CEYE, CEYHZ, CEYHX, CEZE, CEZHX, CEZHY,
  CHXH, CHXEY, CHXEZ, CHYH, CHYEZ, CHYEX, CHZH, CHZEX, CHZEY
};

static void field_shape(const xo_sim *s, int comp, size_t sh[3]) {
  /* simulation/simulation_init.cpp:28-35 */
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  switch (comp) {
    case XO_EX: sh[0] = nx; sh[1] = ny + 1; sh[2] = nz + 1; break;
    case XO_EY: sh[0] = nx + 1; sh[1] = ny; sh[2] = nz + 1; break;
    case XO_EZ: sh[0] = nx + 1; sh[1] = ny + 1; sh[2] = nz; break;
    case XO_HX: sh[0] = nx + 1; sh[1] = ny; sh[2] = nz; break;
    case XO_HY: sh[0] = nx; sh[1] = ny + 1; sh[2] = nz; break;
    default: sh[0] = nx; sh[1] = ny; sh[2] = nz + 1; break;
  }
}

xo_sim *xo_sim_create(double cell_size, float cfl) {
  xo_sim *s = (xo_sim *)calloc(1, sizeof(xo_sim));
  if (!s) return NULL;
  s->dl = cell_size;
  s->cfl = cfl;
  /* grid/grid_space.cpp:14-27 */
  for (int a = 0; a < 3; ++a) {
    s->minc[a] = DBL_MAX;
    s->maxc[a] = DBL_MIN;
    s->n[a] = 0;
  }
  return s;
}

void xo_sim_destroy(xo_sim *s) {
  if (!s) return;
  for (int a = 0; a < 3; ++a) {
    free(s->ec[a]); free(s->hc[a]); free(s->es[a]); free(s->hs[a]);
  }
  for (int q = 0; q < 6; ++q) a3_free(&s->f[q]);
  for (int q = 0; q < 18; ++q) a3_free(&s->c[q]);
  for (int b = 0; b < s->npml; ++b) {
    pml_t *p = &s->pml[b];
    free(p->be); free(p->ae); free(p->bm); free(p->am); free(p->ke); free(p->km);
    for (int q = 0; q < 4; ++q) { free(p->psi[q]); free(p->cpsi[q]); }
  }
  free(s->pml);
  free(s->objs);
  tfsf_t *t = &s->tf;
  free(t->e_inc); free(t->h_inc); free(t->wave);
  free(t->px_full); free(t->py_full); free(t->pz_full);
  free(t->px_half); free(t->py_half); free(t->pz_half);
  nffft_t *nf = &s->nf;
  free(nf->freqs); free(nf->theta); free(nf->phi); free(nf->k);
  free(nf->acc); free(nf->tj); free(nf->tm);
  free(s->bb.samples);
  for (int q = 0; q < s->nsrc; ++q) {
    free(s->src[q].wave); free(s->src[q].coff_i);
    free(s->src[q].jc); free(s->src[q].cecjc); free(s->src[q].cjcec);
  }
  free(s->src);
  for (int q = 0; q < s->nlmon; ++q) free(s->lmon[q].value);
  free(s->lmon);
  free(s->probes);
  free(s->probe_buf);
  free(s->design_eps);
  free(s->dft_acc);
  free(s->dmask);
  free(s->dk); free(s->dbeta); free(s->dca); free(s->dcb);
  for (int q = 0; q < 3; ++q) {
    a3_free(&s->J[q]);
    a3_free(&s->Jp[q]);
    a3_free(&s->Ep[q]);
  }
  free(s->dxi);
  free(s->dcc);
  free(s);
}

/* ---------------------------------------------------------------- builder */

static int add_obj(xo_sim *s, obj_t o) {
  if (s->inited) FAIL("simulation already initialised");
  obj_t *n = (obj_t *)realloc(s->objs, (size_t)(s->nobj + 1) * sizeof(obj_t));
  if (!n) FAIL("out of memory");
  s->objs = n;
  s->objs[s->nobj++] = o;
  return 0;
}

int xo_sim_add_cube(xo_sim *s, const double origin[3], const double size[3],
                    float eps_r, float mu_r, double sigma_e, double sigma_m) {
  obj_t o;
  memset(&o, 0, sizeof o);
  o.kind = 0;
  for (int a = 0; a < 3; ++a) { o.p[a] = origin[a]; o.s[a] = size[a]; }
  o.eps_r = eps_r; o.mu_r = mu_r; o.sigma_e = sigma_e; o.sigma_m = sigma_m;
  return add_obj(s, o);
}

int xo_sim_set_drude(xo_sim *s, double eps_inf, double omega_p, double nu) {
  if (s->inited) FAIL("simulation already initialised");
  if (s->nobj == 0) FAIL("no object to make dispersive");
  obj_t *o = &s->objs[s->nobj - 1];
  /* LinearDispersiveMaterial(name, 1, 1, sigma_e, sigma_m)
     (material/dispersive_material.h:21-24, dispersive_material.cpp:118-124) */
  o->eps_r = 1.0f;
  o->mu_r = 1.0f;
  o->drude = 1;
  o->eps_inf = eps_inf;
  o->omega_p = omega_p;
  o->nu = nu;
  s->has_drude = 1;
  return 0;
}

/* DebyMedium(name, eps_s, eps_inf, tau, sigma_e, sigma_m) and
   LorentzMedium(name, eps_s, eps_inf, omega_p, nv, sigma_e, sigma_m)
   (material/dispersive_material.h:59-76, 137-141): the last added object
   becomes a LinearDispersiveMaterial of that kind */
int xo_sim_set_debye(xo_sim *s, double eps_s, double eps_inf, double tau) {
  if (s->inited) FAIL("simulation already initialised");
  if (s->nobj == 0) FAIL("no object to make dispersive");
  obj_t *o = &s->objs[s->nobj - 1];
  o->eps_r = 1.0f;
  o->mu_r = 1.0f;
  o->drude = 2;
  o->eps_s = eps_s;
  o->eps_inf = eps_inf;
  o->tau = tau;
  s->has_drude = 1;
  return 0;
}

int xo_sim_set_lorentz(xo_sim *s, double eps_s, double eps_inf, double omega_p, double nu) {
  if (s->inited) FAIL("simulation already initialised");
  if (s->nobj == 0) FAIL("no object to make dispersive");
  obj_t *o = &s->objs[s->nobj - 1];
  o->eps_r = 1.0f;
  o->mu_r = 1.0f;
  o->drude = 3;
  o->eps_s = eps_s;
  o->eps_inf = eps_inf;
  o->omega_p = omega_p;
  o->nu = nu;
  s->has_drude = 1;
  s->has_lorentz = 1;
  return 0;
}

int xo_sim_add_sphere(xo_sim *s, const double center[3], double radius,
                      float eps_r, float mu_r, double sigma_e, double sigma_m) {
  obj_t o;
  memset(&o, 0, sizeof o);
  o.kind = 1;
  for (int a = 0; a < 3; ++a) o.p[a] = center[a];
  o.r = radius;
  o.eps_r = eps_r; o.mu_r = mu_r; o.sigma_e = sigma_e; o.sigma_m = sigma_m;
  return add_obj(s, o);
}

/* shape/cylinder.cpp:11-15 */
int xo_sim_add_cylinder(xo_sim *s, int axis, const double center[3], double radius, double height,
                        float eps_r, float mu_r, double sigma_e, double sigma_m) {
  if (axis < 0 || axis > 2) FAIL("bad cylinder axis");
  obj_t o;
  memset(&o, 0, sizeof o);
  o.kind = 2;
  o.axis = axis;
  for (int a = 0; a < 3; ++a) o.p[a] = center[a];
  o.r = radius;
  o.h = height;
  o.eps_r = eps_r; o.mu_r = mu_r; o.sigma_e = sigma_e; o.sigma_m = sigma_m;
  return add_obj(s, o);
}

int xo_sim_add_pml(xo_sim *s, int orientation, int thickness, int order,
                   double sigma_ratio, double alpha_min, double alpha_max,
                   double kappa_max) {
  if (s->inited) FAIL("simulation already initialised");
  if (orientation < 0 || orientation > 5 || thickness <= 0) FAIL("bad PML");
  pml_t *n = (pml_t *)realloc(s->pml, (size_t)(s->npml + 1) * sizeof(pml_t));
  if (!n) FAIL("out of memory");
  s->pml = n;
  pml_t *p = &s->pml[s->npml++];
  memset(p, 0, sizeof *p);
  p->ori = orientation; p->d = thickness; p->order = order;
  p->sigma_ratio = sigma_ratio; p->alpha_min = alpha_min;
  p->alpha_max = alpha_max; p->kappa_max = kappa_max;
  return 0;
}

int xo_sim_set_tfsf(xo_sim *s, const size_t distance[3], double e0,
                    double theta, double phi, double psi, double amplitude,
                    double tau, double t0) {
  if (s->inited) FAIL("simulation already initialised");
  tfsf_t *t = &s->tf;
  memset(t, 0, sizeof *t);
  for (int a = 0; a < 3; ++a) t->dist[a] = distance[a];
  t->e0 = e0; t->theta = theta; t->phi = phi; t->psi = psi;
  t->amp = amplitude; t->tau = tau; t->t0 = t0;
  s->has_tfsf = 1;
  return 0;
}

/* TFSF2D(distance_x, distance_y, phi_inc, ez_i0, waveform)
   (tfsf/tfsf_2d.cpp:10-24): a TFSF with theta = pi/2, psi = 0 and no z
   faces.  wp = (amplitude, tau or frequency, t0 or phase, modulation
   frequency) of waveform kind wkind. */
int xo_sim_set_tfsf2d(xo_sim *s, size_t distance_x, size_t distance_y, double phi, double ez_i0, int wkind,
                      const double wp[4]) {
  if (s->inited) FAIL("simulation already initialised");
  if (wkind < 0 || wkind > 2) FAIL("bad waveform kind");
  tfsf_t *t = &s->tf;
  memset(t, 0, sizeof *t);
  t->dist[0] = distance_x; t->dist[1] = distance_y; t->dist[2] = 0;
  t->e0 = ez_i0; t->theta = XO_PI / 2; t->phi = phi; t->psi = 0;
  t->wkind = wkind;
  t->amp = wp[0]; t->tau = wp[1]; t->t0 = wp[2]; t->wp3 = wp[3];
  t->dim2 = 1;
  s->has_tfsf = 1;
  return 0;
}

int xo_sim_set_nffft(xo_sim *s, const size_t distance[3], const double *freqs,
                     int nf, const double *theta, int ntheta,
                     const double *phi, int nphi) {
  if (s->inited) FAIL("simulation already initialised");
  nffft_t *f = &s->nf;
  memset(f, 0, sizeof *f);
  for (int a = 0; a < 3; ++a) f->dist[a] = distance[a];
  f->nf = nf; f->nt = ntheta; f->np = nphi;
  f->freqs = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
  f->k = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
  f->theta = (double *)malloc(sizeof(double) * (size_t)(ntheta > 0 ? ntheta : 1));
  f->phi = (double *)malloc(sizeof(double) * (size_t)(nphi > 0 ? nphi : 1));
  for (int q = 0; q < nf; ++q) {
    f->freqs[q] = freqs[q];
    /* nffft/nffft_fd.h: _wave_number{2 * PI * _frequencies / C_0} */
    f->k[q] = 2 * XO_PI * freqs[q] / C0();
  }
  for (int q = 0; q < ntheta; ++q) f->theta[q] = theta[q];
  for (int q = 0; q < nphi; ++q) f->phi[q] = phi[q];
  s->has_nf = 1;
  return 0;
}

int xo_sim_add_lumped(xo_sim *s, int kind, const double origin[3], const double size[3], int dir,
                      double value, int wkind, const double wp[4]) {
  if (s->inited) FAIL("simulation already initialised");
  if (kind < LE_CURRENT || kind > LE_INDUCTOR) FAIL("bad lumped element kind");
  if (kind <= LE_VOLTAGE ? (dir < 0 || dir > 5) : (dir < 0 || dir > 2)) FAIL("bad orientation or axis");
  if (wkind < 0 || wkind > 2) FAIL("bad waveform");
  csrc_t *n = (csrc_t *)realloc(s->src, (size_t)(s->nsrc + 1) * sizeof(csrc_t));
  if (!n) FAIL("out of memory");
  s->src = n;
  csrc_t *c = &s->src[s->nsrc++];
  memset(c, 0, sizeof *c);
  for (int a = 0; a < 3; ++a) { c->origin[a] = origin[a]; c->size[a] = size[a]; }
  c->kind = kind;
  c->ori = kind <= LE_VOLTAGE ? dir : 2 * dir;
  /* current_source.cpp:12-15, voltage_source.cpp:19-22: R = 0 becomes 1e-20 */
  c->R = (kind <= LE_VOLTAGE && value == 0) ? 1e-20 : value;
  c->wkind = wkind;
  if (wp) { c->amp = wp[0]; c->tau = wp[1]; c->t0 = wp[2]; c->wp3 = wp[3]; }
  return 0;
}

int xo_sim_add_current_source(xo_sim *s, const double origin[3],
                              const double size[3], int orientation,
                              double resistance, double amplitude, double tau,
                              double t0) {
  const double wp[4] = {amplitude, tau, t0, 0.0};
  return xo_sim_add_lumped(s, LE_CURRENT, origin, size, orientation, resistance, 0, wp);
}

int xo_sim_add_plane(xo_sim *s, const double origin[3], const double size[3], float eps_r, float mu_r,
                     double sigma_e, double sigma_m) {
  if (xo_sim_add_cube(s, origin, size, eps_r, mu_r, sigma_e, sigma_m)) return -1;
  s->objs[s->nobj - 1].plane = 1;
  return 0;
}

int xo_sim_add_line_monitor(xo_sim *s, int kind, const double origin[3], const double size[3],
                            int orientation) {
  if (s->inited) FAIL("simulation already initialised");
  if (kind < 0 || kind > 1 || orientation < 0 || orientation > 5) FAIL("bad monitor");
  lmon_t *n = (lmon_t *)realloc(s->lmon, (size_t)(s->nlmon + 1) * sizeof(lmon_t));
  if (!n) FAIL("out of memory");
  s->lmon = n;
  lmon_t *m = &s->lmon[s->nlmon++];
  memset(m, 0, sizeof *m);
  m->kind = kind;
  m->ori = orientation;
  for (int a = 0; a < 3; ++a) { m->origin[a] = origin[a]; m->size[a] = size[a]; }
  return s->nlmon - 1;
}

int xo_sim_add_probe(xo_sim *s, int comp, size_t i, size_t j, size_t k) {
  if (s->inited) FAIL("simulation already initialised");
  probe_t *n = (probe_t *)realloc(s->probes, (size_t)(s->nprobe + 1) * sizeof(probe_t));
  if (!n) FAIL("out of memory");
  s->probes = n;
  s->probes[s->nprobe].comp = comp;
  s->probes[s->nprobe].i = i;
  s->probes[s->nprobe].j = j;
  s->probes[s->nprobe].k = k;
  s->nprobe++;
  return 0;
}

/* ------------------------------------------------------------- grid space */

/* grid/grid_space.cpp:576-618 */
static void resize_space(xo_sim *s, const double nmin[3], const double nmax[3]) {
  for (int a = 0; a < 3; ++a) {
    if (s->minc[a] > nmin[a]) s->minc[a] = nmin[a];
    if (s->maxc[a] < nmax[a]) s->maxc[a] = nmax[a];
  }
  for (int a = 0; a < 3; ++a) {
    size_t nn = (size_t)round((s->maxc[a] - s->minc[a]) / s->dl);
    if (s->maxc[a] < s->minc[a]) nn = 1;
    if (nn > s->n[a]) s->n[a] = nn;
  }
}

static void wrapped_box(const obj_t *o, double mn[3], double mx[3]) {
  if (o->kind == 0) {
    /* shape/cube.cpp:80-82; cube getXmax = point + size */
    for (int a = 0; a < 3; ++a) { mn[a] = o->p[a]; mx[a] = o->p[a] + o->s[a]; }
  } else if (o->kind == 2) {
    /* shape/cylinder.cpp:44-60: radius across, height/2 either side along
       the axis; the cube's max is point + size */
    for (int a = 0; a < 3; ++a) {
      double half = (a == o->axis) ? o->h / 2 : o->r, full = (a == o->axis) ? o->h : 2 * o->r;
      mn[a] = o->p[a] - half;
      mx[a] = mn[a] + full;
    }
  } else {
    /* shape/sphere.cpp:32-36 */
    for (int a = 0; a < 3; ++a) {
      mn[a] = o->p[a] - o->r;
      mx[a] = mn[a] + 2 * o->r;
    }
  }
}

static int is_inside(const obj_t *o, const double pt[3]) {
  if (o->kind == 0) {
    /* shape/cube.cpp:27-37 */
    return is_less_or_equal(o->p[0], pt[0]) && is_less_or_equal(o->p[1], pt[1]) &&
           is_less_or_equal(o->p[2], pt[2]) &&
           is_greater_or_equal(o->p[0] + o->s[0], pt[0]) &&
           is_greater_or_equal(o->p[1] + o->s[1], pt[1]) &&
           is_greater_or_equal(o->p[2] + o->s[2], pt[2]);
  }
  if (o->kind == 2) {
    /* shape/cylinder.cpp:21-42: inside the wrapped box, then within the
       radius of the axis */
    double mn[3], mx[3];
    wrapped_box(o, mn, mx);
    for (int a = 0; a < 3; ++a)
      if (!(is_less_or_equal(mn[a], pt[a]) && is_greater_or_equal(mx[a], pt[a]))) return 0;
    int u = o->axis == 0 ? 1 : 0, v = o->axis == 2 ? 1 : 2;
    double du = pt[u] - o->p[u], dv = pt[v] - o->p[v];
    return is_less_or_equal(sqrt(du * du + dv * dv), o->r);
  }
  /* shape/sphere.cpp:27-30 */
  double dx = pt[0] - o->p[0], dy = pt[1] - o->p[1], dz = pt[2] - o->p[2];
  double dist = sqrt(dx * dx + dy * dy + dz * dz);
  return is_less_or_equal(dist, o->r);
}

static int init_grid(xo_sim *s) {
  /* simulation/simulation_init.cpp:105-146 */
  for (int q = 0; q < s->nobj; ++q) {
    double mn[3], mx[3];
    wrapped_box(&s->objs[q], mn, mx);
    resize_space(s, mn, mx);
  }
  for (int b = 0; b < s->npml; ++b) {
    /* grid/grid_space.cpp:364-464 */
    int ori = s->pml[b].ori;
    int a = ori / 2;
    double size = s->pml[b].d * s->dl;
    if (s->n[a] <= 1) FAIL("GridSpace is not extended along axis %d", a);
    double mn[3] = {s->minc[0], s->minc[1], s->minc[2]};
    double mx[3] = {s->maxc[0], s->maxc[1], s->maxc[2]};
    if (ori % 2 == 0) mn[a] = s->minc[a] - size; else mx[a] = s->maxc[a] + size;
    resize_space(s, mn, mx);
  }
  /* grid/grid_space.cpp:501-536: THREE_DIMENSION when every axis has more
     than one cell, TWO_DIMENSION (the TMz branch) when only z has one */
  for (int a = 0; a < 3; ++a)
    if (s->n[a] < 1) s->n[a] = 1;
  s->dim2 = s->n[2] == 1;
  for (int a = 0; a < 2; ++a)
    if (s->n[a] <= 1)
      FAIL("only 3-D grids and 2-D grids with one z cell are supported (axis %d has %zu cells)", a, s->n[a]);
  /* grid/grid_space.cpp:625-638, 737-757 */
  for (int a = 0; a < 3; ++a) {
    size_t n = s->n[a];
    s->ec[a] = (double *)malloc(sizeof(double) * (n + 1));
    s->hc[a] = (double *)malloc(sizeof(double) * n);
    s->es[a] = (double *)malloc(sizeof(double) * n);
    s->hs[a] = (double *)malloc(sizeof(double) * (n + 1));
    for (size_t i = 0; i <= n; ++i) s->ec[a][i] = s->minc[a] + (double)i * s->dl;
    for (size_t i = 0; i < n; ++i) s->hc[a][i] = (s->ec[a][i] + s->ec[a][i + 1]) / 2;
    for (size_t i = 0; i < n; ++i) s->es[a][i] = s->ec[a][i + 1] - s->ec[a][i];
    /* virtual h nodes: [h0 - dl, h..., h_{n-1} + dl] */
    for (size_t i = 0; i <= n; ++i) {
      double lo = (i == 0) ? s->hc[a][0] - s->dl : s->hc[a][i - 1];
      double hi = (i == n) ? s->hc[a][n - 1] + s->dl : s->hc[a][i];
      s->hs[a][i] = hi - lo;
    }
  }
  return 0;
}

/* grid/grid_space.cpp:551-574: first minimum of |e_node - x| */
static int transform(const xo_sim *s, int a, double x, size_t *out) {
  if (x < s->minc[a]) FAIL("coordinate out of range");
  size_t best = 0;
  double bv = fabs(s->ec[a][0] - x);
  for (size_t i = 1; i <= s->n[a]; ++i) {
    double v = fabs(s->ec[a][i] - x);
    if (v < bv) { bv = v; best = i; }
  }
  *out = best;
  return 0;
}

/* ------------------------------------------------------------ coefficients */

/* fdtd_basic_coff/fdtd_basic_coff.cpp:77-91 */
static void calc_coff_e(a3 cece, a3 cecha, a3 cechb, a3 eps, a3 sig, double dt,
                        const double *da, int da_ax, const double *db, int db_ax) {
  for (size_t i = 0; i < eps.n0; ++i)
    for (size_t j = 0; j < eps.n1; ++j)
      for (size_t k = 0; k < eps.n2; ++k) {
        size_t idx[3] = {i, j, k};
        double e = AT(eps, i, j, k), sg = AT(sig, i, j, k);
        double a = da[idx[da_ax]], b = db[idx[db_ax]];
        AT(cece, i, j, k) = (2 * e - dt * sg) / (2 * e + dt * sg);
        AT(cecha, i, j, k) = -(2 * dt / b) / (2 * e + sg * dt);
        AT(cechb, i, j, k) = (2 * dt / a) / (2 * e + sg * dt);
      }
}

static void calc_coff_h(a3 chch, a3 chcea, a3 chceb, a3 mu, a3 sm, double dt,
                        const double *da, int da_ax, const double *db, int db_ax) {
  for (size_t i = 0; i < mu.n0; ++i)
    for (size_t j = 0; j < mu.n1; ++j)
      for (size_t k = 0; k < mu.n2; ++k) {
        size_t idx[3] = {i, j, k};
        double m = AT(mu, i, j, k), sg = AT(sm, i, j, k);
        double a = da[idx[da_ax]], b = db[idx[db_ax]];
        AT(chch, i, j, k) = (2 * m - dt * sg) / (2 * m + dt * sg);
        AT(chcea, i, j, k) = (2 * dt / b) / (2 * m + sg * dt);
        AT(chceb, i, j, k) = -(2 * dt / a) / (2 * m + sg * dt);
      }
}

/* ----------------------------------------------------------------- CPML */

/* which field arrays a face acts on: perfect_match_layer.cpp:98-126
   X: (ea,eb,ha,hb) = (Ey,Ez,Hy,Hz); Y: (Ez,Ex,Hz,Hx); Z: (Ex,Ey,Hx,Hy) */
static const int PML_COMP[3][4] = {{XO_EY, XO_EZ, XO_HY, XO_HZ},
                                   {XO_EZ, XO_EX, XO_HZ, XO_HX},
                                   {XO_EX, XO_EY, XO_HX, XO_HY}};
/* coefficient arrays (ceahb, cebha, chaeb, chbea): perfect_match_layer.cpp:62-95 */
static const int PML_COEF[3][4] = {{CEYHZ, CEZHY, CHYEZ, CHZEY},
                                   {CEZHX, CEXHZ, CHZEX, CHXEZ},
                                   {CEXHY, CEYHX, CHXEY, CHYEX}};

/* element (layer n along axis a, transverse t1,t2) of a field array */
static double *fa_at(a3 *f, int a, size_t n, size_t t1, size_t t2) {
  size_t i, j, k;
  if (a == 0) { i = n; j = t1; k = t2; }
  else if (a == 1) { i = t1; j = n; k = t2; }
  else { i = t1; j = t2; k = n; }
  return &AT(*f, i, j, k);
}
static void transverse_extent(const a3 *f, int a, size_t *t1, size_t *t2) {
  if (a == 0) { *t1 = f->n1; *t2 = f->n2; }
  else if (a == 1) { *t1 = f->n0; *t2 = f->n2; }
  else { *t1 = f->n0; *t2 = f->n1; }
}

static int init_pml(xo_sim *s, pml_t *p) {
  /* perfect_match_layer.cpp:49-96 */
  int a = p->ori / 2;
  int positive = p->ori % 2;
  p->dl = s->dl;
  p->start = positive ? s->n[a] - (size_t)p->d : 0;
  int d = p->d;
  p->be = (double *)malloc(sizeof(double) * (size_t)d);
  p->ae = (double *)malloc(sizeof(double) * (size_t)d);
  p->bm = (double *)malloc(sizeof(double) * (size_t)d);
  p->am = (double *)malloc(sizeof(double) * (size_t)d);
  p->ke = (double *)malloc(sizeof(double) * (size_t)d);
  p->km = (double *)malloc(sizeof(double) * (size_t)d);
  /* perfect_match_layer.cpp:133-155 (P) and 302-325 (N) */
  double e2m = MU0() / EPS0;
  double sigma_max = p->sigma_ratio * (p->order + 1) / (150.0 * XO_PI * p->dl);
  double dt = s->dt;
  for (int i = 0; i < d; ++i) {
    double rho_e, rho_m;
    if (positive) {
      rho_e = (i + 1 - 0.75) / (double)d;
      rho_m = (i + 1 - 0.25) / (double)d;
    } else {
      rho_e = (d - i - 0.75) / (double)d;
      rho_m = (d - i - 0.25) / (double)d;
    }
    double sigma_e = sigma_max * pow(rho_e, p->order);
    double sigma_m = sigma_max * pow(rho_m, p->order);
    sigma_m = sigma_m * e2m;
    double kappa_e = 1 + (p->kappa_max - 1) * pow(rho_e, p->order);
    double kappa_m = 1 + (p->kappa_max - 1) * pow(rho_m, p->order);
    double alpha_e = p->alpha_min + (p->alpha_max - p->alpha_min) * (1 - rho_e);
    double alpha_m = p->alpha_min + (p->alpha_max - p->alpha_min) * (1 - rho_m);
    alpha_m = alpha_m * e2m;
    p->be[i] = exp((-dt / EPS0) * (sigma_e / kappa_e + alpha_e));
    p->ae[i] = (1 / p->dl) * (p->be[i] - 1) * sigma_e /
               (kappa_e * (sigma_e + kappa_e * alpha_e));
    p->bm[i] = exp((-dt / MU0()) * (sigma_m / kappa_m + alpha_m));
    p->am[i] = (1 / p->dl) * (p->bm[i] - 1) * sigma_m /
               (kappa_m * (sigma_m + kappa_m * alpha_m));
    p->ke[i] = kappa_e;
    p->km[i] = kappa_m;
  }
  /* psi arrays (zero) and c_psi capture + kappa division:
     perfect_match_layer.cpp:157-299 (P) and 327-431 (N).  The E arrays sit at
     start+l+1 on N faces and start+l on P faces; H arrays at start+l. */
  for (int q = 0; q < 4; ++q) {
    a3 *f = &s->f[PML_COMP[a][q]];
    a3 *cf = &s->c[PML_COEF[a][q]];
    size_t t1, t2;
    transverse_extent(f, a, &t1, &t2);
    p->tn1[q] = t1;
    p->tn2[q] = t2;
    size_t sz = (size_t)d * t1 * t2;
    p->psi[q] = (double *)calloc(sz ? sz : 1, sizeof(double));
    p->cpsi[q] = (double *)calloc(sz ? sz : 1, sizeof(double));
    if (!p->psi[q] || !p->cpsi[q]) FAIL("out of memory");
    int is_e = q < 2;
    for (int l = 0; l < d; ++l) {
      size_t n = p->start + (size_t)l + ((is_e && !positive) ? 1 : 0);
      double kap = is_e ? p->ke[l] : p->km[l];
      for (size_t u = 0; u < t1; ++u)
        for (size_t v = 0; v < t2; ++v) {
          double *cc = fa_at(cf, a, n, u, v);
          p->cpsi[q][((size_t)l * t1 + u) * t2 + v] = *cc * p->dl;
          *cc = *cc / kap;
        }
    }
  }
  return 0;
}

/* perfect_match_layer.cpp:574-718 */
static void pml_update_e(xo_sim *s, pml_t *p) {
  int a = p->ori / 2;
  int positive = p->ori % 2;
  /* ea uses hb, eb uses ha */
  for (int q = 0; q < 2; ++q) {
    a3 *e = &s->f[PML_COMP[a][q]];
    a3 *h = &s->f[PML_COMP[a][q == 0 ? 3 : 2]];
    size_t t1 = p->tn1[q], t2 = p->tn2[q];
    for (int l = 0; l < p->d; ++l) {
      size_t n = p->start + (size_t)l + (positive ? 0 : 1);
      for (size_t u = 0; u < t1; ++u)
        for (size_t v = 0; v < t2; ++v) {
          size_t pi = ((size_t)l * t1 + u) * t2 + v;
          double *ps = &p->psi[q][pi];
          *ps = p->be[l] * *ps + p->ae[l] * (*fa_at(h, a, n, u, v) - *fa_at(h, a, n - 1, u, v));
          double *ev = fa_at(e, a, n, u, v);
          *ev = *ev + p->cpsi[q][pi] * *ps;
        }
    }
  }
}

/* perfect_match_layer.cpp:434-572 */
static void pml_update_h(xo_sim *s, pml_t *p) {
  int a = p->ori / 2;
  /* ha uses eb, hb uses ea */
  for (int q = 3; q >= 2; --q) {
    a3 *h = &s->f[PML_COMP[a][q]];
    a3 *e = &s->f[PML_COMP[a][q == 3 ? 0 : 1]];
    size_t t1 = p->tn1[q], t2 = p->tn2[q];
    for (int l = 0; l < p->d; ++l) {
      size_t n = p->start + (size_t)l;
      for (size_t u = 0; u < t1; ++u)
        for (size_t v = 0; v < t2; ++v) {
          size_t pi = ((size_t)l * t1 + u) * t2 + v;
          double *ps = &p->psi[q][pi];
          *ps = p->bm[l] * *ps + p->am[l] * (*fa_at(e, a, n + 1, u, v) - *fa_at(e, a, n, u, v));
          double *hv = fa_at(h, a, n, u, v);
          *hv = *hv + p->cpsi[q][pi] * *ps;
        }
    }
  }
}

/* ----------------------------------------------------------------- TFSF */

static double gaussian(double amp, double tau, double t0, double t) {
  /* waveform/gaussian_waveform.cpp:22-24 */
  return amp * exp(-pow((t - t0) / tau, 2));
}

/* a waveform's value at time t: kind 0 Gaussian (amp, tau, t0), 1 sinusoid
   (amp, f = tau, phase = t0), 2 cosine-modulated Gaussian (amp, tau, t0,
   f = p3) */
static double wave_at(int kind, double amp, double tau, double t0, double p3, double t) {
  if (kind == 1) /* waveform/sinusoidal_waveform.cpp:18-21 */
    return amp * sin(2 * XO_PI * tau * t + t0);
  if (kind == 2) /* waveform/cosine_modulated_gaussian_waveform.cpp:29-33 */
    return amp * cos(2 * XO_PI * p3 * (t - t0)) * exp(-pow((t - t0) / tau, 2));
  return gaussian(amp, tau, t0, t);
}

static int init_tfsf(xo_sim *s) {
  tfsf_t *t = &s->tf;
  /* tfsf/tfsf.cpp:15-33, 35-53 */
  double st = sin(t->theta), ct = cos(t->theta), sp = sin(t->phi), cp = cos(t->phi);
  double sps = sin(t->psi), cps = cos(t->psi);
  double kv[3] = {st * cp, st * sp, ct};
  /* tfsf/tfsf_3d.cpp:18-30 */
  double u[3][3] = {{-sp, ct * cp, st * cp}, {cp, ct * sp, st * sp}, {0, -st, ct}};
  double ke[3] = {sps, cps, 0};
  for (int r = 0; r < 3; ++r) t->te[r] = u[r][0] * ke[0] + u[r][1] * ke[1] + u[r][2] * ke[2];
  t->th[0] = kv[1] * t->te[2] - kv[2] * t->te[1];
  t->th[1] = kv[2] * t->te[0] - kv[0] * t->te[2];
  t->th[2] = kv[0] * t->te[1] - kv[1] * t->te[0];
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  if (t->dim2 ? (2 * t->dist[0] >= nx || 2 * t->dist[1] >= ny || t->dist[0] < 1 || t->dist[1] < 1)
              : (2 * t->dist[0] >= nx || 2 * t->dist[1] >= ny || 2 * t->dist[2] >= nz ||
                 t->dist[0] < 1 || t->dist[1] < 1 || t->dist[2] < 1))
    FAIL("TFSF box does not fit the grid");
  t->li = t->dist[0]; t->ri = nx - t->dist[0];
  t->lj = t->dist[1]; t->rj = ny - t->dist[1];
  t->lk = t->dist[2]; t->rk = nz - t->dist[2];
  size_t bx = t->ri - t->li, by = t->rj - t->lj, bz = t->rk - t->lk;
  t->wave = (double *)malloc(sizeof(double) * (s->T ? s->T : 1));
  for (size_t n = 0; n < s->T; ++n)
    t->wave[n] = wave_at(t->wkind, t->amp, t->tau, t->t0, t->wp3, (double)n * s->dt);
  /* tfsf/tfsf_3d.cpp:33-151; tfsf/tfsf_2d.cpp:26-97, whose diagonal takes
     sqrt(nz^2) in place of nz^2 */
  t->ratio = 1 / sqrt(pow(sin(t->theta), 4) * (pow(cos(t->phi), 4) + pow(sin(t->phi), 4)) +
                      pow(cos(t->theta), 4));
  double diag = t->dim2 ? sqrt(pow((double)bx, 2) + pow((double)by, 2) + sqrt(pow((double)bz, 2))) * t->ratio
                        : sqrt(pow((double)bx, 2) + pow((double)by, 2) + pow((double)bz, 2)) * t->ratio;
  t->L = (size_t)ceil(diag) + 4 + 1;
  t->e_inc = (double *)calloc(t->L, sizeof(double));
  t->h_inc = (double *)calloc(t->L - 1, sizeof(double));
  double extra[3] = {2 * kv[0] / t->ratio, 2 * kv[1] / t->ratio, 2 * kv[2] / t->ratio};
  double org[3];
  org[0] = (kv[0] >= 0) ? (double)t->li : (double)t->ri;
  org[1] = (kv[1] >= 0) ? (double)t->lj : (double)t->rj;
  org[2] = (kv[2] >= 0) ? (double)t->lk : (double)t->rk;
  if (t->dim2) org[2] = 0; /* tfsf_2d.cpp:60-77: the origin's z is 0 */
  for (int a = 0; a < 3; ++a) org[a] -= extra[a];
  size_t bn[3] = {bx, by, bz}, st3[3] = {t->li, t->lj, t->lk};
  double *full[3], *half[3];
  for (int a = 0; a < 3; ++a) {
    full[a] = (double *)malloc(sizeof(double) * (bn[a] + 1));
    half[a] = (double *)malloc(sizeof(double) * (bn[a] + 2));
    for (size_t i = 0; i < bn[a] + 1; ++i)
      full[a][i] = ((double)(i + st3[a]) - org[a]) * kv[a] * t->ratio;
    for (size_t i = 0; i < bn[a] + 2; ++i)
      half[a][i] = ((double)(i + st3[a]) - org[a] - 0.5) * kv[a] * t->ratio;
  }
  t->px_full = full[0]; t->py_full = full[1]; t->pz_full = full[2];
  t->px_half = half[0]; t->py_half = half[1]; t->pz_half = half[2];
  t->scaled_dl = s->dl / t->ratio;
  t->ceihi = -(s->dt / (EPS0 * t->scaled_dl));
  t->chiei = -(s->dt / (MU0() * t->scaled_dl));
  t->ceie = 1.0;
  t->chih = 1.0;
  t->a = 0;
  t->b = 0;
  return 0;
}

/* tfsf/tfsf_3d.cpp:153-183 */
static void tfsf_update_incident(xo_sim *s, size_t step) {
  tfsf_t *t = &s->tf;
  double dt = s->dt;
  size_t L = t->L;
  double *e = t->e_inc, *h = t->h_inc;
  e[0] = t->e0 * t->wave[step];
  double x = e[L - 2], y = e[L - 1];
  for (size_t i = 1; i < L - 1; ++i) e[i] = t->ceie * e[i] + t->ceihi * (h[i] - h[i - 1]);
  double c1 = (C0() * dt - t->scaled_dl) / (C0() * dt + t->scaled_dl);
  double c2 = 2 * t->scaled_dl / (C0() * dt + t->scaled_dl);
  e[L - 1] = -t->a + c1 * (e[L - 2] + t->b) + c2 * (x + y);
  t->a = x;
  t->b = y;
  for (size_t i = 0; i < L - 1; ++i) h[i] = t->chih * h[i] + t->chiei * (e[i + 1] - e[i]);
}

/* tfsf/tfsf_3d.cpp:413-483: comp 0..5 = Ex..Hz incident value at (i,j,k) */
static double tfsf_inc(const tfsf_t *t, int comp, size_t i, size_t j, size_t k) {
  double proj;
  const double *arr;
  double scale;
  switch (comp) {
    case XO_EX:
      proj = t->px_half[i - t->li + 1] + t->py_full[j - t->lj] + t->pz_full[k - t->lk];
      arr = t->e_inc; scale = t->te[0]; break;
    case XO_EY:
      proj = t->px_full[i - t->li] + t->py_half[j - t->lj + 1] + t->pz_full[k - t->lk];
      arr = t->e_inc; scale = t->te[1]; break;
    case XO_EZ:
      proj = t->px_full[i - t->li] + t->py_full[j - t->lj] + t->pz_half[k - t->lk + 1];
      arr = t->e_inc; scale = t->te[2]; break;
    case XO_HX:
      proj = t->px_full[i - t->li] + t->py_half[j - t->lj + 1] + t->pz_half[k - t->lk + 1] - 0.5;
      arr = t->h_inc; scale = t->th[0]; break;
    case XO_HY:
      proj = t->px_half[i - t->li + 1] + t->py_full[j - t->lj] + t->pz_half[k - t->lk + 1] - 0.5;
      arr = t->h_inc; scale = t->th[1]; break;
    default:
      proj = t->px_half[i - t->li + 1] + t->py_half[j - t->lj + 1] + t->pz_full[k - t->lk] - 0.5;
      arr = t->h_inc; scale = t->th[2]; break;
  }
  size_t idx = (size_t)proj;
  double w = proj - (double)idx;
  return (1 - w) * (scale * arr[idx]) + w * (scale * arr[idx + 1]);
}

/* TFSF2D::getIncident{Ez,Hx,Hy} (tfsf_2d.cpp:230-257): the 3-D
   interpolation without the z projections */
static double tfsf_inc2d(const tfsf_t *t, int comp, size_t i, size_t j) {
  double proj;
  const double *arr;
  double scale;
  if (comp == XO_EZ) {
    proj = t->px_full[i - t->li] + t->py_full[j - t->lj];
    arr = t->e_inc; scale = t->te[2];
  } else if (comp == XO_HX) {
    proj = t->px_full[i - t->li] + t->py_half[j - t->lj + 1] - 0.5;
    arr = t->h_inc; scale = t->th[0];
  } else {
    proj = t->px_half[i - t->li + 1] + t->py_full[j - t->lj] - 0.5;
    arr = t->h_inc; scale = t->th[1];
  }
  size_t idx = (size_t)proj;
  double w = proj - (double)idx;
  return (1 - w) * (scale * arr[idx]) + w * (scale * arr[idx + 1]);
}

/* TFSF2D::updateH (tfsf_2d.cpp:148-186): Hy on the x faces, Hx on the y
   faces, the incident Ez sampled on the total-field side */
static void tfsf_update_h2d(xo_sim *s) {
  tfsf_t *t = &s->tf;
  a3 *hx = &s->f[XO_HX], *hy = &s->f[XO_HY];
  size_t li = t->li, ri = t->ri, lj = t->lj, rj = t->rj;
  double cbx = s->dt / (MU0() * s->dl), cby = cbx;
  for (size_t j = lj; j < rj + 1; ++j) AT(*hy, li - 1, j, 0) -= cbx * tfsf_inc2d(t, XO_EZ, li, j);
  for (size_t j = lj; j < rj + 1; ++j) AT(*hy, ri, j, 0) += cbx * tfsf_inc2d(t, XO_EZ, ri, j);
  for (size_t i = li; i < ri + 1; ++i) AT(*hx, i, lj - 1, 0) += cby * tfsf_inc2d(t, XO_EZ, i, lj);
  for (size_t i = li; i < ri + 1; ++i) AT(*hx, i, rj, 0) -= cby * tfsf_inc2d(t, XO_EZ, i, rj);
}

/* TFSF2D::updateE (tfsf_2d.cpp:188-228): Ez on all four faces, in xn, xp,
   yn, yp order (the corners take two corrections) */
static void tfsf_update_e2d(xo_sim *s) {
  tfsf_t *t = &s->tf;
  a3 *ez = &s->f[XO_EZ];
  size_t li = t->li, ri = t->ri, lj = t->lj, rj = t->rj;
  double cax = s->dt / (EPS0 * s->dl), cay = cax;
  for (size_t j = lj; j < rj + 1; ++j) AT(*ez, li, j, 0) -= cax * tfsf_inc2d(t, XO_HY, li - 1, j);
  for (size_t j = lj; j < rj + 1; ++j) AT(*ez, ri, j, 0) += cax * tfsf_inc2d(t, XO_HY, ri, j);
  for (size_t i = li; i < ri + 1; ++i) AT(*ez, i, lj, 0) += cay * tfsf_inc2d(t, XO_HX, i, lj - 1);
  for (size_t i = li; i < ri + 1; ++i) AT(*ez, i, rj, 0) -= cay * tfsf_inc2d(t, XO_HX, i, rj);
}

/* tfsf/tfsf_3d.cpp:185-297 */
static void tfsf_update_h(xo_sim *s) {
  tfsf_t *t = &s->tf;
  if (t->dim2) {
    tfsf_update_h2d(s);
    return;
  }
  a3 *hx = &s->f[XO_HX], *hy = &s->f[XO_HY], *hz = &s->f[XO_HZ];
  size_t li = t->li, ri = t->ri, lj = t->lj, rj = t->rj, lk = t->lk, rk = t->rk;
  double cbx = s->dt / (MU0() * s->dl), cby = cbx, cbz = cbx;
  for (size_t j = lj; j < rj; ++j)
    for (size_t k = lk; k < rk + 1; ++k) AT(*hz, li - 1, j, k) += cbx * tfsf_inc(t, XO_EY, li, j, k);
  for (size_t j = lj; j < rj + 1; ++j)
    for (size_t k = lk; k < rk; ++k) AT(*hy, li - 1, j, k) -= cbx * tfsf_inc(t, XO_EZ, li, j, k);
  for (size_t j = lj; j < rj; ++j)
    for (size_t k = lk; k < rk + 1; ++k) AT(*hz, ri, j, k) -= cbx * tfsf_inc(t, XO_EY, ri, j, k);
  for (size_t j = lj; j < rj + 1; ++j)
    for (size_t k = lk; k < rk; ++k) AT(*hy, ri, j, k) += cbx * tfsf_inc(t, XO_EZ, ri, j, k);
  for (size_t k = lk; k < rk; ++k)
    for (size_t i = li; i < ri + 1; ++i) AT(*hx, i, lj - 1, k) += cby * tfsf_inc(t, XO_EZ, i, lj, k);
  for (size_t k = lk; k < rk + 1; ++k)
    for (size_t i = li; i < ri; ++i) AT(*hz, i, lj - 1, k) -= cby * tfsf_inc(t, XO_EX, i, lj, k);
  for (size_t k = lk; k < rk; ++k)
    for (size_t i = li; i < ri + 1; ++i) AT(*hx, i, rj, k) -= cby * tfsf_inc(t, XO_EZ, i, rj, k);
  for (size_t k = lk; k < rk + 1; ++k)
    for (size_t i = li; i < ri; ++i) AT(*hz, i, rj, k) += cby * tfsf_inc(t, XO_EX, i, rj, k);
  for (size_t i = li; i < ri; ++i)
    for (size_t j = lj; j < rj + 1; ++j) AT(*hy, i, j, lk - 1) += cbz * tfsf_inc(t, XO_EX, i, j, lk);
  for (size_t i = li; i < ri + 1; ++i)
    for (size_t j = lj; j < rj; ++j) AT(*hx, i, j, lk - 1) -= cbz * tfsf_inc(t, XO_EY, i, j, lk);
  for (size_t i = li; i < ri; ++i)
    for (size_t j = lj; j < rj + 1; ++j) AT(*hy, i, j, rk) -= cbz * tfsf_inc(t, XO_EX, i, j, rk);
  for (size_t i = li; i < ri + 1; ++i)
    for (size_t j = lj; j < rj; ++j) AT(*hx, i, j, rk) += cbz * tfsf_inc(t, XO_EY, i, j, rk);
}

/* tfsf/tfsf_3d.cpp:299-411 */
static void tfsf_update_e(xo_sim *s) {
  tfsf_t *t = &s->tf;
  if (t->dim2) {
    tfsf_update_e2d(s);
    return;
  }
  a3 *ex = &s->f[XO_EX], *ey = &s->f[XO_EY], *ez = &s->f[XO_EZ];
  size_t li = t->li, ri = t->ri, lj = t->lj, rj = t->rj, lk = t->lk, rk = t->rk;
  double cax = s->dt / (EPS0 * s->dl), cay = cax, caz = cax;
  for (size_t j = lj; j < rj; ++j)
    for (size_t k = lk; k < rk + 1; ++k) AT(*ey, li, j, k) += cax * tfsf_inc(t, XO_HZ, li - 1, j, k);
  for (size_t j = lj; j < rj + 1; ++j)
    for (size_t k = lk; k < rk; ++k) AT(*ez, li, j, k) -= cax * tfsf_inc(t, XO_HY, li - 1, j, k);
  for (size_t j = lj; j < rj; ++j)
    for (size_t k = lk; k < rk + 1; ++k) AT(*ey, ri, j, k) -= cax * tfsf_inc(t, XO_HZ, ri, j, k);
  for (size_t j = lj; j < rj + 1; ++j)
    for (size_t k = lk; k < rk; ++k) AT(*ez, ri, j, k) += cax * tfsf_inc(t, XO_HY, ri, j, k);
  for (size_t k = lk; k < rk; ++k)
    for (size_t i = li; i < ri + 1; ++i) AT(*ez, i, lj, k) += cay * tfsf_inc(t, XO_HX, i, lj - 1, k);
  for (size_t k = lk; k < rk + 1; ++k)
    for (size_t i = li; i < ri; ++i) AT(*ex, i, lj, k) -= cay * tfsf_inc(t, XO_HZ, i, lj - 1, k);
  for (size_t k = lk; k < rk; ++k)
    for (size_t i = li; i < ri + 1; ++i) AT(*ez, i, rj, k) -= cay * tfsf_inc(t, XO_HX, i, rj, k);
  for (size_t k = lk; k < rk + 1; ++k)
    for (size_t i = li; i < ri; ++i) AT(*ex, i, rj, k) += cay * tfsf_inc(t, XO_HZ, i, rj, k);
  for (size_t i = li; i < ri; ++i)
    for (size_t j = lj; j < rj + 1; ++j) AT(*ex, i, j, lk) += caz * tfsf_inc(t, XO_HY, i, j, lk - 1);
  for (size_t i = li; i < ri + 1; ++i)
    for (size_t j = lj; j < rj; ++j) AT(*ey, i, j, lk) -= caz * tfsf_inc(t, XO_HX, i, j, lk - 1);
  for (size_t i = li; i < ri; ++i)
    for (size_t j = lj; j < rj + 1; ++j) AT(*ex, i, j, rk) -= caz * tfsf_inc(t, XO_HY, i, j, rk);
  for (size_t i = li; i < ri + 1; ++i)
    for (size_t j = lj; j < rj; ++j) AT(*ey, i, j, rk) += caz * tfsf_inc(t, XO_HX, i, j, rk);
}

/* ---------------------------------------------------------------- NF-FF */

/* accumulator offset of current c (0..7) of face pair a (0 x, 1 y, 2 z).
   x pair: my_xn, mz_xn, jy_xn, jz_xn, my_xp, mz_xp, jy_xp, jz_xp  (nf, nyb, nzb)
   y pair: mz_yn, mx_yn, jz_yn, jx_yn, mz_yp, mx_yp, jz_yp, jx_yp  (nf, nzb, nxb)
   z pair: mx_zn, my_zn, jx_zn, jy_zn, mx_zp, my_zp, jx_zp, jy_zp  (nf, nxb, nyb) */
static size_t nf_face_cells(const nffft_t *f, int a) {
  if (a == 0) return f->nyb * f->nzb;
  if (a == 1) return f->nzb * f->nxb;
  return f->nxb * f->nyb;
}
static size_t nf_off(const nffft_t *f, int a, int c) {
  size_t off = 0;
  for (int b = 0; b < a; ++b) off += 8 * (size_t)f->nf * nf_face_cells(f, b);
  return off + (size_t)c * (size_t)f->nf * nf_face_cells(f, a);
}

size_t xo_sim_nffft_size(const xo_sim *s) {
  if (!s->has_nf || !s->inited) return 0;
  return nf_off(&s->nf, 3, 0);
}

int xo_sim_set_nffft_broadband(xo_sim *s, const size_t distance[3], double theta, double phi) {
  if (s->inited) FAIL("simulation already initialised");
  bb_t *b = &s->bb;
  free(b->samples);
  memset(b, 0, sizeof *b);
  for (int a = 0; a < 3; ++a) b->dist[a] = distance[a];
  b->theta = theta;
  b->phi = phi;
  s->has_bb = 1;
  return 0;
}

static int init_bb(xo_sim *s) {
  bb_t *b = &s->bb;
  /* nffft/nffft.cpp:22-34 */
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  if (s->dim2 ? (2 * b->dist[0] >= nx || 2 * b->dist[1] >= ny || b->dist[0] < 1 || b->dist[1] < 1 ||
                 b->dist[2] != 0)
              : (2 * b->dist[0] >= nx || 2 * b->dist[1] >= ny || 2 * b->dist[2] >= nz ||
                 b->dist[0] < 1 || b->dist[1] < 1 || b->dist[2] < 1))
    FAIL("NF-FF box does not fit the grid");
  b->li = b->dist[0]; b->ri = nx - b->dist[0];
  b->lj = b->dist[1]; b->rj = ny - b->dist[1];
  b->lk = b->dist[2]; b->rk = nz - b->dist[2];
  b->nxb = b->ri - b->li; b->nyb = b->rj - b->lj; b->nzb = b->rk - b->lk;
  /* on a 2-D grid the z faces are not kept: nffft_broadband.cpp:438-440
     never integrates them (and their H stencils would reach k = -1) */
  b->Q = b->nyb * b->nzb + b->nzb * b->nxb + (s->dim2 ? 0 : b->nxb * b->nyb);
  /* nffft_broadband.cpp:36-91 sizes every series 2T-1; the steps past the
     run's last are never written and are taken as zero */
  b->samples = (double *)calloc((s->T ? s->T : 1) * 8 * b->Q, sizeof(double));
  if (!b->samples) FAIL("out of memory");
  return 0;
}

static int init_nffft(xo_sim *s) {
  nffft_t *f = &s->nf;
  /* nffft/nffft.cpp:22-34 */
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  if (2 * f->dist[0] >= nx || 2 * f->dist[1] >= ny || 2 * f->dist[2] >= nz ||
      f->dist[0] < 1 || f->dist[1] < 1 || f->dist[2] < 1)
    FAIL("NF-FF box does not fit the grid");
  f->li = f->dist[0]; f->ri = nx - f->dist[0];
  f->lj = f->dist[1]; f->rj = ny - f->dist[1];
  f->lk = f->dist[2]; f->rk = nz - f->dist[2];
  f->nxb = f->ri - f->li; f->nyb = f->rj - f->lj; f->nzb = f->rk - f->lk;
  size_t tot = nf_off(f, 3, 0);
  f->acc = (double complex *)calloc(tot ? tot : 1, sizeof(double complex));
  /* nffft/nffft_fd.cpp:91-106 */
  size_t T = s->T;
  f->tj = (double complex *)malloc(sizeof(double complex) * (size_t)f->nf * (T ? T : 1));
  f->tm = (double complex *)malloc(sizeof(double complex) * (size_t)f->nf * (T ? T : 1));
  if (!f->acc || !f->tj || !f->tm) FAIL("out of memory");
  double dt = s->dt;
  const double complex mi2pi = ((-0.0 - 1.0 * I) * 2.0) * XO_PI;
  for (int n = 0; n < f->nf; ++n)
    for (size_t t = 0; t < T; ++t) {
      double argj = f->freqs[n] * ((double)t + 0.5) * dt;
      double argm = f->freqs[n] * (double)t * dt;
      double complex zj = CMPLX(creal(mi2pi) * argj, cimag(mi2pi) * argj);
      double complex zm = CMPLX(creal(mi2pi) * argm, cimag(mi2pi) * argm);
      double complex ej = cexp(zj), em = cexp(zm);
      f->tj[(size_t)n * T + t] = CMPLX(dt * creal(ej), dt * cimag(ej));
      f->tm[(size_t)n * T + t] = CMPLX(dt * creal(em), dt * cimag(em));
    }
  return 0;
}

#define EX(i, j, k) AT(s->f[XO_EX], i, j, k)
#define EY(i, j, k) AT(s->f[XO_EY], i, j, k)
#define EZ(i, j, k) AT(s->f[XO_EZ], i, j, k)
#define HX(i, j, k) AT(s->f[XO_HX], i, j, k)
#define HY(i, j, k) AT(s->f[XO_HY], i, j, k)
#define HZ(i, j, k) AT(s->f[XO_HZ], i, j, k)

static void acc_add(double complex *a, double v, double complex tw) {
  *a = CMPLX(creal(*a) + v * creal(tw), cimag(*a) + v * cimag(tw));
}

/* The eight face-centre currents of one surface cell, with the signs of
   nffft/nffft_fd.cpp:108-203 (and nffft_broadband.cpp:94-155, which has the
   same), and the stencils of
   electromagnetic_field/linear_interpolation_emf.cpp:286-388.
   pair 0 (x faces li, ri): cell (a, b) = (j, k)
   pair 1 (y faces lj, rj): cell (a, b) = (k, i)
   pair 2 (z faces lk, rk): cell (a, b) = (i, j) */
static void nf_face_values(const xo_sim *s, int pair, size_t lo, size_t hi, size_t a, size_t b,
                           double v[8]) {
  if (pair == 0) {
    size_t li = lo, ri = hi, j = a, k = b;
    v[0] = -1.0 * (0.5 * (EZ(li, j, k) + EZ(li, j + 1, k)));
    v[1] = 0.5 * (EY(li, j, k) + EY(li, j, k + 1));
    v[2] = 0.25 * (HZ(li, j, k) + HZ(li, j, k + 1) + HZ(li - 1, j, k) + HZ(li - 1, j, k + 1));
    v[3] = -1.0 * (0.25 * (HY(li, j, k) + HY(li, j + 1, k) + HY(li - 1, j, k) + HY(li - 1, j + 1, k)));
    v[4] = 0.5 * (EZ(ri, j, k) + EZ(ri, j + 1, k));
    v[5] = -1.0 * (0.5 * (EY(ri, j, k) + EY(ri, j, k + 1)));
    v[6] = -1.0 * (0.25 * (HZ(ri, j, k) + HZ(ri, j, k + 1) + HZ(ri - 1, j, k) + HZ(ri - 1, j, k + 1)));
    v[7] = 0.25 * (HY(ri, j, k) + HY(ri, j + 1, k) + HY(ri - 1, j, k) + HY(ri - 1, j + 1, k));
  } else if (pair == 1) {
    size_t lj = lo, rj = hi, k = a, i = b;
    v[0] = -1.0 * (0.5 * (EX(i, lj, k) + EX(i, lj + 1, k)));
    v[1] = 0.5 * (EZ(i, lj, k) + EZ(i + 1, lj, k));
    v[2] = 0.25 * (HX(i + 1, lj, k) + HX(i, lj, k) + HX(i + 1, lj - 1, k) + HX(i, lj - 1, k));
    v[3] = -1.0 * (0.25 * (HZ(i, lj, k) + HZ(i, lj, k + 1) + HZ(i, lj - 1, k) + HZ(i, lj - 1, k + 1)));
    v[4] = 0.5 * (EX(i, rj, k) + EX(i, rj + 1, k));
    v[5] = -1.0 * (0.5 * (EZ(i, rj, k) + EZ(i + 1, rj, k)));
    v[6] = -1.0 * (0.25 * (HX(i + 1, rj, k) + HX(i, rj, k) + HX(i + 1, rj - 1, k) + HX(i, rj - 1, k)));
    v[7] = 0.25 * (HZ(i, rj, k) + HZ(i, rj, k + 1) + HZ(i, rj - 1, k) + HZ(i, rj - 1, k + 1));
  } else {
    size_t lk = lo, rk = hi, i = a, j = b;
    v[0] = -1.0 * (0.5 * (EY(i, j, lk) + EY(i + 1, j, lk)));
    v[1] = 0.5 * (EX(i, j, lk) + EX(i, j, lk + 1));
    v[2] = 0.25 * (HY(i, j, lk) + HY(i, j + 1, lk) + HY(i, j, lk - 1) + HY(i, j + 1, lk - 1));
    v[3] = -1.0 * (0.25 * (HX(i + 1, j, lk) + HX(i, j, lk) + HX(i + 1, j, lk - 1) + HX(i, j, lk - 1)));
    v[4] = 0.5 * (EY(i, j, rk) + EY(i + 1, j, rk));
    v[5] = -1.0 * (0.5 * (EX(i, j, rk) + EX(i, j, rk + 1)));
    v[6] = -1.0 * (0.25 * (HY(i, j, rk) + HY(i, j + 1, rk) + HY(i, j, rk - 1) + HY(i, j + 1, rk - 1)));
    v[7] = 0.25 * (HX(i + 1, j, rk) + HX(i, j, rk) + HX(i + 1, j, rk - 1) + HX(i, j, rk - 1));
  }
}

/* nffft/nffft_fd.cpp:108-203: the running DFT of the face currents */
static void nffft_update(xo_sim *s) {
  nffft_t *f = &s->nf;
  size_t li = f->li, ri = f->ri, lj = f->lj, rj = f->rj, lk = f->lk, rk = f->rk;
  size_t nt = s->cur, T = s->T;
  int NF = f->nf;
  size_t cx = nf_face_cells(f, 0), cy = nf_face_cells(f, 1), cz = nf_face_cells(f, 2);
  for (size_t j = lj; j < rj; ++j)
    for (size_t k = lk; k < rk; ++k) {
      double v[8];
      nf_face_values(s, 0, li, ri, j, k, v);
      size_t cell = (j - lj) * f->nzb + (k - lk);
      for (int n = 0; n < NF; ++n)
        for (int c = 0; c < 8; ++c) {
          double complex tw = ((c % 4) < 2) ? f->tm[(size_t)n * T + nt] : f->tj[(size_t)n * T + nt];
          acc_add(&f->acc[nf_off(f, 0, c) + (size_t)n * cx + cell], v[c], tw);
        }
    }
  for (size_t k = lk; k < rk; ++k)
    for (size_t i = li; i < ri; ++i) {
      double v[8];
      nf_face_values(s, 1, lj, rj, k, i, v);
      size_t cell = (k - lk) * f->nxb + (i - li);
      for (int n = 0; n < NF; ++n)
        for (int c = 0; c < 8; ++c) {
          double complex tw = ((c % 4) < 2) ? f->tm[(size_t)n * T + nt] : f->tj[(size_t)n * T + nt];
          acc_add(&f->acc[nf_off(f, 1, c) + (size_t)n * cy + cell], v[c], tw);
        }
    }
  for (size_t i = li; i < ri; ++i)
    for (size_t j = lj; j < rj; ++j) {
      double v[8];
      nf_face_values(s, 2, lk, rk, i, j, v);
      size_t cell = (i - li) * f->nyb + (j - lj);
      for (int n = 0; n < NF; ++n)
        for (int c = 0; c < 8; ++c) {
          double complex tw = ((c % 4) < 2) ? f->tm[(size_t)n * T + nt] : f->tj[(size_t)n * T + nt];
          acc_add(&f->acc[nf_off(f, 2, c) + (size_t)n * cz + cell], v[c], tw);
        }
    }
}

/* nffft_broadband.cpp:94-155: the step's face currents into the series.
   Surface cell q runs over the x faces (j, k), then the y faces (k, i),
   then the z faces (i, j), the last index fastest. */
static void bb_update(xo_sim *s) {
  bb_t *b = &s->bb;
  double *row = b->samples + s->cur * 8 * b->Q;
  size_t q = 0;
  for (size_t j = b->lj; j < b->rj; ++j)
    for (size_t k = b->lk; k < b->rk; ++k, ++q) {
      double v[8];
      nf_face_values(s, 0, b->li, b->ri, j, k, v);
      for (int c = 0; c < 8; ++c) row[(size_t)c * b->Q + q] = v[c];
    }
  for (size_t k = b->lk; k < b->rk; ++k)
    for (size_t i = b->li; i < b->ri; ++i, ++q) {
      double v[8];
      nf_face_values(s, 1, b->lj, b->rj, k, i, v);
      for (int c = 0; c < 8; ++c) row[(size_t)c * b->Q + q] = v[c];
    }
  if (s->dim2) return;
  for (size_t i = b->li; i < b->ri; ++i)
    for (size_t j = b->lj; j < b->rj; ++j, ++q) {
      double v[8];
      nf_face_values(s, 2, b->lk, b->rk, i, j, v);
      for (int c = 0; c < 8; ++c) row[(size_t)c * b->Q + q] = v[c];
    }
}

/* complex product as evaluated for finite operands (a*c-b*d, a*d+b*c) */
static double complex cmul(double complex x, double complex y) {
  double a = creal(x), b = cimag(x), c = creal(y), d = cimag(y);
  return CMPLX(a * c - b * d, a * d + b * c);
}
static double complex cadd(double complex x, double complex y) {
  return CMPLX(creal(x) + creal(y), cimag(x) + cimag(y));
}
static double complex cscale(double s, double complex x) {
  return CMPLX(creal(x) * s, cimag(x) * s);
}
static double complex cexpi(double d) {
  /* std::exp(1i * d) with a zero real part */
  return cexp(CMPLX(0.0 * d, 1.0 * d));
}

/* nffft/nffft_fd.cpp:379-556.  out layout: 8 complex arrays
   (a_theta, a_phi, f_theta, f_phi, e_theta, e_phi, h_theta, h_phi) then
   2 real arrays (power_theta, power_phi), each nf*ntheta*nphi, C order. */
int xo_sim_farfield(xo_sim *s, double *out) {
  if (!s->inited || !s->has_nf) FAIL("no NF-FF surface");
  nffft_t *f = &s->nf;
  size_t NF = (size_t)f->nf, NT = (size_t)f->nt, NP = (size_t)f->np, M = NF * NT * NP;
  double complex *res = (double complex *)calloc(8 * M, sizeof(double complex));
  double *pw = out + 16 * M;
  double dl = s->dl;
  double eta = ETA0();
  for (size_t n = 0; n < NF; ++n)
    for (size_t t = 0; t < NT; ++t)
      for (size_t p = 0; p < NP; ++p) {
        double cos_t = cos(f->theta[t]), sin_t = sin(f->theta[t]);
        double cos_p = cos(f->phi[p]), sin_p = sin(f->phi[p]);
        double stcp = sin_t * cos_p, stsp = sin_t * sin_p;
        double complex ax = 0, ay = 0, az = 0, fx = 0, fy = 0, fz = 0;
        double dk = dl * f->k[n];
        double area = dl * dl;
        {
          double complex sjy = 0, sjz = 0, smy = 0, smz = 0;
          size_t cx = nf_face_cells(f, 0);
          for (size_t j = f->lj; j < f->rj; ++j)
            for (size_t k = f->lk; k < f->rk; ++k) {
              double dn = dk * ((double)f->li * stcp + ((double)j + 0.5) * stsp + ((double)k + 0.5) * cos_t);
              double dp = dk * ((double)f->ri * stcp + ((double)j + 0.5) * stsp + ((double)k + 0.5) * cos_t);
              double complex pn = cexpi(dn), pp = cexpi(dp);
              size_t cell = n * cx + (j - f->lj) * f->nzb + (k - f->lk);
              double complex *A = f->acc;
              sjy = cadd(sjy, cadd(cmul(A[nf_off(f, 0, 2) + cell], pn), cmul(A[nf_off(f, 0, 6) + cell], pp)));
              sjz = cadd(sjz, cadd(cmul(A[nf_off(f, 0, 3) + cell], pn), cmul(A[nf_off(f, 0, 7) + cell], pp)));
              smy = cadd(smy, cadd(cmul(A[nf_off(f, 0, 0) + cell], pn), cmul(A[nf_off(f, 0, 4) + cell], pp)));
              smz = cadd(smz, cadd(cmul(A[nf_off(f, 0, 1) + cell], pn), cmul(A[nf_off(f, 0, 5) + cell], pp)));
            }
          ay = cadd(ay, cscale(area, sjy)); az = cadd(az, cscale(area, sjz));
          fy = cadd(fy, cscale(area, smy)); fz = cadd(fz, cscale(area, smz));
        }
        {
          double complex sjz = 0, sjx = 0, smz = 0, smx = 0;
          size_t cy = nf_face_cells(f, 1);
          for (size_t k = f->lk; k < f->rk; ++k)
            for (size_t i = f->li; i < f->ri; ++i) {
              double dn = dk * (((double)i + 0.5) * stcp + (double)f->lj * stsp + ((double)k + 0.5) * cos_t);
              double dp = dk * (((double)i + 0.5) * stcp + (double)f->rj * stsp + ((double)k + 0.5) * cos_t);
              double complex pn = cexpi(dn), pp = cexpi(dp);
              size_t cell = n * cy + (k - f->lk) * f->nxb + (i - f->li);
              double complex *A = f->acc;
              sjz = cadd(sjz, cadd(cmul(A[nf_off(f, 1, 2) + cell], pn), cmul(A[nf_off(f, 1, 6) + cell], pp)));
              sjx = cadd(sjx, cadd(cmul(A[nf_off(f, 1, 3) + cell], pn), cmul(A[nf_off(f, 1, 7) + cell], pp)));
              smz = cadd(smz, cadd(cmul(A[nf_off(f, 1, 0) + cell], pn), cmul(A[nf_off(f, 1, 4) + cell], pp)));
              smx = cadd(smx, cadd(cmul(A[nf_off(f, 1, 1) + cell], pn), cmul(A[nf_off(f, 1, 5) + cell], pp)));
            }
          az = cadd(az, cscale(area, sjz)); ax = cadd(ax, cscale(area, sjx));
          fz = cadd(fz, cscale(area, smz)); fx = cadd(fx, cscale(area, smx));
        }
        {
          double complex sjx = 0, sjy = 0, smx = 0, smy = 0;
          size_t cz = nf_face_cells(f, 2);
          for (size_t j = f->lj; j < f->rj; ++j)
            for (size_t i = f->li; i < f->ri; ++i) {
              double dn = dk * (((double)i + 0.5) * stcp + ((double)j + 0.5) * stsp + (double)f->lk * cos_t);
              double dp = dk * (((double)i + 0.5) * stcp + ((double)j + 0.5) * stsp + (double)f->rk * cos_t);
              double complex pn = cexpi(dn), pp = cexpi(dp);
              size_t cell = n * cz + (i - f->li) * f->nyb + (j - f->lj);
              double complex *A = f->acc;
              sjx = cadd(sjx, cadd(cmul(A[nf_off(f, 2, 2) + cell], pn), cmul(A[nf_off(f, 2, 6) + cell], pp)));
              sjy = cadd(sjy, cadd(cmul(A[nf_off(f, 2, 3) + cell], pn), cmul(A[nf_off(f, 2, 7) + cell], pp)));
              smx = cadd(smx, cadd(cmul(A[nf_off(f, 2, 0) + cell], pn), cmul(A[nf_off(f, 2, 4) + cell], pp)));
              smy = cadd(smy, cadd(cmul(A[nf_off(f, 2, 1) + cell], pn), cmul(A[nf_off(f, 2, 5) + cell], pp)));
            }
          ax = cadd(ax, cscale(area, sjx)); ay = cadd(ay, cscale(area, sjy));
          fx = cadd(fx, cscale(area, smx)); fy = cadd(fy, cscale(area, smy));
        }
        double ctcp = cos_t * cos_p, ctsp = cos_t * sin_p;
        /* cos_t_cos_p * a_x + cos_t_sin_p * a_y - sin_t * a_z */
        double complex a_phi = cadd(cscale(-sin_p, ax), cscale(cos_p, ay));
        double complex f_phi = cadd(cscale(-sin_p, fx), cscale(cos_p, fy));
        double complex a_theta = CMPLX(creal(cscale(ctcp, ax)) + creal(cscale(ctsp, ay)) - sin_t * creal(az),
                        cimag(cscale(ctcp, ax)) + cimag(cscale(ctsp, ay)) - sin_t * cimag(az));
        double complex f_theta = CMPLX(creal(cscale(ctcp, fx)) + creal(cscale(ctsp, fy)) - sin_t * creal(fz),
                        cimag(cscale(ctcp, fx)) + cimag(cscale(ctsp, fy)) - sin_t * cimag(fz));
        size_t m = (n * NT + t) * NP + p;
        res[0 * M + m] = a_theta;
        res[1 * M + m] = a_phi;
        res[2 * M + m] = f_theta;
        res[3 * M + m] = f_phi;
        double kk = f->k[n];
        double four_pi = 4 * XO_PI;
        double complex mik = CMPLX(-0.0 * kk, -1.0 * kk), pik = CMPLX(0.0 * kk, 1.0 * kk);
        double complex z1 = cadd(cscale(eta, a_theta), f_phi);
        double complex z2 = cadd(cscale(-eta, a_phi), f_theta);
        double complex z3 = CMPLX(creal(a_phi) - creal(f_theta) / eta, cimag(a_phi) - cimag(f_theta) / eta);
        double complex z4 = CMPLX(creal(a_theta) + creal(f_phi) / eta, cimag(a_theta) + cimag(f_phi) / eta);
        double complex w;
        w = cmul(mik, z1); res[4 * M + m] = CMPLX(creal(w) / four_pi, cimag(w) / four_pi);
        w = cmul(pik, z2); res[5 * M + m] = CMPLX(creal(w) / four_pi, cimag(w) / four_pi);
        w = cmul(pik, z3); res[6 * M + m] = CMPLX(creal(w) / four_pi, cimag(w) / four_pi);
        w = cmul(mik, z4); res[7 * M + m] = CMPLX(creal(w) / four_pi, cimag(w) / four_pi);
        double coff = pow(kk, 2) / (32 * XO_PI * XO_PI * eta);
        pw[0 * M + m] = coff * pow(cabs(z1), 2);
        pw[1 * M + m] = coff * pow(cabs(z2), 2);
      }
  for (size_t q = 0; q < 8 * M; ++q) {
    out[2 * q] = creal(res[q]);
    out[2 * q + 1] = cimag(res[q]);
  }
  free(res);
  return 0;
}

/* ------------------------------------------------------------ Drude ADE */

static int init_drude(xo_sim *s) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  const size_t ncell = nx * ny * nz;
  s->dmask = (int *)calloc(ncell > 0 ? ncell : 1, sizeof(int));
  s->dk = (double *)calloc((size_t)s->nobj, sizeof(double));
  s->dbeta = (double *)calloc((size_t)s->nobj, sizeof(double));
  s->dca = (double *)calloc((size_t)s->nobj, sizeof(double));
  s->dcb = (double *)calloc((size_t)s->nobj, sizeof(double));
  s->dxi = (double *)calloc((size_t)s->nobj, sizeof(double));
  s->dcc = (double *)calloc((size_t)s->nobj, sizeof(double));
  if (!s->dmask || !s->dk || !s->dbeta || !s->dca || !s->dcb || !s->dxi || !s->dcc) FAIL("out of memory");
  for (int q = 0; q < 3; ++q) {
    size_t sh[3];
    field_shape(s, q, sh);
    if (a3_alloc(&s->J[q], sh[0], sh[1], sh[2], 0.0)) FAIL("out of memory");
    if (s->has_lorentz &&
        (a3_alloc(&s->Jp[q], sh[0], sh[1], sh[2], 0.0) || a3_alloc(&s->Ep[q], sh[0], sh[1], sh[2], 0.0)))
      FAIL("out of memory");
  }
  double dt = s->dt, dl = s->dl;
  for (int q = 0; q < s->nobj; ++q) {
    const obj_t *o = &s->objs[q];
    if (!o->drude) continue;
    /* dl is the grid's base size (dispersive_material.h:47-48) */
    double sigma_e = o->sigma_e;
    if (o->drude == 2) {
      /* DebyMedium::init (dispersive_material.cpp:205-222) */
      double delta_eps = o->eps_s - o->eps_inf;
      s->dk[q] = (2 * o->tau - dt) / (2 * o->tau + dt);
      s->dbeta[q] = (2 * delta_eps * EPS0 * dt) / (2 * o->tau + dt);
      double temp = 2 * EPS0 * o->eps_inf + s->dbeta[q] + sigma_e * dt;
      s->dca[q] = (2 * EPS0 * o->eps_inf + s->dbeta[q] - sigma_e * dt) / temp;
      s->dcb[q] = (2 * dt) / (temp * dl);
    } else if (o->drude == 3) {
      /* LorentzMedium::init (dispersive_material.cpp:17-36): dk = alpha,
         dbeta = gamma, dca / dcb the E and previous-E weights, dcc the curl
         weight */
      double delta_eps = o->eps_s - o->eps_inf;
      double t0 = o->nu * dt + 1;
      s->dk[q] = (2 - o->omega_p * o->omega_p * dt * dt) / t0;
      s->dxi[q] = (o->nu * dt - 1) / t0;
      s->dbeta[q] = (EPS0 * delta_eps * o->omega_p * o->omega_p * dt * dt) / t0;
      double t1 = 4 * EPS0 * o->eps_inf + s->dbeta[q] + 2 * sigma_e * dt;
      s->dca[q] = (4 * EPS0 * o->eps_inf - 2 * sigma_e * dt) / t1;
      s->dcb[q] = s->dbeta[q] / t1;
      s->dcc[q] = (4 * dt) / (t1 * dl);
    } else {
      /* DrudeMedium::init (dispersive_material.cpp:127-144) */
      s->dk[q] = (2 - o->nu * dt) / (2 + o->nu * dt);
      s->dbeta[q] = (o->omega_p * o->omega_p * EPS0 * dt) / (2 + o->nu * dt);
      double temp = 2 * EPS0 * o->eps_inf + s->dbeta[q] * dt + sigma_e * dt;
      s->dca[q] = (2 * EPS0 * o->eps_inf - s->dbeta[q] * dt - sigma_e * dt) / temp;
      s->dcb[q] = 2 * dt / (temp * dl);
    }
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j)
        for (size_t k = 0; k < nz; ++k) {
          double pt[3] = {s->hc[0][i], s->hc[1][j], s->hc[2][k]};
          if (is_inside(o, pt)) s->dmask[(i * ny + j) * nz + k] = q + 1;
        }
  }
  return 0;
}

/* One dispersive cell of component c: DrudeMedium::updateE{x,y,z}
   (dispersive_material.cpp:146-193), DebyMedium::updateE{x,y,z} (224-275)
   or LorentzMedium::updateE{x,y,z} (38-117).  curl is the reference's
   -(dH_a) + (dH_b) of the component, in H differences. */
static void disp_cell(xo_sim *s, int q, int c, size_t i, size_t j, size_t k, double curl) {
  const double dl = s->dl, dt = s->dt;
  double *e = &AT(s->f[c], i, j, k), *jc = &AT(s->J[c], i, j, k);
  const double e_old = *e;
  double j_sum = 0;
  if (s->objs[q].drude == 3) {
    double *ep = &AT(s->Ep[c], i, j, k), *jp = &AT(s->Jp[c], i, j, k);
    const double j_old = *jc;
    j_sum += 0.5 * dl * ((1 + s->dk[q]) * *jc + s->dxi[q] * *jp);
    *e = s->dca[q] * *e + s->dcb[q] * *ep + s->dcc[q] * (curl - j_sum);
    *jc = s->dk[q] * *jc + s->dxi[q] * *jp + s->dbeta[q] * (*e - *ep) / (2 * dt);
    *ep = e_old;
    *jp = j_old;
    return;
  }
  j_sum += 0.5 * dl * ((1 + s->dk[q]) * *jc);
  *e = s->dca[q] * *e + s->dcb[q] * (curl - j_sum);
  if (s->objs[q].drude == 2)
    *jc = s->dk[q] * *jc + s->dbeta[q] * (*e - e_old) / dt;
  else
    *jc = s->dk[q] * *jc + s->dbeta[q] * (*e + e_old);
}

/* simulation/simulation_update.cpp:221-280 with the dispersive media's
   updateE{x,y,z} on their cells */
static void update_e_drude(xo_sim *s) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  a3 *c = s->c;
  const int *dm = s->dmask;
#define DRUDE_AT(i, j, k) dm[((i) * ny + (j)) * nz + (k)]
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 1; j < ny; ++j)
      for (size_t k = 1; k < nz; ++k) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          disp_cell(s, d - 1, XO_EX, i, j, k, -(HY(i, j, k) - HY(i, j, k - 1)) + (HZ(i, j, k) - HZ(i, j - 1, k)));
          continue;
        }
        EX(i, j, k) = AT(c[CEXE], i, j, k) * EX(i, j, k) +
                      AT(c[CEXHY], i, j, k) * (HY(i, j, k) - HY(i, j, k - 1)) +
                      AT(c[CEXHZ], i, j, k) * (HZ(i, j, k) - HZ(i, j - 1, k));
      }
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < ny; ++j)
    for (size_t k = 1; k < nz; ++k)
      for (size_t i = 1; i < nx; ++i) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          disp_cell(s, d - 1, XO_EY, i, j, k, -(HZ(i, j, k) - HZ(i - 1, j, k)) + (HX(i, j, k) - HX(i, j, k - 1)));
          continue;
        }
        EY(i, j, k) = AT(c[CEYE], i, j, k) * EY(i, j, k) +
                      AT(c[CEYHX], i, j, k) * (HX(i, j, k) - HX(i, j, k - 1)) +
                      AT(c[CEYHZ], i, j, k) * (HZ(i, j, k) - HZ(i - 1, j, k));
      }
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < nz; ++k)
    for (size_t i = 1; i < nx; ++i)
      for (size_t j = 1; j < ny; ++j) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          disp_cell(s, d - 1, XO_EZ, i, j, k, -(HX(i, j, k) - HX(i, j - 1, k)) + (HY(i, j, k) - HY(i - 1, j, k)));
          continue;
        }
        EZ(i, j, k) = AT(c[CEZE], i, j, k) * EZ(i, j, k) +
                      AT(c[CEZHX], i, j, k) * (HX(i, j, k) - HX(i, j - 1, k)) +
                      AT(c[CEZHY], i, j, k) * (HY(i, j, k) - HY(i - 1, j, k));
      }
#undef DRUDE_AT
}

/* --------------------------------------------------------- current source */

static double lumped_wave(const csrc_t *c, double t) {
  return wave_at(c->wkind, c->amp, c->tau, c->t0, c->wp3, t);
}

/* cell sizes of a lumped element's cell along its axis a: (da, db, dc) of
   the lumped elements' init (da, db: the H-node sizes across, dc: the
   E-node size along).  The reference's Y branches of Resistor, Capacitor
   and Inductor assign _da twice and never _db; the intended (dz, dx) is
   taken, as for CurrentSource. */
static void lumped_sizes(const xo_sim *s, int a, size_t i, size_t j, size_t k, double *da, double *db,
                         double *dc) {
  if (a == 0) { *da = s->hs[1][j]; *db = s->hs[2][k]; *dc = s->es[0][i]; }
  else if (a == 1) { *da = s->hs[2][k]; *db = s->hs[0][i]; *dc = s->es[1][j]; }
  else { *da = s->hs[0][i]; *db = s->hs[1][j]; *dc = s->es[2][k]; }
}

static int init_csrc(xo_sim *s, csrc_t *c) {
  /* the grid box of the element's cube */
  size_t st[3], en[3];
  for (int a = 0; a < 3; ++a) {
    double lo = c->origin[a], hi = c->origin[a] + c->size[a];
    if (transform(s, a, lo, &st[a]) || transform(s, a, hi, &en[a])) return -1;
    if (en[a] < st[a]) FAIL("end < start");
    if (en[a] == st[a]) FAIL("lumped element box has no cells");
  }
  c->is = st[0]; c->ie = en[0]; c->js = st[1]; c->je = en[1]; c->ks = st[2]; c->ke = en[2];
  int a = c->ori / 2;
  size_t ni = c->ie - c->is, nj = c->je - c->js, nk = c->ke - c->ks, ncell = ni * nj * nk;
  double rf = 0, af = c->amp;
  if (c->kind == LE_CURRENT) {
    /* current_source.cpp:34-116 */
    if (a == 0) {
      rf = c->R * (double)nk * (double)nj / (double)ni;
      af = c->amp / (double)(nk * nj);
    } else if (a == 1) {
      rf = c->R * (double)nk * (double)ni / (double)nj;
      af = c->amp / (double)(nk * ni);
    } else {
      rf = c->R * (double)ni * (double)nj / (double)nk;
      af = c->amp / (double)(ni * nj);
    }
    if (c->ori % 2 == 0) af *= -1;
  } else if (c->kind == LE_VOLTAGE) {
    /* voltage_source.cpp:40-107 (its YP branch never runs: the condition
       tests YN twice; taken as intended) */
    if (a == 0) {
      rf = c->R * (double)nj * (double)nk / (double)ni;
      af = c->amp / (double)ni;
    } else if (a == 1) {
      rf = c->R * (double)ni * (double)nk / (double)nj;
      af = c->amp / (double)nj;
    } else {
      rf = c->R * (double)ni * (double)nj / (double)nk;
      af = c->amp / (double)nk;
    }
    if (c->ori % 2 == 0) af *= -1;
  } else if (c->kind == LE_RESISTOR || c->kind == LE_INDUCTOR) {
    /* resistor.cpp:22-76, inductor.cpp:19-68 */
    if (a == 0) rf = c->R * (double)nj * (double)nk / (double)ni;
    else if (a == 1) rf = c->R * (double)ni * (double)nk / (double)nj;
    else rf = c->R * (double)ni * (double)nj / (double)nk;
  } else {
    /* capacitor.cpp:25-61 */
    if (a == 0) rf = c->R * (double)ni / (double)(nj * nk);
    else if (a == 1) rf = c->R * (double)nj / (double)(ni * nk);
    else rf = c->R * (double)nk / (double)(ni * nj);
  }
  c->rf = rf;
  if (c->kind <= LE_VOLTAGE) {
    c->amp = af; /* setAmplitude(factor) before init */
    size_t T = s->T;
    c->wave = (double *)malloc(sizeof(double) * (T ? T : 1));
    if (!c->wave) FAIL("out of memory");
    double start = 0.5 * s->dt, stop = ((double)T - 0.5) * s->dt;
    double step = (stop - start) / fmax(1.0, (double)T - 1.0);
    for (size_t n = 0; n < T; ++n) {
      double tt = (n + 1 == T && T > 1) ? stop : start + step * (double)n;
      c->wave[n] = lumped_wave(c, tt);
    }
    c->coff_i = (double *)calloc(ncell ? ncell : 1, sizeof(double));
    if (!c->coff_i) FAIL("out of memory");
  }
  if (c->kind == LE_INDUCTOR) {
    c->jc = (double *)calloc(ncell ? ncell : 1, sizeof(double));
    c->cecjc = (double *)calloc(ncell ? ncell : 1, sizeof(double));
    c->cjcec = (double *)calloc(ncell ? ncell : 1, sizeof(double));
    if (!c->jc || !c->cecjc || !c->cjcec) FAIL("out of memory");
  }
  return 0;
}

static void csrc_correct_coff(xo_sim *s, csrc_t *c, a3 *eps_arr, a3 *sig_arr) {
  /* correctFDTDCoff of current_source.cpp:118-138, 167-187,
     voltage_source.cpp:109-166, resistor.cpp:78-121, capacitor.cpp:63-122,
     inductor.cpp:70-113 */
  int a = c->ori / 2;
  double rf = c->rf;
  double dt = s->dt;
  int ce, ca, cb;
  if (a == 0) { ce = CEXE; ca = CEXHY; cb = CEXHZ; }
  else if (a == 1) { ce = CEYE; ca = CEYHZ; cb = CEYHX; }
  else { ce = CEZE; ca = CEZHX; cb = CEZHY; }
  size_t nj = c->je - c->js, nk = c->ke - c->ks;
  for (size_t i = c->is; i < c->ie; ++i)
    for (size_t j = c->js; j < c->je; ++j)
      for (size_t k = c->ks; k < c->ke; ++k) {
        double da, db, dc;
        lumped_sizes(s, a, i, j, k, &da, &db, &dc);
        double e = AT(*eps_arr, i, j, k), sg = AT(*sig_arr, i, j, k);
        size_t q = ((i - c->is) * nj + (j - c->js)) * nk + (k - c->ks);
        if (c->kind == LE_INDUCTOR) {
          c->cjcec[q] = dt * dc / (rf * da * db);
          c->cecjc[q] = -(2 * dt) / (2 * e + dt * sg);
          continue;
        }
        double beta;
        if (c->kind == LE_VOLTAGE) {
          double alpha = da * db * rf;
          beta = dt * dc / alpha;
          c->coff_i[q] = -2 * dt / ((2 * e + dt * sg + beta) * alpha);
        } else if (c->kind == LE_CAPACITOR) {
          beta = 2 * rf * dc / (da * db);
          AT(s->c[ce], i, j, k) = (2 * e - dt * sg + beta) / (2 * e + dt * sg + beta);
          AT(s->c[ca], i, j, k) = -(2 * dt) / ((2 * e + dt * sg + beta) * db);
          AT(s->c[cb], i, j, k) = (2 * dt) / ((2 * e + dt * sg + beta) * da);
          continue;
        } else {
          beta = (dt * dc) / (rf * da * db);
          if (c->kind == LE_CURRENT) c->coff_i[q] = -2 * dt / ((2 * e + dt * sg + beta) * (db * da));
        }
        AT(s->c[ce], i, j, k) = (2 * e - dt * sg - beta) / (2 * e + dt * sg + beta);
        AT(s->c[ca], i, j, k) = -2 * dt / ((2 * e + dt * sg + beta) * db);
        AT(s->c[cb], i, j, k) = 2 * dt / ((2 * e + dt * sg + beta) * da);
      }
}

static void csrc_correct_e(xo_sim *s, csrc_t *c) {
  /* correctE of current_source.cpp:140-163, voltage_source.cpp:168-193,
     inductor.cpp:115-146 (resistor and capacitor: none) */
  if (c->kind == LE_RESISTOR || c->kind == LE_CAPACITOR) return;
  int a = c->ori / 2;
  a3 *f = &s->f[a == 0 ? XO_EX : (a == 1 ? XO_EY : XO_EZ)];
  size_t nj = c->je - c->js, nk = c->ke - c->ks;
  if (c->kind == LE_INDUCTOR) {
    for (size_t i = c->is; i < c->ie; ++i)
      for (size_t j = c->js; j < c->je; ++j)
        for (size_t k = c->ks; k < c->ke; ++k) {
          size_t q = ((i - c->is) * nj + (j - c->js)) * nk + (k - c->ks);
          AT(*f, i, j, k) += c->cecjc[q] * c->jc[q];
          c->jc[q] = c->cjcec[q] * AT(*f, i, j, k) + c->jc[q];
        }
    return;
  }
  double w = c->wave[s->cur];
  for (size_t i = c->is; i < c->ie; ++i)
    for (size_t j = c->js; j < c->je; ++j)
      for (size_t k = c->ks; k < c->ke; ++k)
        AT(*f, i, j, k) += c->coff_i[((i - c->is) * nj + (j - c->js)) * nk + (k - c->ks)] * w;
}

/* ------------------------------------------------------- line monitors */

static int init_lmon(xo_sim *s, lmon_t *m) {
  size_t st[3], en[3];
  for (int a = 0; a < 3; ++a) {
    double lo = m->origin[a], hi = m->origin[a] + m->size[a];
    if (transform(s, a, lo, &st[a]) || transform(s, a, hi, &en[a])) return -1;
    if (en[a] < st[a]) FAIL("end < start");
  }
  m->is = st[0]; m->ie = en[0]; m->js = st[1]; m->je = en[1]; m->ks = st[2]; m->ke = en[2];
  int a = m->ori / 2;
  size_t sh[3];
  if (m->kind == 0) {
    /* voltage_monitor.cpp:20-61: E_a over the box */
    field_shape(s, a, sh);
    if (m->ie > sh[0] || m->je > sh[1] || m->ke > sh[2]) FAIL("voltage monitor outside the field array");
  } else {
    /* current_monitor.cpp:20-64: the loop reads H one cell outside the box
       on both sides across the axis */
    const size_t *lo = &st[0], *hi = &en[0];
    if (lo[a] >= s->n[a]) FAIL("current monitor loop outside the grid");
    for (int b = 0; b < 3; ++b)
      if (b != a && (lo[b] < 1 || hi[b] + 1 > s->n[b])) FAIL("current monitor loop outside the grid");
  }
  m->value = (double *)calloc(s->T ? s->T : 1, sizeof(double));
  if (!m->value) FAIL("out of memory");
  return 0;
}

static void lmon_update(xo_sim *s, lmon_t *m) {
  size_t t = s->cur;
  int a = m->ori / 2;
  if (m->kind == 0) {
    /* voltage_monitor.cpp:63-93: value(t) += coff(c) E(i,j,k), with
       coff = -dc / (cells across), negated for the N orientation */
    size_t ni = m->ie - m->is, nj = m->je - m->js, nk = m->ke - m->ks;
    size_t across = a == 0 ? nj * nk : (a == 1 ? ni * nk : ni * nj);
    a3 *f = &s->f[a];
    double v = 0.0;
    for (size_t i = m->is; i < m->ie; ++i)
      for (size_t j = m->js; j < m->je; ++j)
        for (size_t k = m->ks; k < m->ke; ++k) {
          size_t idx = a == 0 ? i : (a == 1 ? j : k);
          double coff = -s->hs[a][idx] / (double)across;
          if (m->ori % 2 == 0) coff *= -1;
          v += coff * AT(*f, i, j, k);
        }
    m->value[t] = v;
    return;
  }
  /* current_monitor.cpp:66-122 */
  const double pos = m->ori % 2 == 0 ? -1.0 : 1.0;
  size_t is = m->is, ie = m->ie, js = m->js, je = m->je, ks = m->ks, ke = m->ke;
  double i1 = 0, i2 = 0;
  if (a == 0) {
    for (size_t j = js; j <= je; ++j) i1 += (HY(is, j, ks - 1) - HY(is, j, ke)) * s->es[1][j];
    for (size_t k = ks; k <= ke; ++k) i2 += (HZ(is, je, k) - HZ(is, js - 1, k)) * s->es[2][k];
    m->value[t] = pos * (i2 + i1);
  } else if (a == 1) {
    for (size_t k = ks; k <= ke; ++k) i1 += (HZ(is - 1, js, k) - HZ(ie, js, k)) * s->es[2][k];
    for (size_t i = is; i <= ie; ++i) i2 += (HX(i, js, ke) - HX(i, js, ks - 1)) * s->es[0][i];
    m->value[t] = pos * (i1 + i2);
  } else {
    for (size_t i = is; i <= ie; ++i) i1 += (HX(i, js - 1, ks) - HX(i, je, ks)) * s->es[0][i];
    for (size_t j = js; j <= je; ++j) i2 += (HY(ie, j, ks) - HY(is - 1, j, ks)) * s->es[1][j];
    m->value[t] = pos * (i1 + i2);
  }
}

int xo_sim_get_line_monitors(const xo_sim *s, double *out) {
  if (!s->inited) FAIL("not initialised");
  for (int q = 0; q < s->nlmon; ++q) memcpy(out + (size_t)q * s->T, s->lmon[q].value, s->T * sizeof(double));
  return 0;
}

/* source waveform table of lumped element q (T values at (n + 1/2) dt) */
int xo_sim_get_lumped_wave(const xo_sim *s, int q, double *out) {
  if (!s->inited || q < 0 || q >= s->nsrc || !s->src[q].wave) FAIL("no such source");
  memcpy(out, s->src[q].wave, s->T * sizeof(double));
  return 0;
}

/* inductor currents of lumped element q (its box, C order) */
int xo_sim_get_inductor_current(const xo_sim *s, int q, double *out) {
  if (!s->inited || q < 0 || q >= s->nsrc || s->src[q].kind != LE_INDUCTOR) FAIL("no such inductor");
  const csrc_t *c = &s->src[q];
  memcpy(out, c->jc, (c->ie - c->is) * (c->je - c->js) * (c->ke - c->ks) * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------------ init */

int xo_sim_init(xo_sim *s, size_t total_steps) {
  if (s->inited) FAIL("already initialised");
  s->T = total_steps;
  s->cur = 0;
  if (s->nobj == 0) FAIL("no objects");
  /* simulation_init.cpp:10-22 */
  if (init_grid(s)) return -1;
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  /* fdtd_basic_coff.cpp:11-29, 342-346: cfl arrives as float */
  double cfl = (double)s->cfl;
  double dx = s->dl, dy = s->dl, dz = s->dl;
  if (s->dim2) /* fdtd_basic_coff.cpp:13-19, 347-349: no z term when nz <= 1 */
    s->dt = cfl / (C0() * sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy)));
  else
    s->dt = cfl / (C0() * sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)));
  if (s->dim2 && (s->has_nf || s->nsrc || s->nlmon || s->has_design || s->has_dft || s->has_drude))
    FAIL("a 2-D grid takes objects, PML on x and y, TFSF2D, NffftBroadBand and probes only");
  if (s->has_tfsf && s->tf.dim2 != s->dim2)
    FAIL(s->dim2 ? "a 2-D grid needs TFSF2D" : "TFSF2D needs a 2-D grid");
  for (int q = 0; q < 6; ++q) {
    size_t sh[3];
    field_shape(s, q, sh);
    if (a3_alloc(&s->f[q], sh[0], sh[1], sh[2], 0.0)) FAIL("out of memory");
  }
  /* fdtd_basic_coff.cpp:313-339: eps/mu/sigma arrays per component */
  a3 eps[3], mu[3], sg[3], sm[3];
  for (int q = 0; q < 3; ++q) {
    size_t she[3], shh[3];
    field_shape(s, q, she);
    field_shape(s, q + 3, shh);
    if (a3_alloc(&eps[q], she[0], she[1], she[2], EPS0) ||
        a3_alloc(&sg[q], she[0], she[1], she[2], 0.0) ||
        a3_alloc(&mu[q], shh[0], shh[1], shh[2], MU0()) ||
        a3_alloc(&sm[q], shh[0], shh[1], shh[2], 0.0))
      FAIL("out of memory");
  }
  /* object/object.cpp:53-73: mask at cell centres, applied at the same
     (i,j,k) index of every staggered array */
  for (int q = 0; q < s->nobj; ++q) {
    const obj_t *o = &s->objs[q];
    double e = (double)o->eps_r * EPS0, m = (double)o->mu_r * MU0();
    if (o->plane) {
      /* object/object_plane.cpp:8-50: sigma_e of the two tangential E
         components on the plane, nothing else */
      size_t st[3], en[3];
      for (int a = 0; a < 3; ++a)
        if (transform(s, a, o->p[a], &st[a]) || transform(s, a, o->p[a] + o->s[a], &en[a])) return -1;
      size_t is = st[0], ie = en[0], js = st[1], je = en[1], ks = st[2], ke = en[2];
      if (is == ie) {
        for (size_t j = js; j < je; ++j)
          for (size_t k = ks; k <= ke; ++k) AT(sg[1], is, j, k) = o->sigma_e;
        for (size_t j = js; j <= je; ++j)
          for (size_t k = ks; k < ke; ++k) AT(sg[2], is, j, k) = o->sigma_e;
      } else if (js == je) {
        for (size_t i = is; i < ie; ++i)
          for (size_t k = ks; k <= ke; ++k) AT(sg[0], i, js, k) = o->sigma_e;
        for (size_t i = is; i <= ie; ++i)
          for (size_t k = ks; k < ke; ++k) AT(sg[2], i, js, k) = o->sigma_e;
      } else if (ks == ke) {
        for (size_t i = is; i < ie; ++i)
          for (size_t j = js; j <= je; ++j) AT(sg[0], i, j, ks) = o->sigma_e;
        for (size_t i = is; i <= ie; ++i)
          for (size_t j = js; j < je; ++j) AT(sg[1], i, j, ks) = o->sigma_e;
      }
      continue;
    }
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j)
        for (size_t k = 0; k < nz; ++k) {
          /* grid/grid_space.cpp:241-247, 284-292: z of a 2-D cell centre is 0 */
          double pt[3] = {s->hc[0][i], s->hc[1][j], s->dim2 ? 0.0 : s->hc[2][k]};
          if (!is_inside(o, pt)) continue;
          for (int c = 0; c < 3; ++c) {
            AT(eps[c], i, j, k) = e;
            AT(mu[c], i, j, k) = m;
            AT(sg[c], i, j, k) = o->sigma_e;
            AT(sm[c], i, j, k) = o->sigma_m;
          }
        }
  }
  if (s->has_drude && init_drude(s)) return -1;
  if (s->has_design) {
    const size_t *b = s->dbox;
    if (b[1] > nx || b[3] > ny || b[5] > nz) FAIL("design box outside the grid");
    size_t q = 0;
    for (size_t i = b[0]; i < b[1]; ++i)
      for (size_t j = b[2]; j < b[3]; ++j)
        for (size_t k = b[4]; k < b[5]; ++k, ++q)
          for (int c = 0; c < 3; ++c) {
            AT(eps[c], i, j, k) = (double)s->design_eps[q] * EPS0;
            AT(mu[c], i, j, k) = MU0();
            AT(sg[c], i, j, k) = 0.0;
            AT(sm[c], i, j, k) = 0.0;
          }
  }
  if (s->has_dft) {
    size_t sh[3];
    field_shape(s, s->dft_comp, sh);
    size_t lim1 = sh[s->dft_axis == 0 ? 1 : 0], lim2 = sh[s->dft_axis == 2 ? 1 : 2];
    if (s->dft_index >= sh[s->dft_axis] || s->dft_hi1 > lim1 || s->dft_hi2 > lim2 ||
        s->dft_lo1 >= s->dft_hi1 || s->dft_lo2 >= s->dft_hi2)
      FAIL("DFT plane outside the component array");
    s->dft_acc = (double complex *)calloc((s->dft_hi1 - s->dft_lo1) * (s->dft_hi2 - s->dft_lo2),
                                          sizeof(double complex));
    if (!s->dft_acc) FAIL("out of memory");
  }
  if (s->has_tfsf && init_tfsf(s)) return -1;
  if (s->has_nf && init_nffft(s)) return -1;
  if (s->has_bb && init_bb(s)) return -1;
  for (int q = 0; q < s->nsrc; ++q)
    if (init_csrc(s, &s->src[q])) return -1;
  for (int q = 0; q < s->nlmon; ++q)
    if (init_lmon(s, &s->lmon[q])) return -1;
  /* fdtd_basic_coff.cpp:31-122 */
  for (int q = 0; q < 18; ++q) {
    size_t sh[3];
    int comp = q / 3;
    field_shape(s, comp, sh);
    if (a3_alloc(&s->c[q], sh[0], sh[1], sh[2], 0.0)) FAIL("out of memory");
  }
  double dt = s->dt;
  double **es = s->es, **hs = s->hs;
  calc_coff_e(s->c[CEXE], s->c[CEXHY], s->c[CEXHZ], eps[0], sg[0], dt, hs[1], 1, hs[2], 2);
  calc_coff_e(s->c[CEYE], s->c[CEYHZ], s->c[CEYHX], eps[1], sg[1], dt, hs[2], 2, hs[0], 0);
  calc_coff_e(s->c[CEZE], s->c[CEZHX], s->c[CEZHY], eps[2], sg[2], dt, hs[0], 0, hs[1], 1);
  calc_coff_h(s->c[CHXH], s->c[CHXEY], s->c[CHXEZ], mu[0], sm[0], dt, es[1], 1, es[2], 2);
  calc_coff_h(s->c[CHYH], s->c[CHYEZ], s->c[CHYEX], mu[1], sm[1], dt, es[2], 2, es[0], 0);
  calc_coff_h(s->c[CHZH], s->c[CHZEX], s->c[CHZEY], mu[2], sm[2], dt, es[0], 0, es[1], 1);
  for (int q = 0; q < s->nsrc; ++q) {
    int a = s->src[q].ori / 2;
    csrc_correct_coff(s, &s->src[q], &eps[a], &sg[a]);
  }
  for (int q = 0; q < 3; ++q) {
    a3_free(&eps[q]); a3_free(&mu[q]); a3_free(&sg[q]); a3_free(&sm[q]);
  }
  /* simulation_init.cpp:93-97 */
  for (int b = 0; b < s->npml; ++b)
    if (init_pml(s, &s->pml[b])) return -1;
  if (s->nprobe) {
    s->probe_buf = (double *)calloc((size_t)s->nprobe * (s->T ? s->T : 1), sizeof(double));
    for (int p = 0; p < s->nprobe; ++p) {
      size_t sh[3];
      const probe_t *pr = &s->probes[p];
      if (pr->comp < 0 || pr->comp > 5) FAIL("bad probe component");
      field_shape(s, pr->comp, sh);
      if (pr->i >= sh[0] || pr->j >= sh[1] || pr->k >= sh[2]) FAIL("probe outside the field array");
    }
  }
  s->inited = 1;
  return 0;
}

/* ------------------------------------------------------ DFT plane monitor */

/* element (u, v) of the monitor plane: axis a fixed at dft_index, the other
   two axes in increasing order */
static double *dft_cell(a3 *f, int a, size_t idx, size_t u, size_t v) {
  if (a == 0) return &AT(*f, idx, u, v);
  if (a == 1) return &AT(*f, u, idx, v);
  return &AT(*f, u, v, idx);
}

/* weight of the sample taken after step t: the magnetic-current weight of
   nffft_fd.cpp:102-103, dt exp(-i 2 pi f t dt) */
static double complex dft_weight(const xo_sim *s, size_t t) {
  double ph = -2 * XO_PI * s->dft_freq * ((double)t * s->dt);
  return s->dt * (cos(ph) + I * sin(ph));
}

static void dft_update(xo_sim *s, size_t t) {
  double complex w = dft_weight(s, t);
  size_t n2 = s->dft_hi2 - s->dft_lo2;
  for (size_t u = s->dft_lo1; u < s->dft_hi1; ++u)
    for (size_t v = s->dft_lo2; v < s->dft_hi2; ++v)
      s->dft_acc[(u - s->dft_lo1) * n2 + (v - s->dft_lo2)] +=
          *dft_cell(&s->f[s->dft_comp], s->dft_axis, s->dft_index, u, v) * w;
}

/* ----------------------------------------------------------------- step */

/* simulation/simulation_update.cpp:170-196 */
static void update_e(xo_sim *s) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  a3 *c = s->c;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 1; j < ny; ++j)
      for (size_t k = 1; k < nz; ++k)
        EX(i, j, k) = AT(c[CEXE], i, j, k) * EX(i, j, k) +
                      AT(c[CEXHY], i, j, k) * (HY(i, j, k) - HY(i, j, k - 1)) +
                      AT(c[CEXHZ], i, j, k) * (HZ(i, j, k) - HZ(i, j - 1, k));
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < ny; ++j)
    for (size_t k = 1; k < nz; ++k)
      for (size_t i = 1; i < nx; ++i)
        EY(i, j, k) = AT(c[CEYE], i, j, k) * EY(i, j, k) +
                      AT(c[CEYHX], i, j, k) * (HX(i, j, k) - HX(i, j, k - 1)) +
                      AT(c[CEYHZ], i, j, k) * (HZ(i, j, k) - HZ(i - 1, j, k));
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < nz; ++k)
    for (size_t i = 1; i < nx; ++i)
      for (size_t j = 1; j < ny; ++j)
        EZ(i, j, k) = AT(c[CEZE], i, j, k) * EZ(i, j, k) +
                      AT(c[CEZHX], i, j, k) * (HX(i, j, k) - HX(i, j - 1, k)) +
                      AT(c[CEZHY], i, j, k) * (HY(i, j, k) - HY(i - 1, j, k));
}

/* simulation/simulation_update.cpp:84-100 */
static void update_h(xo_sim *s) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  a3 *c = s->c;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j)
      for (size_t k = 0; k < nz; ++k) {
        HX(i, j, k) = AT(c[CHXH], i, j, k) * HX(i, j, k) +
                      AT(c[CHXEY], i, j, k) * (EY(i, j, k + 1) - EY(i, j, k)) +
                      AT(c[CHXEZ], i, j, k) * (EZ(i, j + 1, k) - EZ(i, j, k));
        HY(i, j, k) = AT(c[CHYH], i, j, k) * HY(i, j, k) +
                      AT(c[CHYEZ], i, j, k) * (EZ(i + 1, j, k) - EZ(i, j, k)) +
                      AT(c[CHYEX], i, j, k) * (EX(i, j, k + 1) - EX(i, j, k));
        HZ(i, j, k) = AT(c[CHZH], i, j, k) * HZ(i, j, k) +
                      AT(c[CHZEX], i, j, k) * (EX(i, j + 1, k) - EX(i, j, k)) +
                      AT(c[CHZEY], i, j, k) * (EY(i + 1, j, k) - EY(i, j, k));
      }
}

int xo_sim_step(xo_sim *s, size_t nsteps) {
  if (!s->inited) FAIL("not initialised");
  if (s->cur + nsteps > s->T) FAIL("step beyond total_steps");
  for (size_t q = 0; q < nsteps; ++q) {
    /* simulation/simulation_update.cpp:9-36 */
    if (s->has_drude)
      update_e_drude(s);
    else
      update_e(s);
    for (int c = 0; c < s->nsrc; ++c) csrc_correct_e(s, &s->src[c]);
    if (s->has_tfsf) tfsf_update_e(s);
    for (int b = 0; b < s->npml; ++b) pml_update_e(s, &s->pml[b]);
    update_h(s);
    if (s->has_tfsf) tfsf_update_incident(s, s->cur);
    if (s->has_tfsf) tfsf_update_h(s);
    for (int b = 0; b < s->npml; ++b) pml_update_h(s, &s->pml[b]);
    if (s->has_nf) nffft_update(s);
    if (s->has_bb) bb_update(s);
    if (s->has_dft) dft_update(s, s->cur);
    for (int p = 0; p < s->nprobe; ++p) {
      const probe_t *pr = &s->probes[p];
      s->probe_buf[(size_t)p * s->T + s->cur] = AT(s->f[pr->comp], pr->i, pr->j, pr->k);
    }
    for (int m = 0; m < s->nlmon; ++m) lmon_update(s, &s->lmon[m]);
    s->cur++;
  }
  return 0;
}

/* ------------------------------------------------ broadband far field */

/* nffft_broadband.cpp:320-324: f = linspace(-fs/2, fs/2, 2T-1) with
   fs = 1/dt (xtensor's start + q * step, step = (stop - start) / (n - 1)),
   k = 2 pi f / c0 */
int xo_sim_bb_frequencies(const xo_sim *s, double *freq, double *k) {
  if (!s->inited || !s->has_bb) FAIL("no broadband NF-FF surface");
  size_t Ns = 2 * s->T - 1;
  double fs = 1 / s->dt;
  double start = -fs / 2, stop = fs / 2;
  double step = (stop - start) / (double)(Ns > 1 ? Ns - 1 : 1);
  for (size_t q = 0; q < Ns; ++q) {
    double f = start + step * (double)q;
    if (freq) freq[q] = f;
    if (k) k[q] = 2 * XO_PI * f / C0();
  }
  return 0;
}

/* One surface cell of calculateAuxiliary (nffft_broadband.cpp:441-466):
   for the cell's four series (j_a, j_b, m_a, m_b) add
   ds * exp(i delay k) * fftshift(fft(series)) to (n_a, n_b, l_a, l_b).
   fft is the forward DFT over Ns = 2T-1 points of the series, zero past
   step T-1: bin q holds m = q - (T-1), X(m) = sum_n x[n] e^{-2 pi i m n / Ns}
   and X(-m) = conj(X(m)) for the real series. */
static void bb_cell(const xo_sim *s, const double *k, const double complex *tw, double complex *X,
                    double delay, double ds, const size_t series[4], double complex *const tgt[4]) {
  const bb_t *b = &s->bb;
  size_t T = s->T, Ns = 2 * T - 1;
  for (int u = 0; u < 4; ++u) {
    const double *x = b->samples + series[u];
    for (size_t m = 0; m < T; ++m) {
      double complex acc = 0;
      for (size_t n = 0; n < T; ++n) {
        double v = x[n * 8 * b->Q];
        double complex w = tw[(m * n) % Ns];
        acc = CMPLX(creal(acc) + v * creal(w), cimag(acc) + v * cimag(w));
      }
      X[m] = acc;
    }
    for (size_t q = 0; q < Ns; ++q) {
      long long m = (long long)q - (long long)(T - 1);
      double complex xq = m >= 0 ? X[m] : conj(X[-m]);
      double complex ph = cscale(ds, cexpi(delay * k[q]));
      tgt[u][q] = cadd(tgt[u][q], cmul(ph, xq));
    }
  }
}

/* nffft_broadband.cpp:316-439 (calculateLNXYZ): out holds the six
   surface integrals n_x, n_y, n_z, l_x, l_y, l_z over the 2T-1 frequency
   bins, each interleaved re,im.  k: the bins' wave numbers, or NULL for
   xo_sim_bb_frequencies'. */
int xo_sim_bb_sums(xo_sim *s, const double *k_in, double *out) {
  if (!s->inited || !s->has_bb) FAIL("no broadband NF-FF surface");
  const bb_t *b = &s->bb;
  size_t T = s->T, Ns = 2 * T - 1;
  double *k = (double *)malloc(Ns * sizeof(double));
  double complex *tw = (double complex *)malloc(Ns * sizeof(double complex));
  double complex *X = (double complex *)malloc(T * sizeof(double complex));
  double complex *acc = (double complex *)calloc(6 * Ns, sizeof(double complex));
  if (!k || !tw || !X || !acc) {
    free(k); free(tw); free(X); free(acc);
    FAIL("out of memory");
  }
  if (k_in)
    memcpy(k, k_in, Ns * sizeof(double));
  else
    xo_sim_bb_frequencies(s, NULL, k);
  for (size_t r = 0; r < Ns; ++r) {
    double a = 2 * XO_PI * (double)r / (double)Ns;
    tw[r] = CMPLX(cos(a), -sin(a));
  }
  /* nffft_broadband.cpp:27-28 */
  const double rx = sin(b->theta) * cos(b->phi), ry = sin(b->theta) * sin(b->phi), rz = cos(b->theta);
  /* grid/grid_box.cpp:101-105 */
  const double cx = (double)(b->li + b->nxb / 2), cy = (double)(b->lj + b->nyb / 2),
               cz = (double)(b->lk + b->nzb / 2);
  const double dl = s->dl, ds = dl * dl;
  const size_t Q = b->Q, qy = b->nyb * b->nzb, qz = qy + b->nzb * b->nxb;
  double complex *nx_ = acc, *ny_ = acc + Ns, *nz_ = acc + 2 * Ns;
  double complex *lx_ = acc + 3 * Ns, *ly_ = acc + 4 * Ns, *lz_ = acc + 5 * Ns;
#define BB_DELAY(I, J, K, OX, OY, OZ)                                                   \
  ((((double)(I) + (OX)) - cx) * rx * dl + (((double)(J) + (OY)) - cy) * ry * dl) + \
      (((double)(K) + (OZ)) - cz) * rz * dl
  /* x faces: (n_y, n_z, l_y, l_z) <- (jy, jz, my, mz) = currents 2, 3, 0, 1 */
  for (int side = 0; side < 2; ++side) {
    size_t i = side ? b->ri : b->li;
    double complex *const tgt[4] = {ny_, nz_, ly_, lz_};
    for (size_t j = b->lj; j < b->rj; ++j)
      for (size_t kk = b->lk; kk < b->rk; ++kk) {
        size_t cell = (j - b->lj) * b->nzb + (kk - b->lk), c0 = 4 * (size_t)side;
        size_t ser[4] = {(c0 + 2) * Q + cell, (c0 + 3) * Q + cell, c0 * Q + cell, (c0 + 1) * Q + cell};
        bb_cell(s, k, tw, X, BB_DELAY(i, j, kk, 0.0, 0.5, 0.5), ds, ser, tgt);
      }
  }
  /* y faces: (n_x, n_z, l_x, l_z) <- (jx, jz, mx, mz) = currents 3, 2, 1, 0 */
  for (int side = 0; side < 2; ++side) {
    size_t j = side ? b->rj : b->lj;
    double complex *const tgt[4] = {nx_, nz_, lx_, lz_};
    for (size_t i = b->li; i < b->ri; ++i)
      for (size_t kk = b->lk; kk < b->rk; ++kk) {
        size_t cell = qy + (kk - b->lk) * b->nxb + (i - b->li), c0 = 4 * (size_t)side;
        size_t ser[4] = {(c0 + 3) * Q + cell, (c0 + 2) * Q + cell, (c0 + 1) * Q + cell, c0 * Q + cell};
        bb_cell(s, k, tw, X, BB_DELAY(i, j, kk, 0.5, 0.0, 0.5), ds, ser, tgt);
      }
  }
  /* z faces (only for nz > 1): (n_x, n_y, l_x, l_y) <- (jx, jy, mx, my) = currents 2, 3, 0, 1 */
  if (b->nzb > 1)
    for (int side = 0; side < 2; ++side) {
      size_t kk = side ? b->rk : b->lk;
      double complex *const tgt[4] = {nx_, ny_, lx_, ly_};
      for (size_t i = b->li; i < b->ri; ++i)
        for (size_t j = b->lj; j < b->rj; ++j) {
          size_t cell = qz + (i - b->li) * b->nyb + (j - b->lj), c0 = 4 * (size_t)side;
          size_t ser[4] = {(c0 + 2) * Q + cell, (c0 + 3) * Q + cell, c0 * Q + cell, (c0 + 1) * Q + cell};
          bb_cell(s, k, tw, X, BB_DELAY(i, j, kk, 0.5, 0.5, 0.0), ds, ser, tgt);
        }
    }
#undef BB_DELAY
  for (size_t q = 0; q < 6 * Ns; ++q) {
    out[2 * q] = creal(acc[q]);
    out[2 * q + 1] = cimag(acc[q]);
  }
  free(k); free(tw); free(X); free(acc);
  return 0;
}

size_t xo_sim_bb_samples_size(const xo_sim *s) {
  return s->inited && s->has_bb ? s->T * 8 * s->bb.Q : 0;
}

int xo_sim_get_bb_samples(const xo_sim *s, double *out) {
  if (!s->inited || !s->has_bb) FAIL("no broadband NF-FF surface");
  memcpy(out, s->bb.samples, s->T * 8 * s->bb.Q * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------- accessors */

#ifdef _OPENMP
#include <omp.h>
#endif
/* threads for the loops that are Jacobi-style (bitwise independent of the
   thread count); 1 reproduces the reference's single-threaded run */
int xo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int xo_sim_set_drude_j(xo_sim *s, int comp, const double *in) {
  if (!s->inited || !s->has_drude) FAIL("no Drude cells");
  if (comp < 0 || comp > 2) FAIL("bad component");
  a3 *a = &s->J[comp];
  memcpy(a->d, in, a->n0 * a->n1 * a->n2 * sizeof(double));
  return 0;
}

int xo_sim_get_drude_j(xo_sim *s, int comp, double *out) {
  if (!s->inited || !s->has_drude) FAIL("no Drude cells");
  if (comp < 0 || comp > 2) FAIL("bad component");
  const a3 *a = &s->J[comp];
  memcpy(out, a->d, a->n0 * a->n1 * a->n2 * sizeof(double));
  return 0;
}

int xo_sim_dims(const xo_sim *s, size_t out[3]) {
  for (int a = 0; a < 3; ++a) out[a] = s->n[a];
  return s->inited ? 0 : -1;
}
double xo_sim_dt(const xo_sim *s) { return s->dt; }
size_t xo_sim_current_step(const xo_sim *s) { return s->cur; }

size_t xo_sim_field_size(const xo_sim *s, int comp) {
  if (!s->inited || comp < 0 || comp > 5) return 0;
  return s->f[comp].n0 * s->f[comp].n1 * s->f[comp].n2;
}
int xo_sim_get_field(const xo_sim *s, int comp, double *out) {
  if (!s->inited || comp < 0 || comp > 5) FAIL("bad field request");
  memcpy(out, s->f[comp].d, xo_sim_field_size(s, comp) * sizeof(double));
  return 0;
}
int xo_sim_set_field(xo_sim *s, int comp, const double *in) {
  if (!s->inited || comp < 0 || comp > 5) FAIL("bad field request");
  memcpy(s->f[comp].d, in, xo_sim_field_size(s, comp) * sizeof(double));
  return 0;
}
int xo_sim_get_coeff(const xo_sim *s, int which, double *out) {
  if (!s->inited || which < 0 || which >= 18) FAIL("bad coefficient request");
  const a3 *a = &s->c[which];
  memcpy(out, a->d, a->n0 * a->n1 * a->n2 * sizeof(double));
  return 0;
}
int xo_sim_get_pml_profile(const xo_sim *s, int b, double *out) {
  if (!s->inited || b < 0 || b >= s->npml) FAIL("bad boundary index");
  const pml_t *p = &s->pml[b];
  int d = p->d;
  for (int i = 0; i < d; ++i) {
    out[0 * d + i] = p->be[i]; out[1 * d + i] = p->ae[i];
    out[2 * d + i] = p->bm[i]; out[3 * d + i] = p->am[i];
    out[4 * d + i] = p->ke[i]; out[5 * d + i] = p->km[i];
  }
  return 0;
}
size_t xo_sim_tfsf_aux_size(const xo_sim *s) {
  return (s->inited && s->has_tfsf) ? s->tf.L : 0;
}
int xo_sim_get_tfsf_aux(const xo_sim *s, double *e_inc, double *h_inc) {
  if (!s->inited || !s->has_tfsf) FAIL("no TFSF source");
  memcpy(e_inc, s->tf.e_inc, s->tf.L * sizeof(double));
  memcpy(h_inc, s->tf.h_inc, (s->tf.L - 1) * sizeof(double));
  return 0;
}
int xo_sim_get_waveform(const xo_sim *s, double *out) {
  if (!s->inited || !s->has_tfsf) FAIL("no TFSF source");
  memcpy(out, s->tf.wave, s->T * sizeof(double));
  return 0;
}
int xo_sim_get_nffft(const xo_sim *s, double *out) {
  if (!s->inited || !s->has_nf) FAIL("no NF-FF surface");
  size_t n = xo_sim_nffft_size(s);
  for (size_t q = 0; q < n; ++q) {
    out[2 * q] = creal(s->nf.acc[q]);
    out[2 * q + 1] = cimag(s->nf.acc[q]);
  }
  return 0;
}
int xo_sim_get_probes(const xo_sim *s, double *out) {
  if (!s->inited) FAIL("not initialised");
  memcpy(out, s->probe_buf, (size_t)s->nprobe * s->T * sizeof(double));
  return 0;
}

/* ---------------------------------------------------------------- adjoint
 * Discrete adjoint of the run loop with respect to the fields, and the
 * gradient of a DFT-plane objective with respect to the permittivity of a
 * design box.  The reference has no adjoint; this is the exact transpose of
 * the forward operations above, undone one by one in reverse order:
 *   forward step t: update_e, sources, TFSF-E, CPML-E (faces in order),
 *                   update_h, incident grid, TFSF-H, CPML-H (faces in order),
 *                   NF-FF, DFT tap
 *   reverse step t: DFT tap^T (inject), CPML-H^T (faces reversed), update_h^T,
 *                   CPML-E^T (faces reversed), gradient, update_e^T
 * Sources, TFSF corrections and the incident grid add terms that do not
 * depend on the fields, so their transposes are empty.
 */

typedef struct {
  a3 f[6];       /* adjoint of Ex..Hz */
  double **psi;  /* [npml][4] adjoint of each face's psi arrays */
  a3 J[3];       /* adjoint of the Drude currents (when there are Drude cells) */
} adj_t;

static int adj_alloc(xo_sim *s, adj_t *a) {
  memset(a, 0, sizeof *a);
  for (int q = 0; q < 6; ++q) {
    size_t sh[3];
    field_shape(s, q, sh);
    if (a3_alloc(&a->f[q], sh[0], sh[1], sh[2], 0.0)) FAIL("out of memory");
  }
  if (s->has_drude)
    for (int q = 0; q < 3; ++q) {
      size_t sh[3];
      field_shape(s, q, sh);
      if (a3_alloc(&a->J[q], sh[0], sh[1], sh[2], 0.0)) FAIL("out of memory");
    }
  a->psi = (double **)calloc((size_t)(s->npml ? s->npml : 1) * 4, sizeof(double *));
  if (!a->psi) FAIL("out of memory");
  for (int b = 0; b < s->npml; ++b)
    for (int q = 0; q < 4; ++q) {
      size_t sz = (size_t)s->pml[b].d * s->pml[b].tn1[q] * s->pml[b].tn2[q];
      a->psi[4 * b + q] = (double *)calloc(sz ? sz : 1, sizeof(double));
      if (!a->psi[4 * b + q]) FAIL("out of memory");
    }
  return 0;
}

static void adj_free(xo_sim *s, adj_t *a) {
  for (int q = 0; q < 6; ++q) a3_free(&a->f[q]);
  for (int q = 0; q < 3; ++q) a3_free(&a->J[q]);
  if (a->psi)
    for (int b = 0; b < 4 * s->npml; ++b) free(a->psi[b]);
  free(a->psi);
}

/* transpose of pml_update_h: psi' = bm psi + am (e(n+1) - e(n)); h(n) += cpsi psi' */
static void adj_pml_h(xo_sim *s, pml_t *p, double **pbar, adj_t *ad) {
  int a = p->ori / 2;
  for (int q = 2; q <= 3; ++q) {
    a3 *hb = &ad->f[PML_COMP[a][q]];
    a3 *eb = &ad->f[PML_COMP[a][q == 3 ? 0 : 1]];
    size_t t1 = p->tn1[q], t2 = p->tn2[q];
    for (int l = 0; l < p->d; ++l) {
      size_t n = p->start + (size_t)l;
      for (size_t u = 0; u < t1; ++u)
        for (size_t v = 0; v < t2; ++v) {
          size_t pi = ((size_t)l * t1 + u) * t2 + v;
          double chi = pbar[q][pi] + p->cpsi[q][pi] * *fa_at(hb, a, n, u, v);
          *fa_at(eb, a, n + 1, u, v) += p->am[l] * chi;
          *fa_at(eb, a, n, u, v) -= p->am[l] * chi;
          pbar[q][pi] = p->bm[l] * chi;
        }
    }
  }
  (void)s;
}

/* transpose of pml_update_e: psi' = be psi + ae (h(n) - h(n-1)); e(n) += cpsi psi' */
static void adj_pml_e(xo_sim *s, pml_t *p, double **pbar, adj_t *ad) {
  int a = p->ori / 2;
  int positive = p->ori % 2;
  for (int q = 0; q < 2; ++q) {
    a3 *eb = &ad->f[PML_COMP[a][q]];
    a3 *hb = &ad->f[PML_COMP[a][q == 0 ? 3 : 2]];
    size_t t1 = p->tn1[q], t2 = p->tn2[q];
    for (int l = 0; l < p->d; ++l) {
      size_t n = p->start + (size_t)l + (positive ? 0 : 1);
      for (size_t u = 0; u < t1; ++u)
        for (size_t v = 0; v < t2; ++v) {
          size_t pi = ((size_t)l * t1 + u) * t2 + v;
          double chi = pbar[q][pi] + p->cpsi[q][pi] * *fa_at(eb, a, n, u, v);
          *fa_at(hb, a, n, u, v) += p->ae[l] * chi;
          *fa_at(hb, a, n - 1, u, v) -= p->ae[l] * chi;
          pbar[q][pi] = p->be[l] * chi;
        }
    }
  }
  (void)s;
}

#define AEX(i, j, k) AT(ad->f[XO_EX], i, j, k)
#define AEY(i, j, k) AT(ad->f[XO_EY], i, j, k)
#define AEZ(i, j, k) AT(ad->f[XO_EZ], i, j, k)
#define AHX(i, j, k) AT(ad->f[XO_HX], i, j, k)
#define AHY(i, j, k) AT(ad->f[XO_HY], i, j, k)
#define AHZ(i, j, k) AT(ad->f[XO_HZ], i, j, k)

/* transpose of update_h (all three components over i<nx, j<ny, k<nz) */
static void adj_update_h(xo_sim *s, adj_t *ad) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  a3 *c = s->c;
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j)
      for (size_t k = 0; k < nz; ++k) {
        double ax = AHX(i, j, k), ay = AHY(i, j, k), az = AHZ(i, j, k);
        double w;
        w = AT(c[CHXEY], i, j, k) * ax; AEY(i, j, k + 1) += w; AEY(i, j, k) -= w;
        w = AT(c[CHXEZ], i, j, k) * ax; AEZ(i, j + 1, k) += w; AEZ(i, j, k) -= w;
        w = AT(c[CHYEZ], i, j, k) * ay; AEZ(i + 1, j, k) += w; AEZ(i, j, k) -= w;
        w = AT(c[CHYEX], i, j, k) * ay; AEX(i, j, k + 1) += w; AEX(i, j, k) -= w;
        w = AT(c[CHZEX], i, j, k) * az; AEX(i, j + 1, k) += w; AEX(i, j, k) -= w;
        w = AT(c[CHZEY], i, j, k) * az; AEY(i + 1, j, k) += w; AEY(i, j, k) -= w;
        AHX(i, j, k) = AT(c[CHXH], i, j, k) * ax;
        AHY(i, j, k) = AT(c[CHYH], i, j, k) * ay;
        AHZ(i, j, k) = AT(c[CHZH], i, j, k) * az;
      }
}

/* transpose of one Drude cell's update (update_e_drude):
     E' = ca E + cb (curl - h kp1 J),  J' = k J + beta (E' + E)
   undone last op first.  ebar, jbar: adjoints of E' and J' in, of E and J
   out; returns t, the adjoint of E' after the J' transpose, whose curl
   transpose the caller scatters with weight cb (curl = -dHa + dHb). */
static double adj_drude_cell(const xo_sim *s, int q, double *ebar, double *jbar) {
  double kk = s->dk[q], beta = s->dbeta[q], ca = s->dca[q], cb = s->dcb[q];
  double hk = 0.5 * s->dl * (1 + kk);
  double jb = *jbar;
  double t = *ebar + beta * jb;   /* J' = ... + beta E' */
  *ebar = beta * jb + ca * t;     /* J' = ... + beta E; E' = ca E + ... */
  *jbar = kk * jb - cb * hk * t;  /* J' = k J; E' = ... - cb h kp1 J */
  return t;
}

/* transpose of update_e (each component over its Yee range) */
static void adj_update_e(xo_sim *s, adj_t *ad) {
  size_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  a3 *c = s->c;
  double w;
  const int *dm = s->dmask;
#define DRUDE_AT(i, j, k) (dm ? dm[((i) * ny + (j)) * nz + (k)] : 0)
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 1; j < ny; ++j)
      for (size_t k = 1; k < nz; ++k) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          double cb = s->dcb[d - 1];
          double t = adj_drude_cell(s, d - 1, &AEX(i, j, k), &AT(ad->J[0], i, j, k));
          w = -cb * t; AHY(i, j, k) += w; AHY(i, j, k - 1) -= w;
          w = cb * t; AHZ(i, j, k) += w; AHZ(i, j - 1, k) -= w;
          continue;
        }
        double a = AEX(i, j, k);
        w = AT(c[CEXHY], i, j, k) * a; AHY(i, j, k) += w; AHY(i, j, k - 1) -= w;
        w = AT(c[CEXHZ], i, j, k) * a; AHZ(i, j, k) += w; AHZ(i, j - 1, k) -= w;
        AEX(i, j, k) = AT(c[CEXE], i, j, k) * a;
      }
  for (size_t j = 0; j < ny; ++j)
    for (size_t k = 1; k < nz; ++k)
      for (size_t i = 1; i < nx; ++i) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          double cb = s->dcb[d - 1];
          double t = adj_drude_cell(s, d - 1, &AEY(i, j, k), &AT(ad->J[1], i, j, k));
          w = -cb * t; AHZ(i, j, k) += w; AHZ(i - 1, j, k) -= w;
          w = cb * t; AHX(i, j, k) += w; AHX(i, j, k - 1) -= w;
          continue;
        }
        double a = AEY(i, j, k);
        w = AT(c[CEYHX], i, j, k) * a; AHX(i, j, k) += w; AHX(i, j, k - 1) -= w;
        w = AT(c[CEYHZ], i, j, k) * a; AHZ(i, j, k) += w; AHZ(i - 1, j, k) -= w;
        AEY(i, j, k) = AT(c[CEYE], i, j, k) * a;
      }
  for (size_t k = 0; k < nz; ++k)
    for (size_t i = 1; i < nx; ++i)
      for (size_t j = 1; j < ny; ++j) {
        int d = DRUDE_AT(i, j, k);
        if (d) {
          double cb = s->dcb[d - 1];
          double t = adj_drude_cell(s, d - 1, &AEZ(i, j, k), &AT(ad->J[2], i, j, k));
          w = -cb * t; AHX(i, j, k) += w; AHX(i, j - 1, k) -= w;
          w = cb * t; AHY(i, j, k) += w; AHY(i - 1, j, k) -= w;
          continue;
        }
        double a = AEZ(i, j, k);
        w = AT(c[CEZHX], i, j, k) * a; AHX(i, j, k) += w; AHX(i, j - 1, k) -= w;
        w = AT(c[CEZHY], i, j, k) * a; AHY(i, j, k) += w; AHY(i - 1, j, k) -= w;
        AEZ(i, j, k) = AT(c[CEZE], i, j, k) * a;
      }
#undef DRUDE_AT
}

/* the H half of reverse step: CPML-H^T (faces reversed), update_h^T */
static void adj_half_h(xo_sim *s, adj_t *ad) {
  for (int b = s->npml - 1; b >= 0; --b) adj_pml_h(s, &s->pml[b], ad->psi + 4 * b, ad);
  adj_update_h(s, ad);
}
/* CPML-E^T (faces reversed); update_e^T follows separately */
static void adj_pml_e_all(xo_sim *s, adj_t *ad) {
  for (int b = s->npml - 1; b >= 0; --b) adj_pml_e(s, &s->pml[b], ad->psi + 4 * b, ad);
}

int xo_sim_set_design(xo_sim *s, const size_t box[6], const float *eps_r) {
  if (s->inited) FAIL("simulation already initialised");
  if (box[1] <= box[0] || box[3] <= box[2] || box[5] <= box[4]) FAIL("empty design box");
  size_t nd = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
  free(s->design_eps);
  s->design_eps = (float *)malloc(nd * sizeof(float));
  if (!s->design_eps) FAIL("out of memory");
  memcpy(s->design_eps, eps_r, nd * sizeof(float));
  memcpy(s->dbox, box, sizeof s->dbox);
  s->has_design = 1;
  return 0;
}

int xo_sim_set_dft_plane(xo_sim *s, int comp, int axis, size_t index, size_t lo1, size_t hi1,
                         size_t lo2, size_t hi2, double freq) {
  if (s->inited) FAIL("simulation already initialised");
  if (comp < 0 || comp > 5 || axis < 0 || axis > 2) FAIL("bad DFT plane");
  s->dft_comp = comp; s->dft_axis = axis; s->dft_index = index;
  s->dft_lo1 = lo1; s->dft_hi1 = hi1; s->dft_lo2 = lo2; s->dft_hi2 = hi2;
  s->dft_freq = freq;
  s->has_dft = 1;
  return 0;
}

int xo_sim_get_dft(const xo_sim *s, double *out) {
  if (!s->inited || !s->has_dft) FAIL("no DFT plane");
  size_t n = (s->dft_hi1 - s->dft_lo1) * (s->dft_hi2 - s->dft_lo2);
  memcpy(out, s->dft_acc, n * sizeof(double complex));
  return 0;
}

static double dft_objective(const xo_sim *s) {
  size_t n = (s->dft_hi1 - s->dft_lo1) * (s->dft_hi2 - s->dft_lo2);
  double J = 0;
  for (size_t q = 0; q < n; ++q) J += creal(s->dft_acc[q]) * creal(s->dft_acc[q]) +
                                     cimag(s->dft_acc[q]) * cimag(s->dft_acc[q]);
  return J;
}

/* dJ/d eps_r of every design cell (C order over the box) and J, for
   J = sum |F|^2 over the DFT plane after total_steps steps.  Runs the
   forward loop from step 0, keeping the design-box E fields of every step,
   then the reverse loop. */
static int only_drude_dispersion(const xo_sim *s) {
  for (int q = 0; q < s->nobj; ++q)
    if (s->objs[q].drude > 1) return 0;
  return 1;
}

int xo_sim_grad(xo_sim *s, double *grad, double *J_out) {
  if (!s->inited || s->cur != 0) FAIL("grad needs a fresh simulation");
  if (!only_drude_dispersion(s)) FAIL("the adjoint covers Drude media only, not Debye or Lorentz");
  if (!s->has_design || !s->has_dft) FAIL("grad needs a design box and a DFT plane");
  const size_t *b = s->dbox;
  size_t nd = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]), T = s->T;
  double *hist = (double *)calloc((T + 1) * nd * 3, sizeof(double));
  if (!hist) FAIL("out of memory");
#define HIST(t, q, c) hist[((size_t)(t) * nd + (q)) * 3 + (c)]
  for (size_t t = 0; t < T; ++t) {
    if (xo_sim_step(s, 1)) { free(hist); return -1; }
    size_t q = 0;
    for (size_t i = b[0]; i < b[1]; ++i)
      for (size_t j = b[2]; j < b[3]; ++j)
        for (size_t k = b[4]; k < b[5]; ++k, ++q)
          for (int c = 0; c < 3; ++c) HIST(t + 1, q, c) = AT(s->f[c], i, j, k);
  }
  *J_out = dft_objective(s);
  adj_t ad;
  if (adj_alloc(s, &ad)) { free(hist); return -1; }
  memset(grad, 0, nd * sizeof(double));
  size_t n2 = s->dft_hi2 - s->dft_lo2;
  for (size_t tt = T; tt-- > 0;) {
    /* DFT tap^T: dJ/dE = 2 Re(conj(F) w_t) on the plane */
    double complex w = dft_weight(s, tt);
    for (size_t u = s->dft_lo1; u < s->dft_hi1; ++u)
      for (size_t v = s->dft_lo2; v < s->dft_hi2; ++v) {
        double complex F = s->dft_acc[(u - s->dft_lo1) * n2 + (v - s->dft_lo2)];
        *dft_cell(&ad.f[s->dft_comp], s->dft_axis, s->dft_index, u, v) += 2 * creal(conj(F) * w);
      }
    adj_half_h(s, &ad);
    adj_pml_e_all(s, &ad);
    /* the design cells' E coefficients depend on eps: with sigma = 0,
       d E'/d eps = -(E' - E)/eps, and d eps/d eps_r = eps_0 */
    size_t q = 0;
    for (size_t i = b[0]; i < b[1]; ++i)
      for (size_t j = b[2]; j < b[3]; ++j)
        for (size_t k = b[4]; k < b[5]; ++k, ++q) {
          double er = (double)s->design_eps[q], acc = 0;
          for (int c = 0; c < 3; ++c)
            acc += AT(ad.f[c], i, j, k) * (HIST(tt + 1, q, c) - HIST(tt, q, c));
          grad[q] -= acc / er;
        }
    adj_update_e(s, &ad);
  }
#undef HIST
  adj_free(s, &ad);
  free(hist);
  return 0;
}

/* Reverse-run `nsteps` steps of the field map from a given adjoint state
   (the six adjoint fields packed in component order, then, when there are
   Drude cells, the three J adjoints in E shapes; psi adjoints zero), with no
   objective; the result overwrites `fields`.  For the identity
   <L x, y> = <x, L^T y> of the step map L. */
int xo_sim_adjoint_run(xo_sim *s, size_t nsteps, double *fields) {
  if (!s->inited) FAIL("not initialised");
  if (!only_drude_dispersion(s)) FAIL("the adjoint covers Drude media only, not Debye or Lorentz");
  adj_t ad;
  if (adj_alloc(s, &ad)) return -1;
  a3 *st[9];
  int nst = s->has_drude ? 9 : 6;
  for (int q = 0; q < 6; ++q) st[q] = &ad.f[q];
  for (int q = 0; q < 3; ++q) st[6 + q] = &ad.J[q];
  size_t off = 0;
  for (int q = 0; q < nst; ++q) {
    size_t n = st[q]->n0 * st[q]->n1 * st[q]->n2;
    memcpy(st[q]->d, fields + off, n * sizeof(double));
    off += n;
  }
  for (size_t t = 0; t < nsteps; ++t) {
    adj_half_h(s, &ad);
    adj_pml_e_all(s, &ad);
    adj_update_e(s, &ad);
  }
  off = 0;
  for (int q = 0; q < nst; ++q) {
    size_t n = st[q]->n0 * st[q]->n1 * st[q]->n2;
    memcpy(fields + off, st[q]->d, n * sizeof(double));
    off += n;
  }
  adj_free(s, &ad);
  return 0;
}
