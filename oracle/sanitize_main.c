/*
 * sanitize_main.c -- TEST INFRASTRUCTURE ONLY: drives the FP64 oracle through
 * small scenes covering every builder call, so that a build of
 * xfdtd_oracle.c with AddressSanitizer and UndefinedBehaviorSanitizer (the
 * Makefile's `sanitize` target) checks each of its loops for out-of-bounds
 * accesses, use after free, leaks, signed overflow and misaligned loads.
 *
 * Scenes (all a few thousand cells, a few dozen steps):
 *   1. 3-D scatter: dielectric sphere, CPML with kappa > 1 on six faces,
 *      oblique TFSF3D, NffftFd at two frequencies, a probe, far field and
 *      every read-back accessor;
 *   2. media and circuits: Drude sphere, Debye and Lorentz cubes, an
 *      ObjectPlane, a current source, voltage source, resistor, capacitor and
 *      inductor, voltage and current line monitors;
 *   3. adjoint: design box, DFT-plane objective, gradient through a Drude
 *      block and the reverse run on packed fields;
 *   4. broadband NF-FF (NffftBroadBand) behind a TFSF3D source;
 *   5. 2-D TMz: PEC cylinder, TFSF2D, a 2-D broadband surface.
 * Exits 0 and prints "sanitized oracle scenes: ok" when every call succeeds
 * and every field stays finite; a sanitizer report ends the run non-zero.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xfdtd_oracle.h"

#define CHK(x)                                                                 \
  do {                                                                         \
    if ((x) != 0) {                                                            \
      fprintf(stderr, "%s:%d %s failed: %s\n", __FILE__, __LINE__, #x,        \
              xo_last_error());                                                \
      exit(1);                                                                 \
    }                                                                          \
  } while (0)

static double *buf(size_t n) {
  double *p = calloc(n ? n : 1, sizeof(double));
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(1);
  }
  return p;
}

static double checksum(const double *v, size_t n, const char *what) {
  double s = 0.0;
  for (size_t q = 0; q < n; ++q) {
    if (!isfinite(v[q])) {
      fprintf(stderr, "%s: non-finite value at %zu\n", what, q);
      exit(1);
    }
    s += fabs(v[q]);
  }
  return s;
}

static double fields_checksum(xo_sim *s, const char *scene) {
  double total = 0.0;
  for (int c = XO_EX; c <= XO_HZ; ++c) {
    size_t n = xo_sim_field_size(s, c);
    double *f = buf(n);
    CHK(xo_sim_get_field(s, c, f));
    total += checksum(f, n, scene);
    free(f);
  }
  return total;
}

static size_t max_field(const xo_sim *s) {
  size_t m = 0;
  for (int c = XO_EX; c <= XO_HZ; ++c) {
    size_t n = xo_sim_field_size(s, c);
    if (n > m) m = n;
  }
  return m;
}

static void add_pml6(xo_sim *s, int n, double kappa) {
  for (int o = XO_XN; o <= XO_ZP; ++o) CHK(xo_sim_add_pml(s, o, n, 3, 1.0, 0.0, 0.05, kappa));
}

static xo_sim *create(double dl, float cfl) {
  xo_sim *s = xo_sim_create(dl, cfl);
  if (!s) {
    fprintf(stderr, "xo_sim_create failed\n");
    exit(1);
  }
  return s;
}

static void scene_scatter(void) {
  const double dl = 5e-3;
  const size_t T = 40;
  xo_sim *s = create(dl, 0.99f);
  const double o[3] = {0, 0, 0}, sz[3] = {24 * dl, 22 * dl, 20 * dl};
  CHK(xo_sim_add_cube(s, o, sz, 1.0f, 1.0f, 0.0, 0.0));
  const double c[3] = {12 * dl, 11 * dl, 10 * dl};
  CHK(xo_sim_add_sphere(s, c, 3.5 * dl, 3.0f, 1.0f, 0.0, 0.0));
  add_pml6(s, 4, 2.0);
  const size_t dist[3] = {6, 6, 6};
  const double tau = 1e-10;
  CHK(xo_sim_set_tfsf(s, dist, 1.0, 0.3, 0.4, 0.2, 1.0, tau, 3 * tau));
  const size_t nd[3] = {5, 5, 5};
  const double fr[2] = {1e9, 3e9}, th[3] = {0.1, 1.5, 3.0}, ph[2] = {0.0, 1.2};
  CHK(xo_sim_set_nffft(s, nd, fr, 2, th, 3, ph, 2));
  CHK(xo_sim_add_probe(s, XO_EX, 12, 11, 10));
  CHK(xo_sim_add_probe(s, XO_HZ, 10, 12, 9));
  CHK(xo_sim_init(s, T));
  CHK(xo_sim_step(s, T / 2));
  CHK(xo_sim_step(s, T - T / 2));

  double *co = buf(max_field(s));
  for (int w = 0; w < 18; ++w) CHK(xo_sim_get_coeff(s, w, co));
  free(co);
  double *prof = buf(6 * 256);
  for (int b = 0; b < 6; ++b) CHK(xo_sim_get_pml_profile(s, b, prof));
  free(prof);
  size_t L = xo_sim_tfsf_aux_size(s);
  double *ei = buf(L), *hi = buf(L);
  CHK(xo_sim_get_tfsf_aux(s, ei, hi));
  checksum(ei, L, "e_inc");
  free(ei);
  free(hi);
  double *wf = buf(T);
  CHK(xo_sim_get_waveform(s, wf));
  free(wf);
  size_t nn = xo_sim_nffft_size(s);
  double *acc = buf(2 * nn);
  CHK(xo_sim_get_nffft(s, acc));
  checksum(acc, 2 * nn, "nffft");
  free(acc);
  double *ff = buf(10 * 2 * 12);
  CHK(xo_sim_farfield(s, ff));
  checksum(ff, 10 * 2 * 12, "far field");
  free(ff);
  double *pr = buf(2 * T);
  CHK(xo_sim_get_probes(s, pr));
  free(pr);
  /* a field written back is read back unchanged */
  size_t n = xo_sim_field_size(s, XO_EY);
  double *f = buf(n), *g = buf(n);
  CHK(xo_sim_get_field(s, XO_EY, f));
  CHK(xo_sim_set_field(s, XO_EY, f));
  CHK(xo_sim_get_field(s, XO_EY, g));
  if (memcmp(f, g, n * sizeof(double)) != 0) {
    fprintf(stderr, "set_field / get_field round trip differs\n");
    exit(1);
  }
  free(f);
  free(g);
  printf("scatter: fields %.6e, step %zu\n", fields_checksum(s, "scatter"), xo_sim_current_step(s));
  xo_sim_destroy(s);
}

static void scene_media(void) {
  const double dl = 5e-3;
  const size_t T = 30;
  xo_sim *s = create(dl, 0.99f);
  const double o[3] = {0, 0, 0}, sz[3] = {26 * dl, 24 * dl, 24 * dl};
  CHK(xo_sim_add_cube(s, o, sz, 1.0f, 1.0f, 0.0, 0.0));
  const double c[3] = {9 * dl, 9 * dl, 12 * dl};
  CHK(xo_sim_add_sphere(s, c, 2.5 * dl, 1.0f, 1.0f, 0.0, 0.0));
  CHK(xo_sim_set_drude(s, 1.0, 2e10, 1e9));
  const double db[3] = {15 * dl, 6 * dl, 6 * dl}, dbs[3] = {3 * dl, 3 * dl, 3 * dl};
  CHK(xo_sim_add_cube(s, db, dbs, 1.0f, 1.0f, 0.0, 0.0));
  CHK(xo_sim_set_debye(s, 4.0, 2.0, 1e-11));
  const double lb[3] = {15 * dl, 14 * dl, 14 * dl};
  CHK(xo_sim_add_cube(s, lb, dbs, 1.0f, 1.0f, 0.0, 0.0));
  CHK(xo_sim_set_lorentz(s, 3.0, 1.5, 2e10, 1e9));
  const double po[3] = {6 * dl, 15 * dl, 6 * dl}, ps[3] = {4 * dl, 0.0, 4 * dl};
  CHK(xo_sim_add_plane(s, po, ps, 1.0f, 1.0f, 1e3, 0.0));
  add_pml6(s, 4, 1.0);
  const double tau = 1e-10, gw[4] = {1.0, tau, 3 * tau, 0.0}, sw[4] = {1.0, 2e9, 0.1, 0.0},
               cw[4] = {1.0, tau, 3 * tau, 3e9};
  const double cs[3] = {12 * dl, 12 * dl, 8 * dl}, css[3] = {dl, dl, 2 * dl};
  CHK(xo_sim_add_lumped(s, 0, cs, css, XO_ZP, 50.0, 0, gw));
  const double vs[3] = {18 * dl, 18 * dl, 8 * dl};
  CHK(xo_sim_add_lumped(s, 1, vs, css, XO_ZP, 50.0, 1, sw));
  const double rs[3] = {8 * dl, 18 * dl, 8 * dl}, rss[3] = {2 * dl, dl, dl};
  CHK(xo_sim_add_lumped(s, 2, rs, rss, 0, 100.0, 0, gw));
  const double ca[3] = {18 * dl, 8 * dl, 16 * dl}, cas[3] = {dl, 2 * dl, dl};
  CHK(xo_sim_add_lumped(s, 3, ca, cas, 1, 1e-12, 0, gw));
  const double ind[3] = {12 * dl, 6 * dl, 16 * dl};
  CHK(xo_sim_add_lumped(s, 4, ind, css, 2, 1e-9, 2, cw));
  const double ms[3] = {12 * dl, 12 * dl, 8 * dl};
  if (xo_sim_add_line_monitor(s, 0, ms, css, XO_ZP) < 0 || xo_sim_add_line_monitor(s, 1, ms, css, XO_ZP) < 0) {
    fprintf(stderr, "line monitor: %s\n", xo_last_error());
    exit(1);
  }
  CHK(xo_sim_init(s, T));
  CHK(xo_sim_step(s, T));
  double *lm = buf(2 * T);
  CHK(xo_sim_get_line_monitors(s, lm));
  checksum(lm, 2 * T, "line monitors");
  free(lm);
  double *w = buf(T);
  CHK(xo_sim_get_inductor_current(s, 4, w));
  CHK(xo_sim_get_lumped_wave(s, 0, w));
  free(w);
  size_t n = max_field(s);
  double *j = buf(n);
  for (int comp = 0; comp < 3; ++comp) {
    CHK(xo_sim_get_drude_j(s, comp, j));
    CHK(xo_sim_set_drude_j(s, comp, j));
  }
  free(j);
  CHK(xo_sim_step(s, 0));
  printf("media: fields %.6e\n", fields_checksum(s, "media"));
  xo_sim_destroy(s);
}

static void scene_adjoint(void) {
  const double dl = 5e-3;
  const size_t T = 36;
  xo_sim *s = create(dl, 0.99f);
  const double o[3] = {0, 0, 0}, sz[3] = {20 * dl, 20 * dl, 20 * dl};
  CHK(xo_sim_add_cube(s, o, sz, 1.0f, 1.0f, 0.0, 0.0));
  const double xb[3] = {12 * dl, 12 * dl, 12 * dl}, xs[3] = {2 * dl, 2 * dl, 2 * dl};
  CHK(xo_sim_add_cube(s, xb, xs, 1.0f, 1.0f, 0.0, 0.0));
  CHK(xo_sim_set_drude(s, 1.0, 2e10, 1e9));
  add_pml6(s, 4, 1.0);
  const double cs[3] = {7 * dl, 7 * dl, 9 * dl}, css[3] = {dl, dl, 2 * dl};
  const double tau = 1e-10;
  CHK(xo_sim_add_current_source(s, cs, css, XO_ZP, 50.0, 1.0, tau, 3 * tau));
  const size_t box[6] = {8, 11, 8, 11, 8, 11};
  float eps[27];
  for (int q = 0; q < 27; ++q) eps[q] = 1.0f + 0.1f * (float)(q % 7);
  CHK(xo_sim_set_design(s, box, eps));
  CHK(xo_sim_set_dft_plane(s, XO_EZ, 0, 13, 5, 14, 5, 14, 3e9));
  CHK(xo_sim_init(s, T));
  double grad[27], J = 0.0;
  CHK(xo_sim_grad(s, grad, &J));
  checksum(grad, 27, "gradient");
  double *d = buf(2 * 81);
  CHK(xo_sim_get_dft(s, d));
  free(d);
  size_t packed = 0;
  for (int c = XO_EX; c <= XO_HZ; ++c) packed += xo_sim_field_size(s, c);
  double *adj = buf(4 * packed);
  adj[packed / 2] = 1.0;
  CHK(xo_sim_adjoint_run(s, 5, adj));
  free(adj);
  printf("adjoint: J %.6e, |grad| %.6e\n", J, checksum(grad, 27, "gradient"));
  xo_sim_destroy(s);
}

static void scene_broadband(void) {
  const double dl = 5e-3;
  const size_t T = 30;
  xo_sim *s = create(dl, 0.99f);
  const double o[3] = {0, 0, 0}, sz[3] = {22 * dl, 22 * dl, 22 * dl};
  CHK(xo_sim_add_cube(s, o, sz, 1.0f, 1.0f, 0.0, 0.0));
  const double c[3] = {11 * dl, 11 * dl, 11 * dl};
  CHK(xo_sim_add_sphere(s, c, 3 * dl, 1.0f, 1.0f, 1e24, 0.0));
  add_pml6(s, 4, 1.0);
  const size_t dist[3] = {6, 6, 6}, nd[3] = {5, 5, 5};
  const double tau = 1e-10;
  CHK(xo_sim_set_tfsf(s, dist, 1.0, 0.0, 0.0, 0.0, 1.0, tau, 3 * tau));
  CHK(xo_sim_set_nffft_broadband(s, nd, 3.14159265358979, 0.0));
  CHK(xo_sim_init(s, T));
  CHK(xo_sim_step(s, T));
  const size_t nb = 2 * T - 1;
  double *f = buf(nb), *k = buf(nb), *sums = buf(12 * nb);
  CHK(xo_sim_bb_frequencies(s, f, k));
  CHK(xo_sim_bb_sums(s, NULL, sums));
  checksum(sums, 12 * nb, "broadband sums");
  size_t ns = xo_sim_bb_samples_size(s);
  double *smp = buf(ns);
  CHK(xo_sim_get_bb_samples(s, smp));
  free(smp);
  printf("broadband: sums %.6e\n", checksum(sums, 12 * nb, "broadband sums"));
  free(f);
  free(k);
  free(sums);
  xo_sim_destroy(s);
}

static void scene_tmz(void) {
  const double dl = 1e-3;
  const size_t T = 60;
  xo_sim *s = create(dl, 0.8f);
  const double o[3] = {0, 0, 0}, sz[3] = {40 * dl, 36 * dl, 0.0};
  CHK(xo_sim_add_cube(s, o, sz, 1.0f, 1.0f, 0.0, 0.0));
  const double c[3] = {20 * dl, 18 * dl, 0.0};
  CHK(xo_sim_add_cylinder(s, 2, c, 5 * dl, 0.0, 1.0f, 1.0f, 1e24, 0.0));
  for (int q = XO_XN; q <= XO_YP; ++q) CHK(xo_sim_add_pml(s, q, 5, 3, 1.0, 0.0, 0.05, 1.0));
  const double tau = 1.7 / 8e9, wp[4] = {1.0, tau, 0.8 * tau, 12e9};
  CHK(xo_sim_set_tfsf2d(s, 10, 10, 0.785398163397448, 1.0, 2, wp));
  const size_t nd[3] = {8, 8, 0};
  CHK(xo_sim_set_nffft_broadband(s, nd, 1.5707963267949, 3.92699081698724));
  CHK(xo_sim_init(s, T));
  CHK(xo_sim_step(s, T));
  const size_t nb = 2 * T - 1;
  double *sums = buf(12 * nb);
  CHK(xo_sim_bb_sums(s, NULL, sums));
  printf("tmz: fields %.6e, sums %.6e\n", fields_checksum(s, "tmz"), checksum(sums, 12 * nb, "tmz sums"));
  free(sums);
  xo_sim_destroy(s);
}

int main(void) {
  xo_set_threads(2); /* returns the thread count */
  scene_scatter();
  scene_media();
  scene_adjoint();
  scene_broadband();
  scene_tmz();
  printf("sanitized oracle scenes: ok\n");
  return 0;
}
