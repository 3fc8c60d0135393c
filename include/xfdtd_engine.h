/*
 * xfdtd_engine.h -- C-ABI of the MI355X (gfx950) FDTD field engine.
 *
 * This is the device boundary of the drop-in: the C++ host classes
 * (include/xfdtd/xfdtd.hpp) do the reference's set-up on the CPU and hand the
 * per-step work to these entry points.  Every call replaces a piece of the
 * reference's run loop, cited by file:line under the reference's src/:
 *
 *   xf_engine_step          Simulation::run loop body        simulation/simulation_update.cpp:9-36
 *                           (updateE 130-197, correctE 282-290, updateBoundaryE 124-128,
 *                            updateH 50-108, updateTFSFIncidentField 42-48, correctH 292-300,
 *                            updateBoundaryH 118-122, updateNFFFT 207-213, updateMonitor 215-219)
 *   xf_engine_set_materials FDTDBasicCoff arrays             fdtd_basic_coff/fdtd_basic_coff.cpp:31-122
 *                           + Object::init masks             object/object.cpp:53-73
 *   xf_engine_set_cpml      PML::init                        boundary/perfect_match_layer.cpp:22-96, 133-432
 *   xf_engine_set_tfsf      TFSF3D::init                     tfsf/tfsf_3d.cpp:33-151
 *   xf_engine_set_nffft     NffftFd::init / initDFT          nffft/nffft_fd.cpp:17-106
 *   xf_engine_add_current_source  CurrentSource::correctE    lumped_element/current_source.cpp:140-163
 *                           (also VoltageSource::correctE   lumped_element/voltage_source.cpp:168-193)
 *   xf_engine_add_inductor  Inductor::correctE               lumped_element/inductor.cpp:115-146
 *   xf_engine_add_line_monitor  Voltage/CurrentMonitor::update  monitor/voltage_monitor.cpp:63-93,
 *                                                            monitor/current_monitor.cpp:66-122
 *   xf_engine_add_probe     TimeDomainFieldMonitor sample    monitor/field_monitor.cpp:29-122
 *   xf_engine_get_field     EMF::getEx..getHz                include/simulation/simulation.h:92-97
 *   xf_engine_get_nffft     NffftFd accumulators             nffft/nffft_fd.cpp:126-202
 *   xf_engine_farfield_sums NffftFd::calculateFarfieldX/Y/Z  nffft/nffft_fd.cpp:423-556
 *   xf_engine_set_dispersive  Drude/Debye/LorentzMedium::updateE  material/dispersive_material.cpp:7-275
 *   xf_engine_set_nffft_broadband  NffftBroadBand::init / update  nffft/nffft_broadband.cpp:30-155
 *   xf_engine_broadband_sums       NffftBroadBand::calculateLNXYZ nffft/nffft_broadband.cpp:348-466
 *
 * Conventions: status codes, never exceptions; the engine owns device memory,
 * the caller owns every host buffer; one HIP stream per engine; an engine is
 * not thread-safe.  Field arrays crossing the boundary are FP64, C order, in
 * the reference's shapes (simulation/simulation_init.cpp:28-35):
 *   Ex (nx,ny+1,nz+1)  Ey (nx+1,ny,nz+1)  Ez (nx+1,ny+1,nz)
 *   Hx (nx+1,ny,nz)    Hy (nx,ny+1,nz)    Hz (nx,ny,nz+1)
 */
#ifndef XFDTD_ENGINE_H
#define XFDTD_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xf_engine xf_engine;

typedef enum {
  XF_OK = 0,
  XF_EINVAL = 1,  /* bad argument or shape */
  XF_ENOMEM = 2,  /* device allocation failed */
  XF_EDEVICE = 3, /* HIP runtime error, or no usable gfx950 device */
  XF_ESTATE = 4   /* call out of order (e.g. step before set-up) */
} xf_status;

enum { XF_EX = 0, XF_EY, XF_EZ, XF_HX, XF_HY, XF_HZ };

/* curl-coefficient tables, each indexed [material][index along the
   derivative axis].  The reference's 18 per-cell arrays
   (fdtd_basic_coff.cpp:93-121) are functions of the cell's material and of
   one node size along the derivative axis, so 6 self terms per material plus
   these 12 one-dimensional tables reproduce them exactly. */
enum {
  XF_CB_EXHY = 0, /* cexhy, z: -(2dt/hz[k])/(2eps+sigma dt)   */
  XF_CB_EXHZ,     /* cexhz, y:  (2dt/hy[j])/(2eps+sigma dt)   */
  XF_CB_EYHZ,     /* ceyhz, x: -(2dt/hx[i])/(...)             */
  XF_CB_EYHX,     /* ceyhx, z:  (2dt/hz[k])/(...)             */
  XF_CB_EZHX,     /* cezhx, y: -(2dt/hy[j])/(...)             */
  XF_CB_EZHY,     /* cezhy, x:  (2dt/hx[i])/(...)             */
  XF_CB_HXEY,     /* chxey, z:  (2dt/ez[k])/(2mu+sigma_m dt)  */
  XF_CB_HXEZ,     /* chxez, y: -(2dt/ey[j])/(...)             */
  XF_CB_HYEZ,     /* chyez, x:  (2dt/ex[i])/(...)             */
  XF_CB_HYEX,     /* chyex, z: -(2dt/ez[k])/(...)             */
  XF_CB_HZEX,     /* chzex, y:  (2dt/ey[j])/(...)             */
  XF_CB_HZEY,     /* chzey, x: -(2dt/ex[i])/(...)             */
  XF_CB_COUNT
};

typedef struct {
  int64_t nx, ny, nz; /* cell counts; nz = 1: a 2-D (TMz) grid, stepped by the
                          same Yee updates (simulation_update.cpp:157-168) */
  int precision;      /* 64: FP64 fields (parity mode); 32: FP32 fields */
  int n_materials;    /* 1..255 */
} xf_grid_desc;

xf_status xf_engine_create(const xf_grid_desc *g, xf_engine **out);
void xf_engine_destroy(xf_engine *e);
/* thread-local message of the last failing call */
const char *xf_last_error(void);

/* index: material per cell over (nx+1)*(ny+1)*(nz+1), C order; the entries
   at i==nx, j==ny or k==nz are the padded default cells of the reference's
   staggered arrays (fdtd_basic_coff.cpp:313-339).
   ca: [6][n_materials] self coefficients cexe,ceye,ceze,chxh,chyh,chzh.
   cb: XF_CB_COUNT tables, table t of length n_materials*(n_axis(t)+1), with
       the CPML kappa division already applied (perfect_match_layer.cpp:187-208).
   cpsi: same shapes, the coefficient times dl captured before that division
       (the reference's _c_psi_* arrays); read only inside CPML slabs. */
xf_status xf_engine_set_materials(xf_engine *e, const uint8_t *index,
                                  const double *ca, const double *const *cb,
                                  const double *const *cpsi);

/* CPML along `axis` (0 x, 1 y, 2 z): d_n layers on the N face, d_p on the P
   face (0 = no face).  b_e/a_e are indexed by the E-node index along the
   axis (n in [1,d_n] and [N-d_p, N-1]); b_m/a_m by the H index
   (n in [0,d_n) and [N-d_p, N)).  Length N+1 each.  Faces are applied in
   x, y, z order, the XN..ZP insertion order of every reference example. */
xf_status xf_engine_set_cpml(xf_engine *e, int axis, int d_n, int d_p,
                             const double *b_e, const double *a_e,
                             const double *b_m, const double *a_m);

typedef struct {
  int64_t li, ri, lj, rj, lk, rk; /* total-field box: tfsf/tfsf.cpp:45-49 */
  int64_t aux_size;               /* L: tfsf/tfsf_3d.cpp:53-57 */
  const double *proj_full[3];     /* lengths (r-l)+1 */
  const double *proj_half[3];     /* lengths (r-l)+2 */
  double transform_e[3], transform_h[3];
  double ceie, ceihi, chih, chiei; /* 1-D grid coefficients */
  double abc_c1, abc_c2;           /* tfsf/tfsf_3d.cpp:161-165 */
  double ca[3], cb[3];             /* dt/(eps0 d), dt/(mu0 d) per axis */
  const double *source;            /* e_0 * g(n dt), n < total_steps */
  int64_t total_steps;
  /* 1: TFSF2D (tfsf/tfsf_2d.cpp:26-259) on a grid with nz = 1: the box is
     [li,ri] x [lj,rj] x [0,1), only Ez, Hx and Hy take corrections and the
     incident projections have no z term */
  int64_t two_d;
} xf_tfsf_desc;
xf_status xf_engine_set_tfsf(xf_engine *e, const xf_tfsf_desc *d);

typedef struct {
  int64_t li, ri, lj, rj, lk, rk; /* output box: nffft/nffft.cpp:29-33 */
  int nf;
  int64_t total_steps;
  const double *tj; /* interleaved complex [nf][total_steps], (t+1/2) dt */
  const double *tm; /* interleaved complex [nf][total_steps], t dt */
} xf_nffft_desc;
xf_status xf_engine_set_nffft(xf_engine *e, const xf_nffft_desc *d);
/* size (complex entries) of the accumulator block returned by get_nffft */
int64_t xf_engine_nffft_size(const xf_engine *e);
/* 24 accumulators, interleaved complex, in the order
   x pair: my_xn,mz_xn,jy_xn,jz_xn,my_xp,mz_xp,jy_xp,jz_xp  [nf][nyb][nzb]
   y pair: mz_yn,mx_yn,jz_yn,jx_yn,mz_yp,mx_yp,jz_yp,jx_yp  [nf][nzb][nxb]
   z pair: mx_zn,my_zn,jx_zn,jy_zn,mx_zp,my_zp,jx_zp,jy_zp  [nf][nxb][nyb] */
xf_status xf_engine_get_nffft(xf_engine *e, double *out);
/* Far-field face sums for every direction (f, theta, phi), FP64 on device:
   out [nf][ntheta][nphi][3 pairs][4] interleaved complex, where pair x holds
   (sum_jy, sum_jz, sum_my, sum_mz), pair y (sum_jz, sum_jx, sum_mz, sum_mx),
   pair z (sum_jx, sum_jy, sum_mx, sum_my) of nffft_fd.cpp:423-556 before the
   area factor.  wave_number[nf] = 2 pi f / c0; dl = cell size. */
xf_status xf_engine_farfield_sums(xf_engine *e, double dl,
                                  const double *wave_number,
                                  const double *theta, int ntheta,
                                  const double *phi, int nphi, double *out);
/* The same far-field face sums without any device allocation.  This entry
   point has no counterpart in the reference: it is for callers that cannot
   call hipMalloc (serving loops, harnesses that forbid allocation).  The
   direction table and the output block live in one file-scope device array
   of XF_FARFIELD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
   a request needs 3 ntheta nphi + nf doubles (rounded up to 256 bytes) plus
   192 nf ntheta nphi bytes, and a larger one is refused with XF_EINVAL.
   Every call on a GPU shares that GPU's array: calls that overlap on one GPU
   (from several threads or engines) race.  Output equals
   xf_engine_farfield_sums bit for bit. */
#define XF_FARFIELD_NOALLOC_SCRATCH_BYTES (8u << 20)
xf_status xf_engine_farfield_sums_noalloc(xf_engine *e, double dl,
                                          const double *wave_number,
                                          const double *theta, int ntheta,
                                          const double *phi, int nphi, double *out);

/* Broadband NF-FF surface (NffftBroadBand, nffft/nffft_broadband.cpp:30-91,
   the monostatic RCS path): after every step t < total_steps the face
   currents of every surface cell of box {li,ri,lj,rj,lk,rk} (nffft/nffft.cpp:
   29-33) are stored, with the signs and stencils of NffftFd
   (nffft_broadband.cpp:94-155).  A batch refuses engines that have one.
   On a 2-D grid (nz = 1) the box is {li,ri,lj,rj,0,1} and the z faces are
   not kept (nffft_broadband.cpp:438-440 never integrates them). */
xf_status xf_engine_set_nffft_broadband(xf_engine *e, const int64_t box[6], int64_t total_steps);
/* the stored series, FP64, [total_steps][8][Q]: current c in the order of the
   get_nffft accumulators, Q surface cells (x faces (j,k), then y faces
   (k,i), then z faces (i,j), the last index fastest) */
int64_t xf_engine_broadband_samples_size(const xf_engine *e);
xf_status xf_engine_get_broadband_samples(xf_engine *e, double *out);
/* Surface integrals of the broadband far field in direction (theta, phi)
   (nffft_broadband.cpp:348-466): for Ns = 2 total_steps - 1 bins,
   bin q = m + total_steps - 1 of n_a += ds e^{i delay k_q} fftshift(fft(j_a))
   and l_a likewise from the magnetic currents, where fft is the forward DFT
   over Ns points of the series (zero past the last step) and delay is the
   cell's projection from the box centre (grid/grid_box.cpp:101-105).
   wave_number[Ns] = k_q (2 pi f_q / c0 for the reference's bins); dl = cell
   size; out = n_x, n_y, n_z, l_x, l_y, l_z, each [Ns] interleaved complex.
   The DFTs run as batched FP64 real-to-complex hipFFT transforms. */
xf_status xf_engine_broadband_sums(xf_engine *e, double dl, const double *wave_number, double theta,
                                   double phi, double *out);
/* device scratch the far field may take per batch of transforms (rows and
   spectra; default 1 GiB): larger batches mean fewer launches.  No
   reference counterpart (the reference transforms one series at a time). */
xf_status xf_engine_set_broadband_scratch(xf_engine *e, int64_t bytes);

/* E += coff[cell] * wave[t] over the box [i0,i1)x[j0,j1)x[k0,k1) of component
   comp (XF_EX/EY/EZ); coff has (i1-i0)(j1-j0)(k1-k0) entries, wave total_steps. */
xf_status xf_engine_add_current_source(xf_engine *e, int comp, int64_t i0,
                                       int64_t i1, int64_t j0, int64_t j1,
                                       int64_t k0, int64_t k1,
                                       const double *coff, const double *wave,
                                       int64_t total_steps);

/* the run's total_time_steps (Simulation::run(T), simulation_update.cpp:5-8):
   stepping beyond it is refused, and probe buffers are sized by it.  Tables
   with their own length (sources, TFSF, NF-FF, DFT) lower it further. */
xf_status xf_engine_set_horizon(xf_engine *e, int64_t total_steps);

/* Inductor::correctE (lumped_element/inductor.cpp:70-146) on the box
   [i0,i1) x [j0,j1) x [k0,k1) of component comp (0..2): after the E
   update, E += cecjc J, then J = cjcec E + J, per cell (C order over the
   box).  Lumped corrections run in the order they were added, sources and
   inductors alike.  J is loop state (checkpointed); gradients refuse
   engines with inductors. */
xf_status xf_engine_add_inductor(xf_engine *e, int comp, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
                                 int64_t k0, int64_t k1, const double *cecjc, const double *cjcec);
/* VoltageMonitor / CurrentMonitor::update (monitor/voltage_monitor.cpp:63-93,
   monitor/current_monitor.cpp:66-122) as a term list: after every step
   t < total_steps, value[t] = sign * (S0 + S1), S_g the FP64 sum in term
   order of w[q] (f[cell_a] - f[cell_b]) over the terms of group g, with f
   component comp[q] and cell_b[3q] < 0 meaning f[cell_a] alone.  Cells are
   (i, j, k) triples.  *index receives the monitor's number. */
xf_status xf_engine_add_line_monitor(xf_engine *e, int nterms, const int *comp, const int64_t *cell_a,
                                     const int64_t *cell_b, const double *w, const int *group, double sign,
                                     int64_t total_steps, int *index);
int xf_engine_line_monitor_count(const xf_engine *e);
/* values [monitor][total_steps] */
xf_status xf_engine_get_line_monitors(xf_engine *e, double *out, int64_t total_steps);

/* point sample of comp at (i,j,k) after every step */
xf_status xf_engine_add_probe(xf_engine *e, int comp, int64_t i, int64_t j,
                              int64_t k);
/* out [n_probes][total_steps] */
xf_status xf_engine_get_probes(xf_engine *e, double *out, int64_t total_steps);

/* Run n steps of the reference's time loop from the current step; the
   launches are enqueued on the engine's stream and the call returns at once. */
xf_status xf_engine_step(xf_engine *e, int64_t n);
xf_status xf_engine_sync(xf_engine *e);
int64_t xf_engine_current_step(const xf_engine *e);

xf_status xf_engine_get_field(xf_engine *e, int comp, double *out);
/* comp over the box [i0, i0+ni) x [j0, j0+nj) x [k0, k0+nk), box = {i0, ni,
   j0, nj, k0, nk}, in C order: the slice a TimeDomainFieldMonitor dumps
   (monitor/field_monitor.cpp:99-121); only the box crosses to the host. */
xf_status xf_engine_get_box(xf_engine *e, int comp, const int64_t *box, double *out);
xf_status xf_engine_set_field(xf_engine *e, int comp, const double *in);

/* average device time (ms) of the E and H update kernels over the steps run
   since the last reset, measured with HIP events on the engine's stream */
xf_status xf_engine_kernel_times(xf_engine *e, double *ms_e, double *ms_h,
                                 int64_t *samples);
xf_status xf_engine_set_timing(xf_engine *e, int enabled);
/* launch geometry of the curl kernels: tz threads along z per block
   (multiple of 64), ty rows along y (0 = flat runs of tz cells over the
   padded (y,z) plane), tx x-planes marched per thread (0 = whole axis), and
   flags: XF_TILE_GLOBAL_TABLES reads coefficient tables from global memory
   instead of LDS, XF_TILE_X_FASTEST orders tiles along x first,
   XF_TILE_XCD_RUNS gives each XCD one contiguous run of tiles instead of
   handing them out in workgroup-id order, XF_TILE_XCD_COLS keeps all k tiles
   of a (j, i) column on one XCD.  Every geometry gives identical results.
   Default: 64, 4, 8, 0; 256, 0, 4, 0 for grids of at most 81^3 nodes and
   64, 0, 4, 0 for 2-D grids. */
enum { XF_TILE_GLOBAL_TABLES = 2, XF_TILE_X_FASTEST = 4, XF_TILE_XCD_RUNS = 8, XF_TILE_XCD_COLS = 16 };
xf_status xf_engine_set_tiling(xf_engine *e, int tz, int ty, int tx, int flags);
/* Time-step scheme (no counterpart in the reference, whose loop runs
   updateE and updateH as two sweeps; simulation_update.cpp:9-36).  mode 0
   (the default): the split E and H kernels of xf_engine_set_tiling.  mode
   1: one fused kernel per step computes E and then H in a single march over
   x, reading the step's fields once and writing the next step's into a
   second copy of the state (fields, psi and the incident line; the engine
   swaps the copies after each step).  It applies to 3-D engines without
   design boxes, dispersive media or lumped elements that are not in a
   batch; others keep the split kernels.  rows: j rows per thread of the
   fused kernel (1, 2 or 4; 0: 2); planes: x planes per workgroup (1 to 32;
   0: 16).  mode | 16: one contiguous run of tiles per XCD instead of tile
   order by workgroup id.  Results are bit-identical for every choice. */
xf_status xf_engine_set_sweep(xf_engine *e, int mode, int rows, int planes);
/* 1 when the next step runs the fused kernel, else 0 */
int xf_engine_sweep_active(const xf_engine *e);

/* Divergence check (no counterpart in the reference, whose time loop runs
   on whatever the fields become; simulation_update.cpp:9-36).  The H kernels flag the
   first step whose H comes out non-finite (NaN or infinite); every `every`
   steps (default 100, 0: never) xf_engine_step and xf_batch_step read the
   flag, a sync point, and if it is set return XF_ESTATE naming that step.
   The engine then refuses to step until xf_engine_restore. */
xf_status xf_engine_set_divergence_check(xf_engine *e, int every);
/* the step at which non-finite values were detected, or -1 */
int64_t xf_engine_diverged_step(const xf_engine *e);

/* Gradient schedule (no counterpart in the reference, which has no
   adjoint): xf_engine_grad_eps replays each checkpointed segment and then
   sweeps it backwards; with overlap the sweep of segment s runs on a second
   stream while segment s - 1 is replayed.  mode -1 (the default): overlap in
   FP32 only; 0: never; 1: always.  The gradient is bit-identical either way. */
xf_status xf_engine_set_adjoint_overlap(xf_engine *e, int mode);

/* Drude cells (DrudeMedium, material/dispersive_material.cpp:118-193, and
   the dispersive branch of Simulation::updateE, simulation_update.cpp:
   221-280).  box[6] = [i0,i1)x[j0,j1)x[k0,k1) holds every Drude cell and lies
   outside every CPML slab; coef[m][6] = ca, cb, 1 + k, k,
   beta, dl/2 of material m as DrudeMedium::init computes them
   (dispersive_material.cpp:127-144), cb = 0 for a material that is not a
   Drude medium.  A Drude cell's E components inside the Yee ranges take
   E' = ca E + cb (curl - dl/2 (1+k) J), J' = k J + beta (E' + E); its H
   components use the standard tables.  The J arrays start at zero and are
   part of the checkpointed state.  Call after xf_engine_set_materials. */
xf_status xf_engine_set_drude(xf_engine *e, const int64_t *box, const double *coef);
/* The dispersive media of the box, one record of XF_ADE_STRIDE doubles per
   material: [XF_ADE_KIND] = 0 none, 1 DrudeMedium, 2 DebyMedium, 3
   LorentzMedium (material/dispersive_material.cpp:7-275, by the tagging
   xf_engine_set_drude describes); [5] = dl/2 and
     Drude   [0..4] = ca, cb, 1 + k, k, beta:  as xf_engine_set_drude
     Debye   [0..4] = ca, cb, 1 + k, k, beta, [9] = dt:
             E' = ca E + cb (curl - dl/2 (1+k) J), J' = k J + beta (E' - E) / dt
     Lorentz [0..4] = ca, cb, 1 + alpha, alpha, gamma, [7] = xi, [8] = cc,
             [10] = 2 dt:  with the previous step's E_p and J_p,
             E' = ca E + cb E_p + cc (curl - dl/2 ((1+alpha) J + xi J_p)),
             J' = alpha J + xi J_p + gamma (E' - E_p) / (2 dt), then E_p = E, J_p = J.
   J (and for Lorentz E_p, J_p) start at zero and are checkpointed state.
   Gradients refuse Debye and Lorentz media.  Call after
   xf_engine_set_materials, in place of xf_engine_set_drude. */
#define XF_ADE_STRIDE 12
#define XF_ADE_KIND 6
xf_status xf_engine_set_dispersive(xf_engine *e, const int64_t *box, const double *coef);
/* the Drude box, and J of component comp (0..2) over it, C order */
xf_status xf_engine_drude_box(xf_engine *e, int64_t *box);
xf_status xf_engine_get_drude_j(xf_engine *e, int comp, double *out);
xf_status xf_engine_set_drude_j(xf_engine *e, int comp, const double *in);

/* ---- Differentiation (no counterpart in the reference, which runs forward
   only).  A design box of cells whose permittivity is a parameter, a DFT
   monitor whose energy is the objective, state checkpoints, and the
   discrete adjoint. */

/* Design box [i0,i1)x[j0,j1)x[k0,k1) (box[6]) inside the grid and outside
   every CPML slab (call after xf_engine_set_cpml).  coef: [cell][9] E
   coefficients of each design cell in the order cexe, ceye, ceze, cexhy,
   ceyhx, cexhz, cezhx, ceyhz, cezhy (the reference's per-cell arrays,
   fdtd_basic_coff.cpp:77-91, for a lossless cell); eps_r: the cells'
   relative permittivity (for the gradient).  Cells in C order. */
xf_status xf_engine_set_design(xf_engine *e, const int64_t *box, const double *coef, const double *eps_r);
/* New coefficients and permittivity for the cells of the design box already
   set (same layout as xf_engine_set_design), for an optimisation loop that
   keeps one engine; the fields are left as they are (rewind with
   xf_engine_restore to a checkpoint taken at step 0). */
xf_status xf_engine_update_design(xf_engine *e, const double *coef, const double *eps_r);
/* DFT monitor of component comp on the plane axis = index, the other two
   axes (increasing order) over [lo1,hi1) x [lo2,hi2): after step t,
   F += f * w[t] in FP64.  w: interleaved complex [total_steps]. */
xf_status xf_engine_set_dft_plane(xf_engine *e, int comp, int axis, int64_t index, int64_t lo1, int64_t hi1,
                                  int64_t lo2, int64_t hi2, const double *w, int64_t total_steps);
xf_status xf_engine_get_dft(xf_engine *e, double *out /* interleaved complex */);
/* copy of the whole time-loop state (fields, psi, incident grid,
   accumulators, step counter) into / back from a numbered slot */
xf_status xf_engine_checkpoint(xf_engine *e, int slot);
xf_status xf_engine_restore(xf_engine *e, int slot);
/* J = sum |F|^2 after the full run and dJ/d eps_r of every design cell
   (C order).  Runs the forward loop from step 0 with a checkpoint every
   `checkpoint_every` steps, then the reverse loop segment by segment: each
   segment is recomputed from its checkpoint, keeping the design-box E
   fields (each recomputed step covers only the tiles that can still reach
   the design box by the segment's end), and the adjoint kernels step back
   through it (over the planes the adjoint can have reached from the
   objective plane); the last segment's
   design fields are kept during the forward run, so an interval >= the run
   length recomputes nothing.  checkpoint_every = 0 chooses: the whole run
   when its design fields fit in a quarter of the free device memory (at
   least 4 GiB), else 100.  The engine is left at the
   end of the forward run.  Design cells must be lossless. */
xf_status xf_engine_grad_eps(xf_engine *e, int checkpoint_every, double *grad, double *J);
/* the transpose of nsteps steps of the field map applied to six adjoint
   fields (packed in component order, reference shapes), in place; the psi
   adjoints start at zero.  With a Drude box the three J adjoints follow in
   the E shapes (cells outside the box pass through).  For checking
   <L x, y> = <x, L^T y>. */
xf_status xf_engine_adjoint_run(xf_engine *e, int64_t nsteps, double *fields);

/* Batches: several set-up engines of one precision, all at the same step,
   stepped together.  Each per-step kernel (curl + CPML, TFSF faces, 1-D
   incident grids, NF-FF) is one launch over every instance, so a batch of
   small, uneven problems fills the GPU where one engine at a time cannot.
   Results are bit-identical to stepping each engine alone.  The batch has its
   own stream; the member engines' other calls (get_field, get_nffft, step,
   ...) wait for it.  Destroy the batch before its engines.  No counterpart in
   the reference, which runs one Simulation at a time. */
typedef struct xf_batch xf_batch;
xf_status xf_batch_create(xf_engine *const *engines, int n, xf_batch **out);
/* n steps of every member (enqueued; returns at once) */
xf_status xf_batch_step(xf_batch *b, int64_t n);
xf_status xf_batch_sync(xf_batch *b);
int xf_batch_size(const xf_batch *b);
void xf_batch_destroy(xf_batch *b);
/* xf_engine_grad_eps for every member at once (all members at step 0, with
   a design box, a DFT plane and the same step horizon): the forward run and
   the recomputed segments step as a batch, the reverse steps are one launch
   per kernel over all members.  grads[q] receives member q's dJ/d eps_r,
   J[q] its objective; the results equal each member's own
   xf_engine_grad_eps.  checkpoint_every = 0: the whole run when all members'
   design fields fit in a quarter of the free device memory (at least
   4 GiB), else 100. */
xf_status xf_batch_grad_eps(xf_batch *b, int checkpoint_every, double *const *grads, double *J);

#ifdef __cplusplus
}
#endif
#endif
