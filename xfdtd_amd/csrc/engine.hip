// engine.hip -- C-ABI of the gfx950 FDTD field engine (include/xfdtd_engine.h).
//
// Device state of one simulation and the per-step launch sequence of the
// reference's run loop (simulation/simulation_update.cpp:9-36):
//   E curl + CPML-E (fused) -> current sources -> TFSF E faces
//   -> H curl + CPML-H (fused) -> 1-D incident grid step -> TFSF H faces
//   -> NF-FF running DFT -> probes
// The reference applies CPML after the lumped and TFSF corrections; the host
// validates that the lumped cells and TFSF faces lie outside every CPML slab,
// where the two orders touch disjoint cells and give identical bits.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "engine_kernels.hpp"
#include "sweep_kernels.hpp"
#include "xfdtd_engine.h"

using namespace xfk;

static thread_local std::string g_err;
extern "C" const char *xf_last_error(void) { return g_err.c_str(); }

#define FAILS(code, ...)                                   \
  do {                                                     \
    char b_[512];                                          \
    snprintf(b_, sizeof b_, __VA_ARGS__);                  \
    g_err = b_;                                            \
    return code;                                           \
  } while (0)
#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) FAILS(XF_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

namespace {

// a lumped E correction on a box: kind 0 adds coff * wave[t] (current and
// voltage sources), kind 1 is an inductor (E += coff J; J = c2 E + J)
struct SourceDev {
  int comp, i0, i1, j0, j1, k0, k1;
  int kind = 0;
  void *coff = nullptr, *wave = nullptr, *c2 = nullptr, *jc = nullptr;
};

// voltage / current monitor: value = sign (sum of group 0 + sum of group 1)
// of terms w (f[ca] - f[cb]) (f[ca] alone when cb < 0), in term order
struct LineMonDev {
  void *terms = nullptr;
  int n = 0;
  double sign = 1;
  double *buf = nullptr;  // [T]
  long long T = 0;
};

}  // namespace

struct xf_engine {
  xf_grid_desc g{};
  int nx = 0, ny = 0, nz = 0, pz = 0, nmat = 0, prec = 64;
  size_t esz = 8;
  long long si = 0, sj = 0, ncell = 0;
  hipStream_t stream = nullptr;
  void *f[6] = {};
  uint8_t *mat = nullptr;
  // coefficient records per side (0: E, 1: H), see Grid in engine_kernels.hpp
  void *self[2] = {}, *kt[2] = {}, *jt[2] = {}, *it[2] = {}, *kp[2] = {}, *jp[2] = {}, *ip[2] = {};
  bool h_uni = false;  // every material has material 0's H coefficients
  bool materials_set = false;
  int dn[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
  void *be[3] = {}, *ae[3] = {}, *bm[3] = {}, *am[3] = {};
  void *psi_e[6] = {}, *psi_h[6] = {};
  // TFSF
  bool has_tfsf = false;
  int tli = 0, tri = 0, tlj = 0, trj = 0, tlk = 0, trk = 0, L = 0;
  void *pf[3] = {}, *ph[3] = {}, *e_inc = nullptr, *h_inc = nullptr, *ab = nullptr, *src = nullptr;
  double te[3] = {}, th[3] = {}, ceie = 0, ceihi = 0, chih = 0, chiei = 0, c1 = 0, c2 = 0;
  double tca[3] = {}, tcb[3] = {};
  long long tfsf_T = 0;
  bool tfsf_two_d = false;
  // NF-FF
  bool has_nf = false;
  Nffft nf{};
  double2 *acc = nullptr, *tj = nullptr, *tm = nullptr;
  long long nf_size = 0;
  // face values of up to NF_RING steps not yet added to the accumulators
  // (empty between API calls) and the step of the first
  void *nf_ring = nullptr;
  void *nf_zrec = nullptr;  // NF-FF z-face records (NfRec)
  int nf_pending = 0;
  long long nf_step0 = 0;
  // broadband NF-FF: the face values of every step, [step][8][surface cell]
  bool has_bb = false;
  Nffft bb{};
  void *bb_hist = nullptr;
  long long bb_T = 0;
  long long bb_scratch = 1LL << 30;  // bytes of rows + spectra per far-field transform
  // sources, probes
  std::vector<SourceDev> sources;
  std::vector<LineMonDev> line_mons;
  long long src_T = 0;
  std::vector<int> probe_comp;
  std::vector<long long> probe_cell;
  int *d_probe_comp = nullptr;
  long long *d_probe_cell = nullptr;
  double *probe_buf = nullptr;
  long long probe_T = 0;
  bool probes_dirty = true;
  long long cur = 0;
  long long total = -1;  // smallest total_steps among the attached tables
  // tiling
  int tz = 64, ty = 4, tx = 8, tile_flags = 0;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;  // 4 events per timed step: E start/end, H start/end
  size_t ev_used = 0;
  std::vector<void *> owned;
  // stream of the batch this engine belongs to (work on it precedes any use
  // of the engine's own stream)
  hipStream_t batch_stream = nullptr;
  // every buffer that is state of the time loop (fields, psi, incident
  // grid, accumulators): what a checkpoint copies.  Buffers that the fused
  // step double-buffers are named by the engine member that points at the
  // current copy.
  struct StateBuf {
    void *p;
    void **slot;
    size_t n;
    void *ptr() const { return slot ? *slot : p; }
  };
  std::vector<StateBuf> state;
  std::vector<std::vector<void *>> ckpt;
  std::vector<long long> ckpt_cur;
  size_t psi_bytes[3] = {0, 0, 0};  // bytes of each psi array per axis
  // design box with per-cell E coefficients
  bool has_design = false;
  int db[6] = {0, 0, 0, 0, 0, 0};
  void *dcoef = nullptr;
  double *d_eps = nullptr;
  int nd = 0;
  // DFT plane monitor
  bool has_dft = false;
  DftPlane dft{};
  long long dft_T = 0;
  // Drude box: per-material ADE coefficients and the box's J arrays
  bool has_drude = false;
  int xb[6] = {0, 0, 0, 0, 0, 0};
  void *drc = nullptr;
  void *dj[3] = {nullptr, nullptr, nullptr};
  void *djp[3] = {nullptr, nullptr, nullptr}, *dep[3] = {nullptr, nullptr, nullptr};  // Lorentz state
  bool ade_non_drude = false;  // a Debye or Lorentz record (no adjoint)
  int dev = 0;                 // HIP device the engine's memory lives on
  // fused E+H step (k_sweep): off by default (the split kernels are faster
  // on MI355X); rows of j per thread, x planes per workgroup (0: defaults)
  int sweep = 0, sweep_rows = 0, sweep_tx = 0;
  int sweep_flags = 0;  // 1: each XCD takes a contiguous run of tiles
  // divergence check: the H kernels lower *div to the first step whose H
  // came out non-finite; the host reads it every div_every steps (0: never)
  // and then refuses to step on (div_step >= 0) until a restore
  void *div = nullptr;
  int div_every = 100;
  long long div_step = -1;
  // re-running a segment from a checkpoint for the adjoint: the observers
  // (NF-FF, DFT plane, broadband series, probes, line monitors) are skipped;
  // their state is restored from the checkpoint at the end of the run
  bool replay = false;
  // a replay step's curl launches cover only the tiles that meet the cell
  // window [win[0], win[1]) x [win[2], win[3]) x [win[4], win[5]) (i, j, k)
  bool win_on = false;
  int win[6] = {0, 0, 0, 0, 0, 0};
  // a gradient's design-box history slot the split E kernel fills as it
  // steps (null: none)
  void *snap = nullptr;
  // gradient: run each segment's adjoint sweep on a second stream beside the
  // replay of the segment before it (-1: in FP32 only, where the kernels are
  // latency-bound at the sizes a gradient takes; 0: never; 1: always)
  int adj_overlap = -1;
  // second copies of the double-buffered state (allocated at the first
  // fused step): the fused step reads e->f, psi, the incident line and
  // writes these, then the two are swapped
  bool alt_ready = false;
  void *f_alt[6] = {};
  void *psi_e_alt[6] = {}, *psi_h_alt[6] = {};
  // for the fused step both psi sets live in one block each, slots at
  // psi_n: the elements of one slot (a pair takes two)
  void *psi_blk = nullptr, *psi_blk_alt = nullptr;
  long long psi_n = 0;
  void *e_inc_alt = nullptr, *h_inc_alt = nullptr;
  TfTerms tf_terms{};              // the step's TFSF face terms (fused step)
  // the TFSF z-face terms the curl kernels add (ZFaces), 8 FP64 segments
  void *zterms = nullptr;
  // fused-step tile lists, one run per kind (SW_CORE, SW_ZONLY, SW_ANY) in
  // one device array, built for rows/planes (sp_R, sp_tx)
  void *sp_tiles = nullptr;
  int sp_R = 0, sp_tx = 0, sp_off[3] = {}, sp_cnt[3] = {};
};

// wait for everything enqueued for this engine, on its own stream and on
// the stream of a batch it belongs to
static hipError_t engine_wait(xf_engine *e) {
  hipError_t r = hipSuccess;
  if (e->batch_stream) r = hipStreamSynchronize(e->batch_stream);
  if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
  return r;
}

static xf_status dev_alloc(xf_engine *e, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t r = hipMalloc(p, bytes);
  if (r != hipSuccess) FAILS(XF_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(r));
  e->owned.push_back(*p);
  // zero on the engine stream and wait, so that later synchronous uploads on
  // the null stream cannot be overtaken by the fill
  r = hipMemsetAsync(*p, 0, bytes, e->stream);
  if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
  if (r != hipSuccess) FAILS(XF_EDEVICE, "hipMemsetAsync: %s", hipGetErrorString(r));
  return XF_OK;
}

// ---------------------------------------------------------------- allocation-free far field
// xf_engine_farfield_sums_noalloc keeps its direction table, wave numbers and
// the output block of k_farfield in this file-scope array instead of device
// buffers allocated per call.  Each device holds its own copy for the life of
// the process; its address on a device is taken once, with that device
// current, when the first engine is created there.  The host-side name is
// not a device address and is never passed to a kernel or a copy.
static constexpr size_t kFarScratchBytes = (size_t)XF_FARFIELD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double2 scratch[kFarScratchBytes / sizeof(double2)];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices] = {};
// host staging of the direction table (never larger than the device block)
static double g_far_host[kFarScratchBytes / sizeof(double)];

static xf_status scratch_setup(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) FAILS(XF_EDEVICE, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  if (g_scratch_dev[dev]) return XF_OK;
  void *p = nullptr;
  const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (r != hipSuccess || p == nullptr)
    FAILS(XF_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(r));
  g_scratch_dev[dev] = p;
  return XF_OK;
}

// scratch device buffers of one call, zeroed on the stream that uses them;
// alloc waits for the fill, so neither that stream's earlier work nor a
// synchronous copy on the null stream can see the fill land late (a fill on
// the null stream is not ordered with the engines' non-blocking streams)
struct AdjBufs {
  // device buffers of one adjoint or gradient run, carved from a few large
  // allocations (slabs of geometrically growing size): a run takes dozens of
  // buffers, and one hipMalloc / hipFree each would cost more than the
  // run's kernels on small grids
  hipStream_t st = nullptr;
  std::vector<void *> mem;
  char *cur = nullptr;
  size_t used = 0, cap = 0, next = 16u << 20;
  explicit AdjBufs(hipStream_t s = nullptr) : st(s) {}
  AdjBufs(const AdjBufs &) = delete;
  AdjBufs &operator=(const AdjBufs &) = delete;
  // `bytes` of device memory (256-byte aligned), zeroed unless zero = false
  void *alloc(size_t bytes, bool zero = true) {
    if (bytes == 0) bytes = 16;
    const size_t need = (bytes + 255) & ~size_t(255);
    if (!cur || used + need > cap) {
      const size_t sz = std::max(need, next);
      void *p = nullptr;
      if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
      mem.push_back(p);
      cur = (char *)p;
      cap = sz;
      used = 0;
      next = std::min<size_t>(2 * sz, (size_t)4 << 30);
    }
    void *p = cur + used;
    used += need;
    if (zero && (hipMemsetAsync(p, 0, bytes, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess))
      return nullptr;
    return p;
  }
  ~AdjBufs() {
    if (st) hipStreamSynchronize(st);
    for (void *p : mem) hipFree(p);
  }
};

// upload FP64 host data as the engine's working precision
static xf_status upload_real(xf_engine *e, void **dst, const double *src, size_t n) {
  xf_status s = dev_alloc(e, dst, n * e->esz);
  if (s != XF_OK) return s;
  if (n == 0) return XF_OK;
  if (e->prec == 64) {
    HIPCHK(hipMemcpy(*dst, src, n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(n);
    for (size_t q = 0; q < n; ++q) tmp[q] = (float)src[q];
    HIPCHK(hipMemcpy(*dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return XF_OK;
}

// upload FP64 host data as FP64 whatever the working precision
static xf_status upload_f64(xf_engine *e, void **dst, const double *src, size_t n) {
  xf_status s = dev_alloc(e, dst, n * sizeof(double));
  if (s != XF_OK) return s;
  if (n > 0) HIPCHK(hipMemcpy(*dst, src, n * sizeof(double), hipMemcpyHostToDevice));
  return XF_OK;
}

static void comp_shape(const xf_engine *e, int comp, int sh[3]) {
  const int nx = e->nx, ny = e->ny, nz = e->nz;
  switch (comp) {
    case XF_EX: sh[0] = nx; sh[1] = ny + 1; sh[2] = nz + 1; break;
    case XF_EY: sh[0] = nx + 1; sh[1] = ny; sh[2] = nz + 1; break;
    case XF_EZ: sh[0] = nx + 1; sh[1] = ny + 1; sh[2] = nz; break;
    case XF_HX: sh[0] = nx + 1; sh[1] = ny; sh[2] = nz; break;
    case XF_HY: sh[0] = nx; sh[1] = ny + 1; sh[2] = nz; break;
    default: sh[0] = nx; sh[1] = ny; sh[2] = nz + 1; break;
  }
}

// clear the divergence flag (a blocking copy: no memset node for a stream
// that may be captured)
static xf_status reset_divergence(xf_engine *e) {
  const unsigned long long none = ~0ull;
  HIPCHK(hipMemcpy(e->div, &none, sizeof none, hipMemcpyHostToDevice));
  e->div_step = -1;
  return XF_OK;
}

// read the divergence flag after the work enqueued on `st` (a sync point)
static xf_status check_divergence(xf_engine *e, hipStream_t st, int member = -1) {
  unsigned long long h = ~0ull;
  HIPCHK(hipMemcpyAsync(&h, e->div, sizeof h, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h == ~0ull) return XF_OK;
  e->div_step = (long long)h;
  if (member >= 0)
    FAILS(XF_ESTATE, "batch member %d: non-finite H values at step %lld, the run diverged (time step above the "
          "Courant limit or an unstable medium)", member, e->div_step);
  FAILS(XF_ESTATE, "non-finite H values at step %lld, the run diverged (time step above the Courant limit or an "
        "unstable medium)", e->div_step);
}

extern "C" xf_status xf_engine_create(const xf_grid_desc *g, xf_engine **out) {
  if (!g || !out) FAILS(XF_EINVAL, "null argument");
  *out = nullptr;
  if (g->nx < 2 || g->ny < 2 || g->nz < 1)
    FAILS(XF_EINVAL, "grid must be 3-D or 2-D with nz = 1 (got %lld,%lld,%lld)", (long long)g->nx,
          (long long)g->ny, (long long)g->nz);
  if (g->nx > 100000 || g->ny > 100000 || g->nz > 100000) FAILS(XF_EINVAL, "grid too large");
  if (g->precision != 32 && g->precision != 64) FAILS(XF_EINVAL, "precision must be 32 or 64");
  if (g->n_materials < 1 || g->n_materials > 255) FAILS(XF_EINVAL, "n_materials must be 1..255");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) FAILS(XF_EDEVICE, "no HIP device");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, 0));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    FAILS(XF_EDEVICE, "device is %s; this engine is built for gfx950", prop.gcnArchName);
  xf_engine *e = new xf_engine();
  e->g = *g;
  e->nx = (int)g->nx; e->ny = (int)g->ny; e->nz = (int)g->nz;
  e->prec = g->precision;
  e->esz = e->prec == 64 ? 8 : 4;
  e->nmat = g->n_materials;
  // z rows padded to 32 cells; a 2-D grid (nz == 1) keeps its two z slots
  // unpadded and runs the curl kernels on flat runs of the (j, k) plane
  e->pz = e->nz == 1 ? 2 : ((e->nz + 1) + 31) / 32 * 32;
  if (e->nz == 1) {
    e->tz = 64;
    e->ty = 0;
    e->tx = 4;
  } else if ((long long)(e->nx + 1) * (e->ny + 1) * (e->nz + 1) <= 81LL * 81 * 81) {
    // small grids: flat runs of 256 cells marching 4 planes give several
    // times the workgroups of 64 x 4 tiles over 8 planes, and the short
    // kernels are latency-bound (64^3: 58.9 against 69.5 us per FP64 step;
    // at 96^3 FP64 the 64 x 4 tiles are faster again)
    e->tz = 256;
    e->ty = 0;
    e->tx = 4;
  }
  e->sj = e->pz;
  e->si = (long long)(e->ny + 1) * e->pz;
  e->ncell = (long long)(e->nx + 1) * e->si;
  // a curl-kernel tile addresses its planes with 32-bit byte offsets from a
  // 64-bit base at its first plane: it needs two y-z planes under 4 GiB
  if (2ULL * (unsigned long long)e->si * e->esz >= (1ULL << 32)) {
    const long long ps = e->si;
    delete e;
    FAILS(XF_EINVAL, "a y-z plane of %lld padded cells: two planes exceed the 4 GiB a tile addresses", ps);
  }
  if (hipGetDevice(&e->dev) != hipSuccess) {
    delete e;
    FAILS(XF_EDEVICE, "hipGetDevice failed");
  }
  if (scratch_setup(e->dev) != XF_OK) {
    delete e;
    return XF_EDEVICE;
  }
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    FAILS(XF_EDEVICE, "hipStreamCreate failed");
  }
  for (int q = 0; q < 6; ++q) {
    xf_status s = dev_alloc(e, &e->f[q], (size_t)e->ncell * e->esz);
    if (s != XF_OK) { xf_engine_destroy(e); return s; }
    e->state.push_back({nullptr, &e->f[q], (size_t)e->ncell * e->esz});
  }
  xf_status s = dev_alloc(e, &e->div, sizeof(unsigned long long));
  if (s == XF_OK) s = reset_divergence(e);
  if (s != XF_OK) { xf_engine_destroy(e); return s; }
  *out = e;
  return XF_OK;
}

extern "C" void xf_engine_destroy(xf_engine *e) {
  if (!e) return;
  if (e->stream) engine_wait(e);
  for (void *p : e->owned) hipFree(p);
  for (hipEvent_t ev : e->ev) hipEventDestroy(ev);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

extern "C" xf_status xf_engine_set_materials(xf_engine *e, const uint8_t *index, const double *ca,
                                             const double *const *cb, const double *const *cpsi) {
  if (!e || !index || !ca || !cb || !cpsi) FAILS(XF_EINVAL, "null argument");
  const int nx = e->nx, ny = e->ny, nz = e->nz, nm = e->nmat;
  // material index into the padded layout
  std::vector<uint8_t> m((size_t)e->ncell, 0);
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j <= ny; ++j) {
      const uint8_t *src = index + ((size_t)i * (ny + 1) + j) * (nz + 1);
      uint8_t *dst = m.data() + (size_t)i * e->si + (size_t)j * e->sj;
      for (int k = 0; k <= nz; ++k) {
        if (src[k] >= nm) FAILS(XF_EINVAL, "material index %d >= n_materials %d", src[k], nm);
        dst[k] = src[k];
      }
    }
  void *pm = nullptr;
  xf_status s = dev_alloc(e, &pm, (size_t)e->ncell);
  if (s != XF_OK) return s;
  HIPCHK(hipMemcpy(pm, m.data(), (size_t)e->ncell, hipMemcpyHostToDevice));
  e->mat = (uint8_t *)pm;
  const int nax[3] = {nx, ny, nz};
  // do all materials share material 0's H coefficients (self terms, curl and
  // psi tables)?  Then the H kernels need no material byte.
  {
    static const int htab[6] = {XF_CB_HXEY, XF_CB_HYEX, XF_CB_HXEZ, XF_CB_HZEX, XF_CB_HYEZ, XF_CB_HZEY};
    static const int hax[6] = {2, 2, 1, 1, 0, 0};
    bool uni = true;
    for (int m = 1; m < nm && uni; ++m) {
      for (int c = 0; c < 3; ++c) uni = uni && ca[(3 + c) * (size_t)nm + m] == ca[(3 + c) * (size_t)nm];
      for (int t = 0; t < 6 && uni; ++t) {
        const size_t n1 = (size_t)nax[hax[t]] + 1;
        for (size_t n = 0; n < n1 && uni; ++n)
          uni = cb[htab[t]][m * n1 + n] == cb[htab[t]][n] && cpsi[htab[t]][m * n1 + n] == cpsi[htab[t]][n];
      }
    }
    e->h_uni = uni;
  }
  // self terms of each side as 4-wide records [nmat][4]
  for (int side = 0; side < 2; ++side) {
    std::vector<double> rec(4 * (size_t)nm, 0.0);
    for (int m = 0; m < nm; ++m)
      for (int c = 0; c < 3; ++c) rec[4 * (size_t)m + c] = ca[(3 * side + c) * (size_t)nm + m];
    if ((s = upload_real(e, &e->self[side], rec.data(), rec.size())) != XF_OK) return s;
  }
  // per side and axis, the two tables whose derivative runs along that axis,
  // interleaved as [n][material][2] (the kernels read both with one load):
  //   E: z (cexhy, ceyhx), y (cexhz, cezhx), x (ceyhz, cezhy)
  //   H: z (chxey, chyex), y (chxez, chzex), x (chyez, chzey)
  static const int pairs[2][3][2] = {
      {{XF_CB_EXHY, XF_CB_EYHX}, {XF_CB_EXHZ, XF_CB_EZHX}, {XF_CB_EYHZ, XF_CB_EZHY}},
      {{XF_CB_HXEY, XF_CB_HYEX}, {XF_CB_HXEZ, XF_CB_HZEX}, {XF_CB_HYEZ, XF_CB_HZEY}}};
  static const int pair_axis[3] = {2, 1, 0};
  for (int side = 0; side < 2; ++side)
    for (int q = 0; q < 3; ++q) {
      const int ax = pair_axis[q], n1 = nax[ax] + 1;
      const int t0 = pairs[side][q][0], t1 = pairs[side][q][1];
      std::vector<double> cp(2 * (size_t)n1 * nm), pp(cp.size());
      for (int n = 0; n < n1; ++n)
        for (int m = 0; m < nm; ++m) {
          const size_t d = 2 * ((size_t)n * nm + m), src = (size_t)m * n1 + n;
          cp[d] = cb[t0][src];
          cp[d + 1] = cb[t1][src];
          pp[d] = cpsi[t0][src];
          pp[d + 1] = cpsi[t1][src];
        }
      void **dc = q == 0 ? &e->kt[side] : q == 1 ? &e->jt[side] : &e->it[side];
      void **dp = q == 0 ? &e->kp[side] : q == 1 ? &e->jp[side] : &e->ip[side];
      if ((s = upload_real(e, dc, cp.data(), cp.size())) != XF_OK) return s;
      if ((s = upload_real(e, dp, pp.data(), pp.size())) != XF_OK) return s;
    }
  // default (absent) CPML tables
  for (int a = 0; a < 3; ++a) {
    if (!e->be[a]) {
      std::vector<double> z(nax[a] + 1, 0.0);
      if ((s = upload_real(e, &e->be[a], z.data(), z.size())) != XF_OK) return s;
      if ((s = upload_real(e, &e->ae[a], z.data(), z.size())) != XF_OK) return s;
      if ((s = upload_real(e, &e->bm[a], z.data(), z.size())) != XF_OK) return s;
      if ((s = upload_real(e, &e->am[a], z.data(), z.size())) != XF_OK) return s;
    }
  }
  for (int q = 0; q < 6; ++q)
    if (!e->psi_e[q]) {
      if ((s = dev_alloc(e, &e->psi_e[q], 16)) != XF_OK) return s;
      if ((s = dev_alloc(e, &e->psi_h[q], 16)) != XF_OK) return s;
    }
  e->materials_set = true;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_cpml(xf_engine *e, int axis, int d_n, int d_p, const double *b_e,
                                        const double *a_e, const double *b_m, const double *a_m) {
  if (!e || axis < 0 || axis > 2 || !b_e || !a_e || !b_m || !a_m) FAILS(XF_EINVAL, "bad argument");
  const int nax[3] = {e->nx, e->ny, e->nz};
  const int N = nax[axis];
  if (d_n < 0 || d_p < 0 || d_n + d_p >= N) FAILS(XF_EINVAL, "CPML layers %d+%d do not fit %d cells", d_n, d_p, N);
  e->dn[axis] = d_n;
  e->dp[axis] = d_p;
  xf_status s;
  if ((s = upload_real(e, &e->be[axis], b_e, (size_t)N + 1)) != XF_OK) return s;
  if ((s = upload_real(e, &e->ae[axis], a_e, (size_t)N + 1)) != XF_OK) return s;
  if ((s = upload_real(e, &e->bm[axis], b_m, (size_t)N + 1)) != XF_OK) return s;
  if ((s = upload_real(e, &e->am[axis], a_m, (size_t)N + 1)) != XF_OK) return s;
  const int Lay = d_n + d_p;
  const size_t nx1 = e->nx + 1, ny1 = e->ny + 1;
  size_t n;
  // psi arrays of the two E and two H components this axis acts on
  if (axis == 0) n = (size_t)Lay * ny1 * e->pz;       // [L][ny+1][pz]
  else if (axis == 1) n = nx1 * Lay * e->pz;          // [nx+1][L][pz]
  else n = nx1 * ny1 * Lay;                           // [nx+1][ny+1][L]
  // the axis' two arrays interleaved in one allocation per side (slot s0 at
  // the even elements, s1 at the odd), so a kernel reaches both with one
  // two-wide access
  static const int e_slots[3][2] = {{P_BX, P_CX}, {P_AY, P_CY}, {P_AZ, P_BZ}};
  const int s0 = e_slots[axis][0], s1 = e_slots[axis][1];
  if ((s = dev_alloc(e, &e->psi_e[s0], 2 * n * e->esz)) != XF_OK) return s;
  if ((s = dev_alloc(e, &e->psi_h[s0], 2 * n * e->esz)) != XF_OK) return s;
  e->psi_e[s1] = (char *)e->psi_e[s0] + e->esz;
  e->psi_h[s1] = (char *)e->psi_h[s0] + e->esz;
  e->state.push_back({nullptr, &e->psi_e[s0], 2 * n * e->esz});
  e->state.push_back({nullptr, &e->psi_h[s0], 2 * n * e->esz});
  e->psi_bytes[axis] = n * e->esz;
  return XF_OK;
}

static size_t drude_cells(const xf_engine *e) {
  return (size_t)(e->xb[1] - e->xb[0]) * (e->xb[3] - e->xb[2]) * (e->xb[5] - e->xb[4]);
}

// bytes the automatic checkpoint interval lets the design history of a
// gradient run take: a quarter of the free device memory, at least 4 GiB
static double history_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  return std::max(4.0 * (1 << 30), 0.25 * (double)free_b);
}

static void set_total(xf_engine *e, long long T) {
  if (e->total < 0 || T < e->total) e->total = T;
}

extern "C" xf_status xf_engine_set_horizon(xf_engine *e, int64_t total_steps) {
  if (!e || total_steps < 0) FAILS(XF_EINVAL, "bad argument");
  set_total(e, (long long)total_steps);
  return XF_OK;
}

extern "C" xf_status xf_engine_set_tfsf(xf_engine *e, const xf_tfsf_desc *d) {
  if (!e || !d) FAILS(XF_EINVAL, "null argument");
  const bool two_d = d->two_d != 0;
  if (two_d != (e->nz == 1)) FAILS(XF_EINVAL, two_d ? "TFSF2D needs a grid with nz = 1" : "a grid with nz = 1 needs TFSF2D");
  const bool zbad = two_d ? (d->lk != 0 || d->rk != 1) : (d->lk < 1 || d->rk >= e->nz || d->rk <= d->lk);
  if (d->li < 1 || d->lj < 1 || d->ri >= e->nx || d->rj >= e->ny || d->ri <= d->li || d->rj <= d->lj || zbad)
    FAILS(XF_EINVAL, "TFSF box outside the grid");
  if (d->aux_size < 3) FAILS(XF_EINVAL, "TFSF aux size too small");
  e->tli = (int)d->li; e->tri = (int)d->ri; e->tlj = (int)d->lj; e->trj = (int)d->rj;
  e->tlk = (int)d->lk; e->trk = (int)d->rk; e->L = (int)d->aux_size;
  const int bn[3] = {e->tri - e->tli, e->trj - e->tlj, e->trk - e->tlk};
  xf_status s;
  // the incident line, its projections and source table are FP64 at either
  // field precision
  for (int a = 0; a < 3; ++a) {
    if ((s = upload_f64(e, &e->pf[a], d->proj_full[a], (size_t)bn[a] + 1)) != XF_OK) return s;
    if ((s = upload_f64(e, &e->ph[a], d->proj_half[a], (size_t)bn[a] + 2)) != XF_OK) return s;
    e->te[a] = d->transform_e[a];
    e->th[a] = d->transform_h[a];
    e->tca[a] = d->ca[a];
    e->tcb[a] = d->cb[a];
  }
  // projections must stay inside the aux arrays for every face cell
  e->ceie = d->ceie; e->ceihi = d->ceihi; e->chih = d->chih; e->chiei = d->chiei;
  e->c1 = d->abc_c1; e->c2 = d->abc_c2;
  if ((s = dev_alloc(e, &e->e_inc, (size_t)e->L * sizeof(double))) != XF_OK) return s;
  if ((s = dev_alloc(e, &e->h_inc, (size_t)(e->L - 1) * sizeof(double))) != XF_OK) return s;
  if ((s = dev_alloc(e, &e->ab, 2 * sizeof(double))) != XF_OK) return s;
  e->state.push_back({nullptr, &e->e_inc, (size_t)e->L * sizeof(double)});
  e->state.push_back({nullptr, &e->h_inc, (size_t)(e->L - 1) * sizeof(double)});
  e->state.push_back({e->ab, nullptr, 2 * sizeof(double)});
  if ((s = upload_f64(e, &e->src, d->source, (size_t)d->total_steps)) != XF_OK) return s;
  e->tfsf_T = d->total_steps;
  set_total(e, d->total_steps);
  e->tfsf_two_d = two_d;
  e->has_tfsf = true;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_nffft(xf_engine *e, const xf_nffft_desc *d) {
  if (!e || !d || !d->tj || !d->tm) FAILS(XF_EINVAL, "null argument");
  if (d->li < 1 || d->lj < 1 || d->lk < 1 || d->ri >= e->nx || d->rj >= e->ny || d->rk >= e->nz ||
      d->ri <= d->li || d->rj <= d->lj || d->rk <= d->lk || d->nf < 1)
    FAILS(XF_EINVAL, "NF-FF box outside the grid");
  Nffft &n = e->nf;
  n.li = (int)d->li; n.ri = (int)d->ri; n.lj = (int)d->lj; n.rj = (int)d->rj;
  n.lk = (int)d->lk; n.rk = (int)d->rk; n.nf = d->nf;
  n.nxb = n.ri - n.li; n.nyb = n.rj - n.lj; n.nzb = n.rk - n.lk;
  n.zf = 1;
  n.T_total = d->total_steps;
  const long long cells[3] = {(long long)n.nyb * n.nzb, (long long)n.nzb * n.nxb, (long long)n.nxb * n.nyb};
  n.base[0] = 0;
  n.base[1] = 8LL * n.nf * cells[0];
  n.base[2] = n.base[1] + 8LL * n.nf * cells[1];
  e->nf_size = n.base[2] + 8LL * n.nf * cells[2];
  xf_status s;
  void *p;
  if ((s = dev_alloc(e, &p, (size_t)e->nf_size * sizeof(double2))) != XF_OK) return s;
  e->acc = (double2 *)p;
  e->state.push_back({p, nullptr, (size_t)e->nf_size * sizeof(double2)});
  const size_t tw = (size_t)d->nf * d->total_steps;
  if ((s = dev_alloc(e, &p, tw * sizeof(double2))) != XF_OK) return s;
  e->tj = (double2 *)p;
  HIPCHK(hipMemcpy(e->tj, d->tj, tw * sizeof(double2), hipMemcpyHostToDevice));
  if ((s = dev_alloc(e, &p, tw * sizeof(double2))) != XF_OK) return s;
  e->tm = (double2 *)p;
  HIPCHK(hipMemcpy(e->tm, d->tm, tw * sizeof(double2), hipMemcpyHostToDevice));
  n.tj = e->tj; n.tm = e->tm; n.acc = e->acc;
  set_total(e, d->total_steps);
  e->has_nf = true;
  return XF_OK;
}

extern "C" int64_t xf_engine_nffft_size(const xf_engine *e) { return e && e->has_nf ? e->nf_size : 0; }

// The device keeps the y-face cells (i, k), k fastest; the interface orders
// them (k, i).  Reorder `blocks` consecutive y-face blocks of nxb * nzb
// entries of `w` doubles each, in place.
static void y_faces_to_interface(double *y, long long blocks, int nxb, int nzb, int w) {
  std::vector<double> tmp((size_t)nxb * nzb * w);
  for (long long b = 0; b < blocks; ++b) {
    double *blk = y + b * (long long)nxb * nzb * w;
    for (int i = 0; i < nxb; ++i)
      for (int k = 0; k < nzb; ++k)
        for (int c = 0; c < w; ++c) tmp[((size_t)k * nxb + i) * w + c] = blk[((size_t)i * nzb + k) * w + c];
    std::memcpy(blk, tmp.data(), tmp.size() * sizeof(double));
  }
}

extern "C" xf_status xf_engine_get_nffft(xf_engine *e, double *out) {
  if (!e || !out || !e->has_nf) FAILS(XF_ESTATE, "no NF-FF surface");
  HIPCHK(engine_wait(e));
  HIPCHK(hipMemcpy(out, e->acc, (size_t)e->nf_size * sizeof(double2), hipMemcpyDeviceToHost));
  // y pair: 8 currents x nf frequencies, each a [nxb][nzb] complex block
  const Nffft &n = e->nf;
  y_faces_to_interface(out + 2 * n.base[1], 8LL * n.nf, n.nxb, n.nzb, 2);
  return XF_OK;
}

// direction cosines on the host with the same libm as the reference
// (nffft_fd.cpp:383-388): tab[0, nd) sin(t) cos(p), [nd, 2 nd) sin(t) sin(p),
// [2 nd, 3 nd) cos(t), then dl * k per frequency
static void far_table(const xf_engine *e, double dl, const double *wave_number, const double *theta, int ntheta,
                      const double *phi, int nphi, double *tab) {
  const int nd = ntheta * nphi;
  for (int t = 0; t < ntheta; ++t)
    for (int p = 0; p < nphi; ++p) {
      const double cos_t = cos(theta[t]), sin_t = sin(theta[t]);
      const double cos_p = cos(phi[p]), sin_p = sin(phi[p]);
      tab[t * nphi + p] = sin_t * cos_p;
      tab[nd + t * nphi + p] = sin_t * sin_p;
      tab[2 * nd + t * nphi + p] = cos_t;
    }
  for (int f = 0; f < e->nf.nf; ++f) tab[3 * nd + f] = dl * wave_number[f];
}

static void far_launch(xf_engine *e, const double *dtab, int ntheta, int nphi, double2 *dout) {
  const int nd = ntheta * nphi;
  FarDir fd;
  fd.stcp = dtab; fd.stsp = dtab + nd; fd.ct = dtab + 2 * nd; fd.dk = dtab + 3 * nd; fd.nt = ntheta; fd.np = nphi;
  hipLaunchKernelGGL(k_farfield, dim3(nd, e->nf.nf), dim3(256), 0, e->stream, e->nf, fd, dout);
}

extern "C" xf_status xf_engine_farfield_sums(xf_engine *e, double dl, const double *wave_number,
                                             const double *theta, int ntheta, const double *phi,
                                             int nphi, double *out) {
  if (!e || !out || !e->has_nf || !wave_number || !theta || !phi || ntheta < 1 || nphi < 1)
    FAILS(XF_EINVAL, "bad far-field request");
  const int nd = ntheta * nphi;
  std::vector<double> tab(3 * (size_t)nd + e->nf.nf);
  far_table(e, dl, wave_number, theta, ntheta, phi, nphi, tab.data());
  double *dtab = nullptr;
  double2 *dout = nullptr;
  const size_t outn = (size_t)e->nf.nf * nd * 12;
  HIPCHK(hipMalloc((void **)&dtab, tab.size() * sizeof(double)));
  HIPCHK(hipMalloc((void **)&dout, outn * sizeof(double2)));
  HIPCHK(hipMemcpy(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  far_launch(e, dtab, ntheta, nphi, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(engine_wait(e));
  HIPCHK(hipMemcpy(out, dout, outn * sizeof(double2), hipMemcpyDeviceToHost));
  hipFree(dtab);
  hipFree(dout);
  return XF_OK;
}

extern "C" xf_status xf_engine_farfield_sums_noalloc(xf_engine *e, double dl, const double *wave_number,
                                                     const double *theta, int ntheta, const double *phi,
                                                     int nphi, double *out) {
  if (!e || !out || !e->has_nf || !wave_number || !theta || !phi || ntheta < 1 || nphi < 1)
    FAILS(XF_EINVAL, "bad far-field request");
  const size_t nd = (size_t)ntheta * (size_t)nphi, nf = (size_t)e->nf.nf;
  const size_t tab_n = 3 * nd + nf;
  const size_t tab_bytes = (tab_n * sizeof(double) + 255) / 256 * 256;
  const size_t out_bytes = nf * nd * 12 * sizeof(double2);
  if (tab_bytes + out_bytes > kFarScratchBytes)
    FAILS(XF_EINVAL,
          "far field of %zu directions x %zu frequencies needs %zu bytes of scratch; the allocation-free entry "
          "point holds %zu",
          nd, nf, tab_bytes + out_bytes, kFarScratchBytes);
  void *base = g_scratch_dev[e->dev];
  if (!base) FAILS(XF_ESTATE, "no scratch address for device %d", e->dev);
  double *dtab = (double *)base;
  double2 *dout = (double2 *)((char *)base + tab_bytes);
  HIPCHK(engine_wait(e));
  far_table(e, dl, wave_number, theta, ntheta, phi, nphi, g_far_host);
  HIPCHK(hipMemcpyAsync(dtab, g_far_host, tab_n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  far_launch(e, dtab, ntheta, nphi, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return XF_OK;
}

// ---------------------------------------------------------------- broadband NF-FF

static long long bb_cells(const Nffft &n) { return nf_cells(n); }

extern "C" xf_status xf_engine_set_nffft_broadband(xf_engine *e, const int64_t box[6], int64_t total_steps) {
  if (!e || !box) FAILS(XF_EINVAL, "null argument");
  if (e->has_bb) FAILS(XF_ESTATE, "broadband NF-FF surface already set");
  // a 2-D grid (nz == 1) takes the box's one z layer [0, 1) and keeps no
  // z faces (nffft_broadband.cpp:438-440 never integrates them)
  const bool two_d = e->nz == 1;
  const bool zbad = two_d ? (box[4] != 0 || box[5] != 1) : (box[4] < 1 || box[5] >= e->nz || box[5] <= box[4]);
  if (box[0] < 1 || box[2] < 1 || box[1] >= e->nx || box[3] >= e->ny || box[1] <= box[0] || box[3] <= box[2] ||
      zbad || total_steps < 1)
    FAILS(XF_EINVAL, "broadband NF-FF box outside the grid or no steps");
  Nffft &n = e->bb;
  n = Nffft{};
  n.li = (int)box[0]; n.ri = (int)box[1]; n.lj = (int)box[2]; n.rj = (int)box[3];
  n.lk = (int)box[4]; n.rk = (int)box[5];
  n.nxb = n.ri - n.li; n.nyb = n.rj - n.lj; n.nzb = n.rk - n.lk;
  n.zf = two_d ? 0 : 1;
  n.T_total = total_steps;
  const double bytes = (double)total_steps * 8 * bb_cells(n) * e->esz;
  xf_status s = dev_alloc(e, &e->bb_hist, (size_t)bytes);
  if (s != XF_OK) return s;
  e->bb_T = total_steps;
  set_total(e, total_steps);
  e->has_bb = true;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_broadband_scratch(xf_engine *e, int64_t bytes) {
  if (!e || bytes < 1) FAILS(XF_EINVAL, "bad scratch size");
  e->bb_scratch = bytes;
  return XF_OK;
}

extern "C" int64_t xf_engine_broadband_samples_size(const xf_engine *e) {
  return e && e->has_bb ? e->bb_T * 8 * bb_cells(e->bb) : 0;
}

extern "C" xf_status xf_engine_get_broadband_samples(xf_engine *e, double *out) {
  if (!e || !out || !e->has_bb) FAILS(XF_ESTATE, "no broadband NF-FF surface");
  const size_t n = (size_t)xf_engine_broadband_samples_size(e);
  HIPCHK(engine_wait(e));
  if (e->prec == 64) {
    HIPCHK(hipMemcpy(out, e->bb_hist, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), e->bb_hist, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) out[q] = (double)tmp[q];
  }
  // [step][8][Q]: the y faces of each (step, current) row in interface order
  const Nffft &b = e->bb;
  const long long Q = bb_cells(b), cx = (long long)b.nyb * b.nzb;
  for (long long r = 0; r < e->bb_T * 8; ++r) y_faces_to_interface(out + r * Q + cx, 1, b.nxb, b.nzb, 1);
  return XF_OK;
}

// the surface integrals of the broadband far field: series of `chunk`
// consecutive (current, cell) pairs are transposed into FP64 rows of Ns
// points (zero past step T-1), transformed by one batched real-to-complex
// FFT, and folded into per-slice partial sums of every bin; the partials
// are added in slice order at the end
extern "C" xf_status xf_engine_broadband_sums(xf_engine *e, double dl, const double *wave_number, double theta,
                                              double phi, double *out) {
  if (!e || !wave_number || !out || !e->has_bb) FAILS(XF_EINVAL, "bad broadband far-field request");
  const long long T = e->bb_T, Ns = 2 * T - 1, Q = bb_cells(e->bb), S = 8 * Q;
  if (Ns > 0x7fffffff) FAILS(XF_EINVAL, "too many steps for one transform");
  HIPCHK(engine_wait(e));
  BbFar d{};
  d.n = e->bb;
  d.T = T;
  d.Ns = Ns;
  d.Q = Q;
  // direction and centre on the host with the reference's libm
  // (nffft_broadband.cpp:27-28, grid/grid_box.cpp:101-105)
  d.rx = sin(theta) * cos(phi);
  d.ry = sin(theta) * sin(phi);
  d.rz = cos(theta);
  d.cx = (double)(d.n.li + d.n.nxb / 2);
  d.cy = (double)(d.n.lj + d.n.nyb / 2);
  d.cz = (double)(d.n.lk + d.n.nzb / 2);
  d.dl = dl;
  d.ds = dl * dl;
  d.zfaces = d.n.nzb > 1;
  // chunk of series per transform: bb_scratch bytes of rows and spectra
  const long long per = Ns * 8 + T * 16;
  long long chunk = std::max<long long>(1, std::min<long long>(S, e->bb_scratch / per));
  chunk = std::min<long long>(chunk, 65535LL * 64);  // grid y of the row transpose
  const int slices = (int)std::min<long long>(256, chunk);
  AdjBufs B(e->stream);
  double *rows = (double *)B.alloc((size_t)chunk * Ns * sizeof(double));
  double2 *spec = (double2 *)B.alloc((size_t)chunk * T * sizeof(double2));
  double2 *part = (double2 *)B.alloc((size_t)slices * 6 * Ns * sizeof(double2));
  double *dk = (double *)B.alloc((size_t)Ns * sizeof(double));
  double2 *dout = (double2 *)B.alloc((size_t)6 * Ns * sizeof(double2));
  if (!rows || !spec || !part || !dk || !dout) FAILS(XF_ENOMEM, "broadband far-field buffers");
  HIPCHK(hipMemcpy(dk, wave_number, (size_t)Ns * sizeof(double), hipMemcpyHostToDevice));
  d.k = dk;
  hipfftHandle plan;
  int nfft = (int)Ns;
  if (hipfftPlanMany(&plan, 1, &nfft, nullptr, 1, (int)Ns, nullptr, 1, (int)T, HIPFFT_D2Z, (int)chunk) !=
      HIPFFT_SUCCESS)
    FAILS(XF_EDEVICE, "hipfftPlanMany(%d, batch %lld) failed", nfft, chunk);
  struct PlanFree {
    hipfftHandle p;
    ~PlanFree() { hipfftDestroy(p); }
  } plan_free{plan};
  if (hipfftSetStream(plan, e->stream) != HIPFFT_SUCCESS) FAILS(XF_EDEVICE, "hipfftSetStream failed");
  for (long long s0 = 0; s0 < S; s0 += chunk) {
    const long long n = std::min<long long>(chunk, S - s0);
    const dim3 tg((unsigned)((Ns + 63) / 64), (unsigned)((chunk + 63) / 64));
    if (e->prec == 64)
      hipLaunchKernelGGL(k_bb_rows<double>, tg, dim3(64, 4), 0, e->stream, (const double *)e->bb_hist, S, T, Ns,
                         s0, n, chunk, rows);
    else
      hipLaunchKernelGGL(k_bb_rows<float>, tg, dim3(64, 4), 0, e->stream, (const float *)e->bb_hist, S, T, Ns, s0,
                         n, chunk, rows);
    HIPCHK(hipGetLastError());
    if (hipfftExecD2Z(plan, rows, reinterpret_cast<hipfftDoubleComplex *>(spec)) != HIPFFT_SUCCESS)
      FAILS(XF_EDEVICE, "hipfftExecD2Z failed");
    hipLaunchKernelGGL(k_bb_fold, dim3((unsigned)((Ns + 255) / 256), (unsigned)slices), dim3(256), 0, e->stream, d,
                       (const double2 *)spec, s0, n, slices, part, s0 == 0 ? 1 : 0);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_bb_sum_slices, dim3((unsigned)((6 * Ns + 255) / 256)), dim3(256), 0, e->stream,
                     (const double2 *)part, slices, 6 * Ns, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(out, dout, (size_t)6 * Ns * sizeof(double2), hipMemcpyDeviceToHost));
  return XF_OK;
}

extern "C" xf_status xf_engine_add_current_source(xf_engine *e, int comp, int64_t i0, int64_t i1,
                                                  int64_t j0, int64_t j1, int64_t k0, int64_t k1,
                                                  const double *coff, const double *wave,
                                                  int64_t total_steps) {
  if (!e || !coff || !wave || comp < XF_EX || comp > XF_EZ) FAILS(XF_EINVAL, "bad source");
  int sh[3];
  comp_shape(e, comp, sh);
  if (i0 < 0 || j0 < 0 || k0 < 0 || i1 > sh[0] || j1 > sh[1] || k1 > sh[2] || i1 <= i0 ||
      j1 <= j0 || k1 <= k0)
    FAILS(XF_EINVAL, "source box outside the component array");
  SourceDev sd;
  sd.comp = comp; sd.i0 = (int)i0; sd.i1 = (int)i1; sd.j0 = (int)j0; sd.j1 = (int)j1;
  sd.k0 = (int)k0; sd.k1 = (int)k1;
  const size_t nc = (size_t)(i1 - i0) * (j1 - j0) * (k1 - k0);
  xf_status s;
  if ((s = upload_real(e, &sd.coff, coff, nc)) != XF_OK) return s;
  if ((s = upload_real(e, &sd.wave, wave, (size_t)total_steps)) != XF_OK) return s;
  e->sources.push_back(sd);
  set_total(e, total_steps);
  return XF_OK;
}

extern "C" xf_status xf_engine_add_inductor(xf_engine *e, int comp, int64_t i0, int64_t i1, int64_t j0,
                                            int64_t j1, int64_t k0, int64_t k1, const double *cecjc,
                                            const double *cjcec) {
  if (!e || !cecjc || !cjcec || comp < XF_EX || comp > XF_EZ) FAILS(XF_EINVAL, "bad inductor");
  int sh[3];
  comp_shape(e, comp, sh);
  if (i0 < 0 || j0 < 0 || k0 < 0 || i1 > sh[0] || j1 > sh[1] || k1 > sh[2] || i1 <= i0 || j1 <= j0 || k1 <= k0)
    FAILS(XF_EINVAL, "inductor box outside the component array");
  SourceDev sd;
  sd.comp = comp; sd.i0 = (int)i0; sd.i1 = (int)i1; sd.j0 = (int)j0; sd.j1 = (int)j1;
  sd.k0 = (int)k0; sd.k1 = (int)k1;
  sd.kind = 1;
  const size_t nc = (size_t)(i1 - i0) * (j1 - j0) * (k1 - k0);
  xf_status s;
  if ((s = upload_real(e, &sd.coff, cecjc, nc)) != XF_OK) return s;
  if ((s = upload_real(e, &sd.c2, cjcec, nc)) != XF_OK) return s;
  if ((s = dev_alloc(e, &sd.jc, nc * e->esz)) != XF_OK) return s;
  e->state.push_back({sd.jc, nullptr, nc * e->esz});
  e->sources.push_back(sd);
  return XF_OK;
}

extern "C" xf_status xf_engine_add_line_monitor(xf_engine *e, int nterms, const int *comp, const int64_t *cell_a,
                                                const int64_t *cell_b, const double *w, const int *group,
                                                double sign, int64_t total_steps, int *index) {
  if (!e || nterms < 1 || !comp || !cell_a || !cell_b || !w || !group || !index || total_steps < 1)
    FAILS(XF_EINVAL, "bad line monitor");
  std::vector<LineTerm> t((size_t)nterms);
  for (int q = 0; q < nterms; ++q) {
    if (comp[q] < 0 || comp[q] > 5 || (group[q] != 0 && group[q] != 1))
      FAILS(XF_EINVAL, "bad line monitor term %d", q);
    int sh[3];
    comp_shape(e, comp[q], sh);
    const int64_t *a = cell_a + 3 * q, *b = cell_b + 3 * q;
    if (a[0] < 0 || a[1] < 0 || a[2] < 0 || a[0] >= sh[0] || a[1] >= sh[1] || a[2] >= sh[2])
      FAILS(XF_EINVAL, "line monitor term %d outside the component array", q);
    const bool hb = b[0] >= 0;
    if (hb && (b[1] < 0 || b[2] < 0 || b[0] >= sh[0] || b[1] >= sh[1] || b[2] >= sh[2]))
      FAILS(XF_EINVAL, "line monitor term %d outside the component array", q);
    t[q].ca = a[0] * e->si + a[1] * e->sj + a[2];
    t[q].cb = hb ? b[0] * e->si + b[1] * e->sj + b[2] : -1;
    t[q].w = w[q];
    t[q].comp = comp[q];
    t[q].group = group[q];
  }
  LineMonDev m;
  xf_status s;
  if ((s = dev_alloc(e, &m.terms, t.size() * sizeof(LineTerm))) != XF_OK) return s;
  HIPCHK(hipMemcpy(m.terms, t.data(), t.size() * sizeof(LineTerm), hipMemcpyHostToDevice));
  void *p;
  if ((s = dev_alloc(e, &p, (size_t)total_steps * sizeof(double))) != XF_OK) return s;
  m.buf = (double *)p;
  m.n = nterms;
  m.sign = sign;
  m.T = total_steps;
  set_total(e, total_steps);
  e->line_mons.push_back(m);
  *index = (int)e->line_mons.size() - 1;
  return XF_OK;
}

extern "C" xf_status xf_engine_get_line_monitors(xf_engine *e, double *out, int64_t total_steps) {
  if (!e || !out) FAILS(XF_EINVAL, "null argument");
  HIPCHK(engine_wait(e));
  for (size_t q = 0; q < e->line_mons.size(); ++q) {
    if (e->line_mons[q].T != total_steps) FAILS(XF_ESTATE, "line monitor %zu holds %lld steps", q, e->line_mons[q].T);
    HIPCHK(hipMemcpy(out + q * (size_t)total_steps, e->line_mons[q].buf, (size_t)total_steps * sizeof(double),
                     hipMemcpyDeviceToHost));
  }
  return XF_OK;
}

extern "C" int xf_engine_line_monitor_count(const xf_engine *e) { return e ? (int)e->line_mons.size() : 0; }

extern "C" xf_status xf_engine_add_probe(xf_engine *e, int comp, int64_t i, int64_t j, int64_t k) {
  if (!e || comp < 0 || comp > 5) FAILS(XF_EINVAL, "bad probe");
  int sh[3];
  comp_shape(e, comp, sh);
  if (i < 0 || j < 0 || k < 0 || i >= sh[0] || j >= sh[1] || k >= sh[2])
    FAILS(XF_EINVAL, "probe outside the component array");
  e->probe_comp.push_back(comp);
  e->probe_cell.push_back((long long)i * e->si + (long long)j * e->sj + k);
  e->probes_dirty = true;
  return XF_OK;
}

static xf_status ensure_probes(xf_engine *e, long long T) {
  if (e->probe_comp.empty()) return XF_OK;
  if (!e->probes_dirty && e->probe_T >= T) return XF_OK;
  const size_t n = e->probe_comp.size();
  void *p;
  xf_status s;
  if ((s = dev_alloc(e, &p, n * sizeof(int))) != XF_OK) return s;
  e->d_probe_comp = (int *)p;
  if ((s = dev_alloc(e, &p, n * sizeof(long long))) != XF_OK) return s;
  e->d_probe_cell = (long long *)p;
  if ((s = dev_alloc(e, &p, n * (size_t)T * sizeof(double))) != XF_OK) return s;
  e->probe_buf = (double *)p;
  HIPCHK(hipMemcpy(e->d_probe_comp, e->probe_comp.data(), n * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->d_probe_cell, e->probe_cell.data(), n * sizeof(long long), hipMemcpyHostToDevice));
  e->probe_T = T;
  e->probes_dirty = false;
  return XF_OK;
}

extern "C" xf_status xf_engine_get_probes(xf_engine *e, double *out, int64_t total_steps) {
  if (!e || !out) FAILS(XF_EINVAL, "null argument");
  if (e->probe_comp.empty()) return XF_OK;
  if (e->probe_buf == nullptr || total_steps != e->probe_T) FAILS(XF_ESTATE, "probe buffer holds %lld steps", e->probe_T);
  HIPCHK(engine_wait(e));
  HIPCHK(hipMemcpy(out, e->probe_buf, e->probe_comp.size() * (size_t)total_steps * sizeof(double),
                   hipMemcpyDeviceToHost));
  return XF_OK;
}

template <typename T>
static Grid<T> make_grid(xf_engine *e) {
  Grid<T> g;
  for (int q = 0; q < 6; ++q) g.f[q] = (T *)e->f[q];
  g.mat = e->mat;
  for (int side = 0; side < 2; ++side) {
    g.self[side] = (const T *)e->self[side];
    g.kt[side] = (const T *)e->kt[side];
    g.jt[side] = (const T *)e->jt[side];
    g.it[side] = (const T *)e->it[side];
    g.kp[side] = (const T *)e->kp[side];
    g.jp[side] = (const T *)e->jp[side];
    g.ip[side] = (const T *)e->ip[side];
  }
  for (int a = 0; a < 3; ++a) {
    g.dn[a] = e->dn[a]; g.dp[a] = e->dp[a];
    g.be[a] = (const T *)e->be[a]; g.ae[a] = (const T *)e->ae[a];
    g.bm[a] = (const T *)e->bm[a]; g.am[a] = (const T *)e->am[a];
  }
  for (int q = 0; q < 6; ++q) {
    g.psi_e[q] = (T *)e->psi_e[q];
    g.psi_h[q] = (T *)e->psi_h[q];
  }
  g.nx = e->nx; g.ny = e->ny; g.nz = e->nz; g.pz = e->pz; g.nmat = e->nmat;
  g.h_uni = e->h_uni ? 1 : 0;
  g.sj = e->sj; g.si = e->si;
  for (int a = 0; a < 6; ++a) g.db[a] = e->db[a];
  g.dcoef = (const T *)e->dcoef;
  g.snap = (T *)e->snap;
  for (int a = 0; a < 6; ++a) g.xb[a] = e->xb[a];
  g.drc = (const T *)e->drc;
  for (int q = 0; q < 3; ++q) {
    g.dj[q] = (T *)e->dj[q];
    g.djp[q] = (T *)e->djp[q];
    g.dep[q] = (T *)e->dep[q];
  }
  g.div = (unsigned long long *)e->div;
  return g;
}

// Do the curl kernels apply the TFSF z-face corrections (ZFaces)?  On 3-D
// grids without lumped E corrections, which the reference applies before the
// TFSF ones (the face kernels run after them); batches keep every face in the
// face kernels.
static bool tfsf_zc(const xf_engine *e) { return e->has_tfsf && !e->tfsf_two_d && e->sources.empty(); }

// Cell counts of the face kernels' segments (tfsf_e_body, tfsf_h_body).
// TFSF2D (tfsf_2d.cpp:148-228) corrects Ez on the four faces of its one z
// layer, Hy on the x faces and Hx on the y faces.
static void tfsf_segs(const xf_engine *e, bool zc, TfsfESeg &es, TfsfHSeg &hs) {
  const int ni = e->tri - e->tli, nj = e->trj - e->tlj, nk = e->trk - e->tlk;
  const bool two_d = e->tfsf_two_d;
  int ce[6] = {2 * nj * (nk + 1), zc ? 0 : 2 * (ni - 1) * nj, 2 * (nj + 1) * nk, 2 * (ni - 1) * nk,
               2 * ni * (nk + 1), zc ? 0 : 2 * ni * (nj - 1)};
  int ch[12] = {nj * (nk + 1), (nj + 1) * nk, nj * (nk + 1), (nj + 1) * nk,
                nk * (ni + 1), (nk + 1) * ni, nk * (ni + 1), (nk + 1) * ni,
                ni * (nj + 1), (ni + 1) * nj, ni * (nj + 1), (ni + 1) * nj};
  if (two_d) {
    ce[0] = ce[1] = ce[4] = ce[5] = 0;
    for (int q : {0, 2, 5, 7, 8, 9, 10, 11}) ch[q] = 0;
  }
  if (zc)
    for (int q = 8; q < 12; ++q) ch[q] = 0;
  es.off[0] = 0;
  for (int q = 0; q < 6; ++q) es.off[q + 1] = es.off[q] + ce[q];
  hs.off[0] = 0;
  for (int q = 0; q < 12; ++q) hs.off[q + 1] = hs.off[q] + ch[q];
}

static Tfsf make_tfsf(xf_engine *e) {
  Tfsf t;
  std::memset(&t, 0, sizeof t);  // padding included: records are compared bytewise (upload_tz)
  t.li = e->tli; t.ri = e->tri; t.lj = e->tlj; t.rj = e->trj; t.lk = e->tlk; t.rk = e->trk; t.L = e->L;
  t.two_d = e->tfsf_two_d ? 1 : 0;
  for (int a = 0; a < 3; ++a) {
    t.pf[a] = (const double *)e->pf[a];
    t.ph[a] = (const double *)e->ph[a];
    t.te[a] = e->te[a];
    t.th[a] = e->th[a];
    t.ca[a] = e->tca[a];
    t.cb[a] = e->tcb[a];
  }
  t.ceie = e->ceie; t.ceihi = e->ceihi; t.chih = e->chih; t.chiei = e->chiei;
  t.c1 = e->c1; t.c2 = e->c2;
  t.e_inc = (double *)e->e_inc; t.h_inc = (double *)e->h_inc; t.ab = (double *)e->ab;
  t.src = (const double *)e->src;
  return t;
}



// launch geometry of the curl kernels from the engine's tiling settings
static Tiling make_tiling(const xf_engine *e) {
  Tiling tl;
  tl.tz = e->tz;
  tl.ty = e->ty;
  tl.tx = e->tx > 0 ? e->tx : e->nx + 1;
  // a tile's tx + 1 planes stay within the 4 GiB its 32-bit offsets reach
  // (results do not depend on tx)
  const long long cap = (long long)(((1ULL << 32) - 1) / ((unsigned long long)e->si * e->esz)) - 1;
  if (tl.tx > cap) tl.tx = (int)std::max(1LL, cap);
  if (tl.ty == 0) {
    // flat runs of tz cells over the padded (j,k) plane
    const long long plane = (long long)(e->ny + 1) * e->pz;
    tl.ntz = (int)((plane + tl.tz - 1) / tl.tz);
    tl.nty = 1;
    tl.krows = tl.tz;
    tl.jrows = (tl.tz + e->pz - 1) / e->pz + 1;
  } else {
    tl.ntz = (e->nz + 1 + tl.tz - 1) / tl.tz;
    tl.nty = (e->ny + 1 + tl.ty - 1) / tl.ty;
    tl.krows = tl.tz;
    tl.jrows = tl.ty;
  }
  tl.ntx = (e->nx + 1 + tl.tx - 1) / tl.tx;
  tl.nblocks = tl.ntz * tl.nty * tl.ntx;
  tl.xcd_cols = (e->tile_flags & XF_TILE_XCD_COLS) != 0;
  if (tl.xcd_cols) tl.nblocks = 8 * tl.ntz * ((tl.nty * tl.ntx + 7) / 8);
  tl.x_fastest = (e->tile_flags & XF_TILE_X_FASTEST) != 0;
  tl.xcd_runs = (e->tile_flags & XF_TILE_XCD_RUNS) != 0;
  tl.oz = tl.oy = tl.ox = 0;
  return tl;
}

// Narrow a tiling to the box of tiles that meets the engine's cell window
// (flat tiles: the runs that meet its j rows, every k).  Results inside the
// window do not depend on the tiles, so only the cells outside it change.
static void window_tiling(const xf_engine *e, Tiling &tl) {
  const int *w = e->win;
  auto span = [](long long lo, long long hi, int size, int count, int &first, int &n) {
    const long long a = std::max(0LL, lo / size), b = std::min((long long)count, (hi + size - 1) / size);
    first = (int)a;
    n = (int)std::max(1LL, b - a);
  };
  span(w[0], w[1], tl.tx, tl.ntx, tl.ox, tl.ntx);
  if (tl.ty == 0) {
    tl.oy = 0;
    span((long long)w[2] * e->pz, (long long)w[3] * e->pz, tl.tz, tl.ntz, tl.oz, tl.ntz);
  } else {
    span(w[2], w[3], tl.ty, tl.nty, tl.oy, tl.nty);
    span(w[4], w[5], tl.tz, tl.ntz, tl.oz, tl.ntz);
  }
  tl.nblocks = tl.xcd_cols ? 8 * tl.ntz * ((tl.nty * tl.ntx + 7) / 8) : tl.ntz * tl.nty * tl.ntx;
}

// Whether the split E kernel can write a step's design-box history itself:
// nothing launched after it changes E in the box (no TFSF face plane, with
// a cell to spare, and no lumped element meets the box) and the step is
// not the fused one.
static bool fused_ok(const xf_engine *e);
static bool e_kernel_snaps(const xf_engine *e) {
  if (fused_ok(e) || !e->has_design) return false;
  auto meets = [](int lo, int hi, int a, int b) { return lo <= b && hi > a; };  // [lo, hi) and [a, b]
  const int *d = e->db;
  if (e->has_tfsf) {
    const int f[3][2] = {{e->tli, e->tri}, {e->tlj, e->trj}, {e->tlk, e->trk}};
    for (int a = 0; a < 3; ++a)
      for (int s = 0; s < 2; ++s)
        if (meets(d[2 * a], d[2 * a + 1], f[a][s] - 1, f[a][s] + 1)) return false;
  }
  for (const SourceDev &sd : e->sources)
    if (meets(d[0], d[1], sd.i0 - 1, sd.i1) && meets(d[2], d[3], sd.j0 - 1, sd.j1) &&
        meets(d[4], d[5], sd.k0 - 1, sd.k1))
      return false;
  return true;
}

// The cells a replayed gradient step has to get right, one to spare: the
// design box grown by m cells along each axis, clipped to the grid.
static void replay_window(const xf_engine *e, int m, int w[6]) {
  const int hi[3] = {e->nx + 1, e->ny + 1, e->nz + 1};
  for (int d = 0; d < 3; ++d) {
    w[2 * d] = std::max(0, e->db[2 * d] - m);
    w[2 * d + 1] = std::min(hi[d], e->db[2 * d + 1] + m);
  }
}

// the smallest m whose replay window is the whole grid
static int replay_window_full(const xf_engine *e) {
  const int hi[3] = {e->nx + 1, e->ny + 1, e->nz + 1};
  int m = 0;
  for (int d = 0; d < 3; ++d) m = std::max(m, std::max(e->db[2 * d], hi[d] - e->db[2 * d + 1]));
  return m;
}

// Whether workgroup b of a curl launch over tiling tl (no tile origin)
// covers a cell of the window w; flat tiles count by their j rows.
static bool block_meets_window(const xf_engine *e, const Tiling &tl, int b, const int w[6]) {
  int bz, by, bx;
  if (!tile_of_block(tl, b, bz, by, bx)) return false;
  const long long i0 = (long long)bx * tl.tx;
  if (i0 >= w[1] || i0 + tl.tx <= w[0]) return false;
  if (tl.ty == 0) {
    const long long p0 = (long long)bz * tl.tz, j0 = p0 / e->pz, j1 = (p0 + tl.tz - 1) / e->pz;
    return j0 < w[3] && j1 >= w[2];
  }
  const long long j0 = (long long)by * tl.ty, k0 = (long long)bz * tl.tz;
  return j0 < w[3] && j0 + tl.ty > w[2] && k0 < w[5] && k0 + tl.tz > w[4];
}

static int tiling_threads(const Tiling &tl) { return tl.ty == 0 ? tl.tz : tl.tz * tl.ty; }

static bool has_non_drude(const xf_engine *e) { return e->ade_non_drude; }

static bool has_inductor(const xf_engine *e) {
  for (const SourceDev &sd : e->sources)
    if (sd.kind == 1) return true;
  return false;
}

// lumped E corrections in insertion order (simulation_update.cpp:282-290)
template <typename T>
static void launch_lumped(const xf_engine *e, const SourceDev &sd, T *f, hipStream_t st, long long step) {
  const int nj = sd.j1 - sd.j0, nk = sd.k1 - sd.k0;
  const int nc = (sd.i1 - sd.i0) * nj * nk;
  if (sd.kind == 1)
    hipLaunchKernelGGL(k_inductor<T>, dim3((nc + 255) / 256), dim3(256), 0, st, f, e->si, e->sj, sd.i0, nj, nk,
                       sd.j0, sd.k0, nc, (const T *)sd.coff, (const T *)sd.c2, (T *)sd.jc);
  else
    hipLaunchKernelGGL(k_source<T>, dim3((nc + 255) / 256), dim3(256), 0, st, f, e->si, e->sj, sd.i0, nj, nk,
                       sd.j0, sd.k0, nc, (const T *)sd.coff, (const T *)sd.wave, step);
}

// voltage and current monitors of the step (simulation_update.cpp:215-219)
template <typename T>
static void launch_line_monitors(const xf_engine *e, const Grid<T> &g, hipStream_t st, long long step) {
  for (const LineMonDev &m : e->line_mons)
    if (step < m.T)
      hipLaunchKernelGGL(k_line_monitor<T>, dim3(1), dim3(64), 0, st, g, (const LineTerm *)m.terms, m.n, m.sign,
                         m.buf, step);
}

// can this engine take the fused E+H step?  It covers Yee, CPML and TFSF;
// design boxes, dispersive media and lumped elements need the split
// kernels, and so do 2-D grids, engines in a batch (a batch's launches hold
// the member's field pointers) and fields over 4 GiB (the fused kernel
// addresses a whole component with 32-bit offsets).
static int sweep_rows(const xf_engine *e) {
  if (e->sweep_rows > 0) return e->sweep_rows;
  return 2;
}

static int sweep_tx(const xf_engine *e) { return e->sweep_tx > 0 ? e->sweep_tx : 16; }

// material slots per LDS record of the fused step (4 or 8; more materials
// take the split kernels)
static int sweep_ms(const xf_engine *e) { return e->nmat <= 4 ? 4 : 8; }

// LDS bytes of a fused-step workgroup (its coefficient and CPML tables)
static size_t sweep_lds_bytes(const xf_engine *e) {
  const int R = sweep_rows(e), ms = sweep_ms(e);
  int t;
  if (ms == 4) t = R == 1 ? SweepLds<1, 4>::total : R == 2 ? SweepLds<2, 4>::total : SweepLds<4, 4>::total;
  else t = R == 1 ? SweepLds<1, 8>::total : R == 2 ? SweepLds<2, 8>::total : SweepLds<4, 8>::total;
  return (size_t)t * e->esz;
}

static bool fused_ok(const xf_engine *e) {
  return e->sweep == 1 && e->nz > 1 && (unsigned long long)e->ncell * e->esz < (1ULL << 32) && !e->has_design && !e->has_drude && e->sources.empty() &&
         e->batch_stream == nullptr && e->nmat <= 8 && sweep_tx(e) <= SW_TXMAX &&
         sweep_lds_bytes(e) <= 60 * 1024;
}

// the second copy of the incident line
static xf_status alloc_lines_alt(xf_engine *e) {
  if (!e->has_tfsf || e->e_inc_alt) return XF_OK;
  xf_status s;
  if ((s = dev_alloc(e, &e->e_inc_alt, (size_t)e->L * sizeof(double))) != XF_OK) return s;
  if ((s = dev_alloc(e, &e->h_inc_alt, (size_t)(e->L - 1) * sizeof(double))) != XF_OK) return s;
  return XF_OK;
}

// Does index n of an axis of N cells need the CPML or TFSF code of the fused
// step: as an E index (E layers [1, dn] and [N-dp, N-1], TFSF E faces lo and
// hi) or as an H index (H layers [0, dn) and [N-dp, N), TFSF H faces lo - 1
// and hi)?
static bool e_special(int n, int N, int dn, int dp, bool tf, int lo, int hi) {
  return (n >= 1 && n <= dn) || (dp > 0 && n >= N - dp && n <= N - 1) || (tf && (n == lo || n == hi));
}
static bool h_special(int n, int N, int dn, int dp, bool tf, int lo, int hi) {
  return (n >= 0 && n < dn) || (dp > 0 && n >= N - dp && n < N) || (tf && (n == lo - 1 || n == hi));
}

// Chunks of at most `len` indices over [0, n): runs of indices that need the
// special code (as E or H index, or whose successor needs it as an E index,
// since a chunk also computes E one index past its end) and runs that do not
// are chunked separately, so that plain chunks stay plain.
static std::vector<std::pair<int, int>> sweep_chunks(int n, int len, int N, int dn, int dp, bool tf, int lo, int hi) {
  std::vector<std::pair<int, int>> out;
  auto bad = [&](int q) {
    return e_special(q, N, dn, dp, tf, lo, hi) || h_special(q, N, dn, dp, tf, lo, hi) ||
           e_special(q + 1, N, dn, dp, tf, lo, hi);
  };
  int a = 0;
  while (a < n) {
    const bool b0 = bad(a);
    int b = a + 1;
    while (b < n && bad(b) == b0) ++b;
    const int m = (b - a + len - 1) / len;  // chunks of this run, balanced
    for (int q = 0; q < m; ++q) out.push_back({a + (int)((long long)(b - a) * q / m), a + (int)((long long)(b - a) * (q + 1) / m)});
    a = b;
  }
  return out;
}

// Build the fused step's tile lists for R rows per wave and tx planes per
// chunk: x chunks, j groups of SW_WAVES * R rows and 64-cell k chunks, each
// tile sorted into the kind whose code it needs.
static xf_status build_sweep_tiles(xf_engine *e) {
  const int R = sweep_rows(e), tx = sweep_tx(e);
  if (e->sp_tiles && e->sp_R == R && e->sp_tx == tx) return XF_OK;
  const bool tf = e->has_tfsf;
  const int G = SW_WAVES * R;
  const auto xc = sweep_chunks(e->nx + 1, tx, e->nx, e->dn[0], e->dp[0], tf, e->tli, e->tri);
  const auto yc = sweep_chunks(e->ny + 1, G, e->ny, e->dn[1], e->dp[1], tf, e->tlj, e->trj);
  const int nkc = (e->nz + 1 + 63) / 64;
  std::vector<SweepTile> lists[3];
  for (const auto &x : xc)
    for (const auto &y : yc) {
      // x: E planes [i0, i1] (the last one a halo plane), H planes [i0, i1);
      // y: E rows up to the top wave's halo row, H rows below it
      bool sx = false, sy = false;
      for (int p = x.first; p <= x.second; ++p)
        sx |= e_special(p, e->nx, e->dn[0], e->dp[0], tf, e->tli, e->tri) ||
              (p < x.second && h_special(p, e->nx, e->dn[0], e->dp[0], tf, e->tli, e->tri));
      const int rows = y.second - y.first, top = y.first + (rows + R - 1) / R * R;
      for (int j = y.first; j <= top; ++j)
        sy |= e_special(j, e->ny, e->dn[1], e->dp[1], tf, e->tlj, e->trj) ||
              (j < top && h_special(j, e->ny, e->dn[1], e->dp[1], tf, e->tlj, e->trj));
      for (int kc = 0; kc < nkc; ++kc) {
        bool sz = false;
        for (int k = 64 * kc; k <= 64 * kc + 64; ++k)
          sz |= e_special(k, e->nz, e->dn[2], e->dp[2], tf, e->tlk, e->trk) ||
                (k < 64 * kc + 64 && h_special(k, e->nz, e->dn[2], e->dp[2], tf, e->tlk, e->trk));
        const int kind = (sx || sy) ? SW_ANY : sz ? SW_ZONLY : SW_CORE;
        lists[kind].push_back(SweepTile{x.first, x.second, y.first, kc | rows << 16});
      }
    }
  std::vector<SweepTile> all;
  for (int q = 0; q < 3; ++q) {
    e->sp_off[q] = (int)all.size();
    e->sp_cnt[q] = (int)lists[q].size();
    all.insert(all.end(), lists[q].begin(), lists[q].end());
  }
  if (e->sp_tiles) {
    for (auto &p : e->owned)
      if (p == e->sp_tiles) p = nullptr;
    HIPCHK(hipFree(e->sp_tiles));
    e->sp_tiles = nullptr;
  }
  xf_status s = dev_alloc(e, &e->sp_tiles, all.size() * sizeof(SweepTile));
  if (s != XF_OK) return s;
  HIPCHK(hipMemcpy(e->sp_tiles, all.data(), all.size() * sizeof(SweepTile), hipMemcpyHostToDevice));
  e->sp_R = R;
  e->sp_tx = tx;
  return XF_OK;
}

// the second copies of the state the fused step double-buffers, zeroed
static xf_status alloc_alt(xf_engine *e) {
  if (e->alt_ready) return XF_OK;
  xf_status s;
  for (int q = 0; q < 6; ++q)
    if ((s = dev_alloc(e, &e->f_alt[q], (size_t)e->ncell * e->esz)) != XF_OK) return s;
  // both psi sets as one block each: per side the pairs y (AY, CY), z (AZ,
  // BZ) and x (BX, CX), each in 2 * slot elements interleaved as in their own
  // allocations (slot = the largest slab array), so the kernel finds pair r
  // of side h at (3 h + r) * 2 * slot; the current values move into the
  // first block
  static const int pfirst[3] = {P_AY, P_AZ, P_BX}, psecond[3] = {P_CY, P_BZ, P_CX}, paxis[3] = {1, 2, 0};
  long long slot = 1;
  for (int ax = 0; ax < 3; ++ax) slot = std::max(slot, (long long)(e->psi_bytes[ax] / e->esz));
  const long long n = 12 * slot;
  e->psi_n = slot;
  if ((s = dev_alloc(e, &e->psi_blk, (size_t)n * e->esz)) != XF_OK) return s;
  if ((s = dev_alloc(e, &e->psi_blk_alt, (size_t)n * e->esz)) != XF_OK) return s;
  for (int side = 0; side < 2; ++side)
    for (int r = 0; r < 3; ++r) {
      const size_t nb = 2 * e->psi_bytes[paxis[r]];
      void **cur = side ? e->psi_h : e->psi_e, **alt = side ? e->psi_h_alt : e->psi_e_alt;
      const size_t off = (size_t)(3 * side + r) * 2 * slot * e->esz;
      const int a = pfirst[r], b = psecond[r];
      // axes without CPML keep their placeholders (never addressed)
      if (nb) {
        char *dst = (char *)e->psi_blk + off;
        HIPCHK(hipMemcpyAsync(dst, cur[a], nb, hipMemcpyDeviceToDevice, e->stream));
        cur[a] = dst;
        cur[b] = dst + e->esz;
        alt[a] = (char *)e->psi_blk_alt + off;
        alt[b] = (char *)alt[a] + e->esz;
      } else {
        alt[a] = cur[a];
        alt[b] = cur[b];
      }
    }
  HIPCHK(hipStreamSynchronize(e->stream));
  if (e->has_tfsf) {
    if ((s = alloc_lines_alt(e)) != XF_OK) return s;
    // the 24 face-term segments (k_tfsf_terms): sizes over each face's two
    // in-face index ranges
    const long long ni = e->tri - e->tli, nj = e->trj - e->tlj, nk = e->trk - e->tlk;
    TfTerms &w = e->tf_terms;
    w.stride = 1;
    for (int q = 0; q < 24; ++q) w.stride = std::max(w.stride, tf_seg_size(q, ni, nj, nk));
    void *p = nullptr;
    if ((s = dev_alloc(e, &p, (size_t)(24 * w.stride) * sizeof(double))) != XF_OK) return s;
    w.buf = (double *)p;
  }
  e->alt_ready = true;
  return XF_OK;
}

static void swap_sets(xf_engine *e) {
  for (int q = 0; q < 6; ++q) {
    std::swap(e->f[q], e->f_alt[q]);
    std::swap(e->psi_e[q], e->psi_e_alt[q]);
    std::swap(e->psi_h[q], e->psi_h_alt[q]);
  }
  std::swap(e->psi_blk, e->psi_blk_alt);
  if (e->has_tfsf) {
    std::swap(e->e_inc, e->e_inc_alt);
    std::swap(e->h_inc, e->h_inc_alt);
  }
}

// SweepArgs of the two directions: block d reads the set that is current
// after an even (d = 0) or odd (d = 1) number of fused steps and writes the
// other; rebuilt at every stepping call (a few hundred bytes)
// arguments of the fused step from the current set to the other one
template <typename T>
static SweepArgs<T> sweep_args(xf_engine *e) {
  SweepArgs<T> h{};
  for (int q = 0; q < 6; ++q) {
    h.src[q] = (const T *)e->f[q];
    h.dst[q] = (T *)e->f_alt[q];
  }
  h.mat = e->mat;
  for (int side = 0; side < 2; ++side) {
    h.self[side] = (const T *)e->self[side];
    h.kt[side] = (const T *)e->kt[side];
    h.jt[side] = (const T *)e->jt[side];
    h.it[side] = (const T *)e->it[side];
    h.kp[side] = (const T *)e->kp[side];
    h.jp[side] = (const T *)e->jp[side];
    h.ip[side] = (const T *)e->ip[side];
  }
  for (int ax = 0; ax < 3; ++ax) {
    h.dn[ax] = e->dn[ax]; h.dp[ax] = e->dp[ax];
    h.be[ax] = (const T *)e->be[ax]; h.ae[ax] = (const T *)e->ae[ax];
    h.bm[ax] = (const T *)e->bm[ax]; h.am[ax] = (const T *)e->am[ax];
  }
  h.psiA = (const T *)e->psi_blk;
  h.psiB = (T *)e->psi_blk_alt;
  h.psi_slot = (unsigned)e->psi_n;
  h.terms = e->tf_terms.buf;
  h.term_stride = (unsigned)e->tf_terms.stride;
  h.nx = e->nx; h.ny = e->ny; h.nz = e->nz; h.pz = e->pz; h.nmat = e->nmat;
  h.sj = (unsigned)e->sj; h.si = (unsigned)e->si;
  h.tf = e->has_tfsf ? 1 : 0;
  if (e->has_tfsf) {
    h.tb[0] = e->tli; h.tb[1] = e->tri; h.tb[2] = e->tlj; h.tb[3] = e->trj; h.tb[4] = e->tlk; h.tb[5] = e->trk;
  }
  h.st.tiles = nullptr;
  h.st.nblocks = 0;
  h.st.waves = SW_WAVES;
  h.st.xcd_runs = (e->sweep_flags & 1) ? 1 : 0;
  h.div = (unsigned long long *)e->div;
  h.step = e->cur;
  return h;
}

template <typename T, int R, int KIND>
static void launch_sweep_kind(xf_engine *e, SweepArgs<T> &a) {
  if (e->sp_cnt[KIND] == 0) return;
  a.st.tiles = (const SweepTile *)e->sp_tiles + e->sp_off[KIND];
  a.st.nblocks = e->sp_cnt[KIND];
  const dim3 grd(a.st.nblocks), blk(64 * SW_WAVES);
  const size_t lds = sweep_lds_bytes(e);
  if (sweep_ms(e) == 4)
    hipLaunchKernelGGL((k_sweep<T, R, 4, KIND>), grd, blk, lds, e->stream, a);
  else
    hipLaunchKernelGGL((k_sweep<T, R, 8, KIND>), grd, blk, lds, e->stream, a);
}

// the fused step: one launch per tile kind (all read set A and write set B,
// so their order does not matter)
template <typename T, int R>
static void launch_sweep(xf_engine *e) {
  SweepArgs<T> a = sweep_args<T>(e);
  launch_sweep_kind<T, R, SW_ANY>(e, a);
  launch_sweep_kind<T, R, SW_ZONLY>(e, a);
  launch_sweep_kind<T, R, SW_CORE>(e, a);
}

template <typename T>
static xf_status run_steps(xf_engine *e, long long n) {
  const bool fused = fused_ok(e);
  if (fused) {
    xf_status s = alloc_alt(e);
    if (s == XF_OK) s = build_sweep_tiles(e);
    if (s != XF_OK) return s;
  }
  const int R = sweep_rows(e);
  if (fused && R != 1 && R != 2 && R != 4)
    FAILS(XF_EINVAL, "fused step with %d rows per thread (1, 2 or 4)", R);
  Tiling tl = make_tiling(e);
  if (e->replay && e->win_on && !fused) window_tiling(e, tl);
  dim3 blk(tiling_threads(tl));
  dim3 grd(tl.nblocks);
  // TFSF z faces in the curl kernels (split scheme)
  const bool zc = !fused && tfsf_zc(e);
  const int zper = (e->tri - e->tli + 1) * (e->trj - e->tlj + 1);
  if (zc && !e->zterms) {
    xf_status s = dev_alloc(e, &e->zterms, (size_t)8 * zper * sizeof(double));
    if (s != XF_OK) return s;
  }
  // coefficient tables of a tile go to LDS when they fit in 32 KB
  const size_t lds = lds_table_bytes<T>(e->nmat, tl);
  const bool use_lds = lds <= 32768 && !(e->tile_flags & XF_TILE_GLOBAL_TABLES);
  // design-box and Drude planes need the E kernel's general variant
  const bool special = e->has_design || e->has_drude;
  long long nfQ = 0;
  if (e->has_nf) {
    nfQ = nf_cells(e->nf);
    if (!e->nf_ring) {
      xf_status s = dev_alloc(e, &e->nf_ring, (size_t)NF_RING * 8 * nfQ * e->esz);
      if (s != XF_OK) return s;
    }
  }
  // The NF-FF z faces are sampled from records the H kernel leaves (split
  // scheme, 3-D surface, tiles of whole 64-lane k runs with each face plane
  // and the plane below it in one wave), unless the TFSF face kernel, which
  // runs after the H kernel, corrects H on one of the record planes (k in
  // [tlk - 1, trk]).
  Nffft nfs = e->nf;
  nfs.zrec = nullptr;
  NfRec nr{};
  if (e->has_nf && e->nf.zf && !fused && !e->replay && e->nf.rk - 1 > e->nf.lk && tl.ty > 0 && tl.tz % 64 == 0 &&
      e->nf.lk % 64 != 0 && e->nf.rk % 64 != 0) {
    auto clear = [&](int k) { return !e->has_tfsf || k < e->tlk - 1 || k > e->trk; };
    const Nffft &n = e->nf;
    if (clear(n.lk - 1) && clear(n.lk) && clear(n.rk - 1) && clear(n.rk)) {
      if (!e->nf_zrec) {
        xf_status s = dev_alloc(e, &e->nf_zrec, (size_t)2 * (n.nxb + 1) * (n.nyb + 1) * 8 * e->esz);
        if (s != XF_OK) return s;
      }
      nfs.zrec = e->nf_zrec;
      nr.rec = e->nf_zrec;
      nr.li = n.li; nr.lj = n.lj;
      nr.nxr = n.nxb + 1; nr.nyr = n.nyb + 1;
      nr.k0 = n.lk; nr.k1 = n.rk;
    }
  }
  auto nf_flush = [&]() {
    if (e->nf_pending == 0) return;
    hipLaunchKernelGGL(k_nffft_flush<T>, dim3((unsigned)((8 * nfQ + 255) / 256)), dim3(256), 0, e->stream, e->nf,
                       (const T *)e->nf_ring, e->nf_pending, e->nf_step0);
    e->nf_pending = 0;
  };
  for (long long q = 0; q < n; ++q) {
    const long long step = e->cur;
    hipEvent_t *ev = nullptr;
    if (e->timing) {
      if (e->ev_used + 4 > e->ev.size()) {
        for (int r = 0; r < 64; ++r) {
          hipEvent_t x;
          HIPCHK(hipEventCreate(&x));
          e->ev.push_back(x);
        }
      }
      ev = &e->ev[e->ev_used];
      e->ev_used += 4;
    }
    if (fused) {
      // IFA into the second copy of the line, then the whole step in one
      // kernel; the event pair brackets the fused kernel, the H pair is empty
      if (e->has_tfsf) {
        const Tfsf t0 = make_tfsf(e);
        hipLaunchKernelGGL(k_tfsf_ifa_pp, dim3(1), dim3(1024), 0, e->stream, t0, (double *)e->e_inc_alt,
                           (double *)e->h_inc_alt, step);
        // face terms: E from the H line before the IFA, H from the new E line
        Tfsf t1 = t0;
        t1.e_inc = (double *)e->e_inc_alt;
        hipLaunchKernelGGL(k_tfsf_terms, dim3((unsigned)((e->tf_terms.stride + 255) / 256), 24), dim3(256), 0,
                           e->stream, t1, e->tf_terms);
      }
      if (ev) hipEventRecord(ev[0], e->stream);
      if (R == 1)
        launch_sweep<T, 1>(e);
      else if (R == 2)
        launch_sweep<T, 2>(e);
      else
        launch_sweep<T, 4>(e);
      if (ev) {
        hipEventRecord(ev[1], e->stream);
        hipEventRecord(ev[2], e->stream);
        hipEventRecord(ev[3], e->stream);
      }
      swap_sets(e);
    } else {
      const Grid<T> g = make_grid<T>(e);
      Tfsf t{};
      TfsfESeg es{};
      TfsfHSeg hs{};
      ZFaces zf{};
      if (e->has_tfsf) {
        t = make_tfsf(e);
        tfsf_segs(e, zc, es, hs);
        if (zc) {
          zf.terms = (const double *)e->zterms;
          zf.li = e->tli; zf.ri = e->tri; zf.lj = e->tlj; zf.rj = e->trj; zf.lk = e->tlk; zf.rk = e->trk;
          // the E terms of the first step of this call (later steps' come
          // with the H terms of the step before)
          if (q == 0)
            hipLaunchKernelGGL(k_tfsf_zterms, dim3((4 * zper + 255) / 256), dim3(256), 0, e->stream, t,
                               (double *)e->zterms, 0, 4);
        }
      }
      if (ev) hipEventRecord(ev[0], e->stream);
      if (special) {
        if (use_lds)
          hipLaunchKernelGGL((k_update_e<T, true, true>), grd, blk, lds, e->stream, g, tl, zf);
        else
          hipLaunchKernelGGL((k_update_e<T, false, true>), grd, blk, 0, e->stream, g, tl, zf);
      } else {
        if (use_lds)
          hipLaunchKernelGGL((k_update_e<T, true, false>), grd, blk, lds, e->stream, g, tl, zf);
        else
          hipLaunchKernelGGL((k_update_e<T, false, false>), grd, blk, 0, e->stream, g, tl, zf);
      }
      if (ev) hipEventRecord(ev[1], e->stream);
      for (const SourceDev &sd : e->sources) launch_lumped<T>(e, sd, g.f[sd.comp], e->stream, step);
      if (e->has_tfsf) {
        if (es.off[6] > 0)
          hipLaunchKernelGGL(k_tfsf_e<T>, dim3((es.off[6] + 255) / 256), dim3(256), 0, e->stream, g, t, es);
        // the incident line steps before the H kernel, whose z-face terms
        // come from the new line (the E corrections above used the old one);
        // the same line gives the next step's E terms
        hipLaunchKernelGGL(k_tfsf_ifa, dim3(1), dim3(1024), 0, e->stream, t, step);
        if (zc)
          hipLaunchKernelGGL(k_tfsf_zterms, dim3((8 * zper + 255) / 256), dim3(256), 0, e->stream, t,
                             (double *)e->zterms, 0, 8);
      }
      if (ev) hipEventRecord(ev[2], e->stream);
      if (use_lds)
        hipLaunchKernelGGL((k_update_h<T, true>), grd, blk, lds, e->stream, g, tl, zf, nr, step);
      else
        hipLaunchKernelGGL((k_update_h<T, false>), grd, blk, 0, e->stream, g, tl, zf, nr, step);
      if (ev) hipEventRecord(ev[3], e->stream);
      if (e->has_tfsf && hs.off[12] > 0)
        hipLaunchKernelGGL(k_tfsf_h<T>, dim3((hs.off[12] + 255) / 256), dim3(256), 0, e->stream, g, t, hs);
    }
    // after the H half-step: NF-FF, DFT plane, broadband series, probes and
    // line monitors read the step's fields
    const Grid<T> g = make_grid<T>(e);
    if (e->replay) {
      e->cur++;
      continue;
    }
    if (e->has_nf) {
      if (e->nf_pending == 0) e->nf_step0 = step;
      hipLaunchKernelGGL(k_nffft_sample<T>, dim3((unsigned)((nfQ + 255) / 256)), dim3(256), 0, e->stream, g, nfs,
                         (T *)e->nf_ring, e->nf_pending);
      if (++e->nf_pending == NF_RING) nf_flush();
    }
    if (e->has_dft) {
      const int n = (e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
      hipLaunchKernelGGL(k_dft_plane<T>, dim3((n + 255) / 256), dim3(256), 0, e->stream,
                         (const T *)e->f[e->dft.comp], e->si, e->sj, e->dft, step);
    }
    if (e->has_bb && step < e->bb_T) {
      const long long Q = bb_cells(e->bb);
      hipLaunchKernelGGL(k_nffft_sample<T>, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0, e->stream, g, e->bb,
                         (T *)e->bb_hist, (int)step);
    }
    if (!e->probe_comp.empty()) {
      const int np = (int)e->probe_comp.size();
      hipLaunchKernelGGL(k_probe<T>, dim3((np + 63) / 64), dim3(64), 0, e->stream, g, e->d_probe_comp,
                         e->d_probe_cell, np, e->probe_buf, e->probe_T, step);
    }
    launch_line_monitors<T>(e, g, e->stream, step);
    HIPCHK(hipGetLastError());
    e->cur++;
    if (e->div_every > 0 && e->cur % e->div_every == 0) {
      nf_flush();
      xf_status s = check_divergence(e, e->stream);
      if (s != XF_OK) return s;
    }
  }
  nf_flush();
  HIPCHK(hipGetLastError());
  return XF_OK;
}

extern "C" xf_status xf_engine_step(xf_engine *e, int64_t n) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  if (!e->materials_set) FAILS(XF_ESTATE, "materials not set");
  if (n < 0) FAILS(XF_EINVAL, "negative step count");
  if (e->div_step >= 0)
    FAILS(XF_ESTATE, "the run diverged at step %lld; restore a checkpoint to go on", e->div_step);
  if (e->total >= 0 && e->cur + n > e->total)
    FAILS(XF_ESTATE, "step %lld + %lld beyond the %lld steps the source tables hold", e->cur, (long long)n,
          e->total);
  if (!e->probe_comp.empty()) {
    if (e->total < 0) FAILS(XF_ESTATE, "probes need a step horizon (a source or NF-FF table)");
    xf_status s = ensure_probes(e, e->total);
    if (s != XF_OK) return s;
  }
  if (e->batch_stream) HIPCHK(hipStreamSynchronize(e->batch_stream));
  return e->prec == 64 ? run_steps<double>(e, n) : run_steps<float>(e, n);
}

extern "C" xf_status xf_engine_sync(xf_engine *e) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  HIPCHK(engine_wait(e));
  return XF_OK;
}

extern "C" int64_t xf_engine_current_step(const xf_engine *e) { return e ? e->cur : -1; }

extern "C" xf_status xf_engine_get_field(xf_engine *e, int comp, double *out) {
  if (!e || !out || comp < 0 || comp > 5) FAILS(XF_EINVAL, "bad field request");
  int sh[3];
  comp_shape(e, comp, sh);
  HIPCHK(engine_wait(e));
  std::vector<unsigned char> buf((size_t)e->ncell * e->esz);
  HIPCHK(hipMemcpy(buf.data(), e->f[comp], buf.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < sh[0]; ++i)
    for (int j = 0; j < sh[1]; ++j) {
      const size_t base = (size_t)i * e->si + (size_t)j * e->sj;
      double *o = out + ((size_t)i * sh[1] + j) * sh[2];
      if (e->prec == 64) {
        const double *src = (const double *)buf.data() + base;
        for (int k = 0; k < sh[2]; ++k) o[k] = src[k];
      } else {
        const float *src = (const float *)buf.data() + base;
        for (int k = 0; k < sh[2]; ++k) o[k] = (double)src[k];
      }
    }
  return XF_OK;
}

template <typename T>
static xf_status get_box(xf_engine *e, int comp, const int64_t *box, double *out) {
  const long long n = box[1] * box[3] * box[5];
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, (size_t)n * sizeof(double)));
  hipLaunchKernelGGL(k_box_gather<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream,
                     (const T *)e->f[comp], e->si, e->sj, (int)box[0], (int)box[2], (int)box[4], (int)box[3],
                     (int)box[5], n, d);
  hipError_t r = hipGetLastError();
  if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
  if (r == hipSuccess) r = hipMemcpy(out, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(d);
  if (r != hipSuccess) FAILS(XF_EDEVICE, "box read: %s", hipGetErrorString(r));
  return XF_OK;
}

extern "C" xf_status xf_engine_get_box(xf_engine *e, int comp, const int64_t *box, double *out) {
  if (!e || !box || !out || comp < 0 || comp > 5) FAILS(XF_EINVAL, "bad box request");
  int sh[3];
  comp_shape(e, comp, sh);
  for (int a = 0; a < 3; ++a)
    if (box[2 * a] < 0 || box[2 * a + 1] < 1 || box[2 * a] + box[2 * a + 1] > sh[a])
      FAILS(XF_EINVAL, "box [%lld, +%lld) on axis %d outside component %d of extent %d", (long long)box[2 * a],
            (long long)box[2 * a + 1], a, comp, sh[a]);
  HIPCHK(engine_wait(e));
  return e->prec == 64 ? get_box<double>(e, comp, box, out) : get_box<float>(e, comp, box, out);
}

extern "C" xf_status xf_engine_set_field(xf_engine *e, int comp, const double *in) {
  if (!e || !in || comp < 0 || comp > 5) FAILS(XF_EINVAL, "bad field request");
  int sh[3];
  comp_shape(e, comp, sh);
  HIPCHK(engine_wait(e));
  std::vector<unsigned char> buf((size_t)e->ncell * e->esz, 0);
  for (int i = 0; i < sh[0]; ++i)
    for (int j = 0; j < sh[1]; ++j) {
      const size_t base = (size_t)i * e->si + (size_t)j * e->sj;
      const double *src = in + ((size_t)i * sh[1] + j) * sh[2];
      if (e->prec == 64) {
        double *o = (double *)buf.data() + base;
        for (int k = 0; k < sh[2]; ++k) o[k] = src[k];
      } else {
        float *o = (float *)buf.data() + base;
        for (int k = 0; k < sh[2]; ++k) o[k] = (float)src[k];
      }
    }
  HIPCHK(hipMemcpy(e->f[comp], buf.data(), buf.size(), hipMemcpyHostToDevice));
  return XF_OK;
}

extern "C" xf_status xf_engine_kernel_times(xf_engine *e, double *ms_e, double *ms_h, int64_t *samples) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  HIPCHK(engine_wait(e));
  double se = 0, sh = 0;
  const size_t n = e->ev_used / 4;
  for (size_t q = 0; q < n; ++q) {
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, e->ev[4 * q], e->ev[4 * q + 1]));
    HIPCHK(hipEventElapsedTime(&b, e->ev[4 * q + 2], e->ev[4 * q + 3]));
    se += a;
    sh += b;
  }
  const double d = n > 0 ? (double)n : 1.0;
  if (ms_e) *ms_e = se / d;
  if (ms_h) *ms_h = sh / d;
  if (samples) *samples = (int64_t)n;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_timing(xf_engine *e, int enabled) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  HIPCHK(engine_wait(e));
  e->timing = enabled != 0;
  e->ev_used = 0;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_tiling(xf_engine *e, int tz, int ty, int tx, int flags) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  if (tz < 64 || tz % 64 != 0 || ty < 0 || tx < 0 || (ty > 0 ? tz * ty : tz) > 256 || (flags & ~30) != 0)
    FAILS(XF_EINVAL, "tiling %d x %d x %d (flags %d): need tz a multiple of 64 and at most 256 threads", tz,
          ty, tx, flags);
  e->tz = tz; e->ty = ty; e->tx = tx;
  e->tile_flags = flags;
  return XF_OK;
}

extern "C" xf_status xf_engine_set_sweep(xf_engine *e, int mode, int rows, int planes) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  const int flags = mode >> 4;
  mode &= 15;
  if (mode < 0 || mode > 1 || !(rows == 0 || rows == 1 || rows == 2 || rows == 4) || planes < 0 || planes > 4096 ||
      (flags & ~1) != 0)
    FAILS(XF_EINVAL, "sweep mode %d, %d rows, %d planes: need mode 0 or 1, rows 0, 1, 2 or 4, planes >= 0", mode,
          rows, planes);
  HIPCHK(engine_wait(e));
  e->sweep_flags = flags;
  e->sweep = mode;
  e->sweep_rows = rows;
  e->sweep_tx = planes;
  return XF_OK;
}

extern "C" int xf_engine_sweep_active(const xf_engine *e) { return e && fused_ok(e) ? 1 : 0; }

extern "C" xf_status xf_engine_set_divergence_check(xf_engine *e, int every) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  if (every < 0) FAILS(XF_EINVAL, "divergence check every %d steps: need >= 0", every);
  e->div_every = every;
  return XF_OK;
}

extern "C" int64_t xf_engine_diverged_step(const xf_engine *e) { return e ? e->div_step : -1; }

extern "C" xf_status xf_engine_set_adjoint_overlap(xf_engine *e, int mode) {
  if (!e) FAILS(XF_EINVAL, "null engine");
  if (mode < -1 || mode > 1) FAILS(XF_EINVAL, "adjoint overlap %d: need -1, 0 or 1", mode);
  e->adj_overlap = mode;
  return XF_OK;
}

// ---------------------------------------------------------------------------
// Batches: many engines of one precision stepped together.  Each per-step
// kernel is one launch over all instances; a workgroup finds its instance and
// local block in a per-launch map.  Point probes and lumped sources, which
// are a handful of cells, keep per-instance launches on the batch stream.
struct xf_batch {
  std::vector<xf_engine *> e;
  int prec = 64;
  hipStream_t stream = nullptr;
  void *inst = nullptr;  // Inst<T>[n] on the device
  int2 *map_curl = nullptr, *map_te = nullptr, *map_th = nullptr, *map_ifa = nullptr, *map_nf = nullptr;
  int2 *map_dft = nullptr, *map_nff = nullptr;
  int n_curl = 0, n_te = 0, n_th = 0, n_ifa = 0, n_nf = 0, n_dft = 0, n_nff = 0;
  int threads = 256;
  size_t lds = 0;
  bool use_lds = true;
  bool special = false;  // some member has a design box or Drude cells
  std::vector<Tiling> tls;  // each member's curl tiling, as uploaded
  std::vector<void *> owned;
};

static xf_status batch_upload(xf_batch *b, void **dst, const void *src, size_t bytes) {
  if (bytes == 0) bytes = 16;
  HIPCHK(hipMalloc(dst, bytes));
  b->owned.push_back(*dst);
  if (src) HIPCHK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return XF_OK;
}

template <typename T>
static xf_status batch_build(xf_batch *b) {
  const int n = (int)b->e.size();
  std::vector<Inst<T>> in(n);
  std::vector<int2> mc, mte, mth, mifa, mnf, mnff, mdft;
  b->tls.clear();
  b->lds = 0;
  b->use_lds = true;
  b->special = false;
  for (int q = 0; q < n; ++q) {
    xf_engine *e = b->e[q];
    b->special = b->special || e->has_design || e->has_drude;
    Inst<T> &a = in[q];
    std::memset(&a, 0, sizeof a);
    a.g = make_grid<T>(e);
    a.tl = make_tiling(e);
    b->tls.push_back(a.tl);
    const size_t lds = lds_table_bytes<T>(e->nmat, a.tl);
    b->lds = std::max(b->lds, lds);
    if (lds > 32768 || (e->tile_flags & XF_TILE_GLOBAL_TABLES)) b->use_lds = false;
    for (int w = 0; w < a.tl.nblocks; ++w) mc.push_back(make_int2(q, w));
    if (e->has_tfsf) {
      a.t = make_tfsf(e);
      tfsf_segs(e, false, a.es, a.hs);
      const int te = (a.es.off[6] + 255) / 256, th = (a.hs.off[12] + 255) / 256;
      for (int w = 0; w < te; ++w) mte.push_back(make_int2(q, w));
      for (int w = 0; w < th; ++w) mth.push_back(make_int2(q, w));
      mifa.push_back(make_int2(q, 0));
    }
    if (e->has_nf) {
      a.nf = e->nf;
      // the member's own sample ring (its plain steps leave it flushed)
      const long long Q = nf_cells(e->nf);
      if (!e->nf_ring) {
        xf_status s = dev_alloc(e, &e->nf_ring, (size_t)NF_RING * 8 * Q * e->esz);
        if (s != XF_OK) return s;
      }
      a.ring = (T *)e->nf_ring;
      for (long long w = 0; w < (Q + 255) / 256; ++w) mnf.push_back(make_int2(q, (int)w));
      for (long long w = 0; w < (8 * Q + 255) / 256; ++w) mnff.push_back(make_int2(q, (int)w));
    }
    if (e->has_dft) {
      a.dft = e->dft;
      const int nn = (e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
      for (int w = 0; w < (nn + 255) / 256; ++w) mdft.push_back(make_int2(q, w));
    }
  }
  if (!b->use_lds) b->lds = 0;
  if (mc.size() > 0x7fffffffu) FAILS(XF_EINVAL, "batch too large");
  xf_status s;
  if ((s = batch_upload(b, &b->inst, in.data(), in.size() * sizeof(Inst<T>))) != XF_OK) return s;
  void *p;
  if ((s = batch_upload(b, &p, mc.data(), mc.size() * sizeof(int2))) != XF_OK) return s;
  b->map_curl = (int2 *)p; b->n_curl = (int)mc.size();
  if ((s = batch_upload(b, &p, mte.data(), mte.size() * sizeof(int2))) != XF_OK) return s;
  b->map_te = (int2 *)p; b->n_te = (int)mte.size();
  if ((s = batch_upload(b, &p, mth.data(), mth.size() * sizeof(int2))) != XF_OK) return s;
  b->map_th = (int2 *)p; b->n_th = (int)mth.size();
  if ((s = batch_upload(b, &p, mifa.data(), mifa.size() * sizeof(int2))) != XF_OK) return s;
  b->map_ifa = (int2 *)p; b->n_ifa = (int)mifa.size();
  if ((s = batch_upload(b, &p, mnf.data(), mnf.size() * sizeof(int2))) != XF_OK) return s;
  b->map_nf = (int2 *)p; b->n_nf = (int)mnf.size();
  if ((s = batch_upload(b, &p, mnff.data(), mnff.size() * sizeof(int2))) != XF_OK) return s;
  b->map_nff = (int2 *)p; b->n_nff = (int)mnff.size();
  if ((s = batch_upload(b, &p, mdft.data(), mdft.size() * sizeof(int2))) != XF_OK) return s;
  b->map_dft = (int2 *)p; b->n_dft = (int)mdft.size();
  return XF_OK;
}

extern "C" void xf_batch_destroy(xf_batch *b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  for (xf_engine *e : b->e)
    if (e->batch_stream == b->stream) e->batch_stream = nullptr;
  for (void *p : b->owned) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

extern "C" xf_status xf_batch_create(xf_engine *const *engines, int n, xf_batch **out) {
  if (!engines || !out || n < 1) FAILS(XF_EINVAL, "need at least one engine");
  *out = nullptr;
  const xf_engine *e0 = engines[0];
  if (!e0) FAILS(XF_EINVAL, "null engine");
  int threads = -1;
  for (int q = 0; q < n; ++q) {
    xf_engine *e = engines[q];
    if (!e) FAILS(XF_EINVAL, "null engine %d", q);
    if (!e->materials_set) FAILS(XF_ESTATE, "engine %d is not set up", q);
    if (e->batch_stream) FAILS(XF_ESTATE, "engine %d already belongs to a batch", q);
    if (e->has_bb) FAILS(XF_EINVAL, "engine %d has a broadband NF-FF surface, which batches do not step", q);
    if (e->nz == 1) FAILS(XF_EINVAL, "engine %d is a 2-D grid, which batches do not step", q);
    if (e->prec != e0->prec) FAILS(XF_EINVAL, "engine %d has precision %d, engine 0 has %d", q, e->prec, e0->prec);
    if (e->cur != e0->cur) FAILS(XF_ESTATE, "engine %d is at step %lld, engine 0 at %lld", q, e->cur, e0->cur);
    for (int r = 0; r < q; ++r)
      if (engines[r] == e) FAILS(XF_EINVAL, "engine %d appears twice", q);
    const int th = tiling_threads(make_tiling(e));
    if (threads >= 0 && th != threads)
      FAILS(XF_EINVAL, "engine %d tiles with %d threads per block, engine 0 with %d", q, th, threads);
    threads = th;
  }
  xf_batch *b = new xf_batch();
  b->e.assign(engines, engines + n);
  b->prec = e0->prec;
  b->threads = threads;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    FAILS(XF_EDEVICE, "hipStreamCreateWithFlags failed");
  }
  for (xf_engine *e : b->e) {
    if (!e->probe_comp.empty() && e->total >= 0) {
      xf_status s = ensure_probes(e, e->total);
      if (s != XF_OK) { xf_batch_destroy(b); return s; }
    }
    if (engine_wait(e) != hipSuccess) { xf_batch_destroy(b); FAILS(XF_EDEVICE, "engine stream failed"); }
  }
  xf_status s = b->prec == 64 ? batch_build<double>(b) : batch_build<float>(b);
  if (s != XF_OK) { xf_batch_destroy(b); return s; }
  for (xf_engine *e : b->e) e->batch_stream = b->stream;
  *out = b;
  return XF_OK;
}

// replay: re-running a segment for the adjoint, without the observers
template <typename T>
// map_curl / n_curl: the curl kernels' workgroups (default: every tile of
// every member; a replay passes the tiles of its window)
static xf_status batch_run(xf_batch *b, long long n, bool replay = false, const int2 *map_curl = nullptr,
                           int n_curl = 0) {
  const Inst<T> *in = (const Inst<T> *)b->inst;
  const dim3 blk(b->threads);
  const int2 *mc = map_curl ? map_curl : b->map_curl;
  const dim3 nc((unsigned)(map_curl ? n_curl : b->n_curl));
  // NF-FF samples go to each member's ring, flushed every NF_RING steps and
  // at the end of the call
  int nf_pending = 0;
  long long nf_step0 = 0;
  auto nf_flush = [&]() {
    if (nf_pending == 0) return;
    hipLaunchKernelGGL(kb_nffft_flush<T>, dim3(b->n_nff), dim3(256), 0, b->stream, in, b->map_nff, nf_pending,
                       nf_step0);
    nf_pending = 0;
  };
  for (long long q = 0; q < n; ++q) {
    const long long step = b->e[0]->cur;
    if (b->special) {
      if (b->use_lds)
        hipLaunchKernelGGL((kb_update_e<T, true, true>), nc, blk, b->lds, b->stream, in, mc);
      else
        hipLaunchKernelGGL((kb_update_e<T, false, true>), nc, blk, 0, b->stream, in, mc);
    } else {
      if (b->use_lds)
        hipLaunchKernelGGL((kb_update_e<T, true, false>), nc, blk, b->lds, b->stream, in, mc);
      else
        hipLaunchKernelGGL((kb_update_e<T, false, false>), nc, blk, 0, b->stream, in, mc);
    }
    for (xf_engine *e : b->e)
      for (const SourceDev &sd : e->sources) launch_lumped<T>(e, sd, (T *)e->f[sd.comp], b->stream, step);
    if (b->n_te > 0) hipLaunchKernelGGL(kb_tfsf_e<T>, dim3(b->n_te), dim3(256), 0, b->stream, in, b->map_te);
    // the incident lines step before the H kernel (its z-face corrections)
    if (b->n_ifa > 0)
      hipLaunchKernelGGL(kb_tfsf_ifa<T>, dim3(b->n_ifa), dim3(1024), 0, b->stream, in, b->map_ifa, step);
    if (b->use_lds)
      hipLaunchKernelGGL((kb_update_h<T, true>), nc, blk, b->lds, b->stream, in, mc, step);
    else
      hipLaunchKernelGGL((kb_update_h<T, false>), nc, blk, 0, b->stream, in, mc, step);
    if (b->n_th > 0) hipLaunchKernelGGL(kb_tfsf_h<T>, dim3(b->n_th), dim3(256), 0, b->stream, in, b->map_th);
    if (!replay) {
      if (b->n_nf > 0) {
        if (nf_pending == 0) nf_step0 = step;
        hipLaunchKernelGGL(kb_nffft_sample<T>, dim3(b->n_nf), dim3(256), 0, b->stream, in, b->map_nf, nf_pending);
        if (++nf_pending == NF_RING) nf_flush();
      }
      if (b->n_dft > 0) hipLaunchKernelGGL(kb_dft_plane<T>, dim3(b->n_dft), dim3(256), 0, b->stream, in, b->map_dft, step);
      for (xf_engine *e : b->e)
        if (!e->probe_comp.empty()) {
          const int np = (int)e->probe_comp.size();
          hipLaunchKernelGGL(k_probe<T>, dim3((np + 63) / 64), dim3(64), 0, b->stream, make_grid<T>(e),
                             e->d_probe_comp, e->d_probe_cell, np, e->probe_buf, e->probe_T, step);
        }
      for (xf_engine *e : b->e)
        if (!e->line_mons.empty()) launch_line_monitors<T>(e, make_grid<T>(e), b->stream, step);
    }
    HIPCHK(hipGetLastError());
    for (xf_engine *e : b->e) e->cur++;
    const int every = b->e[0]->div_every;
    if (every > 0 && b->e[0]->cur % every == 0) {
      nf_flush();
      for (size_t m = 0; m < b->e.size(); ++m) {
        xf_status s = check_divergence(b->e[m], b->stream, (int)m);
        if (s != XF_OK) return s;
      }
    }
  }
  nf_flush();
  HIPCHK(hipGetLastError());
  return XF_OK;
}

extern "C" xf_status xf_batch_step(xf_batch *b, int64_t n) {
  if (!b) FAILS(XF_EINVAL, "null batch");
  if (n < 0) FAILS(XF_EINVAL, "negative step count");
  for (size_t q = 0; q < b->e.size(); ++q) {
    xf_engine *e = b->e[q];
    if (e->cur != b->e[0]->cur) FAILS(XF_ESTATE, "engine %zu was stepped outside the batch", q);
    if (e->div_step >= 0)
      FAILS(XF_ESTATE, "engine %zu diverged at step %lld; restore a checkpoint to go on", q, e->div_step);
    if (e->total >= 0 && e->cur + n > e->total)
      FAILS(XF_ESTATE, "engine %zu: step %lld + %lld beyond the %lld steps its tables hold", q, e->cur,
            (long long)n, e->total);
    if (!e->probe_comp.empty() && (e->probe_buf == nullptr || e->probes_dirty))
      FAILS(XF_ESTATE, "engine %zu: probes added after the batch was made", q);
    // work enqueued on the engine's own stream goes first
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  return b->prec == 64 ? batch_run<double>(b, n) : batch_run<float>(b, n);
}

extern "C" xf_status xf_batch_sync(xf_batch *b) {
  if (!b) FAILS(XF_EINVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  return XF_OK;
}

extern "C" int xf_batch_size(const xf_batch *b) { return b ? (int)b->e.size() : 0; }

// ---------------------------------------------------------------------------
// Design box, DFT plane, checkpoints and the adjoint (no reference
// counterpart: the reference runs forward only).

extern "C" xf_status xf_engine_set_design(xf_engine *e, const int64_t *box, const double *coef,
                                          const double *eps_r) {
  if (!e || !box || !coef || !eps_r) FAILS(XF_EINVAL, "null argument");
  const int N[3] = {e->nx, e->ny, e->nz};
  for (int a = 0; a < 3; ++a) {
    const int64_t lo = box[2 * a], hi = box[2 * a + 1];
    // inside the grid and outside every CPML slab (E layers 1..dn, N-dp..N-1)
    if (lo < 1 || hi <= lo || hi > N[a] - 1 || lo < e->dn[a] + 1 || hi > N[a] - e->dp[a])
      FAILS(XF_EINVAL, "design box [%lld,%lld) on axis %d must lie inside the grid and outside the CPML",
            (long long)lo, (long long)hi, a);
  }
  if (e->has_design) FAILS(XF_ESTATE, "design box already set");
  const size_t nd = (size_t)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
  xf_status s;
  if ((s = upload_real(e, &e->dcoef, coef, 9 * nd)) != XF_OK) return s;
  void *p;
  if ((s = dev_alloc(e, &p, nd * sizeof(double))) != XF_OK) return s;
  e->d_eps = (double *)p;
  HIPCHK(hipMemcpy(e->d_eps, eps_r, nd * sizeof(double), hipMemcpyHostToDevice));
  for (int a = 0; a < 6; ++a) e->db[a] = (int)box[a];
  e->nd = (int)nd;
  e->has_design = true;
  return XF_OK;
}

extern "C" xf_status xf_engine_update_design(xf_engine *e, const double *coef, const double *eps_r) {
  if (!e || !coef || !eps_r) FAILS(XF_EINVAL, "null argument");
  if (!e->has_design) FAILS(XF_ESTATE, "no design box");
  const size_t nd = (size_t)e->nd;
  HIPCHK(engine_wait(e));
  if (e->prec == 64) {
    HIPCHK(hipMemcpy(e->dcoef, coef, 9 * nd * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(9 * nd);
    for (size_t q = 0; q < 9 * nd; ++q) tmp[q] = (float)coef[q];
    HIPCHK(hipMemcpy(e->dcoef, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(e->d_eps, eps_r, nd * sizeof(double), hipMemcpyHostToDevice));
  return XF_OK;
}

extern "C" xf_status xf_engine_set_drude(xf_engine *e, const int64_t *box, const double *coef) {
  if (!e || !box || !coef) FAILS(XF_EINVAL, "null argument");
  // the Drude records of the dispersive layout
  std::vector<double> rec(XF_ADE_STRIDE * (size_t)std::max(e->nmat, 0), 0.0);
  for (int m = 0; m < e->nmat; ++m) {
    for (int c = 0; c < 6; ++c) rec[XF_ADE_STRIDE * m + c] = coef[6 * m + c];
    rec[XF_ADE_STRIDE * m + XF_ADE_KIND] = coef[6 * m + 1] != 0 ? 1 : 0;
  }
  return xf_engine_set_dispersive(e, box, rec.data());
}

extern "C" xf_status xf_engine_set_dispersive(xf_engine *e, const int64_t *box, const double *coef) {
  if (!e || !box || !coef) FAILS(XF_EINVAL, "null argument");
  if (!e->materials_set) FAILS(XF_ESTATE, "set the materials before the Drude box");
  if (e->has_drude) FAILS(XF_ESTATE, "Drude box already set");
  const int N[3] = {e->nx, e->ny, e->nz};
  for (int a = 0; a < 3; ++a) {
    const int64_t lo = box[2 * a], hi = box[2 * a + 1];
    // inside the grid and outside every CPML slab, like the design box
    if (lo < (e->dn[a] > 0 ? e->dn[a] + 1 : 0) || hi <= lo || hi > N[a] - e->dp[a])
      FAILS(XF_EINVAL, "Drude box [%lld,%lld) on axis %d must lie inside the grid and outside the CPML",
            (long long)lo, (long long)hi, a);
  }
  bool lorentz = false;
  for (int m = 0; m < e->nmat; ++m) {
    const double *c = coef + XF_ADE_STRIDE * m;
    const double kind = c[XF_ADE_KIND];
    if (kind != 0 && kind != 1 && kind != 2 && kind != 3)
      FAILS(XF_EINVAL, "bad dispersive kind %g of material %d", kind, m);
    // Drude and Debye: a finite ca and a positive cb; Lorentz: a positive
    // curl weight and a positive 2 dt; Debye: a positive dt
    if ((kind == 1 || kind == 2) && !(c[0] == c[0] && c[1] > 0))
      FAILS(XF_EINVAL, "bad dispersive coefficients of material %d", m);
    if ((kind == 2 && !(c[9] > 0)) || (kind == 3 && !(c[0] == c[0] && c[8] > 0 && c[10] > 0)))
      FAILS(XF_EINVAL, "bad dispersive coefficients of material %d", m);
    lorentz = lorentz || kind == 3;
    e->ade_non_drude = e->ade_non_drude || kind == 2 || kind == 3;
  }
  const size_t n = (size_t)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
  xf_status s;
  if ((s = upload_real(e, &e->drc, coef, XF_ADE_STRIDE * (size_t)e->nmat)) != XF_OK) return s;
  for (int q = 0; q < 3; ++q) {
    if ((s = dev_alloc(e, &e->dj[q], n * e->esz)) != XF_OK) return s;
    e->state.push_back({e->dj[q], nullptr, n * e->esz});
    if (lorentz) {
      if ((s = dev_alloc(e, &e->djp[q], n * e->esz)) != XF_OK) return s;
      if ((s = dev_alloc(e, &e->dep[q], n * e->esz)) != XF_OK) return s;
      e->state.push_back({e->djp[q], nullptr, n * e->esz});
      e->state.push_back({e->dep[q], nullptr, n * e->esz});
    }
  }
  for (int a = 0; a < 6; ++a) e->xb[a] = (int)box[a];
  e->has_drude = true;
  return XF_OK;
}

// Drude J (box layout) <-> E-shaped host arrays: `to_dev` copies the box part
// of `full` into dev, else the box part of dev into `full` (other cells of
// `full` are left alone)
template <typename T>
static xf_status drude_box_copy(xf_engine *e, int comp, void *dev, double *full, bool to_dev) {
  int sh[3];
  comp_shape(e, comp, sh);
  const int *b = e->xb;
  const size_t n = drude_cells(e);
  std::vector<T> box(n);
  if (!to_dev) HIPCHK(hipMemcpy(box.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
  size_t q = 0;
  for (int i = b[0]; i < b[1]; ++i)
    for (int j = b[2]; j < b[3]; ++j)
      for (int k = b[4]; k < b[5]; ++k, ++q) {
        double &f = full[((size_t)i * sh[1] + j) * sh[2] + k];
        if (to_dev)
          box[q] = (T)f;
        else
          f = (double)box[q];
      }
  if (to_dev) HIPCHK(hipMemcpy(dev, box.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return XF_OK;
}

extern "C" xf_status xf_engine_set_drude_j(xf_engine *e, int comp, const double *in) {
  if (!e || !in || comp < 0 || comp > 2) FAILS(XF_EINVAL, "bad argument");
  if (!e->has_drude) FAILS(XF_ESTATE, "no Drude box");
  HIPCHK(engine_wait(e));
  const size_t n = drude_cells(e);
  if (e->prec == 64) {
    HIPCHK(hipMemcpy(e->dj[comp], in, n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(in, in + n);
    HIPCHK(hipMemcpy(e->dj[comp], tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return XF_OK;
}

extern "C" xf_status xf_engine_drude_box(xf_engine *e, int64_t *box) {
  if (!e || !box) FAILS(XF_EINVAL, "null argument");
  if (!e->has_drude) FAILS(XF_ESTATE, "no Drude box");
  for (int a = 0; a < 6; ++a) box[a] = e->xb[a];
  return XF_OK;
}

extern "C" xf_status xf_engine_get_drude_j(xf_engine *e, int comp, double *out) {
  if (!e || !out || comp < 0 || comp > 2) FAILS(XF_EINVAL, "bad argument");
  if (!e->has_drude) FAILS(XF_ESTATE, "no Drude box");
  HIPCHK(engine_wait(e));
  const size_t n = (size_t)(e->xb[1] - e->xb[0]) * (e->xb[3] - e->xb[2]) * (e->xb[5] - e->xb[4]);
  if (e->prec == 64) {
    HIPCHK(hipMemcpy(out, e->dj[comp], n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), e->dj[comp], n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) out[q] = tmp[q];
  }
  return XF_OK;
}

extern "C" xf_status xf_engine_set_dft_plane(xf_engine *e, int comp, int axis, int64_t index, int64_t lo1,
                                             int64_t hi1, int64_t lo2, int64_t hi2, const double *w,
                                             int64_t total_steps) {
  if (!e || !w || comp < 0 || comp > 5 || axis < 0 || axis > 2) FAILS(XF_EINVAL, "bad DFT plane");
  if (e->has_dft) FAILS(XF_ESTATE, "DFT plane already set");
  int sh[3];
  comp_shape(e, comp, sh);
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  if (index < 0 || index >= sh[axis] || lo1 < 0 || hi1 > sh[a1] || lo1 >= hi1 || lo2 < 0 || hi2 > sh[a2] ||
      lo2 >= hi2)
    FAILS(XF_EINVAL, "DFT plane outside the component array");
  DftPlane &d = e->dft;
  d.comp = comp; d.axis = axis; d.index = (int)index;
  d.lo1 = (int)lo1; d.hi1 = (int)hi1; d.lo2 = (int)lo2; d.hi2 = (int)hi2;
  const size_t n = (size_t)(hi1 - lo1) * (hi2 - lo2);
  xf_status s;
  void *p;
  if ((s = dev_alloc(e, &p, n * sizeof(double2))) != XF_OK) return s;
  d.acc = (double2 *)p;
  e->state.push_back({p, nullptr, n * sizeof(double2)});
  if ((s = dev_alloc(e, &p, (size_t)total_steps * sizeof(double2))) != XF_OK) return s;
  HIPCHK(hipMemcpy(p, w, (size_t)total_steps * sizeof(double2), hipMemcpyHostToDevice));
  d.w = (const double2 *)p;
  e->dft_T = total_steps;
  set_total(e, total_steps);
  e->has_dft = true;
  return XF_OK;
}

extern "C" xf_status xf_engine_get_dft(xf_engine *e, double *out) {
  if (!e || !out || !e->has_dft) FAILS(XF_ESTATE, "no DFT plane");
  HIPCHK(engine_wait(e));
  const size_t n = (size_t)(e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
  HIPCHK(hipMemcpy(out, e->dft.acc, n * sizeof(double2), hipMemcpyDeviceToHost));
  return XF_OK;
}

// copies of every state buffer, owned by the caller of these helpers
// copy the engine's state into `dst` (allocated at the first save: from
// `arena` when one is given, else one hipMalloc per array)
static xf_status state_save(xf_engine *e, std::vector<void *> &dst, hipStream_t st = nullptr,
                            AdjBufs *arena = nullptr) {
  if (!st) st = e->stream;
  if (dst.empty()) {
    for (auto &b : e->state) {
      void *p = nullptr;
      if (arena) {
        if (!(p = arena->alloc(b.n, false))) FAILS(XF_ENOMEM, "checkpoint buffers");
      } else {
        HIPCHK(hipMalloc(&p, b.n));
      }
      dst.push_back(p);
    }
  }
  for (size_t q = 0; q < e->state.size(); ++q)
    HIPCHK(hipMemcpyAsync(dst[q], e->state[q].ptr(), e->state[q].n, hipMemcpyDeviceToDevice, st));
  return XF_OK;
}
static xf_status state_load(xf_engine *e, const std::vector<void *> &src, hipStream_t st = nullptr) {
  if (!st) st = e->stream;
  for (size_t q = 0; q < e->state.size(); ++q)
    HIPCHK(hipMemcpyAsync(e->state[q].ptr(), src[q], e->state[q].n, hipMemcpyDeviceToDevice, st));
  return XF_OK;
}
static void state_free(std::vector<void *> &v) {
  for (void *p : v) hipFree(p);
  v.clear();
}

extern "C" xf_status xf_engine_checkpoint(xf_engine *e, int slot) {
  if (!e || slot < 0 || slot > 4095) FAILS(XF_EINVAL, "bad checkpoint slot");
  HIPCHK(engine_wait(e));
  if ((int)e->ckpt.size() <= slot) {
    e->ckpt.resize(slot + 1);
    e->ckpt_cur.resize(slot + 1, -1);
  }
  if (!e->ckpt[slot].empty() && e->ckpt[slot].size() != e->state.size()) state_free(e->ckpt[slot]);
  xf_status s = state_save(e, e->ckpt[slot]);
  if (s != XF_OK) return s;
  e->ckpt_cur[slot] = e->cur;
  return XF_OK;
}

extern "C" xf_status xf_engine_restore(xf_engine *e, int slot) {
  if (!e || slot < 0 || slot >= (int)e->ckpt.size() || e->ckpt[slot].size() != e->state.size())
    FAILS(XF_EINVAL, "no checkpoint in slot %d", slot);
  HIPCHK(engine_wait(e));
  xf_status s = state_load(e, e->ckpt[slot]);
  if (s != XF_OK) return s;
  e->cur = e->ckpt_cur[slot];
  return reset_divergence(e);
}

// device buffers of one adjoint run

template <typename T>
static xf_status adj_setup(xf_engine *e, AdjBufs &B, Adj<T> &a) {
  std::memset(&a, 0, sizeof a);
  const size_t fb = (size_t)e->ncell * sizeof(T);
  for (int q = 0; q < 3; ++q) {
    if (!(a.Ea[q] = (T *)B.alloc(fb)) || !(a.Ha[q] = (T *)B.alloc(fb))) FAILS(XF_ENOMEM, "adjoint fields");
  }
  // psi adjoints, two copies each: the same layouts as psi (slots of the y,
  // z, x axes)
  static const int slot_axis[6] = {1, 2, 0, 2, 0, 1};  // P_AY, P_AZ, P_BX, P_BZ, P_CX, P_CY
  for (int c = 0; c < 2; ++c)
    for (int q = 0; q < 6; ++q) {
      const size_t nb = e->psi_bytes[slot_axis[q]];
      if (!(a.Pe[c][q] = (T *)B.alloc(nb)) || !(a.Ph[c][q] = (T *)B.alloc(nb))) FAILS(XF_ENOMEM, "adjoint psi");
    }
  if (e->has_drude)
    for (int q = 0; q < 3; ++q)
      if (!(a.Ja[q] = (T *)B.alloc(drude_cells(e) * sizeof(T)))) FAILS(XF_ENOMEM, "adjoint Drude currents");
  return XF_OK;
}

template <typename T>
static xf_status grad_run(xf_engine *e, int K, double *grad, double *J) {
  const long long T_ = e->total;
  const int nseg = (int)((T_ + K - 1) / K);
  std::vector<std::vector<void *>> ck(nseg + 1);
  AdjBufs CK(e->stream);  // checkpoint storage
  AdjBufs B(e->stream);
  // two design-history buffers: segment s uses hist[s & 1], so the replay of
  // segment s - 1 can run while the adjoint sweeps segment s
  const size_t hb = (size_t)(K + 1) * e->nd * 3 * sizeof(T);
  T *hist[2] = {(T *)B.alloc(hb), nseg > 1 ? (T *)B.alloc(hb) : nullptr};
  double *g = (double *)B.alloc((size_t)e->nd * sizeof(double));
  if (!hist[0] || (nseg > 1 && !hist[1]) || !g) FAILS(XF_ENOMEM, "adjoint buffers");
  const Grid<T> gr = make_grid<T>(e);
  const unsigned nsnap = (unsigned)((e->nd + 255) / 256);
  // forward run with a checkpoint at the start of every segment but the
  // last, whose design-box E fields are kept as it runs (it is never
  // recomputed), and a checkpoint at the end.  The E kernel writes each
  // step's design fields itself when nothing after it changes E in the box.
  const bool esnap = e_kernel_snaps(e);
  T *hl = hist[(nseg - 1) & 1];
  for (int sgi = 0; sgi < nseg; ++sgi) {
    const long long n = std::min<long long>(K, T_ - (long long)sgi * K);
    xf_status s;
    if (sgi + 1 < nseg) {
      if ((s = state_save(e, ck[sgi], nullptr, &CK)) != XF_OK) return s;
      if ((s = run_steps<T>(e, n)) != XF_OK) return s;
      continue;
    }
    hipLaunchKernelGGL(k_design_snap<T>, dim3(nsnap), dim3(256), 0, e->stream, gr, hl, 0);
    for (long long l = 0; l < n; ++l) {
      if (esnap) e->snap = hl + (size_t)(l + 1) * e->nd * 3;
      s = run_steps<T>(e, 1);
      e->snap = nullptr;
      if (s != XF_OK) return s;
      if (!esnap) hipLaunchKernelGGL(k_design_snap<T>, dim3(nsnap), dim3(256), 0, e->stream, gr, hl, (int)(l + 1));
    }
  }
  xf_status s = state_save(e, ck[nseg], nullptr, &CK);
  if (s != XF_OK) return s;
  // objective from the final DFT accumulators
  const DftPlane dpl = e->dft;
  const size_t np = (size_t)(dpl.hi1 - dpl.lo1) * (dpl.hi2 - dpl.lo2);
  double2 *Ffin = (double2 *)B.alloc(np * sizeof(double2));
  if (!Ffin) FAILS(XF_ENOMEM, "adjoint buffers");
  HIPCHK(hipMemcpyAsync(Ffin, dpl.acc, np * sizeof(double2), hipMemcpyDeviceToDevice, e->stream));
  std::vector<double2> F(np);
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(F.data(), Ffin, np * sizeof(double2), hipMemcpyDeviceToHost));
  double Jv = 0;
  for (const double2 &f : F) Jv += f.x * f.x + f.y * f.y;
  *J = Jv;
  Adj<T> a;
  if ((s = adj_setup<T>(e, B, a)) != XF_OK) return s;
  a.dft = dpl;
  a.dft.acc = Ffin;
  a.inject = dpl.comp < 3;
  a.eps_r = e->d_eps;
  a.grad = g;
  a.grad_on = 1;
  // the adjoint sweeps run on a stream of their own (default priority) when
  // they overlap the replays, which stay on the engine's stream.  Events
  // order each sweep after its segment's replay and each replay after the
  // sweep that last used its history buffer.
  struct Side {
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev;
    ~Side() {
      if (st) hipStreamSynchronize(st);
      for (hipEvent_t x : ev) hipEventDestroy(x);
      if (st && own) hipStreamDestroy(st);
    }
    bool own = false;
  } side;
  const bool overlap = e->adj_overlap < 0 ? e->prec == 32 : e->adj_overlap != 0;
  if (overlap) {
    HIPCHK(hipStreamCreateWithFlags(&side.st, hipStreamNonBlocking));
    side.own = true;
  }
  hipStream_t as = overlap ? side.st : e->stream;
  side.ev.resize(2 * (size_t)nseg, nullptr);
  for (hipEvent_t &x : side.ev) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  hipEvent_t *ready = side.ev.data(), *done = side.ev.data() + nseg;  // replay of s ready, sweep of s done
  HIPCHK(hipEventRecord(ready[nseg - 1], e->stream));
  int par = 0;  // which psi-adjoint copy holds the current values
  // The adjoints start at zero and are driven only on the objective plane.
  // Along i, k_adj_e moves them up one plane per reverse step (it reads H
  // adjoints at i and i - 1) and k_adj_h down one (E adjoints at i and
  // i + 1); psi and Drude adjoints are per cell.  Before reverse step q
  // (counted from the run's end) they are therefore zero beyond q planes of
  // the plane's i extent [pi0, pi1): step q launches only the workgroups
  // over the planes within q + 2 of it, and the cells beyond, zero in every
  // array, are left as they are (a full launch would compute zeros there).
  const long long pi0 = dpl.axis == 0 ? dpl.index : dpl.lo1, pi1 = dpl.axis == 0 ? dpl.index + 1 : dpl.hi1;
  long long q = 0;
  auto rev_window = [&](long long &b0) {
    const long long lo = std::max(0LL, pi0 - (q + 2)), hi = std::min((long long)e->nx + 1, pi1 + q + 2);
    b0 = lo * e->si / 256;
    return (unsigned)(adj_blocks(hi * e->si) - b0);
  };
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    // replay segment sgi - 1 into the other history buffer, once the sweep
    // of segment sgi + 1 (its last user) is done
    if (sgi >= 1) {
      const int r = sgi - 1;
      const long long r0 = (long long)r * K, rn = std::min<long long>(K, T_ - r0);
      T *hr = hist[r & 1];
      if (sgi + 1 < nseg) HIPCHK(hipStreamWaitEvent(e->stream, done[sgi + 1], 0));
      if ((s = state_load(e, ck[r])) != XF_OK) return s;
      e->cur = r0;
      hipLaunchKernelGGL(k_design_snap<T>, dim3(nsnap), dim3(256), 0, e->stream, gr, hr, 0);
      e->replay = true;
      e->win_on = true;
      for (long long l = 0; l < rn && s == XF_OK; ++l) {
        // Only the design box's E after each step is kept, and a step moves
        // information by at most one cell along each axis (E reads H at c
        // and c - 1, H reads E at c and c + 1), so step l only has to be
        // right within rn - 1 - l cells of the box; the window keeps one
        // cell more.  The cells left outside go stale and are never read
        // by a cell inside a later window; the next replay starts from a
        // checkpoint of the whole state.
        replay_window(e, (int)(rn - l), e->win);  // rn <= K, an int
        if (esnap) e->snap = hr + (size_t)(l + 1) * e->nd * 3;
        s = run_steps<T>(e, 1);
        e->snap = nullptr;
        if (s == XF_OK && !esnap)
          hipLaunchKernelGGL(k_design_snap<T>, dim3(nsnap), dim3(256), 0, e->stream, gr, hr, (int)(l + 1));
      }
      e->win_on = false;
      e->replay = false;
      if (s != XF_OK) return s;
      HIPCHK(hipEventRecord(ready[r], e->stream));
    }
    // reverse steps of segment sgi
    const long long t0 = (long long)sgi * K, n = std::min<long long>(K, T_ - t0);
    HIPCHK(hipStreamWaitEvent(as, ready[sgi], 0));
    a.hist = hist[sgi & 1];
    for (long long l = n - 1; l >= 0; --l, ++q) {
      long long b0 = 0;
      const unsigned nw = rev_window(b0);
      hipLaunchKernelGGL(k_adj_e<T>, dim3(nw), dim3(256), 0, as, gr, a, t0 + l, (int)l, par, (int)ADJ_STEP, b0);
      hipLaunchKernelGGL(k_adj_h<T>, dim3(nw), dim3(256), 0, as, gr, a, par, b0);
      par ^= 1;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(done[sgi], as));
  }
  HIPCHK(hipStreamSynchronize(as));
  // leave the engine at the end of the forward run
  if ((s = state_load(e, ck[nseg])) != XF_OK) return s;
  e->cur = T_;
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(hipMemcpy(grad, g, (size_t)e->nd * sizeof(double), hipMemcpyDeviceToHost));
  return XF_OK;
}

extern "C" xf_status xf_engine_grad_eps(xf_engine *e, int checkpoint_every, double *grad, double *J) {
  if (!e || !grad || !J || checkpoint_every < 0) FAILS(XF_EINVAL, "bad argument");
  if (!e->materials_set || !e->has_design || !e->has_dft)
    FAILS(XF_ESTATE, "the gradient needs materials, a design box and a DFT plane");
  if (has_inductor(e)) FAILS(XF_EINVAL, "the gradient does not transpose inductors");
  if (has_non_drude(e)) FAILS(XF_EINVAL, "the gradient covers Drude media only, not Debye or Lorentz");
  if (e->cur != 0) FAILS(XF_ESTATE, "the gradient runs from step 0 (engine is at step %lld)", e->cur);
  if (e->total < 1) FAILS(XF_ESTATE, "no step horizon");
  if (!e->probe_comp.empty()) {
    xf_status s = ensure_probes(e, e->total);
    if (s != XF_OK) return s;
  }
  HIPCHK(engine_wait(e));
  int K = checkpoint_every;
  if (K == 0) {
    // automatic: keep the design fields of the whole run when they fit in
    // the history budget (no recompute), else a checkpoint every 100 steps
    const double bytes = (double)(e->total + 1) * e->nd * 3 * e->esz;
    K = bytes <= history_budget() ? (int)std::min<long long>(e->total, 1 << 30) : 100;
  }
  return e->prec == 64 ? grad_run<double>(e, K, grad, J) : grad_run<float>(e, K, grad, J);
}

template <typename T>
static xf_status adjoint_run(xf_engine *e, long long nsteps, double *fields) {
  AdjBufs B(e->stream);
  Adj<T> a;
  xf_status s = adj_setup<T>(e, B, a);
  if (s != XF_OK) return s;
  // upload the adjoint fields into the padded layout
  std::vector<T> buf((size_t)e->ncell);
  size_t off = 0;
  for (int q = 0; q < 6; ++q) {
    int sh[3];
    comp_shape(e, q, sh);
    std::fill(buf.begin(), buf.end(), T(0));
    for (int i = 0; i < sh[0]; ++i)
      for (int j = 0; j < sh[1]; ++j)
        for (int k = 0; k < sh[2]; ++k) buf[(size_t)i * e->si + (size_t)j * e->sj + k] = (T)fields[off++];
    HIPCHK(hipMemcpy(q < 3 ? (void *)a.Ea[q] : (void *)a.Ha[q - 3], buf.data(), buf.size() * sizeof(T),
                     hipMemcpyHostToDevice));
  }
  // the Drude J adjoints follow in E shapes; only their box part is state
  size_t joff[3] = {0, 0, 0};
  if (e->has_drude)
    for (int q = 0; q < 3; ++q) {
      int sh[3];
      comp_shape(e, q, sh);
      joff[q] = off;
      off += (size_t)sh[0] * sh[1] * sh[2];
      if ((s = drude_box_copy<T>(e, q, a.Ja[q], fields + joff[q], true)) != XF_OK) return s;
    }
  const Grid<T> gr = make_grid<T>(e);
  const unsigned nb = (unsigned)adj_blocks(e->ncell);
  // the given adjoints are the gathered, unscaled values the kernels pass
  // between them (the first E kernel takes its E adjoint as it is); a last
  // E kernel applies the pending E scaling
  int par = 0;
  for (long long t = 0; t < nsteps; ++t) {
    hipLaunchKernelGGL(k_adj_e<T>, dim3(nb), dim3(256), 0, e->stream, gr, a, 0LL, 0, par,
                       (int)(t == 0 ? ADJ_FIRST : ADJ_STEP), 0LL);
    hipLaunchKernelGGL(k_adj_h<T>, dim3(nb), dim3(256), 0, e->stream, gr, a, par, 0LL);
    par ^= 1;
  }
  if (nsteps > 0)
    hipLaunchKernelGGL(k_adj_e<T>, dim3(nb), dim3(256), 0, e->stream, gr, a, 0LL, 0, par, (int)ADJ_FINAL, 0LL);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  off = 0;
  for (int q = 0; q < 6; ++q) {
    int sh[3];
    comp_shape(e, q, sh);
    HIPCHK(hipMemcpy(buf.data(), q < 3 ? (void *)a.Ea[q] : (void *)a.Ha[q - 3], buf.size() * sizeof(T),
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < sh[0]; ++i)
      for (int j = 0; j < sh[1]; ++j)
        for (int k = 0; k < sh[2]; ++k) fields[off++] = (double)buf[(size_t)i * e->si + (size_t)j * e->sj + k];
  }
  if (e->has_drude)
    for (int q = 0; q < 3; ++q)
      if ((s = drude_box_copy<T>(e, q, a.Ja[q], fields + joff[q], false)) != XF_OK) return s;
  return XF_OK;
}

extern "C" xf_status xf_engine_adjoint_run(xf_engine *e, int64_t nsteps, double *fields) {
  if (!e || !fields || nsteps < 1) FAILS(XF_EINVAL, "bad argument");
  if (!e->materials_set) FAILS(XF_ESTATE, "materials not set");
  if (has_inductor(e)) FAILS(XF_EINVAL, "the adjoint does not transpose inductors");
  if (has_non_drude(e)) FAILS(XF_EINVAL, "the adjoint covers Drude media only, not Debye or Lorentz");
  HIPCHK(engine_wait(e));
  return e->prec == 64 ? adjoint_run<double>(e, nsteps, fields) : adjoint_run<float>(e, nsteps, fields);
}

// ---------------------------------------------------------------------------
// Batched gradient: xf_engine_grad_eps for every member of a batch at once.
// The forward and recomputed segments step through batch_run; the design
// snapshots and the reverse steps are one launch per kernel over all members.
template <typename T>
static xf_status batch_grad_run(xf_batch *b, int K, double *const *grads, double *Js) {
  const int n = (int)b->e.size();
  const long long T_ = b->e[0]->total;
  const int nseg = (int)((T_ + K - 1) / K);
  hipStream_t st = b->stream;
  std::vector<std::vector<std::vector<void *>>> ck(n, std::vector<std::vector<void *>>(nseg + 1));
  AdjBufs CK(st);  // checkpoint storage of every member
  AdjBufs B(st);  // every member's adjoint buffers
  std::vector<InstAdj<T>> ia(n);
  std::vector<double *> gdev(n);
  std::vector<double2 *> Ffin(n);
  std::vector<int2> msnap, mcell;
  for (int q = 0; q < n; ++q) {
    xf_engine *e = b->e[q];
    InstAdj<T> &x = ia[q];
    std::memset(&x, 0, sizeof x);
    x.g = make_grid<T>(e);
    x.hist = (T *)B.alloc((size_t)(K + 1) * e->nd * 3 * sizeof(T));
    gdev[q] = (double *)B.alloc((size_t)e->nd * sizeof(double));
    const size_t np = (size_t)(e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
    Ffin[q] = (double2 *)B.alloc(np * sizeof(double2));
    if (!x.hist || !gdev[q] || !Ffin[q]) FAILS(XF_ENOMEM, "adjoint buffers of member %d", q);
    xf_status s = adj_setup<T>(e, B, x.a);
    if (s != XF_OK) return s;
    x.a.dft = e->dft;
    x.a.dft.acc = Ffin[q];
    x.a.inject = e->dft.comp < 3;
    x.a.hist = x.hist;
    x.a.eps_r = e->d_eps;
    x.a.grad = gdev[q];
    x.a.grad_on = 1;
    for (int w = 0; w < (e->nd + 255) / 256; ++w) msnap.push_back(make_int2(q, w));
    for (long long w = 0; w < adj_blocks(e->ncell); ++w) mcell.push_back(make_int2(q, (int)w));
  }
  if (mcell.size() > 0x7fffffffu) FAILS(XF_EINVAL, "batch too large");
  AdjBufs D(st);
  InstAdj<T> *d_ia = (InstAdj<T> *)D.alloc(n * sizeof(InstAdj<T>));
  int2 *d_snap = (int2 *)D.alloc(msnap.size() * sizeof(int2));
  int2 *d_cell = (int2 *)D.alloc(mcell.size() * sizeof(int2));
  if (!d_ia || !d_snap || !d_cell) FAILS(XF_ENOMEM, "batch adjoint tables");
  HIPCHK(hipMemcpy(d_ia, ia.data(), n * sizeof(InstAdj<T>), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_snap, msnap.data(), msnap.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_cell, mcell.data(), mcell.size() * sizeof(int2), hipMemcpyHostToDevice));
  const unsigned nsnap = (unsigned)msnap.size(), ncell = (unsigned)mcell.size();
  auto snap = [&](int slot) {
    hipLaunchKernelGGL(kb_design_snap<T>, dim3(nsnap), dim3(256), 0, st, d_ia, d_snap, slot);
  };
  // forward, with member checkpoints at segment starts and the last
  // segment's design fields kept
  for (int sgi = 0; sgi < nseg; ++sgi) {
    const long long m = std::min<long long>(K, T_ - (long long)sgi * K);
    xf_status s;
    if (sgi + 1 < nseg) {
      for (int q = 0; q < n; ++q)
        if ((s = state_save(b->e[q], ck[q][sgi], st, &CK)) != XF_OK) return s;
      if ((s = batch_run<T>(b, m)) != XF_OK) return s;
      continue;
    }
    snap(0);
    for (long long l = 0; l < m; ++l) {
      if ((s = batch_run<T>(b, 1)) != XF_OK) return s;
      snap((int)(l + 1));
    }
  }
  for (int q = 0; q < n; ++q) {
    xf_engine *e = b->e[q];
    xf_status s = state_save(e, ck[q][nseg], st, &CK);
    if (s != XF_OK) return s;
    const size_t np = (size_t)(e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
    HIPCHK(hipMemcpyAsync(Ffin[q], e->dft.acc, np * sizeof(double2), hipMemcpyDeviceToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  for (int q = 0; q < n; ++q) {
    xf_engine *e = b->e[q];
    const size_t np = (size_t)(e->dft.hi1 - e->dft.lo1) * (e->dft.hi2 - e->dft.lo2);
    std::vector<double2> F(np);
    HIPCHK(hipMemcpy(F.data(), Ffin[q], np * sizeof(double2), hipMemcpyDeviceToHost));
    double Jv = 0;
    for (const double2 &f : F) Jv += f.x * f.x + f.y * f.y;
    Js[q] = Jv;
  }
  // The replays' curl launches cover only the tiles within reach of each
  // member's design box by the segment's end, as a single engine's do:
  // replay step l of m takes the map of margin m - l, which lists every
  // member's tiles that meet its box grown by that many cells.  Margins
  // from mfull on cover every grid and take the batch's own map.
  int mfull = 0;
  for (int q = 0; q < n; ++q) mfull = std::max(mfull, replay_window_full(b->e[q]));
  const int mw = nseg > 1 ? std::min(K, mfull - 1) : 0;  // margins 1 .. mw get a map
  std::vector<long long> woff((size_t)mw + 2, 0);
  int2 *d_wmap = nullptr;
  if (mw > 0) {
    std::vector<int2> wmap;
    const std::vector<Tiling> &tls = b->tls;
    for (int mg = 1; mg <= mw; ++mg) {
      woff[mg] = (long long)wmap.size();
      for (int q = 0; q < n; ++q) {
        int w[6];
        replay_window(b->e[q], mg, w);
        for (int x = 0; x < tls[q].nblocks; ++x)
          if (block_meets_window(b->e[q], tls[q], x, w)) wmap.push_back(make_int2(q, x));
      }
    }
    woff[mw + 1] = (long long)wmap.size();
    if (!(d_wmap = (int2 *)B.alloc(wmap.size() * sizeof(int2), false))) FAILS(XF_ENOMEM, "replay maps");
    HIPCHK(hipMemcpy(d_wmap, wmap.data(), wmap.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  // The reverse steps launch only the workgroups the adjoints can have
  // reached, as a single engine's do: a member's cell block w can hold a
  // nonzero adjoint from reverse step enter(w) on, the first step whose
  // window (the planes within step + 2 of the objective plane's i extent)
  // meets the block's planes.  A second map lists every member's blocks
  // sorted by that step, so reverse step r launches its first ready[r]
  // entries; from step rfull on every block is in and the launches take
  // the map in its own order.  The kernels remap a launch's workgroups into
  // contiguous runs per XCD over whatever map and count they get.
  std::vector<long long> ready;
  int2 *d_rev = nullptr;
  {
    std::vector<std::pair<int, int2>> ent;
    ent.reserve(mcell.size());
    for (int q = 0; q < n; ++q) {
      const xf_engine *e = b->e[q];
      const DftPlane &d = e->dft;
      const long long pi0 = d.axis == 0 ? d.index : d.lo1, pi1 = d.axis == 0 ? d.index + 1 : d.hi1;
      const long long nb = adj_blocks(e->ncell);
      for (long long w = 0; w < nb; ++w) {
        const long long ilo = w * 256 / e->si, ihi = std::min((long long)e->nx, (w * 256 + 255) / e->si);
        const long long r = std::max(0LL, std::max(pi0 - 2 - ihi, ilo - pi1 - 1));
        ent.push_back({(int)r, make_int2(q, (int)w)});
      }
    }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const std::pair<int, int2> &x, const std::pair<int, int2> &y) { return x.first < y.first; });
    const int rfull = ent.empty() ? 0 : ent.back().first;
    if (rfull > 0) {
      // ready[r]: the entries with enter <= r
      ready.assign((size_t)rfull, 0);
      std::vector<int2> rev(ent.size());
      for (size_t x = 0; x < ent.size(); ++x) {
        rev[x] = ent[x].second;
        if (ent[x].first < rfull) ++ready[ent[x].first];
      }
      for (int r = 1; r < rfull; ++r) ready[r] += ready[r - 1];
      if (!(d_rev = (int2 *)D.alloc(rev.size() * sizeof(int2), false))) FAILS(XF_ENOMEM, "batch adjoint tables");
      HIPCHK(hipMemcpy(d_rev, rev.data(), rev.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
  }
  long long rstep = 0;  // reverse steps taken
  // reverse, segment by segment
  int par = 0;  // which psi-adjoint copy holds the current values
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    const long long t0 = (long long)sgi * K, m = std::min<long long>(K, T_ - t0);
    xf_status s;
    if (sgi + 1 < nseg) {
      for (int q = 0; q < n; ++q) {
        if ((s = state_load(b->e[q], ck[q][sgi], st)) != XF_OK) return s;
        b->e[q]->cur = t0;
      }
      snap(0);
      for (long long l = 0; l < m; ++l) {
        const long long mg = m - l;
        if (mg <= mw)
          s = batch_run<T>(b, 1, true, d_wmap + woff[mg], (int)(woff[mg + 1] - woff[mg]));
        else
          s = batch_run<T>(b, 1, true);
        if (s != XF_OK) return s;
        snap((int)(l + 1));
      }
    }
    for (long long l = m - 1; l >= 0; --l, ++rstep) {
      const bool part = rstep < (long long)ready.size();
      const int2 *mp = part ? d_rev : d_cell;
      const dim3 nw((unsigned)(part ? ready[rstep] : ncell));
      hipLaunchKernelGGL(kb_adj_e<T>, nw, dim3(256), 0, st, d_ia, mp, t0 + l, (int)l, par);
      hipLaunchKernelGGL(kb_adj_h<T>, nw, dim3(256), 0, st, d_ia, mp, par);
      par ^= 1;
    }
    HIPCHK(hipGetLastError());
  }
  // leave every member at the end of the forward run
  for (int q = 0; q < n; ++q) {
    xf_status s = state_load(b->e[q], ck[q][nseg], st);
    if (s != XF_OK) return s;
    b->e[q]->cur = T_;
  }
  HIPCHK(hipStreamSynchronize(st));
  for (int q = 0; q < n; ++q)
    HIPCHK(hipMemcpy(grads[q], gdev[q], (size_t)b->e[q]->nd * sizeof(double), hipMemcpyDeviceToHost));
  return XF_OK;
}

extern "C" xf_status xf_batch_grad_eps(xf_batch *b, int checkpoint_every, double *const *grads, double *J) {
  if (!b || !grads || !J || checkpoint_every < 0) FAILS(XF_EINVAL, "bad argument");
  double bytes = 0;
  for (size_t q = 0; q < b->e.size(); ++q) {
    xf_engine *e = b->e[q];
    if (!grads[q]) FAILS(XF_EINVAL, "null gradient buffer %zu", q);
    if (!e->has_design || !e->has_dft)
      FAILS(XF_ESTATE, "member %zu: the gradient needs a design box and a DFT plane", q);
    if (has_inductor(e)) FAILS(XF_EINVAL, "member %zu: the gradient does not transpose inductors", q);
    if (has_non_drude(e))
      FAILS(XF_EINVAL, "member %zu: the gradient covers Drude media only, not Debye or Lorentz", q);
    if (e->cur != 0) FAILS(XF_ESTATE, "member %zu is at step %lld; the gradient runs from step 0", q, e->cur);
    if (e->total < 1 || e->total != b->e[0]->total)
      FAILS(XF_ESTATE, "member %zu: every member needs the same step horizon", q);
    HIPCHK(hipStreamSynchronize(e->stream));
    bytes += (double)(e->total + 1) * e->nd * 3 * e->esz;
  }
  int K = checkpoint_every;
  if (K == 0) K = bytes <= history_budget() ? (int)std::min<long long>(b->e[0]->total, 1 << 30) : 100;
  return b->prec == 64 ? batch_grad_run<double>(b, K, grads, J) : batch_grad_run<float>(b, K, grads, J);
}
