// engine_kernels.hpp -- gfx950 kernels of the FDTD field engine.
//
// Layout in HBM: every field component and the material index use one padded
// (nx+1) x (ny+1) x pz grid, z fastest, pz a multiple of 32 elements so each
// z row starts on a 128-byte line; the x and y CPML psi arrays use the same
// z rows over their slabs (the two arrays of an axis interleaved).  The six
// components share one index, so a cell's six values and its material byte
// are read with one address computation.  A component's reference shape
// (simulation/simulation_init.cpp:28-35) is a sub-box of this grid; entries
// outside it stay zero and are never read.
//
// Arithmetic follows the reference expression by expression and the file is
// compiled with -ffp-contract=off, so the FP64 instantiation rounds exactly
// like the reference's generic x86-64 build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xfdtd_engine.h"

namespace xfk {

enum { EX = 0, EY, EZ, HX, HY, HZ };
enum {
  CB_EXHY = 0, CB_EXHZ, CB_EYHZ, CB_EYHX, CB_EZHX, CB_EZHY,
  CB_HXEY, CB_HXEZ, CB_HYEZ, CB_HYEX, CB_HZEX, CB_HZEY
};
// psi arrays: E: ex_y, ex_z, ey_x, ey_z, ez_x, ez_y; H: hx_y, hx_z, hy_x, hy_z, hz_x, hz_y
enum { P_AY = 0, P_AZ, P_BX, P_BZ, P_CX, P_CY };

template <typename T>
struct Grid {
  T *f[6];
  const uint8_t *mat;
  // coefficient records per side (0: E, 1: H): self terms [nmat][4]; curl
  // and psi coefficient pairs per axis [n_axis+1][nmat][2]
  const T *self[2];
  const T *kt[2], *jt[2], *it[2];
  const T *kp[2], *jp[2], *ip[2];
  // CPML per axis
  int dn[3], dp[3];
  const T *be[3], *ae[3], *bm[3], *am[3];
  T *psi_e[6], *psi_h[6];
  int nx, ny, nz, pz, nmat;
  // 1: every material has material 0's H coefficients (no magnetic media),
  // so the H kernels use material 0 and read no material byte
  int h_uni;
  long long sj, si;  // strides of j and i in elements
  // design box [db0,db1) x [db2,db3) x [db4,db5) with per-cell E coefficients
  // dcoef[cell][9] = cexe, ceye, ceze, cexhy, ceyhx, cexhz, cezhx, ceyhz,
  // cezhy (an empty box when db0 == db1)
  int db[6];
  const T *dcoef;
  // when set, the E kernel also writes each design cell's new (Ex, Ey, Ez)
  // to snap[design cell][3] (a gradient's design-box history of the step)
  T *snap;
  // Drude box [xb0,xb1) x [xb2,xb3) x [xb4,xb5) (empty when xb0 == xb1):
  // per-material ADE records drc[m][XF_ADE_STRIDE] (kind at XF_ADE_KIND:
  // 0 none, 1 Drude, 2 Debye, 3 Lorentz; the layout of xf_engine_set_dispersive)
  // (cb = 0 for a material that is not a Drude medium) and the polarisation
  // currents of the box's cells, dj[component][cell]
  int xb[6];
  const T *drc;
  T *dj[3];
  T *djp[3], *dep[3];  // Lorentz: the previous step's J and E (null without Lorentz media)
  // divergence flag: the H kernel lowers it to the index of a step whose H
  // came out non-finite (all ones while every value is finite)
  unsigned long long *div;
};

template <typename T>
__device__ __forceinline__ bool in_design(const Grid<T> &g, int i, int j, int k) {
  return i >= g.db[0] && i < g.db[1] && j >= g.db[2] && j < g.db[3] && k >= g.db[4] && k < g.db[5];
}
template <typename T>
__device__ __forceinline__ int design_cell(const Grid<T> &g, int i, int j, int k) {
  return ((i - g.db[0]) * (g.db[3] - g.db[2]) + (j - g.db[2])) * (g.db[5] - g.db[4]) + (k - g.db[4]);
}
template <typename T>
__device__ __forceinline__ bool in_drude(const Grid<T> &g, int i, int j, int k) {
  return i >= g.xb[0] && i < g.xb[1] && j >= g.xb[2] && j < g.xb[3] && k >= g.xb[4] && k < g.xb[5];
}
template <typename T>
__device__ __forceinline__ int drude_cell(const Grid<T> &g, int i, int j, int k) {
  return ((i - g.xb[0]) * (g.xb[3] - g.xb[2]) + (j - g.xb[2])) * (g.xb[5] - g.xb[4]) + (k - g.xb[4]);
}

// DrudeMedium::updateE{x,y,z} (dispersive_material.cpp:146-193) for one
// component of one cell: `curl` is the reference's bracket without j_sum,
// e.g. -(Hy(k) - Hy(k-1)) + (Hz(j) - Hz(j-1)) for Ex.  Returns the new E and
// advances J in place.
template <typename T>
__device__ __forceinline__ T drude_e(const T *c, T *jp, T e_old, T curl) {
  const T jv = *jp;
  const T j_sum = T(0) + c[5] * (c[2] * jv);
  const T e_new = c[0] * e_old + c[1] * (curl - j_sum);
  *jp = c[3] * jv + c[4] * (e_new + e_old);
  return e_new;
}

// one dispersive E cell: DrudeMedium / DebyMedium / LorentzMedium::updateE
// (dispersive_material.cpp:146-193, 224-275, 38-117) in the reference's
// expression order; jq, eq are the Lorentz previous J and E
template <typename T>
__device__ __forceinline__ T ade_e(const T *c, T *jp, T *jq, T *eq, T e_old, T curl) {
  const T kind = c[XF_ADE_KIND];
  if (kind == T(3)) {
    const T jv = *jp, jpv = *jq, epv = *eq;
    const T j_sum = T(0) + c[5] * (c[2] * jv + c[7] * jpv);
    const T e_new = c[0] * e_old + c[1] * epv + c[8] * (curl - j_sum);
    *jp = c[3] * jv + c[7] * jpv + c[4] * (e_new - epv) / c[10];
    *eq = e_old;
    *jq = jv;
    return e_new;
  }
  if (kind == T(2)) {
    const T jv = *jp;
    const T j_sum = T(0) + c[5] * (c[2] * jv);
    const T e_new = c[0] * e_old + c[1] * (curl - j_sum);
    *jp = c[3] * jv + c[4] * (e_new - e_old) / c[9];
    return e_new;
  }
  return drude_e(c, jp, e_old, curl);
}

__device__ __forceinline__ int e_layer(int n, int N, int dn, int dp) {
  if (n >= 1 && n <= dn) return n - 1;
  if (dp > 0 && n >= N - dp && n <= N - 1) return dn + n - (N - dp);
  return -1;
}
__device__ __forceinline__ int h_layer(int n, int N, int dn, int dp) {
  if (n < dn) return n;
  if (dp > 0 && n >= N - dp && n < N) return dn + n - (N - dp);
  return -1;
}

// Launch tiling of the curl kernels: a 1-D grid of tiles, each TZ x TY
// threads over (k, j) marching over TX x-planes.  The hardware hands
// consecutive workgroup ids to the 8 XCDs round-robin, and each XCD has its
// own L2.  By default tile ids follow workgroup ids (z, y, x order): the eight
// XCDs stream through neighbouring tiles together, and a tile's y-1 row and
// x-1 plane, read by another XCD, come back through the fabric.  xcd_runs
// instead gives each XCD one contiguous run of tiles, which keeps those
// neighbours in its L2 (about 20% less read traffic) but measures about 15%
// slower on the 400^3 grid; so does giving each XCD a band of y rows.
struct Tiling {
  int tz, ty, tx;           // threads along k, j (ty = 0: flat); planes along i
  int ntz, nty, ntx;        // tile counts
  int nblocks;
  int krows, jrows;         // coefficient-table slots of a tile along k and j
  int x_fastest;            // tile order: x chunks first (else z, y, x)
  int xcd_runs;             // one contiguous run of tile ids per XCD
  int xcd_cols;             // all k tiles of a (j, i) column on one XCD
  int oz, oy, ox;           // first tile along k, j, i (a launch over a box of tiles)
};

// tile of workgroup b; false for a padding workgroup of the column order
// (also called on the host, to list the tiles of a batched launch that meet
// a window)
__host__ __device__ __forceinline__ bool tile_of_block(const Tiling &t, int b, int &bz, int &by, int &bx) {
  const int q = t.nblocks >> 3, r = t.nblocks & 7;
  const int xcd = b & 7, local = b >> 3;
  if (t.xcd_cols) {
    bz = local % t.ntz;
    const int col = (local / t.ntz) * 8 + xcd;
    by = col % t.nty;
    bx = col / t.nty;
    return bx < t.ntx;
  }
  const int id = t.xcd_runs ? xcd * q + (xcd < r ? xcd : r) + local : b;
  if (t.x_fastest) {
    bx = id % t.ntx;
    const int rest = id / t.ntx;
    bz = rest % t.ntz;
    by = rest / t.ntz;
    return true;
  }
  bz = id % t.ntz;
  const int rest = id / t.ntz;
  by = rest % t.nty;
  bx = rest / t.nty;
  return true;
}

// Cell of a thread.  Two tile shapes: a TZ x TY block of (k, j), or (TY = 0)
// a flat run of TZ consecutive cells of the padded (j, k) plane, which keeps
// each wave on one contiguous 256-byte (FP32) piece of every field and lets
// a workgroup cover 1 KB of a plane row.  ks / js are the thread's slots in
// the tile's staged k and j coefficient tables.
struct Cell {
  int j, k, ks, js, j_first;
};

__device__ __forceinline__ Cell thread_cell(const Tiling &tl, int bz, int by, int pz) {
  Cell c;
  if (tl.ty == 0) {
    const int p = bz * tl.tz + (int)threadIdx.x;
    c.j = p / pz;
    c.k = p - c.j * pz;
    c.ks = (int)threadIdx.x;
    c.j_first = (bz * tl.tz) / pz;
    c.js = c.j - c.j_first;
  } else {
    c.ks = (int)(threadIdx.x % tl.tz);
    c.js = (int)(threadIdx.x / tl.tz);
    c.k = bz * tl.tz + c.ks;
    c.j_first = by * tl.ty;
    c.j = c.j_first + c.js;
  }
  return c;
}

// ---------------------------------------------------------------------------
// Coefficient tables of one tile, staged in LDS when they fit.  The curl
// coefficients of a cell are functions of its material and of one node index
// (fdtd_basic_coff.cpp:93-121), so a tile marching over TX planes needs, per
// material: 3 self terms, 2 tables over its k slots, 2 over its j rows and 2
// over its i planes.
// Coefficient records.  For each half-step (side 0: E, 1: H) the engine
// keeps, per material, the three self terms in one 4-wide record, and per
// axis a table of pairs [n][material] holding the two curl coefficients of
// the two components whose update differentiates along that axis (and the
// same for the CPML psi coefficients), so one LDS read returns both.
template <typename T>
struct alignas(2 * sizeof(T)) P2 {
  T x, y;
};
template <typename T>
struct alignas(4 * sizeof(T)) P4 {
  T x, y, z, w;
};

// The tables of one tile: staged in LDS, or views of the global tables.
template <typename T>
struct TileTables {
  const P4<T> *self;          // [nmat]
  const P2<T> *k, *j, *i;     // curl pairs [slot][nmat]
  const P2<T> *pk, *pj, *pi;  // psi pairs [slot][nmat]
  int kcol, jcol, i0;         // this thread's k slot * nmat, j slot * nmat; tile's first plane
};

template <typename T>
__host__ __device__ inline size_t lds_table_bytes(int nmat, const Tiling &tl) {
  return sizeof(T) * (size_t)nmat * (4 + 4 * (size_t)tl.krows + 4 * (size_t)tl.jrows + 4 * (size_t)tl.tx);
}

// does the index range [lo, hi] meet a CPML slab of an axis of N cells?  The
// E layers are n in [1, dn] and [N-dp, N-1], the H layers [0, dn) and
// [N-dp, N); both lie inside [0, dn] and [N-dp-1, N].
__device__ __forceinline__ bool meets_slab(int lo, int hi, int N, int dn, int dp) {
  return (dn > 0 && lo <= dn) || (dp > 0 && hi >= N - dp - 1);
}

// Stage the tables of one tile in LDS (cooperative; ends with a barrier).
// The psi tables of an axis are staged only when the tile meets a slab of
// that axis: a thread reads them only for a cell inside one.
template <typename T>
__device__ __forceinline__ TileTables<T> stage_tables(unsigned char *lds, const Grid<T> &g, int side,
                                                     const Tiling &tl, const Cell &cell, int i0,
                                                     bool psi_k, bool psi_j, bool psi_i) {
  const int nmat = g.nmat;
  const int nk = nmat * tl.krows, nj = nmat * tl.jrows, ni = nmat * tl.tx;
  P4<T> *sself = reinterpret_cast<P4<T> *>(lds);
  P2<T> *p = reinterpret_cast<P2<T> *>(sself + nmat);
  P2<T> *sk = p; p += nk;
  P2<T> *qk = p; p += nk;
  P2<T> *sj = p; p += nj;
  P2<T> *qj = p; p += nj;
  P2<T> *si = p; p += ni;
  P2<T> *qi = p;
  const P4<T> *gself = reinterpret_cast<const P4<T> *>(g.self[side]);
  const P2<T> *gk = reinterpret_cast<const P2<T> *>(g.kt[side]);
  const P2<T> *gj = reinterpret_cast<const P2<T> *>(g.jt[side]);
  const P2<T> *gi = reinterpret_cast<const P2<T> *>(g.it[side]);
  const P2<T> *hk = reinterpret_cast<const P2<T> *>(g.kp[side]);
  const P2<T> *hj = reinterpret_cast<const P2<T> *>(g.jp[side]);
  const P2<T> *hi = reinterpret_cast<const P2<T> *>(g.ip[side]);
  const int nth = blockDim.x;
  for (int q = threadIdx.x; q < nmat; q += nth) sself[q] = gself[q];
  if (cell.js == 0 || tl.ty == 0) {
    // each k slot is written by the thread that owns it (row 0 of a 2-D tile)
    const int src = min(cell.k, g.nz) * nmat, dst = cell.ks * nmat;
    for (int m = 0; m < nmat; ++m) {
      sk[dst + m] = gk[src + m];
      if (psi_k) qk[dst + m] = hk[src + m];
    }
  }
  for (int q = threadIdx.x; q < nj; q += nth) {
    const int n = min(cell.j_first + q / nmat, g.ny), o = n * nmat + q % nmat;
    sj[q] = gj[o];
    if (psi_j) qj[q] = hj[o];
  }
  for (int q = threadIdx.x; q < ni; q += nth) {
    const int n = min(i0 + q / nmat, g.nx), o = n * nmat + q % nmat;
    si[q] = gi[o];
    if (psi_i) qi[q] = hi[o];
  }
  __syncthreads();
  TileTables<T> tt;
  tt.self = sself;
  tt.k = sk; tt.j = sj; tt.i = si;
  tt.pk = qk; tt.pj = qj; tt.pi = qi;
  tt.kcol = cell.ks * nmat;
  tt.jcol = cell.js * nmat;
  tt.i0 = i0;
  return tt;
}

// the same view over the global tables (no staging)
template <typename T>
__device__ __forceinline__ TileTables<T> global_tables(const Grid<T> &g, int side, const Cell &cell) {
  TileTables<T> tt;
  tt.self = reinterpret_cast<const P4<T> *>(g.self[side]);
  tt.k = reinterpret_cast<const P2<T> *>(g.kt[side]);
  tt.j = reinterpret_cast<const P2<T> *>(g.jt[side]);
  tt.i = reinterpret_cast<const P2<T> *>(g.it[side]);
  tt.pk = reinterpret_cast<const P2<T> *>(g.kp[side]);
  tt.pj = reinterpret_cast<const P2<T> *>(g.jp[side]);
  tt.pi = reinterpret_cast<const P2<T> *>(g.ip[side]);
  tt.kcol = min(cell.k, g.nz) * g.nmat;
  tt.jcol = min(cell.j, g.ny) * g.nmat;
  tt.i0 = 0;
  return tt;
}

// block-uniform test: does this tile meet a CPML slab along each axis?
template <typename T>
__device__ __forceinline__ void tile_slabs(const Grid<T> &g, const Tiling &tl, int bz, const Cell &cell, int i0,
                                           int i1, bool &sk, bool &sj, bool &si) {
  if (tl.ty == 0)
    sk = g.dn[2] + g.dp[2] > 0;
  else
    sk = meets_slab(bz * tl.tz, bz * tl.tz + tl.tz - 1, g.nz, g.dn[2], g.dp[2]);
  sj = meets_slab(cell.j_first, cell.j_first + tl.jrows - 1, g.ny, g.dn[1], g.dp[1]);
  si = meets_slab(i0, i1 - 1, g.nx, g.dn[0], g.dp[0]);
}

// Field loads and stores by 32-bit byte offset from a wave-uniform base: all
// six components and the material map share one offset per cell, which
// lets the compiler use the SGPR-base + VGPR-offset addressing form.
template <typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T *base, unsigned off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}
__device__ __forceinline__ unsigned ldm(const uint8_t *base, unsigned cell) { return base[cell]; }

// psi element offsets.  Slab arrays of the y axis are [nx+1][Ly][pz], of the
// z axis [nx+1][ny+1][Lz], of the x axis [Lx][ny+1][pz]; every offset fits 32
// bits because every slab array is smaller than a field.
__device__ __forceinline__ unsigned psi_y(int i, int ly, int k, int Ly, int pz) {
  return ((unsigned)i * (unsigned)Ly + (unsigned)ly) * (unsigned)pz + (unsigned)k;
}
__device__ __forceinline__ unsigned psi_z(int i, int j, int lz, int nm1y, int Lz) {
  return ((unsigned)i * (unsigned)nm1y + (unsigned)j) * (unsigned)Lz + (unsigned)lz;
}
// The two psi arrays of an axis share one allocation, interleaved: x (BX,
// CX), y (AY, CY), z (AZ, BZ), the first at even elements; psi_e[first] and
// psi_e[second] point at its element 0 and 1, so slot q's cell idx is at
// psi_e[q][2 idx], and one two-wide access reaches both.
template <typename T>
__device__ __forceinline__ P2<T> ld_pair(const T *first, unsigned idx) {
  return reinterpret_cast<const P2<T> *>(first)[idx];
}
template <typename T>
__device__ __forceinline__ void st_pair(T *first, unsigned idx, T a, T b) {
  P2<T> v;
  v.x = a;
  v.y = b;
  reinterpret_cast<P2<T> *>(first)[idx] = v;
}
__device__ __forceinline__ unsigned psi_x(int lx, int j, int k, int nm1y, int pz) {
  return ((unsigned)lx * (unsigned)nm1y + (unsigned)j) * (unsigned)pz + (unsigned)k;
}

// CPML state of one (j, k) column: its layer along y and z (-1 outside a
// slab) and the b/a coefficients of those layers, read once per column.
template <typename T>
struct ColCpml {
  int ly, lz;
  T by, ay, bz, az;
};

template <typename T>
__device__ __forceinline__ ColCpml<T> col_cpml(const T *const *b, const T *const *a, int ly, int lz, int j,
                                               int k) {
  ColCpml<T> cc;
  cc.ly = ly;
  cc.lz = lz;
  cc.by = cc.ay = cc.bz = cc.az = T(0);
  if (ly >= 0) {
    cc.by = b[1][j];
    cc.ay = a[1][j];
  }
  if (lz >= 0) {
    cc.bz = b[2][k];
    cc.az = a[2][k];
  }
  return cc;
}

// One x-plane of a column: the loads of the cell (fields, material, the
// in-plane neighbours and, on the CPML path, the psi values), issued
// together so that a plane costs one memory round trip.
template <typename T>
struct Plane {
  unsigned m;
  T f0, f1, f2;  // the three components this half-step updates
  T g0, g1, g2;  // the three components it reads
  T n0, n1, n2, n3, n4, n5;  // neighbours (see the kernels)
  int lx;        // layer along x (-1 outside the x slabs)
  T bx, ax;      // its CPML b/a (wave-uniform)
  T ps[6];       // psi values of the cell, by P_* slot
  double z0, z1; // TFSF z-face terms of the cell (E: Ey, Ex; H: Hy, Hx), 0 if none
};

// Per-column constants of a curl kernel.
template <typename T>
struct Col {
  T *__restrict__ u0;  // updated components (Ex, Ey, Ez or Hx, Hy, Hz)
  T *__restrict__ u1;
  T *__restrict__ u2;
  const T *__restrict__ r0;  // read components (Hx, Hy, Hz or Ex, Ey, Ez)
  const T *__restrict__ r1;
  const T *__restrict__ r2;
  const uint8_t *__restrict__ mat;
  // byte offsets of the k and j neighbours; 0 where the neighbour does not
  // exist, in which case every term that would use it is discarded
  unsigned dk, dj;
  bool kin, jin;  // column inside the Yee ranges along k and j
  ColCpml<T> cp;
  int j, k;
};

// ---------------------------------------------------------------------------
// TFSF (tfsf/tfsf_3d.cpp).  The 1-D incident line, its projections and the
// face-correction arithmetic are FP64 for both field precisions: the line is
// a few hundred cells, and an FP32 field then takes one rounding per
// correction, of the FP64 sum, instead of carrying an FP32 line's rounding
// into the scattered-field region.
struct Tfsf {
  int li, ri, lj, rj, lk, rk, L;
  int two_d;  // TFSF2D: Ez, Hx, Hy only, projections without z
  const double *pf[3], *ph[3];
  double te[3], th[3];
  double ceie, ceihi, chih, chiei, c1, c2;
  double ca[3], cb[3];
  double *e_inc, *h_inc, *ab;
  const double *src;
};

// TFSF3D::getIncident{Ex..Hz} (tfsf_3d.cpp:413-483)
__device__ __forceinline__ double tfsf_inc(const Tfsf &t, int comp, int i, int j, int k) {
  double proj, s;
  const double *arr;
  switch (comp) {
    case EX: proj = t.ph[0][i - t.li + 1] + t.pf[1][j - t.lj] + t.pf[2][k - t.lk];
             arr = t.e_inc; s = t.te[0]; break;
    case EY: proj = t.pf[0][i - t.li] + t.ph[1][j - t.lj + 1] + t.pf[2][k - t.lk];
             arr = t.e_inc; s = t.te[1]; break;
    // TFSF2D::getIncident{Ez,Hx,Hy} (tfsf_2d.cpp:230-257) leave out the z
    // projection
    case EZ: proj = t.pf[0][i - t.li] + t.pf[1][j - t.lj];
             if (!t.two_d) proj = proj + t.ph[2][k - t.lk + 1];
             arr = t.e_inc; s = t.te[2]; break;
    case HX: proj = t.pf[0][i - t.li] + t.ph[1][j - t.lj + 1];
             if (!t.two_d) proj = proj + t.ph[2][k - t.lk + 1];
             proj = proj - 0.5;
             arr = t.h_inc; s = t.th[0]; break;
    case HY: proj = t.ph[0][i - t.li + 1] + t.pf[1][j - t.lj];
             if (!t.two_d) proj = proj + t.ph[2][k - t.lk + 1];
             proj = proj - 0.5;
             arr = t.h_inc; s = t.th[1]; break;
    default: proj = t.ph[0][i - t.li + 1] + t.ph[1][j - t.lj + 1] + t.pf[2][k - t.lk] - 0.5;
             arr = t.h_inc; s = t.th[2]; break;
  }
  const long long idx = (long long)proj;
  const double w = proj - (double)idx;
  return (1.0 - w) * (s * arr[idx]) + w * (s * arr[idx + 1]);
}

// The z faces in the curl kernels (ZFaces).  The cells of a z-face column
// (k = lk or rk for E, lk - 1 or rk for H) that no x or y face corrects take
// one correction each, so adding it at the end of the cell's update (after
// the CPML terms) gives the same operations in the same order as the face
// kernels.  The products coefficient * incident value are computed once per
// step by k_tfsf_zterms (tfsf_3d.cpp:368-411 for E, :259-297 for H) into
// FP64 segments [seg][i - li][j - lj] of (ni + 1) x (nj + 1) entries: E
// segments 0..3 (Ey at lk and rk: ca_z * Hx_inc at lk - 1 and rk; Ex at lk
// and rk: ca_z * Hy_inc), H segments 4..7 (Hy at lk - 1 and rk: cb_z * Ex_inc
// at lk and rk, sampled on the total-field side; Hx: cb_z * Ey_inc).  A
// curl-kernel lane loads its cell's terms with the plane's other loads.
struct ZFaces {
  const double *terms;  // null: the face kernels apply every correction
  int li, ri, lj, rj, lk, rk;
};

// DPP lane moves over the whole wave: every lane takes `src` of its
// neighbour, the end lane that has none takes `fill`.  One v_mov_dpp per
// 32 bits with `fill` as the old value, so no lane is masked off while the
// move reads: a move under a partial EXEC mask would read the masked lanes as
// out of range.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float src, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(src), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double src, double fill) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(src);
  const unsigned long long f = (unsigned long long)__double_as_longlong(fill);
  const int l2 = __builtin_amdgcn_update_dpp((int)(unsigned)(f & 0xffffffffull), (int)(unsigned)(b & 0xffffffffull),
                                             CTRL, 0xF, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp((int)(unsigned)(f >> 32), (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2));
}
// the value held by the lane below (k - 1; lane 0 takes `fill`) and above
// (k + 1; lane 63 takes `fill`)
template <typename T>
__device__ __forceinline__ T from_below(T v, T fill) { return dpp_move<0x138>(v, fill); }
template <typename T>
__device__ __forceinline__ T from_above(T v, T fill) { return dpp_move<0x130>(v, fill); }

// NF-FF z faces (NffftFd surface planes k = k0 and k1): the H kernel leaves
// the field values their stencils read in 8-value records
// rec[face][i - li][j - lj] (i < li + nxr, j < lj + nyr), so the sample
// kernel reads no z-face cache line of the fields (rec null: it reads the
// fields).  Record: Ey, Ex, Ex(k+1), Hy at kf, then Hx at kf, 0, Hy and Hx
// at kf - 1.  The threads of (i, j, kf) and (i, j, kf - 1), adjacent lanes of
// one wave, store the two halves with one four-wide store.
struct NfRec {
  void *rec;
  int li, lj, nxr, nyr, k0, k1;
};
template <typename T>
__device__ __forceinline__ void nf_record(const NfRec &nr, int i, int j, int k, T ey, T ex, T ex1, T hx, T hy) {
  const T hx_up = from_above(hx, T(0));  // lane kf - 1 takes Hx(kf)
  const int face = k >= nr.k1 - 1 ? 1 : 0;
  const int kf = face ? nr.k1 : nr.k0;
  if ((k != kf && k != kf - 1) || j < nr.lj || j >= nr.lj + nr.nyr) return;
  const bool top = k == kf;
  T *rec = (T *)nr.rec + ((face * nr.nxr + (i - nr.li)) * nr.nyr + (j - nr.lj)) * 8 + (top ? 0 : 4);
  P4<T> v;
  v.x = top ? ey : hx_up;
  v.y = top ? ex : T(0);
  v.z = top ? ex1 : hy;
  v.w = top ? hy : hx;
  *reinterpret_cast<P4<T> *>(rec) = v;
}

__device__ __forceinline__ int zseg_size(const ZFaces &z) { return (z.ri - z.li + 1) * (z.rj - z.lj + 1); }
__device__ __forceinline__ int zterm_index(const ZFaces &z, int seg, int i, int j) {
  return seg * zseg_size(z) + (i - z.li) * (z.rj - z.lj + 1) + (j - z.lj);
}
__global__ void __launch_bounds__(256) k_tfsf_zterms(Tfsf t, double *buf, int s0, int s1) {
  const int nj1 = t.rj - t.lj + 1, per = (t.ri - t.li + 1) * nj1;
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= (s1 - s0) * per) return;
  const int s = s0 + q / per, r = q % per;
  const int i = t.li + r / nj1, j = t.lj + r % nj1;
  double v = 0.0;
  switch (s) {
    case 0: if (i > t.li && i < t.ri && j < t.rj) v = t.ca[2] * tfsf_inc(t, HX, i, j, t.lk - 1); break;
    case 1: if (i > t.li && i < t.ri && j < t.rj) v = t.ca[2] * tfsf_inc(t, HX, i, j, t.rk); break;
    case 2: if (i < t.ri && j > t.lj && j < t.rj) v = t.ca[2] * tfsf_inc(t, HY, i, j, t.lk - 1); break;
    case 3: if (i < t.ri && j > t.lj && j < t.rj) v = t.ca[2] * tfsf_inc(t, HY, i, j, t.rk); break;
    case 4: if (i < t.ri) v = t.cb[2] * tfsf_inc(t, EX, i, j, t.lk); break;
    case 5: if (i < t.ri) v = t.cb[2] * tfsf_inc(t, EX, i, j, t.rk); break;
    case 6: if (j < t.rj) v = t.cb[2] * tfsf_inc(t, EY, i, j, t.lk); break;
    default: if (j < t.rj) v = t.cb[2] * tfsf_inc(t, EY, i, j, t.rk); break;
  }
  buf[q + s0 * per] = v;
}

// ---------------------------------------------------------------------------
// E half-step: Simulation::updateE (simulation_update.cpp:170-196) fused with
// PML::updateE of every face (perfect_match_layer.cpp:574-718), faces in x,
// y, z order.  One thread owns a (j,k) column and marches over TX x-planes.
// Planes that meet no CPML slab take the SLAB = false path, which has no psi
// code at all.  Every plane stores all three components: cells outside the
// Yee and CPML ranges get their own value back (each cell is owned by one
// thread, so this is exact), which keeps the stores free of branches.
//
// Plane fields: f = Ex, Ey, Ez; g = Hx, Hy, Hz; n0 = Hy(k-1), n1 = Hx(k-1),
// n2 = Hz(j-1), n3 = Hx(j-1).
template <typename T, bool SLAB>
__device__ __forceinline__ void load_plane_e(const Grid<T> &g, const Col<T> &q, unsigned cell, int i,
                                             Plane<T> &p, const ZFaces &zf, bool zplane) {
  const unsigned o = cell * (unsigned)sizeof(T);
  p.m = ldm(q.mat, cell);
  p.f0 = ldo(q.u0, o);
  p.f1 = ldo(q.u1, o);
  p.f2 = ldo(q.u2, o);
  p.g0 = ldo(q.r0, o);
  p.g1 = ldo(q.r1, o);
  p.g2 = ldo(q.r2, o);
  p.n0 = ldo(q.r1, o - q.dk);
  p.n1 = ldo(q.r0, o - q.dk);
  p.n2 = ldo(q.r2, o - q.dj);
  p.n3 = ldo(q.r0, o - q.dj);
  if constexpr (SLAB) {
    p.lx = e_layer(i, g.nx, g.dn[0], g.dp[0]);
    p.bx = p.ax = T(0);
    for (int s = 0; s < 6; ++s) p.ps[s] = T(0);
    const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx, j = q.j, k = q.k;
    const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = g.ny + 1;
    if (lx >= 0) {
      p.bx = g.be[0][i];
      p.ax = g.ae[0][i];
    }
    // psi pairs, each one two-wide load: x (BX for Ey, CX for Ez), y (AY for
    // Ex, CY for Ez), z (AZ for Ex, BZ for Ey)
    if (lx >= 0 && (j < g.ny || k < g.nz)) {
      const P2<T> v = ld_pair(g.psi_e[P_BX], psi_x(lx, j, k, nm1y, g.pz));
      p.ps[P_BX] = v.x;
      p.ps[P_CX] = v.y;
    }
    if (ly >= 0 && (i < g.nx || k < g.nz)) {
      const P2<T> v = ld_pair(g.psi_e[P_AY], psi_y(i, ly, k, Ly, g.pz));
      p.ps[P_AY] = v.x;
      p.ps[P_CY] = v.y;
    }
    if (lz >= 0 && (i < g.nx || j < g.ny)) {
      const P2<T> v = ld_pair(g.psi_e[P_AZ], psi_z(i, j, lz, nm1y, Lz));
      p.ps[P_AZ] = v.x;
      p.ps[P_BZ] = v.y;
    }
    // TFSF z faces: Ey off the x faces, Ex off the y faces
    p.z0 = p.z1 = 0.0;
    if (zplane && (k == zf.lk || k == zf.rk)) {
      const int up = k == zf.rk;
      if (j >= zf.lj && j < zf.rj && i > zf.li && i < zf.ri) p.z0 = zf.terms[zterm_index(zf, up, i, j)];
      if (i < zf.ri && j > zf.lj && j < zf.rj) p.z1 = zf.terms[zterm_index(zf, 2 + up, i, j)];
    }
  }
}

template <typename T, bool SLAB, bool SPECIAL>
__device__ __forceinline__ void compute_plane_e(const Grid<T> &g, const TileTables<T> &tt, const Col<T> &q,
                                                const Plane<T> &p, int i, unsigned cell, T hz_prev,
                                                T hy_prev, const ZFaces &zf, bool zt) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, j = q.j, k = q.k;
  const int m = (int)p.m;
  const T ex = p.f0, ey = p.f1, ez = p.f2, hx = p.g0, hy = p.g1, hz = p.g2;
  const T dhy_k = hy - p.n0;     // Hy(k) - Hy(k-1)
  const T dhz_j = hz - p.n2;     // Hz(j) - Hz(j-1)
  const T dhx_k = hx - p.n1;     // Hx(k) - Hx(k-1)
  const T dhz_i = hz - hz_prev;  // Hz(i) - Hz(i-1)
  const T dhx_j = hx - p.n3;     // Hx(j) - Hx(j-1)
  const T dhy_i = hy - hy_prev;  // Hy(i) - Hy(i-1)
  const int mi = (i - tt.i0) * g.nmat + m;
  P4<T> cs = tt.self[m];
  P2<T> ck = tt.k[tt.kcol + m], cj = tt.j[tt.jcol + m], ci = tt.i[mi];
  if constexpr (SLAB && SPECIAL) {
    // design cells carry their own coefficients (outside every CPML slab)
    if (in_design(g, i, j, k)) {
      const T *d = g.dcoef + 9 * design_cell(g, i, j, k);
      cs.x = d[0]; cs.y = d[1]; cs.z = d[2];
      ck.x = d[3]; ck.y = d[4]; cj.x = d[5]; cj.y = d[6]; ci.x = d[7]; ci.y = d[8];
    }
  }
  // cexe*Ex + cexhy*dHy/dz + cexhz*dHz/dy, and likewise for Ey, Ez
  const T yx = cs.x * ex + ck.x * dhy_k + cj.x * dhz_j;
  const T yy = cs.y * ey + ck.y * dhx_k + ci.x * dhz_i;
  const T yz = cs.z * ez + cj.y * dhx_j + ci.y * dhy_i;
  const bool ix = i >= 1 && i < nx;
  // Ex (nx, ny+1, nz+1): Yee i < nx, j in [1,ny), k in [1,nz)
  // Ey (nx+1, ny, nz+1): Yee i in [1,nx), j < ny, k in [1,nz)
  // Ez (nx+1, ny+1, nz): Yee i in [1,nx), j in [1,ny), k < nz
  T vx = (i < nx && q.jin && q.kin) ? yx : ex;
  T vy = (ix && j < ny && q.kin) ? yy : ey;
  T vz = (ix && q.jin && k < nz) ? yz : ez;
  if constexpr (SLAB && SPECIAL) {
    // Drude cells (outside every CPML slab) replace the Yee update
    // (simulation_update.cpp:221-280)
    if (in_drude(g, i, j, k)) {
      const T *dc = g.drc + XF_ADE_STRIDE * m;
      if (dc[XF_ADE_KIND] != T(0)) {
        const int dq = drude_cell(g, i, j, k);
        // the Lorentz state pointers are only read for a Lorentz record
        T *jq[3], *eq[3];
        for (int c = 0; c < 3; ++c) {
          jq[c] = g.djp[c] ? g.djp[c] + dq : nullptr;
          eq[c] = g.dep[c] ? g.dep[c] + dq : nullptr;
        }
        if (i < nx && q.jin && q.kin) vx = ade_e(dc, g.dj[0] + dq, jq[0], eq[0], ex, -dhy_k + dhz_j);
        if (ix && j < ny && q.kin) vy = ade_e(dc, g.dj[1] + dq, jq[1], eq[1], ey, -dhz_i + dhx_k);
        if (ix && q.jin && k < nz) vz = ade_e(dc, g.dj[2] + dq, jq[2], eq[2], ez, -dhx_j + dhy_i);
      }
    }
  }
  if constexpr (SLAB) {
    const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx;
    const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = ny + 1;
    // the new psi of each pair (the old value where the component lies
    // outside its range), stored with one two-wide store; then the terms,
    // added in the order of the reference's updates
    const bool ux = i < nx, uy = j < ny, uz = k < nz;
    T n_ay = p.ps[P_AY], n_az = p.ps[P_AZ], n_bx = p.ps[P_BX], n_bz = p.ps[P_BZ], n_cx = p.ps[P_CX],
      n_cy = p.ps[P_CY];
    if (lx >= 0 && (uy || uz)) {
      if (uy) n_bx = p.bx * p.ps[P_BX] + p.ax * dhz_i;
      if (uz) n_cx = p.bx * p.ps[P_CX] + p.ax * dhy_i;
      st_pair(g.psi_e[P_BX], psi_x(lx, j, k, nm1y, g.pz), n_bx, n_cx);
    }
    if (ly >= 0 && (ux || uz)) {
      if (ux) n_ay = q.cp.by * p.ps[P_AY] + q.cp.ay * dhz_j;
      if (uz) n_cy = q.cp.by * p.ps[P_CY] + q.cp.ay * dhx_j;
      st_pair(g.psi_e[P_AY], psi_y(i, ly, k, Ly, g.pz), n_ay, n_cy);
    }
    if (lz >= 0 && (ux || uy)) {
      if (ux) n_az = q.cp.bz * p.ps[P_AZ] + q.cp.az * dhy_k;
      if (uy) n_bz = q.cp.bz * p.ps[P_BZ] + q.cp.az * dhx_k;
      st_pair(g.psi_e[P_AZ], psi_z(i, j, lz, nm1y, Lz), n_az, n_bz);
    }
    if (ux) {
      if (ly >= 0) vx = vx + tt.pj[tt.jcol + m].x * n_ay;
      if (lz >= 0) vx = vx + tt.pk[tt.kcol + m].x * n_az;
    }
    if (uy) {
      if (lx >= 0) vy = vy + tt.pi[mi].x * n_bx;
      if (lz >= 0) vy = vy + tt.pk[tt.kcol + m].y * n_bz;
    }
    if (uz) {
      if (lx >= 0) vz = vz + tt.pi[mi].y * n_cx;
      if (ly >= 0) vz = vz + tt.pj[tt.jcol + m].y * n_cy;
    }
    // TFSF z faces (wave-uniform zt: the wave holds a z-face column and the
    // plane lies in [li, ri]); Ey -/+ and Ex +/- at lk / rk
    if (zt && (k == zf.lk || k == zf.rk)) {
      const bool up = k == zf.rk;
      if (j >= zf.lj && j < zf.rj && i > zf.li && i < zf.ri) {
        const double v = vy;
        vy = (T)(up ? v + p.z0 : v - p.z0);
      }
      if (i >= zf.li && i < zf.ri && j > zf.lj && j < zf.rj) {
        const double v = vx;
        vx = (T)(up ? v - p.z1 : v + p.z1);
      }
    }
  }
  const unsigned o = cell * (unsigned)sizeof(T);
  sto(q.u0, o, vx);
  sto(q.u1, o, vy);
  sto(q.u2, o, vz);
  if constexpr (SLAB && SPECIAL) {
    if (g.snap && in_design(g, i, j, k)) {
      T *h = g.snap + (long long)design_cell(g, i, j, k) * 3;
      h[0] = vx;
      h[1] = vy;
      h[2] = vz;
    }
  }
}

// SPECIAL: the engine has a design box or Drude cells, whose planes take
// the general path; without them that code is not compiled in.
// zf: the TFSF z faces this kernel applies (zf.terms null: none).
template <typename T, bool LDS, bool SPECIAL>
__device__ __forceinline__ void update_e_body(const Grid<T> &g, const Tiling &tl, const ZFaces &zf, int block) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int bz, by, bx;
  if (!tile_of_block(tl, block, bz, by, bx)) return;
  bz += tl.oz;
  by += tl.oy;
  bx += tl.ox;
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  const bool live = k <= nz && j <= ny;
  // the tile's field and material bases start at its first plane (i0 - 1
  // for the H(i - 1) loads), so 32-bit byte offsets reach tx + 1 planes of
  // any grid
  const long long pb = (long long)(i0 >= 1 ? i0 - 1 : 0) * g.si;
  Col<T> q;
  q.u0 = g.f[EX] + pb; q.u1 = g.f[EY] + pb; q.u2 = g.f[EZ] + pb;
  q.r0 = g.f[HX] + pb; q.r1 = g.f[HY] + pb; q.r2 = g.f[HZ] + pb;
  q.mat = g.mat + pb;
  q.dk = k >= 1 ? (unsigned)sizeof(T) : 0u;
  q.dj = j >= 1 ? (unsigned)(g.sj * (long long)sizeof(T)) : 0u;
  q.kin = k >= 1 && k < nz;
  q.jin = j >= 1 && j < ny;
  q.j = j;
  q.k = k;
  q.cp = col_cpml(g.be, g.ae, live ? e_layer(j, ny, g.dn[1], g.dp[1]) : -1,
                  live ? e_layer(k, nz, g.dn[2], g.dp[2]) : -1, j, k);
  // a wave that holds a TFSF z-face column (k = lk or rk, j in [lj, rj])
  // applies those corrections on the planes [li, ri], on the slab path
  const bool zwave =
      zf.terms && __ballot(live && (k == zf.lk || k == zf.rk) && j >= zf.lj && j <= zf.rj) != 0ull;
  // a wave whose column meets no y or z slab runs the plain path on every
  // plane outside the x slabs
  const bool col_slab = zwave || __ballot(q.cp.ly >= 0 || q.cp.lz >= 0) != 0ull;
  // a wave whose column crosses the design box or the Drude box takes the
  // general path on those boxes' planes
  const bool col_design =
      SPECIAL && __ballot(live && j >= g.db[2] && j < g.db[3] && k >= g.db[4] && k < g.db[5]) != 0ull;
  const bool col_drude =
      SPECIAL && __ballot(live && j >= g.xb[2] && j < g.xb[3] && k >= g.xb[4] && k < g.xb[5]) != 0ull;
  unsigned c = (unsigned)((long long)i0 * g.si - pb + (long long)j * g.sj + k);
  // x-neighbours are carried in registers: Hz(i-1), Hy(i-1)
  T hz_prev = T(0), hy_prev = T(0);
  Plane<T> A;
  bool slab = col_slab || e_layer(i0, nx, g.dn[0], g.dp[0]) >= 0 ||
              (col_design && i0 >= g.db[0] && i0 < g.db[1]) || (col_drude && i0 >= g.xb[0] && i0 < g.xb[1]);
  // the first plane's loads go out before the tables are staged
  if (live) {
    if (i0 >= 1) {
      hz_prev = ldo(q.r2, (c - (unsigned)g.si) * (unsigned)sizeof(T));
      hy_prev = ldo(q.r1, (c - (unsigned)g.si) * (unsigned)sizeof(T));
    }
    if (slab)
      load_plane_e<T, true>(g, q, c, i0, A, zf, zwave && i0 >= zf.li && i0 <= zf.ri);
    else
      load_plane_e<T, false>(g, q, c, i0, A, zf, false);
  }
  TileTables<T> tt;
  if (LDS) {
    bool sk, sj, si;
    tile_slabs(g, tl, bz, cell, i0, i1, sk, sj, si);
    tt = stage_tables<T>(lds_raw, g, 0, tl, cell, i0, sk, sj, si);
  } else {
    tt = global_tables<T>(g, 0, cell);
  }
  if (!live) return;
  for (int i = i0;;) {
    if (slab)
      compute_plane_e<T, true, SPECIAL>(g, tt, q, A, i, c, hz_prev, hy_prev, zf, zwave && i >= zf.li && i <= zf.ri);
    else
      compute_plane_e<T, false, SPECIAL>(g, tt, q, A, i, c, hz_prev, hy_prev, zf, false);
    hz_prev = A.g2;
    hy_prev = A.g1;
    if (++i >= i1) break;
    c += (unsigned)g.si;
    slab = col_slab || e_layer(i, nx, g.dn[0], g.dp[0]) >= 0 || (col_design && i >= g.db[0] && i < g.db[1]) ||
           (col_drude && i >= g.xb[0] && i < g.xb[1]);
    if (slab)
      load_plane_e<T, true>(g, q, c, i, A, zf, zwave && i >= zf.li && i <= zf.ri);
    else
      load_plane_e<T, false>(g, q, c, i, A, zf, false);
  }
}

// ---------------------------------------------------------------------------
// H half-step: Simulation::updateH (simulation_update.cpp:84-100) fused with
// PML::updateH (perfect_match_layer.cpp:434-572).  Same plane scheme as the
// E kernel; the i+1 values Ez(i+1), Ey(i+1) loaded for one plane are the
// Ez(i), Ey(i) of the next.
//
// Plane fields: f = Hx, Hy, Hz; g0 = Ex; n0 = Ey(k+1), n1 = Ex(k+1),
// n2 = Ez(j+1), n3 = Ex(j+1), n4 = Ez(i+1), n5 = Ey(i+1).
template <typename T, bool SLAB>
__device__ __forceinline__ void load_plane_h(const Grid<T> &g, const Col<T> &q, unsigned cell, int i,
                                             unsigned di, Plane<T> &p, const ZFaces &zf, bool zplane) {
  const unsigned o = cell * (unsigned)sizeof(T);
  p.m = 0;
  if (!g.h_uni) p.m = ldm(q.mat, cell);
  p.f0 = ldo(q.u0, o);
  p.f1 = ldo(q.u1, o);
  p.f2 = ldo(q.u2, o);
  p.g0 = ldo(q.r0, o);
  p.n0 = ldo(q.r1, o + q.dk);
  p.n1 = ldo(q.r0, o + q.dk);
  p.n2 = ldo(q.r2, o + q.dj);
  p.n3 = ldo(q.r0, o + q.dj);
  p.n4 = ldo(q.r2, o + di);
  p.n5 = ldo(q.r1, o + di);
  if constexpr (SLAB) {
    p.lx = h_layer(i, g.nx, g.dn[0], g.dp[0]);
    p.bx = p.ax = T(0);
    for (int s = 0; s < 6; ++s) p.ps[s] = T(0);
    const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx, j = q.j, k = q.k;
    const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = g.ny + 1;
    if (lx >= 0) {
      p.bx = g.bm[0][i];
      p.ax = g.am[0][i];
    }
    // psi pairs, each one two-wide load: x (BX for Hy, CX for Hz), y (AY for
    // Hx, CY for Hz), z (AZ for Hx, BZ for Hy)
    const bool sx = q.jin && q.kin, sy = i < g.nx && q.kin, sz = i < g.nx && q.jin;
    if (lx >= 0 && (sy || sz)) {
      const P2<T> v = ld_pair(g.psi_h[P_BX], psi_x(lx, j, k, nm1y, g.pz));
      p.ps[P_BX] = v.x;
      p.ps[P_CX] = v.y;
    }
    if (ly >= 0 && (sx || sz)) {
      const P2<T> v = ld_pair(g.psi_h[P_AY], psi_y(i, ly, k, Ly, g.pz));
      p.ps[P_AY] = v.x;
      p.ps[P_CY] = v.y;
    }
    if (lz >= 0 && (sx || sy)) {
      const P2<T> v = ld_pair(g.psi_h[P_AZ], psi_z(i, j, lz, nm1y, Lz));
      p.ps[P_AZ] = v.x;
      p.ps[P_BZ] = v.y;
    }
    // TFSF z faces: Hy and Hx at lk - 1 and rk
    p.z0 = p.z1 = 0.0;
    if (zplane && (k == zf.lk - 1 || k == zf.rk)) {
      const int up = k == zf.rk;
      if (i < zf.ri && j >= zf.lj && j <= zf.rj) p.z0 = zf.terms[zterm_index(zf, 4 + up, i, j)];
      if (j >= zf.lj && j < zf.rj) p.z1 = zf.terms[zterm_index(zf, 6 + up, i, j)];
    }
  }
}

template <typename T, bool SLAB>
__device__ __forceinline__ bool compute_plane_h(const Grid<T> &g, const TileTables<T> &tt, const Col<T> &q,
                                                const Plane<T> &p, int i, unsigned cell, T ez_c, T ey_c,
                                                const ZFaces &zf, bool zt, const NfRec &nr, bool nfz) {
  const int nx = g.nx, j = q.j, k = q.k;
  const bool kin = q.kin, jin = q.jin;
  const int m = (int)p.m;
  const T hx = p.f0, hy = p.f1, hz = p.f2, ex = p.g0;
  const T dey_k = p.n0 - ey_c;  // Ey(k+1) - Ey(k)
  const T dez_j = p.n2 - ez_c;  // Ez(j+1) - Ez(j)
  const T dez_i = p.n4 - ez_c;  // Ez(i+1) - Ez(i)
  const T dex_k = p.n1 - ex;    // Ex(k+1) - Ex(k)
  const T dex_j = p.n3 - ex;    // Ex(j+1) - Ex(j)
  const T dey_i = p.n5 - ey_c;  // Ey(i+1) - Ey(i)
  const int mi = (i - tt.i0) * g.nmat + m;
  const P4<T> cs = tt.self[m];
  const P2<T> ck = tt.k[tt.kcol + m], cj = tt.j[tt.jcol + m], ci = tt.i[mi];
  // chxh*Hx + chxey*dEy/dz + chxez*dEz/dy, and likewise for Hy, Hz
  const T yx = cs.x * hx + ck.x * dey_k + cj.x * dez_j;
  const T yy = cs.y * hy + ci.x * dez_i + ck.y * dex_k;
  const T yz = cs.z * hz + cj.y * dex_j + ci.y * dey_i;
  // all three H components update over i < nx, j < ny, k < nz
  const bool yee = (i < nx) && jin && kin;
  T vx = yee ? yx : hx;
  T vy = yee ? yy : hy;
  T vz = yee ? yz : hz;
  if constexpr (SLAB) {
    const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx;
    const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = g.ny + 1;
    // CPML over Hx (nx+1, ny, nz): j < ny, k < nz; Hy (nx, ny+1, nz): i < nx,
    // k < nz; Hz (nx, ny, nz+1): i < nx, j < ny.  The new psi of each pair
    // (the old value where the component lies outside its range), stored
    // with one two-wide store; then the terms in the reference's order.
    const bool ux = jin && kin, uy = i < nx && kin, uz = i < nx && jin;
    T n_ay = p.ps[P_AY], n_az = p.ps[P_AZ], n_bx = p.ps[P_BX], n_bz = p.ps[P_BZ], n_cx = p.ps[P_CX],
      n_cy = p.ps[P_CY];
    if (lx >= 0 && (uy || uz)) {
      if (uy) n_bx = p.bx * p.ps[P_BX] + p.ax * dez_i;
      if (uz) n_cx = p.bx * p.ps[P_CX] + p.ax * dey_i;
      st_pair(g.psi_h[P_BX], psi_x(lx, j, k, nm1y, g.pz), n_bx, n_cx);
    }
    if (ly >= 0 && (ux || uz)) {
      if (ux) n_ay = q.cp.by * p.ps[P_AY] + q.cp.ay * dez_j;
      if (uz) n_cy = q.cp.by * p.ps[P_CY] + q.cp.ay * dex_j;
      st_pair(g.psi_h[P_AY], psi_y(i, ly, k, Ly, g.pz), n_ay, n_cy);
    }
    if (lz >= 0 && (ux || uy)) {
      if (ux) n_az = q.cp.bz * p.ps[P_AZ] + q.cp.az * dey_k;
      if (uy) n_bz = q.cp.bz * p.ps[P_BZ] + q.cp.az * dex_k;
      st_pair(g.psi_h[P_AZ], psi_z(i, j, lz, nm1y, Lz), n_az, n_bz);
    }
    if (ux) {
      if (ly >= 0) vx = vx + tt.pj[tt.jcol + m].x * n_ay;
      if (lz >= 0) vx = vx + tt.pk[tt.kcol + m].x * n_az;
    }
    if (uy) {
      if (lx >= 0) vy = vy + tt.pi[mi].x * n_bx;
      if (lz >= 0) vy = vy + tt.pk[tt.kcol + m].y * n_bz;
    }
    if (uz) {
      if (lx >= 0) vz = vz + tt.pi[mi].y * n_cx;
      if (ly >= 0) vz = vz + tt.pj[tt.jcol + m].y * n_cy;
    }
    // TFSF z faces: Hy +/- and Hx -/+ at lk - 1 / rk
    if (zt && (k == zf.lk - 1 || k == zf.rk)) {
      const bool up = k == zf.rk;
      if (i >= zf.li && i < zf.ri && j >= zf.lj && j <= zf.rj) {
        const double v = vy;
        vy = (T)(up ? (v - p.z0) : (v + p.z0));
      }
      if (i >= zf.li && i <= zf.ri && j >= zf.lj && j < zf.rj) {
        const double v = vx;
        vx = (T)(up ? (v + p.z1) : (v - p.z1));
      }
    }
  }
  const unsigned o = cell * (unsigned)sizeof(T);
  sto(q.u0, o, vx);
  sto(q.u1, o, vy);
  sto(q.u2, o, vz);
  if (nfz) nf_record<T>(nr, i, j, k, ey_c, ex, p.n1, vx, vy);
  // non-finite output: the run diverged
  return !(__builtin_isfinite(vx) && __builtin_isfinite(vy) && __builtin_isfinite(vz));
}

template <typename T, bool LDS>
__device__ __forceinline__ void update_h_body(const Grid<T> &g, const Tiling &tl, const ZFaces &zf, const NfRec &nr,
                                              int block, long long step) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int bz, by, bx;
  if (!tile_of_block(tl, block, bz, by, bx)) return;
  bz += tl.oz;
  by += tl.oy;
  bx += tl.ox;
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  const bool live = k <= nz && j <= ny;
  const unsigned bsi = (unsigned)(g.si * (long long)sizeof(T));
  // bases at the tile's first plane: offsets reach its tx planes and the
  // E(i + 1) plane after them
  const long long pb = (long long)i0 * g.si;
  Col<T> q;
  q.u0 = g.f[HX] + pb; q.u1 = g.f[HY] + pb; q.u2 = g.f[HZ] + pb;
  q.r0 = g.f[EX] + pb; q.r1 = g.f[EY] + pb; q.r2 = g.f[EZ] + pb;
  q.mat = g.mat + pb;
  q.dk = k < nz ? (unsigned)sizeof(T) : 0u;
  q.dj = j < ny ? (unsigned)(g.sj * (long long)sizeof(T)) : 0u;
  q.kin = k < nz;
  q.jin = j < ny;
  q.j = j;
  q.k = k;
  q.cp = col_cpml(g.bm, g.am, live ? h_layer(j, ny, g.dn[1], g.dp[1]) : -1,
                  live ? h_layer(k, nz, g.dn[2], g.dp[2]) : -1, j, k);
  // TFSF z-face columns (k = lk - 1 or rk, j in [lj, rj]), on the slab path
  const bool zwave =
      zf.terms && __ballot(live && (k == zf.lk - 1 || k == zf.rk) && j >= zf.lj && j <= zf.rj) != 0ull;
  const bool col_slab = zwave || __ballot(q.cp.ly >= 0 || q.cp.lz >= 0) != 0ull;
  // a wave that holds cells of the NF-FF z-face records
  const bool nfwave = nr.rec && __ballot(live && (k == nr.k0 - 1 || k == nr.k0 || k == nr.k1 - 1 || k == nr.k1) &&
                                         j >= nr.lj && j < nr.lj + nr.nyr) != 0ull;
  unsigned c = (unsigned)((long long)j * g.sj + k);
  T ez_c = T(0), ey_c = T(0);
  Plane<T> A;
  bool slab = col_slab || h_layer(i0, nx, g.dn[0], g.dp[0]) >= 0;
  // the first plane's loads go out before the tables are staged
  if (live) {
    ez_c = ldo(q.r2, c * (unsigned)sizeof(T));
    ey_c = ldo(q.r1, c * (unsigned)sizeof(T));
    const unsigned di = i0 < nx ? bsi : 0u;
    if (slab)
      load_plane_h<T, true>(g, q, c, i0, di, A, zf, zwave && i0 >= zf.li && i0 <= zf.ri);
    else
      load_plane_h<T, false>(g, q, c, i0, di, A, zf, false);
  }
  TileTables<T> tt;
  if (LDS) {
    bool sk, sj, si;
    tile_slabs(g, tl, bz, cell, i0, i1, sk, sj, si);
    tt = stage_tables<T>(lds_raw, g, 1, tl, cell, i0, sk, sj, si);
  } else {
    tt = global_tables<T>(g, 1, cell);
  }
  if (!live) return;
  bool bad = false;
  for (int i = i0;;) {
    const bool nfz = nfwave && i >= nr.li && i < nr.li + nr.nxr;
    if (slab)
      bad |= compute_plane_h<T, true>(g, tt, q, A, i, c, ez_c, ey_c, zf, zwave && i >= zf.li && i <= zf.ri, nr,
                                      nfz);
    else
      bad |= compute_plane_h<T, false>(g, tt, q, A, i, c, ez_c, ey_c, zf, false, nr, nfz);
    ez_c = A.n4;
    ey_c = A.n5;
    if (++i >= i1) break;
    c += (unsigned)g.si;
    slab = col_slab || h_layer(i, nx, g.dn[0], g.dp[0]) >= 0;
    const unsigned di = i < nx ? bsi : 0u;
    if (slab)
      load_plane_h<T, true>(g, q, c, i, di, A, zf, zwave && i >= zf.li && i <= zf.ri);
    else
      load_plane_h<T, false>(g, q, c, i, di, A, zf, false);
  }
  if (bad) atomicMin(g.div, (unsigned long long)step);
}

// The FP32 E kernel without design or Drude code schedules about 1.5% faster
// with a waves-per-EU hint at the occupancy it reaches anyway (7); the same
// hint made the H kernel slower and left FP64 unchanged.
template <typename T, bool LDS, bool SPECIAL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !SPECIAL ? 7 : 1)))
k_update_e(Grid<T> g, Tiling tl, ZFaces zf) {
  update_e_body<T, LDS, SPECIAL>(g, tl, zf, blockIdx.x);
}
template <typename T, bool LDS>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, Tiling tl, ZFaces zf, NfRec nr, long long step) {
  update_h_body<T, LDS>(g, tl, zf, nr, blockIdx.x, step);
}

// TFSF3D::updateIncidentField (tfsf_3d.cpp:153-183): one workgroup
__device__ __forceinline__ void tfsf_ifa_body(const Tfsf &t, long long step) {
  const int L = t.L;
  double *e = t.e_inc, *h = t.h_inc;
  const double x = e[L - 2], y = e[L - 1];
  __syncthreads();
  if (threadIdx.x == 0) e[0] = t.src[step];
  for (int i = 1 + threadIdx.x; i < L - 1; i += blockDim.x)
    e[i] = t.ceie * e[i] + t.ceihi * (h[i] - h[i - 1]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double a = t.ab[0], b = t.ab[1];
    e[L - 1] = -a + t.c1 * (e[L - 2] + b) + t.c2 * (x + y);
    t.ab[0] = x;
    t.ab[1] = y;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < L - 1; i += blockDim.x)
    h[i] = t.chih * h[i] + t.chiei * (e[i + 1] - e[i]);
}
__global__ void __launch_bounds__(1024) k_tfsf_ifa(Tfsf t, long long step) {
  tfsf_ifa_body(t, step);
}

// TFSF3D::updateE (tfsf_3d.cpp:299-411).  The face cells come in six
// segments, each enumerated with k fastest where k varies, so that a wave's
// cells share cache lines: Ey on the x faces (i = li, ri; j in [lj, rj); k in
// [lk, rk]), Ey on the z faces away from the x faces, Ez on the x faces (j in
// [lj, rj], k in [lk, rk)), Ez on the y faces away from the x faces, Ex on the
// y faces (i in [li, ri), k in [lk, rk]), Ex on the z faces away from the y
// faces.  The two z-face segments are empty when the curl kernels apply them
// (ZFaces).  Edge cells take two corrections; one thread owns each target cell
// and applies them in the reference's xn, xp, yn, yp, zn, zp order.
struct TfsfESeg {
  int off[7];  // prefix offsets of the six segments
};
template <typename T>
__device__ __forceinline__ void tfsf_e_body(const Grid<T> &g, const Tfsf &t, const TfsfESeg &s, int block) {
  const int q = block * (int)blockDim.x + (int)threadIdx.x;
  if (q >= s.off[6]) return;
  int seg = 0;
  while (q >= s.off[seg + 1]) ++seg;
  const int r = q - s.off[seg];
  const int li = t.li, ri = t.ri, lj = t.lj, rj = t.rj, lk = t.lk, rk = t.rk;
  const int ni = ri - li, nj = rj - lj, nk = rk - lk;
  int i, j, k, comp;
  switch (seg) {
    case 0: { const int w = r / (nk + 1); k = lk + r % (nk + 1); j = lj + w % nj; i = w / nj ? ri : li; comp = EY; break; }
    case 1: { const int w = r / nj; j = lj + r % nj; i = li + 1 + w % (ni - 1); k = w / (ni - 1) ? rk : lk; comp = EY; break; }
    case 2: { const int w = r / nk; k = lk + r % nk; j = lj + w % (nj + 1); i = w / (nj + 1) ? ri : li; comp = EZ; break; }
    case 3: { const int w = r / nk; k = lk + r % nk; i = li + 1 + w % (ni - 1); j = w / (ni - 1) ? rj : lj; comp = EZ; break; }
    case 4: { const int w = r / (nk + 1); k = lk + r % (nk + 1); i = li + w % ni; j = w / ni ? rj : lj; comp = EX; break; }
    default: { const int w = r / (nj - 1); j = lj + 1 + r % (nj - 1); i = li + w % ni; k = w / ni ? rk : lk; comp = EX; break; }
  }
  T *p = &g.f[comp][(long long)i * g.si + (long long)j * g.sj + k];
  double val = *p;
  if (comp == EY) {
    if (i == li) val = val + t.ca[0] * tfsf_inc(t, HZ, li - 1, j, k);
    if (i == ri) val = val - t.ca[0] * tfsf_inc(t, HZ, ri, j, k);
    if (k == lk) val = val - t.ca[2] * tfsf_inc(t, HX, i, j, lk - 1);
    if (k == rk) val = val + t.ca[2] * tfsf_inc(t, HX, i, j, rk);
  } else if (comp == EZ) {
    if (i == li) val = val - t.ca[0] * tfsf_inc(t, HY, li - 1, j, k);
    if (i == ri) val = val + t.ca[0] * tfsf_inc(t, HY, ri, j, k);
    if (j == lj) val = val + t.ca[1] * tfsf_inc(t, HX, i, lj - 1, k);
    if (j == rj) val = val - t.ca[1] * tfsf_inc(t, HX, i, rj, k);
  } else {
    if (j == lj) val = val - t.ca[1] * tfsf_inc(t, HZ, i, lj - 1, k);
    if (j == rj) val = val + t.ca[1] * tfsf_inc(t, HZ, i, rj, k);
    if (k == lk) val = val + t.ca[2] * tfsf_inc(t, HY, i, j, lk - 1);
    if (k == rk) val = val - t.ca[2] * tfsf_inc(t, HY, i, j, rk);
  }
  *p = (T)val;
}
template <typename T>
__global__ void k_tfsf_e(Grid<T> g, Tfsf t, TfsfESeg s) {
  tfsf_e_body<T>(g, t, s, blockIdx.x);
}

// TFSF3D::updateH (tfsf_3d.cpp:185-297).  No H cell takes two corrections.
// Twelve segments (x faces, y faces, z faces; k fastest where k varies); the
// four z-face segments are empty when the curl kernels apply them (ZFaces).
struct TfsfHSeg {
  int off[13];  // prefix offsets of the 12 face loops
};
template <typename T>
__device__ __forceinline__ void tfsf_h_body(const Grid<T> &g, const Tfsf &t, const TfsfHSeg &s,
                                            int block) {
  const int q = block * (int)blockDim.x + (int)threadIdx.x;
  if (q >= s.off[12]) return;
  int seg = 0;
  while (q >= s.off[seg + 1]) ++seg;
  const int r = q - s.off[seg];
  const int li = t.li, ri = t.ri, lj = t.lj, rj = t.rj, lk = t.lk, rk = t.rk;
  int i = 0, j = 0, k = 0, comp = HX, src = EX;
  double coef = 0;
  bool add = true;
  switch (seg) {
    case 0: { const int nk = rk + 1 - lk; j = lj + r / nk; k = lk + r % nk; i = li - 1;
              comp = HZ; src = EY; coef = t.cb[0]; add = true; break; }   // xn hz += cbx*eyi(li)
    case 1: { const int nk = rk - lk; j = lj + r / nk; k = lk + r % nk; i = li - 1;
              comp = HY; src = EZ; coef = t.cb[0]; add = false; break; }  // xn hy -= cbx*ezi(li)
    case 2: { const int nk = rk + 1 - lk; j = lj + r / nk; k = lk + r % nk; i = ri;
              comp = HZ; src = EY; coef = t.cb[0]; add = false; break; }
    case 3: { const int nk = rk - lk; j = lj + r / nk; k = lk + r % nk; i = ri;
              comp = HY; src = EZ; coef = t.cb[0]; add = true; break; }
    case 4: { const int nk = rk - lk; i = li + r / nk; k = lk + r % nk; j = lj - 1;
              comp = HX; src = EZ; coef = t.cb[1]; add = true; break; }   // yn hx += cby*ezi(lj)
    case 5: { const int nk = rk + 1 - lk; i = li + r / nk; k = lk + r % nk; j = lj - 1;
              comp = HZ; src = EX; coef = t.cb[1]; add = false; break; }
    case 6: { const int nk = rk - lk; i = li + r / nk; k = lk + r % nk; j = rj;
              comp = HX; src = EZ; coef = t.cb[1]; add = false; break; }
    case 7: { const int nk = rk + 1 - lk; i = li + r / nk; k = lk + r % nk; j = rj;
              comp = HZ; src = EX; coef = t.cb[1]; add = true; break; }
    case 8: { const int nj = rj + 1 - lj; i = li + r / nj; j = lj + r % nj; k = lk - 1;
              comp = HY; src = EX; coef = t.cb[2]; add = true; break; }   // zn hy += cbz*exi(lk)
    case 9: { const int nj = rj - lj; i = li + r / nj; j = lj + r % nj; k = lk - 1;
              comp = HX; src = EY; coef = t.cb[2]; add = false; break; }
    case 10: { const int nj = rj + 1 - lj; i = li + r / nj; j = lj + r % nj; k = rk;
               comp = HY; src = EX; coef = t.cb[2]; add = false; break; }
    default: { const int nj = rj - lj; i = li + r / nj; j = lj + r % nj; k = rk;
               comp = HX; src = EY; coef = t.cb[2]; add = true; break; }
  }
  // the incident E is sampled on the total-field side: li / lj / lk for the
  // N faces (the H cell sits at index l-1), ri / rj / rk for the P faces
  const int si_ = (seg == 0 || seg == 1) ? li : i;
  const int sj_ = (seg == 4 || seg == 5) ? lj : j;
  const int sk_ = (seg == 8 || seg == 9) ? lk : k;
  const double inc = tfsf_inc(t, src, si_, sj_, sk_);
  T *p = &g.f[comp][(long long)i * g.si + (long long)j * g.sj + k];
  const double v = *p;
  *p = (T)(add ? (v + coef * inc) : (v - coef * inc));
}
template <typename T>
__global__ void k_tfsf_h(Grid<T> g, Tfsf t, TfsfHSeg s) {
  tfsf_h_body<T>(g, t, s, blockIdx.x);
}

// ---------------------------------------------------------------------------
// CurrentSource::correctE (lumped_element/current_source.cpp:140-163)
template <typename T>
__global__ void k_source(T *f, long long si, long long sj, int i0, int nj, int nk,
                         int j0, int k0, int ncell, const T *coff, const T *wave,
                         long long step) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncell) return;
  const int i = i0 + q / (nj * nk), j = j0 + (q / nk) % nj, k = k0 + q % nk;
  T *p = &f[(long long)i * si + (long long)j * sj + k];
  *p = *p + coff[q] * wave[step];
}

// Inductor::correctE (lumped_element/inductor.cpp:115-146): E += cecjc J,
// then J = cjcec E + J with the corrected E
template <typename T>
__global__ void k_inductor(T *f, long long si, long long sj, int i0, int nj, int nk, int j0, int k0, int ncell,
                           const T *cecjc, const T *cjcec, T *jc) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncell) return;
  const int i = i0 + q / (nj * nk), j = j0 + (q / nk) % nj, k = k0 + q % nk;
  T *p = &f[(long long)i * si + (long long)j * sj + k];
  const T e = *p + cecjc[q] * jc[q];
  *p = e;
  jc[q] = cjcec[q] * e + jc[q];
}

// VoltageMonitor / CurrentMonitor::update (monitor/voltage_monitor.cpp:63-93,
// monitor/current_monitor.cpp:66-122): one thread sums the terms in the
// reference's loop order, in FP64
struct LineTerm {
  long long ca, cb;  // cells (cb < 0: none)
  double w;
  int comp, group;
};
template <typename T>
__global__ void k_line_monitor(Grid<T> g, const LineTerm *terms, int n, double sign, double *buf, long long step) {
  if (threadIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0;
  for (int q = 0; q < n; ++q) {
    const LineTerm t = terms[q];
    const T *f = g.f[t.comp];
    double x = (double)f[t.ca];
    if (t.cb >= 0) x = x - (double)f[t.cb];
    const double v = x * t.w;
    if (t.group == 0)
      s0 = s0 + v;
    else
      s1 = s1 + v;
  }
  buf[step] = sign * (s0 + s1);
}

// point probes (monitor update slot, simulation_update.cpp:215-219)
template <typename T>
__global__ void k_probe(Grid<T> g, const int *comp, const long long *cell, int n,
                        double *buf, long long T_total, long long step) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  buf[(long long)q * T_total + step] = (double)g.f[comp[q]][cell[q]];
}

// a box of one component gathered in C order as FP64 (field monitors,
// field_monitor.cpp:99-121)
template <typename T>
__global__ void k_box_gather(const T *f, long long si, long long sj, int i0, int j0, int k0, int nj, int nk,
                             long long n, double *out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long long i = i0 + q / ((long long)nj * nk), r = q % ((long long)nj * nk);
  const long long j = j0 + r / nk, k = k0 + r % nk;
  out[q] = (double)f[i * si + j * sj + k];
}

// ---------------------------------------------------------------------------
// NffftFd::update (nffft_fd.cpp:108-203) with the face-centre stencils of
// linear_interpolation_emf.cpp:286-388.  One thread per face cell of a face
// pair; FP64 accumulators in every precision.  Device order of the surface
// cells: x faces (j, k), y faces (i, k), z faces (i, j), the last index
// fastest, so that the threads of a wave read consecutive k wherever k varies
// (the reference, and get_nffft, order the y faces (k, i); the host permutes
// them on the way out).
struct Nffft {
  int li, ri, lj, rj, lk, rk, nf;
  int nxb, nyb, nzb;
  int zf;                  // 1: the z faces are kept (0 on a 2-D grid)
  long long T_total;
  const double2 *tj, *tm;  // [nf][T]
  double2 *acc;
  long long base[3];       // accumulator offset of each face pair
  // z-face records left by the H kernel (NfRec) or null
  const void *zrec;
};

// surface cells: x faces, y faces and (when kept) z faces
__host__ __device__ inline long long nf_cells(const Nffft &n) {
  return (long long)n.nyb * n.nzb + (long long)n.nzb * n.nxb + (n.zf ? (long long)n.nxb * n.nyb : 0LL);
}

// the eight face-centre values of surface cell q (faces x, then y, then z;
// both faces of a pair): false past the last cell
template <typename T>
__device__ __forceinline__ bool nffft_values(const Grid<T> &g, const Nffft &n, int q, T *v, int &pair, int &cells,
                                             int &cell) {
  const int cx = n.nyb * n.nzb, cy = n.nzb * n.nxb, cz = n.zf ? n.nxb * n.nyb : 0;
  const T *Ex = g.f[EX], *Ey = g.f[EY], *Ez = g.f[EZ];
  const T *Hx = g.f[HX], *Hy = g.f[HY], *Hz = g.f[HZ];
  const long long si = g.si, sj = g.sj;
#define C3(i, j, k) ((long long)(i) * si + (long long)(j) * sj + (k))
  if (q < cx) {
    pair = 0; cells = cx; cell = q;
    const int j = n.lj + q / n.nzb, k = n.lk + q % n.nzb;
    const int li = n.li, ri = n.ri;
    v[0] = T(-1) * (T(0.5) * (Ez[C3(li, j, k)] + Ez[C3(li, j + 1, k)]));
    v[1] = T(0.5) * (Ey[C3(li, j, k)] + Ey[C3(li, j, k + 1)]);
    v[2] = T(0.25) * (Hz[C3(li, j, k)] + Hz[C3(li, j, k + 1)] + Hz[C3(li - 1, j, k)] + Hz[C3(li - 1, j, k + 1)]);
    v[3] = T(-1) * (T(0.25) * (Hy[C3(li, j, k)] + Hy[C3(li, j + 1, k)] + Hy[C3(li - 1, j, k)] + Hy[C3(li - 1, j + 1, k)]));
    v[4] = T(0.5) * (Ez[C3(ri, j, k)] + Ez[C3(ri, j + 1, k)]);
    v[5] = T(-1) * (T(0.5) * (Ey[C3(ri, j, k)] + Ey[C3(ri, j, k + 1)]));
    v[6] = T(-1) * (T(0.25) * (Hz[C3(ri, j, k)] + Hz[C3(ri, j, k + 1)] + Hz[C3(ri - 1, j, k)] + Hz[C3(ri - 1, j, k + 1)]));
    v[7] = T(0.25) * (Hy[C3(ri, j, k)] + Hy[C3(ri, j + 1, k)] + Hy[C3(ri - 1, j, k)] + Hy[C3(ri - 1, j + 1, k)]);
  } else if ((q -= cx) < cy) {
    pair = 1; cells = cy; cell = q;
    const int i = n.li + q / n.nzb, k = n.lk + q % n.nzb;
    const int lj = n.lj, rj = n.rj;
    v[0] = T(-1) * (T(0.5) * (Ex[C3(i, lj, k)] + Ex[C3(i, lj + 1, k)]));
    v[1] = T(0.5) * (Ez[C3(i, lj, k)] + Ez[C3(i + 1, lj, k)]);
    v[2] = T(0.25) * (Hx[C3(i + 1, lj, k)] + Hx[C3(i, lj, k)] + Hx[C3(i + 1, lj - 1, k)] + Hx[C3(i, lj - 1, k)]);
    v[3] = T(-1) * (T(0.25) * (Hz[C3(i, lj, k)] + Hz[C3(i, lj, k + 1)] + Hz[C3(i, lj - 1, k)] + Hz[C3(i, lj - 1, k + 1)]));
    v[4] = T(0.5) * (Ex[C3(i, rj, k)] + Ex[C3(i, rj + 1, k)]);
    v[5] = T(-1) * (T(0.5) * (Ez[C3(i, rj, k)] + Ez[C3(i + 1, rj, k)]));
    v[6] = T(-1) * (T(0.25) * (Hx[C3(i + 1, rj, k)] + Hx[C3(i, rj, k)] + Hx[C3(i + 1, rj - 1, k)] + Hx[C3(i, rj - 1, k)]));
    v[7] = T(0.25) * (Hz[C3(i, rj, k)] + Hz[C3(i, rj, k + 1)] + Hz[C3(i, rj - 1, k)] + Hz[C3(i, rj - 1, k + 1)]);
  } else if ((q -= cy) < cz && n.zrec) {
    // the same sums over the records' copies of the fields
    pair = 2; cells = cz; cell = q;
    // record: [0] Ey, [1] Ex, [2] Ex(k+1), [3] Hy, [4] Hx at lk (rk); [6] Hy,
    // [7] Hx at lk - 1 (rk - 1)
    const int a = q / n.nyb, b = q % n.nyb, nyr = n.nyb + 1;
    const T *lo = (const T *)n.zrec + ((long long)a * nyr + b) * 8;
    const T *hi = lo + (long long)(n.nxb + 1) * nyr * 8;
    const long long di = (long long)nyr * 8, dj = 8;
    v[0] = T(-1) * (T(0.5) * (lo[0] + lo[di]));
    v[1] = T(0.5) * (lo[1] + lo[2]);
    v[2] = T(0.25) * (lo[3] + lo[dj + 3] + lo[6] + lo[dj + 6]);
    v[3] = T(-1) * (T(0.25) * (lo[di + 4] + lo[4] + lo[di + 7] + lo[7]));
    v[4] = T(0.5) * (hi[0] + hi[di]);
    v[5] = T(-1) * (T(0.5) * (hi[1] + hi[2]));
    v[6] = T(-1) * (T(0.25) * (hi[3] + hi[dj + 3] + hi[6] + hi[dj + 6]));
    v[7] = T(0.25) * (hi[di + 4] + hi[4] + hi[di + 7] + hi[7]);
  } else if (q < cz) {
    pair = 2; cells = cz; cell = q;
    const int i = n.li + q / n.nyb, j = n.lj + q % n.nyb;
    const int lk = n.lk, rk = n.rk;
    v[0] = T(-1) * (T(0.5) * (Ey[C3(i, j, lk)] + Ey[C3(i + 1, j, lk)]));
    v[1] = T(0.5) * (Ex[C3(i, j, lk)] + Ex[C3(i, j, lk + 1)]);
    v[2] = T(0.25) * (Hy[C3(i, j, lk)] + Hy[C3(i, j + 1, lk)] + Hy[C3(i, j, lk - 1)] + Hy[C3(i, j + 1, lk - 1)]);
    v[3] = T(-1) * (T(0.25) * (Hx[C3(i + 1, j, lk)] + Hx[C3(i, j, lk)] + Hx[C3(i + 1, j, lk - 1)] + Hx[C3(i, j, lk - 1)]));
    v[4] = T(0.5) * (Ey[C3(i, j, rk)] + Ey[C3(i + 1, j, rk)]);
    v[5] = T(-1) * (T(0.5) * (Ex[C3(i, j, rk)] + Ex[C3(i, j, rk + 1)]));
    v[6] = T(-1) * (T(0.25) * (Hy[C3(i, j, rk)] + Hy[C3(i, j + 1, rk)] + Hy[C3(i, j, rk - 1)] + Hy[C3(i, j + 1, rk - 1)]));
    v[7] = T(0.25) * (Hx[C3(i + 1, j, rk)] + Hx[C3(i, j, rk)] + Hx[C3(i + 1, j, rk - 1)] + Hx[C3(i, j, rk - 1)]);
  } else {
    return false;
  }
#undef C3
  return true;
}

// The running DFT of NffftFd::update (nffft_fd.cpp:108-203) over a window
// of steps: each step stores its face values in ring slot s ([slot][8][cells
// of all faces]), and k_nffft_flush adds the window's terms to every
// accumulator in step order -- the same operations as accumulating every
// step, with the accumulator read and written once per window.
constexpr int NF_RING = 16;

template <typename T>
__device__ __forceinline__ void nffft_sample_body(const Grid<T> &g, const Nffft &n, T *ring, int slot, int block) {
  const int q = block * (int)blockDim.x + (int)threadIdx.x;
  const long long Q = nf_cells(n);
  T v[8];
  int pair, cells, cell;
  if (!nffft_values<T>(g, n, q, v, pair, cells, cell)) return;
  for (int c = 0; c < 8; ++c) ring[((long long)slot * 8 + c) * Q + q] = v[c];
}
template <typename T>
__global__ void k_nffft_sample(Grid<T> g, Nffft n, T *ring, int slot) {
  nffft_sample_body<T>(g, n, ring, slot, blockIdx.x);
}

// one thread per (value c, surface cell): its 8 x nf accumulators' share
template <typename T>
__device__ __forceinline__ void nffft_flush_body(const Nffft &n, const T *__restrict__ ring, int nslots,
                                                 long long step0, int block) {
  const int cx = n.nyb * n.nzb, cy = n.nzb * n.nxb, cz = n.zf ? n.nxb * n.nyb : 0;
  const long long Q = (long long)cx + cy + cz;
  const long long t = (long long)block * blockDim.x + threadIdx.x;
  if (t >= 8 * Q) return;
  const int c = (int)(t / Q);
  int q = (int)(t - (long long)c * Q);
  int pair, cells, cell;
  if (q < cx) { pair = 0; cells = cx; cell = q; }
  else if ((q -= cx) < cy) { pair = 1; cells = cy; cell = q; }
  else { q -= cy; pair = 2; cells = cz; cell = q; }
  T v[NF_RING];
#pragma unroll
  for (int s = 0; s < NF_RING; ++s)
    if (s < nslots) v[s] = ring[((long long)s * 8 + c) * Q + (t - (long long)c * Q)];
  const double2 *tw = (c % 4) < 2 ? n.tm : n.tj;
  for (int f = 0; f < n.nf; ++f) {
    double2 *a = &n.acc[n.base[pair] + ((long long)c * n.nf + f) * cells + cell];
    double2 x = *a;
    const double2 *w = tw + (long long)f * n.T_total + step0;
#pragma unroll
    for (int s = 0; s < NF_RING; ++s)
      if (s < nslots) {
        const double vs = (double)v[s];
        const double2 ws = w[s];
        x.x = x.x + vs * ws.x;
        x.y = x.y + vs * ws.y;
      }
    *a = x;
  }
}
template <typename T>
__global__ void __launch_bounds__(256) k_nffft_flush(Nffft n, const T *__restrict__ ring, int nslots,
                                                    long long step0) {
  nffft_flush_body<T>(n, ring, nslots, step0, blockIdx.x);
}

// ---------------------------------------------------------------------------
// NffftFd::calculateFarfieldX/Y/Z (nffft_fd.cpp:423-556): one workgroup per
// direction (f, theta, phi), FP64 partial sums reduced with wavefront
// shuffles and LDS.
struct FarDir {
  const double *stcp, *stsp, *ct;  // per (theta,phi)
  const double *dk;                // dl * k per frequency
  int nt, np;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}

__global__ void __launch_bounds__(256) k_farfield(Nffft n, FarDir d, double2 *out) {
  const int dir = blockIdx.x;  // (t, p)
  const int f = blockIdx.y;
  const double stcp = d.stcp[dir], stsp = d.stsp[dir], ct = d.ct[dir];
  const double dk = d.dk[f];
  double2 s[12];
  for (int q = 0; q < 12; ++q) s[q] = make_double2(0.0, 0.0);
  const int cells[3] = {n.nyb * n.nzb, n.nzb * n.nxb, n.nxb * n.nyb};
  for (int pair = 0; pair < 3; ++pair) {
    const int cc = cells[pair];
    const double2 *A = n.acc + n.base[pair];
    for (int q = threadIdx.x; q < cc; q += blockDim.x) {
      double dn, dp;
      if (pair == 0) {
        const int j = n.lj + q / n.nzb, k = n.lk + q % n.nzb;
        const double r = ((double)j + 0.5) * stsp, z = ((double)k + 0.5) * ct;
        dn = dk * ((double)n.li * stcp + r + z);
        dp = dk * ((double)n.ri * stcp + r + z);
      } else if (pair == 1) {
        const int i = n.li + q / n.nzb, k = n.lk + q % n.nzb;
        const double x = ((double)i + 0.5) * stcp, z = ((double)k + 0.5) * ct;
        dn = dk * (x + (double)n.lj * stsp + z);
        dp = dk * (x + (double)n.rj * stsp + z);
      } else {
        const int i = n.li + q / n.nyb, j = n.lj + q % n.nyb;
        const double x = ((double)i + 0.5) * stcp, y = ((double)j + 0.5) * stsp;
        dn = dk * (x + y + (double)n.lk * ct);
        dp = dk * (x + y + (double)n.rk * ct);
      }
      double sn, cn, sp, cp;
      sincos(dn, &sn, &cn);
      sincos(dp, &sp, &cp);
      const double2 pn = make_double2(cn, sn), pp = make_double2(cp, sp);
      // currents (c%4): pair x: my,mz,jy,jz; pair y: mz,mx,jz,jx; pair z: mx,my,jx,jy.
      // sums: [J_a, J_b, M_a, M_b] = c 2,3,0,1 with N face c, P face c+4
      const long long stride = (long long)n.nf * cc;
      const long long o = (long long)f * cc + q;
      const int order[4] = {2, 3, 0, 1};
      for (int u = 0; u < 4; ++u) {
        const int c = order[u];
        const double2 an = A[c * stride + o], ap = A[(c + 4) * stride + o];
        s[pair * 4 + u] = cadd(s[pair * 4 + u], cadd(cmul(an, pn), cmul(ap, pp)));
      }
    }
  }
  // block reduction
  __shared__ double red[256 / 64][24];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 12; ++q) {
    double re = s[q].x, im = s[q].y;
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_xor(re, off, 64);
      im += __shfl_xor(im, off, 64);
    }
    if (lane == 0) { red[wid][2 * q] = re; red[wid][2 * q + 1] = im; }
  }
  __syncthreads();
  if (threadIdx.x < 24) {
    double acc = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) acc += red[w][threadIdx.x];
    const long long base = ((long long)f * d.nt * d.np + dir) * 12;
    double *o = reinterpret_cast<double *>(out + base);
    o[threadIdx.x] = acc;
  }
}

// ---------------------------------------------------------------------------
// NffftBroadBand::calculateLNXYZ / calculateAuxiliary
// (nffft_broadband.cpp:348-466).  A series is one face current c (0..7, the
// order of the NF-FF accumulators) of one surface cell, sampled every step
// into [step][c][cell] by k_nffft_sample.
struct BbFar {
  Nffft n;
  long long T, Ns, Q;
  double rx, ry, rz;  // observation direction
  double cx, cy, cz;  // box centre (grid/grid_box.cpp:101-105)
  double dl, ds;
  int zfaces;         // z faces only when the box has nz > 1
  const double *k;    // wave number of every bin
};

// series s0 .. s0+chunk-1 as FP64 rows of Ns points, zero past step T-1 and
// for series past s0+n: 64 x 64 tiles through LDS, coalesced on both sides
template <typename T>
__global__ void __launch_bounds__(256) k_bb_rows(const T *__restrict__ hist, long long S, long long Tn, long long Ns,
                                                 long long s0, long long n, long long chunk,
                                                 double *__restrict__ rows) {
  __shared__ double tile[64][65];
  const long long t0 = (long long)blockIdx.x * 64, c0 = (long long)blockIdx.y * 64;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int r = ty; r < 64; r += 4) {
    const long long t = t0 + r, c = c0 + tx;
    double v = 0.0;
    if (t < Tn && c < n) v = (double)hist[t * S + s0 + c];
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const long long c = c0 + r, t = t0 + tx;
    if (c < chunk && t < Ns) rows[c * Ns + t] = tile[tx][r];
  }
}

// one thread per (bin q, slice of the chunk's series): adds
// ds e^{i delay k_q} X_s(m), m = q - (T-1), to the bin's sum of the series'
// target (n_x, n_y, n_z, l_x, l_y, l_z).  spec holds the non-negative half of
// each real series' DFT; X(-m) = conj(X(m)).
__global__ void __launch_bounds__(256) k_bb_fold(BbFar d, const double2 *__restrict__ spec, long long s0, long long n,
                                                 int slices, double2 *part, int first) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int slice = blockIdx.y;
  if (q >= d.Ns) return;
  const long long m = q - (d.T - 1), am = m < 0 ? -m : m;
  const double kq = d.k[q];
  double2 acc[6];
#pragma unroll
  for (int u = 0; u < 6; ++u)
    acc[u] = first ? make_double2(0.0, 0.0) : part[((long long)slice * 6 + u) * d.Ns + q];
  const long long per = (n + slices - 1) / slices;
  const long long a = (long long)slice * per, b = a + per < n ? a + per : n;
  const Nffft &f = d.n;
  const long long cxq = (long long)f.nyb * f.nzb, cyq = (long long)f.nzb * f.nxb;
  for (long long sl = a; sl < b; ++sl) {
    const long long s = s0 + sl;
    const int c = (int)(s / d.Q);
    long long cell = s - (long long)c * d.Q;
    const bool side = c >= 4;
    const int cur = c & 3;
    double x, y, z;
    int tgt;
    if (cell < cxq) {
      const int i = side ? f.ri : f.li, j = f.lj + (int)(cell / f.nzb), k = f.lk + (int)(cell % f.nzb);
      x = ((double)i + 0.0) - d.cx; y = ((double)j + 0.5) - d.cy; z = ((double)k + 0.5) - d.cz;
      tgt = cur == 0 ? 4 : cur == 1 ? 5 : cur == 2 ? 1 : 2;  // my, mz, jy, jz
    } else if ((cell -= cxq) < cyq) {
      const int i = f.li + (int)(cell / f.nzb), k = f.lk + (int)(cell % f.nzb), j = side ? f.rj : f.lj;
      x = ((double)i + 0.5) - d.cx; y = ((double)j + 0.0) - d.cy; z = ((double)k + 0.5) - d.cz;
      tgt = cur == 0 ? 5 : cur == 1 ? 3 : cur == 2 ? 2 : 0;  // mz, mx, jz, jx
    } else {
      if (!d.zfaces) continue;
      cell -= cyq;
      const int i = f.li + (int)(cell / f.nyb), j = f.lj + (int)(cell % f.nyb), k = side ? f.rk : f.lk;
      x = ((double)i + 0.5) - d.cx; y = ((double)j + 0.5) - d.cy; z = ((double)k + 0.0) - d.cz;
      tgt = cur == 0 ? 3 : cur == 1 ? 4 : cur == 2 ? 0 : 1;  // mx, my, jx, jy
    }
    const double delay = (x * d.rx * d.dl + y * d.ry * d.dl) + z * d.rz * d.dl;
    double2 X = spec[sl * d.T + am];
    if (m < 0) X.y = -X.y;
    double sn, cs;
    sincos(delay * kq, &sn, &cs);
    const double2 t = cmul(make_double2(d.ds * cs, d.ds * sn), X);
    switch (tgt) {
      case 0: acc[0] = cadd(acc[0], t); break;
      case 1: acc[1] = cadd(acc[1], t); break;
      case 2: acc[2] = cadd(acc[2], t); break;
      case 3: acc[3] = cadd(acc[3], t); break;
      case 4: acc[4] = cadd(acc[4], t); break;
      default: acc[5] = cadd(acc[5], t); break;
    }
  }
#pragma unroll
  for (int u = 0; u < 6; ++u) part[((long long)slice * 6 + u) * d.Ns + q] = acc[u];
}

// out[i] = sum over slices of part[slice][i], in slice order
__global__ void __launch_bounds__(256) k_bb_sum_slices(const double2 *__restrict__ part, int slices, long long total,
                                                       double2 *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double2 a = make_double2(0.0, 0.0);
  for (int s = 0; s < slices; ++s) a = cadd(a, part[(long long)s * total + i]);
  out[i] = a;
}

// ---------------------------------------------------------------------------
// DFT monitor of one component on a plane (the adjoint's objective): after
// step t, F(u, v) += f(cell) * w[t], FP64.  (u, v) run over the two axes
// other than `axis`, in increasing axis order.
struct DftPlane {
  int comp, axis, index, lo1, hi1, lo2, hi2;
  double2 *acc;       // [hi1-lo1][hi2-lo2]
  const double2 *w;   // [T]
};

__device__ __forceinline__ long long plane_cell(const DftPlane &d, long long si, long long sj, int u, int v) {
  if (d.axis == 0) return (long long)d.index * si + (long long)u * sj + v;
  if (d.axis == 1) return (long long)u * si + (long long)d.index * sj + v;
  return (long long)u * si + (long long)v * sj + d.index;
}

template <typename T>
__device__ __forceinline__ void dft_plane_body(const T *f, long long si, long long sj, const DftPlane &d,
                                               long long step, int q) {
  const int n2 = d.hi2 - d.lo2, n = (d.hi1 - d.lo1) * n2;
  if (q >= n) return;
  const int u = d.lo1 + q / n2, v = d.lo2 + q % n2;
  const double x = (double)f[plane_cell(d, si, sj, u, v)];
  const double2 w = d.w[step];
  double2 a = d.acc[q];
  a.x = a.x + x * w.x;
  a.y = a.y + x * w.y;
  d.acc[q] = a;
}
template <typename T>
__global__ void k_dft_plane(const T *f, long long si, long long sj, DftPlane d, long long step) {
  dft_plane_body<T>(f, si, sj, d, step, blockIdx.x * blockDim.x + threadIdx.x);
}

// E of every design cell into hist[slot][cell][3] (for the gradient)
template <typename T>
__device__ __forceinline__ void design_snap_body(const Grid<T> &g, T *hist, int slot, int q) {
  const int nj = g.db[3] - g.db[2], nk = g.db[5] - g.db[4];
  const int nd = (g.db[1] - g.db[0]) * nj * nk;
  if (q >= nd) return;
  const int i = g.db[0] + q / (nj * nk), j = g.db[2] + (q / nk) % nj, k = g.db[4] + q % nk;
  const long long c = (long long)i * g.si + (long long)j * g.sj + k;
  T *h = hist + ((long long)slot * nd + q) * 3;
  h[0] = g.f[EX][c];
  h[1] = g.f[EY][c];
  h[2] = g.f[EZ][c];
}
template <typename T>
__global__ void k_design_snap(Grid<T> g, T *hist, int slot) {
  design_snap_body<T>(g, hist, slot, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the step map: the transpose of the E and H half-steps
// with their CPML terms and Drude currents, in the adjoint fields Ea, Ha,
// the psi adjoints Pe, Ph and the Drude J adjoints Ja.
//
// A half-step's transpose scales each updated cell's own adjoint by the
// cell's self coefficient and adds to every cell whose difference the update
// took a "difference weight": the curl coefficient times the adjoint, plus
// the CPML a-coefficient times the psi adjoint chi = p + q * adjoint (p then
// becomes b * chi).  The weights are not stored: a kernel evaluates the
// weights of its own cell and of the neighbours it gathers from out of the
// neighbours' adjoints and coefficients.  So between kernels the fields hold
// gathered adjoints whose own-cell scaling is still to come:
//   Ea: after the H transpose's gather and the objective injection, before
//       the E transpose's scaling;
//   Ha: after the E transpose's gather, before the H transpose's scaling.
// k_adj_e applies the pending E scaling and gathers the H-side weights W;
// k_adj_h applies the pending H scaling and gathers the E-side weights V.
// A psi adjoint is read (old) by the cells that gather its weight and
// updated by the cell it belongs to, so the two copies Pe[par], Pe[par ^ 1]
// (and Ph) alternate from step to step.  A Drude J adjoint is read by the H
// gathers of a step and updated by the next E scaling.
//
//   H-side weights (W): Wxy chxey (Ey k-diff), Wxz chxez (Ez j), Wyz chyez
//     (Ez i), Wyx chyex (Ex k), Wzx chzex (Ex j), Wzy chzey (Ey i)
//   E-side weights (V): Vxy cexhy (Hy k-diff), Vxz cexhz (Hz j), Vyz ceyhz
//     (Hz i), Vyx ceyhx (Hx k), Vzx cezhx (Hx j), Vzy cezhy (Hy i)
enum { W_XY = 0, W_XZ, W_YZ, W_YX, W_ZX, W_ZY };
// k_adj_e modes: a full step; the first step of a run whose E adjoint is
// given unscaled; the final E scaling alone
enum { ADJ_STEP = 0, ADJ_FIRST = 1, ADJ_FINAL = 2 };

template <typename T>
struct Adj {
  T *Ea[3], *Ha[3];
  T *Pe[2][6], *Ph[2][6];  // psi adjoints (two copies each), psi layouts
  T *Ja[3];                // Drude J adjoints, Drude-box layout
  // objective injection (plane of comp < 3) and design gradient
  DftPlane dft;       // dft.acc = final forward F
  int inject;
  const T *hist;      // design E history of the segment [slot][cell][3]
  const double *eps_r;  // design cells
  double *grad;
  int grad_on;
};

template <typename T>
__device__ __forceinline__ P2<T> pair_at(const T *tab, int n, int nmat, int m) {
  return reinterpret_cast<const P2<T> *>(tab)[n * nmat + m];
}

// H-side weights of cell (i, j, k) (flat index c) from its gathered H adjoint
// (hx, hy, hz).  The psi terms are formed for the components in `mask` (bit
// W_*); own: this thread also updates the cell's psi adjoints (all six).
template <typename T>
__device__ __forceinline__ void adj_w(const Grid<T> &g, const Adj<T> &a, int par, long long c, int i, int j,
                                      int k, T hx, T hy, T hz, unsigned mask, bool own, T w[6]) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const bool sx = j < ny && k < nz, sy = i < nx && k < nz, sz = i < nx && j < ny;
  const int m = g.mat[c];
  const P2<T> pk = pair_at(g.kt[1], k, nmat, m), pj = pair_at(g.jt[1], j, nmat, m),
              pi = pair_at(g.it[1], i, nmat, m);
  const bool yee = i < nx && j < ny && k < nz;
  w[W_XY] = yee ? pk.x * hx : T(0);
  w[W_XZ] = yee ? pj.x * hx : T(0);
  w[W_YZ] = yee ? pi.x * hy : T(0);
  w[W_YX] = yee ? pk.y * hy : T(0);
  w[W_ZX] = yee ? pj.y * hz : T(0);
  w[W_ZY] = yee ? pi.y * hz : T(0);
  const int lx = h_layer(i, nx, g.dn[0], g.dp[0]), ly = h_layer(j, ny, g.dn[1], g.dp[1]),
            lz = h_layer(k, nz, g.dn[2], g.dp[2]);
  if (lx < 0 && ly < 0 && lz < 0) return;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = ny + 1;
  const P2<T> qk = pair_at(g.kp[1], k, nmat, m), qj = pair_at(g.jp[1], j, nmat, m),
              qi = pair_at(g.ip[1], i, nmat, m);
  T *const *P = a.Ph[par];
  T *const *Q = a.Ph[par ^ 1];
  auto term = [&](bool on, int q, int slot, long long o, T qc, T h, T am, T bm) {
    if (!on || !(own || ((mask >> q) & 1u))) return;
    const T chi = P[slot][o] + qc * h;
    w[q] = w[q] + am * chi;
    if (own) Q[slot][o] = bm * chi;
  };
  if (ly >= 0) {
    const long long o = psi_y(i, ly, k, Ly, g.pz);
    term(sx, W_XZ, P_AY, o, qj.x, hx, g.am[1][j], g.bm[1][j]);
    term(sz, W_ZX, P_CY, o, qj.y, hz, g.am[1][j], g.bm[1][j]);
  }
  if (lz >= 0) {
    const long long o = psi_z(i, j, lz, nm1y, Lz);
    term(sx, W_XY, P_AZ, o, qk.x, hx, g.am[2][k], g.bm[2][k]);
    term(sy, W_YX, P_BZ, o, qk.y, hy, g.am[2][k], g.bm[2][k]);
  }
  if (lx >= 0) {
    const long long o = psi_x(lx, j, k, nm1y, g.pz);
    term(sy, W_YZ, P_BX, o, qi.x, hy, g.am[0][i], g.bm[0][i]);
    term(sz, W_ZY, P_CX, o, qi.y, hz, g.am[0][i], g.bm[0][i]);
  }
}

// E-side weights of cell (i, j, k) from its gathered E adjoint (ex, ey, ez):
// design cells carry their own coefficients; a Drude cell's weights are
// -cb t and +cb t of t = e + beta jb, with the J adjoint jb not yet updated
// for this adjoint (the transpose of E' = ca E + cb (curl - (dl/2)(1+k) J),
// J' = k J + beta (E' + E), taken J' first).
template <typename T>
__device__ __forceinline__ void adj_v(const Grid<T> &g, const Adj<T> &a, int par, long long c, int i, int j,
                                      int k, T ex, T ey, T ez, unsigned mask, bool own, T v[6]) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const bool sx = i < nx, sy = j < ny, sz = k < nz;  // inside Ex / Ey / Ez shapes
  const int m = g.mat[c];
  P2<T> pk = pair_at(g.kt[0], k, nmat, m), pj = pair_at(g.jt[0], j, nmat, m), pi = pair_at(g.it[0], i, nmat, m);
  if (in_design(g, i, j, k)) {
    const T *d = g.dcoef + 9 * design_cell(g, i, j, k);
    pk.x = d[3]; pk.y = d[4]; pj.x = d[5]; pj.y = d[6]; pi.x = d[7]; pi.y = d[8];
  }
  const bool ix = i >= 1 && i < nx, jin = j >= 1 && j < ny, kin = k >= 1 && k < nz;
  const bool yx = sx && jin && kin, yy = ix && sy && kin, yz = ix && jin && sz;
  v[W_XY] = yx ? pk.x * ex : T(0);
  v[W_XZ] = yx ? pj.x * ex : T(0);
  v[W_YZ] = yy ? pi.x * ey : T(0);
  v[W_YX] = yy ? pk.y * ey : T(0);
  v[W_ZX] = yz ? pj.y * ez : T(0);
  v[W_ZY] = yz ? pi.y * ez : T(0);
  const int lx = e_layer(i, nx, g.dn[0], g.dp[0]), ly = e_layer(j, ny, g.dn[1], g.dp[1]),
            lz = e_layer(k, nz, g.dn[2], g.dp[2]);
  if (lx >= 0 || ly >= 0 || lz >= 0) {
    const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = ny + 1;
    const P2<T> qk = pair_at(g.kp[0], k, nmat, m), qj = pair_at(g.jp[0], j, nmat, m),
                qi = pair_at(g.ip[0], i, nmat, m);
    T *const *P = a.Pe[par];
    T *const *Q = a.Pe[par ^ 1];
    auto term = [&](bool on, int q, int slot, long long o, T qc, T e, T ae, T be) {
      if (!on || !(own || ((mask >> q) & 1u))) return;
      const T chi = P[slot][o] + qc * e;
      v[q] = v[q] + ae * chi;
      if (own) Q[slot][o] = be * chi;
    };
    if (ly >= 0) {
      const long long o = psi_y(i, ly, k, Ly, g.pz);
      term(sx, W_XZ, P_AY, o, qj.x, ex, g.ae[1][j], g.be[1][j]);
      term(sz, W_ZX, P_CY, o, qj.y, ez, g.ae[1][j], g.be[1][j]);
    }
    if (lz >= 0) {
      const long long o = psi_z(i, j, lz, nm1y, Lz);
      term(sx, W_XY, P_AZ, o, qk.x, ex, g.ae[2][k], g.be[2][k]);
      term(sy, W_YX, P_BZ, o, qk.y, ey, g.ae[2][k], g.be[2][k]);
    }
    if (lx >= 0) {
      const long long o = psi_x(lx, j, k, nm1y, g.pz);
      term(sy, W_YZ, P_BX, o, qi.x, ey, g.ae[0][i], g.be[0][i]);
      term(sz, W_ZY, P_CX, o, qi.y, ez, g.ae[0][i], g.be[0][i]);
    }
  }
  // Drude cells lie outside every CPML slab
  if (in_drude(g, i, j, k)) {
    const T *dc = g.drc + XF_ADE_STRIDE * m;
    const T cb = dc[1];
    if (cb != T(0)) {
      const int dq = drude_cell(g, i, j, k);
      const T beta = dc[4];
      if (yx) {
        const T t = ex + beta * a.Ja[0][dq];
        v[W_XY] = -cb * t;  // -(Hy(k) - Hy(k-1))
        v[W_XZ] = cb * t;   // +(Hz(j) - Hz(j-1))
      }
      if (yy) {
        const T t = ey + beta * a.Ja[1][dq];
        v[W_YZ] = -cb * t;  // -(Hz(i) - Hz(i-1))
        v[W_YX] = cb * t;   // +(Hx(k) - Hx(k-1))
      }
      if (yz) {
        const T t = ez + beta * a.Ja[2][dq];
        v[W_ZX] = -cb * t;  // -(Hx(j) - Hx(j-1))
        v[W_ZY] = cb * t;   // +(Hy(i) - Hy(i-1))
      }
    }
  }
}

// the E transpose's own-cell part at cell (i, j, k): ca e on Yee cells; on
// Drude cells beta jb + ca t with t = e + beta jb, and the J adjoint becomes
// k jb - cb (dl/2)(1+k) t
template <typename T>
__device__ __forceinline__ void adj_scale_e(const Grid<T> &g, const Adj<T> &a, long long c, int i, int j, int k,
                                            T &ex, T &ey, T &ez) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int m = g.mat[c];
  const T *self = g.self[0] + 4 * m;
  T ca_x = self[0], ca_y = self[1], ca_z = self[2];
  if (in_design(g, i, j, k)) {
    const T *d = g.dcoef + 9 * design_cell(g, i, j, k);
    ca_x = d[0]; ca_y = d[1]; ca_z = d[2];
  }
  const bool sx = i < nx, sy = j < ny, sz = k < nz;
  const bool ix = i >= 1 && i < nx, jin = j >= 1 && j < ny, kin = k >= 1 && k < nz;
  const bool yx = sx && jin && kin, yy = ix && sy && kin, yz = ix && jin && sz;
  T ox = yx ? ca_x * ex : ex, oy = yy ? ca_y * ey : ey, oz = yz ? ca_z * ez : ez;
  if (in_drude(g, i, j, k)) {
    const T *dc = g.drc + XF_ADE_STRIDE * m;
    const T cb = dc[1];
    if (cb != T(0)) {
      const int dq = drude_cell(g, i, j, k);
      const T ca = dc[0], hk = dc[5] * dc[2], kk = dc[3], beta = dc[4];
      if (yx) {
        const T jb = a.Ja[0][dq], t = ex + beta * jb;
        ox = beta * jb + ca * t;
        a.Ja[0][dq] = kk * jb - cb * hk * t;
      }
      if (yy) {
        const T jb = a.Ja[1][dq], t = ey + beta * jb;
        oy = beta * jb + ca * t;
        a.Ja[1][dq] = kk * jb - cb * hk * t;
      }
      if (yz) {
        const T jb = a.Ja[2][dq], t = ez + beta * jb;
        oz = beta * jb + ca * t;
        a.Ja[2][dq] = kk * jb - cb * hk * t;
      }
    }
  }
  ex = ox;
  ey = oy;
  ez = oz;
}

// The adjoint kernels run one thread per cell of the padded layout, 64
// consecutive cells per wave, so a cell's k neighbour is the adjacent lane:
// every cell takes the two weights it gathers from its k neighbour from that
// lane by DPP instead of forming them again; only the wave's end lane (lane 0
// for k - 1, lane 63 for k + 1) forms its neighbour's itself.  A launch needs
// adj_blocks(ncell) workgroups of 256.
__host__ __device__ inline long long adj_blocks(long long ncell) { return (ncell + 255) / 256; }

// Workgroup ids go round-robin over the 8 XCDs, each with its own L2.  The
// batched adjoint kernels hand each XCD one contiguous run of blocks instead,
// so the cells whose j + 1 / i + 1 (j - 1 / i - 1) neighbours a block reads
// are read again, as their own, by a block of the same XCD shortly after (or
// before).  (A single engine's adjoint at 128^3 keeps the plain order: there
// the whole state sits in the Infinity Cache and the runs were slower.)
__device__ __forceinline__ long long adj_block(long long b, long long n) {
  const long long q = n >> 3, r = n & 7, xcd = b & 7, local = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

// per E cell: the pending E scaling, the gather of the H-side weights of the
// cell and of its k-1, j-1, i-1 neighbours, the objective injection and the
// gradient; t is the thread's index in the launch
template <typename T>
__device__ __forceinline__ void adj_e_body(const Grid<T> &g, const Adj<T> &a, long long step, int slot, int par,
                                           int mode, long long t) {
  const long long c = t;
  const int lane = (int)(t & 63), wave = (int)((t >> 6) & 3);
  const int i = (int)(c / g.si), r = (int)(c - (long long)i * g.si), j = r / (int)g.sj, k = r - j * (int)g.sj;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool valid = i <= nx && j <= ny && k <= nz;
  if (mode == ADJ_FINAL) {
    if (!valid) return;
    T ex = a.Ea[0][c], ey = a.Ea[1][c], ez = a.Ea[2][c];
    adj_scale_e(g, a, c, i, j, k, ex, ey, ez);
    a.Ea[0][c] = ex;
    a.Ea[1][c] = ey;
    a.Ea[2][c] = ez;
    return;
  }
  const long long cj = c - g.sj, ci = c - g.si;
  const bool sx = i < nx, sy = j < ny, sz = k < nz;  // inside Ex / Ey / Ez shapes
  // gather: transpose of the H curl differences (forward differences of E)
  T ex = T(0), ey = T(0), ez = T(0);
  T w[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, wj[6], wi[6];
  if (valid) {
    ex = a.Ea[0][c];
    ey = a.Ea[1][c];
    ez = a.Ea[2][c];
    if (mode != ADJ_FIRST) adj_scale_e(g, a, c, i, j, k, ex, ey, ez);
    adj_w(g, a, par, c, i, j, k, a.Ha[0][c], a.Ha[1][c], a.Ha[2][c], 0u, true, w);
  }
  // the k - 1 cell's weights: from the lane below (whose cell it is when
  // k >= 1); lane 0 takes lane 63's of the wave below through LDS, and lane
  // 0 of a workgroup's first wave forms them
  T wk_xy = from_below(w[W_XY], T(0)), wk_yx = from_below(w[W_YX], T(0));
  __shared__ P2<T> edge[4];
  if (lane == 63) {
    edge[wave].x = w[W_XY];
    edge[wave].y = w[W_YX];
  }
  __syncthreads();
  if (!valid) return;
  if (lane == 0 && k >= 1 && (sx || sy)) {
    if (wave > 0) {
      wk_xy = edge[wave - 1].x;
      wk_yx = edge[wave - 1].y;
    } else {
      T wk[6];
      adj_w(g, a, par, c - 1, i, j, k - 1, a.Ha[0][c - 1], a.Ha[1][c - 1], T(0), (1u << W_XY) | (1u << W_YX), false,
            wk);
      wk_xy = wk[W_XY];
      wk_yx = wk[W_YX];
    }
  }
  if (j >= 1 && (sx || sz))
    adj_w(g, a, par, cj, i, j - 1, k, a.Ha[0][cj], T(0), a.Ha[2][cj], (1u << W_XZ) | (1u << W_ZX), false, wj);
  if (i >= 1 && (sy || sz))
    adj_w(g, a, par, ci, i - 1, j, k, T(0), a.Ha[1][ci], a.Ha[2][ci], (1u << W_YZ) | (1u << W_ZY), false, wi);
  if (sx) ex = ex + (((k >= 1 ? wk_yx : T(0)) - w[W_YX]) + ((j >= 1 ? wj[W_ZX] : T(0)) - w[W_ZX]));
  if (sy) ey = ey + (((k >= 1 ? wk_xy : T(0)) - w[W_XY]) + ((i >= 1 ? wi[W_ZY] : T(0)) - w[W_ZY]));
  if (sz) ez = ez + (((j >= 1 ? wj[W_XZ] : T(0)) - w[W_XZ]) + ((i >= 1 ? wi[W_YZ] : T(0)) - w[W_YZ]));
  // objective: dJ/dE = 2 Re(conj(F) w_t) on the monitor plane
  if (a.inject) {
    const DftPlane &d = a.dft;
    int u, v, w;
    if (d.axis == 0) { w = i; u = j; v = k; } else if (d.axis == 1) { u = i; w = j; v = k; } else { u = i; v = j; w = k; }
    if (w == d.index && u >= d.lo1 && u < d.hi1 && v >= d.lo2 && v < d.hi2) {
      const double2 F = d.acc[(u - d.lo1) * (d.hi2 - d.lo2) + (v - d.lo2)];
      const double2 wt = d.w[step];
      const T inj = (T)(2 * (F.x * wt.x + F.y * wt.y));
      if (d.comp == EX) ex = ex + inj;
      if (d.comp == EY) ey = ey + inj;
      if (d.comp == EZ) ez = ez + inj;
    }
  }
  const bool des = in_design(g, i, j, k);
  const int dq = des ? design_cell(g, i, j, k) : 0;
  // gradient: with sigma = 0 in the design box, dE'/d eps = -(E' - E)/eps
  if (a.grad_on && des) {
    const int nd = (g.db[1] - g.db[0]) * (g.db[3] - g.db[2]) * (g.db[5] - g.db[4]);
    const T *h1 = a.hist + ((long long)(slot + 1) * nd + dq) * 3;
    const T *h0 = a.hist + ((long long)slot * nd + dq) * 3;
    double acc = 0;
    acc = acc + (double)ex * ((double)h1[0] - (double)h0[0]);
    acc = acc + (double)ey * ((double)h1[1] - (double)h0[1]);
    acc = acc + (double)ez * ((double)h1[2] - (double)h0[2]);
    a.grad[dq] = a.grad[dq] - acc / a.eps_r[dq];
  }
  a.Ea[0][c] = ex;
  a.Ea[1][c] = ey;
  a.Ea[2][c] = ez;
}
// Waves-per-EU hints at the occupancy the kernels reach anyway (8 in FP32,
// 7 in FP64) schedule them faster: about 5% in FP32, 1% in FP64 (a hint of 8
// would make the FP64 kernels spill)
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 7)))
k_adj_e(Grid<T> g, Adj<T> a, long long step, int slot, int par, int mode, long long b0) {
  adj_e_body<T>(g, a, step, slot, par, mode, (b0 + (long long)blockIdx.x) * blockDim.x + threadIdx.x);
}

// per H cell: the pending H scaling, then the gather of the E-side weights of
// the cell and of its k+1, j+1, i+1 neighbours; t is the thread's index in
// the launch
template <typename T>
__device__ __forceinline__ void adj_h_body(const Grid<T> &g, const Adj<T> &a, int par, long long t) {
  const long long c = t;
  const int lane = (int)(t & 63), wave = (int)((t >> 6) & 3);
  const int i = (int)(c / g.si), r = (int)(c - (long long)i * g.si), j = r / (int)g.sj, k = r - j * (int)g.sj;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool valid = i <= nx && j <= ny && k <= nz;
  // inside the Hx (nx+1, ny, nz), Hy (nx, ny+1, nz), Hz (nx, ny, nz+1) shapes
  const bool sx = j < ny && k < nz, sy = i < nx && k < nz, sz = i < nx && j < ny;
  const long long cj = c + g.sj, ci = c + g.si;
  // gather: transpose of the E curl differences (backward differences of H)
  T v[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, vj[6], vi[6];
  if (valid) adj_v(g, a, par, c, i, j, k, a.Ea[0][c], a.Ea[1][c], a.Ea[2][c], 0u, true, v);
  // the k + 1 cell's weights: from the lane above (whose cell it is when
  // k < nz); lane 63 takes lane 0's of the wave above through LDS, and lane
  // 63 of a workgroup's last wave forms them
  T vk_xy = from_above(v[W_XY], T(0)), vk_yx = from_above(v[W_YX], T(0));
  __shared__ P2<T> edge[4];
  if (lane == 0) {
    edge[wave].x = v[W_XY];
    edge[wave].y = v[W_YX];
  }
  __syncthreads();
  if (!valid) return;
  if (lane == 63 && k < nz && (sx || sy)) {
    if (wave < 3) {
      vk_xy = edge[wave + 1].x;
      vk_yx = edge[wave + 1].y;
    } else {
      T vk[6];
      adj_v(g, a, par, c + 1, i, j, k + 1, a.Ea[0][c + 1], a.Ea[1][c + 1], T(0), (1u << W_XY) | (1u << W_YX), false,
            vk);
      vk_xy = vk[W_XY];
      vk_yx = vk[W_YX];
    }
  }
  const bool yee = i < nx && j < ny && k < nz;
  const T *self = g.self[1] + 4 * (int)g.mat[c];
  T hx = a.Ha[0][c], hy = a.Ha[1][c], hz = a.Ha[2][c];
  hx = yee ? self[0] * hx : hx;
  hy = yee ? self[1] * hy : hy;
  hz = yee ? self[2] * hz : hz;
  if (j < ny && (sx || sz))
    adj_v(g, a, par, cj, i, j + 1, k, a.Ea[0][cj], T(0), a.Ea[2][cj], (1u << W_XZ) | (1u << W_ZX), false, vj);
  if (i < nx && (sy || sz))
    adj_v(g, a, par, ci, i + 1, j, k, T(0), a.Ea[1][ci], a.Ea[2][ci], (1u << W_YZ) | (1u << W_ZY), false, vi);
  // vk: zero where k = nz (no cell above; the term is then unused)
  const T vk_yx_ = k < nz ? vk_yx : T(0), vk_xy_ = k < nz ? vk_xy : T(0);
  if (sx) hx = hx + ((v[W_YX] - vk_yx_) + (v[W_ZX] - vj[W_ZX]));
  if (sy) hy = hy + ((v[W_XY] - vk_xy_) + (v[W_ZY] - vi[W_ZY]));
  if (sz) hz = hz + ((v[W_XZ] - vj[W_XZ]) + (v[W_YZ] - vi[W_YZ]));
  a.Ha[0][c] = hx;
  a.Ha[1][c] = hy;
  a.Ha[2][c] = hz;
}
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 7)))
k_adj_h(Grid<T> g, Adj<T> a, int par, long long b0) {
  adj_h_body<T>(g, a, par, (b0 + (long long)blockIdx.x) * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Batched launches: many independent simulations of one precision stepped
// together, each per-step kernel one launch over all of them.  A launch's
// workgroup b handles local block map[b].y of instance map[b].x, whose
// arguments are read from device memory.
template <typename T>
struct Inst {
  Grid<T> g;
  Tiling tl;
  Tfsf t;
  TfsfESeg es;
  TfsfHSeg hs;
  Nffft nf;
  T *ring;  // NF-FF sample ring ([NF_RING][8][surface cells])
  DftPlane dft;
};

// an instance of a batched gradient: its grid, adjoint buffers and the
// segment's design-box E history
template <typename T>
struct InstAdj {
  Grid<T> g;
  Adj<T> a;
  T *hist;
};

template <typename T, bool LDS, bool SPECIAL>
__global__ void __launch_bounds__(256) kb_update_e(const Inst<T> *__restrict__ in, const int2 *__restrict__ map) {
  const int2 w = map[blockIdx.x];
  update_e_body<T, LDS, SPECIAL>(in[w.x].g, in[w.x].tl, ZFaces{}, w.y);
}
template <typename T, bool LDS>
__global__ void __launch_bounds__(256) kb_update_h(const Inst<T> *__restrict__ in, const int2 *__restrict__ map,
                                                   long long step) {
  const int2 w = map[blockIdx.x];
  update_h_body<T, LDS>(in[w.x].g, in[w.x].tl, ZFaces{}, NfRec{}, w.y, step);
}
template <typename T>
__global__ void kb_tfsf_e(const Inst<T> *__restrict__ in, const int2 *__restrict__ map) {
  const int2 w = map[blockIdx.x];
  const Inst<T> &a = in[w.x];
  tfsf_e_body<T>(a.g, a.t, a.es, w.y);
}
template <typename T>
__global__ void kb_tfsf_h(const Inst<T> *__restrict__ in, const int2 *__restrict__ map) {
  const int2 w = map[blockIdx.x];
  const Inst<T> &a = in[w.x];
  tfsf_h_body<T>(a.g, a.t, a.hs, w.y);
}
// one workgroup per instance with a TFSF source: map[b].x
template <typename T>
__global__ void __launch_bounds__(1024) kb_tfsf_ifa(const Inst<T> *__restrict__ in, const int2 *__restrict__ map,
                                                    long long step) {
  tfsf_ifa_body(in[map[blockIdx.x].x].t, step);
}
template <typename T>
__global__ void kb_nffft_sample(const Inst<T> *__restrict__ in, const int2 *__restrict__ map, int slot) {
  const int2 w = map[blockIdx.x];
  nffft_sample_body<T>(in[w.x].g, in[w.x].nf, in[w.x].ring, slot, w.y);
}
template <typename T>
__global__ void __launch_bounds__(256) kb_nffft_flush(const Inst<T> *__restrict__ in, const int2 *__restrict__ map,
                                                     int nslots, long long step0) {
  const int2 w = map[blockIdx.x];
  nffft_flush_body<T>(in[w.x].nf, in[w.x].ring, nslots, step0, w.y);
}
// 256-thread blocks over each instance's DFT plane
template <typename T>
__global__ void kb_dft_plane(const Inst<T> *__restrict__ in, const int2 *__restrict__ map, long long step) {
  const int2 w = map[blockIdx.x];
  const Inst<T> &a = in[w.x];
  dft_plane_body<T>(a.g.f[a.dft.comp], a.g.si, a.g.sj, a.dft, step, w.y * 256 + threadIdx.x);
}
// 256-thread blocks over each instance's design box / padded grid
template <typename T>
__global__ void kb_design_snap(const InstAdj<T> *__restrict__ in, const int2 *__restrict__ map, int slot) {
  const int2 w = map[blockIdx.x];
  design_snap_body<T>(in[w.x].g, in[w.x].hist, slot, w.y * 256 + threadIdx.x);
}
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 7)))
kb_adj_e(const InstAdj<T> *__restrict__ in, const int2 *__restrict__ map, long long step, int slot, int par) {
  const int2 w = map[adj_block(blockIdx.x, gridDim.x)];
  adj_e_body<T>(in[w.x].g, in[w.x].a, step, slot, par, ADJ_STEP, (long long)w.y * 256 + threadIdx.x);
}
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 7)))
kb_adj_h(const InstAdj<T> *__restrict__ in, const int2 *__restrict__ map, int par) {
  const int2 w = map[adj_block(blockIdx.x, gridDim.x)];
  adj_h_body<T>(in[w.x].g, in[w.x].a, par, (long long)w.y * 256 + threadIdx.x);
}

}  // namespace xfk
