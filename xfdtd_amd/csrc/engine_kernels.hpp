// engine_kernels.hpp -- gfx950 kernels of the FDTD field engine.
//
// Layout in HBM: every field component, the material index and the x/y CPML
// psi arrays use one padded (nx+1) x (ny+1) x pz grid, z fastest, pz a
// multiple of 32 elements so each z row starts on a 128-byte line.  The six
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
  const T *ca;      // [6][nmat]
  const T *cb[12];  // [nmat][n_axis+1]
  const T *cpsi[12];
  // CPML per axis
  int dn[3], dp[3];
  const T *be[3], *ae[3], *bm[3], *am[3];
  T *psi_e[6], *psi_h[6];
  int nx, ny, nz, pz, nmat;
  long long sj, si;  // strides of j and i in elements
};

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
};

// tile of this workgroup; false for a padding workgroup of the column order
__device__ __forceinline__ bool tile_of_block(const Tiling &t, int &bz, int &by, int &bx) {
  const int b = blockIdx.x;
  const int q = t.nblocks >> 3, r = t.nblocks & 7;
  const int xcd = b & 7, local = b >> 3;
  if (t.xcd_cols) {
    bz = local % t.ntz;
    const int col = (local / t.ntz) * 8 + xcd;
    by = col % t.nty;
    bx = col / t.nty;
    return bx < t.ntx;
  }
  const int id = t.xcd_runs ? xcd * q + min(xcd, r) + local : b;
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
template <typename T>
struct TileTables {
  const T *ca;                                 // [3][nmat] self terms
  const T *tk0, *tk1, *tj0, *tj1, *ti0, *ti1;  // curl coefficients [nmat][slots]
  const T *pk0, *pk1, *pj0, *pj1, *pi0, *pi1;  // CPML psi coefficients, same slots
  int sk, sj, si;                              // slots per material along k, j, i
  int ks, js, i0;                              // this thread's k and j slots, tile's first plane
};

// The global tables one curl kernel reads: three self terms from ca_base, and
// per axis two curl tables and the two psi tables of the same components,
// each [nmat][n + 1].
template <typename T>
struct TableSrc {
  int ca_base;
  const T *k0, *k1, *j0, *j1, *i0, *i1;
  const T *pk0, *pk1, *pj0, *pj1, *pi0, *pi1;
  int nk1, nj1, ni1;
};

template <typename T>
__host__ __device__ inline size_t lds_table_bytes(int nmat, const Tiling &tl) {
  return sizeof(T) * (size_t)nmat * (3 + 4 * tl.krows + 4 * tl.jrows + 4 * tl.tx);
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
__device__ __forceinline__ TileTables<T> stage_tables(T *lds, const T *ca, int nmat, const TableSrc<T> &s,
                                                     const Tiling &tl, const Cell &cell, int i0,
                                                     bool psi_k, bool psi_j, bool psi_i) {
  const int nk = nmat * tl.krows, nj = nmat * tl.jrows, ni = nmat * tl.tx;
  T *p = lds;
  T *sca = p; p += 3 * nmat;
  T *sk0 = p; p += nk;
  T *sk1 = p; p += nk;
  T *qk0 = p; p += nk;
  T *qk1 = p; p += nk;
  T *sj0 = p; p += nj;
  T *sj1 = p; p += nj;
  T *qj0 = p; p += nj;
  T *qj1 = p; p += nj;
  T *si0 = p; p += ni;
  T *si1 = p; p += ni;
  T *qi0 = p; p += ni;
  T *qi1 = p;
  const int nth = blockDim.x;
  for (int q = threadIdx.x; q < 3 * nmat; q += nth) sca[q] = ca[(s.ca_base + q / nmat) * nmat + q % nmat];
  if (cell.js == 0 || tl.ty == 0) {
    // each k slot is written by the thread that owns it (row 0 of a 2-D tile)
    const int n = min(cell.k, s.nk1 - 1);
    for (int m = 0; m < nmat; ++m) {
      const int d = m * tl.krows + cell.ks, o = m * s.nk1 + n;
      sk0[d] = s.k0[o];
      sk1[d] = s.k1[o];
      if (psi_k) {
        qk0[d] = s.pk0[o];
        qk1[d] = s.pk1[o];
      }
    }
  }
  for (int q = threadIdx.x; q < nj; q += nth) {
    const int o = (q / tl.jrows) * s.nj1 + min(cell.j_first + q % tl.jrows, s.nj1 - 1);
    sj0[q] = s.j0[o];
    sj1[q] = s.j1[o];
    if (psi_j) {
      qj0[q] = s.pj0[o];
      qj1[q] = s.pj1[o];
    }
  }
  for (int q = threadIdx.x; q < ni; q += nth) {
    const int o = (q / tl.tx) * s.ni1 + min(i0 + q % tl.tx, s.ni1 - 1);
    si0[q] = s.i0[o];
    si1[q] = s.i1[o];
    if (psi_i) {
      qi0[q] = s.pi0[o];
      qi1[q] = s.pi1[o];
    }
  }
  __syncthreads();
  TileTables<T> tt;
  tt.ca = sca;
  tt.tk0 = sk0; tt.tk1 = sk1; tt.tj0 = sj0; tt.tj1 = sj1; tt.ti0 = si0; tt.ti1 = si1;
  tt.pk0 = qk0; tt.pk1 = qk1; tt.pj0 = qj0; tt.pj1 = qj1; tt.pi0 = qi0; tt.pi1 = qi1;
  tt.sk = tl.krows; tt.sj = tl.jrows; tt.si = tl.tx;
  tt.ks = cell.ks; tt.js = cell.js; tt.i0 = i0;
  return tt;
}

// the same view over the global tables (no staging)
template <typename T>
__device__ __forceinline__ TileTables<T> global_tables(const T *ca, int nmat, const TableSrc<T> &s,
                                                      const Cell &cell) {
  TileTables<T> tt;
  tt.ca = ca + s.ca_base * nmat;
  tt.tk0 = s.k0; tt.tk1 = s.k1; tt.tj0 = s.j0; tt.tj1 = s.j1; tt.ti0 = s.i0; tt.ti1 = s.i1;
  tt.pk0 = s.pk0; tt.pk1 = s.pk1; tt.pj0 = s.pj0; tt.pj1 = s.pj1; tt.pi0 = s.pi0; tt.pi1 = s.pi1;
  tt.sk = s.nk1; tt.sj = s.nj1; tt.si = s.ni1;
  tt.ks = min(cell.k, s.nk1 - 1); tt.js = min(cell.j, s.nj1 - 1); tt.i0 = 0;
  return tt;
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
__device__ __forceinline__ ColCpml<T> col_cpml(const Grid<T> &g, const T *const *b, const T *const *a,
                                               int ly, int lz, int j, int k) {
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

// ---------------------------------------------------------------------------
// E half-step: Simulation::updateE (simulation_update.cpp:170-196) fused with
// PML::updateE of every face (perfect_match_layer.cpp:574-718), faces in x,
// y, z order.  One thread owns a (j,k) column and marches over TX x-planes.
// A plane's loads -- the six field values, the material byte, the four
// in-plane neighbours and, inside a CPML slab, the psi values of the cell --
// are issued together, so a plane costs one memory round trip.  With PF two
// plane buffers (A, B) alternate so that the loads of plane i+1 are in flight
// while plane i is computed.
template <typename T>
struct PlaneE {
  unsigned m;
  T ex, ey, ez, hx, hy, hz, hyk, hxk, hzj, hxj;
  int lx;     // layer along x (-1 outside the x slabs)
  T bx, ax;   // its CPML b/a (wave-uniform)
  T ps[6];    // psi values of the cell, by P_* slot
};

template <typename T>
struct ColE {
  T *__restrict__ Ex;
  T *__restrict__ Ey;
  T *__restrict__ Ez;
  const T *__restrict__ Hx;
  const T *__restrict__ Hy;
  const T *__restrict__ Hz;
  const uint8_t *__restrict__ mat;
  unsigned bsj;  // byte stride of j
  bool k1, j1;   // k >= 1, j >= 1
  ColCpml<T> cp;
};

template <typename T>
__device__ __forceinline__ void load_plane_e(const Grid<T> &g, const ColE<T> &q, long long c, int i, int j,
                                             int k, PlaneE<T> &p) {
  const unsigned cell = (unsigned)c, o = cell * (unsigned)sizeof(T);
  p.m = ldm(q.mat, cell);
  p.ex = ldo(q.Ex, o);
  p.ey = ldo(q.Ey, o);
  p.ez = ldo(q.Ez, o);
  p.hx = ldo(q.Hx, o);
  p.hy = ldo(q.Hy, o);
  p.hz = ldo(q.Hz, o);
  p.hyk = q.k1 ? ldo(q.Hy, o - (unsigned)sizeof(T)) : T(0);
  p.hxk = q.k1 ? ldo(q.Hx, o - (unsigned)sizeof(T)) : T(0);
  p.hzj = q.j1 ? ldo(q.Hz, o - q.bsj) : T(0);
  p.hxj = q.j1 ? ldo(q.Hx, o - q.bsj) : T(0);
  p.lx = e_layer(i, g.nx, g.dn[0], g.dp[0]);
  p.bx = p.ax = T(0);
  for (int s = 0; s < 6; ++s) p.ps[s] = T(0);
  const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx;
  if (lx < 0 && ly < 0 && lz < 0) return;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = g.ny + 1;
  if (lx >= 0) {
    p.bx = g.be[0][i];
    p.ax = g.ae[0][i];
  }
  if (i < g.nx) {
    if (ly >= 0) p.ps[P_AY] = g.psi_e[P_AY][psi_y(i, ly, k, Ly, g.pz)];
    if (lz >= 0) p.ps[P_AZ] = g.psi_e[P_AZ][psi_z(i, j, lz, nm1y, Lz)];
  }
  if (j < g.ny) {
    if (lx >= 0) p.ps[P_BX] = g.psi_e[P_BX][psi_x(lx, j, k, nm1y, g.pz)];
    if (lz >= 0) p.ps[P_BZ] = g.psi_e[P_BZ][psi_z(i, j, lz, nm1y, Lz)];
  }
  if (k < g.nz) {
    if (lx >= 0) p.ps[P_CX] = g.psi_e[P_CX][psi_x(lx, j, k, nm1y, g.pz)];
    if (ly >= 0) p.ps[P_CY] = g.psi_e[P_CY][psi_y(i, ly, k, Ly, g.pz)];
  }
}

template <typename T>
__device__ __forceinline__ void compute_plane_e(const Grid<T> &g, const TileTables<T> &tt, const ColE<T> &q,
                                                const PlaneE<T> &p, int i, int j, int k, long long c,
                                                bool jin, bool kin, T hz_prev, T hy_prev) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int nm1y = ny + 1;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const int m = (int)p.m;
  const int lx = p.lx, ly = q.cp.ly, lz = q.cp.lz;
  const T dhy_k = p.hy - p.hyk;     // Hy(k) - Hy(k-1)
  const T dhz_j = p.hz - p.hzj;     // Hz(j) - Hz(j-1)
  const T dhx_k = p.hx - p.hxk;     // Hx(k) - Hx(k-1)
  const T dhz_i = p.hz - hz_prev;   // Hz(i) - Hz(i-1)
  const T dhx_j = p.hx - p.hxj;     // Hx(j) - Hx(j-1)
  const T dhy_i = p.hy - hy_prev;   // Hy(i) - Hy(i-1)
  const int mk = m * tt.sk + tt.ks, mj = m * tt.sj + tt.js, mi = m * tt.si + (i - tt.i0);
  const unsigned o = (unsigned)c * (unsigned)sizeof(T);
  // Yee terms for all three components, evaluated unconditionally so the
  // plane's loads are consumed on every path (one wait per plane)
  const T yx = tt.ca[m] * p.ex + tt.tk0[mk] * dhy_k + tt.tj0[mj] * dhz_j;
  const T yy = tt.ca[nmat + m] * p.ey + tt.tk1[mk] * dhx_k + tt.ti0[mi] * dhz_i;
  const T yz = tt.ca[2 * nmat + m] * p.ez + tt.tj1[mj] * dhx_j + tt.ti1[mi] * dhy_i;
  const bool ix = i >= 1 && i < nx;
  // Ex: shape (nx, ny+1, nz+1); Yee j in [1,ny), k in [1,nz)
  {
    const bool yee = (i < nx) && jin && kin;
    T v = yee ? yx : p.ex;
    const bool slab = (i < nx) && (ly >= 0 || lz >= 0);
    if (slab) {
      if (ly >= 0) {
        const T ps = q.cp.by * p.ps[P_AY] + q.cp.ay * dhz_j;
        g.psi_e[P_AY][psi_y(i, ly, k, Ly, g.pz)] = ps;
        v = v + tt.pj0[mj] * ps;
      }
      if (lz >= 0) {
        const T ps = q.cp.bz * p.ps[P_AZ] + q.cp.az * dhy_k;
        g.psi_e[P_AZ][psi_z(i, j, lz, nm1y, Lz)] = ps;
        v = v + tt.pk0[mk] * ps;
      }
    }
    if (yee || slab) sto(q.Ex, o, v);
  }
  // Ey: shape (nx+1, ny, nz+1); Yee k in [1,nz), i in [1,nx)
  {
    const bool yee = ix && kin && (j < ny);
    T v = yee ? yy : p.ey;
    const bool slab = (j < ny) && (lx >= 0 || lz >= 0);
    if (slab) {
      if (lx >= 0) {
        const T ps = p.bx * p.ps[P_BX] + p.ax * dhz_i;
        g.psi_e[P_BX][psi_x(lx, j, k, nm1y, g.pz)] = ps;
        v = v + tt.pi0[mi] * ps;
      }
      if (lz >= 0) {
        const T ps = q.cp.bz * p.ps[P_BZ] + q.cp.az * dhx_k;
        g.psi_e[P_BZ][psi_z(i, j, lz, nm1y, Lz)] = ps;
        v = v + tt.pk1[mk] * ps;
      }
    }
    if (yee || slab) sto(q.Ey, o, v);
  }
  // Ez: shape (nx+1, ny+1, nz); Yee i in [1,nx), j in [1,ny)
  {
    const bool yee = ix && jin && (k < nz);
    T v = yee ? yz : p.ez;
    const bool slab = (k < nz) && (lx >= 0 || ly >= 0);
    if (slab) {
      if (lx >= 0) {
        const T ps = p.bx * p.ps[P_CX] + p.ax * dhy_i;
        g.psi_e[P_CX][psi_x(lx, j, k, nm1y, g.pz)] = ps;
        v = v + tt.pi1[mi] * ps;
      }
      if (ly >= 0) {
        const T ps = q.cp.by * p.ps[P_CY] + q.cp.ay * dhx_j;
        g.psi_e[P_CY][psi_y(i, ly, k, Ly, g.pz)] = ps;
        v = v + tt.pj1[mj] * ps;
      }
    }
    if (yee || slab) sto(q.Ez, o, v);
  }
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

template <typename T, bool LDS, bool PF>
__global__ void __launch_bounds__(256) k_update_e(Grid<T> g, Tiling tl) {
  extern __shared__ unsigned char lds_raw[];
  int bz, by, bx;
  if (!tile_of_block(tl, bz, by, bx)) return;
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  const bool live = k <= nz && j <= ny;
  ColE<T> q;
  q.Ex = g.f[EX]; q.Ey = g.f[EY]; q.Ez = g.f[EZ];
  q.Hx = g.f[HX]; q.Hy = g.f[HY]; q.Hz = g.f[HZ];
  q.mat = g.mat;
  q.bsj = (unsigned)(g.sj * (long long)sizeof(T));
  q.k1 = k >= 1;
  q.j1 = j >= 1;
  const bool kin = (k >= 1 && k < nz), jin = (j >= 1 && j < ny);
  long long c = (long long)i0 * g.si + (long long)j * g.sj + k;
  // x-neighbours are carried in registers: Hz(i-1), Hy(i-1)
  T hz_prev = T(0), hy_prev = T(0);
  PlaneE<T> A, B;
  // the first plane's loads go out before the tables are staged
  if (live) {
    q.cp = col_cpml(g, g.be, g.ae, e_layer(j, ny, g.dn[1], g.dp[1]), e_layer(k, nz, g.dn[2], g.dp[2]), j, k);
    if (i0 >= 1) {
      hz_prev = ldo(q.Hz, (unsigned)(c - g.si) * (unsigned)sizeof(T));
      hy_prev = ldo(q.Hy, (unsigned)(c - g.si) * (unsigned)sizeof(T));
    }
    load_plane_e(g, q, c, i0, j, k, A);
  }
  TableSrc<T> src;
  src.ca_base = EX;
  src.k0 = g.cb[CB_EXHY]; src.k1 = g.cb[CB_EYHX];
  src.j0 = g.cb[CB_EXHZ]; src.j1 = g.cb[CB_EZHX];
  src.i0 = g.cb[CB_EYHZ]; src.i1 = g.cb[CB_EZHY];
  src.pk0 = g.cpsi[CB_EXHY]; src.pk1 = g.cpsi[CB_EYHX];
  src.pj0 = g.cpsi[CB_EXHZ]; src.pj1 = g.cpsi[CB_EZHX];
  src.pi0 = g.cpsi[CB_EYHZ]; src.pi1 = g.cpsi[CB_EZHY];
  src.nk1 = nz + 1; src.nj1 = ny + 1; src.ni1 = nx + 1;
  TileTables<T> tt;
  if (LDS) {
    bool sk, sj, si;
    tile_slabs(g, tl, bz, cell, i0, i1, sk, sj, si);
    tt = stage_tables<T>(reinterpret_cast<T *>(lds_raw), g.ca, nmat, src, tl, cell, i0, sk, sj, si);
  } else {
    tt = global_tables<T>(g.ca, nmat, src, cell);
  }
  if (!live) return;
  if (!PF) {
    for (int i = i0;;) {
      compute_plane_e(g, tt, q, A, i, j, k, c, jin, kin, hz_prev, hy_prev);
      hz_prev = A.hz;
      hy_prev = A.hy;
      if (++i >= i1) break;
      c += g.si;
      load_plane_e(g, q, c, i, j, k, A);
    }
    return;
  }
  for (int i = i0; i < i1; i += 2, c += 2 * g.si) {
    if (i + 1 < i1) load_plane_e(g, q, c + g.si, i + 1, j, k, B);
    compute_plane_e(g, tt, q, A, i, j, k, c, jin, kin, hz_prev, hy_prev);
    hz_prev = A.hz;
    hy_prev = A.hy;
    if (i + 1 >= i1) break;
    if (i + 2 < i1) load_plane_e(g, q, c + 2 * g.si, i + 2, j, k, A);
    compute_plane_e(g, tt, q, B, i + 1, j, k, c + g.si, jin, kin, hz_prev, hy_prev);
    hz_prev = B.hz;
    hy_prev = B.hy;
  }
}

// ---------------------------------------------------------------------------
// H half-step: Simulation::updateH (simulation_update.cpp:84-100) fused with
// PML::updateH (perfect_match_layer.cpp:434-572).  Same plane scheme as the
// E kernel; Ez(i), Ey(i) are the (i+1) values loaded for the previous plane.
template <typename T>
struct PlaneH {
  unsigned m;
  T hx, hy, hz, ex, eyk, exk, ezj, exj, ezi, eyi;
  int lx;
  T bx, ax;
  T ps[6];
};

template <typename T>
struct ColH {
  const T *__restrict__ Ex;
  const T *__restrict__ Ey;
  const T *__restrict__ Ez;
  T *__restrict__ Hx;
  T *__restrict__ Hy;
  T *__restrict__ Hz;
  const uint8_t *__restrict__ mat;
  unsigned bsj, bsi;  // byte strides of j and i
  bool kin, jin;      // k < nz, j < ny
  ColCpml<T> cp;
};

template <typename T>
__device__ __forceinline__ void load_plane_h(const Grid<T> &g, const ColH<T> &q, long long c, int i, int j,
                                             int k, PlaneH<T> &p) {
  const unsigned cell = (unsigned)c, o = cell * (unsigned)sizeof(T);
  p.m = ldm(q.mat, cell);
  p.hx = ldo(q.Hx, o);
  p.hy = ldo(q.Hy, o);
  p.hz = ldo(q.Hz, o);
  p.ex = ldo(q.Ex, o);
  p.eyk = q.kin ? ldo(q.Ey, o + (unsigned)sizeof(T)) : T(0);  // Ey(k+1)
  p.exk = q.kin ? ldo(q.Ex, o + (unsigned)sizeof(T)) : T(0);  // Ex(k+1)
  p.ezj = q.jin ? ldo(q.Ez, o + q.bsj) : T(0);                // Ez(j+1)
  p.exj = q.jin ? ldo(q.Ex, o + q.bsj) : T(0);                // Ex(j+1)
  p.ezi = (i < g.nx) ? ldo(q.Ez, o + q.bsi) : T(0);           // Ez(i+1)
  p.eyi = (i < g.nx) ? ldo(q.Ey, o + q.bsi) : T(0);           // Ey(i+1)
  p.lx = h_layer(i, g.nx, g.dn[0], g.dp[0]);
  p.bx = p.ax = T(0);
  for (int s = 0; s < 6; ++s) p.ps[s] = T(0);
  const int ly = q.cp.ly, lz = q.cp.lz, lx = p.lx;
  if (lx < 0 && ly < 0 && lz < 0) return;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2], nm1y = g.ny + 1;
  if (lx >= 0) {
    p.bx = g.bm[0][i];
    p.ax = g.am[0][i];
  }
  if (q.jin && q.kin) {
    if (ly >= 0) p.ps[P_AY] = g.psi_h[P_AY][psi_y(i, ly, k, Ly, g.pz)];
    if (lz >= 0) p.ps[P_AZ] = g.psi_h[P_AZ][psi_z(i, j, lz, nm1y, Lz)];
  }
  if (i < g.nx && q.kin) {
    if (lx >= 0) p.ps[P_BX] = g.psi_h[P_BX][psi_x(lx, j, k, nm1y, g.pz)];
    if (lz >= 0) p.ps[P_BZ] = g.psi_h[P_BZ][psi_z(i, j, lz, nm1y, Lz)];
  }
  if (i < g.nx && q.jin) {
    if (lx >= 0) p.ps[P_CX] = g.psi_h[P_CX][psi_x(lx, j, k, nm1y, g.pz)];
    if (ly >= 0) p.ps[P_CY] = g.psi_h[P_CY][psi_y(i, ly, k, Ly, g.pz)];
  }
}

template <typename T>
__device__ __forceinline__ void compute_plane_h(const Grid<T> &g, const TileTables<T> &tt, const ColH<T> &q,
                                                const PlaneH<T> &p, int i, int j, int k, long long c,
                                                T ez_c, T ey_c) {
  const int nx = g.nx, ny = g.ny, nmat = g.nmat;
  const int nm1y = ny + 1;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const bool kin = q.kin, jin = q.jin;
  const int m = (int)p.m;
  const int lx = p.lx, ly = q.cp.ly, lz = q.cp.lz;
  const T dey_k = p.eyk - ey_c;  // Ey(k+1) - Ey(k)
  const T dez_j = p.ezj - ez_c;  // Ez(j+1) - Ez(j)
  const T dez_i = p.ezi - ez_c;  // Ez(i+1) - Ez(i)
  const T dex_k = p.exk - p.ex;  // Ex(k+1) - Ex(k)
  const T dex_j = p.exj - p.ex;  // Ex(j+1) - Ex(j)
  const T dey_i = p.eyi - ey_c;  // Ey(i+1) - Ey(i)
  const int mk = m * tt.sk + tt.ks, mj = m * tt.sj + tt.js, mi = m * tt.si + (i - tt.i0);
  const unsigned o = (unsigned)c * (unsigned)sizeof(T);
  const T yx = tt.ca[m] * p.hx + tt.tk0[mk] * dey_k + tt.tj0[mj] * dez_j;
  const T yy = tt.ca[nmat + m] * p.hy + tt.ti0[mi] * dez_i + tt.tk1[mk] * dex_k;
  const T yz = tt.ca[2 * nmat + m] * p.hz + tt.tj1[mj] * dex_j + tt.ti1[mi] * dey_i;
  // Hx: shape (nx+1, ny, nz); Yee i < nx
  {
    const bool in = jin && kin;
    const bool yee = in && (i < nx);
    T v = yee ? yx : p.hx;
    const bool slab = in && (ly >= 0 || lz >= 0);
    if (slab) {
      if (ly >= 0) {
        const T ps = q.cp.by * p.ps[P_AY] + q.cp.ay * dez_j;
        g.psi_h[P_AY][psi_y(i, ly, k, Ly, g.pz)] = ps;
        v = v + tt.pj0[mj] * ps;
      }
      if (lz >= 0) {
        const T ps = q.cp.bz * p.ps[P_AZ] + q.cp.az * dey_k;
        g.psi_h[P_AZ][psi_z(i, j, lz, nm1y, Lz)] = ps;
        v = v + tt.pk0[mk] * ps;
      }
    }
    if (yee || slab) sto(q.Hx, o, v);
  }
  // Hy: shape (nx, ny+1, nz); Yee j < ny
  {
    const bool in = (i < nx) && kin;
    const bool yee = in && jin;
    T v = yee ? yy : p.hy;
    const bool slab = in && (lx >= 0 || lz >= 0);
    if (slab) {
      if (lx >= 0) {
        const T ps = p.bx * p.ps[P_BX] + p.ax * dez_i;
        g.psi_h[P_BX][psi_x(lx, j, k, nm1y, g.pz)] = ps;
        v = v + tt.pi0[mi] * ps;
      }
      if (lz >= 0) {
        const T ps = q.cp.bz * p.ps[P_BZ] + q.cp.az * dex_k;
        g.psi_h[P_BZ][psi_z(i, j, lz, nm1y, Lz)] = ps;
        v = v + tt.pk1[mk] * ps;
      }
    }
    if (yee || slab) sto(q.Hy, o, v);
  }
  // Hz: shape (nx, ny, nz+1); Yee k < nz
  {
    const bool in = (i < nx) && jin;
    const bool yee = in && kin;
    T v = yee ? yz : p.hz;
    const bool slab = in && (lx >= 0 || ly >= 0);
    if (slab) {
      if (lx >= 0) {
        const T ps = p.bx * p.ps[P_CX] + p.ax * dey_i;
        g.psi_h[P_CX][psi_x(lx, j, k, nm1y, g.pz)] = ps;
        v = v + tt.pi1[mi] * ps;
      }
      if (ly >= 0) {
        const T ps = q.cp.by * p.ps[P_CY] + q.cp.ay * dex_j;
        g.psi_h[P_CY][psi_y(i, ly, k, Ly, g.pz)] = ps;
        v = v + tt.pj1[mj] * ps;
      }
    }
    if (yee || slab) sto(q.Hz, o, v);
  }
}

template <typename T, bool LDS, bool PF>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, Tiling tl) {
  extern __shared__ unsigned char lds_raw[];
  int bz, by, bx;
  if (!tile_of_block(tl, bz, by, bx)) return;
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  const bool live = k <= nz && j <= ny;
  ColH<T> q;
  q.Ex = g.f[EX]; q.Ey = g.f[EY]; q.Ez = g.f[EZ];
  q.Hx = g.f[HX]; q.Hy = g.f[HY]; q.Hz = g.f[HZ];
  q.mat = g.mat;
  q.bsj = (unsigned)(g.sj * (long long)sizeof(T));
  q.bsi = (unsigned)(g.si * (long long)sizeof(T));
  q.kin = k < nz;
  q.jin = j < ny;
  long long c = (long long)i0 * g.si + (long long)j * g.sj + k;
  T ez_c = T(0), ey_c = T(0);
  PlaneH<T> A, B;
  // the first plane's loads go out before the tables are staged
  if (live) {
    q.cp = col_cpml(g, g.bm, g.am, h_layer(j, ny, g.dn[1], g.dp[1]), h_layer(k, nz, g.dn[2], g.dp[2]), j, k);
    ez_c = ldo(q.Ez, (unsigned)c * (unsigned)sizeof(T));
    ey_c = ldo(q.Ey, (unsigned)c * (unsigned)sizeof(T));
    load_plane_h(g, q, c, i0, j, k, A);
  }
  TableSrc<T> src;
  src.ca_base = HX;
  src.k0 = g.cb[CB_HXEY]; src.k1 = g.cb[CB_HYEX];
  src.j0 = g.cb[CB_HXEZ]; src.j1 = g.cb[CB_HZEX];
  src.i0 = g.cb[CB_HYEZ]; src.i1 = g.cb[CB_HZEY];
  src.pk0 = g.cpsi[CB_HXEY]; src.pk1 = g.cpsi[CB_HYEX];
  src.pj0 = g.cpsi[CB_HXEZ]; src.pj1 = g.cpsi[CB_HZEX];
  src.pi0 = g.cpsi[CB_HYEZ]; src.pi1 = g.cpsi[CB_HZEY];
  src.nk1 = nz + 1; src.nj1 = ny + 1; src.ni1 = nx + 1;
  TileTables<T> tt;
  if (LDS) {
    bool sk, sj, si;
    tile_slabs(g, tl, bz, cell, i0, i1, sk, sj, si);
    tt = stage_tables<T>(reinterpret_cast<T *>(lds_raw), g.ca, nmat, src, tl, cell, i0, sk, sj, si);
  } else {
    tt = global_tables<T>(g.ca, nmat, src, cell);
  }
  if (!live) return;
  if (!PF) {
    for (int i = i0;;) {
      compute_plane_h(g, tt, q, A, i, j, k, c, ez_c, ey_c);
      ez_c = A.ezi;
      ey_c = A.eyi;
      if (++i >= i1) break;
      c += g.si;
      load_plane_h(g, q, c, i, j, k, A);
    }
    return;
  }
  for (int i = i0; i < i1; i += 2, c += 2 * g.si) {
    if (i + 1 < i1) load_plane_h(g, q, c + g.si, i + 1, j, k, B);
    compute_plane_h(g, tt, q, A, i, j, k, c, ez_c, ey_c);
    ez_c = A.ezi;
    ey_c = A.eyi;
    if (i + 1 >= i1) break;
    if (i + 2 < i1) load_plane_h(g, q, c + 2 * g.si, i + 2, j, k, A);
    compute_plane_h(g, tt, q, B, i + 1, j, k, c + g.si, ez_c, ey_c);
    ez_c = B.ezi;
    ey_c = B.eyi;
  }
}
// ---------------------------------------------------------------------------
// TFSF (tfsf/tfsf_3d.cpp)
template <typename T>
struct Tfsf {
  int li, ri, lj, rj, lk, rk, L;
  const T *pf[3], *ph[3];
  T te[3], th[3];
  T ceie, ceihi, chih, chiei, c1, c2;
  T ca[3], cb[3];
  T *e_inc, *h_inc, *ab;
  const T *src;
};

// TFSF3D::getIncident{Ex..Hz} (tfsf_3d.cpp:413-483)
template <typename T>
__device__ __forceinline__ T tfsf_inc(const Tfsf<T> &t, int comp, int i, int j, int k) {
  T proj, s;
  const T *arr;
  switch (comp) {
    case EX: proj = t.ph[0][i - t.li + 1] + t.pf[1][j - t.lj] + t.pf[2][k - t.lk];
             arr = t.e_inc; s = t.te[0]; break;
    case EY: proj = t.pf[0][i - t.li] + t.ph[1][j - t.lj + 1] + t.pf[2][k - t.lk];
             arr = t.e_inc; s = t.te[1]; break;
    case EZ: proj = t.pf[0][i - t.li] + t.pf[1][j - t.lj] + t.ph[2][k - t.lk + 1];
             arr = t.e_inc; s = t.te[2]; break;
    case HX: proj = t.pf[0][i - t.li] + t.ph[1][j - t.lj + 1] + t.ph[2][k - t.lk + 1] - T(0.5);
             arr = t.h_inc; s = t.th[0]; break;
    case HY: proj = t.ph[0][i - t.li + 1] + t.pf[1][j - t.lj] + t.ph[2][k - t.lk + 1] - T(0.5);
             arr = t.h_inc; s = t.th[1]; break;
    default: proj = t.ph[0][i - t.li + 1] + t.ph[1][j - t.lj + 1] + t.pf[2][k - t.lk] - T(0.5);
             arr = t.h_inc; s = t.th[2]; break;
  }
  const long long idx = (long long)proj;
  const T w = proj - (T)idx;
  return (T(1) - w) * (s * arr[idx]) + w * (s * arr[idx + 1]);
}

// TFSF3D::updateIncidentField (tfsf_3d.cpp:153-183): one workgroup
template <typename T>
__global__ void __launch_bounds__(1024) k_tfsf_ifa(Tfsf<T> t, long long step) {
  const int L = t.L;
  T *e = t.e_inc, *h = t.h_inc;
  const T x = e[L - 2], y = e[L - 1];
  __syncthreads();
  if (threadIdx.x == 0) e[0] = t.src[step];
  for (int i = 1 + threadIdx.x; i < L - 1; i += blockDim.x)
    e[i] = t.ceie * e[i] + t.ceihi * (h[i] - h[i - 1]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const T a = t.ab[0], b = t.ab[1];
    e[L - 1] = -a + t.c1 * (e[L - 2] + b) + t.c2 * (x + y);
    t.ab[0] = x;
    t.ab[1] = y;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < L - 1; i += blockDim.x)
    h[i] = t.chih * h[i] + t.chiei * (e[i + 1] - e[i]);
}

__device__ __forceinline__ void perimeter(int p, int A, int B, int &u, int &v) {
  if (p < A) { u = p; v = 0; return; }
  if (p < 2 * A) { u = p - A; v = B - 1; return; }
  const int q = p - 2 * A;
  if (q < B - 2) { u = 0; v = 1 + q; return; }
  u = A - 1; v = 1 + (q - (B - 2));
}

// TFSF3D::updateE (tfsf_3d.cpp:299-411).  Edge cells take two corrections;
// one thread owns each target cell and applies them in the reference's
// xn, xp, yn, yp, zn, zp order.
template <typename T>
__global__ void k_tfsf_e(Grid<T> g, Tfsf<T> t, int nEy, int nEz, int nEx) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int li = t.li, ri = t.ri, lj = t.lj, rj = t.rj, lk = t.lk, rk = t.rk;
  if (q < nEy) {
    const int A = ri - li + 1, B = rk - lk + 1, P = 2 * A + 2 * (B - 2);
    const int j = lj + q / P;
    int u, v;
    perimeter(q % P, A, B, u, v);
    const int i = li + u, k = lk + v;
    T *p = &g.f[EY][(long long)i * g.si + (long long)j * g.sj + k];
    T val = *p;
    if (i == li) val = val + t.ca[0] * tfsf_inc(t, HZ, li - 1, j, k);
    if (i == ri) val = val - t.ca[0] * tfsf_inc(t, HZ, ri, j, k);
    if (k == lk) val = val - t.ca[2] * tfsf_inc(t, HX, i, j, lk - 1);
    if (k == rk) val = val + t.ca[2] * tfsf_inc(t, HX, i, j, rk);
    *p = val;
    return;
  }
  q -= nEy;
  if (q < nEz) {
    const int A = ri - li + 1, B = rj - lj + 1, P = 2 * A + 2 * (B - 2);
    const int k = lk + q / P;
    int u, v;
    perimeter(q % P, A, B, u, v);
    const int i = li + u, j = lj + v;
    T *p = &g.f[EZ][(long long)i * g.si + (long long)j * g.sj + k];
    T val = *p;
    if (i == li) val = val - t.ca[0] * tfsf_inc(t, HY, li - 1, j, k);
    if (i == ri) val = val + t.ca[0] * tfsf_inc(t, HY, ri, j, k);
    if (j == lj) val = val + t.ca[1] * tfsf_inc(t, HX, i, lj - 1, k);
    if (j == rj) val = val - t.ca[1] * tfsf_inc(t, HX, i, rj, k);
    *p = val;
    return;
  }
  q -= nEz;
  if (q < nEx) {
    const int A = rj - lj + 1, B = rk - lk + 1, P = 2 * A + 2 * (B - 2);
    const int i = li + q / P;
    int u, v;
    perimeter(q % P, A, B, u, v);
    const int j = lj + u, k = lk + v;
    T *p = &g.f[EX][(long long)i * g.si + (long long)j * g.sj + k];
    T val = *p;
    if (j == lj) val = val - t.ca[1] * tfsf_inc(t, HZ, i, lj - 1, k);
    if (j == rj) val = val + t.ca[1] * tfsf_inc(t, HZ, i, rj, k);
    if (k == lk) val = val + t.ca[2] * tfsf_inc(t, HY, i, j, lk - 1);
    if (k == rk) val = val - t.ca[2] * tfsf_inc(t, HY, i, j, rk);
    *p = val;
  }
}

// TFSF3D::updateH (tfsf_3d.cpp:185-297).  No H cell takes two corrections.
struct TfsfHSeg {
  int off[13];  // prefix offsets of the 12 face loops
};
template <typename T>
__global__ void k_tfsf_h(Grid<T> g, Tfsf<T> t, TfsfHSeg s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= s.off[12]) return;
  int seg = 0;
  while (q >= s.off[seg + 1]) ++seg;
  const int r = q - s.off[seg];
  const int li = t.li, ri = t.ri, lj = t.lj, rj = t.rj, lk = t.lk, rk = t.rk;
  int i = 0, j = 0, k = 0, comp = HX, src = EX;
  T coef = 0;
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
    case 4: { const int ni = ri + 1 - li; k = lk + r / ni; i = li + r % ni; j = lj - 1;
              comp = HX; src = EZ; coef = t.cb[1]; add = true; break; }   // yn hx += cby*ezi(lj)
    case 5: { const int ni = ri - li; k = lk + r / ni; i = li + r % ni; j = lj - 1;
              comp = HZ; src = EX; coef = t.cb[1]; add = false; break; }
    case 6: { const int ni = ri + 1 - li; k = lk + r / ni; i = li + r % ni; j = rj;
              comp = HX; src = EZ; coef = t.cb[1]; add = false; break; }
    case 7: { const int ni = ri - li; k = lk + r / ni; i = li + r % ni; j = rj;
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
  const T inc = tfsf_inc(t, src, si_, sj_, sk_);
  T *p = &g.f[comp][(long long)i * g.si + (long long)j * g.sj + k];
  *p = add ? (*p + coef * inc) : (*p - coef * inc);
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

// point probes (monitor update slot, simulation_update.cpp:215-219)
template <typename T>
__global__ void k_probe(Grid<T> g, const int *comp, const long long *cell, int n,
                        double *buf, long long T_total, long long step) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  buf[(long long)q * T_total + step] = (double)g.f[comp[q]][cell[q]];
}

// ---------------------------------------------------------------------------
// NffftFd::update (nffft_fd.cpp:108-203) with the face-centre stencils of
// linear_interpolation_emf.cpp:286-388.  One thread per face cell of a face
// pair; FP64 accumulators in every precision.
struct Nffft {
  int li, ri, lj, rj, lk, rk, nf;
  int nxb, nyb, nzb;
  long long T_total;
  const double2 *tj, *tm;  // [nf][T]
  double2 *acc;
  long long base[3];       // accumulator offset of each face pair
};

__device__ __forceinline__ void acc_add(double2 *a, double v, double2 tw) {
  double2 x = *a;
  x.x = x.x + v * tw.x;
  x.y = x.y + v * tw.y;
  *a = x;
}

template <typename T>
__global__ void k_nffft(Grid<T> g, Nffft n, long long step) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int cx = n.nyb * n.nzb, cy = n.nzb * n.nxb, cz = n.nxb * n.nyb;
  const T *Ex = g.f[EX], *Ey = g.f[EY], *Ez = g.f[EZ];
  const T *Hx = g.f[HX], *Hy = g.f[HY], *Hz = g.f[HZ];
  const long long si = g.si, sj = g.sj;
#define C3(i, j, k) ((long long)(i) * si + (long long)(j) * sj + (k))
  T v[8];
  int pair, cells, cell;
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
    const int k = n.lk + q / n.nxb, i = n.li + q % n.nxb;
    const int lj = n.lj, rj = n.rj;
    v[0] = T(-1) * (T(0.5) * (Ex[C3(i, lj, k)] + Ex[C3(i, lj + 1, k)]));
    v[1] = T(0.5) * (Ez[C3(i, lj, k)] + Ez[C3(i + 1, lj, k)]);
    v[2] = T(0.25) * (Hx[C3(i + 1, lj, k)] + Hx[C3(i, lj, k)] + Hx[C3(i + 1, lj - 1, k)] + Hx[C3(i, lj - 1, k)]);
    v[3] = T(-1) * (T(0.25) * (Hz[C3(i, lj, k)] + Hz[C3(i, lj, k + 1)] + Hz[C3(i, lj - 1, k)] + Hz[C3(i, lj - 1, k + 1)]));
    v[4] = T(0.5) * (Ex[C3(i, rj, k)] + Ex[C3(i, rj + 1, k)]);
    v[5] = T(-1) * (T(0.5) * (Ez[C3(i, rj, k)] + Ez[C3(i + 1, rj, k)]));
    v[6] = T(-1) * (T(0.25) * (Hx[C3(i + 1, rj, k)] + Hx[C3(i, rj, k)] + Hx[C3(i + 1, rj - 1, k)] + Hx[C3(i, rj - 1, k)]));
    v[7] = T(0.25) * (Hz[C3(i, rj, k)] + Hz[C3(i, rj, k + 1)] + Hz[C3(i, rj - 1, k)] + Hz[C3(i, rj - 1, k + 1)]);
  } else if ((q -= cy) < cz) {
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
    return;
  }
#undef C3
  for (int f = 0; f < n.nf; ++f) {
    const double2 twm = n.tm[(long long)f * n.T_total + step];
    const double2 twj = n.tj[(long long)f * n.T_total + step];
    for (int c = 0; c < 8; ++c) {
      double2 *a = &n.acc[n.base[pair] + ((long long)c * n.nf + f) * cells + cell];
      acc_add(a, (double)v[c], (c % 4) < 2 ? twm : twj);
    }
  }
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
        const int k = n.lk + q / n.nxb, i = n.li + q % n.nxb;
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

}  // namespace xfk
