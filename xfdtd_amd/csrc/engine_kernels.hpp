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
// consecutive workgroup ids to the 8 XCDs round-robin; tile ids are remapped
// so that each XCD works on one contiguous run of tiles, which keeps a tile's
// y-1 neighbour row in the same XCD's L2.
struct Tiling {
  int tz, ty, tx;           // threads along k, j (ty = 0: flat); planes along i
  int ntz, nty, ntx;        // tile counts
  int nblocks;
  int krows, jrows;         // coefficient-table slots of a tile along k and j
  int x_fastest;            // tile order: x chunks first (else z, y, x)
};

__device__ __forceinline__ void tile_of_block(const Tiling &t, int &bz, int &by, int &bx) {
  const int b = blockIdx.x;
  const int q = t.nblocks >> 3, r = t.nblocks & 7;
  const int xcd = b & 7, local = b >> 3;
  const int id = xcd * q + min(xcd, r) + local;
  if (t.x_fastest) {
    bx = id % t.ntx;
    const int rest = id / t.ntx;
    bz = rest % t.ntz;
    by = rest / t.ntz;
    return;
  }
  bz = id % t.ntz;
  const int rest = id / t.ntz;
  by = rest % t.nty;
  bx = rest / t.nty;
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
  const T *ca;    // [3][nmat]
  const T *tk0, *tk1;  // [nmat][krows]
  const T *tj0, *tj1;  // [nmat][jrows]
  const T *ti0, *ti1;  // [nmat][TX]
  int sk, sj, si;      // entries per material of the k/j/i tables
  int ks, js, i0;      // this thread's k and j slots, tile's first plane
};

template <typename T>
__host__ __device__ inline size_t lds_table_bytes(int nmat, const Tiling &tl) {
  return sizeof(T) * (size_t)nmat * (3 + 2 * tl.krows + 2 * tl.jrows + 2 * tl.tx);
}

// stage the tables of one tile (cooperative; ends with a barrier).
// self: ca entries ca_base..ca_base+2; k/j/i: global tables with lengths
// (n+1) per material.
template <typename T>
__device__ __forceinline__ TileTables<T> stage_tables(T *lds, const T *ca, int ca_base, int nmat,
                                                     const T *gk0, const T *gk1, int nk1,
                                                     const T *gj0, const T *gj1, int nj1,
                                                     const T *gi0, const T *gi1, int ni1,
                                                     const Tiling &tl, const Cell &cell, int i0) {
  TileTables<T> tt;
  T *p = lds;
  T *sca = p; p += 3 * nmat;
  T *sk0 = p; p += nmat * tl.krows;
  T *sk1 = p; p += nmat * tl.krows;
  T *sj0 = p; p += nmat * tl.jrows;
  T *sj1 = p; p += nmat * tl.jrows;
  T *si0 = p; p += nmat * tl.tx;
  T *si1 = p;
  const int nth = blockDim.x;
  for (int q = threadIdx.x; q < 3 * nmat; q += nth) sca[q] = ca[(ca_base + q / nmat) * nmat + q % nmat];
  if (cell.js == 0 || tl.ty == 0) {
    // each k slot is written by the thread that owns it (row 0 of a 2-D tile)
    const int n = min(cell.k, nk1 - 1);
    for (int m = 0; m < nmat; ++m) {
      sk0[m * tl.krows + cell.ks] = gk0[m * nk1 + n];
      sk1[m * tl.krows + cell.ks] = gk1[m * nk1 + n];
    }
  }
  for (int q = threadIdx.x; q < nmat * tl.jrows; q += nth) {
    const int m = q / tl.jrows, n = min(cell.j_first + q % tl.jrows, nj1 - 1);
    sj0[q] = gj0[m * nj1 + n];
    sj1[q] = gj1[m * nj1 + n];
  }
  for (int q = threadIdx.x; q < nmat * tl.tx; q += nth) {
    const int m = q / tl.tx, n = min(i0 + q % tl.tx, ni1 - 1);
    si0[q] = gi0[m * ni1 + n];
    si1[q] = gi1[m * ni1 + n];
  }
  __syncthreads();
  tt.ca = sca;
  tt.tk0 = sk0; tt.tk1 = sk1; tt.tj0 = sj0; tt.tj1 = sj1; tt.ti0 = si0; tt.ti1 = si1;
  tt.sk = tl.krows; tt.sj = tl.jrows; tt.si = tl.tx;
  tt.ks = cell.ks; tt.js = cell.js; tt.i0 = i0;
  return tt;
}

// the same view over the global tables (no staging)
template <typename T>
__device__ __forceinline__ TileTables<T> global_tables(const T *ca, int ca_base, int nmat,
                                                      const T *gk0, const T *gk1, int nk1,
                                                      const T *gj0, const T *gj1, int nj1,
                                                      const T *gi0, const T *gi1, int ni1,
                                                      const Cell &cell) {
  TileTables<T> tt;
  tt.ca = ca + ca_base * nmat;
  tt.tk0 = gk0; tt.tk1 = gk1; tt.tj0 = gj0; tt.tj1 = gj1; tt.ti0 = gi0; tt.ti1 = gi1;
  tt.sk = nk1; tt.sj = nj1; tt.si = ni1;
  tt.ks = min(cell.k, nk1 - 1); tt.js = min(cell.j, nj1 - 1); tt.i0 = 0;
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

// ---------------------------------------------------------------------------
// E half-step: Simulation::updateE (simulation_update.cpp:170-196) fused with
// PML::updateE of every face (perfect_match_layer.cpp:574-718), faces in x,
// y, z order.  One thread owns a (j,k) column and marches over TX x-planes.
// Two plane buffers (A, B) alternate so that the loads of plane i+1 are in
// flight while plane i is computed, without register copies between them.
template <typename T>
struct PlaneE {
  unsigned m;
  T ex, ey, ez, hx, hy, hz, hyk, hxk, hzj, hxj;
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
  long long sj;
  unsigned bsj;  // byte stride of j
  bool k1, j1;   // k >= 1, j >= 1
};

template <typename T>
__device__ __forceinline__ void load_plane_e(const ColE<T> &q, long long c, PlaneE<T> &p) {
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
}

template <typename T>
__device__ __forceinline__ void compute_plane_e(const Grid<T> &g, const TileTables<T> &tt, const ColE<T> &q,
                                                const PlaneE<T> &p, int i, int j, int k, long long c,
                                                int ly, int lz, bool jin, bool kin, T hz_prev,
                                                T hy_prev) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int nm1y = ny + 1, nm1z = nz + 1, nm1x = nx + 1;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const int m = (int)p.m;
  const int lx = e_layer(i, nx, g.dn[0], g.dp[0]);
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
    if (i < nx && (ly >= 0 || lz >= 0)) {
      if (ly >= 0) {
        T *ps_ = &g.psi_e[P_AY][((long long)i * Ly + ly) * g.pz + k];
        const T ps = g.be[1][j] * (*ps_) + g.ae[1][j] * dhz_j;
        *ps_ = ps;
        v = v + g.cpsi[CB_EXHZ][m * nm1y + j] * ps;
      }
      if (lz >= 0) {
        T *ps_ = &g.psi_e[P_AZ][((long long)i * nm1y + j) * Lz + lz];
        const T ps = g.be[2][k] * (*ps_) + g.ae[2][k] * dhy_k;
        *ps_ = ps;
        v = v + g.cpsi[CB_EXHY][m * nm1z + k] * ps;
      }
    }
    if (yee || (i < nx && (ly >= 0 || lz >= 0))) sto(q.Ex, o, v);
  }
  // Ey: shape (nx+1, ny, nz+1); Yee k in [1,nz), i in [1,nx)
  {
    const bool yee = ix && kin && (j < ny);
    T v = yee ? yy : p.ey;
    const bool slab = (j < ny) && (lx >= 0 || lz >= 0);
    if (slab) {
      if (lx >= 0) {
        T *ps_ = &g.psi_e[P_BX][((long long)lx * nm1y + j) * g.pz + k];
        const T ps = g.be[0][i] * (*ps_) + g.ae[0][i] * dhz_i;
        *ps_ = ps;
        v = v + g.cpsi[CB_EYHZ][m * nm1x + i] * ps;
      }
      if (lz >= 0) {
        T *ps_ = &g.psi_e[P_BZ][((long long)i * nm1y + j) * Lz + lz];
        const T ps = g.be[2][k] * (*ps_) + g.ae[2][k] * dhx_k;
        *ps_ = ps;
        v = v + g.cpsi[CB_EYHX][m * nm1z + k] * ps;
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
        T *ps_ = &g.psi_e[P_CX][((long long)lx * nm1y + j) * g.pz + k];
        const T ps = g.be[0][i] * (*ps_) + g.ae[0][i] * dhy_i;
        *ps_ = ps;
        v = v + g.cpsi[CB_EZHY][m * nm1x + i] * ps;
      }
      if (ly >= 0) {
        T *ps_ = &g.psi_e[P_CY][((long long)i * Ly + ly) * g.pz + k];
        const T ps = g.be[1][j] * (*ps_) + g.ae[1][j] * dhx_j;
        *ps_ = ps;
        v = v + g.cpsi[CB_EZHX][m * nm1y + j] * ps;
      }
    }
    if (yee || slab) sto(q.Ez, o, v);
  }
}

template <typename T, bool LDS, bool PF>
__global__ void __launch_bounds__(256) k_update_e(Grid<T> g, Tiling tl) {
  extern __shared__ unsigned char lds_raw[];
  int bz, by, bx;
  tile_of_block(tl, bz, by, bx);
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  TileTables<T> tt;
  if (LDS)
    tt = stage_tables<T>(reinterpret_cast<T *>(lds_raw), g.ca, EX, nmat, g.cb[CB_EXHY], g.cb[CB_EYHX],
                         nz + 1, g.cb[CB_EXHZ], g.cb[CB_EZHX], ny + 1, g.cb[CB_EYHZ], g.cb[CB_EZHY],
                         nx + 1, tl, cell, i0);
  else
    tt = global_tables<T>(g.ca, EX, nmat, g.cb[CB_EXHY], g.cb[CB_EYHX], nz + 1, g.cb[CB_EXHZ],
                          g.cb[CB_EZHX], ny + 1, g.cb[CB_EYHZ], g.cb[CB_EZHY], nx + 1, cell);
  if (k > nz || j > ny) return;
  ColE<T> q;
  q.Ex = g.f[EX]; q.Ey = g.f[EY]; q.Ez = g.f[EZ];
  q.Hx = g.f[HX]; q.Hy = g.f[HY]; q.Hz = g.f[HZ];
  q.mat = g.mat;
  q.sj = g.sj;
  q.bsj = (unsigned)(g.sj * (long long)sizeof(T));
  q.k1 = k >= 1;
  q.j1 = j >= 1;
  const int ly = e_layer(j, ny, g.dn[1], g.dp[1]);
  const int lz = e_layer(k, nz, g.dn[2], g.dp[2]);
  const bool kin = (k >= 1 && k < nz), jin = (j >= 1 && j < ny);
  long long c = (long long)i0 * g.si + (long long)j * g.sj + k;
  // x-neighbours are carried in registers: Hz(i-1), Hy(i-1)
  T hz_prev = (i0 >= 1) ? ldo(q.Hz, (unsigned)(c - g.si) * (unsigned)sizeof(T)) : T(0);
  T hy_prev = (i0 >= 1) ? ldo(q.Hy, (unsigned)(c - g.si) * (unsigned)sizeof(T)) : T(0);
  if (!PF) {
    // one plane at a time: its loads are issued together, latency is hidden
    // by the other waves on the CU
    for (int i = i0; i < i1; ++i, c += g.si) {
      PlaneE<T> A;
      load_plane_e(q, c, A);
      compute_plane_e(g, tt, q, A, i, j, k, c, ly, lz, jin, kin, hz_prev, hy_prev);
      hz_prev = A.hz;
      hy_prev = A.hy;
    }
    return;
  }
  PlaneE<T> A, B;
  load_plane_e(q, c, A);
  for (int i = i0; i < i1; i += 2, c += 2 * g.si) {
    if (i + 1 < i1) load_plane_e(q, c + g.si, B);
    compute_plane_e(g, tt, q, A, i, j, k, c, ly, lz, jin, kin, hz_prev, hy_prev);
    hz_prev = A.hz;
    hy_prev = A.hy;
    if (i + 1 >= i1) break;
    if (i + 2 < i1) load_plane_e(q, c + 2 * g.si, A);
    compute_plane_e(g, tt, q, B, i + 1, j, k, c + g.si, ly, lz, jin, kin, hz_prev, hy_prev);
    hz_prev = B.hz;
    hy_prev = B.hy;
  }
}

// ---------------------------------------------------------------------------
// H half-step: Simulation::updateH (simulation_update.cpp:84-100) fused with
// PML::updateH (perfect_match_layer.cpp:434-572).  Same A/B plane buffers;
// Ez(i), Ey(i) are the (i+1) values loaded for the previous plane.
template <typename T>
struct PlaneH {
  unsigned m;
  T hx, hy, hz, ex, eyk, exk, ezj, exj, ezi, eyi;
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
  long long sj, si;
  unsigned bsj, bsi;  // byte strides of j and i
  bool kin, jin;
  int nx;
};

template <typename T>
__device__ __forceinline__ void load_plane_h(const ColH<T> &q, long long c, int i, PlaneH<T> &p) {
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
  p.ezi = (i < q.nx) ? ldo(q.Ez, o + q.bsi) : T(0);           // Ez(i+1)
  p.eyi = (i < q.nx) ? ldo(q.Ey, o + q.bsi) : T(0);           // Ey(i+1)
}

template <typename T>
__device__ __forceinline__ void compute_plane_h(const Grid<T> &g, const TileTables<T> &tt, const ColH<T> &q,
                                                const PlaneH<T> &p, int i, int j, int k, long long c,
                                                int ly, int lz, T ez_c, T ey_c) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int nm1y = ny + 1, nm1z = nz + 1, nm1x = nx + 1;
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const bool kin = q.kin, jin = q.jin;
  const int m = (int)p.m;
  const int lx = h_layer(i, nx, g.dn[0], g.dp[0]);
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
        T *ps_ = &g.psi_h[P_AY][((long long)i * Ly + ly) * g.pz + k];
        const T ps = g.bm[1][j] * (*ps_) + g.am[1][j] * dez_j;
        *ps_ = ps;
        v = v + g.cpsi[CB_HXEZ][m * nm1y + j] * ps;
      }
      if (lz >= 0) {
        T *ps_ = &g.psi_h[P_AZ][((long long)i * nm1y + j) * Lz + lz];
        const T ps = g.bm[2][k] * (*ps_) + g.am[2][k] * dey_k;
        *ps_ = ps;
        v = v + g.cpsi[CB_HXEY][m * nm1z + k] * ps;
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
        T *ps_ = &g.psi_h[P_BX][((long long)lx * nm1y + j) * g.pz + k];
        const T ps = g.bm[0][i] * (*ps_) + g.am[0][i] * dez_i;
        *ps_ = ps;
        v = v + g.cpsi[CB_HYEZ][m * nm1x + i] * ps;
      }
      if (lz >= 0) {
        T *ps_ = &g.psi_h[P_BZ][((long long)i * nm1y + j) * Lz + lz];
        const T ps = g.bm[2][k] * (*ps_) + g.am[2][k] * dex_k;
        *ps_ = ps;
        v = v + g.cpsi[CB_HYEX][m * nm1z + k] * ps;
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
        T *ps_ = &g.psi_h[P_CX][((long long)lx * nm1y + j) * g.pz + k];
        const T ps = g.bm[0][i] * (*ps_) + g.am[0][i] * dey_i;
        *ps_ = ps;
        v = v + g.cpsi[CB_HZEY][m * nm1x + i] * ps;
      }
      if (ly >= 0) {
        T *ps_ = &g.psi_h[P_CY][((long long)i * Ly + ly) * g.pz + k];
        const T ps = g.bm[1][j] * (*ps_) + g.am[1][j] * dex_j;
        *ps_ = ps;
        v = v + g.cpsi[CB_HZEX][m * nm1y + j] * ps;
      }
    }
    if (yee || slab) sto(q.Hz, o, v);
  }
}

template <typename T, bool LDS, bool PF>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, Tiling tl) {
  extern __shared__ unsigned char lds_raw[];
  int bz, by, bx;
  tile_of_block(tl, bz, by, bx);
  const Cell cell = thread_cell(tl, bz, by, g.pz);
  const int k = cell.k, j = cell.j;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nmat = g.nmat;
  const int i0 = bx * tl.tx;
  const int i1 = min(i0 + tl.tx, nx + 1);
  TileTables<T> tt;
  if (LDS)
    tt = stage_tables<T>(reinterpret_cast<T *>(lds_raw), g.ca, HX, nmat, g.cb[CB_HXEY], g.cb[CB_HYEX],
                         nz + 1, g.cb[CB_HXEZ], g.cb[CB_HZEX], ny + 1, g.cb[CB_HYEZ], g.cb[CB_HZEY],
                         nx + 1, tl, cell, i0);
  else
    tt = global_tables<T>(g.ca, HX, nmat, g.cb[CB_HXEY], g.cb[CB_HYEX], nz + 1, g.cb[CB_HXEZ],
                          g.cb[CB_HZEX], ny + 1, g.cb[CB_HYEZ], g.cb[CB_HZEY], nx + 1, cell);
  if (k > nz || j > ny) return;
  ColH<T> q;
  q.Ex = g.f[EX]; q.Ey = g.f[EY]; q.Ez = g.f[EZ];
  q.Hx = g.f[HX]; q.Hy = g.f[HY]; q.Hz = g.f[HZ];
  q.mat = g.mat;
  q.sj = g.sj;
  q.si = g.si;
  q.bsj = (unsigned)(g.sj * (long long)sizeof(T));
  q.bsi = (unsigned)(g.si * (long long)sizeof(T));
  q.kin = k < nz;
  q.jin = j < ny;
  q.nx = nx;
  const int ly = h_layer(j, ny, g.dn[1], g.dp[1]);
  const int lz = h_layer(k, nz, g.dn[2], g.dp[2]);
  long long c = (long long)i0 * g.si + (long long)j * g.sj + k;
  T ez_c = ldo(q.Ez, (unsigned)c * (unsigned)sizeof(T)), ey_c = ldo(q.Ey, (unsigned)c * (unsigned)sizeof(T));
  if (!PF) {
    for (int i = i0; i < i1; ++i, c += g.si) {
      PlaneH<T> A;
      load_plane_h(q, c, i, A);
      compute_plane_h(g, tt, q, A, i, j, k, c, ly, lz, ez_c, ey_c);
      ez_c = A.ezi;
      ey_c = A.eyi;
    }
    return;
  }
  PlaneH<T> A, B;
  load_plane_h(q, c, i0, A);
  for (int i = i0; i < i1; i += 2, c += 2 * g.si) {
    if (i + 1 < i1) load_plane_h(q, c + g.si, i + 1, B);
    compute_plane_h(g, tt, q, A, i, j, k, c, ly, lz, ez_c, ey_c);
    ez_c = A.ezi;
    ey_c = A.eyi;
    if (i + 1 >= i1) break;
    if (i + 2 < i1) load_plane_h(q, c + 2 * g.si, i + 2, A);
    compute_plane_h(g, tt, q, B, i + 1, j, k, c + g.si, ly, lz, ez_c, ey_c);
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
