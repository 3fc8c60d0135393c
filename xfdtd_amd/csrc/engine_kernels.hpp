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

// ---------------------------------------------------------------------------
// E half-step: Simulation::updateE (simulation_update.cpp:170-196) fused with
// PML::updateE of every face (perfect_match_layer.cpp:574-718), faces in x,
// y, z order.  One thread owns a (j,k) column and marches over an x range.
template <typename T>
__global__ void __launch_bounds__(256) k_update_e(Grid<T> g, int tx) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (k > nz || j > ny) return;
  const int i0 = blockIdx.z * tx;
  const int i1 = min(i0 + tx, nx + 1);
  const int nm1y = ny + 1, nm1z = nz + 1, nm1x = nx + 1;
  T *Ex = g.f[EX], *Ey = g.f[EY], *Ez = g.f[EZ];
  const T *Hx = g.f[HX], *Hy = g.f[HY], *Hz = g.f[HZ];
  const int Lx = g.dn[0] + g.dp[0], Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const int ly = e_layer(j, ny, g.dn[1], g.dp[1]);
  const int lz = e_layer(k, nz, g.dn[2], g.dp[2]);
  for (int i = i0; i < i1; ++i) {
    const long long c = (long long)i * g.si + (long long)j * g.sj + k;
    const int m = g.mat[c];
    const int lx = e_layer(i, nx, g.dn[0], g.dp[0]);
    // Ex: shape (nx, ny+1, nz+1); Yee j in [1,ny), k in [1,nz)
    if (i < nx) {
      const bool yee = (j >= 1 && j < ny && k >= 1 && k < nz);
      if (yee || ly >= 0 || lz >= 0) {
        T v = Ex[c];
        const T dhy = (k >= 1) ? (Hy[c] - Hy[c - 1]) : T(0);
        const T dhz = (j >= 1) ? (Hz[c] - Hz[c - g.sj]) : T(0);
        if (yee)
          v = g.ca[EX * g.nmat + m] * v + g.cb[CB_EXHY][m * nm1z + k] * dhy +
              g.cb[CB_EXHZ][m * nm1y + j] * dhz;
        if (ly >= 0) {
          T *p = &g.psi_e[P_AY][((long long)i * Ly + ly) * g.pz + k];
          const T ps = g.be[1][j] * (*p) + g.ae[1][j] * dhz;
          *p = ps;
          v = v + g.cpsi[CB_EXHZ][m * nm1y + j] * ps;
        }
        if (lz >= 0) {
          T *p = &g.psi_e[P_AZ][((long long)i * nm1y + j) * Lz + lz];
          const T ps = g.be[2][k] * (*p) + g.ae[2][k] * dhy;
          *p = ps;
          v = v + g.cpsi[CB_EXHY][m * nm1z + k] * ps;
        }
        Ex[c] = v;
      }
    }
    // Ey: shape (nx+1, ny, nz+1); Yee k in [1,nz), i in [1,nx)
    if (j < ny) {
      const bool yee = (i >= 1 && i < nx && k >= 1 && k < nz);
      if (yee || lx >= 0 || lz >= 0) {
        T v = Ey[c];
        const T dhx = (k >= 1) ? (Hx[c] - Hx[c - 1]) : T(0);
        const T dhz = (i >= 1) ? (Hz[c] - Hz[c - g.si]) : T(0);
        if (yee)
          v = g.ca[EY * g.nmat + m] * v + g.cb[CB_EYHX][m * nm1z + k] * dhx +
              g.cb[CB_EYHZ][m * nm1x + i] * dhz;
        if (lx >= 0) {
          T *p = &g.psi_e[P_BX][((long long)lx * nm1y + j) * g.pz + k];
          const T ps = g.be[0][i] * (*p) + g.ae[0][i] * dhz;
          *p = ps;
          v = v + g.cpsi[CB_EYHZ][m * nm1x + i] * ps;
        }
        if (lz >= 0) {
          T *p = &g.psi_e[P_BZ][((long long)i * nm1y + j) * Lz + lz];
          const T ps = g.be[2][k] * (*p) + g.ae[2][k] * dhx;
          *p = ps;
          v = v + g.cpsi[CB_EYHX][m * nm1z + k] * ps;
        }
        Ey[c] = v;
      }
    }
    // Ez: shape (nx+1, ny+1, nz); Yee i in [1,nx), j in [1,ny)
    if (k < nz) {
      const bool yee = (i >= 1 && i < nx && j >= 1 && j < ny);
      if (yee || lx >= 0 || ly >= 0) {
        T v = Ez[c];
        const T dhx = (j >= 1) ? (Hx[c] - Hx[c - g.sj]) : T(0);
        const T dhy = (i >= 1) ? (Hy[c] - Hy[c - g.si]) : T(0);
        if (yee)
          v = g.ca[EZ * g.nmat + m] * v + g.cb[CB_EZHX][m * nm1y + j] * dhx +
              g.cb[CB_EZHY][m * nm1x + i] * dhy;
        if (lx >= 0) {
          T *p = &g.psi_e[P_CX][((long long)lx * nm1y + j) * g.pz + k];
          const T ps = g.be[0][i] * (*p) + g.ae[0][i] * dhy;
          *p = ps;
          v = v + g.cpsi[CB_EZHY][m * nm1x + i] * ps;
        }
        if (ly >= 0) {
          T *p = &g.psi_e[P_CY][((long long)i * Ly + ly) * g.pz + k];
          const T ps = g.be[1][j] * (*p) + g.ae[1][j] * dhx;
          *p = ps;
          v = v + g.cpsi[CB_EZHX][m * nm1y + j] * ps;
        }
        Ez[c] = v;
      }
    }
  }
  (void)Lx;
}

// ---------------------------------------------------------------------------
// H half-step: Simulation::updateH (simulation_update.cpp:84-100) fused with
// PML::updateH (perfect_match_layer.cpp:434-572).
template <typename T>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, int tx) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (k > nz || j > ny) return;
  const int i0 = blockIdx.z * tx;
  const int i1 = min(i0 + tx, nx + 1);
  const int nm1y = ny + 1, nm1z = nz + 1, nm1x = nx + 1;
  const T *Ex = g.f[EX], *Ey = g.f[EY], *Ez = g.f[EZ];
  T *Hx = g.f[HX], *Hy = g.f[HY], *Hz = g.f[HZ];
  const int Ly = g.dn[1] + g.dp[1], Lz = g.dn[2] + g.dp[2];
  const int ly = h_layer(j, ny, g.dn[1], g.dp[1]);
  const int lz = h_layer(k, nz, g.dn[2], g.dp[2]);
  for (int i = i0; i < i1; ++i) {
    const long long c = (long long)i * g.si + (long long)j * g.sj + k;
    const int m = g.mat[c];
    const int lx = h_layer(i, nx, g.dn[0], g.dp[0]);
    // Hx: shape (nx+1, ny, nz); Yee i < nx
    if (j < ny && k < nz) {
      const bool yee = (i < nx);
      if (yee || ly >= 0 || lz >= 0) {
        T v = Hx[c];
        const T dey = Ey[c + 1] - Ey[c];
        const T dez = Ez[c + g.sj] - Ez[c];
        if (yee)
          v = g.ca[HX * g.nmat + m] * v + g.cb[CB_HXEY][m * nm1z + k] * dey +
              g.cb[CB_HXEZ][m * nm1y + j] * dez;
        if (ly >= 0) {
          T *p = &g.psi_h[P_AY][((long long)i * Ly + ly) * g.pz + k];
          const T ps = g.bm[1][j] * (*p) + g.am[1][j] * dez;
          *p = ps;
          v = v + g.cpsi[CB_HXEZ][m * nm1y + j] * ps;
        }
        if (lz >= 0) {
          T *p = &g.psi_h[P_AZ][((long long)i * nm1y + j) * Lz + lz];
          const T ps = g.bm[2][k] * (*p) + g.am[2][k] * dey;
          *p = ps;
          v = v + g.cpsi[CB_HXEY][m * nm1z + k] * ps;
        }
        Hx[c] = v;
      }
    }
    // Hy: shape (nx, ny+1, nz); Yee j < ny
    if (i < nx && k < nz) {
      const bool yee = (j < ny);
      if (yee || lx >= 0 || lz >= 0) {
        T v = Hy[c];
        const T dez = Ez[c + g.si] - Ez[c];
        const T dex = Ex[c + 1] - Ex[c];
        if (yee)
          v = g.ca[HY * g.nmat + m] * v + g.cb[CB_HYEZ][m * nm1x + i] * dez +
              g.cb[CB_HYEX][m * nm1z + k] * dex;
        if (lx >= 0) {
          T *p = &g.psi_h[P_BX][((long long)lx * nm1y + j) * g.pz + k];
          const T ps = g.bm[0][i] * (*p) + g.am[0][i] * dez;
          *p = ps;
          v = v + g.cpsi[CB_HYEZ][m * nm1x + i] * ps;
        }
        if (lz >= 0) {
          T *p = &g.psi_h[P_BZ][((long long)i * nm1y + j) * Lz + lz];
          const T ps = g.bm[2][k] * (*p) + g.am[2][k] * dex;
          *p = ps;
          v = v + g.cpsi[CB_HYEX][m * nm1z + k] * ps;
        }
        Hy[c] = v;
      }
    }
    // Hz: shape (nx, ny, nz+1); Yee k < nz
    if (i < nx && j < ny) {
      const bool yee = (k < nz);
      if (yee || lx >= 0 || ly >= 0) {
        T v = Hz[c];
        const T dex = Ex[c + g.sj] - Ex[c];
        const T dey = Ey[c + g.si] - Ey[c];
        if (yee)
          v = g.ca[HZ * g.nmat + m] * v + g.cb[CB_HZEX][m * nm1y + j] * dex +
              g.cb[CB_HZEY][m * nm1x + i] * dey;
        if (lx >= 0) {
          T *p = &g.psi_h[P_CX][((long long)lx * nm1y + j) * g.pz + k];
          const T ps = g.bm[0][i] * (*p) + g.am[0][i] * dey;
          *p = ps;
          v = v + g.cpsi[CB_HZEY][m * nm1x + i] * ps;
        }
        if (ly >= 0) {
          T *p = &g.psi_h[P_CY][((long long)i * Ly + ly) * g.pz + k];
          const T ps = g.bm[1][j] * (*p) + g.am[1][j] * dex;
          *p = ps;
          v = v + g.cpsi[CB_HZEX][m * nm1y + j] * ps;
        }
        Hz[c] = v;
      }
    }
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
