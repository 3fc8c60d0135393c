// sweep_kernels.hpp -- one kernel per time step: the E half-step and the H
// half-step of Simulation::updateE/updateH (simulation_update.cpp:50-108,
// 130-196) with the CPML psi recursions of every face (perfect_match_layer.cpp:
// 434-718) and the TFSF face corrections (tfsf_3d.cpp:185-411) in one march
// over x.
//
// The split kernels read all six components twice per step (once for E, once
// for H).  Here a workgroup reads the step's state once from set A and writes
// the next state to set B (fields and psi, double-buffered by the engine):
//
//   k (z, contiguous in memory) runs over the 64 lanes of a wave, which owns
//   the 64 cells k0..k0+63 (k0 a multiple of 64, so its stores are whole
//   cache lines).  The k-1 neighbour an E update needs and the k+1 neighbour
//   an H update needs come from the adjacent lane by DPP (wave_shr /
//   wave_shl).  The two wave ends are handled packed: lane r (r <= R) loads H
//   at k0 - 1 for row j0 + r and computes the new Ex and Ey of the cell
//   k0 + 64 above the wave for that row, so each plane costs one extra cell
//   update per wave instead of one per row; lane 0 and lane 63 take row r's
//   value from lane r.
//   j runs over R rows held in registers by each thread, plus one halo row
//   above (j0 + R) whose Ex and Ez are recomputed, and the H values of the row
//   below (j0 - 1).
//   x is marched: at plane i the kernel loads plane i + 1, computes E(i + 1)
//   from H(i), H(i + 1), then H(i) from E(i), E(i + 1), and stores E(i), H(i).
//   The first pass of a chunk (i = i0 - 1) only computes E(i0); the last
//   computes E(i1) as a halo plane.
//
// Tiles come in three kinds, each compiled on its own so that the plain one
// carries none of the CPML and TFSF code (whose registers would otherwise
// lower every tile's occupancy): SW_CORE tiles meet no CPML slab and no TFSF
// face (their halo cells included), SW_ZONLY tiles meet them only along z
// (the per-lane z-slab psi and the z faces), SW_ANY tiles take everything.
// The host sorts the tiles into one list per kind and launches each list.
//
// The coefficient records, the CPML b/a values of the tile's planes and rows
// and the addresses the CPML and TFSF branches need are staged in LDS once per
// workgroup, so the march waits on one memory round trip per plane.  The TFSF
// corrections of the step are precomputed per face cell (k_tfsf_terms) and
// only added here.
//
// Every halo value is computed with the same expressions, from the same set-A
// inputs, as by the thread that owns the cell, so it has the same bits.  The
// arithmetic follows compute_plane_e / compute_plane_h and tfsf_e_body /
// tfsf_h_body (and the reference) expression by expression; with
// -ffp-contract=off the FP64 instantiation is bit-identical to the split
// kernels and the oracle.
//
// Loads outside the grid are clamped to the nearest valid slot; the values
// they give only enter terms that the reference's index-range tests discard,
// exactly as in the split kernels.
#pragma once

namespace xfk {

// Tile kinds of the fused step (see above).
enum { SW_CORE = 0, SW_ZONLY = 1, SW_ANY = 2 };

// One workgroup's tile: x planes [i0, i1), rows [j0, j0 + rows) in waves of R
// rows from j0, k chunk kc (cells 64 kc .. 64 kc + 63); kr = kc | rows << 16.
struct alignas(16) SweepTile {
  int i0, i1, j0, kr;
};

// launch geometry of one tile list: workgroups of `waves` waves.  The
// hardware deals workgroup ids to the 8 XCDs round-robin and each XCD has its
// own L2.  By default tiles follow workgroup ids (k chunk fastest, then j
// group, then x chunk), so the whole chip works on neighbouring tiles at a
// time; with xcd_runs each XCD takes one contiguous run of the list instead,
// which keeps the rows and cache lines that neighbouring tiles share in one
// L2 but measures 10-15% slower on the 400^3 grid.
struct SweepTiling {
  const SweepTile *tiles;
  int nblocks;
  int waves;  // waves per workgroup, each a strip of R rows
  int xcd_runs;
};

// TFSF face terms of one step, precomputed by k_tfsf_terms: for each of the
// twelve E segments and twelve H segments of TFSF3D::updateE / updateH
// (tfsf_3d.cpp:185-411) the product coefficient * incident value of every
// cell, as an FP64 array over the face's two in-face indices.  E segments
// 0..11: Ey@x, Ey@z, Ez@x, Ez@y, Ex@y, Ex@z (N face, then P face); H
// segments 12..23: Hz@x, Hz@y, Hy@x, Hy@z, Hx@y, Hx@z.  Segment s starts at
// element s * stride of one buffer (stride: the largest segment; sizes:
// tf_seg_size).
struct TfTerms {
  double *buf;
  long long stride;
};

__host__ __device__ inline long long tf_seg_size(int s, long long ni, long long nj, long long nk) {
  switch (s >> 1) {
    case 0: return nj * (nk + 1);
    case 1: return (ni + 1) * nj;
    case 2: return (nj + 1) * nk;
    case 3: return (ni + 1) * nk;
    case 4: return ni * (nk + 1);
    case 5: return ni * (nj + 1);
    case 6: return nj * (nk + 1);
    case 7: return ni * (nk + 1);
    case 8: return (nj + 1) * nk;
    case 9: return ni * (nj + 1);
    case 10: return (ni + 1) * nk;
    default: return (ni + 1) * nj;
  }
}

// Arguments of one fused step (one direction of the ping-pong).
// element of slot q (P_AY..P_CY) at psi index idx within one side of a psi
// block: pair region y, z or x of 2 * psl elements, the slot's half of it
__device__ __forceinline__ unsigned sw_psi(int q, unsigned idx, unsigned psl) {
  const int region = (q == P_AY || q == P_CY) ? 0 : (q == P_AZ || q == P_BZ) ? 1 : 2;
  const int half = (q == P_CY || q == P_BZ || q == P_CX) ? 1 : 0;
  return (unsigned)region * 2u * psl + 2u * idx + (unsigned)half;
}

template <typename T>
struct SweepArgs {
  const T *src[6];  // fields of set A (read)
  T *dst[6];        // fields of set B (written)
  const uint8_t *mat;
  const T *self[2], *kt[2], *jt[2], *it[2];  // curl coefficient records (E, H)
  const T *kp[2], *jp[2], *ip[2];            // psi coefficient pairs
  const T *be[3], *ae[3], *bm[3], *am[3];
  // psi of set A: one block; per side (E, then H) the interleaved pairs y
  // (AY, CY), z (AZ, BZ), x (BX, CX) of 2 * psi_slot elements each (sw_psi)
  const T *psiA;
  T *psiB;  // psi of set B, same layout
  unsigned psi_slot;
  const double *terms;  // TFSF face terms of the step (segment s at s * term_stride)
  unsigned term_stride;
  int nx, ny, nz, pz, nmat;
  unsigned sj, si;
  int dn[3], dp[3];
  int tf, tb[6];  // TFSF box li, ri, lj, rj, lk, rk
  SweepTiling st;
  unsigned long long *div;  // lowered to `step` where an H value comes out non-finite
  long long step;
};

// CPML layer of index n on one axis and side (-1 outside the slabs)
__device__ __forceinline__ int cpml_layer(bool e_side, int n, int N, int dn, int dp) {
  return e_side ? e_layer(n, N, dn, dp) : h_layer(n, N, dn, dp);
}

// LDS layout of a workgroup's tables, in units of T, with MS material slots:
// the self records of both sides ([side][MS][4]); for each side and table kind
// (curl, psi) the pairs over the k slots (the 64 lanes and the cell above) as
// [MS][SW_KS], so that consecutive lanes read consecutive pairs, over the j
// slots (the workgroup's rows and the top halo row) as [MS][js] and over the
// i slots (up to SW_TXMAX planes and the halo plane) as [MS][is]; then the x
// and y CPML b/a values.  All offsets are compile-time constants.
constexpr int SW_KS = 65;
constexpr int SW_WAVES = 4;
constexpr int SW_TXMAX = 32;
template <int R, int MS>
struct SweepLds {
  static constexpr int js = SW_WAVES * R + 1, is = SW_TXMAX + 1;
  static constexpr int per = 2 * MS * (SW_KS + js + is);
  static constexpr int self(int side) { return side * 4 * MS; }
  static constexpr int kpair(int side, int psi) { return 8 * MS + (2 * side + psi) * per; }
  static constexpr int jpair(int side, int psi) { return kpair(side, psi) + 2 * MS * SW_KS; }
  static constexpr int ipair(int side, int psi) { return jpair(side, psi) + 2 * MS * js; }
  static constexpr int ba_x = 8 * MS + 4 * per;  // [4][is]: be ae bm am
  static constexpr int ba_y = ba_x + 4 * is;     // [4][js]
  static constexpr int total = ba_y + 4 * js;
};

// CPML state of one cell along z: layer, b, a (E side and H side)
template <typename T>
struct ZLay {
  int le, lh;
  T be, ae, bh, ah;
};

// value of v in lane `lane` (a compile-time constant after unrolling)
__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// What the cell updates of one wave share: grid sizes, CPML extents, the
// staged tables, the psi blocks and the TFSF box with the step's face terms.
template <typename T, int R, int MS>
struct SwCtx {
  using L = SweepLds<R, MS>;
  int nx, ny, nz, pz;
  int dn1, dp1, Ly, Lz, nm1y;
  int i0, jwg;
  const T *lds;
  const T *__restrict__ pA;
  T *__restrict__ pB;
  unsigned psl;
  int li, ri, lj, rj, lk, rk;
  long long nj, nk;
  const double *terms;
  unsigned tstr;
  __device__ const P4<T> &self(int side, int m) const { return reinterpret_cast<const P4<T> *>(lds + L::self(side))[m]; }
  __device__ const P2<T> &kp(int side, int psi, int m, int s) const {
    return reinterpret_cast<const P2<T> *>(lds + L::kpair(side, psi))[m * SW_KS + s];
  }
  __device__ const P2<T> &jp(int side, int psi, int m, int s) const {
    return reinterpret_cast<const P2<T> *>(lds + L::jpair(side, psi))[m * L::js + s];
  }
  __device__ const P2<T> &ip(int side, int psi, int m, int s) const {
    return reinterpret_cast<const P2<T> *>(lds + L::ipair(side, psi))[m * L::is + s];
  }
  __device__ const double *seg(int s) const { return terms + (size_t)s * tstr; }
};

// One E cell (p, j, kk) of material m: Simulation::updateE + PML::updateE +
// TFSF3D::updateE, in the order of compute_plane_e and tfsf_e_body, for the
// components DX, DY, DZ.  d*: the curl differences; kslot: the cell's k slot
// in the staged tables; own: this thread stores the cell's psi; slab: the
// wave may meet a CPML slab, tfp: the plane may hold TFSF E faces (both
// wave-uniform); lxe: the plane's E layer along x.  A SW_ZONLY tile has no
// x or y slab and no x or y face, so only the z terms are compiled in.
template <typename T, int R, int MS, int KIND, bool DX, bool DY, bool DZ>
__device__ __forceinline__ void sw_e_cell(const SwCtx<T, R, MS> &c, int p, int j, int kk, int kslot, int m,
                                          const ZLay<T> &z, T exv, T eyv, T ezv, T dhy_k, T dhz_j, T dhx_k,
                                          T dhz_i, T dhx_j, T dhy_i, bool own, bool slab, bool tfp, int lxe, T &vx,
                                          T &vy, T &vz) {
  using L = SweepLds<R, MS>;
  const int is = p - c.i0, js = j - c.jwg;
  const bool kin = kk >= 1 && kk < c.nz, klt = kk < c.nz;
  const bool jin = j >= 1 && j < c.ny, jlt = j < c.ny;
  const bool ix = p >= 1 && p < c.nx, plt = p < c.nx;
  const P4<T> cs = c.self(0, m);
  P2<T> ck{}, cj{}, ci{};
  if (DX || DY) ck = c.kp(0, 0, m, kslot);
  if (DX || DZ) cj = c.jp(0, 0, m, js);
  if (DY || DZ) ci = c.ip(0, 0, m, is);
  if (DX) {
    const T yx = cs.x * exv + ck.x * dhy_k + cj.x * dhz_j;
    vx = (plt && jin && kin) ? yx : exv;
  }
  if (DY) {
    const T yy = cs.y * eyv + ck.y * dhx_k + ci.x * dhz_i;
    vy = (ix && jlt && kin) ? yy : eyv;
  }
  if (DZ) {
    const T yz = cs.z * ezv + cj.y * dhx_j + ci.y * dhy_i;
    vz = (ix && jin && klt) ? yz : ezv;
  }
  if (KIND != SW_CORE && slab) {
    const int pl = min(p, c.nx);  // tables and psi of a plane beyond nx are never used
    const unsigned jc = (unsigned)min(j, c.ny);
    const unsigned psl = c.psl;
    if constexpr (KIND == SW_ANY) {
      const int lye = j <= c.ny ? cpml_layer(true, j, c.ny, c.dn1, c.dp1) : -1;
      const T bye = c.lds[L::ba_y + js], aye = c.lds[L::ba_y + L::js + js];
      const T bxe = c.lds[L::ba_x + is], axe = c.lds[L::ba_x + L::is + is];
      const unsigned kq = (unsigned)min(kk, c.pz - 1);
      if (DX && plt) {
        if (lye >= 0) {
          const unsigned o = sw_psi(0, psi_y(pl, lye, kq, c.Ly, c.pz), psl);
          const T ps = bye * c.pA[o] + aye * dhz_j;
          if (own) c.pB[o] = ps;
          vx = vx + c.jp(0, 1, m, js).x * ps;
        }
        if (z.le >= 0) {
          const unsigned o = sw_psi(1, psi_z(pl, jc, z.le, c.nm1y, c.Lz), psl);
          const T ps = z.be * c.pA[o] + z.ae * dhy_k;
          if (own) c.pB[o] = ps;
          vx = vx + c.kp(0, 1, m, kslot).x * ps;
        }
      }
      if (DY && jlt) {
        if (lxe >= 0) {
          const unsigned o = sw_psi(2, psi_x(lxe, jc, kq, c.nm1y, c.pz), psl);
          const T ps = bxe * c.pA[o] + axe * dhz_i;
          if (own) c.pB[o] = ps;
          vy = vy + c.ip(0, 1, m, is).x * ps;
        }
        if (z.le >= 0) {
          const unsigned o = sw_psi(3, psi_z(pl, jc, z.le, c.nm1y, c.Lz), psl);
          const T ps = z.be * c.pA[o] + z.ae * dhx_k;
          if (own) c.pB[o] = ps;
          vy = vy + c.kp(0, 1, m, kslot).y * ps;
        }
      }
      if (DZ && klt) {
        if (lxe >= 0) {
          const unsigned o = sw_psi(4, psi_x(lxe, jc, kq, c.nm1y, c.pz), psl);
          const T ps = bxe * c.pA[o] + axe * dhy_i;
          if (own) c.pB[o] = ps;
          vz = vz + c.ip(0, 1, m, is).y * ps;
        }
        if (lye >= 0) {
          const unsigned o = sw_psi(5, psi_y(pl, lye, kq, c.Ly, c.pz), psl);
          const T ps = bye * c.pA[o] + aye * dhx_j;
          if (own) c.pB[o] = ps;
          vz = vz + c.jp(0, 1, m, js).y * ps;
        }
      }
    } else {
      // z slab only: Ex and Ey take their z psi
      if (z.le >= 0) {
        if (DX && plt) {
          const unsigned o = sw_psi(1, psi_z(pl, jc, z.le, c.nm1y, c.Lz), psl);
          const T ps = z.be * c.pA[o] + z.ae * dhy_k;
          if (own) c.pB[o] = ps;
          vx = vx + c.kp(0, 1, m, kslot).x * ps;
        }
        if (DY && jlt) {
          const unsigned o = sw_psi(3, psi_z(pl, jc, z.le, c.nm1y, c.Lz), psl);
          const T ps = z.be * c.pA[o] + z.ae * dhx_k;
          if (own) c.pB[o] = ps;
          vy = vy + c.kp(0, 1, m, kslot).y * ps;
        }
      }
    }
  }
  // TFSF E corrections (tfsf_e_body), in FP64, on face cells only
  if (KIND != SW_CORE && tfp) {
    const int li = c.li, ri = c.ri, lj = c.lj, rj = c.rj, lk = c.lk, rk = c.rk;
    const long long nj = c.nj, nk = c.nk;
    if (j >= lj && j <= rj) {
      const bool kb = kk >= lk && kk <= rk;
      if constexpr (KIND == SW_ANY) {
        // Ey: j in [lj, rj), (i, k) on the perimeter: x faces then z faces
        if (DY && j < rj && kb && (p == li || p == ri || kk == lk || kk == rk)) {
          double v = vy;
          const long long ox = (long long)(j - lj) * (nk + 1) + (kk - lk), oz = (long long)(p - li) * nj + (j - lj);
          if (p == li) v = v + c.seg(0)[ox];
          if (p == ri) v = v - c.seg(1)[ox];
          if (kk == lk) v = v - c.seg(2)[oz];
          if (kk == rk) v = v + c.seg(3)[oz];
          vy = (T)v;
        }
        // Ez: k in [lk, rk), (i, j) on the perimeter: x faces then y faces
        if (DZ && kk >= lk && kk < rk && (p == li || p == ri || j == lj || j == rj)) {
          double v = vz;
          const long long ox = (long long)(j - lj) * nk + (kk - lk), oy = (long long)(p - li) * nk + (kk - lk);
          if (p == li) v = v - c.seg(4)[ox];
          if (p == ri) v = v + c.seg(5)[ox];
          if (j == lj) v = v + c.seg(6)[oy];
          if (j == rj) v = v - c.seg(7)[oy];
          vz = (T)v;
        }
        // Ex: i in [li, ri), (j, k) on the perimeter: y faces then z faces
        if (DX && p < ri && kb && (j == lj || j == rj || kk == lk || kk == rk)) {
          double v = vx;
          const long long oy = (long long)(p - li) * (nk + 1) + (kk - lk),
                          oz = (long long)(p - li) * (nj + 1) + (j - lj);
          if (j == lj) v = v - c.seg(8)[oy];
          if (j == rj) v = v + c.seg(9)[oy];
          if (kk == lk) v = v + c.seg(10)[oz];
          if (kk == rk) v = v - c.seg(11)[oz];
          vx = (T)v;
        }
      } else {
        // z faces only (p is neither li nor ri, j neither lj nor rj)
        if (kk == lk || kk == rk) {
          if (DY && j < rj) {
            double v = vy;
            const long long oz = (long long)(p - li) * nj + (j - lj);
            if (kk == lk) v = v - c.seg(2)[oz];
            if (kk == rk) v = v + c.seg(3)[oz];
            vy = (T)v;
          }
          if (DX && p < ri) {
            double v = vx;
            const long long oz = (long long)(p - li) * (nj + 1) + (j - lj);
            if (kk == lk) v = v + c.seg(10)[oz];
            if (kk == rk) v = v - c.seg(11)[oz];
            vx = (T)v;
          }
        }
      }
    }
  }
}

// One H cell (i, j, k) of material m: Simulation::updateH + PML::updateH +
// TFSF3D::updateH in the order of compute_plane_h and tfsf_h_body.
template <typename T, int R, int MS, int KIND>
__device__ __forceinline__ void sw_h_cell(const SwCtx<T, R, MS> &c, int i, int j, int k, int lane, int m,
                                          const ZLay<T> &z, T hxv, T hyv, T hzv, T dey_k, T dez_j, T dez_i, T dex_k,
                                          T dex_j, T dey_i, bool own, bool slab, bool tfh, int lxh, T &vx, T &vy,
                                          T &vz) {
  using L = SweepLds<R, MS>;
  const int is = i - c.i0, js = j - c.jwg;
  const bool ilt = i < c.nx, jlt = j < c.ny, klt = k < c.nz;
  const P4<T> cs = c.self(1, m);
  const P2<T> ck = c.kp(1, 0, m, lane), cj = c.jp(1, 0, m, js), ci = c.ip(1, 0, m, is);
  const T yx = cs.x * hxv + ck.x * dey_k + cj.x * dez_j;
  const T yy = cs.y * hyv + ci.x * dez_i + ck.y * dex_k;
  const T yz = cs.z * hzv + cj.y * dex_j + ci.y * dey_i;
  const bool yee = ilt && jlt && klt;
  vx = yee ? yx : hxv;
  vy = yee ? yy : hyv;
  vz = yee ? yz : hzv;
  if (KIND != SW_CORE && slab) {
    const unsigned jc = (unsigned)min(j, c.ny);
    const unsigned psl = c.psl, oh = 6 * psl;
    if constexpr (KIND == SW_ANY) {
      const int lyh = j <= c.ny ? cpml_layer(false, j, c.ny, c.dn1, c.dp1) : -1;
      const T byh = c.lds[L::ba_y + 2 * L::js + js], ayh = c.lds[L::ba_y + 3 * L::js + js];
      const T bxh = c.lds[L::ba_x + 2 * L::is + is], axh = c.lds[L::ba_x + 3 * L::is + is];
      const unsigned kl = (unsigned)min(k, c.pz - 1);
      if (jlt && klt) {
        if (lyh >= 0) {
          const unsigned o = oh + sw_psi(0, psi_y(i, lyh, kl, c.Ly, c.pz), psl);
          const T ps = byh * c.pA[o] + ayh * dez_j;
          if (own) c.pB[o] = ps;
          vx = vx + c.jp(1, 1, m, js).x * ps;
        }
        if (z.lh >= 0) {
          const unsigned o = oh + sw_psi(1, psi_z(i, jc, z.lh, c.nm1y, c.Lz), psl);
          const T ps = z.bh * c.pA[o] + z.ah * dey_k;
          if (own) c.pB[o] = ps;
          vx = vx + c.kp(1, 1, m, lane).x * ps;
        }
      }
      if (ilt && klt) {
        if (lxh >= 0) {
          const unsigned o = oh + sw_psi(2, psi_x(lxh, jc, kl, c.nm1y, c.pz), psl);
          const T ps = bxh * c.pA[o] + axh * dez_i;
          if (own) c.pB[o] = ps;
          vy = vy + c.ip(1, 1, m, is).x * ps;
        }
        if (z.lh >= 0) {
          const unsigned o = oh + sw_psi(3, psi_z(i, jc, z.lh, c.nm1y, c.Lz), psl);
          const T ps = z.bh * c.pA[o] + z.ah * dex_k;
          if (own) c.pB[o] = ps;
          vy = vy + c.kp(1, 1, m, lane).y * ps;
        }
      }
      if (ilt && jlt) {
        if (lxh >= 0) {
          const unsigned o = oh + sw_psi(4, psi_x(lxh, jc, kl, c.nm1y, c.pz), psl);
          const T ps = bxh * c.pA[o] + axh * dey_i;
          if (own) c.pB[o] = ps;
          vz = vz + c.ip(1, 1, m, is).y * ps;
        }
        if (lyh >= 0) {
          const unsigned o = oh + sw_psi(5, psi_y(i, lyh, kl, c.Ly, c.pz), psl);
          const T ps = byh * c.pA[o] + ayh * dex_j;
          if (own) c.pB[o] = ps;
          vz = vz + c.jp(1, 1, m, js).y * ps;
        }
      }
    } else {
      // z slab only: Hx and Hy take their z psi
      if (z.lh >= 0 && klt) {
        if (jlt) {
          const unsigned o = oh + sw_psi(1, psi_z(i, jc, z.lh, c.nm1y, c.Lz), psl);
          const T ps = z.bh * c.pA[o] + z.ah * dey_k;
          if (own) c.pB[o] = ps;
          vx = vx + c.kp(1, 1, m, lane).x * ps;
        }
        if (ilt) {
          const unsigned o = oh + sw_psi(3, psi_z(i, jc, z.lh, c.nm1y, c.Lz), psl);
          const T ps = z.bh * c.pA[o] + z.ah * dex_k;
          if (own) c.pB[o] = ps;
          vy = vy + c.kp(1, 1, m, lane).y * ps;
        }
      }
    }
  }
  // TFSF H corrections (tfsf_h_body), in FP64; no H cell takes two
  if (KIND != SW_CORE && tfh) {
    const int li = c.li, ri = c.ri, lj = c.lj, rj = c.rj, lk = c.lk, rk = c.rk;
    const long long nj = c.nj, nk = c.nk;
    if (j >= lj - 1 && j <= rj) {
      if constexpr (KIND == SW_ANY) {
        // Hz: x faces (j in [lj, rj), k in [lk, rk]), y faces (i in [li, ri), k in [lk, rk])
        if (k >= lk && k <= rk) {
          if (j >= lj && j < rj && (i == li - 1 || i == ri)) {
            const double t = c.seg(i == ri ? 13 : 12)[(long long)(j - lj) * (nk + 1) + (k - lk)];
            vz = (T)(i == ri ? (double)vz - t : (double)vz + t);
          } else if (i >= li && i < ri && (j == lj - 1 || j == rj)) {
            const double t = c.seg(j == rj ? 15 : 14)[(long long)(i - li) * (nk + 1) + (k - lk)];
            vz = (T)(j == rj ? (double)vz + t : (double)vz - t);
          }
        }
        // Hy: x faces (j in [lj, rj], k in [lk, rk)), z faces (i in [li, ri), j in [lj, rj])
        if (j >= lj) {
          if (k >= lk && k < rk && (i == li - 1 || i == ri)) {
            const double t = c.seg(i == ri ? 17 : 16)[(long long)(j - lj) * nk + (k - lk)];
            vy = (T)(i == ri ? (double)vy + t : (double)vy - t);
          } else if (i >= li && i < ri && (k == lk - 1 || k == rk)) {
            const double t = c.seg(k == rk ? 19 : 18)[(long long)(i - li) * (nj + 1) + (j - lj)];
            vy = (T)(k == rk ? (double)vy - t : (double)vy + t);
          }
        }
        // Hx: y faces (k in [lk, rk), i in [li, ri]), z faces (i in [li, ri], j in [lj, rj))
        if (i >= li && i <= ri) {
          if (k >= lk && k < rk && (j == lj - 1 || j == rj)) {
            const double t = c.seg(j == rj ? 21 : 20)[(long long)(i - li) * nk + (k - lk)];
            vx = (T)(j == rj ? (double)vx - t : (double)vx + t);
          } else if (j >= lj && j < rj && (k == lk - 1 || k == rk)) {
            const double t = c.seg(k == rk ? 23 : 22)[(long long)(i - li) * nj + (j - lj)];
            vx = (T)(k == rk ? (double)vx + t : (double)vx - t);
          }
        }
      } else if (k == lk - 1 || k == rk) {
        // z faces only (i is neither li - 1 nor ri, j neither lj - 1 nor rj)
        if (j >= lj && i >= li && i < ri) {
          const double t = c.seg(k == rk ? 19 : 18)[(long long)(i - li) * (nj + 1) + (j - lj)];
          vy = (T)(k == rk ? (double)vy - t : (double)vy + t);
        }
        if (i >= li && i <= ri && j >= lj && j < rj) {
          const double t = c.seg(k == rk ? 23 : 22)[(long long)(i - li) * nj + (j - lj)];
          vx = (T)(k == rk ? (double)vx + t : (double)vx - t);
        }
      }
    }
  }
}

template <typename T, int R, int MS, int KIND>
__device__ __forceinline__ void sweep_body(const SweepArgs<T> &a, int block) {
  extern __shared__ __align__(16) unsigned char sw_lds[];
  using L = SweepLds<R, MS>;
  const int nx = a.nx, ny = a.ny, nz = a.nz, pz = a.pz, nmat = a.nmat;
  const SweepTiling st = a.st;
  int tile = block;
  if (st.xcd_runs) {
    const int q = st.nblocks >> 3, rr = st.nblocks & 7, xcd = block & 7;
    tile = xcd * q + min(xcd, rr) + (block >> 3);
  }
  const SweepTile tl = st.tiles[tile];
  const int lane = (int)(threadIdx.x & 63u);
  const int wv = (int)(threadIdx.x >> 6);
  const int jwg = tl.j0;  // first row of the workgroup
  const int jend = jwg + (tl.kr >> 16);
  const int j0 = jwg + wv * R;
  const int i0 = tl.i0;
  const int i1 = tl.i1;
  const int k0 = (tl.kr & 0xffff) * 64;
  const int k = k0 + lane;
  const int dn0 = a.dn[0], dp0 = a.dp[0], dn1 = a.dn[1], dp1 = a.dp[1], dn2 = a.dn[2], dp2 = a.dp[2];
  const int Lx = dn0 + dp0, Ly = dn1 + dp1, Lz = dn2 + dp2;

  // ---- stage the tile's tables in LDS
  T *lds = reinterpret_cast<T *>(sw_lds);
  // psi pairs only in tiles that meet a slab; CPML b/a values of x and y only
  // in SW_ANY tiles
  bool tile_slab = false;
  if constexpr (KIND == SW_ANY)
    tile_slab = (Lx > 0 && meets_slab(i0, i1, nx, dn0, dp0)) ||
                (Ly > 0 && meets_slab(jwg, jwg + L::js - 1, ny, dn1, dp1)) ||
                (Lz > 0 && meets_slab(k0 - 1, k0 + 64, nz, dn2, dp2));
  else if constexpr (KIND == SW_ZONLY)
    tile_slab = Lz > 0 && meets_slab(k0 - 1, k0 + 64, nz, dn2, dp2);
  {
    const int nth = (int)blockDim.x, t0 = (int)threadIdx.x;
    for (int q = t0; q < 8 * nmat; q += nth) {
      const int side = q / (4 * nmat), r = q % (4 * nmat);
      lds[L::self(side) + r] = a.self[side][r];
    }
    // global records are [slot][material][2]; LDS records [material][slot][2]
    const int npsi = tile_slab ? 2 : 1, n2 = 2 * nmat;
    for (int side = 0; side < 2; ++side)
      for (int psi = 0; psi < npsi; ++psi) {
        const T *tk = psi ? a.kp[side] : a.kt[side];
        for (int q = t0; q < n2 * SW_KS; q += nth) {
          const int sl = q / n2, r = q % n2;
          lds[L::kpair(side, psi) + ((r >> 1) * SW_KS + sl) * 2 + (r & 1)] = tk[min(k0 + sl, nz) * n2 + r];
        }
        if (psi && KIND != SW_ANY) continue;  // z psi pairs only
        const T *tj = psi ? a.jp[side] : a.jt[side];
        const T *ti = psi ? a.ip[side] : a.it[side];
        for (int q = t0; q < n2 * L::js; q += nth) {
          const int sl = q / n2, r = q % n2;
          lds[L::jpair(side, psi) + ((r >> 1) * L::js + sl) * 2 + (r & 1)] = tj[min(jwg + sl, ny) * n2 + r];
        }
        for (int q = t0; q < n2 * L::is; q += nth) {
          const int sl = q / n2, r = q % n2;
          lds[L::ipair(side, psi) + ((r >> 1) * L::is + sl) * 2 + (r & 1)] = ti[min(i0 + sl, nx) * n2 + r];
        }
      }
    if (KIND == SW_ANY && tile_slab) {
      for (int q = t0; q < 4 * L::is; q += nth) {
        const int w = q / L::is, n = min(i0 + q % L::is, nx);
        const T *src = w == 0 ? a.be[0] : w == 1 ? a.ae[0] : w == 2 ? a.bm[0] : a.am[0];
        lds[L::ba_x + q] = src[n];
      }
      for (int q = t0; q < 4 * L::js; q += nth) {
        const int w = q / L::js, n = min(jwg + q % L::js, ny);
        const T *src = w == 0 ? a.be[1] : w == 1 ? a.ae[1] : w == 2 ? a.bm[1] : a.am[1];
        lds[L::ba_y + q] = src[n];
      }
    }
  }
  __syncthreads();
  if (j0 > ny || j0 >= jend) return;  // the whole wave lies beyond the grid or the tile

  SwCtx<T, R, MS> c;
  c.nx = nx; c.ny = ny; c.nz = nz; c.pz = pz;
  c.dn1 = dn1; c.dp1 = dp1; c.Ly = Ly; c.Lz = Lz; c.nm1y = ny + 1;
  c.i0 = i0; c.jwg = jwg;
  c.lds = lds;
  c.pA = a.psiA; c.pB = a.psiB; c.psl = a.psi_slot;
  c.li = c.ri = c.lj = c.rj = c.lk = c.rk = 0;
  c.terms = a.terms; c.tstr = a.term_stride;
  // TFSF: the E corrections touch cells of [li, ri] x [lj, rj] x [lk, rk]
  // on its faces, the H corrections cells of [li-1, ri] x [lj-1, rj] x
  // [lk-1, rk] on its outer faces.  Wave-uniform: does the wave's (j, k)
  // footprint (the halo row and cell included) meet those boxes?
  bool tf_e = false, tf_h = false;
  if (KIND != SW_CORE && a.tf) {
    c.li = a.tb[0]; c.ri = a.tb[1]; c.lj = a.tb[2]; c.rj = a.tb[3]; c.lk = a.tb[4]; c.rk = a.tb[5];
    tf_e = j0 <= c.rj && j0 + R >= c.lj && k0 <= c.rk && k0 + 64 >= c.lk;
    tf_h = j0 <= c.rj && j0 + R - 1 >= c.lj - 1 && k0 <= c.rk && k0 + 63 >= c.lk - 1;
  }
  c.nj = c.rj - c.lj;
  c.nk = c.rk - c.lk;
  const int li = c.li, ri = c.ri;

  // ---- per-lane constants: this lane's cell k.  The k - 1 neighbour of
  // lane 0 and the k0 + 64 cell above lane 63 are handled packed: lane r
  // (r <= R) loads and computes them for row j0 + r, and lane 0 / lane 63
  // take row r's value from lane r.
  const int kl = min(k, pz - 1);
  const int kh = k0 + 64, khl = min(kh, pz - 1);
  const int kbl = max(k0 - 1, 0), k63 = min(k0 + 63, pz - 1);
  const bool kown = k <= nz;
  auto zlayer = [&](int kk) {
    ZLay<T> z;
    z.le = z.lh = -1;
    z.be = z.ae = z.bh = z.ah = T(0);
    if (KIND == SW_CORE || kk > nz) return z;
    z.le = cpml_layer(true, kk, nz, dn2, dp2);
    z.lh = cpml_layer(false, kk, nz, dn2, dp2);
    if (z.le >= 0) {
      z.be = a.be[2][kk];
      z.ae = a.ae[2][kk];
    }
    if (z.lh >= 0) {
      z.bh = a.bm[2][kk];
      z.ah = a.am[2][kk];
    }
    return z;
  };
  const ZLay<T> zc = zlayer(k);
  const ZLay<T> zt = zlayer(kh);  // the halo cell
  const unsigned sj = a.sj, si = a.si;
  const T *__restrict__ fEx = a.src[EX];
  const T *__restrict__ fEy = a.src[EY];
  const T *__restrict__ fEz = a.src[EZ];
  const T *__restrict__ fHx = a.src[HX];
  const T *__restrict__ fHy = a.src[HY];
  const T *__restrict__ fHz = a.src[HZ];
  const uint8_t *__restrict__ mat = a.mat;
  // does any row of this wave, or any lane (or the halo cell), meet a y / z
  // slab?  (wave-uniform)
  const bool any_z = KIND != SW_CORE && Lz > 0 && __ballot(zc.le >= 0 || zc.lh >= 0 || zt.le >= 0) != 0ull;
  const bool any_y = KIND == SW_ANY && Ly > 0 && meets_slab(j0, j0 + R, ny, dn1, dp1);

  // element offset of row r - 1 within a plane (clamped), r in [0, R + 1]
  unsigned roff[R + 2];
#pragma unroll
  for (int r = -1; r <= R; ++r) roff[r + 1] = (unsigned)min(max(j0 + r, 0), ny) * sj;
  // this lane's packed row j0 + hr and the row below it
  const int hr = min(lane, R);
  const int hj = j0 + hr;
  const unsigned hro = (unsigned)min(hj, ny) * sj;
  const unsigned hrom = (unsigned)min(max(hj - 1, 0), ny) * sj;
  constexpr unsigned S = (unsigned)sizeof(T);

  // carried between planes (plane i): E(i) of rows 0..R (Ey rows 0..R-1),
  // H(i) of rows 0..R-1 (Hy also row R), the material of rows 0..R-1; packed:
  // the halo cell's new Ex, Ey and old Hz
  T hx[R], hy[R + 1], hz[R];
  T ex[R + 1], ey[R], ez[R + 1];
  unsigned mcur[R];
  T gx = T(0), gy = T(0), ghz = T(0);
#pragma unroll
  for (int r = 0; r <= R; ++r) {
    hy[r] = ex[r] = ez[r] = T(0);
    if (r < R) hz[r] = hx[r] = ey[r] = T(0), mcur[r] = 0;
  }
  // before the first plane: H(i0 - 1), zero before plane 0
  if (i0 >= 1) {
    const unsigned pbase = (unsigned)(i0 - 1) * si;
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      const unsigned cc = pbase + roff[r + 1];
      hy[r] = ldo(fHy, (cc + kl) * S);
      if (r < R) hz[r] = ldo(fHz, (cc + kl) * S);
    }
    ghz = ldo(fHz, (pbase + hro + khl) * S);
  }
  bool bad = false;  // a non-finite H value was stored
  // i = i0 - 1 only computes E(i0); from i0 on, each pass computes E(i + 1)
  // and H(i) and stores plane i
  for (int i = i0 - 1; i < i1; ++i) {
    const int p = i + 1;
    // ---- loads of plane p (clamped to nx) from set A
    T hxq[R + 2], hyq[R + 1], hzq[R + 2];
    T ax[R + 1], ay[R], az[R + 1];
    unsigned mq[R + 1];
    T qhyb, qhxb, qex, qey, qhx, qhy, qhz, qhzm, qhy63, qhx63;
    unsigned qm;
    {
      const unsigned pbase = (unsigned)min(p, nx) * si;
#pragma unroll
      for (int r = -1; r <= R; ++r) {
        const unsigned o = (pbase + roff[r + 1] + kl) * S;
        hxq[r + 1] = ldo(fHx, o);
        hzq[r + 1] = ldo(fHz, o);
        if (r >= 0) {
          hyq[r] = ldo(fHy, o);
          ax[r] = ldo(fEx, o);
          az[r] = ldo(fEz, o);
          mq[r] = ldm(mat, pbase + roff[r + 1] + kl);
          if (r < R) ay[r] = ldo(fEy, o);
        }
      }
      const unsigned ob = (pbase + hro + kbl) * S, oh = (pbase + hro + khl) * S;
      const unsigned o63 = (pbase + hro + k63) * S;
      qhyb = ldo(fHy, ob);
      qhxb = ldo(fHx, ob);
      qex = ldo(fEx, oh);
      qey = ldo(fEy, oh);
      qhx = ldo(fHx, oh);
      qhy = ldo(fHy, oh);
      qhz = ldo(fHz, oh);
      qhzm = ldo(fHz, (pbase + hrom + khl) * S);
      qhy63 = ldo(fHy, o63);
      qhx63 = ldo(fHx, o63);
      qm = ldm(mat, pbase + hro + khl);
    }
    // ---- E of plane p: rows 0..R of the wave's cells (row R: Ex, Ez), and
    // packed, the halo cell of rows 0..R-1 (Ex, Ey)
    {
      const int lxe = (KIND == SW_ANY && p <= nx) ? cpml_layer(true, p, nx, dn0, dp0) : -1;
      const bool slab = lxe >= 0 || any_y || any_z;
      const bool tfp = tf_e && p >= li && p <= ri;
#pragma unroll
      for (int r = 0; r <= R; ++r) {
        const int j = j0 + r;
        const T hxv = hxq[r + 1], hyv = hyq[r], hzv = hzq[r + 1];
        const T hyk = from_below(hyv, lane_value(qhyb, r));
        const bool own = kown && r < R && j <= ny && j < jend && p < i1;
        T vx, vy = T(0), vz;
        if (r < R) {
          const T hxk = from_below(hxv, lane_value(qhxb, r));
          sw_e_cell<T, R, MS, KIND, true, true, true>(c, p, j, k, lane, (int)mq[r], zc, ax[r], ay[r], az[r],
                                                      hyv - hyk, hzv - hzq[r], hxv - hxk, hzv - hz[r],
                                                      hxv - hxq[r], hyv - hy[r], own, slab, tfp, lxe, vx, vy, vz);
          ay[r] = vy;
        } else {
          sw_e_cell<T, R, MS, KIND, true, false, true>(c, p, j, k, lane, (int)mq[r], zc, ax[r], T(0), az[r],
                                                       hyv - hyk, hzv - hzq[r], T(0), T(0), hxv - hxq[r],
                                                       hyv - hy[r], false, slab, tfp, lxe, vx, vy, vz);
        }
        ax[r] = vx;
        az[r] = vz;
      }
      // the halo cell (j0 + hr, k0 + 64), meaningful in lanes 0..R-1
      T wx, wy, wz = T(0);
      sw_e_cell<T, R, MS, KIND, true, true, false>(c, p, hj, kh, SW_KS - 1, (int)qm, zt, qex, qey, T(0),
                                                   qhy - qhy63, qhz - qhzm, qhx - qhx63, qhz - ghz, T(0), T(0),
                                                   false, slab, tfp, lxe, wx, wy, wz);
      qex = wx;
      qey = wy;
    }
    if (i >= i0) {
      // ---- H of plane i from H(i) (set A), E(i) and E(i + 1) (set B)
      const int lxh = KIND == SW_ANY ? cpml_layer(false, i, nx, dn0, dp0) : -1;
      const bool slab = lxh >= 0 || any_y || any_z;
      const bool tfh = tf_h && i >= li - 1 && i <= ri;
      const unsigned pbase = (unsigned)i * si;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const T exv = ex[r], eyv = ey[r], ezv = ez[r];
        const T eyk = from_above(eyv, lane_value(gy, r));
        const T exk = from_above(exv, lane_value(gx, r));
        const bool own = kown && j <= ny && j < jend;
        T vx, vy, vz;
        sw_h_cell<T, R, MS, KIND>(c, i, j, k, lane, (int)mcur[r], zc, hx[r], hy[r], hz[r], eyk - eyv,
                                  ez[r + 1] - ezv, az[r] - ezv, exk - exv, ex[r + 1] - exv, ay[r] - eyv, own, slab,
                                  tfh, lxh, vx, vy, vz);
        if (own) {
          bad |= !(__builtin_isfinite(vx) && __builtin_isfinite(vy) && __builtin_isfinite(vz));
          const unsigned o = (pbase + roff[r + 1] + kl) * S;
          sto(a.dst[EX], o, exv);
          sto(a.dst[EY], o, eyv);
          sto(a.dst[EZ], o, ezv);
          sto(a.dst[HX], o, vx);
          sto(a.dst[HY], o, vy);
          sto(a.dst[HZ], o, vz);
        }
      }
    }
    // ---- plane i + 1 becomes plane i
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      ex[r] = ax[r];
      ez[r] = az[r];
      hy[r] = hyq[r];
      if (r < R) {
        ey[r] = ay[r];
        hx[r] = hxq[r + 1];
        hz[r] = hzq[r + 1];
        mcur[r] = mq[r];
      }
    }
    gx = qex;
    gy = qey;
    ghz = qhz;
  }
  if (bad) atomicMin(a.div, (unsigned long long)a.step);
}

template <typename T, int R, int MS, int KIND>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs<T> a) {
  sweep_body<T, R, MS, KIND>(a, (int)blockIdx.x);
}

// TFSF3D::updateIncidentField (tfsf_3d.cpp:153-183) from the line (e, h) into
// (e2, h2), leaving the old line for the fused step's E corrections; the
// ABC state is updated in place.  One workgroup.
__global__ void __launch_bounds__(1024) k_tfsf_ifa_pp(Tfsf t, double *e2, double *h2, long long step) {
  const int L = t.L;
  const double *e = t.e_inc, *h = t.h_inc;
  const double x = e[L - 2], y = e[L - 1];
  if (threadIdx.x == 0) e2[0] = t.src[step];
  for (int i = 1 + (int)threadIdx.x; i < L - 1; i += (int)blockDim.x)
    e2[i] = t.ceie * e[i] + t.ceihi * (h[i] - h[i - 1]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double a = t.ab[0], b = t.ab[1];
    e2[L - 1] = -a + t.c1 * (e2[L - 2] + b) + t.c2 * (x + y);
    t.ab[0] = x;
    t.ab[1] = y;
  }
  __syncthreads();
  for (int i = (int)threadIdx.x; i < L - 1; i += (int)blockDim.x)
    h2[i] = t.chih * h[i] + t.chiei * (e2[i + 1] - e2[i]);
}

// The TFSF face terms of one step (TfTerms): t.h_inc is the H line before
// this step's IFA (for the E terms), t.e_inc the E line after it (for the H
// terms).  One thread per term, the products exactly as tfsf_e_body and
// tfsf_h_body form them.
__global__ void __launch_bounds__(256) k_tfsf_terms(Tfsf t, TfTerms w) {
  const int s = (int)blockIdx.y;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int li = t.li, ri = t.ri, lj = t.lj, rj = t.rj, lk = t.lk, rk = t.rk;
  const int nj = rj - lj, nk = rk - lk;
  if (r >= tf_seg_size(s, ri - li, nj, nk)) return;
  const long long q = (long long)s * w.stride + r;
  const int side = s & 1;
  double v = 0;
  switch (s) {
    // E: Ey@x [nj][nk+1]: HZ at (li-1 | ri, j, k)
    case 0: case 1: { const int j = lj + (int)(r / (nk + 1)), k = lk + (int)(r % (nk + 1));
                      v = t.ca[0] * tfsf_inc(t, HZ, side ? ri : li - 1, j, k); break; }
    // Ey@z [ni+1][nj]: HX at (i, j, lk-1 | rk)
    case 2: case 3: { const int i = li + (int)(r / nj), j = lj + (int)(r % nj);
                      v = t.ca[2] * tfsf_inc(t, HX, i, j, side ? rk : lk - 1); break; }
    // Ez@x [nj+1][nk]: HY at (li-1 | ri, j, k)
    case 4: case 5: { const int j = lj + (int)(r / nk), k = lk + (int)(r % nk);
                      v = t.ca[0] * tfsf_inc(t, HY, side ? ri : li - 1, j, k); break; }
    // Ez@y [ni+1][nk]: HX at (i, lj-1 | rj, k)
    case 6: case 7: { const int i = li + (int)(r / nk), k = lk + (int)(r % nk);
                      v = t.ca[1] * tfsf_inc(t, HX, i, side ? rj : lj - 1, k); break; }
    // Ex@y [ni][nk+1]: HZ at (i, lj-1 | rj, k)
    case 8: case 9: { const int i = li + (int)(r / (nk + 1)), k = lk + (int)(r % (nk + 1));
                      v = t.ca[1] * tfsf_inc(t, HZ, i, side ? rj : lj - 1, k); break; }
    // Ex@z [ni][nj+1]: HY at (i, j, lk-1 | rk)
    case 10: case 11: { const int i = li + (int)(r / (nj + 1)), j = lj + (int)(r % (nj + 1));
                        v = t.ca[2] * tfsf_inc(t, HY, i, j, side ? rk : lk - 1); break; }
    // H: Hz@x [nj][nk+1]: cb0 * EY at (li | ri, j, k)
    case 12: case 13: { const int j = lj + (int)(r / (nk + 1)), k = lk + (int)(r % (nk + 1));
                        v = t.cb[0] * tfsf_inc(t, EY, side ? ri : li, j, k); break; }
    // Hz@y [ni][nk+1]: cb1 * EX at (i, lj | rj, k)
    case 14: case 15: { const int i = li + (int)(r / (nk + 1)), k = lk + (int)(r % (nk + 1));
                        v = t.cb[1] * tfsf_inc(t, EX, i, side ? rj : lj, k); break; }
    // Hy@x [nj+1][nk]: cb0 * EZ at (li | ri, j, k)
    case 16: case 17: { const int j = lj + (int)(r / nk), k = lk + (int)(r % nk);
                        v = t.cb[0] * tfsf_inc(t, EZ, side ? ri : li, j, k); break; }
    // Hy@z [ni][nj+1]: cb2 * EX at (i, j, lk | rk)
    case 18: case 19: { const int i = li + (int)(r / (nj + 1)), j = lj + (int)(r % (nj + 1));
                        v = t.cb[2] * tfsf_inc(t, EX, i, j, side ? rk : lk); break; }
    // Hx@y [ni+1][nk]: cb1 * EZ at (i, lj | rj, k)
    case 20: case 21: { const int i = li + (int)(r / nk), k = lk + (int)(r % nk);
                        v = t.cb[1] * tfsf_inc(t, EZ, i, side ? rj : lj, k); break; }
    // Hx@z [ni+1][nj]: cb2 * EY at (i, j, lk | rk)
    default: { const int i = li + (int)(r / nj), j = lj + (int)(r % nj);
               v = t.cb[2] * tfsf_inc(t, EY, i, j, side ? rk : lk); break; }
  }
  w.buf[q] = v;
}

}  // namespace xfk
