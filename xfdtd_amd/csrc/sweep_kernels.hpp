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
//   wave_shl); at the ends, lane 0 loads H at k0 - 1 and lane 63 recomputes
//   Ex and Ey at k0 + 64 (the next wave's first cell) from its own loads.
//   j runs over R rows held in registers by each thread, plus one halo row
//   above (j0 + R) whose Ex and Ez are recomputed, and the H values of the row
//   below (j0 - 1).
//   x is marched: at plane i the kernel loads plane i + 1, computes E(i + 1)
//   from H(i), H(i + 1), then H(i) from E(i), E(i + 1), and stores E(i), H(i).
//   The first pass of a chunk (i = i0 - 1) only computes E(i0); the last
//   computes E(i1) as a halo plane.
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

// launch geometry: workgroups of `waves` waves over tiles numbered k chunk
// fastest, then j group, then i chunk.  The hardware deals workgroup ids to
// the 8 XCDs round-robin and each XCD has its own L2; with xcd_runs each XCD
// takes one contiguous run of tile numbers instead, so the tiles that share
// halo rows and the cache lines at k-chunk ends run on the same L2 at about
// the same time.
struct SweepTiling {
  int nkc, njg, nic;
  int tx;     // x planes per chunk
  int waves;  // waves per workgroup, each a strip of R rows
  int nblocks;
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
template <typename T>
struct SweepArgs {
  const T *src[6];  // fields of set A (read)
  T *dst[6];        // fields of set B (written)
  const uint8_t *mat;
  const T *self[2], *kt[2], *jt[2], *it[2];  // curl coefficient records (E, H)
  const T *kp[2], *jp[2], *ip[2];            // psi coefficient pairs
  const T *be[3], *ae[3], *bm[3], *am[3];
  const T *psiA;  // psi of set A: one block, slot q (P_AY..P_CY of E, then H) at q * psi_slot
  T *psiB;        // psi of set B, same layout
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

// CPML layer of index n on one axis and side (-1 outside the slabs)
__device__ __forceinline__ int cpml_layer(bool e_side, int n, int N, int dn, int dp) {
  return e_side ? e_layer(n, N, dn, dp) : h_layer(n, N, dn, dp);
}

// LDS layout of a workgroup's tables, in units of T, with MS material slots
// per record: self records of both sides, curl pairs over the k slots (the
// 64 lanes and the cell above), j slots (the workgroup's rows and the top
// halo row) and i slots (up to SW_TXMAX planes and the halo plane), the same
// for the psi pairs, then the x and y CPML b/a values.  All offsets are
// compile-time constants.
constexpr int SW_KS = 65;
constexpr int SW_WAVES = 4;
constexpr int SW_TXMAX = 32;
template <int R, int MS>
struct SweepLds {
  static constexpr int js = SW_WAVES * R + 1, is = SW_TXMAX + 1;
  static constexpr int per = 2 * MS * (SW_KS + js + is);
  static constexpr int self(int side) { return side * 4 * MS; }
  static constexpr int pair(int side, int psi, int axis) {
    return 8 * MS + (2 * side + psi) * per + (axis >= 1 ? 2 * MS * SW_KS : 0) + (axis >= 2 ? 2 * MS * js : 0);
  }
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

template <typename T, int R, int MS>
__device__ __forceinline__ void sweep_body(const SweepArgs<T> &a, int block) {
  extern __shared__ __align__(16) unsigned char sw_lds[];
  const int nx = a.nx, ny = a.ny, nz = a.nz, pz = a.pz, nmat = a.nmat;
  const SweepTiling st = a.st;
  int tile = block;
  if (st.xcd_runs) {
    const int q = st.nblocks >> 3, rr = st.nblocks & 7, xcd = block & 7;
    tile = xcd * q + min(xcd, rr) + (block >> 3);
  }
  const int kc = tile % st.nkc;
  const int rest = tile / st.nkc;
  const int jg = rest % st.njg;
  const int ic = rest / st.njg;
  const int lane = (int)(threadIdx.x & 63u);
  const int wv = (int)(threadIdx.x >> 6);
  const int jwg = jg * st.waves * R;  // first row of the workgroup
  const int j0 = jwg + wv * R;
  const int i0 = ic * st.tx;
  const int i1 = min(i0 + st.tx, nx + 1);
  const int k0 = kc * 64;
  const int k = k0 + lane;
  const int dn0 = a.dn[0], dp0 = a.dp[0], dn1 = a.dn[1], dp1 = a.dp[1], dn2 = a.dn[2], dp2 = a.dp[2];
  const int Lx = dn0 + dp0, Ly = dn1 + dp1, Lz = dn2 + dp2, nm1y = ny + 1;
  const bool tf = a.tf != 0;

  // ---- stage the tile's tables in LDS
  using L = SweepLds<R, MS>;
  T *lds = reinterpret_cast<T *>(sw_lds);
  // psi pairs and CPML values only when the tile meets a slab
  const bool tile_slab = (Lx > 0 && meets_slab(i0, i1, nx, dn0, dp0)) ||
                         (Ly > 0 && meets_slab(jwg, jwg + L::js - 1, ny, dn1, dp1)) ||
                         (Lz > 0 && meets_slab(k0 - 1, k0 + 64, nz, dn2, dp2));
  {
    const int nth = (int)blockDim.x, t0 = (int)threadIdx.x;
    for (int q = t0; q < 8 * nmat; q += nth) {
      const int side = q / (4 * nmat), r = q % (4 * nmat);
      lds[L::self(side) + (r / 4) * 4 + r % 4] = a.self[side][r];
    }
    // records are [slot][material][2] with MS material slots per k, j, i slot
    const int npsi = tile_slab ? 2 : 1;
    for (int side = 0; side < 2; ++side)
      for (int psi = 0; psi < npsi; ++psi) {
        const T *tk = psi ? a.kp[side] : a.kt[side];
        const T *tj = psi ? a.jp[side] : a.jt[side];
        const T *ti = psi ? a.ip[side] : a.it[side];
        for (int q = t0; q < 2 * nmat * SW_KS; q += nth) {
          const int sl = q / (2 * nmat), r = q % (2 * nmat);
          lds[L::pair(side, psi, 0) + sl * 2 * MS + r] = tk[min(k0 + sl, nz) * 2 * nmat + r];
        }
        for (int q = t0; q < 2 * nmat * L::js; q += nth) {
          const int sl = q / (2 * nmat), r = q % (2 * nmat);
          lds[L::pair(side, psi, 1) + sl * 2 * MS + r] = tj[min(jwg + sl, ny) * 2 * nmat + r];
        }
        for (int q = t0; q < 2 * nmat * L::is; q += nth) {
          const int sl = q / (2 * nmat), r = q % (2 * nmat);
          lds[L::pair(side, psi, 2) + sl * 2 * MS + r] = ti[min(i0 + sl, nx) * 2 * nmat + r];
        }
      }
    if (tile_slab) {
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
  if (j0 > ny) return;  // the whole wave lies beyond the grid

  // ---- per-lane constants: this lane's cell k, and for lane 63 the cell
  // above the wave (kh), for lane 0 the cell below (kb)
  const int kl = min(k, pz - 1);
  const int kh = k0 + 64, khl = min(kh, pz - 1);
  const int kbl = max(k0 - 1, 0);
  const bool top = lane == 63, bot = lane == 0;
  const bool kown = k <= nz;
  const bool kin_e = k >= 1 && k < nz, klt = k < nz;
  const bool kin_eh = kh < nz;  // kh >= 64
  auto zlayer = [&](int kk) {
    ZLay<T> z;
    const bool v = kk <= nz;
    z.le = v ? cpml_layer(true, kk, nz, dn2, dp2) : -1;
    z.lh = v ? cpml_layer(false, kk, nz, dn2, dp2) : -1;
    z.be = z.ae = z.bh = z.ah = T(0);
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
  const ZLay<T> zt = zlayer(kh);  // the halo cell (used by lane 63)
  const unsigned sj = a.sj, si = a.si;
  const T *__restrict__ fEx = a.src[EX];
  const T *__restrict__ fEy = a.src[EY];
  const T *__restrict__ fEz = a.src[EZ];
  const T *__restrict__ fHx = a.src[HX];
  const T *__restrict__ fHy = a.src[HY];
  const T *__restrict__ fHz = a.src[HZ];
  const uint8_t *__restrict__ mat = a.mat;
  const P4<T> *selfE = reinterpret_cast<const P4<T> *>(lds + L::self(0));
  const P4<T> *selfH = reinterpret_cast<const P4<T> *>(lds + L::self(1));
  const P2<T> *kE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 0, 0));
  const P2<T> *jE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 0, 1));
  const P2<T> *iE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 0, 2));
  const P2<T> *kH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 0, 0));
  const P2<T> *jH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 0, 1));
  const P2<T> *iH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 0, 2));
  const P2<T> *kpE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 1, 0));
  const P2<T> *jpE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 1, 1));
  const P2<T> *ipE = reinterpret_cast<const P2<T> *>(lds + L::pair(0, 1, 2));
  const P2<T> *kpH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 1, 0));
  const P2<T> *jpH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 1, 1));
  const P2<T> *ipH = reinterpret_cast<const P2<T> *>(lds + L::pair(1, 1, 2));
  const T *bax = lds + L::ba_x;
  const T *bay = lds + L::ba_y;
  // psi slot offsets (elements) in a block
  const unsigned psl = a.psi_slot;
  const unsigned o_ay = 0, o_az = psl, o_bx = 2 * psl, o_bz = 3 * psl, o_cx = 4 * psl, o_cy = 5 * psl, o_h = 6 * psl;
  const T *__restrict__ pA = a.psiA;
  T *__restrict__ pB = a.psiB;
  // does any row of this wave, or any lane (or the halo cell), meet a y / z
  // slab?  (wave-uniform)
  const bool any_z = Lz > 0 && __ballot(zc.le >= 0 || zc.lh >= 0 || zt.le >= 0) != 0ull;
  const bool any_y = Ly > 0 && meets_slab(j0, j0 + R, ny, dn1, dp1);
  // TFSF: the E corrections touch cells of [li, ri] x [lj, rj] x [lk, rk]
  // on its faces, the H corrections cells of [li-1, ri] x [lj-1, rj] x
  // [lk-1, rk] on its outer faces.  Wave-uniform: does the wave's (j, k)
  // footprint (the halo cell included) meet those boxes?
  int li = 0, ri = 0, lj = 0, rj = 0, lk = 0, rk = 0;
  bool tf_e = false, tf_h = false;
  if (tf) {
    li = a.tb[0]; ri = a.tb[1]; lj = a.tb[2]; rj = a.tb[3]; lk = a.tb[4]; rk = a.tb[5];
    tf_e = j0 <= rj && j0 + R >= lj && k0 <= rk && k0 + 64 >= lk;
    tf_h = j0 <= rj && j0 + R - 1 >= lj - 1 && k0 <= rk && k0 + 63 >= lk - 1;
  }
  const long long nj = rj - lj, nk = rk - lk;
  // TFSF term segment s of this step
  const unsigned tstr = a.term_stride;
  auto seg = [&](int sg) { return a.terms + (size_t)sg * tstr; };

  // element offset of row r - 1 within a plane (clamped), r in [0, R + 1]
  unsigned roff[R + 2];
#pragma unroll
  for (int r = -1; r <= R; ++r) roff[r + 1] = (unsigned)min(max(j0 + r, 0), ny) * sj;

  // One E cell (p, j, kk): Simulation::updateE + PML::updateE + TFSF3D::
  // updateE, in the order of compute_plane_e and tfsf_e_body.  d*: the
  // curl differences; (ps, own): psi slot offset within a plane of the cell,
  // whether this thread stores its psi; yrow: Ey exists (not the halo row);
  // zrow: Ez is wanted (not the halo cell).
  auto e_cell = [&](int p, int j, int kk, int kslot, int m, const ZLay<T> &z, bool kin, bool klt_, T exv, T eyv,
                    T ezv, T dhy_k, T dhz_j, T dhx_k, T dhz_i, T dhx_j, T dhy_i, bool own, bool yrow, bool zrow,
                    bool slab, bool tfp, int lxe, T &vx, T &vy, T &vz) {
    const int pl = min(p, nx);  // tables and psi of a plane beyond nx are never used
    const int is = p - i0, js = j - jwg;
    const bool jvalid = j <= ny;
    const bool jin = j >= 1 && j < ny, jlt = j < ny;
    const bool ix = p >= 1 && p < nx, plt = p < nx;
    const P4<T> cs = selfE[m];
    const P2<T> ck = kE[kslot * MS + m], cj = jE[js * MS + m], ci = iE[is * MS + m];
    const T yx = cs.x * exv + ck.x * dhy_k + cj.x * dhz_j;
    const T yy = cs.y * eyv + ck.y * dhx_k + ci.x * dhz_i;
    const T yz = cs.z * ezv + cj.y * dhx_j + ci.y * dhy_i;
    vx = (plt && jin && kin) ? yx : exv;
    vy = (ix && jlt && kin) ? yy : eyv;
    vz = (ix && jin && klt_) ? yz : ezv;
    if (slab) {
      const int lye = jvalid ? cpml_layer(true, j, ny, dn1, dp1) : -1;
      const T bye = bay[js], aye = bay[L::js + js];
      const T bxe = bax[is], axe = bax[L::is + is];
      const unsigned jc = (unsigned)min(j, ny);
      const unsigned kq = (unsigned)min(kk, pz - 1);
      if (plt) {
        if (lye >= 0) {
          const unsigned o = o_ay + psi_y(pl, lye, kq, Ly, pz);
          const T ps = bye * pA[o] + aye * dhz_j;
          if (own) pB[o] = ps;
          vx = vx + jpE[js * MS + m].x * ps;
        }
        if (z.le >= 0) {
          const unsigned o = o_az + psi_z(pl, jc, z.le, nm1y, Lz);
          const T ps = z.be * pA[o] + z.ae * dhy_k;
          if (own) pB[o] = ps;
          vx = vx + kpE[kslot * MS + m].x * ps;
        }
      }
      if (jlt && yrow) {
        if (lxe >= 0) {
          const unsigned o = o_bx + psi_x(lxe, jc, kq, nm1y, pz);
          const T ps = bxe * pA[o] + axe * dhz_i;
          if (own) pB[o] = ps;
          vy = vy + ipE[is * MS + m].x * ps;
        }
        if (z.le >= 0) {
          const unsigned o = o_bz + psi_z(pl, jc, z.le, nm1y, Lz);
          const T ps = z.be * pA[o] + z.ae * dhx_k;
          if (own) pB[o] = ps;
          vy = vy + kpE[kslot * MS + m].y * ps;
        }
      }
      if (klt_ && zrow) {
        if (lxe >= 0) {
          const unsigned o = o_cx + psi_x(lxe, jc, kq, nm1y, pz);
          const T ps = bxe * pA[o] + axe * dhy_i;
          if (own) pB[o] = ps;
          vz = vz + ipE[is * MS + m].y * ps;
        }
        if (lye >= 0) {
          const unsigned o = o_cy + psi_y(pl, lye, kq, Ly, pz);
          const T ps = bye * pA[o] + aye * dhx_j;
          if (own) pB[o] = ps;
          vz = vz + jpE[js * MS + m].y * ps;
        }
      }
    }
    // TFSF E corrections (tfsf_e_body), in FP64, on face cells only
    if (tfp && j >= lj && j <= rj) {
      const bool kb = kk >= lk && kk <= rk;
      // Ey: j in [lj, rj), (i, k) on the perimeter: x faces then z faces
      if (yrow && j < rj && kb && (p == li || p == ri || kk == lk || kk == rk)) {
        double v = vy;
        const long long ox = (long long)(j - lj) * (nk + 1) + (kk - lk), oz = (long long)(p - li) * nj + (j - lj);
        if (p == li) v = v + seg(0)[ox];
        if (p == ri) v = v - seg(1)[ox];
        if (kk == lk) v = v - seg(2)[oz];
        if (kk == rk) v = v + seg(3)[oz];
        vy = (T)v;
      }
      // Ez: k in [lk, rk), (i, j) on the perimeter: x faces then y faces
      if (zrow && kk >= lk && kk < rk && (p == li || p == ri || j == lj || j == rj)) {
        double v = vz;
        const long long ox = (long long)(j - lj) * nk + (kk - lk), oy = (long long)(p - li) * nk + (kk - lk);
        if (p == li) v = v - seg(4)[ox];
        if (p == ri) v = v + seg(5)[ox];
        if (j == lj) v = v + seg(6)[oy];
        if (j == rj) v = v - seg(7)[oy];
        vz = (T)v;
      }
      // Ex: i in [li, ri), (j, k) on the perimeter: y faces then z faces
      if (p < ri && kb && (j == lj || j == rj || kk == lk || kk == rk)) {
        double v = vx;
        const long long oy = (long long)(p - li) * (nk + 1) + (kk - lk), oz = (long long)(p - li) * (nj + 1) + (j - lj);
        if (j == lj) v = v - seg(8)[oy];
        if (j == rj) v = v + seg(9)[oy];
        if (kk == lk) v = v + seg(10)[oz];
        if (kk == rk) v = v - seg(11)[oz];
        vx = (T)v;
      }
    }
  };

  // carried between planes (plane i): H of rows 0..R-1 (Hy also row R), E
  // of rows 0..R (Ey rows 0..R-1), and for lane 63 the halo cell's E (Ex,
  // Ey) and H Hz of rows 0..R-1
  T hx[R], hy[R + 1], hz[R];
  T ex[R + 1], ey[R], ez[R + 1];
  T gx[R], gy[R], ghz[R];
  unsigned mcur[R];
  // plane i + 1
  T ax[R + 1], ay[R], az[R + 1];
  T hxq[R + 2], hyq[R + 2], hzq[R + 2];
  T hyb[R + 1], hxb[R];             // lane 0: Hy, Hx at k0 - 1
  T egx[R], egy[R], hgx[R], hgy[R], hgz[R + 1];  // lane 63: the halo cell (hgz from row -1)
  unsigned mq[R + 1], mg[R];
#pragma unroll
  for (int r = 0; r <= R; ++r) {
    hy[r] = T(0);
    if (r < R) hz[r] = hx[r] = ghz[r] = gx[r] = gy[r] = T(0);
  }
  // before the first plane: H(i0 - 1), zero before plane 0
  if (i0 >= 1) {
    const unsigned pbase = (unsigned)(i0 - 1) * si;
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      const unsigned c = pbase + roff[r + 1];
      hy[r] = ldo(fHy, (c + kl) * (unsigned)sizeof(T));
      if (r < R) {
        hz[r] = ldo(fHz, (c + kl) * (unsigned)sizeof(T));
        if (top) ghz[r] = ldo(fHz, (c + khl) * (unsigned)sizeof(T));
      }
    }
  }
  bool bad = false;  // a non-finite H value was stored
  // i = i0 - 1 only computes E(i0); from i0 on, each pass computes E(i + 1)
  // and H(i) and stores plane i
  for (int i = i0 - 1; i < i1; ++i) {
    const int p = i + 1;
    // ---- loads of plane p (clamped to nx) from set A
    {
      const unsigned pbase = (unsigned)min(p, nx) * si;
#pragma unroll
      for (int r = -1; r <= R; ++r) {
        const unsigned c = pbase + roff[r + 1];
        const unsigned o = (c + kl) * (unsigned)sizeof(T);
        hxq[r + 1] = ldo(fHx, o);
        hzq[r + 1] = ldo(fHz, o);
        if (top && r < R) hgz[r + 1] = ldo(fHz, (c + khl) * (unsigned)sizeof(T));
        if (r >= 0) {
          hyq[r + 1] = ldo(fHy, o);
          ax[r] = ldo(fEx, o);
          az[r] = ldo(fEz, o);
          mq[r] = ldm(mat, c + kl);
          hyb[r] = T(0);
          if (bot) hyb[r] = ldo(fHy, (c + kbl) * (unsigned)sizeof(T));
          if (r < R) {
            ay[r] = ldo(fEy, o);
            hxb[r] = T(0);
            egx[r] = egy[r] = hgx[r] = hgy[r] = T(0);
            mg[r] = 0;
            if (bot) hxb[r] = ldo(fHx, (c + kbl) * (unsigned)sizeof(T));
            if (top) {
              const unsigned og = (c + khl) * (unsigned)sizeof(T);
              egx[r] = ldo(fEx, og);
              egy[r] = ldo(fEy, og);
              hgx[r] = ldo(fHx, og);
              hgy[r] = ldo(fHy, og);
              mg[r] = ldm(mat, c + khl);
            }
          }
        }
      }
    }
    // ---- E of plane p: rows 0..R of the wave's cells, rows 0..R-1 of the
    // halo cell (lane 63)
    {
      const int lxe = p <= nx ? cpml_layer(true, p, nx, dn0, dp0) : -1;
      const bool slab = lxe >= 0 || any_y || any_z;
      const bool tfp = tf_e && p >= li && p <= ri;
#pragma unroll
      for (int r = 0; r <= R; ++r) {
        const int j = j0 + r;
        const T hxv = hxq[r + 1], hyv = hyq[r + 1], hzv = hzq[r + 1];
        const T hyk = from_below(hyv, hyb[r]);
        const T hxk = from_below(hxv, r < R ? hxb[r] : T(0));
        const T eyv = r < R ? ay[r] : T(0);
        const T dhz_i = r < R ? hzv - hz[r] : T(0);
        const bool own = kown && r < R && j <= ny && p < i1;
        T vx, vy, vz;
        e_cell(p, j, k, lane, (int)mq[r], zc, kin_e, klt, ax[r], eyv, az[r], hyv - hyk, hzv - hzq[r], hxv - hxk,
               dhz_i, hxv - hxq[r], hyv - hy[r], own, r < R, true, slab, tfp, lxe, vx, vy, vz);
        if (r < R) {
          // the halo cell k0 + 64 of this row (meaningful in lane 63)
          T wx, wy, wz;
          const T hzjm = hgz[r];  // Hz(j - 1) at the halo cell
          e_cell(p, j, kh, 64, (int)mg[r], zt, kin_eh, kh < nz, egx[r], egy[r], T(0), hgy[r] - hyv,
                 hgz[r + 1] - hzjm, hgx[r] - hxv, hgz[r + 1] - ghz[r], T(0), T(0), false, true, false, slab, tfp,
                 lxe, wx, wy, wz);
          egx[r] = wx;
          egy[r] = wy;
          ay[r] = vy;
        }
        ax[r] = vx;
        az[r] = vz;
      }
    }
    if (i >= i0) {
      // ---- H of plane i from H(i) (set A), E(i) and E(i + 1) (set B)
      const int lxh = cpml_layer(false, i, nx, dn0, dp0);
      const bool slab = lxh >= 0 || any_y || any_z;
      const bool ilt = i < nx;
      const int is = i - i0;
      const bool tfh = tf_h && i >= li - 1 && i <= ri;
      const unsigned pbase = (unsigned)i * si;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r;
        const bool jvalid = j <= ny;
        const bool jlt = j < ny;
        const int m = (int)mcur[r];
        const int js = j - jwg;
        const T hxv = hx[r], hyv = hy[r], hzv = hz[r];
        const T exv = ex[r], eyv = ey[r], ezv = ez[r];
        const T eyk = from_above(eyv, gy[r]);
        const T exk = from_above(exv, gx[r]);
        const T dey_k = eyk - eyv;
        const T dez_j = ez[r + 1] - ezv;
        const T dez_i = az[r] - ezv;
        const T dex_k = exk - exv;
        const T dex_j = ex[r + 1] - exv;
        const T dey_i = ay[r] - eyv;
        const P4<T> cs = selfH[m];
        const P2<T> ck = kH[lane * MS + m], cj = jH[js * MS + m], ci = iH[is * MS + m];
        const T yx = cs.x * hxv + ck.x * dey_k + cj.x * dez_j;
        const T yy = cs.y * hyv + ci.x * dez_i + ck.y * dex_k;
        const T yz = cs.z * hzv + cj.y * dex_j + ci.y * dey_i;
        const bool yee = ilt && jlt && klt;
        T vx = yee ? yx : hxv;
        T vy = yee ? yy : hyv;
        T vz = yee ? yz : hzv;
        const bool own = kown && jvalid;
        if (slab) {
          const int lyh = jvalid ? cpml_layer(false, j, ny, dn1, dp1) : -1;
          const T byh = bay[2 * L::js + js], ayh = bay[3 * L::js + js];
          const T bxh = bax[2 * L::is + is], axh = bax[3 * L::is + is];
          const unsigned jc = (unsigned)min(j, ny);
          if (jlt && klt) {
            if (lyh >= 0) {
              const unsigned o = o_h + o_ay + psi_y(i, lyh, kl, Ly, pz);
              const T ps = byh * pA[o] + ayh * dez_j;
              if (own) pB[o] = ps;
              vx = vx + jpH[js * MS + m].x * ps;
            }
            if (zc.lh >= 0) {
              const unsigned o = o_h + o_az + psi_z(i, jc, zc.lh, nm1y, Lz);
              const T ps = zc.bh * pA[o] + zc.ah * dey_k;
              if (own) pB[o] = ps;
              vx = vx + kpH[lane * MS + m].x * ps;
            }
          }
          if (ilt && klt) {
            if (lxh >= 0) {
              const unsigned o = o_h + o_bx + psi_x(lxh, jc, kl, nm1y, pz);
              const T ps = bxh * pA[o] + axh * dez_i;
              if (own) pB[o] = ps;
              vy = vy + ipH[is * MS + m].x * ps;
            }
            if (zc.lh >= 0) {
              const unsigned o = o_h + o_bz + psi_z(i, jc, zc.lh, nm1y, Lz);
              const T ps = zc.bh * pA[o] + zc.ah * dex_k;
              if (own) pB[o] = ps;
              vy = vy + kpH[lane * MS + m].y * ps;
            }
          }
          if (ilt && jlt) {
            if (lxh >= 0) {
              const unsigned o = o_h + o_cx + psi_x(lxh, jc, kl, nm1y, pz);
              const T ps = bxh * pA[o] + axh * dey_i;
              if (own) pB[o] = ps;
              vz = vz + ipH[is * MS + m].y * ps;
            }
            if (lyh >= 0) {
              const unsigned o = o_h + o_cy + psi_y(i, lyh, kl, Ly, pz);
              const T ps = byh * pA[o] + ayh * dex_j;
              if (own) pB[o] = ps;
              vz = vz + jpH[js * MS + m].y * ps;
            }
          }
        }
        // TFSF H corrections (tfsf_h_body), in FP64; no H cell takes two
        if (tfh && j >= lj - 1 && j <= rj) {
          // Hz: x faces (j in [lj, rj), k in [lk, rk]), y faces (i in [li, ri), k in [lk, rk])
          if (k >= lk && k <= rk) {
            if (j >= lj && j < rj && (i == li - 1 || i == ri)) {
              const double t = seg(i == ri ? 13 : 12)[(long long)(j - lj) * (nk + 1) + (k - lk)];
              vz = (T)(i == ri ? (double)vz - t : (double)vz + t);
            } else if (i >= li && i < ri && (j == lj - 1 || j == rj)) {
              const double t = seg(j == rj ? 15 : 14)[(long long)(i - li) * (nk + 1) + (k - lk)];
              vz = (T)(j == rj ? (double)vz + t : (double)vz - t);
            }
          }
          // Hy: x faces (j in [lj, rj], k in [lk, rk)), z faces (i in [li, ri), j in [lj, rj])
          if (j >= lj) {
            if (k >= lk && k < rk && (i == li - 1 || i == ri)) {
              const double t = seg(i == ri ? 17 : 16)[(long long)(j - lj) * nk + (k - lk)];
              vy = (T)(i == ri ? (double)vy + t : (double)vy - t);
            } else if (i >= li && i < ri && (k == lk - 1 || k == rk)) {
              const double t = seg(k == rk ? 19 : 18)[(long long)(i - li) * (nj + 1) + (j - lj)];
              vy = (T)(k == rk ? (double)vy - t : (double)vy + t);
            }
          }
          // Hx: y faces (k in [lk, rk), i in [li, ri]), z faces (i in [li, ri], j in [lj, rj))
          if (i >= li && i <= ri) {
            if (k >= lk && k < rk && (j == lj - 1 || j == rj)) {
              const double t = seg(j == rj ? 21 : 20)[(long long)(i - li) * nk + (k - lk)];
              vx = (T)(j == rj ? (double)vx - t : (double)vx + t);
            } else if (j >= lj && j < rj && (k == lk - 1 || k == rk)) {
              const double t = seg(k == rk ? 23 : 22)[(long long)(i - li) * nj + (j - lj)];
              vx = (T)(k == rk ? (double)vx + t : (double)vx - t);
            }
          }
        }
        if (own) {
          bad |= !(__builtin_isfinite(vx) && __builtin_isfinite(vy) && __builtin_isfinite(vz));
          const unsigned o = (pbase + roff[r + 1] + kl) * (unsigned)sizeof(T);
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
      hy[r] = hyq[r + 1];
      if (r < R) {
        ey[r] = ay[r];
        hx[r] = hxq[r + 1];
        hz[r] = hzq[r + 1];
        mcur[r] = mq[r];
        gx[r] = egx[r];
        gy[r] = egy[r];
        ghz[r] = hgz[r + 1];
      }
    }
  }
  if (bad) atomicMin(a.div, (unsigned long long)a.step);
}

template <typename T, int R, int MS>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs<T> a) {
  sweep_body<T, R, MS>(a, (int)blockIdx.x);
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
