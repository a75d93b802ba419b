// s3g_kernels.hip -- gfx950 kernels of the pSi3D time step (everything fd,
// settrap, settrap2 and save do except the free-surface solve).
//
// One lane owns one wet column c and walks it in k; a 256-lane workgroup covers
// 256 consecutive columns.  Column-local recurrences (Thomas sweeps, the
// cumulative baroclinic term, continuity's bottom-up integration) stay in
// registers; the forward-sweep factors go to a per-column scratch plane stack
// in the same coalesced layout and are read back by the same lane.
#include <algorithm>
#include <type_traits>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "s3g_internal.h"

namespace s3g {

#define IX(k, c) ((size_t)(k) * (size_t)D.LP + (size_t)(c))

constexpr double G = 9.806, RHOAIR = 1.3, KAPPAS = 0.41, PI = 3.1415926535897932;
constexpr double SED_H = 0.5, AV_MOL = 1.3e-6, DV_MOL = 1.4e-7;
constexpr double Q2_MIN = 2.e-6, Q2L_MIN = 6.80236723501e-9, GH_MAX = 2.9e-2, GH_MIN = -2.8e-1;
constexpr double B_1 = 16.6, E_1 = 1.80, E_2 = 1.33, E_3 = 0.25, SQ = 0.20;

__device__ __forceinline__ int col() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }

// Per-cell kernels: a lane owns one column and a chunk of `kc` consecutive
// layers (column-invariant index loads are shared by the chunk).  The tile
// word packs kc (low 8 bits) and the order of the 1-D grid of (column block,
// layer chunk) tiles: order 0 runs layer chunk by layer chunk (all column
// blocks of a chunk are in flight together, so the E/W neighbour rows of a
// plane are read by concurrently running workgroups); orders 1-3 remap the
// grid so that each of the 8 XCDs (workgroups are dispatched round-robin over
// them) runs a contiguous range of column blocks with all their layer chunks,
// either block by block (1) or in panels of 8 (2) or 16 (3) column blocks
// walked chunk by chunk, so that a workgroup's east/west neighbour rows were
// fetched into the same XCD's L2 by a workgroup a few slots earlier.  The
// autotuner picks both per kernel; results do not depend on them.
// Each kernel has a chunked instantiation (CH) and a single-layer one used for
// chunk size 1: a loop lets the compiler hoist the column's address arithmetic,
// which can cost more registers (occupancy) than it saves.
// The scalar a transport kernel advances: salinity/temperature (salpp -> sal)
// or one tracer (tracerpp -> tracer).
template <typename R>
struct Scal {
  const R *cpp;
  R *cn;
  const R *src;  // source added to the right-hand side (HeatSource), or null
};

struct CellTile {
  int c, k0, k1;  // column, layers [k0, k1)
};

__device__ __forceinline__ bool cell_tile(int lm, int k_first, int nk, int tile, CellTile &t) {
  const int kc = tile & 255, order = tile >> 8;
  const int nkc = (nk + kc - 1) / kc;
  const int nbx = (lm + kBlock - 1) / kBlock;
  const int pid = blockIdx.x;
  int bx, kq;
  if (order == 0) {
    kq = pid / nbx;
    bx = pid - kq * nbx;
  } else {
    const int total = nbx * nkc;
    const int xcd = pid & 7, idx = pid >> 3;
    const int per = total >> 3, rem = total & 7;
    const int logical = (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    if (order == 1) {
      bx = logical / nkc;
      kq = logical - bx * nkc;
    } else {
      // panels of P column blocks, walked layer chunk by layer chunk
      const int P = (order == 2) ? 8 : 16;
      const int panel = logical / (P * nkc);
      const int r = logical - panel * P * nkc;
      const int Pp = min(P, nbx - panel * P);
      kq = r / Pp;
      bx = panel * P + (r - kq * Pp);
    }
  }
  t.c = bx * kBlock + threadIdx.x;
  t.k0 = k_first + kq * kc;
  t.k1 = min(t.k0 + kc, k_first + nk);
  return t.c < lm;
}

// UNESCO/Gill EOS with the fixed-point pressure iteration (si3d_utils.f90:3865-3949),
// evaluated in FP64 in every precision mode.  salinity**(3/2) is salinity**1 in the
// reference (integer exponent 3/2).
__device__ double densty(double t, double s, double elev) {
  double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  double densw = 999.842594 + 6.793952e-2 * t - 9.095290e-3 * t2 + 1.001685e-4 * t3 -
                 1.120083e-6 * t4 + 6.536332e-9 * t5;
  double densws = densw + s * (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t2 - 8.2467e-7 * t3 + 5.3875e-9 * t4) +
                  s * (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t2) + 4.8314e-4 * s * s;
  double kw = 19652.21 + 148.4206 * t - 2.327105 * t2 + 1.360477e-2 * t3 - 5.155288e-5 * t4;
  double ks = kw + s * (54.6746 - 0.603459 * t + 1.09987e-2 * t2 - 6.1670e-5 * t3) +
              s * (7.944e-2 + 1.6483e-2 * t - 5.3009e-4 * t2);
  double kp1 = 3.239908 + 1.43713e-3 * t + 1.16092e-4 * t2 - 5.77905e-7 * t3 +
               s * (2.2838e-3 - 1.0981e-5 * t - 1.6078e-6 * t2) + 1.91075e-4 * s;
  double kp2 = (8.50935e-5 - 6.12293e-6 * t + 5.2787e-8 * t2) + s * (-9.9348e-7 + 2.0816e-8 * t + 9.1697e-10 * t2);
  double guess = densws, rho = densws, delta = 1.0;
  for (int it = 0; delta > 1e-6 && it < 10000; ++it) {
    double p = 1e-5 * (guess * 9.806 * elev);
    double k = ks + p * kp1 + p * p * kp2;
    rho = densws / (1.0 - p / k);
    delta = fabs(rho - guess);
    guess = rho;
  }
  return rho;
}

// Per-lake data of an ensemble context (D.lake != null); one lake otherwise.
template <typename R>
__device__ __forceinline__ const double *zlv(const Dev<R> &D, int c) {
  return D.lake ? D.zlevel + (size_t)D.lake[c] * D.KPz : D.zlevel;
}
template <typename R>
__device__ __forceinline__ int km1c(const Dev<R> &D, int c) {
  return D.lake ? D.lake_km1[D.lake[c]] : D.km1;
}
struct Wind {
  double uair, vair, cdw;
};
template <typename R>
__device__ __forceinline__ Wind wind(const Dev<R> &D, int c) {
  if (D.wind) return {D.wind[3 * c], D.wind[3 * c + 1], D.wind[3 * c + 2]};  // met-station modes
  if (D.force) {
    const double *f = D.force + (D.lake ? 3 * D.lake[c] : 0);
    return {f[0], f[1], f[2]};
  }
  return {D.uair, D.vair, D.cdw};
}

template <typename R>
__device__ __forceinline__ double zdens(const Dev<R> &D, int k, int c) {
  double zl = zlv(D, c)[k];
  return (zl == -100.0) ? 0.5 * (double)D.hp[IX(k, c)] : zl + 0.5 * (double)D.hp[IX(k, c)];
}

// ---------------------------------------------------------------------------
// Horizontal eddy coefficient of the momentum stencil at column w, layer k
// (exmom's haxpp / haypp, si3d_procedures.f90:632-690 and 966-1026), evaluated
// on the fly.  X: x-momentum (u-faces), else y-momentum.  DX: the x-direction
// coefficient (haxpp), else haypp.
template <typename R, bool X, bool DX>
__device__ __forceinline__ double hcoef(const Dev<R> &D, int k, int w) {
  const int S = D.lm;
  if (w == S) return 0.0;
  const int wn = X ? D.lE[w] : D.lN[w];  // the face neighbour of w
  if (wn == S) return 0.0;
  const int k1q = X ? D.k1u[w] : D.k1v[w];
  const int kmq = min(D.kmz[wn], D.kmz[w]);
  if (k < k1q || k > kmq) return 0.0;
  const R *hq = X ? D.hupp : D.hvpp;
  const int other = DX ? D.lE[w] : D.lN[w];
  double hmin = dmin((double)hq[IX(k, other)], (double)hq[IX(k, w)]);
  if (D.P.ihd == 1) return (DX ? D.P.Ax0 : D.P.Ay0) * hmin;
  if (D.P.ihd > 1) {
    const bool single = (X && DX) || (!X && !DX);
    if (single) return (double)D.kh[IX(k, other)] * hmin;
    double k4 = ((double)D.kh[IX(k, D.lE[w])] + (double)D.kh[IX(k, w)] + (double)D.kh[IX(k, D.lN[w])] +
                 (double)D.kh[IX(k, D.lE[D.lN[w]])]) * 0.25;
    return k4 * hmin;
  }
  return 0.0;
}

// exmom: explicit advection, Coriolis and horizontal diffusion of the x (X) or
// y momentum equation, si3d_procedures.f90:578-1173.  One lane per (column,
// layer, k = 2..km in chunks, see cell_tile).  ex can only
// ever be non-zero on k1 <= k <= kmq (kmq = bottom of the face, fixed), so
// layers below kmq are never touched (they stay zero from allocation) and the
// zeroing of the reference (645-656, 821-827) reduces to k1 <= k < k1q.
// Explicit term of the momentum face of column c at level k1q <= k <= kmq
// (exmom, 795-950 / 1031-1160).
template <typename R, bool X>
__device__ __forceinline__ double exmom_value(const Dev<R> &D, int c, int k, int k1q, int kmq, double twodt1) {
  const int lE = D.lE[c], ln = X ? lE : D.lN[c];
  const int lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const R *q = X ? D.up : D.vp;
  const R *qpp = X ? D.upp : D.vpp;
  const R *hq = X ? D.hup : D.hvp;
  const int l2 = X ? D.lS[lE] : D.lW[lN];  // SE corner (x) / NW corner (y)
  const double f = D.P.f;
  {
    if ((double)D.hp[IX(k, c)] <= 0.0 || (double)D.hp[IX(k, ln)] <= 0.0)
      return X ? (double)D.uhpp[IX(k, c)] : (double)D.vhpp[IX(k, c)];
    double cor, uE, uW, vN, vS, wU, wD;
    if (X) {
      cor = 0.25 * f * ((double)D.vhp[IX(k, lE)] + (double)D.vhp[IX(k, c)] + (double)D.vhp[IX(k, l2)] + (double)D.vhp[IX(k, lS)]);
      uE = (double)D.uhp[IX(k, lE)] + (double)D.uhp[IX(k, c)];
      uW = (double)D.uhp[IX(k, c)] + (double)D.uhp[IX(k, lW)];
      vN = (double)D.vhp[IX(k, lE)] + (double)D.vhp[IX(k, c)];
      vS = (double)D.vhp[IX(k, l2)] + (double)D.vhp[IX(k, lS)];
      wU = (double)D.wp[IX(k, lE)] + (double)D.wp[IX(k, c)];
      wD = (double)D.wp[IX(k + 1, lE)] + (double)D.wp[IX(k + 1, c)];
    } else {
      cor = 0.25 * f * ((double)D.uhp[IX(k, lN)] + (double)D.uhp[IX(k, c)] + (double)D.uhp[IX(k, l2)] + (double)D.uhp[IX(k, lW)]);
      uE = (double)D.uhp[IX(k, lN)] + (double)D.uhp[IX(k, c)];
      uW = (double)D.uhp[IX(k, l2)] + (double)D.uhp[IX(k, lW)];
      vN = (double)D.vhp[IX(k, lN)] + (double)D.vhp[IX(k, c)];
      vS = (double)D.vhp[IX(k, c)] + (double)D.vhp[IX(k, lS)];
      wU = (double)D.wp[IX(k, lN)] + (double)D.wp[IX(k, c)];
      wD = (double)D.wp[IX(k + 1, lN)] + (double)D.wp[IX(k + 1, c)];
    }
    if (k == k1q) wU = 0.0;
    if (k == kmq) wD = 0.0;
    const double qc = q[IX(k, c)];
    double adv = 0.0;
    const int it = D.P.itrmom;
    if (it == 1 || it == 3) {
      double scE = (double)q[IX(k, lE)] + qc, scW = (double)q[IX(k, lW)] + qc;
      double scN = (double)q[IX(k, lN)] + qc, scS = (double)q[IX(k, lS)] + qc;
      adv = (uE * scE - uW * scW) / D.P.fourdx + (vN * scN - vS * scS) / D.P.fourdy;
    } else {
      const double pc = qpp[IX(k, c)];
      adv = ((uE + fabs(uE)) * pc + (uE - fabs(uE)) * (double)qpp[IX(k, lE)] -
             (uW + fabs(uW)) * (double)qpp[IX(k, lW)] - (uW - fabs(uW)) * pc) / D.P.fourdx +
            ((vN + fabs(vN)) * pc + (vN - fabs(vN)) * (double)qpp[IX(k, lN)] -
             (vS + fabs(vS)) * (double)qpp[IX(k, lS)] - (vS - fabs(vS)) * pc) / D.P.fourdy;
    }
    if (it == 1 || it == 4) {
      // vertical weights th/th1 (680-689): th(k1q)=0, th(kmq+1)=0.5
      const double hkm = hq[IX(k - 1, c)], hk = hq[IX(k, c)], hkp = hq[IX(k + 1, c)];
      double th = (k == k1q) ? 0.0 : hkm / (hkm + hk);
      double th1 = (k == k1q) ? 0.0 : 1.0 - th;
      double thp = (k == kmq) ? 0.5 : hk / (hk + hkp);
      double th1p = (k == kmq) ? 0.5 : 1.0 - thp;
      adv += (wU * (th * qc + th1 * (double)q[IX(k - 1, c)]) - wD * (thp * (double)q[IX(k + 1, c)] + th1p * qc)) / 2.;
    } else if (it == 2) {
      const double pc = qpp[IX(k, c)];
      adv += ((wU + fabs(wU)) * pc + (wU - fabs(wU)) * (double)qpp[IX(k - 1, c)]) / 4. -
             ((wD + fabs(wD)) * (double)qpp[IX(k + 1, c)] + (wD - fabs(wD)) * pc) / 4.;
    } else {  // it == 3
      const double hkm = hq[IX(k - 1, c)], hk = hq[IX(k, c)], hkp = hq[IX(k + 1, c)];
      double scU = (qc * hk + (double)q[IX(k - 1, c)] * hkm) / (hk + hkm);
      double scD = (qc * hk + (double)q[IX(k + 1, c)] * hkp) / (hk + hkp);
      adv += (wU * scU - wD * scD) / 2.;
    }
    double hd = 0.0;
    if (D.P.ihd >= 1) {
      const double pc = qpp[IX(k, c)];
      const double axc = hcoef<R, X, true>(D, k, c), axw = hcoef<R, X, true>(D, k, lW);
      const double ayc = hcoef<R, X, false>(D, k, c), ays = hcoef<R, X, false>(D, k, lS);
      if (D.P.ihd == 1) {
        hd = (ayc * ((double)qpp[IX(k, lN)] - pc) - ays * (pc - (double)qpp[IX(k, lS)])) / D.P.dydy +
             (axc * ((double)qpp[IX(k, lE)] - pc) - axw * (pc - (double)qpp[IX(k, lW)])) / D.P.dxdx;
      } else if (X) {
        hd = 2. * (axc * ((double)D.upp[IX(k, lE)] - pc) - axw * (pc - (double)D.upp[IX(k, lW)])) / D.P.dxdx +
             (ayc * ((double)D.upp[IX(k, lN)] - pc) - ays * (pc - (double)D.upp[IX(k, lS)])) / D.P.dydy +
             (ayc * ((double)D.vpp[IX(k, lE)] - (double)D.vpp[IX(k, c)]) -
              ays * ((double)D.vpp[IX(k, D.lE[lS])] - (double)D.vpp[IX(k, lS)])) / D.P.dxdy;
      } else {
        hd = 2. * (ayc * ((double)D.vpp[IX(k, lN)] - pc) - ays * (pc - (double)D.vpp[IX(k, lS)])) / D.P.dydy +
             (axc * ((double)D.vpp[IX(k, lE)] - pc) - axw * (pc - (double)D.vpp[IX(k, lW)])) / D.P.dxdx +
             (axc * ((double)D.upp[IX(k, lN)] - (double)D.upp[IX(k, c)]) -
              axw * ((double)D.upp[IX(k, D.lN[lW])] - (double)D.upp[IX(k, lW)])) / D.P.dxdy;
      }
    }
    if (X) return (double)D.uhpp[IX(k, c)] - twodt1 * (adv * D.P.iadv - cor - hd);
    return (double)D.vhpp[IX(k, c)] - twodt1 * (adv * D.P.iadv + cor - hd);
  }
}

template <typename R, bool X>
__device__ __forceinline__ void exmom_cell(const Dev<R> &D, int c, int k, double twodt1) {
  const int S = D.lm;
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == S) return;  // no face: ex stays zero
  const int kmq = min(D.kmz[ln], D.kmz[c]);
  if (k > kmq) return;
  const int k1q = X ? D.k1u[c] : D.k1v[c];
  R *ex = X ? D.ex : D.ey;
  ex[IX(k, c)] = (k < k1q) ? R(0) : (R)exmom_value<R, X>(D, c, k, k1q, kmq, twodt1);
}

template <typename R, bool X, bool CH>
__device__ __forceinline__ void exmom_tile(const Dev<R> &D, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) exmom_cell<R, X>(D, t.c, k, twodt1);
}

// Chunked instantiation; the single-layer one below asks the compiler for 8
// waves per SIMD (it fits 64 VGPRs without spilling, against 70-72 and 7
// waves when left free).
template <typename R, bool X, bool CH>
__global__ void __launch_bounds__(kBlock) k_exmom(Dev<R> D, double twodt1, int nk, int tile) {
  exmom_tile<R, X, CH>(D, twodt1, nk, tile);
}

template <typename R, bool X>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_exmom1(Dev<R> D, double twodt1, int nk, int tile) {
  exmom_tile<R, X, false>(D, twodt1, nk, tile);
}

// Buffer-resource loads of the per-cell kernels.  In a per-cell kernel the
// layer k is the same for the whole workgroup, so each 3-D field is read
// through a buffer resource of the whole field (scalar registers), a scalar
// plane offset k*LP*sizeof(R) shared by every field, and a 32-bit byte offset
// per neighbour column shared by every field: one VGPR per neighbour instead
// of a 64-bit address (and its carry arithmetic) per field and neighbour.
// Fields must be below 4 GiB (checked on the host before these kernels run).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const T *a, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)a, (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ double bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, T v, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((__vector_size__(8))) unsigned, v), r,
                                          voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// Memory access of the column kernels at (layer k, column c), where each lane
// has its own k: through plain pointers (PtrIO, any field size), or through
// buffer resources with one 32-bit byte offset per access (BufIO, fields below
// 4 GiB), so that an address takes one VGPR instead of a 64-bit pair, and the
// compiler cannot turn each field's address into a 64-bit pointer stepped down
// the column.
struct PtrIO {
  int LP;
  template <typename T>
  __device__ __forceinline__ T ld(const T *a, int k, int c) const {
    return a[(size_t)k * LP + c];
  }
  template <typename T>
  __device__ __forceinline__ void st(T *a, int k, int c, T v) const {
    a[(size_t)k * LP + c] = v;
  }
};
struct BufIO {
  unsigned LP, KP;
  template <typename T>
  __device__ __forceinline__ T ld(const T *a, int k, int c) const {
    const auto r = rsrc(a, KP * LP * (unsigned)sizeof(T));
    const unsigned off = ((unsigned)k * LP + (unsigned)c) * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    else if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
  template <typename T>
  __device__ __forceinline__ void st(T *a, int k, int c, T v) const {
    const auto r = rsrc(a, KP * LP * (unsigned)sizeof(T));
    const unsigned off = ((unsigned)k * LP + (unsigned)c) * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
    else if constexpr (sizeof(T) == 8)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((__vector_size__(8))) unsigned, v), r,
                                            off, 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((__vector_size__(16))) unsigned, v), r,
                                             off, 0, 0);
  }
};

template <typename IO, typename R>
__device__ __forceinline__ IO make_io(const Dev<R> &D) {
  if constexpr (std::is_same<IO, BufIO>::value)
    return BufIO{(unsigned)D.LP, (unsigned)D.KP};
  else
    return PtrIO{D.LP};
}

// exmom for the centred scheme (itrmom 1) with constant-coefficient or no
// horizontal diffusion (ihd 0, 1): the same value as exmom_value, evaluated
// in the same order (so the two kernels agree bit for bit), for a chunk of
// layers per lane.  The column's neighbour indices, face bottoms (kmu / kmv)
// and top indices are loaded once per chunk; the vertical neighbours k-1,
// k+1 of q, its face thickness and wp at k+1 slide down the chunk in
// registers.  Relies on symmetric neighbours: the east neighbour of the west
// neighbour is the column itself (lE[lW[c]] = c), and likewise north of south.
template <typename R, bool X>
__global__ void __launch_bounds__(kBlock) k_exmom_w(Dev<R> D, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const int ln = X ? lE : lN;
  if (ln == S) return;  // no face: ex stays zero
  const int *kmf = X ? D.kmu : D.kmv;
  const int *k1f = X ? D.k1u : D.k1v;
  const int kmq = kmf[c];
  if (t.k0 > kmq) return;
  const int k1q = k1f[c];
  const int l2 = X ? D.lS[lE] : D.lW[lN];  // SE corner (x) / NW corner (y)
  // diffusion coefficient of the west face (x) and south face (y) of the cell:
  // non-zero on the neighbour face's own wet range
  const int k1W = k1f[lW], kmW = kmf[lW], k1S = k1f[lS], kmS = kmf[lS];
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz, oE = (unsigned)lE * sz, oW = (unsigned)lW * sz, oN = (unsigned)lN * sz,
                 oS = (unsigned)lS * sz, o2 = (unsigned)l2 * sz, on = (unsigned)ln * sz;
  const auto q = rsrc(X ? D.up : D.vp, nb), qpp = rsrc(X ? D.upp : D.vpp, nb);
  const auto hq = rsrc(X ? D.hup : D.hvp, nb), hqq = rsrc(X ? D.hupp : D.hvpp, nb);
  const auto qhpp = rsrc(X ? D.uhpp : D.vhpp, nb), ex = rsrc(X ? D.ex : D.ey, nb);
  const auto uhp = rsrc(D.uhp, nb), vhp = rsrc(D.vhp, nb), wp = rsrc(D.wp, nb), hp = rsrc(D.hp, nb);
  const double f = D.P.f, iadv = D.P.iadv, Ax0 = D.P.Ax0, Ay0 = D.P.Ay0;
  const int ihd = D.P.ihd;
  unsigned pk = (unsigned)t.k0 * pb;  // byte offset of plane k
  double qm = bld<R>(q, oc, pk - pb), q0 = bld<R>(q, oc, pk);
  double hm = bld<R>(hq, oc, pk - pb), h0 = bld<R>(hq, oc, pk);
  double w0c = bld<R>(wp, oc, pk), w0n = bld<R>(wp, on, pk);
  for (int k = t.k0; k < t.k1; ++k, pk += pb) {
    if (k > kmq) break;
    const unsigned pk1 = pk + pb;
    const double qp = bld<R>(q, oc, pk1), hpk = bld<R>(hq, oc, pk1);
    const double w1c = bld<R>(wp, oc, pk1), w1n = bld<R>(wp, on, pk1);
    double val = 0.0;
    if (k >= k1q) {
      if (bld<R>(hp, oc, pk) <= 0.0 || bld<R>(hp, on, pk) <= 0.0) {
        val = bld<R>(qhpp, oc, pk);
      } else {
        double cor, uE, uW, vN, vS;
        if (X) {
          const double vhE = bld<R>(vhp, oE, pk), vhc = bld<R>(vhp, oc, pk), vh2 = bld<R>(vhp, o2, pk),
                       vhS = bld<R>(vhp, oS, pk);
          const double uhc = bld<R>(uhp, oc, pk);
          cor = 0.25 * f * (vhE + vhc + vh2 + vhS);
          uE = bld<R>(uhp, oE, pk) + uhc;
          uW = uhc + bld<R>(uhp, oW, pk);
          vN = vhE + vhc;
          vS = vh2 + vhS;
        } else {
          const double uhN = bld<R>(uhp, oN, pk), uhc = bld<R>(uhp, oc, pk), uh2 = bld<R>(uhp, o2, pk),
                       uhW = bld<R>(uhp, oW, pk);
          const double vhc = bld<R>(vhp, oc, pk);
          cor = 0.25 * f * (uhN + uhc + uh2 + uhW);
          uE = uhN + uhc;
          uW = uh2 + uhW;
          vN = bld<R>(vhp, oN, pk) + vhc;
          vS = vhc + bld<R>(vhp, oS, pk);
        }
        double wU = w0n + w0c, wD = w1n + w1c;
        if (k == k1q) wU = 0.0;
        if (k == kmq) wD = 0.0;
        const double qc = q0;
        const double scE = bld<R>(q, oE, pk) + qc, scW = bld<R>(q, oW, pk) + qc;
        const double scN = bld<R>(q, oN, pk) + qc, scS = bld<R>(q, oS, pk) + qc;
        double adv = (uE * scE - uW * scW) / D.P.fourdx + (vN * scN - vS * scS) / D.P.fourdy;
        const double hkm = hm, hk = h0, hkp = hpk;
        const double th = (k == k1q) ? 0.0 : hkm / (hkm + hk);
        const double th1 = (k == k1q) ? 0.0 : 1.0 - th;
        const double thp = (k == kmq) ? 0.5 : hk / (hk + hkp);
        const double th1p = (k == kmq) ? 0.5 : 1.0 - thp;
        adv += (wU * (th * qc + th1 * qm) - wD * (thp * qp + th1p * qc)) / 2.;
        double hd = 0.0;
        if (ihd >= 1) {
          const double pc = bld<R>(qpp, oc, pk), hqc = bld<R>(hqq, oc, pk);
          const double axc = Ax0 * dmin(bld<R>(hqq, oE, pk), hqc);
          const double axw = (k < k1W || k > kmW) ? 0.0 : Ax0 * dmin(hqc, bld<R>(hqq, oW, pk));
          const double ayc = Ay0 * dmin(bld<R>(hqq, oN, pk), hqc);
          const double ays = (k < k1S || k > kmS) ? 0.0 : Ay0 * dmin(hqc, bld<R>(hqq, oS, pk));
          hd = (ayc * (bld<R>(qpp, oN, pk) - pc) - ays * (pc - bld<R>(qpp, oS, pk))) / D.P.dydy +
               (axc * (bld<R>(qpp, oE, pk) - pc) - axw * (pc - bld<R>(qpp, oW, pk))) / D.P.dxdx;
        }
        if (X)
          val = bld<R>(qhpp, oc, pk) - twodt1 * (adv * iadv - cor - hd);
        else
          val = bld<R>(qhpp, oc, pk) - twodt1 * (adv * iadv + cor - hd);
      }
    }
    bst<R>(ex, (R)val, oc, pk);
    qm = q0;
    q0 = qp;
    hm = h0;
    h0 = hpk;
    w0c = w1c;
    w0n = w1n;
  }
}

// bclnc_km: baroclinic term of the bottom layer on a horizontal plane,
// si3d_procedures.f90:1633-1746.
template <typename R, bool X>
__device__ double bclnc_km(const Dev<R> &D, int c, int ln, int kb, double plain) {
  const double eps = FLT_EPSILON;  // EPSILON of a default REAL
  const double hq = X ? (double)D.hup[IX(kb, c)] : (double)D.hvp[IX(kb, c)];
  const double hplus = hq + eps, hminus = hq - eps;
  bool ce = false, cw = false;
  double rn = D.rhop[IX(kb, ln)], rw = D.rhop[IX(kb, c)];
  const double hpn = D.hp[IX(kb, ln)], hpc = D.hp[IX(kb, c)];
  if (hpn > hplus || hpn < hminus) {
    ce = true;
    double h1 = D.hp[IX(kb - 1, ln)];
    double dz1 = 0.5 * (h1 + hpn), dz2 = 0.5 * (h1 + hq);
    double r1 = D.rhop[IX(kb - 1, ln)];
    rn = r1 + (dz2 / dz1) * ((double)D.rhop[IX(kb, ln)] - r1);
  }
  if (hpc > hplus || hpc < hminus) {
    cw = true;
    double h1 = D.hp[IX(kb - 1, c)];
    double dz1 = 0.5 * (h1 + hpc), dz2 = 0.5 * (h1 + hq);
    double r1 = D.rhop[IX(kb - 1, c)];
    rw = r1 + (dz2 / dz1) * ((double)D.rhop[IX(kb, c)] - r1);
  }
  if (ce || cw) return (X ? D.P.gdtdx : D.P.gdtdy) * hq * (rn - rw);
  return plain;
}

// matmom: per-face tridiagonal momentum system and its two-RHS elimination,
// si3d_procedures.f90:1176-1630 with trid 3210-3289.  Pass A: column-mean
// density; pass B: assemble each row and eliminate downward, storing the
// factors (c1 -> cfx/cfy, d1 -> agx/agy, e1 -> arx/ary).  The back
// substitution (trid's upward pass) is left to vel, the only consumer of the
// per-layer solution; the column sums eagx/earx that matcon needs come out of
// the downward pass itself: with x = U^-1 d (U unit upper bidiagonal with
// superdiagonal c1), sum_k x_k = w . d where U^T w = 1, i.e. w_k1 = 1,
// w_k = 1 - c1_(k-1) w_(k-1), a downward recurrence.  The weights and terms
// use the factors as stored, so the sums are those of the solution vel
// back-substitutes.
template <typename R, bool X>
__global__ void __launch_bounds__(kBlock) k_matmom(Dev<R> D, double tz) {
  const int c = col();
  if (c >= D.lm) return;
  const int S = D.lm;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const int ln = X ? lE : lN;
  if (ln == S) return;
  const int kmq = min(D.kmz[ln], D.kmz[c]);
  const int k1q = X ? D.k1u[c] : D.k1v[c];
  const int nwl = kmq - k1q + 1;
  if (nwl < 1) return;
  const double twodt1 = D.P.twodt * tz, theta = D.P.theta;
  const R *hq = X ? D.hup : D.hvp;
  const R *qpp = X ? D.upp : D.vpp;
  R *AG = X ? D.agx : D.agy;
  R *AR = X ? D.arx : D.ary;
  const double gq = X ? D.P.gdtdx : D.P.gdtdy;

  double rsum = 0.0;
  for (int k = k1q; k <= kmq; ++k) rsum += 1000. + (double)D.rhop[IX(k, c)];
  const double rhoq = rsum / nwl, irhoq = 1.0 / rhoq;
  const double ws0 = rhoq * gq * (D.spp[ln] - D.spp[c]);

  // wind stress at the top face (1348-1360 / 1541-1553)
  double usurf, vsurf, wind_g;
  if (X) {
    const int lSE = D.lS[lE];
    usurf = D.up[IX(k1q, c)];
    vsurf = ((double)D.vp[IX(D.k1v[c], c)] + (double)D.vp[IX(D.k1v[lS], lS)] + (double)D.vp[IX(D.k1v[lE], lE)] +
             (double)D.vp[IX(D.k1v[lSE], lSE)]) / 4.;
  } else {
    const int lNW = D.lW[lN];
    vsurf = D.vp[IX(k1q, c)];
    usurf = ((double)D.up[IX(D.k1u[c], c)] + (double)D.up[IX(D.k1u[lW], lW)] + (double)D.up[IX(D.k1u[lN], lN)] +
             (double)D.up[IX(D.k1u[lNW], lNW)]) / 4.;
  }
  const Wind wd = wind(D, c);
  const double cwq = wd.cdw * RHOAIR * sqrt((wd.vair - vsurf) * (wd.vair - vsurf) + (wd.uair - usurf) * (wd.uair - usurf));
  wind_g = cwq / rhoq * (X ? wd.uair : wd.vair) * twodt1;
  const double wind_a = cwq / rhoq / (double)hq[IX(k1q, c)] * twodt1;
  // quadratic bottom stress (1364-1369 / 1557-1562)
  double taub;
  if (X) {
    double vb = 0.25 * ((double)D.vhpp[IX(kmq, lE)] + (double)D.vhpp[IX(kmq, c)] + (double)D.vhpp[IX(kmq, lS)] +
                        (double)D.vhpp[IX(kmq, D.lS[lE])]);
    double ub = D.uhpp[IX(kmq, c)];
    double hb = D.hup[IX(kmq, c)];
    taub = D.P.cd * sqrt(ub * ub + vb * vb) / (hb * hb);
  } else {
    double ub = 0.25 * ((double)D.uhpp[IX(kmq, lN)] + (double)D.uhpp[IX(kmq, D.lW[lN])] + (double)D.uhpp[IX(kmq, lW)] +
                        (double)D.uhpp[IX(kmq, c)]);
    double vb = D.vhpp[IX(kmq, c)];
    double hb = D.hvp[IX(kmq, c)];
    taub = D.P.cd * sqrt(vb * vb + ub * ub) / (hb * hb);
  }
  const double bot_a = taub * twodt1;

  // The level arithmetic runs in the storage precision C = R (FP64, or FP32
  // in mixed mode); the per-column terms above and the column sums below are
  // FP64.
  using C = R;
  const C twodt1c = (C)twodt1, thetac = (C)theta, tzc = (C)tz, gqc = (C)gq, ws0c = (C)ws0, irhoqc = (C)irhoq;
  const C twodtc = (C)D.P.twodt;
  C bcl = 0, hdr_prev = 0, c1p = 0, d1p = 0, e1p = 0;
  R *CF = X ? D.cfx : D.cfy;
  double w = 1.0, sg = 0.0, sr = 0.0, c1s = 0.0;  // column-sum weight, sums, c1 as stored
  // thicknesses (and 1/h), face viscosity, qpp and the viscous flux of levels
  // k-1, k (k+1) carried down the sweep.  The sweep is bound by its
  // arithmetic: divisions by a quantity used more than once are
  // multiplications by its reciprocal (1/h, 1/rho, the pivot), which moves
  // results by rounding only; the viscous flux through the top of layer k is
  // the one through the bottom of layer k-1.
  C hkm = hq[IX(k1q - 1, c)], hk = hq[IX(k1q, c)], Avk = 0, rhk = C(1) / hk, dkp_prev = 0;
  C qkm = 0, qk = (nwl > 1) ? (C)qpp[IX(k1q, c)] : C(0);
  for (int k = k1q; k <= kmq; ++k) {
    const C hkp = hq[IX(k + 1, c)];
    const C Avkp = (k < kmq) ? C(0.5) * ((C)D.Av[IX(k + 1, ln)] + (C)D.Av[IX(k + 1, c)]) : C(0);
    const C qkp = (nwl > 1 && k < kmq) ? (C)qpp[IX(k + 1, c)] : C(0);
    if (D.P.ibc != 0) {
      C hdr = gqc * hk * ((C)D.rhop[IX(k, ln)] - (C)D.rhop[IX(k, c)]);
      if (k == kmq && kmq > k1q) hdr = (C)bclnc_km<R, X>(D, c, ln, kmq, (double)hdr);
      bcl = (k == k1q) ? hdr : bcl + (hdr_prev + hdr);
      hdr_prev = hdr;
    }
    C vd = 0;
    if (nwl > 1) {
      const C dk = dkp_prev;
      const C dkp = (k < kmq) ? Avkp * (qk - qkp) / (hk + hkp) : C(0);
      vd = twodtc * (dk - dkp) * C(2) * (C(1) - thetac);
      dkp_prev = dkp;
    }
    const C hh = hk * irhoqc;
    C gg = (C)(X ? D.ex : D.ey)[IX(k, c)] - hh * (bcl + ws0c) * tzc + vd * tzc;
    C a1 = 0, a2 = 1, a3 = 0;
    if (nwl > 1) {
      a3 = (k < kmq) ? -twodt1c * Avkp / (hkp * (hk + hkp)) * C(2) * thetac : C(0);
      a1 = (k > k1q) ? -twodt1c * Avk / (hkm * (hkm + hk)) * C(2) * thetac : C(0);
      a2 = C(1) - (hkm * rhk) * a1 - (hkp * rhk) * a3;
    }
    if (k == k1q) {
      gg += (C)wind_g;
      a2 += (C)wind_a;
    }
    if (k == kmq) a2 += (C)bot_a;
    C c1, d1, e1;
    if (k == k1q) {
      const C rd = C(1) / a2;
      c1 = a3 * rd;
      d1 = gg * rd;
      e1 = hh * rd;
    } else {
      const C rd = C(1) / (a2 - a1 * c1p);
      c1 = a3 * rd;
      d1 = (gg - a1 * d1p) * rd;
      e1 = (hh - a1 * e1p) * rd;
    }
    const R c1r = (R)c1, d1r = (R)d1, e1r = (R)e1;
    CF[IX(k, c)] = c1r;
    AG[IX(k, c)] = d1r;
    AR[IX(k, c)] = e1r;
    if (k > k1q) w = 1.0 - c1s * w;
    sg += w * (double)d1r;
    sr += w * (double)e1r;
    c1s = (double)c1r;
    c1p = c1;
    d1p = d1;
    e1p = e1;
    hkm = hk;
    hk = hkp;
    rhk = C(1) / hkp;
    Avk = Avkp;
    qkm = qk;
    qk = qkp;
  }
  if (X) {
    D.eagx[c] = sg;
    D.earx[c] = sr;
  } else {
    D.eagy[c] = sg;
    D.eary[c] = sr;
  }
}

// matcon: pentadiagonal free-surface coefficients, si3d_procedures.f90:1749-1867.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_matcon(Dev<R> D, double tz) {
  const int c = col();
  if (c >= D.lm) return;
  const int S = D.lm;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const double cx = D.P.gdt2dx2 * tz * tz * 1000., cy = D.P.gdt2dy2 * tz * tz * 1000.;
  const double sxc = (lE != S) ? cx * D.earx[c] : 0.0;
  const double syc = (lN != S) ? cy * D.eary[c] : 0.0;
  const double sxw = (lW != S) ? cx * D.earx[lW] : 0.0;  // E of lW is c: wet
  const double sys = (lS != S) ? cy * D.eary[lS] : 0.0;
  double sum = 0.0;
  const int kms = D.kmz[c];
  for (int k = D.k1; k <= kms; ++k)
    sum += ((double)D.uhpp[IX(k, c)] - (double)D.uhpp[IX(k, lW)]) / D.P.dx +
           ((double)D.vhpp[IX(k, c)] - (double)D.vhpp[IX(k, lS)]) / D.P.dy;
  const double dd = D.P.dt * tz * sum;
  const double eW = (lW != S) ? D.eagx[lW] : 0.0, eS = (lS != S) ? D.eagy[lS] : 0.0;
  D.sx[c] = sxc;
  D.sy[c] = syc;
  D.dd[c] = dd;
  D.qq[c] = D.spp[c] - D.P.dtdx * tz * (D.eagx[c] - eW) - D.P.dtdy * tz * (D.eagy[c] - eS) - dd;
  D.rr[c] = 1.0 + sxc + sxw + syc + sys;
}

// slayer_h: new surface-layer thicknesses after the zeta solve (drying only),
// si3d_procedures.f90:2070-2175, including the re-test of the same u-layer (2136).
// One lane per (column, layer k = 1..km1).  Below kmz+1 (the sediment cell) h,
// hp, hu, hup, hv, hvp are all zero, so only k <= kmz+1 is visited.
template <typename R>
__device__ __forceinline__ void slayer_h_cell(const Dev<R> &D, int c, int k) {
  if (k > D.kmz[c] + 1) return;
  const int S = D.lm, km1 = km1c(D, c);
  const int k1z = D.k1z[c], k1u = D.k1u[c], k1v = D.k1v[c];
  const double sc = D.s[c];
  // zeta points: drying of the top cell moves the computation one layer down
  {
    const double haux0 = dmin(zlv(D, c)[k1z + 1], D.hhs[c]) + sc;
    if (k < k1z || k >= k1z + 2) {
      D.h[IX(k, c)] = D.hp[IX(k, c)];
    } else if (k == k1z) {
      D.h[IX(k, c)] = (haux0 <= 0.0) ? R(0) : (R)haux0;
    } else if (haux0 <= 0.0) {  // k == k1z + 1 after the top cell dried
      const double haux = dmin(zlv(D, c)[k + 1], D.hhs[c]) + sc;
      D.h[IX(k, c)] = (haux <= 0.0) ? R(0) : (R)haux;
    }
  }
  const int lE = D.lE[c], lN = D.lN[c];
  // u points: the reference re-tests the same layer after drying (2136)
  if (k != k1u) {
    D.hu[IX(k, c)] = D.hup[IX(k, c)];
  } else if (lE != S) {
    const double haux = dmin(zlv(D, c)[k + 1], D.hhu[c]) + dmax(sc, D.s[lE]);
    D.hu[IX(k, c)] = (haux <= 0.0) ? R(0) : (R)haux;
  } else {
    D.hu[IX(k, c)] = D.hup[IX(k, c)];
  }
  // v points
  if (lN == S) {
    if (k < k1v || k >= k1v + 2 || k <= min(k1v + 1, km1)) D.hv[IX(k, c)] = D.hvp[IX(k, c)];
  } else {
    const double sm = dmax(sc, D.s[lN]);
    const double haux0 = dmin(zlv(D, c)[k1v + 1], D.hhv[c]) + sm;
    if (k < k1v || k >= k1v + 2) {
      D.hv[IX(k, c)] = D.hvp[IX(k, c)];
    } else if (k == k1v) {
      D.hv[IX(k, c)] = (haux0 <= 0.0) ? R(0) : (R)haux0;
    } else if (haux0 <= 0.0) {
      const double haux = dmin(zlv(D, c)[k + 1], D.hhv[c]) + sm;
      D.hv[IX(k, c)] = (haux <= 0.0) ? R(0) : (R)haux;
    }
  }
}

template <typename R, bool CH>
__global__ void __launch_bounds__(kBlock) k_slayer_h(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 1, nk, tile, t)) return;
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) slayer_h_cell<R>(D, t.c, k);
}

// Windowed form of k_slayer_h (same values, same order): the column's top
// indices, levels and surface elevations are read once per chunk of layers
// and the 3-D fields through buffer resources with a scalar plane offset.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_slayer_h_w(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 1, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const int kend = min(t.k1, D.kmz[c] + 2);
  if (t.k0 >= kend) return;
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz;
  const int km1 = km1c(D, c), k1z = D.k1z[c], k1u = D.k1u[c], k1v = D.k1v[c];
  const int lE = D.lE[c], lN = D.lN[c];
  const double sc = D.s[c], hhs = D.hhs[c], hhu = D.hhu[c], hhv = D.hhv[c];
  const double *zl = zlv(D, c);
  const double smE = (lE != S) ? dmax(sc, D.s[lE]) : 0.0;
  const double smN = (lN != S) ? dmax(sc, D.s[lN]) : 0.0;
  const double haux0z = dmin(zl[k1z + 1], hhs) + sc;
  const double haux0v = (lN != S) ? dmin(zl[k1v + 1], hhv) + smN : 0.0;
  const auto h = rsrc(D.h, nb), hu = rsrc(D.hu, nb), hv = rsrc(D.hv, nb);
  const auto hp = rsrc(D.hp, nb), hup = rsrc(D.hup, nb), hvp = rsrc(D.hvp, nb);
  for (int k = t.k0; k < kend; ++k) {
    const unsigned pk = (unsigned)k * pb;
    if (k < k1z || k >= k1z + 2) {
      bst<R>(h, (R)bld<R>(hp, oc, pk), oc, pk);
    } else if (k == k1z) {
      bst<R>(h, (haux0z <= 0.0) ? R(0) : (R)haux0z, oc, pk);
    } else if (haux0z <= 0.0) {
      const double haux = dmin(zl[k + 1], hhs) + sc;
      bst<R>(h, (haux <= 0.0) ? R(0) : (R)haux, oc, pk);
    }
    if (k != k1u || lE == S) {
      bst<R>(hu, (R)bld<R>(hup, oc, pk), oc, pk);
    } else {
      const double haux = dmin(zl[k + 1], hhu) + smE;
      bst<R>(hu, (haux <= 0.0) ? R(0) : (R)haux, oc, pk);
    }
    if (lN == S) {
      if (k < k1v || k >= k1v + 2 || k <= min(k1v + 1, km1)) bst<R>(hv, (R)bld<R>(hvp, oc, pk), oc, pk);
    } else if (k < k1v || k >= k1v + 2) {
      bst<R>(hv, (R)bld<R>(hvp, oc, pk), oc, pk);
    } else if (k == k1v) {
      bst<R>(hv, (haux0v <= 0.0) ? R(0) : (R)haux0v, oc, pk);
    } else if (haux0v <= 0.0) {
      const double haux = dmin(zl[k + 1], hhv) + smN;
      bst<R>(hv, (haux <= 0.0) ? R(0) : (R)haux, oc, pk);
    }
  }
}

// The sentinel column: h := hp, hu := hup, hv := hvp (2087-2089).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_slayer_h_sentinel(Dev<R> D) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > D.km1 + 1) return;
  const int c = D.lm;
  D.h[IX(k, c)] = D.hp[IX(k, c)];
  D.hu[IX(k, c)] = D.hup[IX(k, c)];
  D.hv[IX(k, c)] = D.hvp[IX(k, c)];
}

// vel: fluxes from the new zeta plus the top-cell drying merge,
// si3d_procedures.f90:2430-2545, with matmom's back substitution (trid
// 3282-3288) in the same upward sweep: ag, ar of layer k from the stored
// factors, uh = ag - g dt/dx 1000 (s_E - s) ar.  One lane per column and
// face.  The drying merge moves the top flux one layer down: when the top
// face cell k0 is dry its flux joins layer k0+1's (or, with a single wet
// layer, the layer below the face) and k0's becomes zero.
template <typename R, bool X>
__device__ __forceinline__ void vel_column(const Dev<R> &D, int c, double tz) {
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == D.lm) return;
  const int kmq = X ? D.kmu[c] : D.kmv[c];
  const int k0 = X ? D.k1u[c] : D.k1v[c];
  if (k0 > kmq) return;
  const R *cf = X ? D.cfx : D.cfy;
  const R *ag = X ? D.agx : D.agy;
  const R *ar = X ? D.arx : D.ary;
  R *qh = X ? D.uh : D.vh;
  const R *hq = X ? D.hu : D.hv;
  const double cq = (X ? D.P.gdtdx : D.P.gdtdy) * tz * 1000. * (D.s[ln] - D.s[c]);
  const bool dry_top = (double)hq[IX(k0, c)] <= 0.0;
  double xg = 0.0, xr = 0.0;
  R v1 = R(0);  // flux of layer k0+1, held until k0's is known (dry top)
  // Layers in groups of four, the group's factors loaded together before its
  // substitutions and stores (as in k_continuity); layers of a group above k0
  // read a valid layer (clamped at 0) and are discarded.
  for (int kg = kmq; kg >= k0; kg -= 4) {
    R fc[4], fg[4], fr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = max(kg - j, 0);
      fc[j] = cf[IX(kk, c)];
      fg[j] = ag[IX(kk, c)];
      fr[j] = ar[IX(kk, c)];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = kg - j;
      if (k < k0) break;
      if (k < kmq) {
        const double c1 = fc[j];
        xg = (double)fg[j] - c1 * xg;
        xr = (double)fr[j] - c1 * xr;
      } else {
        xg = fg[j];
        xr = fr[j];
      }
      const R v = (R)((double)(R)xg - cq * (double)(R)xr);
      if (!dry_top || k > k0 + 1) {
        qh[IX(k, c)] = v;
      } else if (k == k0 + 1) {
        v1 = v;
      } else {  // the dry top cell: its flux lands one layer below (2490-2497, 2535-2541)
        qh[IX(k0 + 1, c)] = (k0 + 1 > kmq) ? v + qh[IX(k0 + 1, c)] : v + v1;
        qh[IX(k0, c)] = R(0);
      }
    }
  }
}

// Both faces of a column; the top indices move in k_continuity<R,1>.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_vel(Dev<R> D, double tz) {
  const int c = col();
  if (c >= D.lm) return;
  vel_column<R, true>(D, c, tz);
  vel_column<R, false>(D, c, tz);
}

// The per-layer solution of one face's momentum system (what vel
// back-substitutes), written to ag_out / ar_out on the face's wet layers and
// zero elsewhere: the reference's agx/arx (agy/ary) for callers that read
// them after matmom.
template <typename R, bool X>
__global__ void __launch_bounds__(kBlock) k_momentum_solution(Dev<R> D, R *ag_out, R *ar_out) {
  const int c = col();
  if (c >= D.lm) return;
  for (int k = 0; k < D.KP; ++k) ag_out[IX(k, c)] = ar_out[IX(k, c)] = R(0);
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == D.lm) return;
  const int kmq = X ? D.kmu[c] : D.kmv[c];
  const int k0 = X ? D.k1u[c] : D.k1v[c];
  const R *cf = X ? D.cfx : D.cfy;
  const R *ag = X ? D.agx : D.agy;
  const R *ar = X ? D.arx : D.ary;
  double xg = 0.0, xr = 0.0;
  for (int k = kmq; k >= k0; --k) {
    if (k == kmq) {
      xg = ag[IX(k, c)];
      xr = ar[IX(k, c)];
    } else {
      const double c1 = cf[IX(k, c)];
      xg = (double)ag[IX(k, c)] - c1 * xg;
      xr = (double)ar[IX(k, c)] - c1 * xr;
    }
    ag_out[IX(k, c)] = (R)xg;
    ar_out[IX(k, c)] = (R)xr;
  }
}

// continuity: vertical velocity by bottom-up integration of the horizontal
// flux divergence, si3d_procedures.f90:2635-2745 (case 1 from uh and uhpp, case
// 2 from uhp), accumulated in FP64.  Case 1 also updates k1z for drying (vel,
// 2547-2561).
template <typename R, int IST>
__global__ void __launch_bounds__(kBlock) k_continuity(Dev<R> D) {
  const int c = col();
  if (c >= D.lm) return;
  const int kms = D.kmz[c], k1s = D.k1z[c], lW = D.lW[c], lS = D.lS[c];
  if (k1s <= kms) {
    auto div = [&](int k) -> double {
      if (IST == 1)
        return ((double)D.uh[IX(k, c)] - (double)D.uh[IX(k, lW)] + (double)D.uhpp[IX(k, c)] - (double)D.uhpp[IX(k, lW)]) / D.P.twodx +
               ((double)D.vh[IX(k, c)] - (double)D.vh[IX(k, lS)] + (double)D.vhpp[IX(k, c)] - (double)D.vhpp[IX(k, lS)]) / D.P.twody;
      return ((double)D.uhp[IX(k, c)] - (double)D.uhp[IX(k, lW)]) / D.P.dx +
             ((double)D.vhp[IX(k, c)] - (double)D.vhp[IX(k, lS)]) / D.P.dy;
    };
    for (int k = 0; k <= D.km1 + 1; ++k)
      if (k < k1s || k > kms) D.wp[IX(k, c)] = R(0);
    // Levels in groups of four: the group's loads are issued together before
    // its sums and stores (a level at a time, each level's loads wait for the
    // previous level's store in program order).  Levels of a group above k1s
    // read a valid level (clamped at 0) and are discarded.
    double w = 0.0;
    for (int k = kms; k >= k1s; k -= 4) {
      double d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) d[j] = div(max(k - j, 0));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (k - j < k1s) break;
        w -= d[j];
        if (k - j > k1s) D.wp[IX(k - j, c)] = (R)w;
      }
    }
    for (int k = D.k1; k <= k1s - 1; ++k) w -= div(k);
    D.wp[IX(k1s, c)] = (R)w;
  }
  if (IST == 1) {
    // top indices after drying (vel 2490-2497, 2535-2541, 2547-2561)
    const int S = D.lm;
    const int lE = D.lE[c], lN = D.lN[c];
    if (lE != S) {
      const int k0 = D.k1u[c];
      if (k0 <= min(D.kmz[lE], kms) && (double)D.hu[IX(k0, c)] <= 0.0) D.k1u[c] = k0 + 1;
    }
    if (lN != S) {
      const int k0 = D.k1v[c];
      if (k0 <= min(D.kmz[lN], kms) && (double)D.hv[IX(k0, c)] <= 0.0) D.k1v[c] = k0 + 1;
    }
    const int k0 = D.k1z[c];
    if ((double)D.h[IX(k0, c)] <= 0.0) D.k1z[c] = min(k0 + 1, km1c(D, c));
  }
}

// Superbee limiter C_f (si3d_procedures.f90:2928) and the TVD face flux.
__device__ __forceinline__ double tvd_flux(double vel, double s1, double s2, double s3, double s4,
                                           double cfdef, double courant) {
  double cf = cfdef;
  const double d23 = s3 - s2;
  if (d23 != 0.0) {
    double ratio = (vel >= 0.0) ? (s2 - s1) / d23 : (s4 - s3) / d23;
    cf = dmax(0.0, dmax(dmin(1.0, 2. * ratio), dmin(2.0, ratio)));
  }
  return vel / 2. * (s3 + s2) - ((1. - cf) * fabs(vel) + vel * vel * courant * cf) * d23 / 2.;
}

// exsal (part 1): east-west and north-south TVD fluxes at the faces of each
// column, si3d_procedures.f90:2885-2975.  Only the wet range of each column is
// rewritten; the arrays persist across calls like the reference's Bth2/Bth4.
template <typename R>
__device__ __forceinline__ void exsal_hflux_cell(const Dev<R> &D, int c, int k, double twodt1, const R *cpp) {
  const int kms = D.kmz[c], k1s = D.k1z[c];
  if (k < k1s || k > kms) return;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  double fe = 0.0, fn = 0.0;
  if ((double)D.hup[IX(k, c)] > 0.0) {
    const int lEE = D.lE[lE];
    const double vel = ((double)D.uhpp[IX(k, c)] + (double)D.uh[IX(k, c)]) / 2.;
    const double s2 = cpp[IX(k, c)], s3 = cpp[IX(k, lE)];
    const double s1 = ((double)D.hpp[IX(k, lW)] <= 0.0) ? s2 : (double)cpp[IX(k, lW)];
    const double s4 = ((double)D.hpp[IX(k, lEE)] <= 0.0) ? s3 : (double)cpp[IX(k, lEE)];
    fe = tvd_flux(vel, s1, s2, s3, s4, 0.0, twodt1 / D.P.dx);
  }
  if ((double)D.hvp[IX(k, c)] > 0.0) {
    const int lNN = D.lN[lN];
    const double vel = ((double)D.vhpp[IX(k, c)] + (double)D.vh[IX(k, c)]) / 2.;
    const double s2 = cpp[IX(k, c)], s3 = cpp[IX(k, lN)];
    const double s1 = ((double)D.hpp[IX(k, lS)] <= 0.0) ? s2 : (double)cpp[IX(k, lS)];
    const double s4 = ((double)D.hpp[IX(k, lNN)] <= 0.0) ? s3 : (double)cpp[IX(k, lNN)];
    fn = tvd_flux(vel, s1, s2, s3, s4, 0.0, twodt1 / D.P.dy);
  }
  D.fx_ew[IX(k, c)] = (R)fe;
  D.fx_ns[IX(k, c)] = (R)fn;
}

template <typename R, bool CH>
__global__ void __launch_bounds__(kBlock) k_exsal_hflux(Dev<R> D, Scal<R> sv, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) exsal_hflux_cell<R>(D, t.c, k, twodt1, sv.cpp);
}

// Vertical TVD flux through the top of cell k (si3d_procedures.f90:2999-3035).
template <typename R>
__device__ __forceinline__ double vflux(const Dev<R> &D, int c, int k, int k1s, double twodt1, const R *cpp) {
  if (!((double)D.hp[IX(k - 1, c)] > 0.0)) return 0.0;
  const double vel = (k == k1s) ? 0.0 : (double)D.wp[IX(k, c)];
  const double s2 = cpp[IX(k, c)], s3 = cpp[IX(k - 1, c)];
  const double s4 = ((double)D.hpp[IX(k - 2, c)] <= 0.0) ? s3 : (double)cpp[IX(k - 2, c)];
  const double s1 = ((double)D.hpp[IX(k + 1, c)] <= 0.0) ? s2 : (double)cpp[IX(k + 1, c)];
  const double delz = ((double)D.hp[IX(k, c)] + (double)D.hp[IX(k - 1, c)]) / 2.;
  return tvd_flux(vel, s1, s2, s3, s4, 1.0, twodt1 / delz);
}

// exsal (part 2) fused with imsal: flux divergence + horizontal diffusion form
// the right-hand side (3038-3067), then the implicit vertical diffusion column
// solve (imsal, 3079-3207, trid1 3292-3352) produces sal.  With DO_SOLVE false
// only Bex is written (phase-level checking of exsal alone).
template <typename R, bool DO_EX, bool DO_SOLVE, bool TR>
__global__ void __launch_bounds__(kBlock) k_exsal_imsal(Dev<R> D, Scal<R> sv, double twodt1) {
  const int c = col();
  if (c >= D.lm) return;
  const int S = D.lm;
  const int kms = D.kmz[c], k1s = D.k1z[c];
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const int nwl = kms - k1s + 1;
  // Right-hand side of layer k (3038-3067); fud_k is the vertical flux through
  // the top of k, carried from the previous layer, fud_kp the one below.
  auto bh = [&](const R *hq, double a0, int w, int k) -> double {
    return (w == S || k > D.kmz[w]) ? 0.0 : a0 * (double)hq[IX(k, w)];
  };
  auto rhs = [&](int k, double fud_k, double fud_kp) -> double {
    const double sc = sv.cpp[IX(k, c)];
    double b = (double)D.hpp[IX(k, c)] * sc / twodt1 -
               ((double)D.fx_ew[IX(k, c)] - (lW == S ? 0.0 : (double)D.fx_ew[IX(k, lW)])) / D.P.dx -
               ((double)D.fx_ns[IX(k, c)] - (lS == S ? 0.0 : (double)D.fx_ns[IX(k, lS)])) / D.P.dy -
               (fud_k - fud_kp);
    if (D.P.ihd > 0) {
      b += (bh(D.hupp, D.P.Ax0, c, k) * ((double)sv.cpp[IX(k, lE)] - sc) -
            bh(D.hupp, D.P.Ax0, lW, k) * (sc - (double)sv.cpp[IX(k, lW)])) / D.P.dxdx +
           (bh(D.hvpp, D.P.Ay0, c, k) * ((double)sv.cpp[IX(k, lN)] - sc) -
            bh(D.hvpp, D.P.Ay0, lS, k) * (sc - (double)sv.cpp[IX(k, lS)])) / D.P.dydy;
    }
    return (double)(R)b;  // the reference stores Bex in the state precision
  };
  if (DO_EX && !DO_SOLVE) {  // exsal alone: Bex for every layer of the column
    for (int k = 1; k <= D.km1; ++k)
      if (k < k1s || k > kms) D.bex[IX(k, c)] = R(0);
    double fud_k = (k1s <= kms) ? vflux(D, c, k1s, k1s, twodt1, sv.cpp) : 0.0;
    for (int k = k1s; k <= kms; ++k) {
      const double fud_kp = (k + 1 <= kms) ? vflux(D, c, k + 1, k1s, twodt1, sv.cpp) : 0.0;
      D.bex[IX(k, c)] = (R)rhs(k, fud_k, fud_kp);
      fud_k = fud_kp;
    }
    return;
  }
  if (nwl < 1) {
    if (TR) {
      const double bsed = (double)(R)((double)D.hpp[IX(kms + 1, c)] * (double)sv.cpp[IX(kms + 1, c)] / twodt1);
      sv.cn[IX(kms + 1, c)] = (R)(bsed / ((double)D.hpp[IX(kms + 1, c)] / twodt1));
      for (int k = D.k1 - 1; k <= kms + 1; ++k)
        if ((double)sv.cn[IX(k, c)] < 0.0) sv.cn[IX(k, c)] = R(0);
    }
    return;
  }
  // Bex of layer k: computed in the sweep (fused) or read back (imsal alone),
  // plus the heat source (dsal = Bex + HeatSource, 3140, 3167)
  double fud = (DO_EX && k1s <= kms) ? vflux(D, c, k1s, k1s, twodt1, sv.cpp) : 0.0;
  auto bex = [&](int k) -> double {
    double b;
    if (!DO_EX) {
      b = (double)D.bex[IX(k, c)];
    } else {
      const double fud_kp = (k + 1 <= kms) ? vflux(D, c, k + 1, k1s, twodt1, sv.cpp) : 0.0;
      b = rhs(k, fud, fud_kp);
      fud = fud_kp;
    }
    if (!TR && sv.src) b += (double)sv.src[IX(k, c)];
    return b;
  };
  // imsal: layer thicknesses at time n with the surface-cell corrections (3124-3132)
  auto hn = [&](int k) -> double {
    if (k == k1s) return twodt1 * (double)D.wp[IX(k1s, c)] + (double)D.hpp[IX(k1s, c)];
    if (k == k1s + 1 && (double)D.hpp[IX(k1s, c)] <= 0.0) return D.hpp[IX(k1s + 1, c)];
    return D.h[IX(k, c)];
  };
  // tracers: Bex of the sediment layer (exTracer 4342) and the tail of ImTracer
  // (4449-4457): the sediment value from its own Bex, negatives clipped
  auto tracer_tail = [&]() {
    const double bsed = (double)(R)((double)D.hpp[IX(kms + 1, c)] * (double)sv.cpp[IX(kms + 1, c)] / twodt1);
    sv.cn[IX(kms + 1, c)] = (R)(bsed / ((double)D.hpp[IX(kms + 1, c)] / twodt1));
    for (int k = D.k1 - 1; k <= kms + 1; ++k)
      if ((double)sv.cn[IX(k, c)] < 0.0) sv.cn[IX(k, c)] = R(0);
  };
  if (nwl == 1) {
    double a = hn(k1s) / twodt1;
    double v = bex(k1s) / a;
    if (!TR && (double)D.h[IX(k1s, c)] < 1.E-2) v = sv.cpp[IX(k1s, c)];
    sv.cn[IX(k1s, c)] = (R)v;
    if (TR) tracer_tail();
    return;
  }
  // forward e/f sweep: e -> t0, f -> sal (in place), then back substitution.
  // The coupling through the top of layer k (a1) is the one through the
  // bottom of layer k-1 (a3), carried; 1/twodt1 and the pivots are multiplied
  // by as reciprocals (a change of rounding only).
  const double itw = 1.0 / twodt1;
  double hk = hn(k1s), hkp = hn(k1s + 1);
  double ep = 0.0, fp = 0.0, a3p = 0.0;
  for (int k = k1s; k <= kms; ++k) {
    const double a1 = a3p;
    const double a3 = (k < kms) ? -(double)D.Dv[IX(k + 1, c)] / (hk + hkp) * 2. : 0.0;
    const double a2 = hk * itw - a1 - a3;
    const double d = bex(k);
    if (k == k1s) {
      const double rd = 1.0 / a2;
      ep = -a3 * rd;
      fp = d * rd;
    } else if (k < kms) {
      const double rd = 1.0 / (a2 + a1 * ep);
      ep = -a3 * rd;
      fp = (d - a1 * fp) * rd;
    } else {
      fp = (d - a1 * fp) / (a2 + a1 * ep);  // bottom value
      ep = 0.0;
    }
    D.t0[IX(k, c)] = (R)ep;
    sv.cn[IX(k, c)] = (R)fp;
    a3p = a3;
    hk = hkp;
    hkp = (k + 2 <= kms) ? hn(k + 2) : 0.0;
  }
  // back substitution, the factors of four layers loaded together (layers
  // above k1s in a group read a valid layer, clamped at 0, and are discarded)
  double x = fp;
  for (int kg = kms - 1; kg >= k1s; kg -= 4) {
    R te[4], tf[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kk = max(kg - j, 0);
      te[j] = D.t0[IX(kk, c)];
      tf[j] = sv.cn[IX(kk, c)];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (kg - j < k1s) break;
      x = (double)te[j] * x + (double)tf[j];
      sv.cn[IX(kg - j, c)] = (R)x;
    }
  }
  for (int k = D.k1; k <= k1s - 1; ++k) sv.cn[IX(k, c)] = (R)x;
  if (TR) {
    sv.cn[IX(D.k1 - 1, c)] = (R)x;
    tracer_tail();
  } else {
    sv.cn[IX(kms + 1, c)] = sv.cn[IX(kms, c)];
  }
}

// exsal (part 2) per cell: Bex of layer k (3038-3067), with the vertical TVD
// fluxes through its top and bottom faces; the column solve (imsal) then reads
// it back.  One lane per (column, layer) keeps the flux arithmetic off the
// column sweep's serial path.
template <typename R, bool CH, bool TR>
__global__ void __launch_bounds__(kBlock) k_exsal_rhs(Dev<R> D, Scal<R> sv, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const int kms = D.kmz[c], k1s = D.k1z[c];
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const int kb = max(t.k0, k1s), ke = CH ? min(t.k1, kms + 1) : min(min(t.k1, t.k0 + 1), kms + 1);
  double fud_kp = (kb < ke) ? vflux(D, c, kb, k1s, twodt1, sv.cpp) : 0.0;  // carried down the chunk
  for (int k = kb; k < ke; ++k) {
    const double fud_k = fud_kp;
    fud_kp = (k + 1 <= kms) ? vflux(D, c, k + 1, k1s, twodt1, sv.cpp) : 0.0;
    auto bh = [&](const R *hq, double a0, int w) -> double {
      return (w == S || k > D.kmz[w]) ? 0.0 : a0 * (double)hq[IX(k, w)];
    };
    const double sc = sv.cpp[IX(k, c)];
    double b = (double)D.hpp[IX(k, c)] * sc / twodt1 -
               ((double)D.fx_ew[IX(k, c)] - (lW == S ? 0.0 : (double)D.fx_ew[IX(k, lW)])) / D.P.dx -
               ((double)D.fx_ns[IX(k, c)] - (lS == S ? 0.0 : (double)D.fx_ns[IX(k, lS)])) / D.P.dy -
               (fud_k - fud_kp);
    if (D.P.ihd > 0) {
      b += (bh(D.hupp, D.P.Ax0, c) * ((double)sv.cpp[IX(k, lE)] - sc) -
            bh(D.hupp, D.P.Ax0, lW) * (sc - (double)sv.cpp[IX(k, lW)])) / D.P.dxdx +
           (bh(D.hvpp, D.P.Ay0, c) * ((double)sv.cpp[IX(k, lN)] - sc) -
            bh(D.hvpp, D.P.Ay0, lS) * (sc - (double)sv.cpp[IX(k, lS)])) / D.P.dydy;
    }
    D.bex[IX(k, c)] = (R)b;
  }
  if (TR && kms >= t.k0 && kms < t.k1)  // Bex of the sediment layer (exTracer 4342)
    D.bex[IX(kms + 1, c)] = (R)((double)D.hpp[IX(kms + 1, c)] * (double)sv.cpp[IX(kms + 1, c)] / twodt1);
}

// Windowed forms of exsal's two per-cell kernels: the same values in the same
// order as k_exsal_hflux / k_exsal_rhs, with the chunk's layers walked
// together by the whole workgroup (k is uniform, so the fields are read
// through buffer resources and one 32-bit offset per neighbour column), the
// column's indices loaded once per chunk, and, in the right-hand side, the
// vertical neighbours k-2 .. k+2 of the scalar and of hpp, k-1 .. k+1 of hp
// and wp at k, k+1 carried down the chunk in registers.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_exsal_hflux_w(Dev<R> D, Scal<R> sv, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c;
  const int kms = D.kmz[c], k1s = D.k1z[c];
  const int kb = max(t.k0, k1s), ke = min(t.k1, kms + 1);
  if (kb >= ke) return;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  const int lEE = D.lE[lE], lNN = D.lN[lN];
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz, oE = (unsigned)lE * sz, oW = (unsigned)lW * sz, oN = (unsigned)lN * sz,
                 oS = (unsigned)lS * sz, oEE = (unsigned)lEE * sz, oNN = (unsigned)lNN * sz;
  const auto cpp = rsrc(sv.cpp, nb), hpp = rsrc(D.hpp, nb), hup = rsrc(D.hup, nb), hvp = rsrc(D.hvp, nb);
  const auto uhpp = rsrc(D.uhpp, nb), uh = rsrc(D.uh, nb), vhpp = rsrc(D.vhpp, nb), vh = rsrc(D.vh, nb);
  const auto few = rsrc(D.fx_ew, nb), fns = rsrc(D.fx_ns, nb);
  const double cx = twodt1 / D.P.dx, cy = twodt1 / D.P.dy;
  unsigned pk = (unsigned)t.k0 * pb;
  for (int k = t.k0; k < t.k1; ++k, pk += pb) {
    if (k >= ke) break;
    if (k < kb) continue;
    double fe = 0.0, fn = 0.0;
    const double s2 = bld<R>(cpp, oc, pk);
    if (bld<R>(hup, oc, pk) > 0.0) {
      const double vel = (bld<R>(uhpp, oc, pk) + bld<R>(uh, oc, pk)) / 2.;
      const double s3 = bld<R>(cpp, oE, pk);
      const double s1 = (bld<R>(hpp, oW, pk) <= 0.0) ? s2 : bld<R>(cpp, oW, pk);
      const double s4 = (bld<R>(hpp, oEE, pk) <= 0.0) ? s3 : bld<R>(cpp, oEE, pk);
      fe = tvd_flux(vel, s1, s2, s3, s4, 0.0, cx);
    }
    if (bld<R>(hvp, oc, pk) > 0.0) {
      const double vel = (bld<R>(vhpp, oc, pk) + bld<R>(vh, oc, pk)) / 2.;
      const double s3 = bld<R>(cpp, oN, pk);
      const double s1 = (bld<R>(hpp, oS, pk) <= 0.0) ? s2 : bld<R>(cpp, oS, pk);
      const double s4 = (bld<R>(hpp, oNN, pk) <= 0.0) ? s3 : bld<R>(cpp, oNN, pk);
      fn = tvd_flux(vel, s1, s2, s3, s4, 0.0, cy);
    }
    bst<R>(few, (R)fe, oc, pk);
    bst<R>(fns, (R)fn, oc, pk);
  }
}

template <typename R, bool TR>
__global__ void __launch_bounds__(kBlock) k_exsal_rhs_w(Dev<R> D, Scal<R> sv, double twodt1, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const int kms = D.kmz[c], k1s = D.k1z[c];
  const int kb = max(t.k0, k1s), ke = min(t.k1, kms + 1);
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz;
  const auto cpp = rsrc(sv.cpp, nb), hpp = rsrc(D.hpp, nb), bex = rsrc(D.bex, nb);
  if (kb < ke) {
    const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
    const int kmW = D.kmz[lW], kmS = D.kmz[lS];
    const unsigned oE = (unsigned)lE * sz, oW = (unsigned)lW * sz, oN = (unsigned)lN * sz, oS = (unsigned)lS * sz;
    const auto hp = rsrc(D.hp, nb), wp = rsrc(D.wp, nb), few = rsrc(D.fx_ew, nb), fns = rsrc(D.fx_ns, nb);
    const auto hupp = rsrc(D.hupp, nb), hvpp = rsrc(D.hvpp, nb);
    const bool ihd = D.P.ihd > 0, wW = lW != S, wS = lS != S;
    // vertical TVD flux through the top of layer j (vflux), from the window
    auto vf = [&](double hp_above, double vel, double s1, double s2, double s3, double s4, double hp_j) -> double {
      if (!(hp_above > 0.0)) return 0.0;
      const double delz = (hp_j + hp_above) / 2.;
      return tvd_flux(vel, s1, s2, s3, s4, 1.0, twodt1 / delz);
    };
    unsigned pk = (unsigned)t.k0 * pb;
    double cm2 = bld<R>(cpp, oc, pk - 2 * pb), cm1 = bld<R>(cpp, oc, pk - pb), c0 = bld<R>(cpp, oc, pk),
           c1 = bld<R>(cpp, oc, pk + pb);
    double hm2 = bld<R>(hpp, oc, pk - 2 * pb), hm1 = bld<R>(hpp, oc, pk - pb), h0 = bld<R>(hpp, oc, pk),
           h1 = bld<R>(hpp, oc, pk + pb);
    double pm1 = bld<R>(hp, oc, pk - pb), p0 = bld<R>(hp, oc, pk);
    double w0 = bld<R>(wp, oc, pk);
    double fud_kp = 0.0;
    for (int k = t.k0; k < t.k1; ++k, pk += pb) {
      if (k >= ke) break;
      const double c2 = bld<R>(cpp, oc, pk + 2 * pb), h2 = bld<R>(hpp, oc, pk + 2 * pb);
      const double p1 = bld<R>(hp, oc, pk + pb), w1 = bld<R>(wp, oc, pk + pb);
      if (k >= kb) {
        // flux through the top of k: carried, or evaluated at the first layer
        const double fud_k =
            (k == kb) ? vf(pm1, (k == k1s) ? 0.0 : w0, (h1 <= 0.0) ? c0 : c1, c0, cm1, (hm2 <= 0.0) ? cm1 : cm2, p0)
                      : fud_kp;
        fud_kp = (k + 1 <= kms) ? vf(p0, w1, (h2 <= 0.0) ? c1 : c2, c1, c0, (hm1 <= 0.0) ? c0 : cm1, p1) : 0.0;
        const double sc = c0;
        double b = h0 * sc / twodt1 -
                   (bld<R>(few, oc, pk) - (wW ? bld<R>(few, oW, pk) : 0.0)) / D.P.dx -
                   (bld<R>(fns, oc, pk) - (wS ? bld<R>(fns, oS, pk) : 0.0)) / D.P.dy - (fud_k - fud_kp);
        if (ihd) {
          const double bc = (k > kms) ? 0.0 : D.P.Ax0 * bld<R>(hupp, oc, pk);
          const double bw = (!wW || k > kmW) ? 0.0 : D.P.Ax0 * bld<R>(hupp, oW, pk);
          const double bn = (k > kms) ? 0.0 : D.P.Ay0 * bld<R>(hvpp, oc, pk);
          const double bs = (!wS || k > kmS) ? 0.0 : D.P.Ay0 * bld<R>(hvpp, oS, pk);
          b += (bc * (bld<R>(cpp, oE, pk) - sc) - bw * (sc - bld<R>(cpp, oW, pk))) / D.P.dxdx +
               (bn * (bld<R>(cpp, oN, pk) - sc) - bs * (sc - bld<R>(cpp, oS, pk))) / D.P.dydy;
        }
        bst<R>(bex, (R)b, oc, pk);
      }
      cm2 = cm1, cm1 = c0, c0 = c1, c1 = c2;
      hm2 = hm1, hm1 = h0, h0 = h1, h1 = h2;
      pm1 = p0, p0 = p1;
      w0 = w1;
    }
  }
  if (TR && kms >= t.k0 && kms < t.k1) {  // Bex of the sediment layer (exTracer 4342)
    const unsigned ps = (unsigned)(kms + 1) * pb;
    bst<R>(bex, (R)(bld<R>(hpp, oc, ps) * bld<R>(cpp, oc, ps) / twodt1), oc, ps);
  }
}

// vel2: layer_h (si3d_procedures.f90:2181-2231) then the top-cell flux split for
// wetting (2570-2631).  One lane per (column, layer k = 2..km); the
// lane of the top face layer performs the split for both layers it touches,
// evaluating layer_h for the layer above itself.
template <typename R, bool X>
__device__ __forceinline__ void vel2_face(const Dev<R> &D, int c, int k, double sc) {
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == D.lm) return;
  const int kmq = min(D.kmz[c], D.kmz[ln]);
  const double sm = dmax(sc, D.s[ln]);
  const double hh = X ? D.hhu[c] : D.hhv[c];
  auto hlay = [&](int kk) {
    const double *zl = zlv(D, c);
    double v = dmin(zl[kk + 1], hh) - dmax(zl[kk], -sm);
    return (v <= 0.0) ? 0.0 : v;
  };
  R *hq = X ? D.hu : D.hv;
  R *qh = X ? D.uh : D.vh;
  if (k <= kmq) hq[IX(k, c)] = (R)hlay(k);
  const int kt = X ? D.k1u[c] : D.k1v[c];
  if (k != kt) return;
  // thickness of the layer above the top face after layer_h (kt-1 >= k1 is recomputed,
  // layer 1 keeps its value)
  const double hm = (kt - 1 >= D.k1 && kt - 1 <= kmq) ? (double)(R)hlay(kt - 1) : (double)hq[IX(kt - 1, c)];
  if (hm > 0.0) {
    const double hk = (kt <= kmq) ? (double)(R)hlay(kt) : (double)hq[IX(kt, c)];
    const double t = (double)qh[IX(kt, c)] / (hk + hm);
    qh[IX(kt, c)] = (R)(t * hk);
    qh[IX(kt - 1, c)] = (R)(t * hm);
  }
}

template <typename R, bool CH>
__global__ void __launch_bounds__(kBlock) k_vel2(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c;
  const double sc = D.s[c];
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) {
    if (k <= D.kmz[c]) {
      double v = dmin(zlv(D, c)[k + 1], D.hhs[c]) - dmax(zlv(D, c)[k], -sc);
      D.h[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
    }
    vel2_face<R, true>(D, c, k, sc);
    vel2_face<R, false>(D, c, k, sc);
  }
}

// Windowed form of k_vel2 (same values, same order): the column's levels,
// face bottoms and neighbour surface elevations are read once per chunk of
// layers, the level pair zl(k), zl(k+1) is carried down the chunk, and the
// thicknesses are stored through buffer resources with a scalar plane offset.
// The top-face flux split (one lane per column and face) keeps plain loads.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_vel2_w(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 2, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz;
  const int kmz = D.kmz[c], k1 = D.k1;
  const int lE = D.lE[c], lN = D.lN[c];
  const bool fx = lE != S, fy = lN != S;
  const int kmu = fx ? min(kmz, D.kmz[lE]) : -1, kmv = fy ? min(kmz, D.kmz[lN]) : -1;
  const int k1u = D.k1u[c], k1v = D.k1v[c];
  const double sc = D.s[c], hhs = D.hhs[c], hhu = D.hhu[c], hhv = D.hhv[c];
  const double smE = fx ? dmax(sc, D.s[lE]) : 0.0, smN = fy ? dmax(sc, D.s[lN]) : 0.0;
  const double *zl = zlv(D, c);
  const auto h = rsrc(D.h, nb), hu = rsrc(D.hu, nb), hv = rsrc(D.hv, nb);
  auto lay = [](double zt, double zb, double hh, double sm) {
    const double v = dmin(zt, hh) - dmax(zb, -sm);
    return (v <= 0.0) ? 0.0 : v;
  };
  // the top-cell flux split of one face (vel2_face after its thickness)
  auto split = [&](R *hq, R *qh, int kt, int kmq, double hh, double sm) {
    const double hm = (kt - 1 >= k1 && kt - 1 <= kmq) ? (double)(R)lay(zl[kt], zl[kt - 1], hh, sm)
                                                      : (double)hq[IX(kt - 1, c)];
    if (hm > 0.0) {
      const double hk = (kt <= kmq) ? (double)(R)lay(zl[kt + 1], zl[kt], hh, sm) : (double)hq[IX(kt, c)];
      const double tq = (double)qh[IX(kt, c)] / (hk + hm);
      qh[IX(kt, c)] = (R)(tq * hk);
      qh[IX(kt - 1, c)] = (R)(tq * hm);
    }
  };
  // nothing below the column's bed other than a top face's split
  const int kend = min(t.k1, max(kmz, max(fx ? k1u : -1, fy ? k1v : -1)) + 1);
  if (t.k0 >= kend) return;
  double zk = zl[t.k0];
  for (int k = t.k0; k < kend; ++k) {
    const unsigned pk = (unsigned)k * pb;
    const double zk1 = zl[k + 1];
    if (k <= kmz) {
      const double v = dmin(zk1, hhs) - dmax(zk, -sc);
      bst<R>(h, (v <= 0.0) ? R(0) : (R)v, oc, pk);
    }
    if (fx) {
      if (k <= kmu) bst<R>(hu, (R)lay(zk1, zk, hhu, smE), oc, pk);
      if (k == k1u) split(D.hu, D.uh, k1u, kmu, hhu, smE);
    }
    if (fy) {
      if (k <= kmv) bst<R>(hv, (R)lay(zk1, zk, hhv, smN), oc, pk);
      if (k == k1v) split(D.hv, D.vh, k1v, kmv, hhv, smN);
    }
    zk = zk1;
  }
}

// Fraction of the surface shortwave flux still present at depth z below the
// free surface, Henderson-Sellers (Engineering Limnology, eq. 2.25) as
// SolarFr (si3d_boundaryconditions.f90:4786-4816).
__device__ __forceinline__ double solar_fraction(double eta, double z) {
  if (eta >= 0.1) {
    const double beta = 0.265 * log(eta) + 0.614;
    return (z < 0.60) ? 1. - beta * z / 0.60 : (1. - beta) * exp(-eta * (z - 0.60));
  }
  return exp(-eta * z);
}

// Surface heat budget per column (ifsurfbc 1-3): the shortwave fraction of
// each wet layer (DistributeQswH, si3d_boundaryconditions.f90:4672-4727) and
// HeatSource in temperature units (DistributeMomentumHeatSourcesH, 5144-5485):
// the non-penetrative fluxes enter the surface layer, shortwave is spread over
// the column.  HeatSource is zero above and below the wet layers.
// HeatSource of one column from its non-penetrative surface flux `surf` (W/m2)
// and its shortwave Qsw spread over the wet layers with attenuation eta.
template <typename R>
__device__ void heat_column(const Dev<R> &D, int c, double eta, double Qsw, double surf) {
  constexpr double kCp = 4181.6;
  const int kms = D.kmz[c], k1s = D.k1z[c], nwl = kms - k1s + 1;
  for (int k = 1; k <= D.km1; ++k) D.heat[IX(k, c)] = R(0);
  if (nwl < 1) {  // the surface term alone, as the reference leaves it
    D.heat[IX(k1s, c)] = (R)surf;
    return;
  }
  // shortwave fractions, with what reaches the bottom shared evenly (remFr)
  double rem = 0.0;
  if (nwl >= 2) {
    double z = 0.0, sum = 0.0;
    for (int k = k1s; k <= kms; ++k) {
      const double zn = z + (double)D.hp[IX(k, c)];
      sum += solar_fraction(eta, z) - solar_fraction(eta, zn);
      z = zn;
    }
    rem = 1. - sum;
    rem = (rem > 1.e-6) ? rem / nwl : 0.0;
  }
  double z = 0.0;
  for (int k = k1s; k <= kms; ++k) {
    double fr = 1.0;
    if (nwl >= 2) {
      const double zn = z + (double)D.hp[IX(k, c)];
      fr = solar_fraction(eta, z) - solar_fraction(eta, zn) + rem;
      z = zn;
    }
    double hs = ((k == k1s ? surf : 0.0) + Qsw * fr) / (((double)D.rhop[IX(k, c)] + 1000.) * kCp);
    if (Qsw < -100.) hs = 0.0;
    D.heat[IX(k, c)] = (R)hs;
  }
}

// Longwave, latent and sensible flux into the surface layer (W/m2) for the
// budget modes (si3d_boundaryconditions.f90:5268-5316): Qbra the absorbed
// atmospheric longwave, ea the air's vapour pressure, cE / cH the Kondo
// coefficients at wind speed ws.
template <typename R>
__device__ double surface_flux(const Dev<R> &D, int c, double Qbra, double ea, double Ta, double Pa, double ws,
                               double cE, double cH) {
  constexpr double kStefanBoltzmann = 5.6697e-8, kEmWater = 0.97, kCpAir = 1012.0;
  const double ts = D.salp[IX(D.k1z[c], c)];
  const double T = ts + 273.;
  const double esw = 2.1718e10 * exp(-4157. / (T - 33.91));
  const double lv = 1.91846e6 * (T / (T - 33.91)) * (T / (T - 33.91));
  return Qbra - kEmWater * kStefanBoltzmann * T * T * T * T - RHOAIR * lv * cE * (esw - ea) * 0.622 / Pa * ws -
         RHOAIR * kCpAir * cH * (ts - Ta) * ws;
}

// Shortwave fractions of each wet layer (DistributeQswH,
// si3d_boundaryconditions.f90:4672-4727) and HeatSource in temperature units
// (DistributeMomentumHeatSourcesH, 5144-5485) for the spatially uniform
// budget modes 1-3: the non-penetrative fluxes enter the surface layer,
// shortwave is spread over the column; HeatSource is zero above and below
// the wet layers.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_heat_source(Dev<R> D, HeatForce H) {
  const int c = col();
  if (c >= D.lm) return;
  const double surf = H.mode == 1 ? H.Qn - H.Qsw
                                  : surface_flux(D, c, H.Qbra, H.ea, H.Ta, H.Pa, H.ws, H.coeffE, H.coeffH);
  heat_column(D, c, H.eta, H.Qsw, surf);
}

// Kondo (1975) neutral transfer coefficients of vapour (cE) and heat (cH) at
// wind speed ws: the five wind ranges of si3d_boundaryconditions.f90:5166-5178.
__device__ void kondo(double ws, double &cE, double &cH) {
  const double lim[6] = {0.0, 2.2, 5.0, 8.0, 25.0, 50.0};
  const double ah[5] = {0.000, 0.9270, 1.1500, 1.1700, 1.652}, bh[5] = {1.185, 0.0521, 0.0100, 0.0075, -0.017},
               ch[5] = {0.000, 0.0000, 0.0000, -4.5E-4, 0.000}, ph[5] = {-0.157, 1., 1., 1., 1.};
  const double ae[5] = {0.000, 0.9690, 1.1800, 1.1960, 1.680}, be[5] = {1.230, 0.0521, 0.0100, 0.0080, -0.016},
               ce[5] = {0.000, 0.0000, 0.0000, -4.0E-4, 0.000}, pe[5] = {-0.160, 1., 1., 1., 1.};
  int r = 4;
  for (int q = 0; q < 4; ++q)
    if (ws > lim[q] && ws <= lim[q + 1]) {
      r = q;
      break;
    }
  cH = (ah[r] + bh[r] * pow(ws, ph[r]) + ch[r] * (ws - 8.) * (ws - 8.)) * 1.e-3;
  cE = (ae[r] + be[r] * pow(ws, pe[r]) + ce[r] * (ws - 8.) * (ws - 8.)) * 1.e-3;
}

// The met-station modes per wet column (ifsurfbc 10, 11; InterpMetData and
// DistributeMomentumHeatSourcesH cases 10 / 11, si3d_boundaryconditions.f90:
// 5331-5483, 5703-5748): the station values weighted by the column's weights
// W[m][c] (the two-pass scheme divides by their sum and adds the residuals
// weighted by W^(1/gammaB), normalised), the column's wind and its drag (icw
// 0: cw; 1: Amorocho & DeVries; 2: Andreas et al., whose square-root term
// the reference raises to (1/2) = 0) into D.wind, and HeatSource.  Mode 10
// computes the air's longwave from a cloud cover of 0 (the module value the
// reference reads there, which that mode never sets).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_met_heat(Dev<R> D, MetArgs M, const double *W, double *wind) {
  const int c = col();
  if (c >= D.lm) return;
  double x[6] = {0, 0, 0, 0, 0, 0};
  double sw = 0.0, sg = 0.0;
  for (int m = 0; m < M.nmet; ++m) {
    const double w = W[(size_t)m * D.LP + c];
    sw += w;
    for (int q = 0; q < 6; ++q) x[q] += w * M.st[6 * m + q];
  }
  if (M.iinterp == 2) {
    for (int q = 0; q < 6; ++q) x[q] /= sw;
    for (int m = 0; m < M.nmet; ++m) sg += pow(W[(size_t)m * D.LP + c], 1.0 / M.gamma);
    for (int m = 0; m < M.nmet; ++m) {
      const double g = pow(W[(size_t)m * D.LP + c], 1.0 / M.gamma) / sg;
      for (int q = 0; q < 6; ++q) x[q] += g * M.res[6 * m + q];
    }
  }
  const double ua = x[4], va = x[5], ws = sqrt(ua * ua + va * va);
  double cdw;
  if (M.icw == 0) cdw = M.cw;
  else if (M.icw == 1) cdw = 0.0015 / (1.0 + exp((12.5 - ws) / 1.56)) + 0.00104;
  else if (ws == 0.0) cdw = M.cw;
  else {
    const double us = (0.239 + 0.0433 * ((ws - 8.271) + 1.0)) / ws;
    cdw = us * us;
  }
  wind[3 * c] = ua;
  wind[3 * c + 1] = va;
  wind[3 * c + 2] = cdw;
  constexpr double kStefanBoltzmann = 5.6697e-8, kAlbedoLw = 0.03;
  const double Qsw = x[0], Ta = x[1], TaK = Ta + 273.;
  const double ea = 2.1718e10 * exp(-4157. / (TaK - 33.91)) * x[2];
  double Qbra;
  if (M.mode == 10) {
    const double em_air = 0.642 * pow(ea / TaK, 0.14285714);  // cloud cover 0
    Qbra = em_air * kStefanBoltzmann * TaK * TaK * TaK * TaK * (1. - kAlbedoLw);
  } else {
    Qbra = x[3] * (1. - kAlbedoLw);
  }
  double cE, cH;
  kondo(ws, cE, cH);
  heat_column(D, c, M.eta, Qsw, surface_flux(D, c, Qbra, ea, Ta, M.Pa, ws, cE, cH));
}

// UpdateHorizontalMixingCoefficients (Smagorinsky), si3d_mixing.f90:894-966.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_hmix(Dev<R> D) {
  const int c = col();
  if (c >= D.lm) return;
  const int lE = D.lE[c], lW = D.lW[c], lN = D.lN[c], lS = D.lS[c];
  for (int k = D.k1z[c]; k <= D.kmz[c]; ++k) {
    double khv;
    if ((double)D.hp[IX(k, lW)] <= 0.0 || (double)D.hp[IX(k, lE)] <= 0.0 || (double)D.hp[IX(k, lS)] <= 0.0 ||
        (double)D.hp[IX(k, lN)] <= 0.0) {
      khv = D.P.Ax0;
    } else {
      double dudx = ((double)D.upp[IX(k, c)] - (double)D.upp[IX(k, lW)]) / D.P.dx;
      double dvdy = ((double)D.vpp[IX(k, c)] - (double)D.vpp[IX(k, lS)]) / D.P.dy;
      double dudy = (((double)D.upp[IX(k, lN)] + (double)D.upp[IX(k, D.lN[lW])]) -
                     ((double)D.upp[IX(k, lS)] + (double)D.upp[IX(k, D.lS[lW])])) / D.P.fourdy;
      double dvdx = (((double)D.vpp[IX(k, lE)] + (double)D.vpp[IX(k, D.lE[lS])]) -
                     ((double)D.vpp[IX(k, lW)] + (double)D.vpp[IX(k, D.lW[lS])])) / D.P.fourdx;
      khv = dmax(D.P.Ax0, 0.03 * D.P.dx * D.P.dy * sqrt(dudx * dudx + dvdy * dvdy + .5 * (dudy + dvdx) * (dudy + dvdx)));
    }
    D.kh[IX(k, c)] = (R)khv;
  }
}

// The p-level values of level k of column m that the interface terms of
// levels k and k+1 share: hp, rhop, up at m and its west face, vp at m and
// its south face.  The column sweep loads each level once and carries it.
// A pair of values of the storage type (one 16- or 8-byte access).
template <typename R>
struct Pair;
template <>
struct Pair<double> {
  using T = double2;
  static __device__ __forceinline__ T make(double a, double b) { return make_double2(a, b); }
};
template <>
struct Pair<float> {
  using T = float2;
  static __device__ __forceinline__ T make(double a, double b) { return make_float2((float)a, (float)b); }
};

struct TurbRaw {
  double h, rho, um, uw, vm, vs;
};

template <typename R, typename IO>
__device__ __forceinline__ TurbRaw turb_raw(const Dev<R> &D, const IO &io, int m, int k, int lW, int lS) {
  return TurbRaw{(double)io.ld(D.hp, k, m), (double)io.ld(D.rhop, k, m), (double)io.ld(D.up, k, m),
                 (double)io.ld(D.up, k, lW), (double)io.ld(D.vp, k, m),  (double)io.ld(D.vp, k, lS)};
}

// Interface quantities of level k (k1s < k <= kms) of column m from the
// p-level fields (a = level k-1, b = level k): the limited length scale, shear
// and buoyancy production (the TKE source terms, si3d_mixing.f90:411-440) and
// the new eddy coefficients (245-290), plus q2p(k) and q2lp(k).  Av(k)/Dv(k)
// are read for the productions before the new values replace them.
struct TurbLev {
  double xl, sp, bp, av, dv, q2p, q2lp, rq, sq;  // rq = 1 / q2p, sq = sqrt(q2p)
};

template <typename R, typename IO>
__device__ __forceinline__ TurbLev turb_level(const Dev<R> &D, const IO &io, int m, int k, const TurbRaw &a,
                                             const TurbRaw &b) {
  TurbLev L;
  // The closure is bound by FP64 arithmetic (about 30 divisions per level as
  // written): divisions by a quantity used more than once are multiplications
  // by its reciprocal (1/dz, 1/q2p here; 1/twodt1 and the elimination pivots
  // in the sweep), which moves results by rounding only.
  const double hkm = a.h, hk = b.h;
  const double dz = 0.5 * (hkm + hk), idz = 1.0 / dz;
  const double rkm = a.rho, rk = b.rho;
  const double drdz = (rkm - rk) * idz, rden = 0.5 * (rkm + rk) + 1000.;
  const double N2 = -G / rden * drdz;
  const double dudz = 0.5 * (a.um + a.uw - b.um - b.uw) * idz;
  const double dvdz = 0.5 * (a.vm + a.vs - b.vm - b.vs) * idz;
  const double q2pk = io.ld(D.q2p, k, m), q2lpk = io.ld(D.q2lp, k, m);
  const double rq = 1.0 / q2pk;
  L.q2p = q2pk;
  L.q2lp = q2lpk;
  L.rq = rq;
  L.sq = sqrt(q2pk);
  const double xl0 = q2lpk * rq;
  double xl = xl0;
  if (N2 > 0.0) xl = dmin(xl, 0.53 * sqrt(q2pk / N2));
  L.xl = xl;
  L.sp = (double)io.ld(D.Av, k, m) * (dudz * dudz + dvdz * dvdz);
  L.bp = -(double)io.ld(D.Dv, k, m) * N2;
  // eddy coefficients (UpdateMixingCoefficients evaluates N^2 in this order)
  const double N2b = -G * drdz / rden;
  double xlb = xl0;
  if (N2b > 0.0) xlb = dmin(xlb, 0.53 * sqrt(q2pk / N2b));
  double Gh = -N2b * xlb * xlb * rq;
  Gh = dmax(GH_MIN, dmin(GH_MAX, Gh));
  const double Sh = D.P.t_1 / (1. - Gh * D.P.t_2);
  const double Sm = (D.P.t_3 + Sh * Gh * D.P.t_4) / (1. - D.P.t_5 * Gh);
  const double Rt = L.sq * xlb;
  L.av = Rt * Sm + AV_MOL;
  L.dv = Rt * Sh + DV_MOL;
  return L;
}

// Per cell (split variant): the interface quantities of level k into the
// FP64 scratch planes (length scale -> d0, shear production -> d1, buoyancy
// production -> d2) and the new Av/Dv in place; Av/Dv are zero outside
// k1s < k <= kms.  Columns with at most one wet layer are left alone.
template <typename R, bool CH>
__global__ void __launch_bounds__(kBlock) k_turb_cell(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm, 1, nk, tile, t)) return;
  const int m = t.c;
  const int kms = D.kmz[m], k1s = D.k1z[m];
  if (kms - k1s + 1 <= 1) return;
  const int lW = D.lW[m], lS = D.lS[m];
  const PtrIO io = make_io<PtrIO>(D);
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) {
    if (k <= k1s || k > kms) {
      io.st(D.Av, k, m, R(0)), io.st(D.Dv, k, m, R(0));
      continue;
    }
    const TurbLev L = turb_level(D, io, m, k, turb_raw(D, io, m, k - 1, lW, lS), turb_raw(D, io, m, k, lW, lS));
    io.st(D.d0, k, m, (R)L.xl);
    io.st(D.d1, k, m, (R)L.sp);
    io.st(D.d2, k, m, (R)L.bp);
    double av = L.av, dv = L.dv;
    if (k == k1s + 1 && (double)io.ld(D.hp, k1s, m) < 1.E-2) av = dmax((double)(R)av, 1.E-4), dv = dmax((double)(R)dv, 1.E-4);
    io.st(D.Av, k, m, (R)av);
    io.st(D.Dv, k, m, (R)dv);
  }
}

// Two-equation turbulence closure: SolveTKEandLengthScale (si3d_mixing.f90:
// 301-551, tridT 554-612) and the eddy coefficients of UpdateMixingCoefficients
// (245-290).  FP64 throughout, as the reference's REAL*8 turbulence state.
//
// SPLIT: the interface quantities and Av/Dv come from k_turb_cell (the sweep
// reads them from the scratch planes and writes its factors over them);
// otherwise they are computed one level ahead inside the sweep.
//
// With `budget` (trapezoidal calls, si3d_mixing.f90:466-475) the column also
// returns its share of the basin energy budget over levels k1s+1..kms, each
// weighted by rho h = (rhop + 1000) hp: shear and buoyancy production,
// dissipation q2 sqrt(q2p) / (B1 l) and q2, the new q2 after its floor.  The
// weights of the last two wait for the back substitution in single precision
// (the reference's rhoxh is single precision too); the sums are FP64.
struct TkeSums {
  double sh, by, di, tk;
};

template <typename R, bool SPLIT, typename IO>
__device__ __forceinline__ TkeSums turb_column(const Dev<R> &D, const IO &io, int m, int istep, bool budget) {
  TkeSums bs{0.0, 0.0, 0.0, 0.0};
  const int kms = D.kmz[m], k1s = D.k1z[m], nwl = kms - k1s + 1;
  if (nwl <= 1) return bs;
  const double twodt1 = D.P.twodt / istep, r3 = 3. / twodt1, itw = 1.0 / twodt1;
  const int lW = D.lW[m], lS = D.lS[m];
  // forward-sweep factors of the two systems: split, E1/E2 in d0/d1 and the
  // right-hand sides in q2/q2l; fused, (e1, e2) and (f1, f2) as pairs in d0 and
  // d1 (2 n3 doubles each), one 16-byte access per level and direction
  R *E1 = D.d0, *E2 = D.d1;
  using R2 = typename Pair<R>::T;
  R2 *EP = reinterpret_cast<R2 *>(D.d0), *FP = reinterpret_cast<R2 *>(D.d1);
  double hbelow = 0.0;
  for (int k = k1s + 1; k <= kms; ++k) hbelow += io.ld(D.hp, k, m);
  // interface quantities of level k: read back (split) or computed from the
  // carried raw values of level k-1 (a) and the new ones of level k (b)
  auto level = [&](int k, const TurbRaw &a, const TurbRaw &b) -> TurbLev {
    if (SPLIT) {
      TurbLev L{};
      L.xl = io.ld(D.d0, k, m);
      L.sp = io.ld(D.d1, k, m);
      L.bp = io.ld(D.d2, k, m);
      L.q2p = io.ld(D.q2p, k, m);
      L.q2lp = io.ld(D.q2lp, k, m);
      L.rq = 1.0 / L.q2p;
      L.sq = sqrt(L.q2p);
      return L;
    }
    return turb_level(D, io, m, k, a, b);
  };
  // boundary values (445-457)
  const double b1_23 = pow(B_1, 2. / 3.);
  double ub = 0.5 * ((double)io.ld(D.u, kms, m) + (double)io.ld(D.u, kms, lW));
  double vb = 0.5 * ((double)io.ld(D.v, kms, m) + (double)io.ld(D.v, kms, lS));
  const double ds_bot = 3. * b1_23 * (D.P.cd * (ub * ub + vb * vb)) / twodt1;
  const Wind wd = wind(D, m);
  double us = 0.5 * ((double)io.ld(D.up, D.k1u[m], m) + (double)io.ld(D.up, D.k1u[lW], lW));
  double vs = 0.5 * ((double)io.ld(D.vp, D.k1v[m], m) + (double)io.ld(D.vp, D.k1v[lS], lS));
  const double ds_top = 3. * b1_23 *
                        (wd.cdw * RHOAIR * ((wd.uair - us) * (wd.uair - us) + (wd.vair - vs) * (wd.vair - vs)) / 1000.) / twodt1;
  const bool thin_top = (double)io.ld(D.hp, k1s, m) < 1.E-2;
  // One forward sweep over rows k1s..kms of both systems (row k1s and row
  // kms+1 are the Dirichlet rows: a1 = a3 = 0, a2 = 3/twodt1), with the
  // interface quantities of level k+1 computed one level ahead.
  double e1 = 0.0, f1 = ds_top / r3, e2 = 0.0, f2 = 0.0;  // row k1s
  if (SPLIT) {
    io.st(D.q2, k1s, m, (R)f1);
    io.st(D.q2l, k1s, m, (R)f2);
    io.st(E1, k1s, m, R(0));
    io.st(E2, k1s, m, R(0));
  } else {
    io.st(EP, k1s, m, Pair<R>::make(0.0, 0.0));
    io.st(FP, k1s, m, Pair<R>::make(f1, f2));
  }
  TurbRaw rcur{};  // raw values of level k (unused when split)
  if (!SPLIT) rcur = turb_raw(D, io, m, k1s + 1, lW, lS);
  TurbLev cur = level(k1s + 1, SPLIT ? rcur : turb_raw(D, io, m, k1s, lW, lS), rcur);
  double dco_prev = SQ * 0.5 * (0.0 + cur.xl) * sqrt(0.5 * (io.ld(D.q2p, k1s, m) + cur.q2p));
  double zfromt = 0.0;
  double hkm = io.ld(D.hp, k1s, m);
  for (int k = k1s + 1; k <= kms; ++k) {
    const bool more = k + 1 <= kms;
    TurbRaw rnxt{};
    TurbLev nxt{};
    if (more) {
      if (!SPLIT) rnxt = turb_raw(D, io, m, k + 1, lW, lS);
      nxt = level(k + 1, rcur, rnxt);
    }
    const double xlk = cur.xl, xlp = more ? nxt.xl : 0.0;
    const double q2pk = cur.q2p;
    const double q2p_next = more ? nxt.q2p : io.ld(D.q2p, k + 1, m);
    const double dco = SQ * 0.5 * (xlk + xlp) * sqrt(0.5 * (q2pk + q2p_next));
    const double hk = SPLIT ? (double)io.ld(D.hp, k, m) : rcur.h;
    const double a3 = -dco / (hk * 0.5 * (hk + hkm));
    const double a1 = -dco_prev / (hkm * 0.5 * (hk + hkm));
    const double shp = cur.sp, byp = cur.bp, rq = cur.rq;
    const double sqb = cur.sq / (B_1 * xlk);  // sqrt(q2p) / (B1 l)
    // TKE row
    double a2 = r3 - a1 - a3 + 2. * sqb;
    double ds = (4. * q2pk - io.ld(D.q2pp, k, m)) * itw;
    if (byp >= 0.0) {
      ds += 2. * (shp + byp);
    } else {
      ds += 2. * shp;
      a2 -= byp * 2. * rq;
    }
    // length-scale row (wall function of M&Y)
    zfromt += hkm;
    const double zfromb = hbelow;  // sum of hp(k..kms)
    hbelow -= hk;
    const double r1 = xlk / (KAPPAS * zfromb), r2 = xlk / (KAPPAS * zfromt);
    double a2l = r3 - a1 - a3 + sqb * (1. + E_2 * r1 * r1 + E_3 * r2 * r2);
    double dsl;
    if (D.P.iturb == 1) {
      if (byp >= 0) {
        dsl = xlk * (E_1 * shp + E_1 * byp);
      } else {
        dsl = xlk * (E_1 * shp);
        a2l -= E_1 * byp * rq;
      }
    } else {
      if (byp >= 0) {
        dsl = xlk * (E_1 * shp + 2.00 * byp);
      } else {
        dsl = xlk * (E_1 * shp);
        a2l -= 5.06 * byp * rq;
      }
    }
    dsl += (4. * cur.q2lp - io.ld(D.q2lpp, k, m)) * itw;
    // forward sweeps
    double rd = 1.0 / (a2 + a1 * e1);
    e1 = -a3 * rd;
    f1 = (ds - a1 * f1) * rd;
    rd = 1.0 / (a2l + a1 * e2);
    e2 = -a3 * rd;
    f2 = (dsl - a1 * f2) * rd;
    if (SPLIT) {
      io.st(E1, k, m, (R)e1);
      io.st(E2, k, m, (R)e2);
      io.st(D.q2, k, m, (R)f1);
      io.st(D.q2l, k, m, (R)f2);
    } else {
      io.st(EP, k, m, Pair<R>::make(e1, e2));
      io.st(FP, k, m, Pair<R>::make(f1, f2));
    }
    if (budget) {
      const double rhoxh = ((SPLIT ? (double)io.ld(D.rhop, k, m) : rcur.rho) + 1000.) * hk;
      bs.sh += shp * rhoxh;
      bs.by += byp * rhoxh;
      io.st(D.tkb, k, m, make_float2((float)(sqb * rhoxh), (float)rhoxh));
    }
    // new eddy coefficients of interface k (the surface minimum, 281-284)
    if (!SPLIT) {
      double av = cur.av, dv = cur.dv;
      if (k == k1s + 1 && thin_top) av = dmax((double)(R)av, 1.E-4), dv = dmax((double)(R)dv, 1.E-4);
      io.st(D.Av, k, m, (R)av);
      io.st(D.Dv, k, m, (R)dv);
    }
    dco_prev = dco;
    hkm = hk;
    cur = nxt;
    rcur = rnxt;
  }
  // bottom Dirichlet row kms+1 (a1 = 0) and back substitution
  double x1 = ds_bot / r3, x2 = 0.0;
  io.st(D.q2, kms + 1, m, (R)dmax(Q2_MIN, x1));
  io.st(D.q2l, kms + 1, m, (R)dmax(Q2L_MIN, x2));
  for (int k = kms; k >= k1s; --k) {
    if (SPLIT) {
      x1 = (double)io.ld(E1, k, m) * x1 + (double)io.ld(D.q2, k, m);
      x2 = (double)io.ld(E2, k, m) * x2 + (double)io.ld(D.q2l, k, m);
    } else {
      const R2 e = io.ld(EP, k, m), f = io.ld(FP, k, m);
      x1 = (double)e.x * x1 + (double)f.x;
      x2 = (double)e.y * x2 + (double)f.y;
    }
    const double q2n = (double)(R)dmax(Q2_MIN, x1);
    io.st(D.q2, k, m, (R)q2n);
    io.st(D.q2l, k, m, (R)dmax(Q2L_MIN, x2));
    if (budget && k > k1s) {
      const float2 w = io.ld(D.tkb, k, m);
      bs.di += q2n * (double)w.x;
      bs.tk += q2n * (double)w.y;
    }
  }
  for (int k = D.k1; k <= k1s - 1; ++k) {
    io.st(D.q2, k, m, io.ld(D.q2, k1s, m));
    io.st(D.q2l, k, m, io.ld(D.q2l, k1s, m));
  }
  if (!SPLIT) {
    for (int k = 1; k <= k1s; ++k) io.st(D.Av, k, m, R(0)), io.st(D.Dv, k, m, R(0));
    for (int k = kms + 1; k <= D.km1; ++k) io.st(D.Av, k, m, R(0)), io.st(D.Dv, k, m, R(0));
  }
  return bs;
}

// Sum of v over the workgroup in a fixed order (wavefront shuffles, then the
// wavefronts' sums in LDS added by lane 0 of the workgroup).
__device__ __forceinline__ double wg_sum(double v, double *lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) t += lds[i];
  return t;
}

// One lane per column (every lane reaches the workgroup reduction, so the
// columns without work contribute zeros); with `budget` each workgroup
// writes its four sums to tke_part[4 blockIdx.x ..].
template <typename R, bool SPLIT, bool BUF>
__global__ void __launch_bounds__(kBlock) k_turb(Dev<R> D, int istep, int budget) {
  __shared__ double lds[kBlock / 64];
  const int m = col();
  TkeSums bs{0.0, 0.0, 0.0, 0.0};
  if (m < D.lm) {
    if constexpr (BUF)
      bs = turb_column<R, SPLIT>(D, make_io<BufIO>(D), m, istep, budget != 0);
    else
      bs = turb_column<R, SPLIT>(D, make_io<PtrIO>(D), m, istep, budget != 0);
  }
  if (budget) {
    const double v[4] = {bs.sh, bs.by, bs.di, bs.tk};
    for (int i = 0; i < 4; ++i) {
      const double t = wg_sum(v[i], lds);
      if (threadIdx.x == 0) D.tke_part[4 * blockIdx.x + i] = t;
    }
  }
}

// The budget state after a closure call (si3d_mixing.f90:333-338, 466-475,
// 539-548): a trapezoidal call replaces the four sums by this call's, added
// over the workgroups in a fixed order; a leapfrog call scales them by
// dx dy dt (TKinE by dx dy dt / 2) and adds the first three to the
// cumulative ShearCum, BuocyCum, DisspCum.  One workgroup.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_tke_budget(Dev<R> D, int nblk, int istep) {
  __shared__ double lds[kBlock / 64];
  if (istep > 1) {
    for (int i = 0; i < 4; ++i) {
      double v = 0.0;
      for (int b = threadIdx.x; b < nblk; b += kBlock) v += D.tke_part[4 * b + i];
      const double t = wg_sum(v, lds);
      if (threadIdx.x == 0) D.tke[i] = t;
    }
  } else if (threadIdx.x == 0) {
    const double a = D.P.dx * D.P.dy * D.P.dt;
    for (int i = 0; i < 3; ++i) {
      D.tke[i] *= a;
      D.tke[4 + i] += D.tke[i];
    }
    D.tke[3] = 0.5 * D.tke[3] * a;
  }
}

// ---------------------------------------------------------------------------
// Time-level shuffles.

// 2-D zeta levels: settrap (spp := sp, sp := (s+spp)/2, 3447-3450), settrap2
// (sp := (s+spp)/2, 3951-3962), smooth's Asselin filter (3366-3372).
template <typename R, int MODE>
__global__ void __launch_bounds__(kBlock) k_zeta_levels(Dev<R> D) {
  const int c = col();
  if (c > D.lm) return;
  if (MODE == 0) {  // settrap
    D.spp[c] = D.sp[c];
    if (c < D.lm) D.sp[c] = 0.5 * (D.s[c] + D.spp[c]);
  } else if (MODE == 1) {  // settrap2
    D.sp[c] = 0.5 * (D.s[c] + D.spp[c]);
  } else {  // smooth
    if (c < D.lm) D.sp[c] = D.sp[c] + D.P.beta2 * (D.s[c] - 2. * D.sp[c] + D.spp[c]);
  }
}

// Time-level shuffles as per-cell kernels (every (column, layer) lane is
// independent) followed by small per-column kernels for what needs the whole
// column: the top-layer indices and the corrections at the top cell.

// Layer thickness of a zeta-, u- or v-cell at level k for surface elevation
// `eta` and depth `hh` (layer_hp2 / layer_h, si3d_procedures.f90:2181-2293).
__device__ __forceinline__ double cell_h(const double *zlevel, int k, double hh, double eta) {
  const double v = dmin(zlevel[k + 1], hh) - dmax(zlevel[k], -eta);
  return v;
}

// Pass 1 of the u/v fluxes at n+1/2 (settrap 3488-3531, settrap2 3984-4023)
// for one face cell.  UPD_PP: settrap moves the p level into pp first.
template <typename R, bool X, bool UPD_PP>
__device__ __forceinline__ void half_level_flux_cell(const Dev<R> &D, int c, int k) {
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == D.lm || k < D.k1 || k > min(D.kmz[c], D.kmz[ln])) return;
  R *qhp = X ? D.uhp : D.vhp;
  R *qhpp = X ? D.uhpp : D.vhpp;
  R *qp = X ? D.up : D.vp;
  R *qpp = X ? D.upp : D.vpp;
  const R *qh = X ? D.uh : D.vh;
  const R *hq = X ? D.hup : D.hvp;
  if (UPD_PP) {
    qhpp[IX(k, c)] = qhp[IX(k, c)];
    qpp[IX(k, c)] = qp[IX(k, c)];
  }
  const double hk = hq[IX(k, c)];
  if (hk > 0.0) {
    const double f = 0.5 * ((double)qh[IX(k, c)] + (double)qhpp[IX(k, c)]);
    qhp[IX(k, c)] = (R)f;
    qp[IX(k, c)] = (R)(f / hk);
  } else {
    qhp[IX(k, c)] = R(0);
    qp[IX(k, c)] = R(0);
  }
}

// Top-cell corrections of the n+1/2 fluxes (3532-3543, 4024-4035).
template <typename R, bool X>
__device__ __forceinline__ void half_level_flux_top(const Dev<R> &D, int c, int k) {
  const int ln = X ? D.lE[c] : D.lN[c];
  if (ln == D.lm) return;
  R *qhp = X ? D.uhp : D.vhp;
  R *qp = X ? D.up : D.vp;
  const R *qh = X ? D.uh : D.vh;
  const R *qhpp = X ? D.uhpp : D.vhpp;
  const R *hq = X ? D.hup : D.hvp;
  const R *hqn = X ? D.hu : D.hv;
  const R *hqpp = X ? D.hupp : D.hvpp;
  if ((double)hqn[IX(k - 1, c)] > 0.0) {
    double f = (double)qhp[IX(k, c)] + (double)qh[IX(k - 1, c)] / 2.;
    qhp[IX(k, c)] = (R)f;
    qp[IX(k, c)] = (R)(f / (double)hq[IX(k, c)]);
  }
  if ((double)hqpp[IX(k - 1, c)] > 0.0) {
    double f = (double)qhp[IX(k, c)] + (double)qhpp[IX(k - 1, c)] / 2.;
    qhp[IX(k, c)] = (R)f;
    qp[IX(k, c)] = (R)(f / (double)hq[IX(k, c)]);
  }
}

// TopLayerIndexp2 (2368-2427): first wet layer of the zeta-, u- and v-cells of
// column c from hp / hup / hvp, km1 when there is none.
template <typename R>
__device__ __forceinline__ void top_layer_indices(const Dev<R> &D, int c) {
  const int S = D.lm;
  const int km1 = km1c(D, c);
  int k1z = km1, k1u = km1, k1v = km1;
  for (int k = D.k1; k <= D.kmz[c]; ++k)
    if ((double)D.hp[IX(k, c)] > 0.0) {
      k1z = k;
      break;
    }
  const int lE = D.lE[c], lN = D.lN[c];
  if (lE != S)
    for (int k = D.k1; k <= min(D.kmz[c], D.kmz[lE]); ++k)
      if ((double)D.hup[IX(k, c)] > 0.0) {
        k1u = k;
        break;
      }
  if (lN != S)
    for (int k = D.k1; k <= min(D.kmz[c], D.kmz[lN]); ++k)
      if ((double)D.hvp[IX(k, c)] > 0.0) {
        k1v = k;
        break;
      }
  D.k1z[c] = k1z;
  D.k1u[c] = k1u;
  D.k1v[c] = k1v;
}

// settrap (MODE 0, 3418-3618) / settrap2 (MODE 1, 3944-4087), per cell
// (c = 0..lm, k = 0..km1+1): pp := p for the thicknesses (settrap), layer_hp2's
// thicknesses from the n+1/2 zeta (2238-2293), scalar and density at n+1/2
// (3471-3486), pass 1 of the fluxes, turbulence levels (settrap_2EqTVars).
template <typename R, int MODE, bool CH>
__global__ void __launch_bounds__(kBlock) k_settrap_cell(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm + (MODE == 0 ? 1 : 0), 0, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const double spc = (c < S) ? D.sp[c] : 0.0;
  const int kbot = (c < S) ? D.kmz[c] + 1 : D.km1 + 2;  // below the sediment cell nothing ever changes
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) {
    if (k > kbot) break;
    if (MODE == 0) {
      D.hpp[IX(k, c)] = D.hp[IX(k, c)];
      D.hupp[IX(k, c)] = D.hup[IX(k, c)];
      D.hvpp[IX(k, c)] = D.hvp[IX(k, c)];
    }
    if (c == S) continue;
    const int kmz = D.kmz[c];
    const bool wet = k >= D.k1 && k <= kmz;
    if (wet) {
      const double v = cell_h(zlv(D, c), k, D.hhs[c], spc);
      D.hp[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
    }
    const int lE = D.lE[c], lN = D.lN[c];
    if (lE != S && k >= D.k1 && k <= min(kmz, D.kmz[lE])) {
      const double v = cell_h(zlv(D, c), k, D.hhu[c], dmax(spc, D.sp[lE]));
      D.hup[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
    }
    if (lN != S && k >= D.k1 && k <= min(kmz, D.kmz[lN])) {
      const double v = cell_h(zlv(D, c), k, D.hhv[c], dmax(spc, D.sp[lN]));
      D.hvp[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
    }
    if (wet) {
      if (MODE == 0) D.salpp[IX(k, c)] = D.salp[IX(k, c)];
      const double tm = 0.5 * ((double)D.sal[IX(k, c)] + (double)D.salpp[IX(k, c)]);
      D.salp[IX(k, c)] = (R)tm;
      D.rhop[IX(k, c)] = (R)(densty((double)(R)tm, 0.00004, zdens(D, k, c)) - 1000.);
    }
    half_level_flux_cell<R, true, MODE == 0>(D, c, k);
    half_level_flux_cell<R, false, MODE == 0>(D, c, k);
    if (D.P.iturb > 0 && wet) {
      if (MODE == 0) {
        D.q2pp[IX(k, c)] = D.q2p[IX(k, c)];
        D.q2lpp[IX(k, c)] = D.q2lp[IX(k, c)];
      }
      D.q2p[IX(k, c)] = (R)(0.5 * ((double)D.q2[IX(k, c)] + (double)D.q2pp[IX(k, c)]));
      D.q2lp[IX(k, c)] = (R)(0.5 * ((double)D.q2l[IX(k, c)] + (double)D.q2lpp[IX(k, c)]));
    }
  }
}

// Windowed form of k_settrap_cell (same values, same order): the chunk's
// layers are walked by the whole workgroup together, the fields are read and
// written through buffer resources with one 32-bit offset per column, and the
// column's indices, depths, neighbour elevations and zlevel pointer are loaded
// once per chunk.  The thicknesses written in a layer are reused from
// registers by the fluxes and the density of the same layer.
template <typename R, int MODE>
__global__ void __launch_bounds__(kBlock) k_settrap_cell_w(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm + (MODE == 0 ? 1 : 0), 0, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const bool sent = c == S;
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz;
  const auto hp = rsrc(D.hp, nb), hup = rsrc(D.hup, nb), hvp = rsrc(D.hvp, nb);
  const auto hpp = rsrc(D.hpp, nb), hupp = rsrc(D.hupp, nb), hvpp = rsrc(D.hvpp, nb);
  const int kbot = sent ? D.km1 + 2 : D.kmz[c] + 1;  // below the sediment cell nothing ever changes
  const int kend = min(t.k1, kbot + 1);
  if (sent) {
    if (MODE == 0)
      for (unsigned k = t.k0, pk = k * pb; (int)k < kend; ++k, pk += pb) {
        bst<R>(hpp, (R)bld<R>(hp, oc, pk), oc, pk);
        bst<R>(hupp, (R)bld<R>(hup, oc, pk), oc, pk);
        bst<R>(hvpp, (R)bld<R>(hvp, oc, pk), oc, pk);
      }
    return;
  }
  const int k1 = D.k1, kmz = D.kmz[c], lE = D.lE[c], lN = D.lN[c];
  const int kmu = (lE != S) ? min(kmz, D.kmz[lE]) : -1, kmv = (lN != S) ? min(kmz, D.kmz[lN]) : -1;
  const double spc = D.sp[c], hhs = D.hhs[c];
  const double etau = (lE != S) ? dmax(spc, D.sp[lE]) : 0.0, etav = (lN != S) ? dmax(spc, D.sp[lN]) : 0.0;
  const double hhu = D.hhu[c], hhv = D.hhv[c];
  const double *zl = zlv(D, c);
  const bool turb = D.P.iturb > 0;
  const auto sal = rsrc(D.sal, nb), salp = rsrc(D.salp, nb), salpp = rsrc(D.salpp, nb), rhop = rsrc(D.rhop, nb);
  const auto uh = rsrc(D.uh, nb), uhp = rsrc(D.uhp, nb), uhpp = rsrc(D.uhpp, nb), up = rsrc(D.up, nb),
             upp = rsrc(D.upp, nb);
  const auto vh = rsrc(D.vh, nb), vhp = rsrc(D.vhp, nb), vhpp = rsrc(D.vhpp, nb), vp = rsrc(D.vp, nb),
             vpp = rsrc(D.vpp, nb);
  // pass 1 of the n+1/2 fluxes of one face (half_level_flux_cell), with the
  // face thickness hk of this layer
  auto flux = [&](const auto &qh, const auto &qhp, const auto &qhpp, const auto &qp, const auto &qpp, double hk,
                  unsigned pk) {
    double fpp;  // the flux at pp, after settrap moved p into it
    if (MODE == 0) {
      fpp = bld<R>(qhp, oc, pk);
      bst<R>(qhpp, (R)fpp, oc, pk);
      bst<R>(qpp, (R)bld<R>(qp, oc, pk), oc, pk);
    } else {
      fpp = bld<R>(qhpp, oc, pk);
    }
    if (hk > 0.0) {
      const double f = 0.5 * (bld<R>(qh, oc, pk) + fpp);
      bst<R>(qhp, (R)f, oc, pk);
      bst<R>(qp, (R)(f / hk), oc, pk);
    } else {
      bst<R>(qhp, R(0), oc, pk);
      bst<R>(qp, R(0), oc, pk);
    }
  };
  for (unsigned k = t.k0, pk = k * pb; (int)k < kend; ++k, pk += pb) {
    const int kk = (int)k;
    if (MODE == 0) {
      bst<R>(hpp, (R)bld<R>(hp, oc, pk), oc, pk);
      bst<R>(hupp, (R)bld<R>(hup, oc, pk), oc, pk);
      bst<R>(hvpp, (R)bld<R>(hvp, oc, pk), oc, pk);
    }
    const bool wet = kk >= k1 && kk <= kmz;
    R hpk = R(0);
    if (wet) {
      const double v = cell_h(zl, kk, hhs, spc);
      hpk = (v <= 0.0) ? R(0) : (R)v;
      bst<R>(hp, hpk, oc, pk);
    }
    const bool fu = kk >= k1 && kk <= kmu, fv = kk >= k1 && kk <= kmv;
    double hu = 0.0, hv = 0.0;
    if (fu) {
      const double v = cell_h(zl, kk, hhu, etau);
      const R h = (v <= 0.0) ? R(0) : (R)v;
      bst<R>(hup, h, oc, pk);
      hu = (double)h;
    }
    if (fv) {
      const double v = cell_h(zl, kk, hhv, etav);
      const R h = (v <= 0.0) ? R(0) : (R)v;
      bst<R>(hvp, h, oc, pk);
      hv = (double)h;
    }
    if (wet) {
      double spp;
      if (MODE == 0) {
        spp = bld<R>(salp, oc, pk);
        bst<R>(salpp, (R)spp, oc, pk);
      } else {
        spp = bld<R>(salpp, oc, pk);
      }
      const double tm = 0.5 * (bld<R>(sal, oc, pk) + spp);
      bst<R>(salp, (R)tm, oc, pk);
      const double zk = zl[kk];
      const double zd = (zk == -100.0) ? 0.5 * (double)hpk : zk + 0.5 * (double)hpk;
      bst<R>(rhop, (R)(densty((double)(R)tm, 0.00004, zd) - 1000.), oc, pk);
    }
    if (fu) flux(uh, uhp, uhpp, up, upp, hu, pk);
    if (fv) flux(vh, vhp, vhpp, vp, vpp, hv, pk);
    if (turb && wet) {
      const auto q2 = rsrc(D.q2, nb), q2p = rsrc(D.q2p, nb), q2pp = rsrc(D.q2pp, nb);
      const auto q2l = rsrc(D.q2l, nb), q2lp = rsrc(D.q2lp, nb), q2lpp = rsrc(D.q2lpp, nb);
      double qpp, qlpp;
      if (MODE == 0) {
        qpp = bld<R>(q2p, oc, pk);
        qlpp = bld<R>(q2lp, oc, pk);
        bst<R>(q2pp, (R)qpp, oc, pk);
        bst<R>(q2lpp, (R)qlpp, oc, pk);
      } else {
        qpp = bld<R>(q2pp, oc, pk);
        qlpp = bld<R>(q2lpp, oc, pk);
      }
      bst<R>(q2p, (R)(0.5 * (bld<R>(q2, oc, pk) + qpp)), oc, pk);
      bst<R>(q2lp, (R)(0.5 * (bld<R>(q2l, oc, pk) + qlpp)), oc, pk);
    }
  }
}

// settrap / settrap2 per column, after k_settrap_cell: top-layer indices, the
// top-cell flux corrections and the turbulence fix of a dry top cell.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_settrap_top(Dev<R> D) {
  const int c = col();
  if (c >= D.lm) return;
  top_layer_indices(D, c);
  half_level_flux_top<R, true>(D, c, D.k1u[c]);
  half_level_flux_top<R, false>(D, c, D.k1v[c]);
  if (D.P.iturb > 0) {
    const int k = D.k1z[c];
    if ((double)D.hpp[IX(k, c)] <= 0.0) {
      D.q2lp[IX(k + 1, c)] = D.q2lp[IX(k + 2, c)];
      D.q2p[IX(k + 1, c)] = D.q2p[IX(k + 2, c)];
    }
  }
}

// save(istep), si3d_procedures.f90:3621-3941 (+ save_2EqTvars), per cell
// (c = 0..lm, k = 0..km1+1; the reference loops whole columns, 3678-3703).
template <typename R, int ISTEP, bool CH>
__global__ void __launch_bounds__(kBlock) k_save_cell(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm + 1, 0, nk, tile, t)) return;
  const int c = t.c;
  // Below a column's sediment cell every array holds its initial value and
  // the copies are no-ops (the reference still loops whole columns).
  const int kbot = (c < D.lm) ? D.kmz[c] + 1 : D.km1 + 2;
  for (int k = t.k0; k < (CH ? t.k1 : min(t.k1, t.k0 + 1)); ++k) {
    if (k > kbot) break;
    if (ISTEP == 1) {
      D.hpp[IX(k, c)] = D.hp[IX(k, c)];
      D.hupp[IX(k, c)] = D.hup[IX(k, c)];
      D.hvpp[IX(k, c)] = D.hvp[IX(k, c)];
    }
    D.hp[IX(k, c)] = D.h[IX(k, c)];
    D.hup[IX(k, c)] = D.hu[IX(k, c)];
    D.hvp[IX(k, c)] = D.hv[IX(k, c)];
    if (D.P.iturb > 0) {
      if (ISTEP == 1) {
        D.q2pp[IX(k, c)] = D.q2p[IX(k, c)];
        D.q2lpp[IX(k, c)] = D.q2lp[IX(k, c)];
      }
      D.q2p[IX(k, c)] = D.q2[IX(k, c)];
      D.q2lp[IX(k, c)] = D.q2l[IX(k, c)];
    }
    if (c == D.lm || k < D.k1 || k > km1c(D, c) - 1) continue;
    if (ISTEP == 1) {
      D.salpp[IX(k, c)] = D.salp[IX(k, c)];
      D.uhpp[IX(k, c)] = D.uhp[IX(k, c)];
      D.vhpp[IX(k, c)] = D.vhp[IX(k, c)];
      D.upp[IX(k, c)] = D.up[IX(k, c)];
      D.vpp[IX(k, c)] = D.vp[IX(k, c)];
    }
    const R tm = D.sal[IX(k, c)];
    D.salp[IX(k, c)] = tm;
    D.rhop[IX(k, c)] = (R)(densty((double)tm, 0.00004, zdens(D, k, c)) - 1000.);
    const double hu = D.hup[IX(k, c)];
    if (hu > 0.0) {
      const R f = D.uh[IX(k, c)];
      D.uhp[IX(k, c)] = f;
      D.up[IX(k, c)] = D.u[IX(k, c)] = (R)((double)f / hu);
    } else {
      D.uhp[IX(k, c)] = D.up[IX(k, c)] = D.u[IX(k, c)] = R(0);
    }
    const double hv = D.hvp[IX(k, c)];
    if (hv > 0.0) {
      const R f = D.vh[IX(k, c)];
      D.vhp[IX(k, c)] = f;
      D.vp[IX(k, c)] = D.v[IX(k, c)] = (R)((double)f / hv);
    } else {
      D.vhp[IX(k, c)] = D.vp[IX(k, c)] = D.v[IX(k, c)] = R(0);
    }
  }
}

// Windowed form of k_save_cell (same values, same order; see k_settrap_cell_w).
template <typename R, int ISTEP>
__global__ void __launch_bounds__(kBlock) k_save_cell_w(Dev<R> D, int nk, int tile) {
  CellTile t;
  if (!cell_tile(D.lm + 1, 0, nk, tile, t)) return;
  const int c = t.c, S = D.lm;
  const bool sent = c == S;
  const unsigned sz = sizeof(R), nb = (unsigned)D.KP * (unsigned)D.LP * sz, pb = (unsigned)D.LP * sz;
  const unsigned oc = (unsigned)c * sz;
  // below a column's sediment cell every array holds its initial value
  const int kbot = sent ? D.km1 + 2 : D.kmz[c] + 1;
  const int kend = min(t.k1, kbot + 1), kwet = sent ? -1 : km1c(D, c) - 1, k1 = D.k1;
  const double *zl = sent ? nullptr : zlv(D, c);
  const bool turb = D.P.iturb > 0;
  const auto h = rsrc(D.h, nb), hu = rsrc(D.hu, nb), hv = rsrc(D.hv, nb);
  const auto hp = rsrc(D.hp, nb), hup = rsrc(D.hup, nb), hvp = rsrc(D.hvp, nb);
  const auto hpp = rsrc(D.hpp, nb), hupp = rsrc(D.hupp, nb), hvpp = rsrc(D.hvpp, nb);
  const auto sal = rsrc(D.sal, nb), salp = rsrc(D.salp, nb), salpp = rsrc(D.salpp, nb), rhop = rsrc(D.rhop, nb);
  const auto uh = rsrc(D.uh, nb), uhp = rsrc(D.uhp, nb), uhpp = rsrc(D.uhpp, nb), u = rsrc(D.u, nb),
             up = rsrc(D.up, nb), upp = rsrc(D.upp, nb);
  const auto vh = rsrc(D.vh, nb), vhp = rsrc(D.vhp, nb), vhpp = rsrc(D.vhpp, nb), v = rsrc(D.v, nb),
             vp = rsrc(D.vp, nb), vpp = rsrc(D.vpp, nb);
  // the new face flux and velocity from the transport and the new thickness
  auto face = [&](const auto &qh, const auto &qhp, const auto &q, const auto &qp, double hq, unsigned pk) {
    if (hq > 0.0) {
      const R f = (R)bld<R>(qh, oc, pk);
      bst<R>(qhp, f, oc, pk);
      const R w = (R)((double)f / hq);
      bst<R>(qp, w, oc, pk);
      bst<R>(q, w, oc, pk);
    } else {
      bst<R>(qhp, R(0), oc, pk);
      bst<R>(qp, R(0), oc, pk);
      bst<R>(q, R(0), oc, pk);
    }
  };
  for (unsigned k = t.k0, pk = k * pb; (int)k < kend; ++k, pk += pb) {
    const int kk = (int)k;
    if (ISTEP == 1) {
      bst<R>(hpp, (R)bld<R>(hp, oc, pk), oc, pk);
      bst<R>(hupp, (R)bld<R>(hup, oc, pk), oc, pk);
      bst<R>(hvpp, (R)bld<R>(hvp, oc, pk), oc, pk);
    }
    const R hk = (R)bld<R>(h, oc, pk), huk = (R)bld<R>(hu, oc, pk), hvk = (R)bld<R>(hv, oc, pk);
    bst<R>(hp, hk, oc, pk);
    bst<R>(hup, huk, oc, pk);
    bst<R>(hvp, hvk, oc, pk);
    if (turb) {
      const auto q2 = rsrc(D.q2, nb), q2p = rsrc(D.q2p, nb), q2pp = rsrc(D.q2pp, nb);
      const auto q2l = rsrc(D.q2l, nb), q2lp = rsrc(D.q2lp, nb), q2lpp = rsrc(D.q2lpp, nb);
      if (ISTEP == 1) {
        bst<R>(q2pp, (R)bld<R>(q2p, oc, pk), oc, pk);
        bst<R>(q2lpp, (R)bld<R>(q2lp, oc, pk), oc, pk);
      }
      bst<R>(q2p, (R)bld<R>(q2, oc, pk), oc, pk);
      bst<R>(q2lp, (R)bld<R>(q2l, oc, pk), oc, pk);
    }
    if (sent || kk < k1 || kk > kwet) continue;
    if (ISTEP == 1) {
      bst<R>(salpp, (R)bld<R>(salp, oc, pk), oc, pk);
      bst<R>(uhpp, (R)bld<R>(uhp, oc, pk), oc, pk);
      bst<R>(vhpp, (R)bld<R>(vhp, oc, pk), oc, pk);
      bst<R>(upp, (R)bld<R>(up, oc, pk), oc, pk);
      bst<R>(vpp, (R)bld<R>(vp, oc, pk), oc, pk);
    }
    const R tm = (R)bld<R>(sal, oc, pk);
    bst<R>(salp, tm, oc, pk);
    const double zk = zl[kk];
    const double zd = (zk == -100.0) ? 0.5 * (double)hk : zk + 0.5 * (double)hk;
    bst<R>(rhop, (R)(densty((double)tm, 0.00004, zd) - 1000.), oc, pk);
    face(uh, uhp, u, up, (double)huk, pk);
    face(vh, vhp, v, vp, (double)hvk, pk);
  }
}

// save per column, after k_save_cell: zeta levels, top-layer indices from the
// new hp, and the turbulence fix of a dry top cell.
template <typename R, int ISTEP>
__global__ void __launch_bounds__(kBlock) k_save_top(Dev<R> D) {
  const int c = col();
  if (c > D.lm) return;
  if (ISTEP == 1) D.spp[c] = (c == D.lm) ? D.spp[c] : D.sp[c];
  D.sp[c] = D.s[c];
  if (c == D.lm) return;
  top_layer_indices(D, c);
  if (D.P.iturb > 0) {
    const int k1s = D.k1z[c];
    const bool guard = (ISTEP == 1) ? true : (km1c(D, c) - k1s > 1);
    if (guard && (double)D.hpp[IX(k1s, c)] <= 0.0) {
      D.q2lp[IX(k1s + 1, c)] = D.q2lp[IX(k1s + 2, c)];
      D.q2p[IX(k1s + 1, c)] = D.q2p[IX(k1s + 2, c)];
    }
  }
}

// smooth (Asselin filter on the leapfrog step), si3d_procedures.f90:3356-3415.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_smooth_col(Dev<R> D) {
  const int c = col();
  if (c >= D.lm) return;
  const int S = D.lm;
  const double spc = D.sp[c];
  for (int k = D.k1; k <= D.kmz[c]; ++k) {
    double v = dmin(zlv(D, c)[k + 1], D.hhs[c]) - dmax(zlv(D, c)[k], -spc);
    D.hp[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
  }
  for (int dir = 0; dir < 2; ++dir) {
    const bool X = (dir == 0);
    const int ln = X ? D.lE[c] : D.lN[c];
    if (ln == S) continue;
    R *hq = X ? D.hup : D.hvp;
    const double *hh = X ? D.hhu : D.hhv;
    R *qh = X ? D.uh : D.vh;
    R *qhp = X ? D.uhp : D.vhp;
    R *qhpp = X ? D.uhpp : D.vhpp;
    R *qp = X ? D.up : D.vp;
    const double sm = dmax(spc, D.sp[ln]);
    for (int k = D.k1; k <= min(D.kmz[c], D.kmz[ln]); ++k) {
      double v = dmin(zlv(D, c)[k + 1], hh[c]) - dmax(zlv(D, c)[k], -sm);
      hq[IX(k, c)] = (v <= 0.0) ? R(0) : (R)v;
    }
    for (int k = D.k1; k <= min(D.kmz[c], D.kmz[ln]); ++k) {
      const double hk = hq[IX(k, c)];
      if (hk <= 0.0) continue;
      double f = (double)qhp[IX(k, c)] + D.P.beta2 * ((double)qh[IX(k, c)] - 2. * (double)qhp[IX(k, c)] + (double)qhpp[IX(k, c)]);
      qhp[IX(k, c)] = (R)f;
      qp[IX(k, c)] = (R)(f / hk);
    }
  }
}

// Density of the initial state at every column, sentinel included
// (InitializeScalarFields, si3d_procedures.f90:368-377).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_init_density(Dev<R> D) {
  const int c = col();
  if (c > D.lm) return;
  const int km1 = km1c(D, c);
  for (int k = D.k1; k <= km1; ++k)
    D.rhop[IX(k, c)] = (R)(densty((double)D.salp[IX(k, c)], 0.00004, zdens(D, k, c)) - 1000.);
}

// ---------------------------------------------------------------------------
// Host-side launch sequences.

template <typename R> Dev<R> &dev(s3g_ctx *c);
template <> Dev<double> &dev<double>(s3g_ctx *c) { return c->d64; }
template <> Dev<float> &dev<float>(s3g_ctx *c) { return c->d32; }

static inline dim3 grid_for(int n) { return dim3((n + kBlock - 1) / kBlock); }

#define LAUNCH(name, kern, n, ...)                                          \
  do {                                                                      \
    int _slot = -1;                                                         \
    if (c->timing) timer_begin(c, name, &_slot);                            \
    hipLaunchKernelGGL(kern, grid_for(n), dim3(kBlock), 0, c->stream, __VA_ARGS__); \
    if (c->timing) timer_end(c, _slot);                                     \
  } while (0)

// Per-cell kernel over layers k_first .. k_first+nk-1 in tiles described by the
// tuned word c->kc[tid] (chunk size | order << 8); the kernel takes (..., nk,
// tile) as its last arguments.
#define LAUNCHC(name, kern1, kernc, tid, n, nk, ...)                        \
  do {                                                                      \
    int _slot = -1;                                                         \
    const int _tile = c->kc[tid], _kc = _tile & 255;                        \
    const int _nb = ((n) + kBlock - 1) / kBlock * (((nk) + _kc - 1) / _kc); \
    if (c->timing) timer_begin(c, name, &_slot);                            \
    if (_kc == 1)                                                           \
      hipLaunchKernelGGL(kern1, dim3(_nb), dim3(kBlock), 0, c->stream, __VA_ARGS__, (int)(nk), _tile); \
    else                                                                    \
      hipLaunchKernelGGL(kernc, dim3(_nb), dim3(kBlock), 0, c->stream, __VA_ARGS__, (int)(nk), _tile); \
    if (c->timing) timer_end(c, _slot);                                     \
  } while (0)

// Windowed per-cell kernels (buffer-resource loads, one chunk of layers per
// lane): used unless S3G_WINDOW=0, for fields below 4 GiB; exmom's also needs
// centred momentum advection, ihd 0 or 1 and symmetric neighbours.
static inline bool use_window(const s3g_ctx *c) {
  const size_t field = (size_t)c->KP * c->LP * (c->precision == S3G_FP64 ? 8 : 4);
  return c->window && field < ((size_t)1 << 32);
}
static inline bool exmom_window(const s3g_ctx *c) {
  return use_window(c) && c->nb_symmetric && c->P.itrmom == 1 && c->P.ihd <= 1;
}
#define LAUNCH_EXMOM(name, X)                                                                              \
  do {                                                                                                     \
    if (exmom_window(c))                                                                                   \
      LAUNCHC(name, (k_exmom_w<R, X>), (k_exmom_w<R, X>), (X) ? TILE_EXMOM_X : TILE_EXMOM_Y, L, c->km - 1, \
              D, twodt1);                                                                                  \
    else                                                                                                   \
      LAUNCHC(name, (k_exmom1<R, X>), (k_exmom<R, X, true>), (X) ? TILE_EXMOM_X : TILE_EXMOM_Y, L,         \
              c->km - 1, D, twodt1);                                                                       \
  } while (0)

#define LAUNCH_SLAYER_H(name)                                                                         \
  do {                                                                                                \
    if (use_window(c))                                                                                \
      LAUNCHC(name, (k_slayer_h_w<R>), (k_slayer_h_w<R>), TILE_SLAYER_H, L, c->km1, D);               \
    else                                                                                              \
      LAUNCHC(name, (k_slayer_h<R, false>), (k_slayer_h<R, true>), TILE_SLAYER_H, L, c->km1, D);      \
  } while (0)

#define LAUNCH_VEL2(name)                                                                             \
  do {                                                                                                \
    if (use_window(c))                                                                                \
      LAUNCHC(name, (k_vel2_w<R>), (k_vel2_w<R>), TILE_VEL2, L, c->km - 1, D);                        \
    else                                                                                              \
      LAUNCHC(name, (k_vel2<R, false>), (k_vel2<R, true>), TILE_VEL2, L, c->km - 1, D);               \
  } while (0)

template <typename R>
static void sync_forcing(s3g_ctx *c) {
  Dev<R> &D = dev<R>(c);
  D.uair = c->uair;
  D.vair = c->vair;
  D.cdw = c->cdw;
  if (c->nlakes == 1 && D.force) {  // the kernels read the wind from D.force
    c->lake_force.assign({c->uair, c->vair, c->cdw});
    set_forcing(c);
  }
}

template <typename R>
static Scal<R> sal_of(Dev<R> &D) {
  return Scal<R>{D.salpp, D.sal, D.heat};
}

// Explicit + implicit transport of one scalar (exsal/imsal or exTracer/
// ImTracer): face fluxes, then the right-hand side per cell + column solve,
// or both fused in the column sweep (autotuned).
// With open flow boundaries the boundary fluxes join Bex between the two
// (q: the boundary series of this scalar, 1 active scalar, 2.. tracers).
template <typename R, bool TR>
static void transport(s3g_ctx *c, Scal<R> S, double twodt1, int q) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  const bool tr = TR;
  if (use_window(c))
    LAUNCHC(tr ? "tracer_hflux" : "exsal_hflux", (k_exsal_hflux_w<R>), (k_exsal_hflux_w<R>), TILE_HFLUX, L,
            c->km - 1, D, S, twodt1);
  else
    LAUNCHC(tr ? "tracer_hflux" : "exsal_hflux", (k_exsal_hflux<R, false>), (k_exsal_hflux<R, true>), TILE_HFLUX,
            L, c->km - 1, D, S, twodt1);
  if (c->exsal_split || c->obc.any_flow) {
    if (use_window(c))
      LAUNCHC(tr ? "tracer_rhs" : "exsal_rhs", (k_exsal_rhs_w<R, TR>), (k_exsal_rhs_w<R, TR>), TILE_RHS, L,
              c->km - 1, D, S, twodt1);
    else
      LAUNCHC(tr ? "tracer_rhs" : "exsal_rhs", (k_exsal_rhs<R, false, TR>), (k_exsal_rhs<R, true, TR>), TILE_RHS,
              L, c->km - 1, D, S, twodt1);
    obc_exscalar<R>(c, S.cpp, q);
    LAUNCH(tr ? "imtracer" : "imsal", (k_exsal_imsal<R, false, true, TR>), L, D, S, twodt1);
  } else {
    LAUNCH(tr ? "tracer_imsal" : "exsal_imsal", (k_exsal_imsal<R, true, true, TR>), L, D, S, twodt1);
  }
  obc_scalar<R>(c, S.cn, q);
}

template <typename R>
static void turbulence(s3g_ctx *c, int istep) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  // the basin energy budget is a single lake's (an ensemble keeps none)
  const bool track = c->tke_budget && c->nlakes == 1;
  const int budget = (track && istep > 1) ? 1 : 0;
  if (c->turb_split) {
    LAUNCHC("turb_cell", (k_turb_cell<R, false>), (k_turb_cell<R, true>), TILE_TURB, L, c->km1, D);
    LAUNCH("turbulence", (k_turb<R, true, false>), L, D, istep, budget);
  } else if (use_window(c) && 2 * sizeof(R) * c->KP * (size_t)c->LP < ((size_t)1 << 32)) {
    LAUNCH("turbulence", (k_turb<R, false, true>), L, D, istep, budget);  // factor pairs: 2 planes per level
  } else {
    LAUNCH("turbulence", (k_turb<R, false, false>), L, D, istep, budget);
  }
  if (track) hipLaunchKernelGGL((k_tke_budget<R>), dim3(1), dim3(kBlock), 0, c->stream, D, (int)grid_for(L).x, istep);
}

// surfbc on the host, then HeatSource on the device when the heat budget is
// due (ifsurfbc 1-3 on leapfrog steps and step 1).
template <typename R>
static void surface_forcing(s3g_ctx *c, int n, int istep) {
  surfbc(c, n, istep);
  if (c->heat_due) {
    c->heat_due = false;
    LAUNCH("heat_source", (k_heat_source<R>), c->lm, dev<R>(c), c->heat_force);
  }
  if (c->met_due) {
    c->met_due = false;
    LAUNCH("met_heat", (k_met_heat<R>), c->lm, dev<R>(c), c->met, (const double *)c->d_wmet, c->d_wind);
  }
}

template <typename R>
int phase(s3g_ctx *c, int ph, int n, int istep) {
  const double tz = 1.0 / istep, twodt1 = c->P.twodt * tz;
  if (ph == 0) {
    obc_uvh<R>(c);
    surface_forcing<R>(c, n, istep);
    return 0;
  }
  sync_forcing<R>(c);
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  switch (ph) {
    case 1:
      if (c->P.ihd > 1) LAUNCH("hmix", (k_hmix<R>), L, D);
      break;
    case 2:
      LAUNCH_EXMOM("exmom_x", true);
      obc_exmom<R>(c, true, twodt1);
      break;
    case 3: LAUNCH("matmom_x", (k_matmom<R, true>), L, D, tz); break;
    case 4:
      LAUNCH_EXMOM("exmom_y", false);
      obc_exmom<R>(c, false, twodt1);
      break;
    case 5: LAUNCH("matmom_y", (k_matmom<R, false>), L, D, tz); break;
    case 6:
      LAUNCH("matcon", (k_matcon<R>), L, D, tz);
      obc_qq<R>(c, tz);
      break;
    case 7: {
      int rc = solve_free_surface<R>(c, tz);
      LAUNCH_SLAYER_H("slayer_h");
      LAUNCH("slayer_h_sentinel", (k_slayer_h_sentinel<R>), c->km1 + 2, D);
      obc_uvh<R>(c);
      return rc;
    }
    case 8:
      LAUNCH("vel", (k_vel<R>), L, D, tz);
      LAUNCH("continuity1", (k_continuity<R, 1>), L, D);
      obc_vel<R>(c);
      break;
    case 9:
      if (use_window(c))
        LAUNCHC("exsal_hflux", (k_exsal_hflux_w<R>), (k_exsal_hflux_w<R>), TILE_HFLUX, L, c->km - 1, D, sal_of(D),
                twodt1);
      else
        LAUNCHC("exsal_hflux", (k_exsal_hflux<R, false>), (k_exsal_hflux<R, true>), TILE_HFLUX, L, c->km - 1, D,
                sal_of(D), twodt1);
      LAUNCH("exsal_div", (k_exsal_imsal<R, true, false, false>), L, D, sal_of(D), twodt1);
      obc_exscalar<R>(c, D.salpp, 1);
      break;
    case 10:
      LAUNCH("imsal", (k_exsal_imsal<R, false, true, false>), L, D, sal_of(D), twodt1);
      obc_scalar<R>(c, D.sal, 1);
      break;
    case 11: LAUNCH_VEL2("vel2"); break;
    case 12:
      LAUNCH("smooth_zeta", (k_zeta_levels<R, 2>), L + 1, D);
      LAUNCH("smooth_col", (k_smooth_col<R>), L, D);
      break;
    case 13:
      if (c->P.iturb >= 1) turbulence<R>(c, istep);
      break;
    default: return S3G_ERR_ARG;
  }
  return hip_check(hipGetLastError(), "phase launch");
}

// fd, si3d_procedures.f90:382-575.
// Autotuner: for each per-cell kernel, time every tile (layer-chunk size and
// grid order) on the current state (one untimed launch, then the best of
// three) and keep the fastest; then choose between the split (per-cell +
// column sweep) and fused forms of exsal/imsal and of the turbulence closure.
// The outputs of every timed launch are saved before and restored after, so
// tuning leaves the state exactly as it found it.
template <typename R>
int autotune(s3g_ctx *c) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  const double tz = 1.0, twodt1 = c->P.twodt;
  const size_t n3 = (size_t)c->KP * c->LP;
  const int cand[] = {1, 2, 4, 8, 16, 32, 256 + 1, 256 + 4, 512 + 1, 512 + 2, 512 + 4, 768 + 1, 768 + 2, 768 + 4};
  const bool timing = c->timing;
  c->timing = false;
  char *save = nullptr;
  // the largest set of outputs saved at once: five state arrays, or q2, q2l,
  // Av, Dv
  const size_t save_bytes = std::max(5 * sizeof(R), 2 * sizeof(double) + 2 * sizeof(R)) * n3;
  if (hipMalloc(&save, save_bytes) != hipSuccess) {  // no room: keep the default tiles
    (void)hipGetLastError();
    c->timing = timing;
    if (c->debug) std::fprintf(stderr, "[s3g] autotune skipped: no device memory for the output snapshot\n");
    return 0;
  }
  // Each timed run starts from cold caches, as in a step where the inputs were
  // written by earlier kernels (timing back-to-back repeats would let the 256 MB
  // Infinity Cache hold the inputs and favour candidates that re-read them).
  void *scrub = nullptr;
  const size_t scrub_bytes = (size_t)512 << 20;
  if (hipMalloc(&scrub, scrub_bytes) != hipSuccess) scrub = nullptr;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  struct Out {
    void *p;
    size_t bytes;
  };
  const size_t bR = n3 * sizeof(R), bD = n3 * sizeof(double);
  auto snapshot = [&](const std::vector<Out> &outs, bool restore) {
    size_t o = 0;
    for (const Out &x : outs) {
      if (restore)
        hipMemcpyAsync(x.p, save + o, x.bytes, hipMemcpyDeviceToDevice, c->stream);
      else
        hipMemcpyAsync(save + o, x.p, x.bytes, hipMemcpyDeviceToDevice, c->stream);
      o += x.bytes;
    }
  };
  auto time_it = [&](auto &&launch, int reps) {
    float best = 1e30f;
    launch();
    for (int rep = 0; rep < reps; ++rep) {
      if (scrub) hipMemsetAsync(scrub, rep, scrub_bytes, c->stream);
      hipEventRecord(e0, c->stream);
      launch();
      hipEventRecord(e1, c->stream);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      best = std::min(best, ms);
    }
    return best;
  };
  for (int tid = 0; tid < TILE_TUNED; ++tid) {
    if ((tid == TILE_HFLUX || tid == TILE_RHS) && c->P.isal == 0) continue;
    if (tid == TILE_TURB && c->P.iturb < 1) continue;
    if (tid == TILE_VEL) continue;  // vel is a column kernel (it back-substitutes matmom's systems)
    std::vector<Out> outs;
    switch (tid) {
      case TILE_EXMOM_X: outs = {{D.ex, bR}}; break;
      case TILE_EXMOM_Y: outs = {{D.ey, bR}}; break;
      case TILE_SLAYER_H: outs = {{D.h, bR}, {D.hu, bR}, {D.hv, bR}}; break;
      case TILE_HFLUX: outs = {{D.fx_ew, bR}, {D.fx_ns, bR}}; break;
      case TILE_RHS: outs = {{D.bex, bR}}; break;
      case TILE_TURB: outs = {{D.Av, bR}, {D.Dv, bR}}; break;  // d0..d2 are scratch
      default: outs = {{D.h, bR}, {D.hu, bR}, {D.hv, bR}, {D.uh, bR}, {D.vh, bR}}; break;
    }
    snapshot(outs, false);
    auto launch = [&]() {
      switch (tid) {
        case TILE_EXMOM_X: LAUNCH_EXMOM("", true); break;
        case TILE_EXMOM_Y: LAUNCH_EXMOM("", false); break;
        case TILE_SLAYER_H: LAUNCH_SLAYER_H(""); break;
        case TILE_HFLUX:
          if (use_window(c))
            LAUNCHC("", (k_exsal_hflux_w<R>), (k_exsal_hflux_w<R>), TILE_HFLUX, L, c->km - 1, D, sal_of(D), twodt1);
          else
            LAUNCHC("", (k_exsal_hflux<R, false>), (k_exsal_hflux<R, true>), TILE_HFLUX, L, c->km - 1, D, sal_of(D),
                    twodt1);
          break;
        case TILE_RHS:
          if (use_window(c))
            LAUNCHC("", (k_exsal_rhs_w<R, false>), (k_exsal_rhs_w<R, false>), TILE_RHS, L, c->km - 1, D, sal_of(D),
                    twodt1);
          else
            LAUNCHC("", (k_exsal_rhs<R, false, false>), (k_exsal_rhs<R, true, false>), TILE_RHS, L, c->km - 1, D,
                    sal_of(D), twodt1);
          break;
        case TILE_TURB: LAUNCHC("", (k_turb_cell<R, false>), (k_turb_cell<R, true>), TILE_TURB, L, c->km1, D); break;
        default: LAUNCH_VEL2(""); break;
      }
    };
    int best_kc = c->kc[tid];
    float best = 1e30f;
    for (int kc : cand) {
      c->kc[tid] = kc;
      const float ms = time_it(launch, 3);
      if (ms < best) best = ms, best_kc = kc;
    }
    c->kc[tid] = best_kc;
    snapshot(outs, true);
    if (c->debug) std::fprintf(stderr, "[s3g] tile %d: kc=%d order=%d (%.3f ms)\n", tid, best_kc & 255, best_kc >> 8, best);
  }
  // The split or fused forms of exsal/imsal and of the turbulence closure are
  // not chosen by timing: the split exsal stores Bex in the state precision
  // (FP32 in mixed mode) where the fused one keeps it in registers, so the two
  // round differently, and results must not depend on which was faster on a
  // given run.  They are fixed per context (split exsal, fused closure, the
  // faster pair on every lake measured) and can be set with S3G_EXSAL /
  // S3G_TURB = split | fused.
  const int e = hip_check(hipStreamSynchronize(c->stream), "autotune");
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(save);
  if (scrub) hipFree(scrub);
  c->timing = timing;
  return e;
}


// Segments of a step whose kernel sequence and arguments are fixed by
// (istep, lastiter): captured into a HIP graph on first use, replayed after.
enum { SEG_FD_PRE = 0, SEG_FD_POST = 1, SEG_SETTRAP = 2, SEG_SETTRAP2 = 3, SEG_SAVE = 4 };

template <typename F>
static int run_segment(s3g_ctx *c, int seg, int istep, int lastiter, F &&launch) {
  if (!c->step_graph || c->timing || c->obc.n > 0) {
    launch();
    return 0;
  }
  hipGraphExec_t &ex = c->seg_exec[seg][istep == 1 ? 0 : 1][lastiter ? 1 : 0];
  if (!ex) {
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      launch();
      ok = hipStreamEndCapture(c->stream, &g) == hipSuccess && g;
    }
    if (ok) ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) hipGraphDestroy(g);
    if (!ok) {  // launch kernel by kernel from now on
      (void)hipGetLastError();
      ex = nullptr;
      c->step_graph = false;
      if (c->debug) std::fprintf(stderr, "[s3g] step graph capture failed; launching kernels one by one\n");
      launch();
      return 0;
    }
  }
  ++c->seg_replays;
  return hip_check(hipGraphLaunch(ex, c->stream), "step graph launch");
}

template <typename R>
int fd(s3g_ctx *c, int n, int istep, int lastiter) {
  if (!c->tuned) {
    c->tuned = true;
    if (c->tune) {
      const int e = autotune<R>(c);
      if (e < 0) return e;
    }
  }
  const double tz = 1.0 / istep, twodt1 = c->P.twodt * tz;
  obc_uvh<R>(c);  // openbcUVH (423)
  surface_forcing<R>(c, n, istep);
  sync_forcing<R>(c);
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  int e = run_segment(c, SEG_FD_PRE, istep, 0, [&] {
    if (c->P.ihd > 1) LAUNCH("hmix", (k_hmix<R>), L, D);
    LAUNCH_EXMOM("exmom_x", true);
    obc_exmom<R>(c, true, twodt1);
    LAUNCH("matmom_x", (k_matmom<R, true>), L, D, tz);
    LAUNCH_EXMOM("exmom_y", false);
    obc_exmom<R>(c, false, twodt1);
    LAUNCH("matmom_y", (k_matmom<R, false>), L, D, tz);
    LAUNCH("matcon", (k_matcon<R>), L, D, tz);
    obc_qq<R>(c, tz);
  });
  if (e < 0) return e;
  int rc = solve_free_surface<R>(c, tz);
  if (rc < 0) return rc;
  e = run_segment(c, SEG_FD_POST, istep, lastiter, [&] {
    LAUNCH_SLAYER_H("slayer_h");
    LAUNCH("slayer_h_sentinel", (k_slayer_h_sentinel<R>), c->km1 + 2, D);
    obc_uvh<R>(c);  // openbcUVH after the solve (482)
    LAUNCH("vel", (k_vel<R>), L, D, tz);
    LAUNCH("continuity1", (k_continuity<R, 1>), L, D);
    obc_vel<R>(c);
    if (c->P.isal != 0) {
      transport<R, false>(c, sal_of(D), twodt1, 1);
    }
    // non-active scalars on the last iteration only (521-547); sources and
    // sinks are zero (srcsnk00)
    if (c->ntr > 0 && c->P.niter > 0 && lastiter == 1) {
      const size_t n3 = (size_t)c->KP * c->LP;
      for (int nt = 0; nt < c->ntr; ++nt)
        transport<R, true>(c, Scal<R>{D.trpp + nt * n3, D.tr + nt * n3, nullptr}, twodt1, 2 + nt);
    }
    LAUNCH_VEL2("vel2");
    if (c->P.ismooth >= 1 && istep == 1) {
      LAUNCH("smooth_zeta", (k_zeta_levels<R, 2>), L + 1, D);
      LAUNCH("smooth_col", (k_smooth_col<R>), L, D);
    }
    if (c->P.iturb >= 1) turbulence<R>(c, istep);
  });
  if (e < 0) return e;
  e = hip_check(hipGetLastError(), "fd launch");
  return e < 0 ? e : rc;
}

template <typename R>
int settrap(s3g_ctx *c) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  const int e = run_segment(c, SEG_SETTRAP, 2, 0, [&] {
    LAUNCH("settrap_zeta", (k_zeta_levels<R, 0>), L + 1, D);
    if (use_window(c))
      LAUNCHC("settrap_cell", (k_settrap_cell_w<R, 0>), (k_settrap_cell_w<R, 0>), TILE_SETTRAP, L + 1, c->km1 + 2, D);
    else
      LAUNCHC("settrap_cell", (k_settrap_cell<R, 0, false>), (k_settrap_cell<R, 0, true>), TILE_SETTRAP, L + 1,
              c->km1 + 2, D);
    LAUNCH("settrap_top", (k_settrap_top<R>), L, D);
    LAUNCH("continuity2", (k_continuity<R, 2>), L, D);
    obc_vel<R>(c);
  });
  if (e < 0) return e;
  obc_shift(c, 0);
  return hip_check(hipGetLastError(), "settrap launch");
}

template <typename R>
int settrap2(s3g_ctx *c) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  const int e = run_segment(c, SEG_SETTRAP2, 2, 0, [&] {
    LAUNCH("settrap2_zeta", (k_zeta_levels<R, 1>), L + 1, D);
    if (use_window(c))
      LAUNCHC("settrap2_cell", (k_settrap_cell_w<R, 1>), (k_settrap_cell_w<R, 1>), TILE_SETTRAP, L, c->km1 + 2, D);
    else
      LAUNCHC("settrap2_cell", (k_settrap_cell<R, 1, false>), (k_settrap_cell<R, 1, true>), TILE_SETTRAP, L,
              c->km1 + 2, D);
    LAUNCH("settrap_top", (k_settrap_top<R>), L, D);
    LAUNCH("continuity2", (k_continuity<R, 2>), L, D);
    obc_vel<R>(c);
  });
  if (e < 0) return e;
  return hip_check(hipGetLastError(), "settrap2 launch");
}

template <typename R>
int save(s3g_ctx *c, int istep) {
  Dev<R> &D = dev<R>(c);
  const int L = c->lm;
  const int e = run_segment(c, SEG_SAVE, istep, 0, [&] {
    if (istep == 1) {
      if (use_window(c))
        LAUNCHC("save_cell", (k_save_cell_w<R, 1>), (k_save_cell_w<R, 1>), TILE_SAVE, L + 1, c->km1 + 2, D);
      else
        LAUNCHC("save_cell", (k_save_cell<R, 1, false>), (k_save_cell<R, 1, true>), TILE_SAVE, L + 1, c->km1 + 2, D);
      LAUNCH("save_top", (k_save_top<R, 1>), L + 1, D);
    } else {
      if (use_window(c))
        LAUNCHC("save_cell", (k_save_cell_w<R, 2>), (k_save_cell_w<R, 2>), TILE_SAVE, L + 1, c->km1 + 2, D);
      else
        LAUNCHC("save_cell", (k_save_cell<R, 2, false>), (k_save_cell<R, 2, true>), TILE_SAVE, L + 1, c->km1 + 2, D);
      LAUNCH("save_top", (k_save_top<R, 2>), L + 1, D);
    }
    LAUNCH("continuity2", (k_continuity<R, 2>), L, D);
    obc_vel<R>(c);
    if (c->ntr > 0)  // tracerpp = tracer (3772-3774, 3927-3929)
      hipMemcpyAsync(D.trpp, D.tr, (size_t)c->ntr * c->KP * c->LP * sizeof(R), hipMemcpyDeviceToDevice,
                     c->stream);
  });
  if (e < 0) return e;
  obc_shift(c, istep == 1 ? 1 : 2);
  return hip_check(hipGetLastError(), "save launch");
}

// Per-lake forcing of an ensemble (uair, vair, cdw per lake), passed by value
// in chunks of 64 lakes.
struct ForceChunk {
  double v[3 * 64];
};
__global__ void k_set_forcing(ForceChunk f, int first, int n, double *force) {
  const int i = threadIdx.x;
  if (i < 3 * n) force[3 * first + i] = f.v[i];
}

int set_forcing(s3g_ctx *c) {
  for (int first = 0; first < c->nlakes; first += 64) {
    const int n = std::min(64, c->nlakes - first);
    ForceChunk f{};
    for (int i = 0; i < 3 * n; ++i) f.v[i] = c->lake_force[3 * first + i];
    hipLaunchKernelGGL(k_set_forcing, dim3(1), dim3(192), 0, c->stream, f, first, n, c->d_force);
  }
  return hip_check(hipGetLastError(), "set_forcing");
}

int init_density(s3g_ctx *c) {
  if (c->precision == S3G_FP64)
    hipLaunchKernelGGL(k_init_density<double>, grid_for(c->lm + 1), dim3(kBlock), 0, c->stream, c->d64);
  else
    hipLaunchKernelGGL(k_init_density<float>, grid_for(c->lm + 1), dim3(kBlock), 0, c->stream, c->d32);
  return hip_check(hipStreamSynchronize(c->stream), "init_density");
}

// agx/arx (x) or agy/ary (y) as the reference holds them after matmom: the
// per-layer solution, back-substituted from the stored factors into ag / ar
// (device buffers of the field size, zeroed by the caller).
template <typename R>
int momentum_solution(s3g_ctx *c, bool x, void *ag, void *ar) {
  Dev<R> &D = dev<R>(c);
  if (x)
    hipLaunchKernelGGL((k_momentum_solution<R, true>), grid_for(c->lm), dim3(kBlock), 0, c->stream, D, (R *)ag,
                       (R *)ar);
  else
    hipLaunchKernelGGL((k_momentum_solution<R, false>), grid_for(c->lm), dim3(kBlock), 0, c->stream, D, (R *)ag,
                       (R *)ar);
  return hip_check(hipGetLastError(), "momentum solution");
}
template int momentum_solution<double>(s3g_ctx *, bool, void *, void *);
template int momentum_solution<float>(s3g_ctx *, bool, void *, void *);

template int fd<double>(s3g_ctx *, int, int, int);
template int fd<float>(s3g_ctx *, int, int, int);
template int phase<double>(s3g_ctx *, int, int, int);
template int phase<float>(s3g_ctx *, int, int, int);
template int settrap<double>(s3g_ctx *);
template int settrap<float>(s3g_ctx *);
template int settrap2<double>(s3g_ctx *);
template int settrap2<float>(s3g_ctx *);
template int save<double>(s3g_ctx *, int);
template int save<float>(s3g_ctx *, int);

}  // namespace s3g
