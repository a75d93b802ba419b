// s3g_openbc.hip -- open boundaries of types 1 (water surface elevation) and
// 2 (free-surface flow) on the four sides, si3d_boundaryconditions.f90.
//
// A boundary is a run of wet edge cells (a column for W/E, a row for N/S).
// Every step touches only those cells: one launch per boundary, in input
// order (the reference's single-thread order, which matters where two ranges
// overlap), over its cells or cells x layers.  The boundary transports and
// thicknesses keep the reference's (k, j) / (k, i) rows at three time levels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "s3g_internal.h"

namespace s3g {

#define IX(k, c) ((size_t)(k) * (size_t)D.LP + (size_t)(c))

constexpr int kWave = 64;  // CDNA wavefront

__device__ __forceinline__ double &BQ(const ObcArgs &A, int sd, int lev, int k, int a) {
  return A.bq[(size_t)(sd * 3 + lev) * A.SZ + (size_t)a * A.KB + k];
}
__device__ __forceinline__ double &BH(const ObcArgs &A, int sd, int lev, int k, int a) {
  return A.bh[(size_t)(sd * 3 + lev) * A.SZ + (size_t)a * A.KB + k];
}

// Each kernel handles one boundary nn over a grid of lanes (one per cell, or
// per cell and layer); the host launches the boundaries in input order, so
// where two boundaries touch the same cells, qq or rows, the later one sees
// the earlier one's result as in the reference's loop.

// Sum over the workgroup (kBlock lanes); lane 0 gets the result.
__device__ double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  return red[0];
}

// Water level of flow boundary nn's section: the column one past the
// section's end (the Fortran loop index after the loop, 985/1021/1080/1137)
// against the deepest cell of the section.
template <typename R>
__device__ __forceinline__ double flow_level(const Dev<R> &D, const ObcArgs &A, int nn) {
  const int sd = A.side[nn], last = A.cells[A.first[nn] + A.count[nn] - 1];
  const int lpast = (sd == 0 || sd == 2) ? D.lN[last] : D.lE[last];
  return fmax(D.s[lpast], -A.hbc[nn] + A.dzmin);
}

// openbcUVH (si3d_boundaryconditions.f90:900-1398) and the t = 0 values of
// openbc0 (23-897, first = 1), type 1: zeta (at t = 0: sp) of the cells.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_level(Dev<R> D, ObcArgs A, int nn, int first) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < A.count[nn]) (first ? D.sp : D.s)[A.cells[A.first[nn] + i]] = A.val[nn];
}

// Type 2, part 1: the wet thickness of every layer of the section below its
// water level (rows cleared first; the north/south thickness only at t = 0,
// 1089/1146), and per-workgroup partial sums of the area (dy on every side,
// 990-1170; the south side sums its own thicknesses where the reference sums
// the north side's, 1170).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_section(Dev<R> D, ObcArgs A, int nn, int first) {
  __shared__ double red[kBlock];
  const int e = blockIdx.x * kBlock + threadIdx.x, K = D.km1, sd = A.side[nn];
  const bool we = (sd == 0 || sd == 2);
  double area = 0.0;
  if (e < A.count[nn] * K) {
    const int i = e / K, k = 1 + e % K, c = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i];
    BQ(A, sd, 0, k, a) = 0.0;
    if (k >= D.k1 && k <= D.kmz[c]) {
      double hh = fmin(D.zlevel[k + 1], D.hhs[c]) - fmax(D.zlevel[k], -flow_level(D, A, nn));
      if (hh <= 0.0) hh = 0.0;
      BH(A, sd, 0, k, a) = hh;
      area = hh * D.P.dy;
    } else if (we || first) {
      BH(A, sd, 0, k, a) = 0.0;
    }
  }
  const double tot = block_sum(area, red);
  if (threadIdx.x == 0) A.part[blockIdx.x] = tot;
}

// Type 2, part 2: the flow spread uniformly over the section's area (every
// workgroup adds up the partial sums itself).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_flow(Dev<R> D, ObcArgs A, int nn, int first, int nparts) {
  __shared__ double red[kBlock];
  double v = 0.0;
  for (int j = threadIdx.x; j < nparts; j += kBlock) v += A.part[j];
  const double areatot = block_sum(v, red);
  const int e = blockIdx.x * kBlock + threadIdx.x, K = D.km1, sd = A.side[nn];
  if (e >= A.count[nn] * K) return;
  const int i = e / K, k = 1 + e % K, c = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i];
  const int kf = (first && sd == 1) ? D.k1z[c] : D.k1;  // openbc0 north from k1z (600)
  if (k >= kf && k <= D.kmz[c]) BQ(A, sd, 0, k, a) = A.val[nn] / areatot * BH(A, sd, 0, k, a);
}

// MODqqddrr4openBC, si3d_boundaryconditions.f90:1508-1690.  Type 1 (one lane
// per cell) moves the coupling to the known zeta into qq of the inner
// neighbour and drops it from the matrix; type 2 (one wavefront per cell,
// lanes over the layers) adds the boundary inflow at n+1 and n-1 to qq.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_qq(Dev<R> D, ObcArgs A, int nn, double tz) {
  const int S = D.lm, sd = A.side[nn];
  if (A.type[nn] == 1) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.count[nn]) return;
    const int c = A.cells[A.first[nn] + i];
    const int li = sd == 0 ? D.lE[c] : sd == 1 ? D.lS[c] : sd == 2 ? D.lW[c] : D.lN[c];
    double *cp = sd == 0 ? &D.sx[c] : sd == 1 ? &D.sy[li] : sd == 2 ? &D.sx[li] : &D.sy[c];
    if (li != S) D.qq[li] = D.qq[li] + *cp * D.s[c];
    if (sd == 0 || sd == 3 || li != S) *cp = 0.0;
    return;
  }
  const int i = (blockIdx.x * kBlock + threadIdx.x) / kWave, lane = threadIdx.x % kWave;
  if (i >= A.count[nn]) return;  // whole wavefronts leave together
  const int c = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i];
  double s0 = 0.0, s2 = 0.0;
  for (int k = D.k1 + lane; k <= D.kmz[c]; k += kWave) {
    s0 += BQ(A, sd, 0, k, a);
    s2 += BQ(A, sd, 2, k, a);
  }
  for (int m = kWave / 2; m > 0; m >>= 1) {
    s0 += __shfl_xor(s0, m, kWave);
    s2 += __shfl_xor(s2, m, kWave);
  }
  if (lane != 0) return;
  const double dtdx1 = D.P.dtdx * tz, dtdy1 = D.P.dtdy * tz;
  double q = D.qq[c];
  switch (sd) {
    case 0: q = q + dtdx1 * s0 + dtdx1 * s2; break;
    case 1: q = q - dtdy1 * s0 - dtdy1 * s2; break;
    case 2: q = q - dtdx1 * s0 - dtdx1 * s2; break;
    default: q = q + dtdy1 * s0 + dtdy1 * s2; break;
  }
  D.qq[c] = q;
}

// MODexmom4openBCX / Y, si3d_boundaryconditions.f90:1801-2936, one lane per
// cell and layer: the explicit term of the first momentum face inside the
// boundary from upwind advection only, the boundary transport (type 2) being
// the outer neighbour.  The face is the boundary cell's own (W, S) or its
// inner neighbour's (E, N); the east water-level face keeps ex = 0 (its
// assignment is commented out, 2070).
template <typename R, bool X>
__global__ void __launch_bounds__(kBlock) k_obc_exmom(Dev<R> D, ObcArgs A, int nn, double twodt1) {
  const int e = blockIdx.x * kBlock + threadIdx.x, S = D.lm, K = D.km1;
  if (e >= A.count[nn] * K) return;
  const int sd = A.side[nn], ty = A.type[nn];
  const int i = e / K, k = 1 + e % K, cb = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i];
  const int l = (sd == 0 || sd == 3) ? cb : (sd == 2 ? D.lW[cb] : D.lS[cb]);
  if (l == S) return;
  if (ty == 2 && (sd == 0 ? D.lE[l] == S : sd == 3 ? D.lN[l] == S : false)) return;
  const int lE = D.lE[l], lW = D.lW[l], lN = D.lN[l], lS = D.lS[l];
  const int kmq = X ? min(D.kmz[lE], D.kmz[l]) : min(D.kmz[lN], D.kmz[l]);
  const int k1q = X ? D.k1u[l] : D.k1v[l];
  if (kmq - k1q + 1 < 1) return;
  R *ex = X ? D.ex : D.ey;
  if (k < k1q || k > kmq || (ty == 1 && sd == 2)) {
    ex[IX(k, l)] = R(0);
    return;
  }
  auto P = [](double v) { return v + fabs(v); };
  auto M = [](double v) { return v - fabs(v); };
  auto g = [&](const R *f3, int kk, int cc) { return (double)f3[IX(kk, cc)]; };
  double bq0 = 0.0, bdry = 0.0;
  if (ty == 2) {
    bq0 = (BQ(A, sd, 0, k, a) + BQ(A, sd, 2, k, a)) / 2.;
    bdry = (BH(A, sd, 2, k, a) > 0.0) ? BQ(A, sd, 2, k, a) / BH(A, sd, 2, k, a) : 0.0;
  }
  double adv;
  if (X) {
    const double vN = g(D.vhp, k, lE) + g(D.vhp, k, l), vS = g(D.vhp, k, D.lS[lE]) + g(D.vhp, k, lS);
    const double wU = g(D.wp, k, lE) + g(D.wp, k, l), wD = g(D.wp, k + 1, lE) + g(D.wp, k + 1, l);
    double hx;
    if (sd == 0) {
      const double uE = g(D.uhp, k, lE) + g(D.uhp, k, l);
      const double uW = g(D.uhp, k, l) + (ty == 2 ? bq0 : g(D.uhp, k, l));
      const double uout = (ty == 2) ? bdry : g(D.upp, k, l);
      hx = P(uE) * g(D.upp, k, l) + M(uE) * g(D.upp, k, lE) - P(uW) * uout - M(uW) * g(D.upp, k, l);
    } else {
      const double uE = g(D.uhp, k, l) + bq0;
      const double uW = g(D.uhp, k, lW) + g(D.uhp, k, l);
      hx = P(uE) * g(D.upp, k, l) + M(uE) * bdry - P(uW) * g(D.upp, k, lW) - M(uW) * g(D.upp, k, l);
    }
    adv = hx / D.P.fourdx + (P(vN) * g(D.upp, k, l) + M(vN) * g(D.upp, k, lN) - P(vS) * g(D.upp, k, lS) -
                             M(vS) * g(D.upp, k, l)) / D.P.fourdy;
    adv = adv + (P(wU) * g(D.upp, k, l) + M(wU) * g(D.upp, k - 1, l)) / 4. -
          (P(wD) * g(D.upp, k + 1, l) + M(wD) * g(D.upp, k, l)) / 4.;
    ex[IX(k, l)] = (R)(g(D.uhpp, k, l) - twodt1 * (adv * D.P.iadv));
  } else {
    const double uE = g(D.uhp, k, lN) + g(D.uhp, k, l), uW = g(D.uhp, k, D.lW[lN]) + g(D.uhp, k, lW);
    const double wU = g(D.wp, k, lN) + g(D.wp, k, l), wD = g(D.wp, k + 1, lN) + g(D.wp, k + 1, l);
    double hy;
    if (sd == 1) {
      const double vN = g(D.vhp, k, l) + (ty == 2 ? bq0 : g(D.vhp, k, l));
      const double vS = g(D.vhp, k, l) + g(D.vhp, k, lS);
      const double vout = (ty == 2) ? bdry : g(D.vpp, k, l);
      hy = P(vN) * g(D.vpp, k, l) + M(vN) * vout - P(vS) * g(D.vpp, k, lS) - M(vS) * g(D.vpp, k, l);
    } else {
      const double vN = g(D.vhp, k, lN) + g(D.vhp, k, l);
      const double vS = g(D.vhp, k, l) + (ty == 2 ? bq0 : g(D.vhp, k, l));
      const double vout = (ty == 2) ? bdry : g(D.vpp, k, l);
      hy = P(vN) * g(D.vpp, k, l) + M(vN) * g(D.vpp, k, lN) - P(vS) * vout - M(vS) * g(D.vpp, k, l);
    }
    adv = (P(uE) * g(D.vpp, k, l) + M(uE) * g(D.vpp, k, lE) - P(uW) * g(D.vpp, k, lW) - M(uW) * g(D.vpp, k, l)) /
              D.P.fourdx + hy / D.P.fourdy;
    adv = adv + (P(wU) * g(D.vpp, k, l) + M(wU) * g(D.vpp, k - 1, l)) / 4. -
          (P(wD) * g(D.vpp, k + 1, l) + M(wD) * g(D.vpp, k, l)) / 4.;
    ex[IX(k, l)] = (R)(g(D.vhpp, k, l) - twodt1 * (adv * D.P.iadv));
  }
}

// MODvel4openBC, si3d_boundaryconditions.f90:2941-3154, after every
// continuity call.  Type 1 (one lane per cell and layer) clears the
// tangential transport of the boundary cells and copies w from the inner
// neighbour; type 2 (one wavefront per cell) adds the boundary inflow to w
// from the bottom up: w(k) += the inflow of layers k..kmz, a suffix sum over
// the lanes carried from one 64-layer chunk to the next.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_vel(Dev<R> D, ObcArgs A, int nn) {
  const int K = D.km1, sd = A.side[nn];
  const bool we = (sd == 0 || sd == 2);
  if (A.type[nn] == 1) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= A.count[nn] * K) return;
    const int c = A.cells[A.first[nn] + e / K], k = 1 + e % K;
    const int ln = sd == 0 ? D.lE[c] : sd == 1 ? D.lS[c] : sd == 2 ? D.lW[c] : D.lN[c];
    (we ? D.vh : D.uh)[IX(k, c)] = R(0);
    D.wp[IX(k, c)] = D.wp[IX(k, ln)];
    return;
  }
  const int i = (blockIdx.x * kBlock + threadIdx.x) / kWave, lane = threadIdx.x % kWave;
  if (i >= A.count[nn]) return;  // whole wavefronts leave together
  const int c = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i], kms = D.kmz[c];
  const double den = we ? D.P.twodx : D.P.twody;
  const double sgn = (sd == 0 || sd == 3) ? 1.0 : -1.0;
  if (lane == 0) D.wp[IX(kms + 1, c)] = R(0);
  double carry = 0.0;
  for (int top = kms; top >= D.k1; top -= kWave) {
    const int k = top - lane;  // lane 0 holds the deepest layer of the chunk
    double f = 0.0;
    if (k >= D.k1) f = sgn * (BQ(A, sd, 0, k, a) + BQ(A, sd, 2, k, a)) / den;
    for (int m = 1; m < kWave; m <<= 1) {  // inclusive scan from the deepest layer up
      const double o = __shfl_up(f, m, kWave);
      if (lane >= m) f += o;
    }
    const double dru = carry + f;
    if (k >= D.k1) D.wp[IX(k, c)] = (R)((double)D.wp[IX(k, c)] + dru);
    carry = __shfl(dru, kWave - 1, kWave);
  }
}

// MODexsal4openbc / MODexTracer4openbc, si3d_boundaryconditions.f90:3314-3658,
// 3809-4137 (type 2): the advective flux through the boundary face into Bex,
// the boundary value on inflow, the cell's own value (n-1) on outflow.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_exscalar(Dev<R> D, ObcArgs A, int nn, const R *cpp) {
  const int e = blockIdx.x * kBlock + threadIdx.x, K = D.km1;
  if (e >= A.count[nn] * K) return;
  const int sd = A.side[nn], i = e / K, k = 1 + e % K;
  const int c = A.cells[A.first[nn] + i], a = A.along[A.first[nn] + i];
  if (k < D.k1z[c] || k > D.kmz[c]) return;
  const double sbc = A.sc[nn];
  const double fl = BQ(A, sd, 0, k, a) + BQ(A, sd, 2, k, a);
  const double own = cpp[IX(k, c)];
  double b = D.bex[IX(k, c)];
  switch (sd) {
    case 0: b = b + fl * ((fl >= 0.0 && sbc > 0.0) ? sbc : own) / D.P.twodx; break;
    case 1: b = b - fl * ((fl < 0.0 && sbc > 0.0) ? sbc : own) / D.P.twody; break;
    case 2: b = b - fl * ((fl < 0.0 && sbc > 0.0) ? sbc : own) / D.P.twodx; break;
    default: b = b + fl * ((fl > 0.0 && sbc > 0.0) ? sbc : own) / D.P.twody; break;
  }
  D.bex[IX(k, c)] = (R)b;
}

// openbcSCA / openbcTracer, si3d_boundaryconditions.f90:3157-3311, 3661-3806
// (type 1): the boundary value where the flow enters, the inner neighbour's
// value elsewhere.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_obc_scalar(Dev<R> D, ObcArgs A, int nn, R *cn) {
  const int e = blockIdx.x * kBlock + threadIdx.x, K = D.km1;
  if (e >= A.count[nn] * K) return;
  const int sd = A.side[nn], c = A.cells[A.first[nn] + e / K], k = 1 + e % K;
  if (k < D.k1z[c] || k > D.kmz[c]) return;
  const double sbc = A.sc[nn];
  bool in;
  int ln;
  switch (sd) {
    case 0: in = (double)D.uh[IX(k, c)] > 0.0; ln = D.lE[c]; break;
    case 1: in = (double)D.vh[IX(k, D.lS[c])] < 0.0; ln = D.lS[c]; break;
    case 2: in = (double)D.uh[IX(k, D.lW[c])] < 0.0; ln = D.lW[c]; break;
    default: in = (double)D.vh[IX(k, c)] > 0.0; ln = D.lN[c]; break;
  }
  cn[IX(k, c)] = (in && sbc > 0.0) ? (R)sbc : cn[IX(k, ln)];
}

// Time levels of the boundary rows: settrap (mode 0, 3558-3604), save with
// istep 1 (mode 1, 3872-3918) and istep 2 (mode 2, 3722-3768).  A boundary
// shifts both rows of its orientation (W and E, or N and S) over its range,
// so overlapping ranges shift twice, as in the reference.
__global__ void __launch_bounds__(kBlock) k_obc_shift(ObcArgs A, int nn, int km1, int mode) {
  const int e = blockIdx.x * kBlock + threadIdx.x, cnt = A.count[nn], sd = A.side[nn];
  if (e >= 2 * cnt * km1) return;
  const bool we = (sd == 0 || sd == 2);
  const int side = (e / (cnt * km1)) == 0 ? (we ? 2 : 1) : (we ? 0 : 3);
  const int r = e % (cnt * km1), a = A.along[A.first[nn] + r / km1], k = 1 + r % km1;
  if (mode <= 1) {
    BQ(A, side, 2, k, a) = BQ(A, side, 1, k, a);
    BH(A, side, 2, k, a) = BH(A, side, 1, k, a);
  }
  if (mode == 0) {
    BQ(A, side, 1, k, a) = (BQ(A, side, 2, k, a) + BQ(A, side, 0, k, a)) / 2.;
    BH(A, side, 1, k, a) = (BH(A, side, 1, k, a) + BH(A, side, 0, k, a)) / 2.;
  } else {
    BQ(A, side, 1, k, a) = BQ(A, side, 0, k, a);
    BH(A, side, 1, k, a) = BH(A, side, 0, k, a);
  }
}

// ---------------------------------------------------------------------------
// host side

template <typename R> static Dev<R> &devR(s3g_ctx *c);
template <> Dev<double> &devR<double>(s3g_ctx *c) { return c->d64; }
template <> Dev<float> &devR<float>(s3g_ctx *c) { return c->d32; }

// Value q of boundary nn's series at time thrs (hours): parab over records
// dtsecOpenBC apart (938-939, 3187-3188, si3d_utils.f90:895-920); past the
// last record the end parabola is extended.
static double obc_value(const OpenBc &B, int nn, int q, double thrs) {
  const double *fx = &B.vals[((size_t)nn * B.nv + q) * (B.np + 1)];
  const double dxs = B.dtsec / 3600.;
  int m = (int)(thrs / dxs);
  double theta = (thrs - (double)m * dxs) / dxs;
  if (m == 0) {
    m = 2;
    theta -= 1.0;
  } else {
    m = m + 1;
  }
  if (m + 1 > B.np) m = B.np - 1;
  return fx[m] + 0.5 * theta * (fx[m + 1] - fx[m - 1] + theta * (fx[m + 1] + fx[m - 1] - 2.0 * fx[m]));
}

// Kernel arguments of one call; q < 0 leaves val/sc unset, q = 0 fills val
// (elevation or flow), q >= 1 fills sc from value q.
static ObcArgs obc_args(s3g_ctx *c, int q) {
  const OpenBc &B = c->obc;
  ObcArgs A{};
  A.n = B.n;
  A.KB = B.KB;
  A.SZ = B.NA * B.KB;
  for (int nn = 0; nn < B.n; ++nn) {
    A.side[nn] = B.side[nn];
    A.type[nn] = B.type[nn];
    A.first[nn] = B.first[nn];
    A.count[nn] = B.count[nn];
    if (q == 0) A.val[nn] = obc_value(B, nn, 0, c->thrs);
    if (q >= 1) A.sc[nn] = obc_value(B, nn, q, c->thrs);
  }
  for (int nn = 0; nn < B.n; ++nn) A.hbc[nn] = B.hbc[nn];
  A.cells = B.cells;
  A.along = B.along;
  A.bq = B.bq;
  A.bh = B.bh;
  A.part = B.part;
  A.dzmin = B.dzmin;
  return A;
}

#define OBC_LAUNCH(name, kern, n, ...)                                              \
  do {                                                                                \
    int _slot = -1;                                                                   \
    if (c->timing) timer_begin(c, name, &_slot);                                      \
    hipLaunchKernelGGL(kern, dim3(((n) + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, __VA_ARGS__); \
    if (c->timing) timer_end(c, _slot);                                               \
  } while (0)

// openbcUVH for every boundary (first = 1: the t = 0 values of openbc0 from
// the given values).
template <typename R>
static void uvh(s3g_ctx *c, const ObcArgs &A, int first) {
  const OpenBc &B = c->obc;
  for (int nn = 0; nn < B.n; ++nn) {
    if (B.type[nn] == 1) {
      OBC_LAUNCH("openbc", (k_obc_level<R>), B.count[nn], devR<R>(c), A, nn, first);
    } else {
      const int n = B.count[nn] * c->km1, nparts = (n + kBlock - 1) / kBlock;
      OBC_LAUNCH("openbc", (k_obc_section<R>), n, devR<R>(c), A, nn, first);
      OBC_LAUNCH("openbc", (k_obc_flow<R>), n, devR<R>(c), A, nn, first, nparts);
    }
  }
}

template <typename R>
void obc_uvh(s3g_ctx *c) {
  if (c->obc.n == 0) return;
  uvh<R>(c, obc_args(c, 0), 0);
}

template <typename R>
void obc_exmom(s3g_ctx *c, bool x, double twodt1) {
  const OpenBc &B = c->obc;
  if (B.n == 0) return;
  const ObcArgs A = obc_args(c, -1);
  for (int nn = 0; nn < B.n; ++nn) {
    if (x != (B.side[nn] == 0 || B.side[nn] == 2)) continue;
    if (x)
      OBC_LAUNCH("openbc", (k_obc_exmom<R, true>), B.count[nn] * c->km1, devR<R>(c), A, nn, twodt1);
    else
      OBC_LAUNCH("openbc", (k_obc_exmom<R, false>), B.count[nn] * c->km1, devR<R>(c), A, nn, twodt1);
  }
}

template <typename R>
void obc_qq(s3g_ctx *c, double tz) {
  const OpenBc &B = c->obc;
  if (B.n == 0) return;
  const ObcArgs A = obc_args(c, -1);
  for (int nn = 0; nn < B.n; ++nn)
    OBC_LAUNCH("openbc", (k_obc_qq<R>), B.type[nn] == 1 ? B.count[nn] : B.count[nn] * kWave, devR<R>(c), A, nn, tz);
}

template <typename R>
void obc_vel(s3g_ctx *c) {
  const OpenBc &B = c->obc;
  if (B.n == 0) return;
  const ObcArgs A = obc_args(c, -1);
  for (int nn = 0; nn < B.n; ++nn)
    OBC_LAUNCH("openbc", (k_obc_vel<R>), B.type[nn] == 1 ? B.count[nn] * c->km1 : B.count[nn] * kWave, devR<R>(c), A,
               nn);
}

template <typename R>
void obc_exscalar(s3g_ctx *c, const R *cpp, int q) {
  const OpenBc &B = c->obc;
  if (B.n == 0 || !B.any_flow) return;
  const ObcArgs A = obc_args(c, q);
  for (int nn = 0; nn < B.n; ++nn)
    if (B.type[nn] == 2) OBC_LAUNCH("openbc", (k_obc_exscalar<R>), B.count[nn] * c->km1, devR<R>(c), A, nn, cpp);
}

template <typename R>
void obc_scalar(s3g_ctx *c, R *cn, int q) {
  const OpenBc &B = c->obc;
  if (B.n == 0) return;
  const ObcArgs A = obc_args(c, q);
  for (int nn = 0; nn < B.n; ++nn)
    if (B.type[nn] == 1) OBC_LAUNCH("openbc", (k_obc_scalar<R>), B.count[nn] * c->km1, devR<R>(c), A, nn, cn);
}

void obc_shift(s3g_ctx *c, int mode) {
  const OpenBc &B = c->obc;
  if (B.n == 0) return;
  const ObcArgs A = obc_args(c, -1);
  for (int nn = 0; nn < B.n; ++nn)
    OBC_LAUNCH("openbc", k_obc_shift, 2 * B.count[nn] * c->km1, A, nn, c->km1, mode);
}

template void obc_uvh<double>(s3g_ctx *);
template void obc_uvh<float>(s3g_ctx *);
template void obc_exmom<double>(s3g_ctx *, bool, double);
template void obc_exmom<float>(s3g_ctx *, bool, double);
template void obc_qq<double>(s3g_ctx *, double);
template void obc_qq<float>(s3g_ctx *, double);
template void obc_vel<double>(s3g_ctx *);
template void obc_vel<float>(s3g_ctx *);
template void obc_exscalar<double>(s3g_ctx *, const double *, int);
template void obc_exscalar<float>(s3g_ctx *, const float *, int);
template void obc_scalar<double>(s3g_ctx *, double *, int);
template void obc_scalar<float>(s3g_ctx *, float *, int);

}  // namespace s3g

using namespace s3g;

// openbc0 (si3d_boundaryconditions.f90:23-897) for types 1 and 2: checks the
// boundaries like the reference (wet edge cells with a dry outer neighbour,
// no dry water-level cell), stores the series and sets the t = 0 values.
extern "C" int s3g_set_open_boundaries(s3g_ctx *c, const s3g_openbc *bc) {
  if (!c || !bc) {
    set_error("s3g_set_open_boundaries: null argument");
    return S3G_ERR_ARG;
  }
  if (c->nlakes != 1) {
    set_error("s3g_set_open_boundaries: ensembles take no open boundaries");
    return S3G_ERR_UNSUPPORTED;
  }
  if (c->obc.n > 0) {
    set_error("s3g_set_open_boundaries: open boundaries are already set");
    return S3G_ERR_ARG;
  }
  if (bc->nopen <= 0) return S3G_OK;
  if (bc->nopen > kMaxOpen || !bc->iside || !bc->itype || !bc->isbc || !bc->jsbc || !bc->iebc || !bc->jebc ||
      !bc->values || bc->npts < 3 || bc->nvar != c->ntr + 2 || !(bc->dtsec > 0.0)) {
    set_error("s3g_set_open_boundaries: need 1..20 boundaries, >= 3 records of ntr+2 values, dtsec > 0");
    return S3G_ERR_ARG;
  }
  // (i, j) -> column map from the host copy of l2i / l2j
  const int W = c->jm1 + 2, S = c->lm;
  std::vector<int> ij((size_t)(c->im1 + 2) * W, S);
  for (int l = 0; l < c->lm; ++l) ij[(size_t)c->l2i[l] * W + c->l2j[l]] = l;
  auto at = [&](int i, int j) { return (i < 0 || j < 0 || i > c->im1 + 1 || j > c->jm1 + 1) ? S : ij[(size_t)i * W + j]; };
  std::vector<int> kmz(c->lm), k1z(c->lm);
  std::vector<double> hhs(c->lm);
  const bool f64 = (c->precision == S3G_FP64);
  if (hipMemcpy(kmz.data(), f64 ? c->d64.kmz : c->d32.kmz, c->lm * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(k1z.data(), f64 ? c->d64.k1z : c->d32.k1z, c->lm * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(hhs.data(), f64 ? c->d64.hhs : c->d32.hhs, c->lm * sizeof(double), hipMemcpyDeviceToHost) !=
          hipSuccess)
    return hip_check(hipGetLastError(), "s3g_set_open_boundaries copy");
  OpenBc B;
  B.n = bc->nopen;
  B.np = bc->npts;
  B.nv = bc->nvar;
  B.dtsec = bc->dtsec;
  B.dzmin = bc->dzmin;
  std::vector<int> cells, along;
  for (int nn = 0; nn < B.n; ++nn) {
    const int sd = bc->iside[nn] - 1, ty = bc->itype[nn];
    const int is = bc->isbc[nn], js = bc->jsbc[nn], ie = bc->iebc[nn], je = bc->jebc[nn];
    if (sd < 0 || sd > 3) {
      set_error("s3g_set_open_boundaries: iside must be 1 (W), 2 (N), 3 (E) or 4 (S)");
      return S3G_ERR_ARG;
    }
    if (ty != 1 && ty != 2) {
      set_error("s3g_set_open_boundaries: only boundary types 1 (elevation) and 2 (flow) are supported");
      return S3G_ERR_UNSUPPORTED;
    }
    const bool we = (sd == 0 || sd == 2);
    if (ie < is || je < js || is < 1 || js < 1 || ie > c->im1 || je > c->jm1) {
      set_error("s3g_set_open_boundaries: boundary range outside the grid");
      return S3G_ERR_ARG;
    }
    B.side.push_back(sd);
    B.type.push_back(ty);
    B.first.push_back((int)cells.size());
    for (int a = (we ? js : is); a <= (we ? je : ie); ++a) {
      const int i = we ? is : a, j = we ? a : js;
      const int io = i + (sd == 0 ? -1 : sd == 2 ? 1 : 0), jo = j + (sd == 1 ? 1 : sd == 3 ? -1 : 0);
      const int l = at(i, j);
      if (l == S || at(io, jo) != S) {
        set_error("s3g_set_open_boundaries: boundary cell (" + std::to_string(i) + "," + std::to_string(j) +
                  ") is not a wet edge cell of its side");
        return S3G_ERR_ARG;
      }
      if (ty == 1 && kmz[l] - k1z[l] + 1 < 1) {
        set_error("s3g_set_open_boundaries: dry point on a water-level boundary");
        return S3G_ERR_ARG;
      }
      cells.push_back(l);
      along.push_back(a);
    }
    B.count.push_back((int)cells.size() - B.first.back());
    B.any_flow = B.any_flow || ty == 2;
    double hm = -1.e6;  // the section's deepest cell (hbc, 978-982); the bathymetry does not change
    for (int i = B.first.back(); i < (int)cells.size(); ++i) hm = std::max(hm, hhs[cells[i]]);
    B.hbc.push_back(hm);
  }
  B.vals.assign((size_t)B.n * B.nv * (B.np + 1), 0.0);
  for (int nn = 0; nn < B.n; ++nn)
    for (int r = 0; r < B.np; ++r)
      for (int q = 0; q < B.nv; ++q)
        B.vals[((size_t)nn * B.nv + q) * (B.np + 1) + r + 1] = bc->values[((size_t)nn * B.np + r) * B.nv + q];
  B.KB = c->km1 + 2;
  B.NA = std::max(c->im1, c->jm1) + 2;
  const size_t nb = (size_t)12 * B.NA * B.KB * sizeof(double);
  void *p = nullptr;
  if (hipMalloc(&p, cells.size() * sizeof(int)) != hipSuccess) return hip_check(hipGetLastError(), "openbc alloc");
  c->allocs.push_back(p);
  B.cells = (int *)p;
  if (hipMalloc(&p, along.size() * sizeof(int)) != hipSuccess) return hip_check(hipGetLastError(), "openbc alloc");
  c->allocs.push_back(p);
  B.along = (int *)p;
  if (hipMalloc(&p, nb) != hipSuccess) return hip_check(hipGetLastError(), "openbc alloc");
  c->allocs.push_back(p);
  B.bq = (double *)p;
  if (hipMalloc(&p, nb) != hipSuccess) return hip_check(hipGetLastError(), "openbc alloc");
  c->allocs.push_back(p);
  B.bh = (double *)p;
  const size_t nparts = ((size_t)cells.size() * c->km1 + kBlock - 1) / kBlock;
  if (hipMalloc(&p, nparts * sizeof(double)) != hipSuccess) return hip_check(hipGetLastError(), "openbc alloc");
  c->allocs.push_back(p);
  B.part = (double *)p;
  // ordered on the library's stream before the t = 0 kernels below
  if (hipMemcpyAsync(B.cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(B.along, along.data(), along.size() * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemsetAsync(B.bq, 0, nb, c->stream) != hipSuccess || hipMemsetAsync(B.bh, 0, nb, c->stream) != hipSuccess)
    return hip_check(hipGetLastError(), "openbc upload");
  c->obc = B;
  // t = 0 values from the first record (varsOpenBC(nn,1,1), 337-647)
  ObcArgs A = obc_args(c, -1);
  for (int nn = 0; nn < B.n; ++nn) A.val[nn] = B.vals[((size_t)nn * B.nv) * (B.np + 1) + 1];
  if (c->precision == S3G_FP64)
    uvh<double>(c, A, 1);
  else
    uvh<float>(c, A, 1);
  int e = hip_check(hipGetLastError(), "openbc0 launch");
  if (e < 0) return e;
  return hip_check(hipStreamSynchronize(c->stream), "openbc0");
}
