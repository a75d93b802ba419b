// s3g_pcg.hip -- free-surface solve A zeta = qq on gfx950.
//
// Replaces SolverSparse's NSPCG call (si3d_procedures.f90:1870-2065).  A is the
// pentadiagonal SPD matrix assembled by matcon: row c holds rr(c) on the
// diagonal, -sx(W) and -sy(S) for the lower neighbours and -sy(c), -sx(c) for
// the upper ones (1954-1967); it is applied matrix-free from rr/sx/sy and the
// neighbour index arrays.  The initial guess is sp (1970).
//
// Conjugate gradients, templated on the vector type T.  Every reduction is
// FP64 and deterministic: each workgroup writes one partial (64-lane wavefront
// shuffles, then LDS across the four wavefronts) and the next kernel
// re-reduces the partials in a fixed order, so every workgroup derives the
// same alpha, beta and convergence decision without a host round trip.
// Iterations are enqueued in chunks; once converged the remaining kernels of a
// chunk return at entry.
//
// FP64 mode: one CG with T = double, stopped on ||r|| <= tol/2 ||qq|| and
// confirmed by an FP64 true-residual recompute.
// Mixed mode: iterative refinement.  The outer loop forms the FP64 true
// residual r = qq - A zeta; an inner CG with T = float (FP32 matrix copy,
// vectors and SpMV; FP64 dots) reduces it by `kInnerTol`, or on the last sweep
// only as far as the outer tolerance needs; zeta += d in FP64.
//
// Preconditioners: multigrid (below), Jacobi, or an exact tridiagonal solve along each north-south
// line of wet cells (the wet cells of one i-row are consecutive in l, so a line
// is a contiguous range of columns; one lane per line runs the Thomas sweep,
// factors computed once per matrix).
#include <cmath>
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <utility>
#include <vector>

#include "s3g_internal.h"

namespace s3g {

constexpr double kInnerTol = 1e-5;  // largest inner FP32 reduction per refinement sweep

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum; result valid in thread 0.
__device__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += lds[i];
  __syncthreads();
  return s;
}

// N reductions of nblk partials at once (slots[q] * nblk + i); every lane of
// every block gets the same N values (fixed summation order).
template <int N>
__device__ __forceinline__ void reduce_slots(const double *part, const int (&slots)[N], int nblk, double *lds,
                                             double (&out)[N]) {
  double s[N];
#pragma unroll
  for (int q = 0; q < N; ++q) s[q] = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x)
#pragma unroll
    for (int q = 0; q < N; ++q) s[q] += part[slots[q] * nblk + i];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    s[q] = wave_sum(s[q]);
    if (lane == 0) lds[w * N + q] = s[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double t = 0.0;
    for (int v = 0; v < nw; ++v) t += lds[v * N + q];
    out[q] = t;
  }
  __syncthreads();
}

// Partial sums layout: [RZ0 | RZ1 | RR0 | RR1 | PQ | BB | TR | ZR | RY | YAY] x nblk
// (ZR: (z_i, r_{i-1}) of flexible CG; RY, YAY: (r, y) and (y, A y) of the
// learned correction y)
enum { P_RZ0 = 0, P_RZ1 = 1, P_RR0 = 2, P_RR1 = 3, P_PQ = 4, P_BB = 5, P_TR = 6, P_ZR = 7, P_RY = 8, P_YAY = 9,
       P_NUM = 10 };

template <typename T>
struct CgArgs {
  const int *lE, *lW, *lN, *lS;
  const T *rr, *sx, *sy;  // matrix (sx(S) = sy(S) = 0)
  int n, nblk;
  T *x, *r, *r1, *z, *p0, *p1, *q;  // r / r1: residual entering even / odd iterations
  const T *b;
  double *part;
  int *iters;
  double tol2;          // stop when (r,r) <= tol2 * (b,b); the kernels read it from
  const double *tol2p;  // this device copy, set at the start of every CG run
  int nlines;
  const int *line_start;
  const T *lc, *ld;
  int flex;  // flexible CG: Polak-Ribiere beta from the extra ZR partial
  T *yc, *ry;  // correction stage: the network's y and r - A y
};

template <typename T>
__device__ __forceinline__ double Ax(const CgArgs<T> &A, const T *v, int c) {
  const int lW = A.lW[c], lS = A.lS[c], lN = A.lN[c], lE = A.lE[c];
  return (double)A.rr[c] * (double)v[c] - (double)A.sx[lW] * (double)v[lW] - (double)A.sy[lS] * (double)v[lS] -
         (double)A.sy[c] * (double)v[lN] - (double)A.sx[c] * (double)v[lE];
}

// r = b - A x; Jacobi z; partials (r,z), (r,r), (b,b).
template <typename T, bool JACOBI>
__global__ void __launch_bounds__(kBlock) k_cg_residual(CgArgs<T> A) {
  __shared__ double lds[16];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    const double b = A.b[c];
    const double r = b - Ax(A, A.x, c);
    A.r[c] = (T)r;
    if (JACOBI) {
      const double z = (double)(T)r / (double)A.rr[c];
      A.z[c] = (T)z;
      rz += (double)(T)r * (double)(T)z;
    }
    rr += r * r;
    bb += b * b;
  }
  rz = block_sum(rz, lds);
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    if (JACOBI) A.part[P_RZ0 * A.nblk + blockIdx.x] = rz;
    A.part[P_RR0 * A.nblk + blockIdx.x] = rr;
    A.part[P_BB * A.nblk + blockIdx.x] = bb;
    if (blockIdx.x == 0) A.iters[0] = 0;
  }
}

// Line preconditioner z = T^-1 r (tridiagonal part of A along each N-S line),
// partial (r,z) into SLOT.  Launched with nblk blocks: blocks beyond the line
// count write zero partials.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_line_solve(CgArgs<T> A, const T *r, int slot) {
  __shared__ double lds[16];
  double rz = 0.0;
  for (int ln = blockIdx.x * blockDim.x + threadIdx.x; ln < A.nlines; ln += gridDim.x * blockDim.x) {
    const int c0 = A.line_start[ln], c1 = A.line_start[ln + 1];
    double y = 0.0;
    for (int c = c0; c < c1; ++c) {
      const double off = (c > c0) ? (double)A.sy[c - 1] : 0.0;
      y = ((double)r[c] + off * y) * (double)A.ld[c];
      A.z[c] = (T)y;
    }
    double zn = 0.0;
    for (int c = c1 - 1; c >= c0; --c) {
      const double zc = (c < c1 - 1) ? (double)A.z[c] - (double)A.lc[c] * zn : (double)A.z[c];
      A.z[c] = (T)zc;
      zn = zc;
      rz += (double)r[c] * (double)(T)zc;
    }
  }
  rz = block_sum(rz, lds);
  if (threadIdx.x == 0) A.part[slot * A.nblk + blockIdx.x] = rz;
}

// Iteration i, step 1: beta, p_new = z + beta p_old at the column and its four
// neighbours (double-buffered p), q = A p_new, partial (p,q).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_dir(CgArgs<T> A, int i) {
  __shared__ double lds[32];
  const int cur = i & 1;
  double v[5];
  if (A.flex) {
    reduce_slots<5>(A.part, {P_RR0 + cur, P_BB, P_RZ0 + cur, P_RZ0 + (cur ^ 1), P_ZR}, A.nblk, lds, v);
  } else {
    double u[4];
    reduce_slots<4>(A.part, {P_RR0 + cur, P_BB, P_RZ0 + cur, P_RZ0 + (cur ^ 1)}, A.nblk, lds, u);
    for (int q = 0; q < 4; ++q) v[q] = u[q];
    v[4] = 0.0;
  }
  if (v[0] <= *A.tol2p * v[1]) return;  // converged: every block takes the same branch
  // Fletcher-Reeves, or Polak-Ribiere for flexible CG (the two agree while
  // the preconditioner is fixed, where (z_i, r_{i-1}) = 0)
  double beta = (i > 0) ? (v[2] - v[4]) / v[3] : 0.0;
  if (A.flex && !isfinite(beta)) beta = 0.0;  // restart along z
  const T *pold = cur ? A.p0 : A.p1;
  T *pnew = cur ? A.p1 : A.p0;
  double pq = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    const int lW = A.lW[c], lS = A.lS[c], lN = A.lN[c], lE = A.lE[c];
    auto pn = [&](int j) { return (double)(T)((double)A.z[j] + beta * (double)pold[j]); };  // z(S)=p(S)=0
    const double pc = pn(c);
    pnew[c] = (T)pc;
    const double q = (double)A.rr[c] * pc - (double)A.sx[lW] * pn(lW) - (double)A.sy[lS] * pn(lS) -
                     (double)A.sy[c] * pn(lN) - (double)A.sx[c] * pn(lE);
    A.q[c] = (T)q;
    pq += pc * (double)(T)q;
  }
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) A.part[P_PQ * A.nblk + blockIdx.x] = pq;
}

// Iteration i, step 2: alpha, x += alpha p, r -= alpha q, Jacobi z, partials
// (r,z) and (r,r) for iteration i+1.
template <typename T, bool JACOBI>
__global__ void __launch_bounds__(kBlock) k_cg_update(CgArgs<T> A, int i) {
  __shared__ double lds[32];
  const int cur = i & 1, nxt = cur ^ 1;
  double v[4];
  reduce_slots<4>(A.part, {P_RR0 + cur, P_BB, P_RZ0 + cur, P_PQ}, A.nblk, lds, v);
  if (v[0] <= *A.tol2p * v[1]) {
    // Converged: carry this residual partial into the slot iteration i+1 reads,
    // so every later kernel of the chunk also sees convergence and returns.
    if (threadIdx.x == 0) A.part[(P_RR0 + nxt) * A.nblk + blockIdx.x] = A.part[(P_RR0 + cur) * A.nblk + blockIdx.x];
    return;
  }
  const double alpha = v[2] / v[3];
  const T *p = cur ? A.p1 : A.p0;
  const T *ro = cur ? A.r1 : A.r;
  T *rn = cur ? A.r : A.r1;
  double rzn = 0.0, rrn = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    A.x[c] = (T)((double)A.x[c] + alpha * (double)p[c]);
    const double r = (double)(T)((double)ro[c] - alpha * (double)A.q[c]);
    rn[c] = (T)r;
    if (JACOBI) {
      const double z = (double)(T)(r / (double)A.rr[c]);
      A.z[c] = (T)z;
      rzn += r * z;
    }
    rrn += r * r;
  }
  rzn = block_sum(rzn, lds);
  rrn = block_sum(rrn, lds);
  if (threadIdx.x == 0) {
    if (JACOBI) A.part[(P_RZ0 + nxt) * A.nblk + blockIdx.x] = rzn;
    A.part[(P_RR0 + nxt) * A.nblk + blockIdx.x] = rrn;
    if (blockIdx.x == 0) A.iters[0] += 1;  // iterations since the run's first residual
  }
}

// ---- flexible CG: the learned correction stage ----------------------------
// (z_i, r_{i-1}) partial for the Polak-Ribiere beta; rprev is the residual
// entering the previous iteration.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_flex_zr(CgArgs<T> A, const T *rprev) {
  __shared__ double lds[16];
  double s = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x)
    s += (double)A.z[c] * (double)rprev[c];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) A.part[P_ZR * A.nblk + blockIdx.x] = s;
}

// y = sigma * out(pix) (the network's estimate of A^-1 r) into yc, A y into
// ry, partials (r, y) and (y, A y).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_flex_apply(CgArgs<T> A, const T *r, const int *pix, const float *out,
                                                      const double *sig) {
  __shared__ double lds[16];
  const double s = sig[0];
  auto y = [&](int j) { return j == A.n ? 0.0 : (double)(T)(s * (double)out[pix[j]]); };
  double ry = 0.0, yay = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    const int lW = A.lW[c], lS = A.lS[c], lN = A.lN[c], lE = A.lE[c];
    const double yc = y(c);
    A.yc[c] = (T)yc;
    const double ay = (double)A.rr[c] * yc - (double)A.sx[lW] * y(lW) - (double)A.sy[lS] * y(lS) -
                      (double)A.sy[c] * y(lN) - (double)A.sx[c] * y(lE);
    A.ry[c] = (T)ay;
    ry += (double)r[c] * yc;
    yay += yc * ay;
  }
  ry = block_sum(ry, lds);
  yay = block_sum(yay, lds);
  if (threadIdx.x == 0) {
    A.part[P_RY * A.nblk + blockIdx.x] = ry;
    A.part[P_YAY * A.nblk + blockIdx.x] = yay;
  }
}

// Energy line search on y: gamma = (r, y) / (y, A y) minimises the A-norm of
// the error after the step gamma y (0 if y is zero or not finite); yc :=
// gamma y, ry := r - gamma A y.  The step cannot raise the error's A-norm and
// the preconditioner that follows (multigrid, line or Jacobi, each an A-norm
// contraction here) lowers it further, so z = gamma y + M^-1 (r - gamma A y)
// satisfies 2 (r, z) > (z, A z) > 0: a descent direction whatever the network
// returns.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_flex_scale(CgArgs<T> A, const T *r) {
  __shared__ double lds[16];
  double v[2];
  reduce_slots<2>(A.part, {P_RY, P_YAY}, A.nblk, lds, v);
  double gamma = v[1] > 0.0 ? v[0] / v[1] : 0.0;
  if (!isfinite(gamma)) gamma = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    A.yc[c] = (T)(gamma * (double)A.yc[c]);
    A.ry[c] = (T)((double)r[c] - gamma * (double)A.ry[c]);
  }
}

// z = y + M^-1 (r - A y): z already holds M^-1 (r - A y), or with JACOBI it is
// (r - A y) / rr here; partials (r, z) into SLOT and (r_prev, z) into ZR
// (zero without a previous residual).
template <typename T, bool JACOBI>
__global__ void __launch_bounds__(kBlock) k_flex_finish(CgArgs<T> A, const T *r, const T *rprev, int slot) {
  __shared__ double lds[16];
  double rz = 0.0, zr = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.n; c += gridDim.x * blockDim.x) {
    const double m = JACOBI ? (double)(T)((double)A.ry[c] / (double)A.rr[c]) : (double)A.z[c];
    const double z = (double)(T)((double)A.yc[c] + m);
    A.z[c] = (T)z;
    rz += (double)r[c] * z;
    if (rprev) zr += (double)rprev[c] * z;
  }
  rz = block_sum(rz, lds);
  zr = block_sum(zr, lds);
  if (threadIdx.x == 0) {
    A.part[slot * A.nblk + blockIdx.x] = rz;
    A.part[P_ZR * A.nblk + blockIdx.x] = zr;
  }
}

// Line factors of tridiag(-sy(c-1), rr(c), -sy(c)) per line: ld = 1/pivot,
// lc = modified super-diagonal.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_line_factor(CgArgs<T> A, T *lc, T *ld) {
  for (int ln = blockIdx.x * blockDim.x + threadIdx.x; ln < A.nlines; ln += gridDim.x * blockDim.x) {
    const int c0 = A.line_start[ln], c1 = A.line_start[ln + 1];
    double cprev = 0.0;
    for (int c = c0; c < c1; ++c) {
      const double a = (c > c0) ? -(double)A.sy[c - 1] : 0.0;
      const double inv = 1.0 / ((double)A.rr[c] - a * cprev);
      ld[c] = (T)inv;
      const double csup = (c < c1 - 1) ? -(double)A.sy[c] * inv : 0.0;
      lc[c] = (T)csup;
      cprev = csup;
    }
  }
}


// ---- aggregation multigrid (preconditioner S3G_PREC_MG) --------------------
// V(1,1) cycle with damped Jacobi smoothing (omega = kMgOmega) on levels of 2x2
// aggregates with piecewise-constant transfers.  The coarse operators stay
// 5-point: diagonal = the children's mass plus the couplings that leave the
// aggregate, off-diagonals = the couplings that leave it eastward / northward,
// each scaled by s3g_ctx::mg_cs.  The Galerkin product P^T A P (mg_cs = 1)
// counts two fine faces per coarse face and so doubles the coarse diffusion;
// scaling it back towards the coarse grid's own discretisation (mg_cs = 1/2)
// restores the coarse-grid correction of smooth errors.  The coarsest level
// (<= 1024 cells) is smoothed by a few damped Jacobi sweeps
// (s3g_ctx::mg_sweeps) inside one workgroup.  Every step is a fixed linear
// operator, every coarse operator is SPD and pre/post smoothing are identical,
// so the preconditioner is symmetric positive definite and plain CG applies.
constexpr double kMgOmega = 0.8;
constexpr int kMgCoarseMax = 1024;

template <typename T>
struct MgPair {  // a fine level f and the next coarser level c
  int nf, nc;
  const int *fE, *fW, *fN, *fS;
  const T *frr, *fsx, *fsy, *fb;
  T *fx;
  const int *ch, *parent;
  T *crr, *csx, *csy, *cb;
  const T *cx;
};

// cs scales the couplings that leave the aggregate (diagonal = children's
// mass + the scaled outward couplings): cs = 1 is the Galerkin operator,
// cs = 1/2 the 5-point operator of the coarse grid's own discretisation (a
// coarse face crosses two fine faces).  Either keeps A_c SPD.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_galerkin(MgPair<T> M, double cs) {
  for (int C = blockIdx.x * blockDim.x + threadIdx.x; C < M.nc; C += gridDim.x * blockDim.x) {
    double rr = 0.0, inner = 0.0, sx = 0.0, sy = 0.0, ext = 0.0;
    for (int s = 0; s < 4; ++s) {
      const int f = M.ch[4 * C + s];
      if (f == M.nf) continue;
      rr += (double)M.frr[f];
      const int e = M.fE[f], n = M.fN[f], w = M.fW[f], so = M.fS[f];
      if (e != M.nf) {
        if (M.parent[e] == C) inner += (double)M.fsx[f];
        else sx += (double)M.fsx[f];
      }
      if (n != M.nf) {
        if (M.parent[n] == C) inner += (double)M.fsy[f];
        else sy += (double)M.fsy[f];
      }
      if (w != M.nf && M.parent[w] != C) ext += (double)M.fsx[w];
      if (so != M.nf && M.parent[so] != C) ext += (double)M.fsy[so];
    }
    ext += sx + sy;
    M.crr[C] = (T)(rr - 2.0 * inner - (1.0 - cs) * ext);
    M.csx[C] = (T)(cs * sx);
    M.csy[C] = (T)(cs * sy);
  }
}

// The restriction kernels give each child slot of an aggregate its own lane:
// lane 4C + s handles child s of aggregate C (ch[4C + s], the finer level's
// sentinel when absent), so the four children of an aggregate sit in four
// consecutive lanes of one wavefront and a wavefront reads two runs of 32
// consecutive fine cells instead of 64 cells two apart.  The aggregate's sum
// is formed by the first lane of the four in child order (the order of a
// sequential loop over the children), with the other three values fetched by
// in-wavefront shuffles.
__device__ __forceinline__ double child_sum(double v, bool valid) {
  const double v1 = __shfl(v, 1, 4), v2 = __shfl(v, 2, 4), v3 = __shfl(v, 3, 4);
  const int m = __shfl((int)valid, 0, 4) | (__shfl((int)valid, 1, 4) << 1) | (__shfl((int)valid, 2, 4) << 2) |
                (__shfl((int)valid, 3, 4) << 3);
  double acc = 0.0;
  if (m & 1) acc += v;
  if (m & 2) acc += v1;
  if (m & 4) acc += v2;
  if (m & 8) acc += v3;
  return acc;
}

// Pre-smoothing from a zero guess (x = omega D^-1 b), residual, restriction.
// Launched with a multiple of 4 lanes per workgroup; every group of four
// lanes runs the loop the same number of times (4 nc is a multiple of 4).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_down(MgPair<T> M) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 4 * M.nc; q += gridDim.x * blockDim.x) {
    const int f = M.ch[q];
    const bool valid = f != M.nf;
    double t = 0.0;
    if (valid) {
      auto xs = [&](int j) { return kMgOmega * (double)M.fb[j] / (double)M.frr[j]; };  // fb(S)=0, frr(S)=1
      const int w = M.fW[f], so = M.fS[f], n = M.fN[f], e = M.fE[f];
      const double ax = (double)M.frr[f] * xs(f) - (double)M.fsx[w] * xs(w) - (double)M.fsy[so] * xs(so) -
                        (double)M.fsy[f] * xs(n) - (double)M.fsx[f] * xs(e);
      t = (double)M.fb[f] - ax;
    }
    const double acc = child_sum(t, valid);
    if ((q & 3) == 0) M.cb[q >> 2] = (T)acc;
  }
}

// Prolongation + post-smoothing: y = omega D^-1 b + P x_c, x = y + omega D^-1 (b - A y).
// With SLOT >= 0 (finest level) also writes the partial of (b, x) = (r, z).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_mg_up(MgPair<T> M, double *part, int nblk, int slot) {
  __shared__ double lds[16];
  double rz = 0.0;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < M.nf; f += gridDim.x * blockDim.x) {
    auto ys = [&](int j) {  // j = sentinel -> 0 (fb = 0, parent -> coarse sentinel with cx = 0)
      return kMgOmega * (double)M.fb[j] / (double)M.frr[j] + (double)M.cx[M.parent[j]];
    };
    const int w = M.fW[f], so = M.fS[f], n = M.fN[f], e = M.fE[f];
    const double yf = ys(f);
    const double ay = (double)M.frr[f] * yf - (double)M.fsx[w] * ys(w) - (double)M.fsy[so] * ys(so) -
                      (double)M.fsy[f] * ys(n) - (double)M.fsx[f] * ys(e);
    const double x = (double)(T)(yf + kMgOmega * ((double)M.fb[f] - ay) / (double)M.frr[f]);
    M.fx[f] = (T)x;
    rz += (double)M.fb[f] * x;
  }
  if (slot >= 0) {
    rz = block_sum(rz, lds);
    if (threadIdx.x == 0) part[slot * nblk + blockIdx.x] = rz;
  }
}

// CG iteration i, step 2, fused with the first restriction of the V-cycle:
// alpha, x += alpha p and r_new = r - alpha q for the children of each level-1
// aggregate, (r_new, r_new) partial, and the restricted residual of the
// pre-smoothed level 0 (as k_mg_down, with r_new at the neighbours recomputed
// from the previous residual buffer).  Launched with nblk blocks.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_update_mg(CgArgs<T> A, MgPair<T> M, int i) {
  __shared__ double lds[32];
  const int cur = i & 1, nxt = cur ^ 1;
  double v[4];
  reduce_slots<4>(A.part, {P_RR0 + cur, P_BB, P_RZ0 + cur, P_PQ}, A.nblk, lds, v);
  if (v[0] <= *A.tol2p * v[1]) {
    if (threadIdx.x == 0) A.part[(P_RR0 + nxt) * A.nblk + blockIdx.x] = A.part[(P_RR0 + cur) * A.nblk + blockIdx.x];
    return;
  }
  const double alpha = v[2] / v[3];
  const T *p = cur ? A.p1 : A.p0;
  const T *ro = cur ? A.r1 : A.r;
  T *rn = cur ? A.r : A.r1;
  auto rnew = [&](int j) { return (double)(T)((double)ro[j] - alpha * (double)A.q[j]); };  // r(S) = q(S) = 0
  double rrn = 0.0;
  // one lane per child slot of a level-1 aggregate (see k_mg_down)
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 4 * M.nc; q += gridDim.x * blockDim.x) {
    const int f = M.ch[q];
    const bool valid = f != M.nf;
    double t = 0.0;
    if (valid) {
      A.x[f] = (T)((double)A.x[f] + alpha * (double)p[f]);
      const double rf = rnew(f);
      rn[f] = (T)rf;
      rrn += rf * rf;
      auto xs = [&](int j) { return kMgOmega * rnew(j) / (double)M.frr[j]; };
      const int w = M.fW[f], so = M.fS[f], n = M.fN[f], e = M.fE[f];
      const double ax = (double)M.frr[f] * (kMgOmega * rf / (double)M.frr[f]) - (double)M.fsx[w] * xs(w) -
                        (double)M.fsy[so] * xs(so) - (double)M.fsy[f] * xs(n) - (double)M.fsx[f] * xs(e);
      t = rf - ax;
    }
    const double acc = child_sum(t, valid);
    if ((q & 3) == 0) M.cb[q >> 2] = (T)acc;
  }
  rrn = block_sum(rrn, lds);
  if (threadIdx.x == 0) {
    A.part[(P_RR0 + nxt) * A.nblk + blockIdx.x] = rrn;
    if (blockIdx.x == 0) A.iters[0] += 1;  // iterations since the run's first residual
  }
}

// The small levels of the hierarchy (each at most s3g_ctx::mg_small cells) in one
// workgroup of kMgCoarseMax lanes:
// restriction sweeps, the coarsest smoothing and the prolongation sweeps, with
// a workgroup barrier between levels.  Same arithmetic as k_mg_down /
// k_mg_up.
constexpr int kMgMaxLevels = 12;

template <typename T>
struct MgChain {
  int nlev;                       // levels in the chain (first = the finest handled here)
  MgPair<T> pr[kMgMaxLevels];     // pr[i]: level i (fine) -> level i+1 (coarse), i < nlev-1
  int nc;                         // coarsest size
  int sweeps;                     // damped Jacobi sweeps on the coarsest level
  const int *cE, *cW, *cN, *cS;
  const T *crr, *csx, *csy, *cb;
  T *cx;
};

template <typename T>
__global__ void __launch_bounds__(1024) k_mg_chain(MgChain<T> Ch) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int i = 0; i + 1 < Ch.nlev; ++i) {
    const MgPair<T> &M = Ch.pr[i];
    for (int C = tid; C < M.nc; C += nt) {
      double acc = 0.0;
      for (int s2 = 0; s2 < 4; ++s2) {
        const int f = M.ch[4 * C + s2];
        if (f == M.nf) continue;
        auto xs = [&](int j) { return kMgOmega * (double)M.fb[j] / (double)M.frr[j]; };
        const int w = M.fW[f], so = M.fS[f], n = M.fN[f], e = M.fE[f];
        const double ax = (double)M.frr[f] * xs(f) - (double)M.fsx[w] * xs(w) - (double)M.fsy[so] * xs(so) -
                          (double)M.fsy[f] * xs(n) - (double)M.fsx[f] * xs(e);
        acc += (double)M.fb[f] - ax;
      }
      M.cb[C] = (T)acc;
    }
    __syncthreads();
  }
  // coarsest (at most blockDim.x cells): damped Jacobi sweeps in LDS
  {
    __shared__ double xs[kMgCoarseMax + 1];
    const int n = Ch.nc, cc = tid;
    const bool act = cc < n;
    if (cc == 0) xs[n] = 0.0;
    double bc = 0, rc = 1, sxc = 0, syc = 0, sxw = 0, sys = 0;
    int w = n, so = n, no = n, e = n;
    if (act) {
      bc = Ch.cb[cc]; rc = Ch.crr[cc]; sxc = Ch.csx[cc]; syc = Ch.csy[cc];
      w = Ch.cW[cc]; so = Ch.cS[cc]; no = Ch.cN[cc]; e = Ch.cE[cc];
      sxw = Ch.csx[w]; sys = Ch.csy[so];
      xs[cc] = kMgOmega * bc / rc;
    }
    __syncthreads();
    for (int it = 1; it < Ch.sweeps; ++it) {
      double xn = 0.0;
      if (act) {
        const double ax = rc * xs[cc] - sxw * xs[w] - sys * xs[so] - syc * xs[no] - sxc * xs[e];
        xn = xs[cc] + kMgOmega * (bc - ax) / rc;
      }
      __syncthreads();
      if (act) xs[cc] = xn;
      __syncthreads();
    }
    if (act) Ch.cx[cc] = (T)xs[cc];
    __syncthreads();
  }
  for (int i = Ch.nlev - 2; i >= 0; --i) {
    const MgPair<T> &M = Ch.pr[i];
    for (int f = tid; f < M.nf; f += nt) {
      auto ys = [&](int j) { return kMgOmega * (double)M.fb[j] / (double)M.frr[j] + (double)M.cx[M.parent[j]]; };
      const int w = M.fW[f], so = M.fS[f], n = M.fN[f], e = M.fE[f];
      const double yf = ys(f);
      const double ay = (double)M.frr[f] * yf - (double)M.fsx[w] * ys(w) - (double)M.fsy[so] * ys(so) -
                        (double)M.fsy[f] * ys(n) - (double)M.fsx[f] * ys(e);
      M.fx[f] = (T)(yf + kMgOmega * ((double)M.fb[f] - ay) / (double)M.frr[f]);
    }
    __syncthreads();
  }
}

// ---- FP64 outer level (zeta, qq, true residual) ---------------------------
// x = s, or with `extrap` the linear extrapolation 2 sp - spp of the two
// latest levels (sp = zeta^n, spp = zeta^{n-1} before a leapfrog solve).
__global__ void __launch_bounds__(kBlock) k_start_guess(double *x, const double *s, const double *sp,
                                                        const double *spp, int n, int extrap) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= n; c += gridDim.x * blockDim.x)
    x[c] = (c < n) ? (extrap ? 2.0 * sp[c] - spp[c] : s[c]) : 0.0;
}

// FP64 true residual; optionally stores it as the right-hand side (cast) of the
// inner solve and clears the inner vectors.  Partials: TR = (r,r), BB = (b,b).
template <typename T, bool STORE>
__global__ void __launch_bounds__(kBlock) k_true_residual(Outer O, T *rin, T *d, T *p0, T *p1, T *z) {
  __shared__ double lds[16];
  double rr = 0.0, bb = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
    const int lW = O.lW[c], lS = O.lS[c], lN = O.lN[c], lE = O.lE[c];
    const double b = O.b[c];
    const double r = b - (O.rr[c] * O.x[c] - O.sx[lW] * O.x[lW] - O.sy[lS] * O.x[lS] - O.sy[c] * O.x[lN] -
                          O.sx[c] * O.x[lE]);
    if (STORE) {
      rin[c] = (T)r;
      d[c] = T(0);
      p0[c] = p1[c] = z[c] = T(0);
    }
    rr += r * r;
    bb += b * b;
  }
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    O.part[P_TR * O.nblk + blockIdx.x] = rr;
    O.part[P_BB * O.nblk + blockIdx.x] = bb;
  }
}

// Per-lake true residual of an ensemble: workgroup (slice, lake) sums r^2 and
// b^2 over its share of the lake's contiguous columns off[lake] ..
// off[lake+1]-1 into out[(lake * slices + slice) * 2 + {0, 1}].
constexpr int kLakeSlices = 8;
__global__ void __launch_bounds__(kBlock) k_lake_residual(Outer O, const int *off, double *out) {
  __shared__ double lds[16];
  const int lake = blockIdx.y;
  const int a = off[lake], e = off[lake + 1];
  double rr = 0.0, bb = 0.0;
  for (int c = a + blockIdx.x * blockDim.x + threadIdx.x; c < e; c += gridDim.x * blockDim.x) {
    const int lW = O.lW[c], lS = O.lS[c], lN = O.lN[c], lE = O.lE[c];
    const double b = O.b[c];
    const double r = b - (O.rr[c] * O.x[c] - O.sx[lW] * O.x[lW] - O.sy[lS] * O.x[lS] - O.sy[c] * O.x[lN] -
                          O.sx[c] * O.x[lE]);
    rr += r * r;
    bb += b * b;
  }
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    out[((size_t)lake * gridDim.x + blockIdx.x) * 2] = rr;
    out[((size_t)lake * gridDim.x + blockIdx.x) * 2 + 1] = bb;
  }
}

// x += d (FP64 accumulation of the inner correction).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_add_correction(double *x, const T *d, int n) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) x[c] += (double)d[c];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_cast_matrix(const double *rr, const double *sx, const double *sy, T *rrT,
                                                        T *sxT, T *syT, int n) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= n; c += gridDim.x * blockDim.x) {
    rrT[c] = (T)rr[c];
    sxT[c] = (T)sx[c];
    syT[c] = (T)sy[c];
  }
}

__global__ void k_copy_to_s(double *s, const double *x, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) s[c] = x[c];
}

// ---------------------------------------------------------------------------

template <typename T>
static T *mg_alloc(s3g_ctx *c, size_t n, double fill_last) {
  void *p = nullptr;
  if (hipMalloc(&p, (n + 1) * sizeof(T)) != hipSuccess) return nullptr;
  c->allocs.push_back(p);
  std::vector<T> h(n + 1, T(0));
  h[n] = (T)fill_last;
  hipMemcpy(p, h.data(), (n + 1) * sizeof(T), hipMemcpyHostToDevice);
  return (T *)p;
}

static int *upload_ints(s3g_ctx *c, const std::vector<int> &v) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(int)) != hipSuccess) return nullptr;
  c->allocs.push_back(p);
  hipMemcpy(p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice);
  return (int *)p;
}

// Hierarchy of 2x2 aggregates built once from the (i,j) position of each wet
// column; ordering i outer / j inner on every level.
static int mg_setup(s3g_ctx *c, const int *l2i, const int *l2j, bool fp32) {
  Pcg &P = c->pcg;
  P.mg.clear();
  P.mg.emplace_back();  // level 0 = the lake grid
  std::vector<int> ci(c->lm), cj(c->lm);
  for (int k = 0; k < c->lm; ++k) ci[k] = l2i[k] - 2, cj[k] = l2j[k] - 2;
  int nf = c->lm;
  while ((nf > kMgCoarseMax || P.mg.size() == 1) && P.mg.size() < 24) {
    std::vector<std::pair<long, int>> key(nf);
    for (int f = 0; f < nf; ++f) key[f] = {((long)(ci[f] >> 1) << 32) | (long)(cj[f] >> 1), f};
    std::sort(key.begin(), key.end());
    std::vector<int> parent(nf + 1), ni, nj;
    std::vector<int> ch;
    int nc = 0;
    for (int a = 0; a < nf;) {
      int b = a;
      while (b < nf && key[b].first == key[a].first) ++b;
      ni.push_back((int)(key[a].first >> 32));
      nj.push_back((int)(key[a].first & 0xffffffffL));
      for (int s2 = 0; s2 < 4; ++s2) ch.push_back(a + s2 < b ? key[a + s2].second : nf);
      for (int q = a; q < b; ++q) parent[key[q].second] = nc;
      ++nc;
      a = b;
    }
    parent[nf] = nc;
    // coarse neighbours through a dense (I,J) table
    int imax = 0, jmax = 0;
    for (int C = 0; C < nc; ++C) imax = std::max(imax, ni[C]), jmax = std::max(jmax, nj[C]);
    std::vector<int> tab((size_t)(imax + 3) * (jmax + 3), nc);
    auto T_ = [&](int I, int J) -> int & { return tab[(size_t)(I + 1) * (jmax + 3) + (J + 1)]; };
    for (int C = 0; C < nc; ++C) T_(ni[C], nj[C]) = C;
    std::vector<int> E(nc + 1, nc), W(nc + 1, nc), N(nc + 1, nc), S(nc + 1, nc);
    for (int C = 0; C < nc; ++C) {
      E[C] = T_(ni[C] + 1, nj[C]);
      W[C] = T_(ni[C] - 1, nj[C]);
      N[C] = T_(ni[C], nj[C] + 1);
      S[C] = T_(ni[C], nj[C] - 1);
    }
    MgLevel L;
    L.n = nc;
    L.lE = upload_ints(c, E);
    L.lW = upload_ints(c, W);
    L.lN = upload_ints(c, N);
    L.lS = upload_ints(c, S);
    L.ch = upload_ints(c, ch);
    L.parent_of_fine = upload_ints(c, parent);
    if (fp32) {
      L.rr = mg_alloc<float>(c, nc, 1.0);
      L.sx = mg_alloc<float>(c, nc, 0.0);
      L.sy = mg_alloc<float>(c, nc, 0.0);
      L.b = mg_alloc<float>(c, nc, 0.0);
      L.x = mg_alloc<float>(c, nc, 0.0);
    } else {
      L.rr = mg_alloc<double>(c, nc, 1.0);
      L.sx = mg_alloc<double>(c, nc, 0.0);
      L.sy = mg_alloc<double>(c, nc, 0.0);
      L.b = mg_alloc<double>(c, nc, 0.0);
      L.x = mg_alloc<double>(c, nc, 0.0);
    }
    if (!L.lE || !L.ch || !L.parent_of_fine || !L.rr || !L.x) return S3G_ERR_HIP;
    P.mg.push_back(L);
    ci = ni;
    cj = nj;
    nf = nc;
  }
  if (nf > kMgCoarseMax) {
    set_error("multigrid: coarsest level has " + std::to_string(nf) + " cells");
    return S3G_ERR_UNSUPPORTED;
  }
  return hip_check(hipGetLastError(), "mg setup");
}

int pcg_setup(s3g_ctx *c, const int *lN0, const int *l2i, const int *l2j) {
  // North-south lines: maximal runs of consecutive columns c-1 -> c with
  // lN(c-1) = c (lN0 is 0-based, sentinel = lm).
  Pcg &P = c->pcg;
  P.n = c->lm;
  P.nblk = std::max(1, std::min((P.n + kBlock - 1) / kBlock, 512));
  std::vector<int> starts;
  for (int cc = 0; cc < c->lm; ++cc)
    if (cc == 0 || lN0[cc - 1] != cc) starts.push_back(cc);
  starts.push_back(c->lm);
  P.nlines = (int)starts.size() - 1;
  P.line_start = upload_ints(c, starts);
  if (!P.line_start) return S3G_ERR_HIP;
  if (c->prec_kind == S3G_PREC_MG) return mg_setup(c, l2i, l2j, c->precision == S3G_MIXED);
  return 0;
}

// (b,b) and (r,r) of the FP64 true residual from their partials (adjacent
// slots P_BB, P_TR) with one copy and one synchronisation; with `pending` the
// copy was already enqueued (behind the last CG chunk) and only the sync and
// the sums remain.
static bool true_residual_sums(s3g_ctx *c, double &bb, double &tr, bool pending) {
  const int nb = c->pcg.nblk;
  if (!pending)
    hipMemcpyAsync(c->h_scal, c->pcg.part + P_BB * nb, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return false;
  bb = tr = 0.0;
  for (int k = 0; k < nb; ++k) bb += c->h_scal[k];
  for (int k = 0; k < nb; ++k) tr += c->h_scal[nb + k];
  return true;
}

// z = M^-1 r for the line and multigrid preconditioners, with the partial of
// (r, z) written into `slot` (Jacobi is fused into the CG kernels instead).
// `rc` is the residual buffer; with `down0_done` the level-0 restriction has
// already run (fused into k_cg_update_mg).
template <typename T>
static void apply_prec(s3g_ctx *c, CgArgs<T> &A, const T *rc, int slot, bool down0_done) {
  const dim3 b(kBlock);
  if (c->prec_kind == S3G_PREC_LINE) {
    hipLaunchKernelGGL((k_line_solve<T>), dim3(A.nblk), b, 0, c->stream, A, rc, slot);
    return;
  }
  // multigrid V-cycle: level 0 is the CG system itself
  std::vector<MgLevel> &Lv = c->pcg.mg;
  const int nl = (int)Lv.size();
  auto pair = [&](int l) {  // level l (fine) and l+1 (coarse)
    MgPair<T> M{};
    if (l == 0) {
      M.nf = A.n;
      M.fE = A.lE; M.fW = A.lW; M.fN = A.lN; M.fS = A.lS;
      M.frr = A.rr; M.fsx = A.sx; M.fsy = A.sy; M.fb = rc; M.fx = A.z;
    } else {
      M.nf = Lv[l].n;
      M.fE = Lv[l].lE; M.fW = Lv[l].lW; M.fN = Lv[l].lN; M.fS = Lv[l].lS;
      M.frr = (const T *)Lv[l].rr; M.fsx = (const T *)Lv[l].sx; M.fsy = (const T *)Lv[l].sy;
      M.fb = (const T *)Lv[l].b; M.fx = (T *)Lv[l].x;
    }
    M.nc = Lv[l + 1].n;
    M.ch = Lv[l + 1].ch;
    M.parent = Lv[l + 1].parent_of_fine;
    M.crr = (T *)Lv[l + 1].rr; M.csx = (T *)Lv[l + 1].sx; M.csy = (T *)Lv[l + 1].sy;
    M.cb = (T *)Lv[l + 1].b; M.cx = (const T *)Lv[l + 1].x;
    return M;
  };
  auto grid = [&](int n) { return dim3(std::max(1, std::min((n + kBlock - 1) / kBlock, 2048))); };
  // levels 0 .. lc-1 run one grid-wide kernel per level; lc .. nl-1 (each at
  // most mg_small cells) run in one workgroup.  lc >= 1: level 0 is the CG
  // system, whose prolongation also writes the (r,z) partials.
  int lc = 1;
  while (lc < nl - 1 && Lv[lc].n > c->mg_small) ++lc;
  if (nl - lc > kMgMaxLevels) lc = nl - kMgMaxLevels;
  for (int l = down0_done ? 1 : 0; l < lc; ++l) {
    MgPair<T> M = pair(l);
    hipLaunchKernelGGL((k_mg_down<T>), grid(4 * M.nc), b, 0, c->stream, M);
  }
  {
    MgChain<T> Ch{};
    Ch.nlev = nl - lc;
    for (int i = 0; i + 1 < Ch.nlev; ++i) Ch.pr[i] = pair(lc + i);
    const MgLevel &Lc = Lv[nl - 1];
    Ch.nc = Lc.n;
    Ch.sweeps = c->mg_sweeps;
    Ch.cE = Lc.lE; Ch.cW = Lc.lW; Ch.cN = Lc.lN; Ch.cS = Lc.lS;
    Ch.crr = (const T *)Lc.rr; Ch.csx = (const T *)Lc.sx; Ch.csy = (const T *)Lc.sy; Ch.cb = (const T *)Lc.b;
    Ch.cx = (T *)Lc.x;
    hipLaunchKernelGGL((k_mg_chain<T>), dim3(1), dim3(1024), 0, c->stream, Ch);
  }
  for (int l = lc - 1; l >= 0; --l) {
    MgPair<T> M = pair(l);
    if (l == 0)
      hipLaunchKernelGGL((k_mg_up<T>), dim3(A.nblk), b, 0, c->stream, M, A.part, A.nblk, slot);
    else
      hipLaunchKernelGGL((k_mg_up<T>), grid(M.nf), b, 0, c->stream, M, A.part, A.nblk, -1);
  }
}

// Galerkin coarse operators for the current matrix (once per solve).
template <typename T>
static void mg_build(s3g_ctx *c, CgArgs<T> &A) {
  std::vector<MgLevel> &Lv = c->pcg.mg;
  for (int l = 0; l + 1 < (int)Lv.size(); ++l) {
    MgPair<T> M{};
    if (l == 0) {
      M.nf = A.n; M.fE = A.lE; M.fW = A.lW; M.fN = A.lN; M.fS = A.lS;
      M.frr = A.rr; M.fsx = A.sx; M.fsy = A.sy;
    } else {
      M.nf = Lv[l].n; M.fE = Lv[l].lE; M.fW = Lv[l].lW; M.fN = Lv[l].lN; M.fS = Lv[l].lS;
      M.frr = (const T *)Lv[l].rr; M.fsx = (const T *)Lv[l].sx; M.fsy = (const T *)Lv[l].sy;
    }
    M.nc = Lv[l + 1].n;
    M.ch = Lv[l + 1].ch;
    M.parent = Lv[l + 1].parent_of_fine;
    M.crr = (T *)Lv[l + 1].rr; M.csx = (T *)Lv[l + 1].sx; M.csy = (T *)Lv[l + 1].sy;
    hipLaunchKernelGGL((k_mg_galerkin<T>), dim3(std::max(1, std::min((M.nc + kBlock - 1) / kBlock, 2048))),
                       dim3(kBlock), 0, c->stream, M, c->mg_cs);
  }
}

__global__ void k_set_scalar(double *p, double v) { *p = v; }

// Runs CG iterations on CgArgs A (x, r, b set up) until (r,r) <= tol2 (b,b) or
// maxit; returns the number of completed iterations (< 0 on HIP error).
// `tail`, if given, is enqueued behind every chunk of iterations, before the
// iteration count is copied back (the FP64 solve computes its true residual
// there, so a converged solve needs a single synchronisation).
template <typename T, typename Tail>
static int run_cg(s3g_ctx *c, CgArgs<T> &A, int maxit, int hint, Tail tail) {
  const dim3 g(A.nblk), b(kBlock);
  const bool jac = (c->prec_kind == S3G_PREC_JACOBI) || (c->prec_kind == S3G_PREC_LINE && A.nlines == 0);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, c->pcg.tol2, A.tol2);
  const bool mg = c->prec_kind == S3G_PREC_MG;
  // flexible CG with the learned correction stage on iterations 0, period, ...
  const int period = A.flex ? learn_flex_period(c) : 0;
  bool corr_ok = true;
  // the stage on the residual rc entering an iteration whose (r,z) partial
  // goes to `slot` (rprev: the residual entering the iteration before)
  auto correction = [&](const T *rc, const T *rprev, int slot) {
    LearnOut lo{};
    if (learn_correction(c, rc, sizeof(T) == 4, A.part + (P_RR0 + slot - P_RZ0) * A.nblk, A.nblk, &lo) < 0) {
      corr_ok = false;
      return;
    }
    hipLaunchKernelGGL((k_flex_apply<T>), g, b, 0, c->stream, A, rc, lo.pix, lo.out, lo.sig);
    hipLaunchKernelGGL((k_flex_scale<T>), g, b, 0, c->stream, A, rc);
    if (jac) {
      hipLaunchKernelGGL((k_flex_finish<T, true>), g, b, 0, c->stream, A, rc, rprev, slot);
    } else {
      apply_prec(c, A, (const T *)A.ry, P_ZR, false);  // its (r - A y, z) partial is replaced below
      hipLaunchKernelGGL((k_flex_finish<T, false>), g, b, 0, c->stream, A, rc, rprev, slot);
    }
  };
  bool capturing = false;  // the graph of iterations 2 and 3 never holds the stage
  auto corrects = [&](int i) { return !capturing && period > 0 && i % period == 0; };  // iteration i starts with it
  if (!jac) {
    hipLaunchKernelGGL((k_cg_residual<T, false>), g, b, 0, c->stream, A);
    if (corrects(0))
      correction((const T *)A.r, (const T *)nullptr, P_RZ0);
    else
      apply_prec(c, A, (const T *)A.r, P_RZ0, false);
  } else {
    hipLaunchKernelGGL((k_cg_residual<T, true>), g, b, 0, c->stream, A);
    if (corrects(0)) correction((const T *)A.r, (const T *)nullptr, P_RZ0);
  }
  MgPair<T> M0{};
  if (mg) {  // level 0 -> 1 of the hierarchy; fb is recomputed inside the fused kernel
    const MgLevel &L1 = c->pcg.mg[1];
    M0.nf = A.n;
    M0.fE = A.lE; M0.fW = A.lW; M0.fN = A.lN; M0.fS = A.lS;
    M0.frr = A.rr; M0.fsx = A.sx; M0.fsy = A.sy;
    M0.nc = L1.n; M0.ch = L1.ch; M0.parent = L1.parent_of_fine;
    M0.cb = (T *)L1.b;
  }
  auto iteration = [&](int i) {
    const T *rnext = ((i + 1) & 1) ? A.r1 : A.r;
    const T *rcur = (i & 1) ? A.r1 : A.r;
    hipLaunchKernelGGL((k_cg_dir<T>), g, b, 0, c->stream, A, i);
    if (corrects(i + 1)) {
      hipLaunchKernelGGL((k_cg_update<T, false>), g, b, 0, c->stream, A, i);
      correction(rnext, rcur, P_RZ0 + ((i + 1) & 1));
      return;
    }
    if (mg) {
      hipLaunchKernelGGL((k_cg_update_mg<T>), g, b, 0, c->stream, A, M0, i);
      apply_prec(c, A, rnext, P_RZ0 + ((i + 1) & 1), true);
    } else if (!jac) {
      hipLaunchKernelGGL((k_cg_update<T, false>), g, b, 0, c->stream, A, i);
      apply_prec(c, A, rnext, P_RZ0 + ((i + 1) & 1), false);
    } else {
      hipLaunchKernelGGL((k_cg_update<T, true>), g, b, 0, c->stream, A, i);
    }
    if (A.flex) hipLaunchKernelGGL((k_flex_zr<T>), g, b, 0, c->stream, A, rcur);
  };
  // The kernels depend on the iteration index only through i & 1 and i > 0,
  // so iterations (2m, 2m+1), m >= 1, are one launch of a graph of
  // iterations 2 and 3.  The graph is captured on first use and again when
  // the arguments change (a new tolerance of the mixed solve's inner CG).
  hipGraphExec_t pair = nullptr;
  if (c->cg_graph) {
    const int gi = sizeof(T) == 8 ? 0 : 1;
    std::vector<char> key(sizeof(A) + sizeof(M0));
    CgArgs<T> Ak = A;
    Ak.tol2 = 0.0;  // read from the device copy: not part of the graph
    std::memcpy(key.data(), &Ak, sizeof(Ak));
    std::memcpy(key.data() + sizeof(A), &M0, sizeof(M0));
    if (!c->cg_exec[gi] || key != c->cg_key[gi]) {
      if (c->cg_exec[gi]) hipGraphExecDestroy(c->cg_exec[gi]);
      c->cg_exec[gi] = nullptr;
      hipGraph_t gr = nullptr;
      bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        capturing = true;
        iteration(2);
        iteration(3);
        capturing = false;
        ok = hipStreamEndCapture(c->stream, &gr) == hipSuccess && gr;
      }
      if (ok) ok = hipGraphInstantiate(&c->cg_exec[gi], gr, nullptr, nullptr, 0) == hipSuccess;
      if (gr) hipGraphDestroy(gr);
      if (!ok) {  // launch kernel by kernel from now on
        (void)hipGetLastError();
        c->cg_exec[gi] = nullptr;
        c->cg_graph = false;
        if (c->debug) std::fprintf(stderr, "[s3g] CG graph capture failed; launching iterations one by one\n");
      }
      c->cg_key[gi] = key;
    }
    pair = c->cg_exec[gi];
  }
  int it = 0, slot = -1;
  int chunk = std::max(4, std::min(hint + 2, 512));
  while (it < maxit) {
    int nl = std::min(chunk, maxit - it);
    // a chunk ends before an iteration that runs the correction stage, so
    // that the host sees convergence before it launches the network
    if (period > 0) nl = std::min(nl, ((it + 1) / period + 1) * period - 1 - it);
    if (c->timing) timer_begin(c, sizeof(T) == 8 ? "pcg_chunk_f64" : "pcg_chunk_f32", &slot);
    int j = 0;
    while (j < nl) {
      const int i = it + j;
      if (pair && i >= 2 && (i & 1) == 0 && j + 1 < nl && !corrects(i + 1) && !corrects(i + 2)) {
        hipGraphLaunch(pair, c->stream);
        j += 2;
      } else {
        iteration(i);
        j += 1;
      }
    }
    if (c->timing) timer_end(c, slot);
    if (!corr_ok) return -1;
    it += nl;
    tail();
    hipMemcpyAsync(c->h_iters, A.iters, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (hip_check(hipStreamSynchronize(c->stream), "pcg sync") < 0) return -1;
    if (c->h_iters[0] < it) break;  // fewer updates than launched: converged
    chunk = 4;
  }
  return c->h_iters[0];
}

template <typename T>
static CgArgs<T> cg_args(s3g_ctx *c, const int *lE, const int *lW, const int *lN, const int *lS, const T *rr,
                         const T *sx, const T *sy, T *x, T *r, T *r1, T *z, T *p0, T *p1, T *q, const T *b,
                         const T *lc,
                         const T *ld, double tol) {
  Pcg &P = c->pcg;
  CgArgs<T> A{};
  A.lE = lE;
  A.lW = lW;
  A.lN = lN;
  A.lS = lS;
  A.rr = rr;
  A.sx = sx;
  A.sy = sy;
  A.n = P.n;
  A.nblk = P.nblk;
  A.x = x;
  A.r = r;
  A.r1 = r1;
  A.z = z;
  A.p0 = p0;
  A.p1 = p1;
  A.q = q;
  A.b = b;
  A.part = P.part;
  A.iters = P.iters;
  A.tol2 = tol * tol;
  A.tol2p = P.tol2;
  A.nlines = P.nlines;
  A.line_start = P.line_start;
  A.lc = lc;
  A.ld = ld;
  return A;
}

// Ensembles: enqueue the per-lake residual sums and their copy to the host.
static int lake_residual_enqueue(s3g_ctx *c, const Outer &O) {
  Pcg &P = c->pcg;
  const int nl = c->nlakes;
  if (!P.lake_off) {
    if (hipMalloc(&P.lake_off, sizeof(int) * (nl + 1)) != hipSuccess ||
        hipMalloc(&P.lake_part, sizeof(double) * 2 * kLakeSlices * nl) != hipSuccess ||
        hipHostMalloc((void **)&P.h_lake, sizeof(double) * 2 * kLakeSlices * nl, hipHostMallocDefault) != hipSuccess)
      return hip_check(hipErrorOutOfMemory, "ensemble residual buffers");
    hipMemcpy(P.lake_off, c->lake_off.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice);
  }
  hipLaunchKernelGGL(k_lake_residual, dim3(kLakeSlices, nl), dim3(kBlock), 0, c->stream, O, P.lake_off, P.lake_part);
  hipMemcpyAsync(P.h_lake, P.lake_part, sizeof(double) * 2 * kLakeSlices * nl, hipMemcpyDeviceToHost, c->stream);
  return 0;
}

// After the stream synchronised: the largest relative true residual of the
// lakes (the absolute one for a lake with a zero right-hand side).
static double lake_relres(const s3g_ctx *c) {
  double worst = 0.0;
  for (int i = 0; i < c->nlakes; ++i) {
    double rr = 0.0, bb = 0.0;
    for (int k = 0; k < kLakeSlices; ++k) {
      rr += c->pcg.h_lake[(i * kLakeSlices + k) * 2];
      bb += c->pcg.h_lake[(i * kLakeSlices + k) * 2 + 1];
    }
    const double r = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    worst = std::isfinite(r) ? std::max(worst, r) : r;
    if (!std::isfinite(r)) break;
  }
  return worst;
}

// Buffers of the learned correction stage (y and r - A y), zeroed, kept
// until the context is destroyed.
static int flex_buffers(s3g_ctx *c) {
  Pcg &P = c->pcg;
  const bool fp32 = c->precision == S3G_MIXED;
  if (fp32 ? (P.ycf != nullptr) : (P.yc != nullptr)) return 0;
  const size_t bytes = (size_t)(P.n + 1) * (fp32 ? sizeof(float) : sizeof(double));
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    if (a) hipFree(a);
    return hip_check(hipErrorOutOfMemory, "correction stage buffers");
  }
  hipMemsetAsync(a, 0, bytes, c->stream);
  hipMemsetAsync(b, 0, bytes, c->stream);
  c->allocs.push_back(a);
  c->allocs.push_back(b);
  if (fp32)
    P.ycf = (float *)a, P.ryf = (float *)b;
  else
    P.yc = (double *)a, P.ry = (double *)b;
  return 0;
}

template <typename R>
int solve_free_surface(s3g_ctx *c, double tz) {
  Pcg &P = c->pcg;
  Dev<R> &D = *reinterpret_cast<Dev<R> *>(sizeof(R) == 8 ? (void *)&c->d64 : (void *)&c->d32);
  const bool line = (c->prec_kind == S3G_PREC_LINE) && P.nlines > 0;
  const double tol = c->pcg_tol;
  const dim3 g(P.nblk), b(kBlock);
  const dim3 gl(std::max(1, (P.nlines + kBlock - 1) / kBlock));
  int rc = 0, total = 0;
  double relres = 0.0;
  Outer O{D.lE, D.lW, D.lN, D.lS, D.rr, D.sx, D.sy, D.qq, P.x, P.n, P.nblk, P.part};
  // size the first chunk of iterations from the latest solve of the same kind
  const int slot = std::min(std::max(c->solve_slot, 0), 3);
  const int hint = c->cg_hint[slot] > 0 ? c->cg_hint[slot] : c->stats.pcg_iters_last;
  // Initial guess: the latest free surface s.  The reference starts from sp
  // (1970); s is the same field on leapfrog steps and the previous estimate of
  // zeta^{n+1} on trapezoidal iterations, so it only shortens the iteration.
  const int extrap = (c->pcg_guess == 1 && tz == 1.0) ? 1 : 0;
  hipLaunchKernelGGL(k_start_guess, g, b, 0, c->stream, P.x, (const double *)D.s, (const double *)D.sp,
                     (const double *)D.spp, P.n, extrap);
  if (c->learn && learn_before_solve(c, &O, P.x) < 0) return S3G_ERR_HIP;
  const bool flex = c->learn && learn_flex_period(c) > 0;
  if (flex && flex_buffers(c) < 0) return S3G_ERR_HIP;

  if (c->precision == S3G_FP64) {
    CgArgs<double> A = cg_args<double>(c, D.lE, D.lW, D.lN, D.lS, D.rr, D.sx, D.sy, P.x, P.r, P.r1, P.z, P.p0, P.p1, P.q,
                                       D.qq, P.lc, P.ld, 0.5 * tol);
    if (flex) A.flex = 1, A.yc = P.yc, A.ry = P.ry;
    if (line) hipLaunchKernelGGL((k_line_factor<double>), gl, b, 0, c->stream, A, P.lc, P.ld);
    if (c->prec_kind == S3G_PREC_MG) mg_build(c, A);
    const bool lakes = c->nlakes > 1;
    auto true_res = [&]() {  // behind each chunk: FP64 true residual, partials to the host
      hipLaunchKernelGGL((k_true_residual<double, false>), g, b, 0, c->stream, O, nullptr, nullptr, nullptr,
                         nullptr, nullptr);
      hipMemcpyAsync(c->h_scal, P.part + P_BB * P.nblk, 2 * P.nblk * sizeof(double), hipMemcpyDeviceToHost,
                     c->stream);
      if (lakes) lake_residual_enqueue(c, O);
    };
    for (int restart = 0; restart < 3; ++restart) {
      int it = run_cg(c, A, c->pcg_maxit - total, hint, true_res);
      if (it < 0) return S3G_ERR_HIP;
      total += it;
      double tr = 0.0, bb = 0.0;
      if (!true_residual_sums(c, bb, tr, true)) return S3G_ERR_HIP;
      relres = bb > 0 ? std::sqrt(tr / bb) : std::sqrt(tr);
      if (lakes && std::isfinite(relres)) relres = lake_relres(c);
      if (relres <= tol || !std::isfinite(relres)) break;
      if (total >= c->pcg_maxit) {
        rc = S3G_WARN_ITMAX;
        break;
      }
    }
  } else {
    // iterative refinement with an FP32 inner CG
    hipLaunchKernelGGL((k_cast_matrix<float>), g, b, 0, c->stream, (const double *)D.rr, (const double *)D.sx,
                       (const double *)D.sy, P.rrf, P.sxf, P.syf, P.n);
    CgArgs<float> A = cg_args<float>(c, D.lE, D.lW, D.lN, D.lS, P.rrf, P.sxf, P.syf, P.df, P.rf, P.rf1, P.zf, P.pf0, P.pf1,
                                     P.qf, P.bf, P.lcf, P.ldf, kInnerTol);
    if (flex) A.flex = 1, A.yc = P.ycf, A.ry = P.ryf;
    if (line) hipLaunchKernelGGL((k_line_factor<float>), gl, b, 0, c->stream, A, P.lcf, P.ldf);
    if (c->prec_kind == S3G_PREC_MG) mg_build(c, A);
    for (int sweep = 0; sweep < 12; ++sweep) {
      hipLaunchKernelGGL((k_true_residual<float, true>), g, b, 0, c->stream, O, P.bf, P.df, P.pf0, P.pf1, P.zf);
      if (c->nlakes > 1 && lake_residual_enqueue(c, O) < 0) return S3G_ERR_HIP;
      double tr = 0.0, bb = 0.0;
      if (!true_residual_sums(c, bb, tr, false)) return S3G_ERR_HIP;
      relres = bb > 0 ? std::sqrt(tr / bb) : std::sqrt(tr);
      if (c->nlakes > 1 && std::isfinite(relres)) relres = lake_relres(c);
      if (relres <= tol || !std::isfinite(relres)) break;
      if (total >= c->pcg_maxit) {
        rc = S3G_WARN_ITMAX;
        break;
      }
      // the inner solve only needs to close the gap to tol (with a margin),
      // and FP32 holds about 1e-5 of reduction per sweep
      const double t_in = std::min(0.5, std::max(kInnerTol, 0.3 * tol / relres));
      A.tol2 = t_in * t_in;
      int it = run_cg(c, A, c->pcg_maxit - total, hint / 2, [] {});
      if (it < 0) return S3G_ERR_HIP;
      total += it;
      hipLaunchKernelGGL((k_add_correction<float>), g, b, 0, c->stream, P.x, (const float *)P.df, P.n);
    }
  }
  if (!std::isfinite(relres)) {
    set_error("free-surface solve produced a non-finite residual");
    return S3G_ERR_SOLVER;
  }
  if (c->learn && learn_after_solve(c, &O) < 0) return S3G_ERR_HIP;
  hipLaunchKernelGGL(k_copy_to_s, dim3((P.n + 255) / 256), b, 0, c->stream, D.s, P.x, P.n);
  if (c->debug)
    std::fprintf(stderr, "[s3g] solve %ld: %s %s iters=%d relres=%.3e rc=%d\n", c->stats.solves,
                 c->precision == S3G_FP64 ? "fp64" : "mixed",
                 c->prec_kind == S3G_PREC_MG ? "mg" : (line ? "line" : "jacobi"), total, relres, rc);
  c->stats.pcg_iters_last = total;
  c->cg_hint[slot] = total;
  c->stats.pcg_iters_total += total;
  c->stats.pcg_relres_last = relres;
  c->stats.solves += 1;
  return rc;
}

template int solve_free_surface<double>(s3g_ctx *, double);
template int solve_free_surface<float>(s3g_ctx *, double);

}  // namespace s3g
