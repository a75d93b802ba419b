// s3g_pcg.hip -- free-surface solve A zeta = qq on gfx950.
//
// Replaces SolverSparse's NSPCG call (si3d_procedures.f90:1870-2065).  A is the
// pentadiagonal SPD matrix assembled by matcon: row c holds rr(c) on the
// diagonal, -sx(W) and -sy(S) for the lower neighbours and -sy(c), -sx(c) for
// the upper ones (1954-1967); it is applied matrix-free from rr/sx/sy and the
// neighbour index arrays.  The initial guess is sp (1970).
//
// Conjugate gradients with FP64 vectors and dots.  Every reduction is a
// deterministic two-level sum: each workgroup writes one partial (64-lane
// wavefront shuffles, then LDS across the four wavefronts), and the next kernel
// re-reduces the partials in a fixed order, so all workgroups derive the same
// alpha/beta and the same convergence decision without a host round trip.
// Iterations are enqueued in chunks; once converged, the remaining kernels of a
// chunk return at entry.  The stopping test is the FP64 relative residual
// ||qq - A zeta|| <= tol ||qq||, confirmed by a true-residual recompute.
//
// Preconditioners: Jacobi, or an exact tridiagonal solve along each north-south
// line of wet cells (the wet cells of one i-row are consecutive in l, so a line
// is a contiguous range of columns: one lane per line runs the Thomas sweep;
// its factors are computed once per matrix).
//
// Mixed precision (S3G_MIXED): iterative refinement.  The outer loop computes
// the true residual in FP64; an inner CG in FP32 (FP32 SpMV and vectors, FP64
// dots) solves the correction to a loose tolerance; zeta is updated in FP64.
#include <cmath>

#include "s3g_internal.h"

namespace s3g {

#define SIX(c) (c)

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

// Deterministic reduction of nblk partials; every block gets the same value.
__device__ double reduce_partials(const double *part, int nblk, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) lds[8] = s;
  __syncthreads();
  double r = lds[8];
  __syncthreads();
  return r;
}

struct Mat {
  const int *lE, *lW, *lN, *lS;
  const double *rr, *sx, *sy;
  int n, S;
};

template <typename T>
__device__ __forceinline__ double apply_A(const Mat &M, const T *v, int c) {
  const int lW = M.lW[c], lS = M.lS[c], lN = M.lN[c], lE = M.lE[c];
  double acc = M.rr[c] * (double)v[c];
  acc -= M.sx[lW] * (double)v[lW];  // sx(S)=0 and v(S)=0 at the sentinel
  acc -= M.sy[lS] * (double)v[lS];
  acc -= M.sy[c] * (double)v[lN];
  acc -= M.sx[c] * (double)v[lE];
  return acc;
}

// Partial sums layout in Pcg::part: [RZ0 | RZ1 | RR0 | RR1 | PQ | BB | TR] x nblk
enum { P_RZ0 = 0, P_RZ1 = 1, P_RR0 = 2, P_RR1 = 3, P_PQ = 4, P_BB = 5, P_TR = 6, P_NUM = 7 };

struct PcgArgs {
  Mat M;
  int nblk;
  double *x, *r, *z, *p0, *p1, *q, *b;
  const double *diag;
  double *part;
  int *iters;
  double tol2;  // (0.5*tol)^2
  // line preconditioner
  int nlines;
  const int *line_start;
  const double *lc, *ld;
};

// z = M^-1 r for one column (Jacobi) -- the line preconditioner runs separately.
__device__ __forceinline__ double jacobi(const PcgArgs &A, int c, double r) { return r / A.diag[c]; }

__global__ void __launch_bounds__(kBlock) k_pcg_start(PcgArgs A, const double *sp) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= A.M.n; c += gridDim.x * blockDim.x) {
    A.x[c] = (c < A.M.n) ? sp[c] : 0.0;
    A.p0[c] = A.p1[c] = A.z[c] = A.r[c] = A.q[c] = 0.0;
  }
}

// r = b - A x; z = M^-1 r (Jacobi); partials of (r,z), (r,r), (b,b).
template <bool JACOBI>
__global__ void __launch_bounds__(kBlock) k_pcg_residual(PcgArgs A) {
  __shared__ double lds[16];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.M.n; c += gridDim.x * blockDim.x) {
    const double b = A.b[c];
    const double r = b - apply_A(A.M, A.x, c);
    A.r[c] = r;
    if (JACOBI) {
      const double z = jacobi(A, c, r);
      A.z[c] = z;
      rz += r * z;
    }
    rr += r * r;
    bb += b * b;
  }
  rz = block_sum(rz, lds);
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    A.part[P_RZ0 * A.nblk + blockIdx.x] = rz;
    A.part[P_RR0 * A.nblk + blockIdx.x] = rr;
    A.part[P_BB * A.nblk + blockIdx.x] = bb;
    A.part[P_TR * A.nblk + blockIdx.x] = rr;
  }
}

// Line preconditioner: z = T^-1 r with T the tridiagonal part of A along each
// north-south line (diagonal rr, off-diagonals -sy).  Factors: ld = 1/pivot,
// lc = modified super-diagonal.  One lane per line.  SLOT >= 0 also writes the
// partial of (r,z) for that slot.
template <int SLOT>
__global__ void __launch_bounds__(kBlock) k_line_solve(PcgArgs A) {
  __shared__ double lds[16];
  double rz = 0.0;
  for (int ln = blockIdx.x * blockDim.x + threadIdx.x; ln < A.nlines; ln += gridDim.x * blockDim.x) {
    const int c0 = A.line_start[ln], c1 = A.line_start[ln + 1];
    // forward: y(c) = (r(c) + sy(c-1) y(c-1)) * ld(c)
    double y = 0.0;
    for (int c = c0; c < c1; ++c) {
      const double off = (c > c0) ? A.M.sy[c - 1] : 0.0;
      y = (A.r[c] + off * y) * A.ld[c];
      A.z[c] = y;
    }
    // backward: z(c) = y(c) - lc(c) z(c+1)
    double zn = 0.0;
    for (int c = c1 - 1; c >= c0; --c) {
      const double zc = (c < c1 - 1) ? A.z[c] - A.lc[c] * zn : A.z[c];
      A.z[c] = zc;
      zn = zc;
      if (SLOT >= 0) rz += A.r[c] * zc;
    }
  }
  if (SLOT >= 0) {
    rz = block_sum(rz, lds);
    if (threadIdx.x == 0) A.part[SLOT * A.nblk + blockIdx.x] = rz;
  }
}

// Zero the partials a line-solve block range does not cover.
__global__ void k_zero_part(double *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Iteration i, step 1: beta from (r,z) partials, p_new = z + beta p_old at the
// column and its neighbours, q = A p_new, partial (p,q).
__global__ void __launch_bounds__(kBlock) k_pcg_dir(PcgArgs A, int i) {
  __shared__ double lds[16];
  const int cur = i & 1;
  const double rr = reduce_partials(A.part + (P_RR0 + cur) * A.nblk, A.nblk, lds);
  const double bb = reduce_partials(A.part + P_BB * A.nblk, A.nblk, lds);
  if (rr <= A.tol2 * bb) return;  // converged: every block takes the same branch
  const double rz = reduce_partials(A.part + (P_RZ0 + cur) * A.nblk, A.nblk, lds);
  double beta = 0.0;
  if (i > 0) beta = rz / reduce_partials(A.part + (P_RZ0 + (cur ^ 1)) * A.nblk, A.nblk, lds);
  const double *pold = cur ? A.p0 : A.p1;
  double *pnew = cur ? A.p1 : A.p0;
  double pq = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.M.n; c += gridDim.x * blockDim.x) {
    const int lW = A.M.lW[c], lS = A.M.lS[c], lN = A.M.lN[c], lE = A.M.lE[c];
    auto pn = [&](int j) { return A.z[j] + beta * pold[j]; };  // z(S) = p(S) = 0
    const double pc = pn(c);
    pnew[c] = pc;
    const double q = A.M.rr[c] * pc - A.M.sx[lW] * pn(lW) - A.M.sy[lS] * pn(lS) - A.M.sy[c] * pn(lN) -
                     A.M.sx[c] * pn(lE);
    A.q[c] = q;
    pq += pc * q;
  }
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) A.part[P_PQ * A.nblk + blockIdx.x] = pq;
}

// Iteration i, step 2: alpha, x += alpha p, r -= alpha q, z = M^-1 r (Jacobi),
// partials (r,z) and (r,r) for iteration i+1.
template <bool JACOBI>
__global__ void __launch_bounds__(kBlock) k_pcg_update(PcgArgs A, int i) {
  __shared__ double lds[16];
  const int cur = i & 1, nxt = cur ^ 1;
  const double rr = reduce_partials(A.part + (P_RR0 + cur) * A.nblk, A.nblk, lds);
  const double bb = reduce_partials(A.part + P_BB * A.nblk, A.nblk, lds);
  if (rr <= A.tol2 * bb) return;
  const double rz = reduce_partials(A.part + (P_RZ0 + cur) * A.nblk, A.nblk, lds);
  const double pq = reduce_partials(A.part + P_PQ * A.nblk, A.nblk, lds);
  const double alpha = rz / pq;
  const double *p = cur ? A.p1 : A.p0;
  double rzn = 0.0, rrn = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.M.n; c += gridDim.x * blockDim.x) {
    A.x[c] += alpha * p[c];
    const double r = A.r[c] - alpha * A.q[c];
    A.r[c] = r;
    if (JACOBI) {
      const double z = jacobi(A, c, r);
      A.z[c] = z;
      rzn += r * z;
    }
    rrn += r * r;
  }
  rzn = block_sum(rzn, lds);
  rrn = block_sum(rrn, lds);
  if (threadIdx.x == 0) {
    if (JACOBI) A.part[(P_RZ0 + nxt) * A.nblk + blockIdx.x] = rzn;
    A.part[(P_RR0 + nxt) * A.nblk + blockIdx.x] = rrn;
    if (blockIdx.x == 0) A.iters[0] = i + 1;
  }
}

// True residual partials: ||b - A x||^2 into slot TR.
__global__ void __launch_bounds__(kBlock) k_pcg_true_residual(PcgArgs A) {
  __shared__ double lds[16];
  double rr = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.M.n; c += gridDim.x * blockDim.x) {
    const double r = A.b[c] - apply_A(A.M, A.x, c);
    rr += r * r;
  }
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) A.part[P_TR * A.nblk + blockIdx.x] = rr;
}

// Line factors: for each line, Thomas pivots of tridiag(-sy(c-1), rr(c), -sy(c)).
__global__ void __launch_bounds__(kBlock) k_line_factor(PcgArgs A, double *lc, double *ld) {
  for (int ln = blockIdx.x * blockDim.x + threadIdx.x; ln < A.nlines; ln += gridDim.x * blockDim.x) {
    const int c0 = A.line_start[ln], c1 = A.line_start[ln + 1];
    double cprev = 0.0;
    for (int c = c0; c < c1; ++c) {
      const double a = (c > c0) ? -A.M.sy[c - 1] : 0.0;  // sub-diagonal
      const double piv = A.M.rr[c] - a * cprev;
      const double inv = 1.0 / piv;
      ld[c] = inv;
      const double csup = (c < c1 - 1) ? -A.M.sy[c] * inv : 0.0;
      lc[c] = csup;
      cprev = csup;
    }
  }
}

__global__ void k_copy_to_s(double *s, const double *x, int n) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) s[c] = x[c];
}

// ---------------------------------------------------------------------------

int pcg_setup(s3g_ctx *c, const int *lN0) {
  // lN0: 0-based north neighbour of each column (sentinel = lm).  North-south
  // lines are maximal runs of consecutive columns c-1 -> c with lN(c-1) = c.
  Pcg &P = c->pcg;
  P.n = c->lm;
  P.nblk = std::min((P.n + kBlock - 1) / kBlock, 1024);
  if (P.nblk < 1) P.nblk = 1;
  std::vector<int> starts;
  for (int cc = 0; cc < c->lm; ++cc)
    if (cc == 0 || lN0[cc - 1] != cc) starts.push_back(cc);
  starts.push_back(c->lm);
  P.nlines = (int)starts.size() - 1;
  if (hip_check(hipMalloc(&P.line_start, starts.size() * sizeof(int)), "alloc lines") < 0) return S3G_ERR_HIP;
  c->allocs.push_back(P.line_start);
  return hip_check(hipMemcpy(P.line_start, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice),
                   "upload lines");
}

template <typename R>
int solve_free_surface(s3g_ctx *c, double tz) {
  (void)tz;
  Pcg &P = c->pcg;
  Dev<R> &D = (sizeof(R) == 8) ? *reinterpret_cast<Dev<R> *>(&c->d64) : *reinterpret_cast<Dev<R> *>(&c->d32);
  PcgArgs A{};
  A.M = Mat{D.lE, D.lW, D.lN, D.lS, D.rr, D.sx, D.sy, P.n, c->lm};
  A.nblk = P.nblk;
  A.x = P.x;
  A.r = P.r;
  A.z = P.z;
  A.p0 = P.p0;
  A.p1 = P.p1;
  A.q = P.q;
  A.b = D.qq;
  A.diag = D.rr;
  A.part = P.part;
  A.iters = P.iters;
  const double tol = c->pcg_tol;
  A.tol2 = (0.5 * tol) * (0.5 * tol);
  A.nlines = P.nlines;
  A.line_start = P.line_start;
  A.lc = P.lc;
  A.ld = P.ld;
  const bool line = (c->prec_kind == S3G_PREC_LINE) && P.nlines > 0;
  const dim3 g(P.nblk), b(kBlock);
  const dim3 gl((P.nlines + kBlock - 1) / kBlock > 0 ? (P.nlines + kBlock - 1) / kBlock : 1);
  int slot = -1;
  hipLaunchKernelGGL(k_pcg_start, g, b, 0, c->stream, A, (const double *)D.sp);
  if (line) hipLaunchKernelGGL(k_line_factor, gl, b, 0, c->stream, A, P.lc, P.ld);
  int total = 0, rc = 0;
  double relres = 0.0;
  for (int restart = 0; restart < 3; ++restart) {
    if (line) {
      hipLaunchKernelGGL(k_pcg_residual<false>, g, b, 0, c->stream, A);
      hipLaunchKernelGGL(k_zero_part, dim3((P.nblk + 255) / 256), b, 0, c->stream, A.part + P_RZ0 * P.nblk, P.nblk);
      // line solve writes partials indexed by its own blocks (<= nblk of them)
      hipLaunchKernelGGL(k_line_solve<P_RZ0>, dim3(std::min((int)gl.x, P.nblk)), b, 0, c->stream, A);
    } else {
      hipLaunchKernelGGL(k_pcg_residual<true>, g, b, 0, c->stream, A);
    }
    hipMemsetAsync(P.iters, 0, sizeof(int), c->stream);
    int done = 0, it = 0;
    int chunk = std::max(8, std::min(c->stats.pcg_iters_last + 4, 256));
    while (!done && it < c->pcg_maxit) {
      int nlaunch = std::min(chunk, c->pcg_maxit - it);
      if (c->timing) timer_begin(c, "pcg_iter", &slot);
      for (int j = 0; j < nlaunch; ++j) {
        const int i = it + j;
        hipLaunchKernelGGL(k_pcg_dir, g, b, 0, c->stream, A, i);
        if (line) {
          hipLaunchKernelGGL(k_pcg_update<false>, g, b, 0, c->stream, A, i);
          const int nxt = (i + 1) & 1;
          hipLaunchKernelGGL(k_zero_part, dim3((P.nblk + 255) / 256), b, 0, c->stream,
                             A.part + (P_RZ0 + nxt) * P.nblk, P.nblk);
          if (nxt)
            hipLaunchKernelGGL(k_line_solve<P_RZ1>, dim3(std::min((int)gl.x, P.nblk)), b, 0, c->stream, A);
          else
            hipLaunchKernelGGL(k_line_solve<P_RZ0>, dim3(std::min((int)gl.x, P.nblk)), b, 0, c->stream, A);
        } else {
          hipLaunchKernelGGL(k_pcg_update<true>, g, b, 0, c->stream, A, i);
        }
      }
      if (c->timing) timer_end(c, slot);
      it += nlaunch;
      hipMemcpyAsync(c->h_iters, P.iters, sizeof(int), hipMemcpyDeviceToHost, c->stream);
      if (hip_check(hipStreamSynchronize(c->stream), "pcg sync") < 0) return S3G_ERR_HIP;
      done = (c->h_iters[0] < it);  // fewer completed updates than launched: converged
      chunk = 8;
    }
    total += c->h_iters[0];
    hipLaunchKernelGGL(k_pcg_true_residual, g, b, 0, c->stream, A);
    std::vector<double> h(2 * P.nblk);
    hipMemcpyAsync(c->h_scal, A.part + P_TR * P.nblk, P.nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (hip_check(hipStreamSynchronize(c->stream), "pcg residual") < 0) return S3G_ERR_HIP;
    double tr = 0.0;
    for (int k = 0; k < P.nblk; ++k) tr += c->h_scal[k];
    hipMemcpyAsync(c->h_scal, A.part + P_BB * P.nblk, P.nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (hip_check(hipStreamSynchronize(c->stream), "pcg bb") < 0) return S3G_ERR_HIP;
    double bb = 0.0;
    for (int k = 0; k < P.nblk; ++k) bb += c->h_scal[k];
    relres = (bb > 0) ? std::sqrt(tr / bb) : std::sqrt(tr);
    if (relres <= tol || !std::isfinite(relres)) break;
    if (it >= c->pcg_maxit) {
      rc = S3G_WARN_ITMAX;
      break;
    }
  }
  if (!std::isfinite(relres)) {
    set_error("free-surface solve produced a non-finite residual");
    return S3G_ERR_SOLVER;
  }
  hipLaunchKernelGGL(k_copy_to_s, dim3((P.n + 255) / 256), b, 0, c->stream, D.s, P.x, P.n);
  c->stats.pcg_iters_last = total;
  c->stats.pcg_iters_total += total;
  c->stats.pcg_relres_last = relres;
  c->stats.solves += 1;
  return rc;
}

template int solve_free_surface<double>(s3g_ctx *, double);
template int solve_free_surface<float>(s3g_ctx *, double);

}  // namespace s3g
