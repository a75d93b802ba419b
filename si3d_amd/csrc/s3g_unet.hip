// s3g_unet.hip -- learned free-surface solution operator for the zeta solve.
//
// A 4-level convolutional U-Net (base 32 channels, 3x3 convolutions + ReLU,
// 2x2 average pooling, nearest-neighbour upsampling with skip concatenation,
// 1x1 output) maps the normalised residual of the plain initial guess,
// r0 / sigma with r0 = qq - A zeta_guess and sigma = rms(r0), together with the
// coupling field (rr-1)/rr and the wet mask, to the normalised correction
// (zeta - zeta_guess) / sigma.  The system is linear in qq, so the normalised
// map does not depend on the forcing amplitude.  It is trained on pairs that
// the plain solver itself records (s3g_learn_setup), in the energy norm of
// each pair's own matrix (the norm CG minimises, so that high-frequency
// output noise, which A amplifies, is penalised), by the forward and
// backward kernels below, and used as the initial guess of the free-surface
// CG, which still converges to the same FP64 true-residual tolerance: the
// plain CG is the fallback and decides when the error is small enough.
//
// Tensors are NHWC bf16 on the lake's (i, j) box (H x W padded to multiples
// of 8), channels contiguous.  Convolutions are implicit GEMMs on the gfx950
// matrix cores: v_mfma_f32_32x32x16_bf16, one wavefront per 32 pixels x 32
// output channels, FP32 accumulation; bias, ReLU and the output cast are
// fused.  The input gradient is the same kernel on flipped, transposed
// weights with the producer's ReLU mask fused; the weight gradient is an MFMA
// GEMM over pixels with per-slice partials reduced in a fixed order.  Master
// weights, gradients and Adam moments are FP32.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "s3g_internal.h"

namespace s3g {

typedef unsigned short bf16;
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float((unsigned)v << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
static inline bf16 f2bf_host(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

// A view of a channel slice of an NHWC tensor: element (p, c) at
// ptr[p * stride + off + c].
struct TView {
  bf16 *ptr;
  int stride, off;
};

// ---------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding.
//   out(p, co) = epilogue( bias(co) + sum_{t, ci} in(p + d(t), ci) * Wt(t, co, ci) )
// Taps t = 3 * (dy + 1) + (dx + 1).  Wt is [9][Cout][Cin] (ci contiguous).
// Epilogue: RELU -> max(0, .); MASK -> 0 where mask(p, co) <= 0 (ReLU mask of
// the tensor whose gradient this is).  Cin % 16 == 0, Cout % 32 == 0.
// Workgroup: 4 wavefronts = 128 consecutive pixels x 32 output channels.
// MFMA 32x32x16: lane l holds A[m = l%32][k = 8(l/32) .. +7] (8 channels of
// one shifted pixel, 16 bytes), B[k = 8(l/32) .. +7][n = l%32] (8 input
// channels of one output channel), and D[m][n] for n = l%32,
// m = 8(i/4) + 4(l/32) + i%4, i = 0..15.
template <bool RELU, bool MASK>
__global__ void __launch_bounds__(256) k_conv3x3(TView in, int Cin, const bf16 *__restrict__ Wt,
                                                 const float *__restrict__ bias, int Cout, TView out, TView mask,
                                                 int H, int W) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 31, kg = lane >> 5;
  const int npx = H * W;
  const int p0 = (blockIdx.x * 4 + wv) * 32;
  if (p0 >= npx) return;
  const int co0 = blockIdx.y * 32;
  const int p = p0 + m;
  const bool pin = p < npx;
  const int ph = pin ? p / W : 0, pw = pin ? p - ph * W : 0;
  f32x16 acc = {};
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int qh = ph + dy, qw = pw + dx;
    const bool ok = pin && qh >= 0 && qh < H && qw >= 0 && qw < W;
    const bf16 *a_ptr = in.ptr + (size_t)(ok ? qh * W + qw : 0) * in.stride + in.off + 8 * kg;
    const bf16 *b_ptr = Wt + ((size_t)t * Cout + co0 + m) * Cin + 8 * kg;
    for (int ci = 0; ci < Cin; ci += 16) {
      u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
      if (ok) a = *(const u16x8 *)(a_ptr + ci);
      const u16x8 b = *(const u16x8 *)(b_ptr + ci);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                    acc, 0, 0, 0);
    }
  }
  const int co = co0 + m;
  const float bc = bias ? bias[co] : 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int pr = p0 + 8 * (i >> 2) + 4 * kg + (i & 3);
    if (pr >= npx) continue;
    float v = acc[i] + bc;
    if (RELU) v = fmaxf(v, 0.f);
    if (MASK && !(bf2f(mask.ptr[(size_t)pr * mask.stride + mask.off + co]) > 0.f)) v = 0.f;
    out.ptr[(size_t)pr * out.stride + out.off + co] = f2bf(v);
  }
}

// The same convolution with both operands staged in LDS, for images whose
// width is a multiple of 256.  Workgroup: 4 wavefronts and an output tile of
// 256 consecutive pixels of one image row x TC output channels; each
// wavefront owns 64 pixels x TC channels, i.e. 2 x TC/32 accumulators, so an
// A fragment (8 channels of one shifted pixel) read from LDS feeds TC/32
// MFMAs and a B fragment (8 input channels of one output channel) feeds 2.
// Per chunk of 16 input channels the workgroup stages image rows h-1 .. h+1,
// pixels w0-1 .. w0+256 (zeros outside the image) and the chunk of all nine
// taps of its TC output channels; every fragment read is 64 lanes x 16 bytes
// of one contiguous kilobyte.  The FP32 sums run over chunks, then taps, then
// channels (k_conv3x3: taps, then channels), so outputs agree with
// k_conv3x3 up to the rounding of single bf16 values.
// Requires W % 256 == 0, Cin % 16 == 0, Cout % TC == 0.
template <bool RELU, bool MASK, int TC>
__global__ void __launch_bounds__(256) k_conv3x3_lds(TView in, int Cin, const bf16 *__restrict__ Wt,
                                                     const float *__restrict__ bias, int Cout, TView out,
                                                     TView mask, int H, int W) {
  constexpr int TP = 256, PW = TP + 2, NT = TC / 32;
  __shared__ u16x8 sA[3 * PW * 2];  // [row 0..2][pixel 0..PW-1][channel half]
  __shared__ u16x8 sB[9 * TC * 2];  // [tap][output channel][input channel half]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, m = lane & 31, kg = lane >> 5;
  const int tiles_per_row = W / TP;
  const int h = blockIdx.x / tiles_per_row, w0 = (blockIdx.x - h * tiles_per_row) * TP;
  const int co0 = blockIdx.y * TC;
  f32x16 acc[2][NT];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x16{};
  for (int ci0 = 0; ci0 < Cin; ci0 += 16) {
    __syncthreads();  // the previous chunk's fragments have been read
    for (int v = tid; v < 3 * PW * 2; v += 256) {
      const int half = v & 1, q = v >> 1, r = q / PW, px = q - r * PW;
      const int qh = h + r - 1, qw = w0 + px - 1;
      u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
      if (qh >= 0 && qh < H && qw >= 0 && qw < W)
        a = *(const u16x8 *)(in.ptr + (size_t)(qh * W + qw) * in.stride + in.off + ci0 + 8 * half);
      sA[v] = a;
    }
    for (int v = tid; v < 9 * TC * 2; v += 256) {
      const int half = v & 1, q = v >> 1, t = q / TC, co = q - t * TC;
      sB[v] = *(const u16x8 *)(Wt + ((size_t)t * Cout + co0 + co) * Cin + ci0 + 8 * half);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int r = t / 3, dx = t - 3 * r;  // image row h + r - 1, pixel offset dx - 1
      bf16x8 b[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) b[j] = __builtin_bit_cast(bf16x8, sB[(t * TC + 32 * j + m) * 2 + kg]);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const bf16x8 a = __builtin_bit_cast(bf16x8, sA[(r * PW + wv * 64 + 32 * i + m + dx) * 2 + kg]);
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  const int prow = h * W + w0 + wv * 64;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = co0 + 32 * j + m;
    const float bc = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int pr = prow + 32 * i + 8 * (e >> 2) + 4 * kg + (e & 3);
        float v = acc[i][j][e] + bc;
        if (RELU) v = fmaxf(v, 0.f);
        if (MASK && !(bf2f(mask.ptr[(size_t)pr * mask.stride + mask.off + co]) > 0.f)) v = 0.f;
        out.ptr[(size_t)pr * out.stride + out.off + co] = f2bf(v);
      }
  }
}

// 3x3 convolution: the LDS-staged kernel where the image width allows it,
// the direct one otherwise.
template <bool RELU, bool MASK>
static void conv3x3(hipStream_t s, TView in, int Cin, const bf16 *Wt, const float *bias, int Cout, TView out,
                    TView mask, int H, int W) {
  const int npx = H * W;
  static const bool direct = [] {
    const char *e = std::getenv("S3G_CONV");
    return e && std::strcmp(e, "direct") == 0;
  }();
  if (!direct && W % 256 == 0 && Cin % 16 == 0 && Cout % 32 == 0) {
    if (Cout % 64 == 0)
      hipLaunchKernelGGL((k_conv3x3_lds<RELU, MASK, 64>), dim3(npx / 256, Cout / 64), dim3(256), 0, s, in, Cin, Wt,
                         bias, Cout, out, mask, H, W);
    else
      hipLaunchKernelGGL((k_conv3x3_lds<RELU, MASK, 32>), dim3(npx / 256, Cout / 32), dim3(256), 0, s, in, Cin, Wt,
                         bias, Cout, out, mask, H, W);
    return;
  }
  hipLaunchKernelGGL((k_conv3x3<RELU, MASK>), dim3((npx + 127) / 128, Cout / 32), dim3(256), 0, s, in, Cin, Wt, bias,
                     Cout, out, mask, H, W);
}

// Weight gradient of a 3x3 convolution:
//   dW(t, co, ci) = sum_p dy(p, co) * x(p + d(t), ci)
// as an MFMA GEMM with M = co, N = ci, K = pixels.  Grid: (slices,
// 9 * Cout/32 * ceil(Cin/32)); one wavefront per workgroup walks its pixel
// slice 16 pixels per MFMA and writes its partial tile to
// part[slice][t][co][ci]; k_sum_slices reduces the slices in order.
__global__ void __launch_bounds__(64) k_conv3x3_dw(TView dy, int Cout, TView x, int Cin, int H, int W,
                                                  int px_per_slice, float *__restrict__ part) {
  const int lane = threadIdx.x;
  const int m = lane & 31, kg = lane >> 5;
  const int nci = (Cin + 31) / 32, nco = Cout / 32;
  int r = blockIdx.y;
  const int t = r / (nco * nci);
  r -= t * nco * nci;
  const int cb = r / nci, ib = r - cb * nci;
  const int co = cb * 32 + m, ci = ib * 32 + m;
  const bool ci_ok = ci < Cin;
  const int dyo = t / 3 - 1, dxo = t % 3 - 1;
  const int npx = H * W;
  const int s0 = blockIdx.x * px_per_slice, s1 = min(npx, s0 + px_per_slice);
  f32x16 acc = {};
  for (int pb = s0; pb < s1; pb += 16) {
    u16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = pb + 8 * kg + j;
      bf16 av = 0, bv = 0;
      if (p < s1) {
        av = dy.ptr[(size_t)p * dy.stride + dy.off + co];
        const int ph = p / W, pw = p - ph * W;
        const int qh = ph + dyo, qw = pw + dxo;
        if (ci_ok && qh >= 0 && qh < H && qw >= 0 && qw < W)
          bv = x.ptr[(size_t)(qh * W + qw) * x.stride + x.off + ci];
      }
      a[j] = av;
      b[j] = bv;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                  0, 0, 0);
  }
  // D[m'][n] with m' = output channel row, n = lane%32 = input channel column
  const size_t base = (size_t)blockIdx.x * 9 * Cout * Cin + (size_t)t * Cout * Cin;
  if (ci_ok) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int cor = cb * 32 + 8 * (i >> 2) + 4 * kg + (i & 3);
      part[base + (size_t)cor * Cin + ci] = acc[i];
    }
  }
}

// out[i] (+)= sum_s part[s * n + i], s = 0 .. ns-1 in order.
__global__ void k_sum_slices_acc(const float *__restrict__ part, int ns, long n, float *__restrict__ out, int acc) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < ns; ++q) s += part[(size_t)q * n + i];
    out[i] = acc ? out[i] + s : s;
  }
}

// Head gradient from the per-block partials of k_head_bwd (C = 32).
__global__ void k_sum_head(const float *part, int nb, float *gw, float *gb, int acc) {
  const int c = threadIdx.x;
  if (c > 32) return;
  float s = 0.f;
  for (int q = 0; q < nb; ++q) s += part[(size_t)q * 33 + c];
  if (c < 32)
    gw[c] = acc ? gw[c] + s : s;
  else
    gb[0] = acc ? gb[0] + s : s;
}

__global__ void k_scale(float *x, long n, float s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] *= s;
}

// Bias gradient partials: part[slice][c] = sum over the slice's pixels of
// dy(p, c).  C divides 256; 256 threads = (256/C) pixel lanes x C channels.
__global__ void __launch_bounds__(256) k_bias_grad(TView dy, int C, int npx, int px_per_slice, float *part) {
  __shared__ float lds[256];
  const int c = threadIdx.x % C, pl = threadIdx.x / C, npl = 256 / C;
  const int s0 = blockIdx.x * px_per_slice, s1 = min(npx, s0 + px_per_slice);
  float s = 0.f;
  for (int p = s0 + pl; p < s1; p += npl) s += bf2f(dy.ptr[(size_t)p * dy.stride + dy.off + c]);
  lds[threadIdx.x] = s;
  __syncthreads();
  if (pl == 0) {
    float t = 0.f;
    for (int q = 0; q < npl; ++q) t += lds[q * C + c];
    part[(size_t)blockIdx.x * C + c] = t;
  }
}

// 2x2 average pooling of a channel slice: out (H/2 x W/2 x C).
__global__ void k_pool2(TView in, int H, int W, int C, TView out) {
  const int H2 = H / 2, W2 = W / 2;
  const long n = (long)H2 * W2 * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long q = i / C;
    const int h = (int)(q / W2), w = (int)(q % W2);
    const size_t p00 = (size_t)(2 * h) * W + 2 * w;
    const float s = bf2f(in.ptr[p00 * in.stride + in.off + c]) + bf2f(in.ptr[(p00 + 1) * in.stride + in.off + c]) +
                    bf2f(in.ptr[(p00 + W) * in.stride + in.off + c]) +
                    bf2f(in.ptr[(p00 + W + 1) * in.stride + in.off + c]);
    out.ptr[(size_t)q * out.stride + out.off + c] = f2bf(0.25f * s);
  }
}

// Backward of pooling, added to the skip gradient already in d (a slice of
// the concatenated decoder-input gradient) and masked by the skip tensor's
// ReLU: d(p, c) = (d(p, c) + dlow(parent(p), c) / 4) * [s(p, c) > 0].
__global__ void k_pool2_bwd_add(TView d, TView dlow, TView s, int H, int W, int C) {
  const long n = (long)H * W * C;
  const int W2 = W / 2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long p = i / C;
    const int h = (int)(p / W), w = (int)(p % W);
    const size_t q = (size_t)(h / 2) * W2 + w / 2;
    float v = bf2f(d.ptr[(size_t)p * d.stride + d.off + c]) + 0.25f * bf2f(dlow.ptr[q * dlow.stride + dlow.off + c]);
    if (!(bf2f(s.ptr[(size_t)p * s.stride + s.off + c]) > 0.f)) v = 0.f;
    d.ptr[(size_t)p * d.stride + d.off + c] = f2bf(v);
  }
}

// Nearest-neighbour 2x upsampling of low (H/2 x W/2 x C) into a slice of out.
__global__ void k_up2(TView low, int H, int W, int C, TView out) {
  const long n = (long)H * W * C;
  const int W2 = W / 2;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long p = i / C;
    const int h = (int)(p / W), w = (int)(p % W);
    const size_t q = (size_t)(h / 2) * W2 + w / 2;
    out.ptr[(size_t)p * out.stride + out.off + c] = low.ptr[q * low.stride + low.off + c];
  }
}

// Backward of upsampling: dlow(q, c) = sum of d over the 2x2 children.
__global__ void k_up2_bwd(TView d, int H, int W, int C, TView dlow) {
  const int H2 = H / 2, W2 = W / 2;
  const long n = (long)H2 * W2 * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long q = i / C;
    const int h = (int)(q / W2), w = (int)(q % W2);
    const size_t p00 = (size_t)(2 * h) * W + 2 * w;
    const float s = bf2f(d.ptr[p00 * d.stride + d.off + c]) + bf2f(d.ptr[(p00 + 1) * d.stride + d.off + c]) +
                    bf2f(d.ptr[(p00 + W) * d.stride + d.off + c]) + bf2f(d.ptr[(p00 + W + 1) * d.stride + d.off + c]);
    dlow.ptr[(size_t)q * dlow.stride + dlow.off + c] = f2bf(s);
  }
}

// 1x1 output layer: out(p) = b + sum_c x(p, c) w(c), FP32.
__global__ void k_head(TView x, int C, const float *w, const float *b, float *out, int npx) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npx; p += gridDim.x * blockDim.x) {
    float s = b[0];
    for (int c = 0; c < C; ++c) s += bf2f(x.ptr[(size_t)p * x.stride + x.off + c]) * w[c];
    out[p] = s;
  }
}

// Masked mean-square loss and its gradient: dout = 2 (out - target) mask / nmask;
// per-block partials of sum mask (out - target)^2 in part[blockIdx.x].
__global__ void __launch_bounds__(256) k_loss(const float *out, const float *target, const float *mask, int npx,
                                              float inv_n, float *dout, double *part) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npx; p += gridDim.x * blockDim.x) {
    const float e = (out[p] - target[p]) * mask[p];
    dout[p] = 2.f * e * inv_n;
    s += (double)e * e;
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// Backward of the head: dx(p, c) = dout(p) w(c) [x(p, c) > 0]; partials of
// dw(c) = sum_p dout(p) x(p, c) and db = sum_p dout(p) per block
// (part[blk * (C + 1) + c]).  C = 32, 256 threads = 8 pixel lanes x 32.
__global__ void __launch_bounds__(256) k_head_bwd(TView x, int C, const float *w, const float *dout, int npx,
                                                  int px_per_slice, TView dx, float *part) {
  __shared__ float lds[256], ldb[256];
  const int c = threadIdx.x % C, pl = threadIdx.x / C, npl = 256 / C;
  const int s0 = blockIdx.x * px_per_slice, s1 = min(npx, s0 + px_per_slice);
  float sw = 0.f, sb = 0.f;
  for (int p = s0 + pl; p < s1; p += npl) {
    const float xv = bf2f(x.ptr[(size_t)p * x.stride + x.off + c]);
    const float g = dout[p];
    sw += g * xv;
    if (c == 0) sb += g;
    dx.ptr[(size_t)p * dx.stride + dx.off + c] = f2bf(xv > 0.f ? g * w[c] : 0.f);
  }
  lds[threadIdx.x] = sw;
  ldb[threadIdx.x] = sb;
  __syncthreads();
  if (pl == 0) {
    float t = 0.f, tb = 0.f;
    for (int q = 0; q < npl; ++q) t += lds[q * C + c], tb += ldb[q * C + c];
    part[(size_t)blockIdx.x * (C + 1) + c] = t;
    if (c == 0) part[(size_t)blockIdx.x * (C + 1) + C] = tb;
  }
}

// Adam on FP32 master weights; refreshes the bf16 copy [9][Cout][Cin] and the
// flipped transposed copy [9][Cin][Cout] used by the input-gradient pass.
// For the head (conv == 0) only the FP32 weights are used.
__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, int conv, int Cout, int Cin, bf16 *wb, bf16 *wt) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (conv) {
      const int ci = (int)(i % Cin);
      const long r = i / Cin;
      const int co = (int)(r % Cout), t = (int)(r / Cout);
      wb[i] = f2bf(wi);
      wt[((size_t)(8 - t) * Cin + ci) * Cout + co] = f2bf(wi);
    }
  }
}

__global__ void k_refresh_bf16(const float *w, long n, int Cout, int Cin, bf16 *wb, bf16 *wt) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % Cin);
    const long r = i / Cin;
    const int co = (int)(r % Cout), t = (int)(r / Cout);
    wb[i] = f2bf(w[i]);
    wt[((size_t)(8 - t) * Cin + ci) * Cout + co] = f2bf(w[i]);
  }
}

// Dense input image (H x W x 16, channels 0..2 used) from up to three FP32
// planes; the channels beyond stay zero.
__global__ void k_pack_dense(const float *x3, int npx, bf16 *img) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npx; p += gridDim.x * blockDim.x)
    for (int ch = 0; ch < 3; ++ch) img[(size_t)p * 16 + ch] = f2bf(x3[(size_t)ch * npx + p]);
}

// ---------------------------------------------------------------------------
// The network.

struct ConvLayer {
  int Cin, Cout, level;
  long nw;  // 9 * Cout * Cin
  float *w, *b, *gw, *gb, *mw, *vw, *mb, *vb;
  bf16 *wb, *wt;
};

struct UNet {
  int H = 0, W = 0;  // level-0 size (multiples of 8)
  int Hl[4], Wl[4];
  std::vector<ConvLayer> conv;  // 14 3x3 layers
  float *hw = nullptr, *hb = nullptr, *hgw = nullptr, *hgb = nullptr, *hmw = nullptr, *hvw = nullptr, *hmb = nullptr,
        *hvb = nullptr;  // head (32 -> 1)
  // activations (bf16 NHWC) and their gradients
  bf16 *X0, *A0, *C0, *X1, *A1, *C1, *X2, *A2, *C2, *X3, *A3, *B3, *A4, *B2, *A5, *B1, *A6, *B0;
  bf16 *dA0, *dC0, *dX1, *dA1, *dC1, *dX2, *dA2, *dC2, *dX3, *dA3, *dB3, *dA4, *dB2, *dA5, *dB1, *dA6, *dB0;
  float *out = nullptr, *dout = nullptr;
  float *part = nullptr;  // slice partials (weight and bias gradients)
  size_t part_n = 0;
  double *lpart = nullptr;  // loss partials
  std::vector<void *> allocs;
  long adam_t = 0;
  // Arena mode: buffers are carved out of caller-provided device memory
  // (256-byte aligned) instead of allocated, and the network is built for
  // evaluation only (no Adam moments, no random initial weights).
  unsigned char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;

  template <typename T>
  T *alloc(size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (arena) {
      const size_t b = (bytes + 255) & ~(size_t)255;
      if (arena_used + b > arena_cap) return nullptr;
      void *p = arena + arena_used;
      arena_used += b;
      hipMemset(p, 0, bytes);
      return (T *)p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    hipMemset(p, 0, bytes);
    allocs.push_back(p);
    return (T *)p;
  }
  ~UNet() {
    for (void *p : allocs) hipFree(p);
  }
  size_t px(int l) const { return (size_t)Hl[l] * Wl[l]; }

  bool create(int h, int w, unsigned seed) {
    H = h;
    W = w;
    for (int l = 0; l < 4; ++l) Hl[l] = H >> l, Wl[l] = W >> l;
    const int cfg[14][3] = {{16, 32, 0},  {32, 32, 0},  {32, 64, 1},  {64, 64, 1},  {64, 128, 2},
                            {128, 128, 2}, {128, 256, 3}, {256, 256, 3}, {384, 128, 2}, {128, 128, 2},
                            {192, 64, 1},  {64, 64, 1},  {96, 32, 0},  {32, 32, 0}};
    std::mt19937 rng(seed);
    size_t maxpart = 0;
    for (auto &c : cfg) {
      ConvLayer L{};
      L.Cin = c[0];
      L.Cout = c[1];
      L.level = c[2];
      L.nw = 9L * L.Cout * L.Cin;
      const bool adam = arena == nullptr;
      L.w = alloc<float>(L.nw);
      L.gw = alloc<float>(L.nw);
      L.mw = adam ? alloc<float>(L.nw) : nullptr;
      L.vw = adam ? alloc<float>(L.nw) : nullptr;
      L.b = alloc<float>(L.Cout);
      L.gb = alloc<float>(L.Cout);
      L.mb = adam ? alloc<float>(L.Cout) : nullptr;
      L.vb = adam ? alloc<float>(L.Cout) : nullptr;
      L.wb = alloc<bf16>(L.nw);
      L.wt = alloc<bf16>(L.nw);
      if (!L.w || !L.gw || !L.b || !L.gb || !L.wb || !L.wt) return false;
      conv.push_back(L);
      const int nsl = slices_for(9 * (L.Cout / 32) * ((L.Cin + 31) / 32), px(L.level));
      maxpart = std::max(maxpart, (size_t)nsl * L.nw);
      maxpart = std::max(maxpart, (size_t)bias_slices(px(L.level)) * L.Cout);
      if (!adam) continue;
      // He initialisation (the input layer's unused channels 3..15 get zeros)
      const int fan_in = 9 * (L.Cin == 16 ? 3 : L.Cin);
      std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan_in));
      std::vector<float> hwv(L.nw);
      for (long i = 0; i < L.nw; ++i) {
        const int ci = (int)(i % L.Cin);
        hwv[i] = (L.Cin == 16 && ci >= 3) ? 0.f : nd(rng);
      }
      hipMemcpy(L.w, hwv.data(), L.nw * sizeof(float), hipMemcpyHostToDevice);
    }
    hw = alloc<float>(32);
    hb = alloc<float>(1);
    hgw = alloc<float>(32);
    hgb = alloc<float>(1);
    if (!arena) {
      hmw = alloc<float>(32);
      hvw = alloc<float>(32);
      hmb = alloc<float>(1);
      hvb = alloc<float>(1);
    }
    if (!arena) {
      std::normal_distribution<float> nd(0.f, std::sqrt(1.f / 32.f));
      std::vector<float> h(32);
      for (auto &v : h) v = 0.1f * nd(rng);
      hipMemcpy(hw, h.data(), 32 * sizeof(float), hipMemcpyHostToDevice);
    }
    maxpart = std::max(maxpart, (size_t)bias_slices(px(0)) * 33);
    part_n = maxpart;
    part = alloc<float>(part_n);
    lpart = alloc<double>(1024);
    const size_t p0 = px(0), p1 = px(1), p2 = px(2), p3 = px(3);
    X0 = alloc<bf16>(p0 * 16);
    A0 = alloc<bf16>(p0 * 32);
    C0 = alloc<bf16>(p0 * 96);
    X1 = alloc<bf16>(p1 * 32);
    A1 = alloc<bf16>(p1 * 64);
    C1 = alloc<bf16>(p1 * 192);
    X2 = alloc<bf16>(p2 * 64);
    A2 = alloc<bf16>(p2 * 128);
    C2 = alloc<bf16>(p2 * 384);
    X3 = alloc<bf16>(p3 * 128);
    A3 = alloc<bf16>(p3 * 256);
    B3 = alloc<bf16>(p3 * 256);
    A4 = alloc<bf16>(p2 * 128);
    B2 = alloc<bf16>(p2 * 128);
    A5 = alloc<bf16>(p1 * 64);
    B1 = alloc<bf16>(p1 * 64);
    A6 = alloc<bf16>(p0 * 32);
    B0 = alloc<bf16>(p0 * 32);
    dA0 = alloc<bf16>(p0 * 32);
    dC0 = alloc<bf16>(p0 * 96);
    dX1 = alloc<bf16>(p1 * 32);
    dA1 = alloc<bf16>(p1 * 64);
    dC1 = alloc<bf16>(p1 * 192);
    dX2 = alloc<bf16>(p2 * 64);
    dA2 = alloc<bf16>(p2 * 128);
    dC2 = alloc<bf16>(p2 * 384);
    dX3 = alloc<bf16>(p3 * 128);
    dA3 = alloc<bf16>(p3 * 256);
    dB3 = alloc<bf16>(p3 * 256);
    dA4 = alloc<bf16>(p2 * 128);
    dB2 = alloc<bf16>(p2 * 128);
    dA5 = alloc<bf16>(p1 * 64);
    dB1 = alloc<bf16>(p1 * 64);
    dA6 = alloc<bf16>(p0 * 32);
    dB0 = alloc<bf16>(p0 * 32);
    out = alloc<float>(p0);
    dout = alloc<float>(p0);
    if (!dB0 || !dout || !part) return false;
    refresh(nullptr);
    return hipDeviceSynchronize() == hipSuccess;
  }

  static int slices_for(int tiles, size_t npx) {
    int s = std::max(1, 2048 / std::max(1, tiles));
    s = (int)std::min<size_t>((size_t)s, (npx + 255) / 256);
    return std::max(1, s);
  }
  static int bias_slices(size_t npx) { return (int)std::max<size_t>(1, std::min<size_t>(512, (npx + 1023) / 1024)); }

  long num_weights() const {
    long n = 33;
    for (auto &L : conv) n += L.nw + L.Cout;
    return n;
  }

  void refresh(hipStream_t s) {
    for (auto &L : conv)
      hipLaunchKernelGGL(k_refresh_bf16, dim3(std::min<long>(1024, (L.nw + 255) / 256)), dim3(256), 0, s, L.w, L.nw,
                         L.Cout, L.Cin, L.wb, L.wt);
  }

  // weights <-> flat host array: per layer w (9*Cout*Cin) then b (Cout); head w (32), b (1)
  void get_weights(float *h) {
    for (auto &L : conv) {
      hipMemcpy(h, L.w, L.nw * sizeof(float), hipMemcpyDeviceToHost);
      h += L.nw;
      hipMemcpy(h, L.b, L.Cout * sizeof(float), hipMemcpyDeviceToHost);
      h += L.Cout;
    }
    hipMemcpy(h, hw, 32 * sizeof(float), hipMemcpyDeviceToHost);
    hipMemcpy(h + 32, hb, sizeof(float), hipMemcpyDeviceToHost);
  }
  void set_weights(const float *h) {
    for (auto &L : conv) {
      hipMemcpy(L.w, h, L.nw * sizeof(float), hipMemcpyHostToDevice);
      h += L.nw;
      hipMemcpy(L.b, h, L.Cout * sizeof(float), hipMemcpyHostToDevice);
      h += L.Cout;
    }
    hipMemcpy(hw, h, 32 * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(hb, h + 32, sizeof(float), hipMemcpyHostToDevice);
    refresh(nullptr);
  }
  void get_grads(float *h) {
    for (auto &L : conv) {
      hipMemcpy(h, L.gw, L.nw * sizeof(float), hipMemcpyDeviceToHost);
      h += L.nw;
      hipMemcpy(h, L.gb, L.Cout * sizeof(float), hipMemcpyDeviceToHost);
      h += L.Cout;
    }
    hipMemcpy(h, hgw, 32 * sizeof(float), hipMemcpyDeviceToHost);
    hipMemcpy(h + 32, hgb, sizeof(float), hipMemcpyDeviceToHost);
  }

  static dim3 ew_grid(size_t n) { return dim3((unsigned)std::max<size_t>(1, std::min<size_t>(8192, (n + 255) / 256))); }

  void conv_fwd(hipStream_t s, const ConvLayer &L, TView in, TView out) {
    const int l = L.level;
    conv3x3<true, false>(s, in, L.Cin, (const bf16 *)L.wb, (const float *)L.b, L.Cout, out, TView{nullptr, 0, 0},
                         Hl[l], Wl[l]);
  }
  // dx = conv_T(dy) masked by the ReLU of `mask`; then dW, db from (dy, x).
  void conv_bwd(hipStream_t s, const ConvLayer &L, TView dy, TView x, TView dx, TView mask, bool need_dx,
                bool accumulate) {
    const int l = L.level;
    const int npx = Hl[l] * Wl[l];
    if (need_dx) {
      if (mask.ptr)
        conv3x3<false, true>(s, dy, L.Cout, (const bf16 *)L.wt, nullptr, L.Cin, dx, mask, Hl[l], Wl[l]);
      else
        conv3x3<false, false>(s, dy, L.Cout, (const bf16 *)L.wt, nullptr, L.Cin, dx, TView{nullptr, 0, 0}, Hl[l],
                              Wl[l]);
    }
    const int tiles = 9 * (L.Cout / 32) * ((L.Cin + 31) / 32);
    const int ns = slices_for(tiles, npx);
    const int pps = ((npx + ns - 1) / ns + 15) / 16 * 16;
    hipLaunchKernelGGL(k_conv3x3_dw, dim3(ns, tiles), dim3(64), 0, s, dy, L.Cout, x, L.Cin, Hl[l], Wl[l], pps, part);
    float *gw_dst = L.gw;
    // accumulate: sum the slices into gw (+= via a second buffer pass)
    hipLaunchKernelGGL(k_sum_slices_acc, ew_grid(L.nw), dim3(256), 0, s, (const float *)part, ns, L.nw, gw_dst,
                       accumulate ? 1 : 0);
    const int nb = bias_slices(npx);
    const int bpps = (npx + nb - 1) / nb;
    hipLaunchKernelGGL(k_bias_grad, dim3(nb), dim3(256), 0, s, dy, L.Cout, npx, bpps, part);
    hipLaunchKernelGGL(k_sum_slices_acc, ew_grid(L.Cout), dim3(256), 0, s, (const float *)part, nb, (long)L.Cout,
                       L.gb, accumulate ? 1 : 0);
  }

  void forward(hipStream_t s) {
    auto V = [](bf16 *p, int st, int off) { return TView{p, st, off}; };
    conv_fwd(s, conv[0], V(X0, 16, 0), V(A0, 32, 0));
    conv_fwd(s, conv[1], V(A0, 32, 0), V(C0, 96, 64));
    hipLaunchKernelGGL(k_pool2, ew_grid(px(1) * 32), dim3(256), 0, s, V(C0, 96, 64), Hl[0], Wl[0], 32, V(X1, 32, 0));
    conv_fwd(s, conv[2], V(X1, 32, 0), V(A1, 64, 0));
    conv_fwd(s, conv[3], V(A1, 64, 0), V(C1, 192, 128));
    hipLaunchKernelGGL(k_pool2, ew_grid(px(2) * 64), dim3(256), 0, s, V(C1, 192, 128), Hl[1], Wl[1], 64,
                       V(X2, 64, 0));
    conv_fwd(s, conv[4], V(X2, 64, 0), V(A2, 128, 0));
    conv_fwd(s, conv[5], V(A2, 128, 0), V(C2, 384, 256));
    hipLaunchKernelGGL(k_pool2, ew_grid(px(3) * 128), dim3(256), 0, s, V(C2, 384, 256), Hl[2], Wl[2], 128,
                       V(X3, 128, 0));
    conv_fwd(s, conv[6], V(X3, 128, 0), V(A3, 256, 0));
    conv_fwd(s, conv[7], V(A3, 256, 0), V(B3, 256, 0));
    hipLaunchKernelGGL(k_up2, ew_grid(px(2) * 256), dim3(256), 0, s, V(B3, 256, 0), Hl[2], Wl[2], 256,
                       V(C2, 384, 0));
    conv_fwd(s, conv[8], V(C2, 384, 0), V(A4, 128, 0));
    conv_fwd(s, conv[9], V(A4, 128, 0), V(B2, 128, 0));
    hipLaunchKernelGGL(k_up2, ew_grid(px(1) * 128), dim3(256), 0, s, V(B2, 128, 0), Hl[1], Wl[1], 128,
                       V(C1, 192, 0));
    conv_fwd(s, conv[10], V(C1, 192, 0), V(A5, 64, 0));
    conv_fwd(s, conv[11], V(A5, 64, 0), V(B1, 64, 0));
    hipLaunchKernelGGL(k_up2, ew_grid(px(0) * 64), dim3(256), 0, s, V(B1, 64, 0), Hl[0], Wl[0], 64, V(C0, 96, 0));
    conv_fwd(s, conv[12], V(C0, 96, 0), V(A6, 32, 0));
    conv_fwd(s, conv[13], V(A6, 32, 0), V(B0, 32, 0));
    hipLaunchKernelGGL(k_head, ew_grid(px(0)), dim3(256), 0, s, V(B0, 32, 0), 32, (const float *)hw,
                       (const float *)hb, out, (int)px(0));
  }

  // Gradients of every weight from dout (already in `dout`).  accumulate:
  // add to the gradient buffers instead of overwriting them.
  void backward(hipStream_t s, bool accumulate) {
    auto V = [](bf16 *p, int st, int off) { return TView{p, st, off}; };
    const TView none{nullptr, 0, 0};
    {
      const int npx = (int)px(0);
      const int nb = bias_slices(px(0));
      const int pps = (npx + nb - 1) / nb;
      hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, s, V(B0, 32, 0), 32, (const float *)hw,
                         (const float *)dout, npx, pps, V(dB0, 32, 0), part);
      hipLaunchKernelGGL(k_sum_head, dim3(1), dim3(64), 0, s, (const float *)part, nb, hgw, hgb, accumulate ? 1 : 0);
    }
    conv_bwd(s, conv[13], V(dB0, 32, 0), V(A6, 32, 0), V(dA6, 32, 0), V(A6, 32, 0), true, accumulate);
    conv_bwd(s, conv[12], V(dA6, 32, 0), V(C0, 96, 0), V(dC0, 96, 0), V(C0, 96, 0), true, accumulate);
    hipLaunchKernelGGL(k_up2_bwd, ew_grid(px(1) * 64), dim3(256), 0, s, V(dC0, 96, 0), Hl[0], Wl[0], 64,
                       V(dB1, 64, 0));
    conv_bwd(s, conv[11], V(dB1, 64, 0), V(A5, 64, 0), V(dA5, 64, 0), V(A5, 64, 0), true, accumulate);
    conv_bwd(s, conv[10], V(dA5, 64, 0), V(C1, 192, 0), V(dC1, 192, 0), V(C1, 192, 0), true, accumulate);
    hipLaunchKernelGGL(k_up2_bwd, ew_grid(px(2) * 128), dim3(256), 0, s, V(dC1, 192, 0), Hl[1], Wl[1], 128,
                       V(dB2, 128, 0));
    conv_bwd(s, conv[9], V(dB2, 128, 0), V(A4, 128, 0), V(dA4, 128, 0), V(A4, 128, 0), true, accumulate);
    conv_bwd(s, conv[8], V(dA4, 128, 0), V(C2, 384, 0), V(dC2, 384, 0), V(C2, 384, 0), true, accumulate);
    hipLaunchKernelGGL(k_up2_bwd, ew_grid(px(3) * 256), dim3(256), 0, s, V(dC2, 384, 0), Hl[2], Wl[2], 256,
                       V(dB3, 256, 0));
    conv_bwd(s, conv[7], V(dB3, 256, 0), V(A3, 256, 0), V(dA3, 256, 0), V(A3, 256, 0), true, accumulate);
    conv_bwd(s, conv[6], V(dA3, 256, 0), V(X3, 128, 0), V(dX3, 128, 0), none, true, accumulate);
    hipLaunchKernelGGL(k_pool2_bwd_add, ew_grid(px(2) * 128), dim3(256), 0, s, V(dC2, 384, 256), V(dX3, 128, 0),
                       V(C2, 384, 256), Hl[2], Wl[2], 128);
    conv_bwd(s, conv[5], V(dC2, 384, 256), V(A2, 128, 0), V(dA2, 128, 0), V(A2, 128, 0), true, accumulate);
    conv_bwd(s, conv[4], V(dA2, 128, 0), V(X2, 64, 0), V(dX2, 64, 0), none, true, accumulate);
    hipLaunchKernelGGL(k_pool2_bwd_add, ew_grid(px(1) * 64), dim3(256), 0, s, V(dC1, 192, 128), V(dX2, 64, 0),
                       V(C1, 192, 128), Hl[1], Wl[1], 64);
    conv_bwd(s, conv[3], V(dC1, 192, 128), V(A1, 64, 0), V(dA1, 64, 0), V(A1, 64, 0), true, accumulate);
    conv_bwd(s, conv[2], V(dA1, 64, 0), V(X1, 32, 0), V(dX1, 32, 0), none, true, accumulate);
    hipLaunchKernelGGL(k_pool2_bwd_add, ew_grid(px(0) * 32), dim3(256), 0, s, V(dC0, 96, 64), V(dX1, 32, 0),
                       V(C0, 96, 64), Hl[0], Wl[0], 32);
    conv_bwd(s, conv[1], V(dC0, 96, 64), V(A0, 32, 0), V(dA0, 32, 0), V(A0, 32, 0), true, accumulate);
    conv_bwd(s, conv[0], V(dA0, 32, 0), V(X0, 16, 0), none, none, false, accumulate);
  }

  // loss of `out` against target (masked); writes dout; returns the loss
  // partial count (the loss itself is summed on the host by the caller).
  int loss(hipStream_t s, const float *target, const float *mask, float inv_n) {
    const int npx = (int)px(0);
    const int nb = std::min(1024, (npx + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, (const float *)out, target, mask, npx, inv_n, dout, lpart);
    return nb;
  }

  void adam(hipStream_t s, float lr, float scale) {
    ++adam_t;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
    for (auto &L : conv) {
      if (scale != 1.f) {
        hipLaunchKernelGGL(k_scale, ew_grid(L.nw), dim3(256), 0, s, L.gw, L.nw, scale);
        hipLaunchKernelGGL(k_scale, ew_grid(L.Cout), dim3(256), 0, s, L.gb, (long)L.Cout, scale);
      }
      hipLaunchKernelGGL(k_adam, ew_grid(L.nw), dim3(256), 0, s, L.w, (const float *)L.gw, L.mw, L.vw, L.nw, lr, b1,
                         b2, eps, c1, c2, 1, L.Cout, L.Cin, L.wb, L.wt);
      hipLaunchKernelGGL(k_adam, ew_grid(L.Cout), dim3(256), 0, s, L.b, (const float *)L.gb, L.mb, L.vb,
                         (long)L.Cout, lr, b1, b2, eps, c1, c2, 0, 0, 1, (bf16 *)nullptr, (bf16 *)nullptr);
    }
    if (scale != 1.f) {
      hipLaunchKernelGGL(k_scale, dim3(1), dim3(64), 0, s, hgw, 32L, scale);
      hipLaunchKernelGGL(k_scale, dim3(1), dim3(64), 0, s, hgb, 1L, scale);
    }
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(64), 0, s, hw, (const float *)hgw, hmw, hvw, 32L, lr, b1, b2, eps, c1,
                       c2, 0, 0, 1, (bf16 *)nullptr, (bf16 *)nullptr);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(64), 0, s, hb, (const float *)hgb, hmb, hvb, 1L, lr, b1, b2, eps, c1, c2,
                       0, 0, 1, (bf16 *)nullptr, (bf16 *)nullptr);
  }
};

}  // namespace s3g

// ---------------------------------------------------------------------------
// Learner: training pairs recorded by the plain solver, training, inference.

namespace s3g {

// Pair slot layout: planes of lm floats: r0/sigma, (rr-1)/rr, target, and the
// matrix of that solve (rr, sx, sy) for the energy-norm training loss.
enum { PL_R0 = 0, PL_COEF = 1, PL_TGT = 2, PL_RR = 3, PL_SX = 4, PL_SY = 5, PL_NUM = 6 };

__global__ void __launch_bounds__(256) k_learn_residual(Outer O, float *slot, double *part) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
    const int lW = O.lW[c], lS = O.lS[c], lN = O.lN[c], lE = O.lE[c];
    const double r = O.b[c] - (O.rr[c] * O.x[c] - O.sx[lW] * O.x[lW] - O.sy[lS] * O.x[lS] - O.sy[c] * O.x[lN] -
                               O.sx[c] * O.x[lE]);
    s += r * r;
    slot[(size_t)PL_R0 * O.n + c] = (float)r;  // normalised by k_learn_normalise
    slot[(size_t)PL_COEF * O.n + c] = (float)((O.rr[c] - 1.0) / O.rr[c]);
    slot[(size_t)PL_RR * O.n + c] = (float)O.rr[c];
    slot[(size_t)PL_SX * O.n + c] = (float)O.sx[c];
    slot[(size_t)PL_SY * O.n + c] = (float)O.sy[c];
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// sigma = rms(r0) (fixed-order sum of the block partials); one workgroup.
__global__ void k_learn_sigma(const double *part, int nb, int n, double *sig) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sig[0] = fmax(sqrt(lds[0] / fmax(1.0, (double)n)), 1e-300);
}

__global__ void k_learn_normalise(float *slot, int n, const double *sig, const double *x, double *xg) {
  const double inv = 1.0 / sig[0];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    slot[(size_t)PL_R0 * n + c] = (float)((double)slot[(size_t)PL_R0 * n + c] * inv);
    xg[c] = x[c];
  }
}

__global__ void k_learn_target(float *slot, int n, const double *sig, const double *x, const double *xg) {
  const double inv = 1.0 / sig[0];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    slot[(size_t)PL_TGT * n + c] = (float)((x[c] - xg[c]) * inv);
}

// Scatter a slot's planes onto the dense image planes (input channels 0..2 =
// r0/sigma, coupling, wet mask) and the dense target.
__global__ void k_learn_scatter(const float *slot, int n, const int *pix, int npx, float *in3, float *tgt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int p = pix[c];
    in3[p] = slot[(size_t)PL_R0 * n + c];
    in3[(size_t)npx + p] = slot[(size_t)PL_COEF * n + c];
    in3[(size_t)2 * npx + p] = 1.f;
    if (tgt) tgt[p] = slot[(size_t)PL_TGT * n + c];
  }
}

// Energy-norm errors of the plain guess xg and the learned guess xl against
// the converged x: partials of e^T A e for e = x - xg (part[blk]) and
// e = x - xl (part[nb + blk]).
__global__ void __launch_bounds__(256) k_learn_anorm(Outer O, const double *xg, const double *xl, double *part) {
  __shared__ double lds[2][256];
  double sg = 0.0, sl = 0.0;
  auto eg = [&](int j) { return j == O.n ? 0.0 : O.x[j] - xg[j]; };
  auto el = [&](int j) { return j == O.n ? 0.0 : O.x[j] - xl[j]; };
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
    const int lW = O.lW[c], lS = O.lS[c], lN = O.lN[c], lE = O.lE[c];
    const double ag = O.rr[c] * eg(c) - O.sx[lW] * eg(lW) - O.sy[lS] * eg(lS) - O.sy[c] * eg(lN) - O.sx[c] * eg(lE);
    const double al = O.rr[c] * el(c) - O.sx[lW] * el(lW) - O.sy[lS] * el(lS) - O.sy[c] * el(lN) - O.sx[c] * el(lE);
    sg += eg(c) * ag;
    sl += el(c) * al;
  }
  lds[0][threadIdx.x] = sg;
  lds[1][threadIdx.x] = sl;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + o];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = lds[0][0];
    part[gridDim.x + blockIdx.x] = lds[1][0];
  }
}

// acc[0] += ||x - xl||_A / ||x - xg||_A, acc[1] += 1.  One workgroup.
__global__ void k_learn_gain(const double *part, int nb, double *acc) {
  __shared__ double lds[2][256];
  double sg = 0.0, sl = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) sg += part[i], sl += part[nb + i];
  lds[0][threadIdx.x] = sg;
  lds[1][threadIdx.x] = sl;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + o];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && lds[0][0] > 0.0) {
    acc[0] += sqrt(fmax(lds[1][0], 0.0) / lds[0][0]);
    acc[1] += 1.0;
  }
}

// Change of the energy functional CG minimises, phi(x) = x^T A x / 2 - b^T x,
// from the plain guess xg to the learned guess x: with d = x - xg and
// r = b - A xg, phi(x) - phi(xg) = d^T A d / 2 - d^T r, which is half the change
// of the squared A-norm error; negative exactly when the learned guess is
// closer to the solution.  Partials of d^T A d (part[blk]) and d^T r
// (part[nb + blk]).
__global__ void __launch_bounds__(256) k_learn_functional(Outer O, const double *xg, const double *x, double *part) {
  __shared__ double lds[2][256];
  double sa = 0.0, sr = 0.0;
  auto dv = [&](int j) { return j == O.n ? 0.0 : x[j] - xg[j]; };
  auto gv = [&](int j) { return j == O.n ? 0.0 : xg[j]; };
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
    const int lW = O.lW[c], lS = O.lS[c], lN = O.lN[c], lE = O.lE[c];
    const double ad = O.rr[c] * dv(c) - O.sx[lW] * dv(lW) - O.sy[lS] * dv(lS) - O.sy[c] * dv(lN) - O.sx[c] * dv(lE);
    const double r = O.b[c] - (O.rr[c] * gv(c) - O.sx[lW] * gv(lW) - O.sy[lS] * gv(lS) - O.sy[c] * gv(lN) -
                               O.sx[c] * gv(lE));
    sa += dv(c) * ad;
    sr += dv(c) * r;
  }
  lds[0][threadIdx.x] = sa;
  lds[1][threadIdx.x] = sr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + o];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = lds[0][0];
    part[gridDim.x + blockIdx.x] = lds[1][0];
  }
}

// Keeps the learned guess only if it lowers the energy functional: every
// workgroup adds the partials of k_learn_functional in the same order, so all
// take the same decision; otherwise (or if the change is not finite) the plain
// guess goes back into x and acc[2] counts the rejection.
__global__ void __launch_bounds__(256) k_learn_keep(const double *part, int nb, const double *xg, double *x, int n,
                                                    double *acc) {
  __shared__ double lds[2][256];
  double sa = 0.0, sr = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) sa += part[i], sr += part[nb + i];
  lds[0][threadIdx.x] = sa;
  lds[1][threadIdx.x] = sr;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + o];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  const double dphi = 0.5 * lds[0][0] - lds[1][0];
  if (dphi < 0.0) return;  // closer: keep it (NaN compares false and is rejected)
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) x[c] = xg[c];
  if (blockIdx.x == 0 && threadIdx.x == 0) acc[2] += 1.0;
}

__global__ void k_copy_f64(double *dst, const double *src, int n) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) dst[c] = src[c];
}

// Energy-norm training loss of one pair, the norm CG minimises:
//   loss = d^T A d / n,  d = out(pix) - target on the wet columns,
// with A the pair's own matrix; dout(pix) = 2 A d / n (A is symmetric).  The
// neighbour values of d are recomputed from out and the target.
__global__ void __launch_bounds__(256) k_learn_loss(const float *out, const float *slot, const int *pix, const int *lE,
                                                    const int *lW, const int *lN, const int *lS, int n, float inv_n,
                                                    float *dout, double *part) {
  __shared__ double lds[256];
  const float *tg = slot + (size_t)PL_TGT * n, *rr = slot + (size_t)PL_RR * n, *sx = slot + (size_t)PL_SX * n,
              *sy = slot + (size_t)PL_SY * n;
  auto d = [&](int j) -> double { return j == n ? 0.0 : (double)out[pix[j]] - (double)tg[j]; };
  auto sxv = [&](int j) -> double { return j == n ? 0.0 : (double)sx[j]; };
  auto syv = [&](int j) -> double { return j == n ? 0.0 : (double)sy[j]; };
  double s = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
    const int w = lW[c], so = lS[c], no = lN[c], e = lE[c];
    const double dc = d(c);
    const double ad = (double)rr[c] * dc - sxv(w) * d(w) - syv(so) * d(so) - (double)sy[c] * d(no) -
                      (double)sx[c] * d(e);
    s += dc * ad;
    dout[pix[c]] = (float)(2.0 * ad * inv_n);
  }
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// x += sigma * out(pix) on the wet columns.
__global__ void k_learn_apply(double *x, int n, const int *pix, const float *out, const double *sig) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    x[c] += sig[0] * (double)out[pix[c]];
}

// Input plane 0 of the correction stage: r / sigma on the wet pixels, with
// sigma = rms(r) from the CG's (r,r) partials (every workgroup adds them in
// the same order); block 0 stores sigma.  The coupling and mask planes are the
// ones the solve scattered before its first iteration.
template <typename T>
__global__ void __launch_bounds__(256) k_flex_input(const T *r, const double *part, int nblk, int n, const int *pix,
                                                   float *in0, double *sig) {
  __shared__ double lds[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  const double sg = fmax(sqrt(lds[0] / fmax(1.0, (double)n)), 1e-300);
  if (blockIdx.x == 0 && threadIdx.x == 0) sig[0] = sg;
  const double inv = 1.0 / sg;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    in0[pix[c]] = (float)((double)r[c] * inv);
}

struct Learner {
  UNet net;
  int H = 0, W = 0, npx = 0, n = 0;
  int cap = 0, count = 0, head = 0;
  int mode = 0;
  bool recording = true;
  long steps = 0;
  double last_loss = NAN;
  int *pix = nullptr;
  float *pairs = nullptr, *scratch_slot = nullptr, *in3 = nullptr, *tgt = nullptr, *mask = nullptr;
  double *sig = nullptr, *xg = nullptr, *xl = nullptr, *part = nullptr, *gain = nullptr;
  bool used = false;  // the current solve started from the learned guess
  float *cur_slot = nullptr;
  int flex_period = 0;      // correction stage every flex_period CG iterations (0: off)
  bool flex_ready = false;  // this solve scattered the coupling and mask planes
  long corrections = 0;     // correction stages run since the mode was set
  std::mt19937 rng;
};

static int learn_fail(const char *m) {
  set_error(m);
  return S3G_ERR_HIP;
}

// Device memory of s3g_unet_check_static, the variant that allocates nothing:
// a file-scope array every GPU holds from the library's load to the end of
// the process, big enough for the network's weights, activations, gradients
// and partial sums and the inputs of one dense field of at most 4096 cells.
// Its host-side name is not a device address: each device's copy is reached
// through the address hipGetSymbolAddress returns with that device current.
constexpr size_t kStaticScratchBytes = (size_t)40 << 20;
constexpr long kStaticMaxCells = 4096;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static unsigned char *scratch_addr[kMaxDevices];

// The current device's copy of the scratch array (looked up once per device).
static int static_scratch(unsigned char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return learn_fail("s3g_unet_check_static: no usable current device");
  }
  if (!scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return learn_fail("s3g_unet_check_static: hipGetSymbolAddress of the scratch array failed");
    }
    scratch_addr[dev] = static_cast<unsigned char *>(p);
  }
  *out = scratch_addr[dev];
  return 0;
}

// Before the CG: record r0 / sigma and the coupling field of this solve and,
// in mode 1, move the guess by the network's correction.
int learn_before_solve(s3g_ctx *c, const void *outer_v, double *x) {
  Learner *L = c->learn;
  if (!L) return 0;
  const Outer &O = *(const Outer *)outer_v;
  const bool use = L->mode >= 1 && L->steps > 0;
  const bool flex = L->flex_period > 0 && L->steps > 0;
  L->flex_ready = false;
  if (!L->recording && !use && !flex) {
    L->cur_slot = nullptr;
    return 0;
  }
  L->cur_slot = (L->recording && L->cap > 0) ? L->pairs + (size_t)L->head * PL_NUM * L->n : L->scratch_slot;
  const int nb = std::min(1024, (L->n + 255) / 256);
  hipLaunchKernelGGL(k_learn_residual, dim3(nb), dim3(256), 0, c->stream, O, L->cur_slot, L->part);
  hipLaunchKernelGGL(k_learn_sigma, dim3(1), dim3(256), 0, c->stream, (const double *)L->part, nb, L->n, L->sig);
  hipLaunchKernelGGL(k_learn_normalise, dim3(nb), dim3(256), 0, c->stream, L->cur_slot, L->n, (const double *)L->sig,
                     (const double *)x, L->xg);
  if (use || flex) {  // input planes: r0/sigma, coupling, mask (the last two also serve the correction stage)
    hipLaunchKernelGGL(k_learn_scatter, dim3(nb), dim3(256), 0, c->stream, (const float *)L->cur_slot, L->n,
                       (const int *)L->pix, L->npx, L->in3, (float *)nullptr);
    L->flex_ready = flex;
  }
  if (use) {
    hipLaunchKernelGGL(k_pack_dense, UNet::ew_grid(L->npx), dim3(256), 0, c->stream, (const float *)L->in3, L->npx,
                       L->net.X0);
    L->net.forward(c->stream);
    hipLaunchKernelGGL(k_learn_apply, dim3(nb), dim3(256), 0, c->stream, x, L->n, (const int *)L->pix,
                       (const float *)L->net.out, (const double *)L->sig);
    // the plain guess is the fallback: keep the learned one only if it is closer
    // to the solution in the energy norm (the CG then still runs to the same
    // tolerance either way)
    hipLaunchKernelGGL(k_learn_functional, dim3(nb), dim3(256), 0, c->stream, O, (const double *)L->xg,
                       (const double *)x, L->part);
    hipLaunchKernelGGL(k_learn_keep, dim3(nb), dim3(256), 0, c->stream, (const double *)L->part, nb,
                       (const double *)L->xg, x, L->n, L->gain);
    hipLaunchKernelGGL(k_copy_f64, dim3(nb), dim3(256), 0, c->stream, L->xl, (const double *)x, L->n);
  }
  L->used = use;
  return hip_check(hipGetLastError(), "learned initial guess");
}

// After the CG: the target (zeta - guess) / sigma completes the pair.
int learn_flex_period(const s3g_ctx *c) {
  const Learner *L = c->learn;
  return (L && L->flex_ready) ? L->flex_period : 0;
}

int learn_correction(s3g_ctx *c, const void *r, bool fp32, const double *rr_part, int nblk, LearnOut *out) {
  Learner *L = c->learn;
  if (!L || !L->flex_ready) return learn_fail("learned correction stage without a prepared solve");
  const int nb = std::min(1024, (L->n + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL((k_flex_input<float>), dim3(nb), dim3(256), 0, c->stream, (const float *)r, rr_part, nblk,
                       L->n, (const int *)L->pix, L->in3, L->sig + 1);
  else
    hipLaunchKernelGGL((k_flex_input<double>), dim3(nb), dim3(256), 0, c->stream, (const double *)r, rr_part, nblk,
                       L->n, (const int *)L->pix, L->in3, L->sig + 1);
  hipLaunchKernelGGL(k_pack_dense, UNet::ew_grid(L->npx), dim3(256), 0, c->stream, (const float *)L->in3, L->npx,
                     L->net.X0);
  L->net.forward(c->stream);
  L->corrections += 1;
  out->out = L->net.out;
  out->pix = L->pix;
  out->sig = L->sig + 1;
  return hip_check(hipGetLastError(), "learned correction stage");
}

int learn_after_solve(s3g_ctx *c, const void *outer_v) {
  Learner *L = c->learn;
  if (!L || !L->cur_slot) return 0;
  const Outer &O = *(const Outer *)outer_v;
  const double *x = O.x;
  const int nb = std::min(1024, (L->n + 255) / 256);
  if (L->used) {
    hipLaunchKernelGGL(k_learn_anorm, dim3(nb), dim3(256), 0, c->stream, O, (const double *)L->xg,
                       (const double *)L->xl, L->part);
    hipLaunchKernelGGL(k_learn_gain, dim3(1), dim3(256), 0, c->stream, (const double *)L->part, nb, L->gain);
  }
  hipLaunchKernelGGL(k_learn_target, dim3(nb), dim3(256), 0, c->stream, L->cur_slot, L->n, (const double *)L->sig, x,
                     (const double *)L->xg);
  if (L->recording && L->cap > 0 && L->cur_slot != L->scratch_slot) {
    L->head = (L->head + 1) % L->cap;
    L->count = std::min(L->count + 1, L->cap);
  }
  L->cur_slot = nullptr;
  return hip_check(hipGetLastError(), "learn record");
}

void learn_destroy(s3g_ctx *c) {
  if (c && c->learn) {
    hipDeviceSynchronize();
    for (void *p : c->learn_allocs) hipFree(p);
    c->learn_allocs.clear();
    delete c->learn;
    c->learn = nullptr;
  }
}

}  // namespace s3g

using namespace s3g;

extern "C" {

long s3g_learn_num_weights(void) {
  // 14 convolutions (w + b) and the 1x1 head, in the order of UNet::create
  const int cfg[14][2] = {{16, 32}, {32, 32}, {32, 64}, {64, 64}, {64, 128}, {128, 128}, {128, 256},
                          {256, 256}, {384, 128}, {128, 128}, {192, 64}, {64, 64}, {96, 32}, {32, 32}};
  long n = 33;
  for (auto &q : cfg) n += 9L * q[0] * q[1] + q[1];
  return n;
}

int s3g_learn_setup(s3g_ctx *c, int capacity, unsigned seed) {
  if (!c || capacity < 0) return S3G_ERR_ARG;
  learn_destroy(c);
  auto *L = new Learner();
  c->learn = L;
  L->n = c->lm;
  L->cap = capacity;
  L->rng.seed(seed);
  // dense box of the wet columns: h = i - 2, w = j - 2, padded to multiples of 8
  int imax = 0, jmax = 0;
  const std::vector<int> &hi = c->l2i, &hj = c->l2j;
  for (int k = 0; k < c->lm; ++k) imax = std::max(imax, hi[k] - 2), jmax = std::max(jmax, hj[k] - 2);
  L->H = (imax + 1 + 7) / 8 * 8;
  L->W = (jmax + 1 + 7) / 8 * 8;
  L->npx = L->H * L->W;
  std::vector<int> pix(c->lm);
  std::vector<float> mask(L->npx, 0.f);
  for (int k = 0; k < c->lm; ++k) {
    pix[k] = (hi[k] - 2) * L->W + (hj[k] - 2);
    mask[pix[k]] = 1.f;
  }
  auto dal = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    hipMemset(p, 0, std::max<size_t>(bytes, 16));
    c->learn_allocs.push_back(p);
    return p;
  };
  L->pix = (int *)dal(c->lm * sizeof(int));
  L->pairs = (float *)dal((size_t)std::max(1, capacity) * PL_NUM * c->lm * sizeof(float));
  L->scratch_slot = (float *)dal((size_t)PL_NUM * c->lm * sizeof(float));
  L->in3 = (float *)dal((size_t)3 * L->npx * sizeof(float));
  L->tgt = (float *)dal((size_t)L->npx * sizeof(float));
  L->mask = (float *)dal((size_t)L->npx * sizeof(float));
  L->sig = (double *)dal(2 * sizeof(double));  // [0] initial residual, [1] correction stage
  L->xg = (double *)dal((size_t)(c->lm + 1) * sizeof(double));
  L->xl = (double *)dal((size_t)(c->lm + 1) * sizeof(double));
  L->part = (double *)dal(2048 * sizeof(double));
  L->gain = (double *)dal(3 * sizeof(double));
  if (!L->pix || !L->pairs || !L->part) {
    learn_destroy(c);
    return learn_fail("s3g_learn_setup: device allocation failed");
  }
  hipMemcpy(L->pix, pix.data(), c->lm * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(L->mask, mask.data(), L->npx * sizeof(float), hipMemcpyHostToDevice);
  if (!L->net.create(L->H, L->W, seed)) {
    learn_destroy(c);
    return learn_fail("s3g_learn_setup: network allocation failed");
  }
  return hip_check(hipDeviceSynchronize(), "s3g_learn_setup");
}

int s3g_learn_record(s3g_ctx *c, int on) {
  if (!c || !c->learn) return S3G_ERR_ARG;
  c->learn->recording = on != 0;
  return 0;
}

int s3g_learn_mode(s3g_ctx *c, int mode) {
  if (!c || !c->learn || mode < 0 || mode > 1) return S3G_ERR_ARG;
  c->learn->mode = mode;
  hipMemsetAsync(c->learn->gain, 0, 3 * sizeof(double), c->stream);
  return hip_check(hipStreamSynchronize(c->stream), "s3g_learn_mode");
}

int s3g_learn_flexible(s3g_ctx *c, int period) {
  if (!c || !c->learn || period < 0) return S3G_ERR_ARG;
  c->learn->flex_period = period;
  c->learn->corrections = 0;
  return 0;
}

int s3g_learn_train(s3g_ctx *c, int steps, int batch, double lr, double *loss_out) {
  if (!c || !c->learn || steps < 0 || batch < 1) return S3G_ERR_ARG;
  Learner *L = c->learn;
  if (L->count == 0) {
    set_error("s3g_learn_train: no recorded pairs");
    return S3G_ERR_ARG;
  }
  hipStream_t s = c->stream;
  const int nb = std::min(1024, (L->n + 255) / 256);
  double loss_sum = 0.0;
  int loss_n = 0;
  for (int it = 0; it < steps; ++it) {
    const bool last = it == steps - 1;
    for (int b = 0; b < batch; ++b) {
      const int idx = (int)(L->rng() % (unsigned)L->count);
      const float *slot = L->pairs + (size_t)idx * PL_NUM * L->n;
      hipLaunchKernelGGL(k_learn_scatter, dim3(nb), dim3(256), 0, s, slot, L->n, (const int *)L->pix, L->npx, L->in3,
                         (float *)nullptr);
      hipLaunchKernelGGL(k_pack_dense, UNet::ew_grid(L->npx), dim3(256), 0, s, (const float *)L->in3, L->npx,
                         L->net.X0);
      L->net.forward(s);
      const Dev<double> &D = c->d64;
      const int *lE = c->precision == S3G_FP64 ? D.lE : c->d32.lE, *lW = c->precision == S3G_FP64 ? D.lW : c->d32.lW,
                *lN = c->precision == S3G_FP64 ? D.lN : c->d32.lN, *lS = c->precision == S3G_FP64 ? D.lS : c->d32.lS;
      const int lb = nb;
      hipLaunchKernelGGL(k_learn_loss, dim3(nb), dim3(256), 0, s, (const float *)L->net.out, slot,
                         (const int *)L->pix, lE, lW, lN, lS, L->n, 1.f / (float)L->n, L->net.dout, L->net.lpart);
      L->net.backward(s, b > 0);
      if (last) {
        std::vector<double> h(lb);
        hipMemcpyAsync(h.data(), L->net.lpart, lb * sizeof(double), hipMemcpyDeviceToHost, s);
        if (hip_check(hipStreamSynchronize(s), "s3g_learn_train") < 0) return S3G_ERR_HIP;
        double t = 0.0;
        for (double v : h) t += v;
        loss_sum += t / L->n;
        ++loss_n;
      }
    }
    L->net.adam(s, (float)lr, 1.f / batch);
    ++L->steps;
  }
  if (hip_check(hipStreamSynchronize(s), "s3g_learn_train") < 0) return S3G_ERR_HIP;
  if (loss_n) L->last_loss = loss_sum / loss_n;
  if (loss_out) *loss_out = L->last_loss;
  return 0;
}

int s3g_learn_get_weights(s3g_ctx *c, float *host, long n) {
  if (!c || !c->learn || !host || n != s3g_learn_num_weights()) return S3G_ERR_ARG;
  c->learn->net.get_weights(host);
  return hip_check(hipDeviceSynchronize(), "s3g_learn_get_weights");
}

int s3g_learn_set_weights(s3g_ctx *c, const float *host, long n) {
  if (!c || !c->learn || !host || n != s3g_learn_num_weights()) return S3G_ERR_ARG;
  c->learn->net.set_weights(host);
  if (c->learn->steps == 0) c->learn->steps = 1;  // usable as trained
  return hip_check(hipDeviceSynchronize(), "s3g_learn_set_weights");
}

int s3g_learn_info(s3g_ctx *c, s3g_learn_stats *out) {
  if (!c || !c->learn || !out) return S3G_ERR_ARG;
  const Learner *L = c->learn;
  out->pairs = L->count;
  out->capacity = L->cap;
  out->train_steps = L->steps;
  out->last_loss = L->last_loss;
  out->height = L->H;
  out->width = L->W;
  out->mode = L->mode;
  double g[3] = {0.0, 0.0, 0.0};
  hipMemcpy(g, L->gain, 3 * sizeof(double), hipMemcpyDeviceToHost);
  out->guess_solves = (long)g[1];
  out->guess_error_ratio = g[1] > 0 ? g[0] / g[1] : NAN;
  out->guess_rejected = (long)g[2];
  out->corrections = L->corrections;
  out->flex_period = L->flex_period;
  return hip_check(hipGetLastError(), "s3g_learn_info");
}

// Standalone forward + backward of the network on one dense H x W input
// (3 planes), for gradient checks: loss = mean over mask of (out - target)^2
// with 1/n = 1 / sum(mask).
int s3g_unet_check(int H, int W, const float *weights, const float *x3, const float *target, const float *mask,
                   float *out, float *grads, double *loss) {
  if (H < 8 || W < 8 || H % 8 || W % 8 || !weights || !x3 || !target || !mask) return S3G_ERR_ARG;
  const long nw = s3g_learn_num_weights();
  auto net = std::make_unique<UNet>();
  if (!net->create(H, W, 1)) return learn_fail("s3g_unet_check: allocation failed");
  net->set_weights(weights);
  const int npx = H * W;
  float *d_in3 = nullptr, *d_t = nullptr, *d_m = nullptr;
  hipMalloc(&d_in3, 3 * (size_t)npx * sizeof(float));
  hipMalloc(&d_t, (size_t)npx * sizeof(float));
  hipMalloc(&d_m, (size_t)npx * sizeof(float));
  hipMemcpy(d_in3, x3, 3 * (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_t, target, (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_m, mask, (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  double nm = 0.0;
  for (int p = 0; p < npx; ++p) nm += mask[p];
  hipLaunchKernelGGL(k_pack_dense, UNet::ew_grid(npx), dim3(256), 0, 0, (const float *)d_in3, npx, net->X0);
  net->forward(0);
  const int lb = net->loss(0, d_t, d_m, (float)(1.0 / std::max(1.0, nm)));
  net->backward(0, false);
  int rc = hip_check(hipDeviceSynchronize(), "s3g_unet_check");
  if (rc == 0) {
    if (out) hipMemcpy(out, net->out, (size_t)npx * sizeof(float), hipMemcpyDeviceToHost);
    if (grads) {
      std::vector<float> g(nw);
      net->get_grads(g.data());
      std::memcpy(grads, g.data(), nw * sizeof(float));
    }
    if (loss) {
      std::vector<double> h(lb);
      hipMemcpy(h.data(), net->lpart, lb * sizeof(double), hipMemcpyDeviceToHost);
      double t = 0.0;
      for (double v : h) t += v;
      *loss = t / std::max(1.0, nm);
    }
  }
  hipFree(d_in3);
  hipFree(d_t);
  hipFree(d_m);
  return rc;
}


int s3g_unet_check_static(int H, int W, const float *weights, const float *x3, const float *target,
                          const float *mask, float *out, float *grads, double *loss) {
  if (H < 8 || W < 8 || H % 8 || W % 8 || !weights || !x3 || !target || !mask) return S3G_ERR_ARG;
  if ((long)H * W > kStaticMaxCells)
    return learn_fail("s3g_unet_check_static: the input has more than 4096 cells (the static scratch holds "
                      "a network of at most 64 x 64)");
  unsigned char *base = nullptr;
  if (int e = static_scratch(&base)) return e;
  UNet net;
  net.arena = base;
  net.arena_cap = kStaticScratchBytes;
  if (!net.create(H, W, 1)) return learn_fail("s3g_unet_check_static: the network does not fit the static scratch");
  const int npx = H * W;
  float *d_in3 = net.alloc<float>(3 * (size_t)npx), *d_t = net.alloc<float>(npx), *d_m = net.alloc<float>(npx);
  if (!d_in3 || !d_t || !d_m) return learn_fail("s3g_unet_check_static: the input does not fit the static scratch");
  net.set_weights(weights);
  hipMemcpy(d_in3, x3, 3 * (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_t, target, (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_m, mask, (size_t)npx * sizeof(float), hipMemcpyHostToDevice);
  double nm = 0.0;
  for (int p = 0; p < npx; ++p) nm += mask[p];
  hipLaunchKernelGGL(k_pack_dense, UNet::ew_grid(npx), dim3(256), 0, 0, (const float *)d_in3, npx, net.X0);
  net.forward(0);
  const int lb = net.loss(0, d_t, d_m, (float)(1.0 / std::max(1.0, nm)));
  net.backward(0, false);
  int rc = hip_check(hipDeviceSynchronize(), "s3g_unet_check_static");
  if (rc == 0 && out) rc = hip_check(hipMemcpy(out, net.out, (size_t)npx * sizeof(float), hipMemcpyDeviceToHost), "out");
  if (rc == 0 && grads) net.get_grads(grads);
  if (rc == 0 && loss) {
    double t = 0.0, h[1024];
    rc = hip_check(hipMemcpy(h, net.lpart, lb * sizeof(double), hipMemcpyDeviceToHost), "loss");
    for (int k = 0; k < lb; ++k) t += h[k];
    *loss = t / std::max(1.0, nm);
  }
  return rc;
}

}  // extern "C"
