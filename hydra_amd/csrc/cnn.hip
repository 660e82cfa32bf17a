// cnn.hip -- the learned sub-grid velocity correction: a 4-layer 3x3
// convolutional network on the coarse grid (inputs u, v, q; 32 channels;
// GELU), whose 2-channel output (du, dv) is sampled bilinearly at the
// contour nodes.  No reference counterpart (the reference has no learned
// part); plain bilinear velint stays the fallback.
//
// Every convolution is an implicit GEMM on the CDNA4 matrix cores
// (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation, K = 32 per
// instruction, one 3x3 tap of 32 channels):
//   forward      M = pixels, N = output channels, K = 9 taps x input channels
//   input grads  M = pixels, N = input channels,  K = 9 taps x output channels
//   weight grads M = output channels, N = 9 taps x input channels (+ a
//                column of ones for the bias), K = pixels
// A workgroup of four wavefronts owns a tile of 4 columns x 64 rows of
// pixels; the tile with its one-pixel halo (6 x 66 pixels x channels) is
// staged in LDS once with 16-byte loads, and every operand of the tile's
// matrix products is read from there.  The weights of a layer sit in
// registers.  Weight gradients are reduced per workgroup in registers and
// summed over the workgroups in a fixed order (deterministic).
// The grid is doubly periodic (pixel p = ix * n + iy, neighbours wrap).
// Activations and gradients between layers are BF16; pre-activations, the
// network output, master weights and the Adam moments are FP32.
#include <math.h>

#include "hyd_internal.h"

namespace hyd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// D += A B for one 16x16 block with K = 32.  Lane l: A row l % 16 and B
// column l % 16, k = 8 (l / 16) .. + 7; D rows 4 (l / 16) + i, column l % 16.
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
  union {
    uint4 u;
    bf16x8 b;
  } x;
  x.u = v;
  return x.b;
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// tiles: TW columns (ix) x TH rows (iy) of pixels per workgroup, one
// wavefront per column
constexpr int kTW = 4, kTH = 64, kHW = kTW + 2, kHH = kTH + 2;

// Stage the tile (ix0.., iy0..) of a [pixels][C] BF16 tensor with its halo
// into LDS as [hx][hy][C] (hx, hy = 0 .. kHW-1, kHH-1; wrapping).
template <int C>
__device__ __forceinline__ void stage_halo(const uint16_t *__restrict__ in, int n, int ix0,
                                           int iy0, uint16_t *tile) {
  constexpr int UPP = C * 2 / 8;  // 8-byte units per pixel (C = 4: 1, C = 32: 8)
  constexpr int NU = kHW * kHH * UPP;
  for (int u = threadIdx.x; u < NU; u += blockDim.x) {
    const int q = u / UPP, part = u - q * UPP;
    const int hx = q / kHH, hy = q - hx * kHH;
    int gx = ix0 - 1 + hx, gy = iy0 - 1 + hy;
    gx = gx < 0 ? gx + n : (gx >= n ? gx - n : gx);
    gy %= n;
    gy = gy < 0 ? gy + n : gy;
    const uint2 v = *(const uint2 *)(in + ((int64_t)gx * n + gy) * C + part * 4);
    *(uint2 *)(tile + (hx * kHH + hy) * C + part * 4) = v;
  }
}

// A operand of 16 pixels (rows r0 + l % 16 of column cx of the tile) for
// K block kb of a tensor with C channels, shifted by sgn * off(tap).
template <int C>
__device__ __forceinline__ bf16x8 a_operand(const uint16_t *tile, int cx, int r, int kb, int g,
                                            int sgn) {
  if (C == 32) {
    const int tap = kb, dx = sgn * (tap / 3 - 1), dy = sgn * (tap % 3 - 1);
    const uint16_t *src = tile + ((cx + 1 + dx) * kHH + (r + 1 + dy)) * 32 + 8 * g;
    return as_bf16x8(*(const uint4 *)src);
  } else {  // C == 4: the lane's eight k are two taps of four channels
    uint4 v = make_uint4(0, 0, 0, 0);
    const int ta = (kb * 32 + 8 * g) / 4;
    if (ta < 9) {
      const int dx = sgn * (ta / 3 - 1), dy = sgn * (ta % 3 - 1);
      const uint2 w = *(const uint2 *)(tile + ((cx + 1 + dx) * kHH + (r + 1 + dy)) * 4);
      v.x = w.x;
      v.y = w.y;
    }
    if (ta + 1 < 9) {
      const int dx = sgn * ((ta + 1) / 3 - 1), dy = sgn * ((ta + 1) % 3 - 1);
      const uint2 w = *(const uint2 *)(tile + ((cx + 1 + dx) * kHH + (r + 1 + dy)) * 4);
      v.z = w.x;
      v.w = w.y;
    }
    return as_bf16x8(v);
  }
}

// 3x3 convolution of a [pixels][CIN] BF16 tensor with packed weights w
// [NB*16][KP] (KP = 9 CIN padded to 32; row = output column of the GEMM):
//   MODE 0: forward, z = conv + bias (FP32 [pixels][cout]) and y = GELU(z)
//   MODE 1: forward of the last layer, y = z
//   MODE 2: input gradient (transposed convolution: taps shifted by -off),
//           out = conv * GELU'(zprev) as BF16 [pixels][cout]
template <int CIN, int NB, int MODE>
__global__ void __launch_bounds__(256) cnn_conv(const uint16_t *__restrict__ in,
                                                const uint16_t *__restrict__ w,
                                                const float *__restrict__ bias, int n, int cout,
                                                float *__restrict__ z, uint16_t *__restrict__ y,
                                                const float *__restrict__ zprev) {
  constexpr int KP = (9 * CIN + 31) / 32 * 32, KB = KP / 32;
  __shared__ __align__(16) uint16_t tile[kHW * kHH * CIN];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4;
  const int iy0 = blockIdx.x * kTH, ix0 = blockIdx.y * kTW;
  stage_halo<CIN>(in, n, ix0, iy0, tile);
  bf16x8 wr[KB][NB];
#pragma unroll
  for (int kb = 0; kb < KB; kb++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
      wr[kb][nb] = as_bf16x8(
          *(const uint4 *)(w + (int64_t)(nb * 16 + (l & 15)) * KP + kb * 32 + 8 * g));
  __syncthreads();
  const int sgn = MODE == 2 ? -1 : 1;
  f32x4 acc[4][NB];
#pragma unroll
  for (int mb = 0; mb < 4; mb++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[mb][nb] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < KB; kb++)
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      const bf16x8 a = a_operand<CIN>(tile, wv, mb * 16 + (l & 15), kb, g, sgn);
#pragma unroll
      for (int nb = 0; nb < NB; nb++) acc[mb][nb] = mfma32(a, wr[kb][nb], acc[mb][nb]);
    }
  const int ix = ix0 + wv;
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    const int co = nb * 16 + (l & 15);
    if (co >= cout) continue;
    const float b0 = MODE == 2 ? 0.f : bias[co];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int iy = iy0 + mb * 16 + 4 * g + i;
        if (iy >= n) continue;
        const int64_t o = ((int64_t)ix * n + iy) * cout + co;
        const float v = acc[mb][nb][i];
        if (MODE == 2) {
          y[o] = f2bf(v * gelu_grad(zprev[o]));
        } else {
          const float zz = v + b0;
          if (z) z[o] = zz;  // pre-activations are kept for training only
          if (y) y[o] = f2bf(MODE == 1 ? zz : gelu(zz));
        }
      }
  }
}

// Weight gradients: dW[co][k] = sum_p dz[p][co] x[p + off(tap(k))][ci(k)]
// for k = tap CIN + ci < 9 CIN, and the bias gradient sum_p dz[p][co] in
// column 9 CIN (a column of ones).  Each workgroup sums over its tiles
// (grid-stride) with the wavefronts sharing the N blocks, and writes its
// partial [MB*16][NK*16] block; cnn_wgrad_sum adds the partials in order.
template <int CIN, int COT, int MB, int NK>
__global__ void __launch_bounds__(256) cnn_wgrad(const uint16_t *__restrict__ dz,
                                                 const uint16_t *__restrict__ x, int n,
                                                 int ntiles, float *__restrict__ part) {
  __shared__ __align__(16) uint16_t xt[kHW * kHH * CIN];
  __shared__ __align__(16) uint16_t dzt[kTW * kTH * COT];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, g = l >> 4;
  constexpr int NPW = (NK + 3) / 4;  // N blocks per wavefront
  f32x4 acc[MB][NPW];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int j = 0; j < NPW; j++) acc[mb][j] = {0.f, 0.f, 0.f, 0.f};
  const int tiles_y = (n + kTH - 1) / kTH;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int iy0 = (tile % tiles_y) * kTH, ix0 = (tile / tiles_y) * kTW;
    __syncthreads();  // the previous tile's operands are no longer read
    stage_halo<CIN>(x, n, ix0, iy0, xt);
    // dz of the tile itself (no halo): kTW columns x kTH rows
    constexpr int UPP = COT * 2 / 8;
    for (int u = threadIdx.x; u < kTW * kTH * UPP; u += blockDim.x) {
      const int q = u / UPP, part8 = u - q * UPP;
      const int cx = q / kTH, r = q - cx * kTH;
      const int gy = (iy0 + r) % n;
      const uint2 v = (iy0 + r < n)
                          ? *(const uint2 *)(dz + ((int64_t)(ix0 + cx) * n + gy) * COT + part8 * 4)
                          : make_uint2(0, 0);  // rows past n contribute nothing
      *(uint2 *)(dzt + q * COT + part8 * 4) = v;
    }
    __syncthreads();
    // K = the tile's 256 pixels, 32 at a time: pixel index q = cx kTH + r
#pragma unroll 1
    for (int q0 = 0; q0 < kTW * kTH; q0 += 32) {
      bf16x8 a[MB];
#pragma unroll
      for (int mb = 0; mb < MB; mb++) {
        const int co = mb * 16 + (l & 15);
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const uint16_t v = co < COT ? dzt[(q0 + 8 * g + j) * COT + co] : (uint16_t)0;
          a[mb][j] = __builtin_bit_cast(__bf16, v);
        }
      }
#pragma unroll
      for (int jn = 0; jn < NPW; jn++) {
        const int nb = wv + 4 * jn;
        if (nb >= NK) continue;
        const int k = nb * 16 + (l & 15);
        const int tap = k / CIN, ci = k - tap * CIN;
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int q = q0 + 8 * g + j, cx = q / kTH, r = q - cx * kTH;
          uint16_t v;
          if (tap < 9) {
            const int dx = tap / 3 - 1, dy = tap % 3 - 1;
            v = xt[((cx + 1 + dx) * kHH + (r + 1 + dy)) * CIN + ci];
          } else {
            v = (k == 9 * CIN) ? (uint16_t)0x3f80 : (uint16_t)0;  // ones: the bias column
          }
          b[j] = __builtin_bit_cast(__bf16, v);
        }
#pragma unroll
        for (int mb = 0; mb < MB; mb++) acc[mb][jn] = mfma32(a[mb], b, acc[mb][jn]);
      }
    }
  }
  float *pw = part + (int64_t)blockIdx.x * (MB * 16) * (NK * 16);
#pragma unroll
  for (int jn = 0; jn < NPW; jn++) {
    const int nb = wv + 4 * jn;
    if (nb >= NK) continue;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        pw[(int64_t)(mb * 16 + 4 * g + i) * (NK * 16) + nb * 16 + (l & 15)] = acc[mb][jn][i];
  }
}

// Sum the workgroups' partials in a fixed order (8 strided groups of
// workgroups, then the 8 group sums in order); unpack into the master
// layout W[co][ci][tap] and the biases of one layer.  32 weights per
// workgroup of 256 threads: lane e % 32, group threadIdx.x / 32.
constexpr int kSumLanes = 32, kSumGroups = 8;
__global__ void __launch_bounds__(256) cnn_wgrad_sum(const float *__restrict__ part, int nparts,
                                                     int rows, int cols, int ci_n, int co_n,
                                                     float *__restrict__ gW,
                                                     float *__restrict__ gb) {
  __shared__ float red[kSumGroups][kSumLanes];
  const int lane = threadIdx.x % kSumLanes, grp = threadIdx.x / kSumLanes;
  const int e = blockIdx.x * kSumLanes + lane;
  const int cnt = co_n * (9 * ci_n + 1);
  const int co = e / (9 * ci_n + 1), k = e - co * (9 * ci_n + 1);
  float s = 0.f;
  if (e < cnt)
    for (int pidx = grp; pidx < nparts; pidx += kSumGroups)
      s += part[((int64_t)pidx * rows + co) * cols + k];
  red[grp][lane] = s;
  __syncthreads();
  if (grp != 0 || e >= cnt) return;
  s = red[0][lane];
#pragma unroll
  for (int j = 1; j < kSumGroups; j++) s += red[j][lane];
  if (k == 9 * ci_n) {
    gb[co] = s;
  } else {
    const int tap = k / ci_n, ci = k - tap * ci_n;
    gW[((int64_t)co * ci_n + ci) * 9 + tap] = s;
  }
}

// ---- inputs, sampling, loss ------------------------------------------------
__global__ void cnn_pack_input(const float *__restrict__ in3, int P, uint16_t *__restrict__ x) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  for (int c = 0; c < 3; c++) x[(int64_t)p * 4 + c] = f2bf(in3[(int64_t)p * 3 + c]);
  x[(int64_t)p * 4 + 3] = 0;
}

struct CnnGeom {
  double xmin, ymin, gli;
  int n;
};

__device__ __forceinline__ void bilinear(double x, double y, const CnnGeom &g, int idx[4],
                                         float w[4]) {
  const double xx = g.gli * (x - g.xmin), yy = g.gli * (y - g.ymin);
  int ix0 = (int)floor(xx), iy0 = (int)floor(yy);
  const float px = (float)(xx - ix0), py = (float)(yy - iy0);
  ix0 = ((ix0 % g.n) + g.n) % g.n;
  iy0 = ((iy0 % g.n) + g.n) % g.n;
  const int ix1 = ix0 + 1 == g.n ? 0 : ix0 + 1, iy1 = iy0 + 1 == g.n ? 0 : iy0 + 1;
  idx[0] = ix0 * g.n + iy0;
  idx[1] = ix1 * g.n + iy0;
  idx[2] = ix0 * g.n + iy1;
  idx[3] = ix1 * g.n + iy1;
  w[0] = (1.f - px) * (1.f - py);
  w[1] = px * (1.f - py);
  w[2] = (1.f - px) * py;
  w[3] = px * py;
}

// out [pixels][2] (FP32) sampled at the nodes; with targets, the loss
// sum (pred - t)^2 / m accumulated in *loss and its gradient scattered into
// gout [pixels][2].
__global__ void cnn_sample(const float *__restrict__ out, const double *__restrict__ xn,
                           const double *__restrict__ yn, int m, CnnGeom g,
                           float *__restrict__ pu, float *__restrict__ pv,
                           const float *__restrict__ tu, const float *__restrict__ tv,
                           float *__restrict__ gout, float *__restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (i < m) {
    int idx[4];
    float w[4];
    bilinear(xn[i], yn[i], g, idx, w);
    float u = 0.f, v = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      u += w[j] * out[(int64_t)idx[j] * 2];
      v += w[j] * out[(int64_t)idx[j] * 2 + 1];
    }
    if (pu) pu[i] = u;
    if (pv) pv[i] = v;
    if (tu) {
      const float du = u - tu[i], dv = v - tv[i];
      e = du * du + dv * dv;
      const float s = 2.f / (float)m;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        atomicAdd(&gout[(int64_t)idx[j] * 2], s * du * w[j]);
        atomicAdd(&gout[(int64_t)idx[j] * 2 + 1], s * dv * w[j]);
      }
    }
  }
  if (tu) {
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, e / (float)m);
  }
}

__global__ void cnn_to_bf16(const float *__restrict__ a, int64_t cnt, uint16_t *__restrict__ b) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < cnt) b[k] = f2bf(a[k]);
}

__global__ void cnn_zero(float *__restrict__ a, int64_t cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < cnt) a[k] = 0.f;
}

// Adam on the FP32 master weights (all layers in one array), then the BF16
// operand layouts: wf[co][tap CI + ci] (forward) and wb[ci][tap CO + co]
// (input gradients) of every layer.
struct LayerDims {
  int ci, co, cop, off;  // channels in/out, out padded to 16, offset in the master array
};
struct CnnLayers {
  LayerDims L[4];
  int wf_off[4], wb_off[4];
};

__global__ void cnn_adam(float *__restrict__ w, float *__restrict__ m1, float *__restrict__ m2,
                         const float *__restrict__ gw, int64_t cnt, float lr, float b1, float b2,
                         float c1, float c2, float eps) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const float gk = gw[k];
  const float a = b1 * m1[k] + (1.f - b1) * gk;
  const float v = b2 * m2[k] + (1.f - b2) * gk * gk;
  m1[k] = a;
  m2[k] = v;
  w[k] = w[k] - lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// master layout per layer: W[co][ci][tap] then bias[co].  Matrix-core
// operand layouts (rows padded to 16, K padded to 32 with zeros):
//   wf[co][k], k = tap CI + ci            (forward, B = W^T)
//   wb[ci][k], k = tap COT + co, COT = co padded to 4   (input gradients)
__global__ void cnn_pack_weights(const float *__restrict__ w, CnnLayers d, uint16_t *wf,
                                 uint16_t *wb) {
  const int layer = blockIdx.y;
  const LayerDims D = d.L[layer];
  const float *W = w + D.off;
  const int Kf = 9 * D.ci, KPf = (Kf + 31) / 32 * 32, totf = D.cop * KPf;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < totf; e += gridDim.x * blockDim.x) {
    const int co = e / KPf, k = e - co * KPf, tap = k / D.ci, ci = k - tap * D.ci;
    const float v = (co < D.co && k < Kf) ? W[((int64_t)co * D.ci + ci) * 9 + tap] : 0.f;
    wf[d.wf_off[layer] + e] = f2bf(v);
  }
  const int cot = (D.co + 3) / 4 * 4, Kb = 9 * cot, KPb = (Kb + 31) / 32 * 32;
  const int totb = ((D.ci + 15) / 16 * 16) * KPb;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < totb; e += gridDim.x * blockDim.x) {
    const int ci = e / KPb, k = e - ci * KPb, tap = k / cot, co = k - tap * cot;
    const float v =
        (ci < D.ci && k < Kb && co < D.co) ? W[((int64_t)co * D.ci + ci) * 9 + tap] : 0.f;
    wb[d.wb_off[layer] + e] = f2bf(v);
  }
}

// Widen a [pixels][2] gradient to [pixels][4] BF16 (the last layer's output
// channels padded to 4 for the input-gradient GEMM).
__global__ void cnn_pack_dz_last(const float *__restrict__ g2, int P, uint16_t *__restrict__ d4) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  d4[(int64_t)p * 4] = f2bf(g2[(int64_t)p * 2]);
  d4[(int64_t)p * 4 + 1] = f2bf(g2[(int64_t)p * 2 + 1]);
  d4[(int64_t)p * 4 + 2] = 0;
  d4[(int64_t)p * 4 + 3] = 0;
}

// Probe of the matrix-core operand layout (used by the tests): D = A B for
// one 16x16x32 block, A row-major [16][32], B N-major [16][32] (b[n][k]),
// D row-major [16][16], read and written as the kernels above do.
__global__ void cnn_mfma_probe(const uint16_t *a, const uint16_t *b, float *d) {
  const int l = threadIdx.x & 63, g = l >> 4;
  bf16x8 av, bv;
  for (int j = 0; j < 8; j++) {
    av[j] = __builtin_bit_cast(__bf16, a[(l & 15) * 32 + 8 * g + j]);
    bv[j] = __builtin_bit_cast(__bf16, b[(l & 15) * 32 + 8 * g + j]);
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma32(av, bv, acc);
  for (int i = 0; i < 4; i++) d[(4 * g + i) * 16 + (l & 15)] = acc[i];
}

// ---- host side --------------------------------------------------------------
struct Cnn {
  int n = 0, P = 0, dev = 0;
  hipStream_t s = nullptr;
  CnnLayers dims{};
  int64_t nw = 0;                     // master weights incl. biases
  float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *gw = nullptr;
  uint16_t *wf = nullptr, *wb = nullptr;
  float *part = nullptr;              // per-workgroup weight-gradient partials
  int nparts = 0;                     // workgroups of the weight-gradient kernels
  uint16_t *x0 = nullptr;             // packed input [P][4]
  float *z[4] = {};                   // pre-activations
  uint16_t *y[4] = {};                // activations (BF16)
  uint16_t *dz[4] = {};               // pre-activation gradients (BF16)
  float *gout = nullptr;              // d loss / d output [P][2]
  float *in3 = nullptr;               // staging of the 3-channel input
  double *xn = nullptr, *yn = nullptr;
  float *tu = nullptr, *tv = nullptr, *pu = nullptr, *pv = nullptr, *loss = nullptr;
  int cap_nodes = 0;
  int64_t t = 0;                      // Adam step count
  CnnGeom geom{};
};

}  // namespace hyd

struct hyd_cnn {
  hyd::Cnn c;
};

using namespace hyd;

static const int kCi[4] = {4, 32, 32, 32}, kCo[4] = {32, 32, 32, 2};

static void cnn_free(Cnn &c) {
  void *ptrs[] = {c.w, c.m1, c.m2, c.gw, c.wf, c.wb, c.part, c.x0, c.gout, c.in3, c.xn, c.yn,
                  c.tu, c.tv, c.pu, c.pv, c.loss};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (int i = 0; i < 4; i++) {
    if (c.z[i]) hipFree(c.z[i]);
    if (c.y[i]) hipFree(c.y[i]);
    if (c.dz[i]) hipFree(c.dz[i]);
  }
  if (c.s) hipStreamDestroy(c.s);
}

template <class T>
static int cmalloc(T **p, int64_t n) {
  HYD_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
  return HYD_OK;
}

static int cnn_nodes_cap(Cnn &c, int m) {
  if (m <= c.cap_nodes) return HYD_OK;
  void *ptrs[] = {c.xn, c.yn, c.tu, c.tv, c.pu, c.pv};
  for (void *p : ptrs)
    if (p) hipFree(p);
  c.xn = c.yn = nullptr;
  c.tu = c.tv = c.pu = c.pv = nullptr;
  c.cap_nodes = 0;
  HYD_CHECK(cmalloc(&c.xn, m));
  HYD_CHECK(cmalloc(&c.yn, m));
  HYD_CHECK(cmalloc(&c.tu, m));
  HYD_CHECK(cmalloc(&c.tv, m));
  HYD_CHECK(cmalloc(&c.pu, m));
  HYD_CHECK(cmalloc(&c.pv, m));
  c.cap_nodes = m;
  return HYD_OK;
}

extern "C" int hyd_cnn_create(int32_t n, double ell, uint64_t seed, int32_t device, hyd_cnn **out) {
  if (!out || n < 4 || n % 4 != 0 || n > 4096 || !(ell > 0.0)) {
    set_error("cnn_create: n must be a multiple of 4 in [4, 4096], ell > 0");
    return HYD_EARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    set_error("cnn_create: no HIP device");
    return HYD_EHIP;
  }
  HYD_HIP(hipSetDevice(device));
  hyd_cnn *h = new hyd_cnn();
  Cnn &c = h->c;
  c.n = n;
  c.P = n * n;
  c.dev = device;
  c.geom.n = n;
  c.geom.xmin = -0.5 * ell;
  c.geom.ymin = -0.5 * ell;
  c.geom.gli = n / ell;
  int64_t off = 0, wfo = 0, wbo = 0;
  for (int i = 0; i < 4; i++) {
    LayerDims &D = c.dims.L[i];
    D.ci = kCi[i];
    D.co = kCo[i];
    D.cop = (D.co + 15) / 16 * 16;
    D.off = (int)off;
    off += (int64_t)D.co * D.ci * 9 + D.co;  // weights, then biases
    c.dims.wf_off[i] = (int)wfo;
    wfo += (int64_t)D.cop * ((9 * D.ci + 31) / 32 * 32);
    c.dims.wb_off[i] = (int)wbo;
    wbo += (int64_t)((D.ci + 15) / 16 * 16) * ((9 * ((D.co + 3) / 4 * 4) + 31) / 32 * 32);
  }
  c.nw = off;
  // weight gradients: at most 256 workgroups, each a partial of at most
  // 32 x 304 values (layers 2-3: 32 output channels, 9 x 32 + 1 columns
  // padded to 16)
  const int ntiles = ((n + kTH - 1) / kTH) * (n / kTW);
  c.nparts = std::min(ntiles, 256);
  int rc = HYD_OK;
  auto T = [&](int r) {
    if (rc == HYD_OK) rc = r;
  };
  if (hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking) != hipSuccess) T(HYD_EHIP);
  T(cmalloc(&c.w, c.nw));
  T(cmalloc(&c.m1, c.nw));
  T(cmalloc(&c.m2, c.nw));
  T(cmalloc(&c.gw, c.nw));
  T(cmalloc(&c.wf, wfo));
  T(cmalloc(&c.wb, wbo));
  T(cmalloc(&c.part, (int64_t)c.nparts * 32 * 304));
  T(cmalloc(&c.x0, (int64_t)c.P * 4));
  T(cmalloc(&c.gout, (int64_t)c.P * 2));
  T(cmalloc(&c.in3, (int64_t)c.P * 3));
  T(cmalloc(&c.loss, 1));
  for (int i = 0; i < 4 && rc == HYD_OK; i++) {
    const int co = kCo[i];
    const int co4 = (co + 3) / 4 * 4;
    T(cmalloc(&c.z[i], (int64_t)c.P * co));
    T(cmalloc(&c.y[i], (int64_t)c.P * co));
    T(cmalloc(&c.dz[i], (int64_t)c.P * co4));
  }
  if (rc != HYD_OK) {
    cnn_free(c);
    delete h;
    return rc;
  }
  // He initialisation (normal, std sqrt(2 / fan_in)), zero biases; a
  // splitmix/Box-Muller stream from the seed
  std::vector<float> w(c.nw, 0.f);
  uint64_t st = seed ? seed : 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    st += 0x9e3779b97f4a7c15ull;
    uint64_t zz = st;
    zz = (zz ^ (zz >> 30)) * 0xbf58476d1ce4e5b9ull;
    zz = (zz ^ (zz >> 27)) * 0x94d049bb133111ebull;
    zz ^= zz >> 31;
    return ((zz >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < 4; i++) {
    const LayerDims &D = c.dims.L[i];
    const int fan_in = 9 * (i == 0 ? 3 : D.ci);
    const double sd = (i == 3 ? 0.1 : 1.0) * sqrt(2.0 / fan_in);
    for (int co = 0; co < D.co; co++)
      for (int ci = 0; ci < D.ci; ci++)
        for (int tap = 0; tap < 9; tap++) {
          const double u1 = next(), u2 = next();
          const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
          // the fourth input channel of layer 1 is padding
          w[D.off + ((int64_t)co * D.ci + ci) * 9 + tap] =
              (i == 0 && ci == 3) ? 0.f : (float)(sd * g);
        }
  }
  rc = HYD_OK;
  if (hipMemcpy(c.w, w.data(), sizeof(float) * c.nw, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c.m1, 0, sizeof(float) * c.nw) != hipSuccess ||
      hipMemset(c.m2, 0, sizeof(float) * c.nw) != hipSuccess)
    rc = HYD_EHIP;
  if (rc == HYD_OK) {
    hipLaunchKernelGGL(cnn_pack_weights, dim3(64, 4), dim3(256), 0, c.s, c.w, c.dims, c.wf, c.wb);
    if (hipStreamSynchronize(c.s) != hipSuccess) rc = HYD_EHIP;
  }
  if (rc != HYD_OK) {
    set_error("cnn_create: device set-up failed");
    cnn_free(c);
    delete h;
    return rc;
  }
  *out = h;
  return HYD_OK;
}

extern "C" int hyd_cnn_destroy(hyd_cnn *h) {
  if (!h) return HYD_OK;
  hipSetDevice(h->c.dev);
  cnn_free(h->c);
  delete h;
  return HYD_OK;
}

static int cnn_forward(Cnn &c) {
  const int P = c.P;
  hipLaunchKernelGGL(cnn_pack_input, dim3((P + 255) / 256), dim3(256), 0, c.s, c.in3, P, c.x0);
  const uint16_t *wf = c.wf;
  const float *w = c.w;
  auto bias = [&](int i) { return w + c.dims.L[i].off + (int64_t)c.dims.L[i].co * c.dims.L[i].ci * 9; };
  const dim3 grid((c.n + kTH - 1) / kTH, c.n / kTW);
  hipLaunchKernelGGL((cnn_conv<4, 2, 0>), grid, dim3(256), 0, c.s, c.x0, wf + c.dims.wf_off[0],
                     bias(0), c.n, 32, c.z[0], c.y[0], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 2, 0>), grid, dim3(256), 0, c.s, c.y[0],
                     wf + c.dims.wf_off[1], bias(1), c.n, 32, c.z[1], c.y[1], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 2, 0>), grid, dim3(256), 0, c.s, c.y[1],
                     wf + c.dims.wf_off[2], bias(2), c.n, 32, c.z[2], c.y[2], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 1, 1>), grid, dim3(256), 0, c.s, c.y[2],
                     wf + c.dims.wf_off[3], bias(3), c.n, 2, c.z[3], c.y[3], nullptr);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static int cnn_upload(Cnn &c, const float *in3, const double *xn, const double *yn, int m,
                      const float *tu, const float *tv) {
  HYD_CHECK(cnn_nodes_cap(c, m));
  HYD_HIP(hipMemcpyAsync(c.in3, in3, sizeof(float) * 3 * c.P, hipMemcpyHostToDevice, c.s));
  if (m > 0) {
    HYD_HIP(hipMemcpyAsync(c.xn, xn, sizeof(double) * m, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(c.yn, yn, sizeof(double) * m, hipMemcpyHostToDevice, c.s));
    if (tu) {
      HYD_HIP(hipMemcpyAsync(c.tu, tu, sizeof(float) * m, hipMemcpyHostToDevice, c.s));
      HYD_HIP(hipMemcpyAsync(c.tv, tv, sizeof(float) * m, hipMemcpyHostToDevice, c.s));
    }
  }
  return HYD_OK;
}

// Loss and gradients of all weights (FP32, master layout) into c.gw.
static int cnn_backward(Cnn &c, int m) {
  const int P = c.P, nblk = (P / 16 + 3) / 4;
  hipLaunchKernelGGL(cnn_zero, dim3((2 * P + 255) / 256), dim3(256), 0, c.s, c.gout,
                     (int64_t)2 * P);
  hipLaunchKernelGGL(cnn_zero, dim3(1), dim3(64), 0, c.s, c.loss, (int64_t)1);
  hipLaunchKernelGGL(cnn_zero, dim3((unsigned)((c.nw + 255) / 256)), dim3(256), 0, c.s, c.gw,
                     c.nw);
  if (m > 0)
    hipLaunchKernelGGL(cnn_sample, dim3((m + 255) / 256), dim3(256), 0, c.s, c.z[3], c.xn, c.yn,
                       m, c.geom, c.pu, c.pv, c.tu, c.tv, c.gout, c.loss);
  hipLaunchKernelGGL(cnn_pack_dz_last, dim3((P + 255) / 256), dim3(256), 0, c.s, c.gout, P,
                     c.dz[3]);
  const uint16_t *wb = c.wb;
  const dim3 grid((c.n + kTH - 1) / kTH, c.n / kTW);
  // input gradients back through layers 4..2 (layer 1's input needs none)
  hipLaunchKernelGGL((cnn_conv<4, 2, 2>), grid, dim3(256), 0, c.s, c.dz[3],
                     wb + c.dims.wb_off[3], nullptr, c.n, 32, nullptr, c.dz[2], c.z[2]);
  hipLaunchKernelGGL((cnn_conv<32, 2, 2>), grid, dim3(256), 0, c.s, c.dz[2],
                     wb + c.dims.wb_off[2], nullptr, c.n, 32, nullptr, c.dz[1], c.z[1]);
  hipLaunchKernelGGL((cnn_conv<32, 2, 2>), grid, dim3(256), 0, c.s, c.dz[1],
                     wb + c.dims.wb_off[1], nullptr, c.n, 32, nullptr, c.dz[0], c.z[0]);
  // weight and bias gradients: partials per workgroup, summed in order
  const int ntiles = ((c.n + kTH - 1) / kTH) * (c.n / kTW);
  const uint16_t *xin[4] = {c.x0, c.y[0], c.y[1], c.y[2]};
  for (int i = 0; i < 4; i++) {
    const LayerDims &D = c.dims.L[i];
    int rows = 0, cols = 0;
    if (i == 0) {
      hipLaunchKernelGGL((cnn_wgrad<4, 32, 2, 3>), dim3(c.nparts), dim3(256), 0, c.s, c.dz[0],
                         xin[0], c.n, ntiles, c.part);
      rows = 32, cols = 48;
    } else if (i < 3) {
      hipLaunchKernelGGL((cnn_wgrad<32, 32, 2, 19>), dim3(c.nparts), dim3(256), 0, c.s, c.dz[i],
                         xin[i], c.n, ntiles, c.part);
      rows = 32, cols = 304;
    } else {
      hipLaunchKernelGGL((cnn_wgrad<32, 4, 1, 19>), dim3(c.nparts), dim3(256), 0, c.s, c.dz[3],
                         xin[3], c.n, ntiles, c.part);
      rows = 16, cols = 304;
    }
    const int cnt = D.co * (9 * D.ci + 1);
    hipLaunchKernelGGL(cnn_wgrad_sum, dim3((cnt + kSumLanes - 1) / kSumLanes),
                       dim3(kSumLanes * kSumGroups), 0, c.s, c.part,
                       c.nparts, rows, cols, D.ci, D.co, c.gw + D.off,
                       c.gw + D.off + (int64_t)D.co * D.ci * 9);
  }
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static Cnn *cnn_of(hyd_cnn *h) {
  if (h) hipSetDevice(h->c.dev);
  return h ? &h->c : nullptr;
}

extern "C" int hyd_cnn_forward(hyd_cnn *h, const float *in3, const double *xn,
                               const double *yn, int32_t m, float *du, float *dv) {
  Cnn *c = cnn_of(h);
  if (!c || !in3 || m < 0 || (m > 0 && (!xn || !yn || !du || !dv))) {
    set_error("cnn_forward: bad arguments");
    return HYD_EARG;
  }
  HYD_CHECK(cnn_upload(*c, in3, xn, yn, m, nullptr, nullptr));
  HYD_CHECK(cnn_forward(*c));
  if (m > 0) {
    hipLaunchKernelGGL(cnn_sample, dim3((m + 255) / 256), dim3(256), 0, c->s, c->z[3], c->xn,
                       c->yn, m, c->geom, c->pu, c->pv, nullptr, nullptr, nullptr, nullptr);
    HYD_HIP(hipMemcpyAsync(du, c->pu, sizeof(float) * m, hipMemcpyDeviceToHost, c->s));
    HYD_HIP(hipMemcpyAsync(dv, c->pv, sizeof(float) * m, hipMemcpyDeviceToHost, c->s));
  }
  HYD_HIP(hipStreamSynchronize(c->s));
  return HYD_OK;
}

extern "C" int hyd_cnn_gradients(hyd_cnn *h, const float *in3, const double *xn,
                                 const double *yn, int32_t m, const float *tu, const float *tv,
                                 float *grad, float *loss) {
  Cnn *c = cnn_of(h);
  if (!c || !in3 || m <= 0 || !xn || !yn || !tu || !tv || !grad || !loss) {
    set_error("cnn_gradients: bad arguments");
    return HYD_EARG;
  }
  HYD_CHECK(cnn_upload(*c, in3, xn, yn, m, tu, tv));
  HYD_CHECK(cnn_forward(*c));
  HYD_CHECK(cnn_backward(*c, m));
  HYD_HIP(hipMemcpyAsync(grad, c->gw, sizeof(float) * c->nw, hipMemcpyDeviceToHost, c->s));
  HYD_HIP(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->s));
  HYD_HIP(hipStreamSynchronize(c->s));
  return HYD_OK;
}

extern "C" int hyd_cnn_train_step(hyd_cnn *h, const float *in3, const double *xn,
                                  const double *yn, int32_t m, const float *tu, const float *tv,
                                  float lr, float *loss) {
  Cnn *c = cnn_of(h);
  if (!c || !in3 || m <= 0 || !xn || !yn || !tu || !tv || !(lr > 0.f)) {
    set_error("cnn_train_step: bad arguments");
    return HYD_EARG;
  }
  HYD_CHECK(cnn_upload(*c, in3, xn, yn, m, tu, tv));
  HYD_CHECK(cnn_forward(*c));
  HYD_CHECK(cnn_backward(*c, m));
  c->t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)c->t), c2 = 1.f - powf(b2, (float)c->t);
  hipLaunchKernelGGL(cnn_adam, dim3((unsigned)((c->nw + 255) / 256)), dim3(256), 0, c->s, c->w,
                     c->m1, c->m2, c->gw, c->nw, lr, b1, b2, c1, c2, 1e-8f);
  hipLaunchKernelGGL(cnn_pack_weights, dim3(64, 4), dim3(256), 0, c->s, c->w, c->dims, c->wf,
                     c->wb);
  HYD_HIP(hipGetLastError());
  if (loss) HYD_HIP(hipMemcpyAsync(loss, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->s));
  HYD_HIP(hipStreamSynchronize(c->s));
  return HYD_OK;
}

extern "C" int hyd_cnn_num_weights(hyd_cnn *h, int64_t *nw) {
  Cnn *c = cnn_of(h);
  if (!c || !nw) {
    set_error("cnn_num_weights: bad arguments");
    return HYD_EARG;
  }
  *nw = c->nw;
  return HYD_OK;
}

extern "C" int hyd_cnn_get_weights(hyd_cnn *h, float *w) {
  Cnn *c = cnn_of(h);
  if (!c || !w) {
    set_error("cnn_get_weights: bad arguments");
    return HYD_EARG;
  }
  HYD_HIP(hipMemcpy(w, c->w, sizeof(float) * c->nw, hipMemcpyDeviceToHost));
  return HYD_OK;
}

extern "C" int hyd_cnn_set_weights(hyd_cnn *h, const float *w) {
  Cnn *c = cnn_of(h);
  if (!c || !w) {
    set_error("cnn_set_weights: bad arguments");
    return HYD_EARG;
  }
  HYD_HIP(hipMemcpy(c->w, w, sizeof(float) * c->nw, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(cnn_pack_weights, dim3(64, 4), dim3(256), 0, c->s, c->w, c->dims, c->wf,
                     c->wb);
  HYD_HIP(hipGetLastError());
  HYD_HIP(hipStreamSynchronize(c->s));
  c->t = 0;
  HYD_HIP(hipMemset(c->m1, 0, sizeof(float) * c->nw));
  HYD_HIP(hipMemset(c->m2, 0, sizeof(float) * c->nw));
  return HYD_OK;
}

// ---- the correction inside the time step ------------------------------------
namespace hyd {

template <class T>
__global__ void cnn_pack_fields(const T *__restrict__ uu, const T *__restrict__ vv,
                                const double *__restrict__ q, int P, double sui, double sqi,
                                uint16_t *__restrict__ x) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  x[(int64_t)p * 4] = f2bf((float)((double)uu[p] * sui));
  x[(int64_t)p * 4 + 1] = f2bf((float)((double)vv[p] * sui));
  x[(int64_t)p * 4 + 2] = f2bf((float)((double)q[p] * sqi));
  x[(int64_t)p * 4 + 3] = 0;
}

__global__ void cnn_sample_scaled(const float *__restrict__ out, const double *__restrict__ xn,
                                  const double *__restrict__ yn, int m, CnnGeom g, float scale,
                                  float *__restrict__ du, float *__restrict__ dv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int idx[4];
  float w[4];
  bilinear(xn[i], yn[i], g, idx, w);
  float u = 0.f, v = 0.f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    u += w[j] * out[(int64_t)idx[j] * 2];
    v += w[j] * out[(int64_t)idx[j] * 2 + 1];
  }
  du[i] = scale * u;
  dv[i] = scale * v;
}

}  // namespace hyd

// The correction (du, dv) at m nodes from the gridded u, v and PV q of a
// time-step stage (physical layout f[ix*n + iy]; u, v FP64 or FP32, q
// FP64), inputs
// scaled by 1/su and 1/sq as in training, the output by sd; everything on
// stream s (the caller's), using the network's own work buffers.
int hyd::cnn_correction(hyd_cnn *h, hipStream_t s, const void *uu, const void *vv,
                        const double *q, bool f32, double su, double sq, double sd, const double *xn,
                        const double *yn, int m, float *du, float *dv) {
  Cnn &c = h->c;
  const int P = c.P;
  const double sui = 1.0 / su, sqi = 1.0 / sq;
  if (f32)
    hipLaunchKernelGGL(cnn_pack_fields<float>, dim3((P + 255) / 256), dim3(256), 0, s,
                       (const float *)uu, (const float *)vv, q, P, sui, sqi, c.x0);
  else
    hipLaunchKernelGGL(cnn_pack_fields<double>, dim3((P + 255) / 256), dim3(256), 0, s,
                       (const double *)uu, (const double *)vv, q, P, sui, sqi, c.x0);
  const uint16_t *wf = c.wf;
  const float *w = c.w;
  auto bias = [&](int i) { return w + c.dims.L[i].off + (int64_t)c.dims.L[i].co * c.dims.L[i].ci * 9; };
  const dim3 grid((c.n + kTH - 1) / kTH, c.n / kTW);
  // inference: the hidden pre-activations and the last layer's BF16 copy
  // are not needed
  hipLaunchKernelGGL((cnn_conv<4, 2, 0>), grid, dim3(256), 0, s, c.x0, wf + c.dims.wf_off[0],
                     bias(0), c.n, 32, nullptr, c.y[0], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 2, 0>), grid, dim3(256), 0, s, c.y[0], wf + c.dims.wf_off[1],
                     bias(1), c.n, 32, nullptr, c.y[1], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 2, 0>), grid, dim3(256), 0, s, c.y[1], wf + c.dims.wf_off[2],
                     bias(2), c.n, 32, nullptr, c.y[2], nullptr);
  hipLaunchKernelGGL((cnn_conv<32, 1, 1>), grid, dim3(256), 0, s, c.y[2], wf + c.dims.wf_off[3],
                     bias(3), c.n, 2, c.z[3], nullptr, nullptr);
  if (m > 0)
    hipLaunchKernelGGL(cnn_sample_scaled, dim3((m + 255) / 256), dim3(256), 0, s, c.z[3], xn, yn,
                       m, c.geom, (float)sd, du, dv);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int hyd::cnn_grid(hyd_cnn *h) { return h ? h->c.n : 0; }

extern "C" int hyd_cnn_mfma_probe(const uint16_t *a, const uint16_t *b, float *d) {
  if (!a || !b || !d) {
    set_error("cnn_mfma_probe: bad arguments");
    return HYD_EARG;
  }
  uint16_t *da, *db;
  float *dd;
  HYD_HIP(hipMalloc((void **)&da, 1024));
  HYD_HIP(hipMalloc((void **)&db, 1024));
  HYD_HIP(hipMalloc((void **)&dd, 1024));
  int rc = HYD_OK;
  if (hipMemcpy(da, a, 1024, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, 1024, hipMemcpyHostToDevice) != hipSuccess)
    rc = HYD_EHIP;
  if (rc == HYD_OK) {
    hipLaunchKernelGGL(cnn_mfma_probe, dim3(1), dim3(64), 0, 0, da, db, dd);
    if (hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost) != hipSuccess) rc = HYD_EHIP;
  }
  hipFree(da);
  hipFree(db);
  hipFree(dd);
  if (rc != HYD_OK) set_error("cnn_mfma_probe: device error");
  return rc;
}
