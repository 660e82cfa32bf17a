// cnn.hip -- the learned sub-grid velocity correction: a 4-layer 3x3
// convolutional network on the coarse grid (inputs u, v, q; 32 channels;
// GELU), whose 2-channel output (du, dv) is sampled bilinearly at the
// contour nodes.  No reference counterpart (the reference has no learned
// part); plain bilinear velint stays the fallback.
//
// Every convolution is an implicit GEMM on the matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation), one
// wavefront per 16x16 output block:
//   forward      M = pixels, N = output channels, K = 9 taps x input channels
//   input grads  M = pixels, N = input channels,  K = 9 taps x output channels
//   weight grads M = output channels, N = 9 taps x input channels, K = pixels
// The grid is doubly periodic (pixel p = ix * n + iy, neighbours wrap).
// Activations and gradients between layers are BF16; pre-activations, the
// network output, master weights and the Adam moments are FP32.
#include <math.h>

#include "hyd_internal.h"

namespace hyd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// pixel p shifted by tap (0..8: dx = tap / 3 - 1 along ix, dy = tap % 3 - 1
// along iy), times sgn (+1 forward, -1 for the transposed convolution)
__device__ __forceinline__ int tap_pixel(int p, int tap, int sgn, int n) {
  const int ix = p / n, iy = p - ix * n;
  int jx = ix + sgn * (tap / 3 - 1), jy = iy + sgn * (tap % 3 - 1);
  jx = jx < 0 ? jx + n : (jx >= n ? jx - n : jx);
  jy = jy < 0 ? jy + n : (jy >= n ? jy - n : jy);
  return jx * n + jy;
}

// One 16x16 block of C = A B with A (M x K) gathered by a functor and B
// (K x N) stored N-major (b[n * ldb + k]).  Lane l: A row l % 16, B column
// l % 16, k = k0 + 4 (l / 16) .. + 3; D rows 4 (l / 16) + i, column l % 16.

// Forward convolution: in [pixels][CI] (BF16), wf [co][9 CI] (BF16, co
// padded to a multiple of 16), out z [pixels][CO] (FP32) and y = GELU(z)
// (or z when last) as BF16 [pixels][CO].
template <int CI, int CO, bool LAST>
__global__ void __launch_bounds__(256) cnn_conv_fwd(const uint16_t *__restrict__ in,
                                                    const uint16_t *__restrict__ wf,
                                                    const float *__restrict__ bias, int n,
                                                    float *__restrict__ z,
                                                    uint16_t *__restrict__ y) {
  constexpr int K = 9 * CI, NB = (CO + 15) / 16;
  const int l = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);  // 16-pixel block
  const int P = n * n;
  if (blk * 16 >= P) return;
  const int p_a = blk * 16 + (l & 15);  // A row of this lane
  const int g = l >> 4;
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int co_b = nb * 16 + (l & 15);  // B column of this lane
    for (int k0 = 0; k0 < K; k0 += 16) {
      const int k = k0 + 4 * g;  // 4 consecutive k of one tap (CI % 4 == 0)
      bf16x4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
      if (k < K) {
        const int tap = k / CI, ci = k - tap * CI;
        const uint16_t *src = in + (int64_t)tap_pixel(p_a, tap, 1, n) * CI + ci;
        const uint16_t *wsrc = wf + (int64_t)co_b * K + k;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          a[j] = (short)src[j];
          b[j] = (short)wsrc[j];
        }
      }
      acc = mfma16(a, b, acc);
    }
    const int co = nb * 16 + (l & 15);
    if (co < CO) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int p = blk * 16 + 4 * g + i;
        const float v = acc[i] + bias[co];
        z[(int64_t)p * CO + co] = v;
        y[(int64_t)p * CO + co] = f2bf(LAST ? v : gelu(v));
      }
    }
  }
}

// Input gradients: dx[p][ci] = sum_{tap,co} W[co][ci][tap] dz[p - off(tap)][co]
// with wb [ci][9 CO] (BF16); dz [pixels][CO] (BF16).  Writes dy_prev =
// dx * GELU'(z_prev) as BF16 (the previous layer's pre-activation grads).
template <int CI, int CO>
__global__ void __launch_bounds__(256) cnn_conv_bwd_data(const uint16_t *__restrict__ dz,
                                                         const uint16_t *__restrict__ wb, int n,
                                                         const float *__restrict__ z_prev,
                                                         uint16_t *__restrict__ dz_prev) {
  constexpr int K = 9 * CO, NB = (CI + 15) / 16;
  const int l = threadIdx.x & 63;
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int P = n * n;
  if (blk * 16 >= P) return;
  const int p_a = blk * 16 + (l & 15);
  const int g = l >> 4;
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int ci_b = nb * 16 + (l & 15);
    for (int k0 = 0; k0 < K; k0 += 16) {
      const int k = k0 + 4 * g;
      bf16x4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
      if (k < K) {
        const int tap = k / CO, co = k - tap * CO;
        const uint16_t *src = dz + (int64_t)tap_pixel(p_a, tap, -1, n) * CO + co;
        const uint16_t *wsrc = wb + (int64_t)ci_b * K + k;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          a[j] = (short)src[j];
          b[j] = (short)wsrc[j];
        }
      }
      acc = mfma16(a, b, acc);
    }
    const int ci = nb * 16 + (l & 15);
    if (ci < CI) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int p = blk * 16 + 4 * g + i;
        const float zz = z_prev[(int64_t)p * CI + ci];
        dz_prev[(int64_t)p * CI + ci] = f2bf(acc[i] * gelu_grad(zz));
      }
    }
  }
}

// Weight gradients: dW[co][k] += sum_p dz[p][co] x[p + off(tap(k))][ci(k)],
// one wavefront per (16 co) x (16 k) block and pixel slice; FP32 atomics
// into dw [co][9 CI]; block (0, *) also accumulates the bias gradient.  dz
// rows hold CO channels of which the first co_n are real.
template <int CI, int CO>
__global__ void __launch_bounds__(64) cnn_conv_bwd_weight(const uint16_t *__restrict__ dz,
                                                          const uint16_t *__restrict__ x, int n,
                                                          int pix_per_wave, int co_n,
                                                          float *__restrict__ dw,
                                                          float *__restrict__ db) {
  constexpr int K = 9 * CI;
  const int l = threadIdx.x & 63, g = l >> 4;
  const int kb = blockIdx.x, cb = blockIdx.y;
  const int P = n * n;
  const int p_begin = blockIdx.z * pix_per_wave, p_end = min(P, p_begin + pix_per_wave);
  const int co_a = cb * 16 + (l & 15);  // A row (output channel)
  const int k_b = kb * 16 + (l & 15);   // B column (tap, input channel)
  const int tap = k_b / CI, ci = k_b - (k_b / CI) * CI;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int p0 = p_begin; p0 < p_end; p0 += 16) {
    bf16x4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int p = p0 + 4 * g + j;
      if (p < p_end) {
        if (co_a < CO) a[j] = (short)dz[(int64_t)p * CO + co_a];
        if (k_b < K) b[j] = (short)x[(int64_t)tap_pixel(p, tap, 1, n) * CI + ci];
      }
    }
    if (kb == 0 && co_a < CO)
#pragma unroll
      for (int j = 0; j < 4; j++) bsum += bf2f((uint16_t)a[j]);
    acc = mfma16(a, b, acc);
  }
  const int k_out = kb * 16 + (l & 15);
  if (k_out < K)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = cb * 16 + 4 * g + i;
      if (co < co_n) atomicAdd(&dw[(int64_t)co * K + k_out], acc[i]);
    }
  if (kb == 0) {
    // lanes l, l+16, l+32, l+48 hold the same channel (every lane takes
    // part in the shuffles; padded channels hold zeros)
    bsum += __shfl_down(bsum, 32);
    bsum += __shfl_down(bsum, 16);
    if (g == 0 && co_a < co_n) atomicAdd(&db[co_a], bsum);
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

// master layout per layer: W[co][ci][tap] then bias[co]
__global__ void cnn_pack_weights(const float *__restrict__ w, CnnLayers d, uint16_t *wf,
                                 uint16_t *wb) {
  const int layer = blockIdx.y;
  const LayerDims D = d.L[layer];
  const int K = 9 * D.ci, total = D.cop * K;
  const float *W = w + D.off;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const int co = e / K, k = e - co * K, tap = k / D.ci, ci = k - tap * D.ci;
    const float v = co < D.co ? W[((int64_t)co * D.ci + ci) * 9 + tap] : 0.f;
    wf[d.wf_off[layer] + e] = f2bf(v);
  }
  // wb[ci][tap * COP4 + co], COP4 = co padded to 4 (k chunks of 4 in one tap)
  const int cop4 = (D.co + 3) / 4 * 4, Kb = 9 * cop4, totb = ((D.ci + 15) / 16 * 16) * Kb;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < totb; e += gridDim.x * blockDim.x) {
    const int ci = e / Kb, k = e - ci * Kb, tap = k / cop4, co = k - tap * cop4;
    const float v = (ci < D.ci && co < D.co) ? W[((int64_t)co * D.ci + ci) * 9 + tap] : 0.f;
    wb[d.wb_off[layer] + e] = f2bf(v);
  }
}

// Repack the weight gradient of one layer from the GEMM layout
// dw[co][tap CI + ci] into the master layout W[co][ci][tap].
__global__ void cnn_unpack_grad(const float *__restrict__ dwk, int ci_n, int co_n,
                                float *__restrict__ g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int K = 9 * ci_n;
  if (e >= co_n * K) return;
  const int co = e / K, k = e - co * K, tap = k / ci_n, ci = k - tap * ci_n;
  g[((int64_t)co * ci_n + ci) * 9 + tap] = dwk[e];
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
// one 16x16x16 block, A row-major [16][16], B N-major [16][16] (b[n][k]),
// D row-major, all read and written as the kernels above read them.
__global__ void cnn_mfma_probe(const uint16_t *a, const uint16_t *b, float *d) {
  const int l = threadIdx.x & 63, g = l >> 4;
  bf16x4 av, bv;
  for (int j = 0; j < 4; j++) {
    av[j] = (short)a[(l & 15) * 16 + 4 * g + j];
    bv[j] = (short)b[(l & 15) * 16 + 4 * g + j];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma16(av, bv, acc);
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
  float *dwk = nullptr;               // GEMM-layout weight gradient (largest layer)
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
  void *ptrs[] = {c.w, c.m1, c.m2, c.gw, c.wf, c.wb, c.dwk, c.x0, c.gout, c.in3, c.xn, c.yn,
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
  int64_t off = 0, wfo = 0, wbo = 0, maxk = 0;
  for (int i = 0; i < 4; i++) {
    LayerDims &D = c.dims.L[i];
    D.ci = kCi[i];
    D.co = kCo[i];
    D.cop = (D.co + 15) / 16 * 16;
    D.off = (int)off;
    off += (int64_t)D.co * D.ci * 9 + D.co;  // weights, then biases
    c.dims.wf_off[i] = (int)wfo;
    wfo += (int64_t)D.cop * 9 * D.ci;
    c.dims.wb_off[i] = (int)wbo;
    wbo += (int64_t)((D.ci + 15) / 16 * 16) * 9 * ((D.co + 3) / 4 * 4);
    maxk = std::max<int64_t>(maxk, (int64_t)D.co * 9 * D.ci);
  }
  c.nw = off;
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
  T(cmalloc(&c.dwk, maxk));
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
  const int P = c.P, nblk = (P / 16 + 3) / 4;
  hipLaunchKernelGGL(cnn_pack_input, dim3((P + 255) / 256), dim3(256), 0, c.s, c.in3, P, c.x0);
  const uint16_t *wf = c.wf;
  const float *w = c.w;
  auto bias = [&](int i) { return w + c.dims.L[i].off + (int64_t)c.dims.L[i].co * c.dims.L[i].ci * 9; };
  hipLaunchKernelGGL((cnn_conv_fwd<4, 32, false>), dim3(nblk), dim3(256), 0, c.s, c.x0,
                     wf + c.dims.wf_off[0], bias(0), c.n, c.z[0], c.y[0]);
  hipLaunchKernelGGL((cnn_conv_fwd<32, 32, false>), dim3(nblk), dim3(256), 0, c.s, c.y[0],
                     wf + c.dims.wf_off[1], bias(1), c.n, c.z[1], c.y[1]);
  hipLaunchKernelGGL((cnn_conv_fwd<32, 32, false>), dim3(nblk), dim3(256), 0, c.s, c.y[1],
                     wf + c.dims.wf_off[2], bias(2), c.n, c.z[2], c.y[2]);
  hipLaunchKernelGGL((cnn_conv_fwd<32, 2, true>), dim3(nblk), dim3(256), 0, c.s, c.y[2],
                     wf + c.dims.wf_off[3], bias(3), c.n, c.z[3], c.y[3]);
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
  // input gradients back through layers 4..2 (layer 1's input needs none)
  hipLaunchKernelGGL((cnn_conv_bwd_data<32, 4>), dim3(nblk), dim3(256), 0, c.s, c.dz[3],
                     wb + c.dims.wb_off[3], c.n, c.z[2], c.dz[2]);
  hipLaunchKernelGGL((cnn_conv_bwd_data<32, 32>), dim3(nblk), dim3(256), 0, c.s, c.dz[2],
                     wb + c.dims.wb_off[2], c.n, c.z[1], c.dz[1]);
  hipLaunchKernelGGL((cnn_conv_bwd_data<32, 32>), dim3(nblk), dim3(256), 0, c.s, c.dz[1],
                     wb + c.dims.wb_off[1], c.n, c.z[0], c.dz[0]);
  // weight gradients
  const int slices = std::max(1, std::min(256, P / 256));
  const int ppw = ((P + slices - 1) / slices + 15) / 16 * 16;
  const uint16_t *xin[4] = {c.x0, c.y[0], c.y[1], c.y[2]};
  for (int i = 0; i < 4; i++) {
    const LayerDims &D = c.dims.L[i];
    const int K = 9 * D.ci;
    hipLaunchKernelGGL(cnn_zero, dim3((unsigned)((D.co * K + 255) / 256)), dim3(256), 0, c.s,
                       c.dwk, (int64_t)D.co * K);
    float *db = c.gw + D.off + (int64_t)D.co * D.ci * 9;
    const dim3 grid((K + 15) / 16, (D.co + 15) / 16, (P + ppw - 1) / ppw);
    if (i == 0)
      hipLaunchKernelGGL((cnn_conv_bwd_weight<4, 32>), grid, dim3(64), 0, c.s, c.dz[0], xin[0],
                         c.n, ppw, D.co, c.dwk, db);
    else if (i < 3)
      hipLaunchKernelGGL((cnn_conv_bwd_weight<32, 32>), grid, dim3(64), 0, c.s, c.dz[i],
                         xin[i], c.n, ppw, D.co, c.dwk, db);
    else
      hipLaunchKernelGGL((cnn_conv_bwd_weight<32, 4>), grid, dim3(64), 0, c.s, c.dz[3], xin[3],
                         c.n, ppw, D.co, c.dwk, db);
    hipLaunchKernelGGL(cnn_unpack_grad, dim3((D.co * K + 255) / 256), dim3(256), 0, c.s, c.dwk,
                       D.ci, D.co, c.gw + D.off);
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

extern "C" int hyd_cnn_mfma_probe(const uint16_t *a, const uint16_t *b, float *d) {
  if (!a || !b || !d) {
    set_error("cnn_mfma_probe: bad arguments");
    return HYD_EARG;
  }
  uint16_t *da, *db;
  float *dd;
  HYD_HIP(hipMalloc((void **)&da, 512));
  HYD_HIP(hipMalloc((void **)&db, 512));
  HYD_HIP(hipMalloc((void **)&dd, 1024));
  int rc = HYD_OK;
  if (hipMemcpy(da, a, 512, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, 512, hipMemcpyHostToDevice) != hipSuccess)
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
