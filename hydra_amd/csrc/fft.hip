// fft.hip -- stafft-faithful real FFTs on gfx950.
//
// Each launch transforms a batch of m vectors of length n (a power of two)
// entirely in LDS: load once (optionally transposed and with a fused
// spectral derivative / revfft pre-scaling), run every radix stage between
// two LDS buffers, store once (optionally transposed, with the 1/sqrt(n) or
// 2/sqrt(n) normalisation and a sign folded in).
//
// The butterflies are those of forrdx2/forrdx4 (stafft.f90:890-997,
// 1080-1130) and revrdx4/revrdx2 (stafft.f90:1463-1571, 1656-1706), applied
// in the factor order of forfft/revfft (stafft.f90:212-271, 331-390), with
// the host-built trig table, evaluated with the same expression order and
// no FMA contraction.  The result is therefore bit-identical to the
// reference, while the data movement is one HBM read and one HBM write per
// 1D transform batch.
//
// Layouts: "strided" means element e of vector v is x[v + m*e] (the
// reference's x(m,n) with the vector index fastest); "rows" means
// x[e + n*v].  ptospc = x-transform reading strided (ny vectors) and writing
// rows (this is the transpose of sta2dfft.f90:122-126), then an in-place
// strided y-transform.  spctop = strided y-inverse into a work array, then
// an x-inverse reading rows and writing strided (sta2dfft.f90:152-162).
#include <hip/hip_ext.h>

#include "hyd_internal.h"

namespace hyd {

// Up to kMaxF fields of the same shape and direction per launch
// (blockIdx.y selects the field); each field has its own input, output,
// fused derivative and scaling.
static constexpr int kMaxF = 4;
static constexpr int kMaxV = 8;  // vectors per block: 1, 2, 4 or 8
// threads per block of every FFT launch (a constant, so that the item loops
// of a stage have compile-time trip counts)
static constexpr int kFftThreads = 512;
// Twiddles read from the (L1/L2-resident) global table rather than staged in
// LDS: the smaller LDS footprint lets more blocks share a CU.
static constexpr bool kTrigInLds = false;
static size_t fft_lds_bytes(int n, int V, size_t tsize) {
  return tsize * (2 * (size_t)V * n + (kTrigInLds ? 2 * (size_t)n : 0));
}
struct FftArgs {
  const void *in[kMaxF];    // TI
  void *out[kMaxF];         // TO
  const double *dk[kMaxF];  // filtered wavenumbers for a fused derivative
  int pre[kMaxF];           // 0 none, 1 derivative across vectors, 2 along elements
  double post[kMaxF];       // normalisation (and sign) folded into the store
  const void *trig;         // T, 2n: cos then -sin (stafft.f90:103-122)
  const double *sub;        // forward: subtract *sub / subdiv from every input value
  double subdiv;
  int n, m, V, lgV;
  int in_rows, out_rows;
  int n2, n4;
};

// ---------------------------------------------------------------------------
// Radix stages.  The butterflies of one stage fall into three kinds (k = 0,
// the pairs k / lv-k, and the middle k = lv/2 for even lv); each kind has
// its own loop so that a wavefront never steps through another kind's code.
// Items of a kind are numbered with the offset (irem / ic) fastest, then k,
// then the vector.

// forward stages (decimation in time)

template <class T>
__device__ __forceinline__ void fwd_r2(const T *__restrict__ a, T *__restrict__ b,
                                       int n, int V, int rem, int lv,
                                       const T *__restrict__ cs,
                                       const T *__restrict__ sn) {
  const int nh = (lv - 1) / 2;
#define A(j, k) av[irem + rem * ((j) + 2 * (k))]
#define B(k, j) bv[irem + rem * ((k) + lv * (j))]
  for (int it = threadIdx.x; it < V * rem; it += kFftThreads) {
    const int v = it / rem, irem = it - v * rem;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const T a00 = A(0, 0), a10 = A(1, 0);
    B(0, 0) = a00 + a10;
    B(0, 1) = a00 - a10;
  }
  const int per = rem * nh;
  for (int it = threadIdx.x; it < V * per; it += kFftThreads) {
    const int v = it / per, r = it - v * per;
    const int kq = r / rem, irem = r - kq * rem;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const int k = 1 + kq, kc = lv - k;
    const T c1k = cs[k], s1k = sn[k];
    const T a1k = A(1, k), a1kc = A(1, kc), a0k = A(0, k), a0kc = A(0, kc);
    const T x1 = c1k * a1k - s1k * a1kc;
    const T y1 = c1k * a1kc + s1k * a1k;
    B(k, 0) = a0k + x1;
    B(kc, 0) = a0k - x1;
    B(k, 1) = y1 - a0kc;
    B(kc, 1) = a0kc + y1;
  }
#undef A
#undef B
}

template <class T>
__device__ __forceinline__ void fwd_r4(const T *__restrict__ a, T *__restrict__ b,
                                       int n, int V, int rem, int lv,
                                       const T *__restrict__ cs,
                                       const T *__restrict__ sn) {
  const T rtf12 = (T)0.7071067811865475244008443621048490392848359376884740365883398689953662;
  const int nh = (lv - 1) / 2;
  const int even = (lv & 1) == 0;
#define A(j, k) av[irem + rem * ((j) + 4 * (k))]
#define B(k, j) bv[irem + rem * ((k) + lv * (j))]
  for (int it = threadIdx.x; it < V * rem; it += kFftThreads) {
    const int v = it / rem, irem = it - v * rem;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const T a0 = A(0, 0), a1 = A(1, 0), a2 = A(2, 0), a3 = A(3, 0);
    const T t1r = a0 + a2;
    const T t2r = a1 + a3;
    B(0, 0) = t1r + t2r;
    B(0, 1) = a0 - a2;
    B(0, 2) = t1r - t2r;
    B(0, 3) = a3 - a1;
  }
  const int per = rem * nh;
  for (int it = threadIdx.x; it < V * per; it += kFftThreads) {
    const int v = it / per, r = it - v * per;
    const int kq = r / rem, irem = r - kq * rem;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const int k = 1 + kq, kc = lv - k;
    const T c1k = cs[k], s1k = sn[k];
    const T c2k = cs[k + lv], s2k = sn[k + lv];
    const T c3k = cs[k + 2 * lv], s3k = sn[k + 2 * lv];
    const T a0k = A(0, k), a0kc = A(0, kc), a1k = A(1, k), a1kc = A(1, kc);
    const T a2k = A(2, k), a2kc = A(2, kc), a3k = A(3, k), a3kc = A(3, kc);
    const T x1p = c1k * a1k - s1k * a1kc;
    const T y1p = c1k * a1kc + s1k * a1k;
    const T x2p = c2k * a2k - s2k * a2kc;
    const T y2p = c2k * a2kc + s2k * a2k;
    const T x3p = c3k * a3k - s3k * a3kc;
    const T y3p = c3k * a3kc + s3k * a3k;
    const T t1r = a0k + x2p;
    const T t1i = a0kc + y2p;
    const T t2r = x1p + x3p;
    const T t2i = y1p + y3p;
    const T t3r = a0k - x2p;
    const T t3i = a0kc - y2p;
    const T t4r = x3p - x1p;
    const T t4i = y1p - y3p;
    B(k, 0) = t1r + t2r;
    B(kc, 0) = t3r - t4i;
    B(k, 1) = t3r + t4i;
    B(kc, 1) = t1r - t2r;
    B(k, 2) = t2i - t1i;
    B(kc, 2) = t3i + t4r;
    B(k, 3) = t4r - t3i;
    B(kc, 3) = t1i + t2i;
  }
  if (even) {
    const int l2 = lv / 2;
    for (int it = threadIdx.x; it < V * rem; it += kFftThreads) {
      const int v = it / rem, irem = it - v * rem;
      const T *av = a + v * n;
      T *bv = b + v * n;
      const T a0 = A(0, l2), a1 = A(1, l2), a2 = A(2, l2), a3 = A(3, l2);
      const T q1 = rtf12 * (a1 - a3);
      const T q2 = rtf12 * (a1 + a3);
      B(l2, 0) = a0 + q1;
      B(l2, 1) = a0 - q1;
      B(l2, 2) = a2 - q2;
      B(l2, 3) = -a2 - q2;
    }
  }
#undef A
#undef B
}

// ---------------------------------------------------------------------------
// reverse stages (decimation in frequency); nv = cum, lv = rem

template <class T>
__device__ __forceinline__ void rev_r4(const T *__restrict__ a, T *__restrict__ b,
                                       int n, int V, int cum, int lv,
                                       const T *__restrict__ cs,
                                       const T *__restrict__ sn) {
  const T rtf12 = (T)0.7071067811865475244008443621048490392848359376884740365883398689953662;
  const int nh = (lv - 1) / 2;
  const int even = (lv & 1) == 0;
#define A(k, j) av[ic + cum * ((k) + lv * (j))]
#define B(j, k) bv[ic + cum * ((j) + 4 * (k))]
  for (int it = threadIdx.x; it < V * cum; it += kFftThreads) {
    const int v = it / cum, ic = it - v * cum;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const T a0 = A(0, 0), a1 = A(0, 1), a2 = A(0, 2), a3 = A(0, 3);
    const T t1r = a0 + a2;
    const T t2r = a1;
    const T t3r = a0 - a2;
    const T t4r = a3;
    B(0, 0) = t1r + t2r;
    B(1, 0) = t3r + t4r;
    B(2, 0) = t1r - t2r;
    B(3, 0) = t3r - t4r;
  }
  const int per = cum * nh;
  for (int it = threadIdx.x; it < V * per; it += kFftThreads) {
    const int v = it / per, r = it - v * per;
    const int kq = r / cum, ic = r - kq * cum;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const int k = 1 + kq, kc = lv - k;
    const T c1k = cs[k], s1k = sn[k];
    const T c2k = cs[k + lv], s2k = sn[k + lv];
    const T c3k = cs[k + 2 * lv], s3k = sn[k + 2 * lv];
    const T ak0 = A(k, 0), ak1 = A(k, 1), ak2 = A(k, 2), ak3 = A(k, 3);
    const T akc0 = A(kc, 0), akc1 = A(kc, 1), akc2 = A(kc, 2), akc3 = A(kc, 3);
    const T t1r = ak0 + akc1;
    const T t1i = akc3 - ak2;
    const T t2r = ak1 + akc0;
    const T t2i = akc2 - ak3;
    const T t3r = ak0 - akc1;
    const T t3i = akc3 + ak2;
    const T t4r = ak1 - akc0;
    const T t4i = akc2 + ak3;
    const T x1p = t3r + t4i;
    const T y1p = t3i - t4r;
    const T x2p = t1r - t2r;
    const T y2p = t1i - t2i;
    const T x3p = t3r - t4i;
    const T y3p = t3i + t4r;
    B(0, k) = t1r + t2r;
    B(0, kc) = t1i + t2i;
    B(1, k) = c1k * x1p - s1k * y1p;
    B(1, kc) = c1k * y1p + s1k * x1p;
    B(2, k) = c2k * x2p - s2k * y2p;
    B(2, kc) = c2k * y2p + s2k * x2p;
    B(3, k) = c3k * x3p - s3k * y3p;
    B(3, kc) = c3k * y3p + s3k * x3p;
  }
  if (even) {
    const int l2 = lv / 2;
    for (int it = threadIdx.x; it < V * cum; it += kFftThreads) {
      const int v = it / cum, ic = it - v * cum;
      const T *av = a + v * n;
      T *bv = b + v * n;
      const T a0 = A(l2, 0), a1 = A(l2, 1), a2 = A(l2, 2), a3 = A(l2, 3);
      B(0, l2) = a0 + a1;
      B(2, l2) = a3 - a2;
      const T t3r = a0 - a1;
      const T t4r = a3 + a2;
      B(1, l2) = rtf12 * (t3r + t4r);
      B(3, l2) = rtf12 * (t4r - t3r);
    }
  }
#undef A
#undef B
}

template <class T>
__device__ __forceinline__ void rev_r2(const T *__restrict__ a, T *__restrict__ b,
                                       int n, int V, int cum, int lv,
                                       const T *__restrict__ cs,
                                       const T *__restrict__ sn) {
  const int nh = (lv - 1) / 2;
#define A(k, j) av[ic + cum * ((k) + lv * (j))]
#define B(j, k) bv[ic + cum * ((j) + 2 * (k))]
  for (int it = threadIdx.x; it < V * cum; it += kFftThreads) {
    const int v = it / cum, ic = it - v * cum;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const T a0 = A(0, 0), a1 = A(0, 1);
    B(0, 0) = a0 + a1;
    B(1, 0) = a0 - a1;
  }
  const int per = cum * nh;
  for (int it = threadIdx.x; it < V * per; it += kFftThreads) {
    const int v = it / per, r = it - v * per;
    const int kq = r / cum, ic = r - kq * cum;
    const T *av = a + v * n;
    T *bv = b + v * n;
    const int k = 1 + kq, kc = lv - k;
    const T c1k = cs[k], s1k = sn[k];
    const T ak0 = A(k, 0), akc0 = A(kc, 0), ak1 = A(k, 1), akc1 = A(kc, 1);
    const T x1p = ak0 - akc0;
    const T y1p = akc1 + ak1;
    B(0, k) = ak0 + akc0;
    B(0, kc) = akc1 - ak1;
    B(1, k) = c1k * x1p - s1k * y1p;
    B(1, kc) = c1k * y1p + s1k * x1p;
  }
#undef A
#undef B
}

// ---------------------------------------------------------------------------

// N (the transform length) is a compile-time constant so that the stage
// loops unroll and every index division becomes a shift or a multiply; the
// batch VT (vectors per block, 1..kMaxV) is one as well.  Blocks are mapped
// so that the blocks one XCD runs hold neighbouring vectors (shared L2 lines
// on the strided side).
template <int N, bool FWD, class TI, class T, class TO, int VT>
__global__ void __launch_bounds__(kFftThreads) fft_kernel(FftArgs g) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = (T *)lds_raw;
  const int f = blockIdx.y;
  const TI *gin = (const TI *)g.in[f];
  TO *gout = (TO *)g.out[f];
  const double *gdk = g.dk[f];
  const int gpre = g.pre[f];
  const T *gtrig = (const T *)g.trig;
  constexpr int L2N = __builtin_ctz(N);
  constexpr int n2 = L2N & 1, n4 = L2N >> 1;
  const int m = g.m;
  constexpr int V = VT, lv = __builtin_ctz(VT);
  const int nb = gridDim.x;
  const int bid = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int v0 = bid * V;
  T *bufA = lds, *bufB = lds + V * N;
  const T *tcs = kTrigInLds ? lds + 2 * V * N : gtrig, *tsn = tcs + N;
  const int tot = V * N;
  constexpr int nhalf = N / 2;
  // ---- load.  Every value a thread needs is fetched before any is used
  // (branch-free, from clamped in-bounds addresses), so each thread has all
  // of its global loads in flight at once.
  constexpr int PER = (V * N + kFftThreads - 1) / kFftThreads;  // values per thread
  constexpr bool FULL = (V * N) % kFftThreads == 0;
  const int tid = threadIdx.x;
  TI raw[PER];
  double dkr[PER];
  const bool rows_in = g.in_rows != 0;
  const int mh = m / 2;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int idx = tid + kFftThreads * j;
    const bool live = FULL || idx < tot;
    int v, e;
    if (rows_in) {
      v = idx / N;
      e = idx - v * N;
    } else {
      e = idx >> lv;
      v = idx & (V - 1);
    }
    const int vg = v0 + v;
    // the source element: mirrored for a fused derivative (sta2dfft.f90:169-245)
    int vs = vg, es = e;
    if (!FWD && gpre == 2) es = (e == 0) ? 0 : N - e;
    if (!FWD && gpre == 1) vs = (vg == 0) ? 0 : m - vg;
    raw[j] = live ? (rows_in ? gin[(int64_t)vs * N + es] : gin[vs + (int64_t)m * es]) : (TI)0;
    // the derivative's wavenumber, filtered (index 0 where the result is 0)
    if (!FWD && gpre == 2) {
      const int kk = (e < nhalf) ? e : N - e;
      dkr[j] = gdk[(e == 0 || e == nhalf) ? 0 : 2 * kk - 1];
    } else if (!FWD && gpre == 1) {
      const int kk = (vg < mh) ? vg : m - vg;
      dkr[j] = gdk[(vg == 0 || vg == mh) ? 0 : 2 * kk - 1];
    }
  }
  if (kTrigInLds) {
    constexpr int TPER = (2 * N + kFftThreads - 1) / kFftThreads;
    T tr[TPER];
#pragma unroll
    for (int j = 0; j < TPER; j++) {
      const int i = tid + kFftThreads * j;
      tr[j] = (2 * N % kFftThreads == 0 || i < 2 * N) ? gtrig[i] : (T)0;
    }
#pragma unroll
    for (int j = 0; j < TPER; j++) {
      const int i = tid + kFftThreads * j;
      if (2 * N % kFftThreads == 0 || i < 2 * N) lds[2 * V * N + i] = tr[j];
    }
  }
  const double subv = (FWD && g.sub) ? *g.sub / g.subdiv : 0.0;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int idx = tid + kFftThreads * j;
    if (!(FULL || idx < tot)) continue;
    int v, e;
    if (rows_in) {
      v = idx / N;
      e = idx - v * N;
    } else {
      e = idx >> lv;
      v = idx & (V - 1);
    }
    const int vg = v0 + v;
    T x;
    if (FWD) {
      // con2grid's mean removal (contours.f90:338-345) fused into the load
      x = g.sub ? (T)(raw[j] - subv) : (T)raw[j];
    } else {
      if (gpre == 2) {
        // along elements: der(e) from var(n-e)
        const T dkv = (T)dkr[j];
        const T var = (T)raw[j];
        x = (e == 0 || e == nhalf) ? (T)0.0 : ((e < nhalf) ? -dkv * var : dkv * var);
      } else if (gpre == 1) {
        // across vectors (vector index is the derivative's wavenumber, m = its length)
        const T dkv = (T)dkr[j];
        const T var = (T)raw[j];
        x = (vg == 0 || vg == mh) ? (T)0.0 : ((vg < mh) ? -dkv * var : dkv * var);
      } else {
        x = (T)raw[j];
      }
      // revfft pre-processing, stafft.f90:311-325
      if (e >= nhalf + 1) x = -x;
      if (e == 0 || e == nhalf) x = (T)0.5 * x;
    }
    bufA[v * N + e] = x;
  }
  __syncthreads();
  // ---- stages
  T *a = bufA, *b = bufB;
  if (FWD) {
    int rem = N, cum = 1;
#pragma unroll
    for (int s = 0; s < n2; s++) {
      rem /= 2;
      const int iloc = (rem - 1) * 2 * cum;
      fwd_r2(a, b, N, V, rem, cum, tcs + iloc, tsn + iloc);
      __syncthreads();
      T *t = a; a = b; b = t;
      cum *= 2;
    }
#pragma unroll
    for (int s = 0; s < n4; s++) {
      rem /= 4;
      const int iloc = (rem - 1) * 4 * cum;
      fwd_r4(a, b, N, V, rem, cum, tcs + iloc, tsn + iloc);
      __syncthreads();
      T *t = a; a = b; b = t;
      cum *= 4;
    }
  } else {
    int cum = 1, rem = N;
#pragma unroll
    for (int s = 0; s < n4; s++) {
      rem /= 4;
      const int iloc = (cum - 1) * 4 * rem;
      rev_r4(a, b, N, V, cum, rem, tcs + iloc, tsn + iloc);
      __syncthreads();
      T *t = a; a = b; b = t;
      cum *= 4;
    }
#pragma unroll
    for (int s = 0; s < n2; s++) {
      rem /= 2;
      const int iloc = (cum - 1) * 2 * rem;
      rev_r2(a, b, N, V, cum, rem, tcs + iloc, tsn + iloc);
      __syncthreads();
      T *t = a; a = b; b = t;
      cum *= 2;
    }
  }
  // ---- store (normalisation folded in, stafft.f90:275-284 / 394-403)
  const T post = (T)g.post[f];
  const bool rows_out = g.out_rows != 0;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int idx = tid + kFftThreads * j;
    if (!(FULL || idx < tot)) continue;
    int v, e;
    if (rows_out) {
      v = idx / N;
      e = idx - v * N;
      gout[(int64_t)(v0 + v) * N + e] = (TO)(a[v * N + e] * post);
    } else {
      e = idx >> lv;
      v = idx & (V - 1);
      gout[(v0 + v) + (int64_t)m * e] = (TO)(a[v * N + e] * post);
    }
  }
}

template <int N, class TI, class T, class TO, int VT>
static int launch_nv(bool fwd, dim3 grid, size_t shm, hipStream_t s, const FftArgs &g,
                     hipEvent_t ea, hipEvent_t eb) {
  // more than 64 KiB of LDS per workgroup must be requested (gfx950: 160 KiB)
  static bool raised = false;
  if (!raised && shm > 65536) {
    HYD_HIP(hipFuncSetAttribute((const void *)fft_kernel<N, true, TI, T, TO, VT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HYD_HIP(hipFuncSetAttribute((const void *)fft_kernel<N, false, TI, T, TO, VT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    raised = true;
  }
  // with timing events the launch records them at the dispatch's own start
  // and end; without, an ordinary launch (the one stream capture records)
  if (ea && eb) {
    if (fwd)
      hipExtLaunchKernelGGL((fft_kernel<N, true, TI, T, TO, VT>), grid, dim3(kFftThreads), (uint32_t)shm,
                            s, ea, eb, 0u, g);
    else
      hipExtLaunchKernelGGL((fft_kernel<N, false, TI, T, TO, VT>), grid, dim3(kFftThreads), (uint32_t)shm,
                            s, ea, eb, 0u, g);
  } else if (fwd) {
    hipLaunchKernelGGL((fft_kernel<N, true, TI, T, TO, VT>), grid, dim3(kFftThreads), shm, s, g);
  } else {
    hipLaunchKernelGGL((fft_kernel<N, false, TI, T, TO, VT>), grid, dim3(kFftThreads), shm, s, g);
  }
  return HYD_OK;
}

template <int N, class TI, class T, class TO>
static int launch_n(bool fwd, dim3 grid, size_t shm, hipStream_t s, const FftArgs &g,
                    hipEvent_t ea, hipEvent_t eb) {
  switch (g.V) {
    case 1: return launch_nv<N, TI, T, TO, 1>(fwd, grid, shm, s, g, ea, eb);
    case 2: return launch_nv<N, TI, T, TO, 2>(fwd, grid, shm, s, g, ea, eb);
    case 4: return launch_nv<N, TI, T, TO, 4>(fwd, grid, shm, s, g, ea, eb);
    case 8: return launch_nv<N, TI, T, TO, 8>(fwd, grid, shm, s, g, ea, eb);
    default: set_error("fft: unsupported batch"); return HYD_EARG;
  }
}

template <class TI, class T, class TO>
static int fft_launch_t(Ctx &c, bool fwd, int F, const TI *const *in, TO *const *out, int n,
                        int m, int in_rows, int out_rows, const int *pre,
                        const double *const *dk, const double *post, const T *trig,
                        const double *sub = nullptr, double subdiv = 1.0) {
  if (F < 1 || F > kMaxF) {
    set_error("fft: batch of 1..4 fields");
    return HYD_EARG;
  }
  FftArgs g;
  for (int f = 0; f < kMaxF; f++) {
    const int q = f < F ? f : 0;
    g.in[f] = in[q];
    g.out[f] = out[q];
    g.dk[f] = dk ? dk[q] : nullptr;
    g.pre[f] = pre ? pre[q] : 0;
    g.post[f] = post[q];
  }
  g.trig = trig;
  g.sub = sub;
  g.subdiv = subdiv;
  g.n = n;
  g.m = m;
  // batch per workgroup: the autotuned value of this layout class when it
  // divides m, else min(4096/n, kMaxV)
  const int cls = fwd ? (out_rows ? 0 : 1) : (in_rows ? 3 : 2);
  // tuned for the context's length, per arithmetic type
  int V = 0;
  if (n == c.c.nx) {
    if (sizeof(T) == 4) V = c.fft_v32[cls];
    else if (F == 1 && c.fft_v1[cls]) V = c.fft_v1[cls];
    else if (F >= 3 && c.fft_v4[cls]) V = c.fft_v4[cls];
    else V = c.fft_v[cls];
  }
  if (V < 1 || V > kMaxV || V > m || m % V) {
    // the largest power of two <= 4096/n and kMaxV that divides m
    V = std::max(1, std::min(4096 / n, kMaxV));
    while (V > 1 && (V > m || m % V)) V /= 2;
  }
  g.V = V;
  g.lgV = __builtin_ctz(V);
  g.in_rows = in_rows;
  g.out_rows = out_rows;
  // factor counts of this length (stafft.f90:146-178: 4s, then one 2)
  const int lg = __builtin_ctz((unsigned)n);
  g.n2 = lg & 1;
  g.n4 = lg >> 1;
  const size_t shm = fft_lds_bytes(n, V, sizeof(T));
  const dim3 grid(m / V, F);
  // one launch reads and writes F x m vectors of n values
  KTimer tk(c, sizeof(T) == 8 ? (fwd ? "k:fft_fwd" : "k:fft_rev") : (fwd ? "k:fft32_fwd" : "k:fft32_rev"),
            (int64_t)(sizeof(TI) + sizeof(TO)) * n * m * F, true);
  switch (n) {
    case 8: HYD_CHECK((launch_n<8, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 16: HYD_CHECK((launch_n<16, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 32: HYD_CHECK((launch_n<32, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 64: HYD_CHECK((launch_n<64, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 128: HYD_CHECK((launch_n<128, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 256: HYD_CHECK((launch_n<256, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 512: HYD_CHECK((launch_n<512, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 1024: HYD_CHECK((launch_n<1024, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 2048: HYD_CHECK((launch_n<2048, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    case 4096: HYD_CHECK((launch_n<4096, TI, T, TO>(fwd, grid, shm, c.s, g, tk.start(), tk.stop()))); break;
    default:
      set_error("fft: transform length must be a power of two in [8, 4096]");
      return HYD_EFACTOR;
  }
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static int fft_launch(Ctx &c, bool fwd, const double *in, double *out, int n, int m,
                      int in_rows, int out_rows, int pre, const double *dk, double post,
                      const double *trig) {
  return fft_launch_t<double, double, double>(c, fwd, 1, &in, &out, n, m, in_rows, out_rows,
                                              &pre, &dk, &post, trig);
}

template <class TI, class T, class TO>
static int fft_launch1(Ctx &c, bool fwd, const TI *in, TO *out, int n, int m, int in_rows,
                       int out_rows, int pre, const double *dk, double post, const T *trig,
                       const double *sub = nullptr, double subdiv = 1.0) {
  return fft_launch_t<TI, T, TO>(c, fwd, 1, &in, &out, n, m, in_rows, out_rows, &pre, &dk, &post,
                                 trig, sub, subdiv);
}

// Autotuner: for each layout class (forward strided->rows, forward
// strided->strided, reverse strided->strided, reverse rows->strided) time
// the candidate batches V on scratch fields and keep the fastest.  Every V
// gives bit-identical results (each vector's arithmetic is the same).  The
// FP64 passes are timed as the step launches most of them, two fields per
// launch, over three rounds that visit the candidates in turn (so that a
// slow spell of the device does not favour one), each candidate keeping
// its best round.
int fft_autotune(Ctx &c) {
  const int n = c.c.nx;
  hipEvent_t e0, e1;
  HYD_HIP(hipEventCreate(&e0));
  HYD_HIP(hipEventCreate(&e1));
  const bool fw[4] = {true, true, false, false};
  const int inr[4] = {0, 0, 0, 1}, outr[4] = {1, 0, 0, 0};
  int rc = HYD_OK;
  // two-field launches first (fft_v), then one-field ones (fft_v1: half
  // the grid, so a smaller batch may fill the device better) and four-field
  // ones (fft_v4; outputs in the fine-grid scratch)
  for (int pass = 0; pass < 3 * 4 && rc == HYD_OK; pass++) {
    const int cls = pass % 4, F = pass < 4 ? 2 : (pass < 8 ? 1 : 4);
    int *tab = F == 2 ? c.fft_v : (F == 1 ? c.fft_v1 : c.fft_v4);
    float best[kMaxV + 1];
    for (int V = 0; V <= kMaxV; V++) best[V] = 1e30f;
    const double *trig = fw[cls] ? c.xtrig : c.ytrig;
    const double *in[4] = {c.wa, c.wc, c.wb, c.wd};
    double *out[4] = {F == 4 ? c.fq : c.wb, F == 4 ? c.fq + c.N : c.wd, c.fq + 2 * c.N,
                      c.fq + 3 * c.N};
    const double post[4] = {1.0, 1.0, 1.0, 1.0};
    auto run = [&]() {
      return fft_launch_t<double, double, double>(c, fw[cls], F, in, out, n, n, inr[cls],
                                                  outr[cls], nullptr, nullptr, post, trig);
    };
    for (int round = 0; round < 3 && rc == HYD_OK; round++) {
      for (int V = 1; V <= kMaxV && V <= n; V *= 2) {
        if (fft_lds_bytes(n, V, sizeof(double)) > 160 * 1024) break;
        tab[cls] = V;
        rc = run();
        if (rc != HYD_OK) break;
        if (hipEventRecord(e0, c.s) != hipSuccess) { rc = HYD_EHIP; break; }
        for (int r = 0; r < 8 && rc == HYD_OK; r++) rc = run();
        if (rc != HYD_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, c.s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = HYD_EHIP;
          break;
        }
        best[V] = std::min(best[V], ms);
      }
    }
    int bestV = 1;
    for (int V = 2; V <= kMaxV; V *= 2)
      if (best[V] < best[bestV]) bestV = V;
    tab[cls] = bestV;
  }
  // the FP32 passes of the mixed step (half the LDS per vector), timed the
  // same way
  for (int cls = 0; cls < 4 && rc == HYD_OK && c.p.precision == 1; cls++) {
    float best[kMaxV + 1];
    for (int V = 0; V <= kMaxV; V++) best[V] = 1e30f;
    const float *trig = fw[cls] ? c.xtrig32 : c.ytrig32;
    const float *in[2] = {c.fa32, c.fc32};
    float *out[2] = {c.fb32, c.fd32};
    const double post[2] = {1.0, 1.0};
    auto run = [&]() {
      return fft_launch_t<float, float, float>(c, fw[cls], 2, in, out, n, n, inr[cls], outr[cls],
                                               nullptr, nullptr, post, trig);
    };
    for (int round = 0; round < 3 && rc == HYD_OK; round++) {
      for (int V = 1; V <= kMaxV && V <= n; V *= 2) {
        if (fft_lds_bytes(n, V, sizeof(float)) > 160 * 1024) break;
        c.fft_v32[cls] = V;
        rc = run();
        if (rc != HYD_OK) break;
        if (hipEventRecord(e0, c.s) != hipSuccess) { rc = HYD_EHIP; break; }
        for (int r = 0; r < 8 && rc == HYD_OK; r++) rc = run();
        if (rc != HYD_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, c.s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = HYD_EHIP;
          break;
        }
        best[V] = std::min(best[V], ms);
      }
    }
    int bestV = 1;
    for (int V = 2; V <= kMaxV; V *= 2)
      if (best[V] < best[bestV]) bestV = V;
    c.fft_v32[cls] = bestV;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != HYD_OK) set_error("fft autotuning failed");
  return rc;
}

// ptospc, sta2dfft.f90:103-132: x-transform (ny strided vectors) written as
// rows = svar(kx,iy); then the y-transform of svar in place.
// ptospc of rvar - *sum / ndiv (the subtraction done while loading).
int fft_ptospc_mean(Ctx &c, const double *rvar, const double *sum, double ndiv, double *svar) {
  const int n = c.c.nx;
  KTimer t(c, "fft_ptospc", 4 * 8 * c.N);
  const double *in[1] = {rvar};
  double *out[1] = {svar};
  const double p[1] = {c.plan.fnorm};
  HYD_CHECK((fft_launch_t<double, double, double>(c, true, 1, in, out, n, n, 0, 1, nullptr,
                                                   nullptr, p, c.xtrig, sum, ndiv)));
  HYD_CHECK(fft_launch(c, true, svar, svar, n, n, 0, 0, 0, nullptr, c.plan.fnorm, c.ytrig));
  return HYD_OK;
}

int fft_ptospc(Ctx &c, const double *rvar, double *svar) {
  const int n = c.c.nx;
  KTimer t(c, "fft_ptospc", 4 * 8 * c.N);
  HYD_CHECK(fft_launch(c, true, rvar, svar, n, n, 0, 1, 0, nullptr, c.plan.fnorm, c.xtrig));
  HYD_CHECK(fft_launch(c, true, svar, svar, n, n, 0, 0, 0, nullptr, c.plan.fnorm, c.ytrig));
  return HYD_OK;
}

// spctop, sta2dfft.f90:136-165, with an optional fused xderiv / yderiv of
// the input (spectral.f90:328-332, 359-363) and a sign folded into the last
// store (spectral.f90:335-339).  tmp is an (nx,ny) work array.
int fft_spctop(Ctx &c, const double *svar, double *rvar, int pre, double post_sign, double *tmp) {
  const int n = c.c.nx;
  KTimer t(c, "fft_spctop", 4 * 8 * c.N);
  const double *dk = pre == PRE_XDERIV ? c.hrkx : (pre == PRE_YDERIV ? c.hrky : nullptr);
  // y-inverse: vectors kx (strided), elements ky.  xderiv acts across the
  // vectors (pre=1), yderiv along the elements (pre=2).
  HYD_CHECK(fft_launch(c, false, svar, tmp, n, n, 0, 0, pre, dk, c.plan.rnorm, c.ytrig));
  HYD_CHECK(fft_launch(c, false, tmp, rvar, n, n, 1, 0, 0, nullptr, c.plan.rnorm * post_sign,
                       c.xtrig));
  return HYD_OK;
}

// forfft(m, n, x) / revfft(m, n, x) on a strided (m, n) array in place
// (stafft.f90:187-406): scaled by 1/sqrt(n) forward, 2/sqrt(n) backward,
// revfft's halving of the real-only entries inside the pass.
int fft_1d(Ctx &c, bool fwd, int m, int n, double *x, const double *trig) {
  const double scale = fwd ? 1.0 / sqrt((double)n) : 2.0 / sqrt((double)n);
  return fft_launch(c, fwd, x, x, n, m, 0, 0, 0, nullptr, scale, trig);
}

// forfft(m, ng, x) / revfft(m, ng, x) on a strided (m, ng) array in place.
// Mixed-precision transforms (FP32 arithmetic, FP64 spectral state):
// ptospc from an FP32 physical field to FP64 spectral coefficients, spctop
// from FP64 coefficients (optionally differentiated) to an FP32 field.
// ptospc32 of rvar - *mean (the subtraction done in double while loading,
// as con2grid_mixed's mean removal).
int fft_ptospc32_mean(Ctx &c, const float *rvar, const double *mean, double *svar) {
  const int n = c.c.nx;
  HYD_CHECK((fft_launch1<float, float, float>(c, true, rvar, c.tmp32, n, n, 0, 1, 0, nullptr,
                                                c.plan.fnorm, c.xtrig32, mean, 1.0)));
  HYD_CHECK((fft_launch1<float, float, double>(c, true, c.tmp32, svar, n, n, 0, 0, 0, nullptr,
                                                 c.plan.fnorm, c.ytrig32)));
  return HYD_OK;
}

int fft_ptospc32(Ctx &c, const float *rvar, double *svar) {
  const int n = c.c.nx;
  HYD_CHECK((fft_launch1<float, float, float>(c, true, rvar, c.tmp32, n, n, 0, 1, 0, nullptr,
                                                c.plan.fnorm, c.xtrig32)));
  HYD_CHECK((fft_launch1<float, float, double>(c, true, c.tmp32, svar, n, n, 0, 0, 0, nullptr,
                                                 c.plan.fnorm, c.ytrig32)));
  return HYD_OK;
}

int fft_spctop32(Ctx &c, const double *svar, float *rvar, int pre, double post_sign) {
  const int n = c.c.nx;
  const double *dk = pre == PRE_XDERIV ? c.hrkx : (pre == PRE_YDERIV ? c.hrky : nullptr);
  HYD_CHECK((fft_launch1<double, float, float>(c, false, svar, c.tmp32, n, n, 0, 0, pre, dk,
                                                 c.plan.rnorm, c.ytrig32)));
  HYD_CHECK((fft_launch1<float, float, float>(c, false, c.tmp32, rvar, n, n, 1, 0, 0, nullptr,
                                                c.plan.rnorm * post_sign, c.xtrig32)));
  return HYD_OK;
}

// Batches of F <= 4 fields: two launches whatever F is.
int fft_spctop_n(Ctx &c, int F, const double *const *svar, double *const *rvar, const int *pre,
                 const double *post_sign, double *const *tmp) {
  const int n = c.c.nx;
  KTimer t(c, "fft_spctop", 4 * 8 * c.N * F);
  const double *dk[kMaxF];
  double p1[kMaxF], p2[kMaxF];
  for (int f = 0; f < F; f++) {
    dk[f] = pre[f] == PRE_XDERIV ? c.hrkx : (pre[f] == PRE_YDERIV ? c.hrky : nullptr);
    p1[f] = c.plan.rnorm;
    p2[f] = c.plan.rnorm * post_sign[f];
  }
  const int zero[kMaxF] = {0, 0, 0, 0};
  HYD_CHECK((fft_launch_t<double, double, double>(c, false, F, svar, tmp, n, n, 0, 0, pre, dk, p1,
                                                   c.ytrig)));
  HYD_CHECK((fft_launch_t<double, double, double>(c, false, F, (const double *const *)tmp, rvar, n,
                                                   n, 1, 0, zero, nullptr, p2, c.xtrig)));
  return HYD_OK;
}

int fft_ptospc_n(Ctx &c, int F, const double *const *rvar, double *const *svar) {
  const int n = c.c.nx;
  KTimer t(c, "fft_ptospc", 4 * 8 * c.N * F);
  const double p[kMaxF] = {c.plan.fnorm, c.plan.fnorm, c.plan.fnorm, c.plan.fnorm};
  HYD_CHECK((fft_launch_t<double, double, double>(c, true, F, rvar, svar, n, n, 0, 1, nullptr,
                                                   nullptr, p, c.xtrig)));
  HYD_CHECK((fft_launch_t<double, double, double>(c, true, F, (const double *const *)svar, svar, n,
                                                   n, 0, 0, nullptr, nullptr, p, c.ytrig)));
  return HYD_OK;
}

int fft_spctop32_n(Ctx &c, int F, const double *const *svar, float *const *rvar, const int *pre,
                   const double *post_sign, float *const *tmp) {
  const int n = c.c.nx;
  const double *dk[kMaxF];
  double p1[kMaxF], p2[kMaxF];
  for (int f = 0; f < F; f++) {
    dk[f] = pre[f] == PRE_XDERIV ? c.hrkx : (pre[f] == PRE_YDERIV ? c.hrky : nullptr);
    p1[f] = c.plan.rnorm;
    p2[f] = c.plan.rnorm * post_sign[f];
  }
  const int zero[kMaxF] = {0, 0, 0, 0};
  HYD_CHECK((fft_launch_t<double, float, float>(c, false, F, svar, tmp, n, n, 0, 0, pre, dk, p1,
                                                 c.ytrig32)));
  HYD_CHECK((fft_launch_t<float, float, float>(c, false, F, (const float *const *)tmp, rvar, n, n,
                                                1, 0, zero, nullptr, p2, c.xtrig32)));
  return HYD_OK;
}

int fft_ptospc32_n(Ctx &c, int F, const float *const *rvar, double *const *svar,
                   float *const *tmp) {
  const int n = c.c.nx;
  const double p[kMaxF] = {c.plan.fnorm, c.plan.fnorm, c.plan.fnorm, c.plan.fnorm};
  HYD_CHECK((fft_launch_t<float, float, float>(c, true, F, rvar, tmp, n, n, 0, 1, nullptr, nullptr,
                                                p, c.xtrig32)));
  HYD_CHECK((fft_launch_t<float, float, double>(c, true, F, (const float *const *)tmp, svar, n, n,
                                                 0, 0, nullptr, nullptr, p, c.ytrig32)));
  return HYD_OK;
}

int fft_forward_1d(Ctx &c, double *x, int m) {
  return fft_launch(c, true, x, x, c.c.nx, m, 0, 0, 0, nullptr, c.plan.fnorm, c.xtrig);
}
int fft_reverse_1d(Ctx &c, double *x, int m) {
  return fft_launch(c, false, x, x, c.c.nx, m, 0, 0, 0, nullptr, c.plan.rnorm, c.xtrig);
}

}  // namespace hyd
