// seqsum.hip -- the exact value of a left-to-right floating-point sum.
//
// gfortran evaluates SUM(a) and the explicit accumulation loops of the
// reference (contours.f90:330-336, evolution.f90:553) as
// s = ((a(1) + a(2)) + a(3)) + ..., each addition rounded.  Those sums feed
// the state (the PV mean removed by con2grid, the rms vorticity inside the
// integrating factors), and a one-ulp change there alters surgery decisions
// within a hundred steps, so the GPU must produce the very same double.
//
// Method (speculate, then verify every accepted step):
//   1. The array is cut into chunks of kChunk values.  Each chunk is summed
//      sequentially from a guessed start g_c (double-double prefix of the
//      chunk totals).  If every partial sum of that run stays in the binade
//      of g_c, each addition rounds on the same grid u, so
//      f_c(s) = s + D_c for every start s in that binade whose run also
//      stays inside it (D_c = f_c(g_c) - g_c, exact).  Exact ties make the
//      rounding depend on the parity of s/u; they are recorded.
//   2. One wavefront chains 64 chunks at a time: lane l forms the candidate
//      start S_l = S + (D_0 + ... + D_{l-1}) with a lane scan, and accepts the
//      chunk only if S_l is in the same binade as the group's first chunk,
//      the run from S_l stays inside it with a 4-ulp margin, and (when the
//      chunk had a tie) (S_l - g_l)/u is even.  While every earlier lane is
//      accepted, all scan partials are differences of values of one binade,
//      hence exact, so the first rejected lane's candidate is the true
//      sequential value; that chunk is summed sequentially from it.
//   The result is the sequential sum bit for bit on any input.
#include "hyd_internal.h"

namespace hyd {

static constexpr int kChunk = 256;

struct ChunkInfo {
  double total;   // sequential sum of the chunk from 0 (for the guess)
  double g;       // guessed start
  double d;       // f_c(g) - g
  double lo, hi;  // extremes of the exact pre-rounding sums, relative to g
  int32_t ok;     // run from g stayed in one binade
  int32_t tie;    // run from g met an exact tie
};

__device__ __forceinline__ int binade(double x) {
  return (int)((__double_as_longlong(x) >> 52) & 0x7ff);
}

__device__ __forceinline__ double pow2_field(int eb) {  // 2^(eb-1023)
  return __longlong_as_double((int64_t)eb << 52);
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__global__ void seq_chunk_totals(const double *__restrict__ a, int64_t n, ChunkInfo *ci,
                                 int64_t nch) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const int64_t b = c * kChunk, e = min(n, b + kChunk);
  double s = 0.0;
  for (int64_t i = b; i < e; i++) s = s + a[i];
  ci[c].total = s;
}

// Guesses: double-double exclusive prefix of the chunk totals (one wave,
// lanes scan 64 totals at a time).
__global__ void seq_guess(ChunkInfo *ci, int64_t nch) {
  const int l = threadIdx.x;
  double hi = 0.0, lo = 0.0;  // running prefix (uniform)
  for (int64_t c0 = 0; c0 < nch; c0 += 64) {
    const int64_t c = c0 + l;
    double th = (c < nch) ? ci[c].total : 0.0, tl = 0.0;
    // inclusive double-double scan across lanes
    for (int off = 1; off < 64; off <<= 1) {
      const double oh = __shfl_up(th, off), ol = __shfl_up(tl, off);
      if (l >= off) {
        double s, e;
        two_sum(oh, th, s, e);
        e = e + (ol + tl);
        two_sum(s, e, th, tl);
      }
    }
    // exclusive = inclusive of lane l-1
    double eh = __shfl_up(th, 1), el = __shfl_up(tl, 1);
    if (l == 0) { eh = 0.0; el = 0.0; }
    double s, e;
    two_sum(hi, eh, s, e);
    e = e + (lo + el);
    double gh, gl;
    two_sum(s, e, gh, gl);
    if (c < nch) ci[c].g = gh;
    // advance the running prefix by the group's total (lane 63 inclusive)
    const double gth = __shfl(th, 63), gtl = __shfl(tl, 63);
    two_sum(hi, gth, s, e);
    e = e + (lo + gtl);
    two_sum(s, e, hi, lo);
  }
}

__global__ void seq_simulate(const double *__restrict__ a, int64_t n, ChunkInfo *ci, int64_t nch) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const int64_t b = c * kChunk, e = min(n, b + kChunk);
  const double g = ci[c].g;
  const int eb = binade(g);
  int ok = (eb >= 2 && eb <= 2045) ? 1 : 0;
  int tie = 0;
  const double half = pow2_field(eb) * 0x1p-53;  // half the grid spacing of g's binade
  const bool neg = g < 0.0;
  double s = g, lo = 0.0, hi = 0.0;
  for (int64_t i = b; i < e; i++) {
    const double x = a[i];
    double t, err;
    two_sum(s, x, t, err);
    const double rel = (s - g) + x;  // exact to < 1 ulp of the binade
    lo = fmin(lo, rel);
    hi = fmax(hi, rel);
    if (fabs(err) == half) tie = 1;
    s = t;
    if (binade(s) != eb || (s < 0.0) != neg) ok = 0;
  }
  ci[c].d = s - g;
  ci[c].lo = lo;
  ci[c].hi = hi;
  ci[c].ok = ok;
  ci[c].tie = tie;
}

// Is a run from start x (same binade eb and sign as g) strictly inside the
// binade, with a 4-ulp margin, given the extremes [lo,hi] relative to g?
__device__ __forceinline__ bool run_inside(double x, int eb, double lo, double hi) {
  const double bl = pow2_field(eb), bh = pow2_field(eb + 1), m = 4.0 * bl * 0x1p-52;
  const double ax = fabs(x);
  const double rlo = (x < 0.0) ? -hi : lo;
  const double rhi = (x < 0.0) ? -lo : hi;
  return ax + rlo > bl + m && ax + rhi < bh - m;
}

__global__ void seq_chain(const double *__restrict__ a, int64_t n, const ChunkInfo *ci,
                          int64_t nch, double *out) {
  const int l = threadIdx.x;
  double S = 0.0;
  int64_t c0 = 0;
  // the first chunk always runs sequentially from the +0 accumulator
  {
    const int64_t e = min(n, (int64_t)kChunk);
    double s = 0.0;
    if (l == 0)
      for (int64_t i = 0; i < e; i++) s = s + a[i];
    S = __shfl(s, 0);
    c0 = 1;
  }
  while (c0 < nch) {
    const int64_t c = c0 + l;
    const bool valid = c < nch;
    ChunkInfo k;
    if (valid) {
      k = ci[c];
    } else {
      k.total = 0.0; k.g = 0.0; k.d = 0.0; k.lo = 0.0; k.hi = 0.0; k.ok = 0; k.tie = 0;
    }
    // inclusive scan of d across lanes
    double incl = k.d;
    for (int off = 1; off < 64; off <<= 1) {
      const double o = __shfl_up(incl, off);
      if (l >= off) incl = o + incl;
    }
    double excl = __shfl_up(incl, 1);
    if (l == 0) excl = 0.0;
    const double Sl = S + excl;
    const int eb0 = __shfl(binade(k.g), 0);
    const bool neg0 = __shfl(k.g < 0.0 ? 1 : 0, 0) != 0;
    const int eb = binade(k.g);
    bool acc = valid && k.ok && eb == eb0 && ((k.g < 0.0) == neg0) && binade(Sl) == eb &&
               ((Sl < 0.0) == (k.g < 0.0)) && run_inside(Sl, eb, k.lo, k.hi) &&
               run_inside(k.g, eb, k.lo, k.hi);
    if (acc && k.tie) {
      const double u = pow2_field(eb) * 0x1p-52;
      const double q = (Sl - k.g) / u;  // exact integer (same binade, multiples of u)
      acc = fmod(q, 2.0) == 0.0;
    }
    const uint64_t rej = __ballot(!acc);
    if (rej == 0) {
      S = __shfl(Sl + k.d, 63);  // exact: true end of the group's last chunk
      c0 += 64;
      continue;
    }
    const int f = __ffsll((unsigned long long)rej) - 1;
    const double Sf = __shfl(Sl, f);
    const int64_t cf = c0 + f;
    if (cf >= nch) {
      S = Sf;
      break;
    }
    const int64_t b = cf * kChunk, e = min(n, b + kChunk);
    double s = Sf;
    if (l == 0)
      for (int64_t i = b; i < e; i++) s = s + a[i];
    S = __shfl(s, 0);
    c0 = cf + 1;
  }
  if (l == 0) *out = S;
}

int seqsum(Ctx &c, const double *a, int64_t n, double *dout) {
  const int64_t nch = (n + kChunk - 1) / kChunk;
  if (nch > c.seq_cap) {
    set_error("seqsum: array longer than the context's work space");
    return HYD_ECAPACITY;
  }
  ChunkInfo *buf = (ChunkInfo *)c.seq_work;
  KTimer t(c, "seqsum", 8 * n);
  const int nb = (int)((nch + 255) / 256);
  hipLaunchKernelGGL(seq_chunk_totals, dim3(nb), dim3(256), 0, c.s, a, n, buf, nch);
  hipLaunchKernelGGL(seq_guess, dim3(1), dim3(64), 0, c.s, buf, nch);
  hipLaunchKernelGGL(seq_simulate, dim3(nb), dim3(256), 0, c.s, a, n, buf, nch);
  hipLaunchKernelGGL(seq_chain, dim3(1), dim3(64), 0, c.s, a, n, buf, nch, dout);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int64_t seqsum_work_chunks(int64_t n) { return (n + kChunk - 1) / kChunk + 64; }

int64_t seqsum_work_bytes(int64_t n) { return (int64_t)sizeof(ChunkInfo) * seqsum_work_chunks(n); }

}  // namespace hyd
