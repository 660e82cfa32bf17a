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
//   * Translation property.  If every partial sum of a run stays inside one
//     binade [2^E, 2^(E+1)) (one sign), each addition rounds on the same grid
//     u = 2^(E-52), so s_(j+1) = s_j + RN_u(x_j) and the run adds the same
//     D = sum RN_u(x_j) to any start of that binade whose run also stays
//     inside it.  Exact ties make the rounding depend on the parity of s/u;
//     a tie is accepted only if the parity of the two starts agrees.
//   * Level 1: chunks of kChunk values are simulated in parallel from a
//     guessed start g (double-double prefix of the chunk totals), recording
//     D, the extreme pre-rounding partial sums and ties.
//   * Level 2: one wavefront chains 64 chunks at a time: lane l takes the
//     candidate start S + (D_0 + ... + D_(l-1)) from a lane scan and accepts
//     its chunk if the candidate lies in the group's binade and the run from
//     it stays inside with a 4-ulp margin.  While all earlier lanes are
//     accepted every scan partial is a difference of two values of one
//     binade, hence exact, so the first rejected lane's candidate is the true
//     sequential value.  A rejected chunk is summed exactly by the same
//     speculate-and-verify scheme one level down (4 values per lane).
// Every accepted translation is exact by construction and every rejected
// piece is summed in order, so the result is the sequential sum bit for bit.
#include "hyd_internal.h"

namespace hyd {

static constexpr int kChunk = 256;   // values per chunk
static constexpr int kTile = 16;     // columns staged through LDS per pass
static constexpr int kLook = 4;      // chain look-ahead (groups of 64 chunks)

struct ChunkInfo {
  double total;   // sequential sum of the chunk from 0 (for the guess)
  double g;       // guessed start
  double d;       // f_c(g) - g
  double lo, hi;  // extremes of the pre-rounding sums, relative to g
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

__device__ __forceinline__ void dd_add(double ah, double al, double bh, double bl, double &sh,
                                       double &sl) {
  double s, e;
  two_sum(ah, bh, s, e);
  e = e + (al + bl);
  two_sum(s, e, sh, sl);
}

// Simulate a run of values from start g: returns D and fills extremes/flags.
struct Run {
  double d, lo, hi;
  int ok, tie;
};

__device__ __forceinline__ void run_step(double &s, double x, double g, int eb, bool neg,
                                         double half, Run &r) {
  double t, err;
  two_sum(s, x, t, err);
  const double rel = (s - g) + x;  // pre-rounding sum relative to g, < 1 ulp off
  r.lo = fmin(r.lo, rel);
  r.hi = fmax(r.hi, rel);
  if (fabs(err) == half) r.tie = 1;
  s = t;
  if (binade(s) != eb || (s < 0.0) != neg) r.ok = 0;
}

// Is a run from start x (binade eb, same sign as g) strictly inside the
// binade with a 4-ulp margin, given extremes [lo,hi] relative to g?
__device__ __forceinline__ bool run_inside(double x, int eb, double lo, double hi) {
  const double bl = pow2_field(eb), bh = pow2_field(eb + 1), m = 4.0 * bl * 0x1p-52;
  const double ax = fabs(x);
  const double rlo = (x < 0.0) ? -hi : lo;
  const double rhi = (x < 0.0) ? -lo : hi;
  return ax + rlo > bl + m && ax + rhi < bh - m;
}

__device__ __forceinline__ bool accept(double S, double g, int eb0, bool neg0, const Run &r) {
  const int eb = binade(g);
  bool acc = r.ok && eb >= 2 && eb <= 2045 && eb == eb0 && ((g < 0.0) == neg0) &&
             binade(S) == eb && ((S < 0.0) == (g < 0.0)) && run_inside(S, eb, r.lo, r.hi) &&
             run_inside(g, eb, r.lo, r.hi);
  if (acc && r.tie) {
    const double u = pow2_field(eb) * 0x1p-52;
    const double q = (S - g) / u;  // exact integer (same binade)
    acc = fmod(q, 2.0) == 0.0;
  }
  return acc;
}

// ---- level 1 --------------------------------------------------------------
// One workgroup of 256 threads owns 256 consecutive chunks; values are staged
// through LDS in kTile-wide column tiles so that global loads are coalesced.
template <bool SIM>
__global__ void __launch_bounds__(256) seq_chunks(const double *__restrict__ a, int64_t n,
                                                  ChunkInfo *ci, int64_t nch) {
  __shared__ double tile[256][kTile + 1];
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 256;
  const int64_t c = c0 + t;
  double s = 0.0, g = 0.0, half = 0.0;
  int eb = 0;
  bool neg = false;
  Run r;
  r.lo = 0.0;
  r.hi = 0.0;
  r.ok = 1;
  r.tie = 0;
  if (SIM && c < nch) {
    g = ci[c].g;
    s = g;
    eb = binade(g);
    neg = g < 0.0;
    half = pow2_field(eb) * 0x1p-53;
    r.ok = (eb >= 2 && eb <= 2045) ? 1 : 0;
  }
  for (int col0 = 0; col0 < kChunk; col0 += kTile) {
    // cooperative coalesced load of 256 rows x kTile values
    for (int idx = t; idx < 256 * kTile; idx += 256) {
      const int row = idx / kTile, j = idx - row * kTile;
      const int64_t gi = (c0 + row) * kChunk + col0 + j;
      tile[row][j] = gi < n ? a[gi] : 0.0;
    }
    __syncthreads();
    if (c < nch) {
      const int64_t base = c * kChunk + col0;
#pragma unroll
      for (int j = 0; j < kTile; j++) {
        if (base + j < n) {
          if (SIM) run_step(s, tile[t][j], g, eb, neg, half, r);
          else s = s + tile[t][j];
        }
      }
    }
    __syncthreads();
  }
  if (c < nch) {
    if (SIM) {
      ci[c].d = s - g;
      ci[c].lo = r.lo;
      ci[c].hi = r.hi;
      ci[c].ok = r.ok;
      ci[c].tie = r.tie;
    } else {
      ci[c].total = s;
    }
  }
}

// Guesses: double-double exclusive prefix of the chunk totals (one wave).
__global__ void seq_guess(ChunkInfo *ci, int64_t nch) {
  const int l = threadIdx.x;
  double hi = 0.0, lo = 0.0;
  for (int64_t c0 = 0; c0 < nch; c0 += 64) {
    const int64_t c = c0 + l;
    double th = (c < nch) ? ci[c].total : 0.0, tl = 0.0;
    for (int off = 1; off < 64; off <<= 1) {
      const double oh = __shfl_up(th, off), ol = __shfl_up(tl, off);
      if (l >= off) dd_add(oh, ol, th, tl, th, tl);
    }
    double eh = __shfl_up(th, 1), el = __shfl_up(tl, 1);
    if (l == 0) {
      eh = 0.0;
      el = 0.0;
    }
    double gh, gl;
    dd_add(hi, lo, eh, el, gh, gl);
    if (c < nch) ci[c].g = gh;
    dd_add(hi, lo, __shfl(th, 63), __shfl(tl, 63), hi, lo);
  }
}

// Exact sequential sum of a[b, e) (e - b <= 256) starting from S, by the
// whole wavefront: lanes hold 4 consecutive values each, speculate from
// double-double guesses and verify as in the chunk chain.
__device__ double wave_seq(const double *__restrict__ a, int64_t b, int64_t e, double S,
                           bool plus_zero) {
  const int l = threadIdx.x & 63;
  double x[4];
  int cntl = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int64_t i = b + 4 * l + j;
    x[j] = i < e ? a[i] : 0.0;
    cntl += i < e ? 1 : 0;
  }
  // the reference's accumulator starts at +0: the first addition is exact
  // except for the sign of a zero, which the plain loop below keeps
  int first = 0;  // first lane not yet summed
  if (plus_zero) {
    // lane 0's values from +0 sequentially
    double s = S;
    if (l == 0)
      for (int j = 0; j < cntl; j++) s = s + x[j];
    S = __shfl(s, 0);
    first = 1;
  }
  while (first < 64) {
    // guesses: S + double-double prefix of the lane sums of lanes >= first
    double ls = 0.0;
    if (l >= first)
      for (int j = 0; j < cntl; j++) ls = ls + x[j];
    double th = l >= first ? ls : 0.0, tl = 0.0;
    for (int off = 1; off < 64; off <<= 1) {
      const double oh = __shfl_up(th, off), ol = __shfl_up(tl, off);
      if (l >= off) dd_add(oh, ol, th, tl, th, tl);
    }
    double eh = __shfl_up(th, 1), el = __shfl_up(tl, 1);
    if (l == 0) {
      eh = 0.0;
      el = 0.0;
    }
    double gh, gl;
    dd_add(S, 0.0, eh, el, gh, gl);
    const double g = gh;
    const int eb = binade(g);
    Run r;
    r.lo = 0.0;
    r.hi = 0.0;
    r.tie = 0;
    r.ok = (eb >= 2 && eb <= 2045) ? 1 : 0;
    const double half = pow2_field(eb) * 0x1p-53;
    double s = g;
    for (int j = 0; j < cntl; j++) run_step(s, x[j], g, eb, g < 0.0, half, r);
    const double d = s - g;
    // candidates from the exact-while-accepted scan of d over lanes >= first
    double incl = l >= first ? d : 0.0;
    for (int off = 1; off < 64; off <<= 1) {
      const double o = __shfl_up(incl, off);
      if (l >= off) incl = o + incl;
    }
    double excl = __shfl_up(incl, 1);
    if (l == 0) excl = 0.0;
    const double Sl = S + excl;
    const int eb0 = __shfl(eb, first);
    const bool neg0 = __shfl(g < 0.0 ? 1 : 0, first) != 0;
    const bool acc = l < first || cntl == 0 || accept(Sl, g, eb0, neg0, r);
    const uint64_t rej = __ballot(!acc);
    if (rej == 0) {
      // all remaining lanes translate exactly: end value is the last active
      // lane's candidate plus its d
      S = __shfl(Sl + d, 63);
      break;
    }
    const int f = __ffsll((unsigned long long)rej) - 1;
    double sf = __shfl(Sl, f);
    // lane f summed in order from its exact start
    double s2 = sf;
    if (l == f)
      for (int j = 0; j < cntl; j++) s2 = s2 + x[j];
    S = __shfl(s2, f);
    first = f + 1;
  }
  return S;
}

// Level 2: the chain over chunks (one wavefront).
__global__ void __launch_bounds__(64) seq_chain(const double *__restrict__ a, int64_t n,
                                                const ChunkInfo *__restrict__ ci, int64_t nch,
                                                double *out) {
  const int l = threadIdx.x;
  // the first chunk from the +0 accumulator
  double S = wave_seq(a, 0, min(n, (int64_t)kChunk), 0.0, true);
  int64_t c0 = 1;
  // look-ahead buffer of chunk records (lane l holds chunk c0 + 64*q + l)
  ChunkInfo buf[kLook];
  int64_t bufc0 = c0;
#pragma unroll
  for (int q = 0; q < kLook; q++) {
    const int64_t c = bufc0 + 64 * q + l;
    if (c < nch) buf[q] = ci[c];
  }
  while (c0 < nch) {
    if (c0 - bufc0 >= 64 * kLook || c0 < bufc0) {
      // refill from c0
      bufc0 = c0;
#pragma unroll
      for (int q = 0; q < kLook; q++) {
        const int64_t c = bufc0 + 64 * q + l;
        if (c < nch) buf[q] = ci[c];
      }
    }
    // records of chunks c0 .. c0+63: they live in buf at offset (c0-bufc0)
    const int64_t off = c0 - bufc0;
    ChunkInfo k;
    const int64_t c = c0 + l;
    const bool valid = c < nch;
    if ((off & 63) == 0 && off + 64 <= 64 * kLook) {
      const int q = (int)(off >> 6);
      k = buf[0];
#pragma unroll
      for (int qq = 1; qq < kLook; qq++)
        if (qq == q) k = buf[qq];
    } else if (valid) {
      k = ci[c];
    }
    if (!valid) {
      k.total = 0.0; k.g = 0.0; k.d = 0.0; k.lo = 0.0; k.hi = 0.0; k.ok = 0; k.tie = 0;
    }
    double incl = k.d;
    for (int o = 1; o < 64; o <<= 1) {
      const double v = __shfl_up(incl, o);
      if (l >= o) incl = v + incl;
    }
    double excl = __shfl_up(incl, 1);
    if (l == 0) excl = 0.0;
    const double Sl = S + excl;
    const int eb0 = __shfl(binade(k.g), 0);
    const bool neg0 = __shfl(k.g < 0.0 ? 1 : 0, 0) != 0;
    Run r;
    r.d = k.d;
    r.lo = k.lo;
    r.hi = k.hi;
    r.ok = k.ok;
    r.tie = k.tie;
    const bool acc = valid && accept(Sl, k.g, eb0, neg0, r);
    const uint64_t rej = __ballot(!acc);
    if (rej == 0) {
      S = __shfl(Sl + k.d, 63);
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
    S = wave_seq(a, b, e, Sf, false);
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
  hipLaunchKernelGGL(seq_chunks<false>, dim3(nb), dim3(256), 0, c.s, a, n, buf, nch);
  hipLaunchKernelGGL(seq_guess, dim3(1), dim3(64), 0, c.s, buf, nch);
  hipLaunchKernelGGL(seq_chunks<true>, dim3(nb), dim3(256), 0, c.s, a, n, buf, nch);
  hipLaunchKernelGGL(seq_chain, dim3(1), dim3(64), 0, c.s, a, n, buf, nch, dout);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int64_t seqsum_work_chunks(int64_t n) { return (n + kChunk - 1) / kChunk + 64; }

int64_t seqsum_work_bytes(int64_t n) { return (int64_t)sizeof(ChunkInfo) * seqsum_work_chunks(n); }

}  // namespace hyd
