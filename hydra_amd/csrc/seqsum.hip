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
//   * Translation property.  While the partial sums stay inside one binade
//     [2^E, 2^(E+1)) of one sign, every addition rounds on the grid
//     u = 2^(E-52) and s_(j+1) = s_j + RN_u(x_j): a run adds the same D to
//     any start on the grid whose run stays inside, provided the starts
//     agree mod 2u (exact ties round to even).  Simulating the run from g
//     and from g+u gives both translations D0, D1.
//   * Level 1: chunks of kChunk values are simulated in parallel from a
//     guessed start (double-double prefix of the chunk totals).  The guessed
//     run is cut into segments wherever its binade changes; a segment starts
//     with that "head" addition, which the chain performs exactly.
//   * Level 2: one wavefront walks the segments 64 at a time.  Between heads
//     the lanes take their start parities from a GF(2) scan, their
//     candidate starts from a scan of the chosen translations, and accept
//     while the candidate lies on the guessed run's grid with a 4-ulp
//     margin to the binade edges.  While all earlier lanes are accepted
//     every scan partial is a difference of two values of one binade, hence
//     exact, so the first rejected lane's candidate is the true sequential
//     value; that segment is summed in order (a wavefront-level version of
//     the same scheme).
// Every accepted translation is exact by construction and every rejected
// piece is summed in order, so the result is the sequential sum bit for bit.
#include "hyd_internal.h"

namespace hyd {

static constexpr int kChunk = 256;   // values per chunk
static constexpr int kTile = 16;     // columns staged through LDS per pass

struct ChunkInfo {
  double total;   // sequential sum of the chunk from 0 (for the guess)
  double g;       // guessed start
};

// A segment of a chunk's guessed run: an optional head addition, then
// values whose guessed partial sums all share the binade of g.
struct SegRec {
  double g;       // guessed value after the head (or the chunk's guess)
  double d0, d1;  // translations from g and from g + u
  double lo, hi;  // extremes of the pre-rounding sums of both runs, rel. to g
  double xh;      // the head value
  int32_t chunk;
  int32_t meta;   // b | e << 9 | head << 18 | replay << 19 (b, e: offsets in chunk)
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
// MODE 0: chunk totals from +0.  MODE 1: segments of the guessed run.
__device__ __forceinline__ int segkey(double x) {
  return (int)((uint64_t)__double_as_longlong(x) >> 52);  // sign and exponent
}

struct SegState {
  double g, s0, s1, lo, hi, xh, u;
  int b, head, bad, key;
};

__device__ __forceinline__ void seg_open(SegState &z, int b, int head, double xh, double g) {
  const int eb = binade(g);
  z.g = g;
  z.s0 = g;
  z.key = segkey(g);
  z.bad = (eb < 2 || eb > 2045) ? 1 : 0;
  z.u = z.bad ? 0.0 : pow2_field(eb) * 0x1p-52;
  z.s1 = g + z.u;
  if (segkey(z.s1) != z.key) z.bad = 1;
  z.lo = 0.0;
  z.hi = 0.0;
  z.xh = xh;
  z.b = b;
  z.head = head;
}

__device__ __forceinline__ void seg_close(const SegState &z, int e, int64_t c, SegRec *out,
                                          int &cnt, int cap) {
  if (!(z.head || e > z.b)) return;  // empty leading segment
  if (cnt < cap) {
    SegRec r;
    r.g = z.g;
    r.d0 = z.s0 - z.g;
    r.d1 = z.s1 - (z.g + z.u);
    r.lo = z.lo;
    r.hi = z.hi;
    r.xh = z.xh;
    r.chunk = (int32_t)c;
    r.meta = z.b | (e << 9) | (z.head << 18) | (z.bad << 19);
    out[cnt] = r;
  }
  cnt++;
}

template <int MODE>
__global__ void __launch_bounds__(256) seq_chunks(const double *__restrict__ a, int64_t n,
                                                  ChunkInfo *ci, int64_t nch, SegRec *segs,
                                                  int32_t *segcnt, int cap) {
  __shared__ double tile[256][kTile + 1];
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 256;
  const int64_t c = c0 + t;
  double s = 0.0;
  SegState z;
  int cnt = 0;
  SegRec *out = segs + c * cap;
  if (MODE == 1 && c < nch) seg_open(z, 0, 0, 0.0, ci[c].g);
  for (int col0 = 0; col0 < kChunk; col0 += kTile) {
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
          const double x = tile[t][j];
          if (MODE == 0) {
            s = s + x;
          } else {
            const double t0 = z.s0 + x;
            if (segkey(t0) != z.key) {
              seg_close(z, col0 + j, c, out, cnt, cap);
              seg_open(z, col0 + j, 1, x, t0);
            } else {
              const double r0 = (z.s0 - z.g) + x;
              const double r1 = (z.s1 - z.g) + x;
              z.lo = fmin(z.lo, fmin(r0, r1));
              z.hi = fmax(z.hi, fmax(r0, r1));
              z.s0 = t0;
              z.s1 = z.s1 + x;
              if (segkey(z.s1) != z.key) z.bad = 1;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (c < nch) {
    if (MODE == 0) {
      ci[c].total = s;
    } else {
      const int e = (int)min((int64_t)kChunk, n - c * kChunk);
      seg_close(z, e, c, out, cnt, cap);
      if (cnt > cap) {
        // too many binade changes: the whole chunk is summed in order
        SegRec r;
        r.g = 0.0;
        r.d0 = r.d1 = r.lo = r.hi = r.xh = 0.0;
        r.chunk = (int32_t)c;
        r.meta = 0 | (e << 9) | (1 << 19);
        out[0] = r;
        cnt = 1;
      }
      segcnt[c] = cnt;
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

// Compaction of the per-chunk segment lists (one workgroup): exclusive scan
// of the counts, then each thread copies its chunks' records.
__global__ void __launch_bounds__(1024) seq_compact(const SegRec *__restrict__ segs,
                                                    const int32_t *__restrict__ segcnt,
                                                    int64_t nch, int cap, SegRec *out,
                                                    int64_t *total) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nch + 1023) / 1024;
  const int64_t b = min(nch, per * t), e = min(nch, b + per);
  int64_t s = 0;
  for (int64_t c = b; c < e; c++) s += segcnt[c];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t o = part[t] - s;
  for (int64_t c = b; c < e; c++) {
    const int k = segcnt[c];
    for (int j = 0; j < k; j++) out[o + j] = segs[c * cap + j];
    o += k;
  }
  if (t == 1023) *total = part[1023];
}

// Exact sequential sum of a[b, e) (e - b <= 256) starting from S, by the
// whole wavefront: lanes hold 4 consecutive values each, speculate from
// double-double guesses and verify one level down.
__device__ double wave_seq(const double *__restrict__ a, int64_t b, int64_t e, double S) {
  const int l = threadIdx.x & 63;
  double x[4];
  int cntl = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int64_t i = b + 4 * l + j;
    x[j] = i < e ? a[i] : 0.0;
    cntl += i < e ? 1 : 0;
  }
  int first = 0;  // first lane not yet summed
  while (first < 64) {
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
      S = __shfl(Sl + d, 63);
      break;
    }
    const int f = __ffsll((unsigned long long)rej) - 1;
    double sf = __shfl(Sl, f);
    double s2 = sf;
    if (l == f)
      for (int j = 0; j < cntl; j++) s2 = s2 + x[j];
    S = __shfl(s2, f);
    first = f + 1;
  }
  return S;
}

__device__ __forceinline__ int parity_of(double num, double u) {
  return fmod(num / u, 2.0) != 0.0 ? 1 : 0;  // num is an exact multiple of u
}

// Level 2: the chain over the compacted segments (one wavefront).
__global__ void __launch_bounds__(64) seq_chain(const double *__restrict__ a, int64_t n,
                                                const SegRec *__restrict__ sg,
                                                const int64_t *total, double *out) {
  const int l = threadIdx.x;
  const int64_t M = *total;
  double S = 0.0;  // the reference's accumulator starts at +0
  SegRec nxt;
  if (l < M) nxt = sg[l];
  for (int64_t base = 0; base < M; base += 64) {
    const SegRec R = nxt;
    if (base + 64 + l < M) nxt = sg[base + 64 + l];  // prefetch the next group
    const int nv = (int)min((int64_t)64, M - base);
    const bool valid = l < nv;
    const int head = valid ? (R.meta >> 18) & 1 : 0;
    const int rep = valid ? (R.meta >> 19) & 1 : 0;
    const int key = segkey(R.g);
    const double gn = __shfl_down(R.g, 1);  // next lane's guess
    int q = 0;
    while (q < nv) {
      const int rep_q = __shfl(rep, q);
      const int64_t chunk_q = (int64_t)__shfl(R.chunk, q);
      const int meta_q = __shfl(R.meta, q);
      if (rep_q) {
        const int64_t b = chunk_q * kChunk + (meta_q & 511);
        const int64_t e = chunk_q * kChunk + ((meta_q >> 9) & 511);
        S = wave_seq(a, b, e, S);
        q++;
        continue;
      }
      if (__shfl(head, q)) S = S + __shfl(R.xh, q);
      const int key_q = __shfl(key, q);
      // the run: lanes q .. h-1 without a head, replay or binade change
      const uint64_t stop = __ballot(valid && l > q && (head || rep || key != key_q));
      const int h = stop ? __ffsll((unsigned long long)stop) - 1 : nv;
      const bool in = l >= q && l < h;
      const double u = pow2_field(binade(__shfl(R.g, q))) * 0x1p-52;
      // start parities: p_(l+1) = b_l ? a_l : p_l ^ a_l, a GF(2) affine scan
      int mk = 1, mc = 0;  // map p -> mk ? p ^ mc : mc
      if (in && l + 1 < h) {
        const int al = parity_of((R.g + R.d0) - gn, u);
        const int bl = parity_of(R.d1 - R.d0, u);
        mk = bl ? 0 : 1;
        mc = al;
      }
      // inclusive scan of maps (earlier maps applied first)
      for (int off = 1; off < 64; off <<= 1) {
        const int ok_ = __shfl_up(mk, off), oc = __shfl_up(mc, off);
        if (l >= off) {
          // compose: this after other
          mc = mk ? (oc ^ mc) : mc;
          mk = mk & ok_;
        }
      }
      int ek = __shfl_up(mk, 1), ec = __shfl_up(mc, 1);
      if (l == 0) {
        ek = 1;
        ec = 0;
      }
      // maps of lanes < q are identities, so each prefix starts at lane q
      const int key_ok_q = segkey(S) == key_q;
      const int p_q = key_ok_q ? parity_of(S - __shfl(R.g, q), u) : 0;
      const int p = ek ? (p_q ^ ec) : ec;
      const double d = in ? (p ? R.d1 : R.d0) : 0.0;
      double incl = d;
      for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(incl, off);
        if (l >= off) incl = o + incl;
      }
      // lanes < q add zeros, so lane l's exclusive partial covers [q, l)
      double excl = __shfl_up(incl, 1);
      if (l == 0) excl = 0.0;
      const double Sl = S + excl;
      bool acc = false;
      if (in) {
        const int eb = binade(R.g);
        acc = segkey(Sl) == key && eb >= 2 && eb <= 2045 && run_inside(Sl, eb, R.lo, R.hi) &&
              run_inside(R.g, eb, R.lo, R.hi) && parity_of(Sl - R.g, u) == p;
      }
      const uint64_t rej = __ballot(in && !acc);
      if (rej == 0) {
        S = __shfl(Sl + d, h - 1);
        q = h;
        continue;
      }
      const int f = __ffsll((unsigned long long)rej) - 1;
      const double Sf = __shfl(Sl, f);
      const int64_t chunk_f = (int64_t)__shfl(R.chunk, f);
      const int meta_f = __shfl(R.meta, f);
      const int64_t b = chunk_f * kChunk + (meta_f & 511) + ((meta_f >> 18) & 1);
      const int64_t e = chunk_f * kChunk + ((meta_f >> 9) & 511);
      S = wave_seq(a, b, e, Sf);
      q = f + 1;
    }
  }
  if (l == 0) *out = S;
}

__global__ void seq_zero(double *out) {
  if (threadIdx.x == 0) *out = 0.0;
}

static int seg_cap(int64_t nch) { return nch <= 65536 ? 8 : 4; }

struct SeqLayout {
  int64_t nch, ci, segs, cnt, comp, tot, bytes;
  int cap;
};

static SeqLayout seq_layout(int64_t n) {
  SeqLayout L;
  L.nch = (n + kChunk - 1) / kChunk;
  L.cap = seg_cap(L.nch);
  auto al = [](int64_t x) { return (x + 255) / 256 * 256; };
  L.ci = 0;
  L.segs = al(L.ci + (int64_t)sizeof(ChunkInfo) * L.nch);
  L.cnt = al(L.segs + (int64_t)sizeof(SegRec) * L.nch * L.cap);
  L.comp = al(L.cnt + 4 * L.nch);
  L.tot = al(L.comp + (int64_t)sizeof(SegRec) * L.nch * L.cap);
  L.bytes = al(L.tot + 8);
  return L;
}

int seqsum(Ctx &c, const double *a, int64_t n, double *dout) {
  const SeqLayout L = seq_layout(n);
  if (L.nch > c.seq_cap || seq_layout((int64_t)c.seq_cap * kChunk).bytes < L.bytes) {
    set_error("seqsum: array longer than the context's work space");
    return HYD_ECAPACITY;
  }
  char *w = (char *)c.seq_work;
  ChunkInfo *ci = (ChunkInfo *)(w + L.ci);
  SegRec *segs = (SegRec *)(w + L.segs);
  int32_t *cnt = (int32_t *)(w + L.cnt);
  SegRec *comp = (SegRec *)(w + L.comp);
  int64_t *tot = (int64_t *)(w + L.tot);
  KTimer t(c, "seqsum", 8 * n);
  if (n == 0) {
    hipLaunchKernelGGL(seq_zero, dim3(1), dim3(64), 0, c.s, dout);
    HYD_HIP(hipGetLastError());
    return HYD_OK;
  }
  const int nb = (int)((L.nch + 255) / 256);
  hipLaunchKernelGGL(seq_chunks<0>, dim3(nb), dim3(256), 0, c.s, a, n, ci, L.nch, segs, cnt,
                     L.cap);
  hipLaunchKernelGGL(seq_guess, dim3(1), dim3(64), 0, c.s, ci, L.nch);
  hipLaunchKernelGGL(seq_chunks<1>, dim3(nb), dim3(256), 0, c.s, a, n, ci, L.nch, segs, cnt,
                     L.cap);
  hipLaunchKernelGGL(seq_compact, dim3(1), dim3(1024), 0, c.s, segs, cnt, L.nch, L.cap, comp,
                     tot);
  hipLaunchKernelGGL(seq_chain, dim3(1), dim3(64), 0, c.s, a, n, comp, tot, dout);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int64_t seqsum_work_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

int64_t seqsum_work_bytes(int64_t n) { return seq_layout(n).bytes; }

}  // namespace hyd
