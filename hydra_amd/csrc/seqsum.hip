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

static constexpr int kChunk = 128;   // values per chunk
static constexpr int kTile = 32;     // columns staged through LDS per pass

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

// Is a run from start x (binade eb, same sign as g) strictly inside the
// binade with a 4-ulp margin, given extremes [lo,hi] relative to g?
__device__ __forceinline__ bool run_inside(double x, int eb, double lo, double hi) {
  const double bl = pow2_field(eb), bh = pow2_field(eb + 1), m = 4.0 * bl * 0x1p-52;
  const double ax = fabs(x);
  const double rlo = (x < 0.0) ? -hi : lo;
  const double rhi = (x < 0.0) ? -lo : hi;
  return ax + rlo > bl + m && ax + rhi < bh - m;
}

// ---- wavefront primitives -------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = dpp_i<CTRL>(__double2loint(x));
  const int hi = dpp_i<CTRL>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Inclusive wavefront scan of doubles (row shifts, then row broadcasts).
__device__ __forceinline__ double wave_incl_add(double v) {
  const int l = lane_id(), rl = l & 15;
  double t;
  t = dpp_d<0x111>(v); if (rl >= 1) v = t + v;
  t = dpp_d<0x112>(v); if (rl >= 2) v = t + v;
  t = dpp_d<0x114>(v); if (rl >= 4) v = t + v;
  t = dpp_d<0x118>(v); if (rl >= 8) v = t + v;
  t = dpp_d<0x142>(v); if ((l & 31) >= 16) v = t + v;
  t = dpp_d<0x143>(v); if (l >= 32) v = t + v;
  return v;
}
// value of lane l-1 (0 for lane 0)
__device__ __forceinline__ double wave_prev_d(double v) {
  const double t = dpp_d<0x138>(v);
  return lane_id() == 0 ? 0.0 : t;
}

// GF(2) affine maps p -> k ? p ^ c : c packed as (k << 1) | c; "then" applies
// the earlier map t first.
__device__ __forceinline__ int map_then(int t, int v) {
  const int kv = v >> 1, cv = v & 1, kt = t >> 1, ct = t & 1;
  return ((kv & kt) << 1) | (kv ? (ct ^ cv) : cv);
}
__device__ __forceinline__ int wave_incl_map(int v) {
  const int l = lane_id(), rl = l & 15;
  int t;
  t = dpp_i<0x111>(v); if (rl >= 1) v = map_then(t, v);
  t = dpp_i<0x112>(v); if (rl >= 2) v = map_then(t, v);
  t = dpp_i<0x114>(v); if (rl >= 4) v = map_then(t, v);
  t = dpp_i<0x118>(v); if (rl >= 8) v = map_then(t, v);
  t = dpp_i<0x142>(v); if ((l & 31) >= 16) v = map_then(t, v);
  t = dpp_i<0x143>(v); if (l >= 32) v = map_then(t, v);
  return v;
}

// ---- level 1 --------------------------------------------------------------
// One wavefront owns 64 consecutive chunks (a thread per chunk); values are
// staged through LDS in kTile-wide column tiles, the next tile's loads in
// flight while the current one is summed.
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
  // tiny binades are left to the in-order path (the chain uses 1/u = 2^(52-E))
  z.bad = (eb < 64 || eb > 2045) ? 1 : 0;
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
__global__ void __launch_bounds__(64) seq_chunks(const double *__restrict__ a, int64_t n,
                                                 ChunkInfo *ci, int64_t nch, SegRec *segs,
                                                 int32_t *segcnt, int cap) {
  __shared__ double tile[64][kTile + 1];
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 64;
  const int64_t c = c0 + t;
  double s = 0.0;
  SegState z;
  int cnt = 0;
  SegRec *out = segs + c * cap;
  if (MODE == 1 && c < nch) seg_open(z, 0, 0, 0.0, ci[c].g);
  constexpr int per = 64 * kTile / 64;  // values each thread stages per tile
  double nxt[per];
  auto fetch = [&](int col0) {
#pragma unroll
    for (int j = 0; j < per; j++) {
      const int idx = t + 64 * j, row = idx / kTile, col = idx - row * kTile;
      const int64_t gi = (c0 + row) * kChunk + col0 + col;
      nxt[j] = gi < n ? a[gi] : 0.0;
    }
  };
  fetch(0);
  for (int col0 = 0; col0 < kChunk; col0 += kTile) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < per; j++) {
      const int idx = t + 64 * j, row = idx / kTile, col = idx - row * kTile;
      tile[row][col] = nxt[j];
    }
    __syncthreads();
    if (col0 + kTile < kChunk) fetch(col0 + kTile);
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

// Guesses: double-double exclusive prefix of the chunk totals (one
// workgroup: per-thread runs, then a scan of the run totals).
__global__ void __launch_bounds__(1024) seq_guess(ChunkInfo *ci, int64_t nch) {
  __shared__ double ph[1024], pl[1024];
  const int t = threadIdx.x;
  const int64_t per = (nch + 1023) / 1024;
  const int64_t b = min(nch, per * t), e = min(nch, b + per);
  double hi = 0.0, lo = 0.0;
  for (int64_t c = b; c < e; c++) dd_add(hi, lo, ci[c].total, 0.0, hi, lo);
  ph[t] = hi;
  pl[t] = lo;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    double oh = 0.0, ol = 0.0;
    if (t >= off) {
      oh = ph[t - off];
      ol = pl[t - off];
    }
    __syncthreads();
    if (t >= off) {
      double sh, sl;
      dd_add(oh, ol, ph[t], pl[t], sh, sl);
      ph[t] = sh;
      pl[t] = sl;
    }
    __syncthreads();
  }
  // exclusive start of this thread's run
  double gh = 0.0, gl = 0.0;
  if (t > 0) {
    gh = ph[t - 1];
    gl = pl[t - 1];
  }
  for (int64_t c = b; c < e; c++) {
    ci[c].g = gh + gl;
    dd_add(gh, gl, ci[c].total, 0.0, gh, gl);
  }
}

// Exclusive scan of the segment counts (one workgroup).
__global__ void __launch_bounds__(1024) seq_offsets(const int32_t *__restrict__ segcnt,
                                                    int64_t nch, int64_t *off, int64_t *total) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nch + 1023) / 1024;
  const int64_t b = min(nch, per * t), e = min(nch, b + per);
  int64_t s = 0;
  for (int64_t c = b; c < e; c++) s += segcnt[c];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t acc = part[t] - s;
  for (int64_t c = b; c < e; c++) {
    off[c] = acc;
    acc += segcnt[c];
  }
  if (t == 1023) *total = part[1023];
}

// Compaction: each thread moves one chunk's records.
__global__ void seq_compact(const SegRec *__restrict__ segs, const int32_t *__restrict__ segcnt,
                            const int64_t *__restrict__ off, int64_t nch, int cap, SegRec *out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const int k = segcnt[c];
  for (int j = 0; j < k; j++) out[off[c] + j] = segs[c * cap + j];
}

// Exact sequential sum of a[b, e) (e - b <= 2 * 64) starting from S: the
// wavefront loads the values, then adds them in order (the values reach
// the adder through readlane, so the chain is one dependent add per value).
__device__ double lane_seq(const double *__restrict__ a, int64_t b, int64_t e, double S) {
  const int l = lane_id();
  const int cnt = (int)(e - b);
  const double x0 = l < cnt ? a[b + l] : 0.0;
  const double x1 = 64 + l < cnt ? a[b + 64 + l] : 0.0;
  const int c0 = cnt < 64 ? cnt : 64;
  for (int i = 0; i < c0; i++) S = S + readlane_d(x0, i);
  for (int i = 64; i < cnt; i++) S = S + readlane_d(x1, i - 64);
  return S;
}

__device__ __forceinline__ int parity_of(double num, double uinv) {
  // num is an exact multiple of u = 1/uinv (a power of two), |num/u| < 2^53
  return (int)((long long)(num * uinv) & 1);
}

// Level 2: the chain over the compacted segments.  Wave 0 runs the chain;
// waves 1-3 stream the next block of records into LDS meanwhile (double
// buffered, one barrier per block).
static constexpr int kBlockRecs = 512;  // records per LDS buffer (8 groups)

__device__ void chain_group(const double *__restrict__ a, const SegRec &R, int nv, double &S) {
  const int l = lane_id();
  const bool valid = l < nv;
  const int head = valid ? (R.meta >> 18) & 1 : 0;
  const int rep = valid ? (R.meta >> 19) & 1 : 0;
  const int key = segkey(R.g);
  const double gn = dpp_d<0x130>(R.g);  // next lane's guess (wave_shl:1)
  int q = 0;
  while (q < nv) {
    const int meta_q = __builtin_amdgcn_readlane(R.meta, q);
    const int64_t chunk_q = (int64_t)__builtin_amdgcn_readlane(R.chunk, q);
    if ((meta_q >> 19) & 1) {
      const int64_t b = chunk_q * kChunk + (meta_q & 511);
      const int64_t e = chunk_q * kChunk + ((meta_q >> 9) & 511);
      S = lane_seq(a, b, e, S);
      q++;
      continue;
    }
    if ((meta_q >> 18) & 1) S = S + readlane_d(R.xh, q);
    const int key_q = __builtin_amdgcn_readlane(key, q);
    // the run: lanes q .. h-1 without a head, replay or binade change
    const uint64_t stop = __ballot(valid && l > q && (head || rep || key != key_q));
    const int h = stop ? __ffsll((unsigned long long)stop) - 1 : nv;
    const bool in = l >= q && l < h;
    const double g_q = readlane_d(R.g, q);
    const int eq = binade(g_q);
    // 1/u = 2^(52 - E) for the binade of g_q (exact; only used while in range)
    const double ui = (eq >= 2 && eq <= 2045) ? pow2_field(2098 - eq) : 0.0;
    // start parities: p_(l+1) = b_l ? a_l : p_l ^ a_l, a GF(2) affine scan
    int mp = 2;  // identity map
    if (in && l + 1 < h) {
      const int al = parity_of((R.g + R.d0) - gn, ui);
      const int bl = parity_of(R.d1 - R.d0, ui);
      mp = ((bl ? 0 : 1) << 1) | al;
    }
    const int incl_m = wave_incl_map(mp);
    int ex_m = dpp_i<0x138>(incl_m);
    if (l == 0) ex_m = 2;
    // maps of lanes < q are identities, so each prefix starts at lane q
    const int p_q = segkey(S) == key_q ? parity_of(S - g_q, ui) : 0;
    const int p = (ex_m >> 1) ? (p_q ^ (ex_m & 1)) : (ex_m & 1);
    const double d = in ? (p ? R.d1 : R.d0) : 0.0;
    // lanes < q add zeros, so lane l's exclusive partial covers [q, l)
    const double excl = wave_prev_d(wave_incl_add(d));
    const double Sl = S + excl;
    bool acc = false;
    if (in) {
      const int eb = binade(R.g);
      acc = segkey(Sl) == key && eb >= 2 && eb <= 2045 && run_inside(Sl, eb, R.lo, R.hi) &&
            run_inside(R.g, eb, R.lo, R.hi) && parity_of(Sl - R.g, ui) == p;
    }
    const uint64_t rej = __ballot(in && !acc);
    if (rej == 0) {
      S = readlane_d(Sl + d, h - 1);
      q = h;
      continue;
    }
    const int f = __ffsll((unsigned long long)rej) - 1;
    const double Sf = readlane_d(Sl, f);
    const int64_t chunk_f = (int64_t)__builtin_amdgcn_readlane(R.chunk, f);
    const int meta_f = __builtin_amdgcn_readlane(R.meta, f);
    const int64_t b = chunk_f * kChunk + (meta_f & 511) + ((meta_f >> 18) & 1);
    const int64_t e = chunk_f * kChunk + ((meta_f >> 9) & 511);
    S = lane_seq(a, b, e, Sf);
    q = f + 1;
  }
}

__global__ void __launch_bounds__(256) seq_chain(const double *__restrict__ a, int64_t n,
                                                 const SegRec *__restrict__ sg,
                                                 const int64_t *total, double *out) {
  __shared__ SegRec buf[2][kBlockRecs];
  const int t = threadIdx.x, w = t >> 6;
  const int64_t M = *total;
  const int64_t nblk = (M + kBlockRecs - 1) / kBlockRecs;
  double S = 0.0;  // the reference's accumulator starts at +0
  // block 0 by every wave
  for (int j = t; j < kBlockRecs; j += 256)
    if (j < M) buf[0][j] = sg[j];
  __syncthreads();
  for (int64_t k = 0; k < nblk; k++) {
    if (w != 0) {
      const int64_t b1 = (k + 1) * kBlockRecs;
      for (int j = t - 64; j < kBlockRecs; j += 192)
        if (b1 + j < M) buf[(k + 1) & 1][j] = sg[b1 + j];
    } else {
      const int64_t b0 = k * kBlockRecs;
      for (int g0 = 0; g0 < kBlockRecs && b0 + g0 < M; g0 += 64) {
        const int nv = (int)min((int64_t)64, M - b0 - g0);
        SegRec R;
        if (t < nv) R = buf[k & 1][g0 + t];
        else R.meta = 0, R.g = 0.0, R.d0 = R.d1 = R.lo = R.hi = R.xh = 0.0, R.chunk = 0;
        chain_group(a, R, nv, S);
      }
    }
    __syncthreads();
  }
  if (t == 0) *out = S;
}

__global__ void seq_zero(double *out) {
  if (threadIdx.x == 0) *out = 0.0;
}

static int seg_cap(int64_t nch) { return nch <= 65536 ? 8 : 2; }

struct SeqLayout {
  int64_t nch, ci, segs, cnt, off, comp, tot, bytes;
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
  L.off = al(L.cnt + 4 * L.nch);
  L.comp = al(L.off + 8 * L.nch);
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
  int64_t *off = (int64_t *)(w + L.off);
  SegRec *comp = (SegRec *)(w + L.comp);
  int64_t *tot = (int64_t *)(w + L.tot);
  KTimer t(c, "seqsum", 8 * n);
  if (n == 0) {
    hipLaunchKernelGGL(seq_zero, dim3(1), dim3(64), 0, c.s, dout);
    HYD_HIP(hipGetLastError());
    return HYD_OK;
  }
  const int nb = (int)((L.nch + 63) / 64);
  hipLaunchKernelGGL(seq_chunks<0>, dim3(nb), dim3(64), 0, c.s, a, n, ci, L.nch, segs, cnt,
                     L.cap);
  hipLaunchKernelGGL(seq_guess, dim3(1), dim3(1024), 0, c.s, ci, L.nch);
  hipLaunchKernelGGL(seq_chunks<1>, dim3(nb), dim3(64), 0, c.s, a, n, ci, L.nch, segs, cnt,
                     L.cap);
  hipLaunchKernelGGL(seq_offsets, dim3(1), dim3(1024), 0, c.s, cnt, L.nch, off, tot);
  hipLaunchKernelGGL(seq_compact, dim3((unsigned)((L.nch + 255) / 256)), dim3(256), 0, c.s, segs,
                     cnt, off, L.nch, L.cap, comp);
  hipLaunchKernelGGL(seq_chain, dim3(1), dim3(256), 0, c.s, a, n, comp, tot, dout);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int64_t seqsum_work_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

int64_t seqsum_work_bytes(int64_t n) { return seq_layout(n).bytes; }

}  // namespace hyd
