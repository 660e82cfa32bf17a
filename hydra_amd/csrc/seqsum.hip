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
//     with that "head" addition, which the chain performs exactly.  Windows
//     of kWin chunks whose runs are single segments fold into composites
//     (translations and admissible start intervals per start parity).
//   * Level 2: one wavefront walks the windows 64 at a time.  Lanes take
//     their start parities from a GF(2) scan, their candidate starts from a
//     scan of the chosen translations, and accept while the candidate lies
//     on the guessed run's grid inside the window's admissible interval.
//     While all earlier lanes are accepted every scan partial is a
//     difference of two values of one binade, hence exact, so the first
//     rejected lane's candidate is the true sequential value.  That window,
//     and every window that did not fold, is walked record by record the
//     same way (heads added in order, rejected segments summed in order).
// Every accepted translation is exact by construction and every rejected
// piece is summed in order, so the result is the sequential sum bit for bit.
#include <mutex>

#include "hyd_internal.h"

namespace hyd {

static constexpr int kChunk = 64;    // values per chunk

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
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)x, l);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
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
// Inclusive scans inside each DPP row of 16 lanes.
__device__ __forceinline__ double row_incl_add(double v) {
  const int rl = lane_id() & 15;
  double t;
  t = dpp_d<0x111>(v); if (rl >= 1) v = t + v;
  t = dpp_d<0x112>(v); if (rl >= 2) v = t + v;
  t = dpp_d<0x114>(v); if (rl >= 4) v = t + v;
  t = dpp_d<0x118>(v); if (rl >= 8) v = t + v;
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

__device__ __forceinline__ int row_incl_map(int v) {
  const int rl = lane_id() & 15;
  int t;
  t = dpp_i<0x111>(v); if (rl >= 1) v = map_then(t, v);
  t = dpp_i<0x112>(v); if (rl >= 2) v = map_then(t, v);
  t = dpp_i<0x114>(v); if (rl >= 4) v = map_then(t, v);
  t = dpp_i<0x118>(v); if (rl >= 8) v = map_then(t, v);
  return v;
}

// ---- segments of a guessed run ---------------------------------------------
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

__device__ __forceinline__ int parity_of(double num, double uinv) {
  // num is an exact multiple of u = 1/uinv (a power of two), |num/u| < 2^53
  return (int)((long long)(num * uinv) & 1);
}

// ---- composites -----------------------------------------------------------
// A window of kWin consecutive chunks whose guessed runs each form a single
// segment (no binade change, no replay) folds into one composite.  For a
// start S_1 = g_1 + delta u on the grid of that binade the members are pure
// translations: with p = parity(delta), member i starts at S_1 + off_i(p)
// (off_i the sum of the earlier members' chosen translations, parities
// following the GF(2) maps) and the window adds D(p) = off_(K+1)(p).  The
// member chain would accept member i iff its start lies on the grid, inside
// the binade with the 4-ulp margins around its run; those conditions are
// intervals in delta, intersected per parity (level 1) with one extra ulp for
// the rounding of the margin tests.  A start inside its interval is
// therefore accepted by every member and gives the member chain's value;
// any other window is summed by the member chain.
static constexpr int kWin = 16;

struct CompRec {
  double g;                   // first member's guess
  double d0, d1;              // window translations for start parity 0 / 1
  double lo0, hi0, lo1, hi1;  // admissible delta (units of u, magnitude) per parity
  int32_t win;                // window index
  int32_t meta;               // nm | ok << 5 | (map k << 1 | map c) << 6
  uint64_t cnts;              // record count of member chunk i in bits 4i..4i+3
};

// ---- level 1 --------------------------------------------------------------
// One wavefront owns 64 consecutive chunks (a thread per chunk); the block's
// 64 x 64 values are staged through LDS in one coalesced pass and each
// thread walks its chunk's row, reading eight values ahead of its adds.
// MODE 0: chunk totals from +0 (tot) and the block's double-double total
//         (blk).
// MODE 1: guesses (double-double prefix of the chunk totals: earlier blocks,
//         then the block's own scan), segments of the guessed runs, and the
//         composites of the block's kWin-chunk windows.
__device__ __forceinline__ void dd_shfl_xor(double &h, double &l, int m) {
  const double oh = __shfl_xor(h, m), ol = __shfl_xor(l, m);
  dd_add(h, l, oh, ol, h, l);
}

template <int MODE>
__global__ void __launch_bounds__(64) seq_chunks(const double *__restrict__ a, int64_t n,
                                                 int64_t nch, double *tot, double2 *blk,
                                                 SegRec *segs, int32_t *segcnt, CompRec *cps,
                                                 int cap) {
  __shared__ double tile[64][kChunk + 1];
  __shared__ SegRec first[64];
  __shared__ int32_t fcnt[64];
  __shared__ double gnx;
  const int t = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 64;
  const int64_t c = c0 + t;
  // the block's 64 chunks, staged whole: thread t loads value t of every
  // chunk (coalesced) before the prologue, so the loads overlap it
  double v[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; j++) {
    const int64_t gi = (c0 + j) * kChunk + t;
    v[j] = gi < n ? a[gi] : 0.0;
  }
  double s = 0.0;
  SegState z;
  int cnt = 0;
  SegRec *out = segs + c * cap;
  double gch = 0.0;
  // MODE 1 walks the two runs (from g and g + u) branch-free first; a chunk
  // whose run from g changes binade is walked again, value by value, into
  // its segments
  double f0 = 0.0, f1 = 0.0, fmn = 0.0, fmx = 0.0;
  bool chg = false, fbad = false;
  if (MODE == 1) {
    // offset: earlier blocks' totals (any order: the guess needs no exactness)
    double oh = 0.0, ol = 0.0;
    for (int64_t k = t; k < (int64_t)blockIdx.x; k += 64) {
      const double2 v = blk[k];
      dd_add(oh, ol, v.x, v.y, oh, ol);
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) dd_shfl_xor(oh, ol, m);
    // inclusive scan of this block's chunk totals
    double ih = c < nch ? tot[c] : 0.0, il = 0.0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double ph = __shfl_up(ih, o), pl = __shfl_up(il, o);
      if (t >= o) dd_add(ph, pl, ih, il, ih, il);
    }
    double eh = __shfl_up(ih, 1), el = __shfl_up(il, 1);
    if (t == 0) eh = el = 0.0;
    double gh, gl;
    dd_add(oh, ol, eh, el, gh, gl);
    gch = gh + gl;
    if (c < nch) seg_open(z, 0, 0, 0.0, gch);
    f0 = z.s0;
    f1 = z.s1;
    fmn = fmx = z.g;
    if (t == 63) {
      dd_add(oh, ol, ih, il, gh, gl);
      gnx = gh + gl;  // guess of the next block's first chunk
    }
  }
#pragma unroll
  for (int j = 0; j < kChunk; j++) tile[j][t] = v[j];
  __syncthreads();
  if (c < nch) {
    const double *row = tile[t];
    auto step = [&](double x) {
      if (MODE == 0) {
        s = s + x;
      } else {
        // the rounded partial sums bound the exact ones within half an ulp,
        // as the pre-rounding offsets would
        const double t0 = f0 + x, t1 = f1 + x;
        chg = chg || segkey(t0) != z.key;
        fbad = fbad || segkey(t1) != z.key;
        const double lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
        fmn = lo < fmn ? lo : fmn;
        fmx = hi > fmx ? hi : fmx;
        f0 = t0;
        f1 = t1;
      }
    };
    if ((c + 1) * kChunk <= n) {
      // eight values read ahead of the chain that consumes them
#pragma unroll
      for (int j0 = 0; j0 < kChunk; j0 += 8) {
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = row[j0 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) step(x[j]);
      }
    } else {
      const int e = (int)(n - c * kChunk);
      for (int j = 0; j < e; j++) step(row[j]);
    }
  }
  if (MODE == 0) {
    if (c < nch) tot[c] = s;
    double h = s, l = 0.0;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) dd_shfl_xor(h, l, m);
    if (t == 0) blk[blockIdx.x] = make_double2(h, l);
    return;
  }
  if (c < nch) {
    const int e = (int)min((int64_t)kChunk, n - c * kChunk);
    if (!chg) {
      // one segment: the fast walk's sums (offsets exact within one binade)
      z.s0 = f0;
      z.s1 = f1;
      z.lo = fmin(0.0, fmn - z.g);
      z.hi = fmax(0.0, fmx - z.g);
      if (fbad) z.bad = 1;
    } else {
      // the same walk, cut into segments where the run from g changes
      // binade (values read eight ahead)
      const double *row = tile[t];
      double mn = z.g, mx = z.g;
      for (int j0 = 0; j0 < e; j0 += 8) {
        double x8[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x8[j] = j0 + j < e ? row[j0 + j] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          if (j0 + j >= e) break;
          const double x = x8[j];
          const double t0 = z.s0 + x;
          if (segkey(t0) != z.key) {
            z.lo = fmin(0.0, mn - z.g);
            z.hi = fmax(0.0, mx - z.g);
            seg_close(z, j0 + j, c, out, cnt, cap);
            seg_open(z, j0 + j, 1, x, t0);
            mn = mx = z.g;
          } else {
            const double t1 = z.s1 + x;
            const double lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            mn = lo < mn ? lo : mn;
            mx = hi > mx ? hi : mx;
            z.s0 = t0;
            z.s1 = t1;
            if (segkey(t1) != z.key) z.bad = 1;
          }
        }
      }
      z.lo = fmin(0.0, mn - z.g);
      z.hi = fmax(0.0, mx - z.g);
    }
    seg_close(z, e, c, out, cnt, cap);
    if (cnt > cap) {
      // too many binade changes: the whole chunk is summed in order (the
      // record keeps the chunk's guess for the window's class table)
      SegRec r;
      r.g = gch;
      r.d0 = r.d1 = r.lo = r.hi = r.xh = 0.0;
      r.chunk = (int32_t)c;
      r.meta = 0 | (e << 9) | (1 << 19);
      out[0] = r;
      cnt = 1;
    }
    segcnt[c] = cnt;
    first[t] = out[0];
    fcnt[t] = cnt;
  }
  __syncthreads();
  // composites: the 16 lanes of DPP row w fold window w of the block, lane
  // i taking member i: start parities by a row scan of the members' maps,
  // translations by row scans (exact sums), intervals by row max/min
  const int wl = t / kWin, i = t % kWin;
  const int64_t w0 = c0 + (int64_t)wl * kWin;
  const int nm = w0 < nch ? (int)min((int64_t)kWin, nch - w0) : 0;
  const bool mem = i < nm;
  const int fc = mem ? fcnt[t] : 1;
  const bool single = __ballot(!(fc == 1)) >> (wl * kWin) & 0xffff ? false : true;
  uint64_t cn = mem ? (uint64_t)fc << (4 * i) : 0;
#pragma unroll
  for (int m = 1; m < kWin; m <<= 1) cn |= __shfl_xor(cn, m);
  const SegRec A = first[wl * kWin];
  const SegRec Ri = first[t];
  const int key = segkey(A.g), eb = binade(A.g);
  const bool ok0 = !((A.meta >> 18) & 3) && eb >= 64 && eb <= 2045;
  const double ui = ok0 ? pow2_field(2098 - eb) : 0.0;
  const double sig = A.g < 0.0 ? -1.0 : 1.0;
  const double G1 = ok0 ? (fabs(A.g) - pow2_field(eb)) * ui : 0.0;
  bool okm = true;
  int mapm = 2;  // identity
  double d0m = 0.0, d1m = 0.0, Li = -0x1p60, Ui = 0x1p60;
  if (mem && single && ok0) {
    okm = (i == 0 || (((Ri.meta >> 18) & 3) == 0 && segkey(Ri.g) == key)) &&
          run_inside(Ri.g, eb, Ri.lo, Ri.hi);
    const double rlo = sig > 0.0 ? Ri.lo : -Ri.hi;
    const double rhi = sig > 0.0 ? Ri.hi : -Ri.lo;
    Li = ceil(5.0 - rlo * ui) + 1.0;
    Ui = floor(0x1p52 - 5.0 - rhi * ui) - 1.0;
    const bool last = i + 1 == nm;
    const bool has_next = w0 + nm < nch;
    const double gn = !last ? first[t + 1].g : ((wl + 1) * kWin < 64 ? first[(wl + 1) * kWin].g : gnx);
    const int am = (!last || has_next) ? parity_of((Ri.g + Ri.d0) - gn, ui) : 0;
    const int bm = parity_of(Ri.d1 - Ri.d0, ui);
    mapm = ((bm ? 0 : 1) << 1) | am;  // p -> bm ? am : p ^ am
    d0m = Ri.d0;
    d1m = Ri.d1;
  }
  // start parity of member i for window start parity 0 and 1
  const int rl = t & 15;
  const int incl = row_incl_map(mapm);
  int ex = dpp_i<0x111>(incl);
  if (rl == 0) ex = 2;
  const int pa = (ex >> 1) ? (0 ^ (ex & 1)) : (ex & 1);
  const int pb = (ex >> 1) ? (1 ^ (ex & 1)) : (ex & 1);
  const double da = mem ? (pa ? d1m : d0m) : 0.0, db = mem ? (pb ? d1m : d0m) : 0.0;
  const double ia = row_incl_add(da), ib = row_incl_add(db);
  double xa = dpp_d<0x111>(ia), xb = dpp_d<0x111>(ib);
  if (rl == 0) xa = xb = 0.0;
  // admissible start offsets: member i's start is G1 + delta + sig D / u
  double lo0 = -0x1p60, hi0 = 0x1p60, lo1 = -0x1p60, hi1 = 0x1p60;
  if (mem && single && ok0) {
    const double Oa = sig * xa * ui, Ob = sig * xb * ui;
    lo0 = Li - G1 - Oa;
    hi0 = Ui - G1 - Oa;
    lo1 = Li - G1 - Ob;
    hi1 = Ui - G1 - Ob;
    // partial translations stay well inside 2^53 u, so every sum is exact
    if (fabs(ia) * ui > 0x1p51 || fabs(ib) * ui > 0x1p51) okm = false;
  }
#pragma unroll
  for (int m = 1; m < kWin; m <<= 1) {
    lo0 = fmax(lo0, __shfl_xor(lo0, m));
    hi0 = fmin(hi0, __shfl_xor(hi0, m));
    lo1 = fmax(lo1, __shfl_xor(lo1, m));
    hi1 = fmin(hi1, __shfl_xor(hi1, m));
  }
  const bool ok = single && ok0 && ((__ballot(!okm) >> (wl * kWin)) & 0xffff) == 0;
  // totals and the whole window's map sit in member nm - 1's lane
  const int src = wl * kWin + (nm > 0 ? nm - 1 : 0);
  const double Da = __shfl(ia, src), Db = __shfl(ib, src);
  const int mw = __shfl(incl, src);
  if (i == 0 && nm > 0) {
    CompRec C;
    if (single) {
      const int pn0 = (mw >> 1) ? (0 ^ (mw & 1)) : (mw & 1);
      const int pn1 = (mw >> 1) ? (1 ^ (mw & 1)) : (mw & 1);
      C.g = A.g;
      C.d0 = Da;
      C.d1 = Db;
      C.lo0 = lo0;
      C.hi0 = hi0;
      C.lo1 = lo1;
      C.hi1 = hi1;
      C.meta = nm | ((ok ? 1 : 0) << 5) | ((((pn0 ^ pn1) << 1) | pn0) << 6);
    } else {
      C.g = C.d0 = C.d1 = C.lo0 = C.hi0 = C.lo1 = C.hi1 = 0.0;
      C.meta = nm;
    }
    C.win = (int32_t)(w0 / kWin);
    C.cnts = cn;
    cps[w0 / kWin] = C;
  }
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

// The member chain over up to 64 compacted records (one per lane).  Runs
// of at most kWalk records are walked one record at a time (a handful of
// dependent operations each); longer runs take the wavefront scans.
static constexpr int kWalk = 4;

__device__ void chain_group(const double *__restrict__ a, const SegRec &R, int nv, double &S) {
  const int l = lane_id();
  const bool valid = l < nv;
  const int head = valid ? (R.meta >> 18) & 1 : 0;
  const int rep = valid ? (R.meta >> 19) & 1 : 0;
  const int key = segkey(R.g);
  const double gn = dpp_d<0x130>(R.g);  // next lane's guess (wave_shl:1)
  // Everything that depends on the lane's own record only, computed once:
  // 1/u = 2^(52 - E) of its binade (exact; used only while in range), its
  // parity map p -> b ? a : p ^ a towards the next lane (used only when the
  // next lane continues the run, hence shares the binade), and whether the
  // guessed run itself keeps the margins.
  const int eb = binade(R.g);
  const bool eb_ok = eb >= 2 && eb <= 2045;
  const double ui = eb_ok ? pow2_field(2098 - eb) : 0.0;
  const int map_own = ((parity_of(R.d1 - R.d0, ui) ? 0 : 1) << 1) |
                      parity_of((R.g + R.d0) - gn, ui);
  const bool g_inside = eb_ok && run_inside(R.g, eb, R.lo, R.hi);
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
    if (h - q <= kWalk) {
      // a short run (the records between binade crossings): each record is
      // checked and translated on its own, S being exact at its start
      for (int r = q; r < h; r++) {
        const double g_r = readlane_d(R.g, r), d0_r = readlane_d(R.d0, r);
        const double d1_r = readlane_d(R.d1, r), lo_r = readlane_d(R.lo, r);
        const double hi_r = readlane_d(R.hi, r), ui_r = readlane_d(ui, r);
        const int gin_r = __builtin_amdgcn_readlane((int)g_inside, r);
        if (gin_r && segkey(S) == key_q && run_inside(S, binade(g_r), lo_r, hi_r)) {
          S = S + (parity_of(S - g_r, ui_r) ? d1_r : d0_r);
        } else {
          const int64_t chunk_r = (int64_t)__builtin_amdgcn_readlane(R.chunk, r);
          const int meta_r = __builtin_amdgcn_readlane(R.meta, r);
          const int64_t b = chunk_r * kChunk + (meta_r & 511) + ((meta_r >> 18) & 1);
          const int64_t e = chunk_r * kChunk + ((meta_r >> 9) & 511);
          S = lane_seq(a, b, e, S);
        }
      }
      q = h;
      continue;
    }
    const bool in = l >= q && l < h;
    const double g_q = readlane_d(R.g, q);
    const double ui_q = readlane_d(ui, q);
    // start parities: p_(l+1) = b_l ? a_l : p_l ^ a_l, a GF(2) affine scan
    const int mp = (in && l + 1 < h) ? map_own : 2;  // 2: identity
    const int incl_m = wave_incl_map(mp);
    int ex_m = dpp_i<0x138>(incl_m);
    if (l == 0) ex_m = 2;
    // maps of lanes < q are identities, so each prefix starts at lane q
    const int p_q = segkey(S) == key_q ? parity_of(S - g_q, ui_q) : 0;
    const int p = (ex_m >> 1) ? (p_q ^ (ex_m & 1)) : (ex_m & 1);
    const double d = in ? (p ? R.d1 : R.d0) : 0.0;
    // lanes < q add zeros, so lane l's exclusive partial covers [q, l)
    const double excl = wave_prev_d(wave_incl_add(d));
    const double Sl = S + excl;
    bool acc = false;
    if (in) {
      acc = g_inside && segkey(Sl) == key && run_inside(Sl, eb, R.lo, R.hi) &&
            parity_of(Sl - R.g, ui) == p;
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

// Records of window win (chunks kWin win .. + nm, counts packed in cnts),
// one per lane in order, summed by the member chain from S.
__device__ void member_level(const double *__restrict__ a, const SegRec *__restrict__ sg, int cap,
                             int64_t win, int nm, uint64_t cnts, double &S) {
  const int l = lane_id();
  const int64_t c0 = win * kWin;
  // lane r takes record r: chunk i with off_i <= r < off_i + cnt_i
  int ch = 0, k = l, tot = 0;
  for (int i = 0; i < nm; i++) {
    if (l >= tot) {
      ch = i;
      k = l - tot;
    }
    tot += (cnts >> (4 * i)) & 15;
  }
  SegRec R;
  if (l < tot) {
    R = sg[(c0 + ch) * cap + k];
  } else {
    R.meta = 0, R.g = 0.0, R.d0 = R.d1 = R.lo = R.hi = R.xh = 0.0, R.chunk = 0;
  }
  chain_group(a, R, tot, S);
}

// ---- class tables of the windows that did not fold -------------------------
// A window whose guessed run crosses binades (heads) is a translation only
// per residue class of the start: for a start S = G_c + Delta, G_c = g + c u
// (g the window's first guess, u the grid of g's binade, c < M) and Delta a
// multiple of M u = 2 u_max (u_max the coarsest grid the run visits), every
// rounding of the run from S happens on the grid it happens on from G_c,
// displaced by Delta, as long as every partial sum stays in its binade:
// the run then ends exactly at E_c + Delta.  seq_classes walks the window's
// records from the 32 starts G_0..G_31 (one lane each, heads added, rejected
// records summed in order), intersecting the admissible Delta interval over
// every step (4-ulp margins, as run_inside).  The chain then resolves such a
// window with one lookup: c from the low mantissa bits of S - g, Delta =
// S - G_c, a range check, S = E_c + Delta; a start outside its interval, or
// a window whose run climbs more than 4 binades (M > 32), takes the member
// chain.
static constexpr int kClasses = 32;

struct ClassTab {
  double e[kClasses], gs[kClasses], lo[kClasses], hi[kClasses];
};

// Delta keeps y + Delta in y's binade (margins of 4 ulps); a power of two or
// a value outside the normal range admits Delta = 0 only.
__device__ __forceinline__ void keep_binade(double y, double &dlo, double &dhi, int &emax) {
  const uint64_t bits = (uint64_t)__double_as_longlong(y);
  const int eb = (int)((bits >> 52) & 0x7ff);
  if (eb < 64 || eb > 2045 || (bits & 0xfffffffffffffull) == 0) {
    dlo = fmax(dlo, 0.0);
    dhi = fmin(dhi, 0.0);
    emax = max(emax, eb);
    return;
  }
  const double bl = pow2_field(eb), bh = pow2_field(eb + 1), m = pow2_field(eb - 50);
  const double L = y > 0.0 ? bl + m : -bh + m, U = y > 0.0 ? bh - m : -bl - m;
  dlo = fmax(dlo, L - y);
  dhi = fmin(dhi, U - y);
  emax = max(emax, eb);
}

// The values a[b, e) added in order to S, with the binade constraints.
__device__ __forceinline__ double ordered_steps(const double *__restrict__ a, int64_t b, int64_t e,
                                               double S, double &dlo, double &dhi, int &emax) {
  for (int64_t i = b; i < e; i++) {
    S = S + a[i];
    keep_binade(S, dlo, dhi, emax);
  }
  return S;
}

__global__ void __launch_bounds__(64) seq_classes(const double *__restrict__ a,
                                                  const SegRec *__restrict__ sg, int cap,
                                                  CompRec *__restrict__ cp, int64_t nw,
                                                  int32_t *__restrict__ tab_m,
                                                  ClassTab *__restrict__ tab) {
  const int64_t w = blockIdx.x;
  const CompRec C = cp[w];
  if ((C.meta >> 5) & 1) return;  // folded: no table needed
  const int l = threadIdx.x, nm = C.meta & 31;
  const int64_t c0 = w * kWin;
  // lane r holds record r of the window (as the member chain)
  int ch = 0, k = l, tot = 0;
  for (int i = 0; i < nm; i++) {
    if (l >= tot) {
      ch = i;
      k = l - tot;
    }
    tot += (C.cnts >> (4 * i)) & 15;
  }
  SegRec R;
  if (l < tot) {
    R = sg[(c0 + ch) * cap + k];
  } else {
    R.meta = 0, R.g = 0.0, R.d0 = R.d1 = R.lo = R.hi = R.xh = 0.0, R.chunk = 0;
  }
  const double g = readlane_d(R.g, 0);
  const int e0 = binade(g);
  // (a replay record carries its chunk's guess, so a table can start there)
  bool ok = tot > 0 && e0 >= 64 && e0 <= 2045;
  const double u = ok ? pow2_field(e0) * 0x1p-52 : 0.0;
  const int c = l < kClasses ? l : 0;
  double S = g + (double)c * u;  // exact: c < 2^5 ulps
  ok = ok && segkey(S) == segkey(g);
  double dlo = -0x1p1000, dhi = 0x1p1000;
  int emax = e0;
  // per-record quantities that do not depend on the start
  const int ebr = binade(R.g);
  const bool ebr_ok = ebr >= 2 && ebr <= 2045;
  const double uir = ebr_ok ? pow2_field(2098 - ebr) : 0.0;
  const bool gin = ebr_ok && run_inside(R.g, ebr, R.lo, R.hi);
  for (int r = 0; r < tot && __ballot(ok); r++) {
    const int meta = __builtin_amdgcn_readlane(R.meta, r);
    const int64_t base = (int64_t)__builtin_amdgcn_readlane(R.chunk, r) * kChunk;
    const int64_t b = base + (meta & 511), e = base + ((meta >> 9) & 511);
    if ((meta >> 19) & 1) {
      S = ordered_steps(a, b, e, S, dlo, dhi, emax);
      continue;
    }
    const int head = (meta >> 18) & 1;
    if (head) {
      S = S + readlane_d(R.xh, r);
      keep_binade(S, dlo, dhi, emax);
    }
    const double gr = readlane_d(R.g, r), lor = readlane_d(R.lo, r), hir = readlane_d(R.hi, r);
    const int eb = binade(gr);
    if (__builtin_amdgcn_readlane((int)gin, r) && segkey(S) == segkey(gr) &&
        run_inside(S, eb, lor, hir)) {
      // the record's run from S + Delta is its run from S displaced: the
      // range of run_inside, relative to S
      const double bl = pow2_field(eb), bh = pow2_field(eb + 1), m = 4.0 * bl * 0x1p-52;
      const double rlo = S < 0.0 ? -hir : lor, rhi = S < 0.0 ? -lor : hir;
      const double ax = fabs(S);
      const double alo = (bl + m - rlo) - ax, ahi = (bh - m - rhi) - ax;  // on |S|
      dlo = fmax(dlo, S < 0.0 ? -ahi : alo);
      dhi = fmin(dhi, S < 0.0 ? -alo : ahi);
      emax = max(emax, eb);
      const int p = parity_of(S - gr, readlane_d(uir, r));
      S = S + (p ? readlane_d(R.d1, r) : readlane_d(R.d0, r));
    } else {
      S = ordered_steps(a, b + head, e, S, dlo, dhi, emax);
    }
  }
  // M u = 2 u_max over every lane's run
  int em = emax;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) em = max(em, __shfl_xor(em, o));
  const int lgm = 1 + em - e0;
  const int M = (lgm >= 1 && lgm <= 5) ? (1 << lgm) : 0;
  const bool all_ok = __ballot(l < M && !ok) == 0;
  if (l < kClasses) {
    tab[w].e[l] = S;
    tab[w].gs[l] = g + (double)c * u;
    tab[w].lo[l] = dlo;
    tab[w].hi[l] = dhi;
  }
  const bool valid = M > 0 && all_ok && __ballot(ok) != 0;
  if (l == 0) tab_m[w] = valid ? M : 0;
  // Two classes whose runs end in g's binade: the window is a composite of
  // the ordinary kind (start parity = class, translations E_c - G_c exact
  // in one binade, the class intervals in units of u), so the chain's scan
  // takes it without a break.
  if (valid && M == 2) {
    const double E0 = readlane_d(S, 0), E1 = readlane_d(S, 1);
    const double lo0 = readlane_d(dlo, 0), hi0 = readlane_d(dhi, 0);
    const double lo1 = readlane_d(dlo, 1), hi1 = readlane_d(dhi, 1);
    const int key = segkey(g);
    if (segkey(E0) == key && segkey(E1) == key && l == 0) {
      const double ui = pow2_field(2098 - e0);  // 1/u of g's binade
      const double g1 = g + u;
      int pn0 = 0, pn1 = 0;
      if (w + 1 < nw) {
        const double gn = sg[(w + 1) * kWin * cap].g;
        if (segkey(gn) == key) {
          pn0 = parity_of(E0 - gn, ui);
          pn1 = parity_of(E1 - gn, ui);
        }
      }
      CompRec D = C;
      D.g = g;
      D.d0 = E0 - g;
      D.d1 = E1 - g1;
      // delta = sign(g) (S - g) / u is an integer: the class intervals,
      // tightened to integers
      if (g > 0.0) {
        D.lo0 = ceil(lo0 * ui);
        D.hi0 = floor(hi0 * ui);
        D.lo1 = ceil(1.0 + lo1 * ui);
        D.hi1 = floor(1.0 + hi1 * ui);
      } else {
        D.lo0 = ceil(-(hi0 * ui));
        D.hi0 = floor(-(lo0 * ui));
        D.lo1 = ceil(-(1.0 + hi1 * ui));
        D.hi1 = floor(-(1.0 + lo1 * ui));
      }
      D.meta = nm | (1 << 5) | ((((pn0 ^ pn1) << 1) | pn0) << 6);
      cp[w] = D;
    }
  }
}

// The window's run from S by its class table; false when the table does
// not cover S (the caller then takes the member chain).
__device__ __forceinline__ bool class_lookup(const int32_t *__restrict__ tab_m,
                                             const ClassTab *__restrict__ tab,
                                             const SegRec *__restrict__ sg, int cap, int64_t win,
                                             double &S) {
  const int M = tab_m[win];
  if (M == 0) return false;
  const int l = lane_id();
  const ClassTab &T = tab[win];
  const double e = l < kClasses ? T.e[l] : 0.0, gs = l < kClasses ? T.gs[l] : 0.0;
  const double lo = l < kClasses ? T.lo[l] : 0.0, hi = l < kClasses ? T.hi[l] : 0.0;
  const double g = readlane_d(gs, 0);
  if (segkey(S) != segkey(g)) return false;
  // k = (S - g) / u: the difference of the mantissas, negated for g < 0
  const int dm = (int)((uint32_t)__double2loint(S) - (uint32_t)__double2loint(g));
  const int c = (g < 0.0 ? -dm : dm) & (M - 1);
  const double delta = S - readlane_d(gs, c);  // exact: one binade
  if (!(delta >= readlane_d(lo, c) && delta <= readlane_d(hi, c))) return false;
  S = readlane_d(e, c) + delta;
  return true;
}

// A window that did not fold: its class table, else the member chain.
__device__ void window_level(const double *__restrict__ a, const SegRec *__restrict__ sg, int cap,
                             const int32_t *__restrict__ tab_m, const ClassTab *__restrict__ tab,
                             int64_t win, int nm, uint64_t cnts, double &S) {
  if (class_lookup(tab_m, tab, sg, cap, win, S)) return;
  member_level(a, sg, cap, win, nm, cnts, S);
}

// The chain over 64 windows: as chain_group, with each composite's parity
// map, translations and admissible intervals; windows that did not fold,
// and the first rejected one, go to the member chain.
__device__ void chain_comp(const double *__restrict__ a, const SegRec *__restrict__ sg, int cap,
                           const int32_t *__restrict__ tab_m, const ClassTab *__restrict__ tab,
                           const CompRec &C, int nv, double &S) {
  const int l = lane_id();
  const bool valid = l < nv;
  const int ok = valid ? (C.meta >> 5) & 1 : 0;
  const int key = segkey(C.g);
  int q = 0;
  while (q < nv) {
    const int meta_q = __builtin_amdgcn_readlane(C.meta, q);
    const int win_q = __builtin_amdgcn_readlane(C.win, q);
    if (!((meta_q >> 5) & 1)) {
      window_level(a, sg, cap, tab_m, tab, win_q, meta_q & 31, readlane_u64(C.cnts, q), S);
      q++;
      continue;
    }
    const int key_q = __builtin_amdgcn_readlane(key, q);
    const uint64_t stop = __ballot(valid && l > q && (!ok || key != key_q));
    const int h = stop ? __ffsll((unsigned long long)stop) - 1 : nv;
    const bool in = l >= q && l < h;
    const double g_q = readlane_d(C.g, q);
    const int eq = binade(g_q);
    const double ui = (eq >= 2 && eq <= 2045) ? pow2_field(2098 - eq) : 0.0;
    const int mp = (in && l + 1 < h) ? (C.meta >> 6) & 3 : 2;
    const int incl_m = wave_incl_map(mp);
    int ex_m = dpp_i<0x138>(incl_m);
    if (l == 0) ex_m = 2;
    const int p_q = segkey(S) == key_q ? parity_of(S - g_q, ui) : 0;
    const int p = (ex_m >> 1) ? (p_q ^ (ex_m & 1)) : (ex_m & 1);
    const double d = in ? (p ? C.d1 : C.d0) : 0.0;
    const double excl = wave_prev_d(wave_incl_add(d));
    const double Sl = S + excl;
    bool acc = false;
    if (in && segkey(Sl) == key) {
      const double off = Sl - C.g;  // exact: one binade
      const double delta = (C.g < 0.0 ? -off : off) * ui;
      acc = parity_of(off, ui) == p && delta >= (p ? C.lo1 : C.lo0) &&
            delta <= (p ? C.hi1 : C.hi0);
    }
    const uint64_t rej = __ballot(in && !acc);
    if (rej == 0) {
      S = readlane_d(Sl + d, h - 1);
      q = h;
      continue;
    }
    const int f = __ffsll((unsigned long long)rej) - 1;
    S = readlane_d(Sl, f);
    const int meta_f = __builtin_amdgcn_readlane(C.meta, f);
    const int win_f = __builtin_amdgcn_readlane(C.win, f);
    window_level(a, sg, cap, tab_m, tab, win_f, meta_f & 31, readlane_u64(C.cnts, f), S);
    q = f + 1;
  }
}

// Level 2: the chain over the windows.  Wave 0 runs the chain; waves 1-3
// stream the next block of composites into LDS meanwhile (double buffered,
// one barrier per block).
static constexpr int kBlockComps = 256;

__global__ void __launch_bounds__(256) seq_chain(const double *__restrict__ a,
                                                 const SegRec *__restrict__ sg, int cap,
                                                 const CompRec *__restrict__ cp, int64_t M,
                                                 const int32_t *__restrict__ tab_m,
                                                 const ClassTab *__restrict__ tab, double *out) {
  __shared__ CompRec buf[2][kBlockComps];
  __shared__ int sink;
  const int t = threadIdx.x, w = t >> 6;
  const int64_t nblk = (M + kBlockComps - 1) / kBlockComps;
  double S = 0.0;  // the reference's accumulator starts at +0
  for (int j = t; j < kBlockComps; j += 256)
    if (j < M) buf[0][j] = cp[j];
  __syncthreads();
  for (int64_t k = 0; k < nblk; k++) {
    if (w != 0) {
      const int64_t b1 = (k + 1) * kBlockComps;
      uint64_t touched = 0;
      for (int j = t - 64; j < kBlockComps; j += 192) {
        if (b1 + j >= M) break;
        const CompRec C = cp[b1 + j];
        buf[(k + 1) & 1][j] = C;
        if (!((C.meta >> 5) & 1)) {
          // the member chain will read this window's records: one load per
          // 128-byte line brings them into this CU's caches meanwhile
          const char *base = (const char *)(sg + (int64_t)C.win * kWin * cap);
          const int bytes = (C.meta & 31) * cap * (int)sizeof(SegRec);
          for (int o = 0; o < bytes; o += 128) touched ^= *(const uint64_t *)(base + o);
        }
      }
      if (touched == 0x7ff4dead0badf00dull) sink = 1;  // keeps the loads
    } else {
      const int64_t b0 = k * kBlockComps;
      for (int g0 = 0; g0 < kBlockComps && b0 + g0 < M; g0 += 64) {
        const int nv = (int)min((int64_t)64, M - b0 - g0);
        CompRec C;
        if (t < nv) {
          C = buf[k & 1][g0 + t];
        } else {
          C.g = C.d0 = C.d1 = C.lo0 = C.hi0 = C.lo1 = C.hi1 = 0.0;
          C.win = 0;
          C.meta = 0;
          C.cnts = 0;
        }
        chain_comp(a, sg, cap, tab_m, tab, C, nv, S);
      }
    }
    __syncthreads();
  }
  if (t == 0) *out = S;
}

__global__ void seq_zero(double *out) {
  if (threadIdx.x == 0) *out = 0.0;
}

// records per chunk (a window of kWin chunks must fit one wavefront); very
// long arrays keep 2 to bound the work space
static int seg_cap(int64_t nch) { return nch <= 65536 ? 64 / kWin : 2; }

struct SeqLayout {
  int64_t nch, nb, nw, tot, blk, segs, cnt, cps, tabm, tab, bytes;
  int cap;
};

static SeqLayout seq_layout(int64_t n) {
  SeqLayout L;
  L.nch = (n + kChunk - 1) / kChunk;
  L.nb = (L.nch + 63) / 64;
  L.nw = (L.nch + kWin - 1) / kWin;
  L.cap = seg_cap(L.nch);
  auto al = [](int64_t x) { return (x + 255) / 256 * 256; };
  L.tot = 0;
  L.blk = al(L.tot + 8 * L.nch);
  L.segs = al(L.blk + 16 * L.nb);
  L.cnt = al(L.segs + (int64_t)sizeof(SegRec) * L.nch * L.cap);
  L.cps = al(L.cnt + 4 * L.nch);
  L.tabm = al(L.cps + (int64_t)sizeof(CompRec) * L.nw);
  L.tab = al(L.tabm + 4 * L.nw);
  L.bytes = al(L.tab + (int64_t)sizeof(ClassTab) * L.nw);
  return L;
}

// The four launches of one sum over a work area laid out by seq_layout(n).
static int seqsum_launch(const double *a, int64_t n, double *dout, char *w, const SeqLayout &L,
                         hipStream_t s) {
  double *tot = (double *)(w + L.tot);
  double2 *blk = (double2 *)(w + L.blk);
  SegRec *segs = (SegRec *)(w + L.segs);
  int32_t *cnt = (int32_t *)(w + L.cnt);
  CompRec *cps = (CompRec *)(w + L.cps);
  int32_t *tabm = (int32_t *)(w + L.tabm);
  ClassTab *tab = (ClassTab *)(w + L.tab);
  if (n == 0) {
    hipLaunchKernelGGL(seq_zero, dim3(1), dim3(64), 0, s, dout);
    HYD_HIP(hipGetLastError());
    return HYD_OK;
  }
  hipLaunchKernelGGL(seq_chunks<0>, dim3((unsigned)L.nb), dim3(64), 0, s, a, n, L.nch, tot, blk,
                     segs, cnt, cps, L.cap);
  hipLaunchKernelGGL(seq_chunks<1>, dim3((unsigned)L.nb), dim3(64), 0, s, a, n, L.nch, tot, blk,
                     segs, cnt, cps, L.cap);
  hipLaunchKernelGGL(seq_classes, dim3((unsigned)L.nw), dim3(64), 0, s, a, segs, L.cap, cps, L.nw,
                     tabm, tab);
  hipLaunchKernelGGL(seq_chain, dim3(1), dim3(256), 0, s, a, segs, L.cap, cps, L.nw, tabm, tab,
                     dout);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int seqsum(Ctx &c, const double *a, int64_t n, double *dout) {
  const SeqLayout L = seq_layout(n);
  if (L.nch > c.seq_cap || seq_layout((int64_t)c.seq_cap * kChunk).bytes < L.bytes) {
    set_error("seqsum: array longer than the context's work space");
    return HYD_ECAPACITY;
  }
  KTimer t(c, "seqsum", 8 * n);
  return seqsum_launch(a, n, dout, (char *)c.seq_work, L, c.s);
}

// ---- the sequential sum without allocation ---------------------------------
// Work area in a file-scope device array: each GPU the library runs on holds
// its own copy (19.06 MiB used of 24 MiB for the largest sum, 2^22 values)
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a host-side stand-in): each device's
// address comes from hipGetSymbolAddress, taken once per device, with that
// device current, on the first call there.
static constexpr int64_t kNoAllocMaxN = (int64_t)1 << 22;
static constexpr int64_t kNoAllocBytes = (int64_t)24 << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

static int scratch_address(void **out) {
  int dev = 0;
  HYD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("sum_sequential_noalloc: device ordinal beyond the library's table");
    return HYD_EARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HYD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("sum_sequential_noalloc: no device address for the work area");
      return HYD_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HYD_OK;
}

int64_t seqsum_work_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

int64_t seqsum_work_bytes(int64_t n) { return seq_layout(n).bytes; }

}  // namespace hyd

extern "C" int hyd_sum_sequential_noalloc(const double *a, int64_t n, double *sum, void *stream) {
  using namespace hyd;
  if (n < 0 || n > kNoAllocMaxN) {
    set_error("sum_sequential_noalloc: length outside [0, 2^22]");
    return n < 0 ? HYD_EARG : HYD_ECAPACITY;
  }
  if (!sum || (n > 0 && !a)) {
    set_error("sum_sequential_noalloc: null device pointer");
    return HYD_EARG;
  }
  const SeqLayout L = seq_layout(n);
  if (L.bytes > kNoAllocBytes) {
    set_error("sum_sequential_noalloc: work area larger than the static scratch");
    return HYD_ECAPACITY;
  }
  void *w = nullptr;
  HYD_CHECK(scratch_address(&w));
  return seqsum_launch(a, n, sum, (char *)w, L, (hipStream_t)stream);
}
