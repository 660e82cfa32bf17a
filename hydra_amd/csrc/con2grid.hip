// con2grid.hip -- contour -> grid conversion, contours.f90:173-346, on gfx950.
//
// The reference accumulates +/-qjump into fine cells in node order, sweeps
// each fine column sequentially in y, averages twice with a nine-point
// stencil and removes the sequential mean.  Every one of those sums is
// reproduced bit for bit:
//
//   * crossings are found segment-parallel; each one adds to a packed int32
//     counter of its fine cell (low 16 bits: +qjump count, high 16: -qjump).
//     A cell's sequential node-order sum depends only on the counts unless it
//     holds both signs with at least three of one sign; those rare cells are
//     flagged on the fly and re-summed in node order from per-cell buckets.
//   * the y sweep runs one wavefront per fine column over 64-row segments;
//     only cells with crossings change the running value, so the serial
//     chain is the column's crossings, in row order, exactly as the
//     reference adds them.
//   * the bottom row qbot (contours.f90:234-238) is the same chain over the
//     x-crossings of y=ymin.
//   * the nine-point averages (contours.f90:290-327) are pointwise with the
//     reference's expression order.
//   * the mean (contours.f90:330-336) is the exact sequential sum (seqsum).
//
// Fine-grid layout here: the swept field is column-major by fine column,
// rows iy = 1..nyf, index (ix-1)*nyf + (iy-1), which is also the key of a
// cell.  The counters are stored by groups of 4 fine columns, the group's
// 4 counters of a row side by side (cnt_at): a 128-byte line holds 8 rows
// of 4 columns, so the crossings a contour makes in neighbouring columns
// share lines, which halves the cost of the counter atomics.  nxf % 4 == 0.
#include "hyd_internal.h"

namespace hyd {

// Order-dependent cells ("flagged": both signs, at least three of one) are
// re-summed in node order from per-cell buckets: an open-addressing table
// maps the cell key to a bucket; a claim pass gives every flagged cell a
// bucket of exactly p+q slots, a fill pass drops each crossing's (node, sign)
// into its cell's bucket, and one thread per bucket orders its few entries by
// node and adds them as the reference does.  Nothing is sorted globally and
// the table is left empty again at the end of the sweep.
struct Buckets {
  int32_t *tkey;    // table keys, -1 = empty
  int32_t *tid;     // bucket of each table slot
  int32_t *list;    // per bucket: table slot, first entry, size, fill count
  int32_t *slots;   // entries: node << 1 | (sign > 0)
  double *vals;     // node-order sum of each bucket
  uint32_t tmask;   // table size - 1 (power of two)
  int32_t bcap;     // buckets
  int64_t scap;     // entries
};

struct C2GArgs {
  const double *xq, *yq;
  const int32_t *nextq;
  int n;
  int ultra;         // 0: con2grid conventions, 1: con2ugrid (congen.f90:353-416)
  int nxf, nyf;
  double ellx, elly, hlx, hly, hlxi, hlyi, xmin, ymin, glxfi, glyfi, small, qjump;
  const double *xgf;
  int32_t *cnt;      // nyf*nxf packed counters
  int32_t *qjx;      // nxf packed counters (bottom row)
  // [0] flagged cells, [1] flagged qjx, [2] bucket entries used, [4] buckets
  // used, [5] dropped edge crossings; [8] sticky error (kept across sweeps)
  int32_t *flags;
  Buckets bk;
};

// Address of the counter of fine cell (column cx, row cy), both 0-based.
__device__ __forceinline__ int64_t cnt_at(int cx, int cy, int nyf) {
  return (int64_t)(cx >> 2) * 4 * nyf + 4 * cy + (cx & 3);
}

__device__ __forceinline__ bool is_flagged(int32_t w) {
  const int p = w & 0xffff, q = (w >> 16) & 0xffff;
  return p > 0 && q > 0 && (p >= 3 || q >= 3);
}

__device__ __forceinline__ void add_count(int32_t *slot, int sign, int32_t *flag) {
  const int32_t d = sign > 0 ? 1 : (1 << 16);
  const int32_t old = atomicAdd(slot, d);
  if (!is_flagged(old) && is_flagged(old + d)) atomicAdd(flag, 1);
}

__device__ __forceinline__ uint32_t bk_hash(int32_t key, uint32_t mask) {
  return ((uint32_t)key * 2654435761u) & mask;
}

// table slot holding key, or -1
__device__ __forceinline__ int64_t bk_find(const Buckets &b, int32_t key) {
  uint32_t h = bk_hash(key, b.tmask);
  for (uint32_t probe = 0; probe <= b.tmask; probe++) {
    const int32_t k = b.tkey[h];
    if (k == key) return h;
    if (k == -1) return -1;
    h = (h + 1) & b.tmask;
  }
  return -1;
}

// The segment from node i to its successor: fine cell of the node,
// periodic increments, x-crossing count (contours.f90:206-231, congen.f90:
// 354-383).  Both node passes evaluate it the same way, so the bucket pass
// recomputes it rather than reading it back.
struct SegGeom {
  double xi, yi, dx, dy, yy;
  int ixci, nxc;
};
__device__ __forceinline__ SegGeom seg_geom(const C2GArgs &a, int i) {
  SegGeom s;
  s.xi = a.xq[i];
  s.yi = a.yq[i];
  const int ia = a.nextq[i] - 1;
  const double xa = a.xq[ia], ya = a.yq[ia];
  // fine cell of each node: contours.f90:207 (1-based) / congen.f90:354 (0-based)
  const int base = a.ultra ? 0 : 1;
  s.ixci = base + (int)(a.glxfi * (s.xi - a.xmin));
  const int ixca = base + (int)(a.glxfi * (xa - a.xmin));
  const double xx = xa - s.xi;
  s.dx = xx - a.ellx * (double)(int)(xx * a.hlxi);
  s.yy = ya - s.yi;
  s.dy = s.yy - a.elly * (double)(int)(s.yy * a.hlyi);
  const int ixdif = ixca - s.ixci;
  s.nxc = ixdif - a.nxf * ((2 * ixdif) / a.nxf);
  return s;
}

// Node phase: contours.f90:206-231 and the crossing loop 249-269.
static constexpr int kCrossBatch = 4;
__global__ void c2g_nodes(C2GArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const SegGeom sg = seg_geom(a, i);
  const double xi = sg.xi, yi = sg.yi, dx = sg.dx, dy = sg.dy, yy = sg.yy;
  const int ixci = sg.ixci, nxc = sg.nxc;
  if (fabs(yy) > a.hly) {
    const double cc = copysign(1.0, dy);
    const double p = -(yi - a.hly * cc) / (dy + a.small);
    const int ix = 1 + (int)(a.glxfi * (fmod(a.ellx + a.hlx + xi + p * dx, a.ellx)));
    // ix = nxf+1 can only come from rounding at the domain edge; the
    // reference would index past qjx there, so it is dropped and flagged
    if (ix >= 1 && ix <= a.nxf) add_count(&a.qjx[ix - 1], cc > 0.0 ? -1 : 1, &a.flags[1]);
    else atomicAdd(&a.flags[5], 1);
  }
  if (nxc != 0) {
    const int jump = nxc >= 0 ? 1 : -1;
    // contours.f90:251-253 vs congen.f90:398-400
    const int ixbeg = a.ultra ? ixci + (1 + jump) / 2 + a.nxf : ixci + (jump - 1) / 2 + a.nxf;
    const int sgn = a.ultra ? jump : (copysign(1.0, dx) > 0.0 ? 1 : -1);
    // crossings in groups of kCrossBatch: the cells of a group are found
    // first and their counter atomics issued together, so a node's
    // crossings cost one atomic round trip per group, not one each
    const int32_t d = sgn > 0 ? 1 : (1 << 16);
    int ncr = 0;
    while (ncr != nxc) {
      int64_t cell[kCrossBatch];
      int ixs[kCrossBatch];
      double xg[kCrossBatch];
      const int nb = min(kCrossBatch, jump * (nxc - ncr));
#pragma unroll
      for (int b = 0; b < kCrossBatch; b++) {
        ixs[b] = 1 + (ixbeg + ncr + jump * b) % a.nxf;
        xg[b] = b < nb ? a.xgf[ixs[b] - 1] : 0.0;
      }
#pragma unroll
      for (int b = 0; b < kCrossBatch; b++) {
        cell[b] = -1;
        if (b < nb) {
          const int ix = ixs[b];
          const double x2 = xg[b] - xi;
          const double px0 = (x2 - a.ellx * (double)(int)(x2 * a.hlxi)) / dx;
          const double y2 = yi + px0 * dy;
          const int iy = 2 + (int)(a.glyfi * (y2 - a.elly * (double)(int)(y2 * a.hlyi) - a.ymin));
          // row nyf+1 is never swept (contours.f90:274): its jumps are dropped
          if (iy >= 2 && iy <= a.nyf) cell[b] = cnt_at(ix - 1, iy - 1, a.nyf);
        }
      }
      ncr += jump * nb;
      int32_t old[kCrossBatch];
#pragma unroll
      for (int b = 0; b < kCrossBatch; b++) old[b] = cell[b] >= 0 ? atomicAdd(&a.cnt[cell[b]], d) : 0;
#pragma unroll
      for (int b = 0; b < kCrossBatch; b++)
        if (cell[b] >= 0 && !is_flagged(old[b]) && is_flagged(old[b] + d)) atomicAdd(&a.flags[0], 1);
    }
  }
}

// Claim (PASS 0) and fill (PASS 1) of the buckets of flagged cells; both
// revisit the crossings of node i exactly as c2g_nodes found them.
template <int PASS>
__device__ __forceinline__ void bk_visit(const C2GArgs &a, int32_t key, int32_t w, int node,
                                         int sgn) {
  const Buckets &b = a.bk;
  if (PASS == 0) {
    uint32_t h = bk_hash(key, b.tmask);
    for (uint32_t probe = 0;; probe++) {
      if (probe > b.tmask) {
        atomicOr(&a.flags[8], 1);
        return;
      }
      const int32_t prev = atomicCAS(&b.tkey[h], -1, key);
      if (prev == key) return;
      if (prev == -1) {
        const int sz = (w & 0xffff) + ((w >> 16) & 0xffff);
        const int nb = atomicAdd(&a.flags[4], 1);
        const int base = atomicAdd(&a.flags[2], sz);
        if (nb >= b.bcap || (int64_t)base + sz > b.scap) {
          b.tid[h] = -1;
          atomicOr(&a.flags[8], 1);
          return;
        }
        b.tid[h] = nb;
        b.list[4 * nb] = (int32_t)h;
        b.list[4 * nb + 1] = base;
        b.list[4 * nb + 2] = sz;
        b.list[4 * nb + 3] = 0;
        return;
      }
      h = (h + 1) & b.tmask;
    }
  } else {
    const int64_t h = bk_find(b, key);
    const int nb = h >= 0 ? b.tid[h] : -1;
    if (nb < 0) {
      atomicOr(&a.flags[8], 1);
      return;
    }
    const int pos = atomicAdd(&b.list[4 * nb + 3], 1);
    if (pos >= b.list[4 * nb + 2]) {
      atomicOr(&a.flags[8], 2);
      return;
    }
    b.slots[b.list[4 * nb + 1] + pos] = (node << 1) | (sgn > 0 ? 1 : 0);
  }
}

template <int PASS>
__device__ __forceinline__ void bk_node(const C2GArgs &a, int i) {
  const SegGeom sg = seg_geom(a, i);
  const double xi = sg.xi, yi = sg.yi, dx = sg.dx, dy = sg.dy, yy = sg.yy;
  const int nxc = sg.nxc, ixci = sg.ixci;
  const int64_t nkey = (int64_t)a.nxf * a.nyf;
  if (a.flags[1] != 0 && fabs(yy) > a.hly) {
    const double cc = copysign(1.0, dy);
    const double p = -(yi - a.hly * cc) / (dy + a.small);
    const int ix = 1 + (int)(a.glxfi * (fmod(a.ellx + a.hlx + xi + p * dx, a.ellx)));
    if (ix >= 1 && ix <= a.nxf) {
      const int32_t w = a.qjx[ix - 1];
      if (is_flagged(w)) bk_visit<PASS>(a, (int32_t)(nkey + ix - 1), w, i, cc > 0.0 ? -1 : 1);
    }
  }
  if (a.flags[0] != 0 && nxc != 0) {
    const int jump = nxc >= 0 ? 1 : -1;
    const int ixbeg = a.ultra ? ixci + (1 + jump) / 2 + a.nxf : ixci + (jump - 1) / 2 + a.nxf;
    const int sgn = a.ultra ? jump : (copysign(1.0, dx) > 0.0 ? 1 : -1);
    for (int ncr = 0; ncr != nxc; ncr += jump) {
      const int ix = 1 + (ixbeg + ncr) % a.nxf;
      const double x2 = a.xgf[ix - 1] - xi;
      const double px0 = (x2 - a.ellx * (double)(int)(x2 * a.hlxi)) / dx;
      const double y2 = yi + px0 * dy;
      const int iy = 2 + (int)(a.glyfi * (y2 - a.elly * (double)(int)(y2 * a.hlyi) - a.ymin));
      if (iy < 2 || iy > a.nyf) continue;
      const int64_t cell = (int64_t)(ix - 1) * a.nyf + (iy - 1);
      const int32_t w = a.cnt[cnt_at(ix - 1, iy - 1, a.nyf)];
      if (is_flagged(w)) bk_visit<PASS>(a, (int32_t)cell, w, i, sgn);
    }
  }
}

// Grid-stride over the nodes: the launch is small, since it usually finds
// no flagged cell and returns at once.
template <int PASS>
__global__ void c2g_buckets(C2GArgs a) {
  if (a.flags[0] == 0 && a.flags[1] == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x)
    bk_node<PASS>(a, i);
}

// Bucket nb: order the entries by node and add them in that order (the
// reference's node loop).
__device__ void bucket_sum(const Buckets &b, int nb, double qjump) {
  const int base = b.list[4 * nb + 1], sz = b.list[4 * nb + 2];
  int32_t *e = b.slots + base;
  for (int t = 1; t < sz; t++) {
    const int32_t v = e[t];
    int u = t - 1;
    while (u >= 0 && (e[u] >> 1) > (v >> 1)) {
      e[u + 1] = e[u];
      u--;
    }
    e[u + 1] = v;
  }
  double s = 0.0;
  for (int t = 0; t < sz; t++) s = s + ((e[t] & 1) ? qjump : -qjump);
  b.vals[nb] = s;
}

// Empty the table again (after the sweep has read the bucket sums; the
// bucket count is the copy c2g_qbot kept in flags[6]) and zero the n4 int4
// of counters for the next sweep, in whole lines.
__global__ void c2g_bucket_clear(Buckets b, int32_t *flags, int4 *cnt, int64_t n4) {
  const int nb_used = min(flags[6], b.bcap);
  const int stride = gridDim.x * blockDim.x;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (int nb = tid; nb < nb_used; nb += stride) {
    b.tkey[b.list[4 * nb]] = -1;
    if (b.list[4 * nb + 3] != b.list[4 * nb + 2]) atomicOr(&flags[8], 2);
  }
  for (int64_t k = tid; k < n4; k += stride) cnt[k] = make_int4(0, 0, 0, 0);
}

// Value of a cell's node-order sum from its counts.
__device__ __forceinline__ double cell_value(int32_t w, int64_t key, double qjump,
                                             const Buckets &b) {
  const int p = w & 0xffff, q = (w >> 16) & 0xffff;
  if (q == 0) {
    double s = 0.0;
    for (int t = 0; t < p; t++) s = s + qjump;
    return s;
  }
  if (p == 0) {
    double s = 0.0;
    for (int t = 0; t < q; t++) s = s + (-qjump);
    return s;
  }
  if (p <= 2 && q <= 2) return (double)(p - q) * qjump;
  const int64_t h = bk_find(b, (int32_t)key);
  const int nb = h >= 0 ? b.tid[h] : -1;
  return nb >= 0 ? b.vals[nb] : 0.0 / 0.0;
}

// cell_value for the common cells first: a single crossing of either sign
// (0 + qjump and 0 - qjump are exact).
__device__ __forceinline__ double cell_value_fast(int32_t w, int64_t key, double qjump,
                                                  const Buckets &b) {
  if (w == 1) return qjump;
  if (w == (1 << 16)) return -qjump;
  return cell_value(w, key, qjump, b);
}

// Inclusive wavefront scan of ints (row shifts, then row broadcasts).
template <int CTRL>
__device__ __forceinline__ int c2g_dpp(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ int wave_scan_i32(int v) {
  const int l = threadIdx.x & 63, rl = l & 15;
  int t;
  t = c2g_dpp<0x111>(v); if (rl >= 1) v += t;
  t = c2g_dpp<0x112>(v); if (rl >= 2) v += t;
  t = c2g_dpp<0x114>(v); if (rl >= 4) v += t;
  t = c2g_dpp<0x118>(v); if (rl >= 8) v += t;
  t = c2g_dpp<0x142>(v); if ((l & 31) >= 16) v += t;
  t = c2g_dpp<0x143>(v); if (l >= 32) v += t;
  return v;
}

// qbot (contours.f90:234-238): qbot(1) = 0, qbot(ix+1) = qbot(ix) + qjx(ix)
// over the non-empty bottom-row counters, in order.  Per pass of kQbotPart
// counters the block evaluates the counters' values and compacts the
// non-empty ones into LDS in order (a block scan of their count); one lane
// then runs the chain over the compacted list, eight values loaded ahead of
// each group of eight dependent adds, overwriting each value with the chain
// after it; every counter then reads the value after the last non-empty one
// at or before it.  The additions are the reference's, in its order.
// Any nxf (passes of kQbotPart).  The bucket sums of the order-dependent cells (usually
// none) are made first, one thread per bucket.  Being the last reader of
// the counters and of flags[0..5], it leaves them zero for the next sweep
// (the bucket count moves to flags[6] for c2g_bucket_clear).
static constexpr int kQbotPart = 4096;  // counters per pass, four per thread
static constexpr int kQbotPer = kQbotPart / 1024;
__global__ void __launch_bounds__(1024) c2g_qbot(int32_t *qjx, double *qbot, int nxf,
                                                 int64_t nkey, double qjump, Buckets bk,
                                                 int32_t *flags) {
  __shared__ double comp[kQbotPart + 8];  // non-empty values, then the chain after each
  __shared__ int wsum[16];
  __shared__ double vin;  // the chain's value entering the pass
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
  const int nb_used = min(flags[4], bk.bcap);
  for (int nb = tid; nb < nb_used; nb += blockDim.x) bucket_sum(bk, nb, qjump);
  __syncthreads();
  double v = 0.0;  // qbot(1); thread 0 carries the chain across passes
  if (tid == 0) qbot[0] = 0.0;
  for (int part = 0; part < nxf - 1; part += kQbotPart) {
    const int m = min(kQbotPart, nxf - 1 - part);
    const int t0 = kQbotPer * tid;
    int32_t w[kQbotPer];
#pragma unroll
    for (int u = 0; u < kQbotPer; u++) w[u] = t0 + u < m ? qjx[part + t0 + u] : 0;
    double cv[kQbotPer];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < kQbotPer; u++) {
      if (w[u]) qjx[part + t0 + u] = 0;
      cv[u] = w[u] ? cell_value(w[u], nkey + part + t0 + u, qjump, bk) : 0.0;
      cnt += w[u] != 0;
    }
    const int incl = wave_scan_i32(cnt);
    if (l == 63) wsum[wv] = incl;
    if (tid == 0) vin = v;
    __syncthreads();
    int off = 0, total = 0;
    for (int k = 0; k < 16; k++) {
      const int ws = wsum[k];
      if (k < wv) off += ws;
      total += ws;
    }
    const int ex = off + incl - cnt;  // non-empty counters before t0
    {
      int e = ex;
#pragma unroll
      for (int u = 0; u < kQbotPer; u++)
        if (w[u]) comp[e++] = cv[u];
    }
    __syncthreads();
    if (tid == 0) {
      double x[8], y[8];
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = comp[u];
      for (int j0 = 0; j0 < total; j0 += 8) {
        // the next group's values are read before this group's adds (the
        // reads beyond total hit the padding and are not used)
#pragma unroll
        for (int u = 0; u < 8; u++) y[u] = comp[min(j0 + 8 + u, kQbotPart + 7)];
        const int nk = min(8, total - j0);
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (u < nk) {
            v = x[u] + v;
            comp[j0 + u] = v;
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = y[u];
      }
    }
    __syncthreads();
    const double v0 = vin;
    int e = ex;
#pragma unroll
    for (int u = 0; u < kQbotPer; u++) {
      e += w[u] != 0;  // non-empty counters up to t0 + u
      if (t0 + u < m) qbot[part + t0 + u + 1] = e > 0 ? comp[e - 1] : v0;
    }
    __syncthreads();
  }
  if (tid == 0) {
    qjx[nxf - 1] = 0;  // the last counter is never part of the chain
    flags[6] = flags[4];
#pragma unroll
    for (int k = 0; k < 6; k++) flags[k] = 0;
    flags[7] = 0;
  }
}

// y sweep (contours.f90:272-277): one wavefront per fine column.
__global__ void c2g_colscan(const int32_t *cnt, const double *qbot, double *fq, int nxf, int nyf,
                            double qjump, Buckets bk) {
  const int wpb = blockDim.x / 64;
  const int col = blockIdx.x * wpb + threadIdx.x / 64;
  const int l = threadIdx.x & 63;
  if (col >= nxf) return;
  double *out = fq + (int64_t)col * nyf;
  double v = qbot[col];
  constexpr int U = 8;  // segments of 64 rows loaded ahead
  for (int base = 0; base < nyf; base += 64 * U) {
    int32_t w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = base + 64 * u + l;
      w[u] = (r < nyf && r >= 1) ? cnt[cnt_at(col, r, nyf)] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = base + 64 * u + l;  // iy-1
      const double cv =
          w[u] ? cell_value(w[u], (int64_t)col * nyf + r, qjump, bk) : 0.0;
      uint64_t mask = __ballot(w[u] != 0);
      double mine = v;
      while (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        v = readlane_f64(cv, b) + v;
        if (l >= b) mine = v;
      }
      if (r < nyf) out[r] = mine;
    }
  }
}

// Nine-point average of an (nh2 x nh2) column-major field f to (nh x nh),
// contours.f90:298-326; out layout column-major with leading dim nh.
template <class T>
__global__ void c2g_avg9(const T *__restrict__ f, T *__restrict__ out, int nh) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ix = blockIdx.y + 1;
  if (iy > nh) return;
  const int n2 = 2 * nh;
#define F(r, c) f[(int64_t)((c) == 0 ? n2 - 1 : (c)-1) * n2 + ((r) == 0 ? n2 - 1 : (r)-1)]
  const int r1 = 2 * iy, r0 = 2 * iy - 2, c1 = 2 * ix, c0 = 2 * ix - 2;
  const T qev0_a = F(r1, c1), qev0_b = F(r0, c1), qev2_a = F(r1, c0), qev2_b = F(r0, c0);
  const T qev1_a = F(r1, c1 - 1), qev1_b = F(r0, c1 - 1);
  const T qod0 = F(r1 - 1, c1), qod2 = F(r1 - 1, c0), qod1 = F(r1 - 1, c1 - 1);
#undef F
  out[(int64_t)(ix - 1) * nh + (iy - 1)] =
      (T)0.0625 * (qev0_a + qev0_b + qev2_a + qev2_b) +
      (T)0.125 * (qev1_a + qev1_b + qod0 + qod2) + (T)0.25 * qod1;
}

// y sweep fused with the first nine-point average (contours.f90:272-277
// and 298-326, the 4ng -> 2ng step): a workgroup of 2K+1 wavefronts sweeps
// the fine columns 2 i0 - 1 .. 2 i0 + 2K - 1 (one each, in the order of the
// y sweep, 64 rows at a time) and averages them into output columns
// i0 .. i0+K-1 through LDS, so the fine field never reaches HBM.  Output
// row 0 needs the last fine row (periodic) and is written at the end.
// Needs n2 % 64 == 0 and (n2 / 2) % K == 0.
static constexpr int kAvgK = 4;
static constexpr int kAvgW = 2 * kAvgK + 1;

__global__ void __launch_bounds__(64 * kAvgW)
    c2g_colscan_avg9(const int32_t *__restrict__ cnt, const double *__restrict__ qbot,
                     double *__restrict__ out, int n2, double qjump, Buckets bk) {
  constexpr int TPS = 4;  // tiles of 64 rows per stage (one barrier each)
  __shared__ double tile[2][kAvgW][64 * TPS + 1];  // [buffer][column][1 + row], 0: carry
  __shared__ double edge[kAvgW][3];                // rows 0, 1 and n2 - 1 of each column
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nh = n2 / 2, ntiles = n2 / 64;
  const int i0 = blockIdx.x * kAvgK;
  int col = 2 * i0 - 1 + w;
  if (col < 0) col += n2;
  const int32_t *cc = cnt + cnt_at(col, 0, n2);  // row r at cc[4 r]
  double v = qbot[col];
  const int tid = threadIdx.x, oc = tid >> 7, jr = tid & 127;
  // counts of the next stage load while the current one is swept
  int32_t wv[TPS], wn[TPS];
#pragma unroll
  for (int u = 0; u < TPS; u++) {
    const int r = 64 * u + l;
    wn[u] = (u < ntiles && r >= 1) ? cc[4 * r] : 0;
  }
  for (int st = 0; st * TPS < ntiles; st++) {
    const int t0 = st * TPS, b = st & 1;
#pragma unroll
    for (int u = 0; u < TPS; u++) {
      wv[u] = wn[u];
      wn[u] = t0 + TPS + u < ntiles ? cc[4 * (64 * (t0 + TPS + u) + l)] : 0;
    }
    if (l == 0) tile[b][w][0] = v;  // row 64 t0 - 1 (unused for st = 0)
#pragma unroll
    for (int u = 0; u < TPS; u++) {
      if (t0 + u < ntiles) {
        const int r = 64 * (t0 + u) + l;
        const double cv = wv[u] ? cell_value_fast(wv[u], (int64_t)col * n2 + r, qjump, bk) : 0.0;
        uint64_t mask = __ballot(wv[u] != 0);
        double mine = v;
        while (mask) {
          const int bb = __ffsll((unsigned long long)mask) - 1;
          mask &= mask - 1;
          v = readlane_f64(cv, bb) + v;
          if (l >= bb) mine = v;
        }
        tile[b][w][1 + 64 * u + l] = mine;
        if (st == 0 && u == 0 && l < 2) edge[w][l] = mine;
      }
    }
    __syncthreads();
    const int nrows = 32 * min(TPS, ntiles - t0);
    if (tid < 128 * kAvgK && jr < nrows && (st > 0 || jr > 0)) {
      const double(*T)[64 * TPS + 1] = tile[b];
      const int c0 = 2 * oc, c1 = c0 + 1, c2 = c0 + 2, r0 = 2 * jr, r1 = r0 + 1, r2 = r0 + 2;
      const double A = T[c2][r2], B = T[c2][r0], C = T[c0][r2], D = T[c0][r0];
      const double E = T[c1][r2], F = T[c1][r0], G = T[c2][r1], H = T[c0][r1];
      const double I = T[c1][r1];
      out[(int64_t)(i0 + oc) * nh + 32 * t0 + jr] =
          0.0625 * (A + B + C + D) + 0.125 * (E + F + G + H) + 0.25 * I;
    }
  }
  if (l == 0) edge[w][2] = v;  // row n2 - 1
  __syncthreads();
  if (tid < kAvgK) {
    const int o = tid, c0 = 2 * o, c1 = c0 + 1, c2 = c0 + 2;
    const double A = edge[c2][1], B = edge[c2][2], C = edge[c0][1], D = edge[c0][2];
    const double E = edge[c1][1], F = edge[c1][2], G = edge[c2][0], H = edge[c0][0];
    const double I = edge[c1][0];
    out[(int64_t)(i0 + o) * nh] = 0.0625 * (A + B + C + D) + 0.125 * (E + F + G + H) + 0.25 * I;
  }
}

__global__ void c2g_subtract(const double *qa, double *qc, int64_t n, const double *qsum,
                             double ndiv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double qavg0 = *qsum / ndiv;
  qc[k] = qa[k] - qavg0;
}

__global__ void fill_zero(double *a, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.0;
}

// Crossing accumulation and y sweep of the fine (or ultra-fine) grid:
// contours.f90:206-277 (ultra=0) / congen.f90:353-423 (ultra=1).  Leaves the
// swept field in fq (column-major, nyf rows).
bool colscan_avg9_fits(int nxf, int nyf) {
  return nxf == nyf && nyf % 64 == 0 && (nxf / 2) % kAvgK == 0;
}

// With avg non-null (and colscan_avg9_fits) the sweep writes the first
// nine-point average of the fine field to avg instead of the field to fq.
int fine_sweep(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n, int ultra,
               int nxf, int nyf, double glxfi, double glyfi, const double *xg, int32_t *cnt,
               int32_t *qjx, double *qbot, double *fq, double *avg) {
  if (avg && !colscan_avg9_fits(nxf, nyf)) {
    set_error("fine_sweep: grid shape unsuitable for the fused average");
    return HYD_EARG;
  }
  const Consts &k = c.c;
  const int64_t nfine = (int64_t)nxf * nyf;
  if (nfine + nxf > INT32_MAX) {
    set_error("con2grid: fine grid too large for int32 cell keys");
    return HYD_EARG;
  }
  if (nxf % 4) {
    set_error("con2grid: fine grid must have a multiple of 4 columns");
    return HYD_EARG;
  }
  // the context's own counters come back zero from the previous sweep
  HYD_CHECK(sweep_clear_join(c));
  const bool own = cnt == c.cnt && qjx == c.qjx_cnt;
  if (!(own && c.sweep_clean)) {
    HYD_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * nfine, c.s));
    HYD_HIP(hipMemsetAsync(qjx, 0, sizeof(int32_t) * nxf, c.s));
    HYD_HIP(hipMemsetAsync(c.flags, 0, sizeof(int32_t) * 8, c.s));
  }
  c.sweep_clean = false;
  C2GArgs a;
  a.xq = xq;
  a.yq = yq;
  a.nextq = nextq;
  a.n = n;
  a.ultra = ultra;
  a.nxf = nxf;
  a.nyf = nyf;
  a.ellx = k.ellx;
  a.elly = k.elly;
  a.hlx = k.hlx;
  a.hly = k.hly;
  a.hlxi = k.hlxi;
  a.hlyi = k.hlyi;
  a.xmin = k.xmin;
  a.ymin = k.ymin;
  a.glxfi = glxfi;
  a.glyfi = glyfi;
  a.small = k.small;
  a.qjump = c.hsc->qjump;
  a.xgf = xg;
  a.cnt = cnt;
  a.qjx = qjx;
  a.flags = c.flags;
  a.bk.tkey = c.rp_tkey;
  a.bk.tid = c.rp_tid;
  a.bk.list = c.rp_list;
  a.bk.slots = c.replay;
  a.bk.vals = c.replay_val;
  a.bk.tmask = (uint32_t)(c.rp_tsize - 1);
  a.bk.bcap = (int32_t)c.rp_bcap;
  a.bk.scap = c.replay_cap;
  const int nb = (n + 255) / 256;
  if (n > 0) {
    hipLaunchKernelGGL(c2g_nodes, dim3(nb), dim3(256), 0, c.s, a);
    HYD_HIP(hipGetLastError());
    const int nbk = std::min(nb, 1024);
    hipLaunchKernelGGL(c2g_buckets<0>, dim3(nbk), dim3(256), 0, c.s, a);
    hipLaunchKernelGGL(c2g_buckets<1>, dim3(nbk), dim3(256), 0, c.s, a);
  }
  hipLaunchKernelGGL(c2g_qbot, dim3(1), dim3(1024), 0, c.s, qjx, qbot, nxf, nfine, a.qjump, a.bk,
                     c.flags);
  if (avg)
    hipLaunchKernelGGL(c2g_colscan_avg9, dim3(nxf / 2 / kAvgK), dim3(64 * kAvgW), 0, c.s, cnt,
                       qbot, avg, nxf, a.qjump, a.bk);
  else
    hipLaunchKernelGGL(c2g_colscan, dim3((nxf + 3) / 4), dim3(256), 0, c.s, cnt, qbot, fq, nxf,
                       nyf, a.qjump, a.bk);
  c.clr_cnt = cnt;
  c.clr_n4 = nfine / 4;
  c.clear_pending = true;
  if (!c.defer_clear) HYD_CHECK(sweep_clear(c));
  HYD_HIP(hipGetLastError());
  c.sweep_clean = own;
  return HYD_OK;
}

static Buckets ctx_buckets(const Ctx &c) {
  Buckets b;
  b.tkey = c.rp_tkey;
  b.tid = c.rp_tid;
  b.list = c.rp_list;
  b.slots = c.replay;
  b.vals = c.replay_val;
  b.tmask = (uint32_t)(c.rp_tsize - 1);
  b.bcap = (int32_t)c.rp_bcap;
  b.scap = c.replay_cap;
  return b;
}

// The pending clearing of the last sweep, on the current stream c.s (on
// the second stream it leaves an event for the next sweep to wait on).
int sweep_clear(Ctx &c) {
  if (!c.clear_pending) return HYD_OK;
  c.clear_pending = false;
  hipLaunchKernelGGL(c2g_bucket_clear, dim3(1024), dim3(256), 0, c.s, ctx_buckets(c), c.flags,
                     (int4 *)c.clr_cnt, c.clr_n4);
  HYD_HIP(hipGetLastError());
  if (c.s == c.s2) {
    HYD_HIP(hipEventRecord(c.ev_cleared, c.s));
    c.clear_event = true;
  }
  return HYD_OK;
}

// Before a sweep (on c.s): the previous sweep's clearing has run.
int sweep_clear_join(Ctx &c) {
  if (c.clear_pending) return sweep_clear(c);  // nobody ran it: here, in order
  if (c.clear_event) {
    HYD_HIP(hipStreamWaitEvent(c.s, c.ev_cleared, 0));
    c.clear_event = false;
  }
  return HYD_OK;
}

// Reports (and clears) an overflow of the order-dependent-cell buckets of any
// sweep since the last check; the bucket table is emptied so that the next
// sweep starts clean.
int c2g_status(Ctx &c) {
  int32_t e = 0;
  HYD_HIP(hipMemcpyAsync(&e, c.flags + 8, sizeof(int32_t), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  if (e == 0) return HYD_OK;
  HYD_HIP(hipMemsetAsync(c.rp_tkey, 0xff, sizeof(int32_t) * c.rp_tsize, c.s));
  HYD_HIP(hipMemsetAsync(c.flags + 8, 0, sizeof(int32_t), c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  set_error("con2grid: too many order-dependent cells for the bucket table");
  return HYD_ECAPACITY;
}

// defer_mean (n > 0, no counts_out): the averaged grid is left in c.c2g_w
// and its sequential sum in c.dsum; the caller removes the mean (the
// forward FFT does it while loading) and qc is not written.
int con2grid(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n,
             double *qc, int32_t *counts_out, bool defer_mean) {
  const Consts &k = c.c;
  const int64_t N = c.N;
  if (defer_mean && (n == 0 || counts_out)) {
    set_error("con2grid: deferred mean needs nodes and no counts");
    return HYD_EARG;
  }
  if (n == 0) {
    hipLaunchKernelGGL(fill_zero, dim3((N + 255) / 256), dim3(256), 0, c.s, qc, N);
    HYD_HIP(hipGetLastError());
    return HYD_OK;
  }
  if (n > c.cap) {
    set_error("con2grid: node count exceeds capacity");
    return HYD_ECAPACITY;
  }
  const int64_t nfine = (int64_t)k.nxf * k.nyf;
  KTimer tm(c, "con2grid", (int64_t)n * 24 + 8 * N);
  // 4ng -> 2ng (fused with the sweep unless the fine field is wanted) -> ng
  const int nh1 = k.nxf / 2, nh2 = k.nx;
  const bool fused = !counts_out && colscan_avg9_fits(k.nxf, k.nyf);
  HYD_CHECK(fine_sweep(c, xq, yq, nextq, n, 0, k.nxf, k.nyf, k.glxfi, k.glyfi, c.xgf, c.cnt,
                       c.qjx_cnt, c.qbot, c.fq, fused ? c.mq : nullptr));
  if (!fused)
    hipLaunchKernelGGL(c2g_avg9<double>, dim3((nh1 + 255) / 256, nh1), dim3(256), 0, c.s, c.fq,
                       c.mq, nh1);
  hipLaunchKernelGGL(c2g_avg9<double>, dim3((nh2 + 255) / 256, nh2), dim3(256), 0, c.s, c.mq,
                     c.c2g_w, nh2);
  HYD_HIP(hipGetLastError());
  if (c.after_sweep) {
    auto fn = std::move(c.after_sweep);
    c.after_sweep = nullptr;
    HYD_CHECK(fn());
  }
  // sequential mean, contours.f90:330-336
  HYD_CHECK(seqsum(c, c.c2g_w, N, c.dsum));
  if (defer_mean) return HYD_OK;
  hipLaunchKernelGGL(c2g_subtract, dim3((N + 255) / 256), dim3(256), 0, c.s, c.c2g_w, qc, N, c.dsum,
                     (double)N);
  HYD_HIP(hipGetLastError());
  if (counts_out) {
    // integer level counts of the swept fine field (diagnostic path)
    std::vector<double> f(nfine);
    HYD_HIP(hipMemcpyAsync(f.data(), c.fq, sizeof(double) * nfine, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    const double q = c.hsc->qjump;
    for (int64_t t = 0; t < nfine; t++) counts_out[t] = (int32_t)llround(f[t] / q);
  }
  return HYD_OK;
}

// ---- mixed precision ------------------------------------------------------
// The fine field is an integer PV level count times qjump: the bottom row
// and every column are integer prefix sums of the crossing counts (order
// free), converted once; averages in FP32; the mean by a parallel
// reduction.  No order-dependent cell needs special treatment.
__device__ __forceinline__ int level_jump(int32_t w) { return (w & 0xffff) - ((w >> 16) & 0xffff); }

__global__ void c2g_fill_i32(int32_t *a, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0;
}

// Zero n ints (n % 4 == 0, a 16-byte aligned), 16 bytes per store.
__global__ void c2g_fill_i32x4(int4 *a, int64_t n4) {
  const int4 z = make_int4(0, 0, 0, 0);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4;
       k += (int64_t)gridDim.x * blockDim.x)
    a[k] = z;
}

// Zero the bottom-row counters and the per-sweep flags (flags[0..7]).
__global__ void c2g_reset_small(int32_t *qjx, int nxf, int32_t *flags) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nxf; k += gridDim.x * blockDim.x)
    qjx[k] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 8) flags[threadIdx.x] = 0;
}

// qb[ix] = sum_{j < ix} jump(qjx[j]) (one workgroup)
__global__ void __launch_bounds__(1024) c2g_qbot_int(const int32_t *qjx, int32_t *qb, int nxf) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (nxf + 1023) / 1024;
  const int b = min(nxf, per * t), e = min(nxf, b + per);
  int32_t s = 0;
  for (int j = b; j < e; j++) s += level_jump(qjx[j]);
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int32_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t acc = part[t] - s;
  for (int j = b; j < e; j++) {
    qb[j] = acc;
    acc += level_jump(qjx[j]);
  }
}

// y sweep as integer prefix sums, one wavefront per fine column
__global__ void c2g_colscan_int(const int32_t *cnt, const int32_t *qb, float *fq, int nxf,
                                int nyf, double qjump) {
  const int wpb = blockDim.x / 64;
  const int col = blockIdx.x * wpb + threadIdx.x / 64;
  const int l = threadIdx.x & 63;
  if (col >= nxf) return;
  float *out = fq + (int64_t)col * nyf;
  int32_t run = qb[col];
  for (int base = 0; base < nyf; base += 64) {
    const int r = base + l;
    int32_t v = (r < nyf && r >= 1) ? level_jump(cnt[cnt_at(col, r, nyf)]) : 0;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t u = __shfl_up(v, o);
      if (l >= o) v += u;
    }
    if (r < nyf) out[r] = (float)((double)(run + v) * qjump);
    run += __shfl(v, 63);
  }
}


// Integer y sweep fused with the first FP32 nine-point average (4ng ->
// 2ng): the blocking of c2g_colscan_avg9, with each 64-row tile of a column
// summed by a wavefront scan (integer sums are order free) and converted to
// (level * qjump) in FP32 as c2g_colscan_int does.
__global__ void __launch_bounds__(64 * kAvgW)
    c2g_colscan_int_avg9(const int32_t *__restrict__ cnt, const int32_t *__restrict__ qb,
                         float *__restrict__ out, int n2, double qjump) {
  constexpr int TPS = 4;
  __shared__ float tile[2][kAvgW][64 * TPS + 1];
  __shared__ float edge[kAvgW][3];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nh = n2 / 2, ntiles = n2 / 64;
  const int i0 = blockIdx.x * kAvgK;
  int col = 2 * i0 - 1 + w;
  if (col < 0) col += n2;
  const int32_t *cc = cnt + cnt_at(col, 0, n2);  // row r at cc[4 r]
  int32_t run = qb[col];
  const int tid = threadIdx.x, oc = tid >> 7, jr = tid & 127;
  // counts of the next stage load while the current one is scanned
  int32_t wv[TPS], wn[TPS];
#pragma unroll
  for (int u = 0; u < TPS; u++) {
    const int r = 64 * u + l;
    wn[u] = (u < ntiles && r >= 1) ? cc[4 * r] : 0;
  }
  for (int st = 0; st * TPS < ntiles; st++) {
    const int t0 = st * TPS, b = st & 1;
#pragma unroll
    for (int u = 0; u < TPS; u++) {
      wv[u] = level_jump(wn[u]);
      wn[u] = t0 + TPS + u < ntiles ? cc[4 * (64 * (t0 + TPS + u) + l)] : 0;
    }
    if (l == 0) tile[b][w][0] = (float)((double)run * qjump);  // row 64 t0 - 1 (unused for st = 0)
#pragma unroll
    for (int u = 0; u < TPS; u++) {
      if (t0 + u < ntiles) {
        const int32_t v = wave_scan_i32(wv[u]);
        const float f = (float)((double)(run + v) * qjump);
        tile[b][w][1 + 64 * u + l] = f;
        if (st == 0 && u == 0 && l < 2) edge[w][l] = f;
        run += __builtin_amdgcn_readlane(v, 63);
      }
    }
    __syncthreads();
    const int nrows = 32 * min(TPS, ntiles - t0);
    if (tid < 128 * kAvgK && jr < nrows && (st > 0 || jr > 0)) {
      const float(*T)[64 * TPS + 1] = tile[b];
      const int c0 = 2 * oc, c1 = c0 + 1, c2 = c0 + 2, r0 = 2 * jr, r1 = r0 + 1, r2 = r0 + 2;
      const float A = T[c2][r2], B = T[c2][r0], C = T[c0][r2], D = T[c0][r0];
      const float E = T[c1][r2], F = T[c1][r0], G = T[c2][r1], H = T[c0][r1];
      const float I = T[c1][r1];
      out[(int64_t)(i0 + oc) * nh + 32 * t0 + jr] =
          0.0625f * (A + B + C + D) + 0.125f * (E + F + G + H) + 0.25f * I;
    }
  }
  if (l == 0) edge[w][2] = (float)((double)run * qjump);  // row n2 - 1
  __syncthreads();
  if (tid < kAvgK) {
    const int o = tid, c0 = 2 * o, c1 = c0 + 1, c2 = c0 + 2;
    const float A = edge[c2][1], B = edge[c2][2], C = edge[c0][1], D = edge[c0][2];
    const float E = edge[c1][1], F = edge[c1][2], G = edge[c2][0], H = edge[c0][0];
    const float I = edge[c1][0];
    out[(int64_t)(i0 + o) * nh] =
        0.0625f * (A + B + C + D) + 0.125f * (E + F + G + H) + 0.25f * I;
  }
}

__global__ void c2g_sum_partial(const float *a, int64_t n, double *part) {
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[k];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

// part[1024] <- mean of the 1024 partial sums' total over ndiv (one
// workgroup, fixed-order tree: deterministic)
__global__ void __launch_bounds__(1024) c2g_mean(double *part, double ndiv) {
  __shared__ double s[1024];
  const int t = threadIdx.x;
  s[t] = part[t];
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) s[t] += s[t + o];
    __syncthreads();
  }
  if (t == 0) part[1024] = s[0] / ndiv;
}

__global__ void c2g_subtract_mean(const float *qa, float *qc, int64_t n, const double *mean) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) qc[k] = (float)((double)qa[k] - *mean);
}

// defer_mean: the averaged grid is left in c.fc32 and its mean in
// c.part[1024]; qc is not written (the forward FFT removes the mean).
int con2grid_mixed(Ctx &c, float *qc, bool defer_mean) {
  const Consts &k = c.c;
  const int64_t N = c.N;
  const int n = c.nptq;
  const int nxf = k.nxf, nyf = k.nyf;
  const int64_t nfine = (int64_t)nxf * nyf;
  KTimer tm(c, "con2grid32", (int64_t)n * 24 + 4 * N);
  HYD_CHECK(sweep_clear_join(c));  // no earlier clearing may land inside this sweep
  c.sweep_clean = false;  // this sweep clears at its start and leaves counts behind
  if (nfine % 4 == 0)
    hipLaunchKernelGGL(c2g_fill_i32x4, dim3(4096), dim3(256), 0, c.s, (int4 *)c.cnt, nfine / 4);
  else
    hipLaunchKernelGGL(c2g_fill_i32, dim3((unsigned)((nfine + 255) / 256)), dim3(256), 0, c.s,
                       c.cnt, nfine);
  hipLaunchKernelGGL(c2g_reset_small, dim3((nxf + 255) / 256), dim3(256), 0, c.s, c.qjx_cnt, nxf,
                     c.flags);
  if (n > 0) {
    C2GArgs a;
    a.xq = c.xq;
    a.yq = c.yq;
    a.nextq = c.nextq;
    a.n = n;
    a.ultra = 0;
    a.nxf = nxf;
    a.nyf = nyf;
    a.ellx = k.ellx;
    a.elly = k.elly;
    a.hlx = k.hlx;
    a.hly = k.hly;
    a.hlxi = k.hlxi;
    a.hlyi = k.hlyi;
    a.xmin = k.xmin;
    a.ymin = k.ymin;
    a.glxfi = k.glxfi;
    a.glyfi = k.glyfi;
    a.small = k.small;
    a.qjump = 0.0;
    a.xgf = c.xgf;
    a.cnt = c.cnt;
    a.qjx = c.qjx_cnt;
    a.flags = c.flags;
    hipLaunchKernelGGL(c2g_nodes, dim3((n + 255) / 256), dim3(256), 0, c.s, a);
  }
  const double qjump = c.hsc->qjump;  // fixed between recontourings
  hipLaunchKernelGGL(c2g_qbot_int, dim3(1), dim3(1024), 0, c.s, c.qjx_cnt, c.qb32, nxf);
  const int nh1 = nxf / 2, nh2 = k.nx;
  if (colscan_avg9_fits(nxf, nyf)) {
    hipLaunchKernelGGL(c2g_colscan_int_avg9, dim3(nxf / 2 / kAvgK), dim3(64 * kAvgW), 0, c.s,
                       c.cnt, c.qb32, c.mq32, nxf, qjump);
  } else {
    hipLaunchKernelGGL(c2g_colscan_int, dim3((nxf + 3) / 4), dim3(256), 0, c.s, c.cnt, c.qb32,
                       c.fq32, nxf, nyf, qjump);
    hipLaunchKernelGGL(c2g_avg9<float>, dim3((nh1 + 255) / 256, nh1), dim3(256), 0, c.s, c.fq32,
                       c.mq32, nh1);
  }
  hipLaunchKernelGGL(c2g_avg9<float>, dim3((nh2 + 255) / 256, nh2), dim3(256), 0, c.s, c.mq32,
                     c.fc32, nh2);
  hipLaunchKernelGGL(c2g_sum_partial, dim3(1024), dim3(256), 0, c.s, c.fc32, N, c.part);
  hipLaunchKernelGGL(c2g_mean, dim3(1), dim3(1024), 0, c.s, c.part, (double)N);
  if (defer_mean) {
    HYD_HIP(hipGetLastError());
    return HYD_OK;
  }
  hipLaunchKernelGGL(c2g_subtract_mean, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.s,
                     c.fc32, qc, N, c.part + 1024);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
