// surgery.hip -- contour surgery, contours.f90:350-683, on gfx950.
//
// The reference visits the nodes of each PV level in order and, whenever a
// node comes within dm of a segment listed in its box, moves the node,
// rewires two links and continues -- later tests see the modified state, so
// the decisions are order dependent.  The device version keeps that order
// exactly while doing almost all of the work in parallel:
//
//   1. increments dx,dy,dsq per node (contours.f90:397-415);
//   2. boxing per level (428-515): every segment is listed in the boxes its
//      extent spans; a radix sort of (box, node) keys gives the box lists in
//      the reference's order (by node within a box);
//   3. a parallel pre-pass runs every node's box test on the initial state;
//      a node without a hit can only change its outcome if something it reads
//      is modified first;
//   4. one wavefront per level replays the level in node order: lanes test 64
//      consecutive nodes for "needs a look" (pre-pass hit, or its link, its
//      position or its box was touched by an earlier reconnection); those are
//      processed one by one exactly as the reference does, the box's segments
//      tested 64 at a time; every reconnection marks what it touched;
//   5. contours are rebuilt by pointer jumping (cycle minimum = start node,
//      list ranking = position along the contour), renoded with the batched
//      renode kernel, and contours of three or fewer nodes are dropped
//      (contours.f90:619-680).
#include <math.h>

#include <hipcub/hipcub.hpp>

#include "hyd_internal.h"

namespace hyd {

struct SgConst {
  double ellx, elly, hlxi, hlyi, xmin, ymin, oms, d4small, dm2;
};

__device__ __forceinline__ double wrp(double v, double L, double hli) {
  return v - L * (double)(int)(v * hli);
}

// per-level parameters (device arrays)
struct SgLevels {
  const int32_t *ibeg, *iend;   // 1-based node range of each level
  const int32_t *nbx, *nby;
  const double *bwxi, *bwyi;
  const int64_t *boxbase;       // first global box of each level
  const int32_t *lev_of_node;   // level of node (0-based), indexed by node-1
};

// 1. increments
__global__ void sg_incr(const double *x, const double *y, const int32_t *nextq, int n, double *dx,
                        double *dy, double *dsq, SgConst k, int32_t *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // box indices below assume nodes inside the periodic domain
  if (!(x[i] >= k.xmin && x[i] < -k.xmin && y[i] >= k.ymin && y[i] < -k.ymin)) atomicAdd(err, 1);
  const int a = nextq[i] - 1;
  const double d1 = wrp(x[a] - x[i], k.ellx, k.hlxi);
  const double d2 = wrp(y[a] - y[i], k.elly, k.hlyi);
  dx[i] = d1;
  dy[i] = d2;
  dsq[i] = d1 * d1 + d2 * d2;
}

// 2a. box ranges of every segment (slot value ib lists segment nextq(ib))
__global__ void sg_box_range(const double *x, const double *y, const int32_t *nextq,
                             const double *dx, const double *dy, int n, SgLevels L, SgConst k,
                             int32_t *rng, int32_t *cnt, int32_t *err) {
  const int ib0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib0 >= n) return;
  const int lv = L.lev_of_node[ib0];
  const int nbx = L.nbx[lv], nby = L.nby[lv];
  const double bwxi = L.bwxi[lv], bwyi = L.bwyi[lv];
  const double dnbx = (double)nbx, dnby = (double)nby;
  const int i = nextq[ib0] - 1;
  int mbx1, mbx2, mby1, mby2;
  if (dx[i] > 0.0) {
    mbx1 = (int)(dnbx + bwxi * (x[i] - k.xmin));
    mbx2 = (int)(dnbx + bwxi * (x[i] + dx[i] - k.xmin));
  } else {
    mbx1 = (int)(dnbx + bwxi * (x[i] + dx[i] - k.xmin));
    mbx2 = (int)(dnbx + bwxi * (x[i] - k.xmin));
  }
  if (dy[i] > 0.0) {
    mby1 = (int)(dnby + bwyi * (y[i] - k.ymin));
    mby2 = (int)(dnby + bwyi * (y[i] + dy[i] - k.ymin));
  } else {
    mby1 = (int)(dnby + bwyi * (y[i] + dy[i] - k.ymin));
    mby2 = (int)(dnby + bwyi * (y[i] - k.ymin));
  }
  int c = (mbx2 - mbx1 + 1) * (mby2 - mby1 + 1);
  if (mbx2 < mbx1 || mby2 < mby1 || c > 4096) {
    atomicAdd(err, 1);
    c = 0;
  }
  rng[4 * ib0] = mbx1;
  rng[4 * ib0 + 1] = mbx2;
  rng[4 * ib0 + 2] = mby1;
  rng[4 * ib0 + 3] = mby2;
  cnt[ib0] = c;
}

// 2b. (global box, node) keys
__global__ void sg_keys(const int32_t *rng, const int32_t *cnt, const int64_t *off, int n,
                        SgLevels L, uint64_t *keys) {
  const int ib0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib0 >= n || cnt[ib0] == 0) return;
  const int lv = L.lev_of_node[ib0];
  const int nbx = L.nbx[lv], nby = L.nby[lv];
  int64_t o = off[ib0];
  for (int mbx = rng[4 * ib0]; mbx <= rng[4 * ib0 + 1]; mbx++) {
    const int mmbx = mbx % nbx;
    for (int mby = rng[4 * ib0 + 2]; mby <= rng[4 * ib0 + 3]; mby++) {
      const int mb = nby * mmbx + mby % nby + 1;
      const uint64_t gbox = (uint64_t)(L.boxbase[lv] + mb - 1);
      keys[o++] = (gbox << 32) | (uint64_t)(ib0 + 1);
    }
  }
}

// 2c. box boundaries from the sorted keys: kb1 (first slot) per global box
__global__ void sg_box_bounds(const uint64_t *keys, int64_t nk, int64_t *kb1, int64_t *kb2,
                              int32_t *node) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk) return;
  const uint64_t k = keys[t];
  const int64_t box = (int64_t)(k >> 32);
  node[t] = (int32_t)(k & 0xffffffffu);
  if (t == 0 || (int64_t)(keys[t - 1] >> 32) != box) kb1[box] = t;
  if (t == nk - 1 || (int64_t)(keys[t + 1] >> 32) != box) kb2[box] = t;
}

__device__ __forceinline__ int64_t node_box(double xi, double yi, int lv, const SgLevels &L,
                                            const SgConst &k) {
  const int nby = L.nby[lv];
  const int mb = nby * (int)(L.bwxi[lv] * (xi - k.xmin)) + (int)(L.bwyi[lv] * (yi - k.ymin)) + 1;
  return L.boxbase[lv] + mb - 1;
}

// The test of contours.f90:537-563 for node i (successor of ib) against the
// segment listed by slot value isb.  Returns 1 on a reconnection.
__device__ __forceinline__ int sg_test(const double *x, const double *y, const int32_t *nextq,
                                       const double *dx, const double *dy, const double *dsq,
                                       int ib, int i, int isb, const SgConst &k, double &delx,
                                       double &dely, double &aa) {
  const int is = nextq[isb - 1];
  if (is == ib || is == i) return 0;
  delx = wrp(x[is - 1] - x[i - 1], k.ellx, k.hlxi);
  dely = wrp(y[is - 1] - y[i - 1], k.elly, k.hlyi);
  aa = delx * dx[is - 1] + dely * dy[is - 1];
  if (aa * (aa + dsq[is - 1]) + k.d4small < 0.0) {
    const double t = delx * dy[is - 1] - dely * dx[is - 1];
    const double cc = t * t;
    if (cc * (cc - k.dm2 * dsq[is - 1]) < 0.0) return 1;
  }
  return 0;
}

// 3. pre-pass on the initial state
__global__ void sg_prepass(const double *x, const double *y, const int32_t *nextq,
                           const double *dx, const double *dy, const double *dsq, int n,
                           SgLevels L, SgConst k, const int64_t *kb1, const int64_t *kb2,
                           const int32_t *node, int8_t *hit) {
  const int ib0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib0 >= n) return;
  const int lv = L.lev_of_node[ib0];
  const int ib = ib0 + 1;
  const int i = nextq[ib0];
  const int64_t box = node_box(x[i - 1], y[i - 1], lv, L, k);
  int h = 0;
  if (kb1[box] >= 0)
    for (int64_t s = kb1[box]; s <= kb2[box]; s++) {
      double a, b, c;
      if (sg_test(x, y, nextq, dx, dy, dsq, ib, i, node[s], k, a, b, c)) {
        h = 1;
        break;
      }
    }
  hit[ib0] = (int8_t)h;
}

struct SgState {
  double *x, *y, *dx, *dy, *dsq;
  int32_t *nextq, *prev;
  int8_t *hit, *nextdirty, *posdirty, *boxdirty;
  const int64_t *kb1, *kb2;
  const int32_t *node, *rng;
  int32_t *nhit;
};

__device__ void mark_boxes(const SgState &s, int j, int lv, const SgLevels &L) {
  const int nbx = L.nbx[lv], nby = L.nby[lv];
  const int32_t *r = s.rng + 4 * (j - 1);
  for (int mbx = r[0]; mbx <= r[1]; mbx++) {
    const int mmbx = mbx % nbx;
    for (int mby = r[2]; mby <= r[3]; mby++) {
      const int mb = nby * mmbx + mby % nby + 1;
      s.boxdirty[L.boxbase[lv] + mb - 1] = 1;
    }
  }
}

// 4. ordered replay, one wavefront per level.
__global__ void __launch_bounds__(64) sg_replay(SgState s, SgLevels L, SgConst k, int nlev) {
  const int lv = blockIdx.x;
  if (lv >= nlev) return;
  const int l = threadIdx.x;
  const int ibeg = L.ibeg[lv], iend = L.iend[lv];
  for (int base = ibeg; base <= iend; base += 64) {
    int last = -1;  // lanes <= last are done
    for (;;) {
      const int ib = base + l;
      bool want = false;
      if (l > last && ib <= iend) {
        const int i = s.nextq[ib - 1];
        want = s.hit[ib - 1] || s.nextdirty[ib - 1] || s.posdirty[i - 1] ||
               s.boxdirty[node_box(s.x[i - 1], s.y[i - 1], lv, L, k)];
      }
      const uint64_t m = __ballot(want);
      if (m == 0) break;
      const int b = __ffsll((unsigned long long)m) - 1;
      last = b;
      // process node ib_b exactly as contours.f90:529-612
      const int ibb = base + b;
      int i = s.nextq[ibb - 1];
      const int64_t box = node_box(s.x[i - 1], s.y[i - 1], lv, L, k);
      const int64_t k1 = s.kb1[box], k2 = s.kb2[box];
      if (k1 < 0) continue;
      int64_t kk = k1;
      while (kk <= k2) {
        const int64_t slot = kk + l;
        int hl = 0;
        double delx = 0.0, dely = 0.0, aa = 0.0;
        if (slot <= k2)
          hl = sg_test(s.x, s.y, s.nextq, s.dx, s.dy, s.dsq, ibb, i, s.node[slot], k, delx, dely,
                       aa);
        const uint64_t hm = __ballot(hl != 0);
        if (hm == 0) {
          kk += 64;
          continue;
        }
        const int h = __ffsll((unsigned long long)hm) - 1;
        // reconnection (contours.f90:569-608), done by lane h
        int inew = 0;
        if (l == h) {
          int isb = s.node[kk + h];
          int is = s.nextq[isb - 1];
          double dxa, dya;
          if (s.dsq[is - 1] + 2.0 * aa < 0.0) {
            dxa = 0.5 * (delx + s.dx[is - 1]);
            dya = 0.5 * (dely + s.dy[is - 1]);
            isb = is;
            is = s.nextq[is - 1];
          } else {
            dxa = 0.5 * delx;
            dya = 0.5 * dely;
          }
          double xx = s.x[i - 1] + dxa;
          s.x[i - 1] = k.oms * wrp(xx, k.ellx, k.hlxi);
          double yy = s.y[i - 1] + dya;
          s.y[i - 1] = k.oms * wrp(yy, k.elly, k.hlyi);
          s.dx[i - 1] = s.dx[i - 1] - dxa;
          s.dy[i - 1] = s.dy[i - 1] - dya;
          s.dsq[i - 1] = s.dx[i - 1] * s.dx[i - 1] + s.dy[i - 1] * s.dy[i - 1];
          s.nextq[isb - 1] = i;
          s.dx[isb - 1] = s.dx[isb - 1] - dxa;
          s.dy[isb - 1] = s.dy[isb - 1] - dya;
          s.dsq[isb - 1] = s.dx[isb - 1] * s.dx[isb - 1] + s.dy[isb - 1] * s.dy[isb - 1];
          s.x[is - 1] = s.x[i - 1];
          s.y[is - 1] = s.y[i - 1];
          s.dx[is - 1] = s.dx[is - 1] + dxa;
          s.dy[is - 1] = s.dy[is - 1] + dya;
          s.dsq[is - 1] = s.dx[is - 1] * s.dx[is - 1] + s.dy[is - 1] * s.dy[is - 1];
          s.nextq[ibb - 1] = is;
          s.dx[ibb - 1] = s.dx[ibb - 1] + dxa;
          s.dy[ibb - 1] = s.dy[ibb - 1] + dya;
          s.dsq[ibb - 1] = s.dx[ibb - 1] * s.dx[ibb - 1] + s.dy[ibb - 1] * s.dy[ibb - 1];
          // what later tests read and this reconnection touched
          s.nextdirty[isb - 1] = 1;
          s.nextdirty[ibb - 1] = 1;
          s.posdirty[i - 1] = 1;
          s.posdirty[is - 1] = 1;
          s.prev[i - 1] = isb;
          s.prev[is - 1] = ibb;
          mark_boxes(s, isb, lv, L);
          mark_boxes(s, ibb, lv, L);
          mark_boxes(s, s.prev[isb - 1], lv, L);
          mark_boxes(s, s.prev[ibb - 1], lv, L);
          atomicAdd(s.nhit, 1);
          inew = is;
          __threadfence();
        }
        i = __shfl(inew, h);
        __threadfence();
        if (i == ibb) break;
        kk = kk + h + 1;
      }
    }
  }
}

// 5. rebuild: cycle minimum and position along the contour by pointer jumping
__global__ void sg_jump_init(const int32_t *nextq, int n, int32_t *m, int32_t *p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  m[j] = j + 1;
  p[j] = nextq[j];
}
__global__ void sg_jump_min(const int32_t *m, const int32_t *p, int n, int32_t *m2, int32_t *p2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q = p[j] - 1;
  m2[j] = min(m[j], m[q]);
  p2[j] = p[q];
}
// distance to the end of the contour opened at its minimum node
__global__ void sg_rank_init(const int32_t *nextq, const int32_t *head, int n, int32_t *d,
                             int32_t *p) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q = nextq[j];
  const bool end = q == head[j];
  d[j] = end ? 0 : 1;
  p[j] = end ? 0 : q;  // 0 = end of list
}
__global__ void sg_rank_jump(const int32_t *d, const int32_t *p, int n, int32_t *d2, int32_t *p2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int q = p[j];
  if (q == 0) {
    d2[j] = d[j];
    p2[j] = 0;
  } else {
    d2[j] = d[j] + d[q - 1];
    p2[j] = p[q - 1];
  }
}
// scatter node j to its place in the walk-ordered buffer
__global__ void sg_scatter(const double *x, const double *y, const int32_t *head,
                           const int32_t *dist, const int32_t *cid, const int64_t *coff,
                           const int32_t *clen, int n, double *xd, double *yd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int h = head[j] - 1;
  const int c = cid[h];
  const int pos = clen[c] - 1 - dist[j];
  xd[coff[c] + pos] = x[j];
  yd[coff[c] + pos] = y[j];
}

__global__ void sg_fill8(int8_t *a, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0;
}
__global__ void sg_fill64(int64_t *a, int64_t n, int64_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}
__global__ void sg_prev(const int32_t *nextq, int n, int32_t *prev) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  prev[nextq[j] - 1] = j + 1;
}
// the new contour store: contour c copied from src[c] to off[c], closed links
__global__ void sg_store(const double *xr, const double *yr, const int64_t *src,
                         const int64_t *off, const int32_t *np, int ncont, double *xq, double *yq,
                         int32_t *nextq) {
  const int c = blockIdx.x;
  if (c >= ncont) return;
  const int64_t o = off[c], s = src[c];
  const int n = np[c];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    xq[o + i] = xr[s + i];
    yq[o + i] = yr[s + i];
    nextq[o + i] = (int32_t)(i + 1 < n ? o + i + 2 : o + 1);
  }
}

// ---- device-side bookkeeping of the rebuild (integer scans replace the
// host loops over all nodes)
// level of every node: levels are contiguous ascending node ranges
__global__ void sg_level_of_node(const int32_t *ib1, int nlev, int n, int32_t *lon) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int node = j + 1;
  int lo = 0, hi = nlev - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (ib1[mid] <= node) lo = mid;
    else hi = mid - 1;
  }
  lon[j] = lo;
}
__global__ void sg_widen(const int32_t *a, int n, int64_t *b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) b[j] = a[j];
}
__global__ void sg_start_flags(const int32_t *head, int n, int32_t *flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) flag[j] = head[j] == j + 1 ? 1 : 0;
}
// contour c (numbered by start node, which is the reference's order: by
// level, then by start node) starts at node j: its length and level
__global__ void sg_contours(const int32_t *head, const int32_t *dist, const int32_t *incl,
                            const int32_t *lon, const int32_t *levt, int n, int32_t *cid,
                            int32_t *clen, int32_t *clev) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || head[j] != j + 1) return;
  const int c = incl[j] - 1;
  cid[j] = c;
  clen[c] = dist[j] + 1;
  clev[c] = levt[lon[j]];
}
__global__ void sg_renode_flags(const int32_t *clen, const int32_t *ncont, int32_t *flag, int cap) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < cap) flag[c] = (c < *ncont && clen[c] > 3) ? 1 : 0;
}
// the contours renode gets (more than three nodes, contours.f90:645-647)
__global__ void sg_renode_list(const int32_t *flag, const int32_t *incl, const int64_t *coff,
                               const int32_t *clen, const int32_t *clev, int cap, int64_t *roff,
                               int32_t *rlen, int32_t *rlev) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cap || !flag[c]) return;
  const int r = incl[c] - 1;
  roff[r] = coff[c];
  rlen[r] = clen[c];
  rlev[r] = clev[c];
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <class T>
static int scan_incl(Ctx &c, const T *in, T *out, int n, const char *name) {
  size_t tb = 0;
  HYD_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out, n, c.s));
  char *tmp;
  HYD_CHECK(dbuf(c, name, (int64_t)std::max<size_t>(tb, 1), &tmp));
  HYD_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out, n, c.s));
  return HYD_OK;
}

// The node-sized work buffers of surgery, allocated ahead of the loop for
// half as many nodes again as the store holds (bytes per node: element
// size times count per node), so that a surgery inside the time loop
// allocates nothing.
int surgery_reserve(Ctx &c, int64_t nodes) {
  struct Buf {
    const char *name;
    int bytes;  // per node
  };
  static const Buf bufs[] = {
      {"sg_lon", 4},    {"sg_dx", 8},     {"sg_dy", 8},    {"sg_dsq", 8},    {"sg_rng", 16},
      {"sg_cnt", 4},    {"sg_off", 8},    {"sg_cnt64", 8}, {"sg_keys", 24},  {"sg_keys2", 24},
      {"sg_node", 12},  {"sg_hit", 1},    {"sg_nd", 1},    {"sg_pd", 1},     {"sg_prevarr", 4},
      {"sg_m1", 4},     {"sg_p1", 4},     {"sg_m2", 4},    {"sg_p2", 4},     {"sg_d1", 4},
      {"sg_q1", 4},     {"sg_d2", 4},     {"sg_q2", 4},    {"sg_flag", 4},   {"sg_incl", 4},
      {"sg_cid", 4},    {"sg_clen", 4},   {"sg_clev", 4},  {"sg_coff", 8},   {"sg_clen64", 8},
      {"sg_xd", 8},     {"sg_yd", 8},     {"sg_flag2", 4}, {"sg_incl2", 4},  {"sg_roff", 8},
      {"sg_ooff", 8},   {"sg_rlen", 4},   {"sg_rlev", 4},  {"sg_npr", 4},    {"sg_xr", 16},
      {"sg_yr", 16}};
  const int64_t n = nodes + nodes / 2 + 1024;
  for (const Buf &b : bufs) {
    char *p;
    HYD_CHECK(dbuf(c, b.name, n * b.bytes, &p));
  }
  return HYD_OK;
}

int surgery(Ctx &c) {
  if (c.nq == 0) return HYD_OK;
  const Consts &K = c.c;
  const int n = c.nptq, nq = c.nq;
  KTimer tm(c, "surgery", (int64_t)n * 72);
  SgConst k;
  k.ellx = K.ellx;
  k.elly = K.elly;
  k.hlxi = K.hlxi;
  k.hlyi = K.hlyi;
  k.xmin = K.xmin;
  k.ymin = K.ymin;
  k.oms = K.oms;
  k.d4small = K.d4small;
  k.dm2 = K.dm2;
  // ---- levels (contours.f90:374-394) and box counts (428-454)
  std::vector<int32_t> ib1, ib2, nbx, nby, levt;
  std::vector<double> bwx, bwy;
  std::vector<int64_t> bbase;
  {
    int levp = c.indq[0];
    int j1 = 0;
    auto close = [&](int jlast) {
      const int ibeg = c.i1q[j1], iend = c.i2q[jlast];
      ib1.push_back(ibeg);
      ib2.push_back(iend);
      levt.push_back(c.indq[j1]);
      const double fnq = (double)(iend - ibeg + 1);
      const double aspect = K.ellx / K.elly;
      const double fnbx = 1.2 * sqrt(fnq * aspect);
      const double fnby = fnbx / aspect;
      const int bx = std::max(std::min((int)llround(fnbx), K.nx), 1);
      const int by = std::max(std::min((int)llround(fnby), K.ny), 1);
      nbx.push_back(bx);
      nby.push_back(by);
      bwx.push_back(K.oms * (double)bx / K.ellx);
      bwy.push_back(K.oms * (double)by / K.elly);
    };
    for (int j = 1; j < nq; j++) {
      if (c.indq[j] > levp) {
        close(j - 1);
        j1 = j;
        levp = c.indq[j];
      }
    }
    close(nq - 1);
    int64_t acc = 0;
    for (size_t l = 0; l < nbx.size(); l++) {
      bbase.push_back(acc);
      acc += (int64_t)nbx[l] * nby[l];
    }
    bbase.push_back(acc);
  }
  const int nlev = (int)ib1.size();
  const int64_t nbox = bbase[nlev];
  // device copies
  int32_t *d_ib1, *d_ib2, *d_nbx, *d_nby, *d_lon;
  double *d_bwx, *d_bwy;
  int64_t *d_bb;
  HYD_CHECK(dbuf(c, "sg_ib1", nlev, &d_ib1));
  HYD_CHECK(dbuf(c, "sg_ib2", nlev, &d_ib2));
  HYD_CHECK(dbuf(c, "sg_nbx", nlev, &d_nbx));
  HYD_CHECK(dbuf(c, "sg_nby", nlev, &d_nby));
  HYD_CHECK(dbuf(c, "sg_bwx", nlev, &d_bwx));
  HYD_CHECK(dbuf(c, "sg_bwy", nlev, &d_bwy));
  HYD_CHECK(dbuf(c, "sg_bb", nlev + 1, &d_bb));
  HYD_CHECK(dbuf(c, "sg_lon", n, &d_lon));
  HYD_HIP(hipMemcpyAsync(d_ib1, ib1.data(), 4 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_ib2, ib2.data(), 4 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_nbx, nbx.data(), 4 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_nby, nby.data(), 4 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_bwx, bwx.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_bwy, bwy.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_bb, bbase.data(), 8 * (nlev + 1), hipMemcpyHostToDevice, c.s));
  int32_t *d_levt;
  HYD_CHECK(dbuf(c, "sg_levt", nlev, &d_levt));
  HYD_HIP(hipMemcpyAsync(d_levt, levt.data(), 4 * nlev, hipMemcpyHostToDevice, c.s));
  hipLaunchKernelGGL(sg_level_of_node, g1(n), dim3(256), 0, c.s, d_ib1, nlev, n, d_lon);
  SgLevels L;
  L.ibeg = d_ib1;
  L.iend = d_ib2;
  L.nbx = d_nbx;
  L.nby = d_nby;
  L.bwxi = d_bwx;
  L.bwyi = d_bwy;
  L.boxbase = d_bb;
  L.lev_of_node = d_lon;
  // ---- 1. increments (x, y modified in place below)
  double *dx, *dy, *dsq;
  HYD_CHECK(dbuf(c, "sg_dx", n, &dx));
  HYD_CHECK(dbuf(c, "sg_dy", n, &dy));
  HYD_CHECK(dbuf(c, "sg_dsq", n, &dsq));
  int32_t *errs;  // [0] box range, [1] reconnections, [2] nodes outside the domain
  HYD_CHECK(dbuf(c, "sg_err", 4, &errs));
  HYD_HIP(hipMemsetAsync(errs, 0, 16, c.s));
  hipLaunchKernelGGL(sg_incr, g1(n), dim3(256), 0, c.s, c.xq, c.yq, c.nextq, n, dx, dy, dsq, k,
                     errs + 2);
  // ---- 2. boxing
  int32_t *rng, *cnt;
  int64_t *off;
  HYD_CHECK(dbuf(c, "sg_rng", 4 * (int64_t)n, &rng));
  HYD_CHECK(dbuf(c, "sg_cnt", n, &cnt));
  HYD_CHECK(dbuf(c, "sg_off", n + 1, &off));
  hipLaunchKernelGGL(sg_box_range, g1(n), dim3(256), 0, c.s, c.xq, c.yq, c.nextq, dx, dy, n, L, k,
                     rng, cnt, errs);
  HYD_HIP(hipGetLastError());
  // off[0] = 0, off[i + 1] = off[i] + cnt[i]
  int64_t *cnt64;
  HYD_CHECK(dbuf(c, "sg_cnt64", n, &cnt64));
  hipLaunchKernelGGL(sg_widen, g1(n), dim3(256), 0, c.s, cnt, n, cnt64);
  HYD_HIP(hipMemsetAsync(off, 0, 8, c.s));
  HYD_CHECK(scan_incl<int64_t>(c, cnt64, off + 1, n, "sg_scan1"));
  int32_t herr[4];
  int64_t nk = 0;
  HYD_HIP(hipMemcpyAsync(herr, errs, 16, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(&nk, off + n, 8, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  if (herr[2]) {
    set_error("surgery: contour nodes outside the periodic domain");
    return HYD_EARG;
  }
  if (herr[0]) {
    set_error("surgery: a segment spans an invalid box range");
    return HYD_EARG;
  }
  uint64_t *keys, *keys2;
  HYD_CHECK(dbuf(c, "sg_keys", nk, &keys));
  HYD_CHECK(dbuf(c, "sg_keys2", nk, &keys2));
  hipLaunchKernelGGL(sg_keys, g1(n), dim3(256), 0, c.s, rng, cnt, off, n, L, keys);
  {
    int end_bit = 33;
    while (end_bit < 64 && ((int64_t)1 << (end_bit - 32)) < nbox) end_bit++;
    size_t tb = 0;
    HYD_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys, keys2, (int)nk, 0, end_bit, c.s));
    char *tmp;
    HYD_CHECK(dbuf(c, "sg_sorttmp", (int64_t)tb, &tmp));
    HYD_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb, keys, keys2, (int)nk, 0, end_bit, c.s));
  }
  int64_t *kb1, *kb2;
  int32_t *node;
  HYD_CHECK(dbuf(c, "sg_kb1", nbox, &kb1));
  HYD_CHECK(dbuf(c, "sg_kb2", nbox, &kb2));
  HYD_CHECK(dbuf(c, "sg_node", nk, &node));
  hipLaunchKernelGGL(sg_fill64, g1(nbox), dim3(256), 0, c.s, kb1, nbox, (int64_t)-1);
  hipLaunchKernelGGL(sg_fill64, g1(nbox), dim3(256), 0, c.s, kb2, nbox, (int64_t)-2);
  if (nk > 0) hipLaunchKernelGGL(sg_box_bounds, g1(nk), dim3(256), 0, c.s, keys2, nk, kb1, kb2, node);
  HYD_HIP(hipGetLastError());
  // ---- 3. pre-pass
  int8_t *hit, *nd, *pd, *bd;
  int32_t *prev;
  HYD_CHECK(dbuf(c, "sg_hit", n, &hit));
  HYD_CHECK(dbuf(c, "sg_nd", n, &nd));
  HYD_CHECK(dbuf(c, "sg_pd", n, &pd));
  HYD_CHECK(dbuf(c, "sg_bd", nbox, &bd));
  HYD_CHECK(dbuf(c, "sg_prevarr", n, &prev));
  hipLaunchKernelGGL(sg_prepass, g1(n), dim3(256), 0, c.s, c.xq, c.yq, c.nextq, dx, dy, dsq, n, L,
                     k, kb1, kb2, node, hit);
  hipLaunchKernelGGL(sg_fill8, g1(n), dim3(256), 0, c.s, nd, (int64_t)n);
  hipLaunchKernelGGL(sg_fill8, g1(n), dim3(256), 0, c.s, pd, (int64_t)n);
  hipLaunchKernelGGL(sg_fill8, g1(nbox), dim3(256), 0, c.s, bd, nbox);
  hipLaunchKernelGGL(sg_prev, g1(n), dim3(256), 0, c.s, c.nextq, n, prev);
  HYD_HIP(hipGetLastError());
  // ---- 4. ordered replay
  SgState st;
  st.x = c.xq;
  st.y = c.yq;
  st.dx = dx;
  st.dy = dy;
  st.dsq = dsq;
  st.nextq = c.nextq;
  st.prev = prev;
  st.hit = hit;
  st.nextdirty = nd;
  st.posdirty = pd;
  st.boxdirty = bd;
  st.kb1 = kb1;
  st.kb2 = kb2;
  st.node = node;
  st.rng = rng;
  st.nhit = errs + 1;
  hipLaunchKernelGGL(sg_replay, dim3(nlev), dim3(64), 0, c.s, st, L, k, nlev);
  HYD_HIP(hipGetLastError());
  // ---- 5. rebuild: cycle minimum (start node) by pointer jumping
  int32_t *m1, *p1, *m2, *p2;
  HYD_CHECK(dbuf(c, "sg_m1", n, &m1));
  HYD_CHECK(dbuf(c, "sg_p1", n, &p1));
  HYD_CHECK(dbuf(c, "sg_m2", n, &m2));
  HYD_CHECK(dbuf(c, "sg_p2", n, &p2));
  int maxlev = 1;
  for (int l = 0; l < nlev; l++) maxlev = std::max(maxlev, ib2[l] - ib1[l] + 1);
  int rounds = 0;
  while ((1 << rounds) < maxlev) rounds++;
  hipLaunchKernelGGL(sg_jump_init, g1(n), dim3(256), 0, c.s, c.nextq, n, m1, p1);
  for (int r = 0; r < rounds; r++) {
    hipLaunchKernelGGL(sg_jump_min, g1(n), dim3(256), 0, c.s, m1, p1, n, m2, p2);
    std::swap(m1, m2);
    std::swap(p1, p2);
  }
  // m1 = start node of each node's contour; distance to the contour's end
  int32_t *d1, *q1, *d2, *q2;
  HYD_CHECK(dbuf(c, "sg_d1", n, &d1));
  HYD_CHECK(dbuf(c, "sg_q1", n, &q1));
  HYD_CHECK(dbuf(c, "sg_d2", n, &d2));
  HYD_CHECK(dbuf(c, "sg_q2", n, &q2));
  hipLaunchKernelGGL(sg_rank_init, g1(n), dim3(256), 0, c.s, c.nextq, m1, n, d1, q1);
  for (int r = 0; r < rounds; r++) {
    hipLaunchKernelGGL(sg_rank_jump, g1(n), dim3(256), 0, c.s, d1, q1, n, d2, q2);
    std::swap(d1, d2);
    std::swap(q1, q2);
  }
  HYD_HIP(hipGetLastError());
  // contours in the reference's order: by level, then by start node (the
  // node order, levels being ascending node ranges); numbered by a scan of
  // the start flags
  int32_t *flag, *incl, *d_cid, *d_clen, *d_clev, *d_ncont;
  int64_t *d_coff, *clen64;
  double *xd, *yd;
  HYD_CHECK(dbuf(c, "sg_flag", n, &flag));
  HYD_CHECK(dbuf(c, "sg_incl", n, &incl));
  HYD_CHECK(dbuf(c, "sg_cid", n, &d_cid));
  HYD_CHECK(dbuf(c, "sg_clen", n, &d_clen));
  HYD_CHECK(dbuf(c, "sg_clev", n, &d_clev));
  HYD_CHECK(dbuf(c, "sg_coff", n, &d_coff));
  HYD_CHECK(dbuf(c, "sg_clen64", n, &clen64));
  HYD_CHECK(dbuf(c, "sg_xd", n, &xd));
  HYD_CHECK(dbuf(c, "sg_yd", n, &yd));
  hipLaunchKernelGGL(sg_start_flags, g1(n), dim3(256), 0, c.s, m1, n, flag);
  HYD_CHECK(scan_incl<int32_t>(c, flag, incl, n, "sg_scan2"));
  d_ncont = incl + (n - 1);  // contours in all
  hipLaunchKernelGGL(sg_contours, g1(n), dim3(256), 0, c.s, m1, d1, incl, d_lon, d_levt, n, d_cid,
                     d_clen, d_clev);
  // coff[c] = sum of the earlier contours' lengths (unused slots hold 0)
  hipLaunchKernelGGL(sg_fill64, g1(n), dim3(256), 0, c.s, clen64, (int64_t)n, (int64_t)0);
  hipLaunchKernelGGL(sg_widen, g1(n), dim3(256), 0, c.s, d_clen, n, clen64);
  HYD_HIP(hipMemsetAsync(d_coff, 0, 8, c.s));
  if (n > 1) HYD_CHECK(scan_incl<int64_t>(c, clen64, d_coff + 1, n - 1, "sg_scan3"));
  hipLaunchKernelGGL(sg_scatter, g1(n), dim3(256), 0, c.s, c.xq, c.yq, m1, d1, d_cid, d_coff,
                     d_clen, n, xd, yd);
  HYD_HIP(hipGetLastError());
  // renode the contours of more than three nodes (contours.f90:645-647)
  int32_t *flag2, *incl2, *d_rlen, *d_npr, *d_rlev;
  int64_t *d_roff, *d_ooff;
  HYD_CHECK(dbuf(c, "sg_flag2", n, &flag2));
  HYD_CHECK(dbuf(c, "sg_incl2", n, &incl2));
  HYD_CHECK(dbuf(c, "sg_roff", n, &d_roff));
  HYD_CHECK(dbuf(c, "sg_ooff", n, &d_ooff));
  HYD_CHECK(dbuf(c, "sg_rlen", n, &d_rlen));
  HYD_CHECK(dbuf(c, "sg_rlev", n, &d_rlev));
  HYD_CHECK(dbuf(c, "sg_npr", n, &d_npr));
  hipLaunchKernelGGL(sg_renode_flags, g1(n), dim3(256), 0, c.s, d_clen, d_ncont, flag2, n);
  HYD_CHECK(scan_incl<int32_t>(c, flag2, incl2, n, "sg_scan4"));
  hipLaunchKernelGGL(sg_renode_list, g1(n), dim3(256), 0, c.s, flag2, incl2, d_coff, d_clen,
                     d_clev, n, d_roff, d_rlen, d_rlev);
  HYD_HIP(hipGetLastError());
  int32_t nr = 0;
  HYD_HIP(hipMemcpyAsync(&nr, incl2 + (n - 1), 4, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  double *xr, *yr;
  int64_t capo = std::max<int64_t>(2 * (int64_t)n + 1024, 4096);
  HYD_CHECK(dbuf(c, "sg_xr", capo, &xr));
  HYD_CHECK(dbuf(c, "sg_yr", capo, &yr));
  int64_t nout = 0;
  int rc = renode_batch(c, xd, yd, d_roff, d_rlen, nr, n, xr, yr, d_ooff, d_npr, &nout, capo);
  if (rc == HYD_ECAPACITY) {
    HYD_CHECK(dbuf(c, "sg_xr", nout + 64, &xr));
    HYD_CHECK(dbuf(c, "sg_yr", nout + 64, &yr));
    rc = renode_batch(c, xd, yd, d_roff, d_rlen, nr, n, xr, yr, d_ooff, d_npr, &nout, nout + 64);
  }
  HYD_CHECK(rc);
  std::vector<int64_t> ooff(nr);
  std::vector<int32_t> npr(nr), rlev(nr);
  HYD_HIP(hipMemcpyAsync(ooff.data(), d_ooff, 8 * (size_t)nr, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(npr.data(), d_npr, 4 * (size_t)nr, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(rlev.data(), d_rlev, 4 * (size_t)nr, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  // keep contours with more than three nodes after renoding (648-660)
  std::vector<int32_t> i1, i2, np, ind;
  std::vector<int64_t> src;
  int64_t npt = 0;
  for (int j = 0; j < nr; j++) {
    if (npr[j] > 3) {
      i1.push_back((int32_t)(npt + 1));
      npt += npr[j];
      i2.push_back((int32_t)npt);
      np.push_back(npr[j]);
      ind.push_back(rlev[j]);
      src.push_back(ooff[j]);
    }
  }
  if (npt > c.cap) {
    // The reconnections above already rewired xq, yq and nextq in place, so
    // the store no longer matches its contour table: the context drops its
    // state and every step call reports HYD_ESTATE until a new field or
    // contour set is loaded.
    c.have_field = false;
    c.nptq = 0;
    c.nq = 0;
    c.i1q.clear();
    c.i2q.clear();
    c.npq.clear();
    c.indq.clear();
    set_error("surgery: node count exceeds the context capacity; the context's contours are "
              "dropped (set_field or set_contours to continue)");
    return HYD_ECAPACITY;
  }
  const int nqn = (int)np.size();
  if (nqn > 0) {
    std::vector<int64_t> o1(nqn);
    for (int j = 0; j < nqn; j++) o1[j] = i1[j] - 1;
    int64_t *d_o, *d_src;
    int32_t *d_n;
    HYD_CHECK(dbuf(c, "sg_o1", nqn, &d_o));
    HYD_CHECK(dbuf(c, "sg_src", nqn, &d_src));
    HYD_CHECK(dbuf(c, "sg_n1", nqn, &d_n));
    HYD_HIP(hipMemcpyAsync(d_o, o1.data(), 8 * (size_t)nqn, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(d_src, src.data(), 8 * (size_t)nqn, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(d_n, np.data(), 4 * (size_t)nqn, hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(sg_store, dim3(nqn), dim3(256), 0, c.s, xr, yr, d_src, d_o, d_n, nqn, c.xq,
                       c.yq, c.nextq);
    HYD_HIP(hipGetLastError());
  }
  HYD_HIP(hipStreamSynchronize(c.s));
  c.i1q = i1;
  c.i2q = i2;
  c.npq = np;
  c.indq = ind;
  c.nq = nqn;
  c.nptq = (int)npt;
  return HYD_OK;
}

}  // namespace hyd
