// congen.hip -- recontouring, congen.f90:24-441, on gfx950.
//
//   1. ultra-fine field U on the 16ng grid (congen.f90:40-106): the contour
//      part by the same crossing / sweep kernels as con2grid (con2ugrid
//      conventions) with the exact sequential mean, plus the bilinear
//      interpolation of the coarse residual, in the reference's order.
//   2. level range from max/min of U (congen.f90:163-173).
//   3. crossings of every level with every x grid line and y grid line
//      (congen.f90:201-254), numbered exactly as the reference numbers them
//      (per level: x-line crossings by (ix, iy), then y-line crossings by
//      (iy, ix)).  One pass counts crossings per (level, line), a scan gives
//      offsets, a second pass writes them in order; both passes run a
//      wavefront per grid line with an ordered ballot loop.
//   4. successor of each crossing: the crossing leaving the box it enters
//      (congen.f90:276-292), found from the box's four edges.  Boxes with two
//      exits (saddles) are resolved by the walk below with the reference's
//      last-in-first-out rule (icrtab/noctab).
//   5. the walk (congen.f90:262-293): one thread per level follows the
//      successor pointers from the lowest free crossing, exactly as the
//      reference, and records each closed contour.
//   6. renode three times per contour (congen.f90:296-310) with the batched
//      renode kernel; contours that vanish are dropped as in the reference.
#include <math.h>
#include <string.h>

#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "hyd_internal.h"

namespace hyd {

int fine_sweep(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n, int ultra,
               int nxf, int nyf, double glxfi, double glyfi, const double *xg, int32_t *cnt,
               int32_t *qjx, double *qbot, double *fq, double *avg);
int renode_batch(Ctx &c, const double *xd, const double *yd, const int64_t *off_in,
                 const int32_t *npd, int ncont, int64_t nin, double *xr, double *yr,
                 int64_t *off_out, int32_t *npr, int64_t *ntot_out, int64_t cap_out);

struct CgGeom {
  int nx, ny, nxu, nyu, mgu;
  double ellx, elly, hlxi, hlyi, oms, glxu, glyu, xmin, ymin;
};

// U = (sweep - qavg) + bilinear(qq)  or  U = bilinear(qq)  (congen.f90:45-95)
__global__ void cg_build_u(double *__restrict__ u, const double *__restrict__ sweep,
                           const double *__restrict__ qsum, double ndiv,
                           const double *__restrict__ qq, const double *__restrict__ w,
                           CgGeom g, int with_contours) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ix = blockIdx.y + 1;
  if (iy > g.nyu) return;
  const int mgu = g.mgu;
  const int ixx = (ix - 1) / mgu, iyy = (iy - 1) / mgu;
  const int ix0 = 1 + ixx, iy0 = 1 + iyy;
  const int ix1 = 2 + ixx - g.nx * (ix0 / g.nx), iy1 = 2 + iyy - g.ny * (iy0 / g.ny);
  const int ixf = ix - mgu * ixx, iyf = iy - mgu * iyy;
  const int wk = (ixf - 1) * mgu + (iyf - 1);
  const double w00 = w[wk], w10 = w[256 + wk], w01 = w[512 + wk], w11 = w[768 + wk];
#define Q(a, b) qq[(int64_t)((b)-1) * g.ny + (a)-1]
  const double p00 = w00 * Q(iy0, ix0), p10 = w10 * Q(iy1, ix0);
  const double p01 = w01 * Q(iy0, ix1), p11 = w11 * Q(iy1, ix1);
#undef Q
  const int64_t k = (int64_t)(ix - 1) * g.nyu + (iy - 1);
  double v;
  if (with_contours) {
    const double qavg = *qsum / ndiv;
    const double a = sweep[k] - qavg;
    v = a + p00 + p10 + p01 + p11;
  } else {
    v = p00 + p10 + p01 + p11;
  }
  u[k] = v;
}

// column-major (nyu x nxu) -> transposed (nxu x nyu)
__global__ void cg_transpose(const double *__restrict__ a, double *__restrict__ b, int nr, int nc) {
  __shared__ double t[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + threadIdx.x;
    if (c < nc && r < nr) t[j][threadIdx.x] = a[(int64_t)c * nr + r];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + threadIdx.x;
    if (c < nc && r < nr) b[(int64_t)r * nc + c] = t[threadIdx.x][j];
  }
}

__global__ void cg_minmax(const double *__restrict__ u, int64_t n, double *out) {
  __shared__ double smx[256], smn[256];
  double mx = -1.0 / 0.0, mn = 1.0 / 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double v = u[k];
    mx = v > mx ? v : mx;
    mn = v < mn ? v : mn;
  }
  smx[threadIdx.x] = mx;
  smn[threadIdx.x] = mn;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + s]);
      smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = smx[0];
    out[2 * blockIdx.x + 1] = smn[0];
  }
}

// Level value of lev (congen.f90:186) and the sign test of congen.f90:206:
// isy = sign(1, qa - qtmp) is +1 exactly when qa >= qtmp (qtmp != 0).
struct Levels {
  int levbeg, nlev;
  double qjump, qoff;
};
__device__ __forceinline__ double qlev(const Levels &L, int lev) {
  return ((double)lev - 0.5) * L.qjump - L.qoff;
}
// levels whose value q satisfies (a >= q) != (b >= q): q in (min, max]
__device__ __forceinline__ void level_range(const Levels &L, double a, double b, int &l1, int &l2) {
  const double lo = a < b ? a : b, hi = a < b ? b : a;
  if (!(lo < hi)) {
    l1 = 1;
    l2 = 0;
    return;
  }
  // first lev with q > lo, last lev with q <= hi (monotone in lev); start
  // from a float estimate and correct exactly with qlev
  int s = (int)floor((lo + L.qoff) / L.qjump + 0.5) + 1;
  s = max(s, L.levbeg);
  while (s > L.levbeg && qlev(L, s - 1) > lo) s--;
  while (s < L.levbeg + L.nlev && !(qlev(L, s) > lo)) s++;
  int e = (int)floor((hi + L.qoff) / L.qjump + 0.5);
  e = min(e, L.levbeg + L.nlev - 1);
  while (e < L.levbeg + L.nlev - 1 && qlev(L, e + 1) <= hi) e++;
  while (e >= L.levbeg && !(qlev(L, e) <= hi)) e--;
  l1 = s;
  l2 = e;
}

// Pass over lines: a wavefront per grid line (x-line = column of U, or
// y-line = column of U^T).  mode 0: count crossings per (level, line);
// mode 1: write them in order.
struct LineArgs {
  const double *f;   // column-major, nrow rows, nline lines
  int nrow, nline;
  Levels L;
  double gl;         // grid spacing along the line (glyu for x-lines, glxu for y-lines)
  int32_t *cnt;      // [nlev * nline] counts (mode 0) / running offsets (mode 1)
  // mode 1 outputs (global crossing index)
  int32_t *c_row;    // row index (1-based) of the crossing within its line
  int32_t *c_line;   // line index (1-based)
  double *c_pos;     // gl*qd(r)/(qd(r+1)-qd(r)) (congen.f90:221, 249)
  int8_t *c_sgn;     // isy / isx of the lower point (+1 / -1)
};

__global__ void cg_lines(LineArgs a, int mode) {
  const int wpb = blockDim.x / 64;
  const int line = blockIdx.x * wpb + threadIdx.x / 64;
  const int l = threadIdx.x & 63;
  if (line >= a.nline) return;
  const double *col = a.f + (int64_t)line * a.nrow;
  const double first = col[0];
  for (int base = 0; base < a.nrow; base += 64) {
    const int r = base + l;  // 0-based row; edge (r, r+1 wrapping)
    int l1 = 1, l2 = 0;
    double v0 = 0.0, v1 = 0.0;
    if (r < a.nrow) {
      v0 = col[r];
      v1 = (r + 1 < a.nrow) ? col[r + 1] : first;
      level_range(a.L, v0, v1, l1, l2);
    }
    uint64_t mask = __ballot(l1 <= l2);
    while (mask) {
      const int b = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      const int bl1 = __shfl(l1, b), bl2 = __shfl(l2, b);
      const double b0 = __shfl(v0, b), b1 = __shfl(v1, b);
      // lanes split the levels of this edge; each level's slot is private
      // to this line, so the order within the line is the row order
      for (int lev = bl1 + l; lev <= bl2; lev += 64) {
        int32_t *slot = &a.cnt[(int64_t)(lev - a.L.levbeg) * a.nline + line];
        if (mode == 0) {
          *slot += 1;
        } else {
          const int32_t g = (*slot)++;
          const double q = qlev(a.L, lev);
          const double d0 = b0 - q, d1 = b1 - q;
          a.c_row[g] = base + b + 1;
          a.c_line[g] = line + 1;
          a.c_pos[g] = a.gl * d0 / (d1 - d0);
          a.c_sgn[g] = d0 >= 0.0 ? 1 : -1;
        }
      }
    }
  }
}

// Crossing geometry and box entered/left (congen.f90:213-222, 241-250).
struct XingArgs {
  int64_t nx_tot;        // x-line crossings (all levels) come first in the arrays
  const int32_t *row, *line;
  const double *pos;
  const int8_t *sgn;
  const int32_t *lev_of;  // level (0-based) of each crossing
  const int64_t *lev_base;  // first crossing id of each level
  const int64_t *lev_nx;    // number of x-line crossings of each level
  int64_t ntot;
  CgGeom g;
  const double *xgu, *ygu;
  double *xcr, *ycr;
  int64_t *kib, *kob;
};

__device__ __forceinline__ int ibx_of(int j, int n) { return j == 1 ? n - 1 : j - 2; }

// Per-level crossing numbering: the arrays produced by cg_lines hold x-line
// crossings of all levels first (level-major), then y-line crossings; this
// kernel maps them into per-level order (x then y) and computes geometry.
__global__ void cg_xing(XingArgs a, const int64_t *xoff, const int64_t *yoff, int64_t *perm) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntot) return;
  const bool isx = t < a.nx_tot;
  const int lv = a.lev_of[t];
  const int64_t within = isx ? t - xoff[lv] : a.lev_nx[lv] + (t - a.nx_tot - yoff[lv]);
  const int64_t gid = a.lev_base[lv] + within;  // per-level order: x then y
  perm[t] = gid;
  const CgGeom &g = a.g;
  const int r = a.row[t], ln = a.line[t];
  const int s = a.sgn[t];
  const int inc = (1 - s) / 2;
  if (isx) {
    // x-line ix = ln, edge between iy = r and r+1
    const int ix = ln, iy = r;
    const int64_t kaa = (int64_t)(iy - 1) * g.nxu + 1;
    a.kib[gid] = kaa + ibx_of(ix + inc, g.nxu);
    a.kob[gid] = kaa + ibx_of(ix + 1 - inc, g.nxu);
    a.xcr[gid] = a.xgu[ix - 1];
    const double yy = a.ygu[iy - 1] - a.pos[t];
    a.ycr[gid] = g.oms * (yy - g.elly * (double)(int)(yy * g.hlyi));
  } else {
    // y-line iy = ln, edge between ix = r and r+1
    const int iy = ln, ix = r;
    a.kib[gid] = (int64_t)g.nxu * ibx_of(iy + 1 - inc, g.nyu) + ix;
    a.kob[gid] = (int64_t)g.nxu * ibx_of(iy + inc, g.nyu) + ix;
    a.ycr[gid] = a.ygu[iy - 1];
    const double xx = a.xgu[ix - 1] - a.pos[t];
    a.xcr[gid] = g.oms * (xx - g.ellx * (double)(int)(xx * g.hlxi));
  }
}

// Exits of every box: sort-free lookup through a hash of kob.  Each box has
// at most two exits per level; they are stored in crossing order, which is
// the icrtab order of the reference.
struct HashArgs {
  int64_t *keys;     // (level << 40) | kob, empty = -1
  int64_t *v1, *v2;  // exit crossing ids (v1 < v2), -1 when absent
  int64_t cap;       // power of two
};

__device__ __forceinline__ uint64_t hmix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

__global__ void cg_hash_insert(const int64_t *kob, const int32_t *lev_of_gid, int64_t ntot,
                               HashArgs h, int32_t *overflow) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ntot) return;
  const int64_t key = ((int64_t)lev_of_gid[gid] << 40) | kob[gid];
  uint64_t p = hmix((uint64_t)key) & (uint64_t)(h.cap - 1);
  for (int64_t probe = 0; probe < h.cap; probe++) {
    const unsigned long long prev =
        atomicCAS((unsigned long long *)&h.keys[p], (unsigned long long)-1LL,
                  (unsigned long long)key);
    if (prev == (unsigned long long)-1LL || (long long)prev == key) {
      // slot owned by this key: record the exit (two slots, min/max order)
      const unsigned long long old =
          atomicCAS((unsigned long long *)&h.v1[p], (unsigned long long)-1LL,
                    (unsigned long long)gid);
      if (old != (unsigned long long)-1LL) {
        h.v2[p] = gid;  // at most two exits per box and level
      }
      return;
    }
    p = (p + 1) & (uint64_t)(h.cap - 1);
  }
  atomicAdd(overflow, 1);
}

__global__ void cg_successor(const int64_t *kib, const int32_t *lev_of_gid, int64_t ntot,
                             HashArgs h, int64_t *succ, int64_t *sad2, int32_t *err) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ntot) return;
  const int64_t key = ((int64_t)lev_of_gid[gid] << 40) | kib[gid];
  uint64_t p = hmix((uint64_t)key) & (uint64_t)(h.cap - 1);
  for (int64_t probe = 0; probe < h.cap; probe++) {
    const int64_t k = h.keys[p];
    if (k == key) {
      int64_t a = h.v1[p], b = h.v2[p];
      if (b >= 0 && b < a) {
        const int64_t t = a;
        a = b;
        b = t;
      }
      if (b < 0) {
        succ[gid] = a;   // one exit
        sad2[gid] = -1;
      } else {
        succ[gid] = a;   // icrtab(k,1)
        sad2[gid] = b;   // icrtab(k,2): taken first (last in, first out)
      }
      return;
    }
    if (k == -1) break;
    p = (p + 1) & (uint64_t)(h.cap - 1);
  }
  succ[gid] = -1;
  atomicAdd(err, 1);
}

// congen.f90:258-293, the walk.  A box with one exit per level fixes the
// successor of the crossing that enters it, so the crossings form
// deterministic chains that end where a crossing enters a box with two
// exits (a saddle, whose exit depends on the order of visits: icrtab is
// last-in-first-out and noctab counts down, not decremented when the walk
// closes, congen.f90:283).  Pointer jumping finds every chain, its length and
// each crossing's position; a thread per level then walks the chains in the
// reference's order -- start at the lowest unused crossing, which is the
// lowest crossing of the first unused chain -- taking only the saddle
// decisions sequentially; a last pass writes every crossing to its place in
// the walk.  Chains without a saddle are closed loops on their own.

// chain end (the saddle entry) and distance to it
__global__ void cw_init(const int64_t *succ, const int64_t *sad2, int64_t n, int32_t *P,
                        int32_t *E, int32_t *D) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (sad2[c] >= 0) {
    P[c] = -1;
    E[c] = (int32_t)c;
    D[c] = 0;
  } else {
    P[c] = (int32_t)succ[c];
    E[c] = -1;
    D[c] = 1;
  }
}
__global__ void cw_jump(const int32_t *P, const int32_t *E, const int32_t *D, int64_t n,
                        int32_t *P2, int32_t *E2, int32_t *D2) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int32_t q = P[c];
  if (q < 0) {
    P2[c] = -1;
    E2[c] = E[c];
    D2[c] = D[c];
  } else {
    P2[c] = P[q];
    E2[c] = E[q];
    D2[c] = D[c] + D[q];
  }
}
// closed loops without a saddle: minimum crossing by pointer jumping
__global__ void cw_cyc_init(const int64_t *succ, const int32_t *E, int64_t n, int32_t *M,
                            int32_t *Q) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  M[c] = (int32_t)c;
  Q[c] = E[c] < 0 ? (int32_t)succ[c] : (int32_t)c;
}
__global__ void cw_cyc_min(const int32_t *M, const int32_t *Q, int64_t n, int32_t *M2,
                           int32_t *Q2) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int32_t q = Q[c];
  M2[c] = min(M[c], M[q]);
  Q2[c] = Q[q];
}
// distance to the loop's last crossing (the one before its minimum)
__global__ void cw_cyc_rank_init(const int64_t *succ, const int32_t *E, const int32_t *M,
                                 int64_t n, int32_t *R, int32_t *Q) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (E[c] >= 0) {
    R[c] = 0;
    Q[c] = -1;
    return;
  }
  const int32_t s = (int32_t)succ[c];
  const bool last = s == M[c];
  R[c] = last ? 0 : 1;
  Q[c] = last ? -1 : s;
}
__global__ void cw_rank_jump(const int32_t *R, const int32_t *Q, int64_t n, int32_t *R2,
                             int32_t *Q2) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int32_t q = Q[c];
  if (q < 0) {
    R2[c] = R[c];
    Q2[c] = -1;
  } else {
    R2[c] = R[c] + R[q];
    Q2[c] = Q[q];
  }
}
// chain key (its end crossing, or a loop's minimum) and per-key tables
struct ChainTab {
  int32_t *key, *pos;       // per crossing
  int32_t *hd, *len, *cmin;  // per key: head distance, length, lowest crossing
  int8_t *loop, *used;       // per key
  int64_t *aoff, *boff;      // per key: output offsets of the (up to two) pieces
  int32_t *afrom;            // per key: first position of piece a
};
__global__ void cw_keys(const int32_t *E, const int32_t *D, const int32_t *M, const int32_t *R,
                        int64_t n, ChainTab t) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const bool lp = E[c] < 0;
  const int32_t k = lp ? M[c] : E[c];
  t.key[c] = k;
  if (lp) {
    if (k == c) {
      t.len[k] = R[c] + 1;
      t.loop[k] = 1;
    }
  } else {
    atomicMax(&t.hd[k], D[c]);
    if (k == c) t.loop[k] = 0;
  }
  atomicMin(&t.cmin[k], (int32_t)c);
}
__global__ void cw_pos(const int32_t *D, const int32_t *R, int64_t n, ChainTab t, int8_t *start,
                       const int32_t *lev_of_gid, int32_t *levcnt) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int32_t k = t.key[c];
  if (t.loop[k]) {
    t.pos[c] = t.len[k] - 1 - R[c];
  } else {
    t.pos[c] = t.hd[k] - D[c];
    if (k == c) t.len[k] = t.hd[k] + 1;
  }
  const bool st = t.cmin[k] == (int32_t)c;
  start[c] = st ? 1 : 0;
  if (st) atomicAdd(&levcnt[lev_of_gid[c]], 1);
}
__global__ void cw_fill_tab(ChainTab t, int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  t.hd[c] = 0;
  t.len[c] = 0;
  t.cmin[c] = 0x7fffffff;
  t.loop[c] = 0;
  t.used[c] = 0;
}
// starts (chain minima) in crossing order; soff[lv] = first start of level lv
__global__ void cw_walk(const int32_t *starts, const int64_t *soff, const int32_t *scnt,
                        const int64_t *succ, const int64_t *sad2, int32_t *sadcnt,
                        const int64_t *lev_base, const int64_t *lev_n, int nlev, ChainTab t,
                        int64_t *cyc_start, int32_t *cyc_len, int32_t *cyc_count, int32_t *err) {
  const int lv = blockIdx.x * blockDim.x + threadIdx.x;
  if (lv >= nlev) return;
  const int64_t b = lev_base[lv], nn = lev_n[lv];
  int64_t w = b;
  int32_t nc = 0;
  for (int32_t i = 0; i < scnt[lv]; i++) {
    const int32_t s = starts[soff[lv] + i];
    const int32_t h = t.key[s];
    if (t.used[h]) continue;
    t.used[h] = 1;
    const int64_t start = w;
    if (t.loop[h]) {
      t.afrom[h] = 0;
      t.aoff[h] = w;
      w += t.len[h];
    } else {
      const int32_t ps = t.pos[s];
      t.afrom[h] = ps;
      t.aoff[h] = w;
      w += t.len[h] - ps;
      int32_t cur = h;
      for (int64_t guard = 0;; guard++) {
        const int64_t e2 = sad2[cur];
        const int32_t noc = sadcnt[e2];
        const int64_t nxt = noc == 2 ? e2 : succ[cur];
        if (nxt == s) break;
        sadcnt[e2] = noc - 1;
        if (nxt < b || nxt >= b + nn || guard > nn || t.pos[nxt] != 0) {
          atomicAdd(err, 1);
          break;
        }
        const int32_t kn = t.key[nxt];
        if (kn == h) {
          t.boff[h] = w;
          w += ps;
          break;
        }
        if (t.used[kn] || t.loop[kn]) {
          atomicAdd(err, 1);
          break;
        }
        t.used[kn] = 1;
        t.afrom[kn] = 0;
        t.aoff[kn] = w;
        w += t.len[kn];
        cur = kn;
      }
    }
    cyc_start[b + nc] = start;
    cyc_len[b + nc] = (int32_t)(w - start);
    nc++;
  }
  if (w != b + nn) atomicAdd(err, 1);
  cyc_count[lv] = nc;
}
__global__ void cw_expand(ChainTab t, int64_t n, const int32_t *err, int64_t *order) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || *err) return;  // a failed walk leaves offsets unset
  const int32_t k = t.key[c], p = t.pos[c];
  const int64_t o = p >= t.afrom[k] ? t.aoff[k] + (p - t.afrom[k]) : t.boff[k] + p;
  if (o >= 0 && o < n) order[o] = c;
}

__global__ void cg_gather(const int64_t *order, const double *xcr, const double *ycr, int64_t n,
                          double *xd, double *yd) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = order[t];
  xd[t] = xcr[g];
  yd[t] = ycr[g];
}

__global__ void cg_fill_i32(int32_t *a, int64_t n, int32_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}
__global__ void cg_fill_i64(int64_t *a, int64_t n, int64_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}
__global__ void cg_nextq(int32_t *nextq, const int64_t *off, const int32_t *np, int ncont) {
  const int c = blockIdx.x;
  if (c >= ncont) return;
  const int64_t o = off[c];
  const int n = np[c];
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    nextq[o + i] = (int32_t)(i + 1 < n ? o + i + 2 : o + 1);
}

template <class T>
static int grow(T **p, int64_t &cap, int64_t need) {
  if (need <= cap && *p) return HYD_OK;
  if (*p) HYD_HIP(hipFree(*p));
  *p = nullptr;
  int64_t c = std::max<int64_t>(need, 1024);
  HYD_HIP(hipMalloc((void **)p, sizeof(T) * c));
  cap = c;
  return HYD_OK;
}

// congen.f90:24-128 on the device state.  qq is the residual (or the full
// field when there are no contours), a device (ny,nx) array.
int recontour(Ctx &c, const double *qq) {
  const Consts &k = c.c;
  const int nxu = k.nxu, nyu = k.nyu;
  const int64_t NU = (int64_t)nxu * nyu;
  const double qjump = c.hsc->qjump;
  CgGeom g;
  g.nx = k.nx;
  g.ny = k.ny;
  g.nxu = nxu;
  g.nyu = nyu;
  g.mgu = 16;
  g.ellx = k.ellx;
  g.elly = k.elly;
  g.hlxi = k.hlxi;
  g.hlyi = k.hlyi;
  g.oms = k.oms;
  g.glxu = k.glxu;
  g.glyu = k.glyu;
  g.xmin = k.xmin;
  g.ymin = k.ymin;
  KTimer tm(c, "recontour", 8 * NU * 4);
  double *U, *UT, *xgu, *ygu, *wts;
  HYD_CHECK(dbuf(c, "U", NU, &U));
  HYD_CHECK(dbuf(c, "UT", NU, &UT));
  HYD_CHECK(dbuf(c, "xgu", nxu, &xgu));
  HYD_CHECK(dbuf(c, "ygu", nyu, &ygu));
  HYD_CHECK(dbuf(c, "w", 1024, &wts));
  {
    std::vector<double> hx(nxu), hy(nyu), hw(1024);
    for (int i = 1; i <= nxu; i++) hx[i - 1] = k.xmin + k.glxu * (double)(i - 1);
    for (int i = 1; i <= nyu; i++) hy[i - 1] = k.ymin + k.glyu * (double)(i - 1);
    // contours.f90:100-115
    const double fac = 1.0 / 16.0;
    for (int ixf = 1; ixf <= 16; ixf++) {
      const double pxf = fac * (double)(ixf - 1), pxc = 1.0 - pxf;
      for (int iyf = 1; iyf <= 16; iyf++) {
        const double pyf = fac * (double)(iyf - 1), pyc = 1.0 - pyf;
        const int w = (ixf - 1) * 16 + (iyf - 1);
        hw[w] = pyc * pxc;
        hw[256 + w] = pyf * pxc;
        hw[512 + w] = pyc * pxf;
        hw[768 + w] = pyf * pxf;
      }
    }
    HYD_HIP(hipMemcpyAsync(xgu, hx.data(), sizeof(double) * nxu, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(ygu, hy.data(), sizeof(double) * nyu, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(wts, hw.data(), sizeof(double) * 1024, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
  }
  // ---- 1. ultra-fine field
  const bool withc = c.nq > 0;
  if (withc) {
    int32_t *ucnt, *uqjx;
    double *uqbot, *usweep;
    HYD_CHECK(dbuf(c, "ucnt", NU, &ucnt));
    HYD_CHECK(dbuf(c, "uqjx", nxu, &uqjx));
    HYD_CHECK(dbuf(c, "uqbot", nxu, &uqbot));
    HYD_CHECK(dbuf(c, "usweep", NU, &usweep));
    HYD_CHECK(fine_sweep(c, c.xq, c.yq, c.nextq, c.nptq, 1, nxu, nyu, k.glxui, k.glyui, xgu, ucnt,
                         uqjx, uqbot, usweep, nullptr));
    // congen.f90:426-432: sequential mean over the ultra-fine grid
    void *saved_w = c.seq_work;
    const int64_t saved_cap = c.seq_cap;
    void *w2;
    HYD_CHECK(dbuf(c, "seqw", seqsum_work_bytes(NU), (char **)&w2));
    c.seq_work = w2;
    c.seq_cap = seqsum_work_chunks(NU);
    const int rc = seqsum(c, usweep, NU, c.dsum + 2);
    c.seq_work = saved_w;
    c.seq_cap = saved_cap;
    HYD_CHECK(rc);
    hipLaunchKernelGGL(cg_build_u, dim3((nyu + 255) / 256, nxu), dim3(256), 0, c.s, U, usweep,
                       c.dsum + 2, (double)NU, qq, wts, g, 1);
  } else {
    // congen.f90:67-69: nothing to contour if the L1 norm vanishes
    std::vector<double> h(c.N);
    HYD_HIP(hipMemcpyAsync(h.data(), qq, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    double s = 0.0;
    for (int64_t t = 0; t < c.N; t++) s = s + fabs(h[t]);
    if (k.garea * s < k.small) {
      c.nq = 0;
      c.nptq = 0;
      return HYD_OK;
    }
    hipLaunchKernelGGL(cg_build_u, dim3((nyu + 255) / 256, nxu), dim3(256), 0, c.s, U, nullptr,
                       nullptr, 1.0, qq, wts, g, 0);
  }
  HYD_HIP(hipGetLastError());
  // ---- 2. level range
  double *mm;
  HYD_CHECK(dbuf(c, "mm", 2 * 1024, &mm));
  hipLaunchKernelGGL(cg_minmax, dim3(1024), dim3(256), 0, c.s, U, NU, mm);
  std::vector<double> hmm(2048);
  HYD_HIP(hipMemcpyAsync(hmm.data(), mm, sizeof(double) * 2048, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  double qamax = hmm[0], qamin = hmm[1];
  for (int b = 1; b < 1024; b++) {
    qamax = hmm[2 * b] > qamax ? hmm[2 * b] : qamax;
    qamin = hmm[2 * b + 1] < qamin ? hmm[2 * b + 1] : qamin;
  }
  const int nlevm = 2000;
  const double qjumpi = 1.0 / qjump;
  const double qoff = qjump * (double)nlevm;
  const int levbeg = (int)((qoff + qamin) * qjumpi + 0.5) + 1;
  const int levend = (int)((qoff + qamax) * qjumpi + 0.5);
  if (levbeg > levend) {
    c.nq = 0;
    c.nptq = 0;
    return HYD_OK;
  }
  const int nlev = levend - levbeg + 1;
  Levels L;
  L.levbeg = levbeg;
  L.nlev = nlev;
  L.qjump = qjump;
  L.qoff = qoff;
  // ---- 3. crossings
  dim3 tg(32, 8);
  hipLaunchKernelGGL(cg_transpose, dim3((nxu + 31) / 32, (nyu + 31) / 32), tg, 0, c.s, U, UT, nyu,
                     nxu);
  int32_t *cntx, *cnty;
  HYD_CHECK(dbuf(c, "cntx", (int64_t)nlev * nxu, &cntx));
  HYD_CHECK(dbuf(c, "cnty", (int64_t)nlev * nyu, &cnty));
  HYD_HIP(hipMemsetAsync(cntx, 0, sizeof(int32_t) * nlev * nxu, c.s));
  HYD_HIP(hipMemsetAsync(cnty, 0, sizeof(int32_t) * nlev * nyu, c.s));
  LineArgs lx;
  memset(&lx, 0, sizeof(lx));
  lx.f = U;
  lx.nrow = nyu;
  lx.nline = nxu;
  lx.L = L;
  lx.cnt = cntx;
  lx.gl = k.glyu;
  LineArgs ly = lx;
  ly.gl = k.glxu;
  ly.f = UT;
  ly.nrow = nxu;
  ly.nline = nyu;
  ly.cnt = cnty;
  hipLaunchKernelGGL(cg_lines, dim3((nxu + 3) / 4), dim3(256), 0, c.s, lx, 0);
  hipLaunchKernelGGL(cg_lines, dim3((nyu + 3) / 4), dim3(256), 0, c.s, ly, 0);
  HYD_HIP(hipGetLastError());
  std::vector<int32_t> hcx((size_t)nlev * nxu), hcy((size_t)nlev * nyu);
  HYD_HIP(hipMemcpyAsync(hcx.data(), cntx, sizeof(int32_t) * hcx.size(), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(hcy.data(), cnty, sizeof(int32_t) * hcy.size(), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  // exclusive scans: x-crossings of all levels first (level-major, line-major)
  std::vector<int64_t> xoff(nlev), yoff(nlev), lev_nx(nlev), lev_n(nlev), lev_base(nlev);
  int64_t tot = 0;
  for (int lv = 0; lv < nlev; lv++) {
    xoff[lv] = tot;
    for (int j = 0; j < nxu; j++) {
      const int32_t v = hcx[(size_t)lv * nxu + j];
      hcx[(size_t)lv * nxu + j] = (int32_t)tot;
      tot += v;
    }
    lev_nx[lv] = tot - xoff[lv];
  }
  const int64_t nx_tot = tot;
  for (int lv = 0; lv < nlev; lv++) {
    yoff[lv] = tot - nx_tot;
    for (int j = 0; j < nyu; j++) {
      const int32_t v = hcy[(size_t)lv * nyu + j];
      hcy[(size_t)lv * nyu + j] = (int32_t)tot;
      tot += v;
    }
    lev_n[lv] = lev_nx[lv] + (tot - nx_tot - yoff[lv]);
  }
  const int64_t ntot = tot;
  if (ntot >= ((int64_t)1 << 31)) {
    set_error("recontour: more than 2^31 crossings");
    return HYD_ECAPACITY;
  }
  int64_t acc = 0;
  for (int lv = 0; lv < nlev; lv++) {
    lev_base[lv] = acc;
    acc += lev_n[lv];
  }
  HYD_HIP(hipMemcpyAsync(cntx, hcx.data(), sizeof(int32_t) * hcx.size(), hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(cnty, hcy.data(), sizeof(int32_t) * hcy.size(), hipMemcpyHostToDevice, c.s));
  int32_t *c_row, *c_line, *lev_of, *lev_of_gid;
  double *c_pos, *xcr, *ycr;
  int8_t *c_sgn;
  int64_t *kib, *kob, *perm;
  const int64_t nalloc = std::max<int64_t>(ntot, 1);
  HYD_CHECK(dbuf(c, "c_row", nalloc, &c_row));
  HYD_CHECK(dbuf(c, "c_line", nalloc, &c_line));
  HYD_CHECK(dbuf(c, "c_pos", nalloc, &c_pos));
  HYD_CHECK(dbuf(c, "c_sgn", nalloc, &c_sgn));
  HYD_CHECK(dbuf(c, "lev_of", nalloc, &lev_of));
  HYD_CHECK(dbuf(c, "lev_of_gid", nalloc, &lev_of_gid));
  HYD_CHECK(dbuf(c, "xcr", nalloc, &xcr));
  HYD_CHECK(dbuf(c, "ycr", nalloc, &ycr));
  HYD_CHECK(dbuf(c, "kib", nalloc, &kib));
  HYD_CHECK(dbuf(c, "kob", nalloc, &kob));
  HYD_CHECK(dbuf(c, "perm", nalloc, &perm));
  lx.c_row = ly.c_row = c_row;
  lx.c_line = ly.c_line = c_line;
  lx.c_pos = ly.c_pos = c_pos;
  lx.c_sgn = ly.c_sgn = c_sgn;
  hipLaunchKernelGGL(cg_lines, dim3((nxu + 3) / 4), dim3(256), 0, c.s, lx, 1);
  hipLaunchKernelGGL(cg_lines, dim3((nyu + 3) / 4), dim3(256), 0, c.s, ly, 1);
  HYD_HIP(hipGetLastError());
  {
    // level of each crossing in the line-pass order (x block then y block)
    std::vector<int32_t> hl(ntot);
    for (int lv = 0; lv < nlev; lv++) {
      for (int64_t t = xoff[lv]; t < xoff[lv] + lev_nx[lv]; t++) hl[t] = lv;
      const int64_t ny_l = lev_n[lv] - lev_nx[lv];
      for (int64_t t = nx_tot + yoff[lv]; t < nx_tot + yoff[lv] + ny_l; t++) hl[t] = lv;
    }
    std::vector<int32_t> hg(ntot);
    for (int lv = 0; lv < nlev; lv++)
      for (int64_t t = lev_base[lv]; t < lev_base[lv] + lev_n[lv]; t++) hg[t] = lv;
    HYD_HIP(hipMemcpyAsync(lev_of, hl.data(), sizeof(int32_t) * ntot, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(lev_of_gid, hg.data(), sizeof(int32_t) * ntot, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
  }
  int64_t *d_xoff, *d_yoff, *d_lev_nx, *d_lev_base, *d_lev_n;
  HYD_CHECK(dbuf(c, "xoff", nlev, &d_xoff));
  HYD_CHECK(dbuf(c, "yoff", nlev, &d_yoff));
  HYD_CHECK(dbuf(c, "lev_nx", nlev, &d_lev_nx));
  HYD_CHECK(dbuf(c, "lev_base", nlev, &d_lev_base));
  HYD_CHECK(dbuf(c, "lev_n", nlev, &d_lev_n));
  HYD_HIP(hipMemcpyAsync(d_xoff, xoff.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_yoff, yoff.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_lev_nx, lev_nx.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_lev_base, lev_base.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(d_lev_n, lev_n.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  XingArgs xa;
  xa.nx_tot = nx_tot;
  xa.row = c_row;
  xa.line = c_line;
  xa.pos = c_pos;
  xa.sgn = c_sgn;
  xa.lev_of = lev_of;
  xa.lev_base = d_lev_base;
  xa.lev_nx = d_lev_nx;
  xa.ntot = ntot;
  xa.g = g;
  xa.xgu = xgu;
  xa.ygu = ygu;
  xa.xcr = xcr;
  xa.ycr = ycr;
  xa.kib = kib;
  xa.kob = kob;
  const dim3 gb((unsigned)((ntot + 255) / 256));
  if (ntot > 0) hipLaunchKernelGGL(cg_xing, gb, dim3(256), 0, c.s, xa, d_xoff, d_yoff, perm);
  HYD_HIP(hipGetLastError());
  // ---- 4. successors via the box hash
  int64_t hcap = 1024;
  while (hcap < 2 * ntot) hcap <<= 1;
  HashArgs h;
  HYD_CHECK(dbuf(c, "hkeys", hcap, &h.keys));
  HYD_CHECK(dbuf(c, "hv1", hcap, &h.v1));
  HYD_CHECK(dbuf(c, "hv2", hcap, &h.v2));
  h.cap = hcap;
  hipLaunchKernelGGL(cg_fill_i64, dim3((hcap + 255) / 256), dim3(256), 0, c.s, h.keys, hcap, -1);
  hipLaunchKernelGGL(cg_fill_i64, dim3((hcap + 255) / 256), dim3(256), 0, c.s, h.v1, hcap, -1);
  hipLaunchKernelGGL(cg_fill_i64, dim3((hcap + 255) / 256), dim3(256), 0, c.s, h.v2, hcap, -1);
  int64_t *succ, *sad2;
  int32_t *sadcnt, *errs;
  HYD_CHECK(dbuf(c, "succ", nalloc, &succ));
  HYD_CHECK(dbuf(c, "sad2", nalloc, &sad2));
  HYD_CHECK(dbuf(c, "sadcnt", nalloc, &sadcnt));
  HYD_CHECK(dbuf(c, "errs", 4, &errs));
  HYD_HIP(hipMemsetAsync(errs, 0, 16, c.s));
  if (ntot > 0) {
    hipLaunchKernelGGL(cg_fill_i32, gb, dim3(256), 0, c.s, sadcnt, ntot, 2);
    hipLaunchKernelGGL(cg_hash_insert, gb, dim3(256), 0, c.s, kob, lev_of_gid, ntot, h, errs);
    hipLaunchKernelGGL(cg_successor, gb, dim3(256), 0, c.s, kib, lev_of_gid, ntot, h, succ, sad2,
                       errs + 1);
  }
  HYD_HIP(hipGetLastError());
  // ---- 5. walk: chains between saddles by pointer jumping, then the
  // saddle decisions in the reference's order (one thread per level)
  int64_t *order, *cyc_start, *d_cyc_base;
  int32_t *cyc_len, *cyc_count;
  HYD_CHECK(dbuf(c, "order", nalloc, &order));
  HYD_CHECK(dbuf(c, "cyc_start", nalloc, &cyc_start));
  HYD_CHECK(dbuf(c, "cyc_len", nalloc, &cyc_len));
  HYD_CHECK(dbuf(c, "cyc_base", nlev, &d_cyc_base));
  HYD_CHECK(dbuf(c, "cyc_count", nlev, &cyc_count));
  HYD_HIP(hipMemcpyAsync(d_cyc_base, lev_base.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
  {
    int32_t *P1, *E1, *D1, *P2, *E2, *D2, *M1, *Q1, *M2, *Q2, *R1, *R2;
    HYD_CHECK(dbuf(c, "cw_p1", nalloc, &P1));
    HYD_CHECK(dbuf(c, "cw_e1", nalloc, &E1));
    HYD_CHECK(dbuf(c, "cw_d1", nalloc, &D1));
    HYD_CHECK(dbuf(c, "cw_p2", nalloc, &P2));
    HYD_CHECK(dbuf(c, "cw_e2", nalloc, &E2));
    HYD_CHECK(dbuf(c, "cw_d2", nalloc, &D2));
    HYD_CHECK(dbuf(c, "cw_m1", nalloc, &M1));
    HYD_CHECK(dbuf(c, "cw_q1", nalloc, &Q1));
    HYD_CHECK(dbuf(c, "cw_m2", nalloc, &M2));
    HYD_CHECK(dbuf(c, "cw_q2", nalloc, &Q2));
    HYD_CHECK(dbuf(c, "cw_r1", nalloc, &R1));
    HYD_CHECK(dbuf(c, "cw_r2", nalloc, &R2));
    ChainTab tb;
    HYD_CHECK(dbuf(c, "cw_key", nalloc, &tb.key));
    HYD_CHECK(dbuf(c, "cw_pos", nalloc, &tb.pos));
    HYD_CHECK(dbuf(c, "cw_hd", nalloc, &tb.hd));
    HYD_CHECK(dbuf(c, "cw_len", nalloc, &tb.len));
    HYD_CHECK(dbuf(c, "cw_cmin", nalloc, &tb.cmin));
    HYD_CHECK(dbuf(c, "cw_loop", nalloc, &tb.loop));
    HYD_CHECK(dbuf(c, "cw_used", nalloc, &tb.used));
    HYD_CHECK(dbuf(c, "cw_aoff", nalloc, &tb.aoff));
    HYD_CHECK(dbuf(c, "cw_boff", nalloc, &tb.boff));
    HYD_CHECK(dbuf(c, "cw_afrom", nalloc, &tb.afrom));
    int8_t *sflag;
    int32_t *levcnt, *starts, *nsel;
    int64_t *d_soff;
    HYD_CHECK(dbuf(c, "cw_sflag", nalloc, &sflag));
    HYD_CHECK(dbuf(c, "cw_levcnt", nlev, &levcnt));
    HYD_CHECK(dbuf(c, "cw_starts", nalloc, &starts));
    HYD_CHECK(dbuf(c, "cw_nsel", 1, &nsel));
    HYD_CHECK(dbuf(c, "cw_soff", nlev, &d_soff));
    int64_t maxn = 1;
    for (int lv = 0; lv < nlev; lv++) maxn = std::max(maxn, lev_n[lv]);
    int rounds = 1;
    while (((int64_t)1 << rounds) < maxn + 1) rounds++;
    HYD_HIP(hipMemsetAsync(levcnt, 0, 4 * nlev, c.s));
    if (ntot > 0) {
      hipLaunchKernelGGL(cw_init, gb, dim3(256), 0, c.s, succ, sad2, ntot, P1, E1, D1);
      for (int r = 0; r < rounds; r++) {
        hipLaunchKernelGGL(cw_jump, gb, dim3(256), 0, c.s, P1, E1, D1, ntot, P2, E2, D2);
        std::swap(P1, P2);
        std::swap(E1, E2);
        std::swap(D1, D2);
      }
      hipLaunchKernelGGL(cw_cyc_init, gb, dim3(256), 0, c.s, succ, E1, ntot, M1, Q1);
      for (int r = 0; r < rounds; r++) {
        hipLaunchKernelGGL(cw_cyc_min, gb, dim3(256), 0, c.s, M1, Q1, ntot, M2, Q2);
        std::swap(M1, M2);
        std::swap(Q1, Q2);
      }
      hipLaunchKernelGGL(cw_cyc_rank_init, gb, dim3(256), 0, c.s, succ, E1, M1, ntot, R1, Q1);
      for (int r = 0; r < rounds; r++) {
        hipLaunchKernelGGL(cw_rank_jump, gb, dim3(256), 0, c.s, R1, Q1, ntot, R2, Q2);
        std::swap(R1, R2);
        std::swap(Q1, Q2);
      }
      hipLaunchKernelGGL(cw_fill_tab, gb, dim3(256), 0, c.s, tb, ntot);
      hipLaunchKernelGGL(cw_keys, gb, dim3(256), 0, c.s, E1, D1, M1, R1, ntot, tb);
      hipLaunchKernelGGL(cw_pos, gb, dim3(256), 0, c.s, D1, R1, ntot, tb, sflag, lev_of_gid,
                         levcnt);
      HYD_HIP(hipGetLastError());
      // starts (chain minima) in crossing order
      size_t tbytes = 0;
      hipcub::CountingInputIterator<int32_t> ids(0);
      HYD_HIP(hipcub::DeviceSelect::Flagged(nullptr, tbytes, ids, sflag, starts, nsel, (int)ntot,
                                            c.s));
      char *tmp;
      HYD_CHECK(dbuf(c, "cw_seltmp", (int64_t)tbytes, &tmp));
      HYD_HIP(hipcub::DeviceSelect::Flagged(tmp, tbytes, ids, sflag, starts, nsel, (int)ntot,
                                            c.s));
    }
    std::vector<int32_t> hlc(nlev);
    HYD_HIP(hipMemcpyAsync(hlc.data(), levcnt, 4 * nlev, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    std::vector<int64_t> hso(nlev);
    int64_t so = 0;
    for (int lv = 0; lv < nlev; lv++) {
      hso[lv] = so;
      so += hlc[lv];
    }
    HYD_HIP(hipMemcpyAsync(d_soff, hso.data(), 8 * nlev, hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(cw_walk, dim3((nlev + 63) / 64), dim3(64), 0, c.s, starts, d_soff, levcnt,
                       succ, sad2, sadcnt, d_lev_base, d_lev_n, nlev, tb, cyc_start, cyc_len,
                       cyc_count, errs + 2);
    if (ntot > 0) hipLaunchKernelGGL(cw_expand, gb, dim3(256), 0, c.s, tb, ntot, errs + 2, order);
    HYD_HIP(hipGetLastError());
  }
  std::vector<int32_t> hcc(nlev), herr(4);
  HYD_HIP(hipMemcpyAsync(hcc.data(), cyc_count, 4 * nlev, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(herr.data(), errs, 16, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  if (herr[0] || herr[1] || herr[2]) {
    set_error("recontour: inconsistent crossing tables (hash " + std::to_string(herr[0]) +
              ", successor " + std::to_string(herr[1]) + ", walk " + std::to_string(herr[2]) + ")");
    return HYD_ECAPACITY;
  }
  // contour list in (level, walk) order
  std::vector<int64_t> cs_all, cl_all;
  std::vector<int32_t> lev_all;
  {
    int64_t nc_tot = 0;
    for (int lv = 0; lv < nlev; lv++) nc_tot += hcc[lv];
    std::vector<int64_t> cs(ntot ? ntot : 1);
    std::vector<int32_t> cl(ntot ? ntot : 1);
    HYD_HIP(hipMemcpyAsync(cs.data(), cyc_start, 8 * std::max<int64_t>(ntot, 1), hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipMemcpyAsync(cl.data(), cyc_len, 4 * std::max<int64_t>(ntot, 1), hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    cs_all.reserve(nc_tot);
    for (int lv = 0; lv < nlev; lv++)
      for (int j = 0; j < hcc[lv]; j++) {
        cs_all.push_back(cs[lev_base[lv] + j]);
        cl_all.push_back(cl[lev_base[lv] + j]);
        lev_all.push_back(levbeg + lv);
      }
  }
  const int ncont = (int)cs_all.size();
  // ---- 6. renode x3 (congen.f90:295-310)
  double *xd, *yd;
  HYD_CHECK(dbuf(c, "cg_xd", nalloc, &xd));
  HYD_CHECK(dbuf(c, "cg_yd", nalloc, &yd));
  if (ntot > 0) hipLaunchKernelGGL(cg_gather, gb, dim3(256), 0, c.s, order, xcr, ycr, ntot, xd, yd);
  std::vector<int64_t> off(ncont);
  std::vector<int32_t> np(ncont);
  for (int j = 0; j < ncont; j++) {
    off[j] = cs_all[j];
    np[j] = (int32_t)cl_all[j];
  }
  std::vector<int> alive(ncont);
  for (int j = 0; j < ncont; j++) alive[j] = j;
  double *bx[2] = {xd, nullptr}, *by[2] = {yd, nullptr};
  HYD_CHECK(dbuf(c, "cg_xa", nalloc + 64, &bx[1]));
  HYD_CHECK(dbuf(c, "cg_ya", nalloc + 64, &by[1]));
  int64_t *d_off_in, *d_off_out;
  int32_t *d_np, *d_npr;
  HYD_CHECK(dbuf(c, "rn_off_in", std::max(ncont, 1), &d_off_in));
  HYD_CHECK(dbuf(c, "rn_off_out", std::max(ncont, 1), &d_off_out));
  HYD_CHECK(dbuf(c, "rn_np", std::max(ncont, 1), &d_np));
  HYD_CHECK(dbuf(c, "rn_npr", std::max(ncont, 1), &d_npr));
  int cur = 0;
  int64_t nin = ntot;
  for (int round = 0; round < 3; round++) {
    const int na = (int)alive.size();
    if (na == 0) break;
    std::vector<int64_t> oi(na);
    std::vector<int32_t> ni(na);
    for (int j = 0; j < na; j++) {
      oi[j] = off[alive[j]];
      ni[j] = np[alive[j]];
    }
    HYD_HIP(hipMemcpyAsync(d_off_in, oi.data(), 8 * na, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(d_np, ni.data(), 4 * na, hipMemcpyHostToDevice, c.s));
    const int nxt = 1 - cur;
    int64_t ntot_out = 0;
    const int64_t capo = std::min(c.cg_bufs[nxt == 1 ? "cg_xa" : "cg_xd"].cap,
                                  c.cg_bufs[nxt == 1 ? "cg_ya" : "cg_yd"].cap) / 8;
    int rc = renode_batch(c, bx[cur], by[cur], d_off_in, d_np, na, nin, bx[nxt], by[nxt],
                          d_off_out, d_npr, &ntot_out, capo);
    if (rc == HYD_ECAPACITY) {
      // grow the output pair and retry once
      const char *nx_name = nxt == 1 ? "cg_xa" : "cg_xd";
      const char *ny_name = nxt == 1 ? "cg_ya" : "cg_yd";
      // keep the input alive: only the output names are regrown
      HYD_CHECK(dbuf(c, nx_name, ntot_out + 64, &bx[nxt]));
      HYD_CHECK(dbuf(c, ny_name, ntot_out + 64, &by[nxt]));
      rc = renode_batch(c, bx[cur], by[cur], d_off_in, d_np, na, nin, bx[nxt], by[nxt], d_off_out,
                        d_npr, &ntot_out, ntot_out + 64);
    }
    HYD_CHECK(rc);
    std::vector<int64_t> oo(na);
    std::vector<int32_t> npr(na);
    HYD_HIP(hipMemcpyAsync(oo.data(), d_off_out, 8 * na, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipMemcpyAsync(npr.data(), d_npr, 4 * na, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    std::vector<int> next_alive;
    for (int j = 0; j < na; j++) {
      off[alive[j]] = oo[j];
      np[alive[j]] = npr[j];
      if (npr[j] > 0) next_alive.push_back(alive[j]);
    }
    alive.swap(next_alive);
    cur = nxt;
    nin = ntot_out;
  }
  // ---- 7. new contour store (congen.f90:312-322, 112-125)
  const int nq = (int)alive.size();
  std::vector<int64_t> src_off(nq);
  std::vector<int32_t> src_np(nq);
  int64_t npt = 0;
  c.i1q.resize(nq);
  c.i2q.resize(nq);
  c.npq.resize(nq);
  c.indq.resize(nq);
  for (int j = 0; j < nq; j++) {
    const int a = alive[j];
    src_off[j] = off[a];
    src_np[j] = np[a];
    c.i1q[j] = (int32_t)(npt + 1);
    npt += np[a];
    c.i2q[j] = (int32_t)npt;
    c.npq[j] = np[a];
    const int lev = lev_all[a];
    c.indq[j] = lev - nlevm + (lev - 1) / nlevm - 1;
  }
  if (npt > c.cap) {
    set_error("recontour: node count exceeds the context capacity");
    return HYD_ECAPACITY;
  }
  for (int j = 0; j < nq; j++) {
    const int64_t dst = c.i1q[j] - 1;
    HYD_HIP(hipMemcpyAsync(c.xq + dst, bx[cur] + src_off[j], 8 * src_np[j], hipMemcpyDeviceToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(c.yq + dst, by[cur] + src_off[j], 8 * src_np[j], hipMemcpyDeviceToDevice, c.s));
  }
  {
    std::vector<int64_t> o1(nq);
    for (int j = 0; j < nq; j++) o1[j] = c.i1q[j] - 1;
    int64_t *d_o;
    int32_t *d_n;
    HYD_CHECK(dbuf(c, "nq_off", std::max(nq, 1), &d_o));
    HYD_CHECK(dbuf(c, "nq_np", std::max(nq, 1), &d_n));
    HYD_HIP(hipMemcpyAsync(d_o, o1.data(), 8 * nq, hipMemcpyHostToDevice, c.s));
    HYD_HIP(hipMemcpyAsync(d_n, c.npq.data(), 4 * nq, hipMemcpyHostToDevice, c.s));
    if (nq > 0) hipLaunchKernelGGL(cg_nextq, dim3(nq), dim3(256), 0, c.s, c.nextq, d_o, d_n, nq);
    HYD_HIP(hipGetLastError());
    HYD_HIP(hipStreamSynchronize(c.s));
  }
  c.nq = nq;
  c.nptq = (int)npt;
  return HYD_OK;
}

void free_congen(Ctx &c) {
  for (auto &kv : c.cg_bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  c.cg_bufs.clear();
}

}  // namespace hyd
