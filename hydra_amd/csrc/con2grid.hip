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
//     flagged on the fly and re-summed in node order from a replay list.
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
// Fine-grid layout here: column-major by fine column, rows iy = 1..nyf,
// index (ix-1)*nyf + (iy-1) for both the counters and the swept field.
#include "hyd_internal.h"

namespace hyd {

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
  int32_t *flags;    // [0] flagged cells, [1] flagged qjx, [2] replay entries, [3] overflow
  int32_t *replay;   // triples (key, node, sign)
  int64_t replay_cap;
  int32_t *ixc, *nxc;
  double *ndx, *ndy;
};

__device__ __forceinline__ bool is_flagged(int32_t w) {
  const int p = w & 0xffff, q = (w >> 16) & 0xffff;
  return p > 0 && q > 0 && (p >= 3 || q >= 3);
}

__device__ __forceinline__ void add_count(int32_t *slot, int sign, int32_t *flag) {
  const int32_t d = sign > 0 ? 1 : (1 << 16);
  const int32_t old = atomicAdd(slot, d);
  if (!is_flagged(old) && is_flagged(old + d)) atomicAdd(flag, 1);
}

// Node phase: contours.f90:206-231 and the crossing loop 249-269.
__global__ void c2g_nodes(C2GArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double xi = a.xq[i], yi = a.yq[i];
  const int ia = a.nextq[i] - 1;
  const double xa = a.xq[ia], ya = a.yq[ia];
  // fine cell of each node: contours.f90:207 (1-based) / congen.f90:354 (0-based)
  const int base = a.ultra ? 0 : 1;
  const int ixci = base + (int)(a.glxfi * (xi - a.xmin));
  const int ixca = base + (int)(a.glxfi * (xa - a.xmin));
  const double xx = xa - xi;
  const double dx = xx - a.ellx * (double)(int)(xx * a.hlxi);
  const double yy = ya - yi;
  const double dy = yy - a.elly * (double)(int)(yy * a.hlyi);
  const int ixdif = ixca - ixci;
  const int nxc = ixdif - a.nxf * ((2 * ixdif) / a.nxf);
  a.ixc[i] = ixci;
  a.nxc[i] = nxc;
  a.ndx[i] = dx;
  a.ndy[i] = dy;
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
    for (int ncr = 0; ncr != nxc; ncr += jump) {
      const int ix = 1 + (ixbeg + ncr) % a.nxf;
      const double x2 = a.xgf[ix - 1] - xi;
      const double px0 = (x2 - a.ellx * (double)(int)(x2 * a.hlxi)) / dx;
      const double y2 = yi + px0 * dy;
      const int iy = 2 + (int)(a.glyfi * (y2 - a.elly * (double)(int)(y2 * a.hlyi) - a.ymin));
      // row nyf+1 is never swept (contours.f90:274): its jumps are dropped
      if (iy >= 2 && iy <= a.nyf)
        add_count(&a.cnt[(int64_t)(ix - 1) * a.nyf + (iy - 1)], sgn, &a.flags[0]);
    }
  }
}

// Rare path: list the contributions to flagged cells in node order.
__global__ void c2g_replay_collect(C2GArgs a) {
  if (a.flags[0] == 0 && a.flags[1] == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double xi = a.xq[i], yi = a.yq[i];
  const double dx = a.ndx[i], dy = a.ndy[i];
  const int nxc = a.nxc[i], ixci = a.ixc[i];
  const int ia = a.nextq[i] - 1;
  const double yy = a.yq[ia] - yi;
  const int64_t nkey = (int64_t)a.nxf * a.nyf;
  if (a.flags[1] != 0 && fabs(yy) > a.hly) {
    const double cc = copysign(1.0, dy);
    const double p = -(yi - a.hly * cc) / (dy + a.small);
    const int ix = 1 + (int)(a.glxfi * (fmod(a.ellx + a.hlx + xi + p * dx, a.ellx)));
    if (ix >= 1 && ix <= a.nxf && is_flagged(a.qjx[ix - 1])) {
      const int64_t k = atomicAdd(&a.flags[2], 1);
      if (k < a.replay_cap) {
        a.replay[3 * k] = (int32_t)(nkey + ix - 1);
        a.replay[3 * k + 1] = i;
        a.replay[3 * k + 2] = cc > 0.0 ? -1 : 1;
      } else {
        a.flags[3] = 1;
      }
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
      if (is_flagged(a.cnt[cell])) {
        const int64_t k = atomicAdd(&a.flags[2], 1);
        if (k < a.replay_cap) {
          a.replay[3 * k] = (int32_t)cell;
          a.replay[3 * k + 1] = i;
          a.replay[3 * k + 2] = sgn;
        } else {
          a.flags[3] = 1;
        }
      }
    }
  }
}

// Rare path: sort the replay list by (key, node) and sum each key in node
// order.  Single thread; the list is empty in all but pathological cases.
__global__ void c2g_replay_resolve(int32_t *flags, int32_t *replay, double *vals, int64_t cap,
                                   double qjump) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t k = flags[2];
  if (k == 0) return;
  if (k > cap) k = cap;
  // insertion sort on (key, node)
  for (int64_t a = 1; a < k; a++) {
    const int32_t k0 = replay[3 * a], k1 = replay[3 * a + 1], k2 = replay[3 * a + 2];
    int64_t b = a - 1;
    while (b >= 0 && (replay[3 * b] > k0 || (replay[3 * b] == k0 && replay[3 * b + 1] > k1))) {
      replay[3 * (b + 1)] = replay[3 * b];
      replay[3 * (b + 1) + 1] = replay[3 * b + 1];
      replay[3 * (b + 1) + 2] = replay[3 * b + 2];
      b--;
    }
    replay[3 * (b + 1)] = k0;
    replay[3 * (b + 1) + 1] = k1;
    replay[3 * (b + 1) + 2] = k2;
  }
  // the value of a key is stored at the position of its last entry
  double s = 0.0;
  for (int64_t a = 0; a < k; a++) {
    if (a == 0 || replay[3 * a] != replay[3 * (a - 1)]) s = 0.0;
    s = s + (replay[3 * a + 2] > 0 ? qjump : -qjump);
    vals[a] = s;
  }
  flags[4] = (int32_t)k;
}

// Value of a cell's node-order sum from its counts.
__device__ __forceinline__ double cell_value(int32_t w, int64_t key, double qjump,
                                             const int32_t *replay, const double *vals,
                                             int32_t nrep) {
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
  // flagged: binary search for the last replay entry of this key
  int lo = 0, hi = nrep - 1, last = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int32_t km = replay[3 * mid];
    if (km <= key) {
      if (km == key) last = mid;
      lo = mid + 1;
    } else {
      hi = mid - 1;
    }
  }
  return last >= 0 ? vals[last] : 0.0 / 0.0;
}

// qbot (contours.f90:234-238): one wavefront, chain over the non-empty
// bottom-row counters.
// The counters and their values are staged in LDS by the whole block; wave 0
// runs the chain.  nxf <= kQbotMax.
static constexpr int kQbotMax = 32768;
__global__ void __launch_bounds__(1024) c2g_qbot(const int32_t *qjx, double *qbot, int nxf,
                                                 int64_t nkey, double qjump, const int32_t *flags,
                                                 const int32_t *replay, const double *vals) {
  __shared__ double cvs[kQbotMax / 4];   // values of one quarter at a time
  __shared__ uint8_t nz[kQbotMax / 4];
  const int nrep = flags[4];
  const int q4 = kQbotMax / 4;
  double v = 0.0;  // qbot(1)
  if (threadIdx.x == 0) qbot[0] = 0.0;
  // qbot(ix+1) = qbot(ix) + qjx(ix), ix = 1..nxf-1
  for (int part = 0; part < nxf - 1; part += q4) {
    const int m = min(q4, nxf - 1 - part);
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      const int32_t w = qjx[part + t];
      nz[t] = w != 0;
      cvs[t] = w ? cell_value(w, nkey + part + t, qjump, replay, vals, nrep) : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int l = threadIdx.x;
      for (int base = 0; base < m; base += 64) {
        const int t = base + l;
        const bool ok = t < m;
        const bool w = ok && nz[t];
        const double cv = w ? cvs[t] : 0.0;
        uint64_t mask = __ballot(w);
        double mine = v;
        while (mask) {
          const int b = __ffsll((unsigned long long)mask) - 1;
          mask &= mask - 1;
          v = __shfl(cv, b) + v;
          if (l >= b) mine = v;
        }
        if (ok) qbot[part + t + 1] = mine;
      }
    }
    __syncthreads();
  }
}

// y sweep (contours.f90:272-277): one wavefront per fine column.
__global__ void c2g_colscan(const int32_t *cnt, const double *qbot, double *fq, int nxf, int nyf,
                            double qjump, const int32_t *flags, const int32_t *replay,
                            const double *vals) {
  const int wpb = blockDim.x / 64;
  const int col = blockIdx.x * wpb + threadIdx.x / 64;
  const int l = threadIdx.x & 63;
  if (col >= nxf) return;
  const int nrep = flags[4];
  const int32_t *cc = cnt + (int64_t)col * nyf;
  double *out = fq + (int64_t)col * nyf;
  double v = qbot[col];
  constexpr int U = 8;  // segments of 64 rows loaded ahead
  for (int base = 0; base < nyf; base += 64 * U) {
    int32_t w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = base + 64 * u + l;
      w[u] = (r < nyf && r >= 1) ? cc[r] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = base + 64 * u + l;  // iy-1
      const double cv =
          w[u] ? cell_value(w[u], (int64_t)col * nyf + r, qjump, replay, vals, nrep) : 0.0;
      uint64_t mask = __ballot(w[u] != 0);
      double mine = v;
      while (mask) {
        const int b = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        v = __shfl(cv, b) + v;
        if (l >= b) mine = v;
      }
      if (r < nyf) out[r] = mine;
    }
  }
}

// Nine-point average of an (nh2 x nh2) column-major field f to (nh x nh),
// contours.f90:298-326; out layout column-major with leading dim nh.
__global__ void c2g_avg9(const double *__restrict__ f, double *__restrict__ out, int nh) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int ix = blockIdx.y + 1;
  if (iy > nh) return;
  const int n2 = 2 * nh;
#define F(r, c) f[(int64_t)((c) == 0 ? n2 - 1 : (c)-1) * n2 + ((r) == 0 ? n2 - 1 : (r)-1)]
  const int r1 = 2 * iy, r0 = 2 * iy - 2, c1 = 2 * ix, c0 = 2 * ix - 2;
  const double qev0_a = F(r1, c1), qev0_b = F(r0, c1), qev2_a = F(r1, c0), qev2_b = F(r0, c0);
  const double qev1_a = F(r1, c1 - 1), qev1_b = F(r0, c1 - 1);
  const double qod0 = F(r1 - 1, c1), qod2 = F(r1 - 1, c0), qod1 = F(r1 - 1, c1 - 1);
#undef F
  out[(int64_t)(ix - 1) * nh + (iy - 1)] =
      0.0625 * (qev0_a + qev0_b + qev2_a + qev2_b) +
      0.125 * (qev1_a + qev1_b + qod0 + qod2) + 0.25 * qod1;
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
int fine_sweep(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n, int ultra,
               int nxf, int nyf, double glxfi, double glyfi, const double *xg, int32_t *cnt,
               int32_t *qjx, double *qbot, double *fq) {
  const Consts &k = c.c;
  const int64_t nfine = (int64_t)nxf * nyf;
  HYD_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * nfine, c.s));
  HYD_HIP(hipMemsetAsync(qjx, 0, sizeof(int32_t) * nxf, c.s));
  HYD_HIP(hipMemsetAsync(c.flags, 0, sizeof(int32_t) * 8, c.s));
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
  a.replay = c.replay;
  a.replay_cap = c.replay_cap;
  a.ixc = c.ixc;
  a.nxc = c.nxc;
  a.ndx = c.ndx;
  a.ndy = c.ndy;
  const int nb = (n + 255) / 256;
  if (n > 0) {
    hipLaunchKernelGGL(c2g_nodes, dim3(nb), dim3(256), 0, c.s, a);
    HYD_HIP(hipGetLastError());
    hipLaunchKernelGGL(c2g_replay_collect, dim3(nb), dim3(256), 0, c.s, a);
  }
  hipLaunchKernelGGL(c2g_replay_resolve, dim3(1), dim3(64), 0, c.s, c.flags, c.replay,
                     c.replay_val, c.replay_cap, a.qjump);
  hipLaunchKernelGGL(c2g_qbot, dim3(1), dim3(1024), 0, c.s, qjx, qbot, nxf, nfine, a.qjump,
                     c.flags, c.replay, c.replay_val);
  hipLaunchKernelGGL(c2g_colscan, dim3((nxf + 3) / 4), dim3(256), 0, c.s, cnt, qbot, fq, nxf,
                     nyf, a.qjump, c.flags, c.replay, c.replay_val);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int con2grid(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n,
             double *qc, int32_t *counts_out) {
  const Consts &k = c.c;
  const int64_t N = c.N;
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
  HYD_CHECK(fine_sweep(c, xq, yq, nextq, n, 0, k.nxf, k.nyf, k.glxfi, k.glyfi, c.xgf, c.cnt,
                       c.qjx_cnt, c.qbot, c.fq));
  // 4ng -> 2ng -> ng
  const int nh1 = k.nxf / 2, nh2 = k.nx;
  hipLaunchKernelGGL(c2g_avg9, dim3((nh1 + 255) / 256, nh1), dim3(256), 0, c.s, c.fq, c.mq, nh1);
  hipLaunchKernelGGL(c2g_avg9, dim3((nh2 + 255) / 256, nh2), dim3(256), 0, c.s, c.mq, c.wd, nh2);
  HYD_HIP(hipGetLastError());
  // sequential mean, contours.f90:330-336
  HYD_CHECK(seqsum(c, c.wd, N, c.dsum));
  hipLaunchKernelGGL(c2g_subtract, dim3((N + 255) / 256), dim3(256), 0, c.s, c.wd, qc, N, c.dsum,
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

}  // namespace hyd
