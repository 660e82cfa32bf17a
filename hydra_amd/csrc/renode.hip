// renode.hip -- node redistribution, contours.f90:687-933, for a batch of
// closed contours, one workgroup per contour.
//
// The per-node geometry (increments, cubic coefficients, corner test,
// curvature-based density) is evaluated lane-parallel in the reference's
// expression order.  The two sequential pieces -- the sum of the fractional
// node counts (contours.f90:799-802, whose nint decides how many nodes the
// contour keeps) and the accumulator walk that places the new nodes
// (813-930) -- run as wavefront-uniform chains: 64 consecutive values are
// staged in LDS and every lane reads each one by broadcast, so the serial
// chain is the floating-point additions alone.  The walk is written per old
// node: acc + e(i) is the only rounded operation (acc - 1 is exact for acc
// >= 1), so acc = s - floor(s) after s = acc + e(i) is the reference's
// value after its floor(s) subtractions, and the nodes placed from old node
// i carry acc values s - 1, s - 2, ...; each lane recovers the s of its old
// node from LDS and emits its nodes in parallel.  Positions are then
// evaluated lane-parallel.
//
// Two launches: RN_COUNT computes the geometry, the sum and npr per contour
// (0 = delete) and keeps them; after an exclusive scan of npr, RN_PLACE
// scales, walks and writes the nodes.
#include "hyd_internal.h"

namespace hyd {

struct RenodeArgs {
  // input contours: nodes xd/yd, contour c has npd[c] nodes from off_in[c]
  const double *xd, *yd;
  const int64_t *off_in;
  const int32_t *npd;
  int32_t ncont;
  // scratch, indexed like the input nodes
  double *sdx, *sdy, *sa, *sb, *sc, *sd, *se, *su, *sv, *stm;
  int32_t *wi;     // walk: old node index per new node (indexed like output)
  double *wacc;    // walk: accumulator per new node
  // count pass results per contour, read by the place pass
  double *csum;
  int32_t *cnc, *cfirst;
  // output
  int32_t *npr;    // per contour
  const int64_t *off_out;
  double *xr, *yr;
  // constants
  double ellx, elly, hlxi, hlyi, oms, small, small3, dmsq, elf, dmi, densf;
};

#define F12 0.5
#define F13 (1.0 / 3.0)
#define F16 (1.0 / 6.0)

__device__ __forceinline__ double wrapc(double v, double L, double hli) {
  return v - L * (double)(int)(v * hli);
}

// Shared geometry of contours.f90:717-796 for contour c (all lanes).
__device__ void renode_geometry(const RenodeArgs &g, int64_t o, int n) {
  const int l = threadIdx.x, nt = blockDim.x;
  // increments, v, e
  for (int i = l; i < n; i += nt) {
    const int ia = i + 1 < n ? i + 1 : 0;
    const double xx = g.xd[o + ia] - g.xd[o + i];
    const double dx = wrapc(xx, g.ellx, g.hlxi);
    const double yy = g.yd[o + ia] - g.yd[o + i];
    const double dy = wrapc(yy, g.elly, g.hlyi);
    g.sdx[o + i] = dx;
    g.sdy[o + i] = dy;
    const double v = dx * dx + dy * dy;
    g.sv[o + i] = v;
    g.se[o + i] = sqrt(v);
  }
  __syncthreads();
  // u(i)=v(ib), a(i)=-dx(ib), c(i)=-dy(ib); corner test; b(i)
  for (int i = l; i < n; i += nt) {
    const int ib = i > 0 ? i - 1 : n - 1;
    const double u = g.sv[o + ib], a = -g.sdx[o + ib], c = -g.sdy[o + ib];
    const double dx = g.sdx[o + i], dy = g.sdy[o + i], v = g.sv[o + i];
    const bool corner = dx * a + dy * c > 0.0;
    double b;
    if (corner) {
      b = 0.0;
    } else {
      const double t1 = a * v - dx * u;
      const double t2 = c * v - dy * u;
      b = (dx * c - a * dy) / sqrt(t1 * t1 + t2 * t2 + g.small3);
    }
    g.sb[o + i] = b;
    g.sd[o + i] = corner ? 1.0 : 0.0;  // corner flag (d is reused below)
  }
  __syncthreads();
  // u=e*(b(ia)+b), c=e*(b(ia)-b); a,b,c cubic coefficients; curvature terms
  for (int i = l; i < n; i += nt) {
    const int ia = i + 1 < n ? i + 1 : 0;
    const double e = g.se[o + i], bi = g.sb[o + i], ba = g.sb[o + ia];
    double u = e * (ba + bi);
    double c = e * (ba - bi);
    g.sa[o + i] = F16 * c - F12 * u;
    const double bnew = F12 * (u - c);
    c = F13 * c;
    g.sc[o + i] = c;
    const double v = g.sv[o + i];
    const double ww = 1.0 / (v + g.dmsq);
    u = ww * sqrt(g.elf * v + u * u);
    g.su[o + i] = u;
    g.sv[o + i] = ww * e;
    // b is needed by the neighbour (ba) above: write it to a separate slot
    g.stm[o + i] = bnew;
  }
  __syncthreads();
  for (int i = l; i < n; i += nt) g.sb[o + i] = g.stm[o + i];
  __syncthreads();
  // d(i) = (u(ib)+u(i))/(v(ib)+v(i)+small), kept in stm
  for (int i = l; i < n; i += nt) {
    const int ib = i > 0 ? i - 1 : n - 1;
    g.stm[o + i] = (g.su[o + ib] + g.su[o + i]) / (g.sv[o + ib] + g.sv[o + i] + g.small);
  }
  __syncthreads();
  // u(i) = f12*(d(i)+d(ia)); e(i) = e(i)*min(dmi, densf*sqrt(u)+u)
  for (int i = l; i < n; i += nt) {
    const int ia = i + 1 < n ? i + 1 : 0;
    const double u = F12 * (g.stm[o + i] + g.stm[o + ia]);
    const double t = g.densf * sqrt(u) + u;
    g.se[o + i] = g.se[o + i] * (g.dmi < t ? g.dmi : t);
  }
  __syncthreads();
}

__device__ __forceinline__ double rl_d(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

// Orders the calling wavefront's LDS accesses (LDS serves one wavefront's
// accesses in issue order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// e of the q-th node of a walk that starts at node ibeg and wraps (q < n).
__device__ __forceinline__ double e_at(const double *se, int64_t o, int n, int ibeg, int q) {
  int i = ibeg + q;
  if (i >= n) i -= n;
  return se[o + i];
}

// Sequential sum of e over len nodes from ibeg (wrapping: next() order),
// contours.f90:799-802 and 880-884.  Wavefront-uniform (every lane returns
// the same value); lds holds 64 doubles of the calling wavefront.
__device__ double seq_sum_e(const double *se, int64_t o, int n, int ibeg, int len,
                            double *lds) {
  const int l = threadIdx.x & 63;
  double s = 0.0;
  double v = l < len ? e_at(se, o, n, ibeg, l) : 0.0;
  for (int q0 = 0; q0 < len; q0 += 64) {
    const int qn = q0 + 64 + l;
    const double vn = qn < len ? e_at(se, o, n, ibeg, qn) : 0.0;  // next block in flight
    lds[l] = v;
    wave_sync();
    const int m = len - q0 < 64 ? len - q0 : 64;
    if (m == 64) {
#pragma unroll
      for (int j = 0; j < 64; j++) s = s + lds[j];
    } else {
      for (int j = 0; j < m; j++) s = s + lds[j];
    }
    wave_sync();
    v = vn;
  }
  return s;
}

// The accumulator walk (contours.f90:826-838, 858-868, 907-920) over the
// nodes ibeg, next(ibeg), ... for `want` new nodes, written to out slots
// oo+1 .. oo+want (old node index, acc after the subtraction).  acc starts
// at 0; at most maxlen old nodes are visited.  Wavefront 0 only.
__device__ void walk_nodes(const RenodeArgs &g, int64_t o, int n, int ibeg, int maxlen,
                           int want, int64_t oo, double *lds_e, double *lds_s) {
  const int l = threadIdx.x & 63;
  double acc = 0.0;
  int done = 0;
  double v = l < maxlen ? e_at(g.se, o, n, ibeg, l % n) : 0.0;
  for (int q0 = 0; q0 < maxlen && done < want; q0 += 64) {
    const int qn = q0 + 64 + l;
    const double vn = qn < maxlen ? e_at(g.se, o, n, ibeg, qn % n) : 0.0;
    lds_e[l] = v;
    wave_sync();
    const int m = maxlen - q0 < 64 ? maxlen - q0 : 64;
    if (m == 64) {
#pragma unroll 16
      for (int j = 0; j < 64; j++) {
        const double sj = acc + lds_e[j];
        lds_s[j] = sj;
        acc = sj - floor(sj);
      }
    } else {
      for (int j = 0; j < m; j++) {
        const double sj = acc + lds_e[j];
        lds_s[j] = sj;
        acc = sj - floor(sj);
      }
    }
    wave_sync();
    // old node q0+l places floor(s) new nodes, capped at the total wanted
    const double sl = l < m ? lds_s[l] : 0.0;
    const int k = l < m ? (int)floor(sl) : 0;
    int incl = k;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d, 64);
      if (l >= d) incl += t;
    }
    const int excl = incl - k;
    int i = ibeg + ((q0 + l) % n);
    if (i >= n) i -= n;
    for (int r = 1; r <= k; r++) {
      const int slot = done + excl + r;  // 1-based new node after the first
      if (slot > want) break;
      g.wi[oo + slot] = i;
      g.wacc[oo + slot] = sl - (double)r;
    }
    done += __shfl(incl, 63, 64);
    wave_sync();
    v = vn;
  }
  // (sum(e) = the wanted count, so the walk always ends with done = want;
  // should rounding ever leave slots, they get the last node, never an
  // index outside the contour)
  for (int slot = done + 1 + l; slot <= want; slot += 64) {
    int i = ibeg + (maxlen - 1) % n;
    if (i >= n) i -= n;
    g.wi[oo + slot] = i;
    g.wacc[oo + slot] = 0.0;
  }
}

__device__ __forceinline__ int nint_d(double x) { return (int)llround(x); }

// Place one new node from old node i (0-based) and accumulator acc
// (contours.f90:833-838).
__device__ __forceinline__ void place(const RenodeArgs &g, int64_t o, int i, double acc,
                                      double &xo, double &yo) {
  const double e = g.se[o + i];
  const double p = 1.0 - acc / e;
  const double eta = p * (g.sa[o + i] + p * (g.sb[o + i] + p * g.sc[o + i]));
  const double xx = g.xd[o + i] + p * g.sdx[o + i] - eta * g.sdy[o + i];
  xo = g.oms * wrapc(xx, g.ellx, g.hlxi);
  const double yy = g.yd[o + i] + p * g.sdy[o + i] + eta * g.sdx[o + i];
  yo = g.oms * wrapc(yy, g.elly, g.hlyi);
}

// One workgroup of 256 threads per contour: the lane-parallel phases use
// all of them, the sequential sums and walks run on wavefront 0.
// RN_COUNT (PLACE = false): geometry, corners, sums, npr; RN_PLACE: the
// scaling, the walk and the node positions from the count pass's geometry.
template <bool PLACE>
__global__ void __launch_bounds__(256) renode_kernel(RenodeArgs g) {
  const int c = blockIdx.x;
  if (c >= g.ncont) return;
  const int n = g.npd[c];
  const int64_t o = g.off_in[c];
  const int t = threadIdx.x, l = t & 63, nt = blockDim.x;
  const bool w0 = t < 64;
  __shared__ double lds_e[64], lds_s[64];
  __shared__ int s_ncorn, s_first, s_tot;
  __shared__ double s_sum;
  if (n <= 0) {
    if (!PLACE && t == 0) g.npr[c] = 0;
    return;
  }
  if (!PLACE) {
    renode_geometry(g, o, n);
    // corners (d holds the flag): count, first; the total of e
    // (contours.f90:799-809)
    if (w0) {
      int ncorn = 0, first = -1;
      for (int b0 = 0; b0 < n; b0 += 64) {
        const uint64_t m = __ballot(b0 + l < n && g.sd[o + b0 + l] != 0.0);
        if (m && first < 0) first = b0 + __ffsll((unsigned long long)m) - 1;
        ncorn += __popcll(m);
      }
      const double sum = seq_sum_e(g.se, o, n, 0, n, lds_e);
      if (l == 0) {
        s_ncorn = ncorn;
        s_first = first;
        s_sum = sum;
        g.csum[c] = sum;
        g.cnc[c] = ncorn;
        g.cfirst[c] = first;
      }
    }
    __syncthreads();
  } else if (t == 0) {
    s_sum = g.csum[c];
    s_ncorn = g.cnc[c];
    s_first = g.cfirst[c];
  }
  if (PLACE) __syncthreads();
  const int ncorn = s_ncorn, first = s_first;
  const double sum = s_sum;
  const int npr = nint_d(sum) + 1;
  if (npr < 3) {
    if (!PLACE && t == 0) g.npr[c] = 0;
    return;
  }
  if (ncorn <= 1) {
    if (!PLACE) {
      if (t == 0) g.npr[c] = npr;
      return;
    }
    const double fac = (double)npr / sum;
    for (int i = t; i < n; i += nt) g.se[o + i] = fac * g.se[o + i];
    __syncthreads();
    const int64_t oo = g.off_out[c];
    const int start = ncorn == 0 ? 0 : first;
    if (w0) {
      if (l == 0) {
        g.xr[oo] = g.xd[o + start];
        g.yr[oo] = g.yd[o + start];
      }
      // ncorn = 0 walks i = 1, 2, ...; ncorn = 1 starts at the corner and
      // wraps (contours.f90:826-868)
      walk_nodes(g, o, n, start, n, npr - 1, oo, lds_e, lds_s);
    }
    __syncthreads();
    for (int im = 1 + t; im < npr; im += nt) {
      double xo, yo;
      place(g, o, g.wi[oo + im], g.wacc[oo + im], xo, yo);
      g.xr[oo + im] = xo;
      g.yr[oo + im] = yo;
    }
    return;
  }
  // several corners: renode each stretch between consecutive corners
  // (contours.f90:873-929); wavefront 0 does the sums, scalings and walks.
  if (w0) {
    const int64_t oo = PLACE ? g.off_out[c] : 0;
    int tot = 0;
    int ibeg = first;
    for (int kk = 0; kk < ncorn; kk++) {
      // next corner after ibeg (wrapping to the first one at the end)
      int iend;
      if (kk == ncorn - 1) {
        iend = first;
      } else {
        int j = ibeg + 1;
        for (;;) {
          const int jb = j & ~63;
          const uint64_t m = __ballot(jb + l >= j && jb + l < n && g.sd[o + jb + l] != 0.0);
          if (m) {
            j = jb + __ffsll((unsigned long long)m) - 1;
            break;
          }
          j = jb + 64;
        }
        iend = j;
      }
      // sum over the stretch ibeg .. iend (exclusive), with next() wrapping
      const int len = iend > ibeg ? iend - ibeg : iend + n - ibeg;
      const double sm = seq_sum_e(g.se, o, n, ibeg, len, lds_e);
      if (sm > g.small) {
        const int npseg = nint_d(sm) + 1;
        if (PLACE) {
          const double fac = (double)npseg / sm;
          for (int q0 = l; q0 < len; q0 += 64) {
            const int q = ibeg + q0 < n ? ibeg + q0 : ibeg + q0 - n;
            g.se[o + q] = fac * g.se[o + q];
          }
          __threadfence();  // the walk reads the scaled values (other lanes' stores)
          if (l == 0) {
            g.xr[oo + tot] = g.xd[o + ibeg];
            g.yr[oo + tot] = g.yd[o + ibeg];
            g.wi[oo + tot] = -1;
          }
          if (npseg > 1) walk_nodes(g, o, n, ibeg, n, npseg - 1, oo + tot, lds_e, lds_s);
        }
        tot += npseg;
      }
      ibeg = iend;
    }
    if (l == 0) s_tot = tot;
  }
  __syncthreads();
  if (!PLACE) {
    if (t == 0) g.npr[c] = s_tot;
    return;
  }
  const int64_t oo = g.off_out[c];
  const int tot = s_tot;
  for (int im = t; im < tot; im += nt) {
    const int i = g.wi[oo + im];
    if (i < 0) continue;  // corner copied above
    double xo, yo;
    place(g, o, i, g.wacc[oo + im], xo, yo);
    g.xr[oo + im] = xo;
    g.yr[oo + im] = yo;
  }
}

}  // namespace hyd

namespace hyd {

// Renode a batch of contours: count, exclusive scan, place.  Returns
// HYD_ECAPACITY (with *ntot_out set) when the output would exceed cap_out.
int renode_batch(Ctx &c, const double *xd, const double *yd, const int64_t *off_in,
                 const int32_t *npd, int ncont, int64_t nin, double *xr, double *yr,
                 int64_t *off_out, int32_t *npr, int64_t *ntot_out, int64_t cap_out) {
  *ntot_out = 0;
  if (ncont <= 0) return HYD_OK;
  KTimer tm(c, "renode", 16 * nin * 3);
  RenodeArgs g;
  const Consts &k = c.c;
  g.xd = xd;
  g.yd = yd;
  g.off_in = off_in;
  g.npd = npd;
  g.ncont = ncont;
  const int64_t ns = std::max<int64_t>(nin, 1);
  double **scr[] = {&g.sdx, &g.sdy, &g.sa, &g.sb, &g.sc, &g.sd, &g.se, &g.su, &g.sv, &g.stm};
  static const char *names[] = {"rn_dx", "rn_dy", "rn_a", "rn_b", "rn_c",
                                "rn_d",  "rn_e",  "rn_u", "rn_v", "rn_t"};
  for (int t = 0; t < 10; t++) HYD_CHECK(dbuf(c, names[t], ns, scr[t]));
  g.npr = npr;
  g.off_out = off_out;
  g.xr = xr;
  g.yr = yr;
  g.ellx = k.ellx;
  g.elly = k.elly;
  g.hlxi = k.hlxi;
  g.hlyi = k.hlyi;
  g.oms = k.oms;
  g.small = k.small;
  g.small3 = k.small3;
  g.dmsq = k.dmsq;
  g.elf = k.elf;
  g.dmi = k.dmi;
  g.densf = k.densf;
  g.wi = nullptr;
  g.wacc = nullptr;
  HYD_CHECK(dbuf(c, "rn_csum", ncont, &g.csum));
  HYD_CHECK(dbuf(c, "rn_cnc", ncont, &g.cnc));
  HYD_CHECK(dbuf(c, "rn_cfirst", ncont, &g.cfirst));
  hipLaunchKernelGGL(renode_kernel<false>, dim3(ncont), dim3(256), 0, c.s, g);
  HYD_HIP(hipGetLastError());
  std::vector<int32_t> hn(ncont);
  HYD_HIP(hipMemcpyAsync(hn.data(), npr, 4 * ncont, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  std::vector<int64_t> ho(ncont);
  int64_t tot = 0;
  for (int j = 0; j < ncont; j++) {
    ho[j] = tot;
    tot += hn[j];
  }
  *ntot_out = tot;
  if (tot > cap_out) {
    set_error("renode: output exceeds buffer");
    return HYD_ECAPACITY;
  }
  HYD_HIP(hipMemcpyAsync(off_out, ho.data(), 8 * ncont, hipMemcpyHostToDevice, c.s));
  HYD_CHECK(dbuf(c, "rn_wi", std::max<int64_t>(tot, 1), &g.wi));
  HYD_CHECK(dbuf(c, "rn_wacc", std::max<int64_t>(tot, 1), &g.wacc));
  hipLaunchKernelGGL(renode_kernel<true>, dim3(ncont), dim3(256), 0, c.s, g);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
