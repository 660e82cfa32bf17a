// renode.hip -- node redistribution, contours.f90:687-933, for a batch of
// closed contours, one wavefront per contour.
//
// The per-node geometry (increments, cubic coefficients, corner test,
// curvature-based density) is evaluated lane-parallel in the reference's
// expression order.  The two sequential pieces -- the sum of the fractional
// node counts (contours.f90:799-802, whose nint decides how many nodes the
// contour keeps) and the accumulator walk that places the new nodes
// (813-930) -- run on lane 0 exactly as written, so node counts and
// positions are identical to the reference.  The walk only records (old
// segment, accumulator) per new node; the positions are then evaluated
// lane-parallel.
//
// Two launches: RN_COUNT computes npr per contour (0 = delete); after an
// exclusive scan of npr, RN_PLACE writes the nodes.
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
  const int l = threadIdx.x;
  // increments, v, e
  for (int i = l; i < n; i += 64) {
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
  for (int i = l; i < n; i += 64) {
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
  for (int i = l; i < n; i += 64) {
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
  for (int i = l; i < n; i += 64) g.sb[o + i] = g.stm[o + i];
  __syncthreads();
  // d(i) = (u(ib)+u(i))/(v(ib)+v(i)+small), kept in stm
  for (int i = l; i < n; i += 64) {
    const int ib = i > 0 ? i - 1 : n - 1;
    g.stm[o + i] = (g.su[o + ib] + g.su[o + i]) / (g.sv[o + ib] + g.sv[o + i] + g.small);
  }
  __syncthreads();
  // u(i) = f12*(d(i)+d(ia)); e(i) = e(i)*min(dmi, densf*sqrt(u)+u)
  for (int i = l; i < n; i += 64) {
    const int ia = i + 1 < n ? i + 1 : 0;
    const double u = F12 * (g.stm[o + i] + g.stm[o + ia]);
    const double t = g.densf * sqrt(u) + u;
    g.se[o + i] = g.se[o + i] * (g.dmi < t ? g.dmi : t);
  }
  __syncthreads();
}

// Sequential sum of e over [b, e) in node order (lane 0).
__device__ __forceinline__ double seq_e(const double *se, int64_t o, int b, int e) {
  double s = 0.0;
  for (int i = b; i < e; i++) s = s + se[o + i];
  return s;
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

template <bool PLACE>
__global__ void __launch_bounds__(64) renode_kernel(RenodeArgs g) {
  const int c = blockIdx.x;
  if (c >= g.ncont) return;
  const int n = g.npd[c];
  const int64_t o = g.off_in[c];
  const int l = threadIdx.x;
  if (n <= 0) {
    if (!PLACE && l == 0) g.npr[c] = 0;
    return;
  }
  renode_geometry(g, o, n);
  // corners (d holds the flag); count them and find the first
  __shared__ int s_ncorn, s_first;
  if (l == 0) {
    int nc = 0, first = -1;
    for (int i = 0; i < n; i++)
      if (g.sd[o + i] != 0.0) {
        if (first < 0) first = i;
        nc++;
      }
    s_ncorn = nc;
    s_first = first;
  }
  __syncthreads();
  const int ncorn = s_ncorn;
  // total and node count (contours.f90:799-809)
  __shared__ double s_sum;
  if (l == 0) s_sum = seq_e(g.se, o, 0, n);
  __syncthreads();
  const double sum = s_sum;
  int npr = nint_d(sum) + 1;
  if (npr < 3) {
    if (!PLACE && l == 0) g.npr[c] = 0;
    return;
  }
  if (ncorn <= 1) {
    if (!PLACE) {
      if (l == 0) g.npr[c] = npr;
      return;
    }
    const double fac = (double)npr / sum;
    for (int i = l; i < n; i += 64) g.se[o + i] = fac * g.se[o + i];
    __syncthreads();
    const int64_t oo = g.off_out[c];
    const int start = ncorn == 0 ? 0 : s_first;
    if (l == 0) {
      g.xr[oo] = g.xd[o + start];
      g.yr[oo] = g.yd[o + start];
      // accumulator walk; ncorn=0 walks i = 1,2,.. ; ncorn=1 starts after
      // the corner with next() wrapping (contours.f90:826-868)
      double acc = 0.0;
      int i = ncorn == 0 ? -1 : start - 1;
      for (int im = 1; im < npr; im++) {
        while (acc < 1.0) {
          i = (ncorn == 0) ? i + 1 : (i + 1 < n ? i + 1 : 0);
          if (ncorn == 0 && i >= n) { i = n - 1; break; }  // unreachable: sum(e)=npr
          acc = acc + g.se[o + i];
        }
        acc = acc - 1.0;
        g.wi[oo + im] = i;
        g.wacc[oo + im] = acc;
      }
    }
    __syncthreads();
    for (int im = 1 + l; im < npr; im += 64) {
      double xo, yo;
      place(g, o, g.wi[oo + im], g.wacc[oo + im], xo, yo);
      g.xr[oo + im] = xo;
      g.yr[oo + im] = yo;
    }
    return;
  }
  // several corners: renode each stretch between consecutive corners
  // (contours.f90:873-929).  Lane 0 does the sequential sums and walks.
  if (l == 0) {
    const int64_t oo = PLACE ? g.off_out[c] : 0;
    int tot = 0;
    // list corners in order
    int k = 0;
    int ibeg = s_first;
    int corner_idx = s_first;
    for (int kk = 0; kk < ncorn; kk++) {
      // next corner after corner_idx (wrapping to the first one at the end)
      int iend;
      if (kk == ncorn - 1) {
        iend = s_first;
      } else {
        int j = corner_idx + 1;
        while (g.sd[o + j] == 0.0) j++;
        iend = j;
      }
      // sum over the stretch ibeg .. iend (exclusive), with next() wrapping
      double sm = 0.0;
      int i = ibeg;
      do {
        sm = sm + g.se[o + i];
        i = i + 1 < n ? i + 1 : 0;
      } while (i != iend);
      if (sm > g.small) {
        const int npseg = nint_d(sm) + 1;
        if (PLACE) {
          const double fac = (double)npseg / sm;
          i = ibeg;
          do {
            g.se[o + i] = fac * g.se[o + i];
            i = i + 1 < n ? i + 1 : 0;
          } while (i != iend);
          g.xr[oo + tot] = g.xd[o + ibeg];
          g.yr[oo + tot] = g.yd[o + ibeg];
          g.wi[oo + tot] = -1;
          if (npseg > 1) {
            double acc = 0.0;
            i = ibeg - 1;
            for (int im = tot + 1; im < tot + npseg; im++) {
              while (acc < 1.0) {
                i = i + 1 < n ? i + 1 : 0;
                acc = acc + g.se[o + i];
              }
              acc = acc - 1.0;
              g.wi[oo + im] = i;
              g.wacc[oo + im] = acc;
            }
          }
        }
        tot += npseg;
      }
      ibeg = iend;
      corner_idx = iend;
      k++;
    }
    (void)k;
    if (!PLACE) g.npr[c] = tot;
  }
  if (!PLACE) return;
  __syncthreads();
  const int64_t oo = g.off_out[c];
  const int tot = g.npr[c];
  for (int im = l; im < tot; im += 64) {
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
  hipLaunchKernelGGL(renode_kernel<false>, dim3(ncont), dim3(64), 0, c.s, g);
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
  hipLaunchKernelGGL(renode_kernel<true>, dim3(ncont), dim3(64), 0, c.s, g);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
