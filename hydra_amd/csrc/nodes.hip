// nodes.hip -- per-node kernels: velint (contours.f90:136-169) and the RK4
// node updates of advance (evolution.f90:214-225, 247-256, 274-284, 304-310).
//
// Every expression keeps the reference's evaluation order (no contraction),
// so node positions are bit-identical.  velint is a gather of eight grid
// values per node; the update kernels are pure streams.
#include "hyd_internal.h"

namespace hyd {

struct NodeGeom {
  double xmin, ymin, glxi, glyi, ellx, elly, hlxi, hlyi, oms;
  int nx, ny;
};

__device__ __forceinline__ double wrapx(double xx, const NodeGeom &g) {
  return g.oms * (xx - g.ellx * (double)(int)(xx * g.hlxi));
}
__device__ __forceinline__ double wrapy(double yy, const NodeGeom &g) {
  return g.oms * (yy - g.elly * (double)(int)(yy * g.hlyi));
}

// Bilinear interpolation of (uu,vv) at node i; ixp/iyp of contours.f90:58-66
// are the wrap ix0 == nx -> 1.
template <class G>
__device__ __forceinline__ void interp(const G *__restrict__ uu, const G *__restrict__ vv,
                                       double x, double y, const NodeGeom &g, double &uo,
                                       double &vo) {
  // positions are kept in [-L/2, L/2) by the oms wrap, so 1 <= ix0 <= nx;
  // the clamps only keep a corrupted input inside the arrays
  const double xx = g.glxi * (x - g.xmin);
  const int ix0 = min(max(1 + (int)xx, 1), g.nx);
  const double pxc = (double)ix0 - xx;
  const double px = 1.0 - pxc;
  const int ix1 = ix0 == g.nx ? 1 : ix0 + 1;
  const double yy = g.glyi * (y - g.ymin);
  const int iy0 = min(max(1 + (int)yy, 1), g.ny);
  const double pyc = (double)iy0 - yy;
  const double py = 1.0 - pyc;
  const int iy1 = iy0 == g.ny ? 1 : iy0 + 1;
  const int64_t c00 = (int64_t)(ix0 - 1) * g.ny + (iy0 - 1);
  const int64_t c01 = (int64_t)(ix1 - 1) * g.ny + (iy0 - 1);
  const int64_t c10 = (int64_t)(ix0 - 1) * g.ny + (iy1 - 1);
  const int64_t c11 = (int64_t)(ix1 - 1) * g.ny + (iy1 - 1);
  uo = pyc * (pxc * (double)uu[c00] + px * (double)uu[c01]) +
       py * (pxc * (double)uu[c10] + px * (double)uu[c11]);
  vo = pyc * (pxc * (double)vv[c00] + px * (double)vv[c01]) +
       py * (pxc * (double)vv[c10] + px * (double)vv[c11]);
}

__global__ void velint_kernel(const double *__restrict__ uu, const double *__restrict__ vv,
                              const double *__restrict__ xq, const double *__restrict__ yq,
                              int n, double *__restrict__ uq, double *__restrict__ vq,
                              NodeGeom g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double u, v;
  interp(uu, vv, xq[i], yq[i], g, u, v);
  uq[i] = u;
  vq[i] = v;
}

// Fused velint + RK stage update.  stage 1..4 as in evolution.f90.  G is the
// grid type (double in the plain step, float in the mixed-precision step).
template <class G>
__global__ void rk_nodes_kernel(const G *__restrict__ uu, const G *__restrict__ vv,
                                double *__restrict__ xq, double *__restrict__ yq,
                                double *__restrict__ xqi, double *__restrict__ yqi,
                                double *__restrict__ xqf, double *__restrict__ yqf, int n,
                                int stage, const Scalars *__restrict__ sc, NodeGeom g,
                                const float *__restrict__ du, const float *__restrict__ dv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = sc->dt, dt2 = sc->dt2, dt3 = sc->dt3, dt6 = sc->dt6;
  double uq, vq;
  interp(uu, vv, xq[i], yq[i], g, uq, vq);
  if (du) {  // the learned correction, when a network is attached
    uq = uq + (double)du[i];
    vq = vq + (double)dv[i];
  }
  if (stage == 1) {
    const double x0 = xq[i], y0 = yq[i];
    xqi[i] = x0;
    yqi[i] = y0;
    xq[i] = wrapx(x0 + dt2 * uq, g);
    yq[i] = wrapy(y0 + dt2 * vq, g);
    xqf[i] = wrapx(x0 + dt6 * uq, g);
    yqf[i] = wrapy(y0 + dt6 * vq, g);
  } else if (stage == 2) {
    xq[i] = wrapx(xqi[i] + dt2 * uq, g);
    yq[i] = wrapy(yqi[i] + dt2 * vq, g);
    xqf[i] = wrapx(xqf[i] + dt3 * uq, g);
    yqf[i] = wrapy(yqf[i] + dt3 * vq, g);
  } else if (stage == 3) {
    xq[i] = wrapx(xqi[i] + dt * uq, g);
    yq[i] = wrapy(yqi[i] + dt * vq, g);
    xqf[i] = wrapx(xqf[i] + dt3 * uq, g);
    yqf[i] = wrapy(yqf[i] + dt3 * vq, g);
  } else {
    xq[i] = wrapx(xqf[i] + dt6 * uq, g);
    yq[i] = wrapy(yqf[i] + dt6 * vq, g);
  }
}

static NodeGeom geom(const Ctx &c) {
  NodeGeom g;
  g.xmin = c.c.xmin;
  g.ymin = c.c.ymin;
  g.glxi = c.c.glxi;
  g.glyi = c.c.glyi;
  g.ellx = c.c.ellx;
  g.elly = c.c.elly;
  g.hlxi = c.c.hlxi;
  g.hlyi = c.c.hlyi;
  g.oms = c.c.oms;
  g.nx = c.c.nx;
  g.ny = c.c.ny;
  return g;
}

int velint(Ctx &c, const double *uu, const double *vv, const double *xq, const double *yq, int n,
           double *uq, double *vq) {
  if (n <= 0) return HYD_OK;
  KTimer t(c, "velint", (int64_t)n * 32 + 0);
  hipLaunchKernelGGL(velint_kernel, dim3((n + 255) / 256), dim3(256), 0, c.s, uu, vv, xq, yq, n,
                     uq, vq, geom(c));
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

// The learned correction of this stage at the current node positions:
// physical PV q = spctop(qq) of the stage's inversion, then the network on
// (u, v, q) sampled at the nodes (cnn.hip), into c.cdu, c.cdv.
int correction(Ctx &c) {
  const int64_t N = c.N;
  double *q, *tmp;
  HYD_CHECK(dbuf(c, "cnn_q", N, &q));
  HYD_CHECK(dbuf(c, "cnn_tmp", N, &tmp));
  // sized by the node capacity: allocated once, outside any captured step
  HYD_CHECK(dbuf(c, "cnn_du", c.cap, &c.cdu));
  HYD_CHECK(dbuf(c, "cnn_dv", c.cap, &c.cdv));
  HYD_CHECK(fft_spctop(c, c.qq, q, PRE_NONE, 1.0, tmp));
  const bool f32 = c.p.precision == 1;
  return cnn_correction(c.net, c.s, f32 ? (const void *)c.uu32 : (const void *)c.uu,
                        f32 ? (const void *)c.vv32 : (const void *)c.vv, q, f32, c.net_su,
                        c.net_sq, c.net_sd, c.xq, c.yq, c.nptq, c.cdu, c.cdv);
}

int rk_nodes(Ctx &c, int stage) {
  const int n = c.nptq;
  if (n <= 0 || !c.contex) return HYD_OK;
  static const int64_t words[5] = {0, 8, 10, 10, 6};
  const float *du = nullptr, *dv = nullptr;
  if (c.net) {
    HYD_CHECK(correction(c));
    du = c.cdu;
    dv = c.cdv;
  }
  KTimer t(c, "rk_nodes", (int64_t)n * 8 * words[stage] + 16 * c.N);
  if (c.p.precision == 1)
    hipLaunchKernelGGL(rk_nodes_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, c.s, c.uu32,
                       c.vv32, c.xq, c.yq, c.xqi, c.yqi, c.xqf, c.yqf, n, stage, c.dsc, geom(c),
                       du, dv);
  else
    hipLaunchKernelGGL(rk_nodes_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, c.s, c.uu,
                       c.vv, c.xq, c.yq, c.xqi, c.yqi, c.xqf, c.yqf, n, stage, c.dsc, geom(c),
                       du, dv);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
