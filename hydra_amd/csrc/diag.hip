// diag.hip -- the reference's diagnostics on gfx950, bit for bit:
//
//   * energy (evolution.f90:612-644): ekin = 1/2 garea sum(uu^2 + vv^2),
//     epot = 1/2 garea kd^2 sum(spctop(pp)^2) (0 when kd = 0);
//   * l1norm / l2norm (generic.f90:30-64): garea sum(|qq|), garea sum(qq^2);
//   * spec1d (spectral.f90:370-424): the 1D spectrum of a spectral field.
//
// The field sums are the reference's left-to-right sums, done by seqsum.
// spec1d adds each mode's contribution to its shell k = kmag(kx,ky) in the
// reference's loop order ((1,1); kx = 2..nx at ky = 1; ky = 2..ny at
// kx = 1; then ky outer, kx inner), so every shell is an independent
// sequential sum over a fixed list of modes: the host lists them once per
// context (a stable counting sort by shell), and one wavefront per shell
// adds its list in order, the terms evaluated 64 at a time in parallel and
// added through readlane.
#include "hyd_internal.h"

namespace hyd {

// entry: mode index (< 2^29) | weight code << 29; code 0: ss^2, 1: f12*ss^2,
// 2: f14*ss^2
static constexpr int kShellCodeShift = 29;

__global__ void __launch_bounds__(64) k_spec_shells(const double *__restrict__ ss,
                                                    const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ ent,
                                                    double *__restrict__ spec) {
  const int k = blockIdx.x, l = threadIdx.x;
  const int b = ptr[k], e = ptr[k + 1];
  double s = 0.0;  // spec(k) = zero, spectral.f90:395-397
  for (int j0 = b; j0 < e; j0 += 64) {
    const int j = j0 + l;
    double term = 0.0;
    if (j < e) {
      const int32_t w = ent[j];
      const double v = ss[w & ((1 << kShellCodeShift) - 1)];
      const int code = (int)((uint32_t)w >> kShellCodeShift);
      const double sq = v * v;
      term = code == 2 ? 0.25 * sq : (code == 1 ? 0.5 * sq : sq);
    }
    const int cnt = min(64, e - j0);
    for (int t = 0; t < cnt; t++) s = s + readlane_f64(term, t);
  }
  if (l == 0) spec[k] = s;
}

__global__ void k_scale_rksq(double *ss, const double *rksq, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) ss[k] = ss[k] * rksq[k];
}

// out = a^2 + b^2 (b may be null: out = a^2); mode 1: out = |a|
__global__ void k_sq(const double *a, const double *b, double *out, int64_t n, int mode) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double x = a[k];
  if (mode == 1)
    out[k] = fabs(x);
  else
    out[k] = b ? x * x + b[k] * b[k] : x * x;
}

// Shell lists, built on first use (host tables are the reference's kmag).
static int diag_setup(Ctx &c) {
  if (c.shell_ptr) return HYD_OK;
  const int nx = c.c.nx, ny = c.c.ny, kmax = c.ht.kmax;
  const int64_t N = c.N;
  if (N >= ((int64_t)1 << kShellCodeShift)) {
    set_error("spec1d: grid too large for the shell lists");
    return HYD_EARG;
  }
  std::vector<int32_t> order, code;
  order.reserve(N);
  code.reserve(N);
  order.push_back(0);
  code.push_back(2);
  for (int kx = 1; kx < nx; kx++) {
    order.push_back(kx);
    code.push_back(1);
  }
  for (int ky = 1; ky < ny; ky++) {
    order.push_back(ky * nx);
    code.push_back(1);
  }
  for (int ky = 1; ky < ny; ky++)
    for (int kx = 1; kx < nx; kx++) {
      order.push_back(ky * nx + kx);
      code.push_back(0);
    }
  std::vector<int32_t> ptr(kmax + 2, 0), ent(N);
  for (int64_t i = 0; i < N; i++) {
    const int k = c.ht.kmag[order[i]];
    if (k < 0 || k > kmax) {
      set_error("spec1d: shell index out of range");
      return HYD_EARG;
    }
    ptr[k + 1]++;
  }
  for (int k = 0; k <= kmax; k++) ptr[k + 1] += ptr[k];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t i = 0; i < N; i++) {
    const int k = c.ht.kmag[order[i]];
    ent[fill[k]++] = order[i] | (code[i] << kShellCodeShift);
  }
  HYD_HIP(hipMalloc((void **)&c.shell_ptr, sizeof(int32_t) * (kmax + 2)));
  HYD_HIP(hipMalloc((void **)&c.shell_ent, sizeof(int32_t) * N));
  HYD_HIP(hipMalloc((void **)&c.rksq, sizeof(double) * N));
  HYD_HIP(hipMalloc((void **)&c.diag_out, sizeof(double) * 8));
  HYD_HIP(hipMemcpyAsync(c.shell_ptr, ptr.data(), sizeof(int32_t) * (kmax + 2),
                         hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(c.shell_ent, ent.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice,
                         c.s));
  HYD_HIP(hipMemcpyAsync(c.rksq, c.ht.rksq.data(), sizeof(double) * N, hipMemcpyHostToDevice,
                         c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

void diag_free(Ctx &c) {
  for (void *p : {(void *)c.shell_ptr, (void *)c.shell_ent, (void *)c.rksq, (void *)c.diag_out})
    if (p) hipFree(p);
  c.shell_ptr = c.shell_ent = nullptr;
  c.rksq = c.diag_out = nullptr;
}

// sequential sum of f(a[, b]) over N values into c.diag_out[slot]
static int field_sum(Ctx &c, const double *a, const double *b, int mode, double *work, int slot) {
  const int64_t N = c.N;
  hipLaunchKernelGGL(k_sq, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c.s, a, b, work, N,
                     mode);
  HYD_HIP(hipGetLastError());
  return seqsum(c, work, N, c.diag_out + slot);
}

static int read_out(Ctx &c, double *h, int n) {
  HYD_HIP(hipMemcpyAsync(h, c.diag_out, sizeof(double) * n, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

// evolution.f90:612-644 from the context's uu, vv (physical) and pp
// (spectral), as left by the last inversion.
int diag_energy(Ctx &c, double *ekin, double *epot) {
  HYD_CHECK(diag_setup(c));
  HYD_CHECK(field_sum(c, c.uu, c.vv, 0, c.wa, 0));
  if (!c.c.btropic) {
    HYD_CHECK(fft_spctop(c, c.pp, c.wc, PRE_NONE, 1.0, c.wd));
    HYD_CHECK(field_sum(c, c.wc, nullptr, 0, c.wa, 1));
  }
  double h[2] = {0.0, 0.0};
  HYD_CHECK(read_out(c, h, 2));
  *ekin = 0.5 * c.c.garea * h[0];
  *epot = c.c.btropic ? 0.0 : 0.5 * c.c.garea * c.c.kdsq * h[1];
  return HYD_OK;
}

// generic.f90:30-64 of a physical (ny,nx) field already on the device
int diag_norms(Ctx &c, const double *qq, double *ql1, double *ql2) {
  HYD_CHECK(diag_setup(c));
  HYD_CHECK(field_sum(c, qq, nullptr, 1, c.wd, 0));
  HYD_CHECK(field_sum(c, qq, nullptr, 0, c.wd, 1));
  double h[2];
  HYD_CHECK(read_out(c, h, 2));
  *ql1 = c.c.garea * h[0];
  *ql2 = c.c.garea * h[1];
  return HYD_OK;
}

// spectral.f90:370-424 of a spectral (nx,ny) field on the device (ss is
// multiplied by rksq in place when iopt = 1); spec (device) gets 0..kmax.
int diag_spec1d(Ctx &c, double *ss, double *spec, int iopt) {
  HYD_CHECK(diag_setup(c));
  if (iopt == 1) {
    hipLaunchKernelGGL(k_scale_rksq, dim3((unsigned)((c.N + 255) / 256)), dim3(256), 0, c.s, ss,
                       c.rksq, c.N);
    HYD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_spec_shells, dim3(c.ht.kmax + 1), dim3(64), 0, c.s, ss, c.shell_ptr,
                     c.shell_ent, spec);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
