// step.hip -- the RK4 step of evolution.f90 on the device-resident state.
//
//   inversion  evolution.f90:494-520   con2grid -> ptospc -> qq -> main_invert
//   adapt      evolution.f90:524-608   spctop(qq+kd^2 pp) -> sums / maxima -> dt
//   source     evolution.f90:432-490   gradients -> products -> ptospc
//   advance    evolution.f90:166-326   four stages of the above plus the
//                                      node and spectral RK updates
//
// Pointwise kernels keep the reference's expression order.  The two
// order-sensitive reductions (qavg0 inside con2grid, zzrms here) use the
// exact sequential sum; the maxima are order-free.  The integrating factor
// exp(dt/2*zzrms*qdiss) is evaluated on the host with the same libm as the
// reference, once per distinct qdiss value, and expanded on the device.
#include <math.h>

#include <condition_variable>
#include <mutex>
#include <thread>

#include "hyd_internal.h"

namespace hyd {

int rk_nodes(Ctx &c, int stage);

#define PW_KERNEL(name, params, body)                                   \
  __global__ void name params {                                         \
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   \
    if (k >= n) return;                                                 \
    body;                                                               \
  }

// qq = flo*(qs-qc)+qc+qd  (evolution.f90:514)
PW_KERNEL(k_combine_qq,
          (double *__restrict__ qq, const double *__restrict__ flo, const double *__restrict__ qs,
           const double *__restrict__ qc, const double *__restrict__ qd, int64_t n),
          { qq[k] = flo[k] * (qs[k] - qc[k]) + qc[k] + qd[k]; })

// qq as k_combine_qq, then pp = green*qq (spectral.f90:321-325), one pass
PW_KERNEL(k_combine_green,
          (double *__restrict__ qq, double *__restrict__ pp, const double *__restrict__ flo,
           const double *__restrict__ green, const double *__restrict__ qs,
           const double *__restrict__ qc, const double *__restrict__ qd, int64_t n),
          {
            const double q = flo[k] * (qs[k] - qc[k]) + qc[k] + qd[k];
            qq[k] = q;
            pp[k] = green[k] * q;
          })

// pp = green*qq  (spectral.f90:321-325)
PW_KERNEL(k_green, (double *__restrict__ pp, const double *__restrict__ green,
                    const double *__restrict__ qq, int64_t n),
          { pp[k] = green[k] * qq[k]; })

// qs = flo*(qs-qc)+qc+qd ; qd = fhi*(qs-qc)  (evolution.f90:199-200, 126)
PW_KERNEL(k_reset, (double *__restrict__ qs, double *__restrict__ qd,
                    const double *__restrict__ qc, const double *__restrict__ flo,
                    const double *__restrict__ fhi, int64_t n),
          {
            const double s = flo[k] * (qs[k] - qc[k]) + qc[k] + qd[k];
            qs[k] = s;
            qd[k] = fhi[k] * (s - qc[k]);
          })

// qd = fhi*(qs-qc)  (evolution.f90:126)
PW_KERNEL(k_init_qd, (double *__restrict__ qd, const double *__restrict__ qs,
                      const double *__restrict__ qc, const double *__restrict__ fhi, int64_t n),
          { qd[k] = fhi[k] * (qs[k] - qc[k]); })

// ss = qq + kdsq*pp  (evolution.f90:546)
PW_KERNEL(k_qq_kd, (double *__restrict__ ss, const double *__restrict__ qq,
                    const double *__restrict__ pp, double kdsq, int64_t n),
          { ss[k] = qq[k] + kdsq * pp[k]; })

// zz = zz**2 and the maxima of zz**2 and uu**2+vv**2 (evolution.f90:552-556),
// one pair of maxima per block in mx[2 b], mx[2 b + 1] (k_adapt_scalars
// takes the maximum over the blocks: no same-address atomics)
__global__ void __launch_bounds__(256) k_adapt_reduce(double *__restrict__ zz,
                                                      const double *__restrict__ uu,
                                                      const double *__restrict__ vv, int64_t n,
                                                      unsigned long long *__restrict__ mx) {
  double z2m = 0.0, u2m = 0.0;
  // four elements per thread and pass, all loads in flight before any use
  constexpr int U = 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k0 < n; k0 += U * stride) {
    double z[U], u[U], v[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t k = k0 + j * stride;
      const bool in = k < n;
      z[j] = in ? zz[k] : 0.0;
      u[j] = in ? uu[k] : 0.0;
      v[j] = in ? vv[k] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t k = k0 + j * stride;
      const double z2 = z[j] * z[j];
      if (k < n) zz[k] = z2;
      z2m = fmax(z2m, z2);
      u2m = fmax(u2m, u[j] * u[j] + v[j] * v[j]);
    }
  }
  // block maxima (non-negative doubles order like their bit patterns)
  for (int off = 32; off > 0; off >>= 1) {
    z2m = fmax(z2m, __shfl_down(z2m, off));
    u2m = fmax(u2m, __shfl_down(u2m, off));
  }
  __shared__ double sz[4], su[4];
  if ((threadIdx.x & 63) == 0) {
    sz[threadIdx.x >> 6] = z2m;
    su[threadIdx.x >> 6] = u2m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      z2m = fmax(z2m, sz[w]);
      u2m = fmax(u2m, su[w]);
    }
    mx[2 * blockIdx.x] = (unsigned long long)__double_as_longlong(z2m);
    mx[2 * blockIdx.x + 1] = (unsigned long long)__double_as_longlong(u2m);
  }
}

struct AdaptArgs {
  double dsumi, small, glx, srwfm, dt_fixed, tgsave, tcsave;
};

// evolution.f90:553-578: scalars of adapt.  mx holds nparts pairs of
// maxima (of zz**2 and of uu**2+vv**2, as bit patterns of non-negative
// doubles); one wavefront.
__global__ void k_adapt_scalars(Scalars *sc, const double *sum, const unsigned long long *mx,
                                int nparts, AdaptArgs a, const int32_t *sweep_err) {
  double zm2 = 0.0, um2 = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 64) {
    zm2 = fmax(zm2, __longlong_as_double((long long)mx[2 * i]));
    um2 = fmax(um2, __longlong_as_double((long long)mx[2 * i + 1]));
  }
  for (int off = 32; off > 0; off >>= 1) {
    zm2 = fmax(zm2, __shfl_down(zm2, off));
    um2 = fmax(um2, __shfl_down(um2, off));
  }
  if (threadIdx.x != 0) return;
  sc->sweep_err = sweep_err ? (double)*sweep_err : 0.0;
  const double dtfac = 3.141592653589793238462643383279502884197169399375105820974944592307816 / 10.0;
  const double cflmax = 0.7;
  const double zzrms = sqrt(a.dsumi * sum[0]);
  const double zzmax = sqrt(zm2) + a.small;
  const double uumax = sqrt(um2) + a.small;
  const double t1 = a.glx * cflmax / uumax;
  const double zm = zzmax > a.srwfm ? zzmax : a.srwfm;
  const double t2 = dtfac / zm;
  const double dtacc = t1 < t2 ? t1 : t2;
  double dt = sc->dt;
  if (a.dt_fixed > 0.0) {
    dt = a.dt_fixed;
  } else if (dt > 0.0) {
    dt = dtacc < sc->dtmax ? dtacc : sc->dtmax;
  } else {
    const double tm = a.tgsave < a.tcsave ? a.tgsave : a.tcsave;
    sc->dtmax = 0.2 * tm;
    dt = dtacc < sc->dtmax ? dtacc : sc->dtmax;
  }
  sc->dt = dt;
  sc->dt2 = dt * 0.5;
  sc->dt3 = dt * (1.0 / 3.0);
  sc->dt6 = dt * (1.0 / 6.0);
  sc->twist = sc->twist + dt * zzmax;
  sc->zzrms = zzrms;
  sc->zzmax = zzmax;
  sc->uumax = uumax;
  sc->dtacc = dtacc;
}

// epq = exp(dfac*qdiss) from the host-evaluated distinct values; emq = 1/epq
PW_KERNEL(k_expand_epq, (double *__restrict__ epq, double *__restrict__ emq,
                         const int32_t *__restrict__ map, const double *__restrict__ e,
                         int64_t n),
          {
            const double v = e[map[k]];
            epq[k] = v;
            emq[k] = 1.0 / v;
          })

// qqx = -uu*qqx - vv*qqy  (evolution.f90:449, 460) / with beta (458)
PW_KERNEL(k_advect_prod, (double *__restrict__ qqx, const double *__restrict__ qqy,
                          const double *__restrict__ uu, const double *__restrict__ vv,
                          double beta, int withbeta, int64_t n),
          {
            if (withbeta)
              qqx[k] = -uu[k] * qqx[k] - vv[k] * (qqy[k] + beta);
            else
              qqx[k] = -uu[k] * qqx[k] - vv[k] * qqy[k];
          })

// sqd = sqd - rekman*(qq+kdsq*pp), then the integrating factor (476, 483-487)
PW_KERNEL(k_source_finish, (double *__restrict__ sqd, const double *__restrict__ qq,
                            const double *__restrict__ pp, const double *__restrict__ epq,
                            double rekman, double kdsq, int lev, int64_t n),
          {
            double s = sqd[k] - rekman * (qq[k] + kdsq * pp[k]);
            if (lev == 1) s = epq[k] * s;
            else if (lev == 2) s = epq[k] * epq[k] * s;
            sqd[k] = s;
          })

// spectral RK updates, evolution.f90:228-233, 259-262, 287-291, 313-314
__global__ void k_rk_fields(double *__restrict__ qs, double *__restrict__ qd,
                            double *__restrict__ qsi, double *__restrict__ qsf,
                            double *__restrict__ qdi, double *__restrict__ qdf,
                            const double *__restrict__ sqs, const double *__restrict__ sqd,
                            double *__restrict__ emq, int stage, const Scalars *__restrict__ sc,
                            int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double dt = sc->dt, dt2 = sc->dt2, dt3 = sc->dt3, dt6 = sc->dt6;
  if (stage == 1) {
    const double s0 = qs[k], d0 = qd[k];
    qsi[k] = s0;
    qs[k] = s0 + dt2 * sqs[k];
    qsf[k] = s0 + dt6 * sqs[k];
    qdi[k] = d0;
    qd[k] = emq[k] * (d0 + dt2 * sqd[k]);
    qdf[k] = d0 + dt6 * sqd[k];
  } else if (stage == 2) {
    qs[k] = qsi[k] + dt2 * sqs[k];
    qsf[k] = qsf[k] + dt3 * sqs[k];
    qd[k] = emq[k] * (qdi[k] + dt2 * sqd[k]);
    qdf[k] = qdf[k] + dt3 * sqd[k];
  } else if (stage == 3) {
    qs[k] = qsi[k] + dt * sqs[k];
    qsf[k] = qsf[k] + dt3 * sqs[k];
    const double e = emq[k] * emq[k];
    emq[k] = e;
    qd[k] = e * (qdi[k] + dt * sqd[k]);
    qdf[k] = qdf[k] + dt3 * sqd[k];
  } else {
    qs[k] = qsf[k] + dt6 * sqs[k];
    qd[k] = emq[k] * (qdf[k] + dt6 * sqd[k]);
  }
}

static inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// The contour part of inversion: qc = ptospc(con2grid) (evolution.f90:507-508).
static int contour_grid(Ctx &c) {
  // with per-launch timing (one stream) the deferred work runs first, so
  // that the contour grid's records hold its own kernels only
  if ((c.nptq == 0 || c.timing) && c.after_sweep) {
    auto fn = std::move(c.after_sweep);
    c.after_sweep = nullptr;
    HYD_CHECK(fn());
  }
  if (c.nptq == 0) {
    HYD_CHECK(con2grid(c, c.xq, c.yq, c.nextq, c.nptq, c.c2g_q, nullptr));
    return fft_ptospc(c, c.c2g_q, c.qc);
  }
  HYD_CHECK(con2grid(c, c.xq, c.yq, c.nextq, c.nptq, nullptr, nullptr, true));
  return fft_ptospc_mean(c, c.c2g_w, c.dsum, (double)c.N, c.qc);
}

// The field part of inversion: qq, pp and the velocity (evolution.f90:510-520).
static int inversion_fields(Ctx &c) {
  const int64_t N = c.N;
  {
    KTimer t(c, "combine_qq", 56 * N);
    hipLaunchKernelGGL(k_combine_green, grid1(N), dim3(256), 0, c.s, c.qq, c.pp, c.flo, c.green,
                       c.qs, c.qc, c.qd, N);
  }
  HYD_HIP(hipGetLastError());
  // main_invert: vv = spctop(xderiv(pp)), uu = -spctop(yderiv(pp)), one batch
  {
    const double *in[2] = {c.pp, c.pp};
    double *out[2] = {c.vv, c.uu};
    const int pre[2] = {PRE_XDERIV, PRE_YDERIV};
    const double sg[2] = {1.0, -1.0};
    HYD_CHECK(fft_spctop_n(c, 2, in, out, pre, sg, c.t4));
  }
  return HYD_OK;
}

int inversion(Ctx &c) {
  HYD_CHECK(contour_grid(c));
  return inversion_fields(c);
}

// main_invert (spectral.f90:304-342) on device arrays; tmp is a work array.
int main_invert_dev(Ctx &c, const double *qq, double *uu, double *vv, double *pp, double *tmp) {
  const int64_t N = c.N;
  hipLaunchKernelGGL(k_green, grid1(N), dim3(256), 0, c.s, pp, c.green, qq, N);
  HYD_HIP(hipGetLastError());
  HYD_CHECK(fft_spctop(c, pp, vv, PRE_XDERIV, 1.0, tmp));
  HYD_CHECK(fft_spctop(c, pp, uu, PRE_YDERIV, -1.0, tmp));
  return HYD_OK;
}

// gradient (spectral.f90:345-366) on device arrays.
int gradient_dev(Ctx &c, const double *ff, double *fx, double *fy, double *tmp) {
  HYD_CHECK(fft_spctop(c, ff, fx, PRE_XDERIV, 1.0, tmp));
  HYD_CHECK(fft_spctop(c, ff, fy, PRE_YDERIV, 1.0, tmp));
  return HYD_OK;
}

// exp(dfac*qdiss_u[j]) on the host (the reference's libm exp), split over
// persistent worker threads: creating threads every step would cost more
// than the table itself.  One process-wide pool; concurrent callers take
// turns.
class ExpPool {
 public:
  static ExpPool &get() {
    static ExpPool p;
    return p;
  }
  void run(const double *q, size_t n, double dfac, double *out) {
    std::lock_guard<std::mutex> turn(submit_);
    if (nw_ == 0 || n < 4096) {
      for (size_t j = 0; j < n; j++) out[j] = exp(dfac * q[j]);
      return;
    }
    {
      std::lock_guard<std::mutex> l(m_);
      q_ = q;
      n_ = n;
      dfac_ = dfac;
      out_ = out;
      pending_ = nw_;
      gen_++;
    }
    cv_.notify_all();
    slice(nw_, q, n, dfac, out);  // the caller takes the last share
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return pending_ == 0; });
  }

 private:
  ExpPool() {
    const unsigned hc = std::thread::hardware_concurrency();
    nw_ = hc > 8 ? 7u : (hc > 1 ? hc - 1 : 0u);
    for (unsigned i = 0; i < nw_; i++) th_.emplace_back([this, i] { loop(i); });
  }
  ~ExpPool() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void slice(unsigned i, const double *q, size_t n, double dfac, double *out) const {
    const size_t parts = nw_ + 1, per = (n + parts - 1) / parts;
    const size_t b = i * per, e = std::min(n, b + per);
    for (size_t j = b; j < e; j++) out[j] = exp(dfac * q[j]);
  }
  void loop(unsigned i) {
    uint64_t seen = 0;
    for (;;) {
      const double *q;
      double *out;
      size_t n;
      double dfac;
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        q = q_;
        n = n_;
        dfac = dfac_;
        out = out_;
      }
      slice(i, q, n, dfac, out);
      std::lock_guard<std::mutex> l(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex submit_, m_;
  std::condition_variable cv_, done_;
  std::vector<std::thread> th_;
  unsigned nw_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  const double *q_ = nullptr;
  double *out_ = nullptr;
  size_t n_ = 0;
  double dfac_ = 0.0;
};

static void host_exp_table(const std::vector<double> &qu, double dfac, double *out) {
  ExpPool::get().run(qu.data(), qu.size(), dfac, out);
}

PW_KERNEL(k_interp, (double *__restrict__ out, const double *__restrict__ a,
                     const double *__restrict__ b, double pa, double pb, int64_t n),
          { out[k] = pa * a[k] + pb * b[k]; })

PW_KERNEL(k_sub, (double *__restrict__ out, const double *__restrict__ a,
                  const double *__restrict__ b, int64_t n),
          { out[k] = a[k] - b[k]; })

// savegrid's and savecont's inputs (evolution.f90:648-668, 735-778) at the
// start of a step, after its inversion: the PV and energies interpolated to
// tgrid from the kept qspre/ekinpre/epotpre and this inversion; the
// contours and the residual spctop(qq - qc).  Kept on the host for the
// caller, who writes the files; the record counters advance as the
// reference's do.  Main stream, before the second stream forks.
int save_slots(Ctx &c, bool grid, bool cont) {
  const int64_t N = c.N;
  const double t = c.hsc->t;
  if (grid) {
    Ctx::GridSave g;
    const double pt = (t - c.tgrid) / c.save_dt, ptc = 1.0 - pt;
    double *qqs;
    HYD_CHECK(dbuf(c, "qqs_slot", N, &qqs));
    hipLaunchKernelGGL(k_interp, grid1(N), dim3(256), 0, c.s, qqs, c.qspre, c.qq, pt, ptc, N);
    HYD_HIP(hipGetLastError());
    double ekinpost = 0.0, epotpost = 0.0;
    HYD_CHECK(diag_energy(c, &ekinpost, &epotpost));
    g.tgrid = c.tgrid;
    g.ekin = pt * c.ekinpre + ptc * ekinpost;
    g.epot = pt * c.epotpre + ptc * epotpost;
    g.igrec = c.igrec;
    g.qqs.resize(N);
    HYD_HIP(hipMemcpyAsync(g.qqs.data(), qqs, sizeof(double) * N, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    c.grid_saves.push_back(std::move(g));
    c.igrec += 1;
    c.gsave = 0;
  }
  if (cont) {
    Ctx::ContSave s;
    s.t = t;
    s.qjump = c.hsc->qjump;
    s.icrec = c.icrec;
    s.contex = c.contex ? 1 : 0;
    double *ss, *qa;
    HYD_CHECK(dbuf(c, "res_ss", N, &ss));
    HYD_CHECK(dbuf(c, "res_qa", N, &qa));
    hipLaunchKernelGGL(k_sub, grid1(N), dim3(256), 0, c.s, ss, c.qq, c.qc, N);
    HYD_HIP(hipGetLastError());
    HYD_CHECK(fft_spctop(c, ss, qa, PRE_NONE, 1.0, c.wc));
    s.res.resize(N);
    HYD_HIP(hipMemcpyAsync(s.res.data(), qa, sizeof(double) * N, hipMemcpyDeviceToHost, c.s));
    if (s.contex) {
      s.xq.resize(c.nptq);
      s.yq.resize(c.nptq);
      HYD_HIP(hipMemcpyAsync(s.xq.data(), c.xq, sizeof(double) * c.nptq, hipMemcpyDeviceToHost,
                             c.s));
      HYD_HIP(hipMemcpyAsync(s.yq.data(), c.yq, sizeof(double) * c.nptq, hipMemcpyDeviceToHost,
                             c.s));
      s.npq = c.npq;
      s.i1q = c.i1q;
      s.indq = c.indq;
    }
    HYD_HIP(hipStreamSynchronize(c.s));
    c.cont_saves.push_back(std::move(s));
    c.icrec += 1;
    c.csave = 0;
  }
  return HYD_OK;
}

// monitor.asc row of adapt (evolution.f90:581-584), from the scalars it set
void monitor_record(Ctx &c, const Scalars &s) {
  const double cfl = s.uumax * s.dt / c.c.glx;
  const double row[7] = {s.t, cfl, 0.5 * (s.zzrms * s.zzrms), s.zzrms, s.zzmax, s.uumax, s.twist};
  c.rec_monitor.insert(c.rec_monitor.end(), row, row + 7);
}

// evolution.f90:587-599: a grid save falls in this step when t + dt >= tgrid;
// then qs (reset at the step's start) and the energy of this inversion are
// kept for the interpolation to tgrid, which the caller makes after the
// step; a contour save falls in it when t + dt >= tcont.  Runs on the main
// stream before the second stream forks (the energy's sums use the shared
// sequential-sum work area).
int adapt_copy(Ctx &c);

static int adapt_saves(Ctx &c) {
  HYD_HIP(hipEventSynchronize(c.ev_sc));
  const double t = c.hsc->t, dt = c.hsc->dt;
  c.save_dt = dt;
  const double tgrid = c.tinit + c.p.tgsave * (double)c.igrec;
  c.gsave = t + dt >= tgrid ? 1 : 0;
  if (c.gsave) {
    c.tgrid = tgrid;
    HYD_CHECK(dbuf(c, "qspre", c.N, &c.qspre));
    HYD_HIP(hipMemcpyAsync(c.qspre, c.qs, sizeof(double) * c.N, hipMemcpyDeviceToDevice, c.s));
    HYD_CHECK(diag_energy(c, &c.ekinpre, &c.epotpre));
  }
  const double tcont = c.tinit + c.p.tcsave * (double)c.icrec;
  c.csave = t + dt >= tcont ? 1 : 0;
  return HYD_OK;
}

int adapt(Ctx &c) {
  const int64_t N = c.N;
  double *ss = c.wa, *zz = c.wb;
  const double *src = c.qq;
  if (!c.c.btropic) {
    hipLaunchKernelGGL(k_qq_kd, grid1(N), dim3(256), 0, c.s, ss, c.qq, c.pp, c.c.kdsq, N);
    src = ss;
  }
  HYD_CHECK(fft_spctop(c, src, zz, PRE_NONE, 1.0, c.wc));
  const int64_t nb = std::min<int64_t>((N + 255) / 256, 1024);
  {
    KTimer t(c, "adapt_reduce", 32 * N);
    hipLaunchKernelGGL(k_adapt_reduce, dim3((unsigned)nb), dim3(256), 0, c.s, zz, c.uu, c.vv, N,
                       c.amax);
  }
  HYD_CHECK(seqsum(c, zz, N, c.dsum + 1));
  AdaptArgs a;
  a.dsumi = c.c.dsumi;
  a.small = c.c.small;
  a.glx = c.c.glx;
  a.srwfm = c.c.srwfm;
  a.dt_fixed = c.p.dt_fixed;
  a.tgsave = c.p.tgsave;
  a.tcsave = c.p.tcsave;
  hipLaunchKernelGGL(k_adapt_scalars, dim3(1), dim3(64), 0, c.s, c.dsc, c.dsum + 1, c.amax,
                     (int)nb, a, c.flags + 8);
  HYD_HIP(hipGetLastError());
  // the integrating factors need dt and zzrms on the host: the copy is
  // started now (with saves, whose decision needs them at once) or at the
  // head of the second stream's work (adapt_copy), and the exponentials are
  // evaluated in adapt_finish while the GPU computes the sources of stage 1
  // (which do not use them)
  if (c.saves_on) {
    HYD_CHECK(adapt_copy(c));
    HYD_CHECK(adapt_saves(c));
  }
  return HYD_OK;
}

// adapt's scalars to the host mirror, on the current stream; ev_sc marks
// their arrival.
int adapt_copy(Ctx &c) {
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipEventRecord(c.ev_sc, c.s));
  return HYD_OK;
}

// integrating factors, evolution.f90:603-605: epq = exp(dt/2*zzrms*qdiss)
// evaluated on the host (the reference's libm exp) over the distinct qdiss
// values, expanded on the device; emq = 1/epq.
int adapt_finish(Ctx &c) {
  const int64_t N = c.N;
  HYD_HIP(hipEventSynchronize(c.ev_sc));
  monitor_record(c, *c.hsc);
  const double dfac = c.hsc->dt2 * c.hsc->zzrms;
  host_exp_table(c.ht.qdiss_u, dfac, c.exp_host);
  HYD_HIP(hipMemcpyAsync(c.qdiss_exp, c.exp_host, sizeof(double) * c.ht.qdiss_u.size(),
                         hipMemcpyHostToDevice, c.s));
  hipLaunchKernelGGL(k_expand_epq, grid1(N), dim3(256), 0, c.s, c.epq, c.emq, c.qdiss_map,
                     c.qdiss_exp, N);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

int source(Ctx &c, int lev) {
  const int64_t N = c.N;
  const int wb = c.c.beffect ? 1 : 0;
  // the four gradients (qd and qs, x and y) in one batch
  {
    const double *in[4] = {c.qd, c.qd, c.qs, c.qs};
    double *out[4] = {c.wa, c.wb, c.wc, c.wd};
    const int pre[4] = {PRE_XDERIV, PRE_YDERIV, PRE_XDERIV, PRE_YDERIV};
    const double sg[4] = {1.0, 1.0, 1.0, 1.0};
    HYD_CHECK(fft_spctop_n(c, 4, in, out, pre, sg, c.t4));
  }
  {
    KTimer t(c, "advect_prod", 80 * N);
    hipLaunchKernelGGL(k_advect_prod, grid1(N), dim3(256), 0, c.s, c.wa, c.wb, c.uu, c.vv,
                       c.p.beta, 0, N);
    hipLaunchKernelGGL(k_advect_prod, grid1(N), dim3(256), 0, c.s, c.wc, c.wd, c.uu, c.vv,
                       c.p.beta, wb, N);
  }
  {
    const double *in[2] = {c.wa, c.wc};
    double *out[2] = {c.sqd, c.sqs};
    HYD_CHECK(fft_ptospc_n(c, 2, in, out));
  }
  {
    KTimer t(c, "source_finish", 40 * N);
    hipLaunchKernelGGL(k_source_finish, grid1(N), dim3(256), 0, c.s, c.sqd, c.qq, c.pp, c.epq,
                       c.p.rekman, c.c.kdsq, lev, N);
  }
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static int rk_fields(Ctx &c, int stage) {
  const int64_t N = c.N;
  static const int64_t words[5] = {0, 11, 10, 12, 6};
  KTimer t(c, "rk_fields", words[stage] * 8 * N);
  hipLaunchKernelGGL(k_rk_fields, grid1(N), dim3(256), 0, c.s, c.qs, c.qd, c.qsi, c.qsf, c.qdi,
                     c.qdf, c.sqs, c.sqd, c.emq, stage, c.dsc, N);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

// Advance the host mirror of t by dt/2 (t=t+dt2, evolution.f90:237, 295).
__global__ void k_tick(Scalars *sc) {
  if (threadIdx.x == 0) sc->t = sc->t + sc->dt2;
}

// One RK4 step (evolution.f90:166-326).  Stage k's source and field update
// (which need the velocity of stage k) run on the second stream while the
// main stream moves the nodes and builds the contour grid of stage k+1 (which
// needs only the moved nodes); the field part of stage k+1's inversion waits
// for the update.  Every kernel sees the same inputs as in the reference's
// serial order.
// Runs fn on the second stream (c.s points at it meanwhile, so that every
// launch and timer inside goes there).
template <class Fn>
static int on_s2(Ctx &c, Fn fn) {
  hipStream_t main = c.s;
  // with per-launch timing on, everything stays on the main stream, in
  // program order (a valid order of the same work), so that each timed
  // launch runs alone
  c.s = c.timing ? main : c.s2;
  const int rc = fn();
  c.s = main;
  return rc;
}

// A contour grid whose sweep leaves its clearing to the caller (who runs
// it on the second stream, off the main stream's path).
static int contour_grid_deferred(Ctx &c) {
  c.defer_clear = true;
  const int rc = contour_grid(c);
  c.defer_clear = false;
  return rc;
}

// Fork: the second stream continues after everything enqueued so far on
// the main stream.
static int fork_s2(Ctx &c) {
  HYD_HIP(hipEventRecord(c.ev_fork, c.s));
  HYD_HIP(hipStreamWaitEvent(c.s2, c.ev_fork, 0));
  return HYD_OK;
}

int advance(Ctx &c) {
  const int64_t N = c.N;
  c.after_sweep = nullptr;
  c.n_advances += 1;
  c.n_node_updates += c.nptq;
  auto fields = [&](int stage) {
    return on_s2(c, [&]() {
      HYD_CHECK(rk_fields(c, stage));
      HYD_HIP(hipEventRecord(c.ev_join, c.s));
      return HYD_OK;
    });
  };
  // stage k's source and field update, started once the next contour
  // grid's sweep is queued (con2grid runs c.after_sweep)
  auto source_fields_later = [&](int stage) {
    c.after_sweep = [&c, stage]() {
      HYD_CHECK(fork_s2(c));
      return on_s2(c, [&]() {
        HYD_CHECK(sweep_clear(c));  // the sweep just queued
        HYD_CHECK(source(c, 1));
        HYD_CHECK(rk_fields(c, stage));
        HYD_HIP(hipEventRecord(c.ev_join, c.s));
        return HYD_OK;
      });
    };
  };
  // stage 1 (its sweep's clearing runs on the second stream after adapt)
  HYD_CHECK(contour_grid_deferred(c));
  HYD_CHECK(inversion_fields(c));
  hipLaunchKernelGGL(k_reset, grid1(N), dim3(256), 0, c.s, c.qs, c.qd, c.qc, c.flo, c.fhi, N);
  // saves that fall at this step's start (set by the previous adapt)
  if (c.saves_on && (c.gsave || c.csave)) HYD_CHECK(save_slots(c, c.gsave != 0, c.csave != 0));
  HYD_CHECK(adapt(c));
  HYD_CHECK(fork_s2(c));
  HYD_CHECK(on_s2(c, [&]() {
    if (!c.saves_on) HYD_CHECK(adapt_copy(c));  // off the main stream's path
    HYD_CHECK(sweep_clear(c));
    return source(c, 0);
  }));
  HYD_CHECK(rk_nodes(c, 1));
  // stage 2's contour grid needs only the stage-1 nodes: it is queued
  // before the host evaluates the integrating factors, so that the GPU has
  // it to run meanwhile; the factors are uploaded and expanded on the
  // second stream, ahead of the stage-1 field update that needs them
  // (the tick changes t: the scalars' copy on the second stream has read them)
  if (!c.saves_on) HYD_HIP(hipStreamWaitEvent(c.s, c.ev_sc, 0));
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, c.s, c.dsc);
  HYD_CHECK(contour_grid(c));
  HYD_CHECK(on_s2(c, [&]() { return adapt_finish(c); }));
  // an overflow of an earlier sweep (the scalars are on the host now):
  // reported, and the table reset, behind the main stream's queued sweep
  if (c.hsc->sweep_err != 0.0) return c2g_status(c);
  HYD_CHECK(fields(1));
  // stages 2-4
  for (int stage = 2; stage <= 4; stage++) {
    if (stage > 2) {
      if (stage != 3) hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, c.s, c.dsc);
      HYD_CHECK(contour_grid_deferred(c));  // clearing in the after-sweep work
    }
    HYD_HIP(hipStreamWaitEvent(c.s, c.ev_join, 0));
    HYD_CHECK(inversion_fields(c));
    if (stage < 4) {
      source_fields_later(stage);
      HYD_CHECK(rk_nodes(c, stage));
    } else {
      HYD_CHECK(fork_s2(c));
      HYD_CHECK(on_s2(c, [&]() { return source(c, 2); }));
      HYD_CHECK(rk_nodes(c, stage));
      HYD_CHECK(fields(stage));
    }
  }
  c.after_sweep = nullptr;
  HYD_HIP(hipStreamWaitEvent(c.s, c.ev_join, 0));
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

// qs = flo*(qs-qc)+qc+qd ; qd = qs-qc  (evolution.f90:155-156)
PW_KERNEL(k_prepare, (double *__restrict__ qs, double *__restrict__ qd,
                      const double *__restrict__ qc, const double *__restrict__ flo, int64_t n),
          {
            const double s = flo[k] * (qs[k] - qc[k]) + qc[k] + qd[k];
            qs[k] = s;
            qd[k] = s - qc[k];
          })

// evolution.f90:135-162 (prepare): qr = spctop(qs - qc) for recontouring.
int prepare(Ctx &c) {
  const int64_t N = c.N;
  HYD_CHECK(con2grid(c, c.xq, c.yq, c.nextq, c.nptq, c.wa, nullptr));
  HYD_CHECK(fft_ptospc(c, c.wa, c.qc));
  hipLaunchKernelGGL(k_prepare, grid1(N), dim3(256), 0, c.s, c.qs, c.qd, c.qc, c.flo, N);
  HYD_HIP(hipGetLastError());
  HYD_CHECK(fft_spctop(c, c.qd, c.qr, PRE_NONE, 1.0, c.wc));
  return HYD_OK;
}

// evolution.f90:96-131 (init): qc from the contours and qd = fhi*(qs-qc).
int evolution_init(Ctx &c) {
  const int64_t N = c.N;
  c.contex = c.nptq > 0;
  HYD_CHECK(con2grid(c, c.xq, c.yq, c.nextq, c.nptq, c.wa, nullptr));
  HYD_CHECK(fft_ptospc(c, c.wa, c.qc));
  hipLaunchKernelGGL(k_init_qd, grid1(N), dim3(256), 0, c.s, c.qd, c.qs, c.qc, c.fhi, N);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

// ---- mixed precision --------------------------------------------------------
// FP32 physical fields and transforms, FP64 spectral state, node positions and
// RK accumulation.  No host round trip: the integrating factors use the
// device exp, the sums are parallel reductions, so the step can be captured
// in a graph.

// qqx = -uu*qqx - vv*qqy in FP32
PW_KERNEL(k_advect_prod32, (float *__restrict__ qqx, const float *__restrict__ qqy,
                            const float *__restrict__ uu, const float *__restrict__ vv,
                            float beta, int withbeta, int64_t n),
          {
            if (withbeta)
              qqx[k] = -uu[k] * qqx[k] - vv[k] * (qqy[k] + beta);
            else
              qqx[k] = -uu[k] * qqx[k] - vv[k] * qqy[k];
          })

// sum of zz^2 per block (FP64 accumulation) and the maxima of zz^2, uu^2+vv^2
// (one pair per block in mx[2 b], mx[2 b + 1])
__global__ void __launch_bounds__(256) k_adapt_reduce32(const float *__restrict__ zz,
                                                        const float *__restrict__ uu,
                                                        const float *__restrict__ vv, int64_t n,
                                                        double *__restrict__ part,
                                                        unsigned long long *__restrict__ mx) {
  double s = 0.0, z2m = 0.0, u2m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double z = (double)zz[k];
    const double z2 = z * z;
    s += z2;
    const double u = (double)uu[k], v = (double)vv[k];
    z2m = fmax(z2m, z2);
    u2m = fmax(u2m, u * u + v * v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    z2m = fmax(z2m, __shfl_down(z2m, off));
    u2m = fmax(u2m, __shfl_down(u2m, off));
  }
  __shared__ double ss[4], sz[4], su[4];
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s;
    sz[threadIdx.x >> 6] = z2m;
    su[threadIdx.x >> 6] = u2m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      s += ss[w];
      z2m = fmax(z2m, sz[w]);
      u2m = fmax(u2m, su[w]);
    }
    part[blockIdx.x] = s;
    mx[2 * blockIdx.x] = (unsigned long long)__double_as_longlong(z2m);
    mx[2 * blockIdx.x + 1] = (unsigned long long)__double_as_longlong(u2m);
  }
}

// total of 1024 partial sums (one workgroup, fixed-order tree)
__global__ void __launch_bounds__(1024) k_sum_parts(const double *part, double *out) {
  __shared__ double s[1024];
  const int t = threadIdx.x;
  s[t] = part[t];
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) s[t] += s[t + o];
    __syncthreads();
  }
  if (t == 0) *out = s[0];
}

// exp(dt/2*zzrms*qdiss) for the distinct qdiss values, on the device
__global__ void k_exp_table(const double *__restrict__ qu, int nqu, const Scalars *sc,
                            double *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nqu) return;
  out[j] = exp(sc->dt2 * sc->zzrms * qu[j]);
}

static int adapt_mixed(Ctx &c) {
  const int64_t N = c.N;
  const double *src = c.qq;
  if (!c.c.btropic) {
    hipLaunchKernelGGL(k_qq_kd, grid1(N), dim3(256), 0, c.s, c.wa, c.qq, c.pp, c.c.kdsq, N);
    src = c.wa;
  }
  HYD_CHECK(fft_spctop32(c, src, c.fc32, PRE_NONE, 1.0));
  hipLaunchKernelGGL(k_adapt_reduce32, dim3(1024), dim3(256), 0, c.s, c.fc32, c.uu32, c.vv32, N,
                     c.part, c.amax);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, c.s, c.part, c.dsum + 1);
  AdaptArgs a;
  a.dsumi = c.c.dsumi;
  a.small = c.c.small;
  a.glx = c.c.glx;
  a.srwfm = c.c.srwfm;
  a.dt_fixed = c.p.dt_fixed;
  a.tgsave = c.p.tgsave;
  a.tcsave = c.p.tcsave;
  hipLaunchKernelGGL(k_adapt_scalars, dim3(1), dim3(64), 0, c.s, c.dsc, c.dsum + 1, c.amax, 1024,
                     a, nullptr);
  hipLaunchKernelGGL(k_exp_table, dim3((c.nqu + 255) / 256), dim3(256), 0, c.s, c.qdiss_u_dev,
                     c.nqu, c.dsc, c.qdiss_exp);
  hipLaunchKernelGGL(k_expand_epq, grid1(N), dim3(256), 0, c.s, c.epq, c.emq, c.qdiss_map,
                     c.qdiss_exp, N);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static int source_mixed(Ctx &c, int lev) {
  const int64_t N = c.N;
  const int wb = c.c.beffect ? 1 : 0;
  {
    const double *in[4] = {c.qd, c.qd, c.qs, c.qs};
    float *out[4] = {c.fa32, c.fb32, c.fd32, c.fe32};
    const int pre[4] = {PRE_XDERIV, PRE_YDERIV, PRE_XDERIV, PRE_YDERIV};
    const double sg[4] = {1.0, 1.0, 1.0, 1.0};
    HYD_CHECK(fft_spctop32_n(c, 4, in, out, pre, sg, c.t32));
  }
  hipLaunchKernelGGL(k_advect_prod32, grid1(N), dim3(256), 0, c.s, c.fa32, c.fb32, c.uu32,
                     c.vv32, (float)c.p.beta, 0, N);
  hipLaunchKernelGGL(k_advect_prod32, grid1(N), dim3(256), 0, c.s, c.fd32, c.fe32, c.uu32,
                     c.vv32, (float)c.p.beta, wb, N);
  {
    const float *in[2] = {c.fa32, c.fd32};
    double *out[2] = {c.sqd, c.sqs};
    HYD_CHECK(fft_ptospc32_n(c, 2, in, out, c.t32));
  }
  hipLaunchKernelGGL(k_source_finish, grid1(N), dim3(256), 0, c.s, c.sqd, c.qq, c.pp, c.epq,
                     c.p.rekman, c.c.kdsq, lev, N);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

static int contour_grid_mixed(Ctx &c) {
  HYD_CHECK(con2grid_mixed(c, nullptr, true));
  return fft_ptospc32_mean(c, c.fc32, c.part + 1024, c.qc);
}

static int inversion_fields_mixed(Ctx &c) {
  const int64_t N = c.N;
  hipLaunchKernelGGL(k_combine_qq, grid1(N), dim3(256), 0, c.s, c.qq, c.flo, c.qs, c.qc, c.qd, N);
  hipLaunchKernelGGL(k_green, grid1(N), dim3(256), 0, c.s, c.pp, c.green, c.qq, N);
  HYD_HIP(hipGetLastError());
  {
    const double *in[2] = {c.pp, c.pp};
    float *out[2] = {c.vv32, c.uu32};
    const int pre[2] = {PRE_XDERIV, PRE_YDERIV};
    const double sg[2] = {1.0, -1.0};
    HYD_CHECK(fft_spctop32_n(c, 2, in, out, pre, sg, c.t32));
  }
  return HYD_OK;
}

// The same stage overlap as the plain step (advance); the fork and join are
// events, so a stream capture of the main stream records both streams.
// The mixed step in two parts: the head (stage 1's inversion and adapt,
// after which the step's scalars are final but for t) and the tail (the
// four stages), so that a loop can read the scalars back between the two
// and decide the next step while the tail runs.
int advance_mixed_head(Ctx &c) {
  const int64_t N = c.N;
  HYD_CHECK(contour_grid_mixed(c));
  HYD_CHECK(inversion_fields_mixed(c));
  hipLaunchKernelGGL(k_reset, grid1(N), dim3(256), 0, c.s, c.qs, c.qd, c.qc, c.flo, c.fhi, N);
  return adapt_mixed(c);
}

int advance_mixed(Ctx &c) {
  HYD_CHECK(advance_mixed_head(c));
  return advance_mixed_tail(c);
}

int advance_mixed_tail(Ctx &c) {
  auto fields = [&](int stage) {
    return on_s2(c, [&]() {
      HYD_CHECK(rk_fields(c, stage));
      HYD_HIP(hipEventRecord(c.ev_join, c.s));
      return HYD_OK;
    });
  };
  for (int stage = 1; stage <= 4; stage++) {
    if (stage > 1) {
      if (stage != 3) hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, c.s, c.dsc);
      HYD_CHECK(contour_grid_mixed(c));
      HYD_HIP(hipStreamWaitEvent(c.s, c.ev_join, 0));
      HYD_CHECK(inversion_fields_mixed(c));
    }
    HYD_CHECK(fork_s2(c));
    const int lev = stage == 1 ? 0 : (stage == 4 ? 2 : 1);
    HYD_CHECK(on_s2(c, [&]() { return source_mixed(c, lev); }));
    HYD_CHECK(rk_nodes(c, stage));
    HYD_CHECK(fields(stage));
  }
  HYD_HIP(hipStreamWaitEvent(c.s, c.ev_join, 0));
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

}  // namespace hyd
