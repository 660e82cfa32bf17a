// capi.hip -- the extern "C" boundary (include/hydra_amd.h): context life
// cycle, host-array entry points that mirror the reference routines one to
// one, and the device-resident step.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>

#include <hip/hip_ext.h>

#include "hyd_internal.h"

namespace hyd {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

int inversion(Ctx &c);

KTimer::KTimer(Ctx &ctx, const char *name, int64_t bytes, bool disp) : c(ctx), dispatch(disp) {
  if (!c.timing || (c.timing == 1 && strncmp(name, "k:", 2) != 0)) return;
  Ctx::KRec r;
  r.name = name;
  r.bytes = bytes;
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
  if (!dispatch) hipEventRecord(r.a, c.s);
  c.krec.push_back(r);
  idx = (int)c.krec.size() - 1;
}
KTimer::~KTimer() {
  if (idx >= 0 && !dispatch) hipEventRecord(c.krec[idx].b, c.s);
}
hipEvent_t KTimer::start() const { return idx >= 0 ? c.krec[idx].a : nullptr; }
hipEvent_t KTimer::stop() const { return idx >= 0 ? c.krec[idx].b : nullptr; }

__global__ void k_xderiv(const double *__restrict__ var, double *__restrict__ der,
                         const double *__restrict__ dk, int nx, int ny) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based
  const int ky = blockIdx.y;
  if (kx >= nx) return;
  const int h = nx / 2;
  double v;
  if (kx == 0 || kx == h) v = 0.0;
  else if (kx < h) v = -dk[2 * kx - 1] * var[(int64_t)ky * nx + (nx - kx)];
  else v = dk[2 * (nx - kx) - 1] * var[(int64_t)ky * nx + (nx - kx)];
  der[(int64_t)ky * nx + kx] = v;
}

__global__ void k_yderiv(const double *__restrict__ var, double *__restrict__ der,
                         const double *__restrict__ dk, int nx, int ny) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ky = blockIdx.y;
  if (kx >= nx) return;
  const int h = ny / 2;
  double v;
  if (ky == 0 || ky == h) v = 0.0;
  else if (ky < h) v = -dk[2 * ky - 1] * var[(int64_t)(ny - ky) * nx + kx];
  else v = dk[2 * (ny - ky) - 1] * var[(int64_t)(ny - ky) * nx + kx];
  der[(int64_t)ky * nx + kx] = v;
}

int launch_xderiv(Ctx &c, const double *var, double *der) {
  hipLaunchKernelGGL(k_xderiv, dim3((c.c.nx + 255) / 256, c.c.ny), dim3(256), 0, c.s, var, der,
                     c.hrkx, c.c.nx, c.c.ny);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}
int launch_yderiv(Ctx &c, const double *var, double *der) {
  hipLaunchKernelGGL(k_yderiv, dim3((c.c.nx + 255) / 256, c.c.ny), dim3(256), 0, c.s, var, der,
                     c.hrky, c.c.nx, c.c.ny);
  HYD_HIP(hipGetLastError());
  return HYD_OK;
}

template <class T>
static int dalloc(T **p, int64_t n) {
  HYD_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
  HYD_HIP(hipMemset(*p, 0, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
  return HYD_OK;
}

template <class T>
static int upload(T *d, const T *h, int64_t n, hipStream_t s) {
  HYD_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
  return HYD_OK;
}

static void free_ctx(Ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->green, c->flo, c->fhi, c->hrkx, c->hrky, c->xtrig, c->ytrig, c->xgf,
                  c->qdiss_map, c->qdiss_exp, c->qs, c->qd, c->qc, c->qq, c->pp, c->uu, c->vv,
                  c->qr, c->emq, c->epq, c->qsi, c->qsf, c->sqs, c->qdi, c->qdf, c->sqd, c->wa,
                  c->wb, c->wc, c->wd, c->cnt, c->fq, c->mq, c->qjx_cnt, c->qbot,
                  c->flags, c->replay, c->replay_val, c->rp_tkey,
                  c->rp_tid, c->rp_list, c->xq, c->yq, c->xqi, c->yqi, c->xqf, c->yqf, c->uq,
                  c->vq, c->nextq, c->nxt_tmp, c->dsc, c->dsum, c->seq_work, c->uu32,
                  c->vv32, c->fa32, c->fb32, c->fc32, c->tmp32, c->fq32, c->mq32, c->xtrig32,
                  c->ytrig32, c->qb32, c->qdiss_u_dev, c->part, c->c2g_q, c->c2g_w, c->c2g_q32,
                  c->amax};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (int k = 0; k < 4; k++) {
    if (c->t4[k]) hipFree(c->t4[k]);
    if (c->t32[k]) hipFree(c->t32[k]);
  }
  if (c->fd32) hipFree(c->fd32);
  if (c->fe32) hipFree(c->fe32);
  if (c->hsc) hipHostFree(c->hsc);
  if (c->exp_host) hipHostFree(c->exp_host);
  if (c->ev_sc) hipEventDestroy(c->ev_sc);
  for (hipEvent_t e : {c->ev_fork, c->ev_join, c->ev_cleared})
    if (e) hipEventDestroy(e);
  free_congen(*c);
  diag_free(*c);
  for (auto &r : c->krec) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  if (c->step_graph) hipGraphExecDestroy(c->step_graph);
  if (c->step_graph_tail) hipGraphExecDestroy(c->step_graph_tail);
  if (c->s2) hipStreamDestroy(c->s2);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
}

static int create(const hyd_params *p, hyd_ctx **out) {
  if (!p || !out) {
    set_error("null argument");
    return HYD_EARG;
  }
  Ctx *c = new Ctx();
  c->p = *p;
  int rc = build_consts(*p, c->c);
  if (rc == HYD_OK) rc = build_tables(*p, c->c, c->ht);
  if (rc != HYD_OK) {
    delete c;
    return rc;
  }
  const int n = p->ng;
  int log2n = 0;
  while ((1 << log2n) < n) log2n++;
  c->plan.n = n;
  c->plan.n2 = log2n & 1;
  c->plan.n4 = log2n / 2;
  c->plan.fnorm = 1.0 / sqrt((double)n);
  c->plan.rnorm = 2.0 / sqrt((double)n);
  c->N = (int64_t)n * n;
  c->cap = p->node_capacity > 0 ? p->node_capacity : std::max<int64_t>(64 * c->N, 1 << 20);
  if (c->cap > INT32_MAX / 2) {  // node numbers are int32 (nextq)
    if (p->node_capacity > 0) {
      set_error("node_capacity exceeds the int32 node numbering");
      delete c;
      return HYD_EARG;
    }
    c->cap = INT32_MAX / 2;
  }
  if (p->precision != 0 && p->precision != 1) {
    set_error("precision must be 0 (plain FP64) or 1 (mixed FP32 grid / FP64 state)");
    delete c;
    return HYD_EARG;
  }
  c->dev = p->device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_error("no HIP device available: the hydra_amd library runs on the GPU only");
    delete c;
    return HYD_EHIP;
  }
  if (hipSetDevice(c->dev) != hipSuccess) {
    set_error("hipSetDevice failed");
    delete c;
    return HYD_EHIP;
  }
#define TRY(x)             \
  do {                     \
    int r_ = (x);          \
    if (r_ != HYD_OK) {    \
      free_ctx(c);         \
      return r_;           \
    }                      \
  } while (0)
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete c;
    return HYD_EHIP;
  }
  const int64_t N = c->N;
  double **fields[] = {&c->qs,  &c->qd,  &c->qc,  &c->qq,  &c->pp,  &c->uu, &c->vv,
                       &c->qr,  &c->emq, &c->epq, &c->qsi, &c->qsf, &c->sqs, &c->qdi,
                       &c->qdf, &c->sqd, &c->wa,  &c->wb,  &c->wc,  &c->wd, &c->c2g_q, &c->c2g_w};
  for (double **f : fields) TRY(dalloc(f, N));
  for (int k = 0; k < 4; k++) TRY(dalloc(&c->t4[k], N));
  TRY(dalloc(&c->green, N));
  TRY(dalloc(&c->flo, N));
  TRY(dalloc(&c->fhi, N));
  TRY(dalloc(&c->hrkx, n));
  TRY(dalloc(&c->hrky, n));
  TRY(dalloc(&c->xtrig, 2 * n));
  TRY(dalloc(&c->ytrig, 2 * n));
  TRY(dalloc(&c->xgf, c->c.nxf));
  TRY(dalloc(&c->qdiss_map, N));
  c->nqu = (int)c->ht.qdiss_u.size();
  TRY(dalloc(&c->qdiss_exp, c->nqu));
  const int64_t nfine = (int64_t)c->c.nxf * c->c.nyf;
  TRY(dalloc(&c->cnt, nfine));
  TRY(dalloc(&c->fq, nfine));
  TRY(dalloc(&c->mq, nfine / 4));
  TRY(dalloc(&c->qjx_cnt, c->c.nxf));
  TRY(dalloc(&c->qbot, c->c.nxf));
  TRY(dalloc(&c->flags, 16));
  c->replay_cap = 1 << 22;
  c->rp_bcap = 1 << 20;
  c->rp_tsize = 1 << 21;
  TRY(dalloc(&c->replay, c->replay_cap));
  TRY(dalloc(&c->replay_val, c->rp_bcap));
  TRY(dalloc(&c->rp_list, 4 * c->rp_bcap));
  TRY(dalloc(&c->rp_tid, c->rp_tsize));
  TRY(dalloc(&c->rp_tkey, c->rp_tsize));
  if (hipMemset(c->rp_tkey, 0xff, sizeof(int32_t) * c->rp_tsize) != hipSuccess ||
      hipMemset(c->flags, 0, sizeof(int32_t) * 16) != hipSuccess) {
    set_error("hipMemset failed");
    free_ctx(c);
    return HYD_EHIP;
  }
  const int64_t P = c->cap;
  double **nodes[] = {&c->xq, &c->yq, &c->xqi, &c->yqi, &c->xqf, &c->yqf, &c->uq, &c->vq};
  for (double **f : nodes) TRY(dalloc(f, P));
  TRY(dalloc(&c->nextq, P));
  TRY(dalloc(&c->nxt_tmp, P));
  TRY(dalloc(&c->dsc, 1));
  TRY(dalloc(&c->dsum, 8));
  TRY(dalloc(&c->amax, 2048));
  c->seq_cap = (std::max<int64_t>(N, 4 * N) + 255) / 256 + 64;
  {
    void *w = nullptr;
    if (hipMalloc(&w, seqsum_work_bytes(4 * N)) != hipSuccess) {
      set_error("hipMalloc seqsum work");
      free_ctx(c);
      return HYD_ENOMEM;
    }
    c->seq_work = w;
    c->seq_cap = seqsum_work_chunks(4 * N);
  }
  if (hipHostMalloc((void **)&c->hsc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    free_ctx(c);
    return HYD_ENOMEM;
  }
  memset(c->hsc, 0, sizeof(Scalars));
  if (hipHostMalloc((void **)&c->exp_host, sizeof(double) * std::max<size_t>(c->ht.qdiss_u.size(), 1),
                    hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_sc, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_cleared, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipHostMalloc / hipEventCreate / hipStreamCreate failed");
    free_ctx(c);
    return HYD_ENOMEM;
  }
  TRY(upload(c->green, c->ht.green.data(), N, c->s));
  TRY(upload(c->flo, c->ht.flo.data(), N, c->s));
  TRY(upload(c->fhi, c->ht.fhi.data(), N, c->s));
  TRY(upload(c->hrkx, c->ht.hrkx.data(), n, c->s));
  TRY(upload(c->hrky, c->ht.hrky.data(), n, c->s));
  TRY(upload(c->xtrig, c->ht.xtrig.data(), 2 * n, c->s));
  TRY(upload(c->ytrig, c->ht.ytrig.data(), 2 * n, c->s));
  TRY(upload(c->qdiss_map, c->ht.qdiss_map.data(), N, c->s));
  {
    std::vector<double> xgf(c->c.nxf);
    for (int ix = 1; ix <= c->c.nxf; ix++) xgf[ix - 1] = c->c.xmin + c->c.glxf * (double)(ix - 1);
    TRY(upload(c->xgf, xgf.data(), c->c.nxf, c->s));
    if (hipStreamSynchronize(c->s) != hipSuccess) {
      set_error("stream sync failed");
      free_ctx(c);
      return HYD_EHIP;
    }
  }
  if (p->precision == 1) {
    float **f32[] = {&c->uu32, &c->vv32, &c->fa32,   &c->fb32,   &c->fc32,  &c->tmp32,
                     &c->fd32, &c->fe32, &c->t32[0], &c->t32[1], &c->t32[2], &c->t32[3],
                     &c->c2g_q32};
    for (float **f : f32) TRY(dalloc(f, N));
    TRY(dalloc(&c->fq32, nfine));
    TRY(dalloc(&c->mq32, nfine / 4));
    TRY(dalloc(&c->qb32, c->c.nxf));
    TRY(dalloc(&c->xtrig32, 2 * n));
    TRY(dalloc(&c->ytrig32, 2 * n));
    TRY(dalloc(&c->qdiss_u_dev, c->nqu));
    TRY(dalloc(&c->part, 1024 + 8));
    std::vector<float> tx(2 * n), ty(2 * n);
    for (int i = 0; i < 2 * n; i++) {
      tx[i] = (float)c->ht.xtrig[i];
      ty[i] = (float)c->ht.ytrig[i];
    }
    TRY(upload(c->xtrig32, tx.data(), 2 * n, c->s));
    TRY(upload(c->ytrig32, ty.data(), 2 * n, c->s));
    TRY(upload(c->qdiss_u_dev, c->ht.qdiss_u.data(), c->nqu, c->s));
    if (hipStreamSynchronize(c->s) != hipSuccess) {
      set_error("stream sync failed");
      free_ctx(c);
      return HYD_EHIP;
    }
  }
  TRY(fft_autotune(*c));
#undef TRY
  *out = (hyd_ctx *)c;
  return HYD_OK;
}

static Ctx *C(hyd_ctx *h) { return (Ctx *)h; }

// Staging helpers for the host-array entry points.
static int h2d(Ctx &c, double *d, const double *h, int64_t n) {
  HYD_HIP(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c.s));
  return HYD_OK;
}
static int d2h(Ctx &c, double *h, const double *d, int64_t n) {
  HYD_HIP(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

// Surgery and recontouring with their wall time (each ends with the stream
// drained) accumulated per context: the regularisation costs that a run
// amortises over its steps.
static double wall_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static int timed_surgery(Ctx &c) {
  const double t0 = wall_now();
  HYD_CHECK(surgery(c));
  HYD_HIP(hipStreamSynchronize(c.s));
  c.phase_n[0] += 1;
  c.phase_s[0] += wall_now() - t0;
  return HYD_OK;
}

// complexity.asc row (evolution.f90:72, 111): t, nq, nptq
static void complexity_record(Ctx &c) {
  const double row[3] = {c.hsc->t, (double)c.nq, (double)c.nptq};
  c.rec_complexity.insert(c.rec_complexity.end(), row, row + 3);
}

static int timed_recontour(Ctx &c) {
  const double t0 = wall_now();
  HYD_CHECK(recontour(c, c.qr));
  HYD_HIP(hipStreamSynchronize(c.s));
  c.phase_n[1] += 1;
  c.phase_s[1] += wall_now() - t0;
  return HYD_OK;
}

}  // namespace hyd

using namespace hyd;

extern "C" {

const char *hyd_last_error(void) { return g_err.c_str(); }

int hyd_set_saves(hyd_ctx *h, int32_t on, double tinit, int32_t igrec, int32_t icrec) {
  Ctx &c = *C(h);
  if (on && c.p.precision != 0) {
    set_error("set_saves: the save schedule needs the plain precision (precision 0)");
    return HYD_ESTATE;
  }
  if (igrec < 0 || icrec < 1) {
    set_error("set_saves: igrec >= 0 and icrec >= 1");
    return HYD_EARG;
  }
  c.saves_on = on != 0;
  c.tinit = tinit;
  c.igrec = igrec;
  c.icrec = icrec;
  c.gsave = c.csave = 0;
  return HYD_OK;
}

int hyd_save_counters(hyd_ctx *h, int32_t *igrec, int32_t *icrec) {
  Ctx &c = *C(h);
  if (!igrec || !icrec) {
    set_error("save_counters: null pointer");
    return HYD_EARG;
  }
  *igrec = c.igrec;
  *icrec = c.icrec;
  return HYD_OK;
}

int hyd_pending_saves(hyd_ctx *h, int32_t *grid, int32_t *cont) {
  Ctx &c = *C(h);
  if (!grid || !cont) {
    set_error("pending_saves: null pointer");
    return HYD_EARG;
  }
  *grid = (int32_t)c.grid_saves.size();
  *cont = (int32_t)c.cont_saves.size();
  return HYD_OK;
}

int hyd_take_grid_save(hyd_ctx *h, double *info, double *qqs) {
  Ctx &c = *C(h);
  if (c.grid_saves.empty()) {
    set_error("take_grid_save: no grid save waits");
    return HYD_ESTATE;
  }
  if (!info || !qqs) {
    set_error("take_grid_save: null pointer");
    return HYD_EARG;
  }
  const Ctx::GridSave &g = c.grid_saves.front();
  info[0] = g.tgrid;
  info[1] = g.ekin;
  info[2] = g.epot;
  info[3] = g.igrec;
  std::copy(g.qqs.begin(), g.qqs.end(), qqs);
  c.grid_saves.erase(c.grid_saves.begin());
  return HYD_OK;
}

int hyd_contour_save_size(hyd_ctx *h, int32_t *nq, int32_t *nptq) {
  Ctx &c = *C(h);
  if (c.cont_saves.empty()) {
    set_error("contour_save_size: no contour save waits");
    return HYD_ESTATE;
  }
  if (!nq || !nptq) {
    set_error("contour_save_size: null pointer");
    return HYD_EARG;
  }
  *nq = (int32_t)c.cont_saves.front().npq.size();
  *nptq = (int32_t)c.cont_saves.front().xq.size();
  return HYD_OK;
}

int hyd_take_contour_save(hyd_ctx *h, double *info, double *residual, double *xq, double *yq,
                          int32_t *npq, int32_t *i1q, int32_t *indq) {
  Ctx &c = *C(h);
  if (c.cont_saves.empty()) {
    set_error("take_contour_save: no contour save waits");
    return HYD_ESTATE;
  }
  const Ctx::ContSave &s = c.cont_saves.front();
  if (!info || !residual || (!s.xq.empty() && (!xq || !yq)) ||
      (!s.npq.empty() && (!npq || !i1q || !indq))) {
    set_error("take_contour_save: null pointer");
    return HYD_EARG;
  }
  info[0] = s.t;
  info[1] = s.qjump;
  info[2] = s.icrec;
  info[3] = s.contex;
  std::copy(s.res.begin(), s.res.end(), residual);
  std::copy(s.xq.begin(), s.xq.end(), xq);
  std::copy(s.yq.begin(), s.yq.end(), yq);
  std::copy(s.npq.begin(), s.npq.end(), npq);
  std::copy(s.i1q.begin(), s.i1q.end(), i1q);
  std::copy(s.indq.begin(), s.indq.end(), indq);
  c.cont_saves.erase(c.cont_saves.begin());
  return HYD_OK;
}

// The end of advect (evolution.f90:85-89): saves at tsim itself, after an
// inversion of the current state.
int hyd_save_now(hyd_ctx *h, int32_t grid, int32_t cont) {
  Ctx &c = *C(h);
  if (!c.saves_on || !c.have_field) {
    set_error("save_now: needs a field and the save schedule (hyd_set_saves)");
    return HYD_ESTATE;
  }
  if (grid && !c.qspre) {
    set_error("save_now: no step has kept qs for a grid save yet");
    return HYD_ESTATE;
  }
  if (!grid && !cont) return HYD_OK;
  HYD_CHECK(inversion(c));
  if (grid) c.tgrid = c.tinit + c.p.tgsave * (double)c.igrec;
  HYD_CHECK(save_slots(c, grid != 0, cont != 0));
  return HYD_OK;
}

int hyd_take_records(hyd_ctx *h, int32_t kind, int32_t max, double *out) {
  Ctx &c = *C(h);
  std::vector<double> *v = kind == 0 ? &c.rec_monitor : (kind == 1 ? &c.rec_complexity : nullptr);
  if (!v || max < 0 || (max > 0 && !out)) {
    set_error("take_records: kind 0 (monitor) or 1 (complexity), max >= 0");
    return HYD_EARG;
  }
  const int w = kind == 0 ? 7 : 3;
  const int64_t have = (int64_t)v->size() / w;
  const int64_t n = std::min<int64_t>(have, max);
  if (n > 0) {
    std::copy(v->begin(), v->begin() + n * w, out);
    v->erase(v->begin(), v->begin() + n * w);
  }
  return (int)n;
}

int hyd_set_correction(hyd_ctx *h, hyd_cnn *net, double su, double sq, double sd) {
  Ctx &c = *C(h);
  if (net && (cnn_grid(net) != c.c.nx || !(su > 0.0) || !(sq > 0.0) || !(sd > 0.0))) {
    set_error("set_correction: the network's grid must be the context's, scales > 0");
    return HYD_EARG;
  }
  if (net) {  // the stage's work buffers, allocated here and never inside a captured step
    double *q, *tmp;
    HYD_CHECK(dbuf(c, "cnn_q", c.N, &q));
    HYD_CHECK(dbuf(c, "cnn_tmp", c.N, &tmp));
    HYD_CHECK(dbuf(c, "cnn_du", c.cap, &c.cdu));
    HYD_CHECK(dbuf(c, "cnn_dv", c.cap, &c.cdv));
  }
  c.net = net;
  c.net_su = su;
  c.net_sq = sq;
  c.net_sd = sd;
  c.graph_nptq = -1;  // the captured step changes
  return HYD_OK;
}

int hyd_phase_times(hyd_ctx *h, double *out, int32_t reset) {
  Ctx &c = *C(h);
  if (out) {
    out[0] = c.phase_n[0];
    out[1] = c.phase_s[0];
    out[2] = c.phase_n[1];
    out[3] = c.phase_s[1];
  }
  if (reset) {
    c.phase_n[0] = c.phase_n[1] = 0;
    c.phase_s[0] = c.phase_s[1] = 0.0;
  }
  return HYD_OK;
}

int hyd_advance_counts(hyd_ctx *h, int64_t *advances, int64_t *node_updates, int32_t reset) {
  Ctx &c = *C(h);
  if (advances) *advances = c.n_advances;
  if (node_updates) *node_updates = c.n_node_updates;
  if (reset) c.n_advances = c.n_node_updates = 0;
  return HYD_OK;
}

int hyd_init(const hyd_params *p, hyd_ctx **ctx) { return create(p, ctx); }

int hyd_destroy(hyd_ctx *ctx) {
  free_ctx(C(ctx));
  return HYD_OK;
}

int hyd_ptospc(hyd_ctx *h, const double *rvar, double *svar) {
  Ctx &c = *C(h);
  HYD_CHECK(h2d(c, c.wa, rvar, c.N));
  HYD_CHECK(fft_ptospc(c, c.wa, c.wb));
  return d2h(c, svar, c.wb, c.N);
}

int hyd_spctop(hyd_ctx *h, const double *svar, double *rvar) {
  Ctx &c = *C(h);
  HYD_CHECK(h2d(c, c.wa, svar, c.N));
  HYD_CHECK(fft_spctop(c, c.wa, c.wb, PRE_NONE, 1.0, c.wc));
  return d2h(c, rvar, c.wb, c.N);
}

int hyd_initfft(int32_t n, int32_t *factors, double *trig) {
  if (!factors || !trig) {
    set_error("initfft: null array");
    return HYD_EARG;
  }
  if (n < 8 || n > 4096 || (n & (n - 1))) {
    set_error("initfft: n must be a power of two in [8, 4096] (radix 2 and 4 only)");
    return HYD_EFACTOR;
  }
  int f[5];
  std::vector<double> t;
  HYD_CHECK(initfft_table(n, f, t));
  for (int k = 0; k < 5; k++) factors[k] = f[k];
  memcpy(trig, t.data(), sizeof(double) * 2 * (size_t)n);
  return HYD_OK;
}

static int fft_1d_host(Ctx &c, bool fwd, int32_t m, int32_t n, double *x, const double *trig,
                       const int32_t *factors) {
  if (!x || !trig || !factors || m < 1) {
    set_error("forfft/revfft: m >= 1 and non-null arrays");
    return HYD_EARG;
  }
  if (n < 8 || n > 4096 || (n & (n - 1))) {
    set_error("forfft/revfft: n must be a power of two in [8, 4096]");
    return HYD_EFACTOR;
  }
  int f[5];
  std::vector<double> t;
  HYD_CHECK(initfft_table(n, f, t));
  for (int k = 0; k < 5; k++)
    if (factors[k] != f[k]) {
      set_error("forfft/revfft: factors are not initfft's for n");
      return HYD_EFACTOR;
    }
  const int64_t len = (int64_t)m * n;
  double *dx, *dt;
  HYD_CHECK(dbuf(c, "api_fft_x", len, &dx));
  HYD_CHECK(dbuf(c, "api_fft_trig", 2 * (int64_t)n, &dt));
  HYD_HIP(hipMemcpyAsync(dx, x, sizeof(double) * len, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(dt, trig, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c.s));
  HYD_CHECK(fft_1d(c, fwd, m, n, dx, dt));
  return d2h(c, x, dx, len);
}

int hyd_forfft(hyd_ctx *h, int32_t m, int32_t n, double *x, const double *trig,
               const int32_t *factors) {
  return fft_1d_host(*C(h), true, m, n, x, trig, factors);
}

int hyd_revfft(hyd_ctx *h, int32_t m, int32_t n, double *x, const double *trig,
               const int32_t *factors) {
  return fft_1d_host(*C(h), false, m, n, x, trig, factors);
}

int hyd_xderiv(hyd_ctx *h, const double *var, double *der) {
  Ctx &c = *C(h);
  HYD_CHECK(h2d(c, c.wa, var, c.N));
  HYD_CHECK(launch_xderiv(c, c.wa, c.wb));
  return d2h(c, der, c.wb, c.N);
}

int hyd_yderiv(hyd_ctx *h, const double *var, double *der) {
  Ctx &c = *C(h);
  HYD_CHECK(h2d(c, c.wa, var, c.N));
  HYD_CHECK(launch_yderiv(c, c.wa, c.wb));
  return d2h(c, der, c.wb, c.N);
}

int hyd_main_invert(hyd_ctx *h, const double *qq, double *uu, double *vv, double *pp) {
  Ctx &c = *C(h);
  // wa..wd are scratch for the host-array entry points
  HYD_CHECK(h2d(c, c.wa, qq, c.N));
  HYD_CHECK(main_invert_dev(c, c.wa, c.wb, c.wc, c.wd, c.mq));
  HYD_CHECK(d2h(c, uu, c.wb, c.N));
  HYD_CHECK(d2h(c, vv, c.wc, c.N));
  return d2h(c, pp, c.wd, c.N);
}

int hyd_spec1d(hyd_ctx *h, double *ss, double *spec, int32_t iopt) {
  Ctx &c = *C(h);
  if (!ss || !spec || (iopt != 0 && iopt != 1)) {
    set_error("spec1d: null array or iopt not 0/1");
    return HYD_EARG;
  }
  HYD_CHECK(h2d(c, c.wb, ss, c.N));
  HYD_CHECK(diag_spec1d(c, c.wb, c.wc, iopt));
  HYD_CHECK(d2h(c, spec, c.wc, c.ht.kmax + 1));
  if (iopt == 1) HYD_CHECK(d2h(c, ss, c.wb, c.N));
  return HYD_OK;
}

int hyd_norms(hyd_ctx *h, const double *qq, double *ql1, double *ql2) {
  Ctx &c = *C(h);
  if (!qq || !ql1 || !ql2) {
    set_error("norms: null argument");
    return HYD_EARG;
  }
  HYD_CHECK(h2d(c, c.wb, qq, c.N));
  return diag_norms(c, c.wb, ql1, ql2);
}

int hyd_energy(hyd_ctx *h, double *ekin, double *epot) {
  Ctx &c = *C(h);
  if (!ekin || !epot) {
    set_error("energy: null argument");
    return HYD_EARG;
  }
  if (!c.have_field || c.p.precision != 0) {
    set_error("energy: needs a field and the plain FP64 fields (precision 0)");
    return HYD_ESTATE;
  }
  return diag_energy(c, ekin, epot);
}

int hyd_gradient(hyd_ctx *h, const double *ff, double *ffx, double *ffy) {
  Ctx &c = *C(h);
  HYD_CHECK(h2d(c, c.wa, ff, c.N));
  HYD_CHECK(gradient_dev(c, c.wa, c.wb, c.wc, c.mq));
  HYD_CHECK(d2h(c, ffx, c.wb, c.N));
  return d2h(c, ffy, c.wc, c.N);
}

// The tables of init_contours (contours.f90:48-132), indices as doubles:
// 18 xgf, 19 ygf (fine grid lines), 20 xgu, 21 ygu (ultra-fine), 22 the
// congen weights w00, w10, w01, w11 (each mgu x mgu, iyf fastest), 23 ixp,
// 24 iyp, 25 ixfw, ix0w, ix1w, 26 iyfw, iy0w, iy1w (nxu / nyu each),
// 27 ibx (nxu+1), 28 iby (nyu+1).
static int contour_table(const Consts &k, int id, double *out) {
  const int mgu = k.nxu / k.nx;
  switch (id) {
    case 18:
      for (int i = 1; i <= k.nxf; i++) out[i - 1] = k.xmin + k.glxf * (double)(i - 1);
      return HYD_OK;
    case 19:
      for (int i = 1; i <= k.nyf; i++) out[i - 1] = k.ymin + k.glyf * (double)(i - 1);
      return HYD_OK;
    case 20:
      for (int i = 1; i <= k.nxu; i++) out[i - 1] = k.xmin + k.glxu * (double)(i - 1);
      return HYD_OK;
    case 21:
      for (int i = 1; i <= k.nyu; i++) out[i - 1] = k.ymin + k.glyu * (double)(i - 1);
      return HYD_OK;
    case 22: {
      const double fac = 1.0 / (double)mgu;
      const int w = mgu * mgu;
      for (int ixf = 1; ixf <= mgu; ixf++) {
        const double pxf = fac * (double)(ixf - 1), pxc = 1.0 - pxf;
        for (int iyf = 1; iyf <= mgu; iyf++) {
          const double pyf = fac * (double)(iyf - 1), pyc = 1.0 - pyf;
          const int j = (ixf - 1) * mgu + (iyf - 1);
          out[j] = pyc * pxc;
          out[w + j] = pyf * pxc;
          out[2 * w + j] = pyc * pxf;
          out[3 * w + j] = pyf * pxf;
        }
      }
      return HYD_OK;
    }
    case 23:
    case 24: {
      const int n = id == 23 ? k.nx : k.ny;
      for (int i = 1; i <= n; i++) out[i - 1] = (double)(i < n ? i + 1 : 1);
      return HYD_OK;
    }
    case 25:
    case 26: {
      const int nu = id == 25 ? k.nxu : k.nyu, n = id == 25 ? k.nx : k.ny;
      for (int i = 1; i <= nu; i++) {
        const int q = (i - 1) / mgu;
        const int i0 = 1 + q;
        out[i - 1] = (double)(i - mgu * q);          // ixfw
        out[nu + i - 1] = (double)i0;                 // ix0w
        out[2 * nu + i - 1] = (double)(2 + q - n * (i0 / n));  // ix1w
      }
      return HYD_OK;
    }
    case 27:
    case 28: {
      const int nu = id == 27 ? k.nxu : k.nyu;
      out[0] = (double)(nu - 1);
      for (int i = 2; i <= nu + 1; i++) out[i - 1] = (double)(i - 2);
      return HYD_OK;
    }
  }
  set_error("unknown contour table id");
  return HYD_EARG;
}

int hyd_get_table(hyd_ctx *h, int32_t id, double *out) {
  Ctx &c = *C(h);
  const HostTables &t = c.ht;
  const std::vector<double> *v = nullptr;
  switch (id) {
    case 1: v = &t.green; break;
    case 2: v = &t.flo; break;
    case 3: v = &t.fhi; break;
    case 4: v = &t.qdiss; break;
    case 5: v = &t.rksq; break;
    case 6: v = &t.hrkx; break;
    case 7: v = &t.rkx; break;
    case 8: v = &t.hrky; break;
    case 9: v = &t.rky; break;
    case 10: v = &t.xtrig; break;
    case 11: v = &t.ytrig; break;
    case 12: v = &t.spmf; break;
    case 13: v = &t.alk; break;
    case 14:
      out[0] = (double)t.kmax;
      out[1] = (double)t.kmaxred;
      out[2] = c.c.dsumi;
      return HYD_OK;
    case 15: v = &t.bety; break;
    case 16: v = &t.rk; break;
    case 17:
      for (size_t k = 0; k < t.kmag.size(); k++) out[k] = (double)t.kmag[k];
      return HYD_OK;
    default:
      if (id >= 18 && id <= 28) return contour_table(c.c, id, out);
      set_error("unknown table id");
      return HYD_EARG;
  }
  memcpy(out, v->data(), sizeof(double) * v->size());
  return HYD_OK;
}

static int check_links(const int32_t *nextq, int32_t n) {
  for (int32_t i = 0; i < n; i++)
    if (nextq[i] < 1 || nextq[i] > n) {
      set_error("nextq holds a node number outside 1..nptq");
      return HYD_EARG;
    }
  return HYD_OK;
}

// Every cell index the kernels form from a position stays on its grid:
// 1+int(glxi*(x-xmin)) (velint, contours.f90:146), 1+int(glxfi*(x-xmin))
// (con2grid, :207) and int(glxui*(x-xmin)) (con2ugrid, congen.f90:354) all
// lie in 0..n-1 (+1) exactly when these do; NaN and infinities fail.
static bool position_ok(double x, double xmin, double g0, int n0, double g1, int n1, double g2,
                        int n2) {
  const double d = x - xmin;
  if (!(d >= 0.0)) return false;
  const double a = g0 * d, b = g1 * d, e = g2 * d;
  return a < (double)n0 && b < (double)n1 && e < (double)n2;
}

static int check_positions(const Consts &k, const double *xq, const double *yq, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (!position_ok(xq[i], k.xmin, k.glxi, k.nx, k.glxfi, k.nxf, k.glxui, k.nxu) ||
        !position_ok(yq[i], k.ymin, k.glyi, k.ny, k.glyfi, k.nyf, k.glyui, k.nyu)) {
      set_error("a node lies outside the periodic domain [xmin,xmax) x [ymin,ymax)");
      return HYD_EARG;
    }
  return HYD_OK;
}

// The contour store's invariants (contours.f90:12-15, common.f90:78-88):
// contours numbered consecutively from node 1 (i1q(1) = 1, i1q(j+1) =
// i2q(j)+1, i2q(nq) = nptq, npq = i2q-i1q+1 >= 1), levels ascending (indq,
// which surgery relies on, contours.f90:372-394), and nextq one closed
// cycle through each contour's own nodes.  Surgery, renode and congen index
// device arrays with these numbers, so a table breaking them is rejected
// before anything reaches the device.
static int check_table(const Consts &k, int32_t nptq, int32_t nq, const double *xq,
                       const double *yq, const int32_t *nextq, const int32_t *i1q,
                       const int32_t *i2q, const int32_t *npq, const int32_t *indq) {
  if (nptq < 0 || nq < 0 || (nq == 0) != (nptq == 0)) {
    set_error("contour table: nq and nptq must both be zero or both positive");
    return HYD_EARG;
  }
  if (nq == 0) return HYD_OK;
  if (!xq || !yq || !nextq || !i1q || !i2q || !npq || !indq) {
    set_error("contour table: null array");
    return HYD_EARG;
  }
  int64_t expect = 1;
  for (int32_t j = 0; j < nq; j++) {
    if (i1q[j] != expect || i2q[j] < i1q[j] || i2q[j] > nptq || npq[j] != i2q[j] - i1q[j] + 1 ||
        (j > 0 && indq[j] < indq[j - 1])) {
      set_error("contour table: i1q/i2q/npq not consecutive from node 1, or indq descending");
      return HYD_EARG;
    }
    expect = (int64_t)i2q[j] + 1;
  }
  if (expect != (int64_t)nptq + 1) {
    set_error("contour table: i2q(nq) differs from nptq");
    return HYD_EARG;
  }
  HYD_CHECK(check_links(nextq, nptq));
  for (int32_t j = 0; j < nq; j++) {
    int32_t i = i1q[j];
    for (int32_t m = 0; m < npq[j]; m++) {
      const int32_t nx = nextq[i - 1];
      if (nx < i1q[j] || nx > i2q[j] || (nx == i1q[j]) != (m == npq[j] - 1)) {
        set_error("contour table: nextq is not one closed cycle through each contour");
        return HYD_EARG;
      }
      i = nx;
    }
  }
  return check_positions(k, xq, yq, nptq);
}

int hyd_velint(hyd_ctx *h, const double *uu, const double *vv, const double *xq, const double *yq,
               int32_t nptq, double *uq, double *vq) {
  Ctx &c = *C(h);
  if (nptq < 0 || nptq > c.cap) {
    set_error("velint: node count outside [0, capacity]");
    return HYD_ECAPACITY;
  }
  if (nptq == 0) return HYD_OK;
  HYD_CHECK(check_positions(c.c, xq, yq, nptq));
  HYD_CHECK(h2d(c, c.wa, uu, c.N));
  HYD_CHECK(h2d(c, c.wb, vv, c.N));
  HYD_CHECK(h2d(c, c.xqi, xq, nptq));
  HYD_CHECK(h2d(c, c.yqi, yq, nptq));
  HYD_CHECK(velint(c, c.wa, c.wb, c.xqi, c.yqi, nptq, c.uq, c.vq));
  HYD_CHECK(d2h(c, uq, c.uq, nptq));
  return d2h(c, vq, c.vq, nptq);
}

int hyd_con2grid(hyd_ctx *h, const double *xq, const double *yq, const int32_t *nextq,
                 int32_t nptq, double qjump, double *qc, int32_t *counts) {
  Ctx &c = *C(h);
  if (nptq < 0 || nptq > c.cap) {
    set_error("con2grid: node count outside [0, capacity]");
    return HYD_ECAPACITY;
  }
  HYD_CHECK(check_links(nextq, nptq));
  HYD_CHECK(check_positions(c.c, xq, yq, nptq));
  if (nptq > 0) {
    HYD_CHECK(h2d(c, c.xqi, xq, nptq));
    HYD_CHECK(h2d(c, c.yqi, yq, nptq));
    HYD_HIP(hipMemcpyAsync(c.nxt_tmp, nextq, sizeof(int32_t) * nptq, hipMemcpyHostToDevice, c.s));
  }
  const double saved = c.hsc->qjump;
  c.hsc->qjump = qjump;
  int rc = con2grid(c, c.xqi, c.yqi, c.nxt_tmp, nptq, c.wa, counts);
  c.hsc->qjump = saved;
  HYD_CHECK(rc);
  HYD_CHECK(d2h(c, qc, c.wa, c.N));
  return c2g_status(c);
}

// The velocity that the contours' PV field has on this context's grid:
// con2grid, ptospc, main_invert (contours.f90:173-346, spectral.f90:304-342)
// on work buffers (the context's state is untouched), then velint at m
// points.  fields (may be NULL) receives u, v and the PV grid q as FP32
// [pixels][3], pixel = ix*ny + iy.
int hyd_contour_velocity(hyd_ctx *h, int32_t nptq, const double *xq, const double *yq,
                         const int32_t *nextq, double qjump, int32_t m, const double *xs,
                         const double *ys, double *us, double *vs, float *fields) {
  Ctx &c = *C(h);
  if (nptq < 1 || nptq > c.cap || m < 0 || m > c.cap) {
    set_error("contour_velocity: node or point count outside [1, capacity]");
    return HYD_ECAPACITY;
  }
  HYD_CHECK(check_links(nextq, nptq));
  HYD_CHECK(check_positions(c.c, xq, yq, nptq));
  if (m > 0) HYD_CHECK(check_positions(c.c, xs, ys, m));
  const int64_t N = c.N;
  double *q, *qs, *u, *v, *pp, *tmp, *dxs, *dys, *dus, *dvs;
  HYD_CHECK(dbuf(c, "cv_q", N, &q));
  HYD_CHECK(dbuf(c, "cv_qs", N, &qs));
  HYD_CHECK(dbuf(c, "cv_u", N, &u));
  HYD_CHECK(dbuf(c, "cv_v", N, &v));
  HYD_CHECK(dbuf(c, "cv_pp", N, &pp));
  HYD_CHECK(dbuf(c, "cv_tmp", N, &tmp));
  HYD_CHECK(h2d(c, c.xqi, xq, nptq));
  HYD_CHECK(h2d(c, c.yqi, yq, nptq));
  HYD_HIP(hipMemcpyAsync(c.nxt_tmp, nextq, sizeof(int32_t) * nptq, hipMemcpyHostToDevice, c.s));
  const double saved = c.hsc->qjump;
  c.hsc->qjump = qjump;
  int rc = con2grid(c, c.xqi, c.yqi, c.nxt_tmp, nptq, q, nullptr);
  c.hsc->qjump = saved;
  HYD_CHECK(rc);
  HYD_CHECK(fft_ptospc(c, q, qs));
  HYD_CHECK(main_invert_dev(c, qs, u, v, pp, tmp));
  if (m > 0) {
    HYD_CHECK(dbuf(c, "cv_xs", m, &dxs));
    HYD_CHECK(dbuf(c, "cv_ys", m, &dys));
    HYD_CHECK(dbuf(c, "cv_us", m, &dus));
    HYD_CHECK(dbuf(c, "cv_vs", m, &dvs));
    HYD_CHECK(h2d(c, dxs, xs, m));
    HYD_CHECK(h2d(c, dys, ys, m));
    HYD_CHECK(velint(c, u, v, dxs, dys, m, dus, dvs));
    HYD_CHECK(d2h(c, us, dus, m));
    HYD_CHECK(d2h(c, vs, dvs, m));
  }
  if (fields) {
    std::vector<double> hu(N), hv(N), hq(N);
    HYD_CHECK(d2h(c, hu.data(), u, N));
    HYD_CHECK(d2h(c, hv.data(), v, N));
    HYD_CHECK(d2h(c, hq.data(), q, N));
    for (int64_t k = 0; k < N; k++) {
      fields[3 * k] = (float)hu[k];
      fields[3 * k + 1] = (float)hv[k];
      fields[3 * k + 2] = (float)hq[k];
    }
  }
  return c2g_status(c);
}

int hyd_sum_sequential(hyd_ctx *h, const double *a, int64_t n, double *sum) {
  Ctx &c = *C(h);
  if (n < 0 || n > 4 * c.N) {
    set_error("sum_sequential: length outside [0, 4*ng*ng]");
    return HYD_EARG;
  }
  if (n == 0) {
    *sum = 0.0;
    return HYD_OK;
  }
  double *d = c.fq;  // 4N doubles of scratch
  HYD_CHECK(h2d(c, d, a, n));
  HYD_CHECK(seqsum(c, d, n, c.dsum));
  return d2h(c, sum, c.dsum, 1);
}

int hyd_set_field(hyd_ctx *h, const double *qq, double t0) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  if (c.c.beffect) {
    set_error("beta != 0 is not supported by the device step");
    return HYD_EARG;
  }
  // contint, generic.f90:66-79
  double mx = qq[0], mn = qq[0];
  for (int64_t k = 0; k < c.N; k++) {
    mx = qq[k] > mx ? qq[k] : mx;
    mn = qq[k] < mn ? qq[k] : mn;
  }
  memset(c.hsc, 0, sizeof(Scalars));
  c.hsc->qjump = (mx - mn) / (double)c.p.ncontq;
  c.hsc->t = t0;
  c.hsc->dt = 0.0;
  HYD_CHECK(h2d(c, c.qr, qq, c.N));
  HYD_CHECK(fft_ptospc(c, c.qr, c.qs));
  HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  c.nptq = 0;
  c.nq = 0;
  c.phase = 0;
  c.have_field = true;
  return HYD_OK;
}

int hyd_set_contours(hyd_ctx *h, int32_t nptq, int32_t nq, const double *xq, const double *yq,
                     const int32_t *nextq, const int32_t *i1q, const int32_t *i2q,
                     const int32_t *npq, const int32_t *indq, double qjump) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  if (nptq < 0 || nptq > c.cap || nq < 0) {
    set_error("set_contours: counts outside capacity");
    return HYD_ECAPACITY;
  }
  HYD_CHECK(check_table(c.c, nptq, nq, xq, yq, nextq, i1q, i2q, npq, indq));
  if (nptq > 0) {
    HYD_CHECK(h2d(c, c.xq, xq, nptq));
    HYD_CHECK(h2d(c, c.yq, yq, nptq));
    HYD_HIP(hipMemcpyAsync(c.nextq, nextq, sizeof(int32_t) * nptq, hipMemcpyHostToDevice, c.s));
  }
  c.i1q.assign(i1q, i1q + nq);
  c.i2q.assign(i2q, i2q + nq);
  c.npq.assign(npq, npq + nq);
  c.indq.assign(indq, indq + nq);
  c.nptq = nptq;
  c.nq = nq;
  c.phase = 0;  // a restart recontours first (caps.f90 recont)
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  c.hsc->qjump = qjump;
  HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

int hyd_check_contours(const hyd_params *p, int32_t nptq, int32_t nq, const double *xq,
                       const double *yq, const int32_t *nextq, const int32_t *i1q,
                       const int32_t *i2q, const int32_t *npq, const int32_t *indq) {
  if (!p) {
    set_error("check_contours: null parameters");
    return HYD_EARG;
  }
  Consts k;
  HYD_CHECK(build_consts(*p, k));
  return check_table(k, nptq, nq, xq, yq, nextq, i1q, i2q, npq, indq);
}

int hyd_get_counts(hyd_ctx *h, int32_t *nptq, int32_t *nq) {
  Ctx &c = *C(h);
  *nptq = c.nptq;
  *nq = c.nq;
  return HYD_OK;
}

int hyd_get_contours(hyd_ctx *h, double *xq, double *yq, int32_t *nextq, int32_t *i1q,
                     int32_t *i2q, int32_t *npq, int32_t *indq) {
  Ctx &c = *C(h);
  if (c.nptq > 0) {
    HYD_HIP(hipMemcpyAsync(xq, c.xq, sizeof(double) * c.nptq, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipMemcpyAsync(yq, c.yq, sizeof(double) * c.nptq, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipMemcpyAsync(nextq, c.nextq, sizeof(int32_t) * c.nptq, hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
  }
  std::copy(c.i1q.begin(), c.i1q.end(), i1q);
  std::copy(c.i2q.begin(), c.i2q.end(), i2q);
  std::copy(c.npq.begin(), c.npq.end(), npq);
  std::copy(c.indq.begin(), c.indq.end(), indq);
  return HYD_OK;
}

static double *field_ptr(Ctx &c, int id) {
  double *f[] = {c.qs, c.qd, c.qc, c.qq, c.pp, c.uu, c.vv, c.qr, c.epq, c.emq};
  return (id >= 0 && id < 10) ? f[id] : nullptr;
}

int hyd_get_field(hyd_ctx *h, int32_t id, double *out) {
  Ctx &c = *C(h);
  double *f = field_ptr(c, id);
  if (!f) {
    set_error("unknown field id");
    return HYD_EARG;
  }
  return d2h(c, out, f, c.N);
}

int hyd_put_field(hyd_ctx *h, int32_t id, const double *in) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  double *f = field_ptr(c, id);
  if (!f) {
    set_error("unknown field id");
    return HYD_EARG;
  }
  HYD_CHECK(h2d(c, f, in, c.N));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

static double *scalar_slot(Scalars &s, int id) {
  double *v[] = {&s.t, &s.dt, &s.twist, &s.zzrms, &s.zzmax, &s.uumax, &s.qjump, &s.dtmax};
  return (id >= 0 && id < 8) ? v[id] : nullptr;
}

int hyd_get_scalar(hyd_ctx *h, int32_t id, double *out) {
  Ctx &c = *C(h);
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  double *v = scalar_slot(*c.hsc, id);
  if (!v) {
    set_error("unknown scalar id");
    return HYD_EARG;
  }
  *out = *v;
  return HYD_OK;
}

int hyd_set_scalar(hyd_ctx *h, int32_t id, double val) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // the captured step holds scalars (qjump) as launch arguments
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  double *v = scalar_slot(*c.hsc, id);
  if (!v) {
    set_error("unknown scalar id");
    return HYD_EARG;
  }
  *v = val;
  HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

// Start of advect, evolution.f90:44-51: init, then twist = 0 and no
// regularisations counted yet.
static int advect_start(Ctx &c) {
  HYD_CHECK(evolution_init(c));
  // surgery's work buffers, allocated here rather than inside the loop
  if (c.nptq > 0) HYD_CHECK(surgery_reserve(c, c.nptq));
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  c.hsc->twist = 0.0;
  HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
  c.ireg = 0;
  c.phase = 2;
  c.gsave = c.csave = 0;  // init drops pending saves (evolution.f90:115-116)
  if (c.saves_on) {       // and keeps qspre = qs (evolution.f90:126)
    HYD_CHECK(dbuf(c, "qspre", c.N, &c.qspre));
    HYD_HIP(hipMemcpyAsync(c.qspre, c.qs, sizeof(double) * c.N, hipMemcpyDeviceToDevice, c.s));
  }
  complexity_record(c);
  return HYD_OK;
}

int hyd_evolution_init(hyd_ctx *h) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  if (!c.have_field) {
    set_error("evolution_init before set_field");
    return HYD_ESTATE;
  }
  HYD_CHECK(advect_start(c));
  return c2g_status(c);
}

int hyd_prepare(hyd_ctx *h) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  HYD_CHECK(prepare(c));
  c.phase = 0;
  return c2g_status(c);
}

int hyd_recontour(hyd_ctx *h) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  if (!c.have_field) {
    set_error("recontour before set_field");
    return HYD_ESTATE;
  }
  HYD_CHECK(timed_recontour(c));
  c.phase = 1;
  return c2g_status(c);
}

int hyd_renode(hyd_ctx *h, const double *xd, const double *yd, int32_t npd, double *xr,
               double *yr, int32_t cap_out, int32_t *npr) {
  Ctx &c = *C(h);
  *npr = 0;
  if (npd <= 0) return HYD_OK;
  if (npd > c.cap) {
    set_error("renode: node count exceeds capacity");
    return HYD_ECAPACITY;
  }
  double *dxd, *dyd, *dxr, *dyr;
  int64_t *doff_in, *doff_out;
  int32_t *dnp, *dnpr;
  HYD_CHECK(dbuf(c, "api_xd", npd, &dxd));
  HYD_CHECK(dbuf(c, "api_yd", npd, &dyd));
  HYD_CHECK(dbuf(c, "api_off", 2, &doff_in));
  HYD_CHECK(dbuf(c, "api_offo", 2, &doff_out));
  HYD_CHECK(dbuf(c, "api_np", 2, &dnp));
  HYD_CHECK(dbuf(c, "api_npr", 2, &dnpr));
  const int64_t cap = std::max<int32_t>(cap_out, 1);
  HYD_CHECK(dbuf(c, "api_xr", cap, &dxr));
  HYD_CHECK(dbuf(c, "api_yr", cap, &dyr));
  const int64_t zero = 0;
  HYD_HIP(hipMemcpyAsync(dxd, xd, 8 * (size_t)npd, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(dyd, yd, 8 * (size_t)npd, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(doff_in, &zero, 8, hipMemcpyHostToDevice, c.s));
  HYD_HIP(hipMemcpyAsync(dnp, &npd, 4, hipMemcpyHostToDevice, c.s));
  int64_t ntot = 0;
  HYD_CHECK(renode_batch(c, dxd, dyd, doff_in, dnp, 1, npd, dxr, dyr, doff_out, dnpr, &ntot, cap));
  HYD_HIP(hipMemcpyAsync(xr, dxr, 8 * (size_t)ntot, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipMemcpyAsync(yr, dyr, 8 * (size_t)ntot, hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  *npr = (int32_t)ntot;
  return HYD_OK;
}

static int capture_graph(Ctx &c, int (*part)(Ctx &), hipGraphExec_t *out) {
  if (*out) {
    HYD_HIP(hipGraphExecDestroy(*out));
    *out = nullptr;
  }
  hipGraph_t g = nullptr;
  HYD_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
  const int rc = part(c);
  const hipError_t e = hipStreamEndCapture(c.s, &g);
  if (rc != HYD_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  HYD_HIP(e);
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  HYD_HIP(ei);
  return HYD_OK;
}

// The mixed-precision step has no host round trip: after one eager call it
// is captured as two graphs (head and tail) and replayed; the capture is
// redone whenever the node count changed (surgery, recontouring) and
// skipped while timing.  With copy_scalars the step's scalars are copied to
// the host mirror between head and tail (ev_sc; sc_pending set).
static int advance_mixed_graph(Ctx &c, bool copy_scalars = false) {
  c.sweep_clean = false;  // its sweeps leave their counts behind
  c.n_advances += 1;
  c.n_node_updates += c.nptq;
  const bool eager = c.timing || c.mixed_calls++ == 0;
  if (!eager && (!c.step_graph || !c.step_graph_tail || c.graph_nptq != c.nptq)) {
    HYD_CHECK(capture_graph(c, advance_mixed_head, &c.step_graph));
    HYD_CHECK(capture_graph(c, advance_mixed_tail, &c.step_graph_tail));
    c.graph_nptq = c.nptq;
  }
  if (eager) HYD_CHECK(advance_mixed_head(c));
  else HYD_HIP(hipGraphLaunch(c.step_graph, c.s));
  if (copy_scalars) {
    HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipEventRecord(c.ev_sc, c.s));
    c.sc_pending = true;
  }
  if (eager) return advance_mixed_tail(c);
  HYD_HIP(hipGraphLaunch(c.step_graph_tail, c.s));
  return HYD_OK;
}

// A lazily run mixed step's scalars (adapt's, t at the step's start): on the
// host, and its monitor row recorded.
static int mixed_scalars_in(Ctx &c) {
  if (!c.sc_pending) return HYD_OK;
  c.sc_pending = false;
  HYD_HIP(hipEventSynchronize(c.ev_sc));
  monitor_record(c, *c.hsc);
  return HYD_OK;
}

int hyd_inversion(hyd_ctx *h) {
  Ctx &c = *C(h);
  HYD_CHECK(inversion(c));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

int hyd_advance(hyd_ctx *h) {
  Ctx &c = *C(h);
  if (!c.have_field) {
    set_error("advance before set_field");
    return HYD_ESTATE;
  }
  const double t0 = c.hsc->t;
  if (c.p.precision == 1)
    HYD_CHECK(advance_mixed_graph(c));
  else
    HYD_CHECK(advance(c));
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_CHECK(c2g_status(c));
  if (c.p.precision == 1) {  // adapt ran on the device: its row from the scalars
    Scalars s = *c.hsc;
    s.t = t0;
    monitor_record(c, s);
  }
  return HYD_OK;
}

int hyd_surgery(hyd_ctx *h) {
  Ctx &c = *C(h);
  c.graph_nptq = -1;  // state changed outside the captured step
  HYD_CHECK(timed_surgery(c));
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

// advect loop body, evolution.f90:55-80.  lazy: the plain step leaves its
// scalars on the device (no read-back and no wait at its end); the next
// step's adapt brings them, and any sweep overflow, to the host, and the
// caller reads them back once its loop ends.
static int advect_step(Ctx &c, bool lazy = false) {
  const double twistmax = 2.5;
  const int nregmax = 20;
  HYD_CHECK(mixed_scalars_in(c));  // the previous step's, if it ran lazily
  if (c.hsc->twist > twistmax) {
    c.ireg = c.ireg + 1;
    if (c.ireg == nregmax) {
      HYD_CHECK(prepare(c));
      HYD_HIP(hipStreamSynchronize(c.s));
      c.phase = 0;
      return 1;
    }
    HYD_CHECK(timed_surgery(c));
    HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
    HYD_HIP(hipStreamSynchronize(c.s));
    complexity_record(c);
    c.hsc->twist = c.hsc->twist - twistmax;
    HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
  }
  // the save slots read the step's start time from the host mirror, so a
  // run with saves keeps the read-back
  if (lazy && !c.saves_on && c.have_field) {
    if (c.p.precision == 0) return advance(c);
    return advance_mixed_graph(c, true);  // scalars read back by the next step
  }
  return hyd_advance((hyd_ctx *)&c);
}

int hyd_advect_step(hyd_ctx *h) {
  Ctx &c = *C(h);
  if (!c.have_field) {
    set_error("advect_step before set_field");
    return HYD_ESTATE;
  }
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  HYD_HIP(hipStreamSynchronize(c.s));
  return advect_step(c);
}

// Where a run resumes (c.phase): contours rebuilt from qr first (caps.f90
// recont), then advect's start, or straight on inside advect's loop.
static int resume_cycle(Ctx &c) {
  if (c.phase == 0) {
    HYD_CHECK(timed_recontour(c));
    c.phase = 1;
  }
  if (c.phase == 1) HYD_CHECK(advect_start(c));
  return HYD_OK;
}

// caps.f90:83-91 with a step count in place of the final time.  A run
// resumes where the previous call stopped, so run(a) then run(b) is run(a+b).
int hyd_run(hyd_ctx *h, int32_t nsteps) {
  Ctx &c = *C(h);
  if (!c.have_field) {
    set_error("run before set_field");
    return HYD_ESTATE;
  }
  int done = 0;
  while (done < nsteps) {
    HYD_CHECK(resume_cycle(c));
    while (done < nsteps) {
      const int r = advect_step(c, true);
      if (r < 0) return r;
      if (r == 1) break;
      done++;
    }
  }
  HYD_CHECK(mixed_scalars_in(c));
  HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
  return c2g_status(c);  // waits for the stream
}

// An ensemble of independent flows, each running hyd_run's loop for nsteps
// advances.  Rounds: every flow that needs it recontours / regularises /
// does surgery (host-orchestrated, one flow at a time), then the RK4 steps
// of all flows are enqueued on their own streams and run concurrently (a
// mixed-precision flow replays its graph), then one wait per flow reads its
// scalars back.  Each flow's sequence of operations, hence its result, is
// the one hyd_run would produce.
int hyd_run_batch(hyd_ctx **hs, int32_t n, int32_t nsteps) {
  if (!hs || n < 0 || nsteps < 0) {
    set_error("run_batch: bad arguments");
    return HYD_EARG;
  }
  for (int i = 0; i < n; i++) {
    if (!hs[i] || !C(hs[i])->have_field) {
      set_error("run_batch: every flow needs a context with a field");
      return HYD_ESTATE;
    }
    for (int j = 0; j < i; j++)
      if (hs[j] == hs[i]) {
        set_error("run_batch: a context appears twice");
        return HYD_EARG;
      }
  }
  const double twistmax = 2.5;
  const int nregmax = 20;
  std::vector<int> done(n, 0), stepped(n, 0);
  std::vector<double> tstart(n, 0.0);
  for (;;) {
    bool any = false;
    for (int i = 0; i < n; i++) {
      stepped[i] = 0;
      if (done[i] >= nsteps) continue;
      any = true;
      Ctx &c = *C(hs[i]);
      HYD_CHECK(resume_cycle(c));
      if (c.hsc->twist > twistmax) {
        c.ireg = c.ireg + 1;
        if (c.ireg == nregmax) {
          HYD_CHECK(prepare(c));
          HYD_HIP(hipStreamSynchronize(c.s));
          c.phase = 0;  // recontour next round; this round has no advance
          continue;
        }
        HYD_CHECK(timed_surgery(c));
        HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
        HYD_HIP(hipStreamSynchronize(c.s));
        complexity_record(c);
        c.hsc->twist = c.hsc->twist - twistmax;
        HYD_HIP(hipMemcpyAsync(c.dsc, c.hsc, sizeof(Scalars), hipMemcpyHostToDevice, c.s));
      }
      tstart[i] = c.hsc->t;
      if (c.p.precision == 1)
        HYD_CHECK(advance_mixed_graph(c));
      else
        HYD_CHECK(advance(c));
      HYD_HIP(hipMemcpyAsync(c.hsc, c.dsc, sizeof(Scalars), hipMemcpyDeviceToHost, c.s));
      stepped[i] = 1;
    }
    if (!any) break;
    for (int i = 0; i < n; i++) {
      if (!stepped[i]) continue;
      Ctx &c = *C(hs[i]);
      HYD_CHECK(c2g_status(c));  // waits for the flow's stream
      if (c.p.precision == 1) {
        Scalars sc = *c.hsc;
        sc.t = tstart[i];
        monitor_record(c, sc);
      }
      done[i]++;
    }
  }
  for (int i = 0; i < n; i++) HYD_HIP(hipStreamSynchronize(C(hs[i])->s));
  return HYD_OK;
}

__global__ void k_copy_calib(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    b[k] = a[k];
}

__global__ void k_empty() {}

// The cost of the per-launch timing itself: reps launches of an empty
// kernel, each with its own start/stop events as the timed kernels have
// them (hipExtLaunchKernel); the average event interval in ms.
int hyd_event_overhead(hyd_ctx *h, int32_t reps, double *avg_ms) {
  Ctx &c = *C(h);
  if (reps < 1 || !avg_ms) {
    set_error("event_overhead: reps >= 1 and a result pointer");
    return HYD_EARG;
  }
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  int rc = HYD_OK;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = HYD_EHIP;
  for (int r = 0; r < reps && rc == HYD_OK; r++) {
    hipExtLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0u, c.s, ev[2 * r], ev[2 * r + 1], 0u);
    if (hipGetLastError() != hipSuccess) rc = HYD_EHIP;
  }
  double tot = 0.0;
  if (rc == HYD_OK && hipStreamSynchronize(c.s) != hipSuccess) rc = HYD_EHIP;
  for (int r = 0; r < reps && rc == HYD_OK; r++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]) != hipSuccess) rc = HYD_EHIP;
    tot += ms;
  }
  for (auto &e : ev)
    if (e) hipEventDestroy(e);
  if (rc != HYD_OK) {
    set_error("event_overhead: HIP event or launch failed");
    return rc;
  }
  *avg_ms = tot / reps;
  return HYD_OK;
}

int hyd_copy_calibration(hyd_ctx *h, int64_t nbytes) {
  Ctx &c = *C(h);
  const int64_t n = nbytes / 8;
  if (n <= 0 || n > 16 * c.N) {
    set_error("copy_calibration: size outside (0, 128 ng^2] bytes");
    return HYD_EARG;
  }
  // reads the 16 ng^2-double fine-grid buffer, writes a work buffer
  double *dst;
  HYD_CHECK(dbuf(c, "calib_dst", n, &dst));
  hipLaunchKernelGGL(k_copy_calib, dim3(2048), dim3(256), 0, c.s, c.fq, dst, n);
  HYD_HIP(hipGetLastError());
  HYD_HIP(hipStreamSynchronize(c.s));
  return HYD_OK;
}

int hyd_get_tuning(hyd_ctx *h, int32_t *v) {
  Ctx &c = *C(h);
  for (int k = 0; k < 4; k++) v[k] = c.fft_v[k];
  return HYD_OK;
}

int hyd_sync(hyd_ctx *h) {
  HYD_HIP(hipStreamSynchronize(C(h)->s));
  return HYD_OK;
}

int hyd_set_timing(hyd_ctx *h, int32_t on) {
  Ctx &c = *C(h);
  HYD_HIP(hipStreamSynchronize(c.s));
  for (auto &r : c.krec) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  c.krec.clear();
  c.timing = on < 0 ? 0 : (on > 2 ? 2 : on);
  return HYD_OK;
}

// names: max*32 bytes (NUL padded); returns the number of records.
int hyd_kernel_times(hyd_ctx *h, int32_t max, char *names, double *ms, int64_t *bytes) {
  Ctx &c = *C(h);
  HYD_HIP(hipStreamSynchronize(c.s));
  int n = 0;
  for (auto &r : c.krec) {
    if (n >= max) break;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) t = -1.0f;
    strncpy(names + 32 * n, r.name.c_str(), 31);
    names[32 * n + 31] = 0;
    ms[n] = t;
    bytes[n] = r.bytes;
    n++;
  }
  return n;
}

}  // extern "C"
