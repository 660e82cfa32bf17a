// hyd_internal.h -- shared definitions of the MI355X contour-advection library.
//
// Device layouts (all float64 unless noted):
//   physical field  (ny,nx) column-major, iy fastest     -> f[ix*ny + iy]
//   spectral field  (nx,ny) column-major, kx fastest     -> s[ky*nx + kx]
//   fine grid of con2grid, column-major by fine column ix, rows iy=1..nyf:
//     cnt[(ix-1)*nyf + (iy-1)] (packed int32 +/- crossing counts) and the
//     swept field fq[(ix-1)*nyf + (iy-1)] (float64)
//   contour nodes: structure of arrays xq,yq,nextq (1-based successor)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <functional>

#include <map>
#include <string>
#include <vector>

#include "../../include/hydra_amd.h"

namespace hyd {

void set_error(const std::string &msg);

// Lane l's double, broadcast to the wavefront (l wavefront-uniform).
__device__ __forceinline__ double readlane_f64(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

#define HYD_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::hyd::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
      return HYD_EHIP;                                                             \
    }                                                                              \
  } while (0)

#define HYD_CHECK(expr)          \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != HYD_OK) return rc_; \
  } while (0)

// Host-side constants: the run-time restatement of constants.f90 and the
// parameter block of contours.f90:17-42.
struct Consts {
  int nx, ny, nxf, nyf, nxu, nyu;
  double ellx, elly, hlx, hlxi, xmin, xmax, hly, hlyi, ymin, ymax;
  double kdsq, srwfm, glx, glxi, gly, glyi, garea, dsumi;
  double glxf, glxfi, glyf, glyfi, glxu, glxui, glyu, glyui;
  double amu, ell, dm, dm2, d4small, dmsq, dmi, elf, densf;
  double oms, small, small3;
  bool btropic, beffect;
};

// Spectral tables built on the host (tables.cpp).
struct HostTables {
  std::vector<double> green, flo, fhi, qdiss, rksq, rk;   // (nx,ny)
  std::vector<double> rkx, hrkx, rky, hrky, bety;          // 1-d
  std::vector<double> xtrig, ytrig;                        // 2n
  int xfactors[5], yfactors[5];
  // distinct values of qdiss and the map cell -> distinct index (for the
  // host-evaluated integrating factor, evolution.f90:603-605)
  std::vector<double> qdiss_u;
  std::vector<int32_t> qdiss_map;
  // spectrum shells, spectral.f90:110-133: kmag (nx,ny), kmax, kmaxred,
  // spmf and alk (0..max(nx,ny))
  std::vector<int32_t> kmag;
  int kmax = 0, kmaxred = 0;
  std::vector<double> spmf, alk;
};

int build_consts(const hyd_params &p, Consts &c);
int build_tables(const hyd_params &p, const Consts &c, HostTables &t);
// initfft's factors and trig table for a power-of-two length n
// (stafft.f90:63-125); HYD_EFACTOR for other lengths.
int initfft_table(int n, int f[5], std::vector<double> &trig);

// FFT plan for one length n = 2^p (radix-2 stage then radix-4 stages, the
// factor order of stafft.f90:146-178 and 212-271).
struct FftPlan {
  int n;
  int n2;  // number of radix-2 factors (0 or 1)
  int n4;  // number of radix-4 factors
  double fnorm, rnorm;  // 1/sqrt(n) and 2/sqrt(n)
};

// Named, grow-only device work buffers (recontouring, renode).
struct DBuf {
  void *p = nullptr;
  int64_t cap = 0;  // bytes
};

// Device context.
struct Scalars {  // lives in device memory, mirrored on the host
  double t, dt, dt2, dt3, dt6, twist, zzrms, zzmax, uumax, qjump, dtmax, dtacc;
  // con2grid's sticky error flag as adapt saw it (nonzero: a sweep since the
  // last check overflowed), so a loop that skips the end-of-step read-back
  // still learns of it at the next step's adapt
  double sweep_err;
};

struct Ctx {
  hyd_params p;
  Consts c;
  HostTables ht;
  FftPlan plan;
  int fft_v[4] = {0, 0, 0, 0};    // autotuned FFT batch per layout class (FP64 passes)
  int fft_v1[4] = {0, 0, 0, 0};   // the same for one-field FP64 launches (a smaller grid)
  int fft_v4[4] = {0, 0, 0, 0};   // and for three- and four-field FP64 launches
  int fft_v32[4] = {0, 0, 0, 0};  // the same for the FP32 passes of the mixed step
  hipStream_t s = nullptr;
  int dev = 0;
  int64_t N = 0;        // ng*ng
  int64_t cap = 0;      // node capacity
  // tables on device
  double *green = nullptr, *flo = nullptr, *fhi = nullptr, *hrkx = nullptr, *hrky = nullptr;
  double *xtrig = nullptr, *ytrig = nullptr, *xgf = nullptr;
  int32_t *qdiss_map = nullptr;
  double *qdiss_exp = nullptr;  // exp(dfac*qdiss_u) for the current step
  int nqu = 0;
  // fields
  double *qs = nullptr, *qd = nullptr, *qc = nullptr, *qq = nullptr, *pp = nullptr;
  double *uu = nullptr, *vv = nullptr, *qr = nullptr, *emq = nullptr, *epq = nullptr;
  double *qsi = nullptr, *qsf = nullptr, *sqs = nullptr, *qdi = nullptr, *qdf = nullptr;
  double *sqd = nullptr, *wa = nullptr, *wb = nullptr, *wc = nullptr, *wd = nullptr;
  // con2grid's own grids (the plain step builds the next stage's contour
  // grid while the source of the current one runs on s2)
  double *c2g_q = nullptr, *c2g_w = nullptr;
  // con2grid work
  int32_t *cnt = nullptr;      // (nyf+1) x nxf packed counts
  double *fq = nullptr;        // nyf x nxf swept fine field
  double *mq = nullptr;        // (2ng)^2 mid grid
  int32_t *qjx_cnt = nullptr;  // nxf packed bottom-row counts
  double *qbot = nullptr;      // nxf
  int32_t *flags = nullptr;    // small device counters; [8] sticky overflow
  // cnt, qjx_cnt and flags[0..7] are all zero: every FP64 sweep clears what
  // it used on its way out, so the next one skips the clearing fills
  bool sweep_clean = false;
  // order-dependent cells: key -> bucket table, bucket list, entries, sums
  int32_t *rp_tkey = nullptr, *rp_tid = nullptr, *rp_list = nullptr;
  int64_t rp_tsize = 0, rp_bcap = 0;
  int32_t *replay = nullptr;   // bucket entries (node << 1 | sign)
  double *replay_val = nullptr;
  int64_t replay_cap = 0;
  // mixed precision (precision == 1): FP32 physical fields and twiddles
  float *uu32 = nullptr, *vv32 = nullptr, *fa32 = nullptr, *fb32 = nullptr, *fc32 = nullptr;
  float *c2g_q32 = nullptr;  // con2grid_mixed's output grid
  float *tmp32 = nullptr, *fq32 = nullptr, *mq32 = nullptr;
  float *fd32 = nullptr, *fe32 = nullptr, *t32[4] = {nullptr, nullptr, nullptr, nullptr};
  double *t4[4] = {nullptr, nullptr, nullptr, nullptr};  // FFT temporaries of a batch
  float *xtrig32 = nullptr, *ytrig32 = nullptr;
  int32_t *qb32 = nullptr;        // bottom-row level counts
  double *qdiss_u_dev = nullptr;  // distinct qdiss values (device exp)
  double *part = nullptr;         // per-block partial sums (1024)
  // nodes
  double *xq = nullptr, *yq = nullptr, *xqi = nullptr, *yqi = nullptr, *xqf = nullptr;
  double *yqf = nullptr, *uq = nullptr, *vq = nullptr;
  int32_t *nextq = nullptr;
  int32_t *nxt_tmp = nullptr;  // staging for the host-array entry points
  std::vector<int32_t> i1q, i2q, npq, indq;  // contour table (host copy)
  int nptq = 0, nq = 0;
  bool contex = false;
  int ireg = 0;  // regularisations since the last recontouring
  // diagnostics records (monitor.asc rows, evolution.f90:581-584: t, cfl,
  // zzrms^2/2, zzrms, zzmax, uumax, twist; complexity.asc rows,
  // evolution.f90:72 and 111: t, nq, nptq) and the save schedule of adapt
  // (evolution.f90:587-599): when a grid save falls in the step, qs and the
  // energies at adapt are kept for the interpolation (evolution.f90:660-668)
  std::vector<double> rec_monitor, rec_complexity;
  bool saves_on = false;
  double tinit = 0.0;
  int32_t igrec = 0, icrec = 1;
  int32_t gsave = 0, csave = 0;
  double tgrid = 0.0, save_dt = 0.0, ekinpre = 0.0, epotpre = 0.0;
  double *qspre = nullptr;  // named work buffer "qspre" once a grid save occurs
  // saves made at the start of a step (evolution.f90:202-204), waiting for
  // the caller: grid saves (tgrid, ekin, epot, igrec; qqs spectral) and
  // contour saves (t, qjump, icrec, contex; the residual spctop(qq - qc),
  // npq, i1q, indq and the nodes)
  struct GridSave {
    double tgrid, ekin, epot;
    int32_t igrec;
    std::vector<double> qqs;
  };
  struct ContSave {
    double t, qjump;
    int32_t icrec, contex;
    std::vector<double> res, xq, yq;
    std::vector<int32_t> npq, i1q, indq;
  };
  std::vector<GridSave> grid_saves;
  std::vector<ContSave> cont_saves;
  // learned velocity correction added to velint at the nodes in every
  // stage (no reference counterpart; off unless a network is attached)
  hyd_cnn *net = nullptr;
  double net_su = 1.0, net_sq = 1.0, net_sd = 1.0;  // input and output scales
  float *cdu = nullptr, *cdv = nullptr;  // named buffers "cnn_du", "cnn_dv"
  // where a step-count run resumes: 0 = contours to rebuild from qr
  // (caps.f90 recont), 1 = contours built, advect's start to run,
  // 2 = inside advect's loop
  int phase = 0;
  // surgery (0) and recontouring (1) calls and their wall seconds
  int64_t phase_n[2] = {0, 0};
  double phase_s[2] = {0.0, 0.0};
  int64_t n_advances = 0, n_node_updates = 0;  // RK4 advances taken, and their nodes
  bool have_field = false;
  // scalars
  Scalars *dsc = nullptr;   // device
  Scalars *hsc = nullptr;   // pinned host mirror
  double *dsum = nullptr;   // device scratch for sequential sums (8 slots)
  unsigned long long *amax = nullptr;  // adapt's per-block maxima (2 x 1024)
  double *exp_host = nullptr;    // pinned: host-evaluated exp table of the step
  hipEvent_t ev_sc = nullptr;    // adapt's scalars have reached the host mirror
  // second stream of the plain step (default priority) and its fork/join events
  // diagnostics (diag.hip): spectrum shells in the reference's mode order
  int32_t *shell_ptr = nullptr;  // kmax + 2
  int32_t *shell_ent = nullptr;  // N: mode index | weight code << 29
  double *rksq = nullptr;        // (nx,ny), spec1d's iopt = 1
  double *diag_out = nullptr;    // 8 scratch results
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // The clearing that ends a sweep (counters zeroed, order-dependent cell
  // table emptied) may be left to the caller (defer_clear), who runs it
  // later on either stream (sweep_clear); the next sweep first waits for it
  // (sweep_clear_join), or runs it itself if nobody has.
  bool defer_clear = false, clear_pending = false, clear_event = false;
  int32_t *clr_cnt = nullptr;
  int64_t clr_n4 = 0;
  hipEvent_t ev_cleared = nullptr;
  // work that the plain step starts on its second stream once the contour
  // grid's node and sweep kernels are queued (so that those run alone on
  // the GPU and the second stream's work overlaps the single-workgroup
  // sequential mean instead); con2grid runs and clears it
  std::function<int()> after_sweep;
  void *seq_work = nullptr; // chunk table of seqsum
  int64_t seq_cap = 0;      // chunks it holds
  // work buffers of recontour / renode, allocated on first use
  std::map<std::string, DBuf> cg_bufs;
  // captured mixed-precision step, head and tail (replayed while the node
  // count is unchanged)
  hipGraphExec_t step_graph = nullptr, step_graph_tail = nullptr;
  bool sc_pending = false;  // a mixed step's scalars are on their way to the host (ev_sc)
  int graph_nptq = -1;
  int mixed_calls = 0;
  // timing
  int timing = 0;  // 0 off, 1 single-kernel ("k:") records only, 2 all records
  struct KRec { std::string name; hipEvent_t a, b; int64_t bytes; };
  std::vector<KRec> krec;
};

// ---- launchers (each enqueues on ctx.s) ----------------------------------
// FFT: ptospc / spctop with optional fused spectral pre-operations.
enum SpecPre { PRE_NONE = 0, PRE_XDERIV = 1, PRE_YDERIV = 2 };
int fft_ptospc(Ctx &c, const double *rvar, double *svar);
int fft_ptospc_mean(Ctx &c, const double *rvar, const double *sum, double ndiv, double *svar);
int fft_spctop(Ctx &c, const double *svar, double *rvar, int pre, double post_sign,
               double *tmp);
int fft_forward_1d(Ctx &c, double *x, int m);   // forfft(m, ng, x)
// One batched 1D pass in stafft's layout (m vectors of length n, the
// vector index fastest) in place on device memory: forfft or revfft
// (stafft.f90:187-406) with the caller's trig table (device copy).
int fft_1d(Ctx &c, bool fwd, int m, int n, double *x, const double *trig);
int fft_ptospc32(Ctx &c, const float *rvar, double *svar);
int fft_spctop_n(Ctx &c, int F, const double *const *svar, double *const *rvar, const int *pre,
                 const double *post_sign, double *const *tmp);
int fft_ptospc_n(Ctx &c, int F, const double *const *rvar, double *const *svar);
int fft_spctop32_n(Ctx &c, int F, const double *const *svar, float *const *rvar, const int *pre,
                   const double *post_sign, float *const *tmp);
int fft_ptospc32_n(Ctx &c, int F, const float *const *rvar, double *const *svar,
                   float *const *tmp);
int fft_spctop32(Ctx &c, const double *svar, float *rvar, int pre, double post_sign);
int con2grid_mixed(Ctx &c, float *qc, bool defer_mean = false);
int fft_ptospc32_mean(Ctx &c, const float *rvar, const double *mean, double *svar);
int advance_mixed(Ctx &c);
int advance_mixed_head(Ctx &c);
int advance_mixed_tail(Ctx &c);
int fft_autotune(Ctx &c);
int fft_reverse_1d(Ctx &c, double *x, int m);   // revfft(m, ng, x)

int main_invert_dev(Ctx &c, const double *qq, double *uu, double *vv, double *pp, double *tmp);
int gradient_dev(Ctx &c, const double *ff, double *fx, double *fy, double *tmp);
int launch_xderiv(Ctx &c, const double *var, double *der);
int launch_yderiv(Ctx &c, const double *var, double *der);

int seqsum(Ctx &c, const double *a, int64_t n, double *dout);
// diagnostics (diag.hip)
int diag_energy(Ctx &c, double *ekin, double *epot);
void monitor_record(Ctx &c, const Scalars &s);
int save_slots(Ctx &c, bool grid, bool cont);
// the learned correction at the nodes (cnn.hip)
int cnn_correction(hyd_cnn *h, hipStream_t s, const void *uu, const void *vv, const double *q,
                   bool f32, double su, double sq, double sd, const double *xn, const double *yn, int m,
                   float *du, float *dv);
int cnn_grid(hyd_cnn *h);
int correction(Ctx &c);
int diag_norms(Ctx &c, const double *qq, double *ql1, double *ql2);
int diag_spec1d(Ctx &c, double *ss, double *spec, int iopt);
void diag_free(Ctx &c);
int64_t seqsum_work_bytes(int64_t n);
int64_t seqsum_work_chunks(int64_t n);

int velint(Ctx &c, const double *uu, const double *vv, const double *xq, const double *yq,
           int n, double *uq, double *vq);
int con2grid(Ctx &c, const double *xq, const double *yq, const int32_t *nextq, int n,
             double *qc, int32_t *counts_out, bool defer_mean = false);
int c2g_status(Ctx &c);
int sweep_clear(Ctx &c);
int sweep_clear_join(Ctx &c);

template <class T>
int dbuf(Ctx &c, const char *name, int64_t n, T **out) {
  DBuf &b = c.cg_bufs[name];
  if (b.cap < n * (int64_t)sizeof(T) || !b.p) {
    // a buffer that has to grow (node counts rise between surgeries) grows
    // by half again at least, so that reallocation stays rare
    const int64_t grown = b.p ? b.cap + b.cap / 2 : 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    const int64_t bytes = std::max<int64_t>(std::max<int64_t>(n * (int64_t)sizeof(T), grown), 256);
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      set_error(std::string("out of device memory for work buffer ") + name);
      return HYD_ENOMEM;
    }
    b.cap = bytes;
  }
  *out = (T *)b.p;
  return HYD_OK;
}

int recontour(Ctx &c, const double *qq);
int surgery(Ctx &c);
int surgery_reserve(Ctx &c, int64_t nodes);
int prepare(Ctx &c);
int evolution_init(Ctx &c);
int advance(Ctx &c);
void free_congen(Ctx &c);
int renode_batch(Ctx &c, const double *xd, const double *yd, const int64_t *off_in,
                 const int32_t *npd, int ncont, int64_t nin, double *xr, double *yr,
                 int64_t *off_out, int32_t *npr, int64_t *ntot_out, int64_t cap_out);

// timing helpers
struct KTimer {
  Ctx &c;
  int idx = -1;
  bool dispatch = false;
  // dispatch = true: the caller launches one kernel with start()/stop() as
  // the launch's own events (hipExtLaunchKernel), which time the dispatch
  // itself rather than markers placed around it on the stream
  KTimer(Ctx &ctx, const char *name, int64_t bytes, bool dispatch = false);
  ~KTimer();
  hipEvent_t start() const;
  hipEvent_t stop() const;
};

}  // namespace hyd
