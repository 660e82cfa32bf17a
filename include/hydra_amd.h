/*
 * hydra_amd.h -- C ABI of the MI355X contour-advection step.
 *
 * A drop-in for the doubly periodic CAPS step of Hydra (ca/plane/qg/caps and
 * lib/stafft).  Every entry point below replaces one reference routine,
 * cited as file:line (paths relative to basic_hydra/).  All state lives in an
 * opaque context; arrays crossing the boundary are plain pointers and sizes.
 *
 * Conventions, identical to the reference's Fortran so that a host can pass
 * its module arrays unchanged:
 *   - physical fields are (ny,nx) column-major (iy fastest), length ng*ng
 *   - spectral fields are (nx,ny) column-major (kx fastest) in the stafft
 *     half-complex layout
 *   - contour index arrays (nextq, i1q, i2q) hold 1-based node numbers;
 *     contours of one level are contiguous and levels ascend (indq)
 *   - "host" entry points take host pointers and copy in/out; the
 *     device-resident step (hyd_advance / hyd_advect_step) keeps the whole
 *     state in HBM between calls
 *
 * Every function returns 0 on success or a negative HYD_E* code; the text
 * of the last error is available from hyd_last_error().  (The reference
 * stops the program on the same conditions: stafft.f90:85-93,
 * sta2dfft.f90:70-93.)
 */
#ifndef HYDRA_AMD_H
#define HYDRA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYD_OK 0
#define HYD_EARG (-1)      /* invalid argument                             */
#define HYD_EFACTOR (-2)   /* grid length not a power of two (radix 2/4)  */
#define HYD_ECAPACITY (-3) /* node / contour / crossing capacity exceeded */
#define HYD_EHIP (-4)      /* HIP runtime error                            */
#define HYD_ENOMEM (-5)
#define HYD_ESTATE (-6)    /* call out of order (no contours, no field)   */

typedef struct hyd_ctx hyd_ctx;

/* Run-time form of the compile-time parameters of parameters.f90:14-36 and
   the capacities of constants.f90:34-41. */
typedef struct hyd_params {
  int32_t ng;            /* nx = ny, a power of two >= 8                   */
  int32_t ncontq;        /* PV contour intervals (parameters.f90:17)       */
  double ellx, elly;     /* domain lengths (parameters.f90:29)             */
  double kd, beta;       /* parameters.f90:30                              */
  double rtherm, rekman; /* parameters.f90:31 (rtherm must be 0)           */
  double cdamp, nnu;     /* parameters.f90:32                              */
  double tgsave, tcsave; /* parameters.f90:23 (only dtmax depends on them) */
  double dt_fixed;       /* > 0: fixed time step instead of adapt's choice */
  int64_t node_capacity; /* device node capacity (0: max(64*ng*ng, 2^20))   */
  int32_t precision;     /* 0: plain FP64, operation-order faithful;
                            1: mixed -- FP32 grid fields and transforms,
                            FP64 spectral state, nodes and RK accumulation
                            (no reference counterpart; not bit-exact)     */
  int32_t device;        /* HIP device ordinal                             */
  double ell_cells;      /* surgery/renode length scale ell in grid lengths
                            glx; 0 means the reference's 6.25
                            (contours.f90:38: ell=6.25*glx, dm=amu**2*ell/4
                            with amu=0.2, so dm = 0.0625*glx); 1.0 gives
                            the dm = 0.01*glx that the ng=1024 workload
                            names (about four times the nodes)            */
} hyd_params;

const char *hyd_last_error(void);

/* init_spectral + init_contours: spectral.f90:42-158, contours.f90:48-132,
   sta2dfft.f90:40-97, stafft.f90:63-125.  Tables are built on the host with
   the same libm calls as the reference and uploaded once. */
int hyd_init(const hyd_params *p, hyd_ctx **ctx);
int hyd_destroy(hyd_ctx *ctx);

/* ---- lib/stafft/stafft.f90 (1D real FFT) ---------------------------------- */
/* initfft(n,factors,trig) stafft.f90:63-125, on the host without a context:
   factors[5] = the counts of the radices 6, 4, 2, 3, 5 (the reference's
   order), trig[2n] = cos, then -sin, of its angle table.  n must be a power
   of two in [8, 4096] (only radix 4 and 2 occur); HYD_EFACTOR otherwise, as
   the reference stops (stafft.f90:85-93). */
int hyd_initfft(int32_t n, int32_t *factors, double *trig);
/* forfft(m,n,x,trig,factors) stafft.f90:187-287 and revfft stafft.f90:
   290-406: the m vectors of length n in x(0:m*n-1) (vector index fastest)
   transformed in place -- forward to the half-complex layout scaled by
   1/sqrt(n), backward scaled by 2/sqrt(n) -- on the context's device, with
   the caller's trig table and factors (those initfft gives for n). */
int hyd_forfft(hyd_ctx *ctx, int32_t m, int32_t n, double *x, const double *trig,
               const int32_t *factors);
int hyd_revfft(hyd_ctx *ctx, int32_t m, int32_t n, double *x, const double *trig,
               const int32_t *factors);

/* ---- lib/stafft/sta2dfft.f90 (doubly periodic part) ---------------------- */
/* ptospc(nx,ny,rvar,svar,...) sta2dfft.f90:103-132.  rvar(ny,nx) -> svar(nx,ny).
   (The reference destroys rvar; here it is left unchanged.) */
int hyd_ptospc(hyd_ctx *ctx, const double *rvar, double *svar);
/* spctop(nx,ny,svar,rvar,...) sta2dfft.f90:136-165 */
int hyd_spctop(hyd_ctx *ctx, const double *svar, double *rvar);
/* xderiv / yderiv with the module's filtered wavenumbers hrkx / hrky,
   sta2dfft.f90:169-245 */
int hyd_xderiv(hyd_ctx *ctx, const double *var, double *der);
int hyd_yderiv(hyd_ctx *ctx, const double *var, double *der);

/* ---- ca/plane/qg/caps/spectral.f90 --------------------------------------- */
/* main_invert(qq,uu,vv,pp) spectral.f90:304-342 */
int hyd_main_invert(hyd_ctx *ctx, const double *qq, double *uu, double *vv, double *pp);
/* gradient(ff,ffx,ffy) spectral.f90:345-366 */
int hyd_gradient(hyd_ctx *ctx, const double *ff, double *ffx, double *ffy);
/* spec1d(ss,spec,iopt) spectral.f90:370-424: the 1D spectrum of the
   spectral field ss (nx,ny), each shell summed in the reference's mode
   order; spec[0..kmax] written (spec holds max(nx,ny)+1 values, the rest
   untouched); iopt = 1 first multiplies ss by rksq and returns it so. */
int hyd_spec1d(hyd_ctx *ctx, double *ss, double *spec, int32_t iopt);

/* ---- ca/plane/qg/caps/generic.f90 ---------------------------------------- */
/* l1norm(qq,ql1) generic.f90:30-46 and l2norm(qq,ql2) generic.f90:48-64 of a
   physical field qq (ny,nx): garea*sum(|qq|), garea*sum(qq**2), the sums
   left to right as the reference's. */
int hyd_norms(hyd_ctx *ctx, const double *qq, double *ql1, double *ql2);

/* ---- ca/plane/qg/caps/evolution.f90 (diagnostics) ------------------------ */
/* energy(ekin,epot) evolution.f90:612-644 from the context's uu, vv, pp as
   left by the last inversion (plain precision only: HYD_ESTATE otherwise). */
int hyd_energy(hyd_ctx *ctx, double *ekin, double *epot);

/* Spectral tables of init_spectral (spectral.f90:14-22): id 1 green, 2 flo,
   3 fhi, 4 qdiss, 5 rksq (ng*ng each); 6 hrkx, 7 rkx, 8 hrky, 9 rky (ng);
   10 xtrig, 11 ytrig (2*ng); spectrum shells (spectral.f90:110-133): 12 spmf,
   13 alk (ng+1 each, alk(0) unused), 14 kmax, kmaxred, dsumi (3 values);
   15 bety (ng); 16 rk, 17 kmag (ng*ng; kmag as whole numbers).
   Tables of init_contours (contours.f90:48-132), integers as whole
   numbers: 18 xgf (4ng), 19 ygf (4ng), 20 xgu (16ng), 21 ygu (16ng), 22
   the congen weights w00, w10, w01, w11 one after another (16x16 each, iyf
   fastest), 23 ixp (ng), 24 iyp (ng), 25 ixfw, ix0w, ix1w (16ng each), 26
   iyfw, iy0w, iy1w (16ng each), 27 ibx (16ng+1), 28 iby (16ng+1). */
int hyd_get_table(hyd_ctx *ctx, int32_t id, double *out);

/* ---- ca/plane/qg/caps/contours.f90 ---------------------------------------- */
/* velint(uu,vv,uq,vq) contours.f90:136-169 over nodes (xq,yq) */
int hyd_velint(hyd_ctx *ctx, const double *uu, const double *vv, const double *xq,
               const double *yq, int32_t nptq, double *uq, double *vq);
/* con2grid(qc) contours.f90:173-346.  The module contour arrays are passed
   explicitly.  If counts != NULL the integer PV level counts of the 4ng
   fine grid after the y sweep (round(qa/qjump), (4ng,4ng) column-major) are
   returned as well. */
int hyd_con2grid(hyd_ctx *ctx, const double *xq, const double *yq, const int32_t *nextq,
                 int32_t nptq, double qjump, double *qc, int32_t *counts);

/* renode(xd,yd,npd,xr,yr,npr) contours.f90:687-933.  xr/yr must hold
   cap_out nodes; *npr = 0 means the contour is too small and is deleted. */
int hyd_renode(hyd_ctx *ctx, const double *xd, const double *yd, int32_t npd, double *xr,
               double *yr, int32_t cap_out, int32_t *npr);

/* surgery contours.f90:350-683 on the device contour store: nodes closer
   than dm to a segment of the same PV level are reconnected (in the
   reference's node order), contours are rebuilt, renoded, and contours of
   three or fewer nodes are removed. */
int hyd_surgery(hyd_ctx *ctx);

/* ---- ca/plane/qg/caps/congen.f90 ------------------------------------------ */
/* recontour(qr) congen.f90:24-128 on the device state: the contours and the
   residual qr (the full PV when no contours exist yet) are turned into new
   contours on the 16ng grid; replaces the device contour store. */
int hyd_recontour(hyd_ctx *ctx);

/* ---- exact sequential reduction ------------------------------------------ */
/* The value gfortran's SUM(a) (and the explicit loops of contours.f90:330-336)
   produce: ((a1 + a2) + a3) + ... in storage order, bit for bit.  Used for
   qavg0 (contours.f90:331-336) and zzrms (evolution.f90:553). */
int hyd_sum_sequential(hyd_ctx *ctx, const double *a, int64_t n, double *sum);
/* The same sequential sum without any allocation and without a context:
   a (n doubles) and sum (one double) are device pointers, the work runs on
   stream (a hipStream_t; NULL = the null stream) and the call returns
   without waiting.  No counterpart in the reference: it is for callers that
   cannot allocate -- code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc.  The work area is a static device array
   (24 MiB per GPU, held until the process ends); n is limited to 2^22
   (HYD_ECAPACITY beyond).  All calls on one GPU share that array, so calls
   that overlap on one GPU (different streams, or threads) race. */
int hyd_sum_sequential_noalloc(const double *a, int64_t n, double *sum, void *stream);

/* ---- device-resident state and the RK4 step ------------------------------ */
/* caps.f90:143-183 (loopinit=0): read the full PV qq(ny,nx) at time t0,
   set qjump (contint, generic.f90:66-79), qs = ptospc(qq), no contours. */
int hyd_set_field(hyd_ctx *ctx, const double *qq, double t0);
/* Replace the device contour store (contours.f90:13-15). */
int hyd_set_contours(hyd_ctx *ctx, int32_t nptq, int32_t nq, const double *xq, const double *yq,
                     const int32_t *nextq, const int32_t *i1q, const int32_t *i2q,
                     const int32_t *npq, const int32_t *indq, double qjump);
/* The checks hyd_set_contours makes, without a context or a device (a host
   can validate a restart file before it opens one): contours numbered
   consecutively from node 1 (i1q, i2q, npq as common.f90:78-88 rebuilds
   them), levels ascending in indq (contours.f90:372-394 relies on it),
   nextq one closed cycle through each contour's nodes, every node inside
   [xmin,xmax) x [ymin,ymax).  HYD_EARG, with the reason in
   hyd_last_error(), when one fails. */
int hyd_check_contours(const hyd_params *p, int32_t nptq, int32_t nq, const double *xq,
                       const double *yq, const int32_t *nextq, const int32_t *i1q,
                       const int32_t *i2q, const int32_t *npq, const int32_t *indq);
int hyd_get_counts(hyd_ctx *ctx, int32_t *nptq, int32_t *nq);
int hyd_get_contours(hyd_ctx *ctx, double *xq, double *yq, int32_t *nextq, int32_t *i1q,
                     int32_t *i2q, int32_t *npq, int32_t *indq);
/* Fields: 0 qs, 1 qd, 2 qc, 3 qq, 4 pp, 5 uu, 6 vv, 7 qr, 8 epq, 9 emq */
int hyd_get_field(hyd_ctx *ctx, int32_t id, double *out);
int hyd_put_field(hyd_ctx *ctx, int32_t id, const double *in);
/* Scalars: 0 t, 1 dt, 2 twist, 3 zzrms, 4 zzmax, 5 uumax, 6 qjump, 7 dtmax */
int hyd_get_scalar(hyd_ctx *ctx, int32_t id, double *out);
int hyd_set_scalar(hyd_ctx *ctx, int32_t id, double v);

/* init (evolution.f90:96-131): qc from the contours, qd = fhi*(qs-qc); then
   the start of advect (evolution.f90:44-51): twist = 0, regularisation
   count 0.  Call after every recontouring. */
int hyd_evolution_init(hyd_ctx *ctx);
/* prepare (evolution.f90:135-162): fold the contour PV into qs and leave the
   residual qr = spctop(qs-qc) for recontouring. */
int hyd_prepare(hyd_ctx *ctx);
/* inversion (evolution.f90:494-520) on the device state. */
int hyd_inversion(hyd_ctx *ctx);
/* advance (evolution.f90:166-326): one RK4 step of nodes and fields. */
int hyd_advance(hyd_ctx *ctx);
/* One pass of the advect loop body, evolution.f90:55-80: when twist exceeds
   2.5, count a regularisation; at the 20th, prepare and return 1 (the caller
   recontours, caps.f90:83-91), otherwise surgery and twist -= 2.5.  Then one
   advance.  Returns 0 after an advance, 1 after prepare, < 0 on error. */
int hyd_advect_step(hyd_ctx *ctx);
/* caps.f90:83-91 driven by a step count: recontour, init, then advect
   steps, recontouring whenever the regularisation count reaches 20, until
   nsteps advances have been taken.  A call resumes where the previous one
   stopped: inside the advect loop it continues the loop (run(a) then
   run(b) is run(a+b)); after set_field, set_contours or prepare it starts
   with the recontouring, after recontour with advect's start.  Between the
   steps of one call the host does not wait for the GPU (each step's adapt
   brings its scalars back); the call returns once all of them are done. */
int hyd_run(hyd_ctx *ctx, int32_t nsteps);
/* Regularisation costs of the runs so far (no reference counterpart; the
   reference's phases are timed from outside): out[0] surgery calls, out[1]
   their wall seconds, out[2] recontourings, out[3] their wall seconds (each
   call timed to its stream's drain).  reset != 0 zeroes the counters after
   reading; out may be NULL. */
int hyd_phase_times(hyd_ctx *ctx, double *out, int32_t reset);
/* RK4 advances taken so far and their node-updates (the sum over advances
   of the node count each one moved), for throughput accounting; no
   reference counterpart.  reset != 0 zeroes both after reading; either
   pointer may be NULL. */
int hyd_advance_counts(hyd_ctx *ctx, int64_t *advances, int64_t *node_updates, int32_t reset);

/* ---- diagnostics records and saves (evolution.f90:72, 111, 581-599) ------ */
/* The rows the reference writes each step and at each regularisation,
   oldest first, removed as they are taken: kind 0 monitor.asc (adapt,
   evolution.f90:581-584): t, cfl = uumax*dt/glx, zzrms**2/2, zzrms, zzmax,
   uumax, twist (7 values a row); kind 1 complexity.asc (evolution.f90:72,
   111): t, nq, nptq (3 values).  Returns the number of rows copied (at
   most max). */
int hyd_take_records(hyd_ctx *ctx, int32_t kind, int32_t max, double *out);
/* The save schedule of adapt (evolution.f90:587-599, caps.f90:206-238
   counters): on != 0 enables it from time tinit with grid record igrec and
   contour record icrec (plain precision only).  Each step's adapt then
   decides gsave (t+dt >= tinit+tgsave*igrec) and csave (t+dt >=
   tinit+tcsave*icrec); with gsave it keeps qs and the energies of that
   inversion.  The saves fall at the start of the next step, after its
   inversion (evolution.f90:193-204): the library computes their contents
   there -- qq and the energies interpolated to tgrid (evolution.f90:
   660-668); the contours and the residual spctop(qq - qc) (evolution.f90:
   750-766) -- and keeps them, in order, for the caller, who writes the
   files; the record counters advance as the reference's. */
int hyd_set_saves(hyd_ctx *ctx, int32_t on, double tinit, int32_t igrec, int32_t icrec);
/* The record counters: the next grid record igrec (0-based, qq.r4 writes
   igrec+1) and the next contour record icrec (1-based). */
int hyd_save_counters(hyd_ctx *ctx, int32_t *igrec, int32_t *icrec);
/* Numbers of grid and contour saves waiting. */
int hyd_pending_saves(hyd_ctx *ctx, int32_t *grid, int32_t *cont);
/* The oldest grid save: info[4] = tgrid, ekin, epot, igrec (0-based record
   before the increment); qqs = the PV interpolated to tgrid (spectral,
   ng*ng).  HYD_ESTATE when none waits. */
int hyd_take_grid_save(hyd_ctx *ctx, double *info, double *qqs);
/* The oldest contour save's sizes, then its contents: info[4] = t, qjump,
   icrec, contex; residual (ng*ng physical, spctop(qq - qc)); the nodes and
   npq, i1q, indq as savecont writes them. */
int hyd_contour_save_size(hyd_ctx *ctx, int32_t *nq, int32_t *nptq);
int hyd_take_contour_save(hyd_ctx *ctx, double *info, double *residual, double *xq, double *yq,
                          int32_t *npq, int32_t *i1q, int32_t *indq);
/* The end of advect (evolution.f90:85-89): grid and/or contour saves of the
   current state, after an inversion (the grid save interpolates to the next
   save time from the last kept qs). */
int hyd_save_now(hyd_ctx *ctx, int32_t grid, int32_t cont);

/* An ensemble of n independent flows (distinct contexts, any sizes and
 * precisions), each advanced by hyd_run's loop for nsteps steps.  The flows'
 * RK4 steps of one round run concurrently on their own streams; each flow's
 * result equals hyd_run on that flow alone.  No reference counterpart: the
 * batched ensemble of many flows (one flow per program in the reference,
 * caps.f90). */
int hyd_run_batch(hyd_ctx **ctxs, int32_t n, int32_t nsteps);

/* Synchronise the context's stream (timing helpers). */
int hyd_sync(hyd_ctx *ctx);
/* Kernel-time accounting of the dominant kernels, measured with HIP events
   on the context's stream: returns the number of records written. */
/* No reference counterpart: the FFT batch (vectors per workgroup) that the
   autotuner picked at hyd_init for the layout classes forward
   strided->rows, forward strided->strided, reverse strided->strided and
   reverse rows->strided (v[0..3]). */
int hyd_get_tuning(hyd_ctx *ctx, int32_t *v);
/* No reference counterpart (measurement aid): one plain streaming copy
   kernel of nbytes (<= 128 ng^2) read and nbytes written, the reference
   point that scales hardware traffic counters. */
int hyd_copy_calibration(hyd_ctx *ctx, int64_t nbytes);
/* No reference counterpart (measurement aid): the average interval the
   per-launch timing events report for an empty kernel launched the same
   way (reps launches): the launch and stamping cost that every record of
   hyd_kernel_times includes. */
int hyd_event_overhead(hyd_ctx *ctx, int32_t reps, double *avg_ms);
/* No reference counterpart: per-launch timing with HIP events on the
   library's stream.  Level 0 off, 1 the single kernels whose records start
   with "k:" (one record per launch), 2 every timed group as well.
   hyd_kernel_times returns the records (names: max*32 bytes). */
int hyd_kernel_times(hyd_ctx *ctx, int32_t max, char *names, double *ms, int64_t *bytes);
int hyd_set_timing(hyd_ctx *ctx, int32_t level);

/* ---- learned velocity correction (no reference counterpart) ------------
 * A 4-layer 3x3 convolutional network (inputs u, v, q on an n x n doubly
 * periodic grid of side ell, 32 channels, GELU; BF16 weights and
 * activations on the matrix cores with FP32 accumulation) whose 2-channel
 * output (du, dv) is sampled bilinearly at node positions: a sub-grid
 * correction to velint's velocities.  Plain bilinear velint stays the
 * fallback.  Inputs in3 are FP32 [ix*n + iy][3] (channel fastest). */
typedef struct hyd_cnn hyd_cnn;
int hyd_cnn_create(int32_t n, double ell, uint64_t seed, int32_t device, hyd_cnn **out);
int hyd_cnn_destroy(hyd_cnn *cnn);
/* corrections at m nodes (host arrays) */
int hyd_cnn_forward(hyd_cnn *cnn, const float *in3, const double *xn, const double *yn,
                    int32_t m, float *du, float *dv);
/* loss mean_i |c_i - t_i|^2 and its gradient w.r.t. every weight (FP32,
 * layer by layer: W[co][ci][tap] then bias[co]) */
int hyd_cnn_gradients(hyd_cnn *cnn, const float *in3, const double *xn, const double *yn,
                      int32_t m, const float *tu, const float *tv, float *grad, float *loss);
/* one Adam step (lr) on that loss; *loss is the loss before the step */
int hyd_cnn_train_step(hyd_cnn *cnn, const float *in3, const double *xn, const double *yn,
                       int32_t m, const float *tu, const float *tv, float lr, float *loss);
int hyd_cnn_num_weights(hyd_cnn *cnn, int64_t *nw);
int hyd_cnn_get_weights(hyd_cnn *cnn, float *w);
int hyd_cnn_set_weights(hyd_cnn *cnn, const float *w);
/* One 16x16x32 BF16 matrix-core product D = A B (A [16][32] row-major, B
 * column by column [16][32], D [16][16] row-major; 512 BF16 each, 256
 * FP32): the operand layout check of the network's kernels. */
int hyd_cnn_mfma_probe(const uint16_t *a, const uint16_t *b, float *d);
/* Attach a trained network to a flow (NULL detaches): every RK4 stage then
 * adds the network's (du, dv) at the nodes to velint's velocity.  Inputs
 * are the stage's u/su, v/su and PV q/sq on the context's grid (the
 * network's n must be ng), the output is scaled by sd.  The step is then no
 * longer the reference's; plain velint is the default.  One network per
 * running flow (its work buffers are used on the flow's stream). */
int hyd_set_correction(hyd_ctx *ctx, hyd_cnn *net, double su, double sq, double sd);
/* The velocity that the PV of the contours (xq, yq, nextq, qjump) has on
 * this context's grid -- con2grid, ptospc, main_invert (contours.f90:
 * 173-346, spectral.f90:304-342) -- at m points (xs, ys) by velint
 * (contours.f90:136-169); the context's own state is not touched.  On a
 * grid 4x finer than a flow's, this is the judge of the flow's velint and
 * the target of the learned correction.  fields (may be NULL) receives u,
 * v and the PV grid as FP32 [ng*ng][3]. */
int hyd_contour_velocity(hyd_ctx *ctx, int32_t nptq, const double *xq, const double *yq,
                         const int32_t *nextq, double qjump, int32_t m, const double *xs,
                         const double *ys, double *us, double *vs, float *fields);

#ifdef __cplusplus
}
#endif
#endif
