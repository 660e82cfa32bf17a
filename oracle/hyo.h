/*
 * hyo.h -- the oracle: a plain-C, operation-order-faithful restatement of the
 * Hydra contour-advection step (ca/plane/qg/caps + lib/stafft).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (hydra_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Conventions kept from the reference so that results are bit-comparable:
 *   - physical fields are (ny,nx) column-major, iy fastest:  f[(ix-1)*ny + iy-1]
 *   - spectral fields are (nx,ny) column-major, kx fastest:   s[(ky-1)*nx + kx-1]
 *     in the stafft "half-complex x half-complex" layout
 *   - contour index arrays (nextq, i1q, i2q) hold 1-based node numbers
 *   - every routine evaluates its expressions in the reference's order;
 *     the library is built with -ffp-contract=off (no FMA contraction)
 *
 * Pinning status of each routine is listed in the header comment of the
 * .c file that holds it.
 */
#ifndef HYO_H
#define HYO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYO_OK 0
#define HYO_ERR_FACTOR 1      /* FFT length not a product of 2 and 4      */
#define HYO_ERR_CAPACITY 2    /* node / contour capacity exceeded          */
#define HYO_ERR_ARG 3         /* bad argument                              */
#define HYO_ERR_ALLOC 4

/* Run-time copy of the compile-time parameters of parameters.f90:14-36. */
typedef struct hyo_params {
  int nx, ny;          /* inversion grid                                 */
  int ncontq;          /* number of PV contour intervals                 */
  double ellx, elly;   /* domain lengths                                 */
  double kd, beta;     /* deformation wavenumber, planetary PV gradient  */
  double rtherm, rekman;
  double cdamp, nnu;   /* hyperdiffusion prefactor and power             */
  double tgsave, tcsave;
  double dt_fixed;     /* >0: fixed time step (the only patch vs adapt)  */
  int64_t npm;         /* node capacity (0 -> default)                   */
  double ell_cells;    /* ell in grid lengths (0 -> 6.25, contours.f90:38) */
} hyo_params;

typedef struct hyo_model {
  hyo_params p;
  /* constants.f90 */
  int nx, ny, nxf, nyf, nxu, nyu, mgf, mgu, nlevm;
  double hlx, hlxi, xmin, xmax, hly, hlyi, ymin, ymax;
  double kdsq, srwfm, therm, domarea, aspect;
  double glx, glxi, gly, glyi, garea, dsumi;
  int beffect, btropic, heating, friction;
  /* contours.f90:17-42 */
  double glxf, glxfi, glyf, glyfi, glxu, glxui, glyu, glyui;
  double amu, ell, dm, dm2, d4small, dmsq, dmi, elf, densf;
  int *ixp, *iyp;
  double *xgf, *ygf, *xgu, *ygu;
  double w00[256], w10[256], w01[256], w11[256];
  int *ixfw, *ix0w, *ix1w, *iyfw, *iy0w, *iy1w, *ibx, *iby;
  /* spectral.f90:14-32 */
  double *rksq, *qdiss, *green, *flo, *fhi, *rk;
  double *rkx, *hrkx, *rky, *hrky, *bety;
  double *xtrig, *ytrig;
  int xfactors[5], yfactors[5];
  double *spmf, *alk;
  int *kmag, kmax, kmaxred;
  /* contour store, contours.f90:13-15 */
  int64_t npm, nm;
  int nq, nptq;
  double qjump;
  double *xq, *yq;
  int *nextq, *indq, *npq, *i1q, *i2q;
  /* evolution / common state, evolution.f90:10-16, common.f90:13-22 */
  double *qs, *qd, *qc, *qq, *pp, *uu, *vv, *qr, *emq, *epq, *qspre;
  double t, dt, dt2, dt3, dt6, twist, dtmax, tinit;
  int contex;
  int ireg;
  /* where a step-count run resumes: 0 = contours to rebuild (recont),
   * 1 = contours built, advect's start to run, 2 = inside advect's loop */
  int phase;
  /* diagnostics of the last adapt call */
  double zzrms, zzmax, uumax, dtacc;
  int err;
} hyo_model;

/* ---- lifecycle ---------------------------------------------------------- */
hyo_model *hyo_create(const hyo_params *p);
void hyo_destroy(hyo_model *m);
int hyo_error(const hyo_model *m);

/* ---- stafft.f90 ---------------------------------------------------------- */
int hyo_factorisen(int n, int factors[5]);
int hyo_initfft(int n, int factors[5], double *trig);
void hyo_forfft(int m, int n, double *x, const double *trig, const int factors[5]);
void hyo_revfft(int m, int n, double *x, const double *trig, const int factors[5]);

/* ---- sta2dfft.f90 (doubly periodic part) --------------------------------- */
void hyo_ptospc(const hyo_model *m, double *rvar, double *svar);
void hyo_spctop(const hyo_model *m, double *svar, double *rvar);
void hyo_xderiv(int nx, int ny, const double *rkx, const double *var, double *der);
void hyo_yderiv(int nx, int ny, const double *rky, const double *var, double *der);

/* ---- spectral.f90 -------------------------------------------------------- */
void hyo_main_invert(const hyo_model *m, const double *qq, double *uu, double *vv, double *pp);
void hyo_gradient(const hyo_model *m, const double *ff, double *ffx, double *ffy);
void hyo_spec1d(const hyo_model *m, double *ss, double *spec, int iopt);

/* ---- contours.f90 -------------------------------------------------------- */
void hyo_velint(const hyo_model *m, const double *uu, const double *vv, double *uq, double *vq);
void hyo_con2grid(const hyo_model *m, double *qc);
/* con2grid returning the integer level counts of the 4ng fine grid as well
   (round(qa/qjump), rows 1..nyf) -- the exact integer contract. */
void hyo_con2grid_counts(const hyo_model *m, double *qc, int32_t *counts);
int hyo_surgery(hyo_model *m);
int hyo_renode_cap(const hyo_model *m, const double *xd, const double *yd, int npd,
                   double *xr, double *yr, int64_t cap, int *npr);
int hyo_renode(const hyo_model *m, const double *xd, const double *yd, int npd,
               double *xr, double *yr, int *npr);

/* ---- congen.f90 ---------------------------------------------------------- */
int hyo_recontour(hyo_model *m, double *qq);
void hyo_advect_start(hyo_model *m);

/* ---- evolution.f90 ------------------------------------------------------- */
int hyo_init_from_field(hyo_model *m, const double *qq_init, double t0);
void hyo_evolution_init(hyo_model *m);
void hyo_prepare(hyo_model *m);
void hyo_inversion(hyo_model *m);
void hyo_adapt(hyo_model *m);
void hyo_source(hyo_model *m, double *sqs, double *sqd, int lev);
int hyo_advance(hyo_model *m);
void hyo_energy(const hyo_model *m, double *ekin, double *epot);
double hyo_l1norm(const hyo_model *m, const double *qq);
double hyo_l2norm(const hyo_model *m, const double *qq);
/* One pass of the advect loop body (evolution.f90:55-80): surgery when
   twist>twistmax, then advance.  Returns 1 when nregmax regularisations were
   reached and prepare() has been called (caller must recontour). */
int hyo_advect_step(hyo_model *m);
/* Full cycle helper: recontour from qr, evolution_init, then nsteps advect
   steps (recontouring whenever advect asks for it). */
int hyo_run(hyo_model *m, int nsteps);

/* ---- generic.f90 --------------------------------------------------------- */
double hyo_sum_seq(const double *a, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
