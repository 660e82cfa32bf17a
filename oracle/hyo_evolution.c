/*
 * hyo_evolution.c -- oracle restatement of evolution.f90 (advect 28-92,
 * init 96-131, prepare 135-162, advance 166-326, source 432-490,
 * inversion 494-520, adapt 524-608, energy 612-644) and of the loopinit=0
 * branch of caps.f90 initialise (143-183).
 *
 * TEST INFRASTRUCTURE ONLY (see hyo.h).
 *
 * Differences from the reference program, all of them outside the state
 * trajectory:
 *   - no file output (savegrid/savecont/monitor): they read the state only;
 *   - dt_fixed > 0 replaces dt=min(dtacc,dtmax) in adapt (the fixed-step
 *     patch); every other statement of adapt still runs;
 *   - stochastic and narrow-band forcing and thermal damping are not
 *     supported (ivor=0, fbase=0, rtherm=0 in every configuration).
 *
 * Pinning (evolution.f90 itself does not compile here, so these routines
 * cannot be run against the reference build): adapt, source and energy
 * against closed forms on one- and two-mode fields; advance against
 * textbook RK4 (no damping) and Lawson's integrating-factor RK4 driven
 * stage by stage through the oracle's own right-hand side; init, prepare
 * and inversion as compositions of the pinned con2grid, ptospc, spctop and
 * main_invert.  The RHS pieces those use (con2grid, velint, FFTs,
 * main_invert, gradient) are compared with the reference build.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hyo.h"

static const double small = 1.0e-12;
static const double oms = 1.0 - 1.0e-12;
#define FINT(x) ((int)(x))
#define PI_D 3.141592653589793238462643383279502884197169399375105820974944592307816

static double fmax_f(double a, double b) { return a > b ? a : b; }
static double fmin_f(double a, double b) { return a < b ? a : b; }

/* caps.f90:143-183 (loopinit = 0). qq_init(ny,nx) is the full PV field. */
int hyo_init_from_field(hyo_model *m, const double *qq_init, double t0) {
  long N = (long)m->nx * m->ny;
  if (m->heating) return m->err = HYO_ERR_ARG;
  memcpy(m->qr, qq_init, sizeof(double) * N);
  if (m->beffect) {
    for (int ix = 0; ix < m->nx; ix++)
      for (int iy = 0; iy < m->ny; iy++) m->qr[(long)ix * m->ny + iy] -= m->bety[iy];
    m->qjump = m->p.beta * m->p.elly / (double)m->p.ncontq;
  } else {
    /* generic.f90:66-79 contint */
    double mx = m->qr[0], mn = m->qr[0];
    for (long k = 0; k < N; k++) {
      mx = fmax_f(mx, m->qr[k]);
      mn = fmin_f(mn, m->qr[k]);
    }
    m->qjump = (mx - mn) / (double)m->p.ncontq;
  }
  double *ff = (double *)malloc(sizeof(double) * N);
  memcpy(ff, m->qr, sizeof(double) * N);
  hyo_ptospc(m, ff, m->qs);
  free(ff);
  m->nq = 0;
  m->nptq = 0;
  m->dt = 0.0;
  m->tinit = t0;
  m->t = t0;
  m->twist = 0.0;
  m->phase = 0;
  return HYO_OK;
}

/* evolution.f90:96-131 */
void hyo_evolution_init(hyo_model *m) {
  long N = (long)m->nx * m->ny;
  double *qa = (double *)malloc(sizeof(double) * N);
  m->contex = m->nptq > 0;
  hyo_con2grid(m, qa);
  hyo_ptospc(m, qa, m->qc);
  for (long k = 0; k < N; k++) {
    m->qd[k] = m->fhi[k] * (m->qs[k] - m->qc[k]);
    m->qspre[k] = m->qs[k];
  }
  free(qa);
}

/* evolution.f90:135-162 */
void hyo_prepare(hyo_model *m) {
  long N = (long)m->nx * m->ny;
  double *qa = (double *)malloc(sizeof(double) * N);
  hyo_con2grid(m, qa);
  hyo_ptospc(m, qa, m->qc);
  for (long k = 0; k < N; k++) {
    m->qs[k] = m->flo[k] * (m->qs[k] - m->qc[k]) + m->qc[k] + m->qd[k];
    m->qd[k] = m->qs[k] - m->qc[k];
  }
  hyo_spctop(m, m->qd, m->qr);
  free(qa);
  m->phase = 0;
}

/* Start of advect, evolution.f90:44-51: init, twist = 0, no regularisations
 * counted yet; the run is then inside advect's loop. */
void hyo_advect_start(hyo_model *m) {
  hyo_evolution_init(m);
  m->twist = 0.0;
  m->ireg = 0;
  m->phase = 2;
}

/* evolution.f90:494-520 */
void hyo_inversion(hyo_model *m) {
  long N = (long)m->nx * m->ny;
  double *qa = (double *)malloc(sizeof(double) * N);
  hyo_con2grid(m, qa);
  hyo_ptospc(m, qa, m->qc);
  for (long k = 0; k < N; k++) m->qq[k] = m->flo[k] * (m->qs[k] - m->qc[k]) + m->qc[k] + m->qd[k];
  hyo_main_invert(m, m->qq, m->uu, m->vv, m->pp);
  free(qa);
}

/* evolution.f90:524-608 (file output omitted; dt_fixed patch) */
void hyo_adapt(hyo_model *m) {
  long N = (long)m->nx * m->ny;
  const double dtfac = PI_D / 10.0, cflmax = 0.7;
  double *ss = (double *)malloc(sizeof(double) * N);
  double *zz = (double *)malloc(sizeof(double) * N);
  if (m->btropic) memcpy(ss, m->qq, sizeof(double) * N);
  else
    for (long k = 0; k < N; k++) ss[k] = m->qq[k] + m->kdsq * m->pp[k];
  hyo_spctop(m, ss, zz);
  for (long k = 0; k < N; k++) zz[k] = zz[k] * zz[k];
  double zzrms = sqrt(m->dsumi * hyo_sum_seq(zz, N));
  double mx = zz[0];
  for (long k = 1; k < N; k++) mx = fmax_f(mx, zz[k]);
  double zzmax = sqrt(mx) + small;
  for (long k = 0; k < N; k++) zz[k] = m->uu[k] * m->uu[k] + m->vv[k] * m->vv[k];
  mx = zz[0];
  for (long k = 1; k < N; k++) mx = fmax_f(mx, zz[k]);
  double uumax = sqrt(mx) + small;
  double dtacc = fmin_f(m->glx * cflmax / uumax, dtfac / fmax_f(zzmax, m->srwfm));
  if (m->p.dt_fixed > 0.0) {
    m->dt = m->p.dt_fixed;
  } else if (m->dt > 0.0) {
    m->dt = fmin_f(dtacc, m->dtmax);
  } else {
    m->dtmax = 0.2 * fmin_f(m->p.tgsave, m->p.tcsave);
    m->dt = fmin_f(dtacc, m->dtmax);
  }
  m->dt2 = m->dt * 0.5;
  m->dt3 = m->dt * (1.0 / 3.0);
  m->dt6 = m->dt * (1.0 / 6.0);
  m->twist = m->twist + m->dt * zzmax;
  double dfac = m->dt2 * zzrms;
  for (long k = 0; k < N; k++) {
    m->epq[k] = exp(dfac * m->qdiss[k]);
    m->emq[k] = 1.0 / m->epq[k];
  }
  m->zzrms = zzrms;
  m->zzmax = zzmax;
  m->uumax = uumax;
  m->dtacc = dtacc;
  free(ss);
  free(zz);
}

/* evolution.f90:432-490 (heating and forcing branches not supported) */
void hyo_source(hyo_model *m, double *sqs, double *sqd, int lev) {
  long N = (long)m->nx * m->ny;
  double *qqx = (double *)malloc(sizeof(double) * N);
  double *qqy = (double *)malloc(sizeof(double) * N);
  hyo_gradient(m, m->qd, qqx, qqy);
  for (long k = 0; k < N; k++) qqx[k] = -m->uu[k] * qqx[k] - m->vv[k] * qqy[k];
  hyo_ptospc(m, qqx, sqd);
  hyo_gradient(m, m->qs, qqx, qqy);
  if (m->beffect)
    for (long k = 0; k < N; k++) qqx[k] = -m->uu[k] * qqx[k] - m->vv[k] * (qqy[k] + m->p.beta);
  else
    for (long k = 0; k < N; k++) qqx[k] = -m->uu[k] * qqx[k] - m->vv[k] * qqy[k];
  hyo_ptospc(m, qqx, sqs);
  for (long k = 0; k < N; k++) sqd[k] = sqd[k] - m->p.rekman * (m->qq[k] + m->kdsq * m->pp[k]);
  free(qqx);
  free(qqy);
  if (lev == 0) return;
  if (lev == 1)
    for (long k = 0; k < N; k++) sqd[k] = m->epq[k] * sqd[k];
  else
    for (long k = 0; k < N; k++) sqd[k] = m->epq[k] * m->epq[k] * sqd[k];
}

#define WRAPX(xx) (oms * ((xx)-ellx * (double)FINT((xx)*hlxi)))
#define WRAPY(yy) (oms * ((yy)-elly * (double)FINT((yy)*hlyi)))

/* evolution.f90:166-326 */
int hyo_advance(hyo_model *m) {
  long N = (long)m->nx * m->ny;
  int n = m->nptq;
  double ellx = m->p.ellx, elly = m->p.elly, hlxi = m->hlxi, hlyi = m->hlyi;
  double *buf = (double *)malloc(sizeof(double) * N * 6);
  double *qsi = buf, *qsf = buf + N, *sqs = buf + 2 * N, *qdi = buf + 3 * N, *qdf = buf + 4 * N,
         *sqd = buf + 5 * N;
  double *nb = (double *)malloc(sizeof(double) * ((size_t)n * 6 + 1));
  double *xqi = nb, *yqi = nb + n, *xqf = nb + 2 * n, *yqf = nb + 3 * n, *uq = nb + 4 * n,
         *vq = nb + 5 * n;
  double *xq = m->xq, *yq = m->yq;
  double *qs = m->qs, *qd = m->qd, *qc = m->qc;
  /* stage 1 */
  hyo_inversion(m);
  for (long k = 0; k < N; k++) {
    qs[k] = m->flo[k] * (qs[k] - qc[k]) + qc[k] + qd[k];
    qd[k] = m->fhi[k] * (qs[k] - qc[k]);
  }
  hyo_adapt(m);
  double dt = m->dt, dt2 = m->dt2, dt3 = m->dt3, dt6 = m->dt6;
  hyo_source(m, sqs, sqd, 0);
  if (m->contex) {
    hyo_velint(m, m->uu, m->vv, uq, vq);
    for (int i = 0; i < n; i++) {
      xqi[i] = xq[i];
      yqi[i] = yq[i];
      double xx = xqi[i] + dt2 * uq[i], yy = yqi[i] + dt2 * vq[i];
      xq[i] = WRAPX(xx);
      yq[i] = WRAPY(yy);
      xx = xqi[i] + dt6 * uq[i];
      yy = yqi[i] + dt6 * vq[i];
      xqf[i] = WRAPX(xx);
      yqf[i] = WRAPY(yy);
    }
  }
  for (long k = 0; k < N; k++) {
    qsi[k] = qs[k];
    qs[k] = qsi[k] + dt2 * sqs[k];
    qsf[k] = qsi[k] + dt6 * sqs[k];
    qdi[k] = qd[k];
    qd[k] = m->emq[k] * (qdi[k] + dt2 * sqd[k]);
    qdf[k] = qdi[k] + dt6 * sqd[k];
  }
  /* stage 2 */
  m->t = m->t + dt2;
  hyo_inversion(m);
  hyo_source(m, sqs, sqd, 1);
  if (m->contex) {
    hyo_velint(m, m->uu, m->vv, uq, vq);
    for (int i = 0; i < n; i++) {
      double xx = xqi[i] + dt2 * uq[i], yy = yqi[i] + dt2 * vq[i];
      xq[i] = WRAPX(xx);
      yq[i] = WRAPY(yy);
      xx = xqf[i] + dt3 * uq[i];
      yy = yqf[i] + dt3 * vq[i];
      xqf[i] = WRAPX(xx);
      yqf[i] = WRAPY(yy);
    }
  }
  for (long k = 0; k < N; k++) {
    qs[k] = qsi[k] + dt2 * sqs[k];
    qsf[k] = qsf[k] + dt3 * sqs[k];
    qd[k] = m->emq[k] * (qdi[k] + dt2 * sqd[k]);
    qdf[k] = qdf[k] + dt3 * sqd[k];
  }
  /* stage 3 */
  hyo_inversion(m);
  hyo_source(m, sqs, sqd, 1);
  if (m->contex) {
    hyo_velint(m, m->uu, m->vv, uq, vq);
    for (int i = 0; i < n; i++) {
      double xx = xqi[i] + dt * uq[i], yy = yqi[i] + dt * vq[i];
      xq[i] = WRAPX(xx);
      yq[i] = WRAPY(yy);
      xx = xqf[i] + dt3 * uq[i];
      yy = yqf[i] + dt3 * vq[i];
      xqf[i] = WRAPX(xx);
      yqf[i] = WRAPY(yy);
    }
  }
  for (long k = 0; k < N; k++) {
    qs[k] = qsi[k] + dt * sqs[k];
    qsf[k] = qsf[k] + dt3 * sqs[k];
    m->emq[k] = m->emq[k] * m->emq[k];
    qd[k] = m->emq[k] * (qdi[k] + dt * sqd[k]);
    qdf[k] = qdf[k] + dt3 * sqd[k];
  }
  /* stage 4 */
  m->t = m->t + dt2;
  hyo_inversion(m);
  hyo_source(m, sqs, sqd, 2);
  if (m->contex) {
    hyo_velint(m, m->uu, m->vv, uq, vq);
    for (int i = 0; i < n; i++) {
      double xx = xqf[i] + dt6 * uq[i], yy = yqf[i] + dt6 * vq[i];
      xq[i] = WRAPX(xx);
      yq[i] = WRAPY(yy);
    }
  }
  for (long k = 0; k < N; k++) {
    qs[k] = qsf[k] + dt6 * sqs[k];
    qd[k] = m->emq[k] * (qdf[k] + dt6 * sqd[k]);
  }
  free(buf);
  free(nb);
  return m->err;
}

/* evolution.f90:612-644 */
void hyo_energy(const hyo_model *m, double *ekin, double *epot) {
  long N = (long)m->nx * m->ny;
  double *zz = (double *)malloc(sizeof(double) * N);
  for (long k = 0; k < N; k++) zz[k] = m->uu[k] * m->uu[k] + m->vv[k] * m->vv[k];
  *ekin = 0.5 * m->garea * hyo_sum_seq(zz, N);
  if (m->btropic) {
    *epot = 0.0;
  } else {
    double *ss = (double *)malloc(sizeof(double) * N);
    memcpy(ss, m->pp, sizeof(double) * N);
    hyo_spctop(m, ss, zz);
    for (long k = 0; k < N; k++) zz[k] = zz[k] * zz[k];
    *epot = 0.5 * m->garea * m->kdsq * hyo_sum_seq(zz, N);
    free(ss);
  }
  free(zz);
}

/* generic.f90:30-46 and 48-64: garea * sum(|qq|) and garea * sum(qq**2),
 * sequential sums over the (ny,nx) field in storage order. */
double hyo_l1norm(const hyo_model *m, const double *qq) {
  const long N = (long)m->nx * m->ny;
  double *z = (double *)malloc(sizeof(double) * N);
  for (long k = 0; k < N; k++) z[k] = fabs(qq[k]);
  const double r = m->garea * hyo_sum_seq(z, N);
  free(z);
  return r;
}

double hyo_l2norm(const hyo_model *m, const double *qq) {
  const long N = (long)m->nx * m->ny;
  double *z = (double *)malloc(sizeof(double) * N);
  for (long k = 0; k < N; k++) z[k] = qq[k] * qq[k];
  const double r = m->garea * hyo_sum_seq(z, N);
  free(z);
  return r;
}

/* evolution.f90:55-80, one iteration of the time loop. */
int hyo_advect_step(hyo_model *m) {
  const double twistmax = 2.5;
  const int nregmax = 20;
  if (m->twist > twistmax) {
    m->ireg = m->ireg + 1;
    if (m->ireg == nregmax) {
      hyo_prepare(m);
      return 1;
    }
    if (hyo_surgery(m)) return -1;
    m->twist = m->twist - twistmax;
  }
  if (hyo_advance(m)) return -1;
  return 0;
}

/* caps.f90:83-91 loop driven by a step count instead of tsim.  A run
 * resumes where the previous one stopped (phase): inside advect's loop it
 * continues the loop, so run(a) then run(b) equals run(a + b); recont
 * (recontour of qr) happens only at the start of a cycle. */
int hyo_run(hyo_model *m, int nsteps) {
  int done = 0;
  while (done < nsteps) {
    if (m->phase == 0 && hyo_recontour(m, m->qr)) return -1;
    if (m->phase <= 1) hyo_advect_start(m);
    while (done < nsteps) {
      int r = hyo_advect_step(m);
      if (r < 0) return -1;
      if (r == 1) break;
      done++;
    }
  }
  return 0;
}
