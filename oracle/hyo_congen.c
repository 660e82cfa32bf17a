/*
 * hyo_congen.c -- oracle restatement of congen.f90: recontour (24-128),
 * ugrid2con (132-334) and con2ugrid (338-441).
 *
 * TEST INFRASTRUCTURE ONLY (see hyo.h).
 *
 * Pinning: a single circular patch recontours into one contour per level
 * at the analytic radii (closed form, to the ultra-fine grid spacing);
 * contour count / node count / positions bit-for-bit vs oracle/_ref.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hyo.h"

static const double small = 1.0e-12;
static const double oms = 1.0 - 1.0e-12;
#define FINT(x) ((int)(x))

/* congen.f90:338-441, qa(nyu+1,nxu) column-major */
static void con2ugrid(const hyo_model *m, double *qa) {
  int nxu = m->nxu, nyu = m->nyu, nptq = m->nptq;
  long lda = nyu + 1;
#define QA(iy, ix) qa[(long)((ix)-1) * lda + (iy)-1]
  double ellx = m->p.ellx, elly = m->p.elly, hlx = m->hlx, hly = m->hly;
  double hlxi = m->hlxi, hlyi = m->hlyi, xmin = m->xmin, ymin = m->ymin;
  double qjump = m->qjump;
  const double *xq = m->xq, *yq = m->yq;
  double *qjx = (double *)calloc(nxu, sizeof(double));
  double *dx = (double *)malloc(sizeof(double) * (nptq + 1));
  double *dy = (double *)malloc(sizeof(double) * (nptq + 1));
  int *ixc = (int *)malloc(sizeof(int) * (nptq + 1));
  int *nxc = (int *)malloc(sizeof(int) * (nptq + 1));
  for (int i = 0; i < nptq; i++) ixc[i] = FINT(m->glxui * (xq[i] - xmin));
  for (int i = 0; i < nptq; i++) {
    int ia = m->nextq[i] - 1;
    double xx = xq[ia] - xq[i];
    dx[i] = xx - ellx * (double)FINT(xx * hlxi);
    double yy = yq[ia] - yq[i];
    dy[i] = yy - elly * (double)FINT(yy * hlyi);
    int ixdif = ixc[ia] - ixc[i];
    nxc[i] = ixdif - nxu * ((2 * ixdif) / nxu);
    if (fabs(yy) > hly) {
      double cc = copysign(1.0, dy[i]);
      double p = -(yq[i] - hly * cc) / (dy[i] + small);
      int ix = 1 + FINT(m->glxui * (fmod(ellx + hlx + xq[i] + p * dx[i], ellx)));
      qjx[ix - 1] = qjx[ix - 1] - qjump * cc;
    }
  }
  QA(1, 1) = 0.0;
  for (int ix = 1; ix <= nxu - 1; ix++) QA(1, ix + 1) = QA(1, ix) + qjx[ix - 1];
  for (int ix = 1; ix <= nxu; ix++)
    for (int iy = 2; iy <= nyu + 1; iy++) QA(iy, ix) = 0.0;
  for (int i = 0; i < nptq; i++) {
    if (nxc[i] != 0) {
      int jump = nxc[i] >= 0 ? 1 : -1;
      int ixbeg = ixc[i] + (1 + jump) / 2 + nxu;
      double sqjump = qjump * (double)jump;
      int ncr = 0;
      while (ncr != nxc[i]) {
        int ix = 1 + (ixbeg + ncr) % nxu;
        double xx = m->xgu[ix - 1] - xq[i];
        double px0 = (xx - ellx * (double)FINT(xx * hlxi)) / dx[i];
        double yy = yq[i] + px0 * dy[i];
        int iy = 2 + FINT(m->glyui * (yy - elly * (double)FINT(yy * hlyi) - ymin));
        QA(iy, ix) = QA(iy, ix) + sqjump;
        ncr = ncr + jump;
      }
    }
  }
  for (int ix = 1; ix <= nxu; ix++)
    for (int iy = 2; iy <= nyu; iy++) QA(iy, ix) = QA(iy, ix) + QA(iy - 1, ix);
  double qavg = 0.0;
  for (int ix = 1; ix <= nxu; ix++)
    for (int iy = 1; iy <= nyu; iy++) qavg = qavg + QA(iy, ix);
  qavg = qavg / (double)((long)nxu * nyu);
  for (int ix = 1; ix <= nxu; ix++)
    for (int iy = 1; iy <= nyu; iy++) QA(iy, ix) = QA(iy, ix) - qavg;
#undef QA
  free(qjx); free(dx); free(dy); free(ixc); free(nxc);
}

/* congen.f90:132-334.  Appends the new contours to (xa,ya,...) of the
   recontour work store. */
typedef struct {
  double *xa, *ya;
  int *inda, *npa, *i1a, *i2a;
  int na, npta;
  int64_t npm, nm;
} work_store;

static int ugrid2con(hyo_model *m, const double *qa, work_store *w) {
  int nxu = m->nxu, nyu = m->nyu, nlevm = m->nlevm;
  long lda = nyu + 1;
#define QA(iy, ix) qa[(long)((ix)-1) * lda + (iy)-1]
  double elly = m->p.elly, ellx = m->p.ellx, hlxi = m->hlxi, hlyi = m->hlyi;
  double glxu = m->glxu, glyu = m->glyu, qjump = m->qjump;
  int64_t nxny = (int64_t)nxu * nyu;
  int status = HYO_OK;
  double qjumpi = 1.0 / qjump;
  double qoff = qjump * (double)nlevm;
  double qamax = QA(1, 1), qamin = QA(1, 1);
  for (int ix = 1; ix <= nxu; ix++)
    for (int iy = 1; iy <= nyu; iy++) {
      double v = QA(iy, ix);
      /* Fortran max/min */
      qamax = v > qamax ? v : qamax;
      qamin = v < qamin ? v : qamin;
    }
  int levbeg = FINT((qoff + qamin) * qjumpi + 0.5) + 1;
  int levend = FINT((qoff + qamax) * qjumpi + 0.5);
  if (levbeg > levend) return HYO_OK;
  int64_t ncrm = 4 * nxny / 8 + 1024; /* grown on demand */
  int64_t *kib = (int64_t *)malloc(sizeof(int64_t) * (ncrm + 1));
  double *xcr = (double *)malloc(sizeof(double) * (ncrm + 1));
  double *ycr = (double *)malloc(sizeof(double) * (ncrm + 1));
  char *freec = (char *)malloc(ncrm + 1);
  double *qdx = (double *)malloc(sizeof(double) * (nxu + 2));
  double *qdy = (double *)malloc(sizeof(double) * (nyu + 2));
  int *isx = (int *)malloc(sizeof(int) * (nxu + 2));
  int *isy = (int *)malloc(sizeof(int) * (nyu + 2));
  int32_t *icrtab = (int32_t *)malloc(sizeof(int32_t) * 2 * nxny);
  signed char *noctab = (signed char *)malloc(nxny + 1);
  int64_t nprm = m->npm;
  double *xd = (double *)malloc(sizeof(double) * (nprm + 1));
  double *yd = (double *)malloc(sizeof(double) * (nprm + 1));
#define ICR(k, c) icrtab[((int64_t)(c)-1) * nxny + (k)-1]
#define NOC(k) noctab[(k)-1]
  for (int lev = levbeg; lev <= levend; lev++) {
    int levt = lev - nlevm + (lev - 1) / nlevm - 1;
    int64_t ncr = 0;
    double qtmp = ((double)lev - 0.5) * qjump - qoff;
    memset(noctab, 0, nxny);
    for (int ix = 1; ix <= nxu; ix++) {
      double xgt = m->xgu[ix - 1];
      for (int iy = 1; iy <= nyu; iy++) {
        qdy[iy] = QA(iy, ix) - qtmp;
        isy[iy] = (int)copysign(1.0, qdy[iy]);
      }
      qdy[nyu + 1] = qdy[1];
      isy[nyu + 1] = isy[1];
      for (int iy = 1; iy <= nyu; iy++) {
        if (isy[iy] != isy[iy + 1]) {
          ncr = ncr + 1;
          if (ncr > ncrm) {
            ncrm *= 2;
            kib = (int64_t *)realloc(kib, sizeof(int64_t) * (ncrm + 1));
            xcr = (double *)realloc(xcr, sizeof(double) * (ncrm + 1));
            ycr = (double *)realloc(ycr, sizeof(double) * (ncrm + 1));
            freec = (char *)realloc(freec, ncrm + 1);
          }
          int inc = (1 - isy[iy]) / 2;
          int64_t kaa = (int64_t)(iy - 1) * nxu + 1;
          kib[ncr] = kaa + m->ibx[ix + inc - 1];
          int64_t kob = kaa + m->ibx[ix + 1 - inc - 1];
          NOC(kob) = NOC(kob) + 1;
          ICR(kob, NOC(kob)) = (int32_t)ncr;
          xcr[ncr] = xgt;
          double yy = m->ygu[iy - 1] - glyu * qdy[iy] / (qdy[iy + 1] - qdy[iy]);
          ycr[ncr] = oms * (yy - elly * (double)FINT(yy * hlyi));
        }
      }
    }
    for (int iy = 1; iy <= nyu; iy++) {
      double ygt = m->ygu[iy - 1];
      for (int ix = 1; ix <= nxu; ix++) {
        qdx[ix] = QA(iy, ix) - qtmp;
        isx[ix] = (int)copysign(1.0, qdx[ix]);
      }
      qdx[nxu + 1] = qdx[1];
      isx[nxu + 1] = isx[1];
      for (int ix = 1; ix <= nxu; ix++) {
        if (isx[ix] != isx[ix + 1]) {
          ncr = ncr + 1;
          if (ncr > ncrm) {
            ncrm *= 2;
            kib = (int64_t *)realloc(kib, sizeof(int64_t) * (ncrm + 1));
            xcr = (double *)realloc(xcr, sizeof(double) * (ncrm + 1));
            ycr = (double *)realloc(ycr, sizeof(double) * (ncrm + 1));
            freec = (char *)realloc(freec, ncrm + 1);
          }
          int inc = (1 - isx[ix]) / 2;
          kib[ncr] = (int64_t)nxu * m->iby[iy + 1 - inc - 1] + ix;
          int64_t kob = (int64_t)nxu * m->iby[iy + inc - 1] + ix;
          NOC(kob) = NOC(kob) + 1;
          ICR(kob, NOC(kob)) = (int32_t)ncr;
          ycr[ncr] = ygt;
          double xx = m->xgu[ix - 1] - glxu * qdx[ix] / (qdx[ix + 1] - qdx[ix]);
          xcr[ncr] = oms * (xx - ellx * (double)FINT(xx * hlxi));
        }
      }
    }
    for (int64_t icr = 1; icr <= ncr; icr++) freec[icr] = 1;
    for (int64_t icr = 1; icr <= ncr; icr++) {
      if (!freec[icr]) continue;
      w->na = w->na + 1;
      if (w->na > w->nm) { status = HYO_ERR_CAPACITY; goto out; }
      int na = w->na;
      w->inda[na - 1] = levt;
      int ibeg = w->npta + 1;
      w->i1a[na - 1] = ibeg;
      int npd = 1;
      xd[0] = xcr[icr];
      yd[0] = ycr[icr];
      int64_t k = kib[icr];
      int noc = NOC(k);
      int64_t icrn = ICR(k, noc);
      while (icrn != icr) {
        NOC(k) = (signed char)(noc - 1);
        npd = npd + 1;
        if (npd > nprm) { status = HYO_ERR_CAPACITY; goto out; }
        xd[npd - 1] = xcr[icrn];
        yd[npd - 1] = ycr[icrn];
        freec[icrn] = 0;
        k = kib[icrn];
        noc = NOC(k);
        icrn = ICR(k, noc);
      }
      int keep = 0, npa = 0;
      double *xo = w->xa + (ibeg - 1), *yo = w->ya + (ibeg - 1);
      const int64_t capo = w->npm - ((int64_t)ibeg - 1);
      do {
        if ((status = hyo_renode_cap(m, xd, yd, npd, xo, yo, capo, &npa))) goto out;
        if (npa == 0) break;
        if ((status = hyo_renode_cap(m, xo, yo, npa, xd, yd, nprm, &npd))) goto out;
        if (npd == 0) break;
        if ((status = hyo_renode_cap(m, xd, yd, npd, xo, yo, capo, &npa))) goto out;
        if (npa == 0) break;
        keep = 1;
      } while (0);
      w->npa[na - 1] = npa;
      if (keep) {
        w->npta = w->npta + npa;
        int iend = ibeg + npa - 1;
        w->i2a[na - 1] = iend;
        for (int i = ibeg; i <= iend - 1; i++) m->nextq[i - 1] = i + 1;
        m->nextq[iend - 1] = ibeg;
      } else {
        w->na = w->na - 1;
      }
      freec[icr] = 0;
    }
  }
out:
  free(kib); free(xcr); free(ycr); free(freec); free(qdx); free(qdy); free(isx); free(isy);
  free(icrtab); free(noctab); free(xd); free(yd);
#undef QA
#undef ICR
#undef NOC
  return status;
}

/* congen.f90:24-128.  qq(ny,nx) is the residual (or the full field when
   there are no contours); it is modified when beta != 0 as in the reference. */
int hyo_recontour(hyo_model *m, double *qq) {
  int nx = m->nx, ny = m->ny, nxu = m->nxu, nyu = m->nyu, mgu = m->mgu;
  long lda = nyu + 1;
  double *qa = (double *)calloc((size_t)lda * nxu, sizeof(double));
  if (!qa) return HYO_ERR_ALLOC;
#define QA(iy, ix) qa[(long)((ix)-1) * lda + (iy)-1]
#define Q(iy, ix) qq[(long)((ix)-1) * ny + (iy)-1]
#define W(w, iyf, ixf) m->w[((ixf)-1) * mgu + (iyf)-1]
  work_store w;
  memset(&w, 0, sizeof(w));
  w.npm = m->npm;
  w.nm = m->nm;
  int status = HYO_OK;
  if (m->nq > 0) {
    con2ugrid(m, qa);
    for (int ix = 1; ix <= nxu; ix++) {
      int ixf = m->ixfw[ix - 1], ix0 = m->ix0w[ix - 1], ix1 = m->ix1w[ix - 1];
      for (int iy = 1; iy <= nyu; iy++) {
        int iyf = m->iyfw[iy - 1], iy0 = m->iy0w[iy - 1], iy1 = m->iy1w[iy - 1];
        QA(iy, ix) = QA(iy, ix) + W(w00, iyf, ixf) * Q(iy0, ix0) + W(w10, iyf, ixf) * Q(iy1, ix0) +
                     W(w01, iyf, ixf) * Q(iy0, ix1) + W(w11, iyf, ixf) * Q(iy1, ix1);
      }
    }
  } else {
    double s = 0.0;
    for (long k = 0; k < (long)nx * ny; k++) s = s + fabs(qq[k]);
    double qql1 = m->garea * s;
    if (qql1 < small) {
      free(qa);
      m->phase = 1;
      return HYO_OK;
    }
    if (m->beffect)
      for (int ix = 1; ix <= nx; ix++)
        for (int iy = 1; iy <= ny; iy++)
          Q(iy, ix) = Q(iy, ix) - m->p.beta * (m->ymin + m->gly * (double)(iy - 1));
    for (int ix = 1; ix <= nxu; ix++) {
      int ixf = m->ixfw[ix - 1], ix0 = m->ix0w[ix - 1], ix1 = m->ix1w[ix - 1];
      for (int iy = 1; iy <= nyu; iy++) {
        int iyf = m->iyfw[iy - 1], iy0 = m->iy0w[iy - 1], iy1 = m->iy1w[iy - 1];
        QA(iy, ix) = W(w00, iyf, ixf) * Q(iy0, ix0) + W(w10, iyf, ixf) * Q(iy1, ix0) +
                     W(w01, iyf, ixf) * Q(iy0, ix1) + W(w11, iyf, ixf) * Q(iy1, ix1);
      }
    }
    if (m->beffect)
      for (int ix = 1; ix <= nxu; ix++)
        for (int iy = 1; iy <= nyu; iy++) QA(iy, ix) = QA(iy, ix) + m->p.beta * m->ygu[iy - 1];
  }
#undef Q
#undef W
#undef QA
  w.xa = (double *)malloc(sizeof(double) * w.npm);
  w.ya = (double *)malloc(sizeof(double) * w.npm);
  w.inda = (int *)malloc(sizeof(int) * w.nm);
  w.npa = (int *)malloc(sizeof(int) * w.nm);
  w.i1a = (int *)malloc(sizeof(int) * w.nm);
  w.i2a = (int *)malloc(sizeof(int) * w.nm);
  status = ugrid2con(m, qa, &w);
  if (status == HYO_OK) {
    memcpy(m->xq, w.xa, sizeof(double) * w.npta);
    memcpy(m->yq, w.ya, sizeof(double) * w.npta);
    for (int j = 0; j < w.na; j++) {
      m->i1q[j] = w.i1a[j];
      m->i2q[j] = w.i2a[j];
      m->npq[j] = w.npa[j];
      m->indq[j] = w.inda[j];
    }
    m->nq = w.na;
    m->nptq = w.npta;
    m->phase = 1;
  } else {
    m->err = status;
  }
  free(w.xa); free(w.ya); free(w.inda); free(w.npa); free(w.i1a); free(w.i2a);
  free(qa);
  return status;
}
