/*
 * hyo_contours.c -- oracle restatement of contours.f90: velint (136-169),
 * con2grid (173-346), surgery (350-683) and renode (687-933).
 *
 * TEST INFRASTRUCTURE ONLY (see hyo.h).
 *
 * Pinning:
 *   velint    -- exact for fields bilinear in a cell (closed form), periodic
 *                wrap at the last cell; bit-for-bit vs oracle/_ref.
 *   con2grid  -- integer level counts vs a brute-force point-in-polygon
 *                count on the 4ng grid; sum(qc) == 0; bit-for-bit vs
 *                oracle/_ref on random closed contours.
 *   surgery   -- special cases (two parallel segments closer than dm
 *                reconnect, farther do not; a pinched loop splits);
 *                bit-for-bit (topology and positions) vs oracle/_ref.
 *   renode    -- circle of radius R gives nint(sum e)+1 nodes from the
 *                closed-form curvature; corners preserved; bit-for-bit vs
 *                oracle/_ref.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hyo.h"

static const double small = 1.0e-12;
static const double small3 = 1.0e-12 * 1.0e-12 * 1.0e-12;
static const double oms = 1.0 - 1.0e-12;

/* Fortran int() of a double: truncation toward zero. */
#define FINT(x) ((int)(x))

/* contours.f90:136-169 */
void hyo_velint(const hyo_model *m, const double *uu, const double *vv, double *uq, double *vq) {
  int ny = m->ny;
  for (int i = 0; i < m->nptq; i++) {
    double xx = m->glxi * (m->xq[i] - m->xmin);
    int ix0 = 1 + FINT(xx);
    double pxc = (double)ix0 - xx;
    double px = 1.0 - pxc;
    int ix1 = m->ixp[ix0 - 1];
    double yy = m->glyi * (m->yq[i] - m->ymin);
    int iy0 = 1 + FINT(yy);
    double pyc = (double)iy0 - yy;
    double py = 1.0 - pyc;
    int iy1 = m->iyp[iy0 - 1];
#define F(a, iy, ix) a[(long)((ix)-1) * ny + (iy)-1]
    uq[i] = pyc * (pxc * F(uu, iy0, ix0) + px * F(uu, iy0, ix1)) +
            py * (pxc * F(uu, iy1, ix0) + px * F(uu, iy1, ix1));
    vq[i] = pyc * (pxc * F(vv, iy0, ix0) + px * F(vv, iy0, ix1)) +
            py * (pxc * F(vv, iy1, ix0) + px * F(vv, iy1, ix1));
#undef F
  }
}

/* contours.f90:173-346.  If counts != NULL the integer level counts
   nint(qa/qjump) of the fine grid after the y sweep are stored there,
   (nyf,nxf) column-major, before any averaging. */
void hyo_con2grid_counts(const hyo_model *m, double *qc, int32_t *counts) {
  int nx = m->nx, ny = m->ny, nxf = m->nxf, nyf = m->nyf, nptq = m->nptq;
  double ellx = m->p.ellx, elly = m->p.elly, hlx = m->hlx, hly = m->hly;
  double hlxi = m->hlxi, hlyi = m->hlyi, xmin = m->xmin, ymin = m->ymin;
  double qjump = m->qjump;
  const double *xq = m->xq, *yq = m->yq;
  if (nptq == 0) {
    for (long k = 0; k < (long)nx * ny; k++) qc[k] = 0.0;
    if (counts) memset(counts, 0, sizeof(int32_t) * (size_t)nxf * nyf);
    return;
  }
  long lda = nyf + 1;
  double *qa = (double *)calloc((size_t)lda * nxf, sizeof(double));
#define QA(iy, ix) qa[(long)((ix)-1) * lda + (iy)-1]
  double *qjx = (double *)calloc(nxf, sizeof(double));
  double *qbot = (double *)calloc(nxf, sizeof(double));
  double *dx = (double *)malloc(sizeof(double) * nptq);
  double *dy = (double *)malloc(sizeof(double) * nptq);
  int *ixc = (int *)malloc(sizeof(int) * nptq);
  int *nxc = (int *)malloc(sizeof(int) * nptq);
  for (int i = 0; i < nptq; i++) ixc[i] = 1 + FINT(m->glxfi * (xq[i] - xmin));
  for (int i = 0; i < nptq; i++) {
    int ia = m->nextq[i] - 1;
    double xx = xq[ia] - xq[i];
    dx[i] = xx - ellx * (double)FINT(xx * hlxi);
    double yy = yq[ia] - yq[i];
    dy[i] = yy - elly * (double)FINT(yy * hlyi);
    int ixdif = ixc[ia] - ixc[i];
    nxc[i] = ixdif - nxf * ((2 * ixdif) / nxf);
    if (fabs(yy) > hly) {
      double cc = copysign(1.0, dy[i]);
      double p = -(yq[i] - hly * cc) / (dy[i] + small);
      int ix = 1 + FINT(m->glxfi * (fmod(ellx + hlx + xq[i] + p * dx[i], ellx)));
      qjx[ix - 1] = qjx[ix - 1] - qjump * cc;
    }
  }
  qbot[0] = 0.0;
  for (int ix = 1; ix <= nxf - 1; ix++) qbot[ix] = qbot[ix - 1] + qjx[ix - 1];
  for (int i = 0; i < nptq; i++) {
    if (nxc[i] != 0) {
      int jump = nxc[i] >= 0 ? 1 : -1;
      int ixbeg = ixc[i] + (jump - 1) / 2 + nxf;
      double sqjump = qjump * copysign(1.0, dx[i]);
      int ncr = 0;
      while (ncr != nxc[i]) {
        int ix = 1 + (ixbeg + ncr) % nxf;
        double xx = m->xgf[ix - 1] - xq[i];
        double px0 = (xx - ellx * (double)FINT(xx * hlxi)) / dx[i];
        double yy = yq[i] + px0 * dy[i];
        int iy = 2 + FINT(m->glyfi * (yy - elly * (double)FINT(yy * hlyi) - ymin));
        QA(iy, ix) = QA(iy, ix) + sqjump;
        ncr = ncr + jump;
      }
    }
  }
  for (int ix = 1; ix <= nxf; ix++) {
    QA(1, ix) = qbot[ix - 1];
    for (int iy = 2; iy <= nyf; iy++) QA(iy, ix) = QA(iy, ix) + QA(iy - 1, ix);
  }
  if (m->beffect)
    for (int ix = 1; ix <= nxf; ix++)
      for (int iy = 1; iy <= nyf; iy++) QA(iy, ix) = QA(iy, ix) - m->p.beta * m->ygf[iy - 1];
  if (counts)
    for (int ix = 1; ix <= nxf; ix++)
      for (int iy = 1; iy <= nyf; iy++)
        counts[(long)(ix - 1) * nyf + iy - 1] = (int32_t)lround(QA(iy, ix) / qjump);
  /* nine-point averaging 4ng -> 2ng -> ng, contours.f90:290-327 */
  int nh = nyf / 2 + 1;
  double *qod0 = (double *)malloc(sizeof(double) * nh), *qod1 = (double *)malloc(sizeof(double) * nh);
  double *qod2 = (double *)malloc(sizeof(double) * nh), *qev0 = (double *)malloc(sizeof(double) * nh);
  double *qev1 = (double *)malloc(sizeof(double) * nh), *qev2 = (double *)malloc(sizeof(double) * nh);
  int nxh = nxf, nyh = nyf;
  while (nxh > nx) {
    int nxff = nxh;
    (void)nxff;
    nxh = nxh / 2;
    int nyff = nyh;
    nyh = nyh / 2;
    /* qod*(iy) stored at [iy-1] for iy=1..nyh; qev*(iy) at [iy] for iy=0..nyh */
    for (int iy = 1; iy <= nyh; iy++) {
      int miy = 2 * iy;
      qod2[iy - 1] = QA(miy - 1, nyff);
      qev2[iy] = QA(miy, nyff);
    }
    qev2[0] = QA(nyff, nyff);
    for (int ix = 1; ix <= nxh; ix++) {
      int mix = 2 * ix, mixm1 = mix - 1;
      for (int iy = 1; iy <= nyh; iy++) {
        int miy = 2 * iy;
        qod1[iy - 1] = QA(miy - 1, mixm1);
        qod0[iy - 1] = QA(miy - 1, mix);
        qev1[iy] = QA(miy, mixm1);
        qev0[iy] = QA(miy, mix);
      }
      qev1[0] = qev1[nyh];
      qev0[0] = qev0[nyh];
      for (int iy = 1; iy <= nyh; iy++)
        QA(iy, ix) = 0.0625 * (qev0[iy] + qev0[iy - 1] + qev2[iy] + qev2[iy - 1]) +
                     0.125 * (qev1[iy] + qev1[iy - 1] + qod0[iy - 1] + qod2[iy - 1]) +
                     0.25 * qod1[iy - 1];
      for (int iy = 1; iy <= nyh; iy++) {
        qod2[iy - 1] = qod0[iy - 1];
        qev2[iy] = qev0[iy];
      }
      qev2[0] = qev0[0];
    }
  }
  /* contours.f90:330-342 */
  double qavg0 = 0.0;
  for (int ix = 1; ix <= nx; ix++)
    for (int iy = 1; iy <= ny; iy++) qavg0 = qavg0 + QA(iy, ix);
  qavg0 = qavg0 / (double)(nx * ny);
  for (int ix = 1; ix <= nx; ix++)
    for (int iy = 1; iy <= ny; iy++) qc[(long)(ix - 1) * ny + iy - 1] = QA(iy, ix) - qavg0;
#undef QA
  free(qod0); free(qod1); free(qod2); free(qev0); free(qev1); free(qev2);
  free(qa); free(qjx); free(qbot); free(dx); free(dy); free(ixc); free(nxc);
}

void hyo_con2grid(const hyo_model *m, double *qc) { hyo_con2grid_counts(m, qc, NULL); }

/* contours.f90:687-933.  xd,yd have npd entries; xr,yr hold at most cap
   nodes (HYO_ERR_CAPACITY if the result would not fit; the reference writes
   into fixed arrays without a check).  *npr = 0 means "delete". */
int hyo_renode_cap(const hyo_model *m, const double *xd0, const double *yd0, int npd,
                   double *xr0, double *yr0, int64_t cap, int *npr_out) {
  double ellx = m->p.ellx, elly = m->p.elly, hlxi = m->hlxi, hlyi = m->hlyi;
  double dmsq = m->dmsq, elf = m->elf, dmi = m->dmi, densf = m->densf;
  const double f12 = 0.5, f13 = 1.0 / 3.0, f16 = 1.0 / 6.0;
  size_t n1 = (size_t)npd + 2;
  double *buf = (double *)malloc(sizeof(double) * n1 * 9);
  double *dx = buf, *dy = buf + n1, *a = buf + 2 * n1, *b = buf + 3 * n1, *c = buf + 4 * n1;
  double *d = buf + 5 * n1, *e = buf + 6 * n1, *u = buf + 7 * n1, *v = buf + 8 * n1;
  int *node = (int *)malloc(sizeof(int) * n1);
  int *next = (int *)malloc(sizeof(int) * n1);
  char *corner = (char *)malloc(n1);
  const double *xd = xd0 - 1, *yd = yd0 - 1; /* 1-based views */
  double *xr = xr0 - 1, *yr = yr0 - 1;
  int i, ia, ib;
  for (i = 0; i <= npd - 1; i++) next[i] = i + 1;
  next[npd] = 1;
  for (i = 1; i <= npd; i++) {
    ia = next[i];
    double xx = xd[ia] - xd[i];
    dx[i] = xx - ellx * (double)FINT(xx * hlxi);
    double yy = yd[ia] - yd[i];
    dy[i] = yy - elly * (double)FINT(yy * hlyi);
  }
  for (i = 1; i <= npd; i++) {
    v[i] = dx[i] * dx[i] + dy[i] * dy[i];
    e[i] = sqrt(v[i]);
  }
  for (ib = 1; ib <= npd; ib++) {
    i = next[ib];
    u[i] = v[ib];
    a[i] = -dx[ib];
    c[i] = -dy[ib];
  }
  int ncorn = 0;
  for (i = 1; i <= npd; i++) {
    corner[i] = dx[i] * a[i] + dy[i] * c[i] > 0.0;
    if (corner[i]) {
      ncorn = ncorn + 1;
      node[ncorn] = i;
      b[i] = 0.0;
    } else {
      double t1 = a[i] * v[i] - dx[i] * u[i];
      double t2 = c[i] * v[i] - dy[i] * u[i];
      b[i] = (dx[i] * c[i] - a[i] * dy[i]) / sqrt(t1 * t1 + t2 * t2 + small3);
    }
  }
  for (i = 1; i <= npd; i++) {
    ia = next[i];
    u[i] = e[i] * (b[ia] + b[i]);
    c[i] = e[i] * (b[ia] - b[i]);
  }
  for (i = 1; i <= npd; i++) {
    a[i] = f16 * c[i] - f12 * u[i];
    b[i] = f12 * (u[i] - c[i]);
    c[i] = f13 * c[i];
  }
  for (i = 1; i <= npd; i++) {
    double ww = 1.0 / (v[i] + dmsq);
    u[i] = ww * sqrt(elf * v[i] + u[i] * u[i]);
    v[i] = ww * e[i];
  }
  for (ib = 1; ib <= npd; ib++) {
    i = next[ib];
    d[i] = (u[ib] + u[i]) / (v[ib] + v[i] + small);
  }
  for (i = 1; i <= npd; i++) {
    ia = next[i];
    u[i] = f12 * (d[i] + d[ia]);
  }
  for (i = 1; i <= npd; i++) {
    double t = densf * sqrt(u[i]) + u[i];
    e[i] = e[i] * (dmi < t ? dmi : t);
  }
  double sum = 0.0;
  for (i = 1; i <= npd; i++) sum = sum + e[i];
  int npr = (int)lround(sum) + 1;
  int status = HYO_OK;
  if (npr < 3) {
    npr = 0;
    goto done;
  }
  if (ncorn <= 1 && npr > cap) {
    npr = 0;
    status = HYO_ERR_CAPACITY;
    goto done;
  }
#define PLACE(im)                                                        \
  do {                                                                   \
    double p = 1.0 - acc / e[i];                                         \
    double eta = p * (a[i] + p * (b[i] + p * c[i]));                     \
    double xx = xd[i] + p * dx[i] - eta * dy[i];                         \
    xr[im] = oms * (xx - ellx * (double)FINT(xx * hlxi));                \
    double yy = yd[i] + p * dy[i] + eta * dx[i];                         \
    yr[im] = oms * (yy - elly * (double)FINT(yy * hlyi));                \
  } while (0)
  if (ncorn == 0) {
    double fac = (double)npr / sum;
    for (i = 1; i <= npd; i++) e[i] = fac * e[i];
    xr[1] = xd[1];
    yr[1] = yd[1];
    double acc = 0.0;
    i = 0;
    for (int im = 2; im <= npr; im++) {
      while (acc < 1.0) {
        i = i + 1;
        if (i > npd) { status = HYO_ERR_ARG; goto done; }
        acc = acc + e[i];
      }
      acc = acc - 1.0;
      PLACE(im);
    }
  } else if (ncorn == 1) {
    double fac = (double)npr / sum;
    for (i = 1; i <= npd; i++) e[i] = fac * e[i];
    i = node[1];
    xr[1] = xd[i];
    yr[1] = yd[i];
    double acc = 0.0;
    i = i - 1;
    for (int im = 2; im <= npr; im++) {
      while (acc < 1.0) {
        i = next[i];
        acc = acc + e[i];
      }
      acc = acc - 1.0;
      PLACE(im);
    }
  } else {
    npr = 0;
    node[ncorn + 1] = node[1];
    int ibeg = node[1];
    for (int k = 1; k <= ncorn; k++) {
      int iend = node[k + 1];
      i = ibeg;
      double sm = 0.0;
      while (i != iend) {
        sm = sm + e[i];
        i = next[i];
      }
      if (sm > small) {
        int npseg = (int)lround(sm) + 1;
        if ((int64_t)npr + npseg > cap) {
          npr = 0;
          status = HYO_ERR_CAPACITY;
          goto done;
        }
        double fac = (double)npseg / sm;
        i = ibeg;
        while (i != iend) {
          e[i] = fac * e[i];
          i = next[i];
        }
        xr[npr + 1] = xd[ibeg];
        yr[npr + 1] = yd[ibeg];
        if (npseg > 1) {
          double acc = 0.0;
          i = ibeg - 1;
          for (int im = npr + 2; im <= npr + npseg; im++) {
            while (acc < 1.0) {
              i = next[i];
              acc = acc + e[i];
            }
            acc = acc - 1.0;
            PLACE(im);
          }
        }
        npr = npr + npseg;
      }
      ibeg = iend;
    }
  }
#undef PLACE
done:
  *npr_out = npr;
  free(buf);
  free(node);
  free(next);
  free(corner);
  return status;
}

int hyo_renode(const hyo_model *m, const double *xd, const double *yd, int npd, double *xr,
               double *yr, int *npr) {
  return hyo_renode_cap(m, xd, yd, npd, xr, yr, INT64_MAX, npr);
}

/* contours.f90:350-683 */
int hyo_surgery(hyo_model *m) {
  if (m->nq == 0) return HYO_OK;
  int nx = m->nx, ny = m->ny, nq0 = m->nq, nptq0 = m->nptq;
  double ellx = m->p.ellx, elly = m->p.elly, hlxi = m->hlxi, hlyi = m->hlyi;
  double xmin = m->xmin, ymin = m->ymin, aspect = m->aspect;
  double d4small = m->d4small, dm2 = m->dm2;
  /* 1-based views of the contour store */
  double *xq = m->xq - 1, *yq = m->yq - 1;
  int *nextq = m->nextq - 1, *indq = m->indq - 1, *npq = m->npq - 1, *i1q = m->i1q - 1,
      *i2q = m->i2q - 1;
  int64_t npm = m->npm;
  double *xa = (double *)malloc(sizeof(double) * npm) - 1;
  double *ya = (double *)malloc(sizeof(double) * npm) - 1;
  int *nexta = (int *)malloc(sizeof(int) * npm) - 1;
  double *xd = (double *)malloc(sizeof(double) * (nptq0 + 1));
  double *yd = (double *)malloc(sizeof(double) * (nptq0 + 1));
  double *dx = (double *)malloc(sizeof(double) * (nptq0 + 1));
  double *dy = (double *)malloc(sizeof(double) * (nptq0 + 1));
  double *dsq = (double *)malloc(sizeof(double) * (nptq0 + 1));
  char *avail = (char *)malloc(nptq0 + 1);
  /* surgery may split contours: the new count is bounded by nm, not nq0 */
  int *i1a = (int *)malloc(sizeof(int) * (m->nm + 1));
  int *i2a = (int *)malloc(sizeof(int) * (m->nm + 1));
  int *jq1 = (int *)malloc(sizeof(int) * (nq0 + 1)), *jq2 = (int *)malloc(sizeof(int) * (nq0 + 1));
  int *iq1 = (int *)malloc(sizeof(int) * (nq0 + 1)), *iq2 = (int *)malloc(sizeof(int) * (nq0 + 1));
  int *levq = (int *)malloc(sizeof(int) * (nq0 + 1));
  int ngbs = nx * ny;
  int *nspb = (int *)malloc(sizeof(int) * (ngbs + 1)), *kb1 = (int *)malloc(sizeof(int) * (ngbs + 1));
  int *kb2 = (int *)malloc(sizeof(int) * (ngbs + 1));
  /* segment-box incidences: at most (range of boxes per segment) per node */
  int64_t cap = (int64_t)4 * nptq0 + 1024;
  int *loc = (int *)malloc(sizeof(int) * (cap + 1)), *list = (int *)malloc(sizeof(int) * (cap + 1));
  int *node = (int *)malloc(sizeof(int) * (cap + 1));
  int status = HYO_OK;
  int i, j, is, ie, ib, isb;

  int levp = indq[1], nlev = 1;
  jq1[1] = 1;
  levq[1] = levp;
  for (j = 2; j <= nq0; j++) {
    int lev = indq[j];
    if (lev > levp) {
      jq2[nlev] = j - 1;
      nlev = nlev + 1;
      jq1[nlev] = j;
      levq[nlev] = lev;
      levp = lev;
    }
  }
  jq2[nlev] = nq0;
  for (int lev = 1; lev <= nlev; lev++) {
    iq1[lev] = i1q[jq1[lev]];
    iq2[lev] = i2q[jq2[lev]];
  }
  for (i = 1; i <= nptq0 - 1; i++) {
    double xx = xq[i + 1] - xq[i];
    dx[i] = xx - ellx * (double)FINT(xx * hlxi);
    double yy = yq[i + 1] - yq[i];
    dy[i] = yy - elly * (double)FINT(yy * hlyi);
    dsq[i] = dx[i] * dx[i] + dy[i] * dy[i];
    avail[i] = 1;
  }
  avail[nptq0] = 1;
  for (j = 1; j <= nq0; j++) {
    ie = i2q[j];
    is = i1q[j];
    double xx = xq[is] - xq[ie];
    dx[ie] = xx - ellx * (double)FINT(xx * hlxi);
    double yy = yq[is] - yq[ie];
    dy[ie] = yy - elly * (double)FINT(yy * hlyi);
    dsq[ie] = dx[ie] * dx[ie] + dy[ie] * dy[ie];
  }

  int nq = 0, nptq = 0;
  for (int lev = 1; lev <= nlev; lev++) {
    int ibeg = iq1[lev], iend = iq2[lev];
    int nptlev = iend - ibeg + 1;
    double fnq = (double)nptlev;
    double fnbx = 1.2 * sqrt(fnq * aspect);
    double fnby = fnbx / aspect;
    int nbx = (int)lround(fnbx);
    if (nbx > nx) nbx = nx;
    if (nbx < 1) nbx = 1;
    int nby = (int)lround(fnby);
    if (nby > ny) nby = ny;
    if (nby < 1) nby = 1;
    int nbox = nbx * nby;
    double bwxi = oms * (double)nbx / ellx;
    double bwyi = oms * (double)nby / elly;
    for (int mb = 1; mb <= nbox; mb++) nspb[mb] = 0;
    int64_t nseg = 0;
    double dnbx = (double)nbx, dnby = (double)nby;
    for (ib = ibeg; ib <= iend; ib++) {
      i = nextq[ib];
      int mbx1, mbx2, mby1, mby2;
      if (dx[i] > 0.0) {
        mbx1 = FINT(dnbx + bwxi * (xq[i] - xmin));
        mbx2 = FINT(dnbx + bwxi * (xq[i] + dx[i] - xmin));
      } else {
        mbx1 = FINT(dnbx + bwxi * (xq[i] + dx[i] - xmin));
        mbx2 = FINT(dnbx + bwxi * (xq[i] - xmin));
      }
      if (dy[i] > 0.0) {
        mby1 = FINT(dnby + bwyi * (yq[i] - ymin));
        mby2 = FINT(dnby + bwyi * (yq[i] + dy[i] - ymin));
      } else {
        mby1 = FINT(dnby + bwyi * (yq[i] + dy[i] - ymin));
        mby2 = FINT(dnby + bwyi * (yq[i] - ymin));
      }
      for (int mbx = mbx1; mbx <= mbx2; mbx++) {
        int mmbx = mbx % nbx;
        for (int mby = mby1; mby <= mby2; mby++) {
          int mb = nby * mmbx + mby % nby + 1;
          nspb[mb] = nspb[mb] + 1;
          nseg = nseg + 1;
          if (nseg > cap) {
            cap = cap * 2;
            loc = (int *)realloc(loc, sizeof(int) * (cap + 1));
            list = (int *)realloc(list, sizeof(int) * (cap + 1));
            node = (int *)realloc(node, sizeof(int) * (cap + 1));
          }
          loc[nseg] = mb;
          list[nseg] = ib;
        }
      }
    }
    kb1[1] = 1;
    for (int mb = 1; mb <= nbox - 1; mb++) kb1[mb + 1] = kb1[mb] + nspb[mb];
    for (int mb = 1; mb <= nbox; mb++) kb2[mb] = kb1[mb] - 1;
    for (int64_t k = 1; k <= nseg; k++) {
      int mb = loc[k];
      int ks = kb2[mb] + 1;
      node[ks] = list[k];
      kb2[mb] = ks;
    }
    for (ib = ibeg; ib <= iend; ib++) {
      i = nextq[ib];
      int mb = nby * FINT(bwxi * (xq[i] - xmin)) + FINT(bwyi * (yq[i] - ymin)) + 1;
      for (int k = kb1[mb]; k <= kb2[mb]; k++) {
        isb = node[k];
        is = nextq[isb];
        /* (is-ib)*(is-i) == 0 in the reference, i.e. is is ib or i */
        if (is == ib || is == i) continue;
        double xx = xq[is] - xq[i];
        double delx = xx - ellx * (double)FINT(xx * hlxi);
        double yy = yq[is] - yq[i];
        double dely = yy - elly * (double)FINT(yy * hlyi);
        double aa = delx * dx[is] + dely * dy[is];
        if (aa * (aa + dsq[is]) + d4small < 0.0) {
          double t = delx * dy[is] - dely * dx[is];
          double cc = t * t;
          if (cc * (cc - dm2 * dsq[is]) < 0.0) {
            double dxa, dya;
            if (dsq[is] + 2.0 * aa < 0.0) {
              dxa = 0.5 * (delx + dx[is]);
              dya = 0.5 * (dely + dy[is]);
              isb = is;
              is = nextq[is];
            } else {
              dxa = 0.5 * delx;
              dya = 0.5 * dely;
            }
            xx = xq[i] + dxa;
            xq[i] = oms * (xx - ellx * (double)FINT(xx * hlxi));
            yy = yq[i] + dya;
            yq[i] = oms * (yy - elly * (double)FINT(yy * hlyi));
            dx[i] = dx[i] - dxa;
            dy[i] = dy[i] - dya;
            dsq[i] = dx[i] * dx[i] + dy[i] * dy[i];
            nextq[isb] = i;
            dx[isb] = dx[isb] - dxa;
            dy[isb] = dy[isb] - dya;
            dsq[isb] = dx[isb] * dx[isb] + dy[isb] * dy[isb];
            xq[is] = xq[i];
            yq[is] = yq[i];
            dx[is] = dx[is] + dxa;
            dy[is] = dy[is] + dya;
            dsq[is] = dx[is] * dx[is] + dy[is] * dy[is];
            nextq[ib] = is;
            dx[ib] = dx[ib] + dxa;
            dy[ib] = dy[ib] + dya;
            dsq[ib] = dx[ib] * dx[ib] + dy[ib] * dy[ib];
            i = is;
            if (i == ib) break;
          }
        }
      }
    }
    /* rebuild the contours of this level, contours.f90:619-663 */
    int levt = levq[lev];
    for (i = ibeg; i <= iend; i++) {
      if (!avail[i]) continue;
      nq = nq + 1;
      if (nq > m->nm) { status = HYO_ERR_CAPACITY; goto out; }
      int npd = 1;
      xd[0] = xq[i];
      yd[0] = yq[i];
      avail[i] = 0;
      is = nextq[i];
      while (is != i) {
        npd = npd + 1;
        xd[npd - 1] = xq[is];
        yd[npd - 1] = yq[is];
        avail[is] = 0;
        is = nextq[is];
      }
      int npnew = npd;
      if (npd > 3) {
        int rs = hyo_renode_cap(m, xd, yd, npd, &xa[nptq + 1], &ya[nptq + 1], npm - nptq, &npnew);
        if (rs) { status = rs; goto out; }
      }
      npq[nq] = npnew;
      if (npq[nq] > 3) {
        i1a[nq] = nptq + 1;
        nptq = nptq + npq[nq];
        if (nptq > npm) { status = HYO_ERR_CAPACITY; goto out; }
        i2a[nq] = nptq;
        indq[nq] = levt;
        for (is = i1a[nq]; is <= nptq - 1; is++) nexta[is] = is + 1;
        nexta[nptq] = i1a[nq];
      } else {
        nq = nq - 1;
      }
    }
  }
  for (i = 1; i <= nptq; i++) {
    xq[i] = xa[i];
    yq[i] = ya[i];
    nextq[i] = nexta[i];
  }
  for (j = 1; j <= nq; j++) {
    i1q[j] = i1a[j];
    i2q[j] = i2a[j];
  }
  m->nq = nq;
  m->nptq = nptq;
out:
  free(xa + 1); free(ya + 1); free(nexta + 1); free(xd); free(yd); free(dx); free(dy); free(dsq);
  free(avail); free(i1a); free(i2a); free(jq1); free(jq2); free(iq1); free(iq2); free(levq);
  free(nspb); free(kb1); free(kb2); free(loc); free(list); free(node);
  if (status) m->err = status;
  return status;
}
