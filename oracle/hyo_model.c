/*
 * hyo_model.c -- oracle: constants (constants.f90), init_contours
 * (contours.f90:48-132), init_spectral / main_invert / gradient / spec1d
 * (spectral.f90:42-158, 304-424) and generic.f90 reductions.
 *
 * TEST INFRASTRUCTURE ONLY (see hyo.h).
 *
 * Pinning: the tables are compared bit-for-bit with the amdflang build of
 * spectral.f90/contours.f90 (oracle/_ref); main_invert and gradient are
 * checked on single Fourier modes against the closed form
 * psi = -q/(k^2+kd^2), u = -psi_y, v = psi_x with the Hou-Li filter factor.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hyo.h"

#define PI_D 3.141592653589793238462643383279502884197169399375105820974944592307816

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* glx**4 as gfortran folds a constant integer power: one correctly rounded
   result.  Computed in binary128 and rounded once. */
static double pow4_folded(double x) {
  __float128 q = (__float128)x;
  q = q * q;
  q = q * q;
  return (double)q;
}

static void init_contours(hyo_model *m) {
  int nx = m->nx, ny = m->ny, mgu = m->mgu;
  /* contours.f90:58-66 */
  for (int ix = 1; ix <= nx - 1; ix++) m->ixp[ix - 1] = ix + 1;
  m->ixp[nx - 1] = 1;
  for (int iy = 1; iy <= ny - 1; iy++) m->iyp[iy - 1] = iy + 1;
  m->iyp[ny - 1] = 1;
  /* contours.f90:69-85 */
  for (int ix = 1; ix <= m->nxf; ix++) m->xgf[ix - 1] = m->xmin + m->glxf * (double)(ix - 1);
  for (int iy = 1; iy <= m->nyf; iy++) m->ygf[iy - 1] = m->ymin + m->glyf * (double)(iy - 1);
  for (int ix = 1; ix <= m->nxu; ix++) m->xgu[ix - 1] = m->xmin + m->glxu * (double)(ix - 1);
  for (int iy = 1; iy <= m->nyu; iy++) m->ygu[iy - 1] = m->ymin + m->glyu * (double)(iy - 1);
  /* contours.f90:88-95 */
  m->ibx[0] = m->nxu - 1;
  for (int ix = 2; ix <= m->nxu + 1; ix++) m->ibx[ix - 1] = ix - 2;
  m->iby[0] = m->nyu - 1;
  for (int iy = 2; iy <= m->nyu + 1; iy++) m->iby[iy - 1] = iy - 2;
  /* contours.f90:100-115; w(iyf,ixf) column-major 16x16 */
  double fac = 1.0 / (double)mgu;
  for (int ixf = 1; ixf <= mgu; ixf++) {
    double pxf = fac * (double)(ixf - 1);
    double pxc = 1.0 - pxf;
    for (int iyf = 1; iyf <= mgu; iyf++) {
      double pyf = fac * (double)(iyf - 1);
      double pyc = 1.0 - pyf;
      int k = (ixf - 1) * mgu + (iyf - 1);
      m->w00[k] = pyc * pxc;
      m->w10[k] = pyf * pxc;
      m->w01[k] = pyc * pxf;
      m->w11[k] = pyf * pxf;
    }
  }
  /* contours.f90:117-129 */
  for (int ix = 1; ix <= m->nxu; ix++) {
    int ixx = (ix - 1) / mgu;
    m->ix0w[ix - 1] = 1 + ixx;
    m->ix1w[ix - 1] = 2 + ixx - nx * (m->ix0w[ix - 1] / nx);
    m->ixfw[ix - 1] = ix - mgu * ixx;
  }
  for (int iy = 1; iy <= m->nyu; iy++) {
    int iyy = (iy - 1) / mgu;
    m->iy0w[iy - 1] = 1 + iyy;
    m->iy1w[iy - 1] = 2 + iyy - ny * (m->iy0w[iy - 1] / ny);
    m->iyfw[iy - 1] = iy - mgu * iyy;
  }
}

static int init_spectral(hyo_model *m) {
  int nx = m->nx, ny = m->ny, nwx = nx / 2, nwy = ny / 2;
  double *rkxf = (double *)xcalloc(nx, sizeof(double));
  double *rkyf = (double *)xcalloc(ny, sizeof(double));
  /* init2dfft, sta2dfft.f90:40-97 */
  if (hyo_initfft(nx, m->xfactors, m->xtrig) || hyo_initfft(ny, m->yfactors, m->ytrig)) {
    free(rkxf);
    free(rkyf);
    return HYO_ERR_FACTOR;
  }
  double sc = PI_D / m->p.ellx;
  for (int k = 1; k <= nx; k++) m->hrkx[k - 1] = sc * (double)k;
  sc = PI_D / m->p.elly;
  for (int k = 1; k <= ny; k++) m->hrky[k - 1] = sc * (double)k;
  /* spectral.f90:57-86 */
  m->rkx[0] = 0.0;
  for (int kx = 1; kx <= nwx - 1; kx++) {
    m->rkx[kx] = m->hrkx[2 * kx - 1];
    m->rkx[nx - kx] = m->hrkx[2 * kx - 1];
  }
  m->rkx[nwx] = m->hrkx[nx - 1];
  double scx = 2.0 * PI_D / m->p.ellx;
  double rkxmax = scx * (double)nwx;
  for (int kx = 0; kx < nx; kx++) {
    rkxf[kx] = m->rkx[kx] * exp(-36.0 * pow(m->rkx[kx] / rkxmax, 36.0));
    m->hrkx[kx] = m->hrkx[kx] * exp(-36.0 * pow(m->hrkx[kx] / rkxmax, 36.0));
  }
  m->rky[0] = 0.0;
  for (int ky = 1; ky <= nwy - 1; ky++) {
    m->rky[ky] = m->hrky[2 * ky - 1];
    m->rky[ny - ky] = m->hrky[2 * ky - 1];
  }
  m->rky[nwy] = m->hrky[ny - 1];
  double scy = 2.0 * PI_D / m->p.elly;
  double rkymax = scy * (double)nwy;
  for (int ky = 0; ky < ny; ky++) {
    rkyf[ky] = m->rky[ky] * exp(-36.0 * pow(m->rky[ky] / rkymax, 36.0));
    m->hrky[ky] = m->hrky[ky] * exp(-36.0 * pow(m->hrky[ky] / rkymax, 36.0));
  }
  /* spectral.f90:95-106 */
  double fac = 36.0 / (double)(nx * ny);
  double a2 = rkxmax * rkxmax, b2 = rkymax * rkymax;
  double rkmsi = 1.0 / (a2 > b2 ? a2 : b2);
  for (int ky = 0; ky < ny; ky++) {
    for (int kx = 0; kx < nx; kx++) {
      long k = (long)ky * nx + kx;
      double rksqu = m->rkx[kx] * m->rkx[kx] + m->rky[ky] * m->rky[ky];
      m->rk[k] = sqrt(rksqu);
      m->qdiss[k] = m->p.cdamp * pow(rkmsi * rksqu, m->p.nnu);
      double t = fac * rksqu;
      m->flo[k] = 1.0 / (1.0 + t * t);
      m->fhi[k] = 1.0 - m->flo[k];
      m->rksq[k] = rkxf[kx] * rkxf[kx] + rkyf[ky] * rkyf[ky];
    }
  }
  /* spectral.f90:110-133 */
  double delk = sqrt(scx * scx + scy * scy);
  double delki = 1.0 / delk;
  m->kmax = (int)lround(sqrt(a2 + b2) * delki);
  m->spmf = (double *)xcalloc((size_t)(nx > ny ? nx : ny) + 1, sizeof(double));
  m->alk = (double *)xcalloc((size_t)(nx > ny ? nx : ny) + 1, sizeof(double));
  m->kmag = (int *)xcalloc((size_t)nx * ny, sizeof(int));
  for (int k = 0; k <= m->kmax; k++) m->spmf[k] = 0.0;
  for (int ky = 0; ky < ny; ky++)
    for (int kx = 0; kx < nx; kx++) {
      int k = (int)lround(sqrt(m->rkx[kx] * m->rkx[kx] + m->rky[ky] * m->rky[ky]) * delki);
      m->kmag[(long)ky * nx + kx] = k;
      m->spmf[k] = m->spmf[k] + 1.0;
    }
  double snorm = 4.0 * PI_D / (double)(nx * ny);
  m->spmf[0] = 0.0;
  for (int k = 1; k <= m->kmax; k++) {
    m->spmf[k] = snorm * (double)k / m->spmf[k];
    m->alk[k] = log10(delk * (double)k);
  }
  m->kmaxred = (int)lround(sqrt((a2 + b2) / 2.0) * delki);
  /* spectral.f90:136-147 */
  m->green[0] = 0.0;
  for (int kx = 2; kx <= nx; kx++) m->green[kx - 1] = -1.0 / (m->rkx[kx - 1] * m->rkx[kx - 1] + m->kdsq);
  for (int ky = 2; ky <= ny; ky++)
    m->green[(long)(ky - 1) * nx] = -1.0 / (m->rky[ky - 1] * m->rky[ky - 1] + m->kdsq);
  for (int ky = 2; ky <= ny; ky++)
    for (int kx = 2; kx <= nx; kx++)
      m->green[(long)(ky - 1) * nx + kx - 1] =
          -1.0 / (m->rkx[kx - 1] * m->rkx[kx - 1] + m->rky[ky - 1] * m->rky[ky - 1] + m->kdsq);
  /* spectral.f90:150-152 */
  for (int iy = 1; iy <= ny; iy++) m->bety[iy - 1] = m->p.beta * (m->gly * (double)(iy - 1) + m->ymin);
  free(rkxf);
  free(rkyf);
  return HYO_OK;
}

hyo_model *hyo_create(const hyo_params *p) {
  if (!p || p->nx < 2 || p->ny < 2 || p->nx != p->ny) return NULL;
  hyo_model *m = (hyo_model *)calloc(1, sizeof(hyo_model));
  m->p = *p;
  int nx = p->nx, ny = p->ny;
  m->nx = nx;
  m->ny = ny;
  m->mgf = 4;
  m->mgu = 16;
  m->nlevm = 2000;
  m->nxf = m->mgf * nx;
  m->nyf = m->mgf * ny;
  m->nxu = m->mgu * nx;
  m->nyu = m->mgu * ny;
  /* constants.f90:44-75 */
  const double small = 1.0e-12;
  m->hlx = 0.5 * p->ellx;
  m->hlxi = 1.0 / m->hlx;
  m->xmin = -m->hlx;
  m->xmax = m->hlx;
  m->hly = 0.5 * p->elly;
  m->hlyi = 1.0 / m->hly;
  m->ymin = -m->hly;
  m->ymax = m->hly;
  m->kdsq = p->kd * p->kd;
  m->srwfm = p->beta * p->kd / (1.0 + m->kdsq);
  m->therm = p->rtherm * m->kdsq;
  m->domarea = p->ellx * p->elly;
  m->aspect = p->ellx / p->elly;
  m->glx = p->ellx / (double)nx;
  m->glxi = (double)nx / p->ellx;
  m->gly = p->elly / (double)ny;
  m->glyi = (double)ny / p->elly;
  m->garea = m->glx * m->gly;
  m->dsumi = 1.0 / (double)(nx * ny);
  m->beffect = p->beta != 0.0;
  m->btropic = p->kd == 0.0;
  m->heating = p->rtherm > 0.0;
  m->friction = p->rekman > 0.0;
  /* contours.f90:19-42 */
  m->glxf = p->ellx / (double)m->nxf;
  m->glxfi = (double)m->nxf / p->ellx;
  m->glyf = p->elly / (double)m->nyf;
  m->glyfi = (double)m->nyf / p->elly;
  m->glxu = p->ellx / (double)m->nxu;
  m->glxui = (double)m->nxu / p->ellx;
  m->glyu = p->elly / (double)m->nyu;
  m->glyui = (double)m->nyu / p->elly;
  m->amu = 0.2;
  m->ell = (p->ell_cells > 0.0 ? p->ell_cells : 6.25) * m->glx;
  m->dm = m->amu * m->amu * m->ell / 4.0;
  m->dm2 = m->dm * m->dm;
  m->d4small = small * pow4_folded(m->glx);
  m->dmsq = 4.0 * m->dm2;
  m->dmi = 2.0 / m->dm;
  m->elf = 1.0 / (m->ell * m->ell);
  m->densf = 1.0 / (m->amu * sqrt(m->ell));

  size_t N = (size_t)nx * ny;
  m->ixp = (int *)xcalloc(nx, sizeof(int));
  m->iyp = (int *)xcalloc(ny, sizeof(int));
  m->xgf = (double *)xcalloc(m->nxf, sizeof(double));
  m->ygf = (double *)xcalloc(m->nyf, sizeof(double));
  m->xgu = (double *)xcalloc(m->nxu, sizeof(double));
  m->ygu = (double *)xcalloc(m->nyu, sizeof(double));
  m->ixfw = (int *)xcalloc(m->nxu, sizeof(int));
  m->ix0w = (int *)xcalloc(m->nxu, sizeof(int));
  m->ix1w = (int *)xcalloc(m->nxu, sizeof(int));
  m->iyfw = (int *)xcalloc(m->nyu, sizeof(int));
  m->iy0w = (int *)xcalloc(m->nyu, sizeof(int));
  m->iy1w = (int *)xcalloc(m->nyu, sizeof(int));
  m->ibx = (int *)xcalloc(m->nxu + 1, sizeof(int));
  m->iby = (int *)xcalloc(m->nyu + 1, sizeof(int));
  double **f[] = {&m->rksq, &m->qdiss, &m->green, &m->flo, &m->fhi, &m->rk,
                  &m->qs,   &m->qd,    &m->qc,    &m->qq,  &m->pp,  &m->uu,
                  &m->vv,   &m->qr,    &m->emq,   &m->epq, &m->qspre};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) *f[i] = (double *)xcalloc(N, sizeof(double));
  m->rkx = (double *)xcalloc(nx, sizeof(double));
  m->hrkx = (double *)xcalloc(nx, sizeof(double));
  m->rky = (double *)xcalloc(ny, sizeof(double));
  m->hrky = (double *)xcalloc(ny, sizeof(double));
  m->bety = (double *)xcalloc(ny, sizeof(double));
  m->xtrig = (double *)xcalloc(2 * (size_t)nx, sizeof(double));
  m->ytrig = (double *)xcalloc(2 * (size_t)ny, sizeof(double));
  /* capacities: constants.f90:34-41 scaled to a run-time default */
  /* constants.f90:35-37 */
  m->npm = p->npm > 0 ? p->npm : (int64_t)500 * nx * ny;
  m->nm = m->npm / 20 + m->npm / 200;
  m->xq = (double *)xcalloc(m->npm, sizeof(double));
  m->yq = (double *)xcalloc(m->npm, sizeof(double));
  m->nextq = (int *)xcalloc(m->npm, sizeof(int));
  m->indq = (int *)xcalloc(m->nm, sizeof(int));
  m->npq = (int *)xcalloc(m->nm, sizeof(int));
  m->i1q = (int *)xcalloc(m->nm, sizeof(int));
  m->i2q = (int *)xcalloc(m->nm, sizeof(int));
  init_contours(m);
  if (init_spectral(m)) {
    hyo_destroy(m);
    return NULL;
  }
  return m;
}

void hyo_destroy(hyo_model *m) {
  if (!m) return;
  void *ptrs[] = {m->ixp,  m->iyp,  m->xgf,  m->ygf,   m->xgu,   m->ygu,  m->ixfw, m->ix0w,
                  m->ix1w, m->iyfw, m->iy0w, m->iy1w,  m->ibx,   m->iby,  m->rksq, m->qdiss,
                  m->green, m->flo, m->fhi,  m->rk,    m->qs,    m->qd,   m->qc,   m->qq,
                  m->pp,   m->uu,   m->vv,   m->qr,    m->emq,   m->epq,  m->qspre, m->rkx,
                  m->hrkx, m->rky,  m->hrky, m->bety,  m->xtrig, m->ytrig, m->spmf, m->alk,
                  m->kmag, m->xq,   m->yq,   m->nextq, m->indq,  m->npq,  m->i1q,  m->i2q};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free(m);
}

int hyo_error(const hyo_model *m) { return m ? m->err : HYO_ERR_ARG; }

/* Sequential sum in storage order, as gfortran (and amdflang -O3) evaluate
   SUM() of a contiguous array (generic.f90:23,40,58; evolution.f90:553,629,640). */
double hyo_sum_seq(const double *a, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s = s + a[i];
  return s;
}

/* spectral.f90:304-342 */
void hyo_main_invert(const hyo_model *m, const double *qq, double *uu, double *vv, double *pp) {
  long N = (long)m->nx * m->ny;
  double *vtmp = (double *)malloc(sizeof(double) * N);
  for (long k = 0; k < N; k++) pp[k] = m->green[k] * qq[k];
  hyo_xderiv(m->nx, m->ny, m->hrkx, pp, vtmp);
  hyo_spctop(m, vtmp, vv);
  hyo_yderiv(m->nx, m->ny, m->hrky, pp, vtmp);
  hyo_spctop(m, vtmp, uu);
  for (long k = 0; k < N; k++) uu[k] = -uu[k];
  free(vtmp);
}

/* spectral.f90:345-366 */
void hyo_gradient(const hyo_model *m, const double *ff, double *ffx, double *ffy) {
  long N = (long)m->nx * m->ny;
  double *vtmp = (double *)malloc(sizeof(double) * N);
  hyo_xderiv(m->nx, m->ny, m->hrkx, ff, vtmp);
  hyo_spctop(m, vtmp, ffx);
  hyo_yderiv(m->nx, m->ny, m->hrky, ff, vtmp);
  hyo_spctop(m, vtmp, ffy);
  free(vtmp);
}

/* spectral.f90:370-424 */
void hyo_spec1d(const hyo_model *m, double *ss, double *spec, int iopt) {
  int nx = m->nx, ny = m->ny;
  if (iopt == 1)
    for (long k = 0; k < (long)nx * ny; k++) ss[k] = ss[k] * m->rksq[k];
  for (int k = 0; k <= m->kmax; k++) spec[k] = 0.0;
  int k = m->kmag[0];
  spec[k] = spec[k] + 0.25 * (ss[0] * ss[0]);
  for (int kx = 2; kx <= nx; kx++) {
    k = m->kmag[kx - 1];
    spec[k] = spec[k] + 0.5 * (ss[kx - 1] * ss[kx - 1]);
  }
  for (int ky = 2; ky <= ny; ky++) {
    k = m->kmag[(long)(ky - 1) * nx];
    spec[k] = spec[k] + 0.5 * (ss[(long)(ky - 1) * nx] * ss[(long)(ky - 1) * nx]);
  }
  for (int ky = 2; ky <= ny; ky++)
    for (int kx = 2; kx <= nx; kx++) {
      long i = (long)(ky - 1) * nx + kx - 1;
      k = m->kmag[i];
      spec[k] = spec[k] + ss[i] * ss[i];
    }
}
