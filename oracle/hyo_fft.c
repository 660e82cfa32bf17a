/*
 * hyo_fft.c -- oracle restatement of lib/stafft/stafft.f90 (radix 2 and 4,
 * the only radices reached for power-of-two lengths) and of the doubly
 * periodic part of lib/stafft/sta2dfft.f90.
 *
 * TEST INFRASTRUCTURE ONLY (see hyo.h).
 *
 * Pinning: forfft/revfft/ptospc/spctop are checked against numpy.fft.rfft
 * re-packed into half-complex order and scaled by 1/sqrt(n) (closed form),
 * against revfft(forfft(x)) == x, and bit-for-bit against the amdflang build
 * of stafft.f90 itself (oracle/_ref).  xderiv/yderiv are checked on single
 * Fourier modes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hyo.h"

/* stafft.f90:128-184 */
int hyo_factorisen(int n, int factors[5]) {
  int rem = n;
  for (int i = 0; i < 5; i++) factors[i] = 0;
  if (n < 2) return 1;
  while (rem % 6 == 0) { factors[0]++; rem /= 6; if (rem == 1) return 0; }
  while (rem % 4 == 0) { factors[1]++; rem /= 4; if (rem == 1) return 0; }
  while (rem % 2 == 0) { factors[2]++; rem /= 2; if (rem == 1) return 0; }
  while (rem % 3 == 0) { factors[3]++; rem /= 3; if (rem == 1) return 0; }
  while (rem % 5 == 0) { factors[4]++; rem /= 5; if (rem == 1) return 0; }
  return 1;
}

/* stafft.f90:63-125.  Only lengths 2^k are accepted (radices 2 and 4). */
int hyo_initfft(int n, int factors[5], double *trig) {
  static const int fac[5] = {6, 4, 2, 3, 5};
  const double twopi = 6.283185307179586476925286766559005768394338798750211641949889184615633;
  if (hyo_factorisen(n, factors)) return HYO_ERR_FACTOR;
  if (factors[0] || factors[3] || factors[4]) return HYO_ERR_FACTOR;
  double ftwopin = twopi / (double)n;
  int rem = n, m = 0;
  for (int i = 0; i < 5; i++) {
    for (int j = 0; j < factors[i]; j++) {
      rem = rem / fac[i];
      for (int k = 1; k <= fac[i] - 1; k++)
        for (int l = 0; l <= rem - 1; l++) trig[m++] = ftwopin * (double)(k * l);
      ftwopin = ftwopin * fac[i];
    }
  }
  for (int i = 0; i < n - 1; i++) {
    trig[n + i] = -sin(trig[i]);
    trig[i] = cos(trig[i]);
  }
  return HYO_OK;
}

/* stafft.f90:890-997, a(0:nv-1,0:3,0:lv-1) -> b(0:nv-1,0:lv-1,0:3) */
static void forrdx4(const double *a, double *b, long nv, long lv, const double *cosine,
                    const double *sine) {
#define A(i, j, k) a[(i) + nv * ((j) + 4 * (k))]
#define B(i, k, j) b[(i) + nv * ((k) + lv * (j))]
#define CS(k, j) cosine[(k) + lv * ((j)-1)]
#define SN(k, j) sine[(k) + lv * ((j)-1)]
  const double rtf12 = 0.7071067811865475244008443621048490392848359376884740365883398689953662;
  for (long i = 0; i < nv; i++) {
    double t1r = A(i, 0, 0) + A(i, 2, 0);
    double t2r = A(i, 1, 0) + A(i, 3, 0);
    B(i, 0, 0) = t1r + t2r;
    B(i, 0, 1) = A(i, 0, 0) - A(i, 2, 0);
    B(i, 0, 2) = t1r - t2r;
    B(i, 0, 3) = A(i, 3, 0) - A(i, 1, 0);
  }
  for (long k = 1; k <= (lv - 1) / 2; k++) {
    long kc = lv - k;
    double c1k = CS(k, 1), s1k = SN(k, 1), c2k = CS(k, 2), s2k = SN(k, 2);
    double c3k = CS(k, 3), s3k = SN(k, 3);
    for (long i = 0; i < nv; i++) {
      double x1p = c1k * A(i, 1, k) - s1k * A(i, 1, kc);
      double y1p = c1k * A(i, 1, kc) + s1k * A(i, 1, k);
      double x2p = c2k * A(i, 2, k) - s2k * A(i, 2, kc);
      double y2p = c2k * A(i, 2, kc) + s2k * A(i, 2, k);
      double x3p = c3k * A(i, 3, k) - s3k * A(i, 3, kc);
      double y3p = c3k * A(i, 3, kc) + s3k * A(i, 3, k);
      double t1r = A(i, 0, k) + x2p;
      double t1i = A(i, 0, kc) + y2p;
      double t2r = x1p + x3p;
      double t2i = y1p + y3p;
      double t3r = A(i, 0, k) - x2p;
      double t3i = A(i, 0, kc) - y2p;
      double t4r = x3p - x1p;
      double t4i = y1p - y3p;
      B(i, k, 0) = t1r + t2r;
      B(i, kc, 0) = t3r - t4i;
      B(i, k, 1) = t3r + t4i;
      B(i, kc, 1) = t1r - t2r;
      B(i, k, 2) = t2i - t1i;
      B(i, kc, 2) = t3i + t4r;
      B(i, k, 3) = t4r - t3i;
      B(i, kc, 3) = t1i + t2i;
    }
  }
  if (lv % 2 == 0) {
    long lvd2 = lv / 2;
    for (long i = 0; i < nv; i++) {
      double q1 = rtf12 * (A(i, 1, lvd2) - A(i, 3, lvd2));
      double q2 = rtf12 * (A(i, 1, lvd2) + A(i, 3, lvd2));
      B(i, lvd2, 0) = A(i, 0, lvd2) + q1;
      B(i, lvd2, 1) = A(i, 0, lvd2) - q1;
      B(i, lvd2, 2) = A(i, 2, lvd2) - q2;
      B(i, lvd2, 3) = -A(i, 2, lvd2) - q2;
    }
  }
#undef A
#undef B
#undef CS
#undef SN
}

/* stafft.f90:1080-1130, a(0:nv-1,0:1,0:lv-1) -> b(0:nv-1,0:lv-1,0:1) */
static void forrdx2(const double *a, double *b, long nv, long lv, const double *cosine,
                    const double *sine) {
#define A(i, j, k) a[(i) + nv * ((j) + 2 * (k))]
#define B(i, k, j) b[(i) + nv * ((k) + lv * (j))]
  for (long i = 0; i < nv; i++) {
    B(i, 0, 0) = A(i, 0, 0) + A(i, 1, 0);
    B(i, 0, 1) = A(i, 0, 0) - A(i, 1, 0);
  }
  for (long k = 1; k <= (lv - 1) / 2; k++) {
    long kc = lv - k;
    double c1k = cosine[k], s1k = sine[k];
    for (long i = 0; i < nv; i++) {
      double x1 = c1k * A(i, 1, k) - s1k * A(i, 1, kc);
      double y1 = c1k * A(i, 1, kc) + s1k * A(i, 1, k);
      B(i, k, 0) = A(i, 0, k) + x1;
      B(i, kc, 0) = A(i, 0, k) - x1;
      B(i, k, 1) = y1 - A(i, 0, kc);
      B(i, kc, 1) = A(i, 0, kc) + y1;
    }
  }
#undef A
#undef B
}

/* stafft.f90:1463-1571, a(0:nv-1,0:lv-1,0:3) -> b(0:nv-1,0:3,0:lv-1) */
static void revrdx4(const double *a, double *b, long nv, long lv, const double *cosine,
                    const double *sine) {
#define A(i, k, j) a[(i) + nv * ((k) + lv * (j))]
#define B(i, j, k) b[(i) + nv * ((j) + 4 * (k))]
#define CS(k, j) cosine[(k) + lv * ((j)-1)]
#define SN(k, j) sine[(k) + lv * ((j)-1)]
  const double rtf12 = 0.7071067811865475244008443621048490392848359376884740365883398689953662;
  for (long i = 0; i < nv; i++) {
    double t1r = A(i, 0, 0) + A(i, 0, 2);
    double t2r = A(i, 0, 1);
    double t3r = A(i, 0, 0) - A(i, 0, 2);
    double t4r = A(i, 0, 3);
    B(i, 0, 0) = t1r + t2r;
    B(i, 1, 0) = t3r + t4r;
    B(i, 2, 0) = t1r - t2r;
    B(i, 3, 0) = t3r - t4r;
  }
  for (long k = 1; k <= (lv - 1) / 2; k++) {
    long kc = lv - k;
    double c1k = CS(k, 1), s1k = SN(k, 1), c2k = CS(k, 2), s2k = SN(k, 2);
    double c3k = CS(k, 3), s3k = SN(k, 3);
    for (long i = 0; i < nv; i++) {
      double t1r = A(i, k, 0) + A(i, kc, 1);
      double t1i = A(i, kc, 3) - A(i, k, 2);
      double t2r = A(i, k, 1) + A(i, kc, 0);
      double t2i = A(i, kc, 2) - A(i, k, 3);
      double t3r = A(i, k, 0) - A(i, kc, 1);
      double t3i = A(i, kc, 3) + A(i, k, 2);
      double t4r = A(i, k, 1) - A(i, kc, 0);
      double t4i = A(i, kc, 2) + A(i, k, 3);
      double x1p = t3r + t4i;
      double y1p = t3i - t4r;
      double x2p = t1r - t2r;
      double y2p = t1i - t2i;
      double x3p = t3r - t4i;
      double y3p = t3i + t4r;
      B(i, 0, k) = t1r + t2r;
      B(i, 0, kc) = t1i + t2i;
      B(i, 1, k) = c1k * x1p - s1k * y1p;
      B(i, 1, kc) = c1k * y1p + s1k * x1p;
      B(i, 2, k) = c2k * x2p - s2k * y2p;
      B(i, 2, kc) = c2k * y2p + s2k * x2p;
      B(i, 3, k) = c3k * x3p - s3k * y3p;
      B(i, 3, kc) = c3k * y3p + s3k * x3p;
    }
  }
  if (lv % 2 == 0) {
    long lvd2 = lv / 2;
    for (long i = 0; i < nv; i++) {
      B(i, 0, lvd2) = A(i, lvd2, 0) + A(i, lvd2, 1);
      B(i, 2, lvd2) = A(i, lvd2, 3) - A(i, lvd2, 2);
      double t3r = A(i, lvd2, 0) - A(i, lvd2, 1);
      double t4r = A(i, lvd2, 3) + A(i, lvd2, 2);
      B(i, 1, lvd2) = rtf12 * (t3r + t4r);
      B(i, 3, lvd2) = rtf12 * (t4r - t3r);
    }
  }
#undef A
#undef B
#undef CS
#undef SN
}

/* stafft.f90:1656-1706, a(0:nv-1,0:lv-1,0:1) -> b(0:nv-1,0:1,0:lv-1) */
static void revrdx2(const double *a, double *b, long nv, long lv, const double *cosine,
                    const double *sine) {
#define A(i, k, j) a[(i) + nv * ((k) + lv * (j))]
#define B(i, j, k) b[(i) + nv * ((j) + 2 * (k))]
  for (long i = 0; i < nv; i++) {
    B(i, 0, 0) = A(i, 0, 0) + A(i, 0, 1);
    B(i, 1, 0) = A(i, 0, 0) - A(i, 0, 1);
  }
  for (long k = 1; k <= (lv - 1) / 2; k++) {
    long kc = lv - k;
    double c1k = cosine[k], s1k = sine[k];
    for (long i = 0; i < nv; i++) {
      double x1p = A(i, k, 0) - A(i, kc, 0);
      double y1p = A(i, kc, 1) + A(i, k, 1);
      B(i, 0, k) = A(i, k, 0) + A(i, kc, 0);
      B(i, 0, kc) = A(i, kc, 1) - A(i, k, 1);
      B(i, 1, k) = c1k * x1p - s1k * y1p;
      B(i, 1, kc) = c1k * y1p + s1k * x1p;
    }
  }
#undef A
#undef B
}

/* stafft.f90:187-287 (radix 2 then radix 4 for power-of-two lengths). */
void hyo_forfft(int m, int n, double *x, const double *trig, const int factors[5]) {
  long tot = (long)m * n;
  double *wk = (double *)malloc(sizeof(double) * tot);
  int orig = 1;
  long rem = n, cum = 1;
  for (int i = 0; i < factors[2]; i++) {
    rem /= 2;
    long iloc = (rem - 1) * 2 * cum;
    if (orig) forrdx2(x, wk, m * rem, cum, trig + iloc, trig + n + iloc);
    else forrdx2(wk, x, m * rem, cum, trig + iloc, trig + n + iloc);
    orig = !orig;
    cum *= 2;
  }
  for (int i = 0; i < factors[1]; i++) {
    rem /= 4;
    long iloc = (rem - 1) * 4 * cum;
    if (orig) forrdx4(x, wk, m * rem, cum, trig + iloc, trig + n + iloc);
    else forrdx4(wk, x, m * rem, cum, trig + iloc, trig + n + iloc);
    orig = !orig;
    cum *= 4;
  }
  double normfac = 1.0 / sqrt((double)n);
  if (orig) {
    for (long i = 0; i < tot; i++) x[i] = x[i] * normfac;
  } else {
    for (long i = 0; i < tot; i++) x[i] = wk[i] * normfac;
  }
  free(wk);
}

/* stafft.f90:290-406 (radix 4 then radix 2 for power-of-two lengths). */
void hyo_revfft(int m, int n, double *x, const double *trig, const int factors[5]) {
  long tot = (long)m * n;
  double *wk = (double *)malloc(sizeof(double) * tot);
  for (long i = (long)(n / 2 + 1) * m; i <= tot - 1; i++) x[i] = -x[i];
  for (long i = 0; i < m; i++) x[i] = 0.5 * x[i];
  if (n % 2 == 0) {
    long k = (long)m * n / 2;
    for (long i = 0; i < m; i++) x[k + i] = 0.5 * x[k + i];
  }
  int orig = 1;
  long cum = 1, rem = n;
  for (int i = 0; i < factors[1]; i++) {
    rem /= 4;
    long iloc = (cum - 1) * 4 * rem;
    if (orig) revrdx4(x, wk, m * cum, rem, trig + iloc, trig + n + iloc);
    else revrdx4(wk, x, m * cum, rem, trig + iloc, trig + n + iloc);
    orig = !orig;
    cum *= 4;
  }
  for (int i = 0; i < factors[2]; i++) {
    rem /= 2;
    long iloc = (cum - 1) * 2 * rem;
    if (orig) revrdx2(x, wk, m * cum, rem, trig + iloc, trig + n + iloc);
    else revrdx2(wk, x, m * cum, rem, trig + iloc, trig + n + iloc);
    orig = !orig;
    cum *= 2;
  }
  double normfac = 2.0 / sqrt((double)n);
  if (orig) {
    for (long i = 0; i < tot; i++) x[i] = x[i] * normfac;
  } else {
    for (long i = 0; i < tot; i++) x[i] = wk[i] * normfac;
  }
  free(wk);
}

/* sta2dfft.f90:103-132.  rvar(ny,nx) is destroyed; svar(nx,ny) returned. */
void hyo_ptospc(const hyo_model *m, double *rvar, double *svar) {
  int nx = m->nx, ny = m->ny;
  hyo_forfft(ny, nx, rvar, m->xtrig, m->xfactors);
  for (int kx = 0; kx < nx; kx++)
    for (int iy = 0; iy < ny; iy++) svar[(long)iy * nx + kx] = rvar[(long)kx * ny + iy];
  hyo_forfft(nx, ny, svar, m->ytrig, m->yfactors);
}

/* sta2dfft.f90:136-165.  svar(nx,ny) is destroyed; rvar(ny,nx) returned. */
void hyo_spctop(const hyo_model *m, double *svar, double *rvar) {
  int nx = m->nx, ny = m->ny;
  hyo_revfft(nx, ny, svar, m->ytrig, m->yfactors);
  for (int kx = 0; kx < nx; kx++)
    for (int iy = 0; iy < ny; iy++) rvar[(long)kx * ny + iy] = svar[(long)iy * nx + kx];
  hyo_revfft(ny, nx, rvar, m->xtrig, m->xfactors);
}

/* sta2dfft.f90:169-203 (rkx is the 1-based wavenumber table, rkx[0] = rkx(1)). */
void hyo_xderiv(int nx, int ny, const double *rkx, const double *var, double *der) {
  int nwx = nx / 2, nxp2 = nx + 2;
  for (int ky = 1; ky <= ny; ky++) {
    const double *v = var + (long)(ky - 1) * nx - 1; /* 1-based view */
    double *d = der + (long)(ky - 1) * nx - 1;
    d[1] = 0.0;
    for (int kx = 2; kx <= nx - nwx; kx++) {
      int dkx = 2 * (kx - 1);
      int kxc = nxp2 - kx;
      d[kx] = -rkx[dkx - 1] * v[kxc];
      d[kxc] = rkx[dkx - 1] * v[kx];
    }
  }
  if (nx % 2 == 0) {
    int kxc = nwx + 1;
    for (int ky = 1; ky <= ny; ky++) der[(long)(ky - 1) * nx + kxc - 1] = 0.0;
  }
}

/* sta2dfft.f90:207-245 */
void hyo_yderiv(int nx, int ny, const double *rky, const double *var, double *der) {
  int nwy = ny / 2, nyp2 = ny + 2;
  for (int kx = 0; kx < nx; kx++) der[kx] = 0.0;
  for (int ky = 2; ky <= ny - nwy; ky++) {
    int kyc = nyp2 - ky;
    double fac = rky[2 * (ky - 1) - 1];
    for (int kx = 0; kx < nx; kx++) {
      der[(long)(ky - 1) * nx + kx] = -fac * var[(long)(kyc - 1) * nx + kx];
      der[(long)(kyc - 1) * nx + kx] = fac * var[(long)(ky - 1) * nx + kx];
    }
  }
  if (ny % 2 == 0) {
    int kyc = nwy + 1;
    for (int kx = 0; kx < nx; kx++) der[(long)(kyc - 1) * nx + kx] = 0.0;
  }
}
