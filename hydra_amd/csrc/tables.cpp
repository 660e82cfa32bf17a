// tables.cpp -- host construction of the constants and spectral tables.
//
// These are computed once, on the host, with the same libm calls and in the
// same operation order as the reference, then uploaded:
//   constants.f90:44-75, contours.f90:19-42      (build_consts)
//   stafft.f90:63-184 (initfft/factorisen),
//   sta2dfft.f90:40-97 (init2dfft), spectral.f90:42-158 (build_tables)
// The FFT trig table, the Hou-Li filtered wavenumbers and the Butterworth
// filters enter every step, so they must be identical to the reference's.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>

#include "hyd_internal.h"

namespace hyd {

static const double kPi = 3.141592653589793238462643383279502884197169399375105820974944592307816;

// gfortran folds the constant glx**4 with one correctly rounded result.
static double pow4_folded(double x) {
  __float128 q = (__float128)x;
  q = q * q;
  q = q * q;
  return (double)q;
}

int build_consts(const hyd_params &p, Consts &c) {
  if (p.ng < 8 || (p.ng & (p.ng - 1)) != 0) {
    set_error("ng must be a power of two >= 8");
    return HYD_EFACTOR;
  }
  if (p.ellx <= 0.0 || p.elly <= 0.0) {
    set_error("domain lengths must be positive (sta2dfft.f90:64-93)");
    return HYD_EARG;
  }
  if (p.ellx != p.elly) {
    // con2grid's x-wrap uses the y length (contours.f90:300-303): nx=ny, L_x=L_y
    set_error("only square domains are supported (contours.f90:300-303 assumes nx=ny)");
    return HYD_EARG;
  }
  if (p.rtherm != 0.0) {
    set_error("thermal damping (rtherm>0) is not supported");
    return HYD_EARG;
  }
  memset(&c, 0, sizeof(c));
  c.nx = c.ny = p.ng;
  c.nxf = c.nyf = 4 * p.ng;
  c.nxu = c.nyu = 16 * p.ng;
  c.ellx = p.ellx;
  c.elly = p.elly;
  c.small = 1.0e-12;
  c.small3 = c.small * c.small * c.small;
  c.oms = 1.0 - c.small;
  c.hlx = 0.5 * p.ellx;
  c.hlxi = 1.0 / c.hlx;
  c.xmin = -c.hlx;
  c.xmax = c.hlx;
  c.hly = 0.5 * p.elly;
  c.hlyi = 1.0 / c.hly;
  c.ymin = -c.hly;
  c.ymax = c.hly;
  c.kdsq = p.kd * p.kd;
  c.srwfm = p.beta * p.kd / (1.0 + c.kdsq);
  c.glx = p.ellx / (double)c.nx;
  c.glxi = (double)c.nx / p.ellx;
  c.gly = p.elly / (double)c.ny;
  c.glyi = (double)c.ny / p.elly;
  c.garea = c.glx * c.gly;
  c.dsumi = 1.0 / (double)((int64_t)c.nx * c.ny);
  c.glxf = p.ellx / (double)c.nxf;
  c.glxfi = (double)c.nxf / p.ellx;
  c.glyf = p.elly / (double)c.nyf;
  c.glyfi = (double)c.nyf / p.elly;
  c.glxu = p.ellx / (double)c.nxu;
  c.glxui = (double)c.nxu / p.ellx;
  c.glyu = p.elly / (double)c.nyu;
  c.glyui = (double)c.nyu / p.elly;
  c.amu = 0.2;
  c.ell = (p.ell_cells > 0.0 ? p.ell_cells : 6.25) * c.glx;
  c.dm = c.amu * c.amu * c.ell / 4.0;
  c.dm2 = c.dm * c.dm;
  c.d4small = c.small * pow4_folded(c.glx);
  c.dmsq = 4.0 * c.dm2;
  c.dmi = 2.0 / c.dm;
  c.elf = 1.0 / (c.ell * c.ell);
  c.densf = 1.0 / (c.amu * sqrt(c.ell));
  c.btropic = p.kd == 0.0;
  c.beffect = p.beta != 0.0;
  return HYD_OK;
}

// stafft.f90:128-184 restricted to the radices used by power-of-two lengths.
static int factorise(int n, int f[5]) {
  int rem = n;
  for (int i = 0; i < 5; i++) f[i] = 0;
  while (rem % 6 == 0) { f[0]++; rem /= 6; if (rem == 1) return 0; }
  while (rem % 4 == 0) { f[1]++; rem /= 4; if (rem == 1) return 0; }
  while (rem % 2 == 0) { f[2]++; rem /= 2; if (rem == 1) return 0; }
  while (rem % 3 == 0) { f[3]++; rem /= 3; if (rem == 1) return 0; }
  while (rem % 5 == 0) { f[4]++; rem /= 5; if (rem == 1) return 0; }
  return 1;
}

// stafft.f90:63-125 -- same angle recurrence, then cos / -sin via libm.
static int trig_table(int n, int f[5], std::vector<double> &trig) {
  static const int fac[5] = {6, 4, 2, 3, 5};
  const double twopi = 6.283185307179586476925286766559005768394338798750211641949889184615633;
  if (factorise(n, f) || f[0] || f[3] || f[4]) return HYD_EFACTOR;
  trig.assign(2 * (size_t)n, 0.0);
  double ftwopin = twopi / (double)n;
  int rem = n, m = 0;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < f[i]; j++) {
      rem = rem / fac[i];
      for (int k = 1; k <= fac[i] - 1; k++)
        for (int l = 0; l <= rem - 1; l++) trig[m++] = ftwopin * (double)(k * l);
      ftwopin = ftwopin * fac[i];
    }
  for (int i = 0; i < n - 1; i++) {
    trig[n + i] = -sin(trig[i]);
    trig[i] = cos(trig[i]);
  }
  return HYD_OK;
}

int initfft_table(int n, int f[5], std::vector<double> &trig) { return trig_table(n, f, trig); }

int build_tables(const hyd_params &p, const Consts &c, HostTables &t) {
  const int nx = c.nx, ny = c.ny, nwx = nx / 2, nwy = ny / 2;
  if (trig_table(nx, t.xfactors, t.xtrig) || trig_table(ny, t.yfactors, t.ytrig)) {
    set_error("FFT length must be a power of two");
    return HYD_EFACTOR;
  }
  t.hrkx.resize(nx);
  t.hrky.resize(ny);
  t.rkx.resize(nx);
  t.rky.resize(ny);
  std::vector<double> rkxf(nx), rkyf(ny);
  // init2dfft wavenumbers, sta2dfft.f90:64-94
  double sc = kPi / p.ellx;
  for (int k = 1; k <= nx; k++) t.hrkx[k - 1] = sc * (double)k;
  sc = kPi / p.elly;
  for (int k = 1; k <= ny; k++) t.hrky[k - 1] = sc * (double)k;
  // spectral.f90:57-86
  t.rkx[0] = 0.0;
  for (int kx = 1; kx <= nwx - 1; kx++) {
    t.rkx[kx] = t.hrkx[2 * kx - 1];
    t.rkx[nx - kx] = t.hrkx[2 * kx - 1];
  }
  t.rkx[nwx] = t.hrkx[nx - 1];
  const double scx = 2.0 * kPi / p.ellx, rkxmax = scx * (double)nwx;
  for (int kx = 0; kx < nx; kx++) {
    rkxf[kx] = t.rkx[kx] * exp(-36.0 * pow(t.rkx[kx] / rkxmax, 36.0));
    t.hrkx[kx] = t.hrkx[kx] * exp(-36.0 * pow(t.hrkx[kx] / rkxmax, 36.0));
  }
  t.rky[0] = 0.0;
  for (int ky = 1; ky <= nwy - 1; ky++) {
    t.rky[ky] = t.hrky[2 * ky - 1];
    t.rky[ny - ky] = t.hrky[2 * ky - 1];
  }
  t.rky[nwy] = t.hrky[ny - 1];
  const double scy = 2.0 * kPi / p.elly, rkymax = scy * (double)nwy;
  for (int ky = 0; ky < ny; ky++) {
    rkyf[ky] = t.rky[ky] * exp(-36.0 * pow(t.rky[ky] / rkymax, 36.0));
    t.hrky[ky] = t.hrky[ky] * exp(-36.0 * pow(t.hrky[ky] / rkymax, 36.0));
  }
  // spectral.f90:95-106
  const size_t N = (size_t)nx * ny;
  t.green.assign(N, 0.0);
  t.flo.assign(N, 0.0);
  t.fhi.assign(N, 0.0);
  t.qdiss.assign(N, 0.0);
  t.rksq.assign(N, 0.0);
  t.rk.assign(N, 0.0);
  const double fac = 36.0 / (double)(nx * ny);
  const double a2 = rkxmax * rkxmax, b2 = rkymax * rkymax;
  const double rkmsi = 1.0 / std::max(a2, b2);
  for (int ky = 0; ky < ny; ky++)
    for (int kx = 0; kx < nx; kx++) {
      size_t k = (size_t)ky * nx + kx;
      double rksqu = t.rkx[kx] * t.rkx[kx] + t.rky[ky] * t.rky[ky];
      t.rk[k] = sqrt(rksqu);
      t.qdiss[k] = p.cdamp * pow(rkmsi * rksqu, p.nnu);
      double q = fac * rksqu;
      t.flo[k] = 1.0 / (1.0 + q * q);
      t.fhi[k] = 1.0 - t.flo[k];
      t.rksq[k] = rkxf[kx] * rkxf[kx] + rkyf[ky] * rkyf[ky];
    }
  // spectral.f90:110-133 (spectrum shells)
  const int nm = std::max(nx, ny);
  const double delk = sqrt(scx * scx + scy * scy), delki = 1.0 / delk;
  t.kmax = (int)lround(sqrt(a2 + b2) * delki);
  t.kmaxred = (int)lround(sqrt((a2 + b2) / 2.0) * delki);
  t.kmag.assign(N, 0);
  t.spmf.assign((size_t)nm + 1, 0.0);
  t.alk.assign((size_t)nm + 1, 0.0);
  for (int ky = 0; ky < ny; ky++)
    for (int kx = 0; kx < nx; kx++) {
      const int k = (int)lround(sqrt(t.rkx[kx] * t.rkx[kx] + t.rky[ky] * t.rky[ky]) * delki);
      t.kmag[(size_t)ky * nx + kx] = k;
      t.spmf[k] = t.spmf[k] + 1.0;
    }
  const double snorm = 4.0 * kPi / (double)(nx * ny);
  t.spmf[0] = 0.0;
  for (int k = 1; k <= t.kmax; k++) {
    t.spmf[k] = snorm * (double)k / t.spmf[k];
    t.alk[k] = log10(delk * (double)k);
  }
  // spectral.f90:136-147
  t.green[0] = 0.0;
  for (int kx = 2; kx <= nx; kx++) t.green[kx - 1] = -1.0 / (t.rkx[kx - 1] * t.rkx[kx - 1] + c.kdsq);
  for (int ky = 2; ky <= ny; ky++)
    t.green[(size_t)(ky - 1) * nx] = -1.0 / (t.rky[ky - 1] * t.rky[ky - 1] + c.kdsq);
  for (int ky = 2; ky <= ny; ky++)
    for (int kx = 2; kx <= nx; kx++)
      t.green[(size_t)(ky - 1) * nx + kx - 1] =
          -1.0 / (t.rkx[kx - 1] * t.rkx[kx - 1] + t.rky[ky - 1] * t.rky[ky - 1] + c.kdsq);
  t.bety.resize(ny);
  for (int iy = 1; iy <= ny; iy++) t.bety[iy - 1] = p.beta * (c.gly * (double)(iy - 1) + c.ymin);
  // Distinct qdiss values: exp(dfac*qdiss) is evaluated once per value.
  std::map<double, int32_t> idx;
  for (size_t k = 0; k < N; k++) idx.emplace(t.qdiss[k], 0);
  t.qdiss_u.clear();
  for (auto &kv : idx) {
    kv.second = (int32_t)t.qdiss_u.size();
    t.qdiss_u.push_back(kv.first);
  }
  t.qdiss_map.resize(N);
  for (size_t k = 0; k < N; k++) t.qdiss_map[k] = idx[t.qdiss[k]];
  return HYD_OK;
}

}  // namespace hyd
