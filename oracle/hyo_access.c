/*
 * hyo_access.c -- named accessors for the oracle model, so that the Python
 * test harness (oracle/hyo.py) can view and set state without mirroring the
 * C struct layout.  TEST INFRASTRUCTURE ONLY (see hyo.h).
 */
#include <string.h>

#include "hyo.h"

double *hyo_dfield(hyo_model *m, const char *name) {
#define D(n) \
  if (!strcmp(name, #n)) return m->n;
  D(rksq) D(qdiss) D(green) D(flo) D(fhi) D(rk) D(rkx) D(hrkx) D(rky) D(hrky) D(bety)
  D(xtrig) D(ytrig) D(xq) D(yq) D(qs) D(qd) D(qc) D(qq) D(pp) D(uu) D(vv) D(qr) D(emq)
  D(epq) D(qspre) D(xgf) D(ygf) D(xgu) D(ygu) D(spmf) D(alk)
#undef D
  return 0;
}

int *hyo_ifield(hyo_model *m, const char *name) {
#define I(n) \
  if (!strcmp(name, #n)) return m->n;
  I(nextq) I(indq) I(npq) I(i1q) I(i2q) I(ixp) I(iyp)
#undef I
  return 0;
}

double hyo_get_d(const hyo_model *m, const char *name) {
#define D(n) \
  if (!strcmp(name, #n)) return m->n;
  D(qjump) D(t) D(dt) D(dt2) D(dt3) D(dt6) D(twist) D(dtmax) D(zzrms) D(zzmax) D(uumax)
  D(dtacc) D(glx) D(dm) D(dm2) D(d4small) D(dmsq) D(dmi) D(elf) D(densf) D(ell) D(kdsq)
  D(xmin) D(ymin) D(hlx) D(hly) D(garea) D(dsumi)
#undef D
  return 0.0 / 0.0;
}

void hyo_set_d(hyo_model *m, const char *name, double v) {
#define D(n) \
  if (!strcmp(name, #n)) { m->n = v; return; }
  D(qjump) D(t) D(dt) D(twist) D(dtmax) D(d4small)
#undef D
}

long hyo_get_i(const hyo_model *m, const char *name) {
#define I(n) \
  if (!strcmp(name, #n)) return (long)m->n;
  I(nx) I(ny) I(nq) I(nptq) I(npm) I(nm) I(contex) I(ireg) I(phase) I(err) I(nxf) I(nyf) I(nxu) I(nyu)
  I(kmax) I(kmaxred)
#undef I
  return -1;
}

void hyo_set_i(hyo_model *m, const char *name, long v) {
#define I(n) \
  if (!strcmp(name, #n)) { m->n = (int)v; return; }
  I(nq) I(nptq) I(contex) I(ireg) I(phase)
#undef I
}
