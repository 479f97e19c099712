/* Oracle: CPU restatement of the 21-cm brightness-temperature step
 * (t21.c:116-252) and the background cosmology it uses (auxiliary.c:224-303).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  FP64 throughout, as the
 * reference computes it, with this directory's own FFT. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fft_cpu.h"
#include "oracle.h"

#define ORC_H0 3.24078e-18 /* Input_variables.h:108, h/s */

/* auxiliary.c:224-226 */
double orc_hubble_z(const orc_cosmo *c, double z) {
  return c->hubble * ORC_H0 * sqrt(c->omega_m * (1. + z) * (1. + z) * (1. + z) + c->omega_lambda);
}

/* auxiliary.c:242-256: midpoint sum over [zz, zz + 100) in steps of 1e-3,
 * loop variable accumulated in FP64 */
static double orc_int_growth(const orc_cosmo *c, double zz) {
  const double dz = 0.001, h0 = orc_hubble_z(c, 0.);
  double s = 0., z;
  for (z = zz + dz / 2.; z < zz + 100.; z += dz) {
    const double e = orc_hubble_z(c, z) / h0;
    s += (1. + z) / (e * e * e);
  }
  return s * dz;
}

/* auxiliary.c:262-266 */
double orc_growth(const orc_cosmo *c, double z) {
  return orc_hubble_z(c, z) / orc_hubble_z(c, 0.) * orc_int_growth(c, z) / orc_int_growth(c, 0.);
}

/* auxiliary.c:285-292 with dzdt (:232-234) */
double orc_dgrowth_dt(const orc_cosmo *c, double z) {
  const double dz = 1e-3;
  return (orc_growth(c, z + dz / 2.) - orc_growth(c, z - dz / 2.)) / dz *
         (-orc_hubble_z(c, z) * (1. + z));
}

/* t21.c:121-251 for one redshift.  xc, xalpha, tk all NULL: the TS >> Tcmb
 * branch (t21 = 1 before the density and velocity factors). */
void orc_t21(int n, const orc_cosmo *c, int vel_comp, double z, const float *delta,
             const float *xhii, const float *xc, const float *xalpha, const float *tk,
             float *out, double *mean_t21, double *mean_ts) {
  const long n3 = (long)n * n * n, h = n / 2 + 1;
  double *t21 = (double *)malloc(sizeof(double) * n3);
  double *dv = (double *)malloc(sizeof(double) * n3);
  double *dvk = (double *)malloc(sizeof(double) * 2 * n * n * h);
  double ts = 0., aver = 0.;
  long i;
  const double tcmb = 2.725 * (1. + z);
  if (xc && xalpha && tk) {
    for (i = 0; i < n3; i++) {
      const double xtot = (double)xalpha[i] + (double)xc[i];
      t21[i] = xtot / (1. + xtot);
      t21[i] = t21[i] * (1. - tcmb / (double)tk[i]);
      ts += tcmb / (1. - t21[i]);
    }
  } else {
    for (i = 0; i < n3; i++) t21[i] = 1.0;
  }
  for (i = 0; i < n3; i++) dv[i] = (double)delta[i];
  orc_r2c3d_d(n, dv, dvk);
  /* velocity gradient along the chosen axis: (k_axis/k)^2, k = 0 dropped */
  for (long a = 0; a < n; a++) {
    const long ia = a > n / 2 ? a - n : a;
    for (long j = 0; j < n; j++) {
      const long ij = j > n / 2 ? j - n : j;
      for (long p = 0; p < h; p++) {
        const long idx = (a * n + j) * h + p;
        if (a == 0 && j == 0 && p == 0) {
          dvk[2 * idx] = dvk[2 * idx + 1] = 0.;
          continue;
        }
        const double kk2 = (double)(ia * ia + ij * ij + p * p);
        const double ka = vel_comp == 1 ? (double)ia : (vel_comp == 2 ? (double)ij : (double)p);
        const double f = ka * ka / kk2;
        dvk[2 * idx] *= f;
        dvk[2 * idx + 1] *= f;
      }
    }
  }
  orc_c2r3d_d(n, dvk, dv);
  {
    const double growth = orc_growth(c, z), dgdt = orc_dgrowth_dt(c, z), hz = orc_hubble_z(c, z);
    const double h_ = c->hubble;
    const double cfac = 23.0 / 1000. * (0.7 / h_) * (c->omega_b * h_ * h_ / 0.02) *
                        sqrt((0.15 / c->omega_m / h_ / h_) * (1. + z) / 10.);
    for (i = 0; i < n3; i++) {
      double v = -dgdt * dv[i] / (double)n3 / hz / growth; /* dx^3/L^3 = 1/N^3 */
      if (v > 1.0) v = 1.0;
      else if (v < -0.5) v = -0.5;
      t21[i] = cfac * (1. + (double)delta[i]) * t21[i] / (1. + v);
      t21[i] = t21[i] * (1. - (double)xhii[i]);
      out[i] = (float)t21[i];
      aver += t21[i];
    }
  }
  *mean_t21 = aver / (double)n3;
  if (mean_ts) *mean_ts = (xc && xalpha && tk) ? ts / (double)n3 : 0.0;
  free(t21);
  free(dv);
  free(dvk);
}
