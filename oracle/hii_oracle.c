/* Oracle: CPU restatement of SimFast21's ionization-field step.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h.  Built with gcc -fopenmp.
 *
 * Reference files (SimFast21): get_HIIbubbles.c, user_functions.c,
 * auxiliary.c, Input_variables.c.  Line numbers are cited per function.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include "oracle.h"
#include "fft_cpu.h"

#define CM_PER_MPC 3.0857e24   /* user_functions.c:26 */

/* ------------------------------------------------------------------ */
/* Astrophysical fits: user_functions.c:32-106 (constants as published */
/* there; the formulas are restated, not copied).                       */
/* ------------------------------------------------------------------ */
#define RION_C1 1.575e39
#define RION_C2 9.367608e+07
#define RION_C3 0.44
#define RION_C4 2.27044695e+00
#define V_NORM 1.e+45
#define RREC_R1 9.84815696
#define RREC_R2 1.76105133
#define RREC_R3 0.81722878
#define RREC_R4 5.06773134
#define QION_A 2.05707274e-02
#define QION_B 5.29944413e+01

/* user_functions.c:40-47: ionising rate of a halo of mass m, per V_norm. */
double orc_Rion(float hmass, double z) {
  double t1 = RION_C1 * hmass * pow(1.0 + z, RION_C4);
  double t2 = pow(hmass / RION_C2, RION_C3);
  double t3 = exp(-pow(RION_C2 / hmass, 3.0));
  return (t1 * t2 * t3) / V_NORM;
}

/* user_functions.c:58-65: recombination rate of a cell with 1+delta =
 * overdensity; L_h is the box side in Mpc/h (global_L after
 * Input_variables.c:121), so L_h/hubble is Mpc. */
double orc_Rrec(float overdensity, double z, double L_h, long n, double hubble) {
  double a = 1.e-24 * RREC_R1 * pow(1.0 + z, RREC_R4);
  double x = pow(overdensity / RREC_R2, RREC_R3);
  double y = pow(x / (1. + x), 4.);
  double dvv = pow(L_h * CM_PER_MPC / n / hubble / (1. + z), 3.0);
  return a * y * dvv / V_NORM;
}

/* user_functions.c:84-88 */
double orc_Qion(double z) { return pow(10.0, QION_A * z + QION_B) / V_NORM; }

/* user_functions.c:92-97: recombination coefficient over the Haardt & Madau
 * (2012) background, a degree-5 polynomial in z inside a power of ten. */
double orc_G_H(double z) {
  double e1 = -4.37152818e-05 * pow(z, 5.) + 1.64394635e-03 * pow(z, 4.) - 2.03044756e-02 * pow(z, 3.);
  double e2 = 7.18335221e-02 * pow(z, 2.) - 8.85791388e-02 * z - 1.20887500e+01;
  return 4.19232273531e-13 / pow(10.0, e1 + e2);
}

/* user_functions.c:101-106: Popping et al. (2009) residual neutral fraction,
 * evaluated with the reference's own (cancelling) expression in FP64. */
double orc_XHI(double r) {
  double num, q;
  if (r == 0.0) return 0.0;
  q = 2. * r + 1.;
  num = q - sqrt(q * q - 4. * r * r);
  return num / (2. * r);
}

/* auxiliary.c:350-359: Fourier transform of a unit-volume top hat,
 * 3 j1(x)/x, with the reference's series branch below x = 0.03. */
double orc_W_filter(double x) {
  double j;
  if (x < 0.03) j = 1.0 / 3.0 - x * x / 30.0;
  else j = -(cos(x) - sin(x) / x) / x / x;
  return 3.0 * j;
}

/* auxiliary.c:537-544 */
long orc_check_borders(long x, long n) {
  x = x % n;
  if (x < 0) x += n;
  return x;
}

/* get_HIIbubbles.c:192 (bfactor), :298-299 and :397 (main loop, R >= 2dx),
 * :400-402 and :461 (small loop, R >= dx).  R is divided by bfactor
 * repeatedly in FP64, exactly as the loops do. */
int orc_ladder(double rmax, double dx, int nbins, double *radii, int cap, int *n_main) {
  double bfactor = pow(10.0, log10(rmax / dx) / nbins);
  double R = rmax;
  int k = 0, m;
  while (R >= 2 * dx) { if (k >= cap) return -1; radii[k++] = R; R /= bfactor; }
  m = k;
  while (R >= dx) { if (k >= cap) return -1; radii[k++] = R; R /= bfactor; }
  if (n_main) *n_main = m;
  return k;
}

/* Bound of the ball tests `sqrt(d2)*dx <= R` (get_HIIbubbles.c:366, :443,
 * :447): sqrt and the product are monotone, so the admissible d2 form a
 * prefix of the integers; find its last element exactly. */
long orc_d2max(double R, double dx) {
  long m;
  if (R < 0) return -1;
  m = (long)floor((R / dx) * (R / dx));
  while (m > 0 && !(sqrt((double)m) * dx <= R)) m--;
  while (sqrt((double)(m + 1)) * dx <= R) m++;
  return m;
}

/* get_HIIbubbles.c:216-225 (fresid, rec) and :269-274 (SFR -> emissivity),
 * with the reference's evaluation order and float roundings. */
void orc_prepass_sfr(long n3, const float *delta, const float *sfrd, double z,
                     double fesc, double dx_h, double L_h, long n, double hubble,
                     float *fresid, float *rec, float *ion) {
  long i;
  double gh = orc_G_H(z), qion = orc_Qion(z), z3 = pow(1. + z, 3.0);
  #pragma omp parallel for schedule(static)
  for (i = 0; i < n3; i++) {
    float f = (float)((1. + delta[i]) * 1.881e-7 * z3 * gh);
    fresid[i] = (float)orc_XHI(f);
    rec[i] = (float)orc_Rrec((float)(1.0 + delta[i]), z, L_h, n, hubble);
    ion[i] = (float)(sfrd[i] * dx_h * dx_h * dx_h * qion * fesc);
  }
}

/* auxiliary.c:761-808: cloud-in-cell deposit of v at (x1,y1,z1) (cell
 * units) with periodic wrap.  Per axis: truncate toward zero; a negative
 * remainder steps one cell down and its magnitude becomes the far weight
 * (:764-784).  Weights are formed the way the reference rounds them: the
 * far fraction is first stored back into a float (1 - frac), and each
 * factor 1 - w, their product and the accumulation are FP64, rounded once
 * into the float map. */
void orc_cic(float x1, float y1, float z1, float v, float *map, long n) {
  long x = (long)x1, y = (long)y1, z = (long)z1;
  float f[3];
  double w[3][2];
  int a, b, c, d;
  f[0] = x1 - (float)x; f[1] = y1 - (float)y; f[2] = z1 - (float)z;
  if (f[0] < 0.0f) { x -= 1; f[0] = fabsf(f[0]); }
  if (f[1] < 0.0f) { y -= 1; f[1] = fabsf(f[1]); }
  if (f[2] < 0.0f) { z -= 1; f[2] = fabsf(f[2]); }
  for (d = 0; d < 3; d++) {
    float far = (float)(1.0 - f[d]);
    w[d][0] = 1.0 - f[d];
    w[d][1] = 1.0 - far;
  }
  for (a = 0; a < 2; a++)
    for (b = 0; b < 2; b++)
      for (c = 0; c < 2; c++) {
        long xi = orc_check_borders(x + a, n), yi = orc_check_borders(y + b, n),
             zi = orc_check_borders(z + c, n);
        long o = xi * n * n + yi * n + zi;
        map[o] = (float)(map[o] + w[0][a] * w[1][b] * w[2][c] * v);
      }
}

/* auxiliary.c:643-682: block average of an n^3 box down to ns^3 with
 * sm = round(n/ns), reduced by one if ns*sm > n; FP64 accumulation. */
void orc_smooth_box(const float *box, float *out, long n, long ns) {
  long sm = (long)roundf((1.0f * n) / ns), i;
  double sm3;
  if (ns * sm > n) sm--;
  sm3 = 1.0 * sm * sm * sm;
  #pragma omp parallel for schedule(static)
  for (i = 0; i < ns * ns * ns; i++) {
    long a = i / (ns * ns), b = (i / ns) % ns, c = i % ns, u, v, w;
    double acc = 0.0;
    for (u = 0; u < sm; u++)
      for (v = 0; v < sm; v++)
        for (w = 0; w < sm; w++)
          acc += box[(a * sm + u) * n * n + (b * sm + v) * n + (c * sm + w)];
    out[i] = (float)(acc / sm3);
  }
}

/* get_HIIbubbles.c:229-259: halo-mode emissivity. */
void orc_halo_emissivity(long nhalos, const float *halos, double z, long nh, long n,
                         double dx_halo, double dx_smooth, double fesc, float *ion) {
  float *hm = (float *)calloc((size_t)nh * nh * nh, sizeof(float));
  long i;
  for (i = 0; i < nhalos; i++)
    orc_cic(halos[4 * i + 1], halos[4 * i + 2], halos[4 * i + 3],
            (float)orc_Rion(halos[4 * i], z), hm, nh);
  orc_smooth_box(hm, ion, nh, n);
  for (i = 0; i < n * n * n; i++)
    ion[i] = (float)(ion[i] * pow(dx_smooth / dx_halo, 3) * fesc);
  free(hm);
}

/* ------------------------------------------------------------------ */
/* Exact periodic dilation helpers                                     */
/* ------------------------------------------------------------------ */

/* 1D squared-distance lower envelope (Felzenszwalb & Huttenlocher 2012) of
 * f over a periodic line of length n: d[x] = min_y f[y] + minimg(x-y)^2.
 * The line is unrolled three times so every image within n is seen. */
static void edt1d_periodic(int n, const double *f, double *d, double *fe, int *v, double *zb) {
  int m = 3 * n, q, k = -1, x;
  for (q = 0; q < m; q++) fe[q] = f[q % n];
  for (q = 0; q < m; q++) {
    double s;
    if (!isfinite(fe[q])) continue;
    if (k < 0) { k = 0; v[0] = q; zb[0] = -INFINITY; zb[1] = INFINITY; continue; }
    for (;;) {
      s = ((fe[q] + (double)q * q) - (fe[v[k]] + (double)v[k] * v[k])) / (2.0 * q - 2.0 * v[k]);
      if (s <= zb[k]) { k--; if (k < 0) break; } else break;
    }
    if (k < 0) { k = 0; v[0] = q; zb[0] = -INFINITY; zb[1] = INFINITY; continue; }
    k++; v[k] = q; zb[k] = s; zb[k + 1] = INFINITY;
  }
  if (k < 0) { for (x = 0; x < n; x++) d[x] = INFINITY; return; }
  {
    int kk = 0;
    for (x = n; x < 2 * n; x++) {
      while (zb[kk + 1] < x) kk++;
      d[x - n] = (double)(x - v[kk]) * (x - v[kk]) + fe[v[kk]];
    }
  }
}

/* Squared min-image distance of every cell to the nearest cell with
 * mask != 0 (INFINITY if none). */
static void edt3d_periodic(int n, const unsigned char *mask, double *dist) {
  long n3 = (long)n * n * n, i;
  for (i = 0; i < n3; i++) dist[i] = mask[i] ? 0.0 : INFINITY;
  /* three axis passes: p (stride 1), j (stride n), i (stride n^2) */
  {
    int ax;
    for (ax = 0; ax < 3; ax++) {
      long stride = ax == 0 ? 1 : (ax == 1 ? n : (long)n * n);
      long l;
      #pragma omp parallel
      {
        double *f = (double *)malloc(sizeof(double) * n);
        double *d = (double *)malloc(sizeof(double) * n);
        double *fe = (double *)malloc(sizeof(double) * 3 * n);
        double *zb = (double *)malloc(sizeof(double) * (3 * n + 1));
        int *v = (int *)malloc(sizeof(int) * 3 * n);
        #pragma omp for schedule(static)
        for (l = 0; l < (long)n * n; l++) {
          long base, t;
          if (ax == 0) base = l * n;
          else if (ax == 1) base = (l / n) * n * n + (l % n);
          else base = l;
          for (t = 0; t < n; t++) f[t] = dist[base + t * stride];
          edt1d_periodic(n, f, d, fe, v, zb);
          for (t = 0; t < n; t++) dist[base + t * stride] = d[t];
        }
        free(f); free(d); free(fe); free(zb); free(v);
      }
    }
  }
}

void orc_dilate_weighted_brute(int n, const int *w, unsigned char *out) {
  long n3 = (long)n * n * n, y;
  memset(out, 0, (size_t)n3);
  for (y = 0; y < n3; y++) {
    int yi, yj, yp, a, b, c, r;
    if (w[y] < 0) continue;
    yi = (int)(y / ((long)n * n)); yj = (int)((y / n) % n); yp = (int)(y % n);
    r = (int)floor(sqrt((double)w[y])) + 1;
    if (r > n / 2) r = n / 2;
    for (a = -r; a <= r; a++)
      for (b = -r; b <= r; b++)
        for (c = -r; c <= r; c++) {
          long x = orc_check_borders(yi + a, n) * n * n + orc_check_borders(yj + b, n) * n +
                   orc_check_borders(yp + c, n);
          int da = abs(a), db = abs(b), dc = abs(c);
          if (n - da < da) da = n - da;
          if (n - db < db) db = n - db;
          if (n - dc < dc) dc = n - dc;
          if ((long)da * da + (long)db * db + (long)dc * dc <= w[y]) out[x] = 1;
        }
  }
}

/* Same result as orc_dilate_weighted_brute, by one squared EDT per
 * distinct bound: cells with w[y] == D dilate every x with min-image
 * |x-y|^2 <= D.  O(N^3) per distinct bound, for full-size boxes. */
void orc_dilate_weighted_edt(int n, const int *w, unsigned char *out) {
  const long n3 = (long)n * n * n;
  long i;
  int *vals = NULL, nv = 0, cap = 0;
  unsigned char *mask = (unsigned char *)malloc((size_t)n3);
  double *dist = (double *)malloc(sizeof(double) * n3);
  memset(out, 0, (size_t)n3);
  for (i = 0; i < n3; i++) {
    int k, seen = 0;
    if (w[i] < 0) continue;
    for (k = 0; k < nv; k++)
      if (vals[k] == w[i]) { seen = 1; break; }
    if (seen) continue;
    if (nv == cap) {
      cap = cap ? 2 * cap : 64;
      vals = (int *)realloc(vals, sizeof(int) * cap);
    }
    vals[nv++] = w[i];
  }
  for (int k = 0; k < nv; k++) {
    for (i = 0; i < n3; i++) mask[i] = w[i] == vals[k];
    edt3d_periodic(n, mask, dist);
    for (i = 0; i < n3; i++)
      if (dist[i] <= (double)vals[k]) out[i] = 1;
  }
  free(vals);
  free(mask);
  free(dist);
}

/* One periodic line of the separable max-plus transform:
 * g[x] = max over y with f[y] >= 0 of f[y] - dmin(x, y)^2, or -1 where that
 * maximum is negative or no y is present.  Minimised form: d(x) = min over
 * the images q = y - n, y, y + n of (x - q)^2 + h(q), h = -f, by the lower
 * envelope of parabolas (Felzenszwalb and Huttenlocher, "Distance
 * Transforms of Sampled Functions", 2012, Algorithm 1); for x in [0, n) the
 * three images contain the min-image one. */
static void maxplus_line(int n, const long *f, long *g, int *v, double *z) {
  int k = -1;
  for (int q = 0; q < 3 * n; q++) {
    const long fq = f[q % n];
    if (fq < 0) continue;
    const double cq = (double)(q - n) * (double)(q - n) - (double)fq;
    if (k < 0) {
      k = 0;
      v[0] = q;
      z[0] = -HUGE_VAL;
      z[1] = HUGE_VAL;
      continue;
    }
    /* intersection with the envelope's last parabola, in line positions
     * (image q sits at q - n); z[0] = -inf ends the loop at k = 0 */
    double sx;
    for (;;) {
      const int r = v[k];
      const double cr = (double)(r - n) * (double)(r - n) - (double)f[r % n];
      sx = (cq - cr) / (2.0 * (double)(q - r));
      if (sx <= z[k]) { k--; continue; }
      break;
    }
    k++;
    v[k] = q;
    z[k] = sx;
    z[k + 1] = HUGE_VAL;
  }
  if (k < 0) {
    for (int x = 0; x < n; x++) g[x] = -1;
    return;
  }
  int j = 0;
  for (int x = 0; x < n; x++) {
    while (z[j + 1] < (double)x) j++;
    const long d = (long)(x + n - v[j]);
    const long val = f[v[j] % n] - d * d;
    g[x] = val >= 0 ? val : -1;
  }
}

/* Same map as orc_dilate_weighted_brute: the separable max-plus transform
 * g(x) = max_y (w(y) - |x-y|^2) one axis at a time (min-image squared
 * distance is a sum of per-axis terms), out[x] = g(x) >= 0; a negative
 * partial maximum cannot become non-negative in a later pass and is
 * dropped.  O(N^3) for any number of distinct bounds. */
void orc_dilate_weighted_maxplus(int n, const int *w, unsigned char *out) {
  const long n3 = (long)n * n * n, nn = (long)n * n;
  long *a = (long *)malloc(sizeof(long) * n3);
  for (long i = 0; i < n3; i++) a[i] = w[i] >= 0 ? w[i] : -1;
  for (int axis = 0; axis < 3; axis++) {
    const long stride = axis == 0 ? 1 : (axis == 1 ? n : nn);
#pragma omp parallel
    {
      long *f = (long *)malloc(sizeof(long) * n), *g = (long *)malloc(sizeof(long) * n);
      int *v = (int *)malloc(sizeof(int) * 3 * n);
      double *z = (double *)malloc(sizeof(double) * (3 * n + 1));
#pragma omp for schedule(static)
      for (long line = 0; line < nn; line++) {
        /* line = (outer, inner) over the two other axes */
        const long o = line / n, in = line % n;
        long base;
        if (axis == 0) base = o * nn + in * n;
        else if (axis == 1) base = o * nn + in;
        else base = o * n + in;
        for (int e = 0; e < n; e++) f[e] = a[base + e * stride];
        maxplus_line(n, f, g, v, z);
        for (int e = 0; e < n; e++) a[base + e * stride] = g[e];
      }
      free(f); free(g); free(v); free(z);
    }
  }
  for (long i = 0; i < n3; i++) out[i] = a[i] >= 0;
  free(a);
}

/* ------------------------------------------------------------------ */
/* Per-slice pipeline, once in the reference's typing (float fields,   */
/* FP32 FFT, FP64 W and ratios) and once in FP64 throughout.           */
/* ------------------------------------------------------------------ */
#define T float
#define SUF _f
#define R2C orc_r2c3d_f
#define C2R orc_c2r3d_f
#include "hii_oracle_tmpl.h"
#undef T
#undef SUF
#undef R2C
#undef C2R

#define T double
#define SUF _d
#define R2C orc_r2c3d_d
#define C2R orc_c2r3d_d
#include "hii_oracle_tmpl.h"
#undef T
#undef SUF
#undef R2C
#undef C2R

int orc_hii_slice(const orc_hii_params *p, const float *rec, const float *ion,
                  const float *fresid, float *xhii, double *neutral, int *first_centre) {
  if (p->n < 4 || p->nr < 0) return -1;
  if (p->precision == 64) return hii_slice_d(p, rec, ion, fresid, xhii, neutral, first_centre);
  return hii_slice_f(p, rec, ion, fresid, xhii, neutral, first_centre);
}

void orc_filtered_d(int n, double dk, double R, const float *rec, const float *ion,
                    double *rec_r, double *ion_r) {
  filtered_pair_d(n, dk, R, rec, ion, rec_r, ion_r);
}

void orc_filtered_f(int n, double dk, double R, const float *rec, const float *ion,
                    float *rec_r, float *ion_r) {
  filtered_pair_f(n, dk, R, rec, ion, rec_r, ion_r);
}

/* Spherically averaged power spectrum with the normalisation of
 * tools/power3d.c:172-191 (top-hat window: wsum = n^3; P = <|F_k|^2> dl^3 /
 * wsum over half-spectrum entries, F unnormalised), in nbins linear bins of
 * the mode index |k|/dk between 1 and n/2:
 *   bin b holds 1 + b*(n/2-1)/nbins <= |k|/dk < 1 + (b+1)*(n/2-1)/nbins,
 * the last bin closed at n/2; k = 0 and corner modes beyond n/2 are dropped. */
void orc_power_spectrum(int n, double L, const double *field, int nbins,
                        double *pk, double *kmean, long *counts) {
  int h = n / 2 + 1, i, j, q, b;
  double *spec = (double *)malloc(sizeof(double) * 2 * (size_t)n * n * h);
  double dl = L / n, dk = 2.0 * M_PI / L, width = (n / 2 - 1) / (double)nbins;
  orc_r2c3d_d(n, field, spec);
  for (b = 0; b < nbins; b++) { pk[b] = 0; kmean[b] = 0; counts[b] = 0; }
  for (i = 0; i < n; i++) {
    int ii = i > n / 2 ? i - n : i;
    for (j = 0; j < n; j++) {
      int jj = j > n / 2 ? j - n : j;
      for (q = 0; q < h; q++) {
        double kq = sqrt((double)ii * ii + (double)jj * jj + (double)q * q);
        size_t o = 2 * (((size_t)i * n + j) * h + q);
        if (kq < 1.0 || kq > n / 2) continue;
        b = (int)floor((kq - 1.0) / width);
        if (b >= nbins) b = nbins - 1;
        pk[b] += spec[o] * spec[o] + spec[o + 1] * spec[o + 1];
        kmean[b] += kq * dk;
        counts[b]++;
      }
    }
  }
  for (b = 0; b < nbins; b++) {
    if (counts[b] > 0) {
      pk[b] = pk[b] / counts[b] / ((double)n * n * n) * dl * dl * dl;
      kmean[b] /= counts[b];
    }
  }
  free(spec);
}
