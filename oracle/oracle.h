/* Oracle: CPU restatement of SimFast21's ionization-field step
 * (get_HIIbubbles.c and the helpers it calls).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * drop-in CLI, the Python package) may link, load or call this code; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only as a checker / baseline.
 *
 * Every function cites the reference lines it restates.  Pins (what checks
 * each function independently of its own formula) are listed in the tests:
 * numpy FFTs, mpmath evaluations, brute-force dilation on tiny boxes, closed
 * forms for the ladder and the criterion truth table.
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- astrophysical fits (user_functions.c) ---- */
double orc_Rion(float hmass, double z);                     /* user_functions.c:40-47 */
double orc_Rrec(float overdensity, double z,                 /* user_functions.c:58-65 */
                double L_h, long n, double hubble);
double orc_Qion(double z);                                   /* user_functions.c:84-88 */
double orc_G_H(double z);                                    /* user_functions.c:92-97 */
double orc_XHI(double ratio);                                /* user_functions.c:101-106 */

/* ---- auxiliary.c ---- */
double orc_W_filter(double x);                               /* auxiliary.c:350-359 */
long orc_check_borders(long x, long n);                      /* auxiliary.c:537-544 */

/* ---- radius ladder, get_HIIbubbles.c:192, :298-299, :397, :400-402, :461 ----
 * Writes the radii visited by the two loops (main: R >= 2dx, then small:
 * R >= dx) into radii[0..cap) and returns their count (or -1 if cap is too
 * small); *n_main receives the main-loop count. */
int orc_ladder(double rmax, double dx, int nbins, double *radii, int cap, int *n_main);

/* Largest integer m with sqrt((double)m)*dx <= R: the squared-offset bound of
 * the ball tests get_HIIbubbles.c:366, :443, :447. */
long orc_d2max(double R, double dx);

/* ---- per-cell pre-pass in SFR mode (get_HIIbubbles.c:216-225, :269-274) ----
 * delta, sfrd: n3 floats.  Outputs fresid, rec, ion (n3 floats). */
void orc_prepass_sfr(long n3, const float *delta, const float *sfrd, double z,
                     double fesc, double dx_h, double L_h, long n, double hubble,
                     float *fresid, float *rec, float *ion);

/* ---- halo mode emissivity (get_HIIbubbles.c:229-259) ----
 * CIC (auxiliary.c:761-808) of Rion on an nh^3 grid, smooth_box
 * (auxiliary.c:643-682) down to n^3, times (dx_s/dx_h)^3 * fesc. */
void orc_halo_emissivity(long nhalos, const float *halos /* M,x,y,z */, double z,
                         long nh, long n, double dx_halo, double dx_smooth,
                         double fesc, float *ion);
void orc_cic(float x1, float y1, float z1, float v, float *map, long n);  /* auxiliary.c:761-808 */
void orc_smooth_box(const float *box, float *out, long n, long ns);       /* auxiliary.c:643-682 */

/* ---- per-slice bubble finder (get_HIIbubbles.c:277-479) ---- */
typedef struct {
  int n;                /* N_smoothed */
  double dx;            /* cell size, same unit as radii */
  double dk;            /* 2*pi/L */
  int nr;               /* number of radii */
  const double *radii;  /* descending ladder; R >= 2dx -> main-loop dataflow */
  int precision;        /* 32: reference typing, FP32 FFT; 64: FP64 throughout */
  int dilation;         /* 0: reference dataflow (FFT convolution + stencil)
                           1: exact ball dilation (per-radius squared EDT) */
} orc_hii_params;

/* rec, ion: n^3 floats (Rrec and emissivity boxes).  fresid may be NULL
 * (treated as 0).  Outputs: xhii (n^3 floats), *neutral = mean(1 - xhii) in
 * FP64 index order.  first_centre (nullable, n^3 ints): index of the largest
 * radius at which the cell is a bubble centre, -1 if never. */
int orc_hii_slice(const orc_hii_params *p, const float *rec, const float *ion,
                  const float *fresid, float *xhii, double *neutral,
                  int *first_centre);

/* One radius of the filtered fields (get_HIIbubbles.c:305-325): returns the
 * unnormalised c2r outputs n^3 * (W_R * rec) and n^3 * (W_R * ion).
 * _d: FP64 throughout; _f: reference typing (FP32 FFT, FP64 W). */
void orc_filtered_d(int n, double dk, double R, const float *rec, const float *ion,
                    double *rec_r, double *ion_r);
void orc_filtered_f(int n, double dk, double R, const float *rec, const float *ion,
                    float *rec_r, float *ion_r);

/* Exact periodic dilation of a weighted centre map: out[x] = 1 iff some y
 * with w[y] >= 0 has min-image |x-y|^2 <= w[y].  Brute force over centres;
 * meant for small boxes in tests. */
void orc_dilate_weighted_brute(int n, const int *w, unsigned char *out);
/* The same map by one squared Euclidean distance transform per distinct
 * bound (Felzenszwalb-Huttenlocher lower envelopes); for full-size boxes. */
void orc_dilate_weighted_edt(int n, const int *w, unsigned char *out);
/* The same map by the separable max-plus transform (per-line lower
 * envelopes, three axes); O(N^3) whatever the number of distinct bounds. */
void orc_dilate_weighted_maxplus(int n, const int *w, unsigned char *out);

/* 20-bin spherically averaged power spectrum of a real box (FP64), linear
 * bins in |k| index space: see orc_power_spectrum in hii_oracle.c. */
void orc_power_spectrum(int n, double L, const double *field, int nbins,
                        double *pk, double *kmean, long *counts);

/* ---- the bubble finder's inputs (get_SFR.c, get_nldensity.c) ---- */
void orc_sfrd(long nhalos, const float *halos /* M,x,y,z */, double z, long nh, long n,
              double dx_halo, float *sfrd, double *mean);           /* get_SFR.c:113-148 */
void orc_nldensity(long nh, long n, double L_h, double hubble, double growth_in, double growth,
                   const float *vx, const float *vy, const float *vz, const float *delta0,
                   float *box_halo, float *out);                    /* get_nldensity.c:95-166 */

/* ---- 21-cm step (t21.c) and its background cosmology (auxiliary.c) ---- */
typedef struct orc_cosmo {
  double hubble, omega_m, omega_b, omega_lambda;
} orc_cosmo;
double orc_hubble_z(const orc_cosmo *c, double z);            /* auxiliary.c:224-226 */
double orc_growth(const orc_cosmo *c, double z);              /* auxiliary.c:242-266 */
double orc_dgrowth_dt(const orc_cosmo *c, double z);          /* auxiliary.c:285-292 */
/* t21.c:121-251 for one redshift (xc, xalpha, tk all NULL: TS >> Tcmb);
 * out: the float box written to deltaTb/; *mean_ts only in the spin branch */
void orc_t21(int n, const orc_cosmo *c, int vel_comp, double z, const float *delta,
             const float *xhii, const float *xc, const float *xalpha, const float *tk,
             float *out, double *mean_t21, double *mean_ts);

#ifdef __cplusplus
}
#endif
#endif
