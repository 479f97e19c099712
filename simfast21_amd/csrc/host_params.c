/* Host logic that needs no GPU: parameter file parsing, the radius ladder,
 * the ball bound and the background cosmology of the 21-cm step.  Part of
 * libsf21hii. */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/sf21_hii.h"

#define SF21_PI 3.1415926535897932385E0 /* Input_variables.h:10 */

const char *sf21_strerror(int code) {
  switch (code) {
    case SF21_OK: return "ok";
    case SF21_ERR_ARG: return "invalid argument or unsupported size";
    case SF21_ERR_ALLOC: return "allocation failed";
    case SF21_ERR_HIP: return "HIP runtime error";
    case SF21_ERR_IO: return "file missing, unreadable or short";
    case SF21_ERR_FORMAT: return "malformed parameter file";
    default: return "unknown error";
  }
}

static void set_msg(char *msg, int cap, const char *text) {
  if (msg && cap > 0) {
    strncpy(msg, text, (size_t)cap - 1);
    msg[cap - 1] = 0;
  }
}

/* Input_variables.c:16-141 for the keys get_HIIbubbles reads.
 * - lines whose first character is '#' or '\n' are skipped (:39);
 * - any other line must contain '=' (:40), which becomes a blank, and must
 *   yield two whitespace-separated tokens (:42-43) -- so text after the
 *   value (a trailing comment) is ignored;
 * - unknown keys are ignored; T/F flags accept exactly "T" or "F";
 * - derived values as :121-133. */
int sf21_params_read(const char *base_dir, sf21_params *p, char *msg, int msg_cap) {
  char path[4096], line[256], key[256], val[256], *eq;
  FILE *f;
  if (!base_dir || !p) return SF21_ERR_ARG;
  memset(p, 0, sizeof(*p));
  snprintf(path, sizeof(path), "%s/simfast21.ini", base_dir);
  f = fopen(path, "r");
  if (!f) {
    set_msg(msg, msg_cap, "\nThe parameter file simfast21.ini cannot be open. Exit...\n");
    return SF21_ERR_IO;
  }
  while (fgets(line, sizeof(line), f) != NULL) {
    int nt;
    if (line[0] == '#' || line[0] == '\n') continue;
    eq = strchr(line, '=');
    if (eq) *eq = ' ';
    nt = eq ? sscanf(line, "%255s %255s", key, val) : 0;
    if (!eq || nt != 2) {
      char buf[400];
      snprintf(buf, sizeof(buf), "Wrong format! Exiting...\n%s\n", line);
      set_msg(msg, msg_cap, buf);
      fclose(f);
      return SF21_ERR_FORMAT;
    }
    if (!strcmp(key, "nthreads")) p->nthreads = atoi(val);
    else if (!strcmp(key, "dz")) p->dz = atof(val);
    else if (!strcmp(key, "zmax")) p->zmax = atof(val);
    else if (!strcmp(key, "zmin")) p->zmin = atof(val);
    else if (!strcmp(key, "sim_length")) p->box_len = atof(val);
    else if (!strcmp(key, "N_halo")) p->n_halo = atoi(val);
    else if (!strcmp(key, "N_smoothed")) p->n = atoi(val);
    else if (!strcmp(key, "hubble")) p->hubble = atof(val);
    else if (!strcmp(key, "fesc")) p->fesc = atof(val);
    else if (!strcmp(key, "Ion_cutoff")) p->ion_cutoff = atof(val);
    else if (!strcmp(key, "bubble_Rmax")) p->bubble_rmax = atof(val);
    else if (!strcmp(key, "bubble_Nbins")) p->bubble_nbins = atoi(val);
    else if (!strcmp(key, "use_SFR")) {
      if (!strcmp(val, "T")) p->use_sfr = 1;
      if (!strcmp(val, "F")) p->use_sfr = 0;
    }
    else if (!strcmp(key, "Vel_comp")) p->vel_comp = atoi(val);
    else if (!strcmp(key, "omega_matter")) p->omega_m = atof(val);
    else if (!strcmp(key, "omega_baryon")) p->omega_b = atof(val);
    else if (!strcmp(key, "omega_lambda")) p->omega_lambda = atof(val);
    else if (!strcmp(key, "Zminsfr")) p->zmin_sfr = atof(val);
    else if (!strcmp(key, "use_Lya_xrays")) {
      if (!strcmp(val, "T")) p->use_lya_xrays = 1;
      if (!strcmp(val, "F")) p->use_lya_xrays = 0;
    }
    else if (!strcmp(key, "Original_nldensity_box")) {
      if (!strcmp(val, "T")) p->save_original_nl = 1;
      if (!strcmp(val, "F")) p->save_original_nl = 0;
    }
  }
  fclose(f);
  p->box_len = p->box_len * p->hubble;
  p->bubble_rmax = p->bubble_rmax * p->hubble;
  if (p->bubble_rmax > p->box_len / 2.) p->bubble_rmax = p->box_len / 2.;
  p->dk = 2 * SF21_PI / p->box_len;
  p->dx_halo = p->n_halo ? p->box_len / p->n_halo : 0.0;
  p->dx = p->n ? p->box_len / p->n : 0.0;
  if (p->dz != 0.0) {
    p->zmin = p->zmax - p->dz * (double)ceil((p->zmax - p->zmin) / p->dz);
    /* Input_variables.c:137-139 */
    p->zmin_sfr = p->zmax - p->dz * (double)ceil((p->zmax - p->zmin_sfr) / p->dz);
    if (p->zmin_sfr < p->zmin) p->zmin_sfr = p->zmin;
  }
  set_msg(msg, msg_cap, "");
  return SF21_OK;
}

/* ---- background cosmology (auxiliary.c:224-303) ----
 * H(z) = h H0 sqrt(Om (1+z)^3 + OL) in 1/s, H0 = 3.24078e-18 h/s
 * (Input_variables.h:108); the growth factor is H(z)/H(0) I(z)/I(0) with
 * I(z) the midpoint sum of (1+z')/(H(z')/H(0))^3 over [z, z + 100) in steps
 * of 1e-3 (the loop variable accumulated in FP64 as the reference does);
 * dD/dt the central difference of step 1e-3 in z times dz/dt = -H(z)(1+z). */
#define SF21_H0 3.24078e-18

double sf21_hubble_z(const sf21_params *p, double z) {
  return p->hubble * SF21_H0 * sqrt(p->omega_m * (1. + z) * (1. + z) * (1. + z) + p->omega_lambda);
}

static double growth_integral(const sf21_params *p, double zz) {
  const double dz = 0.001, h0 = sf21_hubble_z(p, 0.);
  double s = 0., z, e;
  for (z = zz + dz / 2.; z < zz + 100.; z += dz) {
    e = sf21_hubble_z(p, z) / h0;
    s += (1. + z) / e / e / e;
  }
  return s * dz;
}

double sf21_growth(const sf21_params *p, double z) {
  return sf21_hubble_z(p, z) / sf21_hubble_z(p, 0.) * growth_integral(p, z) / growth_integral(p, 0.);
}

double sf21_dgrowth_dt(const sf21_params *p, double z) {
  const double dz = 1e-3;
  return (sf21_growth(p, z + dz / 2.) - sf21_growth(p, z - dz / 2.)) / dz *
         (-sf21_hubble_z(p, z) * (1. + z));
}

int sf21_ladder(double rmax, double dx, int nbins, double *radii, int cap, int *n_main) {
  double bfactor, R;
  int k = 0, m;
  if (!(dx > 0) || nbins <= 0 || cap < 0) return SF21_ERR_ARG;
  bfactor = pow(10.0, log10(rmax / dx) / nbins);
  if (!(bfactor > 1.0)) return SF21_ERR_ARG;
  R = rmax;
  while (R >= 2 * dx) {
    if (k >= cap) return SF21_ERR_ARG;
    if (radii) radii[k] = R;
    k++;
    R /= bfactor;
  }
  m = k;
  while (R >= dx) {
    if (k >= cap) return SF21_ERR_ARG;
    if (radii) radii[k] = R;
    k++;
    R /= bfactor;
  }
  if (n_main) *n_main = m;
  return k;
}

long sf21_ball_d2max(double R, double dx) {
  long m;
  if (!(R >= 0) || !(dx > 0)) return -1;
  m = (long)floor((R / dx) * (R / dx));
  while (m > 0 && !(sqrt((double)m) * dx <= R)) m--;
  while (sqrt((double)(m + 1)) * dx <= R) m++;
  return m;
}
