/* Oracle CLI: CPU restatement of the get_HIIbubbles.x program (argv, ini
 * parsing, directory creation, redshift loop, file reads and writes, stdout
 * summary, exit codes).  TEST INFRASTRUCTURE ONLY: the tests run it beside
 * the product CLI on the same base_dir and compare the files.
 *
 *   get_HIIbubbles.c:40-96    argv check, directories
 *   get_HIIbubbles.c:192-195  ladder summary lines
 *   get_HIIbubbles.c:204-488  redshift loop
 *   get_HIIbubbles.c:494-528  high-z zero boxes, text tables
 *   Input_variables.c:16-141  ini parsing and derived quantities
 *
 * Environment (oracle-only knobs): ORC_PRECISION=32|64 (default 32, the
 * reference's typing), ORC_DILATION=0|1 (default 0, the reference's FFT
 * convolution; 1 = exact ball dilation).
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include "oracle.h"

#define ORC_PI 3.1415926535897932385E0   /* Input_variables.h:10 */

typedef struct {
  int nthreads, use_sfr, nbins;
  long n_halo, n;
  double dz, zmax, zmin, L, hubble, fesc, xhlim, rmax;
  /* derived (Input_variables.c:119-136) */
  double dx_halo, dx, dk;
} orc_ini;

/* Input_variables.c:16-141 restated for the keys this step reads.  Lines
 * starting with '#' or '\n' are skipped, every other line needs '=', and
 * the first two whitespace tokens after replacing '=' by ' ' are key and
 * value (so trailing comments are ignored); unknown keys are ignored. */
static void read_ini(const char *basedir, orc_ini *p) {
  char line[256], fname[300], key[100], val[100], *eq;
  FILE *f;
  memset(p, 0, sizeof(*p));
  sprintf(fname, "%s/simfast21.ini", basedir);
  if ((f = fopen(fname, "r")) == NULL) {
    printf("\nThe parameter file simfast21.ini cannot be open. Exit...\n");
    exit(1);
  }
  while (fgets(line, 256, f) != NULL) {
    int nt;
    if (line[0] == '#' || line[0] == '\n') continue;
    if ((eq = strchr(line, '=')) == NULL) { printf("Wrong format! Exiting...\n"); printf("%s\n", line); exit(1); }
    *eq = ' ';
    nt = sscanf(line, "%s %s", key, val);
    if (nt != 2 || strlen(key) == 0 || strlen(val) == 0) { printf("Wrong format! Exiting...\n"); printf("%s\n", line); exit(1); }
    if (!strcmp(key, "nthreads")) p->nthreads = atoi(val);
    else if (!strcmp(key, "dz")) p->dz = atof(val);
    else if (!strcmp(key, "zmax")) p->zmax = atof(val);
    else if (!strcmp(key, "zmin")) p->zmin = atof(val);
    else if (!strcmp(key, "sim_length")) p->L = atof(val);
    else if (!strcmp(key, "N_halo")) p->n_halo = atoi(val);
    else if (!strcmp(key, "N_smoothed")) p->n = atoi(val);
    else if (!strcmp(key, "hubble")) p->hubble = atof(val);
    else if (!strcmp(key, "fesc")) p->fesc = atof(val);
    else if (!strcmp(key, "Ion_cutoff")) p->xhlim = atof(val);
    else if (!strcmp(key, "bubble_Rmax")) p->rmax = atof(val);
    else if (!strcmp(key, "bubble_Nbins")) p->nbins = atoi(val);
    else if (!strcmp(key, "use_SFR")) { if (!strcmp(val, "T")) p->use_sfr = 1; if (!strcmp(val, "F")) p->use_sfr = 0; }
  }
  fclose(f);
  p->L = p->L * p->hubble;                       /* Mpc/h, :121 */
  p->rmax = p->rmax * p->hubble;                 /* :123 */
  if (p->rmax > p->L / 2.) p->rmax = p->L / 2.;  /* :124 */
  p->dk = 2 * ORC_PI / p->L;                     /* :127 */
  p->dx_halo = p->L / p->n_halo;                 /* :128 */
  p->dx = p->L / p->n;                           /* :129 */
  p->zmin = p->zmax - p->dz * (double)ceil((p->zmax - p->zmin) / p->dz);  /* :133 */
}

static void ensure_dir(const char *basedir, const char *sub, const char *msg) {
  char fname[300];
  DIR *d;
  sprintf(fname, "%s/%s", basedir, sub);
  if ((d = opendir(fname)) == NULL) {
    printf("%s", msg);
    if (mkdir(fname, (S_IRWXU | S_IRGRP | S_IXGRP | S_IROTH | S_IXOTH)) != 0) {
      printf("Error creating directory!\n");
      exit(1);
    }
  } else closedir(d);
}

int main(int argc, char *argv[]) {
  orc_ini p;
  char fname[300];
  FILE *fid;
  long n3, i;
  double z, neutral, bfactor, *xhi, *radii;
  float *delta, *src, *fresid, *rec, *ion, *xhii;
  int iz, nr, nmain, cap, prec, dil;
  orc_hii_params hp;
  const char *e;

  if (argc != 2) {
    printf("Generates boxes with ionization fraction for a range of redshifts\n");
    printf("usage: get_HIIbubbles base_dir\n");
    printf("base_dir contains simfast21.ini and directory structure\n");
    exit(1);
  }
  read_ini(argv[1], &p);
  e = getenv("ORC_PRECISION"); prec = e ? atoi(e) : 32;
  e = getenv("ORC_DILATION"); dil = e ? atoi(e) : 0;
  ensure_dir(argv[1], "Ionization", "Creating Ionization directory\n");
  ensure_dir(argv[1], "Output_text_files", "Creating Output_text_files directory\n");

  n3 = p.n * p.n * p.n;
  delta = (float *)calloc(n3, sizeof(float));
  src = (float *)calloc(n3, sizeof(float));
  fresid = (float *)calloc(n3, sizeof(float));
  rec = (float *)calloc(n3, sizeof(float));
  ion = (float *)calloc(n3, sizeof(float));
  xhii = (float *)calloc(n3, sizeof(float));
  xhi = (double *)malloc(((int)((p.zmax - p.zmin) / p.dz + 2)) * sizeof(double));
  cap = 4 * p.nbins + 64;
  radii = (double *)malloc(sizeof(double) * cap);
  bfactor = pow(10.0, log10(p.rmax / p.dx) / p.nbins);
  printf("Bubble radius ratio (bfactor): %f\n", bfactor);
  printf("Number of bubble steps: %d\n", (int)((log10(p.rmax) - log10(p.dx)) / log10(bfactor)));
  printf("Redshift cycle...\n");
  fflush(0);
  nr = orc_ladder(p.rmax, p.dx, p.nbins, radii, cap, &nmain);
  if (nr < 0) { printf("Ladder too long\n"); exit(1); }
  hp.n = (int)p.n; hp.dx = p.dx; hp.dk = p.dk; hp.nr = nr; hp.radii = radii;
  hp.precision = prec; hp.dilation = dil;

  iz = 0;
  neutral = 0.;
  for (z = p.zmin; z < (p.zmax + p.dz / 10) && (neutral < p.xhlim); z += p.dz) {
    printf("z = %f\n", z); fflush(0);
    sprintf(fname, "%s/delta/deltanl_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, p.L / p.hubble);
    fid = fopen(fname, "rb");
    if (fid == NULL) { printf("\nError reading deltanl file... Check path or if the file exists..."); exit(1); }
    if (fread(delta, sizeof(float), n3, fid)) {}
    fclose(fid);
    if (p.use_sfr == 0) {
      long nh;
      float *halos;
      sprintf(fname, "%s/Halos/halonl_z%.3f_N%ld_L%.1f.dat.catalog", argv[1], z, p.n_halo, p.L / p.hubble);
      fid = fopen(fname, "rb");
      if (fid == NULL) { printf("\nError reading %s file... Check path or if the file exists...", fname); exit(1); }
      if (fread(&nh, sizeof(long), 1, fid)) {}
      printf("Reading %ld halos...\n", nh); fflush(0);
      halos = (float *)malloc(sizeof(float) * 4 * (nh > 0 ? nh : 1));
      if (fread(halos, 4 * sizeof(float), nh, fid)) {}
      fclose(fid);
      orc_halo_emissivity(nh, halos, z, p.n_halo, p.n, p.dx_halo, p.dx, p.fesc, src);
      free(halos);
      {
        /* fresid and rec from the pre-pass; emissivity from the halos */
        float *dummy = (float *)calloc(n3, sizeof(float));
        orc_prepass_sfr(n3, delta, dummy, z, p.fesc, p.dx, p.L, p.n, p.hubble, fresid, rec, dummy);
        free(dummy);
      }
      memcpy(ion, src, sizeof(float) * n3);
    } else {
      sprintf(fname, "%s/SFR/sfrd_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, p.L / p.hubble);
      fid = fopen(fname, "rb");
      if (fid == NULL) { printf("\nError reading %s file... Check path or if the file exists...", fname); exit(1); }
      if (fread(src, sizeof(float), n3, fid)) {}
      fclose(fid);
      orc_prepass_sfr(n3, delta, src, z, p.fesc, p.dx, p.L, p.n, p.hubble, fresid, rec, ion);
    }
    printf("Going to smaller R cycle...\n"); fflush(0);
    if (orc_hii_slice(&hp, rec, ion, fresid, xhii, &neutral, NULL) != 0) { printf("Memory Problem...\n"); exit(1); }
    printf("neutral fraction=%lf\n", neutral); fflush(0);
    xhi[iz] = neutral;
    sprintf(fname, "%s/Ionization/xHII_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, p.L / p.hubble);
    if ((fid = fopen(fname, "wb")) == NULL) { printf("Error opening file:%s\n", fname); exit(1); }
    fwrite(xhii, sizeof(float), n3, fid);
    fclose(fid);
    iz++;
  }
  printf("Writing high redshift xHII boxes...\n");
  for (i = 0; i < n3; i++) xhii[i] = 0.0f;
  while (z < (p.zmax + p.dz / 10)) {
    xhi[iz] = 1.0;
    sprintf(fname, "%s/Ionization/xHII_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, p.L / p.hubble);
    if ((fid = fopen(fname, "wb")) == NULL) { printf("Error opening file:%s\n", fname); exit(1); }
    fwrite(xhii, sizeof(float), n3, fid);
    fclose(fid);
    iz++;
    z += p.dz;
  }
  sprintf(fname, "%s/Output_text_files/zsim.txt", argv[1]);
  if ((fid = fopen(fname, "a")) == NULL) { printf("Error opening file:%s\n", fname); exit(1); }
  for (z = p.zmax; z > (p.zmin - p.dz / 10); z -= p.dz) fprintf(fid, "%f\n", z);
  fclose(fid);
  sprintf(fname, "%s/Output_text_files/x_HI_N%ld_L%.1f.dat", argv[1], p.n, p.L / p.hubble);
  if ((fid = fopen(fname, "a")) == NULL) { printf("Error opening file:%s\n", fname); exit(1); }
  for (i = iz - 1; i >= 0; i--) fprintf(fid, "%lf\n", xhi[i]);
  fclose(fid);
  exit(0);
}
