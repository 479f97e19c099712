/* get_SFR -- drop-in replacement for SimFast21's get_SFR.x: SFRD boxes from
 * the non-linear halo catalogs, deposited on an MI355X through libsf21hii.
 *
 * Contract kept from get_SFR.c:
 *   argv            `get_SFR base_dir`; other argc: three usage lines, exit 1
 *   parameters      simfast21.ini; with use_SFR = F and use_Lya_xrays = F it
 *                   prints that no SFR is needed and exits 0 (:51)
 *   redshifts       zmax down to zmin (Zminsim with use_SFR = T, else Zminsfr)
 *   inputs          Halos/halonl_z%.3f_N%ld_L%.1f.dat.catalog (N = N_halo)
 *   outputs         SFR/sfrd_z%.3f_N%ld_L%.1f.dat (N = N_smoothed); appends
 *                   "%f %.8E" (z, mean SFRD) to
 *                   Output_text_files/sfrd_av_N%ld_L%.1f.dat
 * Differences (deliberate): short catalogs are errors; CIC accumulates with
 * FP32 atomics (the reference: serial FP32 adds).
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <hip/hip_runtime_api.h>

#include "sf21_hii.h"

static void make_dir(const char *base, const char *sub, const char *msg) {
  char path[4096];
  DIR *d;
  snprintf(path, sizeof(path), "%s/%s", base, sub);
  if ((d = opendir(path)) == NULL) {
    printf("%s", msg);
    if (mkdir(path, (S_IRWXU | S_IRGRP | S_IXGRP | S_IROTH | S_IXOTH)) != 0) {
      printf("Error creating directory!\n");
      exit(1);
    }
  } else {
    closedir(d);
  }
}

int main(int argc, char *argv[]) {
  sf21_params p;
  char msg[512], fname[4096];
  FILE *fid, *fid_sa;
  double z, zmin, zmax, dz, lmpc, mean;
  long n3, nhalos;
  float *host, *d_halos = NULL, *d_out;
  long d_cap = 0;
  sf21_hii_ctx *ctx = NULL;
  int rc;

  if (argc != 2) {
    printf("Generates SFRD using nonlinear halo catalogue.\n");
    printf("usage: get_SFR base_dir\n");
    printf("base_dir contains simfast21.ini\n");
    exit(1);
  }
  rc = sf21_params_read(argv[1], &p, msg, sizeof(msg));
  if (rc != SF21_OK) {
    printf("%s", msg);
    exit(1);
  }
  if (p.use_lya_xrays == 0 && p.use_sfr == 0) {
    printf("SFR, Lya and xray use set to false - no need to calculate SFR\n");
    exit(0);
  }
  zmin = p.use_sfr ? p.zmin : p.zmin_sfr;
  zmax = p.zmax;
  dz = p.dz;
  lmpc = p.box_len / p.hubble;
  n3 = p.n * p.n * p.n;
  printf("\nCalculating SFRD between z=%f and z=%f with step %f\n", zmin, zmax, dz);
  printf("Using %d threads\n", p.nthreads); /* get_SFR.c:59-62 with -D_OMPTHREAD_ */
  make_dir(argv[1], "SFR", "Creating SFR directory\n");
  make_dir(argv[1], "Output_text_files", "Creating Output_text_files directory\n");
  snprintf(fname, sizeof(fname), "%s/Output_text_files/sfrd_av_N%ld_L%.1f.dat", argv[1], p.n, lmpc);
  if ((fid_sa = fopen(fname, "a")) == NULL) {
    printf("\nError opening output %s file...\n", fname);
    exit(1);
  }
  host = (float *)malloc(sizeof(float) * n3);
  if (!host || hipMalloc((void **)&d_out, sizeof(float) * n3) != hipSuccess) {
    printf("Problem...\n");
    exit(1);
  }
  {
    /* deposits only: any N_smoothed, no transform buffers */
    sf21_hii_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.n = (int)p.n;
    cfg.box_len = p.box_len;
    cfg.input_mode = SF21_INPUT_HALO;
    cfg.n_halo = (int)p.n_halo;
    cfg.path = SF21_PATH_DEPOSIT;
    cfg.hubble = p.hubble;
    cfg.fesc = p.fesc;
    if ((rc = sf21_hii_create(&cfg, &ctx)) != SF21_OK) {
      printf("GPU setup failed: %s\n", sf21_strerror(rc));
      exit(1);
    }
  }
  for (z = zmax; z > (zmin - dz / 10); z -= dz) {
    float *halos;
    snprintf(fname, sizeof(fname), "%s/Halos/halonl_z%.3f_N%ld_L%.1f.dat.catalog", argv[1], z,
             p.n_halo, lmpc);
    fid = fopen(fname, "rb");
    if (fid == NULL || fread(&nhalos, sizeof(long), 1, fid) != 1 || nhalos < 0) {
      printf("\nError reading %s file... Check path or if the file exists...", fname);
      exit(1);
    }
    printf("Redshift: %f. Reading %ld halos...\n", z, nhalos);
    fflush(0);
    halos = (float *)malloc(sizeof(float) * 4 * (size_t)(nhalos > 0 ? nhalos : 1));
    if (!halos || (nhalos > 0 && fread(halos, 4 * sizeof(float), (size_t)nhalos, fid) != (size_t)nhalos)) {
      printf("\nError reading %s file... Check path or if the file exists...", fname);
      exit(1);
    }
    fclose(fid);
    if (nhalos > d_cap) {
      if (d_halos) hipFree(d_halos);
      if (hipMalloc((void **)&d_halos, sizeof(float) * 4 * nhalos) != hipSuccess) {
        printf("Problem - halo...\n");
        exit(1);
      }
      d_cap = nhalos;
    }
    if (nhalos > 0 &&
        hipMemcpy(d_halos, halos, sizeof(float) * 4 * nhalos, hipMemcpyHostToDevice) != hipSuccess) {
      printf("Problem - halo...\n");
      exit(1);
    }
    free(halos);
    rc = sf21_sfrd_dev(ctx, d_halos, nhalos, z, d_out, &mean);
    if (rc != SF21_OK || hipMemcpy(host, d_out, sizeof(float) * n3, hipMemcpyDeviceToHost) != hipSuccess) {
      printf("SFRD failed: %s\n", sf21_strerror(rc != SF21_OK ? rc : SF21_ERR_HIP));
      exit(1);
    }
    printf("Writing SFRD file...\n");
    snprintf(fname, sizeof(fname), "%s/SFR/sfrd_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    fid = fopen(fname, "wb");
    if (fid == NULL || fwrite(host, sizeof(float), (size_t)n3, fid) != (size_t)n3) {
      printf("\nError opening %s file...\n", fname);
      exit(1);
    }
    fclose(fid);
    fprintf(fid_sa, "%f %.8E\n", z, mean);
  }
  fclose(fid_sa);
  sf21_hii_destroy(ctx);
  if (d_halos) hipFree(d_halos);
  hipFree(d_out);
  free(host);
  return 0;
}
