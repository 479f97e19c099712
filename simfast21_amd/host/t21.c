/* t21 -- drop-in replacement for SimFast21's t21.x computing the 21-cm
 * brightness temperature boxes on an MI355X through libsf21hii.
 *
 * Contract kept from t21.c:
 *   argv            `t21 base_dir`; any other argc prints "Usage : t21
 *                   base_dir" and exits 1 (:49-52)
 *   parameters      base_dir/simfast21.ini (Input_variables.c:16-141)
 *   directories     creates deltaTb/ (:67-74)
 *   inputs per z    delta/deltanl, Ionization/xHII and, for z > Zminsfr with
 *                   use_Lya_xrays = T, x_c/xc, Lya/xalpha, xrays/TempX
 *                   (_z%.3f_N%ld_L%.1f.dat), z from zmax down by dz (:116-160)
 *   outputs         deltaTb/deltaTb_z%.3f_N%ld_L%.1f.dat (skipped when it
 *                   exists, :118-122); appends "%f %.8E" lines to
 *                   Output_text_files/t21_av_N%ld_L%.1f.dat and, in the spin
 *                   branch, TS_av_N%ld_L%.1f.dat (:241-244)
 *   exit codes      0 on success, 1 on any error with a message
 * Differences (deliberate): short input files are errors (the reference
 * does not check fread); the arithmetic runs on the GPU in FP32 (the
 * reference: FP64), with FP64 means.
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <hip/hip_runtime_api.h>

#include "sf21_hii.h"

static int read_box(const char *path, float *buf, long n3) {
  FILE *f = fopen(path, "rb");
  size_t got;
  if (!f) return SF21_ERR_IO;
  got = fread(buf, sizeof(float), (size_t)n3, f);
  fclose(f);
  return got == (size_t)n3 ? SF21_OK : SF21_ERR_IO;
}

static void upload(float *dev, const float *host, long n3) {
  if (hipMemcpy(dev, host, sizeof(float) * n3, hipMemcpyHostToDevice) != hipSuccess) {
    printf("GPU copy failed\n");
    exit(1);
  }
}

int main(int argc, char *argv[]) {
  sf21_params p;
  char msg[512], fname[4096];
  FILE *fid, *fidtxt, *fidtxt2;
  DIR *dir;
  double z, lmpc, mean_t21, mean_ts;
  long n3;
  float *host, *d_delta, *d_xhii, *d_xc, *d_xa, *d_tk, *d_out;
  sf21_hii_ctx *ctx = NULL;
  int rc;

  if (argc != 2) {
    printf("Usage : t21 base_dir\n");
    exit(1);
  }
  rc = sf21_params_read(argv[1], &p, msg, sizeof(msg));
  if (rc != SF21_OK) {
    printf("%s", msg);
    exit(1);
  }
  if (p.use_lya_xrays == 0)
    printf("Lya and xray use set to false - assuming TS>>TCMB in t21 calculation.\n");
  printf("Using %d threads\n", p.nthreads); /* t21.c:60-65 with -D_OMPTHREAD_ */
  lmpc = p.box_len / p.hubble;
  n3 = p.n * p.n * p.n;

  snprintf(fname, sizeof(fname), "%s/deltaTb", argv[1]);
  if ((dir = opendir(fname)) == NULL) {
    printf("Creating t21 directory\n");
    if (mkdir(fname, (S_IRWXU | S_IRGRP | S_IXGRP | S_IROTH | S_IXOTH)) != 0) {
      printf("Error creating directory!\n");
      exit(1);
    }
  } else {
    closedir(dir);
  }
  snprintf(fname, sizeof(fname), "%s/Output_text_files/t21_av_N%ld_L%.1f.dat", argv[1], p.n, lmpc);
  if ((fidtxt = fopen(fname, "a")) == NULL) {
    printf("\nError opening output %s file...\n", fname);
    exit(1);
  }
  snprintf(fname, sizeof(fname), "%s/Output_text_files/TS_av_N%ld_L%.1f.dat", argv[1], p.n, lmpc);
  if ((fidtxt2 = fopen(fname, "a")) == NULL) {
    printf("\nError opening output %s file...\n", fname);
    exit(1);
  }
  host = (float *)malloc(sizeof(float) * n3);
  if (!host) {
    printf("Problem...\n");
    exit(1);
  }
  {
    /* FFT tables and scratch of a context at this N (one nominal radius) */
    sf21_hii_config cfg;
    double r0 = p.dx;
    memset(&cfg, 0, sizeof(cfg));
    cfg.n = (int)p.n;
    cfg.box_len = p.box_len;
    cfg.n_radii = 1;
    cfg.radii = &r0;
    cfg.input_mode = SF21_INPUT_DIRECT;
    cfg.path = SF21_PATH_FUSED;
    cfg.radius_batch = 1;
    rc = sf21_hii_create(&cfg, &ctx);
    if (rc != SF21_OK || p.n < 64) {
      printf("GPU setup failed: %s\n", sf21_strerror(rc != SF21_OK ? rc : SF21_ERR_ARG));
      exit(1);
    }
  }
  if (hipMalloc((void **)&d_delta, sizeof(float) * n3) != hipSuccess ||
      hipMalloc((void **)&d_xhii, sizeof(float) * n3) != hipSuccess ||
      hipMalloc((void **)&d_xc, sizeof(float) * n3) != hipSuccess ||
      hipMalloc((void **)&d_xa, sizeof(float) * n3) != hipSuccess ||
      hipMalloc((void **)&d_tk, sizeof(float) * n3) != hipSuccess ||
      hipMalloc((void **)&d_out, sizeof(float) * n3) != hipSuccess) {
    printf("Problem...\n");
    exit(1);
  }

  for (z = p.zmax; z > (p.zmin - p.dz / 10); z -= p.dz) {
    int spin;
    printf("T21 - z = %f\n", z);
    fflush(0);
    snprintf(fname, sizeof(fname), "%s/deltaTb/deltaTb_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if ((fid = fopen(fname, "rb")) != NULL) {
      printf("File:%s already exists - skipping this redshift...\n", fname);
      fclose(fid);
      continue;
    }
    spin = z > (p.zmin_sfr - p.dz / 10) && p.use_lya_xrays == 1;
    if (spin) {
      static const char *sub[3] = {"x_c/xc", "Lya/xalpha", "xrays/TempX"};
      float *dst[3] = {d_xc, d_xa, d_tk};
      for (int k = 0; k < 3; k++) {
        snprintf(fname, sizeof(fname), "%s/%s_z%.3f_N%ld_L%.1f.dat", argv[1], sub[k], z, p.n, lmpc);
        if (read_box(fname, host, n3) != SF21_OK) {
          printf("Error opening file:%s\n", fname);
          exit(1);
        }
        upload(dst[k], host, n3);
      }
    }
    snprintf(fname, sizeof(fname), "%s/delta/deltanl_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if (read_box(fname, host, n3) != SF21_OK) {
      printf("Error reading deltanl file... Check path or if the file exists...");
      exit(1);
    }
    upload(d_delta, host, n3);
    snprintf(fname, sizeof(fname), "%s/Ionization/xHII_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if (read_box(fname, host, n3) != SF21_OK) {
      printf("Error opening file:%s\n", fname);
      exit(1);
    }
    upload(d_xhii, host, n3);
    rc = sf21_t21_dev(ctx, &p, z, d_delta, d_xhii, spin ? d_xc : NULL, spin ? d_xa : NULL,
                      spin ? d_tk : NULL, d_out, &mean_t21, &mean_ts);
    if (rc != SF21_OK ||
        hipMemcpy(host, d_out, sizeof(float) * n3, hipMemcpyDeviceToHost) != hipSuccess) {
      printf("t21 failed: %s\n", sf21_strerror(rc != SF21_OK ? rc : SF21_ERR_HIP));
      exit(1);
    }
    snprintf(fname, sizeof(fname), "%s/deltaTb/deltaTb_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if ((fid = fopen(fname, "wb")) == NULL) {
      printf("Cannot open file:%s...\n", fname);
      exit(1);
    }
    if (fwrite(host, sizeof(float), (size_t)n3, fid) != (size_t)n3) {
      printf("Cannot open file:%s...\n", fname);
      exit(1);
    }
    fclose(fid);
    fprintf(fidtxt, "%f %.8E\n", z, mean_t21);
    if (spin) fprintf(fidtxt2, "%f %.8E\n", z, mean_ts);
  }
  fclose(fidtxt);
  fclose(fidtxt2);
  sf21_hii_destroy(ctx);
  hipFree(d_delta);
  hipFree(d_xhii);
  hipFree(d_xc);
  hipFree(d_xa);
  hipFree(d_tk);
  hipFree(d_out);
  free(host);
  return 0;
}
