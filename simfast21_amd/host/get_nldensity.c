/* get_nldensity -- drop-in replacement for SimFast21's get_nldensity.x:
 * non-linear density boxes by displacing the z = 0 density with the linear
 * velocity field, deposited on an MI355X through libsf21hii.
 *
 * Contract kept from get_nldensity.c:
 *   argv            `get_nldensity base_dir`; other argc: three usage lines,
 *                   exit 1
 *   inputs          Velocity/vel_{x,y,z}_z0_N%ld_L%.1f.dat and
 *                   delta/delta_z0_N%ld_L%.1f.dat (N = N_halo)
 *   outputs         per z from zmax down: delta/deltanl_z%.3f_N%ld_L%.1f.dat
 *                   (N = N_smoothed; existing files are skipped) and, with
 *                   Original_nldensity_box = T, deltanl_o_z..._N<N_halo>
 * Differences (deliberate): short input files are errors; positions are
 * formed and the CIC accumulated in FP32 with atomics (the reference: FP64
 * positions rounded to float, serial FP32 adds).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "sf21_hii.h"

static int read_box(const char *path, float *buf, long n) {
  FILE *f = fopen(path, "rb");
  size_t got;
  if (!f) return SF21_ERR_IO;
  got = fread(buf, sizeof(float), (size_t)n, f);
  fclose(f);
  return got == (size_t)n ? SF21_OK : SF21_ERR_IO;
}

static float *upload(const float *host, long n) {
  float *d = NULL;
  if (hipMalloc((void **)&d, sizeof(float) * n) != hipSuccess ||
      hipMemcpy(d, host, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) {
    printf("Problem...\n");
    exit(1);
  }
  return d;
}

static void write_box(const char *path, const float *buf, long n) {
  FILE *f = fopen(path, "wb");
  if (f == NULL || fwrite(buf, sizeof(float), (size_t)n, f) != (size_t)n) {
    printf("\nError opening output nl_density file... Check path...\n");
    exit(1);
  }
  fclose(f);
}

int main(int argc, char *argv[]) {
  sf21_params p;
  char msg[512], fname[4096];
  FILE *fid;
  double z, zmin, zmax, dz, lmpc;
  long n3, n3h;
  float *host, *d_v[3], *d_delta, *d_box, *d_out;
  sf21_hii_ctx *ctx = NULL;
  static const char *axis[3] = {"x", "y", "z"}, *axis_u[3] = {"X", "Y", "Z"};
  int rc, k;

  if (argc != 2) {
    printf("Generates non-linear density boxes for a range of redshifts\n");
    printf("usage: get_nldensity base_dir\n");
    printf("base_dir contains simfast21.ini and directory structure\n");
    exit(1);
  }
  rc = sf21_params_read(argv[1], &p, msg, sizeof(msg));
  if (rc != SF21_OK) {
    printf("%s", msg);
    exit(1);
  }
  zmin = p.zmin;
  zmax = p.zmax;
  dz = p.dz;
  lmpc = p.box_len / p.hubble;
  n3 = p.n * p.n * p.n;
  n3h = p.n_halo * p.n_halo * p.n_halo;
  printf("Using %d threads\n", p.nthreads); /* get_nldensity.c:51-54 with -D_OMPTHREAD_ */
  printf("\nzmin=%E zmax=%E  dz=%E\n", zmin, zmax, dz);
  printf("Reading velocity boxes...\n");
  fflush(0);
  host = (float *)malloc(sizeof(float) * (n3h > n3 ? n3h : n3));
  if (!host) {
    printf("Problem...\n");
    exit(1);
  }
  for (k = 0; k < 3; k++) {
    snprintf(fname, sizeof(fname), "%s/Velocity/vel_%s_z0_N%ld_L%.1f.dat", argv[1], axis[k], p.n_halo, lmpc);
    if (read_box(fname, host, n3h) != SF21_OK) {
      printf("\nError reading %s velocity file... Check path or if the file exists...", axis_u[k]);
      exit(1);
    }
    d_v[k] = upload(host, n3h);
  }
  printf("Velocity reading done...\n");
  fflush(0);
  snprintf(fname, sizeof(fname), "%s/delta/delta_z0_N%ld_L%.1f.dat", argv[1], p.n_halo, lmpc);
  if (read_box(fname, host, n3h) != SF21_OK) {
    printf("\nError reading density file... Check if the file exists...\n");
    exit(1);
  }
  d_delta = upload(host, n3h);
  if (hipMalloc((void **)&d_box, sizeof(float) * n3h) != hipSuccess ||
      hipMalloc((void **)&d_out, sizeof(float) * n3) != hipSuccess) {
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
    if ((rc = sf21_hii_create(&cfg, &ctx)) != SF21_OK) {
      printf("GPU setup failed: %s\n", sf21_strerror(rc));
      exit(1);
    }
  }
  for (z = zmax; z > (zmin - dz / 10.); z -= dz) {
    printf("Generating delta_nl for redshift: %lf\n", z);
    fflush(0);
    snprintf(fname, sizeof(fname), "%s/delta/deltanl_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if ((fid = fopen(fname, "rb")) != NULL) {
      printf("File %s already exists - skipping...\n", fname);
      fclose(fid);
      continue;
    }
    rc = sf21_nldensity_dev(ctx, &p, z, d_v[0], d_v[1], d_v[2], d_delta, d_box, d_out);
    if (rc != SF21_OK || hipMemcpy(host, d_out, sizeof(float) * n3, hipMemcpyDeviceToHost) != hipSuccess) {
      printf("Non-linear density failed: %s\n", sf21_strerror(rc != SF21_OK ? rc : SF21_ERR_HIP));
      exit(1);
    }
    write_box(fname, host, n3);
    if (p.save_original_nl == 1) {
      if (hipMemcpy(host, d_box, sizeof(float) * n3h, hipMemcpyDeviceToHost) != hipSuccess) {
        printf("Problem...\n");
        exit(1);
      }
      snprintf(fname, sizeof(fname), "%s/delta/deltanl_o_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n_halo, lmpc);
      write_box(fname, host, n3h);
    }
  }
  sf21_hii_destroy(ctx);
  for (k = 0; k < 3; k++) hipFree(d_v[k]);
  hipFree(d_delta);
  hipFree(d_box);
  hipFree(d_out);
  free(host);
  return 0;
}
