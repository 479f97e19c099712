/* get_HIIbubbles -- drop-in replacement for SimFast21's get_HIIbubbles.x
 * running the bubble finder on an MI355X through libsf21hii.
 *
 * Contract kept from get_HIIbubbles.c:
 *   argv            `get_HIIbubbles base_dir`; any other argc prints the
 *                   three usage lines and exits 1 (:63-68)
 *   parameters      base_dir/simfast21.ini (Input_variables.c:16-141)
 *   directories     creates Ionization/ and Output_text_files/ (:80-96)
 *   inputs          delta/deltanl_z%.3f_N%ld_L%.1f.dat and either
 *                   SFR/sfrd_z%.3f_N%ld_L%.1f.dat (use_SFR = T) or
 *                   Halos/halonl_z%.3f_N%ld_L%.1f.dat.catalog (use_SFR = F,
 *                   N = N_halo) per redshift, z ascending from the snapped zmin
 *                   by dz (:206-213, :236-246, :262-267)
 *   outputs         Ionization/xHII_z%.3f_N%ld_L%.1f.dat for every z, all-zero
 *                   boxes above the cutoff (:480-513); appends zsim.txt and
 *                   x_HI_N%ld_L%.1f.dat (:515-528)
 *   stdout          the reference's progress lines and "neutral fraction=%lf"
 *   exit codes      0 on success, 1 on any error with a message
 * Differences (deliberate):
 *   - a short input file is an error here; the reference does not check
 *     fread and would run on stale buffer contents (:212, :266);
 *   - a short halo catalog is an error as well;
 *   - the bubble finder runs on the GPU (fused exact path); nthreads is read
 *     but only sizes the host I/O.
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

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

static int read_box(const char *path, float *buf, long n3) {
  FILE *f = fopen(path, "rb");
  size_t got;
  if (!f) return SF21_ERR_IO;
  got = fread(buf, sizeof(float), (size_t)n3, f);
  fclose(f);
  return got == (size_t)n3 ? SF21_OK : SF21_ERR_IO;
}

/* Halo catalog: long count, then count x {float Mass, x, y, z}
 * (output_files.txt:10-14, auxiliary.h:13-17). */
static float *read_catalog(const char *path, long *count) {
  FILE *f = fopen(path, "rb");
  float *h;
  long n = 0;
  if (!f) return NULL;
  if (fread(&n, sizeof(long), 1, f) != 1 || n < 0) {
    fclose(f);
    return NULL;
  }
  h = (float *)malloc(sizeof(float) * 4 * (size_t)(n > 0 ? n : 1));
  if (!h || (n > 0 && fread(h, 4 * sizeof(float), (size_t)n, f) != (size_t)n)) {
    free(h);
    fclose(f);
    return NULL;
  }
  fclose(f);
  *count = n;
  return h;
}

static void write_box(const char *path, const float *buf, long n3) {
  FILE *f = fopen(path, "wb");
  if (!f) {
    printf("Error opening file:%s\n", path);
    exit(1);
  }
  if (fwrite(buf, sizeof(float), (size_t)n3, f) != (size_t)n3) {
    printf("Error writing file:%s\n", path);
    exit(1);
  }
  fclose(f);
}

int main(int argc, char *argv[]) {
  sf21_params p;
  char msg[512], fname[4096];
  double *radii, *xhi, z, neutral, bfactor, lmpc;
  float *delta, *sfrd, *xhii, *halos = NULL;
  long nhalos = 0;
  long n3, i;
  int nr, nmain, cap, iz, rc;
  sf21_hii_ctx *ctx = NULL;
  FILE *fid;

  if (argc != 2) {
    printf("Generates boxes with ionization fraction for a range of redshifts\n");
    printf("usage: get_HIIbubbles base_dir\n");
    printf("base_dir contains simfast21.ini and directory structure\n");
    exit(1);
  }
  rc = sf21_params_read(argv[1], &p, msg, sizeof(msg));
  if (rc != SF21_OK) {
    printf("%s", msg);
    exit(1);
  }
  /* the reference is built with -D_OMPTHREAD_ (makefile:14) and reports its
   * thread count here (get_HIIbubbles.c:74-79) */
  printf("Using %d threads\n", p.nthreads);
  fflush(0);
  make_dir(argv[1], "Ionization", "Creating Ionization directory\n");
  make_dir(argv[1], "Output_text_files", "Creating Output_text_files directory\n");

  n3 = p.n * p.n * p.n;
  lmpc = p.box_len / p.hubble;
  delta = (float *)malloc(sizeof(float) * n3);
  sfrd = (float *)malloc(sizeof(float) * n3);
  xhii = (float *)malloc(sizeof(float) * n3);
  xhi = (double *)malloc(sizeof(double) * ((int)((p.zmax - p.zmin) / p.dz + 2)));
  cap = 64 * (p.bubble_nbins > 0 ? p.bubble_nbins : 1) + 64;
  radii = (double *)malloc(sizeof(double) * cap);
  if (!delta || !sfrd || !xhii || !xhi || !radii) {
    printf("Memory Problem...\n");
    exit(1);
  }
  bfactor = pow(10.0, log10(p.bubble_rmax / p.dx) / p.bubble_nbins);
  printf("Bubble radius ratio (bfactor): %f\n", bfactor);
  printf("Number of bubble steps: %d\n",
         (int)((log10(p.bubble_rmax) - log10(p.dx)) / log10(bfactor)));
  printf("Redshift cycle...\n");
  fflush(0);
  nr = sf21_ladder(p.bubble_rmax, p.dx, p.bubble_nbins, radii, cap, &nmain);
  if (nr < 1) {
    printf("Invalid bubble radius ladder (bubble_Rmax=%f, bubble_Nbins=%d)\n", p.bubble_rmax,
           p.bubble_nbins);
    exit(1);
  }

  iz = 0;
  neutral = 0.;
  for (z = p.zmin; z < (p.zmax + p.dz / 10) && (neutral < p.ion_cutoff); z += p.dz) {
    printf("z = %f\n", z);
    fflush(0);
    snprintf(fname, sizeof(fname), "%s/delta/deltanl_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    if (read_box(fname, delta, n3) != SF21_OK) {
      printf("\nError reading deltanl file... Check path or if the file exists...");
      exit(1);
    }
    if (p.use_sfr) {
      snprintf(fname, sizeof(fname), "%s/SFR/sfrd_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
      if (read_box(fname, sfrd, n3) != SF21_OK) {
        printf("\nError reading %s file... Check path or if the file exists...", fname);
        exit(1);
      }
    } else {
      snprintf(fname, sizeof(fname), "%s/Halos/halonl_z%.3f_N%ld_L%.1f.dat.catalog", argv[1], z,
               p.n_halo, lmpc);
      free(halos);
      halos = read_catalog(fname, &nhalos);
      if (!halos) {
        printf("\nError reading %s file... Check path or if the file exists...", fname);
        exit(1);
      }
      printf("Reading %ld halos...\n", nhalos);
      fflush(0);
    }
    if (!ctx) {
      sf21_hii_config cfg;
      memset(&cfg, 0, sizeof(cfg));
      cfg.n = (int)p.n;
      cfg.box_len = p.box_len;
      cfg.n_radii = nr;
      cfg.radii = radii;
      cfg.input_mode = p.use_sfr ? SF21_INPUT_SFR : SF21_INPUT_HALO;
      cfg.n_halo = (int)p.n_halo;
      cfg.path = SF21_PATH_FUSED;
      cfg.hubble = p.hubble;
      cfg.fesc = p.fesc;
      rc = sf21_hii_create(&cfg, &ctx);
      if (rc != SF21_OK) {
        printf("GPU setup failed: %s\n", sf21_strerror(rc));
        exit(1);
      }
    }
    printf("Going to smaller R cycle...\n");
    fflush(0);
    if (p.use_sfr)
      rc = sf21_hii_run_slice(ctx, delta, sfrd, NULL, z, xhii, &neutral);
    else
      rc = sf21_hii_run_slice_halo(ctx, delta, halos, nhalos, z, xhii, &neutral);
    if (rc != SF21_OK) {
      printf("Bubble finder failed: %s\n", sf21_strerror(rc));
      exit(1);
    }
    printf("neutral fraction=%lf\n", neutral);
    fflush(0);
    xhi[iz] = neutral;
    snprintf(fname, sizeof(fname), "%s/Ionization/xHII_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    write_box(fname, xhii, n3);
    iz++;
  }

  printf("Writing high redshift xHII boxes...\n");
  for (i = 0; i < n3; i++) xhii[i] = 0.0f;
  while (z < (p.zmax + p.dz / 10)) {
    xhi[iz] = 1.0;
    snprintf(fname, sizeof(fname), "%s/Ionization/xHII_z%.3f_N%ld_L%.1f.dat", argv[1], z, p.n, lmpc);
    write_box(fname, xhii, n3);
    iz++;
    z += p.dz;
  }

  snprintf(fname, sizeof(fname), "%s/Output_text_files/zsim.txt", argv[1]);
  if ((fid = fopen(fname, "a")) == NULL) {
    printf("Error opening file:%s\n", fname);
    exit(1);
  }
  for (z = p.zmax; z > (p.zmin - p.dz / 10); z -= p.dz) fprintf(fid, "%f\n", z);
  fclose(fid);
  snprintf(fname, sizeof(fname), "%s/Output_text_files/x_HI_N%ld_L%.1f.dat", argv[1], p.n, lmpc);
  if ((fid = fopen(fname, "a")) == NULL) {
    printf("Error opening file:%s\n", fname);
    exit(1);
  }
  for (i = iz - 1; i >= 0; i--) fprintf(fid, "%lf\n", xhi[i]);
  fclose(fid);

  if (ctx) sf21_hii_destroy(ctx);
  free(delta);
  free(sfrd);
  free(xhii);
  free(xhi);
  free(radii);
  free(halos);
  return 0;
}
