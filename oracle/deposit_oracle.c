/* Oracle: CPU restatement of the steps that produce the bubble finder's
 * inputs: the SFRD box (get_SFR.c:98-148) and the non-linear density box
 * (get_nldensity.c:95-166).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 * The deposits use orc_cic (auxiliary.c:761-808) and orc_smooth_box
 * (auxiliary.c:643-682); positions and masses are formed in FP64 and
 * rounded to float as the reference does. */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

/* get_SFR.c:113-134: CIC of sfr(M, z) = Rion/Qion (user_functions.c:71-74),
 * smooth_box, / dx_halo^3 */
void orc_sfrd(long nhalos, const float *halos, double z, long nh, long n, double dx_halo,
              float *sfrd, double *mean) {
  float *hm = (float *)calloc((size_t)nh * nh * nh, sizeof(float));
  double s = 0.0;
  long i;
  for (i = 0; i < nhalos; i++)
    orc_cic(halos[4 * i + 1], halos[4 * i + 2], halos[4 * i + 3],
            (float)(orc_Rion(halos[4 * i], z) / orc_Qion(z)), hm, nh);
  orc_smooth_box(hm, sfrd, nh, n);
  for (i = 0; i < n * n * n; i++) {
    sfrd[i] = (float)(sfrd[i] / pow(dx_halo, 3));
    s += sfrd[i];
  }
  if (mean) *mean = s / (double)(n * n * n);
  free(hm);
}

/* get_nldensity.c:95-99 (velocity to cells), :117-124 (mass), :140-157
 * (displace and deposit on -1), :161 (smooth_box) */
void orc_nldensity(long nh, long n, double L_h, double hubble, double growth_in, double growth,
                   const float *vx, const float *vy, const float *vz, const float *delta0,
                   float *box_halo, float *out) {
  const long n3h = nh * nh * nh;
  long i, j, p, ind;
  for (ind = 0; ind < n3h; ind++) box_halo[ind] = -1.0f;
  for (i = 0; i < nh; i++)
    for (j = 0; j < nh; j++)
      for (p = 0; p < nh; p++) {
        ind = (i * nh + j) * nh + p;
        const float sx = (float)(vx[ind] * ((nh / L_h) * hubble));
        const float sy = (float)(vy[ind] * ((nh / L_h) * hubble));
        const float sz = (float)(vz[ind] * ((nh / L_h) * hubble));
        const float mass = (float)(delta0[ind] * growth_in + 1.0);
        orc_cic((float)(i + sx * growth), (float)(j + sy * growth), (float)(p + sz * growth),
                mass, box_halo, nh);
      }
  if (out) orc_smooth_box(box_halo, out, nh, n);
}
