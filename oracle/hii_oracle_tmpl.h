/* Type-generic per-slice pipeline of the oracle (included by hii_oracle.c with
 * T = float / double).  TEST INFRASTRUCTURE ONLY.
 *
 * T = float reproduces the reference's typing: float fields and spectra, FP32
 * FFT, FP64 filter weights and FP64 ratio tests (get_HIIbubbles.c:56-57,
 * :284, :318-319, :341, :391-392, :438, :470).  T = double is the same
 * dataflow in FP64 throughout ("the FP64 run of the same pipeline").
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* FP64 filter weight for every integer |k|^2 = indi^2+indj^2+p^2: the
 * reference evaluates W_filter(dk*sqrt(m)*R) per element
 * (get_HIIbubbles.c:317-319); the value depends on m only. */
static double *FN(w_table)(int n, double dk, double R, long *mmax) {
  long m, mm = 3L * (n / 2) * (n / 2);
  double *w = (double *)malloc(sizeof(double) * (mm + 1));
  for (m = 0; m <= mm; m++) {
    double kk = dk * sqrt((double)m);
    w[m] = orc_W_filter(kk * R);
  }
  *mmax = mm;
  return w;
}

/* get_HIIbubbles.c:305-322 / :406-423: out = in * W(|k| R) on the half
 * layout; the index fold keeps i = n/2 positive (:309-311). */
static void FN(filter_half)(int n, const double *w, const T *in, T *out) {
  int h = n / 2 + 1;
  long i;
  #pragma omp parallel for schedule(static)
  for (i = 0; i < n; i++) {
    long indi = i > n / 2 ? -(n - i) : i, j, q;
    for (j = 0; j < n; j++) {
      long indj = j > n / 2 ? -(n - j) : j;
      for (q = 0; q < h; q++) {
        long m = indi * indi + indj * indj + q * q;
        size_t o = 2 * (((size_t)i * n + j) * h + q);
        out[o] = (T)((double)in[o] * w[m]);
        out[o + 1] = (T)((double)in[o + 1] * w[m]);
      }
    }
  }
}

static void FN(filtered_pair)(int n, double dk, double R, const float *rec, const float *ion,
                              T *rec_r, T *ion_r) {
  long n3 = (long)n * n * n, i, mm;
  size_t nh = (size_t)n * n * (n / 2 + 1);
  T *a = (T *)malloc(sizeof(T) * n3);
  T *ca = (T *)malloc(sizeof(T) * 2 * nh);
  T *cf = (T *)malloc(sizeof(T) * 2 * nh);
  double *w = FN(w_table)(n, dk, R, &mm);
  for (i = 0; i < n3; i++) a[i] = (T)rec[i];
  R2C(n, a, ca);
  FN(filter_half)(n, w, ca, cf);
  C2R(n, cf, rec_r);
  for (i = 0; i < n3; i++) a[i] = (T)ion[i];
  R2C(n, a, ca);
  FN(filter_half)(n, w, ca, cf);
  C2R(n, cf, ion_r);
  free(a); free(ca); free(cf); free(w);
}

/* Per-slice body of get_HIIbubbles.c: quick fill :277-288, forward FFTs
 * :290-291, main radius loop :298-398, small radius loop :400-462, clamp
 * :466-471 and the FP64 mean :473-477. */
static int FN(hii_slice)(const orc_hii_params *p, const float *rec, const float *ion,
                         const float *fresid, float *xhii, double *neutral, int *first_centre) {
  int n = p->n, h = n / 2 + 1, r;
  long n3 = (long)n * n * n, i;
  size_t nh = (size_t)n * n * h;
  double dx = p->dx, acc;
  T *bubblef = (T *)malloc(sizeof(T) * n3);
  T *rec_t = (T *)malloc(sizeof(T) * n3);
  T *ion_t = (T *)malloc(sizeof(T) * n3);
  T *rec_c = (T *)malloc(sizeof(T) * 2 * nh);
  T *ion_c = (T *)malloc(sizeof(T) * 2 * nh);
  T *tmp_c = (T *)malloc(sizeof(T) * 2 * nh);
  T *tmp2_c = (T *)malloc(sizeof(T) * 2 * nh);
  T *rec_r = (T *)malloc(sizeof(T) * n3);
  T *ion_r = (T *)malloc(sizeof(T) * n3);
  T *bubble = (T *)malloc(sizeof(T) * n3);
  unsigned char *centre = (unsigned char *)malloc((size_t)n3);
  double *dist = p->dilation == 1 ? (double *)malloc(sizeof(double) * n3) : NULL;

  if (!bubblef || !rec_t || !ion_t || !rec_c || !ion_c || !tmp_c || !tmp2_c || !rec_r ||
      !ion_r || !bubble || !centre || (p->dilation == 1 && !dist))
    return -2;
  if (first_centre) for (i = 0; i < n3; i++) first_centre[i] = -1;

  /* quick fill, :277-288 */
  for (i = 0; i < n3; i++) {
    double t;
    if (ion[i] > 0.) t = rec[i] > 0. ? (double)ion[i] / rec[i] : 1.0;
    else t = 0.;
    bubblef[i] = t >= 1.0 ? (T)1.0 : (T)t;
    rec_t[i] = (T)rec[i];
    ion_t[i] = (T)ion[i];
  }
  R2C(n, rec_t, rec_c);
  R2C(n, ion_t, ion_c);

  for (r = 0; r < p->nr; r++) {
    double R = p->radii[r];
    long mm, any = 0;
    double *w = FN(w_table)(n, p->dk, R, &mm);
    FN(filter_half)(n, w, rec_c, tmp_c);
    C2R(n, tmp_c, rec_r);
    FN(filter_half)(n, w, ion_c, tmp_c);
    C2R(n, tmp_c, ion_r);
    free(w);

    /* criterion, :335-352 (main) and :437-438 (small) */
    for (i = 0; i < n3; i++) {
      int c = 0;
      if (ion_r[i] > 0.) {
        if (rec_r[i] > 0.) c = ((double)ion_r[i] / rec_r[i] >= 1.0);
        else c = 1;
      }
      centre[i] = (unsigned char)c;
      any |= c;
      if (c && first_centre && first_centre[i] < 0) first_centre[i] = r;
    }
    if (!any) continue;

    if (p->dilation == 1) {
      long d2 = orc_d2max(R, dx);
      edt3d_periodic(n, centre, dist);
      for (i = 0; i < n3; i++)
        if (dist[i] <= (double)d2) bubblef[i] = (T)1.0;
    } else if (R >= 2 * dx) {
      /* main loop: ball window :363-371, two r2c :372-373, product
       * :379-381, c2r :382, scale and threshold :390-393 */
      long a;
      double dx3 = pow(dx, 3);
      #pragma omp parallel for schedule(static)
      for (a = 0; a < n; a++) {
        long b, c, img;
        for (b = 0; b < n; b++)
          for (c = 0; c < n; c++) {
            int in = 0;
            for (img = 0; img < 8 && !in; img++) {
              long u = (img & 4) ? a - n : a, v = (img & 2) ? b - n : b, s = (img & 1) ? c - n : c;
              if (sqrt((double)(u * u + v * v + s * s)) * dx <= R) in = 1;
            }
            rec_r[(a * n + b) * n + c] = in ? (T)1.0 : (T)0.0;   /* top_hat_r */
          }
      }
      for (i = 0; i < n3; i++) bubble[i] = centre[i] ? (T)1.0 : (T)0.0;
      R2C(n, rec_r, tmp_c);     /* top_hat_c */
      R2C(n, bubble, tmp2_c);   /* bubble_c */
      for (i = 0; i < (long)nh; i++) {
        T ar = tmp2_c[2 * i], ai = tmp2_c[2 * i + 1], br = tmp_c[2 * i], bi = tmp_c[2 * i + 1];
        tmp2_c[2 * i] = ar * br - ai * bi;
        tmp2_c[2 * i + 1] = ar * bi + ai * br;
      }
      C2R(n, tmp2_c, bubble);
      for (i = 0; i < n3; i++) {
        bubble[i] = (T)(bubble[i] * dx3 / n3);
        if (bubble[i] > dx3 / 100) bubblef[i] = (T)1.0;
      }
    } else {
      /* small loop: stencil paint :428-458 with periodic wrap */
      long nc = (long)(R / dx), c0;
      for (c0 = 0; c0 < n3; c0++) {
        long ci, cj, ck, u, v, s;
        if (!centre[c0]) continue;
        ci = c0 / ((long)n * n); cj = (c0 / n) % n; ck = c0 % n;
        for (u = -(nc + 1); u <= nc + 1; u++) {
          long a = orc_check_borders(ci + u, n);
          for (v = -(nc + 1); v <= nc + 1; v++) {
            long b;
            if (!(sqrt((double)(u * u + v * v)) * dx <= R)) continue;
            b = orc_check_borders(cj + v, n);
            for (s = -(nc + 1); s <= nc + 1; s++) {
              long c = orc_check_borders(ck + s, n);
              if (sqrt((double)(u * u + v * v + s * s)) * dx <= R)
                bubblef[(a * n + b) * n + c] = (T)1.0;
            }
          }
        }
      }
    }
  }

  /* clamp :466-471 and mean :473-477 */
  acc = 0.0;
  for (i = 0; i < n3; i++) {
    double fr = fresid ? (double)fresid[i] : 0.0;
    if (bubblef[i] > 1.0 - fr) bubblef[i] = (T)(1.0 - fr);
    acc += 1.0 - bubblef[i];
    xhii[i] = (float)bubblef[i];
  }
  *neutral = acc / n3;

  free(bubblef); free(rec_t); free(ion_t); free(rec_c); free(ion_c); free(tmp_c);
  free(tmp2_c); free(rec_r); free(ion_r); free(bubble); free(centre); free(dist);
  return 0;
}

#undef FN
#undef CAT
#undef CAT_
