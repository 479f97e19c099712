/* Type-generic body of the oracle's CPU FFT.  Included twice by fft_cpu.c with
 * T = double / float and SUF = _d / _f.  Test infrastructure only.
 *
 * Algorithm: self-sorting Stockham passes with radices 4, 2, 3, 5 (and a direct
 * DFT for any other prime factor).  Pass with radix R and span Ns (product of
 * the radices already applied), for butterfly j in [0, n/R):
 *     k = j mod Ns
 *     v_r = in[j + r*n/R] * exp(sign*2*pi*i*k*r/(Ns*R)),  r < R
 *     v   = DFT_R(v)
 *     out[(j/Ns)*Ns*R + k + r*Ns] = v_r
 * Input and output are in natural order.  Twiddles are computed in double and
 * rounded to T, so the float build behaves like a single-precision FFT library.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

typedef struct {
  int n;
  int nstage;
  int radix[32];
  int ns[32];
  T *tw[32];   /* per stage: (Ns*R) complex twiddles exp(-2*pi*i*q/(Ns*R)) */
  T *rtw;      /* real-transform twiddles exp(-2*pi*i*k/n), k < n/2+1 */
} FN(orc_plan);

static FN(orc_plan) *FN(plan_cache)[64];
static int FN(plan_count) = 0;

static FN(orc_plan) *FN(make_plan)(int n) {
  FN(orc_plan) *p;
  int i, s, m = n, ns = 1;
  #pragma omp critical(orc_plan_cache)
  {
    p = NULL;
    for (i = 0; i < FN(plan_count); i++)
      if (FN(plan_cache)[i]->n == n) { p = FN(plan_cache)[i]; break; }
    if (!p) {
      p = (FN(orc_plan) *)calloc(1, sizeof(*p));
      p->n = n;
      s = 0;
      while (m > 1) {
        int r;
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
        else if (m % 3 == 0) r = 3;
        else if (m % 5 == 0) r = 5;
        else { r = m; for (i = 7; i * i <= m; i += 2) if (m % i == 0) { r = i; break; } }
        p->radix[s] = r;
        p->ns[s] = ns;
        p->tw[s] = (T *)malloc(sizeof(T) * 2 * (size_t)ns * r);
        for (i = 0; i < ns * r; i++) {
          double a = -2.0 * M_PI * (double)i / (double)(ns * r);
          p->tw[s][2 * i] = (T)cos(a);
          p->tw[s][2 * i + 1] = (T)sin(a);
        }
        ns *= r; m /= r; s++;
      }
      p->nstage = s;
      p->rtw = (T *)malloc(sizeof(T) * 2 * (size_t)(n / 2 + 1));
      for (i = 0; i <= n / 2; i++) {
        double a = -2.0 * M_PI * (double)i / (double)n;
        p->rtw[2 * i] = (T)cos(a);
        p->rtw[2 * i + 1] = (T)sin(a);
      }
      FN(plan_cache)[FN(plan_count)++] = p;
    }
  }
  return p;
}

/* One complex transform of length p->n; x is input, y scratch; result in x. */
static void FN(fft_exec)(const FN(orc_plan) *p, T *x, T *y, int sign) {
  int n = p->n, s, j, r, q;
  T *in = x, *out = y, *tmp;
  for (s = 0; s < p->nstage; s++) {
    int R = p->radix[s], Ns = p->ns[s], nr = n / R;
    const T *tw = p->tw[s];
    T vr[64], vi[64], wr[64], wi[64];
    for (j = 0; j < nr; j++) {
      int k = j % Ns;
      int base = (j / Ns) * Ns * R + k;
      for (r = 0; r < R; r++) {
        T ar = in[2 * (j + r * nr)], ai = in[2 * (j + r * nr) + 1];
        /* exp(sign*2*pi*i*k*r/(Ns*R)); tw holds the forward (negative) angle */
        T cr = tw[2 * (k * r)];
        T ci = (sign < 0) ? tw[2 * (k * r) + 1] : -tw[2 * (k * r) + 1];
        vr[r] = ar * cr - ai * ci;
        vi[r] = ar * ci + ai * cr;
      }
      if (R == 2) {
        T t0r = vr[0] + vr[1], t0i = vi[0] + vi[1];
        T t1r = vr[0] - vr[1], t1i = vi[0] - vi[1];
        wr[0] = t0r; wi[0] = t0i; wr[1] = t1r; wi[1] = t1i;
      } else if (R == 4) {
        T s02r = vr[0] + vr[2], s02i = vi[0] + vi[2];
        T d02r = vr[0] - vr[2], d02i = vi[0] - vi[2];
        T s13r = vr[1] + vr[3], s13i = vi[1] + vi[3];
        T d13r = vr[1] - vr[3], d13i = vi[1] - vi[3];
        /* multiply d13 by sign*i: (a+ib)*(sign*i) = -sign*b + sign*i*a */
        T jr = (T)(-sign) * d13i, ji = (T)sign * d13r;
        wr[0] = s02r + s13r; wi[0] = s02i + s13i;
        wr[2] = s02r - s13r; wi[2] = s02i - s13i;
        wr[1] = d02r + jr;   wi[1] = d02i + ji;
        wr[3] = d02r - jr;   wi[3] = d02i - ji;
      } else {
        /* direct DFT of length R with twiddles exp(sign*2*pi*i*a*b/R) */
        for (q = 0; q < R; q++) {
          T accr = 0, acci = 0;
          for (r = 0; r < R; r++) {
            int e = (q * r) % R;
            double a = (double)sign * 2.0 * M_PI * (double)e / (double)R;
            T c = (T)cos(a), sn = (T)sin(a);
            accr += vr[r] * c - vi[r] * sn;
            acci += vr[r] * sn + vi[r] * c;
          }
          wr[q] = accr; wi[q] = acci;
        }
      }
      for (r = 0; r < R; r++) {
        out[2 * (base + r * Ns)] = wr[r];
        out[2 * (base + r * Ns) + 1] = wi[r];
      }
    }
    tmp = in; in = out; out = tmp;
  }
  if (in != x) memcpy(x, in, sizeof(T) * 2 * (size_t)n);
}

void FN(orc_fft1d)(int n, T *data, int sign) {
  FN(orc_plan) *p = FN(make_plan)(n);
  T *y = (T *)malloc(sizeof(T) * 2 * (size_t)n);
  FN(fft_exec)(p, data, y, sign);
  free(y);
}

/* Real forward transform of one row: x[0..n) -> X[0..n/2] (interleaved). */
static void FN(r2c_row)(const FN(orc_plan) *ph, const FN(orc_plan) *pf, const T *x, T *X, T *z, T *scr) {
  int m = pf->n / 2, k;
  for (k = 0; k < m; k++) { z[2 * k] = x[2 * k]; z[2 * k + 1] = x[2 * k + 1]; }
  FN(fft_exec)(ph, z, scr, -1);
  for (k = 0; k <= m; k++) {
    int a = k % m, b = (m - k) % m;
    T zr = z[2 * a], zi = z[2 * a + 1];
    T cr = z[2 * b], ci = -z[2 * b + 1];
    T er = (T)0.5 * (zr + cr), ei = (T)0.5 * (zi + ci);
    /* O = (Z - conj Z')/(2i) */
    T dr = (T)0.5 * (zr - cr), di = (T)0.5 * (zi - ci);
    T or_ = di, oi = -dr;
    T wr = pf->rtw[2 * k], wi = pf->rtw[2 * k + 1];
    X[2 * k] = er + (wr * or_ - wi * oi);
    X[2 * k + 1] = ei + (wr * oi + wi * or_);
  }
}

/* Real backward transform of one row: X[0..n/2] -> x[0..n), scaled by n. */
static void FN(c2r_row)(const FN(orc_plan) *ph, const FN(orc_plan) *pf, const T *X, T *x, T *z, T *scr) {
  int m = pf->n / 2, k;
  for (k = 0; k < m; k++) {
    T ar = X[2 * k], ai = X[2 * k + 1];
    T br = X[2 * (m - k)], bi = -X[2 * (m - k) + 1];
    if (k == 0) { ai = 0; bi = 0; }
    {
      T sr = ar + br, si = ai + bi;       /* X[k] + conj X[m-k] */
      T dr = ar - br, di = ai - bi;       /* X[k] - conj X[m-k] */
      /* w^{-k} = conj(rtw[k]) */
      T wr = pf->rtw[2 * k], wi = -pf->rtw[2 * k + 1];
      T tr = wr * dr - wi * di, ti = wr * di + wi * dr;
      /* i * t = -ti + i tr */
      z[2 * k] = sr - ti;
      z[2 * k + 1] = si + tr;
    }
  }
  FN(fft_exec)(ph, z, scr, +1);
  for (k = 0; k < m; k++) { x[2 * k] = z[2 * k]; x[2 * k + 1] = z[2 * k + 1]; }
}

/* In-place complex transforms along the two slow axes of an n x n x w array. */
static void FN(c2c_slow_axes)(int n, int w, T *a, int sign) {
  FN(orc_plan) *p = FN(make_plan)(n);
  /* along j (stride w), lines (i, c) */
  #pragma omp parallel
  {
    T *buf = (T *)malloc(sizeof(T) * 4 * (size_t)n);
    T *scr = buf + 2 * (size_t)n;
    long li;
    #pragma omp for schedule(static)
    for (li = 0; li < (long)n * w; li++) {
      long ii = li / w, c = li % w, j;
      T *base = a + 2 * ((size_t)ii * n * w + c);
      for (j = 0; j < n; j++) { buf[2 * j] = base[2 * j * w]; buf[2 * j + 1] = base[2 * j * w + 1]; }
      FN(fft_exec)(p, buf, scr, sign);
      for (j = 0; j < n; j++) { base[2 * j * w] = buf[2 * j]; base[2 * j * w + 1] = buf[2 * j + 1]; }
    }
    #pragma omp barrier
    /* along i (stride n*w), lines (j, c) */
    #pragma omp for schedule(static)
    for (li = 0; li < (long)n * w; li++) {
      long jj = li / w, c = li % w, ii2;
      T *base = a + 2 * ((size_t)jj * w + c);
      size_t st = (size_t)n * w;
      for (ii2 = 0; ii2 < n; ii2++) { buf[2 * ii2] = base[2 * ii2 * st]; buf[2 * ii2 + 1] = base[2 * ii2 * st + 1]; }
      FN(fft_exec)(p, buf, scr, sign);
      for (ii2 = 0; ii2 < n; ii2++) { base[2 * ii2 * st] = buf[2 * ii2]; base[2 * ii2 * st + 1] = buf[2 * ii2 + 1]; }
    }
    free(buf);
  }
}

void FN(orc_r2c3d)(int n, const T *in, T *out) {
  FN(orc_plan) *pf = FN(make_plan)(n);
  FN(orc_plan) *ph = FN(make_plan)(n / 2);
  int h = n / 2 + 1;
  #pragma omp parallel
  {
    T *z = (T *)malloc(sizeof(T) * 2 * (size_t)n);
    T *scr = z + n;
    long row;
    #pragma omp for schedule(static)
    for (row = 0; row < (long)n * n; row++)
      FN(r2c_row)(ph, pf, in + (size_t)row * n, out + 2 * (size_t)row * h, z, scr);
    free(z);
  }
  FN(c2c_slow_axes)(n, h, out, -1);
}

void FN(orc_c2r3d)(int n, const T *in, T *out) {
  FN(orc_plan) *pf = FN(make_plan)(n);
  FN(orc_plan) *ph = FN(make_plan)(n / 2);
  int h = n / 2 + 1;
  size_t nh = (size_t)n * n * h;
  T *tmp = (T *)malloc(sizeof(T) * 2 * nh);
  memcpy(tmp, in, sizeof(T) * 2 * nh);
  FN(c2c_slow_axes)(n, h, tmp, +1);
  #pragma omp parallel
  {
    T *z = (T *)malloc(sizeof(T) * 2 * (size_t)n);
    T *scr = z + n;
    long row;
    #pragma omp for schedule(static)
    for (row = 0; row < (long)n * n; row++)
      FN(c2r_row)(ph, pf, tmp + 2 * (size_t)row * h, out + (size_t)row * n, z, scr);
    free(z);
  }
  free(tmp);
}

void FN(orc_c2c3d)(int n, T *data, int sign) {
  FN(orc_plan) *p = FN(make_plan)(n);
  #pragma omp parallel
  {
    T *buf = (T *)malloc(sizeof(T) * 2 * (size_t)n);
    long row;
    #pragma omp for schedule(static)
    for (row = 0; row < (long)n * n; row++)
      FN(fft_exec)(p, data + 2 * (size_t)row * n, buf, sign);
    free(buf);
  }
  FN(c2c_slow_axes)(n, n, data, sign);
}

#undef FN
#undef CAT
#undef CAT_
