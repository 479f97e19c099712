/* CPU FFT used by the oracle (test infrastructure only; never linked into the
 * product library or CLI).
 *
 * It stands in for FFTW 3 single/double precision, which the reference links
 * (makefile:8-9) and calls in get_HIIbubbles.c:110-183 (plans) and :290-291,
 * :324-325, :372-373, :382, :424-425 (execute).  Semantics follow FFTW's
 * documented conventions:
 *   - unnormalised transforms, forward sign -1, backward sign +1;
 *   - r2c output is the "half" layout n x n x (n/2+1), last index fastest;
 *   - c2r takes the half layout and returns n^3 reals scaled by n^3;
 *     the imaginary parts of the self-conjugate rows (p = 0 and p = n/2) are
 *     ignored, as for any Hermitian-input c2r.
 * Sizes: any n whose prime factors are 2, 3 and 5 (other primes fall back to a
 * direct DFT stage).  Complex data is interleaved (re, im).
 */
#ifndef ORACLE_FFT_CPU_H
#define ORACLE_FFT_CPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* 1D complex transforms, in place, length n, sign -1 (forward) or +1. */
void orc_fft1d_d(int n, double *data, int sign);
void orc_fft1d_f(int n, float *data, int sign);

/* 3D real transforms, n^3 real <-> n*n*(n/2+1) complex (interleaved). */
void orc_r2c3d_d(int n, const double *in, double *out);
void orc_c2r3d_d(int n, const double *in, double *out); /* in is not modified */
void orc_r2c3d_f(int n, const float *in, float *out);
void orc_c2r3d_f(int n, const float *in, float *out);

/* 3D complex transform, n^3 complex interleaved, in place. */
void orc_c2c3d_d(int n, double *data, int sign);
void orc_c2c3d_f(int n, float *data, int sign);

#ifdef __cplusplus
}
#endif
#endif
