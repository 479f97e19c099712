// Kernels of the plain per-radius GPU path -- the reference's dataflow
// (get_HIIbubbles.c:216-479) one step per launch -- and of the FFTW-like
// 3D real/complex transforms (half layout n x n x (n/2+1), unnormalised,
// forward sign -1), which replace fftwf_plan_dft_{r2c,c2r}_3d + fftwf_execute
// (get_HIIbubbles.c:110-183, :290-291, :324-325, :372-373, :382).
#include "fft_core.h"
#include "kernels.h"
#include "cells.h"

namespace sf21 {

// ------------------------------------------------------------------------
// Real rows: an L-point real transform is an L/2-point complex transform of
// z[n] = x[2n] + i x[2n+1] plus a split step:
//   E = (Z[k] + conj Z[L/2-k])/2,  O = (Z[k] - conj Z[L/2-k])/(2i),
//   X[k] = E + exp(-2 pi i k/L) O,  k = 0..L/2
// and the inverse builds Z'[k] = (X[k] + conj X[L/2-k])
//   + i exp(+2 pi i k/L) (X[k] - conj X[L/2-k]),  whose L/2-point inverse
// is L*(x[2n] + i x[2n+1]).  Imaginary parts of X[0], X[L/2] are ignored.
// ------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(256) void k_r2c_rows(const float* __restrict__ in,
                                                  float2* __restrict__ out,
                                                  const float2* __restrict__ twh,
                                                  const float2* __restrict__ twf) {
  constexpr int H = L / 2, E = Plan<H>::E, TPL = H / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<H>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<H>::kFloats;
  float* sim = sre + LdsLine<H>::kFloats;
  const float2* __restrict__ rin = reinterpret_cast<const float2*>(in + row * L);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = rin[t + TPL * m];
  line_fft<H, -1>(v, sre, sim, t, twh);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int o = lds_pad(t + TPL * m);
    sre[o] = v[m].x;
    sim[o] = v[m].y;
  }
  lds_pin();
  __syncthreads();
  float2* __restrict__ rout = out + row * (H + 1);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int kk = t + TPL * m;
    const int ob = lds_pad(kk == 0 ? 0 : H - kk);
    const int oa = lds_pad(kk);
    const float zr = sre[oa], zi = sim[oa], cr = sre[ob], ci = -sim[ob];
    const float er = 0.5f * (zr + cr), ei = 0.5f * (zi + ci);
    const float dr = 0.5f * (zr - cr), di = 0.5f * (zi - ci);
    const float orr = di, oi = -dr;  // (d)/(i)
    const float2 w = twf[kk];
    rout[kk] = make_float2(er + (w.x * orr - w.y * oi), ei + (w.x * oi + w.y * orr));
  }
  if (t == 0) {
    const float zr = sre[0], zi = sim[0];
    rout[H] = make_float2(zr - zi, 0.0f);
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_c2r_rows(const float2* __restrict__ in,
                                                  float* __restrict__ out,
                                                  const float2* __restrict__ twh,
                                                  const float2* __restrict__ twf) {
  constexpr int H = L / 2, E = Plan<H>::E, TPL = H / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<H>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<H>::kFloats;
  float* sim = sre + LdsLine<H>::kFloats;
  const float2* __restrict__ rin = in + row * (H + 1);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int kk = t + TPL * m;
    float2 xa = rin[kk], xb = rin[H - kk];
    xb.y = -xb.y;
    if (kk == 0) {
      xa.y = 0.0f;
      xb.y = 0.0f;
    }
    const float sr = xa.x + xb.x, si = xa.y + xb.y;
    const float dr = xa.x - xb.x, di = xa.y - xb.y;
    const float2 w = twf[kk];  // exp(-2 pi i k/L); use its conjugate
    const float tr = w.x * dr + w.y * di, ti = w.x * di - w.y * dr;
    v[m] = make_float2(sr - ti, si + tr);
  }
  line_fft<H, +1>(v, sre, sim, t, twh);
  float2* __restrict__ rout = reinterpret_cast<float2*>(out + row * L);
#pragma unroll
  for (int m = 0; m < E; ++m) rout[t + TPL * m] = v[m];
}

template <int L, int SIGN>
__global__ __launch_bounds__(256) void k_c2c_rows(float2* __restrict__ data,
                                                  const float2* __restrict__ tw) {
  constexpr int E = Plan<L>::E, TPL = L / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<L>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = data[row * L + t + TPL * m];
  line_fft<L, SIGN>(v, sre, sim, t, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) data[row * L + t + TPL * m] = v[m];
}

// ------------------------------------------------------------------------
// Plain-path elementwise steps
// ------------------------------------------------------------------------

// get_HIIbubbles.c:216-225, :269-274, :277-288: cell fields and quick fill.
__global__ void k_cell_fields(long n3, const float* __restrict__ a, const float* __restrict__ b,
                              const float* __restrict__ c, CellConsts k, float* __restrict__ rec,
                              float* __restrict__ ion, float* __restrict__ fresid,
                              float* __restrict__ bubblef) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    float r, o, f;
    cell_fields(k, a, b, c, i, r, o, f);
    rec[i] = r;
    ion[i] = o;
    fresid[i] = f;
    float q = 0.0f;
    if (o > 0.0f) q = r > 0.0f ? fminf(o / r, 1.0f) : 1.0f;
    bubblef[i] = q;
  }
}

// get_HIIbubbles.c:305-322 on the half layout, both spectra in one pass.
__global__ void k_filter_half(int n, const float2* __restrict__ in, float2* __restrict__ out,
                              long nh, const float* __restrict__ w) {
  const int h = n / 2 + 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % h);
    const long r = i / h;
    const int jj = (int)(r % n), ii = (int)(r / n);
    const float wv = w[fold_sq(ii, n) + fold_sq(jj, n) + q * q];
    float2 a = in[i], b = in[i + nh];
    out[i] = make_float2(a.x * wv, a.y * wv);
    out[i + nh] = make_float2(b.x * wv, b.y * wv);
  }
}

// get_HIIbubbles.c:335-352: centre box of 0/1 and the any-centre flag.
__global__ void k_criterion(long n3, const float* __restrict__ rec, const float* __restrict__ ion,
                            float* __restrict__ bubble, int* __restrict__ flag) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float r = rec[i], o = ion[i];
    const bool c = (o > 0.0f) && (!(r > 0.0f) || o >= r);
    bubble[i] = c ? 1.0f : 0.0f;
    if (c) flag[0] = 1;
  }
}

// get_HIIbubbles.c:363-371: periodic ball of radius R (min image over the 8
// corner images), as 0/1; d2 is the host-computed integer bound of
// sqrt(d^2)*dx <= R.
__global__ void k_ball(int n, int d2, float* __restrict__ top) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    int a = (int)(i / ((long)n * n)), b = (int)((i / n) % n), c = (int)(i % n);
    a = a < n - a ? a : n - a;
    b = b < n - b ? b : n - b;
    c = c < n - c ? c : n - c;
    top[i] = ((long)a * a + (long)b * b + (long)c * c <= d2) ? 1.0f : 0.0f;
  }
}

// get_HIIbubbles.c:379-381
__global__ void k_cmul(long count, float2* __restrict__ a, const float2* __restrict__ b) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    a[i] = cmul(a[i], b[i]);
}

// get_HIIbubbles.c:390-393: count * dx^3/N^3 > dx^3/100.
__global__ void k_threshold(long n3, const float* __restrict__ count, float scale, double thr,
                            float* __restrict__ bubblef) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float v = count[i] * scale;
    if ((double)v > thr) bubblef[i] = 1.0f;
  }
}

// get_HIIbubbles.c:428-458 as a gather (race-free): a cell is painted iff a
// centre lies at an offset d with |d_a| <= reach and |d|^2 <= d2.
__global__ void k_stencil(int n, const float* __restrict__ centres, int d2, int reach,
                          float* __restrict__ bubblef) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i / ((long)n * n)), b = (int)((i / n) % n), c = (int)(i % n);
    bool hit = false;
    for (int u = -reach; u <= reach && !hit; ++u) {
      const int x = (a - u + n) % n;
      for (int v = -reach; v <= reach && !hit; ++v) {
        const int y = (b - v + n) % n;
        for (int w = -reach; w <= reach; ++w) {
          if (u * u + v * v + w * w > d2) continue;
          const int z = (c - w + n) % n;
          if (centres[((long)x * n + y) * n + z] != 0.0f) {
            hit = true;
            break;
          }
        }
      }
    }
    if (hit) bubblef[i] = 1.0f;
  }
}

// get_HIIbubbles.c:466-477: clamp and FP64 partial sums of 1 - x.
__global__ __launch_bounds__(256) void k_clamp_sum(long n3, float* __restrict__ bubblef,
                                                   const float* __restrict__ fresid,
                                                   double* __restrict__ partials) {
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float v = fminf(bubblef[i], 1.0f - fresid[i]);
    bubblef[i] = v;
    acc += 1.0 - (double)v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------
static unsigned ew_blocks(long count) {
  long b = (count + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

#define SF21_REAL_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(512)

void launch_r2c_rows(int n, const float* in, float2* out, const float2* tw_half,
                     const float2* twf, hipStream_t s);
void launch_c2r_rows(int n, const float2* in, float* out, const float2* tw_half,
                     const float2* twf, hipStream_t s);

void launch_r2c_rows(int n, const float* in, float2* out, const float2* tw_half,
                     const float2* twf, hipStream_t s) {
#define X(L)                                                                             \
  if (n == L) {                                                                          \
    constexpr int LPB = 256 / ((L / 2) / Plan<L / 2>::E);                                \
    hipLaunchKernelGGL(k_r2c_rows<L>, dim3((L * L) / LPB), dim3(256), 0, s, in, out, tw_half, twf); \
    return;                                                                              \
  }
  SF21_REAL_SIZES(X)
#undef X
}

void launch_c2r_rows(int n, const float2* in, float* out, const float2* tw_half,
                     const float2* twf, hipStream_t s) {
#define X(L)                                                                             \
  if (n == L) {                                                                          \
    constexpr int LPB = 256 / ((L / 2) / Plan<L / 2>::E);                                \
    hipLaunchKernelGGL(k_c2r_rows<L>, dim3((L * L) / LPB), dim3(256), 0, s, in, out, tw_half, twf); \
    return;                                                                              \
  }
  SF21_REAL_SIZES(X)
#undef X
}

void launch_c2c_rows(int n, int sign, float2* data, const float2* tw, hipStream_t s) {
#define X(L)                                                                             \
  if (n == L) {                                                                          \
    constexpr int LPB = 256 / (L / Plan<L>::E);                                          \
    if (sign < 0)                                                                        \
      hipLaunchKernelGGL((k_c2c_rows<L, -1>), dim3((L * L) / LPB), dim3(256), 0, s, data, tw); \
    else                                                                                 \
      hipLaunchKernelGGL((k_c2c_rows<L, +1>), dim3((L * L) / LPB), dim3(256), 0, s, data, tw); \
    return;                                                                              \
  }
  X(32) X(64) X(96) X(128) X(192) X(256) X(512)
#undef X
}

void launch_cell_fields(long n3, const float* a, const float* b, const float* c, CellConsts k,
                        float* rec, float* ion, float* fresid, float* bubblef, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_fields, dim3(ew_blocks(n3)), dim3(256), 0, s, n3, a, b, c, k, rec, ion,
                     fresid, bubblef);
}

void launch_filter_half(int n, const float2* in, float2* out, const float* wlut, hipStream_t s) {
  const long nh = (long)n * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_filter_half, dim3(ew_blocks(nh)), dim3(256), 0, s, n, in, out, nh, wlut);
}

void launch_criterion(long n3, const float* rec, const float* ion, float* bubble, int* flag,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_criterion, dim3(ew_blocks(n3)), dim3(256), 0, s, n3, rec, ion, bubble, flag);
}

void launch_ball(int n, int d2, float* top_hat, hipStream_t s) {
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_ball, dim3(ew_blocks(n3)), dim3(256), 0, s, n, d2, top_hat);
}

void launch_cmul(long count, float2* a, const float2* b, hipStream_t s) {
  hipLaunchKernelGGL(k_cmul, dim3(ew_blocks(count)), dim3(256), 0, s, count, a, b);
}

void launch_threshold(long n3, const float* count, float scale, double thr, float* bubblef,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_threshold, dim3(ew_blocks(n3)), dim3(256), 0, s, n3, count, scale, thr,
                     bubblef);
}

void launch_stencil(int n, const float* centres, int d2, int reach, float* bubblef,
                    hipStream_t s) {
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_stencil, dim3(ew_blocks(n3)), dim3(256), 0, s, n, centres, d2, reach,
                     bubblef);
}

int clamp_partials_count(long n3) { return (int)ew_blocks(n3); }

void launch_clamp_sum(long n3, float* bubblef, const float* fresid, double* partials,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_clamp_sum, dim3(ew_blocks(n3)), dim3(256), 0, s, n3, bubblef, fresid,
                     partials);
}

}  // namespace sf21

namespace sf21 {

// ------------------------------------------------------------------------
// Binned power spectrum of a half spectrum (tools/power3d.c:172-191
// normalisation): per block FP64 sums of |F|^2, |k| and counts per bin in
// LDS, written as partials and reduced in a fixed order.  full: the n^3
// spectrum of a complex transform of the real box (generic sides); only its
// half p <= n/2 is binned, the same modes as the half spectrum's.
// ------------------------------------------------------------------------
constexpr int kPkMaxBins = 64;

__global__ __launch_bounds__(256) void k_pk_partial(int n, const float2* __restrict__ spec,
                                                    const short* __restrict__ bin_of_m,
                                                    const float* __restrict__ kabs_of_m,
                                                    int nbins, double* __restrict__ part,
                                                    bool full) {
  __shared__ double s_p[kPkMaxBins], s_k[kPkMaxBins], s_c[kPkMaxBins];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_p[b] = s_k[b] = s_c[b] = 0.0;
  __syncthreads();
  const int h = full ? n : n / 2 + 1;
  const long nh = (long)n * n * h;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % h);
    if (q > n / 2) continue;
    const long r = i / h;
    const int m = fold_sq((int)(r / n), n) + fold_sq((int)(r % n), n) + q * q;
    const int b = bin_of_m[m];
    if (b < 0) continue;
    const float2 f = spec[i];
    atomicAdd(&s_p[b], (double)f.x * f.x + (double)f.y * f.y);
    atomicAdd(&s_k[b], (double)kabs_of_m[m]);
    atomicAdd(&s_c[b], 1.0);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    part[(3L * blockIdx.x) * nbins + b] = s_p[b];
    part[(3L * blockIdx.x + 1) * nbins + b] = s_k[b];
    part[(3L * blockIdx.x + 2) * nbins + b] = s_c[b];
  }
}

__global__ void k_pk_final(const double* __restrict__ part, int nblocks, int nbins,
                           double norm, double* __restrict__ out) {
  const int b = threadIdx.x;
  if (b >= nbins) return;
  double p = 0.0, k = 0.0, c = 0.0;
  for (int i = 0; i < nblocks; ++i) {
    p += part[(3L * i) * nbins + b];
    k += part[(3L * i + 1) * nbins + b];
    c += part[(3L * i + 2) * nbins + b];
  }
  out[b] = c > 0 ? p / c * norm : 0.0;
  out[nbins + b] = c > 0 ? k / c : 0.0;
  out[2 * nbins + b] = c;
}

int pk_partials_count(int n) {
  long nh = (long)n * n * (n / 2 + 1);
  long b = (nh + 255) / 256;
  return (int)(b > 2048 ? 2048 : b);
}

void launch_power_spectrum(int n, const float2* spec, const short* bin_of_m, const float* kabs_of_m,
                           int nbins, double norm, double* part, double* out, hipStream_t s,
                           bool full) {
  const int nb = pk_partials_count(n);
  hipLaunchKernelGGL(k_pk_partial, dim3(nb), dim3(256), 0, s, n, spec, bin_of_m, kabs_of_m, nbins,
                     part, full);
  hipLaunchKernelGGL(k_pk_final, dim3(1), dim3(kPkMaxBins), 0, s, part, nb, nbins, norm, out);
}

}  // namespace sf21
