// 21-cm brightness temperature (t21.c:116-252): the velocity gradient of
// the density field along one axis, (k_axis/k)^2 delta_k transformed back,
// and one per-cell pass for the spin factor, the velocity and density
// factors and the neutral fraction.  FP32 arithmetic (the reference is FP64;
// the FP64 is in the per-redshift scalars, computed on the host, and in the
// means).
#include "fft_core.h"
#include "kernels.h"

namespace sf21 {

static unsigned blocks_t21(long count) {
  long b = (count + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Half spectrum of delta times (ind_axis^2 / |ind|^2), k = 0 set to zero
// (t21.c:166-194; the dk factors cancel).  axis: 1 = i, 2 = j, other = p.
__global__ void k_vgrad_filter(int n, float2* __restrict__ spec, int axis) {
  const int h = n / 2 + 1;
  const long nh = (long)n * n * h;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % h);
    const long r = i / h;
    const int a2 = fold_sq((int)(r / n), n), j2 = fold_sq((int)(r % n), n), q2 = q * q;
    const int m = a2 + j2 + q2;
    const int num = axis == 1 ? a2 : (axis == 2 ? j2 : q2);
    const float f = m == 0 ? 0.0f : (float)num / (float)m;
    const float2 v = spec[i];
    spec[i] = make_float2(v.x * f, v.y * f);
  }
}

// Full-spectrum form for box sides without a register plan (complex
// transform of the real box): the same factor with the p index folded like
// i and j.  The spectrum stays Hermitian, so the inverse is real.
__global__ void k_vgrad_filter_full(int n, float2* __restrict__ spec, int axis) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int p = (int)(i % n);
    const long r = i / n;
    const int a2 = fold_sq((int)(r / n), n), j2 = fold_sq((int)(r % n), n), q2 = fold_sq(p, n);
    const int m = a2 + j2 + q2;
    const int num = axis == 1 ? a2 : (axis == 2 ? j2 : q2);
    const float f = m == 0 ? 0.0f : (float)num / (float)m;
    const float2 v = spec[i];
    spec[i] = make_float2(v.x * f, v.y * f);
  }
}

__global__ void k_real_part(long n3, const float2* __restrict__ in, float* __restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x)
    out[i] = in[i].x;
}

// Per cell: spin factor x/(1+x) (1 - Tcmb/TK), x = x_alpha + x_c (t21.c:121-154;
// 1 without the spin boxes), dv = clamp(vfac * dv_raw, -0.5, 1) (:196-206),
// t21 = c21 (1 + delta) spin/(1 + dv) (:208-212) times (1 - xHII) (:215-223).
// FP64 partial sums of t21 and of TS = Tcmb/(1 - spin), the latter written
// Tcmb (1 + x)/(1 + x Tcmb/TK) to avoid the cancellation.
__global__ __launch_bounds__(256) void k_t21_cells(long n3, const float* __restrict__ delta,
                                                   const float* __restrict__ xhii,
                                                   const float* __restrict__ dv_raw,
                                                   const float* __restrict__ xc,
                                                   const float* __restrict__ xa,
                                                   const float* __restrict__ tk, float vfac,
                                                   float c21, float tcmb,
                                                   float* __restrict__ out,
                                                   double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  double st = 0.0, sts = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    float spin = 1.0f;
    if (xc) {
      const float x = xa[i] + xc[i];
      const float ratio = tcmb / tk[i];
      spin = x / (1.0f + x) * (1.0f - ratio);
      sts += (double)(tcmb * (1.0f + x) / (1.0f + x * ratio));
    }
    float dv = vfac * dv_raw[i];
    dv = dv > 1.0f ? 1.0f : (dv < -0.5f ? -0.5f : dv);
    const float t = c21 * (1.0f + delta[i]) * spin / (1.0f + dv) * (1.0f - xhii[i]);
    out[i] = t;
    st += (double)t;
  }
  r1[threadIdx.x] = st;
  r2[threadIdx.x] = sts;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r1[threadIdx.x] += r1[threadIdx.x + s];
      r2[threadIdx.x] += r2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

__global__ __launch_bounds__(256) void k_t21_final(const double* __restrict__ part, int np,
                                                   double* __restrict__ out) {
  __shared__ double r1[256], r2[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r1[threadIdx.x] += r1[threadIdx.x + s];
      r2[threadIdx.x] += r2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r1[0];
    out[1] = r2[0];
  }
}

int t21_partials_count(long n3) { return (int)blocks_t21(n3); }

void launch_vgrad_filter(int n, float2* spec, int axis, hipStream_t s) {
  const long nh = (long)n * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_vgrad_filter, dim3(blocks_t21(nh)), dim3(256), 0, s, n, spec, axis);
}

void launch_vgrad_filter_full(int n, float2* spec, int axis, hipStream_t s) {
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_vgrad_filter_full, dim3(blocks_t21(n3)), dim3(256), 0, s, n, spec, axis);
}

void launch_real_part(long n3, const float2* in, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_real_part, dim3(blocks_t21(n3)), dim3(256), 0, s, n3, in, out);
}

void launch_t21_cells(long n3, const float* delta, const float* xhii, const float* dv_raw,
                      const float* xc, const float* xa, const float* tk, float vfac, float c21,
                      float tcmb, float* out, double* part, double* sums, hipStream_t s) {
  const unsigned nb = blocks_t21(n3);
  hipLaunchKernelGGL(k_t21_cells, dim3(nb), dim3(256), 0, s, n3, delta, xhii, dv_raw, xc, xa, tk,
                     vfac, c21, tcmb, out, part);
  hipLaunchKernelGGL(k_t21_final, dim3(1), dim3(256), 0, s, part, (int)nb, sums);
}

}  // namespace sf21
