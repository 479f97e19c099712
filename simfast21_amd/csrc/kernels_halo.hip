// Halo-catalog emissivity (use_SFR = F, get_HIIbubbles.c:229-259): cloud-in-
// cell deposit of each halo's ionising rate Rion(M, z) (user_functions.c:40-47)
// on the N_halo^3 grid (auxiliary.c:761-808), then the block average down to
// N_smoothed^3 (smooth_box, auxiliary.c:643-682) times (dx_s/dx_h)^3 fesc.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace sf21 {

// Cloud-in-cell deposit of v at (x1, y1, z1) in grid-cell units
// (auxiliary.c:761-808): per axis truncate toward zero and, for a negative
// remainder, step one cell down and use |remainder| as the far weight
// (:764-784); the 8 cells wrap periodically.  FP32 atomics.
__device__ __forceinline__ void cic_deposit(float x1, float y1, float z1, float v, int ng,
                                            float* __restrict__ grid) {
  int x = (int)x1, y = (int)y1, z = (int)z1;
  float fx = x1 - (float)x, fy = y1 - (float)y, fz = z1 - (float)z;
  if (fx < 0.0f) { x -= 1; fx = -fx; }
  if (fy < 0.0f) { y -= 1; fy = -fy; }
  if (fz < 0.0f) { z -= 1; fz = -fz; }
  auto wrap = [ng](int q) { q %= ng; return q < 0 ? q + ng : q; };
  for (int a = 0; a < 2; ++a) {
    const int xi = wrap(x + a);
    const float wx = a ? fx : 1.0f - fx;
    for (int b = 0; b < 2; ++b) {
      const int yi = wrap(y + b);
      const float wy = b ? fy : 1.0f - fy;
      for (int c = 0; c < 2; ++c) {
        const int zi = wrap(z + c);
        const float wz = c ? fz : 1.0f - fz;
        atomicAdd(&grid[((long)xi * ng + yi) * ng + zi], wx * wy * wz * v);
      }
    }
  }
}

// One thread per halo.  Positions are in N_halo cell units; the deposited
// value is Rion(M, z) (user_functions.c:40-47) times `per_rion` (1 for the
// emissivity, 1/Qion(z) for the SFR of user_functions.c:71-74).
__global__ void k_cic_rion(long nhalos, const float* __restrict__ halos, int ng, float k_rion,
                           float c2, float c3, float* __restrict__ grid) {
  for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nhalos;
       h += (long)gridDim.x * blockDim.x) {
    const float m = halos[4 * h], x1 = halos[4 * h + 1], y1 = halos[4 * h + 2],
                z1 = halos[4 * h + 3];
    // Rion = c1 M (1+z)^c4 (M/c2)^c3 exp(-(c2/M)^3) / V_norm, k_rion = c1 (1+z)^c4 / V_norm
    const float q = c2 / m;
    const float rion = k_rion * m * powf(m / c2, c3) * expf(-(q * q * q));
    cic_deposit(x1, y1, z1, rion, ng, grid);
  }
}

// Non-linear density (get_nldensity.c:140-166): every cell (i, j, p) of the
// N_halo^3 grid moves its mass 1 + delta0 D_in to (i, j, p) + v s (D(z) -
// D_in) and deposits it by CIC on a grid that starts at -1.  v: the z = 0
// velocity boxes (Mpc), s = N_halo h / L_h turns them into cells.
__global__ void k_fill_value(float* __restrict__ p, float value, long count) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void k_displace_deposit(int ng, const float* __restrict__ vx,
                                   const float* __restrict__ vy, const float* __restrict__ vz,
                                   const float* __restrict__ delta0, float vscale,
                                   float growth_in, float growth, float* __restrict__ grid) {
  const long g3 = (long)ng * ng * ng;
  for (long ind = (long)blockIdx.x * blockDim.x + threadIdx.x; ind < g3;
       ind += (long)gridDim.x * blockDim.x) {
    const int i = (int)(ind / ((long)ng * ng)), j = (int)((ind / ng) % ng), p = (int)(ind % ng);
    const float mass = fmaf(delta0[ind], growth_in, 1.0f);
    const float x1 = fmaf(vx[ind] * vscale, growth, (float)i);
    const float y1 = fmaf(vy[ind] * vscale, growth, (float)j);
    const float z1 = fmaf(vz[ind] * vscale, growth, (float)p);
    cic_deposit(x1, y1, z1, mass, ng, grid);
  }
}

// Block average of sm^3 cells (compensated FP32 sum), then the scale.
__global__ void k_smooth_scale(int ng, int ns, int sm, const float* __restrict__ grid,
                               float scale, float* __restrict__ out) {
  const long n3 = (long)ns * ns * ns;
  const float inv = 1.0f / ((float)sm * sm * sm);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int a = (int)(i / ((long)ns * ns)), b = (int)((i / ns) % ns), c = (int)(i % ns);
    float s = 0.0f, comp = 0.0f;
    for (int u = 0; u < sm; ++u)
      for (int v = 0; v < sm; ++v) {
        const float* __restrict__ row = grid + ((long)(a * sm + u) * ng + (b * sm + v)) * ng + c * sm;
        for (int w = 0; w < sm; ++w) {
          const float yv = row[w] - comp;
          const float t = s + yv;
          comp = (t - s) - yv;
          s = t;
        }
      }
    out[i] = (s * inv) * scale;
  }
}

__global__ void k_zero_f32(float* __restrict__ p, long count) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

static unsigned blocks_for(long count) {
  long b = (count + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// FP64 sum of a float box: block partials, then one block.
__global__ __launch_bounds__(256) void k_sum_f32(long count, const float* __restrict__ x,
                                                 double* __restrict__ part) {
  __shared__ double r[256];
  double a = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    a += (double)x[i];
  r[threadIdx.x] = a;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) r[threadIdx.x] += r[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

__global__ __launch_bounds__(256) void k_sum_final(const double* __restrict__ part, int np,
                                                   double* __restrict__ out) {
  __shared__ double r[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) a += part[i];
  r[threadIdx.x] = a;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) r[threadIdx.x] += r[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = r[0];
}

int sum_partials_count(long count) { return (int)blocks_for(count); }

void launch_sum_f32(long count, const float* x, double* part, double* out, hipStream_t s) {
  const unsigned nb = blocks_for(count);
  hipLaunchKernelGGL(k_sum_f32, dim3(nb), dim3(256), 0, s, count, x, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, part, (int)nb, out);
}

void launch_nldensity(int ng, int ns, int sm, const float* vx, const float* vy, const float* vz,
                      const float* delta0, float vscale, float growth_in, float growth,
                      float* grid, float* out, hipStream_t s) {
  const long g3 = (long)ng * ng * ng;
  hipLaunchKernelGGL(k_fill_value, dim3(blocks_for(g3)), dim3(256), 0, s, grid, -1.0f, g3);
  hipLaunchKernelGGL(k_displace_deposit, dim3(blocks_for(g3)), dim3(256), 0, s, ng, vx, vy, vz,
                     delta0, vscale, growth_in, growth, grid);
  if (out)
    hipLaunchKernelGGL(k_smooth_scale, dim3(blocks_for((long)ns * ns * ns)), dim3(256), 0, s, ng,
                       ns, sm, grid, 1.0f, out);
}

void launch_halo_emissivity(long nhalos, const float* halos, int ng, int ns, int sm, float k_rion,
                            float c2, float c3, float scale, float* grid, float* out,
                            hipStream_t s) {
  const long g3 = (long)ng * ng * ng;
  hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(g3)), dim3(256), 0, s, grid, g3);
  if (nhalos > 0)
    hipLaunchKernelGGL(k_cic_rion, dim3(blocks_for(nhalos)), dim3(256), 0, s, nhalos, halos, ng,
                       k_rion, c2, c3, grid);
  hipLaunchKernelGGL(k_smooth_scale, dim3(blocks_for((long)ns * ns * ns)), dim3(256), 0, s, ng, ns,
                     sm, grid, scale, out);
}

}  // namespace sf21
