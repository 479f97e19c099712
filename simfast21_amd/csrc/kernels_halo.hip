// Halo-catalog emissivity (use_SFR = F, get_HIIbubbles.c:229-259): cloud-in-
// cell deposit of each halo's ionising rate Rion(M, z) (user_functions.c:40-47)
// on the N_halo^3 grid (auxiliary.c:761-808), then the block average down to
// N_smoothed^3 (smooth_box, auxiliary.c:643-682) times (dx_s/dx_h)^3 fesc.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace sf21 {

// One thread per halo.  Positions are in N_halo cell units.  Per axis the
// reference truncates toward zero and, for a negative remainder, steps one
// cell down and uses |remainder| as the far weight (auxiliary.c:764-784);
// the cells wrap periodically.
__global__ void k_cic_rion(long nhalos, const float* __restrict__ halos, int ng, float k_rion,
                           float c2, float c3, float* __restrict__ grid) {
  for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < nhalos;
       h += (long)gridDim.x * blockDim.x) {
    const float m = halos[4 * h], x1 = halos[4 * h + 1], y1 = halos[4 * h + 2],
                z1 = halos[4 * h + 3];
    // Rion = c1 M (1+z)^c4 (M/c2)^c3 exp(-(c2/M)^3) / V_norm, k_rion = c1 (1+z)^c4 / V_norm
    const float q = c2 / m;
    const float rion = k_rion * m * powf(m / c2, c3) * expf(-(q * q * q));
    int x = (int)x1, y = (int)y1, z = (int)z1;
    float fx = x1 - (float)x, fy = y1 - (float)y, fz = z1 - (float)z;
    if (fx < 0.0f) { x -= 1; fx = -fx; }
    if (fy < 0.0f) { y -= 1; fy = -fy; }
    if (fz < 0.0f) { z -= 1; fz = -fz; }
    auto wrap = [ng](int v) { v %= ng; return v < 0 ? v + ng : v; };
    for (int a = 0; a < 2; ++a) {
      const int xi = wrap(x + a);
      const float wx = a ? fx : 1.0f - fx;
      for (int b = 0; b < 2; ++b) {
        const int yi = wrap(y + b);
        const float wy = b ? fy : 1.0f - fy;
        for (int c = 0; c < 2; ++c) {
          const int zi = wrap(z + c);
          const float wz = c ? fz : 1.0f - fz;
          atomicAdd(&grid[((long)xi * ng + yi) * ng + zi], wx * wy * wz * rion);
        }
      }
    }
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
