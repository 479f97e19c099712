// Kernels of the differentiable mode.  The forward stays the exact bubble
// finder; the backward differentiates a tau-relaxed model of it:
//   Delta = exp(g)/<exp(g)>,  rec = Delta^2,  ion = zeta E
//   s_R   = min(I_R - R_R, I_R) / max(R_R, rho),  rho = 1e-3 <rec>
//           (R_R, I_R: rec and ion filtered by the top hat W_R; the sign of
//            s_R is the centre test of get_HIIbubbles.c:339-349)
//   phi_R = softplus(u) - softplus(u - c), u = s_R / tau, c = kPhiCap:
//           a softplus bounded by c (one covering centre gives 1 - exp(-c));
//           unbounded, the centres' phi (up to ~1e6) would set the FP32
//           error of the ball convolution far above the small A values of
//           the transition region
//   A     = sum_R B_R (*) phi_R   (B_R: the lattice ball of radius R, (*)
//           periodic convolution)
//   x~    = q + (1 - q)(1 - exp(-A))            (q: quick fill, :281-288)
//   T     = (1 - x~) Delta,   observables <x~> and the binned P(k) of T.
//
// Box layout as in the bubble finder: index (i*n + j)*n + p.  "(p,j)-space"
// means transformed along p and j and real along i (the layout of the
// forward spectrum).  Unnormalised inverse transforms carry a factor N = n^3.
//
// Backward, recomputing the per-radius fields from the forward spectrum Z:
//   sweep 1 (per radius batch): W_R Z -> inverse -> (N R_R, s N I_R) ->
//     phi_R, two radii packed per complex line (phi_a + i phi_b) ->
//     forward -> C += (B_a - i B_b) FFT(phi_a + i phi_b); Re IFFT(C)/N = A.
//     (B_R real and even: Re IFFT((B_a - i B_b) FFT(phi_a + i phi_b)) =
//      N (B_a (*) phi_a + B_b (*) phi_b).)
//   observables and their adjoints (Abar, qbar, direct Deltabar);
//   sweep 2 (per radius batch): IFFT((B_a + i B_b) FFT(Abar)) = N (phibar_a
//     + i phibar_b); recomputed pair -> seeds (Rbar_R + i Ibar_R) -> forward
//     -> G += W_R FFT(seed_R); IFFT(G)/N = recbar + i ionbar.
#include "fft_core.h"
#include "kernels.h"

namespace sf21 {

static unsigned eb(long count) {
  long b = (count + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ---------------------------------------------------------------------------
// FP64 block reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_reduce2(double a, double b, double* r1, double* r2) {
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
}

// out[slot0] = sum part[2i], out[slot1] = sum part[2i+1]
__global__ __launch_bounds__(256) void k_final2(const double* __restrict__ part, int np,
                                                double* __restrict__ out, int slot0, int slot1) {
  __shared__ double r1[256], r2[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  block_reduce2(a, b, r1, r2);
  if (threadIdx.x == 0) {
    if (slot0 >= 0) out[slot0] = r1[0];
    if (slot1 >= 0) out[slot1] = r2[0];
  }
}

__global__ __launch_bounds__(256) void k_sum_exp(long n3, const float* __restrict__ g,
                                                 double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x)
    s += (double)expf(g[i]);
  block_reduce2(s, 0.0, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// Fingerprint of the inputs (g, E): per element a 64-bit mix (the
// splitmix64 finaliser) of the value's bits with its index, summed modulo
// 2^64.  The sum is exact and order-free, so the result does not depend on
// the launch shape; a changed value, or two values swapped between cells,
// changes it except with probability ~2^-64.  Tells whether the fields kept
// by a forward pass belong to the inputs of the following backward pass.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ void block_reduce2_u64(unsigned long long a, unsigned long long b,
                                                  unsigned long long* r1, unsigned long long* r2) {
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
}

__global__ __launch_bounds__(256) void k_fingerprint(long n3, const float* __restrict__ g,
                                                     const float* __restrict__ emis,
                                                     unsigned long long* __restrict__ part) {
  __shared__ unsigned long long r1[256], r2[256];
  unsigned long long a = 0, b = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const unsigned long long pos = (unsigned long long)i << 32;
    a += mix64(pos | __float_as_uint(g[i]));
    b += mix64((pos | __float_as_uint(emis[i])) ^ 0xD6E8FEB86659FD93ull);
  }
  block_reduce2_u64(a, b, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

__global__ __launch_bounds__(256) void k_fingerprint_final(const unsigned long long* __restrict__ part,
                                                           int np, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long r1[256], r2[256];
  unsigned long long a = 0, b = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  block_reduce2_u64(a, b, r1, r2);
  if (threadIdx.x == 0) {
    out[0] = r1[0];
    out[1] = r2[0];
  }
}

// Delta = exp(g) N / sum exp(g), rec = Delta^2, ion = zeta E; sum(rec) partials.
__global__ __launch_bounds__(256) void k_make_fields(long n3, const float* __restrict__ g,
                                                     const float* __restrict__ emis, float zf,
                                                     const double* __restrict__ st,
                                                     float* __restrict__ delta1,
                                                     float* __restrict__ rec,
                                                     float* __restrict__ ion,
                                                     double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  const float inv_mu = (float)((double)n3 / st[DS_SUM_EXP]);
  double sr = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float d = expf(g[i]) * inv_mu;
    delta1[i] = d;
    const float r = d * d;
    rec[i] = r;
    ion[i] = zf * emis[i];
    sr += (double)r;
  }
  block_reduce2(sr, 0.0, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// T = (1 - x) Delta (exact forward), sum(x) partials.
__global__ __launch_bounds__(256) void k_t_field(long n3, const float* __restrict__ x,
                                                 const float* __restrict__ delta1,
                                                 float* __restrict__ t,
                                                 double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  double sx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float xv = x[i];
    t[i] = (1.0f - xv) * delta1[i];
    sx += (double)xv;
  }
  block_reduce2(sx, 0.0, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Ball spectra: B_R = FFT of the indicator of {d : fold(d_i)^2 + fold(d_j)^2
// + fold(d_p)^2 <= d2max(R)} (the exact ball of the dilation, each periodic
// cell once); real and even in each axis, kept folded as [fi][fj][fp] with
// f in [0, n/2].
// ---------------------------------------------------------------------------
__global__ void k_ball_box(int n, int d2, float2* __restrict__ out) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int p = (int)(i % n), j = (int)((i / n) % n), a = (int)(i / ((long)n * n));
    const int m = fold_sq(a, n) + fold_sq(j, n) + fold_sq(p, n);
    out[i] = make_float2(m <= d2 ? 1.0f : 0.0f, 0.0f);
  }
}

__global__ void k_ball_fold(int n, const float2* __restrict__ spec, float* __restrict__ table) {
  const int H = n / 2 + 1;
  const long h3 = (long)H * H * H;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < h3;
       i += (long)gridDim.x * blockDim.x) {
    const int fp = (int)(i % H), fj = (int)((i / H) % H), fa = (int)(i / ((long)H * H));
    table[i] = spec[((long)fa * n + fj) * n + fp].x;
  }
}

__device__ __forceinline__ int fold_idx(int i, int n) { return i > (n >> 1) ? n - i : i; }

// ---------------------------------------------------------------------------
// Relaxation helpers
// ---------------------------------------------------------------------------
// phi(u) = softplus(u) - softplus(u - c) and phi'(u) = sigmoid(u) -
// sigmoid(u - c), c = 20, each with one exponential t = e^u:
//   phi  = log1p(t) - log1p(t e^-c), which is t (1 - e^-c) to FP32 precision
//          below u = -17 (log1p(t) = t (1 - t/2 ...), t < 4.2e-8), and c
//          above u = 2c (both softplus terms linear);
//   phi' = t (1 - e^-c) / ((1 + t)(1 + t e^-c)) -- the difference of the two
//          sigmoids without cancellation -- and e^(c-u) (1 - e^-c) above
//          u = 2c, where t would overflow.
constexpr float kPhiCap = 20.0f;
constexpr float kExpNegCap = 2.0611536224385583e-09f;  // e^-20
// Contraction off: the same operations in every kernel that inlines them, so
// the stored-field and recomputing backward passes agree bit for bit.
__device__ __forceinline__ float phi_of(float u) {
#pragma clang fp contract(off)
  if (u > 2.0f * kPhiCap) return kPhiCap;
  const float t = expf(u);
  if (u < -17.0f) return t * (1.0f - kExpNegCap);
  return log1pf(t) - log1pf(t * kExpNegCap);
}
__device__ __forceinline__ float dphi_of(float u) {
#pragma clang fp contract(off)
  if (u > 2.0f * kPhiCap) return expf(kPhiCap - u) * (1.0f - kExpNegCap);
  const float t = expf(u);
  const float den = (1.0f + t) * fmaf(t, kExpNegCap, 1.0f);
  return t * (1.0f - kExpNegCap) / den;
}

// rho in unnormalised units: N rho = 1e-3 sum(rec)
__device__ __forceinline__ float rho_u_of(const double* st) {
  return (float)(1e-3 * st[DS_SUM_REC]);
}

// phi_R of one cell from its pair in unnormalised units (ru = N R_R, iu =
// N I_R)
__device__ __forceinline__ float phi_cell(float ru, float iu, float rho_u, float inv_tau) {
  return phi_of(fminf(iu - ru, iu) / fmaxf(ru, rho_u) * inv_tau);
}

// Seed of one cell (the adjoints of the normalised fields, Rbar + i Ibar)
// from its pair (ru, iu) and N phibar:
//   sbar = phibar phi'(s/tau)/tau,  phi'(u) = sigmoid(u) - sigmoid(u - c),
//   Rbar = sbar ds/dR,  Ibar = sbar ds/dI,
//   ds/dI = 1/den, ds/dR = -[R >= 0]/den - [R > rho] num/den^2,
//   num = min(I - R, I), den = max(R, rho);
// the rho term sbar ds/drho = -sbar [R <= rho] num/rho^2 is added to rsum.
// In unnormalised units the factors N of ds/d(.) and of phibar cancel.
__device__ __forceinline__ float2 seed_cell(float ru, float iu, float phib, float rho_u,
                                            float inv_tau, double& rsum) {
  const float na = iu - ru;
  const bool use_a = na <= iu;
  const float num = use_a ? na : iu;
  const bool den_r = ru > rho_u;
  const float den = den_r ? ru : rho_u;
  const float inv_den = 1.0f / den;
  const float s = num * inv_den;
  const float sb = phib * dphi_of(s * inv_tau) * inv_tau;
  const float q = num * inv_den * inv_den;
  const float dR = (use_a ? -inv_den : 0.0f) - (den_r ? q : 0.0f);
  if (!den_r) rsum -= (double)sb * (double)q;
  return make_float2(sb * dR, sb * inv_den);
}

// ---------------------------------------------------------------------------
// Sweep 1, rows: inverse along p of each radius' pair, phi_R, pairs packed
// and transformed forward along p, written in place to slot r/2 (slot r/2's
// row of this block was read at an earlier or the same iteration).
// ---------------------------------------------------------------------------
// KEPT: the batch's real-space pairs come from pairs (kept by the exact
// forward pass) instead of an inverse transform of u; u only receives the
// phi pairs.
template <int L, bool KEPT = false>
__global__ __launch_bounds__(256) void k_rows_phi(float2* u, int nrad, long ustride,
                                                  const float* __restrict__ bal,
                                                  const double* __restrict__ st, float inv_tau,
                                                  const float2* __restrict__ tw,
                                                  float2* __restrict__ pairs, long pstride) {
  constexpr int E = Plan<L>::E, TPL = L / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<L>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const float inv_s = 1.0f / bal[0];
  const float rho_u = rho_u_of(st);
  float pa[E];
#pragma unroll
  for (int m = 0; m < E; ++m) pa[m] = 0.0f;
#pragma unroll 1
  for (int r = 0; r < nrad; ++r) {
    const int tl = opaque_lane(t);
    float2 v[E];
    if constexpr (KEPT) {
      const float2* src = pairs + (long)r * pstride + row * L;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ld_stream(src + tl + TPL * m);
    } else {
      const float2* src = u + (long)r * ustride + row * L;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = src[tl + TPL * m];
      line_fft<L, +1>(v, sre, sim, tl, tw);
    }
    if (pairs && !KEPT) {
      // stored for the second sweep (field-storage mode)
      float2* dst = pairs + (long)r * pstride + row * L;
#pragma unroll
      for (int m = 0; m < E; ++m) dst[tl + TPL * m] = v[m];
    }
    float ph[E];
#pragma unroll
    for (int m = 0; m < E; ++m) ph[m] = phi_cell(v[m].x, v[m].y * inv_s, rho_u, inv_tau);
    if ((r & 1) == 0 && r + 1 < nrad) {
#pragma unroll
      for (int m = 0; m < E; ++m) pa[m] = ph[m];
      continue;
    }
    float2 w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = (r & 1) ? make_float2(pa[m], ph[m]) : make_float2(ph[m], 0.0f);
    line_fft<L, -1>(w, sre, sim, tl, tw);
    float2* dst = u + (long)(r >> 1) * ustride + row * L;
#pragma unroll
    for (int m = 0; m < E; ++m) dst[tl + TPL * m] = w[m];
  }
}

// ---------------------------------------------------------------------------
// Ball passes along i.  bt: folded ball tables of the batch's radii (table r
// at bt + r*h3).  Sweep 1 (ACC): per pair, forward along i, times
// (B_a - i B_b), summed over the batch's pairs, one inverse along i, added to
// out ((p,j)-space accumulator).  Sweep 2 (!ACC): the (p,j)-space Abar line
// forward along i once, per pair times (B_a + i B_b), inverse along i, to
// slot pr of out.
// ---------------------------------------------------------------------------
template <int L, int TILE, bool ACC>
__global__ __launch_bounds__(TILE*(L / Plan<L>::E)) void k_ball_i(
    const float2* in, long istride, float2* out, long ostride, int nrad,
    const float* __restrict__ bt, long h3, const float2* __restrict__ tw) {
  constexpr int E = Plan<L>::E, TPL = L / E, NT = L / TILE, H = L / 2 + 1;
  __shared__ float lds[TILE * 2 * LdsLine<L>::kFloats];
  const int c = threadIdx.x % TILE, t = threadIdx.x / TILE;
  const int kj = blockIdx.x / NT;
  const int kp = (blockIdx.x % NT) * TILE + c;
  const long off = (long)kj * L + kp;
  float* sre = lds + c * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const int bjp = fold_idx(kj, L) * H + fold_idx(kp, L);
  const int npair = (nrad + 1) / 2;
  float2 z[E];
  if constexpr (ACC) {
#pragma unroll
    for (int m = 0; m < E; ++m) z[m] = make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) z[m] = in[off + (t + TPL * m) * (L * L)];
    line_fft<L, -1>(z, sre, sim, t, tw);
  }
#pragma unroll 1
  for (int pr = 0; pr < npair; ++pr) {
    const int tl = opaque_lane(t);
    const float* __restrict__ ta = bt + (long)(2 * pr) * h3;
    const float* __restrict__ tb = bt + (long)(2 * pr + 1) * h3;
    const bool two = 2 * pr + 1 < nrad;
    float2 v[E];
    if constexpr (ACC) {
      const float2* src = in + (long)pr * istride + off;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = src[(tl + TPL * m) * (L * L)];
      line_fft<L, -1>(v, sre, sim, tl, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ti = fold_idx(tl + TPL * m, L) * (H * H) + bjp;
        const float ba = ta[ti], bb = two ? tb[ti] : 0.0f;
        z[m].x += ba * v[m].x + bb * v[m].y;
        z[m].y += ba * v[m].y - bb * v[m].x;
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ti = fold_idx(tl + TPL * m, L) * (H * H) + bjp;
        const float ba = ta[ti], bb = two ? tb[ti] : 0.0f;
        v[m] = make_float2(ba * z[m].x - bb * z[m].y, ba * z[m].y + bb * z[m].x);
      }
      line_fft<L, +1>(v, sre, sim, tl, tw);
      float2* dst = out + (long)pr * ostride + off;
#pragma unroll
      for (int m = 0; m < E; ++m) dst[(tl + TPL * m) * (L * L)] = v[m];
    }
  }
  if constexpr (ACC) {
    line_fft<L, +1>(z, sre, sim, t, tw);
    float2* dst = out + off;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const float2 o = dst[(t + TPL * m) * (L * L)];
      dst[(t + TPL * m) * (L * L)] = make_float2(o.x + z[m].x, o.y + z[m].y);
    }
  }
}

// ---------------------------------------------------------------------------
// Sweep 2, rows: inverse along p of the recomputed pair (N R_R, s N I_R) and
// of the ball pair (N phibar_a, N phibar_b); the seeds (seed_cell) packed
// Rbar + i Ibar, forward along p, in place over the pair.  The rho terms are
// summed in FP64 into rho_acc[block] (accumulated over the batches).
// ---------------------------------------------------------------------------
template <int L, bool STORED>
__global__ __launch_bounds__(256) void k_rows_seed(float2* u, const float2* v, int nrad,
                                                   long ustride, long vstride,
                                                   const float* __restrict__ bal,
                                                   const double* __restrict__ st, float inv_tau,
                                                   double* __restrict__ rho_acc,
                                                   const float2* __restrict__ tw,
                                                   const float2* __restrict__ pairs, long pstride) {
  constexpr int E = Plan<L>::E, TPL = L / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<L>::kFloats];
  __shared__ double r1[256], r2[256];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const float inv_s = 1.0f / bal[0];
  const float rho_u = rho_u_of(st);
  double rsum = 0.0;
  const int npair = (nrad + 1) / 2;
#pragma unroll 1
  for (int pr = 0; pr < npair; ++pr) {
    float2 pb[E];
    {
      const int tl = opaque_lane(t);
      const float2* bs = v + (long)pr * vstride + row * L;
#pragma unroll
      for (int m = 0; m < E; ++m) pb[m] = bs[tl + TPL * m];
      line_fft<L, +1>(pb, sre, sim, tl, tw);
    }
    const int nh = (2 * pr + 1 < nrad) ? 2 : 1;
#pragma unroll 1
    for (int hh = 0; hh < nh; ++hh) {
      const int tl = opaque_lane(t);
      float2 y[E];
      float2* src = u + (long)(2 * pr + hh) * ustride + row * L;
      if constexpr (STORED) {
        // the pair as the first sweep left it (already inverse-transformed)
        const float2* ps = pairs + (long)(2 * pr + hh) * pstride + row * L;
#pragma unroll
        for (int m = 0; m < E; ++m) y[m] = ps[tl + TPL * m];
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m) y[m] = src[tl + TPL * m];
        line_fft<L, +1>(y, sre, sim, tl, tw);
      }
#pragma unroll
      for (int m = 0; m < E; ++m)
        y[m] = seed_cell(y[m].x, y[m].y * inv_s, hh ? pb[m].y : pb[m].x, rho_u, inv_tau, rsum);
      line_fft<L, -1>(y, sre, sim, tl, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) src[tl + TPL * m] = y[m];
    }
  }
  block_reduce2(rsum, 0.0, r1, r2);
  if (threadIdx.x == 0) rho_acc[blockIdx.x] += r1[0];
}

// ---------------------------------------------------------------------------
// Sweep 2, along i: forward of each seed line, times W_R, summed over the
// batch, one inverse along i, added to acc ((p,j)-space).
// ---------------------------------------------------------------------------
template <int L, int TILE, bool WTAB>
__global__ __launch_bounds__(TILE*(L / Plan<L>::E)) void k_filter_acc_i(
    const float2* in, long istride, int nrad, const float* __restrict__ wlut, long mlut,
    const float2* __restrict__ kappa, float2* acc, const float2* __restrict__ tw) {
  constexpr int E = Plan<L>::E, TPL = L / E, NT = L / TILE;
  __shared__ float lds[TILE * 2 * LdsLine<L>::kFloats];
  // mirror pairs (t, TPL - t) in adjacent slots: lanes ^ TILE share W
  const int c = threadIdx.x % TILE, t = paired_slot_to_t(threadIdx.x / TILE, TPL);
  const int kj = blockIdx.x / NT;
  const int kp = (blockIdx.x % NT) * TILE + c;
  const long off = (long)kj * L + kp;
  float* sre = lds + c * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const int mjp = fold_sq(kj, L) + fold_sq(kp, L);
  int mm[E / 2 + 1];
  SqrtPair sq[E / 2 + 1];
#pragma unroll
  for (int m = 0; m <= E / 2; ++m) {
    mm[m] = mjp + fold_sq((t + TPL * m) % L, L);
    sq[m] = sqrt_pair(mm[m]);
  }
  float2 z[E];
#pragma unroll
  for (int m = 0; m < E; ++m) z[m] = make_float2(0.f, 0.f);
#pragma unroll 1
  for (int r = 0; r < nrad; ++r) {
    const int tl = opaque_lane(t);
    const float2* src = in + (long)r * istride + off;
    float2 v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = src[(tl + TPL * m) * (L * L)];
    line_fft<L, -1>(v, sre, sim, tl, tw);
    // the same W_R as the forward pass (k_filter_i): table or FP32 evaluation
    float wv[E];
    if constexpr (WTAB) {
      const float* __restrict__ w = wlut + (long)r * mlut;
#pragma unroll
      for (int m = 0; m < E; ++m) wv[m] = w[mjp + fold_sq(tl + TPL * m, L)];
    } else {
      w_line_shared<L, E, TILE>(kappa[r], sq, mm, t, wv);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
      z[m].x += wv[m] * v[m].x;
      z[m].y += wv[m] * v[m].y;
    }
  }
  line_fft<L, +1>(z, sre, sim, t, tw);
  float2* dst = acc + off;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const float2 o = dst[(t + TPL * m) * (L * L)];
    dst[(t + TPL * m) * (L * L)] = make_float2(o.x + z[m].x, o.y + z[m].y);
  }
}

// ---------------------------------------------------------------------------
// Generic sides (no register plan): the per-radius steps of both sweeps as
// passes over full n^3 complex boxes; the 3D transforms between them run as
// separate line passes (k_gen_pass), and one radius pair shares a complex
// transform as on the plan sides.
// ---------------------------------------------------------------------------
// phi of one radius from its real-space pair (N R_R, s N I_R) into the real
// (half 0: imaginary part cleared) or imaginary part (half 1) of slot.
__global__ void k_gen_phi(long n3, const float2* __restrict__ pair, float2* __restrict__ slot,
                          int half, const float* __restrict__ bal, const double* __restrict__ st,
                          float inv_tau) {
  const float inv_s = 1.0f / bal[0];
  const float rho_u = rho_u_of(st);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float2 v = pair[i];
    const float ph = phi_cell(v.x, v.y * inv_s, rho_u, inv_tau);
    slot[i] = half ? make_float2(slot[i].x, ph) : make_float2(ph, 0.0f);
  }
}

// Ball products on the full spectrum (tb null: the pair's second radius is
// absent).  ACC (sweep 1): out += (B_a - i B_b) in; otherwise (sweep 2):
// out = (B_a + i B_b) in -- the products of k_ball_i.
template <bool ACC>
__global__ void k_gen_ball(int n, const float2* __restrict__ in, float2* __restrict__ out,
                           const float* __restrict__ ta, const float* __restrict__ tb) {
  const long n3 = (long)n * n * n;
  const int H = n / 2 + 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int p = (int)(i % n), j = (int)((i / n) % n), a = (int)(i / ((long)n * n));
    const long ti = ((long)fold_idx(a, n) * H + fold_idx(j, n)) * H + fold_idx(p, n);
    const float ba = ta[ti], bb = tb ? tb[ti] : 0.0f;
    const float2 v = in[i];
    if constexpr (ACC) {
      const float2 o = out[i];
      out[i] = make_float2(o.x + (ba * v.x + bb * v.y), o.y + (ba * v.y - bb * v.x));
    } else {
      out[i] = make_float2(ba * v.x - bb * v.y, ba * v.y + bb * v.x);
    }
  }
}

// Seeds of one radius in place over its real-space pair; phibar from the real
// (half 0) or imaginary part (half 1) of v = N (phibar_a + i phibar_b); the
// rho terms summed in FP64 into rho_acc[block].
__global__ __launch_bounds__(256) void k_gen_seed(long n3, float2* __restrict__ u,
                                                  const float2* __restrict__ v, int half,
                                                  const float* __restrict__ bal,
                                                  const double* __restrict__ st, float inv_tau,
                                                  double* __restrict__ rho_acc) {
  __shared__ double r1[256], r2[256];
  const float inv_s = 1.0f / bal[0];
  const float rho_u = rho_u_of(st);
  double rsum = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float2 y = u[i], b = v[i];
    u[i] = seed_cell(y.x, y.y * inv_s, half ? b.y : b.x, rho_u, inv_tau, rsum);
  }
  block_reduce2(rsum, 0.0, r1, r2);
  if (threadIdx.x == 0) rho_acc[blockIdx.x] += r1[0];
}

// acc += W_R(|k|) in on the full spectrum: the weight of the forward filter
// pass (k_gen_pass mode 1), the same function of the same |k|^2.
__global__ void k_gen_filter_acc(int n, const float2* __restrict__ in, float2* __restrict__ acc,
                                 float2 kappa) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int p = (int)(i % n), j = (int)((i / n) % n), a = (int)(i / ((long)n * n));
    const float w = w_tophat_k(kappa, fold_sq(a, n) + fold_sq(j, n) + fold_sq(p, n));
    const float2 v = in[i], o = acc[i];
    acc[i] = make_float2(o.x + w * v.x, o.y + w * v.y);
  }
}

// P(k) adjoint on the full spectrum F of T (the binned modes are the p <=
// n/2 half): H = 2 c_b F there, 0 elsewhere; the real part of the
// unnormalised inverse of H is dL/dT = sum_half 2 c_b Re(F e^{ikx}).
__global__ void k_pk_adjoint_full(int n, float2* __restrict__ spec,
                                  const short* __restrict__ bin_of_m,
                                  const double* __restrict__ coef) {
  const long n3 = (long)n * n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const int p = (int)(i % n);
    const long r = i / n;
    float c = 0.0f;
    if (p <= n / 2) {
      const int m = fold_sq((int)(r / n), n) + fold_sq((int)(r % n), n) + p * p;
      const int b = bin_of_m[m];
      c = b >= 0 ? 2.0f * (float)coef[b] : 0.0f;
    }
    const float2 v = spec[i];
    spec[i] = make_float2(v.x * c, v.y * c);
  }
}

// ---------------------------------------------------------------------------
// Elementwise passes
// ---------------------------------------------------------------------------
__device__ __forceinline__ float quick_fill_d(float rec, float ion) {
  if (ion > 0.0f) return rec > 0.0f ? fminf(ion / rec, 1.0f) : 1.0f;
  return 0.0f;
}

// From C = N A + i(...): x~, T = (1 - x~) Delta, dx~/dq = exp(-A),
// dx~/dA = (1 - q) exp(-A); sum(x~) partials.
__global__ __launch_bounds__(256) void k_relaxed_state(long n3, const float2* __restrict__ cacc,
                                                       const float* __restrict__ rec,
                                                       const float* __restrict__ ion,
                                                       const float* __restrict__ delta1,
                                                       float* __restrict__ xt,
                                                       float* __restrict__ t,
                                                       float* __restrict__ fq,
                                                       float* __restrict__ fa,
                                                       double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  const float inv_n3 = (float)(1.0 / (double)n3);
  double sx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float A = cacc[i].x * inv_n3;
    const float e = expf(-A);
    const float q = quick_fill_d(rec[i], ion[i]);
    const float x = q + (1.0f - q) * (1.0f - e);
    xt[i] = x;
    t[i] = (1.0f - x) * delta1[i];
    fq[i] = e;
    fa[i] = (1.0f - q) * e;
    sx += (double)x;
  }
  block_reduce2(sx, 0.0, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// P(k) adjoint on the half spectrum F of T: H = c_b F with c_b = dL/dP_b *
// norm / count_b, doubled on the planes p = 0 and p = n/2, so that the
// unnormalised C2R of H is dL/dT.
__global__ void k_pk_adjoint(int n, float2* __restrict__ spec, const short* __restrict__ bin_of_m,
                             const double* __restrict__ coef) {
  const int h = n / 2 + 1;
  const long nh = (long)n * n * h;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % h);
    const long r = i / h;
    const int m = fold_sq((int)(r / n), n) + fold_sq((int)(r % n), n) + q * q;
    const int b = bin_of_m[m];
    float c = b >= 0 ? (float)coef[b] : 0.0f;
    if (q == 0 || q == n / 2) c *= 2.0f;
    const float2 v = spec[i];
    spec[i] = make_float2(v.x * c, v.y * c);
  }
}

// Adjoint coefficients of the binned P(k): c_b = dL/dP_b norm / count_b
// (counts from the P(k) pass just run, pk[2 nbins + b]).
__global__ void k_pk_coef(const double* __restrict__ gp, const double* __restrict__ pk, int nbins,
                          double norm, double* __restrict__ coef) {
  const int b = threadIdx.x;
  if (b >= nbins) return;
  const double cnt = pk[2 * nbins + b];
  coef[b] = cnt > 0 ? gp[b] * norm / cnt : 0.0;
}

// xbar = gx - Tbar Delta (gx = dL/dX / N);  Abar = xbar (1 - q) exp(-A) (as a complex
// box);  qbar = xbar exp(-A);  direct Deltabar = Tbar (1 - x~).
__global__ void k_seed_a(long n3, const float* __restrict__ tbar, const float* __restrict__ delta1,
                         const float* __restrict__ xt, const float* __restrict__ fq,
                         const float* __restrict__ fa, float gx,
                         float2* __restrict__ abar, float* __restrict__ qbar,
                         float* __restrict__ dbar) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float tb = tbar[i];
    const float xb = gx - tb * delta1[i];
    abar[i] = make_float2(xb * fa[i], 0.0f);
    qbar[i] = xb * fq[i];
    dbar[i] = tb * (1.0f - xt[i]);
  }
}

// Per cell: recbar, ionbar from G = N (recbar + i ionbar) plus the quick-fill
// and rho terms, Deltabar = direct + 2 Delta recbar; FP64 partials of
// sum(ionbar E) (-> dL/dzeta) and sum(Deltabar Delta).
__global__ __launch_bounds__(256) void k_grad_cells(long n3, const float2* __restrict__ gacc,
                                                    const float* __restrict__ qbar,
                                                    const float* __restrict__ rec,
                                                    const float* __restrict__ ion,
                                                    const float* __restrict__ emis,
                                                    const float* __restrict__ delta1,
                                                    const float* __restrict__ dbar,
                                                    const double* __restrict__ st,
                                                    float* __restrict__ dd_out,
                                                    double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  const float inv_n3 = (float)(1.0 / (double)n3);
  // rho = 1e-3 sum(rec)/N: drho/drec = 1e-3/N per cell
  const float rho_coef = (float)(st[DS_RHO_BAR] * 1e-3 / (double)n3);
  double sz = 0.0, sd = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    const float2 gs = gacc[i];
    float rb = gs.x * inv_n3 + rho_coef;
    float ib = gs.y * inv_n3;
    const float r = rec[i], o = ion[i];
    if (o > 0.0f && r > 0.0f && o < r) {
      const float qb = qbar[i];
      ib += qb / r;
      rb -= qb * o / (r * r);
    }
    const float d = delta1[i];
    const float ddel = dbar[i] + 2.0f * d * rb;
    dd_out[i] = ddel;
    sz += (double)ib * (double)emis[i];
    sd += (double)ddel * (double)d;
  }
  block_reduce2(sz, sd, r1, r2);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

// gbar = Delta (Deltabar - S/N), S = sum(Deltabar Delta)  (Delta = e^g/mean e^g)
__global__ void k_grad_g(long n3, const float* __restrict__ delta1, const float* __restrict__ dd,
                         const double* __restrict__ st, float* __restrict__ gout) {
  const float sbar = (float)(st[DS_S] / (double)n3);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x)
    gout[i] = delta1[i] * (dd[i] - sbar);
}

__global__ void k_fill_c(float2* __restrict__ p, long count) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = make_float2(0.f, 0.f);
}
__global__ void k_fill_d(double* __restrict__ p, long count) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}
__global__ void k_to_complex(long n3, const float* __restrict__ in, float2* __restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x)
    out[i] = make_float2(in[i], 0.0f);
}

// sum of the per-block rho accumulators -> st[DS_RHO_BAR]
__global__ __launch_bounds__(256) void k_sum_rho(const double* __restrict__ acc, int np,
                                                 double* __restrict__ st) {
  __shared__ double r1[256], r2[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) a += acc[i];
  block_reduce2(a, 0.0, r1, r2);
  if (threadIdx.x == 0) st[DS_RHO_BAR] = r1[0];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int diff_partials_count(long n3) { return (int)eb(n3); }

void launch_sum_exp(long n3, const float* g, double* part, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_exp, dim3(eb(n3)), dim3(256), 0, s, n3, g, part);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, part, (int)eb(n3), st, DS_SUM_EXP, -1);
}
void launch_make_fields(long n3, const float* g, const float* emis, float zeta, double* st,
                        float* delta1, float* rec, float* ion, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_make_fields, dim3(eb(n3)), dim3(256), 0, s, n3, g, emis, zeta, st, delta1,
                     rec, ion, part);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, part, (int)eb(n3), st, DS_SUM_REC, -1);
}
void launch_t_field(long n3, const float* x, const float* delta1, float* t, double* part,
                    double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_t_field, dim3(eb(n3)), dim3(256), 0, s, n3, x, delta1, t, part);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, part, (int)eb(n3), st, DS_SUM_X, -1);
}
void launch_ball_box(int n, int d2, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_ball_box, dim3(eb((long)n * n * n)), dim3(256), 0, s, n, d2, out);
}
void launch_ball_fold(int n, const float2* spec, float* table, hipStream_t s) {
  const long h = n / 2 + 1;
  hipLaunchKernelGGL(k_ball_fold, dim3(eb(h * h * h)), dim3(256), 0, s, n, spec, table);
}

template <int L>
static void rows_phi_t(int n, float2* u, int nrad, long ustride, const float* bal, const double* st,
                       float inv_tau, const float2* tw, float2* pairs, long pstride, bool kept,
                       hipStream_t s) {
  constexpr int LPB = 256 / (L / Plan<L>::E);
  const dim3 g((unsigned)((long)n * n / LPB));
  if (kept)
    hipLaunchKernelGGL((k_rows_phi<L, true>), g, dim3(256), 0, s, u, nrad, ustride, bal, st,
                       inv_tau, tw, pairs, pstride);
  else
    hipLaunchKernelGGL((k_rows_phi<L, false>), g, dim3(256), 0, s, u, nrad, ustride, bal, st,
                       inv_tau, tw, pairs, pstride);
}
template <int L>
static void rows_seed_t(int n, float2* u, const float2* v, int nrad, long ustride, long vstride,
                        const float* bal, const double* st, float inv_tau, double* rho_acc,
                        const float2* tw, const float2* pairs, long pstride, hipStream_t s) {
  constexpr int LPB = 256 / (L / Plan<L>::E);
  const dim3 g((unsigned)((long)n * n / LPB));
  if (pairs)
    hipLaunchKernelGGL((k_rows_seed<L, true>), g, dim3(256), 0, s, u, v, nrad, ustride, vstride, bal,
                       st, inv_tau, rho_acc, tw, pairs, pstride);
  else
    hipLaunchKernelGGL((k_rows_seed<L, false>), g, dim3(256), 0, s, u, v, nrad, ustride, vstride,
                       bal, st, inv_tau, rho_acc, tw, pairs, pstride);
}
template <int L, bool ACC>
static void ball_i_t(const float2* in, long istride, float2* out, long ostride, int nrad,
                     const float* bt, const float2* tw, hipStream_t s) {
  constexpr int TILE = 16, TPL = L / Plan<L>::E;
  const long H = L / 2 + 1;
  hipLaunchKernelGGL((k_ball_i<L, TILE, ACC>), dim3((unsigned)((long)L * (L / TILE))),
                     dim3(TILE * TPL), 0, s, in, istride, out, ostride, nrad, bt, H * H * H, tw);
}
template <int L>
static void filter_acc_t(const float2* in, long istride, int nrad, const float* wlut, long mlut,
                         const float2* kappa, int wmode, float2* acc, const float2* tw,
                         hipStream_t s) {
  constexpr int TILE = 16, TPL = L / Plan<L>::E;
  const dim3 g((unsigned)((long)L * (L / TILE))), b(TILE * TPL);
  if (wmode == 0)
    hipLaunchKernelGGL((k_filter_acc_i<L, TILE, true>), g, b, 0, s, in, istride, nrad, wlut, mlut,
                       kappa, acc, tw);
  else
    hipLaunchKernelGGL((k_filter_acc_i<L, TILE, false>), g, b, 0, s, in, istride, nrad, wlut, mlut,
                       kappa, acc, tw);
}

#define SF21_DIFF_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(512)

bool diff_size_supported(int n) {
  switch (n) {
#define C(L) case L: return true;
    SF21_DIFF_SIZES(C)
#undef C
  }
  return false;
}
int rows_blocks(int n) {
  switch (n) {
#define C(L) case L: return (int)((long)n * n / (256 / (L / Plan<L>::E)));
    SF21_DIFF_SIZES(C)
#undef C
  }
  return 0;
}
void launch_fingerprint(long n3, const float* g, const float* emis, double* part,
                        unsigned long long* out, hipStream_t s) {
  // part holds 2 eb(n3) doubles; the same bytes carry the 64-bit partials
  auto* pu = reinterpret_cast<unsigned long long*>(part);
  hipLaunchKernelGGL(k_fingerprint, dim3(eb(n3)), dim3(256), 0, s, n3, g, emis, pu);
  hipLaunchKernelGGL(k_fingerprint_final, dim3(1), dim3(256), 0, s, pu, (int)eb(n3), out);
}
void launch_rows_phi(int n, float2* u, int nrad, long ustride, const float* bal, const double* st,
                     float inv_tau, const float2* tw, float2* pairs, long pstride, hipStream_t s,
                     bool kept) {
  switch (n) {
#define C(L) case L: rows_phi_t<L>(n, u, nrad, ustride, bal, st, inv_tau, tw, pairs, pstride, kept, s); break;
    SF21_DIFF_SIZES(C)
#undef C
  }
}
void launch_rows_seed(int n, float2* u, const float2* v, int nrad, long ustride, long vstride,
                      const float* bal, const double* st, float inv_tau, double* rho_acc,
                      const float2* tw, const float2* pairs, long pstride, hipStream_t s) {
  switch (n) {
#define C(L) case L: rows_seed_t<L>(n, u, v, nrad, ustride, vstride, bal, st, inv_tau, rho_acc, tw, pairs, pstride, s); break;
    SF21_DIFF_SIZES(C)
#undef C
  }
}
void launch_ball_i(int n, bool acc, const float2* in, long istride, float2* out, long ostride,
                   int nrad, const float* bt, const float2* tw, hipStream_t s) {
  switch (n) {
#define C(L)                                                                   \
  case L:                                                                      \
    if (acc) ball_i_t<L, true>(in, istride, out, ostride, nrad, bt, tw, s);    \
    else ball_i_t<L, false>(in, istride, out, ostride, nrad, bt, tw, s);       \
    break;
    SF21_DIFF_SIZES(C)
#undef C
  }
}
void launch_filter_acc_i(int n, const float2* in, long istride, int nrad, const float* wlut,
                         long mlut, const float2* kappa, int wmode, float2* acc, const float2* tw,
                         hipStream_t s) {
  switch (n) {
#define C(L) case L: filter_acc_t<L>(in, istride, nrad, wlut, mlut, kappa, wmode, acc, tw, s); break;
    SF21_DIFF_SIZES(C)
#undef C
  }
}

void launch_relaxed_state(long n3, const float2* cacc, const float* rec, const float* ion,
                          const float* delta1, float* xt, float* t, float* fq, float* fa,
                          double* part, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_relaxed_state, dim3(eb(n3)), dim3(256), 0, s, n3, cacc, rec, ion, delta1, xt,
                     t, fq, fa, part);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, part, (int)eb(n3), st, DS_SUM_XT, -1);
}
void launch_pk_adjoint(int n, float2* spec, const short* bin_of_m, const double* coef,
                       hipStream_t s) {
  const long nh = (long)n * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_pk_adjoint, dim3(eb(nh)), dim3(256), 0, s, n, spec, bin_of_m, coef);
}
void launch_pk_coef(const double* gp, const double* pk, int nbins, double norm, double* coef,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_pk_coef, dim3(1), dim3(64), 0, s, gp, pk, nbins, norm, coef);
}
void launch_seed_a(long n3, const float* tbar, const float* delta1, const float* xt,
                   const float* fq, const float* fa, float gx_n, float2* abar, float* qbar,
                   float* dbar, hipStream_t s) {
  hipLaunchKernelGGL(k_seed_a, dim3(eb(n3)), dim3(256), 0, s, n3, tbar, delta1, xt, fq, fa, gx_n,
                     abar, qbar, dbar);
}
void launch_sum_rho(const double* acc, int np, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_rho, dim3(1), dim3(256), 0, s, acc, np, st);
}
void launch_grad_cells(long n3, const float2* gacc, const float* qbar, const float* rec,
                       const float* ion, const float* emis, const float* delta1,
                       const float* dbar, double* st, float* dd_out, double* part,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_grad_cells, dim3(eb(n3)), dim3(256), 0, s, n3, gacc, qbar, rec, ion, emis,
                     delta1, dbar, st, dd_out, part);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(256), 0, s, part, (int)eb(n3), st, DS_ZETA_BAR, DS_S);
}
void launch_grad_g(long n3, const float* delta1, const float* dd, const double* st, float* gout,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_grad_g, dim3(eb(n3)), dim3(256), 0, s, n3, delta1, dd, st, gout);
}
void launch_fill_c(float2* p, long count, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_c, dim3(eb(count)), dim3(256), 0, s, p, count);
}
void launch_fill_d(double* p, long count, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_d, dim3(eb(count)), dim3(256), 0, s, p, count);
}
void launch_to_complex(long n3, const float* in, float2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_to_complex, dim3(eb(n3)), dim3(256), 0, s, n3, in, out);
}

int gen_seed_blocks(long n3) { return (int)eb(n3); }
void launch_gen_phi(long n3, const float2* pair, float2* slot, int half, const float* bal,
                    const double* st, float inv_tau, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_phi, dim3(eb(n3)), dim3(256), 0, s, n3, pair, slot, half, bal, st,
                     inv_tau);
}
void launch_gen_ball(int n, bool acc, const float2* in, float2* out, const float* ta,
                     const float* tb, hipStream_t s) {
  const long n3 = (long)n * n * n;
  if (acc) hipLaunchKernelGGL(k_gen_ball<true>, dim3(eb(n3)), dim3(256), 0, s, n, in, out, ta, tb);
  else hipLaunchKernelGGL(k_gen_ball<false>, dim3(eb(n3)), dim3(256), 0, s, n, in, out, ta, tb);
}
void launch_gen_seed(long n3, float2* u, const float2* v, int half, const float* bal,
                     const double* st, float inv_tau, double* rho_acc, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_seed, dim3(eb(n3)), dim3(256), 0, s, n3, u, v, half, bal, st, inv_tau,
                     rho_acc);
}
void launch_gen_filter_acc(int n, const float2* in, float2* acc, float2 kappa, hipStream_t s) {
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_gen_filter_acc, dim3(eb(n3)), dim3(256), 0, s, n, in, acc, kappa);
}
void launch_pk_adjoint_full(int n, float2* spec, const short* bin_of_m, const double* coef,
                            hipStream_t s) {
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_pk_adjoint_full, dim3(eb(n3)), dim3(256), 0, s, n, spec, bin_of_m, coef);
}

}  // namespace sf21
