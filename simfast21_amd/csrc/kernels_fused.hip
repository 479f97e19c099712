// Fused exact-path kernels of the bubble finder (get_HIIbubbles.c:216-479,
// restructured for one GPU):
//
//   k_fwd_rows      per slice: cell pre-pass (:216-225, :269-274 or direct
//                   boxes) fused into the forward transform along p of the
//                   packed field z = rec + i*ion (both r2c of :290-291 in one
//                   complex transform).
//   k_cols          forward along j (per slice), inverse along j (per radius).
//   k_filter_i      per radius batch: forward along i, then for each radius
//                   the top-hat filter (:305-322) and the inverse along i.
//   k_inv_rows_crit per radius: inverse along p and the bubble-centre test
//                   (:327-352, :437-438); centres record the ball bound of
//                   the radius in a per-cell max map.
//   k_dt_*          exact periodic dilation of all centres by their balls at
//                   once (replaces :358-393 and :428-458): a separable
//                   max-plus parabola transform of the weight map; the last
//                   pass is fused with the quick fill (:277-288), the clamp
//                   (:466-471) and the FP64 mean (:473-477).
#include "fft_core.h"
#include "kernels.h"
#include "cells.h"
#include <climits>

namespace sf21 {

// ------------------------------------------------------------------------
// Channel balance of the packed transform.  rec and ion share one complex
// transform (z = rec + i*s*ion); FP32 rounding of a complex FFT is spread
// over both channels in proportion to |z|, so without balancing the smaller
// field inherits the larger one's absolute error.  s = 2^round(log2(
// rms(rec)/rms(ion))) gives both channels the same RMS; a power of two keeps
// the scaling exact (the centre test becomes Im >= s*Re, and Im/s is the
// unscaled ion field bit for bit).  Sums in FP64.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_balance_partial(long n3, const float* __restrict__ a,
                                                         const float* __restrict__ b,
                                                         const float* __restrict__ c, CellConsts k,
                                                         double* __restrict__ part) {
  double sr = 0.0, si = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    float rec, ion, fr;
    cell_fields(k, a, b, c, i, rec, ion, fr);
    sr += (double)rec * rec;
    si += (double)ion * ion;
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = sr;
  r2[threadIdx.x] = si;
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

__global__ __launch_bounds__(256) void k_balance_final(const double* __restrict__ part, int np,
                                                       float* __restrict__ scale) {
  __shared__ double r1[256], r2[256];
  double sr = 0.0, si = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    sr += part[2 * i];
    si += part[2 * i + 1];
  }
  r1[threadIdx.x] = sr;
  r2[threadIdx.x] = si;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r1[threadIdx.x] += r1[threadIdx.x + s];
      r2[threadIdx.x] += r2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 0;
    if (r1[0] > 0.0 && r2[0] > 0.0) {
      e = (int)rint(0.5 * log2(r1[0] / r2[0]));
      e = e < -60 ? -60 : (e > 60 ? 60 : e);
    }
    scale[0] = ldexpf(1.0f, e);
  }
}

// ------------------------------------------------------------------------
// Forward along p with the fused cell pre-pass.
// ------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(256) void k_fwd_rows(const float* __restrict__ a,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ c, CellConsts k,
                                                  const float* __restrict__ bal,
                                                  float2* __restrict__ y,
                                                  const float2* __restrict__ tw) {
  constexpr int E = Plan<L>::E, TPL = L / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<L>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const float sc = bal[0];
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    float rec, ion, fr;
    cell_fields(k, a, b, c, row * L + t + TPL * m, rec, ion, fr);
    v[m] = make_float2(rec, ion * sc);
  }
  line_fft<L, -1>(v, sre, sim, t, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) y[row * L + t + TPL * m] = v[m];
}

// ------------------------------------------------------------------------
// Strided complex transform along j (axis 0) or i (axis 1) of an
// L x L x width array; TILE adjacent columns per block (128 B runs).
// blockIdx.y selects one of several arrays spaced bstride apart.
// ------------------------------------------------------------------------
template <int L, int SIGN, int TILE, int E = Plan<L>::E, bool NT = false>
__global__ __launch_bounds__(TILE*(L / E)) void k_cols(const float2* in, float2* out,
                                                       int width, int axis,
                                                       long bstride,
                                                       const float2* __restrict__ tw) {
  constexpr int TPL = L / E;
  __shared__ float lds[TILE * 2 * LdsLine<L>::kFloats];
  const int c = threadIdx.x % TILE, t = threadIdx.x / TILE;
  const int ntile = (width + TILE - 1) / TILE;
  const int outer = blockIdx.x / ntile;
  const int col = (blockIdx.x % ntile) * TILE + c;
  const bool ok = col < width;
  long base, stride;
  if (axis == 0) {
    base = (long)outer * L * width + col;
    stride = width;
  } else if (axis == 1) {
    base = (long)outer * width + col;
    stride = (long)L * width;
  } else {  // 2: along j in the 16-column tile layout (width L)
    base = (long)outer * L * width + (long)(col >> 4) * (L * 16) + (col & 15);
    stride = 16;
  }
  base += (long)blockIdx.y * bstride;
  float* sre = lds + c * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  float2 v[E];
  if (ok) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const float2* p = in + base + (long)(t + TPL * m) * stride;
      v[m] = NT ? ld_stream(p) : *p;
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = make_float2(0.f, 0.f);
  }
  line_fft_e<L, E, SIGN>(v, sre, sim, t, tw);
  if (ok) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float2* p = out + base + (long)(t + TPL * m) * stride;
      if (NT) st_stream(p, v[m]);
      else *p = v[m];
    }
  }
}

// The same pass with each block walking `per_block` consecutive tiles (tile
// q: batch q / (ntile L), outer (q / ntile) mod L, column tile q mod ntile)
// and loading tile q + 1 into a second register set before transforming and
// storing tile q, so that one resident block per CU keeps loads in flight
// while it computes.  Tiles are disjoint, so the pass may run in place.
template <int L, int SIGN, int TILE, int E, bool NT>
__global__ __launch_bounds__(TILE*(L / E)) void k_cols_pipe(const float2* in, float2* out,
                                                            int width, int axis, long bstride,
                                                            const float2* __restrict__ tw,
                                                            long ntiles, int per_block) {
  constexpr int TPL = L / E;
  __shared__ float lds[TILE * 2 * LdsLine<L>::kFloats];
  const int c = threadIdx.x % TILE;
  const int ntile = (width + TILE - 1) / TILE;
  const long q0 = (long)blockIdx.x * per_block;
  const long q1 = q0 + per_block < ntiles ? q0 + per_block : ntiles;
  float* sre = lds + c * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  // axis 2: along j in the 16-column tile layout (width L)
  const long stride = axis == 0 ? (long)width : (axis == 1 ? (long)L * width : 16L);
  // first element of column c of tile q, -1 past the last column
  auto origin = [&](long q) -> long {
    const long y = q / ((long)ntile * L);
    const long rem = q - y * ((long)ntile * L);
    const int outer = (int)(rem / ntile);
    const int col = (int)(rem % ntile) * TILE + c;
    if (col >= width) return -1;
    const long b = axis == 0   ? (long)outer * L * width + col
                   : axis == 1 ? (long)outer * width + col
                               : (long)outer * L * width + (long)(col >> 4) * (L * 16) + (col & 15);
    return b + y * bstride;
  };
  float2 cur[E], nxt[E];
  long cb = origin(q0);
  {
    const int t = (int)threadIdx.x / TILE;
#pragma unroll
    for (int m = 0; m < E; ++m)
      cur[m] = cb < 0 ? make_float2(0.f, 0.f)
                      : (NT ? ld_stream(in + cb + (long)(t + TPL * m) * stride)
                            : in[cb + (long)(t + TPL * m) * stride]);
  }
#pragma unroll 1
  for (long q = q0; q < q1; ++q) {
    const int t = opaque_lane((int)threadIdx.x / TILE);
    long nb = -1;
    if (q + 1 < q1) {
      nb = origin(q + 1);
#pragma unroll
      for (int m = 0; m < E; ++m)
        nxt[m] = nb < 0 ? make_float2(0.f, 0.f)
                        : (NT ? ld_stream(in + nb + (long)(t + TPL * m) * stride)
                              : in[nb + (long)(t + TPL * m) * stride]);
    }
    line_fft_e<L, E, SIGN>(cur, sre, sim, t, tw);
    if (cb >= 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        float2* p = out + cb + (long)(t + TPL * m) * stride;
        if (NT) st_stream(p, cur[m]);
        else *p = cur[m];
      }
    }
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = nxt[m];
    cb = nb;
  }
}

// ------------------------------------------------------------------------
// Per radius batch: forward along i of the (p, j)-transformed spectrum,
// top-hat filter W(|k| R) for each radius (get_HIIbubbles.c:305-322; W from
// an FP64 table indexed by the integer |k|^2, rounded to float), inverse
// along i.  The forward spectrum along i never leaves the registers.
// ------------------------------------------------------------------------
// JPAIR: a block takes the lines at kj and L - kj (kj = 1 + block row, up
// to L/2 - 1) of the same p columns, which share every W value (|k|^2
// depends on kj^2 only): one W evaluation serves two lines, whose inverse
// transforms run one after the other through the same LDS region.  Without
// JPAIR a block takes the line kj = block row * kj_step (kj_step L/2 covers
// the unpaired lines 0 and L/2).
template <int L, int TILE, bool WTAB, bool JPAIR = false>
__global__ __launch_bounds__(TILE*(L / Plan<L>::E)) void k_filter_i(
    const float2* __restrict__ spec, float2* __restrict__ u, int nrad, long ustride,
    const float* __restrict__ wlut, long mlut, const float2* __restrict__ kappa,
    const float2* __restrict__ tw, int kj_step, int t16) {
  constexpr int E = Plan<L>::E, TPL = L / E, NT = L / TILE;
  constexpr int NL = JPAIR ? 2 : 1;
  __shared__ float lds[TILE * 2 * LdsLine<L>::kFloats];
  // mirror pairs (t, TPL - t) in adjacent slots: lanes ^ TILE share W
  const int c = threadIdx.x % TILE, t = paired_slot_to_t(threadIdx.x / TILE, TPL);
  const int row = blockIdx.x / NT;
  int kjs[NL];
  kjs[0] = JPAIR ? 1 + row : row * kj_step;
  if constexpr (JPAIR) kjs[1] = L - 1 - row;
  const int kp = (blockIdx.x % NT) * TILE + c;
  float* sre = lds + c * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  float2 z[NL][E];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    // element e of this line sits at line[e * L * L] (32-bit offsets: L^3 < 2^31)
    const float2* __restrict__ src = spec + (long)kjs[l] * L + kp;
#pragma unroll
    for (int m = 0; m < E; ++m) z[l][m] = src[(t + TPL * m) * (L * L)];
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) line_fft<L, -1>(z[l], sre, sim, t, tw);
  const int mjp = fold_sq(kjs[0], L) + fold_sq(kp, L);
  // |k|^2 and its square root for the lower-half positions (radius-independent)
  int mm[E / 2 + 1];
  SqrtPair sq[E / 2 + 1];
#pragma unroll
  for (int m = 0; m <= E / 2; ++m) {
    mm[m] = mjp + fold_sq((t + TPL * m) % L, L);
    sq[m] = sqrt_pair(mm[m]);
  }
#pragma unroll 1
  for (int r = 0; r < nrad; ++r) {
    const int tl = opaque_lane(t);
    float wv[E];
    if constexpr (WTAB) {
      const float* __restrict__ w = wlut + (long)r * mlut;
#pragma unroll
      for (int m = 0; m < E; ++m) wv[m] = w[mjp + fold_sq(tl + TPL * m, L)];
    } else {
      w_line_shared<L, E, TILE>(kappa[r], sq, mm, t, wv);  // dk R as hi + lo
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      float2 v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = make_float2(z[l][m].x * wv[m], z[l][m].y * wv[m]);
      line_fft<L, +1>(v, sre, sim, tl, tw);
      // t16: the j pass's layout -- 16-column tiles of the (j, p) plane stored
      // as contiguous L x 16 blocks, element (j, p) at (p/16) 16 L + 16 j + p%16
      float2* __restrict__ dst =
          u + (long)r * ustride +
          (t16 ? (long)(kp >> 4) * (L * 16) + (long)kjs[l] * 16 + (kp & 15) : (long)kjs[l] * L + kp);
#pragma unroll
      for (int m = 0; m < E; ++m) st_stream(dst + (tl + TPL * m) * (L * L), v[m]);
    }
  }
}

// ------------------------------------------------------------------------
// Per radius batch: inverse along p -> (N^3 rec_R, s N^3 ion_R) for each of
// the nrad radii, centre test, and for every cell the largest ball bound
// among the batch's radii at which it is a centre.  Batches are issued in
// ascending radius, so a batch's bounds exceed every earlier batch's and a
// plain store (only where some centre was found) leaves in wmap the bound of
// the largest radius -- no atomics and no reads of the map.  dbg (nullable)
// receives the unscaled filtered pairs, keep (nullable) the pairs as
// computed (N^3 rec_R, s N^3 ion_R), for the differentiable mode's backward.
// ------------------------------------------------------------------------
template <int L, int E = Plan<L>::E>
__global__ __launch_bounds__(256) void k_inv_rows_crit(const float2* __restrict__ vin,
                                                       int nrad, long vstride,
                                                       const int* __restrict__ d2max,
                                                       const float* __restrict__ bal,
                                                       int* __restrict__ wmap,
                                                       float2* __restrict__ dbg,
                                                       float2* __restrict__ keep,
                                                       const float2* __restrict__ tw, int t16) {
  constexpr int TPL = L / E, LPB = 256 / TPL;
  __shared__ float lds[LPB * 2 * LdsLine<L>::kFloats];
  const int lw = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const long row = (long)blockIdx.x * LPB + lw;
  float* sre = lds + lw * 2 * LdsLine<L>::kFloats;
  float* sim = sre + LdsLine<L>::kFloats;
  const float sc = bal[0];
  int wbest[E];
#pragma unroll
  for (int m = 0; m < E; ++m) wbest[m] = -1;
#pragma unroll 1
  for (int r = 0; r < nrad; ++r) {
    const int tl = opaque_lane(t);
    float2 v[E];
    if (t16) {
      // 16-column tile layout: element p of row (i, j) at i L^2 + (p/16) 16 L
      // + 16 j + p%16
      const float2* __restrict__ src =
          vin + (long)r * vstride + (row / L) * ((long)L * L) + (row % L) * 16;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int p = tl + TPL * m;
        v[m] = ld_stream(src + (p >> 4) * (L * 16) + (p & 15));
      }
    } else {
      const float2* __restrict__ src = vin + (long)r * vstride + row * L;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ld_stream(src + tl + TPL * m);
    }
    line_fft_e<L, E, +1>(v, sre, sim, tl, tw);
    const int w = d2max[r];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      // Re = N^3 rec_R, Im = s N^3 ion_R: centre test on (s rec_R, s ion_R)
      if (is_centre(v[m].x * sc, v[m].y)) wbest[m] = max(wbest[m], w);
      if (dbg) dbg[(long)r * vstride + row * L + tl + TPL * m] = make_float2(v[m].x, v[m].y / sc);
      if (keep) st_stream(keep + (long)r * vstride + row * L + tl + TPL * m, v[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < E; ++m)
    if (wbest[m] >= 0) wmap[row * L + t + TPL * m] = wbest[m];
}

__global__ void k_fill_i32(int* __restrict__ p, int value, long count) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = value;
}

// ------------------------------------------------------------------------
// Exact dilation: g(x) = max_y (w(y) - minimg|x-y|^2) over cells with
// w(y) >= 0, computed one axis at a time (the max-plus transform with a
// parabola is separable).  A cell is ionised iff g(x) >= 0, i.e. iff some
// centre's ball {|d|^2 <= d2max(R)} covers it.  Negative intermediates can
// never become non-negative again and are stored as -1.
//
// 1D pass over a periodic line of n values, one wavefront per line, the line
// in LDS.  On the unrolled line, h(x, y) = f(y) - (x-y)^2 is supermodular
// (h(x1,y1) + h(x2,y2) - h(x1,y2) - h(x2,y1) = 2(x2-x1)(y2-y1) > 0), so the
// maximisers of x are non-decreasing in x.  Outputs are found in three
// levels: every 8*Q-th x over its whole window (8 lanes splitting it), every
// Q-th x inside the bracket of its level-0 neighbours, the rest inside the
// bracket of their level-1 neighbours.  A
// neighbour xa that found no non-negative value bounds the search too: a
// y < xa with f(y) >= (x-y)^2 > (xa-y)^2 would contradict it.  Only values
// >= 0 are ever needed, so levels 0 and 1 never look farther than
// floor(sqrt(fmax)); level 2 scans its bracket once for all of a lane's
// cells.
// ------------------------------------------------------------------------
struct Found {
  int best;  // -1: nothing >= 0
  int lo, hi;
};

template <class F>
__device__ __forceinline__ void consider(int x, int y, F f, Found& r) {
  const int v = f(y);
  if (v < 0) return;
  const int d = x - y;
  const int c = v - d * d;
  if (c > r.best) {
    r.best = c;
    r.lo = y;
    r.hi = y;
  } else if (c == r.best) {
    r.lo = min(r.lo, y);
    r.hi = max(r.hi, y);
  }
}

template <class F>
__device__ __forceinline__ Found scan_range(int x, int lo, int hi, int hl, F f) {
  Found r{-1, 0, 0};
  lo = max(lo, x - hl);
  hi = min(hi, x + hl);
  for (int y = lo; y <= hi; ++y) consider(x, y, f, r);
  return r;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int isqrt_floor(int v) {
  int r = (int)sqrtf((float)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

// Transform one line held in LDS (element e at col[e + (e >> 3)]); each of
// the P1 = n/Q active lanes returns the outputs x = Q*lane + q, q < Q.
template <int QMAX>
__device__ __forceinline__ void dt_line_wave(const int* col, int n, int Q, int hcap, int lane,
                                             int (&out)[QMAX]) {
  const int P1 = n / Q;
  const bool act = lane < P1;
  auto f = [&](int y) {
    y = y < 0 ? y + n : (y >= n ? y - n : y);
    return col[y + (y >> 3)];
  };
  int lmax = -1;
  if (act)
    for (int q = 0; q < Q; ++q) lmax = max(lmax, col[Q * lane + q + ((Q * lane + q) >> 3)]);
  const int fmax = wave_max(lmax);
  if (fmax < 0) {
#pragma unroll
    for (int q = 0; q < QMAX; ++q) out[q] = -1;
    return;
  }
  const int hl = min(hcap, isqrt_floor(fmax));
  const int x1 = Q * lane;
  // level 0: the points x = 8*Q*m, one per group of 8 lanes (8m .. 8m + 7):
  // lane k of the group takes the candidates x + d, d = -hl + k (mod 8), of
  // the window [x - hl, x + hl], then a 3-step butterfly leaves the group's
  // (best, lo, hi) in every lane of it
  Found l0{-1, 0, 0};
  {
    const int g0 = lane & ~7;
    if (g0 < P1) {
      const int x0 = Q * g0;
      for (int d = -hl + (lane & 7); d <= hl; d += 8) consider(x0, x0 + d, f, l0);
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
      const int ob = __shfl_xor(l0.best, o, 64);
      const int olo = __shfl_xor(l0.lo, o, 64), ohi = __shfl_xor(l0.hi, o, 64);
      if (ob > l0.best) {
        l0.best = ob;
        l0.lo = olo;
        l0.hi = ohi;
      } else if (ob == l0.best && ob >= 0) {
        l0.lo = min(l0.lo, olo);
        l0.hi = max(l0.hi, ohi);
      }
    }
  }
  // level 1: bracket [xa, xb] of level-0 points, xb = xa + 8Q (wrapping to 0 + n)
  const int ga = lane & ~7;
  const int gb = ga + 8;
  const int wb = gb < P1 ? gb : 0;
  const int a_best = __shfl(l0.best, ga, 64), a_lo = __shfl(l0.lo, ga, 64);
  const int b_best = __shfl(l0.best, wb, 64), b_hi0 = __shfl(l0.hi, wb, 64);
  Found l1 = l0;
  if (act && (lane & 7) != 0) {
    const int xa = Q * ga, xb = Q * gb;
    const int lo = a_best >= 0 ? a_lo : xa;
    const int hi = b_best >= 0 ? b_hi0 + (gb < P1 ? 0 : n) : xb;
    l1 = scan_range(x1, lo, hi, hl, f);
  }
  // level 2: the cells x1 + q, 0 < q < Q, all inside the bracket [lo, hi]
  // between x1's and the next lane's maximisers, in one pass over it.  No
  // window clip: with h = min(floor(sqrt(max w)), n/2) a candidate outside
  // x's window is either farther than sqrt(f(y)) (negative value) or an
  // unrolled image farther than n/2 whose value is below that of its own
  // cell at the distance n - d, so it never raises a non-negative maximum.
  const int nl = lane + 1 < P1 ? lane + 1 : 0;
  const int n_best = __shfl(l1.best, nl, 64), n_hi0 = __shfl(l1.hi, nl, 64);
  out[0] = l1.best;
#pragma unroll
  for (int q = 1; q < QMAX; ++q) out[q] = -1;
  if (act && Q > 1) {
    const int lo = l1.best >= 0 ? l1.lo : x1;
    const int hi = n_best >= 0 ? n_hi0 + (lane + 1 < P1 ? 0 : n) : x1 + Q;
    for (int y = lo; y <= hi; ++y) {
      int v = f(y);
      v = v < 0 ? -(1 << 28) : v;
#pragma unroll
      for (int q = 1; q < QMAX; ++q) {
        const int d = x1 + q - y;
        out[q] = max(out[q], v - d * d);
      }
    }
  }
}

// Coverage only (g(x) >= 0) of one line held in LDS, for the last axis.
// Cell x is covered iff some y with f(y) >= 0 lies within r(y) = min(hcap,
// floor(sqrt(f(y)))) of it on the ring: d^2 <= f(y) <=> |d| <= floor(sqrt
// f(y)), and the transform above never looks past hl <= hcap.  As r <= n/2,
// the images of y that can reach x in [0, n) are y - n, y and y + n, so x is
// covered from the left iff max(max_{y <= x} (y + r), max_y (y + r) - n) >= x
// and from the right iff min(min_{y >= x} (y - r), min_y (y - r) + n) <= x:
// one prefix max and one suffix min over the line (in-lane over the lane's Q
// cells, then across the wave).  out[q] = 1 for covered cells x = Q*lane + q.
template <int QMAX>
__device__ __forceinline__ void cover_line_wave(const int* col, int n, int Q, int hcap, int lane,
                                                int (&out)[QMAX]) {
  constexpr int kNone = 1 << 29;
  const bool act = lane < n / Q;
  int rr[QMAX], ll[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    rr[q] = -kNone;
    ll[q] = kNone;
    if (act && q < Q) {
      const int y = Q * lane + q;
      const int v = col[y + (y >> 3)];
      if (v >= 0) {
        const int r = min(hcap, isqrt_floor(v));
        rr[q] = y + r;
        ll[q] = y - r;
      }
    }
  }
#pragma unroll
  for (int q = 1; q < QMAX; ++q) rr[q] = max(rr[q], rr[q - 1]);
#pragma unroll
  for (int q = QMAX - 2; q >= 0; --q) ll[q] = min(ll[q], ll[q + 1]);
  // inclusive scans of the lane totals: max from lane 0 up, min from lane 63 down
  int sr = rr[QMAX - 1], sl = ll[0];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int ur = __shfl_up(sr, o, 64);
    const int dl = __shfl_down(sl, o, 64);
    if (lane >= o) sr = max(sr, ur);
    if (lane + o < 64) sl = min(sl, dl);
  }
  const int rall = __shfl(sr, 63, 64), lall = __shfl(sl, 0, 64);
  int er = __shfl_up(sr, 1, 64), el = __shfl_down(sl, 1, 64);
  if (lane == 0) er = -kNone;
  if (lane == 63) el = kNone;
  const int cr = max(er, rall - n), cl = min(el, lall + n);
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    const int x = Q * lane + q;
    out[q] = (max(rr[q], cr) >= x || min(ll[q], cl) <= x) ? 1 : 0;
  }
}

constexpr int kDtCols = 16;  // columns per block in the strided passes

__host__ __device__ constexpr int dt_col_floats(int n) { return n + n / 8 + 1; }

// Strided pass: lines along i (axis 1: outer = j) or j (axis 0: outer = i)
// for a tile of kDtCols adjacent p columns; in place through LDS.
// QMAX >= a.q: the per-lane result array of the line pass.  The last tile
// of a row may be partial (n a multiple of 8): its missing columns load as
// empty (-1) lines and are not stored.
template <int QMAX>
__global__ __launch_bounds__(256) void k_dt_tile(DtArgs a, int axis, const int* in, int* out) {
  extern __shared__ int sm[];
  const int n = a.n, Q = a.q, cs = dt_col_floats(n);
  const int ntile = (n + kDtCols - 1) / kDtCols;
  const int outer = blockIdx.x / ntile, p0 = (blockIdx.x % ntile) * kDtCols;
  const int ncol = n - p0 < kDtCols ? n - p0 : kDtCols;
  long base, stride;
  if (axis == 0) {
    base = (long)outer * n * n + p0;
    stride = n;
  } else {
    base = (long)outer * n + p0;
    stride = (long)n * n;
  }
  for (int t = threadIdx.x; t < n * kDtCols; t += blockDim.x) {
    const int e = t / kDtCols, c = t % kDtCols;
    sm[c * cs + e + (e >> 3)] = c < ncol ? in[base + e * stride + c] : -1;
  }
  lds_pin();
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = wave; c < kDtCols; c += blockDim.x >> 6) {
    int* col = sm + c * cs;
    int o[QMAX];
    dt_line_wave<QMAX>(col, n, Q, a.h, lane, o);
    if (lane < n / Q)
      for (int q = 0; q < Q; ++q) {
        const int e = Q * lane + q;
        col[e + (e >> 3)] = o[q];
      }
  }
  lds_pin();
  __syncthreads();
  for (int t = threadIdx.x; t < n * kDtCols; t += blockDim.x) {
    const int e = t / kDtCols, c = t % kDtCols;
    if (c < ncol) out[base + e * stride + c] = sm[c * cs + e + (e >> 3)];
  }
}

// Pass along p, one row per wave, fused with the final per-cell value
// (quick fill :277-288, union of balls, clamp :466-471) and FP64 partial
// sums of 1 - xHII (:473-477).
template <int QMAX>
__global__ __launch_bounds__(256) void k_dt_rows_final(DtArgs a, const int* in,
                                                       const float* __restrict__ pa,
                                                       const float* __restrict__ pb,
                                                       const float* __restrict__ pc, CellConsts k,
                                                       float* __restrict__ xhii,
                                                       double* __restrict__ partials) {
  extern __shared__ int sm[];
  const int n = a.n, Q = a.q, cs = dt_col_floats(n);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  int* col = sm + wave * cs;
  double acc = 0.0;
  const long rows = (long)n * n;
  // each wave walks its own rows through its own LDS line: only its lanes
  // exchange data there, and the LDS operations of one wave complete in
  // issue order, so compiler fences suffice and the waves run independently
  const long stride = (long)gridDim.x * nw;
  for (long row = (long)blockIdx.x * nw + wave; row < rows; row += stride) {
    const int* src = in + row * n;
    for (int e = lane; e < n; e += 64) col[e + (e >> 3)] = src[e];
    lds_pin();
    __builtin_amdgcn_wave_barrier();
    int o[QMAX];
    cover_line_wave<QMAX>(col, n, Q, a.h, lane, o);
    lds_pin();
    __builtin_amdgcn_wave_barrier();
    // the lane's Q consecutive flags go back into the wave's own line (its
    // reads are complete), then lane takes cells lane, lane + 64, ...:
    // coalesced loads of the fields and stores of x
    if (lane < n / Q)
      for (int q = 0; q < Q; ++q) {
        const int e = Q * lane + q;
        col[e + (e >> 3)] = o[q];
      }
    lds_pin();
    __builtin_amdgcn_wave_barrier();
    for (int e = lane; e < n; e += 64) {
      const long idx = row * n + e;
      float rec, ion, fres;
      cell_fields(k, pa, pb, pc, idx, rec, ion, fres);
      float v = col[e + (e >> 3)] ? 1.0f : quick_fill(rec, ion);
      v = fminf(v, 1.0f - fres);
      xhii[idx] = v;
      acc += 1.0 - (double)v;
    }
    lds_pin();
    __builtin_amdgcn_wave_barrier();
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

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials,
                                                      int np, double scale,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partials[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// ------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------
#define SF21_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(512)

int balance_partials_count(long n3) {
  // 2048 blocks: 8 waves per SIMD on 256 CUs, enough loads in flight
  long b = (n3 + 255) / 256;
  return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

void launch_balance(long n3, const float* a, const float* b, const float* c, CellConsts k,
                    double* part, float* scale, hipStream_t s) {
  const int nb = balance_partials_count(n3);
  hipLaunchKernelGGL(k_balance_partial, dim3(nb), dim3(256), 0, s, n3, a, b, c, k, part);
  hipLaunchKernelGGL(k_balance_final, dim3(1), dim3(256), 0, s, part, nb, scale);
}

void launch_fwd_rows(int n, const float* a, const float* b, const float* c, CellConsts k,
                     const float* bal, float2* y, const float2* tw, hipStream_t s) {
#define X(L)                                                                          \
  if (n == L) {                                                                       \
    constexpr int LPB = 256 / (L / Plan<L>::E);                                       \
    hipLaunchKernelGGL(k_fwd_rows<L>, dim3((L * L) / LPB), dim3(256), 0, s, a, b, c, k, bal, y, tw); \
    return;                                                                           \
  }
  SF21_SIZES(X)
#undef X
}

template <int L, int TILE, int E, bool NT>
static void cols_t(int width, int axis, int sign, const float2* in, float2* out, int nbatch,
                   long bstride, const float2* tw, hipStream_t s, bool pipe) {
  constexpr int NTH = TILE * (L / E);
  if (pipe) {
    constexpr int kPer = 8;
    const long ntiles = (long)L * ((width + TILE - 1) / TILE) * nbatch;
    const dim3 g((unsigned)((ntiles + kPer - 1) / kPer));
    if (sign < 0)
      hipLaunchKernelGGL((k_cols_pipe<L, -1, TILE, E, NT>), g, dim3(NTH), 0, s, in, out, width, axis,
                         bstride, tw, ntiles, kPer);
    else
      hipLaunchKernelGGL((k_cols_pipe<L, +1, TILE, E, NT>), g, dim3(NTH), 0, s, in, out, width, axis,
                         bstride, tw, ntiles, kPer);
    return;
  }
  dim3 g(L * ((width + TILE - 1) / TILE), nbatch);
  if (sign < 0)
    hipLaunchKernelGGL((k_cols<L, -1, TILE, E, NT>), g, dim3(NTH), 0, s, in, out, width, axis, bstride, tw);
  else
    hipLaunchKernelGGL((k_cols<L, +1, TILE, E, NT>), g, dim3(NTH), 0, s, in, out, width, axis, bstride, tw);
}

// tile: columns per block, 16 (default) or 8; variant bit 0: 16 values per
// thread (sizes 128-512; tile 16 or 32); bit 1: blocks walk 8 tiles with the
// next tile's loads issued before the current tile's transform (tile 16, or
// 32 with bit 0); bit 2: non-temporal loads and stores (whole streamed
// buffers with width a multiple of the tile; the per-radius passes)
void launch_cols(int n, int width, int axis, int sign, const float2* in, float2* out,
                 int nbatch, long bstride, const float2* tw, hipStream_t s, int tile, int variant) {
  const bool wide = variant & 1, pipe = (variant & 2) != 0, nt = (variant & 4) != 0;
  if ((variant & 8) && axis == 0 && width == n) axis = 2;  // the 16-column tile layout
#define X(L)                                                                              \
  if (n == L) {                                                                           \
    constexpr int E0 = Plan<L>::E;                                                        \
    if (tile == 8)                                                                        \
      cols_t<L, 8, E0, false>(width, axis, sign, in, out, nbatch, bstride, tw, s, false); \
    else if (nt)                                                                          \
      cols_t<L, 16, E0, true>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe);  \
    else                                                                                  \
      cols_t<L, 16, E0, false>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe); \
    return;                                                                               \
  }
#define XW(L)                                                                             \
  if (n == L && wide) {                                                                   \
    if (tile == 32 && nt)                                                                 \
      cols_t<L, 32, 16, true>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe);  \
    else if (tile == 32)                                                                  \
      cols_t<L, 32, 16, false>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe); \
    else if (nt)                                                                          \
      cols_t<L, 16, 16, true>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe);   \
    else                                                                                  \
      cols_t<L, 16, 16, false>(width, axis, sign, in, out, nbatch, bstride, tw, s, pipe);  \
    return;                                                                               \
  }
  XW(128) XW(256) XW(512)
  SF21_SIZES(X)
#undef XW
#undef X
}

void launch_filter_i(int n, const float2* spec, float2* u, int nrad, long ustride,
                     const float* wlut, long m_lut, const float2* kappa, const float2* tw,
                     hipStream_t s, int tile, int wmode, bool jpair, bool t16) {
  const int lay = t16 ? 1 : 0;
#define X(L, TL)                                                                          \
  if (n == L && tile == TL) {                                                             \
    constexpr int NT = TL * (L / Plan<L>::E);                                             \
    if (jpair && TL >= 8 && wmode != 0) {                                                 \
      /* mirror pairs (kj, L - kj), kj = 1 .. L/2 - 1, then the lines 0 and L/2 */        \
      const dim3 g((L / 2 - 1) * (L / TL)), g0(2 * (L / TL));                             \
      hipLaunchKernelGGL((k_filter_i<L, TL, false, true>), g, dim3(NT), 0, s, spec, u,    \
                         nrad, ustride, wlut, m_lut, kappa, tw, 0, lay);                  \
      hipLaunchKernelGGL((k_filter_i<L, TL, false>), g0, dim3(NT), 0, s, spec, u, nrad,   \
                         ustride, wlut, m_lut, kappa, tw, L / 2, lay);                    \
      return;                                                                             \
    }                                                                                     \
    if (wmode == 0)                                                                       \
      hipLaunchKernelGGL((k_filter_i<L, TL, true>), dim3(L * (L / TL)), dim3(NT), 0, s, spec, u, \
                         nrad, ustride, wlut, m_lut, kappa, tw, 1, lay);                  \
    else                                                                                  \
      hipLaunchKernelGGL((k_filter_i<L, TL, false>), dim3(L * (L / TL)), dim3(NT), 0, s, spec, u, \
                         nrad, ustride, wlut, m_lut, kappa, tw, 1, lay);                  \
    return;                                                                               \
  }
#define Y(L) X(L, 4) X(L, 8) X(L, 16)
  Y(32) Y(64) Y(96) Y(128) Y(192) Y(256) Y(512)
#undef Y
#undef X
}

void launch_inv_rows_crit(int n, const float2* v, int nrad, long vstride, const int* d2max,
                          const float* bal, int* wmap, float2* dbg, hipStream_t s,
                          const float2* tw, bool wide, float2* keep, bool t16) {
  const int lay = t16 ? 1 : 0;
#define XW(L)                                                                             \
  if (n == L && wide) {                                                                   \
    constexpr int LPB = 256 / (L / 16);                                                   \
    hipLaunchKernelGGL((k_inv_rows_crit<L, 16>), dim3((L * L) / LPB), dim3(256), 0, s, v, nrad, \
                       vstride, d2max, bal, wmap, dbg, keep, tw, lay);                          \
    return;                                                                               \
  }
#define X(L)                                                                              \
  if (n == L) {                                                                           \
    constexpr int LPB = 256 / (L / Plan<L>::E);                                           \
    hipLaunchKernelGGL(k_inv_rows_crit<L>, dim3((L * L) / LPB), dim3(256), 0, s, v, nrad, vstride, \
                       d2max, bal, wmap, dbg, keep, tw, lay);                                   \
    return;                                                                               \
  }
  XW(128) XW(256) XW(512)
  SF21_SIZES(X)
#undef XW
#undef X
}

void launch_fill_i32(int* p, int value, long count, hipStream_t s) {
  long blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, value, count);
}

int dt_partials_count(int n) {
  const long rows = (long)n * n;
  const long blocks = (rows + 3) / 4;
  return (int)(blocks < 2048 ? blocks : 2048);
}

int dt_q(int n) {
  // cells per lane: n/q lanes, a multiple of 8 and at most 64 (0: none)
  for (int q = 1; q <= 16; ++q)
    if (n % q == 0 && n / q <= 64 && (n / q) % 8 == 0) return q;
  return 0;
}

void launch_dt_rows(const DtArgs& a, const int* in, int* out, hipStream_t s) {
  // axis i: tiles (j, p-block)
  const size_t lds = sizeof(int) * kDtCols * dt_col_floats(a.n);
  const dim3 g((unsigned)(a.n * ((a.n + kDtCols - 1) / kDtCols)));
  if (a.q <= 8) hipLaunchKernelGGL(k_dt_tile<8>, g, dim3(256), lds, s, a, 1, in, out);
  else hipLaunchKernelGGL(k_dt_tile<16>, g, dim3(256), lds, s, a, 1, in, out);
}
void launch_dt_cols(const DtArgs& a, const int* in, int* out, hipStream_t s) {
  // axis j: tiles (i, p-block)
  const size_t lds = sizeof(int) * kDtCols * dt_col_floats(a.n);
  const dim3 g((unsigned)(a.n * ((a.n + kDtCols - 1) / kDtCols)));
  if (a.q <= 8) hipLaunchKernelGGL(k_dt_tile<8>, g, dim3(256), lds, s, a, 0, in, out);
  else hipLaunchKernelGGL(k_dt_tile<16>, g, dim3(256), lds, s, a, 0, in, out);
}
void launch_dt_final(const DtArgs& a, const int* in, const float* pa, const float* pb,
                     const float* pc, CellConsts k, float* xhii, double* partials,
                     int npartials, hipStream_t s) {
  const size_t lds = sizeof(int) * 4 * dt_col_floats(a.n);
  if (a.q <= 8)
    hipLaunchKernelGGL(k_dt_rows_final<8>, dim3(npartials), dim3(256), lds, s, a, in, pa, pb, pc, k,
                       xhii, partials);
  else
    hipLaunchKernelGGL(k_dt_rows_final<16>, dim3(npartials), dim3(256), lds, s, a, in, pa, pb, pc,
                       k, xhii, partials);
}
void launch_sum_partials(const double* partials, int np, double scale, double* out,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partials, np, scale, out);
}

}  // namespace sf21
