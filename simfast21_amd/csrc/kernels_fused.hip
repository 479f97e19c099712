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
#include <climits>

namespace sf21 {

// ------------------------------------------------------------------------
// Cell fields
// ------------------------------------------------------------------------

// rec, ion and residual neutral fraction of one cell.  SFR mode restates
// get_HIIbubbles.c:220-222 and :273 in FP32: Rrec (user_functions.c:58-65)
// as k_rec * (x/(1+x))^4 with x = (D/r2)^r3, D = fl32(1+delta); XHI
// (user_functions.c:101-106) in its cancellation-free form
// 2r/(2r+1+sqrt(4r+1)), algebraically identical to the reference's.
__device__ __forceinline__ void cell_fields(const CellConsts& k, const float* __restrict__ a,
                                            const float* __restrict__ b,
                                            const float* __restrict__ c, long i, float& rec,
                                            float& ion, float& fres) {
  if (k.mode == 0) {
    const float d = a[i];
    const float D = 1.0f + d;
    const float r = fmaf(d, k.c_fres, k.c_fres);
    fres = 2.0f * r / (2.0f * r + 1.0f + sqrtf(4.0f * r + 1.0f));
    const float x = powf(D / k.r2, k.r3);
    const float y = x / (1.0f + x);
    const float y2 = y * y;
    rec = k.k_rec * (y2 * y2);
    ion = b[i] * k.k_ion;
  } else {
    rec = a[i];
    ion = b[i];
    fres = c ? c[i] : 0.0f;
  }
}

// get_HIIbubbles.c:281-288: partial ionisation of a cell from its own
// emissivity and recombination rate.
__device__ __forceinline__ float quick_fill(float rec, float ion) {
  if (ion > 0.0f) {
    if (rec > 0.0f) {
      const float q = ion / rec;
      return q >= 1.0f ? 1.0f : q;
    }
    return 1.0f;
  }
  return 0.0f;
}

// get_HIIbubbles.c:339-349: centre iff ion > 0 and (rec <= 0 or
// ion/rec >= 1); for rec > 0 the FP64 quotient test equals ion >= rec.
__device__ __forceinline__ bool is_centre(float rec, float ion) {
  return (ion > 0.0f) && (!(rec > 0.0f) || ion >= rec);
}

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
// levels: every 8*Q-th x by an outward scan with early exit (stop once
// d^2 > fmax - best), every Q-th x inside the bracket of its level-0
// neighbours, the rest inside the bracket of their level-1 neighbours.  A
// neighbour xa that found no non-negative value bounds the search too: a
// y < xa with f(y) >= (x-y)^2 > (xa-y)^2 would contradict it.  Only values
// >= 0 are ever needed, so candidates farther than floor(sqrt(fmax)) are
// never visited.
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
__device__ __forceinline__ Found scan_out(int x, int hl, int fmax, F f) {
  Found r{-1, 0, 0};
  for (int d = 0; d <= hl; ++d) {
    if (r.best >= 0 && d * d > fmax - r.best) break;
    consider(x, x - d, f, r);
    if (d) consider(x, x + d, f, r);
  }
  return r;
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
  // level 0: lanes 0, 8, 16, ... (x = 8*Q*m)
  Found l0{-1, 0, 0};
  if (act && (lane & 7) == 0) l0 = scan_out(x1, hl, fmax, f);
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
  // level 2: between x1 and the next lane's level-1 point
  const int nl = lane + 1 < P1 ? lane + 1 : 0;
  const int n_best = __shfl(l1.best, nl, 64), n_hi0 = __shfl(l1.hi, nl, 64);
  out[0] = l1.best;
#pragma unroll
  for (int q = 1; q < QMAX; ++q) {
    out[q] = -1;
    if (act && q < Q) {
      const int lo = l1.best >= 0 ? l1.lo : x1;
      const int hi = n_best >= 0 ? n_hi0 + (lane + 1 < P1 ? 0 : n) : x1 + Q;
      out[q] = scan_range(x1 + q, lo, hi, hl, f).best;
    }
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
  // block-uniform trip count: every wave reaches both barriers each time
  for (long r0 = (long)blockIdx.x * nw; r0 < rows; r0 += (long)gridDim.x * nw) {
    const long row = r0 + wave;
    const bool valid = row < rows;
    if (valid) {
      const int* src = in + row * n;
      for (int e = lane; e < n; e += 64) col[e + (e >> 3)] = src[e];
    }
    lds_pin();
    __syncthreads();
    int o[QMAX];
    if (valid) dt_line_wave<QMAX>(col, n, Q, a.h, lane, o);
    if (valid) {
      // the lane's Q consecutive results go back into the wave's own column
      // (the transform's reads of it are complete), then lane takes cells
      // lane, lane + 64, ...: coalesced loads of the fields and stores of x
      if (lane < n / Q)
        for (int q = 0; q < Q; ++q) {
          const int e = Q * lane + q;
          col[e + (e >> 3)] = o[q] >= 0 ? 1 : 0;
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
    }
    lds_pin();
    __syncthreads();
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
// Generic sizes (box sides with no register plan, e.g. 600 = 8*3*5*5):
// whole-line mixed-radix Stockham transforms, 8 lines per 256-thread block.
// Stage of radix R and span NS (the product of the earlier radices),
// butterfly j < L/R with k = j mod NS: a_r = x[j + r L/R] w^(k r), w =
// exp(-+2 pi i / (NS R)); DFT_R; y[(j/NS) NS R + k + r NS] = a_r -- the
// self-sorting stages of the register transforms.  The first stage reads its
// inputs from HBM and the last writes its outputs there (positions j + r L/R
// in both, so runs of consecutive j stay coalesced); the stages in between
// exchange through one padded LDS line per line (a thread's butterflies are
// read into registers, barrier, written, barrier).  Lanes: rows (axis 0)
// take 32 consecutive butterflies of one line per half wave; the column
// passes (axes 1, 2) take the 8 lines -- adjacent p columns, 64 B runs --
// fastest.  32-bit indices (L^3 <= 2^30).
// ------------------------------------------------------------------------
constexpr int kGenLines = 8;
constexpr int kGenMaxL = 768;

// floor(a / d) for 0 <= a, d <= 2^11 from rd = 1/d rounded to float:
// (a + 1/2)/d lies at least 1/(2d) >= 2^-12 inside an integer interval and
// the product's rounding error is below 2^-12.
__device__ __forceinline__ int gen_div(int a, float rd) {
  return (int)(((float)a + 0.5f) * rd);
}

// twiddles w^(k r), r = 1..R-1, of a stage's butterfly
template <int R, int SIGN>
__device__ __forceinline__ void gen_twiddle(float2* a, int k, int tstep,
                                            const float2* __restrict__ tw) {
#pragma unroll
  for (int r = 1; r < R; ++r) {
    float2 w = tw[k * r * tstep];
    if (SIGN > 0) w.y = -w.y;
    a[r] = cmul(a[r], w);
  }
}

// First stage (radix 8, NS = 1): inputs from HBM (times W(|k| R) in mode 1).
// Output position o = 8 j + r is kept transposed, at slot r L/8 + j, so that
// the stores of consecutive butterflies are consecutive (a stride-8 store
// pattern conflicts in the LDS banks); the next stage reads position o at
// gen_t1_slot(o).
__device__ __forceinline__ int gen_t1_slot(int o, int lr0) { return (o & 7) * lr0 + (o >> 3); }

// LC (and NSC): the side (and the stage span) as compile-time constants for
// the specialised sides, 0 for the runtime plan.
template <int R, int SIGN, int MODE, int LC>
__device__ __forceinline__ void gen_first(const GenPlan& P, int jj, float2* line,
                                          const float2* src, unsigned base, unsigned stride,
                                          int msq, float2 kappa) {
  const int L = LC ? LC : P.L, Lr = L / R;
#pragma unroll 2
  for (int j = jj; j < Lr; j += 32) {
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = j + r * Lr;
      float2 v = src[base + (unsigned)e * stride];
      if constexpr (MODE == 1) {
        const float w = w_tophat_k(kappa, msq + fold_sq(e, L));
        v = make_float2(v.x * w, v.y * w);
      }
      a[r] = v;
    }
    Dft<R, SIGN>::run(a);
#pragma unroll
    for (int r = 0; r < R; ++r) line[r * Lr + j] = a[r];
  }
  lds_pin();
  __syncthreads();
}

// Filtered first stage (mode 1) of the i-axis pass: the W of the block's 8
// lines depends on |k|^2 = msq_l + fold(e)^2, so it is evaluated once per
// (line, |fold(e)|) -- 8 (L/2 + 1) values instead of 8 L -- into a table in
// the line buffers, which the stage's stores overwrite only after every
// thread has multiplied its inputs.  The same function of the same |k|^2 as
// the direct evaluation: identical values.  The first radix is 8, so a
// thread holds at most (768/8 + 31)/32 = 3 butterflies.
template <int SIGN, int LC>
__device__ __forceinline__ void gen_first_wtab(const GenPlan& P, float2* gsm, float2* line,
                                               int l, int jj, const float2* src, unsigned base,
                                               unsigned stride, int mfix, int p0, float2 kappa) {
  constexpr int R = 8, MAXIT = ((LC ? LC : kGenMaxL) / R + 31) / 32;
  const int L = LC ? LC : P.L, Lr = L / R, H = L / 2 + 1;
  float* wt = reinterpret_cast<float*>(gsm);
  for (int t = threadIdx.x; t < kGenLines * H; t += blockDim.x) {
    const int ll = t / H, f = t - ll * H;
    wt[t] = w_tophat_k(kappa, mfix + fold_sq(p0 + ll, L) + f * f);
  }
  lds_pin();
  __syncthreads();
  float2 a[MAXIT][R];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int j = jj + 32 * it;
    if (j < Lr) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = j + r * Lr;
        const float2 v = src[base + (unsigned)e * stride];
        const float w = wt[l * H + (e <= L / 2 ? e : L - e)];
        a[it][r] = make_float2(v.x * w, v.y * w);
      }
    }
  }
  lds_pin();
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int j = jj + 32 * it;
    if (j < Lr) {
      Dft<R, SIGN>::run(a[it]);
#pragma unroll
      for (int r = 0; r < R; ++r) line[r * Lr + j] = a[it][r];
    }
  }
  lds_pin();
  __syncthreads();
}

// Middle stage: all of a thread's butterflies are read before the barrier,
// then written back into the same line.  T1: the stage after the first
// (inputs in the transposed slots).
template <int R, int SIGN, bool T1, int LC, int NSC>
__device__ __forceinline__ void gen_mid(const GenPlan& P, int st, int jj, float2* line,
                                        const float2* __restrict__ tw) {
  constexpr int MAXIT = ((LC ? LC : kGenMaxL) / R + 31) / 32;
  const int L = LC ? LC : P.L, NS = LC ? NSC : P.ns[st], Lr = L / R, tstep = L / (NS * R);
  const float rns = P.rns[st];
  auto quot = [&](int j) { return LC ? (int)((unsigned)j / (unsigned)NSC) : gen_div(j, rns); };
  float2 a[MAXIT][R];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int j = jj + 32 * it;
    if (j < Lr) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int o = j + r * Lr;
        a[it][r] = line[T1 ? gen_t1_slot(o, L >> 3) : o];
      }
      gen_twiddle<R, SIGN>(a[it], j - quot(j) * NS, tstep, tw);
      Dft<R, SIGN>::run(a[it]);
    }
  }
  lds_pin();
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int j = jj + 32 * it;
    if (j < Lr) {
      const int q = quot(j), k = j - q * NS;
      float2* y = line + q * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) y[r * NS] = a[it][r];
    }
  }
  lds_pin();
  __syncthreads();
}

// Last stage (NS R = L, k = j): outputs at j + r L/R to HBM (mode 0), or in
// mode 2 the centre test (dst null) or the unscaled pair to dst (the
// inspection hook: one kernel for both, so that the hook sees the values the
// centre test saw, bit for bit).
template <int R, int SIGN, int MODE, bool T1, int LC, int NSC>
__device__ __forceinline__ void gen_last(const GenPlan& P, int st, int jj, const float2* line,
                                         float2* dst, unsigned base, unsigned stride,
                                         const float2* __restrict__ tw, float sc,
                                         int* __restrict__ wmap, int d2) {
  const int L = LC ? LC : P.L, NS = LC ? NSC : P.ns[st], Lr = L / R, tstep = L / (NS * R);
#pragma unroll 2
  for (int j = jj; j < Lr; j += 32) {
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = j + r * Lr;
      a[r] = line[T1 ? gen_t1_slot(o, L >> 3) : o];
    }
    gen_twiddle<R, SIGN>(a, j, tstep, tw);
    Dft<R, SIGN>::run(a);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const unsigned idx = base + (unsigned)(j + r * Lr) * stride;
      const float2 v = a[r];
      if constexpr (MODE == 2) {
        if (dst)
          dst[idx] = make_float2(v.x, v.y / sc);
        else if (is_centre(v.x * sc, v.y))
          wmap[idx] = d2;
      } else {
        dst[idx] = v;
      }
    }
  }
}

// 4 waves per SIMD (128 VGPRs): the LDS line buffers allow 4 blocks per CU
// at 600, and the bound (a few spilled registers in the column variants) ran
// 5% faster than 130 VGPRs at 3 waves
// LC = 600: the sample configuration's side with its plan 8, 3, 5, 5 (spans
// 1, 8, 24, 120) and LDS line stride 616 as constants; 0: the runtime plan.
template <int SIGN, int MODE, bool ROWS, int LC>
__global__ __launch_bounds__(256, 4) void k_gen_pass(const float2* src, float2* dst, GenPlan plan,
                                                  int axis, const float2* __restrict__ tw,
                                                  float2 kappa, const float* __restrict__ bal,
                                                  int* __restrict__ wmap, int d2) {
  extern __shared__ float2 gsm[];
  const unsigned L = LC ? (unsigned)LC : (unsigned)plan.L, L2 = L * L;
  // i-axis passes: consecutive tiles (which share 128 B lines) on one XCD --
  // workgroups are dealt round-robin over the 8 XCDs, and the grid (L^2/8
  // blocks, L a multiple of 8) is a multiple of 8.  Measured at 600: the
  // filtered i pass 1.30 -> 1.15 ms; the j pass 0.80 -> 0.81 ms, so it keeps
  // the plain order.
  const unsigned blk = (!ROWS && axis == 2)
                           ? (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3)
                           : blockIdx.x;
  const int l = ROWS ? (int)(threadIdx.x >> 5) : (int)(threadIdx.x & 7);
  const int jj = ROWS ? (int)(threadIdx.x & 31) : (int)(threadIdx.x >> 3);
  unsigned base, stride;
  int msq = 0;  // |k|^2 of the two coordinates fixed along the line
  if (ROWS) {
    const unsigned row = blk * kGenLines + l;
    base = row * L;
    stride = 1;
    if (MODE == 1) msq = fold_sq((int)(row / L), (int)L) + fold_sq((int)(row % L), (int)L);
  } else {
    const unsigned ntile = L / kGenLines;
    const unsigned outer = blk / ntile;
    const unsigned p = (blk % ntile) * kGenLines + l;
    base = axis == 1 ? outer * L2 + p : outer * L + p;
    stride = axis == 1 ? L : L2;
    if (MODE == 1) msq = fold_sq((int)outer, (int)L) + fold_sq((int)p, (int)L);
  }
  const float sc = MODE >= 2 ? bal[0] : 1.0f;
  float2* line = gsm + l * (LC == 600 ? 616 : plan.lp);
  // the first radix is 8 on every generic side (multiples of 8)
  if constexpr (MODE == 1 && !ROWS) {
    // the filtered pass runs along i only (axis 2: outer = j)
    const unsigned ntile = L / kGenLines;
    gen_first_wtab<SIGN, LC>(plan, gsm, line, l, jj, src, base, stride,
                             fold_sq((int)(blk / ntile), (int)L),
                             (int)((blk % ntile) * kGenLines), kappa);
  } else {
    gen_first<8, SIGN, MODE, LC>(plan, jj, line, src, base, stride, msq, kappa);
  }
  if constexpr (LC == 600) {
    gen_mid<3, SIGN, true, 600, 8>(plan, 1, jj, line, tw);
    gen_mid<5, SIGN, false, 600, 24>(plan, 2, jj, line, tw);
    gen_last<5, SIGN, MODE, false, 600, 120>(plan, 3, jj, line, dst, base, stride, tw, sc, wmap,
                                             d2);
    return;
  }
  const int last = plan.nst - 1;
#define GMS(R, T) gen_mid<R, SIGN, T, 0, 0>(plan, st, jj, line, tw)
#define GLS(R, T) \
  gen_last<R, SIGN, MODE, T, 0, 0>(plan, last, jj, line, dst, base, stride, tw, sc, wmap, d2)
  if (last == 1) {
    switch (plan.radix[1]) {
      case 8: GLS(8, true); break;
      case 4: GLS(4, true); break;
      case 2: GLS(2, true); break;
      case 3: GLS(3, true); break;
      default: GLS(5, true); break;
    }
    return;
  }
  {
    const int st = 1;
    switch (plan.radix[1]) {
      case 8: GMS(8, true); break;
      case 4: GMS(4, true); break;
      case 2: GMS(2, true); break;
      case 3: GMS(3, true); break;
      default: GMS(5, true); break;
    }
  }
  for (int st = 2; st < last; ++st) {
    switch (plan.radix[st]) {
      case 8: GMS(8, false); break;
      case 4: GMS(4, false); break;
      case 2: GMS(2, false); break;
      case 3: GMS(3, false); break;
      default: GMS(5, false); break;
    }
  }
  switch (plan.radix[last]) {
    case 8: GLS(8, false); break;
    case 4: GLS(4, false); break;
    case 2: GLS(2, false); break;
    case 3: GLS(3, false); break;
    default: GLS(5, false); break;
  }
#undef GMS
#undef GLS
}

__global__ void k_gen_pack(long n3, const float* __restrict__ a, const float* __restrict__ b,
                           const float* __restrict__ c, CellConsts k,
                           const float* __restrict__ bal, float2* __restrict__ spec) {
  const float sc = bal[0];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
       i += (long)gridDim.x * blockDim.x) {
    float rec, ion, fr;
    cell_fields(k, a, b, c, i, rec, ion, fr);
    spec[i] = make_float2(rec, ion * sc);
  }
}

// ------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------
#define SF21_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(512)

int balance_partials_count(long n3) {
  long b = (n3 + 255) / 256;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
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
    else                                                                                  \
      cols_t<L, 16, 16, false>(width, axis, sign, in, out, nbatch, bstride, tw, s, false); \
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

bool gen_plan_make(int n, GenPlan* p) {
  if (n < 16 || n > kGenMaxL || n % kGenLines != 0) return false;
  int m = n, ns = 1, st = 0;
  while (m > 1) {
    int r;
    if (m % 8 == 0) r = 8;
    else if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else return false;
    if (st >= 16) return false;
    p->radix[st] = r;
    p->ns[st] = ns;
    p->rns[st] = 1.0f / (float)ns;
    ns *= r;
    m /= r;
    ++st;
  }
  p->L = n;
  p->nst = st;
  if (p->radix[0] != 8 || st < 2) return false;  // the kernels' first stage is radix 8
  // line stride = 8 (mod 32) float2: the 8 lines of a column-pass wave
  // start in different bank groups
  p->lp = n + ((8 - n % 32) + 32) % 32;
  return true;
}

size_t gen_pass_lds(const GenPlan& p) { return sizeof(float2) * kGenLines * (size_t)p.lp; }

#define GEN_KERNELS_L(X, LC)                                                               \
  X(-1, 0, true, LC) X(-1, 0, false, LC) X(+1, 0, true, LC) X(+1, 0, false, LC)             \
  X(+1, 1, true, LC) X(+1, 1, false, LC) X(+1, 2, true, LC) X(+1, 2, false, LC)
#define GEN_KERNELS(X) GEN_KERNELS_L(X, 0) GEN_KERNELS_L(X, 600)

bool gen_prepare(const GenPlan& p) {
  // LDS of the transform blocks above 64 KiB is requested per kernel
  const size_t lds = gen_pass_lds(p);
  if (lds <= 65536) return true;
  const int v = (int)lds;
  const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
  bool ok = true;
#define GA(SG, MD, RW, LC) \
  ok = ok && hipFuncSetAttribute((const void*)k_gen_pass<SG, MD, RW, LC>, at, v) == hipSuccess;
  GEN_KERNELS(GA)
#undef GA
  return ok;
}

void launch_gen_pass(const GenPlan& plan, int axis, int sign, int mode, const float2* src,
                     float2* dst, const float2* tw, float2 kappa, const float* bal, int* wmap,
                     int d2, hipStream_t s) {
  const unsigned L = (unsigned)plan.L;
  const dim3 g(L * L / kGenLines), b(256);
  const size_t lds = gen_pass_lds(plan);
  const bool rows = axis == 0;
  // mode 3 (the unscaled pair) runs the centre-test kernel with dst set
  const int sg = sign < 0 ? -1 : +1, md = sign < 0 ? 0 : (mode == 3 ? 2 : mode);
  // the specialised side's plan is the one gen_plan_make builds for it
  const int lc = (plan.L == 600 && plan.nst == 4 && plan.lp == 616) ? 600 : 0;
#define GL(SG, MD, RW, LC)                                                                     \
  if (sg == SG && md == MD && rows == RW && lc == LC)                                          \
    hipLaunchKernelGGL((k_gen_pass<SG, MD, RW, LC>), g, b, lds, s, src, dst, plan, axis, tw,    \
                       kappa, bal, wmap, d2);
  GEN_KERNELS(GL)
#undef GL
}

void launch_gen_pack(long n3, const float* a, const float* b, const float* c, CellConsts k,
                     const float* bal, float2* spec, hipStream_t s) {
  long blocks = (n3 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_gen_pack, dim3((unsigned)blocks), dim3(256), 0, s, n3, a, b, c, k, bal, spec);
}

}  // namespace sf21
