// Fused finder on box sides without a register transform plan: whole-line
// mixed-radix Stockham transforms (first and last stage in HBM, the others
// through LDS), the cell pre-pass packing, and their launchers.
#include "fft_core.h"
#include "kernels.h"
#include "cells.h"

namespace sf21 {

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
  // column passes with bit `axis` of plan.xcd: consecutive tiles (which
  // share 128 B lines) on one XCD -- workgroups are dealt round-robin over
  // the 8 XCDs, and the grid (L^2/8 blocks, L a multiple of 8) is a multiple
  // of 8, so the remap is a permutation.  Measured at 600: the filtered i
  // pass 1.30 -> 1.15 ms, the j pass 0.80 -> 0.81 ms (default: i passes
  // only; the autotuner searches the four masks).
  const unsigned blk = (!ROWS && ((plan.xcd >> axis) & 1))
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
  p->xcd = 1 << 2;
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
