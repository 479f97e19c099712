// Device building blocks of the CDNA4 FFT used by the bubble finder.
//
// A line of length L is spread over TPL = L/E threads, each holding E
// complex values in registers.  Stages are self-sorting Stockham passes
// (radix R, span NS = product of the radices already applied).  For
// butterfly j in [0, L/R):
//     k   = j mod NS
//     a_r = x[j + r*L/R] * exp(sign*2*pi*i*k*r/(NS*R))
//     a   = DFT_R(a)
//     y[(j/NS)*NS*R + k + r*NS] = a_r
// Thread t owns butterflies j = t + TPL*b (b < E/R), which makes its input
// slots "element t + TPL*m" for every stage; the last stage's outputs land
// in the same slots, so loads and stores of a whole line are both in the
// pattern element(t + TPL*m) -> coalesced for contiguous lines and, with
// lanes mapped across columns, for strided ones.  Only the exchanges
// between stages go through LDS (split re/im arrays, padded against bank
// conflicts).
#pragma once
#include <hip/hip_runtime.h>

namespace sf21 {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by sign*i
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// In-register DFTs, y_k = sum_n x_n exp(SIGN*2*pi*i*n*k/R).
template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  }
};

template <int SIGN>
struct Dft<4, SIGN> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = mul_si<SIGN>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int SIGN>
struct Dft<8, SIGN> {
  static __device__ __forceinline__ void run(float2* a) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, SIGN>::run(e);
    Dft<4, SIGN>::run(o);
    // o[k] *= exp(SIGN*2*pi*i*k/8)
    float2 o1 = o[1], o3 = o[3];
    o[1] = make_float2(h * (o1.x - SIGN * o1.y), h * (o1.y + SIGN * o1.x));
    o[2] = mul_si<SIGN>(o[2]);
    o[3] = make_float2(h * (-o3.x - SIGN * o3.y), h * (-o3.y + SIGN * o3.x));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 4] = csub(e[k], o[k]);
    }
  }
};

template <int SIGN>
struct Dft<3, SIGN> {
  static __device__ __forceinline__ void run(float2* a) {
    const float c = -0.5f, s = SIGN * 0.86602540378443864676f;
    float2 t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x + c * t1.x, a[0].y + c * t1.y);
    float2 r = make_float2(-s * t2.y, s * t2.x);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int SIGN>
struct Dft<5, SIGN> {
  static __device__ __forceinline__ void run(float2* a) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = SIGN * 0.95105651629515357212f, s2 = SIGN * 0.58778525229247312917f;
    float2 b1 = cadd(a[1], a[4]), b4 = csub(a[1], a[4]);
    float2 b2 = cadd(a[2], a[3]), b3 = csub(a[2], a[3]);
    float2 m1 = make_float2(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
    float2 m2 = make_float2(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
    // i*(s1*b4 + s2*b3) and i*(s2*b4 - s1*b3)
    float2 n1 = make_float2(-(s1 * b4.y + s2 * b3.y), s1 * b4.x + s2 * b3.x);
    float2 n2 = make_float2(-(s2 * b4.y - s1 * b3.y), s2 * b4.x - s1 * b3.x);
    a[0] = cadd(a[0], cadd(b1, b2));
    a[1] = cadd(m1, n1);
    a[4] = csub(m1, n1);
    a[2] = cadd(m2, n2);
    a[3] = csub(m2, n2);
  }
};

// Streaming accesses of the per-radius fields: each element is read or
// written once per pass and not again before the next pass has streamed the
// whole (K n^3) buffer, so the accesses carry the non-temporal hint.
typedef float sf21_v2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  const sf21_v2 v = __builtin_nontemporal_load(reinterpret_cast<const sf21_v2*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float2* p, float2 a) {
  sf21_v2 v;
  v.x = a.x;
  v.y = a.y;
  __builtin_nontemporal_store(v, reinterpret_cast<sf21_v2*>(p));
}

// Compiler-level fence for the LDS exchanges.  __syncthreads() orders the
// hardware, but the compiler was observed to sink a stage's LDS reads past
// the following barrier into a later basic block (k_filter_i: the forward
// transform's last reads landed after the barrier, racing with the inverse
// transform's first LDS writes of another wave).  An empty asm with a memory
// clobber pins every LDS access of an exchange on its side of the barrier.
__device__ __forceinline__ void lds_pin() { asm volatile("" ::: "memory"); }

// Returns v unchanged but opaque to the optimiser.  Used on the lane index
// at the top of radius loops: without it the compiler hoists every FFT
// stage's LDS and twiddle addresses out of the loop and keeps them live in
// VGPRs (124+ VGPRs, spills at 1024-thread blocks); recomputing them costs a
// few VALU per iteration and frees half the register file.
__device__ __forceinline__ int opaque_lane(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// LDS index padding: one slot every 8 breaks the stride-R write pattern of
// the first stage.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 3); }
__device__ __forceinline__ unsigned lds_pad_u(unsigned i) { return i + (i >> 3); }
template <int L>
struct LdsLine {
  static constexpr int kFloats = L + L / 8 + 1;  // odd stride between lines
};

// Radix plans.  E = values per thread; the radices all divide E.
template <int L>
struct Plan;
template <> struct Plan<32>  { static constexpr int E = 8; };
template <> struct Plan<48>  { static constexpr int E = 12; };
template <> struct Plan<64>  { static constexpr int E = 8; };
template <> struct Plan<96>  { static constexpr int E = 12; };
template <> struct Plan<128> { static constexpr int E = 8; };
template <> struct Plan<192> { static constexpr int E = 12; };
template <> struct Plan<256> { static constexpr int E = 8; };
template <> struct Plan<512> { static constexpr int E = 8; };

template <int... Rs>
struct RadixList {};

template <int L> struct Radices;
template <> struct Radices<32>  { using type = RadixList<8, 4>; };
template <> struct Radices<48>  { using type = RadixList<4, 4, 3>; };
template <> struct Radices<64>  { using type = RadixList<8, 8>; };
template <> struct Radices<96>  { using type = RadixList<4, 4, 2, 3>; };
template <> struct Radices<128> { using type = RadixList<8, 8, 2>; };
template <> struct Radices<192> { using type = RadixList<4, 4, 4, 3>; };
template <> struct Radices<256> { using type = RadixList<8, 8, 4>; };
template <> struct Radices<512> { using type = RadixList<8, 8, 8>; };

// LDS slot offset of r*NS past a write base: the padding of base + r*NS
// equals pad(base) + r*NS + (r*NS >> 3) whenever the low three bits of base
// cannot carry into the next multiple of 8 -- NS a multiple of 8, or NS | 8
// with base = k (mod 8), k < NS (the base's other term a multiple of 8).
template <int NS, int R>
struct StageSlots {
  static constexpr bool kConst = (NS % 8 == 0) || ((NS * R) % 8 == 0 && 8 % NS == 0);
  static constexpr int off(int r) { return r * NS + ((r * NS) >> 3); }
};

// One Stockham stage; recurses into the remaining radices.
// tw: table exp(-2*pi*i*q/L), q < L.  sre/sim: this line's LDS region.
// Index arithmetic is unsigned (lanes are non-negative), and where the
// thread's butterflies b share k = t mod NS (TPL a multiple of NS) the LDS
// slots of every b and r are one per-thread base plus compile-time offsets.
template <int L, int E, int SIGN, int NS, int R, int... REST>
__device__ __forceinline__ void stages(float2 (&v)[E], float* sre, float* sim, int t,
                                       const float2* __restrict__ tw) {
  constexpr int TPL = L / E;
  constexpr int NB = E / R;
  constexpr bool LAST = sizeof...(REST) == 0;
  static_assert(E % R == 0, "radix must divide the per-thread count");
  constexpr bool kUniformK = TPL % NS == 0;
  constexpr bool kFastW = kUniformK && StageSlots<NS, R>::kConst && (TPL * R) % 8 == 0;
  const unsigned ut = (unsigned)t;
  const unsigned k0 = ut % NS;
  const unsigned pbase0 = lds_pad_u((ut / NS) * NS * R + k0);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const unsigned j = ut + TPL * b;
    const unsigned k = kUniformK ? k0 : j % NS;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[b + r * NB];
    if constexpr (NS > 1) {
      const unsigned ks = k * (unsigned)(L / (NS * R));
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[ks * r];
        if (SIGN > 0) w.y = -w.y;
        a[r] = cmul(a[r], w);
      }
    }
    Dft<R, SIGN>::run(a);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = a[r];
    } else if constexpr (kFastW) {
      const unsigned pb = pbase0 + (unsigned)(b * (TPL * R + (TPL * R) / 8));
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned o = pb + StageSlots<NS, R>::off(r);
        sre[o] = a[r].x;
        sim[o] = a[r].y;
      }
    } else {
      const unsigned base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const unsigned o = lds_pad_u(base + r * NS);
        sre[o] = a[r].x;
        sim[o] = a[r].y;
      }
    }
  }
  if constexpr (!LAST) {
    lds_pin();
    __syncthreads();
    const unsigned pt = lds_pad_u(ut);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned o = (TPL % 8 == 0) ? pt + (unsigned)(m * (TPL + TPL / 8)) : lds_pad_u(ut + TPL * m);
      v[m] = make_float2(sre[o], sim[o]);
    }
    lds_pin();
    __syncthreads();
    stages<L, E, SIGN, NS * R, REST...>(v, sre, sim, t, tw);
  }
}

template <int L, int E, int SIGN, int... Rs>
__device__ __forceinline__ void line_fft_impl(float2 (&v)[E], float* sre, float* sim, int t,
                                              const float2* __restrict__ tw, RadixList<Rs...>) {
  stages<L, E, SIGN, 1, Rs...>(v, sre, sim, t, tw);
}

// Line transform with E values per thread (every radix of the size's plan
// must divide E): E = 16 halves the threads per line and doubles the work
// between the exchanges' barriers.
template <int L, int E, int SIGN>
__device__ __forceinline__ void line_fft_e(float2 (&v)[E], float* sre, float* sim, int t,
                                           const float2* __restrict__ tw) {
  line_fft_impl<L, E, SIGN>(v, sre, sim, t, tw, typename Radices<L>::type{});
}

// Full line transform.  Every thread of the block must call it (it contains
// block barriers), including threads whose line is out of range.
template <int L, int SIGN>
__device__ __forceinline__ void line_fft(float2 (&v)[Plan<L>::E], float* sre, float* sim, int t,
                                         const float2* __restrict__ tw) {
  line_fft_impl<L, Plan<L>::E, SIGN>(v, sre, sim, t, tw, typename Radices<L>::type{});
}

// Signed frequency index as the reference folds it (get_HIIbubbles.c:309-315):
// i > n/2 -> i - n; i = n/2 stays positive.  Only its square is used.
__device__ __forceinline__ int fold_sq(int i, int n) {
  int f = i > (n >> 1) ? i - n : i;
  return f * f;
}

// Top-hat window W(x) = 3 j1(x)/x = 3 (sin x - x cos x)/x^3, x = |k| R
// (auxiliary.c:350-359), in FP32 with a phase accurate to ~1e-7 absolute:
//   - x = kappa sqrt(m), kappa = dk R as a float pair (hi, lo) rounded from
//     FP64, m = integer |k|^2 in dk units: kept as a double-float p + x_lo
//     (Dekker's exact product via fmaf; sqrt(m) completed by its exact
//     residual m - s^2, so the hardware square root's last bit is
//     irrelevant);
//   - below x = 1.25: the series 3 sum_k (-1)^(k+1) 2k x^(2k-2)/(2k+1)! to
//     x^12 (truncation < 1e-11), where the closed form cancels;
//   - above: x - n pi/2 with a three-part Cody-Waite constant (12-bit first
//     parts: n*C1 and n*C2 exact for n < 4096, i.e. x < 6400), then Taylor
//     sin to r^9 and cos to r^10 on |r| <= pi/4 (truncation < 2e-9) and the
//     quadrant n mod 4.
// A plain FP32 x = kappa*sqrtf(m) carries ~x*6e-8 of phase error, i.e. a
// relative W error of ~5e-5 at x ~ 700 (R = 100 Mpc at 512^3); the library
// sincosf would also carry its large-argument reduction code.
// sqrt(m) as s + s_lo (s_lo from the exact residual m - s^2): independent of
// the radius, computed once per element.
struct SqrtPair {
  float s, s_lo;
};
__device__ __forceinline__ SqrtPair sqrt_pair(int m) {
  const float fm = (float)m;
  const float s = __builtin_amdgcn_sqrtf(fm);
  return SqrtPair{s, fmaf(-s, s, fm) * (0.5f * __builtin_amdgcn_rcpf(s))};
}

__device__ __forceinline__ float w_tophat_sq(float2 kap, SqrtPair sq, int m) {
  const float p = kap.x * sq.s;
  const float x_lo = fmaf(kap.x, sq.s, -p) + fmaf(kap.x, sq.s_lo, kap.y * sq.s);
  const float x = p + x_lo;
  const float nq = rintf(p * 0.6366197466850281f);
  float r = fmaf(-nq, 1.5703125f, p);
  r = fmaf(-nq, 4.837512969970703e-4f, r);
  r = fmaf(-nq, 7.549790126404332e-8f, r);
  r += x_lo;
  const float r2 = r * r;
  const float sr = r * (1.0f + r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f +
                   r2 * (1.0f / 362880.0f)))));
  const float cr = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f +
                   r2 * (1.0f / 40320.0f + r2 * (-1.0f / 3628800.0f)))));
  const int qd = (int)nq & 3;
  const float sn = (qd & 1) ? cr : sr, cn = (qd & 1) ? sr : cr;
  const float sx = (qd & 2) ? -sn : sn;                   // sin(r + qd pi/2)
  const float cx = ((qd + 1) & 2) ? -cn : cn;             // cos(r + qd pi/2)
  const float inv = __builtin_amdgcn_rcpf(x);
  float w = 3.0f * (sx - x * cx) * (inv * inv * inv);
  if (x < 1.25f) {  // rare (low |k|): skipped by whole waves
    const float q = x * x;
    w = 1.0f + q * (-0.1f + q * (1.0f / 280.0f + q * (-1.0f / 15120.0f +
        q * (1.0f / 1330560.0f + q * (-1.0f / 172972800.0f + q * (1.0f / 31135104000.0f))))));
  }
  return m == 0 ? 1.0f : w;
}

__device__ __forceinline__ float w_tophat_k(float2 kap, int m) {
  return w_tophat_sq(kap, sqrt_pair(m), m);
}

// Lanes of the i-axis passes.  W depends on |k|^2 only, so the line
// positions k and L - k share it.  Thread slot q (threadIdx.x / TILE) holds
// line position t with the pairs (t, TPL - t) in adjacent slots -- q = 2j +
// i: j = 0 -> t = 0, TPL/2 (their mirrors are their own); j >= 1 -> t = j,
// TPL - j -- so a thread's mirror partner is lane ^ TILE of the same wave.
// Memory coalescing is unchanged (columns stay the fastest lane index).
__device__ __forceinline__ int paired_slot_to_t(int q, int tpl) {
  const int j = q >> 1, i = q & 1;
  return j == 0 ? (i ? tpl / 2 : 0) : (i ? tpl - j : j);
}

// W for the E positions t + TPL m of a thread, evaluating only the lower
// half (m <= E/2) and taking the upper half from the mirror partner:
// position t + TPL m mirrors to (TPL - t) + TPL (E - 1 - m) for 0 < t < TPL,
// to TPL (E - m) for t = 0 and to TPL/2 + TPL (E - 1 - m) for t = TPL/2.
template <int L, int E, int TILE>
__device__ __forceinline__ void w_line_shared(float2 kap, const SqrtPair (&sq)[E / 2 + 1],
                                              const int (&mm)[E / 2 + 1], int t,
                                              float (&wv)[E]) {
  constexpr int TPL = L / E;
  float wl[E / 2 + 1];
#pragma unroll
  for (int m = 0; m <= E / 2; ++m) wl[m] = w_tophat_sq(kap, sq[m], mm[m]);
#pragma unroll
  for (int m = 0; m < E / 2; ++m) wv[m] = wl[m];
#pragma unroll
  for (int m = E / 2; m < E; ++m) {
    const float other = __shfl_xor(wl[E - 1 - m], TILE);
    const float own0 = (m == E / 2) ? wl[E / 2] : wl[(E - m) < E / 2 ? (E - m) : 0];
    wv[m] = t == 0 ? own0 : (t == TPL / 2 ? wl[E - 1 - m] : other);
  }
}

}  // namespace sf21
