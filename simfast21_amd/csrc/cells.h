// Per-cell device helpers shared by the fused, generic-side and plain
// paths: the cell pre-pass, the quick fill and the centre test.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

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

}  // namespace sf21
