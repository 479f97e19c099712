// Kernel declarations and launch helpers shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sf21 {

// Per-slice constants of the SFR-mode cell pre-pass (get_HIIbubbles.c:216-225,
// :269-274), evaluated on the host in FP64 and rounded to float.
struct CellConsts {
  int mode;         // 0: a = delta, b = SFRD ; 1: a = rec, b = ion, c = fresid (nullable)
  float c_fres;     // 1.881e-7 (1+z)^3 G_H(z)
  float k_rec;      // 1e-24 r1 (1+z)^r4 DVV / V_norm
  float r2, r3;     // Rrec fit constants
  float k_ion;      // dx^3 Qion(z) fesc
};

struct DtArgs {
  int n;  // box side
  int h;  // reach cap: min(floor(sqrt(max weight)), n/2)
};

// ---- fused exact path ----
int balance_partials_count(long n3);
void launch_balance(long n3, const float* a, const float* b, const float* c, CellConsts k,
                    double* part, float* scale, hipStream_t s);
void launch_fwd_rows(int n, const float* a, const float* b, const float* c, CellConsts k,
                     const float* bal, float2* y, const float2* tw, hipStream_t s);
void launch_cols(int n, int width, int axis, int sign, const float2* in, float2* out,
                 int nbatch, long bstride, const float2* tw, hipStream_t s);
void launch_filter_i(int n, const float2* spec, float2* u, int nrad, long ustride,
                     const float* wlut, long m_lut, const float* kappa, const float2* tw,
                     hipStream_t s, int tile, int wmode);
void launch_inv_rows_crit(int n, const float2* v, int nrad, long vstride, const int* d2max,
                          const float* bal, int* wmap, float2* dbg, hipStream_t s,
                          const float2* tw);
void launch_fill_i32(int* p, int value, long count, hipStream_t s);
void launch_dt_rows(const DtArgs& a, const int* in, int* out, hipStream_t s);
void launch_dt_cols(const DtArgs& a, const int* in, int* out, hipStream_t s);
void launch_dt_final(const DtArgs& a, const int* in, const float* pa, const float* pb,
                     const float* pc, CellConsts k, float* xhii, double* partials,
                     int npartials, hipStream_t s);
void launch_sum_partials(const double* partials, int np, double scale, double* out, hipStream_t s);
int dt_partials_count(int n);

// ---- plain path (reference dataflow) and the FFTW-like API ----
void launch_r2c_rows(int n, const float* in, float2* out, const float2* tw_half,
                     const float2* twf, hipStream_t s);
void launch_c2r_rows(int n, const float2* in, float* out, const float2* tw_half,
                     const float2* twf, hipStream_t s);
void launch_c2c_rows(int n, int sign, float2* data, const float2* tw, hipStream_t s);
void launch_cell_fields(long n3, const float* a, const float* b, const float* c, CellConsts k,
                        float* rec, float* ion, float* fresid, float* bubblef, hipStream_t s);
void launch_filter_half(int n, const float2* in, float2* out, const float* wlut, hipStream_t s);
void launch_criterion(long n3, const float* rec, const float* ion, float* bubble, int* flag,
                      hipStream_t s);
void launch_ball(int n, int d2, float* top_hat, hipStream_t s);
void launch_cmul(long count, float2* a, const float2* b, hipStream_t s);
void launch_threshold(long n3, const float* count, float scale, double thr, float* bubblef,
                      hipStream_t s);
void launch_stencil(int n, const float* centres, int d2, int reach, float* bubblef,
                    hipStream_t s);
int clamp_partials_count(long n3);
void launch_clamp_sum(long n3, float* bubblef, const float* fresid, double* partials,
                      hipStream_t s);

// ---- binned power spectrum ----
int pk_partials_count(int n);
void launch_power_spectrum(int n, const float2* spec, const short* bin_of_m, const float* kabs_of_m,
                           int nbins, double norm, double* part, double* out, hipStream_t s);

// ---- halo-catalog emissivity ----
void launch_halo_emissivity(long nhalos, const float* halos, int ng, int ns, int sm, float k_rion,
                            float c2, float c3, float scale, float* grid, float* out,
                            hipStream_t s);

}  // namespace sf21
