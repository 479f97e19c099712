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
  int q;  // cells per lane in the line passes (n/q lanes, a multiple of 8, <= 64)
};

// Mixed-radix plan of a line length without a register plan (generic
// sizes): Stockham stages of radix 8, 4, 2, 3, 5, span ns[s] = product of
// the earlier radices.
struct GenPlan {
  int L;
  int nst;
  int lp;          // LDS line stride (float2)
  int xcd;         // bit a (a = 1: j, 2: i): the passes along axis a deal runs of
                   // consecutive blocks to one XCD (an autotuner choice)
  int radix[16];
  int ns[16];
  float rns[16];   // 1 / ns in FP32 (exact quotients of the lane indices)
};

// ---- generic sizes (lengths without a register plan) ----
// Whole-line transforms (first and last stage in HBM, the others through
// LDS), 8 lines per 256-thread block: axis 0
// along p (contiguous rows), 1 along j, 2 along i (8 adjacent p columns).
// mode 0: in place (src == dst allowed); 1 (axis 2 only): inverse-filter
// input -- each element of src times W(|k| R) (kappa = dk R as hi + lo)
// before the transform, result to dst; 2: centre test after the transform (Re = n^3
// rec_R, Im = s n^3 ion_R, s = bal[0]): wmap = d2 where a centre, nothing
// stored (dst null); 3: dst = (Re, Im / s), the unscaled filtered pair (the
// same kernel as mode 2).  LDS: gen_pass_lds(plan) bytes.
bool gen_plan_make(int n, GenPlan* p);
size_t gen_pass_lds(const GenPlan& p);
// allows the transform blocks' LDS on the current device (false: refused)
bool gen_prepare(const GenPlan& p);
void launch_gen_pass(const GenPlan& plan, int axis, int sign, int mode, const float2* src,
                     float2* dst, const float2* tw, float2 kappa, const float* bal, int* wmap,
                     int d2, hipStream_t s);
// spec = (rec, s ion) of the cell pre-pass, s = bal[0]
void launch_gen_pack(long n3, const float* a, const float* b, const float* c, CellConsts k,
                     const float* bal, float2* spec, hipStream_t s);

// ---- fused exact path ----
int balance_partials_count(long n3);
void launch_balance(long n3, const float* a, const float* b, const float* c, CellConsts k,
                    double* part, float* scale, hipStream_t s);
void launch_fwd_rows(int n, const float* a, const float* b, const float* c, CellConsts k,
                     const float* bal, float2* y, const float2* tw, hipStream_t s);
// tile: columns per block (8, 16; 16 or 32 when wide); variant bit 0: 16
// values per thread instead of the size's default (sizes 128, 256, 512);
// bit 1: blocks walk 8 tiles, prefetching the next (tile 16, or 16 and 32 wide);
// bit 2: non-temporal loads and stores; bit 3 (axis 0, width n): the array
// is in the 16-column tile layout -- each (i, 16 p columns) tile an n x 16
// block, element (i, j, p) at i n^2 + (p/16) 16 n + 16 j + p%16
void launch_cols(int n, int width, int axis, int sign, const float2* in, float2* out,
                 int nbatch, long bstride, const float2* tw, hipStream_t s, int tile = 16,
                 int variant = 0);
// jpair: one block per mirror pair of i-lines (kj, n - kj), sharing W
// (tile 8 or 16, weights evaluated in the kernel); t16: u in the 16-column
// tile layout of launch_cols (the p pass then reads it with t16 as well)
void launch_filter_i(int n, const float2* spec, float2* u, int nrad, long ustride,
                     const float* wlut, long m_lut, const float2* kappa, const float2* tw,
                     hipStream_t s, int tile, int wmode, bool jpair = false, bool t16 = false);
void launch_inv_rows_crit(int n, const float2* v, int nrad, long vstride, const int* d2max,
                          const float* bal, int* wmap, float2* dbg, hipStream_t s,
                          const float2* tw, bool wide = false,
                          float2* keep = nullptr, bool t16 = false);
void launch_fill_i32(int* p, int value, long count, hipStream_t s);
// cells per lane of the dilation line passes for box side n (0: n not usable)
int dt_q(int n);
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
// full: spec is the n^3 spectrum of a complex transform (its p <= n/2 half binned)
void launch_power_spectrum(int n, const float2* spec, const short* bin_of_m, const float* kabs_of_m,
                           int nbins, double norm, double* part, double* out, hipStream_t s,
                           bool full = false);

// ---- differentiable mode (exact forward, tau-relaxed adjoint) ----
// slots of the FP64 device state array
enum DiffSlot {
  DS_SUM_EXP = 0,  // sum exp(g)
  DS_SUM_REC,      // sum rec (rho = 1e-3 sum rec / N)
  DS_SUM_X,        // sum x (exact forward)
  DS_SUM_XT,       // sum x~ (relaxed model)
  DS_RHO_BAR,      // dL/drho
  DS_ZETA_BAR,     // sum ionbar E = dL/dzeta
  DS_S,            // sum Deltabar Delta
  DS_COUNT
};
bool diff_size_supported(int n);
int diff_partials_count(long n3);
int rows_blocks(int n);
void launch_sum_exp(long n3, const float* g, double* part, double* st, hipStream_t s);
void launch_make_fields(long n3, const float* g, const float* emis, float zeta, double* st,
                        float* delta1, float* rec, float* ion, double* part, hipStream_t s);
void launch_t_field(long n3, const float* x, const float* delta1, float* t, double* part,
                    double* st, hipStream_t s);
void launch_ball_box(int n, int d2, float2* out, hipStream_t s);
void launch_ball_fold(int n, const float2* spec, float* table, hipStream_t s);
// pairs (nullable): per-radius filtered pairs, stored by rows_phi and read
// by rows_seed instead of a recomputation (field-storage mode)
// kept: pairs holds the batch's real-space pairs (kept by the forward pass);
// no inverse transform of u and no store into pairs
void launch_rows_phi(int n, float2* u, int nrad, long ustride, const float* bal, const double* st,
                     float inv_tau, const float2* tw, float2* pairs, long pstride, hipStream_t s,
                     bool kept = false);
// 64-bit position-mixed hashes of g and emis into out[0], out[1] (part:
// eb(n3) pairs of doubles, reused as 64-bit partials)
void launch_fingerprint(long n3, const float* g, const float* emis, double* part,
                        unsigned long long* out, hipStream_t s);
void launch_rows_seed(int n, float2* u, const float2* v, int nrad, long ustride, long vstride,
                      const float* bal, const double* st, float inv_tau, double* rho_acc,
                      const float2* tw, const float2* pairs, long pstride, hipStream_t s);
void launch_ball_i(int n, bool acc, const float2* in, long istride, float2* out, long ostride,
                   int nrad, const float* bt, const float2* tw, hipStream_t s);
void launch_filter_acc_i(int n, const float2* in, long istride, int nrad, const float* wlut,
                         long mlut, const float2* kappa, int wmode, float2* acc, const float2* tw,
                         hipStream_t s);
void launch_relaxed_state(long n3, const float2* cacc, const float* rec, const float* ion,
                          const float* delta1, float* xt, float* t, float* fq, float* fa,
                          double* part, double* st, hipStream_t s);
void launch_pk_adjoint(int n, float2* spec, const short* bin_of_m, const double* coef,
                       hipStream_t s);
void launch_pk_coef(const double* gp, const double* pk, int nbins, double norm, double* coef,
                    hipStream_t s);
void launch_seed_a(long n3, const float* tbar, const float* delta1, const float* xt,
                   const float* fq, const float* fa, float gx_n, float2* abar, float* qbar,
                   float* dbar, hipStream_t s);
void launch_sum_rho(const double* acc, int np, double* st, hipStream_t s);
void launch_grad_cells(long n3, const float2* gacc, const float* qbar, const float* rec,
                       const float* ion, const float* emis, const float* delta1,
                       const float* dbar, double* st, float* dd_out, double* part,
                       hipStream_t s);
void launch_grad_g(long n3, const float* delta1, const float* dd, const double* st, float* gout,
                   hipStream_t s);
// generic sides: full n^3 complex boxes, transforms by launch_gen_pass
int gen_seed_blocks(long n3);  // rho accumulators written by launch_gen_seed
void launch_gen_phi(long n3, const float2* pair, float2* slot, int half, const float* bal,
                    const double* st, float inv_tau, hipStream_t s);
void launch_gen_ball(int n, bool acc, const float2* in, float2* out, const float* ta,
                     const float* tb, hipStream_t s);
void launch_gen_seed(long n3, float2* u, const float2* v, int half, const float* bal,
                     const double* st, float inv_tau, double* rho_acc, hipStream_t s);
void launch_gen_filter_acc(int n, const float2* in, float2* acc, float2 kappa, hipStream_t s);
void launch_pk_adjoint_full(int n, float2* spec, const short* bin_of_m, const double* coef,
                            hipStream_t s);
void launch_fill_c(float2* p, long count, hipStream_t s);
void launch_fill_d(double* p, long count, hipStream_t s);
void launch_to_complex(long n3, const float* in, float2* out, hipStream_t s);

// ---- 21-cm brightness temperature ----
int t21_partials_count(long n3);
void launch_vgrad_filter(int n, float2* spec, int axis, hipStream_t s);
// the same on the full n^3 spectrum of a complex transform (generic sides)
void launch_vgrad_filter_full(int n, float2* spec, int axis, hipStream_t s);
void launch_real_part(long n3, const float2* in, float* out, hipStream_t s);
void launch_t21_cells(long n3, const float* delta, const float* xhii, const float* dv_raw,
                      const float* xc, const float* xa, const float* tk, float vfac, float c21,
                      float tcmb, float* out, double* part, double* sums, hipStream_t s);

// ---- halo-catalog emissivity, SFRD and non-linear density deposits ----
int sum_partials_count(long count);
void launch_sum_f32(long count, const float* x, double* part, double* out, hipStream_t s);
void launch_nldensity(int ng, int ns, int sm, const float* vx, const float* vy, const float* vz,
                      const float* delta0, float vscale, float growth_in, float growth,
                      float* grid, float* out, hipStream_t s);
void launch_halo_emissivity(long nhalos, const float* halos, int ng, int ns, int sm, float k_rion,
                            float c2, float c3, float scale, float* grid, float* out,
                            hipStream_t s);

}  // namespace sf21
