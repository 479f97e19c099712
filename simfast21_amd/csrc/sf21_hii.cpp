// libsf21hii: context, launch sequences and C-ABI of the MI355X bubble
// finder.  See include/sf21_hii.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sf21_hii.h"
#include "kernels.h"

using namespace sf21;

namespace {

// ---- astrophysical fits of the SFR-mode pre-pass (user_functions.c) ----
constexpr double kCmPerMpc = 3.0857e24;  // user_functions.c:26
constexpr double kVnorm = 1.e+45;        // user_functions.c:38
constexpr double kR1 = 9.84815696, kR2 = 1.76105133, kR3 = 0.81722878, kR4 = 5.06773134;
constexpr double kQa = 2.05707274e-02, kQb = 5.29944413e+01;
constexpr double kPi = 3.1415926535897932385E0;  // Input_variables.h:10

// user_functions.c:92-97
double fit_G_H(double z) {
  const double e = -4.37152818e-05 * std::pow(z, 5.) + 1.64394635e-03 * std::pow(z, 4.) -
                   2.03044756e-02 * std::pow(z, 3.) + 7.18335221e-02 * std::pow(z, 2.) -
                   8.85791388e-02 * z - 1.20887500e+01;
  return 4.19232273531e-13 / std::pow(10.0, e);
}
// user_functions.c:84-88
double fit_Qion(double z) { return std::pow(10.0, kQa * z + kQb) / kVnorm; }

// auxiliary.c:350-359
double w_filter(double x) {
  double j;
  if (x < 0.03)
    j = 1.0 / 3.0 - x * x / 30.0;
  else
    j = -(std::cos(x) - std::sin(x) / x) / x / x;
  return 3.0 * j;
}

// Box sides with register transform plans (32: complex transforms only).
bool plan_size(int n) {
  switch (n) {
    case 64: case 96: case 128: case 192: case 256: case 512: return true;
    default: return false;
  }
}

// Other sides of the fused finder: mixed-radix (2, 3, 5) transforms through
// LDS, and dilation line passes whose lanes tile the side.
bool generic_size(int n) {
  GenPlan g;
  return n >= 16 && n <= 768 && gen_plan_make(n, &g) && dt_q(n) > 0;
}

bool size_supported(int n, int path) {
  if (n == 32) return path == SF21_PATH_FUSED;
  if (plan_size(n)) return true;
  return path == SF21_PATH_FUSED && generic_size(n);
}

struct KStat {
  std::string name;
  long launches = 0;
  double ms = 0.0;
};

struct Pending {
  int stat;
  hipEvent_t a, b;
};

struct GraphEntry {
  std::vector<const void*> ptrs;
  std::vector<double> zs;
  // multi-box graphs: (id, generation) of every context the graph uses
  std::vector<unsigned long long> ids;
  hipGraphExec_t exec = nullptr;
};

// Context ids are never reused, so a graph that names a destroyed context
// matches no later call, even when a new context lands at the same address.
std::atomic<unsigned long long> g_next_ctx_id{1};

}  // namespace

// Column-pass variant bit for whole streamed complex buffers (width n): the
// loads and stores carry the non-temporal hint.  The half-spectrum passes
// (width n/2 + 1, tiles straddling cache lines) do not use it.
constexpr int kColsStream = 4;
// Column-pass variant bit: the per-radius buffer of the forward finder is in
// the 16-column tile layout (the filter pass writes it, the j pass runs on
// contiguous n x 16 blocks, the p pass reads 128 B pieces of its rows).
constexpr int kColsT16 = 8;

struct sf21_hii_ctx {
  sf21_hii_config cfg{};
  std::vector<double> radii;
  std::vector<int> d2;
  int n = 0, h = 0, K = 0;
  int tile_filter = 16;
  int cols_tile = 16;  // columns per block of the per-radius passes along j
  int cols_wide = 0;  // per-radius passes along j: bit 0 16 values per thread, bit 1 prefetching blocks
  bool rows_wide = false;  // 16 values per thread in the centre-test pass
  int filter_w = 1;  // 1: FP32 on the fly, 0: FP64 table per |k|^2
  int filter_pair = 0;  // 1: the filter pass takes mirror pairs of i-lines
  bool generic = false;  // fused finder on a side without a register plan
  GenPlan gplan{};
  std::vector<float2> kap;  // dk R per radius as a float pair (host copy)
  float2* d_kappa = nullptr;  // dk R per radius as a float pair (hi, lo)
  long n3 = 0, nh = 0;
  double L = 0, dx = 0, dk = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // tables
  float2 *tw = nullptr, *twh = nullptr;
  float* wlut = nullptr;
  long mlut = 0;
  int* d_d2 = nullptr;
  // fused path
  float2 *spec = nullptr, *ubuf = nullptr;
  int *wmap = nullptr, *gbuf = nullptr;
  double* bal_part = nullptr;
  float* bal = nullptr;
  DtArgs dt{};
  double* partials = nullptr;
  int npart = 0;
  double* neutral_dev = nullptr;
  int neutral_cap = 0;
  double* neutral_host = nullptr;
  // plain path
  float *rec = nullptr, *ion = nullptr, *fres = nullptr, *bubblef = nullptr;
  float *box_a = nullptr, *box_b = nullptr, *bubble = nullptr;
  float2 *spec2 = nullptr, *tmp2 = nullptr, *th = nullptr;
  int* flag_dev = nullptr;
  int* flag_host = nullptr;
  // power spectrum
  short* pk_bin = nullptr;
  float* pk_kabs = nullptr;
  double* pk_part = nullptr;
  double* pk_out = nullptr;
  int pk_nbins = 0;
  // differentiable mode (allocated on first use)
  float *df_delta = nullptr, *df_rec = nullptr, *df_ion = nullptr, *df_x = nullptr;
  float *df_t = nullptr, *df_fq = nullptr, *df_fa = nullptr, *df_qbar = nullptr;
  float *df_dbar = nullptr, *df_dd = nullptr;
  float2 *df_v = nullptr, *df_c = nullptr, *df_g = nullptr, *df_half = nullptr;
  float* df_ball = nullptr;  // folded ball spectra, (n/2+1)^3 per radius
  double *df_st = nullptr, *df_part = nullptr, *df_rho = nullptr, *df_gp = nullptr;
  double* df_coef = nullptr;
  double* df_host = nullptr;  // pinned staging
  float2* df_pairs = nullptr; // field-storage mode: n_R x n^3 filtered pairs
  int df_store = 2;           // 0 recompute, 1 store, 2 store if it fits in half the free HBM,
                              // 3 the exact forward keeps them for the backward
  unsigned long long* df_fp = nullptr;  // input hashes (device): [0..1] kept, [2..3] current
  bool df_kept = false;       // df_pairs hold the fields of the inputs fingerprinted below
  unsigned long long df_kept_fp[2] = {0, 0};
  unsigned long long* df_fp_host = nullptr;  // pinned staging of the four hashes
  double df_kept_zeta = 0.0;
  bool df_last_kept = false;  // the last backward read forward-kept fields
  bool df_ready = false;
  // 21-cm step
  float* t21_dv = nullptr;      // velocity gradient (unnormalised C2R output)
  double* t21_part = nullptr;   // partial sums, then the two totals at the end
  // halo mode
  float* halo_grid = nullptr;   // N_halo^3 deposit grid
  float* halo_buf = nullptr;    // catalog upload (host entry point)
  double* sum_part = nullptr;   // FP64 partial sums of a box, then the total
  long halo_cap = 0;
  // profiling
  bool prof = false;
  std::vector<KStat> stats;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  // graphs
  std::vector<GraphEntry> graphs;
  bool capturing = false;
  // identity for graphs that span several contexts: id is unique per
  // context; gen changes whenever buffers, stream or kernel choices change
  unsigned long long id = 0, gen = 0;
  long captures = 0;  // graphs captured with this context as owner
};

struct sf21_fft_plan {
  int n = 0, kind = 0;
  float2 *tw = nullptr, *twh = nullptr;
};

namespace {

template <class T>
int dev_alloc(T** p, size_t count) {
  if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return SF21_ERR_ALLOC;
  }
  return SF21_OK;
}

template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// exp(-2 pi i q/len), q < len, rounded from FP64.
int make_twiddles(int len, float2** out) {
  std::vector<float2> h(len);
  for (int q = 0; q < len; ++q) {
    const double a = -2.0 * kPi * (double)q / (double)len;
    h[q] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  int rc = dev_alloc(out, (size_t)len);
  if (rc) return rc;
  if (hipMemcpy(*out, h.data(), sizeof(float2) * len, hipMemcpyHostToDevice) != hipSuccess)
    return SF21_ERR_HIP;
  return SF21_OK;
}

// Per-slice constants of the SFR-mode pre-pass (get_HIIbubbles.c:220-222,
// :273 with user_functions.c:58-65, :84-97), FP64 on the host.
CellConsts cell_consts(const sf21_hii_ctx* c, double z) {
  CellConsts k{};
  // halo mode shares the SFR pre-pass (rec, fresid from delta) with the
  // emissivity box taken as given (k_ion = 1)
  k.mode = (c->cfg.input_mode == SF21_INPUT_SFR || c->cfg.input_mode == SF21_INPUT_HALO) ? 0 : 1;
  if (k.mode == 0) {
    const double h = c->cfg.hubble;
    k.c_fres = (float)(1.881e-7 * std::pow(1. + z, 3.0) * fit_G_H(z));
    const double dvv = std::pow(c->L * kCmPerMpc / c->n / h / (1. + z), 3.0);
    k.k_rec = (float)(1.e-24 * kR1 * std::pow(1.0 + z, kR4) * dvv / kVnorm);
    k.r2 = (float)kR2;
    k.r3 = (float)kR3;
    k.k_ion = c->cfg.input_mode == SF21_INPUT_HALO
                  ? 1.0f
                  : (float)(c->dx * c->dx * c->dx * fit_Qion(z) * c->cfg.fesc);
  }
  return k;
}

// ---- profiling: HIP event pairs around each launch on the ctx stream ----
int stat_index(sf21_hii_ctx* c, const char* name) {
  for (size_t i = 0; i < c->stats.size(); ++i)
    if (c->stats[i].name == name) return (int)i;
  c->stats.push_back(KStat{name, 0, 0.0});
  return (int)c->stats.size() - 1;
}

hipEvent_t get_event(sf21_hii_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

template <class F>
void launch(sf21_hii_ctx* c, const char* name, F&& f) {
  if (c->prof && !c->capturing) {
    Pending p{stat_index(c, name), get_event(c), get_event(c)};
    (void)hipEventRecord(p.a, c->stream);
    f();
    (void)hipEventRecord(p.b, c->stream);
    c->pending.push_back(p);
  } else {
    f();
  }
}

void collect(sf21_hii_ctx* c) {
  for (auto& p : c->pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(p.b);
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->stats[p.stat].ms += ms;
      c->stats[p.stat].launches += 1;
    }
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

int check_hip() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SF21_OK : SF21_ERR_HIP;
}

// ---- fused exact path on a generic side: one radius per step ------------
// Forward transforms of z = rec + i s ion along p, j, i; per radius in
// ascending order, the filtered inverse along i, the inverse along j and the
// inverse along p with the centre test; then the same dilation and final
// pass as the register-plan sides.
void enqueue_generic(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
                     float* xhii, double* neutral_slot, int only_radius, float2* dbg) {
  const long n3 = c->n3;
  const CellConsts k = cell_consts(c, z);
  hipStream_t s = c->stream;
  const GenPlan& g = c->gplan;
  const float2 none = make_float2(0.f, 0.f);
  launch(c, "balance", [&] { launch_balance(n3, a, b, cc, k, c->bal_part, c->bal, s); });
  launch(c, "gen_pack", [&] { launch_gen_pack(n3, a, b, cc, k, c->bal, c->spec, s); });
  for (int axis = 0; axis < 3; ++axis)
    launch(c, "gen_fwd", [&] {
      launch_gen_pass(g, axis, -1, 0, c->spec, c->spec, c->tw, none, nullptr, nullptr, 0, s);
    });
  launch(c, "fill_map", [&] { launch_fill_i32(c->wmap, -1, n3, s); });
  const int r_lo = only_radius >= 0 ? only_radius : 0;
  const int r_hi = only_radius >= 0 ? only_radius + 1 : (int)c->radii.size();
  for (int r = r_hi - 1; r >= r_lo; --r) {
    launch(c, "gen_filter_i", [&] {
      launch_gen_pass(g, 2, +1, 1, c->spec, c->ubuf, c->tw, c->kap[r], nullptr, nullptr, 0, s);
    });
    launch(c, "gen_inv_j", [&] {
      launch_gen_pass(g, 1, +1, 0, c->ubuf, c->ubuf, c->tw, none, nullptr, nullptr, 0, s);
    });
    launch(c, "gen_inv_p_crit", [&] {
      if (dbg)
        launch_gen_pass(g, 0, +1, 3, c->ubuf, dbg, c->tw, none, c->bal, nullptr, 0, s);
      else
        launch_gen_pass(g, 0, +1, 2, c->ubuf, nullptr, c->tw, none, c->bal, c->wmap, c->d2[r], s);
    });
  }
  if (only_radius >= 0) return;
  launch(c, "dt_rows", [&] { launch_dt_rows(c->dt, c->wmap, c->gbuf, s); });
  launch(c, "dt_cols", [&] { launch_dt_cols(c->dt, c->gbuf, c->gbuf, s); });
  launch(c, "dt_final", [&] {
    launch_dt_final(c->dt, c->gbuf, a, b, cc, k, xhii, c->partials, c->npart, s);
  });
  launch(c, "sum", [&] {
    launch_sum_partials(c->partials, c->npart, 1.0 / (double)n3, neutral_slot, s);
  });
}

// ---- fused exact path: enqueue one slice -------------------------------
// keep (nullable): the filtered pairs of every radius, radius r at keep + r n^3
void enqueue_fused(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
                   float* xhii, double* neutral_slot, int only_radius, float2* dbg,
                   float2* keep = nullptr) {
  if (c->generic) {  // the differentiable mode (keep) refuses generic sides
    enqueue_generic(c, a, b, cc, z, xhii, neutral_slot, only_radius, dbg);
    return;
  }
  const int n = c->n;
  const long n3 = c->n3;
  const CellConsts k = cell_consts(c, z);
  hipStream_t s = c->stream;
  launch(c, "balance", [&] { launch_balance(n3, a, b, cc, k, c->bal_part, c->bal, s); });
  launch(c, "fwd_rows", [&] { launch_fwd_rows(n, a, b, cc, k, c->bal, c->spec, c->tw, s); });
  launch(c, "cols_fwd_j", [&] { launch_cols(n, n, 0, -1, c->spec, c->spec, 1, 0, c->tw, s, 16, kColsStream); });
  launch(c, "fill_map", [&] { launch_fill_i32(c->wmap, -1, n3, s); });
  // radius batches in ascending radius (from the end of the descending
  // ladder): each batch's ball bounds exceed the earlier batches', which the
  // centre-map update relies on
  const int r_lo = only_radius >= 0 ? only_radius : 0;
  const int r_hi = only_radius >= 0 ? only_radius + 1 : (int)c->radii.size();
  for (int re = r_hi; re > r_lo; re -= c->K) {
    const int rb = std::max(r_lo, re - c->K);
    const int kk = re - rb;
    launch(c, "filter_i", [&] {
      launch_filter_i(n, c->spec, c->ubuf, kk, n3, c->wlut + (long)rb * c->mlut, c->mlut,
                      c->d_kappa + rb, c->tw, s, c->tile_filter, c->filter_w, c->filter_pair,
                      true);
    });
    launch(c, "cols_inv_j", [&] {
      launch_cols(n, n, 0, +1, c->ubuf, c->ubuf, kk, n3, c->tw, s, c->cols_tile,
                  c->cols_wide | kColsStream | kColsT16);
    });
    launch(c, "inv_rows_crit", [&] {
      launch_inv_rows_crit(n, c->ubuf, kk, n3, c->d_d2 + rb, c->bal, c->wmap, dbg, s, c->tw, c->rows_wide,
                           keep ? keep + (long)rb * n3 : nullptr, true);
    });
  }
  if (only_radius >= 0) return;
  // the centre map stays intact for inspection; each line pass reads its
  // whole line before writing it, so the later passes run in place
  launch(c, "dt_rows", [&] { launch_dt_rows(c->dt, c->wmap, c->gbuf, s); });
  launch(c, "dt_cols", [&] { launch_dt_cols(c->dt, c->gbuf, c->gbuf, s); });
  launch(c, "dt_final", [&] {
    launch_dt_final(c->dt, c->gbuf, a, b, cc, k, xhii, c->partials, c->npart, s);
  });
  launch(c, "sum", [&] {
    launch_sum_partials(c->partials, c->npart, 1.0 / (double)n3, neutral_slot, s);
  });
}

// ---- plain path helpers: FFTW-like 3D real transforms ------------------
void r2c3d(sf21_hii_ctx* c, const float* in, float2* out) {
  const int n = c->n, h = n / 2 + 1;
  hipStream_t s = c->stream;
  launch(c, "r2c_rows", [&] { launch_r2c_rows(n, in, out, c->twh, c->tw, s); });
  launch(c, "cols_half", [&] { launch_cols(n, h, 0, -1, out, out, 1, 0, c->tw, s); });
  launch(c, "cols_half", [&] { launch_cols(n, h, 1, -1, out, out, 1, 0, c->tw, s); });
}

void c2r3d(sf21_hii_ctx* c, float2* in, float* out) {
  const int n = c->n, h = n / 2 + 1;
  hipStream_t s = c->stream;
  launch(c, "cols_half", [&] { launch_cols(n, h, 1, +1, in, in, 1, 0, c->tw, s); });
  launch(c, "cols_half", [&] { launch_cols(n, h, 0, +1, in, in, 1, 0, c->tw, s); });
  launch(c, "c2r_rows", [&] { launch_c2r_rows(n, in, out, c->twh, c->tw, s); });
}

int run_plain(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
              float* xhii, double* neutral) {
  const int n = c->n;
  const long n3 = c->n3, nh = c->nh;
  const CellConsts k = cell_consts(c, z);
  hipStream_t s = c->stream;
  launch(c, "cell_fields", [&] {
    launch_cell_fields(n3, a, b, cc, k, c->rec, c->ion, c->fres, c->bubblef, s);
  });
  r2c3d(c, c->rec, c->spec2);
  r2c3d(c, c->ion, c->spec2 + nh);
  const double dx3 = c->dx * c->dx * c->dx;
  for (size_t r = 0; r < c->radii.size(); ++r) {
    const double R = c->radii[r];
    launch(c, "filter_half",
           [&] { launch_filter_half(n, c->spec2, c->tmp2, c->wlut + (long)r * c->mlut, s); });
    c2r3d(c, c->tmp2, c->box_a);
    c2r3d(c, c->tmp2 + nh, c->box_b);
    launch(c, "fill_flag", [&] { launch_fill_i32(c->flag_dev, 0, 1, s); });
    launch(c, "criterion",
           [&] { launch_criterion(n3, c->box_a, c->box_b, c->bubble, c->flag_dev, s); });
    if (hipMemcpyAsync(c->flag_host, c->flag_dev, sizeof(int), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
      return SF21_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return SF21_ERR_HIP;
    if (!c->flag_host[0]) continue;
    if (R >= 2 * c->dx) {
      launch(c, "ball", [&] { launch_ball(n, c->d2[r], c->box_a, s); });
      r2c3d(c, c->box_a, c->th);
      r2c3d(c, c->bubble, c->tmp2);
      launch(c, "cmul", [&] { launch_cmul(nh, c->tmp2, c->th, s); });
      c2r3d(c, c->tmp2, c->box_b);
      launch(c, "threshold", [&] {
        launch_threshold(n3, c->box_b, (float)(dx3 / (double)n3), dx3 / 100, c->bubblef, s);
      });
    } else {
      const int reach = (int)(R / c->dx) + 1;
      launch(c, "stencil", [&] { launch_stencil(n, c->bubble, c->d2[r], reach, c->bubblef, s); });
    }
  }
  const int np = clamp_partials_count(n3);
  launch(c, "clamp_sum", [&] { launch_clamp_sum(n3, c->bubblef, c->fres, c->partials, s); });
  launch(c, "sum", [&] {
    launch_sum_partials(c->partials, np, 1.0 / (double)n3, c->neutral_dev, s);
  });
  if (hipMemcpyAsync(xhii, c->bubblef, sizeof(float) * n3, hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return SF21_ERR_HIP;
  if (hipMemcpyAsync(c->neutral_host, c->neutral_dev, sizeof(double), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return SF21_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return SF21_ERR_HIP;
  if (neutral) *neutral = c->neutral_host[0];
  return check_hip();
}

void free_all(sf21_hii_ctx* c) {
  for (auto& g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  dev_free(c->tw); dev_free(c->twh); dev_free(c->wlut); dev_free(c->d_d2); dev_free(c->d_kappa);
  dev_free(c->spec); dev_free(c->ubuf); dev_free(c->wmap); dev_free(c->gbuf);
  dev_free(c->bal_part); dev_free(c->bal); dev_free(c->partials); dev_free(c->neutral_dev);
  dev_free(c->rec); dev_free(c->ion); dev_free(c->fres); dev_free(c->bubblef);
  dev_free(c->box_a); dev_free(c->box_b); dev_free(c->bubble);
  dev_free(c->spec2); dev_free(c->tmp2); dev_free(c->th); dev_free(c->flag_dev);
  dev_free(c->halo_grid); dev_free(c->halo_buf);
  dev_free(c->t21_dv); dev_free(c->t21_part); dev_free(c->sum_part);
  dev_free(c->pk_bin); dev_free(c->pk_kabs); dev_free(c->pk_part); dev_free(c->pk_out);
  dev_free(c->df_delta); dev_free(c->df_rec); dev_free(c->df_ion); dev_free(c->df_x);
  dev_free(c->df_t); dev_free(c->df_fq); dev_free(c->df_fa); dev_free(c->df_qbar);
  dev_free(c->df_dbar); dev_free(c->df_dd); dev_free(c->df_v); dev_free(c->df_c);
  dev_free(c->df_g); dev_free(c->df_half); dev_free(c->df_ball); dev_free(c->df_st);
  dev_free(c->df_part); dev_free(c->df_rho); dev_free(c->df_gp); dev_free(c->df_coef);
  dev_free(c->df_fp);
  dev_free(c->df_pairs);
  if (c->df_host) (void)hipHostFree(c->df_host);
  c->df_host = nullptr;
  if (c->df_fp_host) (void)hipHostFree(c->df_fp_host);
  c->df_fp_host = nullptr;
  c->df_ready = false;
  if (c->flag_host) (void)hipHostFree(c->flag_host);
  if (c->neutral_host) (void)hipHostFree(c->neutral_host);
  c->flag_host = nullptr;
  c->neutral_host = nullptr;
  for (auto e : c->pool) (void)hipEventDestroy(e);
  c->pool.clear();
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

// ---- binned power spectrum: tables and launch ---------------------------
// Bin of every integer |k|^2 (in dk units): linear bins of |k| over [1, n/2],
// width (n/2 - 1)/nbins, the last closed; k = 0 and the corners beyond n/2
// are dropped (-1).
int ensure_pk(sf21_hii_ctx* c, int nbins) {
  if (c->pk_nbins == nbins) return SF21_OK;
  const int hh = c->n / 2;
  const long mmax = 3L * hh * hh;
  std::vector<short> bins(mmax + 1);
  std::vector<float> kab(mmax + 1);
  const double width = (hh - 1) / (double)nbins;
  for (long m = 0; m <= mmax; ++m) {
    const double kq = std::sqrt((double)m);
    int b = -1;
    if (m >= 1 && kq <= hh) {
      b = (int)std::floor((kq - 1.0) / width);
      if (b >= nbins) b = nbins - 1;
    }
    bins[m] = (short)b;
    kab[m] = (float)(kq * c->dk);
  }
  int rc;
  dev_free(c->pk_bin); dev_free(c->pk_kabs); dev_free(c->pk_part); dev_free(c->pk_out);
  c->pk_nbins = 0;
  if ((rc = dev_alloc(&c->pk_bin, bins.size())) || (rc = dev_alloc(&c->pk_kabs, kab.size())) ||
      (rc = dev_alloc(&c->pk_part, (size_t)3 * nbins * pk_partials_count(c->n))) ||
      (rc = dev_alloc(&c->pk_out, (size_t)3 * nbins)))
    return rc;
  if (hipMemcpy(c->pk_bin, bins.data(), sizeof(short) * bins.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->pk_kabs, kab.data(), sizeof(float) * kab.size(), hipMemcpyHostToDevice) != hipSuccess)
    return SF21_ERR_HIP;
  c->pk_nbins = nbins;
  return SF21_OK;
}

// P(k) of a real box (tools/power3d.c:172-191 normalisation, dl^3/N) into
// c->pk_out: [0, nbins) power, [nbins, 2 nbins) mean |k|, then counts.
// half: scratch of n*n*(n/2+1) complex values, left holding the spectrum.
// half: the half spectrum, or on generic sides the full n^3 spectrum
void enqueue_pk(sf21_hii_ctx* c, const float* field, float2* half, int nbins) {
  hipStream_t s = c->stream;
  if (c->generic) {
    const float2 none = make_float2(0.f, 0.f);
    launch(c, "to_complex", [&] { launch_to_complex(c->n3, field, half, s); });
    for (int axis = 0; axis < 3; ++axis)
      launch(c, "gen_fwd", [&] {
        launch_gen_pass(c->gplan, axis, -1, 0, half, half, c->tw, none, nullptr, nullptr, 0, s);
      });
  } else {
    r2c3d(c, field, half);
  }
  const double dl = c->dx;
  const double norm = dl * dl * dl / ((double)c->n3);
  launch(c, "power_spectrum", [&] {
    launch_power_spectrum(c->n, half, c->pk_bin, c->pk_kabs, nbins, norm, c->pk_part, c->pk_out,
                          s, c->generic);
  });
}

// ---- differentiable mode ----------------------------------------------
// Unnormalised 3D complex transform in place on a generic side (line passes
// along p, j, i).
void gen_fft3(sf21_hii_ctx* c, float2* x, int sign) {
  const float2 none = make_float2(0.f, 0.f);
  for (int axis = 0; axis < 3; ++axis)
    launch(c, sign < 0 ? "gen_fwd" : "gen_inv", [&] {
      launch_gen_pass(c->gplan, axis, sign, 0, x, x, c->tw, none, nullptr, nullptr, 0, c->stream);
    });
}

int ensure_diff(sf21_hii_ctx* c) {
  if (c->df_ready) return SF21_OK;
  const size_t n3 = (size_t)c->n3, nh = (size_t)c->nh;
  const size_t h = (size_t)(c->n / 2 + 1);
  const size_t npair = (size_t)(c->K + 1) / 2;
  int rc;
  if ((rc = dev_alloc(&c->df_delta, n3)) || (rc = dev_alloc(&c->df_rec, n3)) ||
      (rc = dev_alloc(&c->df_ion, n3)) || (rc = dev_alloc(&c->df_x, n3)) ||
      (rc = dev_alloc(&c->df_t, n3)) || (rc = dev_alloc(&c->df_fq, n3)) ||
      (rc = dev_alloc(&c->df_fa, n3)) || (rc = dev_alloc(&c->df_qbar, n3)) ||
      (rc = dev_alloc(&c->df_dbar, n3)) || (rc = dev_alloc(&c->df_dd, n3)) ||
      (rc = dev_alloc(&c->df_v, npair * n3)) || (rc = dev_alloc(&c->df_c, n3)) ||
      (rc = dev_alloc(&c->df_g, n3)) ||
      // the P(k) spectrum: the half spectrum, or on generic sides the full one
      (rc = dev_alloc(&c->df_half, c->generic ? n3 : nh)) ||
      (rc = dev_alloc(&c->df_ball, h * h * h * c->radii.size())) ||
      (rc = dev_alloc(&c->df_st, (size_t)DS_COUNT)) ||
      (rc = dev_alloc(&c->df_part, (size_t)2 * diff_partials_count(c->n3))) ||
      (rc = dev_alloc(&c->df_rho, (size_t)std::max(rows_blocks(c->n), gen_seed_blocks(c->n3)))) ||
      (rc = dev_alloc(&c->df_gp, 64)) ||
      (rc = dev_alloc(&c->df_coef, 64)) || (rc = dev_alloc(&c->df_fp, 4)))
    return rc;
  if (hipHostMalloc(reinterpret_cast<void**>(&c->df_host), sizeof(double) * 256) != hipSuccess) {
    c->df_host = nullptr;
    return SF21_ERR_ALLOC;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->df_fp_host), sizeof(unsigned long long) * 4) !=
      hipSuccess) {
    c->df_fp_host = nullptr;
    return SF21_ERR_ALLOC;
  }
  // ball spectra of every radius: FFT of the periodic ball indicator, folded
  hipStream_t s = c->stream;
  const int n = c->n;
  for (size_t r = 0; r < c->radii.size(); ++r) {
    launch_ball_box(n, c->d2[r], c->df_c, s);
    if (c->generic) {
      gen_fft3(c, c->df_c, -1);
    } else {
      launch_c2c_rows(n, -1, c->df_c, c->tw, s);
      launch_cols(n, n, 0, -1, c->df_c, c->df_c, 1, 0, c->tw, s, 16, kColsStream);
      launch_cols(n, n, 1, -1, c->df_c, c->df_c, 1, 0, c->tw, s, 16, kColsStream);
    }
    launch_ball_fold(n, c->df_c, c->df_ball + r * h * h * h, s);
  }
  if (hipStreamSynchronize(s) != hipSuccess) return SF21_ERR_HIP;
  if ((rc = check_hip())) return rc;
  c->df_ready = true;
  return SF21_OK;
}

// Delta, rec, ion from (g, emis, zeta) and the FP64 sums they need.
void enqueue_diff_fields(sf21_hii_ctx* c, const float* g, const float* emis, double zeta) {
  hipStream_t s = c->stream;
  launch(c, "diff_fields", [&] {
    launch_sum_exp(c->n3, g, c->df_part, c->df_st, s);
    launch_make_fields(c->n3, g, emis, (float)zeta, c->df_st, c->df_delta, c->df_rec, c->df_ion,
                       c->df_part, s);
  });
}

// Exact forward: x (bubble finder on rec = Delta^2, ion = zeta emis, direct
// inputs), T = (1 - x) Delta, sum(x), P(k) of T.
void enqueue_diff_forward(sf21_hii_ctx* c, const float* g, const float* emis, double zeta,
                          int nbins, float* x, bool keep) {
  hipStream_t s = c->stream;
  enqueue_diff_fields(c, g, emis, zeta);
  enqueue_fused(c, c->df_rec, c->df_ion, nullptr, 0.0, x, c->neutral_dev, -1, nullptr,
                keep ? c->df_pairs : nullptr);
  if (keep) launch(c, "fingerprint", [&] { launch_fingerprint(c->n3, g, emis, c->df_part, c->df_fp, s); });
  launch(c, "diff_t", [&] { launch_t_field(c->n3, x, c->df_delta, c->df_t, c->df_part, c->df_st, s); });
  enqueue_pk(c, c->df_t, c->df_half, nbins);
}

// Relaxed adjoint.  Expects df_gp (dL/dP_b) on the device.
// kept: df_pairs hold this input's filtered pairs (kept by the forward);
// sweep 1 reads them instead of recomputing them from the spectrum
void enqueue_diff_backward(sf21_hii_ctx* c, const float* g, const float* emis, double zeta,
                           double tau, int nbins, double gx, float* grad_g, bool kept) {
  hipStream_t s = c->stream;
  const int n = c->n;
  const long n3 = c->n3;
  const int nr = (int)c->radii.size();
  const long h = n / 2 + 1, h3 = h * h * h;
  const float inv_tau = (float)(1.0 / tau);
  const CellConsts k = cell_consts(c, 0.0);  // direct inputs
  enqueue_diff_fields(c, g, emis, zeta);
  // forward spectrum Z of rec + i s ion, (p,j)-space
  launch(c, "balance", [&] { launch_balance(n3, c->df_rec, c->df_ion, nullptr, k, c->bal_part, c->bal, s); });
  if (!kept) {
    launch(c, "fwd_rows", [&] { launch_fwd_rows(n, c->df_rec, c->df_ion, nullptr, k, c->bal, c->spec, c->tw, s); });
    launch(c, "cols_fwd_j", [&] { launch_cols(n, n, 0, -1, c->spec, c->spec, 1, 0, c->tw, s, 16, kColsStream); });
  }
  // sweep 1: C = sum_R B_R FFT(phi_R)
  launch(c, "diff_fill", [&] { launch_fill_c(c->df_c, n3, s); });
  for (int rb = 0; rb < nr; rb += c->K) {
    const int kk = std::min(c->K, nr - rb);
    const int kp = (kk + 1) / 2;
    if (kept) {
      launch(c, "rows_phi", [&] {
        launch_rows_phi(n, c->ubuf, kk, n3, c->bal, c->df_st, inv_tau, c->tw, c->df_pairs + (long)rb * n3,
                        n3, s, true);
      });
    } else {
      launch(c, "filter_i", [&] {
        launch_filter_i(n, c->spec, c->ubuf, kk, n3, c->wlut + (long)rb * c->mlut, c->mlut,
                        c->d_kappa + rb, c->tw, s, c->tile_filter, c->filter_w, c->filter_pair);
      });
      launch(c, "cols_inv_j", [&] { launch_cols(n, n, 0, +1, c->ubuf, c->ubuf, kk, n3, c->tw, s, c->cols_tile, c->cols_wide | kColsStream); });
      launch(c, "rows_phi", [&] {
        launch_rows_phi(n, c->ubuf, kk, n3, c->bal, c->df_st, inv_tau, c->tw,
                        c->df_pairs ? c->df_pairs + (long)rb * n3 : nullptr, n3, s);
      });
    }
    launch(c, "cols_fwd_j", [&] { launch_cols(n, n, 0, -1, c->ubuf, c->ubuf, kp, n3, c->tw, s, c->cols_tile, c->cols_wide | kColsStream); });
    launch(c, "ball_acc_i", [&] {
      launch_ball_i(n, true, c->ubuf, n3, c->df_c, 0, kk, c->df_ball + (long)rb * h3, c->tw, s);
    });
  }
  launch(c, "cols_inv_j", [&] { launch_cols(n, n, 0, +1, c->df_c, c->df_c, 1, 0, c->tw, s, 16, kColsStream); });
  launch(c, "c2c_rows", [&] { launch_c2c_rows(n, +1, c->df_c, c->tw, s); });
  // relaxed state, its P(k) and the observables' adjoints
  launch(c, "relaxed_state", [&] {
    launch_relaxed_state(n3, c->df_c, c->df_rec, c->df_ion, c->df_delta, c->df_x, c->df_t,
                         c->df_fq, c->df_fa, c->df_part, c->df_st, s);
  });
  enqueue_pk(c, c->df_t, c->df_half, nbins);
  const double dl = c->dx;
  const double norm = dl * dl * dl / ((double)n3);
  launch(c, "pk_adjoint", [&] {
    launch_pk_coef(c->df_gp, c->pk_out, nbins, norm, c->df_coef, s);
    launch_pk_adjoint(n, c->df_half, c->pk_bin, c->df_coef, s);
  });
  c2r3d(c, c->df_half, c->df_t);  // df_t = dL/dT
  launch(c, "seed_a", [&] {
    launch_seed_a(n3, c->df_t, c->df_delta, c->df_x, c->df_fq, c->df_fa, (float)(gx / (double)n3),
                  c->df_c, c->df_qbar, c->df_dbar, s);
  });
  launch(c, "c2c_rows", [&] { launch_c2c_rows(n, -1, c->df_c, c->tw, s); });
  launch(c, "cols_fwd_j", [&] { launch_cols(n, n, 0, -1, c->df_c, c->df_c, 1, 0, c->tw, s, 16, kColsStream); });
  // sweep 2: G = sum_R W_R FFT(seed_R)
  launch(c, "diff_fill", [&] {
    launch_fill_c(c->df_g, n3, s);
    launch_fill_d(c->df_rho, rows_blocks(n), s);
  });
  for (int rb = 0; rb < nr; rb += c->K) {
    const int kk = std::min(c->K, nr - rb);
    const int kp = (kk + 1) / 2;
    launch(c, "ball_i", [&] {
      launch_ball_i(n, false, c->df_c, 0, c->df_v, n3, kk, c->df_ball + (long)rb * h3, c->tw, s);
    });
    launch(c, "cols_inv_j", [&] { launch_cols(n, n, 0, +1, c->df_v, c->df_v, kp, n3, c->tw, s, c->cols_tile, c->cols_wide | kColsStream); });
    if (!c->df_pairs) {  // recompute the batch's filtered pairs
      launch(c, "filter_i", [&] {
        launch_filter_i(n, c->spec, c->ubuf, kk, n3, c->wlut + (long)rb * c->mlut, c->mlut,
                        c->d_kappa + rb, c->tw, s, c->tile_filter, c->filter_w, c->filter_pair);
      });
      launch(c, "cols_inv_j", [&] { launch_cols(n, n, 0, +1, c->ubuf, c->ubuf, kk, n3, c->tw, s, c->cols_tile, c->cols_wide | kColsStream); });
    }
    launch(c, "rows_seed", [&] {
      launch_rows_seed(n, c->ubuf, c->df_v, kk, n3, n3, c->bal, c->df_st, inv_tau, c->df_rho, c->tw,
                       c->df_pairs ? c->df_pairs + (long)rb * n3 : nullptr, n3, s);
    });
    launch(c, "cols_fwd_j", [&] { launch_cols(n, n, 0, -1, c->ubuf, c->ubuf, kk, n3, c->tw, s, c->cols_tile, c->cols_wide | kColsStream); });
    launch(c, "filter_acc_i", [&] {
      launch_filter_acc_i(n, c->ubuf, n3, kk, c->wlut + (long)rb * c->mlut, c->mlut, c->d_kappa + rb,
                          c->filter_w, c->df_g, c->tw, s);
    });
  }
  launch(c, "cols_inv_j", [&] { launch_cols(n, n, 0, +1, c->df_g, c->df_g, 1, 0, c->tw, s, 16, kColsStream); });
  launch(c, "c2c_rows", [&] { launch_c2c_rows(n, +1, c->df_g, c->tw, s); });
  launch(c, "grad_cells", [&] {
    launch_sum_rho(c->df_rho, rows_blocks(n), c->df_st, s);
    launch_grad_cells(n3, c->df_g, c->df_qbar, c->df_rec, c->df_ion, emis, c->df_delta, c->df_dbar,
                      c->df_st, c->df_dd, c->df_part, s);
    launch_grad_g(n3, c->df_delta, c->df_dd, c->df_st, grad_g, s);
  });
}

// The same backward on a generic side: full n^3 complex boxes, 3D transforms
// by the line passes (gen_fft3), the per-radius steps one radius pair at a
// time (phi and the ball products share a complex transform per pair), the
// per-radius fields always recomputed from the forward spectrum Z.
void enqueue_diff_backward_generic(sf21_hii_ctx* c, const float* g, const float* emis,
                                   double zeta, double tau, int nbins, double gx, float* grad_g) {
  hipStream_t s = c->stream;
  const int n = c->n;
  const long n3 = c->n3;
  const int nr = (int)c->radii.size();
  const long h = n / 2 + 1, h3 = h * h * h;
  const float inv_tau = (float)(1.0 / tau);
  const CellConsts k = cell_consts(c, 0.0);  // direct inputs
  const GenPlan& gp = c->gplan;
  const float2 none = make_float2(0.f, 0.f);
  enqueue_diff_fields(c, g, emis, zeta);
  // forward spectrum Z of rec + i s ion
  launch(c, "balance", [&] { launch_balance(n3, c->df_rec, c->df_ion, nullptr, k, c->bal_part, c->bal, s); });
  launch(c, "gen_pack", [&] { launch_gen_pack(n3, c->df_rec, c->df_ion, nullptr, k, c->bal, c->spec, s); });
  gen_fft3(c, c->spec, -1);
  // the real-space pair (N R_R, s N I_R) of radius r into ubuf
  auto pair_of = [&](int r) {
    launch(c, "gen_filter_i", [&] {
      launch_gen_pass(gp, 2, +1, 1, c->spec, c->ubuf, c->tw, c->kap[r], nullptr, nullptr, 0, s);
    });
    launch(c, "gen_inv_j", [&] { launch_gen_pass(gp, 1, +1, 0, c->ubuf, c->ubuf, c->tw, none, nullptr, nullptr, 0, s); });
    launch(c, "gen_inv_p", [&] { launch_gen_pass(gp, 0, +1, 0, c->ubuf, c->ubuf, c->tw, none, nullptr, nullptr, 0, s); });
  };
  // sweep 1: C = sum_R B_R FFT(phi_R)
  launch(c, "diff_fill", [&] { launch_fill_c(c->df_c, n3, s); });
  for (int r = 0; r < nr; r += 2) {
    const bool two = r + 1 < nr;
    for (int hh = 0; hh < (two ? 2 : 1); ++hh) {
      pair_of(r + hh);
      launch(c, "gen_phi", [&] { launch_gen_phi(n3, c->ubuf, c->df_v, hh, c->bal, c->df_st, inv_tau, s); });
    }
    gen_fft3(c, c->df_v, -1);
    launch(c, "gen_ball_acc", [&] {
      launch_gen_ball(n, true, c->df_v, c->df_c, c->df_ball + (long)r * h3,
                      two ? c->df_ball + (long)(r + 1) * h3 : nullptr, s);
    });
  }
  gen_fft3(c, c->df_c, +1);
  // relaxed state, its P(k) and the observables' adjoints
  launch(c, "relaxed_state", [&] {
    launch_relaxed_state(n3, c->df_c, c->df_rec, c->df_ion, c->df_delta, c->df_x, c->df_t,
                         c->df_fq, c->df_fa, c->df_part, c->df_st, s);
  });
  enqueue_pk(c, c->df_t, c->df_half, nbins);
  const double dl = c->dx;
  const double norm = dl * dl * dl / ((double)n3);
  launch(c, "pk_adjoint", [&] {
    launch_pk_coef(c->df_gp, c->pk_out, nbins, norm, c->df_coef, s);
    launch_pk_adjoint_full(n, c->df_half, c->pk_bin, c->df_coef, s);
  });
  gen_fft3(c, c->df_half, +1);
  launch(c, "real_part", [&] { launch_real_part(n3, c->df_half, c->df_t, s); });  // dL/dT
  launch(c, "seed_a", [&] {
    launch_seed_a(n3, c->df_t, c->df_delta, c->df_x, c->df_fq, c->df_fa, (float)(gx / (double)n3),
                  c->df_c, c->df_qbar, c->df_dbar, s);
  });
  gen_fft3(c, c->df_c, -1);
  // sweep 2: G = sum_R W_R FFT(seed_R)
  const int nrho = gen_seed_blocks(n3);
  launch(c, "diff_fill", [&] {
    launch_fill_c(c->df_g, n3, s);
    launch_fill_d(c->df_rho, nrho, s);
  });
  for (int r = 0; r < nr; r += 2) {
    const bool two = r + 1 < nr;
    launch(c, "gen_ball", [&] {
      launch_gen_ball(n, false, c->df_c, c->df_v, c->df_ball + (long)r * h3,
                      two ? c->df_ball + (long)(r + 1) * h3 : nullptr, s);
    });
    gen_fft3(c, c->df_v, +1);  // N (phibar_a + i phibar_b)
    for (int hh = 0; hh < (two ? 2 : 1); ++hh) {
      pair_of(r + hh);
      launch(c, "gen_seed", [&] {
        launch_gen_seed(n3, c->ubuf, c->df_v, hh, c->bal, c->df_st, inv_tau, c->df_rho, s);
      });
      gen_fft3(c, c->ubuf, -1);
      launch(c, "gen_filter_acc", [&] { launch_gen_filter_acc(n, c->ubuf, c->df_g, c->kap[r + hh], s); });
    }
  }
  gen_fft3(c, c->df_g, +1);
  launch(c, "grad_cells", [&] {
    launch_sum_rho(c->df_rho, nrho, c->df_st, s);
    launch_grad_cells(n3, c->df_g, c->df_qbar, c->df_rec, c->df_ion, emis, c->df_delta, c->df_dbar,
                      c->df_st, c->df_dd, c->df_part, s);
    launch_grad_g(n3, c->df_delta, c->df_dd, c->df_st, grad_g, s);
  });
}

int diff_check(sf21_hii_ctx* c, const float* g, const float* emis, int nbins) {
  if (!c || !g || !emis || nbins < 1 || nbins > 64) return SF21_ERR_ARG;
  if (c->cfg.input_mode != SF21_INPUT_DIRECT || c->cfg.path != SF21_PATH_FUSED) return SF21_ERR_ARG;
  if (c->n < 64 || !(c->generic || diff_size_supported(c->n))) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  if ((rc = ensure_pk(c, nbins))) return rc;
  if ((rc = ensure_diff(c))) return rc;
  if (c->df_store == 2 && !c->df_pairs && !c->generic) {
    // automatic field storage: keep the pairs when they take at most half
    // of the free device memory
    size_t free_b = 0, total_b = 0;
    const size_t need = sizeof(float2) * (size_t)c->n3 * c->radii.size();
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need <= free_b / 2 &&
        dev_alloc(&c->df_pairs, (size_t)c->n3 * c->radii.size()) != SF21_OK)
      c->df_pairs = nullptr;
    (void)hipGetLastError();
  }
  return SF21_OK;
}

// ---- autotuner: choices of the fused path, persisted per (arch, n, radii) ----
// Only choices that change launch shapes and stage order are searched; the
// top-hat weight mode (in-kernel FP32 or the FP64 table) changes values and
// stays an explicit configuration setting.
struct TuneChoice {
  int tile = 16, K = 16, ct = 16, cw = 0, rw = 0, fp = 0;
};

std::string tune_key(const sf21_hii_ctx* c) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, c->cfg.device) == hipSuccess) {
    arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
  }
  // FNV-1a over the ladder's bytes: a choice is reused only for the same
  // box side, length unit and radii
  unsigned long long hsh = 1469598103934665603ull;
  for (double R : c->radii) {
    unsigned char b[sizeof(double)];
    std::memcpy(b, &R, sizeof(double));
    for (unsigned char x : b) hsh = (hsh ^ x) * 1099511628211ull;
  }
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s n=%d L=%.17g radii=%zu ladder=%016llx", arch.c_str(), c->n,
                c->L, c->radii.size(), hsh);
  return buf;
}

const char* tune_path(const char* given) {
  if (given && *given) return given;
  const char* e = std::getenv("SF21_TUNE_CACHE");
  return (e && *e) ? e : nullptr;
}

// Cache lines: "<key> | tile=T K=K ct=C cw=0..3 rw=0|1 fp=0|1 | <ms> ms";
// the last match wins.
bool tune_lookup(const char* path, const std::string& key, TuneChoice* out) {
  FILE* f = path ? std::fopen(path, "r") : nullptr;
  if (!f) return false;
  char line[512];
  bool found = false;
  while (std::fgets(line, sizeof(line), f)) {
    const char* bar = std::strchr(line, '|');
    if (!bar) continue;
    std::string k(line, (size_t)(bar - line));
    while (!k.empty() && k.back() == ' ') k.pop_back();
    TuneChoice t;
    if (k == key && std::sscanf(bar + 1, " tile=%d K=%d ct=%d cw=%d rw=%d fp=%d", &t.tile,
                                &t.K, &t.ct, &t.cw, &t.rw, &t.fp) == 6) {
      *out = t;
      found = true;
    }
  }
  std::fclose(f);
  return found;
}

void tune_store(const char* path, const std::string& key, const TuneChoice& t, double ms) {
  FILE* f = path ? std::fopen(path, "a") : nullptr;
  if (!f) return;
  std::fprintf(f, "%s | tile=%d K=%d ct=%d cw=%d rw=%d fp=%d | %.3f ms\n", key.c_str(),
               t.tile, t.K, t.ct, t.cw, t.rw, t.fp, ms);
  std::fclose(f);
}

// Sides without a register plan: "<key> | gx=M | <ms> ms", M the XCD-order
// mask of the line passes (GenPlan::xcd); the last match wins.
bool tune_lookup_gen(const char* path, const std::string& key, int* gx) {
  FILE* f = path ? std::fopen(path, "r") : nullptr;
  if (!f) return false;
  char line[512];
  bool found = false;
  while (std::fgets(line, sizeof(line), f)) {
    const char* bar = std::strchr(line, '|');
    if (!bar) continue;
    std::string k(line, (size_t)(bar - line));
    while (!k.empty() && k.back() == ' ') k.pop_back();
    int m = 0;
    if (k == key && std::sscanf(bar + 1, " gx=%d", &m) == 1 && m >= 0 && m <= 6 && (m & 1) == 0) {
      *gx = m;
      found = true;
    }
  }
  std::fclose(f);
  return found;
}

void tune_store_gen(const char* path, const std::string& key, int gx, double ms) {
  FILE* f = path ? std::fopen(path, "a") : nullptr;
  if (!f) return;
  std::fprintf(f, "%s | gx=%d | %.3f ms\n", key.c_str(), gx, ms);
  std::fclose(f);
}

// Called whenever a context's buffers, stream or kernel choices change: its
// own graphs go, and the new generation keeps multi-box graphs cached on
// other contexts from replaying the old state.
void drop_graphs(sf21_hii_ctx* c) {
  for (auto& g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  c->gen++;
}

// Change the radius batch: the K n^3 pair buffer (and the differentiable
// mode's (K+1)/2 n^3 ball-pair buffer) are reallocated.
int set_radius_batch(sf21_hii_ctx* c, int K) {
  if (K > (int)c->radii.size()) K = (int)c->radii.size();
  if (K < 1) return SF21_ERR_ARG;
  if (K == c->K) return SF21_OK;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  drop_graphs(c);
  dev_free(c->ubuf);
  int rc = dev_alloc(&c->ubuf, (size_t)c->n3 * K);
  if (rc) {
    // keep the context usable with its previous batch
    if (dev_alloc(&c->ubuf, (size_t)c->n3 * c->K)) return SF21_ERR_ALLOC;
    return rc;
  }
  c->K = K;
  if (c->df_v) {
    dev_free(c->df_v);
    if ((rc = dev_alloc(&c->df_v, (size_t)(c->K + 1) / 2 * c->n3))) {
      c->df_ready = false;
      return rc;
    }
  }
  return SF21_OK;
}

bool wide_ok(int n) { return n == 128 || n == 256 || n == 512; }


// Column-pass variant that exists for this size and tile: 16 values per
// thread only for wide_ok sizes (tile 16 or 32), prefetching blocks only
// with tile 16 (8 values) or tile 32 (16 values).
int cols_variant(int n, int tile, int cw) {
  const bool wide = (cw & 1) && wide_ok(n) && (tile == 16 || tile == 32);
  const bool pipe = (cw & 2) && (tile == 16 || (wide && tile == 32));
  return (wide ? 1 : 0) | (pipe ? 2 : 0);
}

void apply_choice(sf21_hii_ctx* c, const TuneChoice& t) {
  c->tile_filter = t.tile;
  c->cols_tile = t.ct;
  c->cols_wide = cols_variant(c->n, t.ct, t.cw);
  c->rows_wide = t.rw && wide_ok(c->n);
  c->filter_pair = t.fp && t.tile >= 8 && c->filter_w;
  c->df_kept = false;  // kept fields were computed with the previous choices
  drop_graphs(c);
}

// Deposit-only context: the stream and the parameters the deposit entry
// points read; halo grid and sums are allocated on first use.
int create_deposit(const sf21_hii_config* cfg, sf21_hii_ctx** out) {
  if (cfg->n < 1 || cfg->n_halo < cfg->n || cfg->n_halo > 4096 || !(cfg->box_len > 0))
    return SF21_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    (void)hipGetLastError();
    return SF21_ERR_HIP;
  }
  if (hipSetDevice(cfg->device) != hipSuccess) return SF21_ERR_HIP;
  auto* c = new sf21_hii_ctx();
  c->id = g_next_ctx_id++;
  c->cfg = *cfg;
  c->cfg.radii = nullptr;
  c->cfg.n_radii = 0;
  c->n = cfg->n;
  c->n3 = (long)c->n * c->n * c->n;
  c->nh = (long)c->n * c->n * (c->n / 2 + 1);
  c->L = cfg->box_len;
  c->dx = c->L / c->n;
  c->dk = 2 * kPi / c->L;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SF21_ERR_HIP;
  }
  c->own_stream = true;
  *out = c;
  return SF21_OK;
}

// Entry points that need the transforms and the bubble finder's buffers.
bool is_deposit(const sf21_hii_ctx* c) { return c && c->cfg.path == SF21_PATH_DEPOSIT; }
// Sides without a register plan: the fused finder runs one radius per step
// through the line passes; the autotuner searches their block order, and the
// differentiable mode recomputes its fields there.
bool is_generic(const sf21_hii_ctx* c) { return c && c->generic; }

// Autotuner on a side without a register plan: the XCD-order mask of the
// line passes (block order only: results bit-identical), the current mask
// first, then the others; one warm-up and two timed slices each.
int autotune_generic(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
                     const char* cache_path, double* best_ms) {
  float* xout = nullptr;
  int rc = dev_alloc(&xout, (size_t)c->n3);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    dev_free(xout);
    return SF21_ERR_HIP;
  }
  const bool prof = c->prof;
  c->prof = false;
  std::vector<int> masks{c->gplan.xcd};
  for (int m : {0, 2, 4, 6})
    if (m != c->gplan.xcd) masks.push_back(m);
  int best = c->gplan.xcd;
  double best_t = 1e300;
  for (int m : masks) {
    c->gplan.xcd = m;
    drop_graphs(c);
    enqueue_fused(c, a, b, cc, z, xout, c->neutral_dev, -1, nullptr);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < 2; ++r) enqueue_fused(c, a, b, cc, z, xout, c->neutral_dev, -1, nullptr);
    (void)hipEventRecord(e1, c->stream);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = SF21_ERR_HIP;
      break;
    }
    if (ms / 2 < best_t) {
      best_t = ms / 2;
      best = m;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(xout);
  c->prof = prof;
  c->gplan.xcd = best;
  drop_graphs(c);
  if (rc) return rc;
  tune_store_gen(tune_path(cache_path), tune_key(c), best, best_t);
  if (best_ms) *best_ms = best_t;
  return check_hip();
}

}  // namespace

extern "C" {

int sf21_hii_create(const sf21_hii_config* cfg, sf21_hii_ctx** out) {
  if (!cfg || !out) return SF21_ERR_ARG;
  *out = nullptr;
  if (cfg->path == SF21_PATH_DEPOSIT) return create_deposit(cfg, out);
  if (!cfg->radii || cfg->n_radii < 1 || !(cfg->box_len > 0)) return SF21_ERR_ARG;
  if (!size_supported(cfg->n, cfg->path)) return SF21_ERR_ARG;
  if (cfg->input_mode != SF21_INPUT_SFR && cfg->input_mode != SF21_INPUT_DIRECT &&
      cfg->input_mode != SF21_INPUT_HALO)
    return SF21_ERR_ARG;
  if (cfg->input_mode == SF21_INPUT_HALO && (cfg->n_halo < cfg->n || cfg->n_halo > 4096))
    return SF21_ERR_ARG;
  if (cfg->path != SF21_PATH_FUSED && cfg->path != SF21_PATH_PLAIN) return SF21_ERR_ARG;
  // the ladder is descending (get_HIIbubbles.c:298-461 divides R each step):
  // the fused path walks it in ascending batches and stores each batch's
  // ball bound over the previous one, which holds only in that order
  for (int r = 0; r < cfg->n_radii; ++r)
    if (!(cfg->radii[r] > 0) || (r > 0 && !(cfg->radii[r] <= cfg->radii[r - 1])))
      return SF21_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
    (void)hipGetLastError();
    return SF21_ERR_HIP;
  }
  if (hipSetDevice(cfg->device) != hipSuccess) return SF21_ERR_HIP;
  auto* c = new sf21_hii_ctx();
  c->id = g_next_ctx_id++;
  c->cfg = *cfg;
  c->radii.assign(cfg->radii, cfg->radii + cfg->n_radii);
  c->cfg.radii = c->radii.data();
  c->n = cfg->n;
  c->n3 = (long)c->n * c->n * c->n;
  c->nh = (long)c->n * c->n * (c->n / 2 + 1);
  c->L = cfg->box_len;
  c->dx = c->L / c->n;
  c->dk = 2 * kPi / c->L;
  c->generic = cfg->path == SF21_PATH_FUSED && !plan_size(c->n) && c->n != 32;
  if (c->generic && (!gen_plan_make(c->n, &c->gplan) || !gen_prepare(c->gplan))) {
    (void)hipGetLastError();
    delete c;
    return SF21_ERR_HIP;
  }
  int rc = SF21_OK;
  long wmax = -1;
  for (double R : c->radii) {
    long d = sf21_ball_d2max(R, c->dx);
    c->d2.push_back((int)d);
    if (d > wmax) wmax = d;
  }
  c->h = (int)std::floor(std::sqrt((double)wmax));
  while ((long)(c->h + 1) * (c->h + 1) <= wmax) c->h++;
  while ((long)c->h * c->h > wmax) c->h--;
  if (c->h > c->n / 2) c->h = c->n / 2;
  c->K = cfg->radius_batch > 0 ? cfg->radius_batch : 16;
  bool tuned = false;
  TuneChoice tc;
  if (cfg->path == SF21_PATH_FUSED && !c->generic && cfg->radius_batch <= 0 &&
      cfg->filter_tile == 0 && cfg->filter_w == 0)
    tuned = tune_lookup(tune_path(nullptr), tune_key(c), &tc);
  if (tuned) c->K = tc.K;
  if (c->generic && cfg->filter_tile == 0 && cfg->filter_w == 0) {
    int gx = 0;
    if (tune_lookup_gen(tune_path(nullptr), tune_key(c), &gx)) c->gplan.xcd = gx;
  }
  if (cfg->filter_tile == 4 || cfg->filter_tile == 8 || cfg->filter_tile == 16)
    c->tile_filter = cfg->filter_tile;
  c->filter_w = cfg->filter_w == 2 ? 0 : 1;
  // untuned: mirror-paired filter blocks where the pairs still fill the GPU
  c->filter_pair = c->n >= 256 && c->tile_filter >= 8 && c->filter_w;
  if (tuned) {
    c->tile_filter = tc.tile;
    c->cols_tile = tc.ct;
    c->cols_wide = cols_variant(c->n, tc.ct, tc.cw);
    c->rows_wide = tc.rw && wide_ok(c->n);
    c->filter_pair = tc.fp && tc.tile >= 8 && c->filter_w;
  }
  if (c->K > (int)c->radii.size()) c->K = (int)c->radii.size();
  if (c->generic) c->K = 1;  // one radius per step, in-kernel FP32 weights
  if (c->generic) c->filter_w = 1;

  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SF21_ERR_HIP;
  }
  c->own_stream = true;
  // tables
  if ((rc = make_twiddles(c->n, &c->tw)) || (rc = make_twiddles(c->n / 2, &c->twh))) goto fail;
  {
    const int hh = c->n / 2;
    c->mlut = 3L * hh * hh + 1;
    std::vector<float> w((size_t)c->mlut * c->radii.size());
    for (size_t r = 0; r < c->radii.size(); ++r)
      for (long m = 0; m < c->mlut; ++m)
        w[r * c->mlut + m] = (float)w_filter(c->dk * std::sqrt((double)m) * c->radii[r]);
    if ((rc = dev_alloc(&c->wlut, w.size()))) goto fail;
    if (hipMemcpy(c->wlut, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice) != hipSuccess) {
      rc = SF21_ERR_HIP;
      goto fail;
    }
    std::vector<float2> kap(c->radii.size());
    for (size_t r = 0; r < c->radii.size(); ++r) {
      const double k = c->dk * c->radii[r];
      const float hi = (float)k;
      kap[r] = make_float2(hi, (float)(k - (double)hi));
    }
    c->kap = kap;
    if ((rc = dev_alloc(&c->d_kappa, kap.size()))) goto fail;
    if (hipMemcpy(c->d_kappa, kap.data(), sizeof(float2) * kap.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
      rc = SF21_ERR_HIP;
      goto fail;
    }
    if ((rc = dev_alloc(&c->d_d2, c->d2.size()))) goto fail;
    if (hipMemcpy(c->d_d2, c->d2.data(), sizeof(int) * c->d2.size(), hipMemcpyHostToDevice) !=
        hipSuccess) {
      rc = SF21_ERR_HIP;
      goto fail;
    }
  }
  c->neutral_cap = 256;
  if ((rc = dev_alloc(&c->neutral_dev, (size_t)c->neutral_cap))) goto fail;
  if (hipHostMalloc(reinterpret_cast<void**>(&c->neutral_host), sizeof(double) * c->neutral_cap) !=
      hipSuccess) {
    rc = SF21_ERR_ALLOC;
    goto fail;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->flag_host), sizeof(int)) != hipSuccess) {
    rc = SF21_ERR_ALLOC;
    goto fail;
  }
  // fused-path buffers (also used by the filtered-field inspection hook)
  {
    c->dt.n = c->n;
    c->dt.h = c->h;
    c->dt.q = dt_q(c->n);
    c->npart = dt_partials_count(c->n);
    // the plain path uses the batch buffer only for the inspection hook
    if (cfg->path == SF21_PATH_PLAIN) c->K = 1;
    if ((rc = dev_alloc(&c->spec, (size_t)c->n3))) goto fail;
    // automatic batch: halve K until the K n^3 complex buffer fits
    while ((rc = dev_alloc(&c->ubuf, (size_t)c->n3 * c->K)) && cfg->radius_batch <= 0 && c->K > 1)
      c->K /= 2;
    if (rc) goto fail;
    if ((rc = dev_alloc(&c->wmap, (size_t)c->n3)) || (rc = dev_alloc(&c->gbuf, (size_t)c->n3)) ||
        (rc = dev_alloc(&c->bal_part, (size_t)2 * balance_partials_count(c->n3))) ||
        (rc = dev_alloc(&c->bal, 1)) ||
        (rc = dev_alloc(&c->partials, (size_t)std::max(c->npart, clamp_partials_count(c->n3)))))
      goto fail;
  }
  if (cfg->path == SF21_PATH_PLAIN) {
    if ((rc = dev_alloc(&c->rec, (size_t)c->n3)) || (rc = dev_alloc(&c->ion, (size_t)c->n3)) ||
        (rc = dev_alloc(&c->fres, (size_t)c->n3)) || (rc = dev_alloc(&c->bubblef, (size_t)c->n3)) ||
        (rc = dev_alloc(&c->box_a, (size_t)c->n3)) || (rc = dev_alloc(&c->box_b, (size_t)c->n3)) ||
        (rc = dev_alloc(&c->bubble, (size_t)c->n3)) || (rc = dev_alloc(&c->spec2, (size_t)(2 * c->nh))) ||
        (rc = dev_alloc(&c->tmp2, (size_t)(2 * c->nh))) || (rc = dev_alloc(&c->th, (size_t)c->nh)) ||
        (rc = dev_alloc(&c->flag_dev, 1)))
      goto fail;
  }
  *out = c;
  return SF21_OK;
fail:
  free_all(c);
  delete c;
  return rc;
}

int sf21_hii_destroy(sf21_hii_ctx* c) {
  if (!c) return SF21_ERR_ARG;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return SF21_OK;
}

int sf21_hii_set_stream(sf21_hii_ctx* c, void* stream) {
  if (!c) return SF21_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  c->own_stream = false;
  c->stream = reinterpret_cast<hipStream_t>(stream);
  return SF21_OK;
}

int sf21_hii_run_slice_dev(sf21_hii_ctx* c, const float* a, const float* b, const float* cc,
                           double z, float* xhii, double* neutral) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !a || !b || !xhii) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  if (c->cfg.path == SF21_PATH_PLAIN) {
    rc = run_plain(c, a, b, cc, z, xhii, neutral);
  } else {
    enqueue_fused(c, a, b, cc, z, xhii, c->neutral_dev, -1, nullptr);
    if (hipMemcpyAsync(c->neutral_host, c->neutral_dev, sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess)
      return SF21_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
    if (neutral) *neutral = c->neutral_host[0];
    rc = check_hip();
  }
  if (c->prof) collect(c);
  return rc;
}

int sf21_hii_run_slice(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
                       float* xhii, double* neutral) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !a || !b || !xhii) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  const size_t bytes = sizeof(float) * (size_t)c->n3;
  float *da = nullptr, *db = nullptr, *dc = nullptr, *dx = nullptr;
  int rc = SF21_OK;
  if ((rc = dev_alloc(&da, c->n3)) || (rc = dev_alloc(&db, c->n3)) || (rc = dev_alloc(&dx, c->n3)) ||
      (cc && (rc = dev_alloc(&dc, c->n3)))) {
    dev_free(da); dev_free(db); dev_free(dc); dev_free(dx);
    return rc;
  }
  if (hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      (cc && hipMemcpy(dc, cc, bytes, hipMemcpyHostToDevice) != hipSuccess))
    rc = SF21_ERR_HIP;
  if (!rc) rc = sf21_hii_run_slice_dev(c, da, db, dc, z, dx, neutral);
  if (!rc && hipMemcpy(xhii, dx, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SF21_ERR_HIP;
  dev_free(da); dev_free(db); dev_free(dc); dev_free(dx);
  return rc;
}

int sf21_hii_run_batch_dev(sf21_hii_ctx* c, int ns, const float* const* a, const float* const* b,
                           const float* const* cc, const double* z, float* const* xhii,
                           double* neutral) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || ns < 1 || !a || !b || !z || !xhii || ns > c->neutral_cap) return SF21_ERR_ARG;
  for (int s = 0; s < ns; ++s)
    if (!a[s] || !b[s] || !xhii[s]) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  if (c->cfg.path == SF21_PATH_PLAIN) {
    for (int s = 0; s < ns; ++s) {
      int rc = run_plain(c, a[s], b[s], cc ? cc[s] : nullptr, z[s], xhii[s], neutral ? &neutral[s] : nullptr);
      if (rc) return rc;
    }
    if (c->prof) collect(c);
    return SF21_OK;
  }
  if (c->prof) {
    for (int s = 0; s < ns; ++s)
      enqueue_fused(c, a[s], b[s], cc ? cc[s] : nullptr, z[s], xhii[s], c->neutral_dev + s, -1,
                    nullptr);
  } else {
    // one hipGraph per distinct (pointers, redshifts) tuple, captured once
    std::vector<const void*> key;
    std::vector<double> zs(z, z + ns);
    for (int s = 0; s < ns; ++s) {
      key.push_back(a[s]);
      key.push_back(b[s]);
      key.push_back(cc ? cc[s] : nullptr);
      key.push_back(xhii[s]);
    }
    GraphEntry* ge = nullptr;
    for (auto& g : c->graphs)
      if (g.ptrs == key && g.zs == zs) ge = &g;
    if (!ge) {
      hipGraph_t graph = nullptr;
      if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return SF21_ERR_HIP;
      c->capturing = true;
      for (int s = 0; s < ns; ++s)
        enqueue_fused(c, a[s], b[s], cc ? cc[s] : nullptr, z[s], xhii[s], c->neutral_dev + s, -1,
                      nullptr);
      c->capturing = false;
      if (hipStreamEndCapture(c->stream, &graph) != hipSuccess) return SF21_ERR_HIP;
      GraphEntry g;
      g.ptrs = key;
      g.zs = zs;
      hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      c->captures++;
      if (e != hipSuccess) return SF21_ERR_HIP;
      if (c->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(c->graphs.front().exec);
        c->graphs.erase(c->graphs.begin());
      }
      c->graphs.push_back(g);
      ge = &c->graphs.back();
    }
    if (hipGraphLaunch(ge->exec, c->stream) != hipSuccess) return SF21_ERR_HIP;
  }
  if (hipMemcpyAsync(c->neutral_host, c->neutral_dev, sizeof(double) * ns, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (neutral)
    for (int s = 0; s < ns; ++s) neutral[s] = c->neutral_host[s];
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_run_multi_dev(int nbox, sf21_hii_ctx* const* ctx, const float* const* a,
                           const float* const* b, const float* const* cc, const double* z,
                           float* const* xhii, double* neutral) {
  if (nbox < 1 || !ctx || !a || !b || !z || !xhii) return SF21_ERR_ARG;
  sf21_hii_ctx* c0 = ctx[0];
  // per-context neutral slots, in box order
  std::vector<sf21_hii_ctx*> uniq;
  std::vector<int> slot(nbox);
  for (int i = 0; i < nbox; ++i) {
    sf21_hii_ctx* c = ctx[i];
    if (!c || !a[i] || !b[i] || !xhii[i] || c->cfg.path != SF21_PATH_FUSED ||
        c->cfg.device != c0->cfg.device)
      return SF21_ERR_ARG;
    int used = 0;
    for (int k = 0; k < i; ++k) used += ctx[k] == c;
    if (used >= c->neutral_cap) return SF21_ERR_ARG;
    slot[i] = used;
    if (std::find(uniq.begin(), uniq.end(), c) == uniq.end()) uniq.push_back(c);
  }
  if (hipSetDevice(c0->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  hipStream_t s = c0->stream;
  std::vector<hipStream_t> saved;
  for (auto* c : uniq) saved.push_back(c->stream);
  bool prof = false;
  for (auto* c : uniq) prof = prof || c->prof;
  int rc = SF21_OK;
  // boxes of one context run in box order on that context's stream (they
  // share its buffers); different contexts share nothing and run as
  // parallel branches of the graph
  auto enqueue_ctx = [&](sf21_hii_ctx* c) {
    for (int i = 0; i < nbox; ++i)
      if (ctx[i] == c)
        enqueue_fused(c, a[i], b[i], cc ? cc[i] : nullptr, z[i], xhii[i], c->neutral_dev + slot[i],
                      -1, nullptr);
  };
  if (prof) {
    // profiled calls: every context on the first one's stream, in box order
    for (auto* c : uniq) c->stream = s;
    for (int i = 0; i < nbox; ++i)
      enqueue_fused(ctx[i], a[i], b[i], cc ? cc[i] : nullptr, z[i], xhii[i],
                    ctx[i]->neutral_dev + slot[i], -1, nullptr);
  } else {
    // one hipGraph for the whole set, cached on the first context by
    // (contexts, pointers, redshifts)
    std::vector<const void*> key;
    std::vector<double> zs(z, z + nbox);
    std::vector<unsigned long long> ids;
    for (int i = 0; i < nbox; ++i) {
      ids.push_back(ctx[i]->id);
      ids.push_back(ctx[i]->gen);
      key.push_back(a[i]);
      key.push_back(b[i]);
      key.push_back(cc ? cc[i] : nullptr);
      key.push_back(xhii[i]);
    }
    GraphEntry* ge = nullptr;
    for (auto& g : c0->graphs)
      if (g.ptrs == key && g.zs == zs && g.ids == ids) ge = &g;
    if (!ge) {
      hipGraph_t graph = nullptr;
      std::vector<hipEvent_t> ev(uniq.size() + 1, nullptr);
      for (auto& e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) rc = SF21_ERR_HIP;
      if (!rc && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = SF21_ERR_HIP;
      } else if (!rc) {
        for (auto* c : uniq) c->capturing = true;
        // fork: every other context's stream joins the capture after s's
        // fork point, runs its boxes, and s waits for it before the end
        hipError_t e = hipEventRecord(ev[0], s);
        for (size_t k = 1; k < uniq.size() && e == hipSuccess; ++k)
          e = hipStreamWaitEvent(uniq[k]->stream, ev[0], 0);
        if (e == hipSuccess) {
          for (auto* c : uniq) enqueue_ctx(c);
          for (size_t k = 1; k < uniq.size() && e == hipSuccess; ++k) {
            e = hipEventRecord(ev[k], uniq[k]->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, ev[k], 0);
          }
        }
        for (auto* c : uniq) c->capturing = false;
        if (hipStreamEndCapture(s, &graph) != hipSuccess || e != hipSuccess) rc = SF21_ERR_HIP;
      }
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      if (!rc) {
        GraphEntry g;
        g.ptrs = key;
        g.zs = zs;
        g.ids = ids;
        hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        c0->captures++;
        if (e != hipSuccess) {
          rc = SF21_ERR_HIP;
        } else {
          if (c0->graphs.size() >= 8) {
            (void)hipGraphExecDestroy(c0->graphs.front().exec);
            c0->graphs.erase(c0->graphs.begin());
          }
          c0->graphs.push_back(g);
          ge = &c0->graphs.back();
        }
      }
    }
    if (!rc && hipGraphLaunch(ge->exec, s) != hipSuccess) rc = SF21_ERR_HIP;
  }
  if (!rc) {
    for (auto* c : uniq) {
      int used = 0;
      for (int i = 0; i < nbox; ++i) used += ctx[i] == c;
      if (hipMemcpyAsync(c->neutral_host, c->neutral_dev, sizeof(double) * used,
                         hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = SF21_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = SF21_ERR_HIP;
  }
  for (size_t k = 0; k < uniq.size(); ++k) uniq[k]->stream = saved[k];
  if (rc) return rc;
  if (neutral)
    for (int i = 0; i < nbox; ++i) neutral[i] = ctx[i]->neutral_host[slot[i]];
  for (auto* c : uniq)
    if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_weight_map(sf21_hii_ctx* c, int* host_out) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !host_out) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (hipMemcpy(host_out, c->wmap, sizeof(int) * c->n3, hipMemcpyDeviceToHost) != hipSuccess)
    return SF21_ERR_HIP;
  return SF21_OK;
}

int sf21_hii_dilate_dev(sf21_hii_ctx* c, const int* host_map, const float* a, const float* b,
                        const float* cc, double z, float* xhii, double* neutral) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !host_map || !a || !b || !xhii || c->cfg.path != SF21_PATH_FUSED) return SF21_ERR_ARG;
  // the line passes' reach cap h = min(floor(sqrt(largest d2max)), n/2) is
  // exact only for bounds up to the ladder's largest
  const int wmax = *std::max_element(c->d2.begin(), c->d2.end());
  for (long i = 0; i < c->n3; ++i)
    if (host_map[i] < -1 || host_map[i] > wmax) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  hipStream_t s = c->stream;
  const long n3 = c->n3;
  const CellConsts k = cell_consts(c, z);
  if (hipMemcpyAsync(c->wmap, host_map, sizeof(int) * n3, hipMemcpyHostToDevice, s) != hipSuccess)
    return SF21_ERR_HIP;
  launch(c, "dt_rows", [&] { launch_dt_rows(c->dt, c->wmap, c->gbuf, s); });
  launch(c, "dt_cols", [&] { launch_dt_cols(c->dt, c->gbuf, c->gbuf, s); });
  launch(c, "dt_final", [&] {
    launch_dt_final(c->dt, c->gbuf, a, b, cc, k, xhii, c->partials, c->npart, s);
  });
  launch(c, "sum", [&] {
    launch_sum_partials(c->partials, c->npart, 1.0 / (double)n3, c->neutral_dev, s);
  });
  if (hipMemcpyAsync(c->neutral_host, c->neutral_dev, sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return SF21_ERR_HIP;
  if (neutral) *neutral = c->neutral_host[0];
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_graph_captures(const sf21_hii_ctx* c, long* captures) {
  if (!c || !captures) return SF21_ERR_ARG;
  *captures = c->captures;
  return SF21_OK;
}

int sf21_hii_filtered_dev(sf21_hii_ctx* c, const float* a, const float* b, const float* cc,
                          double z, int radius_index, float* out) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !a || !b || !out || radius_index < 0 || radius_index >= (int)c->radii.size())
    return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  enqueue_fused(c, a, b, cc, z, nullptr, nullptr, radius_index, reinterpret_cast<float2*>(out));
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_autotune(sf21_hii_ctx* c, const float* a, const float* b, const float* cc, double z,
                      const char* cache_path, double* best_ms) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !a || !b || c->cfg.path != SF21_PATH_FUSED) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  if (is_generic(c)) return autotune_generic(c, a, b, cc, z, cache_path, best_ms);
  float* xout = nullptr;
  int rc = dev_alloc(&xout, (size_t)c->n3);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    dev_free(xout);
    return SF21_ERR_HIP;
  }
  const bool prof = c->prof;
  c->prof = false;
  // time one candidate: 1 warm-up + 2 timed slices
  auto timed = [&](const TuneChoice& t, double* ms_out) -> int {
    if (set_radius_batch(c, t.K)) return 1;  // does not fit: skip
    apply_choice(c, t);
    enqueue_fused(c, a, b, cc, z, xout, c->neutral_dev, -1, nullptr);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < 2; ++r) enqueue_fused(c, a, b, cc, z, xout, c->neutral_dev, -1, nullptr);
    (void)hipEventRecord(e1, c->stream);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return -1;
    *ms_out = ms / 2;
    return 0;
  };
  // coordinate search from the defaults: each dimension in turn, the others
  // held at the best so far
  const int nr = (int)c->radii.size();
  TuneChoice best;
  best.K = std::min(16, nr);
  double best_t = 1e300;
  {
    double ms;
    int r0 = timed(best, &ms);
    if (r0 < 0) rc = SF21_ERR_HIP;
    else if (r0 == 0) best_t = ms;
  }
  auto consider = [&](TuneChoice t) {
    if (rc) return;
    double ms;
    int r1 = timed(t, &ms);
    if (r1 < 0) rc = SF21_ERR_HIP;
    else if (r1 == 0 && ms < best_t) {
      best_t = ms;
      best = t;
    }
  };
  // larger batches read the spectrum fewer times per slice; candidates that
  // do not fit in device memory are skipped
  for (int K : {8, 12, 24, 32, nr}) {
    TuneChoice t = best;
    t.K = std::min(K, nr);
    if (t.K != best.K) consider(t);
  }
  {
    TuneChoice t = best;
    t.tile = 8;
    consider(t);
  }
  {
    TuneChoice t = best;
    t.fp = 1 - t.fp;
    consider(t);
  }
  {
    const TuneChoice base = best;
    const int cols_opts[7][2] = {{8, 0}, {16, 0}, {16, 2}, {16, 1}, {16, 3}, {32, 1}, {32, 3}};
    for (auto& o : cols_opts) {
      if ((o[1] & 1) && !wide_ok(c->n)) continue;
      TuneChoice t = base;
      t.ct = o[0];
      t.cw = o[1];
      if (t.ct != base.ct || t.cw != base.cw) consider(t);
    }
  }
  if (wide_ok(c->n)) {
    TuneChoice t = best;
    t.rw = 1;
    consider(t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(xout);
  c->prof = prof;
  if (rc) return rc;
  if (best_t >= 1e300) return SF21_ERR_ALLOC;
  if ((rc = set_radius_batch(c, best.K))) return rc;
  apply_choice(c, best);
  tune_store(tune_path(cache_path), tune_key(c), best, best_t);
  if (best_ms) *best_ms = best_t;
  return check_hip();
}

int sf21_hii_tuning(const sf21_hii_ctx* c, int* filter_tile, int* radius_batch, int* filter_w,
                    int* cols_tile, int* cols_wide, int* rows_wide, int* filter_pair) {
  if (!c) return SF21_ERR_ARG;
  if (filter_pair) *filter_pair = c->filter_pair;
  if (filter_tile) *filter_tile = c->tile_filter;
  if (radius_batch) *radius_batch = c->K;
  if (filter_w) *filter_w = c->filter_w ? 1 : 2;
  if (cols_tile) *cols_tile = c->cols_tile;
  if (cols_wide) *cols_wide = c->cols_wide;
  if (rows_wide) *rows_wide = c->rows_wide ? 1 : 0;
  return SF21_OK;
}

int sf21_hii_radius_batch(const sf21_hii_ctx* c) { return c ? c->K : SF21_ERR_ARG; }

int sf21_hii_line_order(const sf21_hii_ctx* c) {
  return is_generic(c) ? c->gplan.xcd : SF21_ERR_ARG;
}

int sf21_hii_profile(sf21_hii_ctx* c, int enable) {
  if (!c) return SF21_ERR_ARG;
  c->prof = enable != 0;
  return SF21_OK;
}

int sf21_hii_kernel_stats(sf21_hii_ctx* c, int index, char* name, int cap, long* launches,
                          double* total_ms) {
  if (!c) return SF21_ERR_ARG;
  if (index < 0) return (int)c->stats.size();
  if (index >= (int)c->stats.size()) return SF21_ERR_ARG;
  if (name && cap > 0) {
    std::strncpy(name, c->stats[index].name.c_str(), (size_t)cap - 1);
    name[cap - 1] = 0;
  }
  if (launches) *launches = c->stats[index].launches;
  if (total_ms) *total_ms = c->stats[index].ms;
  return SF21_OK;
}

int sf21_hii_reset_stats(sf21_hii_ctx* c) {
  if (!c) return SF21_ERR_ARG;
  c->stats.clear();
  return SF21_OK;
}

// ---- binned power spectrum ------------------------------------------------
int sf21_hii_power_spectrum_dev(sf21_hii_ctx* c, const float* field, int nbins, double* pk,
                                double* kmean, double* counts) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !field || nbins < 1 || nbins > 64 || c->n < 64) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  if ((rc = ensure_pk(c, nbins))) return rc;
  // the spectrum goes to the context's spectrum buffer (n^3 >= n*n*(n/2+1))
  enqueue_pk(c, field, c->spec, nbins);
  std::vector<double> host(3 * nbins);
  if (hipMemcpyAsync(host.data(), c->pk_out, sizeof(double) * 3 * nbins, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  for (int b = 0; b < nbins; ++b) {
    if (pk) pk[b] = host[b];
    if (kmean) kmean[b] = host[nbins + b];
    if (counts) counts[b] = host[2 * nbins + b];
  }
  if (c->prof) collect(c);
  return check_hip();
}

// ---- halo-catalog emissivity (get_HIIbubbles.c:229-259) -----------------
}  // extern "C"

// Scratch of sf21_hii_halo_emissivity_noalloc: the N_halo^3 CIC deposit
// grid, at most kNoAllocGrid floats (64 MiB: N_halo <= 256), in a
// file-scope device array that every GPU the library runs on holds from
// module load until the process ends.  In host code the array's name is a
// host-side stand-in, not a device address: each device's address comes
// from hipGetSymbolAddress, taken once per device (with that device
// current) on the entry point's first call there, and only that address is
// passed to kernels.  Calls on one GPU share its array: each call zeroes the
// grid before depositing, and calls that overlap on one GPU race.
constexpr long kNoAllocGrid = 1L << 24;
static __device__ __align__(256) float scratch[kNoAllocGrid];

namespace {
// Launch constants of the halo-mode emissivity: smooth_box's block size
// (auxiliary.c:654-656), the volume ratio times fesc (get_HIIbubbles.c:253-258)
// and Rion's redshift factor (user_functions.c:40-47: c1 (1+z)^c4 / V_norm).
struct HaloConsts {
  long sm;
  float scale, k_rion;
};
int halo_consts(const sf21_hii_ctx* c, double z, HaloConsts* k) {
  const int ng = c->cfg.n_halo, ns = c->n;
  k->sm = (long)roundf((1.0f * ng) / ns);
  if ((long)ns * k->sm > ng) k->sm--;
  if (k->sm < 1) return SF21_ERR_ARG;
  const double dx_h = c->L / ng;
  k->scale = (float)(std::pow(c->dx / dx_h, 3) * c->cfg.fesc);
  k->k_rion = (float)(1.575e39 * std::pow(1.0 + z, 2.27044695e+00) / kVnorm);
  return SF21_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

// The device address of this device's scratch array (the device must be
// current).
int scratch_address(int device, float** out) {
  if (device < 0 || device >= kMaxDevices) return SF21_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return SF21_ERR_HIP;
    }
    g_scratch_dev[device] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[device];
  return SF21_OK;
}
}  // namespace

extern "C" {

int sf21_hii_halo_emissivity_dev(sf21_hii_ctx* c, const float* halos, long nhalos, double z,
                                 float* ion_out) {
  if (!c || (!halos && nhalos > 0) || nhalos < 0 || !ion_out) return SF21_ERR_ARG;
  if (c->cfg.input_mode != SF21_INPUT_HALO) return SF21_ERR_ARG;
  HaloConsts k;
  if (halo_consts(c, z, &k)) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  const int ng = c->cfg.n_halo, ns = c->n;
  if (!c->halo_grid) {
    int rc = dev_alloc(&c->halo_grid, (size_t)ng * ng * ng);
    if (rc) return rc;
  }
  launch(c, "halo_emissivity", [&] {
    launch_halo_emissivity(nhalos, halos, ng, ns, (int)k.sm, k.k_rion, (float)9.367608e+07, 0.44f,
                           k.scale, c->halo_grid, ion_out, c->stream);
  });
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_halo_emissivity_noalloc(sf21_hii_ctx* c, const float* halos, long nhalos, double z,
                                     float* ion_out) {
  if (!c || (!halos && nhalos > 0) || nhalos < 0 || !ion_out) return SF21_ERR_ARG;
  if (c->cfg.input_mode != SF21_INPUT_HALO) return SF21_ERR_ARG;
  const long ng = c->cfg.n_halo;
  if (ng * ng * ng > kNoAllocGrid) return SF21_ERR_ARG;
  HaloConsts k;
  if (halo_consts(c, z, &k)) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  float* grid = nullptr;
  int rc = scratch_address(c->cfg.device, &grid);
  if (rc) return rc;
  // enqueued on the context's stream, no host synchronisation (capturable)
  launch_halo_emissivity(nhalos, halos, (int)ng, c->n, (int)k.sm, k.k_rion, (float)9.367608e+07,
                         0.44f, k.scale, grid, ion_out, c->stream);
  return check_hip();
}

int sf21_hii_run_slice_halo(sf21_hii_ctx* c, const float* delta, const float* halos, long nhalos,
                            double z, float* xhii, double* neutral) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !delta || (!halos && nhalos > 0) || nhalos < 0 || !xhii) return SF21_ERR_ARG;
  if (c->cfg.input_mode != SF21_INPUT_HALO) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  const size_t bytes = sizeof(float) * (size_t)c->n3;
  float *dd = nullptr, *dion = nullptr, *dx = nullptr;
  int rc;
  if (nhalos > c->halo_cap) {
    dev_free(c->halo_buf);
    if ((rc = dev_alloc(&c->halo_buf, (size_t)4 * nhalos))) return rc;
    c->halo_cap = nhalos;
  }
  if ((rc = dev_alloc(&dd, c->n3)) || (rc = dev_alloc(&dion, c->n3)) || (rc = dev_alloc(&dx, c->n3))) {
    dev_free(dd); dev_free(dion); dev_free(dx);
    return rc;
  }
  if (hipMemcpy(dd, delta, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      (nhalos > 0 && hipMemcpy(c->halo_buf, halos, sizeof(float) * 4 * nhalos,
                               hipMemcpyHostToDevice) != hipSuccess))
    rc = SF21_ERR_HIP;
  if (!rc) rc = sf21_hii_halo_emissivity_dev(c, c->halo_buf, nhalos, z, dion);
  if (!rc) rc = sf21_hii_run_slice_dev(c, dd, dion, nullptr, z, dx, neutral);
  if (!rc && hipMemcpy(xhii, dx, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SF21_ERR_HIP;
  dev_free(dd); dev_free(dion); dev_free(dx);
  return rc;
}

// ---- 21-cm brightness temperature (t21.c:116-252) -------------------------
int sf21_t21_dev(sf21_hii_ctx* c, const sf21_params* p, double z, const float* delta,
                 const float* xhii, const float* xc, const float* xalpha, const float* tk,
                 float* out, double* mean_t21, double* mean_ts) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || !p || !delta || !xhii || !out || c->n < 64) return SF21_ERR_ARG;
  const bool spin = xc || xalpha || tk;
  if (spin && !(xc && xalpha && tk)) return SF21_ERR_ARG;
  if (!(p->hubble > 0) || !(p->omega_m > 0)) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  const int n = c->n;
  const long n3 = c->n3;
  if (!c->t21_dv && ((rc = dev_alloc(&c->t21_dv, (size_t)n3)) ||
                     (rc = dev_alloc(&c->t21_part, (size_t)2 * t21_partials_count(n3) + 2))))
    return rc;
  // per-redshift scalars in FP64 (auxiliary.c:224-292, t21.c:197-212)
  const double growth = sf21_growth(p, z), dgdt = sf21_dgrowth_dt(p, z), hz = sf21_hubble_z(p, z);
  const double h = p->hubble;
  const double c21 = 23.0 / 1000. * (0.7 / h) * (p->omega_b * h * h / 0.02) *
                     std::sqrt((0.15 / p->omega_m / h / h) * (1. + z) / 10.);
  // -dD/dt / (H D) times the unnormalised transform's 1/N^3 (dx^3/L^3)
  const double vfac = -dgdt / hz / growth / (double)n3;
  const double tcmb = 2.725 * (1. + z);
  if (c->generic) {
    // complex transforms of the real box through the generic line passes
    hipStream_t s = c->stream;
    const float2 none = make_float2(0.f, 0.f);
    launch(c, "to_complex", [&] { launch_to_complex(n3, delta, c->spec, s); });
    for (int axis = 0; axis < 3; ++axis)
      launch(c, "gen_fwd", [&] {
        launch_gen_pass(c->gplan, axis, -1, 0, c->spec, c->spec, c->tw, none, nullptr, nullptr, 0, s);
      });
    launch(c, "vgrad_filter", [&] { launch_vgrad_filter_full(n, c->spec, p->vel_comp, s); });
    for (int axis = 0; axis < 3; ++axis)
      launch(c, "gen_inv", [&] {
        launch_gen_pass(c->gplan, axis, +1, 0, c->spec, c->spec, c->tw, none, nullptr, nullptr, 0, s);
      });
    launch(c, "real_part", [&] { launch_real_part(n3, c->spec, c->t21_dv, s); });
  } else {
    r2c3d(c, delta, c->spec);
    launch(c, "vgrad_filter", [&] { launch_vgrad_filter(n, c->spec, p->vel_comp, c->stream); });
    c2r3d(c, c->spec, c->t21_dv);
  }
  double* sums = c->t21_part + 2 * t21_partials_count(n3);
  launch(c, "t21_cells", [&] {
    launch_t21_cells(n3, delta, xhii, c->t21_dv, xc, xalpha, tk, (float)vfac, (float)c21,
                     (float)tcmb, out, c->t21_part, sums, c->stream);
  });
  double host[2];
  if (hipMemcpyAsync(host, sums, sizeof(host), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  if (mean_t21) *mean_t21 = host[0] / (double)n3;
  if (mean_ts) *mean_ts = spin ? host[1] / (double)n3 : 0.0;
  if (c->prof) collect(c);
  return check_hip();
}

}  // extern "C"

// ---- the step before: SFRD and non-linear density deposits --------------
namespace {
// smooth_box's block size (auxiliary.c:654-656)
long smooth_block(int ng, int ns) {
  long sm = (long)roundf((1.0f * ng) / ns);
  if ((long)ns * sm > ng) sm--;
  return sm;
}
int ensure_halo_grid(sf21_hii_ctx* c) {
  const long ng = c->cfg.n_halo;
  if (c->halo_grid) return SF21_OK;
  return dev_alloc(&c->halo_grid, (size_t)ng * ng * ng);
}
int box_mean(sf21_hii_ctx* c, const float* box, long count, double* mean) {
  if (!c->sum_part) {
    int rc = dev_alloc(&c->sum_part, (size_t)sum_partials_count(c->n3 > count ? c->n3 : count) + 1);
    if (rc) return rc;
  }
  double* tot = c->sum_part + sum_partials_count(c->n3 > count ? c->n3 : count);
  launch_sum_f32(count, box, c->sum_part, tot, c->stream);
  double h = 0.0;
  if (hipMemcpyAsync(&h, tot, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  *mean = h / (double)count;
  return SF21_OK;
}
}  // namespace

extern "C" {

int sf21_sfrd_dev(sf21_hii_ctx* c, const float* halos, long nhalos, double z, float* out,
                  double* mean) {
  if (!c || (!halos && nhalos > 0) || nhalos < 0 || !out) return SF21_ERR_ARG;
  if (c->cfg.n_halo < c->n || c->cfg.n_halo > 4096) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  if ((rc = ensure_halo_grid(c))) return rc;
  const int ng = c->cfg.n_halo, ns = c->n;
  const long sm = smooth_block(ng, ns);
  if (sm < 1) return SF21_ERR_ARG;
  const double dx_h = c->L / ng;
  // sfr = Rion/Qion (user_functions.c:71-74); Msun/yr per halo, then the
  // block average divided by dx_halo^3 (get_SFR.c:126-139)
  const float k_sfr = (float)(1.575e39 * std::pow(1.0 + z, 2.27044695e+00) / kVnorm / fit_Qion(z));
  launch(c, "sfrd", [&] {
    launch_halo_emissivity(nhalos, halos, ng, ns, (int)sm, k_sfr, (float)9.367608e+07, 0.44f,
                           (float)(1.0 / (dx_h * dx_h * dx_h)), c->halo_grid, out, c->stream);
  });
  if (mean && (rc = box_mean(c, out, c->n3, mean))) return rc;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_nldensity_dev(sf21_hii_ctx* c, const sf21_params* p, double z, const float* vx,
                       const float* vy, const float* vz, const float* delta0, float* box_halo,
                       float* out) {
  if (!c || !p || !vx || !vy || !vz || !delta0 || (!out && !box_halo)) return SF21_ERR_ARG;
  if (c->cfg.n_halo < c->n || c->cfg.n_halo > 4096 || !(p->box_len > 0) || !(p->hubble > 0))
    return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  (void)hipGetLastError();
  int rc;
  float* grid = box_halo;
  if (!grid) {
    if ((rc = ensure_halo_grid(c))) return rc;
    grid = c->halo_grid;
  }
  const int ng = c->cfg.n_halo, ns = c->n;
  const long sm = smooth_block(ng, ns);
  if (sm < 1) return SF21_ERR_ARG;
  // get_nldensity.c:95-99, :117, :132: velocities to cells (box_len in
  // Mpc/h, the files in Mpc), the z = 0 mass 1 + delta0 D(300), the
  // displacement factor D(z) - D(300)
  const double growth_in = sf21_growth(p, 300.0);
  const double growth = sf21_growth(p, z) - growth_in;
  const double vscale = (double)ng / p->box_len * p->hubble;
  launch(c, "nldensity", [&] {
    launch_nldensity(ng, ns, (int)sm, vx, vy, vz, delta0, (float)vscale, (float)growth_in,
                     (float)growth, grid, out, c->stream);
  });
  if (hipStreamSynchronize(c->stream) != hipSuccess) return SF21_ERR_HIP;
  if (c->prof) collect(c);
  return check_hip();
}

}  // extern "C"

extern "C" {

// ---- FFTW-like transforms ------------------------------------------------
int sf21_fft_plan_3d(int n, int kind, sf21_fft_plan** out) {
  if (!out || kind < SF21_FFT_R2C || kind > SF21_FFT_C2C_BACKWARD) return SF21_ERR_ARG;
  const bool real = kind == SF21_FFT_R2C || kind == SF21_FFT_C2R;
  if (!plan_size(n) && !(n == 32 && !real)) return SF21_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return SF21_ERR_HIP;
  }
  auto* p = new sf21_fft_plan();
  p->n = n;
  p->kind = kind;
  int rc;
  if ((rc = make_twiddles(n, &p->tw)) || (rc = make_twiddles(n / 2, &p->twh))) {
    dev_free(p->tw);
    dev_free(p->twh);
    delete p;
    return rc;
  }
  *out = p;
  return SF21_OK;
}

int sf21_fft_execute(sf21_fft_plan* p, void* in, void* out, void* stream) {
  if (!p || !in || !out) return SF21_ERR_ARG;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int n = p->n, h = n / 2 + 1;
  (void)hipGetLastError();
  switch (p->kind) {
    case SF21_FFT_R2C: {
      float2* o = reinterpret_cast<float2*>(out);
      launch_r2c_rows(n, reinterpret_cast<const float*>(in), o, p->twh, p->tw, s);
      launch_cols(n, h, 0, -1, o, o, 1, 0, p->tw, s);
      launch_cols(n, h, 1, -1, o, o, 1, 0, p->tw, s);
      break;
    }
    case SF21_FFT_C2R: {
      float2* i2 = reinterpret_cast<float2*>(in);
      launch_cols(n, h, 1, +1, i2, i2, 1, 0, p->tw, s);
      launch_cols(n, h, 0, +1, i2, i2, 1, 0, p->tw, s);
      launch_c2r_rows(n, i2, reinterpret_cast<float*>(out), p->twh, p->tw, s);
      break;
    }
    default: {
      const int sign = p->kind == SF21_FFT_C2C_FORWARD ? -1 : +1;
      float2* o = reinterpret_cast<float2*>(out);
      if (in != out &&
          hipMemcpyAsync(o, in, sizeof(float2) * (size_t)n * n * n, hipMemcpyDeviceToDevice, s) !=
              hipSuccess)
        return SF21_ERR_HIP;
      launch_c2c_rows(n, sign, o, p->tw, s);
      launch_cols(n, n, 0, sign, o, o, 1, 0, p->tw, s);
      launch_cols(n, n, 1, sign, o, o, 1, 0, p->tw, s);
    }
  }
  return check_hip();
}

int sf21_fft_destroy(sf21_fft_plan* p) {
  if (!p) return SF21_ERR_ARG;
  dev_free(p->tw);
  dev_free(p->twh);
  delete p;
  return SF21_OK;
}


// ---- differentiable mode ---------------------------------------------------
int sf21_hii_diff_store_fields(sf21_hii_ctx* c, int mode) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  if (!c || mode < 0 || mode > 3) return SF21_ERR_ARG;
  // generic sides always recompute: automatic (2) stores nothing there
  if (is_generic(c) && mode != 0 && mode != 2) return SF21_ERR_ARG;
  if (hipSetDevice(c->cfg.device) != hipSuccess) return SF21_ERR_HIP;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->df_store = mode;
  c->df_kept = false;
  if (mode == 0) {
    dev_free(c->df_pairs);
    return SF21_OK;
  }
  if (c->df_pairs || mode == 2) return SF21_OK;  // automatic: decided at the next backward
  return dev_alloc(&c->df_pairs, (size_t)c->n3 * c->radii.size());
}

int sf21_hii_diff_last_kept(const sf21_hii_ctx* c) {
  if (!c) return SF21_ERR_ARG;
  return c->df_last_kept ? 1 : 0;
}

int sf21_hii_diff_forward_dev(sf21_hii_ctx* c, const float* g, const float* emis, double zeta,
                              int nbins, float* x_out, double* mean_x, double* pk) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  int rc = diff_check(c, g, emis, nbins);
  if (rc) return rc;
  float* x = x_out ? x_out : c->df_x;
  const bool keep = c->df_store == 3 && c->df_pairs;
  c->df_kept = false;
  enqueue_diff_forward(c, g, emis, zeta, nbins, x, keep);
  if (hipMemcpyAsync(c->df_host, c->df_st, sizeof(double) * DS_COUNT, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess ||
      hipMemcpyAsync(c->df_host + DS_COUNT, c->pk_out, sizeof(double) * nbins,
                     hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      (keep && hipMemcpyAsync(c->df_fp_host, c->df_fp, sizeof(unsigned long long) * 2,
                              hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  if (keep) {
    c->df_kept = true;
    c->df_kept_fp[0] = c->df_fp_host[0];
    c->df_kept_fp[1] = c->df_fp_host[1];
    c->df_kept_zeta = zeta;
  }
  if (mean_x) *mean_x = c->df_host[DS_SUM_X] / (double)c->n3;
  if (pk)
    for (int b = 0; b < nbins; ++b) pk[b] = c->df_host[DS_COUNT + b];
  if (c->prof) collect(c);
  return check_hip();
}

int sf21_hii_diff_backward_dev(sf21_hii_ctx* c, const float* g, const float* emis, double zeta,
                               double tau, int nbins, double dl_dx, const double* dl_dpk,
                               float* grad_g, double* grad_zeta, double* relaxed_mean_x,
                               double* relaxed_pk) {
  if (is_deposit(c)) return SF21_ERR_ARG;
  int rc = diff_check(c, g, emis, nbins);
  if (rc) return rc;
  if (!grad_g || !(tau > 0.0)) return SF21_ERR_ARG;
  for (int b = 0; b < nbins; ++b) c->df_host[b] = dl_dpk ? dl_dpk[b] : 0.0;
  if (hipMemcpyAsync(c->df_gp, c->df_host, sizeof(double) * nbins, hipMemcpyHostToDevice,
                     c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  // forward-kept fields: used when this call's (g, emis, zeta) fingerprint as
  // the last forward's did; after a backward in mode 3 the pairs are this
  // input's either way (sweep 1 rewrites them when it recomputes)
  bool kept = false;
  if (c->df_store == 3 && c->df_pairs) {
    launch_fingerprint(c->n3, g, emis, c->df_part, c->df_fp + 2, c->stream);
    if (hipMemcpyAsync(c->df_fp_host + 2, c->df_fp + 2, sizeof(unsigned long long) * 2,
                       hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return SF21_ERR_HIP;
    kept = c->df_kept && zeta == c->df_kept_zeta && c->df_fp_host[2] == c->df_kept_fp[0] &&
           c->df_fp_host[3] == c->df_kept_fp[1];
    c->df_kept_fp[0] = c->df_fp_host[2];
    c->df_kept_fp[1] = c->df_fp_host[3];
    c->df_kept_zeta = zeta;
  }
  c->df_kept = false;  // set again below once the pairs are this input's
  if (c->generic)
    enqueue_diff_backward_generic(c, g, emis, zeta, tau, nbins, dl_dx, grad_g);
  else
    enqueue_diff_backward(c, g, emis, zeta, tau, nbins, dl_dx, grad_g, kept);
  // the staging buffer is reused for the results only after the stream has
  // consumed the upload above (same stream, ordered)
  if (hipMemcpyAsync(c->df_host + 64, c->df_st, sizeof(double) * DS_COUNT, hipMemcpyDeviceToHost,
                     c->stream) != hipSuccess ||
      hipMemcpyAsync(c->df_host + 64 + DS_COUNT, c->pk_out, sizeof(double) * nbins,
                     hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return SF21_ERR_HIP;
  c->df_kept = c->df_store == 3 && c->df_pairs;
  c->df_last_kept = kept;
  const double* hs = c->df_host + 64;
  if (grad_zeta) *grad_zeta = hs[DS_ZETA_BAR];
  if (relaxed_mean_x) *relaxed_mean_x = hs[DS_SUM_XT] / (double)c->n3;
  if (relaxed_pk)
    for (int b = 0; b < nbins; ++b) relaxed_pk[b] = hs[DS_COUNT + b];
  if (c->prof) collect(c);
  return check_hip();
}

}  // extern "C"

