/* sf21_hii.h -- C-ABI of the MI355X (gfx950) ionization-field step of
 * SimFast21: the excursion-set HII bubble finder of get_HIIbubbles.c.
 *
 * The reference has no library interface: its boundary is the program
 * `get_HIIbubbles.x base_dir` (get_HIIbubbles.c:40-68) plus the files under
 * base_dir, and inside it the FFTW-threaded loop get_HIIbubbles.c:98-488.
 * This header is the thin C boundary that replaces that loop.  It is called
 * by the drop-in CLI (simfast21_amd/host/get_HIIbubbles.c), by the Python
 * package through ctypes, and by the benchmark.
 *
 * Conventions
 *   - Every function returns int: SF21_OK (0) or a negative SF21_ERR_*; the
 *     library never calls exit().  sf21_strerror() names a code.
 *   - Boxes are float32, n^3, index i*n*n + j*n + p (get_HIIbubbles.c:338,
 *     output_files.txt:21-25).  "_dev" entry points take device pointers
 *     (hipMalloc'd memory on the context's device); the others host pointers.
 *   - Lengths: the box side and the radii share one unit (the reference
 *     works in Mpc/h: Input_variables.c:121-129).
 *   - A context is single-threaded and bound to one HIP stream; separate
 *     contexts may run concurrently.  The caller owns its buffers; the
 *     context owns device scratch, tables and captured graphs.
 */
#ifndef SF21_HII_H
#define SF21_HII_H

#ifdef __cplusplus
extern "C" {
#endif

#define SF21_OK 0
#define SF21_ERR_ARG (-1)         /* invalid argument or unsupported size */
#define SF21_ERR_ALLOC (-2)       /* device or host allocation failed */
#define SF21_ERR_HIP (-3)         /* HIP runtime error (no device, launch failure) */
#define SF21_ERR_IO (-4)          /* file missing, unreadable or short */
#define SF21_ERR_FORMAT (-5)      /* malformed simfast21.ini line */

/* Input modes of a slice. */
#define SF21_INPUT_SFR 0     /* a = delta (deltanl box), b = SFRD box: the reference's
                                use_SFR = T path, get_HIIbubbles.c:216-225, :262-275 */
#define SF21_INPUT_DIRECT 1  /* a = recombination box, b = emissivity box, c = residual
                                neutral fraction box (nullable -> 0) */
#define SF21_INPUT_HALO 2    /* a = delta, b = emissivity box deposited from a halo catalog
                                (sf21_hii_halo_emissivity_dev): the reference's
                                use_SFR = F path, get_HIIbubbles.c:216-259 */

/* Execution paths. */
#define SF21_PATH_FUSED 0    /* fused exact finder: filtered fields of several radii per
                                pass, exact ball dilation of all centres at once */
#define SF21_PATH_PLAIN 1    /* the reference's dataflow on the GPU, one radius at a time:
                                two real inverse transforms per radius, ball FFT
                                convolution (main loop) and stencil (small loop), the
                                any-centre flag read back each radius */
#define SF21_PATH_DEPOSIT 2  /* deposits only (sf21_sfrd_dev, sf21_nldensity_dev,
                                sf21_hii_halo_emissivity_dev): any n >= 1 (n_halo a
                                whole number of n-blocks as smooth_box takes it), no
                                transforms and no bubble-finder buffers; radii may be
                                NULL.  Every other entry point rejects such a context. */

typedef struct sf21_hii_ctx sf21_hii_ctx;
typedef struct sf21_params sf21_params;

typedef struct sf21_hii_config {
  int n;                 /* N_smoothed: 64, 96, 128, 192, 256, 512 (32 fused only); fused
                            path also any other multiple of 8 in 16..768 with factors
                            2, 3, 5 and n/q a multiple of 8, at most 64, for some q <= 16
                            (e.g. 80, 120, 200, 360, 600); there the autotuner
                            searches the line passes' block order and the
                            differentiable mode stores no fields */
  double box_len;        /* box side L (unit of the radii) */
  int n_radii;           /* ladder length */
  const double *radii;   /* descending ladder (get_HIIbubbles.c:298-461), every radius > 0
                            and none larger than the one before (else SF21_ERR_ARG);
                            radii >= 2*dx are the main loop, [dx, 2*dx) the small loop */
  int input_mode;        /* SF21_INPUT_* */
  int path;              /* SF21_PATH_* */
  double hubble;         /* h, SFR mode (Rrec's cell volume, user_functions.c:62) */
  double fesc;           /* escape fraction, SFR mode (get_HIIbubbles.c:273) */
  int radius_batch;      /* radii per fused filter pass, 0 = automatic (16, fewer if the
                            K n^3 complex buffer does not fit); sides without a register
                            plan run one radius per step */
  int device;            /* HIP device ordinal */
  int filter_tile;       /* columns per block of the fused filter pass: 4, 8, 16 or 0 (auto) */
  int filter_w;          /* top-hat weights W(|k| R): 0 = automatic (FP32 in the kernel),
                            1 = FP32 evaluation in the kernel (series below |k|R = 1.25,
                            3 (sin x - x cos x)/x^3 above), 2 = FP64 table per integer
                            |k|^2 rounded to float (sides without a register plan always
                            evaluate in the kernel) */
  int n_halo;            /* N_halo, halo mode: the catalog grid (>= n) */
} sf21_hii_config;

const char *sf21_strerror(int code);

/* Context lifetime: replaces the allocations and FFTW plans of
 * get_HIIbubbles.c:98-191. */
int sf21_hii_create(const sf21_hii_config *cfg, sf21_hii_ctx **out);
int sf21_hii_destroy(sf21_hii_ctx *ctx);
/* Use an external hipStream_t (e.g. torch's) instead of the context's own. */
int sf21_hii_set_stream(sf21_hii_ctx *ctx, void *hip_stream);

/* One redshift slice: replaces get_HIIbubbles.c:216-479 (everything between
 * reading the input boxes and writing the xHII box).  z is used in SFR mode
 * only.  *neutral receives mean(1 - xHII) reduced in FP64 (:473-477). */
int sf21_hii_run_slice(sf21_hii_ctx *ctx, const float *a, const float *b, const float *c,
                       double z, float *xhii, double *neutral);
int sf21_hii_run_slice_dev(sf21_hii_ctx *ctx, const float *a, const float *b,
                           const float *c, double z, float *xhii, double *neutral);

/* Several slices of the same box size in one submission (device pointers);
 * the whole sequence is captured once into a hipGraph and replayed.
 * neutral: host array of n_slices doubles. */
int sf21_hii_run_batch_dev(sf21_hii_ctx *ctx, int n_slices, const float *const *a,
                           const float *const *b, const float *const *c, const double *z,
                           float *const *xhii, double *neutral);

/* Boxes of different sizes in one submission (device pointers): box i runs
 * on context ctx[i] (fused path, one device; boxes may share a context of
 * their size and ladder, e.g. one per distinct N).  The whole set is
 * captured into one hipGraph on ctx[0]'s stream, cached on ctx[0] by
 * (contexts, pointers, redshifts), and replayed.  neutral: host array of
 * nbox doubles.  No counterpart in the reference (one box per run there);
 * each box is get_HIIbubbles.c:216-479 on its own inputs. */
int sf21_hii_run_multi_dev(int nbox, sf21_hii_ctx *const *ctx, const float *const *a,
                           const float *const *b, const float *const *c, const double *z,
                           float *const *xhii, double *neutral);

/* Halo mode.  halos: n_halos records {float Mass, x, y, z} with positions
 * in N_halo cell units (auxiliary.h:13-17, output_files.txt:10-14).
 * Emissivity box = smooth_box(CIC(Rion(M, z))) * (dx_s/dx_h)^3 * fesc
 * (get_HIIbubbles.c:233-259); _dev takes device pointers.  run_slice_halo
 * takes host delta and catalog and runs the whole slice. */
int sf21_hii_halo_emissivity_dev(sf21_hii_ctx *ctx, const float *halos, long n_halos, double z,
                                 float *ion_out);
/* Allocation-free variant of sf21_hii_halo_emissivity_dev: no counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and result (bit for bit: the same kernels);
 * the N_halo^3 deposit grid lives in a library-owned device array of at
 * most 2^24 floats (64 MiB, N_halo <= 256; larger grids -> SF21_ERR_ARG)
 * that each GPU holds until the process ends, instead of the context's
 * lazily allocated grid.  Enqueued on the context's stream without host
 * synchronisation.  Every call on one GPU shares that GPU's array: calls
 * that overlap in time on one GPU (other streams, other contexts, other
 * threads) race. */
int sf21_hii_halo_emissivity_noalloc(sf21_hii_ctx *ctx, const float *halos, long n_halos,
                                     double z, float *ion_out);
int sf21_hii_run_slice_halo(sf21_hii_ctx *ctx, const float *delta, const float *halos,
                            long n_halos, double z, float *xhii, double *neutral);

/* Spherically averaged power spectrum of a real n^3 box (device pointer),
 * normalised as tools/power3d.c:172-191 does for a top-hat window
 * (P = <|F_k|^2> dl^3 / n^3 over half-spectrum entries, dl = box_len/n), in
 * nbins (<= 64) linear bins of |k|/dk over [1, n/2] (width (n/2-1)/nbins,
 * last bin closed; k = 0 and corners beyond n/2 dropped).  Sums in FP64.
 * Host outputs: pk, kmean (mean |k| per bin), counts (modes per bin); any
 * may be NULL.  n >= 64.  Uses the context's spectrum buffer. */
int sf21_hii_power_spectrum_dev(sf21_hii_ctx *ctx, const float *field, int nbins, double *pk,
                                double *kmean, double *counts);

/* ---- 21-cm brightness temperature: the step after (t21.c:116-252) --------
 * One redshift on device boxes (n^3 floats; the context supplies n, its FFT
 * tables and scratch; n >= 64): the velocity gradient of delta along
 * p->vel_comp (1 = i, 2 = j, other = p) in k-space, (k_axis/k)^2, scaled by
 * -dD/dt/(H D) and clamped to [-0.5, 1]; t21 = 23 mK (1+delta) S/(1+dv)
 * (0.7/h)(Ob h^2/0.02) sqrt((0.15/Om/h^2)(1+z)/10) (1 - xHII) in kelvin,
 * with S = x/(1+x) (1 - Tcmb/TK), x = x_alpha + x_c, when the three spin
 * boxes are given (the reference's z > Zminsfr with use_Lya_xrays = T), else
 * S = 1.  Means in FP64: *mean_t21, and *mean_ts of Tcmb/(1 - S) (spin
 * branch; 0 otherwise). */
int sf21_t21_dev(sf21_hii_ctx *ctx, const sf21_params *p, double z, const float *delta,
                 const float *xhii, const float *xc, const float *xalpha, const float *tk,
                 float *out, double *mean_t21, double *mean_ts);

/* ---- The step before: the bubble finder's inputs -------------------------
 * SFRD box of one redshift from a halo catalog (get_SFR.c:98-148; device
 * pointers): CIC of sfr(M, z) = Rion(M, z)/Qion(z) (user_functions.c:71-74)
 * on the N_halo^3 grid, smooth_box to n^3, divided by dx_halo^3 (Msun per
 * year per (Mpc/h)^3).  The context needs n_halo >= n.  *mean (nullable):
 * the box average in FP64 (the sfrd_av table). */
int sf21_sfrd_dev(sf21_hii_ctx *ctx, const float *halos, long n_halos, double z, float *out,
                  double *mean);
/* Non-linear density of one redshift (get_nldensity.c:117-166; device
 * pointers): each N_halo^3 cell carries mass 1 + delta0 D(300) to its
 * position plus v N_halo h/L_h (D(z) - D(300)) (v: the z = 0 velocity boxes
 * in Mpc) and deposits it by CIC on a grid that starts at -1; box_halo
 * (nullable: context scratch) receives that N_halo^3 box and out (nullable)
 * its smooth_box average to n^3 -- the deltanl box of the bubble finder. */
int sf21_nldensity_dev(sf21_hii_ctx *ctx, const sf21_params *p, double z, const float *vx,
                       const float *vy, const float *vz, const float *delta0, float *box_halo,
                       float *out);

/* ---- Differentiable mode -------------------------------------------------
 * No counterpart in the reference (SimFast21 has no gradients); it wraps the
 * bubble finder of get_HIIbubbles.c:277-479 for inference work.  Context:
 * input_mode SF21_INPUT_DIRECT, path SF21_PATH_FUSED, n >= 64; sides
 * without a register plan (such as 600) run the same model through their
 * line transforms, one radius pair at a time, always recomputing the
 * per-radius fields.  Inputs
 * (device, n^3 floats): g, a Gaussian field, and emis, an emissivity
 * pattern; zeta scales it.  The model:
 *   Delta = exp(g)/<exp(g)>, rec = Delta^2, ion = zeta*emis, no residual
 *   neutral fraction; x = the exact bubble finder on (rec, ion);
 *   T = (1 - x) Delta; observables <x> and P_b(T) in nbins bins as
 *   sf21_hii_power_spectrum_dev defines them.
 * diff_forward runs the exact model: x_out (device, nullable), *mean_x and
 * pk (host, nbins, nullable).
 * diff_backward returns the gradient of L = dl_dx <x~> + sum_b dl_dpk[b] P_b
 * of the tau-relaxed model (the centre test min(I_R - R_R, I_R) >= 0 becomes
 * phi_R = softplus(u) - softplus(u - 20), u = s_R/tau, with
 * s_R = min(I_R - R_R, I_R)/max(R_R, 1e-3 <rec>);
 * the union of balls becomes x~ = q + (1 - q)(1 - exp(-sum_R B_R * phi_R)),
 * B_R the exact ball, q the quick fill) with respect to g (grad_g, device
 * n^3) and zeta (*grad_zeta), plus the relaxed <x~> and P_b (host,
 * nullable).  Per-radius fields are recomputed in the backward unless
 * sf21_hii_diff_store_fields asks otherwise; tau -> 0 gives back the exact
 * forward.  Sums in FP64. */
/* Field storage: the backward's first sweep recomputes every radius'
 * filtered pair from the forward spectrum; mode 1 keeps them (n_R n^3
 * complex floats, 47 GB at 512^3 with 47 radii) for the second sweep, mode 0
 * recomputes them there as well, mode 2 (default) keeps them when they take
 * at most half of the free device memory.  Mode 3: the exact forward keeps
 * them (its centre-test pass writes them, +8 B per voxel and radius) and the
 * next backward reads them instead of recomputing its first sweep -- when
 * its (g, emis, zeta) carry the same fingerprints as the forward's (64-bit
 * sums of a hash of each value's bits mixed with its index: an input
 * changed in place between the two calls, values swapped between cells
 * included, is detected except with probability ~2^-64); otherwise it
 * recomputes as in mode 1.  The gradient is identical in every mode;
 * SF21_ERR_ALLOC if mode 1 or 3 does not fit.  Sides without a register
 * plan accept modes 0 and 2 only (SF21_ERR_ARG otherwise) and store
 * nothing. */
int sf21_hii_diff_store_fields(sf21_hii_ctx *ctx, int mode);
/* 1 if the last diff_backward read fields kept by the forward (mode 3), 0
 * otherwise. */
int sf21_hii_diff_last_kept(const sf21_hii_ctx *ctx);
int sf21_hii_diff_forward_dev(sf21_hii_ctx *ctx, const float *g, const float *emis, double zeta,
                              int nbins, float *x_out, double *mean_x, double *pk);
int sf21_hii_diff_backward_dev(sf21_hii_ctx *ctx, const float *g, const float *emis,
                               double zeta, double tau, int nbins, double dl_dx,
                               const double *dl_dpk, float *grad_g, double *grad_zeta,
                               double *relaxed_mean_x, double *relaxed_pk);

/* Inspection for tests (fused path): per-cell ball bound d2max(R) of the
 * largest radius at which the cell was a bubble centre in the last slice, -1
 * if never (host output, n^3 ints). */
int sf21_hii_weight_map(sf21_hii_ctx *ctx, int *host_out);
/* Inspection for tests (fused path): the exact periodic dilation and the
 * final per-cell pass of a slice (quick fill, clamp, FP64 mean) on a given
 * map instead of the finder's.  host_map (n^3 ints): per cell -1 or a ball
 * bound d2 -- every x with min-image |x - y|^2 <= d2 of such a cell y is
 * ionised -- each at most the ladder's largest d2max (SF21_ERR_ARG
 * otherwise); a, b, c, z as for sf21_hii_run_slice_dev, xhii (device) and
 * *neutral as there.  No counterpart in the reference, which dilates per
 * radius (get_HIIbubbles.c:358-393, :428-458). */
int sf21_hii_dilate_dev(sf21_hii_ctx *ctx, const int *host_map, const float *a, const float *b,
                        const float *c, double z, float *xhii, double *neutral);
/* Number of hipGraphs captured so far with ctx as their owner (batches on
 * ctx, multi-box submissions whose first context is ctx).  A cached graph
 * is replayed only while every context it uses keeps its buffers, stream
 * and kernel choices; after sf21_hii_autotune, sf21_hii_set_stream or a
 * context's destruction, the next call captures again. */
int sf21_hii_graph_captures(const sf21_hii_ctx *ctx, long *captures);
/* Filtered pair of one ladder radius for a slice: out (device, 2*n^3
 * floats, interleaved) = n^3 * (W_R rec, W_R ion) as the unnormalised c2r of
 * get_HIIbubbles.c:324-325 gives them. */
int sf21_hii_filtered_dev(sf21_hii_ctx *ctx, const float *a, const float *b, const float *c,
                          double z, int radius_index, float *out);

/* Autotuner (fused path; no counterpart in the reference, whose FFTW plans
 * are tuned by FFTW_MEASURE at get_HIIbubbles.c:110-183).  Times the slice
 * (a, b, c, z; device pointers) by a coordinate search over launch shapes
 * and stage orders: radius batch 16, 8, 12, 24, 32 or the whole ladder
 * (capped by the ladder; skipped when it does not fit), filter tile 16/8, filter blocks over single
 * i-lines or mirror pairs (kj, n - kj) sharing W, the per-radius column
 * passes (8 values per thread with 8 or 16 columns per block, or 16 values
 * per thread with 16 or 32 columns, the latter for n = 128, 256, 512; with
 * 16 or 32 columns also as blocks that walk 8 tiles and prefetch the next)
 * and the centre-test pass (8 or 16 values per thread).  The top-hat weight
 * mode (filter_w) is never searched: it stays what the configuration says.
 * Keeps the fastest in the context and appends it to the cache file
 * (cache_path, else $SF21_TUNE_CACHE, else not recorded) keyed by
 * (device arch, n, box_len, the radii).  A context created with
 * radius_batch, filter_tile and filter_w all 0 takes the cached choice for
 * its key from $SF21_TUNE_CACHE.  *best_ms (nullable): the winner's time per
 * slice.  Sides without a register plan (one radius per step through line
 * transforms) have one choice: which column passes (along j, along i) deal
 * runs of consecutive blocks to one XCD (sf21_hii_line_order); the four
 * masks are timed and the winner is cached the same way.  Every choice
 * changes launch shapes or block order only: the results are identical. */
int sf21_hii_autotune(sf21_hii_ctx *ctx, const float *a, const float *b, const float *c,
                      double z, const char *cache_path, double *best_ms);
/* Current choices: filter tile, radius batch, top-hat weights (1 in the
 * kernel, 2 table), column-pass tile, column-pass variant (bit 0: 16 values
 * per thread, bit 1: prefetching blocks), 16 values per thread in the
 * centre-test pass (0/1), filter blocks over mirror pairs of i-lines that
 * share the weights (0/1); any pointer may be NULL. */
int sf21_hii_tuning(const sf21_hii_ctx *ctx, int *filter_tile, int *radius_batch, int *filter_w,
                    int *cols_tile, int *cols_wide, int *rows_wide, int *filter_pair);

/* Effective radii per fused filter pass (the automatic choice depends on
 * free device memory); negative on error. */
int sf21_hii_radius_batch(const sf21_hii_ctx *ctx);
/* Sides without a register plan: the XCD-order mask of the line passes
 * (bit 1: passes along j, bit 2: along i; default 4); SF21_ERR_ARG on other
 * sides.  No counterpart in the reference (a launch choice of this
 * implementation). */
int sf21_hii_line_order(const sf21_hii_ctx *ctx);

/* Kernel timing with HIP events on the context's stream. */
int sf21_hii_profile(sf21_hii_ctx *ctx, int enable);
int sf21_hii_kernel_stats(sf21_hii_ctx *ctx, int index, char *name, int name_cap,
                          long *launches, double *total_ms);
int sf21_hii_reset_stats(sf21_hii_ctx *ctx);

/* ---- FFTW-like 3D transforms on device memory --------------------------
 * Replace fftwf_plan_dft_r2c_3d / fftwf_plan_dft_c2r_3d / fftwf_execute as
 * used at get_HIIbubbles.c:110, :128, :137, :146, :159, :176, :180, :290-291,
 * :324-325, :372-373, :382.  Unnormalised, forward sign -1, half layout
 * n x n x (n/2+1) interleaved complex.  C2R destroys its input, like FFTW's
 * multi-dimensional c2r. */
#define SF21_FFT_R2C 0
#define SF21_FFT_C2R 1
#define SF21_FFT_C2C_FORWARD 2
#define SF21_FFT_C2C_BACKWARD 3
typedef struct sf21_fft_plan sf21_fft_plan;
int sf21_fft_plan_3d(int n, int kind, sf21_fft_plan **out);
int sf21_fft_execute(sf21_fft_plan *plan, void *in, void *out, void *hip_stream);
int sf21_fft_destroy(sf21_fft_plan *plan);

/* ---- host logic (no GPU needed) ---------------------------------------- */

/* The parameters get_HIIbubbles reads from base_dir/simfast21.ini, parsed as
 * Input_variables.c:16-141 does, with its derived values. */
struct sf21_params {
  int nthreads;
  long n_halo, n;          /* N_halo, N_smoothed */
  double zmin, zmax, dz;   /* zmin snapped to zmax - dz*ceil((zmax-zmin)/dz) (:133) */
  double box_len;          /* sim_length * hubble, Mpc/h (:121) */
  double hubble;
  double fesc, ion_cutoff;
  double bubble_rmax;      /* bubble_Rmax * hubble, capped at box_len/2 (:123-124) */
  int bubble_nbins;
  int use_sfr;             /* 0 unless use_SFR = T */
  double dx, dx_halo, dk;  /* box_len/n, box_len/n_halo, 2*pi/box_len (:127-129) */
  /* keys of the 21-cm step (t21.c) */
  double omega_m, omega_b, omega_lambda;
  int vel_comp;            /* Vel_comp: 1 = i, 2 = j, other = p axis of the gradient */
  int use_lya_xrays;       /* 0 unless use_Lya_xrays = T */
  double zmin_sfr;         /* Zminsfr snapped like zmin and raised to it (:137-139) */
  int save_original_nl;    /* Original_nldensity_box = T: keep the N_halo^3 box too */
};
/* msg receives the reference's message for SF21_ERR_IO / SF21_ERR_FORMAT. */
int sf21_params_read(const char *base_dir, sf21_params *out, char *msg, int msg_cap);

/* Background cosmology of the 21-cm step (auxiliary.c:224-303): H(z) in
 * 1/s, the growth factor D(z) (D(0) = 1) and dD/dt in 1/s, in FP64 with the
 * reference's step sizes. */
double sf21_hubble_z(const sf21_params *p, double z);
double sf21_growth(const sf21_params *p, double z);
double sf21_dgrowth_dt(const sf21_params *p, double z);

/* Radius ladder of get_HIIbubbles.c:192, :298-299, :397, :400-402, :461:
 * bfactor = 10^(log10(rmax/dx)/nbins), R divided by bfactor in FP64 while
 * R >= 2dx (main), then while R >= dx (small).  Returns the count or
 * SF21_ERR_ARG if cap is too small; *n_main = main-loop count. */
int sf21_ladder(double rmax, double dx, int nbins, double *radii, int cap, int *n_main);
/* Largest integer m with sqrt((double)m)*dx <= R: the ball bound of the
 * tests at get_HIIbubbles.c:366, :443, :447. */
long sf21_ball_d2max(double R, double dx);

#ifdef __cplusplus
}
#endif
#endif /* SF21_HII_H */
