/*
 * gmilos.h -- C ABI of the MI355X-native G-MILOS inversion engine.
 *
 * The reference exposes its per-pixel Levenberg-Marquardt engine as the CUDA
 * kernel lm_mils / lm_mils_11 (milosUtils.cuh:54-60) plus sixteen __constant__
 * symbols set with cudaMemcpyToSymbol (gmilos.cu:46-68, 1184-1198), driven by
 * the host code of gmilos.cu:914-1322.  This header replaces that surface with
 * a context object that carries the same constant set (gm_setup) and entry
 * points taking plain device pointers and sizes.
 *
 * Conventions
 *  - Every buffer is caller-owned device memory unless stated otherwise.
 *  - Every call is stream-ordered on the hipStream_t passed as `void *stream`
 *    (NULL = the default stream) and returns a gm_status; nothing calls exit().
 *  - A context is reentrant with respect to other contexts; calls on one
 *    context must not overlap in time (they share its workspace).
 *  - Layouts follow the reference: observed / synthetic spectra are float
 *    [npix][4][nlambda] (I block, Q, U, V; utilsFits.c:605-608), models are
 *    double [npix][11] in Init_Model order eta0, B, vlos, dopp, aa, gm, az,
 *    S0, S1, mac, alfa (defines.h:94-106).
 */
#ifndef GMILOS_H
#define GMILOS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GM_MAX_LAMBDA  256   /* wavelengths per pixel                       */
#define GM_MAX_LINES   4     /* spectral lines sharing one atmosphere        */
#define GM_MAX_COMP    10    /* Zeeman sub-components per type and line      */
#define GM_NMODEL      11    /* Init_Model fields                            */
#define GM_NFREE       9     /* response-function slots solved (0..8)        */

typedef enum {
    GM_OK = 0,
    GM_ERR_INVALID = 1,      /* bad argument or inconsistent setup          */
    GM_ERR_HIP = 2,          /* a HIP runtime call failed                   */
    GM_ERR_NOMEM = 3,        /* device allocation failed                    */
    GM_ERR_UNSUPPORTED = 4,  /* valid for the reference, not built here     */
    GM_ERR_NOT_FOUND = 5,    /* requested item absent (e.g. LINES index)     */
    GM_CANCELLED = 6         /* serving-loop tile cancelled by the caller    */
} gm_status;

typedef enum {
    GM_PREC_FP32 = 0,        /* everything FP32 incl. solve and models      */
    GM_PREC_MIXED = 1        /* FP32 synthesis/RF/PSF/normal equations; FP64
                                solve, chi^2 accumulation and accept test,
                                FP64 models                                  */
} gm_precision;

typedef enum {
    GM_SOLVE_SVD = 0,        /* the reference's SVD pseudo-inverse (|w| > 1e-12
                                rule): LDL^T where trace(A^-1) < 1e12 proves
                                every eigenvalue > 1e-12, Jacobi SVD otherwise */
    GM_SOLVE_LDL = 1,        /* LDL^T only (zero rows dropped)               */
    GM_SOLVE_JACOBI = 2,     /* always the Jacobi SVD                        */
    GM_SOLVE_CORDIC = 3      /* the reference's own 10-term solve, result for
                                result: 18 steps of its FP32 CORDIC Jacobi
                                (mil_svdf + svdcordicf, milosUtils.cu:328-365,
                                svdcordic.cu:423-537), unconverged as there;
                                the 11-term system (mac free) keeps the FP64
                                SVD, as lm_mils_11 does                     */
} gm_solver;

typedef enum {
    GM_INIT_CE = 0,          /* init model, then classical estimates overwrite
                                B, vlos, gamma, phi (milosUtils.cu:615-628)  */
    GM_INIT_MODEL = 1,       /* init model as given for every pixel          */
    GM_INIT_PER_PIXEL = 2    /* caller passes a per-pixel start model        */
} gm_init_mode;

/* Zeeman pattern of one line, the reference's Cuantic (defines.h:108-123). */
typedef struct {
    int32_t n_pi, n_sig;
    double  nub[GM_MAX_COMP], nup[GM_MAX_COMP], nur[GM_MAX_COMP];
    double  web[GM_MAX_COMP], wep[GM_MAX_COMP], wer[GM_MAX_COMP];
    double  gl, gu, geff, fo;
} gm_zeeman;

/* The constant set the reference uploads with cudaMemcpyToSymbol
 * (gmilos.cu:1184-1198) plus the new engine's choices.                      */
typedef struct {
    int32_t nlambda;
    double  lambda[GM_MAX_LAMBDA];   /* absolute wavelengths (A)  d_lambda_const */
    int32_t nlines;
    double  lambda0[GM_MAX_LINES];   /* line centres (A)         d_wlines_const */
    gm_zeeman zee[GM_MAX_LINES];     /*                           d_cuantic_const */
    int32_t use_psf;                 /*                    d_use_convolution_const */
    double  psf[GM_MAX_LAMBDA];      /* nlambda taps, centre at nlambda/2       */
    double  weight[4];               /*                           d_weight_const */
    double  noise;                   /* sigma^2                   d_sigma_const  */
    double  ilambda;                 /* initial Marquardt lambda  d_ilambda_const */
    double  toplim;                  /*                           d_toplim_const */
    int32_t miter;                   /* max LM iterations         d_miter_const  */
    double  mu;                      /* cos(theta)                d_ah_const     */
    int32_t logclambda;              /*                           d_logclambda_const */
    int32_t fix[11];                 /* 1 = free, Init_Model order d_fix_const  */
    double  init_model[GM_NMODEL];   /*                           d_initModel_const */
    int32_t init_mode;               /* gm_init_mode                             */
    int32_t ce_line;                 /* line used by the classical estimates     */
    int32_t precision;               /* gm_precision                             */
    int32_t solver;                  /* gm_solver                                */
    int32_t rf_reference_rules;      /* 1 = reference response-function scaling  */
    double  psf_tail_tol;            /* drop PSF taps whose total tail mass is
                                        below this fraction (0 keeps all taps)  */
    int32_t use_slight;              /* mix a stray-light profile:
                                        I = alfa I_mag + (1 - alfa) I_stray
                                        (mil_sinrf.cu:133-140; slight of lm_mils) */
    float   slight[4 * GM_MAX_LAMBDA]; /* stray light [4][nlambda], I,Q,U,V      */
} gm_setup;

/* Macroturbulence and stray light.  A context takes the reference's 11-term
 * physics (circular convolution with the Gaussian macroturbulence kernel for
 * pixels whose mac exceeds 1e-4, mil_sinrf.cu:101-116 and me_der.cu:119-156;
 * stray-light mixing with alfa-scaled responses, me_der.cu:191-211) when its
 * set-up has use_slight, fix[9], or an initial mac above 1e-4.  Other
 * contexts run the 9-parameter kernels, which take mac = 0 and alfa = 1
 * whatever a caller's model rows hold.
 * A free filling factor (fix[10], which needs use_slight) solves eleven
 * slots in FP64: 0-8 as the 9-parameter system, 9 = mac (free only with
 * fix[9], the reference's lm_mils_11; otherwise the 10-term lm_mils), 10 =
 * alfa.  nfree is 4 nlambda - NTERMS (10 or 11).  Pieces of such a context
 * use alpha [npix][66], beta [npix][11], delta [npix][11] and responses
 * [npix][4][NTERMS][nlambda] in the reference's slot order.                 */

/* Outputs of an inversion (caller-owned device memory; synth and naccs may
 * be NULL).  Same content as the reference's d_vModels, d_vChisqrf,
 * d_vNumIter and d_spectraAdjusted (gmilos.cu:1176-1179).                   */
typedef struct {
    double  *model;                  /* [npix][11] best model                   */
    float   *chisqr;                 /* [npix] best chi^2                        */
    int32_t *iters;                  /* [npix] LM iterations                     */
    float   *synth;                  /* [npix][4][nlambda] last evaluated synth  */
    int32_t *naccs;                  /* [npix] accepted steps that rebuilt the
                                        normal equations (for work accounting)  */
} gm_out;

/* Per-call statistics (host memory) for measurement. */
typedef struct {
    int64_t npix;
    int64_t synth_evals;             /* pixel-evaluations of synth+chi^2        */
    int64_t rf_evals;                /* pixel-evaluations of RF+normal eqs       */
    int64_t solves;                  /* pixel step solves                        */
    int64_t svd_solves;              /* ... of which by the Jacobi SVD           */
    int32_t lm_rounds;               /* LM rounds executed                       */
    int32_t psf_taps;                /* PSF taps applied per sample              */
    float   ms_total;                /* wall time of the device work (events)    */
    float   ms_synth, ms_rf, ms_solve, ms_other; /* per-kernel-class totals
                                        (only when timing is enabled)            */
} gm_stats;

typedef struct gm_ctx gm_ctx;

/* ---- context ------------------------------------------------------------ */
/* Replaces the constant uploads of gmilos.cu:1184-1198 and the occupancy /
 * heap set-up of gmilos.cu:1008-1050.                                       */
gm_status gm_ctx_create(const gm_setup *setup, gm_ctx **ctx);
gm_status gm_ctx_destroy(gm_ctx *ctx);
const char *gm_status_string(gm_status st);
/* Per-pixel stray light (mapStrayLight of lm_mils, gmilos.cu:1212-1226):
 * map is device memory [npix][4][nlambda] that replaces the set-up's
 * profile for pixel p of every later gm_invert / gm_synth / gm_synth_rf /
 * gm_normal_equations call on this context (calls with more pixels than the
 * map covers are refused); NULL reverts to the profile.  The context's
 * set-up must have use_slight.  The caller keeps the map alive meanwhile. */
gm_status gm_ctx_set_stray_map(gm_ctx *ctx, const float *map, int64_t npix);
/* Enables per-kernel-class event timing in gm_stats (adds syncs). */
gm_status gm_ctx_set_timing(gm_ctx *ctx, int32_t enable);
/* Statistics of the context's last gm_invert; waits for that call to
 * finish on its stream. */
gm_status gm_last_stats(const gm_ctx *ctx, gm_stats *stats);

/* ---- inversion ---------------------------------------------------------- */
/* Inverts npix pixels: the work of lm_mils (milosUtils.cu:574-710) launched
 * by gmilos.cu:1233-1240.  init is [npix][11] when init_mode is
 * GM_INIT_PER_PIXEL, otherwise ignored (may be NULL).  Returns once the work
 * is enqueued on `stream` (no host wait, except when the context's state
 * must grow for a larger npix): the outputs are complete when the stream
 * reaches this point, so copies or kernels queued behind it on the same
 * stream see them, and host work or other streams overlap it.  Calls on one
 * context must use one stream at a time.                                   */
gm_status gm_invert(gm_ctx *ctx, const float *obs, int64_t npix,
                    const double *init, const gm_out *out, void *stream);

/* gm_invert without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (work under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The LM state lives
 * in a fixed device array of about 49 MB per GPU, held from the first
 * context created on that GPU until the process ends, so npix is at most
 * GM_NOALLOC_MAX_PIXELS (GM_ERR_INVALID beyond).  Same kernels and results
 * as gm_invert.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race and must be
 * serialised by the caller.  Per-kernel timing is not recorded.            */
#define GM_NOALLOC_MAX_PIXELS 65536
gm_status gm_invert_noalloc(gm_ctx *ctx, const float *obs, int64_t npix,
                            const double *init, const gm_out *out, void *stream);

/* ---- path pieces, exposed for parity tests and the serving loop --------- */
/* Synthesis of npix models (mil_sinrf, mil_sinrf.cu:27-142, with the PSF). */
gm_status gm_synth(gm_ctx *ctx, const double *models, int64_t npix,
                   float *spectra, void *stream);
/* Synthesis + response functions (me_der, me_der.cu:81-215) in the
 * reference layout rf[npix][4][10][nlambda] after the PSF and the zeroing of
 * fixed parameters (milosUtils.cu:243-271).                                */
gm_status gm_synth_rf(gm_ctx *ctx, const double *models, int64_t npix,
                      float *spectra, float *rf, void *stream);
/* Normal equations and chi^2 at the given models (covarmf + fchisqr,
 * lib.cu:65-155): alpha packed upper triangle [npix][45] over slots 0..8,
 * beta [npix][9], chisqr [npix] (already divided by nfree).                */
gm_status gm_normal_equations(gm_ctx *ctx, const double *models,
                              const float *obs, int64_t npix, float *alpha,
                              float *beta, double *chisqr, void *stream);
/* Damped step solve (milosUtils.cu:647-653 + mil_svd, milosUtils.cu:280-325)
 * for npix systems: delta[npix][9] = solve(alpha with diagonal*(1+flambda),
 * beta) using the context's precision and solver.                          */
gm_status gm_step_solve(gm_ctx *ctx, const float *alpha, const float *beta,
                        const double *flambda, int64_t npix, double *delta,
                        void *stream);
/* Classical estimates (estimacionesClasicas, milosUtils.cu:386-563) applied
 * to the init model.  for_initial_use = 1: the LM start (85/95 degree
 * inclination guard, NaN guards of milosUtils.cu:621-628); 0: the estimates
 * mode of gmilos.cu:309-431 (B_los stored in S0, no guards).               */
gm_status gm_classical_estimates(gm_ctx *ctx, const float *obs, int64_t npix,
                                 int32_t for_initial_use, double *models,
                                 void *stream);

/* ---- serving loop ---------------------------------------------------------- */
/* A persistent inversion service.  No reference counterpart: the reference
 * inverts one cube per process (gmilos.cu:1200-1256).  Tiles of any size are
 * submitted as they arrive, each for one registered setup; a worker thread
 * admits them into a device-resident LM state pool of that setup while the
 * summed footprint of in-flight tiles stays within the HBM budget, appends
 * their pixels to the running LM loop at the next round (continuous
 * batching: pixels of several tiles share every launch), retires converged
 * pixels into the tile's outputs and compacts the survivors.  Setups run on
 * separate streams and overlap on the device.  Every pixel's result equals
 * the one gm_invert gives for it: the LM state is per pixel.               */
typedef struct gm_server gm_server;
typedef struct {
    int64_t hbm_budget_bytes;        /* in-flight footprint limit (0: 8 GiB)    */
    int64_t pool_pixels;             /* state slots per setup (0: budget-sized) */
    int32_t max_tiles;               /* in-flight tiles per setup (0: 1024)     */
    int32_t device;                  /* HIP device ordinal                      */
    int32_t aging_ms;                /* admission: a queued tile that does not
                                        fit yet is passed by later tiles that
                                        do, until it has waited this long;
                                        then nothing passes it (0: 1000 ms,
                                        <0: strict arrival order)              */
    int32_t scan;                    /* queued tiles examined per admission
                                        pass (0: 64)                           */
} gm_server_config;
typedef struct {
    uint64_t ticket;
    int32_t  setup_id;
    int32_t  status;                 /* gm_status                               */
    int64_t  npix;
    double   t_submit_ms;            /* server clock (gm_server_clock_ms)       */
    double   t_admit_ms;
    double   t_done_ms;              /* completion observed by the worker       */
} gm_tile_result;
typedef struct {
    int64_t tiles_done, pixels_done;
    int64_t lm_rounds;               /* rounds launched, all setups              */
    int64_t pending_tiles, inflight_tiles;
    int64_t inflight_bytes, peak_inflight_bytes;
} gm_server_stats;

gm_status gm_server_create(const gm_server_config *cfg, gm_server **out);
/* Registers a setup (validated as gm_ctx_create); returns its id. */
gm_status gm_server_add_setup(gm_server *srv, const gm_setup *setup, int32_t *setup_id);
/* Queues a tile.  obs [npix][4][nlambda] and the outputs in *out are device
 * memory the caller keeps valid until the tile's result has been polled;
 * init [npix][11] when the setup's init_mode is GM_INIT_PER_PIXEL.  npix
 * must not exceed the setup's pool.  Thread-safe.                          */
gm_status gm_server_submit(gm_server *srv, int32_t setup_id, const float *obs, int64_t npix,
                           const double *init, const gm_out *out, uint64_t *ticket);
/* Moves up to max completed tiles into res (host); waits up to wait_ms for
 * at least one when none is ready (0: no wait, <0: until one completes or
 * nothing is outstanding). */
gm_status gm_server_poll(gm_server *srv, gm_tile_result *res, int32_t max, int32_t *n,
                         int32_t wait_ms);
/* Cancels a submitted tile.  A queued tile leaves the queue; a tile in
 * flight leaves the LM loop at the next round (its outputs then hold the
 * models reached so far, iterations past the cap).  Either way its result
 * arrives through gm_server_poll with status GM_CANCELLED, and its buffers
 * stay in use until then.  GM_ERR_NOT_FOUND: unknown ticket, or the tile
 * completed (or was cancelled) already.  Thread-safe.                      */
gm_status gm_server_cancel(gm_server *srv, uint64_t ticket);
gm_status gm_server_get_stats(gm_server *srv, gm_server_stats *stats);
double    gm_server_clock_ms(const gm_server *srv);
/* Stops the worker after the outstanding tiles complete, frees the pools. */
gm_status gm_server_destroy(gm_server *srv);

/* ---- learned initial guess ----------------------------------------------- */
/* A perceptron (4 x nlambda -> 256 -> 256 -> 256 -> 11, GELU) trained and run
 * with BF16 MFMA kernels; its prediction replaces the fixed init.mod start
 * (GM_INIT_PER_PIXEL).  No reference counterpart: the reference always starts
 * from init.mod + classical estimates (milosUtils.cu:615-628).              */
typedef struct gm_mlp gm_mlp;
gm_status gm_mlp_create(int32_t nlambda, uint64_t seed, gm_mlp **out);
gm_status gm_mlp_destroy(gm_mlp *net);
int64_t   gm_mlp_num_params(const gm_mlp *net);
/* FP32 master parameters, layer by layer: W[in][out] row-major then b[out]. */
gm_status gm_mlp_get_params(const gm_mlp *net, float *host, int64_t n);
gm_status gm_mlp_set_params(gm_mlp *net, const float *host, int64_t n);
/* One AdamW step on a batch (b a multiple of 64) of observations
 * [b][4][nlambda] with true models [b][11]; the mean-squared loss of the
 * batch (before the update) is written to *loss (host). */
gm_status gm_mlp_train_step(gm_mlp *net, const float *obs, const double *models, int64_t b,
                            float lr, float weight_decay, float *loss, void *stream);
/* Loss and gradients (host [num_params]) without updating: for checks. */
gm_status gm_mlp_gradients(gm_mlp *net, const float *obs, const double *models, int64_t b,
                           float *grads, float *loss, void *stream);
/* Start models [npix][11] for any npix: network outputs decoded onto `base`
 * (host [11], supplies mac / alfa), clamped to the LM's ranges. */
gm_status gm_mlp_predict(gm_mlp *net, const float *obs, int64_t npix, const double *base,
                         double *models, void *stream);

/* Training data and the guided inversion.  gm_prior: uniform ranges per
 * model field (lo == hi fixes a field), plus a fraction of pixels whose B is
 * redrawn below weak_b_max.  gm_prior_default gives eta0 1-20, B 0-3000 G,
 * vlos -3..3 km/s, dopp 20-60 mA, a 0.01-0.5, gamma and phi 0-180 deg,
 * S0 0.1-0.4, S1 0.5-0.9, mac 0, alpha 1.                                   */
typedef struct {
    double lo[GM_NMODEL], hi[GM_NMODEL];
    double weak_fraction, weak_b_max;
} gm_prior;
void      gm_prior_default(gm_prior *prior);
/* models [npix][11] (device) for pixel indices offset .. offset+npix-1; a
 * pixel's draw depends only on (seed, index). */
gm_status gm_draw_models(const gm_prior *prior, uint64_t seed, int64_t offset, int64_t npix,
                         double *models, void *stream);
/* Adds N(0, (sigma Ic)^2), Ic = S0 + S1 mu, to every sample of spectra. */
gm_status gm_add_noise(const double *models, float *spectra, int64_t npix, int32_t nlambda,
                       double sigma, double mu, uint64_t seed, void *stream);
/* Draws n_train models from the prior, synthesises them with synth_ctx
 * (its line, grid and PSF), adds noise sigma, and runs `steps` AdamW steps
 * of `batch` rows (linear warm-up, cosine decay from lr to lr/20).
 * loss_curve (host, [steps]) receives each step's loss when not NULL.      */
gm_status gm_mlp_fit(gm_mlp *net, gm_ctx *synth_ctx, const gm_prior *prior, int64_t n_train,
                     int64_t batch, int64_t steps, float lr, float weight_decay, double sigma,
                     uint64_t seed, float *loss_curve, void *stream);
typedef struct {
    int64_t npix;
    int64_t n_fallback;              /* pixels re-inverted from the fixed guess */
    int64_t n_replaced;              /* ... whose fallback chi^2 was lower      */
    int64_t iters_learned;           /* LM iterations, learned start (sum)      */
    int64_t iters_fallback;          /* LM iterations of the fallback (sum)     */
    float   ms_predict, ms_lm, ms_fallback, ms_total;
} gm_guided_stats;
/* Learned start + LM (learned: init_mode GM_INIT_PER_PIXEL), then every
 * pixel with chi^2 > fallback_chisqr is inverted again by `plain` (its
 * fixed guess, classical estimates as configured) and takes that result
 * when its chi^2 is lower.  The network's decode uses plain's init model
 * for the fields it does not predict (mac, alpha).                          */
gm_status gm_invert_guided(gm_ctx *learned, gm_ctx *plain, gm_mlp *net, const float *obs,
                           int64_t npix, double fallback_chisqr, const gm_out *out,
                           gm_guided_stats *stats, void *stream);

/* ---- host-side set-up helpers (no device work) -------------------------- */
/* Zeeman pattern from quantum numbers (create_cuantic.c:31-215). */
gm_status gm_zeeman_from_quantum(double sl, double ll, double jl, double su,
                                 double lu, double ju, double fo,
                                 gm_zeeman *out);
/* Parses one LINES record with index `line_index` (readConfig.c:388-492);
 * returns the central wavelength and the quantum numbers. */
gm_status gm_parse_lines_file(const char *path, int32_t line_index,
                              double *lambda0, double qn[6]);
/* Gaussian PSF of FWHM (mA) on the grid, centred at lambda[nlambda/2]
 * (fgaussC.c:14-87, gmilos.cu:244-245). */
gm_status gm_psf_gaussian(double fwhm_mA, const double *lambda,
                          int32_t nlambda, double *psf);
/* Tabulated PSF (mA, value) resampled onto the grid offsets (mA) with the
 * reference's centring offset and unit-sum normalisation (gmilos.cu:247-291,
 * interpolLineal.c:46-98). */
gm_status gm_psf_from_table(const double *dl_mA, const double *val,
                            int32_t ntab, const double *offsets_mA,
                            int32_t nlambda, double step_mA, double *psf);
/* Fills a gm_setup with the reference defaults (readConfig.c:704-744). */
void gm_setup_defaults(gm_setup *s);

#ifdef __cplusplus
}
#endif
#endif /* GMILOS_H */
