/*
 * gm_internal.h -- device-side set-up tables and kernel launchers shared by
 * the C ABI (gm_capi.cpp) and the kernels (gm_kernels.hip).
 */
#ifndef GM_INTERNAL_H
#define GM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gmilos.h"

#define GM_ALPHA_N   45          /* packed upper triangle of the 9x9 system */
#define GM_ALPHA_MAX 66          /* ... of the 11x11 system (free filling factor) */
#define GM_BETA_MAX  11
#define GM_AB_REC9   64          /* per-pixel alpha/beta record: 45 + 9 at 48 */
#define GM_AB_REC11  80          /* ... 66 + 11 at 68                          */
#define GM_AB_BETA9  48
#define GM_AB_BETA11 68
#define GM_NPRE      19          /* per-pixel FP32 synthesis constants      */
#define GM_PRE_REC   20          /* ... stored as 80-byte per-pixel records  */
#define GM_MODEL_REC 12          /* LM state models: 96-byte per-pixel records */
#define GM_RED_N     55          /* 45 alpha + 9 beta + 1 chi^2             */
#define GM_SLOTS     36          /* staged values per wavelength: per Stokes
                                    [synth, eta0, B, vlos, dopp, a, gamma, phi,
                                    S1]                                       */
#define GM_CPT       (GM_MAX_LINES * GM_MAX_COMP)

/* PSF of the 9-parameter response kernel (k_rf) at up to 32 wavelengths on
 * the matrix cores when at least GM_MFMA_MIN_TAPS taps are applied.  The
 * dense product costs about as much as the VALU correlation with 15 taps
 * (FP32 MFMA and FP32 VALU work do not overlap on a SIMD: measured, the
 * kernel's time is the sum of both), so fewer taps stay on the VALU.
 * GM_MFMA_RF 0 keeps every set-up on the VALU.                            */
#ifndef GM_MFMA_RF
#define GM_MFMA_RF 1
#endif
#ifndef GM_MFMA_MIN_TAPS
#define GM_MFMA_MIN_TAPS 16
#endif

/* Device copy of the set-up, built once per context (read-only). */
struct gm_dsetup {
    int32_t nlambda;
    int32_t nlines;
    int32_t ncomp[GM_MAX_LINES][3];          /* sub-components per type     */
    float   comp_nu[GM_MAX_LINES][3][GM_MAX_COMP];
    float   comp_w[GM_MAX_LINES][3][GM_MAX_COMP];
    float   l0[GM_MAX_LINES];                /* line centres (A)            */
    float   fo[GM_MAX_LINES];                /* relative line strengths     */
    float   dl[GM_MAX_LINES][GM_MAX_LAMBDA]; /* lambda - lambda0 (A)        */
    int32_t use_psf;
    int32_t psf_mfma;                        /* k_rf's PSF on the matrix cores */
    int32_t psf_dlo, psf_dhi;                /* applied tap offsets d=j-c   */
    float   psf_tap[GM_MAX_LAMBDA];          /* tap for offset d: [d-dlo]   */
    float   psf_edge[GM_MAX_LAMBDA];         /* 1 - sum of in-range taps    */
    float   weight[4];
    float   wsig[4];                         /* weight / noise              */
    int32_t wsig_equal;                      /* all four wsig the same       */
    float   mu;
    float   shi_scale;                       /* 1 reference rule, 2 exact   */
    uint32_t freemask;                       /* bit p: RF slot p free (0-8) */
    int32_t fix[11];
    int32_t nfree_chi;                       /* 4*nlambda - 10              */
    double  noise;
    double  ilambda, toplim;
    int32_t miter, logclambda;
    double  init_model[GM_NMODEL];
    /* classical estimates */
    int32_t ce_lo, ce_hi;
    double  ce_l0, ce_geff;
    double  lambda[GM_MAX_LAMBDA];           /* absolute wavelengths (A)    */
    /* macroturbulence / stray light (kernels built with EXT only) */
    int32_t ext;                             /* 11-term physics active       */
    int32_t np;                              /* solved slots: 9, or 11 with a
                                                free filling factor (slot 9
                                                mac, slot 10 alfa)          */
    int32_t nt_ref;                          /* reference NTERMS: 10 or 11   */
    int32_t use_slight;
    double  mac_l0;                          /* lambda0 of the first line    */
    float   mac_tail_tol;                    /* combined mac kernel: taps
                                                outside the window holding
                                                all those above tol * sum / n
                                                are skipped (0: every tap) */
    float   psf_full[GM_MAX_LAMBDA];         /* all nlambda PSF taps         */
    float   slight[4 * GM_MAX_LAMBDA];       /* stray light [4][nlambda]     */
    const float *slight_map;                 /* per-pixel [npix][4][nlambda]
                                                or NULL (mapStrayLight)     */
};

/* Per-pixel LM state, structure-of-arrays over the pixel capacity. */
struct gm_state {
    int64_t  cap;
    double  *best;        /* [cap][GM_MODEL_REC] best accepted model        */
    double  *trial;       /* [cap][GM_MODEL_REC] last proposed model        */
    float   *alpha;       /* [cap][ab_rec] per-pixel records: packed normal
                             matrix at best (45 or 66 floats) ...            */
    float   *beta;        /* ... and the gradient at alpha + ab_beta         */
    int32_t  ab_rec;      /* record stride in floats (64, or 80 for 11 slots) */
    float   *pre_best;    /* [cap][GM_PRE_REC] synthesis constants of best   */
    float   *pre_trial;   /* [cap][GM_PRE_REC] ... of trial                  */
    double  *chi_trial;   /* [cap] chi^2 of trial                           */
    double  *ochisqr;     /* [cap] chi^2 of best                            */
    double  *chisqr0;     /* [cap] initial chi^2                            */
    double  *flambda;     /* [cap] Marquardt lambda                          */
    double  *factor;      /* [cap] PARBETA_FACTOR (logclambda)              */
    int32_t *iter;        /* [cap]                                           */
    int32_t *nacc;        /* [cap]                                           */
    int32_t *list[4];     /* two active lists (ping-pong), accepted list and
                             SVD fallback list                               */
    int32_t *counts;      /* [4]: counts of the four lists                   */
    int32_t *hist;        /* [3 * (max rounds + 1)] per-round counts          */
};

/* Output slot of one tile of the serving loop (device table). */
struct gm_tile_dev {
    double  *model;       /* [npix][11]                                      */
    float   *chisqr;
    int32_t *iters;
    int32_t *naccs;
    float   *synth;       /* [npix][4][nlambda] or NULL                      */
    int64_t  slot0;       /* first state slot of the tile                    */
    int32_t  remaining;   /* pixels still in the LM loop                     */
    int32_t  pad;
};

/* Lane mapping of the wavelength-parallel kernels. */
struct gm_lanecfg {
    int lpp;   /* lanes per pixel (32 or 64)                                 */
    int nlpl;  /* wavelengths per lane                                       */
    int ppw;   /* pixels per 64-lane workgroup                               */
    int rows;  /* staged rows per pixel incl. zero padding                   */
    int padlo; /* zero rows before wavelength 0                              */
    int ext;   /* macroturbulence / stray-light kernels                      */
    int np;    /* 9 or 11 solved slots                                       */
    int mfma;  /* k_rf's PSF on the matrix cores (gm_dsetup::psf_mfma): its
                  staged rows need no zero padding                         */
};

gm_lanecfg gm_lane_config(const gm_dsetup &h);
/* Dynamic LDS per workgroup of the wavelength-parallel kernels. */
size_t gm_lds_bytes(const gm_lanecfg &lc);

/* Launchers (gm_kernels.hip).  `list` may be NULL (identity) with the item
 * count given by *count (device) or by n_direct when count is NULL.  `pre`,
 * when not NULL, holds the pixels' synthesis constants (SoA, same stride as
 * the models); otherwise they are derived from the models.                 */
void gm_launch_rf(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                  const float *pre, int64_t model_stride, const float *obs, const int32_t *list,
                  const int32_t *count, int64_t n_direct, int64_t grid_items,
                  float *alpha, float *beta, int64_t ab_stride, double *chi_out,
                  float *synth_out, float *rf_out, int mixed, hipStream_t st,
                  int alpha_first = 1, double *model_dst = nullptr,
                  float *pre_dst = nullptr);
void gm_launch_synth(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                     const float *pre, int64_t model_stride, const float *obs, const int32_t *list,
                     const int32_t *count, int64_t n_direct, int64_t grid_items,
                     double *chi_out, float *synth_out, int mixed, hipStream_t st);
void gm_launch_solve(const gm_dsetup *dS, gm_state st, const int32_t *list,
                     const int32_t *count, int64_t grid_items, int mixed, int solver,
                     int32_t *reset0, int32_t *reset1, int32_t *hist,
                     int32_t *fb_list, int32_t *fb_count, hipStream_t stream, int np = 9,
                     int cordic10 = 0);
void gm_launch_step_solve(const float *alpha, const float *beta, const double *flambda,
                          int64_t n, double *delta, int mixed, int solver,
                          hipStream_t stream, int np = 9, int cordic10 = 0);
/* eleven-slot contexts with mac fixed take the reference's 10-term CORDIC
 * solve under GM_SOLVE_CORDIC (with mac free the reference solves the
 * 11-term system with an FP64 SVD) */
inline int gm_cordic10(const gm_setup &s, int np)
{
    return s.solver == GM_SOLVE_CORDIC && np == 11 && !s.fix[9];
}
void gm_launch_decide(const gm_dsetup *dS, gm_state st, const int32_t *list,
                      const int32_t *count, int64_t grid_items, int32_t *next_list,
                      int32_t *next_count, int32_t *acc_list, int32_t *acc_count,
                      int32_t *reset_fb, int32_t *done_list, int32_t *done_count, int mixed,
                      hipStream_t stream, int defer_copy = 0, int32_t *fb_hist = nullptr,
                      int small_groups = 0);
void gm_launch_init_state(const gm_dsetup *dS, gm_state st, const float *obs,
                          const double *init, int init_mode, int nlambda, int64_t slot0,
                          int64_t npix, int32_t *list0, hipStream_t stream);
void gm_launch_init_chi(gm_state st, int64_t slot0, int64_t npix, hipStream_t stream);
void gm_launch_append_range(int32_t *list, int32_t *count, int64_t slot0, int64_t n,
                            hipStream_t stream);
void gm_launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t stream);
void gm_launch_retire(gm_state st, const int32_t *done, const int32_t *done_count,
                      int64_t grid_items, const float *synth, int nl4, const int32_t *slot_tile,
                      gm_tile_dev *tiles, int32_t *completed, hipStream_t stream);
void gm_launch_finalize(gm_state st, int64_t npix, const gm_out &out,
                        hipStream_t stream);
void gm_launch_ce(const gm_dsetup *dS, const float *obs, int64_t npix, int initial,
                  double *models, hipStream_t stream);
void gm_launch_set_count(int32_t *count, int32_t value, hipStream_t stream);
void gm_launch_copy_count(const int32_t *src, int32_t *dst, hipStream_t stream);

/* LM state of gm_invert_noalloc: a fixed layout for GM_NOALLOC_MAX_PIXELS
 * pixels (gm_state_bytes of that capacity, rounded up) in a file-scope
 * device array (gm_noalloc.hip), one per device.                          */
#define GM_NOALLOC_SCRATCH_BYTES ((size_t)49 << 20)
#define GM_MAX_ROUNDS 1024   /* LM iteration cap of a set-up (miter) */
/* Bytes of the LM state for `cap` pixels (cap a multiple of 256): per pixel
 * two model records, five doubles, the alpha/beta record, two constants
 * records, two ints and four list entries; then the list counters and the
 * per-round counts, each array on a 256-byte boundary.                    */
constexpr size_t gm_state_bytes(int64_t cap)
{
    return (size_t)cap * (sizeof(double) * (2 * GM_MODEL_REC + 5) +
                          sizeof(float) * (GM_AB_REC11 + 2 * GM_PRE_REC) +
                          sizeof(int32_t) * (2 + 4)) +
           256 + ((sizeof(int32_t) * 3 * (GM_MAX_ROUNDS + 2) + 255) / 256) * 256;
}
static_assert(gm_state_bytes(GM_NOALLOC_MAX_PIXELS) <= GM_NOALLOC_SCRATCH_BYTES,
              "no-alloc scratch smaller than its state");
/* Lays the LM state for `cap` pixels out from `base` (256-byte aligned). */
void gm_carve_state(void *base, int64_t cap, int np, gm_state &st);
/* Looks the array's device address up for `device`, which must be the
 * current device; once per device (later calls return at once).          */
gm_status gm_scratch_init(int device);
/* The device address found by gm_scratch_init, or NULL before it ran.    */
void *gm_scratch_ptr(int device);

#define GM_TRY(x)                                                              \
    do {                                                                       \
        gm_status _s = (x);                                                    \
        if (_s != GM_OK) return _s;                                            \
    } while (0)

/* Context internals shared with the serving loop (gm_capi.cpp). */
const gm_dsetup *gm_ctx_device_setup(const gm_ctx *c);
const gm_dsetup &gm_ctx_host_setup(const gm_ctx *c);
const gm_lanecfg &gm_ctx_lanes(const gm_ctx *c);
const gm_setup &gm_ctx_public_setup(const gm_ctx *c);
int gm_ctx_device(const gm_ctx *c);

#endif
