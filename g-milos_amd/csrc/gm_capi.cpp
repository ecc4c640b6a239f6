/*
 * gm_capi.cpp -- C ABI of the inversion engine (gmilos.h): contexts, device
 * set-up tables, workspace and the host-side LM round loop.
 *
 * The reference drives its engine from gmilos.cu:914-1322: constants go to
 * __constant__ symbols, the cube is split over streams and threads, and one
 * kernel runs the whole per-pixel LM with device-side child launches.  Here a
 * context holds the set-up in device memory, the LM state lives in HBM as
 * per-pixel records (models, synthesis constants, normal equations) beside
 * per-pixel scalar arrays, and each LM round is four stream-ordered launches
 * over compacted pixel lists: step solve -> synthesis + chi^2 ->
 * accept/reject -> response functions + normal equations for accepted
 * pixels.  gm_invert only enqueues; gm_last_stats reads the round counters.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gm_internal.h"

namespace {

constexpr int kMaxRounds = GM_MAX_ROUNDS;

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
};

} // namespace

struct gm_ctx {
    gm_setup setup;
    gm_dsetup host_ds;
    gm_dsetup *dS = nullptr;
    gm_lanecfg lc;
    int device = 0;
    /* LM state workspace */
    gm_state st;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    /* scratch for the piecewise API */
    float *scr_beta = nullptr;
    int64_t scr_cap = 0;
    /* round bookkeeping */
    int32_t *h_counts = nullptr;     /* pinned [kMaxRounds][3]                */
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int timing = 0;
    std::vector<Events> kev;         /* per-kernel events when timing is on   */
    std::vector<int> kclass;         /* their classes: 0 synth 1 rf 2 solve 3 other */
    gm_stats stats;
    /* gm_invert returns once its work is enqueued; the round counters and
     * event times of the last call are read back on the first gm_last_stats */
    int stats_pending = 0;
    int pending_miter = 0;
    int64_t slight_map_npix = 0;     /* pixels covered by the stray-light map */
};

static gm_status hip_ok(hipError_t e)
{
    if (e == hipSuccess) return GM_OK;
    return (e == hipErrorOutOfMemory) ? GM_ERR_NOMEM : GM_ERR_HIP;
}

#define GM_HIP(x) GM_TRY(hip_ok(x))

extern "C" const char *gm_status_string(gm_status st)
{
    switch (st) {
    case GM_OK: return "ok";
    case GM_ERR_INVALID: return "invalid argument";
    case GM_ERR_HIP: return "HIP runtime error";
    case GM_ERR_NOMEM: return "out of device memory";
    case GM_ERR_UNSUPPORTED: return "unsupported configuration";
    case GM_ERR_NOT_FOUND: return "not found";
    case GM_CANCELLED: return "cancelled";
    }
    return "unknown status";
}

gm_lanecfg gm_lane_config(const gm_dsetup &h)
{
    gm_lanecfg lc;
    const int n = h.nlambda;
    if (n <= 32) { lc.lpp = 32; lc.nlpl = 1; }
    else if (n <= 64) { lc.lpp = 64; lc.nlpl = 1; }
    else if (n <= 128) { lc.lpp = 64; lc.nlpl = 2; }
    else { lc.lpp = 64; lc.nlpl = 4; }
    lc.ppw = 64 / lc.lpp;
    lc.ext = h.ext;
    lc.np = h.np;
    lc.mfma = (h.use_psf && h.psf_mfma && lc.lpp == 32 && lc.nlpl == 1 && !h.ext) ? 1 : 0;
    const int padlo = h.psf_dlo < 0 ? -h.psf_dlo : 0;
    const int padhi = h.psf_dhi > 0 ? h.psf_dhi : 0;
    lc.padlo = padlo;
    lc.rows = lc.lpp * lc.nlpl + padlo + padhi;
    return lc;
}

/* Builds the device tables.  The PSF correlation of the reference
 * (convolution.cu:153-241) applies all nlambda taps centred at nlambda/2;
 * taps whose combined tail mass is below psf_tail_tol are dropped, which
 * changes results by less than that fraction of the signal.               */
static gm_status build_dsetup(const gm_setup &s, gm_dsetup &d)
{
    std::memset(&d, 0, sizeof d);
    const int n = s.nlambda;
    if (n < 2 || n > GM_MAX_LAMBDA) return GM_ERR_INVALID;
    if (s.nlines < 1 || s.nlines > GM_MAX_LINES) return GM_ERR_INVALID;
    if (!(s.noise > 0)) return GM_ERR_INVALID;
    if (s.miter < 0 || s.miter > kMaxRounds) return GM_ERR_INVALID;
    /* a free filling factor needs the stray light it is the fraction of
     * (the reference reads the profile only then, gmilos.cu:688-716) */
    if (s.fix[10] && !s.use_slight) return GM_ERR_INVALID;
    if (s.use_slight)
        for (int k = 0; k < 4 * n; ++k)
            if (!std::isfinite(s.slight[k])) return GM_ERR_INVALID;
    if (s.ce_line < 0 || s.ce_line >= s.nlines) return GM_ERR_INVALID;
    if (s.precision != GM_PREC_FP32 && s.precision != GM_PREC_MIXED) return GM_ERR_INVALID;
    if (s.solver != GM_SOLVE_SVD && s.solver != GM_SOLVE_LDL && s.solver != GM_SOLVE_JACOBI &&
        s.solver != GM_SOLVE_CORDIC)
        return GM_ERR_INVALID;
    d.nlambda = n;
    d.nlines = s.nlines;
    for (int L = 0; L < s.nlines; ++L) {
        const gm_zeeman &z = s.zee[L];
        if (z.n_pi < 0 || z.n_pi > GM_MAX_COMP || z.n_sig < 0 || z.n_sig > GM_MAX_COMP)
            return GM_ERR_INVALID;
        d.ncomp[L][0] = z.n_pi;
        d.ncomp[L][1] = z.n_sig;
        d.ncomp[L][2] = z.n_sig;
        for (int i = 0; i < z.n_pi; ++i) {
            d.comp_nu[L][0][i] = (float)z.nup[i];
            d.comp_w[L][0][i] = (float)z.wep[i];
        }
        for (int i = 0; i < z.n_sig; ++i) {
            d.comp_nu[L][1][i] = (float)z.nub[i];
            d.comp_w[L][1][i] = (float)z.web[i];
            /* red shifts: the reference mirrors the blue ones
             * (nurB[i] = -nubB[N_SIG-1-i], mil_sinrf.cu:74-76) */
            d.comp_nu[L][2][i] = (float)(-z.nub[z.n_sig - 1 - i]);
            d.comp_w[L][2][i] = (float)z.wer[i];
        }
        d.l0[L] = (float)s.lambda0[L];
        d.fo[L] = (float)z.fo;
        for (int k = 0; k < n; ++k) d.dl[L][k] = (float)(s.lambda[k] - s.lambda0[L]);
    }
    d.use_psf = s.use_psf ? 1 : 0;
    const int c = n / 2;
    if (d.use_psf) {
        double tot = 0;
        for (int j = 0; j < n; ++j) tot += std::fabs(s.psf[j]);
        if (!(tot > 0)) return GM_ERR_INVALID;
        int jlo = 0, jhi = n - 1;
        double budget = s.psf_tail_tol > 0 ? s.psf_tail_tol * tot : 0.0;
        double cut = 0;
        while (jlo < jhi) {
            const double lo = std::fabs(s.psf[jlo]), hi = std::fabs(s.psf[jhi]);
            const double take = lo <= hi ? lo : hi;
            if (cut + take > budget) break;
            cut += take;
            if (lo <= hi) ++jlo; else --jhi;
        }
        /* many taps, 9 parameters, up to 32 wavelengths: k_rf correlates on
         * the matrix cores, where the dense product costs the same for all
         * taps as for a few, so every kernel then applies all of them */
        const bool ext = s.use_slight || s.fix[9] || s.init_model[9] > 0.0001;
        d.psf_mfma = (GM_MFMA_RF && n <= 32 && !ext && jhi - jlo + 1 >= GM_MFMA_MIN_TAPS) ? 1 : 0;
        if (d.psf_mfma) { jlo = 0; jhi = n - 1; }
        d.psf_dlo = jlo - c;
        d.psf_dhi = jhi - c;
        for (int j = jlo; j <= jhi; ++j) d.psf_tap[j - jlo] = (float)s.psf[j];
        for (int k = 0; k < n; ++k) {
            double r = 0;
            for (int j = 0; j < n; ++j) {
                const int m = k + j - c;
                if (m >= 0 && m < n) r += s.psf[j];
            }
            d.psf_edge[k] = (float)(1.0 - r);
        }
    } else {
        d.psf_dlo = d.psf_dhi = 0;
    }
    for (int i = 0; i < 4; ++i) {
        d.weight[i] = (float)s.weight[i];
        d.wsig[i] = (float)(s.weight[i] / s.noise);
    }
    d.wsig_equal = d.wsig[0] == d.wsig[1] && d.wsig[0] == d.wsig[2] && d.wsig[0] == d.wsig[3];
    d.mu = (float)s.mu;
    d.shi_scale = s.rf_reference_rules ? 1.0f : 2.0f;
    d.freemask = 0;
    for (int p = 0; p < 9; ++p)
        if (s.fix[p]) d.freemask |= (1u << p);
    for (int p = 0; p < 11; ++p) d.fix[p] = s.fix[p];
    d.nfree_chi = 4 * n - 10;   /* nfree = nlambda*NPARMS - NTERMS, milosUtils.cu:585 */
    d.noise = s.noise;
    d.ilambda = s.ilambda;
    d.toplim = s.toplim;
    d.miter = s.miter;
    d.logclambda = s.logclambda;
    for (int f = 0; f < GM_NMODEL; ++f) d.init_model[f] = s.init_model[f];
    for (int k = 0; k < n; ++k) d.lambda[k] = s.lambda[k];
    d.use_slight = s.use_slight ? 1 : 0;
    d.ext = (d.use_slight || s.fix[9] || s.init_model[9] > 0.0001) ? 1 : 0;
    /* free filling factor: eleven solved slots; the reference runs
     * lm_mils_11 when mac is free too, else the 10-term lm_mils
     * (gmilos.cu:1234-1239), which sets nfree and the RF layout */
    d.np = s.fix[10] ? 11 : 9;
    d.nt_ref = (s.fix[9] && s.fix[10]) ? 11 : 10;
    d.nfree_chi = 4 * n - d.nt_ref;
    if (d.np == 11) {
        if (d.nt_ref == 11) d.freemask |= (1u << 9);
        d.freemask |= (1u << 10);
    }
    d.mac_l0 = s.lambda0[0];
    d.mac_tail_tol = s.psf_tail_tol > 0 ? (float)s.psf_tail_tol : 0.f;
    for (int k = 0; k < n; ++k) d.psf_full[k] = d.use_psf ? (float)s.psf[k] : 0.f;
    if (d.use_slight)
        for (int k = 0; k < 4 * n; ++k) d.slight[k] = s.slight[k];
    /* classical-estimate range: the whole grid for one line, else the
     * samples nearest to the chosen line */
    d.ce_lo = 0;
    d.ce_hi = n;
    if (s.nlines > 1) {
        int first = -1, last = -1;
        for (int k = 0; k < n; ++k) {
            int best = 0;
            for (int L = 1; L < s.nlines; ++L)
                if (std::fabs(s.lambda[k] - s.lambda0[L]) < std::fabs(s.lambda[k] - s.lambda0[best]))
                    best = L;
            if (best == s.ce_line) {
                if (first < 0) first = k;
                last = k;
            }
        }
        if (first >= 0) { d.ce_lo = first; d.ce_hi = last + 1; }
    }
    d.ce_l0 = s.lambda0[s.ce_line];
    d.ce_geff = s.zee[s.ce_line].geff;
    return GM_OK;
}

extern "C" gm_status gm_ctx_create(const gm_setup *setup, gm_ctx **out)
{
    if (!setup || !out) return GM_ERR_INVALID;
    *out = nullptr;
    gm_ctx *c = new (std::nothrow) gm_ctx();
    if (!c) return GM_ERR_NOMEM;
    c->setup = *setup;
    gm_status s = build_dsetup(*setup, c->host_ds);
    if (s != GM_OK) { delete c; return s; }
    c->lc = gm_lane_config(c->host_ds);
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return GM_ERR_HIP; }
    /* a wide PSF table applied with every tap at 129-256 wavelengths stages
     * more rows per pixel than a workgroup's LDS holds */
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
        hipSuccess) {
        delete c;
        return GM_ERR_HIP;
    }
    if (gm_lds_bytes(c->lc) > (size_t)lds_max) { delete c; return GM_ERR_UNSUPPORTED; }
    /* device address of this device's gm_invert_noalloc state (once per
     * device) */
    gm_status ss = gm_scratch_init(c->device);
    if (ss != GM_OK) { delete c; return ss; }
    if (hipMalloc(&c->dS, sizeof(gm_dsetup)) != hipSuccess) { delete c; return GM_ERR_NOMEM; }
    if (hipMemcpy(c->dS, &c->host_ds, sizeof(gm_dsetup), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(c->dS);
        delete c;
        return GM_ERR_HIP;
    }
    if (hipHostMalloc((void **)&c->h_counts, sizeof(int32_t) * 3 * (kMaxRounds + 1)) != hipSuccess) {
        hipFree(c->dS);
        delete c;
        return GM_ERR_NOMEM;
    }
    hipEventCreate(&c->ev_start);
    hipEventCreate(&c->ev_stop);
    std::memset(&c->st, 0, sizeof c->st);
    std::memset(&c->stats, 0, sizeof c->stats);
    *out = c;
    return GM_OK;
}

extern "C" gm_status gm_ctx_destroy(gm_ctx *c)
{
    if (!c) return GM_ERR_INVALID;
    hipDeviceSynchronize();
    if (c->ws) hipFree(c->ws);
    if (c->scr_beta) hipFree(c->scr_beta);
    if (c->dS) hipFree(c->dS);
    if (c->h_counts) hipHostFree(c->h_counts);
    for (auto &e : c->kev) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    if (c->ev_start) hipEventDestroy(c->ev_start);
    if (c->ev_stop) hipEventDestroy(c->ev_stop);
    delete c;
    return GM_OK;
}

extern "C" gm_status gm_ctx_set_stray_map(gm_ctx *c, const float *map, int64_t npix)
{
    if (!c || (map && npix <= 0)) return GM_ERR_INVALID;
    if (map && !c->host_ds.use_slight) return GM_ERR_INVALID;
    /* the device copy of the set-up is read by work in flight */
    GM_HIP(hipDeviceSynchronize());
    GM_HIP(hipMemcpy(&c->dS->slight_map, &map, sizeof(map), hipMemcpyHostToDevice));
    c->host_ds.slight_map = map;
    c->slight_map_npix = map ? npix : 0;
    return GM_OK;
}

/* A stray-light map must cover every pixel of a call. */
static bool map_covers(const gm_ctx *c, int64_t npix)
{
    return !c->host_ds.slight_map || npix <= c->slight_map_npix;
}

extern "C" gm_status gm_ctx_set_timing(gm_ctx *c, int32_t enable)
{
    if (!c) return GM_ERR_INVALID;
    c->timing = enable ? 1 : 0;
    return GM_OK;
}

/* Waits for the last gm_invert and turns its round counters and event
 * times into gm_stats. */
static gm_status finish_stats(gm_ctx *c)
{
    if (!c->stats_pending) return GM_OK;
    GM_HIP(hipEventSynchronize(c->ev_stop));
    c->stats_pending = 0;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    c->stats.ms_total = ms;
    const int64_t npix = c->stats.npix;
    c->stats.synth_evals = npix;
    c->stats.rf_evals = npix;
    int rounds = 0;
    for (int r = 0; r < c->pending_miter; ++r) {
        const int32_t *hc = c->h_counts + 3 * r;
        if (hc[0] > 0) rounds = r + 1;
        c->stats.solves += hc[0];
        c->stats.svd_solves += hc[1];
        c->stats.synth_evals += hc[0];
        c->stats.rf_evals += hc[2];
    }
    c->stats.lm_rounds = rounds;
    if (c->timing) {
        for (size_t i = 0; i < c->kclass.size(); ++i) {
            float t = 0;
            (void)hipEventElapsedTime(&t, c->kev[i].a, c->kev[i].b);
            switch (c->kclass[i]) {
            case 0: c->stats.ms_synth += t; break;
            case 1: c->stats.ms_rf += t; break;
            case 2: c->stats.ms_solve += t; break;
            default: c->stats.ms_other += t; break;
            }
        }
    }
    return GM_OK;
}

extern "C" gm_status gm_last_stats(const gm_ctx *c, gm_stats *s)
{
    if (!c || !s) return GM_ERR_INVALID;
    GM_TRY(finish_stats(const_cast<gm_ctx *>(c)));
    *s = c->stats;
    return GM_OK;
}

/* Accessors for the serving loop and the learned-guess driver. */
const gm_dsetup *gm_ctx_device_setup(const gm_ctx *c) { return c->dS; }
const gm_dsetup &gm_ctx_host_setup(const gm_ctx *c) { return c->host_ds; }
const gm_lanecfg &gm_ctx_lanes(const gm_ctx *c) { return c->lc; }
const gm_setup &gm_ctx_public_setup(const gm_ctx *c) { return c->setup; }
int gm_ctx_device(const gm_ctx *c) { return c->device; }

void gm_carve_state(void *base, int64_t cap, int np, gm_state &st)
{
    char *p = (char *)base;
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 255) & ~(size_t)255;
        return (void *)r;
    };
    st.cap = cap;
    st.best = (double *)take(sizeof(double) * GM_MODEL_REC * cap);
    st.trial = (double *)take(sizeof(double) * GM_MODEL_REC * cap);
    const bool p11 = np == 11;
    st.ab_rec = p11 ? GM_AB_REC11 : GM_AB_REC9;
    st.alpha = (float *)take(sizeof(float) * GM_AB_REC11 * cap);
    st.beta = st.alpha + (p11 ? GM_AB_BETA11 : GM_AB_BETA9);
    st.pre_best = (float *)take(sizeof(float) * GM_PRE_REC * cap);
    st.pre_trial = (float *)take(sizeof(float) * GM_PRE_REC * cap);
    st.chi_trial = (double *)take(sizeof(double) * cap);
    st.ochisqr = (double *)take(sizeof(double) * cap);
    st.chisqr0 = (double *)take(sizeof(double) * cap);
    st.flambda = (double *)take(sizeof(double) * cap);
    st.factor = (double *)take(sizeof(double) * cap);
    st.iter = (int32_t *)take(sizeof(int32_t) * cap);
    st.nacc = (int32_t *)take(sizeof(int32_t) * cap);
    for (int i = 0; i < 4; ++i) st.list[i] = (int32_t *)take(sizeof(int32_t) * cap);
    st.counts = (int32_t *)take(sizeof(int32_t) * 4);
    st.hist = (int32_t *)take(sizeof(int32_t) * 3 * (kMaxRounds + 2));
}

/* The LM state of gm_invert: one allocation sized for `cap` pixels, grown
 * (after a device synchronisation) when a call needs more. */
static gm_status ensure_state(gm_ctx *c, int64_t npix)
{
    if (npix <= c->st.cap) return GM_OK;
    const int64_t cap = ((npix + 255) / 256) * 256;
    const size_t bytes = gm_state_bytes(cap);
    if (c->ws) {
        hipDeviceSynchronize();
        hipFree(c->ws);
        c->ws = nullptr;
        c->st.cap = 0;
    }
    GM_HIP(hipMalloc(&c->ws, bytes));
    c->ws_bytes = bytes;
    gm_carve_state(c->ws, cap, c->lc.np, c->st);
    return GM_OK;
}

static Events &kernel_events(gm_ctx *c, size_t i)
{
    while (c->kev.size() <= i) {
        Events e;
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        c->kev.push_back(e);
    }
    return c->kev[i];
}

/* Argument checks shared by gm_invert and gm_invert_noalloc; resets the
 * statistics.  *done: nothing to launch (empty call). */
static gm_status invert_prologue(gm_ctx *c, const float *obs, int64_t npix, const double *init,
                                 const gm_out *out, bool *done)
{
    *done = true;
    if (!c || !out || npix < 0) return GM_ERR_INVALID;
    if (npix == 0) {
        std::memset(&c->stats, 0, sizeof c->stats);
        c->stats_pending = 0;
        return GM_OK;
    }
    if (!obs || npix > (int64_t)0x7fffffff) return GM_ERR_INVALID;
    if (c->setup.init_mode == GM_INIT_PER_PIXEL && !init) return GM_ERR_INVALID;
    if (!map_covers(c, npix)) return GM_ERR_INVALID;
    std::memset(&c->stats, 0, sizeof c->stats);
    c->stats.npix = npix;
    c->stats.psf_taps = c->host_ds.use_psf ? (c->host_ds.psf_dhi - c->host_ds.psf_dlo + 1) : 0;
    *done = false;
    return GM_OK;
}

/* The launches of one inversion over the LM state `st` (capacity >= npix):
 * start state and first normal equations, miter LM rounds, synthetic output
 * of the last evaluated models, outputs.  With `timing`, per-kernel events
 * (created on first use) time each launch. */
static gm_status invert_launch(gm_ctx *c, const gm_state &st, const float *obs, int64_t npix,
                               const double *init, const gm_out *out, hipStream_t s, bool timing)
{
    const int mixed = c->setup.precision == GM_PREC_MIXED;
    const int solver = c->setup.solver;
    const gm_lanecfg &lc = c->lc;
    int32_t *cnt = st.counts;
    c->kclass.clear();
    c->stats_pending = 0;
    size_t kidx = 0;
    auto kbegin = [&](int cls) {
        if (!timing) return;
        hipEventRecord(kernel_events(c, kidx).a, s);
        c->kclass.push_back(cls);
    };
    auto kend = [&]() {
        if (!timing) return;
        hipEventRecord(kernel_events(c, kidx).b, s);
        ++kidx;
    };

    GM_HIP(hipEventRecord(c->ev_start, s));
    kbegin(3);
    gm_launch_init_state(c->dS, st, obs, init, c->setup.init_mode, c->setup.nlambda, 0, npix,
                         st.list[0], s);
    gm_launch_set_count(&cnt[0], (int32_t)npix, s);
    gm_launch_set_count(&cnt[3], 0, s);
    kend();
    kbegin(1);
    gm_launch_rf(lc, c->dS, st.best, st.pre_best, -GM_MODEL_REC, obs, st.list[0], &cnt[0], 0, npix, st.alpha,
                 st.beta, -st.ab_rec, st.ochisqr, nullptr, nullptr, mixed, s);
    kend();
    kbegin(3);
    gm_launch_init_chi(st, 0, npix, s);
    kend();
    int cur = 0;
    const int miter = c->setup.miter > 0 ? c->setup.miter : 1;  /* do-while */
    /* hist row r: [active in round r, SVD solves in round r, accepted in
     * round r]; the solve kernel of round r writes hist[3r] and the accepted
     * count of round r-1 into hist[3r - 1], the decision kernel hist[3r + 1];
     * empty trailing rounds cost four short launches */
    int32_t *hist = st.hist + 3;
    for (int r = 0; r < miter; ++r) {
        const int nxt = 1 - cur;
        kbegin(2);
        gm_launch_solve(c->dS, st, st.list[cur], &cnt[cur], npix, mixed, solver, &cnt[nxt],
                        &cnt[2], hist + 3 * r, st.list[3], &cnt[3], s, lc.np,
                        gm_cordic10(c->setup, lc.np));
        kend();
        kbegin(0);
        gm_launch_synth(lc, c->dS, st.trial, st.pre_trial, -GM_MODEL_REC, obs, st.list[cur], &cnt[cur], 0,
                        npix, st.chi_trial, nullptr, mixed, s);
        kend();
        kbegin(3);
        gm_launch_decide(c->dS, st, st.list[cur], &cnt[cur], npix, st.list[nxt], &cnt[nxt],
                         st.list[2], &cnt[2], &cnt[3], nullptr, nullptr, mixed, s, 1,
                         hist + 3 * r + 1);
        kend();
        kbegin(1);
        /* alpha_first = 0: the filling-factor response of the reference's
         * loop (milosUtils.cu:673) */
        /* accepted pixels: normal equations at the trial, which becomes the
         * best state (copied by the RF kernel) */
        gm_launch_rf(lc, c->dS, st.trial, st.pre_trial, -GM_MODEL_REC, obs, st.list[2], &cnt[2], 0, npix,
                     st.alpha, st.beta, -st.ab_rec, nullptr, nullptr, nullptr, mixed, s, 0, st.best,
                     st.pre_best);
        kend();
        cur = nxt;
    }
    gm_launch_copy_count(&cnt[2], hist + 3 * miter - 1, s);
    /* the synthetic output is the last evaluated model's spectrum (the
     * reference's spectra buffer after its loop): one synthesis of every
     * pixel's last trial, instead of writing every trial's spectrum */
    if (out->synth) {
        kbegin(0);
        gm_launch_synth(lc, c->dS, st.trial, st.pre_trial, -GM_MODEL_REC, nullptr, nullptr, nullptr,
                        npix, npix, nullptr, out->synth, mixed, s);
        kend();
    }
    kbegin(3);
    gm_launch_finalize(st, npix, *out, s);
    kend();
    GM_HIP(hipMemcpyAsync(c->h_counts, hist, sizeof(int32_t) * 3 * miter, hipMemcpyDeviceToHost, s));
    GM_HIP(hipEventRecord(c->ev_stop, s));
    GM_HIP(hipGetLastError());
    c->stats_pending = 1;
    c->pending_miter = miter;
    return GM_OK;
}

extern "C" gm_status gm_invert(gm_ctx *c, const float *obs, int64_t npix,
                               const double *init, const gm_out *out, void *stream)
{
    bool done = false;
    GM_TRY(invert_prologue(c, obs, npix, init, out, &done));
    if (done) return GM_OK;
    GM_TRY(ensure_state(c, npix));
    return invert_launch(c, c->st, obs, npix, init, out, (hipStream_t)stream, c->timing != 0);
}

/* gm_invert over this device's fixed state array (gm_noalloc.hip), whose
 * address gm_ctx_create looked up: no allocation of any kind.  The state's
 * layout is fixed at GM_NOALLOC_MAX_PIXELS pixels.  Everything the kernels
 * need at zero or at a start value is set by the launches themselves (list
 * counts, per-round counts, the start state), so the contents left by an
 * earlier call do not matter. */
extern "C" gm_status gm_invert_noalloc(gm_ctx *c, const float *obs, int64_t npix,
                                       const double *init, const gm_out *out, void *stream)
{
    if (c && npix > GM_NOALLOC_MAX_PIXELS) return GM_ERR_INVALID;
    bool done = false;
    GM_TRY(invert_prologue(c, obs, npix, init, out, &done));
    if (done) return GM_OK;
    void *base = gm_scratch_ptr(c->device);
    if (!base) return GM_ERR_HIP;
    gm_state st;
    std::memset(&st, 0, sizeof st);
    gm_carve_state(base, GM_NOALLOC_MAX_PIXELS, c->lc.np, st);
    return invert_launch(c, st, obs, npix, init, out, (hipStream_t)stream, false);
}

/* ---- path pieces ---------------------------------------------------------- */
extern "C" gm_status gm_synth(gm_ctx *c, const double *models, int64_t npix, float *spectra,
                              void *stream)
{
    if (!c || npix < 0) return GM_ERR_INVALID;
    if (!map_covers(c, npix)) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!models || !spectra) return GM_ERR_INVALID;
    gm_launch_synth(c->lc, c->dS, models, nullptr, 0, nullptr, nullptr, nullptr, npix, npix, nullptr,
                    spectra, 0, (hipStream_t)stream);
    return hip_ok(hipGetLastError());
}

extern "C" gm_status gm_synth_rf(gm_ctx *c, const double *models, int64_t npix, float *spectra,
                                 float *rf, void *stream)
{
    if (!c || npix < 0) return GM_ERR_INVALID;
    if (!map_covers(c, npix)) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!models) return GM_ERR_INVALID;
    gm_launch_rf(c->lc, c->dS, models, nullptr, 0, nullptr, nullptr, nullptr, npix, npix, nullptr, nullptr,
                 0, nullptr, spectra, rf, 0, (hipStream_t)stream);
    return hip_ok(hipGetLastError());
}

extern "C" gm_status gm_normal_equations(gm_ctx *c, const double *models, const float *obs,
                                         int64_t npix, float *alpha, float *beta, double *chisqr,
                                         void *stream)
{
    if (!c || npix < 0) return GM_ERR_INVALID;
    if (!map_covers(c, npix)) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!models || !obs) return GM_ERR_INVALID;
    const int mixed = c->setup.precision == GM_PREC_MIXED;
    gm_launch_rf(c->lc, c->dS, models, nullptr, 0, obs, nullptr, nullptr, npix, npix, alpha, beta, 0,
                 chisqr, nullptr, nullptr, mixed, (hipStream_t)stream);
    return hip_ok(hipGetLastError());
}

extern "C" gm_status gm_step_solve(gm_ctx *c, const float *alpha, const float *beta,
                                   const double *flambda, int64_t npix, double *delta,
                                   void *stream)
{
    if (!c || npix < 0) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!alpha || !beta || !flambda || !delta) return GM_ERR_INVALID;
    gm_launch_step_solve(alpha, beta, flambda, npix, delta, c->setup.precision == GM_PREC_MIXED,
                         c->setup.solver, (hipStream_t)stream, c->lc.np,
                         gm_cordic10(c->setup, c->lc.np));
    return hip_ok(hipGetLastError());
}

extern "C" gm_status gm_classical_estimates(gm_ctx *c, const float *obs, int64_t npix,
                                            int32_t for_initial_use, double *models,
                                            void *stream)
{
    if (!c || npix < 0) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!obs || !models) return GM_ERR_INVALID;
    gm_launch_ce(c->dS, obs, npix, for_initial_use ? 1 : 0, models, (hipStream_t)stream);
    return hip_ok(hipGetLastError());
}
