/*
 * gm_guided.hip -- training-set generation, network training and the
 * learned-guess inversion with the fixed-guess fallback (gmilos.h).
 *
 * No reference counterpart: the reference always starts the LM from
 * init.mod + classical estimates (milosUtils.cu:615-628).  Here
 *   gm_draw_models / gm_add_noise  draw true models and noisy profiles on the
 *                                  device (counter-based hashing, so a draw
 *                                  depends only on seed and pixel index);
 *   gm_mlp_fit                     synthesises a training set with the
 *                                  engine's own synthesis and trains the
 *                                  network with AdamW;
 *   gm_invert_guided               predicts start models, runs the LM from
 *                                  them, and re-inverts from the fixed guess
 *                                  every pixel whose chi^2 exceeds the
 *                                  fallback threshold, keeping the lower
 *                                  chi^2 of the two.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "gm_internal.h"
#include "gmilos.h"

namespace {

/* splitmix64 finaliser (Steele, Lea & Flood 2014) as a counter-based hash */
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t idx, uint32_t k)
{
    const uint64_t h = mix64(mix64(seed ^ (idx * 0x2545f4914f6cdd1dull)) + k);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);   /* [0, 1) */
}

__global__ void k_draw_models(gm_prior pr, uint64_t seed, int64_t offset, int64_t n,
                              double *__restrict__ models)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t idx = (uint64_t)(offset + p);
        double *m = models + p * GM_NMODEL;
        for (int f = 0; f < GM_NMODEL; ++f)
            m[f] = pr.lo[f] + (pr.hi[f] - pr.lo[f]) * uniform01(seed, idx, (uint32_t)f);
        if (pr.weak_fraction > 0 && uniform01(seed, idx, 20) < pr.weak_fraction)
            m[1] = pr.lo[1] + (pr.weak_b_max - pr.lo[1]) * uniform01(seed, idx, 21);
    }
}

/* Gaussian noise sigma * Ic (Ic = S0 + S1 mu) on all Stokes samples,
 * Box-Muller on hashed uniforms. */
__global__ void k_add_noise(const double *__restrict__ models, float *__restrict__ spectra,
                            int64_t n, int nl4, double sigma, double mu, uint64_t seed)
{
    const int64_t total = n * nl4;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = i / nl4;
        const double ic = models[p * GM_NMODEL + 7] + models[p * GM_NMODEL + 8] * mu;
        const double u1 = 1.0 - uniform01(seed, (uint64_t)i, 101);
        const double u2 = uniform01(seed, (uint64_t)i, 102);
        const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        spectra[i] = (float)((double)spectra[i] + sigma * ic * g);
    }
}

/* idx <- pixels with chisqr > thr (order not preserved); *count. */
__global__ void k_select_above(const float *__restrict__ chisqr, int64_t n, float thr,
                               int32_t *__restrict__ idx, int32_t *__restrict__ count)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const float c = chisqr[p];
        if (!(c <= thr)) idx[atomicAdd(count, 1)] = (int32_t)p;
    }
}

__global__ void k_gather_rows(const float *__restrict__ src, const int32_t *__restrict__ idx,
                              int64_t n, int row, float *__restrict__ dst)
{
    const int64_t total = n * row;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = i / row;
        const int r = (int)(i % row);
        dst[i] = src[(int64_t)idx[k] * row + r];
    }
}

/* Keep the fallback result where its chi^2 is lower; count replacements and
 * add the fallback's iterations to the total. */
__global__ void k_merge(const int32_t *__restrict__ idx, int64_t n, int nl4, gm_out fb, gm_out out,
                        unsigned long long *__restrict__ acc)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx[k];
        atomicAdd(&acc[1], (unsigned long long)fb.iters[k]);
        if (!(fb.chisqr[k] < out.chisqr[p])) continue;
        atomicAdd(&acc[2], 1ull);
        for (int f = 0; f < GM_NMODEL; ++f) out.model[p * GM_NMODEL + f] = fb.model[k * GM_NMODEL + f];
        out.chisqr[p] = fb.chisqr[k];
        out.iters[p] = fb.iters[k];
        if (out.naccs && fb.naccs) out.naccs[p] = fb.naccs[k];
        if (out.synth && fb.synth)
            for (int j = 0; j < nl4; ++j) out.synth[p * nl4 + j] = fb.synth[k * nl4 + j];
    }
}

__global__ void k_sum_iters(const int32_t *__restrict__ it, int64_t n,
                            unsigned long long *__restrict__ acc)
{
    unsigned long long s = 0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x)
        s += (unsigned long long)it[p];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&acc[0], s);
}

int grid1(int64_t n, int b)
{
    int64_t g = (n + b - 1) / b;
    return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    gm_status alloc(size_t b) { return hipMalloc(&p, b) == hipSuccess ? GM_OK : GM_ERR_NOMEM; }
};

} // namespace

extern "C" void gm_prior_default(gm_prior *pr)
{
    if (!pr) return;
    const double lo[GM_NMODEL] = {1.0, 0.0, -3.0, 0.020, 0.01, 0.0, 0.0, 0.1, 0.5, 0.0, 1.0};
    const double hi[GM_NMODEL] = {20.0, 3000.0, 3.0, 0.060, 0.5, 180.0, 180.0, 0.4, 0.9, 0.0, 1.0};
    for (int f = 0; f < GM_NMODEL; ++f) {
        pr->lo[f] = lo[f];
        pr->hi[f] = hi[f];
    }
    pr->weak_fraction = 0.0;
    pr->weak_b_max = 100.0;
}

extern "C" gm_status gm_draw_models(const gm_prior *pr, uint64_t seed, int64_t offset, int64_t npix,
                                    double *models, void *stream)
{
    if (!pr || npix < 0 || offset < 0) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!models) return GM_ERR_INVALID;
    hipLaunchKernelGGL(k_draw_models, dim3(grid1(npix, 256)), dim3(256), 0, (hipStream_t)stream, *pr,
                       seed, offset, npix, models);
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_add_noise(const double *models, float *spectra, int64_t npix, int32_t nlambda,
                                  double sigma, double mu, uint64_t seed, void *stream)
{
    if (npix < 0 || nlambda < 1 || nlambda > GM_MAX_LAMBDA || !(sigma >= 0)) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    if (!models || !spectra) return GM_ERR_INVALID;
    hipLaunchKernelGGL(k_add_noise, dim3(grid1(npix * 4 * nlambda, 256)), dim3(256), 0,
                       (hipStream_t)stream, models, spectra, npix, 4 * nlambda, sigma, mu, seed);
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_mlp_fit(gm_mlp *net, gm_ctx *synth_ctx, const gm_prior *pr, int64_t n_train,
                                int64_t batch, int64_t steps, float lr, float weight_decay,
                                double sigma, uint64_t seed, float *loss_curve, void *stream)
{
    if (!net || !synth_ctx || !pr || n_train < 64 || batch < 64 || (batch % 64) || steps < 1 ||
        batch > n_train || !(lr > 0))
        return GM_ERR_INVALID;
    const gm_setup &su = gm_ctx_public_setup(synth_ctx);
    const int nl4 = 4 * su.nlambda;
    hipStream_t s = (hipStream_t)stream;
    DevBuf obs, models;
    GM_TRY(obs.alloc(sizeof(float) * nl4 * (size_t)n_train));
    GM_TRY(models.alloc(sizeof(double) * GM_NMODEL * (size_t)n_train));
    double *m = (double *)models.p;
    float *o = (float *)obs.p;
    GM_TRY(gm_draw_models(pr, seed, 0, n_train, m, s));
    GM_TRY(gm_synth(synth_ctx, m, n_train, o, s));
    GM_TRY(gm_add_noise(m, o, n_train, su.nlambda, sigma, su.mu, seed ^ 0x5bd1e995ull, s));
    const int64_t nb = n_train / batch;
    const int64_t warm = steps < 500 ? steps / 10 : 50;
    for (int64_t t = 0; t < steps; ++t) {
        /* linear warm-up, cosine decay to 5 % */
        const double fr = (double)t / (double)steps;
        double lt = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * fr)));
        if (t < warm) lt *= (double)(t + 1) / (double)warm;
        const int64_t b0 = (t % nb) * batch;
        float loss = 0.f;
        GM_TRY(gm_mlp_train_step(net, o + b0 * nl4, m + b0 * GM_NMODEL, batch, (float)lt,
                                 weight_decay, loss_curve ? &loss : nullptr, s));
        if (loss_curve) loss_curve[t] = loss;
    }
    return hipStreamSynchronize(s) == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_invert_guided(gm_ctx *learned, gm_ctx *plain, gm_mlp *net, const float *obs,
                                      int64_t npix, double fallback_chisqr, const gm_out *out,
                                      gm_guided_stats *stats, void *stream)
{
    if (!learned || !plain || !net || !out || npix < 0) return GM_ERR_INVALID;
    const gm_setup &sl = gm_ctx_public_setup(learned);
    const gm_setup &sp = gm_ctx_public_setup(plain);
    if (sl.init_mode != GM_INIT_PER_PIXEL || sp.init_mode == GM_INIT_PER_PIXEL ||
        sl.nlambda != sp.nlambda)
        return GM_ERR_INVALID;
    /* the fallback inverts a gathered subset: per-pixel stray-light maps
     * would be indexed by the wrong pixel */
    if (gm_ctx_host_setup(learned).slight_map || gm_ctx_host_setup(plain).slight_map)
        return GM_ERR_INVALID;
    gm_guided_stats st{};
    st.npix = npix;
    if (npix == 0) {
        if (stats) *stats = st;
        return GM_OK;
    }
    if (!obs || !out->model || !out->chisqr || !out->iters) return GM_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    const int nl4 = 4 * sl.nlambda;
    hipEvent_t ev[4];
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) return GM_ERR_HIP;
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard()
        {
            for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    DevBuf guess, misc;
    GM_TRY(guess.alloc(sizeof(double) * GM_NMODEL * (size_t)npix));
    GM_TRY(misc.alloc(sizeof(int32_t) * ((size_t)npix + 64) + 64));
    unsigned long long *acc = (unsigned long long *)misc.p;        /* [4] */
    int32_t *count = (int32_t *)((char *)misc.p + 32);
    int32_t *idx = (int32_t *)((char *)misc.p + 64);
    if (hipMemsetAsync(misc.p, 0, 64, s) != hipSuccess) return GM_ERR_HIP;
    (void)hipEventRecord(ev[0], s);
    GM_TRY(gm_mlp_predict(net, obs, npix, sp.init_model, (double *)guess.p, s));
    (void)hipEventRecord(ev[1], s);
    GM_TRY(gm_invert(learned, obs, npix, (const double *)guess.p, out, s));
    (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(k_sum_iters, dim3(grid1(npix, 256)), dim3(256), 0, s, out->iters, npix, acc);
    hipLaunchKernelGGL(k_select_above, dim3(grid1(npix, 256)), dim3(256), 0, s, out->chisqr, npix,
                       (float)fallback_chisqr, idx, count);
    int32_t nf = 0;
    if (hipMemcpyAsync(&nf, count, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GM_ERR_HIP;
    st.n_fallback = nf;
    if (nf > 0) {
        DevBuf fobs, fout;
        GM_TRY(fobs.alloc(sizeof(float) * nl4 * (size_t)nf));
        const size_t per = sizeof(double) * GM_NMODEL + sizeof(float) + 2 * sizeof(int32_t) +
                           (out->synth ? sizeof(float) * nl4 : 0);
        GM_TRY(fout.alloc(per * (size_t)nf + 1024));
        char *q = (char *)fout.p;
        gm_out fo{};
        fo.model = (double *)q;
        q += sizeof(double) * GM_NMODEL * nf;
        fo.chisqr = (float *)q;
        q += sizeof(float) * nf;
        fo.iters = (int32_t *)q;
        q += sizeof(int32_t) * nf;
        fo.naccs = (int32_t *)q;
        q += sizeof(int32_t) * nf;
        fo.synth = out->synth ? (float *)q : nullptr;
        hipLaunchKernelGGL(k_gather_rows, dim3(grid1((int64_t)nf * nl4, 256)), dim3(256), 0, s, obs, idx,
                           (int64_t)nf, nl4, (float *)fobs.p);
        GM_TRY(gm_invert(plain, (const float *)fobs.p, nf, nullptr, &fo, s));
        hipLaunchKernelGGL(k_merge, dim3(grid1(nf, 256)), dim3(256), 0, s, idx, (int64_t)nf, nl4, fo,
                           *out, acc);
        if (hipStreamSynchronize(s) != hipSuccess) return GM_ERR_HIP;
    }
    (void)hipEventRecord(ev[3], s);
    unsigned long long h[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GM_ERR_HIP;
    st.iters_learned = (int64_t)h[0];
    st.iters_fallback = (int64_t)h[1];
    st.n_replaced = (int64_t)h[2];
    (void)hipEventElapsedTime(&st.ms_predict, ev[0], ev[1]);
    (void)hipEventElapsedTime(&st.ms_lm, ev[1], ev[2]);
    (void)hipEventElapsedTime(&st.ms_fallback, ev[2], ev[3]);
    (void)hipEventElapsedTime(&st.ms_total, ev[0], ev[3]);
    if (stats) *stats = st;
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}
