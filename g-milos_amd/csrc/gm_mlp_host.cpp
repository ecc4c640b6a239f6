/*
 * gm_mlp_host.cpp -- C ABI of the learned-guess network (gmilos.h): parameter
 * storage, forward / backward / AdamW sequencing over the BF16 MFMA kernels
 * of gm_mlp.hip, and batched prediction of start models.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "gm_mlp.h"
#include "gmilos.h"

namespace {

struct Layer {
    int in, out;
    int64_t woff, boff;   /* offsets into the flat FP32 parameter arrays */
};

template <typename T> T *dalloc(size_t n)
{
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return (T *)p;
}

} // namespace

struct gm_mlp {
    int nl, kin;
    Layer L[GM_MLP_LAYERS];
    int64_t nparam = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   /* FP32 flat */
    uint16_t *wb = nullptr;                                          /* BF16 copy */
    uint16_t *wt = nullptr;            /* BF16 weights transposed per layer (inference) */
    int64_t step = 0;
    /* activations for a batch of `cap` rows */
    int64_t cap = 0;
    uint16_t *x = nullptr;                 /* [cap][kin]                 */
    float *z[3] = {nullptr, nullptr, nullptr};   /* pre-activations [cap][256] */
    uint16_t *a[3] = {nullptr, nullptr, nullptr};
    float *y = nullptr;                    /* [cap][NOUT]                */
    float *t = nullptr;                    /* targets [cap][NOUT]        */
    uint16_t *dy = nullptr;                /* [cap][NOUT]                */
    uint16_t *dz[3] = {nullptr, nullptr, nullptr};
    float *loss = nullptr;
    double *base = nullptr;
};

static void free_batch(gm_mlp *n)
{
    void *ps[] = {n->x, n->y, n->t, n->dy, n->z[0], n->z[1], n->z[2], n->a[0], n->a[1], n->a[2],
                  n->dz[0], n->dz[1], n->dz[2]};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    n->x = nullptr; n->y = nullptr; n->t = nullptr; n->dy = nullptr;
    for (int i = 0; i < 3; ++i) { n->z[i] = nullptr; n->a[i] = nullptr; n->dz[i] = nullptr; }
    n->cap = 0;
}

static gm_status ensure_batch(gm_mlp *n, int64_t rows)
{
    if (rows <= n->cap) return GM_OK;
    (void)hipDeviceSynchronize();
    free_batch(n);
    const int64_t cap = ((rows + 127) / 128) * 128;
    const int H = GM_MLP_HIDDEN;
    n->x = dalloc<uint16_t>(cap * n->kin);
    n->y = dalloc<float>(cap * GM_MLP_NOUT);
    n->t = dalloc<float>(cap * GM_MLP_NOUT);
    n->dy = dalloc<uint16_t>(cap * GM_MLP_NOUT);
    for (int i = 0; i < 3; ++i) {
        n->z[i] = dalloc<float>(cap * H);
        n->a[i] = dalloc<uint16_t>(cap * H);
        n->dz[i] = dalloc<uint16_t>(cap * H);
    }
    bool ok = n->x && n->y && n->t && n->dy;
    for (int i = 0; i < 3; ++i) ok = ok && n->z[i] && n->a[i] && n->dz[i];
    if (!ok) { free_batch(n); return GM_ERR_NOMEM; }
    (void)hipMemset(n->x, 0, cap * n->kin * sizeof(uint16_t));
    n->cap = cap;
    return GM_OK;
}

extern "C" gm_status gm_mlp_create(int32_t nlambda, uint64_t seed, gm_mlp **out)
{
    if (!out || nlambda < 2 || nlambda > GM_MAX_LAMBDA) return GM_ERR_INVALID;
    gm_mlp *n = new (std::nothrow) gm_mlp();
    if (!n) return GM_ERR_NOMEM;
    n->nl = nlambda;
    n->kin = ((4 * nlambda + 63) / 64) * 64;
    const int dims[GM_MLP_LAYERS + 1] = {n->kin, GM_MLP_HIDDEN, GM_MLP_HIDDEN, GM_MLP_HIDDEN,
                                         GM_MLP_NOUT};
    int64_t off = 0;
    for (int l = 0; l < GM_MLP_LAYERS; ++l) {
        n->L[l].in = dims[l];
        n->L[l].out = dims[l + 1];
        n->L[l].woff = off;
        off += (int64_t)dims[l] * dims[l + 1];
        n->L[l].boff = off;
        off += dims[l + 1];
    }
    n->nparam = off;
    /* He-style initialisation, zero biases; unused input rows / outputs 0 */
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < GM_MLP_LAYERS; ++l) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / (l == 0 ? 4 * nlambda : dims[l])));
        for (int i = 0; i < dims[l]; ++i)
            for (int j = 0; j < dims[l + 1]; ++j) {
                float w = nd(rng);
                if (l == 0 && i >= 4 * nlambda) w = 0.f;
                if (l == GM_MLP_LAYERS - 1) w = (j < 11) ? w * 0.5f : 0.f;
                h[n->L[l].woff + (int64_t)i * dims[l + 1] + j] = w;
            }
    }
    n->w = dalloc<float>(off);
    n->g = dalloc<float>(off);
    n->m = dalloc<float>(off);
    n->v = dalloc<float>(off);
    n->wb = dalloc<uint16_t>(off);
    n->wt = dalloc<uint16_t>(off);
    n->loss = dalloc<float>(1);
    n->base = dalloc<double>(GM_NMODEL);
    if (!n->w || !n->g || !n->m || !n->v || !n->wb || !n->wt || !n->loss || !n->base) {
        gm_mlp_destroy(n);
        return GM_ERR_NOMEM;
    }
    (void)hipMemcpy(n->w, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
    (void)hipMemset(n->m, 0, off * sizeof(float));
    (void)hipMemset(n->v, 0, off * sizeof(float));
    gm_mlp_to_bf16(n->w, n->wb, off, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) { gm_mlp_destroy(n); return GM_ERR_HIP; }
    *out = n;
    return GM_OK;
}

extern "C" gm_status gm_mlp_destroy(gm_mlp *n)
{
    if (!n) return GM_ERR_INVALID;
    (void)hipDeviceSynchronize();
    free_batch(n);
    void *ps[] = {n->w, n->g, n->m, n->v, n->wb, n->wt, n->loss, n->base};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete n;
    return GM_OK;
}

extern "C" int64_t gm_mlp_num_params(const gm_mlp *n) { return n ? n->nparam : 0; }

extern "C" gm_status gm_mlp_get_params(const gm_mlp *n, float *host, int64_t cnt)
{
    if (!n || !host || cnt < n->nparam) return GM_ERR_INVALID;
    return hipMemcpy(host, n->w, n->nparam * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
               ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_mlp_set_params(gm_mlp *n, const float *host, int64_t cnt)
{
    if (!n || !host || cnt < n->nparam) return GM_ERR_INVALID;
    if (hipMemcpy(n->w, host, n->nparam * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return GM_ERR_HIP;
    gm_mlp_to_bf16(n->w, n->wb, n->nparam, nullptr);
    return hipDeviceSynchronize() == hipSuccess ? GM_OK : GM_ERR_HIP;
}

/* Forward pass over `rows` (a multiple of 128) already in n->x; the FP32
 * pre-activations are kept only for a following backward pass. */
static void forward(gm_mlp *n, int64_t rows, hipStream_t s, bool keep_z = true)
{
    const uint16_t *in = n->x;
    int ld = n->kin;
    for (int l = 0; l < 3; ++l) {
        const Layer &L = n->L[l];
        gm_mlp_gemm(0, 0, GM_EPI_BIAS_GELU, in, ld, n->wb + L.woff, L.out, (int)rows, L.out, L.in,
                    n->w + L.boff, nullptr, 0, keep_z ? n->z[l] : nullptr, n->a[l], L.out, s);
        in = n->a[l];
        ld = L.out;
    }
    const Layer &L = n->L[3];
    gm_mlp_gemm(0, 0, GM_EPI_BIAS, in, ld, n->wb + L.woff, L.out, (int)rows, L.out, L.in,
                n->w + L.boff, nullptr, 0, n->y, nullptr, L.out, s);
}

/* Backward pass from n->dy: gradients into n->g. */
static void backward(gm_mlp *n, int64_t rows, hipStream_t s)
{
    const uint16_t *dcur = n->dy;
    int ldcur = GM_MLP_NOUT;
    for (int l = 3; l >= 0; --l) {
        const Layer &L = n->L[l];
        const uint16_t *xin = (l == 0) ? n->x : n->a[l - 1];
        const int ldx = (l == 0) ? n->kin : GM_MLP_HIDDEN;
        /* dW = X^T . dZ  ([in x rows] . [rows x out]) */
        gm_mlp_gemm(1, 0, GM_EPI_STORE, xin, ldx, dcur, ldcur, L.in, L.out, (int)rows, nullptr,
                    nullptr, 0, n->g + L.woff, nullptr, L.out, s);
        gm_mlp_bias_grad(dcur, rows, L.out, n->g + L.boff, s);
        if (l > 0) {
            /* dZ_prev = (dZ . W^T) * GELU'(Z_prev) */
            gm_mlp_gemm(0, 1, GM_EPI_GELU_GRAD, dcur, ldcur, n->wb + L.woff, L.out, (int)rows, L.in,
                        L.out, nullptr, n->z[l - 1], GM_MLP_HIDDEN, nullptr, n->dz[l - 1],
                        GM_MLP_HIDDEN, s);
            dcur = n->dz[l - 1];
            ldcur = GM_MLP_HIDDEN;
        }
    }
}

static gm_status loss_and_grads(gm_mlp *n, const float *obs, const double *models, int64_t b,
                                float *loss_host, hipStream_t s)
{
    if (!n || !obs || !models || b <= 0 || (b % 64)) return GM_ERR_INVALID;
    gm_status st = ensure_batch(n, b);
    if (st != GM_OK) return st;
    gm_mlp_prep_input(obs, b, n->nl, n->kin, n->x, s);
    gm_mlp_prep_target(models, b, n->t, s);
    forward(n, b, s);
    if (hipMemsetAsync(n->loss, 0, sizeof(float), s) != hipSuccess) return GM_ERR_HIP;
    gm_mlp_loss(n->y, n->t, b, n->dy, n->loss, s);
    backward(n, b, s);
    if (loss_host) {
        if (hipMemcpyAsync(loss_host, n->loss, sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
            return GM_ERR_HIP;
        if (hipStreamSynchronize(s) != hipSuccess) return GM_ERR_HIP;
    }
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_mlp_train_step(gm_mlp *n, const float *obs, const double *models,
                                       int64_t b, float lr, float wd, float *loss, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    gm_status st = loss_and_grads(n, obs, models, b, loss, s);
    if (st != GM_OK) return st;
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)n->step), bc2 = 1.f - std::pow(b2, (float)n->step);
    gm_mlp_adamw(n->w, n->g, n->m, n->v, n->wb, n->nparam, lr, b1, b2, 1e-8f, wd, bc1, bc2, s);
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_mlp_gradients(gm_mlp *n, const float *obs, const double *models, int64_t b,
                                      float *grads, float *loss, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    gm_status st = loss_and_grads(n, obs, models, b, loss, s);
    if (st != GM_OK) return st;
    if (hipMemcpyAsync(grads, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
        return GM_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? GM_OK : GM_ERR_HIP;
}

extern "C" gm_status gm_mlp_predict(gm_mlp *n, const float *obs, int64_t npix, const double *base,
                                    double *models, void *stream)
{
    if (!n || !obs || !base || !models || npix < 0) return GM_ERR_INVALID;
    if (npix == 0) return GM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(n->base, base, sizeof(double) * GM_NMODEL, hipMemcpyHostToDevice, s) != hipSuccess)
        return GM_ERR_HIP;
    if (n->kin <= GM_MLP_HIDDEN && ((uintptr_t)obs & 15) == 0) {
        /* one fused kernel: no batch buffers, activations stay on chip */
        gm_mlp_offsets of{};
        for (int l = 0; l < GM_MLP_LAYERS; ++l) {
            gm_mlp_transpose(n->wb + n->L[l].woff, n->L[l].in, n->L[l].out, n->wt + n->L[l].woff, s);
            of.woff[l] = n->L[l].woff;
            of.boff[l] = n->L[l].boff;
        }
        gm_mlp_infer_fused(obs, npix, n->nl, n->kin, n->wt, n->w, of, n->base, models, s);
        return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
    }
    const int64_t chunk = 1 << 18;
    gm_status st = ensure_batch(n, npix < chunk ? npix : chunk);
    if (st != GM_OK) return st;
    for (int64_t p0 = 0; p0 < npix; p0 += chunk) {
        const int64_t rows = (npix - p0 < chunk) ? npix - p0 : chunk;
        const int64_t padded = ((rows + 127) / 128) * 128;
        gm_mlp_prep_input(obs + p0 * 4 * n->nl, rows, n->nl, n->kin, n->x, s);
        forward(n, padded, s, false);
        gm_mlp_decode(n->y, rows, n->base, models + p0 * GM_NMODEL, s);
    }
    return hipGetLastError() == hipSuccess ? GM_OK : GM_ERR_HIP;
}
