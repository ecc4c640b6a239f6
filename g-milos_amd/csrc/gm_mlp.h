/*
 * gm_mlp.h -- launchers of the learned-guess network kernels (gm_mlp.hip),
 * used by the network's C ABI (gm_mlp_host.cpp).
 *
 * Output map (standardised over the training prior):
 *   0 (B-1500)/866        1 cos(gamma)          2 2 sin(gamma) - 1
 *   3 cos(2 phi)          4 sin(2 phi)          5 vlos/1.732
 *   6 (dopp-0.04)/0.01155 7 (a-0.255)/0.1414    8 (eta0-10.5)/5.485
 *   9 (S0/Ic-0.25)/0.1    10 (Ic-1)/0.2  with Ic = S0 + S1
 */
#ifndef GM_MLP_H
#define GM_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GM_MLP_NOUT   64     /* padded output width (11 used)   */
#define GM_MLP_HIDDEN 256
#define GM_MLP_LAYERS 4

enum { GM_EPI_BIAS_GELU = 0, GM_EPI_BIAS = 1, GM_EPI_GELU_GRAD = 2, GM_EPI_STORE = 3 };

void gm_mlp_gemm(int ta, int tb, int epi, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                 int M, int N, int K, const float *bias, const float *zprev, int ldz, float *Cf,
                 uint16_t *Cb, int ldc, hipStream_t s);
void gm_mlp_prep_input(const float *obs, int64_t npix, int nl, int kin, uint16_t *x, hipStream_t s);
void gm_mlp_prep_target(const double *models, int64_t npix, float *t, hipStream_t s);
void gm_mlp_decode(const float *y, int64_t npix, const double *base, double *models, hipStream_t s);
void gm_mlp_loss(const float *y, const float *t, int64_t b, uint16_t *dy, float *loss, hipStream_t s);
void gm_mlp_bias_grad(const uint16_t *dy, int64_t b, int n, float *db, hipStream_t s);
void gm_mlp_adamw(float *w, const float *g, float *m, float *v, uint16_t *wb, int64_t n, float lr,
                  float b1, float b2, float eps, float wd, float bc1, float bc2, hipStream_t s);
void gm_mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);

/* Fused inference (kin <= GM_MLP_HIDDEN, kin % 64 == 0, obs 16-byte
 * aligned): obs [npix][4][nl]
 * -> start models [npix][11] in one kernel.  wt holds every layer's BF16
 * weights transposed to [out][in] at the same offsets as the [in][out]
 * copy; w the FP32 parameters (biases read from it). */
struct gm_mlp_offsets {
    int64_t woff[GM_MLP_LAYERS], boff[GM_MLP_LAYERS];
};
void gm_mlp_infer_fused(const float *obs, int64_t npix, int nl, int kin, const uint16_t *wt,
                        const float *w, const gm_mlp_offsets &of, const double *base, double *models,
                        hipStream_t s);
void gm_mlp_transpose(const uint16_t *w, int in, int out, uint16_t *wt, hipStream_t s);

#endif
