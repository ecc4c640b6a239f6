/*
 * gm_kernels.hip -- gfx950 kernels of the per-pixel Levenberg-Marquardt path.
 *
 * Two lane mappings:
 *  - wavelength-parallel ("lambda lanes"): 32 or 64 lanes own one pixel, each
 *    lane one or more wavelengths.  Used for synthesis, response functions,
 *    PSF correlation and normal equations, whose per-wavelength state is large
 *    and whose cross-wavelength coupling (PSF, sums) goes through LDS.
 *  - pixel-parallel ("pixel lanes"): one lane owns one pixel.  Used for the
 *    step solve, the accept/reject decision, the classical estimates and state
 *    bookkeeping, which are serial per pixel.
 * LM state lives in HBM as structure-of-arrays between kernels; active pixels
 * are carried in compacted lists so retired pixels cost nothing.
 *
 * Reference routines restated here: mil_sinrf.cu:27-433 (synthesis),
 * me_der.cu:35-376 (response functions), convolution.cu:153-241 (PSF),
 * lib.cu:65-155 (normal equations, chi^2), milosUtils.cu:30-238 (step
 * application and clamps), milosUtils.cu:280-365 (pseudo-inverse solve),
 * milosUtils.cu:386-563 (classical estimates), milosUtils.cu:574-710 (LM).
 */
#include "gm_internal.h"
#include <type_traits>

namespace {

constexpr float kVLIGHT = 2.99792458e+5f;   /* defines.h:80 */
constexpr float kCTE = 4.6686411e-13f;      /* defines.h:82 */
constexpr double kPI = 3.14159265358979323846264338327950288;

/* Minimum waves per SIMD the RF kernel is compiled for.  At 2 the 30-lambda
 * kernel still fits 3 (161 VGPRs) but is scheduled more freely: per 1M-pixel
 * inversion 30.1 -> 28.7 ms against 3; the 2- and 4-wavelength-per-lane
 * kernels drop their scratch spills. */
/* The same for two wavelengths per lane (nlambda 65-128, the two-line
 * set-up's 112): 184 VGPRs at 2; at 3 (168 VGPRs, 52 B of spills) the
 * two-line inversion's k_rf takes 32.0 -> 34.0 ms per 262 144 pixels. */
#ifndef GM_RF_WAVES_2L
#define GM_RF_WAVES_2L 2
#endif
#ifndef GM_RF_WAVES
#define GM_RF_WAVES 2
#endif

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

/* The first Zeeman sub-component assigns the per-wavelength sums instead of
 * adding to zeros (response kernel / synthesis kernel). */
#ifndef GM_PEEL_RF
#define GM_PEEL_RF 1
#endif
#ifndef GM_PEEL_SYNTH
#define GM_PEEL_SYNTH 1
#endif

/* ------------------------------------------------------------------------ */
/* Voigt and Faraday-Voigt functions, FP32: Hui, Armstrong & Wray (1978)
 * rational approximation w(z) ~ P(Z)/Q(Z), Z = a - i u.  The real-coefficient
 * polynomials are evaluated at the complex point with Knuth's second-order
 * recurrence (TAOCP vol. 2, 4.6.4): with r = 2 Re Z and s = |Z|^2,
 *   A <- r A + B,  B <- c_j - s A_old,   p(Z) = Z A + B,
 * two FMAs per coefficient instead of a complex multiply-add; then a scaled
 * complex division (no |Q|^2 overflow for large |u|).                      */
constexpr double kHuiP[7] = {122.607931777104326, 214.382388694706425, 181.928533092181549,
                             93.155580458138441,  30.180142196210589,  5.912626209773153,
                             0.564189583562615};
constexpr double kHuiQ[8] = {122.60793177387535, 352.730625110963558, 457.334478783897737,
                             348.703917719495792, 170.354001821091472, 53.992906912940207,
                             10.479857114260399, 1.0};
constexpr double kRSPI = 0.56418958354775628695;   /* 1/sqrt(pi) */

template <int N>
__device__ __forceinline__ void poly_knuth(const double (&c)[N + 1], float x, float y, float r,
                                           float s, float &re, float &im)
{
    float A = (float)c[N], B = (float)c[N - 1];
#pragma unroll
    for (int j = N - 2; j >= 0; --j) {
        const float t = A;
        A = fmaf(r, A, B);
        B = fmaf(-s, t, (float)c[j]);
    }
    re = fmaf(x, A, B);
    im = y * A;
}

/* R(Z) = Q(Z)/sqrt(pi) - Z P(Z), coefficients formed in double. */
struct HuiR {
    double c[8];
    constexpr HuiR() : c{}
    {
        c[0] = kHuiQ[0] * kRSPI;
        for (int k = 1; k < 8; ++k) c[k] = kHuiQ[k] * kRSPI - kHuiP[k - 1];
    }
};
constexpr HuiR kHuiR{};

__device__ __forceinline__ void voigt32(float a, float u, float &H, float &F)
{
    const float x = a, y = -u;
    const float r = 2.f * a, s2 = fmaf(a, a, u * u);
    float pr, pi, qr, qi;
    poly_knuth<6>(kHuiP, x, y, r, s2, pr, pi);
    poly_knuth<7>(kHuiQ, x, y, r, s2, qr, qi);
    const float s = frcp(fmaxf(fabsf(qr), fabsf(qi)));
    const float qrs = qr * s, qis = qi * s;
    const float inv = frcp(qrs * qrs + qis * qis) * s;
    H = (pr * qrs + pi * qis) * inv;
    F = 0.5f * (pi * qrs - pr * qis) * inv;
}

/* Voigt, Faraday-Voigt and their u-derivatives.  The derivatives are the
 * identities dH/du = 4aF - 2uH, dF/du = 1/sqrt(pi) - aH - 2uF
 * (me_der.cu) applied to the rational W = P/Q, i.e. 2 Im(Z W) and
 * 1/sqrt(pi) - Re(Z W).  Since Z P/Q = 1/sqrt(pi) - R/Q with
 * R(Z) = Q(Z)/sqrt(pi) - Z P(Z), they equal -2 Im(R/Q) and Re(R/Q) exactly;
 * evaluating R avoids the cancellation of the two terms in the line wings,
 * where |u|^2 times the FP32 rounding of H and F would otherwise dominate.
 * (The reference's RR constant differs from 1/sqrt(pi) by 5e-11, below FP32
 * resolution of every use.)                                                */
__device__ __forceinline__ void voigt32_d(float a, float u, float &H, float &F, float &dH,
                                          float &dF)
{
    const float x = a, y = -u;
    const float r = 2.f * a, s2 = fmaf(a, a, u * u);
    float pr, pi, qr, qi, rr, ri;
    poly_knuth<6>(kHuiP, x, y, r, s2, pr, pi);
    poly_knuth<7>(kHuiQ, x, y, r, s2, qr, qi);
    poly_knuth<7>(kHuiR.c, x, y, r, s2, rr, ri);
    const float s = frcp(fmaxf(fabsf(qr), fabsf(qi)));
    const float qrs = qr * s, qis = qi * s;
    const float inv = frcp(qrs * qrs + qis * qis) * s;
    H = (pr * qrs + pi * qis) * inv;
    F = 0.5f * (pi * qrs - pr * qis) * inv;
    dF = (rr * qrs + ri * qis) * inv;
    dH = -2.f * (ri * qrs - rr * qis) * inv;
}

/* Model-dependent constants shared by all wavelengths of a pixel
 * (mil_sinrf.cu:38-60).  gamma/phi derivatives are per degree.              */
struct PixPre {
    float B, aa, inv_dopp, vlos, S0, S1, E0, E0_2;
    float sinis, cosi, cosis_2, sina, cosa, sinis_cosa, sinis_sina;
    float sinda, cosda, sindi, cosdi;
};

static_assert(sizeof(PixPre) == GM_NPRE * sizeof(float), "PixPre layout");

/* PixPre records of the LM state, [pix][GM_PRE_REC] (80 bytes, 16-byte
 * aligned): a pixel's constants are one contiguous read however sparse the
 * active pixels are. */
__device__ __forceinline__ void store_pre(float *__restrict__ pre, int64_t pix, const PixPre &p)
{
    const float *v = reinterpret_cast<const float *>(&p);
    float4 *r = reinterpret_cast<float4 *>(pre + pix * GM_PRE_REC);
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    r[4] = make_float4(v[16], v[17], v[18], 0.f);
}
__device__ __forceinline__ PixPre load_pre(const float *__restrict__ pre, int64_t pix)
{
    PixPre p;
    float *v = reinterpret_cast<float *>(&p);
    const float4 *r = reinterpret_cast<const float4 *>(pre + pix * GM_PRE_REC);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const float4 x = r[q];
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        if (q < 4) v[4 * q + 3] = x.w;
    }
    return p;
}

__device__ __forceinline__ PixPre make_pre(const double *m)
{
    PixPre p;
    const float CC = (float)(kPI / 180.0);
    float sin_gm, cosi, sina, cosa;
    sincosf((float)(m[5] * (kPI / 180.0)), &sin_gm, &cosi);
    sincosf((float)(2.0 * m[6] * (kPI / 180.0)), &sina, &cosa);
    p.B = (float)m[1];
    p.aa = (float)m[4];
    p.inv_dopp = (float)(1.0 / m[3]);
    p.vlos = (float)m[2];
    p.S0 = (float)m[7];
    p.S1 = (float)m[8];
    p.E0 = (float)m[0];
    p.E0_2 = p.E0 * 0.5f;
    p.sinis = sin_gm * sin_gm;
    p.cosi = cosi;
    p.cosis_2 = (1.f + cosi * cosi) * 0.5f;
    p.sina = sina;
    p.cosa = cosa;
    p.sinis_cosa = p.sinis * cosa;
    p.sinis_sina = p.sinis * sina;
    p.sinda = cosa * (2.f * CC);
    p.cosda = -sina * (2.f * CC);
    p.sindi = cosi * sin_gm * (2.f * CC);
    p.cosdi = -sin_gm * CC;
    return p;
}

/* Unconvolved synthesis (and response functions) at one wavelength k.
 * out[s*9 + j], s = Stokes, j = 0 synth, 1..6 d/d{eta0,B,vlos,dopp,a,gamma},
 * 7 d/dphi, 8 d/dS1.  Follows mil_sinrf.cu:144-433 and me_der.cu:35-376; the
 * Voigt derivatives are summed over the sub-components of each type.        */
template <bool RF>
__device__ __forceinline__ void eval_lambda(const gm_dsetup *__restrict__ S,
                                            const PixPre &pp, int k, float *out, float dl0)
{
    /* profiles and their derivatives summed over lines and sub-components
     * per component type t = 0 pi, 1 sigma_b, 2 sigma_r; the first
     * sub-component of the first line assigns the sums */
    float fi[3], shi[3];
    float dfi[3][4], dshi[3][4];
    const float shs = S->shi_scale;
    constexpr bool PEEL = RF ? GM_PEEL_RF : GM_PEEL_SYNTH;
    auto component = [&](int L, int t, int i, float ulos_ub, float parcial, float auxCte, float fo,
                         bool first) {
        const float nu = parcial * S->comp_nu[L][t][i];
        const float w = S->comp_w[L][t][i] * fo;
        const float u = ulos_ub - nu * pp.B;
        float H, F, dHv = 0.f, dFv = 0.f;
        if (RF)
            voigt32_d(pp.aa, u, H, F, dHv, dFv);
        else
            voigt32(pp.aa, u, H, F);
        fi[t] = first ? w * H : fmaf(w, H, fi[t]);
        shi[t] = first ? (2.f * w) * F : fmaf(2.f * w, F, shi[t]);
        if (RF) {
            const float dH_u = dHv * w;
            const float dF_u = dFv * w;
            const float uu = -u * pp.inv_dopp;
            const float sdF = shs * dF_u;
            /* d/d{B, vlos, dopp, a} = (factor) x (per-u derivative) */
            const float a0[4] = {dH_u, dH_u, dH_u, dF_u};
            const float f0[4] = {-nu, auxCte, uu, -2.f};
            const float a1[4] = {sdF, sdF, sdF, shs * dH_u};
            const float f1[4] = {-nu, auxCte, uu, 0.5f};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dfi[t][j] = first ? a0[j] * f0[j] : fmaf(a0[j], f0[j], dfi[t][j]);
                dshi[t][j] = first ? a1[j] * f1[j] : fmaf(a1[j], f1[j], dshi[t][j]);
            }
        }
    };
    /* one line; the first is peeled with its lambda - lambda0 from the
     * caller (dl0 = S->dl[0][k], read once per kernel), so that a
     * single-line set-up issues no vector load here: a load waited for
     * in the middle of the pixel would also wait for every memory
     * operation issued before it */
    auto line = [&](int L, float dlk) {
        const float l0 = S->l0[L], fo = S->fo[L];
        const float ulos = pp.vlos * l0 * pp.inv_dopp * (1.f / kVLIGHT);
        const float parcial = l0 * l0 * pp.inv_dopp * kCTE;
        const float auxCte = -l0 * pp.inv_dopp * (1.f / kVLIGHT);
        const float ub = dlk * pp.inv_dopp - ulos;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int nc = S->ncomp[L][t];
            int i0 = 0;
            if (L == 0 && !PEEL) {
                fi[t] = 0.f;
                shi[t] = 0.f;
                if (RF)
#pragma unroll
                    for (int j = 0; j < 4; ++j) { dfi[t][j] = 0.f; dshi[t][j] = 0.f; }
            } else if (L == 0) {
                if (nc > 0) {
                    component(0, t, 0, ub, parcial, auxCte, fo, true);
                    i0 = 1;
                } else {
                    fi[t] = 0.f;
                    shi[t] = 0.f;
#pragma unroll
                    for (int j = 0; j < 4; ++j) { dfi[t][j] = 0.f; dshi[t][j] = 0.f; }
                }
            }
            for (int i = i0; i < nc; ++i) component(L, t, i, ub, parcial, auxCte, fo, false);
        }
    };
    line(0, dl0);
    const int nlines = S->nlines;
    for (int L = 1; L < nlines; ++L) line(L, S->dl[L][k]);
    const float E0_2 = pp.E0_2;
    const float parcial1 = fi[1] + fi[2];
    const float parcial2 = E0_2 * (fi[0] - 0.5f * parcial1);
    const float parcial3 = E0_2 * (shi[0] - 0.5f * (shi[1] + shi[2]));
    const float etain = E0_2 * (fi[0] * pp.sinis + parcial1 * pp.cosis_2);
    const float etaq = parcial2 * pp.sinis_cosa;
    const float etau = parcial2 * pp.sinis_sina;
    const float etav = (fi[2] - fi[1]) * (E0_2 * pp.cosi);
    const float rhoq = parcial3 * pp.sinis_cosa;
    const float rhou = parcial3 * pp.sinis_sina;
    const float rhov = (shi[2] - shi[1]) * (E0_2 * pp.cosi);
    const float etai = 1.f + etain;
    const float etai_2 = etai * etai;
    const float auxq = rhoq * rhoq, auxu = rhou * rhou, auxv = rhov * rhov;
    const float gp1 = etai_2 - etaq * etaq - etau * etau - etav * etav + auxq + auxu + auxv;
    const float gp3 = etai_2 + auxq + auxu + auxv;
    const float gp2 = etaq * rhoq + etau * rhou + etav * rhov;
    const float dt = etai_2 * gp1 - gp2 * gp2;
    const float dti = frcp(dt);
    const float gp4 = etai_2 * etaq + etai * (etav * rhou - etau * rhov);
    const float gp5 = etai_2 * etau + etai * (etaq * rhov - etav * rhoq);
    const float gp6 = etai_2 * etav + etai * (etau * rhoq - etaq * rhou);
    const float g4 = gp4 + rhoq * gp2;
    const float g5 = gp5 + rhou * gp2;
    const float g6 = gp6 + rhov * gp2;
    const float ah = S->mu;
    const float dtiaux = dti * (-(pp.S1 * ah));
    out[0] = pp.S0 - dtiaux * etai * gp3;
    out[9] = dtiaux * g4;
    out[18] = dtiaux * g5;
    out[27] = dtiaux * g6;
    if (!RF) return;

    /* d(eta, rho)/dp: me_der.cu:254-312 */
    float d_ei[7], d_eq[7], d_eu[7], d_ev[7], d_rq[7], d_ru[7], d_rv[7];
    const float E0 = pp.E0;
    const float iE0 = frcp(E0);
    d_ei[0] = etain * iE0; d_eq[0] = etaq * iE0; d_eu[0] = etau * iE0;
    d_ev[0] = etav * iE0; d_rq[0] = rhoq * iE0; d_ru[0] = rhou * iE0;
    d_rv[0] = rhov * iE0;
    /* for B, vlos, dopp and a (j = 1..4) the seven derivatives are fixed
     * combinations of the six per-type profile derivatives dfi[t], dshi[t]
     * (me_der.cu:254-312):
     *   d eta_I = A dfi0 + B (dfi1 + dfi2),  d eta_Q,U = (sc, ss) (dfi0 -
     *   (dfi1 + dfi2) / 2),  d eta_V = E (dfi2 - dfi1), and the same for
     *   rho_Q,U,V with dshi; they enter the chain rule below folded into its
     *   coefficients */
    const float sinis_E0_2 = pp.sinis * E0_2, cosis_2_E0_2 = pp.cosis_2 * E0_2;
    const float sc = pp.sinis_cosa * E0_2, ss = pp.sinis_sina * E0_2;
    const float cosi_2_E0 = pp.cosi * E0_2;
    const float cosi_cosdi = pp.cosi * pp.cosdi * E0_2, sindi_E0_2 = pp.sindi * E0_2;
    const float cosdi_E0_2 = E0_2 * pp.cosdi;
    d_ei[5] = fi[0] * sindi_E0_2 + parcial1 * cosi_cosdi;
    d_eq[5] = parcial2 * (pp.sindi * pp.cosa);
    d_eu[5] = parcial2 * (pp.sindi * pp.sina);
    d_ev[5] = (fi[2] - fi[1]) * cosdi_E0_2;
    d_rq[5] = parcial3 * (pp.sindi * pp.cosa);
    d_ru[5] = parcial3 * (pp.sindi * pp.sina);
    d_rv[5] = (shi[2] - shi[1]) * cosdi_E0_2;
    d_eq[6] = parcial2 * (pp.sinis * pp.cosda);
    d_eu[6] = parcial2 * (pp.sinis * pp.sinda);
    d_rq[6] = parcial3 * (pp.sinis * pp.cosda);
    d_ru[6] = parcial3 * (pp.sinis * pp.sinda);

    /* chain rule through the ME solution (me_der.cu:317-373).  Every
     * response is linear in its parameter's seven absorption-matrix
     * derivatives (d eta_I,Q,U,V, d rho_Q,U,V), with coefficients that
     * depend only on the wavelength: out[s][p] = sum_m C[s][m] d_m[p].  The
     * coefficients are formed once per wavelength from the reference's
     * expressions for dgp1..dgp6 and d(dt), then applied to the seven
     * parameters.                                                          */
    const float dtaux = (-(pp.S1 * ah)) * dti * dti;
    const float etai_gp3 = etai * gp3;
    const float a2 = 2.f * etai;
    const float ext1 = a2 * etaq + etav * rhou - etau * rhov;
    const float ext2 = a2 * etau + etaq * rhov - etav * rhoq;
    const float ext3 = a2 * etav + etau * rhoq - etaq * rhou;
    /* d(dt) = sum_m cdt[m] d_m, from ext4 ei + etai^2 dgp1 - 2 gp2 dgp2 */
    const float t2e = 2.f * etai_2, t2g = 2.f * gp2;
    float cdt[7];
    cdt[0] = a2 * gp1 + t2e * etai;
    cdt[1] = -(t2e * etaq + t2g * rhoq);
    cdt[2] = -(t2e * etau + t2g * rhou);
    cdt[3] = -(t2e * etav + t2g * rhov);
    cdt[4] = t2e * rhoq - t2g * etaq;
    cdt[5] = t2e * rhou - t2g * etau;
    cdt[6] = t2e * rhov - t2g * etav;
    /* the Delta-multiplied numerators' coefficients, per Stokes parameter */
    float ca[4][7];
    ca[0][0] = gp3 + a2 * etai;  ca[0][1] = 0.f;  ca[0][2] = 0.f;  ca[0][3] = 0.f;
    ca[0][4] = a2 * rhoq;        ca[0][5] = a2 * rhou;  ca[0][6] = a2 * rhov;
    ca[1][0] = ext1;
    ca[1][1] = etai_2 + rhoq * rhoq;
    ca[1][2] = rhoq * rhou - etai * rhov;
    ca[1][3] = rhoq * rhov + etai * rhou;
    ca[1][4] = gp2 + rhoq * etaq;
    ca[1][5] = rhoq * etau + etai * etav;
    ca[1][6] = rhoq * etav - etai * etau;
    ca[2][0] = ext2;
    ca[2][1] = rhou * rhoq + etai * rhov;
    ca[2][2] = etai_2 + rhou * rhou;
    ca[2][3] = rhou * rhov - etai * rhoq;
    ca[2][4] = rhou * etaq - etai * etav;
    ca[2][5] = gp2 + rhou * etau;
    ca[2][6] = rhou * etav + etai * etaq;
    ca[3][0] = ext3;
    ca[3][1] = rhov * rhoq - etai * rhou;
    ca[3][2] = rhov * rhou + etai * rhoq;
    ca[3][3] = etai_2 + rhov * rhov;
    ca[3][4] = rhov * etaq + etai * etau;
    ca[3][5] = rhov * etau - etai * etaq;
    ca[3][6] = gp2 + rhov * etav;
    /* out = dtaux (Delta a - g d(dt)); I carries the opposite sign */
    const float gq[4] = {etai_gp3, g4, g5, g6};
    const float sg[4] = {-1.f, 1.f, 1.f, 1.f};
    float C[4][7];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const float u = sg[st] * dt * dtaux, v = sg[st] * gq[st] * dtaux;
#pragma unroll
        for (int m = 0; m < 7; ++m)
            C[st][m] = (st == 0 && m >= 1 && m <= 3) ? -(v * cdt[m])     /* ca[0][1..3] = 0 */
                                                     : u * ca[st][m] - v * cdt[m];
    }
    /* eta0 and gamma: all seven absorption-matrix derivatives */
#pragma unroll
    for (int p = 0; p < 6; p += 5) {
        const float dm[7] = {d_ei[p], d_eq[p], d_eu[p], d_ev[p], d_rq[p], d_ru[p], d_rv[p]};
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            float o = C[st][0] * dm[0];
#pragma unroll
            for (int m = 1; m < 7; ++m) o = fmaf(C[st][m], dm[m], o);
            out[st * 9 + 1 + p] = o;
        }
    }
    /* B, vlos, dopp, a: out = sum_t D[t] dfi[t] + D[3 + t] dshi[t], with the
     * combinations above folded into D once per wavelength (six products
     * per response instead of seven after forming the seven derivatives) */
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const float q1 = fmaf(C[st][1], sc, C[st][2] * ss);   /* eta_Q,U */
        const float q2 = fmaf(C[st][4], sc, C[st][5] * ss);   /* rho_Q,U */
        const float h1 = fmaf(C[st][0], cosis_2_E0_2, -0.5f * q1);
        const float h2 = -0.5f * q2;
        const float e3 = C[st][3] * cosi_2_E0, e6 = C[st][6] * cosi_2_E0;
        const float D[6] = {fmaf(C[st][0], sinis_E0_2, q1), h1 - e3, h1 + e3, q2, h2 - e6, h2 + e6};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float o = D[0] * dfi[0][j];
            o = fmaf(D[1], dfi[1][j], o);
            o = fmaf(D[2], dfi[2][j], o);
            o = fmaf(D[3], dshi[0][j], o);
            o = fmaf(D[4], dshi[1][j], o);
            o = fmaf(D[5], dshi[2][j], o);
            out[st * 9 + 2 + j] = o;
        }
    }
    /* phi changes eta_Q,U and rho_Q,U only, and d/dphi of I and V is zero
     * (me_der.cu:372-373) */
#pragma unroll
    for (int st = 1; st < 3; ++st)
        out[st * 9 + 7] = fmaf(C[st][1], d_eq[6],
                               fmaf(C[st][2], d_eu[6], fmaf(C[st][4], d_rq[6], C[st][5] * d_ru[6])));
    const float dti2 = -(dti * ah);
    out[0 + 8] = -dti2 * etai_gp3;
    out[9 + 8] = dti2 * g4;
    out[18 + 8] = dti2 * g5;
    out[27 + 8] = dti2 * g6;
    out[0 + 7] = 0.f;   /* d I / d phi = 0 (me_der.cu:372) */
    out[27 + 7] = 0.f;  /* d V / d phi = 0 (me_der.cu:373) */
}

__device__ __forceinline__ void load_model(const double *__restrict__ models,
                                           int64_t fstride, int64_t pstride,
                                           int64_t pix, double *m)
{
#pragma unroll
    for (int f = 0; f < GM_NMODEL; ++f) m[f] = models[f * fstride + pix * pstride];
}

/* Global -> LDS copy of 16 bytes per active lane (global_load_lds_dwordx4):
 * lane L's bytes land at LDS byte address base + 16 L (M0 holds the base;
 * checked on the device: lanes 0-4 and 32-36 of a wave fill slots 0-4 and
 * 32-36).  Written as inline assembly: issued through the builtin, the
 * compiler makes every later LDS access wait for the copy, which would
 * stall the staging writes of the pixel being computed.  The copy has no
 * register destination; the reader waits for it with lds_copy_wait(). */
__device__ __forceinline__ void lds_copy16(const float *src, const float4 *dst_lds)
{
    const uint32_t base = (uint32_t)reinterpret_cast<uintptr_t>(
        (__attribute__((address_space(3))) const void *)dst_lds);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                 :
                 : "s"(base), "v"(src)
                 : "memory", "m0");
}
/* s_waitcnt vmcnt(0): every vector memory operation of the wave, the LDS
 * copies included, has completed */
__device__ __forceinline__ void lds_copy_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

/* Pixel pairs of a grid-stride loop (PPW = 64 / LPP pixels per wave).  List
 * entries are loaded ahead and kept 32 bits wide: widened to 64 bits at
 * once, a load would be waited for right there.  With REC (the LM loop's
 * per-pixel records) the constants record of the next pair is also copied
 * global -> LDS while this pair is computed (lanes 0-4 of each pixel, one
 * float4 each, to slot `lane` of the other stage of `buf`), so its memory
 * latency is off the critical path; that needs the list two pairs ahead. */
template <int LPP, bool REC>
struct PairFeed {
    static constexpr int PPW = 64 / LPP;
    static constexpr int PFS = (PPW - 1) * LPP + GM_PRE_REC / 4;   /* LDS slots per stage */
    const int32_t *list;
    const float *pre;
    float4 *buf;
    int64_t cnt, stride;
    int h, k0;
    bool has_list;
    int32_t next = 0, after = 0;
    int stage = 0;

    __device__ int32_t fetch(int64_t b) const { return list[(b + h < cnt) ? b + h : b]; }
    __device__ void copy(int32_t e, int st) const
    {
        if (k0 < GM_PRE_REC / 4) lds_copy16(pre + (int64_t)e * GM_PRE_REC + 4 * k0, buf + st * PFS);
    }
    __device__ void start(int64_t first)
    {
        if (has_list && first < cnt) next = fetch(first);
        if constexpr (REC) {
            if (first + stride < cnt) after = fetch(first + stride);
            if (first < cnt) copy(next, 0);
        }
    }
    /* this pair's pixel: the list entry, or the item itself */
    __device__ int64_t pixel(int64_t base) const
    {
        return has_list ? (int64_t)next : ((base + h < cnt) ? base + h : base);
    }
    /* REC: this pair's constants record; then the loads for the next pairs.
     * The copy into this stage was issued an iteration ago; it is waited
     * for before anything else is issued (the wait covers every outstanding
     * vector memory operation). */
    __device__ PixPre advance_rec(int64_t base)
    {
        lds_copy_wait();
        PixPre pp;
        float *v = reinterpret_cast<float *>(&pp);
        const float4 *r = buf + stage * PFS + h * LPP;
#pragma unroll
        for (int q = 0; q < GM_PRE_REC / 4; ++q) {
            const float4 x = r[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            if (q < GM_PRE_REC / 4 - 1) v[4 * q + 3] = x.w;
        }
        const int32_t e2 = (base + 2 * stride < cnt) ? fetch(base + 2 * stride) : 0;
        if (base + stride < cnt) copy(after, stage ^ 1);
        next = after;
        after = e2;
        stage ^= 1;
        return pp;
    }
    /* without records: the next list entry */
    __device__ void advance(int64_t base)
    {
        if (has_list && base + stride < cnt) next = fetch(base + stride);
    }
};

/* Reference continuum rule for the I-variant of the PSF: the larger end
 * sample (mil_sinrf.cu:120-123, me_der.cu:171-174).                         */
__device__ __forceinline__ float end_max(float x0, float xn) { return (x0 > xn) ? x0 : xn; }

/* Folds two per-lane partial sums across the lane halves of an H-lane
 * pair of groups (H = 16: lanes L and L ^ 16; H = 32: L and L ^ 32): the
 * lower lane of each pair gets a's total, the upper one b's.  One
 * v_permlane16/32_swap (a VALU instruction that trades the upper group of
 * its first operand with the lower group of its second) and one add per
 * two sums, instead of a ds_bpermute per sum through the LDS pipeline,
 * which the PSF and the reduction already keep busy. */
template <int H>
__device__ __forceinline__ float fold_pair(float a, float b)
{
    static_assert(H == 16 || H == 32, "fold width");
    unsigned lo, hi;
    if constexpr (H == 16) {
        const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b),
                                                         false, false);
        lo = sw[0];
        hi = sw[1];
    } else {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b),
                                                         false, false);
        lo = sw[0];
        hi = sw[1];
    }
    return __uint_as_float(lo) + __uint_as_float(hi);
}

/* Sum over the W-lane group (W = 16, 32 or 64), valid in the group's first
 * lane (in every lane of its even 16-lane rows): four DPP steps give every
 * lane its 16-lane row's sum (partners l^8 by row rotation, l^7 by the
 * half-row mirror, l^2 and l^1 by quad permutations: together they span all
 * sixteen), then the rows are folded with lane swaps -- VALU instructions
 * only, where __shfl_xor would issue a ds_bpermute per step. */
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) { return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    return __hiloint2double((int)dpp_u32<CTRL>((uint32_t)__double2hiint(v)),
                            (int)dpp_u32<CTRL>((uint32_t)__double2loint(v)));
}
__device__ __forceinline__ float upper_row_partner(float v, bool half32)
{
    const auto sw = half32 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false)
                           : __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(sw[1]);
}
__device__ __forceinline__ double upper_row_partner(double v, bool half32)
{
    return __hiloint2double(__float_as_int(upper_row_partner(__int_as_float(__double2hiint(v)), half32)),
                            __float_as_int(upper_row_partner(__int_as_float(__double2loint(v)), half32)));
}
template <typename T, int W>
__device__ __forceinline__ T group_sum_first(T v)
{
    static_assert(W == 16 || W == 32 || W == 64, "group width");
    v += dpp_move<0x128>(v);            /* row_ror:8       -> l ^ 8       */
    v += dpp_move<0x141>(v);            /* row_half_mirror -> l ^ 7       */
    v += dpp_move<0x4E>(v);             /* quad_perm 2301  -> l ^ 2       */
    v += dpp_move<0xB1>(v);             /* quad_perm 1032  -> l ^ 1       */
    if constexpr (W >= 32) v += upper_row_partner(v, false);
    if constexpr (W == 64) v += upper_row_partner(v, true);
    return v;
}

template <typename T, int W>
__device__ __forceinline__ T group_sum(T v)
{
#pragma unroll
    for (int m = W / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

/* Macroturbulence kernel of one pixel on the wavelength grid
 * (fgauss.cu:16-63): exp(-x^2/2), x = (lambda - lambda[n/2]) / sigma,
 * sigma = lambda0 mac / c, normalised to unit sum (FP64 as the reference),
 * then correlated with all nlambda PSF taps when the PSF is on
 * (direct_convolution_double, mil_sinrf.cu:104-107).  Called by every lane
 * of the workgroup (two barriers); gk, gc are the pixel group's LDS buffers.
 * gc receives the combined kernel; returns its sum.                         */
template <int LPP, int NLPL, bool DER = false>
__device__ float mac_kernel(const gm_dsetup *__restrict__ S, double mac, float *gk, float *gc,
                            int k0, float *gdk = nullptr, float *gd = nullptr)
{
    const int n = S->nlambda;
    const double macv = mac > 1e-4 ? mac : 1e-4;
    const double ild = S->mac_l0 * macv / 2.99792458e5;
    const double lc = S->lambda[n / 2];
    double g[NLPL], x2[NLPL], part = 0.0;
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        double v = 0.0;
        x2[s] = 0.0;
        if (k < n) {
            const double x = (S->lambda[k] - lc) / ild;
            x2[s] = x * x;
            const double t = 0.5 * x2[s];
            v = (t < 1e30) ? exp(-t) : 0.0;
        }
        g[s] = v;
        part += v;
    }
    const double sum = group_sum<double, LPP>(part);
    if constexpr (DER) {
        /* d/dmac of the normalised kernel: G (x^2 - 1) / mac as
         * me_der.cu:124-131, or G (x^2 - sum G x^2) / mac exactly */
        double m2 = 1.0;
        if (S->shi_scale != 1.0f) {
            double pm = 0.0;
#pragma unroll
            for (int s = 0; s < NLPL; ++s) pm += g[s] / sum * x2[s];
            m2 = group_sum<double, LPP>(pm);
        }
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
            if (k < n) gdk[k] = (float)(g[s] / sum / macv * (x2[s] - m2));
        }
    }
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        if (k < n) gk[k] = (float)(g[s] / sum);
    }
    __syncthreads();
    const int c = n / 2;
    float tot = 0.f;
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        if (k < n) {
            float v = 0.f, vd = 0.f;
            if (S->use_psf) {
                /* the PSF's applied taps (offsets psf_dlo..psf_dhi: all of
                 * them with psf_tail_tol 0) */
                const int dlo = S->psf_dlo, dhi = S->psf_dhi;
                for (int d = dlo; d <= dhi; ++d) {
                    const int i = k + d;
                    if (i >= 0 && i < n) {
                        const float tp = S->psf_tap[d - dlo];
                        v += tp * gk[i];
                        if constexpr (DER) vd += tp * gdk[i];
                    }
                }
            } else {
                v = gk[k];
                if constexpr (DER) vd = gdk[k];
            }
            gc[k] = v;
            if constexpr (DER) gd[k] = vd;
            tot += v;
        }
    }
    __syncthreads();
    return group_sum<float, LPP>(tot);
}

/* Circular correlation row index of convolution.cu:90-118: output m takes
 * tap j from sample (m + n/2 - j) mod n.                                     */
__device__ __forceinline__ int circ_row(int m, int c, int j, int n)
{
    int r = m + c - j;
    r += (r < 0) ? n : 0;
    r -= (r >= n) ? n : 0;
    return r;
}

/* Tap window [lo, hi] of this lane group's combined mac kernel gc (n taps,
 * sum `total`) that holds every tap of at least tol * total / n: the taps
 * outside it, skipped by the circular correlation, carry at most tol of the
 * kernel's mass (tol = psf_tail_tol as for the PSF; 0 keeps every tap, the
 * reference's direct correlation).  Called by every lane of the wave. */
template <int LPP, int NLPL>
__device__ __forceinline__ void mac_window(const gm_dsetup *__restrict__ S, const float *gc, float total,
                                           int k0, int h, int &lo, int &hi)
{
    const int n = S->nlambda;
    lo = 0;
    hi = n - 1;
    const float tol = S->mac_tail_tol;
    if (!(tol > 0.f)) return;
    const float thr = tol * total / (float)n;
    int l = n, u = -1;
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        const bool sig = k < n && gc[k] >= thr;
        const uint64_t m = __ballot(sig);
        const uint64_t mh = (LPP == 64) ? m : ((m >> (32 * h)) & 0xffffffffull);
        if (mh) {
            l = min(l, s * LPP + __builtin_ctzll(mh));
            u = max(u, s * LPP + 63 - __builtin_clzll(mh));
        }
    }
    if (u >= 0) {
        lo = l;
        hi = u;
    }
}

/* The 30 of the 36 evaluated values (per Stokes: synth, d/d{eta0, B, vlos,
 * dopp, a, gamma, phi}, d/dS1) that k_rf convolves: all but d/dphi of I and
 * V (zero) and the four d/dS1. */
__constant__ constexpr int kStagedSlot[30] = {0,  1,  2,  3,  4,  5,  6,                 /* I */
                                              9,  10, 11, 12, 13, 14, 15, 16,          /* Q */
                                              18, 19, 20, 21, 22, 23, 24, 25,          /* U */
                                              27, 28, 29, 30, 31, 32, 33};             /* V */

/* RF slot (reference layout) -> staged slot j within a Stokes block; -1 is
 * the unconvolved S0 response, -2 the always-zero slot 9.                   */
__device__ __forceinline__ int staged_slot(int p)
{
    return (p <= 6) ? p + 1 : (p == 7 ? -1 : (p == 8 ? 8 : -2));
}

/* ------------------------------------------------------------------------ */
/* The instrument PSF as a matrix product on the matrix cores (up to 32
 * wavelengths, two pixels per wave).  The reference correlates every
 * staged vector with all nlambda taps centred at nlambda/2
 * (convolution.cu:153-241): z[k] = sum_m psf[m - k + n/2] x[m], i.e.
 * Z = X^T P^T per pixel with P[k][m] the banded Toeplitz matrix of the taps.
 * Applied densely over m = 0..31 (zero rows and zero table entries past the
 * grid), all taps cost one fixed product, 32 MFMAs per wave: cheaper than
 * the VALU correlation from about 16 taps on, dearer below (FP32 matrix and
 * FP32 VALU work do not overlap on a SIMD), so set-ups with fewer taps
 * keep the VALU correlation.
 *
 * Tap table (LDS, 64 floats, the same for every pixel): tab[u] is the tap
 * for m - k = u - 31, zero outside the kernel's support.                   */
__device__ __forceinline__ float psf_tab_entry(const gm_dsetup *__restrict__ S, int u)
{
    const int n = S->nlambda;
    const int j = u - 31 + n / 2;
    return (u < 63 && j >= 0 && j < n) ? S->psf_full[j] : 0.f;
}

typedef float gm_f32x32 __attribute__((ext_vector_type(32)));

/* 32 staged values per wavelength (rows of 36 floats, one per wavelength,
 * rows 0-31 of each pixel's region; xs = the lane's pixel region + k0),
 * two pixels per wave: v_mfma_f32_32x32x1_2b_f32 with block b = pixel b,
 * A_b[i] = X_b[m][i] (lane 32b + i), B_b[k] = P[k][m] (lane 32b + k), summed
 * over m.  D lane L, register r holds block r/16, output k = L%32, value
 * i = 8((r%16)/4) + 4(L/32) + r%4; one permlane32 swap per register pair
 * hands each half-wave its own pixel: afterwards registers r and 16 + r of
 * every lane hold values 8(r/4) + r%4 and 8(r/4) + 4 + r%4 of its pixel at
 * its own wavelength.                                                      */
__device__ __forceinline__ void psf_mfma32(const float *__restrict__ xs,
                                           const float *__restrict__ ts, float (&z)[32])
{
    gm_f32x32 acc = {};
#pragma unroll
    for (int m = 0; m < 32; ++m)
        acc = __builtin_amdgcn_mfma_f32_32x32x1f32(xs[m * 36], ts[m], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[r]),
                                                         __float_as_uint(acc[16 + r]), false,
                                                         false);
        z[8 * (r / 4) + (r % 4)] = __uint_as_float(sw[0]);
        z[8 * (r / 4) + 4 + (r % 4)] = __uint_as_float(sw[1]);
    }
}

/* Normal-equation entry e of the packed upper triangle (n x n, n = 9 or
 * 11) -> its parameter pair, for masking fixed parameters after the
 * reduction. */
struct PairTable {
    int8_t p[66], q[66];
    constexpr PairTable(int nn) : p{}, q{}
    {
        int e = 0;
        for (int a = 0; a < nn; ++a)
            for (int b = a; b < nn; ++b) { p[e] = (int8_t)a; q[e] = (int8_t)b; ++e; }
    }
};
__constant__ constexpr PairTable kPairs9{9};
__constant__ constexpr PairTable kPairs11{11};

/* Responses that vanish identically: d/dS0 of Q, U, V (me_der.cu:365-368)
 * and d/dphi of I and V (me_der.cu:372-373); their normal-equation
 * products are skipped at compile time. */
__host__ __device__ constexpr bool rf_nonzero(int st, int p)
{
    return !((st != 0 && p == 7) || ((st == 0 || st == 3) && p == 6));
}

/* One accept / reject decision (milosUtils.cu:660-698): stop tests on the
 * chi^2 change and value, lambda update (/5 on accept, x10 on reject, with
 * the logclambda factor), lambda bounds and the iteration cap.          */
template <typename T> struct LmStep {
    bool accept, cont;
    T flambda;
};
template <typename T>
__device__ __forceinline__ LmStep<T> lm_step(const gm_dsetup *__restrict__ S, T chisqr, T ochisqr,
                                             T flambda, T factor, int it)
{
    bool clanda = (fabs((ochisqr - chisqr) * T(100) / chisqr) < (T)S->toplim) ||
                  (chisqr < T(0.0001));
    const bool accept = chisqr - ochisqr < T(0);
    flambda = accept ? flambda / (T(5.0) * factor) : flambda * T(10.0) * factor;
    if ((flambda > T(1e+7)) || (flambda < T(1e-25))) clanda = true;
    LmStep<T> r;
    r.accept = accept;
    r.cont = (it + 1 < S->miter) && !clanda;
    r.flambda = flambda;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Synthesis + response functions + PSF + normal equations + chi^2 for a list
 * of pixels; wavelength-parallel mapping.                                   */
template <int LPP, int NLPL, bool MIXED, int XM, bool REC>
__global__ __launch_bounds__(64, NLPL == 2 ? GM_RF_WAVES_2L : GM_RF_WAVES) void k_rf(
    const gm_dsetup *__restrict__ S, const double *__restrict__ models,
    const float *__restrict__ pre, int64_t mfs, int64_t mps, const float *__restrict__ obs,
    const int32_t *__restrict__ list, const int32_t *__restrict__ count,
    int64_t n_direct, float *__restrict__ alpha, float *__restrict__ beta,
    int64_t aes, int64_t aps, int64_t bes, int64_t bps, double *__restrict__ chi_out,
    float *__restrict__ synth_out, float *__restrict__ rf_out, int rows, int padlo,
    int alpha_first, double *__restrict__ model_dst, float *__restrict__ pre_dst)
{
    /* XM: 0 the 9-parameter kernel, 1 with macroturbulence / stray light,
     * 2 the eleven-slot system (slot 9 mac, slot 10 filling factor) */
    constexpr bool EXT = XM >= 1;
    constexpr bool P11 = XM == 2;
    constexpr int NP = P11 ? 11 : 9;
    constexpr int NA = NP * (NP + 1) / 2;
    constexpr int NRED = NA + NP + 1;            /* alpha, beta, chi^2         */
    constexpr int PPW = 64 / LPP;
    constexpr int RSTR = P11 ? 80 : 60;        /* reduction row stride (floats) */
    /* REC: the LM loop's calls — models and constants in per-pixel records,
     * normal equations into per-pixel records, a device pixel list,
     * observations, no response output.  Fixed at compile time, the strides
     * and null tests need no scalar registers: with them at run time the
     * kernel spills about 130 SGPRs to VGPR lanes and reads them back with
     * one v_readlane (a VALU issue) each per pixel pair. */
    if constexpr (REC) {
        mfs = 1;
        mps = GM_MODEL_REC;
        aes = 1;
        bes = 1;
        aps = P11 ? GM_AB_REC11 : GM_AB_REC9;
        bps = aps;
    }
    const bool has_pre = REC || pre != nullptr;
    const bool has_list = REC || list != nullptr;
    const bool has_obs = REC || obs != nullptr;
    const bool has_alpha = REC || alpha != nullptr;
    const bool has_beta = REC || beta != nullptr;
    const bool has_rf = !REC && rf_out != nullptr;
    const bool want_chi = chi_out != nullptr;
    extern __shared__ float4 lds4[];
    const int lane = threadIdx.x;
    const int h = lane / LPP;
    const int k0 = lane % LPP;
    const int n = S->nlambda;
    const int64_t cnt = (REC || count) ? (int64_t)(*count) : n_direct;
    const int region4 = max(rows * 9, (LPP / 2) * (RSTR / 4));
    float4 *conv = lds4 + h * region4;
    float *red = reinterpret_cast<float *>(conv);
    /* EXT: macroturbulence kernel buffers after the PPW regions */
    float *gk = reinterpret_cast<float *>(lds4 + PPW * region4) + h * (P11 ? 4 : 2) * (LPP * NLPL);
    float *gc = gk + LPP * NLPL;
    float *gdk = gc + LPP * NLPL;                /* P11: d/dmac kernel          */
    float *gd = gdk + LPP * NLPL;
    const int use_psf = S->use_psf;
    const int dlo = S->psf_dlo, dhi = S->psf_dhi;
    const uint32_t fm = S->freemask;
    const int nfree = S->nfree_chi;
    /* up to 32 wavelengths: the PSF on the matrix cores, all taps; the tap
     * table after the pixel regions (ordered by the first staging barrier) */
    constexpr bool MFMA_OK = GM_MFMA_RF && LPP == 32 && NLPL == 1 && XM == 0;
    const bool mfma = MFMA_OK && use_psf && S->psf_mfma;
    float *ptab = reinterpret_cast<float *>(lds4 + PPW * region4);
    if (mfma) ptab[lane] = psf_tab_entry(S, lane);

    /* per-lane set-up values that are the same for every pixel, read once:
     * inside the pixel loop each is a vector load waited for at once, and
     * that wait also waits for every memory operation issued before it */
    float dl0[NLPL], edge_k[NLPL];
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        dl0[s] = S->dl[0][k < n ? k : n - 1];
        edge_k[s] = k < n ? S->psf_edge[k] : 0.f;
    }
    /* equal Stokes weights (the usual case): the normal-equation products
     * are summed unweighted and scaled once after the reduction.  A lane
     * past the grid must then contribute zero by itself: its PSF taps, edge
     * weight and S0 terms are scaled by zero (tapscale) -- the weighted sums
     * give it a zero weight instead.  One wavelength per lane, VALU PSF. */
    const bool uw = NLPL == 1 && XM == 0 && S->wsig_equal && !mfma;
    const float tapscale = (uw && k0 >= n) ? 0.f : 1.f;
    const float uwscale = uw ? S->wsig[0] : 1.f;
    /* normal-equation entries this lane writes (e = k0, k0 + LPP, ...):
     * bit i set when entry k0 + i LPP pairs two free parameters */
    uint32_t keep_e = 0;
    for (int e = k0, it = 0; e < NA + NP; e += LPP, ++it) {
        const int pp_ = (e < NA) ? (P11 ? kPairs11.p[e] : kPairs9.p[e]) : e - NA;
        const int qq_ = (e < NA) ? (P11 ? kPairs11.q[e] : kPairs9.q[e]) : e - NA;
        keep_e |= ((fm >> pp_) & (fm >> qq_) & 1u) << it;
    }

    /* grid-stride loop over pixel pairs (PairFeed: list entries and, for
     * REC, constants records prefetched) */
    const int64_t stride = (int64_t)gridDim.x * PPW;
    __shared__ float4 pfb[REC ? 2 * PairFeed<LPP, REC>::PFS : 1];
    PairFeed<LPP, REC> feed{list, pre, pfb, cnt, stride, h, k0, has_list};
    const int64_t first = (int64_t)blockIdx.x * PPW;
    feed.start(first);
    for (int64_t base = first; base < cnt; base += stride) {
        const int64_t item = base + h;
        const bool valid = item < cnt;
        const int64_t pix = feed.pixel(base);
        PixPre pp;
        if constexpr (REC) {
            pp = feed.advance_rec(base);
        } else {
            feed.advance(base);
            if (has_pre) {
                pp = load_pre(pre, pix);
            } else {
                double m[GM_NMODEL];
                load_model(models, mfs, mps, pix, m);
                pp = make_pre(m);
            }
        }
        /* the accepted trial becomes the best state here rather than in
         * k_decide: lanes 0-10 load its model and constants now and store
         * them at the end of the pixel, behind the computation */
        const bool copy_rec = model_dst && valid && k0 < GM_MODEL_REC / 2 + GM_PRE_REC / 4;
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        if (copy_rec)
            rec = (k0 < GM_MODEL_REC / 2)
                      ? reinterpret_cast<const float4 *>(models + pix * GM_MODEL_REC)[k0]
                      : reinterpret_cast<const float4 *>(pre + pix * GM_PRE_REC)[k0 - GM_MODEL_REC / 2];
        /* observed samples of this lane's wavelengths, loaded before the
         * synthesis so that their latency is hidden behind it */
        float ob[NLPL][4];
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
#pragma unroll
            for (int st = 0; st < 4; ++st)
                ob[s][st] = (has_obs && k < n) ? obs[pix * 4 * n + st * n + k] : 0.f;
        }

        /* zero rows: the padding and the wavelengths past the grid (the
         * reduction buffer aliases them) */
        for (int r = k0; r < rows; r += LPP)
            if (r < padlo || r >= padlo + n)
#pragma unroll
                for (int j = 0; j < 9; ++j) conv[r * 9 + j] = make_float4(0.f, 0.f, 0.f, 0.f);
        /* stage unconvolved values, one row per wavelength of the grid */
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
            float v[GM_SLOTS];
            eval_lambda<true>(S, pp, k < n ? k : n - 1, v, dl0[s]);
            /* 30 of the 36 values: d/dphi of I and V is zero and the S1
             * responses follow from the convolved synthesis (below) */
            float c[32];
#pragma unroll
            for (int i = 0; i < 30; ++i) c[i] = v[kStagedSlot[i]];
            c[30] = 0.f;
            c[31] = 0.f;
            if (k < n) {
                float4 *row = conv + (padlo + k) * 9;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    row[j] = make_float4(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3]);
            }
        }
        __syncthreads();
        /* EXT: this pixel's macroturbulence and filling factor */
        bool mac_on = false;
        float alv = 1.f, gsum = 1.f;
        int mlo = 0, mhi = n - 1;
        if (EXT) {
            const double mac = models[9 * mfs + pix * mps];
            alv = (float)models[10 * mfs + pix * mps];
            mac_on = mac > 0.0001;
            gsum = mac_kernel<LPP, NLPL, P11>(S, mac, gk, gc, k0, gdk, gd);
            mac_window<LPP, NLPL>(S, gc, gsum, k0, h, mlo, mhi);
        }
        const float s0c = (EXT && mac_on) ? gsum : 1.f;
        const float s0rf = (EXT && S->use_slight) ? alv : 1.f;
        const float inv_s1 = frcp(pp.S1);

        /* INIT1 (one wavelength per lane, no stray light): the first Stokes
         * parameter's products initialise the normal-equation sums, and a
         * lane past the grid contributes with weight zero instead of being
         * skipped */
        constexpr bool INIT1 = NLPL == 1 && XM == 0;
        float acc[NRED];
        if constexpr (!INIT1) {
#pragma unroll
            for (int e = 0; e < NRED; ++e) acc[e] = 0.f;
        } else if constexpr (MIXED) {
            acc[NRED - 1] = 0.f;
        }
        double chi = 0.0;
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
            float z[32];
            float zm[4] = {0.f, 0.f, 0.f, 0.f};  /* P11: d/dmac per Stokes */
            if (EXT && mac_on) {
                /* macroturbulence: circular correlation with the combined
                 * kernel, no Ic rule (me_der.cu:143-152) */
#pragma unroll
                for (int j = 0; j < 32; ++j) z[j] = 0.f;
                const int kk = k < n ? k : n - 1;
                for (int j = mlo; j <= mhi; ++j) {
                    const float tap = gc[j];
                    const float4 *row = conv + (padlo + circ_row(kk, n / 2, j, n)) * 9;
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const float4 x = row[q];
                        z[4 * q] += tap * x.x;
                        z[4 * q + 1] += tap * x.y;
                        z[4 * q + 2] += tap * x.z;
                        z[4 * q + 3] += tap * x.w;
                    }
                    if (P11) {
                        /* unconvolved spectra (staged positions 0, 7, 15,
                         * 23) with the kernel derivative, me_der.cu:139-141 */
                        const float td = gd[j];
                        zm[0] += td * row[0].x;
                        zm[1] += td * row[1].w;
                        zm[2] += td * row[3].w;
                        zm[3] += td * row[5].w;
                    }
                }
            } else if (MFMA_OK && mfma) {
                psf_mfma32(reinterpret_cast<const float *>(conv + padlo * 9) + k0, ptab + 31 - k0, z);
                /* I-variant (Ic trick) on the Stokes-I block, as below */
                const float edge = edge_k[s];
                const float4 *r0 = conv + padlo * 9;
                const float4 *rn = conv + (padlo + n - 1) * 9;
                const float4 a0 = r0[0], a1 = r0[1];
                const float4 b0 = rn[0], b1 = rn[1];
                z[0] += end_max(a0.x, b0.x) * edge;
                z[1] += end_max(a0.y, b0.y) * edge;
                z[2] += end_max(a0.z, b0.z) * edge;
                z[3] += end_max(a0.w, b0.w) * edge;
                z[4] += end_max(a1.x, b1.x) * edge;
                z[5] += end_max(a1.y, b1.y) * edge;
                z[6] += end_max(a1.z, b1.z) * edge;
            } else if (use_psf) {
                /* the first tap initialises the sums (dlo <= dhi whenever
                 * the PSF is on) */
                {
                    const float tap = S->psf_tap[0] * tapscale;
                    const float4 *row = conv + (padlo + k + dlo) * 9;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float4 x = row[j];
                        z[4 * j] = tap * x.x;
                        z[4 * j + 1] = tap * x.y;
                        z[4 * j + 2] = tap * x.z;
                        z[4 * j + 3] = tap * x.w;
                    }
                }
                for (int d = dlo + 1; d <= dhi; ++d) {
                    const float tap = S->psf_tap[d - dlo] * tapscale;
                    const float4 *row = conv + (padlo + k + d) * 9;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float4 x = row[j];
                        z[4 * j] += tap * x.x;
                        z[4 * j + 1] += tap * x.y;
                        z[4 * j + 2] += tap * x.z;
                        z[4 * j + 3] += tap * x.w;
                    }
                }
                /* I-variant (Ic trick) on the Stokes-I block: adds
                 * Ic * (1 - sum of in-range taps), convolution.cu:200-241 */
                const float edge = edge_k[s];
                const float4 *r0 = conv + padlo * 9;
                const float4 *rn = conv + (padlo + n - 1) * 9;
                const float4 a0 = r0[0], a1 = r0[1];
                const float4 b0 = rn[0], b1 = rn[1];
                z[0] += end_max(a0.x, b0.x) * edge;
                z[1] += end_max(a0.y, b0.y) * edge;
                z[2] += end_max(a0.z, b0.z) * edge;
                z[3] += end_max(a0.w, b0.w) * edge;
                z[4] += end_max(a1.x, b1.x) * edge;
                z[5] += end_max(a1.y, b1.y) * edge;
                z[6] += end_max(a1.z, b1.z) * edge;
            } else {
                const float4 *row = conv + (padlo + k) * 9;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float4 x = row[j];
                    z[4 * j] = x.x; z[4 * j + 1] = x.y; z[4 * j + 2] = x.z; z[4 * j + 3] = x.w;
                }
            }
            /* back to the 36 slots.  With I = S0 + S1 f and Q, U, V = S1 g
             * (mil_sinrf.cu:346-424) the S1 responses are f and g, and by
             * linearity of the correlation (and of the Ic rule, whose end
             * maximum of f is that of I minus S0, over S1) their convolved
             * values are (conv(I) - S0)/S1 and conv(Q, U, V)/S1.           */
            float y[GM_SLOTS];
#pragma unroll
            for (int i = 0; i < 30; ++i) y[kStagedSlot[i]] = z[i];
            y[7] = 0.f;
            y[34] = 0.f;
            y[8] = (y[0] - pp.S0 * (s0c * tapscale)) * inv_s1;
            y[17] = y[9] * inv_s1;
            y[26] = y[18] * inv_s1;
            y[35] = y[27] * inv_s1;
            const bool in_grid = k < n;
            if (INIT1 || in_grid) {
                float ra[4] = {0.f, 0.f, 0.f, 0.f};  /* P11: d/dalfa per Stokes */
                if (EXT && S->use_slight) {
                    /* stray light (mil_sinrf.cu:133-140, me_der.cu:191-207):
                     * d/dalfa from the magnetic spectrum at the first
                     * evaluation and from the mixed one afterwards, as the
                     * reference's LM loop passes it (milosUtils.cu:673) */
#pragma unroll
                    for (int st = 0; st < 4; ++st) {
                        const float ls = S->slight_map ? S->slight_map[pix * 4 * n + st * n + k]
                                                       : S->slight[st * n + k];
                        const float pure = y[st * 9];
                        y[st * 9] = alv * pure + (1.f - alv) * ls;
                        if (P11) ra[st] = (alpha_first ? pure : y[st * 9]) - ls;
#pragma unroll
                        for (int p = 1; p < 9; ++p) y[st * 9 + p] *= alv;
                        if (P11) zm[st] *= alv;
                    }
                }
                /* responses of every parameter, fixed ones included: their
                 * rows and columns are zeroed after the reduction
                 * (FijaACeroDerivadasNoNecesarias, milosUtils.cu:243-271) */
                float g[4][NP];
                float r[4];
#pragma unroll
                for (int st = 0; st < 4; ++st) {
                    r[st] = has_obs ? y[st * 9] - ob[s][st] : 0.f;
#pragma unroll
                    for (int p = 0; p < 9; ++p) {
                        const int js = staged_slot(p);
                        g[st][p] = (js >= 0) ? y[st * 9 + js] : (st == 0 ? s0rf * tapscale : 0.f);
                    }
                    if constexpr (P11) {
                        g[st][9] = zm[st];
                        g[st][10] = ra[st];
                    }
                }
                auto accumulate = [&](auto weighted) {
                    constexpr bool W = decltype(weighted)::value;
#pragma unroll
                    for (int st = 0; st < 4 && has_obs; ++st) {
                        const float ws = (INIT1 && !in_grid) ? 0.f : S->wsig[st];
                        const bool first = INIT1 && st == 0;
                        int e = 0;
#pragma unroll
                        for (int p = 0; p < NP; ++p) {
                            const float wg = W ? ws * g[st][p] : g[st][p];
#pragma unroll
                            for (int q = p; q < NP; ++q, ++e) {
                                if (!rf_nonzero(st, p) || !rf_nonzero(st, q)) {
                                    if (first) acc[e] = 0.f;
                                    continue;
                                }
                                acc[e] = first ? wg * g[st][q] : fmaf(wg, g[st][q], acc[e]);
                            }
                        }
                        const float wr = W ? ws * r[st] : r[st];
#pragma unroll
                        for (int p = 0; p < NP; ++p) {
                            if (!rf_nonzero(st, p)) {
                                if (first) acc[NA + p] = 0.f;
                                continue;
                            }
                            acc[NA + p] = first ? wr * g[st][p] : fmaf(wr, g[st][p], acc[NA + p]);
                        }
                    }
                };
                if (uw)
                    accumulate(std::false_type{});
                else
                    accumulate(std::true_type{});
                /* chi^2 only where it is asked for: the LM loop's calls
                 * take it from the trial synthesis */
                if (want_chi) {
#pragma unroll
                    for (int st = 0; st < 4 && has_obs; ++st) {
                        const float ws = (INIT1 && !in_grid) ? 0.f : S->wsig[st];
                        if (MIXED)
                            chi += (double)ws * ((double)r[st] * (double)r[st]);
                        else
                            acc[NRED - 1] = (INIT1 && st == 0)
                                                ? (r[st] * r[st]) * ws
                                                : fmaf(r[st] * r[st], ws, acc[NRED - 1]);
                    }
                }
                if (in_grid && synth_out && valid) {
                    float *so = synth_out + pix * 4 * n + k;
#pragma unroll
                    for (int st = 0; st < 4; ++st) so[st * n] = y[st * 9];
                }
                if (in_grid && has_rf && valid) {
                    if constexpr (P11) {
                        /* reference layout [4][NTERMS][n]: NTERMS 10 keeps
                         * the filling factor in slot 9, 11 has mac in 9 and
                         * alfa in 10; fixed parameters' responses are zero */
                        const int nto = S->nt_ref;
                        float *ro = rf_out + pix * 4 * nto * n + k;
#pragma unroll
                        for (int st = 0; st < 4; ++st) {
#pragma unroll
                            for (int p = 0; p < 9; ++p)
                                ro[(st * nto + p) * n] = ((fm >> p) & 1u) ? g[st][p] : 0.f;
                            const float g9 = ((fm >> 9) & 1u) ? g[st][9] : 0.f;
                            const float g10 = ((fm >> 10) & 1u) ? g[st][10] : 0.f;
                            ro[(st * nto + 9) * n] = (nto == 11) ? g9 : g10;
                            if (nto == 11) ro[(st * nto + 10) * n] = g10;
                        }
                    } else {
                        float *ro = rf_out + pix * 4 * 10 * n + k;
#pragma unroll
                        for (int st = 0; st < 4; ++st)
#pragma unroll
                            for (int p = 0; p < 10; ++p)
                                ro[(st * 10 + p) * n] =
                                    (p < 9 && ((fm >> p) & 1u)) ? g[st][p] : 0.f;
                    }
                }
            }
        }
        /* cross-lane reduction of the normal-equation partial sums: fold
         * the two lane halves of the pixel pairwise (lane l < HL keeps the
         * even entries, lane l + HL the odd ones), then sum the HL rows of
         * each half through LDS (aliasing the staged rows): entry e lives
         * at red[l * RSTR + (e & 1) * HALF + e / 2] */
        constexpr int HL = LPP / 2;
        constexpr int NSUM = NRED - 1;                  /* alpha and beta entries */
        constexpr int NPAIR = (NSUM + 1) / 2;
        constexpr int HALF = (NPAIR + 3) / 4 * 4;
        static_assert(2 * HALF <= RSTR, "reduction row");
        const float chi_part = acc[NRED - 1];    /* FP32 chi^2 term of this lane */
        float fs[NPAIR];
#pragma unroll
        for (int j = 0; j < NPAIR; ++j)
            fs[j] = fold_pair<HL>(acc[2 * j], (2 * j + 1 < NSUM) ? acc[2 * j + 1] : 0.f);
        __syncthreads();
        {
            float4 *rrow = reinterpret_cast<float4 *>(red + (k0 % HL) * RSTR + (k0 / HL) * HALF);
#pragma unroll
            for (int j = 0; j < HALF / 4; ++j)
                rrow[j] = make_float4((4 * j < NPAIR) ? fs[4 * j] : 0.f,
                                      (4 * j + 1 < NPAIR) ? fs[4 * j + 1] : 0.f,
                                      (4 * j + 2 < NPAIR) ? fs[4 * j + 2] : 0.f,
                                      (4 * j + 3 < NPAIR) ? fs[4 * j + 3] : 0.f);
        }
        __syncthreads();
        for (int e = k0, it = 0; e < NSUM; e += LPP, ++it) {
            const float *col = red + (e & 1) * HALF + (e >> 1);
            float sum = 0.f;
#pragma unroll 8
            for (int l = 0; l < HL; ++l) sum += col[l * RSTR];
            /* rows and columns of fixed parameters are zero */
            sum = ((keep_e >> it) & 1u) ? sum * uwscale : 0.f;
            if (valid) {
                if (e < NA) {
                    if (has_alpha) alpha[e * aes + pix * aps] = sum;
                } else {
                    if (has_beta) beta[(e - NA) * bes + pix * bps] = sum;
                }
            }
        }
        if (want_chi) {
            if (MIXED) {
                chi = group_sum_first<double, LPP>(chi) / (double)nfree;
            } else {
                const float c = group_sum_first<float, LPP>(chi_part);
                chi = (double)(c / (float)nfree);
            }
            if (valid && k0 == 0) chi_out[pix] = chi;
        }
        if (copy_rec) {
            if (k0 < GM_MODEL_REC / 2)
                reinterpret_cast<float4 *>(model_dst + pix * GM_MODEL_REC)[k0] = rec;
            else
                reinterpret_cast<float4 *>(pre_dst + pix * GM_PRE_REC)[k0 - GM_MODEL_REC / 2] = rec;
        }
        __syncthreads();
    }
}

/* Synthesis + PSF + chi^2 only (rejected steps need nothing more). */
/* Minimum waves per SIMD k_synth is compiled for (1: no constraint; the
 * 30-lambda kernel takes 68 VGPRs, 7 waves).  At 8 it fits 64 VGPRs with
 * 12 B of spills: 8.25 -> 8.65 ms of synthesis per 1M-pixel inversion. */
#ifndef GM_SYNTH_WAVES
#define GM_SYNTH_WAVES 1
#endif
template <int LPP, int NLPL, bool MIXED, bool EXT, bool REC>
__global__ __launch_bounds__(64, GM_SYNTH_WAVES) void k_synth(
    const gm_dsetup *__restrict__ S, const double *__restrict__ models,
    const float *__restrict__ pre, int64_t mfs, int64_t mps, const float *__restrict__ obs,
    const int32_t *__restrict__ list, const int32_t *__restrict__ count,
    int64_t n_direct, double *__restrict__ chi_out, float *__restrict__ synth_out,
    int rows, int padlo)
{
    constexpr int PPW = 64 / LPP;
    /* REC: the LM loop's call (per-pixel model records with constants, a
     * device pixel list, observations, a chi^2 output), fixed at compile
     * time as in k_rf */
    if constexpr (REC) {
        mfs = 1;
        mps = GM_MODEL_REC;
    }
    const bool has_pre = REC || pre != nullptr;
    const bool has_list = REC || list != nullptr;
    const bool has_obs = REC || obs != nullptr;
    const bool has_chi = REC || chi_out != nullptr;
    extern __shared__ float4 lds4[];
    const int lane = threadIdx.x;
    const int h = lane / LPP;
    const int k0 = lane % LPP;
    const int n = S->nlambda;
    const int64_t cnt = (REC || count) ? (int64_t)(*count) : n_direct;
    float4 *conv = lds4 + h * rows;
    /* EXT: macroturbulence kernel buffers after the PPW row blocks */
    float *gk = reinterpret_cast<float *>(lds4 + PPW * rows) + h * 2 * (LPP * NLPL);
    float *gc = gk + LPP * NLPL;
    const int use_psf = S->use_psf;
    const int dlo = S->psf_dlo, dhi = S->psf_dhi;
    const int nfree = S->nfree_chi;

    /* zero rows, written once: the padding and the wavelengths past the
     * grid (no lane stages those) */
    for (int r = k0; r < rows; r += LPP)
        if (r < padlo || r >= padlo + n) conv[r] = make_float4(0.f, 0.f, 0.f, 0.f);

    /* per-lane set-up values that are the same for every pixel, read once:
     * inside the pixel loop each is a vector load waited for at once, and
     * that wait also waits for every memory operation issued before it */
    float dl0[NLPL], edge_k[NLPL];
#pragma unroll
    for (int s = 0; s < NLPL; ++s) {
        const int k = k0 + s * LPP;
        dl0[s] = S->dl[0][k < n ? k : n - 1];
        edge_k[s] = k < n ? S->psf_edge[k] : 0.f;
    }

    /* grid-stride loop over pixel pairs (PairFeed: list entries and, for
     * REC, constants records prefetched) */
    const int64_t stride = (int64_t)gridDim.x * PPW;
    __shared__ float4 pfb[REC ? 2 * PairFeed<LPP, REC>::PFS : 1];
    PairFeed<LPP, REC> feed{list, pre, pfb, cnt, stride, h, k0, has_list};
    const int64_t first = (int64_t)blockIdx.x * PPW;
    feed.start(first);
    for (int64_t base = first; base < cnt; base += stride) {
        const int64_t item = base + h;
        const bool valid = item < cnt;
        const int64_t pix = feed.pixel(base);
        PixPre pp;
        if constexpr (REC) {
            pp = feed.advance_rec(base);
        } else {
            feed.advance(base);
            if (has_pre) {
                pp = load_pre(pre, pix);
            } else {
                double m[GM_NMODEL];
                load_model(models, mfs, mps, pix, m);
                pp = make_pre(m);
            }
        }
        /* observed samples of this lane's wavelengths, loaded before the
         * synthesis so that their latency is hidden behind it */
        float ob[NLPL][4];
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
#pragma unroll
            for (int st = 0; st < 4; ++st)
                ob[s][st] = (has_obs && k < n) ? obs[pix * 4 * n + st * n + k] : 0.f;
        }
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
            float v[GM_SLOTS];
            eval_lambda<false>(S, pp, k < n ? k : n - 1, v, dl0[s]);
            if (k < n) conv[padlo + k] = make_float4(v[0], v[9], v[18], v[27]);
        }
        __syncthreads();
        bool mac_on = false;
        float alv = 1.f;
        int mlo = 0, mhi = n - 1;
        if (EXT) {
            const double mac = models[9 * mfs + pix * mps];
            alv = (float)models[10 * mfs + pix * mps];
            mac_on = mac > 0.0001;
            const float gsum = mac_kernel<LPP, NLPL>(S, mac, gk, gc, k0);
            mac_window<LPP, NLPL>(S, gc, gsum, k0, h, mlo, mhi);
        }
        double chi = 0.0;
        float chif = 0.f;
#pragma unroll
        for (int s = 0; s < NLPL; ++s) {
            const int k = k0 + s * LPP;
            float4 y;
            if (EXT && mac_on) {
                /* macroturbulence (mil_sinrf.cu:101-116): circular, no Ic rule */
                y = make_float4(0.f, 0.f, 0.f, 0.f);
                const int kk = k < n ? k : n - 1;
                for (int j = mlo; j <= mhi; ++j) {
                    const float tap = gc[j];
                    const float4 x = conv[padlo + circ_row(kk, n / 2, j, n)];
                    y.x += tap * x.x;
                    y.y += tap * x.y;
                    y.z += tap * x.z;
                    y.w += tap * x.w;
                }
            } else if (use_psf) {
                /* the first tap initialises the sums (dlo <= dhi) */
                {
                    const float tap = S->psf_tap[0];
                    const float4 x = conv[padlo + k + dlo];
                    y = make_float4(tap * x.x, tap * x.y, tap * x.z, tap * x.w);
                }
                for (int d = dlo + 1; d <= dhi; ++d) {
                    const float tap = S->psf_tap[d - dlo];
                    const float4 x = conv[padlo + k + d];
                    y.x += tap * x.x;
                    y.y += tap * x.y;
                    y.z += tap * x.z;
                    y.w += tap * x.w;
                }
                const float edge = edge_k[s];
                y.x += end_max(conv[padlo].x, conv[padlo + n - 1].x) * edge;
            } else {
                y = conv[padlo + k];
            }
            if (EXT && S->use_slight && k < n) {
                const float *sl = S->slight_map ? S->slight_map + pix * 4 * n : S->slight;
                y.x = alv * y.x + (1.f - alv) * sl[k];
                y.y = alv * y.y + (1.f - alv) * sl[n + k];
                y.z = alv * y.z + (1.f - alv) * sl[2 * n + k];
                y.w = alv * y.w + (1.f - alv) * sl[3 * n + k];
            }
            if (k < n && has_obs) {
                const float r0 = y.x - ob[s][0], r1 = y.y - ob[s][1], r2 = y.z - ob[s][2],
                            r3 = y.w - ob[s][3];
                if (MIXED) {
                    chi += (double)S->wsig[0] * ((double)r0 * r0) + (double)S->wsig[1] * ((double)r1 * r1) +
                           (double)S->wsig[2] * ((double)r2 * r2) + (double)S->wsig[3] * ((double)r3 * r3);
                } else {
                    chif += (r0 * r0) * S->wsig[0] + (r1 * r1) * S->wsig[1] +
                            (r2 * r2) * S->wsig[2] + (r3 * r3) * S->wsig[3];
                }
            }
            if (k < n && synth_out && valid) {
                float *so = synth_out + pix * 4 * n + k;
                so[0] = y.x; so[n] = y.y; so[2 * n] = y.z; so[3 * n] = y.w;
            }
        }
        if (MIXED) {
            chi = group_sum_first<double, LPP>(chi) / (double)nfree;
        } else {
            chif = group_sum_first<float, LPP>(chif);
            chi = (double)(chif / (float)nfree);
        }
        if (has_chi && valid && k0 == 0) chi_out[pix] = chi;
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------ */
/* 9x9 symmetric step solves, pixel-parallel, all indices static.           */
__host__ __device__ constexpr int ut(int i, int j)
{
    return (i <= j) ? (i * 9 - (i * (i - 1)) / 2 + (j - i))
                    : (j * 9 - (j * (j - 1)) / 2 + (i - j));
}

template <typename T> __device__ __forceinline__ T rsq(T x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double rsq<double>(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    return y * (1.5 - 0.5 * x * y * y);      /* one Newton step */
}

/* Rotation that annihilates a_pq of the symmetric 2x2 block
 * [[a_pp, a_pq], [a_pq, a_qq]] (the smaller of the two angles, |t| <= 1):
 * t = sign(d) 2 a_pq / (|d| + r), d = a_qq - a_pp, r = sqrt(d^2 + 4 a_pq^2),
 * c = 1 / sqrt(1 + t^2), s = t c.  In FP64 it is formed from two
 * reciprocal square roots instead of a square root and a division: with
 * y = 1/r and h = (1 + |d| y) / 2 (= c^2), z = 1/sqrt(h) gives c = h z,
 * s = sign(d) a_pq y z and t = s z; c^2 + s^2 = h z^2 = 1.  Each hardware
 * estimate takes two Newton steps, so the result is the IEEE formula's to
 * a few units in the last place.  r = 0 (a zero block) gives the identity. */
__device__ __forceinline__ double rsq_nr2(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    return y * (1.5 - 0.5 * x * y * y);
}

template <typename T>
__device__ __forceinline__ void jacobi_angle(T app, T aqq, T apq, T &c, T &sn, T &t)
{
    const T d = aqq - app;
    if constexpr (sizeof(T) == 8) {
        const double r2 = d * d + 4.0 * apq * apq;
        /* r2 = 0 only for a_pq = 0 = d: h = 1 gives c = 1, s = t = 0 */
        const bool rot = r2 > 0.0;
        const double y = rsq_nr2(rot ? r2 : 1.0);
        const double h = rot ? 0.5 + 0.5 * (fabs(d) * y) : 1.0;
        const double z = rsq_nr2(h);
        double sv = (apq * y) * z;
        sv = (d < 0.0) ? -sv : sv;
        c = h * z;
        sn = sv;
        t = sv * z;
    } else {
        const T den = fabs(d) + sqrt(d * d + T(4) * apq * apq);
        T ti = (den > T(0)) ? (T(2) * apq) / den : T(0);
        ti = (d < T(0)) ? -ti : ti;
        t = ti;
        c = rsq<T>(T(1) + ti * ti);
        sn = ti * c;
    }
}

/* SVD of the symmetric damped matrix (singular values = |eigenvalues|) by
 * Jacobi rotations, then the reference's pseudo-inverse
 * x = V diag(1/w, |w| > 1e-12) V^T b (milosUtils.cu:296-322).
 *
 * Rotations follow the round-robin (parallel) ordering: at step K the nine
 * indices form the four disjoint pairs ((K+i) mod 9, (K-i) mod 9), i = 1..4,
 * and index K waits, so a sweep is nine steps of four independent rotations
 * (all 36 pairs once) whose angle computations and updates interleave
 * instead of forming one 36-long dependency chain.  A' = J^T A J is applied
 * block-wise on the packed upper triangle: the pair's own 2x2 block is
 * diagonalised exactly, the 2x2 blocks between two pairs get G_i^T M G_j,
 * the idle index's entries get the column rotation, and V <- V J.          */
template <typename T, int K>
__device__ __forceinline__ void jacobi_step(T (&a)[45], T (&v)[81])
{
    constexpr int P[4] = {(K + 1) % 9, (K + 2) % 9, (K + 3) % 9, (K + 4) % 9};
    constexpr int Q[4] = {(K + 8) % 9, (K + 7) % 9, (K + 6) % 9, (K + 5) % 9};
    T c[4], sn[4], t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T app = a[ut(P[i], P[i])], aqq = a[ut(Q[i], Q[i])], apq = a[ut(P[i], Q[i])];
        jacobi_angle<T>(app, aqq, apq, c[i], sn[i], t[i]);
    }
    /* blocks between two pairs: rows of pair i, columns of pair j */
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j) {
            const T m00 = a[ut(P[i], P[j])], m01 = a[ut(P[i], Q[j])];
            const T m10 = a[ut(Q[i], P[j])], m11 = a[ut(Q[i], Q[j])];
            const T n00 = c[j] * m00 - sn[j] * m01, n01 = sn[j] * m00 + c[j] * m01;
            const T n10 = c[j] * m10 - sn[j] * m11, n11 = sn[j] * m10 + c[j] * m11;
            a[ut(P[i], P[j])] = c[i] * n00 - sn[i] * n10;
            a[ut(P[i], Q[j])] = c[i] * n01 - sn[i] * n11;
            a[ut(Q[i], P[j])] = sn[i] * n00 + c[i] * n10;
            a[ut(Q[i], Q[j])] = sn[i] * n01 + c[i] * n11;
        }
    /* the waiting index K against each pair */
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T akp = a[ut(K, P[i])], akq = a[ut(K, Q[i])];
        a[ut(K, P[i])] = c[i] * akp - sn[i] * akq;
        a[ut(K, Q[i])] = sn[i] * akp + c[i] * akq;
    }
    /* each pair's own block */
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T apq = a[ut(P[i], Q[i])];
        a[ut(P[i], P[i])] -= t[i] * apq;
        a[ut(Q[i], Q[i])] += t[i] * apq;
        a[ut(P[i], Q[i])] = T(0);
    }
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const T vp = v[r * 9 + P[i]], vq = v[r * 9 + Q[i]];
            v[r * 9 + P[i]] = c[i] * vp - sn[i] * vq;
            v[r * 9 + Q[i]] = sn[i] * vp + c[i] * vq;
        }
}

template <typename T, int MAXSWEEP>
__device__ __forceinline__ void solve_svd(T (&a)[45], const T (&b)[9], T (&x)[9])
{
    T v[81];
#pragma unroll
    for (int i = 0; i < 81; ++i) v[i] = (i % 10 == 0) ? T(1) : T(0);
    const T eps2 = (sizeof(T) == 8) ? T(1e-30) : T(1e-14);
    bool done = false;
    for (int sweep = 0; sweep < MAXSWEEP; ++sweep) {
        T off = 0, dg = 0;
#pragma unroll
        for (int p = 0; p < 9; ++p) {
            dg += a[ut(p, p)] * a[ut(p, p)];
#pragma unroll
            for (int q = p + 1; q < 9; ++q) off += a[ut(p, q)] * a[ut(p, q)];
        }
        done = done || (off <= eps2 * dg);
        if (__all(done)) break;
        if (!done) {
            jacobi_step<T, 0>(a, v);
            jacobi_step<T, 1>(a, v);
            jacobi_step<T, 2>(a, v);
            jacobi_step<T, 3>(a, v);
            jacobi_step<T, 4>(a, v);
            jacobi_step<T, 5>(a, v);
            jacobi_step<T, 6>(a, v);
            jacobi_step<T, 7>(a, v);
            jacobi_step<T, 8>(a, v);
        }
    }
    T aux[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) sum += b[k] * v[k * 9 + j];
        const T w = fabs(a[ut(j, j)]);
        aux[j] = (w > T(1e-12)) ? sum / w : T(0);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) sum += v[i * 9 + k] * aux[k];
        x[i] = sum;
    }
}

/* The same Jacobi SVD by one whole wave for one N x N matrix (N = 9, or 11
 * for the eleven-slot systems): element e of the list [A's packed upper
 * triangle, N(N+1)/2 entries; V row-major, N^2 entries] belongs to lane
 * e % 64, slot e / 64 (N = 9: A entry l or V entry l - 45 in slot 0, V entry
 * l + 19 in slot 1; N = 11: three slots), so a rotation step costs each lane
 * two or three element updates instead of every lane repeating the whole
 * step.  The elements live in registers and in LDS (A in full and V,
 * row-major); a step reads the three partner elements each update needs,
 * lanes 0 .. (N-1)/2 - 1 form the step's rotations and publish (c, s, t),
 * every lane reads the rotations of its elements' rows and columns, updates
 * and writes back.  The lanes of one wave reach LDS in program order, so no
 * barrier is needed between the phases, only a fence that keeps the
 * compiler from reordering them.  The rotations and update formulas are
 * those of jacobi_step; the sweep test sums the same squares (in another
 * order).                                                                   */
template <typename T, int N>
struct WaveSvd {
    static constexpr int P = (N - 1) / 2;   /* rotations per step */
    T rot[P + 1][4];       /* c, s, t of pair m (1..P); pair 0: the waiting index */
    T a[N * N];
    T v[N * N];
    T b[N], aux[N], x[N];
};

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

template <typename T>
__device__ __forceinline__ T wave_uniform_sum(T v)
{
    v = group_sum_first<T, 64>(v);
    if constexpr (sizeof(T) == 8)
        return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                                __builtin_amdgcn_readfirstlane(__double2loint(v)));
    else
        return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <int N>
__host__ __device__ constexpr int wpk(int i, int j) { return i * N - (i * (i - 1)) / 2 + (j - i); }

/* One slot of a lane: its element (row i, column j of A or V), the LDS
 * offsets it is written to (o, and ot = the mirrored entry of A), and
 * whether it exists (the last slot has more lanes than elements). */
struct WaveElem {
    int i, j, o, ot;
    bool isA, valid;
};
template <int N>
struct WaveLayout {
    static constexpr int NA = N * (N + 1) / 2, E = NA + N * N, S = (E + 63) / 64;
    /* 0: every element of slot s is in V, 1: every one in A, 2: both */
    static constexpr int kind(int s) { return 64 * s >= NA ? 0 : (64 * (s + 1) <= NA ? 1 : 2); }
};
template <int N>
__device__ __forceinline__ WaveElem wave_elem(int e)
{
    constexpr int NA = WaveLayout<N>::NA, E = WaveLayout<N>::E;
    WaveElem r;
    r.valid = e < E;
    r.isA = e < NA;
    if (r.isA) {
        int i = 0;
#pragma unroll
        for (int q = 1; q < N; ++q) i += (e >= wpk<N>(q, q)) ? 1 : 0;
        r.i = i;
        r.j = i + e - (i * N - (i * (i - 1)) / 2);
    } else {
        const int f = r.valid ? e - NA : N * N - 1;
        r.i = f / N;
        r.j = f - N * r.i;
    }
    r.o = r.i * N + r.j;
    r.ot = r.isA ? r.j * N + r.i : r.o;
    return r;
}

/* Step K of the round-robin ordering: index i is paired with (2K - i) mod N;
 * with m = (i - K) mod N it is the p of pair m (m = 1..P, sigma = -s) or the
 * q of pair N - m (m = P+1..N-1, sigma = +s); m = 0 is the waiting index.  */
template <typename T, int N, int K>
__device__ __forceinline__ void wave_jacobi_step(WaveSvd<T, N> &L, int lane,
                                                 const WaveElem (&el)[WaveLayout<N>::S],
                                                 T (&val)[WaveLayout<N>::S])
{
    constexpr int S = WaveLayout<N>::S, P = WaveSvd<T, N>::P, K2 = (2 * K) % N;
    const auto partner = [](int i) { const int r = K2 - i; return (r < 0) ? r + N : r; };
    const auto pairno = [](int i, bool &pside) {
        int m = i - K;
        m += (m < 0) ? N : 0;
        pside = m <= P;
        return pside ? m : N - m;
    };
    /* lanes 0 .. P-1: the rotation of pair mm (the other lanes repeat one) */
    const int mm = (lane % P) + 1;
    int p = K + mm, q = K - mm;
    p -= (p >= N) ? N : 0;
    q += (q < 0) ? N : 0;
    const T app = L.a[p * (N + 1)], aqq = L.a[q * (N + 1)], apq = L.a[p * N + q];
    int pi[S], pj[S];
    T y[S], z[S], w[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        T *const B = (WaveLayout<N>::kind(s) == 0) ? L.v
                     : (WaveLayout<N>::kind(s) == 1) ? L.a
                     : (el[s].isA ? L.a : L.v);
        pi[s] = partner(el[s].i);
        pj[s] = partner(el[s].j);
        y[s] = B[el[s].i * N + pj[s]];
        if (WaveLayout<N>::kind(s) != 0) {
            z[s] = B[pi[s] * N + el[s].j];
            w[s] = B[pi[s] * N + pj[s]];
        }
    }
    T c, sn, t;
    jacobi_angle<T>(app, aqq, apq, c, sn, t);
    if (lane < P) {
        L.rot[mm][0] = c;
        L.rot[mm][1] = sn;
        L.rot[mm][2] = t;
    }
    wave_fence();
    T nv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        bool ps_i, ps_j;
        const int ni = pairno(el[s].i, ps_i), nj = pairno(el[s].j, ps_j);
        T cj = L.rot[nj][0], sj = L.rot[nj][1];
        if (WaveLayout<N>::kind(s) == 0) {
            asm volatile("" : "+v"(cj), "+v"(sj));
            const T gj = ps_j ? -sj : sj;
            nv[s] = cj * val[s] + gj * y[s];
        } else {
            T ci = L.rot[ni][0], si = L.rot[ni][1], ti = L.rot[ni][2];
            /* all five loaded here, together: otherwise the compiler moves
             * some of them into the branches it makes of the selects below,
             * behind a wait for the others */
            asm volatile("" : "+v"(ci), "+v"(si), "+v"(ti), "+v"(cj), "+v"(sj));
            const T gi = ps_i ? -si : si, gj = ps_j ? -sj : sj;
            /* A' = J^T A J, V' = V J: the column rotation of this element's
             * row, then (A) the row rotation; a pair's diagonal gets
             * a_pp - t a_pq, a_qq + t a_pq and its own off-diagonal element 0 */
            const T row = cj * val[s] + gj * y[s];
            const T gen = ci * row + gi * (cj * z[s] + gj * w[s]);
            const T dia = val[s] + (ps_i ? -ti : ti) * y[s];
            const bool pair_off = (el[s].j == pi[s]) && (el[s].i != el[s].j);
            T n = (el[s].i == el[s].j) ? dia : (pair_off ? T(0) : gen);
            if (WaveLayout<N>::kind(s) == 2) n = el[s].isA ? n : row;
            nv[s] = n;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        T *const B = (WaveLayout<N>::kind(s) == 0) ? L.v
                     : (WaveLayout<N>::kind(s) == 1) ? L.a
                     : (el[s].isA ? L.a : L.v);
        if (el[s].valid) {
            B[el[s].o] = nv[s];
            if (WaveLayout<N>::kind(s) != 0) B[el[s].ot] = nv[s];
        }
        val[s] = nv[s];
    }
    wave_fence();
}

template <typename T, int N>
__device__ __forceinline__ void wave_sweep(WaveSvd<T, N> &L, int lane,
                                           const WaveElem (&el)[WaveLayout<N>::S],
                                           T (&val)[WaveLayout<N>::S])
{
    wave_jacobi_step<T, N, 0>(L, lane, el, val);
    wave_jacobi_step<T, N, 1>(L, lane, el, val);
    wave_jacobi_step<T, N, 2>(L, lane, el, val);
    wave_jacobi_step<T, N, 3>(L, lane, el, val);
    wave_jacobi_step<T, N, 4>(L, lane, el, val);
    wave_jacobi_step<T, N, 5>(L, lane, el, val);
    wave_jacobi_step<T, N, 6>(L, lane, el, val);
    wave_jacobi_step<T, N, 7>(L, lane, el, val);
    wave_jacobi_step<T, N, 8>(L, lane, el, val);
    if constexpr (N > 9) {
        wave_jacobi_step<T, N, 9>(L, lane, el, val);
        wave_jacobi_step<T, N, 10>(L, lane, el, val);
    }
}

/* Pseudo-inverse step of one pixel's damped N x N system by the whole wave
 * (all 64 lanes active, the same pointers in every lane): al = packed
 * N(N+1)/2-entry matrix, be = right-hand side, damp = 1 + lambda on the
 * diagonal.  x is returned in every lane.                                  */
template <typename T, int N, int MAXSWEEP>
__device__ __forceinline__ void wave_svd_solve(WaveSvd<T, N> &L, const float *__restrict__ al,
                                               const float *__restrict__ be, T damp, T (&x)[N])
{
    static_assert(N == 9 || N == 11, "wave Jacobi size");
    constexpr int S = WaveLayout<N>::S;
    const int lane = threadIdx.x & 63;
    WaveElem el[S];
    T val[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        el[s] = wave_elem<N>(64 * s + lane);
        if (el[s].isA) {
            val[s] = (T)al[64 * s + lane];
            if (el[s].i == el[s].j) val[s] = val[s] * damp;
        } else {
            val[s] = (el[s].i == el[s].j) ? T(1) : T(0);
        }
        T *const B = el[s].isA ? L.a : L.v;
        if (el[s].valid) {
            B[el[s].o] = val[s];
            B[el[s].ot] = val[s];
        }
    }
    if (lane < 4) L.rot[0][lane] = (lane == 0) ? T(1) : T(0);
    if (lane < N) L.b[lane] = (T)be[lane];
    wave_fence();
    const T eps2 = (sizeof(T) == 8) ? T(1e-30) : T(1e-14);
    for (int sweep = 0; sweep < MAXSWEEP; ++sweep) {
        T off = 0, dg = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (WaveLayout<N>::kind(s) == 0) continue;
            const T sq = (el[s].isA && el[s].valid) ? val[s] * val[s] : T(0);
            off += (el[s].i != el[s].j) ? sq : T(0);
            dg += (el[s].i == el[s].j) ? sq : T(0);
        }
        off = wave_uniform_sum<T>(off);
        dg = wave_uniform_sum<T>(dg);
        if (off <= eps2 * dg) break;
        wave_sweep<T, N>(L, lane, el, val);
    }
    /* x = V diag(1/w, w > 1e-12) V^T b, w = |a_jj|: lane j forms entry j of
     * each product */
    if (lane < N) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) sum += L.b[k] * L.v[k * N + lane];
        const T w = fabs(L.a[lane * (N + 1)]);
        L.aux[lane] = (w > T(1e-12)) ? sum / w : T(0);
    }
    wave_fence();
    if (lane < N) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) sum += L.v[lane * N + k] * L.aux[k];
        L.x[lane] = sum;
    }
    wave_fence();
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = L.x[i];
    wave_fence();
}

/* LDL^T solve of the damped matrix (symmetric positive semi-definite by
 * construction).  A pivot of an exactly zero row (a fixed parameter) drops
 * that component, as the pseudo-inverse does.  Returns whether the result
 * provably equals the reference's pseudo-inverse solution: every other pivot
 * is positive and trace(A^-1) < 1e12, so every eigenvalue exceeds 1e-12 and
 * V diag(1/w) V^T = A^-1 on the remaining block.  The trace comes from the
 * factors: trace(A^-1) = sum_j (1/d_j) sum_k (L^-1)_jk^2.                   */
template <typename T>
__device__ __forceinline__ bool solve_ldl(const T (&a)[45], const T (&b)[9], T (&x)[9])
{
    T L[45], d[9], id[9];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const T ajj = a[ut(j, j)];
        T dj = ajj;
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= L[ut(j, k)] * L[ut(j, k)] * d[k];
        const bool null_row = (ajj == T(0));
        const bool pos = dj > T(0);
        ok = ok && (null_row || pos);
        const bool use = pos && !null_row;
        d[j] = use ? dj : T(0);
        id[j] = use ? T(1) / dj : T(0);
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            T lij = a[ut(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) lij -= L[ut(i, k)] * L[ut(j, k)] * d[k];
            L[ut(i, j)] = lij * id[j];
        }
    }
    T z[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        T zi = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) zi -= L[ut(i, k)] * z[k];
        z[i] = zi;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) z[i] *= id[i];
#pragma unroll
    for (int i = 8; i >= 0; --i) {
        T xi = z[i];
#pragma unroll
        for (int k = i + 1; k < 9; ++k) xi -= L[ut(k, i)] * x[k];
        x[i] = xi;
    }
    /* L^-1 in place (unit lower triangular), column by column */
#pragma unroll
    for (int j = 0; j < 9; ++j)
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            T mij = L[ut(i, j)];
#pragma unroll
            for (int k = j + 1; k < i; ++k) mij += L[ut(i, k)] * L[ut(k, j)];
            L[ut(i, j)] = -mij;
        }
    T tr = T(0);
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        T row = T(1);
#pragma unroll
        for (int k = 0; k < j; ++k) row += L[ut(j, k)] * L[ut(j, k)];
        tr += id[j] * row;
    }
    return ok && (tr < T(1e12));
}

/* Damped normal matrix alpha_ii (1 + lambda) (milosUtils.cu:648-652). */
template <typename T>
__device__ __forceinline__ void load_damped(const float *__restrict__ alpha,
                                            const float *__restrict__ beta, int64_t es,
                                            T flambda, T (&a)[45], T (&b)[9])
{
#pragma unroll
    for (int e = 0; e < 45; ++e) a[e] = (T)alpha[e * es];
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        b[p] = (T)beta[p * es];
        a[ut(p, p)] = a[ut(p, p)] * (T(1) + flambda);
    }
}

/* AplicaDelta + check (milosUtils.cu:30-238), double model. */
template <typename T, int N = 9>
__device__ __forceinline__ void apply_delta_check(const gm_dsetup *__restrict__ S,
                                                  const double *best, const T (&d)[N],
                                                  double *m)
{
    const int32_t *fix = S->fix;
    T aux;
    if (fix[0]) m[0] = best[0] - d[0];
    if (fix[1]) {
        aux = d[1];
        aux = aux < T(-300) ? T(-300) : (aux > T(300) ? T(300) : aux);
        m[1] = best[1] - aux;
    }
    if (fix[2]) m[2] = best[2] - d[2];
    if (fix[3]) m[3] = best[3] - d[3];
    if (fix[4]) m[4] = best[4] - d[4];
    if (fix[5]) {
        aux = d[5];
        aux = aux < T(-30) ? T(-30) : (aux > T(30) ? T(30) : aux);
        m[5] = best[5] - aux;
    }
    if (fix[6]) {
        aux = d[6];
        aux = aux < T(-30) ? T(-30) : (aux > T(30) ? T(30) : aux);
        m[6] = best[6] - aux;
    }
    if (fix[7]) m[7] = best[7] - d[7];
    if (fix[8]) m[8] = best[8] - d[8];
    if constexpr (N == 11) {
        /* 11 slots: mac in 9, filling factor in 10 (AplicaDelta,
         * milosUtils.cu:67-71; the 10-term AplicaDeltaf takes the filling
         * factor from its slot 9, milosUtils.cu:131-138, which is slot 10
         * here) */
        if (fix[9]) m[9] = best[9] - d[9];
        if (fix[10]) m[10] = best[10] - d[10];
    }
    /* check(): milosUtils.cu:160-238 */
    if (m[1] < 0) { m[1] = -m[1]; m[5] = 180.0 - m[5]; }
    if (m[1] > 5000) m[1] = 5000;
    if (m[5] < 0) m[5] = -m[5];
    if (m[5] > 180) m[5] = 360.0 - m[5];
    if (m[6] < 0) m[6] = 180 + m[6];
    if (m[6] > 180) m[6] = m[6] - 180.0;
    m[0] = m[0] < 1 ? 1 : (m[0] > 2500 ? 2500 : m[0]);
    m[2] = m[2] < -20 ? -20 : (m[2] > 20 ? 20 : m[2]);
    m[3] = m[3] < 0.0001 ? 0.0001 : (m[3] > 0.6 ? 0.6 : m[3]);
    m[4] = m[4] < 0.0001 ? 0.0001 : (m[4] > 10.0 ? 10.0 : m[4]);
    m[7] = m[7] < 0.0001 ? 0.0001 : (m[7] > 2.0 ? 2.0 : m[7]);
    m[8] = m[8] < 0.0001 ? 0.0001 : (m[8] > 2.0 ? 2.0 : m[8]);
    m[9] = m[9] < 0 ? 0 : (m[9] > 4 ? 4 : m[9]);
    m[10] = m[10] < 0 ? 0.0 : (m[10] > 1.0 ? 1.0 : m[10]);
}

/* the same with the best and trial records already in registers (m is
 * updated in place) */
template <typename T>
__device__ __forceinline__ void apply_loaded_to_trial(const gm_dsetup *__restrict__ S,
                                                      const gm_state &st, int64_t pix,
                                                      const T (&x)[9], const double (&best)[GM_NMODEL],
                                                      double (&m)[GM_NMODEL])
{
    apply_delta_check<T>(S, best, x, m);
#pragma unroll
    for (int f = 0; f < GM_NMODEL; ++f) st.trial[pix * GM_MODEL_REC + f] = m[f];
    store_pre(st.pre_trial, pix, make_pre(m));
}

template <typename T>
__device__ __forceinline__ void apply_to_trial(const gm_dsetup *__restrict__ S, const gm_state &st,
                                               int64_t pix, const T (&x)[9])
{
    /* free fields of the trial are best - delta; fixed ones keep the
     * trial's values, which check() may have changed (AplicaDeltaf writes
     * only free fields into the persisting model, milosUtils.cu:93-159) */
    double best[GM_NMODEL], m[GM_NMODEL];
#pragma unroll
    for (int f = 0; f < GM_NMODEL; ++f) {
        best[f] = st.best[pix * GM_MODEL_REC + f];
        m[f] = st.trial[pix * GM_MODEL_REC + f];
    }
    apply_loaded_to_trial<T>(S, st, pix, x, best, m);
}

/* ---- eleven-slot systems (free filling factor) ----------------------------
 * Slots 0-8 as above, 9 = mac, 10 = filling factor.  The reference solves
 * them as its 10-term (lm_mils, alfa free) or 11-term (lm_mils_11, mac and
 * alfa free) system; a slot of a fixed parameter holds a zero row, which the
 * pseudo-inverse drops, so both are one 11-slot solve here.               */
template <int N>
__host__ __device__ constexpr int utn(int i, int j)
{
    return (i <= j) ? (i * N - (i * (i - 1)) / 2 + (j - i))
                    : (j * N - (j * (j - 1)) / 2 + (i - j));
}

template <typename T, int N>
__device__ __forceinline__ void load_damped_n(const float *__restrict__ alpha,
                                              const float *__restrict__ beta, int64_t es,
                                              T flambda, T (&a)[N * (N + 1) / 2], T (&b)[N])
{
#pragma unroll
    for (int e = 0; e < N * (N + 1) / 2; ++e) a[e] = (T)alpha[e * es];
#pragma unroll
    for (int p = 0; p < N; ++p) {
        b[p] = (T)beta[p * es];
        a[utn<N>(p, p)] = a[utn<N>(p, p)] * (T(1) + flambda);
    }
}

/* LDL^T with the trace(A^-1) < 1e12 certificate (see solve_ldl), N slots. */
template <typename T, int N>
__device__ __forceinline__ bool solve_ldl_n(const T (&a)[N * (N + 1) / 2], const T (&b)[N],
                                            T (&x)[N])
{
    T L[N * (N + 1) / 2], d[N], id[N];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const T ajj = a[utn<N>(j, j)];
        T dj = ajj;
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= L[utn<N>(j, k)] * L[utn<N>(j, k)] * d[k];
        const bool null_row = (ajj == T(0));
        const bool pos = dj > T(0);
        ok = ok && (null_row || pos);
        const bool use = pos && !null_row;
        d[j] = use ? dj : T(0);
        id[j] = use ? T(1) / dj : T(0);
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            T lij = a[utn<N>(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) lij -= L[utn<N>(i, k)] * L[utn<N>(j, k)] * d[k];
            L[utn<N>(i, j)] = lij * id[j];
        }
    }
    T z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T zi = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) zi -= L[utn<N>(i, k)] * z[k];
        z[i] = zi;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] *= id[i];
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        T xi = z[i];
#pragma unroll
        for (int k = i + 1; k < N; ++k) xi -= L[utn<N>(k, i)] * x[k];
        x[i] = xi;
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            T mij = L[utn<N>(i, j)];
#pragma unroll
            for (int k = j + 1; k < i; ++k) mij += L[utn<N>(i, k)] * L[utn<N>(k, j)];
            L[utn<N>(i, j)] = -mij;
        }
    T tr = T(0);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        T row = T(1);
#pragma unroll
        for (int k = 0; k < j; ++k) row += L[utn<N>(j, k)] * L[utn<N>(j, k)];
        tr += id[j] * row;
    }
    return ok && (tr < T(1e12));
}

/* Cyclic Jacobi eigen-decomposition in round-robin order for odd N (step K:
 * pairs ((K+i) mod N, (K-i) mod N), i = 1..(N-1)/2, index K waits), then the
 * pseudo-inverse x = V diag(1/w, |w| > 1e-12) V^T b (milosUtils.cu:296-322).
 * Same scheme as jacobi_step / solve_svd; used for the pixels whose LDL^T
 * certificate fails.                                                       */
template <typename T, int N, int K>
__device__ __forceinline__ void jacobi_step_n(T (&a)[N * (N + 1) / 2], T (&v)[N * N])
{
    constexpr int H = (N - 1) / 2;
    T c[H], sn[H], t[H];
    int P[H], Q[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        P[i] = (K + 1 + i) % N;
        Q[i] = (K + N - 1 - i) % N;
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const T app = a[utn<N>(P[i], P[i])], aqq = a[utn<N>(Q[i], Q[i])];
        const T apq = a[utn<N>(P[i], Q[i])];
        jacobi_angle<T>(app, aqq, apq, c[i], sn[i], t[i]);
    }
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = i + 1; j < H; ++j) {
            const T m00 = a[utn<N>(P[i], P[j])], m01 = a[utn<N>(P[i], Q[j])];
            const T m10 = a[utn<N>(Q[i], P[j])], m11 = a[utn<N>(Q[i], Q[j])];
            const T n00 = c[j] * m00 - sn[j] * m01, n01 = sn[j] * m00 + c[j] * m01;
            const T n10 = c[j] * m10 - sn[j] * m11, n11 = sn[j] * m10 + c[j] * m11;
            a[utn<N>(P[i], P[j])] = c[i] * n00 - sn[i] * n10;
            a[utn<N>(P[i], Q[j])] = c[i] * n01 - sn[i] * n11;
            a[utn<N>(Q[i], P[j])] = sn[i] * n00 + c[i] * n10;
            a[utn<N>(Q[i], Q[j])] = sn[i] * n01 + c[i] * n11;
        }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const T akp = a[utn<N>(K, P[i])], akq = a[utn<N>(K, Q[i])];
        a[utn<N>(K, P[i])] = c[i] * akp - sn[i] * akq;
        a[utn<N>(K, Q[i])] = sn[i] * akp + c[i] * akq;
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const T apq = a[utn<N>(P[i], Q[i])];
        a[utn<N>(P[i], P[i])] -= t[i] * apq;
        a[utn<N>(Q[i], Q[i])] += t[i] * apq;
        a[utn<N>(P[i], Q[i])] = T(0);
    }
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int i = 0; i < H; ++i) {
            const T vp = v[r * N + P[i]], vq = v[r * N + Q[i]];
            v[r * N + P[i]] = c[i] * vp - sn[i] * vq;
            v[r * N + Q[i]] = sn[i] * vp + c[i] * vq;
        }
}

template <typename T, int N, int K>
struct JacobiSweep {
    static __device__ __forceinline__ void run(T (&a)[N * (N + 1) / 2], T (&v)[N * N])
    {
        jacobi_step_n<T, N, K>(a, v);
        JacobiSweep<T, N, K + 1>::run(a, v);
    }
};
template <typename T, int N>
struct JacobiSweep<T, N, N> {
    static __device__ __forceinline__ void run(T (&)[N * (N + 1) / 2], T (&)[N * N]) {}
};

template <typename T, int N, int MAXSWEEP>
__device__ __forceinline__ void solve_svd_n(T (&a)[N * (N + 1) / 2], const T (&b)[N], T (&x)[N])
{
    T v[N * N];
#pragma unroll
    for (int i = 0; i < N * N; ++i) v[i] = (i % (N + 1) == 0) ? T(1) : T(0);
    const T eps2 = (sizeof(T) == 8) ? T(1e-30) : T(1e-14);
    for (int sweep = 0; sweep < MAXSWEEP; ++sweep) {
        T off = 0, dg = 0;
#pragma unroll
        for (int p = 0; p < N; ++p) {
            dg += a[utn<N>(p, p)] * a[utn<N>(p, p)];
#pragma unroll
            for (int q = p + 1; q < N; ++q) off += a[utn<N>(p, q)] * a[utn<N>(p, q)];
        }
        if (off <= eps2 * dg) break;
        JacobiSweep<T, N, 0>::run(a, v);
    }
    T aux[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) sum += b[k] * v[k * N + j];
        const T w = fabs(a[utn<N>(j, j)]);
        aux[j] = (w > T(1e-12)) ? sum / w : T(0);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T sum = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) sum += v[i * N + k] * aux[k];
        x[i] = sum;
    }
}

/* ---- the reference's 10-term step solve ---------------------------------
 * mil_svdf (milosUtils.cu:328-365) over svdcordicf (svdcordic.cu:423-537):
 * NUM_ITER_SVD_CORDIC = 18 steps (definesCuda.cuh:119) of a Brent-Luk
 * parallel Jacobi eigen-decomposition of the damped 10 x 10 system in FP32,
 * then x = V diag(1/w, |w| > 1e-12) V^T b.  A step takes the five diagonal
 * 2 x 2 blocks of index pairs (2k, 2k+1) and their angles
 * theta_k = atan(2 a12 / (a22 - a11)) / 2 (pi/4 when a11 = a22, clamped to
 * +-pi/4), applies G_i^T B G_j (rows first, then columns) to every block
 * (i, j) with G = [[c, s], [-s, c]], zeroes the off-diagonal entry of the
 * diagonal blocks, rotates V's column pairs by G, and moves index i to
 * CP[i] (the systolic exchange, gmilos.cu:108-128).  No value moves here:
 * SIG.s[k][i] is the register slot that holds index i after k steps, known
 * at compile time; CP is a 9-cycle with 0 fixed, so every nine steps each
 * index is back in its own slot after the reference's 18 steps.          */
namespace cordic {
constexpr int CP[10] = {0, 2, 4, 1, 6, 3, 8, 5, 9, 7};
struct Frames {
    int s[9][10];
};
constexpr Frames make_frames()
{
    Frames f{};
    int inv[10] = {};
    for (int i = 0; i < 10; ++i) inv[CP[i]] = i;
    for (int i = 0; i < 10; ++i) f.s[0][i] = i;
    for (int k = 1; k < 9; ++k)
        for (int i = 0; i < 10; ++i) f.s[k][i] = f.s[k - 1][inv[i]];
    return f;
}
constexpr Frames SIG = make_frames();
constexpr float QPI = 0.78539816339744830962f;    /* pi / 4 */

/* (r0 r1) x (c0 c1) block of the symmetric matrix: rows rotated by
 * (cr, sr), then columns by (cc, sc), in the reference's operation order. */
__device__ __forceinline__ void rot_block(float &b00, float &b01, float &b10, float &b11, float cr,
                                          float sr, float cc, float sc)
{
    const float o11 = b00 * cr - b10 * sr, o21 = b00 * sr + b10 * cr;
    const float o12 = b01 * cr - b11 * sr, o22 = b01 * sr + b11 * cr;
    b00 = o11 * cc - o12 * sc;
    b01 = o11 * sc + o12 * cc;
    b10 = o21 * cc - o22 * sc;
    b11 = o21 * sc + o22 * cc;
}

/* Step K18 (0-17) of the sweep pair: the angles of the five diagonal blocks
 * (kept for the back-substitution), the two-sided block rotations, and the
 * same rotations applied to z = V^T b.  V itself is never formed: with V the
 * product of the step rotations G_0 ... G_17 (in register-slot terms), V^T b
 * is the rotations applied to b in order and V u is their transposes applied
 * in reverse order (rev). */
template <int K18>
__device__ __forceinline__ void fwd(float (&m)[55], float (&z)[10], volatile float *th)
{
    constexpr int K = K18 % 9;
    float c[5], s[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int i = SIG.s[K][2 * k], j = SIG.s[K][2 * k + 1];
        const float a11 = m[utn<10>(i, i)], a12 = m[utn<10>(i, j)], a22 = m[utn<10>(j, j)];
        float t;
        if (a11 == a22) {
            t = QPI;
        } else {
            /* __fdividef of the reference: a reciprocal and a multiply */
            t = 0.5f * atanf(2.f * a12 * __builtin_amdgcn_rcpf(a22 - a11));
            t = t > QPI ? QPI : (t < -QPI ? -QPI : t);
        }
        th[(K18 * 5 + k) * 64] = t;
        __sincosf(t, &s[k], &c[k]);
    }
#pragma unroll
    for (int bi = 0; bi < 5; ++bi)
#pragma unroll
        for (int bj = bi; bj < 5; ++bj) {
            const int r0 = SIG.s[K][2 * bi], r1 = SIG.s[K][2 * bi + 1];
            const int c0 = SIG.s[K][2 * bj], c1 = SIG.s[K][2 * bj + 1];
            float b00 = m[utn<10>(r0, c0)], b01 = m[utn<10>(r0, c1)];
            float b10 = m[utn<10>(r1, c0)], b11 = m[utn<10>(r1, c1)];
            rot_block(b00, b01, b10, b11, c[bi], s[bi], c[bj], s[bj]);
            m[utn<10>(r0, c0)] = b00;
            m[utn<10>(r1, c1)] = b11;
            if (bi == bj) {
                m[utn<10>(r0, c1)] = 0.f;
            } else {
                m[utn<10>(r0, c1)] = b01;
                m[utn<10>(r1, c0)] = b10;
            }
        }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int q0 = SIG.s[K][2 * k], q1 = SIG.s[K][2 * k + 1];
        const float z0 = z[q0], z1 = z[q1];
        z[q0] = z0 * c[k] - z1 * s[k];
        z[q1] = z0 * s[k] + z1 * c[k];
    }
    if constexpr (K18 < 17) fwd<K18 + 1>(m, z, th);
}

template <int K18>
__device__ __forceinline__ void rev(float (&u)[10], volatile float *th)
{
    constexpr int K = K18 % 9;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        float s, c;
        __sincosf(th[(K18 * 5 + k) * 64], &s, &c);
        const int q0 = SIG.s[K][2 * k], q1 = SIG.s[K][2 * k + 1];
        const float u0 = u[q0], u1 = u[q1];
        u[q0] = u0 * c + u1 * s;
        u[q1] = u1 * c - u0 * s;
    }
    if constexpr (K18 > 0) rev<K18 - 1>(u, th);
}

/* m: packed damped 10 x 10 upper triangle (destroyed); b, x: 10; th: this
 * lane's column of a [90][64] LDS array for the step angles (kept out of
 * registers, which the two matrices fill).  After the 18 steps every index
 * is back in its own register slot (CP^18 = I), so the diagonal of m is w in
 * index order. */
constexpr int TH_FLOATS = 90 * 64;
__device__ __forceinline__ void solve10(float (&m)[55], const float (&b)[10], float (&x)[10],
                                        volatile float *th)
{
    float z[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) z[i] = b[i];
    fwd<0>(m, z, th);
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        const float w = m[utn<10>(j, j)];
        x[j] = z[j] * ((fabsf(w) > 1e-12f) ? 1.f / w : 0.f);
    }
    rev<17>(x, th);
}

/* The 9-slot system (packed 45) as the reference's 10 x 10 with slot 9 (the
 * filling factor's response, zero while alfa is fixed) a zero row. */
template <typename T>
__device__ __forceinline__ void solve9(const T (&a)[45], const T (&b)[9], T (&x)[9],
                                       volatile float *th)
{
    float m[55], bb[10], xx[10];
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
        for (int j = i; j < 10; ++j) m[utn<10>(i, j)] = (j < 9) ? (float)a[ut(i, j)] : 0.f;
#pragma unroll
    for (int p = 0; p < 9; ++p) bb[p] = (float)b[p];
    bb[9] = 0.f;
    solve10(m, bb, xx, th);
#pragma unroll
    for (int p = 0; p < 9; ++p) x[p] = (T)xx[p];
}

/* The 11-slot system with mac fixed (slot 9 a zero row) as the reference's
 * 10-term one: slots 0-8, then the filling factor in slot 9. */
__device__ __forceinline__ void solve11_as10(const double (&a)[66], const double (&b)[11],
                                             double (&x)[11], volatile float *th)
{
    constexpr int MAP[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 10};
    float m[55], bb[10], xx[10];
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
        for (int j = i; j < 10; ++j) m[utn<10>(i, j)] = (float)a[utn<11>(MAP[i], MAP[j])];
#pragma unroll
    for (int p = 0; p < 10; ++p) bb[p] = (float)b[MAP[p]];
    solve10(m, bb, xx, th);
#pragma unroll
    for (int p = 0; p < 9; ++p) x[p] = xx[p];
    x[9] = 0.0;
    x[10] = xx[9];
}
} // namespace cordic

/* Step solve of the eleven-slot system for the active list, FP64 (the
 * reference's 11-term solve is FP64 svdcmp): LDL^T where it provably equals
 * the pseudo-inverse, then AplicaDelta + check; the other pixels go to the
 * fallback list, solved by k_solve_svd<., 11>.  (With the Jacobi inline, a
 * lane's matrix copy and V held the kernel at 256 VGPRs + 256 AGPRs, 804 B
 * of spills and one wave per SIMD.) */
template <bool CORDIC10>
__global__ __launch_bounds__(64) void k_solve11(const gm_dsetup *__restrict__ S, gm_state st,
                                                const int32_t *__restrict__ list,
                                                const int32_t *__restrict__ count,
                                                int32_t *__restrict__ reset0,
                                                int32_t *__restrict__ reset1,
                                                int32_t *__restrict__ hist,
                                                int32_t *__restrict__ fb_list,
                                                int32_t *__restrict__ fb_count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (hist) {
            hist[0] = *count;
            if (reset1) hist[-1] = *reset1;
        }
        if (reset0) *reset0 = 0;
        if (reset1) *reset1 = 0;
    }
    const int64_t cnt = *count;
    const int64_t cap = st.cap;
    for (int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x; i < cnt;
         i += (int64_t)gridDim.x * 64) {
        const int64_t pix = list[i];
        double a[66], b[11], x[11];
        load_damped_n<double, 11>(st.alpha + pix * st.ab_rec, st.beta + pix * st.ab_rec, 1,
                                  st.flambda[pix], a, b);
        if constexpr (CORDIC10) {
            __shared__ float th[cordic::TH_FLOATS];
            cordic::solve11_as10(a, b, x, th + threadIdx.x);
        } else if (!solve_ldl_n<double, 11>(a, b, x)) {
            /* no certificate: the wave Jacobi of k_solve_svd<., 11> */
            fb_list[atomicAdd(fb_count, 1)] = (int32_t)pix;
            continue;
        }
        double best[GM_NMODEL], m[GM_NMODEL];
#pragma unroll
        for (int f = 0; f < GM_NMODEL; ++f) {
            best[f] = st.best[pix * GM_MODEL_REC + f];
            m[f] = st.trial[pix * GM_MODEL_REC + f];
        }
        apply_delta_check<double, 11>(S, best, x, m);
#pragma unroll
        for (int f = 0; f < GM_NMODEL; ++f) st.trial[pix * GM_MODEL_REC + f] = m[f];
        store_pre(st.pre_trial, pix, make_pre(m));
    }
}

/* Piece entry: delta [n][11] from alpha [n][66], beta [n][11]. */
__global__ __launch_bounds__(64) void k_step_solve11(const float *__restrict__ alpha,
                                                     const float *__restrict__ beta,
                                                     const double *__restrict__ flambda,
                                                     int64_t n, double *__restrict__ delta,
                                                     int cordic10)
{
    for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
        const int64_t i = base + threadIdx.x;
        bool need_svd = false;
        if (i < n) {
            double a[66], b[11], x[11];
            load_damped_n<double, 11>(alpha + i * 66, beta + i * 11, 1, flambda[i], a, b);
            if (cordic10) {
                __shared__ float th[cordic::TH_FLOATS];
                cordic::solve11_as10(a, b, x, th + threadIdx.x);
            } else {
                need_svd = !solve_ldl_n<double, 11>(a, b, x);
            }
            if (!need_svd) {
#pragma unroll
                for (int p = 0; p < 11; ++p) delta[i * 11 + p] = x[p];
            }
        }
        /* the uncertified systems: the Jacobi SVD of the whole wave */
        uint64_t need = __ballot(need_svd);
        if (need) {
            __shared__ WaveSvd<double, 11> L;
            while (need) {
                const int l = __builtin_ctzll(need);
                need &= need - 1;
                const int64_t fi = base + l;
                double x[11];
                wave_svd_solve<double, 11, 14>(L, alpha + fi * 66, beta + fi * 11, 1.0 + flambda[fi], x);
                if ((int)threadIdx.x == l) {
#pragma unroll
                    for (int p = 0; p < 11; ++p) delta[fi * 11 + p] = x[p];
                }
            }
        }
    }
}

#ifndef GM_SOLVE_PREFETCH
#define GM_SOLVE_PREFETCH 1
#endif
/* GM_SOLVE_TRIAL_ALL = 1: k_solve prefetches the whole trial record; 0:
 * only mac and the filling factor, and a fixed parameter's field after the
 * solve (16 of 88 bytes with nine free parameters): solves 6.30 -> 6.15 ms
 * per 1M-pixel inversion. */
#ifndef GM_SOLVE_TRIAL_ALL
#define GM_SOLVE_TRIAL_ALL 0
#endif
template <typename T>
constexpr int svd_sweeps() { return sizeof(T) == 8 ? 12 : 10; }

/* Step solve for the active list.  SOLVER = GM_SOLVE_SVD: LDL^T where it
 * provably equals the SVD pseudo-inverse, otherwise the pixel goes to the
 * fallback list solved by k_solve_svd; GM_SOLVE_LDL: LDL^T only;
 * GM_SOLVE_JACOBI: always the Jacobi SVD, here; GM_SOLVE_CORDIC: the
 * reference's 18-step FP32 CORDIC solve (cordic::solve10).  fb_count
 * counts the Jacobi solves of both Jacobi modes. */
/* Minimum waves per SIMD k_solve is compiled for: 3 fits it in 166 VGPRs
 * without spills (178 unconstrained, 2 waves); per 1M-pixel inversion the
 * solves take 6.28 -> 6.20 ms; at 4 (128 VGPRs, 108 B of spills) 7.61 ms.
 * The Jacobi-always instantiation (a pixel's whole matrix and V in one
 * lane) is left unconstrained. */
#ifndef GM_SOLVE_WAVES
#define GM_SOLVE_WAVES 3
#endif
template <bool MIXED, int SOLVER>
__global__ __launch_bounds__(64, SOLVER == GM_SOLVE_JACOBI ? 1 : GM_SOLVE_WAVES) void k_solve(const gm_dsetup *__restrict__ S,
                                              gm_state st, const int32_t *__restrict__ list,
                                              const int32_t *__restrict__ count,
                                              int32_t *__restrict__ reset0,
                                              int32_t *__restrict__ reset1,
                                              int32_t *__restrict__ hist,
                                              int32_t *__restrict__ fb_list,
                                              int32_t *__restrict__ fb_count)
{
    /* hist[0] = active pixels this round, hist[-1] = accepted pixels of the
     * previous round (read before that list's counter is reset); hist[1],
     * the fallback count, is written by k_decide */
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (hist) {
            hist[0] = *count;
            if (reset1) hist[-1] = *reset1;
        }
        if (reset0) *reset0 = 0;
        if (reset1) *reset1 = 0;
    }
    using T = typename std::conditional<MIXED, double, float>::type;
    const int64_t cnt = *count;
    const int64_t cap = st.cap;
    for (int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x; i < cnt;
         i += (int64_t)gridDim.x * 64) {
        const int64_t pix = list[i];
        T a[45], b[9], x[9];
        /* the best and trial records are loaded together with the normal
         * equations, so their latency overlaps the solve's instead of
         * following it */
        constexpr bool PREF = GM_SOLVE_PREFETCH && (SOLVER == GM_SOLVE_SVD || SOLVER == GM_SOLVE_LDL);
        double best_m[GM_NMODEL], trial_m[GM_NMODEL];
        if constexpr (PREF) {
#pragma unroll
            for (int f = 0; f < GM_NMODEL; ++f) {
                best_m[f] = st.best[pix * GM_MODEL_REC + f];
                /* of the trial only the fields the step leaves alone are
                 * read: mac and filling factor here, a fixed parameter's
                 * below (a free one is overwritten by best - delta) */
                trial_m[f] = (GM_SOLVE_TRIAL_ALL || f >= 9) ? st.trial[pix * GM_MODEL_REC + f] : 0.0;
            }
        }
        load_damped<T>(st.alpha + pix * st.ab_rec, st.beta + pix * st.ab_rec, 1,
                       (T)st.flambda[pix], a, b);
        if constexpr (SOLVER == GM_SOLVE_CORDIC) {
            __shared__ float th[cordic::TH_FLOATS];
            cordic::solve9<T>(a, b, x, th + threadIdx.x);
            apply_to_trial<T>(S, st, pix, x);
            continue;
        }
        if (SOLVER == GM_SOLVE_JACOBI) {
            /* every pixel through the Jacobi SVD: one pixel per lane (the
             * wave Jacobi of the fallback list is one pixel per wave, for
             * the few uncertified pixels of the two-tier solve) */
            solve_svd<T, svd_sweeps<T>()>(a, b, x);
            atomicAdd(fb_count, 1);
            apply_to_trial<T>(S, st, pix, x);
            continue;
        }
        const bool proven = solve_ldl<T>(a, b, x);
        if (SOLVER == GM_SOLVE_SVD && !proven) {
            fb_list[atomicAdd(fb_count, 1)] = (int32_t)pix;
            continue;
        }
        if constexpr (PREF) {
            if (!GM_SOLVE_TRIAL_ALL) {
#pragma unroll
                for (int f = 0; f < 9; ++f)
                    if (!S->fix[f]) trial_m[f] = st.trial[pix * GM_MODEL_REC + f];
            }
            apply_loaded_to_trial<T>(S, st, pix, x, best_m, trial_m);
        } else {
            apply_to_trial<T>(S, st, pix, x);
        }
    }
}

/* Jacobi SVD of the fallback list of the two-tier solve, one wave per pixel
 * (wave_svd_solve).  Per 1M-pixel inversion the fallback pixels' Jacobi
 * took 1.5 ms with every lane repeating the rotations of the whole matrix. */
template <bool MIXED, int N>
__global__ __launch_bounds__(64) void k_solve_svd(const gm_dsetup *__restrict__ S, gm_state st,
                                                  const int32_t *__restrict__ fb_list,
                                                  const int32_t *__restrict__ fb_count)
{
    /* the eleven-slot systems are solved in FP64 in both precision modes */
    using T = typename std::conditional<MIXED || N == 11, double, float>::type;
    __shared__ WaveSvd<T, N> L;
    const int64_t cnt = *fb_count;
    const int64_t cap = st.cap;
    for (int64_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        const int64_t pix = fb_list[i];
        T x[N];
        wave_svd_solve<T, N, (N == 11 ? 14 : svd_sweeps<T>())>(
            L, st.alpha + pix * st.ab_rec, st.beta + pix * st.ab_rec, T(1) + (T)st.flambda[pix], x);
        if (threadIdx.x == 0) {
            if constexpr (N == 9) {
                apply_to_trial<T>(S, st, pix, x);
            } else {
                double best[GM_NMODEL], m[GM_NMODEL];
#pragma unroll
                for (int f = 0; f < GM_NMODEL; ++f) {
                    best[f] = st.best[pix * GM_MODEL_REC + f];
                    m[f] = st.trial[pix * GM_MODEL_REC + f];
                }
                apply_delta_check<double, 11>(S, best, x, m);
#pragma unroll
                for (int f = 0; f < GM_NMODEL; ++f) st.trial[pix * GM_MODEL_REC + f] = m[f];
                store_pre(st.pre_trial, pix, make_pre(m));
            }
        }
    }
}

template <bool MIXED, int SOLVER>
__global__ __launch_bounds__(64) void k_step_solve(const float *__restrict__ alpha,
                                                   const float *__restrict__ beta,
                                                   const double *__restrict__ flambda,
                                                   int64_t n, double *__restrict__ delta)
{
    using T = typename std::conditional<MIXED, double, float>::type;
    for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
        const int64_t i = base + threadIdx.x;
        bool need_svd = false;
        if (i < n) {
            /* caller layout: alpha [n][45], beta [n][9] */
            T a[45], b[9], x[9];
            load_damped<T>(alpha + i * 45, beta + i * 9, 1, (T)flambda[i], a, b);
            if constexpr (SOLVER == GM_SOLVE_CORDIC) {
                __shared__ float th[cordic::TH_FLOATS];
                cordic::solve9<T>(a, b, x, th + threadIdx.x);
            } else if (SOLVER == GM_SOLVE_JACOBI) {
                solve_svd<T, svd_sweeps<T>()>(a, b, x);
            } else {
                need_svd = !solve_ldl<T>(a, b, x) && SOLVER == GM_SOLVE_SVD;
            }
            if (!need_svd) {
#pragma unroll
                for (int p = 0; p < 9; ++p) delta[i * 9 + p] = (double)x[p];
            }
        }
        /* the uncertified systems: the Jacobi SVD of the whole wave, one
         * system after the other, as in the LM loop */
        if constexpr (SOLVER == GM_SOLVE_SVD) {
            uint64_t need = __ballot(need_svd);
            if (need) {
                __shared__ WaveSvd<T, 9> L;
                while (need) {
                    const int l = __builtin_ctzll(need);
                    need &= need - 1;
                    const int64_t fi = base + l;
                    T x[9];
                    wave_svd_solve<T, 9, svd_sweeps<T>()>(L, alpha + fi * 45, beta + fi * 9,
                                                          T(1) + (T)flambda[fi], x);
                    if ((int)threadIdx.x == l) {
#pragma unroll
                        for (int p = 0; p < 9; ++p) delta[fi * 9 + p] = (double)x[p];
                    }
                }
            }
        }
    }
}

/* Workgroup size and grid cap of k_decide.  Per 1M-pixel inversion the
 * decisions take the same 2.5 ms (with the other small kernels) with caps of
 * 256, 512 or 1024 workgroups of 1024; workgroups of 512 or 256 (more of
 * them, so more list atomics) take 3.1 and 4.6 ms.  The three list
 * reservations of a workgroup from three lanes at once instead of one after
 * the other from one lane: 1.44 -> 1.15 ms per inversion. */
#ifndef GM_DECIDE_THREADS
#define GM_DECIDE_THREADS 1024
#endif
#ifndef GM_DECIDE_BLOCKS
#define GM_DECIDE_BLOCKS 512
#endif
/* Accept / reject and LM bookkeeping (milosUtils.cu:660-698); builds the
 * next active list and the list of accepted pixels that need new normal
 * equations, with one atomic per workgroup and list.  THREADS: workgroup
 * size (GM_DECIDE_THREADS for an inversion; the serving loop, whose rounds
 * run beside other set-ups' kernels, takes 256, which fit partly occupied
 * CUs). */
template <bool MIXED, int THREADS>
__global__ __launch_bounds__(THREADS) void k_decide(const gm_dsetup *__restrict__ S, gm_state st,
                                                 const int32_t *__restrict__ list,
                                                 const int32_t *__restrict__ count,
                                                 int32_t *__restrict__ next_list,
                                                 int32_t *__restrict__ next_count,
                                                 int32_t *__restrict__ acc_list,
                                                 int32_t *__restrict__ acc_count,
                                                 int32_t *__restrict__ reset_fb,
                                                 int32_t *__restrict__ done_list,
                                                 int32_t *__restrict__ done_count,
                                                 int defer_copy, int32_t *__restrict__ fb_hist)
{
    using T = typename std::conditional<MIXED, double, float>::type;
    __shared__ int wc[THREADS / 64], wa[THREADS / 64], wd[THREADS / 64];
    __shared__ int base_c, base_a, base_d;
    /* the solve kernel of this round is complete: its SVD count is final */
    if (blockIdx.x == 0 && threadIdx.x == 0 && reset_fb) {
        if (fb_hist) *fb_hist = *reset_fb;
        *reset_fb = 0;
    }
    const int64_t cnt = *count;
    const int64_t cap = st.cap;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t base = (int64_t)blockIdx.x * THREADS; base < cnt;
         base += (int64_t)gridDim.x * THREADS) {
        const int64_t i = base + tid;
        const bool valid = i < cnt;
        int32_t pix = valid ? list[i] : 0;
        bool cont = false, accn = false;
        if (valid) {
            const T chisqr = (T)st.chi_trial[pix];
            T ochisqr = (T)st.ochisqr[pix];
            const int it = st.iter[pix];
            const LmStep<T> d = lm_step<T>(S, chisqr, ochisqr, (T)st.flambda[pix],
                                           (T)st.factor[pix], it);
            /* with defer_copy the RF kernel copies the accepted trial of
             * a continuing pixel; finished pixels are copied here */
            if (d.accept && !(defer_copy && d.cont)) {
#pragma unroll
                for (int q = 0; q < GM_MODEL_REC / 2; ++q)
                    reinterpret_cast<double2 *>(st.best + pix * GM_MODEL_REC)[q] =
                        reinterpret_cast<const double2 *>(st.trial + pix * GM_MODEL_REC)[q];
#pragma unroll
                for (int q = 0; q < GM_PRE_REC / 4; ++q)
                    reinterpret_cast<float4 *>(st.pre_best + pix * GM_PRE_REC)[q] =
                        reinterpret_cast<const float4 *>(st.pre_trial + pix * GM_PRE_REC)[q];
            }
            if (d.accept) ochisqr = chisqr;
            if (S->logclambda) {
                const T c0 = (T)st.chisqr0[pix];
                st.factor[pix] = (double)(log10(chisqr) / log10(c0));
            }
            st.flambda[pix] = (double)d.flambda;
            st.ochisqr[pix] = (double)ochisqr;
            st.iter[pix] = it + 1;
            cont = d.cont;
            accn = d.accept && cont;
            if (accn) st.nacc[pix] += 1;
        }
        const bool fin = valid && !cont && done_list;
        const uint64_t mc = __ballot(cont);
        const uint64_t ma = __ballot(accn);
        const uint64_t md = __ballot(fin);
        if (lane == 0) {
            wc[wave] = (int)__popcll(mc);
            wa[wave] = (int)__popcll(ma);
            wd[wave] = (int)__popcll(md);
        }
        __syncthreads();
        /* lanes 0, 1, 2: prefix sums of the next, accepted and done
         * counts over the waves, and one list reservation each; the three
         * atomics go out as one instruction (issued from a single lane,
         * each one's return was waited for before the next was issued) */
        if (tid < 3) {
            int *wl = (tid == 0) ? wc : (tid == 1) ? wa : wd;
            int32_t *ctr = (tid == 0) ? next_count : (tid == 1) ? acc_count : done_count;
            int sum = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                const int c = wl[w];
                wl[w] = sum;
                sum += c;
            }
            const int b = (sum && ctr) ? atomicAdd(ctr, sum) : 0;
            int *dst = (tid == 0) ? &base_c : (tid == 1) ? &base_a : &base_d;
            *dst = b;
        }
        __syncthreads();
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (cont) next_list[base_c + wc[wave] + __popcll(mc & below)] = pix;
        if (accn) acc_list[base_a + wa[wave] + __popcll(ma & below)] = pix;
        if (fin) done_list[base_d + wd[wave] + __popcll(md & below)] = pix;
        __syncthreads();
    }
}

/* Classical estimates, FP64 (milosUtils.cu:386-563), pixel-parallel. */
__device__ void classical_estimates(const gm_dsetup *__restrict__ S, const float *sp,
                                    double *model, int initial)
{
    const int n = S->nlambda;
    const int lo = S->ce_lo, hi = S->ce_hi;
    const double lambda_0 = S->ce_l0;
    const float *sI = sp, *sQ = sp + n, *sU = sp + 2 * n, *sV = sp + 3 * n;
    const double *lambda = S->lambda;
    int i;
    int begin = lo;
    for (i = lo; i < hi; i++) { if (sI[i] < 0) begin++; else break; }
    int end = hi;
    for (i = hi - 1; i >= lo; i--) { if (sI[i] < 0) end--; else break; }
    /* all-negative I: the reference would read I[-1]; use the first sample */
    const double Ic = (end > lo) ? sI[end - 1] : sI[lo];
    /* The reference's four sample loops as one loop over [lo, hi) whose
     * bounds are the same in every lane (the line window), each sum masked
     * to its own range and conditions: the wavelength loads are then
     * uniform and the samples of all four loops are read once.  Each sum
     * still adds its terms in ascending wavelength order.                  */
    double x = 0, y = 0, x_vlos = 0, y_vlos = 0, xm = 0, ym = 0, xg = 0, yg = 0;
    double sum_u = 0.0, sum_q = 0.0;
    for (i = lo; i < hi; i++) {
        const double dl = lambda[i] - lambda_0;
        const float fI = sI[i], fQ = sQ[i], fU = sU[i], fV = sV[i];
        const bool in = (i >= begin) && (i < end - 1);
        if (in && fI > -1 && fV > -1) {
            const double aux = (Ic - (double)(fI + fV));
            const double aux_vlos = (Ic - fI);
            x += aux * dl;
            x_vlos += aux_vlos * dl;
            y += aux;
            y_vlos += aux_vlos;
            const double auxm = (Ic - (double)(fI - fV));
            xm += auxm * dl;
            ym += auxm;
        }
        if (in && fQ > -1 && fU > -1 && fV > -1) {
            const double L = fabs(sqrtf(fQ * fQ + fU * fU));
            const double mm = fabs((4 * dl * L));
            xg = xg + fabsf(fV) * mm;
            yg = yg + fabsf(fV) * fabsf(fV);
        }
        if (fU > -1 && fQ > -1) {
            if (fabs(fU) > 0.0001 || fabs(fQ) > 0.0001) {
                sum_u += fU;
                sum_q += fQ;
            }
        }
    }
    const double lp = (fabs(y) > 1e-15) ? x / y : 0;
    const double lm = (fabs(ym) > 1e-15) ? xm / ym : 0;
    const double C = (4.6686411e-13 * (lambda_0 * lambda_0) * S->ce_geff);
    const double Blos = (1 / C) * ((lp - lm) / 2);
    double Vlos = (2.99792458e+5 / lambda_0) * ((x_vlos / y_vlos) / 2);
    yg = yg * fabs((3 * C * Blos));
    const double tan_gamma = fabs(sqrt(xg / yg));
    double gamma_rad = atan(tan_gamma);
    double gamma = gamma_rad * (180 / kPI);
    if (initial) {
        if (gamma >= 85 && gamma <= 90) gamma_rad = 85 * (kPI / 180);
        if (gamma > 90 && gamma <= 95) gamma_rad = 95 * (kPI / 180);
    }
    if (Blos < 0) gamma = (180) - gamma;
    const double tan2phi = sum_u / sum_q;
    double phi = (atan(tan2phi) * 180 / kPI) / 2;
    if (sum_u > 0 && sum_q > 0) phi = phi;
    else if (sum_u < 0 && sum_q > 0) phi = phi + 180;
    else if (sum_u < 0 && sum_q < 0) phi = phi + 90;
    else if (sum_u > 0 && sum_q < 0) phi = phi + 90;
    const double B_aux = fabs(Blos / cos(gamma_rad));
    if (Vlos < (-4)) Vlos = -4;
    if (Vlos > (4)) Vlos = 4;
    model[1] = (B_aux > 4000 ? 4000 : B_aux);
    model[2] = Vlos;
    model[5] = gamma;
    model[6] = phi;
    if (!initial) {
        model[7] = Blos;
        return;
    }
    /* NaN guards: milosUtils.cu:621-628 */
    if (isnan(model[1])) model[1] = 1;
    if (isnan(model[2])) model[2] = 1e-3;
    if (isnan(model[5])) model[5] = 1;
    if (isnan(model[6])) model[6] = 1;
}

__global__ __launch_bounds__(64) void k_ce(const gm_dsetup *__restrict__ S,
                                           const float *__restrict__ obs, int64_t npix,
                                           int initial, double *__restrict__ models)
{
    const int n = S->nlambda;
    for (int64_t p = (int64_t)blockIdx.x * 64 + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * 64) {
        double m[GM_NMODEL];
#pragma unroll
        for (int f = 0; f < GM_NMODEL; ++f) m[f] = S->init_model[f];
        classical_estimates(S, obs + p * 4 * n, m, initial);
#pragma unroll
        for (int f = 0; f < GM_NMODEL; ++f) models[p * GM_NMODEL + f] = m[f];
    }
}

/* LM start state of state slots slot0 .. slot0+npix-1 (milosUtils.cu:
 * 595-643): start model (fixed, classical estimates or per pixel), lambda,
 * counters, and the slot list for the first normal-equation evaluation.
 * obs is indexed by slot, init by p.  For the classical estimates the
 * block's ppb consecutive spectra are first copied to LDS with coalesced
 * loads (row stride 4 nlambda + 1 floats against bank conflicts).          */
__global__ __launch_bounds__(64) void k_init_state(const gm_dsetup *__restrict__ S, gm_state st,
                                                   const float *__restrict__ obs,
                                                   const double *__restrict__ init,
                                                   int init_mode, int64_t slot0, int64_t npix,
                                                   int32_t *__restrict__ list0, int ppb)
{
    extern __shared__ float spec[];
    const int n = S->nlambda;
    const int n4 = 4 * n, stride = 4 * n + 1;
    const int64_t cap = st.cap;
    const int tid = threadIdx.x;
    const bool vec4 = (reinterpret_cast<uintptr_t>(obs) & 15) == 0;
    for (int64_t base = (int64_t)blockIdx.x * ppb; base < npix; base += (int64_t)gridDim.x * ppb) {
        const int nb = (int)((npix - base < ppb) ? npix - base : ppb);
        if (init_mode == GM_INIT_CE) {
            const float *src = obs + (slot0 + base) * n4;
            const int total = nb * n4;
            if (vec4) {
                /* float4 loads, eight in flight per lane before their LDS
                 * stores (one at a time, each store waited for its load:
                 * about 60 serial load latencies per block); n4 is a
                 * multiple of 4, so a float4 never straddles two spectra */
                for (int i0 = 4 * tid; i0 < total; i0 += 8 * 256) {
                    float4 v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int i = i0 + 256 * u;
                        if (i < total) v[u] = *reinterpret_cast<const float4 *>(src + i);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int i = i0 + 256 * u;
                        if (i < total) {
                            const int r = i / n4;
                            float *d = spec + r * stride + (i - r * n4);
                            d[0] = v[u].x;
                            d[1] = v[u].y;
                            d[2] = v[u].z;
                            d[3] = v[u].w;
                        }
                    }
                }
            } else {
                for (int i = tid; i < total; i += 64) spec[(i / n4) * stride + (i % n4)] = src[i];
            }
            __syncthreads();
        }
        if (tid < nb) {
            const int64_t p = base + tid;
            const int64_t q = slot0 + p;
            double m[GM_NMODEL];
            if (init_mode == GM_INIT_PER_PIXEL) {
#pragma unroll
                for (int f = 0; f < GM_NMODEL; ++f) m[f] = init[p * GM_NMODEL + f];
                /* the 9-parameter kernels have no macroturbulence or stray
                 * light: such contexts keep the set-up's mac and alfa */
                if (!S->ext) {
                    m[9] = S->init_model[9];
                    m[10] = S->init_model[10];
                }
            } else {
#pragma unroll
                for (int f = 0; f < GM_NMODEL; ++f) m[f] = S->init_model[f];
                if (init_mode == GM_INIT_CE) classical_estimates(S, spec + tid * stride, m, 1);
            }
#pragma unroll
            for (int f = 0; f < GM_NMODEL; ++f) {
                st.best[q * GM_MODEL_REC + f] = m[f];
                st.trial[q * GM_MODEL_REC + f] = m[f];
            }
            store_pre(st.pre_best, q, make_pre(m));
            st.flambda[q] = S->ilambda;
            st.factor[q] = 1.0;
            st.iter[q] = 0;
            st.nacc[q] = 0;
            list0[p] = (int32_t)q;
        }
        if (init_mode == GM_INIT_CE) __syncthreads();
    }
}

__global__ void k_init_chi(gm_state st, int64_t slot0, int64_t npix)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x)
        st.chisqr0[slot0 + p] = st.ochisqr[slot0 + p];
}

/* list[*count + i] = slot0 + i for i < n; k_add_count then advances *count. */
__global__ void k_append_range(int32_t *__restrict__ list, const int32_t *__restrict__ count,
                               int64_t slot0, int64_t n)
{
    const int64_t base = *count;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        list[base + i] = (int32_t)(slot0 + i);
}
__global__ void k_add_count(int32_t *c, int32_t v) { *c += v; }

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

/* Retire pixels that left the LM loop: copy the best model, chi^2,
 * iteration and acceptance counts and the last evaluated spectrum from the
 * state slot to the owning tile's output, one wave per pixel; the last pixel
 * of a tile appends the tile to the completed list (completed[0] = count). */
/* Copies retired pixels into their tiles' outputs (one wave per pixel) and
 * counts them off their tiles: one lane per pixel, and one atomic per wave
 * when the wave's 64 consecutive retired pixels share a tile (the common
 * case: a tile is a contiguous slot range and pixels retire in slot order),
 * so a large tile finishing in one round does not serialise on one counter.
 * The host reads the completed list after the launch, so no fence orders
 * the copies before the counts. */
__global__ __launch_bounds__(256) void k_retire(gm_state st, const int32_t *__restrict__ done,
                                                const int32_t *__restrict__ done_count,
                                                const float *__restrict__ synth, int nl4,
                                                const int32_t *__restrict__ slot_tile,
                                                gm_tile_dev *__restrict__ tiles,
                                                int32_t *__restrict__ completed)
{
    const int64_t cnt = *done_count;
    const int64_t cap = st.cap;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = wave; i < cnt; i += nwaves) {
        const int64_t slot = done[i];
        const gm_tile_dev &g = tiles[slot_tile[slot]];
        const int64_t idx = slot - g.slot0;
        if (g.synth)
            for (int k = lane; k < nl4; k += 64) g.synth[idx * nl4 + k] = synth[slot * nl4 + k];
        if (lane < GM_NMODEL && g.model) g.model[idx * GM_NMODEL + lane] = st.best[slot * GM_MODEL_REC + lane];
        if (lane == 0) {
            if (g.chisqr) g.chisqr[idx] = (float)st.ochisqr[slot];
            if (g.iters) g.iters[idx] = st.iter[slot];
            if (g.naccs) g.naccs[idx] = st.nacc[slot];
        }
    }
    for (int64_t b0 = wave * 64; b0 < cnt; b0 += nwaves * 64) {
        const int n = (cnt - b0 < 64) ? (int)(cnt - b0) : 64;
        const int t = (lane < n) ? slot_tile[done[b0 + lane]] : -1;
        const int t0 = __shfl(t, 0, 64);
        if (__all(lane >= n || t == t0)) {
            if (lane == 0 && atomicSub(&tiles[t0].remaining, n) == n)
                completed[1 + atomicAdd(&completed[0], 1)] = t0;
        } else if (lane < n) {
            if (atomicSub(&tiles[t].remaining, 1) == 1)
                completed[1 + atomicAdd(&completed[0], 1)] = t;
        }
    }
}

__global__ void k_finalize(gm_state st, int64_t npix, double *__restrict__ model,
                           float *__restrict__ chisqr, int32_t *__restrict__ iters,
                           int32_t *__restrict__ naccs)
{
    const int64_t cap = st.cap;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        if (model)
            for (int f = 0; f < GM_NMODEL; ++f) model[p * GM_NMODEL + f] = st.best[p * GM_MODEL_REC + f];
        if (chisqr) chisqr[p] = (float)st.ochisqr[p];
        if (iters) iters[p] = st.iter[p];
        if (naccs) naccs[p] = st.nacc[p];
    }
}

__global__ void k_set_count(int32_t *c, int32_t v) { *c = v; }
__global__ void k_copy_count(const int32_t *src, int32_t *dst) { *dst = *src; }

inline int grid_for(int64_t items, int per_block, int cap_blocks)
{
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap_blocks) g = cap_blocks;
    return (int)g;
}

/* Grid cap of the grid-stride kernels.  The wavelength-parallel kernels
 * hold about 6 000 one-wave workgroups at once; with a static stride, a
 * grid just above that leaves a second, mostly empty wave of workgroups
 * running at low occupancy at the end of every launch.  Four or more
 * workgroups per resident slot even it out: per 1M-pixel inversion
 * 56.2 ms with 8 192, 52.7 with 12 288, 52.0 with 24 576-65 536, 56.0 with
 * one workgroup per pixel pair (262 144). */
constexpr int kMaxBlocks = 32768;

/* Dynamic LDS of k_rf and k_synth: the staged rows of each pixel of the
 * workgroup (or the normal-equation reduction, whichever is larger), plus
 * the macroturbulence kernel rows of the eleven-slot variants.            */
/* Staged rows per pixel and zero rows before wavelength 0 of each kernel:
 * the matrix-core PSF reads the 32 rows of the grid only, the VALU
 * correlation reads rows shifted by every applied tap offset. */
int rf_rows(const gm_lanecfg &lc) { return lc.mfma ? 32 : lc.rows; }
int rf_padlo(const gm_lanecfg &lc) { return lc.mfma ? 0 : lc.padlo; }
int synth_rows(const gm_lanecfg &lc) { return lc.rows; }
int synth_padlo(const gm_lanecfg &lc) { return lc.padlo; }

size_t rf_lds_bytes(const gm_lanecfg &lc)
{
    const int xm = lc.np == 11 ? 2 : (lc.ext ? 1 : 0);
    const int red4 = (lc.lpp / 2) * (xm == 2 ? 20 : 15);
    const int region4 = (rf_rows(lc) * 9 > red4) ? rf_rows(lc) * 9 : red4;
    return (size_t)lc.ppw * region4 * sizeof(float4) +
           (size_t)lc.ppw * (xm == 2 ? 4 : (xm == 1 ? 2 : 0)) * lc.lpp * lc.nlpl * sizeof(float) +
           (lc.mfma ? 64 * sizeof(float) : 0);          /* PSF tap table */
}

size_t synth_lds_bytes(const gm_lanecfg &lc)
{
    return (size_t)lc.ppw * synth_rows(lc) * sizeof(float4) +
           (lc.ext ? (size_t)lc.ppw * 2 * lc.lpp * lc.nlpl * sizeof(float) : 0);
}

template <int LPP, int NLPL>
void launch_rf_t(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                 const float *pre, int64_t mfs, int64_t mps, const float *obs, const int32_t *list,
                 const int32_t *count, int64_t n_direct, int64_t grid_items, float *alpha,
                 float *beta, int64_t aes, int64_t aps, int64_t bes, int64_t bps, double *chi,
                 float *synth, float *rf, int mixed, hipStream_t s, int alpha_first,
                 double *model_dst, float *pre_dst)
{
    const int xm = lc.np == 11 ? 2 : (lc.ext ? 1 : 0);
    const size_t shm = rf_lds_bytes(lc);
    const int rows = rf_rows(lc), padlo = rf_padlo(lc);
    const int g = grid_for(grid_items, lc.ppw, kMaxBlocks);
    /* the LM loop's calls take the specialised kernel (fixed record strides) */
    const int64_t abrec = xm == 2 ? GM_AB_REC11 : GM_AB_REC9;
    const bool rec = mfs == 1 && mps == GM_MODEL_REC && aes == 1 && bes == 1 && aps == abrec &&
                     bps == abrec && pre && list && count && obs && alpha && beta && !rf;
#define GM_RF_CASE(MX, X)                                                                         \
    do {                                                                                          \
        if (rec)                                                                                  \
            hipLaunchKernelGGL((k_rf<LPP, NLPL, MX, X, true>), dim3(g), dim3(64), shm, s, dS,     \
                               models, pre, mfs, mps, obs, list, count, n_direct, alpha, beta,    \
                               aes, aps, bes, bps, chi, synth, rf, rows, padlo, alpha_first,     \
                               model_dst, pre_dst);                                               \
        else                                                                                      \
            hipLaunchKernelGGL((k_rf<LPP, NLPL, MX, X, false>), dim3(g), dim3(64), shm, s, dS,    \
                               models, pre, mfs, mps, obs, list, count, n_direct, alpha, beta,    \
                               aes, aps, bes, bps, chi, synth, rf, rows, padlo, alpha_first,     \
                               model_dst, pre_dst);                                               \
    } while (0)
    if (xm == 2) {
        if (mixed) GM_RF_CASE(true, 2);
        else GM_RF_CASE(false, 2);
    } else if (xm == 1) {
        if (mixed) GM_RF_CASE(true, 1);
        else GM_RF_CASE(false, 1);
    } else {
        if (mixed) GM_RF_CASE(true, 0);
        else GM_RF_CASE(false, 0);
    }
#undef GM_RF_CASE
}

template <int LPP, int NLPL>
void launch_synth_t(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                    const float *pre, int64_t mfs, int64_t mps, const float *obs, const int32_t *list,
                    const int32_t *count, int64_t n_direct, int64_t grid_items, double *chi,
                    float *synth, int mixed, hipStream_t s)
{
    const size_t shm = synth_lds_bytes(lc);
    const int rows = synth_rows(lc), padlo = synth_padlo(lc);
    const int g = grid_for(grid_items, lc.ppw, kMaxBlocks);
    const bool rec = mfs == 1 && mps == GM_MODEL_REC && pre && list && count && obs && chi;
#define GM_SYNTH_CASE(MX, EX)                                                                     \
    do {                                                                                          \
        if (rec)                                                                                  \
            hipLaunchKernelGGL((k_synth<LPP, NLPL, MX, EX, true>), dim3(g), dim3(64), shm, s, dS, \
                               models, pre, mfs, mps, obs, list, count, n_direct, chi, synth,     \
                               rows, padlo);                                                \
        else                                                                                      \
            hipLaunchKernelGGL((k_synth<LPP, NLPL, MX, EX, false>), dim3(g), dim3(64), shm, s,    \
                               dS, models, pre, mfs, mps, obs, list, count, n_direct, chi, synth, \
                               rows, padlo);                                                \
    } while (0)
    if (lc.ext) {
        if (mixed) GM_SYNTH_CASE(true, true);
        else GM_SYNTH_CASE(false, true);
    } else {
        if (mixed) GM_SYNTH_CASE(true, false);
        else GM_SYNTH_CASE(false, false);
    }
#undef GM_SYNTH_CASE
}

} // namespace

size_t gm_lds_bytes(const gm_lanecfg &lc)
{
    const size_t a = rf_lds_bytes(lc), b = synth_lds_bytes(lc);
    return a > b ? a : b;
}

/* Model addressing: the LM state is SoA [11][cap] (field stride cap, pixel
 * stride 1); caller arrays are AoS [npix][11] (field stride 1, pixel 11).   */
static void split_strides(int64_t model_stride, int64_t &mfs, int64_t &mps)
{
    if (model_stride > 0) { mfs = model_stride; mps = 1; }
    else if (model_stride < 0) { mfs = 1; mps = -model_stride; }   /* state records */
    else { mfs = 1; mps = GM_NMODEL; }
}

void gm_launch_rf(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                  const float *pre, int64_t model_stride, const float *obs, const int32_t *list,
                  const int32_t *count, int64_t n_direct, int64_t grid_items, float *alpha,
                  float *beta, int64_t ab_stride, double *chi_out, float *synth_out,
                  float *rf_out, int mixed, hipStream_t st, int alpha_first,
                  double *model_dst, float *pre_dst)
{
    int64_t mfs, mps;
    split_strides(model_stride, mfs, mps);
    /* alpha/beta: SoA over the state capacity (entry stride ab_stride, pixel
     * stride 1), per-pixel records of -ab_stride floats (the LM state: one
     * pixel's normal equations are written and read as contiguous lines,
     * however sparse the active pixels), or, with ab_stride 0, the caller's
     * AoS [npix][45] / [npix][9]
     * ([npix][66] / [npix][11] for eleven slots) */
    const int np = lc.np == 11 ? 11 : 9;
    const int64_t aes = ab_stride > 0 ? ab_stride : 1;
    const int64_t aps = ab_stride > 0 ? 1 : (ab_stride < 0 ? -ab_stride : np * (np + 1) / 2);
    const int64_t bes = ab_stride > 0 ? ab_stride : 1;
    const int64_t bps = ab_stride > 0 ? 1 : (ab_stride < 0 ? -ab_stride : np);
    if (lc.lpp == 32 && lc.nlpl == 1)
        launch_rf_t<32, 1>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items, alpha,
                           beta, aes, aps, bes, bps, chi_out, synth_out, rf_out, mixed, st, alpha_first,
                           model_dst, pre_dst);
    else if (lc.lpp == 64 && lc.nlpl == 1)
        launch_rf_t<64, 1>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items, alpha,
                           beta, aes, aps, bes, bps, chi_out, synth_out, rf_out, mixed, st, alpha_first,
                           model_dst, pre_dst);
    else if (lc.lpp == 64 && lc.nlpl == 2)
        launch_rf_t<64, 2>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items, alpha,
                           beta, aes, aps, bes, bps, chi_out, synth_out, rf_out, mixed, st, alpha_first,
                           model_dst, pre_dst);
    else
        launch_rf_t<64, 4>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items, alpha,
                           beta, aes, aps, bes, bps, chi_out, synth_out, rf_out, mixed, st, alpha_first,
                           model_dst, pre_dst);
}

void gm_launch_synth(const gm_lanecfg &lc, const gm_dsetup *dS, const double *models,
                     const float *pre, int64_t model_stride, const float *obs, const int32_t *list,
                     const int32_t *count, int64_t n_direct, int64_t grid_items,
                     double *chi_out, float *synth_out, int mixed, hipStream_t st)
{
    int64_t mfs, mps;
    split_strides(model_stride, mfs, mps);
    if (lc.lpp == 32 && lc.nlpl == 1)
        launch_synth_t<32, 1>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items,
                              chi_out, synth_out, mixed, st);
    else if (lc.lpp == 64 && lc.nlpl == 1)
        launch_synth_t<64, 1>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items,
                              chi_out, synth_out, mixed, st);
    else if (lc.lpp == 64 && lc.nlpl == 2)
        launch_synth_t<64, 2>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items,
                              chi_out, synth_out, mixed, st);
    else
        launch_synth_t<64, 4>(lc, dS, models, pre, mfs, mps, obs, list, count, n_direct, grid_items,
                              chi_out, synth_out, mixed, st);
}

void gm_launch_solve(const gm_dsetup *dS, gm_state st, const int32_t *list,
                     const int32_t *count, int64_t grid_items, int mixed, int solver,
                     int32_t *reset0, int32_t *reset1, int32_t *hist, int32_t *fb_list,
                     int32_t *fb_count, hipStream_t s, int np, int cordic10)
{
    const int g = grid_for(grid_items, 64, kMaxBlocks);
    if (np == 11) {
        /* eleven slots: FP64 LDL^T, the uncertified systems by the wave
         * Jacobi; with mac fixed the reference runs its 10-term CORDIC solve */
        if (cordic10) {
            hipLaunchKernelGGL((k_solve11<true>), dim3(g), dim3(64), 0, s, dS, st, list, count, reset0,
                               reset1, hist, fb_list, fb_count);
        } else {
            hipLaunchKernelGGL((k_solve11<false>), dim3(g), dim3(64), 0, s, dS, st, list, count, reset0,
                               reset1, hist, fb_list, fb_count);
            hipLaunchKernelGGL((k_solve_svd<true, 11>), dim3(grid_for(grid_items, 1, 1024)), dim3(64), 0, s,
                               dS, st, fb_list, fb_count);
        }
        return;
    }
#define GM_SOLVE_CASE(MX, SV)                                                                    \
    hipLaunchKernelGGL((k_solve<MX, SV>), dim3(g), dim3(64), 0, s, dS, st, list, count, reset0,  \
                       reset1, hist, fb_list, fb_count)
    if (mixed) {
        if (solver == GM_SOLVE_SVD) GM_SOLVE_CASE(true, GM_SOLVE_SVD);
        else if (solver == GM_SOLVE_LDL) GM_SOLVE_CASE(true, GM_SOLVE_LDL);
        else if (solver == GM_SOLVE_CORDIC) GM_SOLVE_CASE(true, GM_SOLVE_CORDIC);
        else GM_SOLVE_CASE(true, GM_SOLVE_JACOBI);
    } else {
        if (solver == GM_SOLVE_SVD) GM_SOLVE_CASE(false, GM_SOLVE_SVD);
        else if (solver == GM_SOLVE_LDL) GM_SOLVE_CASE(false, GM_SOLVE_LDL);
        else if (solver == GM_SOLVE_CORDIC) GM_SOLVE_CASE(false, GM_SOLVE_CORDIC);
        else GM_SOLVE_CASE(false, GM_SOLVE_JACOBI);
    }
#undef GM_SOLVE_CASE
    if (solver == GM_SOLVE_SVD) {
        const int gf = grid_for(grid_items, 1, 1024);
        if (mixed)
            hipLaunchKernelGGL((k_solve_svd<true, 9>), dim3(gf), dim3(64), 0, s, dS, st, fb_list, fb_count);
        else
            hipLaunchKernelGGL((k_solve_svd<false, 9>), dim3(gf), dim3(64), 0, s, dS, st, fb_list, fb_count);
    }
}

void gm_launch_step_solve(const float *alpha, const float *beta, const double *flambda,
                          int64_t n, double *delta, int mixed, int solver, hipStream_t s,
                          int np, int cordic10)
{
    const int g = grid_for(n, 64, kMaxBlocks);
    if (np == 11) {
        hipLaunchKernelGGL(k_step_solve11, dim3(g), dim3(64), 0, s, alpha, beta, flambda, n, delta,
                           cordic10);
        return;
    }
#define GM_STEP_CASE(MX, SV)                                                                     \
    hipLaunchKernelGGL((k_step_solve<MX, SV>), dim3(g), dim3(64), 0, s, alpha, beta, flambda, n, delta)
    if (mixed) {
        if (solver == GM_SOLVE_SVD) GM_STEP_CASE(true, GM_SOLVE_SVD);
        else if (solver == GM_SOLVE_LDL) GM_STEP_CASE(true, GM_SOLVE_LDL);
        else if (solver == GM_SOLVE_CORDIC) GM_STEP_CASE(true, GM_SOLVE_CORDIC);
        else GM_STEP_CASE(true, GM_SOLVE_JACOBI);
    } else {
        if (solver == GM_SOLVE_SVD) GM_STEP_CASE(false, GM_SOLVE_SVD);
        else if (solver == GM_SOLVE_LDL) GM_STEP_CASE(false, GM_SOLVE_LDL);
        else if (solver == GM_SOLVE_CORDIC) GM_STEP_CASE(false, GM_SOLVE_CORDIC);
        else GM_STEP_CASE(false, GM_SOLVE_JACOBI);
    }
#undef GM_STEP_CASE
}

void gm_launch_decide(const gm_dsetup *dS, gm_state st, const int32_t *list,
                      const int32_t *count, int64_t grid_items, int32_t *next_list,
                      int32_t *next_count, int32_t *acc_list, int32_t *acc_count,
                      int32_t *reset_fb, int32_t *done_list, int32_t *done_count, int mixed,
                      hipStream_t s, int defer_copy, int32_t *fb_hist, int small_groups)
{
#define GM_DECIDE_CASE(MX, TH)                                                                         \
    hipLaunchKernelGGL((k_decide<MX, TH>), dim3(grid_for(grid_items, TH, GM_DECIDE_BLOCKS * (GM_DECIDE_THREADS / TH))), \
                       dim3(TH), 0, s, dS, st, list, count, next_list, next_count, acc_list, acc_count,   \
                       reset_fb, done_list, done_count, defer_copy, fb_hist)
    if (small_groups) {
        if (mixed) GM_DECIDE_CASE(true, 256);
        else GM_DECIDE_CASE(false, 256);
    } else {
        if (mixed) GM_DECIDE_CASE(true, GM_DECIDE_THREADS);
        else GM_DECIDE_CASE(false, GM_DECIDE_THREADS);
    }
#undef GM_DECIDE_CASE
}

#ifndef GM_INIT_PPB
#define GM_INIT_PPB 32
#endif
void gm_launch_init_state(const gm_dsetup *dS, gm_state st, const float *obs,
                          const double *init, int init_mode, int nlambda, int64_t slot0,
                          int64_t npix, int32_t *list0, hipStream_t s)
{
    /* up to 32 pixels per block (at most 32 KB of staged spectra): at
     * 30 wavelengths 15.5 KB, ten blocks per CU, whose staging loads
     * overlap (64 pixels, five blocks per CU: 0.23 ms slower per 1M px) */
    const int stride = 4 * nlambda + 1;
    int ppb = 8192 / stride;
    ppb = ppb < 1 ? 1 : (ppb > GM_INIT_PPB ? GM_INIT_PPB : ppb);
    const size_t shm = (init_mode == GM_INIT_CE) ? (size_t)ppb * stride * sizeof(float) : 0;
    const int g = grid_for(npix, ppb, kMaxBlocks);
    hipLaunchKernelGGL(k_init_state, dim3(g), dim3(64), shm, s, dS, st, obs, init, init_mode, slot0,
                       npix, list0, ppb);
}

void gm_launch_init_chi(gm_state st, int64_t slot0, int64_t npix, hipStream_t s)
{
    const int g = grid_for(npix, 256, kMaxBlocks);
    hipLaunchKernelGGL(k_init_chi, dim3(g), dim3(256), 0, s, st, slot0, npix);
}

void gm_launch_append_range(int32_t *list, int32_t *count, int64_t slot0, int64_t n,
                            hipStream_t s)
{
    const int g = grid_for(n, 256, kMaxBlocks);
    hipLaunchKernelGGL(k_append_range, dim3(g), dim3(256), 0, s, list, count, slot0, n);
    hipLaunchKernelGGL(k_add_count, dim3(1), dim3(1), 0, s, count, (int32_t)n);
}

void gm_launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s)
{
    const int g = grid_for(n, 256, kMaxBlocks);
    hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, p, n, v);
}

void gm_launch_retire(gm_state st, const int32_t *done, const int32_t *done_count,
                      int64_t grid_items, const float *synth, int nl4, const int32_t *slot_tile,
                      gm_tile_dev *tiles, int32_t *completed, hipStream_t s)
{
    const int g = grid_for(grid_items, 4, kMaxBlocks);
    hipLaunchKernelGGL(k_retire, dim3(g), dim3(256), 0, s, st, done, done_count, synth, nl4,
                       slot_tile, tiles, completed);
}

void gm_launch_finalize(gm_state st, int64_t npix, const gm_out &out, hipStream_t s)
{
    const int g = grid_for(npix, 256, kMaxBlocks);
    hipLaunchKernelGGL(k_finalize, dim3(g), dim3(256), 0, s, st, npix, out.model, out.chisqr,
                       out.iters, out.naccs);
}

void gm_launch_ce(const gm_dsetup *dS, const float *obs, int64_t npix, int initial,
                  double *models, hipStream_t s)
{
    const int g = grid_for(npix, 64, kMaxBlocks);
    hipLaunchKernelGGL(k_ce, dim3(g), dim3(64), 0, s, dS, obs, npix, initial, models);
}

void gm_launch_set_count(int32_t *count, int32_t value, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, s, count, value);
}

void gm_launch_copy_count(const int32_t *src, int32_t *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, s, src, dst);
}
