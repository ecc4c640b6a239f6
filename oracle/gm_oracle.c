/*
 * gm_oracle.c -- FP64 host restatement of the G-MILOS per-pixel inversion.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle and CPU baseline).  Only the
 * test-suite, the smoke check and the benchmark's CPU-baseline leg link it.
 *
 * Each function restates a reference routine in double precision; the
 * reference file:line it follows is cited at the function.  Nothing here is
 * tuned for speed beyond OpenMP over pixels.
 */
#include "gm_oracle.h"

#include <math.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GMO_PI      3.14159265358979323846264338327950288
#define GMO_VLIGHT  2.99792458e+5        /* km/s            defines.h:80 */
#define GMO_CTE     4.6686411e-13        /* Zeeman constant defines.h:82 */
#define GMO_RR      0.5641895836         /* 1/sqrt(pi)      defines.h:78 */

/* ------------------------------------------------------------------------ */
/* Zeeman pattern.  Follows create_cuantic.c:31-215 (Landé factors, effective
 * factor, pi / sigma-blue / sigma-red shifts and strengths normalised per
 * component type).  M values use the reference's integer truncation of J.   */
static double lande_g(double s, double l, double j)
{
    if (j == 0) return 0.0;
    return (3.0 * j * (j + 1.0) + s * (s + 1.0) - l * (l + 1.0)) /
           (2.0 * j * (j + 1.0));
}

int gmo_cuanten(double sl, double ll, double jl, double su, double lu,
                double ju, double fos, gmo_zeeman *z)
{
    memset(z, 0, sizeof(*z));
    double g1 = lande_g(sl, ll, jl);
    double g2 = lande_g(su, lu, ju);
    z->gl = g1;
    z->gu = g2;
    z->geff = (g1 + g2) / 2 + (g1 - g2) * (jl * (jl + 1.) - ju * (ju + 1.)) / 4;
    z->fo = fos;
    int nlo = (int)(2 * jl + 1), nup = (int)(2 * ju + 1);
    int jj = (int)(ju - jl);
    int ip = 0, ib = 0, ir = 0;
    for (int a = 0; a < nlo; a++) {
        double ml = a - (int)jl;
        for (int b = 0; b < nup; b++) {
            double mu = b - (int)ju;
            int dm = (int)(mu - ml);
            double shift = g1 * ml - g2 * mu;
            if (dm == 0) {
                if (ip >= GMO_MAXCOMP) return -1;
                double s = 0;
                if (jj == -1) s = jl * jl - ml * ml;
                else if (jj == 0) s = ml * ml;
                else if (jj == 1) s = (jl + 1) * (jl + 1) - ml * ml;
                z->wep[ip] = s; z->nup[ip] = shift; ip++;
            } else if (dm == 1) {
                if (ib >= GMO_MAXCOMP) return -1;
                double s = 0;
                if (jj == -1) s = (jl - ml) * (jl - ml - 1) / 4;
                else if (jj == 0) s = (jl - ml) * (jl + ml + 1) / 4;
                else if (jj == 1) s = (jl + ml + 1) * (jl + ml + 2) / 4;
                z->web[ib] = s; z->nub[ib] = shift; ib++;
            } else if (dm == -1) {
                if (ir >= GMO_MAXCOMP) return -1;
                double s = 0;
                if (jj == -1) s = (jl + ml) * (jl + ml - 1) / 4;
                else if (jj == 0) s = (jl + ml) * (jl - ml + 1) / 4;
                else if (jj == 1) s = (jl - ml + 1) * (jl - ml + 2) / 4;
                z->wer[ir] = s; z->nur[ir] = shift; ir++;
            }
        }
    }
    z->n_pi = (int)(2 * (jl < ju ? jl : ju)) + 1;
    z->n_sig = (int)(jl + ju);
    double sp = 0, sb = 0, sr = 0;
    for (int i = 0; i < z->n_pi; i++) sp += z->wep[i];
    for (int i = 0; i < z->n_pi; i++) z->wep[i] /= sp;
    for (int i = 0; i < z->n_sig; i++) sb += z->web[i];
    for (int i = 0; i < z->n_sig; i++) z->web[i] /= sb;
    for (int i = 0; i < z->n_sig; i++) sr += z->wer[i];
    for (int i = 0; i < z->n_sig; i++) z->wer[i] /= sr;
    return 0;
}

/* Gaussian instrument profile sampled on the grid.  fgaussC.c:14-87:
 * sigma = FWHM * 0.42466090 / 1000 (mA -> A), exp(-x^2/2) at every sample,
 * normalised to unit sum; centre = lambda_central.                          */
void gmo_fgauss_wl(double fwhm_mA, double step, double lambda_first,
                   double lambda_central, int n, double *out)
{
    double sigma = fwhm_mA * 0.42466090 / 1000.0;
    double sum = 0;
    for (int i = 0; i < n; i++) {
        double x = ((lambda_first + i * step) - lambda_central) / sigma;
        out[i] = exp(-(x * x) / 2);
        sum += out[i];
    }
    for (int i = 0; i < n; i++) out[i] /= sum;
}

/* Tabulated PSF resampled on the grid offsets (mA), shifted by offset_mA,
 * by textbook linear interpolation, then normalised to unit sum.
 * interpolLineal.c:46-98 (offset >= 0 branch) with the offset rule of
 * gmilos.cu:271-288.  Returns -1 where the reference's GSL call would fail
 * (a sample beyond the table's upper end).                                  */
int gmo_psf_interp(const double *dl, const double *psf, int npsf,
                   const double *offsets_mA, int n, double offset_mA,
                   double *out)
{
    for (int i = 0; i < n; i++) {
        double x = offsets_mA[i] - offset_mA;
        out[i] = 0.0;
        if (x < dl[0]) continue;
        if (x > dl[npsf - 1]) return -1;
        int lo = 0, hi = npsf - 1;
        while (hi - lo > 1) {           /* largest lo with dl[lo] <= x */
            int mid = (lo + hi) / 2;
            if (dl[mid] > x) hi = mid; else lo = mid;
        }
        double t = (x - dl[lo]) / (dl[hi] - dl[lo]);
        out[i] = psf[lo] + t * (psf[hi] - psf[lo]);
    }
    double sum = 0;
    for (int i = 0; i < n; i++) sum += out[i];
    for (int i = 0; i < n; i++) out[i] /= sum;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Voigt H and Faraday-Voigt F via the rational approximation of Hui,
 * Armstrong & Wray (1978): w(z) ~ P(Z)/Q(Z), Z = a - i u.  H = Re w,
 * F = Im(w)/2.  The reference evaluates the same rational function with an
 * unrolled Horner chain (mil_sinrf.cu:172-332; coefficients
 * definesCuda.cuh:88-110).                                                  */
static const double HUI_P[7] = {
    122.607931777104326, 214.382388694706425, 181.928533092181549,
    93.155580458138441, 30.180142196210589, 5.912626209773153,
    0.564189583562615};
static const double HUI_Q[8] = {
    122.60793177387535, 352.730625110963558, 457.334478783897737,
    348.703917719495792, 170.354001821091472, 53.992906912940207,
    10.479857114260399, 1.0};

void gmo_voigt(double a, double u, double *H, double *F)
{
    /* Z = a - i u */
    double zr = a, zi = -u;
    double pr = HUI_P[6], pi = 0.0;
    for (int k = 5; k >= 0; k--) {
        double tr = pr * zr - pi * zi + HUI_P[k];
        double ti = pr * zi + pi * zr;
        pr = tr; pi = ti;
    }
    double qr = HUI_Q[7], qi = 0.0;
    for (int k = 6; k >= 0; k--) {
        double tr = qr * zr - qi * zi + HUI_Q[k];
        double ti = qr * zi + qi * zr;
        qr = tr; qi = ti;
    }
    double den = qr * qr + qi * qi;
    *H = (pr * qr + pi * qi) / den;
    *F = 0.5 * (pi * qr - pr * qi) / den;
}

/* ------------------------------------------------------------------------ */
/* PSF correlation on the grid with zero padding: y_k = sum_j h_j x_{k+j-n/2}
 * over j = 0..n-1 (convolution.cu:153-198, dirConvPar zero-filled at
 * milosUtils.cu:890-891).  The _ic variant convolves (Ic - x) and returns
 * Ic - y (convolution.cu:200-241).                                          */
void gmo_conv(double *x, int n, const double *h)
{
    double tmp[GMO_MAXL];
    int c = n / 2;
    for (int k = 0; k < n; k++) {
        double acc = 0;
        for (int j = 0; j < n; j++) {
            int m = k + j - c;
            if (m >= 0 && m < n) acc += h[j] * x[m];
        }
        tmp[k] = acc;
    }
    memcpy(x, tmp, sizeof(double) * n);
}

void gmo_conv_ic(double *x, int n, const double *h, double ic)
{
    double tmp[GMO_MAXL];
    for (int k = 0; k < n; k++) tmp[k] = ic - x[k];
    gmo_conv(tmp, n, h);
    for (int k = 0; k < n; k++) x[k] = ic - tmp[k];
}

/* Circular convolution centred on n/2 (d_convCircular / convCircular,
 * convolution.cu:90-118, 251-280): the reference forms
 * a_i = sum_j h_j x_{(i-j) mod n} and stores a_i at index (i - n/2) mod n,
 * i.e. y_m = sum_j h_j x_{(m + n/2 - j) mod n}.                            */
void gmo_conv_circ(double *x, int n, const double *h)
{
    double tmp[GMO_MAXL];
    const int c = n / 2;
    for (int m = 0; m < n; m++) {
        double acc = 0;
        for (int j = 0; j < n; j++) {
            int k = (m + c - j) % n;
            if (k < 0) k += n;
            acc += h[j] * x[k];
        }
        tmp[m] = acc;
    }
    memcpy(x, tmp, sizeof(double) * n);
}

/* Macroturbulence kernel (fgauss.cu:16-63): a Gaussian of width
 * sigma = lambda0 * mac / c (mac in km/s) on the grid, centred on
 * lambda[n/2] and normalised to unit sum.  Its mac derivative follows
 * me_der.cu:124-131, G (x^2 - 1) / mac, which drops the derivative of the
 * normalisation; with rf_reference_rules = 0 it is the exact derivative of
 * the normalised kernel, G (x^2 - sum_j G_j x_j^2) / mac.  With the PSF on
 * both are correlated with it (mil_sinrf.cu:104-107, me_der.cu:132-138).    */
void gmo_mac_kernels(const gmo_setup *s, double mac, double *G, double *Gd)
{
    const int n = s->nlambda;
    const double ild = (s->lambda0[0] * mac) / 2.99792458e5;
    double x2[GMO_MAXL], sum = 0;
    for (int i = 0; i < n; i++) {
        const double x = (s->lambda[i] - s->lambda[n / 2]) / ild;
        x2[i] = x * x;
        const double term = x2[i] / 2;
        G[i] = (term < 1e30) ? exp(-term) : 0;
        sum += G[i];
    }
    for (int i = 0; i < n; i++) G[i] /= sum;
    if (Gd) {
        double m2 = 1.0;
        if (!s->rf_reference_rules) {
            m2 = 0;
            for (int i = 0; i < n; i++) m2 += G[i] * x2[i];
        }
        for (int i = 0; i < n; i++) Gd[i] = G[i] / mac * (x2[i] - m2);
        if (s->use_psf) gmo_conv(Gd, n, s->psf);
    }
    if (s->use_psf) gmo_conv(G, n, s->psf);
}

/* The continuum used by the Ic variant: the larger end sample
 * (mil_sinrf.cu:118-124, me_der.cu:170-175).                                */
static double end_max(const double *x, int n)
{
    return (x[0] > x[n - 1]) ? x[0] : x[n - 1];
}

/* ------------------------------------------------------------------------ */
/* Synthesis and response functions.
 * Synthesis follows mil_sinrf.cu:27-142 and the per-wavelength ME solution
 * mil_sinrf.cu:334-433.  Response functions follow me_der.cu:35-77 (Voigt
 * derivatives), me_der.cu:81-215 (driver + PSF, Ic variant on dI) and
 * me_der.cu:218-376 (chain rule through the ME solution).  Angles are in
 * degrees, so gamma/phi derivatives are per degree.
 *
 * Multi-line extension: every line contributes FO_l-weighted profiles to the
 * same absorption matrix; with one line and FO = 1 the formulas are the
 * reference's.  For lines with several sub-components per type the Voigt
 * derivatives are summed over sub-components (the synthesis sums them too).
 *
 * rf_reference_rules = 1 reproduces the reference's magneto-optical Voigt
 * derivative scale (me_der.cu:71-74 store w*dF where shi = 2*w*F, i.e. half
 * the derivative of shi); 0 gives the exact derivative.                     */
/* NTERMS of the reference kernel the set-up corresponds to: lm_mils_11
 * (11) runs when macroturbulence and filling factor are both free
 * (gmilos.cu:1234-1239), lm_mils (10) otherwise.                           */
int gmo_nterms(const gmo_setup *s)
{
    return s->nterms == 11 ? 11 : 10;
}

static void synth_core(const gmo_setup *s, const double *m, double *spectra,
                       double *rf, int alpha_rf_mixed)
{
    const int n = s->nlambda;
    const double CC = GMO_PI / 180.0;
    const double CC_2 = CC * 2;
    double sin_gm = sin(m[GMO_GM] * CC), cosi = cos(m[GMO_GM] * CC);
    double sinis = sin_gm * sin_gm;
    double cosis = cosi * cosi;
    double cosis_2 = (1 + cosis) / 2;
    double sina = sin(2 * m[GMO_AZ] * CC), cosa = cos(2 * m[GMO_AZ] * CC);
    double sinda = cosa * CC_2, cosda = -sina * CC_2;
    double sindi = cosi * sin_gm * CC_2, cosdi = -sin_gm * CC;
    double sinis_cosa = sinis * cosa, sinis_sina = sinis * sina;
    double E0 = m[GMO_ETA0];
    double E0_2 = E0 / 2.0;
    double dopp = m[GMO_DOPP], aa = m[GMO_AA], B = m[GMO_B];
    double ah = s->mu;
    double S0 = m[GMO_S0], S1 = m[GMO_S1];
    double shi_scale = s->rf_reference_rules ? 1.0 : 2.0;

    double I[GMO_MAXL] = {0}, Q[GMO_MAXL] = {0}, U[GMO_MAXL] = {0}, V[GMO_MAXL] = {0};
    const int NT = gmo_nterms(s);

    for (int l = 0; l < n; l++) {
        /* FO-weighted profiles per component type t = 0 pi, 1 blue, 2 red */
        double fi[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
        double dfi[3][4], dshi[3][4];
        memset(dfi, 0, sizeof dfi);
        memset(dshi, 0, sizeof dshi);
        for (int L = 0; L < s->nlines; L++) {
            const gmo_zeeman *z = &s->zee[L];
            double l0 = s->lambda0[L];
            double fo = z->fo;
            double ulos = (m[GMO_VLOS] * l0) / (GMO_VLIGHT * dopp);
            double parcial = ((l0 * l0) / dopp) * GMO_CTE;
            double auxCte = -l0 / (GMO_VLIGHT * dopp);
            for (int t = 0; t < 3; t++) {
                int nc = (t == 0) ? z->n_pi : z->n_sig;
                for (int i = 0; i < nc; i++) {
                    double nu, w;
                    if (t == 0) { nu = parcial * z->nup[i]; w = z->wep[i]; }
                    else if (t == 1) { nu = parcial * z->nub[i]; w = z->web[i]; }
                    else { nu = -parcial * z->nub[z->n_sig - (i + 1)]; w = z->wer[i]; }
                    double u = (s->lambda[l] - l0) / dopp - ulos - nu * B;
                    double H, F;
                    gmo_voigt(aa, u, &H, &F);
                    fi[t] += fo * w * H;
                    shi[t] += fo * w * F * 2;
                    if (rf) {
                        double dH_u = ((4 * aa * F) - (2 * u * H)) * w * fo;
                        double dF_u = (GMO_RR - aa * H - 2 * u * F) * w * fo;
                        double uu = -u / dopp;
                        dfi[t][0] += dH_u * (-nu);
                        dfi[t][1] += dH_u * auxCte;
                        dfi[t][2] += dH_u * uu;
                        dfi[t][3] += -2.0 * dF_u;
                        dshi[t][0] += shi_scale * dF_u * (-nu);
                        dshi[t][1] += shi_scale * dF_u * auxCte;
                        dshi[t][2] += shi_scale * dF_u * uu;
                        dshi[t][3] += shi_scale * dH_u / 2.0;
                    }
                }
            }
        }
        /* ME solution, mil_sinrf.cu:346-393 */
        double parcial1 = fi[1] + fi[2];
        double parcial2 = E0_2 * (fi[0] - parcial1 / 2);
        double parcial3 = E0_2 * (shi[0] - (shi[1] + shi[2]) / 2);
        double etain = E0_2 * (fi[0] * sinis + parcial1 * cosis_2);
        double etaq = parcial2 * sinis_cosa;
        double etau = parcial2 * sinis_sina;
        double etav = (fi[2] - fi[1]) * (E0_2 * cosi);
        double rhoq = parcial3 * sinis_cosa;
        double rhou = parcial3 * sinis_sina;
        double rhov = (shi[2] - shi[1]) * (E0_2 * cosi);
        double etai = 1.0 + etain;
        double etai_2 = etai * etai;
        double auxq = rhoq * rhoq, auxu = rhou * rhou, auxv = rhov * rhov;
        double gp1 = etai_2 - etaq * etaq - etau * etau - etav * etav + auxq + auxu + auxv;
        double gp3 = etai_2 + auxq + auxu + auxv;
        double gp2 = etaq * rhoq + etau * rhou + etav * rhov;
        double dt = etai_2 * gp1 - gp2 * gp2;
        double dti = 1.0 / dt;
        double gp4 = etai_2 * etaq + etai * (etav * rhou - etau * rhov);
        double gp5 = etai_2 * etau + etai * (etaq * rhov - etav * rhoq);
        double gp6 = etai_2 * etav + etai * (etau * rhoq - etaq * rhou);
        double g4 = gp4 + rhoq * gp2;
        double g5 = gp5 + rhou * gp2;
        double g6 = gp6 + rhov * gp2;
        double dtiaux = dti * (-(S1 * ah));
        I[l] = S0 - dtiaux * etai * gp3;
        Q[l] = dtiaux * g4;
        U[l] = dtiaux * g5;
        V[l] = dtiaux * g6;
        if (!rf) continue;

        /* d(eta,rho)/dp for p = eta0,B,vlos,dopp,a,gamma,phi: me_der.cu:254-312 */
        double d_ei[7], d_eq[7], d_eu[7], d_ev[7], d_rq[7], d_ru[7], d_rv[7];
        memset(d_ei, 0, sizeof d_ei); memset(d_eq, 0, sizeof d_eq);
        memset(d_eu, 0, sizeof d_eu); memset(d_ev, 0, sizeof d_ev);
        memset(d_rq, 0, sizeof d_rq); memset(d_ru, 0, sizeof d_ru);
        memset(d_rv, 0, sizeof d_rv);
        d_ei[0] = etain / E0; d_eq[0] = etaq / E0; d_eu[0] = etau / E0;
        d_ev[0] = etav / E0; d_rq[0] = rhoq / E0; d_ru[0] = rhou / E0;
        d_rv[0] = rhov / E0;
        double sinis_E0_2 = sinis * E0_2, cosis_2_E0_2 = cosis_2 * E0_2;
        double sc = E0 * (sinis * cosa) / 2, ss = E0 * (sinis * sina) / 2;
        double cosi_2_E0 = E0 * cosi / 2.0;
        for (int j = 1; j < 5; j++) {
            double dfisum = dfi[1][j - 1] + dfi[2][j - 1];
            d_ei[j] = dfi[0][j - 1] * sinis_E0_2 + dfisum * cosis_2_E0_2;
            double aux = dfi[0][j - 1] - dfisum / 2;
            d_eq[j] = aux * sc;
            d_eu[j] = aux * ss;
            d_ev[j] = (dfi[2][j - 1] - dfi[1][j - 1]) * cosi_2_E0;
            double aux2 = dshi[0][j - 1] - (dshi[1][j - 1] + dshi[2][j - 1]) / 2.0;
            d_rq[j] = aux2 * sc;
            d_ru[j] = aux2 * ss;
            d_rv[j] = (dshi[2][j - 1] - dshi[1][j - 1]) * cosi_2_E0;
        }
        double cosi_cosdi = cosi * cosdi * E0_2, sindi_E0_2 = sindi * E0_2;
        double cosdi_E0_2 = E0_2 * cosdi;
        d_ei[5] = fi[0] * sindi_E0_2 + parcial1 * cosi_cosdi;
        d_eq[5] = parcial2 * (sindi * cosa);
        d_eu[5] = parcial2 * (sindi * sina);
        d_ev[5] = (fi[2] - fi[1]) * cosdi_E0_2;
        d_rq[5] = parcial3 * (sindi * cosa);
        d_ru[5] = parcial3 * (sindi * sina);
        d_rv[5] = (shi[2] - shi[1]) * cosdi_E0_2;
        d_eq[6] = parcial2 * (sinis * cosda);
        d_eu[6] = parcial2 * (sinis * sinda);
        d_rq[6] = parcial3 * (sinis * cosda);
        d_ru[6] = parcial3 * (sinis * sinda);

        /* chain rule through the ME solution, me_der.cu:317-373 */
        double dtaux = (-(S1 * ah)) / (dt * dt);
        double etai_gp3 = etai * gp3;
        double a2 = 2 * etai;
        double ext1 = a2 * etaq + etav * rhou - etau * rhov;
        double ext2 = a2 * etau + etaq * rhov - etav * rhoq;
        double ext3 = a2 * etav + etau * rhoq - etaq * rhou;
        double ext4 = a2 * gp1;
        for (int p = 0; p < 7; p++) {
            double ei = d_ei[p], eq = d_eq[p], eu = d_eu[p], ev = d_ev[p];
            double rq = d_rq[p], ru = d_ru[p], rv = d_rv[p];
            double dgp1 = 2.0 * (etai * ei - etaq * eq - etau * eu - etav * ev +
                                 rhoq * rq + rhou * ru + rhov * rv);
            double dgp2 = rhoq * eq + etaq * rq + rhou * eu + etau * ru + rhov * ev + etav * rv;
            double d_dt = ext4 * ei + etai_2 * dgp1 - 2 * gp2 * dgp2;
            double dgp3 = 2.0 * (etai * ei + rhoq * rq + rhou * ru + rhov * rv);
            double dgp4 = ei * ext1 + etai_2 * eq + etai * (rhou * ev + etav * ru - rhov * eu - etau * rv);
            double dgp5 = ei * ext2 + etai_2 * eu + etai * (rhov * eq + etaq * rv - rhoq * ev - etav * rq);
            double dgp6 = ei * ext3 + etai_2 * ev + etai * (rhoq * eu + etau * rq - rhou * eq - etaq * ru);
            rf[0 * NT * n + p * n + l] = -(((ei * gp3 + etai * dgp3) * dt - d_dt * etai_gp3) * dtaux);
            rf[1 * NT * n + p * n + l] = ((dgp4 + rq * gp2 + rhoq * dgp2) * dt - d_dt * g4) * dtaux;
            rf[2 * NT * n + p * n + l] = ((dgp5 + ru * gp2 + rhou * dgp2) * dt - d_dt * g5) * dtaux;
            rf[3 * NT * n + p * n + l] = ((dgp6 + rv * gp2 + rhov * dgp2) * dt - d_dt * g6) * dtaux;
        }
        double dti2 = -(dti * ah);
        rf[0 * NT * n + 8 * n + l] = -dti2 * etai_gp3;
        rf[1 * NT * n + 8 * n + l] = dti2 * g4;
        rf[2 * NT * n + 8 * n + l] = dti2 * g5;
        rf[3 * NT * n + 8 * n + l] = dti2 * g6;
        rf[0 * NT * n + 7 * n + l] = 1;
        rf[1 * NT * n + 7 * n + l] = 0;
        rf[2 * NT * n + 7 * n + l] = 0;
        rf[3 * NT * n + 7 * n + l] = 0;
        rf[0 * NT * n + 6 * n + l] = 0;
        rf[3 * NT * n + 6 * n + l] = 0;
        for (int st = 0; st < 4; st++)
            for (int p = 9; p < NT; p++) rf[st * NT * n + p * n + l] = 0;
    }

    /* macroturbulence (mac > 1e-4): circular convolution with the Gaussian
     * (PSF-correlated when the PSF is on) instead of the PSF, no Ic rule;
     * slot 9 holds d/dmac, the unconvolved spectra convolved with the
     * kernel's derivative (mil_sinrf.cu:101-116, me_der.cu:119-156)        */
    const double mac = m[GMO_MAC];
    if (mac > 0.0001) {
        double G[GMO_MAXL], Gd[GMO_MAXL];
        gmo_mac_kernels(s, mac, G, rf ? Gd : NULL);
        if (rf) {
            const double *src[4] = {I, Q, U, V};
            for (int st = 0; st < 4; st++) {
                double *x = rf + st * NT * n + 9 * n;
                memcpy(x, src[st], sizeof(double) * n);
                gmo_conv_circ(x, n, Gd);
                for (int p = 0; p < 9; p++)
                    if (p != 7) gmo_conv_circ(rf + st * NT * n + p * n, n, G);
            }
        }
        gmo_conv_circ(I, n, G);
        gmo_conv_circ(Q, n, G);
        gmo_conv_circ(U, n, G);
        gmo_conv_circ(V, n, G);
    } else if (s->use_psf) {
        /* instrument PSF: mil_sinrf.cu:118-130 and me_der.cu:162-189 */
        gmo_conv_ic(I, n, s->psf, end_max(I, n));
        gmo_conv(Q, n, s->psf);
        gmo_conv(U, n, s->psf);
        gmo_conv(V, n, s->psf);
        if (rf) {
            for (int p = 0; p < NT; p++) {
                if (p == 7) continue;
                double *x = rf + 0 * NT * n + p * n;
                gmo_conv_ic(x, n, s->psf, end_max(x, n));
            }
            for (int st = 1; st < 4; st++)
                for (int p = 0; p < NT; p++) {
                    if (p == 7) continue;
                    gmo_conv(rf + st * NT * n + p * n, n, s->psf);
                }
        }
    }
    for (int l = 0; l < n; l++) {
        spectra[l] = I[l];
        spectra[n + l] = Q[l];
        spectra[2 * n + l] = U[l];
        spectra[3 * n + l] = V[l];
    }
    /* stray light: I = alfa I_mag + (1 - alfa) I_stray for every Stokes
     * parameter (mil_sinrf.cu:133-140); every response is scaled by alfa and
     * the last slot (9 for NTERMS 10, 10 for 11) becomes d/dalfa =
     * I_mag - I_stray (me_der.cu:191-211).  The
     * reference's LM loop passes the mixed spectrum as "I_mag" after the
     * first evaluation (milosUtils.cu:673), which gives alfa (I_mag - I_stray);
     * alpha_rf_mixed reproduces that.                                       */
    if (s->use_slight) {
        const double al = m[GMO_ALFA];
        double pure[4 * GMO_MAXL];
        memcpy(pure, spectra, sizeof(double) * 4 * n);
        for (int k = 0; k < 4 * n; k++) spectra[k] = pure[k] * al + s->slight[k] * (1.0 - al);
        if (rf) {
            const double *base = alpha_rf_mixed ? spectra : pure;
            for (int st = 0; st < 4; st++)
                for (int p = 0; p < NT; p++)
                    for (int l = 0; l < n; l++) {
                        double *r = rf + st * NT * n + p * n + l;
                        *r = (p == NT - 1) ? base[st * n + l] - s->slight[st * n + l] : *r * al;
                    }
        }
    }
    /* zero the RFs of fixed parameters: milosUtils.cu:243-271 (NTERMS=10:
     * slot 9 follows the filling-factor flag fix[10]; 11: slot p follows
     * fix[p]).                                                              */
    if (rf) {
        for (int p = 0; p < NT; p++) {
            int free_p = (p < 9 || NT == 11) ? s->fix[p] : s->fix[10];
            if (free_p) continue;
            for (int st = 0; st < 4; st++)
                for (int l = 0; l < n; l++) rf[st * NT * n + p * n + l] = 0;
        }
    }
}

void gmo_synth(const gmo_setup *s, const double *model, double *spectra)
{
    synth_core(s, model, spectra, NULL, 0);
}

void gmo_synth_rf(const gmo_setup *s, const double *model, double *spectra,
                  double *rf)
{
    synth_core(s, model, spectra, rf, 0);
}

void gmo_synth_rf2(const gmo_setup *s, const double *model, double *spectra,
                   double *rf, int alpha_rf_mixed)
{
    synth_core(s, model, spectra, rf, alpha_rf_mixed);
}

/* Normal equations.  lib.cu:12-59 / 65-118: alpha = sum_s (w_s/sigma^2)
 * J_s^T J_s, beta = sum_s (w_s/sigma^2) J_s^T (synth - obs).                */
void gmo_covarm(const gmo_setup *s, const float *obs, const double *synth,
                const double *rf, double *alpha, double *beta)
{
    const int n = s->nlambda, NT = gmo_nterms(s);
    double AP[4][11 * 11], BT[4][11];
    for (int j = 0; j < 4; j++) {
        double w = s->weight[j], wd = s->weight[j] / s->noise;
        for (int i = 0; i < NT; i++)
            for (int h = 0; h < NT; h++) {
                double sum = 0, sum2 = 0;
                for (int k = 0; k < n; k++) {
                    double dh = rf[j * n * NT + h * n + k];
                    sum += rf[j * n * NT + i * n + k] * dh;
                    if (i == 0)
                        sum2 += (w * (synth[k + n * j] - (double)obs[k + n * j])) * dh;
                }
                AP[j][NT * i + h] = sum * wd;
                if (i == 0) BT[j][h] = sum2 / s->noise;
            }
    }
    for (int i = 0; i < NT; i++) beta[i] = BT[0][i] + BT[1][i] + BT[2][i] + BT[3][i];
    for (int i = 0; i < NT * NT; i++) alpha[i] = AP[0][i] + AP[1][i] + AP[2][i] + AP[3][i];
}

/* lib.cu:120-155: chi^2 = sum_s w_s sum_l (synth-obs)^2 / sigma^2 / nfree  */
double gmo_chisqr(const gmo_setup *s, const float *obs, const double *synth,
                  int nfree)
{
    const int n = s->nlambda;
    double opa[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 4; j++) {
            double d = synth[i + n * j] - (double)obs[i + n * j];
            opa[j] += ((d * d) * s->weight[j]) / s->noise;
        }
    return (opa[0] + opa[1] + opa[2] + opa[3]) / nfree;
}

/* ------------------------------------------------------------------------ */
/* Step solve.  The reference forms delta = V diag(1/w_i, |w_i| > 1e-12) V^T
 * beta from an SVD of the damped normal matrix (milosUtils.cu:280-325 with
 * svdcmp.cu:37-243).  The damped matrix is symmetric, so its SVD is its
 * eigen-decomposition with w = |eigenvalue|; this oracle obtains it with the
 * cyclic Jacobi method run to convergence in FP64 (Golub & Van Loan,
 * "Matrix Computations", sec. 8.5), then applies the same pseudo-inverse.  */
void gmo_svd_solve(const double *Ain, const double *b, int n, double *x,
                   double *w_out)
{
    double A[11][11], Vm[11][11];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            /* use the upper triangle; the reference matrix is symmetric */
            A[i][j] = (i <= j) ? Ain[i * n + j] : Ain[j * n + i];
            Vm[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p][q] * A[p][q];
        if (off == 0.0) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) /
                           (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0);
                double sn = t * c;
                for (int r = 0; r < n; r++) {
                    double arp = A[r][p], arq = A[r][q];
                    A[r][p] = c * arp - sn * arq;
                    A[r][q] = sn * arp + c * arq;
                }
                for (int r = 0; r < n; r++) {
                    double apr = A[p][r], aqr = A[q][r];
                    A[p][r] = c * apr - sn * aqr;
                    A[q][r] = sn * apr + c * aqr;
                }
                A[p][q] = A[q][p] = 0.0;
                for (int r = 0; r < n; r++) {
                    double vrp = Vm[r][p], vrq = Vm[r][q];
                    Vm[r][p] = c * vrp - sn * vrq;
                    Vm[r][q] = sn * vrp + c * vrq;
                }
            }
    }
    double aux[11];
    for (int j = 0; j < n; j++) {
        double sum = 0;
        for (int k = 0; k < n; k++) sum += b[k] * Vm[k][j];
        double w = fabs(A[j][j]);
        if (w_out) w_out[j] = w;
        aux[j] = sum * ((w > 1e-12) ? (1.0 / w) : 0.0);
    }
    for (int i = 0; i < n; i++) {
        double sum = 0;
        for (int k = 0; k < n; k++) sum += Vm[i][k] * aux[k];
        x[i] = sum;
    }
}

/* The reference's 10-term solve (mil_svdf, milosUtils.cu:328-365, over
 * svdcordicf, svdcordic.cu:423-537; NUM_ITER_SVD_CORDIC = 18,
 * definesCuda.cuh:119), restated in FP64.  A Brent-Luk parallel Jacobi
 * eigen-decomposition: each step takes the five diagonal 2 x 2 blocks of
 * index pairs (2k, 2k+1), their angles theta_k = atan(2 a12/(a22 - a11))/2
 * (pi/4 when a11 == a22, clamped to +-pi/4; svdcordic.cu:83-114), rotates
 * every block (i, j) as G_i^T B G_j -- rows first, then columns, with
 * G = [[c, s], [-s, c]] (ROTACIONf, svdcordic.cu:33-50; PDf, PNDf) -- zeroes
 * the diagonal blocks' off-diagonal entry, rotates V's column pairs by G
 * (PVf), and then moves index i to CP[i] in both the matrix and V (the
 * a_out positions of svdcordic.cu:441-500 and cordicPosFila/Col,
 * gmilos.cu:108-128).  CP is a 9-cycle, so after 9k steps V A_k V^T = A.
 * a_out / v_out (10 x 10, may be NULL) receive the final matrix and V. */
static const int CORDIC_CP[10] = {0, 2, 4, 1, 6, 3, 8, 5, 9, 7};

void gmo_cordic_eig10(const double *Ain, int steps, double *a_out, double *v_out)
{
    double A[10][10], V[10][10], An[10][10], Vn[10][10];
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++) {
            A[i][j] = (i <= j) ? Ain[i * 10 + j] : Ain[j * 10 + i];
            V[i][j] = (i == j) ? 1.0 : 0.0;
        }
    const double qpi = GMO_PI / 4;
    for (int st = 0; st < steps; st++) {
        double c[5], sn[5];
        for (int k = 0; k < 5; k++) {
            const double a11 = A[2 * k][2 * k], a12 = A[2 * k][2 * k + 1];
            const double a22 = A[2 * k + 1][2 * k + 1];
            double th = qpi;
            if (a11 != a22) {
                th = atan(2 * a12 / (a22 - a11)) / 2;
                if (th > qpi) th = qpi;
                if (th < -qpi) th = -qpi;
            }
            c[k] = cos(th);
            sn[k] = sin(th);
        }
        for (int bi = 0; bi < 5; bi++)
            for (int bj = bi; bj < 5; bj++) {
                const int r = 2 * bi, q = 2 * bj;
                /* rows of the block by G_bi^T */
                const double o11 = A[r][q] * c[bi] - A[r + 1][q] * sn[bi];
                const double o21 = A[r][q] * sn[bi] + A[r + 1][q] * c[bi];
                const double o12 = A[r][q + 1] * c[bi] - A[r + 1][q + 1] * sn[bi];
                const double o22 = A[r][q + 1] * sn[bi] + A[r + 1][q + 1] * c[bi];
                /* then its columns by G_bj */
                double m[2][2];
                m[0][0] = o11 * c[bj] - o12 * sn[bj];
                m[0][1] = o11 * sn[bj] + o12 * c[bj];
                m[1][0] = o21 * c[bj] - o22 * sn[bj];
                m[1][1] = o21 * sn[bj] + o22 * c[bj];
                if (bi == bj) m[0][1] = m[1][0] = 0.0;
                for (int u = 0; u < 2; u++)
                    for (int w = 0; w < 2; w++) {
                        const int P = CORDIC_CP[r + u], Q = CORDIC_CP[q + w];
                        An[P][Q] = An[Q][P] = m[u][w];
                    }
            }
        for (int f = 0; f < 10; f++)
            for (int k = 0; k < 5; k++) {
                const double v0 = V[f][2 * k], v1 = V[f][2 * k + 1];
                Vn[CORDIC_CP[f]][CORDIC_CP[2 * k]] = v0 * c[k] - v1 * sn[k];
                Vn[CORDIC_CP[f]][CORDIC_CP[2 * k + 1]] = v0 * sn[k] + v1 * c[k];
            }
        memcpy(A, An, sizeof A);
        memcpy(V, Vn, sizeof V);
    }
    if (a_out) memcpy(a_out, A, sizeof A);
    if (v_out) memcpy(v_out, V, sizeof V);
}

/* delta = V diag(1/w, |w| > 1e-12) V^T b with w = diag of the 18-step
 * matrix, the reference's order of summation (milosUtils.cu:337-364). */
void gmo_cordic_solve10(const double *A, const double *b, double *x, double *w_out)
{
    double a[100], v[100], aux[10];
    gmo_cordic_eig10(A, 18, a, v);
    for (int j = 0; j < 10; j++) {
        double sum = 0;
        for (int k = 0; k < 10; k++) sum += b[k] * v[k * 10 + j];
        const double w = a[j * 10 + j];
        if (w_out) w_out[j] = w;
        aux[j] = sum * ((fabs(w) > 1e-12) ? (1.0 / w) : 0.0);
    }
    for (int i = 0; i < 10; i++) {
        double sum = 0;
        for (int k = 0; k < 10; k++) sum += v[i * 10 + k] * aux[k];
        x[i] = sum;
    }
}

/* Apply the step to the free parameters with the reference's clamps on the
 * step: |dB| <= 300 G, |dgamma|,|dphi| <= 30 deg (milosUtils.cu:93-159,
 * NTERMS=10 rule: slot 9 updates the filling factor when it is free).      */
void gmo_apply_delta(const gmo_setup *s, const double *model,
                     const double *delta, double *out)
{
    const int *fix = s->fix;
    double aux;
    if (fix[0]) out[GMO_ETA0] = model[GMO_ETA0] - delta[0];
    if (fix[1]) {
        aux = delta[1];
        if (aux < -300) aux = -300; else if (aux > 300) aux = 300;
        out[GMO_B] = model[GMO_B] - aux;
    }
    if (fix[2]) out[GMO_VLOS] = model[GMO_VLOS] - delta[2];
    if (fix[3]) out[GMO_DOPP] = model[GMO_DOPP] - delta[3];
    if (fix[4]) out[GMO_AA] = model[GMO_AA] - delta[4];
    if (fix[5]) {
        aux = delta[5];
        if (aux < -30) aux = -30; else if (aux > 30) aux = 30;
        out[GMO_GM] = model[GMO_GM] - aux;
    }
    if (fix[6]) {
        aux = delta[6];
        if (aux < -30) aux = -30; else if (aux > 30) aux = 30;
        out[GMO_AZ] = model[GMO_AZ] - aux;
    }
    if (fix[7]) out[GMO_S0] = model[GMO_S0] - delta[7];
    if (fix[8]) out[GMO_S1] = model[GMO_S1] - delta[8];
    if (fix[9]) out[GMO_MAC] = model[GMO_MAC] - delta[9];
    /* AplicaDelta (11 terms, milosUtils.cu:30-91) / AplicaDeltaf (10,
     * milosUtils.cu:131-138) */
    if (fix[10]) out[GMO_ALFA] = model[GMO_ALFA] - delta[gmo_nterms(s) == 11 ? 10 : 9];
}

/* Physical ranges, milosUtils.cu:160-238. */
void gmo_check(double *m)
{
    if (m[GMO_B] < 0) { m[GMO_B] = -m[GMO_B]; m[GMO_GM] = 180.0 - m[GMO_GM]; }
    if (m[GMO_B] > 5000) m[GMO_B] = 5000;
    if (m[GMO_GM] < 0) m[GMO_GM] = -m[GMO_GM];
    if (m[GMO_GM] > 180) m[GMO_GM] = 360.0 - m[GMO_GM];
    if (m[GMO_AZ] < 0) m[GMO_AZ] = 180 + m[GMO_AZ];
    if (m[GMO_AZ] > 180) m[GMO_AZ] = m[GMO_AZ] - 180.0;
    if (m[GMO_ETA0] < 1) m[GMO_ETA0] = 1;
    if (m[GMO_ETA0] > 2500) m[GMO_ETA0] = 2500;
    if (m[GMO_VLOS] < -20) m[GMO_VLOS] = -20;
    if (m[GMO_VLOS] > 20) m[GMO_VLOS] = 20;
    if (m[GMO_DOPP] < 0.0001) m[GMO_DOPP] = 0.0001;
    if (m[GMO_DOPP] > 0.6) m[GMO_DOPP] = 0.6;
    if (m[GMO_AA] < 0.0001) m[GMO_AA] = 0.0001;
    if (m[GMO_AA] > 10.0) m[GMO_AA] = 10.0;
    if (m[GMO_S0] < 0.0001) m[GMO_S0] = 0.0001;
    if (m[GMO_S0] > 2.00) m[GMO_S0] = 2.00;
    if (m[GMO_S1] < 0.0001) m[GMO_S1] = 0.0001;
    if (m[GMO_S1] > 2.00) m[GMO_S1] = 2.00;
    if (m[GMO_MAC] < 0) m[GMO_MAC] = 0;
    if (m[GMO_MAC] > 4) m[GMO_MAC] = 4;
    if (m[GMO_ALFA] < 0) m[GMO_ALFA] = 0.0;
    if (m[GMO_ALFA] > 1.0) m[GMO_ALFA] = 1.0;
}

/* Wavelength range used by the classical estimates: the whole grid for one
 * line; with several lines, the samples nearest to the chosen line.        */
static void ce_range(const gmo_setup *s, int *lo, int *hi)
{
    *lo = 0; *hi = s->nlambda;
    if (s->nlines <= 1) return;
    int L = s->ce_line, first = -1, last = -1;
    for (int l = 0; l < s->nlambda; l++) {
        int best = 0;
        for (int k = 1; k < s->nlines; k++)
            if (fabs(s->lambda[l] - s->lambda0[k]) < fabs(s->lambda[l] - s->lambda0[best])) best = k;
        if (best == L) { if (first < 0) first = l; last = l; }
    }
    if (first >= 0) { *lo = first; *hi = last + 1; }
}

/* Classical (weak-field / centre-of-gravity) estimates of B, vlos, gamma and
 * phi.  milosUtils.cu:386-563, including the leading/trailing I<0 "ghost"
 * sample trimming, the 85/95 degree inclination guard and the azimuth
 * quadrant table.  spectro is float (the observation), arithmetic double.  */
void gmo_classical_estimates(const gmo_setup *s, const float *spectro,
                             double *model, int for_initial_use)
{
    const int n = s->nlambda;
    int lo, hi;
    ce_range(s, &lo, &hi);
    const double lambda_0 = s->lambda0[s->ce_line];
    const float *sI = spectro, *sQ = spectro + n, *sU = spectro + 2 * n, *sV = spectro + 3 * n;
    const double *lambda = s->lambda;
    int i;
    int begin = lo;
    for (i = lo; i < hi; i++) { if (sI[i] < 0) begin++; else break; }
    int end = hi;
    for (i = hi - 1; i >= lo; i--) { if (sI[i] < 0) end--; else break; }
    /* all-negative I: the reference reads I[-1] (milosUtils.cu:426); the
     * oracle and the engine both use the first sample instead */
    double Ic = (end > lo) ? sI[end - 1] : sI[lo];

    double x = 0, y = 0, x_vlos = 0, y_vlos = 0, aux, aux_vlos;
    for (i = begin; i < end - 1; i++) {
        if (sI[i] > -1 && sV[i] > -1) {
            aux = (Ic - (double)(sI[i] + sV[i]));
            aux_vlos = (Ic - sI[i]);
            x += aux * (lambda[i] - lambda_0);
            x_vlos += aux_vlos * (lambda[i] - lambda_0);
            y += aux;
            y_vlos += aux_vlos;
        }
    }
    double lp = (fabs(y) > 1e-15) ? x / y : 0;
    x = 0; y = 0;
    for (i = begin; i < end - 1; i++) {
        if (sI[i] > -1 && sV[i] > -1) {
            aux = (Ic - (double)(sI[i] - sV[i]));
            x += aux * (lambda[i] - lambda_0);
            y += aux;
        }
    }
    double lm = (fabs(y) > 1e-15) ? x / y : 0;
    double C = (GMO_CTE * (lambda_0 * lambda_0) * s->zee[s->ce_line].geff);
    double Blos = (1 / C) * ((lp - lm) / 2);
    double Vlos = (GMO_VLIGHT / lambda_0) * ((x_vlos / y_vlos) / 2);

    x = 0; y = 0;
    for (i = begin; i < end - 1; i++) {
        if (sQ[i] > -1 && sU[i] > -1 && sV[i] > -1) {
            double L = fabs(sqrtf(sQ[i] * sQ[i] + sU[i] * sU[i]));
            double mm = fabs((4 * (lambda[i] - lambda_0) * L));
            x = x + fabsf(sV[i]) * mm;
            y = y + fabsf(sV[i]) * fabsf(sV[i]);
        }
    }
    y = y * fabs((3 * C * Blos));
    double tan_gamma = fabs(sqrt(x / y));
    double gamma_rad = atan(tan_gamma);
    double gamma = gamma_rad * (180 / GMO_PI);
    if (for_initial_use) {
        if (gamma >= 85 && gamma <= 90) gamma_rad = 85 * (GMO_PI / 180);
        if (gamma > 90 && gamma <= 95) gamma_rad = 95 * (GMO_PI / 180);
    }
    if (Blos < 0) gamma = (180) - gamma;

    double sum_u = 0.0, sum_q = 0.0;
    for (i = lo; i < hi; i++) {
        if (sU[i] > -1 && sQ[i] > -1) {
            if (fabs(sU[i]) > 0.0001 || fabs(sQ[i]) > 0.0001) {
                sum_u += sU[i];
                sum_q += sQ[i];
            }
        }
    }
    double tan2phi = sum_u / sum_q;
    double phi = (atan(tan2phi) * 180 / GMO_PI) / 2;
    if (sum_u > 0 && sum_q > 0) phi = phi;
    else if (sum_u < 0 && sum_q > 0) phi = phi + 180;
    else if (sum_u < 0 && sum_q < 0) phi = phi + 90;
    else if (sum_u > 0 && sum_q < 0) phi = phi + 90;

    double B_aux = fabs(Blos / cos(gamma_rad));
    if (Vlos < (-4)) Vlos = -4;
    if (Vlos > (4)) Vlos = 4;
    model[GMO_B] = (B_aux > 4000 ? 4000 : B_aux);
    model[GMO_VLOS] = Vlos;
    model[GMO_GM] = gamma;
    model[GMO_AZ] = phi;
    if (!for_initial_use) model[GMO_S0] = Blos;
}

/* ------------------------------------------------------------------------ */
/* Per-pixel Levenberg-Marquardt, milosUtils.cu:574-710 (lm_mils).  The
 * synthetic output is the last evaluated (possibly rejected) model's
 * spectrum; the returned model and chi^2 are the best accepted ones.        */
int gmo_lm_pixel(const gmo_setup *s, const float *obs, const double *init,
                 double *model_out, double *chisqr_out, int *iter_out,
                 float *synth_out, int *nacc_out)
{
    const int n = s->nlambda, NT = gmo_nterms(s);
    const int nfree = (n * GMO_NPARMS) - NT;
    double best[GMO_NMODEL], model[GMO_NMODEL];
    double synth[4 * GMO_MAXL];
    double *rf = (double *)malloc(sizeof(double) * 4 * NT * n);
    double alpha[11 * 11], covar[11 * 11], beta[11], delta[11];
    if (!rf) return -1;

    if (init) {
        memcpy(best, init, sizeof best);
    } else {
        memcpy(best, s->init_model, sizeof best);
        if (s->use_ce) {
            gmo_classical_estimates(s, obs, best, 1);
            if (isnan(best[GMO_B])) best[GMO_B] = 1;
            if (isnan(best[GMO_VLOS])) best[GMO_VLOS] = 1e-3;
            if (isnan(best[GMO_GM])) best[GMO_GM] = 1;
            if (isnan(best[GMO_AZ])) best[GMO_AZ] = 1;
        }
    }
    gmo_synth_rf(s, best, synth, rf);
    gmo_covarm(s, obs, synth, rf, alpha, beta);
    double ochisqr = gmo_chisqr(s, obs, synth, nfree);
    double chisqr0 = ochisqr, chisqr;
    memcpy(model, best, sizeof model);
    double flambda = s->ilambda, factor = 1.0;
    int iter = 0, clanda = 0, nacc = 0;
    do {
        memcpy(covar, alpha, sizeof(double) * NT * NT);
        for (int j = 0; j < NT; j++) covar[j * (NT + 1)] = alpha[j * (NT + 1)] * (1.0 + flambda);
        if (s->solver == GMO_SOLVE_CORDIC && NT == 10)
            gmo_cordic_solve10(covar, beta, delta, NULL);
        else
            gmo_svd_solve(covar, beta, NT, delta, NULL);
        gmo_apply_delta(s, best, delta, model);
        gmo_check(model);
        gmo_synth(s, model, synth);
        if (synth_out)
            for (int k = 0; k < 4 * n; k++) synth_out[k] = (float)synth[k];
        chisqr = gmo_chisqr(s, obs, synth, nfree);
        if ((fabs((ochisqr - chisqr) * 100 / chisqr) < s->toplim) || (chisqr < 0.0001))
            clanda = 1;
        if (chisqr - ochisqr < 0.) {
            flambda = flambda / (5.0 * factor);
            memcpy(best, model, sizeof best);
            gmo_synth_rf2(s, best, synth, rf, 1);
            gmo_covarm(s, obs, synth, rf, alpha, beta);
            ochisqr = chisqr;
            nacc++;
        } else {
            flambda = flambda * 10.0 * factor;
        }
        if ((flambda > 1e+7) || (flambda < 1e-25)) clanda = 1;
        iter++;
        if (s->logclambda) factor = log10(chisqr) / log10(chisqr0);
    } while (iter < s->miter && !clanda);

    memcpy(model_out, best, sizeof best);
    *chisqr_out = ochisqr;
    *iter_out = iter;
    if (nacc_out) *nacc_out = nacc;
    free(rf);
    return 0;
}

void gmo_invert(const gmo_setup *s, const float *obs, long npix,
                const double *init_models, double *models, float *chisqr,
                int *iters, float *synth, int *naccs, int nthreads)
{
    const int n = s->nlambda;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
    for (long p = 0; p < npix; p++) {
        double c2;
        int it, na;
        gmo_lm_pixel(s, obs + p * 4 * n,
                     init_models ? init_models + p * GMO_NMODEL : NULL,
                     models + p * GMO_NMODEL, &c2, &it,
                     synth ? synth + p * 4 * n : NULL, &na);
        chisqr[p] = (float)c2;
        iters[p] = it;
        if (naccs) naccs[p] = na;
    }
}

void gmo_synth_batch(const gmo_setup *s, const double *models, long npix,
                     float *spectra, int nthreads)
{
    const int n = s->nlambda;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long p = 0; p < npix; p++) {
        double sp[4 * GMO_MAXL];
        gmo_synth(s, models + p * GMO_NMODEL, sp);
        for (int k = 0; k < 4 * n; k++) spectra[p * 4 * n + k] = (float)sp[k];
    }
}

int gmo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
