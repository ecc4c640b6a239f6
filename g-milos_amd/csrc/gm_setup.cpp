/*
 * gm_setup.cpp -- host-side set-up: Zeeman patterns, LINES records, PSF
 * kernels and reference defaults (no device work).
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gmilos.h"

/* Landé factor of a level; zero for J = 0 (create_cuantic.c:46-54). */
static double lande(double s, double l, double j)
{
    if (j == 0) return 0.0;
    return (3.0 * j * (j + 1.0) + s * (s + 1.0) - l * (l + 1.0)) / (2.0 * j * (j + 1.0));
}

/* Zeeman components of a transition lower (sl,ll,jl) -> upper (su,lu,ju):
 * shifts g_lo*M_lo - g_up*M_up and relative strengths of the pi (dM = 0),
 * sigma-blue (dM = +1) and sigma-red (dM = -1) components, each type
 * normalised to unit sum.  Same construction as create_cuantic.c:31-215,
 * including its integer truncation of J when listing M values.            */
extern "C" gm_status gm_zeeman_from_quantum(double sl, double ll, double jl, double su,
                                            double lu, double ju, double fo, gm_zeeman *z)
{
    if (!z || jl < 0 || ju < 0) return GM_ERR_INVALID;
    std::memset(z, 0, sizeof *z);
    const double g1 = lande(sl, ll, jl), g2 = lande(su, lu, ju);
    z->gl = g1;
    z->gu = g2;
    z->geff = (g1 + g2) / 2 + (g1 - g2) * (jl * (jl + 1.) - ju * (ju + 1.)) / 4;
    z->fo = fo;
    const int nlo = (int)(2 * jl + 1), nup = (int)(2 * ju + 1);
    const int dj = (int)(ju - jl);
    int ip = 0, ib = 0, ir = 0;
    for (int a = 0; a < nlo; ++a) {
        const double ml = a - (int)jl;
        for (int b = 0; b < nup; ++b) {
            const double mu = b - (int)ju;
            const int dm = (int)(mu - ml);
            const double shift = g1 * ml - g2 * mu;
            double w = 0;
            if (dm == 0) {
                if (ip >= GM_MAX_COMP) return GM_ERR_UNSUPPORTED;
                w = dj == -1 ? jl * jl - ml * ml : dj == 0 ? ml * ml : (jl + 1) * (jl + 1) - ml * ml;
                z->wep[ip] = w;
                z->nup[ip++] = shift;
            } else if (dm == 1) {
                if (ib >= GM_MAX_COMP) return GM_ERR_UNSUPPORTED;
                w = dj == -1 ? (jl - ml) * (jl - ml - 1) / 4
                  : dj == 0 ? (jl - ml) * (jl + ml + 1) / 4 : (jl + ml + 1) * (jl + ml + 2) / 4;
                z->web[ib] = w;
                z->nub[ib++] = shift;
            } else if (dm == -1) {
                if (ir >= GM_MAX_COMP) return GM_ERR_UNSUPPORTED;
                w = dj == -1 ? (jl + ml) * (jl + ml - 1) / 4
                  : dj == 0 ? (jl + ml) * (jl - ml + 1) / 4 : (jl - ml + 1) * (jl - ml + 2) / 4;
                z->wer[ir] = w;
                z->nur[ir++] = shift;
            }
        }
    }
    z->n_pi = (int)(2 * (jl < ju ? jl : ju)) + 1;
    z->n_sig = (int)(jl + ju);
    if (z->n_pi > GM_MAX_COMP || z->n_sig > GM_MAX_COMP) return GM_ERR_UNSUPPORTED;
    double sp = 0, sb = 0, sr = 0;
    for (int i = 0; i < z->n_pi; ++i) sp += z->wep[i];
    for (int i = 0; i < z->n_sig; ++i) { sb += z->web[i]; sr += z->wer[i]; }
    for (int i = 0; i < z->n_pi; ++i) z->wep[i] /= sp;
    for (int i = 0; i < z->n_sig; ++i) { z->web[i] /= sb; z->wer[i] /= sr; }
    return GM_OK;
}

/* Orbital quantum number of a spectroscopic term letter (readConfig.c:410-461). */
static int term_l(char c, double *l)
{
    static const char *letters = "SPDFGHJ";
    const char *p = std::strchr(letters, c);
    if (!p || !c) return 0;
    *l = (double)(p - letters);
    return 1;
}

/* One LINES record: "idx=ATOM ion lambda0 ... -loggf 2S+1 L J- 2S+1 L J".
 * Same field layout as readConfig.c:408; S = (multiplicity - 1) / 2 with
 * integer division as there (readConfig.c:462-463).                        */
extern "C" gm_status gm_parse_lines_file(const char *path, int32_t line_index, double *lambda0,
                                         double qn[6])
{
    if (!path || !lambda0 || !qn) return GM_ERR_INVALID;
    FILE *fp = std::fopen(path, "r");
    if (!fp) return GM_ERR_NOT_FOUND;
    char buf[1024];
    gm_status st = GM_ERR_NOT_FOUND;
    while (std::fgets(buf, sizeof buf, fp)) {
        int idx = -1, ion = 0, mlo = 0, mup = 0;
        char atom[64] = {0};
        double lam = 0, damp = 0, pot = 0, loggf = 0, jlo = 0, jup = 0;
        char clo = 0, cup = 0;
        const int got = std::sscanf(buf, "%i=%63s %i %lf %lf %lf -%lf %i%c %lf- %i%c %lf", &idx,
                                    atom, &ion, &lam, &damp, &pot, &loggf, &mlo, &clo, &jlo, &mup,
                                    &cup, &jup);
        if (got < 1 || idx != line_index) continue;
        double llo = 0, lup = 0;
        if (got < 13 || !term_l(clo, &llo) || !term_l(cup, &lup)) {
            st = GM_ERR_INVALID;
            break;
        }
        *lambda0 = lam;
        qn[0] = (double)((mlo - 1) / 2);
        qn[1] = llo;
        qn[2] = jlo;
        qn[3] = (double)((mup - 1) / 2);
        qn[4] = lup;
        qn[5] = jup;
        st = GM_OK;
        break;
    }
    std::fclose(fp);
    return st;
}

/* Gaussian PSF on the grid, fgaussC.c:14-87 as called at gmilos.cu:244-245:
 * sigma = FWHM * 0.42466090 / 1000 A, centre lambda[n/2], unit sum.        */
extern "C" gm_status gm_psf_gaussian(double fwhm_mA, const double *lambda, int32_t n, double *psf)
{
    if (!lambda || !psf || n < 2 || !(fwhm_mA > 0)) return GM_ERR_INVALID;
    const double sigma = fwhm_mA * 0.42466090 / 1000.0;
    const double step = lambda[1] - lambda[0];
    const double centre = lambda[n / 2];
    double sum = 0;
    for (int i = 0; i < n; ++i) {
        const double x = ((lambda[0] + i * step) - centre) / sigma;
        psf[i] = std::exp(-(x * x) / 2);
        sum += psf[i];
    }
    for (int i = 0; i < n; ++i) psf[i] /= sum;
    return GM_OK;
}

/* Tabulated PSF onto the grid (gmilos.cu:247-291, interpolLineal.c:46-98):
 * the grid sample whose offset truncates to 0 is moved to index n/2 by an
 * offset of (n/2 - pos) * step, the table is linearly interpolated at
 * (offset_i - offset) and the result normalised to unit sum.  A table that
 * does not span the grid is rejected as in gmilos.cu:272-275.             */
extern "C" gm_status gm_psf_from_table(const double *dl, const double *val, int32_t ntab,
                                       const double *off, int32_t n, double step_mA, double *psf)
{
    if (!dl || !val || !off || !psf || ntab < 2 || n < 2) return GM_ERR_INVALID;
    if (std::trunc(off[0]) < std::trunc(dl[0]) || std::trunc(off[n - 1]) > std::trunc(dl[ntab - 1]))
        return GM_ERR_INVALID;
    int pos = 0;
    for (int i = 0; i < n && !pos; ++i)
        if (std::fabs(std::trunc(off[i])) == 0) pos = i;
    double offset = 0;
    if (pos != n / 2) offset = (((n / 2) - pos) * (step_mA / 1000.0)) * 1000;
    for (int i = 0; i < n; ++i) {
        const double x = off[i] - offset;
        psf[i] = 0.0;
        if (x < dl[0]) continue;
        if (x > dl[ntab - 1]) return GM_ERR_INVALID;
        int lo = 0, hi = ntab - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (dl[mid] > x) hi = mid; else lo = mid;
        }
        const double t = (x - dl[lo]) / (dl[hi] - dl[lo]);
        psf[i] = val[lo] + t * (val[hi] - val[lo]);
    }
    double sum = 0;
    for (int i = 0; i < n; ++i) sum += psf[i];
    if (!(sum != 0)) return GM_ERR_INVALID;
    for (int i = 0; i < n; ++i) psf[i] /= sum;
    return GM_OK;
}

/* Reference defaults, readConfig.c:704-744 (noise = NOISE_SIGMA^2,
 * ilambda 0.1, toplim 1e-12, mu 1, unit weights, nothing free). */
extern "C" void gm_setup_defaults(gm_setup *s)
{
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    for (int i = 0; i < 4; ++i) s->weight[i] = 1.0;
    s->noise = 1e-3 * 1e-3;
    s->ilambda = 0.1;
    s->toplim = 1e-12;
    s->mu = 1.0;
    s->miter = 50;
    s->init_model[10] = 1.0;
    s->init_mode = GM_INIT_CE;
    s->precision = GM_PREC_MIXED;
    s->solver = GM_SOLVE_SVD;
    s->rf_reference_rules = 1;
    s->psf_tail_tol = 1e-9;
}
