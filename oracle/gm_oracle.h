/*
 * gm_oracle.h -- FP64 host restatement of the G-MILOS per-pixel inversion.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline; it is linked only by the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark.  The product path (the HIP library) never
 * calls it.
 *
 * Every routine restates the reference algorithm in double precision and cites
 * the reference file:line it follows.  Multi-line synthesis (more than one
 * spectral line sharing one atmosphere) is an extension the reference lacks;
 * it reduces exactly to the reference formulas for one line with FO = 1.
 */
#ifndef GM_ORACLE_H
#define GM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define GMO_MAXL      256   /* wavelengths per pixel                       */
#define GMO_MAXLINES  4     /* spectral lines sharing one atmosphere       */
#define GMO_MAXCOMP   10    /* Zeeman sub-components per type (pi/b/r)     */
#define GMO_NT        10    /* response-function slots (reference NTERMS)  */
#define GMO_NPARMS    4     /* Stokes I,Q,U,V                              */

/* Model order (defines.h:94-106): eta0,B,vlos,dopp,aa,gm,az,S0,S1,mac,alfa */
enum { GMO_ETA0 = 0, GMO_B, GMO_VLOS, GMO_DOPP, GMO_AA, GMO_GM, GMO_AZ,
       GMO_S0, GMO_S1, GMO_MAC, GMO_ALFA, GMO_NMODEL };

/* Zeeman pattern of one line (defines.h:108-123, create_cuantic.c:31-215). */
typedef struct {
    int    n_pi, n_sig;
    double nub[GMO_MAXCOMP], nup[GMO_MAXCOMP], nur[GMO_MAXCOMP];
    double web[GMO_MAXCOMP], wep[GMO_MAXCOMP], wer[GMO_MAXCOMP];
    double gl, gu, geff, fo;
} gmo_zeeman;

typedef struct {
    int     nlambda;
    double  lambda[GMO_MAXL];        /* absolute wavelengths, Angstrom      */
    int     nlines;
    double  lambda0[GMO_MAXLINES];   /* line centres, Angstrom              */
    gmo_zeeman zee[GMO_MAXLINES];    /* per-line pattern incl. strength fo  */
    int     use_psf;
    double  psf[GMO_MAXL];           /* nlambda taps, centre at nlambda/2   */
    double  weight[4];               /* WeightForStokes                     */
    double  noise;                   /* sigma^2 (readConfig.c:2001-2008)    */
    double  ilambda;                 /* initial Marquardt lambda            */
    double  toplim;
    int     miter;
    double  mu;                      /* ah                                  */
    int     logclambda;
    int     fix[11];
    double  init_model[GMO_NMODEL];
    int     use_ce;                  /* classical estimates overwrite init  */
    int     ce_line;                 /* line used by classical estimates    */
    int     rf_reference_rules;      /* 1: reproduce reference me_der rules  */
    int     use_slight;              /* stray-light profile mixed in          */
    double  slight[4 * GMO_MAXL];    /* stray light, I,Q,U,V blocks of n      */
    int     nterms;                  /* reference NTERMS: 10 (lm_mils, also 0)
                                        or 11 (lm_mils_11: mac and alfa free) */
    int     solver;                  /* GMO_SOLVE_SVD: the pseudo-inverse from a
                                        converged SVD; GMO_SOLVE_CORDIC: the
                                        reference's 18-step CORDIC solve for
                                        the 10-term system                  */
} gmo_setup;

#define GMO_SOLVE_SVD 0
#define GMO_SOLVE_CORDIC 3

/* ---- atomic / PSF set-up ---------------------------------------------- */
int  gmo_cuanten(double sl, double ll, double jl, double su, double lu,
                 double ju, double fos, gmo_zeeman *z);
void gmo_fgauss_wl(double fwhm_mA, double step, double lambda_first,
                   double lambda_central, int n, double *out);
int  gmo_psf_interp(const double *dl, const double *psf, int npsf,
                    const double *offsets_mA, int n, double offset_mA,
                    double *out);

/* ---- per-pixel physics -------------------------------------------------- */
void gmo_voigt(double a, double u, double *H, double *F);
void gmo_conv(double *x, int n, const double *h);
void gmo_conv_ic(double *x, int n, const double *h, double ic);
/* circular convolution centred on n/2 (macroturbulence) */
void gmo_conv_circ(double *x, int n, const double *h);
/* macroturbulence kernel and its mac derivative on the grid, both
 * correlated with the PSF when it is on (Gd may be NULL) */
void gmo_mac_kernels(const gmo_setup *s, double mac, double *G, double *Gd);
/* synthesis: spectra[4*n] (I block, Q, U, V) */
void gmo_synth(const gmo_setup *s, const double *model, double *spectra);
/* synthesis + response functions rf[s*NT*n + p*n + l], NT = nterms (10
 * unless the set-up selects the 11-term layout) */
void gmo_synth_rf(const gmo_setup *s, const double *model, double *spectra,
                  double *rf);
/* as gmo_synth_rf; alpha_rf_mixed = 1 takes the filling-factor response
 * from the mixed spectrum, as the reference's LM loop does after the
 * first evaluation (milosUtils.cu:673) */
void gmo_synth_rf2(const gmo_setup *s, const double *model, double *spectra,
                   double *rf, int alpha_rf_mixed);
void gmo_covarm(const gmo_setup *s, const float *obs, const double *synth,
                const double *rf, double *alpha, double *beta);
double gmo_chisqr(const gmo_setup *s, const float *obs, const double *synth,
                  int nfree);
int  gmo_nterms(const gmo_setup *s);
/* pseudo-inverse solve of a symmetric n x n system via an SVD (n <= 11) */
void gmo_svd_solve(const double *A, const double *b, int n, double *x,
                   double *w_out);
/* the reference's 10-term CORDIC solve: `steps` Brent-Luk steps of the
 * symmetric 10 x 10 A (upper triangle read), final matrix and V (10 x 10,
 * either may be NULL); the solve runs 18 steps */
void gmo_cordic_eig10(const double *A, int steps, double *a_out, double *v_out);
void gmo_cordic_solve10(const double *A, const double *b, double *x, double *w_out);
void gmo_apply_delta(const gmo_setup *s, const double *model,
                     const double *delta, double *out);
void gmo_check(double *model);
void gmo_classical_estimates(const gmo_setup *s, const float *spectro,
                             double *model, int for_initial_use);

/* ---- LM driver ---------------------------------------------------------- */
/* init: optional per-pixel start model (NULL = setup init model + CE rule) */
int gmo_lm_pixel(const gmo_setup *s, const float *obs, const double *init,
                 double *model_out, double *chisqr_out, int *iter_out,
                 float *synth_out, int *nacc_out);
void gmo_invert(const gmo_setup *s, const float *obs, long npix,
                const double *init_models, double *models, float *chisqr,
                int *iters, float *synth, int *naccs, int nthreads);
void gmo_synth_batch(const gmo_setup *s, const double *models, long npix,
                     float *spectra, int nthreads);
int  gmo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
