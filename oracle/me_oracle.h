/*
 * me_oracle.h -- CPU oracle for the Milne-Eddington per-pixel inversion.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference's per-pixel path (P-MILOS built with REAL=double), written to
 * check the GPU library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as
 * the timed CPU baseline.  The product (libpmilos_hip.so) never links it.
 *
 * Every function cites the reference file:line it follows.  Pinning status
 * of each function (what checks it besides itself) is listed in me_oracle.c.
 */
#ifndef ME_ORACLE_H
#define ME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXL 512   /* max wavelength samples */
#define ORC_MAXC 64    /* max Zeeman components per group */
#define ORC_NPAR 11    /* Init_Model parameters (defines.h:86-98) */

/* Init_Model parameter order (defines.h:86-98). */
enum { ORC_ETA0 = 0, ORC_B, ORC_VLOS, ORC_DOPP, ORC_AA, ORC_GM, ORC_AZ,
       ORC_S0, ORC_S1, ORC_MAC, ORC_ALFA };

/* Zeeman pattern of one line (Cuantic, defines.h:100-115). */
typedef struct {
    int n_pi, n_sig;
    double nup[ORC_MAXC], wep[ORC_MAXC];   /* pi: shifts, strengths */
    double nub[ORC_MAXC], web[ORC_MAXC];   /* blue sigma */
    double nur[ORC_MAXC], wer[ORC_MAXC];   /* red sigma */
    double gl, gu, geff, fo;
} orc_zeeman_t;

/* Everything lm_mils reads besides the pixel (milosUtils.h:54-57 plus the
 * globals NTERMS, G, cuantic, CC listed at milos.c:39-100). */
typedef struct {
    int nlambda;
    int nterms;                 /* NTERMS after milos.c:154-157 */
    double lambda0;             /* wlines[1] */
    double lambda[ORC_MAXL];    /* wavelength axis, Angstrom */
    int use_psf;                /* ConvolveWithPSF */
    double G[ORC_MAXL];         /* PSF kernel, length nlambda */
    orc_zeeman_t z;
    int fix[ORC_NPAR];
    double weight[4];
    double noise;               /* per-point variance 1/SN^2 */
    double ilambda;             /* initial LM damping */
    double toplim;
    int miter;
    double ah;                  /* mu */
    int logclambda;
    int use_slight;             /* a stray-light profile is present */
    double slight[4 * ORC_MAXL];
    /* Test switch, 0 in every reference-parity use.  The reference's
     * magneto-optic derivatives (me_der.c:586-600) are those of sum w*F
     * although shi sums w*2F (mil_sinrf.c:440): with 1 the oracle doubles
     * them so the response functions become the exact derivatives of its
     * own synthesis, which finite differences can then pin. */
    int exact_rho_deriv;
    /* useFFT (readConfig.c:995-999): the PSF enters by circular convolution
     * (the DFT product of mil_sinrf.c:301-386, me_der.c:331-470) instead of
     * the zero-padded correlation with the continuum trick.  Evaluated here
     * as the direct circular sum, which the product equals in exact
     * arithmetic. */
    int use_fft;
} orc_setup_t;

/* ---- setup (run once) ---- */
int  orc_zeeman(double sl, double ll, double jl, double su, double lu,
                double ju, orc_zeeman_t *z);
void orc_gauss_psf(double fwhm_mA, double step, double lambda_start,
                   double lambda_centre, int n, double *G);
int  orc_nterms(const int fix[ORC_NPAR], double guess_mac);

/* ---- primitives ---- */
void orc_fvoigt(double damp, const double *v, int n, double *H, double *F);
void orc_gauss_mac(double mc, const double *lambda, int n, double lambda_line,
                   int deriv, double *out);
void orc_conv_direct(double *x, int nx, const double *h, int nh);
void orc_conv_direct_ic(double *x, int nx, const double *h, int nh, double ic);
void orc_conv_circular(const double *x, const double *h, int n, double *out);

/* ---- synthesis + response functions (milos.c:601-602 synthesis mode) ----
 * spectra[4*L]; rf[4*NT*L] (nullable) in d_spectra layout
 * rf[l + L*param + L*NT*stokes]. */
void orc_synth(const orc_setup_t *s, const double model[ORC_NPAR],
               double *spectra, double *rf);

/* ---- normal equations, chi2, solve, update ---- */
void   orc_covarm(const orc_setup_t *s, const double *sig, const float *obs,
                  const double *spectra, const double *rf,
                  double *beta, double *alpha);
double orc_chisqr(const orc_setup_t *s, const double *spectra,
                  const float *obs, const double *sig, double nfree);
void   orc_pinv_solve(int nt, const double *h, const double *beta,
                      double *delta);
void   orc_eigen_sym(int n, const double *a, double *w, double *v);
void   orc_aplica_delta(const orc_setup_t *s, const double cur[ORC_NPAR],
                        double *delta, double out[ORC_NPAR]);
void   orc_check(double m[ORC_NPAR]);
void   orc_fija_cero(const orc_setup_t *s, double *rf);

/* ---- per-pixel ---- */
void orc_classical(const orc_setup_t *s, const float *spectro,
                   double model[ORC_NPAR]);
/* for_init = 0: the estimates-only mode (B_los in S0, no inclination nudge) */
void orc_classical_ex(const orc_setup_t *s, const float *spectro,
                      double model[ORC_NPAR], int for_init);
/* lm_mils: model is in/out.  fit (4L, nullable) receives the last trial
 * synthesis.  Returns 1, or -1 when nfree == 0 (milosUtils.c:509-512). */
int  orc_lm(const orc_setup_t *s, const float *spectro, double model[ORC_NPAR],
            double *fit, float *chi2, int *iters, int *accepted);
/* The FITS pixel loop of milos.c:958-1009: guess, classical estimates, NaN
 * guards, lm_mils.  models[npix*11], chi2/iters/accepted[npix], fit nullable
 * [npix*4L] (float, as written to the profile cube).  Pixels [p0, p1). */
void orc_invert_range(const orc_setup_t *s, const float *stokes, int p0, int p1,
                      const double guess[ORC_NPAR], double *models, float *chi2,
                      int *iters, int *accepted, float *fit);

#ifdef __cplusplus
}
#endif
#endif
