/*
 * me_oracle.c -- FP64 CPU oracle of the Milne-Eddington per-pixel inversion.
 *
 * TEST INFRASTRUCTURE ONLY (see me_oracle.h).  A restatement of P-MILOS
 * (reference src/ C files, REAL = PRECISION = double) in re-entrant C: the
 * reference's ~70 global scratch vectors become one per-call workspace.
 *
 * Pinning (what checks each function other than itself; tests/ hold them):
 *   orc_zeeman        reference create_cuantic.c built in oracle/_ref; closed-form
 *                     Lande factors of the 5P1-5D0 triplet.
 *   orc_gauss_psf     reference fgauss.c (fgauss_WL) in oracle/_ref.
 *   orc_gauss_mac     reference fgauss.c (fgauss) in oracle/_ref.
 *   orc_conv_*        reference convolution.c in oracle/_ref; numpy correlate/roll.
 *   orc_fvoigt        reference fvoigt.c in oracle/_ref; scipy wofz to the
 *                     rational fit's own error.
 *   orc_synth (I..V)  reference golden test/synthesis_results.per; Stokes
 *                     symmetry identities.
 *   orc_synth (rf)    central finite differences of orc_synth; closed forms for
 *                     the S0, S1 and eta0 rows.
 *   orc_covarm/chisqr numpy sums.
 *   orc_pinv_solve    numpy eigh/pinv with the same 1e-12 eigenvalue cutoff.
 *   orc_aplica_delta, orc_check, orc_fija_cero  table-driven special cases.
 *   orc_classical     weak-field/centre-of-gravity closed forms on shifted
 *                     analytic profiles.
 *   orc_lm            invariants: exact-data exit after one trip, recovery of
 *                     the true model from a perturbed guess.
 */
#include "me_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The arithmetic type of what the reference declares REAL (defines.h:8-36):
 * double by default (USE_DOUBLE_PRECISION); float when built with
 * -DORC_REAL=float, the reference's default build (makefile:21-25), in which
 * PRECISION quantities (the model, lambda, the PSF and macroturbulence
 * kernels, the convolution sums, the solve) stay double and the REAL
 * trigonometric macros are the float functions. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
#define ORC_REAL_IS_FLOAT (sizeof(real) == sizeof(float))
static inline real r_sin(real x) { return ORC_REAL_IS_FLOAT ? (real)sinf((float)x) : (real)sin(x); }
static inline real r_cos(real x) { return ORC_REAL_IS_FLOAT ? (real)cosf((float)x) : (real)cos(x); }
static inline real r_fabs(real x) { return ORC_REAL_IS_FLOAT ? (real)fabsf((float)x) : (real)fabs(x); }

/* Constants, defines.h:66-78. */
#define ORC_PI 3.14159265358979323846264338327950288419716939937510
#define ORC_RR 0.5641895836
#define ORC_VLIGHT 2.99792458e+5
#define ORC_CTE4_6_13 4.6686411e-13

/* ------------------------------------------------------------------------ */
/* Setup                                                                     */
/* ------------------------------------------------------------------------ */

/* Zeeman components of one line: Lande factors, M sublevels, component
 * strengths and shifts, normalisation.  Follows create_cuantic.c:60-234
 * (Cuanten), including the (int) truncation of J when building M. */
int orc_zeeman(double sl, double ll, double jl, double su, double lu,
               double ju, orc_zeeman_t *z)
{
    double g1 = 0.0, g2 = 0.0;
    if (jl != 0)                                             /* :71-74 */
        g1 = (3.0 * jl * (jl + 1) + sl * (sl + 1) - ll * (ll + 1)) /
             (2 * jl * (jl + 1));
    if (ju != 0)                                             /* :76-79 */
        g2 = (3.0 * ju * (ju + 1.0) + su * (su + 1.0) - lu * (lu + 1.0)) /
             (2.0 * ju * (ju + 1.0));
    double geff = (g1 + g2) / 2 + (g1 - g2) * (jl * (jl + 1.) - ju * (ju + 1.)) / 4;

    int nm1 = 0, nm2 = 0;
    int m1[2 * ORC_MAXC + 2], m2[2 * ORC_MAXC + 2];
    for (int i = 0; i < 2 * jl + 1; i++) m1[nm1++] = i - (int)jl;   /* :85-89 */
    for (int i = 0; i < 2 * ju + 1; i++) m2[nm2++] = i - (int)ju;   /* :90-94 */

    int n_pi = (int)((2 * (jl < ju ? jl : ju)) + 1);                 /* :96 */
    int n_sig = (int)(jl + ju);                                      /* :97 */
    if (n_pi > ORC_MAXC || n_sig > ORC_MAXC) return -1;

    memset(z, 0, sizeof(*z));
    int ipi = 0, ib = 0, ir = 0;
    int jj = (int)(ju - jl);                                          /* :116 */
    for (int j = 0; j <= 2 * jl; j++) {
        for (int i = 0; i <= 2 * ju; i++) {
            int im = m2[i] - m1[j];
            double m = m1[j];
            double shift = g1 * m1[j] - g2 * m2[i];
            if (im == 0) {                     /* central component :125-143 */
                double s = 0;
                if (jj == -1) s = jl * jl - m * m;
                else if (jj == 0) s = m * m;
                else if (jj == 1) s = (jl + 1) * (jl + 1) - m * m;
                if (ipi >= ORC_MAXC) return -1;
                z->wep[ipi] = s; z->nup[ipi] = shift; ipi++;
            } else if (im == 1) {              /* blue :144-162 */
                double s = 0;
                if (jj == -1) s = (jl - m) * (jl - m - 1) / 4;
                else if (jj == 0) s = (jl - m) * (jl + m + 1) / 4;
                else if (jj == 1) s = (jl + m + 1) * (jl + m + 2) / 4;
                if (ib >= ORC_MAXC) return -1;
                z->web[ib] = s; z->nub[ib] = shift; ib++;
            } else if (im == -1) {             /* red :163-181 */
                double s = 0;
                if (jj == -1) s = (jl + m) * (jl + m - 1) / 4;
                else if (jj == 0) s = (jl + m) * (jl - m + 1) / 4;
                else if (jj == 1) s = (jl - m + 1) * (jl - m + 2) / 4;
                if (ir >= ORC_MAXC) return -1;
                z->wer[ir] = s; z->nur[ir] = shift; ir++;
            }
        }
    }
    /* normalisation of each group, :188-218 */
    double sp = 0, sb = 0, sr = 0;
    for (int i = 0; i < n_pi; i++) sp = sp + z->wep[i];
    for (int i = 0; i < n_pi; i++) z->wep[i] = z->wep[i] / sp;
    for (int i = 0; i < n_sig; i++) sb = sb + z->web[i];
    for (int i = 0; i < n_sig; i++) z->web[i] = z->web[i] / sb;
    for (int i = 0; i < n_sig; i++) sr = sr + z->wer[i];
    for (int i = 0; i < n_sig; i++) z->wer[i] = z->wer[i] / sr;

    z->n_pi = n_pi; z->n_sig = n_sig;
    z->gl = g1; z->gu = g2; z->geff = geff; z->fo = 1.0;   /* FO=1, :41 */
    return 0;
}

/* Instrumental Gaussian PSF on the grid, normalised to unit sum.
 * fgauss.c:79-154 (fgauss_WL), called as milos.c:269. */
void orc_gauss_psf(double fwhm_mA, double step, double lambda_start,
                   double lambda_centre, int n, double *G)
{
    double sigma = fwhm_mA * 0.42466090 / 1000.0;                  /* :92 */
    double cte = 0;
    for (int i = 0; i < n; i++) {
        double lx = lambda_start + i * step;                       /* :99 */
        double a = (lx - lambda_centre) / sigma;
        double term = (a * a) / 2;
        G[i] = (term < 1e30) ? exp(-term) : 0.0;                   /* :108-125 */
    }
    for (int i = 0; i < n; i++) cte += G[i];                       /* :140-143 */
    for (int i = 0; i < n; i++) G[i] /= cte;                       /* :145-148 */
}

/* NTERMS = 11 - [fix[10]==0] - [mac==0 && fix[9]==0], milos.c:154-157. */
int orc_nterms(const int fix[ORC_NPAR], double guess_mac)
{
    int nt = 11;
    if (fix[10] == 0) nt--;
    if (guess_mac == 0 && fix[9] == 0) nt--;
    return nt;
}

/* ------------------------------------------------------------------------ */
/* Primitives                                                                */
/* ------------------------------------------------------------------------ */

/* Voigt H and Faraday-Voigt F by the 7/8-term complex rational function of
 * fvoigt.c:15-80: w = P(z)/Q(z), z = a - i v; H = Re w, F = Im w / 2.
 * Coefficients fvoigt.c:20-26. */
static const double VA[7] = {122.607931777104326, 214.382388694706425,
                             181.928533092181549, 93.155580458138441,
                             30.180142196210589, 5.912626209773153,
                             0.564189583562615};
static const double VB[8] = {122.60793177387535, 352.730625110963558,
                             457.334478783897737, 348.703917719495792,
                             170.354001821091472, 53.992906912940207,
                             10.479857114260399, 1.};

void orc_fvoigt(double damp, const double *v, int n, double *H, double *F)
{
    for (int i = 0; i < n; i++) {
        double complex z = damp - v[i] * I;                        /* :30-33 */
        double complex num = VA[6];
        for (int j = 5; j >= 0; j--) num = num * z + VA[j];         /* :36-47 */
        double complex den = z + VB[6];
        for (int j = 5; j >= 0; j--) den = den * z + VB[j];         /* :50-61 */
        double complex w = num / den;                              /* :63-66 */
        H[i] = creal(w);                                           /* :69-72 */
        F[i] = cimag(w) * 0.5;                                     /* :74-77 */
    }
}

/* orc_fvoigt on REAL arguments (fvoigt.c:15: vv, h, f are REAL, the
 * rational function is evaluated in complex double). */
static void fvoigt_r(double damp, const real *v, int n, real *H, real *F)
{
    for (int i = 0; i < n; i++) {
        double complex z = damp - v[i] * I;
        double complex num = VA[6];
        for (int j = 5; j >= 0; j--) num = num * z + VA[j];
        double complex den = z + VB[6];
        for (int j = 5; j >= 0; j--) den = den * z + VB[j];
        double complex w = num / den;
        H[i] = creal(w);
        F[i] = cimag(w) * 0.5;
    }
}

/* The correlations of convolution.c on REAL rows (:54-78, :90-113) and the
 * circular convolution (:126-153): PRECISION kernel and sums, REAL data. */
static void conv_direct_r(real *x, int nx, const double *h, int nh)
{
    int half = nh / 2;
    double pad[3 * ORC_MAXL + 2];
    memset(pad, 0, sizeof(double) * (size_t)(nx + nh + 1));
    for (int k = 0; k < nx; k++) pad[k + half] = x[k];
    for (int k = 0; k < nx; k++) {
        double acc = 0;
        for (int j = 0; j < nh; j++) acc += h[j] * pad[j + k];
        x[k] = acc;
    }
}

static void conv_direct_ic_r(real *x, int nx, const double *h, int nh, real ic)
{
    int half = nh / 2;
    double pad[3 * ORC_MAXL + 2];
    memset(pad, 0, sizeof(double) * (size_t)(nx + nh + 1));
    for (int k = 0; k < nx; k++) pad[k + half] = ic - x[k];
    for (int k = 0; k < nx; k++) {
        double acc = 0;
        for (int j = 0; j < nh; j++) acc += h[j] * pad[j + k];
        x[k] = ic - acc;
    }
}

static void conv_circular_r(const real *x, const double *h, int n, real *out)
{
    real res[ORC_MAXL];
    for (int i = 0; i < n; i++) {
        double acc = 0;
        for (int j = 0; j < n; j++) {
            int m = i - j;
            if (m < 0) m += n;
            acc += h[j] * x[m];
        }
        int dst = (i < n / 2) ? (i + n / 2 + (n % 2)) : (i - n / 2);
        res[dst] = acc;
    }
    memcpy(out, res, sizeof(real) * (size_t)n);
}

/* Macroturbulence Gaussian (and its derivative w.r.t. v_mac), centred on
 * lambda[n/2], unit sum.  fgauss.c:16-66. */
void orc_gauss_mac(double mc, const double *lambda, int n, double lambda_line,
                   int deriv, double *out)
{
    double centro = lambda[n / 2];                                 /* :26 */
    double ild = (lambda_line * mc) / 2.99792458e5;                /* :27 */
    for (int i = 0; i < n; i++) {
        double a = (lambda[i] - centro) / ild;
        double term = (a * a) / 2;
        out[i] = (term < 1e30) ? exp(-term) : 0.0;                 /* :37-43 */
    }
    double cte = 0;
    for (int i = 0; i < n; i++) cte += out[i];
    for (int i = 0; i < n; i++) out[i] /= cte;
    if (deriv == 1) {                                              /* :57-63 */
        for (int i = 0; i < n; i++) {
            double a = (lambda[i] - centro) / ild;
            out[i] = out[i] / mc * ((a * a) - 1.0);
        }
    }
}

/* "Same"-size zero-padded correlation y[k] = sum_j h[j] xpad[j+k] with
 * xpad[m + nh/2] = x[m].  convolution.c:21-42 / :54-78. */
void orc_conv_direct(double *x, int nx, const double *h, int nh)
{
    int half = nh / 2;
    double pad[3 * ORC_MAXL + 2];
    memset(pad, 0, sizeof(double) * (size_t)(nx + nh + 1));
    for (int k = 0; k < nx; k++) pad[k + half] = x[k];
    for (int k = 0; k < nx; k++) {
        double acc = 0;
        for (int j = 0; j < nh; j++) acc += h[j] * pad[j + k];
        x[k] = acc;
    }
}

/* Same correlation applied to (ic - x), returning ic - result.
 * convolution.c:90-113. */
void orc_conv_direct_ic(double *x, int nx, const double *h, int nh, double ic)
{
    int half = nh / 2;
    double pad[3 * ORC_MAXL + 2];
    memset(pad, 0, sizeof(double) * (size_t)(nx + nh + 1));
    for (int k = 0; k < nx; k++) pad[k + half] = ic - x[k];
    for (int k = 0; k < nx; k++) {
        double acc = 0;
        for (int j = 0; j < nh; j++) acc += h[j] * pad[j + k];
        x[k] = ic - acc;
    }
}

/* Circular convolution sum_j h[j] x[(i-j) mod n], rotated so that
 * out[m] = conv[(m + n/2) mod n].  convolution.c:126-153. */
void orc_conv_circular(const double *x, const double *h, int n, double *out)
{
    double res[ORC_MAXL];
    for (int i = 0; i < n; i++) {
        double acc = 0;
        for (int j = 0; j < n; j++) {
            int m = i - j;
            if (m < 0) m += n;
            acc += h[j] * x[m];
        }
        int dst = (i < n / 2) ? (i + n / 2 + (n % 2)) : (i - n / 2);  /* :144-147 */
        res[dst] = acc;
    }
    memcpy(out, res, sizeof(double) * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* Synthesis (mil_sinrf.c:46-415) and response functions (me_der.c:48-609)   */
/* ------------------------------------------------------------------------ */

/* The intermediate state mil_sinrf leaves for me_der (the reference keeps it
 * in globals, calculosCompartidos.c:34-154). */
typedef struct {
    int L, ncomp[3];
    real sin_gm, cosi, sinis, cosis, cosis_2, sina, cosa;
    real sinda, cosda, sindi, cosdi, sinis_cosa, sinis_sina;
    double E0;                                  /* PRECISION */
    real shift[3][ORC_MAXC];                  /* nupB, nubB, nurB */
    real uu[3][ORC_MAXC][ORC_MAXL];
    real H[3][ORC_MAXC][ORC_MAXL], F[3][ORC_MAXC][ORC_MAXL];
    real fi[3][ORC_MAXL], shi[3][ORC_MAXL];   /* p, b, r */
    real parcial1[ORC_MAXL], parcial2[ORC_MAXL], parcial3[ORC_MAXL];
    real etain[ORC_MAXL], etaqn[ORC_MAXL], etaun[ORC_MAXL], etavn[ORC_MAXL];
    real rhoqn[ORC_MAXL], rhoun[ORC_MAXL], rhovn[ORC_MAXL];
    real etai[ORC_MAXL], etaq[ORC_MAXL], etau[ORC_MAXL], etav[ORC_MAXL];
    real rhoq[ORC_MAXL], rhou[ORC_MAXL], rhov[ORC_MAXL];
    real etai_2[ORC_MAXL], gp1[ORC_MAXL], gp2[ORC_MAXL], gp3[ORC_MAXL];
    real dt[ORC_MAXL], dti[ORC_MAXL];
    real g4r[ORC_MAXL], g5r[ORC_MAXL], g6r[ORC_MAXL];  /* gpX_gp2_rhoY */
    real spectra_mac[4 * ORC_MAXL];           /* before convolution */
    real spectra_slight[4 * ORC_MAXL];        /* before stray-light mix */
    real d_ei[7][ORC_MAXL], d_eq[7][ORC_MAXL], d_eu[7][ORC_MAXL], d_ev[7][ORC_MAXL];
    real d_rq[7][ORC_MAXL], d_ru[7][ORC_MAXL], d_rv[7][ORC_MAXL];
} orc_ws_t;

/* mil_sinrf.c:46-415.  Writes spectra[4L] and fills ws for orc_rf.  The
 * model (Init_Model) and its locals are PRECISION; the angle terms, the
 * Zeeman splittings and weights, u and every per-wavelength quantity are
 * REAL (milos.c:47-73, mil_sinrf.c:54-61). */
static void synth_ws(const orc_setup_t *s, const double *m, real *spectra,
                     orc_ws_t *ws)
{
    const int L = s->nlambda;
    const orc_zeeman_t *z = &s->z;
    const real CC = ORC_PI / 180.0, CC_2 = CC * 2;                /* milos.c:232-233 */
    double MF = m[ORC_B], VL = m[ORC_VLOS], LD = m[ORC_DOPP], A = m[ORC_AA];
    double B0 = m[ORC_S0], B1 = -((m[ORC_S1]) * s->ah);            /* :71-72 */
    double MC = m[ORC_MAC], ALF = m[ORC_ALFA];
    double AZI = m[ORC_AZ] * CC, GM = m[ORC_GM] * CC;              /* :84-85 */

    ws->L = L;
    ws->sin_gm = r_sin(GM);                                        /* :87-103 */
    ws->cosi = r_cos(GM);
    ws->sinis = ws->sin_gm * ws->sin_gm;
    ws->cosis = ws->cosi * ws->cosi;
    ws->cosis_2 = (1 + ws->cosis) / 2;
    real azi_2 = 2 * AZI;
    ws->sina = r_sin(azi_2);
    ws->cosa = r_cos(azi_2);
    ws->sinda = ws->cosa * CC_2;
    ws->cosda = -ws->sina * CC_2;
    ws->sindi = ws->cosi * ws->sin_gm * CC_2;
    ws->cosdi = -ws->sin_gm * CC;
    ws->sinis_cosa = ws->sinis * ws->cosa;
    ws->sinis_sina = ws->sinis * ws->sina;

    double E0 = m[ORC_ETA0] * z->fo;                               /* :117 */
    ws->E0 = E0;
    real ulos = (VL * s->lambda0) / (ORC_VLIGHT * LD);             /* :120 */
    real u[ORC_MAXL];
    for (int i = 0; i < L; i++) u[i] = ((s->lambda[i] - s->lambda0) / LD) - ulos;

    real parcial = ((s->lambda0 * s->lambda0) / LD) * ORC_CTE4_6_13; /* :153 */
    ws->ncomp[0] = z->n_pi; ws->ncomp[1] = z->n_sig; ws->ncomp[2] = z->n_sig;
    for (int i = 0; i < z->n_pi; i++) ws->shift[0][i] = parcial * (real)z->nup[i];
    for (int i = 0; i < z->n_sig; i++) ws->shift[1][i] = parcial * (real)z->nub[i];
    for (int i = 0; i < z->n_sig; i++)                             /* :164-166 */
        ws->shift[2][i] = -ws->shift[1][z->n_sig - (i + 1)];

    const double *wgt[3] = {z->wep, z->web, z->wer};
    for (int g = 0; g < 3; g++) {           /* funcionComponentFor_sinrf :418-449 */
        for (int i = 0; i < L; i++) { ws->fi[g][i] = 0; ws->shi[g][i] = 0; }
        for (int c = 0; c < ws->ncomp[g]; c++) {
            real *uu = ws->uu[g][c], *H = ws->H[g][c], *F = ws->F[g][c];
            const real w = (real)wgt[g][c];
            for (int j = 0; j < L; j++) uu[j] = u[j] - ws->shift[g][c] * MF;
            fvoigt_r(A, uu, L, H, F);
            for (int j = 0; j < L; j++) ws->fi[g][j] = ws->fi[g][j] + w * H[j];
            for (int j = 0; j < L; j++)
                ws->shi[g][j] = (ws->shi[g][j] + (w * F[j] * 2));
        }
    }
    const real *fi_p = ws->fi[0], *fi_b = ws->fi[1], *fi_r = ws->fi[2];
    const real *shi_p = ws->shi[0], *shi_b = ws->shi[1], *shi_r = ws->shi[2];

    real E0_2 = E0 / 2.0;                                           /* :188-223 */
    for (int i = 0; i < L; i++) {
        ws->parcial1[i] = fi_b[i] + fi_r[i];
        ws->parcial2[i] = (E0_2) * (fi_p[i] - (ws->parcial1[i]) / 2);
        ws->parcial3[i] = (E0_2) * (shi_p[i] - (shi_b[i] + shi_r[i]) / 2);
    }
    real cosi_E0_2 = E0_2 * ws->cosi;
    for (int i = 0; i < L; i++) {
        ws->etain[i] = ((E0_2) * (fi_p[i] * ws->sinis + (ws->parcial1[i]) * ws->cosis_2));
        ws->etaqn[i] = (ws->parcial2[i] * ws->sinis_cosa);
        ws->etaun[i] = (ws->parcial2[i] * ws->sinis_sina);
        ws->etavn[i] = (fi_r[i] - fi_b[i]) * cosi_E0_2;
        ws->rhoqn[i] = (ws->parcial3[i] * ws->sinis_cosa);
        ws->rhoun[i] = (ws->parcial3[i] * ws->sinis_sina);
        ws->rhovn[i] = (shi_r[i] - shi_b[i]) * cosi_E0_2;
        ws->etai[i] = 1.0 + ws->etain[i];
        ws->etaq[i] = ws->etaqn[i]; ws->etau[i] = ws->etaun[i]; ws->etav[i] = ws->etavn[i];
        ws->rhoq[i] = ws->rhoqn[i]; ws->rhou[i] = ws->rhoun[i]; ws->rhov[i] = ws->rhovn[i];
    }

    for (int i = 0; i < L; i++) {                                   /* :229-266 */
        real ei = ws->etai[i], eq = ws->etaq[i], eu = ws->etau[i], ev = ws->etav[i];
        real rq = ws->rhoq[i], ru = ws->rhou[i], rv = ws->rhov[i];
        real ei2 = ei * ei;
        real auxq = rq * rq, auxu = ru * ru, auxv = rv * rv;
        ws->etai_2[i] = ei2;
        ws->gp1[i] = ei2 - eq * eq - eu * eu - ev * ev + auxq + auxu + auxv;
        ws->gp3[i] = ei2 + auxq + auxu + auxv;
        ws->gp2[i] = eq * rq + eu * ru + ev * rv;
        ws->dt[i] = ei2 * ws->gp1[i] - ws->gp2[i] * ws->gp2[i];
        ws->dti[i] = 1.0 / ws->dt[i];
        real gp4 = ei2 * eq + ei * (ev * ru - eu * rv);
        real gp5 = ei2 * eu + ei * (eq * rv - ev * rq);
        real gp6 = (ei2) * ev + ei * (eu * rq - eq * ru);
        ws->g4r[i] = gp4 + rq * ws->gp2[i];
        ws->g5r[i] = gp5 + ru * ws->gp2[i];
        ws->g6r[i] = gp6 + rv * ws->gp2[i];
    }
    for (int i = 0; i < L; i++) {                                   /* :268-285 */
        real dtiaux = ws->dti[i] * (B1);
        spectra[i] = B0 - dtiaux * ws->etai[i] * ws->gp3[i];
        spectra[i + L] = (dtiaux * (ws->g4r[i]));
        spectra[i + 2 * L] = (dtiaux * (ws->g5r[i]));
        spectra[i + 3 * L] = (dtiaux * (ws->g6r[i]));
    }
    memcpy(ws->spectra_mac, spectra, sizeof(real) * 4 * (size_t)L);

    int mac_applied = 0;
    if (MC > 0.0001) {                                              /* :288-357 */
        mac_applied = 1;
        double gmac[ORC_MAXL];
        orc_gauss_mac(MC, s->lambda, L, s->lambda0, 0, gmac);
        if (s->use_psf && s->use_fft)                               /* :302-314 */
            orc_conv_circular(gmac, s->G, L, gmac);
        else if (s->use_psf)
            orc_conv_direct(gmac, L, s->G, L);
        for (int st = 0; st < 4; st++)
            conv_circular_r(spectra + L * st, gmac, L, spectra + L * st);
    }
    if (!mac_applied && s->use_psf && s->use_fft) {                 /* :362-386 */
        for (int st = 0; st < 4; st++)
            conv_circular_r(spectra + L * st, s->G, L, spectra + L * st);
    } else if (!mac_applied && s->use_psf) {                        /* :387-402 */
        real ic = (spectra[0] > spectra[L - 1]) ? spectra[0] : spectra[L - 1];
        conv_direct_ic_r(spectra, L, s->G, L, ic);
        for (int st = 1; st < 4; st++) conv_direct_r(spectra + L * st, L, s->G, L);
    }
    if (s->use_slight) {                                            /* :405-412 */
        for (int i = 0; i < 4 * L; i++) {
            ws->spectra_slight[i] = spectra[i];
            spectra[i] = spectra[i] * ALF + (real)s->slight[i] * (1.0 - ALF);
        }
    }
}

/* me_der.c:48-530.  Needs ws from synth_ws of the same model.  spec_in is
 * the reference's "spectra" argument (the unconvolved synthesis) and
 * spec_slight its "spectra_slight" argument. */
static void rf_ws(const orc_setup_t *s, const double *m, orc_ws_t *ws,
                  const real *spec_in, const real *spec_slight, real *D)
{
    const int L = s->nlambda, NT = s->nterms;
    const double LD = m[ORC_DOPP], A = m[ORC_AA];
    const double B1 = -((m[ORC_S1]) * s->ah), MC = m[ORC_MAC], ALF = m[ORC_ALFA];
    const double E0 = ws->E0;
    real dfi[3][4][ORC_MAXL], dshi[3][4][ORC_MAXL];
    memset(dfi, 0, sizeof(dfi));
    memset(dshi, 0, sizeof(dshi));

    const double *wgt[3] = {s->z.wep, s->z.web, s->z.wer};
    real auxCte = (-s->lambda0) / (ORC_VLIGHT * LD);                /* :544-546 */
    const real rk = s->exact_rho_deriv ? 2.0 : 1.0;
    for (int g = 0; g < 3; g++) {            /* funcionComponentFor :536-609 */
        for (int c = 0; c < ws->ncomp[g]; c++) {     /* last component wins */
            const real *H = ws->H[g][c], *F = ws->F[g][c];
            const real *uu = ws->uu[g][c];
            real nux = ws->shift[g][c];
            const real w = (real)wgt[g][c];
            for (int j = 0; j < L; j++) {
                real dH_u = ((4 * A * F[j]) - (2 * uu[j] * H[j])) * w;
                real dF_u = (ORC_RR - A * H[j] - 2 * uu[j] * F[j]) * w;
                real uun = -uu[j] / LD;
                dfi[g][0][j] = dH_u * (-nux);
                dfi[g][1][j] = dH_u * auxCte;
                dfi[g][2][j] = (dH_u * uun);
                dfi[g][3][j] = (-2 * dF_u);
                dshi[g][0][j] = rk * ((dF_u) * (-nux));
                dshi[g][1][j] = rk * (dF_u * auxCte);
                dshi[g][2][j] = rk * (dF_u * uun);
                dshi[g][3][j] = rk * (dH_u / 2);
            }
        }
    }

    /* derivatives of eta/rho w.r.t. parameters 0..6, :93-199 */
    real (*d_ei)[ORC_MAXL] = ws->d_ei, (*d_eq)[ORC_MAXL] = ws->d_eq;
    real (*d_eu)[ORC_MAXL] = ws->d_eu, (*d_ev)[ORC_MAXL] = ws->d_ev;
    real (*d_rq)[ORC_MAXL] = ws->d_rq, (*d_ru)[ORC_MAXL] = ws->d_ru;
    real (*d_rv)[ORC_MAXL] = ws->d_rv;
    memset(ws->d_ei, 0, sizeof(ws->d_ei)); memset(ws->d_eq, 0, sizeof(ws->d_eq));
    memset(ws->d_eu, 0, sizeof(ws->d_eu)); memset(ws->d_ev, 0, sizeof(ws->d_ev));
    memset(ws->d_rq, 0, sizeof(ws->d_rq)); memset(ws->d_ru, 0, sizeof(ws->d_ru));
    memset(ws->d_rv, 0, sizeof(ws->d_rv));
    for (int i = 0; i < L; i++) {
        d_ei[0][i] = ws->etain[i] / E0; d_eq[0][i] = ws->etaqn[i] / E0;
        d_eu[0][i] = ws->etaun[i] / E0; d_ev[0][i] = ws->etavn[i] / E0;
        d_rq[0][i] = ws->rhoqn[i] / E0; d_ru[0][i] = ws->rhoun[i] / E0;
        d_rv[0][i] = ws->rhovn[i] / E0;
    }
    real sc = E0 * ws->sinis_cosa / 2;                              /* :118-127 */
    real ss = E0 * ws->sinis_sina / 2;
    real E0_2 = E0 / 2.0;
    real sindi_cosa = ws->sindi * ws->cosa, sindi_sina = ws->sindi * ws->sina;
    real cosdi_E0_2 = (E0_2) * ws->cosdi;
    real cosi_2_E0 = E0 * ws->cosi / 2.0;
    real cosis_2_E0_2 = ws->cosis_2 * E0_2;
    real sinis_E0_2 = ws->sinis * E0_2;
    for (int j = 1; j < 5; j++) {                                   /* :130-154 */
        for (int i = 0; i < L; i++) {
            real dfisum = dfi[1][j - 1][i] + dfi[2][j - 1][i];
            d_ei[j][i] = (dfi[0][j - 1][i] * sinis_E0_2 + dfisum * cosis_2_E0_2);
            real aux = dfi[0][j - 1][i] - dfisum / 2;
            d_eq[j][i] = (aux) * sc;
            d_eu[j][i] = (aux) * ss;
            d_ev[j][i] = (dfi[2][j - 1][i] - dfi[1][j - 1][i]) * cosi_2_E0;
            real auxr = dshi[0][j - 1][i] - (dshi[1][j - 1][i] + dshi[2][j - 1][i]) / 2.0;
            d_rq[j][i] = (auxr) * sc;
            d_ru[j][i] = (auxr) * ss;
            d_rv[j][i] = ((dshi[2][j - 1][i] - dshi[1][j - 1][i])) * cosi_2_E0;
        }
    }
    real cosi_cosdi = ws->cosi * ws->cosdi * E0_2;                   /* :157-180 */
    real sindi_E0_2 = ws->sindi * E0_2;
    for (int i = 0; i < L; i++) {
        d_ei[5][i] = ws->fi[0][i] * sindi_E0_2 + (ws->parcial1[i]) * cosi_cosdi;
        d_eq[5][i] = ws->parcial2[i] * sindi_cosa;
        d_eu[5][i] = ws->parcial2[i] * sindi_sina;
        d_ev[5][i] = (ws->fi[2][i] - ws->fi[1][i]) * cosdi_E0_2;
        d_rq[5][i] = ws->parcial3[i] * sindi_cosa;
        d_ru[5][i] = ws->parcial3[i] * sindi_sina;
        d_rv[5][i] = (ws->shi[2][i] - ws->shi[1][i]) * cosdi_E0_2;
    }
    real sinis_cosda = ws->sinis * ws->cosda, sinis_sinda = ws->sinis * ws->sinda;
    for (int i = 0; i < L; i++) {                                   /* :184-199 */
        d_eq[6][i] = ws->parcial2[i] * sinis_cosda;
        d_eu[6][i] = ws->parcial2[i] * sinis_sinda;
        d_rq[6][i] = ws->parcial3[i] * sinis_cosda;
        d_ru[6][i] = ws->parcial3[i] * sinis_sinda;
    }

    /* chain rule through the Unno-Rachkovsky solution, :207-277 */
    for (int i = 0; i < L; i++) {
        real ei = ws->etai[i], eq = ws->etaq[i], eu = ws->etau[i], ev = ws->etav[i];
        real rq = ws->rhoq[i], ru = ws->rhou[i], rv = ws->rhov[i];
        real ei2 = ws->etai_2[i], gp1 = ws->gp1[i], gp2 = ws->gp2[i], gp3 = ws->gp3[i];
        real dt = ws->dt[i];
        real dtaux = (B1) / (dt * dt);
        real etai_gp3 = ei * gp3;
        real aux = 2 * ei;
        real ext1 = aux * eq + ev * ru - eu * rv;
        real ext2 = aux * eu + eq * rv - ev * rq;
        real ext3 = aux * ev + eu * rq - eq * ru;
        real ext4 = aux * gp1;
        for (int il = 0; il < 7; il++) {
            real dei = d_ei[il][i], deq = d_eq[il][i], deu = d_eu[il][i], dev = d_ev[il][i];
            real drq = d_rq[il][i], dru = d_ru[il][i], drv = d_rv[il][i];
            real dgp1 = 2.0 * (ei * dei - eq * deq - eu * deu - ev * dev + rq * drq +
                               ru * dru + rv * drv);
            real dgp2 = rq * deq + eq * drq + ru * deu + eu * dru + rv * dev + ev * drv;
            real d_dt = ext4 * dei + ei2 * dgp1 - 2 * gp2 * dgp2;
            real dgp3 = 2.0 * (ei * dei + rq * drq + ru * dru + rv * drv);
            D[i + L * il] = -(((dei * gp3 + ei * dgp3) * dt - d_dt * etai_gp3) * (dtaux));
            real dgp4 = dei * (ext1) + (ei2) * deq + ei * (ru * dev + ev * dru - rv * deu - eu * drv);
            D[i + L * il + L * NT] = ((dgp4 + drq * gp2 + rq * dgp2) * dt - d_dt * (ws->g4r[i])) * (dtaux);
            real dgp5 = dei * (ext2) + (ei2) * deu + ei * (rv * deq + eq * drv - rq * dev - ev * drq);
            D[i + L * il + L * NT * 2] = ((dgp5 + dru * gp2 + ru * dgp2) * dt - d_dt * (ws->g5r[i])) * (dtaux);
            real dgp6 = dei * (ext3) + (ei2) * dev + ei * (rq * deu + eu * drq - ru * deq - eq * dru);
            D[i + L * il + L * NT * 3] = ((dgp6 + drv * gp2 + rv * dgp2) * dt - d_dt * (ws->g6r[i])) * (dtaux);
        }
        /* S1 (:281-298) uses dti scaled in place by -ah */
        real dtin = -(ws->dti[i] * s->ah);
        D[i + L * 8] = -dtin * etai_gp3;
        D[i + L * 8 + L * NT] = dtin * (ws->g4r[i]);
        D[i + L * 8 + L * NT * 2] = dtin * (ws->g5r[i]);
        D[i + L * 8 + L * NT * 3] = dtin * (ws->g6r[i]);
        /* S0 (:301-306) */
        D[i + L * 7] = 1;
        D[i + L * 7 + L * NT] = 0;
        D[i + L * 7 + L * NT * 2] = 0;
        D[i + L * 7 + L * NT * 3] = 0;
        /* azimuth on I and V (:309-312) */
        D[i + L * 6] = 0;
        D[i + L * 6 + L * NT * 3] = 0;
    }

    int mac_applied = 0;
    /* Without macroturbulence (v_mac <= 1e-4, :317) the reference leaves its
     * v_mac row as the last call that had one computed it -- the previous
     * model, or another pixel -- and convolves it with the PSF again
     * (:471-495, every row but S0's).  Here it is zero, the derivative of
     * the spectra with respect to v_mac at v_mac = 0 (the kernel's width
     * enters at second order); the GPU kernels do the same. */
    if (NT >= 10 && !(MC > 0.0001))
        for (int st = 0; st < 4; st++)
            for (int i = 0; i < L; i++) D[9 * L + L * NT * st + i] = 0;
    if (MC > 0.0001) {                                              /* :317-438 */
        mac_applied = 1;
        double gmac[ORC_MAXL], gder[ORC_MAXL];
        orc_gauss_mac(MC, s->lambda, L, s->lambda0, 0, gmac);
        double ild = (s->lambda0 * MC) / 2.99792458e5;
        double centro = s->lambda[L / 2];
        for (int i = 0; i < L; i++) {
            double a = (s->lambda[i] - centro) / ild;
            gder[i] = (gmac[i] / MC * ((a * a) - 1.0));
        }
        if (s->use_psf && s->use_fft) {                             /* :331-354 */
            orc_conv_circular(gder, s->G, L, gder);
            orc_conv_circular(gmac, s->G, L, gmac);
        } else if (s->use_psf) {
            orc_conv_direct(gder, L, s->G, L);
            orc_conv_direct(gmac, L, s->G, L);
        }
        for (int st = 0; st < 4; st++)       /* run twice in the reference */
            conv_circular_r(spec_in + L * st, gder, L, D + 9 * L + L * NT * st);
        for (int st = 0; st < 4; st++)
            for (int p = 0; p < 9; p++)
                if (p != 7)
                    conv_circular_r(D + L * p + L * NT * st, gmac, L,
                                    D + L * p + L * NT * st);
    }
    if (!mac_applied && s->use_psf && s->use_fft) {                 /* :441-470 */
        for (int st = 0; st < 4; st++)
            for (int p = 0; p < NT; p++)
                if (p != 7)
                    conv_circular_r(D + L * p + L * NT * st, s->G, L,
                                    D + L * p + L * NT * st);
    } else if (!mac_applied && s->use_psf) {                        /* :471-495 */
        for (int p = 0; p < NT; p++) {
            if (p == 7) continue;
            real *row = D + L * p;
            real ic = (row[0] > row[L - 1]) ? row[0] : row[L - 1];
            conv_direct_ic_r(row, L, s->G, L, ic);
        }
        for (int st = 1; st < 4; st++)
            for (int p = 0; p < NT; p++)
                if (p != 7) conv_direct_r(D + L * p + L * NT * st, L, s->G, L);
    }
    if (s->use_slight) {                                            /* :499-525 */
        for (int st = 0; st < 4; st++) {
            for (int p = 0; p < NT; p++) {
                for (int i = 0; i < L; i++) {
                    real *d = D + L * p + L * NT * st + i;
                    *d = *d * ALF;
                    int ff_row = -1;
                    if (NT == 11) ff_row = 10;
                    else if (s->fix[9]) ff_row = 10;
                    else ff_row = 9;
                    if (p == ff_row)
                        *d = spec_slight[L * st + i] - (real)s->slight[L * st + i];
                }
            }
        }
    }
}

void orc_synth(const orc_setup_t *s, const double model[ORC_NPAR],
               double *spectra, double *rf)
{
    const size_t L = s->nlambda, NT = s->nterms;
    orc_ws_t *ws = (orc_ws_t *)malloc(sizeof(orc_ws_t));
    real *sp = (real *)malloc(sizeof(real) * 4 * L);
    real *d = rf ? (real *)calloc(4 * NT * L, sizeof(real)) : NULL;
    synth_ws(s, model, sp, ws);
    if (rf) rf_ws(s, model, ws, ws->spectra_mac, ws->spectra_slight, d);  /* milos.c:601-602 */
    for (size_t i = 0; i < 4 * L; i++) spectra[i] = sp[i];
    if (rf)
        for (size_t i = 0; i < 4 * NT * L; i++) rf[i] = d[i];
    free(d);
    free(sp);
    free(ws);
}

/* ------------------------------------------------------------------------ */
/* Normal equations, chi2, solve, update                                     */
/* ------------------------------------------------------------------------ */

/* covarm2, lib.c:71-116, with the per-Stokes sums of lib.c:237-268: REAL
 * sums (float in the reference's default build). */
static void covarm_r(const orc_setup_t *s, const real *sig, const float *obs,
                     const real *spectra, const real *rf, real *beta, real *alpha)
{
    const int L = s->nlambda, NT = s->nterms;
    real AP[4][ORC_NPAR * ORC_NPAR], BT[4][ORC_NPAR];
    for (int st = 0; st < 4; st++) {
        real opa[ORC_MAXL];
        real w = s->weight[st];
        for (int i = 0; i < L; i++) opa[i] = w * (spectra[i + L * st] - obs[i + L * st]);
        const real *d = rf + st * L * NT;
        for (int i = 0; i < NT; i++) {
            for (int h = 0; h < NT; h++) {
                real sum = 0, sum2 = 0;
                for (int k = 0; k < L; k++) {
                    real dAux = d[h * L + k];
                    sum += (d[i * L + k] * dAux) * (w / sig[L * st + k]);
                    if (i == 0) sum2 += ((opa[k] * dAux)) / sig[L * st + k];
                }
                AP[st][NT * i + h] = sum;
                if (i == 0) BT[st][h] = sum2;
            }
        }
    }
    for (int i = 0; i < NT; i++) {                                  /* lib.c:237-248 */
        real sum = 0;
        for (int st = 0; st < 4; st++) sum += BT[st][i];
        beta[i] = sum;
    }
    for (int i = 0; i < NT * NT; i++) {                             /* lib.c:257-268 */
        real sum = 0;
        for (int st = 0; st < 4; st++) sum += AP[st][i];
        alpha[i] = sum;
    }
}

/* fchisqr, lib.c:126-152 (REAL sums). */
static real chisqr_r(const orc_setup_t *s, const real *spectra, const float *obs,
                     const real *sig, real nfree)
{
    const int L = s->nlambda;
    real o[4] = {0, 0, 0, 0};
    for (int i = 0; i < L; i++) {
        for (int st = 0; st < 4; st++) {
            real d = spectra[i + L * st] - obs[i + L * st];
            o[st] += (((d * d) * (real)s->weight[st]) / (sig[i + L * st]));
        }
    }
    real tot = 0;
    tot += o[0] + o[1] + o[2] + o[3];
    return tot / nfree;
}

void orc_covarm(const orc_setup_t *s, const double *sig, const float *obs,
                const double *spectra, const double *rf, double *beta,
                double *alpha)
{
    const int L = s->nlambda, NT = s->nterms;
    real sg[4 * ORC_MAXL], sp[4 * ORC_MAXL], b[ORC_NPAR], a[ORC_NPAR * ORC_NPAR];
    real *d = (real *)malloc(sizeof(real) * 4 * (size_t)NT * (size_t)L);
    for (int i = 0; i < 4 * L; i++) { sg[i] = sig[i]; sp[i] = spectra[i]; }
    for (int i = 0; i < 4 * NT * L; i++) d[i] = rf[i];
    covarm_r(s, sg, obs, sp, d, b, a);
    for (int i = 0; i < NT; i++) beta[i] = b[i];
    for (int i = 0; i < NT * NT; i++) alpha[i] = a[i];
    free(d);
}

double orc_chisqr(const orc_setup_t *s, const double *spectra, const float *obs,
                  const double *sig, double nfree)
{
    const int L = s->nlambda;
    real sg[4 * ORC_MAXL], sp[4 * ORC_MAXL];
    for (int i = 0; i < 4 * L; i++) { sg[i] = sig[i]; sp[i] = spectra[i]; }
    return chisqr_r(s, sp, obs, sg, nfree);
}

/* Symmetric eigen-decomposition by cyclic Jacobi rotations (own method; the
 * reference calls GSL gsl_eigen_symmv, milosUtils.c:247-250, whose output
 * order is unspecified).  a is n*n row-major (not modified), w[n] the
 * eigenvalues, v[n*n] the eigenvectors as columns (v[i*n+j] = component i of
 * vector j), as GSL returns them. */
void orc_eigen_sym(int n, const double *a_in, double *w, double *v)
{
    double a[ORC_NPAR * ORC_NPAR];
    memcpy(a, a_in, sizeof(double) * (size_t)(n * n));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; i++) {
            diag += a[i * n + i] * a[i * n + i];
            for (int j = i + 1; j < n; j++) off += a[i * n + j] * a[i * n + j];
        }
        if (off == 0.0 || off <= 1e-36 * diag) break;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (apq == 0.0) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) /
                           (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; k++) {               /* columns p, q */
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - sn * akq;
                    a[k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {               /* rows p, q */
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - sn * aqk;
                    a[q * n + k] = sn * apk + c * aqk;
                }
                a[p * n + q] = a[q * n + p] = 0.0;
                for (int k = 0; k < n; k++) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - sn * vkq;
                    v[k * n + q] = sn * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) w[i] = a[i * n + i];
}

/* Eigen pseudo-inverse solve delta = V diag(1/w, |w|>1e-12) V^T beta,
 * milosUtils.c:237-271. */
void orc_pinv_solve(int nt, const double *h, const double *beta, double *delta)
{
    const double eps = 1e-12;                                       /* :240 */
    double w[ORC_NPAR], v[ORC_NPAR * ORC_NPAR], aux[ORC_NPAR];
    orc_eigen_sym(nt, h, w, v);
    for (int j = 0; j < nt; j++) {                                  /* :254-260 */
        double sum = 0;
        for (int i = 0; i < nt; i++) sum += beta[i] * v[i * nt + j];
        aux[j] = sum * ((fabs(w[j]) > eps) ? (1 / w[j]) : 0.0);
    }
    for (int i = 0; i < nt; i++) {                                  /* :263-269 */
        double sum = 0;
        for (int j = 0; j < nt; j++) sum += v[i * nt + j] * aux[j];
        delta[i] = sum;
    }
}

/* AplicaDelta, milosUtils.c:47-115: out = cur - delta for free parameters,
 * with B clipped to +-300 and gamma/azimuth to +-30 (written back into
 * delta). */
void orc_aplica_delta(const orc_setup_t *s, const double cur[ORC_NPAR],
                      double *delta, double out[ORC_NPAR])
{
    const int *f = s->fix;
    if (f[0]) out[ORC_ETA0] = cur[ORC_ETA0] - delta[0];
    if (f[1]) {
        if (delta[1] < -300) delta[1] = -300;
        else if (delta[1] > 300) delta[1] = 300;
        out[ORC_B] = cur[ORC_B] - delta[1];
    }
    if (f[2]) out[ORC_VLOS] = cur[ORC_VLOS] - delta[2];
    if (f[3]) out[ORC_DOPP] = cur[ORC_DOPP] - delta[3];
    if (f[4]) out[ORC_AA] = cur[ORC_AA] - delta[4];
    if (f[5]) {
        if (delta[5] < -30) delta[5] = -30;
        else if (delta[5] > 30) delta[5] = 30;
        out[ORC_GM] = cur[ORC_GM] - delta[5];
    }
    if (f[6]) {
        if (delta[6] < -30) delta[6] = -30;
        else if (delta[6] > 30) delta[6] = 30;
        out[ORC_AZ] = cur[ORC_AZ] - delta[6];
    }
    if (f[7]) out[ORC_S0] = cur[ORC_S0] - delta[7];
    if (f[8]) out[ORC_S1] = cur[ORC_S1] - delta[8];
    if (f[9]) out[ORC_MAC] = cur[ORC_MAC] - delta[9];
    if (f[10]) {
        if (s->nterms == 11) out[ORC_ALFA] = cur[ORC_ALFA] - delta[10];
        else out[ORC_ALFA] = cur[ORC_ALFA] - delta[9];
    }
}

/* Physical clamps, milosUtils.c:118-196. */
void orc_check(double m[ORC_NPAR])
{
    if (m[ORC_B] < 0) { m[ORC_B] = -(m[ORC_B]); m[ORC_GM] = 180.0 - (m[ORC_GM]); }
    if (m[ORC_B] > 5000) m[ORC_B] = 5000;
    if (m[ORC_GM] < 0) m[ORC_GM] = -(m[ORC_GM]);
    if (m[ORC_GM] > 180) m[ORC_GM] = 360.0 - m[ORC_GM];
    if (m[ORC_AZ] < 0) m[ORC_AZ] = 180 + (m[ORC_AZ]);
    if (m[ORC_AZ] > 180) m[ORC_AZ] = m[ORC_AZ] - 180.0;
    if (m[ORC_ETA0] < 1) m[ORC_ETA0] = 1;
    if (m[ORC_ETA0] > 2500) m[ORC_ETA0] = 2500;
    if (m[ORC_VLOS] < (-20)) m[ORC_VLOS] = (-20);
    if (m[ORC_VLOS] > 20) m[ORC_VLOS] = 20;
    if (m[ORC_DOPP] < 0.0001) m[ORC_DOPP] = 0.0001;
    if (m[ORC_DOPP] > 0.6) m[ORC_DOPP] = 0.6;
    if (m[ORC_AA] < 0.0001) m[ORC_AA] = 0.0001;
    if (m[ORC_AA] > 10.0) m[ORC_AA] = 10.0;
    if (m[ORC_S0] < 0.0001) m[ORC_S0] = 0.0001;
    if (m[ORC_S0] > 2.00) m[ORC_S0] = 2.00;
    if (m[ORC_S1] < 0.0001) m[ORC_S1] = 0.0001;
    if (m[ORC_S1] > 2.00) m[ORC_S1] = 2.00;
    if (m[ORC_MAC] < 0) m[ORC_MAC] = 0;
    if (m[ORC_MAC] > 4) m[ORC_MAC] = 4;
    if (m[ORC_ALFA] < 0) m[ORC_ALFA] = 0.0;
    if (m[ORC_ALFA] > 1.0) m[ORC_ALFA] = 1.0;
}

/* FijaACeroDerivadasNoNecesarias, milosUtils.c:199-227. */
static void fija_cero_r(const orc_setup_t *s, real *rf)
{
    const int L = s->nlambda, NT = s->nterms;
    for (int p = 0; p < NT; p++) {
        int zero;
        if (NT == 9 || NT == 11) zero = (s->fix[p] == 0);
        else if (NT == 10) zero = (p < 9) ? (s->fix[p] == 0) : (s->fix[9] == 0 && s->fix[10] == 0);
        else zero = 0;
        if (!zero) continue;
        for (int st = 0; st < 4; st++)
            for (int i = 0; i < L; i++) rf[i + L * p + st * L * NT] = 0;
    }
}

void orc_fija_cero(const orc_setup_t *s, double *rf)
{
    const size_t n = 4 * (size_t)s->nterms * (size_t)s->nlambda;
    real *d = (real *)malloc(sizeof(real) * n);
    for (size_t i = 0; i < n; i++) d[i] = rf[i];
    fija_cero_r(s, d);
    for (size_t i = 0; i < n; i++) rf[i] = d[i];
    free(d);
}

/* ------------------------------------------------------------------------ */
/* Per pixel                                                                 */
/* ------------------------------------------------------------------------ */

/* Classical estimates, milosUtils.c:287-452.  for_init = forInitialUse:
 * 1 for the inversion's starting point, 0 for the estimates-only mode
 * (no 85-95 degree nudge of B's inclination; B_los stored in S0). */
void orc_classical_ex(const orc_setup_t *s, const float *spectro, double model[ORC_NPAR],
                      int for_init)
{
    const int n = s->nlambda;
    const double l0 = s->lambda0;
    const double *lam = s->lambda;
    const float *sI = spectro, *sQ = spectro + n, *sU = spectro + 2 * n, *sV = spectro + 3 * n;

    int begin = 0;                                                  /* :302-325 */
    for (int i = 0; i < n; i++) { if (sI[i] < 0) begin++; else break; }
    int end = n;
    for (int i = n - 1; i >= 0; i--) { if (sI[i] < 0) end--; else break; }

    /* :328; the reference reads spectro[-1] when every I sample is < 0 */
    double Ic = (end >= 1) ? spectro[end - 1] : 0.0;
    double x = 0, y = 0, xv = 0, yv = 0;
    for (int i = begin; i < end - 1; i++) {                         /* :334-344 */
        if (sI[i] > -1 && sV[i] > -1) {
            double aux = (Ic - (sI[i] + sV[i]));
            double auxv = (Ic - sI[i]);
            x += (aux * (lam[i] - l0));
            xv += (auxv * (lam[i] - l0));
            y += aux;
            yv += auxv;
        }
    }
    double lplus = (fabs(y) > 1e-15) ? x / y : 0;                   /* :347-350 */
    x = 0; y = 0;
    for (int i = begin; i < end - 1; i++) {                         /* :354-361 */
        if (sI[i] > -1 && sV[i] > -1) {
            double aux = (Ic - (sI[i] - sV[i]));
            x += (aux * (lam[i] - l0));
            y += aux;
        }
    }
    double lminus = (fabs(y) > 1e-15) ? x / y : 0;
    double C = (ORC_CTE4_6_13 * (l0 * l0) * s->z.geff);             /* :368 */
    double Blos = (1 / C) * ((lplus - lminus) / 2);                  /* :371 */
    double Vlos = (ORC_VLIGHT / (l0)) * ((xv / yv) / 2);             /* :372 */

    x = 0; y = 0;                                                   /* :375-393 */
    for (int i = begin; i < end - 1; i++) {
        if (sQ[i] > -1 && sU[i] > -1 && sV[i] > -1) {
            float qu = sQ[i] * sQ[i] + sU[i] * sU[i];   /* float arithmetic, as in C */
            double Lp = fabs(sqrt(qu));
            double m = fabs((4 * (lam[i] - l0) * Lp));
            x = x + fabs(sV[i]) * m;
            y = y + fabs(sV[i]) * fabs(sV[i]);
        }
    }
    y = y * fabs((3 * C * Blos));
    double tan_gamma = fabs(sqrt(x / y));
    double gamma_rad = atan(tan_gamma);
    double gamma = gamma_rad * (180 / ORC_PI);
    if (for_init) {                                                  /* :395-402 */
        if (gamma >= 85 && gamma <= 90) gamma_rad = 85 * (ORC_PI / 180);
        if (gamma > 90 && gamma <= 95) gamma_rad = 95 * (ORC_PI / 180);
    }
    if (Blos < 0) gamma = (180) - gamma;                            /* :405-406 */

    double su = 0.0, sq = 0.0;                                      /* :412-420 */
    for (int i = 0; i < n; i++) {
        if (sU[i] > -1 && sQ[i] > -1) {
            if (fabs(sU[i]) > 0.0001 || fabs(sQ[i]) > 0.0001) { su += sU[i]; sq += sQ[i]; }
        }
    }
    double tan2phi = su / sq;
    double phi = (atan(tan2phi) * 180 / ORC_PI) / 2;                /* :422-430 */
    if (su > 0 && sq > 0) phi = phi;
    else if (su < 0 && sq > 0) phi = phi + 180;
    else if (su < 0 && sq < 0) phi = phi + 90;
    else if (su > 0 && sq < 0) phi = phi + 90;

    double B_aux = fabs(Blos / cos(gamma_rad));                      /* :435 */
    if (Vlos < (-4)) Vlos = -4;
    if (Vlos > (4)) Vlos = 4;
    model[ORC_B] = (B_aux > 4000 ? 4000 : B_aux);                    /* :444-447 */
    model[ORC_VLOS] = Vlos;
    model[ORC_GM] = gamma;
    model[ORC_AZ] = phi;
    if (!for_init) model[ORC_S0] = Blos;                            /* :449-450 */
}

void orc_classical(const orc_setup_t *s, const float *spectro, double model[ORC_NPAR])
{
    orc_classical_ex(s, spectro, model, 1);
}

/* Test builds only (-DORC_PERTURB=eps): every synthesised sample and
 * response value multiplied by 1 + eps * u, u uniform in [-1, 1) from a
 * fixed-seed generator -- an implementation that differs from this one by
 * rounding-level amounts everywhere.  The spread of the inversions under it
 * measures how far such differences alone carry the LM paths. */
#ifdef ORC_PERTURB
static unsigned long long perturb_state = 0x9E3779B97F4A7C15ull;
static void perturb(real *x, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        perturb_state ^= perturb_state << 13;
        perturb_state ^= perturb_state >> 7;
        perturb_state ^= perturb_state << 17;
        double u = (double)(perturb_state >> 11) * (2.0 / 9007199254740992.0) - 1.0;
        x[i] = x[i] * (1.0 + ORC_PERTURB * u);
    }
}
#define PERTURB(x, n) perturb((x), (n))
#else
#define PERTURB(x, n) ((void)0)
#endif

/* The step of mil_svd (milosUtils.c:237-271).  The reference takes it from
 * GSL's symmetric eigen-solver, which is absent here; the default stands in
 * for it with orc_pinv_solve.  Built with -DORC_SOLVE=1 (test builds only),
 * another method with the same result in exact arithmetic: LDL^T on the
 * active rows, with the pseudo-inverse when the factorisation fails or the
 * smallest eigenvalue may lie within 1000x of the 1e-12 cutoff (1/trace of
 * the inverse below 1e-9).  The spread between the two measures how far the
 * choice of a correct solver moves the inversions. */
#ifndef ORC_SOLVE
#define ORC_SOLVE 0
#endif
static void solve_step(int nt, const double *h, const double *beta, double *delta)
{
    if (ORC_SOLVE == 1) {
        double F[ORC_NPAR][ORC_NPAR], d[ORC_NPAR], y[ORC_NPAR];
        int act[ORC_NPAR], ok = 1;
        for (int i = 0; i < nt; i++) act[i] = h[i * nt + i] != 0.0;
        for (int i = 0; i < nt; i++)
            for (int j = 0; j < nt; j++)
                F[i][j] = (act[i] && act[j]) ? h[i * nt + j] : (i == j ? 1.0 : 0.0);
        for (int j = 0; j < nt && ok; j++) {
            double dj = F[j][j];
            for (int k = 0; k < j; k++) dj -= F[j][k] * F[j][k] * d[k];
            if (!(dj > 0.0)) { ok = 0; break; }
            d[j] = dj;
            for (int i = j + 1; i < nt; i++) {
                double v = F[i][j];
                for (int k = 0; k < j; k++) v -= F[i][k] * F[j][k] * d[k];
                F[i][j] = v / dj;
            }
        }
        if (ok) {
            /* trace of the inverse: sum over columns of L^-1 D^-1 L^-T */
            double Li[ORC_NPAR][ORC_NPAR], tr = 0;
            for (int c = 0; c < nt; c++) {
                for (int i = 0; i < nt; i++) {
                    double v = (i == c) ? 1.0 : 0.0;
                    for (int k = c; k < i; k++) v -= F[i][k] * Li[k][c];
                    Li[i][c] = i < c ? 0.0 : v;
                }
            }
            for (int i = 0; i < nt; i++)
                for (int k = 0; k <= i; k++)
                    if (act[i]) tr += Li[i][k] * Li[i][k] / d[i];
            if (tr < 1e9) {
                for (int i = 0; i < nt; i++) {
                    double v = act[i] ? beta[i] : 0.0;
                    for (int k = 0; k < i; k++) v -= F[i][k] * y[k];
                    y[i] = v;
                }
                for (int i = nt - 1; i >= 0; i--) {
                    double v = y[i] / d[i];
                    for (int k = i + 1; k < nt; k++) v -= F[k][i] * delta[k];
                    delta[i] = act[i] ? v : 0.0;
                }
                return;
            }
        }
    }
    orc_pinv_solve(nt, h, beta, delta);
}

/* lm_mils, milosUtils.c:466-635.  REAL: damping, chi2, factor, alpha,
 * beta, sigma, spectra, response functions; PRECISION: covar, delta, the
 * model and toplim (:475-533). */
int orc_lm(const orc_setup_t *s, const float *spectro, double model[ORC_NPAR],
           double *fit, float *chi2, int *iters, int *accepted)
{
    const int L = s->nlambda, NT = s->nterms;
    const real better = 5.0, worst = 10.0;                          /* :475-477 */
    real factor = 1.0;
    real sig[4 * ORC_MAXL];
    int n_ghosts = 0;
    for (int i = 0; i < 4 * L; i++) {                               /* :484-492 */
        if (spectro[i] <= -1) { sig[i] = 1000000000000000000000.0; n_ghosts++; }
        else sig[i] = s->noise;
    }
    int nfree = (L * 4) - NT;                                       /* lib.c:399-407 */
    nfree = nfree - n_ghosts;
    if (accepted) *accepted = 0;
    if (nfree == 0) return -1;                                      /* :509-512 */

    orc_ws_t *ws = (orc_ws_t *)malloc(sizeof(orc_ws_t));
    real *spectra = (real *)malloc(sizeof(real) * 4 * (size_t)L);
    real *rf = (real *)calloc(4 * (size_t)NT * (size_t)L, sizeof(real));
    real beta[ORC_NPAR], alpha[ORC_NPAR * ORC_NPAR];
    double covar[ORC_NPAR * ORC_NPAR], betad[ORC_NPAR];
    double delta[ORC_NPAR];
    real flambda = s->ilambda;
    int it = 0, nacc = 0, clanda = 0;

    synth_ws(s, model, spectra, ws);                                /* :536-539 */
    PERTURB(spectra, 4 * (size_t)L);
    rf_ws(s, model, ws, ws->spectra_mac, ws->spectra_slight, rf);
    PERTURB(rf, 4 * (size_t)NT * L);
    fija_cero_r(s, rf);
    covarm_r(s, sig, spectro, spectra, rf, beta, alpha);
    for (int i = 0; i < NT * NT; i++) covar[i] = alpha[i];
    real ochisqr = chisqr_r(s, spectra, spectro, sig, nfree);       /* :546-547 */
    real chisqr0 = ochisqr;
    double trial[ORC_NPAR];
    memcpy(trial, model, sizeof(trial));                            /* :549 */

    do {
        for (int i = 0; i < NT; i++) {                              /* :559-563 */
            int ind = i * (NT + 1);
            covar[ind] = alpha[ind] * (1.0 + flambda);
        }
        for (int i = 0; i < NT; i++) betad[i] = beta[i];
        solve_step(NT, covar, betad, delta);                         /* :565 */
        orc_aplica_delta(s, model, delta, trial);                    /* :566 */
        orc_check(trial);                                            /* :568 */
        synth_ws(s, trial, spectra, ws);                             /* :569 */
        PERTURB(spectra, 4 * (size_t)L);
        real chisqr = chisqr_r(s, spectra, spectro, sig, nfree);
        if ((r_fabs((ochisqr - chisqr) * 100 / chisqr) < s->toplim) || (chisqr < 0.0001))
            clanda = 1;                                              /* :579-580 */
        if (chisqr - ochisqr < 0.) {                                 /* :581-606 */
            flambda = flambda / (better * factor);
            memcpy(model, trial, sizeof(trial));
            rf_ws(s, model, ws, ws->spectra_mac, spectra, rf);      /* args as :587 */
            PERTURB(rf, 4 * (size_t)NT * L);
            fija_cero_r(s, rf);
            covarm_r(s, sig, spectro, spectra, rf, beta, alpha);
            for (int i = 0; i < NT * NT; i++) covar[i] = alpha[i];
            ochisqr = chisqr;
            nacc++;
        } else {                                                     /* :607-615 */
            for (int i = 0; i < NT * NT; i++) covar[i] = alpha[i];
            flambda = flambda * worst * factor;
        }
        if ((flambda > 1e+7) || (flambda < 1e-25)) clanda = 1;       /* :617-618 */
        it++;
        if (s->logclambda)                                           /* :621 */
            factor = (real)(log10f((float)chisqr) / log10f((float)chisqr0));
    } while (it < s->miter && !clanda);

    if (chi2) *chi2 = (float)ochisqr;                                /* :625 */
    if (iters) *iters = it;
    if (accepted) *accepted = nacc;
    if (fit)
        for (int i = 0; i < 4 * L; i++) fit[i] = spectra[i];
    free(rf); free(spectra); free(ws);
    return 1;
}

/* Pixel loop of the FITS inversion, milos.c:958-1009. */
void orc_invert_range(const orc_setup_t *s, const float *stokes, int p0, int p1,
                      const double guess[ORC_NPAR], double *models, float *chi2,
                      int *iters, int *accepted, float *fit)
{
    const int L = s->nlambda;
    double *fitd = (double *)malloc(sizeof(double) * 4 * (size_t)L);
    for (int p = p0; p < p1; p++) {
        const float *spec = stokes + (size_t)p * 4 * L;
        double m[ORC_NPAR];
        memcpy(m, guess, sizeof(m));                                 /* :961-972 */
        orc_classical(s, spec, m);                                   /* :975 */
        if (isnan(m[ORC_B])) m[ORC_B] = 1;                           /* :976-983 */
        if (isnan(m[ORC_VLOS])) m[ORC_VLOS] = 1e-3;
        if (isnan(m[ORC_GM])) m[ORC_GM] = 1;
        if (isnan(m[ORC_AZ])) m[ORC_AZ] = 1;
        float c2 = 0.0f;
        int it = p, acc = 0;      /* iters holds the pixel index if lm returns -1 (:997) */
        int rc = orc_lm(s, spec, m, fitd, &c2, &it, &acc);
        memcpy(models + (size_t)p * ORC_NPAR, m, sizeof(m));         /* :1001 */
        if (chi2) chi2[p] = c2;
        if (iters) iters[p] = it;
        if (accepted) accepted[p] = acc;
        if (fit) {
            if (rc == 1)
                for (int k = 0; k < 4 * L; k++) fit[(size_t)p * 4 * L + k] = (float)fitd[k];
            else
                for (int k = 0; k < 4 * L; k++) fit[(size_t)p * 4 * L + k] = 0.0f;
        }
    }
    free(fitd);
}
