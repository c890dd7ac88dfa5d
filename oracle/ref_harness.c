/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Links four numeric files of the reference, compiled unchanged from where
 * they lie (fvoigt.c, create_cuantic.c, fgauss.c, convolution.c), into
 * oracle/_ref/libref_core.so so tests can check the oracle against the
 * reference's own code.  The reference keeps its scratch buffers in globals
 * defined by calculosCompartidos.c/milos.c (which need FFTW and GSL, absent
 * here); this file defines and allocates those few globals and exposes flat
 * wrappers callable through ctypes.
 */
#include <stdlib.h>
#include <string.h>
#include "defines.h"
#include "convolution.h"

_Complex double *z, *zden, *zdiv;       /* fvoigt.c:4 */
PRECISION *GMAC;                          /* fgauss.c:5 */
PRECISION *dirConvPar;                    /* convolution.c:9 */
REAL *resultConv;                         /* convolution.c:10 */

static int cap = 0;

static void ensure(int n)
{
    if (n <= cap) return;
    free(z); free(zden); free(zdiv); free(GMAC); free(dirConvPar); free(resultConv);
    z = malloc(n * sizeof(_Complex double));
    zden = malloc(n * sizeof(_Complex double));
    zdiv = malloc(n * sizeof(_Complex double));
    GMAC = calloc(n, sizeof(PRECISION));
    dirConvPar = calloc(2 * n + 1, sizeof(PRECISION));   /* calculosCompartidos.c:57 */
    resultConv = calloc(n, sizeof(REAL));
    cap = n;
}

int ref_fvoigt(double damp, double *v, int n, double *h, double *f)
{
    ensure(n);
    return fvoigt(damp, v, n, h, f);
}

/* Flattened Cuantic: out[0..3] = n_pi, n_sig, gl, gu; out[4] = geff;
 * then nup[n_pi], wep[n_pi], nub[n_sig], web[n_sig], nur[n_sig], wer[n_sig]. */
int ref_cuantic(double sl, double ll, double jl, double su, double lu, double ju,
                double *out, int nout)
{
    PRECISION dat[7] = {1, sl, ll, jl, su, lu, ju};
    Cuantic *c = create_cuantic(dat, 0);
    int np = (int)c->N_PI, ns = (int)c->N_SIG;
    int need = 5 + 2 * np + 4 * ns;
    if (need > nout) return -1;
    out[0] = np; out[1] = ns; out[2] = c->GL; out[3] = c->GU; out[4] = c->GEFF;
    double *o = out + 5;
    for (int i = 0; i < np; i++) *o++ = c->NUP[i];
    for (int i = 0; i < np; i++) *o++ = c->WEP[i];
    for (int i = 0; i < ns; i++) *o++ = c->NUB[i];
    for (int i = 0; i < ns; i++) *o++ = c->WEB[i];
    for (int i = 0; i < ns; i++) *o++ = c->NUR[i];
    for (int i = 0; i < ns; i++) *o++ = c->WER[i];
    return need;
}

int ref_fgauss_wl(double fwhm, double step, double l0, double lc, int n, double *g)
{
    int size = 0;
    PRECISION *G = fgauss_WL(fwhm, step, l0, lc, n, &size);
    memcpy(g, G, sizeof(double) * n);
    free(G);
    return size;
}

void ref_fgauss(double mc, double *eje, int n, double landa, int deriv, double *out)
{
    ensure(n);
    fgauss(mc, eje, n, landa, deriv);
    memcpy(out, GMAC, sizeof(double) * n);
}

void ref_direct(double *x, int n, double *h, int nh)
{
    ensure(n > nh ? n : nh);
    memset(dirConvPar, 0, sizeof(double) * (2 * cap + 1));
    direct_convolution(x, n, h, nh);
}

void ref_direct_double(double *x, int n, double *h, int nh)
{
    ensure(n > nh ? n : nh);
    memset(dirConvPar, 0, sizeof(double) * (2 * cap + 1));
    direct_convolution_double(x, n, h, nh);
}

void ref_direct_ic(double *x, int n, double *h, int nh, double ic)
{
    ensure(n > nh ? n : nh);
    memset(dirConvPar, 0, sizeof(double) * (2 * cap + 1));
    direct_convolution_ic(x, n, h, nh, ic);
}

void ref_circular(double *x, double *h, int n, double *out)
{
    ensure(n);
    convCircular(x, h, n, out);
}
