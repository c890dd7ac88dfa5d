// pm_setup.h -- the per-context constants (PmConst) from a pm_config:
// the Zeeman pattern of the line, the fixed-parameter masks, the wavelength
// grid and the PSF tap windows.  Host code with no HIP calls; pm_ctx_create
// uploads the result.
#pragma once

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>

#include "pm_device.h"
#include "../../include/pmilos.h"

// The reference's NTERMS from the free parameters (milos.c:154-157).
static inline int nterms_of(const int fix[PM_NPAR], double guess_mac)
{
    int nt = 11;
    if (fix[10] == 0) nt--;
    if (guess_mac == 0 && fix[9] == 0) nt--;
    return nt;
}

// Zeeman pattern of the line: Lande factors, sublevels, strengths and shifts
// per Delta-M group, normalised per group (create_cuantic.c:60-234; M values
// built from (int)J as there).
static inline int zeeman_setup(PmConst &c, double sl, double ll, double jl, double su,
                               double lu, double ju)
{
    double g1 = jl != 0 ? (3.0 * jl * (jl + 1) + sl * (sl + 1) - ll * (ll + 1)) / (2 * jl * (jl + 1)) : 0.0;
    double g2 = ju != 0 ? (3.0 * ju * (ju + 1.0) + su * (su + 1.0) - lu * (lu + 1.0)) / (2.0 * ju * (ju + 1.0)) : 0.0;
    c.geff = (g1 + g2) / 2 + (g1 - g2) * (jl * (jl + 1.) - ju * (ju + 1.)) / 4;
    int npi = (int)((2 * (jl < ju ? jl : ju)) + 1), nsig = (int)(jl + ju);
    if (npi < 1 || nsig < 1 || npi > PM_MAXC || nsig > PM_MAXC) return PM_ENOTSUP;
    int jj = (int)(ju - jl);
    int ip = 0, ib = 0, ir = 0;
    double wer_tmp[PM_MAXC];
    for (int j = 0; j <= 2 * jl; j++) {
        int ml = j - (int)jl;
        for (int i = 0; i <= 2 * ju; i++) {
            int mu = i - (int)ju;
            int dm = mu - ml;
            double m = ml;
            double shift = g1 * ml - g2 * mu;
            double s = 0;
            if (dm == 0) {
                if (jj == -1) s = jl * jl - m * m;
                else if (jj == 0) s = m * m;
                else if (jj == 1) s = (jl + 1) * (jl + 1) - m * m;
                if (ip >= PM_MAXC) return PM_ENOTSUP;
                c.wep[ip] = s; c.nup[ip] = shift; ip++;
            } else if (dm == 1) {
                if (jj == -1) s = (jl - m) * (jl - m - 1) / 4;
                else if (jj == 0) s = (jl - m) * (jl + m + 1) / 4;
                else if (jj == 1) s = (jl + m + 1) * (jl + m + 2) / 4;
                if (ib >= PM_MAXC) return PM_ENOTSUP;
                c.web[ib] = s; c.nub[ib] = shift; ib++;
            } else if (dm == -1) {
                if (jj == -1) s = (jl + m) * (jl + m - 1) / 4;
                else if (jj == 0) s = (jl + m) * (jl - m + 1) / 4;
                else if (jj == 1) s = (jl - m + 1) * (jl - m + 2) / 4;
                if (ir >= PM_MAXC) return PM_ENOTSUP;
                wer_tmp[ir] = s; ir++;
            }
        }
    }
    if (ip != npi || ib != nsig || ir != nsig) return PM_ENOTSUP;
    double sp = 0, sb = 0, sr = 0;
    for (int i = 0; i < npi; i++) sp = sp + c.wep[i];
    for (int i = 0; i < npi; i++) c.wep[i] = c.wep[i] / sp;
    for (int i = 0; i < nsig; i++) sb = sb + c.web[i];
    for (int i = 0; i < nsig; i++) c.web[i] = c.web[i] / sb;
    for (int i = 0; i < nsig; i++) sr = sr + wer_tmp[i];
    for (int i = 0; i < nsig; i++) c.wer[i] = wer_tmp[i] / sr;
    c.npi = npi;
    c.nsig = nsig;
    return PM_OK;
}

// Fills c from cfg (whose lambda, psf and slight arrays have nlambda and
// 4 * nlambda entries); PM_ENOTSUP for a line without a component in some
// group or with more than PM_MAXC.  Leaves the device-side fields of the
// context to the caller.
static inline int pm_const_build(const pm_config *cfg, PmConst &c)
{
    const int L = cfg->nlambda;
    memset(&c, 0, sizeof(c));
    const int rc = zeeman_setup(c, cfg->sl, cfg->ll, cfg->jl, cfg->su, cfg->lu, cfg->ju);
    if (rc != PM_OK) return rc;
    c.L = L;
    c.NT = nterms_of(cfg->fix, cfg->guess[9]);
    c.no_classical = cfg->no_classical ? 1 : 0;
    c.use_fft = cfg->use_fft ? 1 : 0;
    if (cfg->slight) {
        c.use_slight = 1;
        for (int i = 0; i < 4 * L; i++) c.slight[i] = cfg->slight[i];
    }
    c.use_psf = cfg->psf ? 1 : 0;
    c.miter = cfg->miter;
    c.logclambda = cfg->logclambda;
    for (int i = 0; i < 11; i++) { c.fix[i] = cfg->fix[i]; c.guess[i] = cfg->guess[i]; }
    c.lambda0 = cfg->lambda0;
    c.noise = cfg->noise;
    c.inv_noise = 1.0 / cfg->noise;
    c.ilambda = cfg->ilambda;
    c.toplim = cfg->toplim;
    c.toplim_mixed = cfg->precision == PM_MIXED ? PM_MIXED_TOPLIM : 0.0;
    c.mu = cfg->mu;
    for (int i = 0; i < 4; i++) {
        c.weight[i] = cfg->weight[i];
        c.swt[0][i] = sqrt(c.weight[i] * c.inv_noise);
        c.swt[1][i] = sqrt(c.weight[i] * 1e-21);
    }
    // J^T J entries of fixed parameters, zeroed after the sums: the rows of
    // FijaACeroDerivadasNoNecesarias (milosUtils.c:199-227); in the 9-, 11-
    // and 10-row layouts row p is fixed when fix[p] == 0
    for (int lay = 0; lay < 3; lay++) {
        const int nt = lay == 0 ? 9 : lay == 1 ? 11 : 10, np = nt * (nt + 1) / 2;
        for (int i = 0; i < nt; i++) {
            const bool fi = c.fix[i] == 0;
            for (int h = 0; h <= i; h++) {
                const int e = i * (i + 1) / 2 + h;
                if (fi || c.fix[h] == 0) c.zbits[lay][e >> 5] |= 1u << (e & 31);
            }
            if (fi) c.zbits[lay][(np + i) >> 5] |= 1u << ((np + i) & 31);
        }
    }
    for (int i = 0; i < L; i++) {
        c.lam[i] = cfg->lambda[i];
        c.dlam[i] = cfg->lambda[i] - cfg->lambda0;
        c.G[i] = cfg->psf ? cfg->psf[i] : 0.0;
    }
    // PSF tap ranges per form (see PmConst)
    {
        double gmax = 0;
        for (int i = 0; i < L; i++) gmax = fmax(gmax, fabs(c.G[i]));
        int lo64 = L, hi64 = 0, lo32 = L, hi32 = 0;
        for (int i = 0; i < L; i++) {
            double r = gmax > 0 ? fabs(c.G[i]) / gmax : 0;
            if (r >= 1e-16) { lo64 = lo64 < i ? lo64 : i; hi64 = i + 1; }
            if (r >= 1e-8) { lo32 = lo32 < i ? lo32 : i; hi32 = i + 1; }
        }
        if (!c.use_psf) { lo64 = lo32 = 0; hi64 = hi32 = 0; }
        c.glo64 = lo64; c.ghi64 = hi64; c.glo32 = lo32; c.ghi32 = hi32;
        const int half = L / 2;
        c.reach64 = c.reach32 = 0;
        for (int d = -PM_HC; d <= PM_HC; d++) {
            const int j = half + d;
            c.gc64[PM_HC + d] = (j >= lo64 && j < hi64) ? c.G[j] : 0.0;
            c.gc32[PM_HC + d] = (j >= lo32 && j < hi32) ? (float)c.G[j] : 0.0f;
        }
        for (int j = lo64; j < hi64; j++) c.reach64 = std::max(c.reach64, std::abs(j - half));
        for (int j = lo32; j < hi32; j++) c.reach32 = std::max(c.reach32, std::abs(j - half));
        c.reach64 = std::min(c.reach64, PM_HC + 1);
        c.reach32 = std::min(c.reach32, PM_HC + 1);
    }
    return PM_OK;
}
