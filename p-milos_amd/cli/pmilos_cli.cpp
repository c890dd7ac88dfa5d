// pmilos_cli.cpp -- the command line of the MI355X inversion engine, a
// drop-in for the reference's two programs on one GPU:
//
//   pmilos_gpu file.mtrol   as milos.x  (milos.c:117-1099): synthesis
//                           (NumberOfCycles 0), classical estimates only
//                           (< 0), or inversion (> 0) of a .per profile or a
//                           FITS cube
//   pmilos_gpu file.minit   as pmilos.x (pmilos.c): a FITS cube, a numbered
//                           time series t1..t2, every numbered cube of the
//                           directory (t1 = *) and, with t2 = *, new cubes
//                           as they appear; subfield, mask, output prefix
//
// Options: --mixed (FP32 spectra/RFs, FP64 solve), --mixed-toplim T (the
// mixed form's chi2-unchanged threshold, default PM_MIXED_TOPLIM; 0 = the
// reference's toplim test), --autotune (pick the kernel shape on the first
// cube), --watch-timeout S (default 600, the reference's TIMEOUT_FILE),
// --watch-poll S (default 5).
//
// Inputs, outputs and file names follow the reference; the differences
// (all deliberate) are listed in usage().  Every pixel is inverted by
// libpmilos_hip.so through its C ABI; a time series goes through the
// serving loop, so the next cube is read and uploaded while the GPU
// inverts the previous one.
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pmilos.h"
#include "pm_io.h"

namespace {

struct Opts {
    int precision = PM_FP64;
    double mixed_toplim = PM_MIXED_TOPLIM;
    bool autotune = false;
    double watch_timeout = 600.0;
    double watch_poll = 5.0;
};

void usage()
{
    fprintf(stderr,
            "usage: pmilos_gpu <file.mtrol | file.minit> [--mixed] [--mixed-toplim T]\n"
            "                  [--autotune] [--watch-timeout S] [--watch-poll S]\n"
            "  --mixed-toplim T: the mixed form ends its LM loop when chi2 changes by\n"
            "  less than T/100 relative (default %g; 0 = the reference's toplim test\n"
            "  alone, as milos.x)\n"
            "Deliberate differences from milos.x / pmilos.x:\n"
            "  FITS cubes in any axis order are read by their CTYPEs (the reference's\n"
            "  fast path scrambles HPLN,HPLT,WAVE,STOKES cubes); a blank output prefix\n"
            "  is no prefix (not \"!\"); PSF samples outside a .psf file's range are 0;\n"
            "  the directory watch skips every name containing _mod or _stokes.\n",
            PM_MIXED_TOPLIM);
}

std::string no_ext(const std::string &p)
{
    // get_basefilename (readConfig.c:1257): the path minus its extension
    size_t dot = p.rfind('.');
    size_t slash = p.rfind('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return p;
    return p.substr(0, dot);
}

std::string base_of(const std::string &p)
{
    size_t slash = p.rfind('/');
    return slash == std::string::npos ? p : p.substr(slash + 1);
}

std::string dir_of(const std::string &p)
{
    size_t slash = p.rfind('/');
    return slash == std::string::npos ? "." : p.substr(0, slash);
}

std::string ext_of(const std::string &p)
{
    size_t dot = p.rfind('.');
    if (dot == std::string::npos) return "";
    for (size_t i = dot + 1; i < p.size(); i++)
        if (!isalnum((unsigned char)p[i])) return "";
    return p.substr(dot);
}

bool readable(const std::string &p) { return !p.empty() && access(p.c_str(), F_OK) == 0; }

std::string num3(int i)
{
    // the reference's index suffix: 00i, 0i, i (pmilos.c:788-799)
    char b[32];
    if (i >= 0 && i < 10) snprintf(b, sizeof b, "00%d", i);
    else if (i >= 10 && i < 100) snprintf(b, sizeof b, "0%d", i);
    else snprintf(b, sizeof b, "%d", i);
    return b;
}

int fail(const char *what, const std::string &detail, int rc = 1)
{
    fprintf(stderr, "pmilos_gpu: %s%s%s\n", what, detail.empty() ? "" : ": ", detail.c_str());
    return rc;
}

// The per-run setup of milos.c:133-392: model, wavelengths, line, PSF.
struct Setup {
    pmio_control c;
    double guess[11] = {0};
    int L = 0, line_index = 0;
    double lambda0 = 0;
    double dat[7] = {0};
    std::vector<double> lam, off;        // wavelengths (A) and offsets from the line centre (mA)
    std::vector<double> G;               // PSF on the grid, when used
};

int setup(Setup &s)
{
    pmio_control &c = s.c;
    if (c.guess[0]) {
        int rc = pmio_read_model(c.guess, s.guess);
        if (rc < 0) return fail("cannot read the initial model", c.guess);
        if (rc > 0) return fail("initial model parameter out of range, index", std::to_string(rc - 1));
    }
    if (s.guess[10] < 1 && !readable(c.straylight))
        return fail("filling factor < 1 and no stray-light file", c.straylight);
    if (c.use_grid) {
        double a, st, b;
        s.line_index = pmio_read_grid(c.grid, &a, &st, &b);
        if (!s.line_index) return fail("cannot read the wavelength grid", c.grid);
        s.lambda0 = pmio_read_lines(c.lines, s.line_index, s.dat);
        if (s.lambda0 == 0) return fail("line not found in", c.lines);
        int n = pm_grid(s.lambda0, a, st, b, nullptr, 0);
        if (n < 2 || n > PM_MAX_LAMBDA) return fail("unsupported number of wavelengths", std::to_string(n));
        s.L = n;
        s.lam.resize(n);
        pm_grid(s.lambda0, a, st, b, s.lam.data(), n);
        s.off.resize(n);
        s.off[0] = a;                                             // milos.c:176-179
        for (int i = 1; i < n; i++) s.off[i] = s.off[i - 1] + st;
    } else {
        double *l;
        int n, idx;
        if (pmio_fits_read_lambda(c.wavelength, &l, &n, &idx))
            return fail("cannot read the wavelength file", c.wavelength);
        s.line_index = idx;
        s.lambda0 = pmio_read_lines(c.lines, idx, s.dat);
        if (s.lambda0 == 0) { pmio_free(l); return fail("line not found in", c.lines); }
        if (n < 2 || n > PM_MAX_LAMBDA) { pmio_free(l); return fail("unsupported number of wavelengths", std::to_string(n)); }
        s.L = n;
        s.lam.assign(l, l + n);
        pmio_free(l);
        s.off.resize(n);
        for (int i = 0; i < n; i++) s.off[i] = (s.lam[i] - s.lambda0) * 1000.0;
    }
    if (c.use_psf) {
        s.G.resize(s.L);
        if (!c.psf_is_file) {
            pm_psf_gaussian(c.fwhm, s.lam.data(), s.L, s.G.data());
            FILE *f = fopen("gaussian.psf", "w");                   // milos.c:283-297
            if (f) {
                for (int k = 0; k < s.L; k++)
                    fprintf(f, "\t%f\t%e\n", (s.lam[k] - s.lambda0) * 1000, s.G[k]);
                fclose(f);
            }
        } else {
            std::vector<double> dl(100000), ps(100000);
            int n = pmio_read_psf_file(c.psf, dl.data(), ps.data(), (int)dl.size());
            if (n < 2) return fail("cannot read the PSF file", c.psf);
            if (trunc(s.off[0]) < trunc(dl[0]) || trunc(s.off[s.L - 1]) > trunc(dl[n - 1]))
                return fail("the PSF file covers less than the wavelength grid", c.psf);
            // recentre on the grid's middle sample (milos.c:331-340)
            int pos = 0;
            for (int i = 0; i < s.L && !pos; i++)
                if (fabs(trunc(s.off[i])) == 0) pos = i;
            double shift = 0;
            if (pos != s.L / 2) shift = ((s.L / 2 - pos) * ((s.off[1] - s.off[0]) / 1000.0)) * 1000;
            if (pmio_interp_psf(dl.data(), ps.data(), n, s.off.data(), s.L, shift, s.G.data()))
                return fail("the PSF file has no weight on the grid", c.psf);
        }
    }
    return 0;
}

pm_ctx *make_ctx(const Setup &s, const Opts &o, const float *slight, bool classical, int miter)
{
    pm_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.nlambda = s.L;
    cfg.lambda = s.lam.data();
    cfg.lambda0 = s.lambda0;
    cfg.sl = s.dat[1]; cfg.ll = s.dat[2]; cfg.jl = s.dat[3];
    cfg.su = s.dat[4]; cfg.lu = s.dat[5]; cfg.ju = s.dat[6];
    cfg.psf = s.c.use_psf ? s.G.data() : nullptr;
    for (int i = 0; i < 11; i++) { cfg.fix[i] = s.c.fix[i]; cfg.guess[i] = s.guess[i]; }
    for (int i = 0; i < 4; i++) cfg.weight[i] = s.c.weight[i];
    cfg.noise = s.c.noise;
    cfg.ilambda = s.c.ilambda;
    cfg.toplim = 1e-12;
    cfg.miter = miter > 0 ? miter : 1;
    cfg.mu = s.c.mu;
    cfg.logclambda = s.c.logclambda;
    cfg.precision = o.precision;
    cfg.slight = slight;
    cfg.no_classical = classical ? 0 : 1;
    cfg.use_fft = s.c.use_fft ? 1 : 0;
    pm_ctx *ctx = nullptr;
    int rc = pm_ctx_create(&cfg, &ctx);
    if (rc) {
        fail("context creation failed", pm_strerror(rc));
        return nullptr;
    }
    if (o.precision == PM_MIXED && (rc = pm_ctx_set_mixed_toplim(ctx, o.mixed_toplim))) {
        fail("--mixed-toplim", pm_strerror(rc));
        pm_ctx_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

// A stray-light file: a .per profile, a 2-D FITS profile or (cube == true)
// a 4-D FITS cube of per-pixel profiles.
struct Slight {
    std::vector<float> v;
    bool per_pixel = false;
    long npix = 0;
};

int read_slight(const Setup &s, bool cube_ok, int x1, int x2, int y1, int y2, Slight &out)
{
    const std::string p = s.c.straylight;
    out = Slight();
    if (ext_of(p) == ".per") {
        out.v.resize(4 * s.L);
        std::vector<double> off(s.L);
        int n = pmio_read_per(p.c_str(), s.L, out.v.data(), off.data());
        if (n < s.L) return fail("cannot read the stray-light profile", p);
        for (int i = 0; i < s.L; i++)                             // utilsFits.c:857-860
            if (off[i] != trunc(s.off[i]))
                return fail("stray-light wavelengths differ from the grid", p);
        out.npix = 1;
        return 0;
    }
    if (ext_of(p) == ".fits") {
        float *d;
        int pp;
        long n;
        if (pmio_fits_read_straylight(p.c_str(), s.L, x1, x2, y1, y2, &d, &pp, &n))
            return fail("cannot read the stray-light FITS", p);
        if (pp && !cube_ok) { pmio_free(d); out.v.clear(); return 0; }   // milos.c:531-536: ignored
        out.v.assign(d, d + (size_t)n * 4 * s.L);
        pmio_free(d);
        out.per_pixel = pp != 0;
        out.npix = n;
        return 0;
    }
    return fail("unknown stray-light file type", p);
}

// ------------------------------------------------------------ milos.x

int run_milos(Setup &s, const Opts &o)
{
    pmio_control &c = s.c;
    const int L = s.L;
    const std::string obs = c.observed;
    const std::string stem = no_ext(obs.empty() ? std::string(c.guess) : obs);

    if (c.cycles < 0) {                                           // milos.c:399-521
        pm_ctx *ctx = make_ctx(s, o, nullptr, true, 1);
        if (!ctx) return 1;
        int rc = 0;
        if (ext_of(obs) == ".per") {
            std::vector<float> sp(4 * L);
            if (pmio_read_per(obs.c_str(), L, sp.data(), nullptr) != L) rc = fail("profile length differs from the grid", obs);
            double m[11];
            if (!rc && pm_classical(ctx, sp.data(), 1, m) == 0)
                rc = pmio_write_model_text((stem + "_model_ce.mod").c_str(), m, 0, 0.0, 1);
        } else if (ext_of(obs) == ".fits") {
            float *cube;
            pmio_cube_layout lay;
            if (pmio_fits_read_cube(obs.c_str(), L, 0, 0, 0, 0, &cube, &lay)) {
                rc = fail("cannot read the Stokes cube", obs);
            } else {
                const size_t n = (size_t)lay.rows * lay.cols;
                std::vector<double> m(n * 11);
                rc = pm_classical(ctx, cube, (int)n, m.data());
                if (!rc) rc = pmio_fits_write_models((stem + "_model_ce.fits").c_str(), lay.rows, lay.cols, m.data(), nullptr, nullptr);
                pmio_free(cube);
                pmio_free_layout(&lay);
            }
        } else {
            rc = fail("unknown observed-profiles type", obs);
        }
        pm_ctx_destroy(ctx);
        return rc ? 1 : 0;
    }

    if (c.cycles == 0) {                                          // milos.c:522-651
        Slight sl;
        if (readable(c.straylight) && read_slight(s, false, 0, 0, 0, 0, sl)) return 1;
        pm_ctx *ctx = make_ctx(s, o, sl.v.empty() ? nullptr : sl.v.data(), true, 1);
        if (!ctx) return 1;
        std::vector<double> spec(4 * L);
        int rc = pm_synth(ctx, s.guess, 1, spec.data(), nullptr);
        if (!rc) {
            std::vector<double> off(L);
            for (int k = 0; k < L; k++) off[k] = (s.lam[k] - s.lambda0) * 1000;
            rc = pmio_write_per((stem + ".per").c_str(), s.line_index, off.data(), spec.data(), L);
        }
        pm_ctx_destroy(ctx);
        if (rc) return fail("synthesis failed", "");
        printf("synthesis done: %s.per\n", stem.c_str());
        return 0;
    }

    // inversion
    Slight sl;
    if (c.fix[10] && readable(c.straylight) &&
        read_slight(s, ext_of(obs) == ".fits", 0, 0, 0, 0, sl))
        return 1;
    const float *profile = (!sl.v.empty() && !sl.per_pixel) ? sl.v.data() : nullptr;
    if (ext_of(obs) == ".per") {                                  // milos.c:654-840
        std::vector<float> sp(4 * L);
        if (pmio_read_per(obs.c_str(), L, sp.data(), nullptr) != L)
            return fail("profile length differs from the grid", obs);
        pm_ctx *ctx = make_ctx(s, o, profile, false, c.cycles);
        if (!ctx) return 1;
        double m[11];
        float chi2, fit[4 * PM_MAX_LAMBDA];
        int it = 0;
        int rc = pm_invert(ctx, sp.data(), 1, m, &chi2, &it, nullptr, fit);
        pm_ctx_destroy(ctx);
        if (rc) return fail("inversion failed", pm_strerror(rc));
        pmio_write_model_text((stem + "_model.mod").c_str(), m, it, chi2, 0);
        if (c.save_synthesis) {
            std::vector<double> f(fit, fit + 4 * L), off(L);
            for (int k = 0; k < L; k++) off[k] = (s.lam[k] - s.lambda0) * 1000;
            pmio_write_per((stem + "_stokes.per").c_str(), s.line_index, off.data(), f.data(), L);
        }
        printf("inverted model: %s_model.mod (%d iterations, chi2 %e)\n", stem.c_str(), it, chi2);
        return 0;
    }
    if (ext_of(obs) != ".fits") return fail("unknown observed-profiles type", obs);
    float *cube;                                                  // milos.c:841-1052
    pmio_cube_layout lay;
    if (pmio_fits_read_cube(obs.c_str(), L, 0, 0, 0, 0, &cube, &lay))
        return fail("cannot read the Stokes cube", obs);
    const size_t n = (size_t)lay.rows * lay.cols;
    if (sl.per_pixel && (size_t)sl.npix != n) {
        pmio_free(cube);
        pmio_free_layout(&lay);
        return fail("stray-light cube and Stokes cube differ in size", c.straylight);
    }
    pm_ctx *ctx = make_ctx(s, o, profile, true, c.cycles);
    if (!ctx) return 1;
    if (o.autotune) {
        int shape[3];
        if (pm_ctx_autotune_host(ctx, cube, sl.per_pixel ? sl.v.data() : nullptr,
                                 (int)std::min<size_t>(n, 65536), shape) == 0)
            printf("kernel shape: %d lanes x %d samples, %d pixels per workgroup\n", shape[0],
                   shape[1], shape[2]);
    }
    std::vector<double> m(n * 11);
    std::vector<float> chi2(n), fit(c.save_synthesis ? n * 4 * L : 0);
    std::vector<int> it(n);
    int rc = pm_invert_sl(ctx, cube, sl.per_pixel ? sl.v.data() : nullptr, (int)n, m.data(),
                          chi2.data(), it.data(), nullptr, c.save_synthesis ? fit.data() : nullptr);
    pm_ctx_destroy(ctx);
    if (!rc) rc = pmio_fits_write_models((stem + "_mod.fits").c_str(), lay.rows, lay.cols, m.data(), chi2.data(), it.data());
    if (!rc && c.save_synthesis) rc = pmio_fits_write_cube((stem + "_stokes.fits").c_str(), &lay, fit.data());
    pmio_free(cube);
    pmio_free_layout(&lay);
    if (rc) return fail("inversion or output failed", "");
    printf("inverted %zu pixels: %s_mod.fits\n", n, stem.c_str());
    return 0;
}

// ------------------------------------------------------------ pmilos.x

struct Job {
    std::string in, out_mod, out_stokes;
};

// Numbers of the cubes "<observed>NNN.fits" present in the directory.
std::vector<int> numbered(const std::string &observed)
{
    std::vector<int> nums;
    const std::string dir = dir_of(observed), stem = base_of(observed);
    DIR *d = opendir(dir.c_str());
    if (!d) return nums;
    while (struct dirent *e = readdir(d)) {
        std::string name = e->d_name;
        if (name.rfind(stem, 0) != 0 || ext_of(name) != ".fits") continue;
        std::string mid = no_ext(name).substr(stem.size());
        if (mid.empty() || mid.find_first_not_of("0123456789") != std::string::npos) continue;
        nums.push_back(atoi(mid.c_str()));
    }
    closedir(d);
    std::sort(nums.begin(), nums.end());
    return nums;
}

Job job_for(const pmio_control &c, int i)
{
    // pmilos.c:788-820
    const std::string obs = c.observed, idx = num3(i);
    const std::string out = std::string(c.prefix) + base_of(obs);
    return {obs + idx + ".fits", out + idx + "_mod.fits", out + idx + "_stokes.fits"};
}

struct Inflight {
    Job job;
    long long ticket = 0;
    int rows = 0, cols = 0;
    std::vector<int> keep;               // unmasked pixel indices (all when no mask)
    std::vector<float> stokes, fitbuf;
    std::vector<double> m;
    std::vector<float> chi2;
    std::vector<int> it, acc;
    pmio_cube_layout lay{};
    bool ok = false;
};

int run_pmilos(Setup &s, const Opts &o)
{
    pmio_control &c = s.c;
    const int L = s.L;
    std::vector<Job> jobs;
    if (c.invert_directory) {                                     // t1 = *: every numbered cube
        for (int i : numbered(c.observed)) jobs.push_back(job_for(c, i));
    } else if (c.loop_inversion) {                                // from t1 up to the last present
        for (int i : numbered(c.observed))
            if (i >= c.t1) jobs.push_back(job_for(c, i));
    } else if (c.t1 == 0 && c.t2 == 0) {                          // one cube, pmilos.c:763-771
        jobs.push_back({c.observed, no_ext(c.guess) + "_mod.fits", no_ext(c.observed) + "_stokes.fits"});
    } else {
        for (int i = c.t1; i <= c.t2; i++) jobs.push_back(job_for(c, i));
    }
    const bool sub = c.subx1 || c.subx2 || c.suby1 || c.suby2;
    // mask (pmilos.c:348-356): 0 = not inverted
    std::vector<double> mask;
    long mr = 0, mc = 0;
    if (readable(c.mask)) {
        double *d;
        if (pmio_fits_read_2d(c.mask, &d, &mr, &mc)) return fail("cannot read the mask", c.mask);
        mask.assign(d, d + mr * mc);
        pmio_free(d);
        if (sub) {
            if (c.subx2 > mr || c.suby2 > mc) return fail("subfield outside the mask", c.mask);
            std::vector<double> cut;
            for (long y = c.suby1 - 1; y < c.suby2; y++)
                for (long x = c.subx1 - 1; x < c.subx2; x++) cut.push_back(mask[y * mr + x]);
            mask.swap(cut);
            mr = c.subx2 - c.subx1 + 1;
            mc = c.suby2 - c.suby1 + 1;
        }
    }
    Slight sl;
    if (c.fix[10] && readable(c.straylight) &&
        read_slight(s, true, c.subx1, c.subx2, c.suby1, c.suby2, sl))
        return 1;
    pm_ctx *ctx = make_ctx(s, o, (!sl.v.empty() && !sl.per_pixel) ? sl.v.data() : nullptr, true,
                           c.cycles);
    if (!ctx) return 1;
    pm_server_opts so;
    memset(&so, 0, sizeof so);
    pm_server *srv = nullptr;
    if (pm_server_create(&so, &srv)) { pm_ctx_destroy(ctx); return fail("server creation failed", ""); }

    std::set<std::string> seen;
    std::vector<Inflight *> pending;
    int errors = 0;
    bool tuned = !o.autotune;

    auto submit = [&](const Job &j) {
        seen.insert(j.in);
        Inflight *f = new Inflight();
        f->job = j;
        float *cube;
        if (pmio_fits_read_cube(j.in.c_str(), L, c.subx1, c.subx2, c.suby1, c.suby2, &cube, &f->lay)) {
            fail("cannot read the Stokes cube", j.in);
            errors++;
            delete f;
            return;
        }
        f->rows = f->lay.rows;
        f->cols = f->lay.cols;
        const size_t n = (size_t)f->rows * f->cols;
        if (!mask.empty() && (mr != f->rows || mc != f->cols)) {
            fail("mask and cube differ in size", j.in);
        }
        for (size_t p = 0; p < n; p++)
            if (mask.empty() || mr != f->rows || mc != f->cols || mask[p] != 0) f->keep.push_back((int)p);
        const size_t k = f->keep.size();
        f->stokes.resize(k * 4 * L);
        for (size_t q = 0; q < k; q++)
            memcpy(&f->stokes[q * 4 * L], cube + (size_t)f->keep[q] * 4 * L, sizeof(float) * 4 * L);
        pmio_free(cube);
        std::vector<float> slp;
        if (sl.per_pixel) {
            if ((size_t)sl.npix != n) { fail("stray-light cube and Stokes cube differ in size", j.in); errors++; delete f; return; }
            slp.resize(k * 4 * L);
            for (size_t q = 0; q < k; q++)
                memcpy(&slp[q * 4 * L], &sl.v[(size_t)f->keep[q] * 4 * L], sizeof(float) * 4 * L);
        }
        f->m.assign(k * 11, 0.0);
        f->chi2.assign(k, 0.0f);
        f->it.assign(k, 0);
        f->acc.assign(k, 0);
        if (c.save_synthesis) f->fitbuf.assign(k * 4 * L, 0.0f);
        if (!tuned && k > 0) {                                    // the first cube picks the shape
            int shape[3];
            if (pm_ctx_autotune_host(ctx, f->stokes.data(), sl.per_pixel ? slp.data() : nullptr,
                                     (int)std::min<size_t>(k, 65536), shape) == 0)
                printf("kernel shape: %d lanes x %d samples, %d pixels per workgroup\n",
                       shape[0], shape[1], shape[2]);
            tuned = true;
        }
        int rc = pm_submit(srv, ctx, f->stokes.data(), sl.per_pixel ? slp.data() : nullptr, (int)k,
                           f->m.data(), f->chi2.data(), f->it.data(), f->acc.data(),
                           c.save_synthesis ? f->fitbuf.data() : nullptr, &f->ticket);
        if (rc) { fail("submission failed", pm_strerror(rc)); errors++; delete f; return; }
        if (sl.per_pixel) {
            // the server reads the stray light while the scan runs: wait here
            // rather than keep a copy alive (per-pixel stray light is rare)
            double lat;
            rc = pm_wait(srv, f->ticket, &lat);
            f->ticket = -1;
            f->ok = rc == 0;
        }
        pending.push_back(f);
    };

    auto finish = [&](Inflight *f) {
        if (f->ticket >= 0) {
            double lat;
            f->ok = pm_wait(srv, f->ticket, &lat) == 0;
        }
        if (!f->ok) { fail("inversion failed", f->job.in); errors++; }
        const size_t n = (size_t)f->rows * f->cols;
        std::vector<double> m(n * 11, 0.0);
        std::vector<float> chi2(n, 0.0f), fit;
        std::vector<int> it(n, 0);
        if (c.save_synthesis) fit.assign(n * 4 * L, 0.0f);
        for (size_t q = 0; q < f->keep.size(); q++) {
            const size_t p = f->keep[q];
            memcpy(&m[p * 11], &f->m[q * 11], sizeof(double) * 11);
            chi2[p] = f->chi2[q];
            it[p] = f->it[q];
            if (c.save_synthesis) memcpy(&fit[p * 4 * L], &f->fitbuf[q * 4 * L], sizeof(float) * 4 * L);
        }
        if (pmio_fits_write_models(f->job.out_mod.c_str(), f->rows, f->cols, m.data(), chi2.data(), it.data()))
            { fail("cannot write", f->job.out_mod); errors++; }
        else printf("inverted %s -> %s\n", f->job.in.c_str(), f->job.out_mod.c_str());
        if (c.save_synthesis && pmio_fits_write_cube(f->job.out_stokes.c_str(), &f->lay, fit.data()))
            { fail("cannot write", f->job.out_stokes); errors++; }
        pmio_free_layout(&f->lay);
        delete f;
    };

    for (const Job &j : jobs) {
        submit(j);
        // keep at most two cubes on the host ahead of the writer
        while (pending.size() > 2) { finish(pending.front()); pending.erase(pending.begin()); }
    }
    if (c.loop_inversion) {                                       // pmilos.c:2047-2103
        auto t0 = std::chrono::steady_clock::now();
        const std::string dir = dir_of(c.observed);
        while (true) {
            std::string fresh;
            std::vector<std::string> names;
            if (DIR *d = opendir(dir.c_str())) {
                while (struct dirent *e = readdir(d)) names.push_back(e->d_name);
                closedir(d);
            }
            std::sort(names.begin(), names.end());
            for (const std::string &nm : names) {
                if (nm == "." || nm == ".." || nm.find("_mod") != std::string::npos ||
                    nm.find("_stokes") != std::string::npos || ext_of(nm) != ".fits")
                    continue;
                std::string path = dir + "/" + nm;
                struct stat st;
                if (stat(path.c_str(), &st) || S_ISDIR(st.st_mode)) continue;
                bool known = false;
                for (const std::string &x : seen)        // every input comes from this directory
                    if (base_of(x) == nm) known = true;
                if (!known) { fresh = path; break; }
            }
            if (!fresh.empty()) {
                const std::string out = std::string(c.prefix) + base_of(fresh);
                submit({fresh, out + "_mod.fits", out + "_stokes.fits"});
                while (!pending.empty()) { finish(pending.front()); pending.erase(pending.begin()); }
                t0 = std::chrono::steady_clock::now();
                continue;
            }
            while (!pending.empty()) { finish(pending.front()); pending.erase(pending.begin()); }
            double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (waited >= o.watch_timeout) break;
            std::this_thread::sleep_for(std::chrono::duration<double>(o.watch_poll));
        }
    }
    while (!pending.empty()) { finish(pending.front()); pending.erase(pending.begin()); }
    pm_server_destroy(srv);
    pm_ctx_destroy(ctx);
    return errors ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Opts o;
    const char *file = nullptr;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--mixed") o.precision = PM_MIXED;
        else if (a == "--mixed-toplim" && i + 1 < argc) o.mixed_toplim = atof(argv[++i]);
        else if (a == "--autotune") o.autotune = true;
        else if (a == "--watch-timeout" && i + 1 < argc) o.watch_timeout = atof(argv[++i]);
        else if (a == "--watch-poll" && i + 1 < argc) o.watch_poll = atof(argv[++i]);
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else if (!file) file = argv[i];
        else { usage(); return 2; }
    }
    if (!file) { usage(); return 2; }
    Setup s;
    const bool minit = ext_of(file) == ".minit";
    int rc = minit ? pmio_read_minit(file, &s.c) : pmio_read_mtrol(file, &s.c);
    if (rc == -1) return fail("cannot read", file);
    if (rc == -2) fprintf(stderr, "pmilos_gpu: %s is short; the missing fields keep their defaults\n", file);
    if (rc == -3) return fail("the wavelength or PSF file is of an unknown kind", file);
    if (rc == -4) return fail("a subfield needs all of subx1, subx2, suby1, suby2", file);
    if (setup(s)) return 1;
    return minit ? run_pmilos(s, o) : run_milos(s, o);
}
