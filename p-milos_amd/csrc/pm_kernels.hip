// pm_kernels.hip -- host side of libpmilos_hip for MI355X (gfx950):
// contexts, kernel dispatch and autotuning, the batch graph and the C ABI
// declared in include/pmilos.h.  The kernels are in pm_invert.h, the
// context constants' set-up in pm_setup.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <mutex>
#include <vector>
#include <stdlib.h>
#include <string.h>

#include "pm_device.h"
#include "pm_internal.h"
#include "../../include/pmilos.h"

#include "pm_invert.h"
#include "pm_setup.h"

// ============================================================ host side

// The host-pointer inversion runs its pixels in up to PM_HOST_CHUNKS chunks,
// one stream each (upload, kernel and copy-back of a chunk in its stream's
// order, so that copies overlap other chunks' kernels): from
// PM_HOST_SPLIT_MIN pixels on, 1/16, 14/16 and 1/16 of them.  The small
// first chunk's kernel starts after a short upload and runs while the big
// chunk uploads; the small last one runs while the big one's results come
// back.  Equal chunks do worse: every chunk boundary leaves part of the GPU
// idle while a persistent kernel drains (1024^2 pixels x 30 samples, FP64,
// pinned buffers: one chunk 161.6 ms, eight equal ones 162.3 ms, 1/16 +
// 15/16 153.1 ms, 1/16 + 14/16 + 1/16 154.7 ms).  Below 2^20 pixels the
// split costs more than it hides (262,144 pixels at 6 samples: 12% slower).
constexpr int PM_HOST_STREAMS = 3, PM_HOST_CHUNKS = 3, PM_HOST_SPLIT_MIN = 1 << 20;

struct pm_ctx {
    PmConst h;
    PmConst *d = nullptr;
    int *counters = nullptr;
    unsigned long long *fallbacks = nullptr;
    int precision = PM_FP64;
    int device = 0;
    int ncu = 0;
    hipStream_t last_stream = nullptr;
    int variant[2] = {0, 0};   // kernel variant of the 9-row and general forms
    // a shape chosen by pm_ctx_set_shape among all compiled ones, in place of
    // variant[] (valid when forced[] is set)
    bool forced[2] = {false, false};
    int forced_shape[2][3] = {{0, 0, 0}, {0, 0, 0}};
    // recorded after every launch that uses the context's work-queue
    // counters, so that a launch on another stream waits for it (the
    // counters are reset at the start of each launch)
    hipEvent_t done = nullptr;
    bool launched = false;
    void *scratch = nullptr;   // this device's scratch (pm_invert_device_noalloc)
    // the host-pointer calls' device buffers and streams, made on first use
    // and kept (grown, never shrunk) until pm_ctx_destroy
    struct {
        hipStream_t s[PM_HOST_STREAMS] = {};      // chunk k on s[k]
        char *ws = nullptr;
        size_t ws_bytes = 0;
        int *counters = nullptr;            // a set per chunk
        unsigned long long *fallbacks = nullptr;   // one per chunk
        hipEvent_t cleared = nullptr;       // every chunk's counters zeroed
    } host;
};

// Orders a launch that uses the context's counters after the previous one
// when the stream changes (outside graph capture, where the graph's own
// counters are used).
static int ctx_counters_acquire(pm_ctx *ctx, hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) return PM_EDEVICE;
    if (cs != hipStreamCaptureStatusNone) return PM_EINVAL;
    if (ctx->launched && st != ctx->last_stream &&
        hipStreamWaitEvent(st, ctx->done, 0) != hipSuccess)
        return PM_EDEVICE;
    return PM_OK;
}

static int ctx_counters_release(pm_ctx *ctx, hipStream_t st)
{
    ctx->last_stream = st;
    if (hipEventRecord(ctx->done, st) != hipSuccess) return PM_EDEVICE;
    ctx->launched = true;
    return PM_OK;
}

// ------------------------------------------------------------ no-allocation scratch

// Device scratch of pm_invert_device_noalloc: the work-queue counters
// (PM_NSLICE counters PM_CSTRIDE ints apart) followed by the fallback
// counter.  Each GPU holds its own copy for the life of the process.  In
// host code the name stands for no device address: the address of this
// GPU's copy comes from hipGetSymbolAddress (scratch_of).
#define PM_SCRATCH_WORDS 136             // 8-byte words: 1,088 bytes per GPU
static_assert(PM_NSLICE * PM_CSTRIDE * sizeof(int) + sizeof(unsigned long long) <=
                  PM_SCRATCH_WORDS * sizeof(unsigned long long),
              "scratch too small for the counters");
static __device__ __align__(256) unsigned long long scratch[PM_SCRATCH_WORDS];

#define PM_MAX_DEVICES 64
static std::mutex scratch_mu;
static void *scratch_addr[PM_MAX_DEVICES];

// The device address of the current device's scratch, looked up once per
// device (with that device current) and kept.
static int scratch_of(int dev, void **out)
{
    if (dev < 0 || dev >= PM_MAX_DEVICES) return PM_ENOTSUP;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PM_EDEVICE;
        scratch_addr[dev] = p;
    }
    *out = scratch_addr[dev];
    return PM_OK;
}

extern "C" const char *pm_strerror(int status)
{
    switch (status) {
    case PM_OK: return "ok";
    case PM_EINVAL: return "invalid argument";
    case PM_ENOMEM: return "device allocation failed";
    case PM_EDEVICE: return "HIP runtime error";
    case PM_ENOTSUP: return "configuration not supported";
    case PM_ERANGE: return "size above the entry point's maximum";
    default: return "unknown error";
    }
}

extern "C" int pm_grid(double lambda0, double initial_mA, double step_mA, double final_mA,
                       double *lambda, int cap)
{
    if (!(step_mA != 0.0)) return PM_EINVAL;
    int n = (int)(((final_mA - initial_mA) / step_mA) + 1);       // milos.c:173
    if (n <= 0) return PM_EINVAL;
    double step = step_mA / 1000.0;
    double v = lambda0 + (initial_mA / 1000.0);                   // milos.c:194-197
    for (int i = 0; i < n; i++) {
        if (i > 0) v = v + step;
        if (lambda && i < cap) lambda[i] = v;
    }
    return n;
}

extern "C" int pm_psf_gaussian(double fwhm_mA, const double *lambda, int n, double *G)
{
    if (!lambda || !G || n < 2 || !(fwhm_mA > 0)) return PM_EINVAL;
    // fgauss_WL(FWHM, lambda[1]-lambda[0], lambda[0], lambda[n/2], n) (milos.c:269)
    double sigma = fwhm_mA * 0.42466090 / 1000.0;                 // fgauss.c:92
    double step = lambda[1] - lambda[0];
    double centre = lambda[n / 2];
    double sum = 0;
    for (int i = 0; i < n; i++) {
        double a = ((lambda[0] + i * step) - centre) / sigma;
        double term = (a * a) / 2;
        G[i] = term < 1e30 ? exp(-term) : 0.0;
        sum += G[i];
    }
    for (int i = 0; i < n; i++) G[i] /= sum;
    return PM_OK;
}

extern "C" int pm_nterms(const int fix[PM_NPAR], double guess_mac)
{
    return nterms_of(fix, guess_mac);
}

#define PM_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "pmilos: %s failed: %s\n", #call, hipGetErrorString(e_)); \
            return PM_EDEVICE;                                                \
        }                                                                     \
    } while (0)

extern "C" int pm_ctx_create(const pm_config *cfg, pm_ctx **out)
{
    if (!cfg || !out || !cfg->lambda) return PM_EINVAL;
    *out = nullptr;
    const int L = cfg->nlambda;
    if (L < 2 || L > PM_MAX_LAMBDA) return PM_ENOTSUP;
    if (cfg->miter < 1) return PM_EINVAL;
    if (cfg->precision != PM_FP64 && cfg->precision != PM_MIXED) return PM_EINVAL;
    pm_ctx *ctx = new pm_ctx();
    int rc = pm_const_build(cfg, ctx->h);
    if (rc != PM_OK) { delete ctx; return rc; }
    ctx->precision = cfg->precision;
    if (hipGetDevice(&ctx->device) != hipSuccess) { delete ctx; return PM_EDEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) { delete ctx; return PM_EDEVICE; }
    ctx->ncu = prop.multiProcessorCount;
    if (hipMalloc(&ctx->d, sizeof(PmConst)) != hipSuccess ||
        hipMalloc(&ctx->counters, sizeof(int) * PM_NSLICE * PM_CSTRIDE) != hipSuccess ||
        hipMalloc(&ctx->fallbacks, sizeof(unsigned long long)) != hipSuccess) {
        pm_ctx_destroy(ctx);
        return PM_ENOMEM;
    }
    if (hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming) != hipSuccess) {
        pm_ctx_destroy(ctx);
        return PM_EDEVICE;
    }
    if ((rc = scratch_of(ctx->device, &ctx->scratch)) != PM_OK) {
        pm_ctx_destroy(ctx);
        return rc;
    }
    if (hipMemcpy(ctx->d, &ctx->h, sizeof(PmConst), hipMemcpyHostToDevice) != hipSuccess) {
        pm_ctx_destroy(ctx);
        return PM_EDEVICE;
    }
    *out = ctx;
    return PM_OK;
}

extern "C" void pm_ctx_destroy(pm_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->d) (void)hipFree(ctx->d);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->fallbacks) (void)hipFree(ctx->fallbacks);
    if (ctx->done) (void)hipEventDestroy(ctx->done);
    for (auto st : ctx->host.s)
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    if (ctx->host.ws) (void)hipFree(ctx->host.ws);
    if (ctx->host.counters) (void)hipFree(ctx->host.counters);
    if (ctx->host.fallbacks) (void)hipFree(ctx->host.fallbacks);
    if (ctx->host.cleared) (void)hipEventDestroy(ctx->host.cleared);
    delete ctx;
}

extern "C" int pm_ctx_nterms(const pm_ctx *ctx) { return ctx ? ctx->h.NT : PM_EINVAL; }

// ------------------------------------------------------------ dispatch

// share: the fraction of the GPU's resident workgroups the launch may take
// (1 = all of them).
template <typename T, int W, int NLPL, int NT, bool GEN, int BLK>
static int launch_invert(pm_ctx *ctx, const InvArgs &a, hipStream_t st, double share, bool reset)
{
    using LD = Lds<T, W, NLPL, NT, GEN, BLK>;
    auto kern = pm_invert_kernel<T, W, NLPL, NT, GEN, BLK>;
    int per_cu = 0;
    PM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, LD::BLOCK, 0));
    if (per_cu < 1) per_cu = 1;
    long long need = ((long long)a.npix + LD::GPB - 1) / LD::GPB;
    long long grid = (long long)(per_cu * ctx->ncu * (share > 0 && share < 1 ? share : 1.0) + 0.5);
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (reset) {
        hipLaunchKernelGGL(pm_reset_counters, dim3(1), dim3(64), 0, st, a.counters, a.fallbacks);
        PM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(LD::BLOCK), 0, st, a);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

template <typename T, int W, int NLPL, int NT, bool GEN>
static int launch_synth(pm_ctx *ctx, const double *m, int n, double *sp, double *rf,
                        hipStream_t st)
{
    using LD = Lds<T, W, NLPL, NT, GEN>;
    long long need = ((long long)n + LD::GPB - 1) / LD::GPB;
    long long grid = (long long)ctx->ncu * 8;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((pm_synth_kernel<T, W, NLPL, NT, GEN>), dim3((unsigned)grid), dim3(LD::BLOCK),
                       0, st, (const PmConst *)ctx->d, m, n, sp, rf);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

// Kernel variants the autotuner chooses from: lanes per pixel W, samples
// per lane NLPL (W * NLPL >= L) and threads per workgroup (pixels per
// workgroup = threads / W).  The first of each list is the default mapping:
// in the FP64 form the smallest power of two >= L lanes up to a full
// wavefront, then a pixel over two waves (128 lanes, one sample each); in
// the FP32 form two samples per lane where such a shape is listed.  T =
// double (FP64 form) or float (mixed form: spectra, RFs and convolution in
// FP32); the general form (NTERMS 10/11, macroturbulence, stray light) is
// the 10- or 11-row kernel.  The synthesis kernel
// follows the selected variant's lane mapping.
typedef int (*InvLaunch)(pm_ctx *, const InvArgs &, hipStream_t, double, bool);
typedef int (*SynLaunch)(pm_ctx *, const double *, int, double *, double *, hipStream_t);
struct InvVariant {
    int W, NLPL, BLK;
    InvLaunch fn;
    SynLaunch sfn;         // synthesis with the same lane mapping
};
#define PM_V(T, W, NLPL, NT, G, BLK)                                          \
    InvVariant{W, NLPL, BLK, &launch_invert<T, W, NLPL, NT, G, BLK>,          \
               &launch_synth<T, W, NLPL, NT, G>}

template <typename T, int NT, bool G>
static int variants_of(int L, InvVariant *v)
{
    // Several samples per lane in the FP32 form; in the FP64 form, whose
    // two-sample kernels need far more than 256 registers per lane
    // (occupancy 1, never the fastest), only 8x2 after the one-sample shapes
    // and, beyond 64 samples, 64x2 after the two-wave 128x1 shape.  The FP64 8x2 and 16x2 kernels once
    // gave wrong inversions on the GPU; for 8x2 the in-place correlation's
    // missing wait (a block longer than the half-width) explains it and is
    // fixed, for 16x2 no cause was found and it stays out.
    constexpr bool MULTI = !G && sizeof(T) == 4;
    // In the FP32 form the two-sample shapes come first: on a million pixels
    // they invert 14 and 30 samples 1.2x faster than one sample per lane
    // (the solve and LM control are shared by twice the samples per lane).
    int n = 0;
    if (L <= 8) {
        v[n++] = PM_V(T, 8, 1, NT, G, 128);
    } else if (L <= 16) {
        if constexpr (MULTI) v[n++] = PM_V(T, 8, 2, NT, G, 128);
        v[n++] = PM_V(T, 16, 1, NT, G, 256);
        v[n++] = PM_V(T, 16, 1, NT, G, 128);
        if constexpr (!G && sizeof(T) == 8) v[n++] = PM_V(T, 8, 2, NT, G, 128);
    } else if (L <= 32) {
        if constexpr (MULTI) {
            v[n++] = PM_V(T, 16, 2, NT, G, 256);
            v[n++] = PM_V(T, 16, 2, NT, G, 128);
        }
        v[n++] = PM_V(T, 32, 1, NT, G, 256);
        v[n++] = PM_V(T, 32, 1, NT, G, 128);
    } else if (L <= 64) {
        if constexpr (MULTI) v[n++] = PM_V(T, 32, 2, NT, G, 256);
        v[n++] = PM_V(T, 64, 1, NT, G, 256);
        v[n++] = PM_V(T, 64, 1, NT, G, 128);
    } else {
        // FP64: a pixel over two waves, one sample per lane, first: two waves
        // per SIMD where the two-sample kernels fit only one (3% faster on
        // the 112-sample cube with v_mac free; the workgroup barriers of
        // its convolutions and sums take most of the gain)
        if constexpr (sizeof(T) == 8) v[n++] = PM_V(T, 128, 1, NT, G, 128);
        v[n++] = PM_V(T, 64, 2, NT, G, 256);
        v[n++] = PM_V(T, 64, 2, NT, G, 128);
    }
    return n;
}

constexpr int PM_MAXVAR = 8;

// The general form runs the 10-row kernel when the reference's layout has
// 10 rows with v_mac in row 9 and no filling-factor row (v_mac free, no
// stray light, fix[10] == 0), else the 11-row one.  The two lists have the
// same shapes in the same order, so a variant index or a forced shape
// means the same kernel shape in either.
static bool ten_rows(const pm_ctx *ctx, bool slight_cube)
{
    return ctx->h.NT == 10 && ctx->h.fix[10] == 0 && !ctx->h.use_slight && !slight_cube;
}

static int variants_at(const pm_ctx *ctx, bool gen, bool cube, int L, InvVariant *v)
{
    const bool mx = ctx->precision == PM_MIXED;
    if (gen && ten_rows(ctx, cube))
        return mx ? variants_of<float, 10, true>(L, v) : variants_of<double, 10, true>(L, v);
    if (gen)
        return mx ? variants_of<float, 11, true>(L, v) : variants_of<double, 11, true>(L, v);
    return mx ? variants_of<float, 9, false>(L, v) : variants_of<double, 9, false>(L, v);
}

static int variants(const pm_ctx *ctx, bool gen, InvVariant *v, bool cube = false)
{
    return variants_at(ctx, gen, cube, ctx->h.L, v);
}

// Every compiled shape of a form that covers L samples (W * NLPL >= L):
// the variants of each sampling bucket from L up.
static int all_shapes(const pm_ctx *ctx, bool gen, InvVariant *v, int cap, bool cube = false)
{
    int n = 0;
    const int buckets[5] = {8, 16, 32, 64, PM_MAX_LAMBDA};
    for (int b = 0; b < 5; b++) {
        if (buckets[b] < ctx->h.L) continue;
        InvVariant w[PM_MAXVAR];
        const int nw = variants_at(ctx, gen, cube, buckets[b], w);
        for (int i = 0; i < nw && n < cap; i++) v[n++] = w[i];
    }
    return n;
}

// The shape in use for a form: the one pm_ctx_set_shape chose, else the
// autotuner's / pm_ctx_set_variant's index into variants(), else the first.
// cube: the inversion has a per-pixel stray-light cube.
static InvVariant selected(const pm_ctx *ctx, bool gen, bool cube = false)
{
    if (ctx->forced[gen]) {
        InvVariant v[5 * PM_MAXVAR];
        const int n = all_shapes(ctx, gen, v, 5 * PM_MAXVAR, cube);
        for (int i = 0; i < n; i++)
            if (v[i].W == ctx->forced_shape[gen][0] && v[i].NLPL == ctx->forced_shape[gen][1] &&
                v[i].BLK == ctx->forced_shape[gen][2])
                return v[i];
    }
    InvVariant v[PM_MAXVAR];
    const int nv = variants(ctx, gen, v, cube);
    return v[ctx->variant[gen] < nv ? ctx->variant[gen] : 0];
}

// The 9-row kernel covers NTERMS 9 without stray light, with the PSF by
// zero-padded correlation; everything else takes the general one.
static bool needs_general(const pm_ctx *ctx, bool slight_cube)
{
    return ctx->h.NT != 9 || ctx->h.use_slight || slight_cube ||
           (ctx->h.use_fft && ctx->h.use_psf);
}

extern "C" int pm_synth_device(pm_ctx *ctx, const double *d_models, int n, double *d_spectra,
                               double *d_rf, void *stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_models || !d_spectra))) return PM_EINVAL;
    if (n == 0) return PM_OK;
    hipStream_t st = (hipStream_t)stream;
    const bool gen = needs_general(ctx, false);
    return selected(ctx, gen).sfn(ctx, d_models, n, d_spectra, d_rf, st);
}

int pm_invert_launch(pm_ctx *ctx, const float *d_stokes, const float *d_slight, int npix,
                     double *d_models, float *d_chi2, int *d_iters, int *d_accepted,
                     float *d_fit, int *d_counters, unsigned long long *d_fallbacks,
                     hipStream_t st, double share, bool reset)
{
    if (!ctx || npix < 0) return PM_EINVAL;
    if (npix == 0) return PM_OK;
    if (!d_stokes || !d_models || !d_chi2 || !d_iters) return PM_EINVAL;
    InvArgs a;
    a.c = ctx->d;
    a.stokes = d_stokes;
    a.npix = npix;
    a.slice = (npix + PM_NSLICE - 1) / PM_NSLICE;
    a.models = d_models;
    a.chi2 = d_chi2;
    a.iters = d_iters;
    a.accepted = d_accepted;
    a.fit = d_fit;
    a.counters = d_counters;
    a.fallbacks = d_fallbacks;
    a.slight = d_slight;
    const bool gen = needs_general(ctx, d_slight != nullptr);
    return selected(ctx, gen, d_slight != nullptr).fn(ctx, a, st, share, reset);
}

int pm_ctx_nlambda(const pm_ctx *ctx) { return ctx->h.L; }

double pm_ctx_cost(const pm_ctx *ctx)
{
    // per-iteration work of one pixel, relative: the sample-wise terms, the
    // convolutions over the PSF taps (or about half of the samples when
    // macroturbulence is on) for every response row, the NT^3 solve
    const PmConst &c = ctx->h;
    const bool mac = c.guess[9] > 0.0001 || c.fix[9];
    const double taps = mac ? 0.4 * c.L : (c.use_psf ? c.ghi64 - c.glo64 : 0);
    return c.L * (400.0 + 10.0 * c.NT * taps) + 10.0 * c.NT * c.NT * c.NT;
}
int pm_ctx_device(const pm_ctx *ctx) { return ctx->device; }
size_t pm_counters_bytes() { return sizeof(int) * PM_NSLICE * PM_CSTRIDE; }

int pm_reset_launch(int *d_counters, unsigned long long *d_fallbacks, hipStream_t st)
{
    hipLaunchKernelGGL(pm_reset_counters, dim3(1), dim3(64), 0, st, d_counters, d_fallbacks);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

extern "C" int pm_invert_sl_device(pm_ctx *ctx, const float *d_stokes, const float *d_slight,
                                   int npix, double *d_models, float *d_chi2, int *d_iters,
                                   int *d_accepted, float *d_fit, void *stream)
{
    if (!ctx) return PM_EINVAL;
    if (npix == 0) return PM_OK;
    hipStream_t st = (hipStream_t)stream;
    int rc = ctx_counters_acquire(ctx, st);
    if (rc != PM_OK) return rc;
    rc = pm_invert_launch(ctx, d_stokes, d_slight, npix, d_models, d_chi2, d_iters, d_accepted,
                          d_fit, ctx->counters, ctx->fallbacks, st, 1.0);
    if (rc != PM_OK) return rc;
    return ctx_counters_release(ctx, st);
}

extern "C" int pm_invert_device(pm_ctx *ctx, const float *d_stokes, int npix, double *d_models,
                                float *d_chi2, int *d_iters, int *d_accepted, float *d_fit,
                                void *stream)
{
    return pm_invert_sl_device(ctx, d_stokes, nullptr, npix, d_models, d_chi2, d_iters,
                               d_accepted, d_fit, stream);
}

extern "C" int pm_invert_device_noalloc(pm_ctx *ctx, const float *d_stokes, int npix,
                                        double *d_models, float *d_chi2, int *d_iters,
                                        int *d_accepted, float *d_fit, void *stream)
{
    if (!ctx || npix < 0) return PM_EINVAL;
    if (npix > PM_NOALLOC_MAX_NPIX) return PM_ERANGE;
    if (npix == 0) return PM_OK;
    if (!ctx->scratch) return PM_EDEVICE;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PM_EDEVICE;
    if (dev != ctx->device) return PM_EINVAL;
    // counters first, the fallback counter after them; both cleared by the
    // launch's reset kernel (no memset, so the launch can be captured)
    int *cnt = (int *)ctx->scratch;
    unsigned long long *fb = (unsigned long long *)ctx->scratch + PM_NSLICE * PM_CSTRIDE / 2;
    return pm_invert_launch(ctx, d_stokes, nullptr, npix, d_models, d_chi2, d_iters, d_accepted,
                            d_fit, cnt, fb, (hipStream_t)stream, 1.0);
}

extern "C" int pm_classical_device(pm_ctx *ctx, const float *d_stokes, int npix,
                                   double *d_models, void *stream)
{
    if (!ctx || npix < 0) return PM_EINVAL;
    if (npix == 0) return PM_OK;
    if (!d_stokes || !d_models) return PM_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(pm_classical_kernel, dim3((unsigned)((npix + 127) / 128)), dim3(128), 0, st,
                       (const PmConst *)ctx->d, d_stokes, npix, d_models);
    PM_HIP(hipGetLastError());
    return PM_OK;
}

extern "C" int pm_last_fallback_count(pm_ctx *ctx, long long *count)
{
    if (!ctx || !count) return PM_EINVAL;
    unsigned long long v = 0;
    if (ctx->launched) PM_HIP(hipEventSynchronize(ctx->done));
    PM_HIP(hipMemcpy(&v, ctx->fallbacks, sizeof(v), hipMemcpyDeviceToHost));
    *count = (long long)v;
    return PM_OK;
}

// ------------------------------------------------------------ autotuner

extern "C" int pm_ctx_autotune(pm_ctx *ctx, const float *d_stokes, const float *d_slight, int npix,
                               void *stream, int chosen[3])
{
    if (!ctx || !d_stokes || npix < 1) return PM_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    {
        const int rc0 = ctx_counters_acquire(ctx, st);
        if (rc0 != PM_OK) return rc0;
    }
    const bool gen = needs_general(ctx, d_slight != nullptr);
    InvVariant v[PM_MAXVAR];
    const int nv = variants(ctx, gen, v, d_slight != nullptr);
    const size_t n = npix;
    void *buf = nullptr;
    if (hipMalloc(&buf, n * (11 * sizeof(double) + 3 * sizeof(int))) != hipSuccess)
        return PM_ENOMEM;
    double *dm = (double *)buf;
    float *dc = (float *)(dm + 11 * n);
    int *di = (int *)(dc + n), *da = di + n;
    InvArgs a;
    a.c = ctx->d;
    a.stokes = d_stokes;
    a.npix = npix;
    a.slice = (npix + PM_NSLICE - 1) / PM_NSLICE;
    a.models = dm;
    a.chi2 = dc;
    a.iters = di;
    a.accepted = da;
    a.fit = nullptr;
    a.counters = ctx->counters;
    a.fallbacks = ctx->fallbacks;
    a.slight = d_slight;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PM_EDEVICE;
    int best = 0;
    float best_ms = 0;
    for (int k = 0; k < nv && rc == PM_OK; k++) {
        float ms = 0;
        rc = v[k].fn(ctx, a, st, 1.0, true);                 // first run: not timed
        if (rc == PM_OK && hipEventRecord(e0, st) != hipSuccess) rc = PM_EDEVICE;
        for (int rep = 0; rep < 2 && rc == PM_OK; rep++) rc = v[k].fn(ctx, a, st, 1.0, true);
        if (rc == PM_OK && (hipEventRecord(e1, st) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = PM_EDEVICE;
        if (rc == PM_OK && (k == 0 || ms < best_ms)) { best = k; best_ms = ms; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (ctx_counters_release(ctx, st) != PM_OK && rc == PM_OK) rc = PM_EDEVICE;
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (rc != PM_OK) return rc;
    ctx->variant[gen] = best;
    ctx->forced[gen] = false;
    if (chosen) { chosen[0] = v[best].W; chosen[1] = v[best].NLPL; chosen[2] = v[best].BLK / v[best].W; }
    return PM_OK;
}

extern "C" int pm_ctx_variant(const pm_ctx *ctx, int general, int out[3])
{
    if (!ctx || !out) return PM_EINVAL;
    InvVariant v[PM_MAXVAR];
    const int nv = variants(ctx, general != 0, v);
    const InvVariant u = selected(ctx, general != 0);
    out[0] = u.W;
    out[1] = u.NLPL;
    out[2] = u.BLK / u.W;
    return nv;
}

extern "C" int pm_ctx_set_mixed_toplim(pm_ctx *ctx, double toplim_mixed)
{
    if (!ctx || !(toplim_mixed >= 0.0) || !(toplim_mixed < 100.0)) return PM_EINVAL;
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    PM_HIP(hipSetDevice(ctx->device));
    // the context's last kernels read the constants: let them finish first
    // (the copy below is on the null stream, which does not wait for
    // non-blocking streams)
    hipError_t e = ctx->launched ? hipEventSynchronize(ctx->done) : hipSuccess;
    if (e == hipSuccess)
        e = hipMemcpy(&ctx->d->toplim_mixed, &toplim_mixed, sizeof(double), hipMemcpyHostToDevice);
    PM_HIP(hipSetDevice(dev));
    PM_HIP(e);
    ctx->h.toplim_mixed = toplim_mixed;       // the host copy follows the device's
    return PM_OK;
}

extern "C" int pm_ctx_set_variant(pm_ctx *ctx, int general, int index)
{
    if (!ctx) return PM_EINVAL;
    InvVariant v[PM_MAXVAR];
    const int nv = variants(ctx, general != 0, v);
    if (index < 0 || index >= nv) return PM_EINVAL;
    ctx->variant[general != 0] = index;
    ctx->forced[general != 0] = false;
    return PM_OK;
}

extern "C" int pm_ctx_set_shape(pm_ctx *ctx, int general, int lanes, int samples,
                                int pixels_per_workgroup)
{
    if (!ctx || lanes < 1 || samples < 1 || pixels_per_workgroup < 1) return PM_EINVAL;
    const bool gen = general != 0;
    InvVariant v[5 * PM_MAXVAR];
    const int n = all_shapes(ctx, gen, v, 5 * PM_MAXVAR);
    for (int i = 0; i < n; i++)
        if (v[i].W == lanes && v[i].NLPL == samples && v[i].BLK == lanes * pixels_per_workgroup) {
            ctx->forced[gen] = true;
            ctx->forced_shape[gen][0] = lanes;
            ctx->forced_shape[gen][1] = samples;
            ctx->forced_shape[gen][2] = v[i].BLK;
            return PM_OK;
        }
    return PM_ENOTSUP;
}

// ------------------------------------------------------------ one launch for a batch

struct pm_graph {
    int n = 0;
    std::vector<int *> counters;
    std::vector<unsigned long long *> fallbacks;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" void pm_graph_destroy(pm_graph *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto p : g->counters) if (p) (void)hipFree(p);
    for (auto p : g->fallbacks) if (p) (void)hipFree(p);
    delete g;
}

extern "C" int pm_graph_create(const pm_batch_item *items, int n, double share_by_work,
                               pm_graph **out)
{
    if (!items || n < 1 || !out) return PM_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        const pm_batch_item &b = items[i];
        if (!b.ctx || b.npix < 1 || !b.d_stokes || !b.d_models || !b.d_chi2 || !b.d_iters)
            return PM_EINVAL;
        if (b.ctx->device != items[0].ctx->device) return PM_EINVAL;
    }
    pm_graph *g = new pm_graph();
    g->n = n;
    g->counters.assign(n, nullptr);
    g->fallbacks.assign(n, nullptr);
    for (int i = 0; i < n; i++)
        if (hipMalloc(&g->counters[i], sizeof(int) * PM_NSLICE * PM_CSTRIDE) != hipSuccess ||
            hipMalloc(&g->fallbacks[i], sizeof(unsigned long long)) != hipSuccess) {
            pm_graph_destroy(g);
            return PM_ENOMEM;
        }
    // each branch's share of the GPU: its estimated work (pixels x cost per
    // pixel) over the batch's, when asked; else every branch may fill it
    std::vector<double> share(n, 1.0);
    if (share_by_work > 0) {
        double tot = 0;
        for (int i = 0; i < n; i++) tot += items[i].npix * pm_ctx_cost(items[i].ctx);
        for (int i = 0; i < n; i++)
            share[i] = std::max(0.02, items[i].npix * pm_ctx_cost(items[i].ctx) / tot);
    }
    std::vector<hipStream_t> br(n, nullptr);
    std::vector<hipEvent_t> ev(n + 1, nullptr);
    hipStream_t s0 = nullptr;
    int rc = PM_OK;
    if (hipStreamCreateWithFlags(&s0, hipStreamNonBlocking) != hipSuccess) rc = PM_EDEVICE;
    for (int i = 0; i < n && rc == PM_OK; i++)
        if (hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking) != hipSuccess) rc = PM_EDEVICE;
    for (int i = 0; i <= n && rc == PM_OK; i++)
        if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) rc = PM_EDEVICE;
    bool capturing = false;
    if (rc == PM_OK) {
        if (hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) != hipSuccess) rc = PM_EDEVICE;
        else capturing = true;
    }
    if (rc == PM_OK && hipEventRecord(ev[n], s0) != hipSuccess) rc = PM_EDEVICE;
    for (int i = 0; i < n && rc == PM_OK; i++) {
        const pm_batch_item &b = items[i];
        if (hipStreamWaitEvent(br[i], ev[n], 0) != hipSuccess) { rc = PM_EDEVICE; break; }
        rc = pm_invert_launch(b.ctx, b.d_stokes, b.d_slight, b.npix, b.d_models, b.d_chi2,
                              b.d_iters, b.d_accepted, b.d_fit, g->counters[i], g->fallbacks[i],
                              br[i], share[i]);
        if (rc == PM_OK && (hipEventRecord(ev[i], br[i]) != hipSuccess ||
                            hipStreamWaitEvent(s0, ev[i], 0) != hipSuccess))
            rc = PM_EDEVICE;
    }
    if (capturing) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s0, &graph);
        if (e != hipSuccess && rc == PM_OK) rc = PM_EDEVICE;
        g->graph = graph;
    }
    if (rc == PM_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = PM_EDEVICE;
    for (auto e : ev) if (e) (void)hipEventDestroy(e);
    for (auto s : br) if (s) (void)hipStreamDestroy(s);
    if (s0) (void)hipStreamDestroy(s0);
    if (rc != PM_OK) {
        pm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return PM_OK;
}

extern "C" int pm_graph_launch(pm_graph *g, void *stream)
{
    if (!g || !g->exec) return PM_EINVAL;
    PM_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return PM_OK;
}

// ------------------------------------------------------------ host-pointer wrappers

namespace {
// The host calls' streams, counters and per-chunk fallback counters, made
// on the first host call of a context.
int host_setup(pm_ctx *ctx)
{
    auto &h = ctx->host;
    if (h.counters) return PM_OK;
    for (auto &st : h.s)
        if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            st = nullptr;
            return PM_EDEVICE;
        }
    // each step is skipped when an earlier call that failed later already
    // made it, so a retry after an error leaks nothing
    if (!h.fallbacks &&
        hipMalloc(&h.fallbacks, sizeof(unsigned long long) * PM_HOST_CHUNKS) != hipSuccess) {
        h.fallbacks = nullptr;
        return PM_ENOMEM;
    }
    if (!h.cleared && hipEventCreateWithFlags(&h.cleared, hipEventDisableTiming) != hipSuccess) {
        h.cleared = nullptr;
        return PM_EDEVICE;
    }
    if (hipMalloc(&h.counters, pm_counters_bytes() * PM_HOST_CHUNKS) != hipSuccess) {
        h.counters = nullptr;
        return PM_ENOMEM;
    }
    return PM_OK;
}

// The host calls' device buffer, at least `bytes` (the old one is freed
// when it grows: the calls that use it are synchronous, so none is in
// flight).
int host_ws(pm_ctx *ctx, size_t bytes, char **out)
{
    auto &h = ctx->host;
    if (h.ws_bytes < bytes) {
        if (h.ws) (void)hipFree(h.ws);
        h.ws = nullptr;
        h.ws_bytes = 0;
        if (hipMalloc(&h.ws, bytes) != hipSuccess) return PM_ENOMEM;
        h.ws_bytes = bytes;
    }
    *out = h.ws;
    return PM_OK;
}

// Waits for every host stream of the context (on every exit of a host
// call: their buffers are reused by the next one).
struct HostSync {
    pm_ctx *ctx;
    ~HostSync()
    {
        for (auto st : ctx->host.s)
            if (st) (void)hipStreamSynchronize(st);
    }
};

size_t al(size_t n) { return (n + 255) / 256 * 256; }
}  // namespace

extern "C" int pm_synth(pm_ctx *ctx, const double *models, int n, double *spectra, double *rf)
{
    if (!ctx || n < 0 || (n > 0 && (!models || !spectra))) return PM_EINVAL;
    if (n == 0) return PM_OK;
    const size_t L = ctx->h.L, NT = ctx->h.NT, bm = sizeof(double) * 11 * n,
                 bs = sizeof(double) * 4 * L * n, br = rf ? sizeof(double) * 4 * NT * L * n : 0;
    int rc = host_setup(ctx);
    if (rc) return rc;
    HostSync sync{ctx};
    char *w = nullptr;
    if ((rc = host_ws(ctx, al(bm) + al(bs) + al(br), &w))) return rc;
    double *dm = (double *)w, *ds = (double *)(w + al(bm)), *dr = (double *)(w + al(bm) + al(bs));
    hipStream_t st = ctx->host.s[0];
    PM_HIP(hipMemcpyAsync(dm, models, bm, hipMemcpyHostToDevice, st));
    if ((rc = pm_synth_device(ctx, dm, n, ds, rf ? dr : nullptr, st))) return rc;
    PM_HIP(hipMemcpyAsync(spectra, ds, bs, hipMemcpyDeviceToHost, st));
    if (rf) PM_HIP(hipMemcpyAsync(rf, dr, br, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    return PM_OK;
}

// Host buffers: the pixels go in chunks (PM_HOST_CHUNKS), chunk k on the
// context's host stream k in its own slice of the device buffers: upload,
// kernel, copy-back.  Every chunk has its own work-queue counters and
// fallback counter, all zeroed before the first upload: a counter reset
// queued behind another chunk would wait for a CU that the other stream's
// persistent kernel holds, and hold back its own chunk's kernel.  Every
// chunk's kernel is queued before any copy-back, since a copy into pageable
// memory returns only when it is done, i.e. after its chunk's kernel.  A
// pixel's results do not depend on its chunk, so the bits are those of one
// pm_invert_sl_device over all pixels.
extern "C" int pm_invert_sl(pm_ctx *ctx, const float *stokes, const float *slight, int npix,
                            double *models, float *chi2, int *iters, int *accepted, float *fit)
{
    if (!ctx || npix < 0) return PM_EINVAL;
    if (npix == 0) return PM_OK;
    if (!stokes || !models || !chi2 || !iters) return PM_EINVAL;
    const size_t L = ctx->h.L, n = npix, bst = sizeof(float) * 4 * L * n;
    int bound[PM_HOST_CHUNKS + 1] = {0, npix, npix, npix};
    int nch = 1;
    if (npix >= PM_HOST_SPLIT_MIN) {
        const int small = npix / 16;
        bound[1] = small;
        bound[2] = npix - small;
        nch = 3;
    }
    const size_t off_sl = al(bst), off_m = off_sl + (slight ? al(bst) : 0),
                 off_c = off_m + al(sizeof(double) * 11 * n), off_i = off_c + al(sizeof(float) * n),
                 off_a = off_i + al(sizeof(int) * n), off_f = off_a + (accepted ? al(sizeof(int) * n) : 0),
                 total = off_f + (fit ? al(bst) : 0);
    int rc = host_setup(ctx);
    if (rc) return rc;
    HostSync sync{ctx};
    char *w = nullptr;
    if ((rc = host_ws(ctx, total, &w))) return rc;
    float *d_st = (float *)w, *d_sl = slight ? (float *)(w + off_sl) : nullptr;
    double *d_m = (double *)(w + off_m);
    float *d_c = (float *)(w + off_c), *d_f = fit ? (float *)(w + off_f) : nullptr;
    int *d_i = (int *)(w + off_i), *d_a = accepted ? (int *)(w + off_a) : nullptr;
    // the fallback total is written where pm_last_fallback_count reads it:
    // no device-pointer call may still be using it
    if (ctx->launched) PM_HIP(hipEventSynchronize(ctx->done));
    auto &h = ctx->host;
    const size_t cset = pm_counters_bytes() / sizeof(int);
    for (int k = 0; k < nch; k++)
        if ((rc = pm_reset_launch(h.counters + k * cset, h.fallbacks + k, h.s[0]))) return rc;
    PM_HIP(hipEventRecord(h.cleared, h.s[0]));
    for (int k = 1; k < nch; k++) PM_HIP(hipStreamWaitEvent(h.s[k], h.cleared, 0));
    for (int k = 0; k < nch; k++) {
        hipStream_t st = h.s[k];
        const size_t p0 = bound[k], m = bound[k + 1] - bound[k], q0 = p0 * 4 * L;
        const size_t sb = sizeof(float) * 4 * L * m;
        PM_HIP(hipMemcpyAsync(d_st + q0, stokes + q0, sb, hipMemcpyHostToDevice, st));
        if (slight) PM_HIP(hipMemcpyAsync(d_sl + q0, slight + q0, sb, hipMemcpyHostToDevice, st));
        rc = pm_invert_launch(ctx, d_st + q0, slight ? d_sl + q0 : nullptr, (int)m, d_m + p0 * 11,
                              d_c + p0, d_i + p0, d_a ? d_a + p0 : nullptr, d_f ? d_f + q0 : nullptr,
                              h.counters + k * cset, h.fallbacks + k, st, 1.0, false);
        if (rc) return rc;
    }
    for (int k = 0; k < nch; k++) {
        hipStream_t st = h.s[k];
        const size_t p0 = bound[k], m = bound[k + 1] - bound[k], q0 = p0 * 4 * L;
        PM_HIP(hipMemcpyAsync(models + p0 * 11, d_m + p0 * 11, sizeof(double) * 11 * m,
                              hipMemcpyDeviceToHost, st));
        PM_HIP(hipMemcpyAsync(chi2 + p0, d_c + p0, sizeof(float) * m, hipMemcpyDeviceToHost, st));
        PM_HIP(hipMemcpyAsync(iters + p0, d_i + p0, sizeof(int) * m, hipMemcpyDeviceToHost, st));
        if (accepted)
            PM_HIP(hipMemcpyAsync(accepted + p0, d_a + p0, sizeof(int) * m, hipMemcpyDeviceToHost,
                                  st));
        if (fit)
            PM_HIP(hipMemcpyAsync(fit + q0, d_f + q0, sizeof(float) * 4 * L * m,
                                  hipMemcpyDeviceToHost, st));
    }
    for (int k = 0; k < nch; k++) PM_HIP(hipStreamSynchronize(h.s[k]));
    unsigned long long fb[PM_HOST_CHUNKS] = {}, tot = 0;
    PM_HIP(hipMemcpy(fb, h.fallbacks, sizeof(unsigned long long) * nch, hipMemcpyDeviceToHost));
    for (int k = 0; k < nch; k++) tot += fb[k];
    PM_HIP(hipMemcpy(ctx->fallbacks, &tot, sizeof(tot), hipMemcpyHostToDevice));
    return PM_OK;
}

extern "C" int pm_invert(pm_ctx *ctx, const float *stokes, int npix, double *models, float *chi2,
                         int *iters, int *accepted, float *fit)
{
    return pm_invert_sl(ctx, stokes, nullptr, npix, models, chi2, iters, accepted, fit);
}

extern "C" int pm_classical(pm_ctx *ctx, const float *stokes, int npix, double *models)
{
    if (!ctx || npix < 0) return PM_EINVAL;
    if (npix == 0) return PM_OK;
    if (!stokes || !models) return PM_EINVAL;
    const size_t L = ctx->h.L, n = npix, bs = sizeof(float) * 4 * L * n,
                 bm = sizeof(double) * 11 * n;
    int rc = host_setup(ctx);
    if (rc) return rc;
    HostSync sync{ctx};
    char *w = nullptr;
    if ((rc = host_ws(ctx, al(bs) + al(bm), &w))) return rc;
    hipStream_t st = ctx->host.s[0];
    PM_HIP(hipMemcpyAsync(w, stokes, bs, hipMemcpyHostToDevice, st));
    if ((rc = pm_classical_device(ctx, (const float *)w, npix, (double *)(w + al(bs)), st)))
        return rc;
    PM_HIP(hipMemcpyAsync(models, w + al(bs), bm, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    return PM_OK;
}

extern "C" int pm_ctx_autotune_host(pm_ctx *ctx, const float *stokes, const float *slight,
                                    int npix, int chosen[3])
{
    if (!ctx || !stokes || npix < 1) return PM_EINVAL;
    const size_t bs = sizeof(float) * (size_t)ctx->h.L * 4 * npix;
    int rc = host_setup(ctx);
    if (rc) return rc;
    HostSync sync{ctx};
    char *w = nullptr;
    if ((rc = host_ws(ctx, al(bs) * (slight ? 2 : 1), &w))) return rc;
    hipStream_t st = ctx->host.s[0];
    PM_HIP(hipMemcpyAsync(w, stokes, bs, hipMemcpyHostToDevice, st));
    if (slight) PM_HIP(hipMemcpyAsync(w + al(bs), slight, bs, hipMemcpyHostToDevice, st));
    return pm_ctx_autotune(ctx, (const float *)w, slight ? (const float *)(w + al(bs)) : nullptr,
                           npix, st, chosen);
}
