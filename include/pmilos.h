/*
 * pmilos.h -- C ABI of the MI355X Milne-Eddington inversion library
 * (libpmilos_hip.so).
 *
 * Plain C: pointers and sizes only.  Every entry point returns PM_OK (0) or
 * a negative error code (pm_strerror gives the text); none of them exits the
 * process.  A context is used by one host thread at a time; its kernels run
 * asynchronously on the stream given to the *_device entry points (NULL =
 * the default stream), the host-pointer entry points are synchronous.
 * Inversions of one context share its work-queue counters: one issued on a
 * stream other than the previous one's waits for that one to finish (an
 * event dependency, no host wait), so a context's inversions never overlap
 * on the device; use one context per stream for concurrent inversions, or
 * pm_graph / pm_server, which keep counters per branch and per stream.
 * pm_invert_device / pm_invert_sl_device / pm_ctx_autotune refuse a stream
 * under graph capture (PM_EINVAL): capture a batch with pm_graph_create.
 *
 * Reference interfaces replaced (P-MILOS, src/):
 *   pm_ctx_create    the per-run setup milos.c:133-263 + create_cuantic
 *                    (create_cuantic.c:24-45) + fgauss_WL PSF (milos.c:269),
 *                    i.e. the globals NTERMS, cuantic, G, CC that lm_mils and
 *                    mil_sinrf read (milos.c:39-100).
 *   pm_invert        the FITS pixel loop milos.c:958-1009 (guess copy,
 *                    estimacionesClasicas, NaN guards, lm_mils per pixel) and
 *                    its MPI copies pmilos.c:1256-1336; lm_mils itself is
 *                    milosUtils.h:54-57 / milosUtils.c:466-635.
 *   pm_synth         synthesis mode milos.c:601-602: mil_sinrf
 *                    (mil_sinrf.c:46) + me_der (me_der.c:48).
 *   pm_grid          wavelength axis milos.c:173-197 (readMallaGrid output).
 *   pm_psf_gaussian  fgauss_WL, fgauss.c:79-154, as called at milos.c:269.
 *   pm_nterms        NTERMS, milos.c:154-157.
 *   pm_server_*, pm_submit, pm_poll, pm_wait
 *                    the MPI master/worker farm that hands cubes to ranks
 *                    (pmilos.c:1181-1336) and the time-series loop that
 *                    takes cubes as they appear (pmilos.c:2047-2103): scans
 *                    are submitted as they arrive and share one GPU.
 */
#ifndef PMILOS_H
#define PMILOS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_NPAR 11        /* Init_Model: eta0 B vlos dopp aa gm az S0 S1 mac alfa */
#define PM_MAX_LAMBDA 128 /* wavelength samples per pixel supported by the kernels */

enum pm_status {
    PM_OK = 0,
    PM_EINVAL = -22,      /* bad argument */
    PM_ENOMEM = -12,      /* device allocation failed */
    PM_EDEVICE = -5,      /* HIP runtime error */
    PM_ENOTSUP = -95,     /* configuration not supported by the kernels */
    PM_ERANGE = -34       /* size above the entry point's maximum */
};

enum pm_precision {
    PM_FP64 = 0,          /* all arithmetic in double (reference REAL=double) */
    PM_MIXED = 1          /* FP32 spectra/RF/convolution, FP64 solve and chi2 */
};

typedef struct pm_ctx pm_ctx;

/* Run configuration: the fields of ConfigControl (defines.h:223-277) and
 * the line data that lm_mils and mil_sinrf read. */
typedef struct {
    int nlambda;                 /* L, 2 <= L <= PM_MAX_LAMBDA */
    const double *lambda;        /* [L] wavelength axis in Angstrom */
    double lambda0;              /* central wavelength of the line (wlines[1]) */
    double sl, ll, jl;           /* lower level S, L, J (readConfig.c:104-110) */
    double su, lu, ju;           /* upper level S, L, J */
    const double *psf;           /* [L] PSF kernel (ConvolveWithPSF), or NULL */
    int fix[PM_NPAR];            /* 1 = free parameter */
    double weight[4];            /* WeightForStokes */
    double noise;                /* per-point variance, 1/SN^2 */
    double ilambda;              /* InitialDiagonalElement */
    double toplim;               /* 1e-12 in the reference */
    int miter;                   /* NumberOfCycles (> 0) */
    double mu;                   /* heliocentric mu (ah) */
    int logclambda;              /* Diagonal element acceleration */
    double guess[PM_NPAR];       /* initial model (INITIAL_MODEL) */
    int precision;               /* enum pm_precision */
    const float *slight;         /* [4][L] stray-light profile (the .per file of
                                    readPerStrayLightFile, utilsFits.c:835-877),
                                    or NULL */
    int no_classical;            /* 1: start every pixel from the guess itself,
                                    as lm_mils on a .per profile (milos.c:741-775);
                                    0: classical estimates and NaN guards first,
                                    as the FITS pixel loop (milos.c:958-983) */
    int use_fft;                 /* useFFT (readConfig.c:995-999): 1 = the PSF
                                    enters every Stokes row and response row by
                                    circular convolution, and is folded into the
                                    macroturbulence kernel circularly (the DFT
                                    products of mil_sinrf.c:301-386 and
                                    me_der.c:331-470); 0 = the zero-padded
                                    correlation with the continuum trick for I */
} pm_config;

const char *pm_strerror(int status);

/* Wavelength axis: lambda[0] = lambda0 + initial/1000, then accumulated
 * steps of step/1000 (all in mA).  Writes min(n, cap) values, returns n. */
int pm_grid(double lambda0, double initial_mA, double step_mA, double final_mA,
            double *lambda, int cap);

/* Gaussian PSF of the given FWHM (mA) on the axis, unit sum. */
int pm_psf_gaussian(double fwhm_mA, const double *lambda, int nlambda, double *G);

/* NTERMS for a fix vector and guess macroturbulence. */
int pm_nterms(const int fix[PM_NPAR], double guess_mac);

int  pm_ctx_create(const pm_config *cfg, pm_ctx **out);
void pm_ctx_destroy(pm_ctx *ctx);
int  pm_ctx_nterms(const pm_ctx *ctx);

/* Synthesis (and optionally response functions) of n models.
 * models [n][11] double; spectra [n][4][L] double; rf nullable
 * [n][4][NT][L] double in the reference's d_spectra layout.
 * Host pointers, synchronous. */
int pm_synth(pm_ctx *ctx, const double *models, int n, double *spectra, double *rf);

/* The same with device pointers, asynchronous on stream (hipStream_t). */
int pm_synth_device(pm_ctx *ctx, const double *d_models, int n, double *d_spectra,
                    double *d_rf, void *stream);

/* Inversion of npix pixels.  stokes [npix][4][L] float (I,Q,U,V contiguous
 * per pixel, the forParallel layout of utilsFits.c:224,329).
 * Outputs: models [npix][11] double (Init_Model order); chi2 [npix] float
 * (last accepted chi2); iters [npix] int (LM trips, or the pixel index when
 * lm_mils returns -1 for nfree == 0, as milos.c:997 leaves it); accepted
 * nullable [npix] int (chi2-decreasing trips); fit nullable [npix][4][L] float
 * (the last trial synthesis, milosUtils.c:569).  Host pointers, synchronous:
 * from 2^20 pixels on, the pixels go in three chunks (1/16, 14/16, 1/16) on
 * three streams of the context, each chunk's upload and copy-back
 * overlapping the other chunks' kernels (pinned host buffers copy
 * fastest), with the same results as one pm_invert_device call.  The device buffers of the host-pointer calls (pm_synth,
 * pm_invert, pm_invert_sl, pm_classical, pm_ctx_autotune_host) belong to
 * the context: made on first use, grown when a call needs more, kept until
 * pm_ctx_destroy. */
int pm_invert(pm_ctx *ctx, const float *stokes, int npix, double *models,
              float *chi2, int *iters, int *accepted, float *fit);

/* The same with device pointers, asynchronous on stream (hipStream_t). */
int pm_invert_device(pm_ctx *ctx, const float *d_stokes, int npix, double *d_models,
                     float *d_chi2, int *d_iters, int *d_accepted, float *d_fit,
                     void *stream);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): pm_invert_device without the
 * context's work-queue counters.  It allocates nothing and calls no
 * allocating routine; its counters live in a fixed device scratch array of
 * the library (1,088 bytes per GPU, held until the process ends) whose
 * address is looked up once per device by pm_ctx_create.  The counters are
 * cleared by a kernel at the start of every call (no memset), so the call
 * can be captured into a graph.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two streams, or two graph
 * replays in flight) race, so order them.  The current device must be the
 * context's (PM_EINVAL otherwise); npix > PM_NOALLOC_MAX_NPIX gives
 * PM_ERANGE (the work-queue counters are 32-bit).  The eigen-fallback count
 * of these calls is not reported by pm_last_fallback_count.  Otherwise as
 * pm_invert_device (the same kernels, the same results). */
#define PM_NOALLOC_MAX_NPIX (1 << 30)
int pm_invert_device_noalloc(pm_ctx *ctx, const float *d_stokes, int npix, double *d_models,
                             float *d_chi2, int *d_iters, int *d_accepted, float *d_fit,
                             void *stream);

/* Inversion with a stray-light profile per pixel: slight [npix][4][L] float
 * (the 4-D stray-light cube of readFitsStrayLightFile, utilsFits.c:881,
 * indexed per pixel as milos.c:986-996 does).  NULL = the context's profile
 * (pm_config.slight) or none.  Otherwise as pm_invert / pm_invert_device. */
int pm_invert_sl(pm_ctx *ctx, const float *stokes, const float *slight, int npix,
                 double *models, float *chi2, int *iters, int *accepted, float *fit);
int pm_invert_sl_device(pm_ctx *ctx, const float *d_stokes, const float *d_slight, int npix,
                        double *d_models, float *d_chi2, int *d_iters, int *d_accepted,
                        float *d_fit, void *stream);

/* Classical estimates only (NumberOfCycles < 0, milos.c:399-520): per
 * pixel the context's guess with B, vlos, gamma, phi from
 * estimacionesClasicas(forInitialUse = 0) and B_los in the S0 slot
 * (milosUtils.c:449-450); models [npix][11]. */
int pm_classical(pm_ctx *ctx, const float *stokes, int npix, double *models);
int pm_classical_device(pm_ctx *ctx, const float *d_stokes, int npix, double *d_models,
                        void *stream);

/* Number of pixels whose 9x9/10x10/11x11 solve needed the eigen
 * pseudo-inverse fallback in the last inversion on this context (diagnostic;
 * synchronises the context's last stream). */
int pm_last_fallback_count(pm_ctx *ctx, long long *count);

/* ------------------------------------------------------------ tuning and batches */

/* Autotuner: times every inversion-kernel variant this context's sampling
 * allows -- lanes per pixel, wavelength samples per lane, pixels per
 * workgroup -- on npix pixels of d_stokes (device; d_slight as for
 * pm_invert_sl_device) and keeps the fastest for this context's later
 * inversions of that kind (with or without a stray-light cube).  chosen
 * (nullable) = {lanes per pixel, samples per lane, pixels per workgroup}.
 * Synchronises stream.  The pixels per workgroup do not change a pixel's
 * arithmetic; the lane mapping changes the order of the sums over
 * wavelength, so results of two variants agree to rounding (and the
 * inversions to the stated parameter tolerances). */
int pm_ctx_autotune(pm_ctx *ctx, const float *d_stokes, const float *d_slight, int npix,
                    void *stream, int chosen[3]);

/* The same on host buffers (copied to the device for the trials). */
int pm_ctx_autotune_host(pm_ctx *ctx, const float *stokes, const float *slight, int npix,
                         int chosen[3]);

/* The variant in use for the 9-row (general = 0) or general form: out as
 * chosen above; returns the number of variants there are. */
int pm_ctx_variant(const pm_ctx *ctx, int general, int out[3]);

/* Selects variant index (0 = the default mapping) of that form by hand. */
int pm_ctx_set_variant(pm_ctx *ctx, int general, int index);

/* No counterpart in the reference.  Selects a kernel shape of that form by
 * its lanes per pixel, samples per lane and pixels per workgroup, among all
 * the shapes the library has compiled that cover the context's sampling
 * (lanes * samples >= L), including those the autotuner does not offer for
 * this sampling (e.g. 64 x 2 at 30 samples); PM_ENOTSUP if there is none.
 * For testing and tuning; pm_ctx_set_variant and pm_ctx_autotune replace
 * the choice. */
int pm_ctx_set_shape(pm_ctx *ctx, int general, int lanes, int samples,
                     int pixels_per_workgroup);

/* No counterpart in the reference: the mixed form's "chi2 unchanged" exit
 * test (milosUtils.c:579-580, |ochi2 - chi2| * 100 / chi2 < toplim) uses
 * max(toplim, toplim_mixed).  A chi2 summed from FP32 spectra cannot
 * resolve the reference's 1e-14 relative change (the FP32 and FP64 forms'
 * chi2 of one model differ by up to ~1e-4 relative), so near the minimum
 * the reference's test only ends the loop once the damping has made the
 * step smaller than the FP32 rounding of every sample.  Mixed contexts start
 * with PM_MIXED_TOPLIM (|dchi2| / chi2 < 1e-6); 0 restores the reference's
 * test.  The FP64 form ignores it.  It waits for the context's last
 * inversion to finish before it rewrites the context's device constants.
 * 0 <= toplim_mixed < 100, else PM_EINVAL. */
#define PM_MIXED_TOPLIM 1e-4
int pm_ctx_set_mixed_toplim(pm_ctx *ctx, double toplim_mixed);

/* One launch for several inversions (e.g. cubes of different samplings):
 * each item is a pm_invert_sl_device call on device buffers.  They are
 * captured into one HIP graph, one concurrent branch per item with its own
 * work-queue counters, and pm_graph_launch runs the whole batch as one
 * graph launch, as often as wanted (the buffers are fixed at creation).
 * share_by_work > 0 caps each branch at its share of the estimated work of
 * the batch (the branches then run side by side); 0 lets every branch fill
 * the GPU (later branches take over as earlier ones drain). */
typedef struct {
    pm_ctx *ctx;
    const float *d_stokes;
    const float *d_slight;       /* nullable */
    int npix;
    double *d_models;
    float *d_chi2;
    int *d_iters;
    int *d_accepted;             /* nullable */
    float *d_fit;                /* nullable */
} pm_batch_item;

typedef struct pm_graph pm_graph;

int  pm_graph_create(const pm_batch_item *items, int n, double share_by_work, pm_graph **out);
int  pm_graph_launch(pm_graph *g, void *stream);
void pm_graph_destroy(pm_graph *g);

/* ------------------------------------------------------------ serving
 *
 * A server takes scans (Stokes cubes of any size, on any context of its
 * device) as they arrive and inverts them concurrently.  Scans are admitted
 * while the device bytes of the admitted ones stay within hbm_budget (the
 * first queued scan that fits; any scan when none is admitted), uploaded,
 * split into pixel chunks that run on several streams, the next chunk from
 * the admitted scan with the earliest deadline (its submission plus 1.5
 * times its estimated time alone on the GPU; once passed, now plus 1.5
 * times the estimated time of its remaining work), and their results are copied
 * back into the caller's buffers.  A scan's results are the same bits as pm_invert / pm_invert_sl
 * give for it alone.  The caller keeps its buffers alive and unchanged
 * until pm_wait returns for the ticket.  One host thread may submit while
 * others wait.  While a context has scans queued or running, the server's
 * threads use it: make no other call on that context (pm_submit aside)
 * until pm_wait has returned for each of them, and use a second context
 * for other work meanwhile. */
typedef struct pm_server pm_server;

typedef struct {
    size_t hbm_budget;           /* device bytes of admitted scans; 0 = half of the free memory */
    int streams;                 /* compute streams, 1..8 (0 = 3) */
    int chunk_px;                /* chunk size: the work of this many pixels of the 30-sample,
                                    9-parameter workload (0 = 65536) */
} pm_server_opts;

int  pm_server_create(const pm_server_opts *opts, pm_server **out);
/* Finishes every submitted scan, then frees the server. */
void pm_server_destroy(pm_server *srv);

/* Queues a scan; arguments as pm_invert_sl.  PM_ENOMEM when the scan alone
 * needs more than the budget. */
int pm_submit(pm_server *srv, pm_ctx *ctx, const float *stokes, const float *slight, int npix,
              double *models, float *chi2, int *iters, int *accepted, float *fit,
              long long *ticket);

/* 1 when the scan's results are in the caller's buffers, 0 while not. */
int pm_poll(pm_server *srv, long long ticket);

/* Blocks until the scan is done and retires the ticket; latency_s
 * (nullable) = submission to results in host memory, in seconds.  Returns
 * the scan's status; PM_EINVAL for a ticket already retired, also when
 * another thread's pm_wait retired it while this one waited. */
int pm_wait(pm_server *srv, long long ticket, double *latency_s);

#ifdef __cplusplus
}
#endif
#endif
