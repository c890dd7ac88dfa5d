// pm_invert.h -- the inversion, synthesis and estimates kernels (device
// code) of libpmilos_hip: LDS layout, convolutions, the synthesis and
// response-function slots of a pixel group, the persistent work queue and
// the kernels themselves.  Included by pm_kernels.hip, which holds the host
// side (contexts, dispatch, the C ABI).
//
// Kernels
//   pm_invert_kernel  persistent: each group of W lanes pulls pixels from
//                     per-XCD work counters and runs the whole per-pixel path
//                     (classical estimates -> LM loop of lm_mils) on one pixel
//                     at a time, so a converged pixel frees its lanes for the
//                     next one immediately (no padding of short pixels to the
//                     longest one).
//   pm_synth_kernel   synthesis (+ response functions) of given models.
//   pm_classical_kernel  classical estimates only, one thread per pixel.
#pragma once

#include "pm_device.h"

#define PM_BLOCK 256
// threads per workgroup for W-lane pixel groups (W = 8 would need more than
// 64 KiB of LDS for 32 groups; a 128-lane group is its own workgroup)
template <int W> struct BlockOf { static constexpr int value = (W == 8 || W == 128) ? 128 : PM_BLOCK; };
#define PM_CHUNK 4           // pixels taken per work-counter increment
#define PM_NSLICE 8          // one work counter per XCD
#define PM_CSTRIDE 32        // ints between counters (separate cache lines)

// ------------------------------------------------------------ LDS layout

template <typename T, int W, int NLPL, int NT, bool GEN = false, int BLK = BlockOf<W>::value>
struct Lds {
    static constexpr int PADW = 2 * W * NLPL;                 // >= 2L
    // row stride: odd, so that lanes reading the same offset of different
    // rows fall in different LDS banks
    static constexpr int RS = PADW + 1;
    static constexpr int CH = W >= 16 ? 16 : W;
    // lane stride CH+1: no bank conflicts; a 128-lane group adds CH doubles
    // for the exchange of the two waves' sums
    static constexpr int RED_BYTES = ((CH + 1) * W + (W > 64 ? CH : 0)) * 8;
    // rows per PSF pass: 8 (two parameters) when they fit in the space the
    // J^T J reduction needs anyway, else 4
    static constexpr int PAD_ROWS = 8 * RS * (int)sizeof(T) <= RED_BYTES ? 8 : 4;
    static constexpr int PAD_BYTES = PAD_ROWS * RS * (int)sizeof(T);
    static constexpr int EIG_BYTES = (2 * NT * NT + 2 * NT) * 8;
    static constexpr int U0 = PAD_BYTES > EIG_BYTES ? PAD_BYTES : EIG_BYTES;
    // the PSF row buffer, the eigen fallback and the J^T J reduction share
    // one region (the row buffer is re-zeroed after the other two)
    static constexpr int U_BYTES = ((U0 > RED_BYTES ? U0 : RED_BYTES) + 15) / 16 * 16;
    static constexpr int OBS_BYTES = (4 * W * NLPL * 4 + 15) / 16 * 16;
    static constexpr int NP = NT * (NT + 1) / 2;
    static constexpr int AB_BYTES = ((NP + NT) * 8 + 15) / 16 * 16;
    static constexpr int LM_BYTES = 24 * 8;                   // model[11], trial[11]
    // macroturbulence kernel and its v_mac derivative, both PSF-folded
    static constexpr int MAC_BYTES = GEN ? ((2 * W * NLPL * (int)sizeof(T) + 15) / 16 * 16) : 0;
    static constexpr int GROUP_BYTES = U_BYTES + OBS_BYTES + AB_BYTES + LM_BYTES + MAC_BYTES;
    static constexpr int BLOCK = BLK;
    static constexpr int GPB = BLOCK / W;
    static constexpr int BLOCK_BYTES = GPB * GROUP_BYTES;
    static_assert(BLOCK_BYTES <= 65536, "LDS per workgroup above 64 KiB");
    static_assert(W <= 64 || BLK == W, "a 128-lane group is its own workgroup");
};

// ------------------------------------------------------------ convolution

// Lane `src` of the group's value of v (src uniform): a readlane (scalar, no
// LDS trip) when the group is the whole wave, a group shuffle inside a wave,
// through LDS between workgroup barriers for a two-wave group.
template <int W, typename T>
__device__ __forceinline__ T lane_of(T v, int src)
{
    if constexpr (W == 128) {
        __shared__ T xv;
        if ((int)threadIdx.x == src) xv = v;
        __syncthreads();
        const T r = xv;
        __syncthreads();
        return r;
    } else if constexpr (W == 64) {
        if constexpr (sizeof(T) == 8) return readlane_d(v, src);
        else return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
    } else {
        return __shfl(v, src, W);
    }
}

// Zero-padded "same" correlation of R rows x[r][slot] with the PSF, in place
// (convolution.c:54-78).  Rows 0 and (when R == 8) 4 are Stokes I rows and
// use the continuum trick of convolution.c:90-113 with Ic = max(first, last)
// of that row (mil_sinrf.c:390-395, me_der.c:478-483), taken by in-group
// shuffles so the row goes to LDS once.  pad: R rows of 2*W*NLPL values
// (stride one more), zero
// outside [L/2, L/2 + L).  The FP64 form accumulates in double as
// convolution.c:68 does; the mixed form in FP32.
template <typename T, int W, int NLPL, int R, bool IC = true>
__device__ __forceinline__ void convR(PmC c, T *pad, int t, int L,
                                      T x[R][NLPL])
{
    typedef typename AccT<T>::type A;
    constexpr int RS = 2 * W * NLPL + 1;
    const int half = L / 2;
    const int klast = (L - 1) / W, tlast = (L - 1) % W;
    T ic[2] = {T(0), T(0)};
#pragma unroll
    for (int q = 0; q < (IC ? (R == 8 ? 2 : 1) : 0); q++) {
        const int r = 4 * q;
        T last = x[r][0];
#pragma unroll
        for (int k = 1; k < NLPL; k++) if (k == klast) last = x[r][k];
        T a = lane_of<W>(x[r][0], 0), b = lane_of<W>(last, tlast);
        ic[q] = (a > b) ? a : b;
    }
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                T v = x[r][k];
                if (IC && r == 0) v = ic[0] - v;
                if (IC && R == 8 && r == 4) v = ic[1] - v;
                pad[r * RS + half + lam] = v;
            }
        }
    }
    group_sync<W>();
    const int jlo = sizeof(T) == 8 ? c->glo64 : c->glo32;
    const int jhi = sizeof(T) == 8 ? c->ghi64 : c->ghi32;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L) {
            A acc[R];
#pragma unroll
            for (int r = 0; r < R; r++) acc[r] = 0;
            const T *p0 = pad + lam;
            for (int j = jlo; j < jhi; j++) {
                A gj = A(c->G[j]);
#pragma unroll
                for (int r = 0; r < R; r++) acc[r] += gj * (A)p0[r * RS + j];
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                if (IC && r == 0) x[r][k] = T((A)ic[0] - acc[r]);
                else if (IC && R == 8 && r == 4) x[r][k] = T((A)ic[1] - acc[r]);
                else x[r][k] = T(acc[r]);
            }
        }
    }
    group_sync<W>();
}

// Half-width of the register-blocked correlation for rows of NX samples
// (the window start half - HC stays inside the row buffer for L/2 >= HC):
// 10 taps each side in the FP64 form (taps down to 1e-16 of the peak, below
// its unit roundoff), 7 in the FP32 form (down to 1e-8).  That covers
// Gaussian PSFs narrower than 3.01 (FP64) and 3.10 (FP32) samples FWHM; a
// 60 mA Gaussian on a 20 mA grid is 3.0.  Wider PSFs take the tap loop
// (convR), with the same sums.
template <typename T, int NX> struct ConvHC {
    static constexpr int cap = sizeof(T) == 8 ? 10 : 7;
    static constexpr int value = NX / 2 - 1 < cap ? NX / 2 - 1 : cap;
};

// The correlation of convR, register-blocked: each lane of the group takes
// one block of B = R*NX/W consecutive outputs of one row, loads the B + 2*HC
// samples it needs from the row buffer once, and accumulates the taps from
// registers (a tap-outer loop over compile-time offsets, with the taps in
// scalar registers).  Same operations in the same order as convR (taps in
// ascending order, each product added to one accumulator), so the results
// are the same bits.  The outputs go back to the row buffer in place, after
// every lane has loaded its window.
template <typename T, int W, int NLPL, int R>
struct ConvF {
    typedef typename AccT<T>::type A;
    static constexpr int NX = W * NLPL;
    static constexpr int HC = ConvHC<T, NX>::value;
    static constexpr int NB = W / R;                 // blocks per row
    static constexpr int B = NX / NB;                // outputs per lane
    static constexpr int NWIN = B + 2 * HC;
    // used where the window and the accumulators stay well inside the
    // register file
    static constexpr bool ok = W % R == 0 && HC > 0 &&
        (NWIN * (int)sizeof(T) + B * (int)sizeof(A)) / 4 <= 96;
};

template <typename T, int W, int NLPL, int R>
__device__ __forceinline__ bool convF_usable(PmC c, int L)
{
    using CF = ConvF<T, W, NLPL, R>;
    const int reach = sizeof(T) == 8 ? c->reach64 : c->reach32;
    return CF::ok && reach <= CF::HC && CF::HC <= L / 2;
}

template <typename T, int W, int NLPL, int R, bool IC = true>
__device__ __forceinline__ void convF(PmC c, T *pad, int t, int L,
                                      T x[R][NLPL])
{
    using CF = ConvF<T, W, NLPL, R>;
    typedef typename CF::A A;
    constexpr int RS = 2 * W * NLPL + 1, HC = CF::HC, NB = CF::NB, B = CF::B;
    const int half = L / 2;
    const int klast = (L - 1) / W, tlast = (L - 1) % W;
    T ic[2] = {T(0), T(0)};
#pragma unroll
    for (int q = 0; q < (IC ? (R == 8 ? 2 : 1) : 0); q++) {
        const int r = 4 * q;
        T last = x[r][0];
#pragma unroll
        for (int k = 1; k < NLPL; k++) if (k == klast) last = x[r][k];
        T a = lane_of<W>(x[r][0], 0), b = lane_of<W>(last, tlast);
        ic[q] = (a > b) ? a : b;
    }
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                T v = x[r][k];
                if (IC && r == 0) v = ic[0] - v;
                if (IC && R == 8 && r == 4) v = ic[1] - v;
                pad[r * RS + half + lam] = v;
            }
        }
    }
    group_sync<W>();
    {
        // lane t takes block t / R of row t % R: the lanes that read LDS in
        // the same cycle (16 for 8-byte, 32 for 4-byte reads) spread over
        // the rows, whose odd stride puts them in different banks, instead
        // of all taking blocks of one row B elements apart (two banks for
        // B = 8 doubles)
        const int r = t % R, b = t / R;
        const T *src = pad + r * RS + half - HC + b * B;
        T w[CF::NWIN];
#pragma unroll
        for (int i = 0; i < CF::NWIN; i++) w[i] = src[i];
        A acc[B];
#pragma unroll
        for (int o = 0; o < B; o++) acc[o] = 0;
        // every tap of the window; those beyond the PSF's tap range are 0
        // and add exactly 0
        A g[2 * HC + 1];
#pragma unroll
        for (int d = -HC; d <= HC; d++)
            g[HC + d] = sizeof(T) == 8 ? (A)c->gc64[PM_HC + d] : (A)c->gc32[PM_HC + d];
#pragma unroll
        for (int d = -HC; d <= HC; d++)
#pragma unroll
            for (int o = 0; o < B; o++) acc[o] += g[HC + d] * (A)w[o + HC + d];
        group_sync<W>();                       // every window loaded before any store
        const bool irow = IC && (r % 4) == 0;
        const T icr = (R == 8 && r == 4) ? ic[1] : ic[0];
        // one FMA per output: icr - acc on the I rows (the continuum trick,
        // rounded once as the subtraction is), acc + (-0) elsewhere (acc's
        // own bits, signed zeros included)
        const A sg = irow ? A(-1) : A(1), off = irow ? (A)icr : A(-0.0);
        T *dst = pad + r * RS + half + b * B;
        // every output is stored; those at or past L land in the row's
        // zero margin and are cleared again below
#pragma unroll
        for (int o = 0; o < B; o++) dst[o] = T(fma(acc[o], sg, off));
    }
    group_sync<W>();
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L) {
#pragma unroll
            for (int r = 0; r < R; r++) x[r][k] = pad[r * RS + half + lam];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) pad[r * RS + half + lam] = T(0);
        }
    }
    group_sync<W>();
}

// Circular convolution with the kernel h (length L: a kernel in LDS, or the
// PSF of the constants in the FFT mode), rotated so that
// out[m] = conv[(m + L/2) mod L] (convCircular, convolution.c:126-153), over
// the kernel's taps [jlo, jhi): the others are below the rounding of the
// sum (see Taps).  The rows are staged in the middle of the PSF row buffer,
// leaving its zero margins untouched.
template <typename T, int W, int NLPL, int R, typename HP>
__device__ __forceinline__ void convC(T *pad, HP h, int jlo, int jhi, int t, int L,
                                      T x[R][NLPL])
{
    typedef typename AccT<T>::type A;
    constexpr int RS = 2 * W * NLPL + 1;
    const int half = L / 2;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L)
#pragma unroll
            for (int r = 0; r < R; r++) pad[r * RS + half + lam] = x[r][k];
    }
    group_sync<W>();
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int m = t + k * W;
        if (m < L) {
            int i = m + half;
            if (i >= L) i -= L;
            A acc[R];
#pragma unroll
            for (int r = 0; r < R; r++) acc[r] = 0;
            int idx = i - jlo;
            if (idx < 0) idx += L;
            for (int j = jlo; j < jhi; j++) {
                A hj = (A)h[j];
#pragma unroll
                for (int r = 0; r < R; r++) acc[r] += hj * (A)pad[r * RS + half + idx];
                idx = idx == 0 ? L - 1 : idx - 1;
            }
#pragma unroll
            for (int r = 0; r < R; r++) x[r][k] = T(acc[r]);
        }
    }
    group_sync<W>();
}

// The circular convolution of convC, register-blocked: the rows are stored
// twice (sample m at TC + ((m - L/2) mod L) and L further on), so that the
// samples of output m under tap j sit at TC + m - j + L and each lane's
// block of B consecutive outputs reads one contiguous window per chunk of
// TC taps.  The taps come from the pixel's kernel in LDS, in ascending order
// as in convC; the last chunk reads past the window's end, where the kernel
// holds exact zeros (mac_kernels), so the results are the same.
// Every lane loads its windows before any lane stores; the doubled rows are
// cleared afterwards (the other convolutions expect zero margins).
template <typename T, int W, int NLPL, int R>
struct ConvCF {
    typedef typename AccT<T>::type A;
    static constexpr int NX = W * NLPL;
    static constexpr int NB = W / R;
    static constexpr int B = NX / NB;
    static constexpr int TC = 8;
    static constexpr int NWIN = B + TC - 1;
    static constexpr bool ok = W % R == 0 &&
        (NWIN * (int)sizeof(T) + B * (int)sizeof(A)) / 4 <= 64;
};

template <typename T, int W, int NLPL, int R>
__device__ __forceinline__ bool convCF_usable(int L)
{
    // every window stays inside its row of 2*W*NLPL + 1
    return ConvCF<T, W, NLPL, R>::ok && L + ConvCF<T, W, NLPL, R>::TC - 1 <= W * NLPL;
}

template <typename T, int W, int NLPL, int R>
__device__ __forceinline__ void convCF(T *pad, const T *h, int jlo, int jhi, int t, int L,
                                       T x[R][NLPL])
{
    using CF = ConvCF<T, W, NLPL, R>;
    typedef typename CF::A A;
    constexpr int RS = 2 * W * NLPL + 1, NB = CF::NB, B = CF::B, TC = CF::TC;
    const int half = L / 2;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        const int lam = t + k * W;
        if (lam < L) {
            int q = lam - half;
            if (q < 0) q += L;
#pragma unroll
            for (int r = 0; r < R; r++) {
                pad[r * RS + TC + q] = x[r][k];
                pad[r * RS + TC + q + L] = x[r][k];
            }
        }
    }
    group_sync<W>();
    // lane t takes block t / R of row t % R (bank spread, as in convF)
    const int r = t % R, b = t / R, m0 = b * B;
    A acc[B];
#pragma unroll
    for (int o = 0; o < B; o++) acc[o] = 0;
    if (m0 < L) {
        const T *row = pad + r * RS + m0 + L + 1;
        for (int j0 = jlo; j0 < jhi; j0 += TC) {
            const T *src = row - j0;
            T w[CF::NWIN];
#pragma unroll
            for (int i = 0; i < CF::NWIN; i++) w[i] = src[i];
#pragma unroll
            for (int u = 0; u < TC; u++) {
                const A hj = (A)h[j0 + u];
#pragma unroll
                for (int o = 0; o < B; o++) acc[o] += hj * (A)w[o + TC - 1 - u];
            }
        }
    }
    group_sync<W>();
    T *dst = pad + r * RS + TC;
#pragma unroll
    for (int o = 0; o < B; o++)
        if (m0 + o < L) dst[m0 + o] = T(acc[o]);
    group_sync<W>();
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        const int lam = t + k * W;
        if (lam < L)
#pragma unroll
            for (int r2 = 0; r2 < R; r2++) x[r2][k] = pad[r2 * RS + TC + lam];
    }
    group_sync<W>();
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        const int lam = t + k * W;
        if (lam < L) {
            int q = lam - half;
            if (q < 0) q += L;
#pragma unroll
            for (int r2 = 0; r2 < R; r2++) {
                pad[r2 * RS + TC + q] = T(0);
                pad[r2 * RS + TC + q + L] = T(0);
            }
        }
    }
    group_sync<W>();
}

// The macroturbulence convolutions: register-blocked where the rows fit.
template <typename T, int W, int NLPL, int R>
__device__ __forceinline__ void convMac(T *pad, const T *h, int jlo, int jhi, int t, int L,
                                       T x[R][NLPL])
{
    if (convCF_usable<T, W, NLPL, R>(L)) convCF<T, W, NLPL, R>(pad, h, jlo, jhi, t, L, x);
    else convC<T, W, NLPL, R>(pad, h, jlo, jhi, t, L, x);
}

// The PSF applied to R rows: zero-padded correlation with the continuum
// trick on I (direct mode), or circular convolution of every row (FFT
// mode, mil_sinrf.c:362-386, me_der.c:441-470: the DFT product with the
// PSF's transform and the -L/2 rotation is convCircular with the PSF as
// kernel).  The FFT mode is handled by the general form only.  The direct
// correlation is register-blocked where the PSF's taps fit (convF).
template <typename T, int W, int NLPL, int R, bool GEN, bool IC = true>
__device__ __forceinline__ void convPSF(PmC c, T *pad, int t, int L,
                                        T x[R][NLPL])
{
    if (GEN && c->use_fft) {
        const int jlo = sizeof(T) == 8 ? c->glo64 : c->glo32;
        const int jhi = sizeof(T) == 8 ? c->ghi64 : c->ghi32;
        convC<T, W, NLPL, R>(pad, c->G, jlo, jhi, t, L, x);
    } else if (ConvF<T, W, NLPL, R>::ok && convF_usable<T, W, NLPL, R>(c, L)) {
        convF<T, W, NLPL, R, IC>(c, pad, t, L, x);
    } else {
        convR<T, W, NLPL, R, IC>(c, pad, t, L, x);
    }
}

// Tap windows of the two macroturbulence kernels: [lo, hi) holds every tap
// whose magnitude reaches 1e-16 (FP64 form) / 1e-8 (FP32 form) of the
// kernel's largest, below each form's unit roundoff; the rest are below the
// rounding of the sums they enter.
struct Taps {
    int mlo, mhi, dlo, dhi;
    double msum;         // sum of the kernel's taps in [mlo, mhi)
};

template <typename T, int W, int NLPL>
__device__ __forceinline__ void tap_window(const T v[NLPL], int t, int L, int &lo, int &hi)
{
    double mx = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++)
        if (t + k * W < L) mx = fmax(mx, fabs((double)v[k]));
    mx = group_max<W>(mx);
    const double thr = mx * (sizeof(T) == 8 ? 1e-16 : 1e-8);
    int a = L, b = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam < L && fabs((double)v[k]) >= thr) { a = min(a, lam); b = max(b, lam + 1); }
    }
    lo = group_min_int<W>(a);
    hi = group_max_int<W>(b);
    if (lo >= hi) { lo = 0; hi = 0; }
}

// Macroturbulence kernel of model m and its v_mac derivative (fgauss,
// fgauss.c:16-66; the derivative as me_der.c:328-329,410), both correlated
// with the PSF when there is one (mil_sinrf.c:348-350, me_der.c:412-416),
// or circularly convolved with it in the FFT mode (mil_sinrf.c:302-314,
// me_der.c:331-354); written to gm / gd in LDS, with their tap windows.
template <typename T, int W, int NLPL>
__device__ __forceinline__ void mac_kernels(PmC c, const double *m,
                                            T *pad, int t, int L, T *gm, T *gd, Taps &tp)
{
    // in the arithmetic of the form (FP32 in the mixed form, as the
    // reference's default build); the wavelength offset from the centre is
    // formed in FP64 in both
    typedef typename AccT<T>::type A;
    const A MC = A(m[9]);
    const A ild = (A(c->lambda0) * MC) / A(2.99792458e5);
    const double centro = c->lam[L / 2];
    A gv[NLPL], a2[NLPL];
    double part = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        gv[k] = 0;
        a2[k] = 0;
        if (lam < L) {
            A a = A(c->lam[lam] - centro) / ild;
            A term = (a * a) / A(2);
            gv[k] = term < A(1e30) ? exp(-term) : A(0);
            a2[k] = a * a;
            part += (double)gv[k];
        }
    }
    const A sum = A(group_sum<W>(part));
    T x[2][NLPL];
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        A g = gv[k] / sum;
        x[0][k] = T(g);
        x[1][k] = T(g / MC * (a2[k] - A(1)));
    }
    if (c->use_psf) convPSF<T, W, NLPL, 2, true, false>(c, pad, t, L, x);
    tap_window<T, W, NLPL>(x[0], t, L, tp.mlo, tp.mhi);
    tap_window<T, W, NLPL>(x[1], t, L, tp.dlo, tp.dhi);
    // every slot of gm / gd, with exact zeros outside the tap windows: the
    // register-blocked convolution reads whole chunks of taps past a
    // window's end without a test (the convolutions use no tap outside
    // their window)
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        gm[lam] = (lam >= tp.mlo && lam < tp.mhi) ? x[0][k] : T(0);
        gd[lam] = (lam >= tp.dlo && lam < tp.dhi) ? x[1][k] : T(0);
    }
    double tsum = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        if (lam >= tp.mlo && lam < tp.mhi) tsum += (double)x[0][k];
    }
    tp.msum = group_sum<W>(tsum);
    group_sync<W>();
}

// What the general form keeps from a synthesis for its response functions.
template <typename T, int NLPL>
struct Extra {
    T Sraw[4][NLPL];     // before any convolution (spectra_mac)
    T Ssl[4][NLPL];      // after convolution, before the stray-light mix
    bool mac;            // macroturbulence applied (v_mac > 1e-4)
    Taps tp;             // tap windows of the macroturbulence kernels
};

// Response rows that are 0 by construction: the azimuth's on I and V
// (me_der.c:309-312) and S0's off I, never convolved (me_der.c:432).
__device__ __forceinline__ constexpr bool lit0(int p, int s)
{
    return (p == 6 && (s == 0 || s == 3)) || (p == 7 && s != 0);
}

// Synthesis of model m for this lane's wavelength slots: fills the sample
// state (for the RFs) and the emergent Stokes vector S, convolved
// (mil_sinrf.c:287-403) and mixed with the stray light (:405-412).  The
// general form (GEN) also handles macroturbulence and stray light.
template <typename T, int W, int NLPL, bool GEN>
__device__ __forceinline__ void synth_slots(PmC c, const double *m,
                                            const float *slp, T *pad, T *gm, T *gd, int t,
                                            int L, Geom<T> &g, Samp<T> sp[NLPL], T S[4][NLPL],
                                            Extra<T, NLPL> &ex)
{
    make_geom<T>(c, m, g);
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        int lam = t + k * W;
        int li = lam < L ? lam : L - 1;
        profiles<T>(c, g, c->dlam[li], sp[k]);
        UR<T> r;
        unno<T>(g, sp[k], r);
        T s4[4];
        stokes<T>(g, r, s4);
#pragma unroll
        for (int s = 0; s < 4; s++) S[s][k] = s4[s];
    }
    // every field of ex is defined on every call (whatever the branches
    // below take), so that none stays live in a register across the LM loop
    ex.mac = false;
    ex.tp = Taps{0, 0, 0, 0, 0.0};
    if (GEN) {
#pragma unroll
        for (int k = 0; k < NLPL; k++)
#pragma unroll
            for (int s = 0; s < 4; s++) ex.Sraw[s][k] = S[s][k];
        ex.mac = m[9] > 0.0001;                                    // mil_sinrf.c:288
    }
    if (ex.mac) {
        mac_kernels<T, W, NLPL>(c, m, pad, t, L, gm, gd, ex.tp);
        convMac<T, W, NLPL, 4>(pad, gm, ex.tp.mlo, ex.tp.mhi, t, L, S); // :353-354
    } else if (c->use_psf) {
        convPSF<T, W, NLPL, 4, GEN>(c, pad, t, L, S);              // :362-402
    }
    if (GEN) {
#pragma unroll
        for (int k = 0; k < NLPL; k++)
#pragma unroll
            for (int s = 0; s < 4; s++) ex.Ssl[s][k] = S[s][k];
    }
    if (GEN && slp) {                                              // :405-412
        const T alf = T(m[10]);
#pragma unroll
        for (int k = 0; k < NLPL; k++) {
            int lam = t + k * W;
            int li = lam < L ? lam : L - 1;
#pragma unroll
            for (int s = 0; s < 4; s++)
                S[s][k] = S[s][k] * alf + T(slp[s * L + li]) * (T(1) - alf);
        }
    }
}

// Response functions of the synthesis held in sp, convolved like the
// spectra (me_der.c:316-497, the S0 row never), two parameters per pass
// where the row buffer holds 8 rows.  The general form adds the
// macroturbulence row (row 9, me_der.c:418-425) and the stray-light scaling
// and filling-factor row (me_der.c:499-525); ff_src is the reference's
// "spectra_slight" argument of that call: the synthesis before the mix on
// the first call (milosUtils.c:537), the mixed one after an accepted step
// (:587), selected by ff_pre.
template <typename T, int W, int NLPL, int NT, bool GEN>
__device__ __forceinline__ void rf_slots(PmC c, const double *m,
                                         const float *slp, const Geom<T> &g,
                                         const Samp<T> sp[NLPL], T *pad, const T *gm,
                                         const T *gd, int t, int L, const Extra<T, NLPL> &ex,
                                         bool ff_pre, const T S[4][NLPL], T D[NLPL][4][NT])
{
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        UR<T> r;
        unno<T>(g, sp[k], r);
        response<T, NT>(c, g, sp[k], r, D[k]);
    }
    // rows per PSF pass: the buffer's 8 where they fit, 4 in the 9-row FP32
    // form (its register-blocked windows stay small)
    constexpr int R = (!GEN && sizeof(T) == 4) ? 4 : Lds<T, W, NLPL, NT, GEN>::PAD_ROWS;
    constexpr int PPC = R / 4;
    // rows 0..5 in full passes, S0's never; the azimuth's and S1's come
    // from the convolved spectrum (below), S1's in the FFT mode without
    // macroturbulence from a pass of its own
    constexpr int NPS = 6;
    constexpr int NFULL = NPS / PPC;
    const bool mac = GEN && ex.mac;
    if (mac && NT >= 10) {                      // v_mac row from the raw synthesis
        T x[4][NLPL];
#pragma unroll
        for (int k = 0; k < NLPL; k++)
#pragma unroll
            for (int s = 0; s < 4; s++) x[s][k] = ex.Sraw[s][k];
        convMac<T, W, NLPL, 4>(pad, gd, ex.tp.dlo, ex.tp.dhi, t, L, x);
#pragma unroll
        for (int k = 0; k < NLPL; k++)
#pragma unroll
            for (int s = 0; s < 4; s++) D[k][s][NT >= 10 ? 9 : 0] = x[s][k];
    }
    if (mac || c->use_psf) {
#pragma unroll
        for (int q0 = 0; q0 < NFULL * PPC; q0 += PPC) {
            T x[R][NLPL];
#pragma unroll
            for (int u = 0; u < PPC; u++) {
                const int q = q0 + u, p = q < 7 ? q : q + 1;
#pragma unroll
                for (int k = 0; k < NLPL; k++)
#pragma unroll
                    for (int s = 0; s < 4; s++) x[4 * u + s][k] = D[k][s][p];
            }
            if (mac) convMac<T, W, NLPL, R>(pad, gm, ex.tp.mlo, ex.tp.mhi, t, L, x);
            else convPSF<T, W, NLPL, R, GEN>(c, pad, t, L, x);
#pragma unroll
            for (int u = 0; u < PPC; u++) {
                const int q = q0 + u, p = q < 7 ? q : q + 1;
#pragma unroll
                for (int k = 0; k < NLPL; k++)
#pragma unroll
                    for (int s = 0; s < 4; s++) D[k][s][p] = x[4 * u + s][k];
            }
        }
        if (NPS % PPC) {                        // the last parameter alone
            constexpr int q = NFULL * PPC, p = q < 7 ? q : q + 1;
            T x[4][NLPL];
#pragma unroll
            for (int k = 0; k < NLPL; k++)
#pragma unroll
                for (int s = 0; s < 4; s++) x[s][k] = D[k][s][p];
            if (mac) convMac<T, W, NLPL, 4>(pad, gm, ex.tp.mlo, ex.tp.mhi, t, L, x);
            else convPSF<T, W, NLPL, 4, GEN>(c, pad, t, L, x);
#pragma unroll
            for (int k = 0; k < NLPL; k++)
#pragma unroll
                for (int s = 0; s < 4; s++) D[k][s][p] = x[s][k];
        }
        if (GEN && !mac && c->use_fft) {
            T x[4][NLPL];
#pragma unroll
            for (int k = 0; k < NLPL; k++)
#pragma unroll
                for (int s = 0; s < 4; s++) x[s][k] = D[k][s][8];
            convPSF<T, W, NLPL, 4, GEN>(c, pad, t, L, x);
#pragma unroll
            for (int k = 0; k < NLPL; k++)
#pragma unroll
                for (int s = 0; s < 4; s++) D[k][s][8] = x[s][k];
        } else {
            // the S1 row is (I - S0)/S1 on I and Q/S1, U/S1, V/S1 before the
            // convolution, which maps it to the same expressions of the
            // convolved spectrum of this model (before any stray-light mix):
            // with the PSF alone through the continuum trick, with
            // macroturbulence through the circular sum of the taps, which
            // takes a constant c to c times their sum
            typedef typename AccT<T>::type A;
            const A s0 = (A)(m[7] * (mac ? ex.tp.msum : 1.0)), is1 = A(1) / (A)m[8];
#pragma unroll
            for (int k = 0; k < NLPL; k++)
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const T sc = (GEN && slp) ? ex.Ssl[s][k] : S[s][k];
                    D[k][s][8] = T(((A)sc - (s == 0 ? s0 : A(0))) * is1);
                }
        }
    }
    {
        // The azimuth turns the frame of the linear polarisation: Q + iU of
        // the model at azimuth phi is that at 0 times exp(2i phi), the same
        // at every wavelength, so d(Q, U)/d(phi) = (-2U, 2Q) per radian.  The
        // rows of me_der.c:183-199 and :205-277 (through the convolutions,
        // which are linear and act per row) are therefore (-2U, 2Q) pi/180
        // of this model's convolved spectrum before any stray-light mix; the
        // mix then scales them by alpha below like every other row.
        typedef typename AccT<T>::type A;
        const A tw = A(2.0 * PM_PI / 180.0);
#pragma unroll
        for (int k = 0; k < NLPL; k++) {
            const T q = (GEN && slp) ? ex.Ssl[1][k] : S[1][k];
            const T u = (GEN && slp) ? ex.Ssl[2][k] : S[2][k];
            D[k][1][6] = T(-tw * (A)u);
            D[k][2][6] = T(tw * (A)q);
        }
    }
    if (GEN && NT == 11 && slp) {
        const T alf = T(m[10]);
        const int ff = 10;                     // the filling factor's row
#pragma unroll
        for (int k = 0; k < NLPL; k++) {
            int lam = t + k * W;
            int li = lam < L ? lam : L - 1;
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int p = 0; p < NT; p++) {
                    D[k][s][p] = D[k][s][p] * alf;
                    if (p == ff) D[k][s][p] = (ff_pre ? ex.Ssl[s][k] : S[s][k]) - T(slp[s * L + li]);
                }
        }
    }
    // the azimuth rows of I and V are 0 (me_der.c:309-312) and stay 0 through
    // the convolutions and the stray-light scaling; restated so that their
    // J^T J terms drop out
#pragma unroll
    for (int k = 0; k < NLPL; k++) { D[k][0][6] = T(0); D[k][3][6] = T(0); }
}

// ------------------------------------------------------------ work queue

struct InvArgs {
    const PmConst *c;
    const float *stokes;
    int npix;
    int slice;
    double *models;
    float *chi2;
    int *iters;
    int *accepted;
    float *fit;
    int *counters;
    unsigned long long *fallbacks;
    const float *slight;     // per-pixel stray light [npix][4][L], or null
};

__device__ __forceinline__ void fetch_chunk(int *counters, int xcd, int npix, int slice,
                                            int &sl_off, int &lo_out, int &hi_out)
{
    lo_out = -1;
    hi_out = -1;
    while (sl_off < PM_NSLICE) {
        int s = (xcd + sl_off) & (PM_NSLICE - 1);
        int lo = s * slice;
        int hi = min(npix, lo + slice);
        if (lo < hi) {
            int v = atomicAdd(&counters[s * PM_CSTRIDE], PM_CHUNK);
            if (lo + v < hi) {
                lo_out = lo + v;
                hi_out = min(hi, lo + v + PM_CHUNK);
                return;
            }
        }
        sl_off++;
    }
}

__global__ void pm_reset_counters(int *counters, unsigned long long *fallbacks)
{
    int i = threadIdx.x;
    if (i < PM_NSLICE) counters[i * PM_CSTRIDE] = 0;
    if (i == 0) *fallbacks = 0ull;
}

// ------------------------------------------------------------ inversion kernel

// Waves per SIMD each inversion kernel is compiled for (the registers per
// lane it may use: 512 / waves): three for the one-sample 9-row FP32
// kernels (168 registers, where LDS allows three workgroups per CU), two for
// the other FP32 kernels, the one-sample 9-row FP64 ones and the two-wave
// (128-lane) FP64 general one, one for the other FP64 general and
// two-sample kernels (which need more than 256).  The FP32
// general kernels spill a few dozen registers at two waves and are still
// faster than at one (the 112-sample cube 1.6x).
template <typename T, int W, int NLPL, bool GEN>
struct InvOcc {
    static constexpr int value =
        (sizeof(T) == 4 && !GEN && NLPL == 1 && W >= 16) ? 3
        : (sizeof(T) == 4 || (NLPL == 1 && (!GEN || W > 64))) ? 2 : 1;
};

// One pixel per group: lm_mils (milosUtils.c:466-635) preceded by the pixel
// loop's guess copy, classical estimates and NaN guards (milos.c:958-983).
//
// GEN = false: NTERMS 9 with neither macroturbulence nor stray light.
// GEN = true, NT = 11: the reference's NTERMS 9, 10 or 11 in the 11-row
// layout (v_mac in row 9, the filling factor in row 10); rows the
// reference's run does not have are fixed rows here, zeroed like FijaACero
// does, so alpha, beta and delta agree with the reference's NTERMS-sized
// ones on the rows it has.  nfree uses the reference's NTERMS.
// GEN = true, NT = 10: the reference's own 10-row layout of a run with v_mac
// free and no stray light (v_mac in row 9, no filling-factor row): the same
// code without the 11th row in the response functions, J^T J and the solve.
template <typename T, int W, int NLPL, int NT, bool GEN, int BLK>
__global__ void __launch_bounds__(BLK, (InvOcc<T, W, NLPL, GEN>::value))
pm_invert_kernel(InvArgs a)
{
    using LD = Lds<T, W, NLPL, NT, GEN, BLK>;
    constexpr int NP = LD::NP;
    __shared__ __align__(16) char smem[LD::BLOCK_BYTES];
    PmC c = (PmC)a.c;
    const int L = c->L;
    const int grp = threadIdx.x / W, t0 = threadIdx.x % W;
    int t = t0;
    char *gb = smem + grp * LD::GROUP_BYTES;
    T *pad = (T *)gb;
    float *sobs = (float *)(gb + LD::U_BYTES);
    double *ab = (double *)(gb + LD::U_BYTES + LD::OBS_BYTES);
    // the group-uniform model (lm[0..10]) and trial model (lm[11..21]) live
    // in LDS; lane 0 writes them
    double *lm = (double *)(gb + LD::U_BYTES + LD::OBS_BYTES + LD::AB_BYTES);
    double *red = (double *)pad;     // J^T J reduction (shares the row buffer)
    T *gm = (T *)(gb + LD::U_BYTES + LD::OBS_BYTES + LD::AB_BYTES + LD::LM_BYTES);
    T *gd = gm + W * NLPL;
    for (int i = t; i < LD::PAD_ROWS * LD::RS; i += W) pad[i] = T(0);
    group_sync<W>();

    const int xcd = blockIdx.x % PM_NSLICE;
    int sl_off = 0, cur = 0, end = 0;

    double flambda = 0, ochisqr = 0, chisqr0 = 0, factor = 1.0;
    // damping at which the current alpha last passed the eigenvalue bound
    // of the solve (negative: not yet)
    double chk_lambda = -1.0;
    int it = 0, nacc = 0, nfree = 1, pix = -1;
    bool iterating = false;
    // the observed samples stay in LDS (sobs) for the pixel's whole LM loop
    // and are read there where needed, not carried in registers
    T S[4][NLPL];
    Samp<T> sp[NLPL];
    Geom<T> g;
    Extra<T, NLPL> ex;
    const float *slp = nullptr;

    while (true) {
        // lane-derived values recomputed per trip (not in the 9-row FP64
        // kernels, whose occupancy it does not change and which it slows)
        if (GEN || sizeof(T) == 4) t = opaque(t0);
        bool first = false;
        if (!iterating) {
            if (cur >= end) {
                int lo = -1, hi = -1;
                if (t == 0) fetch_chunk(a.counters, xcd, a.npix, a.slice, sl_off, lo, hi);
                if constexpr (W > 64) {
                    lo = lane_of<W>(lo, 0);
                    hi = lane_of<W>(hi, 0);
                    sl_off = lane_of<W>(sl_off, 0);
                } else {
                    lo = __shfl(lo, 0, W);
                    hi = __shfl(hi, 0, W);
                    sl_off = __shfl(sl_off, 0, W);
                }
                if (lo < 0) break;
                cur = lo;
                end = hi;
            }
            pix = cur++;
            const float *src = a.stokes + (size_t)pix * 4 * L;
            if (GEN)                                          // milos.c:986-996
                slp = a.slight ? a.slight + (size_t)pix * 4 * L
                               : (c->use_slight ? (const float *)c->slight : nullptr);
            int ng = 0;
#pragma unroll
            for (int k = 0; k < NLPL; k++) {
                int lam = t + k * W;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    float v = lam < L ? src[s * L + lam] : 0.0f;
                    if (lam < L) {
                        sobs[s * L + lam] = v;
                        ng += (v <= -1.0f);                 // milosUtils.c:484-492
                    }
                }
            }
            group_sync<W>();
            nfree = 4 * L - c->NT - group_sum_int<W>(ng);     // :507-508
            if (c->no_classical) {
                if (t == 0)
                    for (int i = 0; i < 11; i++) lm[i] = lm[11 + i] = c->guess[i];
                group_sync<W>();
            } else {
                Classic e = classical_group<W, NLPL>(c, sobs, true, t);
                group_sync<W>();
                if (t == 0) {
                    for (int i = 0; i < 11; i++) lm[i] = c->guess[i];
                    lm[1] = e.B;
                    lm[2] = e.vlos;
                    lm[5] = e.gm;
                    lm[6] = e.az;
                    for (int i = 0; i < 11; i++) lm[11 + i] = lm[i];
                }
                group_sync<W>();
            }
            if (nfree == 0) {                                 // :509-512, milos.c:997
                if (t == 0) {
                    for (int i = 0; i < 11; i++) a.models[(size_t)pix * 11 + i] = lm[i];
                    a.chi2[pix] = 0.0f;
                    a.iters[pix] = pix;
                    if (a.accepted) a.accepted[pix] = 0;
                }
                if (a.fit) {
                    for (int k = 0; k < NLPL; k++) {
                        int lam = t + k * W;
                        if (lam < L)
                            for (int s = 0; s < 4; s++)
                                a.fit[(size_t)pix * 4 * L + s * L + lam] = 0.0f;
                    }
                }
                continue;
            }
            flambda = c->ilambda;
            it = 0;
            nacc = 0;
            factor = 1.0;
            first = true;
        } else {
            // covar = alpha with the diagonal scaled by (1 + lambda), :559-563
            double delta[NT];
            const bool check = !(chk_lambda >= 0.0 && flambda >= chk_lambda);
            bool ok = false;
            if constexpr (W > 64) {
                // a pixel over two waves: the first solves, the second takes
                // its step and verdict from LDS (the slots of the J^T J
                // cross-wave sum, past the row buffer, idle here) instead of
                // repeating the solve
                static_assert((LD::PAD_BYTES + 7) / 8 <= (LD::CH + 1) * W &&
                              NT + 1 <= LD::CH, "solve exchange slots past the row buffer");
                double *xs = red + (LD::CH + 1) * W;
                if (t < 64) {
                    ok = ldl_solve<NT>(ab, 1.0 + flambda, delta, check);
                    if (t == 0) {
#pragma unroll
                        for (int i = 0; i < NT; i++) xs[i] = delta[i];
                        xs[NT] = ok ? 1.0 : 0.0;
                    }
                }
                group_sync<W>();
                if (t >= 64) {
#pragma unroll
                    for (int i = 0; i < NT; i++) delta[i] = xs[i];
                    ok = xs[NT] != 0.0;
                }
            } else {
                ok = ldl_solve<NT>(ab, 1.0 + flambda, delta, check);
            }
            if (ok && check) chk_lambda = flambda;
            if (!ok) {
                double *ws = (double *)pad;
                for (int e = t; e < NT * NT; e += W) {
                    int i = e / NT, j = e % NT;
                    int ii = i > j ? i : j, jj = i > j ? j : i;
                    double v = ab[ii * (ii + 1) / 2 + jj];
                    ws[e] = (i == j) ? v * (1.0 + flambda) : v;
                }
                for (int i = t; i < NT; i += W) ws[2 * NT * NT + i] = ab[NP + i];
                group_sync<W>();
                jacobi_pinv<W, NT>(ws, t);
#pragma unroll
                for (int i = 0; i < NT; i++) delta[i] = ws[2 * NT * NT + NT + i];
                group_sync<W>();
                for (int i = t; i < LD::PAD_ROWS * LD::RS; i += W) pad[i] = T(0);
                group_sync<W>();
                if (t == 0) atomicAdd(a.fallbacks, 1ull);
            }
            {
                double cur[11], tr[11];
#pragma unroll
                for (int i = 0; i < 11; i++) { cur[i] = lm[i]; tr[i] = lm[11 + i]; }
                aplica_delta(c, NT, cur, delta, tr);          // :566
                check_model(tr);                              // :568
                group_sync<W>();
                if (t == 0)
                    for (int i = 0; i < 11; i++) lm[11 + i] = tr[i];
                group_sync<W>();
            }
        }

        // trial synthesis and chi2, :569-571 (fchisqr, lib.c:126-152)
        synth_slots<T, W, NLPL, GEN>(c, lm + 11, slp, pad, gm, gd, t, L, g, sp, S, ex);
        double part = 0;
#pragma unroll
        for (int k = 0; k < NLPL; k++) {
            int lam = t + k * W;
            if (lam < L) {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const float o = sobs[s * L + lam];
                    double isg = o <= -1.0f ? 1e-21 : c->inv_noise;
                    double d = (double)S[s][k] - (double)o;
                    part += (((d * d) * c->weight[s]) * isg);
                }
            }
        }
        double c2 = group_sum<W>(part) / (double)nfree;

        bool do_rf, finished = false;
        if (first) {
            ochisqr = c2;                                      // :546-547
            chisqr0 = c2;
            do_rf = true;
            iterating = true;
        } else {
            // the mixed form may test against the resolution of its chi2
            // (spectra in FP32) instead of the reference's toplim alone
            const double tl = sizeof(T) == 4 ? fmax(c->toplim, c->toplim_mixed) : c->toplim;
            bool clanda = (fabs((ochisqr - c2) * 100 / c2) < tl) || (c2 < 0.0001);
            bool acc = (c2 - ochisqr < 0.);
            if (acc) {                                         // :581-606
                flambda = flambda / (5.0 * factor);
                for (int i = t; i < 11; i += W) lm[i] = lm[11 + i];
                group_sync<W>();
                ochisqr = c2;
                nacc++;
            } else {                                           // :607-615
                flambda = flambda * 10.0 * factor;
            }
            if ((flambda > 1e+7) || (flambda < 1e-25)) clanda = true;
            it++;
            if (c->logclambda)                                 // :621
                factor = (double)((float)log10((double)(float)c2) /
                                  (float)log10((double)(float)chisqr0));
            finished = !(it < c->miter && !clanda);
            // the reference also recomputes J^T J after the final accepted
            // trip; nothing reads it, so it is skipped
            do_rf = acc && !finished;
        }

        if (do_rf) {
            chk_lambda = -1.0;
            T D[NLPL][4][NT];
            rf_slots<T, W, NLPL, NT, GEN>(c, lm, slp, g, sp, pad, gm, gd, t, L, ex, first, S, D);
            // covarm2 (lib.c:71-116): each alpha/beta entry is a per-lane
            // partial sum followed by the group sum; lane 0 stores it in LDS
            typedef typename AccT<T>::type A;
            // the weights w/sigma^2 are folded into the rows once, as their
            // square roots, so that every J^T J term is one product; the
            // gradient takes sqrt(w/sigma^2)(S - obs).  Slots beyond L get
            // weight 0 (their rows repeat sample L-1, finite, so every term
            // is an exact 0).  The literal zeros of the azimuth I/V rows and
            // of the S0 row off I stay literal, and their products are not
            // formed (each would add an exact 0).  Rows of fixed parameters
            // are zeroed after the sums (zbits).
            A rs[NLPL][4];
#pragma unroll
            for (int k = 0; k < NLPL; k++) {
                const int lam = t + k * W;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const float o = lam < L ? sobs[s * L + lam] : 0.0f;
                    const A sw = lam < L ? A(c->swt[o <= -1.0f ? 1 : 0][s]) : A(0);
                    rs[k][s] = sw * A((double)S[s][k] - (double)o);
#pragma unroll
                    for (int p = 0; p < NT; p++) {
                        if (lit0(p, s)) continue;
                        D[k][s][p] = T((A)D[k][s][p] * sw);
                    }
                }
            }
            // The NP + NT sums are reduced CH at a time through LDS: every lane
            // stores its partials, then lane t adds entry t % CH over its
            // share of the lanes (the shares combined by a cross-row swizzle /
            // shuffle) and one lane per entry writes it for the solve.
            constexpr int CH = LD::CH, NV = NP + NT, NCH = (NV + CH - 1) / CH;
            static_assert((NT == 9 || NT == 10 || NT == 11) && 32 % CH == 0, "J^T J layout");
            constexpr int PARTS = W / CH;
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
#pragma unroll
                for (int i = 0; i < NT; i++) {
#pragma unroll
                    for (int h = 0; h <= i; h++) {
                        const int idx = i * (i + 1) / 2 + h;
                        if (idx / CH != ch) continue;
                        A v = 0;
#pragma unroll
                        for (int k = 0; k < NLPL; k++)
#pragma unroll
                            for (int s = 0; s < 4; s++)
                                v += (A)D[k][s][i] * (A)D[k][s][h];
                        red[t * (CH + 1) + idx % CH] = (double)v;
                    }
                    const int ib = NP + i;
                    if (ib / CH == ch) {
                        A b = 0;
#pragma unroll
                        for (int k = 0; k < NLPL; k++)
#pragma unroll
                            for (int s = 0; s < 4; s++) b += rs[k][s] * (A)D[k][s][i];
                        red[t * (CH + 1) + ib % CH] = (double)b;
                    }
                }
                group_sync<W>();
                const int vv = t % CH, part = t / CH;
                constexpr int SPAN = W / PARTS;
                // pairwise, so the adds form a tree of depth log2(SPAN)
                double v[SPAN];
#pragma unroll
                for (int l = 0; l < SPAN; l++) v[l] = red[(part * SPAN + l) * (CH + 1) + vv];
#pragma unroll
                for (int w = 1; w < SPAN; w *= 2)
#pragma unroll
                    for (int l = 0; l + w < SPAN; l += 2 * w) v[l] += v[l + w];
                double sum = v[0];
                if (PARTS >= 2) sum += swz16_d(sum);
                if (PARTS >= 4) sum += __shfl_xor(sum, 32, W > 64 ? 64 : W);
                if constexpr (W > 64) {
                    // wave 1's totals (parts 4..7) added to wave 0's
                    double *xw = red + (CH + 1) * W;
                    if (part == 4) xw[vv] = sum;
                    group_sync<W>();
                    if (part == 0) sum += xw[vv];
                }
                const int e = ch * CH + vv;
                // a chunk's bits lie in one word (CH divides 32)
                const unsigned zw = c->zbits[NT == 11 ? 1 : NT == 10 ? 2 : 0][(ch * CH) >> 5];
                if ((zw >> (e & 31)) & 1u) sum = 0.0;
                if (part == 0 && e < NV) ab[e] = sum;
                group_sync<W>();
            }
            for (int i = t; i < LD::PAD_ROWS * LD::RS; i += W) pad[i] = T(0);
            group_sync<W>();
            group_sync<W>();
        }

        if (finished) {
            for (int i = t; i < 11; i += W) a.models[(size_t)pix * 11 + i] = lm[i];
            if (t == 0) {
                a.chi2[pix] = (float)ochisqr;                   // :625
                a.iters[pix] = it;
                if (a.accepted) a.accepted[pix] = nacc;
            }
            if (a.fit) {                                        // last trial, :569
#pragma unroll
                for (int k = 0; k < NLPL; k++) {
                    int lam = t + k * W;
                    if (lam < L)
#pragma unroll
                        for (int s = 0; s < 4; s++)
                            a.fit[(size_t)pix * 4 * L + s * L + lam] = (float)S[s][k];
                }
            }
            iterating = false;
        }
    }
}

// ------------------------------------------------------------ synthesis kernel

// RF rows are written in the reference's NTERMS layout (me_der.c); the
// general form maps its 11-row layout back (row 9 of an NTERMS-10 run is
// the filling factor when me_der puts it there, me_der.c:505-520).
template <typename T, int W, int NLPL, int NT, bool GEN>
__global__ void __launch_bounds__(BlockOf<W>::value)
pm_synth_kernel(const PmConst *c0, const double *models, int n, double *spectra, double *rf)
{
    PmC c = (PmC)c0;
    using LD = Lds<T, W, NLPL, NT, GEN>;
    __shared__ __align__(16) char smem[LD::BLOCK_BYTES];
    const int L = c->L;
    const int grp = threadIdx.x / W, t = threadIdx.x % W;
    char *gb = smem + grp * LD::GROUP_BYTES;
    T *pad = (T *)gb;
    T *gm = (T *)(gb + LD::U_BYTES + LD::OBS_BYTES + LD::AB_BYTES + LD::LM_BYTES);
    T *gd = gm + W * NLPL;
    const float *slp = (GEN && c->use_slight) ? (const float *)c->slight : nullptr;
    const int NTR = c->NT;
    const int row9 = (NTR == 10 && slp && !c->fix[9]) ? 10 : 9;
    for (int i = t; i < LD::PAD_ROWS * LD::RS; i += W) pad[i] = T(0);
    group_sync<W>();
    const int ngroups = gridDim.x * LD::GPB;
    for (int pix = blockIdx.x * LD::GPB + grp; pix < n; pix += ngroups) {
        double m[11];
#pragma unroll
        for (int i = 0; i < 11; i++) m[i] = models[(size_t)pix * 11 + i];
        Geom<T> g;
        Samp<T> sp[NLPL];
        T S[4][NLPL];
        Extra<T, NLPL> ex;
        synth_slots<T, W, NLPL, GEN>(c, m, slp, pad, gm, gd, t, L, g, sp, S, ex);
#pragma unroll
        for (int k = 0; k < NLPL; k++) {
            int lam = t + k * W;
            if (lam < L)
#pragma unroll
                for (int s = 0; s < 4; s++)
                    spectra[(size_t)pix * 4 * L + s * L + lam] = (double)S[s][k];
        }
        if (rf) {
            T D[NLPL][4][NT];
            rf_slots<T, W, NLPL, NT, GEN>(c, m, slp, g, sp, pad, gm, gd, t, L, ex, true, S, D);
#pragma unroll
            for (int k = 0; k < NLPL; k++) {
                int lam = t + k * W;
                if (lam < L)
#pragma unroll
                    for (int s = 0; s < 4; s++)
#pragma unroll
                        for (int p = 0; p < NT; p++) {
                            int q = p;                       // row in the reference layout
                            if (GEN && p >= 9) {
                                if (NTR == 9) continue;
                                if (NTR == 10) {
                                    if (p != row9) continue;
                                    q = 9;
                                }
                            }
                            rf[(size_t)pix * 4 * NTR * L + (size_t)s * NTR * L + q * L + lam] =
                                (double)D[k][s][p];
                        }
            }
        }
    }
}

// ------------------------------------------------------------ estimates only

// The classical-estimates mode (NumberOfCycles < 0, milos.c:474-500): the
// guess with B, vlos, gamma, phi from estimacionesClasicas(forInitialUse = 0)
// and B_los in the S0 slot; one thread per pixel.
__global__ void pm_classical_kernel(const PmConst *c0, const float *stokes, int n, double *models)
{
    PmC c = (PmC)c0;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= n) return;
    const Classic e = classical(c, stokes + (size_t)pix * 4 * c->L, false);
    double *m = models + (size_t)pix * 11;
    for (int i = 0; i < 11; i++) m[i] = c->guess[i];
    m[1] = e.B;
    m[2] = e.vlos;
    m[5] = e.gm;
    m[6] = e.az;
    m[7] = e.blos;
}
