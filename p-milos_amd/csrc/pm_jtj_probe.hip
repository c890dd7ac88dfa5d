// J^T J assembly on the matrix cores against the vector form of the
// inversion kernel, in isolation.
//
// covarm2 (lib.c:71-116) forms, per pixel, the packed lower triangle of
// alpha = J^T J (NT = 9 rows of 4L response samples) and beta = J^T r.  The
// inversion kernel keeps each pixel on a group of 32 lanes (one wavelength
// per lane, four Stokes samples per lane) and reduces the 54 per-lane
// partial sums through LDS.  A matrix-core form instead stages the 10 rows
// (9 response rows and the weighted residual) in LDS as a 10 x 4L matrix M,
// padded to 16 rows with zeros, and accumulates M M^T with 4L / 4 = 30
// v_mfma_*_16x16x4 steps per pixel: the 16 x 16 result holds alpha in its
// lower triangle and beta in column 9.
//
// pmp_jtj_run times one of three kernels over the same generated rows:
// mode 0 generates the rows only, mode 1 adds the vector reduction, mode 2
// the matrix-core product.  Each group repeats `reps` pixels; the alpha/beta
// of its last pixel are written out so that the two forms can be compared.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int W = 32;          // lanes per pixel group
constexpr int L = 30;          // wavelengths (one per lane)
constexpr int NT = 9;
constexpr int NP = NT * (NT + 1) / 2, NV = NP + NT;   // 45 + 9
constexpr int BLK = 128;       // 4 groups, 2 waves
constexpr int GPB = BLK / W;
constexpr int CH = 16, NCH = (NV + CH - 1) / CH, SPAN = W / (W / CH);
constexpr int KS = 4 * L;      // staged samples per row

__device__ __forceinline__ float hval(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)(x & 0xffffffu) * (1.0f / 16777216.0f) - 0.5f;
}

// row r (0..8 response, 9 residual) of pixel pix, Stokes s, lane t; zero
// beyond L like the inversion kernel's weighted rows
template <typename T>
__device__ __forceinline__ T gen(uint32_t pix, int r, int s, int t)
{
    if (t >= L) return T(0);
    return T(hval(pix * 2654435761u + (uint32_t)(r * 131 + s * 37 + t) * 97u + 12345u));
}

__device__ __forceinline__ double swz16_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_ds_swizzle(hi, 0x401F),
                            __builtin_amdgcn_ds_swizzle(lo, 0x401F));
}

// the vector reduction of the inversion kernel (same chunking, tree and
// swizzle), products in T, sums in double
template <typename T>
__device__ __forceinline__ void jtj_vector(const T D[4][NT + 1], double *red, double *ab, int t)
{
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
#pragma unroll
        for (int i = 0; i < NT; i++) {
#pragma unroll
            for (int h = 0; h <= i; h++) {
                const int idx = i * (i + 1) / 2 + h;
                if (idx / CH != ch) continue;
                T v = 0;
#pragma unroll
                for (int s = 0; s < 4; s++) v += D[s][i] * D[s][h];
                red[t * (CH + 1) + idx % CH] = (double)v;
            }
            const int ib = NP + i;
            if (ib / CH == ch) {
                T b = 0;
#pragma unroll
                for (int s = 0; s < 4; s++) b += D[s][NT] * D[s][i];
                red[t * (CH + 1) + ib % CH] = (double)b;
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int vv = t % CH, part = t / CH;
        double v[SPAN];
#pragma unroll
        for (int l = 0; l < SPAN; l++) v[l] = red[(part * SPAN + l) * (CH + 1) + vv];
#pragma unroll
        for (int w = 1; w < SPAN; w *= 2)
#pragma unroll
            for (int l = 0; l + w < SPAN; l += 2 * w) v[l] += v[l + w];
        double sum = v[0] + swz16_d(v[0]);
        const int e = ch * CH + vv;
        if (part == 0 && e < NV) ab[e] = sum;
        __builtin_amdgcn_wave_barrier();
    }
}

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    typedef float __attribute__((ext_vector_type(4))) acc;
    static __device__ __forceinline__ acc step(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    typedef double __attribute__((ext_vector_type(4))) acc;
    static __device__ __forceinline__ acc step(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// lane l of the wave supplies M[l % 16][4 kk + l / 16] (0 for the padding
// rows) as both operands of step kk (A = M, B = M^T).  Result element (i, l %
// 16) lands in lane l's v-th accumulator with i = 4 (l / 16) + v for the
// f32 instruction and i = 4 v + l / 16 for the f64 one.
template <typename T>
__device__ __forceinline__ void jtj_mfma(const T *Mg, double *ab, int lane)
{
    typename Mfma<T>::acc c = {0, 0, 0, 0};
    const int r = lane % 16, k0 = lane / 16;
#pragma unroll
    for (int kk = 0; kk < KS / 4; kk++) {
        const T a = r <= NT ? Mg[r * KS + 4 * kk + k0] : T(0);
        c = Mfma<T>::step(a, a, c);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int i = sizeof(T) == 4 ? 4 * (lane / 16) + v : 4 * v + lane / 16, j = lane % 16;
        if (i < NT && j <= i) ab[i * (i + 1) / 2 + j] = (double)c[v];
        if (i < NT && j == NT) ab[NP + i] = (double)c[v];
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(BLK, 2) jtj_kernel(int reps, double *out, double *sink)
{
    __shared__ __align__(16) double red_all[MODE == 1 ? GPB : 1][W * (CH + 1)];
    __shared__ __align__(16) double ab_all[GPB][64];
    __shared__ __align__(16) T M_all[MODE == 2 ? GPB : 1][(NT + 1) * KS];
    const int grp = threadIdx.x / W, t = threadIdx.x % W;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    double *red = red_all[MODE == 1 ? grp : 0], *ab = ab_all[grp];
    double acc = 0;
    const uint32_t base = (blockIdx.x * GPB + grp) * (uint32_t)reps;
    for (int rep = 0; rep < reps; rep++) {
        const uint32_t pix = base + rep;
        T D[4][NT + 1];
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int r = 0; r <= NT; r++) D[s][r] = gen<T>(pix, r, s, t);
        if (MODE == 0) {
            T v = 0;
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int r = 0; r <= NT; r++) v += D[s][r];
            acc += (double)v;
        } else if (MODE == 1) {
            jtj_vector<T>(D, red, ab, t);
            acc += ab[t % NV];
            __builtin_amdgcn_wave_barrier();
        } else {
            T *Mg = M_all[grp];
            if (t < L)
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int r = 0; r <= NT; r++) Mg[r * KS + s * L + t] = D[s][r];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the wave's two groups, one product each, all 64 lanes
#pragma unroll
            for (int g2 = 0; g2 < 2; g2++) {
                const int g = 2 * wave + g2;
                jtj_mfma<T>(M_all[g], ab_all[g], lane);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            acc += ab[t % NV];
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (MODE != 0)
        for (int e = t; e < NV; e += W) out[(size_t)(blockIdx.x * GPB + grp) * NV + e] = ab[e];
    sink[blockIdx.x * BLK + threadIdx.x] = acc;
}

template <typename T>
int launch(int mode, int blocks, int reps, double *out, double *sink, hipStream_t st)
{
    if (mode == 0) hipLaunchKernelGGL((jtj_kernel<T, 0>), dim3(blocks), dim3(BLK), 0, st, reps, out, sink);
    else if (mode == 1) hipLaunchKernelGGL((jtj_kernel<T, 1>), dim3(blocks), dim3(BLK), 0, st, reps, out, sink);
    else hipLaunchKernelGGL((jtj_kernel<T, 2>), dim3(blocks), dim3(BLK), 0, st, reps, out, sink);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

// fp64: 1 for double rows, 0 for float; mode 0/1/2 as above.  out:
// [blocks * 4][54] doubles (alpha packed lower, then beta) of each group's
// last pixel; sink: [blocks * 128] doubles.  *ms = kernel time.
extern "C" int pmp_jtj_run(int fp64, int mode, int blocks, int reps, double *d_out,
                           double *d_sink, float *ms)
{
    if (blocks < 1 || reps < 1 || mode < 0 || mode > 2 || !d_out || !d_sink) return -1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -2;
    hipEventRecord(e0, 0);
    int rc = fp64 ? launch<double>(mode, blocks, reps, d_out, d_sink, 0)
                  : launch<float>(mode, blocks, reps, d_out, d_sink, 0);
    hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) rc = -2;
    if (ms) hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

// the generated rows of group grp's last pixel, [10][4][32] (for the check)
extern "C" void pmp_jtj_rows(int grp, int reps, double *rows)
{
    const uint32_t pix = (uint32_t)grp * (uint32_t)reps + (uint32_t)(reps - 1);
    for (int r = 0; r <= NT; r++)
        for (int s = 0; s < 4; s++)
            for (int t = 0; t < W; t++) {
                float v = 0.0f;
                if (t < L) {
                    uint32_t x = pix * 2654435761u + (uint32_t)(r * 131 + s * 37 + t) * 97u + 12345u;
                    x ^= x >> 16; x *= 0x7feb352du;
                    x ^= x >> 15; x *= 0x846ca68bu;
                    x ^= x >> 16;
                    v = (float)(x & 0xffffffu) * (1.0f / 16777216.0f) - 0.5f;
                }
                rows[(r * 4 + s) * W + t] = v;
            }
}
