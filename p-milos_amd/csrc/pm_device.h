// pm_device.h -- device-side building blocks of the Milne-Eddington
// per-pixel inversion for CDNA4 (gfx950).
//
// Mapping: one pixel is handled by a group of W lanes of a 64-wide
// wavefront (W = 8/16/32/64, a power of two); lane t of the group holds the
// wavelength samples t, t+W, ... (NLPL of them).  Everything that is per
// wavelength (Voigt, Unno-Rachkovsky, response functions) runs across lanes;
// the PSF correlation goes through a zero-padded row buffer in LDS; sums over
// wavelength (chi2, J^T J, gradient) are butterfly reductions inside the
// group, after which every lane holds the same value, so the small LM
// control state (model, damping, chi2) is kept redundantly in every lane of
// the group and the 9x9..11x11 solve runs redundantly in every lane.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define PM_MAXC 16          // Zeeman components per group (pi / blue / red)
#define PM_LMAX 128
#define PM_HC 11            // PSF half-width of the register-blocked correlation
#define PM_PI 3.14159265358979323846264338327950288419716939937510
#define PM_RR 0.5641895836          // defines.h:73 (truncated 1/sqrt(pi))
#define PM_VLIGHT 2.99792458e+5     // defines.h:75
#define PM_CTE4_6_13 4.6686411e-13  // defines.h:77

// Per-context constants resident in device memory (read with scalar loads).
struct PmConst {
    int L, NT;
    int npi, nsig;
    int use_psf;
    int miter, logclambda;
    int fix[11];
    double lambda0, geff;
    double noise, inv_noise, ilambda, toplim, mu;
    // the mixed form's "chi2 unchanged" threshold (the reference's units,
    // |dchi2| * 100 / chi2): the test uses the larger of it and toplim; 0
    // keeps the reference's test
    double toplim_mixed;
    double weight[4];
    // sqrt(weight[s] / sigma^2) of J^T J (lib.c:87-101): [0] with
    // 1/sigma^2 = 1/noise, [1] for a missing sample (1/sigma^2 = 1e-21)
    double swt[2][4];
    // J^T J entries (packed lower alpha, then beta) of fixed parameters, a
    // bit each: [0] for the 9-row layout, [1] for the 11-row one, [2] for
    // the 10-row one (v_mac in row 9)
    unsigned int zbits[3][4];
    double guess[11];
    double nup[PM_MAXC], wep[PM_MAXC];
    double nub[PM_MAXC], web[PM_MAXC], wer[PM_MAXC];
    double lam[PM_LMAX];
    double dlam[PM_LMAX];   // lam - lambda0 in double (mil_sinrf.c:124 subtracts in double)
    double G[PM_LMAX];
    // PSF taps [glo, ghi) whose weight reaches 1e-16 (FP64 form) / 1e-8
    // (FP32 form) of the peak, below each form's unit roundoff: the others
    // are below the rounding of the sum
    int glo64, ghi64, glo32, ghi32;
    // the same taps centred on the kernel centre: gcNN[PM_HC + d] = G[L/2 + d]
    // inside the form's tap range, 0 outside; reachNN = the largest |d| with
    // a tap (PM_HC + 1 when some tap lies beyond PM_HC)
    double gc64[2 * PM_HC + 1];
    float gc32[2 * PM_HC + 1];
    int reach64, reach32;
    // stray-light profile [4][L] mixed in with the filling factor
    // (mil_sinrf.c:405-412), when use_slight
    int use_slight;
    float slight[4 * PM_LMAX];
    // start every pixel from the guess itself, without classical estimates
    // (lm_mils on a .per profile, milos.c:741-775)
    int no_classical;
    // PSF by circular convolution (the reference's FFT mode): Stokes and
    // response rows alike, I without the continuum trick, and folded into
    // the macroturbulence kernel circularly
    int use_fft;
};

// The kernels read the constants through the constant address space: they
// never write them, and reads at uniform addresses become scalar loads.
typedef const __attribute__((address_space(4))) PmConst *PmC;

// ------------------------------------------------------------ group helpers

// Cross-lane moves inside a group.  Within a row of 16 lanes DPP
// quad_perm [1,0,3,2] / [2,3,0,1], row_half_mirror and row_mirror pair every
// lane with a lane of the other half at each level; rows are paired by a
// ds_swizzle xor-16 inside each 32-lane half; the two halves of a wave by
// readlane.  After each level all lanes of the level's span hold the same
// partial sum, so every lane of the group ends with an identical total.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(dpp_i<CTRL>(hi), dpp_i<CTRL>(lo));
}
__device__ __forceinline__ double swz16_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_ds_swizzle(hi, 0x401F),
                            __builtin_amdgcn_ds_swizzle(lo, 0x401F));
}
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_readlane(hi, lane),
                            __builtin_amdgcn_readlane(lo, lane));
}

// Orders LDS traffic between lanes of one wavefront (a group inside one
// wave needs no workgroup barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A group of 128 lanes is the two waves of its workgroup (one pixel per
// workgroup): its LDS traffic is ordered by the workgroup barrier, and its
// sums add the two waves' totals, exchanged through LDS, in a fixed order,
// so that both waves hold the same bits.
template <int W>
__device__ __forceinline__ void group_sync()
{
    if constexpr (W > 64) __syncthreads();
    else wave_sync();
}

// The values of wave 0 and wave 1 of a two-wave workgroup (v uniform within
// each wave).
template <typename V>
__device__ __forceinline__ void wave_pair(V v, V &w0, V &w1)
{
    __shared__ V xs[2];
    if ((threadIdx.x & 63) == 0) xs[threadIdx.x >> 6] = v;
    __syncthreads();
    w0 = xs[0];
    w1 = xs[1];
    __syncthreads();
}

template <int W>
__device__ __forceinline__ double group_sum(double v)
{
    if constexpr (W == 128) {
        double a, b;
        wave_pair(group_sum<64>(v), a, b);
        return a + b;
    }
    static_assert(W == 8 || W == 16 || W == 32 || W == 64 || W == 128, "group width");
    v += dpp_d<0xB1>(v);                 // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);                 // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);                // row_half_mirror
    if (W >= 16) v += dpp_d<0x140>(v);   // row_mirror
    if (W >= 32) v += swz16_d(v);        // rows 0<->1, 2<->3
    if (W >= 64) v = readlane_d(v, 0) + readlane_d(v, 32);
    return v;
}

// 1/x: in FP64 the hardware reciprocal refined by two Newton steps (within
// about an ulp of the division; the arguments here are never 0 or infinite);
// the hardware reciprocal (1 ulp) in FP32
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T> struct AccT { typedef double type; };
template <> struct AccT<float> { typedef float type; };

// Integer sum / min / max and double max over the group, by the same DPP,
// swizzle and readlane levels as group_sum (no LDS round trips): every lane
// ends with the result.  Exact operations, so the order of the levels does
// not change it.
template <int W, typename F>
__device__ __forceinline__ int group_reduce_int(int v, F op)
{
    if constexpr (W == 128) {
        int a, b;
        wave_pair(group_reduce_int<64>(v, op), a, b);
        return op(a, b);
    }
    static_assert(W == 8 || W == 16 || W == 32 || W == 64 || W == 128, "group width");
    v = op(v, dpp_i<0xB1>(v));
    v = op(v, dpp_i<0x4E>(v));
    v = op(v, dpp_i<0x141>(v));
    if (W >= 16) v = op(v, dpp_i<0x140>(v));
    if (W >= 32) v = op(v, __builtin_amdgcn_ds_swizzle(v, 0x401F));
    if (W >= 64) v = op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 32));
    return v;
}

template <int W>
__device__ __forceinline__ int group_sum_int(int v)
{
    return group_reduce_int<W>(v, [](int a, int b) { return a + b; });
}

template <int W>
__device__ __forceinline__ int group_min_int(int v)
{
    return group_reduce_int<W>(v, [](int a, int b) { return a < b ? a : b; });
}

template <int W>
__device__ __forceinline__ int group_max_int(int v)
{
    return group_reduce_int<W>(v, [](int a, int b) { return a > b ? a : b; });
}

template <int W>
__device__ __forceinline__ double group_max(double v)
{
    if constexpr (W == 128) {
        double a, b;
        wave_pair(group_max<64>(v), a, b);
        return fmax(a, b);
    }
    static_assert(W == 8 || W == 16 || W == 32 || W == 64 || W == 128, "group width");
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x141>(v));
    if (W >= 16) v = fmax(v, dpp_d<0x140>(v));
    if (W >= 32) v = fmax(v, swz16_d(v));
    if (W >= 64) v = fmax(readlane_d(v, 0), readlane_d(v, 32));
    return v;
}


// ------------------------------------------------------------ Voigt

// H and F (= Im/2) of the 7/8-term complex rational function P(z)/Q(z),
// z = a - i v, Horner order of fvoigt.c:36-61, coefficients fvoigt.c:20-26.
template <typename T>
__device__ __forceinline__ void voigt(T a, T v, T &H, T &F)
{
    const T A0 = T(122.607931777104326), A1 = T(214.382388694706425),
            A2 = T(181.928533092181549), A3 = T(93.155580458138441),
            A4 = T(30.180142196210589), A5 = T(5.912626209773153),
            A6 = T(0.564189583562615);
    const T B0 = T(122.60793177387535), B1 = T(352.730625110963558),
            B2 = T(457.334478783897737), B3 = T(348.703917719495792),
            B4 = T(170.354001821091472), B5 = T(53.992906912940207),
            B6 = T(10.479857114260399);
    const T zr = a, zi = -v;
    // numerator: a6, then (n*z + a_j)
    T nr = A6, ni = T(0);
#define PM_HORNER(CR, CI, K)                    \
    {                                           \
        T tr = CR * zr - CI * zi + K;           \
        T ti = CR * zi + CI * zr;               \
        CR = tr; CI = ti;                       \
    }
    PM_HORNER(nr, ni, A5) PM_HORNER(nr, ni, A4) PM_HORNER(nr, ni, A3)
    PM_HORNER(nr, ni, A2) PM_HORNER(nr, ni, A1) PM_HORNER(nr, ni, A0)
    // denominator: z + b6, then (d*z + b_j)
    T dr = zr + B6, di = zi;
    PM_HORNER(dr, di, B5) PM_HORNER(dr, di, B4) PM_HORNER(dr, di, B3)
    PM_HORNER(dr, di, B2) PM_HORNER(dr, di, B1) PM_HORNER(dr, di, B0)
#undef PM_HORNER
    // w = n / d: one reciprocal of |d|^2 in FP64; Smith's scaling in FP32,
    // where |d|^2 would overflow for |v| beyond ~2000
    T wr, wi;
    if (sizeof(T) == 8) {
        T inv = recip(dr * dr + di * di);
        wr = (nr * dr + ni * di) * inv;
        wi = (ni * dr - nr * di) * inv;
    } else if (fabs(dr) >= fabs(di)) {
        T r = di * recip(dr), den = dr + di * r;
        T inv = recip(den);
        wr = (nr + ni * r) * inv;
        wi = (ni - nr * r) * inv;
    } else {
        T r = dr * recip(di), den = dr * r + di;
        T inv = recip(den);
        wr = (nr * r + ni) * inv;
        wi = (ni * r - nr) * inv;
    }
    H = wr;
    F = wi * T(0.5);
}

// ------------------------------------------------------------ synthesis

// Angle terms of mil_sinrf.c:84-103 (shared by synthesis and RFs).
template <typename T>
struct Geom {
    T sin_gm, cosi, sinis, cosis, cosis_2, sina, cosa;
    T sinda, cosda, sindi, cosdi, sinis_cosa, sinis_sina;
    T E0, LD, A, MF, B0, B1, ulos, parcial;
    T invE0, invLD;
};

template <typename T>
__device__ __forceinline__ void make_geom(PmC c, const double *m,
                                          Geom<T> &g)
{
    if constexpr (sizeof(T) == 4) {
        // the FP32 form: its spectra are FP32, and so are their angle terms
        // (the reference's default build computes them in float too)
        const float CC = float(PM_PI / 180.0), CC_2 = CC * 2.0f;
        float sgm, cgm, sa, ca;
        sincosf(float(m[5]) * CC, &sgm, &cgm);
        sincosf(2.0f * (float(m[6]) * CC), &sa, &ca);
        const float sinis = sgm * sgm, cosis = cgm * cgm;
        g.sin_gm = sgm; g.cosi = cgm;
        g.sinis = sinis; g.cosis = cosis; g.cosis_2 = (1.0f + cosis) / 2.0f;
        g.sina = sa; g.cosa = ca;
        g.sinda = ca * CC_2; g.cosda = -sa * CC_2;
        g.sindi = cgm * sgm * CC_2; g.cosdi = -sgm * CC;
        g.sinis_cosa = sinis * ca; g.sinis_sina = sinis * sa;
        const float E0 = float(m[0]), LD = float(m[3]);
        g.E0 = E0;
        g.LD = LD; g.A = float(m[4]); g.MF = float(m[1]);
        g.invE0 = 1.0f / E0;
        g.invLD = 1.0f / LD;
        g.B0 = float(m[7]); g.B1 = -(float(m[8]) * float(c->mu));
        g.ulos = (float(m[2]) * float(c->lambda0)) / (float(PM_VLIGHT) * LD);
        g.parcial = (float(c->lambda0 * c->lambda0) / LD) * float(PM_CTE4_6_13);
        return;
    }
    const double CC = PM_PI / 180.0, CC_2 = CC * 2;      // milos.c:232-233
    double GM = m[5] * CC, AZI = m[6] * CC;
    double sgm, cgm, sa, ca;
    sincos(GM, &sgm, &cgm);
    double azi_2 = 2 * AZI;
    sincos(azi_2, &sa, &ca);
    double sinis = sgm * sgm, cosis = cgm * cgm;
    g.sin_gm = T(sgm); g.cosi = T(cgm);
    g.sinis = T(sinis); g.cosis = T(cosis); g.cosis_2 = T((1 + cosis) / 2);
    g.sina = T(sa); g.cosa = T(ca);
    g.sinda = T(ca * CC_2); g.cosda = T(-sa * CC_2);
    g.sindi = T(cgm * sgm * CC_2); g.cosdi = T(-sgm * CC);
    g.sinis_cosa = T(sinis * ca); g.sinis_sina = T(sinis * sa);
    g.E0 = T(m[0]);                                        // FO = 1
    g.LD = T(m[3]); g.A = T(m[4]); g.MF = T(m[1]);
    g.invE0 = T(1.0 / m[0]);
    g.invLD = T(1.0 / m[3]);
    g.B0 = T(m[7]); g.B1 = T(-(m[8] * c->mu));
    g.ulos = T((m[2] * c->lambda0) / (PM_VLIGHT * m[3]));      // mil_sinrf.c:120
    g.parcial = T(((c->lambda0 * c->lambda0) / m[3]) * PM_CTE4_6_13);  // :153
}

// Per-wavelength state of one synthesis kept for the response functions:
// the component sums and the last component of each group (me_der.c keeps
// only the last component's derivative, me_der.c:569-600).
template <typename T>
struct Samp {
    T fi[3], shi[3];          // p, b, r
    T uu[3], H[3], F[3];      // last component of each group
};

// Components and profile sums at one wavelength, mil_sinrf.c:122-183,418-449.
template <typename T>
__device__ __forceinline__ void profiles(PmC c, const Geom<T> &g,
                                         double dl, Samp<T> &s)
{
    // the wavelength offset is formed in double: in FP32 the 0.5 mA spacing
    // of floats near 6000 A would be a percent of a Doppler width
    T u = (T(dl) * g.invLD) - g.ulos;
    // defined whatever the component counts, so that no value of the
    // previous synthesis stays live in a register across the LM loop
#pragma unroll
    for (int q = 0; q < 3; q++) { s.uu[q] = u; s.H[q] = T(0); s.F[q] = T(0); }
    // pi components
    {
        T fi = 0, shi = 0;
        for (int k = 0; k < c->npi; k++) {
            T w = T(c->wep[k]);
            T uu = u - (g.parcial * T(c->nup[k])) * g.MF;
            T H, F;
            voigt(g.A, uu, H, F);
            fi = fi + w * H;
            shi = shi + w * F * T(2);
            s.uu[0] = uu; s.H[0] = H; s.F[0] = F;
        }
        s.fi[0] = fi; s.shi[0] = shi;
    }
    // blue sigma
    {
        T fi = 0, shi = 0;
        for (int k = 0; k < c->nsig; k++) {
            T w = T(c->web[k]);
            T uu = u - (g.parcial * T(c->nub[k])) * g.MF;
            T H, F;
            voigt(g.A, uu, H, F);
            fi = fi + w * H;
            shi = shi + w * F * T(2);
            s.uu[1] = uu; s.H[1] = H; s.F[1] = F;
        }
        s.fi[1] = fi; s.shi[1] = shi;
    }
    // red sigma: shifts mirrored from the blue ones (mil_sinrf.c:164-166)
    {
        T fi = 0, shi = 0;
        for (int k = 0; k < c->nsig; k++) {
            T w = T(c->wer[k]);
            T nur = -(g.parcial * T(c->nub[c->nsig - 1 - k]));
            T uu = u - nur * g.MF;
            T H, F;
            voigt(g.A, uu, H, F);
            fi = fi + w * H;
            shi = shi + w * F * T(2);
            s.uu[2] = uu; s.H[2] = H; s.F[2] = F;
        }
        s.fi[2] = fi; s.shi[2] = shi;
    }
}

// Absorption matrix and Unno-Rachkovsky quantities, mil_sinrf.c:188-266.
template <typename T>
struct UR {
    T etai, etaq, etau, etav, rhoq, rhou, rhov;
    T etai_2, gp1, gp2, gp3, dt, dti, g4r, g5r, g6r;
    T parcial1, parcial2, parcial3;
};

template <typename T>
__device__ __forceinline__ void unno(const Geom<T> &g, const Samp<T> &s, UR<T> &r)
{
    T E0_2 = g.E0 / T(2);
    r.parcial1 = s.fi[1] + s.fi[2];
    r.parcial2 = (E0_2) * (s.fi[0] - (r.parcial1) / T(2));
    r.parcial3 = (E0_2) * (s.shi[0] - (s.shi[1] + s.shi[2]) / T(2));
    T cosi_E0_2 = E0_2 * g.cosi;
    T etain = ((E0_2) * (s.fi[0] * g.sinis + (r.parcial1) * g.cosis_2));
    r.etai = T(1) + etain;
    r.etaq = (r.parcial2 * g.sinis_cosa);
    r.etau = (r.parcial2 * g.sinis_sina);
    r.etav = (s.fi[2] - s.fi[1]) * cosi_E0_2;
    r.rhoq = (r.parcial3 * g.sinis_cosa);
    r.rhou = (r.parcial3 * g.sinis_sina);
    r.rhov = (s.shi[2] - s.shi[1]) * cosi_E0_2;
    T ei = r.etai, eq = r.etaq, eu = r.etau, ev = r.etav;
    T rq = r.rhoq, ru = r.rhou, rv = r.rhov;
    T ei2 = ei * ei;
    T auxq = rq * rq, auxu = ru * ru, auxv = rv * rv;
    r.etai_2 = ei2;
    r.gp1 = ei2 - eq * eq - eu * eu - ev * ev + auxq + auxu + auxv;
    r.gp3 = ei2 + auxq + auxu + auxv;
    r.gp2 = eq * rq + eu * ru + ev * rv;
    r.dt = ei2 * r.gp1 - r.gp2 * r.gp2;
    r.dti = recip(r.dt);
    T gp4 = ei2 * eq + ei * (ev * ru - eu * rv);
    T gp5 = ei2 * eu + ei * (eq * rv - ev * rq);
    T gp6 = (ei2) * ev + ei * (eu * rq - eq * ru);
    r.g4r = gp4 + rq * r.gp2;
    r.g5r = gp5 + ru * r.gp2;
    r.g6r = gp6 + rv * r.gp2;
}

// Emergent Stokes vector before convolution, mil_sinrf.c:268-278.
template <typename T>
__device__ __forceinline__ void stokes(const Geom<T> &g, const UR<T> &r, T S[4])
{
    T dtiaux = r.dti * (g.B1);
    S[0] = g.B0 - dtiaux * r.etai * r.gp3;
    S[1] = (dtiaux * (r.g4r));
    S[2] = (dtiaux * (r.g5r));
    S[3] = (dtiaux * (r.g6r));
}

// Response functions of parameters 0..8 at one wavelength before
// convolution, me_der.c:93-312 (d_spectra rows eta0 B vlos dopp aa gm az S0
// S1).  D[s][p].
template <typename T, int NT>
__device__ __forceinline__ void response(PmC c, const Geom<T> &g,
                                         const Samp<T> &s, const UR<T> &r, T D[4][NT])
{
    const T LD = g.LD, A = g.A, E0 = g.E0;
    // per-group profile derivatives w.r.t. u of the last component
    // (funcionComponentFor, me_der.c:536-609): dfi[q][j] = dH_u[q] * f[q][j],
    // dshi[q][j] = dF_u[q] * f[q][j] for j = B, vlos, dopp (f = -nux, auxCte,
    // -uu/LD); the damping row uses -2 dF_u and dH_u / 2 (me_der.c:581-600)
    T dH_u[3], dF_u[3], fB[3], fD[3];
    const T auxCte = T(-c->lambda0 / PM_VLIGHT) * g.invLD;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        T w, nux;
        if (q == 0) {
            w = T(c->wep[c->npi - 1]);
            nux = g.parcial * T(c->nup[c->npi - 1]);
        } else if (q == 1) {
            w = T(c->web[c->nsig - 1]);
            nux = g.parcial * T(c->nub[c->nsig - 1]);
        } else {
            w = T(c->wer[c->nsig - 1]);
            nux = -(g.parcial * T(c->nub[0]));
        }
        dH_u[q] = ((T(4) * A * s.F[q]) - (T(2) * s.uu[q] * s.H[q])) * w;
        dF_u[q] = (T(PM_RR) - A * s.H[q] - T(2) * s.uu[q] * s.F[q]) * w;
        fB[q] = -nux;
        fD[q] = -s.uu[q] * g.invLD;
    }
    const T E0_2 = E0 / T(2);
    const T cosi_E0_2 = E0_2 * g.cosi;
    const T sc = E0 * g.sinis_cosa / T(2);                       // me_der.c:118-127
    const T ss = E0 * g.sinis_sina / T(2);
    const T sindi_cosa = g.sindi * g.cosa, sindi_sina = g.sindi * g.sina;
    const T cosdi_E0_2 = (E0_2) * g.cosdi;
    const T cosi_2_E0 = E0 * g.cosi / T(2);
    const T cosis_2_E0_2 = g.cosis_2 * E0_2;
    const T sinis_E0_2 = g.sinis * E0_2;
    const T cosi_cosdi = g.cosi * g.cosdi * E0_2;
    const T sindi_E0_2 = g.sindi * E0_2;
    const T sinis_cosda = g.sinis * g.cosda, sinis_sinda = g.sinis * g.sinda;

    // Unno-Rachkovsky terms shared by every parameter (me_der.c:207-220)
    const T ei = r.etai, eq = r.etaq, eu = r.etau, ev = r.etav;
    const T rq = r.rhoq, ru = r.rhou, rv = r.rhov;
    const T ei2 = r.etai_2, gp1 = r.gp1, gp2 = r.gp2, gp3 = r.gp3, dt = r.dt;
    const T dtaux = (g.B1) * (r.dti * r.dti);
    const T etai_gp3 = ei * gp3;
    const T aux2 = T(2) * ei;
    const T ext1 = aux2 * eq + ev * ru - eu * rv;
    const T ext2 = aux2 * eu + eq * rv - ev * rq;
    const T ext3 = aux2 * ev + eu * rq - eq * ru;
    const T ext4 = aux2 * gp1;
#pragma unroll
    for (int il = 0; il < 7; il++) {
        // derivatives of eta_I,Q,U,V and rho_Q,U,V w.r.t. parameter il
        // (me_der.c:93-199)
        T a_ei, a_eq, a_eu, a_ev, a_rq, a_ru, a_rv;
        if (il == 0) {
            a_ei = ((E0_2) * (s.fi[0] * g.sinis + (r.parcial1) * g.cosis_2)) * g.invE0;
            a_eq = (r.parcial2 * g.sinis_cosa) * g.invE0;
            a_eu = (r.parcial2 * g.sinis_sina) * g.invE0;
            a_ev = ((s.fi[2] - s.fi[1]) * cosi_E0_2) * g.invE0;
            a_rq = (r.parcial3 * g.sinis_cosa) * g.invE0;
            a_ru = (r.parcial3 * g.sinis_sina) * g.invE0;
            a_rv = ((s.shi[2] - s.shi[1]) * cosi_E0_2) * g.invE0;
        } else if (il <= 4) {
            T fp, fb, fr, hp, hb, hr;
            if (il == 4) {                                       // damping
                fp = T(-2) * dF_u[0]; fb = T(-2) * dF_u[1]; fr = T(-2) * dF_u[2];
                hp = dH_u[0] / T(2); hb = dH_u[1] / T(2); hr = dH_u[2] / T(2);
            } else {
                T f0, f1, f2;
                if (il == 1) { f0 = fB[0]; f1 = fB[1]; f2 = fB[2]; }
                else if (il == 2) { f0 = auxCte; f1 = auxCte; f2 = auxCte; }
                else { f0 = fD[0]; f1 = fD[1]; f2 = fD[2]; }
                fp = dH_u[0] * f0; fb = dH_u[1] * f1; fr = dH_u[2] * f2;
                hp = dF_u[0] * f0; hb = dF_u[1] * f1; hr = dF_u[2] * f2;
            }
            T dfisum = fb + fr;
            a_ei = (fp * sinis_E0_2 + dfisum * cosis_2_E0_2);
            T aux = fp - dfisum / T(2);
            a_eq = (aux) * sc;
            a_eu = (aux) * ss;
            a_ev = (fr - fb) * cosi_2_E0;
            T auxr = hp - (hb + hr) / T(2);
            a_rq = (auxr) * sc;
            a_ru = (auxr) * ss;
            a_rv = ((hr - hb)) * cosi_2_E0;
        } else if (il == 5) {                                    // gamma
            a_ei = s.fi[0] * sindi_E0_2 + (r.parcial1) * cosi_cosdi;
            a_eq = r.parcial2 * sindi_cosa;
            a_eu = r.parcial2 * sindi_sina;
            a_ev = (s.fi[2] - s.fi[1]) * cosdi_E0_2;
            a_rq = r.parcial3 * sindi_cosa;
            a_ru = r.parcial3 * sindi_sina;
            a_rv = (s.shi[2] - s.shi[1]) * cosdi_E0_2;
        } else {                                                 // azimuth
            a_ei = 0;
            a_eq = r.parcial2 * sinis_cosda;
            a_eu = r.parcial2 * sinis_sinda;
            a_ev = 0;
            a_rq = r.parcial3 * sinis_cosda;
            a_ru = r.parcial3 * sinis_sinda;
            a_rv = 0;
        }
        // chain rule through the Unno-Rachkovsky solution (me_der.c:223-277)
        T dgp1 = T(2) * (ei * a_ei - eq * a_eq - eu * a_eu - ev * a_ev + rq * a_rq +
                         ru * a_ru + rv * a_rv);
        T dgp2 = rq * a_eq + eq * a_rq + ru * a_eu + eu * a_ru + rv * a_ev + ev * a_rv;
        T d_dt = ext4 * a_ei + ei2 * dgp1 - T(2) * gp2 * dgp2;
        T dgp3 = T(2) * (ei * a_ei + rq * a_rq + ru * a_ru + rv * a_rv);
        D[0][il] = -(((a_ei * gp3 + ei * dgp3) * dt - d_dt * etai_gp3) * (dtaux));
        T dgp4 = a_ei * (ext1) + (ei2) * a_eq + ei * (ru * a_ev + ev * a_ru - rv * a_eu - eu * a_rv);
        D[1][il] = ((dgp4 + a_rq * gp2 + rq * dgp2) * dt - d_dt * (r.g4r)) * (dtaux);
        T dgp5 = a_ei * (ext2) + (ei2) * a_eu + ei * (rv * a_eq + eq * a_rv - rq * a_ev - ev * a_rq);
        D[2][il] = ((dgp5 + a_ru * gp2 + ru * dgp2) * dt - d_dt * (r.g5r)) * (dtaux);
        T dgp6 = a_ei * (ext3) + (ei2) * a_ev + ei * (rq * a_eu + eu * a_rq - ru * a_eq - eq * a_ru);
        D[3][il] = ((dgp6 + a_rv * gp2 + rv * dgp2) * dt - d_dt * (r.g6r)) * (dtaux);
    }
    // S1 (me_der.c:281-298), S0 (:301-306), azimuth on I and V (:309-312)
    T dtin = -(r.dti * T(c->mu));
    D[0][8] = -dtin * etai_gp3;
    D[1][8] = dtin * (r.g4r);
    D[2][8] = dtin * (r.g5r);
    D[3][8] = dtin * (r.g6r);
    D[0][7] = T(1); D[1][7] = 0; D[2][7] = 0; D[3][7] = 0;
    D[0][6] = 0; D[3][6] = 0;
#pragma unroll
    for (int p = 9; p < NT; p++) { D[0][p] = 0; D[1][p] = 0; D[2][p] = 0; D[3][p] = 0; }
}

// ------------------------------------------------------------ model update

// AplicaDelta, milosUtils.c:47-115.
__device__ __forceinline__ void aplica_delta(PmC c, int NT,
                                             const double *cur, double *delta, double *out)
{
    const auto *f = c->fix;
    if (f[0]) out[0] = cur[0] - delta[0];
    if (f[1]) {
        if (delta[1] < -300) delta[1] = -300;
        else if (delta[1] > 300) delta[1] = 300;
        out[1] = cur[1] - delta[1];
    }
    if (f[2]) out[2] = cur[2] - delta[2];
    if (f[3]) out[3] = cur[3] - delta[3];
    if (f[4]) out[4] = cur[4] - delta[4];
    if (f[5]) {
        if (delta[5] < -30) delta[5] = -30;
        else if (delta[5] > 30) delta[5] = 30;
        out[5] = cur[5] - delta[5];
    }
    if (f[6]) {
        if (delta[6] < -30) delta[6] = -30;
        else if (delta[6] > 30) delta[6] = 30;
        out[6] = cur[6] - delta[6];
    }
    if (f[7]) out[7] = cur[7] - delta[7];
    if (f[8]) out[8] = cur[8] - delta[8];
    if (NT > 9 && f[9]) out[9] = cur[9] - delta[9];
    if (f[10]) {
        if (NT == 11) out[10] = cur[10] - delta[10];
        else if (NT > 9) out[10] = cur[10] - delta[9];
    }
}

// check, milosUtils.c:118-196.
__device__ __forceinline__ void check_model(double *m)
{
    if (m[1] < 0) { m[1] = -(m[1]); m[5] = 180.0 - (m[5]); }
    if (m[1] > 5000) m[1] = 5000;
    if (m[5] < 0) m[5] = -(m[5]);
    if (m[5] > 180) m[5] = 360.0 - m[5];
    if (m[6] < 0) m[6] = 180 + (m[6]);
    if (m[6] > 180) m[6] = m[6] - 180.0;
    if (m[0] < 1) m[0] = 1;
    if (m[0] > 2500) m[0] = 2500;
    if (m[2] < (-20)) m[2] = (-20);
    if (m[2] > 20) m[2] = 20;
    if (m[3] < 0.0001) m[3] = 0.0001;
    if (m[3] > 0.6) m[3] = 0.6;
    if (m[4] < 0.0001) m[4] = 0.0001;
    if (m[4] > 10.0) m[4] = 10.0;
    if (m[7] < 0.0001) m[7] = 0.0001;
    if (m[7] > 2.00) m[7] = 2.00;
    if (m[8] < 0.0001) m[8] = 0.0001;
    if (m[8] > 2.00) m[8] = 2.00;
    if (m[9] < 0) m[9] = 0;
    if (m[9] > 4) m[9] = 4;
    if (m[10] < 0) m[10] = 0.0;
    if (m[10] > 1.0) m[10] = 1.0;
}

// ------------------------------------------------------------ solve

// delta = A^+ beta for the damped normal matrix (alpha with its diagonal
// scaled by damp = 1 + lambda).  The reference forms the
// eigen pseudo-inverse with an absolute cutoff |w| > 1e-12
// (milosUtils.c:237-271).  For an exactly zero row (a fixed parameter) the
// eigenvalue is 0 and that component is 0; on the remaining block, when an
// LDL^T factorisation succeeds and 1/trace(A^-1) (a lower bound of the
// smallest eigenvalue) exceeds 1e-9, every eigenvalue is far above the cutoff
// and the pseudo-inverse is the inverse.  Otherwise returns false and the
// caller uses the eigen pseudo-inverse itself.
// An inactive row (zero diagonal) gets a unit pivot with its row and column
// cleared, so its component comes out 0; that is done up front (rarely
// needed), and the factorisation itself forms each row's products L_jk d_k
// once for its pivot and the rows below.
// check = false skips the eigenvalue bound: the caller has already passed it
// for the same alpha at a damping no larger (alpha + lambda diag(alpha)
// only grows with lambda, so trace(A^-1) only falls).
template <int N>
__device__ __forceinline__ bool ldl_solve(const double *ab /* LDS: packed lower alpha, then beta */,
                                          double damp, double *x, bool check = true)
{
    constexpr int P = N * (N + 1) / 2;
    double F[P], id[N];
    bool act[N];
    bool all = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int ii = i * (i + 1) / 2;
#pragma unroll
        for (int j = 0; j < i; j++) F[ii + j] = ab[ii + j];
        F[ii + i] = ab[ii + i] * damp;                       // milosUtils.c:559-563
        act[i] = F[ii + i] != 0.0;
        all = all && act[i];
    }
    if (!all) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int ii = i * (i + 1) / 2;
#pragma unroll
            for (int j = 0; j < i; j++) F[ii + j] = (act[i] && act[j]) ? F[ii + j] : 0.0;
            F[ii + i] = act[i] ? F[ii + i] : 1.0;
        }
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        const int jj = j * (j + 1) / 2;
        double w[N > 1 ? N - 1 : 1];
        double dj = F[jj + j];
#pragma unroll
        for (int k = 0; k < j; k++) {
            w[k] = F[jj + k] * F[k * (k + 1) / 2 + k];
            dj -= F[jj + k] * w[k];
        }
        if (!(dj > 0.0)) return false;
        F[jj + j] = dj;
        const double inv_dj = recip(dj);
        id[j] = inv_dj;
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            const int ii = i * (i + 1) / 2;
            double v = F[ii + j];
#pragma unroll
            for (int k = 0; k < j; k++) v -= F[ii + k] * w[k];
            F[ii + j] = v * inv_dj;
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int ii = i * (i + 1) / 2;
        double v = act[i] ? ab[P + i] : 0.0;
#pragma unroll
        for (int k = 0; k < i; k++) v -= F[ii + k] * x[k];
        x[i] = v;
    }
#pragma unroll
    for (int i = 0; i < N; i++) x[i] *= id[i];
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        double v = x[i];
#pragma unroll
        for (int k = i + 1; k < N; k++) v -= F[k * (k + 1) / 2 + i] * x[k];
        x[i] = act[i] ? v : 0.0;
    }
    if (!check) return true;
    // L^-1 in place (column by column), then trace(A^-1) =
    // sum_k (1/d_k) sum_{i<=k} (L^-1)_{ki}^2 over the active rows; 1/trace
    // bounds the smallest eigenvalue from below
#pragma unroll
    for (int j = 0; j < N; j++) {
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            const int ii = i * (i + 1) / 2;
            double v = F[ii + j];
#pragma unroll
            for (int k = j + 1; k < i; k++) v += F[ii + k] * F[k * (k + 1) / 2 + j];
            F[ii + j] = -v;
        }
    }
    double tr = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int kk = k * (k + 1) / 2;
        double s = 1.0;
#pragma unroll
        for (int i = 0; i < k; i++) s += F[kk + i] * F[kk + i];
        tr += act[k] ? s * id[k] : 0.0;
    }
    return tr < 1e9;
}

// v with its value hidden from the optimiser (no instruction is emitted).
// The inversion kernel takes its lane index through this once per trip of its
// persistent loop: otherwise every per-lane address and loop-invariant load
// derived from the lane index is hoisted out of the loop and occupies
// registers for the kernel's whole life (tens of VGPRs, the difference
// between one and two waves per SIMD for the general form).
__device__ __forceinline__ int opaque(int v) { asm volatile("" : "+v"(v)); return v; }

// The lanes of a group (as a bit mask, bit t for lane t) for which pred holds.
template <int W>
__device__ __forceinline__ unsigned long long group_ballot(bool pred)
{
    const unsigned long long b = __ballot(pred);
    if (W == 64) return b;
    const int base = __lane_id() & ~(W - 1);
    return (b >> base) & ((1ull << W) - 1);
}

// Marks v as used here, so that a load feeding it is issued unconditionally
// rather than moved behind the lane predicate that consumes it (a branch
// per load).
__device__ __forceinline__ void pin(double &v) { asm volatile("" : "+v"(v)); }

// LDS doubles the lane-parallel solve uses: one slot per lane of the group
// (pivot, column products, trace terms), 1/d (N), the factor's rows and
// L^-1's rows (N x N each, row-major).
template <int W, int N>
struct LdlWs {
    static constexpr int COL = 0, ID = W, F = W + N, M = F + N * N;
    static constexpr int DOUBLES = M + N * N;
};

// ldl_solve with the factorisation and the trace bound spread over the
// group's lanes: lane i < N holds row i of the damped matrix in registers
// and step k of the factorisation publishes the pivot d_k and the products
// l_jk d_k through LDS, after which every lane updates its own row
// (right-looking).  Each entry of the lower triangle receives the same
// operations in the same order as in ldl_solve (its products subtracted for
// k ascending, then the scaling by 1/d_j), so the factor, delta and the
// bound come out with the same bits.  The upper part of each lane's row
// collects values nothing reads, which keeps the updates free of lane
// predicates; lanes beyond N repeat row N-1.  The substitutions read the
// factor from LDS, every lane running them (each lane needs all of delta);
// L^-1 for the bound is formed a column per lane.  ws:
// LdlWs<W, N>::DOUBLES doubles of LDS that the group may use; they are zero
// again on return.  Needs W >= N.
// Not called by the inversion kernel: the first GPU test run with it in the
// kernel ended in a memory fault whose cause was not found (run on the CPU
// with one thread per lane it is clean under AddressSanitizer and gives
// ldl_solve's bits); tests/test_lanes_cpu.py holds it to those bits.
template <int W, int N>
__device__ __forceinline__ bool ldl_solve_lanes(const double *ab, double damp, double *x,
                                                bool check, double *ws, int t)
{
    static_assert(W >= N && W <= 64, "one lane per row, inside one wave");
    using WS = LdlWs<W, N>;
    constexpr int P = N * (N + 1) / 2;
    const int i = t < N ? t : N - 1;
    const int ii = i * (i + 1) / 2;
    double r[N];
    const double di = ab[ii + i] * damp;                     // milosUtils.c:559-563
    const bool act_i = di != 0.0;
    const unsigned long long actm = group_ballot<W>(t < N && act_i) & ((1ull << N) - 1);
    const bool all = actm == ((1ull << N) - 1);
#pragma unroll
    for (int k = 0; k < N; k++) { r[k] = ab[ii + (k < i ? k : i)]; pin(r[k]); }
    // rows and columns of inactive parameters cleared, unit pivots for them
#pragma unroll
    for (int k = 0; k < N; k++) {
        const bool keep = all || (act_i && ((actm >> k) & 1ull));
        r[k] = k < i ? (keep ? r[k] : 0.0) : (k == i ? ((all || act_i) ? di : 1.0) : 0.0);
    }
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; k++) {
        ws[WS::COL + t] = r[k];
        wave_sync();
        const double dk = ws[WS::COL + k];
        ok = ok && dk > 0.0;
        const double inv = recip(dk);
        ws[WS::ID + k] = inv;
        r[k] = r[k] * inv;                       // l_ik below the pivot
        const double w = r[k] * dk;
        wave_sync();
        ws[WS::COL + t] = w;
        wave_sync();
#pragma unroll
        for (int j = k + 1; j < N; j++) {
            double wj = ws[WS::COL + j];
            pin(wj);
            r[j] -= r[k] * wj;
        }
        wave_sync();
    }
    if (ok) {
#pragma unroll
        for (int k = 0; k < N; k++) ws[WS::F + i * N + k] = r[k];
        wave_sync();
        const double *F = ws + WS::F, *id = ws + WS::ID;
#pragma unroll
        for (int a = 0; a < N; a++) {
            double v = ((actm >> a) & 1ull) ? ab[P + a] : 0.0;
#pragma unroll
            for (int k = 0; k < a; k++) v -= F[a * N + k] * x[k];
            x[a] = v;
        }
#pragma unroll
        for (int a = 0; a < N; a++) x[a] *= id[a];
#pragma unroll
        for (int a = N - 1; a >= 0; a--) {
            double v = x[a];
#pragma unroll
            for (int k = a + 1; k < N; k++) v -= F[k * N + a] * x[k];
            x[a] = ((actm >> a) & 1ull) ? v : 0.0;
        }
        if (check) {
            // column j = t of L^-1: m[a] = -(L_aj + sum_{j<k<a} L_ak m[k])
            // for a > j, k ascending, as ldl_solve forms it in place (the
            // terms k <= j add an exact 0: m[k] = 0 there)
            const int j = i;
            double m[N];
#pragma unroll
            for (int a = 0; a < N; a++) {
                double v = F[a * N + j];
#pragma unroll
                for (int k = 1; k < a; k++) v += F[a * N + k] * m[k];
                m[a] = a > j ? -v : 0.0;
            }
#pragma unroll
            for (int a = 0; a < N; a++) ws[WS::M + a * N + j] = m[a];
            wave_sync();
            // row i: s_i = 1 + sum_{c<i} (L^-1)_ic^2 (zeros beyond the
            // diagonal add nothing), then trace(A^-1) = sum_k s_k / d_k over
            // the active rows, k ascending
            double s = 1.0;
#pragma unroll
            for (int c = 0; c < N - 1; c++) {
                const double q = ws[WS::M + i * N + c];
                s += q * q;
            }
            const double tk = s * id[i];
            wave_sync();
            ws[WS::COL + t] = tk;
            wave_sync();
            double tr = 0.0;
#pragma unroll
            for (int k = 0; k < N; k++) tr += ((actm >> k) & 1ull) ? ws[WS::COL + k] : 0.0;
            ok = tr < 1e9;
        }
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < (WS::DOUBLES + W - 1) / W; q++)
        if (q * W + t < WS::DOUBLES) ws[q * W + t] = 0.0;
    wave_sync();
    return ok;
}

// Eigen pseudo-inverse by cyclic Jacobi rotations on a matrix kept in LDS,
// the lanes of the group sharing the row/column updates.  Same rotation
// formula, sweep order and stopping rule as the CPU restatement.  ws (LDS):
// A = ws[0..N*N) (damped matrix, overwritten), V = ws[N*N..2N*N),
// beta = ws[2N*N..+N), delta written to ws[2N*N+N..+N).  Not inlined: it is
// the rare path and keeps its registers out of the main loop.
template <int W, int N>
__device__ __noinline__ void jacobi_pinv(double *ws, int t)
{
    double *A = ws, *V = ws + N * N;
    const double *beta = ws + 2 * N * N;
    double *delta = ws + 2 * N * N + N;
    for (int i = t; i < N * N; i += W) V[i] = ((i / N) == (i % N)) ? 1.0 : 0.0;
    group_sync<W>();
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < N; i++) {
            diag += A[i * N + i] * A[i * N + i];
            for (int j = i + 1; j < N; j++) off += A[i * N + j] * A[i * N + j];
        }
        if (off == 0.0 || off <= 1e-36 * diag) break;
        for (int p = 0; p < N - 1; p++) {
            for (int q = p + 1; q < N; q++) {
                double apq = A[p * N + q];
                if (apq == 0.0) continue;
                double app = A[p * N + p], aqq = A[q * N + q];
                double theta = (aqq - app) / (2.0 * apq);
                double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
                group_sync<W>();
                for (int k = t; k < N; k += W) {
                    double akp = A[k * N + p], akq = A[k * N + q];
                    A[k * N + p] = cs * akp - sn * akq;
                    A[k * N + q] = sn * akp + cs * akq;
                }
                group_sync<W>();
                for (int k = t; k < N; k += W) {
                    double apk = A[p * N + k], aqk = A[q * N + k];
                    A[p * N + k] = cs * apk - sn * aqk;
                    A[q * N + k] = sn * apk + cs * aqk;
                }
                group_sync<W>();
                if (t == 0) { A[p * N + q] = 0.0; A[q * N + p] = 0.0; }
                for (int k = t; k < N; k += W) {
                    double vkp = V[k * N + p], vkq = V[k * N + q];
                    V[k * N + p] = cs * vkp - sn * vkq;
                    V[k * N + q] = sn * vkp + cs * vkq;
                }
                group_sync<W>();
            }
        }
    }
    // delta_i = sum_j V_ij (V^T beta)_j / w_j, |w_j| > 1e-12 (milosUtils.c:254-269)
    for (int i = t; i < N; i += W) {
        double s = 0;
        for (int j = 0; j < N; j++) {
            double a = 0;
            for (int k = 0; k < N; k++) a += beta[k] * V[k * N + j];
            double w = A[j * N + j];
            s += V[i * N + j] * (a * ((fabs(w) > 1e-12) ? (1 / w) : 0.0));
        }
        delta[i] = s;
    }
    group_sync<W>();
}

// ------------------------------------------------------------ classical estimates

struct Classic {
    double B, vlos, gm, az, blos;
};

// estimacionesClasicas, milosUtils.c:287-452, on one observed profile (in
// LDS when each lane of a pixel group runs it redundantly, in global memory
// for the estimates-only pass), in the reference's summation order and
// operand types, without contracting products into sums (the reference's
// own build has no FMA, makefile:8).  I + V and I - V are float sums
// (milosUtils.c:337,357): formed in double they differ by up to half a
// float ulp per sample, which the LM's first, often ill-conditioned step
// turns into visibly different paths (2% of the pixels of a 112-sample
// cube with v_mac free end outside the FP64 tolerance).  With
// init (forInitialUse = 1, the inversion's starting point) the 85-95 degree
// inclination nudge applies to B and the pixel loop's NaN guards follow
// (milos.c:976-983); without it (the estimates-only mode, milos.c:493) B_los
// is returned for the S0 slot (:449-450) and nothing is guarded.  Not
// inlined: it runs once per pixel.
__device__ __noinline__ Classic classical(PmC c, const float *spec,
                                         bool init)
{
#pragma clang fp contract(off)
    const int n = c->L;
    const double l0 = c->lambda0;
    const float *sI = spec, *sQ = spec + n, *sU = spec + 2 * n, *sV = spec + 3 * n;
    int begin = 0;
    for (int i = 0; i < n; i++) { if (sI[i] < 0) begin++; else break; }
    int end = n;
    for (int i = n - 1; i >= 0; i--) { if (sI[i] < 0) end--; else break; }
    double Ic = (end >= 1) ? (double)spec[end - 1] : 0.0;
    double x = 0, y = 0, xv = 0, yv = 0;
    // I + V and I - V are float sums, as the reference forms them from its
    // float spectro (milosUtils.c:337,357)
    for (int i = begin; i < end - 1; i++) {
        if (sI[i] > -1 && sV[i] > -1) {
            double aux = (Ic - (double)__fadd_rn(sI[i], sV[i]));
            double auxv = (Ic - (double)sI[i]);
            double dl = c->lam[i] - l0;
            x += (aux * dl);
            xv += (auxv * dl);
            y += aux;
            yv += auxv;
        }
    }
    double lplus = (fabs(y) > 1e-15) ? x / y : 0;
    x = 0; y = 0;
    for (int i = begin; i < end - 1; i++) {
        if (sI[i] > -1 && sV[i] > -1) {
            double aux = (Ic - (double)__fsub_rn(sI[i], sV[i]));
            x += (aux * (c->lam[i] - l0));
            y += aux;
        }
    }
    double lminus = (fabs(y) > 1e-15) ? x / y : 0;
    double C = (PM_CTE4_6_13 * (l0 * l0) * c->geff);
    double Blos = (1 / C) * ((lplus - lminus) / 2);
    double Vlos = (PM_VLIGHT / (l0)) * ((xv / yv) / 2);
    x = 0; y = 0;
    for (int i = begin; i < end - 1; i++) {
        if (sQ[i] > -1 && sU[i] > -1 && sV[i] > -1) {
            // float arithmetic as in the reference (float * float)
            float qu = __fadd_rn(__fmul_rn(sQ[i], sQ[i]), __fmul_rn(sU[i], sU[i]));
            double Lp = fabs(sqrt((double)qu));
            double mm = fabs((4 * (c->lam[i] - l0) * Lp));
            double av = fabs((double)sV[i]);
            x = x + av * mm;
            y = y + av * av;
        }
    }
    y = y * fabs((3 * C * Blos));
    double tan_gamma = fabs(sqrt(x / y));
    double gamma_rad = atan(tan_gamma);
    double gamma = gamma_rad * (180 / PM_PI);
    if (init) {
        if (gamma >= 85 && gamma <= 90) gamma_rad = 85 * (PM_PI / 180);
        if (gamma > 90 && gamma <= 95) gamma_rad = 95 * (PM_PI / 180);
    }
    if (Blos < 0) gamma = (180) - gamma;
    double su = 0.0, sq = 0.0;
    for (int i = 0; i < n; i++) {
        if (sU[i] > -1 && sQ[i] > -1) {
            if (fabs((double)sU[i]) > 0.0001 || fabs((double)sQ[i]) > 0.0001) {
                su += sU[i];
                sq += sQ[i];
            }
        }
    }
    double tan2phi = su / sq;
    double phi = (atan(tan2phi) * 180 / PM_PI) / 2;
    if (su > 0 && sq > 0) phi = phi;
    else if (su < 0 && sq > 0) phi = phi + 180;
    else if (su < 0 && sq < 0) phi = phi + 90;
    else if (su > 0 && sq < 0) phi = phi + 90;
    double B_aux = fabs(Blos / cos(gamma_rad));
    if (Vlos < (-4)) Vlos = -4;
    if (Vlos > (4)) Vlos = 4;
    Classic r;
    r.B = (B_aux > 4000 ? 4000 : B_aux);
    r.vlos = Vlos;
    r.gm = gamma;
    r.az = phi;
    r.blos = Blos;
    if (init) {
        // NaN guards of the pixel loop, milos.c:976-983
        if (isnan(r.B)) r.B = 1;
        if (isnan(r.vlos)) r.vlos = 1e-3;
        if (isnan(r.gm)) r.gm = 1;
        if (isnan(r.az)) r.az = 1;
    }
    return r;
}

// The same estimates computed by a pixel group: each lane takes its own
// samples (lambda = t + k*W) and the sums over wavelength are group sums
// (tree order instead of the serial order: the sums' rounding differs from
// the serial ones by a few units in the last place, which parts no more LM
// paths than rebuilding the reference with other flags does), the range of
// the leading and trailing negative I a group min/max.  I + V and I - V are
// float sums as in the reference.  Every lane returns the same result.
template <int W, int NLPL>
__device__ __noinline__ Classic classical_group(PmC c, const float *spec, bool init, int t)
{
    const int n = c->L;
    const double l0 = c->lambda0;
    const float *sI = spec, *sQ = spec + n, *sU = spec + 2 * n, *sV = spec + 3 * n;
    // milosUtils.c:301-325: begin = first sample with I not < 0, end = one
    // past the last
    int lo = n, hi = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        const int i = t + k * W;
        if (i < n && !(sI[i] < 0)) { lo = min(lo, i); hi = max(hi, i + 1); }
    }
    const int begin = group_min_int<W>(lo), end = group_max_int<W>(hi);
    const double Ic = (end >= 1) ? (double)spec[end - 1] : 0.0;
    double x = 0, y = 0, xv = 0, yv = 0, xm = 0, ym = 0, xg = 0, yg = 0, su = 0, sq = 0;
#pragma unroll
    for (int k = 0; k < NLPL; k++) {
        const int i = t + k * W;
        if (i >= n) continue;
        const double dl = c->dlam[i];
        if (i >= begin && i < end - 1) {
            if (sI[i] > -1 && sV[i] > -1) {                       // :330-371
                double aux = (Ic - (double)__fadd_rn(sI[i], sV[i]));
                double auxv = (Ic - (double)sI[i]);
                x += (aux * dl);
                xv += (auxv * dl);
                y += aux;
                yv += auxv;
                double auxm = (Ic - (double)__fsub_rn(sI[i], sV[i]));
                xm += (auxm * dl);
                ym += auxm;
            }
            if (sQ[i] > -1 && sU[i] > -1 && sV[i] > -1) {         // :375-393
                float qu = __fadd_rn(__fmul_rn(sQ[i], sQ[i]), __fmul_rn(sU[i], sU[i]));
                double Lp = fabs(sqrt((double)qu));
                double mm = fabs((4 * dl * Lp));
                double av = fabs((double)sV[i]);
                xg = xg + av * mm;
                yg = yg + av * av;
            }
        }
        if (sU[i] > -1 && sQ[i] > -1) {                           // :409-431
            if (fabs((double)sU[i]) > 0.0001 || fabs((double)sQ[i]) > 0.0001) {
                su += sU[i];
                sq += sQ[i];
            }
        }
    }
    x = group_sum<W>(x); y = group_sum<W>(y); xv = group_sum<W>(xv); yv = group_sum<W>(yv);
    xm = group_sum<W>(xm); ym = group_sum<W>(ym); xg = group_sum<W>(xg); yg = group_sum<W>(yg);
    su = group_sum<W>(su); sq = group_sum<W>(sq);
    const double lplus = (fabs(y) > 1e-15) ? x / y : 0;
    const double lminus = (fabs(ym) > 1e-15) ? xm / ym : 0;
    const double C = (PM_CTE4_6_13 * (l0 * l0) * c->geff);
    const double Blos = (1 / C) * ((lplus - lminus) / 2);
    double Vlos = (PM_VLIGHT / (l0)) * ((xv / yv) / 2);
    yg = yg * fabs((3 * C * Blos));
    const double tan_gamma = fabs(sqrt(xg / yg));
    double gamma_rad = atan(tan_gamma);
    double gamma = gamma_rad * (180 / PM_PI);
    if (init) {
        if (gamma >= 85 && gamma <= 90) gamma_rad = 85 * (PM_PI / 180);
        if (gamma > 90 && gamma <= 95) gamma_rad = 95 * (PM_PI / 180);
    }
    if (Blos < 0) gamma = (180) - gamma;
    const double tan2phi = su / sq;
    double phi = (atan(tan2phi) * 180 / PM_PI) / 2;
    if (su > 0 && sq > 0) phi = phi;
    else if (su < 0 && sq > 0) phi = phi + 180;
    else if (su < 0 && sq < 0) phi = phi + 90;
    else if (su > 0 && sq < 0) phi = phi + 90;
    const double B_aux = fabs(Blos / cos(gamma_rad));
    if (Vlos < (-4)) Vlos = -4;
    if (Vlos > (4)) Vlos = 4;
    Classic r;
    r.B = (B_aux > 4000 ? 4000 : B_aux);
    r.vlos = Vlos;
    r.gm = gamma;
    r.az = phi;
    r.blos = Blos;
    if (init) {
        if (isnan(r.B)) r.B = 1;
        if (isnan(r.vlos)) r.vlos = 1e-3;
        if (isnan(r.gm)) r.gm = 1;
        if (isnan(r.az)) r.az = 1;
    }
    return r;
}
