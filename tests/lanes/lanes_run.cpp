// lanes_run -- runs the library's inversion kernel (pm_invert.h) or its
// group solve on the host with one thread per lane (hip_lanes.h), for the
// CPU tests in tests/test_lanes_cpu.py, built with AddressSanitizer and
// UBSan by tests/lanes/Makefile.  TEST INFRASTRUCTURE.
//
//   lanes_run IN OUT
//
// IN (little-endian): int32 hdr[16], then the mode's arrays (see read_*).
//   mode 1, inversion: hdr = {1, W, NLPL, BLK, kind (0 the 9-row kernel, 1 the
//     11-row general one, 2 the 10-row general one), precision, npix, L,
//     miter, logclambda, use_psf, use_fft, no_classical, slight profile,
//     slight cube, want_fit}; int32 fix[11]; float64 {sl, ll, jl, su, lu, ju,
//     lambda0, noise, ilambda, toplim, mu, weight[4], 0}; float64 guess[11];
//     float64 lambda[L]; float64 psf[L] if use_psf; float32 profile[4L] if
//     slight profile; float32 stokes[npix][4][L]; float32 slight[npix][4][L]
//     if slight cube.
//   OUT: float64 models[npix][11], float32 chi2[npix], int32 iters[npix],
//     int32 accepted[npix], float32 fit[npix][4][L] if want_fit, uint64
//     fallbacks.
//   mode 2, the solve: hdr = {2, W, N, 0, 0, 0, K}; float64 ab[K][N(N+1)/2 + N],
//     damp[K], check[K] (0/1).  OUT per matrix: float64 x[N], ok of the
//     group solve (ldl_solve_lanes, lane 0's copy, then every lane's copy
//     compared), then x[N], ok of ldl_solve; as float64.
#include "hip/hip_runtime.h"

#include <thread>
#include <vector>

#include "pm_invert.h"
#include "pm_setup.h"

namespace {

FILE *fin;

template <typename V>
std::vector<V> rd(size_t n)
{
    std::vector<V> v(n);
    if (n && fread(v.data(), sizeof(V), n, fin) != n) {
        fprintf(stderr, "lanes_run: short input\n");
        exit(2);
    }
    return v;
}

// one workgroup of BLK threads, groups of W lanes, running fn() each
template <int W, int BLK, typename F>
void launch(F fn)
{
    constexpr int G = BLK / W;
    std::vector<std::barrier<> *> bars;
    LanesGroup groups[G];
    for (int g = 0; g < G; g++) {
        bars.push_back(new std::barrier<>(W));
        groups[g].W = W;
        groups[g].bar = bars.back();
    }
    blockDim.x = BLK;
    gridDim.x = 1;
    std::vector<std::thread> th;
    for (int tid = 0; tid < BLK; tid++)
        th.emplace_back([&, tid] {
            threadIdx.x = tid;
            blockIdx.x = 0;
            lanes_grp = &groups[tid / W];
            fn();
        });
    for (auto &t : th) t.join();
    for (auto *b : bars) delete b;
}

template <typename T, int W, int NLPL, int NT, bool GEN, int BLK>
unsigned long long invert(const PmConst *c, const float *stokes, const float *slight, int npix,
                          double *models, float *chi2, int *iters, int *acc, float *fit)
{
    std::vector<int> counters(PM_NSLICE * PM_CSTRIDE, 0);
    unsigned long long fallbacks = 0;
    InvArgs a;
    a.c = c;
    a.stokes = stokes;
    a.npix = npix;
    a.slice = (npix + PM_NSLICE - 1) / PM_NSLICE;
    a.models = models;
    a.chi2 = chi2;
    a.iters = iters;
    a.accepted = acc;
    a.fit = fit;
    a.counters = counters.data();
    a.fallbacks = &fallbacks;
    a.slight = slight;
    launch<W, BLK>([&] { pm_invert_kernel<T, W, NLPL, NT, GEN, BLK>(a); });
    return fallbacks;
}

typedef unsigned long long (*InvFn)(const PmConst *, const float *, const float *, int, double *,
                                    float *, int *, int *, float *);
struct Shape {
    int prec, W, NLPL, BLK, gen;
    InvFn fn;
};
// gen: 0 the 9-row kernel, 1 the 11-row general one, 2 the 10-row general one
#define PM_LS(T, P, W, NLPL, NT, G, BLK, KIND) \
    Shape{P, W, NLPL, BLK, KIND, &invert<T, W, NLPL, NT, G, BLK>}
// the shapes the CPU tests run
const Shape SHAPES[] = {
    PM_LS(double, PM_FP64, 8, 1, 9, false, 128, 0),
    PM_LS(double, PM_FP64, 8, 2, 9, false, 128, 0),
    PM_LS(double, PM_FP64, 16, 2, 9, false, 128, 0),
    PM_LS(double, PM_FP64, 16, 1, 9, false, 128, 0),
    PM_LS(double, PM_FP64, 32, 1, 9, false, 128, 0),
    PM_LS(double, PM_FP64, 64, 2, 11, true, 128, 1),
    PM_LS(double, PM_FP64, 64, 2, 10, true, 128, 2),
    PM_LS(double, PM_FP64, 128, 1, 10, true, 128, 2),
    PM_LS(double, PM_FP64, 128, 1, 11, true, 128, 1),
    PM_LS(double, PM_FP64, 128, 1, 9, false, 128, 0),
    PM_LS(float, PM_MIXED, 16, 2, 9, false, 128, 0),
    PM_LS(float, PM_MIXED, 64, 2, 11, true, 128, 1),
    PM_LS(float, PM_MIXED, 64, 2, 10, true, 128, 2),
};

int mode_invert(const int *h, FILE *fout)
{
    const int W = h[1], NLPL = h[2], BLK = h[3], gen = h[4], prec = h[5], npix = h[6], L = h[7];
    pm_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    auto fix = rd<int>(11);
    auto par = rd<double>(16);
    auto guess = rd<double>(11);
    auto lam = rd<double>(L);
    std::vector<double> psf = h[10] ? rd<double>(L) : std::vector<double>();
    std::vector<float> prof = h[13] ? rd<float>(4 * L) : std::vector<float>();
    auto stokes = rd<float>((size_t)npix * 4 * L);
    std::vector<float> slc = h[14] ? rd<float>((size_t)npix * 4 * L) : std::vector<float>();
    cfg.nlambda = L;
    cfg.lambda = lam.data();
    cfg.sl = par[0]; cfg.ll = par[1]; cfg.jl = par[2];
    cfg.su = par[3]; cfg.lu = par[4]; cfg.ju = par[5];
    cfg.lambda0 = par[6];
    cfg.noise = par[7];
    cfg.ilambda = par[8];
    cfg.toplim = par[9];
    cfg.mu = par[10];
    for (int i = 0; i < 4; i++) cfg.weight[i] = par[11 + i];
    cfg.psf = h[10] ? psf.data() : nullptr;
    for (int i = 0; i < 11; i++) { cfg.fix[i] = fix[i]; cfg.guess[i] = guess[i]; }
    cfg.miter = h[8];
    cfg.logclambda = h[9];
    cfg.precision = prec;
    cfg.slight = h[13] ? prof.data() : nullptr;
    cfg.no_classical = h[12];
    cfg.use_fft = h[11];
    PmConst *c = new PmConst;
    int rc = pm_const_build(&cfg, *c);
    if (rc != PM_OK) {
        fprintf(stderr, "lanes_run: set-up failed (%d)\n", rc);
        return 3;
    }
    const Shape *s = nullptr;
    for (const Shape &x : SHAPES)
        if (x.prec == prec && x.W == W && x.NLPL == NLPL && x.BLK == BLK && x.gen == gen) s = &x;
    if (!s || W * NLPL < L) {
        fprintf(stderr, "lanes_run: shape %d x %d x %d gen %d not built here\n", W, NLPL, BLK, gen);
        return 4;
    }
    std::vector<double> models((size_t)npix * 11);
    std::vector<float> chi2(npix), fit(h[15] ? (size_t)npix * 4 * L : 0);
    std::vector<int> iters(npix), acc(npix);
    const unsigned long long fb =
        s->fn(c, stokes.data(), h[14] ? slc.data() : nullptr, npix, models.data(), chi2.data(),
              iters.data(), acc.data(), h[15] ? fit.data() : nullptr);
    fwrite(models.data(), 8, models.size(), fout);
    fwrite(chi2.data(), 4, chi2.size(), fout);
    fwrite(iters.data(), 4, iters.size(), fout);
    fwrite(acc.data(), 4, acc.size(), fout);
    fwrite(fit.data(), 4, fit.size(), fout);
    fwrite(&fb, 8, 1, fout);
    delete c;
    return 0;
}

template <int W, int N>
void solve_one(const double *ab_in, double damp, bool check, double *out)
{
    constexpr int P = N * (N + 1) / 2;
    static __attribute__((aligned(16))) double lds[P + N + LdlWs<W, N>::DOUBLES];
    double *ab = lds, *ws = lds + P + N;
    memcpy(ab, ab_in, sizeof(double) * (P + N));
    memset(ws, 0, sizeof(double) * LdlWs<W, N>::DOUBLES);
    double xs[W][N];
    bool oks[W];
    launch<W, W>([&] {
        const int t = lanes_in_group();
        double x[N];
        oks[t] = ldl_solve_lanes<W, N>(ab, damp, x, check, ws, t);
        for (int i = 0; i < N; i++) xs[t][i] = x[i];
    });
    // every lane holds the same delta and verdict; the workspace is zero again
    bool same = true;
    for (int t = 1; t < W; t++) {
        same = same && oks[t] == oks[0];
        if (oks[0])
            for (int i = 0; i < N; i++) same = same && memcmp(&xs[t][i], &xs[0][i], 8) == 0;
    }
    for (int e = 0; e < LdlWs<W, N>::DOUBLES; e++) same = same && ws[e] == 0.0;
    double xr[N];
    const bool okr = ldl_solve<N>(ab, damp, xr, check);
    for (int i = 0; i < N; i++) out[i] = xs[0][i];
    out[N] = oks[0] ? 1.0 : 0.0;
    for (int i = 0; i < N; i++) out[N + 1 + i] = xr[i];
    out[2 * N + 1] = okr ? 1.0 : 0.0;
    out[2 * N + 2] = same ? 1.0 : 0.0;
}

template <int W, int N>
int solve_all(int K, FILE *fout)
{
    constexpr int P = N * (N + 1) / 2;
    auto ab = rd<double>((size_t)K * (P + N));
    auto damp = rd<double>(K);
    auto chk = rd<double>(K);
    std::vector<double> out(2 * N + 3);
    for (int k = 0; k < K; k++) {
        solve_one<W, N>(ab.data() + (size_t)k * (P + N), damp[k], chk[k] != 0.0, out.data());
        fwrite(out.data(), 8, out.size(), fout);
    }
    return 0;
}

int mode_solve(const int *h, FILE *fout)
{
    const int W = h[1], N = h[2], K = h[6];
    if (W == 16 && N == 9) return solve_all<16, 9>(K, fout);
    if (W == 32 && N == 9) return solve_all<32, 9>(K, fout);
    if (W == 64 && N == 9) return solve_all<64, 9>(K, fout);
    if (W == 16 && N == 11) return solve_all<16, 11>(K, fout);
    if (W == 64 && N == 11) return solve_all<64, 11>(K, fout);
    fprintf(stderr, "lanes_run: solve %d x %d not built here\n", W, N);
    return 4;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: lanes_run IN OUT\n");
        return 2;
    }
    fin = fopen(argv[1], "rb");
    FILE *fout = fopen(argv[2], "wb");
    if (!fin || !fout) {
        fprintf(stderr, "lanes_run: cannot open files\n");
        return 2;
    }
    auto h = rd<int>(16);
    const int rc = h[0] == 1 ? mode_invert(h.data(), fout)
                 : h[0] == 2 ? mode_solve(h.data(), fout) : 2;
    fclose(fout);
    fclose(fin);
    return rc;
}
