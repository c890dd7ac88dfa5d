"""The library's device code run on the CPU, one thread per lane, under
AddressSanitizer and UBSan (tests/lanes: hip_lanes.h emulates the cross-lane
operations, lanes_run.cpp launches one workgroup of the inversion kernel or
the group solve).

- The inversion kernel of the 9-row form (FP64 8x1, 16x1, 16x2, 8x2, 32x1;
  FP32 16x2) and of the general form with a free v_mac over 112 samples
  (FP64 and FP32 64x2) inverts seeded pixels to the oracle's models, with no
  sanitizer report.  Bits are not compared with the GPU (the emulation
  divides where the GPU takes its hardware reciprocal).
- The lane-parallel LDL^T (ldl_solve_lanes) gives ldl_solve's bits on every
  matrix, every lane holds the same delta and verdict, and its LDS
  workspace is zero again afterwards.
"""
import numpy as np
import pytest

import cases
import lanes_driver as LD
import oracle as O
from fixtures import FE6173, FE6302, GUESS, TRIPLET, add_noise, draw_models

GRIDS = {6: (FE6173, -172.5, 69, 172.5), 14: (FE6173, -260, 40, 260),
         30: (FE6173, -290, 20, 290), 112: (FE6302, -1193.25, 21.5, 1193.25)}


@pytest.fixture(scope="module")
def built():
    return LD.build()


def _case(L, n, seed, mac=False):
    g = GRIDS[L]
    lam = O.grid(*g)
    G = O.gauss_psf(60.0, lam)
    guess, fix = GUESS.copy(), [1] * 9 + [0, 0]
    if mac:
        guess[9], fix = 0.5, [1] * 10 + [0]
    s = O.make_setup(g[0], TRIPLET, lam, psf=G, fix=fix, guess_mac=guess[9])
    m = draw_models(n, seed, mac_range=mac)
    cube = add_noise(np.array([O.synth(s, mm) for mm in m]), m, 1e-3, seed + 1)
    cfg = {"lambda": lam, "quantum": TRIPLET, "lambda0": g[0], "psf": G, "fix": fix,
           "guess": guess}
    return s, cfg, cube, guess


# (shape = precision, lanes, samples per lane, threads, kernel: 0 9-row, 1 11-row
# general, 2 10-row general), sampling,
# pixels, v_mac free
KERNELS = [
    ((0, 8, 1, 128, False), 6, 16, False),
    ((0, 16, 1, 128, False), 14, 16, False),
    ((0, 8, 2, 128, False), 14, 16, False),
    ((0, 16, 2, 128, False), 30, 16, False),
    ((0, 32, 1, 128, False), 30, 16, False),
    ((1, 16, 2, 128, False), 30, 16, False),
    ((0, 64, 2, 128, 1), 112, 4, True),
    ((1, 64, 2, 128, 1), 112, 4, True),
    ((0, 64, 2, 128, 2), 112, 4, True),
    ((1, 64, 2, 128, 2), 112, 4, True),
    ((0, 128, 1, 128, 2), 112, 4, True),
    ((0, 128, 1, 128, 1), 112, 4, True),
    ((0, 128, 1, 128, 0), 112, 4, False),
]


@pytest.mark.parametrize("shape,L,n,mac", KERNELS,
                         ids=["%s-%dx%d-%s-L%d" % ("fp64" if k[0][0] == 0 else "mixed", k[0][1],
                                                   k[0][2], ("9row", "gen11", "gen10")[int(k[0][4])],
                                                   k[1])
                              for k in KERNELS])
def test_kernel_on_cpu_lanes_matches_oracle(built, shape, L, n, mac):
    s, cfg, cube, guess = _case(L, n, 960 + L, mac)
    r = LD.invert(shape, cfg, cube, want_fit=True)
    o = O.invert(s, cube, guess)
    nt = 10 if mac else 9
    d = np.abs(r["model"][:, :nt] - o["model"][:, :nt])
    scale = np.maximum(1.0, np.abs(o["model"][:, :nt]))
    # FP64: the oracle's models to 1e-6 relative on at least 90% of the
    # pixels (rounding-level; the GPU parity floor is tighter still); the
    # mixed form: every pixel within the mixed form's stated tolerance of
    # its sampling (tests/golden/tolerances.json)
    if shape[0] == 0:
        ok = np.all(d <= 1e-6 * scale, axis=1)
        assert ok.mean() >= 0.9, (ok.mean(), d.max(axis=0))
    else:
        tol = cases.tolerance("gen_mac_nt10_l112" if mac else "l30_config1", "mixed")[:nt]
        assert np.all(d <= tol), (d.max(axis=0), tol)
    assert np.all(r["iters"] >= 1) and np.all(r["iters"] <= 50)
    assert np.all(np.isfinite(r["fit"]))


def _matrices(N, K, seed):
    rng = np.random.default_rng(seed)
    out, damp, chk = [], [], []
    for k in range(K):
        J = rng.normal(size=(3 * N, N)) * np.exp(rng.uniform(-4, 4, N))
        if k % 13 == 5:                    # a tiny eigenvalue: the bound 1/trace(A^-1) fails
            J[:, 4] *= 1e-8
        A = J.T @ J
        if k % 5 == 1:                     # a fixed parameter: zero row and column
            A[N - 1, :] = 0
            A[:, N - 1] = 0
        if k % 7 == 2:
            for i in (0, 3):
                A[i, :] = 0
                A[:, i] = 0
        if k % 11 == 4:                    # not positive definite: the factorisation stops
            A[2, 2] = -abs(A[2, 2])
        b = rng.normal(size=N)
        out.append(np.concatenate([np.concatenate([A[i, :i + 1] for i in range(N)]), b]))
        damp.append(1.0 + 10 ** rng.uniform(-6, 1))
        chk.append(1.0 if k % 3 else 0.0)
    return np.array(out), damp, chk


@pytest.mark.parametrize("W,N", [(16, 9), (32, 9), (64, 9), (16, 11), (64, 11)])
def test_lane_parallel_solve_gives_the_replicated_bits(built, W, N):
    ab, damp, chk = _matrices(N, 60, 100 * W + N)
    xg, okg, xr, okr, consistent = LD.solve(W, N, ab, damp, chk)
    assert np.all(consistent)
    assert np.array_equal(okg, okr)
    assert okg.mean() < 1.0 and okg.mean() > 0.5     # both outcomes are exercised
    for k in np.nonzero(okg)[0]:
        assert np.array_equal(xg[k].view(np.int64), xr[k].view(np.int64)), k
