"""The seeded inversion cases of the GPU parity tests, their oracle results
and their stated per-parameter tolerances.

Each case is a float32 Stokes cube synthesised by the FP64 oracle from
seeded models plus seeded noise (the recipe of the benchmark
configurations), the oracle set-up and the library context that invert it,
and the guess.  tests/golden/make_oracle_cases.py inverts every cube with
the FP64 oracle once and stores the results (tests/golden/oracle_cases.npz,
checked against the live oracle by tests/test_cases_fixture.py), and derives
each case's tolerances from how far builds of the oracle that differ only in
rounding end apart on the same cube (tests/golden/tolerances.json; the rule
is in that script's docstring).  A GPU form passes a case when at least
FRAC of the pixels are within the tolerance on every parameter.
"""
import hashlib
import json
import os

import numpy as np

import oracle as O
from fixtures import (FE6173, FE6302, GUESS, TRIPLET, add_noise, draw_models,
                      oracle_cube)

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "oracle_cases.npz")
TOLERANCES = os.path.join(HERE, "golden", "tolerances.json")

GRIDS = {6: (FE6173, -172.5, 69, 172.5), 14: (FE6173, -260, 40, 260),
         30: (FE6173, -290, 20, 290), 31: (FE6173, -290, 20, 310),
         49: (FE6173, -600, 25, 600), 112: (FE6302, -1193.25, 21.5, 1193.25)}
NAMES = ["eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa"]
# the smallest tolerance stated (absolute, Init_Model order): where
# implementations differing only in rounding end closer than this on a case,
# its tolerance is this
FLOOR = np.array([1e-3, 1e-2, 1e-6, 1e-8, 1e-6, 1e-4, 1e-4, 1e-6, 1e-6, 1e-6, 1e-6])
FRAC = 0.995

# (fix, guess v_mac, stray light) of each NTERMS layout the general tests use
LAYOUTS = {
    "mac_nt10": ([1] * 10 + [0], 0.5, False),
    "mac_nt11_sl": ([1] * 11, 0.5, True),
    "ff_nt10_sl": ([1] * 9 + [0, 1], 0.0, True),
    "nt9_sl": ([1] * 9 + [0, 0], 0.0, True),
    "nt9": ([1] * 9 + [0, 0], 0.0, False),
}


def slight_profile(lam, l0):
    """A float32 stray-light profile: a broadened, weak-field line of the
    same grid (what a quiet-Sun average looks like)."""
    s = O.make_setup(l0, TRIPLET, lam)
    m = np.array([6.0, 150.0, 0.3, 0.045, 0.2, 60.0, 20.0, 0.25, 0.75, 0, 1])
    return O.synth(s, m).astype(np.float32)


def general_models(n, seed, layout):
    fix, gmac, use_sl = LAYOUTS[layout]
    m = draw_models(n, seed, mac_range=True)
    rng = np.random.default_rng(seed + 7)
    if use_sl:
        m[:, 10] = rng.uniform(0.6, 1.0, n)
    m[1, 9] = 0.0           # no macroturbulence: the PSF (or nothing) instead
    m[2, 9] = 5e-5          # below the 1e-4 switch of mil_sinrf.c:288
    m[3, 1] = 0.0           # B = 0
    if not fix[9] and gmac == 0.0:
        # NTERMS has no v_mac row: the reference never has v_mac > 0 here
        m[:, 9] = 0.0
    return m


class Case:
    """setup(): the oracle set-up; ctx_kw: pmilos.Context keyword arguments
    (besides the grid); cube, guess; per_pixel: optional per-pixel oracle
    set-ups (a stray-light profile per pixel) with their profiles."""

    def __init__(self, name, L, setup_kw, ctx_kw, cube, guess, per_pixel=None, slight=None):
        self.name, self.L = name, L
        self.l0 = GRIDS[L][0]
        self.lam = O.grid(*GRIDS[L])
        self.setup_kw, self.ctx_kw = setup_kw, ctx_kw
        self.cube, self.guess = cube, guess
        self.per_pixel, self.slight = per_pixel, slight

    def setup(self):
        return O.make_setup(self.l0, TRIPLET, self.lam, **self.setup_kw)

    def oracle(self):
        """The oracle's inversion of the cube with the build in use."""
        if self.per_pixel is None:
            return O.invert(self.setup(), self.cube, self.guess)
        out = {k: [] for k in ("model", "chi2", "iters", "accepted")}
        for i, kw in enumerate(self.per_pixel):
            s = O.make_setup(self.l0, TRIPLET, self.lam, **kw)
            r = O.invert(s, self.cube[i:i + 1], self.guess)
            for k in out:
                out[k].append(r[k][0])
        return {k: np.array(v) for k, v in out.items()}

    def context(self, pmilos, **kw):
        a = dict(self.ctx_kw)
        a.update(kw)
        return pmilos.Context(self.l0, TRIPLET, self.lam, guess=self.guess, **a)


def _plain(name, L, n, mseed, nseed, fwhm=60.0, sigma=1e-3, fix=None, logclambda=0,
           guess=GUESS, edit=None):
    lam = O.grid(*GRIDS[L])
    G = O.gauss_psf(fwhm, lam) if fwhm else None
    kw = dict(psf=G, logclambda=logclambda)
    if fix is not None:
        kw["fix"] = fix
    s = O.make_setup(GRIDS[L][0], TRIPLET, lam, **kw)
    truths = draw_models(n, seed=mseed)
    cube = add_noise(oracle_cube(s, truths), truths, sigma, seed=nseed)
    if edit:
        edit(cube)
    return Case(name, L, kw, dict(kw), cube, guess)


def _edge_edit(cube):
    cube[3, 1, 4] = -5.0                 # ghost point
    cube[4, 0, :3] = -2.0                # leading ghost wavelengths in I
    cube[5, 2, :] = -3.0                 # ghost U
    cube[6] = np.nan                     # NaN pixel


def _general(name, layout, L, psf, n, mseed, nseed, use_fft=False):
    fix, gmac, use_sl = LAYOUTS[layout]
    lam = O.grid(*GRIDS[L])
    G = O.gauss_psf(60.0, lam) if psf else None
    sl = slight_profile(lam, GRIDS[L][0]) if use_sl else None
    skw = dict(psf=G, fix=fix, guess_mac=gmac, use_fft=use_fft,
               slight=None if sl is None else sl.astype(np.float64))
    ckw = dict(psf=G, fix=fix, slight=sl, use_fft=use_fft)
    guess = GUESS.copy()
    guess[9] = gmac
    guess[10] = 1.0 if not use_sl else 0.9
    s = O.make_setup(GRIDS[L][0], TRIPLET, lam, **skw)
    truths = general_models(n, mseed, layout)
    cube = add_noise(oracle_cube(s, truths), truths, 1e-3, seed=nseed)
    return Case(name, L, skw, ckw, cube, guess, slight=sl)


def _per_pixel_slight(name, n=48, L=30):
    layout = "ff_nt10_sl"
    fix, gmac, _ = LAYOUTS[layout]
    lam = O.grid(*GRIDS[L])
    G = O.gauss_psf(60.0, lam)
    sl = slight_profile(lam, FE6173)
    truths = general_models(n, 4000, layout)
    rng = np.random.default_rng(41)
    scl = (sl[None] * rng.uniform(0.8, 1.2, (n, 1, 1))).astype(np.float32)
    cube = np.zeros((n, 4, L), dtype=np.float32)
    per = []
    for i in range(n):
        kw = dict(psf=G, fix=fix, slight=scl[i].astype(np.float64))
        si = O.make_setup(FE6173, TRIPLET, lam, **kw)
        cube[i] = add_noise(O.synth(si, truths[i])[None], truths[i:i + 1], 1e-3, seed=50 + i)[0]
        per.append(kw)
    guess = GUESS.copy()
    guess[10] = 0.9
    c = Case(name, L, dict(psf=G, fix=fix, slight=sl.astype(np.float64)),
             dict(psf=G, fix=fix, slight=sl), cube, guess, per_pixel=per, slight=scl)
    return c


def _mac_variant(name, L, n, mseed, nseed):
    lam = O.grid(*GRIDS[L])
    G = O.gauss_psf(60.0, lam)
    guess, fix = GUESS.copy(), [1] * 10 + [0]
    guess[9] = 0.5
    kw = dict(psf=G, fix=fix)
    s = O.make_setup(GRIDS[L][0], TRIPLET, lam, guess_mac=0.5, **kw)
    m = draw_models(n, mseed, mac_range=True)
    cube = add_noise(oracle_cube(s, m), m, 1e-3, seed=nseed)
    return Case(name, L, dict(kw, guess_mac=0.5), kw, cube, guess)


# name -> (builder, its arguments); the forms each case is inverted in by
# the GPU tests decide which spread a test reads (fp64: build-to-build,
# mixed: float-versus-double)
BUILDERS = {
    "l30_config1": (_plain, dict(L=30, n=4096, mseed=1, nseed=11)),
    "psf20": (_plain, dict(L=30, n=2048, mseed=1301, nseed=1302, fwhm=20.0)),
    "psf120": (_plain, dict(L=30, n=2048, mseed=1301, nseed=1302, fwhm=120.0)),
    "l6": (_plain, dict(L=6, n=4096, mseed=46, nseed=7)),
    "l112": (_plain, dict(L=112, n=2048, mseed=152, nseed=7)),
    "edge": (_plain, dict(L=30, n=1024, mseed=5, nseed=6, fix=[1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0],
                          edit=_edge_edit)),
    "logc": (_plain, dict(L=30, n=2048, mseed=9, nseed=10, logclambda=1)),
    "var_l6": (_plain, dict(L=6, n=4096, mseed=706, nseed=707)),
    "var_l14": (_plain, dict(L=14, n=4096, mseed=714, nseed=715)),
    "var_l30": (_plain, dict(L=30, n=4096, mseed=730, nseed=731)),
    "var_l49": (_plain, dict(L=49, n=2048, mseed=749, nseed=750)),
    "var_l112": (_plain, dict(L=112, n=2048, mseed=812, nseed=813)),
    "gen_mac_nt10_l112": (_general, dict(layout="mac_nt10", L=112, psf=True, n=1024, mseed=3112,
                                         nseed=31)),
    "gen_mac_nt11_sl_l49": (_general, dict(layout="mac_nt11_sl", L=49, psf=False, n=2048,
                                           mseed=3049, nseed=31)),
    "gen_ff_nt10_sl_l30": (_general, dict(layout="ff_nt10_sl", L=30, psf=True, n=2048,
                                          mseed=3030, nseed=31)),
    "gen_nt9_sl_l30": (_general, dict(layout="nt9_sl", L=30, psf=True, n=2048, mseed=3030,
                                      nseed=31)),
    "fft_nt9_l30": (_general, dict(layout="nt9", L=30, psf=True, n=2048, mseed=6030, nseed=37,
                                   use_fft=True)),
    "per_pixel_slight": (_per_pixel_slight, dict(n=1024)),
    "mix_s1e-4": (_plain, dict(L=30, n=4096, mseed=2, nseed=31, sigma=1e-4)),
    "mix_s1e-3": (_plain, dict(L=30, n=4096, mseed=2, nseed=32, sigma=1e-3)),
    "mix_s3e-3": (_plain, dict(L=30, n=4096, mseed=2, nseed=33, sigma=3e-3)),
}


def get(name):
    fn, kw = BUILDERS[name]
    return fn(name, **kw)


def cube_sha256(cube):
    return hashlib.sha256(np.ascontiguousarray(cube).tobytes()).hexdigest()


_fixture = None


def reference(name, cube=None):
    """The FP64 oracle's results for a case (model, iters, chi2) from the
    fixture; with cube, first checks that it is the cube they are for."""
    global _fixture
    if _fixture is None:
        _fixture = dict(np.load(FIXTURE))
    if cube is not None:
        assert cube_sha256(cube) == str(_fixture[name + "|sha256"]), \
            "%s: the regenerated cube differs from the fixture's" % name
    return {k: _fixture["%s|%s" % (name, k)] for k in ("model", "iters", "chi2")}


def tolerance(name, form="fp64"):
    """The stated tolerance (11 parameters, absolute) of a form
    ("fp64" or "mixed") on a case."""
    with open(TOLERANCES) as f:
        t = json.load(f)[name][form]
    return np.array([t[p] for p in NAMES])


def compare(g, name, form="fp64", say=print, frac=FRAC, cube=None):
    """Asserts that the GPU results g (dict with model, iters) are within
    the case's tolerance of the oracle on at least frac of the pixels, on
    every parameter; reports the fractions, the worst parameter and the
    trip-count differences."""
    ref = reference(name, cube)
    tol = tolerance(name, form)
    a, b = g["model"], ref["model"]
    ok = np.all(np.isfinite(a), axis=1) & np.all(np.isfinite(b), axis=1)
    assert np.array_equal(np.all(np.isfinite(a), axis=1), np.all(np.isfinite(b), axis=1))
    d = np.abs(a[ok] - b[ok])
    f = (d <= tol).mean(axis=0)
    di = g["iters"][ok].astype(np.int64) - ref["iters"][ok]
    say("%s %s: fraction within tolerance %s (worst %s %.4f); trips identical on %.4f, "
        "mean |d trips| %.3f, mean trips %.2f (oracle %.2f)"
        % (name, form, np.round(f, 4), NAMES[int(f.argmin())], f.min(), (di == 0).mean(),
           np.abs(di).mean(), g["iters"][ok].mean(), ref["iters"][ok].mean()))
    assert np.all(f >= frac), (name, form, dict(zip(NAMES, np.round(f, 4))))
    return f
