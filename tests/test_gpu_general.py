"""GPU parity of the general form -- NTERMS 10 and 11, macroturbulence,
stray light with its filling factor -- against the CPU oracle on the same
seeded inputs, through the C ABI.  The reference's own golden synthesis
(test/synthesis.mtrol, stray light with alpha 0.85) runs through the GPU."""
import os
import sys

import numpy as np
import pytest

import cases
import oracle as O
import pmilos
from fixtures import (FE6173, FE6302, GUESS, TRIPLET, add_noise, draw_models,
                      oracle_cube, read_fits_2d, read_mod)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRIDS = {30: (FE6173, -290, 20, 290), 49: (FE6173, -600, 25, 600),
         112: (FE6302, -1193.25, 21.5, 1193.25), 31: (FE6173, -290, 20, 310)}

# (fix, guess v_mac, stray light) of each NTERMS layout the reference has
# (milos.c:154-157; me_der.c:418-425 and :499-525 for where rows 9/10 go)
CASES = {
    "mac_nt10": ([1] * 10 + [0], 0.5, False),       # v_mac free, row 9
    "mac_nt11_sl": ([1] * 11, 0.5, True),           # v_mac row 9, filling factor row 10
    "ff_nt10_sl": ([1] * 9 + [0, 1], 0.0, True),    # no v_mac, filling factor row 9
    "macfix_nt10_sl": ([1] * 9 + [0, 0], 0.5, True),  # v_mac fixed but applied; row 9
    "nt9_sl": ([1] * 9 + [0, 0], 0.0, True),        # alpha fixed, stray light mixed in
    "nt9": ([1] * 9 + [0, 0], 0.0, False),          # the 9-parameter layout
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _slight_profile(lam, l0):
    """A float32 stray-light profile: a broadened, weak-field line of the
    same grid (what a quiet-Sun average looks like)."""
    s = O.make_setup(l0, TRIPLET, lam)
    m = np.array([6.0, 150.0, 0.3, 0.045, 0.2, 60.0, 20.0, 0.25, 0.75, 0, 1])
    return O.synth(s, m).astype(np.float32)


def _pair(case, L, psf, precision=pmilos.PM_FP64, **kw):
    fix, gmac, use_sl = CASES[case]
    g = GRIDS[L]
    lam = O.grid(*g)
    G = O.gauss_psf(60.0, lam) if psf else None
    sl = _slight_profile(lam, g[0]) if use_sl else None
    s = O.make_setup(g[0], TRIPLET, lam, psf=G, fix=fix, guess_mac=gmac,
                     slight=None if sl is None else sl.astype(np.float64), **kw)
    guess = GUESS.copy()
    guess[9] = gmac
    guess[10] = 1.0 if not use_sl else 0.9
    ctx = pmilos.Context(g[0], TRIPLET, lam, psf=G, fix=fix, guess=guess, slight=sl,
                         precision=precision, **kw)
    assert ctx.NT == s.nterms
    return s, ctx, guess, sl


def _models(n, seed, case):
    fix, gmac, use_sl = CASES[case]
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


def _err_spectra(a, b):
    floor = 1e-6 * np.abs(b[0]).max()
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), floor)
    return (np.abs(a - b) / scale).max()


def _err_rf(a, b):
    scale = np.maximum(np.abs(b).max(axis=(0, 2), keepdims=True), 1e-5)
    return (np.abs(a - b) / scale).max()


def _golden():
    slight = read_fits_2d(os.path.join(GOLDEN, "straylight.fits"))
    lam = O.grid(FE6173, -600, 25, 600)
    model = read_mod(os.path.join(GOLDEN, "obs.mod"))
    gold = np.loadtxt(os.path.join(GOLDEN, "synthesis_results.per"))[:, 2:].T
    return slight, lam, model, gold


@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_golden_synthesis_gpu(precision):
    """test/synthesis.mtrol -> test/synthesis_results.per: 49 wavelengths,
    no PSF, stray light mixed with filling factor 0.85, NTERMS 10."""
    say("case golden synthesis precision=%d" % precision)
    slight, lam, model, gold = _golden()
    fix = [1] * 9 + [0, 1]
    ctx = pmilos.Context(FE6173, TRIPLET, lam, fix=fix, guess=model, slight=slight,
                         precision=precision)
    s = O.make_setup(FE6173, TRIPLET, lam, fix=fix, guess_mac=model[9], slight=slight)
    assert ctx.NT == s.nterms == 10
    spec, rf = ctx.synth(model, rf=True)
    spec, rf = spec[0], rf[0]
    # the golden file is printed with %e (7 digits) from the float build
    assert np.all(np.abs(spec - gold) <= 3e-7 + 1e-6 * np.abs(gold))
    sp_o, rf_o = O.synth(s, model, rf=True)
    if precision == pmilos.PM_FP64:
        assert _err_spectra(spec, sp_o) < 1e-10
        assert _err_rf(rf, rf_o) < 1e-10
    else:
        assert np.abs(spec - sp_o).max() < 1e-6 * np.abs(sp_o[0]).max()
        assert _err_rf(rf, rf_o) < 2e-3


@pytest.mark.parametrize("psf", [True, False])
@pytest.mark.parametrize("case,L", [("mac_nt10", 112), ("mac_nt10", 30),
                                    ("mac_nt11_sl", 49), ("mac_nt11_sl", 112),
                                    ("ff_nt10_sl", 30), ("macfix_nt10_sl", 49),
                                    ("nt9_sl", 30)])
def test_synthesis_and_rf_general(case, L, psf):
    """Spectra and response functions, rows in the reference's NTERMS layout
    (the filling-factor row included), FP64 within 1e-10 and the mixed form
    within its stated tolerances."""
    say("case general synth %s L=%d psf=%s" % (case, L, psf))
    s, ctx, guess, sl = _pair(case, L, psf)
    _, ctxm, _, _ = _pair(case, L, psf, precision=pmilos.PM_MIXED)
    models = _models(40, 2000 + L, case)
    spec_g, rf_g = ctx.synth(models, rf=True)
    spec_m, rf_m = ctxm.synth(models, rf=True)
    for i, m in enumerate(models):
        sp, rf = O.synth(s, m, rf=True)
        e1, e2 = _err_spectra(spec_g[i], sp), _err_rf(rf_g[i], rf)
        assert e1 < 1e-10, (i, e1)
        assert e2 < 1e-10, (i, e2)
        em = np.abs(spec_m[i] - sp).max() / np.abs(sp[0]).max()
        assert em < 1e-6, (i, em)
        erm = _err_rf(rf_m[i], rf)
        assert erm < 2e-3, (i, erm)


@pytest.mark.parametrize("name", ["gen_mac_nt10_l112", "gen_mac_nt11_sl_l49",
                                  "gen_ff_nt10_sl_l30", "gen_nt9_sl_l30"])
def test_invert_general_fp64(name):
    """Inversions of seeded cubes (sigma 1e-3 Ic) with v_mac ~ U(0, 2) km/s
    and, with stray light, alpha ~ U(0.6, 1): FP64 against the oracle within
    the case's stated tolerance on >= 99.5% of pixels per parameter."""
    say("case general invert %s" % name)
    c = cases.get(name)
    g = c.context(pmilos).invert(c.cube)
    cases.compare(g, name, "fp64", say, cube=c.cube)
    # parameters that are not free keep the guess
    fix = c.ctx_kw["fix"]
    for i in range(9, 11):
        if not fix[i]:
            assert np.all(g["model"][:, i] == c.guess[i])


def test_invert_general_mixed():
    """The mixed form of the 112-sample v_mac inversion against the FP64
    oracle, within its stated mixed tolerance."""
    name = "gen_mac_nt10_l112"
    say("case general invert mixed %s" % name)
    c = cases.get(name)
    g = c.context(pmilos, precision=pmilos.PM_MIXED).invert(c.cube)
    cases.compare(g, name, "mixed", say, cube=c.cube)


def test_invert_per_pixel_stray_light():
    """A stray-light profile per pixel (the 4-D stray-light cube of
    milos.c:986-996): each pixel against the oracle set up with its own
    profile, and a cube repeating the context's profile gives exactly the
    context-profile inversion."""
    name = "per_pixel_slight"
    say("case per-pixel stray light %s" % name)
    c = cases.get(name)
    ctx = c.context(pmilos)
    g = ctx.invert(c.cube, slight=c.slight)
    cases.compare(g, name, "fp64", say, cube=c.cube)
    sl = c.ctx_kw["slight"]
    rep = np.repeat(sl[None], len(c.cube), axis=0)
    a = ctx.invert(c.cube[:256], slight=rep[:256])
    b = ctx.invert(c.cube[:256])
    assert np.array_equal(a["model"], b["model"]) and np.array_equal(a["iters"], b["iters"])
    # host buffers of more than 2^20 pixels (pm_invert_sl splits them over
    # its streams): the cube tiled gives the untiled results tiled, fit
    # included
    reps = -(-(1 << 20) // len(c.cube)) + 1
    say("case per-pixel stray light, %d pixels in chunks" % (reps * len(c.cube)))
    one = ctx.invert(c.cube, slight=c.slight, want_fit=True)
    big = ctx.invert(np.tile(c.cube, (reps, 1, 1)), slight=np.tile(c.slight, (reps, 1, 1)),
                     want_fit=True)
    for k in ("model", "chi2", "iters", "accepted", "fit"):
        t = np.tile(one[k], (reps,) + (1,) * (one[k].ndim - 1))
        assert np.array_equal(big[k], t), k


# the FFT convolution mode (useFFT): even and odd grids, PSF alone and
# folded into the macroturbulence kernel, with stray light
FFT_CASES = [("nt9", 30), ("nt9", 31), ("mac_nt10", 30), ("mac_nt11_sl", 49)]


@pytest.mark.parametrize("case,L", FFT_CASES)
def test_synthesis_and_rf_fft_mode(case, L):
    """useFFT: the PSF by circular convolution of every Stokes and response
    row, and folded into the macroturbulence kernel circularly
    (mil_sinrf.c:301-386, me_der.c:331-470), against the oracle's FFT mode:
    FP64 within 1e-10, the mixed form within its stated tolerances.  The
    PSF-only models differ from the direct mode's near the window edges."""
    say("case fft synth %s L=%d" % (case, L))
    s, ctx, guess, sl = _pair(case, L, True, use_fft=True)
    _, ctxm, _, _ = _pair(case, L, True, precision=pmilos.PM_MIXED, use_fft=True)
    _, ctxd, _, _ = _pair(case, L, True)
    models = _models(40, 5000 + L, case)
    spec_g, rf_g = ctx.synth(models, rf=True)
    spec_m, rf_m = ctxm.synth(models, rf=True)
    spec_d = ctxd.synth(models)
    for i, m in enumerate(models):
        sp, rf = O.synth(s, m, rf=True)
        e1, e2 = _err_spectra(spec_g[i], sp), _err_rf(rf_g[i], rf)
        assert e1 < 1e-10, (i, e1)
        assert e2 < 1e-10, (i, e2)
        em = np.abs(spec_m[i] - sp).max() / np.abs(sp[0]).max()
        assert em < 1e-6, (i, em)
        erm = _err_rf(rf_m[i], rf)
        assert erm < 2e-3, (i, erm)
        if m[9] <= 1e-4:
            assert np.abs(spec_d[i] - spec_g[i]).max() > 1e-9, i


@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_invert_fft_mode(precision):
    """FFT-mode inversions of a seeded 30-sample cube (sigma 1e-3 Ic)
    against the oracle's FFT mode."""
    name = "fft_nt9_l30"
    form = "fp64" if precision == pmilos.PM_FP64 else "mixed"
    say("case fft invert %s %s" % (name, form))
    c = cases.get(name)
    g = c.context(pmilos, precision=precision).invert(c.cube)
    cases.compare(g, name, form, say, cube=c.cube)
