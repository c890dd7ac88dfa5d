"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on
the same seeded inputs.  FP64 synthesis and response functions must agree
within 1e-10 relative; inversions within each case's stated tolerance
(tests/cases.py) on every parameter on >= 99.5% of pixels."""
import sys

import numpy as np
import pytest

import cases
import oracle as O
import pmilos
from fixtures import FE6173, FE6302, GUESS, TRIPLET, draw_models

pytestmark = pytest.mark.gpu

def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grids():
    return {
        6: (FE6173, -172.5, 69, 172.5),
        14: (FE6173, -260, 40, 260),
        30: (FE6173, -290, 20, 290),
        49: (FE6173, -600, 25, 600),
        112: (FE6302, -1193.25, 21.5, 1193.25),
    }


def _pair(L, psf=True, **kw):
    lam = O.grid(*_grids()[L])
    l0 = _grids()[L][0]
    G = O.gauss_psf(60.0, lam) if psf else None
    s = O.make_setup(l0, TRIPLET, lam, psf=G, **{k: v for k, v in kw.items() if k != "guess"})
    ctx = pmilos.Context(l0, TRIPLET, lam, psf=G, guess=kw.get("guess", GUESS),
                         **{k: v for k, v in kw.items() if k in ("fix", "noise", "miter", "logclambda", "ilambda")})
    return s, ctx


def _err_spectra(a, b):
    """max |a-b| relative to each Stokes row's scale, floored at 1e-6 of the
    intensity (Q, U, V are exactly 0 when B = 0 or gamma = 0)."""
    floor = 1e-6 * np.abs(b[0]).max()
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), floor)
    return (np.abs(a - b) / scale).max()


def _err_rf(a, b):
    """max |a-b| relative to each parameter's largest response over the four
    Stokes parameters, floored at 1e-5 per unit parameter: at B = 0 the gamma
    and azimuth responses vanish for all four Stokes and both sides hold only
    rounding noise (~1e-15), while real responses are >= 1e-5."""
    scale = np.maximum(np.abs(b).max(axis=(0, 2), keepdims=True), 1e-5)
    return (np.abs(a - b) / scale).max()


@pytest.mark.parametrize("L", [6, 14, 30, 49, 112])
@pytest.mark.parametrize("psf", [True, False])
def test_synthesis_and_rf_fp64(L, psf):
    say("case synth L=%d psf=%s" % (L, psf))
    s, ctx = _pair(L, psf=psf)
    models = draw_models(48, seed=1000 + L)
    models[0] = [14., 1200., 0.5, 0.07, 0.05, 130., 25., 0.2, 0.75, 0, 1.]
    models[1, 1] = 0.0      # B = 0
    models[2, 5] = 0.0      # gamma = 0
    spec_g, rf_g = ctx.synth(models, rf=True)
    for i, m in enumerate(models):
        sp, rf = O.synth(s, m, rf=True)
        e1 = _err_spectra(spec_g[i], sp)
        e2 = _err_rf(rf_g[i], rf)
        assert e1 < 1e-10, (i, e1)
        assert e2 < 1e-10, (i, e2)


@pytest.mark.parametrize("fwhm", [20.0, 120.0])
@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_synthesis_and_rf_psf_widths(fwhm, precision):
    """30 samples with a narrow PSF (three taps: the register-blocked
    correlation with most of its window zero) and a wide one (taps beyond
    its window: the tap loop), against the oracle."""
    say("case synth psf fwhm=%g precision=%d" % (fwhm, precision))
    lam = O.grid(*_grids()[30])
    G = O.gauss_psf(fwhm, lam)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=G)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, precision=precision)
    models = draw_models(32, seed=1300)
    spec_g, rf_g = ctx.synth(models, rf=True)
    for i, m in enumerate(models):
        sp, rf = O.synth(s, m, rf=True)
        if precision == pmilos.PM_FP64:
            assert _err_spectra(spec_g[i], sp) < 1e-10, i
            assert _err_rf(rf_g[i], rf) < 1e-10, i
        else:
            assert np.abs(spec_g[i] - sp).max() < 1e-6 * np.abs(sp[0]).max(), i
            assert _err_rf(rf_g[i], rf) < 2e-3, i
    if precision == pmilos.PM_FP64:
        name = "psf%d" % fwhm
        say("case invert %s" % name)
        c = cases.get(name)
        cases.compare(c.context(pmilos).invert(c.cube), name, "fp64", say, cube=c.cube)


def _invert_case(name, **kw):
    say("case invert %s" % name)
    c = cases.get(name)
    g = c.context(pmilos, **kw).invert(c.cube)
    cases.compare(g, name, "fp64", say, cube=c.cube)
    return c, g


def test_invert_fp64_config1_field():
    """The CPU-runnable field: 64x64 pixels of seed-1 models, sigma 1e-3 Ic,
    PSF 60 mA, 30 lambda, 9 free, <= 50 iterations; the GPU and the oracle
    invert the same float32 cube."""
    say("case invert l30_config1")
    c = cases.get("l30_config1")
    ctx = c.context(pmilos)
    g = ctx.invert(c.cube, want_fit=True)
    f = cases.compare(g, "l30_config1", "fp64", say, cube=c.cube)
    # at the FP64 floor tolerance alone (every parameter to rounding-level
    # agreement) on more than 99.5% of the pixels
    ref = cases.reference("l30_config1")
    tight = (np.abs(g["model"] - ref["model"]) <= cases.FLOOR).mean(axis=0)
    say("l30_config1 within the floor tolerance %s" % np.round(tight, 4))
    assert np.all(tight >= 0.995)
    assert ctx.fallback_count() >= 0
    # chi2 agrees wherever the models agree at the floor
    both = np.all(np.abs(g["model"] - ref["model"]) <= cases.FLOOR, axis=1)
    np.testing.assert_allclose(g["chi2"][both], ref["chi2"][both], rtol=1e-4)
    # fits are the last trial synthesis of each pixel
    s = c.setup()
    o16 = O.invert(s, c.cube[:16], GUESS, want_fit=True)
    same = g["iters"][:16] == o16["iters"]
    np.testing.assert_allclose(g["fit"][:16][same], o16["fit"][same], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("name", ["l6", "l112"])
def test_invert_fp64_other_samplings(name):
    """6 samples (many pixels end at the 50-trip cap still moving) and 112
    (Fe I 6302.5)."""
    _invert_case(name)


def test_invert_edge_cases():
    """Ghost points (a sample, leading I samples, a whole Stokes row), a NaN
    pixel and a fixed damping."""
    c, g = _invert_case("edge")
    ref = cases.reference("edge")
    assert np.array_equal(np.isfinite(ref["model"][:, 0]), np.isfinite(g["model"][:, 0]))
    assert np.all(g["model"][:, 4] == GUESS[4])          # damping fixed


def test_invert_logclambda():
    """logclambda = 1.  glibc log10f vs the correctly rounded value can move
    the damping factor by 1-2 ulp of float."""
    _invert_case("logc")


def test_nfree_zero_pixel():
    say("case nfree==0")
    lam = O.grid(FE6173, -20, 20, 20)      # L = 3: 12 - 9 = 3 ghosts -> nfree 0
    ctx = pmilos.Context(FE6173, TRIPLET, lam, guess=GUESS)
    s = O.make_setup(FE6173, TRIPLET, lam)
    cube = np.full((5, 4, 3), 0.5, dtype=np.float32)
    cube[2, 1, :] = -2.0                  # 3 ghost samples in Q
    g = ctx.invert(cube)
    o = O.invert(s, cube, GUESS)
    assert g["iters"][2] == 2 and o["iters"][2] == 2      # the pixel index
    assert g["chi2"][2] == 0.0
    np.testing.assert_allclose(g["model"][2], o["model"][2], rtol=1e-12)
