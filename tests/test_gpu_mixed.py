"""Mixed-precision form (FP32 spectra/RFs/convolution, FP64 J^T J sums,
solve and chi2) against the FP64 oracle, within each case's stated mixed
tolerance (tests/cases.py) on >= 99.5% of pixels per parameter."""
import sys

import numpy as np
import pytest

import cases
import oracle as O
import pmilos
from fixtures import FE6173, GUESS, TRIPLET, draw_models

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup():
    lam = O.grid(FE6173, -290, 20, 290)
    G = O.gauss_psf(60.0, lam)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=G)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, precision=pmilos.PM_MIXED)
    return s, ctx


def test_mixed_synthesis_and_rf():
    say("case mixed synth L=30")
    s, ctx = _setup()
    models = draw_models(64, seed=77)
    spec, rf = ctx.synth(models, rf=True)
    for i, m in enumerate(models):
        sp, d = O.synth(s, m, rf=True)
        # FP32 rounding of the intensity (~1) bounds the absolute error of all
        # four Stokes; 1e-6 of the continuum is 1000x below the 1e-3 noise
        assert np.abs(spec[i] - sp).max() < 1e-6 * np.abs(sp[0]).max()
        # per parameter, relative to its largest response over the Stokes
        dsc = np.maximum(np.abs(d).max(axis=(0, 2), keepdims=True), 1e-5)
        assert (np.abs(rf[i] - d) / dsc).max() < 2e-3


@pytest.mark.parametrize("name", ["mix_s1e-4", "mix_s1e-3", "mix_s3e-3"])
def test_mixed_inversion_vs_fp64_oracle(name):
    """The noise sweep sigma in {1e-4, 1e-3, 3e-3} Ic (4096 pixels each):
    the mixed form with its default exit threshold and with the
    reference's test."""
    say("case mixed invert %s" % name)
    c = cases.get(name)
    ctx = c.context(pmilos, precision=pmilos.PM_MIXED)
    cases.compare(ctx.invert(c.cube), name, "mixed", say, cube=c.cube)
    ctx.set_mixed_toplim(0.0)
    cases.compare(ctx.invert(c.cube), name, "mixed", say)


def test_mixed_exit_threshold():
    """Mixed contexts start with PM_MIXED_TOPLIM as their chi2-unchanged
    threshold; 0 restores the reference's toplim test.  The default ends no
    later on average, and the FP64 form ignores the setting (same bits)."""
    say("case mixed exit threshold")
    c = cases.get("mix_s1e-3")
    ctx = c.context(pmilos, precision=pmilos.PM_MIXED)
    g_def = ctx.invert(c.cube)
    ctx.set_mixed_toplim(0.0)
    g_ref = ctx.invert(c.cube)
    assert g_def["iters"].mean() <= g_ref["iters"].mean()
    c64 = c.context(pmilos)
    a = c64.invert(c.cube[:256])
    c64.set_mixed_toplim(1e-2)
    b = c64.invert(c.cube[:256])
    for k in ("model", "iters", "chi2"):
        assert np.array_equal(a[k], b[k]), k
