"""The benchmark field at its full size: 1024 x 1024 pixels of Fe I 6173 at
30 wavelengths, seed-2 models, PSF 60 mA, sigma 1e-3 Ic, 9 free, <= 50
iterations, synthesised by the library's FP64 synthesis (tested against the
oracle to 1e-10 elsewhere).  Checked on every pixel through properties that
do not depend on the size, and against the oracle on pixels spread over the
whole field."""
import numpy as np
import pytest

import cases
import oracle as O
import pmilos
from fixtures import FE6173, GUESS, TRIPLET, draw_models
from test_gpu_fp64 import say

pytestmark = pytest.mark.gpu

NPIX = 1024 * 1024
GRID = (FE6173, -290, 20, 290)

# ranges that check() (milosUtils.c:118-196) keeps every trial model in:
# eta0, B, vlos, dopp, a, gamma, phi, S0, S1
LO = np.array([1.0, 0.0, -20.0, 1e-4, 1e-4, 0.0, 0.0, 1e-4, 1e-4])
HI = np.array([2500.0, 5000.0, 20.0, 0.6, 10.0, 180.0, 180.0, 2.0, 2.0])


@pytest.fixture(scope="module")
def field():
    import torch
    dev = torch.device("cuda:0")
    lam = pmilos.grid(*GRID)
    L = len(lam)
    G = pmilos.psf_gaussian(60.0, lam)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6, miter=50)
    models = torch.from_numpy(draw_models(NPIX, seed=2)).to(dev)
    spec = torch.empty((NPIX, 4, L), dtype=torch.float64, device=dev)
    ctx.synth_ptrs(models.data_ptr(), NPIX, spec.data_ptr())
    gen = torch.Generator(device=dev)
    gen.manual_seed(32)
    ic = (models[:, 7] + models[:, 8]).view(-1, 1, 1)
    stokes = (spec + torch.randn((NPIX, 4, L), dtype=torch.float64, device=dev, generator=gen)
              * 1e-3 * ic).to(torch.float32).contiguous()
    del spec
    torch.cuda.synchronize()
    return dict(torch=torch, dev=dev, ctx=ctx, stokes=stokes, lam=lam, G=G, L=L)


def _run(f, ctx, stokes):
    torch = f["torch"]
    n = stokes.shape[0]
    out = dict(model=torch.empty((n, 11), dtype=torch.float64, device=f["dev"]),
               chi2=torch.empty(n, dtype=torch.float32, device=f["dev"]),
               iters=torch.empty(n, dtype=torch.int32, device=f["dev"]),
               accepted=torch.empty(n, dtype=torch.int32, device=f["dev"]))
    ctx.invert_ptrs(stokes.data_ptr(), n, out["model"].data_ptr(), out["chi2"].data_ptr(),
                    out["iters"].data_ptr(), out["accepted"].data_ptr())
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.fixture(scope="module")
def fp64(field):
    say("case full field fp64 npix=%d" % NPIX)
    return _run(field, field["ctx"], field["stokes"])


def test_every_pixel_inverted_within_the_model_ranges(fp64):
    it, acc, c2, m = fp64["iters"], fp64["accepted"], fp64["chi2"], fp64["model"]
    assert it.min() >= 1 and it.max() <= 50
    assert np.all((acc >= 0) & (acc <= it))
    assert np.all(np.isfinite(c2)) and c2.min() >= 0
    assert np.all(np.isfinite(m))
    moved = acc > 0
    assert moved.mean() > 0.99
    mm = m[moved, :9]
    assert np.all((mm >= LO) & (mm <= HI)), (mm.min(axis=0), mm.max(axis=0))
    say("full field: mean iters %.3f accepted %.3f, median chi2 %.3f"
        % (it.mean(), acc.mean(), np.median(c2)))


def test_pixels_across_the_field_match_the_oracle(field, fp64):
    rng = np.random.default_rng(7)
    idx = np.concatenate([[0, 1, 2, NPIX // 2, NPIX - 2, NPIX - 1],
                          np.sort(rng.choice(NPIX, 506, replace=False))])
    cube = field["stokes"][field["torch"].from_numpy(idx).to(field["dev"])].cpu().numpy()
    lam = O.grid(*GRID)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=O.gauss_psf(60.0, lam))
    o = O.invert(s, cube, GUESS)
    g = {k: v[idx] for k, v in fp64.items()}
    # the tolerance of the 64 x 64 field of the same sampling, noise and
    # model distribution (tests/cases.py, l30_config1)
    tol = cases.tolerance("l30_config1", "fp64")
    f = (np.abs(g["model"] - o["model"]) <= tol).mean(axis=0)
    say("full field sample: fraction within tolerance %s; trips identical on %.4f"
        % (np.round(f, 4), (g["iters"] == o["iters"]).mean()))
    assert np.all(f >= cases.FRAC), f


def test_host_buffers_give_the_device_bits(field, fp64):
    """pm_invert on host buffers (the field in chunks over the context's
    streams, uploads and copy-backs overlapping other chunks' kernels),
    into pinned and into pageable result arrays: the device-pointer
    path's bits on every pixel."""
    from pmilos.serving import _pinned
    host = field["stokes"].cpu().numpy()
    for alloc in (_pinned, np.empty):
        say("case full field host buffers, %s results" % alloc.__name__)
        got = field["ctx"].invert(host, out=pmilos.result_buffers(NPIX, alloc=alloc))
        for k in fp64:
            assert np.array_equal(got[k], fp64[k]), k
    # through the C ABI with the optional outputs left out, as a C caller
    # of pm_invert does
    say("case full field host buffers, no accepted / fit")
    o = pmilos.result_buffers(NPIX)
    rc = pmilos.lib().pm_invert(field["ctx"].h, host.ctypes.data, NPIX, o["model"].ctypes.data,
                                o["chi2"].ctypes.data, o["iters"].ctypes.data, None, None)
    assert rc == 0, pmilos.lib().pm_strerror(rc)
    for k in ("model", "chi2", "iters"):
        assert np.array_equal(o[k], fp64[k]), k


def test_bits_do_not_depend_on_order_or_workgroup(field, fp64):
    """A second pass with the other workgroup size, and a pass over the
    field with its pixels in another order: the same bits per pixel."""
    torch, ctx = field["torch"], field["ctx"]
    (shape, n) = ctx.variant()
    shapes = {}
    for i in range(n):
        ctx.set_variant(i)
        shapes[ctx.variant()[0]] = i
    twin = [i for s, i in shapes.items() if s[:2] == shape[:2] and s != shape]
    assert twin, shapes
    ctx.set_variant(twin[0])
    again = _run(field, ctx, field["stokes"])
    ctx.set_variant(shapes[shape])
    for k in fp64:
        assert np.array_equal(again[k], fp64[k]), k
    perm = np.random.default_rng(3).permutation(NPIX)
    permuted = field["stokes"][torch.from_numpy(perm).to(field["dev"])].contiguous()
    shuffled = _run(field, ctx, permuted)
    for k in fp64:
        assert np.array_equal(shuffled[k], fp64[k][perm]), k


def test_mixed_form_within_tolerance_of_fp64_on_995_of_pixels(field, fp64):
    """The mixed form against the FP64 form over all 1,048,576 pixels, at
    the mixed tolerance of the same models and noise (tests/cases.py,
    mix_s1e-3), on at least 99.5% of the pixels per parameter."""
    say("case full field mixed npix=%d" % NPIX)
    cm = pmilos.Context(FE6173, TRIPLET, field["lam"], psf=field["G"], guess=GUESS,
                        noise=1e-6, miter=50, precision=pmilos.PM_MIXED)
    mx = _run(field, cm, field["stokes"])
    d = np.abs(mx["model"] - fp64["model"])
    frac = (d <= cases.tolerance("mix_s1e-3", "mixed")).mean(axis=0)
    say("full field mixed vs fp64: fraction within tolerance %s" % np.round(frac, 4))
    say("full field mixed: mean iters %.3f (fp64 %.3f), identical on %.4f"
        % (mx["iters"].mean(), fp64["iters"].mean(), (mx["iters"] == fp64["iters"]).mean()))
    assert np.all(frac >= cases.FRAC), frac
