"""GPU tests of the kernel variants, the autotuner and the one-launch batch
(pm_ctx_autotune, pm_ctx_set_variant, pm_graph_*).  Every variant of a
sampling inverts to the oracle's results within the FP64 tolerances; a
graph of several inversions gives, on every replay, the bits the same
inversions give one by one."""
import sys

import numpy as np
import pytest
import torch

import cases
import oracle as O
import pmilos
from fixtures import FE6173, FE6302, GUESS, TRIPLET, add_noise, draw_models, oracle_cube

pytestmark = pytest.mark.gpu

GRIDS = {6: (FE6173, -172.5, 69, 172.5), 14: (FE6173, -260, 40, 260),
         30: (FE6173, -290, 20, 290), 49: (FE6173, -600, 25, 600),
         112: (FE6302, -1193.25, 21.5, 1193.25)}
TOL64 = cases.FLOOR[:9]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(L, mac=False, precision=pmilos.PM_FP64):
    g = GRIDS[L]
    lam = O.grid(*g)
    G = O.gauss_psf(60.0, lam)
    guess, fix = GUESS.copy(), [1] * 9 + [0, 0]
    if mac:
        guess[9], fix = 0.5, [1] * 10 + [0]
    s = O.make_setup(g[0], TRIPLET, lam, psf=G, fix=fix, guess_mac=guess[9])
    ctx = pmilos.Context(g[0], TRIPLET, lam, psf=G, guess=guess, fix=fix, precision=precision)
    return s, ctx, guess


def _err_rel(a, b, axis):
    scale = np.maximum(np.abs(b).max(axis=axis, keepdims=True), 1e-5)
    return (np.abs(a - b) / scale).max()


@pytest.mark.parametrize("L", [6, 14, 30, 49, 112])
@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_every_variant_synthesis_matches_oracle(L, precision):
    """The synthesis kernel follows the selected lane mapping: spectra and
    response functions of every variant against the oracle."""
    s, ctx, guess = _pair(L, precision=precision)
    models = draw_models(24, seed=740 + L)
    ref = [O.synth(s, m, rf=True) for m in models]
    _, nv = ctx.variant()
    for k in range(nv):
        ctx.set_variant(k)
        shape, _ = ctx.variant()
        say("case variant synth L=%d precision=%d index=%d shape=%s" % (L, precision, k, shape))
        sp, rf = ctx.synth(models, rf=True)
        e_rf = max(_err_rel(rf[i], r[1], (0, 2)) for i, r in enumerate(ref))
        if precision == pmilos.PM_FP64:
            e_sp = max(_err_rel(sp[i], r[0], -1) for i, r in enumerate(ref))
            say("  spectra %.3g  rf %.3g" % (e_sp, e_rf))
            assert e_sp < 1e-10 and e_rf < 1e-10
        else:
            # FP32 spectra: absolute, against the intensity scale
            e_sp = max(np.abs(sp[i] - r[0]).max() / np.abs(r[0][0]).max()
                       for i, r in enumerate(ref))
            say("  spectra %.3g  rf %.3g" % (e_sp, e_rf))
            assert e_sp < 1e-6 and e_rf < 2e-3


@pytest.mark.parametrize("L", [6, 14, 30, 49, 112])
@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_every_variant_matches_oracle(L, precision):
    """Every kernel shape of a sampling inverts the case's cube (2048-4096
    pixels) to within the form's stated tolerance of the oracle on >= 99.5%
    of pixels per parameter."""
    name = "var_l%d" % L
    c = cases.get(name)
    ctx = c.context(pmilos, precision=precision)
    form = "fp64" if precision == pmilos.PM_FP64 else "mixed"
    _, nv = ctx.variant()
    for k in range(nv):
        ctx.set_variant(k)
        shape, _ = ctx.variant()
        say("case variant %s %s index=%d shape=%s" % (name, form, k, shape))
        cases.compare(ctx.invert(c.cube), name, form, say, cube=c.cube if k == 0 else None)


@pytest.mark.parametrize("L", [30, 49])
def test_fp64_wide_shapes_at_small_samplings(L):
    """The FP64 shapes offered for more than 64 samples -- two samples per
    lane (64 x 2) and a pixel over two waves (128 x 1, one sample per lane)
    -- forced onto 30 and 49 samples, where one sample per lane in one wave
    is the default: their inversions match the default shape's at the FP64
    floor tolerance (rounding-level agreement on every parameter) on
    >= 99.5% of pixels, and the oracle at the case's tolerance."""
    name = "var_l%d" % L
    c = cases.get(name)
    ctx = c.context(pmilos)
    base = ctx.invert(c.cube)
    for shape in ((64, 2, 4), (64, 2, 2), (128, 1, 1)):
        ctx.set_shape(*shape)
        assert ctx.variant()[0] == shape
        say("case fp64 shape %dx%dx%d %s" % (shape + (name,)))
        g = ctx.invert(c.cube)
        f = (np.abs(g["model"] - base["model"]) <= cases.FLOOR).mean(axis=0)
        say("  within the floor of the default shape %s" % np.round(f, 4))
        assert np.all(f >= cases.FRAC), f
        cases.compare(g, name, "fp64", say)
    # the shapes the library does not compile are refused
    with pytest.raises(pmilos.PmError):
        ctx.set_shape(16, 2, 8)


def test_workgroup_size_does_not_change_bits():
    """Variants that differ only in pixels per workgroup give equal bits."""
    s, ctx, guess = _pair(30)
    truths = draw_models(2000, seed=710)
    cube = add_noise(ctx.synth(truths), truths, 1e-3, seed=711)
    ctx.set_variant(0)
    a = ctx.invert(cube)
    ctx.set_variant(1)
    assert ctx.variant()[0][:2] == (32, 1)
    b = ctx.invert(cube)
    for k in ("model", "chi2", "iters", "accepted"):
        assert np.array_equal(a[k], b[k]), k


def test_autotune_picks_a_variant_and_keeps_results():
    say("case autotune L=30")
    dev = torch.device("cuda:0")
    s, ctx, guess = _pair(30)
    truths = draw_models(16384, seed=720)
    cube = add_noise(ctx.synth(truths), truths, 1e-3, seed=721)
    d = torch.from_numpy(cube).to(dev)
    chosen = ctx.autotune_ptrs(d.data_ptr(), 16384, torch.cuda.current_stream().cuda_stream)
    say("  chosen %s" % (chosen,))
    assert chosen == ctx.variant()[0]
    assert chosen[0] * chosen[1] >= 30 and chosen[2] >= 1
    o = O.invert(s, cube[:512], guess)
    g = ctx.invert(cube[:512])
    ok = np.abs(g["model"][:, :9] - o["model"][:, :9]) <= TOL64
    assert np.all(ok.mean(axis=0) >= 0.995)


def _dev_outputs(n, L, dev):
    return dict(m=torch.empty((n, 11), dtype=torch.float64, device=dev),
                c=torch.empty(n, dtype=torch.float32, device=dev),
                i=torch.empty(n, dtype=torch.int32, device=dev),
                a=torch.empty(n, dtype=torch.int32, device=dev),
                f=torch.empty((n, 4, L), dtype=torch.float32, device=dev))


@pytest.mark.parametrize("share", [0.0, 1.0])
def test_graph_batch_equals_separate_inversions(share):
    say("case graph batch share=%s" % share)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    specs = [(6, False, 3000, 730), (30, False, 4000, 731), (112, True, 600, 732)]
    items, eager, outs = [], [], []
    for L, mac, n, seed in specs:
        _, ctx, _ = _pair(L, mac)
        m = draw_models(n, seed, mac_range=mac)
        cube = add_noise(ctx.synth(m), m, 1e-3, seed + 1)
        eager.append(ctx.invert(cube, want_fit=True))
        d = torch.from_numpy(cube).to(dev)
        o = _dev_outputs(n, L, dev)
        outs.append((ctx, d, o))
        items.append(dict(ctx=ctx, d_stokes=d.data_ptr(), npix=n, d_models=o["m"].data_ptr(),
                          d_chi2=o["c"].data_ptr(), d_iters=o["i"].data_ptr(),
                          d_accepted=o["a"].data_ptr(), d_fit=o["f"].data_ptr()))
    g = pmilos.Graph(items, share_by_work=share)
    try:
        for rep in range(3):
            for _, _, o in outs:
                for t in o.values():
                    t.fill_(-1)
            torch.cuda.synchronize()
            g.launch(st.cuda_stream)
            torch.cuda.synchronize()
            for (ctx, d, o), e in zip(outs, eager):
                assert np.array_equal(o["m"].cpu().numpy(), e["model"]), rep
                assert np.array_equal(o["c"].cpu().numpy(), e["chi2"]), rep
                assert np.array_equal(o["i"].cpu().numpy(), e["iters"]), rep
                assert np.array_equal(o["a"].cpu().numpy(), e["accepted"]), rep
                assert np.array_equal(o["f"].cpu().numpy(), e["fit"]), rep
            # eager work between replays
            _ = (outs[0][1] * 2).sum().item()
    finally:
        g.close()
