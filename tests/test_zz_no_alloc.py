"""pm_invert_device_noalloc (the entry point that allocates nothing) against
pm_invert_device through the C ABI: the same kernels, so the same bits, on
two calls in a row, the second with a smaller input; its size limit is an
error, not a launch."""
import ctypes
import sys

import numpy as np
import pytest

import pmilos
from fixtures import FE6173, GUESS, TRIPLET, add_noise, draw_models

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _outs(torch, n, L, dev):
    return dict(model=torch.full((n, 11), -7.0, dtype=torch.float64, device=dev),
                chi2=torch.full((n,), -7.0, dtype=torch.float32, device=dev),
                iters=torch.full((n,), -7, dtype=torch.int32, device=dev),
                accepted=torch.full((n,), -7, dtype=torch.int32, device=dev),
                fit=torch.full((n, 4, L), -7.0, dtype=torch.float32, device=dev))


def _ptrs(o):
    return (o["model"].data_ptr(), o["chi2"].data_ptr(), o["iters"].data_ptr(),
            o["accepted"].data_ptr(), o["fit"].data_ptr())


@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
def test_noalloc_matches_invert_device(precision):
    import torch
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    lam = pmilos.grid(FE6173, -290, 20, 290)
    L = len(lam)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=pmilos.psf_gaussian(60.0, lam), guess=GUESS,
                         precision=precision)
    truths = draw_models(3000, seed=910)
    cube = add_noise(ctx.synth(truths), truths, 1e-3, seed=911)
    for call, n in enumerate((3000, 1100)):
        say("case noalloc precision=%d call=%d npix=%d" % (precision, call, n))
        d = torch.from_numpy(cube[:n]).to(dev)
        ref, got = _outs(torch, n, L, dev), _outs(torch, n, L, dev)
        ctx.invert_ptrs(d.data_ptr(), n, *_ptrs(ref), stream=st.cuda_stream)
        ctx.invert_ptrs_noalloc(d.data_ptr(), n, *_ptrs(got), stream=st.cuda_stream)
        torch.cuda.synchronize()
        for k in ref:
            a, b = ref[k].cpu().numpy(), got[k].cpu().numpy()
            assert np.array_equal(a, b), (call, k)
        it = got["iters"].cpu().numpy()
        assert it.min() >= 1 and it.max() <= 50


def test_noalloc_size_limit_and_arguments():
    """Sizes beyond PM_NOALLOC_MAX_NPIX and bad arguments are refused on
    the host, before any launch."""
    lam = pmilos.grid(FE6173, -290, 20, 290)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, guess=GUESS)
    L = pmilos.lib()
    fake = ctypes.c_void_p(0x1000)
    rc = L.pm_invert_device_noalloc(ctx.h, fake, (1 << 30) + 1, fake, fake, fake, None, None,
                                    None)
    assert rc == -34
    assert L.pm_strerror(rc).decode() == "size above the entry point's maximum"
    assert L.pm_invert_device_noalloc(None, fake, 10, fake, fake, fake, None, None, None) == -22
    assert L.pm_invert_device_noalloc(ctx.h, fake, -1, fake, fake, fake, None, None, None) == -22
    assert L.pm_invert_device_noalloc(ctx.h, None, 0, None, None, None, None, None, None) == 0
