"""GPU tests of the serving loop (pm_server_*, pm_submit, pm_poll, pm_wait):
scans of different sizes and samplings submitted together give exactly the
bits each gives alone through pm_invert; admission respects the HBM budget;
two waits on one ticket retire it once; a small scan submitted behind a
large one finishes first."""
import ctypes
import sys
import threading

import numpy as np
import pytest
import torch  # noqa: F401  (its HIP runtime loads before the library's)

import oracle as O
import pmilos
from fixtures import FE6173, FE6302, GUESS, TRIPLET, add_noise, draw_models

pytestmark = pytest.mark.gpu

GRIDS = {6: (FE6173, -172.5, 69, 172.5), 14: (FE6173, -260, 40, 260),
         30: (FE6173, -290, 20, 290), 112: (FE6302, -1193.25, 21.5, 1193.25)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ctx(L, precision=pmilos.PM_FP64, mac=False):
    g = GRIDS[L]
    lam = O.grid(*g)
    G = O.gauss_psf(60.0, lam)
    guess = GUESS.copy()
    fix = [1] * 9 + [0, 0]
    if mac:
        guess[9] = 0.5
        fix = [1] * 10 + [0]
    return pmilos.Context(g[0], TRIPLET, lam, psf=G, guess=guess, fix=fix, precision=precision)


def _cube(ctx, n, seed, mac=False):
    m = draw_models(n, seed, mac_range=mac)
    return add_noise(ctx.synth(m), m, 1e-3, seed + 1)


def _same(a, b):
    for k in ("model", "chi2", "iters", "accepted"):
        assert np.array_equal(a[k], b[k]), k
    if a.get("fit") is not None:
        assert np.array_equal(a["fit"], b["fit"])


def test_mix_equals_alone():
    say("case serving mix vs alone")
    ctxs = [_ctx(30), _ctx(6), _ctx(112, pmilos.PM_MIXED, mac=True), _ctx(14, pmilos.PM_MIXED),
            _ctx(30, pmilos.PM_MIXED)]
    sizes = [5000, 900, 300, 2500, 4096]
    cubes = [_cube(c, n, 100 + i, mac=(i == 2)) for i, (c, n) in enumerate(zip(ctxs, sizes))]
    alone = [c.invert(q, want_fit=True) for c, q in zip(ctxs, cubes)]
    srv = pmilos.Server(streams=3, chunk_px=1024)
    try:
        # every scan twice, interleaved, so chunks of all of them share the GPU
        tickets = []
        for rep in range(2):
            for i, (c, q) in enumerate(zip(ctxs, cubes)):
                tickets.append((i, srv.submit(c, q, want_fit=(rep == 0))))
        for i, t in tickets:
            out, lat = srv.wait(t)
            assert lat > 0
            ref = dict(alone[i])
            if out["fit"] is None:
                ref["fit"] = None
            _same(out, ref)
    finally:
        srv.close()


def test_results_into_caller_buffers():
    """submit(out=...): results land in the caller's result set, pinned or
    pageable, reused for a second scan; a set of the wrong size is
    refused before anything is queued."""
    say("case serving caller buffers")
    from pmilos.serving import _pinned
    ctx = _ctx(30, pmilos.PM_MIXED)
    n = 3000
    cubes = [_cube(ctx, n, 300 + i) for i in range(2)]
    alone = [ctx.invert(q) for q in cubes]
    srv = pmilos.Server(streams=2, chunk_px=1024)
    try:
        for alloc in (_pinned, np.empty):
            out = pmilos.Server.result_buffers(n, alloc=alloc)
            for k in ("model", "chi2"):
                out[k][...] = np.nan
            for q, ref in zip(cubes, alone):
                got, _ = srv.wait(srv.submit(ctx, q, out=out))
                assert got["model"] is out["model"]
                _same(got, ref)
        with pytest.raises(ValueError):
            srv.submit(ctx, cubes[0], out=pmilos.Server.result_buffers(n - 1))
    finally:
        srv.close()


def test_budget_admission_and_errors():
    say("case serving budget")
    ctx = _ctx(30)
    cubes = [_cube(ctx, 3000, 200 + i) for i in range(4)]
    alone = [ctx.invert(q) for q in cubes]
    one = 3000 * (4 * 30 * 4 + 88 + 12) + 8 * 256
    srv = pmilos.Server(hbm_budget=int(1.5 * one), streams=2, chunk_px=1024)
    try:
        ts = [srv.submit(ctx, q) for q in cubes]
        assert all(isinstance(t, int) for t in ts)
        for t, ref in zip(ts, alone):
            out, _ = srv.wait(t)
            _same(out, ref)
        with pytest.raises(pmilos.PmError):
            srv.submit(ctx, _cube(ctx, 6000, 300))
        t = srv.submit(ctx, cubes[0][:0])
        out, lat = srv.wait(t)
        assert out["model"].shape == (0, 11)
        with pytest.raises(pmilos.PmError):
            srv.poll(t)
    finally:
        srv.close()


def test_two_waits_on_one_ticket():
    """Two threads in pm_wait on one ticket: one collects the scan with its
    results, the other gets PM_EINVAL (-22) instead of the freed request."""
    say("case serving two waits on one ticket")
    ctx = _ctx(30)
    q = _cube(ctx, 20000, 210)
    alone = ctx.invert(q)
    srv = pmilos.Server()
    try:
        t = srv.submit(ctx, q)
        rcs = []

        def wait_raw():
            lat = ctypes.c_double(0)
            rcs.append(pmilos.lib().pm_wait(srv.h, t, ctypes.byref(lat)))

        th = [threading.Thread(target=wait_raw) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join(timeout=60)
        assert not any(x.is_alive() for x in th)
        assert sorted(rcs) == [-22, 0]
        _, _, out = srv._pending.pop(t)
        _same(out, alone)
        # the binding's own wait on the retired ticket raises PmError
        with pytest.raises(pmilos.PmError):
            srv.wait(t)
    finally:
        srv.close()


def test_small_scan_overtakes_large():
    say("case serving overtake")
    big_ctx, small_ctx = _ctx(30), _ctx(6)
    big = _cube(big_ctx, 200000, 400)
    small = _cube(small_ctx, 1024, 401)
    srv = pmilos.Server(streams=2, chunk_px=8192)
    try:
        tb = srv.submit(big_ctx, big)
        ts = srv.submit(small_ctx, small)
        seen_small_first = False
        while not srv.poll(tb):
            if srv.poll(ts):
                seen_small_first = True
                break
        out_s, lat_s = srv.wait(ts)
        out_b, lat_b = srv.wait(tb)
        say("latency small %.4f s, large %.4f s" % (lat_s, lat_b))
        assert seen_small_first and lat_s < lat_b
        _same(out_s, small_ctx.invert(small))
    finally:
        srv.close()
