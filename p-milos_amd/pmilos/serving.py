"""Streaming request loop on one GPU: scans arrive as a Poisson process and
go through pmilos.Server, which admits them by HBM budget and runs them
concurrently.  Reports pixels/s and p50/p99 per-scan latency (arrival to
results in host memory).

Scan mix: sizes {32x32, 256x256, 1024x1024, 2048x2048} pixels and samplings
{6, 14, 30, 112} wavelengths (Fe I 6173.3356 A at -172.5..172.5 @ 69 mA,
-260..260 @ 40 mA, -290..290 @ 20 mA; Fe I 6302.4936 A at -1193.25..1193.25
@ 21.5 mA with v_mac free, 10 parameters), Gaussian PSF of 60 mA, models
drawn as the benchmark field's, noise 1e-3 Ic, arrivals from seed 50.  Scan
k is synthesised from seed 500 + k (models and noise, made on the GPU) for
the first few scans of each (size, sampling) kind -- four of each kind up to
256^2 pixels, two of the larger ones -- and later scans of a kind reuse
those cubes in turn (a 2048^2 x 112 cube is 7.5 GB of host memory).  The
cubes are in pinned host memory, and so are the results: each scan's
result set comes from a pool of pinned sets of its size, returned to the
pool once the scan has been collected (a serving client's reuse of its
buffers; fresh pageable result arrays cost page faults on every copy).

    python -m pmilos.serving [--scans N] [--util U ...] [--precision mixed|fp64]
    python -m pmilos.serving --per-sampling 25 --small 50   (the benchmark's mix)
"""
import argparse
import json
import sys
import threading
import time

import numpy as np

import pmilos

FE6173, FE6302 = 6173.3356, 6302.4936
TRIPLET = (2, 1, 1.0, 2, 2, 0.0)
GUESS = np.array([20, 1100, 0.2, 0.03, 0.05, 120, 150, 0.35, 0.5, 0, 1.0])
SAMPLINGS = {6: (FE6173, -172.5, 69, 172.5), 14: (FE6173, -260, 40, 260),
             30: (FE6173, -290, 20, 290), 112: (FE6302, -1193.25, 21.5, 1193.25)}
SIZES = (32, 256, 1024, 2048)


def draw_models(n, seed, mac):
    rng = np.random.default_rng(seed)
    B = rng.uniform(0, 2500, n)
    gm = rng.uniform(0, 180, n)
    az = rng.uniform(0, 180, n)
    vl = rng.normal(0, 1.5, n)
    eta = np.exp(rng.uniform(0, np.log(30), n))
    dop = rng.uniform(0.020, 0.050, n)
    aa = rng.uniform(0, 0.5, n)
    S0 = rng.uniform(0.1, 0.4, n)
    S1 = 1 - S0 + rng.uniform(-0.05, 0.05, n)
    vmac = rng.uniform(0, 2, n) if mac else np.zeros(n)
    return np.stack([eta, B, vl, dop, aa, gm, az, S0, S1, vmac, np.ones(n)], 1)


def contexts(precision):
    """One resident context per sampling (line, grid, PSF, Zeeman pattern)."""
    out = {}
    for L, g in SAMPLINGS.items():
        lam = pmilos.grid(*g)
        G = pmilos.psf_gaussian(60.0, lam)
        guess, fix = GUESS.copy(), [1] * 9 + [0, 0]
        if L == 112:
            guess[9], fix = 0.5, [1] * 10 + [0]
        out[L] = pmilos.Context(g[0], TRIPLET, lam, psf=G, guess=guess, fix=fix, miter=50,
                                precision=precision)
    return out


def make_cube(ctx, npix, seed, mac):
    """Synthetic float32 cube [npix][4][L] in pinned host memory (where an
    acquisition pipeline's DMA buffers hold a scan; the server's uploads then
    run at the link's rate instead of through staging copies): the
    library's FP64 synthesis of seeded models plus seeded noise, made on
    the GPU."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    L = ctx.L
    m = torch.from_numpy(draw_models(npix, seed, mac)).to(dev)
    spec = torch.empty((npix, 4, L), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream()
    ctx.synth_ptrs(m.data_ptr(), npix, spec.data_ptr(), 0, st.cuda_stream)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    ic = (m[:, 7] + m[:, 8]).view(-1, 1, 1)
    spec += torch.randn(spec.shape, dtype=torch.float64, device=dev, generator=gen) * 1e-3 * ic
    host = torch.empty((npix, 4, L), dtype=torch.float32, pin_memory=True)
    host.copy_(spec.to(torch.float32))
    del spec, m
    return host.numpy()             # the array keeps the pinned tensor alive


def scan_cubes(ctxs, kinds, distinct_small=4, distinct_large=2):
    """The cube of every scan (host float32 arrays, shared between scans of
    a kind past its first distinct_* ones): scan k's own cube is made from
    seed 500 + k."""
    made = {}
    count = {}
    out = []
    for k, (size, L) in enumerate(kinds):
        lst = made.setdefault((size, L), [])
        cap = distinct_small if size <= 256 else distinct_large
        if len(lst) < cap:
            lst.append(make_cube(ctxs[L], size * size, 500 + k, L == 112))
        j = count.get((size, L), 0)
        count[(size, L)] = j + 1
        out.append(lst[j % len(lst)])
    return out, made


def calibrate(ctxs, made):
    """Pixels/s of each context alone on the first cube of its largest kind
    (a bounded slice, results into a pinned set as in the run), used to size
    the arrival rate for a target utilisation."""
    rate = {}
    for L, ctx in ctxs.items():
        size = max(s for s, l in made if l == L)
        cube = made[(size, L)][0]
        n = min(len(cube), 262144 if L <= 30 else 65536)
        out = pmilos.result_buffers(n, alloc=_pinned)
        ctx.invert(cube[:n], out=out)
        t = time.perf_counter()
        ctx.invert(cube[:n], out=out)
        rate[L] = n / (time.perf_counter() - t)
    return rate


def _pinned(shape, dtype):
    import torch
    t = torch.empty(shape, dtype=getattr(torch, np.dtype(dtype).name), pin_memory=True)
    return t.numpy()


class ResultPool:
    """Pinned result sets by scan size, reused once collected; `ahead` sets
    of every size in `sizes` are made up front."""

    def __init__(self, sizes=(), ahead=2):
        self.free = {}
        self.lock = threading.Lock()
        for n in set(sizes):
            self.free[n] = [pmilos.Server.result_buffers(n, alloc=_pinned) for _ in range(ahead)]

    def take(self, n):
        with self.lock:
            lst = self.free.get(n)
            if lst:
                return lst.pop()
        return pmilos.Server.result_buffers(n, alloc=_pinned)

    def give(self, n, out):
        with self.lock:
            self.free.setdefault(n, []).append(out)


def run(server, ctxs, cubes, kinds, rate_hz, seed=50, progress=False):
    """Submits len(kinds) scans (scan k: cubes[k]) at Poisson arrival times;
    returns per-scan (kind, npix, latency) and the wall time from first
    arrival to last completion.  Results go into pinned sets from a
    ResultPool.  progress: a line on stderr every 50 completed scans."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(kinds))
    tickets = []
    done = threading.Event()
    pool = ResultPool([k[0] * k[0] for k in kinds])

    def submitter():
        t_next = time.perf_counter()
        for k, cube, gap in zip(kinds, cubes, gaps):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            size, L = k
            tickets.append((k, server.submit(ctxs[L], cube, out=pool.take(size * size))))
        done.set()

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    res = []
    i = 0
    while True:
        if i < len(tickets):
            k, t = tickets[i]
            out, lat = server.wait(t)
            pool.give(k[0] * k[0], out)
            res.append((k, k[0] * k[0], lat))
            i += 1
            if progress and i % 50 == 0:
                print("served %d/%d scans, %.1f s" % (i, len(kinds), time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        elif done.is_set() and i >= len(tickets):
            break
        else:
            time.sleep(0.001)
    th.join()
    return res, time.perf_counter() - t0


def summarize(res, wall):
    lat = np.array([r[2] for r in res])
    px = sum(r[1] for r in res)
    out = {"scans": len(res), "pixels_per_s": round(px / wall, 1),
           "p50_s": round(float(np.percentile(lat, 50)), 4),
           "p99_s": round(float(np.percentile(lat, 99)), 4), "per_kind": {}}
    for k in sorted(set(r[0] for r in res)):
        l = np.array([r[2] for r in res if r[0] == k])
        out["per_kind"]["%dx%d_L%d" % (k[0], k[0], k[1])] = {
            "n": len(l), "p50_s": round(float(np.percentile(l, 50)), 4),
            "p99_s": round(float(np.percentile(l, 99)), 4)}
    return out


def bench_kinds(per_sampling=25, small=50, seed=50, samplings=tuple(SAMPLINGS)):
    """The benchmark's scan list: per_sampling scans of 1024^2 pixels at
    every sampling (so that each sampling's p99 rests on that many scans)
    and `small` scans of 32^2 and 256^2 pixels at random samplings, in a
    seeded random order."""
    rng = np.random.default_rng(seed + 1)
    kinds = [(1024, L) for L in samplings for _ in range(per_sampling)]
    kinds += [(int(rng.choice([32, 256])), int(rng.choice(samplings))) for _ in range(small)]
    return [kinds[i] for i in rng.permutation(len(kinds))]


def serve_bench(precision=pmilos.PM_MIXED, scans=48, util=0.5, seed=50, sizes=SIZES,
                samplings=tuple(SAMPLINGS), progress=False, kinds=None, chunk_px=0):
    """A streaming run: `scans` scans of random kinds from sizes x samplings
    (or the given list of (size, L) kinds) arriving as a Poisson process at
    the rate that loads the GPU to `util` of its calibrated capacity.
    chunk_px: the server's chunk size (0: its default)."""
    ctxs = contexts(precision)
    if kinds is None:
        kinds_all = [(s, L) for s in sizes for L in samplings]
        rng = np.random.default_rng(seed)
        kinds = [kinds_all[j] for j in rng.integers(0, len(kinds_all), scans)]
    t0 = time.perf_counter()
    cubes, made = scan_cubes(ctxs, kinds)
    if progress:
        print("%d distinct cubes ready, %.1f s" % (sum(len(v) for v in made.values()),
                                                   time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    rate = calibrate(ctxs, made)
    if progress:
        print("calibrated: %s pixels/s alone" % {L: round(r) for L, r in rate.items()},
              file=sys.stderr, flush=True)
    mean_service = float(np.mean([s * s / rate[L] for s, L in kinds]))
    arrival = util / mean_service
    srv = pmilos.Server(chunk_px=chunk_px)
    try:
        k0 = min(made)                                   # first use of its streams
        srv.wait(srv.submit(ctxs[k0[1]], made[k0][0][:1024]))
        res, wall = run(srv, ctxs, cubes, kinds, arrival, seed, progress)
    finally:
        srv.close()
    out = summarize(res, wall)
    out.update({"arrival_rate_hz": round(arrival, 4), "target_utilisation": util,
                "mean_service_s": round(mean_service, 4),
                "alone_pixels_per_s": {str(L): round(r, 1) for L, r in rate.items()},
                "distinct_cubes": sum(len(v) for v in made.values()),
                "chunk_px": chunk_px or 65536,
                "precision": "mixed" if precision == pmilos.PM_MIXED else "fp64"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scans", type=int, default=48)
    ap.add_argument("--util", type=float, nargs="+", default=[0.5])
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--seed", type=int, default=50)
    ap.add_argument("--per-sampling", type=int, default=0,
                    help="the benchmark's mix: this many 1024^2 scans per sampling")
    ap.add_argument("--small", type=int, default=50,
                    help="with --per-sampling: this many 32^2/256^2 scans besides")
    ap.add_argument("--chunk-px", type=int, default=0,
                    help="the server's chunk size in pixels of the 30-sample workload "
                         "(0: its default, 65536)")
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    prec = pmilos.PM_MIXED if args.precision == "mixed" else pmilos.PM_FP64
    kinds = bench_kinds(args.per_sampling, args.small, args.seed) if args.per_sampling else None
    for u in args.util:
        print(json.dumps(serve_bench(prec, args.scans, u, args.seed, progress=True, kinds=kinds,
                                     chunk_px=args.chunk_px)), flush=True)


if __name__ == "__main__":
    main()
