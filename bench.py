"""Benchmark: inverted pixels/s of the batched Milne-Eddington inversion on
one MI355X.

Workload: a 1024x1024-pixel field of Fe I 6173.3356 A (5P1-5D0, g_eff 2.5)
sampled at 30 wavelengths spaced 20 mA (-290..290 mA), Gaussian PSF of 60 mA
FWHM, models drawn from seed 2 (B ~ U(0,2500) G, gamma, azimuth ~ U(0,180),
vlos ~ N(0,1.5) km/s, eta0 ~ logU(1,30), Doppler width ~ U(20,50) mA,
damping ~ U(0,0.5), S0 ~ U(0.1,0.4), S1 = 1-S0+U(-0.05,0.05)), synthesised
by the library's own FP64 synthesis, noise sigma = 1e-3 Ic (seed 32), stored
as float32.  Inversion: 9 free parameters, <= 50 LM iterations, the run/
guess.mod initial model plus classical estimates per pixel.

One step = one inversion pass over the whole cube, inputs resident in HBM.
Beside it: --steps passes of each form through pm_invert from the cube in
pinned host memory to the results in pinned host memory (PCIe copies
included, overlapped with the kernels by the library), and one pass of each form
with the logclambda damping rule of run/pmilos.mtrol (logclambda = 1).
The same field at noise sigma 1e-4 and 3e-3 Ic gives one more pass of each
form (the mixed form's error against the FP64 form and the change in
iteration count per sigma).  The uneven batch (three cubes of 6, 30 and 112
samples in one graph launch) follows, then a bounded streaming run
(pmilos.serving: Poisson arrivals at half the calibrated capacity of 25
scans of 1024^2 pixels at each of 6, 14, 30 and 112 wavelengths and 50 of
32^2 / 256^2 pixels, mixed form) gives p50/p99 scan latency per kind.  The
CPU baseline (the oracle's restatement, float and double builds, one
process per core) runs first, before the GPU is touched.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--npix N] [--serve-per-sampling S]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "p-milos_amd"))

METRIC = ("pixels/s (30 lambda, 9 free, <=50 it) FP32-mixed vs FP64; "
          "p50/p99 scan latency")
FE6173 = 6173.3356
TRIPLET = (2, 1, 1.0, 2, 2, 0.0)
GUESS = np.array([20, 1100, 0.2, 0.03, 0.05, 120, 150, 0.35, 0.5, 0, 1.0])
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per pixel of one pm_invert_kernel launch on this workload, from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script at
# 1048576 pixels (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the factors
# a plain 512 MiB streaming kernel reads in the same runs); the summaries
# are profiles/pm_invert_kernel_l30_1048576px_hbm.md
TRAFFIC_B_PER_PX = {"fp64": 591.63, "mixed": 595.67}
# lane FLOPs issued per pixel by the same launches (64 x the
# SQ_INSTS_VALU_FLOPS_FP64 / _FP32 counters: every lane slot of an issued
# instruction); the summary is profiles/pm_invert_kernel_l30_1048576px_valu.md
ISSUED_FLOP_PER_PX = {"fp64": {"fp64": 4599999, "fp32": 315},
                      "mixed": {"fp64": 570461, "fp32": 2725551}}
ALGO_B_PER_PX = 4 * 30 * 4 + 11 * 8 + 3 * 4     # Stokes in; model, chi2, iters, accepted out


def draw_models(n, seed):
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
    return np.stack([eta, B, vl, dop, aa, gm, az, S0, S1, np.zeros(n), np.ones(n)], 1)


def _cpu_worker(args):
    """One core: synthesise its block of the workload with the oracle, add the
    same-recipe noise, then time the inversion of the block by the oracle
    build `build`."""
    p0, p1, per_seed, build = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.use("liboracle.so")
    lam = O.grid(FE6173, -290, 20, 290)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=O.gauss_psf(60.0, lam))
    models = draw_models(p1, 2)[p0:p1]
    spec = np.stack([O.synth(s, m) for m in models])
    rng = np.random.default_rng(per_seed)
    ic = (models[:, 7] + models[:, 8])[:, None, None]
    stokes = (spec + rng.normal(0, 1, spec.shape) * 1e-3 * ic).astype(np.float32)
    O.use(build)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=O.gauss_psf(60.0, lam))
    t = time.perf_counter()
    O.invert(s, stokes, GUESS)
    return p1 - p0, time.perf_counter() - t


def cpu_cores():
    """The host CPUs this process may use, at most 16 (a one-GPU job's share
    of the GPU host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(per_core):
    """The oracle's restatement of the per-pixel path, built as the reference
    builds itself (-O3 -march=native), over equal pixel blocks, one process
    per core, like pmilos.c:1181-1205: the REAL = float build (the
    reference's default, liboracle_f32.so; value) and the REAL = double one
    (liboracle_o3.so).  Runs before the GPU is initialised (plain fork)."""
    cores = cpu_cores()
    n = per_core * cores
    out = {}
    for name, build in (("float", "liboracle_f32.so"), ("fp64", "liboracle_o3.so")):
        blocks = [(i * per_core, (i + 1) * per_core, 3200 + i, build) for i in range(cores)]
        with mp.get_context("fork").Pool(cores) as pool:
            res = pool.map(_cpu_worker, blocks)
        slowest = max(r[1] for r in res)
        out[name] = (round(n / slowest, 1), slowest)
    return dict(value=out["float"][0], unit="pixels/s", cores=cores, kind="port",
                per_core=round(out["float"][0] / cores, 1),
                fp64_build={"value": out["fp64"][0], "per_core": round(out["fp64"][0] / cores, 1)},
                sample="first %d pixels of the workload (seed-2 models, same noise recipe), "
                       "%d processes x %d pixels; REAL=float build (the reference's default) "
                       "slowest block %.1f s, REAL=double build %.1f s"
                       % (n, cores, per_core, out["float"][1], out["fp64"][1]))


def uneven_batch(args, torch, pmilos, precision):
    """Three cubes in one launch: 512x512 px at 6 wavelengths (69 mA),
    512x512 px at 30 (20 mA), 256x256 px of Fe I 6302.4936 A at 112 (21.5 mA)
    with v_mac ~ U(0,2) km/s free (10 parameters); model seeds 41/42/43,
    noise 1e-3 Ic.  Each sampling's kernel shape is picked by the autotuner
    on its whole cube; the batch is one HIP graph launch."""
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream()
    specs = [(512 * 512, (FE6173, -172.5, 69, 172.5), False, 41),
             (512 * 512, (FE6173, -290, 20, 290), False, 42),
             (256 * 256, (6302.4936, -1193.25, 21.5, 1193.25), True, 43)]
    items, keep, shapes = [], [], []
    total = 0
    for n, g, mac, seed in specs:
        lam = pmilos.grid(*g)
        L = len(lam)
        guess, fix = GUESS.copy(), [1] * 9 + [0, 0]
        if mac:
            guess[9], fix = 0.5, [1] * 10 + [0]
        ctx = pmilos.Context(g[0], TRIPLET, lam, psf=pmilos.psf_gaussian(60.0, lam), guess=guess,
                             fix=fix, noise=1e-6, miter=50, precision=precision)
        m = draw_models(n, seed)
        if mac:
            m[:, 9] = np.random.default_rng(seed + 1000).uniform(0, 2, n)
        mt = torch.from_numpy(m).to(dev)
        spec = torch.empty((n, 4, L), dtype=torch.float64, device=dev)
        ctx.synth_ptrs(mt.data_ptr(), n, spec.data_ptr(), 0, st.cuda_stream)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + 100)
        ic = (mt[:, 7] + mt[:, 8]).view(-1, 1, 1)
        cube = (spec + torch.randn(spec.shape, dtype=torch.float64, device=dev, generator=gen)
                * 1e-3 * ic).to(torch.float32).contiguous()
        del spec
        o = [torch.empty((n, 11), dtype=torch.float64, device=dev),
             torch.empty(n, dtype=torch.float32, device=dev),
             torch.empty(n, dtype=torch.int32, device=dev),
             torch.empty(n, dtype=torch.int32, device=dev)]
        shapes.append({"nlambda": L, "npix": n, "nterms": ctx.NT,
                       "lanes_per_pixel_x_samples_per_lane_x_pixels_per_workgroup":
                           list(ctx.autotune_ptrs(cube.data_ptr(), n, st.cuda_stream))})
        keep.append((ctx, cube, o))
        items.append(dict(ctx=ctx, d_stokes=cube.data_ptr(), npix=n, d_models=o[0].data_ptr(),
                          d_chi2=o[1].data_ptr(), d_iters=o[2].data_ptr(),
                          d_accepted=o[3].data_ptr()))
        total += n
    graph = pmilos.Graph(items)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(args.steps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    ms_graph = timed(lambda: graph.launch(st.cuda_stream))

    def one_by_one():
        for it in items:
            it["ctx"].invert_ptrs(it["d_stokes"], it["npix"], it["d_models"], it["d_chi2"],
                                  it["d_iters"], it["d_accepted"], 0, st.cuda_stream)
    ms_seq = timed(one_by_one)
    graph.close()
    # each cube alone, for its own kernel time and roofline; the algorithmic
    # work per pixel from its recorded trips (the closed-form FLOP model,
    # with macroturbulence for the 112-sample cube)
    from pmilos import flops
    peak = PEAK_FP64_TFLOPS if precision == pmilos.PM_FP64 else PEAK_FP32_TFLOPS
    total_flop = 0.0
    for k, (it, shape, (n, g, mac, seed)) in enumerate(zip(items, shapes, specs)):
        ms = timed(lambda it=it: it["ctx"].invert_ptrs(
            it["d_stokes"], it["npix"], it["d_models"], it["d_chi2"], it["d_iters"],
            it["d_accepted"], 0, st.cuda_stream))
        iters = keep[k][2][2].cpu().numpy().astype(np.float64)
        acc = keep[k][2][3].cpu().numpy().astype(np.float64)
        fl = float(np.sum(flops.pixel_flops(iters, acc, L=shape["nlambda"], NT=shape["nterms"],
                                            mac=mac)))
        total_flop += fl
        tf = fl / (ms * 1e-3) / 1e12
        shape.update({"mean_iters": round(float(iters.mean()), 3),
                      "mean_accepted": round(float(acc.mean()), 3),
                      "kernel_ms_alone": round(ms, 3),
                      "pixels_per_s_alone": round(n / (ms * 1e-3), 1),
                      "roofline": {"bound": "valu", "achieved": round(tf, 3), "peak": peak,
                                   "unit": "TFLOP/s", "frac": round(tf / peak, 4),
                                   "flop_per_pixel": round(fl / n, 1)}})
    tf = total_flop / (ms_graph * 1e-3) / 1e12
    return {"workload": "uneven batch: 512^2 px x 6 lambda + 512^2 px x 30 lambda + 256^2 px x "
                        "112 lambda (v_mac free), one graph launch",
            "pixels": total, "ms_per_launch": round(ms_graph, 3),
            "pixels_per_s": round(total / (ms_graph * 1e-3), 1),
            "ms_one_by_one": round(ms_seq, 3), "cubes": shapes,
            "roofline": {"bound": "valu", "achieved": round(tf, 3), "peak": peak,
                         "unit": "TFLOP/s", "frac": round(tf / peak, 4),
                         "note": "algorithmic FLOPs of the three cubes (closed-form model "
                                 "over each pixel's recorded trips) / graph launch time"}}


# the mixed form's stated per-parameter tolerance against the FP64 form
# (absolute): eta0, B [G], vlos [km/s], dopp [A], a, gamma [deg], phi [deg],
# S0, S1; the GPU tests hold the mixed form to it against the FP64 oracle
TOL_MIXED = np.array([0.5, 2.0, 2e-3, 2e-4, 5e-3, 0.2, 2.0, 1e-2, 1e-2])


def noise_sweep(done, make_cube, run_both):
    """The field at noise sigma 1e-4, 1e-3 and 3e-3 Ic (noise seeds 31, 32,
    33; the models of the main field): per sigma one FP64 and one mixed pass
    with the kernel shapes already chosen, the mixed form's per-parameter
    error against the FP64 form and the change in iteration count."""
    names = ["eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1"]
    out = {}
    for sigma, seed in ((1e-4, 31), (1e-3, 32), (3e-3, 33)):
        if sigma in done:
            r64, rmx = done[sigma]
        else:
            cube = make_cube(sigma, seed)
            r64, rmx = run_both(cube)
            del cube
        d = np.abs(rmx["models"][:, :9] - r64["models"][:, :9])
        di = rmx["iters"] - r64["iters"]
        out["%g" % sigma] = {
            "fp64_pixels_per_s": round(len(d) / (r64["kernel_ms"] * 1e-3), 1),
            "mixed_pixels_per_s": round(len(d) / (rmx["kernel_ms"] * 1e-3), 1),
            "median_abs_diff": dict(zip(names, [float("%.3g" % v) for v in np.median(d, 0)])),
            "p99_abs_diff": dict(zip(names, [float("%.3g" % v) for v in np.percentile(d, 99, 0)])),
            "frac_within_tol": dict(zip(names, [round(float(v), 5)
                                                for v in (d <= TOL_MIXED).mean(0)])),
            "mean_iters_fp64": round(float(r64["iters"].mean()), 3),
            "mean_iters_mixed": round(float(rmx["iters"].mean()), 3),
            "frac_iters_identical": round(float((di == 0).mean()), 4),
            "mean_abs_iter_change": round(float(np.abs(di).mean()), 3),
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--npix", type=int, default=1024 * 1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--copy-ref-mb", type=int, default=0,
                    help="first run a plain device copy of this many MiB (the reference "
                         "kernel for scaling HBM traffic counters)")
    ap.add_argument("--cpu-per-core", type=int, default=4000)
    ap.add_argument("--serve-per-sampling", type=int, default=25,
                    help="1024^2 scans per sampling in the streaming request loop (0 = skip it)")
    ap.add_argument("--serve-small", type=int, default=50,
                    help="32^2 and 256^2 scans besides them")
    ap.add_argument("--serve-util", type=float, default=0.5)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default kernel shape instead of the autotuner's")
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the host-to-host (pinned) pass and the logclambda=1 passes")
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the noise sweep (sigma 1e-4 and 3e-3 Ic) of the two forms")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the uneven-batch (one graph launch) measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus 1); multi-GPU sharding "
              "is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_per_core)

    import torch
    import pmilos
    from pmilos import flops

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    npix = args.npix
    lam = pmilos.grid(FE6173, -290, 20, 290)
    L = len(lam)
    G = pmilos.psf_gaussian(60.0, lam)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6, miter=50)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    # synthetic cube: library FP64 synthesis, float64 noise, stored float32
    models = torch.from_numpy(draw_models(npix, 2)).to(dev)

    def make_cube(sigma, seed):
        spec = torch.empty((npix, 4, L), dtype=torch.float64, device=dev)
        ctx.synth_ptrs(models.data_ptr(), npix, spec.data_ptr(), 0, sptr)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        ic = (models[:, 7] + models[:, 8]).view(-1, 1, 1)
        noise = torch.randn((npix, 4, L), dtype=torch.float64, device=dev, generator=gen)
        return (spec + noise * sigma * ic).to(torch.float32).contiguous()

    stokes = make_cube(1e-3, 32)
    out_m = torch.empty((npix, 11), dtype=torch.float64, device=dev)
    out_c = torch.empty(npix, dtype=torch.float32, device=dev)
    out_i = torch.empty(npix, dtype=torch.int32, device=dev)
    out_a = torch.empty(npix, dtype=torch.int32, device=dev)
    if args.copy_ref_mb:
        # a plain streaming kernel of known traffic (reads and writes
        # copy_ref_mb MiB once each) for scaling the HBM byte counters
        src = torch.ones((args.copy_ref_mb << 20) // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    torch.cuda.synchronize()

    def run_form(c, peak, cube=None, steps=None, warmup=None, tune=True):
        cube = stokes if cube is None else cube
        steps = args.steps if steps is None else steps
        warmup = args.warmup if warmup is None else warmup
        # kernel shape by the autotuner, on the whole cube (untimed): on a few
        # times the pixels the GPU holds at once the persistent grid's tail
        # hides the differences between the shapes
        shape = c.autotune_ptrs(cube.data_ptr(), npix, sptr) \
            if args.autotune and tune else c.variant()[0]

        def step():
            c.invert_ptrs(cube.data_ptr(), npix, out_m.data_ptr(), out_c.data_ptr(),
                          out_i.data_ptr(), out_a.data_ptr(), 0, sptr)
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern_ms = ev0.elapsed_time(ev1) / steps
        iters = out_i.cpu().numpy().astype(np.float64)
        acc = out_a.cpu().numpy().astype(np.float64)
        fl = float(np.sum(flops.pixel_flops(iters, acc, L=L, NT=9)))
        achieved = fl / (kern_ms * 1e-3) / 1e12
        return dict(value=npix * steps / wall, ms_per_step=wall * 1000.0 / steps,
                    shape=list(shape),
                    kernel_ms=kern_ms, achieved=achieved, frac=achieved / peak,
                    mean_iters=float(iters.mean()), mean_accepted=float(acc.mean()),
                    fallbacks=c.fallback_count(), models=out_m.cpu().numpy(), iters=iters)

    def host_to_host(c):
        """Passes of pm_invert from the float32 cube in pinned host memory to
        the results in pinned host memory (the library's host-buffer path:
        the pixels in chunks over its streams, each chunk's upload and
        copy-back overlapping other chunks' kernels), wall clock averaged
        over --steps passes; and the upload and copy-back alone, back to
        back on one stream."""
        def pinned(shape, dtype):
            return torch.empty(shape, dtype=getattr(torch, np.dtype(dtype).name),
                               pin_memory=True).numpy()

        h_in = pinned(tuple(stokes.shape), np.float32)
        h_in[...] = stokes.cpu().numpy()
        out = pmilos.result_buffers(npix, alloc=pinned)
        d_in = torch.empty_like(stokes)
        outs = (out_m, out_c, out_i, out_a)
        t_in = torch.from_numpy(h_in)
        t_out = [torch.from_numpy(out[k]) for k in ("model", "chi2", "iters", "accepted")]

        def copies():
            d_in.copy_(t_in, non_blocking=True)
            for h, o in zip(t_out, outs):
                h.copy_(o, non_blocking=True)

        def whole():
            c.invert(h_in, out=out)

        res = {}
        for name, fn in (("copies_only", copies), ("whole", whole)):
            fn()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                fn()
            torch.cuda.synchronize()
            res[name] = (time.perf_counter() - t) / args.steps
        del d_in, h_in, out, t_in, t_out
        return {"pixels_per_s": round(npix / res["whole"], 1),
                "ms": round(res["whole"] * 1e3, 3),
                "copies_only_ms": round(res["copies_only"] * 1e3, 3),
                "passes": args.steps}

    def issued(r, form):
        # counter-measured issue rate of this pass (the per-pixel figures are
        # those of the bench field; other --npix values scale them)
        return {k: v * npix / (r["kernel_ms"] * 1e-3) / 1e12
                for k, v in ISSUED_FLOP_PER_PX[form].items()}

    if args.form == "mixed":
        ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6, miter=50,
                             precision=pmilos.PM_MIXED)
    r64 = run_form(ctx, PEAK_FP64_TFLOPS)
    line = {
        "metric": METRIC, "value": round(r64["value"], 1), "unit": "pixels/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (library FP64 synthesis of seeded models + seeded noise)",
        "config": {"workload": "1024x1024 px Fe I 6173.34 A, 30 lambda @ 20 mA, PSF 60 mA, "
                               "sigma 1e-3 Ic, 9 free, <=50 LM it, plain FP64 form",
                   "npix": npix, "nlambda": L, "nterms": 9, "form": "fp64",
                   "lanes_per_pixel_x_samples_per_lane_x_pixels_per_workgroup": r64["shape"]},
        "roofline": {"bound": "valu", "achieved": round(r64["achieved"], 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(r64["frac"], 4),
                     "traffic": round(TRAFFIC_B_PER_PX["fp64"] * npix),
                     "issued_fp64_tflops": round(issued(r64, "fp64")["fp64"], 3),
                     "issued_frac": round(issued(r64, "fp64")["fp64"] / PEAK_FP64_TFLOPS, 4),
                     "note": "FP64 vector ALU work (78.6 TFLOP/s, the MI355X FP64 vector "
                             "peak, equal to its FP64 matrix peak); achieved = algorithmic "
                             "FLOPs of the pass / pm_invert_kernel time; issued = lane "
                             "FLOPs issued per pixel from the SQ_INSTS_VALU_FLOPS "
                             "counters / the same time; traffic = HBM "
                             "bytes per launch from the FETCH_SIZE/WRITE_SIZE counters, "
                             "%.0f B/pixel against %d algorithmic" % (TRAFFIC_B_PER_PX["fp64"],
                                                                     ALGO_B_PER_PX)},
        "kernel_ms": round(r64["kernel_ms"], 3),
        "mean_iters": round(r64["mean_iters"], 3),
        "mean_accepted": round(r64["mean_accepted"], 3),
        "fallback_solves": r64["fallbacks"],
    }
    if args.form == "both":
        cm = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6, miter=50,
                            precision=pmilos.PM_MIXED)
        rmx = run_form(cm, PEAK_FP32_TFLOPS)
        d = np.abs(rmx["models"][:, :9] - r64["models"][:, :9])
        line["mixed"] = {
            "value": round(rmx["value"], 1), "unit": "pixels/s", "dtype": "f32+f64",
            "ms_per_step": round(rmx["ms_per_step"], 3), "kernel_ms": round(rmx["kernel_ms"], 3),
            "speedup_vs_fp64": round(rmx["value"] / r64["value"], 3),
            "lanes_per_pixel_x_samples_per_lane_x_pixels_per_workgroup": rmx["shape"],
            "roofline": {"achieved": round(rmx["achieved"], 3), "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(rmx["frac"], 4),
                         "issued_fp32_tflops": round(issued(rmx, "mixed")["fp32"], 3),
                         "issued_fp64_tflops": round(issued(rmx, "mixed")["fp64"], 3),
                         "traffic": round(TRAFFIC_B_PER_PX["mixed"] * npix)},
            "mean_iters": round(rmx["mean_iters"], 3),
            "median_abs_diff_vs_fp64": dict(zip(pmilos.PARAM_NAMES[:9],
                                                [float("%.3g" % v) for v in np.median(d, 0)])),
            "p99_abs_diff_vs_fp64": dict(zip(pmilos.PARAM_NAMES[:9],
                                             [float("%.3g" % v) for v in np.percentile(d, 99, 0)])),
            "frac_within_tol_vs_fp64": dict(zip(pmilos.PARAM_NAMES[:9],
                                                [round(float(v), 5)
                                                 for v in (d <= TOL_MIXED).mean(0)])),
            "toplim_mixed": pmilos.PM_MIXED_TOPLIM,
        }
        # the same pass with the reference's chi2-unchanged test (toplim alone)
        cm.set_mixed_toplim(0.0)
        r0 = run_form(cm, PEAK_FP32_TFLOPS, steps=1, warmup=1, tune=False)
        cm.set_mixed_toplim(pmilos.PM_MIXED_TOPLIM)
        d0 = np.abs(r0["models"][:, :9] - r64["models"][:, :9])
        line["mixed"]["reference_exit_test"] = {
            "pixels_per_s": round(npix / (r0["kernel_ms"] * 1e-3), 1),
            "kernel_ms": round(r0["kernel_ms"], 3),
            "speedup_vs_fp64": round(r64["kernel_ms"] / r0["kernel_ms"], 3),
            "mean_iters": round(r0["mean_iters"], 3),
            "frac_within_tol_vs_fp64": dict(zip(pmilos.PARAM_NAMES[:9],
                                                [round(float(v), 5)
                                                 for v in (d0 <= TOL_MIXED).mean(0)]))}
        if args.extra:
            line["host_to_host"] = {"fp64": host_to_host(ctx), "mixed": host_to_host(cm)}
            line["logclambda_1"] = {}
            for name, prec, peak in (("fp64", pmilos.PM_FP64, PEAK_FP64_TFLOPS),
                                     ("mixed", pmilos.PM_MIXED, PEAK_FP32_TFLOPS)):
                cl = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6,
                                    miter=50, logclambda=1, precision=prec)
                r = run_form(cl, peak, steps=1, warmup=1)
                line["logclambda_1"][name] = {
                    "pixels_per_s": round(npix / (r["kernel_ms"] * 1e-3), 1),
                    "kernel_ms": round(r["kernel_ms"], 3),
                    "mean_iters": round(r["mean_iters"], 3),
                    "mean_accepted": round(r["mean_accepted"], 3)}
                cl.close()
                if name == "fp64":
                    m64 = r["models"]
                else:
                    d = np.abs(r["models"][:, :9] - m64[:, :9])
                    line["logclambda_1"]["mixed"]["frac_within_tol_vs_fp64"] = dict(
                        zip(pmilos.PARAM_NAMES[:9],
                            [round(float(v), 5) for v in (d <= TOL_MIXED).mean(0)]))
                    line["logclambda_1"]["mixed"]["p99_abs_diff_vs_fp64"] = dict(
                        zip(pmilos.PARAM_NAMES[:9],
                            [float("%.3g" % v) for v in np.percentile(d, 99, 0)]))
        if args.sweep:
            line["mixed"]["noise_sweep"] = noise_sweep(
                {1e-3: (r64, rmx)}, lambda sg, sd: make_cube(sg, sd),
                lambda cube: (run_form(ctx, PEAK_FP64_TFLOPS, cube, 1, 0, False),
                              run_form(cm, PEAK_FP32_TFLOPS, cube, 1, 0, False)))
    if args.batch:
        line["uneven_batch"] = {"fp64": uneven_batch(args, torch, pmilos, pmilos.PM_FP64),
                                "mixed": uneven_batch(args, torch, pmilos, pmilos.PM_MIXED)}
    if args.serve_per_sampling > 0:
        from pmilos import serving
        del stokes, out_m, out_c, out_i, out_a
        torch.cuda.empty_cache()
        line["serving"] = serving.serve_bench(
            pmilos.PM_MIXED, util=args.serve_util,
            kinds=serving.bench_kinds(args.serve_per_sampling, args.serve_small))
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
