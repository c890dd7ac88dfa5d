"""GPU tests of the command line tool (pmilos_gpu), the drop-in for
milos.x / pmilos.x: it runs the reference's own synthesis control file to
the reference's golden output, inverts .per profiles like lm_mils from the
guess, and FITS cubes, time series, subfields, masks and watched
directories with the same bits the library gives when called directly."""
import os
import shutil
import subprocess
import sys
import time

import numpy as np
import pytest

import cases
import fitsutil
import oracle as O
import pmilos
from fixtures import FE6173, GOLDEN, GUESS, TRIPLET, add_noise, draw_models

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(pmilos.LIB_PATH), "pmilos_gpu")
TOL64 = np.array([1e-3, 1e-2, 1e-6, 1e-8, 1e-6, 1e-4, 1e-4, 1e-6, 1e-6])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(args, cwd, timeout=120):
    say("case cli %s in %s" % (" ".join(args), cwd))
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    return r


def mtrol(path, cycles, observed, grid, psf="", straylight="", guess="guess.mod", fix_alfa=0,
          fix_mac=0, sn=1000, logclambda="", use_fft=""):
    """A 42-line control file from the reference's pmilos.mtrol, values set."""
    lines = open(os.path.join(GOLDEN, "pmilos.mtrol")).read().splitlines()
    vals = {0: cycles, 1: observed, 2: straylight, 3: psf, 4: grid, 5: "LINES", 7: guess,
            23: fix_mac, 35: fix_alfa, 37: sn, 40: use_fft, 41: logclambda}
    out = []
    for i, ln in enumerate(lines):
        if i in vals:
            name, rest = ln.split(":", 1)
            comment = rest[rest.index("!"):] if "!" in rest else ""
            ln = "%s:%s\t\t%s" % (name, vals[i], comment)
        out.append(ln)
    open(path, "w").write("\n".join(out) + "\n")


def grid_file(path, a, step, b, index=1):
    src = open(os.path.join(GOLDEN, "malla.grid")).read().splitlines()
    head = [l for l in src if not l.startswith("1 ")]
    head = src[:src.index(next(l for l in src if l.startswith("------"))) + 1]
    head.append("%d                       :        %s,            %s,           %s"
                % (index, a, step, b))
    open(path, "w").write("\n".join(head) + "\n")


@pytest.fixture
def work(tmp_path):
    shutil.copy(os.path.join(GOLDEN, "LINES_run"), tmp_path / "LINES")
    shutil.copy(os.path.join(GOLDEN, "guess.mod"), tmp_path / "guess.mod")
    grid_file(str(tmp_path / "l30.grid"), -290, 20, 290)
    return tmp_path


def _ctx_l30(**kw):
    lam = pmilos.grid(FE6173, -290, 20, 290)
    G = pmilos.psf_gaussian(60.0, lam)
    return lam, G, pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6,
                                  ilambda=float(np.float32(0.1)), miter=50, **kw)


def test_cli_golden_synthesis(tmp_path):
    """The reference's test/synthesis.mtrol (49 wavelengths, stray light with
    alpha 0.85, NTERMS 10) to test/synthesis_results.per."""
    for f in ("synthesis.mtrol", "obs.mod", "malla.grid", "straylight.fits"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    shutil.copy(os.path.join(GOLDEN, "LINES_test"), tmp_path / "LINES")
    run(["synthesis.mtrol"], tmp_path)
    out = np.loadtxt(tmp_path / "obs.per")
    gold = np.loadtxt(os.path.join(GOLDEN, "synthesis_results.per"))
    assert out.shape == gold.shape == (49, 6)
    assert np.array_equal(out[:, 0], gold[:, 0])
    np.testing.assert_allclose(out[:, 1], gold[:, 1], atol=1e-5)
    assert np.all(np.abs(out[:, 2:] - gold[:, 2:]) <= 3e-7 + 1e-6 * np.abs(gold[:, 2:]))


@pytest.mark.parametrize("fft", [0, 1])
def test_cli_per_inversion(work, fft):
    """A .per profile inverted from the guess itself (no classical estimates,
    milos.c:741-775), against the oracle's lm_mils; with the control file's
    FFT switch the PSF is circular (the oracle's FFT mode)."""
    lam = O.grid(FE6173, -290, 20, 290)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=O.gauss_psf(60.0, lam), use_fft=bool(fft))
    truth = np.array([9., 900., 0.8, 0.035, 0.15, 70., 40., 0.25, 0.72, 0, 1.])
    spec = add_noise(O.synth(s, truth)[None], truth[None], 1e-3, 90)[0]
    with open(work / "obs.per", "w") as f:
        for k in range(lam.size):
            f.write("1\t%f\t%e\t%e\t%e\t%e\n" % ((lam[k] - FE6173) * 1000, *spec[:, k]))
    spec32 = np.loadtxt(work / "obs.per")[:, 2:].T.astype(np.float32)
    mtrol(str(work / "p.mtrol"), 50, "obs.per", "l30.grid", psf=60, use_fft=fft)
    run(["p.mtrol"], work)
    txt = open(work / "obs_model.mod").read()
    vals = [float(l.split(":")[1]) for l in txt.splitlines() if ":" in l]
    m_cli, it_cli, chi_cli = np.array(vals[:11]), int(vals[11]), vals[12]
    fit = np.loadtxt(work / "obs_stokes.per")[:, 2:].T
    # the tool is the library: the same call gives the printed values
    # (the model file prints %lf, 6 decimals; the profiles %e)
    _, _, ctx = _ctx_l30(classical=False, use_fft=bool(fft))
    r = ctx.invert(spec32[None], want_fit=True)
    assert np.all(np.abs(m_cli[:9] - r["model"][0, :9]) <= 1e-6)
    assert it_cli == r["iters"][0]
    np.testing.assert_allclose(chi_cli, r["chi2"][0], rtol=1e-6)
    np.testing.assert_allclose(fit, r["fit"][0], rtol=1e-6, atol=1e-12)
    o = O.lm(s, spec32, GUESS)
    say("cli", m_cli[:9], it_cli, chi_cli, "oracle", o["model"][:9], o["iters"], o["chi2"])
    assert np.all(np.abs(m_cli[:9] - o["model"][:9]) <= np.maximum(TOL64, 1e-6))
    np.testing.assert_allclose(chi_cli, o["chi2"], rtol=1e-5)
    # the GPU's sums and solve round differently from the oracle's, and the
    # "chi2 unchanged" exit (|dchi2| * 100 / chi2 < toplim = 1e-12, i.e.
    # about 100 ulps of chi2) may come several trips apart once the model
    # has converged: steps at the rounding level decide it, and each
    # rejection raises the damping tenfold (measured: up to 4 trips here)
    assert abs(it_cli - o["iters"]) <= 6
    np.testing.assert_allclose(fit, o["fit"], rtol=1e-4, atol=1e-6)


def test_cli_fits_inversion_and_estimates(work):
    rows, cols, L = 8, 6, 30
    lam, G, ctx = _ctx_l30()
    models = draw_models(rows * cols, 91)
    cube = add_noise(ctx.synth(models), models, 1e-3, 92)
    fitsutil.write(str(work / "obs.fits"), fitsutil.stokes_cube(cube, rows, cols),
                   ctypes=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"))
    mtrol(str(work / "f.mtrol"), 50, "obs.fits", "l30.grid", psf=60)
    run(["f.mtrol"], work)
    ref = ctx.invert(cube, want_fit=True)
    data, hdr = fitsutil.read(str(work / "obs_mod.fits"))
    planes = data.reshape(13, rows * cols)
    np.testing.assert_array_equal(planes[:11].T, ref["model"].astype(np.float32))
    np.testing.assert_array_equal(planes[11], ref["iters"])
    np.testing.assert_array_equal(planes[12], ref["chi2"])
    fit, _ = fitsutil.read(str(work / "obs_stokes.fits"))
    np.testing.assert_array_equal(fit, fitsutil.stokes_cube(ref["fit"], rows, cols))
    # classical estimates only
    mtrol(str(work / "e.mtrol"), -1, "obs.fits", "l30.grid", psf=60)
    run(["e.mtrol"], work)
    est, _ = fitsutil.read(str(work / "obs_model_ce.fits"))
    np.testing.assert_array_equal(est.reshape(13, -1)[:11].T,
                                  ctx.classical(cube).astype(np.float32))
    s = O.make_setup(FE6173, TRIPLET, lam, psf=G)
    for p in range(0, rows * cols, 7):
        o = O.classical(s, cube[p], GUESS, for_init=False)
        np.testing.assert_allclose(est.reshape(13, -1)[[1, 2, 5, 6, 7], p],
                                   o[[1, 2, 5, 6, 7]].astype(np.float32), rtol=1e-6, atol=1e-5)


def test_cli_psf_file_recentred(work):
    """The PSF from the reference's instrument-profile file
    run/CRISP_6173_25mA.psf on a grid (-300..500 mA, 25 mA, 33 samples)
    whose zero offset is sample 12, not the middle one: the tool recentres
    the profile by (L/2 - 12) steps (milos.c:320-343) and interpolates it
    linearly (interpolationLinearPSF, milosUtils.c:684-738).  The cube is
    synthesised by the oracle with the PSF that numpy interpolates from the
    same file; the tool's inversions match the library called with that PSF
    and the oracle."""
    from test_cli_io import crisp_psf_on_grid
    shutil.copy(os.path.join(GOLDEN, "CRISP_6173_25mA.psf"), work / "crisp.psf")
    grid_file(str(work / "off.grid"), -300, 25, 500)
    off, shift, G = crisp_psf_on_grid(-300, 25, 500)
    assert shift == 100.0 and int(np.argmax(G)) == 16
    lam = O.grid(FE6173, -300, 25, 500)
    s = O.make_setup(FE6173, TRIPLET, lam, psf=G)
    rows, cols = 8, 8
    models = draw_models(rows * cols, 95)
    cube = add_noise(np.stack([O.synth(s, m) for m in models]), models, 1e-3, 96)
    fitsutil.write(str(work / "obs.fits"), fitsutil.stokes_cube(cube, rows, cols),
                   ctypes=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"))
    mtrol(str(work / "f.mtrol"), 50, "obs.fits", "off.grid", psf="crisp.psf")
    run(["f.mtrol"], work)
    planes = fitsutil.read(str(work / "obs_mod.fits"))[0].reshape(13, rows * cols)
    cli = planes[:11].T.astype(np.float64)
    ctx = pmilos.Context(FE6173, TRIPLET, lam, psf=G, guess=GUESS, noise=1e-6,
                         ilambda=float(np.float32(0.1)), miter=50)
    lib = ctx.invert(cube)["model"]
    o = O.invert(s, cube, GUESS)["model"]
    # the model cube holds float32; the PSFs of the tool and of numpy differ
    # in the last bits at most
    tol = np.maximum(cases.FLOOR, 1e-6 * np.abs(o))
    near_lib = np.all(np.abs(cli - lib) <= np.maximum(tol, 1e-6 * np.abs(lib)), axis=1)
    near_orc = np.all(np.abs(cli - o) <= np.maximum(tol, cases.tolerance("l30_config1")), axis=1)
    say("psf file: %d/%d pixels match the library, %d/%d the oracle"
        % (near_lib.sum(), near_lib.size, near_orc.sum(), near_orc.size))
    assert near_lib.mean() >= 63 / 64 and near_orc.mean() >= 63 / 64
    # a grid the file does not cover is refused
    grid_file(str(work / "wide.grid"), -700, 25, 700)
    mtrol(str(work / "w.mtrol"), 50, "obs.fits", "wide.grid", psf="crisp.psf")
    r = subprocess.run([CLI, "w.mtrol"], cwd=work, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "PSF" in (r.stdout + r.stderr)


def test_cli_mixed_and_its_exit_threshold(work):
    """--mixed inverts with the mixed form and its default exit threshold;
    --mixed-toplim 0 restores the reference's toplim test.  Each gives the
    bits of the library called directly with the same setting."""
    rows, cols = 6, 5
    lam, G, ctx = _ctx_l30(precision=pmilos.PM_MIXED)
    models = draw_models(rows * cols, 93)
    cube = add_noise(ctx.synth(models), models, 1e-3, 94)
    fitsutil.write(str(work / "obs.fits"), fitsutil.stokes_cube(cube, rows, cols),
                   ctypes=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"))
    mtrol(str(work / "f.mtrol"), 50, "obs.fits", "l30.grid", psf=60)
    for args, tl in ((["--mixed"], pmilos.PM_MIXED_TOPLIM), (["--mixed", "--mixed-toplim", "0"], 0.0)):
        run(["f.mtrol"] + args, work)
        ctx.set_mixed_toplim(tl)
        ref = ctx.invert(cube)
        planes = fitsutil.read(str(work / "obs_mod.fits"))[0].reshape(13, rows * cols)
        np.testing.assert_array_equal(planes[:11].T, ref["model"].astype(np.float32))
        np.testing.assert_array_equal(planes[11], ref["iters"])
        say("cli %s: mean iters %.2f" % (" ".join(args), planes[11].mean()))


def _minit(path, control, prefix, mask="", t1=1, t2=2, sub=(0, 0, 0, 0), save=1):
    lines = open(os.path.join(GOLDEN, "pmilos.minit")).read().splitlines()
    vals = [control, "fits", "fits", sub[0], sub[1], sub[2], sub[3], prefix, mask, t1, t2, save]
    out = []
    for ln, v in zip(lines, vals):
        name, rest = ln.split(":", 1)
        comment = rest[rest.index("!"):] if "!" in rest else ""
        out.append("%s:%s\t\t%s" % (name, v, comment))
    open(path, "w").write("\n".join(out) + "\n")


def test_cli_time_series_subfield_mask(work):
    rows, cols = 10, 7
    lam, G, ctx = _ctx_l30()
    cubes = {}
    for i in (1, 2):
        models = draw_models(rows * cols, 93 + i)
        cubes[i] = add_noise(ctx.synth(models), models, 1e-3, 95 + i)
        fitsutil.write(str(work / ("obs%03d.fits" % i)), fitsutil.stokes_cube(cubes[i], rows, cols),
                       ctypes=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"))
    sub = (2, 9, 3, 6)                       # rows 2..9, cols 3..6
    nr, nc = 8, 4
    mask = np.ones((cols, rows), dtype=np.int16)
    mask[3, 4] = 0                           # col 3, row 4 (1-based: row 5, col 4)
    mask[5, 2] = 0
    fitsutil.write(str(work / "mask.fits"), mask, bitpix=16)
    os.makedirs(work / "out")
    mtrol(str(work / "s.mtrol"), 50, "obs", "l30.grid", psf=60)
    _minit(str(work / "s.minit"), "s.mtrol", "out/", mask="mask.fits", sub=sub)
    run(["s.minit"], work)
    for i in (1, 2):
        c = cubes[i].reshape(cols, rows, 4, 30)[2:6, 1:9].reshape(-1, 4, 30)
        ref = ctx.invert(c, want_fit=True)
        data, _ = fitsutil.read(str(work / "out" / ("obs%03d_mod.fits" % i)))
        planes = data.reshape(13, nr * nc)
        m = mask[2:6, 1:9].reshape(-1)
        keep = m != 0
        np.testing.assert_array_equal(planes[:11, keep].T, ref["model"][keep].astype(np.float32))
        np.testing.assert_array_equal(planes[11, keep], ref["iters"][keep])
        assert np.all(planes[:, ~keep] == 0) and (~keep).sum() == 2
        fit, _ = fitsutil.read(str(work / "out" / ("obs%03d_stokes.fits" % i)))
        f = ref["fit"].copy()
        f[~keep] = 0
        np.testing.assert_array_equal(fit, fitsutil.stokes_cube(f, nr, nc))


def test_cli_watch_directory(work):
    """t1 = t2 = *: every numbered cube of the directory, then new cubes as
    they appear until the watch times out."""
    rows, cols = 4, 3
    lam, G, ctx = _ctx_l30()
    d = work / "data"
    os.makedirs(d)
    cubes = {}
    for i in (1, 7):
        models = draw_models(rows * cols, 100 + i)
        cubes[i] = add_noise(ctx.synth(models), models, 1e-3, 110 + i)
    cty = ("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN")
    fitsutil.write(str(d / "obs001.fits"), fitsutil.stokes_cube(cubes[1], rows, cols), ctypes=cty)
    mtrol(str(work / "w.mtrol"), 50, "data/obs", "l30.grid", psf=60)
    _minit(str(work / "w.minit"), "w.mtrol", "", t1="*", t2="*", save=0)
    say("case cli watch")
    p = subprocess.Popen([CLI, "w.minit", "--watch-timeout", "4", "--watch-poll", "0.2"],
                         cwd=work, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        t0 = time.time()
        while not (work / "obs001_mod.fits").exists() and time.time() - t0 < 60 \
                and p.poll() is None:
            time.sleep(0.1)
        if not (work / "obs001_mod.fits").exists():
            out, err = p.communicate(timeout=60)
            say(out[-2000:], err[-2000:])
            raise AssertionError("no output for the first cube")
        fitsutil.write(str(d / "obs007.fits.part"), fitsutil.stokes_cube(cubes[7], rows, cols),
                       ctypes=cty)
        os.rename(d / "obs007.fits.part", d / "obs007.fits")
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
    say(out[-1000:], err[-1000:])
    assert p.returncode == 0
    for i, name in ((1, "obs001_mod.fits"), (7, "obs007.fits_mod.fits")):
        data, _ = fitsutil.read(str(work / name))
        ref = ctx.invert(cubes[i])
        np.testing.assert_array_equal(data.reshape(13, -1)[:11].T, ref["model"].astype(np.float32))
