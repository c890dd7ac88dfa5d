"""Host-side driver of tests/lanes/lanes_run (the device code run on CPU
threads, one per lane, under AddressSanitizer and UBSan).  TEST
INFRASTRUCTURE for tests/test_lanes_cpu.py."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LANES = os.path.join(HERE, "lanes")
EXE = os.path.join(LANES, "lanes_run")


def build():
    subprocess.check_call(["make", "-s", "-C", LANES])
    return EXE


def _run(payload, timeout=900):
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fi, "wb") as f:
            f.write(payload)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
                   UBSAN_OPTIONS="print_stacktrace=1")
        res = subprocess.run([EXE, fi, fo], capture_output=True, text=True, timeout=timeout,
                             env=env)
        if res.returncode != 0:
            raise RuntimeError("lanes_run failed (%d):\n%s" % (res.returncode, res.stderr[-6000:]))
        return open(fo, "rb").read(), res.stderr


def invert(shape, cfg, stokes, slight_cube=None, want_fit=False):
    """shape = (precision, W, NLPL, BLK, general); cfg: dict with the
    pm_config fields (lambda, lambda0, quantum, psf, fix, guess, weight,
    noise, ilambda, toplim, miter, mu, logclambda, slight, no_classical,
    use_fft)."""
    prec, W, NLPL, BLK, gen = shape
    st = np.ascontiguousarray(stokes, dtype=np.float32)
    npix, _, L = st.shape
    psf, prof = cfg.get("psf"), cfg.get("slight")
    hdr = np.array([1, W, NLPL, BLK, int(gen), prec, npix, L, cfg.get("miter", 50),
                    cfg.get("logclambda", 0), int(psf is not None), int(cfg.get("use_fft", 0)),
                    int(cfg.get("no_classical", 0)), int(prof is not None),
                    int(slight_cube is not None), int(want_fit)], dtype=np.int32)
    par = np.zeros(16)
    par[0:6] = cfg["quantum"]
    par[6] = cfg["lambda0"]
    par[7] = cfg.get("noise", 1e-6)
    par[8] = cfg.get("ilambda", float(np.float32(0.1)))
    par[9] = cfg.get("toplim", 1e-12)
    par[10] = cfg.get("mu", 1.0)
    par[11:15] = cfg.get("weight", (1, 1, 1, 1))
    parts = [hdr.tobytes(), np.asarray(cfg["fix"], dtype=np.int32).tobytes(), par.tobytes(),
             np.asarray(cfg["guess"], dtype=np.float64).tobytes(),
             np.asarray(cfg["lambda"], dtype=np.float64).tobytes()]
    if psf is not None:
        parts.append(np.asarray(psf, dtype=np.float64).tobytes())
    if prof is not None:
        parts.append(np.asarray(prof, dtype=np.float32).tobytes())
    parts.append(st.tobytes())
    if slight_cube is not None:
        parts.append(np.ascontiguousarray(slight_cube, dtype=np.float32).tobytes())
    out, err = _run(b"".join(parts))
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(out, dtype=dt, count=n, offset=o)
        o += a.nbytes
        return a
    r = dict(model=take(np.float64, npix * 11).reshape(npix, 11), chi2=take(np.float32, npix),
             iters=take(np.int32, npix), accepted=take(np.int32, npix))
    r["fit"] = take(np.float32, npix * 4 * L).reshape(npix, 4, L) if want_fit else None
    r["fallbacks"] = int(take(np.uint64, 1)[0])
    r["stderr"] = err
    return r


def solve(W, N, ab, damp, check):
    """ab [K][N(N+1)/2 + N] (packed lower alpha, then beta); returns per
    matrix (x_group, ok_group, x_replicated, ok_replicated, consistent)."""
    ab = np.ascontiguousarray(ab, dtype=np.float64)
    K = ab.shape[0]
    hdr = np.array([2, W, N, 0, 0, 0, K] + [0] * 9, dtype=np.int32)
    out, _ = _run(hdr.tobytes() + ab.tobytes() + np.asarray(damp, dtype=np.float64).tobytes()
                  + np.asarray(check, dtype=np.float64).tobytes())
    r = np.frombuffer(out, dtype=np.float64).reshape(K, 2 * N + 3)
    return r[:, :N], r[:, N] != 0, r[:, N + 1:2 * N + 1], r[:, 2 * N + 1] != 0, r[:, 2 * N + 2] != 0
