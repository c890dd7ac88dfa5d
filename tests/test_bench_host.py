"""Host-side logic of bench.py (no GPU): the one-GPU refusal and the noise
sweep's statistics."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def _pass(models, iters, ms):
    return {"models": models, "iters": np.asarray(iters, dtype=np.float64), "kernel_ms": ms}


def test_noise_sweep_statistics():
    """Per sigma: the mixed form's |difference| to the FP64 form against the
    stated tolerance, the trip change, and pixels/s from the kernel time."""
    n = 1000
    rng = np.random.default_rng(0)
    m64 = rng.normal(size=(n, 11))
    mixed = m64.copy()
    # parameter 1 (B): 10 pixels beyond its 2 G tolerance, the rest at 0.5 G
    mixed[:, 1] += 0.5
    mixed[:10, 1] += 5.0
    it64 = np.full(n, 40)
    itmx = it64.copy()
    itmx[:250] -= 4
    calls = []

    def make_cube(sigma, seed):
        calls.append((sigma, seed))
        return object()

    def run_both(cube):
        return _pass(m64, it64, 100.0), _pass(mixed, itmx, 50.0)

    out = bench.noise_sweep({1e-3: (_pass(m64, it64, 100.0), _pass(mixed, itmx, 50.0))},
                            make_cube, run_both)
    assert calls == [(1e-4, 31), (3e-3, 33)]          # the given sigma is not rerun
    assert set(out) == {"0.0001", "0.001", "0.003"}
    r = out["0.001"]
    assert r["frac_within_tol"]["B"] == 0.99
    assert r["frac_within_tol"]["eta0"] == 1.0
    assert r["median_abs_diff"]["B"] == 0.5
    assert r["frac_iters_identical"] == 0.75
    assert r["mean_abs_iter_change"] == 1.0
    assert r["mean_iters_fp64"] == 40.0 and r["mean_iters_mixed"] == 39.0
    assert r["fp64_pixels_per_s"] == n / 0.1 and r["mixed_pixels_per_s"] == n / 0.05


def test_flop_model_closed_forms():
    """The frozen per-pixel FLOP terms at the worked configurations:
    30 and 6 samples with the PSF (NT 9), 112 samples with macroturbulence
    and the PSF (NT 10)."""
    sys.path.insert(0, os.path.join(ROOT, "p-milos_amd"))
    from pmilos import flops
    assert flops.terms(30) == {"init": 121592, "iter": 26703, "acc": 100902}
    assert flops.terms(6) == {"init": 14144, "iter": 9783, "acc": 11094}
    assert flops.terms(112, NT=10, mac=True) == {"init": 1300707, "iter": 182214,
                                                  "acc": 1124563}
    assert flops.pixel_flops(25, 15) == 121592 + 25 * 26703 + 15 * 100902
