"""The host code of the path under AddressSanitizer and UBSan
(tests/sanitize/Makefile builds the instrumented copies):

- the oracle's and the file layer's CPU tests (tests/test_oracle.py,
  tests/test_lm_schedule.py, tests/test_cli_io.py) run again in a child
  process against liboracle_asan.so and libpmilos_io_asan.so, with the
  sanitizer runtime preloaded;
- the command line tool, built instrumented, reads the reference's control
  files, LINES, grids, models and FITS stray light up to the point where it
  needs the GPU (context creation fails on a host without one), and takes
  its error paths (no arguments, a truncated control file, a missing file).

Every run must end without a sanitizer report.  The device code itself runs
under the same sanitizers in tests/test_lanes_cpu.py."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "sanitize")
GOLDEN = os.path.join(HERE, "golden")
REPORTS = ("AddressSanitizer", "runtime error:", "LeakSanitizer", "UndefinedBehaviorSanitizer")


def _asan_runtime():
    out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                         text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(os.path.join(ROOT, "p-milos_amd", "pmilos", "libpmilos_hip.so")):
        pytest.fail("libpmilos_hip.so not built: run __graft_entry__.build()")
    subprocess.check_call(["make", "-s", "-C", SAN])
    rt = _asan_runtime()
    if rt is None:
        pytest.fail("g++ has no libasan.so")
    return rt


def _clean(text):
    return not any(r in text for r in REPORTS)


def test_oracle_and_file_layer_tests_under_sanitizers(built):
    env = dict(os.environ, LD_PRELOAD=built, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               ORACLE_LIB=os.path.join(SAN, "liboracle_asan.so"),
               PMIO_LIB=os.path.join(SAN, "libpmilos_io_asan.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", "tests/test_oracle.py",
                        "tests/test_lm_schedule.py", "tests/test_cli_io.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert _clean(out), out[-4000:]
    assert " passed" in r.stdout


def _cli(args, cwd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([os.path.join(SAN, "pmilos_cli_asan")] + args, cwd=cwd, env=env,
                          capture_output=True, text=True, timeout=120)


def test_command_line_tool_under_sanitizers(built, tmp_path):
    for f in ("synthesis.mtrol", "obs.mod", "malla.grid", "straylight.fits", "pmilos.mtrol",
              "pmilos.minit", "guess.mod", "wavelength.grid", "CRISP_6173_25mA.psf"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    shutil.copy(os.path.join(GOLDEN, "LINES_test"), tmp_path / "LINES")
    runs = [[], ["synthesis.mtrol"], ["pmilos.mtrol"], ["pmilos.minit"], ["missing.mtrol"]]
    lines = open(os.path.join(GOLDEN, "synthesis.mtrol")).read().splitlines()
    for n in (5, 20, 41):                                  # truncated control files
        (tmp_path / ("cut%d.mtrol" % n)).write_text("\n".join(lines[:n]) + "\n")
        runs.append(["cut%d.mtrol" % n])
    for args in runs:
        r = _cli(args, tmp_path)
        out = r.stdout + r.stderr
        assert _clean(out), (args, out[-4000:])
        # every run stops with a message of its own: without a GPU the full
        # control file fails at context creation, the others earlier
        assert r.returncode != 0 and ("pmilos_gpu" in out or "usage" in out), (args, out)
    assert "context creation failed" in _cli(["synthesis.mtrol"], tmp_path).stderr
