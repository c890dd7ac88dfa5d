"""Writes the fixtures of the GPU parity tests for every seeded case of
tests/cases.py:

tests/golden/oracle_cases.npz
    per case: the FP64 oracle's inversion of the case's float32 cube
    (models, trip counts, chi2) and the SHA-256 of the cube, so that the GPU
    tests compare against the oracle without running it on the GPU host and
    check that they regenerated the same cube.

tests/golden/tolerances.json
    per case, the stated per-parameter tolerances of the two GPU forms and
    what they were derived from.  The cube is inverted with builds of the
    oracle that differ from liboracle.so only in rounding (oracle/Makefile
    `variants`: the reference's own -O3 -march=native flags, a fast-math
    build, the LDL^T solve in place of the eigen pseudo-inverse, every
    synthesised value perturbed by 4e-16 / 2e-15 relative) and with the
    REAL = float build (the reference's default build):
    - FP64 form:  max(FLOOR, Q-th percentile over pixels of the largest
                  |difference| any rounding variant ends at);
    - mixed form: the larger of that and KMIX x the Q-th percentile of the
                  float build's |difference|.
    Q = 99.9: the distance one pixel in a thousand reaches between two
    implementations of the same arithmetic.  KMIX = 3: the mixed form keeps
    more in FP32 than the reference's float build does (its convolution
    sums and Voigt division are FP32; the float build sums in double and
    evaluates the Voigt in complex double), so its rounding per sample is a
    few times larger.

Run from the repository root (8 processes; a few minutes):

    make -C oracle all variants && python tests/golden/make_oracle_cases.py [case ...]
"""
import json
import multiprocessing as mp
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O    # noqa: E402
import cases          # noqa: E402

CHECKER = "liboracle.so"
VARIANTS = ("liboracle_o3.so", "liboracle_fast.so", "liboracle_ldl.so", "liboracle_pert1.so",
            "liboracle_pert2.so")
FLOAT = "liboracle_f32.so"
Q = 99.9
KMIX = 3.0
BLOCK = 256


def run(job):
    name, build, p0, p1 = job
    O.use(CHECKER)
    c = cases.get(name)
    c.cube = c.cube[p0:p1]
    if c.per_pixel is not None:
        c.per_pixel = c.per_pixel[p0:p1]
    O.use(build)
    r = c.oracle()
    return name, build, p0, r


def spread(a, b):
    ok = np.all(np.isfinite(a), axis=1) & np.all(np.isfinite(b), axis=1)
    return np.abs(a[ok] - b[ok])


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "all", "variants"])
    names = sys.argv[1:] or list(cases.BUILDERS)
    O.use(CHECKER)
    built = {n: cases.get(n) for n in names}
    builds = (CHECKER,) + VARIANTS + (FLOAT,)
    jobs = [(n, b, p0, min(len(built[n].cube), p0 + BLOCK))
            for n in names for b in builds for p0 in range(0, len(built[n].cube), BLOCK)]
    t0 = time.time()
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(run, jobs, chunksize=1)
    print("oracle runs: %.0f s" % (time.time() - t0), flush=True)
    got = {}
    for n, b, p0, r in sorted(res, key=lambda x: (x[0], x[1], x[2])):
        d = got.setdefault((n, b), {k: [] for k in ("model", "iters", "chi2")})
        for k in d:
            d[k].append(r[k])
    got = {k: {kk: np.concatenate(vv) for kk, vv in v.items()} for k, v in got.items()}

    npz_path = os.path.join(HERE, "oracle_cases.npz")
    tol_path = os.path.join(HERE, "tolerances.json")
    arrays = dict(np.load(npz_path)) if os.path.exists(npz_path) else {}
    table = json.load(open(tol_path)) if os.path.exists(tol_path) else {}
    for n in names:
        c = built[n]
        ref = got[(n, CHECKER)]
        arrays[n + "|model"] = ref["model"]
        arrays[n + "|iters"] = ref["iters"].astype(np.int16)
        arrays[n + "|chi2"] = ref["chi2"].astype(np.float32)
        arrays[n + "|sha256"] = np.array(cases.cube_sha256(c.cube))
        qv = np.max([np.percentile(spread(got[(n, v)]["model"], ref["model"]), Q, axis=0)
                     for v in VARIANTS], axis=0)
        qf = np.percentile(spread(got[(n, FLOAT)]["model"], ref["model"]), Q, axis=0)
        t64 = np.maximum(cases.FLOOR, qv)
        tmx = np.maximum(t64, KMIX * qf)
        same = {v: round(float((got[(n, v)]["iters"] == ref["iters"]).mean()), 4)
                for v in VARIANTS + (FLOAT,)}
        table[n] = {
            "pixels": int(len(c.cube)),
            "fp64": dict(zip(cases.NAMES, [float("%.6g" % v) for v in t64])),
            "mixed": dict(zip(cases.NAMES, [float("%.6g" % v) for v in tmx])),
            "variants_q%g" % Q: dict(zip(cases.NAMES, [float("%.4g" % v) for v in qv])),
            "float_q%g" % Q: dict(zip(cases.NAMES, [float("%.4g" % v) for v in qf])),
            "iters_identical": same,
        }
        print("%-22s fp64 %s" % (n, np.array2string(t64, precision=2, max_line_width=250)),
              flush=True)
    np.savez_compressed(npz_path, **arrays)
    with open(tol_path, "w") as f:
        json.dump({"_rule": {"Q": Q, "KMIX": KMIX, "FLOOR": cases.FLOOR.tolist(),
                             "variants": list(VARIANTS), "float": FLOAT}, **table}, f, indent=1,
                  sort_keys=True)


if __name__ == "__main__":
    main()
