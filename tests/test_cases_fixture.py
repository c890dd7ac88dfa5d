"""The fixtures the GPU parity tests compare against (tests/golden/
oracle_cases.npz, tests/golden/tolerances.json, written by
tests/golden/make_oracle_cases.py) are those of the live oracle: every case
regenerates the cube the fixture was made from, and the oracle inverts a
sample of its pixels to the stored bits.  CPU only."""
import json
import zlib

import numpy as np
import pytest

import cases
import oracle as O


@pytest.mark.parametrize("name", list(cases.BUILDERS))
def test_fixture_is_the_live_oracle(name):
    c = cases.get(name)
    ref = cases.reference(name, c.cube)            # asserts the cube's hash
    n = len(c.cube)
    assert ref["model"].shape == (n, 11) and ref["iters"].shape == (n,)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    idx = np.unique(np.concatenate([[0, 1, n // 2, n - 1], rng.choice(n, 8, replace=False)]))
    sub = cases.get(name)
    sub.cube = c.cube[idx]
    if sub.per_pixel is not None:
        sub.per_pixel = [c.per_pixel[i] for i in idx]
    r = sub.oracle()
    assert np.array_equal(r["model"], ref["model"][idx], equal_nan=True)
    assert np.array_equal(r["iters"], ref["iters"][idx])
    assert np.array_equal(r["chi2"], ref["chi2"][idx], equal_nan=True)


def test_tolerances_cover_every_case():
    with open(cases.TOLERANCES) as f:
        t = json.load(f)
    for name in cases.BUILDERS:
        for form in ("fp64", "mixed"):
            tol = cases.tolerance(name, form)
            assert tol.shape == (11,) and np.all(tol >= cases.FLOOR), (name, form)
        assert np.all(cases.tolerance(name, "mixed") >= cases.tolerance(name, "fp64"))
        assert t[name]["pixels"] == len(cases.reference(name)["iters"])
