"""The C-ABI library loads and exports every entry point include/pmilos.h
declares; its host-only setup functions match the oracle.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import pmilos
from fixtures import FE6173

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "pmilos.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pmilos.LIB_PATH):
        pytest.fail("libpmilos_hip.so not built: run __graft_entry__.build()")
    return pmilos.lib()


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert "pm_invert" in names and "pm_ctx_create" in names and len(names) >= 12
    raw = ctypes.CDLL(pmilos.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n


def test_grid_matches_reference_accumulation(lib):
    for args in ((FE6173, -290, 20, 290), (FE6173, -172.5, 69, 172.5),
                 (6302.4936, -1193.25, 21.5, 1193.25), (FE6173, -600, 25, 600),
                 (FE6173, -350, 35, 665)):
        assert np.array_equal(pmilos.grid(*args), O.grid(*args))


def test_psf_matches_oracle(lib):
    lam = O.grid(FE6173, -290, 20, 290)
    assert np.array_equal(pmilos.psf_gaussian(60.0, lam), O.gauss_psf(60.0, lam))
    lam = O.grid(6302.4936, -1193.25, 21.5, 1193.25)
    np.testing.assert_allclose(pmilos.psf_gaussian(60.0, lam), O.gauss_psf(60.0, lam),
                               rtol=1e-15, atol=1e-300)


def test_nterms_table(lib):
    f = lambda *z: (ctypes.c_int * 11)(*z)
    assert lib.pm_nterms(f(*([1] * 9 + [0, 0])), ctypes.c_double(0.0)) == 9
    assert lib.pm_nterms(f(*([1] * 9 + [1, 0])), ctypes.c_double(0.0)) == 10
    assert lib.pm_nterms(f(*([1] * 9 + [0, 0])), ctypes.c_double(0.5)) == 10
    assert lib.pm_nterms(f(*([1] * 11)), ctypes.c_double(0.5)) == 11
    assert lib.pm_nterms(f(*([1] * 9 + [0, 1])), ctypes.c_double(0.0)) == 10


def test_error_codes_without_device_calls(lib):
    assert lib.pm_strerror(-22) == b"invalid argument"
    assert lib.pm_grid(FE6173, 0.0, 0.0, 10.0, None, 0) == -22
    assert lib.pm_psf_gaussian(0.0, None, 0, None) == -22
    lib.pm_ctx_set_mixed_toplim.argtypes = [ctypes.c_void_p, ctypes.c_double]
    assert lib.pm_ctx_set_mixed_toplim(None, 1e-3) == -22

