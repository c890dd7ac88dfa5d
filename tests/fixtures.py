"""Shared test helpers: fixture readers and seeded synthetic workloads.

The FITS reader here only serves the reference's 2-axis float stray-light
fixture (tests/golden/straylight.fits); the model draws follow the seeded
recipe of the benchmark configurations (B, gamma, azimuth, vlos, eta0,
Doppler width, damping, S0, S1 drawn in that order).
"""
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FE6173 = 6173.3356
FE6302 = 6302.4936
TRIPLET = (2, 1, 1.0, 2, 2, 0.0)     # 5P1.0 - 5D0.0 as readFileCuanticLines returns it
GUESS = np.array([20, 1100, 0.2, 0.03, 0.05, 120, 150, 0.35, 0.5, 0, 1.0])  # run/guess.mod


def read_fits_2d(path):
    raw = open(path, "rb").read()
    hdr = {}
    pos = 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            hdr[key] = card[10:].split("/")[0].strip()
    start = ((pos + 2879) // 2880) * 2880
    n1, n2 = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
    assert int(hdr["BITPIX"]) == -32
    return np.frombuffer(raw[start:start + n1 * n2 * 4], dtype=">f4").astype(np.float64).reshape(n2, n1)


def draw_models(n, seed, mac_range=None):
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
    mac = rng.uniform(0, 2, n) if mac_range else np.zeros(n)
    return np.stack([eta, B, vl, dop, aa, gm, az, S0, S1, mac, np.ones(n)], 1)


def add_noise(spec, models, sigma, seed):
    rng = np.random.default_rng(seed)
    ic = (models[:, 7] + models[:, 8])[:, None, None]
    return (spec + rng.normal(0.0, 1.0, spec.shape) * sigma * ic).astype(np.float32)


def setup_l30(psf=True, **kw):
    lam = O.grid(FE6173, -290, 20, 290)
    G = O.gauss_psf(60.0, lam) if psf else None
    return O.make_setup(FE6173, TRIPLET, lam, psf=G, **kw), lam, G


def oracle_cube(s, models):
    return np.stack([O.synth(s, m) for m in models])


def read_mod(path):
    """Init_Model from a .mod file (the reference's model text format)."""
    keys = {"eta_0": 0, "magnetic field": 1, "LOS velocity": 2, "Doppler width": 3,
            "damping": 4, "gamma": 5, "phi": 6, "S_0": 7, "S_1": 8, "v_mac": 9,
            "filling factor": 10}
    m = np.zeros(11)
    for line in open(path):
        if ":" not in line:
            continue
        name, val = line.split(":", 1)
        for k, i in keys.items():
            if k in name:
                m[i] = float(val.split()[0])
    return m
