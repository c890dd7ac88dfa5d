"""Writes ref_core_vectors.npz: outputs of the reference's own fvoigt.c,
create_cuantic.c, fgauss.c and convolution.c (built unchanged into
oracle/_ref/libref_core.so by oracle/Makefile) on fixed inputs.

Run from the repository root after `make -C oracle`:
    python tests/golden/make_ref_vectors.py
The vectors let the oracle be checked where the reference tree is absent.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

QUANTUM = [(2, 1, 1.0, 2, 2, 0.0), (2, 1, 2.0, 2, 2, 2.0), (2, 3, 1.0, 2, 2, 0.0),
           (3, 2, 1.0, 3, 2, 1.0), (2, 1, 2, 2, 1, 3), (2, 2, 0.0, 3, 2, 1.0)]


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_core.so not built")
    out = {}
    rng = np.random.default_rng(20240501)
    v = np.concatenate([np.linspace(-30, 30, 121), rng.uniform(-8, 8, 79)])
    out["voigt_v"] = v
    out["voigt_a"] = np.array([1e-4, 0.01, 0.05, 0.3, 1.0, 4.0, 10.0])
    Hs, Fs = [], []
    for a in out["voigt_a"]:
        H = np.zeros(len(v)); F = np.zeros(len(v))
        R.ref_fvoigt(float(a), v.copy(), len(v), H, F)
        Hs.append(H); Fs.append(F)
    out["voigt_H"] = np.array(Hs)
    out["voigt_F"] = np.array(Fs)
    for k, q in enumerate(QUANTUM):
        buf = np.zeros(512)
        n = R.ref_cuantic(*q, buf, 512)
        out["cuantic_q%d" % k] = np.array(q, dtype=np.float64)
        out["cuantic_out%d" % k] = buf[:n].copy()
    for name, (l0, a, st, b, fw) in {"g30": (6173.3356, -290, 20, 290, 60.0),
                                     "g6": (6173.3356, -172.5, 69, 172.5, 60.0),
                                     "g112": (6302.4936, -1193.25, 21.5, 1193.25, 60.0),
                                     "g31": (6173.3356, -300, 20, 300, 45.0)}.items():
        lam = O.grid(l0, a, st, b)
        n = len(lam)
        G = np.zeros(n)
        R.ref_fgauss_wl(fw, lam[1] - lam[0], lam[0], lam[n // 2], n, G)
        out["psf_lam_" + name] = lam
        out["psf_fwhm_" + name] = np.array([fw])
        out["psf_G_" + name] = G
        for d in (0, 1):
            gm = np.zeros(n)
            R.ref_fgauss(1.7, lam.copy(), n, l0, d, gm)
            out["mac_%s_d%d" % (name, d)] = gm
        x = rng.normal(size=n)
        out["conv_x_" + name] = x
        y = x.copy(); R.ref_direct(y, n, G.copy(), n); out["conv_direct_" + name] = y
        y = x.copy(); R.ref_direct_ic(y, n, G.copy(), n, 1.25); out["conv_ic_" + name] = y
        y = np.zeros(n); R.ref_circular(x.copy(), G.copy(), n, y); out["conv_circ_" + name] = y
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_core_vectors.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
