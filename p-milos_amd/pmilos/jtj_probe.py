"""J^T J on the matrix cores (v_mfma_*_16x16x4) against the inversion
kernel's vector reduction, timed in isolation (libpm_jtj_probe.so).

    python -m pmilos.jtj_probe [--blocks 8192] [--reps 64]

Both forms run on the same generated rows (9 response rows and the weighted
residual over 4 x 30 samples per pixel, 32 lanes per pixel).  Each form's
time net of generating the rows, per pixel, decides which one the inversion
kernel uses; both are checked against numpy sums of the same rows.  Prints
one JSON line.
"""
import argparse
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpm_jtj_probe.so")
NT, NV, W = 9, 54, 32


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError("libpm_jtj_probe.so not built: run python p-milos_amd/pmilos/build.py")
    L = C.CDLL(LIB)
    L.pmp_jtj_run.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                              C.POINTER(C.c_float)]
    L.pmp_jtj_rows.argtypes = [C.c_int, C.c_int, C.c_void_p]
    return L


def _reference(L, grp, reps):
    rows = np.zeros((NT + 1) * 4 * W)
    L.pmp_jtj_rows(grp, reps, rows.ctypes.data)
    M = rows.reshape(NT + 1, 4 * W)
    P = M @ M.T
    ab = [P[i, h] for i in range(NT) for h in range(i + 1)] + [P[NT, i] for i in range(NT)]
    return np.array(ab)


def run(blocks=8192, reps=64):
    """Both forms in both precisions: times, ns per pixel and the largest
    relative difference from numpy over a few groups' last pixels."""
    import torch
    L = _lib()
    dev = torch.device("cuda:0")
    groups = blocks * 4
    out = torch.zeros(groups * NV, dtype=torch.float64, device=dev)
    sink = torch.zeros(blocks * 128, dtype=torch.float64, device=dev)
    npx = groups * reps
    check_groups = [0, 1, 2, 3, groups - 1]
    refs = {g: _reference(L, g, reps) for g in check_groups}
    res = {}
    for fp64 in (0, 1):
        ms = {}
        for mode in (0, 1, 2):
            t = C.c_float(0)
            for _ in range(2):                       # the first launch warms up
                rc = L.pmp_jtj_run(fp64, mode, blocks, reps, out.data_ptr(),
                                   sink.data_ptr(), C.byref(t))
                if rc != 0:
                    raise RuntimeError("pmp_jtj_run failed (%d)" % rc)
            ms[mode] = t.value
            if mode:
                got = out.view(groups, NV).cpu().numpy()
                err = max(float(np.max(np.abs(got[g] - refs[g]) / np.max(np.abs(refs[g]))))
                          for g in check_groups)
                ms["err%d" % mode] = err
        form = "fp64" if fp64 else "fp32"
        res[form] = {
            "ms_rows_only": round(ms[0], 4),
            "ms_vector": round(ms[1], 4), "ms_mfma": round(ms[2], 4),
            "ns_per_pixel_vector": round((ms[1] - ms[0]) * 1e6 / npx, 4),
            "ns_per_pixel_mfma": round((ms[2] - ms[0]) * 1e6 / npx, 4),
            "max_rel_err_vector": float("%.3g" % ms["err1"]),
            "max_rel_err_mfma": float("%.3g" % ms["err2"]),
        }
    return {"pixels_per_launch": npx, "lanes_per_pixel": W, "nlambda": 30, "nterms": NT, **res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=8192)
    ap.add_argument("--reps", type=int, default=64)
    args = ap.parse_args()
    print(json.dumps(run(args.blocks, args.reps)), flush=True)


if __name__ == "__main__":
    main()
