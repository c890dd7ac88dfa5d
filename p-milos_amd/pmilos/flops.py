"""Algorithmic work of one pixel's inversion: the frozen closed-form FLOP and
byte model used for roofline figures.

Counting rules: real + - * / = 1; complex * = 6, complex / = 11, complex
+- real = 1; exp/sin/cos/atan/sqrt = 1; clamps and compares 0.  The counts
follow the reference's per-pixel operation structure (mil_sinrf, me_der,
covarm2, fchisqr, the NT x NT solve, AplicaDelta/check, estimacionesClasicas)
with J^T J counted symmetric and the solve as 9 NT^3 + 4 NT^2 + 2 NT whatever
method computes it.

    F_pixel = F_init + iters * F_iter + accepted * F_acc
"""


def terms(L, NT=9, nc=3, psf=True, mac=False):
    """Per-pixel FLOP terms for L wavelengths, NT free parameters, nc Zeeman
    components; PSF-convolved without macroturbulence, or (mac) with the
    macroturbulence kernel folded with the PSF and applied by circular
    convolution (mil_sinrf.c:287-357, me_der.c:316-438): the kernel (7L),
    its PSF folding (2L^2) and 4 circular rows (8L^2) in the synthesis; the
    kernel and its derivative (13L), both folded (4L^2), the v_mac row
    (8L^2) and 8 rows x 4 Stokes circular (64L^2) in the RFs."""
    if mac:
        f_synth = L * (80 + 105 * nc) + 30 + 10 * L * L + 7 * L
        f_rf = L * (22 * nc + 872) + 76 * L * L + 13 * L
    else:
        conv_s = (8 * L * L + 2 * L) if psf else 0             # 4 rows
        conv_rf = (8 * (NT - 1) * L * L + 2 * (NT - 1) * L) if psf else 0
        f_synth = L * (80 + 105 * nc) + 30 + conv_s
        f_rf = L * (22 * nc + 872) + conv_rf
    f_chi2 = 20 * L
    f_normal = 4 * L * (NT + 2) ** 2 + 1.5 * NT * (NT + 3)
    f_solve = 9 * NT ** 3 + 4 * NT ** 2 + 2 * NT
    f_update = 60
    f_class = 30 * L + 50
    return dict(
        init=f_class + f_synth + f_chi2 + f_rf + f_normal,
        iter=f_solve + f_update + f_synth + f_chi2,
        acc=f_rf + f_normal,
    )


def pixel_flops(iters, accepted, L=30, NT=9, nc=3, psf=True, mac=False):
    """Total algorithmic FLOPs of pixels with the given trip counts (arrays or
    scalars; iters = LM trips, accepted = chi2-decreasing trips)."""
    t = terms(L, NT, nc, psf, mac)
    return t["init"] + iters * t["iter"] + accepted * t["acc"]


def pixel_bytes(L, fit=False):
    """Algorithmic HBM bytes per pixel: float32 IQUV in, 13 float32 model
    planes out (+ the float32 fit)."""
    return 16 * L + 52 + (16 * L if fit else 0)
