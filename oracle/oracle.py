"""ctypes bindings for the CPU oracle (liboracle.so) and oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
The product library never loads this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXL = 512
MAXC = 64
NPAR = 11

PARAM_NAMES = ["eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1",
               "mac", "alfa"]


class Zeeman(C.Structure):
    _fields_ = [("n_pi", C.c_int), ("n_sig", C.c_int),
                ("nup", C.c_double * MAXC), ("wep", C.c_double * MAXC),
                ("nub", C.c_double * MAXC), ("web", C.c_double * MAXC),
                ("nur", C.c_double * MAXC), ("wer", C.c_double * MAXC),
                ("gl", C.c_double), ("gu", C.c_double), ("geff", C.c_double),
                ("fo", C.c_double)]


class Setup(C.Structure):
    _fields_ = [("nlambda", C.c_int), ("nterms", C.c_int),
                ("lambda0", C.c_double), ("lam", C.c_double * MAXL),
                ("use_psf", C.c_int), ("G", C.c_double * MAXL),
                ("z", Zeeman), ("fix", C.c_int * NPAR),
                ("weight", C.c_double * 4), ("noise", C.c_double),
                ("ilambda", C.c_double), ("toplim", C.c_double),
                ("miter", C.c_int), ("ah", C.c_double),
                ("logclambda", C.c_int), ("use_slight", C.c_int),
                ("slight", C.c_double * (4 * MAXL)),
                ("exact_rho_deriv", C.c_int), ("use_fft", C.c_int)]


_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


# the builds of the restatement (oracle/Makefile): the FP64 checker, the
# same at the reference's -O3 -march=native flags (FMA contraction), and
# REAL = float, the reference's default build
BUILDS = ("liboracle.so", "liboracle_o3.so", "liboracle_f32.so")
_loaded = {}


def lib():
    """The oracle build in use: liboracle.so, or the one ORACLE_LIB names,
    or the one use() selected."""
    global _lib
    if _lib is None:
        _lib = _load(os.environ.get("ORACLE_LIB", "liboracle.so"))
    return _lib


def use(name):
    """Selects the oracle build `name` (one of BUILDS) for the calls that
    follow; returns the name of the one in use before."""
    global _lib
    prev = getattr(_lib, "_oracle_name", None) or os.environ.get("ORACLE_LIB", "liboracle.so")
    _lib = _load(name)
    return prev


def _load(name):
    if name in _loaded:
        return _loaded[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run make -C oracle")
    L = C.CDLL(path)
    L._oracle_name = name
    L.orc_zeeman.argtypes = [C.c_double] * 6 + [C.POINTER(Zeeman)]
    L.orc_gauss_psf.argtypes = [C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, _dp]
    L.orc_nterms.argtypes = [_ip, C.c_double]
    L.orc_fvoigt.argtypes = [C.c_double, _dp, C.c_int, _dp, _dp]
    L.orc_gauss_mac.argtypes = [C.c_double, _dp, C.c_int, C.c_double,
                                C.c_int, _dp]
    L.orc_conv_direct.argtypes = [_dp, C.c_int, _dp, C.c_int]
    L.orc_conv_direct_ic.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_double]
    L.orc_conv_circular.argtypes = [_dp, _dp, C.c_int, _dp]
    L.orc_synth.argtypes = [C.POINTER(Setup), _dp, _dp, C.c_void_p]
    L.orc_covarm.argtypes = [C.POINTER(Setup), _dp, _fp, _dp, _dp, _dp, _dp]
    L.orc_chisqr.argtypes = [C.POINTER(Setup), _dp, _fp, _dp, C.c_double]
    L.orc_chisqr.restype = C.c_double
    L.orc_pinv_solve.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_eigen_sym.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_aplica_delta.argtypes = [C.POINTER(Setup), _dp, _dp, _dp]
    L.orc_check.argtypes = [_dp]
    L.orc_fija_cero.argtypes = [C.POINTER(Setup), _dp]
    L.orc_classical.argtypes = [C.POINTER(Setup), _fp, _dp]
    L.orc_classical_ex.argtypes = [C.POINTER(Setup), _fp, _dp, C.c_int]
    L.orc_lm.argtypes = [C.POINTER(Setup), _fp, _dp, C.c_void_p,
                         C.POINTER(C.c_float), C.POINTER(C.c_int),
                         C.POINTER(C.c_int)]
    L.orc_lm.restype = C.c_int
    L.orc_invert_range.argtypes = [C.POINTER(Setup), _fp, C.c_int, C.c_int,
                                   _dp, _dp, _fp, _ip, _ip, C.c_void_p]
    _loaded[name] = L
    return L


def ref():
    """The four reference numeric files built in oracle/_ref, or None."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref_core.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_fvoigt.argtypes = [C.c_double, _dp, C.c_int, _dp, _dp]
        R.ref_cuantic.argtypes = [C.c_double] * 6 + [_dp, C.c_int]
        R.ref_fgauss_wl.argtypes = [C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, _dp]
        R.ref_fgauss.argtypes = [C.c_double, _dp, C.c_int, C.c_double, C.c_int, _dp]
        R.ref_direct.argtypes = [_dp, C.c_int, _dp, C.c_int]
        R.ref_direct_double.argtypes = [_dp, C.c_int, _dp, C.c_int]
        R.ref_direct_ic.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_double]
        R.ref_circular.argtypes = [_dp, _dp, C.c_int, _dp]
        _ref = R
    return _ref


# ---------------------------------------------------------------- setup

def zeeman(sl, ll, jl, su, lu, ju):
    z = Zeeman()
    rc = lib().orc_zeeman(sl, ll, jl, su, lu, ju, C.byref(z))
    if rc != 0:
        raise ValueError("too many Zeeman components")
    return z


def grid(lambda0, initial_mA, step_mA, final_mA):
    """Wavelength axis as milos.c:173-197 builds it (accumulated steps)."""
    n = int((final_mA - initial_mA) / step_mA) + 1
    lam = np.empty(n)
    lam[0] = lambda0 + initial_mA / 1000.0
    step = step_mA / 1000.0
    for i in range(1, n):
        lam[i] = lam[i - 1] + step
    return lam


def gauss_psf(fwhm_mA, lam):
    """PSF kernel as milos.c:269 calls fgauss_WL."""
    n = len(lam)
    G = np.zeros(n)
    lib().orc_gauss_psf(fwhm_mA, lam[1] - lam[0], lam[0], lam[n // 2], n, G)
    return G


def make_setup(lambda0, quantum, lam, psf=None, fix=None, weight=(1, 1, 1, 1),
               noise=1e-6, ilambda=float(np.float32(0.1)), toplim=1e-12,
               miter=50, ah=1.0, logclambda=0, guess_mac=0.0, slight=None, use_fft=False):
    """quantum = (sl, ll, jl, su, lu, ju) as readFileCuanticLines returns."""
    s = Setup()
    L = len(lam)
    assert L <= MAXL
    s.nlambda = L
    s.lambda0 = lambda0
    for i in range(L):
        s.lam[i] = lam[i]
    if psf is not None:
        s.use_psf = 1
        for i in range(L):
            s.G[i] = psf[i]
    s.z = zeeman(*quantum)
    if fix is None:
        fix = [1] * 9 + [0, 0]
    for i in range(NPAR):
        s.fix[i] = int(fix[i])
    s.nterms = lib().orc_nterms(np.asarray(fix, dtype=np.int32), guess_mac)
    for i in range(4):
        s.weight[i] = weight[i]
    s.noise = noise
    s.ilambda = ilambda
    s.toplim = toplim
    s.miter = miter
    s.ah = ah
    s.logclambda = logclambda
    s.use_fft = int(bool(use_fft))
    if slight is not None:
        s.use_slight = 1
        sl = np.asarray(slight, dtype=np.float64).ravel()
        for i in range(4 * L):
            s.slight[i] = sl[i]
    return s


# ---------------------------------------------------------------- calls

def fvoigt(a, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    H = np.zeros_like(v)
    F = np.zeros_like(v)
    lib().orc_fvoigt(a, v, len(v), H, F)
    return H, F


def synth(s, model, rf=False):
    L, NT = s.nlambda, s.nterms
    m = np.ascontiguousarray(model, dtype=np.float64)
    spec = np.zeros(4 * L)
    if rf:
        d = np.zeros(4 * NT * L)
        lib().orc_synth(C.byref(s), m, spec, d.ctypes.data)
        return spec.reshape(4, L), d.reshape(4, NT, L)
    lib().orc_synth(C.byref(s), m, spec, None)
    return spec.reshape(4, L)


def classical(s, spectro, model, for_init=True):
    m = np.array(model, dtype=np.float64)
    lib().orc_classical_ex(C.byref(s), np.ascontiguousarray(spectro, dtype=np.float32).ravel(),
                           m, int(for_init))
    return m


def lm(s, spectro, model):
    m = np.array(model, dtype=np.float64)
    fit = np.zeros(4 * s.nlambda)
    c2 = C.c_float(0)
    it = C.c_int(1)
    acc = C.c_int(0)
    rc = lib().orc_lm(C.byref(s), np.ascontiguousarray(spectro, dtype=np.float32).ravel(),
                      m, fit.ctypes.data, C.byref(c2), C.byref(it), C.byref(acc))
    return dict(rc=rc, model=m, fit=fit.reshape(4, s.nlambda), chi2=c2.value,
                iters=it.value, accepted=acc.value)


def invert(s, stokes, guess, p0=0, p1=None, want_fit=False):
    """stokes: float32 [npix, 4, L].  Returns dict of per-pixel results."""
    st = np.ascontiguousarray(stokes, dtype=np.float32)
    npix = st.shape[0]
    if p1 is None:
        p1 = npix
    models = np.zeros((npix, NPAR))
    chi2 = np.zeros(npix, dtype=np.float32)
    iters = np.zeros(npix, dtype=np.int32)
    acc = np.zeros(npix, dtype=np.int32)
    fit = np.zeros((npix, 4, s.nlambda), dtype=np.float32) if want_fit else None
    lib().orc_invert_range(C.byref(s), st.ravel(), p0, p1,
                           np.ascontiguousarray(guess, dtype=np.float64), models.ravel(),
                           chi2, iters, acc, fit.ctypes.data if want_fit else None)
    return dict(model=models, chi2=chi2, iters=iters, accepted=acc, fit=fit)
