"""Python binding of libpmilos_hip.so (the C ABI of include/pmilos.h).

Host-side mirror of the reference's per-pixel interface: a Context holds what
milos.c sets up once per run (line, Zeeman pattern, wavelength grid, PSF,
control settings); Context.invert replaces the FITS pixel loop around lm_mils
(milos.c:958-1009) and Context.synth the synthesis mode (milos.c:601-602).

The HIP library is required: importing works without it, but every call
raises if libpmilos_hip.so is missing (there is no CPU fallback).

With torch in the same process, import torch before the first call here.
torch ships its own HIP runtime; loaded after the library's (ROCm's
libamdhip64.so.7), torch's finds no GPU.  Loaded first, both work on one
device address space (torch-allocated buffers go straight to the *_device
and *_ptrs calls).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmilos_hip.so")
NPAR = 11
PM_FP64, PM_MIXED = 0, 1
# the mixed form's default chi2-unchanged threshold (PM_MIXED_TOPLIM)
PM_MIXED_TOPLIM = 1e-4
PARAM_NAMES = ["eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa"]

_lib = None


class PmError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [("nlambda", C.c_int), ("lam", C.POINTER(C.c_double)),
                ("lambda0", C.c_double),
                ("sl", C.c_double), ("ll", C.c_double), ("jl", C.c_double),
                ("su", C.c_double), ("lu", C.c_double), ("ju", C.c_double),
                ("psf", C.POINTER(C.c_double)),
                ("fix", C.c_int * NPAR), ("weight", C.c_double * 4),
                ("noise", C.c_double), ("ilambda", C.c_double), ("toplim", C.c_double),
                ("miter", C.c_int), ("mu", C.c_double), ("logclambda", C.c_int),
                ("guess", C.c_double * NPAR), ("precision", C.c_int),
                ("slight", C.POINTER(C.c_float)), ("no_classical", C.c_int),
                ("use_fft", C.c_int)]


class BatchItem(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("d_stokes", C.c_void_p), ("d_slight", C.c_void_p),
                ("npix", C.c_int), ("d_models", C.c_void_p), ("d_chi2", C.c_void_p),
                ("d_iters", C.c_void_p), ("d_accepted", C.c_void_p), ("d_fit", C.c_void_p)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("streams", C.c_int), ("chunk_px", C.c_int)]


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PmError("libpmilos_hip.so not built (run p-milos_amd/pmilos/build.py)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pm_strerror.restype = C.c_char_p
        L.pm_strerror.argtypes = [C.c_int]
        L.pm_grid.argtypes = [C.c_double] * 4 + [vp, C.c_int]
        L.pm_psf_gaussian.argtypes = [C.c_double, vp, C.c_int, vp]
        L.pm_nterms.argtypes = [vp, C.c_double]
        L.pm_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.pm_ctx_destroy.argtypes = [vp]
        L.pm_ctx_destroy.restype = None
        L.pm_ctx_nterms.argtypes = [vp]
        L.pm_synth.argtypes = [vp, vp, C.c_int, vp, vp]
        L.pm_synth_device.argtypes = [vp, vp, C.c_int, vp, vp, vp]
        L.pm_invert.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp]
        L.pm_invert_device.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.pm_invert_device_noalloc.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.pm_invert_sl.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp, vp]
        L.pm_invert_sl_device.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.pm_last_fallback_count.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.pm_ctx_autotune.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_int * 3)]
        L.pm_ctx_variant.argtypes = [vp, C.c_int, C.POINTER(C.c_int * 3)]
        L.pm_ctx_set_variant.argtypes = [vp, C.c_int, C.c_int]
        L.pm_ctx_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.pm_ctx_set_mixed_toplim.argtypes = [vp, C.c_double]
        L.pm_graph_create.argtypes = [C.POINTER(BatchItem), C.c_int, C.c_double, C.POINTER(vp)]
        L.pm_graph_launch.argtypes = [vp, vp]
        L.pm_graph_destroy.argtypes = [vp]
        L.pm_graph_destroy.restype = None
        L.pm_classical.argtypes = [vp, vp, C.c_int, vp]
        L.pm_classical_device.argtypes = [vp, vp, C.c_int, vp, vp]
        L.pm_server_create.argtypes = [C.POINTER(ServerOpts), C.POINTER(vp)]
        L.pm_server_destroy.argtypes = [vp]
        L.pm_server_destroy.restype = None
        L.pm_submit.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp,
                                C.POINTER(C.c_longlong)]
        L.pm_poll.argtypes = [vp, C.c_longlong]
        L.pm_wait.argtypes = [vp, C.c_longlong, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise PmError("pmilos: %s (%d)" % (lib().pm_strerror(rc).decode(), rc))
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data


def grid(lambda0, initial_mA, step_mA, final_mA):
    n = _check(lib().pm_grid(lambda0, initial_mA, step_mA, final_mA, None, 0))
    lam = np.zeros(n)
    lib().pm_grid(lambda0, initial_mA, step_mA, final_mA, lam.ctypes.data, n)
    return lam


def psf_gaussian(fwhm_mA, lam):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    G = np.zeros(len(lam))
    _check(lib().pm_psf_gaussian(fwhm_mA, lam.ctypes.data, len(lam), G.ctypes.data))
    return G


class Context:
    """Per-run state on the device (pm_ctx)."""

    def __init__(self, lambda0, quantum, lam, psf=None, fix=None, weight=(1, 1, 1, 1),
                 noise=1e-6, ilambda=float(np.float32(0.1)), toplim=1e-12, miter=50,
                 mu=1.0, logclambda=0, guess=None, precision=PM_FP64, slight=None,
                 classical=True, use_fft=False):
        self._lam = np.ascontiguousarray(lam, dtype=np.float64)
        self._slight = None if slight is None else \
            np.ascontiguousarray(slight, dtype=np.float32).reshape(4, len(self._lam))
        self._psf = None if psf is None else np.ascontiguousarray(psf, dtype=np.float64)
        cfg = Config()
        cfg.nlambda = len(self._lam)
        cfg.lam = self._lam.ctypes.data_as(C.POINTER(C.c_double))
        cfg.lambda0 = lambda0
        cfg.sl, cfg.ll, cfg.jl, cfg.su, cfg.lu, cfg.ju = [float(q) for q in quantum]
        if self._psf is not None:
            cfg.psf = self._psf.ctypes.data_as(C.POINTER(C.c_double))
        fix = [1] * 9 + [0, 0] if fix is None else fix
        for i in range(NPAR):
            cfg.fix[i] = int(fix[i])
        for i in range(4):
            cfg.weight[i] = weight[i]
        cfg.noise, cfg.ilambda, cfg.toplim = noise, ilambda, toplim
        cfg.miter, cfg.mu, cfg.logclambda = miter, mu, logclambda
        guess = np.zeros(NPAR) if guess is None else guess
        for i in range(NPAR):
            cfg.guess[i] = float(guess[i])
        cfg.precision = precision
        if self._slight is not None:
            cfg.slight = self._slight.ctypes.data_as(C.POINTER(C.c_float))
        cfg.no_classical = 0 if classical else 1
        cfg.use_fft = 1 if use_fft else 0
        h = C.c_void_p()
        _check(lib().pm_ctx_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.L = len(self._lam)
        self.NT = lib().pm_ctx_nterms(h)

    def close(self):
        if getattr(self, "h", None):
            lib().pm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synth(self, models, rf=False):
        m = np.ascontiguousarray(np.atleast_2d(models), dtype=np.float64)
        n = m.shape[0]
        spec = np.zeros((n, 4, self.L))
        d = np.zeros((n, 4, self.NT, self.L)) if rf else None
        _check(lib().pm_synth(self.h, m.ctypes.data, n, spec.ctypes.data, _ptr(d)))
        return (spec, d) if rf else spec

    def invert(self, stokes, want_fit=False, slight=None, out=None):
        """pm_invert_sl on host arrays.  slight: optional per-pixel stray
        light [n][4][L] (else the context's profile, if any).  out: a
        result_buffers() set to write into (pinned host memory copies back
        fastest), else one is allocated."""
        st = np.ascontiguousarray(stokes, dtype=np.float32)
        n = st.shape[0]
        sl = None if slight is None else np.ascontiguousarray(slight, dtype=np.float32)
        if sl is not None and sl.shape != st.shape:
            raise PmError("slight must have the shape of stokes")
        out = _result_set(n, self.L, want_fit, out)
        _check(lib().pm_invert_sl(self.h, st.ctypes.data, _ptr(sl), n, out["model"].ctypes.data,
                                  out["chi2"].ctypes.data, out["iters"].ctypes.data,
                                  out["accepted"].ctypes.data, _ptr(out["fit"])))
        return out

    def invert_ptrs(self, d_stokes, npix, d_models, d_chi2, d_iters, d_accepted=0, d_fit=0,
                    stream=0, d_slight=0):
        if d_slight:
            _check(lib().pm_invert_sl_device(self.h, d_stokes, d_slight, npix, d_models, d_chi2,
                                             d_iters, d_accepted or None, d_fit or None,
                                             stream or None))
        else:
            _check(lib().pm_invert_device(self.h, d_stokes, npix, d_models, d_chi2, d_iters,
                                          d_accepted or None, d_fit or None, stream or None))

    def invert_ptrs_noalloc(self, d_stokes, npix, d_models, d_chi2, d_iters, d_accepted=0,
                            d_fit=0, stream=0):
        """pm_invert_device_noalloc: as invert_ptrs, with the library's fixed
        device scratch instead of the context's counters (allocates nothing;
        calls overlapping on one GPU race)."""
        _check(lib().pm_invert_device_noalloc(self.h, d_stokes, npix, d_models, d_chi2, d_iters,
                                              d_accepted or None, d_fit or None, stream or None))

    def synth_ptrs(self, d_models, n, d_spectra, d_rf=0, stream=0):
        _check(lib().pm_synth_device(self.h, d_models, n, d_spectra, d_rf or None,
                                     stream or None))

    def classical(self, stokes):
        """Classical estimates only: models [n][11] (B_los in the S0 slot)."""
        st = np.ascontiguousarray(stokes, dtype=np.float32)
        models = np.zeros((st.shape[0], NPAR))
        _check(lib().pm_classical(self.h, st.ctypes.data, st.shape[0], models.ctypes.data))
        return models

    def autotune_ptrs(self, d_stokes, npix, stream=0, d_slight=0):
        """Times the kernel variants on npix pixels at device address
        d_stokes and keeps the fastest: (lanes per pixel, samples per lane,
        pixels per workgroup)."""
        out = (C.c_int * 3)()
        _check(lib().pm_ctx_autotune(self.h, d_stokes, d_slight or None, npix, stream or None,
                                     C.byref(out)))
        return tuple(out)

    def variant(self, general=False):
        """((lanes per pixel, samples per lane, pixels per workgroup), number
        of variants) of the form in use."""
        out = (C.c_int * 3)()
        n = _check(lib().pm_ctx_variant(self.h, int(general), C.byref(out)))
        return tuple(out), n

    def set_variant(self, index, general=False):
        _check(lib().pm_ctx_set_variant(self.h, int(general), index))

    def set_shape(self, lanes, samples, pixels_per_workgroup, general=False):
        """Any compiled kernel shape that covers the sampling (pm_ctx_set_shape)."""
        _check(lib().pm_ctx_set_shape(self.h, int(general), lanes, samples,
                                      pixels_per_workgroup))

    def set_mixed_toplim(self, toplim_mixed):
        """The mixed form's chi2-unchanged threshold (reference units:
        |dchi2| * 100 / chi2); mixed contexts start at PM_MIXED_TOPLIM, 0
        restores the reference's toplim test."""
        _check(lib().pm_ctx_set_mixed_toplim(self.h, C.c_double(toplim_mixed)))

    def fallback_count(self):
        v = C.c_longlong(0)
        _check(lib().pm_last_fallback_count(self.h, C.byref(v)))
        return v.value


def result_buffers(n, L=0, want_fit=False, alloc=np.empty):
    """A result set for n pixels (the `out` of Context.invert and
    Server.submit): model [n, 11] f64, chi2 f32, iters and accepted i32, and
    with want_fit the fitted profiles [n, 4, L] f32.  alloc(shape, dtype)
    makes each array (e.g. one that returns pinned host memory)."""
    return dict(model=alloc((n, NPAR), np.float64), chi2=alloc((n,), np.float32),
                iters=alloc((n,), np.int32), accepted=alloc((n,), np.int32),
                fit=alloc((n, 4, L), np.float32) if want_fit else None)


def _result_set(n, L, want_fit, out):
    """out checked against result_buffers(n, L, want_fit) (ValueError if an
    array does not match), or a fresh zeroed set when out is None."""
    if out is None:
        return result_buffers(n, L, want_fit, np.zeros)
    want = result_buffers(n, L, want_fit, lambda sh, dt: (sh, dt))
    for k, v in want.items():
        a = out.get(k)
        if v is None:
            continue
        if (not isinstance(a, np.ndarray) or a.shape != v[0] or a.dtype != v[1]
                or not a.flags.c_contiguous or not a.flags.writeable):
            raise ValueError("out[%r]: need a writable C-contiguous %s array of shape %s"
                             % (k, np.dtype(v[1]).name, v[0]))
    return out if want_fit else dict(out, fit=None)


class Server:
    """Serving loop on the device (pm_server): scans are submitted as they
    arrive, share the GPU, and are collected with wait()."""

    def __init__(self, hbm_budget=0, streams=0, chunk_px=0):
        o = ServerOpts(hbm_budget, streams, chunk_px)
        h = C.c_void_p()
        _check(lib().pm_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self._pending = {}

    result_buffers = staticmethod(lambda *a, **k: result_buffers(*a, **k))

    def submit(self, ctx, stokes, slight=None, want_fit=False, out=None):
        """Queues a scan; returns a ticket.  The results go into `out` (a
        result_buffers() set of the scan's size; pinned host memory copies
        back fastest) or into buffers allocated here; either is kept with
        the ticket and returned by wait()."""
        st = np.ascontiguousarray(stokes, dtype=np.float32)
        n = st.shape[0]
        sl = None if slight is None else np.ascontiguousarray(slight, dtype=np.float32)
        out = _result_set(n, ctx.L, want_fit, out)
        t = C.c_longlong(0)
        _check(lib().pm_submit(self.h, ctx.h, st.ctypes.data, _ptr(sl), n,
                               out["model"].ctypes.data, out["chi2"].ctypes.data,
                               out["iters"].ctypes.data, out["accepted"].ctypes.data,
                               _ptr(out["fit"]), C.byref(t)))
        self._pending[t.value] = (st, sl, out)
        return t.value

    def poll(self, ticket):
        return _check(lib().pm_poll(self.h, ticket)) == 1

    def wait(self, ticket):
        """Blocks until the scan is done: (results dict, latency in s).
        PmError for a ticket already collected."""
        lat = C.c_double(0)
        rc = lib().pm_wait(self.h, ticket, C.byref(lat))
        entry = self._pending.pop(ticket, None)
        _check(rc)
        if entry is None:
            raise PmError("ticket %d has no result set here" % ticket)
        return entry[2], lat.value

    def close(self):
        if getattr(self, "h", None):
            lib().pm_server_destroy(self.h)
            self.h = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Several inversions on device buffers captured as one HIP graph
    (pm_graph): launch() runs them all with one graph launch.  items: dicts
    with ctx, d_stokes, npix, d_models, d_chi2, d_iters and optionally
    d_slight, d_accepted, d_fit (device addresses)."""

    def __init__(self, items, share_by_work=0.0):
        arr = (BatchItem * len(items))()
        self._ctxs = [it["ctx"] for it in items]
        for a, it in zip(arr, items):
            a.ctx = it["ctx"].h.value
            a.d_stokes = it["d_stokes"]
            a.d_slight = it.get("d_slight") or None
            a.npix = it["npix"]
            a.d_models = it["d_models"]
            a.d_chi2 = it["d_chi2"]
            a.d_iters = it["d_iters"]
            a.d_accepted = it.get("d_accepted") or None
            a.d_fit = it.get("d_fit") or None
        h = C.c_void_p()
        _check(lib().pm_graph_create(arr, len(items), share_by_work, C.byref(h)))
        self.h = h

    def launch(self, stream=0):
        _check(lib().pm_graph_launch(self.h, stream or None))

    def close(self):
        if getattr(self, "h", None):
            lib().pm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
