"""The file surface of the command line tool (libpmilos_io.so, host code):
control files, line list, grids and models against the reference's own
fixtures; PSF interpolation against numpy; FITS cubes in every axis order
against an independent numpy writer.  CPU only."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

import fitsutil
import oracle as O
import pmilos
from fixtures import FE6173, FE6302, GOLDEN, GUESS, read_fits_2d, read_mod

# PMIO_LIB: another build of the same library (the sanitizer run of
# tests/test_host_sanitizers.py)
IO_LIB = os.environ.get("PMIO_LIB") or os.path.join(os.path.dirname(pmilos.LIB_PATH),
                                                    "libpmilos_io.so")
P = 4096


class Control(C.Structure):
    _fields_ = [("cycles", C.c_int), ("observed", C.c_char * P), ("straylight", C.c_char * P),
                ("psf", C.c_char * P), ("fwhm", C.c_double), ("use_psf", C.c_int),
                ("psf_is_file", C.c_int), ("wavelength", C.c_char * P), ("grid", C.c_char * P),
                ("use_grid", C.c_int), ("lines", C.c_char * P), ("guess", C.c_char * P),
                ("weight", C.c_double * 4), ("fix", C.c_int * 11), ("invert_filling", C.c_int),
                ("mu", C.c_double), ("sn", C.c_int), ("noise", C.c_double),
                ("ilambda", C.c_double), ("use_fft", C.c_int), ("logclambda", C.c_int),
                ("control", C.c_char * P), ("subx1", C.c_int), ("subx2", C.c_int),
                ("suby1", C.c_int), ("suby2", C.c_int), ("prefix", C.c_char * P),
                ("mask", C.c_char * P), ("t1", C.c_int), ("t2", C.c_int),
                ("invert_directory", C.c_int), ("loop_inversion", C.c_int),
                ("save_synthesis", C.c_int)]


class Layout(C.Structure):
    _fields_ = [("naxis", C.c_int), ("naxes", C.c_long * 4), ("pos_lambda", C.c_int),
                ("pos_stokes", C.c_int), ("pos_row", C.c_int), ("pos_col", C.c_int),
                ("rows", C.c_int), ("cols", C.c_int), ("nlambda", C.c_int), ("ncards", C.c_int),
                ("cards", C.c_void_p)]


@pytest.fixture(scope="module")
def io():
    if not os.path.exists(IO_LIB):
        pytest.fail("libpmilos_io.so not built: run __graft_entry__.build()")
    L = C.CDLL(IO_LIB)
    vp = C.c_void_p
    L.pmio_read_lines.restype = C.c_double
    L.pmio_read_lines.argtypes = [C.c_char_p, C.c_int, vp]
    L.pmio_read_grid.argtypes = [C.c_char_p, vp, vp, vp]
    L.pmio_read_model.argtypes = [C.c_char_p, vp]
    L.pmio_interp_psf.argtypes = [vp, vp, C.c_int, vp, C.c_int, C.c_double, vp]
    L.pmio_read_psf_file.argtypes = [C.c_char_p, vp, vp, C.c_int]
    L.pmio_read_per.argtypes = [C.c_char_p, C.c_int, vp, vp]
    L.pmio_write_per.argtypes = [C.c_char_p, C.c_int, vp, vp, C.c_int]
    L.pmio_fits_read_cube.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.POINTER(C.c_float)), C.POINTER(Layout)]
    L.pmio_fits_write_cube.argtypes = [C.c_char_p, C.POINTER(Layout), vp]
    L.pmio_fits_write_models.argtypes = [C.c_char_p, C.c_int, C.c_int, vp, vp, vp]
    L.pmio_fits_read_2d.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.pmio_fits_read_straylight.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.POINTER(C.POINTER(C.c_float)),
                                            C.POINTER(C.c_int), C.POINTER(C.c_long)]
    L.pmio_fits_read_lambda.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_double)),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pmio_free.argtypes = [vp]
    L.pmio_free_layout.argtypes = [C.POINTER(Layout)]
    return L


@pytest.fixture
def golden_cwd(tmp_path, monkeypatch):
    for f in os.listdir(GOLDEN):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    monkeypatch.chdir(tmp_path)
    return tmp_path


def test_mtrol_pmilos_run(io, golden_cwd):
    c = Control()
    assert io.pmio_read_mtrol(b"pmilos.mtrol", C.byref(c)) == 0
    assert c.cycles == 50
    assert c.observed == b"data/2014.09.28_09:18:00_xtalk_t"
    assert not c.use_psf                       # CRISP_6173_1mA.psf is not there: atof -> 0
    assert c.use_grid and c.grid == b"wavelength.grid"
    assert c.lines == b"LINES" and c.guess == b"guess.mod"
    assert list(c.weight) == [1, 1, 1, 1]
    assert list(c.fix) == [1] * 9 + [0, 0]
    assert c.sn == 1000 and c.noise == pytest.approx(1e-6, rel=1e-15)
    assert c.ilambda == float(np.float32(0.1))
    assert c.logclambda == 1 and c.use_fft == 0 and c.mu == 1.0
    assert c.save_synthesis == 1


def test_mtrol_synthesis_and_blank_fields(io, golden_cwd):
    c = Control()
    assert io.pmio_read_mtrol(b"synthesis.mtrol", C.byref(c)) == 0
    assert c.cycles == 0 and c.observed == b"obs.per"
    assert c.straylight == b"straylight.fits"
    assert not c.use_psf and c.psf == b"!"     # a blank string field reads as "!"
    assert c.grid == b"malla.grid"
    assert c.fix[10] == 1 and c.logclambda == 0
    # a PSF given as a FWHM in mA
    txt = open("synthesis.mtrol").read().replace("PSF file                     :    ",
                                                 "PSF file                     :60  ")
    open("fwhm.mtrol", "w").write(txt)
    c2 = Control()
    assert io.pmio_read_mtrol(b"fwhm.mtrol", C.byref(c2)) == 0
    assert c2.use_psf and c2.fwhm == 60.0 and not c2.psf_is_file
    # a short file keeps the defaults of the lines it lacks
    open("short.mtrol", "w").write("".join(open("synthesis.mtrol").readlines()[:12]))
    c3 = Control()
    assert io.pmio_read_mtrol(b"short.mtrol", C.byref(c3)) == -2
    assert c3.cycles == 0 and c3.noise == pytest.approx(1e-6) and list(c3.fix) == [0] * 11


def test_minit(io, golden_cwd):
    c = Control()
    assert io.pmio_read_minit(b"pmilos.minit", C.byref(c)) == 0
    assert c.control == b"pmilos.mtrol" and c.cycles == 50
    assert (c.subx1, c.subx2, c.suby1, c.suby2) == (0, 0, 0, 0)
    assert c.prefix == b"results/inv_" and c.mask == b"!"
    assert (c.t1, c.t2, c.invert_directory, c.loop_inversion) == (20, 21, 0, 0)
    assert c.save_synthesis == 0
    txt = open("pmilos.minit").read()
    txt = txt.replace("t1                           :20", "t1                           :*")
    txt = txt.replace("t2                           :21", "t2                           :*")
    txt = txt.replace("subx1                        :0", "subx1                        :3")
    open("w.minit", "w").write(txt)
    c2 = Control()
    assert io.pmio_read_minit(b"w.minit", C.byref(c2)) == -4       # subfield bounds incomplete
    assert c2.invert_directory == 1 and c2.loop_inversion == 1


def test_lines_grid_model(io, golden_cwd):
    dat = np.zeros(7)
    assert io.pmio_read_lines(b"LINES_run", 1, dat.ctypes.data) == FE6173
    assert list(dat) == [1, 2, 1, 1.0, 2, 2, 0.0]               # 5P1 - 5D0
    assert io.pmio_read_lines(b"LINES_run", 3, dat.ctypes.data) == FE6302
    assert io.pmio_read_lines(b"LINES_run", 2, dat.ctypes.data) == 6301.5012
    assert list(dat) == [1, 2, 1, 2.0, 2, 2, 2.0]               # 5P2 - 5D2
    assert io.pmio_read_lines(b"LINES_run", 42, dat.ctypes.data) == 0
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    assert io.pmio_read_grid(b"malla.grid", C.byref(a), C.byref(b), C.byref(c)) == 1
    assert (a.value, b.value, c.value) == (-600, 25, 600)
    assert io.pmio_read_grid(b"wavelength.grid", C.byref(a), C.byref(b), C.byref(c)) == 1
    assert (a.value, b.value, c.value) == (-350, 35, 665)
    m = np.zeros(11)
    assert io.pmio_read_model(b"guess.mod", m.ctypes.data) == 0
    assert np.array_equal(m, GUESS)
    m = np.zeros(11)
    assert io.pmio_read_model(b"obs.mod", m.ctypes.data) == 0
    assert np.array_equal(m, read_mod("obs.mod"))
    open("bad.mod", "w").write(open("guess.mod").read().replace(":1100", ":6000"))
    assert io.pmio_read_model(b"bad.mod", np.zeros(11).ctypes.data) == 1 + 1


def test_psf_interpolation_is_linear(io):
    rng = np.random.default_rng(3)
    dl = np.cumsum(rng.uniform(0.5, 3.0, 300)) - 400.0
    psf = np.exp(-0.5 * (dl / 40.0) ** 2) + rng.uniform(0, 1e-3, dl.size)
    off = np.arange(-290.0, 291.0, 20.0)
    for shift in (0.0, 10.0, -7.5):
        G = np.zeros(off.size)
        assert io.pmio_interp_psf(dl.ctypes.data, psf.ctypes.data, dl.size, off.ctypes.data,
                                  off.size, shift, G.ctypes.data) == 0
        want = np.interp(off - shift, dl, psf, left=0.0, right=0.0)
        np.testing.assert_allclose(G, want / want.sum(), rtol=1e-13, atol=1e-16)


def crisp_psf_on_grid(initial, step, final):
    """The reference's instrument profile run/CRISP_6173_25mA.psf (49
    samples, mA, unit peak) on a wavelength grid, recentred as milos.c:331-340
    does when the grid's zero offset is not its middle sample (shifted by
    (L/2 - pos) steps) and normalised, by numpy's linear interpolation:
    (offsets mA, shift mA, G)."""
    dl, ps = np.loadtxt(os.path.join(GOLDEN, "CRISP_6173_25mA.psf")).T
    lam = O.grid(FE6173, initial, step, final)
    off = (lam - FE6173) * 1000.0
    pos = int(np.flatnonzero(np.trunc(off) == 0)[0])
    shift = (lam.size // 2 - pos) * step
    G = np.interp(off - shift, dl, ps, left=0.0, right=0.0)
    return off, shift, G / G.sum()


def test_psf_file_recentred_like_the_reference(io):
    """The file's samples, read by pmio_read_psf_file and interpolated by
    pmio_interp_psf with the recentring shift, against numpy; the peak
    lands on the grid's middle sample."""
    dl = np.zeros(100)
    ps = np.zeros(100)
    n = io.pmio_read_psf_file(os.path.join(GOLDEN, "CRISP_6173_25mA.psf").encode(),
                              dl.ctypes.data, ps.ctypes.data, 100)
    want_dl, want_ps = np.loadtxt(os.path.join(GOLDEN, "CRISP_6173_25mA.psf")).T
    assert n == 49 and np.array_equal(dl[:n], want_dl) and np.array_equal(ps[:n], want_ps)
    off, shift, want = crisp_psf_on_grid(-300, 25, 500)
    assert off.size == 33 and shift == 100.0
    G = np.zeros(off.size)
    assert io.pmio_interp_psf(dl.ctypes.data, ps.ctypes.data, n, off.ctypes.data, off.size,
                              shift, G.ctypes.data) == 0
    np.testing.assert_allclose(G, want, rtol=1e-13, atol=1e-17)
    assert int(np.argmax(G)) == off.size // 2


def test_per_round_trip(io, tmp_path):
    L = 30
    spec = np.random.default_rng(4).normal(size=(4, L))
    off = np.arange(L) * 20.0 - 290.0
    p = str(tmp_path / "x.per").encode()
    assert io.pmio_write_per(p, 1, off.ctypes.data, spec.ctypes.data, L) == 0
    back = np.zeros((4, L), dtype=np.float32)
    offb = np.zeros(L)
    assert io.pmio_read_per(p, L, back.ctypes.data, offb.ctypes.data) == L
    np.testing.assert_allclose(back, spec, rtol=1e-6)            # %e keeps 7 digits
    np.testing.assert_allclose(offb, off, atol=1e-4)


def _read_cube(io, path, L, sub=(0, 0, 0, 0)):
    ptr = C.POINTER(C.c_float)()
    lay = Layout()
    rc = io.pmio_fits_read_cube(path.encode(), L, *sub, C.byref(ptr), C.byref(lay))
    if rc:
        return rc, None, None
    n = lay.rows * lay.cols
    out = np.ctypeslib.as_array(ptr, shape=(n, 4, L)).copy()
    io.pmio_free(ptr)
    return 0, out, lay


ORDERS = [("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"),
          ("HPLN-TAN", "HPLT-TAN", "WAVE-GRI", "STOKES"),
          ("STOKES", "HPLT-TAN", "WAVE-GRI", "HPLN-TAN"),
          ("HPLT-TAN", "WAVE-GRI", "HPLN-TAN", "STOKES")]


@pytest.mark.parametrize("order", ORDERS)
def test_fits_cube_any_axis_order(io, tmp_path, order):
    rows, cols, L = 7, 5, 12
    cube = np.random.default_rng(5).normal(size=(rows * cols, 4, L)).astype(np.float32)
    p = str(tmp_path / "c.fits")
    fitsutil.write(p, fitsutil.stokes_cube(cube, rows, cols, order), ctypes=order,
                   extra=["OBSERVER= 'test'"])
    rc, got, lay = _read_cube(io, p, L)
    assert rc == 0 and (lay.rows, lay.cols) == (rows, cols)
    assert np.array_equal(got, cube)
    # subfield rows 2..5, cols 2..4 (1-based, inclusive)
    rc, sub, lay2 = _read_cube(io, p, L, (2, 5, 2, 4))
    assert rc == 0 and (lay2.rows, lay2.cols) == (4, 3)
    want = cube.reshape(cols, rows, 4, L)[1:4, 1:5].reshape(-1, 4, L)
    assert np.array_equal(sub, want)
    # the fit cube goes back in the input's axis order, its cards kept
    q = str(tmp_path / "o.fits")
    assert io.pmio_fits_write_cube(q.encode(), C.byref(lay), got.ctypes.data) == 0
    data, hdr = fitsutil.read(q)
    orig, _ = fitsutil.read(p)
    assert np.array_equal(data, orig)
    assert [hdr["CTYPE%d" % (k + 1)] for k in range(4)] == list(order)
    assert hdr["OBSERVER"] == "test"
    io.pmio_free_layout(C.byref(lay))
    io.pmio_free_layout(C.byref(lay2))


def test_fits_cube_without_ctypes_and_errors(io, tmp_path):
    rows, cols, L = 9, 6, 5
    cube = np.random.default_rng(6).normal(size=(rows * cols, 4, L)).astype(np.float32)
    p = str(tmp_path / "c.fits")
    # no CTYPEs: the smallest axis is Stokes, the next the wavelengths
    fitsutil.write(p, fitsutil.stokes_cube(cube, rows, cols))
    rc, got, lay = _read_cube(io, p, L)
    assert rc == 0 and np.array_equal(got, cube)
    assert _read_cube(io, p, L + 1)[0] == -4
    assert _read_cube(io, p, L, (1, 2, 0, 0))[0] == -5
    fitsutil.write(p, fitsutil.stokes_cube(cube, rows, cols), bitpix=-64)
    assert _read_cube(io, p, L)[0] == -2


def test_fits_cube_scaled_float(io, tmp_path):
    """A BITPIX -32 cube with BSCALE/BZERO is read as BZERO + BSCALE * stored,
    rounded to float (what cfitsio's TFLOAT read returns, utilsFits.c:212)."""
    rows, cols, L = 4, 3, 6
    stored = np.random.default_rng(8).normal(size=(rows * cols, 4, L)).astype(np.float32)
    p = str(tmp_path / "s.fits")
    fitsutil.write(p, fitsutil.stokes_cube(stored, rows, cols),
                   ctypes=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN"), bzero=0.25, bscale=2.5)
    rc, got, lay = _read_cube(io, p, L)
    assert rc == 0
    want = (0.25 + 2.5 * stored.astype(np.float64)).astype(np.float32)
    assert np.array_equal(got, want)
    io.pmio_free_layout(C.byref(lay))


def test_fits_models_planes(io, tmp_path):
    rows, cols = 6, 4
    n = rows * cols
    m = np.random.default_rng(7).normal(size=(n, 11))
    chi2 = np.arange(n, dtype=np.float32)
    it = np.arange(n, dtype=np.int32) % 50
    p = str(tmp_path / "m.fits")
    assert io.pmio_fits_write_models(p.encode(), rows, cols, m.ctypes.data, chi2.ctypes.data,
                                     it.ctypes.data) == 0
    data, hdr = fitsutil.read(p)
    assert (hdr["NAXIS1"], hdr["NAXIS2"], hdr["NAXIS3"]) == (str(rows), str(cols), "13")
    planes = data.reshape(13, n)                          # each plane in pixel order
    np.testing.assert_array_equal(planes[:11].T, m.astype(np.float32))
    np.testing.assert_array_equal(planes[11], it)
    np.testing.assert_array_equal(planes[12], chi2)


def test_fits_2d_straylight_lambda(io, tmp_path):
    # the reference's own stray-light file, [4][49]
    p = os.path.join(GOLDEN, "straylight.fits").encode()
    ptr = C.POINTER(C.c_float)()
    pp, n = C.c_int(), C.c_long()
    assert io.pmio_fits_read_straylight(p, 49, 0, 0, 0, 0, C.byref(ptr), C.byref(pp),
                                        C.byref(n)) == 0
    got = np.ctypeslib.as_array(ptr, shape=(4, 49)).copy()
    io.pmio_free(ptr)
    assert pp.value == 0 and np.array_equal(got, read_fits_2d(os.path.join(GOLDEN, "straylight.fits")))
    # transposed [49][4] reads the same
    q = str(tmp_path / "t.fits")
    fitsutil.write(q, got.T.copy())
    ptr2 = C.POINTER(C.c_float)()
    assert io.pmio_fits_read_straylight(q.encode(), 49, 0, 0, 0, 0, C.byref(ptr2), C.byref(pp),
                                        C.byref(n)) == 0
    assert np.array_equal(np.ctypeslib.as_array(ptr2, shape=(4, 49)), got)
    io.pmio_free(ptr2)
    # a 2-D mask of 16-bit integers with BZERO, and of bytes
    for bitpix, bz in ((16, 32768), (8, None), (32, None)):
        mask = (np.arange(12).reshape(3, 4) % 3 != 0).astype(np.int64)
        fitsutil.write(q, mask - (bz or 0), bitpix=bitpix, bzero=bz)
        d = C.POINTER(C.c_double)()
        a, b = C.c_long(), C.c_long()
        assert io.pmio_fits_read_2d(q.encode(), C.byref(d), C.byref(a), C.byref(b)) == 0
        assert (a.value, b.value) == (4, 3)
        assert np.array_equal(np.ctypeslib.as_array(d, shape=(3, 4)), mask)
        io.pmio_free(d)
    # a wavelength file: row 1 the line index, row 2 the wavelengths
    lam = FE6173 + np.arange(-5, 6) * 0.02
    fitsutil.write(q, np.stack([np.full(lam.size, 1.0), lam]), bitpix=-64)
    d = C.POINTER(C.c_double)()
    Ln, idx = C.c_int(), C.c_int()
    assert io.pmio_fits_read_lambda(q.encode(), C.byref(d), C.byref(Ln), C.byref(idx)) == 0
    assert (Ln.value, idx.value) == (lam.size, 1)
    assert np.array_equal(np.ctypeslib.as_array(d, shape=(lam.size,)), lam)
    io.pmio_free(d)
