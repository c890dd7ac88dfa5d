"""A minimal FITS primary-image writer/reader in numpy, independent of the
library's own C++ one, for building test inputs and checking outputs."""
import numpy as np

_BITPIX_DTYPE = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _card(key, value):
    if isinstance(value, str):
        v = "'%-8s'" % value
        c = "%-8s= %-20s" % (key, v)
    elif isinstance(value, bool):
        c = "%-8s= %20s" % (key, "T" if value else "F")
    else:
        c = "%-8s= %20s" % (key, value)
    return c.ljust(80)[:80]


def write(path, data, ctypes=None, bitpix=-32, extra=(), bzero=None, bscale=1):
    """data: numpy array in C order, whose LAST axis is NAXIS1 (FITS order
    reversed, as numpy stores it)."""
    data = np.asarray(data)
    naxes = list(reversed(data.shape))
    cards = [_card("SIMPLE", True), _card("BITPIX", bitpix), _card("NAXIS", len(naxes))]
    cards += [_card("NAXIS%d" % (k + 1), n) for k, n in enumerate(naxes)]
    if ctypes:
        cards += [_card("CTYPE%d" % (k + 1), t) for k, t in enumerate(ctypes)]
    if bzero is not None:
        cards += [_card("BSCALE", bscale), _card("BZERO", bzero)]
    cards += [c.ljust(80)[:80] for c in extra]
    cards.append("END".ljust(80))
    hdr = "".join(cards)
    hdr += " " * ((2880 - len(hdr) % 2880) % 2880)
    raw = data.astype(_BITPIX_DTYPE[bitpix]).tobytes()
    raw += b"\0" * ((2880 - len(raw) % 2880) % 2880)
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii"))
        f.write(raw)


def read(path):
    """(data in numpy order, header dict of the cards)."""
    blob = open(path, "rb").read()
    hdr, i = {}, 0
    while True:
        card = blob[i:i + 80].decode("ascii")
        i += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            v = card[10:].split("/")[0].strip()
            hdr[key] = v.strip("'").strip() if v.startswith("'") else v
    start = (i + 2879) // 2880 * 2880
    bitpix = int(hdr["BITPIX"])
    naxes = [int(hdr["NAXIS%d" % (k + 1)]) for k in range(int(hdr["NAXIS"]))]
    n = int(np.prod(naxes))
    dt = np.dtype(_BITPIX_DTYPE[bitpix])
    data = np.frombuffer(blob[start:start + n * dt.itemsize], dtype=dt).reshape(naxes[::-1])
    return data.astype(np.float64), hdr


def stokes_cube(cube, rows, cols, order=("WAVE-GRI", "STOKES", "HPLN-TAN", "HPLT-TAN")):
    """A numpy array for write() holding cube[npix][4][L] (pixel = col*rows +
    row) with the FITS axes in `order` (NAXIS1 first)."""
    npix, four, L = cube.shape
    a = cube.reshape(cols, rows, 4, L)                   # [col][row][stokes][lambda]
    names = {"HPLT-TAN": 0, "HPLN-TAN": 1, "STOKES": 2, "WAVE-GRI": 3}
    perm = [names[o] for o in reversed(order)]          # numpy axis order = reversed FITS order
    return np.ascontiguousarray(a.transpose(perm))
