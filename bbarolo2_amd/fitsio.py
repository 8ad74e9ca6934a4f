"""FITS cubes and ring tables, without astropy.

The reference reads cubes with astropy (Cube.readCube and _readHeaderData, bbarolo/cube.py:182-261: the header,
then the data as float32 (nz, ny, nx), the first plane of a fourth axis, blanks (NaN) set to 0) and writes the
fitted rings as text tables (writeHeader / writeRing and RING_COLUMNS, bbarolo/galfit.py:1633-1705). This module
does the same for the primary HDU of a FITS file (FITS standard 4.0: 2880-byte blocks of 80-character cards,
big-endian data, BSCALE / BZERO / BLANK) and for the columns of the ring tables, so that parameter files, cubes and
outputs stay a drop-in.
"""

import os

import numpy as np

from .header import Header
from .rings import Rings

BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_STRUCTURAL = ("SIMPLE", "BITPIX", "NAXIS", "EXTEND", "END")


def _parse_value(s):
    """The value of a card (the text after '= '), without its comment."""
    s = s.strip()
    if s.startswith("'"):
        # a string: '' inside is a quote; trailing blanks are not significant
        out, i = [], 1
        while i < len(s):
            if s[i] == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(s[i])
            i += 1
        return "".join(out).rstrip()
    v = s.split("/", 1)[0].strip()
    if v in ("T", "F"):
        return v == "T"
    if v == "":
        return None
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E").replace("d", "e"))
    except ValueError:
        return v


def read_header(f):
    """The cards (dict, in file order; COMMENT / HISTORY dropped) of the header at the file position of f, which is
    left at the start of the data."""
    cards = {}
    while True:
        block = f.read(BLOCK)
        if len(block) < BLOCK:
            raise ValueError("FITS: the header ends before its END card")
        for k in range(0, BLOCK, 80):
            card = block[k:k + 80].decode("ascii", errors="replace")
            key = card[:8].strip().upper()
            if key == "END":
                return cards
            if key in ("", "COMMENT", "HISTORY") or card[8:10] != "= ":
                continue
            cards[key] = _parse_value(card[10:])


def read_fits(path):
    """(cards, data) of the primary HDU: data as in the file (NAXISn ... NAXIS1), BSCALE / BZERO applied and
    integer BLANK values set to NaN (float32 for integer data of up to 32 bits, as the reference converts)."""
    with open(path, "rb") as f:
        cards = read_header(f)
        if not cards.get("SIMPLE", False):
            raise ValueError(f"{path}: not a FITS file (SIMPLE is not T)")
        bitpix = int(cards["BITPIX"])
        if bitpix not in _DTYPES:
            raise ValueError(f"{path}: BITPIX = {bitpix}")
        naxis = int(cards.get("NAXIS", 0))
        shape = tuple(int(cards[f"NAXIS{i}"]) for i in range(naxis, 0, -1))
        n = int(np.prod(shape)) if naxis else 0
        raw = np.frombuffer(f.read(n * abs(bitpix) // 8), dtype=_DTYPES[bitpix])
        if raw.size != n:
            raise ValueError(f"{path}: the data are shorter than NAXIS says")
    data = raw.reshape(shape) if naxis else raw
    scale, zero = float(cards.get("BSCALE", 1.0)), float(cards.get("BZERO", 0.0))
    if bitpix > 0:
        blank = cards.get("BLANK")
        ft = np.float32 if bitpix <= 32 else np.float64
        out = data.astype(ft) * ft(scale) + ft(zero) if (scale != 1.0 or zero != 0.0) else data.astype(ft)
        if blank is not None:
            out[data == int(blank)] = np.nan
        return cards, out
    data = data.astype(data.dtype.newbyteorder("="))
    if scale != 1.0 or zero != 0.0:
        data = data * data.dtype.type(scale) + data.dtype.type(zero)
    return cards, data


def _format_card(key, value):
    key = key.upper()[:8]
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20d}"
    elif isinstance(value, (float, np.floating)):
        r = repr(float(value)).upper()
        if "E" not in r and "." not in r:
            r += ".0"
        v = f"{r:>20}"
    else:
        s = str(value).replace("'", "''")
        v = f"'{s:<8}'"
    card = f"{key:<8}= {v}"
    if len(card) > 80:
        raise ValueError(f"FITS card too long: {card}")
    return f"{card:<80}"


def write_fits(path, data, cards=None, overwrite=True):
    """data (float32 or float64, any number of axes, the last one NAXIS1) as the primary HDU, with the cards (dict;
    the structural ones are written from data)."""
    data = np.asarray(data)
    if data.dtype not in (np.float32, np.float64):
        data = data.astype(np.float32)
    if os.path.exists(path) and not overwrite:
        raise FileExistsError(path)
    bitpix = -32 if data.dtype == np.float32 else -64
    head = [_format_card("SIMPLE", True), _format_card("BITPIX", bitpix), _format_card("NAXIS", data.ndim)]
    for i, n in enumerate(reversed(data.shape)):
        head.append(_format_card(f"NAXIS{i + 1}", int(n)))
    for k, v in (cards or {}).items():
        ku = str(k).upper()
        if ku in _STRUCTURAL or (ku.startswith("NAXIS") and ku[5:].isdigit()) or ku in ("BSCALE", "BZERO"):
            continue
        if v is None:
            continue
        head.append(_format_card(ku, v))
    head.append(f"{'END':<80}")
    h = "".join(head).encode("ascii")
    h += b" " * (-len(h) % BLOCK)
    body = np.ascontiguousarray(data, dtype=data.dtype.newbyteorder(">")).tobytes()
    body += b"\0" * (-len(body) % BLOCK)
    with open(path, "wb") as f:
        f.write(h)
        f.write(body)


def load_cube(path):
    """(Header, data) of a cube as the reference's Cube.readCube keeps it (cube.py:182-261): float32 (nz, ny, nx)
    (float64 for 64-bit data), the first plane of a fourth axis (of the third when it is STOKES), NaN set to 0."""
    cards, data = read_fits(path)
    h = dict(cards)
    naxis = int(cards.get("NAXIS", 0))
    for i in range(naxis + 1, 4):
        h[f"NAXIS{i}"] = 1
    stokes3 = naxis > 3 and str(cards.get("CTYPE3", "")).strip().lower() == "stokes"
    if stokes3:
        for k in ("CRPIX", "CRVAL", "CDELT", "CTYPE", "CUNIT", "NAXIS"):
            if f"{k}4" in cards:
                h[f"{k}3"] = cards[f"{k}4"]
    dims = [int(h[f"NAXIS{i + 1}"]) for i in range(3)]
    if data.ndim == 4:
        data = data[:, 0] if stokes3 else data[0]
    n = dims[0] * dims[1] * dims[2]
    data = np.asarray(data).reshape(-1)[:n].reshape(dims[2], dims[1], dims[0])
    data = data.astype(np.float32 if data.dtype.itemsize <= 4 else np.float64, copy=True)
    data[np.isnan(data)] = 0
    return Header(h), np.ascontiguousarray(data)


def save_cube(path, header, data):
    """A (nz, ny, nx) cube with the keywords of header (a Header or a dict)."""
    cards = header.cards if isinstance(header, Header) else dict(header)
    write_fits(path, np.asarray(data), cards)


# ---------------------------------------------------------------------------------------------------------
# Ring tables (rings_final1.txt ...): the columns of RING_COLUMNS, then two error columns per free parameter, then
# FITOK (galfit.py:1633-1705).
RING_COLUMNS = ("RAD(arcs)", "VROT(km/s)", "DISP(km/s)", "INC(deg)", "P.A.(deg)", "Z0(arcs)", "DENS(un)",
                "XPOS(pix)", "YPOS(pix)", "VSYS(km/s)", "VRAD(km/s)")
_RING_KEYS = ("radii", "vrot", "vdisp", "inc", "phi", "z0", "dens", "xpos", "ypos", "vsys", "vrad")
# column widths: 11 for RAD VROT DISP VSYS VRAD, 10 for the others; DENS in %.1e, the others %.3f
_WIDTH = (11, 11, 11, 10, 10, 10, 10, 10, 10, 11, 11)
ERROR_NAMES = ("VROT", "DISP", "DENS", "Z0", "INC", "PA", "XPOS", "YPOS", "VSYS", "VRAD")


def ring_table_header(mpar=None, bad_rings=False):
    """The header line: mpar, the free flags of ERROR_NAMES, adds their error columns."""
    s = "#" + f"{RING_COLUMNS[0]:<{_WIDTH[0] - 1}}" + "".join(f"{c:<{w}}" for c, w in zip(RING_COLUMNS[1:],
                                                                                          _WIDTH[1:]))
    if mpar is not None:
        for free, nm in zip(mpar, ERROR_NAMES):
            if free:
                s += f"{'E_' + nm + '1':<10}{'E_' + nm + '2':<10}"
    if bad_rings:
        s += f"{'FITOK':<10}"
    return s


def ring_table_row(r, i, errors=None, fit_ok=None):
    """Row i of the rings r; errors: (2, nfree) lower and upper errors of the free parameters of this ring."""
    vals = [np.asarray(getattr(r, k), float)[i] for k in _RING_KEYS]
    s = ""
    for k, (v, w) in enumerate(zip(vals, _WIDTH)):
        s += f"{v:<{w}.1e}" if _RING_KEYS[k] == "dens" else f"{v:<{w}.3f}"
    if errors is not None:
        for kk in range(len(errors[0])):
            s += f"{errors[0][kk]:<10.3f}{errors[1][kk]:<10.3f}"
    if fit_ok is not None:
        s += f"{int(fit_ok):<10d}"
    return s


def write_rings(path, r, mpar=None, errors=None, fit_ok=None):
    """The ring table of the rings r. mpar (the free flags of ERROR_NAMES) and errors (nr, 2, nfree) add the error
    columns; fit_ok (nr booleans) adds FITOK."""
    with open(path, "w") as f:
        f.write(ring_table_header(mpar if errors is not None else None, fit_ok is not None) + "\n")
        for i in range(r.nr):
            f.write(ring_table_row(r, i, errors[i] if errors is not None else None,
                                   None if fit_ok is None else fit_ok[i]) + "\n")


def read_rings(path):
    """The rings of a ring table (the columns of RING_COLUMNS; the error and FITOK columns are skipped)."""
    with open(path) as f:
        lines = [ln for ln in f if ln.strip()]
    names = lines[0].lstrip("#").split()
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:] if not ln.lstrip().startswith("#")])
    rows = rows.reshape(-1, len(names))
    vals = {}
    for key, col in zip(_RING_KEYS, RING_COLUMNS):
        if col in names:
            vals[key] = rows[:, names.index(col)]
    return Rings().setRings(**vals)
