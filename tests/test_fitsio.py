"""FITS cubes and ring tables without astropy (bbarolo2_amd.fitsio): round trips, the FITS layout, scaled integer
data and blanks, four-axis cubes, and the reference's own example cubes read against the values its documentation
gives for them (docs/examples/galmod3d.md:8: NGC 2403, 16" pixels, 62 channels of 5.15 km/s, a beam of about 59")."""
import os

import numpy as np
import pytest

from bbarolo2_amd import fitsio
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.rings import Rings

REF_EXAMPLES = "/root/reference/examples"


def test_round_trip_cube_and_cards(tmp_path):
    rng = np.random.default_rng(0)
    data = rng.normal(size=(5, 7, 9)).astype(np.float32)
    data[0, 0, 0] = np.nan
    cards = {"CTYPE1": "RA---SIN", "CDELT1": -0.0027777778, "CRPIX3": 3, "OBJECT": "O'Brien cube", "FLAG": True,
             "BMAJ": 8.3333e-3}
    p = tmp_path / "c.fits"
    fitsio.write_fits(p, data, cards)
    raw = p.read_bytes()
    assert len(raw) % 2880 == 0 and raw[:80].decode() == f"{'SIMPLE  =                    T':<80}"
    c, d = fitsio.read_fits(p)
    assert d.dtype == np.float32 and d.shape == data.shape
    assert np.array_equal(d[~np.isnan(d)], data[~np.isnan(data)]) and np.isnan(d[0, 0, 0])
    for k, v in cards.items():
        assert c[k] == v, (k, c[k], v)
    assert (c["NAXIS1"], c["NAXIS2"], c["NAXIS3"]) == (9, 7, 5)


def _int16_fits(path, arr, bscale, bzero, blank):
    cards = [fitsio._format_card("SIMPLE", True), fitsio._format_card("BITPIX", 16), fitsio._format_card("NAXIS", 3)]
    for i, n in enumerate(reversed(arr.shape)):
        cards.append(fitsio._format_card(f"NAXIS{i + 1}", n))
    cards += [fitsio._format_card("BSCALE", bscale), fitsio._format_card("BZERO", bzero),
              fitsio._format_card("BLANK", blank), f"{'END':<80}"]
    h = "".join(cards).encode()
    h += b" " * (-len(h) % 2880)
    body = arr.astype(">i2").tobytes()
    body += b"\0" * (-len(body) % 2880)
    path.write_bytes(h + body)


def test_scaled_integers_blanks_and_load_cube(tmp_path):
    arr = np.arange(24, dtype=np.int16).reshape(2, 3, 4) - 5
    arr[1, 2, 3] = -32768
    p = tmp_path / "i.fits"
    _int16_fits(p, arr, 0.5, 10.0, -32768)
    c, d = fitsio.read_fits(p)
    want = arr.astype(np.float32) * 0.5 + 10.0
    assert d.dtype == np.float32 and np.isnan(d[1, 2, 3])
    m = arr != -32768
    assert np.array_equal(d[m], want[m])
    h, cube = fitsio.load_cube(p)
    assert cube.shape == (2, 3, 4) and cube[1, 2, 3] == 0 and (h.DimAx(0), h.DimAx(2)) == (4, 2)


def test_four_axis_cubes(tmp_path):
    rng = np.random.default_rng(1)
    d4 = rng.random((2, 6, 5, 4)).astype(np.float32)       # (STOKES, FREQ, y, x)
    p = tmp_path / "s4.fits"
    fitsio.write_fits(p, d4, {"CTYPE3": "VELO", "CTYPE4": "STOKES"})
    h, cube = fitsio.load_cube(p)
    assert np.array_equal(cube, d4[0]) and h.DimAx(2) == 6
    d3s = rng.random((6, 1, 5, 4)).astype(np.float32)      # (FREQ, STOKES, y, x): STOKES third
    p2 = tmp_path / "s3.fits"
    fitsio.write_fits(p2, d3s, {"CTYPE3": "STOKES", "CTYPE4": "VELO", "CDELT4": 5.0})
    h2, cube2 = fitsio.load_cube(p2)
    assert np.array_equal(cube2, d3s[:, 0]) and h2.DimAx(2) == 6 and h2.Cdelt(2) == 5.0


def test_synthetic_header_survives_a_file(tmp_path):
    from bbarolo2_amd.galmod import GalmodSetup
    h = emptyheader([32, 24, 16], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., obj="SYNTH")
    p = tmp_path / "h.fits"
    fitsio.save_cube(p, h, np.zeros((16, 24, 32), np.float32))
    h2, _ = fitsio.load_cube(p)
    a, b = GalmodSetup(h), GalmodSetup(h2)
    assert a.cdelt == b.cdelt and a.crota2 == b.crota2 and np.array_equal(a.cd2i, b.cd2i)
    assert np.array_equal(a.vedges_lines, b.vedges_lines) and a.sig_instr == b.sig_instr


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="the reference's example cubes are not here")
def test_reference_example_cubes():
    h, cube = fitsio.load_cube(os.path.join(REF_EXAMPLES, "ngc2403.fits"))
    assert cube.shape == (62, 150, 150) and cube.dtype == np.float32 and np.all(np.isfinite(cube))
    assert abs(abs(h.Cdelt(0)) * 3600 - 16.0) < 1e-3 and abs(abs(h.Cdelt(2)) / 1000 - 5.15) < 0.01
    assert abs(h.Bmaj() * 3600 - 59) < 1.0
    h2, cube2 = fitsio.load_cube(os.path.join(REF_EXAMPLES, "ugc1281.fits"))
    assert cube2.shape == (h2.DimAx(2), h2.DimAx(1), h2.DimAx(0)) and np.all(np.isfinite(cube2))


def test_ring_table_round_trip(tmp_path):
    R = np.arange(15., 346., 30.)
    r = Rings().setRings(radii=R, xpos=64., ypos=63.5, vsys=12.25, vrot=150 * (2 / np.pi) * np.arctan(R / 60.),
                         vdisp=10., z0=10., dens=1.5e-3, inc=60., phi=120., vrad=-1.5)
    p = tmp_path / "rings_final1.txt"
    errs = np.stack([-np.ones((r.nr, 2)), 2 * np.ones((r.nr, 2))], axis=1)       # (nr, 2, nfree = 2)
    fitsio.write_rings(p, r, mpar=[1, 1, 0, 0, 0, 0, 0, 0, 0, 0], errors=errs, fit_ok=[True] * r.nr)
    lines = p.read_text().splitlines()
    head = lines[0].split()
    assert head[:11] == ["#RAD(arcs)"] + list(fitsio.RING_COLUMNS[1:])
    assert head[11:] == ["E_VROT1", "E_VROT2", "E_DISP1", "E_DISP2", "FITOK"]
    assert lines[1].startswith("15.000     ") and "1.5e-03" in lines[1]
    r2 = fitsio.read_rings(p)
    for k in ("radii", "vrot", "vdisp", "inc", "phi", "z0", "xpos", "ypos", "vsys", "vrad"):
        assert np.allclose(getattr(r2, k), np.round(getattr(r, k), 3), atol=5e-4), k
    assert np.allclose(r2.dens, 1.5e-3)
