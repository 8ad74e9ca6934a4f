"""CPU: the parameter-file reader (bbarolo2_amd/params.py) against the reference's reading rules
(tasks.py:1203-1224, keywords.py:19-152, rings.py:157-360, params.py:1447-1560): keywords any case, comments and a
repeated keyword, the C++-style scalar reads, the defaults and the task defaults of METHOD GLOBAL, the ring forms
(NRADII and RADSEP, a list, range(), Rmin~Rmax:step, one value per ring, file(name,column)), the likelihood
options, the free set of a GLOBAL fit. The parameter files are written here with the values of the reference's
example files (examples/ngc2403_3DFIT.par, ugc1281_3DFIT_GLOBAL.par)."""
import numpy as np
import pytest

from bbarolo2_amd import params as P

NGC2403 = """# Name of the fitsfile
FITSFILE    ngc2403.fits

############## 3DFIT parameters ################
3DFIT       true

// Input rings
NRADII      42
RADSEP      30
VSYS        132.8
XPOS        77
YPOS        77
VROT        100
VRAD        0
VDISP       8
INC         60
pa          123.7
Z0          10

FREE        VROT VDISP PA INC
FLAGERRORS  true
"""

UGC1281 = """FITSFILE        ugc1281.fits
SPECRES         0.425
3DFIT           true
METHOD          GLOBAL
NRADII          8
RADSEP          30
INC             85
FREE            vrot vdisp inc_single phi_single z0_single xpos_single ypos_single vsys_single radmax fracerr
MASK            SEARCH
FLAGERRORS	true
LIKEOPTS        smooth=(vrot=2.3,vdisp=1) dilation=1.5,4 noisecorr=beam
"""


def _write(tmp_path, text, name="fit.par"):
    f = tmp_path / name
    f.write_text(text)
    return str(f)


def test_read_parfile_rules(tmp_path):
    with pytest.warns(UserWarning):
        raw = P.read_parfile(_write(tmp_path, NGC2403 + "VDISP 12   # a later value wins\nNOTAKEY 3\n"))
    assert raw["FITSFILE"] == "ngc2403.fits" and raw["PA"] == "123.7" and raw["3DFIT"] == "true"
    assert raw["VDISP"].split()[0] == "12" and list(raw)[-1] == "VDISP"
    assert "NOTAKEY" not in raw
    with pytest.warns(UserWarning):
        P.read_parfile(_write(tmp_path, "NOTAKEY 3\n", "b.par"))
    with pytest.raises(FileNotFoundError):
        P.read_parfile(str(tmp_path / "missing.par"))


def test_values_and_defaults(tmp_path):
    p = P.ParFile.read(_write(tmp_path, NGC2403))
    assert p.get("3DFIT") is True and p.get("FLAGERRORS") is True and p.get("TWOSTAGE") is True
    assert p.get("NRADII") == 42 and p.get("RADSEP") == 30.0 and p.method() == "CLASSIC"
    gp, fp = p.galmod_par(), p.fit3d_par()
    assert (gp.GALMODTYPE, gp.LTYPE, gp.CMODE, gp.CDENS, gp.ISEED, gp.NV, gp.SM, gp.DENSFLUX) == \
        ("MC", 1, 1, 10.0, -1, -1, True, True)
    assert (fp.FTYPE, fp.WFUNC, fp.NORM, fp.BWEIGHT, fp.TOL, fp.FREE) == (2, 2, "AZIM", 1.0, 1e-3, "VROT VDISP PA INC")
    assert (fp.REVERSE, fp.SWEEPS, fp.RINGREGION) == ("false", 1, "STRIP")       # (params.py:200-202)
    r = P.ParFile({"REVERSE": "AUTO", "SWEEPS": "2", "RINGREGION": "annulus"}).fit3d_par()
    assert (r.REVERSE, r.SWEEPS, r.RINGREGION) == ("auto", 2, "ANNULUS")
    # C++-style reads: the leading number of the first token, '1' or 'true' for a flag
    q = P.ParFile({"NRADII": "12abc", "RADSEP": "2.5e1x", "TWOSTAGE": "0", "SM": "TRUE", "CDENS": "12.7",
                   "RELINT": "1, 0.5 x 3"})
    assert (q.get("NRADII"), q.get("RADSEP"), q.get("TWOSTAGE"), q.get("SM"), q.get("CDENS"), q.get("RELINT")) == \
        (12, 25.0, False, True, 12, [1.0, 0.5])


def test_rings_forms(tmp_path):
    p = P.ParFile.read(_write(tmp_path, NGC2403))
    r = p.rings()
    assert r.nr == 42 and np.allclose(r.radii, (np.arange(42) + 0.5) * 30) and r.radsep == 30
    assert np.all(r.vrot == 100) and np.all(r.xpos == 77) and np.all(r.phi == 123.7) and np.all(r.dens == 1)
    assert np.all(r.vdispthick == -1)
    q = P.ParFile({"RADII": "range(105,400,90)", "XPOS": "33", "YPOS": "33", "VROT": "120 125 130 135"})
    r = q.rings()
    assert np.allclose(r.radii, [105, 195, 285, 375]) and np.allclose(r.vrot, [120, 125, 130, 135])
    assert r.radsep == 90
    assert len(P.radii_range("0~1:0.1")) == 11 and len(P.radii_range("0.03~0.5:0.06")) == 8
    assert P.radii_range("1 2 3") is None
    with pytest.raises(ValueError):
        P.radii_range("range(5,1,1)")
    r = P.ParFile({"RADII": "0.65 1.10 1.55 2.00 2.45 2.90", "XPOS": "5", "YPOS": "6"}).rings()
    assert r.nr == 6 and np.isclose(r.radsep, 0.45)
    # file(name, column): a ring table; the shortest list sets the number of rings
    (tmp_path / "rings.txt").write_text("# RAD VROT\n10 100\n20 110\n30 120\n")
    r = P.ParFile({"RADII": "file(rings.txt,1)", "VROT": "file(rings.txt,2)", "VDISP": "9 8 7 6",
                   "XPOS": "1", "YPOS": "2"}, base=str(tmp_path)).rings()
    assert r.nr == 3 and np.allclose(r.vrot, [100, 110, 120]) and np.allclose(r.vdisp, [9, 8, 7])
    with pytest.raises(NotImplementedError):
        P.ParFile({"NRADII": "3", "RADSEP": "10", "XPOS": "07h36m51s", "YPOS": "65d36m09s"}).rings()


def test_global_defaults_options_and_free(tmp_path):
    p = P.ParFile.read(_write(tmp_path, UGC1281))
    assert p.method() == "GLOBAL" and p.get("GALMODTYPE") == "ANALYTIC" and p.get("SPECRES") == 0.425
    opts = p.like_options()
    assert opts["smooth"] == {"vrot": 2.3, "vdisp": 1.0} and opts["dilation"] == (1.5, 4)
    assert opts["noisecorr"] == "beam" and opts["likemask"] == "dilated" and opts["densprof_smooth"] == 0.3
    assert p.free_spec(8) == [("vrot", "ring"), ("vdisp", "ring"), ("inc", "single"), ("phi", "single"),
                              ("z0", "single"), ("xpos", "single"), ("ypos", "single"), ("vsys", "single"),
                              ("radmax", "single"), ("fracerr", "single")]
    q = P.ParFile({"METHOD": "BAYESIAN"})
    assert q.get("FREE") == P.BAYES_DEFAULT_FREE and q.get("GALMODTYPE") == "ANALYTIC"
    for bad in ("likemask=maybe", "dilation=3", "noisecorr=-1", "radmax=wall", "sigma=0", "colour=red"):
        with pytest.raises(ValueError):
            P.like_options(bad)
    assert P.parse_options("n_live=2000 seed=1 f_live=0.01") == {"n_live": 2000, "seed": 1, "f_live": 0.01}
    assert P.like_options("fracerr=none")["fracerr"] is None
