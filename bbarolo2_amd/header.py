"""The cube header fields that the model-cube likelihood reads, and synthetic cube headers.

Mirrors the accessors of the reference's ``Header`` that ``GalmodSetup`` and ``Smooth3D`` use
(bbarolo/header.py: Cdelt, Crpix, Crval, Cunit, Crota, Bmaj, Bmin, Bpa, Freq0, VelDef, SpectralType,
Redshift), without astropy: a header is built from a plain dict of FITS keywords.
``emptyheader`` follows the conventions of the reference's ``emptyfits`` (bbarolo/utils.py:36-102):
RA---SIN / DEC--SIN / VELO-HEL axes in deg, deg, km/s, CRPIX = (n - 1)/2 + 1, CRVAL = (60, 0, 0).
"""

import math

C_MS = 2.99792458e8          # speed of light, m/s (bbarolo/bbutils.py:23)
C_KMS = C_MS / 1000.

# velocity units -> km/s divisor, as VEL_UNITS of the reference (value in unit / divisor = km/s)
VEL_UNITS = {"km/s": 1., "kms": 1., "m/s": 1E03, "ms": 1E03, "cm/s": 1E05}
# wavelength units -> units per metre (bbarolo/bbutils.py:112-113)
WAVE_UNITS = {"m": 1., "cm": 1E02, "mm": 1E03, "um": 1E06, "mum": 1E06, "micron": 1E06, "nm": 1E09,
              "nanom": 1E09, "a": 1E10, "ang": 1E10, "angstrom": 1E10}
HI_FREQ = 0.1420405751786E10  # default rest frequency (bbarolo/header.py:338-340)


class Header:
    """A FITS-like header of a 3D cube: keywords in a dict (upper-case keys)."""

    def __init__(self, cards):
        self.cards = {str(k).upper(): v for k, v in dict(cards).items()}
        for k in ("NAXIS1", "NAXIS2", "NAXIS3"):
            if k not in self.cards:
                raise ValueError(f"header: {k} is missing")

    def get(self, key, default=None):
        return self.cards.get(key.upper(), default)

    # axis accessors (0-based axis index, as the reference's Header)
    def DimAx(self, i): return int(self.cards[f"NAXIS{i + 1}"])
    def Cdelt(self, i): return float(self.cards[f"CDELT{i + 1}"])
    def Crpix(self, i): return float(self.cards[f"CRPIX{i + 1}"])
    def Crval(self, i): return float(self.cards[f"CRVAL{i + 1}"])
    def Cunit(self, i): return str(self.cards.get(f"CUNIT{i + 1}", "deg" if i < 2 else "km/s"))
    def Ctype(self, i): return str(self.cards.get(f"CTYPE{i + 1}", ""))
    def Crota(self): return float(self.cards.get("CROTA2", self.cards.get("CROTA", 0.)))
    def Bmaj(self): return float(self.cards.get("BMAJ", 0.))
    def Bmin(self): return float(self.cards.get("BMIN", 0.))
    def Bpa(self): return float(self.cards.get("BPA", 0.))
    def Redshift(self): return float(self.cards.get("REDSHIFT", 0.))
    def Name(self): return str(self.cards.get("OBJECT", "NONE"))

    def Freq0(self):
        """FREQ0 / RESTFREQ / RESTFRQ, or the HI line (bbarolo/header.py:330-352, velocity axes)."""
        for k in ("FREQ0", "RESTFREQ", "RESTFRQ"):
            if k in self.cards and float(self.cards[k]) > 0:
                return float(self.cards[k])
        return HI_FREQ

    def VelDef(self):
        """The velocity definition (VELDEF keyword: radio, optical or relativistic; default 'auto', which
        GalmodSetup takes as radio)."""
        v = str(self.cards.get("VELDEF", "auto")).lower()
        for name in ("radio", "optical", "relativistic"):
            if v.startswith(name[:3]):
                return name
        return "auto"

    def SpectralType(self):
        """wave, freq, velo-opt or velo-radio from CTYPE3/CUNIT3 (bbarolo/header.py:319-328)."""
        ct, cu = self.Ctype(2).lower(), self.Cunit(2).lower()
        if "wav" in ct or "um" in cu or "nm" in cu or "ang" in cu or "micr" in cu:
            return "wave"
        if "freq" in ct or "hz" in cu:
            return "freq"
        if "vopt" in ct or "felo" in ct:
            return "velo-opt"
        if "vel" in ct or "vrad" in ct or "m/s" in cu:
            return "velo-radio"
        return "unknown"

    @property
    def shape(self):
        """(nz, ny, nx) of the data array."""
        return (self.DimAx(2), self.DimAx(1), self.DimAx(0))


def arcsconv(cunit):
    """Arcsec per unit of a spatial axis (bbarolo/bbutils.py:227-232)."""
    c = cunit.lower()
    if "deg" in c: return 3600.
    if "arcm" in c: return 60.
    if "arcs" in c: return 1.
    raise ValueError(f"Conversion error (unknown CUNIT for RA-DEC): {cunit}. Cannot convert to ARCSEC.")


def delta_vel(h):
    """Channel width in km/s of a velocity axis (bbarolo/bbutils.py:159-165); for frequency or
    wavelength axes the mean width over the cube (bbutils.py:166-180)."""
    cu = h.Cunit(2).lower()
    if cu in VEL_UNITS:
        return h.Cdelt(2) / VEL_UNITS[cu]
    nz = h.DimAx(2)
    if nz < 2:
        return h.Cdelt(2)
    i0 = (1 - h.Crpix(2)) * h.Cdelt(2) + h.Crval(2)
    i1 = (nz - h.Crpix(2)) * h.Cdelt(2) + h.Crval(2)
    if cu in ("hz", "khz", "mhz", "ghz"):
        fc = {"hz": 1., "khz": 1e3, "mhz": 1e6, "ghz": 1e9}[cu]
        f0 = h.Freq0() / (1 + h.Redshift())
        if h.VelDef() == "radio":
            return -C_KMS * h.Cdelt(2) * fc / f0
        v = lambda f: _freq2vel(f * fc, f0, h.VelDef())
        return (v(i1) - v(i0)) / (nz - 1)
    w0 = WAVE_UNITS.get(cu, 1.)
    rest = h.Crval(2) / w0
    v = lambda w: _freq2vel(C_MS / (w / w0), C_MS / rest, h.VelDef())
    return (v(i1) - v(i0)) / (nz - 1)


def _freq2vel(f, f0, veldef):
    r = f / f0
    if veldef == "radio":
        return C_KMS * (1 - r)
    if veldef == "optical":
        return C_KMS * (1 / r - 1)
    return C_KMS * (1 - r * r) / (1 + r * r)


def emptyheader(dims, cdelts, crpixs=None, crvals=None, ctypes=None, cunits=None, beam=None, obj=None,
                bunit="JY/BEAM", **extra):
    """A header with the conventions of the reference's emptyfits: dims (nx, ny, nz), cdelts in the
    units cunits (default deg, deg, km/s); beam in degrees (a number: circular; (bmaj, bmin, bpa))."""
    ctypes = ["RA---SIN", "DEC--SIN", "VELO-HEL"] if ctypes is None else ctypes
    cunits = ["deg", "deg", "km/s"] if cunits is None else cunits
    crvals = [60, 0, 0] if crvals is None else crvals
    c = {}
    for i, n in enumerate(dims):
        c[f"NAXIS{i + 1}"] = int(n)
        c[f"CRPIX{i + 1}"] = (n - 1) / 2. + 1 if crpixs is None else crpixs[i]
        c[f"CRVAL{i + 1}"] = crvals[i]
        c[f"CDELT{i + 1}"] = cdelts[i]
        c[f"CUNIT{i + 1}"] = cunits[i]
        c[f"CTYPE{i + 1}"] = ctypes[i]
    if beam is not None:
        if isinstance(beam, (int, float)):
            bmaj, bmin, bpa = beam, beam, 0.
        else:
            b = list(beam) + [0., 0.]
            bmaj, bmin, bpa = b[0], b[1] if len(beam) > 1 else 0., b[2] if len(beam) > 2 else 0.
        c["BMAJ"], c["BMIN"], c["BPA"] = float(bmaj), float(bmin), float(bpa)
    c["OBJECT"] = obj or "NONE"
    c["BUNIT"] = bunit
    c.update({k.upper(): v for k, v in extra.items()})
    return Header(c)


def beam_area_pix(h):
    """Area of the Gaussian beam in pixels, pi bmaj bmin / (4 ln 2) / pixel area
    (bbarolo/bayesian/likelihood.py:117-120); 1 if the beam is unknown."""
    bmaj, bmin = abs(h.Bmaj()), abs(h.Bmin())
    pa = abs(h.Cdelt(0) * h.Cdelt(1))
    if bmaj > 0 and bmin > 0 and pa > 0:
        return float(math.pi * bmaj * bmin / (4. * math.log(2.)) / pa)
    return 1.
