"""Parameter files of the fits on the likelihood path.

  * read_parfile (tasks.py:1203-1224): one KEYWORD value per line, case-insensitive keywords, blank lines and lines
    starting with '#' or '/' skipped, a later line overriding an earlier one, unknown keywords skipped with a warning;
  * the keywords the path reads, with the reference's types and defaults (keywords.py:271-500, params.py:137-216,
    315-367) and the task defaults of METHOD GLOBAL / BAYESIAN (keywords.py:695-712: GALMODTYPE ANALYTIC and the
    default FREE); the values are read as the reference reads a parameter file (keywords.py:19-152: the leading
    number of the first token, '1' or 'true' for a flag, ...);
  * the initial rings (readRings, rings.py:262-360): NRADII and RADSEP (radii (i + 1/2) RADSEP), or RADII as a list,
    range(Rmin,Rmax,step) or Rmin~Rmax:step (Rmax included, rings.py:157-174); a ring value given once for all the
    rings, once per ring, or as file(name,column) of a ring table; XPOS and YPOS in pixels;
  * the options of METHOD GLOBAL and BAYESIAN: LIKEOPTS (params.py:1447-1560), SAMPLEROPTS, RUNOPTS.

ParFile turns them into this build's objects: GalmodPar, Fit3DPar, Rings, the FreePars spec of a GLOBAL or BAYESIAN
fit, and the CLASSIC driver (ClassicFit) on the cube of FITSFILE. WCS positions (07h36m51s, 123.4d) are not read.
"""

import math
import os
import re
import warnings

import numpy as np

# ---------------------------------------------------------------------------------------------------------------------
# value readers (the parameter-file semantics of the reference: C++ ``ss >> value`` on the rest of the line)

_INT = re.compile(r"^[+-]?\d+")
_FLOAT = re.compile(r"^[+-]?((\d+\.?\d*)|(\.\d+))([eE][+-]?\d+)?")


def _first(s):
    t = str(s).split()
    return t[0] if t else ""


def read_flag(s):
    """'1' or 'true' (any case) in the first token: True."""
    t = _first(s)
    return t == "1" or t.lower() == "true"


def read_int(s, default=0):
    m = _INT.match(_first(s))
    return int(m.group(0)) if m else default


def read_float(s, default=0.0):
    tok = _first(s)
    m = _FLOAT.match(tok)
    if m:
        return float(m.group(0))
    try:
        return float(tok)              # (nan, inf)
    except ValueError:
        return default


def read_vec(s, typ=float):
    """The numbers of a list (spaces or commas), up to the first that is not a number."""
    out = []
    for t in str(s).replace(",", " ").split():
        try:
            out.append(int(float(t)) if typ is int else float(t))
        except (ValueError, OverflowError):
            break
    return out


def read_beam(s):
    """BMAJ BMIN BPA (arcsec, or degrees with a final d; BPA degrees); one value: a circular beam. [bmaj, bmin, bpa]
    (bpa None if not given) or None for an empty value."""
    toks = str(s).replace(",", " ").split()
    if not toks:
        return None
    if len(toks) > 3:
        raise ValueError(f"BEAM: 'BMAJ BMIN BPA' or one value, not {s!r}")
    vals = []
    for i, t in enumerate(toks):
        deg = i < 2 and t[-1:] in ("d", "D")
        x = float(t[:-1] if deg else t)
        if not math.isfinite(x) or (i < 2 and x <= 0):
            raise ValueError(f"BEAM: the axes must be positive numbers, not {s!r}")
        vals.append(x * 3600. if deg else x)
    return [vals[0], vals[1] if len(vals) > 1 else vals[0], vals[2] if len(vals) > 2 else None]


_READ = {
    "str": lambda s: str(s).strip(),
    "ustr": lambda s: str(s).strip().upper(),
    "lstr": lambda s: str(s).strip().lower(),
    "pos": lambda s: str(s).strip() if "file(" in str(s).lower() else str(s).strip().lower(),
    "flag": read_flag,
    "int": read_int,
    "float": read_float,
    "cdens": lambda s: (lambda x: int(x) if math.isfinite(x) else 0)(read_float(s)),
    "fvec": read_vec,
    "beam": read_beam,
}

# KEYWORD: (type, default, aliases). Ring values are strings (one value, one per ring, file(...)); "-1": not given.
KEYWORDS = {
    # general
    "FITSFILE": ("str", "", ()),
    "OUTFOLDER": ("str", "", ()),
    "VERBOSE": ("int", 1, ()),
    "THREADS": ("int", 16, ()),
    "BEAM": ("beam", None, ()),
    "MASK": ("str", "SEARCH", ()),
    "3DFIT": ("flag", False, ("FIT3D",)),
    "GALMOD": ("flag", False, ()),
    # the model (GALMOD keys of params.py:137-174 and the engine's options)
    "NRADII": ("int", -1, ()),
    "RADII": ("str", "-1", ()),
    "RADSEP": ("float", -1.0, ()),
    "XPOS": ("pos", "-1", ()),
    "YPOS": ("pos", "-1", ()),
    "VSYS": ("str", "-1", ()),
    "VROT": ("str", "-1", ()),
    "VRAD": ("str", "-1", ()),
    "VVERT": ("str", "-1", ()),
    "VDISP": ("str", "-1", ()),
    "INC": ("str", "-1", ()),
    "PA": ("str", "-1", ("PHI",)),
    "Z0": ("str", "-1", ()),
    "DENS": ("str", "-1", ()),
    "DVDZ": ("str", "-1", ()),
    "ZCYL": ("str", "-1", ()),
    "Z0THICK": ("str", "-1", ()),
    "FTHICK": ("str", "-1", ()),
    "VDISPTHICK": ("str", "-1", ()),
    "RADMAXA": ("float", -1.0, ()),
    "RADMAXR": ("float", -1.0, ()),
    "CDENS": ("cdens", 10, ()),
    "CMODE": ("int", 1, ()),
    "ISEED": ("int", -1, ()),
    "NV": ("int", -1, ()),
    "SM": ("flag", True, ()),
    "LTYPE": ("int", 1, ()),
    "REDSHIFT": ("float", 0.0, ()),
    "RESTWAVE": ("fvec", [-1.0], ()),
    "RESTFREQ": ("fvec", [-1.0], ()),
    "RELINT": ("fvec", [1.0], ()),
    "EMPTY": ("flag", True, ()),
    "DENSFLUX": ("flag", True, ()),
    "NOISERMS": ("float", 0.0, ()),
    "GALMODTYPE": ("ustr", "MC", ()),
    "RINGMODE": ("ustr", "CENTRED", ()),
    "SPECRES": ("float", -1.0, ()),
    "LINEAR": ("float", -1.0, ()),
    # 3DFIT (params.py:177-216)
    "DELTAINC": ("float", 5.0, ()),
    "DELTAPA": ("float", 15.0, ("DELTAPHI",)),
    "DELTAVROT": ("float", 1000.0, ()),
    "MINVDISP": ("float", 0.1, ()),
    "MAXVDISP": ("float", 1000.0, ()),
    "FTYPE": ("int", 2, ()),
    "WFUNC": ("int", 2, ()),
    "TOL": ("float", 1e-3, ()),
    "FREE": ("str", "VROT VDISP INC PA", ()),
    "SIDE": ("ustr", "B", ()),
    "BWEIGHT": ("float", 1.0, ()),
    "TWOSTAGE": ("flag", True, ()),
    "REGTYPE": ("lstr", "auto", ()),
    "FLAGERRORS": ("flag", False, ("ERRORS",)),
    "NORM": ("ustr", "AZIM", ()),
    "REVERSE": ("lstr", "false", ()),
    "SWEEPS": ("int", 1, ()),
    "RINGREGION": ("ustr", "STRIP", ()),
    "NORMALCUBE": ("flag", True, ()),
    "STARTRAD": ("int", 0, ()),
    "METHOD": ("ustr", "", ()),
    "SAMPLER": ("ustr", "NAUTILUSBB", ()),
    "PRIORS": ("str", "", ()),
    "SAMPLEROPTS": ("str", "", ("SAMPLER_KWARGS",)),
    "RUNOPTS": ("str", "", ("RUN_KWARGS",)),
    "LIKEOPTS": ("str", "", ("LIKE_KWARGS",)),
}

_LOOKUP = {}
for _name, (_t, _d, _al) in KEYWORDS.items():
    _LOOKUP[_name.lower()] = _name
    for _a in _al:
        _LOOKUP[_a.lower()] = _name

# FREE of 3DFIT with METHOD GLOBAL or BAYESIAN when not given; the models of those methods are ANALYTIC
BAYES_DEFAULT_FREE = "VROT VDISP XPOS_SINGLE YPOS_SINGLE VSYS_SINGLE INC_SINGLE PA_SINGLE Z0_SINGLE RADMAX FRACERR"
_LIKE_DEFAULTS = {"GALMODTYPE": "ANALYTIC", "FREE": BAYES_DEFAULT_FREE}


def lookup(name):
    """The canonical KEYWORD of a name or alias (any case), or None."""
    return _LOOKUP.get(str(name).strip().lower())


def read_parfile(fname):
    """{KEYWORD: value string} of a parameter file, in the order of the file (a keyword given twice: the last value);
    unknown keywords are skipped with a warning."""
    if not os.path.isfile(fname):
        raise FileNotFoundError(f"Could not open parameter file {fname}")
    out = {}
    with open(fname) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] in "#/":
                continue
            parts = line.split(None, 1)
            kw = lookup(parts[0])
            if kw is None:
                warnings.warn(f"parameter file {fname}: keyword {parts[0]} is not read by this build, skipped")
                continue
            out.pop(kw, None)
            out[kw] = parts[1] if len(parts) > 1 else ""
    return out


# ---------------------------------------------------------------------------------------------------------------------
# options: key=value lists (LIKEOPTS, SAMPLEROPTS, RUNOPTS)

def _split_top(s, seps):
    """Split s at the characters seps that are outside parentheses."""
    out, depth, cur = [], 0, []
    for c in s:
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        if depth == 0 and c in seps:
            if "".join(cur).strip():
                out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(c)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def _opt_value(v):
    """A value of an option: (a, b) a tuple, (k=v, ...) a dict, a number, true/false, or the string."""
    v = v.strip()
    if v.startswith("(") and v.endswith(")"):
        items = _split_top(v[1:-1], ",")
        if items and all("=" in it for it in items):
            return {k.strip().lower(): _opt_value(x) for k, x in (it.split("=", 1) for it in items)}
        return tuple(_opt_value(it) for it in items)
    if "," in v and "(" not in v:
        return tuple(_opt_value(it) for it in v.split(","))
    low = v.lower()
    if low in ("true", "false"):
        return low == "true"
    try:
        f = float(v)
        return int(f) if re.fullmatch(r"[+-]?\d+", v) else f
    except ValueError:
        return v


def parse_options(s):
    """'key=value key2=(a,b) key3=(x=1, y=2)' -> {key: value} (keys lower case)."""
    out = {}
    for tok in _split_top(str(s or ""), " \t"):
        if "=" not in tok:
            raise ValueError(f"option {tok!r}: expected key=value")
        k, v = tok.split("=", 1)
        out[k.strip().lower()] = _opt_value(v)
    return out


LIKE_DEFAULTS = dict(likemask="dilated", dilation=(1.0, 3), noisecorr="measured", densprof_smooth=0.3, bbres=False,
                     radmax="edge", radmax_width=None, sigma=None, fracerr=None, smooth=None)
_LIKE_ALIASES = {"usebbres": "bbres"}


def like_options(s):
    """The likelihood options of LIKEOPTS over their defaults (params.py:1447-1560), checked as the reference checks
    them; raises ValueError."""
    out = dict(LIKE_DEFAULTS)
    for k, v in parse_options(s).items():
        key = _LIKE_ALIASES.get(k, k)
        if key not in LIKE_DEFAULTS:
            raise ValueError(f"unknown option '{k}' (options: " + ", ".join(LIKE_DEFAULTS) + ")")
        if isinstance(v, str) and v.lower() == "none" and key != "likemask":
            out[key] = LIKE_DEFAULTS[key]
            continue

        def number(x, positive=False):
            if isinstance(x, bool) or not isinstance(x, (int, float)):
                raise ValueError(f"{key}: expected a number, not {v!r}")
            if not math.isfinite(x) or x < 0 or (positive and x == 0):
                raise ValueError(f"{key}: expected a number {'>' if positive else '>='} 0, not {v!r}")
            return float(x)

        if key == "likemask":
            if str(v).lower() not in ("dilated", "none"):
                raise ValueError(f"likemask: 'dilated' or 'none', not {v!r}")
            out[key] = str(v).lower()
        elif key == "dilation":
            if not (isinstance(v, tuple) and len(v) == 2):
                raise ValueError(f"dilation: (radius in beams, channels), not {v!r}")
            out[key] = (number(v[0]), int(round(number(v[1]))))
        elif key == "noisecorr":
            out[key] = str(v).lower() if str(v).lower() in ("measured", "beam") else number(v, positive=True)
        elif key == "bbres":
            if not isinstance(v, (bool, int)) or v not in (0, 1):
                raise ValueError(f"bbres: true or false, not {v!r}")
            out[key] = bool(v)
        elif key == "radmax":
            if str(v).lower() not in ("ring", "edge", "taper"):
                raise ValueError(f"radmax: 'ring', 'edge' or 'taper', not {v!r}")
            out[key] = str(v).lower()
        elif key == "smooth":
            if not isinstance(v, dict):
                raise ValueError(f"smooth: the bends allowed over a beam, e.g. (vrot=5, vdisp=3), not {v!r}")
            out[key] = {kk: number(vv, positive=True) for kk, vv in v.items()}
        else:                                   # densprof_smooth, sigma, fracerr, radmax_width
            out[key] = number(v, positive=(key == "sigma"))
    return out


# ---------------------------------------------------------------------------------------------------------------------
# rings

_RANGE = re.compile(r"range\(\s*([-+.\deE]+)\s*,\s*([-+.\deE]+)\s*,\s*([-+.\deE]+)\s*\)", re.I)
_TILDE = re.compile(r"([-+.\deE]+)\s*~\s*([-+.\deE]+)\s*:\s*([-+.\deE]+)")
_FILECOL = re.compile(r"file\(\s*([^,)]+?)\s*(?:,\s*(\d+)\s*)?(?:,[^)]*)?\)", re.I)
RING_KEYS = (("xpos", "XPOS"), ("ypos", "YPOS"), ("vsys", "VSYS"), ("vrot", "VROT"), ("vrad", "VRAD"),
             ("vvert", "VVERT"), ("dvdz", "DVDZ"), ("zcyl", "ZCYL"), ("vdisp", "VDISP"), ("z0", "Z0"),
             ("dens", "DENS"), ("inc", "INC"), ("phi", "PA"), ("z0thick", "Z0THICK"), ("fthick", "FTHICK"),
             ("vdispthick", "VDISPTHICK"))


def radii_range(s):
    """RADII as range(Rmin,Rmax,step) or Rmin~Rmax:step: Rmin, Rmin + step, ... up to Rmax included (a tolerance of
    1e-6 step for steps that are not exact in binary); None for another form."""
    s = str(s).strip()
    m = _RANGE.fullmatch(s) or _TILDE.fullmatch(s)
    if m is None:
        return None
    lo, hi, step = (float(x) for x in m.groups())
    if hi < lo or not step > 0:
        raise ValueError(f"RADII {s}: Rmax must not be smaller than Rmin and the step must be positive")
    n = int(np.floor((hi - lo) / step + 1e-6)) + 1
    return list(lo + np.arange(n) * step)


def _ring_tokens(s):
    return str(s).replace(",", " ").split()


def _ring_list(s):
    """One value per ring ('120 125 130' or '120,125,130'), or None for a single value."""
    t = _ring_tokens(s)
    if len(t) < 2:
        return None
    try:
        return [float(x) for x in t]
    except ValueError:
        return None


def _file_column(val, base):
    """file(name,n): column n (1-based, default 1) of a text table (lines starting with # skipped)."""
    m = _FILECOL.search(str(val))
    name, n = m.group(1), int(m.group(2) or 1)
    path = name if os.path.isabs(name) or os.path.exists(name) or not base else os.path.join(base, name)
    out = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            try:
                out.append(float(t[n - 1]))
            except (ValueError, IndexError):
                continue
    return out


# ---------------------------------------------------------------------------------------------------------------------

class ParFile:
    """The parameters of a fit: {KEYWORD: value} (value strings as in a file, or Python values) over the defaults of
    the reference; base: the folder relative file names are looked up in (that of the parameter file)."""

    def __init__(self, params=None, base=""):
        self.raw = {}
        for k, v in (params or {}).items():
            kw = lookup(k)
            if kw is None:
                raise KeyError(f"unknown keyword {k}")
            self.raw[kw] = v
        self.base = base

    @classmethod
    def read(cls, fname):
        return cls(read_parfile(fname), base=os.path.dirname(os.path.abspath(fname)))

    # -- values
    def method(self):
        """METHOD of 3DFIT (default CLASSIC)."""
        m = self.get("METHOD")
        return m or "CLASSIC"

    def get(self, key):
        """The value of a keyword: read from its string, or the task default (METHOD GLOBAL / BAYESIAN: GALMODTYPE
        ANALYTIC and the free set of the likelihood), or the reference's default."""
        kw = lookup(key)
        if kw is None:
            raise KeyError(f"unknown keyword {key}")
        typ, default, _ = KEYWORDS[kw]
        if kw in self.raw:
            v = self.raw[kw]
            return _READ[typ](v) if isinstance(v, str) else v
        if kw in _LIKE_DEFAULTS and kw != "METHOD" and self.method() in ("GLOBAL", "BAYESIAN"):
            return _READ[typ](_LIKE_DEFAULTS[kw])
        return list(default) if isinstance(default, list) else default

    def galmod_par(self):
        from .galmod import GalmodPar
        return GalmodPar(LTYPE=self.get("LTYPE"), CMODE=self.get("CMODE"), CDENS=float(self.get("CDENS")),
                         ISEED=self.get("ISEED"), NV=self.get("NV"), EMPTY=self.get("EMPTY"), SM=self.get("SM"),
                         DENSFLUX=self.get("DENSFLUX"), RESTWAVE=self.get("RESTWAVE"),
                         RESTFREQ=self.get("RESTFREQ"), RELINT=self.get("RELINT"), RADMAXA=self.get("RADMAXA"),
                         RADMAXR=self.get("RADMAXR"), GALMODTYPE=self.get("GALMODTYPE"),
                         RINGMODE=self.get("RINGMODE"), SPECRES=self.get("SPECRES"), LINEAR=self.get("LINEAR"),
                         REDSHIFT=self.get("REDSHIFT"))

    def fit3d_par(self):
        from .galfit import Fit3DPar
        return Fit3DPar(FREE=self.get("FREE"), DELTAVROT=self.get("DELTAVROT"), DELTAINC=self.get("DELTAINC"),
                        DELTAPHI=self.get("DELTAPA"), MINVDISP=self.get("MINVDISP"), MAXVDISP=self.get("MAXVDISP"),
                        FTYPE=self.get("FTYPE"), WFUNC=self.get("WFUNC"), TOL=self.get("TOL"),
                        BWEIGHT=self.get("BWEIGHT"), NORM=self.get("NORM"), SIDE=self.get("SIDE"),
                        REGTYPE=self.get("REGTYPE"), REVERSE=self.get("REVERSE"), SWEEPS=self.get("SWEEPS"),
                        RINGREGION=self.get("RINGREGION"))

    def like_options(self):
        return like_options(self.get("LIKEOPTS"))

    def sampler_options(self):
        return parse_options(self.get("SAMPLEROPTS"))

    def run_options(self):
        return parse_options(self.get("RUNOPTS"))

    def rings(self, header=None):
        """The initial rings (readRings): Rings with every ring key. XPOS and YPOS in pixels (a WCS position is not
        read: NotImplementedError)."""
        from .rings import Rings
        fr = {}
        for key, pkey in (("radii", "RADII"),) + RING_KEYS:
            val = str(self.get(pkey))
            fr[key] = _file_column(val, self.base) if "file(" in val.lower() else None
        for key, pkey in RING_KEYS:
            if fr[key] is None:
                fr[key] = _ring_list(self.get(pkey))
        radii_b = fr["radii"] is not None
        if not radii_b and str(self.get("RADII")) != "-1":
            fr["radii"] = radii_range(self.get("RADII"))
            if fr["radii"] is None:
                fr["radii"] = [float(x) for x in _ring_tokens(self.get("RADII"))]
            radii_b = len(fr["radii"]) > 0
        sizes = [len(v) for v in fr.values() if v is not None and len(v) > 0]
        max_size = min(sizes) if sizes else 2 ** 31

        def pval(pkey, dflt):
            s = str(self.get(pkey))
            if s == "-1" or "file(" in s.lower():
                return dflt
            try:
                return float(_ring_tokens(s)[0])
            except (ValueError, IndexError):
                return dflt

        defaults = dict(vsys=pval("VSYS", 0), vrot=pval("VROT", 0), vrad=pval("VRAD", 0), vvert=pval("VVERT", 0),
                        dvdz=pval("DVDZ", 0), zcyl=pval("ZCYL", 0), vdisp=pval("VDISP", 0), z0=pval("Z0", 0),
                        dens=pval("DENS", 1), inc=pval("INC", 0), phi=pval("PA", 0), z0thick=pval("Z0THICK", 0),
                        fthick=pval("FTHICK", 0), vdispthick=pval("VDISPTHICK", -1))
        for key in ("xpos", "ypos"):
            s = str(self.get(key.upper()))
            if fr[key] is None and s != "-1":
                try:
                    defaults[key] = float(s)
                except ValueError:
                    raise NotImplementedError(f"{key.upper()} {s}: a WCS position is not read by this build; give "
                                              "it in pixels") from None
            elif fr[key] is None:
                defaults[key] = 0.
        nr, radsep = self.get("NRADII"), self.get("RADSEP")
        nr = nr if 0 < nr < max_size else max_size
        if radii_b:
            rr = fr["radii"]
            if nr == 1:
                radsep = rr[0] / 2. if radsep == -1 else radsep
            elif len(rr) > 1:
                radsep = float(np.mean(np.diff(rr)))
        if nr >= 2 ** 31:
            nr = 0
        vals = {}
        for key in ("radii",) + tuple(k for k, _ in RING_KEYS):
            if key == "radii":
                vals[key] = np.asarray(fr["radii"][:nr], float) if radii_b else np.arange(nr) * radsep + radsep / 2.
            else:
                vals[key] = np.asarray(fr[key][:nr], float) if fr[key] is not None else np.full(nr, float(defaults[key]))
        r = Rings().setRings(**vals)
        r.radsep = radsep
        return r

    def free_spec(self, nr):
        """The free parameters of a GLOBAL or BAYESIAN fit as the (key, 'ring' | 'single') list of FreePars: NAME one
        value per ring, NAME_SINGLE one for all, SIGMA and FRACERR single (params.py:1156-1180)."""
        names = {"pa": "phi", "phi": "phi"}
        spec = []
        for tok in str(self.get("FREE")).split():
            t = tok.lower()
            mode = "single" if t.endswith("_single") else "ring"
            base = t[:-7] if mode == "single" else t
            if base.endswith("_func") or "_p" in base:
                raise NotImplementedError(f"FREE {tok}: functions of radius are not built")
            base = names.get(base, base)
            if base in ("sigma", "fracerr", "radmax", "radmaxa", "radmaxr"):
                mode = "single"
            spec.append((base, mode))
        return spec

    def cube(self):
        """(Header, data (nz, ny, nx) float32) of FITSFILE (relative to the parameter file's folder if not found)."""
        from .fitsio import load_cube
        f = str(self.get("FITSFILE"))
        if not f:
            raise ValueError("FITSFILE is not given")
        path = f if os.path.exists(f) or not self.base else os.path.join(self.base, f)
        return load_cube(path)

    def classic_fit(self, header=None, data=None, mask=None):
        """The CLASSIC 3DFIT of the parameters (ClassicFit), on the cube of FITSFILE unless header and data are given;
        mask (nz, ny, nx) bool or None (the stand-in of MASK SEARCH)."""
        from .classic import ClassicFit
        if self.method() != "CLASSIC":
            raise ValueError(f"METHOD {self.method()}: classic_fit is METHOD CLASSIC")
        if header is None:
            header, data = self.cube()
        return ClassicFit(header, data, self.rings(header), par=self.fit3d_par(), gpar=self.galmod_par(), mask=mask,
                          normalcube=self.get("NORMALCUBE"), twostage=self.get("TWOSTAGE"))
