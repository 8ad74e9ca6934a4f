"""Host side of the GALMOD model engine: header-derived constants, ring conventions and sub-rings.

These run once per fit (``GalmodSetup``) or once per evaluation on a few hundred numbers
(``model_rings``, ``subrings``); the per-voxel work runs in the HIP kernels of ``csrc/``.

Reference behaviour followed (bbarolo/galmod.py):
  * GalmodSetup.__init__ / _NHItoRAD / _velocity_edges   galmod.py:613-758   (row A1)
  * model_rings / _half_widths / edge_weights / outer_edge galmod.py:1055-1141 (row A2)
  * GalmodSetup.subrings / side_edges                      galmod.py:761-863   (row A3)
  * DiskModel.box / zstep                                  galmod.py:1186-1240 (row A7)
  * the Phi table of the ANALYTIC line profiles            galmod.py:174-180
The sub-ring table keeps the reference's layout, sr (16, nsub) float64 with rows
  0 radius  1 vrot  2 vrad  3 vvert  4 vdisp  5 dens  6 z0  7 dvdz  8 zcyl  9 inc  10 phi
  11 xpos  12 ypos  13 vsys  14 weight of the receding half  15 weight of the approaching half
in internal units (arcmin, m/s, radians, pixels), because the kernels and the oracle read it.
"""

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np
from scipy.special import erfc as _erfc

from .header import C_MS, VEL_UNITS, WAVE_UNITS, HI_FREQ, delta_vel

ZSAMPLES_MAX = 1024
ZSTEP_FINE = 0.25
ZSTEP_COARSE = 0.75
ZSTEP_BEAM = 3.0

# Tabulated standard normal CDF Phi(t) on a 1/512 grid over [-6, 6] (galmod.py:174-180)
PHI_TMAX = 6.0
PHI_ISTEP = 512.0
PHI_TAB = np.array([0.5 * math.erfc(-(-PHI_TMAX + k / PHI_ISTEP) / math.sqrt(2.0))
                    for k in range(int(2 * PHI_TMAX * PHI_ISTEP) + 2)])


class GalmodError(ValueError):
    """Invalid model input (the reference raises BBaroloError / ValueError before its kernels run)."""


@dataclass
class GalmodPar:
    """The GALMOD keywords of the model engine, with the reference's defaults (bbarolo/params.py:137-174,
    315-367)."""
    LTYPE: int = 1
    CMODE: int = 1
    CDENS: float = 10.
    ISEED: int = -1
    NV: int = -1
    EMPTY: bool = True
    SM: bool = True
    DENSFLUX: bool = True
    RESTWAVE: List[float] = field(default_factory=lambda: [-1.0])
    RESTFREQ: List[float] = field(default_factory=lambda: [-1.0])
    RELINT: List[float] = field(default_factory=lambda: [1.0])
    RADMAXA: float = -1.
    RADMAXR: float = -1.
    GALMODTYPE: str = "MC"
    RINGMODE: str = "CENTRED"
    SPECRES: float = -1.
    LINEAR: float = -1.
    REDSHIFT: float = 0.


def seed32(iseed):
    """ISEED -> the 32-bit seed of the MC models (galmod.py:593-594; -1 -> 0xFFFFFFFF)."""
    return int(iseed) & 0xFFFFFFFF


# ---------------------------------------------------------------------------------------------------------------------
# Header-derived constants (row A1)

# sky-axis units -> arcmin: the first tag found in CUNIT decides
_SKY_ARCMIN = (("deg", 60.), ("arcs", 1. / 60.), ("arcm", 1.))
_FREQ_HZ = {"hz": 1, "khz": 1E03, "mhz": 1E06, "ghz": 1E09}
_VELSYS = {"radio": 1, "optical": 2, "relativistic": 3}
_FWHM_TO_SIGMA = 1. / (2 * math.sqrt(2 * math.log(2)))
# Rayleigh-Jeans conversion of HI column density to surface brightness per axis type: (Boltzmann / Jy, the line
# constant of the axis); the beam-area factor is per square arcminute (galmod.py:727-741)
_RJ = (1.3806488E-23 / 1E-26)
_NHI_CONST = {2: 2.588E04, 3: 3.8475E-06, 4: 1.823E-05}
_ARCMIN2_SR = (math.pi / (180.0 * 60.0)) ** 2


def _sky_arcmin(cunit):
    cu = cunit.lower()
    for tag, f in _SKY_ARCMIN:
        if tag in cu:
            return f
    raise GalmodError(f"GALMOD error (unknown CUNIT for RA-DEC): cannot convert to ARCMIN.\n{cu}")


def _line_weights(relint, nlines):
    if len(relint) != nlines:
        raise GalmodError(f"GALMOD ERROR: {nlines} lines, but RELINT has {len(relint)} values.")
    return np.array(relint, float)


class SpectralAxis:
    """The third axis of a cube in SI units (m, Hz or m/s) with its reference frequency, the lines of the model
    (reference frequency and relative intensity of each) and the velocity convention (galmod.py:640-712).

    axtyp: 2 wavelength, 3 frequency, 4 velocity; velsys: 1 radio, 2 optical, 3 relativistic."""

    def __init__(self, h, gm):
        unit = h.Cunit(2).lower()
        kind = h.SpectralType()
        z1 = 1 + float(gm.REDSHIFT)
        self.velsys = _VELSYS.get(h.VelDef(), 1)
        self.relint = np.array([1.])
        lines = None
        crval, cdelt = h.Crval(2), h.Cdelt(2)
        if kind == "wave":
            if unit not in WAVE_UNITS:
                raise GalmodError(f"GALMOD ERROR: unknown CUNIT3={unit}. Cannot convert to meters.")
            self.axtyp = 2
            to_si = 1. / WAVE_UNITS[unit]
            crval, cdelt = crval * to_si, cdelt * to_si
            rest = gm.RESTWAVE
            self.freq0 = C_MS / (rest[0] * z1 * to_si) if rest[0] != -1 else C_MS / crval
            if len(rest) > 1:
                lines = C_MS / (np.asarray(rest, float) * z1 * to_si)
                self.relint = _line_weights(gm.RELINT, len(rest))
        elif kind == "freq":
            if unit not in _FREQ_HZ:
                raise GalmodError(f"GALMOD ERROR: unknown CUNIT3={unit}. Cannot convert to Hz.")
            self.axtyp = 3
            to_si = _FREQ_HZ[unit]
            crval, cdelt = crval * to_si, cdelt * to_si
            self.freq0 = h.Freq0() / z1 or HI_FREQ
            if len(gm.RESTFREQ) > 1:
                lines = np.asarray(gm.RESTFREQ, float) / z1
                self.relint = _line_weights(gm.RELINT, len(gm.RESTFREQ))
        elif "velo" in kind:
            if unit not in VEL_UNITS:
                raise GalmodError(f"GALMOD ERROR: unknown CUNIT3={unit}. Cannot convert to m/s.")
            self.axtyp = 4
            to_si = 1.0E03 / VEL_UNITS[unit]
            self.freq0 = h.Freq0() / z1 or HI_FREQ
            crval, cdelt = crval * to_si, cdelt * to_si
        else:
            raise GalmodError("GALMOD ERROR: unknown Spectral axis type. Check CTYPE3 and CUNIT3 keywords.")
        self.crval, self.cdelt, self.crpix = crval, cdelt, h.Crpix(2)
        self.lines = np.ascontiguousarray([self.freq0] if lines is None else lines, dtype=np.float64)
        self.relint = np.ascontiguousarray(self.relint, dtype=np.float64)

    def world(self, pix):
        """The world coordinate (SI) at 0-based pixel positions."""
        return (np.asarray(pix, float) + 1 - self.crpix) * self.cdelt + self.crval

    def velocity(self, world, freq0):
        """Velocity (m/s) of world coordinates for a line of reference frequency freq0 (the inverse of _velgrid)."""
        if self.axtyp == 4:
            return np.ascontiguousarray(world, dtype=np.float64)
        ratio = (C_MS / world if self.axtyp == 2 else world) / freq0
        if self.velsys == 1:
            v = C_MS * (1 - ratio)
        elif self.velsys == 2:
            v = C_MS * (1 / ratio - 1)
        else:
            v = C_MS * (1 - ratio * ratio) / (1 + ratio * ratio)
        return np.ascontiguousarray(v, dtype=np.float64)

    def edges(self, nz, freq0=None):
        """Velocity (m/s) of the nz + 1 channel boundaries (pixels k - 1/2) for the line freq0 (default the first)."""
        return self.velocity(self.world(np.arange(nz + 1) - 0.5), self.freq0 if freq0 is None else freq0)

    def column_density_factor(self, nz):
        """Per channel: HI column density to Jy/beam per square arcminute, rounded to float32 (galmod.py:727-741)."""
        w = self.world(np.arange(nz))
        if self.axtyp == 2:
            lam = w
        elif self.axtyp == 3:
            lam = C_MS / w
        else:
            if self.velsys == 2:
                f = self.freq0 * (C_MS / (C_MS + w))
            elif self.velsys == 3:
                f = self.freq0 * np.sqrt((C_MS - w) / (C_MS + w))
            else:
                f = self.freq0 * (1 - w / C_MS)
            lam = C_MS / f
        per_sr = 2 * _RJ / (_NHI_CONST[self.axtyp] * lam * lam) / abs(self.cdelt)
        return (per_sr * _ARCMIN2_SR).astype(np.float32).astype(np.float64)


class GalmodSetup:
    """Header-derived quantities of the models of a cube (row A1, GalmodSetup.__init__, galmod.py:613-726).
    Units: arcmin for the sky, m/s for velocities, radians for angles."""

    def __init__(self, header, par=None, galmodtype=None, densflux=None):
        h = header
        gm = GalmodPar() if par is None else par
        self.header, self.par = h, gm
        self.nsubs = h.DimAx(2)
        self.dimx, self.dimy = h.DimAx(0), h.DimAx(1)
        # sky pixels
        self.cdelt = [h.Cdelt(i) * _sky_arcmin(h.Cunit(i)) for i in (0, 1)]
        self.pixsize = [abs(c) for c in self.cdelt]
        self.pixarea = self.pixsize[0] * self.pixsize[1]
        self.crota2 = h.Crota() * math.pi / 180.
        self.radsep = 0.75 * min(self.pixsize)
        self.zstep = ZSTEP_FINE
        # spectral axis and lines
        ax = SpectralAxis(h, gm)
        self.axis = ax
        self.axtyp, self.velsys, self.freq0 = ax.axtyp, ax.velsys, ax.freq0
        self.crval3, self.cdelt3, self.crpix3 = ax.crval, ax.cdelt, ax.crpix
        self.lfreq0, self.relint = ax.lines, ax.relint
        # channel width and instrumental broadening (SPECRES, else LINEAR, else a channel of FWHM 1)
        self.chwidth = abs(delta_vel(h)) * 1000
        res = next((v for v in (gm.SPECRES, gm.LINEAR) if v != -1), _FWHM_TO_SIGMA)
        self.sig_instr = res * self.chwidth
        # cloud and layer keywords (out-of-range CMODE / CDENS fall back to 1)
        self.ltype = int(gm.LTYPE)
        self.cmode = int(gm.CMODE) if 0 <= gm.CMODE <= 2 else 1
        self.cdens = float(gm.CDENS) if gm.CDENS >= 0 else 1.
        self.iseed, self.empty, self.nvpar = gm.ISEED, gm.EMPTY, gm.NV
        self.radmaxa, self.radmaxr = float(gm.RADMAXA), float(gm.RADMAXR)
        self.densIsFlux = gm.DENSFLUX if densflux is None else densflux
        self.galmodtype = (galmodtype or gm.GALMODTYPE).upper()
        # channel factors: flux per channel, or the column-density conversion
        if self.densIsFlux:
            self.cd2i = np.full(self.nsubs, np.float32(3600. / self.chwidth * 1000.), dtype=np.float64)
        else:
            self.cd2i = ax.column_density_factor(self.nsubs)
        self.vedges = ax.edges(self.nsubs)
        self.vedges_lines = np.ascontiguousarray([ax.edges(self.nsubs, f) for f in self.lfreq0])
        self.chmin = float(np.min(np.abs(np.diff(self.vedges)))) if self.nsubs > 0 else self.chwidth

    # ------------------------------------------------------------------
    def side_edges(self, rings):
        """(approaching, receding) outer edges (arcsec) of the two halves of the disk: the rings' EDGEAPP / EDGEREC,
        else RADMAXA / RADMAXR; None where there is none (negative) (galmod.py:853-863)."""
        def edge(given, default):
            e = default if given is None else float(given)
            return e if e >= 0 else None
        return (edge(getattr(rings, "edgeapp", None), self.radmaxa),
                edge(getattr(rings, "edgerec", None), self.radmaxr))

    def subrings(self, rings):
        """Sub-rings 0.75 pixel wide interpolated from the rings (row A3, galmod.py:761-851).
        Returns (sr (16, nsub) float64, nv (nsub,) int64)."""
        radsep = self.radsep
        rad = np.asarray(rings.radii, float) / 60.
        gaps = rad[1:] - (rad[:-1] + radsep)
        if np.any(gaps <= 0):
            i = int(np.flatnonzero(gaps <= 0)[0]) + 1
            raise GalmodError("GALMOD ERROR: Radii not in increasing order." if rad[i - 1] > rad[i]
                              else "GALMOD ERROR: Radius separation too small.")
        for key, what in (("vdisp", "Negative velocity dispersion not allowed."),
                          ("dens", "Negative column-density or flux are not allowed."),
                          ("z0", "Negative scale height not allowed.")):
            if np.any(np.asarray(getattr(rings, key)) < 0):
                raise GalmodError("GALMOD ERROR: " + what)
        thick = rings.thick()
        if thick:
            f = np.asarray(rings.fthick)
            if np.any((f < 0) | (f > 1)):
                raise GalmodError("GALMOD ERROR: The fraction of gas in the thick layer (FTHICK) must be 0 to 1.")
        V = ring_values(rings, thick)
        empty = self.empty if rings.empty is None else rings.empty
        idx, dr, const, rads = subring_radii(rad, radsep, empty)
        nsub = len(rads)
        sr = np.empty((V.shape[0] + 1, nsub))
        sr[0] = rads
        if nsub:
            sr[1:] = interp_values(V, rad, idx, dr, const)
        return self.finish_table(sr, rings, thick)

    def finish_table(self, sr, rings, thick):
        """The 16-row table of interpolated values sr (14 rows, or 17 with a thick layer) with the disk edge applied to
        the density, the thin and thick layers as two sets of sub-rings, and the weights of the two halves
        (galmod.py:833-851); and the sub-clouds per sub-ring (NV, or the channels; 1 without dispersion)."""
        radsep = self.radsep
        n = sr.shape[1]
        edge = getattr(rings, "edge", None)
        if edge is not None and n:
            sr[5] *= edge_weights(sr[0] * 60., radsep * 60., float(edge), float(rings.edgewidth or 0.))
        if thick:
            sr = split_layers(sr)
        weights = np.ones((2, sr.shape[1]))
        for row, e in zip((0, 1), self.side_edges(rings)[::-1]):          # row 0 receding, 1 approaching
            if e is not None and sr.shape[1]:
                weights[row] = edge_weights(sr[0] * 60., radsep * 60., e)
        out = np.ascontiguousarray(np.vstack([sr[:14], weights]))
        per = self.nvpar if self.nvpar >= 1 else self.nsubs
        return out, np.where(out[4] == 0, 1, per).astype(np.int64)

    def subcloud_flux(self, sr, nv):
        """Mean flux (before cd2i) of one MC sub-cloud over the sub-rings (galmod.py:879-897)."""
        R0, dens = sr[0], sr[5]
        area_pix = 2 * np.pi * R0 * self.radsep / self.pixarea
        with np.errstate(divide="ignore", invalid="ignore"):
            nc = np.floor(self.cdens * dens ** self.cmode * area_pix + 0.5)
            low = np.floor(nc / area_pix + 0.5) < 10.
            nc = np.where(low, np.floor(10 * area_pix + 0.5), nc)
            w = dens * area_pix
            ok = (w > 0) & (nc > 0)
            if not np.any(ok):
                return 0.
            fsc = w[ok] * self.pixarea / (nc[ok] * nv[ok])
            unit = w[ok].sum() / np.sum(w[ok] / fsc)
        ri = self.relint
        return float(unit * np.sum(ri * ri) / np.sum(ri))


def split_layers(sr):
    """A 17-row table (rows 14-16: FTHICK, thick-layer height and dispersion) as the sub-rings of the thin layer
    (density x (1 - FTHICK), no lag) followed by those of the thick one (density x FTHICK, its own height and
    dispersion)."""
    f = sr[14]
    thin, thk = sr[:14].copy(), sr[:14].copy()
    thin[5] *= 1. - f
    thin[7] = 0.
    thk[5] *= f
    thk[6] = sr[15]
    thk[4] = sr[16]
    return np.concatenate([thin, thk], axis=1)


def _thick_height(z0thick, z0):
    """The thick layer's scale height: Z0THICK, or Z0 where Z0THICK is negative (tied to it)."""
    return np.where(z0thick >= 0, z0thick, z0)


# Rows 1-13 of a sub-ring table from the ring values in the reference's units: (key, conversion); rows 14-16 of a
# thick layer follow (FTHICK, its height, its dispersion). galmod.py:793-804.
TABLE_ROWS = (("vrot", lambda v: v * 1000.), ("vrad", lambda v: v * 1000.), ("vvert", lambda v: v * 1000.),
              ("vdisp", lambda v: v * 1000.), ("dens", lambda v: v), ("z0", lambda v: v / 60.),
              ("dvdz", lambda v: v * 1000 * 60), ("zcyl", lambda v: v / 60.), ("inc", np.radians),
              ("phi", np.radians), ("xpos", lambda v: v), ("ypos", lambda v: v), ("vsys", lambda v: v * 1000.))


def table_rows(get, thick):
    """The rows 1.. of a sub-ring table (13, or 16 with a thick layer) from get(key) -> ring values (any shape)."""
    rows = [conv(np.asarray(get(k), float)) for k, conv in TABLE_ROWS]
    if thick:
        vdt = np.asarray(get("vdispthick"), float)
        vdt = np.where(vdt >= 0, vdt, np.asarray(get("vdisp"), float))
        rows += [np.asarray(get("fthick"), float),
                 _thick_height(np.asarray(get("z0thick"), float), np.asarray(get("z0"), float)) / 60., vdt * 1000.]
    return rows


def ring_values(rings, thick):
    """The ring values in internal units, rows 1..13 of the sub-ring table (+3 rows for a thick layer:
    fthick, z0thick, vdispthick)."""
    return np.vstack(table_rows(lambda k: getattr(rings, k), thick)).astype(float)


def subring_radii(rad, radsep, empty):
    """Sub-ring radii (arcmin) accumulated in float32 as the C++ GALMOD (galmod.py:805-819), with the
    interpolation segment of each: idx (ring index i of the segment [i-1, i]), dr (distance from the
    float32 radius of ring i-1), const (True: inner disk with the values of ring 0). Depends on the radii
    only, so it is cached per fit when the radii are fixed."""
    f32 = np.float32
    rads, idx, drs, const = [], [], [], []
    inner = not (rad[0] != 0 and empty)
    r = f32(radsep / 2.0) if inner else f32(rad[0] + radsep / 2.0)
    if inner:
        r0 = f32(rad[0])
        while r < r0:
            rads.append(float(r)); idx.append(0); drs.append(0.); const.append(True)
            r = f32(float(r) + radsep)
    for i in range(1, len(rad)):
        ri = f32(rad[i])
        base = float(f32(rad[i - 1]))
        while r < ri:
            rads.append(float(r)); idx.append(i); drs.append(float(r) - base); const.append(False)
            r = f32(float(r) + radsep)
    return (np.asarray(idx, np.int64), np.asarray(drs, float), np.asarray(const, bool),
            np.asarray(rads, float))


def interp_values(V, rad, idx, dr, const):
    """Linear interpolation of the ring values V (nrow, nr) at the sub-rings (galmod.py:822-832);
    V may carry a leading batch axis (B, nrow, nr)."""
    nr = len(rad)
    nsub = len(idx)
    i1 = np.maximum(idx, 1)
    if nr > 1:
        dur = rad[i1] - rad[i1 - 1]
        lo = V[..., i1 - 1]
        slope = (V[..., i1] - lo) / dur
        vals = lo + slope * dr
    else:
        vals = V[..., np.zeros(nsub, np.int64)] * 1.0
    if np.any(const):
        vals[..., const] = V[..., [0]]
    return vals


# ----------------------------------------------------------------------------------------------
# Ring conventions (row A2)
# ----------------------------------------------------------------------------------------------
RINGMODES = ("CENTRED", "NODES", "SHIFTED")


def ringmode_name(mode):
    m = str(mode).strip().upper()
    m = "CENTRED" if m == "CENTERED" else m
    if m not in RINGMODES:
        raise ValueError(f"Unknown ring convention {mode!r}: RINGMODE can be CENTRED, NODES or SHIFTED.")
    return m


def half_widths(rings):
    """Half widths of the first and last ring of CENTRED (galmod.py:1055-1063)."""
    R = rings.radii
    if rings.nr == 1:
        w = rings.radsep / 2. if rings.radsep > 0 else R[0] / 2.
        return w, w
    return (R[1] - R[0]) / 2., (R[-1] - R[-2]) / 2.


def model_rings(rings, mode="CENTRED", subring=0.):
    """The rings that build the model of a set of rings with a ring convention (galmod.py:1066-1119):
    CENTRED adds the inner and outer half rings with the values of the first and last ring; NODES keeps
    the rings; SHIFTED moves the first and last radius by half of RADSEP."""
    mode = ringmode_name(mode)
    if mode == "NODES" and rings.nr == 1:
        raise GalmodError("RINGMODE NODES needs two radii at least.")
    if mode == "SHIFTED" and rings.nr == 1 and not rings.radsep > 0:
        raise GalmodError("RINGMODE SHIFTED with one ring needs RADSEP > 0.")
    if mode == "NODES" or rings.nr == 0:
        return rings
    r = rings.copy()
    if mode == "SHIFTED":
        if r.nr == 1:
            vals = {k: getattr(r, k)[0] for k in r.as_dict()}
            vals["radii"] = r.radii[0] + r.radsep / 2.
            r.addRing(**vals)
        else:
            r.radii[-1] += r.radsep / 2.
        r.radii[0] = max(r.radii[0] - r.radsep / 2., 0.)
        return r
    R = r.radii
    w0, wn = half_widths(r)
    rin, rout = max(R[0] - w0, 0.), R[-1] + wn
    idx, rad = list(range(r.nr)), list(R)
    if rin / 60. + subring < R[0] / 60.:
        idx.insert(0, 0)
        rad.insert(0, rin)
    if R[-1] / 60. + subring < rout / 60.:
        idx.append(r.nr - 1)
        rad.append(rout)
    out = r.subset(np.array(idx))
    out.radii = np.array(rad, dtype=float)
    return out


def edge_weights(R, dr, edge, width=0.):
    """Density weights of sub-rings at radii R (arcsec, dr wide) for a disk edge at `edge` arcsec
    (galmod.py:1122-1131)."""
    R = np.asarray(R, float)
    if width > 0:
        return 0.5 * _erfc((R - edge) / (math.sqrt(2.) * width))
    return np.clip((edge - (R - dr / 2.)) / dr, 0., 1.) if dr > 0 else (R <= edge).astype(float)


def outer_edge(rings, mode="CENTRED"):
    """Outer radius (arcsec) of the disk of the rings (galmod.py:1134-1141)."""
    mode = ringmode_name(mode)
    if mode == "NODES":
        return float(rings.radii[-1])
    if mode == "SHIFTED":
        return float(rings.radii[-1] + rings.radsep / 2.)
    return float(rings.radii[-1] + half_widths(rings)[1])


def model_box(h, rings, ringmode="CENTRED", final=False):
    """Box [blo, bhi) of the cube (pixels) that holds the model of the rings (galmod.py:1224-1240)."""
    from .header import arcsconv
    conv = arcsconv(h.Cunit(0))
    ext = rings.radii[-1] + rings.radsep / 2.
    if final:
        ext = max(ext, outer_edge(rings, ringmode))
    ext = ext + 3 * float(rings.layer_height()[-1])
    xdis = int(math.ceil(ext / (abs(h.Cdelt(0)) * conv)))
    ydis = int(math.ceil(ext / (abs(h.Cdelt(1)) * conv)))
    xpos = int(math.ceil(rings.xpos[-1]))
    ypos = int(math.ceil(rings.ypos[-1]))
    blo = [max(xpos - xdis, 0), max(ypos - ydis, 0)]
    bhi = [min(xpos + xdis + 1, h.DimAx(0)), min(ypos + ydis + 1, h.DimAx(1))]
    return blo, bhi


def zstep_for(h, setup, smoothed):
    """The projected spacing of the ANALYTIC vertical samples (galmod.py:1186-1198)."""
    if not smoothed:
        return ZSTEP_FINE
    pix = max(setup.pixsize) / 60.
    bmin = min(abs(h.Bmaj()), abs(h.Bmin()))
    return ZSTEP_COARSE if pix > 0 and bmin / pix >= ZSTEP_BEAM else ZSTEP_FINE


def nzq_for(sr):
    """Vertical-sample capacity of the ANALYTIC kernels: 1 for thin disks, else ZSAMPLES_MAX
    (galmod.py:909). sr: one sub-ring table (16, nsub) or a batch of them (B, 16, nsub)."""
    return 1 if not np.any(np.asarray(sr)[..., 6, :] > 0) else ZSAMPLES_MAX


class SubringPlan:
    """The part of GalmodSetup.subrings that depends only on the radii, for batches of rings that share
    them (the radii are not free in most fits): the model rings of the ring convention (model_rings), the
    float32-accumulated sub-ring radii and their interpolation segments. tables(V) then gives the sub-ring
    tables of a batch of ring values with numpy only (no Python loop per parameter vector).

    rings: a Rings giving radii, radsep, empty, edges and the layer switch of the batch."""

    def __init__(self, setup, rings, ringmode="CENTRED", final=True):
        self.setup = setup
        self.thick = rings.thick()
        probe = rings.copy()
        probe.dens = np.arange(rings.nr, dtype=float)       # (to find which input ring each model ring copies)
        mr = model_rings(probe, ringmode, setup.radsep) if final else probe
        self.ring_index = np.asarray(mr.dens, np.int64)
        mr = model_rings(rings, ringmode, setup.radsep) if final else rings
        self.nr_in = rings.nr
        rad = np.asarray(mr.radii, float) / 60.
        for i in range(1, mr.nr):
            if rad[i - 1] + setup.radsep >= rad[i]:
                raise GalmodError("GALMOD ERROR: Radii not in increasing order." if rad[i - 1] > rad[i]
                                  else "GALMOD ERROR: Radius separation too small.")
        self.rad = rad
        empty = setup.empty if mr.empty is None else mr.empty
        self.idx, self.dr, self.const, self.rads = subring_radii(rad, setup.radsep, empty)
        self.nsub = len(self.rads)
        self.proto = mr

    def tables(self, values):
        """values: dict of ring arrays (B, nr) (or (nr,)) in the reference's units for the keys of
        ring_values (missing keys: those of the rings the plan was built with). Returns sr (B, 16, nsub') and
        nv (nsub',), nsub' = 2 nsub with a thick layer."""
        B = None
        for v in values.values():
            a = np.asarray(v)
            if a.ndim == 2:
                B = a.shape[0]
        B = 1 if B is None else B
        p = self.proto
        ri = self.ring_index
        nm = len(ri)

        def get(k):
            """(B, nr_model) values of k: given per input ring (mapped to the model rings), or the plan's."""
            if k in values:
                a = np.asarray(values[k], float)
                if a.ndim == 0:
                    a = np.full(self.nr_in, float(a))
                if a.ndim == 1:
                    a = a[None]
                return np.broadcast_to(a[:, ri], (B, nm))
            return np.broadcast_to(np.asarray(getattr(p, k), float)[None], (B, nm))
        rows = [np.broadcast_to(r, (B, nm)) for r in table_rows(get, self.thick)]
        V = np.stack(rows, axis=1)                                  # (B, nrow, nr_model)
        sr = np.empty((B, V.shape[1] + 1, self.nsub))
        sr[:, 0] = self.rads
        if self.nsub:
            sr[:, 1:] = interp_values(V, self.rad, self.idx, self.dr, self.const)
        out = []
        nv = None
        for b in range(B) if (self.thick or self._edges()) else ():
            s, nv = self.setup.finish_table(sr[b], p, self.thick)
            out.append(s)
        if out:
            return np.ascontiguousarray(np.stack(out)), nv
        full = np.empty((B, 16, self.nsub))
        full[:, :14] = sr[:, :14]
        full[:, 14:] = 1.
        nvtmp = self.setup.nvpar if self.setup.nvpar >= 1 else self.setup.nsubs
        nv = np.where(full[0, 4] == 0, 1, nvtmp).astype(np.int64)
        return full, nv

    def _edges(self):
        p = self.proto
        return getattr(p, "edge", None) is not None or any(e is not None for e in self.setup.side_edges(p))

    def weights(self):
        """(densw, side): the edge weights of the density of the nsub sub-rings of one layer (None without an edge
        of the disk) and the weights (2, nsub_total) of the receding and approaching halves, as finish_table
        applies them (galmod.py:833-851)."""
        p = self.proto
        radsep = self.setup.radsep
        densw = None
        if getattr(p, "edge", None) is not None and self.nsub:
            densw = np.ascontiguousarray(edge_weights(self.rads * 60., radsep * 60., float(p.edge),
                                                      float(p.edgewidth or 0.)))
        R = np.concatenate([self.rads, self.rads]) if self.thick else self.rads
        side = np.ones((2, R.size))
        for k, e in enumerate(self.setup.side_edges(p)[::-1]):          # (receding, approaching)
            if e is not None and R.size:
                side[k] = edge_weights(R * 60., radsep * 60., e)
        return densw, np.ascontiguousarray(side)


# the ring keys of the likelihood context, in the order of its ring table (bbgpu.h BB_K_*)
LIKE_KEYS = ("vrot", "vrad", "vvert", "vdisp", "dens", "z0", "dvdz", "zcyl", "inc", "phi", "xpos", "ypos", "vsys",
             "fthick", "z0thick", "vdispthick", "sigma", "fracerr", "radmax", "radmaxa", "radmaxr")


def ring_table(rings):
    """(len(LIKE_KEYS), nr) float64: the ring values of the keys in the reference's units (the single values sigma,
    fracerr and the edges: 0)."""
    out = np.zeros((len(LIKE_KEYS), rings.nr))
    for k, key in enumerate(LIKE_KEYS[:16]):
        out[k] = np.asarray(getattr(rings, key), float)
    return out
