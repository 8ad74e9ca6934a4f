"""Tilted rings: the input layout of the model-cube likelihood.

Same fields, defaults and helpers as the reference's ``Rings`` (bbarolo/rings.py:15-144): each
parameter is a float64 array of length nr; radii and heights in arcsec, velocities in km/s, angles in
degrees, centres in pixels (0-based).
"""

import numpy as np

RING_KEYS = ("radii", "xpos", "ypos", "vsys", "vrot", "vdisp", "vrad", "vvert",
             "dvdz", "zcyl", "dens", "z0", "inc", "phi", "z0thick", "fthick", "vdispthick")

_DEFAULTS = dict(xpos=0, ypos=0, vsys=0, vrot=0, vdisp=0, vrad=0, vvert=0, dvdz=0, zcyl=0,
                 dens=1, z0=0, inc=0, phi=0, z0thick=0, fthick=0, vdispthick=-1)


class Rings:
    """A set of tilted rings (bbarolo/rings.py:15). Attributes beyond the ring arrays: radsep (width of
    a single ring), edge/edgewidth (outer edge of the disk, RADMAX), edgeapp/edgerec (edges of the
    approaching / receding halves), empty (inner part empty), twolayer (thick layer switch), id."""

    def __init__(self, nr=0):
        self.radsep = 0.
        self.id = 0
        self.edge, self.edgewidth = None, 0.
        self.edgeapp, self.edgerec = None, None
        self.empty = None
        self.twolayer = None
        for k in RING_KEYS:
            setattr(self, k, np.zeros(nr))
        self.vdispthick = np.full(nr, -1.)

    @property
    def nr(self):
        return len(self.radii)

    def setRings(self, **kw):
        """Set all ring arrays; scalars are broadcast to the length of radii (bbarolo/rings.py:71-85)."""
        radii = np.atleast_1d(np.asarray(kw["radii"], dtype=float))
        n = len(radii)
        self.radii = radii.copy()
        for k in RING_KEYS[1:]:
            v = np.asarray(kw.get(k, _DEFAULTS[k]), dtype=float)
            setattr(self, k, np.full(n, float(v)) if v.ndim == 0 else v.astype(float).copy())
        self.radsep = kw["radsep"] if "radsep" in kw else (radii[1] - radii[0] if n > 1 else 0)
        return self

    def _copy_attrs(self, r):
        r.radsep, r.id = self.radsep, self.id
        r.edge, r.edgewidth, r.empty = self.edge, self.edgewidth, self.empty
        r.edgeapp, r.edgerec = self.edgeapp, self.edgerec
        r.twolayer = self.twolayer
        return r

    def copy(self):
        r = self._copy_attrs(Rings())
        for k in RING_KEYS:
            setattr(r, k, np.array(getattr(self, k), dtype=float, copy=True))
        return r

    def subset(self, idx):
        r = self._copy_attrs(Rings())
        for k in RING_KEYS:
            setattr(r, k, np.atleast_1d(np.asarray(getattr(self, k))[idx]).astype(float))
        return r

    def addRing(self, **vals):
        for k in RING_KEYS:
            setattr(self, k, np.append(getattr(self, k), float(vals.get(k, _DEFAULTS.get(k, 0.)))))

    def thick(self):
        """A thick layer: twolayer, or if None, fthick > 0 in some ring (bbarolo/rings.py:123-126)."""
        if self.twolayer is not None:
            return bool(self.twolayer)
        return bool(np.any(np.asarray(self.fthick, float) > 0))

    def thick_z0(self):
        """Scale height of the thick layer: z0thick, or z0 where z0thick < 0."""
        z = np.asarray(self.z0thick, float)
        return np.where(z >= 0, z, np.asarray(self.z0, float))

    def layer_height(self):
        """Scale height of the thicker layer of each ring (bbarolo/rings.py:134-144)."""
        z0 = np.asarray(self.z0, float)
        if not self.thick():
            return z0
        return np.where(np.asarray(self.fthick, float) > 0, np.maximum(z0, self.thick_z0()), z0)

    def as_dict(self):
        return {k: np.array(getattr(self, k)) for k in RING_KEYS}
