"""The uneven multi-galaxy batch: 64 synthetic galaxies (seed 11) within the stated ranges, the same on every
call, each with a velocity range that the cube's channels hold and a disk that its pixels hold."""
import numpy as np

from bbarolo2_amd.synth import multi_galaxy_configs, MULTI_PIX


def test_ranges_and_determinism():
    a = multi_galaxy_configs()
    b = multi_galaxy_configs()
    assert len(a) == 64
    for (h, r, par, noise, seed, spec, desc), (h2, r2, *_) in zip(a, b):
        nx, ny, nz = h.DimAx(0), h.DimAx(1), h.DimAx(2)
        assert nx == ny and 48 <= nx <= 256 and 32 <= nz <= 128
        assert 1 <= r.nr <= 40 and 0.02 <= noise <= 0.10
        assert 30 <= r.inc[0] <= 80 and 0 <= r.phi[0] <= 360 and 5 <= r.vdisp[0] <= 30
        bpix = h.Bmaj() * 3600 / MULTI_PIX
        assert 3 <= bpix <= 6
        # the disk inside the cube, the line-of-sight velocities inside the channels
        assert r.radii[-1] / MULTI_PIX < nx / 2
        assert r.nr == 1 or np.min(np.diff(r.radii)) >= MULTI_PIX * (1 - 1e-9)     # (wider than the sub-rings)
        vext = r.vrot.max() * np.sin(np.radians(r.inc[0])) + 4 * r.vdisp[0]
        assert 2 * vext <= nz * abs(h.Cdelt(2)) * (1 + 1e-9)
        assert np.array_equal(r.vrot, r2.vrot) and h.DimAx(2) == h2.DimAx(2)
        assert par.GALMODTYPE == "ANALYTIC" and [k for k, _ in spec] == ["vrot", "vdisp", "inc", "phi"]
    sides = np.array([c[0].DimAx(0) for c in a])
    assert sides.min() < 80 and sides.max() > 220          # (uneven: both ends of the range are drawn)
    assert len({c[4] for c in a}) == 64                     # (a noise seed per galaxy)
