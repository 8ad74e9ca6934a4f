"""Writes tests/golden/ngc2403_crop.npz from the reference's example cube examples/ngc2403.fits (the data of its
test_3dfit_quick, tests/test_bbarolo.py:306-317): the central 66 x 66 pixels (all 62 channels), which hold the boxes
of every ring of that test (rings 105-375", centre 77, 77), stored as int16 with a scale (peak / 32000: a step of 1/68
of the noise rms) and the header cards the model needs, with CRPIX moved by the crop.

    python tests/golden/make_ngc2403.py /root/reference/examples/ngc2403.fits
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

LO, SIZE = 44, 66


def main(path):
    from bbarolo2_amd.fitsio import read_fits, load_cube
    cards, _ = read_fits(path)
    h, d = load_cube(path)
    sub = d[:, LO:LO + SIZE, LO:LO + SIZE].astype(np.float64)
    scale = float(np.abs(sub).max()) / 32000.
    q = np.round(sub / scale).astype(np.int16)
    keep = {}
    for k, v in cards.items():
        if isinstance(v, (int, float, str, bool)):
            keep[k] = v
    keep["NAXIS1"], keep["NAXIS2"] = SIZE, SIZE
    keep["CRPIX1"] = float(cards["CRPIX1"]) - LO
    keep["CRPIX2"] = float(cards["CRPIX2"]) - LO
    np.savez_compressed(os.path.join(HERE, "ngc2403_crop.npz"), data=q, scale=np.float64(scale), offset=np.int64(LO),
                        cards=json.dumps(keep))
    print(q.shape, scale, os.path.getsize(os.path.join(HERE, "ngc2403_crop.npz")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples/ngc2403.fits")
