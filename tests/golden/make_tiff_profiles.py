"""Extract central profiles of the reference's PSF fixture volumes.

Reads tests/data/psf_NA_0.8_dx_0.1_n_{1.0,1.5}.tif of the reference checkout
(65^3 float32, max-normalised, peak at the centre voxel) and stores their
central x-, y- and z-lines and the central xz plane in
tests/golden/psf_tiff_profiles.npz.  Run in a checkout that has the
reference at /root/reference:  python tests/golden/make_tiff_profiles.py
"""
import os

import numpy as np
from PIL import Image

SRC = "/root/reference/tests/data/psf_NA_0.8_dx_0.1_n_%s.tif"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "psf_tiff_profiles.npz")


def read_stack(path):
    im = Image.open(path)
    frames = []
    i = 0
    while True:
        try:
            im.seek(i)
        except EOFError:
            break
        frames.append(np.array(im, dtype=np.float32))
        i += 1
    return np.stack(frames)


def main():
    out = {}
    for n in ("1.0", "1.5"):
        a = read_stack(SRC % n)
        c = a.shape[0] // 2
        key = n.replace(".", "p")
        out["x_" + key] = a[c, c, :]
        out["y_" + key] = a[c, :, c]
        out["z_" + key] = a[:, c, c]
        out["xz_" + key] = a[:, c, :]
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
