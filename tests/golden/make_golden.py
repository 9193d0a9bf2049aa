#!/usr/bin/env python3
"""Write the golden map fixtures from the reference's own data.

Reads data/converted-maps/C1W.png (colour) and D1.png (height) and
data/custom-maps/flat_height.png from the reference checkout, packs them
the way the reference host loads maps (c/freeimage_futspace.h:47-53:
0xFF000000 | R << 16 | G << 8 | B, rows top to bottom; the height map goes
through the same packing and update_map keeps its low byte,
fut/interactive.fut:189) and stores a 256 x 256 crop of each as
tests/golden/maps.npz.  Run in a checkout that has the reference at
REFERENCE (default /root/reference); the fixture itself is committed, so
nothing at test time reads the reference.
"""
import os
import sys

import numpy as np
from PIL import Image

REFERENCE = os.environ.get("REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
CROP = (slice(384, 640), slice(512, 768))


def packed(path):
    rgb = np.asarray(Image.open(path).convert("RGB"), dtype=np.uint32)
    return (np.uint32(0xFF000000) | (rgb[..., 0] << 16) | (rgb[..., 1] << 8) | rgb[..., 2])


def main():
    d = os.path.join(REFERENCE, "data")
    colour = packed(os.path.join(d, "converted-maps", "C1W.png"))[CROP]
    height = packed(os.path.join(d, "converted-maps", "D1.png"))[CROP].view(np.int32)
    flat = packed(os.path.join(d, "custom-maps", "flat_height.png"))[CROP].view(np.int32)
    out = os.path.join(HERE, "maps.npz")
    np.savez_compressed(out, c1w_colour=colour, d1_height=height, flat_height=flat)
    print(f"wrote {out}: {colour.shape}, heights {np.unique(height & 255).size} levels, "
          f"flat levels {np.unique(flat & 255)}")


if __name__ == "__main__":
    sys.exit(main())
