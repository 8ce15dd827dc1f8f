"""Writes tests/golden/example/ from the reference's only golden data, Docs/Example.

Run in a container where the reference is mounted at /root/reference:
    python tests/golden/make_example_fixture.py

The example is 1196 x 960 x 5 voxels (X = TIFF rows, Y = TIFF columns, Z = files).
What is stored:
  * solid.npz       -- packed solid mask in [z][x][y] order.  The input TIFFs are 8-bit
                       MinIsWhite; a voxel is solid when its raw byte is non-zero.  PIL
                       inverts MinIsWhite on load, so raw = 255 - PIL value.
  * mask.npz        -- the fixture's IsMaximumBall volume (u8), used to feed the fixture's
                       own maximum balls into pore propagation.
  * digests.json    -- SHA-256 of the fixture volumes that are bit-exact
                       (EuclideanDistanceSquares, ContainingBallRadiiSquares,
                       ShortestPathDistances, CavityIds, PoreThroatIds,
                       PoreThroatIdsWatershed) and of the first-occurrence relabelling of
                       PoreThroatIds / PoreThroatIdsWatershed (partitions).
  * *.txt           -- the fixture's text outputs, verbatim.
  * Images/, config.txt -- the example's own input folder (five 8-bit LZW TIFF slices) and
                       config, verbatim: the drop-in folder test runs on them.
The 32-bit fixture volumes are CMYK TIFFs whose four byte samples are the little-endian
bytes of the value.
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np
from PIL import Image

REF = "/root/reference/Docs/Example"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example")
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from tests.volume_utils import canonical_relabel, sha256_array  # noqa: E402


def stack(folder, kind):
    files = sorted(f for f in os.listdir(folder) if f.endswith((".tif", ".tiff")))
    files = sorted(files, key=lambda f: os.path.splitext(f)[0])
    out = []
    for f in files:
        a = np.array(Image.open(os.path.join(folder, f)))
        if kind == "u32":
            a = np.ascontiguousarray(a).view(np.uint32)[..., 0]
        elif kind == "f32":
            a = np.ascontiguousarray(a).view(np.float32)[..., 0]
        elif kind == "raw8":
            a = (255 - a).astype(np.uint8)  # undo PIL's MinIsWhite inversion
        out.append(a)
    return np.stack(out, axis=0)  # [z][x][y]


def main():
    os.makedirs(OUT, exist_ok=True)
    raw = stack(os.path.join(REF, "Images"), "raw8")
    solid = raw != 0
    Z, X, Y = solid.shape
    np.savez_compressed(os.path.join(OUT, "solid.npz"), bits=np.packbits(solid.ravel()),
                        shape=np.array([X, Y, Z]))
    a = os.path.join(REF, "Expected results from -all")
    s = os.path.join(REF, "Expected results from -spd")
    mask = stack(os.path.join(a, "IsMaximumBall"), "raw8")  # same MinIsWhite tags as the input
    np.savez_compressed(os.path.join(OUT, "mask.npz"), mask=mask)
    dig = {
        "shape_xyz": [X, Y, Z],
        "edt2": sha256_array(stack(os.path.join(a, "EuclideanDistanceSquares"), "u32")),
        "cr2": sha256_array(stack(os.path.join(a, "ContainingBallRadiiSquares"), "u32")),
        "spd": sha256_array(stack(os.path.join(s, "ShortestPathDistances"), "f32")),
        "cavity": sha256_array(stack(os.path.join(s, "CavityIds"), "u32")),
        "pt": sha256_array(stack(os.path.join(a, "PoreThroatIds"), "u32")),
        "ptw": sha256_array(stack(os.path.join(a, "PoreThroatIdsWatershed"), "u32")),
        "pt_partition": sha256_array(canonical_relabel(stack(os.path.join(a, "PoreThroatIds"), "u32"))),
        "ptw_partition": sha256_array(canonical_relabel(stack(os.path.join(a, "PoreThroatIdsWatershed"), "u32"))),
        "mask_values": sorted(int(v) for v in np.unique(mask)),
    }
    with open(os.path.join(OUT, "digests.json"), "w") as f:
        json.dump(dig, f, indent=1)
    for t in ["intermediateStatistics.txt", "euclideanDistanceSquareCounts.txt", "pores.txt",
              "sharedPixelTypes.txt", "poresWatershed.txt", "sharedPixelTypesWatershed.txt"]:
        shutil.copyfile(os.path.join(a, t), os.path.join(OUT, t))
    shutil.copyfile(os.path.join(s, "cavityDescriptions.txt"), os.path.join(OUT, "cavityDescriptions.txt"))
    os.makedirs(os.path.join(OUT, "Images"), exist_ok=True)
    for f in os.listdir(os.path.join(REF, "Images")):
        shutil.copyfile(os.path.join(REF, "Images", f), os.path.join(OUT, "Images", f))
    shutil.copyfile(os.path.join(REF, "config.txt"), os.path.join(OUT, "config.txt"))
    print("wrote", OUT, dig)


if __name__ == "__main__":
    main()
