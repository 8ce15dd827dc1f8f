"""Loader for the committed Docs/Example fixture (tests/golden/example)."""
import json
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example")


def solid():
    z = np.load(os.path.join(DIR, "solid.npz"))
    X, Y, Z = (int(v) for v in z["shape"])
    return np.unpackbits(z["bits"])[: X * Y * Z].reshape(Z, X, Y).astype(np.uint8)


def mask():
    return np.load(os.path.join(DIR, "mask.npz"))["mask"]


def digests():
    with open(os.path.join(DIR, "digests.json")) as f:
        return json.load(f)


def text(name):
    """Fixture text with the Windows line endings normalised to LF."""
    with open(os.path.join(DIR, name), "rb") as f:
        return f.read().decode().replace("\r\n", "\n")


def ball_order(mask, e):
    """A processing order of the maximum balls of `mask` that the reference's could have been:
    r2 (the EDT at the centre) non-increasing, and within one r2 the fixture's pore seeds first, in
    pore-ID order, then the other balls in scan order. The fixture's pores.txt lists every pore
    (none was removed: IDs 127, 128, ... without gaps), and pore floods do not depend on each
    other, so every seed is still unlabelled at its turn and every other ball already labelled:
    the pores open at the fixture's seeds with the fixture's IDs, whatever MSVC's tie order was."""
    Z, X, Y = mask.shape
    rows = [r.split() for r in text("pores.txt").splitlines()[1:]]
    seeds = np.array([(int(r[3]) * X + int(r[1])) * Y + int(r[2]) for r in rows], np.int64)  # SeedZ, SeedX, SeedY
    balls = np.flatnonzero(mask.reshape(-1) == 255)
    rank = np.full(balls.size, seeds.size, np.int64)
    where = np.searchsorted(balls, seeds)
    assert np.all(balls[where] == seeds), "a fixture seed is not a maximum ball"
    rank[where] = np.arange(seeds.size)
    r2 = e.reshape(-1)[balls].astype(np.int64)
    return balls[np.lexsort((balls, rank, -r2))].astype(np.uint32)
