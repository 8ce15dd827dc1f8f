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
