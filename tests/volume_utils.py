"""Small numpy helpers shared by the tests (no GPU, no oracle)."""
import hashlib

import numpy as np


def sha256_array(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def canonical_relabel(labels):
    """Relabels a label volume by order of first occurrence in memory order (0 stays 0).

    Two label volumes describe the same partition exactly when their canonical
    relabellings are equal.
    """
    flat = np.ascontiguousarray(labels).ravel()
    uniq, first, inv = np.unique(flat, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(1, len(order) + 1)
    out = rank[inv].astype(np.uint32)
    out[flat == 0] = 0
    # keep 0 -> 0 and renumber the rest densely in first-occurrence order
    nz = out[flat != 0]
    if nz.size:
        u, f2, i2 = np.unique(nz, return_index=True, return_inverse=True)
        o2 = np.argsort(f2, kind="stable")
        r2 = np.empty_like(o2)
        r2[o2] = np.arange(1, len(o2) + 1)
        out[flat != 0] = r2[i2].astype(np.uint32)
    return out.reshape(labels.shape)
