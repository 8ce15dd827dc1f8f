"""Pins the oracle's EDT and containing balls against brute force and scipy on tiny volumes.

A plausible mistake (a dropped boundary term, a wrong start value, a transposed axis, the
wrong tie-break) changes at least one of these volumes.
"""
import numpy as np
import pytest
from scipy.ndimage import distance_transform_edt

import oracle
from tests import volumes


def brute_edt(solid):
    """min over solids and the virtual planes x = -X and x = 2X-1 of the squared distance."""
    Z, X, Y = solid.shape
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    best = np.minimum((xx + X) ** 2, (2 * X - 1 - xx) ** 2).astype(np.int64)
    for z, x, y in zip(*np.nonzero(solid)):
        best = np.minimum(best, (xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2)
    best[solid != 0] = 0
    return best.astype(np.uint32)


def brute_balls(e):
    """Max r2 over covering balls; the first centre in x, y, z scan order among equals."""
    Z, X, Y = e.shape
    cr = np.zeros(e.shape, np.int64)
    best_scan = np.full(e.shape, np.iinfo(np.int64).max)
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    centres = sorted(zip(*np.nonzero(e)), key=lambda t: (t[1], t[2], t[0]))  # x, y, z scan order
    for z, x, y in centres:
        r2 = int(e[z, x, y])
        cov = ((xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2 <= r2) & (e > 0)
        scan = (x * Y + y) * Z + z
        upd = cov & ((r2 > cr) | ((r2 == cr) & (scan < best_scan)))
        cr[upd] = r2
        best_scan[upd] = scan
    return cr.astype(np.uint32), best_scan


CASES = dict(volumes.edge_cases())
CASES["spheres"] = volumes.spheres((10, 14, 12), 6, 1.5, 3.5, 5)
CASES["noise"] = volumes.random_noise((9, 11, 10), 0.2, 9)


@pytest.mark.parametrize("name", sorted(CASES))
def test_edt_brute_force(name):
    solid = CASES[name]
    e, m, h = oracle.edt(solid)
    assert np.array_equal(e, brute_edt(solid))
    assert m == (int(e.max()) if e.size else 0)


def test_edt_matches_scipy_with_virtual_planes():
    solid = volumes.spheres((12, 16, 14), 8, 1.5, 4.0, 11)
    Z, X, Y = solid.shape
    # pad x with X voids on each side, then a solid plane: planes at x = -X - 1 + 1 ... place
    # the planes exactly at x = -X and x = 2X - 1.
    pad = np.zeros((Z, 3 * X + 1, Y), bool)
    pad[:, X:2 * X, :] = solid != 0
    pad[:, 0, :] = True          # x = -X
    pad[:, 3 * X, :] = True      # x = 2X
    pad[:, 3 * X - 1, :] = True  # x = 2X - 1
    d = distance_transform_edt(~pad)
    ref = np.rint(d[:, X:2 * X, :] ** 2).astype(np.uint32)
    e, _, _ = oracle.edt(solid)
    assert np.array_equal(e, ref)


def test_edt_histogram_counts_only_inner_balls():
    solid = volumes.spheres((14, 15, 16), 10, 1.5, 3.5, 3)
    e, m, h = oracle.edt(solid)
    Z, X, Y = e.shape
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    r = np.sqrt(e.astype(np.float64))
    inner = (e > 0) & ~((r > xx) | (r > yy) | (r > zz) | (xx + r > X - 1) | (yy + r > Y - 1) | (zz + r > Z - 1))
    ref = np.bincount(e[inner].astype(np.int64), minlength=m + 1)
    assert np.array_equal(h, ref)


@pytest.mark.parametrize("name", sorted(CASES))
def test_containing_balls_brute_force(name):
    solid = CASES[name]
    e, _, _ = oracle.edt(solid)
    cr, cx, cy, cz, mask, nmax = oracle.containing_balls(e)
    ref_cr, ref_scan = brute_balls(e)
    assert np.array_equal(cr, ref_cr)
    Z, X, Y = e.shape
    void = e > 0
    scan = (cx.astype(np.int64) * Y + cy) * Z + cz
    assert np.array_equal(scan[void], ref_scan[void])
    centres = set(zip(cz[void], cx[void], cy[void]))
    ref_mask = np.where(void, 127, 0).astype(np.uint8)
    for z, x, y in centres:
        ref_mask[z, x, y] = 255
    assert np.array_equal(mask, ref_mask)
    assert nmax == len(centres)


def test_containing_balls_threads_agree():
    solid = volumes.gaussian((20, 22, 24), 2.0, 0.6, 4)
    e, _, _ = oracle.edt(solid)
    one = oracle.containing_balls(e, threads=1)
    many = oracle.containing_balls(e, threads=4)
    for a, b in zip(one[:5], many[:5]):
        assert np.array_equal(a, b)
