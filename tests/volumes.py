"""Seeded test volumes (numpy, CPU) in [z][x][y] layout; 1 = solid, 0 = void."""
import numpy as np


def spheres(shape, n, rmin, rmax, seed):
    Z, X, Y = shape
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n, 3)) * np.array([X, Y, Z])
    r = rng.uniform(rmin, rmax, n)
    z, x, y = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    s = np.zeros(shape, bool)
    for (cx, cy, cz), rr in zip(c, r):
        s |= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= rr * rr
    return s.astype(np.uint8)


def gaussian(shape, sigma, porosity, seed):
    from scipy.ndimage import gaussian_filter

    f = gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma, mode="reflect", truncate=4.0)
    q = np.quantile(f, porosity)
    return (f > q).astype(np.uint8)


def random_noise(shape, p_solid, seed):
    return (np.random.default_rng(seed).uniform(size=shape) < p_solid).astype(np.uint8)


def edge_cases():
    """Small hand-made volumes named by what they exercise."""
    out = {}
    out["all_void"] = np.zeros((5, 7, 9), np.uint8)
    out["all_solid"] = np.ones((4, 6, 5), np.uint8)
    v = np.zeros((7, 9, 8), np.uint8)
    v[3, 4, 5] = 1
    out["single_solid"] = v
    v = np.ones((6, 8, 7), np.uint8)
    v[2, 3, 4] = 0
    out["single_void"] = v
    out["one_slice"] = random_noise((1, 17, 23), 0.3, 1)
    out["one_row"] = random_noise((9, 1, 13), 0.25, 2)
    out["one_column"] = random_noise((11, 7, 1), 0.25, 3)
    v = np.zeros((8, 10, 12), np.uint8)
    v[:, :, 5] = 1
    out["wall_y"] = v
    v = np.zeros((9, 9, 9), np.uint8)
    v[0, :, :] = 1
    v[-1, :, :] = 1
    out["slab_z"] = v
    return out
