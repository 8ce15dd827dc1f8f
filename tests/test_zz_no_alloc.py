"""pma_geodesic_noalloc (the allocation-free geodesic) against pma_geodesic through the C ABI.

Two calls in a row on the same GPU, the second on a smaller volume (the static scratch keeps the
first call's contents), then a volume too large for the scratch, which must be rejected before
any device work.  The test allocates its device buffers with hipMalloc from the HIP runtime the
library is linked against; the entry point itself allocates nothing.
"""
import ctypes

import numpy as np
import pytest

import porousmediaanalysis_amd as pma
from tests import volumes
from tests.gpu_common import log

pytestmark = pytest.mark.gpu


class Cavity(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32),
                ("distance_to_center", ctypes.c_double), ("void_count", ctypes.c_uint64),
                ("reached_walls", ctypes.c_int32), ("reaches_all", ctypes.c_int32)]


def cavity_text(rows):
    """cavityDescriptions.txt from the rows, in the reference's format (Serializer.cpp:205-221)."""
    s = "CavityId SeedX SeedY SeedZ DistanceToCenter VoidPixelsCount ReachedWallsCount IsReachingAllWalls\n"
    for r in rows:
        s += "%u %d %d %d %f %u %d %d\n" % (r.id, r.x, r.y, r.z, r.distance_to_center, r.void_count,
                                            r.reached_walls, r.reaches_all)
    return s


def hip():
    """The HIP runtime libpma.so uses (already loaded with it)."""
    pma.lib()
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class DeviceArray:
    """A hipMalloc'ed copy of a numpy array (or room for one)."""

    def __init__(self, shape, dtype, src=None):
        self.h, self.shape, self.dtype = hip(), shape, np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        if src is not None:
            a = np.ascontiguousarray(src, dtype=self.dtype)
            assert self.h.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.h.hipFree(self.ptr)


def noalloc(solid, d3q7, cap=64):
    L = pma.lib()
    f = L.pma_geodesic_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, pma.Dims, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.POINTER(Cavity), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
    Z, X, Y = solid.shape
    d_solid = DeviceArray(solid.shape, np.uint8, solid)
    d_dist = DeviceArray(solid.shape, np.float32)
    d_cav = DeviceArray(solid.shape, np.uint32)
    rows = (Cavity * cap)()
    count = ctypes.c_size_t()
    try:
        st = f(0, d_solid.ptr, pma.Dims(X, Y, Z), 1 if d3q7 else 0, d_dist.ptr, d_cav.ptr, rows, cap,
               ctypes.byref(count), None)
        if st != 0:
            raise pma.PmaError(st, L.pma_last_error().decode())
        return d_dist.get(), d_cav.get(), cavity_text(rows[:count.value])
    finally:
        for d in (d_solid, d_dist, d_cav):
            d.free()


def with_context(solid, d3q7):
    with pma.Context(0) as ctx:
        ctx.set_stencil(d3q7)
        ctx.load(pma.VOL_SOLID, solid)
        ctx.geodesic()
        return ctx.read(pma.VOL_SPD), ctx.read(pma.VOL_CAVITY), ctx.text(pma.TEXT_CAVITIES)


@pytest.mark.parametrize("d3q7", [False, True])
def test_noalloc_matches_geodesic(d3q7):
    big = volumes.gaussian((72, 80, 64), 3.0, 0.6, 5)       # several cavities
    small = volumes.spheres((24, 30, 28), 25, 2, 5, 8)
    for name, solid in [("big", big), ("small", small)]:
        log(f"noalloc geodesic {name} {solid.shape} d3q7={d3q7}")
        d0, c0, t0 = with_context(solid, d3q7)
        d1, c1, t1 = noalloc(solid, d3q7)
        assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), f"{name}: distances differ"
        assert np.array_equal(c0, c1), f"{name}: cavity ids differ"
        assert t0 == t1, f"{name}: cavity rows differ:\n{t0}\n{t1}"


def test_noalloc_rejects_too_large():
    log("noalloc geodesic: a volume beyond the scratch is rejected on the host")
    L = pma.lib()
    f = L.pma_geodesic_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, pma.Dims, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.POINTER(Cavity), ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
    rows = (Cavity * 1)()
    count = ctypes.c_size_t()
    # 1600^3 voxels = 8e6 blocks: needs ~160 MB of scratch; rejected before any pointer is used
    st = f(0, ctypes.c_void_p(16), pma.Dims(1600, 1600, 1600), 0, ctypes.c_void_p(16), ctypes.c_void_p(16), rows, 1,
           ctypes.byref(count), None)
    assert st == 5, L.pma_last_error().decode()
    assert b"scratch" in L.pma_last_error()
