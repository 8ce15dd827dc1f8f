"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker or the timed CPU baseline.  Volumes are numpy arrays of
shape (Z, X, Y) in C order, i.e. the reference's [z][x][y] layout.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_cpp = ctypes.POINTER(ctypes.c_char_p)


def build():
    """Compiles the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_edt.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u32p, _u64p,
                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int]
        L.oracle_histogram_text.argtypes = [_u64p, ctypes.c_size_t]
        L.oracle_histogram_text.restype = ctypes.c_void_p
        L.oracle_containing_balls.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u32p,
                                              _u32p, _u32p, _u8p, _u64p, ctypes.c_int]
        L.oracle_pores.argtypes = [_u8p, _u32p, _u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _u32p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_pores_in_order.argtypes = [_u32p, _u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            _u32p, ctypes.c_size_t, _u32p, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_remove_empty_pores.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                                ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                                ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_watershed.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_char_p, _u32p, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_geodesic.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _u32p]
        L.oracle_geodesic.restype = ctypes.c_void_p
        L.oracle_sort_desc.argtypes = [_u64p, ctypes.c_size_t]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _take_string(p):
    s = ctypes.string_at(p).decode()
    lib().oracle_free(p)
    return s


def _dims(a):
    Z, X, Y = a.shape
    return X, Y, Z


def edt(solid, threads=1):
    """Squared EDT, max, histogram counts[0..maxE]."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    X, Y, Z = _dims(solid)
    e = np.empty(solid.shape, np.uint32)
    m = ctypes.c_uint32(0)
    n = ctypes.c_size_t(0)
    cap = 1 << 16
    while True:
        h = np.zeros(cap, np.uint64)
        rc = lib().oracle_edt(_ptr(solid, _u8p), X, Y, Z, _ptr(e, _u32p), ctypes.byref(m), _ptr(h, _u64p), cap,
                              ctypes.byref(n), threads)
        if rc == 0:
            return e, int(m.value), h[: n.value].copy()
        cap = n.value


def histogram_text(hist):
    hist = np.ascontiguousarray(hist, dtype=np.uint64)
    return _take_string(lib().oracle_histogram_text(_ptr(hist, _u64p), hist.size))


def containing_balls(e, threads=1):
    e = np.ascontiguousarray(e, dtype=np.uint32)
    X, Y, Z = _dims(e)
    cr, cx, cy, cz = (np.empty(e.shape, np.uint32) for _ in range(4))
    mask = np.empty(e.shape, np.uint8)
    n = ctypes.c_uint64(0)
    lib().oracle_containing_balls(_ptr(e, _u32p), X, Y, Z, _ptr(cr, _u32p), _ptr(cx, _u32p), _ptr(cy, _u32p),
                                  _ptr(cz, _u32p), _ptr(mask, _u8p), ctypes.byref(n), threads)
    return cr, cx, cy, cz, mask, int(n.value)


def pores(mask, e, cr, d3q7=False):
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    e = np.ascontiguousarray(e, dtype=np.uint32)
    cr = np.ascontiguousarray(cr, dtype=np.uint32)
    X, Y, Z = _dims(e)
    pt = np.empty(e.shape, np.uint32)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    lib().oracle_pores(_ptr(mask, _u8p), _ptr(e, _u32p), _ptr(cr, _u32p), X, Y, Z, int(d3q7), _ptr(pt, _u32p),
                       ctypes.byref(a), ctypes.byref(b))
    return pt, _take_string(a.value), _take_string(b.value)


def pores_in_order(e, cr, order, d3q7=False):
    """Pore propagation over the maximum balls in the given order (flat [z][x][y] voxel indices)
    instead of the std::sort permutation."""
    e = np.ascontiguousarray(e, dtype=np.uint32)
    cr = np.ascontiguousarray(cr, dtype=np.uint32)
    order = np.ascontiguousarray(order, dtype=np.uint32)
    X, Y, Z = _dims(e)
    pt = np.empty(e.shape, np.uint32)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    rc = lib().oracle_pores_in_order(_ptr(e, _u32p), _ptr(cr, _u32p), X, Y, Z, int(d3q7), _ptr(order, _u32p),
                                     order.size, _ptr(pt, _u32p), ctypes.byref(a), ctypes.byref(b))
    if rc != 0:
        raise ValueError("an order entry lies outside the volume")
    return pt, _take_string(a.value), _take_string(b.value)


def remove_empty_pores(pt, pores_text, keys_text):
    pt = np.array(pt, dtype=np.uint32, copy=True, order="C")
    X, Y, Z = _dims(pt)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    lib().oracle_remove_empty_pores(_ptr(pt, _u32p), X, Y, Z, pores_text.encode(), keys_text.encode(),
                                    ctypes.byref(a), ctypes.byref(b))
    return pt, _take_string(a.value), _take_string(b.value)


def watershed(cr, max_e, pores_text, d3q7=False):
    cr = np.ascontiguousarray(cr, dtype=np.uint32)
    X, Y, Z = _dims(cr)
    ptw = np.empty(cr.shape, np.uint32)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    rc = lib().oracle_watershed(_ptr(cr, _u32p), X, Y, Z, int(d3q7), max_e, pores_text.encode(), _ptr(ptw, _u32p),
                                ctypes.byref(a), ctypes.byref(b))
    if rc != 0:
        raise ValueError("a pore seed radius exceeds the maximum EDT square")
    return ptw, _take_string(a.value), _take_string(b.value)


def geodesic(solid, d3q7=False):
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    X, Y, Z = _dims(solid)
    dist = np.empty(solid.shape, np.float32)
    cav = np.empty(solid.shape, np.uint32)
    t = lib().oracle_geodesic(_ptr(solid, _u8p), X, Y, Z, int(d3q7), _ptr(dist, _f32p), _ptr(cav, _u32p))
    return dist, cav, _take_string(t)


def sort_desc(records):
    """libstdc++ std::sort of u64 records by their high 32 bits, descending."""
    a = np.array(records, dtype=np.uint64, copy=True)
    lib().oracle_sort_desc(_ptr(a, _u64p), a.size)
    return a


def pipeline(solid, d3q7=False, threads=1):
    """Runs every stage like the reference's -all followed by -spd; returns a dict."""
    e, max_e, hist = edt(solid, threads)
    cr, cx, cy, cz, mask, nmax = containing_balls(e, threads)
    pt, ptxt, ktxt = pores(mask, e, cr, d3q7)
    pt, ptxt, ktxt = remove_empty_pores(pt, ptxt, ktxt)
    ptw, wtxt, wktxt = watershed(cr, max_e, ptxt, d3q7)
    dist, cav, ctxt = geodesic(solid, d3q7)
    return dict(edt2=e, max_e=max_e, hist=hist, cr2=cr, cx=cx, cy=cy, cz=cz, mask=mask, n_max=nmax, pt=pt,
                pores=ptxt, throats=ktxt, ptw=ptw, pores_ws=wtxt, throats_ws=wktxt, dist=dist, cavity=cav,
                cavities=ctxt)
