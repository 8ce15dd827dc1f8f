"""MI355X-native pore-throat + geodesic pipeline (Python view of the C ABI in include/pma/pma.h).

The product is libpma.so: hand-written HIP kernels for gfx950 behind an extern "C" API.
This module only loads that library with ctypes and moves numpy arrays in and out; it has
no compute path of its own, and importing it fails loudly when the library is missing.

Volumes are numpy arrays of shape (Z, X, Y), C order: the reference's [z][x][y] layout.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpma.so")

VOL_SOLID, VOL_EDT2, VOL_CR2, VOL_CX, VOL_CY, VOL_CZ, VOL_MASK, VOL_PT, VOL_PTW, VOL_SPD, VOL_CAVITY = range(11)
TEXT_STATS, TEXT_EDT_COUNTS, TEXT_PORES, TEXT_THROATS, TEXT_PORES_WS, TEXT_THROATS_WS, TEXT_CAVITIES = range(7)
STAGE_EDT, STAGE_BALLS, STAGE_PORES, STAGE_REMOVE, STAGE_WATERSHED, STAGE_GEODESIC = range(6)
STAGE_NAMES = ["edt", "containing_balls", "pores", "remove_empty_pores", "watershed", "geodesic"]
MODE_ALL, MODE_EDT, MODE_RIB, MODE_APT, MODE_SPD = range(5)
_DTYPES = {VOL_SOLID: np.uint8, VOL_MASK: np.uint8, VOL_SPD: np.float32}

STATUS = {0: "PMA_OK", 1: "PMA_ERR_INVALID_ARGUMENT", 2: "PMA_ERR_HIP", 3: "PMA_ERR_OUT_OF_MEMORY",
          4: "PMA_ERR_STATE", 5: "PMA_ERR_CAPACITY", 6: "PMA_ERR_INTERNAL"}

# Every symbol the header declares (checked by the CPU test of the library).
EXPORTS = [
    "pma_create", "pma_destroy", "pma_last_error", "pma_set_stencil", "pma_set_dims", "pma_get_dims",
    "pma_load_volume", "pma_read_volume", "pma_volume_device_ptr", "pma_edt", "pma_containing_balls",
    "pma_pores", "pma_remove_empty_pores", "pma_watershed", "pma_geodesic", "pma_run_all",
    "pma_get_statistics", "pma_get_histogram", "pma_get_text", "pma_set_text", "pma_get_stage_info",
    "pma_generate_spheres", "pma_generate_gaussian", "pma_set_kernel_timing", "pma_get_kernel_time",
    "pma_list_kernel_times", "pma_geodesic_noalloc", "pma_run_folder", "pma_read_tiff_stack",
    "pma_write_tiff_stack", "pma_set_tuning", "pma_get_tuning", "pma_run_batch", "pma_tune_batch",
    "pma_surrogate_create", "pma_surrogate_destroy", "pma_surrogate_param_count", "pma_surrogate_params",
    "pma_surrogate_set_params", "pma_surrogate_train_step", "pma_surrogate_predict",
]
TUNE_GEODESIC_BUCKET, TUNE_FLOOD_GRID, TUNE_REPLAY_MEMORY_MB = range(3)


class PmaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Dims(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32)]


class StageInfo(ctypes.Structure):
    _fields_ = [("device_ms", ctypes.c_double), ("host_ms", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("work_items", ctypes.c_uint64)]


class BatchVolume(ctypes.Structure):
    _fields_ = [("dims", Dims), ("solid", ctypes.c_void_p)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_uint32), ("geodesic_bucket", ctypes.c_double), ("flood_grid", ctypes.c_uint32),
                ("replay_memory_mb", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BATCH_DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int)


def build(jobs=8, force=False):
    """Compiles libpma.so for gfx950 with hipcc and the pma program (Makefile in csrc/); force
    recompiles every source."""
    subprocess.run(["make", "-s", f"-j{jobs}"] + (["-B"] if force else []) + ["-C", os.path.join(_HERE, "csrc")],
                   check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc) first")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.pma_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.pma_destroy.argtypes = [vp]
        L.pma_destroy.restype = None
        L.pma_last_error.restype = ctypes.c_char_p
        L.pma_set_stencil.argtypes = [vp, ctypes.c_int]
        L.pma_set_dims.argtypes = [vp, Dims]
        L.pma_get_dims.argtypes = [vp, ctypes.POINTER(Dims)]
        L.pma_load_volume.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        L.pma_read_volume.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, ctypes.c_int]
        L.pma_volume_device_ptr.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        for f in ["pma_edt", "pma_containing_balls", "pma_pores", "pma_remove_empty_pores", "pma_watershed",
                  "pma_geodesic", "pma_run_all"]:
            getattr(L, f).argtypes = [vp]
        L.pma_get_statistics.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]
        L.pma_get_histogram.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.pma_get_text.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.pma_set_text.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
        L.pma_get_stage_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(StageInfo)]
        L.pma_generate_spheres.argtypes = [vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64]
        L.pma_generate_gaussian.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint64]
        L.pma_run_folder.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.pma_read_tiff_stack.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.POINTER(Dims),
                                          ctypes.POINTER(ctypes.c_int)]
        L.pma_write_tiff_stack.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp, Dims, ctypes.c_int]
        L.pma_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.pma_get_tuning.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.pma_run_batch.argtypes = [ctypes.c_int, ctypes.POINTER(BatchVolume), ctypes.c_size_t, ctypes.c_int,
                                    ctypes.POINTER(BatchConfig), BATCH_DONE, vp]
        L.pma_tune_batch.argtypes = [ctypes.c_int, ctypes.POINTER(BatchVolume), ctypes.c_size_t, ctypes.c_int,
                                     ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_double)]
        L.pma_surrogate_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_uint64]
        L.pma_surrogate_destroy.argtypes = [vp]
        L.pma_surrogate_destroy.restype = None
        L.pma_surrogate_param_count.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.pma_surrogate_params.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
        L.pma_surrogate_set_params.argtypes = [vp, vp, ctypes.c_size_t]
        L.pma_surrogate_train_step.argtypes = [vp, vp, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.pma_surrogate_predict.argtypes = [vp, vp, vp, vp, ctypes.c_int]
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise PmaError(status, lib().pma_last_error().decode())


def read_tiff_stack(folder):
    """A folder of 8-bit or 32-bit TIFF slices (the reference's form) as a (Z, X, Y) array: uint8, or
    uint32 for 32-bit slices (view a float volume with .view(np.float32))."""
    d, bpv = Dims(), ctypes.c_int()
    _check(lib().pma_read_tiff_stack(folder.encode(), None, 0, ctypes.byref(d), ctypes.byref(bpv)))
    out = np.empty((d.z, d.x, d.y), np.uint8 if bpv.value == 1 else np.uint32)
    _check(lib().pma_read_tiff_stack(folder.encode(), out.ctypes.data, out.nbytes, ctypes.byref(d), ctypes.byref(bpv)))
    return out


def write_tiff_stack(folder, like, array):
    """Writes a (Z, X, Y) uint8 or 4-byte volume as TIFF slices named and tagged like folder `like`."""
    a = np.ascontiguousarray(array)
    if a.dtype.itemsize not in (1, 4):
        raise ValueError("1 or 4 bytes per voxel")
    Z, X, Y = a.shape
    _check(lib().pma_write_tiff_stack(folder.encode(), like.encode(), a.ctypes.data, Dims(X, Y, Z), a.dtype.itemsize))


def _batch_volumes(volumes):
    arrs = [np.ascontiguousarray(v, dtype=np.uint8) for v in volumes]
    items = (BatchVolume * max(len(arrs), 1))()
    for i, a in enumerate(arrs):
        Z, X, Y = a.shape
        items[i] = BatchVolume(Dims(X, Y, Z), a.ctypes.data)
    return arrs, items


def run_batch(volumes, d3q7=False, config=None, on_done=None, device=0):
    """The whole pipeline on every (Z, X, Y) solid volume, on a pool of worker contexts.
    on_done(index, ctx) runs on a worker thread for every finished subvolume, with a Context view of
    the worker's context (read volumes and texts there); config: a BatchConfig or None (defaults)."""
    arrs, items = _batch_volumes(volumes)
    errors = []

    def done(_user, index, h, status):
        try:
            if status == 0 and on_done is not None:
                on_done(index, _Borrowed(h, arrs[index].shape))
        except Exception as e:  # an exception must not cross into the worker thread
            errors.append((index, e))

    cb = BATCH_DONE(done) if on_done is not None else BATCH_DONE()
    cfg = ctypes.byref(config) if config is not None else None
    _check(lib().pma_run_batch(device, items, len(arrs), int(d3q7), cfg, cb, None))
    if errors:
        raise errors[0][1]


def tune_batch(sample, d3q7=False, device=0):
    """The autotuner on a sample of subvolumes: (BatchConfig, voxels per second)."""
    arrs, items = _batch_volumes(sample)
    best, vps = BatchConfig(), ctypes.c_double()
    _check(lib().pma_tune_batch(device, items, len(arrs), int(d3q7), ctypes.byref(best), ctypes.byref(vps)))
    return best, vps.value


class Context:
    """One device context (one HIP stream, one resident volume set)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().pma_create(ctypes.byref(h), device))
        self._h = h
        self.shape = None

    def close(self):
        if self._h:
            lib().pma_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration and volumes --------------------------------------------------------
    def set_tuning(self, which, value):
        _check(lib().pma_set_tuning(self._h, which, float(value)))

    def tuning(self, which):
        v = ctypes.c_double()
        _check(lib().pma_get_tuning(self._h, which, ctypes.byref(v)))
        return v.value

    def set_stencil(self, d3q7):
        _check(lib().pma_set_stencil(self._h, 1 if d3q7 else 0))

    def set_shape(self, shape):
        """shape = (Z, X, Y)."""
        Z, X, Y = (int(v) for v in shape)
        _check(lib().pma_set_dims(self._h, Dims(X, Y, Z)))
        self.shape = (Z, X, Y)

    def load(self, which, array):
        dt = _DTYPES.get(which, np.uint32)
        a = np.ascontiguousarray(array, dtype=dt)
        if self.shape != a.shape:
            self.set_shape(a.shape)
        _check(lib().pma_load_volume(self._h, which, a.ctypes.data_as(ctypes.c_void_p), 0))

    def read(self, which):
        dt = _DTYPES.get(which, np.uint32)
        out = np.empty(self.shape, dt)
        _check(lib().pma_read_volume(self._h, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes, 0))
        return out

    def device_ptr(self, which):
        p = ctypes.c_void_p()
        _check(lib().pma_volume_device_ptr(self._h, which, ctypes.byref(p)))
        return p.value

    # -- stages ---------------------------------------------------------------------------
    def edt(self):
        _check(lib().pma_edt(self._h))

    def containing_balls(self):
        _check(lib().pma_containing_balls(self._h))

    def pores(self):
        _check(lib().pma_pores(self._h))

    def remove_empty_pores(self):
        _check(lib().pma_remove_empty_pores(self._h))

    def watershed(self):
        _check(lib().pma_watershed(self._h))

    def geodesic(self):
        _check(lib().pma_geodesic(self._h))

    def run_folder(self, folder, mode):
        """One console pass (MODE_ALL, MODE_EDT, MODE_RIB, MODE_APT, MODE_SPD) over a reference folder."""
        _check(lib().pma_run_folder(self._h, folder.encode(), mode))
        d = Dims()
        _check(lib().pma_get_dims(self._h, ctypes.byref(d)))
        self.shape = (d.z, d.x, d.y)

    def run_all(self):
        _check(lib().pma_run_all(self._h))

    # -- statistics and text ------------------------------------------------------------
    def statistics(self):
        m, n = ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().pma_get_statistics(self._h, ctypes.byref(m), ctypes.byref(n)))
        return int(m.value), int(n.value)

    def histogram(self):
        n = ctypes.c_size_t()
        st = lib().pma_get_histogram(self._h, None, 0, ctypes.byref(n))
        if st not in (0, 5):
            _check(st)
        out = np.zeros(n.value, np.uint64)
        _check(lib().pma_get_histogram(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(n)))
        return out

    def text(self, which):
        n = ctypes.c_size_t()
        st = lib().pma_get_text(self._h, which, None, 0, ctypes.byref(n))
        if st not in (0, 5):
            _check(st)
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().pma_get_text(self._h, which, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def set_text(self, which, text):
        _check(lib().pma_set_text(self._h, which, text.encode()))

    def stage_info(self, stage):
        s = StageInfo()
        _check(lib().pma_get_stage_info(self._h, stage, ctypes.byref(s)))
        return dict(device_ms=s.device_ms, host_ms=s.host_ms, launches=s.launches, work_items=s.work_items)

    # -- synthetic inputs ---------------------------------------------------------------
    def generate_spheres(self, shape, n, rmin, rmax, seed):
        self.set_shape(shape)
        _check(lib().pma_generate_spheres(self._h, n, rmin, rmax, seed))

    def generate_gaussian(self, shape, sigma, porosity, seed):
        self.set_shape(shape)
        _check(lib().pma_generate_gaussian(self._h, sigma, porosity, seed))


class _Borrowed(Context):
    """A view of a context owned by the library (a batch worker's): never destroyed here."""

    def __init__(self, h, shape):
        self._h = ctypes.c_void_p(h)
        self.shape = tuple(shape)

    def close(self):
        self._h = None


class Surrogate:
    """The learned screening network (pma_surrogate_*): train on contexts that hold the exact
    fields, predict radius and distance ratio fields from a context's solid volume and EDT."""

    # (cin, cout, taps, cin_real, cout_real) of the layers, in the parameter vector's order
    LAYERS = [(16, 16, 27, 2, 16), (16, 16, 27, 16, 16), (16, 32, 27, 16, 32), (32, 32, 27, 32, 32),
              (32, 64, 27, 32, 64), (64, 64, 27, 64, 64), (96, 32, 27, 96, 32), (32, 32, 27, 32, 32),
              (48, 16, 27, 48, 16), (16, 16, 27, 16, 16), (16, 16, 1, 16, 2)]

    def __init__(self, device=0, seed=1):
        h = ctypes.c_void_p()
        _check(lib().pma_surrogate_create(ctypes.byref(h), device, seed))
        self._h = h
        n = ctypes.c_size_t()
        _check(lib().pma_surrogate_param_count(self._h, ctypes.byref(n)))
        self.nparam = n.value

    def close(self):
        if self._h:
            lib().pma_surrogate_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def params(self, grads=False):
        out = np.empty(self.nparam, np.float32)
        _check(lib().pma_surrogate_params(self._h, 1 if grads else 0, out.ctypes.data, out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().pma_surrogate_set_params(self._h, p.ctypes.data, p.size))

    def layer_views(self, flat):
        """[(weights [tap][out][in], bias [out])] views of a flat parameter vector."""
        out, off = [], 0
        for cin, cout, taps, _, _ in self.LAYERS:
            w = flat[off:off + taps * cout * cin].reshape(taps, cout, cin)
            off += taps * cout * cin
            out.append((w, flat[off:off + cout]))
            off += cout
        return out

    def train_step(self, ctx, lr=1e-3, update=True):
        loss = ctypes.c_float()
        _check(lib().pma_surrogate_train_step(self._h, ctx._h, lr, 1 if update else 0, ctypes.byref(loss)))
        return loss.value

    def predict(self, ctx):
        radius = np.empty(ctx.shape, np.float32)
        ratio = np.empty(ctx.shape, np.float32)
        _check(lib().pma_surrogate_predict(self._h, ctx._h, radius.ctypes.data, ratio.ctypes.data, 0))
        return radius, ratio
