"""Minimal ctypes binding to the HIP runtime (libamdhip64) for device buffers.

Used by the Python front-end, tests and benchmarks to move arrays to and from
the device pointers that the C ABI takes.  Shares the runtime instance with
libhmglib.so (same process, same library).
"""
import ctypes
import ctypes.util
import os

import numpy as np

_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def _load():
    global _hip
    if _hip is None:
        cands = ["/opt/rocm/lib/libamdhip64.so", ctypes.util.find_library("amdhip64") or ""]
        for c in cands:
            if c and (os.path.exists(c) or not c.startswith("/")):
                try:
                    _hip = ctypes.CDLL(c)
                    break
                except OSError:
                    continue
        if _hip is None:
            raise RuntimeError("libamdhip64 not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                             ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return _hip


def _chk(err, what):
    if err != 0:
        raise RuntimeError(f"{what}: {_load().hipGetErrorString(err).decode()}")


def device_count():
    n = ctypes.c_int(0)
    try:
        err = _load().hipGetDeviceCount(ctypes.byref(n))
    except RuntimeError:
        return 0
    return n.value if err == 0 else 0


def synchronize():
    _chk(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


def trim_pool():
    """Return the unused memory the device's default stream-ordered pool keeps
    cached (the library frees into that pool) and the library's cached large
    buffers to the device."""
    import hmglib_amd
    hmglib_amd.lib().hmg_release_cached_memory()
    pool = ctypes.c_void_p()
    _chk(_load().hipDeviceGetDefaultMemPool(ctypes.byref(pool), 0), "hipDeviceGetDefaultMemPool")
    _chk(_load().hipMemPoolTrimTo(pool, ctypes.c_size_t(0)), "hipMemPoolTrimTo")


def mem_info():
    f = ctypes.c_size_t(0)
    t = ctypes.c_size_t(0)
    _chk(_load().hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)), "hipMemGetInfo")
    return f.value, t.value


class DeviceArray:
    """A device allocation holding a fixed dtype/shape (column-major for 2-D)."""

    def __init__(self, nbytes_or_array, dtype=np.float64, shape=None):
        h = _load()
        if isinstance(nbytes_or_array, np.ndarray):
            arr = nbytes_or_array
            self.dtype = arr.dtype
            self.shape = arr.shape
            nbytes = arr.nbytes
        else:
            self.dtype = np.dtype(dtype)
            self.shape = shape if shape is not None else (nbytes_or_array // self.dtype.itemsize,)
            nbytes = nbytes_or_array
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _chk(h.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if isinstance(nbytes_or_array, np.ndarray):
            self.upload(nbytes_or_array)

    @classmethod
    def empty(cls, shape, dtype=np.float64):
        shape = (shape,) if np.isscalar(shape) else tuple(shape)
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(n, dtype=dtype, shape=shape)

    def upload(self, arr):
        a = np.asfortranarray(arr, dtype=self.dtype) if np.ndim(arr) > 1 else \
            np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _chk(_load().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, hipMemcpyHostToDevice),
             "hipMemcpy H2D")
        _null_stream_sync()

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _chk(_load().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, hipMemcpyDeviceToHost),
             "hipMemcpy D2H")
        return out

    def zero(self):
        _chk(_load().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        _null_stream_sync()

    def free(self):
        if self.ptr:
            _load().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _null_stream_sync():
    # The engine runs on non-blocking streams, which are not ordered after
    # null-stream copies: finish the copy before the engine may read the data.
    _chk(_load().hipStreamSynchronize(None), "hipStreamSynchronize")


def memcpy_htod(dptr, arr):
    a = np.ascontiguousarray(arr)
    _chk(_load().hipMemcpy(dptr, a.ctypes.data, a.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")
    _null_stream_sync()


def memcpy_dtoh(arr, dptr):
    assert arr.flags.c_contiguous or arr.flags.f_contiguous
    _chk(_load().hipMemcpy(arr.ctypes.data, dptr, arr.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
    return arr
