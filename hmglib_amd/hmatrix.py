"""Python front-end over the C ABI, used by tests and benchmarks.

It drives the engine exactly like the reference's drivers drive hmglib
(src/hmglib_test.cu:43-140): init_h_matrix_data, write coordinates into
data.coords_d, set the parameters, setup_h_matrix, attach a kernel object,
precompute_aca / precompute_dense, apply.  Every compute step is a call into
libhmglib.so; Python only moves host arrays to and from device buffers.
"""
import ctypes
import os

import numpy as np

from . import _abi
from .hip import DeviceArray, memcpy_dtoh, memcpy_htod


def rng_uniform(seed, n):
    out = np.empty(n, dtype=np.float64)
    _abi.lib().hmg_rng_uniform(seed, n, out.ctypes.data_as(_abi.c_double_p))
    return out


def rng_normal(seed, n):
    out = np.empty(n, dtype=np.float64)
    _abi.lib().hmg_rng_normal(seed, n, out.ctypes.data_as(_abi.c_double_p))
    return out


class Kernel:
    """A kernel object in device memory (struct system_assembler)."""

    def __init__(self, kind, regularization=0.0, length_scale=1.0):
        self.kind = kind
        self.ptr = _abi.lib().hmg_kernel_new(kind, regularization, length_scale)
        if not self.ptr:
            raise _abi.HMGError(_abi.HMG_ERR_HIP, "hmg_kernel_new",
                                _abi.lib().hmg_last_error().decode())

    def free(self):
        if self.ptr:
            _abi.lib().hmg_kernel_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HMatrix:
    """One H-matrix built from two point sets (SoA arrays of shape (dim, n))."""

    def __init__(self, X1, X2, bits, eta, max_level, c_leaf, k, epsilon=1e-6,
                 root_levels=(0, 0), max_batched_aca_size=1 << 30):
        self.L = _abi.lib()
        X1 = np.ascontiguousarray(X1, dtype=np.float64)
        X2 = np.ascontiguousarray(X2, dtype=np.float64)
        self.dim = X1.shape[0]
        self.n1, self.n2 = X1.shape[1], X2.shape[1]
        self.data = _abi.HMatrixData()
        pc = (ctypes.c_int * 2)(self.n1, self.n2)
        _abi.check(self.L.hmg_init_h_matrix_data(ctypes.byref(self.data), pc, self.dim, bits),
                   "init_h_matrix_data")
        self._alive = True
        d = self.data
        d.eta = eta
        d.max_level = max_level
        d.c_leaf = c_leaf
        d.k = k
        d.epsilon = epsilon
        d.max_batched_dense_size = 8192
        d.max_batched_aca_size = max_batched_aca_size
        d.root_level_set_1 = bytes([root_levels[0] & 0xFF])
        d.root_level_set_2 = bytes([root_levels[1] & 0xFF])
        for s, X in ((0, X1), (1, X2)):
            for k_ in range(self.dim):
                memcpy_htod(d.coords_d[s][k_], X[k_])
        self.kernel = None

    # ---- on-disk format ------------------------------------------------------
    def save(self, path):
        """hmg_save: parameters, kernel, points and the precomputed factors."""
        _abi.check(self.L.hmg_save(ctypes.byref(self.data), os.fsencode(path)), "hmg_save")

    @classmethod
    def load(cls, path):
        """hmg_load: a ready H-matrix (setup rebuilt, factors installed); its kernel
        object belongs to the library."""
        self = cls.__new__(cls)
        self.L = _abi.lib()
        self.data = _abi.HMatrixData()
        _abi.check(self.L.hmg_load(ctypes.byref(self.data), os.fsencode(path)), "hmg_load")
        self._alive = True
        self.dim = self.data.dim
        self.n1, self.n2 = self.data.point_count[0], self.data.point_count[1]
        self.kernel = None
        return self

    # ---- lifecycle ---------------------------------------------------------
    def setup(self):
        _abi.check(self.L.hmg_setup_h_matrix(ctypes.byref(self.data)), "setup_h_matrix")

    def set_kernel(self, kernel):
        self.kernel = kernel
        self.data.assem = kernel.ptr

    def precompute_aca(self):
        _abi.check(self.L.hmg_precompute_aca(ctypes.byref(self.data)), "precompute_aca")

    def precompute_dense(self):
        _abi.check(self.L.hmg_precompute_dense(ctypes.byref(self.data)), "precompute_dense")

    def destroy(self):
        if getattr(self, "_alive", False):
            self.L.hmg_destroy_h_matrix_data(ctypes.byref(self.data))
            self._alive = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ---- products on device buffers ------------------------------------------
    def apply_dev(self, xd, yd, nrhs=1):
        _abi.check(self.L.hmg_apply_h_matrix_mvp_multi(xd.ptr, yd.ptr, nrhs, ctypes.byref(self.data)),
                   "apply_h_matrix_mvp")

    def autotune(self, xd, yd, nrhs=1, reps=2):
        """hmg_autotune_product on device arrays: keeps the fastest launch shape of
        the product and returns it (yd is overwritten)."""
        out = (ctypes.c_int64 * 4)()
        _abi.check(self.L.hmg_autotune_product(ctypes.byref(self.data), xd.ptr, yd.ptr, nrhs, reps, out),
                   "hmg_autotune_product")
        return {"near_lds_bytes": int(out[0]), "one_stream": bool(out[1]), "us_per_product": int(out[2]),
                "lr_workgroup": bool(out[3] & 1), "near_workgroup": bool(out[3] & 2)}

    # ---- products on host arrays ---------------------------------------------
    def apply(self, x):
        x = np.asarray(x, dtype=np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xd = DeviceArray(np.asfortranarray(x.reshape(self.n2, nrhs)))
        yd = DeviceArray.empty((self.n1, nrhs))
        self.apply_dev(xd, yd, nrhs)
        y = yd.download()
        return y[:, 0] if x.ndim == 1 else y

    def apply_without_batching(self, x):
        xd = DeviceArray(np.ascontiguousarray(x, dtype=np.float64))
        yd = DeviceArray.empty(self.n1)
        _abi.check(self.L.hmg_apply_h_matrix_mvp_without_batching(xd.ptr, yd.ptr, ctypes.byref(self.data)),
                   "apply_h_matrix_mvp_without_batching")
        return yd.download()

    def apply_full(self, x):
        xd = DeviceArray(np.ascontiguousarray(x, dtype=np.float64))
        yd = DeviceArray.empty(self.n1)
        _abi.check(self.L.hmg_apply_full_mvp(xd.ptr, yd.ptr, ctypes.byref(self.data)), "apply_full_mvp")
        return yd.download()

    def dense_rows(self, rows, x):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        x = np.asarray(x, dtype=np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xd = DeviceArray(np.asfortranarray(x.reshape(self.n2, nrhs)))
        rd = DeviceArray(rows)
        yd = DeviceArray.empty((rows.size, nrhs))
        _abi.check(self.L.hmg_dense_rows(xd.ptr, nrhs, rd.ptr, rows.size, yd.ptr, ctypes.byref(self.data)),
                   "hmg_dense_rows")
        y = yd.download()
        return y[:, 0] if x.ndim == 1 else y

    def gaussian_rhs(self):
        bd = DeviceArray.empty(self.n1)
        _abi.check(self.L.hmg_set_gaussian_kernel_rhs(bd.ptr, ctypes.byref(self.data)),
                   "set_gaussian_kernel_rhs")
        return bd.download()

    # ---- readers ----------------------------------------------------------------
    def codes(self, s):
        n = self.n1 if s == 0 else self.n2
        out = np.empty(n, dtype=np.uint64)
        return memcpy_dtoh(out, self.data.code_d[s])

    def order(self, s):
        n = self.n1 if s == 0 else self.n2
        out = np.empty(n, dtype=np.uint64)
        return memcpy_dtoh(out, self.data.order[s])

    def point_ids(self, s):
        n = self.n1 if s == 0 else self.n2
        out = np.empty(n, dtype=np.uint32)
        return memcpy_dtoh(out, self.data.point_ids_d[s])

    def sorted_coords(self, s):
        n = self.n1 if s == 0 else self.n2
        out = np.empty((self.dim, n))
        for k in range(self.dim):
            memcpy_dtoh(out[k], self.data.coords_d[s][k])
        return out

    def minmax(self, s):
        mn = np.empty(self.dim)
        mx = np.empty(self.dim)
        memcpy_dtoh(mn, self.data.min_per_dim_d[s])
        memcpy_dtoh(mx, self.data.max_per_dim_d[s])
        return mn, mx

    def blocks(self, dtype):
        n = self.data.mat_vec_data_count
        out = np.zeros(n, dtype=dtype)
        if n:
            memcpy_dtoh(out, ctypes.cast(self.data.mat_vec_data[0], ctypes.c_void_p).value)
        return out

    def schedule(self):
        out = (ctypes.c_int64 * 11)()
        _abi.check(self.L.hmg_get_schedule(ctypes.byref(self.data), out), "hmg_get_schedule")
        return {"near_items": out[0], "lr_items": out[1], "lr_rows": out[2], "lr_t_reads": out[3],
                "lr_m1": out[4], "lr_m2": out[5], "partials": out[6], "near_symmetric": bool(out[7]),
                "col_slots": out[8], "same_sets": bool(out[9]), "near_evals": out[10]}

    def memory(self):
        """hmg_memory: (bytes held now, further bytes the next build/product needs)."""
        out = (ctypes.c_int64 * 2)()
        _abi.check(self.L.hmg_memory(ctypes.byref(self.data), out), "hmg_memory")
        return int(out[0]), int(out[1])

    def set_option(self, option, value):
        """hmg_set_option: option is one of _abi.OPT_* (near-field symmetry takes
        effect at the next setup())."""
        _abi.check(self.L.hmg_set_option(ctypes.byref(self.data), int(option), int(value)),
                   "hmg_set_option")

    def set_precision(self, prec):
        """_abi.PREC_FP64 or _abi.PREC_MIXED: factor precision of the next
        precompute_aca(), and product precision when no factors are stored."""
        self.set_option(_abi.OPT_PRECISION, prec)

    def set_profiling(self, on=True):
        _abi.check(self.L.hmg_set_profiling(ctypes.byref(self.data), 1 if on else 0), "hmg_set_profiling")

    PROFILE_CLASSES = {"gather": 0, "lr_vtx": 1, "near": 2, "aca": 3, "dense_fill": 4, "lr_u": 5,
                       "combine": 6}

    def profile(self):
        out = {}
        for name, cls in self.PROFILE_CLASSES.items():
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            _abi.check(self.L.hmg_get_profile(ctypes.byref(self.data), cls, ctypes.byref(ms),
                                              ctypes.byref(n)), "hmg_get_profile")
            out[name] = (ms.value, n.value)
        return out

    def work(self):
        out = (ctypes.c_int64 * 9)()
        _abi.check(self.L.hmg_get_work(ctypes.byref(self.data), out), "hmg_get_work")
        return {"E_near": out[0], "E_lr": out[1], "dense": out[2], "aca": out[3], "levels": out[4],
                "quirk_levels": out[5], "lr_m1_rank": out[6], "lr_m2_rank": out[7],
                "rank_sum": out[8]}


def _data_array(hms):
    arr = (ctypes.POINTER(_abi.HMatrixData) * len(hms))()
    for i, h in enumerate(hms):
        arr[i] = ctypes.pointer(h.data)
    return arr


def setup_batch(hms, threads=8):
    """hmg_setup_batch: setup() of every HMatrix in `hms`, run concurrently."""
    _abi.check(_abi.lib().hmg_setup_batch(_data_array(hms), len(hms), threads), "hmg_setup_batch")


def precompute_aca_batch(hms, threads=8):
    """hmg_precompute_aca_batch: precompute_aca() of every HMatrix in `hms` (set up,
    kernel set) in one launch set per kernel / dimension / stopping rule /
    precision."""
    _abi.check(_abi.lib().hmg_precompute_aca_batch(_data_array(hms), len(hms), threads),
               "hmg_precompute_aca_batch")


class BatchGraph:
    """hmg_apply_batch_graph_*: the products of `hms` on the device arrays `xds` ->
    `yds` (nrhs[i] columns each) captured once into one hipGraph; launch() replays
    all of them.  The arrays stay bound to the graph; their contents may change."""

    def __init__(self, hms, xds, yds, nrhs):
        n = len(hms)
        self._keep = (hms, xds, yds)
        X = (ctypes.c_void_p * n)(*[x.ptr for x in xds])
        Y = (ctypes.c_void_p * n)(*[y.ptr for y in yds])
        R = (ctypes.c_int * n)(*[int(r) for r in nrhs])
        self._g = ctypes.c_void_p()
        _abi.check(_abi.lib().hmg_apply_batch_graph_create(_data_array(hms), X, Y, R, n, ctypes.byref(self._g)),
                   "hmg_apply_batch_graph_create")

    def launch(self):
        _abi.check(_abi.lib().hmg_apply_batch_graph_launch(self._g), "hmg_apply_batch_graph_launch")

    def destroy(self):
        if self._g:
            _abi.lib().hmg_apply_batch_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def apply_batch(hms, xds, yds, nrhs, threads=8):
    """hmg_apply_batch: yds[i] = H_i xds[i] (device arrays, nrhs[i] columns each),
    the products running concurrently."""
    n = len(hms)
    X = (ctypes.c_void_p * n)(*[x.ptr for x in xds])
    Y = (ctypes.c_void_p * n)(*[y.ptr for y in yds])
    R = (ctypes.c_int * n)(*[int(r) for r in nrhs])
    _abi.check(_abi.lib().hmg_apply_batch(_data_array(hms), X, Y, R, n, threads), "hmg_apply_batch")
