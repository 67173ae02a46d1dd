"""ctypes front-end of the CPU oracle (oracle/libhmg_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg -- never by the
product package hmglib_amd.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhmg_oracle.so")
_lib = None

WT_ACA = 1
WT_DENSE = 2
MAX_DIM = 5

K_GAUSSIAN_REF = 1
K_MATERN_REF = 2
K_GAUSSIAN = 3
K_EXPONENTIAL = 4


class WorkItem(ctypes.Structure):
    """192-byte work item, field order of tree.h:38-55."""
    _fields_ = [
        ("set1_l", ctypes.c_int), ("set1_u", ctypes.c_int),
        ("set2_l", ctypes.c_int), ("set2_u", ctypes.c_int),
        ("work_type", ctypes.c_int),
        ("max1", ctypes.c_double * MAX_DIM), ("min1", ctypes.c_double * MAX_DIM),
        ("max2", ctypes.c_double * MAX_DIM), ("min2", ctypes.c_double * MAX_DIM),
        ("dim", ctypes.c_int),
        ("level_1", ctypes.c_char), ("level_2", ctypes.c_char),
    ]


WORK_ITEM_DTYPE = np.dtype({
    "names": ["set1_l", "set1_u", "set2_l", "set2_u", "work_type", "max1", "min1",
              "max2", "min2", "dim", "level_1", "level_2"],
    "formats": ["<i4"] * 5 + [("<f8", (MAX_DIM,))] * 4 + ["<i4", "i1", "i1"],
    "offsets": [0, 4, 8, 12, 16, 24, 64, 104, 144, 184, 188, 189],
    "itemsize": 192,
})


class Kernel(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dim", ctypes.c_int),
                ("regularization", ctypes.c_double), ("length_scale", ctypes.c_double)]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        c_i64p = P(ctypes.c_int64)
        L.orc_splitmix64_next.restype = ctypes.c_uint64
        L.orc_splitmix64_next.argtypes = [P(ctypes.c_uint64)]
        L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, P(ctypes.c_double)]
        L.orc_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, P(ctypes.c_double)]
        L.orc_stretch.restype = ctypes.c_uint64
        L.orc_stretch.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.orc_morton.argtypes = [P(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                 P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_uint64)]
        L.orc_minmax.argtypes = [P(ctypes.c_double), ctypes.c_int, ctypes.c_int64,
                                 P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_sort.argtypes = [P(ctypes.c_uint64), ctypes.c_int64, P(ctypes.c_uint64)]
        L.orc_find_split.restype = ctypes.c_int
        L.orc_find_split.argtypes = [P(ctypes.c_uint64), ctypes.c_int, ctypes.c_int]
        L.orc_build_ex.restype = ctypes.c_void_p
        L.orc_build_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_build.restype = ctypes.c_void_p
        L.orc_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_free.argtypes = [ctypes.c_void_p]
        for f in ("orc_block_count", "orc_dense_count", "orc_aca_count", "orc_level_count",
                  "orc_num_threads"):
            getattr(L, f).restype = ctypes.c_int
        for f in ("orc_block_count", "orc_dense_count", "orc_aca_count", "orc_level_count"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_get_blocks.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_get_codes.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_uint64)]
        L.orc_get_order.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_uint64)]
        L.orc_get_sorted_coords.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double)]
        L.orc_get_minmax.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double),
                                     P(ctypes.c_double)]
        L.orc_get_diagnostics.argtypes = [ctypes.c_void_p, c_i64p]
        L.orc_kernel_entry.restype = ctypes.c_double
        L.orc_kernel_entry.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, ctypes.c_int]
        L.orc_aca.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, ctypes.c_int]
        L.orc_aca_sample.restype = ctypes.c_int64
        L.orc_aca_sample.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, ctypes.c_int]
        L.orc_aca_matrix.restype = ctypes.c_int
        L.orc_aca_matrix.argtypes = [P(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                     P(ctypes.c_int), P(ctypes.c_int)]
        L.orc_aca_tol.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double]
        L.orc_aca_matrix_tol.restype = ctypes.c_int
        L.orc_aca_matrix_tol.argtypes = [P(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int),
                                         P(ctypes.c_int)]
        L.orc_aca_u_size.restype = ctypes.c_int64
        L.orc_aca_v_size.restype = ctypes.c_int64
        L.orc_aca_u_size.argtypes = [ctypes.c_void_p]
        L.orc_aca_v_size.argtypes = [ctypes.c_void_p]
        L.orc_get_aca.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 8
        L.orc_hmvp.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, P(ctypes.c_double),
                               P(ctypes.c_double)]
        L.orc_hmvp_sample.restype = ctypes.c_int64
        L.orc_hmvp_sample.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, ctypes.c_int,
                                      P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_dense_mvp.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, P(ctypes.c_double),
                                    P(ctypes.c_double)]
        L.orc_dense_rows.argtypes = [ctypes.c_void_p, P(Kernel), ctypes.c_int, c_i64p, ctypes.c_int,
                                     P(ctypes.c_double), P(ctypes.c_double)]
        L.orc_gaussian_rhs.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.orc_work.argtypes = [ctypes.c_void_p, c_i64p]
        L.orc_work_sizes.restype = ctypes.c_int64
        L.orc_work_sizes.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_int), ctypes.c_int,
                                     P(ctypes.c_int), P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _i64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def uniform(seed, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_uniform(seed, n, _dp(out))
    return out


def normal(seed, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_normal(seed, n, _dp(out))
    return out


def splitmix64(seed, count):
    st = ctypes.c_uint64(seed)
    return [lib().orc_splitmix64_next(ctypes.byref(st)) for _ in range(count)]


def stretch(x, dim, bits):
    return lib().orc_stretch(x, dim, bits)


def minmax(X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    dim, n = X.shape
    mn = np.empty(dim)
    mx = np.empty(dim)
    lib().orc_minmax(_dp(X), dim, n, _dp(mn), _dp(mx))
    return mn, mx


def morton(X, bits, mn=None, mx=None):
    X = np.ascontiguousarray(X, dtype=np.float64)
    dim, n = X.shape
    if mn is None:
        mn, mx = minmax(X)
    codes = np.empty(n, dtype=np.uint64)
    lib().orc_morton(_dp(X), dim, bits, n, _dp(np.ascontiguousarray(mn)),
                     _dp(np.ascontiguousarray(mx)), _u64p(codes))
    return codes


def sort_codes(codes):
    c = np.array(codes, dtype=np.uint64, copy=True)
    order = np.empty(c.size, dtype=np.uint64)
    lib().orc_sort(_u64p(c), c.size, _u64p(order))
    return c, order


def find_split(codes, first, last):
    c = np.ascontiguousarray(codes, dtype=np.uint64)
    return lib().orc_find_split(_u64p(c), first, last)


def aca_matrix(A, k, tie_last=True, tol_mode=0, eps=0.0, eta=0.0):
    """ACA (reference pivot rules) of an explicit matrix; returns (U, V, row_piv, col_piv, rank).
    tol_mode 1: the reference's tolerance stop (every 5th step, eps (1-eta)/(1+eps));
    tol_mode 2: relative stop ||u_r|| ||v_r|| <= eps ||U V^T||_F at every step."""
    A = np.asfortranarray(A, dtype=np.float64)
    m1, m2 = A.shape
    kb = max(0, min(k, m1, m2))
    U = np.zeros((m1, max(kb, 1)), order="F")
    V = np.zeros((m2, max(kb, 1)), order="F")
    rp = np.full(max(kb, 1), -1, dtype=np.int32)
    cp = np.full(max(kb, 1), -1, dtype=np.int32)
    r = lib().orc_aca_matrix_tol(_dp(A), m1, m2, k, 1 if tie_last else 0, tol_mode, eps, eta,
                                 _dp(U), _dp(V),
                                 rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 cp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return U[:, :kb], V[:, :kb], rp[:kb], cp[:kb], r


def kernel(kind, dim, regularization=0.0, length_scale=1.0):
    return Kernel(kind, dim, regularization, length_scale)


class HMatrix:
    """Oracle H-matrix: build (Morton, sort, tree, partition), ACA, products."""

    def __init__(self, X1, X2, bits, eta, max_level, c_leaf, root_level_1=0, root_level_2=0,
                 exact_boxes=False):
        X1 = np.ascontiguousarray(X1, dtype=np.float64)
        X2 = np.ascontiguousarray(X2, dtype=np.float64)
        assert X1.shape[0] == X2.shape[0]
        self.dim = X1.shape[0]
        self.n1 = X1.shape[1]
        self.n2 = X2.shape[1]
        self._h = lib().orc_build_ex(self.n1, self.n2, self.dim, bits, _dp(X1), _dp(X2), eta,
                                     max_level, c_leaf, root_level_1, root_level_2, int(bool(exact_boxes)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_free(h)
            self._h = None

    @property
    def nblocks(self):
        return lib().orc_block_count(self._h)

    @property
    def ndense(self):
        return lib().orc_dense_count(self._h)

    @property
    def naca(self):
        return lib().orc_aca_count(self._h)

    @property
    def levels(self):
        return lib().orc_level_count(self._h)

    def blocks(self):
        out = np.zeros(self.nblocks, dtype=WORK_ITEM_DTYPE)
        lib().orc_get_blocks(self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def codes(self, s):
        out = np.empty(self.n1 if s == 0 else self.n2, dtype=np.uint64)
        lib().orc_get_codes(self._h, s, _u64p(out))
        return out

    def order(self, s):
        out = np.empty(self.n1 if s == 0 else self.n2, dtype=np.uint64)
        lib().orc_get_order(self._h, s, _u64p(out))
        return out

    def sorted_coords(self, s):
        n = self.n1 if s == 0 else self.n2
        out = np.empty(self.dim * n, dtype=np.float64)
        lib().orc_get_sorted_coords(self._h, s, _dp(out))
        return out.reshape(self.dim, n)

    def minmax(self, s):
        mn = np.empty(self.dim)
        mx = np.empty(self.dim)
        lib().orc_get_minmax(self._h, s, _dp(mn), _dp(mx))
        return mn, mx

    def diagnostics(self):
        out = np.zeros(4, dtype=np.int64)
        lib().orc_get_diagnostics(self._h, _i64p(out))
        return {"singleton_levels": int(out[0]), "box_differs": int(out[1]),
                "fma_flips": int(out[2]), "undefined_table": int(out[3])}

    def kernel_entry(self, kern, i, j):
        return lib().orc_kernel_entry(self._h, ctypes.byref(kern), i, j)

    def aca(self, kern, k, tie_last=True, tol_mode=0, eps=0.0, eta=0.0):
        lib().orc_aca_tol(self._h, ctypes.byref(kern), k, 1 if tie_last else 0, tol_mode, eps, eta)

    def aca_sample(self, kern, k, stride):
        return lib().orc_aca_sample(self._h, ctypes.byref(kern), k, stride)

    def aca_factors(self):
        na = self.naca
        us = lib().orc_aca_u_size(self._h)
        vs = lib().orc_aca_v_size(self._h)
        kb = np.zeros(na, dtype=np.int32)
        keff = np.zeros(na, dtype=np.int32)
        uoff = np.zeros(na + 1, dtype=np.int64)
        voff = np.zeros(na + 1, dtype=np.int64)
        U = np.zeros(max(us, 1))
        V = np.zeros(max(vs, 1))
        kmax = int(kb.max()) if na else 0
        lib().orc_get_aca(self._h, kb.ctypes.data, keff.ctypes.data, uoff.ctypes.data,
                          voff.ctypes.data, U.ctypes.data, V.ctypes.data, None, None)
        kmax = int(kb.max()) if na else 0
        rp = np.zeros(max(na * max(kmax, 1), 1), dtype=np.int32)
        cp = np.zeros(max(na * max(kmax, 1), 1), dtype=np.int32)
        lib().orc_get_aca(self._h, None, None, None, None, None, None, rp.ctypes.data,
                          cp.ctypes.data)
        return {"kb": kb, "keff": keff, "uoff": uoff, "voff": voff, "U": U[:us], "V": V[:vs],
                "row_piv": rp[:na * max(kmax, 1)].reshape(na, max(kmax, 1)),
                "col_piv": cp[:na * max(kmax, 1)].reshape(na, max(kmax, 1))}

    def hmvp(self, kern, x):
        x = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n2, -1))
        R = x.shape[1]
        y = np.zeros((self.n1, R), order="F")
        lib().orc_hmvp(self._h, ctypes.byref(kern), R, _dp(x), _dp(y))
        return y[:, 0] if R == 1 else y

    def hmvp_sample(self, kern, stride, x):
        """H-MVP over every stride-th block (R = columns of x); returns (y, entries
        touched) -- times a bounded CPU sample."""
        x = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n2, -1))
        R = x.shape[1]
        y = np.zeros((self.n1, R), order="F")
        e = lib().orc_hmvp_sample(self._h, ctypes.byref(kern), stride, R, _dp(x), _dp(y))
        return (y[:, 0] if R == 1 else y), e

    def dense_mvp(self, kern, x):
        x = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n2, -1))
        R = x.shape[1]
        y = np.zeros((self.n1, R), order="F")
        lib().orc_dense_mvp(self._h, ctypes.byref(kern), R, _dp(x), _dp(y))
        return y[:, 0] if R == 1 else y

    def dense_rows(self, kern, rows, x):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        x = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n2, -1))
        R = x.shape[1]
        y = np.zeros((rows.size, R), order="F")
        lib().orc_dense_rows(self._h, ctypes.byref(kern), rows.size, _i64p(rows), R, _dp(x),
                             _dp(y))
        return y[:, 0] if R == 1 else y

    def gaussian_rhs(self):
        b = np.zeros(self.n1)
        lib().orc_gaussian_rhs(self._h, _dp(b))
        return b

    def work(self):
        out = np.zeros(2, dtype=np.int64)
        lib().orc_work(self._h, _i64p(out))
        return int(out[0]), int(out[1])

    def work_sizes(self, max_batched_aca_size):
        """Batch sizing of setup (hmglib.cu:211-219): (dense cap, dense batch sizes,
        ACA batch sizes); each list ends with the 0 the reference records."""
        ci = ctypes.POINTER(ctypes.c_int)
        nd, na = ctypes.c_int(), ctypes.c_int()
        cap = lib().orc_work_sizes(self._h, int(max_batched_aca_size), None, 0, ctypes.byref(nd), None, 0,
                                   ctypes.byref(na))
        dws = np.zeros(max(nd.value, 1), dtype=np.int32)
        aws = np.zeros(max(na.value, 1), dtype=np.int32)
        lib().orc_work_sizes(self._h, int(max_batched_aca_size), dws.ctypes.data_as(ci), nd.value,
                             ctypes.byref(nd), aws.ctypes.data_as(ci), na.value, ctypes.byref(na))
        return int(cap), dws[:nd.value], aws[:na.value]


def num_threads():
    return lib().orc_num_threads()
