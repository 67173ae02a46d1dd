"""Python front-end of the serving loop (include/hmglib_server.h).

Used by tests and the serving benchmark.  The server itself is native: worker
threads in libhmglib.so merge apply requests on the same cached H-matrix into
product passes, run jobs on different H-matrices concurrently on the device and
admit builds against an HBM budget.  Python only packs requests.
"""
import ctypes

import numpy as np

from . import _abi


class Server:
    def __init__(self, budget_bytes=200e9, slots=4, max_rhs_per_pass=32, precision=_abi.PREC_MIXED,
                 product_slots=0):
        self.L = _abi.lib()
        cfg = _abi.ServerConfig(int(budget_bytes), int(slots), int(max_rhs_per_pass), int(precision),
                                int(product_slots))
        h = ctypes.c_void_p()
        _abi.check(self.L.hmg_server_create(ctypes.byref(cfg), ctypes.byref(h)), "hmg_server_create")
        self.h = h
        self._keep = {}   # host / device buffers that must outlive their requests

    def submit_build(self, X1, X2, bits, eta, max_level, c_leaf, k, kind, regularization=0.0,
                     length_scale=1.0, epsilon=0.0, aca_tolerance=0, exact_boxes=False):
        """X1, X2: (dim, n) point arrays (X2 is X1 for one point set).  Returns
        (request id, matrix id).  aca_tolerance: HMG_OPT_ACA_TOLERANCE mode with
        epsilon; exact_boxes: HMG_OPT_EXACT_BOXES."""
        X1 = np.ascontiguousarray(X1, dtype=np.float64)
        same = X2 is X1
        X2 = X1 if same else np.ascontiguousarray(X2, dtype=np.float64)
        b = _abi.BuildRequest()
        b.dim = X1.shape[0]
        b.bits = bits
        b.n[0], b.n[1] = X1.shape[1], X2.shape[1]
        b.coords[0] = X1.ctypes.data_as(_abi.c_double_p)
        b.coords[1] = X2.ctypes.data_as(_abi.c_double_p)
        b.eta, b.max_level, b.c_leaf, b.k = eta, max_level, c_leaf, k
        b.kernel_kind, b.regularization, b.length_scale = kind, regularization, length_scale
        b.epsilon, b.aca_tolerance = epsilon, aca_tolerance
        b.exact_boxes = int(bool(exact_boxes))
        r = ctypes.c_int64()
        m = ctypes.c_int64()
        _abi.check(self.L.hmg_server_submit_build(self.h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(m)),
                   "hmg_server_submit_build")
        return r.value, m.value

    def submit_apply(self, matrix, xd, yd, nrhs):
        """xd, yd: hip.DeviceArray (n2 x nrhs, n1 x nrhs column-major); they must
        stay alive until the request completes (the server keeps a reference)."""
        r = ctypes.c_int64()
        _abi.check(self.L.hmg_server_submit_apply(self.h, int(matrix), xd.ptr, yd.ptr, int(nrhs),
                                                  ctypes.byref(r)), "hmg_server_submit_apply")
        self._keep[r.value] = (xd, yd)
        return r.value

    def submit_apply_ptr(self, matrix, xptr, yptr, nrhs):
        """Like submit_apply on raw device pointers owned by the caller."""
        r = ctypes.c_int64()
        _abi.check(self.L.hmg_server_submit_apply(self.h, int(matrix), xptr, yptr, int(nrhs),
                                                  ctypes.byref(r)), "hmg_server_submit_apply")
        return r.value

    @staticmethod
    def _info(i):
        return {"status": i.status, "t_submit": i.t_submit, "t_start": i.t_start, "t_end": i.t_end,
                "merged": i.merged, "matrix": i.matrix, "message": i.message.decode(errors="replace")}

    def wait(self, request):
        i = _abi.RequestInfo()
        _abi.check(self.L.hmg_server_wait(self.h, int(request), ctypes.byref(i)), "hmg_server_wait")
        self._keep.pop(request, None)
        return self._info(i)

    def query(self, request):
        i = _abi.RequestInfo()
        _abi.check(self.L.hmg_server_query(self.h, int(request), ctypes.byref(i)), "hmg_server_query")
        return self._info(i)

    def drain(self):
        _abi.check(self.L.hmg_server_drain(self.h), "hmg_server_drain")

    def release(self, matrix):
        _abi.check(self.L.hmg_server_release(self.h, int(matrix)), "hmg_server_release")

    def stats(self):
        out = (ctypes.c_int64 * 8)()
        _abi.check(self.L.hmg_server_stats(self.h, out), "hmg_server_stats")
        keys = ("bytes", "cached", "evictions", "completed", "failed", "passes", "shared", "builds")
        return dict(zip(keys, list(out)))

    def resident(self):
        cap = 4096
        ids = (ctypes.c_int64 * cap)()
        n = self.L.hmg_server_resident(self.h, ids, cap)
        return list(ids[:min(n, cap)])

    def destroy(self):
        if getattr(self, "h", None):
            self.L.hmg_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
