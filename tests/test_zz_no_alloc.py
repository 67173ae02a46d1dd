"""The product entry point that allocates nothing (hmg_apply_h_matrix_mvp_noalloc),
through the C ABI, against the allocating entry point (hmg_apply_h_matrix_mvp_multi)
on the same H-matrix: the same kernels in the same order, so the results are
bitwise equal.  Two calls in a row share the library's fixed device scratch; the
second is on a smaller H-matrix, so stale values of the first call sit past the
second call's buffers and inside them."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(gpu, hm, X):
    from hmglib_amd import hip
    L = gpu.lib()
    nrhs = X.shape[1]
    xd = hip.DeviceArray(np.asfortranarray(X))
    yd = hip.DeviceArray.empty((hm.n1, nrhs))
    st = L.hmg_apply_h_matrix_mvp_noalloc(xd.ptr, yd.ptr, nrhs, ctypes.byref(hm.data))
    if st != gpu.HMG_OK:
        msg = L.hmg_last_error()
        xd.free()
        yd.free()
        return st, (msg.decode() if msg else "")
    y = yd.download().reshape(hm.n1, nrhs, order="F")
    xd.free()
    yd.free()
    return st, y


def _built(gpu, n, dim, seed, c_leaf, eta, k, kind, ls):
    X = O.uniform(seed, dim * n).reshape(dim, n)
    hm = gpu.HMatrix(X, X, 32 if dim == 2 else 20, eta, 50, c_leaf, k)
    hm.setup()
    kern = gpu.Kernel(kind, 0.0, ls)
    hm.set_kernel(kern)
    hm.precompute_aca()
    return hm, kern


def test_noalloc_product_equals_allocating_product_twice(gpu):
    # first call: the small plain configuration (N = 4096, 2D, Gaussian l = 0.1,
    # leaf 64, eta 1.5, rank 32); second call: a smaller 3D H-matrix
    cases = [(4096, 2, 1001, 64, 1.5, 32, O.K_GAUSSIAN, 0.1, 3),
             (1500, 3, 1002, 32, 1.0, 16, O.K_EXPONENTIAL, 0.2, 2)]
    for n, dim, seed, c_leaf, eta, k, kind, ls, nrhs in cases:
        hm, kern = _built(gpu, n, dim, seed, c_leaf, eta, k, kind, ls)
        X = O.normal(seed + 7, n * nrhs).reshape(n, nrhs, order="F")
        held = hm.memory()[0]
        st, Y = _noalloc(gpu, hm, X)
        assert st == gpu.HMG_OK, Y
        # the no-alloc product took no buffer of its own from the engine
        assert hm.memory()[0] == held
        # bitwise the allocating entry point's one-right-hand-side-at-a-time path ...
        hm.set_option(gpu.OPT_FP64_BLOCK_RHS, 1 << 20)
        Yref = hm.apply(X)
        np.testing.assert_array_equal(Y, Yref)
        # ... and within rounding of its block path (the default for several right-hand sides)
        hm.set_option(gpu.OPT_FP64_BLOCK_RHS, 2)
        Yblk = hm.apply(X)
        for c in range(nrhs):
            assert np.linalg.norm(Y[:, c] - Yblk[:, c]) / np.linalg.norm(Yblk[:, c]) < 1e-12
        # and the product is the H-matrix's (oracle H-MVP of the same tree and ranks)
        om = O.HMatrix(O.uniform(seed, dim * n).reshape(dim, n), O.uniform(seed, dim * n).reshape(dim, n),
                       32 if dim == 2 else 20, eta, 50, c_leaf)
        ok = O.kernel(kind, dim, 0.0, ls)
        om.aca(ok, k)
        for c in range(nrhs):
            yo = om.hmvp(ok, X[:, c])
            assert np.linalg.norm(Y[:, c] - yo) / np.linalg.norm(yo) < 1e-8
        hm.destroy()
        kern.free()


def test_noalloc_product_refusals(gpu):
    # mixed-precision factors: refused (the no-alloc product is the FP64 one)
    n = 2048
    X = O.uniform(1003, 2 * n).reshape(2, n)
    hm = gpu.HMatrix(X, X, 32, 1.5, 50, 64, 16)
    hm.setup()
    kern = gpu.Kernel(O.K_GAUSSIAN, 0.0, 0.1)
    hm.set_kernel(kern)
    st, msg = _noalloc(gpu, hm, np.ones((n, 1)))
    assert st == gpu.HMG_ERR_NOT_READY and "precompute_aca" in msg
    hm.set_precision(gpu.PREC_MIXED)
    hm.precompute_aca()
    st, msg = _noalloc(gpu, hm, np.ones((n, 1)))
    assert st == gpu.HMG_ERR_NOT_READY
    hm.destroy()
    kern.free()
    # an H-matrix whose per-call buffers exceed the fixed scratch: a clear error
    n = 1000000
    hm, kern = _built(gpu, n, 2, 1004, 64, 1.5, 4, O.K_GAUSSIAN, 0.1)
    st, msg = _noalloc(gpu, hm, np.ones((n, 1)))
    assert st == gpu.HMG_ERR_INVALID and "fixed scratch holds" in msg
    hm.destroy()
    kern.free()
