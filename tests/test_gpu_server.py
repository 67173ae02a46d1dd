"""GPU tests of the serving loop (include/hmglib_server.h): apply requests on the
same cached H-matrix are merged into shared product passes, jobs on different
H-matrices overlap on the device, builds are admitted against an HBM budget.

Every result is checked against the same H-matrix built and applied alone
through the plain C ABI.  Right-hand sides are computed independently of each
other (per-column MFMA accumulation, per-column partial sums), so a request's
result is bitwise the same alone and in the mix.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _points(seed, dim, n):
    return O.uniform(seed, dim * n).reshape(dim, n)


CASES = [
    # (dim, bits, n, eta, c_leaf, k, kind, ls, seed)
    (2, 32, 3000, 1.5, 64, 16, O.K_GAUSSIAN, 0.1, 11),
    (3, 20, 5000, 1.0, 128, 12, O.K_EXPONENTIAL, 0.1, 12),
    (2, 32, 1800, 1.2, 64, 16, O.K_EXPONENTIAL, 0.1, 13),
]


def _standalone(gpu, case, prec):
    dim, bits, n, eta, c_leaf, k, kind, ls, seed = case
    X = _points(seed, dim, n)
    hm = gpu.HMatrix(X, X, bits, eta, 50, c_leaf, k)
    hm.setup()
    kern = gpu.Kernel(kind, 0.0, ls)
    hm.set_kernel(kern)
    hm.set_precision(prec)
    # the server runs every FP64 pass through the block product (a right-hand side's
    # result is then the same in every pass width); so does the reference H-matrix
    hm.set_option(gpu.OPT_FP64_BLOCK_RHS, 1)
    hm.precompute_aca()
    return hm, kern, X


def _submit_build(srv, case, X):
    dim, bits, n, eta, c_leaf, k, kind, ls, seed = case
    return srv.submit_build(X, X, bits, eta, 50, c_leaf, k, kind, 0.0, ls)


@pytest.mark.parametrize("prec_name", ["mixed", "fp64"])
def test_results_same_alone_and_in_the_mix(gpu, prec_name):
    from hmglib_amd import hip
    from hmglib_amd.server import Server
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    srv = Server(budget_bytes=8e9, slots=3, max_rhs_per_pass=32, precision=prec)
    try:
        mats, refs = [], []
        for case in CASES:
            hm, kern, X = _standalone(gpu, case, prec)
            refs.append((hm, kern))
            _, m = _submit_build(srv, case, X)
            mats.append(m)
        # apply requests submitted back to back (the first ones wait for their
        # build), several per H-matrix so that passes are shared
        rng = np.random.default_rng(5)
        reqs = []
        for q in range(24):
            i = q % len(CASES)
            n = CASES[i][2]
            nrhs = int(rng.integers(1, 13))
            Xh = np.asfortranarray(rng.standard_normal((n, nrhs)))
            xd = hip.DeviceArray(Xh)
            yd = hip.DeviceArray.empty((n, nrhs))
            reqs.append((srv.submit_apply(mats[i], xd, yd, nrhs), i, Xh, xd, yd))
        for r, i, Xh, xd, yd in reqs:
            info = srv.wait(r)
            assert info["status"] == gpu.HMG_OK
            assert info["t_submit"] <= info["t_start"] <= info["t_end"]
            Y = yd.download()
            Yref = refs[i][0].apply(Xh)
            np.testing.assert_array_equal(Y.reshape(Yref.shape), Yref)
        st = srv.stats()
        assert st["builds"] == len(CASES) and st["completed"] == len(CASES) + len(reqs)
        assert st["failed"] == 0 and st["evictions"] == 0
        assert st["passes"] <= len(reqs)
        assert sorted(srv.resident()) == sorted(mats)
    finally:
        srv.destroy()


def test_requests_waiting_on_one_matrix_share_a_pass(gpu):
    """One slot, kept busy by a build: apply requests that arrive meanwhile on a
    cached H-matrix are served together in one pass, each result bitwise equal to
    its product alone."""
    from hmglib_amd import hip
    from hmglib_amd.server import Server
    case_a, case_b = CASES[0], (3, 20, 20000, 1.0, 256, 24, O.K_EXPONENTIAL, 0.1, 14)
    hm, kern, Xa = _standalone(gpu, case_a, gpu.PREC_MIXED)
    Xb = _points(case_b[8], case_b[0], case_b[2])
    srv = Server(budget_bytes=16e9, slots=1, max_rhs_per_pass=32, precision=gpu.PREC_MIXED)
    try:
        r, ma = _submit_build(srv, case_a, Xa)
        assert srv.wait(r)["status"] == gpu.HMG_OK
        passes0 = srv.stats()["passes"]
        rb, _ = _submit_build(srv, case_b, Xb)      # occupies the only slot
        rng = np.random.default_rng(9)
        n = case_a[2]
        reqs = []
        for nrhs in (1, 2, 3, 4, 5, 6):
            Xh = np.asfortranarray(rng.standard_normal((n, nrhs)))
            xd = hip.DeviceArray(Xh)
            yd = hip.DeviceArray.empty((n, nrhs))
            reqs.append((srv.submit_apply(ma, xd, yd, nrhs), Xh, xd, yd))
        assert srv.wait(rb)["status"] == gpu.HMG_OK
        for r, Xh, xd, yd in reqs:
            info = srv.wait(r)
            assert info["status"] == gpu.HMG_OK and info["merged"] == len(reqs)
            Yref = hm.apply(Xh)
            np.testing.assert_array_equal(yd.download().reshape(Yref.shape), Yref)
        st = srv.stats()
        assert st["passes"] == passes0 + 1 and st["shared"] == len(reqs)
    finally:
        srv.destroy()


def test_budget_admission_and_eviction(gpu):
    from hmglib_amd import hip
    from hmglib_amd.hmatrix import HMatrix
    from hmglib_amd.server import Server
    case = CASES[1]
    # measure one H-matrix's footprint
    hm, kern, X = _standalone(gpu, case, gpu.PREC_MIXED)
    x = O.normal(21, case[2])
    y_ref = hm.apply(x)
    srv = Server(budget_bytes=1 << 40, slots=2, precision=gpu.PREC_MIXED)
    r, m0 = _submit_build(srv, case, X)
    assert srv.wait(r)["status"] == gpu.HMG_OK
    one = srv.stats()["bytes"]
    srv.destroy()
    assert one > 0
    # a build is admitted at its peak (the tree's bytes plus the estimate for the
    # factors, the FP64 ACA temporaries and the product buffers): the budget holds
    # one built H-matrix beside one build in progress, not two
    dim, bits, n, eta, c_leaf, k, kind, ls, seed = case
    hb = gpu.HMatrix(X, X, bits, eta, 50, c_leaf, k)
    hb.setup()
    hb.set_kernel(kern)
    hb.set_precision(gpu.PREC_MIXED)
    need = sum(hb.memory())
    assert need >= one
    budget = one + need + one // 4
    srv = Server(budget_bytes=budget, slots=2, precision=gpu.PREC_MIXED)
    try:
        ids = []
        for _ in range(4):
            r, m = _submit_build(srv, case, X)
            assert srv.wait(r)["status"] == gpu.HMG_OK
            ids.append(m)
            st = srv.stats()
            assert st["bytes"] <= budget
        st = srv.stats()
        assert st["evictions"] == 2 and st["cached"] == 2
        assert srv.resident() == ids[2:]          # least recently used went first
        xd = hip.DeviceArray(x)
        yd = hip.DeviceArray.empty(case[2])
        with pytest.raises(gpu.HMGError):
            srv.submit_apply(ids[0], xd, yd, 1)   # evicted
        assert srv.wait(srv.submit_apply(ids[3], xd, yd, 1))["status"] == gpu.HMG_OK
        np.testing.assert_array_equal(yd.download(), y_ref)
        # a build that cannot fit even in an empty cache fails, and so do the
        # applies waiting on it
        small = Server(budget_bytes=1 << 20, slots=1, precision=gpu.PREC_MIXED)
        try:
            r, m = _submit_build(small, case, X)
            ra = small.submit_apply(m, xd, yd, 1)
            assert small.wait(r)["status"] == gpu.HMG_ERR_OOM
            assert small.wait(ra)["status"] == gpu.HMG_ERR_OOM
            assert small.stats()["failed"] == 2
        finally:
            small.destroy()
    finally:
        srv.destroy()
