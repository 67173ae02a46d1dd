"""GPU tests at the benchmark configurations' parameters, each result against the
CPU oracle.

  C2  plain FP64 at scale: 3D uniform (seed 2001), exp(-r/0.1), leaf 256, eta 1.0,
      rank 24 -- here at N = 2^15 so that the oracle finishes in seconds; the
      fixed-rank product against the oracle's H-MVP (1e-8) and both against exact
      FP64 rows sampled as the benchmark samples them (seed 2003), and the
      tolerance form of the configuration (1e-7, rank cap 24).
  C4  many clouds: 8-blob Gaussian mixtures, 2D and 3D alternating, Gaussian
      l = 0.1, leaf 64, eta 1.2, relative ACA tolerance 1e-5 with rank cap 32,
      1-8 right-hand sides; ten clouds submitted at once to the server, each
      result against the oracle's tolerance ACA + H-MVP.
  C5  serving: uniform and blob clouds, exponential and Gaussian kernels, 2D and
      3D, several right-hand sides; builds and applies submitted at once, each
      served result against the oracle and bitwise equal to the same H-matrix's
      product alone.
  m = 1 clusters: one-point clusters in DENSE and ACA blocks (the reference's
      batched maps break there, linear_algebra.cu:588-605, 690-707); products
      against the oracle and the dense product.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _sample_rows(seed, n, count):
    u = O.uniform(seed, n)
    return np.argsort(u, kind="stable")[:count].astype(np.int64)


def _blob_cloud(rng, n, dim):
    centres = rng.random((8, dim))
    lab = rng.integers(0, 8, n)
    return np.clip(centres[lab].T + 0.05 * rng.standard_normal((dim, n)), 0.0, 1.0)


# ------------------------------------------------------------------------- C2
@pytest.mark.parametrize("mode", ["fixed_rank", "tolerance"])
def test_c2_parameters_fp64_against_oracle(gpu, mode):
    n = 1 << 15
    X = O.uniform(2001, 3 * n).reshape(3, n)
    eps = 1e-7
    hm = gpu.HMatrix(X, X, 20, 1.0, 50, 256, 24, epsilon=eps)
    hm.setup()
    kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.1)
    hm.set_kernel(kern)
    if mode == "tolerance":
        hm.set_option(gpu.OPT_ACA_TOLERANCE, 2)
    hm.precompute_aca()
    om = O.HMatrix(X, X, 20, 1.0, 50, 256)
    ok = O.kernel(O.K_EXPONENTIAL, 3, 0.0, 0.1)
    g, o = hm.blocks(O.WORK_ITEM_DTYPE), om.blocks()
    for f in ("set1_l", "set1_u", "set2_l", "set2_u", "work_type", "level_1"):
        np.testing.assert_array_equal(g[f], o[f])
    if mode == "tolerance":
        om.aca(ok, 24, tol_mode=2, eps=eps)
    else:
        om.aca(ok, 24)
    keff = om.aca_factors()["keff"]
    w = hm.work()
    if mode == "fixed_rank":
        assert w["rank_sum"] == int(keff.sum()) and w["E_lr"] == om.work()[1]
    else:
        # at these parameters the relative criterion 1e-7 is not met before the
        # rank cap 24 in (nearly) any block, so the ranks are (nearly) the cap
        assert abs(w["rank_sum"] - int(keff.sum())) <= 5e-3 * keff.sum()
        assert keff.sum() <= om.aca_factors()["kb"].sum()
    x = O.normal(2002, n)
    y = hm.apply(x)
    y_o = om.hmvp(ok, x)
    assert _rel(y, y_o) < (1e-8 if mode == "fixed_rank" else 1e-6)
    # exact FP64 rows (the benchmark's error measure), engine and oracle alike
    rows = _sample_rows(2003, n, 2048)
    y_rows = om.dense_rows(ok, rows, x)
    # rows whose sum cancels: relative to the row's sum of |K_ij x_j| scale
    np.testing.assert_allclose(hm.dense_rows(rows, x), y_rows, rtol=1e-12, atol=1e-12 * np.abs(y_rows).max())
    err_g, err_o = _rel(y[rows], y_rows), _rel(y_o[rows], y_rows)
    assert abs(err_g - err_o) < 1e-8
    # the H-approximation error of rank <= 24 at leaf 256, eta 1.0 (the benchmark
    # measures 2.7e-5 at N = 2^20); the tolerance form stops earlier than rank 24 only
    # where the relative criterion 1e-7 holds, so it is no larger than ~1e-4 either
    assert err_g < 1e-4


# ------------------------------------------------------------------------- C4
@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
def test_c4_clouds_batched_through_the_server(gpu, prec_name):
    from hmglib_amd import hip
    from hmglib_amd.server import Server
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    eps = 1e-5
    clouds = []
    for c in range(10):
        rng = np.random.default_rng(4001 + c)
        n = int(2 ** (10 + 3 * rng.random()))        # 2^10 .. 2^13 (the oracle's share)
        dim = 2 if c % 2 == 0 else 3
        X = _blob_cloud(rng, n, dim)
        nrhs = int(rng.integers(1, 9))
        clouds.append((X, dim, n, nrhs, np.asfortranarray(rng.standard_normal((n, nrhs)))))
    srv = Server(budget_bytes=16e9, slots=4, max_rhs_per_pass=32, precision=prec)
    try:
        sub = []
        for X, dim, n, nrhs, Xh in clouds:
            rb, m = srv.submit_build(X, X, 32 if dim == 2 else 20, 1.2, 50, 64, 32, O.K_GAUSSIAN, 0.0, 0.1,
                                     epsilon=eps, aca_tolerance=2)
            xd = hip.DeviceArray(Xh)
            yd = hip.DeviceArray.empty((n, nrhs))
            ra = srv.submit_apply(m, xd, yd, nrhs)
            sub.append((rb, ra, xd, yd))
        for (X, dim, n, nrhs, Xh), (rb, ra, xd, yd) in zip(clouds, sub):
            assert srv.wait(rb)["status"] == gpu.HMG_OK
            assert srv.wait(ra)["status"] == gpu.HMG_OK
            Y = yd.download().reshape(n, nrhs)
            om = O.HMatrix(X, X, 32 if dim == 2 else 20, 1.2, 50, 64)
            ok = O.kernel(O.K_GAUSSIAN, dim, 0.0, 0.1)
            om.aca(ok, 32, tol_mode=2, eps=eps)
            Yo = om.hmvp(ok, Xh).reshape(n, nrhs)
            Yd = om.dense_mvp(ok, Xh).reshape(n, nrhs)
            for c in range(nrhs):
                # the H-approximation error of these clouds is that of the
                # reference's algorithm, not of the tolerance: clipped blobs make
                # one-point clusters whose shifted lookup-table boxes admit blocks
                # that are not separated (up to 0.27 here, oracle and engine alike),
                # and the sequential row pivots can stop a block early
                h_err = _rel(Yo[:, c], Yd[:, c])
                if prec_name == "fp64":
                    # rank decisions at the tolerance may differ in a few blocks
                    # (exp last ulps), each moving the product by about eps there
                    assert _rel(Y[:, c], Yo[:, c]) < 10 * eps
                    assert _rel(Y[:, c], Yd[:, c]) < h_err + 10 * eps
                else:
                    assert _rel(Y[:, c], Yd[:, c]) < h_err + 4e-5
        assert srv.stats()["failed"] == 0
    finally:
        srv.destroy()


# ------------------------------------------------------------------------- C5
C5_CASES = [
    # (dim, blobs, kind, n, seed): tree parameters as the serving workload's
    (2, False, O.K_EXPONENTIAL, 5000, 51),
    (3, True, O.K_GAUSSIAN, 6000, 52),
    (2, True, O.K_GAUSSIAN, 4500, 53),
    (3, False, O.K_EXPONENTIAL, 7000, 54),
]


@pytest.mark.parametrize("exact", [False, True], ids=["reference_boxes", "exact_boxes"])
@pytest.mark.parametrize("prec_name", ["mixed", "fp64"])
def test_c5_served_results_against_oracle(gpu, prec_name, exact):
    from hmglib_amd import hip
    from hmglib_amd.server import Server
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    srv = Server(budget_bytes=16e9, slots=3, max_rhs_per_pass=32, precision=prec, product_slots=1)
    try:
        mats = []
        for dim, blobs, kind, n, seed in C5_CASES:
            rng = np.random.default_rng(seed)
            X = _blob_cloud(rng, n, dim) if blobs else rng.random((dim, n))
            bits, leaf, eta, k = (32, 64, 1.5, 32) if dim == 2 else (20, 256, 1.0, 24)
            rb, m = srv.submit_build(X, X, bits, eta, 50, leaf, k, kind, 0.0, 0.1, exact_boxes=exact)
            mats.append((X, bits, leaf, eta, k, kind, dim, n, rb, m))
        rng = np.random.default_rng(55)
        reqs = []
        for q in range(16):
            i = q % len(mats)
            n = mats[i][7]
            nrhs = int(rng.integers(1, 17))
            Xh = np.asfortranarray(rng.standard_normal((n, nrhs)))
            xd = hip.DeviceArray(Xh)
            yd = hip.DeviceArray.empty((n, nrhs))
            reqs.append((srv.submit_apply(mats[i][9], xd, yd, nrhs), i, Xh, xd, yd))
        for X, bits, leaf, eta, k, kind, dim, n, rb, m in mats:
            assert srv.wait(rb)["status"] == gpu.HMG_OK
        oracles = []
        for X, bits, leaf, eta, k, kind, dim, n, rb, m in mats:
            om = O.HMatrix(X, X, bits, eta, 50, leaf, exact_boxes=exact)
            ok = O.kernel(kind, dim, 0.0, 0.1)
            om.aca(ok, k)
            oracles.append((om, ok))
        for r, i, Xh, xd, yd in reqs:
            assert srv.wait(r)["status"] == gpu.HMG_OK
            n = mats[i][7]
            Y = yd.download().reshape(n, -1)
            om, ok = oracles[i]
            Yo = om.hmvp(ok, Xh).reshape(n, -1)
            for c in range(Y.shape[1]):
                if prec_name == "fp64":
                    assert _rel(Y[:, c], Yo[:, c]) < 1e-6
                else:
                    # the BF16x3 arithmetic bound, relative to sum_j K_ij |x_j| (positive kernels)
                    mag = np.linalg.norm(om.hmvp(ok, np.abs(Xh[:, c])))
                    assert np.linalg.norm(Y[:, c] - Yo[:, c]) / mag < 4e-5
            if exact:
                # exact boxes: the served result is within 1e-4 of the dense product
                Yd = om.dense_mvp(ok, Xh[:, :2]).reshape(n, -1)
                for c in range(Yd.shape[1]):
                    assert _rel(Y[:, c], Yd[:, c]) < 1e-4
        assert srv.stats()["failed"] == 0
    finally:
        srv.destroy()


# --------------------------------------------------------------- m = 1 clusters
@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["quirk", "clustered_leaf1", "clustered_leaf2"])
def test_products_with_one_point_clusters(gpu, prec_name, case):
    from tests.test_oracle import _quirk_points
    if case == "quirk":
        # the size-1 cluster shifts the level's boxes: the 399 x 399 diagonal block
        # gets a one-point box and is taken as admissible (the reference's literal
        # behaviour); rank 400 = full rank keeps the product exact
        X = _quirk_points(400)
        c_leaf, eta, k, max_level = 1, 1.0, 400, 50
    else:
        n = 1500
        X = O.uniform(91 if case.endswith("1") else 92, 2 * n).reshape(2, n) ** 3
        c_leaf, eta, k, max_level = (1, 1.2, 32, 50) if case.endswith("1") else (2, 1.5, 32, 50)
    n = X.shape[1]
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    hm = gpu.HMatrix(X, X, 32, eta, max_level, c_leaf, k)
    hm.setup()
    kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.2)
    hm.set_kernel(kern)
    hm.set_precision(prec)
    hm.precompute_aca()
    om = O.HMatrix(X, X, 32, eta, max_level, c_leaf)
    ok = O.kernel(O.K_EXPONENTIAL, 2, 0.0, 0.2)
    b = om.blocks()
    m1 = b["set1_u"] - b["set1_l"] + 1
    m2 = b["set2_u"] - b["set2_l"] + 1
    one = (m1 == 1) | (m2 == 1)
    assert (one & (b["work_type"] == O.WT_DENSE)).any()
    if case != "quirk":
        assert (one & (b["work_type"] == O.WT_ACA)).any()
    g = hm.blocks(O.WORK_ITEM_DTYPE)
    for f in ("set1_l", "set1_u", "set2_l", "set2_u", "work_type"):
        np.testing.assert_array_equal(g[f], b[f])
    om.aca(ok, k)
    Xr = O.normal(93, n * 3).reshape(n, 3, order="F")
    Y = hm.apply(Xr)
    Yo = om.hmvp(ok, Xr)
    for c in range(3):
        yd = om.dense_mvp(ok, Xr[:, c])
        if prec_name == "fp64":
            # clustered points: late rank steps at rounding level, where a last-ulp
            # exp difference can flip a pivot (the stated FP64 bar is 1e-6)
            assert _rel(Y[:, c], Yo[:, c]) < 1e-6
        else:
            mag = np.linalg.norm(om.dense_mvp(ok, np.abs(Xr[:, c])))
            assert np.linalg.norm(Y[:, c] - Yo[:, c]) / mag < 4e-5
        assert _rel(Yo[:, c], yd) < 1e-6
        assert _rel(Y[:, c], yd) < (1e-6 if prec_name == "fp64" else 1e-4)


# ------------------------------------------------------------ exact cluster boxes
@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
def test_exact_boxes_option_against_oracle(gpu, prec_name):
    """HMG_OPT_EXACT_BOXES on 8-blob clouds whose levels hold one-point clusters
    (where the reference's lookup table shifts boxes): the block list, boxes
    included, bit-exact against the oracle's exact-box mode, which the CPU tests pin
    with a brute-force traversal; the product against the oracle's H-MVP and the
    dense product (the literal boxes admit unseparated blocks on these clouds and
    miss the dense product by up to 0.6)."""
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    for seed, n, dim in ((5204, 5204, 2), (4562, 4562, 3)):
        rng = np.random.default_rng(seed)
        X = _blob_cloud(rng, n, dim)
        bits, leaf, eta, k, kind = (32, 16, 1.5, 32, O.K_EXPONENTIAL) if dim == 2 else (20, 32, 1.0, 24, O.K_GAUSSIAN)
        lit = O.HMatrix(X, X, bits, eta, 50, leaf)
        assert lit.diagnostics()["box_differs"] > 0
        om = O.HMatrix(X, X, bits, eta, 50, leaf, exact_boxes=True)
        hm = gpu.HMatrix(X, X, bits, eta, 50, leaf, k)
        hm.set_option(gpu.OPT_EXACT_BOXES, 1)
        hm.setup()
        g, o = hm.blocks(O.WORK_ITEM_DTYPE), om.blocks()
        assert len(g) == len(o)
        for f in ("set1_l", "set1_u", "set2_l", "set2_u", "work_type", "level_1", "min1", "max1", "min2", "max2"):
            np.testing.assert_array_equal(g[f], o[f])
        kern = gpu.Kernel(kind, 0.0, 0.1)
        hm.set_kernel(kern)
        hm.set_precision(prec)
        hm.precompute_aca()
        ok = O.kernel(kind, dim, 0.0, 0.1)
        om.aca(ok, k)
        x = O.normal(seed + 1, n)
        y = hm.apply(x)
        y_o = om.hmvp(ok, x)
        y_d = om.dense_mvp(ok, x)
        assert _rel(y_o, y_d) < 1e-4
        assert _rel(y, y_o) < (1e-6 if prec_name == "fp64" else 1e-4)
        assert _rel(y, y_d) < (1e-4 if prec_name == "fp64" else 2e-4)
        hm.destroy()
        kern.free()
