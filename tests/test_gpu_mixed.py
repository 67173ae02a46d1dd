"""GPU parity of the mixed-precision H-MVP (FP64 ACA with factors stored as BF16
hi + lo panels, low-rank products and the near field as BF16 hi+lo split MFMA
GEMMs accumulated in FP32, near field evaluated in FP32) against the FP64 engine
on the same H-matrix and against the dense FP64 product of the oracle.

Tolerances (relative 2-norm error per right-hand side):
  mixed vs dense FP64: 1e-4, the stated mixed-precision bar;
  mixed vs the FP64 H-MVP of the same tree: 1e-4 (the ACA is the FP64 one, so
    ranks and pivots are identical; the difference is arithmetic);
  arithmetic alone, where truncation cannot matter (every block dense, or ranks
    high enough that the ACA is exact to ~1e-9), or added to the FP64 H-MVP's own
    error: 4e-5.  Each operand is split into BF16 hi + lo (16 significant bits:
    representation error <= 2^-18 relative) and the lo*lo term is dropped
    (<= 2^-18), so every product carries up to ~3 * 2^-18 = 1.1e-5 relative error
    before FP32 accumulation; sums with cancellation (smooth kernels, random
    right-hand sides) show up to ~2.4e-5 relative to |y|.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL_MIXED_ARITH = 4e-5
TOL_MIXED_VS_FP64 = 1e-4
TOL_MIXED_VS_DENSE = 1e-4


def _points(seed, dim, n):
    return O.uniform(seed, dim * n).reshape(dim, n)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _build(gpu, X1, X2, bits, eta, c_leaf, k, kind, ls, reg=0.0, prec=None, precompute=True):
    hm = gpu.HMatrix(X1, X2, bits, eta, 50, c_leaf, k)
    hm.setup()
    kern = gpu.Kernel(kind, reg, ls)
    hm.set_kernel(kern)
    hm.set_precision(gpu.PREC_MIXED if prec is None else prec)
    if precompute:
        hm.precompute_aca()
    return hm, kern


def _check_columns(Ymx, Yref, tol):
    for c in range(Yref.shape[1]):
        e = _rel(Ymx[:, c], Yref[:, c])
        assert e < tol, f"rhs {c}: rel err {e:.3e} >= {tol:.1e}"


def test_mixed_c1_32_rhs_against_fp64_and_dense(gpu):
    """Config 1 points and kernel (N=4096 uniform [0,1]^2, exp(-r^2/0.02), leaf
    64, eta 1.5, rank 32) with 32 right-hand sides U[-1,1]."""
    n = 4096
    X = _points(1001, 2, n)
    R = 32
    Xr = (O.uniform(1003, n * R) * 2.0 - 1.0).reshape(n, R, order="F")
    h64, _ = _build(gpu, X, X, 32, 1.5, 64, 32, O.K_GAUSSIAN, 0.1, prec=gpu.PREC_FP64)
    hmx, _ = _build(gpu, X, X, 32, 1.5, 64, 32, O.K_GAUSSIAN, 0.1)
    Y64 = h64.apply(Xr)
    Ymx = hmx.apply(Xr)
    _check_columns(Ymx, Y64, TOL_MIXED_ARITH)
    om = O.HMatrix(X, X, 32, 1.5, 50, 64)
    ok = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.1)
    for c in (0, 13, 31):
        yd = om.dense_mvp(ok, Xr[:, c])
        assert _rel(Ymx[:, c], yd) < TOL_MIXED_VS_DENSE
    # bitwise deterministic
    np.testing.assert_array_equal(Ymx, hmx.apply(Xr))
    s = hmx.schedule()
    assert s["near_items"] > 0 and s["near_evals"] >= hmx.work()["E_near"]


@pytest.mark.parametrize("nrhs", [1, 5, 40, 65])
def test_mixed_rhs_counts(gpu, nrhs):
    """Right-hand sides run in passes of 32 (zero-padded); every column matches
    the FP64 product and the single-RHS mixed product."""
    n = 6000
    X = _points(2001, 3, n)
    Xr = O.normal(2002, n * nrhs).reshape(n, nrhs, order="F")
    h64, _ = _build(gpu, X, X, 20, 1.0, 128, 16, O.K_EXPONENTIAL, 0.1, prec=gpu.PREC_FP64)
    hmx, _ = _build(gpu, X, X, 20, 1.0, 128, 16, O.K_EXPONENTIAL, 0.1)
    Ymx = hmx.apply(Xr)
    _check_columns(Ymx, h64.apply(Xr), TOL_MIXED_ARITH)
    for c in sorted({0, nrhs - 1, nrhs // 2}):
        y1 = hmx.apply(Xr[:, c])
        assert _rel(Ymx[:, c], y1) < 1e-6


@pytest.mark.parametrize("kind,ls,reg,same", [
    (O.K_EXPONENTIAL, 0.1, 0.0, False),
    (O.K_GAUSSIAN, 0.1, 0.0, False),
    (O.K_GAUSSIAN_REF, 1.0, 0.0, False),
    (O.K_MATERN_REF, 1.0, 0.0, False),
    (O.K_GAUSSIAN, 0.1, 0.5, True),
    (O.K_EXPONENTIAL, 0.1, 0.25, True),
])
def test_mixed_kernels_and_regularization(gpu, kind, ls, reg, same):
    n1, n2 = 3000, (3000 if same else 2200)
    X1 = _points(3001, 2, n1)
    X2 = X1 if same else _points(3002, 2, n2)
    Xr = O.normal(3003, n2 * 8).reshape(n2, 8, order="F")
    h64, _ = _build(gpu, X1, X2, 32, 1.5, 64, 16, kind, ls, reg=reg, prec=gpu.PREC_FP64)
    hmx, _ = _build(gpu, X1, X2, 32, 1.5, 64, 16, kind, ls, reg=reg)
    Ymx = hmx.apply(Xr)
    Y64 = h64.apply(Xr)
    _check_columns(Ymx, Y64, TOL_MIXED_VS_FP64)
    om = O.HMatrix(X1, X2, 32, 1.5, 50, 64)
    ok = O.kernel(kind, 2, reg, ls)
    for c in (0, 7):
        yd = om.dense_mvp(ok, Xr[:, c])
        # rank 16 leaves an H-approximation error of up to ~1e-4 here in FP64 as
        # well: mixed precision may add at most the arithmetic tolerance to it
        assert _rel(Ymx[:, c], yd) < _rel(Y64[:, c], yd) + TOL_MIXED_ARITH


@pytest.mark.parametrize("kind,ls,reg", [(O.K_EXPONENTIAL, 0.1, 0.0), (O.K_GAUSSIAN, 0.1, 0.3),
                                         (O.K_MATERN_REF, 1.0, 0.0)])
def test_mixed_near_field_arithmetic(gpu, kind, ls, reg):
    """Every block dense (leaf size above N): the product is the near field alone,
    evaluated in FP32 and multiplied as BF16 hi+lo MFMAs.  The error of a sum of
    rounded terms is bounded relative to sum |K_ij| |x_j|, not to |y| (the
    reference Matern kernel is almost constant on the unit cube, so its y cancels
    heavily); every kernel here is single-signed on these inputs (Y_1 < 0 below
    r = 2.19), so |K| |x| = |K |x||."""
    n1, n2 = 2500, 2300
    X1 = _points(6001, 3, n1)
    X2 = X1 if reg else _points(6002, 3, n2)
    n2 = X2.shape[1]
    Xr = O.normal(6003, n2 * 32).reshape(n2, 32, order="F")
    hmx, _ = _build(gpu, X1, X2, 20, 1.0, 4096, 8, kind, ls, reg=reg)
    assert hmx.work()["aca"] == 0
    Ymx = hmx.apply(Xr)
    om = O.HMatrix(X1, X2, 20, 1.0, 50, 4096)
    ok = O.kernel(kind, 3, reg, ls)
    for c in (0, 11, 31):
        yd = om.dense_mvp(ok, Xr[:, c])
        mag = np.linalg.norm(om.dense_mvp(ok, np.abs(Xr[:, c])))
        assert np.linalg.norm(Ymx[:, c] - yd) / mag < TOL_MIXED_ARITH


def test_mixed_low_rank_arithmetic(gpu):
    """Ranks high enough that the ACA truncation error is far below FP32 rounding
    (rank 64 on leaves of 32): mixed and FP64 products differ by FP32 arithmetic
    only."""
    n = 4000
    X = _points(7001, 2, n)
    Xr = O.normal(7002, n * 32).reshape(n, 32, order="F")
    h64, _ = _build(gpu, X, X, 32, 1.5, 32, 64, O.K_EXPONENTIAL, 0.1, prec=gpu.PREC_FP64)
    hmx, _ = _build(gpu, X, X, 32, 1.5, 32, 64, O.K_EXPONENTIAL, 0.1)
    assert hmx.work()["aca"] > 0
    _check_columns(hmx.apply(Xr), h64.apply(Xr), TOL_MIXED_ARITH)


def test_mixed_3d_against_dense(gpu):
    """3D exponential kernel (the scale configuration's kernel and parameters at a
    size the oracle's dense product finishes quickly)."""
    n = 8192
    X = _points(2001, 3, n)
    Xr = O.normal(2002, n * 4).reshape(n, 4, order="F")
    hmx, _ = _build(gpu, X, X, 20, 1.0, 256, 24, O.K_EXPONENTIAL, 0.1)
    Ymx = hmx.apply(Xr)
    om = O.HMatrix(X, X, 20, 1.0, 50, 256)
    ok = O.kernel(O.K_EXPONENTIAL, 3, 0.0, 0.1)
    for c in range(4):
        assert _rel(Ymx[:, c], om.dense_mvp(ok, Xr[:, c])) < TOL_MIXED_VS_DENSE


def test_builds_return_their_device_memory(gpu):
    """Building, applying and destroying H-matrices leaves the device's free memory
    where it was: in mixed precision the FP64 ACA runs into a temporary scratch of
    (m1 + m2) k doubles per block (about 9 GB here) that must be returned, also
    when the build is repeated."""
    from hmglib_amd import hip
    n = 60000
    X = _points(9101, 3, n)
    x = O.normal(9103, n)

    def cycle():
        hm, kern = _build(gpu, X, X, 20, 1.0, 256, 24, O.K_EXPONENTIAL, 0.1)
        hm.apply(x)
        hm.destroy()
        kern.free()

    cycle()
    hip.synchronize()
    hip.trim_pool()
    f0 = hip.mem_info()[0]
    for _ in range(3):
        cycle()
    hip.synchronize()
    hip.trim_pool()
    f1 = hip.mem_info()[0]
    # slack for the pool's chunk granularity and the runtime's own per-stream
    # buffers (about 0.1 GiB over three cycles); the scratch alone is ~27 GiB
    assert f1 >= f0 - (512 << 20), f"{(f0 - f1) / 2**30:.2f} GiB not returned"


@pytest.mark.parametrize("prec_name", ["mixed", "fp64"])
def test_autotune_keeps_the_product(gpu, prec_name):
    """The autotuner times the product's launch shapes on the caller's arrays and
    keeps the fastest; the product's values do not depend on the choice."""
    from hmglib_amd import hip
    n = 20000
    X = _points(9201, 3, n)
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    hm, kern = _build(gpu, X, X, 20, 1.0, 256, 16, O.K_EXPONENTIAL, 0.1, prec=prec)
    R = 32 if prec_name == "mixed" else 2
    xd = hip.DeviceArray(np.asfortranarray(O.normal(9203, n * R).reshape(n, R, order="F")))
    yd = hip.DeviceArray.empty((n, R))
    hm.apply_dev(xd, yd, R)
    y0 = yd.download()
    t = hm.autotune(xd, yd, R, reps=1)
    assert t["near_lds_bytes"] in (0, 16384, 24576, 36864, 49152, 65536, 81920) and t["us_per_product"] > 0
    assert not (t["one_stream"] and t["near_lds_bytes"])
    if prec_name == "fp64":
        # the FP64 product has one form of each; the options are left on
        assert t["lr_workgroup"] and t["near_workgroup"]
    hm.apply_dev(xd, yd, R)
    np.testing.assert_array_equal(yd.download(), y0)
    xd.free()
    yd.free()


def test_mixed_not_precomputed_matches_precomputed(gpu):
    n = 5000
    X = _points(4001, 3, n)
    Xr = O.normal(4002, n * 3).reshape(n, 3, order="F")
    hmx, _ = _build(gpu, X, X, 20, 1.0, 128, 12, O.K_EXPONENTIAL, 0.1, precompute=False)
    Y_fly = hmx.apply(Xr)
    hmx.precompute_aca()
    np.testing.assert_array_equal(Y_fly, hmx.apply(Xr))


def test_switching_precision_back_to_fp64(gpu):
    """Stored factors decide a product's precision; rebuilding them in FP64 gives
    exactly the FP64 engine's product again."""
    n = 4000
    X = _points(5001, 2, n)
    x = O.normal(5002, n)
    h64, _ = _build(gpu, X, X, 32, 1.5, 64, 16, O.K_GAUSSIAN, 0.1, prec=gpu.PREC_FP64)
    hmx, _ = _build(gpu, X, X, 32, 1.5, 64, 16, O.K_GAUSSIAN, 0.1)
    y_mx = hmx.apply(x)
    hmx.set_precision(gpu.PREC_FP64)
    # factors are still FP32: the product stays mixed until precompute_aca runs again
    np.testing.assert_array_equal(hmx.apply(x), y_mx)
    hmx.precompute_aca()
    np.testing.assert_array_equal(hmx.apply(x), h64.apply(x))
    # the unbatched reference path is FP64 and leaves the stored factors usable
    hmx.set_precision(gpu.PREC_MIXED)
    hmx.precompute_aca()
    hmx.apply_without_batching(x)
    np.testing.assert_array_equal(hmx.apply(x), y_mx)


@pytest.mark.parametrize("case", ["c1_2d_gauss", "chunked_2d_eta4", "c3_like_3d_exp"])
def test_lr_workgroup_form_bitwise_equal(gpu, case):
    """The workgroup forms of V^T X, U T and the near field (X / T panel groups
    staged once per workgroup through LDS) run the same arithmetic per output as the
    one-wavefront kernels: the products are bitwise equal.  The eta-4 case has
    column clusters of more than 4096 points (several point chunks per slice:
    partial sums and their reduction)."""
    if case == "c1_2d_gauss":
        n, dim, eta, leaf, k, kind, ls = 4096, 2, 1.5, 64, 32, O.K_GAUSSIAN, 0.1
    elif case == "chunked_2d_eta4":
        n, dim, eta, leaf, k, kind, ls = 50000, 2, 4.0, 64, 16, O.K_EXPONENTIAL, 0.2
    else:
        n, dim, eta, leaf, k, kind, ls = 40000, 3, 1.0, 256, 24, O.K_EXPONENTIAL, 0.1
    X = _points(1201, dim, n)
    hm, kern = _build(gpu, X, X, 32 if dim == 2 else 20, eta, leaf, k, kind, ls)
    if case == "chunked_2d_eta4":
        b = hm.blocks(O.WORK_ITEM_DTYPE)
        aca = b[b["work_type"] == O.WT_ACA]
        assert (aca["set2_u"] - aca["set2_l"] + 1).max() > 4096
    R = 32
    Xr = O.normal(1203, n * R).reshape(n, R, order="F")
    forms = {}
    for lr_wg, near_wg in ((1, 1), (0, 0), (1, 0), (0, 1)):
        hm.set_option(gpu.OPT_LR_WORKGROUP, lr_wg)
        hm.set_option(gpu.OPT_NEAR_WORKGROUP, near_wg)
        forms[(lr_wg, near_wg)] = hm.apply(Xr)
    Y_wg = forms[(1, 1)]
    for Y in forms.values():
        np.testing.assert_array_equal(Y, Y_wg)
    # and both are the H-matrix's product (FP64 H-MVP of the same tree and ranks)
    hm64, k64 = _build(gpu, X, X, 32 if dim == 2 else 20, eta, leaf, k, kind, ls, prec=gpu.PREC_FP64)
    _check_columns(Y_wg, hm64.apply(Xr), TOL_MIXED_VS_FP64)
    hm.destroy()
    kern.free()
    hm64.destroy()
    k64.free()


def test_tuning_options_leave_results_unchanged(gpu):
    """The low-rank issue priority, the batch count of the mixed ACA and the ACA's
    LDS reservation change launch shapes only: ranks, factors and products stay
    bitwise the same (the mixed ACA's batches fill a bounded FP64 scratch in order
    of decreasing block size; each block's ACA is the same arithmetic in any
    batch)."""
    from hmglib_amd import hip
    n = 30000
    X = _points(1301, 3, n)
    R = 32
    Xr = O.normal(1303, n * R).reshape(n, R, order="F")
    ref = None
    for parts, lds, prio in ((16, 0, 0), (1, 0, 1), (4, 49152, 0), (64, 24576, 1)):
        hm = gpu.HMatrix(X, X, 20, 1.0, 50, 256, 24)
        hm.setup()
        kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.1)
        hm.set_kernel(kern)
        hm.set_precision(gpu.PREC_MIXED)
        hm.set_option(gpu.OPT_MIXED_ACA_PARTS, parts)
        hm.set_option(gpu.OPT_ACA_LDS_BYTES, lds)
        hm.set_option(gpu.OPT_LR_PRIORITY, prio)
        hm.precompute_aca()
        out = (hm.work()["rank_sum"], hm.apply(Xr))
        if ref is None:
            ref = out
        else:
            assert out[0] == ref[0]
            np.testing.assert_array_equal(out[1], ref[1])
        hm.destroy()
        kern.free()
    with pytest.raises(gpu.HMGError):
        hm = gpu.HMatrix(X, X, 20, 1.0, 50, 256, 24)
        try:
            hm.set_option(gpu.OPT_MIXED_ACA_PARTS, 0)
        finally:
            hm.destroy()
    # the library's cached large buffers go back to the device on request
    hip.synchronize()
    hip.trim_pool()
    f = hip.mem_info()[0]
    assert f > 0


def test_mixed_aca_in_many_batches_same_as_one(gpu):
    """At N = 2^17 (3D, leaf 256, eta 1.0, rank 24: 1.2e9 FP64 factor elements) the
    mixed build's FP64 ACA runs in 10 batches of at least 2^27 elements by default
    and in one batch with HMG_OPT_MIXED_ACA_PARTS = 1: ranks and the 32-RHS
    products are bitwise the same, and both agree with the FP64 H-matrix to the
    BF16x3 precision."""
    n = 1 << 17
    X = O.uniform(2001, 3 * n).reshape(3, n)
    R = 32
    Xr = O.normal(1403, n * R).reshape(n, R, order="F")
    out = {}
    for parts in (16, 1, 0):
        hm = gpu.HMatrix(X, X, 20, 1.0, 50, 256, 24)
        hm.setup()
        kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.1)
        hm.set_kernel(kern)
        if parts:
            hm.set_precision(gpu.PREC_MIXED)
            hm.set_option(gpu.OPT_MIXED_ACA_PARTS, parts)
        hm.precompute_aca()
        out[parts] = (hm.work()["rank_sum"], hm.apply(Xr))
        hm.destroy()
        kern.free()
    assert out[16][0] == out[1][0] == out[0][0]
    np.testing.assert_array_equal(out[16][1], out[1][1])
    for c in range(R):
        assert _rel(out[16][1][:, c], out[0][1][:, c]) < 4e-5


def test_released_large_buffers_are_reused_and_returned(gpu):
    """Factor buffers of 8 GiB and more stay mapped after their H-matrix is
    destroyed, are reused by the next build, and hmg_release_cached_memory returns
    them: the free device memory after the release is back where it started."""
    from hmglib_amd import hip
    n = 1 << 18
    X = _points(1401, 3, n)
    hip.synchronize()
    hip.trim_pool()
    f0 = hip.mem_info()[0]
    ys = []
    for _ in range(2):
        # FP64 factors of this tree at rank 48: U and V above 8 GiB each
        hm = gpu.HMatrix(X, X, 20, 1.0, 50, 256, 48)
        hm.setup()
        kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.1)
        hm.set_kernel(kern)
        hm.precompute_aca()
        assert sum(hm.memory()) > 0
        ys.append(hm.apply(O.normal(1403, n)))
        hm.destroy()
        kern.free()
    np.testing.assert_array_equal(ys[0], ys[1])
    hip.synchronize()
    f_held = hip.mem_info()[0]
    hip.trim_pool()
    f1 = hip.mem_info()[0]
    assert f1 >= f0 - (512 << 20), f"{(f0 - f1) / 2**30:.2f} GiB not returned"
    assert f1 >= f_held
