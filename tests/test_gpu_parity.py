"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Integer and tree stages are compared bit for bit; products within the stated
relative tolerances.  Inputs come from the seeded generator (SplitMix64 ->
xoshiro256**) with the seeds of the configurations they are drawn from.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# FP64 products: engine vs oracle.  Both evaluate the same blocks with the same
# ranks; exp/sqrt last-ulp and summation-order differences can flip a column
# pivot in late ACA steps, where residuals are dominated by rounding, which moves
# the product by about the ACA truncation error of that step (~1e-10 here).  The
# stated FP64 bar is 1e-6; the test holds the engine to 1e-8.
TOL_FP64 = 1e-8


def _points(seed, dim, n):
    return O.uniform(seed, dim * n).reshape(dim, n)


def _pair(gpu, X1, X2, bits, eta, c_leaf, k, kind, ls=1.0, reg=0.0, max_level=50, precompute=True,
          root_levels=(0, 0)):
    hm = gpu.HMatrix(X1, X2, bits, eta, max_level, c_leaf, k, root_levels=root_levels)
    hm.setup()
    kern = gpu.Kernel(kind, reg, ls)
    hm.set_kernel(kern)
    if precompute:
        hm.precompute_aca()
    om = O.HMatrix(X1, X2, bits, eta, max_level, c_leaf, *root_levels)
    ok = O.kernel(kind, X1.shape[0], reg, ls)
    return hm, kern, om, ok


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


BLOCK_FIELDS = ["set1_l", "set1_u", "set2_l", "set2_u", "work_type", "dim", "level_1", "level_2"]


def _assert_blocks_equal(g, o, dim):
    assert g.shape == o.shape
    for f in BLOCK_FIELDS:
        np.testing.assert_array_equal(g[f], o[f], err_msg=f)
    for f in ("max1", "min1", "max2", "min2"):
        # bit-exact boxes (compare the raw IEEE bits of the first dim entries)
        np.testing.assert_array_equal(g[f][:, :dim].view(np.uint64), o[f][:, :dim].view(np.uint64),
                                      err_msg=f)


@pytest.mark.parametrize("dim,bits,n1,n2,seed,second", [(2, 32, 4096, 4096, 1001, "indep"),
                                                        (3, 20, 5000, 3000, 2001, "indep"),
                                                        (2, 32, 4096, 4096, 1003, "same"),
                                                        (3, 20, 3000, 3000, 2003, "same"),
                                                        (3, 20, 3000, 3000, 2005, "one_differs")])
def test_morton_codes_order_bit_exact(gpu, dim, bits, n1, n2, seed, second):
    """Extrema, codes, order and sorted coordinates of both sets, bit-exact.  With
    both sets bitwise equal the second set's results are copies of the first's;
    one differing coordinate must send the second set through its own sort."""
    X1 = _points(seed, dim, n1)
    if second == "indep":
        X2 = _points(seed + 1, dim, n2)
    else:
        X2 = X1.copy()
        if second == "one_differs":
            X2[1, n2 // 2] = np.nextafter(X2[1, n2 // 2], 2.0)
    hm, _, om, _ = _pair(gpu, X1, X2, bits, 1.5, 64, 16, O.K_GAUSSIAN, 0.1, precompute=False)
    for s in (0, 1):
        mn_g, mx_g = hm.minmax(s)
        mn_o, mx_o = om.minmax(s)
        np.testing.assert_array_equal(mn_g, mn_o)
        np.testing.assert_array_equal(mx_g, mx_o)
        np.testing.assert_array_equal(hm.codes(s), om.codes(s))
        np.testing.assert_array_equal(hm.order(s), om.order(s))
        np.testing.assert_array_equal(hm.point_ids(s), om.order(s).astype(np.uint32))
        np.testing.assert_array_equal(hm.sorted_coords(s), om.sorted_coords(s))


@pytest.mark.parametrize("case", ["c1_same", "c1_indep", "3d_exp", "2d_small_leaf", "unequal_root"])
def test_block_list_bit_exact(gpu, case):
    if case == "c1_same":
        X = _points(1001, 2, 4096)
        args = (X, X, 32, 1.5, 64)
    elif case == "c1_indep":
        args = (_points(1001, 2, 4096), _points(1002, 2, 4096), 32, 1.5, 64)
    elif case == "3d_exp":
        args = (_points(2001, 3, 20000), _points(2001, 3, 20000), 20, 1.0, 256)
    elif case == "2d_small_leaf":
        args = (_points(11, 2, 3000), _points(12, 2, 2500), 32, 0.8, 4)
    else:
        args = (_points(13, 2, 2000), _points(14, 2, 2000), 32, 1.5, 32)
    X1, X2, bits, eta, c_leaf = args
    roots = (0, 2) if case == "unequal_root" else (0, 0)
    hm, _, om, _ = _pair(gpu, X1, X2, bits, eta, c_leaf, 16, O.K_GAUSSIAN, 0.1, precompute=False,
                         root_levels=roots)
    g = hm.blocks(O.WORK_ITEM_DTYPE)
    o = om.blocks()
    assert hm.data.mat_vec_info.dense_count == om.ndense
    assert hm.data.mat_vec_info.aca_count == om.naca
    _assert_blocks_equal(g, o, X1.shape[0])
    assert hm.work()["levels"] == om.levels


def test_code_width_without_interleave_gives_zero_codes(gpu):
    """The reference interleaves only (dim, bits) = (2, 32), (3, 20) and (10, 4)
    (morton.cu:21-72); any other width gives every point code 0, so each cluster
    splits at its middle index and no block is admissible.  Engine and oracle agree
    bit for bit, and the product is the dense one."""
    n = 3000
    X = _points(5, 2, n)
    hm, kern, om, ok = _pair(gpu, X, X, 20, 1.5, 64, 16, O.K_GAUSSIAN, 0.1)
    assert not hm.codes(0).any() and not om.codes(0).any()
    assert om.naca == 0 and om.ndense > 1
    _assert_blocks_equal(hm.blocks(O.WORK_ITEM_DTYPE), om.blocks(), 2)
    x = O.normal(6, n)
    assert _rel(hm.apply(x), om.dense_mvp(ok, x)) < 1e-12


@pytest.mark.parametrize("dim,prec", [(1, "fp64"), (4, "fp64"), (5, "fp64"), (1, "mixed"), (4, "mixed"),
                                      (5, "mixed")])
def test_dimensions_without_morton_interleave(gpu, dim, prec):
    """Dimensions 1, 4 and 5 have no interleave in the reference (morton.cu:21-72):
    codes are 0, every block is DENSE, and the products run the kernels' generic
    dimension paths (coordinates in loops over dim).  Bit-exact block list; the
    FP64 product against the oracle's dense product, mixed precision against it
    within the BF16x3 bound (relative to sum_j K_ij |x_j|, positive kernel)."""
    n = 1500
    X = _points(4601 + dim, dim, n)
    hm = gpu.HMatrix(X, X, 12, 1.5, 50, 64, 8)
    hm.setup()
    kern = gpu.Kernel(O.K_EXPONENTIAL, 0.0, 0.3)
    hm.set_kernel(kern)
    hm.set_precision(gpu.PREC_MIXED if prec == "mixed" else gpu.PREC_FP64)
    hm.precompute_aca()
    om = O.HMatrix(X, X, 12, 1.5, 50, 64)
    ok = O.kernel(O.K_EXPONENTIAL, dim, 0.0, 0.3)
    assert not hm.codes(0).any() and om.naca == 0
    _assert_blocks_equal(hm.blocks(O.WORK_ITEM_DTYPE), om.blocks(), dim)
    Xr = O.normal(4611, n * 2).reshape(n, 2, order="F")
    Y = hm.apply(Xr)
    for c in range(2):
        yd = om.dense_mvp(ok, Xr[:, c])
        if prec == "fp64":
            assert _rel(Y[:, c], yd) < 1e-12
        else:
            mag = np.linalg.norm(om.dense_mvp(ok, np.abs(Xr[:, c])))
            assert np.linalg.norm(Y[:, c] - yd) / mag < 4e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tiny_sets_single_dense_block(gpu, prec):
    """Fewer points than a leaf (5 x 9): one DENSE block, well inside every
    padding the kernels use (8-point groups, 32-row tiles, 128-row items)."""
    X1, X2 = _points(7, 2, 5), _points(8, 2, 9)
    hm = gpu.HMatrix(X1, X2, 32, 1.5, 50, 64, 8)
    hm.setup()
    kern = gpu.Kernel(O.K_GAUSSIAN, 0.0, 0.5)
    hm.set_kernel(kern)
    hm.set_precision(gpu.PREC_MIXED if prec == "mixed" else gpu.PREC_FP64)
    hm.precompute_aca()
    b = hm.blocks(O.WORK_ITEM_DTYPE)
    assert len(b) == 1 and (b["set1_l"][0], b["set1_u"][0], b["set2_l"][0], b["set2_u"][0]) == (0, 4, 0, 8)
    om = O.HMatrix(X1, X2, 32, 1.5, 50, 64)
    ok = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.5)
    X = O.normal(9, 9 * 3).reshape(9, 3, order="F")
    Y = hm.apply(X)
    for c in range(3):
        yd = om.dense_mvp(ok, X[:, c])
        if prec == "fp64":
            assert _rel(Y[:, c], yd) < 1e-14
        else:
            # positive entries: the BF16x3 bound is relative to sum_j K_ij |x_j|
            mag = np.linalg.norm(om.dense_mvp(ok, np.abs(X[:, c])))
            assert np.linalg.norm(Y[:, c] - yd) / mag < 4e-5


@pytest.mark.parametrize("case", ["single_point", "flat_dimension"])
def test_degenerate_extent_is_refused(gpu, case):
    """A point set with max == min in some dimension (one point, or a constant
    coordinate) would divide by zero in the Morton quantisation (morton.cu:119-123);
    setup refuses it with HMG_ERR_INVALID and names the dimension."""
    if case == "single_point":
        X = _points(10, 3, 1)
    else:
        X = _points(11, 2, 100)
        X[1, :] = 0.25
    hm = gpu.HMatrix(X, X, 20 if X.shape[0] == 3 else 32, 1.5, 50, 64, 8)
    with pytest.raises(gpu.HMGError, match="no extent in dimension") as e:
        hm.setup()
    assert e.value.status == gpu.HMG_ERR_INVALID


def test_c1_hmvp_against_oracle_and_dense(gpu):
    """Config 1: N=4096 uniform [0,1]^2, Gaussian exp(-r^2/0.02), leaf 64, eta 1.5, k 32."""
    X = _points(1001, 2, 4096)
    hm, _, om, ok = _pair(gpu, X, X, 32, 1.5, 64, 32, O.K_GAUSSIAN, 0.1)
    x = O.uniform(1003, 4096) * 2.0 - 1.0
    y = hm.apply(x)
    om.aca(ok, 32)
    y_o = om.hmvp(ok, x)
    assert _rel(y, y_o) < TOL_FP64
    y_full = hm.apply_full(x)
    y_dense = om.dense_mvp(ok, x)
    assert _rel(y_full, y_dense) < 1e-13
    # H-approximation error of the configuration (rank 32 fixed), same for both
    err = _rel(y, y_dense)
    assert abs(err - _rel(y_o, y_dense)) < 1e-9
    assert err < 1e-4
    w = hm.work()
    e_near, e_lr = om.work()
    assert w["E_near"] == e_near and w["E_lr"] == e_lr


@pytest.mark.parametrize("kind,ls", [(O.K_GAUSSIAN_REF, 1.0), (O.K_MATERN_REF, 1.0),
                                     (O.K_EXPONENTIAL, 0.1), (O.K_GAUSSIAN, 0.1)])
def test_kernels_independent_sets(gpu, kind, ls):
    # regularization is added where SORTED indices coincide (kernel_system_assembler.h:47);
    # with two different point sets that puts spikes inside admissible blocks, so
    # independent sets are tested without it (as every reference driver does)
    X1 = _points(1001, 2, 3000)
    X2 = _points(1002, 2, 2000)
    hm, _, om, ok = _pair(gpu, X1, X2, 32, 1.5, 64, 16, kind, ls, reg=0.0)
    x = O.normal(1003, 2000)
    y = hm.apply(x)
    om.aca(ok, 16)
    assert _rel(y, om.hmvp(ok, x)) < TOL_FP64
    assert _rel(hm.apply_full(x), om.dense_mvp(ok, x)) < 1e-12


def test_precomputed_dense_matches_on_the_fly(gpu):
    """precompute_dense (hmglib.cu:239-268, linear_algebra.cu:3140-3342): every
    stored DENSE block, column-major m1 x m2, contiguous in block order, against
    the kernel evaluated in numpy on the oracle's sorted points; the product with
    the stored near field against the oracle's H-MVP and the on-the-fly product."""
    from hmglib_amd.hip import memcpy_dtoh
    X = _points(1001, 2, 4096)
    hm, _, om, ok = _pair(gpu, X, X, 32, 1.5, 64, 32, O.K_GAUSSIAN, 0.1, reg=0.3)
    x = O.uniform(1003, 4096) * 2.0 - 1.0
    y_fly = hm.apply(x)
    hm.precompute_dense()
    y_st = hm.apply(x)
    assert _rel(y_st, y_fly) < 1e-14
    om.aca(ok, 32)
    assert _rel(y_st, om.hmvp(ok, x)) < TOL_FP64
    b = om.blocks()[:om.ndense]
    m1 = (b["set1_u"] - b["set1_l"] + 1).astype(np.int64)
    m2 = (b["set2_u"] - b["set2_l"] + 1).astype(np.int64)
    A = np.empty(int((m1 * m2).sum()))
    memcpy_dtoh(A, hm.data.dA)
    Xs = om.sorted_coords(0)
    off = 0
    for q in range(om.ndense):
        l1, l2 = int(b["set1_l"][q]), int(b["set2_l"][q])
        i = np.arange(l1, l1 + m1[q])
        j = np.arange(l2, l2 + m2[q])
        r2 = ((Xs[:, i][:, :, None] - Xs[:, j][:, None, :]) ** 2).sum(axis=0)
        K = np.exp(-r2 / 0.02) + 0.3 * (i[:, None] == j[None, :])
        Ab = A[off:off + m1[q] * m2[q]].reshape(m2[q], m1[q]).T   # column-major m1 x m2
        np.testing.assert_allclose(Ab, K, rtol=1e-14, atol=1e-300)
        off += m1[q] * m2[q]


def test_not_precomputed_aca_matches_precomputed(gpu):
    X = _points(21, 3, 6000)
    hm, _, om, ok = _pair(gpu, X, X, 20, 1.0, 128, 12, O.K_EXPONENTIAL, 0.1, precompute=False)
    x = O.normal(22, 6000)
    y_fly = hm.apply(x)
    hm.precompute_aca()
    y_pre = hm.apply(x)
    np.testing.assert_array_equal(y_fly, y_pre)
    # the ACA recomputed inside the product (linear_algebra.cu:4359-4370) against
    # the oracle's fixed-rank ACA + H-MVP
    om.aca(ok, 12)
    assert _rel(y_fly, om.hmvp(ok, x)) < TOL_FP64


def test_without_batching_first_argmax(gpu):
    X1 = _points(31, 2, 3000)
    X2 = _points(32, 2, 3000)
    hm, _, om, ok = _pair(gpu, X1, X2, 32, 1.5, 64, 16, O.K_GAUSSIAN, 0.1)
    x = O.normal(33, 3000)
    y = hm.apply_without_batching(x)
    om.aca(ok, 16, tie_last=False)
    assert _rel(y, om.hmvp(ok, x)) < TOL_FP64
    # the stored (batched) factors are still in place afterwards
    om.aca(ok, 16, tie_last=True)
    assert _rel(hm.apply(x), om.hmvp(ok, x)) < TOL_FP64


def test_gaussian_rhs_closed_form(gpu):
    X = _points(41, 3, 5000)
    hm, _, om, _ = _pair(gpu, X, X, 20, 1.0, 64, 8, O.K_GAUSSIAN_REF, precompute=False)
    b = hm.gaussian_rhs()
    np.testing.assert_allclose(b, om.gaussian_rhs(), rtol=1e-14, atol=0)


def test_multi_rhs_and_determinism(gpu):
    X = _points(51, 3, 8000)
    hm, _, om, ok = _pair(gpu, X, X, 20, 1.0, 128, 16, O.K_EXPONENTIAL, 0.1)
    Xr = O.normal(52, 8000 * 4).reshape(8000, 4, order="F")
    Y = hm.apply(Xr)
    Y2 = hm.apply(Xr)
    np.testing.assert_array_equal(Y, Y2)
    om.aca(ok, 16)
    Yo = om.hmvp(ok, Xr)
    for c in range(4):
        assert _rel(Y[:, c], Yo[:, c]) < TOL_FP64
        # one right-hand side at a time (symmetric near field): the same to rounding
        assert _rel(hm.apply(Xr[:, c]), Y[:, c]) < 1e-12
    # the block path gives each column bitwise the same result in every pass width:
    # alone (a pass of width 4), in a pass of 4 and in passes of 16 and 32
    hm.set_option(gpu.OPT_FP64_BLOCK_RHS, 1)
    X40 = O.normal(53, 8000 * 40).reshape(8000, 40, order="F")
    X40[:, 7:11] = Xr
    Y40 = hm.apply(X40)
    Y12 = hm.apply(X40[:, :12])
    for c in range(4):
        np.testing.assert_array_equal(hm.apply(Xr[:, c]), Y[:, c])
        np.testing.assert_array_equal(Y40[:, 7 + c], Y[:, c])
        np.testing.assert_array_equal(Y12[:, 7 + c], Y[:, c])
    np.testing.assert_array_equal(Y40[:, 39], hm.apply(X40[:, 39]))


def test_dense_rows_sampled(gpu):
    X = _points(61, 3, 7000)
    hm, _, om, ok = _pair(gpu, X, X, 20, 1.0, 128, 8, O.K_EXPONENTIAL, 0.1, precompute=False)
    x = O.normal(62, 7000)
    rows = np.array([0, 5, 17, 6999, 1234, 4000], dtype=np.int64)
    np.testing.assert_allclose(hm.dense_rows(rows, x), om.dense_rows(ok, rows, x), rtol=1e-12)


@pytest.mark.parametrize("kind", [O.K_GAUSSIAN_REF, O.K_EXPONENTIAL])
def test_regularization_same_sets(gpu, kind):
    X = _points(71, 2, 3000)
    hm, _, om, ok = _pair(gpu, X, X, 32, 1.5, 64, 16, kind, 0.1, reg=0.75)
    x = O.normal(72, 3000)
    om.aca(ok, 16)
    assert _rel(hm.apply(x), om.hmvp(ok, x)) < TOL_FP64
    assert _rel(hm.apply_full(x), om.dense_mvp(ok, x)) < 1e-12


def test_block_list_with_singleton_clusters_bit_exact(gpu):
    """Size-1 clusters change the reference's lookup-table boxes (tree.cu:1000-1024);
    the engine reproduces them literally."""
    from tests.test_oracle import _quirk_points
    X = _quirk_points(64)
    hm, _, om, _ = _pair(gpu, X, X, 32, 1.0, 0, 8, O.K_GAUSSIAN, 0.1, max_level=2, precompute=False)
    assert om.diagnostics()["box_differs"] > 0
    _assert_blocks_equal(hm.blocks(O.WORK_ITEM_DTYPE), om.blocks(), 2)
    assert hm.work()["quirk_levels"] >= 1
    # deeper trees with small leaves on clustered data
    rng_pts = O.uniform(91, 2 * 3000).reshape(2, 3000) ** 3
    hm, _, om, _ = _pair(gpu, rng_pts, rng_pts, 32, 1.2, 2, 8, O.K_GAUSSIAN, 0.1, precompute=False)
    _assert_blocks_equal(hm.blocks(O.WORK_ITEM_DTYPE), om.blocks(), 2)


def _setup_sym(gpu, X, c_leaf, k, kind, ls, sym, eta=1.0, bits=20, reg=0.0):
    hm = gpu.HMatrix(X, X, bits, eta, 50, c_leaf, k)
    hm.set_option(gpu.OPT_NEAR_SYMMETRIC, 1 if sym else 0)
    hm.setup()
    hm.set_kernel(gpu.Kernel(kind, reg, ls))
    hm.precompute_aca()
    return hm


@pytest.mark.parametrize("dim,n,c_leaf,kind,ls,reg", [
    (3, 12000, 128, O.K_EXPONENTIAL, 0.1, 0.0),   # clusters wider than a 64-column tile
    (2, 3000, 8, O.K_GAUSSIAN, 0.1, 0.5),         # many small clusters per tile, regularised
    (2, 4096, 64, O.K_MATERN_REF, 1.0, 0.0),
])
def test_symmetric_near_field(gpu, dim, n, c_leaf, kind, ls, reg):
    """Same points on both sides: the near field evaluates each transposed pair of
    DENSE blocks once.  Entries are bitwise those of the plain form, so the two
    products differ only by summation order; both match the oracle."""
    X = _points(81 + dim, dim, n)
    bits = 20 if dim == 3 else 32
    x = O.normal(83, n)
    hs = _setup_sym(gpu, X, c_leaf, 12, kind, ls, True, bits=bits, reg=reg)
    hp = _setup_sym(gpu, X, c_leaf, 12, kind, ls, False, bits=bits, reg=reg)
    ss, sp = hs.schedule(), hp.schedule()
    assert ss["same_sets"] and ss["near_symmetric"] and ss["col_slots"] > 0
    assert sp["same_sets"] and not sp["near_symmetric"] and sp["col_slots"] == 0
    ys, yp = hs.apply(x), hp.apply(x)
    assert _rel(ys, yp) < 1e-13
    np.testing.assert_array_equal(ys, hs.apply(x))  # deterministic
    om = O.HMatrix(X, X, bits, 1.0, 50, c_leaf, 0, 0)
    ok = O.kernel(kind, dim, reg, ls)
    om.aca(ok, 12)
    assert _rel(ys, om.hmvp(ok, x)) < TOL_FP64


def test_same_sets_detection(gpu):
    X1 = _points(91, 2, 2048)
    X2 = X1.copy()
    X2[1, 7] = np.nextafter(X2[1, 7], 2.0)   # one ulp apart: not the same points
    h = gpu.HMatrix(X1, X2, 32, 1.5, 50, 64, 8)
    h.setup()
    s = h.schedule()
    assert not s["same_sets"] and not s["near_symmetric"] and s["col_slots"] == 0
    h2 = gpu.HMatrix(X1, X1.copy(), 32, 1.5, 50, 64, 8)
    h2.setup()
    assert h2.schedule()["same_sets"]


@pytest.mark.parametrize("k,ls,prec", [(32, 1.0, "fp64"), (100, 0.3, "fp64"), (32, 0.05, "fp64"),
                                       (32, 0.05, "mixed")])
def test_aca_row_scan_same_factors(gpu, k, ls, prec):
    """Gaussian kernels where many admissible blocks reach numerical rank zero
    before step k (smooth ones), or have entries below 1e-13 altogether (narrow
    ones: there the bounding-box test skips the block), so the residual-row
    search rejects long runs of rows.  Those are scanned four rows at a time for rank steps up
    to 64 and one at a time beyond (with k = 100 some blocks stop between steps
    64 and 100); the factors must be bitwise those of the one-row search, so
    ranks and products are equal.  Against the oracle (glibc exp, another
    summation order of the row norms) FP64 rank decisions at the 1e-13 threshold
    may differ in a few blocks only."""
    X = _points(4101, 2, 3000)
    om = O.HMatrix(X, X, 32, 1.5, 50, 64, 0, 0)
    ok = O.kernel(O.K_GAUSSIAN, 2, 0.0, ls)
    om.aca(ok, k)
    f = om.aca_factors()
    keff = f["keff"]
    assert keff.sum() < 0.8 * f["kb"].sum()
    if k > 64:
        assert ((keff > 64) & (keff < f["kb"])).any()
    x = O.normal(4103, 3000)
    res = []
    for scan in (0, 1):
        hm, _, _, _ = _pair(gpu, X, X, 32, 1.5, 64, k, O.K_GAUSSIAN, ls, precompute=False)
        hm.set_option(gpu.OPT_ACA_ROW_SCAN, scan)
        hm.set_precision(gpu.PREC_MIXED if prec == "mixed" else gpu.PREC_FP64)
        hm.precompute_aca()
        res.append((hm.work(), hm.apply(x)))
    (w0, y0), (w1, y1) = res
    assert w0["rank_sum"] == w1["rank_sum"] and w0["E_lr"] == w1["E_lr"]
    np.testing.assert_array_equal(y0, y1)
    if prec == "fp64":
        assert abs(w1["rank_sum"] - int(keff.sum())) <= 1e-3 * keff.sum()
        assert _rel(y1, om.hmvp(ok, x)) < TOL_FP64
    else:
        assert w1["rank_sum"] < 0.95 * f["kb"].sum()
        assert _rel(y1, om.hmvp(ok, x)) < 1e-4


def _blobs(seed, dim, n, sigma=0.05):
    """8-blob Gaussian mixture in [0,1]^dim (centres uniform, clipped)."""
    centres = O.uniform(seed, 8 * dim).reshape(8, dim)
    lab = (O.uniform(seed + 1, n) * 8).astype(np.int64)
    noise = O.normal(seed + 2, dim * n).reshape(dim, n)
    return np.clip(centres[lab].T + sigma * noise, 0.0, 1.0)


@pytest.mark.parametrize("kind,ls", [(O.K_GAUSSIAN, 0.1), (O.K_GAUSSIAN, 0.03), (O.K_EXPONENTIAL, 0.02)])
def test_aca_first_step_row_skip_same_factors(gpu, kind, ls):
    """Clustered points (8 blobs): in many admissible blocks the leading rows sit
    far from the column cluster and every entry of theirs is below the row
    threshold.  The first rank step's search skips rows whose distance to the
    columns' bounding box proves that, unevaluated; factors, ranks and products
    must be bitwise those of the one-row search."""
    n = 6000
    X = _blobs(4301, 2, n)
    x = O.normal(4303, n)
    res = []
    for scan in (0, 1):
        hm, _, om, ok = _pair(gpu, X, X, 32, 1.5, 64, 32, kind, ls, precompute=False)
        hm.set_option(gpu.OPT_ACA_ROW_SCAN, scan)
        hm.precompute_aca()
        res.append((hm.work(), hm.apply(x)))
    (w0, y0), (w1, y1) = res
    assert w0["rank_sum"] == w1["rank_sum"] and w0["E_lr"] == w1["E_lr"]
    np.testing.assert_array_equal(y0, y1)
    om.aca(ok, 32)
    keff = om.aca_factors()["keff"]
    assert keff.sum() < 0.98 * om.aca_factors()["kb"].sum()   # searches reject rows
    assert abs(w1["rank_sum"] - int(keff.sum())) <= 1e-3 * keff.sum()
    # the stated FP64 bar: with clustered points many blocks run late rank steps at
    # rounding level, where a last-ulp exp difference flips pivots (3e-8 observed)
    assert _rel(y1, om.hmvp(ok, x)) < 1e-6


@pytest.mark.parametrize("kind,ls", [(O.K_EXPONENTIAL, 0.1), (O.K_GAUSSIAN, 0.2)])
def test_aca_large_blocks_workgroup_path(gpu, kind, ls):
    """N = 40000 in 2D gives admissible blocks with more than 4096 rows plus
    columns, which run on a workgroup of 4 wavefronts each (or, with the 16384
    threshold lowered to 4096, on 8 wavefronts in a launch beside the rest).
    Ranks and product match the oracle, the 8-wavefront run and the
    all-wavefront run (thresholds raised past every block), in FP64 and in mixed
    precision."""
    n = 40000
    X = _points(4201, 2, n)
    hm, _, om, ok = _pair(gpu, X, X, 32, 1.5, 64, 16, kind, ls)
    om.aca(ok, 16)
    f = om.aca_factors()
    blocks = hm.blocks(O.WORK_ITEM_DTYPE)
    aca = blocks[blocks["work_type"] == O.WT_ACA]
    sizes = (aca["set1_u"] - aca["set1_l"] + 1) + (aca["set2_u"] - aca["set2_l"] + 1)
    assert (sizes > 4096).sum() >= 16
    w = hm.work()
    assert w["rank_sum"] == int(f["keff"].sum()) and w["E_lr"] == om.work()[1]
    x = O.normal(4203, n)
    y = hm.apply(x)
    assert _rel(y, om.hmvp(ok, x)) < TOL_FP64
    # every large block on an 8-wavefront workgroup (launched beside the rest), and
    # none (4-wavefront workgroups only)
    for huge in (4096, 1 << 30):
        hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, huge)
        hm.precompute_aca()
        assert hm.work()["rank_sum"] == w["rank_sum"]
        yh = hm.apply(x)
        assert _rel(yh, y) < TOL_FP64
        if huge == 4096:
            # the 8-wavefront teams' 8-row batched search against their one-row search
            hm.set_option(gpu.OPT_ACA_ROW_SCAN, 0)
            hm.precompute_aca()
            np.testing.assert_array_equal(hm.apply(x), yh)
            hm.set_option(gpu.OPT_ACA_ROW_SCAN, 1)
    hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, 16384)
    hm.set_option(gpu.OPT_ACA_BIG_BLOCK, 1 << 30)
    hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, 1 << 30)
    hm.precompute_aca()
    assert _rel(hm.apply(x), y) < TOL_FP64
    hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, 16384)
    hm.set_precision(gpu.PREC_MIXED)
    hm.precompute_aca()
    ym_wave = hm.apply(x)
    hm.set_option(gpu.OPT_ACA_BIG_BLOCK, 4096)
    hm.precompute_aca()
    ym = hm.apply(x)
    assert _rel(ym, ym_wave) < 1e-6
    hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, 4096)   # the mixed build's batches, 8-wavefront teams
    hm.precompute_aca()
    assert _rel(hm.apply(x), ym_wave) < 1e-6
    hm.set_option(gpu.OPT_ACA_HUGE_BLOCK, 16384)
    # mixed precision: the FP64 ACA's factors rounded to BF16 hi + lo panels; the
    # product is held to the FP64 H-matrix's own approximation error against the
    # dense product plus the BF16x3 arithmetic bound
    y_dense = hm.apply_full(x)
    assert _rel(ym, y_dense) < _rel(y, y_dense) + 4e-5



def _tol_pair(gpu, X, eta, k, kind, ls, mode, eps, prec=None):
    hm = gpu.HMatrix(X, X, 32, eta, 50, 64, k, epsilon=eps)
    hm.setup()
    kern = gpu.Kernel(kind, 0.0, ls)
    hm.set_kernel(kern)
    hm.set_option(gpu.OPT_ACA_TOLERANCE, mode)
    if prec is not None:
        hm.set_precision(prec)
    hm.precompute_aca()
    return hm, kern


@pytest.mark.parametrize("eps", [1e-5, 1e-8])
def test_aca_tolerance_relative_mode_against_oracle(gpu, eps):
    """Relative stop (mode 2) on the batching configuration's kernel (Gaussian,
    length scale 0.1, leaf 64, eta 1.2): per-block ranks as the oracle's (sums
    within 0.5%: the criterion compares norms at rounding level only near the
    threshold), products within the tolerance of each other and of the dense
    product; FP64 and mixed precision share the FP64 ACA, hence the ranks."""
    n = 3000
    X = _points(4401, 2, n)
    om = O.HMatrix(X, X, 32, 1.2, 50, 64)
    ok = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.1)
    om.aca(ok, 32, tol_mode=2, eps=eps)
    f = om.aca_factors()
    keff = f["keff"]
    assert keff.sum() < f["kb"].sum()
    hm, kern = _tol_pair(gpu, X, 1.2, 32, O.K_GAUSSIAN, 0.1, 2, eps)
    w = hm.work()
    assert abs(w["rank_sum"] - int(keff.sum())) <= 5e-3 * keff.sum()
    x = O.normal(4403, n)
    y = hm.apply(x)
    yd = om.dense_mvp(ok, x)
    assert _rel(y, om.hmvp(ok, x)) < 10 * eps
    assert _rel(y, yd) < 100 * eps
    hmx, kx = _tol_pair(gpu, X, 1.2, 32, O.K_GAUSSIAN, 0.1, 2, eps, prec=gpu.PREC_MIXED)
    assert hmx.work()["rank_sum"] == w["rank_sum"]


def test_aca_tolerance_reference_mode(gpu):
    """The reference's criterion (mode 1): with eta >= 1 it never fires, so the
    factors are bitwise the fixed-rank ones; with eta = 0.5 it stops at ranks
    1, 6, 11, ... as the oracle does."""
    n = 3000
    X = _points(4501, 2, n)
    x = O.normal(4503, n)
    h0, k0 = _tol_pair(gpu, X, 1.2, 32, O.K_GAUSSIAN, 0.1, 0, 1e-3)
    h1, k1 = _tol_pair(gpu, X, 1.2, 32, O.K_GAUSSIAN, 0.1, 1, 1e-3)
    assert h0.work() == h1.work()
    np.testing.assert_array_equal(h0.apply(x), h1.apply(x))
    om = O.HMatrix(X, X, 32, 0.5, 50, 64)
    ok = O.kernel(O.K_GAUSSIAN, 2, 0.0, 0.1)
    om.aca(ok, 32, tol_mode=1, eps=1e-4, eta=0.5)
    keff = om.aca_factors()["keff"]
    hm, kern = _tol_pair(gpu, X, 0.5, 32, O.K_GAUSSIAN, 0.1, 1, 1e-4)
    w = hm.work()
    assert abs(w["rank_sum"] - int(keff.sum())) <= 5e-3 * keff.sum()
    assert keff.sum() < 0.9 * om.aca_factors()["kb"].sum()
    assert _rel(hm.apply(x), om.hmvp(ok, x)) < 1e-3


@pytest.mark.parametrize("kind,ls,dim", [(O.K_EXPONENTIAL, 0.1, 3), (O.K_GAUSSIAN, 0.1, 2),
                                         (O.K_EXPONENTIAL, 0.01, 2), (O.K_GAUSSIAN, 0.015, 3)])
def test_product_entries_within_two_ulps_of_glibc(gpu, kind, ls, dim):
    """The FP64 products evaluate Gaussian and exponential entries with their own exp
    (non-positive arguments) and sqrt (non-negative): the dense product of a one-hot
    vector is one column of K with no summation, so each entry must be within 2 ulps
    of the oracle's (glibc exp and sqrt), underflow and the diagonal's exact 1
    included."""
    n = 3000
    X = _points(4901 + dim, dim, n)
    hm, kern, om, ok = _pair(gpu, X, X, 20 if dim == 3 else 32, 1.5, 64, 8, kind, ls=ls, precompute=False)
    for j in (0, 1234, n - 1):
        x = np.zeros(n)
        x[j] = 1.0
        y = hm.apply_full(x)
        yo = om.dense_mvp(ok, x)
        assert yo[j] == 1.0 and y[j] == 1.0
        ulps = np.abs(y - yo) / np.spacing(np.abs(yo))
        assert float(np.max(ulps)) <= 2.0, float(np.max(ulps))
        if kind == O.K_GAUSSIAN and ls < 0.1:
            assert (yo == 0.0).any()  # arguments below -745: both give exactly 0
