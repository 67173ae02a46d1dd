"""GPU tests of the many-cloud path (hmg_setup_batch, hmg_precompute_aca_batch,
hmg_apply_batch): clouds of the batching workload's kind (8-blob Gaussian
mixtures, 2D and 3D, Gaussian l = 0.1, leaf 64, eta 1.2, relative ACA tolerance
1e-5 with rank cap 32, 1-8 right-hand sides) whose admissible blocks all run in
one ACA launch set.  Ranks and products must be bitwise those of each H-matrix
built alone (same kernels, same per-block arithmetic), and each cloud's product
is checked against the oracle's tolerance ACA + H-MVP."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _clouds(n_clouds, seed0=4101, lo=10, span=3):
    out = []
    for c in range(n_clouds):
        rng = np.random.default_rng(seed0 + c)
        n = int(2 ** (lo + span * rng.random()))
        dim = 2 if c % 2 == 0 else 3
        centres = rng.random((8, dim))
        lab = rng.integers(0, 8, n)
        X = np.clip(centres[lab].T + 0.05 * rng.standard_normal((dim, n)), 0.0, 1.0)
        nrhs = int(rng.integers(1, 9))
        out.append((X, dim, n, nrhs, np.asfortranarray(rng.standard_normal((n, nrhs)))))
    return out


def _hm(gpu, X, dim, prec, kind=O.K_GAUSSIAN, ls=0.1, eps=1e-5):
    hm = gpu.HMatrix(X, X, 32 if dim == 2 else 20, 1.2, 50, 64, 32, epsilon=eps)
    hm.set_option(gpu.OPT_ACA_TOLERANCE, 2)
    hm.set_precision(prec)
    kern = gpu.Kernel(kind, 0.0, ls)
    return hm, kern


@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
def test_batched_aca_same_as_alone_and_oracle(gpu, prec_name):
    from hmglib_amd import hip
    from hmglib_amd.hmatrix import apply_batch, precompute_aca_batch, setup_batch
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    eps = 1e-5
    clouds = _clouds(12)
    hms, kerns = [], []
    for X, dim, n, nrhs, Xh in clouds:
        hm, kern = _hm(gpu, X, dim, prec)
        hms.append(hm)
        kerns.append(kern)
    # one cloud with another kernel and precision: a group of its own in the same call
    Xo, dimo, no, nro, Xho = _clouds(1, seed0=4301)[0]
    other, kother = _hm(gpu, Xo, dimo, gpu.PREC_MIXED if prec_name == "fp64" else gpu.PREC_FP64,
                        kind=O.K_EXPONENTIAL, ls=0.2)
    hms.append(other)
    kerns.append(kother)
    clouds.append((Xo, dimo, no, nro, Xho))
    setup_batch(hms, threads=6)
    for hm, kern in zip(hms, kerns):
        hm.set_kernel(kern)
    precompute_aca_batch(hms, threads=6)
    xds = [hip.DeviceArray(c[4]) for c in clouds]
    yds = [hip.DeviceArray.empty((c[2], c[3])) for c in clouds]
    apply_batch(hms, xds, yds, [c[3] for c in clouds], threads=6)
    for i, ((X, dim, n, nrhs, Xh), hm) in enumerate(zip(clouds, hms)):
        Y = yds[i].download().reshape(n, nrhs)
        # alone: the same H-matrix through precompute_aca
        kind, ls = (O.K_EXPONENTIAL, 0.2) if i == len(hms) - 1 else (O.K_GAUSSIAN, 0.1)
        p_i = (gpu.PREC_MIXED if prec_name == "fp64" else gpu.PREC_FP64) if i == len(hms) - 1 else prec
        ha, ka = _hm(gpu, X, dim, p_i, kind=kind, ls=ls)
        ha.setup()
        ha.set_kernel(ka)
        ha.precompute_aca()
        assert ha.work() == hm.work()
        np.testing.assert_array_equal(Y, ha.apply(Xh).reshape(n, nrhs))
        ha.destroy()
        # the oracle
        om = O.HMatrix(X, X, 32 if dim == 2 else 20, 1.2, 50, 64)
        ok = O.kernel(kind, dim, 0.0, ls)
        om.aca(ok, 32, tol_mode=2, eps=eps)
        Yo = om.hmvp(ok, Xh).reshape(n, nrhs)
        Yd = om.dense_mvp(ok, Xh).reshape(n, nrhs)
        for c in range(nrhs):
            if p_i == gpu.PREC_FP64:
                # rank decisions at the tolerance threshold may differ in a few
                # blocks (exp last ulps against glibc)
                assert _rel(Y[:, c], Yo[:, c]) < 10 * eps
            else:
                assert _rel(Y[:, c], Yd[:, c]) < _rel(Yo[:, c], Yd[:, c]) + 4e-5
    for hm in hms:
        hm.destroy()


def test_batch_rejects_repeated_and_unready(gpu):
    from hmglib_amd.hmatrix import precompute_aca_batch
    X, dim, n, nrhs, Xh = _clouds(1)[0]
    hm, kern = _hm(gpu, X, dim, gpu.PREC_FP64)
    with pytest.raises(gpu.HMGError, match="not set up"):
        precompute_aca_batch([hm])
    hm.setup()
    hm.set_kernel(kern)
    with pytest.raises(gpu.HMGError, match="listed twice"):
        precompute_aca_batch([hm, hm])
    precompute_aca_batch([hm])
    assert hm.work()["rank_sum"] > 0
    hm.destroy()


def test_fused_setup_same_as_alone(gpu):
    """hmg_setup_batch runs the setup of clouds that share a dimension and tree
    parameters (one point set against itself, exact cluster boxes) as one set of
    launches per step: all clouds' coordinates side by side, extrema, Morton codes,
    one sort, every tree in one level-by-level traversal, one partition.  Each
    H-matrix must come out bitwise as setup_h_matrix leaves it alone: codes, order,
    point ids, sorted coordinates, extrema of both sets, the complete block list,
    its counts and levels, and the products built on it.  Clouds outside a group
    (independent point sets, the reference's lookup-table boxes) take the
    one-at-a-time path in the same call."""
    from hmglib_amd.hmatrix import setup_batch
    clouds = _clouds(24, seed0=4501, lo=9, span=5)
    made = []
    for i, (X, dim, n, nrhs, Xh) in enumerate(clouds):
        pair = []
        for _ in range(2):
            hm, kern = _hm(gpu, X, dim, gpu.PREC_FP64)
            hm.set_option(gpu.OPT_EXACT_BOXES, 1)
            pair.append((hm, kern))
        made.append(pair)
    # outside the groups: two independent sets, and lookup-table boxes
    rng = np.random.default_rng(4599)
    Xa, Xb = rng.random((2, 3000)), rng.random((2, 2500))
    ind = [gpu.HMatrix(Xa, Xb, 32, 1.2, 50, 64, 32) for _ in range(2)]
    for h in ind:
        h.set_option(gpu.OPT_EXACT_BOXES, 1)
    lut = [_hm(gpu, clouds[0][0], clouds[0][1], gpu.PREC_FP64)[0] for _ in range(2)]
    batch = [p[0][0] for p in made] + [ind[0], lut[0]]
    setup_batch(batch, 8)
    for p in made:
        p[1][0].setup()
    ind[1].setup()
    lut[1].setup()
    pairs = [(p[0][0], p[1][0]) for p in made] + [(ind[0], ind[1]), (lut[0], lut[1])]
    for hb, hs in pairs:
        for s in (0, 1):
            np.testing.assert_array_equal(hb.codes(s), hs.codes(s))
            np.testing.assert_array_equal(hb.order(s), hs.order(s))
            np.testing.assert_array_equal(hb.point_ids(s), hs.point_ids(s))
            np.testing.assert_array_equal(hb.sorted_coords(s), hs.sorted_coords(s))
            for a, b in zip(hb.minmax(s), hs.minmax(s)):
                np.testing.assert_array_equal(a, b)
        bb, bs = hb.blocks(O.WORK_ITEM_DTYPE), hs.blocks(O.WORK_ITEM_DTYPE)
        assert bb.tobytes() == bs.tobytes()
        assert hb.work() == hs.work()
        assert hb.schedule() == hs.schedule()
    # products on the fused setup: bitwise those of the H-matrix set up alone
    for i in (0, 1, 7, 23):
        X, dim, n, nrhs, Xh = clouds[i]
        (hb, kb), (hs, ks) = made[i]
        for hm, kern in ((hb, kb), (hs, ks)):
            hm.set_kernel(kern)
            hm.precompute_aca()
        np.testing.assert_array_equal(hb.apply(Xh), hs.apply(Xh))
    for p in made:
        for hm, kern in p:
            hm.destroy()
            kern.free()
    for h in ind + lut:
        h.destroy()


def test_batch_graph_replays_equal_eager_products(gpu):
    """The products of many H-matrices captured into one hipGraph (FP64 one- and
    several-right-hand-side paths and mixed precision in the same graph) replay
    bitwise the eager products, launch after launch, with new right-hand sides
    and other eager work on the same H-matrices between the launches; a launch
    after one H-matrix's factors were rebuilt is refused."""
    from hmglib_amd import hip
    from hmglib_amd.hmatrix import BatchGraph, precompute_aca_batch, setup_batch
    clouds = _clouds(9, seed0=4701)
    hms, kerns = [], []
    for i, (X, dim, n, nrhs, Xh) in enumerate(clouds):
        hm, kern = _hm(gpu, X, dim, gpu.PREC_MIXED if i % 3 == 2 else gpu.PREC_FP64)
        hm.set_option(gpu.OPT_EXACT_BOXES, 1)
        hms.append(hm)
        kerns.append(kern)
    setup_batch(hms, 4)
    for hm, kern in zip(hms, kerns):
        hm.set_kernel(kern)
    precompute_aca_batch(hms, 4)
    nrhs = [1 if i % 3 == 0 else c[3] for i, c in enumerate(clouds)]
    xds = [hip.DeviceArray(np.asfortranarray(c[4][:, :r])) for c, r in zip(clouds, nrhs)]
    yds = [hip.DeviceArray.empty((c[2], r)) for c, r in zip(clouds, nrhs)]
    g = BatchGraph(hms, xds, yds, nrhs)
    rng = np.random.default_rng(4799)
    try:
        for rep in range(3):
            Xs = [np.asfortranarray(rng.standard_normal((c[2], r))) for c, r in zip(clouds, nrhs)]
            for xd, Xh in zip(xds, Xs):
                xd.upload(Xh)
            # eager products on the same H-matrices between the launches
            for hm, Xh in zip(hms[::2], Xs[::2]):
                hm.apply(Xh)
            g.launch()
            for hm, yd, Xh, r in zip(hms, yds, Xs, nrhs):
                Y = yd.download().reshape(-1, r, order="F")
                np.testing.assert_array_equal(Y, hm.apply(Xh).reshape(-1, r, order="F"))
        hms[1].precompute_aca()
        with pytest.raises(gpu.HMGError):
            g.launch()
    finally:
        g.destroy()
        for a in xds + yds:
            a.free()
        for hm, kern in zip(hms, kerns):
            hm.destroy()
            kern.free()
