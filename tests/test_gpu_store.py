"""GPU tests of the on-disk / wire format of built H-matrices (hmg_save /
hmg_load): a loaded H-matrix rebuilds the identical block list and layout and
its products are bitwise those of the saved one, in FP64 and mixed precision;
a corrupted payload is refused by its checksum."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _points(seed, dim, n):
    return O.uniform(seed, dim * n).reshape(dim, n)


def _built(gpu, X1, X2, bits, eta, c_leaf, k, kind, ls, reg, prec):
    hm = gpu.HMatrix(X1, X2, bits, eta, 50, c_leaf, k)
    hm.setup()
    kern = gpu.Kernel(kind, reg, ls)
    hm.set_kernel(kern)
    hm.set_precision(prec)
    hm.precompute_aca()
    return hm, kern


@pytest.mark.parametrize("prec_name", ["fp64", "mixed"])
@pytest.mark.parametrize("same", [True, False])
def test_save_load_same_products(gpu, tmp_path, prec_name, same):
    prec = gpu.PREC_MIXED if prec_name == "mixed" else gpu.PREC_FP64
    X1 = _points(8101, 3, 6000)
    X2 = X1 if same else _points(8102, 3, 4500)
    hm, kern = _built(gpu, X1, X2, 20, 1.0, 128, 16, O.K_EXPONENTIAL, 0.1, 0.5 if same else 0.0, prec)
    R = 5
    Xr = O.normal(8103, X2.shape[1] * R).reshape(X2.shape[1], R, order="F")
    y = hm.apply(Xr)
    path = tmp_path / "h.hmg"
    hm.save(path)
    w = hm.work()
    hm.destroy()
    kern.free()
    h2 = gpu.HMatrix.load(path)
    assert h2.work() == w
    y2 = h2.apply(Xr)
    np.testing.assert_array_equal(y2, y)
    # the reloaded H-matrix against the oracle (FP64: its H-MVP; mixed: the BF16x3
    # arithmetic bound relative to sum_j K_ij |x_j|, positive kernel)
    om = O.HMatrix(X1, X2, 20, 1.0, 50, 128)
    ok = O.kernel(O.K_EXPONENTIAL, 3, 0.5 if same else 0.0, 0.1)
    om.aca(ok, 16)
    Yo = om.hmvp(ok, Xr)
    for c in range(R):
        if prec_name == "fp64":
            assert np.linalg.norm(y2[:, c] - Yo[:, c]) / np.linalg.norm(Yo[:, c]) < 1e-8
        else:
            mag = np.linalg.norm(om.hmvp(ok, np.abs(Xr[:, c])))
            assert np.linalg.norm(y2[:, c] - Yo[:, c]) / mag < 4e-5
    np.testing.assert_array_equal(h2.blocks(O.WORK_ITEM_DTYPE).view(np.uint8),
                                  _blocks_of(gpu, X1, X2, h2.blocks(O.WORK_ITEM_DTYPE).dtype))
    # saving the loaded H-matrix gives the same bytes
    path2 = tmp_path / "h2.hmg"
    h2.save(path2)
    assert path.read_bytes() == path2.read_bytes()
    h2.destroy()


def _blocks_of(gpu, X1, X2, dtype):
    hm = gpu.HMatrix(X1, X2, 20, 1.0, 50, 128, 16)
    hm.setup()
    b = hm.blocks(dtype).view(np.uint8)
    hm.destroy()
    return b


def test_corrupted_header_is_refused(gpu, tmp_path):
    """The checksum covers the header: a changed kernel length scale (a valid
    value that the loader would otherwise accept) is refused."""
    X = _points(8211, 2, 2500)
    hm, kern = _built(gpu, X, X, 32, 1.5, 64, 16, O.K_GAUSSIAN, 0.1, 0.0, gpu.PREC_FP64)
    path = tmp_path / "h.hmg"
    hm.save(path)
    raw = bytearray(path.read_bytes())
    ls = np.float64(0.1).tobytes()
    at = raw.find(ls, 0, 160)
    assert at > 0
    raw[at:at + 8] = np.float64(0.2).tobytes()
    bad = tmp_path / "bad.hmg"
    bad.write_bytes(bytes(raw))
    with pytest.raises(gpu.HMGError, match="checksum"):
        gpu.HMatrix.load(bad)
    hm.destroy()
    kern.free()


def test_corrupted_payload_is_refused(gpu, tmp_path):
    X = _points(8201, 2, 3000)
    hm, kern = _built(gpu, X, X, 32, 1.5, 64, 16, O.K_GAUSSIAN, 0.1, 0.0, gpu.PREC_FP64)
    path = tmp_path / "h.hmg"
    hm.save(path)
    raw = bytearray(path.read_bytes())
    raw[-100] ^= 0x40                    # one bit of the stored V factor
    bad = tmp_path / "bad.hmg"
    bad.write_bytes(bytes(raw))
    with pytest.raises(gpu.HMGError, match="checksum"):
        gpu.HMatrix.load(bad)
    trunc = tmp_path / "trunc.hmg"
    trunc.write_bytes(bytes(raw[: len(raw) // 2]))
    with pytest.raises(gpu.HMGError, match="truncated"):
        gpu.HMatrix.load(trunc)
    # the good file still loads
    h2 = gpu.HMatrix.load(path)
    x = O.normal(8203, 3000)
    np.testing.assert_array_equal(h2.apply(x), hm.apply(x))
    h2.destroy()
    hm.destroy()
    kern.free()


def test_save_needs_factors(gpu, tmp_path):
    X = _points(8301, 2, 2000)
    hm = gpu.HMatrix(X, X, 32, 1.5, 50, 64, 8)
    hm.setup()
    with pytest.raises(gpu.HMGError, match="precompute_aca"):
        hm.save(tmp_path / "x.hmg")
    hm.destroy()
