"""The reference's own driver sequence through the drop-in C API.

tests/ref_flow/hmglib_flow.c is hmglib_test.cu:25-142 written against
include/hmglib.h: init_h_matrix_data, coordinates written into data.coords_d,
setup_h_matrix, a kernel object created through a DEVICE slot and copied back
(the create_*_kernel_system_assembler_object protocol, hmglib_test.cu:82-90),
precompute_aca, apply_full_mvp, apply_h_matrix_mvp, then (beyond the reference
driver) apply_h_matrix_mvp_without_batching, precompute_dense + apply and
set_gaussian_kernel_rhs, and the destroy calls.  It is a C11 library built by
tests/ref_flow/Makefile and called in-process.  Every output is checked against
the CPU oracle on the same inputs: block counts and batch sizes exactly, the
dense product to 1e-13, the H-matrix products to 1e-8.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FLOW = os.path.join(HERE, "ref_flow", "libhmglib_flow.so")


def _flow():
    if not os.path.exists(FLOW):
        raise RuntimeError(f"{FLOW} is missing: build it with __graft_entry__.build()")
    import hmglib_amd
    hmglib_amd.lib()          # the engine first, so the driver binds to the same instance
    L = ctypes.CDLL(FLOW)
    L.hmglib_flow_run.restype = ctypes.c_int
    L.hmglib_flow_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    return L


def _run(tmp_path, X1, X2, x, bits, k, c_leaf, eta, max_aca, kernel, reg=0.0, max_level=50, eps=1e-6):
    dim, n0 = X1.shape
    n1 = X2.shape[1]
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n0, n1, dim, bits, k, c_leaf, max_level, max_aca, kernel], dtype=np.int32).tobytes())
        f.write(np.array([eta, eps, reg], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(X1, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(X2, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
    rc = _flow().hmglib_flow_run(os.fsencode(inp), os.fsencode(out))
    assert rc == 0, f"hmglib_flow_run returned {rc}"
    raw = out.read_bytes()
    info = np.frombuffer(raw[:24], dtype=np.int32)
    nd, na = int(info[3]), int(info[4])
    off = 24
    dws = np.frombuffer(raw[off:off + 4 * nd], dtype=np.int32)
    off += 4 * nd
    aws = np.frombuffer(raw[off:off + 4 * na], dtype=np.int32)
    off += 4 * na
    rel = float(np.frombuffer(raw[off:off + 8], dtype=np.float64)[0])
    off += 8
    vecs = np.frombuffer(raw[off:off + 8 * 5 * n0], dtype=np.float64).reshape(5, n0)
    off += 8 * 5 * n0
    x_after = np.frombuffer(raw[off:off + 8 * n1], dtype=np.float64)
    return {"dense": int(info[0]), "aca": int(info[1]), "total": int(info[2]), "dense_cap": int(info[5]),
            "dws": dws, "aws": aws, "rel": rel, "y_full": vecs[0], "y": vecs[1], "y_nb": vecs[2],
            "y_st": vecs[3], "b": vecs[4], "x_after": x_after}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", ["gauss_2d_indep", "gauss_2d_same", "matern_3d"])
def test_reference_driver_sequence_against_oracle(gpu, tmp_path, case):
    if case == "gauss_2d_indep":
        # hmglib_test.cu's shape: 2D, independent sets, k 16, leaf 64, eta 1.5, ACA batches of 65536 rows
        dim, bits, n0, n1, k, c_leaf, eta, kind, max_aca = 2, 32, 4096, 3500, 16, 64, 1.5, 1, 65536
        X1 = O.uniform(1001, dim * n0).reshape(dim, n0)
        X2 = O.uniform(1002, dim * n1).reshape(dim, n1)
        reg = 0.0
    elif case == "gauss_2d_same":
        dim, bits, n0, n1, k, c_leaf, eta, kind, max_aca = 2, 32, 3000, 3000, 12, 32, 1.0, 1, 2000
        X1 = O.uniform(1011, dim * n0).reshape(dim, n0)
        X2 = X1
        reg = 0.25
    else:
        dim, bits, n0, n1, k, c_leaf, eta, kind, max_aca = 3, 20, 2500, 2500, 10, 64, 1.5, 2, 65536
        X1 = O.uniform(1021, dim * n0).reshape(dim, n0)
        X2 = X1
        reg = 0.0
    x = O.uniform(1003, n1)
    r = _run(tmp_path, X1, X2, x, bits, k, c_leaf, eta, max_aca, kind, reg=reg)

    om = O.HMatrix(X1, X2, bits, eta, 50, c_leaf)
    ok = O.kernel(O.K_GAUSSIAN_REF if kind == 1 else O.K_MATERN_REF, dim, reg, 1.0)
    # block counts and the batch sizing of setup (hmglib.cu:211-219)
    assert (r["dense"], r["aca"], r["total"]) == (om.ndense, om.naca, om.nblocks)
    cap, dws, aws = om.work_sizes(max_aca)
    assert r["dense_cap"] == cap
    np.testing.assert_array_equal(r["dws"], dws)
    np.testing.assert_array_equal(r["aws"], aws)
    # products: dense, H-matrix (last-argmax ACA), unbatched (first-argmax ACA),
    # stored near field, closed-form right-hand side; x comes back unchanged
    y_dense = om.dense_mvp(ok, x)
    assert _rel(r["y_full"], y_dense) < 1e-13
    om.aca(ok, k)
    y_h = om.hmvp(ok, x)
    assert _rel(r["y"], y_h) < 1e-8
    assert _rel(r["y_st"], y_h) < 1e-8
    assert abs(r["rel"] - _rel(r["y"], r["y_full"])) <= 1e-12
    om.aca(ok, k, tie_last=False)
    assert _rel(r["y_nb"], om.hmvp(ok, x)) < 1e-8
    np.testing.assert_allclose(r["b"], om.gaussian_rhs(), rtol=1e-14, atol=0)
    np.testing.assert_array_equal(r["x_after"], x)
