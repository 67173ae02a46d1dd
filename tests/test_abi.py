"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point that include/*.h declares, and its struct layouts match the ctypes view.
Nothing here launches GPU work."""
import ctypes
import os
import re

import pytest

import hmglib_amd
from hmglib_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names.add(name)
    return sorted(names)


def test_header_declares_the_reference_entry_points():
    names = set(_declared_functions())
    for ref in ["init_h_matrix_data", "setup_h_matrix", "apply_h_matrix_mvp",
                "apply_h_matrix_mvp_without_batching", "destroy_h_matrix_data", "precompute_aca",
                "precompute_dense", "apply_full_mvp", "set_gaussian_kernel_rhs",
                "create_gaussian_kernel_system_assembler_object",
                "destroy_gaussian_kernel_system_assembler_object",
                "create_matern_kernel_system_assembler_object",
                "destroy_matern_kernel_system_assembler_object"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    L = hmglib_amd.lib()
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers the same set
    assert set(names) == set(_abi.SIGNATURES)


def test_struct_layouts_match():
    L = hmglib_amd.lib()
    assert L.hmg_struct_layout(0, -1) == ctypes.sizeof(_abi.HMatrixData)
    for i, (name, _) in enumerate(_abi.HMatrixData._fields_):
        assert getattr(_abi.HMatrixData, name).offset == L.hmg_struct_layout(0, i), name
    assert L.hmg_struct_layout(1, -1) == ctypes.sizeof(_abi.WorkItem) == 192
    for i, (name, _) in enumerate(_abi.WorkItem._fields_):
        assert getattr(_abi.WorkItem, name).offset == L.hmg_struct_layout(1, i), name
    # reference x86-64 layout of work_item (tree.h:38-55): max1 @24, dim @184
    assert _abi.WorkItem.max1.offset == 24 and _abi.WorkItem.dim.offset == 184
    assert L.hmg_struct_layout(4, -1) == ctypes.sizeof(_abi.SystemAssembler) == 24
    assert L.hmg_struct_layout(5, -1) == 12


def test_invalid_arguments_are_rejected_without_touching_the_gpu():
    L = hmglib_amd.lib()
    d = _abi.HMatrixData()
    pc = (ctypes.c_int * 2)(100, 100)
    assert L.hmg_init_h_matrix_data(ctypes.byref(d), pc, 0, 32) == _abi.HMG_ERR_INVALID
    assert b"dim" in L.hmg_last_error()
    assert L.hmg_init_h_matrix_data(ctypes.byref(d), pc, 6, 32) == _abi.HMG_ERR_INVALID
    bad = (ctypes.c_int * 2)(0, 100)
    assert L.hmg_init_h_matrix_data(ctypes.byref(d), bad, 2, 32) == _abi.HMG_ERR_INVALID
    assert L.hmg_init_h_matrix_data(None, pc, 2, 32) == _abi.HMG_ERR_INVALID
    # an uninitialised struct is refused by every stage
    z = _abi.HMatrixData()
    assert L.hmg_setup_h_matrix(ctypes.byref(z)) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_precompute_aca(ctypes.byref(z)) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_precompute_dense(ctypes.byref(z)) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_apply_h_matrix_mvp(None, None, ctypes.byref(z)) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_apply_full_mvp(None, None, ctypes.byref(z)) == _abi.HMG_ERR_NOT_READY
    out = (ctypes.c_int64 * 6)()
    assert L.hmg_get_work(ctypes.byref(z), out) == _abi.HMG_ERR_NOT_READY
    tuned = (ctypes.c_int64 * 4)()
    assert L.hmg_autotune_product(ctypes.byref(z), None, None, 1, 1, tuned) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_destroy_h_matrix_data(ctypes.byref(z)) == _abi.HMG_OK
    assert not L.hmg_kernel_new(99, 0.0, 1.0)


def test_python_front_end_raises_when_library_is_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _abi.lib()


def test_server_struct_layouts_match():
    L = hmglib_amd.lib()
    assert L.hmg_struct_layout(6, -1) == ctypes.sizeof(_abi.ServerConfig)
    assert L.hmg_struct_layout(7, -1) == ctypes.sizeof(_abi.BuildRequest)
    for i, (name, _) in enumerate(_abi.BuildRequest._fields_):
        assert getattr(_abi.BuildRequest, name).offset == L.hmg_struct_layout(7, i), name
    assert L.hmg_struct_layout(8, -1) == ctypes.sizeof(_abi.RequestInfo)
    for i, (name, _) in enumerate(_abi.RequestInfo._fields_):
        assert getattr(_abi.RequestInfo, name).offset == L.hmg_struct_layout(8, i), name


def test_server_host_logic_without_the_gpu():
    """Creating, querying and destroying a server with no requests touches no GPU;
    invalid configurations and requests are refused up front."""
    L = hmglib_amd.lib()
    h = ctypes.c_void_p()
    bad = _abi.ServerConfig(0, 4, 32, _abi.PREC_MIXED)
    assert L.hmg_server_create(ctypes.byref(bad), ctypes.byref(h)) == _abi.HMG_ERR_INVALID
    bad = _abi.ServerConfig(1 << 30, 0, 32, _abi.PREC_MIXED)
    assert L.hmg_server_create(ctypes.byref(bad), ctypes.byref(h)) == _abi.HMG_ERR_INVALID
    bad = _abi.ServerConfig(1 << 30, 2, 32, _abi.PREC_MIXED, 2)   # every slot kept from builds
    assert L.hmg_server_create(ctypes.byref(bad), ctypes.byref(h)) == _abi.HMG_ERR_INVALID
    bad = _abi.ServerConfig(1 << 30, 2, 32, 7)
    assert L.hmg_server_create(ctypes.byref(bad), ctypes.byref(h)) == _abi.HMG_ERR_INVALID
    cfg = _abi.ServerConfig(1 << 30, 3, 32, _abi.PREC_FP64)
    assert L.hmg_server_create(ctypes.byref(cfg), ctypes.byref(h)) == _abi.HMG_OK
    r = ctypes.c_int64()
    dummy = ctypes.c_void_p(16)
    # unknown H-matrix: refused at submit
    assert L.hmg_server_submit_apply(h, 12345, dummy, dummy, 1, ctypes.byref(r)) == _abi.HMG_ERR_NOT_READY
    assert L.hmg_server_submit_apply(h, 1, dummy, dummy, 0, ctypes.byref(r)) == _abi.HMG_ERR_INVALID
    b = _abi.BuildRequest()
    b.dim = 9
    m = ctypes.c_int64()
    assert L.hmg_server_submit_build(h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(m)) == \
        _abi.HMG_ERR_INVALID
    info = _abi.RequestInfo()
    assert L.hmg_server_query(h, 999, ctypes.byref(info)) == _abi.HMG_ERR_INVALID
    out = (ctypes.c_int64 * 8)()
    assert L.hmg_server_stats(h, out) == _abi.HMG_OK
    assert list(out) == [0] * 8
    ids = (ctypes.c_int64 * 4)()
    assert L.hmg_server_resident(h, ids, 4) == 0
    assert L.hmg_server_drain(h) == _abi.HMG_OK
    assert L.hmg_server_destroy(h) == _abi.HMG_OK


def _store_header(**kw):
    """A header of the H-matrix file format (layout of StoreHeader in store.hip)."""
    import struct
    f = dict(magic=b"HMGAMDHM", version=2, header_bytes=160, dim=2, bits=32, n0=10, n1=10, max_level=50,
             c_leaf=4, k=2, precision=0, r1=0, r2=0, kind=3, same=1, nsym=1, pad=0, eta=1.5, eps=0.0,
             reg=0.0, ls=0.1, nblocks=1, ndense=1, naca=0, u=0, v=0, eb=8, checksum=0)
    f.update(kw)
    return struct.pack("<8sII14i4d6qQ", f["magic"], f["version"], f["header_bytes"], f["dim"], f["bits"],
                       f["n0"], f["n1"], f["max_level"], f["c_leaf"], f["k"], f["precision"], f["r1"], f["r2"],
                       f["kind"], f["same"], f["nsym"], f["pad"], f["eta"], f["eps"], f["reg"], f["ls"],
                       f["nblocks"], f["ndense"], f["naca"], f["u"], f["v"], f["eb"], f["checksum"])


def test_load_refuses_bad_files_without_the_gpu(tmp_path):
    """hmg_load checks the header and reads the points before it touches the
    device: missing files, foreign magic, other versions, inconsistent headers and
    truncated point arrays are refused with HMG_ERR_INVALID on a machine without
    a GPU."""
    L = hmglib_amd.lib()
    d = _abi.HMatrixData()
    assert len(_store_header()) == 160
    cases = {
        "missing": None,
        "magic": _store_header(magic=b"NOTHMGLB") + bytes(160),
        "short": _store_header()[:100],
        "version": _store_header(version=1) + bytes(160),
        "precision": _store_header(precision=1, eb=8) + bytes(160),
        "dim": _store_header(dim=9) + bytes(160),
        "points": _store_header() + bytes(80),   # 2 x 10 doubles needed, 80 bytes given
    }
    for name, raw in cases.items():
        p = tmp_path / f"{name}.hmg"
        if raw is not None:
            p.write_bytes(raw)
        st = L.hmg_load(ctypes.byref(d), str(p).encode())
        assert st == _abi.HMG_ERR_INVALID, name
        assert d.engine is None or d.engine == 0, name
