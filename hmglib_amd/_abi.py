"""ctypes view of the C ABI exported by libhmglib.so (declared in include/*.h).

The struct mirrors follow include/hmglib.h field for field; tests compare every
offset against hmg_struct_layout() so a layout drift fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhmglib.so")

MAX_POINT_DIMENSION = 5
WT_ACA = 1
WT_DENSE = 2

KERNEL_GAUSSIAN_REF = 1
KERNEL_MATERN_REF = 2
KERNEL_GAUSSIAN = 3
KERNEL_EXPONENTIAL = 4

OPT_NEAR_SYMMETRIC = 1
OPT_NEAR_LDS_BYTES = 2
OPT_SERIALIZE = 3
OPT_PRECISION = 4
OPT_ACA_ROW_SCAN = 5
OPT_ACA_BIG_BLOCK = 6
OPT_ACA_TOLERANCE = 7
OPT_ACA_HUGE_BLOCK = 8
OPT_LR_WORKGROUP = 9
OPT_EXACT_BOXES = 10
OPT_NEAR_WORKGROUP = 11
OPT_FP64_BLOCK_RHS = 12
OPT_LR_PRIORITY = 13
OPT_MIXED_ACA_PARTS = 14
OPT_ACA_LDS_BYTES = 15

PREC_FP64 = 0
PREC_MIXED = 1

HMG_OK = 0
HMG_ERR_INVALID = 1
HMG_ERR_OOM = 2
HMG_ERR_HIP = 3
HMG_ERR_NOT_READY = 4

c_double_p = ctypes.POINTER(ctypes.c_double)


class MatVecDataInfo(ctypes.Structure):
    _fields_ = [("dense_count", ctypes.c_int), ("aca_count", ctypes.c_int),
                ("total_count", ctypes.c_int)]


class WorkItem(ctypes.Structure):
    _fields_ = [
        ("set1_l", ctypes.c_int), ("set1_u", ctypes.c_int),
        ("set2_l", ctypes.c_int), ("set2_u", ctypes.c_int),
        ("work_type", ctypes.c_int),
        ("max1", ctypes.c_double * MAX_POINT_DIMENSION),
        ("min1", ctypes.c_double * MAX_POINT_DIMENSION),
        ("max2", ctypes.c_double * MAX_POINT_DIMENSION),
        ("min2", ctypes.c_double * MAX_POINT_DIMENSION),
        ("dim", ctypes.c_int),
        ("level_1", ctypes.c_char), ("level_2", ctypes.c_char),
    ]


class SystemAssembler(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("reserved", ctypes.c_int),
                ("regularization", ctypes.c_double), ("length_scale", ctypes.c_double)]


class HMatrixData(ctypes.Structure):
    """struct h_matrix_data (include/hmglib.h; reference src/hmglib.h:29-102)."""
    _fields_ = [
        ("coords_d", ctypes.POINTER(c_double_p) * 2),
        ("point_ids_d", ctypes.POINTER(ctypes.c_uint) * 2),
        ("max_per_dim_d", c_double_p * 2),
        ("min_per_dim_d", c_double_p * 2),
        ("coords_device", ctypes.POINTER(c_double_p) * 2),
        ("code_d", ctypes.POINTER(ctypes.c_uint64) * 2),
        ("points_d", ctypes.c_void_p * 2),
        ("morton_d", ctypes.c_void_p * 2),
        ("order", ctypes.POINTER(ctypes.c_uint64) * 2),
        ("bits", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("eta", ctypes.c_double),
        ("max_level", ctypes.c_int),
        ("c_leaf", ctypes.c_int),
        ("mat_vec_data_count", ctypes.c_int),
        ("max_elements_in_array", ctypes.c_int),
        ("max_elements_in_mat_vec_data_array", ctypes.c_int),
        ("mat_vec_data", ctypes.POINTER(ctypes.POINTER(WorkItem))),
        ("mat_vec_data_array_size", ctypes.c_int),
        ("k", ctypes.c_int),
        ("epsilon", ctypes.c_double),
        ("mat_vec_info", MatVecDataInfo),
        ("point_count", ctypes.c_int * 2),
        ("assem", ctypes.c_void_p),
        ("max_batched_dense_size", ctypes.c_int),
        ("max_batched_aca_size", ctypes.c_int),
        ("dense_batching_ratio", ctypes.c_double),
        ("U", c_double_p),
        ("V", c_double_p),
        ("dA", c_double_p),
        ("magma_queue", ctypes.c_void_p),
        ("dense_work_size", ctypes.POINTER(ctypes.c_int)),
        ("aca_work_size", ctypes.POINTER(ctypes.c_int)),
        ("dense_batch_count", ctypes.c_int),
        ("aca_batch_count", ctypes.c_int),
        ("root_level_set_1", ctypes.c_char),
        ("root_level_set_2", ctypes.c_char),
        ("engine", ctypes.c_void_p),
    ]


class ServerConfig(ctypes.Structure):
    """hmg_server_config (include/hmglib_server.h)."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("slots", ctypes.c_int),
                ("max_rhs_per_pass", ctypes.c_int), ("precision", ctypes.c_int),
                ("product_slots", ctypes.c_int)]


class BuildRequest(ctypes.Structure):
    """hmg_build_request (include/hmglib_server.h)."""
    _fields_ = [("dim", ctypes.c_int), ("bits", ctypes.c_int), ("n", ctypes.c_int * 2),
                ("coords", c_double_p * 2), ("eta", ctypes.c_double), ("max_level", ctypes.c_int),
                ("c_leaf", ctypes.c_int), ("k", ctypes.c_int), ("kernel_kind", ctypes.c_int),
                ("regularization", ctypes.c_double), ("length_scale", ctypes.c_double),
                ("epsilon", ctypes.c_double), ("aca_tolerance", ctypes.c_int), ("exact_boxes", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    """hmg_request_info (include/hmglib_server.h)."""
    _fields_ = [("status", ctypes.c_int), ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_end", ctypes.c_double), ("merged", ctypes.c_int), ("matrix", ctypes.c_int64),
                ("message", ctypes.c_char * 160)]


# name -> (restype, argtypes) for every symbol declared in include/*.h
_P = ctypes.POINTER
_D = ctypes.POINTER(HMatrixData)
_vp = ctypes.c_void_p
SIGNATURES = {
    # drop-in (include/hmglib.h)
    "init_h_matrix_data": (None, [_D, _P(ctypes.c_int), ctypes.c_int, ctypes.c_int]),
    "setup_h_matrix": (None, [_D]),
    "apply_h_matrix_mvp": (None, [_vp, _vp, _D]),
    "apply_h_matrix_mvp_without_batching": (None, [_vp, _vp, _D]),
    "destroy_h_matrix_data": (None, [_D]),
    "precompute_aca": (None, [_D]),
    "precompute_dense": (None, [_D]),
    "apply_full_mvp": (None, [_vp, _vp, _D]),
    "set_gaussian_kernel_rhs": (None, [_vp, _D]),
    # status twins
    "hmg_init_h_matrix_data": (ctypes.c_int, [_D, _P(ctypes.c_int), ctypes.c_int, ctypes.c_int]),
    "hmg_setup_h_matrix": (ctypes.c_int, [_D]),
    "hmg_apply_h_matrix_mvp": (ctypes.c_int, [_vp, _vp, _D]),
    "hmg_apply_h_matrix_mvp_without_batching": (ctypes.c_int, [_vp, _vp, _D]),
    "hmg_destroy_h_matrix_data": (ctypes.c_int, [_D]),
    "hmg_precompute_aca": (ctypes.c_int, [_D]),
    "hmg_precompute_dense": (ctypes.c_int, [_D]),
    "hmg_apply_full_mvp": (ctypes.c_int, [_vp, _vp, _D]),
    "hmg_set_gaussian_kernel_rhs": (ctypes.c_int, [_vp, _D]),
    # extensions
    "hmg_apply_h_matrix_mvp_multi": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _D]),
    "hmg_apply_h_matrix_mvp_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _D]),
    "hmg_autotune_product": (ctypes.c_int, [_D, _vp, _vp, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int64)]),
    "hmg_dense_rows": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _D]),
    "hmg_get_work": (ctypes.c_int, [_D, _P(ctypes.c_int64)]),
    "hmg_get_schedule": (ctypes.c_int, [_D, _P(ctypes.c_int64)]),
    "hmg_set_profiling": (ctypes.c_int, [_D, ctypes.c_int]),
    "hmg_set_option": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_int64]),
    "hmg_get_profile": (ctypes.c_int, [_D, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int64)]),
    "hmg_copy_probe": (ctypes.c_int, [_vp, _vp, ctypes.c_int64]),
    "hmg_release_cached_memory": (None, []),
    "hmg_memory": (ctypes.c_int, [_D, _P(ctypes.c_int64)]),
    "hmg_setup_batch": (ctypes.c_int, [_P(_D), ctypes.c_int, ctypes.c_int]),
    "hmg_apply_batch_graph_create": (ctypes.c_int, [_P(_D), _P(_vp), _P(_vp), _P(ctypes.c_int), ctypes.c_int,
                                                    _P(_vp)]),
    "hmg_apply_batch_graph_launch": (ctypes.c_int, [_vp]),
    "hmg_apply_batch_graph_destroy": (None, [_vp]),
    "hmg_precompute_aca_batch": (ctypes.c_int, [_P(_D), ctypes.c_int, ctypes.c_int]),
    "hmg_apply_batch": (ctypes.c_int, [_P(_D), _P(_vp), _P(_vp), _P(ctypes.c_int), ctypes.c_int,
                                       ctypes.c_int]),
    "hmg_save": (ctypes.c_int, [_D, ctypes.c_char_p]),
    "hmg_load": (ctypes.c_int, [_D, ctypes.c_char_p]),
    # serving (include/hmglib_server.h)
    "hmg_server_create": (ctypes.c_int, [_P(ServerConfig), _P(_vp)]),
    "hmg_server_destroy": (ctypes.c_int, [_vp]),
    "hmg_server_submit_build": (ctypes.c_int, [_vp, _P(BuildRequest), _P(ctypes.c_int64),
                                               _P(ctypes.c_int64)]),
    "hmg_server_submit_apply": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, ctypes.c_int,
                                               _P(ctypes.c_int64)]),
    "hmg_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, _P(RequestInfo)]),
    "hmg_server_query": (ctypes.c_int, [_vp, ctypes.c_int64, _P(RequestInfo)]),
    "hmg_server_drain": (ctypes.c_int, [_vp]),
    "hmg_server_release": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "hmg_server_stats": (ctypes.c_int, [_vp, _P(ctypes.c_int64)]),
    "hmg_server_resident": (ctypes.c_int, [_vp, _P(ctypes.c_int64), ctypes.c_int]),
    "hmg_last_error": (ctypes.c_char_p, []),
    "hmg_struct_layout": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int]),
    "hmg_rng_uniform": (None, [ctypes.c_uint64, ctypes.c_int64, c_double_p]),
    "hmg_rng_normal": (None, [ctypes.c_uint64, ctypes.c_int64, c_double_p]),
    # kernels (include/kernel_system_assembler.h)
    "create_gaussian_kernel_system_assembler_object": (None, [_vp, ctypes.c_double]),
    "destroy_gaussian_kernel_system_assembler_object": (None, [_vp]),
    "create_matern_kernel_system_assembler_object": (None, [_vp, ctypes.c_double]),
    "destroy_matern_kernel_system_assembler_object": (None, [_vp]),
    "hmg_create_kernel_object": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "hmg_destroy_kernel_object": (ctypes.c_int, [_vp]),
    "hmg_kernel_new": (_vp, [ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "hmg_kernel_free": (None, [_vp]),
}

_lib = None


class HMGError(RuntimeError):
    def __init__(self, status, where, msg):
        super().__init__(f"{where} failed with status {status}: {msg}")
        self.status = status


def lib():
    """Load libhmglib.so; raises if the HIP engine has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP engine not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != HMG_OK:
        msg = lib().hmg_last_error()
        raise HMGError(status, where, msg.decode() if msg else "")
