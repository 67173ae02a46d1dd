"""hmglib_amd -- MI355X-native H-matrix engine (drop-in for hmglib's C API).

The product is the shared library libhmglib.so built from csrc/ (HIP kernels for
gfx950 behind the C ABI of include/hmglib.h).  This package is its Python
front-end: a ctypes view of the ABI, device-buffer helpers and a thin wrapper
that drives the ABI like the reference's drivers do.  Importing it fails loudly
when the library is missing; there is no CPU fallback.
"""
from ._abi import (HMG_ERR_HIP, HMG_ERR_INVALID, HMG_ERR_NOT_READY, HMG_ERR_OOM, HMG_OK,
                   KERNEL_EXPONENTIAL, KERNEL_GAUSSIAN, KERNEL_GAUSSIAN_REF,  # noqa: F401
                   KERNEL_MATERN_REF, OPT_ACA_BIG_BLOCK, OPT_ACA_LDS_BYTES, OPT_ACA_HUGE_BLOCK, OPT_ACA_ROW_SCAN, OPT_ACA_TOLERANCE,
                   OPT_EXACT_BOXES, OPT_LR_WORKGROUP, OPT_NEAR_LDS_BYTES, OPT_NEAR_SYMMETRIC, OPT_NEAR_WORKGROUP,
                   OPT_FP64_BLOCK_RHS, OPT_LR_PRIORITY, OPT_MIXED_ACA_PARTS, OPT_PRECISION, OPT_SERIALIZE, PREC_FP64,
                   PREC_MIXED, WT_ACA, WT_DENSE, HMGError, HMatrixData, WorkItem, lib)
from .hmatrix import HMatrix, Kernel, rng_normal, rng_uniform  # noqa: F401
