// kernel_eval.h -- inline kernel-entry evaluation for the HIP kernels.
//
// K(i,j) = k(||x1_i - x2_j||) on Morton-sorted indices, plus the regularization
// when the sorted indices coincide (reference: kernel_system_assembler.h:38-52,
// 82-96).  r^2 is accumulated as r2 = r2 + t*t in dimension order, exactly like
// the CPU restatement; exp/sqrt/yn come from the ROCm device library.
#pragma once
#include "hmg_internal.h"

namespace hmg {

__device__ __forceinline__ double dist2(const Coords& a, int i, const Coords& b, int j, int dim)
{
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
        if (d < dim) {
            double t = a.x[d][i] - b.x[d][j];
            r2 = __dadd_rn(r2, __dmul_rn(t, t));
        }
    }
    return r2;
}

// One kernel kind fixed at compile time: code for one kind only, which keeps the
// register count of kernels that evaluate entries in their inner loops low (the
// Bessel function of the Matern kernel alone needs many registers).
template <int KIND>
__device__ __forceinline__ double kernel_of_r2_k(const KernelParams& kp, double r2)
{
    if (KIND == HMG_KERNEL_GAUSSIAN_REF) {
        const double r = sqrt(r2);
        return exp(-(r * r));
    }
    if (KIND == HMG_KERNEL_MATERN_REF) {
        const double r = sqrt(r2);
        return (yn(1, r + 1.0e-15) * r) / kp.c1;
    }
    if (KIND == HMG_KERNEL_GAUSSIAN) return exp(-(r2 * kp.c0));
    if (KIND == HMG_KERNEL_EXPONENTIAL) return exp(-(sqrt(r2) * kp.c0));
    return __builtin_nan("");
}

__device__ __forceinline__ double kernel_of_r2(const KernelParams& kp, double r2)
{
    switch (kp.kind) {
    case HMG_KERNEL_GAUSSIAN_REF:
        return kernel_of_r2_k<HMG_KERNEL_GAUSSIAN_REF>(kp, r2);
    case HMG_KERNEL_MATERN_REF:
        return kernel_of_r2_k<HMG_KERNEL_MATERN_REF>(kp, r2);
    case HMG_KERNEL_GAUSSIAN:
        return kernel_of_r2_k<HMG_KERNEL_GAUSSIAN>(kp, r2);
    case HMG_KERNEL_EXPONENTIAL:
        return kernel_of_r2_k<HMG_KERNEL_EXPONENTIAL>(kp, r2);
    default:
        return __builtin_nan("");
    }
}

__device__ __forceinline__ double kernel_entry(const KernelParams& kp, const Coords& a, int i,
                                               const Coords& b, int j)
{
    double v = kernel_of_r2(kp, dist2(a, i, b, j, kp.dim));
    if (i == j) v += kp.reg;
    return v;
}

template <int KIND>
__device__ __forceinline__ double kernel_entry_k(const KernelParams& kp, const Coords& a, int i,
                                                 const Coords& b, int j)
{
    double v = kernel_of_r2_k<KIND>(kp, dist2(a, i, b, j, kp.dim));
    if (i == j) v += kp.reg;
    return v;
}

// Host side: runs the statements with `constexpr int KIND` set to the kernel kind
// (kinds are validated when a kernel object is created).
#define HMG_DISPATCH_KIND(kind, ...)                                              \
    switch (kind) {                                                               \
    case HMG_KERNEL_GAUSSIAN_REF: {                                               \
        constexpr int KIND = HMG_KERNEL_GAUSSIAN_REF;                             \
        __VA_ARGS__;                                                              \
    } break;                                                                      \
    case HMG_KERNEL_MATERN_REF: {                                                 \
        constexpr int KIND = HMG_KERNEL_MATERN_REF;                               \
        __VA_ARGS__;                                                              \
    } break;                                                                      \
    case HMG_KERNEL_GAUSSIAN: {                                                   \
        constexpr int KIND = HMG_KERNEL_GAUSSIAN;                                 \
        __VA_ARGS__;                                                              \
    } break;                                                                      \
    default: {                                                                    \
        constexpr int KIND = HMG_KERNEL_EXPONENTIAL;                              \
        __VA_ARGS__;                                                              \
    } break;                                                                      \
    }

}  // namespace hmg
