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

// e^x for x <= 0 (the argument of the product kernels' entries), within 1 ulp:
// Cody-Waite reduction x = n ln2 + r with ln2 split into two doubles (its double
// and the double of the remainder), |r| <= ln2/2, the degree-13 Taylor polynomial
// of e^r (truncation below 5e-18 relative) by Horner's rule, scaled by 2^n.
// Arguments below -800 give 0 (e^-800 is below the smallest double).
__device__ __forceinline__ double exp_nonpos(double x)
{
    x = fmax(x, -800.0);
    const double n = rint(x * 1.4426950408889634);
    double r = __fma_rn(n, -0.6931471805599453, x);
    r = __fma_rn(n, -2.3190468138462996e-17, r);
    double p = 1.6059043836821613e-10;  // 1/13!
    p = __fma_rn(p, r, 2.08767569878681e-09);
    p = __fma_rn(p, r, 2.505210838544172e-08);
    p = __fma_rn(p, r, 2.755731922398589e-07);
    p = __fma_rn(p, r, 2.7557319223985893e-06);
    p = __fma_rn(p, r, 2.48015873015873e-05);
    p = __fma_rn(p, r, 0.0001984126984126984);
    p = __fma_rn(p, r, 0.001388888888888889);
    p = __fma_rn(p, r, 0.008333333333333333);
    p = __fma_rn(p, r, 0.041666666666666664);
    p = __fma_rn(p, r, 0.16666666666666666);
    p = __fma_rn(p, r, 0.5);
    p = __fma_rn(p, r, 1.0);
    p = __fma_rn(p, r, 1.0);
    return ldexp(p, (int)n);
}

// sqrt(x) for finite x >= 0: the hardware reciprocal square root refined by a
// Goldschmidt step and two Newton corrections of the root (g += (x - g^2) h)
__device__ __forceinline__ double sqrt_nonneg(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __fma_rn(-g, h, 0.5);
    g = __fma_rn(g, r, g);
    h = __fma_rn(h, r, h);
    double d = __fma_rn(-g, g, x);
    g = __fma_rn(d, h, g);
    d = __fma_rn(-g, g, x);
    g = __fma_rn(d, h, g);
    return x == 0.0 ? 0.0 : g;
}

// The entries of the FP64 products (near field, dense rows, stored near field):
// the Gaussian and exponential kernels through exp_nonpos / sqrt_nonneg (fewer
// instructions than the general functions: no overflow, scaling or class
// handling), the reference kernels as above.  The ACA keeps kernel_of_r2_k.
template <int KIND>
__device__ __forceinline__ double kernel_of_r2_prod(const KernelParams& kp, double r2)
{
    if (KIND == HMG_KERNEL_GAUSSIAN) return exp_nonpos(-(r2 * kp.c0));
    if (KIND == HMG_KERNEL_EXPONENTIAL) return exp_nonpos(-(sqrt_nonneg(r2) * kp.c0));
    return kernel_of_r2_k<KIND>(kp, r2);
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
