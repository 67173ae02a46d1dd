// morton.hip -- Morton codes with FP64 quantisation, plus the index kernels of
// the Z-order permutation.  Compiled with -ffp-contract=off: quantisation is one
// IEEE division, an exact power-of-two scale, floor and a clamp, so the codes are
// bit-exact to the reference (morton.cu:119-155).
#include "hmg_internal.h"

namespace hmg {

namespace {

// Spread the low `bits` bits of q so that bit b lands at bit dim*b (morton.cu:69-115
// defines exactly the pairs (3,20), (2,32), (10,4); any other pair yields 0).
__device__ __forceinline__ uint64_t spread(uint64_t x, int dim, int bits)
{
    if (dim == 3 && bits == 20) {
        x &= 0xFFFFFull;
        x = (x | (x << 32)) & 0x001F00000000FFFFull;
        x = (x | (x << 16)) & 0x001F0000FF0000FFull;
        x = (x | (x << 8)) & 0x100F00F00F00F00Full;
        x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
        x = (x | (x << 2)) & 0x1249249249249249ull;
        return x;
    }
    if (dim == 2 && bits == 32) {
        x &= 0xFFFFFFFFull;
        x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
        x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
        x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
        x = (x | (x << 2)) & 0x3333333333333333ull;
        x = (x | (x << 1)) & 0x5555555555555555ull;
        return x;
    }
    if (dim == 10 && bits == 4) {
        uint64_t out = 0;
        for (int b = 0; b < 4; b++) out |= ((x >> b) & 1ull) << (10 * b);
        return out;
    }
    return 0;
}

// code of point p of c quantised on the box [dmin, dmax] (morton.cu:127-155)
__device__ __forceinline__ uint64_t morton_code(const Coords& c, int64_t p, int dim, int bits,
                                                const double* __restrict__ dmin, const double* __restrict__ dmax)
{
    const double scale = (double)(1ull << bits);
    const uint64_t cap = (1ull << bits) - 1;
    uint64_t code = 0;
    for (int d = 0; d < dim; d++) {
        const double lo = dmin[d];
        const double mmm = __dsub_rn(dmax[d], lo);
        const double tmp = __ddiv_rn(__dsub_rn(c.x[d][p], lo), mmm);
        uint64_t q = (uint64_t)floor(__dmul_rn(tmp, scale));
        q = q < cap ? q : cap;
        code |= spread(q, dim, bits) << d;
    }
    return code;
}

// several clouds side by side: point p of cloud cid[p] on that cloud's box
__global__ void morton_batched_kernel(Coords c, int dim, int bits, const int* __restrict__ cid,
                                      const double* __restrict__ cmin, const double* __restrict__ cmax,
                                      uint64_t* __restrict__ codes)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n) return;
    const int q = cid[p];
    codes[p] = morton_code(c, p, dim, bits, cmin + (int64_t)q * dim, cmax + (int64_t)q * dim);
}

__global__ void morton_kernel(Coords c, int dim, int bits, const double* __restrict__ dmin,
                              const double* __restrict__ dmax, uint64_t* __restrict__ codes)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= c.n) return;
    codes[p] = morton_code(c, p, dim, bits, dmin, dmax);
}

__global__ void iota_kernel(uint32_t* v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

__global__ void gather_kernel(const double* __restrict__ in, const uint32_t* __restrict__ idx,
                              double* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

__global__ void perm_kernel(const uint32_t* __restrict__ p, uint64_t* __restrict__ order,
                            unsigned* __restrict__ ids, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        order[i] = p[i];
        if (ids) ids[i] = p[i];
    }
}

inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hmg_status morton_codes(const Coords& c, int dim, int bits, const double* dmin, const double* dmax,
                        uint64_t* codes, hipStream_t s)
{
    if (c.n <= 0) return HMG_OK;
    morton_kernel<<<grid_for(c.n, 256), 256, 0, s>>>(c, dim, bits, dmin, dmax, codes);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status morton_codes_batched(const Coords& c, int dim, int bits, const int* cid, const double* cmin,
                                const double* cmax, uint64_t* codes, hipStream_t s)
{
    if (c.n <= 0) return HMG_OK;
    morton_batched_kernel<<<grid_for(c.n, 256), 256, 0, s>>>(c, dim, bits, cid, cmin, cmax, codes);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status iota_u32(uint32_t* v, int64_t n, hipStream_t s)
{
    if (n <= 0) return HMG_OK;
    iota_kernel<<<grid_for(n, 256), 256, 0, s>>>(v, n);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status gather_f64(const double* in, const uint32_t* idx, double* out, int64_t n, hipStream_t s)
{
    if (n <= 0) return HMG_OK;
    gather_kernel<<<grid_for(n, 256), 256, 0, s>>>(in, idx, out, n);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status perm_to_u64(const uint32_t* p, uint64_t* order, unsigned* ids, int64_t n, hipStream_t s)
{
    if (n <= 0) return HMG_OK;
    perm_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, order, ids, n);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

}  // namespace hmg
