// noalloc.hip -- the FP64 H-matrix-vector product without device allocation.
//
// hmg_apply_h_matrix_mvp_noalloc computes what hmg_apply_h_matrix_mvp_multi
// computes for an H-matrix whose FP64 factors are precomputed, with the same
// kernels, but its per-call buffers (x in sorted order, the V^T x partials, t,
// and the partial sums every row adds up) live in a fixed device array of
// kScratchDoubles doubles that the code object holds, instead of in memory the
// engine allocates on first use.  It calls no hipMalloc / hipFree and no
// allocating library routine, so it is usable where allocation is not: under
// stream capture, in serving loops that forbid it.
//
// The array is one per device (each GPU holds its own copy for the life of the
// process); its device address is looked up once per device with
// hipGetSymbolAddress and kept.  Calls on one device share the array: two calls
// that overlap on the same device race (the library serialises nothing here).
#include <mutex>

#include "hmg_internal.h"

namespace hmg {
namespace {

// 32 MiB: a product whose buffers need more is refused
constexpr int64_t kScratchDoubles = (int64_t)1 << 22;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// the current device's copy of `scratch` (looked up on the first call per device)
hmg_status device_scratch(double** out)
{
    int dev = -1;
    HMG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("no-alloc product: device index " + std::to_string(dev) + " out of range");
        return HMG_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("no-alloc product: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return HMG_ERR_HIP;
        }
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return HMG_OK;
}

__global__ void fill_zero_kernel(double* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// elements rounded up to 256-byte boundaries
inline int64_t aligned(int64_t n) { return (n + 31) / 32 * 32; }

}  // namespace
}  // namespace hmg

using namespace hmg;

extern "C" hmg_status hmg_apply_h_matrix_mvp_noalloc(const double* X, double* Y, int nrhs,
                                                     struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    if (!X || !Y || nrhs < 1) {
        set_error("invalid x / y / nrhs");
        return HMG_ERR_INVALID;
    }
    if (data->U == nullptr || e->aca_k < 0 || e->fact_prec != HMG_PREC_FP64 || e->layout_prec != HMG_PREC_FP64) {
        set_error("no-alloc product: needs FP64 factors from precompute_aca (the product allocates nothing, "
                  "so it cannot build them)");
        return HMG_ERR_NOT_READY;
    }
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    const MvpSizes z = mvp_sizes(*e);
    const int64_t need = aligned(z.xs) + aligned(z.Tp) + aligned(z.T) + aligned(z.P);
    if (need > kScratchDoubles) {
        set_error("no-alloc product: this H-matrix needs " + std::to_string(need) +
                  " doubles of per-call buffers; the fixed scratch holds " + std::to_string(kScratchDoubles));
        return HMG_ERR_INVALID;
    }
    double* base;
    HMG_TRY(device_scratch(&base));
    MvpBufs b;
    b.xs = base;
    b.Tp = b.xs + aligned(z.xs);
    b.T = b.Tp + aligned(z.Tp);
    b.P = b.T + aligned(z.T);
    // the previous call's values stay in the array: t is cleared so that slots no
    // V column maps to read as zero (a fill kernel, no memset node)
    const int64_t nt = aligned(z.T);
    fill_zero_kernel<<<(unsigned)std::min<int64_t>((nt + 255) / 256, 1024), 256, 0, e->stream>>>(b.T, nt);
    HMG_HIP(hipGetLastError());
    const bool pre_dense = data->dA != nullptr && e->dense_stored;
    // no profiling events here (creating them is runtime allocation)
    const bool prof_on = e->prof.on;
    e->prof.on = false;
    const hmg_status st = hmvp_with(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0],
                                    data->order[1], X, Y, nrhs, e->U.p, e->V.p, e->keff.p, pre_dense, b);
    e->prof.on = prof_on;
    HMG_TRY(st);
    HMG_HIP(hipStreamSynchronize(e->stream));
    return HMG_OK;
}
