// prims.hip -- device-wide primitives written for gfx950 wave64:
//   * scan_i32       : tiled scan (2048 items per 256-thread block), recursive
//                      over block totals; exclusive or inclusive.
//   * radix_sort_u64 : stable LSD radix sort of (uint64 key, uint32 value) with
//                      8-bit digits.  Per pass: block histograms -> exclusive scan
//                      of the digit-major count matrix -> stable scatter, where
//                      each wavefront ranks equal digits with 8 ballots (a 64-lane
//                      multi-split) and waves are ordered through LDS counters.
//   * minmax_f64     : exact min/max of a double array (two-level reduction).
// These replace the thrust calls of the reference (helper.cu:165 sort_by_key,
// tree.cu:1495-1500 reduce/exclusive_scan, helper.cu:111 minmax_element).
#include <map>
#include <mutex>
#include <utility>

#include "hmg_internal.h"

namespace hmg {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Grow-only scratch buffers keyed by (stream, slot).  Calls on one stream are
// issued by one thread at a time and run in order, so a slot can be reused and
// regrown between them (hipFree waits for the device).  Every H-matrix engine
// owns its streams, so engines driven concurrently from different threads never
// share a buffer; release_scratch drops a stream's buffers when its engine goes.
struct ScratchCache {
    std::mutex mu;
    std::map<std::pair<hipStream_t, int>, DevBuf<unsigned char>> bufs;
};
ScratchCache& scratch_cache()
{
    // never destroyed: freeing device memory at exit can run after the HIP
    // runtime has shut down; the memory is reclaimed with the process
    static ScratchCache* c = new ScratchCache();
    return *c;
}

hmg_status scratch(hipStream_t s, int slot, size_t bytes, void** out)
{
    ScratchCache& c = scratch_cache();
    std::lock_guard<std::mutex> g(c.mu);
    DevBuf<unsigned char>& b = c.bufs[std::make_pair(s, slot)];
    if (b.n < bytes) {
        b.release();
        HMG_TRY(b.alloc(bytes + bytes / 4));
    }
    *out = b.p;
    return HMG_OK;
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        int t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

// Scans one tile of 2048 ints; writes the tile's total to block_sums[blockIdx.x].
// Callers scan in place (in == out): a workgroup reads its whole tile into LDS
// before it writes any of it, and no other workgroup touches that tile.
__global__ void __launch_bounds__(kScanThreads)
scan_tile_kernel(const int* in, int* out, int64_t n, int* block_sums,
                 int inclusive)
{
    __shared__ int tile[kScanTile];
    __shared__ int wsum[kScanThreads / kWave];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        int64_t g = base + r * kScanThreads + tid;
        tile[r * kScanThreads + tid] = g < n ? in[g] : 0;
    }
    __syncthreads();
    int v[kScanItems];
    int run = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = tile[tid * kScanItems + i];
        run += v[i];
    }
    int incl = wave_incl_scan(run);
    const int wid = tid / kWave, lane = tid % kWave;
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    int wave_off = 0;
    for (int w = 0; w < wid; w++) wave_off += wsum[w];
    int excl = wave_off + incl - run;
    int acc = excl;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        int x = v[i];
        tile[tid * kScanItems + i] = inclusive ? acc + x : acc;
        acc += x;
    }
    if (tid == kScanThreads - 1) block_sums[blockIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        int64_t g = base + r * kScanThreads + tid;
        if (g < n) out[g] = tile[r * kScanThreads + tid];
    }
}

__global__ void __launch_bounds__(kScanThreads)
scan_add_kernel(int* __restrict__ out, int64_t n, const int* __restrict__ block_offs)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const int off = block_offs[blockIdx.x];
    if (off == 0) return;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        int64_t g = base + r * kScanThreads + threadIdx.x;
        if (g < n) out[g] += off;
    }
}

hmg_status scan_rec(const int* in, int* out, int64_t n, bool inclusive, int* total,
                    hipStream_t s, int depth)
{
    if (n <= 0) {
        if (total) HMG_HIP(hipMemsetAsync(total, 0, sizeof(int), s));
        return HMG_OK;
    }
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    void* p;
    if (nb == 1) {
        // one tile: its total is the scan's total, written where the caller wants it
        int* sums = total;
        if (!sums) {
            HMG_TRY(scratch(s, 100 + depth, sizeof(int), &p));
            sums = (int*)p;
        }
        scan_tile_kernel<<<1, kScanThreads, 0, s>>>(in, out, n, sums, inclusive ? 1 : 0);
        HMG_HIP(hipGetLastError());
        return HMG_OK;
    }
    HMG_TRY(scratch(s, 100 + depth, sizeof(int) * (size_t)nb, &p));
    int* bsums = (int*)p;
    scan_tile_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, bsums, inclusive ? 1 : 0);
    HMG_HIP(hipGetLastError());
    HMG_TRY(scan_rec(bsums, bsums, nb, false, total, s, depth + 1));
    scan_add_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(out, n, bsums);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

// ---------------------------------------------------------------- radix sort
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kSortWaves = kSortThreads / kWave;

__global__ void __launch_bounds__(kSortThreads)
radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n, int shift, int* __restrict__ counts,
                  int nblocks)
{
    __shared__ int hist[256];
    const int tid = threadIdx.x;
    hist[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortRounds; r++) {
        int64_t g = base + r * kSortThreads + tid;
        if (g < n) atomicAdd(&hist[(keys[g] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[(int64_t)tid * nblocks + blockIdx.x] = hist[tid];
}

__global__ void __launch_bounds__(kSortThreads)
radix_scatter_kernel(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n, int shift,
                     const int* __restrict__ offs_g, int nblocks)
{
    __shared__ int offs[256];
    __shared__ int wcnt[kSortWaves][256];
    const int tid = threadIdx.x;
    const int wid = tid / kWave;
    const int lane = tid % kWave;
    offs[tid] = offs_g[(int64_t)tid * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; r++) {
        const int64_t g = base + r * kSortThreads + tid;
        const bool valid = g < n;
        uint64_t key = valid ? kin[g] : 0ull;
        uint32_t val = valid ? vin[g] : 0u;
        const unsigned digit = (unsigned)((key >> shift) & 255u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wid][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = offs[digit] + rank;
            for (int w = 0; w < wid; w++) pos += wcnt[w][digit];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; w++) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        offs[tid] += add;
        __syncthreads();
    }
}

// ----------------------------------------------------------------- min / max
constexpr int kRedThreads = 256;

__global__ void __launch_bounds__(kRedThreads)
minmax_partial_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ pmin,
                      double* __restrict__ pmax)
{
    __shared__ double smin[kRedThreads], smax[kRedThreads];
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedThreads) {
        double v = x[i];
        lo = v < lo ? v : lo;
        hi = hi < v ? v : hi;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = smin[threadIdx.x + s], b = smax[threadIdx.x + s];
            smin[threadIdx.x] = a < smin[threadIdx.x] ? a : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x] < b ? b : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = smin[0];
        pmax[blockIdx.x] = smax[0];
    }
}

__global__ void __launch_bounds__(kRedThreads)
minmax_final_kernel(const double* __restrict__ pmin, const double* __restrict__ pmax, int nb,
                    double* __restrict__ dmin, double* __restrict__ dmax)
{
    __shared__ double smin[kRedThreads], smax[kRedThreads];
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int i = threadIdx.x; i < nb; i += kRedThreads) {
        lo = pmin[i] < lo ? pmin[i] : lo;
        hi = hi < pmax[i] ? pmax[i] : hi;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = smin[threadIdx.x + s], b = smax[threadIdx.x + s];
            smin[threadIdx.x] = a < smin[threadIdx.x] ? a : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x] < b ? b : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *dmin = smin[0];
        *dmax = smax[0];
    }
}

// bitwise comparison of two double arrays: any difference sets *differ = 1
__global__ void __launch_bounds__(kRedThreads)
differ_kernel(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, int64_t n,
              int* __restrict__ differ)
{
    bool d = false;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedThreads)
        d |= a[i] != b[i];
    if (__any(d) && (threadIdx.x & (kWave - 1)) == 0) *differ = 1;
}

}  // namespace

void release_scratch(hipStream_t s)
{
    ScratchCache& c = scratch_cache();
    std::lock_guard<std::mutex> g(c.mu);
    for (auto it = c.bufs.begin(); it != c.bufs.end();) {
        if (it->first.first == s) {
            it->second.release();
            it = c.bufs.erase(it);
        } else {
            ++it;
        }
    }
}

hmg_status scan_i32(const int* in, int* out, int64_t n, bool inclusive, int* total, hipStream_t s)
{
    return scan_rec(in, out, n, inclusive, total, s, 0);
}

hmg_status radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s)
{
    if (n <= 1) return HMG_OK;
    const int passes = (key_bits + 7) / 8;
    const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
    void *pk, *pv, *pc;
    HMG_TRY(scratch(s, 1, sizeof(uint64_t) * (size_t)n, &pk));
    HMG_TRY(scratch(s, 2, sizeof(uint32_t) * (size_t)n, &pv));
    HMG_TRY(scratch(s, 3, sizeof(int) * (size_t)nblocks * 256, &pc));
    uint64_t* k_alt = (uint64_t*)pk;
    uint32_t* v_alt = (uint32_t*)pv;
    int* counts = (int*)pc;
    uint64_t* ksrc = keys;
    uint32_t* vsrc = vals;
    uint64_t* kdst = k_alt;
    uint32_t* vdst = v_alt;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        radix_hist_kernel<<<nblocks, kSortThreads, 0, s>>>(ksrc, n, shift, counts, nblocks);
        HMG_HIP(hipGetLastError());
        HMG_TRY(scan_i32(counts, counts, (int64_t)nblocks * 256, false, nullptr, s));
        radix_scatter_kernel<<<nblocks, kSortThreads, 0, s>>>(ksrc, vsrc, kdst, vdst, n, shift,
                                                              counts, nblocks);
        HMG_HIP(hipGetLastError());
        std::swap(ksrc, kdst);
        std::swap(vsrc, vdst);
    }
    if (ksrc != keys) {
        HMG_HIP(hipMemcpyAsync(keys, ksrc, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        HMG_HIP(hipMemcpyAsync(vals, vsrc, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    return HMG_OK;
}

hmg_status minmax_f64(const double* x, int64_t n, double* dmin, double* dmax, hipStream_t s)
{
    if (n <= 0) return HMG_ERR_INVALID;
    int nb = (int)std::min<int64_t>(1024, (n + kRedThreads - 1) / kRedThreads);
    void* p;
    HMG_TRY(scratch(s, 4, sizeof(double) * 2 * (size_t)nb, &p));
    double* pmin = (double*)p;
    double* pmax = pmin + nb;
    minmax_partial_kernel<<<nb, kRedThreads, 0, s>>>(x, n, pmin, pmax);
    HMG_HIP(hipGetLastError());
    minmax_final_kernel<<<1, kRedThreads, 0, s>>>(pmin, pmax, nb, dmin, dmax);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status flag_differ_f64(const double* a, const double* b, int64_t n, int* differ, hipStream_t s)
{
    if (n <= 0) return HMG_OK;
    const int nb = (int)std::min<int64_t>(1024, (n + kRedThreads - 1) / kRedThreads);
    differ_kernel<<<nb, kRedThreads, 0, s>>>((const uint64_t*)a, (const uint64_t*)b, n, differ);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

}  // namespace hmg
