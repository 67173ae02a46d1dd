// api.hip -- the C ABI: hmglib's entry points (hmglib.h:104-120,
// kernel_system_assembler.h:66-102) plus hmg_* status twins and extensions.
// Host orchestration only; every computation runs in the HIP kernels of
// prims/morton/tree/aca/mvp.hip on the engine's stream.  All calls are
// synchronous, as in the reference.
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstddef>

#include <mutex>

#include "hmg_internal.h"
#include "../../include/hmglib_server.h"

namespace hmg {

namespace {
thread_local std::string g_error;
}

thread_local hipStream_t g_alloc_stream = nullptr;

namespace {
struct LargeBlock {
    int dev;
    void* p;
    size_t bytes;
};
std::mutex g_large_mu;
std::vector<LargeBlock> g_large;  // released large buffers, still mapped
}  // namespace

namespace {
constexpr size_t kPinChunk = (size_t)4 << 20;
// this thread's pinned, device-mapped staging buffer (host and device addresses)
thread_local void* t_pin = nullptr;
thread_local void* t_pin_dev = nullptr;
thread_local size_t t_pin_bytes = 0;

hmg_status pinned(size_t bytes, void** host, void** dev)
{
    if (t_pin_bytes < bytes) {
        if (t_pin) (void)hipHostFree(t_pin);
        t_pin = t_pin_dev = nullptr;
        t_pin_bytes = 0;
        HMG_HIP(hipHostMalloc(&t_pin, bytes, hipHostMallocMapped | hipHostMallocCoherent));
        HMG_HIP(hipHostGetDevicePointer(&t_pin_dev, t_pin, 0));
        t_pin_bytes = bytes;
    }
    *host = t_pin;
    *dev = t_pin_dev;
    return HMG_OK;
}

// dst[0, n) = src[0, n) bytes; 16 bytes per thread and step when both ends and n
// are 16-byte aligned
__global__ void __launch_bounds__(256) copy_bytes_kernel(const unsigned char* __restrict__ src,
                                                         unsigned char* __restrict__ dst, size_t n, int wide)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x, step = (size_t)gridDim.x * 256;
    if (wide) {
        const uint4* s4 = (const uint4*)src;
        uint4* d4 = (uint4*)dst;
        for (size_t i = t; i < n / 16; i += step) d4[i] = s4[i];
    } else {
        for (size_t i = t; i < n; i += step) dst[i] = src[i];
    }
}

hmg_status copy_bytes(void* dst, const void* src, size_t n, hipStream_t s)
{
    const int wide = (((uintptr_t)dst | (uintptr_t)src | n) & 15) == 0;
    const size_t units = wide ? n / 16 : n;
    const unsigned grid = (unsigned)std::min<size_t>((units + 255) / 256, 1024);
    copy_bytes_kernel<<<grid, 256, 0, s>>>((const unsigned char*)src, (unsigned char*)dst, n, wide);
    return hipGetLastError() == hipSuccess ? HMG_OK : HMG_ERR_HIP;
}
}  // namespace

// The copies between the caller's host memory and the device run as a copy kernel
// reading or writing the thread's device-mapped pinned buffer: the runtime's own
// copy paths for pageable memory (staging, and CPU reads of small device buffers
// through the host mapping) are never used.
hmg_status upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return HMG_OK;
    void *pin, *pin_dev;
    HMG_TRY(pinned(std::min(bytes, kPinChunk), &pin, &pin_dev));
    for (size_t off = 0; off < bytes; off += kPinChunk) {
        const size_t n = std::min(kPinChunk, bytes - off);
        std::memcpy(pin, (const char*)src + off, n);
        const hmg_status ec = copy_bytes((char*)dst + off, pin_dev, n, s);
        const hipError_t es = hipStreamSynchronize(s);  // also on error: the buffer is reused
        if (ec != HMG_OK) {
            set_error("host-to-device copy launch failed");
            return ec;
        }
        HMG_HIP(es);
    }
    return HMG_OK;
}

hmg_status download_bytes(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return HMG_OK;
    void *pin, *pin_dev;
    HMG_TRY(pinned(std::min(bytes, kPinChunk), &pin, &pin_dev));
    for (size_t off = 0; off < bytes; off += kPinChunk) {
        const size_t n = std::min(kPinChunk, bytes - off);
        const hmg_status ec = copy_bytes(pin_dev, (const char*)src + off, n, s);
        const hipError_t es = hipStreamSynchronize(s);
        if (ec != HMG_OK) {
            set_error("device-to-host copy launch failed");
            return ec;
        }
        HMG_HIP(es);
        std::memcpy((char*)dst + off, pin, n);
    }
    return HMG_OK;
}

hipError_t large_alloc(void** p, size_t bytes, size_t* held)
{
    *p = nullptr;
    *held = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        // best fit among the cached buffers of this device no larger than twice the
        // request
        std::lock_guard<std::mutex> lk(g_large_mu);
        size_t best = SIZE_MAX, at = g_large.size();
        for (size_t i = 0; i < g_large.size(); i++) {
            const LargeBlock& b = g_large[i];
            if (b.dev == dev && b.bytes >= bytes && b.bytes / 2 <= bytes && b.bytes < best) {
                best = b.bytes;
                at = i;
            }
        }
        if (at < g_large.size()) {
            *p = g_large[at].p;
            *held = g_large[at].bytes;
            g_large.erase(g_large.begin() + (ptrdiff_t)at);
            return hipSuccess;
        }
    }
    e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        reclaim_device_memory();
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    *held = bytes;
    return hipSuccess;
}

void large_free(void* p, size_t held)
{
    if (!p) return;
    int dev = 0;
    // the buffer may be handed to work on another stream next: wait for the device
    // as hipFree does
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    std::lock_guard<std::mutex> lk(g_large_mu);
    g_large.push_back(LargeBlock{dev, p, held});
}

void reclaim_device_memory()
{
    large_trim(false);
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipDeviceSynchronize() == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
        (void)hipMemPoolTrimTo(pool, 0);
    (void)hipGetLastError();
}

void large_trim(bool every)
{
    int dev = 0;
    if (!every && hipGetDevice(&dev) != hipSuccess) return;
    std::vector<LargeBlock> out;
    {
        std::lock_guard<std::mutex> lk(g_large_mu);
        std::vector<LargeBlock> keep;
        for (const LargeBlock& b : g_large) (every || b.dev == dev ? out : keep).push_back(b);
        g_large.swap(keep);
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (const LargeBlock& b : out) {
        if (b.dev != cur) (void)hipSetDevice(b.dev);
        (void)hipFree(b.p);
        if (b.dev != cur) (void)hipSetDevice(cur);
    }
}

void set_error(const std::string& msg) { g_error = msg; }

Engine* engine_of(const h_matrix_data* d) { return d ? (Engine*)d->engine : nullptr; }

void Profiler::begin(int cls, hipStream_t s)
{
    if (!on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev[cls].push_back(e);
}

void Profiler::end(int cls, hipStream_t s)
{
    if (!on || ev[cls].size() % 2 == 0) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev[cls].push_back(e);
}

void Profiler::clear()
{
    for (auto& v : ev) {
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
        v.clear();
    }
}

KernelParams make_kernel_params(const system_assembler& a, int dim)
{
    KernelParams kp{};
    kp.kind = a.kind;
    kp.dim = dim;
    kp.reg = a.regularization;
    if (a.kind == HMG_KERNEL_GAUSSIAN) kp.c0 = 1.0 / (2.0 * a.length_scale * a.length_scale);
    if (a.kind == HMG_KERNEL_EXPONENTIAL) kp.c0 = 1.0 / a.length_scale;
    kp.c1 = std::pow(2.0, (double)dim / 2.0) * std::tgamma(1.0 + (double)dim / 2.0);
    return kp;
}

Coords coords_of(const h_matrix_data* d, int set)
{
    Coords c{};
    for (int k = 0; k < d->dim && k < MAX_POINT_DIMENSION; k++) c.x[k] = d->coords_d[set][k];
    c.n = d->point_count[set];
    return c;
}

hmg_status read_kernel(const h_matrix_data* d, KernelParams* kp)
{
    if (!d->assem) {
        set_error("h_matrix_data.assem is not set (create a kernel object first)");
        return HMG_ERR_NOT_READY;
    }
    system_assembler a{};
    // on the engine's own (non-blocking) stream: a null-stream copy would wait
    // for every other H-matrix's work on the device
    Engine* e = engine_of(d);
    hipStream_t s = e ? e->stream : nullptr;
    HMG_TRY(download_bytes(&a, d->assem, sizeof(a), s));
    if (a.kind < HMG_KERNEL_GAUSSIAN_REF || a.kind > HMG_KERNEL_EXPONENTIAL) {
        set_error("unknown kernel kind " + std::to_string(a.kind));
        return HMG_ERR_INVALID;
    }
    if ((a.kind == HMG_KERNEL_GAUSSIAN || a.kind == HMG_KERNEL_EXPONENTIAL) && !(a.length_scale > 0)) {
        set_error("kernel length scale must be positive");
        return HMG_ERR_INVALID;
    }
    *kp = make_kernel_params(a, d->dim);
    return HMG_OK;
}

hmg_status check_ready(const h_matrix_data* d, Engine** out)
{
    Engine* e = engine_of(d);
    if (!e) {
        set_error("h_matrix_data not initialised (call init_h_matrix_data)");
        return HMG_ERR_NOT_READY;
    }
    if (!e->built) {
        set_error("H-matrix not set up (call setup_h_matrix)");
        return HMG_ERR_NOT_READY;
    }
    *out = e;
    return HMG_OK;
}

namespace {

// compute_current_dense_work_size_magma (linear_algebra.cu:4059-4108), int64 inside
int dense_work_size_at(const std::vector<int4>& bl, int ndense, int64_t max_size, int start)
{
    int cnt = 0;
    int64_t size = 0;
    while (size <= max_size && start + cnt < ndense) {
        const int4 b = bl[start + cnt];
        size += (int64_t)(b.y - b.x + 1) * (b.w - b.z + 1);
        cnt++;
    }
    if (size > max_size) cnt--;
    return cnt;
}

// compute_current_aca_work_size (linear_algebra.cu:4158-4174)
int aca_work_size_at(const std::vector<int4>& bl, int nblocks, int64_t max_size, int start)
{
    int cnt = 0;
    int64_t size = 0;
    while (size <= max_size && start + cnt < nblocks) {
        const int4 b = bl[start + cnt];
        size += (int64_t)(b.y - b.x + 1);
        cnt++;
    }
    if (size > max_size) cnt--;
    return cnt;
}

}  // namespace

void release_factors(Engine& e, h_matrix_data* d)
{
    e.epoch++;
    e.U.release();
    e.V.release();
    e.Up.release();
    e.Vp.release();
    e.Tpan.release();
    e.tpan_ready = false;
    e.aca_k = -1;
    d->U = nullptr;
    d->V = nullptr;
}

namespace {

void release_build(Engine& e, h_matrix_data* d)
{
    release_factors(e, d);
    e.A.release();
    e.aoff.release();
    e.dense_stored = false;
    d->dA = nullptr;
    e.blocks.release();
    e.blk.release();
    e.h_blk.clear();
    e.kb.release();
    e.keff.release();
    e.aca_list.release();
    e.uoff.release();
    e.voff.release();
    e.toff.release();
    e.h_toff.clear();
    e.h_uoff.clear();
    e.h_voff.clear();
    e.h_kb.clear();
    e.h_dgrp_ptr.clear();
    e.h_dgrp_blk.clear();
    e.h_agrp_ptr.clear();
    e.h_agrp_blk.clear();
    e.h_cgrp_ptr.clear();
    e.h_cgrp_blk.clear();
    e.h_ldv.clear();
    e.ldv.release();
    e.lv_items.release();
    e.lv_slices.release();
    e.tmap.release();
    e.Tp.release();
    e.n_lv_items = e.n_lv_slices = 0;
    e.near_items.release();
    e.near_blk.release();
    e.h_near_items.clear();
    e.fly_items.release();
    e.fly_refs.release();
    e.h_fly_items.clear();
    e.h_fly_refs.clear();
    e.h_col_groups.clear();
    e.n_fly_items = 0;
    e.n_col_slots = 0;
    e.near_sym = false;
    e.lr_items.release();
    e.ldu.release();
    e.h_ldu.clear();
    for (Combine* c : {&e.comb_stored, &e.comb_fly}) {
        c->ptr.release();
        c->idx.release();
        c->groups.release();
        c->total = 0;
        c->slots = 0;
    }
    e.P.release();
    e.xs.release();
    e.T.release();
    for (DevBuf<double>* b : {&e.X64, &e.Tp64, &e.T64, &e.P64, &e.Ys64}) b->release();
    e.n_near_items = e.n_lr_items = 0;
    e.mx_items.release();
    e.mx_wg.release();
    e.h_mx_wg.clear();
    e.mx_refs.release();
    e.h_mx_items.clear();
    e.h_mx_refs.clear();
    e.n_mx_items = 0;
    e.mx_evals = 0;
    e.mx_sched = false;
    e.comb_mx.ptr.release();
    e.comb_mx.idx.release();
    e.comb_mx.groups.release();
    e.comb_mx.total = 0;
    e.Tp32.release();
    e.mxp.release();
    e.P32.release();
    e.Ys32.release();
    e.inv1.release();
    e.inv1_ready = false;
    e.Xp.release();
    e.Xt32.release();
    e.cpan[0].release();
    e.cpan[1].release();
    e.cpan_ready = false;
    e.layout_prec = HMG_PREC_FP64;
    e.built = false;
    if (d->mat_vec_data) {
        free(d->mat_vec_data);
        d->mat_vec_data = nullptr;
    }
    free(d->dense_work_size);
    free(d->aca_work_size);
    d->dense_work_size = nullptr;
    d->aca_work_size = nullptr;
}

}  // namespace

// the tolerance stop's constant from the caller's epsilon and eta
void set_tolerance(Engine& e, const h_matrix_data* d)
{
    e.aca_tol_c = e.aca_tol_mode == 1 ? d->epsilon * (1.0 - d->eta) / (1.0 + d->epsilon) : d->epsilon;
}

namespace {

// FP64 ACA factors into (U, V, keff) for the current block list and rank k

hmg_status build_factors(Engine& e, h_matrix_data* d, const KernelParams& kp, int k, int tie_last,
                         DevBuf<double>& U, DevBuf<double>& V)
{
    set_tolerance(e, d);
    HMG_TRY(aca_layout(e, k, HMG_PREC_FP64));
    const int na = e.naca;
    HMG_TRY(U.alloc(e.h_uoff[na] + 256));
    HMG_TRY(V.alloc(e.h_voff[na] + 256));
    HMG_HIP(hipMemsetAsync(U.p, 0, sizeof(double) * U.n, e.stream));
    HMG_HIP(hipMemsetAsync(V.p, 0, sizeof(double) * V.n, e.stream));
    HMG_TRY(aca_build(e, coords_of(d, 0), coords_of(d, 1), kp, k, tie_last, U.p, V.p, e.keff.p));
    HMG_HIP(hipStreamSynchronize(e.stream));
    return HMG_OK;
}

// mixed precision: the same FP64 ACA, its factors split into the BF16 hi + lo
// panels e.Up / e.Vp
hmg_status build_factors_mixed(Engine& e, h_matrix_data* d, const KernelParams& kp, int k, int tie_last)
{
    set_tolerance(e, d);
    HMG_TRY(aca_layout(e, k, HMG_PREC_MIXED));
    HMG_TRY(e.Up.alloc((size_t)e.up_size));
    HMG_TRY(e.Vp.alloc((size_t)e.vp_size));
    HMG_HIP(hipMemsetAsync(e.Up.p, 0, sizeof(__bf16) * e.Up.n, e.stream));
    HMG_HIP(hipMemsetAsync(e.Vp.p, 0, sizeof(__bf16) * e.Vp.n, e.stream));
    HMG_TRY(aca_build_mixed(e, coords_of(d, 0), coords_of(d, 1), kp, tie_last, e.Up.p, e.Vp.p));
    return HMG_OK;
}

}  // namespace
}  // namespace hmg

using namespace hmg;

extern "C" {

const char* hmg_last_error(void) { return g_error.c_str(); }

void hmg_release_cached_memory(void) { large_trim(true); }

// ------------------------------------------------------------------ init
hmg_status hmg_init_h_matrix_data(struct h_matrix_data* data, int point_count[2], int dim, int bits)
{
    return engine_init(data, point_count, dim, bits, nullptr);
}

}  // extern "C"

namespace hmg {

void engine_bind_stream(Engine* e, hipStream_t s, h_matrix_data* data)
{
    e->stream = e->stream2 = s;
    data->magma_queue = (void*)s;
}

hmg_status engine_init(h_matrix_data* data, int point_count[2], int dim, int bits, hipStream_t ext)
{
    if (!data || !point_count) {
        set_error("null argument");
        return HMG_ERR_INVALID;
    }
    if (dim < 1 || dim > MAX_POINT_DIMENSION) {
        set_error("dim must be in 1..5 (MAX_POINT_DIMENSION)");
        return HMG_ERR_INVALID;
    }
    if (point_count[0] < 1 || point_count[1] < 1) {
        set_error("point counts must be positive");
        return HMG_ERR_INVALID;
    }
    if (bits < 1 || bits > 64) {
        set_error("bits must be in 1..64");
        return HMG_ERR_INVALID;
    }
    Engine* e = new Engine();
    if (ext) {
        // work goes to a stream the caller owns (and may rebind between jobs)
        e->own_streams = false;
        e->stream = e->stream2 = ext;
    } else {
        // non-blocking streams: no implicit ordering with the null stream, so the
        // work of independent H-matrices driven from different host threads overlaps
        HMG_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        HMG_HIP(hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking));
    }
    HMG_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    HMG_HIP(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
    e->dim = dim;
    e->bits = bits;
    // keep freed pool memory for reuse (temporaries of the tree build and the
    // products come and go), up to a bound
    {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = (uint64_t)8 << 30;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    }
    if (const char* v = getenv("HMG_NEAR_LDS_KB")) e->near_lds_pad = std::max(0, atoi(v)) * 1024;
    if (const char* v = getenv("HMG_SERIALIZE")) e->serialize = atoi(v) != 0;
    data->engine = e;
    data->dim = dim;
    data->bits = bits;
    data->magma_queue = (void*)e->stream;
    data->U = data->V = data->dA = nullptr;
    data->mat_vec_data = nullptr;
    data->mat_vec_data_count = 0;
    data->mat_vec_data_array_size = 0;
    data->dense_work_size = data->aca_work_size = nullptr;
    data->dense_batch_count = data->aca_batch_count = 0;
    data->root_level_set_1 = 0;
    data->root_level_set_2 = 0;
    data->assem = nullptr;
    data->mat_vec_info = mat_vec_data_info{0, 0, 0};
    for (int i = 0; i < 2; i++) {
        const int n = point_count[i];
        e->n[i] = n;
        data->point_count[i] = n;
        data->coords_d[i] = (double**)calloc(dim, sizeof(double*));
        for (int d = 0; d < dim; d++)
            HMG_HIP(hipMallocAsync((void**)&data->coords_d[i][d], sizeof(double) * n, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->point_ids_d[i], sizeof(unsigned) * n, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->max_per_dim_d[i], sizeof(double) * dim, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->min_per_dim_d[i], sizeof(double) * dim, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->coords_device[i], sizeof(double*) * dim, e->stream));
        HMG_HIP(hipMemcpyAsync(data->coords_device[i], data->coords_d[i], sizeof(double*) * dim,
                               hipMemcpyHostToDevice, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->code_d[i], sizeof(uint64_t) * n, e->stream));
        HMG_HIP(hipMallocAsync((void**)&data->order[i], sizeof(uint64_t) * n, e->stream));
        point_set ps{data->coords_device[i], dim, n, data->max_per_dim_d[i], data->min_per_dim_d[i],
                     data->point_ids_d[i]};
        HMG_HIP(hipMallocAsync((void**)&data->points_d[i], sizeof(point_set), e->stream));
        HMG_TRY(upload_bytes(data->points_d[i], &ps, sizeof(ps), e->stream));
        morton_code mc{data->code_d[i], n, dim, bits, nullptr, nullptr};
        HMG_HIP(hipMallocAsync((void**)&data->morton_d[i], sizeof(morton_code), e->stream));
        HMG_TRY(upload_bytes(data->morton_d[i], &mc, sizeof(mc), e->stream));
    }
    return HMG_OK;
}

}  // namespace hmg

extern "C" {

// ------------------------------------------------------------------ setup
hmg_status hmg_setup_h_matrix(struct h_matrix_data* data)
{
    Engine* e = engine_of(data);
    if (!e) {
        set_error("h_matrix_data not initialised (call init_h_matrix_data)");
        return HMG_ERR_NOT_READY;
    }
    if (data->max_level < 0 || data->c_leaf < 0) {
        set_error("max_level and c_leaf must be non-negative");
        return HMG_ERR_INVALID;
    }
    AllocScope scope(e->stream);
    release_build(*e, data);
    hipStream_t s = e->stream;
    const int dim = data->dim;
    const bool spread_defined = (dim == 3 && data->bits == 20) || (dim == 2 && data->bits == 32) ||
                                (dim == 10 && data->bits == 4);
    const int key_bits = spread_defined ? std::min(64, dim * data->bits) : 8;
    // Both sets bitwise equal as given (one point cloud against itself): every
    // step below is deterministic, so the second set's extrema, codes, order and
    // sorted coordinates are copies of the first set's.
    bool same_input = false;
    if (data->point_count[0] == data->point_count[1] && data->point_count[0] > 0) {
        DevBuf<int> flag;
        HMG_TRY(flag.alloc(1));
        HMG_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
        for (int d = 0; d < dim; d++)
            HMG_TRY(flag_differ_f64(data->coords_d[0][d], data->coords_d[1][d], data->point_count[0],
                                    flag.p, s));
        int differ = 1;
        HMG_TRY(download_bytes(&differ, flag.p, sizeof(int), s));
        same_input = differ == 0;
    }
    for (int i = 0; i < 2; i++) {
        const int n = data->point_count[i];
        if (i == 1 && same_input) {
            HMG_HIP(hipMemcpyAsync(data->min_per_dim_d[1], data->min_per_dim_d[0], sizeof(double) * dim,
                                   hipMemcpyDeviceToDevice, s));
            HMG_HIP(hipMemcpyAsync(data->max_per_dim_d[1], data->max_per_dim_d[0], sizeof(double) * dim,
                                   hipMemcpyDeviceToDevice, s));
            HMG_HIP(hipMemcpyAsync(data->code_d[1], data->code_d[0], sizeof(uint64_t) * n,
                                   hipMemcpyDeviceToDevice, s));
            HMG_HIP(hipMemcpyAsync(data->order[1], data->order[0], sizeof(uint64_t) * n,
                                   hipMemcpyDeviceToDevice, s));
            HMG_HIP(hipMemcpyAsync(data->point_ids_d[1], data->point_ids_d[0], sizeof(unsigned) * n,
                                   hipMemcpyDeviceToDevice, s));
            for (int d = 0; d < dim; d++)
                HMG_HIP(hipMemcpyAsync(data->coords_d[1][d], data->coords_d[0][d], sizeof(double) * n,
                                       hipMemcpyDeviceToDevice, s));
            break;
        }
        // extremal values (helper.cu:108-138)
        for (int d = 0; d < dim; d++)
            HMG_TRY(minmax_f64(data->coords_d[i][d], n, data->min_per_dim_d[i] + d,
                               data->max_per_dim_d[i] + d, s));
        double mn[MAX_POINT_DIMENSION], mx[MAX_POINT_DIMENSION];
        HMG_TRY(download_bytes(mn, data->min_per_dim_d[i], sizeof(double) * dim, s));
        HMG_TRY(download_bytes(mx, data->max_per_dim_d[i], sizeof(double) * dim, s));
        for (int d = 0; d < dim; d++) {
            if (!(mx[d] > mn[d])) {
                set_error("point set " + std::to_string(i) + " has no extent in dimension " +
                          std::to_string(d) + " (Morton quantisation would divide by zero)");
                return HMG_ERR_INVALID;
            }
        }
        // Morton codes (morton.cu:127-155), stable sort (helper.cu:140-167)
        Coords c = coords_of(data, i);
        HMG_TRY(morton_codes(c, dim, data->bits, data->min_per_dim_d[i], data->max_per_dim_d[i],
                             data->code_d[i], s));
        HMG_TRY(e->perm32[i].alloc(n));
        HMG_TRY(iota_u32(e->perm32[i].p, n, s));
        HMG_TRY(radix_sort_u64(data->code_d[i], e->perm32[i].p, n, key_bits, s));
        HMG_TRY(perm_to_u64(e->perm32[i].p, data->order[i], data->point_ids_d[i], n, s));
        // reorder_point_set (helper.cu:293-328): gather into a temporary, copy back
        DevBuf<double> tmp;
        HMG_TRY(tmp.alloc(n));
        for (int d = 0; d < dim; d++) {
            HMG_TRY(gather_f64(data->coords_d[i][d], e->perm32[i].p, tmp.p, n, s));
            HMG_HIP(hipMemcpyAsync(data->coords_d[i][d], tmp.p, sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, s));
        }
        HMG_HIP(hipStreamSynchronize(s));
        tmp.release();
        e->perm32[i].release();
    }
    // the same point set on both sides (sorted coordinates bitwise equal) makes
    // the kernel matrix bitwise symmetric outside the diagonal
    e->same_sets = same_input;
    if (!same_input && data->point_count[0] == data->point_count[1] && data->point_count[0] > 0) {
        DevBuf<int> flag;
        HMG_TRY(flag.alloc(1));
        HMG_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
        for (int d = 0; d < dim; d++)
            HMG_TRY(flag_differ_f64(data->coords_d[0][d], data->coords_d[1][d], data->point_count[0],
                                    flag.p, s));
        int differ = 1;
        HMG_TRY(download_bytes(&differ, flag.p, sizeof(int), s));
        flag.release();
        e->same_sets = differ == 0;
    }
    // block-cluster tree (tree.cu:1429-1557) and DENSE-first partition
    TreeParams tp{data->eta, data->max_level, data->c_leaf, (int)data->root_level_set_1,
                  (int)data->root_level_set_2, e->exact_boxes};
    HMG_TRY(build_tree(*e, coords_of(data, 0), coords_of(data, 1), data->code_d[0], data->code_d[1], tp));
    return finish_setup(data, e);
}

}  // extern "C"

namespace hmg {

void release_build_state(Engine& e, h_matrix_data* d) { release_build(e, d); }

hmg_status finish_setup(h_matrix_data* data, Engine* e)
{
    AllocScope scope(e->stream);
    e->epoch++;
    data->mat_vec_data = (work_item**)malloc(sizeof(work_item*));
    *data->mat_vec_data = e->blocks.p;
    data->mat_vec_data_count = e->nblocks;
    data->mat_vec_data_array_size = e->nblocks;
    data->max_elements_in_array = -1;
    data->max_elements_in_mat_vec_data_array = -1;
    data->mat_vec_info.dense_count = e->ndense;
    data->mat_vec_info.aca_count = e->naca;
    data->mat_vec_info.total_count = e->nblocks;

    // batch sizing with the reference's semantics (linear_algebra.cu:4176-4288,
    // hmglib.cu:211-219), computed in 64-bit and clamped into the int fields
    int64_t required = 0;
    for (int b = 0; b < e->ndense; b++) {
        const int4 bl = e->h_blk[b];
        required += (int64_t)(bl.y - bl.x + 1) * (bl.w - bl.z + 1);
    }
    const int64_t predicted = (int64_t)((double)required * 1.05);
    data->max_batched_dense_size = (int)std::min<int64_t>(predicted, INT_MAX);
    std::vector<int> dws, aws;
    {
        int idx = 0;
        int w = dense_work_size_at(e->h_blk, e->ndense, predicted, idx);
        dws.push_back(w);
        while (w > 0) {
            idx += w;
            w = dense_work_size_at(e->h_blk, e->ndense, predicted, idx);
            dws.push_back(w);
        }
        idx = e->ndense;
        const int64_t amax = data->max_batched_aca_size;
        w = aca_work_size_at(e->h_blk, e->nblocks, amax, idx);
        aws.push_back(w);
        while (w > 0) {
            idx += w;
            w = aca_work_size_at(e->h_blk, e->nblocks, amax, idx);
            aws.push_back(w);
        }
    }
    data->dense_batch_count = (int)dws.size();
    data->aca_batch_count = (int)aws.size();
    data->dense_work_size = (int*)malloc(sizeof(int) * dws.size());
    data->aca_work_size = (int*)malloc(sizeof(int) * aws.size());
    memcpy(data->dense_work_size, dws.data(), sizeof(int) * dws.size());
    memcpy(data->aca_work_size, aws.data(), sizeof(int) * aws.size());

    // rank layout for k (k_b = min(k,m1,m2)) and the MVP schedule
    HMG_TRY(build_groups(*e));
    HMG_TRY(aca_layout(*e, std::max(data->k, 0)));
    e->built = true;
    return HMG_OK;
}

}  // namespace hmg

extern "C" {

// ------------------------------------------------------------------ precompute
hmg_status hmg_precompute_aca(struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    release_factors(*e, data);
    // The reference skips the whole far field when the first ACA block exceeds
    // max_batched_aca_size (hmglib.cu:232, linear_algebra.cu:4163-4171); the
    // engine has no batch cap and always builds every ACA block.
    if (e->precision == HMG_PREC_MIXED) {
        HMG_TRY(build_factors_mixed(*e, data, kp, std::max(data->k, 0), 1));
        data->U = (double*)e->Up.p;  // BF16 hi + lo panels in mixed precision
        data->V = (double*)e->Vp.p;
    } else {
        HMG_TRY(build_factors(*e, data, kp, std::max(data->k, 0), 1, e->U, e->V));
        data->U = e->U.p;
        data->V = e->V.p;
    }
    e->fact_prec = e->precision;
    e->epoch++;
    return HMG_OK;
}

hmg_status hmg_precompute_dense(struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    int64_t entries = 0;
    for (int b = 0; b < e->ndense; b++) {
        const int4 bl = e->h_blk[b];
        entries += (int64_t)(bl.y - bl.x + 1) * (bl.w - bl.z + 1);
    }
    size_t free_b = 0, total_b = 0;
    HMG_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = 8.0 * (double)entries;
    if (need > (double)free_b * 0.95) {
        set_error("dense blocks need " + std::to_string(need / 1e9) + " GB but only " +
                  std::to_string(free_b / 1e9) + " GB of device memory are free; "
                  "precomputing the dense blocks is impossible (apply evaluates them on the fly)");
        return HMG_ERR_OOM;
    }
    e->A.release();
    HMG_TRY(dense_fill(*e, coords_of(data, 0), coords_of(data, 1), kp));
    data->dA = e->A.p;
    return HMG_OK;
}

// ------------------------------------------------------------------ apply
}  // extern "C"

namespace hmg {

hmg_status apply_enqueue(h_matrix_data* data, Engine* e, const KernelParams& kp, const double* X, double* Y,
                         int nrhs)
{
    if (!(data->U != nullptr && e->aca_k >= 0)) {
        set_error("product of precomputed factors needs precompute_aca first");
        return HMG_ERR_NOT_READY;
    }
    if (e->fact_prec == HMG_PREC_MIXED)
        return hmvp_mixed(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], X, Y,
                          nrhs);
    const bool pre_dense = data->dA != nullptr && e->dense_stored;
    if (nrhs >= e->fp64_block_min && !pre_dense)
        return hmvp_multi(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], X, Y,
                          nrhs, e->U.p, e->V.p);
    return hmvp(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], X, Y, nrhs,
                e->U.p, e->V.p, e->keff.p, pre_dense);
}

std::vector<const void*> product_state(const h_matrix_data* d, const Engine& e)
{
    return {d->U, d->V, d->dA, e.U.p, e.V.p, e.Up.p, e.Vp.p, e.A.p, e.keff.p, e.blk.p, e.lr_items.p,
            e.lv_items.p, e.lv_slices.p, e.tmap.p, e.near_items.p, e.near_blk.p, e.fly_items.p, e.fly_refs.p,
            e.mx_items.p, e.mx_refs.p, e.mx_wg.p, e.vw_items.p, e.uw_items.p, e.comb_stored.idx.p,
            e.comb_fly.idx.p, e.comb_mx.idx.p, e.xs.p, e.Tp.p, e.T.p, e.P.p, e.X64.p, e.Tp64.p, e.T64.p,
            e.P64.p, e.Ys64.p, e.inv1.p, e.Xp.p, e.Xt32.p, e.Tpan.p, e.Tp32.p, e.P32.p, e.Ys32.p,
            e.cpan[0].p, e.cpan[1].p, e.mxp.p, d->assem, (const void*)(uintptr_t)e.epoch};
}

}  // namespace hmg

extern "C" {

hmg_status hmg_apply_h_matrix_mvp_multi(const double* X, double* Y, int nrhs, struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    if (!X || !Y || nrhs < 1) {
        set_error("invalid x / y / nrhs");
        return HMG_ERR_INVALID;
    }
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    const bool pre_aca = data->U != nullptr && e->aca_k >= 0;
    // stored factors decide the product's precision; otherwise the option does
    const int prec = pre_aca ? e->fact_prec : e->precision;
    if (prec == HMG_PREC_MIXED) {
        if (!pre_aca) {
            HMG_TRY(build_factors_mixed(*e, data, kp, std::max(data->k, 0), 1));
            e->fact_prec = HMG_PREC_MIXED;
        }
        HMG_TRY(hmvp_mixed(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1],
                           X, Y, nrhs));
        HMG_HIP(hipStreamSynchronize(e->stream));
        if (!pre_aca) {
            e->Up.release();
            e->Vp.release();
            e->aca_k = -1;
        }
        return HMG_OK;
    }
    const bool pre_dense = data->dA != nullptr && e->dense_stored;
    DevBuf<double> tU, tV;
    const double* U = e->U.p;
    const double* V = e->V.p;
    if (!pre_aca) {
        // no precomputed factors: rebuild the ACA for this product (h_matrix_mvp
        // without precompute, linear_algebra.cu:4359-4370)
        HMG_TRY(build_factors(*e, data, kp, std::max(data->k, 0), 1, tU, tV));
        U = tU.p;
        V = tV.p;
    }
    if (nrhs >= e->fp64_block_min && !pre_dense)
        HMG_TRY(hmvp_multi(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], X, Y,
                           nrhs, U, V));
    else
        HMG_TRY(hmvp(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], X, Y,
                     nrhs, U, V, e->keff.p, pre_dense));
    HMG_HIP(hipStreamSynchronize(e->stream));
    tU.release();
    tV.release();
    if (!pre_aca) e->aca_k = -1;
    return HMG_OK;
}

// Launch shapes of the product timed on the caller's operands (Y is overwritten);
// the fastest is kept in the engine's options.  Candidates: the near field's LDS
// reservation (0 to 80 KB: workgroups per CU beside the low-rank kernels) with
// the two streams overlapped, and both parts on one stream; with mixed-precision
// factors each of these with the low-rank products and the near field each in
// workgroup and in one-wavefront form.  The product's values do not depend on the choice
// (every partial is summed in a fixed order).
hmg_status hmg_autotune_product(struct h_matrix_data* data, const double* X, double* Y, int nrhs, int reps,
                                int64_t out[4])
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    if (!X || !Y || nrhs < 1 || reps < 1 || !out) {
        set_error("invalid x / y / nrhs / reps / out");
        return HMG_ERR_INVALID;
    }
    if (data->U == nullptr || e->aca_k < 0) {
        set_error("autotuning the product needs precompute_aca first");
        return HMG_ERR_NOT_READY;
    }
    struct Cand {
        int lds;
        bool serial;
        bool wg;       // low-rank products in workgroup form
        bool near_wg;  // near field in workgroup form
    };
    const bool mixed = e->fact_prec == HMG_PREC_MIXED;
    std::vector<Cand> cands;
    for (const bool wg : {true, false}) {
        for (const bool nwg : {true, false}) {
            if (!mixed && !(wg && nwg)) continue;  // the FP64 product has one form of each
            for (const int lds : {0, 16 << 10, 24 << 10, 36 << 10, 48 << 10, 64 << 10, 80 << 10})
                cands.push_back(Cand{lds, false, wg, nwg});
            cands.push_back(Cand{0, true, wg, nwg});
        }
    }
    const Cand keep{e->near_lds_pad, e->serialize, e->lr_wg, e->near_wg};
    auto set = [&](const Cand& c) {
        e->near_lds_pad = c.lds;
        e->serialize = c.serial;
        e->lr_wg = c.wg;
        e->near_wg = c.near_wg;
    };
    Cand best_c = keep;
    double best = 1e300;
    for (const Cand& c : cands) {
        set(c);
        hmg_status st = hmg_apply_h_matrix_mvp_multi(X, Y, nrhs, data);  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; st == HMG_OK && r < reps; r++) st = hmg_apply_h_matrix_mvp_multi(X, Y, nrhs, data);
        if (st != HMG_OK) {
            set(keep);
            return st;
        }
        const double us =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / reps;
        if (us < best) {
            best = us;
            best_c = c;
        }
    }
    set(best_c);
    out[0] = best_c.lds;
    out[1] = best_c.serial ? 1 : 0;
    out[2] = (int64_t)best;
    out[3] = (best_c.wg ? 1 : 0) | (best_c.near_wg ? 2 : 0);
    return HMG_OK;
}

hmg_status hmg_apply_h_matrix_mvp(double* x, double* y, struct h_matrix_data* data)
{
    return hmg_apply_h_matrix_mvp_multi(x, y, 1, data);
}

hmg_status hmg_apply_h_matrix_mvp_without_batching(double* x, double* y, struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    if (!x || !y) {
        set_error("invalid x / y");
        return HMG_ERR_INVALID;
    }
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    // ignores precomputed data (hmglib.cu:294); unbatched pivot rule = first argmax
    DevBuf<double> tU, tV;
    const bool had = data->U != nullptr;
    std::vector<int> keep_keff;
    if (had && e->naca > 0) {
        keep_keff.resize(e->naca);
        HMG_HIP(hipMemcpyAsync(keep_keff.data(), e->keff.p, sizeof(int) * e->naca, hipMemcpyDeviceToHost,
                               e->stream));
        HMG_HIP(hipStreamSynchronize(e->stream));
    }
    const int had_k = e->aca_k;
    HMG_TRY(build_factors(*e, data, kp, std::max(data->k, 0), 0, tU, tV));
    HMG_TRY(hmvp(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1], x, y,
                 1, tU.p, tV.p, e->keff.p, false));
    HMG_HIP(hipStreamSynchronize(e->stream));
    tU.release();
    tV.release();
    if (had) {
        // restore the stored factors' layout and ranks
        HMG_TRY(aca_layout(*e, had_k, e->fact_prec));
        if (e->naca > 0)
        {
            HMG_HIP(hipMemcpyAsync(e->keff.p, keep_keff.data(), sizeof(int) * e->naca, hipMemcpyHostToDevice,
                                   e->stream));
            HMG_HIP(hipStreamSynchronize(e->stream));
        }
    } else {
        e->aca_k = -1;
    }
    return HMG_OK;
}

hmg_status hmg_apply_full_mvp(double* x, double* y, struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    HMG_TRY(full_mvp(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[0], data->order[1],
                     x, y, 1));
    HMG_HIP(hipStreamSynchronize(e->stream));
    return HMG_OK;
}

hmg_status hmg_dense_rows(const double* X, int nrhs, const int64_t* rows, int nrows, double* Yrows,
                          struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    if (nrows <= 0) return HMG_OK;
    KernelParams kp;
    HMG_TRY(read_kernel(data, &kp));
    DevBuf<int64_t> inv;
    HMG_TRY(inv.alloc(data->point_count[0]));
    HMG_TRY(inverse_perm(data->order[0], inv.p, data->point_count[0], e->stream));
    HMG_TRY(dense_rows(*e, coords_of(data, 0), coords_of(data, 1), kp, data->order[1], X, nrhs, rows,
                       nrows, Yrows, inv.p));
    HMG_HIP(hipStreamSynchronize(e->stream));
    inv.release();
    return HMG_OK;
}

hmg_status hmg_set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data)
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    AllocScope scope(e->stream);
    HMG_TRY(gaussian_rhs(*e, coords_of(data, 0), data->order[0], b));
    HMG_HIP(hipStreamSynchronize(e->stream));
    return HMG_OK;
}

hmg_status hmg_get_work(const struct h_matrix_data* data, int64_t out[9])
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    int64_t en = 0, el1 = 0, el2 = 0, rsum = 0;
    for (int b = 0; b < e->ndense; b++) {
        const int4 bl = e->h_blk[b];
        en += (int64_t)(bl.y - bl.x + 1) * (bl.w - bl.z + 1);
    }
    std::vector<int> ranks(e->h_kb);
    if (data->U && e->aca_k >= 0 && e->naca > 0)
    {
        HMG_HIP(hipMemcpyAsync(ranks.data(), e->keff.p, sizeof(int) * e->naca, hipMemcpyDeviceToHost, e->stream));
        HMG_HIP(hipStreamSynchronize(e->stream));
    }
    for (int a = 0; a < e->naca; a++) {
        const int4 bl = e->h_blk[e->ndense + a];
        el1 += (int64_t)(bl.y - bl.x + 1) * ranks[a];
        el2 += (int64_t)(bl.w - bl.z + 1) * ranks[a];
        rsum += ranks[a];
    }
    out[0] = en;
    out[1] = el1 + el2;
    out[2] = e->ndense;
    out[3] = e->naca;
    out[4] = e->levels;
    out[5] = e->quirk_levels;
    out[6] = el1;
    out[7] = el2;
    out[8] = rsum;
    return HMG_OK;
}

hmg_status hmg_get_schedule(const struct h_matrix_data* data, int64_t out[11])
{
    Engine* e;
    HMG_TRY(check_ready(data, &e));
    std::vector<LrItem> items((size_t)e->n_lr_items);
    if (!items.empty())
    {
        HMG_HIP(hipMemcpyAsync(items.data(), e->lr_items.p, sizeof(LrItem) * items.size(),
                               hipMemcpyDeviceToHost, e->stream));
        HMG_HIP(hipStreamSynchronize(e->stream));
    }
    int64_t rows = 0, treads = 0, m1s = 0, m2s = 0;
    for (const LrItem& it : items) {
        rows += it.nrows;
        treads += it.K;
    }
    for (int a = 0; a < e->naca; a++) {
        const int4 b = e->h_blk[e->ndense + a];
        m1s += b.y - b.x + 1;
        m2s += b.w - b.z + 1;
    }
    const bool mixed = e->layout_prec == HMG_PREC_MIXED;
    const bool stored = !mixed && data->dA != nullptr && e->dense_stored;
    out[0] = mixed ? e->n_mx_items : stored ? e->n_near_items : e->n_fly_items;
    out[1] = e->n_lr_items;
    out[2] = rows;
    out[3] = treads;
    out[4] = m1s;
    out[5] = m2s;
    const Combine& cb = mixed ? e->comb_mx : stored ? e->comb_stored : e->comb_fly;
    out[6] = cb.total - cb.group_entries + cb.slots;
    out[7] = (!mixed && e->near_sym) ? 1 : 0;
    out[8] = mixed ? 0 : e->n_col_slots;
    out[9] = e->same_sets ? 1 : 0;
    out[10] = mixed ? e->mx_evals : stored ? 0 : e->near_evals;
    return HMG_OK;
}

extern "C++" {
namespace {

template <class T>
int64_t bytes_of(const DevBuf<T>& b)
{
    if (!b.p) return 0;
    return b.held ? (int64_t)b.held : (int64_t)b.n * (int64_t)sizeof(T);
}

int64_t engine_bytes(const Engine& e)
{
    int64_t t = 0;
    t += bytes_of(e.perm32[0]) + bytes_of(e.perm32[1]) + bytes_of(e.blocks) + bytes_of(e.blk);
    t += bytes_of(e.kb) + bytes_of(e.keff) + bytes_of(e.ldv) + bytes_of(e.ldu) + bytes_of(e.aca_list);
    t += bytes_of(e.uoff) + bytes_of(e.voff) + bytes_of(e.toff) + bytes_of(e.U) + bytes_of(e.V);
    t += bytes_of(e.aoff) + bytes_of(e.A) + bytes_of(e.near_items) + bytes_of(e.near_blk);
    t += bytes_of(e.fly_items) + bytes_of(e.fly_refs) + bytes_of(e.lr_items) + bytes_of(e.lv_items);
    t += bytes_of(e.lv_slices) + bytes_of(e.tmap) + bytes_of(e.Tp) + bytes_of(e.P) + bytes_of(e.xs);
    t += bytes_of(e.T) + bytes_of(e.Up) + bytes_of(e.Vp) + bytes_of(e.mx_items) + bytes_of(e.mx_refs);
    t += bytes_of(e.Tpan) + bytes_of(e.Tp32) + bytes_of(e.P32) + bytes_of(e.Xp) + bytes_of(e.Xt32) + bytes_of(e.mxp);
    t += bytes_of(e.Ys32) + bytes_of(e.inv1);
    t += bytes_of(e.cpan[0]) + bytes_of(e.cpan[1]);
    t += bytes_of(e.X64) + bytes_of(e.Tp64) + bytes_of(e.T64) + bytes_of(e.P64) + bytes_of(e.Ys64);
    t += bytes_of(e.vw_items) + bytes_of(e.uw_items) + bytes_of(e.mx_wg);
    for (const Combine* c : {&e.comb_stored, &e.comb_fly, &e.comb_mx})
        t += bytes_of(c->ptr) + bytes_of(c->idx) + bytes_of(c->groups);
    return t;
}

}  // namespace
}  // extern "C++"

hmg_status hmg_memory(const struct h_matrix_data* data, int64_t out[2])
{
    Engine* e = engine_of(data);
    if (!e || !out) {
        set_error("h_matrix_data not initialised");
        return HMG_ERR_NOT_READY;
    }
    // buffers of init_h_matrix_data: coordinates, ids, extents, codes, orders
    int64_t base = 0;
    for (int i = 0; i < 2; i++)
        base += (int64_t)data->point_count[i] * (8 * data->dim + 4 + 8 + 8) + 16 * data->dim + 256;
    out[0] = base + engine_bytes(*e);
    out[1] = 0;
    if (!e->built) return HMG_OK;
    const int na = e->naca;
    const bool have_factors = e->aca_k >= 0 && data->U != nullptr;
    const bool mixed = e->precision == HMG_PREC_MIXED;
    int64_t add = 0;
    if (!have_factors && na > 0) {
        // factor elements of the current layout; the mixed layout pads to 32 floats
        // instead of 16 doubles
        const double elems = (double)(e->h_uoff[na] + e->h_voff[na]);
        // mixed: BF16 hi + lo panels (4 bytes per element, padded to whole MFMA
        // tiles) plus the FP64 scratch of the ACA batches
        add += mixed ? (int64_t)(elems * 4.0 * 1.25) + 8 * aca_mixed_scratch(*e) : (int64_t)(elems * 8.0);
    }
    if (mixed && e->layout_prec != HMG_PREC_MIXED) {
        const int64_t n1 = e->n[0], n2 = e->n[1];
        const int64_t g1 = (n1 + 7) / 8 + 4, g2 = (n2 + 7) / 8 + 4;
        // near-field items: 128-row slices x 4096-column chunks per dense row cluster
        int64_t mx_items = 0;
        const int ng = (int)e->h_dgrp_ptr.size() - 1;
        for (int g = 0; g < ng; g++) {
            int64_t cols = 0;
            int m1 = 0;
            for (int q = e->h_dgrp_ptr[g]; q < e->h_dgrp_ptr[g + 1]; q++) {
                const int4 b = e->h_blk[e->h_dgrp_blk[q]];
                cols += b.w - b.z + 1;
                m1 = b.y - b.x + 1;
            }
            mx_items += (int64_t)((m1 + 127) / 128) * ((cols + 4095) / 4096 + 1);
        }
        // partial V^T X slots of column clusters longer than one 4096-point chunk
        int64_t ntp = 0;
        const int ncg = (int)e->h_cgrp_ptr.size() - 1;
        for (int g = 0; g < ncg; g++) {
            int K = 0;
            for (int q = e->h_cgrp_ptr[g]; q < e->h_cgrp_ptr[g + 1]; q++) K += e->h_kb[e->h_cgrp_blk[q]];
            const int4 b = e->h_blk[e->ndense + e->h_cgrp_blk[e->h_cgrp_ptr[g]]];
            const int64_t nch = (b.w - b.z + 1 + 4095) / 4096;
            if (nch > 1) ntp += nch * ((K + 127) / 128);
        }
        const int64_t slots = (int64_t)e->n_lr_items * 64 + mx_items * 128;
        add += g2 * 2 * kMxRhs * 8 * 2 + (int64_t)e->n[1] * kMxRhs * 4;         // Xp, Xt32
        add += (e->tsize + 32 * (int64_t)ncg) * kMxRhs * 4 + ntp * kMxVCols * kMxRhs * 4;  // Tpan, Tp32
        add += slots * kMxRhs * 4 + slots * 4 + n1 * 4;              // P32, combine list
        add += n1 * kMxRhs * 4 + n1 * 4;                             // Ys32, inv1
        add += (g1 + g2) * e->dim * 8 * 4 + mx_items * (16 + 16);    // panels, items
    }
    if (!mixed && e->X64.n == 0) {
        // FP64 products of up to 32 right-hand sides per pass: sorted right-hand-side
        // rows, V^T X chunk partials, T, row partials, sorted results, inverse order
        const int64_t slots = (int64_t)e->n_lr_items * 64 + (int64_t)e->n_near_items * 128;
        add += ((int64_t)e->n[0] + e->n[1] + e->tsize + slots + (int64_t)e->n_lv_items * 64) * kRbMax * 8 +
               (int64_t)e->n[0] * 4;
    }
    out[1] = add + add / 10 + (64 << 20);
    return HMG_OK;
}

hmg_status hmg_set_option(struct h_matrix_data* data, int option, int64_t value)
{
    Engine* e = engine_of(data);
    if (!e) {
        set_error("h_matrix_data not initialised");
        return HMG_ERR_NOT_READY;
    }
    e->epoch++;
    switch (option) {
    case HMG_OPT_NEAR_SYMMETRIC:
        e->near_sym_enabled = value != 0;
        return HMG_OK;
    case HMG_OPT_NEAR_LDS_BYTES:
        if (value < 0 || value > 128 * 1024) {
            set_error("near-field LDS reservation must be in [0, 131072] bytes");
            return HMG_ERR_INVALID;
        }
        e->near_lds_pad = (int)value;
        return HMG_OK;
    case HMG_OPT_SERIALIZE:
        e->serialize = value != 0;
        return HMG_OK;
    case HMG_OPT_PRECISION:
        if (value != HMG_PREC_FP64 && value != HMG_PREC_MIXED) {
            set_error("precision must be HMG_PREC_FP64 or HMG_PREC_MIXED");
            return HMG_ERR_INVALID;
        }
        e->precision = (int)value;
        return HMG_OK;
    case HMG_OPT_ACA_ROW_SCAN:
        e->aca_row_scan = value != 0;
        return HMG_OK;
    case HMG_OPT_ACA_TOLERANCE:
        if (value < 0 || value > 2) {
            set_error("ACA tolerance mode must be 0, 1 or 2");
            return HMG_ERR_INVALID;
        }
        e->aca_tol_mode = (int)value;
        return HMG_OK;
    case HMG_OPT_ACA_BIG_BLOCK:
        if (value < 0 || value > (int64_t)1 << 30) {
            set_error("ACA workgroup threshold must be in [0, 2^30]");
            return HMG_ERR_INVALID;
        }
        e->aca_big_entries = (int)value;
        return HMG_OK;
    case HMG_OPT_ACA_HUGE_BLOCK:
        if (value < 0 || value > (int64_t)1 << 30) {
            set_error("ACA 8-wavefront threshold must be in [0, 2^30]");
            return HMG_ERR_INVALID;
        }
        e->aca_huge_entries = (int)value;
        return HMG_OK;
    case HMG_OPT_LR_WORKGROUP:
        e->lr_wg = value != 0;
        return HMG_OK;
    case HMG_OPT_EXACT_BOXES:
        e->exact_boxes = value != 0;
        return HMG_OK;
    case HMG_OPT_NEAR_WORKGROUP:
        e->near_wg = value != 0;
        return HMG_OK;
    case HMG_OPT_FP64_BLOCK_RHS:
        if (value < 1 || value > (int64_t)1 << 30) {
            set_error("FP64 block threshold must be at least 1");
            return HMG_ERR_INVALID;
        }
        e->fp64_block_min = (int)value;
        return HMG_OK;
    case HMG_OPT_LR_PRIORITY:
        if (value < 0 || value > 3) {
            set_error("low-rank priority must be 0..3 (bit 0: V^T X, bit 1: U T)");
            return HMG_ERR_INVALID;
        }
        e->lr_prio = (int)value;
        return HMG_OK;
    case HMG_OPT_ACA_LDS_BYTES:
        if (value < 0 || value > 96 * 1024) {
            set_error("ACA LDS reservation must be in 0..98304 bytes");
            return HMG_ERR_INVALID;
        }
        e->aca_lds_pad = (int)value;
        return HMG_OK;
    case HMG_OPT_MIXED_ACA_PARTS:
        if (value < 1 || value > 4096) {
            set_error("mixed ACA parts must be in 1..4096");
            return HMG_ERR_INVALID;
        }
        e->aca_mixed_parts = (int)value;
        return HMG_OK;
    default:
        set_error("unknown option " + std::to_string(option));
        return HMG_ERR_INVALID;
    }
}

hmg_status hmg_set_profiling(struct h_matrix_data* data, int enable)
{
    Engine* e = engine_of(data);
    if (!e) {
        set_error("h_matrix_data not initialised");
        return HMG_ERR_NOT_READY;
    }
    HMG_HIP(hipStreamSynchronize(e->stream));
    e->prof.clear();
    e->prof.on = enable != 0;
    return HMG_OK;
}

hmg_status hmg_get_profile(const struct h_matrix_data* data, int kernel_class, double* total_ms,
                           int64_t* launches)
{
    Engine* e = engine_of(data);
    if (!e) {
        set_error("h_matrix_data not initialised");
        return HMG_ERR_NOT_READY;
    }
    if (kernel_class < 0 || kernel_class >= PROF_NCLASS) {
        set_error("unknown kernel class");
        return HMG_ERR_INVALID;
    }
    HMG_HIP(hipStreamSynchronize(e->stream));
    const auto& v = e->prof.ev[kernel_class];
    double tot = 0.0;
    int64_t n = 0;
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
        float ms = 0.f;
        HMG_HIP(hipEventElapsedTime(&ms, v[i], v[i + 1]));
        tot += ms;
        n++;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return HMG_OK;
}

// ------------------------------------------------------------------ destroy
hmg_status hmg_destroy_h_matrix_data(struct h_matrix_data* data)
{
    Engine* e = engine_of(data);
    if (!e) return HMG_OK;
    // finish the engine's work, then return every buffer on its own stream
    // (stream-ordered frees do not wait for other H-matrices' kernels).  With a
    // caller-owned stream the caller has finished this H-matrix's work already.
    if (e->own_streams) {
        (void)hipStreamSynchronize(e->stream2);
        (void)hipStreamSynchronize(e->stream);
    }
    hipStream_t s = e->stream;
    {
        AllocScope scope(s);
        release_build(*e, data);
        for (int i = 0; i < 2; i++) {
            for (int d = 0; d < data->dim; d++) (void)hipFreeAsync(data->coords_d[i][d], s);
            free(data->coords_d[i]);
            data->coords_d[i] = nullptr;
            for (void* p : {(void*)data->point_ids_d[i], (void*)data->max_per_dim_d[i],
                            (void*)data->min_per_dim_d[i], (void*)data->coords_device[i], (void*)data->code_d[i],
                            (void*)data->order[i], (void*)data->points_d[i], (void*)data->morton_d[i]})
                (void)hipFreeAsync(p, s);
            data->point_ids_d[i] = nullptr;
            data->max_per_dim_d[i] = data->min_per_dim_d[i] = nullptr;
            data->coords_device[i] = nullptr;
            data->code_d[i] = nullptr;
            data->order[i] = nullptr;
            data->points_d[i] = nullptr;
            data->morton_d[i] = nullptr;
            e->perm32[i].release();
        }
        // scratch of caller-owned streams stays with their owner (another job may
        // be using it right now)
        if (e->own_streams) {
            release_scratch(e->stream);
            release_scratch(e->stream2);
        }
    }
    e->prof.clear();
    if (e->own_streams) {
        (void)hipStreamSynchronize(e->stream2);
        (void)hipStreamSynchronize(e->stream);
    }
    (void)hipEventDestroy(e->ev_fork);
    (void)hipEventDestroy(e->ev_join);
    if (e->own_kernel) {
        if (data->assem == e->own_kernel) data->assem = nullptr;
        hmg_kernel_free(e->own_kernel);
    }
    // buffers the engine still holds go back on their streams, which live until then
    const bool own = e->own_streams;
    hipStream_t st = e->stream, st2 = e->stream2;
    delete e;
    data->engine = nullptr;
    data->magma_queue = nullptr;
    if (own) {
        (void)hipStreamSynchronize(st2);
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st2);
        (void)hipStreamDestroy(st);
    }
    return HMG_OK;
}

// ------------------------------------------------------------------ kernels
int hmg_create_kernel_object(struct system_assembler** assem, int kind, double regularization,
                             double length_scale)
{
    if (!assem) {
        set_error("null slot");
        return HMG_ERR_INVALID;
    }
    system_assembler* dptr = hmg_kernel_new(kind, regularization, length_scale);
    if (!dptr) return HMG_ERR_HIP;
    HMG_HIP(hipMemcpy(assem, &dptr, sizeof(dptr), hipMemcpyHostToDevice));
    return HMG_OK;
}

int hmg_destroy_kernel_object(struct system_assembler** assem)
{
    if (!assem) return HMG_OK;
    system_assembler* dptr = nullptr;
    HMG_HIP(hipMemcpy(&dptr, assem, sizeof(dptr), hipMemcpyDeviceToHost));
    hmg_kernel_free(dptr);
    return HMG_OK;
}

struct system_assembler* hmg_kernel_new(int kind, double regularization, double length_scale)
{
    if (kind < HMG_KERNEL_GAUSSIAN_REF || kind > HMG_KERNEL_EXPONENTIAL) {
        set_error("unknown kernel kind");
        return nullptr;
    }
    system_assembler h{};
    h.kind = kind;
    h.regularization = regularization;
    h.length_scale = length_scale;
    system_assembler* dptr = nullptr;
    if (hipMalloc((void**)&dptr, sizeof(h)) != hipSuccess) {
        set_error("hipMalloc of kernel object failed");
        return nullptr;
    }
    if (hipMemcpy(dptr, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dptr);
        set_error("hipMemcpy of kernel object failed");
        return nullptr;
    }
    return dptr;
}

void hmg_kernel_free(struct system_assembler* assem)
{
    if (assem) (void)hipFree(assem);
}

// ------------------------------------------------------------------ layout
int64_t hmg_struct_layout(int which, int field)
{
#define OFF(T, f) (int64_t)offsetof(T, f)
    switch (which) {
    case 0: {
        const int64_t offs[] = {
            OFF(h_matrix_data, coords_d), OFF(h_matrix_data, point_ids_d),
            OFF(h_matrix_data, max_per_dim_d), OFF(h_matrix_data, min_per_dim_d),
            OFF(h_matrix_data, coords_device), OFF(h_matrix_data, code_d),
            OFF(h_matrix_data, points_d), OFF(h_matrix_data, morton_d), OFF(h_matrix_data, order),
            OFF(h_matrix_data, bits), OFF(h_matrix_data, dim), OFF(h_matrix_data, eta),
            OFF(h_matrix_data, max_level), OFF(h_matrix_data, c_leaf),
            OFF(h_matrix_data, mat_vec_data_count), OFF(h_matrix_data, max_elements_in_array),
            OFF(h_matrix_data, max_elements_in_mat_vec_data_array), OFF(h_matrix_data, mat_vec_data),
            OFF(h_matrix_data, mat_vec_data_array_size), OFF(h_matrix_data, k),
            OFF(h_matrix_data, epsilon), OFF(h_matrix_data, mat_vec_info),
            OFF(h_matrix_data, point_count), OFF(h_matrix_data, assem),
            OFF(h_matrix_data, max_batched_dense_size), OFF(h_matrix_data, max_batched_aca_size),
            OFF(h_matrix_data, dense_batching_ratio), OFF(h_matrix_data, U), OFF(h_matrix_data, V),
            OFF(h_matrix_data, dA), OFF(h_matrix_data, magma_queue),
            OFF(h_matrix_data, dense_work_size), OFF(h_matrix_data, aca_work_size),
            OFF(h_matrix_data, dense_batch_count), OFF(h_matrix_data, aca_batch_count),
            OFF(h_matrix_data, root_level_set_1), OFF(h_matrix_data, root_level_set_2),
            OFF(h_matrix_data, engine)};
        if (field < 0) return (int64_t)sizeof(h_matrix_data);
        if (field < (int)(sizeof(offs) / sizeof(offs[0]))) return offs[field];
        return -1;
    }
    case 1: {
        const int64_t offs[] = {OFF(work_item, set1_l), OFF(work_item, set1_u), OFF(work_item, set2_l),
                                OFF(work_item, set2_u), OFF(work_item, work_type), OFF(work_item, max1),
                                OFF(work_item, min1), OFF(work_item, max2), OFF(work_item, min2),
                                OFF(work_item, dim), OFF(work_item, level_1), OFF(work_item, level_2)};
        if (field < 0) return (int64_t)sizeof(work_item);
        if (field < (int)(sizeof(offs) / sizeof(offs[0]))) return offs[field];
        return -1;
    }
    case 2:
        return field < 0 ? (int64_t)sizeof(point_set) : -1;
    case 3:
        return field < 0 ? (int64_t)sizeof(morton_code) : -1;
    case 4:
        return field < 0 ? (int64_t)sizeof(system_assembler) : -1;
    case 5:
        return field < 0 ? (int64_t)sizeof(mat_vec_data_info) : -1;
    case 6:
        return field < 0 ? (int64_t)sizeof(hmg_server_config) : -1;
    case 7: {
        const int64_t offs[] = {OFF(hmg_build_request, dim), OFF(hmg_build_request, bits),
                                OFF(hmg_build_request, n), OFF(hmg_build_request, coords),
                                OFF(hmg_build_request, eta), OFF(hmg_build_request, max_level),
                                OFF(hmg_build_request, c_leaf), OFF(hmg_build_request, k),
                                OFF(hmg_build_request, kernel_kind), OFF(hmg_build_request, regularization),
                                OFF(hmg_build_request, length_scale), OFF(hmg_build_request, epsilon),
                                OFF(hmg_build_request, aca_tolerance), OFF(hmg_build_request, exact_boxes)};
        if (field < 0) return (int64_t)sizeof(hmg_build_request);
        if (field < (int)(sizeof(offs) / sizeof(offs[0]))) return offs[field];
        return -1;
    }
    case 8: {
        const int64_t offs[] = {OFF(hmg_request_info, status), OFF(hmg_request_info, t_submit),
                                OFF(hmg_request_info, t_start), OFF(hmg_request_info, t_end),
                                OFF(hmg_request_info, merged), OFF(hmg_request_info, matrix),
                                OFF(hmg_request_info, message)};
        if (field < 0) return (int64_t)sizeof(hmg_request_info);
        if (field < (int)(sizeof(offs) / sizeof(offs[0]))) return offs[field];
        return -1;
    }
    }
    return -1;
#undef OFF
}

// ------------------------------------------------------------------ generator
static uint64_t sm64(uint64_t* st)
{
    uint64_t z = (*st += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed)
    {
        uint64_t st = seed;
        for (auto& w : s) w = sm64(&st);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

void hmg_rng_uniform(uint64_t seed, int64_t n, double* out)
{
    Xoshiro g(seed);
    for (int64_t i = 0; i < n; i++) out[i] = g.uniform();
}

void hmg_rng_normal(uint64_t seed, int64_t n, double* out)
{
    Xoshiro g(seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i += 2) {
        const double u1 = 1.0 - g.uniform();
        const double u2 = g.uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        out[i] = rad * std::cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = rad * std::sin(two_pi * u2);
    }
}

// ------------------------------------------------------------------ void drop-ins
static void die_on(hmg_status s, const char* where)
{
    if (s != HMG_OK) {
        fprintf(stderr, "hmglib error in %s: %s\n", where, hmg_last_error());
        exit(EXIT_FAILURE);
    }
}

void init_h_matrix_data(struct h_matrix_data* data, int point_count[2], int dim, int bits)
{
    die_on(hmg_init_h_matrix_data(data, point_count, dim, bits), "init_h_matrix_data");
}
void setup_h_matrix(struct h_matrix_data* data) { die_on(hmg_setup_h_matrix(data), "setup_h_matrix"); }
void apply_h_matrix_mvp(double* x, double* y, struct h_matrix_data* data)
{
    die_on(hmg_apply_h_matrix_mvp(x, y, data), "apply_h_matrix_mvp");
}
void apply_h_matrix_mvp_without_batching(double* x, double* y, struct h_matrix_data* data)
{
    die_on(hmg_apply_h_matrix_mvp_without_batching(x, y, data), "apply_h_matrix_mvp_without_batching");
}
void destroy_h_matrix_data(struct h_matrix_data* data)
{
    die_on(hmg_destroy_h_matrix_data(data), "destroy_h_matrix_data");
}
void precompute_aca(struct h_matrix_data* data) { die_on(hmg_precompute_aca(data), "precompute_aca"); }
void precompute_dense(struct h_matrix_data* data)
{
    die_on(hmg_precompute_dense(data), "precompute_dense");
}
void apply_full_mvp(double* x, double* y, struct h_matrix_data* data)
{
    die_on(hmg_apply_full_mvp(x, y, data), "apply_full_mvp");
}
void set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data)
{
    die_on(hmg_set_gaussian_kernel_rhs(b, data), "set_gaussian_kernel_rhs");
}

void create_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem,
                                                    double regularization)
{
    die_on((hmg_status)hmg_create_kernel_object((system_assembler**)assem, HMG_KERNEL_GAUSSIAN_REF,
                                                regularization, 1.0),
           "create_gaussian_kernel_system_assembler_object");
}
void destroy_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem)
{
    die_on((hmg_status)hmg_destroy_kernel_object((system_assembler**)assem),
           "destroy_gaussian_kernel_system_assembler_object");
}
void create_matern_kernel_system_assembler_object(struct matern_kernel_system_assembler** assem,
                                                  double regularization)
{
    die_on((hmg_status)hmg_create_kernel_object((system_assembler**)assem, HMG_KERNEL_MATERN_REF,
                                                regularization, 1.0),
           "create_matern_kernel_system_assembler_object");
}
void destroy_matern_kernel_system_assembler_object(struct matern_kernel_system_assembler** assem)
{
    die_on((hmg_status)hmg_destroy_kernel_object((system_assembler**)assem),
           "destroy_matern_kernel_system_assembler_object");
}

}  // extern "C"
