// batch.hip -- the ACA of many H-matrices (point clouds) in one launch set.
//
// Built one at a time, the H-matrices of many small point clouds leave the GPU
// mostly idle: the ACA of a cloud of a few thousand points is a few hundred
// wavefronts whose rank loops run for milliseconds, and the largest blocks of
// each cloud set the length of its own launch.  hmg_precompute_aca_batch treats
// the admissible blocks of every listed H-matrix that shares a kernel, a point
// dimension and a stopping rule as the blocks of ONE virtual H-matrix:
//   - the clouds' sorted coordinates side by side, one array per dimension and
//     set; cloud c starts at offset o_c in both sets, o_{c+1} = o_c + max(n1, n2),
//     so sorted indices coincide across the two sets exactly where they do inside
//     the cloud (the regularization term stays where it belongs);
//   - their ACA blocks, row and column ranges moved by o_c, their ranks, and
//     where their factors go (FP64: a pointer per block into its H-matrix's own
//     factor layout, written in place; mixed precision: the scratch layout of the
//     mixed ACA in one FP64 buffer of at most kBatchScratch elements).
// The single-H-matrix ACA kernels then run all of these blocks in one launch set
// (the largest on 8-wavefront teams beside the rest), so large blocks of one
// cloud run beside small blocks of the others.  Each H-matrix receives its ranks,
// and in mixed precision its factors split from its slice into its BF16 hi + lo
// panels.  Ranks, pivots and factors are those of precompute_aca on
// each H-matrix alone (same kernels, same per-block arithmetic).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <tuple>

#include "hmg_internal.h"

namespace hmg {
namespace {

// Everything precompute_aca does before its ACA launch, for one H-matrix: the
// kernel, the factor layout for rank k in the requested precision, zeroed factor
// storage.  Runs on the H-matrix's own stream (one host thread per H-matrix).
hmg_status prepare(h_matrix_data* d, KernelParams* kp)
{
    Engine* e;
    HMG_TRY(check_ready(d, &e));
    AllocScope scope(e->stream);
    HMG_TRY(read_kernel(d, kp));
    release_factors(*e, d);
    set_tolerance(*e, d);
    const int k = std::max(d->k, 0);
    if (e->precision == HMG_PREC_MIXED) {
        HMG_TRY(aca_layout(*e, k, HMG_PREC_MIXED));
        HMG_TRY(e->Up.alloc((size_t)e->up_size));
        HMG_TRY(e->Vp.alloc((size_t)e->vp_size));
        HMG_HIP(hipMemsetAsync(e->Up.p, 0, sizeof(__bf16) * e->Up.n, e->stream));
        HMG_HIP(hipMemsetAsync(e->Vp.p, 0, sizeof(__bf16) * e->Vp.n, e->stream));
    } else {
        HMG_TRY(aca_layout(*e, k, HMG_PREC_FP64));
        HMG_TRY(e->U.alloc(e->h_uoff[e->naca] + 256));
        HMG_TRY(e->V.alloc(e->h_voff[e->naca] + 256));
        HMG_HIP(hipMemsetAsync(e->U.p, 0, sizeof(double) * e->U.n, e->stream));
        HMG_HIP(hipMemsetAsync(e->V.p, 0, sizeof(double) * e->V.n, e->stream));
    }
    HMG_HIP(hipStreamSynchronize(e->stream));
    return HMG_OK;
}

// H-matrices whose blocks can share one launch: same kernel (kind, point
// dimension, parameters), stopping rule and precision
using GroupKey = std::tuple<int, int, double, double, double, int, double, int>;

GroupKey key_of(const Engine& e, const KernelParams& kp)
{
    return GroupKey{kp.kind, kp.dim, kp.reg, kp.c0, kp.c1, e.aca_tol_mode, e.aca_tol_c, e.precision};
}

// FP64 elements of the factors of ACA block a of e in the mixed-precision scratch
int64_t scratch_of(const Engine& e, int a)
{
    const int4 b = e.h_blk[e.ndense + a];
    return (int64_t)((b.y - b.x + 1) + (b.w - b.z + 1)) * e.h_kb[a];
}

// The blocks of members[lo, hi) of one group as one virtual H-matrix, its ACA,
// and each member's ranks and factors out of its slice.
hmg_status run_group(h_matrix_data* const* data, const std::vector<int>& members, const KernelParams& kp,
                     hipStream_t s, hipStream_t s2, hipEvent_t ev_fork, hipEvent_t ev_join)
{
    const int nm = (int)members.size();
    Engine* e0 = engine_of(data[members[0]]);
    const bool mixed = e0->precision == HMG_PREC_MIXED;
    const int dim = kp.dim;
    // offsets of the clouds' points, blocks and factors
    std::vector<int64_t> pofs(nm + 1, 0), aofs(nm + 1, 0), fofs(nm + 1, 0);
    for (int m = 0; m < nm; m++) {
        const Engine* e = engine_of(data[members[m]]);
        pofs[m + 1] = pofs[m] + std::max(e->n[0], e->n[1]);
        aofs[m + 1] = aofs[m] + e->naca;
        if (mixed) {
            int64_t t = 0;
            for (int a = 0; a < e->naca; a++) t += scratch_of(*e, a);
            fofs[m + 1] = fofs[m] + t;
        }
    }
    const int64_t npts = pofs[nm], na = aofs[nm];
    if (na == 0) return HMG_OK;
    if (npts > (int64_t)INT32_MAX - 1 || na > (int64_t)INT32_MAX) {
        set_error("batched ACA: too many points or blocks in one launch");
        return HMG_ERR_INVALID;
    }
    AllocScope scope(s);
    Engine v;  // the virtual H-matrix (owns only the buffers below; no streams of its own)
    v.stream = s;
    v.stream2 = s2;
    v.ev_fork = ev_fork;
    v.ev_join = ev_join;
    v.own_streams = false;
    v.dim = dim;
    v.n[0] = v.n[1] = (int)npts;
    v.ndense = 0;
    v.naca = (int)na;
    v.nblocks = (int)na;
    v.aca_row_scan = e0->aca_row_scan;
    v.aca_big_entries = e0->aca_big_entries;
    v.aca_huge_entries = e0->aca_huge_entries;
    v.aca_tol_mode = e0->aca_tol_mode;
    v.aca_tol_c = e0->aca_tol_c;
    v.serialize = false;
    // coordinates side by side (unused tails stay as they are: no block reaches them)
    DevBuf<double> xc[2][MAX_POINT_DIMENSION];
    Coords c[2];
    for (int set = 0; set < 2; set++) {
        c[set] = Coords{};
        c[set].n = (int)npts;
        for (int dd = 0; dd < dim; dd++) {
            HMG_TRY(xc[set][dd].alloc((size_t)npts));
            c[set].x[dd] = xc[set][dd].p;
        }
    }
    v.h_blk.resize((size_t)na);
    std::vector<int> kb((size_t)na), ldu((size_t)na), ldv((size_t)na);
    std::vector<int64_t> uoff((size_t)na, 0), voff((size_t)na, 0);
    // FP64: each block's factors are written in place, in its H-matrix's own layout
    std::vector<double*> ublk(mixed ? 0 : (size_t)na), vblk(mixed ? 0 : (size_t)na);
    HMG_TRY(v.blocks.alloc((size_t)na));
    for (int m = 0; m < nm; m++) {
        const h_matrix_data* d = data[members[m]];
        const Engine* e = engine_of(d);
        const int o = (int)pofs[m];
        for (int set = 0; set < 2; set++)
            for (int dd = 0; dd < dim; dd++)
                HMG_HIP(hipMemcpyAsync(xc[set][dd].p + o, d->coords_d[set][dd], sizeof(double) * e->n[set],
                                       hipMemcpyDeviceToDevice, s));
        if (e->naca > 0)
            HMG_HIP(hipMemcpyAsync(v.blocks.p + aofs[m], e->blocks.p + e->ndense, sizeof(work_item) * e->naca,
                                   hipMemcpyDeviceToDevice, s));
        int64_t run = fofs[m];
        for (int a = 0; a < e->naca; a++) {
            const int64_t q = aofs[m] + a;
            const int4 b = e->h_blk[e->ndense + a];
            v.h_blk[q] = make_int4(b.x + o, b.y + o, b.z + o, b.w + o);
            kb[q] = e->h_kb[a];
            if (mixed) {
                // the mixed ACA's scratch layout: U_b column-major (ld m1), then V_b
                // row-major (ld kb)
                const int m1 = b.y - b.x + 1;
                uoff[q] = run;
                voff[q] = run + (int64_t)m1 * kb[q];
                ldu[q] = m1;
                ldv[q] = std::max(kb[q], 1);
                run += scratch_of(*e, a);
            } else {
                ublk[q] = e->U.p + e->h_uoff[a];
                vblk[q] = e->V.p + e->h_voff[a];
                ldu[q] = e->h_ldu[a];
                ldv[q] = e->h_ldv[a];
            }
        }
    }
    v.h_kb = kb;
    HMG_TRY(upload(v.blk, v.h_blk, s));
    HMG_TRY(upload(v.kb, kb, s));
    HMG_TRY(upload(v.ldu, ldu, s));
    HMG_TRY(upload(v.ldv, ldv, s));
    HMG_TRY(upload(v.uoff, uoff, s));
    HMG_TRY(upload(v.voff, voff, s));
    HMG_TRY(v.keff.alloc((size_t)na));
    HMG_HIP(hipMemsetAsync(v.keff.p, 0, sizeof(int) * na, s));
    // the ACA of every block of the group (last argmax: the batched rule); mixed
    // precision: into an FP64 scratch (bounded by kBatchScratch), then split into
    // each H-matrix's panels
    DevBuf<double> U;
    DevBuf<double*> d_ublk, d_vblk;
    if (mixed) {
        HMG_TRY(U.alloc((size_t)std::max<int64_t>(fofs[nm], 1)));
        HMG_HIP(hipMemsetAsync(U.p, 0, sizeof(double) * U.n, s));
        HMG_TRY(aca_build(v, c[0], c[1], kp, 0, 1, U.p, U.p, v.keff.p));
    } else {
        HMG_TRY(upload(d_ublk, ublk, s));
        HMG_TRY(upload(d_vblk, vblk, s));
        HMG_TRY(aca_build(v, c[0], c[1], kp, 0, 1, nullptr, nullptr, v.keff.p, d_ublk.p, d_vblk.p));
    }
    // ranks and factors to their H-matrices
    DevBuf<int> iota;
    if (mixed) {
        int maxna = 0;
        for (int m = 0; m < nm; m++) maxna = std::max(maxna, engine_of(data[members[m]])->naca);
        std::vector<int> h(maxna);
        for (int q = 0; q < maxna; q++) h[q] = q;
        HMG_TRY(upload(iota, h, s));
    }
    for (int m = 0; m < nm; m++) {
        Engine* e = engine_of(data[members[m]]);
        if (e->naca == 0) continue;
        HMG_HIP(hipMemcpyAsync(e->keff.p, v.keff.p + aofs[m], sizeof(int) * e->naca, hipMemcpyDeviceToDevice, s));
        if (mixed)
            HMG_TRY(aca_to_panels(*e, iota.p, e->naca, v.uoff.p + aofs[m], v.voff.p + aofs[m], U.p, e->Up.p,
                                  e->Vp.p, s));
    }
    HMG_HIP(hipStreamSynchronize(s));
    return HMG_OK;
}

// A stream pair and its fork / join events, destroyed with the object
struct StreamPair {
    hipStream_t s = nullptr, s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ~StreamPair()
    {
        if (s) (void)hipStreamSynchronize(s);
        if (s2) (void)hipStreamSynchronize(s2);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (s2) (void)hipStreamDestroy(s2);
        if (s) (void)hipStreamDestroy(s);
    }
    hmg_status create()
    {
        HMG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HMG_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        HMG_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        HMG_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        return HMG_OK;
    }
};

// ACA launch sets of one hmg_precompute_aca_batch call that run at once
constexpr int kBatchSets = 2;
// the FP64 scratch elements one launch of the mixed path may use (16 GiB)
constexpr int64_t kBatchScratch = (int64_t)1 << 31;

}  // namespace
}  // namespace hmg

using namespace hmg;

namespace hmg {
namespace {

// f(i) for i < count on up to `threads` host threads; the first failure (by
// index) is returned with its message
template <class F>
hmg_status parallel_for(int count, int threads, F f, const char* what)
{
    std::vector<hmg_status> st((size_t)count, HMG_OK);
    std::vector<std::string> err((size_t)count);
    const int nt = std::max(1, std::min(threads > 0 ? threads : 8, count));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < count; i = next++) {
            st[i] = f(i);
            if (st[i] != HMG_OK) err[i] = hmg_last_error();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (int i = 0; i < count; i++)
        if (st[i] != HMG_OK) {
            set_error(std::string(what) + ", H-matrix " + std::to_string(i) + ": " + err[i]);
            return st[i];
        }
    return HMG_OK;
}

// ---- the setup of many clouds in one set of launches ---------------------------
// Cloud q's points sit at [off[q], off[q] + n[q]) of one array per dimension.
constexpr int kBT = 256;

// Both sets' coordinates side by side (the first set's copy), the cloud of every
// point, and whether a cloud's second set differs bitwise from its first
__global__ void __launch_bounds__(kBT)
gather_clouds_kernel(const double* const* __restrict__ src0, const double* const* __restrict__ src1,
                     const int* __restrict__ off, const int* __restrict__ n, int dim, Coords dst,
                     int* __restrict__ cid, int* __restrict__ differ)
{
    const int q = blockIdx.y;
    const int i = blockIdx.x * kBT + threadIdx.x;
    if (i >= n[q]) return;
    const int p = off[q] + i;
    cid[p] = q;
    bool diff = false;
    for (int d = 0; d < dim; d++) {
        const double* a = src0[q * dim + d];
        const double* b = src1[q * dim + d];
        const double v = a[i];
        ((double*)dst.x[d])[p] = v;
        if (a != b && __double_as_longlong(v) != __double_as_longlong(b[i])) diff = true;
    }
    if (diff) differ[q] = 1;
}

// extremal values per cloud and dimension (one workgroup per cloud)
__global__ void __launch_bounds__(kBT)
cloud_minmax_kernel(Coords x, const int* __restrict__ off, const int* __restrict__ n, int dim,
                    double* __restrict__ cmin, double* __restrict__ cmax)
{
    __shared__ double smin[kBT], smax[kBT];
    const int q = blockIdx.x;
    const int o = off[q], m = n[q];
    for (int d = 0; d < dim; d++) {
        double lo = __builtin_inf(), hi = -__builtin_inf();
        for (int i = threadIdx.x; i < m; i += kBT) {
            const double v = x.x[d][o + i];
            lo = v < lo ? v : lo;
            hi = hi < v ? v : hi;
        }
        smin[threadIdx.x] = lo;
        smax[threadIdx.x] = hi;
        __syncthreads();
        for (int st = kBT / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) {
                const double a = smin[threadIdx.x + st], b = smax[threadIdx.x + st];
                smin[threadIdx.x] = a < smin[threadIdx.x] ? a : smin[threadIdx.x];
                smax[threadIdx.x] = smax[threadIdx.x] < b ? b : smax[threadIdx.x];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            cmin[(int64_t)q * dim + d] = smin[0];
            cmax[(int64_t)q * dim + d] = smax[0];
        }
        __syncthreads();
    }
}

__global__ void cloud_key_kernel(const uint32_t* __restrict__ perm, const int* __restrict__ cid, int64_t n,
                                 uint64_t* __restrict__ key)
{
    const int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (i < n) key[i] = (uint64_t)cid[perm[i]];
}

// Sorted position i of cloud q: its point perm[off[q] + i] - off[q] of the cloud.
// Writes the sorted coordinates and codes side by side (xs, codes_s) and into
// both sets of the cloud's own arrays (code, order, point ids, coordinates,
// extrema), as setup_h_matrix leaves them.
struct CloudArrays {
    uint64_t* const* code[2];
    uint64_t* const* order[2];
    unsigned* const* ids[2];
    double* const* x[2];      // [q * dim + d]
    double* const* mn[2];
    double* const* mx[2];
};

__global__ void __launch_bounds__(kBT)
scatter_clouds_kernel(const uint32_t* __restrict__ perm, const uint64_t* __restrict__ codes, Coords xc,
                      const int* __restrict__ off, const int* __restrict__ n, int dim,
                      const double* __restrict__ cmin, const double* __restrict__ cmax, CloudArrays a,
                      Coords xs, uint64_t* __restrict__ codes_s)
{
    const int q = blockIdx.y;
    const int i = blockIdx.x * kBT + threadIdx.x;
    if (i >= n[q]) return;
    const int o = off[q];
    const uint32_t src = perm[o + i];
    const uint64_t code = codes[src];
    const uint64_t local = (uint64_t)(src - (uint32_t)o);
    codes_s[o + i] = code;
    for (int set = 0; set < 2; set++) {
        a.code[set][q][i] = code;
        a.order[set][q][i] = local;
        a.ids[set][q][i] = (unsigned)local;
    }
    for (int d = 0; d < dim; d++) {
        const double v = xc.x[d][src];
        ((double*)xs.x[d])[o + i] = v;
        a.x[0][q * dim + d][i] = v;
        a.x[1][q * dim + d][i] = v;
    }
    if (i < dim) {
        for (int set = 0; set < 2; set++) {
            a.mn[set][q][i] = cmin[(int64_t)q * dim + i];
            a.mx[set][q][i] = cmax[(int64_t)q * dim + i];
        }
    }
}

// The setup of every listed H-matrix in one set of launches per step (coordinates,
// extrema, Morton codes, one sort, the level-by-level traversal of all trees at
// once, the DENSE-first partition), then each H-matrix's schedule on host threads.
// Applies when every H-matrix has one point set against itself (bitwise equal
// sets), exact cluster boxes, root levels 0 and the same dimension, code width and
// tree parameters; *applied = false otherwise (nothing changed).  The result is
// setup_h_matrix's on each H-matrix alone: the same points, codes, stable order,
// boxes and BFS leaf order within each cloud.
hmg_status setup_fused(h_matrix_data* const* data, int count, int threads, bool* applied)
{
    *applied = false;
    if (count < 2 || count > 65535) return HMG_OK;
    const h_matrix_data* d0 = data[0];
    std::vector<Engine*> eng((size_t)count);
    std::vector<int> off((size_t)count), n((size_t)count);
    int64_t total = 0;
    for (int q = 0; q < count; q++) {
        const h_matrix_data* d = data[q];
        Engine* e = engine_of(d);
        if (!e || !e->exact_boxes || d->dim != d0->dim || d->bits != d0->bits || d->eta != d0->eta ||
            d->max_level != d0->max_level || d->c_leaf != d0->c_leaf || d->root_level_set_1 != 0 ||
            d->root_level_set_2 != 0 || d->point_count[0] != d->point_count[1] || d->point_count[0] < 1 ||
            d->max_level < 0 || d->c_leaf < 0)
            return HMG_OK;
        for (int j = 0; j < q; j++)
            if (data[j] == d) return HMG_OK;
        eng[q] = e;
        off[q] = (int)total;
        n[q] = d->point_count[0];
        total += n[q];
    }
    if (total > (int64_t)INT32_MAX - 1) return HMG_OK;
    const int dim = d0->dim, N = (int)total;
    const int nmax = *std::max_element(n.begin(), n.end());
    hipStream_t s = nullptr;
    HMG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard()
        {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{s};
    AllocScope scope(s);
    std::vector<const double*> src0((size_t)count * dim), src1((size_t)count * dim);
    for (int q = 0; q < count; q++)
        for (int k = 0; k < dim; k++) {
            src0[(size_t)q * dim + k] = data[q]->coords_d[0][k];
            src1[(size_t)q * dim + k] = data[q]->coords_d[1][k];
        }
    DevBuf<const double*> dsrc0, dsrc1;
    DevBuf<int> doff, dn, cid, differ;
    HMG_TRY(upload(dsrc0, src0, s));
    HMG_TRY(upload(dsrc1, src1, s));
    HMG_TRY(upload(doff, off, s));
    HMG_TRY(upload(dn, n, s));
    HMG_TRY(cid.alloc(N));
    HMG_TRY(differ.alloc(count));
    HMG_HIP(hipMemsetAsync(differ.p, 0, sizeof(int) * count, s));
    DevBuf<double> xc[MAX_POINT_DIMENSION], xs[MAX_POINT_DIMENSION];
    Coords cc{}, cs{};
    cc.n = cs.n = N;
    for (int k = 0; k < dim; k++) {
        HMG_TRY(xc[k].alloc(N));
        HMG_TRY(xs[k].alloc(N));
        cc.x[k] = xc[k].p;
        cs.x[k] = xs[k].p;
    }
    const dim3 g2((unsigned)((nmax + kBT - 1) / kBT), (unsigned)count);
    gather_clouds_kernel<<<g2, kBT, 0, s>>>(dsrc0.p, dsrc1.p, doff.p, dn.p, dim, cc, cid.p, differ.p);
    HMG_HIP(hipGetLastError());
    std::vector<int> hdiff((size_t)count);
    HMG_TRY(download_bytes(hdiff.data(), differ.p, sizeof(int) * count, s));
    for (int q = 0; q < count; q++)
        if (hdiff[q]) return HMG_OK;  // a cloud with two different sets: the per-cloud path
    // from here on the listed H-matrices are rebuilt
    *applied = true;
    for (int q = 0; q < count; q++) release_build_state(*eng[q], data[q]);
    DevBuf<double> cmin, cmax;
    HMG_TRY(cmin.alloc((size_t)count * dim));
    HMG_TRY(cmax.alloc((size_t)count * dim));
    cloud_minmax_kernel<<<(unsigned)count, kBT, 0, s>>>(cc, doff.p, dn.p, dim, cmin.p, cmax.p);
    HMG_HIP(hipGetLastError());
    {
        std::vector<double> hmin((size_t)count * dim), hmax((size_t)count * dim);
        HMG_TRY(download_bytes(hmin.data(), cmin.p, sizeof(double) * hmin.size(), s));
        HMG_TRY(download_bytes(hmax.data(), cmax.p, sizeof(double) * hmax.size(), s));
        for (int q = 0; q < count; q++)
            for (int k = 0; k < dim; k++)
                if (!(hmax[(size_t)q * dim + k] > hmin[(size_t)q * dim + k])) {
                    set_error("H-matrix " + std::to_string(q) + ": point set 0 has no extent in dimension " +
                              std::to_string(k) + " (Morton quantisation would divide by zero)");
                    return HMG_ERR_INVALID;
                }
    }
    // Morton codes on each cloud's box; one stable sort by code, then one by cloud:
    // within a cloud the order is the cloud's own stable sort by code
    DevBuf<uint64_t> codes, keys;
    DevBuf<uint32_t> perm;
    HMG_TRY(codes.alloc(N));
    HMG_TRY(keys.alloc(N));
    HMG_TRY(perm.alloc(N));
    HMG_TRY(morton_codes_batched(cc, dim, d0->bits, cid.p, cmin.p, cmax.p, codes.p, s));
    HMG_HIP(hipMemcpyAsync(keys.p, codes.p, sizeof(uint64_t) * N, hipMemcpyDeviceToDevice, s));
    HMG_TRY(iota_u32(perm.p, N, s));
    const bool spread_defined = (dim == 3 && d0->bits == 20) || (dim == 2 && d0->bits == 32) ||
                                (dim == 10 && d0->bits == 4);
    const int key_bits = spread_defined ? std::min(64, dim * d0->bits) : 8;
    HMG_TRY(radix_sort_u64(keys.p, perm.p, N, key_bits, s));
    cloud_key_kernel<<<(unsigned)((N + kBT - 1) / kBT), kBT, 0, s>>>(perm.p, cid.p, N, keys.p);
    HMG_HIP(hipGetLastError());
    int cbits = 1;
    while ((1ll << cbits) < count) cbits++;
    HMG_TRY(radix_sort_u64(keys.p, perm.p, N, cbits, s));
    // the sorted clouds side by side and in each H-matrix's own arrays
    std::vector<uint64_t*> pcode[2], pord[2];
    std::vector<unsigned*> pids[2];
    std::vector<double*> px[2], pmn[2], pmx[2];
    for (int set = 0; set < 2; set++) {
        for (int q = 0; q < count; q++) {
            h_matrix_data* d = data[q];
            pcode[set].push_back(d->code_d[set]);
            pord[set].push_back(d->order[set]);
            pids[set].push_back(d->point_ids_d[set]);
            pmn[set].push_back(d->min_per_dim_d[set]);
            pmx[set].push_back(d->max_per_dim_d[set]);
            for (int k = 0; k < dim; k++) px[set].push_back(d->coords_d[set][k]);
        }
    }
    DevBuf<uint64_t*> dcode[2], dord[2];
    DevBuf<unsigned*> dids[2];
    DevBuf<double*> dx[2], dmn[2], dmx[2];
    CloudArrays ca{};
    for (int set = 0; set < 2; set++) {
        HMG_TRY(upload(dcode[set], pcode[set], s));
        HMG_TRY(upload(dord[set], pord[set], s));
        HMG_TRY(upload(dids[set], pids[set], s));
        HMG_TRY(upload(dx[set], px[set], s));
        HMG_TRY(upload(dmn[set], pmn[set], s));
        HMG_TRY(upload(dmx[set], pmx[set], s));
        ca.code[set] = dcode[set].p;
        ca.order[set] = dord[set].p;
        ca.ids[set] = dids[set].p;
        ca.x[set] = dx[set].p;
        ca.mn[set] = dmn[set].p;
        ca.mx[set] = dmx[set].p;
    }
    DevBuf<uint64_t> codes_s;
    HMG_TRY(codes_s.alloc(N));
    scatter_clouds_kernel<<<g2, kBT, 0, s>>>(perm.p, codes.p, cc, doff.p, dn.p, dim, cmin.p, cmax.p, ca, cs,
                                             codes_s.p);
    HMG_HIP(hipGetLastError());
    // all trees in one traversal, each engine's DENSE-first block list
    const TreeParams tp{d0->eta, d0->max_level, d0->c_leaf, 0, 0, true};
    HMG_TRY(build_trees_batched(s, dim, cs, codes_s.p, tp, off, n, cid.p, eng.data()));
    HMG_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < count; q++) eng[q]->same_sets = true;
    // each H-matrix's schedule (host work and its uploads, on its own stream)
    return parallel_for(count, threads, [&](int q) { return finish_setup(data[q], eng[q]); }, "batched setup");
}

}  // namespace
}  // namespace hmg

extern "C" hmg_status hmg_setup_batch(struct h_matrix_data* const* data, int count, int threads)
{
    if (!data || count < 0) {
        set_error("invalid batch");
        return HMG_ERR_INVALID;
    }
    for (int q = 0; q < count; q++)
        if (!data[q] || !engine_of(data[q])) {
            set_error("batched setup: H-matrix " + std::to_string(q) + " not initialised");
            return HMG_ERR_NOT_READY;
        }
    // H-matrices that can share the fused setup, grouped by point dimension, code
    // width and tree parameters; the others are set up one by one on host threads
    using SetupKey = std::tuple<int, int, double, int, int>;
    std::map<SetupKey, std::vector<h_matrix_data*>> groups;
    std::vector<SetupKey> order;
    std::vector<h_matrix_data*> single;
    for (int q = 0; q < count; q++) {
        h_matrix_data* d = data[q];
        const Engine* e = engine_of(d);
        bool repeated = false;
        for (int j = 0; j < q; j++) repeated = repeated || data[j] == d;
        if (repeated || !e->exact_boxes || d->root_level_set_1 != 0 || d->root_level_set_2 != 0 ||
            d->point_count[0] != d->point_count[1]) {
            single.push_back(d);
            continue;
        }
        const SetupKey k{d->dim, d->bits, d->eta, d->max_level, d->c_leaf};
        if (!groups.count(k)) order.push_back(k);
        groups[k].push_back(d);
    }
    for (const SetupKey& k : order) {
        std::vector<h_matrix_data*>& g = groups[k];
        bool applied = false;
        HMG_TRY(setup_fused(g.data(), (int)g.size(), threads, &applied));
        if (!applied) single.insert(single.end(), g.begin(), g.end());
    }
    return parallel_for((int)single.size(), threads, [&](int i) { return hmg_setup_h_matrix(single[i]); },
                        "batched setup");
}

extern "C" hmg_status hmg_apply_batch(struct h_matrix_data* const* data, const double* const* X, double* const* Y,
                                      const int* nrhs, int count, int threads)
{
    if (!data || !X || !Y || !nrhs || count < 0) {
        set_error("invalid batch");
        return HMG_ERR_INVALID;
    }
    return parallel_for(
        count, threads, [&](int i) { return hmg_apply_h_matrix_mvp_multi(X[i], Y[i], nrhs[i], data[i]); },
        "batched apply");
}

extern "C" hmg_status hmg_precompute_aca_batch(struct h_matrix_data* const* data, int count, int threads)
{
    if (!data || count < 0) {
        set_error("invalid batch");
        return HMG_ERR_INVALID;
    }
    if (count == 0) return HMG_OK;
    for (int i = 0; i < count; i++) {
        Engine* e;
        HMG_TRY(check_ready(data[i], &e));
        for (int j = 0; j < i; j++)
            if (data[j] == data[i]) {
                set_error("batched ACA: an H-matrix is listed twice");
                return HMG_ERR_INVALID;
            }
    }
    // layouts and factor storage, one host thread per H-matrix at a time
    std::vector<KernelParams> kps((size_t)count);
    HMG_TRY(parallel_for(count, threads, [&](int i) { return prepare(data[i], &kps[i]); }, "batched ACA"));
    // groups of H-matrices that share a launch, in order of first appearance; the
    // mixed path splits a group where its FP64 scratch would pass kBatchScratch
    std::map<GroupKey, std::vector<int>> groups;
    std::vector<GroupKey> order;
    for (int i = 0; i < count; i++) {
        const GroupKey k = key_of(*engine_of(data[i]), kps[i]);
        if (!groups.count(k)) order.push_back(k);
        groups[k].push_back(i);
    }
    // the launch sets (one per group, or per part of a mixed group whose scratch
    // would pass kBatchScratch) run side by side, each on a stream pair of its own:
    // every set ends with its largest blocks, and the sets' tails overlap
    std::vector<std::vector<int>> jobs;
    for (const GroupKey& k : order) {
        std::vector<int>& g = groups[k];
        const bool mixed = engine_of(data[g[0]])->precision == HMG_PREC_MIXED;
        if (mixed) {
            // largest H-matrices first: a part lasts as long as its longest block,
            // so the clouds with the longest blocks share the first parts
            std::vector<int64_t> need_of(count, 0);
            for (const int i : g) {
                const Engine* e = engine_of(data[i]);
                for (int a = 0; a < e->naca; a++) need_of[i] += scratch_of(*e, a);
            }
            std::stable_sort(g.begin(), g.end(), [&](int x, int y) { return need_of[x] > need_of[y]; });
        }
        std::vector<int> part;
        int64_t used = 0;
        for (size_t q = 0; q <= g.size(); q++) {
            int64_t need = 0;
            if (q < g.size() && mixed) {
                const Engine* e = engine_of(data[g[q]]);
                for (int a = 0; a < e->naca; a++) need += scratch_of(*e, a);
            }
            if (q == g.size() || (!part.empty() && used + need > kBatchScratch)) {
                if (!part.empty()) jobs.push_back(part);
                part.clear();
                used = 0;
            }
            if (q < g.size()) {
                part.push_back(g[q]);
                used += need;
            }
        }
    }
    // at most kBatchSets sets at once (each mixed set holds up to kBatchScratch)
    HMG_TRY(parallel_for(
        (int)jobs.size(), std::min(threads > 0 ? threads : 8, kBatchSets),
        [&](int j) {
            StreamPair sp;
            HMG_TRY(sp.create());
            return run_group(data, jobs[j], kps[jobs[j][0]], sp.s, sp.s2, sp.fork, sp.join);
        },
        "batched ACA launch set"));
    for (int i = 0; i < count; i++) {
        Engine* e = engine_of(data[i]);
        if (e->precision == HMG_PREC_MIXED) {
            data[i]->U = (double*)e->Up.p;
            data[i]->V = (double*)e->Vp.p;
        } else {
            data[i]->U = e->U.p;
            data[i]->V = e->V.p;
        }
        e->fact_prec = e->precision;
        e->epoch++;
    }
    return HMG_OK;
}

// ---- the many-cloud product as one captured hipGraph -----------------------------
// The products of the listed H-matrices (precomputed factors) are captured from
// one stream into one graph: the capture stream forks to every H-matrix's own
// streams through an event, each H-matrix's product is enqueued there (its near
// field on its second stream as usual), and every H-matrix joins back.  A launch
// replays every kernel of every product with one call.  The graph holds the
// device pointers of that moment (operands, factors, schedules, product
// buffers); a launch checks that none of them has changed since (a new setup,
// precompute, precision or a wider product on one of the H-matrices) and refuses
// otherwise.
struct hmg_batch_graph {
    hipStream_t s = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<h_matrix_data*> data;
    std::vector<std::vector<const void*>> state;
    ~hmg_batch_graph()
    {
        if (s) (void)hipStreamSynchronize(s);
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }
};

extern "C" hmg_status hmg_apply_batch_graph_create(struct h_matrix_data* const* data, const double* const* X,
                                                   double* const* Y, const int* nrhs, int count,
                                                   hmg_batch_graph** out)
{
    using namespace hmg;
    if (!out || !data || !X || !Y || !nrhs || count < 1) {
        set_error("invalid batch graph arguments");
        return HMG_ERR_INVALID;
    }
    *out = nullptr;
    std::vector<Engine*> eng((size_t)count);
    std::vector<KernelParams> kp((size_t)count);
    for (int q = 0; q < count; q++) {
        HMG_TRY(check_ready(data[q], &eng[q]));
        for (int j = 0; j < q; j++)
            if (data[j] == data[q]) {
                set_error("batch graph: an H-matrix is listed twice");
                return HMG_ERR_INVALID;
            }
        if (!X[q] || !Y[q] || nrhs[q] < 1) {
            set_error("batch graph: invalid operands of H-matrix " + std::to_string(q));
            return HMG_ERR_INVALID;
        }
        if (!(data[q]->U != nullptr && eng[q]->aca_k >= 0)) {
            set_error("batch graph: H-matrix " + std::to_string(q) + " has no precomputed factors");
            return HMG_ERR_NOT_READY;
        }
        HMG_TRY(read_kernel(data[q], &kp[q]));
    }
    // one eager product each: the product buffers of these widths exist afterwards
    // (and the first-use clears are done), so the captured products allocate nothing
    for (int q = 0; q < count; q++) {
        AllocScope scope(eng[q]->stream);
        HMG_TRY(apply_enqueue(data[q], eng[q], kp[q], X[q], Y[q], nrhs[q]));
    }
    for (int q = 0; q < count; q++) HMG_HIP(hipStreamSynchronize(eng[q]->stream));
    auto* gr = new hmg_batch_graph();
    std::unique_ptr<hmg_batch_graph> own(gr);
    gr->data.assign(data, data + count);
    for (int q = 0; q < count; q++) gr->state.push_back(product_state(data[q], *eng[q]));
    HMG_HIP(hipStreamCreateWithFlags(&gr->s, hipStreamNonBlocking));
    gr->ev.resize((size_t)2 * count, nullptr);
    for (hipEvent_t& e : gr->ev) HMG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    std::vector<bool> prof((size_t)count);
    for (int q = 0; q < count; q++) {
        prof[q] = eng[q]->prof.on;
        eng[q]->prof.on = false;  // no timing events inside the graph
    }
    HMG_HIP(hipStreamBeginCapture(gr->s, hipStreamCaptureModeRelaxed));
    hmg_status st = HMG_OK;
    for (int q = 0; q < count && st == HMG_OK; q++) {
        Engine* e = eng[q];
        AllocScope scope(e->stream);
        if (hipEventRecord(gr->ev[2 * q], gr->s) != hipSuccess ||
            hipStreamWaitEvent(e->stream, gr->ev[2 * q], 0) != hipSuccess) {
            set_error("batch graph: fork into the capture failed");
            st = HMG_ERR_HIP;
            break;
        }
        st = apply_enqueue(data[q], e, kp[q], X[q], Y[q], nrhs[q]);
        if (st == HMG_OK && (hipEventRecord(gr->ev[2 * q + 1], e->stream) != hipSuccess ||
                             hipStreamWaitEvent(gr->s, gr->ev[2 * q + 1], 0) != hipSuccess)) {
            set_error("batch graph: join into the capture failed");
            st = HMG_ERR_HIP;
        }
    }
    const hipError_t ec = hipStreamEndCapture(gr->s, &gr->g);
    for (int q = 0; q < count; q++) eng[q]->prof.on = prof[q];
    HMG_TRY(st);
    HMG_HIP(ec);
    for (int q = 0; q < count; q++)
        if (product_state(data[q], *eng[q]) != gr->state[q]) {
            set_error("batch graph: a product allocated while being captured");
            return HMG_ERR_INVALID;
        }
    HMG_HIP(hipGraphInstantiate(&gr->x, gr->g, nullptr, nullptr, 0));
    HMG_HIP(hipGraphUpload(gr->x, gr->s));  // the first launch then costs what later ones do
    HMG_HIP(hipStreamSynchronize(gr->s));
    *out = own.release();
    return HMG_OK;
}

extern "C" hmg_status hmg_apply_batch_graph_launch(hmg_batch_graph* g)
{
    using namespace hmg;
    if (!g || !g->x) {
        set_error("invalid batch graph");
        return HMG_ERR_INVALID;
    }
    for (size_t q = 0; q < g->data.size(); q++) {
        Engine* e = engine_of(g->data[q]);
        if (!e || product_state(g->data[q], *e) != g->state[q]) {
            set_error("batch graph: H-matrix " + std::to_string(q) +
                      " changed since the capture (capture the graph again)");
            return HMG_ERR_NOT_READY;
        }
    }
    HMG_HIP(hipGraphLaunch(g->x, g->s));
    HMG_HIP(hipStreamSynchronize(g->s));
    return HMG_OK;
}

extern "C" void hmg_apply_batch_graph_destroy(hmg_batch_graph* g) { delete g; }
