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

}  // namespace
}  // namespace hmg

extern "C" hmg_status hmg_setup_batch(struct h_matrix_data* const* data, int count, int threads)
{
    if (!data || count < 0) {
        set_error("invalid batch");
        return HMG_ERR_INVALID;
    }
    return parallel_for(count, threads, [&](int i) { return hmg_setup_h_matrix(data[i]); }, "batched setup");
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
        const std::vector<int>& g = groups[k];
        const bool mixed = engine_of(data[g[0]])->precision == HMG_PREC_MIXED;
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
    }
    return HMG_OK;
}
