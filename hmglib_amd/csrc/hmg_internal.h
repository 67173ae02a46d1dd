// hmg_internal.h -- shared declarations of the HIP engine (gfx950).
//
// Data layout in HBM (all device arrays, library-owned unless stated):
//   coords     per set, dim SoA arrays of N doubles (caller-allocated by init,
//              permuted in place into Morton order by setup)
//   codes      per set, N uint64 Morton codes (sorted), order: N uint64
//   blocks     struct work_item[nblocks] (192 B, DENSE first, then ACA, BFS order)
//   blk        int4 {l1,u1,l2,u2} per block, same order as `blocks`
//   U, V       ACA factors grouped by row cluster (U, column-major) and by column
//              cluster (V, row-major), leading dims padded to 16 doubles
//   A_b        per DENSE block (when precomputed), column-major m1 x m2 at aoff[b]
//   P          partial row / column sums of the MVP work items, combined per row
//              through a CSR list in a fixed order
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hmglib.h"
#include "../../include/kernel_system_assembler.h"

namespace hmg {

constexpr int kWave = 64;
constexpr int kTileRows = 64;     // rows per MVP row tile (one wavefront)

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string& msg);

#define HMG_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::hmg::set_error(std::string(#call) + ": " + hipGetErrorString(e_) +   \
                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
            return HMG_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define HMG_TRY(expr)                       \
    do {                                    \
        hmg_status s_ = (expr);             \
        if (s_ != HMG_OK) return s_;        \
    } while (0)

// ---- kernel descriptor as the device code sees it --------------------------
struct KernelParams {
    int kind;
    int dim;
    double reg;
    double c0;      // Gaussian: 1/(2 l^2); exponential: 1/l
    double c1;      // Matern reference: 1 / (2^{d/2} Gamma(1+d/2))
};

KernelParams make_kernel_params(const system_assembler& a, int dim);

// Coordinates of one set as the device code sees them (SoA, up to 5 dims)
struct Coords {
    const double* x[MAX_POINT_DIMENSION];
    int n;
};

// ---- device buffer helper --------------------------------------------------
// Stream-ordered allocation: inside an AllocScope, DevBuf memory comes from the
// device's memory pool ordered on the scope's stream (hipMallocAsync) and is
// returned on the stream it was allocated on (hipFreeAsync).  hipFree implies a
// device-wide synchronization, which would make every H-matrix served
// concurrently wait for the kernels of all others whenever one frees a buffer.
// Outside a scope DevBuf uses plain hipMalloc / hipFree.
extern thread_local hipStream_t g_alloc_stream;
struct AllocScope {
    hipStream_t prev;
    explicit AllocScope(hipStream_t s) : prev(g_alloc_stream) { g_alloc_stream = s; }
    ~AllocScope() { g_alloc_stream = prev; }
    AllocScope(const AllocScope&) = delete;
    AllocScope& operator=(const AllocScope&) = delete;
};

// Allocations of at least kLargeAlloc bytes bypass the stream-ordered pool: they
// come from hipMalloc and, when released, wait for the device (as hipFree would)
// and stay mapped in a process-wide cache that later large allocations of up to
// twice their size on the same device reuse.  On MI355X a fresh hipMalloc of device
// memory used before costs about 28 ms per GB (tens of GB: seconds); a reused
// buffer costs nothing.  An allocation that fails returns the cached buffers of
// its device and tries again; hmg_release_cached_memory returns them all.
constexpr size_t kLargeAlloc = (size_t)8 << 30;
hipError_t large_alloc(void** p, size_t bytes, size_t* held);
void large_free(void* p, size_t held);
// returns the cached large buffers of the current device (all devices: every = true)
void large_trim(bool every);
// after a failed allocation: returns the cached large buffers and, once the device
// is idle, the stream-ordered pool's free memory of the current device
void reclaim_device_memory();

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;  // allocation stream (nullptr: hipMalloc)
    size_t held = 0;          // bytes of a large buffer (from the large-buffer cache)
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), s(o.s), held(o.held)
    {
        o.p = nullptr;
        o.n = 0;
        o.s = nullptr;
        o.held = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            s = o.s;
            held = o.held;
            o.p = nullptr;
            o.n = 0;
            o.s = nullptr;
            o.held = 0;
        }
        return *this;
    }
    // a buffer still held at scope exit (error paths included) is returned on its
    // allocation stream
    ~DevBuf() { release(); }
    hmg_status alloc(size_t count) {
        if (count <= n && p) return HMG_OK;
        release();
        if (count == 0) count = 1;
        const size_t bytes = count * sizeof(T);
        hipError_t e;
        hipStream_t st = nullptr;
        if (bytes >= kLargeAlloc) {
            e = large_alloc((void**)&p, bytes, &held);
        } else {
            st = g_alloc_stream;
            e = st ? hipMallocAsync((void**)&p, bytes, st) : hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                // cached large buffers or the pool may hold the memory: return it and retry
                (void)hipGetLastError();
                reclaim_device_memory();
                e = st ? hipMallocAsync((void**)&p, bytes, st) : hipMalloc(&p, bytes);
            }
        }
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            held = 0;
            set_error(std::string("device allocation of ") + std::to_string(count * sizeof(T)) +
                      " bytes failed: " + hipGetErrorString(e));
            return HMG_ERR_OOM;
        }
        n = count;
        s = st;
        return HMG_OK;
    }
    void release() {
        if (p) {
            if (held)
                large_free(p, held);
            else if (s)
                (void)hipFreeAsync(p, s);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        s = nullptr;
        held = 0;
    }
};

// ---- per-kernel timing with HIP events on the engine's stream ---------------
enum ProfClass { PROF_GATHER = 0, PROF_LR_VTX = 1, PROF_NEAR = 2, PROF_ACA = 3, PROF_DENSE_FILL = 4,
                 PROF_LR_U = 5, PROF_COMBINE = 6, PROF_NCLASS = 7 };
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> ev[PROF_NCLASS];   // start/stop pairs
    void begin(int cls, hipStream_t s);
    void end(int cls, hipStream_t s);
    void clear();
};

// low-rank work item: rows [row0, row0+nrows) of the row cluster's U matrix
// (K columns, column-major with leading dim ld, U + uoff points at row0), times
// t at T + toff.  Mixed precision: U is the row cluster's BF16 panel (ld = padded
// rows, K = columns padded to 16, uoff = the panel element of row row0 in column
// group 0) and toff the first T panel group (8 slots each).
struct LrItem {
    int row0;
    int nrows;
    int ld;
    int K;
    int64_t uoff;
    int64_t toff;
};

// V^T x work item: columns [0, ncols) of a 64-column slice of a column cluster's
// transposed V (leading dim ld), rows (points) j0 .. j0+nrows-1, V + vbase at row j0.
// Mixed precision: a 128-column slice of the column cluster's BF16 panel (ld =
// columns padded to 32), j0 = first (absolute) point group of 8, nrows = point
// groups (a multiple of 4), vbase = the panel element of the slice's first column
// in group j0
// tcol0: the slice's first column in the tmap order; tp: partial-sum slot of this
// chunk, or -1 when the slice has a single chunk and the item writes T itself
// (mixed precision; the FP64 kernel always writes partials at its item index)
struct LvItem {
    int j0;
    int nrows;
    int ld;
    int ncols;
    int64_t vbase;
    int tcol0;
    int tp;
};
// One wavefront's share of a low-rank workgroup item (mixed precision): W
// wavefronts of a workgroup multiply their own A panel columns / rows by the
// same B panel groups (the same point groups of X for V^T X, the same T slot
// groups for U T), which the workgroup stages once through LDS.
//   aoff: A panel element of the wavefront's first M row in B group b0 (hi half)
//   ld:   A panel leading dimension (Kp of a V panel, m1p of a U panel)
//   m:    valid M (columns of V^T X, rows of U T); 0 for an idle wavefront
//   b0, nb: first B panel group and B groups (a multiple of 4; the same for all
//           wavefronts of the workgroup)
//   o0:   V^T X: first tmap column of the slice; U T: its partial-row slot item
//   tp:   V^T X: partial slot (-1: the result goes to the T panels directly)
struct WgItem {
    int64_t aoff;
    int ld;
    int m;
    int b0;
    int nb;
    int o0;
    int tp;
};
// wavefronts per workgroup of the V^T X and U T workgroup kernels and of the
// near field's workgroup form
constexpr int kVwWaves = 4;
constexpr int kUwWaves = 2;
constexpr int kNwWaves = 2;

// reduction of the row chunks of one column slice: partial slots
// item0 .. item0+nchunks-1, slice columns map to T through tmap[tcol0 + l]
struct LvSlice {
    int item0;
    int nchunks;
    int ncols;
    int tcol0;
};

// Matrix-free near field.  A work item covers up to 128 rows [row0, row0+nrows)
// of one row cluster against the column ranges of its refs.  Refs [ref0, ref_sym)
// contribute row sums only; refs [ref_sym, ref1) are the upper halves of
// transpose pairs (same point set on both sides, disjoint clusters): each entry
// is evaluated once and feeds the row sum of its row and the column sum of its
// column, written to P[cbase + column offset].  ncols_sym = total columns of
// the symmetric refs (they are streamed as one column sequence).
struct NearRef {
    int l2;
    int m2;
    int64_t cbase;
};
struct NearItem {
    int row0;
    int nrows;
    int ref0;
    int ref_sym;
    int ref1;
    int ncols_sym;
};

// symmetric column slots of one column cluster: contribution k of point l2 + j
// sits at P[base + k*m2 + j], k < count
struct ColSlots {
    int l2, m2;
    int64_t base;
    int count;
};

// Mixed-precision near field (many right-hand sides).  Points are handled in
// aligned groups of 8 (the 8 columns one lane feeds to a BF16 MFMA).  A ref is
// a run of groups [g0, g1) (g1 - g0 even: two groups per 16-column step) whose
// columns outside [lo, hi) are masked; an item is up to 128 rows of one row
// cluster against the refs [ref0, ref1).
struct MxRef {
    int g0, g1;
    int lo, hi;
};
struct MxItem {
    int row0;
    int nrows;
    int ref0;
    int ref1;
};

// Number of right-hand sides one mixed-precision pass handles (the N of the
// MFMA tiles); more right-hand sides run in several passes.
constexpr int kMxRhs = 32;
// Columns of one mixed-precision V^T X work item and rows of one mixed-precision
// U T work item (up to four interleaved 32-wide MFMA tiles).
constexpr int kMxVCols = 128;
constexpr int kMxLrRows = 128;
// Rows of one mixed-precision near-field work item (four 32-row MFMA tiles); the
// schedule spaces the items' partial-sum slots by it and the kernel writes them
// at that spacing.
constexpr int kMxNearRows = 128;
// Mixed-precision factor panels (BF16 hi + lo, zero padded): for a row cluster C
// with m1p = rows padded to 32 and its rank columns padded to 32, column group kg
// (8 columns) is [hi: m1p x 8][lo: m1p x 8], U_C[i][8kg+e] at
// 16 m1p kg + 8 i + e (+ 8 m1p for lo).  For a column cluster D with Kp = its rank
// columns padded to 32, absolute point group g (points 8g .. 8g+7; groups from the
// cluster's first to its last, padded to a multiple of 4) is [hi: Kp x 8][lo: Kp x 8],
// V_D[8g+e][l] at 16 Kp (g - g_first) + 8 l + e.  T panels: slot group s/8 is
// [hi: 32 rhs x 8][lo: 32 rhs x 8].  Elements of slack past each panel array let
// the widest tiles read past the last cluster.
constexpr int kMxPanelSlack = 4096;
// placement of one ACA block in the panels
struct MxPanel {
    int64_t ubase;  // U panel of the block's row cluster
    int64_t vbase;  // V panel of the block's column cluster
    int m1p;        // padded rows of that U panel
    int koff;       // the block's first column in the row cluster
    int kp;         // padded columns of that V panel
    int pos;        // the block's first column in the column cluster
    int glo;        // first point group of the column cluster
    int pad;
};

// per-point partial lists of one MVP schedule (CSR over sorted rows): an entry
// v >= 0 is the P slot v, an entry v < 0 the column-slot group -(v+1)
struct Combine {
    DevBuf<int> ptr, idx;
    DevBuf<ColSlots> groups;
    int64_t total = 0;          // list entries
    int64_t group_entries = 0;  // entries naming a column-slot group
    int64_t slots = 0;          // column slots reached through the groups
};

// ---- engine state hung off h_matrix_data::engine ---------------------------
struct Engine {
    Profiler prof;
    hipStream_t stream = nullptr;
    int n[2] = {0, 0};
    int dim = 0;
    int bits = 0;
    bool built = false;
    uint64_t epoch = 0;  // bumped by every change of set-up, factors or options

    DevBuf<uint32_t> perm32[2];                 // sort values
    DevBuf<work_item> blocks;                   // final block list (DENSE first)
    DevBuf<int4> blk;                           // l1,u1,l2,u2 per block
    std::vector<int4> h_blk;
    int nblocks = 0, ndense = 0, naca = 0;
    int levels = 0;
    int quirk_levels = 0;

    // ACA factors (FP64).  U of all ACA blocks that share a row cluster C is one
    // column-major m1(C) x K(C) matrix with leading dim ldu = m1 rounded up to 16
    // (blocks in block order, k_b columns each; U_b[i,l] at uoff[a] + l*ldu[a] + i);
    // V of all ACA blocks that share a column cluster D is one row-major
    // m2(D) x K(D) matrix with leading dim ldv = K rounded up to 16 (V_b[j,l] at
    // voff[a] + j*ldv[a] + l).  Row starts on 128-byte lines keep every 64-lane
    // segment of the MVP streams on 4 lines.  T (V^T x) is laid out in U order;
    // tmap sends V columns to T.
    int aca_k = -1;                             // k used for the stored factors
    std::vector<int> h_kb, h_ldv, h_ldu;
    std::vector<int64_t> h_uoff, h_voff, h_toff; // per ACA block (+ total at [naca])
    DevBuf<int> kb, keff, ldv, ldu;
    DevBuf<int> aca_list;                       // ACA blocks: large (workgroup each) first
    std::vector<int> h_aca_list;
    DevBuf<int64_t> uoff, voff, toff;
    DevBuf<double> U, V;
    int kmax = 0;
    int64_t tsize = 0;

    // dense near field (FP64), per DENSE block contiguous, column-major m1 x m2
    std::vector<int64_t> h_aoff;                // ndense+1
    DevBuf<int64_t> aoff;
    DevBuf<double> A;
    bool dense_stored = false;

    // row-cluster groups: DENSE and ACA blocks grouped by their row cluster
    // (groups in order of first appearance, blocks in block order inside a group)
    std::vector<int> h_dgrp_ptr, h_dgrp_blk;    // dense groups
    std::vector<int> h_agrp_ptr, h_agrp_blk;    // ACA groups (indices into the ACA list)
    std::vector<int> h_cgrp_ptr, h_cgrp_blk;    // ACA blocks grouped by column cluster

    // MVP schedule: work items write partial row (and column) sums into P; every
    // row sums its partials in a fixed order (CSR) and writes y once.  P layout:
    // [low-rank items x 64][near items x 128][symmetric column slots].
    int n_near_items = 0, n_lr_items = 0, n_lv_items = 0, n_lv_slices = 0;
    int n_tp = 0;                               // partial-sum slots of the V^T x items
    // mixed-precision layout: V^T x and U t items grouped by the number of
    // interleaved 32-wide MFMA tiles they use (offsets of the 4-, 2- and 1-tile
    // groups and the end)
    int lv_nt_off[4] = {0, 0, 0, 0};
    int lr_nt_off[4] = {0, 0, 0, 0};
    // the same items as workgroups that share their B operand (kVwWaves / kUwWaves
    // consecutive entries per workgroup); lr_wg selects them
    std::vector<WgItem> h_vw_items, h_uw_items;
    DevBuf<WgItem> vw_items, uw_items;
    bool lr_wg = false;
    // cluster boxes of the tree: false = the reference's lookup table (a one-point
    // cluster shifts the level's boxes), true = each cluster's exact min/max
    bool exact_boxes = false;
    DevBuf<int4> near_items;                    // stored near field: {row0, nrows, b0, b1}
    DevBuf<int> near_blk;
    std::vector<int4> h_near_items;
    // matrix-free near field (symmetric when both sets are the same points)
    bool same_sets = false;                     // sorted coordinates bitwise equal
    bool near_sym_enabled = true;               // option: allow the symmetric form
    bool near_sym = false;                      // the fly schedule uses transpose pairs
    std::vector<ColSlots> h_col_groups;         // column slot groups (bases relative)
    int n_fly_items = 0;
    int64_t n_col_slots = 0;
    int64_t near_evals = 0;                     // kernel entries evaluated per matrix-free pass
    std::vector<NearItem> h_fly_items;
    std::vector<NearRef> h_fly_refs;
    DevBuf<NearItem> fly_items;
    DevBuf<NearRef> fly_refs;
    DevBuf<LrItem> lr_items;
    DevBuf<LvItem> lv_items;
    DevBuf<LvSlice> lv_slices;
    DevBuf<int> tmap;
    DevBuf<double> Tp;                          // V^T x partials per row chunk
    Combine comb_stored, comb_fly;              // stored near field / matrix-free
    DevBuf<double> P, xs, T;
    int64_t p_slots = 0;                        // FP64 partial-sum slots of P
    // FP64 products of blocks of right-hand sides (mvp_multi.hip): the pass's sorted
    // right-hand-side rows, V^T X chunk partials, T, row partials, sorted results;
    // products with at least fp64_block_min right-hand sides take this path
    DevBuf<double> X64, Tp64, T64, P64, Ys64;
    int fp64_block_min = 2;

    hipStream_t stream2 = nullptr;              // near field runs beside the low-rank part
    // tuning: LDS bytes reserved per near-field workgroup (caps its occupancy so
    // the low-rank streams keep wave slots on every CU); run the near field on
    // the main stream instead (no overlap)
    // 64 KB + the kernel's own 12 KB leave room for 2 near-field workgroups per CU
    // (measured on MI355X at N = 2^20, 3D, symmetric near field: 42.8 ms per
    // product vs 46.2 ms with 3 workgroups per CU and 57.9 ms without overlap)
    int near_lds_pad = 64 * 1024;
    bool serialize = false;
    // ACA: scan rejected rows of the residual-row search several at a time;
    // blocks with more rows plus columns than aca_big_entries get a workgroup
    // false: stream == stream2 belongs to the caller (the serving loop's slot
    // streams), which rebinds it per job and keeps its scratch buffers
    bool own_streams = true;
    // kernel object created by hmg_load (freed with the engine)
    system_assembler* own_kernel = nullptr;
    bool aca_row_scan = true;
    // ACA tolerance stop: mode (option) and its constant for the current build
    int aca_tol_mode = 0;
    double aca_tol_c = 0.0;
    int aca_big_entries = 512;
    int aca_huge_entries = 16384;               // above: an 8-wavefront workgroup, launched beside
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    // ---- mixed precision ------------------------------------------------------
    // precision: requested mode for the next precompute_aca / apply (HMG_PREC_*);
    // layout_prec: mode the factor layout and low-rank items were built for;
    // fact_prec: mode of the stored factors (valid when aca_k >= 0).
    int precision = HMG_PREC_FP64;
    int layout_prec = HMG_PREC_FP64;
    int fact_prec = HMG_PREC_FP64;
    DevBuf<__bf16> Up, Vp, Tpan;                // factor and T panels (BF16 hi + lo)
    bool tpan_ready = false;                    // Tpan allocated and zeroed for the layout
    int64_t up_size = 0, vp_size = 0;           // panel elements (with slack)
    std::vector<MxPanel> h_mxp;
    DevBuf<MxPanel> mxp;
    // near field: 64-row items against merged, group-aligned column runs
    std::vector<MxItem> h_mx_items;
    std::vector<MxRef> h_mx_refs;
    DevBuf<MxItem> mx_items;
    DevBuf<MxRef> mx_refs;
    int n_mx_items = 0;
    // workgroup form of the near field: kNwWaves item indices per workgroup (the row
    // slices of one column chunk, -1 = idle); near_wg selects it
    std::vector<int> h_mx_wg;
    DevBuf<int> mx_wg;
    bool near_wg = false;
    // mixed precision: low-rank wavefronts issue at a raised priority (s_setprio):
    // bit 0 V^T X, bit 1 U T
    int lr_prio = 0;
    // mixed-precision ACA: its FP64 scratch holds 1/aca_mixed_parts of the factors
    int aca_mixed_parts = 16;
    // unused LDS reserved by every workgroup of the main ACA launch (caps the blocks
    // in flight per compute unit)
    int aca_lds_pad = 0;
    int64_t mx_evals = 0;                       // kernel entries evaluated per pass (incl. masked)
    bool mx_sched = false;                      // near items built for the current tree
    Combine comb_mx;
    // per pass: X as FP32 rows of kMxRhs values, as BF16 hi/lo panels (per group of
    // 8 points: [hi|lo][rhs][8]), and coordinate panels per group ([dim][8] FP32)
    DevBuf<float> Tp32, P32, Ys32;              // Ys32: sorted result rows of a pass
    DevBuf<int> inv1;                           // sorted position of each original row
    bool inv1_ready = false;
    DevBuf<__bf16> Xp;
    DevBuf<float> Xt32;                         // a pass's right-hand sides as 32-wide FP32 rows
    DevBuf<float> cpan[2];
    bool cpan_ready = false;
    float cpan_scale = 0.f;                     // the scale the coordinate panels were built with
};

Engine* engine_of(const h_matrix_data* d);

// ---- primitives (prims.hip) ------------------------------------------------
// The primitives keep grow-only scratch buffers per (stream, slot); an engine
// releases the scratch of its streams when it is destroyed.
void release_scratch(hipStream_t s);
// exclusive or inclusive scan of int32, n elements, in -> out (may alias);
// optional device pointer `total` receives the sum.
hmg_status scan_i32(const int* in, int* out, int64_t n, bool inclusive, int* total,
                    hipStream_t s);
// stable LSD radix sort of (uint64 key, uint32 value) pairs on the low `key_bits`
// bits; keys/vals sorted in place.
hmg_status radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits,
                          hipStream_t s);
// exact per-dimension min/max of a double array
hmg_status minmax_f64(const double* x, int64_t n, double* dmin, double* dmax, hipStream_t s);
// *differ = 1 if a[i] and b[i] differ bitwise for some i < n (left untouched otherwise)
hmg_status flag_differ_f64(const double* a, const double* b, int64_t n, int* differ, hipStream_t s);

// ---- Morton (morton.hip) ------------------------------------------------------
hmg_status morton_codes(const Coords& c, int dim, int bits, const double* dmin,
                        const double* dmax, uint64_t* codes, hipStream_t s);
// several clouds side by side in c: point p quantised on the box of cloud cid[p]
// (cmin / cmax: dim values per cloud)
hmg_status morton_codes_batched(const Coords& c, int dim, int bits, const int* cid, const double* cmin,
                                const double* cmax, uint64_t* codes, hipStream_t s);
hmg_status iota_u32(uint32_t* v, int64_t n, hipStream_t s);
hmg_status gather_f64(const double* in, const uint32_t* idx, double* out, int64_t n,
                      hipStream_t s);
hmg_status perm_to_u64(const uint32_t* p, uint64_t* order, unsigned* ids, int64_t n,
                       hipStream_t s);

// ---- tree (tree.hip) ---------------------------------------------------------
struct TreeParams {
    double eta;
    int max_level;
    int c_leaf;
    int root_level_1, root_level_2;
    bool exact_boxes;  // cluster boxes from the points themselves (no lookup table)
};
// Builds the DENSE-first block list into eng.blocks / eng.blk / eng.h_blk.
hmg_status build_tree(Engine& eng, const Coords& c1, const Coords& c2, const uint64_t* codes1,
                      const uint64_t* codes2, const TreeParams& tp);
// The trees of several point clouds, each against itself, in one traversal: cloud q's
// sorted points at [off[q], off[q] + n[q]) of c and codes (cid: the cloud of each
// point), exact boxes, the same parameters for all; each engine receives its own
// DENSE-first block list (its own indices) in blocks / blk / h_blk and its counts.
hmg_status build_trees_batched(hipStream_t s, int dim, const Coords& c, const uint64_t* codes,
                               const TreeParams& tp, const std::vector<int>& off, const std::vector<int>& n,
                               const int* cid, Engine* const* eng);
// What setup_h_matrix does after the tree (api.hip): block list pointers and counts,
// batch sizing, row/column groups, the rank layout and the MVP schedule.
hmg_status finish_setup(h_matrix_data* data, Engine* e);
// returns what a previous setup built (the state a new setup starts from; api.hip)
void release_build_state(Engine& e, h_matrix_data* d);
// The product of precomputed factors enqueued on the engine's streams: no host
// synchronisation, and no allocation once the engine's product buffers exist for
// this width (what a hipGraph capture needs; api.hip).
hmg_status apply_enqueue(h_matrix_data* data, Engine* e, const KernelParams& kp, const double* X, double* Y,
                         int nrhs);
// the device pointers an enqueued product reads or writes (schedule, factors,
// product buffers): a captured product is valid while they stay the same
std::vector<const void*> product_state(const h_matrix_data* d, const Engine& e);

// Host <-> device copies of the set-up path (api.hip).  The bytes pass through a
// pinned buffer of the calling thread (at most 4 MiB, allocated once per thread),
// so the runtime never stages pageable memory; each call waits for its copies on
// every path, the error path included, so the host side may go out of scope or
// be freed at once.
hmg_status upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);
hmg_status download_bytes(void* dst, const void* src, size_t bytes, hipStream_t s);

// host vector -> device buffer (grown as needed), on stream s
template <class T>
hmg_status upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s)
{
    HMG_TRY(d.alloc(std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HMG_TRY(upload_bytes(d.p, h.data(), sizeof(T) * h.size(), s));
    return HMG_OK;
}

// ---- engine set-up (api.hip) ---------------------------------------------------
// init_h_matrix_data; ext != nullptr runs all of the H-matrix's work on that
// caller-owned stream (no near-field / low-rank overlap inside one product)
hmg_status engine_init(h_matrix_data* data, int point_count[2], int dim, int bits, hipStream_t ext);
// moves a caller-owned-stream engine to another stream of the caller (between
// jobs, with no work of this H-matrix in flight)
void engine_bind_stream(Engine* e, hipStream_t s, h_matrix_data* data);

// ---- helpers of the C ABI (api.hip) ---------------------------------------------------
// set s of an H-matrix as the kernels see it (its sorted coordinates)
Coords coords_of(const h_matrix_data* d, int set);
// the kernel object behind d->assem, validated
hmg_status read_kernel(const h_matrix_data* d, KernelParams* kp);
// the engine of an H-matrix that has been set up
hmg_status check_ready(const h_matrix_data* d, Engine** out);
// drops stored factors (FP64 and mixed precision)
void release_factors(Engine& e, h_matrix_data* d);
// the tolerance stop's constant of the next ACA from d->epsilon and d->eta
void set_tolerance(Engine& e, const h_matrix_data* d);

// ---- ACA (aca.hip) -------------------------------------------------------------
// tie_last: 1 = last argmax (batched path), 0 = first argmax (unbatched path).
// FP64 factors straight into the FP64 layout.
// ublk / vblk (optional, device arrays of naca pointers): block a's factors start
// at ublk[a] / vblk[a] instead of U + uoff[a] / V + voff[a] (factors that live in
// other H-matrices' buffers).
hmg_status aca_build(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                     int k, int tie_last, double* U, double* V, int* keff, double* const* ublk = nullptr,
                     double* const* vblk = nullptr);
// Mixed precision: the same FP64 ACA, batch by batch through a bounded FP64
// scratch, each batch's factors split into the BF16 hi + lo panels Up / Vp
// (zeroed by the caller).
hmg_status aca_build_mixed(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                           int tie_last, __bf16* Up, __bf16* Vp);
// FP64 scratch elements aca_build_mixed uses for the current layout
int64_t aca_mixed_scratch(const Engine& eng);
// BF16 hi + lo panels (Up / Vp) of ACA blocks list[0 .. n) of eng from FP64
// factors in a scratch: U_b column-major (ld m1) at S + suoff[a], V_b row-major (ld
// kb) at S + svoff[a]; ranks at or past keff[a] become zero columns.  On stream s.
hmg_status aca_to_panels(Engine& eng, const int* list, int n, const int64_t* suoff, const int64_t* svoff,
                         const double* S, __bf16* Up, __bf16* Vp, hipStream_t s);

// ---- schedule (schedule.hip, host) ---------------------------------------------------
// row-cluster groups and near-field work items (once per tree); the matrix-free
// schedule pairs transposed DENSE blocks when eng.same_sets
hmg_status build_groups(Engine& eng);
// ranks k_b = min(k,m1,m2), factor offsets, low-rank work items, per-row partial
// lists, for factors of precision prec (HMG_PREC_*: leading dims are padded to
// 128 bytes; mixed precision also builds its near-field items)
hmg_status aca_layout(Engine& eng, int k, int prec = HMG_PREC_FP64);

// ---- mixed precision (mixed.hip) -------------------------------------------------
// Y = H X for nrhs right-hand sides (X, Y FP64 column-major in the caller's point
// order) with the BF16 hi + lo factor panels: low-rank products and the near field
// (evaluated in FP32) as BF16 hi+lo split MFMA GEMMs accumulated in FP32.
hmg_status hmvp_mixed(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order1, const uint64_t* order2, const double* X, double* Y,
                      int nrhs);

// ---- FP64 products of blocks of right-hand sides (mvp_multi.hip) -------------------
// Y = H X for nrhs right-hand sides with the FP64 factors: up to 32 right-hand
// sides per pass, each factor element read once per pass, the near field
// evaluated on the fly once per entry and pass.  A column's result does not depend
// on the other columns of its pass.
constexpr int kRbMax = 32;
hmg_status hmvp_multi(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order1, const uint64_t* order2, const double* X, double* Y, int nrhs,
                      const double* U, const double* V);

// ---- near field / MVP (mvp.hip) ----------------------------------------------------
hmg_status dense_fill(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp);
hmg_status hmvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                const uint64_t* order1, const uint64_t* order2, const double* x, double* y,
                int nrhs, const double* U, const double* V, const int* keff, bool dense_stored);
// The per-call buffers of the FP64 product: x in sorted order, V^T x partials per
// point chunk, t in U order, and the row / column partial sums every row adds up.
// Every element the product reads is written earlier in the same call.
struct MvpBufs {
    double* xs;
    double* Tp;
    double* T;
    double* P;
};
// their element counts for the engine's current schedule
struct MvpSizes {
    int64_t xs, Tp, T, P;
};
MvpSizes mvp_sizes(const Engine& eng);
// hmvp with caller-provided buffers (allocates nothing)
hmg_status hmvp_with(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                     const uint64_t* order1, const uint64_t* order2, const double* x, double* y, int nrhs,
                     const double* U, const double* V, const int* keff, bool dense_stored, const MvpBufs& b);
hmg_status full_mvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                    const uint64_t* order1, const uint64_t* order2, const double* x, double* y,
                    int nrhs);
hmg_status dense_rows(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order2, const double* x, int nrhs, const int64_t* rows,
                      int nrows, double* yrows, const int64_t* inv1);
hmg_status gaussian_rhs(Engine& eng, const Coords& c1, const uint64_t* order1, double* b);
hmg_status inverse_perm(const uint64_t* order, int64_t* inv, int64_t n, hipStream_t s);

}  // namespace hmg
