// hmg_internal.h -- shared declarations of the HIP engine (gfx950).
//
// Data layout in HBM (all device arrays, library-owned unless stated):
//   coords     per set, dim SoA arrays of N doubles (caller-allocated by init,
//              permuted in place into Morton order by setup)
//   codes      per set, N uint64 Morton codes (sorted), order: N uint64
//   blocks     struct work_item[nblocks] (192 B, DENSE first, then ACA, BFS order)
//   blk        int4 {l1,u1,l2,u2} per block, same order as `blocks`
//   U_b, V_b   per ACA block, column-major m1 x k_b and m2 x k_b, contiguous,
//              at uoff[a] / voff[a] (int64 element offsets)
//   A_b        per DENSE block (when precomputed), column-major m1 x m2 at aoff[b]
//   tiles      row tiles of 64 Morton-ordered rows; CSR list per tile of the blocks
//              whose row range intersects it (block order)
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
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hmg_status alloc(size_t count) {
        if (count <= n && p) return HMG_OK;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                      " bytes failed: " + hipGetErrorString(e));
            return HMG_ERR_OOM;
        }
        n = count;
        return HMG_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ---- per-kernel timing with HIP events on the engine's stream ---------------
enum ProfClass { PROF_GATHER = 0, PROF_LR_VTX = 1, PROF_ROW = 2, PROF_ACA = 3, PROF_DENSE_FILL = 4,
                 PROF_NCLASS = 5 };
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> ev[PROF_NCLASS];   // start/stop pairs
    void begin(int cls, hipStream_t s);
    void end(int cls, hipStream_t s);
    void clear();
};

// ---- engine state hung off h_matrix_data::engine ---------------------------
struct Engine {
    Profiler prof;
    hipStream_t stream = nullptr;
    int n[2] = {0, 0};
    int dim = 0;
    int bits = 0;
    bool built = false;

    DevBuf<uint32_t> perm32[2];                 // sort values
    DevBuf<work_item> blocks;                   // final block list (DENSE first)
    DevBuf<int4> blk;                           // l1,u1,l2,u2 per block
    std::vector<int4> h_blk;
    int nblocks = 0, ndense = 0, naca = 0;
    int levels = 0;
    int quirk_levels = 0;

    // ACA factors (FP64), per ACA block contiguous
    int aca_k = -1;                             // k used for the stored factors
    std::vector<int> h_kb;
    std::vector<int64_t> h_uoff, h_voff;        // naca+1
    DevBuf<int> kb, keff;
    DevBuf<int64_t> uoff, voff;
    DevBuf<double> U, V;
    int kmax = 0;

    // dense near field (FP64), per DENSE block contiguous
    std::vector<int64_t> h_aoff;                // ndense+1
    DevBuf<int64_t> aoff;
    DevBuf<double> A;
    bool dense_stored = false;

    // MVP schedule
    int ntiles = 0;
    DevBuf<int> tile_ptr, tile_blk;
    DevBuf<int64_t> toff;                       // per ACA block offset into T (naca+1)
    std::vector<int64_t> h_toff;
    DevBuf<double> xs, ys, T;                   // permuted x, permuted y, V^T x
    std::vector<int> h_lr_order;                // ACA blocks, largest m2 first
    DevBuf<int> lr_order;

    // host-side batch sizing copies (reference semantics) for dense/aca_work_size
    std::vector<int> dense_ws, aca_ws;
};

Engine* engine_of(const h_matrix_data* d);

// ---- primitives (prims.hip) ------------------------------------------------
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

// ---- Morton (morton.hip) ------------------------------------------------------
hmg_status morton_codes(const Coords& c, int dim, int bits, const double* dmin,
                        const double* dmax, uint64_t* codes, hipStream_t s);
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
};
// Builds the DENSE-first block list into eng.blocks / eng.blk / eng.h_blk.
hmg_status build_tree(Engine& eng, const Coords& c1, const Coords& c2, const uint64_t* codes1,
                      const uint64_t* codes2, const TreeParams& tp);

// ---- ACA (aca.hip) -------------------------------------------------------------
// tie_last: 1 = last argmax (batched path), 0 = first argmax (unbatched path)
hmg_status aca_build(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                     int k, int tie_last, double* U, double* V, int* keff);
hmg_status aca_layout(Engine& eng, int k);

// ---- near field / MVP (mvp.hip) ----------------------------------------------------
hmg_status dense_fill(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp);
hmg_status build_schedule(Engine& eng);
hmg_status hmvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                const uint64_t* order1, const uint64_t* order2, const double* x, double* y,
                int nrhs, const double* U, const double* V, const int* keff, bool dense_stored);
hmg_status full_mvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                    const uint64_t* order1, const uint64_t* order2, const double* x, double* y,
                    int nrhs);
hmg_status dense_rows(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order2, const double* x, int nrhs, const int64_t* rows,
                      int nrows, double* yrows, const int64_t* inv1);
hmg_status gaussian_rhs(Engine& eng, const Coords& c1, const uint64_t* order1, double* b);
hmg_status inverse_perm(const uint64_t* order, int64_t* inv, int64_t n, hipStream_t s);

}  // namespace hmg
