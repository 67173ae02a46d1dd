// mvp_multi.hip -- FP64 H-matrix-vector products for blocks of right-hand sides.
//
// The FP64 product of mvp.hip streams every factor element once per right-hand
// side.  Here up to kRbMax right-hand sides share one pass (reference:
// linear_algebra.cu:4292-4379, 1429-1591, 2817-3137; the same decomposition):
//   gather     X (caller order, column-major) -> Xs[p * RB + c], sorted points,
//              one RB-wide row per point (64 points per workgroup through LDS)
//   lr_v       Tp = V_D^T Xs per (column cluster, 64 columns, <= 1024 points)
//   lr_u       P = U_C T_C per 64 rows of a row cluster
//   near       the near field evaluated on the fly, each entry once for all the
//              pass's right-hand sides
//              all three as FP64 MFMA GEMMs (v_mfma_f64_16x16x4_f64, right-hand
//              sides as the N dimension, 16 or 32 per pass)
//   combine    every row adds its partials in the fixed order of the row's list,
//              then the sorted rows go back to the caller's order through LDS
// An MFMA's output column depends on its own B column only, and the reduction
// steps run in a fixed order, so a right-hand side's result is bitwise the same
// in every pass width and whatever other columns share the pass.
#include "hmg_internal.h"
#include "kernel_eval.h"

namespace hmg {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

constexpr int kRowTile = 64;

// Xs[p*RB + c] = X[order2[p] + (c0 + c) * n2] (0 for c >= nr)
template <int RB>
__global__ void __launch_bounds__(kThreads)
gather_rows_kernel(const double* __restrict__ X, const uint64_t* __restrict__ order2, int n2, int c0, int nr,
                   double* __restrict__ Xs)
{
    __shared__ double tile[RB][kRowTile + 1];
    const int p0 = blockIdx.x * kRowTile;
    for (int e = threadIdx.x; e < kRowTile * RB; e += kThreads) {
        const int c = e / kRowTile, pl = e % kRowTile;
        const int p = p0 + pl;
        tile[c][pl] = (p < n2 && c < nr) ? X[(int64_t)order2[p] + (int64_t)(c0 + c) * n2] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kRowTile * RB; e += kThreads) {
        const int pl = e / RB, c = e % RB;
        const int p = p0 + pl;
        if (p < n2) Xs[(int64_t)p * RB + c] = tile[c][pl];
    }
}

// FP64 MFMA v_mfma_f64_16x16x4_f64 (wave64): for lane l, i = l % 16, g = l / 16:
//   A[i][g] (16 x 4), B[g][i] (4 x 16), D[4r + g][i] in result register r < 4 (16 x 16)
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Tp[(q*64 + col)*RB + c] = sum_{j < nrows} V[vbase + col + j*ld] * Xs[(j0 + j)*RB + c]:
// M = the item's 64 columns (four 16-column tiles), N = right-hand sides (RB / 16
// tiles), K = points, four per step.  Points past nrows enter as zeros.  Four
// wavefronts per SIMD (the accumulators in ordinary registers): the step's loads
// wait on memory, so occupancy pays (2 per SIMD with AGPR accumulators ran 8%
// slower alone and twice as slow beside the near field).
template <int RB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
lr_v_multi_kernel(const LvItem* __restrict__ items, int nitems, const double* __restrict__ V,
                  const double* __restrict__ Xs, double* __restrict__ Tp)
{
    constexpr int NB = RB / 16;
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const LvItem it = items[q];
    const int i16 = lane & 15, g = lane >> 4;
    const int64_t ld = it.ld;
    const double* __restrict__ v = V + it.vbase + i16;
    const double* __restrict__ xb = Xs + (int64_t)it.j0 * RB + i16;
    d4 acc[4][NB];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int h = 0; h < NB; h++) acc[t][h] = d4{0.0, 0.0, 0.0, 0.0};
    // operands of step j0 (point j0 + g; points past nrows load the item's last
    // point and enter as zeros)
    auto load = [&](int j0, double (&a)[4], double (&b)[NB]) {
        const int pt = j0 + g;
        const bool ok = pt < it.nrows;
        const int pc = ok ? pt : it.nrows - 1;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const double x = v[pc * ld + 16 * t];
            a[t] = ok ? x : 0.0;
        }
#pragma unroll
        for (int h = 0; h < NB; h++) {
            const double x = xb[(int64_t)pc * RB + 16 * h];
            b[h] = ok ? x : 0.0;
        }
    };
    double a[4], b[NB];
    load(0, a, b);
    for (int j0 = 0; j0 < it.nrows; j0 += 4) {
        // the next step's operands load while this step's MFMAs run (past the last
        // step the loads repeat it)
        double an[4], bn[NB];
        load(j0 + 4 < it.nrows ? j0 + 4 : j0, an, bn);
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int h = 0; h < NB; h++) acc[t][h] = mfma64(a[t], b[h], acc[t][h]);
#pragma unroll
        for (int t = 0; t < 4; t++) a[t] = an[t];
#pragma unroll
        for (int h = 0; h < NB; h++) b[h] = bn[h];
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int h = 0; h < NB; h++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = 16 * t + 4 * r + g;
                Tp[((int64_t)q * kTileRows + col) * RB + 16 * h + i16] = acc[t][h][r];
            }
}

// T[tmap[tcol0 + lane]*RB + c] = sum of the slice's point chunks, in chunk order
template <int RB>
__global__ void __launch_bounds__(kThreads)
lr_v_finalize_multi_kernel(const LvSlice* __restrict__ slices, int nslices, const double* __restrict__ Tp,
                           const int* __restrict__ tmap, double* __restrict__ T)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nslices) return;
    const LvSlice sl = slices[q];
    if (lane >= sl.ncols) return;
    double s[RB];
#pragma unroll
    for (int c = 0; c < RB; c++) s[c] = 0.0;
    for (int ch = 0; ch < sl.nchunks; ch++) {
        const double* __restrict__ p = Tp + ((int64_t)(sl.item0 + ch) * kTileRows + lane) * RB;
#pragma unroll
        for (int c = 0; c < RB; c++) s[c] += p[c];
    }
    double* __restrict__ t = T + (int64_t)tmap[sl.tcol0 + lane] * RB;
#pragma unroll
    for (int c = 0; c < RB; c++) t[c] = s[c];
}

// P[(q*64 + row)*RB + c] = sum_l U[uoff + row + l*ld] * T[(toff + l)*RB + c]:
// M = the item's 64 rows, N = right-hand sides, K = rank columns, four per step
template <int RB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
lr_u_multi_kernel(const LrItem* __restrict__ items, int nitems, const double* __restrict__ U,
                  const double* __restrict__ T, double* __restrict__ P)
{
    constexpr int NB = RB / 16;
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const LrItem it = items[q];
    const int i16 = lane & 15, g = lane >> 4;
    const int64_t ld = it.ld;
    const double* __restrict__ u = U + it.uoff + i16;
    const double* __restrict__ tb = T + it.toff * RB + i16;
    d4 acc[4][NB];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int h = 0; h < NB; h++) acc[t][h] = d4{0.0, 0.0, 0.0, 0.0};
    // operands of step l0 (rank column l0 + g; columns past K load the cluster's
    // last column and enter as zeros)
    auto load = [&](int l0, double (&a)[4], double (&b)[NB]) {
        const int l = l0 + g;
        const bool ok = l < it.K;
        const int lc = ok ? l : it.K - 1;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const double x = u[lc * ld + 16 * t];
            a[t] = ok ? x : 0.0;
        }
#pragma unroll
        for (int h = 0; h < NB; h++) {
            const double x = tb[(int64_t)lc * RB + 16 * h];
            b[h] = ok ? x : 0.0;
        }
    };
    double a[4], b[NB];
    load(0, a, b);
    for (int l0 = 0; l0 < it.K; l0 += 4) {
        double an[4], bn[NB];
        load(l0 + 4 < it.K ? l0 + 4 : l0, an, bn);
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int h = 0; h < NB; h++) acc[t][h] = mfma64(a[t], b[h], acc[t][h]);
#pragma unroll
        for (int t = 0; t < 4; t++) a[t] = an[t];
#pragma unroll
        for (int h = 0; h < NB; h++) b[h] = bn[h];
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * t + 4 * r + g;
            if (row < it.nrows) {
#pragma unroll
                for (int h = 0; h < NB; h++) P[((int64_t)q * kTileRows + row) * RB + 16 * h + i16] = acc[t][h][r];
            }
        }
}

// Near field on the fly over the stored-near-field schedule (items {row0, nrows <=
// 128, b0, b1} over near_blk), as FP64 MFMA GEMMs: M = 32 rows per pass over the
// columns (two 16-row tiles; four passes for 128 rows), N = right-hand sides, K =
// the item's dense blocks' columns in block order, four per step.  Two tiles per
// pass keep the kernel at 140 registers (three wavefronts per SIMD, launched
// without an LDS reservation); four tiles per pass took 248 (two per SIMD) and
// 86 ms alone at N = 2^20 against 76 ms.  Lane l evaluates the
// entries it feeds as A operand: rows 16t + l % 16 against column l / 16 of the
// step (each entry evaluated once for all right-hand sides).  Columns past a
// block's end enter as zeros.  P[(q*128 + r)*RB + c].
template <int KIND, int DIM, int RB>
__global__ void __launch_bounds__(kThreads)
near_multi_kernel(Coords c1, Coords c2, KernelParams kp, const int4* __restrict__ items, int nitems,
                  const int* __restrict__ near_blk, const int4* __restrict__ blk, const double* __restrict__ Xs,
                  double* __restrict__ P)
{
    constexpr int NB = RB / 16;
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const int4 it = items[q];
    const int dim = DIM > 0 ? DIM : kp.dim;
    const int i16 = lane & 15, g = lane >> 4;
    constexpr int NT = 2;          // 16-row tiles per pass over the columns
    constexpr int PR = 16 * NT;    // rows per pass
    for (int half = 0; half * PR < it.y; half++) {
        int row[NT];
        double xi[NT][MAX_POINT_DIMENSION];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int rr = half * PR + 16 * t + i16;
            row[t] = it.x + (rr < it.y ? rr : 0);  // rows past the item: its first row
#pragma unroll
            for (int d = 0; d < MAX_POINT_DIMENSION; d++) xi[t][d] = d < dim ? c1.x[d][row[t]] : 0.0;
        }
        d4 acc[NT][NB];
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int h = 0; h < NB; h++) acc[t][h] = d4{0.0, 0.0, 0.0, 0.0};
        // steps (block bi, first column j0) four columns at a time; lane l takes
        // column j0 + l / 16 (columns past the block's end load its last column and
        // enter as zeros).  The next step's coordinates and right-hand sides load
        // while this step's entries are evaluated (past the last step they repeat it).
        auto load = [&](int j0, int last, int& cc, bool& ok, double (&xj)[MAX_POINT_DIMENSION],
                        double (&bx)[NB]) {
            const int col = j0 + g;
            ok = col <= last;
            cc = ok ? col : last;
#pragma unroll
            for (int d = 0; d < MAX_POINT_DIMENSION; d++) xj[d] = d < dim ? c2.x[d][cc] : 0.0;
#pragma unroll
            for (int h = 0; h < NB; h++) {
                const double x = Xs[(int64_t)cc * RB + 16 * h + i16];
                bx[h] = ok ? x : 0.0;
            }
        };
        if (it.z < it.w) {  // an item without blocks writes zero partials
            int bi = it.z;
            int4 b = blk[near_blk[bi]];
            int j0 = b.z;
            int cc;
            bool ok;
            double xj[MAX_POINT_DIMENSION], bx[NB];
            load(j0, b.w, cc, ok, xj, bx);
            for (;;) {
                int bin = bi, j0n = j0 + 4;
                int4 bn = b;
                if (j0n > b.w) {
                    bin = bi + 1;
                    if (bin < it.w) {
                        bn = blk[near_blk[bin]];
                        j0n = bn.z;
                    } else {
                        j0n = j0;
                    }
                }
                int ccn;
                bool okn;
                double xjn[MAX_POINT_DIMENSION], bxn[NB];
                load(j0n, bn.w, ccn, okn, xjn, bxn);
                double a[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    double r2 = 0.0;
#pragma unroll
                    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
                        if (d < dim) {
                            const double dd = xi[t][d] - xj[d];
                            r2 = __fma_rn(dd, dd, r2);
                        }
                    }
                    double kv = kernel_of_r2_prod<KIND>(kp, r2);
                    if (row[t] == cc) kv += kp.reg;
                    a[t] = ok ? kv : 0.0;
                }
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int h = 0; h < NB; h++) acc[t][h] = mfma64(a[t], bx[h], acc[t][h]);
                if (bin >= it.w) break;
                bi = bin;
                b = bn;
                j0 = j0n;
                cc = ccn;
                ok = okn;
#pragma unroll
                for (int d = 0; d < MAX_POINT_DIMENSION; d++) xj[d] = xjn[d];
#pragma unroll
                for (int h = 0; h < NB; h++) bx[h] = bxn[h];
            }
        }
        double* __restrict__ out = P + ((int64_t)q * 2 * kTileRows + half * PR) * RB;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rr = 16 * t + 4 * r + g;
                if (half * PR + rr < it.y) {
#pragma unroll
                    for (int h = 0; h < NB; h++) out[(int64_t)rr * RB + 16 * h + i16] = acc[t][h][r];
                }
            }
    }
}

// Ys[i*RB + c] = sum of sorted row i's partials P[v*RB + c], in list order
template <int RB>
__global__ void __launch_bounds__(kThreads)
combine_multi_kernel(const int* __restrict__ ptr, const int* __restrict__ idx, const double* __restrict__ P,
                     double* __restrict__ Ys, int n1)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int i = (int)(t / RB), c = (int)(t % RB);
    if (i >= n1) return;
    double s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; q++) s += P[(int64_t)idx[q] * RB + c];
    Ys[t] = s;
}

// Y[p + (c0 + c)*n1] = Ys[inv1[p]*RB + c] for c < nr, through LDS (64 rows per workgroup)
template <int RB>
__global__ void __launch_bounds__(kThreads)
permute_out_multi_kernel(const double* __restrict__ Ys, const int* __restrict__ inv1, double* __restrict__ Y,
                         int n1, int c0, int nr)
{
    __shared__ double tile[RB][kRowTile + 1];
    const int p0 = blockIdx.x * kRowTile;
    for (int e = threadIdx.x; e < kRowTile * RB; e += kThreads) {
        const int pl = e / RB, c = e % RB;
        const int p = p0 + pl;
        tile[c][pl] = p < n1 ? Ys[(int64_t)inv1[p] * RB + c] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kRowTile * RB; e += kThreads) {
        const int c = e / kRowTile, pl = e % kRowTile;
        const int p = p0 + pl;
        if (c < nr && p < n1) Y[p + (int64_t)(c0 + c) * n1] = tile[c][pl];
    }
}

__global__ void inv_rows_kernel(const uint64_t* __restrict__ order, int* __restrict__ inv, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) inv[order[i]] = i;
}

template <int KIND, int DIM, int RB>
void launch_near_multi(hipStream_t s, const Coords& c1, const Coords& c2, const KernelParams& kp,
                       const Engine& eng, double* Pn)
{
    near_multi_kernel<KIND, DIM, RB><<<grid_for(eng.n_near_items, kWaves), kThreads, 0, s>>>(
        c1, c2, kp, eng.near_items.p, eng.n_near_items, eng.near_blk.p, eng.blk.p, eng.X64.p, Pn);
}

template <int RB>
hmg_status pass(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp, const uint64_t* order2,
                const double* X, double* Y, int c0, int nr, const double* U, const double* V)
{
    hipStream_t s = eng.stream, s2 = eng.serialize ? eng.stream : eng.stream2;
    const int n1 = eng.n[0], n2 = eng.n[1];
    eng.prof.begin(PROF_GATHER, s);
    gather_rows_kernel<RB><<<grid_for(n2, kRowTile), kThreads, 0, s>>>(X, order2, n2, c0, nr, eng.X64.p);
    HMG_HIP(hipGetLastError());
    eng.prof.end(PROF_GATHER, s);
    HMG_HIP(hipEventRecord(eng.ev_fork, s));
    HMG_HIP(hipStreamWaitEvent(s2, eng.ev_fork, 0));
    double* Pn = eng.P64.p + (int64_t)eng.n_lr_items * kTileRows * RB;
    if (eng.n_near_items > 0) {
        eng.prof.begin(PROF_NEAR, s2);
        HMG_DISPATCH_KIND(kp.kind, {
            if (kp.dim == 3)
                launch_near_multi<KIND, 3, RB>(s2, c1, c2, kp, eng, Pn);
            else if (kp.dim == 2)
                launch_near_multi<KIND, 2, RB>(s2, c1, c2, kp, eng, Pn);
            else
                launch_near_multi<KIND, 0, RB>(s2, c1, c2, kp, eng, Pn);
        });
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_NEAR, s2);
    }
    HMG_HIP(hipEventRecord(eng.ev_join, s2));
    if (eng.naca > 0 && eng.n_lr_items > 0) {
        eng.prof.begin(PROF_LR_VTX, s);
        lr_v_multi_kernel<RB><<<grid_for(eng.n_lv_items, kWaves), kThreads, 0, s>>>(eng.lv_items.p, eng.n_lv_items,
                                                                                  V, eng.X64.p, eng.Tp64.p);
        HMG_HIP(hipGetLastError());
        lr_v_finalize_multi_kernel<RB><<<grid_for(eng.n_lv_slices, kWaves), kThreads, 0, s>>>(
            eng.lv_slices.p, eng.n_lv_slices, eng.Tp64.p, eng.tmap.p, eng.T64.p);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_LR_VTX, s);
        eng.prof.begin(PROF_LR_U, s);
        lr_u_multi_kernel<RB><<<grid_for(eng.n_lr_items, kWaves), kThreads, 0, s>>>(eng.lr_items.p, eng.n_lr_items,
                                                                                  U, eng.T64.p, eng.P64.p);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_LR_U, s);
    }
    HMG_HIP(hipStreamWaitEvent(s, eng.ev_join, 0));
    eng.prof.begin(PROF_COMBINE, s);
    combine_multi_kernel<RB><<<grid_for((int64_t)n1 * RB, kThreads), kThreads, 0, s>>>(
        eng.comb_stored.ptr.p, eng.comb_stored.idx.p, eng.P64.p, eng.Ys64.p, n1);
    HMG_HIP(hipGetLastError());
    permute_out_multi_kernel<RB><<<grid_for(n1, kRowTile), kThreads, 0, s>>>(eng.Ys64.p, eng.inv1.p, Y, n1, c0, nr);
    HMG_HIP(hipGetLastError());
    eng.prof.end(PROF_COMBINE, s);
    return HMG_OK;
}

}  // namespace

hmg_status hmvp_multi(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order1, const uint64_t* order2, const double* X, double* Y, int nrhs,
                      const double* U, const double* V)
{
    if (eng.layout_prec != HMG_PREC_FP64) {
        set_error("FP64 block product without an FP64 layout");
        return HMG_ERR_NOT_READY;
    }
    hipStream_t s = eng.stream;
    const int n1 = eng.n[0], n2 = eng.n[1];
    const int rb = nrhs > 16 ? kRbMax : 16;
    // partial-sum slots of the stored-near-field schedule (the combine lists')
    const int64_t slots = (int64_t)eng.n_lr_items * kTileRows + (int64_t)eng.n_near_items * 2 * kTileRows;
    HMG_TRY(eng.X64.alloc((size_t)std::max(n2, 1) * rb));
    HMG_TRY(eng.Tp64.alloc((size_t)std::max(eng.n_lv_items, 1) * kTileRows * rb));
    HMG_TRY(eng.T64.alloc((size_t)std::max<int64_t>(eng.tsize, 1) * rb));
    HMG_TRY(eng.P64.alloc((size_t)std::max<int64_t>(slots, 1) * rb));
    HMG_TRY(eng.Ys64.alloc((size_t)std::max(n1, 1) * rb));
    if (!eng.inv1_ready) {
        HMG_TRY(eng.inv1.alloc(std::max(n1, 1)));
        inv_rows_kernel<<<grid_for(n1, kThreads), kThreads, 0, s>>>(order1, eng.inv1.p, n1);
        HMG_HIP(hipGetLastError());
        eng.inv1_ready = true;
    }
    for (int c0 = 0; c0 < nrhs; c0 += rb) {
        const int nr = std::min(rb, nrhs - c0);
        hmg_status st;
        if (rb == 32)
            st = pass<32>(eng, c1, c2, kp, order2, X, Y, c0, nr, U, V);
        else
            st = pass<16>(eng, c1, c2, kp, order2, X, Y, c0, nr, U, V);
        HMG_TRY(st);
    }
    return HMG_OK;
}

}  // namespace hmg
