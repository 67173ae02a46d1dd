// aca.hip -- batched adaptive cross approximation, one wavefront per admissible
// block, all rank steps inside one launch (no host round trips).
//
// Semantics follow the reference's batched ACA (linear_algebra.cu:3684-4012,
// kernels :918-1147, :963-1018, :2009-2222) with the pivot race at :985 resolved
// to its intended meaning (divide by the pivot value read before scaling):
//   k_b = min(k, m1, m2); row pivots are taken sequentially (0, 1, 2, ...),
//   skipping rows whose residual has 2-norm < 1e-13; a block stops for good when
//   its rows run out; the column pivot is the LAST index of max |v| (tie_last=1)
//   or the FIRST (tie_last=0, the unbatched path linear_algebra.cu:473).
//   Residuals: v_j = K(i_r, j) - sum_{l<r} U[i_r,l] V[j,l], u_i likewise, each
//   accumulated as fma(-s, w, val) in l order (nvcc's contraction of :955/:1143).
// The 2-norm is a wavefront FP64 reduction; the argmax a (|v|, index, v) butterfly.
#include <algorithm>

#include "hmg_internal.h"
#include "kernel_eval.h"

namespace hmg {

namespace {


__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// val - sum_{l < r} p[l*sp] * q[l*sq], accumulated as fma(-p, q, val) in l order
// (the reference's contraction order, linear_algebra.cu:955/:1143).  The loads of
// sixteen (then eight) consecutive l are issued before their FMAs, so a step waits
// for one memory latency per sixteen terms instead of one per term.
template <class T>
__device__ __forceinline__ double residual_update(double val, int r, const T* __restrict__ p, int64_t sp,
                                                  const T* __restrict__ q, int64_t sq)
{
    int l = 0;
    for (; l + 16 <= r; l += 16) {
        T pv[16], qv[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            pv[e] = p[(int64_t)(l + e) * sp];
            qv[e] = q[(int64_t)(l + e) * sq];
        }
#pragma unroll
        for (int e = 0; e < 16; e++) val = __fma_rn(-(double)pv[e], (double)qv[e], val);
    }
    for (; l + 8 <= r; l += 8) {
        T pv[8], qv[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            pv[e] = p[(int64_t)(l + e) * sp];
            qv[e] = q[(int64_t)(l + e) * sq];
        }
#pragma unroll
        for (int e = 0; e < 8; e++) val = __fma_rn(-(double)pv[e], (double)qv[e], val);
    }
    for (; l < r; l++) val = __fma_rn(-(double)p[(int64_t)l * sp], (double)q[(int64_t)l * sq], val);
    return val;
}

// A team runs the ACA of one block: one wavefront (NW = 1) for most blocks, or a
// whole workgroup of NW wavefronts for the few large blocks, whose rank loop is
// otherwise the critical path of the launch.  Team reductions (row norm,
// column-pivot argmax) go through LDS when NW > 1; every thread ends with the
// same value.
template <int NW>
struct Team {
    double* s_v;
    int* s_i;
    double* s_x;
    __device__ int tid() const { return NW == 1 ? (int)(threadIdx.x % kWave) : (int)threadIdx.x; }
    static constexpr int size = NW * kWave;
    __device__ void sync() const
    {
        if (NW == 1)
            __threadfence_block();
        else
            __syncthreads();
    }
    __device__ double sum(double v) const
    {
        v = wave_sum(v);
        if (NW == 1) return v;
        const int w = threadIdx.x / kWave;
        if (threadIdx.x % kWave == 0) s_v[w] = v;
        __syncthreads();
        double t = 0.0;
        for (int q = 0; q < NW; q++) t += s_v[q];
        __syncthreads();
        return t;
    }
    // R sums at once, each equal to sum() of its value
    template <int R>
    __device__ void sum_n(double (&v)[R]) const
    {
#pragma unroll
        for (int c = 0; c < R; c++) v[c] = wave_sum(v[c]);
        if (NW == 1) return;
        const int w = threadIdx.x / kWave;
        if (threadIdx.x % kWave == 0) {
#pragma unroll
            for (int c = 0; c < R; c++) s_v[w * R + c] = v[c];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < R; c++) {
            double t = 0.0;
            for (int q = 0; q < NW; q++) t += s_v[q * R + c];
            v[c] = t;
        }
        __syncthreads();
    }
    // team minimum and maximum (every thread ends with the same value)
    template <bool MAX>
    __device__ double extreme(double v) const
    {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, kWave);
            v = MAX ? fmax(v, o) : fmin(v, o);
        }
        if (NW == 1) return v;
        const int w = threadIdx.x / kWave;
        if (threadIdx.x % kWave == 0) s_v[w] = v;
        __syncthreads();
        double t = s_v[0];
        for (int q = 1; q < NW; q++) t = MAX ? fmax(t, s_v[q]) : fmin(t, s_v[q]);
        __syncthreads();
        return t;
    }
    // (|v|, index, v) maximum with the tie rule: last index (tie_last) or first
    __device__ static bool better(double ob, int oj, double best, int bj, int tie_last)
    {
        return (ob > best) || (ob == best && (tie_last ? (oj > bj) : (oj < bj && oj >= 0)));
    }
    __device__ void argmax(double& best, int& bj, double& bval, int tie_last) const
    {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, kWave);
            const int oj = __shfl_xor(bj, off, kWave);
            const double ov = __shfl_xor(bval, off, kWave);
            if (better(ob, oj, best, bj, tie_last)) {
                best = ob;
                bj = oj;
                bval = ov;
            }
        }
        if (NW == 1) return;
        const int w = threadIdx.x / kWave;
        if (threadIdx.x % kWave == 0) {
            s_v[w] = best;
            s_i[w] = bj;
            s_x[w] = bval;
        }
        __syncthreads();
        best = s_v[0];
        bj = s_i[0];
        bval = s_x[0];
        for (int q = 1; q < NW; q++) {
            if (better(s_v[q], s_i[q], best, bj, tie_last)) {
                best = s_v[q];
                bj = s_i[q];
                bval = s_x[q];
            }
        }
        __syncthreads();
    }
};

// Tolerance stop of the ACA (the reference's batched path has none: its epsilon
// is unused).  mode 1: linear_algebra.cu:498-528 with :379 corrected, checked when
// r % 5 == 0; mode 2: checked at every step.  Stop after rank step r when
// ||u_r|| ||v_r|| <= c ||U V^T||_F over the first r+1 terms, c = eps (1-eta) /
// (1+eps) (mode 1) or eps (mode 2).  ||U V^T||_F^2 is kept as a running sum:
// + 2 sum_{l<r} (u_l.u_r)(v_l.v_r) + ||u_r||^2 ||v_r||^2.
struct AcaTol {
    int mode;
    double c;
};

// Rows scanned at once when the residual-row search rejects a row, and the
// largest rank step that scans (its U terms are staged per wavefront in LDS;
// later steps keep the one-row search).
constexpr int kScanRows = 4;
// Factor columns per pass over the rows of V in the tolerance stop's cross terms
// (each lane reads that many consecutive values of its row; the U side goes
// kScanRows columns at a time, each a coalesced column read).
constexpr int kCrossCols = 8;
static_assert(kCrossCols >= 2 * kScanRows, "team sums share one LDS array");
// Smallest block width whose rejected-row search first bounds the residual rows
// (below it a batch of rows costs about as much as the bounds).
constexpr int kBoundMinCols = 512;
constexpr int kScanL = 64;

// Squared 2-norms of the residual rows i0 .. i0+R-1 at rank step r (rows
// past m1 - 1 give values that the caller ignores), without storing the rows.
// Each entry and each per-thread partial sum is formed exactly as the one-row
// search forms it (same l order of the fma chain, same j order of the squares),
// and the team sums are the same, so a row's norm here equals its norm there
// bitwise.  The R rows are independent chains sharing one load of each
// V entry, and their U terms (the same for every lane) are read from the
// wavefront's LDS slice sU[l][c], so a scan of the many rejected rows of a
// numerically rank-deficient block runs at arithmetic rate rather than one
// memory latency per term.
template <class T, int NW, int KIND, int R>
__device__ void scan_row_norms(const Team<NW>& team, Coords c1, Coords c2, const KernelParams& kp, int l1,
                               int m1, int l2, int m2, int i0, int r, int64_t ld, int64_t lu,
                               const T* __restrict__ Ub, const T* __restrict__ Vb, T* __restrict__ sU,
                               double* __restrict__ sX, double (&nrm)[R])
{
    constexpr int TS = Team<NW>::size;
    const int lane = threadIdx.x % kWave;
    int ic[R];
#pragma unroll
    for (int c = 0; c < R; c++) {
        ic[c] = min(i0 + c, m1 - 1);
        nrm[c] = 0.0;
    }
    // the rows' coordinates, sX[c * MAX_POINT_DIMENSION + d] (read back from LDS
    // instead of keeping R x dim addresses live in registers), and their U terms
    // U[ic[c], l] at sU[l * R + c], l < r <= kScanL
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < R * MAX_POINT_DIMENSION) {
        const int c = lane / MAX_POINT_DIMENSION, d = lane % MAX_POINT_DIMENSION;
        sX[lane] = d < kp.dim ? c1.x[d][l1 + min(i0 + c, m1 - 1)] : 0.0;
    }
    for (int q = lane; q < R * r; q += kWave) {
        const int c = q % R, l = q / R;
        sU[q] = Ub[(int64_t)l * lu + min(i0 + c, m1 - 1)];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int j = team.tid(); j < m2; j += TS) {
        // K(i, j) exactly as kernel_entry_k forms it (same r^2 order), row
        // coordinates from sX
        double xj[MAX_POINT_DIMENSION];
#pragma unroll
        for (int d = 0; d < MAX_POINT_DIMENSION; d++) xj[d] = d < kp.dim ? c2.x[d][l2 + j] : 0.0;
        double val[R];
#pragma unroll
        for (int c = 0; c < R; c++) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
                if (d < kp.dim) {
                    const double t = sX[c * MAX_POINT_DIMENSION + d] - xj[d];
                    r2 = __dadd_rn(r2, __dmul_rn(t, t));
                }
            }
            double v = kernel_of_r2_k<KIND>(kp, r2);
            if (l1 + ic[c] == l2 + j) v += kp.reg;
            val[c] = v;
        }
        const T* __restrict__ vrow = Vb + (int64_t)j * ld;
        int l = 0;
        for (; l + 8 <= r; l += 8) {
            T pv[8];
#pragma unroll
            for (int e = 0; e < 8; e++) pv[e] = vrow[l + e];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const T* u = sU + (l + e) * R;
#pragma unroll
                for (int c = 0; c < R; c++) val[c] = __fma_rn(-(double)pv[e], (double)u[c], val[c]);
            }
        }
        for (; l < r; l++) {
            const double pv = (double)vrow[l];
            const T* u = sU + l * R;
#pragma unroll
            for (int c = 0; c < R; c++) val[c] = __fma_rn(-pv, (double)u[c], val[c]);
        }
#pragma unroll
        for (int c = 0; c < R; c++) nrm[c] = __fma_rn(val[c], val[c], nrm[c]);
    }
    team.template sum_n<R>(nrm);
}

// True when every entry of the block is so small that every row of the first
// rank step has residual norm below the 1e-13 row threshold, i.e. the search would
// reject all m1 rows and the block gets rank 0 after evaluating all m1 x m2
// entries.  Only for kernels that decrease with the distance and when no entry
// carries the regularization (it is added where sorted indices coincide).  The
// tree's boxes only pre-screen (a cluster of one point can carry another
// cluster's box); the decision uses the bounding boxes of the block's points:
// |K(i,j)| <= k(d_min), and sqrt(m2) k(d_min) < 1e-13 / 2 leaves a factor two for
// the rounding of the evaluated norms.
template <int NW, int KIND>
__device__ bool block_negligible(const Team<NW>& team, Coords c1, Coords c2, const KernelParams& kp, int4 b,
                                 const work_item& wi)
{
    if (KIND != HMG_KERNEL_GAUSSIAN && KIND != HMG_KERNEL_EXPONENTIAL && KIND != HMG_KERNEL_GAUSSIAN_REF)
        return false;
    if (kp.reg != 0.0 && b.x <= b.w && b.z <= b.y) return false;
    const double m2 = (double)(b.w - b.z + 1);
    const double thr = 0.5e-13;
    // pre-screen with the tree's boxes, loosely
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
        if (d < kp.dim) {
            const double gap = fmax(0.0, fmax(wi.min1[d] - wi.max2[d], wi.min2[d] - wi.max1[d]));
            r2 += gap * gap;
        }
    }
    if (sqrt(m2) * kernel_of_r2_k<KIND>(kp, 0.5 * r2) >= thr) return false;
    // exact boxes of the points
    r2 = 0.0;
    const int tid = team.tid();
    constexpr int TS = Team<NW>::size;
#pragma unroll
    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
        if (d >= kp.dim) continue;
        double lo1 = 1e300, hi1 = -1e300, lo2 = 1e300, hi2 = -1e300;
        for (int i = b.x + tid; i <= b.y; i += TS) {
            const double x = c1.x[d][i];
            lo1 = fmin(lo1, x);
            hi1 = fmax(hi1, x);
        }
        for (int j = b.z + tid; j <= b.w; j += TS) {
            const double x = c2.x[d][j];
            lo2 = fmin(lo2, x);
            hi2 = fmax(hi2, x);
        }
        lo1 = team.template extreme<false>(lo1);
        hi1 = team.template extreme<true>(hi1);
        lo2 = team.template extreme<false>(lo2);
        hi2 = team.template extreme<true>(hi2);
        const double gap = fmax(0.0, fmax(lo1 - hi2, lo2 - hi1));
        r2 += gap * gap;
    }
    // r2 shrunk by far more than its rounding: the bound stays an upper bound
    return sqrt(m2) * kernel_of_r2_k<KIND>(kp, r2 * (1.0 - 1e-9)) < thr;
}

// Kernels whose entries are bounded by k(distance to a box), k decreasing.
template <int KIND>
constexpr bool kind_monotone()
{
    return KIND == HMG_KERNEL_GAUSSIAN || KIND == HMG_KERNEL_EXPONENTIAL || KIND == HMG_KERNEL_GAUSSIAN_REF;
}

// Columns' bounding box of the block into LDS: low end of dimension d at box[d],
// high end at box[MAX_POINT_DIMENSION + d] (each wavefront of a team writes its
// own copy of the same values).
template <int NW>
__device__ void column_box(const Team<NW>& team, Coords c2, const KernelParams& kp, int4 b, double* __restrict__ box)
{
    const int tid = team.tid();
    const int lane = threadIdx.x % kWave;
    constexpr int TS = Team<NW>::size;
#pragma unroll
    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
        if (d < kp.dim) {
            double lo = 1e300, hi = -1e300;
            for (int j = b.z + tid; j <= b.w; j += TS) {
                const double x = c2.x[d][j];
                lo = fmin(lo, x);
                hi = fmax(hi, x);
            }
            lo = team.template extreme<false>(lo);
            hi = team.template extreme<true>(hi);
            if (lane == 0) {
                box[d] = lo;
                box[MAX_POINT_DIMENSION + d] = hi;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First row i >= start of the block (relative index; m1 when none) whose residual
// at rank step r the search could accept.  Row i's residual is
// K(i,:) - sum_{l<r} U(i,l) V(:,l)^T, so its norm is at most
//   sqrt(m2) k(d_i) + sum_{l<r} |U(i,l)| ||V(:,l)||,
// d_i the distance from point i to the columns' bounding box (vn[l] >= ||V(:,l)||,
// rounded up); rows whose bound is below 1e-13 / 2 have norms below the row
// threshold, so the search rejects them, and they are passed over unevaluated.
// The factor two covers the rounding of the evaluated norms (the same margin as
// block_negligible).
template <class T, int NW, int KIND>
__device__ int first_row_over_bound(const Team<NW>& team, Coords c1, const KernelParams& kp, int4 b, int start,
                                    int r, const T* __restrict__ Ub, int64_t lu, const double* __restrict__ box,
                                    const double* __restrict__ vn)
{
    const int m1 = b.y - b.x + 1;
    const int tid = team.tid();
    constexpr int TS = Team<NW>::size;
    const double sm2 = sqrt((double)(b.w - b.z + 1));
    const double thr = 0.5e-13;
    int found = m1;
    for (int i0 = start; i0 < m1; i0 += TS) {
        double cand = (double)m1;
        const int i = i0 + tid;
        if (i < m1) {
            double r2 = 0.0;
#pragma unroll
            for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
                if (d < kp.dim) {
                    const double x = c1.x[d][b.x + i];
                    const double gap = fmax(0.0, fmax(box[d] - x, x - box[MAX_POINT_DIMENSION + d]));
                    r2 += gap * gap;
                }
            }
            // r2 shrunk by far more than its rounding: the bound stays an upper bound
            double bound = sm2 * kernel_of_r2_k<KIND>(kp, r2 * (1.0 - 1e-9));
            for (int l = 0; l < r; l++) bound = __fma_rn(fabs((double)Ub[(int64_t)l * lu + i]), vn[l], bound);
            if (bound * (1.0 + 1e-9) >= thr) cand = (double)i;
        }
        cand = team.template extreme<false>(cand);
        if (cand < (double)m1) {
            found = (int)cand;
            break;
        }
    }
    return found;
}

// Factor layout: U_b[i, l] at U + uoff[a] + l * ldu[a] + i (column-major, ldu >= m1:
// the row cluster's padded height); V_b is stored transposed, V_b[j, l] at
// V + voff[a] + j * ldv[a] + l, i.e. row j of V_b is contiguous (ldv >= total rank
// of the column cluster's blocks).  T is the storage type of the factors (the
// engine instantiates double only); every residual, norm and pivot division is
// computed in FP64 from the stored values and rounded to T when stored.
// Entry j of a row (i of a column) is always handled by the same thread, so each
// entry is computed exactly as in a single-wavefront team.
template <class T, int NW, int KIND, bool TOL>
__device__ void aca_block(const Team<NW>& team, Coords c1, Coords c2, const KernelParams& kp, int4 b, int k,
                          int64_t ld, int64_t lu, T* __restrict__ Ub, T* __restrict__ Vb, int* keff_out,
                          int tie_last, int scan, T* __restrict__ sU, double* __restrict__ sX,
                          double* __restrict__ box, double* __restrict__ vn, const work_item& wi, AcaTol tol)
{
    const int tid = team.tid();
    constexpr int TS = Team<NW>::size;
    // rows per batch of the rejected-row search: the largest teams scan twice as
    // many at once (one load of each V entry serves all of them)
    constexpr int SR = NW >= 8 ? 2 * kScanRows : kScanRows;
    const int l1 = b.x, m1 = b.y - b.x + 1;
    const int l2 = b.z, m2 = b.w - b.z + 1;
    if (scan && k > 0 && block_negligible<NW, KIND>(team, c1, c2, kp, b, wi)) {
        // what the search ends with after rejecting every row: rank 0, column 0 of
        // V cleared
        for (int j = tid; j < m2; j += TS) Vb[(int64_t)j * ld] = (T)0;
        if (tid == 0) *keff_out = 0;
        return;
    }
    int i_r = -1;
    int ke = 0;
    double fro2 = 0.0;  // ||U V^T||_F^2 of the terms so far (tolerance stop)
    // residual bounds for the rejected-row search (wide blocks, decreasing kernels)
    const bool bound_rows = scan && m2 >= kBoundMinCols && kind_monotone<KIND>();
    bool box_ready = false;
    const int lane = threadIdx.x % kWave;
    for (int r = 0; r < k; r++) {
        bool stop = false;
        // this lane's column-pivot candidate (|v|, index, v) of the accepted row,
        // taken while the row is formed, in the order of the lane's columns
        double best = -1.0, bval = 0.0;
        int bj = -1;
        // residual-row search (compute_current_batched_v_r, :2068-2132)
        for (;;) {
            if (i_r >= m1 - 1) {
                stop = true;
                break;
            }
            i_r++;
            double nrm = 0.0;
            best = -1.0;
            bval = 0.0;
            bj = -1;
            for (int j = tid; j < m2; j += TS) {
                double val = kernel_entry_k<KIND>(kp, c1, l1 + i_r, c2, l2 + j);
                const T* __restrict__ vrow = Vb + (int64_t)j * ld;
                val = residual_update<T>(val, r, vrow, 1, Ub + i_r, lu);
                const T stored = (T)val;
                Vb[(int64_t)j * ld + r] = stored;
                nrm = __fma_rn(val, val, nrm);
                const double x = (double)stored;
                const double ax = fabs(x);
                if (tie_last ? (ax >= best) : (ax > best)) {
                    best = ax;
                    bj = j;
                    bval = x;
                }
            }
            nrm = team.sum(nrm);
            if (sqrt(nrm) >= 1.0e-13) break;
            // row i_r rejected: look for the next acceptable row SR rows at a
            // time; i_r is left just before it (the one-row search above then
            // recomputes and stores it), or at m1 - 1 when every row is rejected
            if (!scan || r > kScanL) continue;
            // wide blocks: rows the residual bound rejects are passed over unevaluated
            // before each batch
            const bool skip = bound_rows && r <= kScanL;
            if (skip && !box_ready) {
                column_box<NW>(team, c2, kp, b, box);
                box_ready = true;
            }
            int found = m1;
            for (int i0 = i_r + 1; i0 < m1; i0 += SR) {
                if (skip) {
                    i0 = first_row_over_bound<T, NW, KIND>(team, c1, kp, b, i0, r, Ub, lu, box, vn);
                    if (i0 >= m1) break;
                }
                double rn[SR];
                scan_row_norms<T, NW, KIND, SR>(team, c1, c2, kp, l1, m1, l2, m2, i0, r, ld, lu, Ub, Vb, sU, sX,
                                                rn);
                for (int c = 0; c < SR && i0 + c < m1; c++) {
                    if (sqrt(rn[c]) >= 1.0e-13) {
                        found = i0 + c;
                        break;
                    }
                }
                if (found < m1) break;
            }
            i_r = found - 1;
        }
        if (stop) {
            // the rank stops here: column r of V (last rejected row) is cleared, column r
            // of U was never written; both stay zero in the product
            for (int j = tid; j < m2; j += TS) Vb[(int64_t)j * ld + r] = (T)0;
            break;
        }
        // column pivot (tuple_absolute_maximum, :991-1018) over the lanes' candidates
        team.argmax(best, bj, bval, tie_last);
        const int jr = bj;
        const double p = bval;
        // v_r /= p and the residual column (batched_fill_kernel_vector_and_scaled_
        // substraction_for_u_r) in one pass: the column reads columns l < r of V only
        double vv = 0.0;
        const T* __restrict__ vpiv = Vb + (int64_t)jr * ld;
        T* __restrict__ u = Ub + r * lu;
        const int mm = m1 > m2 ? m1 : m2;
        for (int t = tid; t < mm; t += TS) {
            if (t < m2) {
                const T v = (T)((double)Vb[(int64_t)t * ld + r] / p);
                Vb[(int64_t)t * ld + r] = v;
                vv = __fma_rn((double)v, (double)v, vv);
            }
            if (t < m1) {
                double val = kernel_entry_k<KIND>(kp, c1, l1 + t, c2, l2 + jr);
                val = residual_update<T>(val, r, Ub + t, lu, vpiv, 1);
                u[t] = (T)val;
            }
        }
        if (bound_rows && r < kScanL) {
            // ||V(:,r)|| for the residual bounds, rounded up
            vv = team.sum(vv);
            if (lane == 0) vn[r] = sqrt(vv) * (1.0 + 1e-9);
        }
        team.sync();
        ke = r + 1;
        if (TOL) {
            const T* __restrict__ ur = Ub + (int64_t)r * lu;
            double nu = 0.0, nv = 0.0;
            for (int i = tid; i < m1; i += TS) nu = __fma_rn((double)ur[i], (double)ur[i], nu);
            for (int j = tid; j < m2; j += TS) {
                const double x = (double)Vb[(int64_t)j * ld + r];
                nv = __fma_rn(x, x, nv);
            }
            nu = team.sum(nu);
            nv = team.sum(nv);
            double cross = 0.0;
            for (int l0 = 0; l0 < r; l0 += kCrossCols) {
                // u_l . u_r of the columns l0 .. l0 + kCrossCols - 1, kScanRows columns
                // per pass over the rows of U (each a coalesced column read), kept in
                // the wavefront's LDS slice
                for (int c0 = l0; c0 < r && c0 < l0 + kCrossCols; c0 += kScanRows) {
                    double pu[kScanRows];
#pragma unroll
                    for (int e = 0; e < kScanRows; e++) pu[e] = 0.0;
                    for (int i = tid; i < m1; i += TS) {
                        const double x = (double)ur[i];
#pragma unroll
                        for (int e = 0; e < kScanRows; e++)
                            if (c0 + e < r) pu[e] = __fma_rn((double)Ub[(int64_t)(c0 + e) * lu + i], x, pu[e]);
                    }
                    team.template sum_n<kScanRows>(pu);
                    if (lane == 0) {
#pragma unroll
                        for (int e = 0; e < kScanRows; e++) sU[c0 - l0 + e] = (T)pu[e];
                    }
                }
                // v_l . v_r of the same columns in one pass over the rows of V (each
                // lane reads kCrossCols consecutive values of its row)
                double pv[kCrossCols];
#pragma unroll
                for (int e = 0; e < kCrossCols; e++) pv[e] = 0.0;
                for (int j = tid; j < m2; j += TS) {
                    const T* __restrict__ vj = Vb + (int64_t)j * ld;
                    const double x = (double)vj[r];
#pragma unroll
                    for (int e = 0; e < kCrossCols; e++)
                        if (l0 + e < r) pv[e] = __fma_rn((double)vj[l0 + e], x, pv[e]);
                }
                team.template sum_n<kCrossCols>(pv);
                __threadfence_block();
#pragma unroll
                for (int e = 0; e < kCrossCols; e++)
                    if (l0 + e < r) cross = __fma_rn((double)sU[e], pv[e], cross);
                __threadfence_block();
            }
            fro2 = fro2 + 2.0 * cross + nu * nv;
            if ((tol.mode == 2 || r % 5 == 0) && sqrt(nu) * sqrt(nv) <= tol.c * sqrt(fmax(fro2, 0.0))) break;
        }
    }
    if (tid == 0) *keff_out = ke;
}

// one wavefront per block of list[] (indices into the ACA blocks)
// One launch for every admissible block: nbig workgroups each run the ACA of one
// large block (list[0 .. nbig)) with all kAcaTeam wavefronts; every other
// workgroup runs kAcaTeam small blocks (list[nbig ..)), one per wavefront.  Three
// wavefronts per SIMD: the rank loop is latency-bound, and 168 registers let each
// wavefront keep sixteen rank terms' loads in flight (and the tolerance stop's
// cross terms) without scratch memory.
constexpr int kAcaTeam = 4;
template <class T, int KIND, bool TOL>
__global__ void __launch_bounds__(kAcaTeam * kWave) __attribute__((amdgpu_waves_per_eu(3)))
aca_kernel(Coords c1, Coords c2, KernelParams kp, const int4* __restrict__ blk,
           const work_item* __restrict__ items, const int* __restrict__ list,
           int nbig, int nsmall, const int* __restrict__ kb, const int64_t* __restrict__ uoff,
           const int* __restrict__ ldu, const int64_t* __restrict__ voff, const int* __restrict__ ldv,
           T* __restrict__ U, T* __restrict__ V, T* const* __restrict__ ublk, T* const* __restrict__ vblk,
           int* __restrict__ keff, int tie_last, int scan, AcaTol tol)
{
    __shared__ double s_v[kAcaTeam * kCrossCols];
    __shared__ int s_i[kAcaTeam];
    __shared__ double s_x[kAcaTeam];
    __shared__ T sU[kAcaTeam][kScanRows * kScanL];
    __shared__ double sX[kAcaTeam][kScanRows * MAX_POINT_DIMENSION];
    __shared__ double sBox[kAcaTeam][2 * MAX_POINT_DIMENSION];
    __shared__ double sVn[kAcaTeam][kScanL];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // large blocks first: their long rank loops start at once
    if ((int)blockIdx.x < nbig) {
        const int a = list[blockIdx.x];
        const Team<kAcaTeam> team{s_v, s_i, s_x};
        aca_block<T, kAcaTeam, KIND, TOL>(team, c1, c2, kp, blk[a], kb[a], ldv[a], ldu[a],
                                          ublk ? ublk[a] : U + uoff[a], vblk ? vblk[a] : V + voff[a],
                                     keff + a, tie_last, scan, sU[w], sX[w], sBox[w], sVn[w], items[a], tol);
    } else {
        const int q = ((int)blockIdx.x - nbig) * kAcaTeam + w;
        if (q >= nsmall) return;
        const int a = list[nbig + q];
        const Team<1> team{nullptr, nullptr, nullptr};
        aca_block<T, 1, KIND, TOL>(team, c1, c2, kp, blk[a], kb[a], ldv[a], ldu[a],
                                   ublk ? ublk[a] : U + uoff[a], vblk ? vblk[a] : V + voff[a],
                              keff + a, tie_last, scan, sU[w], sX[w], sBox[w], sVn[w], items[a], tol);
    }
}

// The largest blocks, one 8-wavefront workgroup each, in a launch of their own
// that runs beside the main one: their rank loops and rejected-row searches
// (m1 x m2 x rank work when most rows are rejected) set the length of the build.
// Two wavefronts per SIMD leave room for every value in registers (no scratch).
constexpr int kAcaHuge = 8;
template <class T, int KIND, bool TOL>
__global__ void __launch_bounds__(kAcaHuge * kWave) __attribute__((amdgpu_waves_per_eu(2)))
aca_huge_kernel(Coords c1, Coords c2, KernelParams kp, const int4* __restrict__ blk,
                const work_item* __restrict__ items, const int* __restrict__ list,
                const int* __restrict__ kb, const int64_t* __restrict__ uoff, const int* __restrict__ ldu,
                const int64_t* __restrict__ voff, const int* __restrict__ ldv, T* __restrict__ U,
                T* __restrict__ V, T* const* __restrict__ ublk, T* const* __restrict__ vblk,
                int* __restrict__ keff, int tie_last, int scan, AcaTol tol)
{
    __shared__ double s_v[kAcaHuge * kCrossCols];
    __shared__ int s_i[kAcaHuge];
    __shared__ double s_x[kAcaHuge];
    __shared__ T sU[kAcaHuge][2 * kScanRows * kScanL];
    __shared__ double sX[kAcaHuge][2 * kScanRows * MAX_POINT_DIMENSION];
    __shared__ double sBox[kAcaHuge][2 * MAX_POINT_DIMENSION];
    __shared__ double sVn[kAcaHuge][kScanL];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int a = list[blockIdx.x];
    const Team<kAcaHuge> team{s_v, s_i, s_x};
    aca_block<T, kAcaHuge, KIND, TOL>(team, c1, c2, kp, blk[a], kb[a], ldv[a], ldu[a],
                                      ublk ? ublk[a] : U + uoff[a], vblk ? vblk[a] : V + voff[a],
                                      keff + a, tie_last, scan, sU[w], sX[w], sBox[w], sVn[w], items[a], tol);
}

}  // namespace

// Block class of the ACA launches: 0 = huge (an 8-wavefront workgroup, launched
// beside the rest), 1 = large (a 4-wavefront workgroup), 2 = one wavefront.
static int aca_class(const Engine& eng, int a)
{
    const int4 b = eng.h_blk[eng.ndense + a];
    const int64_t e = (int64_t)(b.y - b.x + 1) + (b.w - b.z + 1);
    return e > eng.aca_huge_entries ? 0 : e > eng.aca_big_entries ? 1 : 2;
}

// list[0 .. nhuge) huge blocks (their own launch on the second stream, joined
// back), then nbig large and nsmall other blocks in one launch; factor offsets
// uoff / voff (leading dims eng.ldu / eng.ldv) into U / V
template <class T>
static hmg_status aca_launch(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                             int tie_last, const int* list, int nhuge, int nbig, int nsmall, const int64_t* uoff,
                             const int64_t* voff, T* U, T* V, int* keff, T* const* ublk = nullptr,
                             T* const* vblk = nullptr)
{
    const int scan = eng.aca_row_scan ? 1 : 0;
    const int4* blk = eng.blk.p + eng.ndense;
    const AcaTol tol{eng.aca_tol_mode, eng.aca_tol_c};
    hipStream_t s = eng.stream, s2 = eng.serialize ? eng.stream : eng.stream2;
    const bool fork = nhuge > 0 && s2 != s;
    if (nhuge > 0) {
        if (fork) {
            HMG_HIP(hipEventRecord(eng.ev_fork, s));
            HMG_HIP(hipStreamWaitEvent(s2, eng.ev_fork, 0));
        }
        if (tol.mode != 0) {
            HMG_DISPATCH_KIND(kp.kind, aca_huge_kernel<T, KIND, true><<<nhuge, kAcaHuge * kWave, 0, s2>>>(
                                           c1, c2, kp, blk, eng.blocks.p + eng.ndense, list, eng.kb.p, uoff,
                                           eng.ldu.p, voff, eng.ldv.p, U, V, ublk, vblk, keff, tie_last,
                                           scan, tol));
        } else {
            HMG_DISPATCH_KIND(kp.kind, aca_huge_kernel<T, KIND, false><<<nhuge, kAcaHuge * kWave, 0, s2>>>(
                                           c1, c2, kp, blk, eng.blocks.p + eng.ndense, list, eng.kb.p, uoff,
                                           eng.ldu.p, voff, eng.ldv.p, U, V, ublk, vblk, keff, tie_last,
                                           scan, tol));
        }
        HMG_HIP(hipGetLastError());
        if (fork) HMG_HIP(hipEventRecord(eng.ev_join, s2));
        list += nhuge;
    }
    const unsigned grid = (unsigned)(nbig + (nsmall + kAcaTeam - 1) / kAcaTeam);
    if (grid == 0) {
        if (fork) HMG_HIP(hipStreamWaitEvent(s, eng.ev_join, 0));
        return HMG_OK;
    }
    if (tol.mode != 0) {
        HMG_DISPATCH_KIND(kp.kind, aca_kernel<T, KIND, true><<<grid, kAcaTeam * kWave, eng.aca_lds_pad, s>>>(
                                       c1, c2, kp, blk, eng.blocks.p + eng.ndense, list, nbig, nsmall,
                                       eng.kb.p, uoff, eng.ldu.p, voff, eng.ldv.p, U, V, ublk, vblk, keff, tie_last, scan,
                                       tol));
    } else {
        HMG_DISPATCH_KIND(kp.kind, aca_kernel<T, KIND, false><<<grid, kAcaTeam * kWave, eng.aca_lds_pad, s>>>(
                                       c1, c2, kp, blk, eng.blocks.p + eng.ndense, list, nbig, nsmall,
                                       eng.kb.p, uoff, eng.ldu.p, voff, eng.ldv.p, U, V, ublk, vblk, keff, tie_last, scan,
                                       tol));
    }
    HMG_HIP(hipGetLastError());
    if (fork) HMG_HIP(hipStreamWaitEvent(s, eng.ev_join, 0));
    return HMG_OK;
}

hmg_status aca_build(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp, int k,
                     int tie_last, double* U, double* V, int* keff, double* const* ublk, double* const* vblk)
{
    (void)k;
    const int na = eng.naca;
    if (na == 0) return HMG_OK;
    hipStream_t s = eng.stream;
    // huge blocks, then large blocks (a workgroup each), then the rest (a wavefront each)
    std::vector<int>& order = eng.h_aca_list;
    order.clear();
    order.reserve(na);
    int cnt[3] = {0, 0, 0};
    for (int cls = 0; cls < 3; cls++) {
        const int before = (int)order.size();
        for (int a = 0; a < na; a++)
            if (aca_class(eng, a) == cls) order.push_back(a);
        cnt[cls] = (int)order.size() - before;
    }
    HMG_TRY(eng.aca_list.alloc(na));
    HMG_TRY(upload_bytes(eng.aca_list.p, order.data(), sizeof(int) * na, s));
    eng.prof.begin(PROF_ACA, s);
    HMG_TRY(aca_launch<double>(eng, c1, c2, kp, tie_last, eng.aca_list.p, cnt[0], cnt[1], cnt[2], eng.uoff.p,
                               eng.voff.p, U, V, keff, ublk, vblk));
    eng.prof.end(PROF_ACA, s);
    return HMG_OK;
}

namespace {

constexpr int kConvThreads = 256;

// BF16 hi + lo of an FP64 value (the value rounded to FP32 first)
__device__ __forceinline__ void split_bf16(double v, __bf16& hi, __bf16& lo)
{
    const float f = (float)v;
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

// One workgroup per block of list[]: its FP64 factors in the scratch (U_b
// column-major m1 x kb at suoff, V_b row-major m2 x kb at svoff) into the panels.
// Rank columns at or past the block's effective rank are written as zeros (the
// ACA leaves them unwritten).
__global__ void __launch_bounds__(kConvThreads)
mx_panel_kernel(const int* __restrict__ list, const int4* __restrict__ blk, const int* __restrict__ kb,
                const int* __restrict__ keff, const int64_t* __restrict__ suoff, const int64_t* __restrict__ svoff,
                const MxPanel* __restrict__ mxp, const double* __restrict__ S, __bf16* __restrict__ Up,
                __bf16* __restrict__ Vp)
{
    const int a = list[blockIdx.x];
    const int4 b = blk[a];
    const int m1 = b.y - b.x + 1, m2 = b.w - b.z + 1;
    const int k = kb[a], ke = keff[a];
    const MxPanel mp = mxp[a];
    const double* __restrict__ Su = S + suoff[a];
    const double* __restrict__ Sv = S + svoff[a];
    const int64_t n_u = (int64_t)m1 * k;
    for (int64_t q = threadIdx.x; q < n_u; q += kConvThreads) {
        const int i = (int)(q % m1), l = (int)(q / m1);
        const double v = l < ke ? Su[q] : 0.0;
        const int kk = mp.koff + l;
        const int64_t at = mp.ubase + (int64_t)(kk >> 3) * 16 * mp.m1p + (int64_t)i * 8 + (kk & 7);
        __bf16 hi, lo;
        split_bf16(v, hi, lo);
        Up[at] = hi;
        Up[at + 8 * (int64_t)mp.m1p] = lo;
    }
    const int64_t n_v = (int64_t)m2 * k;
    for (int64_t q = threadIdx.x; q < n_v; q += kConvThreads) {
        const int j = (int)(q / k), l = (int)(q % k);
        const double v = l < ke ? Sv[q] : 0.0;
        const int p = b.z + j;
        const int64_t at = mp.vbase + (int64_t)((p >> 3) - mp.glo) * 16 * mp.kp + (int64_t)(mp.pos + l) * 8 + (p & 7);
        __bf16 hi, lo;
        split_bf16(v, hi, lo);
        Vp[at] = hi;
        Vp[at + 8 * (int64_t)mp.kp] = lo;
    }
}

// FP64 elements of one block's factors in the scratch
int64_t block_scratch(const Engine& eng, int a)
{
    const int4 b = eng.h_blk[eng.ndense + a];
    return (int64_t)((b.y - b.x + 1) + (b.w - b.z + 1)) * eng.h_kb[a];
}

}  // namespace

int64_t aca_mixed_scratch(const Engine& eng)
{
    int64_t total = 0;
    for (int a = 0; a < eng.naca; a++) total += block_scratch(eng, a);
    // 1/parts of the factors at a time, at least 2^27 elements (1 GiB): every batch
    // mixes large and small blocks, so more batches cost little GPU time, while a
    // large fresh allocation costs seconds (mapping tens of GB on MI355X)
    const int64_t cap = std::max<int64_t>(total / std::max(eng.aca_mixed_parts, 1), (int64_t)1 << 27);
    return std::min(total, cap);
}

hmg_status aca_to_panels(Engine& eng, const int* list, int n, const int64_t* suoff, const int64_t* svoff,
                         const double* S, __bf16* Up, __bf16* Vp, hipStream_t s)
{
    if (n <= 0) return HMG_OK;
    mx_panel_kernel<<<n, kConvThreads, 0, s>>>(list, eng.blk.p + eng.ndense, eng.kb.p, eng.keff.p, suoff, svoff,
                                              eng.mxp.p, S, Up, Vp);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status aca_build_mixed(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp, int tie_last,
                           __bf16* Up, __bf16* Vp)
{
    const int na = eng.naca;
    if (na == 0) return HMG_OK;
    hipStream_t s = eng.stream;
    const int64_t cap = std::max<int64_t>(aca_mixed_scratch(eng), 1);
    // batches filled in order of decreasing block size: a batch lasts as long as its
    // longest block (the rank loop and rejected-row search of the largest blocks),
    // so the largest blocks run together in the first batches instead of one of
    // them ending every batch
    int64_t largest = 0;
    for (int a = 0; a < na; a++) largest = std::max(largest, block_scratch(eng, a));
    const int64_t scratch = std::max(cap, largest);
    std::vector<int> by_size(na);
    for (int a = 0; a < na; a++) by_size[a] = a;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int x, int y) { return block_scratch(eng, x) > block_scratch(eng, y); });
    std::vector<int64_t> suoff(na), svoff(na);
    std::vector<int> ldu(na), ldv(na);
    std::vector<std::vector<int>> members(1);
    int64_t used = 0;
    for (const int a : by_size) {
        const int4 b = eng.h_blk[eng.ndense + a];
        const int m1 = b.y - b.x + 1, m2 = b.w - b.z + 1, k = eng.h_kb[a];
        const int64_t need = (int64_t)(m1 + m2) * k;
        if (used + need > scratch && !members.back().empty()) {
            members.emplace_back();
            used = 0;
        }
        members.back().push_back(a);
        suoff[a] = used;
        svoff[a] = used + (int64_t)m1 * k;
        ldu[a] = m1;
        ldv[a] = std::max(k, 1);
        used += need;
    }
    const int nb = (int)members.size();
    DevBuf<double> S;
    HMG_TRY(S.alloc((size_t)scratch));
    DevBuf<int64_t> d_suoff, d_svoff;
    DevBuf<int> d_list;
    HMG_TRY(upload(d_suoff, suoff, s));
    HMG_TRY(upload(d_svoff, svoff, s));
    HMG_TRY(upload(eng.ldu, ldu, s));
    HMG_TRY(upload(eng.ldv, ldv, s));
    // per batch: its huge, then large blocks, then the rest (the ACA launches' order)
    std::vector<int> list;
    list.reserve(na);
    std::vector<int> bptr(nb + 1, 0), bhuge(nb, 0), bbig(nb, 0);
    for (int q = 0; q < nb; q++) {
        bptr[q] = (int)list.size();
        for (int cls = 0; cls < 3; cls++) {
            const int before = (int)list.size();
            for (const int a : members[q])
                if (aca_class(eng, a) == cls) list.push_back(a);
            if (cls == 0) bhuge[q] = (int)list.size() - before;
            if (cls == 1) bbig[q] = (int)list.size() - before;
        }
    }
    bptr[nb] = (int)list.size();
    HMG_TRY(upload(d_list, list, s));
    eng.prof.begin(PROF_ACA, s);
    for (int q = 0; q < nb; q++) {
        const int n = bptr[q + 1] - bptr[q];
        if (n == 0) continue;
        HMG_TRY(aca_launch<double>(eng, c1, c2, kp, tie_last, d_list.p + bptr[q], bhuge[q], bbig[q],
                                   n - bhuge[q] - bbig[q], d_suoff.p, d_svoff.p, S.p, S.p, eng.keff.p));
        mx_panel_kernel<<<n, kConvThreads, 0, s>>>(d_list.p + bptr[q], eng.blk.p + eng.ndense, eng.kb.p,
                                                  eng.keff.p, d_suoff.p, d_svoff.p, eng.mxp.p, S.p, Up, Vp);
        HMG_HIP(hipGetLastError());
    }
    eng.prof.end(PROF_ACA, s);
    HMG_HIP(hipStreamSynchronize(s));
    return HMG_OK;
}

}  // namespace hmg
