// mvp.hip -- H-matrix-vector products, near-field fill, dense checker products.
//
// y = P1^T ( sum_dense A_b x~_J + sum_aca U_b (V_b^T x~_J) ),  x~ = P2 x
// (reference: hmglib.cu:271-286, linear_algebra.cu:4292-4379, 1429-1591,
// 2817-3137).  Instead of batched GEMVs and FP64 CAS atomics into y
// (linear_algebra.cu:1152-1181) the product runs as:
//   lr_v      one wavefront per (column cluster, 64 columns, <= 1024 points):
//             t = V_D^T x~_D over the column cluster's row-major V, one lane per
//             column, no cross-lane reductions (HBM stream of V); lr_v_finalize
//             sums the point chunks and scatters t into the U order
//   lr_u      one wavefront per 64 rows of a row cluster: y_C = U_C t_C, one long
//             contiguous GEMV per row cluster (HBM stream of U)
//   near      one wavefront per (row cluster, 64 rows, column chunk): the dense
//             near field, either streamed from the stored blocks or evaluated on
//             the fly with the columns staged through LDS (VALU bound); it runs
//             on a second stream beside the two low-rank kernels
//   combine   every row adds its partials in a fixed order, writes y[order1[i]]
// so results are deterministic and independent of what else runs on the device.
#include "hmg_internal.h"
#include "kernel_eval.h"

namespace hmg {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__global__ void gather_perm_kernel(const double* __restrict__ x, const uint64_t* __restrict__ order,
                                   double* __restrict__ xs, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) xs[p] = x[order[p]];
}

// rocPRIM-style wavefront barrier: LDS traffic between lanes of one wave
__device__ __forceinline__ void wave_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Tp[item*64 + l] = sum_{j < nrows} V[vbase + l + j*ld] * xs[j0 + j]
__global__ void __launch_bounds__(kThreads)
lr_v_kernel(const LvItem* __restrict__ items, int nitems, const double* __restrict__ V,
            const double* __restrict__ xs, double* __restrict__ Tp)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const LvItem it = items[q];
    const bool act = lane < it.ncols;
    const double* __restrict__ v = V + it.vbase + (act ? lane : 0);
    const double* __restrict__ xb = xs + it.j0;
    const int64_t ld = it.ld;
    const int n = it.nrows;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {
        const double v0 = v[(j + 0) * ld], v1 = v[(j + 1) * ld], v2 = v[(j + 2) * ld],
                     v3 = v[(j + 3) * ld], v4 = v[(j + 4) * ld], v5 = v[(j + 5) * ld],
                     v6 = v[(j + 6) * ld], v7 = v[(j + 7) * ld];
        s0 = __fma_rn(v0, xb[j + 0], s0);
        s1 = __fma_rn(v1, xb[j + 1], s1);
        s2 = __fma_rn(v2, xb[j + 2], s2);
        s3 = __fma_rn(v3, xb[j + 3], s3);
        s0 = __fma_rn(v4, xb[j + 4], s0);
        s1 = __fma_rn(v5, xb[j + 5], s1);
        s2 = __fma_rn(v6, xb[j + 6], s2);
        s3 = __fma_rn(v7, xb[j + 7], s3);
    }
    for (; j < n; j++) s0 = __fma_rn(v[j * ld], xb[j], s0);
    Tp[(int64_t)q * kTileRows + lane] = (s0 + s1) + (s2 + s3);
}

// T[tmap[tcol0 + l]] = sum over the slice's point chunks, in chunk order
__global__ void __launch_bounds__(kThreads)
lr_v_finalize_kernel(const LvSlice* __restrict__ slices, int nslices, const double* __restrict__ Tp,
                     const int* __restrict__ tmap, double* __restrict__ T)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nslices) return;
    const LvSlice sl = slices[q];
    if (lane >= sl.ncols) return;
    double s = 0.0;
    for (int c = 0; c < sl.nchunks; c++) s += Tp[(int64_t)(sl.item0 + c) * kTileRows + lane];
    T[tmap[sl.tcol0 + lane]] = s;
}

// y_C rows [row0, row0+nrows) = U_C[rows, :] t_C  ->  P[item * 64 + lane]
__global__ void __launch_bounds__(kThreads)
lr_u_kernel(const LrItem* __restrict__ items, int nitems, const double* __restrict__ U,
            const double* __restrict__ T, double* __restrict__ P)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const LrItem it = items[q];
    const bool act = lane < it.nrows;
    const double* __restrict__ u = U + it.uoff + (act ? lane : 0);
    const double* __restrict__ t = T + it.toff;
    const int64_t m1 = it.ld;
    const int K = it.K;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int l = 0;
    for (; l + 8 <= K; l += 8) {
        const double u0 = u[(l + 0) * m1], u1 = u[(l + 1) * m1], u2 = u[(l + 2) * m1],
                     u3 = u[(l + 3) * m1], u4 = u[(l + 4) * m1], u5 = u[(l + 5) * m1],
                     u6 = u[(l + 6) * m1], u7 = u[(l + 7) * m1];
        s0 = __fma_rn(u0, t[l + 0], s0);
        s1 = __fma_rn(u1, t[l + 1], s1);
        s2 = __fma_rn(u2, t[l + 2], s2);
        s3 = __fma_rn(u3, t[l + 3], s3);
        s0 = __fma_rn(u4, t[l + 4], s0);
        s1 = __fma_rn(u5, t[l + 5], s1);
        s2 = __fma_rn(u6, t[l + 6], s2);
        s3 = __fma_rn(u7, t[l + 7], s3);
    }
    for (; l < K; l++) s0 = __fma_rn(u[l * m1], t[l], s0);
    if (act) P[(int64_t)q * kTileRows + lane] = (s0 + s1) + (s2 + s3);
}

template <int KIND>
__device__ __forceinline__ double kernel_value(const KernelParams& kp, double r2)
{
    return kernel_of_r2_prod<KIND>(kp, r2);
}

// The near field of one work item (up to 128 rows of a row cluster, one or two
// rows per lane) against the columns of its dense blocks, which are staged 64 at a
// time through this wave's LDS slice.  NR rows per lane give NR independent
// evaluation chains per staged column.  DIM = 0 means "dim at run time".
constexpr int kNearRows = 2 * kWave;

// lane i receives lane (i + 1) mod 64 of v (DPP wave_rol:1)
__device__ __forceinline__ double lane_rotate_down(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x134, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x134, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <int KIND, int DIM>
__device__ __forceinline__ double entry_value(const KernelParams& kp, const double* xi, const double* xj)
{
    const int dim = DIM > 0 ? DIM : kp.dim;
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
        if (d < dim) {
            const double t = xi[d] - xj[d];
            r2 = __fma_rn(t, t, r2);
        }
    }
    return kernel_value<KIND>(kp, r2);
}

template <int KIND, int DIM, bool REG, int NR>
__device__ __forceinline__ void near_fly_item(const Coords& c1, const Coords& c2, const KernelParams& kp,
                                              const NearItem& it, const NearRef* __restrict__ refs,
                                              const double* __restrict__ xs,
                                              double (*sh)[kWave], int lane, double* __restrict__ out,
                                              double* __restrict__ P)
{
    constexpr int kS = MAX_POINT_DIMENSION + 1;
    const int dim = DIM > 0 ? DIM : kp.dim;
    bool act[NR];
    int row[NR];
    double xi[NR][MAX_POINT_DIMENSION];
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        act[r] = lane + r * kWave < it.nrows;
        row[r] = it.row0 + (act[r] ? lane + r * kWave : 0);
        acc[r] = 0.0;
#pragma unroll
        for (int d = 0; d < MAX_POINT_DIMENSION; d++) xi[r][d] = d < dim ? c1.x[d][row[r]] : 0.0;
    }
    // symmetric refs: one column stream over all their columns, 64 columns staged
    // at a time (lane L stages column L).  Step s pairs lane L with staged column
    // (L + s) mod 64; the column sums travel one lane down per step, so after 64
    // steps lane L holds the sum of column L over this item's rows.
    if (it.ncols_sym > 0) {
        double xr[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) xr[r] = act[r] ? xs[row[r]] : 0.0;
        int rc = it.ref_sym, rc_start = 0;  // first ref not yet fully staged, its stream offset
        for (int t0 = 0; t0 < it.ncols_sym; t0 += kWave) {
            const int pos = t0 + lane;
            int j = -1;
            int64_t slot = 0;
            int r = rc, rs = rc_start;
            while (r < it.ref1 && rs < t0 + kWave) {
                const NearRef f = refs[r];
                const int re = rs + f.m2;
                if (pos >= rs && pos < re) {
                    j = f.l2 + (pos - rs);
                    slot = f.cbase + (pos - rs);
                }
                if (re <= t0 + kWave) {
                    rc = r + 1;
                    rc_start = re;
                }
                rs = re;
                r++;
            }
            // past the stream's end: any valid point with x = 0 (contributes nothing)
            const int jj = j >= 0 ? j : 0;
            wave_barrier();
#pragma unroll
            for (int d = 0; d < MAX_POINT_DIMENSION; d++)
                if (d < dim) sh[d][lane] = c2.x[d][jj];
            sh[kS - 1][lane] = j >= 0 ? xs[jj] : 0.0;
            wave_barrier();
            double colacc = 0.0;
            for (int s = 0; s < kWave; s++) {
                const int c = (lane + s) & (kWave - 1);
                double xj[MAX_POINT_DIMENSION];
#pragma unroll
                for (int d = 0; d < MAX_POINT_DIMENSION; d++) xj[d] = d < dim ? sh[d][c] : 0.0;
                const double xv = sh[kS - 1][c];
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    const double kv = entry_value<KIND, DIM>(kp, xi[q], xj);
                    acc[q] = __fma_rn(kv, xv, acc[q]);
                    colacc = __fma_rn(kv, xr[q], colacc);
                }
                colacc = lane_rotate_down(colacc);
            }
            if (j >= 0) P[slot] = colacc;
        }
    }
    for (int bi = it.ref0; bi < it.ref_sym; bi++) {
        const NearRef f = refs[bi];
        const int l2 = f.l2, m2 = f.m2;
        for (int j0 = 0; j0 < m2; j0 += kWave) {
            const int cnt = min(kWave, m2 - j0);
            wave_barrier();
            if (lane < cnt) {
#pragma unroll
                for (int d = 0; d < MAX_POINT_DIMENSION; d++)
                    if (d < dim) sh[d][lane] = c2.x[d][l2 + j0 + lane];
                sh[kS - 1][lane] = xs[l2 + j0 + lane];
            }
            wave_barrier();
            for (int jj = 0; jj < cnt; jj++) {
                double xj[MAX_POINT_DIMENSION];
#pragma unroll
                for (int d = 0; d < MAX_POINT_DIMENSION; d++) xj[d] = d < dim ? sh[d][jj] : 0.0;
                const double xv = sh[kS - 1][jj];
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    double r2 = 0.0;
#pragma unroll
                    for (int d = 0; d < MAX_POINT_DIMENSION; d++) {
                        if (d < dim) {
                            const double t = xi[r][d] - xj[d];
                            r2 = __fma_rn(t, t, r2);
                        }
                    }
                    double kv = kernel_value<KIND>(kp, r2);
                    if (REG && row[r] == l2 + j0 + jj) kv += kp.reg;
                    acc[r] = __fma_rn(kv, xv, acc[r]);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++)
        if (act[r]) out[lane + r * kWave] = acc[r];
}

// near field evaluated on the fly: row sums of item q at Pnear[q*128 + row offset],
// symmetric column sums at P[ref cbase + column offset]
template <int KIND, int DIM, bool REG>
__global__ void __launch_bounds__(kThreads)
near_fly_kernel(Coords c1, Coords c2, KernelParams kp, const NearItem* __restrict__ items, int nitems,
                const NearRef* __restrict__ refs, const double* __restrict__ xs,
                double* __restrict__ Pnear, double* __restrict__ P)
{
    __shared__ double sh[kWaves][MAX_POINT_DIMENSION + 1][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = wave_id();
    const int q = blockIdx.x * kWaves + w;
    if (q >= nitems) return;
    const NearItem it = items[q];
    double* out = Pnear + (int64_t)q * kNearRows;
    if (it.nrows > kWave)
        near_fly_item<KIND, DIM, REG, 2>(c1, c2, kp, it, refs, xs, sh[w], lane, out, P);
    else
        near_fly_item<KIND, DIM, REG, 1>(c1, c2, kp, it, refs, xs, sh[w], lane, out, P);
}

// near field from the stored blocks (column-major m1 x m2 per dense block)
__global__ void __launch_bounds__(kThreads)
near_stored_kernel(const int4* __restrict__ items, int nitems, const int* __restrict__ near_blk,
                   const int4* __restrict__ blk, const int64_t* __restrict__ aoff,
                   const double* __restrict__ A, const double* __restrict__ xs,
                   double* __restrict__ P)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const int4 it = items[q];
    for (int half = 0; half * kWave < it.y; half++) {
        const bool act = lane + half * kWave < it.y;
        const int row = it.x + (act ? lane + half * kWave : 0);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int bi = it.z; bi < it.w; bi++) {
            const int bidx = near_blk[bi];
            const int4 b = blk[bidx];
            const int64_t m1 = b.y - b.x + 1;
            const int l2 = b.z, m2 = b.w - b.z + 1;
            const double* __restrict__ a = A + aoff[bidx] + (row - b.x);
            const double* __restrict__ xb = xs + l2;
            int j = 0;
            for (; j + 4 <= m2; j += 4) {
                s0 = __fma_rn(a[(j + 0) * m1], xb[j + 0], s0);
                s1 = __fma_rn(a[(j + 1) * m1], xb[j + 1], s1);
                s2 = __fma_rn(a[(j + 2) * m1], xb[j + 2], s2);
                s3 = __fma_rn(a[(j + 3) * m1], xb[j + 3], s3);
            }
            for (; j < m2; j++) s0 = __fma_rn(a[j * m1], xb[j], s0);
        }
        if (act) P[(int64_t)q * kNearRows + lane + half * kWave] = (s0 + s1) + (s2 + s3);
    }
}

// y[order1[i]] = sum of row i's partials, in list order (a column-slot group
// contributes its count slots of row i in order)
__global__ void combine_kernel(const int* __restrict__ ptr, const int* __restrict__ idx,
                               const ColSlots* __restrict__ groups, const double* __restrict__ P,
                               const uint64_t* __restrict__ order1, double* __restrict__ y, int n1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n1) return;
    double s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; q++) {
        const int v = idx[q];
        if (v >= 0) {
            s += P[v];
        } else {
            const ColSlots g = groups[-v - 1];
            const double* __restrict__ p = P + g.base + (i - g.l2);
            for (int k = 0; k < g.count; k++) s += p[(int64_t)k * g.m2];
        }
    }
    y[order1[i]] = s;
}

// dense near-field fill: one block per dense block, column-major m1 x m2
__global__ void __launch_bounds__(kThreads)
dense_fill_kernel(Coords c1, Coords c2, KernelParams kp, const int4* __restrict__ blk,
                  const int64_t* __restrict__ aoff, double* __restrict__ A)
{
    const int b = blockIdx.x;
    const int4 bl = blk[b];
    const int l1 = bl.x, m1 = bl.y - bl.x + 1, l2 = bl.z, m2 = bl.w - bl.z + 1;
    double* __restrict__ Ab = A + aoff[b];
    const int64_t tot = (int64_t)m1 * m2;
    for (int64_t q = threadIdx.x; q < tot; q += kThreads) {
        const int i = (int)(q % m1), j = (int)(q / m1);
        Ab[q] = kernel_entry(kp, c1, l1 + i, c2, l2 + j);
    }
}

// full dense product (apply_full_mvp), lane per row, all columns
__global__ void __launch_bounds__(kThreads)
full_kernel(Coords c1, Coords c2, KernelParams kp, const double* __restrict__ xs,
            const uint64_t* __restrict__ order1, double* __restrict__ y)
{
    const int row = blockIdx.x * kThreads + threadIdx.x;
    if (row >= c1.n) return;
    double s = 0.0;
    for (int j = 0; j < c2.n; j++) s = __fma_rn(kernel_entry(kp, c1, row, c2, j), xs[j], s);
    y[order1[row]] = s;
}

// exact rows (checker at scale): one wavefront per sampled row
__global__ void __launch_bounds__(kThreads)
rows_kernel(Coords c1, Coords c2, KernelParams kp, const double* __restrict__ xs,
            const int64_t* __restrict__ rows, const int64_t* __restrict__ inv1, int nrows,
            double* __restrict__ yrows)
{
    const int lane = threadIdx.x % kWave;
    const int q = blockIdx.x * kWaves + threadIdx.x / kWave;
    if (q >= nrows) return;
    const int i = (int)inv1[rows[q]];
    double s = 0.0;
    for (int j = lane; j < c2.n; j += kWave) s = __fma_rn(kernel_entry(kp, c1, i, c2, j), xs[j], s);
    s = wave_sum(s);
    if (lane == 0) yrows[q] = s;
}

// gen_gaussian_kernel_rhs (hmglib.cu:403-418) on sorted points, scattered back
__global__ void rhs_kernel(Coords c1, int dim, const uint64_t* __restrict__ order1,
                           double* __restrict__ b)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c1.n) return;
    const double sqpi = sqrt(M_PI);
    double result = 1.0;
    for (int d = 0; d < dim; d++) {
        const double xv = c1.x[d][s];
        result = result * sqpi * (erf(1.0 - xv) - erf(0.0 - xv)) / (2.0);
    }
    b[order1[s]] = result;
}

__global__ void __launch_bounds__(kThreads)
copy_probe_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

__global__ void inv_kernel(const uint64_t* __restrict__ order, int64_t* __restrict__ inv, int64_t n)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[order[s]] = s;
}

}  // namespace

hmg_status dense_fill(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp)
{
    const int nd = eng.ndense;
    eng.h_aoff.assign(nd + 1, 0);
    for (int b = 0; b < nd; b++) {
        const int4 bl = eng.h_blk[b];
        eng.h_aoff[b + 1] = eng.h_aoff[b] + (int64_t)(bl.y - bl.x + 1) * (bl.w - bl.z + 1);
    }
    hipStream_t s = eng.stream;
    HMG_TRY(eng.aoff.alloc(nd + 1));
    HMG_TRY(eng.A.alloc(std::max<int64_t>(eng.h_aoff[nd], 1)));
    HMG_HIP(hipMemcpyAsync(eng.aoff.p, eng.h_aoff.data(), sizeof(int64_t) * (nd + 1),
                           hipMemcpyHostToDevice, s));
    if (nd > 0) {
        eng.prof.begin(PROF_DENSE_FILL, s);
        dense_fill_kernel<<<nd, kThreads, 0, s>>>(c1, c2, kp, eng.blk.p, eng.aoff.p, eng.A.p);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_DENSE_FILL, s);
    }
    HMG_HIP(hipStreamSynchronize(s));
    eng.dense_stored = true;
    eng.epoch++;
    return HMG_OK;
}

namespace {

template <int KIND, int DIM>
void launch_near_fly(bool reg, unsigned grid, hipStream_t s, const Coords& c1, const Coords& c2,
                     const KernelParams& kp, const Engine& eng, const double* xs, double* P)
{
    double* Pnear = P + (int64_t)eng.n_lr_items * kTileRows;
    if (reg)
        near_fly_kernel<KIND, DIM, true><<<grid, kThreads, eng.near_lds_pad, s>>>(
            c1, c2, kp, eng.fly_items.p, eng.n_fly_items, eng.fly_refs.p, xs, Pnear, P);
    else
        near_fly_kernel<KIND, DIM, false><<<grid, kThreads, eng.near_lds_pad, s>>>(
            c1, c2, kp, eng.fly_items.p, eng.n_fly_items, eng.fly_refs.p, xs, Pnear, P);
}

template <int KIND>
void launch_near_fly_dim(bool reg, unsigned grid, hipStream_t s, const Coords& c1, const Coords& c2,
                         const KernelParams& kp, const Engine& eng, const double* xs, double* P)
{
    if (kp.dim == 3)
        launch_near_fly<KIND, 3>(reg, grid, s, c1, c2, kp, eng, xs, P);
    else if (kp.dim == 2)
        launch_near_fly<KIND, 2>(reg, grid, s, c1, c2, kp, eng, xs, P);
    else
        launch_near_fly<KIND, 0>(reg, grid, s, c1, c2, kp, eng, xs, P);
}

}  // namespace

MvpSizes mvp_sizes(const Engine& eng)
{
    MvpSizes z;
    z.xs = std::max(eng.n[1], 1);
    z.Tp = (int64_t)std::max(eng.n_lv_items, 1) * kTileRows;
    z.T = std::max<int64_t>(eng.tsize, 1);
    z.P = std::max<int64_t>(eng.p_slots, 1);
    return z;
}

hmg_status hmvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                const uint64_t* order1, const uint64_t* order2, const double* x, double* y, int nrhs,
                const double* U, const double* V, const int* keff, bool dense_stored)
{
    HMG_TRY(eng.xs.alloc(std::max(eng.n[1], 1)));
    const MvpBufs b{eng.xs.p, eng.Tp.p, eng.T.p, eng.P.p};
    return hmvp_with(eng, c1, c2, kp, order1, order2, x, y, nrhs, U, V, keff, dense_stored, b);
}

hmg_status hmvp_with(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                     const uint64_t* order1, const uint64_t* order2, const double* x, double* y, int nrhs,
                     const double* U, const double* V, const int* keff, bool dense_stored, const MvpBufs& b)
{
    (void)keff;
    hipStream_t s = eng.stream, s2 = eng.serialize ? eng.stream : eng.stream2;
    const int n1 = eng.n[0], n2 = eng.n[1];
    const int na = eng.naca;
    const bool reg = kp.reg != 0.0;
    for (int c = 0; c < nrhs; c++) {
        const double* xc = x + (int64_t)c * n2;
        double* yc = y + (int64_t)c * n1;
        eng.prof.begin(PROF_GATHER, s);
        gather_perm_kernel<<<grid_for(n2, kThreads), kThreads, 0, s>>>(xc, order2, b.xs, n2);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_GATHER, s);
        // near field on the second stream
        HMG_HIP(hipEventRecord(eng.ev_fork, s));
        HMG_HIP(hipStreamWaitEvent(s2, eng.ev_fork, 0));
        const int n_items = dense_stored ? eng.n_near_items : eng.n_fly_items;
        if (n_items > 0) {
            const unsigned g = grid_for(n_items, kWaves);
            eng.prof.begin(PROF_NEAR, s2);
            if (dense_stored) {
                near_stored_kernel<<<g, kThreads, eng.near_lds_pad, s2>>>(
                    eng.near_items.p, eng.n_near_items, eng.near_blk.p, eng.blk.p, eng.aoff.p, eng.A.p,
                    b.xs, b.P + (int64_t)eng.n_lr_items * kTileRows);
            } else {
                switch (kp.kind) {
                case HMG_KERNEL_GAUSSIAN:
                    launch_near_fly_dim<HMG_KERNEL_GAUSSIAN>(reg, g, s2, c1, c2, kp, eng, b.xs, b.P);
                    break;
                case HMG_KERNEL_EXPONENTIAL:
                    launch_near_fly_dim<HMG_KERNEL_EXPONENTIAL>(reg, g, s2, c1, c2, kp, eng, b.xs, b.P);
                    break;
                case HMG_KERNEL_GAUSSIAN_REF:
                    launch_near_fly_dim<HMG_KERNEL_GAUSSIAN_REF>(reg, g, s2, c1, c2, kp, eng, b.xs, b.P);
                    break;
                default:
                    launch_near_fly_dim<HMG_KERNEL_MATERN_REF>(reg, g, s2, c1, c2, kp, eng, b.xs, b.P);
                    break;
                }
            }
            HMG_HIP(hipGetLastError());
            eng.prof.end(PROF_NEAR, s2);
        }
        HMG_HIP(hipEventRecord(eng.ev_join, s2));
        // low-rank part on the main stream
        if (na > 0 && eng.n_lr_items > 0) {
            eng.prof.begin(PROF_LR_VTX, s);
            lr_v_kernel<<<grid_for(eng.n_lv_items, kWaves), kThreads, 0, s>>>(
                eng.lv_items.p, eng.n_lv_items, V, b.xs, b.Tp);
            HMG_HIP(hipGetLastError());
            lr_v_finalize_kernel<<<grid_for(eng.n_lv_slices, kWaves), kThreads, 0, s>>>(
                eng.lv_slices.p, eng.n_lv_slices, b.Tp, eng.tmap.p, b.T);
            HMG_HIP(hipGetLastError());
            eng.prof.end(PROF_LR_VTX, s);
            eng.prof.begin(PROF_LR_U, s);
            lr_u_kernel<<<grid_for(eng.n_lr_items, kWaves), kThreads, 0, s>>>(
                eng.lr_items.p, eng.n_lr_items, U, b.T, b.P);
            HMG_HIP(hipGetLastError());
            eng.prof.end(PROF_LR_U, s);
        }
        HMG_HIP(hipStreamWaitEvent(s, eng.ev_join, 0));
        eng.prof.begin(PROF_COMBINE, s);
        const Combine& cb = dense_stored ? eng.comb_stored : eng.comb_fly;
        combine_kernel<<<grid_for(n1, kThreads), kThreads, 0, s>>>(cb.ptr.p, cb.idx.p, cb.groups.p,
                                                                   b.P, order1, yc, n1);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_COMBINE, s);
    }
    return HMG_OK;
}

hmg_status full_mvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                    const uint64_t* order1, const uint64_t* order2, const double* x, double* y,
                    int nrhs)
{
    hipStream_t s = eng.stream;
    const int n1 = eng.n[0], n2 = eng.n[1];
    HMG_TRY(eng.xs.alloc(std::max(n2, 1)));
    for (int c = 0; c < nrhs; c++) {
        gather_perm_kernel<<<grid_for(n2, kThreads), kThreads, 0, s>>>(x + (int64_t)c * n2, order2,
                                                                       eng.xs.p, n2);
        HMG_HIP(hipGetLastError());
        full_kernel<<<grid_for(n1, kThreads), kThreads, 0, s>>>(c1, c2, kp, eng.xs.p, order1,
                                                                y + (int64_t)c * n1);
        HMG_HIP(hipGetLastError());
    }
    return HMG_OK;
}

hmg_status dense_rows(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order2, const double* x, int nrhs, const int64_t* rows,
                      int nrows, double* yrows, const int64_t* inv1)
{
    hipStream_t s = eng.stream;
    const int n2 = eng.n[1];
    HMG_TRY(eng.xs.alloc(std::max(n2, 1)));
    for (int c = 0; c < nrhs; c++) {
        gather_perm_kernel<<<grid_for(n2, kThreads), kThreads, 0, s>>>(x + (int64_t)c * n2, order2,
                                                                       eng.xs.p, n2);
        HMG_HIP(hipGetLastError());
        rows_kernel<<<grid_for(nrows, kWaves), kThreads, 0, s>>>(c1, c2, kp, eng.xs.p, rows, inv1,
                                                                 nrows, yrows + (int64_t)c * nrows);
        HMG_HIP(hipGetLastError());
    }
    return HMG_OK;
}

hmg_status gaussian_rhs(Engine& eng, const Coords& c1, const uint64_t* order1, double* b)
{
    hipStream_t s = eng.stream;
    rhs_kernel<<<grid_for(c1.n, kThreads), kThreads, 0, s>>>(c1, eng.dim, order1, b);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

hmg_status inverse_perm(const uint64_t* order, int64_t* inv, int64_t n, hipStream_t s)
{
    inv_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(order, inv, n);
    HMG_HIP(hipGetLastError());
    return HMG_OK;
}

}  // namespace hmg

extern "C" hmg_status hmg_copy_probe(const double* src, double* dst, int64_t n)
{
    if (!src || !dst || n < 0) {
        hmg::set_error("invalid copy probe arguments");
        return HMG_ERR_INVALID;
    }
    if (n == 0) return HMG_OK;
    const unsigned grid = (unsigned)std::min<int64_t>((n + hmg::kThreads - 1) / hmg::kThreads, 1 << 16);
    hmg::copy_probe_kernel<<<grid, hmg::kThreads>>>(src, dst, n);
    HMG_HIP(hipGetLastError());
    HMG_HIP(hipDeviceSynchronize());
    return HMG_OK;
}
