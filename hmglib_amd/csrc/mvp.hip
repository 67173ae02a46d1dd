// mvp.hip -- H-matrix-vector products, near-field fill, dense checker products.
//
// y = P1^T ( sum_dense A_b x~_J + sum_aca U_b (V_b^T x~_J) ),  x~ = P2 x
// (reference: hmglib.cu:271-286, linear_algebra.cu:4292-4379, 1429-1591,
// 2817-3137).  Instead of batched GEMVs and FP64 CAS atomics into y
// (linear_algebra.cu:1152-1181) the engine uses row ownership:
//   phase 1  one wavefront per ACA block: t_b = V_b^T x~_J  (stream V once)
//   phase 2  one wavefront per tile of 64 Morton-ordered rows: every lane owns a
//            row and sums, in block order, the dense rows (stored A_b, or
//            evaluated on the fly) and the low-rank rows U_b[i,:] t_b, then writes
//            y[order1[i]] once.
// Every row is summed in a fixed order, so results are deterministic and do not
// depend on what else runs on the device.
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

// phase 1: T[toff[a] + l] = sum_j V_b[j + l*m2] * xs[l2 + j]
__global__ void __launch_bounds__(kThreads)
lr_vtx_kernel(const int4* __restrict__ blk, int naca, const int* __restrict__ keff,
              const int64_t* __restrict__ voff, const int64_t* __restrict__ toff,
              const double* __restrict__ V, const double* __restrict__ xs, double* __restrict__ T)
{
    const int lane = threadIdx.x % kWave;
    const int a = blockIdx.x * kWaves + threadIdx.x / kWave;
    if (a >= naca) return;
    const int4 b = blk[a];
    const int l2 = b.z, m2 = b.w - b.z + 1;
    const int ke = keff[a];
    const double* __restrict__ Vb = V + voff[a];
    const double* __restrict__ xb = xs + l2;
    for (int l = 0; l < ke; l++) {
        const double* __restrict__ col = Vb + (int64_t)l * m2;
        double s0 = 0.0, s1 = 0.0;
        int j = lane;
        for (; j + kWave < m2; j += 2 * kWave) {
            s0 = __fma_rn(col[j], xb[j], s0);
            s1 = __fma_rn(col[j + kWave], xb[j + kWave], s1);
        }
        if (j < m2) s0 = __fma_rn(col[j], xb[j], s0);
        const double s = wave_sum(s0 + s1);
        if (lane == 0) T[toff[a] + l] = s;
    }
}

// phase 2: one wavefront per 64-row tile
__global__ void __launch_bounds__(kThreads)
row_kernel(Coords c1, Coords c2, KernelParams kp, int ntiles, int n1, int ndense,
           const int* __restrict__ tile_ptr, const int* __restrict__ tile_blk,
           const int4* __restrict__ blk, const double* __restrict__ xs,
           const int64_t* __restrict__ aoff, const double* __restrict__ A,
           const int* __restrict__ keff, const int64_t* __restrict__ uoff,
           const int64_t* __restrict__ toff, const double* __restrict__ U,
           const double* __restrict__ T, const uint64_t* __restrict__ order1, double* __restrict__ y)
{
    const int lane = threadIdx.x % kWave;
    const int tile = blockIdx.x * kWaves + threadIdx.x / kWave;
    if (tile >= ntiles) return;
    const int row = tile * kTileRows + lane;
    double acc = 0.0;
    const int p0 = tile_ptr[tile], p1 = tile_ptr[tile + 1];
    for (int p = p0; p < p1; p++) {
        const int bidx = tile_blk[p];
        const int4 b = blk[bidx];
        const int l1 = b.x, u1 = b.y, l2 = b.z;
        const int m1 = u1 - l1 + 1, m2 = b.w - b.z + 1;
        const bool active = row >= l1 && row <= u1;
        if (bidx < ndense) {
            if (A) {
                if (active) {
                    const double* __restrict__ Ab = A + aoff[bidx] + (row - l1);
                    const double* __restrict__ xb = xs + l2;
                    double s0 = 0.0, s1 = 0.0;
                    int j = 0;
                    for (; j + 1 < m2; j += 2) {
                        s0 = __fma_rn(Ab[(int64_t)j * m1], xb[j], s0);
                        s1 = __fma_rn(Ab[(int64_t)(j + 1) * m1], xb[j + 1], s1);
                    }
                    if (j < m2) s0 = __fma_rn(Ab[(int64_t)j * m1], xb[j], s0);
                    acc += s0 + s1;
                }
            } else if (active) {
                double s = 0.0;
                for (int j = 0; j < m2; j++)
                    s = __fma_rn(kernel_entry(kp, c1, row, c2, l2 + j), xs[l2 + j], s);
                acc += s;
            }
        } else if (active) {
            const int a = bidx - ndense;
            const int ke = keff[a];
            const double* __restrict__ Ub = U + uoff[a] + (row - l1);
            const double* __restrict__ t = T + toff[a];
            double s = 0.0;
            for (int l = 0; l < ke; l++) s = __fma_rn(Ub[(int64_t)l * m1], t[l], s);
            acc += s;
        }
    }
    if (row < n1) y[order1[row]] = acc;
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

__global__ void inv_kernel(const uint64_t* __restrict__ order, int64_t* __restrict__ inv, int64_t n)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[order[s]] = s;
}

}  // namespace

hmg_status build_schedule(Engine& eng)
{
    const int n1 = eng.n[0];
    const int nt = (n1 + kTileRows - 1) / kTileRows;
    eng.ntiles = nt;
    std::vector<int> cnt(nt + 1, 0);
    for (int b = 0; b < eng.nblocks; b++) {
        const int4 bl = eng.h_blk[b];
        for (int t = bl.x / kTileRows; t <= bl.y / kTileRows; t++) cnt[t + 1]++;
    }
    for (int t = 0; t < nt; t++) cnt[t + 1] += cnt[t];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    std::vector<int> lst((size_t)cnt[nt]);
    for (int b = 0; b < eng.nblocks; b++) {
        const int4 bl = eng.h_blk[b];
        for (int t = bl.x / kTileRows; t <= bl.y / kTileRows; t++) lst[fill[t]++] = b;
    }
    hipStream_t s = eng.stream;
    HMG_TRY(eng.tile_ptr.alloc(nt + 1));
    HMG_TRY(eng.tile_blk.alloc(std::max<size_t>(lst.size(), 1)));
    HMG_HIP(hipMemcpyAsync(eng.tile_ptr.p, cnt.data(), sizeof(int) * (nt + 1), hipMemcpyHostToDevice, s));
    if (!lst.empty())
        HMG_HIP(hipMemcpyAsync(eng.tile_blk.p, lst.data(), sizeof(int) * lst.size(),
                               hipMemcpyHostToDevice, s));
    HMG_HIP(hipStreamSynchronize(s));
    return HMG_OK;
}

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
    return HMG_OK;
}

hmg_status hmvp(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                const uint64_t* order1, const uint64_t* order2, const double* x, double* y, int nrhs,
                const double* U, const double* V, const int* keff, bool dense_stored)
{
    hipStream_t s = eng.stream;
    const int n1 = eng.n[0], n2 = eng.n[1];
    const int na = eng.naca;
    // T offsets (per ACA block, k_b slots)
    if ((int)eng.h_toff.size() != na + 1) {
        eng.h_toff.assign(na + 1, 0);
        for (int a = 0; a < na; a++) eng.h_toff[a + 1] = eng.h_toff[a] + eng.h_kb[a];
        HMG_TRY(eng.toff.alloc(na + 1));
        HMG_HIP(hipMemcpyAsync(eng.toff.p, eng.h_toff.data(), sizeof(int64_t) * (na + 1),
                               hipMemcpyHostToDevice, s));
    }
    HMG_TRY(eng.xs.alloc(std::max(n2, 1)));
    HMG_TRY(eng.T.alloc(std::max<int64_t>(eng.h_toff[na], 1)));
    for (int c = 0; c < nrhs; c++) {
        const double* xc = x + (int64_t)c * n2;
        double* yc = y + (int64_t)c * n1;
        eng.prof.begin(PROF_GATHER, s);
        gather_perm_kernel<<<grid_for(n2, kThreads), kThreads, 0, s>>>(xc, order2, eng.xs.p, n2);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_GATHER, s);
        if (na > 0) {
            eng.prof.begin(PROF_LR_VTX, s);
            lr_vtx_kernel<<<grid_for(na, kWaves), kThreads, 0, s>>>(
                eng.blk.p + eng.ndense, na, keff, eng.voff.p, eng.toff.p, V, eng.xs.p, eng.T.p);
            HMG_HIP(hipGetLastError());
            eng.prof.end(PROF_LR_VTX, s);
        }
        eng.prof.begin(PROF_ROW, s);
        row_kernel<<<grid_for(eng.ntiles, kWaves), kThreads, 0, s>>>(
            c1, c2, kp, eng.ntiles, n1, eng.ndense, eng.tile_ptr.p, eng.tile_blk.p, eng.blk.p,
            eng.xs.p, dense_stored ? eng.aoff.p : nullptr, dense_stored ? eng.A.p : nullptr, keff,
            eng.uoff.p, eng.toff.p, U, eng.T.p, order1, yc);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_ROW, s);
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
