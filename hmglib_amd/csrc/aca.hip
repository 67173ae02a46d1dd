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
#include "hmg_internal.h"
#include "kernel_eval.h"

namespace hmg {

namespace {

constexpr int kAcaThreads = 256;
constexpr int kAcaWaves = kAcaThreads / kWave;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Factor layout: U_b[i, l] at U + uoff[a] + l * ldu[a] + i (column-major, ldu >= m1:
// the row cluster's padded height); V_b is stored transposed, V_b[j, l] at
// V + voff[a] + j * ldv[a] + l, i.e. row j of V_b is contiguous (ldv >= total rank
// of the column cluster's blocks).  T is the storage type of the factors (double,
// or float for mixed precision): every residual, norm and pivot division is
// computed in FP64 from the stored values and rounded to T once when stored.
template <class T>
__global__ void __launch_bounds__(kAcaThreads)
aca_kernel(Coords c1, Coords c2, KernelParams kp, const int4* __restrict__ blk, int naca,
           const int* __restrict__ kb, const int64_t* __restrict__ uoff, const int* __restrict__ ldu,
           const int64_t* __restrict__ voff, const int* __restrict__ ldv, T* __restrict__ U,
           T* __restrict__ V, int* __restrict__ keff, int tie_last)
{
    const int lane = threadIdx.x % kWave;
    const int a = blockIdx.x * kAcaWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (a >= naca) return;
    const int4 b = blk[a];
    const int l1 = b.x, m1 = b.y - b.x + 1;
    const int l2 = b.z, m2 = b.w - b.z + 1;
    const int k = kb[a];
    const int64_t ld = ldv[a];
    const int64_t lu = ldu[a];
    T* __restrict__ Ub = U + uoff[a];
    T* __restrict__ Vb = V + voff[a];
    int i_r = -1;
    int ke = 0;
    for (int r = 0; r < k; r++) {
        bool stop = false;
        // residual-row search (compute_current_batched_v_r, :2068-2132)
        for (;;) {
            if (i_r >= m1 - 1) {
                stop = true;
                break;
            }
            i_r++;
            double nrm = 0.0;
            for (int j = lane; j < m2; j += kWave) {
                double val = kernel_entry(kp, c1, l1 + i_r, c2, l2 + j);
                const T* __restrict__ vrow = Vb + (int64_t)j * ld;
                for (int l = 0; l < r; l++)
                    val = __fma_rn(-(double)Ub[l * lu + i_r], (double)vrow[l], val);
                Vb[(int64_t)j * ld + r] = (T)val;
                nrm = __fma_rn(val, val, nrm);
            }
            nrm = wave_sum(nrm);
            if (sqrt(nrm) >= 1.0e-13) break;
        }
        if (stop) {
            // the rank stops here: column r of V (last rejected row) is cleared, column r
            // of U was never written; both stay zero in the product
            for (int j = lane; j < m2; j += kWave) Vb[(int64_t)j * ld + r] = (T)0;
            break;
        }
        // column pivot (tuple_absolute_maximum, :991-1018)
        double best = -1.0, bval = 0.0;
        int bj = -1;
        for (int j = lane; j < m2; j += kWave) {
            const double x = (double)Vb[(int64_t)j * ld + r];
            const double ax = fabs(x);
            if (tie_last ? (ax >= best) : (ax > best)) {
                best = ax;
                bj = j;
                bval = x;
            }
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, kWave);
            const int oj = __shfl_xor(bj, off, kWave);
            const double ov = __shfl_xor(bval, off, kWave);
            const bool take = (ob > best) || (ob == best && (tie_last ? (oj > bj) : (oj < bj && oj >= 0)));
            if (take) {
                best = ob;
                bj = oj;
                bval = ov;
            }
        }
        const int jr = bj;
        const double p = bval;
        for (int j = lane; j < m2; j += kWave)
            Vb[(int64_t)j * ld + r] = (T)((double)Vb[(int64_t)j * ld + r] / p);
        __threadfence_block();
        // residual column (batched_fill_kernel_vector_and_scaled_substraction_for_u_r)
        const T* __restrict__ vpiv = Vb + (int64_t)jr * ld;
        T* __restrict__ u = Ub + r * lu;
        for (int i = lane; i < m1; i += kWave) {
            double val = kernel_entry(kp, c1, l1 + i, c2, l2 + jr);
            for (int l = 0; l < r; l++) val = __fma_rn(-(double)vpiv[l], (double)Ub[l * lu + i], val);
            u[i] = (T)val;
        }
        __threadfence_block();
        ke = r + 1;
    }
    if (lane == 0) keff[a] = ke;
}

}  // namespace

template <class T>
hmg_status aca_build(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp, int k,
                     int tie_last, T* U, T* V, int* keff)
{
    (void)k;
    const int na = eng.naca;
    if (na == 0) return HMG_OK;
    hipStream_t s = eng.stream;
    const unsigned grid = (unsigned)((na + kAcaWaves - 1) / kAcaWaves);
    eng.prof.begin(PROF_ACA, s);
    aca_kernel<T><<<grid, kAcaThreads, 0, s>>>(c1, c2, kp, eng.blk.p + eng.ndense, na, eng.kb.p,
                                            eng.uoff.p, eng.ldu.p, eng.voff.p, eng.ldv.p, U, V, keff, tie_last);
    HMG_HIP(hipGetLastError());
    eng.prof.end(PROF_ACA, s);
    return HMG_OK;
}

template hmg_status aca_build<double>(Engine&, const Coords&, const Coords&, const KernelParams&, int,
                                      int, double*, double*, int*);
template hmg_status aca_build<float>(Engine&, const Coords&, const Coords&, const KernelParams&, int,
                                     int, float*, float*, int*);

}  // namespace hmg
