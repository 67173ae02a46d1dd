// mixed.hip -- mixed-precision H-matrix-vector products for many right-hand sides.
//
// Same decomposition as the FP64 product (reference: linear_algebra.cu:4292-4379,
// 1429-1591, 2817-3137), with up to kMxRhs = 32 right-hand sides per pass:
//   prep      X (FP64, caller order) -> BF16 hi/lo panels Xp of the sorted points
//             (groups of 8 points: [hi | lo][32 right-hand sides][8 points])
//   lr_vb     T_D = V_D^T X_D per (column cluster, 128 columns, <= 4096 points):
//             v_mfma_f32_32x32x16_bf16 tiles (M = columns, N = right-hand sides,
//             K = points) over the BF16 hi + lo factor panels and Xp (hi*hi +
//             hi*lo + lo*hi); single-chunk slices write T (split into BF16 hi +
//             lo panels) directly, longer ones FP32 partials that lr_vb_finalize
//             adds in chunk order
//   lr_ub     Y_C = U_C T_C per 128 rows of a row cluster, same MFMA scheme
//   near_mx   the near field evaluated on the fly in FP32: each lane evaluates
//             the 8 kernel entries (its row, 8 consecutive columns) that it feeds
//             to v_mfma_f32_32x32x16_bf16 as A operand, split into BF16 hi + lo;
//             Y += A_hi X_hi + A_hi X_lo + A_lo X_hi, accumulated in FP32.  It
//             runs on the second stream beside the low-rank kernels.
//   combine   every row adds its partials in a fixed order (deterministic)
// MFMA lane layouts (gfx950, wave64), for lane l, h = l / 32, c = l % 32:
//   32x32x2 f32:  A[c][h], B[h][c], D[(r/4)*8 + 4h + r%4][c] for r < 16
//   32x32x16 bf16: A[c][8h + e], B[8h + e][c] for e < 8, D as above
#include "hmg_internal.h"

namespace hmg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kMxRowsItem = kMxNearRows; // rows per near-field item (four 32-row tiles)
constexpr int kMxRowTiles = kMxRowsItem / 32;
constexpr int kGroup = 8;        // points per panel group
inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// kernel parameters in the form the FP32 evaluation uses.  The coordinate panels
// hold the points scaled by cscale, chosen so that the kernel is 2^(-s) with
// s = r'^2 (Gaussian kinds: cscale = sqrt(c0 log2 e)) or s = r' (exponential:
// cscale = c0 log2 e), r' the distance of the scaled points: no multiply per
// entry, and the negation is an operand modifier of v_exp_f32.  Matérn stays FP64
// on unscaled points.
struct KernelParamsF {
    int kind;
    int dim;
    float reg;
    float cscale;
    double c1;
};

KernelParamsF make_params_f(const KernelParams& kp)
{
    const double log2e = 1.4426950408889634073599246810019;
    KernelParamsF f{};
    f.kind = kp.kind;
    f.dim = kp.dim;
    f.reg = (float)kp.reg;
    f.c1 = kp.c1;
    if (kp.kind == HMG_KERNEL_GAUSSIAN)
        f.cscale = (float)std::sqrt(kp.c0 * log2e);
    else if (kp.kind == HMG_KERNEL_EXPONENTIAL)
        f.cscale = (float)(kp.c0 * log2e);
    else if (kp.kind == HMG_KERNEL_GAUSSIAN_REF)
        f.cscale = (float)std::sqrt(log2e);  // reference Gaussian exp(-r^2)
    else
        f.cscale = 1.0f;
    return f;
}

template <int KIND>
__device__ __forceinline__ float kval_f32(const KernelParamsF& kp, float r2)
{
    if (KIND == HMG_KERNEL_GAUSSIAN || KIND == HMG_KERNEL_GAUSSIAN_REF) return __builtin_amdgcn_exp2f(-r2);
    if (KIND == HMG_KERNEL_EXPONENTIAL) return __builtin_amdgcn_exp2f(-__builtin_amdgcn_sqrtf(r2));
    const double r = sqrt((double)r2);
    return (float)((yn(1, r + 1.0e-15) * r) / kp.c1);
}

// ---- per-pass inputs --------------------------------------------------------------
// Xt[i*32 + c] = X[i + (c0+c)*n2] rounded to FP32 (0 past nr): the caller's columns
// turned into 128-byte rows, 64 points per workgroup through LDS (reads and writes
// coalesced), so that the permuted gather below reads whole rows
constexpr int kRowTile = 64;
__global__ void __launch_bounds__(kThreads)
x_rows_kernel(const double* __restrict__ X, int n2, int c0, int nr, float* __restrict__ Xt)
{
    __shared__ float tile[kMxRhs][kRowTile + 1];
    const int i0 = blockIdx.x * kRowTile;
#pragma unroll
    for (int k = 0; k < kRowTile * kMxRhs / kThreads; k++) {
        const int e = threadIdx.x + kThreads * k;
        const int c = e / kRowTile, il = e % kRowTile;
        const int i = i0 + il;
        tile[c][il] = (i < n2 && c < nr) ? (float)X[i + (int64_t)(c0 + c) * n2] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRowTile * kMxRhs / kThreads; k++) {
        const int e = threadIdx.x + kThreads * k;
        const int il = e / kMxRhs, c = e % kMxRhs;
        const int i = i0 + il;
        if (i < n2) Xt[(int64_t)i * kMxRhs + c] = tile[c][il];
    }
}

// Xp panels: group g = p/8 holds [hi | lo][32 right-hand sides][8 points] BF16 of
// Xt[order2[p]*32 + c] (0 past n2)
__global__ void __launch_bounds__(kThreads)
prep_x_kernel(const float* __restrict__ Xt, const uint64_t* __restrict__ order2, int n2, int n2p,
              __bf16* __restrict__ Xp)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (int64_t)n2p * kMxRhs) return;
    const int p = (int)(t / kMxRhs), c = (int)(t % kMxRhs);
    float v = 0.f;
    if (p < n2) v = Xt[(int64_t)order2[p] * kMxRhs + c];
    const __bf16 h = (__bf16)v;
    const __bf16 l = (__bf16)(v - (float)h);
    const int64_t gbase = (int64_t)(p / kGroup) * 2 * kMxRhs * kGroup + c * kGroup + (p % kGroup);
    Xp[gbase] = h;
    Xp[gbase + kMxRhs * kGroup] = l;
}

// coordinate panels: cp[(g*dim + d)*8 + e] = scale * x_d[8g + e] as FP32 (0 past n)
__global__ void __launch_bounds__(kThreads)
coord_panel_kernel(Coords c, int dim, int ngroups, double scale, float* __restrict__ cp)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (int64_t)ngroups * kGroup) return;
    const int p = (int)t;
    const int g = p / kGroup, e = p % kGroup;
    for (int d = 0; d < dim; d++)
        cp[((int64_t)g * dim + d) * kGroup + e] = p < c.n ? (float)(c.x[d][p] * scale) : 0.f;
}

// ---- near field ----------------------------------------------------------------------
template <int KIND, int DIM, bool REG>
__device__ __forceinline__ f32x16 near_tile(const KernelParamsF& kp, int dim, const float* xr,
                                            const f32x8* xc, int row, int col0, bf16x8 bh, bf16x8 bl,
                                            f32x16 acc)
{
    constexpr int DM = DIM > 0 ? DIM : MAX_POINT_DIMENSION;
    f32x8 kv;
#pragma unroll
    for (int e = 0; e < kGroup; e++) {
        float r2 = 0.f;
#pragma unroll
        for (int d = 0; d < DM; d++) {
            if (DIM > 0 || d < dim) {
                const float t = xr[d] - xc[d][e];
                r2 = __builtin_fmaf(t, t, r2);
            }
        }
        float v = kval_f32<KIND>(kp, r2);
        if (REG && row == col0 + e) v += kp.reg;
        kv[e] = v;
    }
    // hi = kv rounded to BF16 (packed pairs); its FP32 value is read back from the
    // packed words (a shift or a mask per value), lo = kv - hi rounded to BF16
    const bf16x8 ah = __builtin_convertvector(kv, bf16x8);
    const u32x4 hw = __builtin_bit_cast(u32x4, ah);
    f32x8 hf;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        hf[2 * p] = __builtin_bit_cast(float, hw[p] << 16);
        hf[2 * p + 1] = __builtin_bit_cast(float, hw[p] & 0xffff0000u);
    }
    const bf16x8 al = __builtin_convertvector(kv - hf, bf16x8);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}

// One wavefront per item: up to 128 rows (four 32-row tiles) of one row cluster
// against the item's column runs, 16 columns (two 8-point groups) per step.
// Lane l evaluates row (l % 32) of each tile against the 8 columns of group
// g + l/32, which is exactly its A operand; its B operand is the same group's
// X panel for right-hand side l % 32, shared by the four tiles.  Columns outside
// a run are masked by zeroing their B values.  P[(q*128 + r)*32 + c] = row r's
// partial for rhs c.
template <int KIND, int DIM, bool REG>
__global__ void __launch_bounds__(kThreads)
near_mx_kernel(const float* __restrict__ cp1, const float* __restrict__ cp2, KernelParamsF kp,
               const MxItem* __restrict__ items, int nitems, const MxRef* __restrict__ refs,
               const bf16x8* __restrict__ Xp, float* __restrict__ P)
{
    constexpr int DM = DIM > 0 ? DIM : MAX_POINT_DIMENSION;
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    const MxItem it = items[q];
    const int dim = DIM > 0 ? DIM : kp.dim;
    const int c32 = lane & 31, h = lane >> 5;
    const int ntile = (it.nrows + 31) >> 5;   // wave-uniform
    float xr[kMxRowTiles][DM];
    int row[kMxRowTiles];
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++) {
        const int rr = t * 32 + c32 < it.nrows ? t * 32 + c32 : 0;  // rows past the item: its first row
        row[t] = it.row0 + rr;
#pragma unroll
        for (int d = 0; d < DM; d++)
            xr[t][d] = (DIM > 0 || d < dim) ? cp1[((int64_t)(row[t] >> 3) * dim + d) * kGroup + (row[t] & 7)] : 0.f;
    }
    f32x16 acc[kMxRowTiles];
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][r] = 0.f;
    // steps (ref rf, first group g) two groups at a time; the next step's X panel
    // values load while this step's entries are evaluated (past the last step the
    // loads repeat its groups)
    int rf = it.ref0;
    int g = refs[rf].g0;
    bf16x8 nbh = Xp[((int64_t)(g + h) * 2 + 0) * kMxRhs + c32];
    bf16x8 nbl = Xp[((int64_t)(g + h) * 2 + 1) * kMxRhs + c32];
    for (;;) {
        const MxRef f = refs[rf];
        const int gg = g + h;
        bf16x8 bh = nbh, bl = nbl;
        int rfn = rf, gn = g + 2;
        if (gn >= f.g1) {
            rfn = rf + 1;
            gn = rfn < it.ref1 ? refs[rfn].g0 : g;
        }
        nbh = Xp[((int64_t)(gn + h) * 2 + 0) * kMxRhs + c32];
        nbl = Xp[((int64_t)(gn + h) * 2 + 1) * kMxRhs + c32];
        const int col0 = gg * kGroup;
        if (col0 < f.lo || col0 + kGroup > f.hi) {
#pragma unroll
            for (int e = 0; e < kGroup; e++) {
                if (col0 + e < f.lo || col0 + e >= f.hi) {
                    bh[e] = (__bf16)0.f;
                    bl[e] = (__bf16)0.f;
                }
            }
        }
        f32x8 xc[DM];
#pragma unroll
        for (int d = 0; d < DM; d++) {
            if (DIM > 0 || d < dim)
                xc[d] = *(const f32x8*)(cp2 + ((int64_t)gg * dim + d) * kGroup);
            else
                xc[d] = f32x8{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int t = 0; t < kMxRowTiles; t++)
            if (t < ntile) acc[t] = near_tile<KIND, DIM, REG>(kp, dim, xr[t], xc, row[t], col0, bh, bl, acc[t]);
        if (rfn >= it.ref1) break;
        rf = rfn;
        g = gn;
    }
    float* __restrict__ out = P + (int64_t)q * kMxRowsItem * kMxRhs + c32;
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rr = t * 32 + (r >> 2) * 8 + h * 4 + (r & 3);
            if (rr < it.nrows) out[(int64_t)rr * kMxRhs] = acc[t][r];
        }
    }
}

// Workgroup form of the near field: kNwWaves wavefronts run the row slices of one
// column chunk (the same refs, so the same column groups in the same order); the
// workgroup stages each step's two X panel groups (2 KB: one 16-byte
// global_load_lds per lane) through LDS, double-buffered, so each group is read
// once per workgroup instead of once per row slice.  Per lane the arithmetic is
// that of near_mx_kernel, in the same order (bitwise the same partials).
// wg[kNwWaves * blockIdx.x + w] = wavefront w's item (-1: idle, it only helps stage).
constexpr int kNwStage = 2 * 2 * kMxRhs;  // bf16x8 entries of one step (two groups, hi + lo)

template <int KIND, int DIM, bool REG>
__global__ void __launch_bounds__(kNwWaves * kWave)
near_mx_wg_kernel(const float* __restrict__ cp1, const float* __restrict__ cp2, KernelParamsF kp,
                  const MxItem* __restrict__ items, const int* __restrict__ wg, const MxRef* __restrict__ refs,
                  const bf16x8* __restrict__ Xp, float* __restrict__ P)
{
    __shared__ bf16x8 sX[2][kNwStage];
    constexpr int DM = DIM > 0 ? DIM : MAX_POINT_DIMENSION;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = wave_id();
    const int qlead = wg[(int64_t)blockIdx.x * kNwWaves];
    const int qme = wg[(int64_t)blockIdx.x * kNwWaves + w];
    const bool active = qme >= 0;
    const MxItem lead = items[qlead];
    const MxItem it = items[active ? qme : qlead];
    const int dim = DIM > 0 ? DIM : kp.dim;
    const int c32 = lane & 31, h = lane >> 5;
    const int ntile = active ? (it.nrows + 31) >> 5 : 0;  // wave-uniform
    float xr[kMxRowTiles][DM];
    int row[kMxRowTiles];
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++) {
        const int rr = t * 32 + c32 < it.nrows ? t * 32 + c32 : 0;  // rows past the item: its first row
        row[t] = it.row0 + rr;
#pragma unroll
        for (int d = 0; d < DM; d++)
            xr[t][d] = (DIM > 0 || d < dim) ? cp1[((int64_t)(row[t] >> 3) * dim + d) * kGroup + (row[t] & 7)] : 0.f;
    }
    f32x16 acc[kMxRowTiles];
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][r] = 0.f;
    // steps: (ref rf, first group g) over the chunk's refs, two groups at a time
    int rf = lead.ref0;
    if (rf >= lead.ref1) return;  // uniform over the workgroup: no barrier is skipped
    int g = refs[rf].g0;
    auto stage = [&](int buf, int gs) {
        __builtin_amdgcn_global_load_lds((const void*)(Xp + (int64_t)gs * (2 * kMxRhs) + tid),
                                         (__attribute__((address_space(3))) void*)(sX[buf] + w * kWave), 16, 0,
                                         0);
    };
    stage(0, g);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int cur = 0;
    for (;;) {
        // the next step
        int rfn = rf, gn = g + 2;
        if (gn >= refs[rf].g1) {
            rfn = rf + 1;
            gn = rfn < lead.ref1 ? refs[rfn].g0 : 0;
        }
        const bool more = rfn < lead.ref1;
        if (more) stage(cur ^ 1, gn);
        if (ntile > 0) {
            const MxRef f = refs[rf];
            const int gg = g + h;
            bf16x8 bh = sX[cur][h * 2 * kMxRhs + c32];
            bf16x8 bl = sX[cur][h * 2 * kMxRhs + kMxRhs + c32];
            const int col0 = gg * kGroup;
            if (col0 < f.lo || col0 + kGroup > f.hi) {
#pragma unroll
                for (int e = 0; e < kGroup; e++) {
                    if (col0 + e < f.lo || col0 + e >= f.hi) {
                        bh[e] = (__bf16)0.f;
                        bl[e] = (__bf16)0.f;
                    }
                }
            }
            f32x8 xc[DM];
#pragma unroll
            for (int d = 0; d < DM; d++) {
                if (DIM > 0 || d < dim)
                    xc[d] = *(const f32x8*)(cp2 + ((int64_t)gg * dim + d) * kGroup);
                else
                    xc[d] = f32x8{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
            for (int t = 0; t < kMxRowTiles; t++)
                if (t < ntile) acc[t] = near_tile<KIND, DIM, REG>(kp, dim, xr[t], xc, row[t], col0, bh, bl, acc[t]);
        }
        if (more) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (!more) break;
        rf = rfn;
        g = gn;
        cur ^= 1;
    }
    if (!active) return;
    float* __restrict__ out = P + (int64_t)qme * kMxRowsItem * kMxRhs + c32;
#pragma unroll
    for (int t = 0; t < kMxRowTiles; t++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rr = t * 32 + (r >> 2) * 8 + h * 4 + (r & 3);
            if (rr < it.nrows) out[(int64_t)rr * kMxRhs] = acc[t][r];
        }
    }
}

// ---- low rank ----------------------------------------------------------------------------
// Both low-rank products are skinny FP32 GEMMs whose A operand streams from HBM:
//   V^T X: A[col][point] = V[point][col]   (V row-major: a point's columns contiguous)
//   U T:   A[row][col]   = U[col][row]     (U column-major: a column's rows contiguous)
// Lane (c, h) of a wavefront loads NT consecutive A values of the contiguous
// direction in one vector load: for k-step k, V[j+h][NT*c .. NT*c+NT-1] or
// ---- low-rank products: BF16 hi + lo panels, v_mfma_f32_32x32x16_bf16 ----------------
// A wavefront owns NT 32-wide output tiles; per 16-long reduction step each lane
// loads, for every tile, the 8 hi and 8 lo BF16 values of its A row (16 B each,
// the 32 lanes of a half-wave reading 512 contiguous bytes) and its B column's 8
// hi and 8 lo values, and issues hi*hi + hi*lo + lo*hi per tile (lo*lo, below
// FP32 resolution, is dropped).  The next step's operands load while the current
// step's MFMAs run; loads are never skipped (past the last step they repeat a
// valid position), so wait counts do not depend on the path.
template <int NT>
struct BfStage {
    bf16x8 ah[NT], al[NT];
    bf16x8 bh, bl;
};

// a: this lane's A row (tile e at + 256 e, lo at + a_lo); b: its B column (lo at + 256)
template <int NT>
__device__ __forceinline__ void bf_load(BfStage<NT>& st, const __bf16* __restrict__ a, int64_t a_lo,
                                        const __bf16* __restrict__ b)
{
#pragma unroll
    for (int e = 0; e < NT; e++) {
        st.ah[e] = *(const bf16x8*)(a + 256 * e);
        st.al[e] = *(const bf16x8*)(a + 256 * e + a_lo);
    }
    st.bh = *(const bf16x8*)b;
    st.bl = *(const bf16x8*)(b + 256);
}

template <int NT>
__device__ __forceinline__ void bf_mfma(f32x16 (&acc)[NT], const BfStage<NT>& st)
{
#pragma unroll
    for (int e = 0; e < NT; e++) {
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(st.ah[e], st.bh, acc[e], 0, 0, 0);
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(st.ah[e], st.bl, acc[e], 0, 0, 0);
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(st.al[e], st.bh, acc[e], 0, 0, 0);
    }
}

// acc[e] += sum over nsteps reduction steps (A advances a_step, B b_step per
// step); nsteps is even (the panels pad every reduction to whole step pairs)
template <int NT>
__device__ __forceinline__ void bf_body(f32x16 (&acc)[NT], const __bf16* __restrict__ a, int64_t a_step,
                                        int64_t a_lo, const __bf16* __restrict__ b, int64_t b_step, int nsteps)
{
    if (nsteps <= 0) return;
    BfStage<NT> A, B;
    bf_load<NT>(A, a, a_lo, b);
    for (int s = 0; s < nsteps; s += 2) {
        bf_load<NT>(B, a + (s + 1) * a_step, a_lo, b + (s + 1) * b_step);
        __builtin_amdgcn_sched_barrier(0);
        bf_mfma<NT>(acc, A);
        __builtin_amdgcn_sched_barrier(0);
        const int s2 = min(s + 2, nsteps - 1);
        bf_load<NT>(A, a + s2 * a_step, a_lo, b + s2 * b_step);
        __builtin_amdgcn_sched_barrier(0);
        bf_mfma<NT>(acc, B);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// T panel element of slot t, right-hand side c (hi; lo at + 256)
__device__ __forceinline__ int64_t tpan_at(int t, int c)
{
    return (int64_t)(t >> 3) * (2 * kMxRhs * kGroup) + c * kGroup + (t & 7);
}

__device__ __forceinline__ void store_split(__bf16* __restrict__ T, int64_t at, float v)
{
    const __bf16 h = (__bf16)v;
    T[at] = h;
    T[at + kMxRhs * kGroup] = (__bf16)(v - (float)h);
}

// T_D = V_D^T X for the item's columns [32 NT) of its slice over its point groups
// (M = columns, N = right-hand sides, K = points):
//   tp < 0:  T panel slot tmap[tcol0 + col] = result (single-chunk slice)
//   tp >= 0: Tp[(tp*128 + col)*32 + c] = partial for this chunk of points
template <int NT>
__global__ void __launch_bounds__(kThreads)
lr_vb_kernel(const LvItem* __restrict__ items, int nitems, const __bf16* __restrict__ Vp,
             const __bf16* __restrict__ Xp, const int* __restrict__ tmap, __bf16* __restrict__ Tpan,
             float* __restrict__ Tp, int prio)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    if (prio) __builtin_amdgcn_s_setprio(2);
    const LvItem it = items[q];
    const int c32 = lane & 31, h = lane >> 5;
    const int64_t gstride = 16 * (int64_t)it.ld;
    f32x16 acc[NT];
#pragma unroll
    for (int e = 0; e < NT; e++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[e][r] = 0.f;
    bf_body<NT>(acc, Vp + it.vbase + h * gstride + c32 * 8, 2 * gstride, 8 * (int64_t)it.ld,
                Xp + (int64_t)(it.j0 + h) * (2 * kMxRhs * kGroup) + c32 * kGroup, 2 * (2 * kMxRhs * kGroup),
                it.nrows / 2);
#pragma unroll
    for (int e = 0; e < NT; e++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int col = 32 * e + (r >> 2) * 8 + h * 4 + (r & 3);
            if (col < it.ncols) {
                if (it.tp < 0)
                    store_split(Tpan, tpan_at(tmap[it.tcol0 + col], c32), acc[e][r]);
                else
                    Tp[((int64_t)it.tp * kMxVCols + col) * kMxRhs + c32] = acc[e][r];
            }
        }
    }
}

// T panel slot tmap[tcol0 + col] = sum of the slice's point chunks, in chunk order
__global__ void __launch_bounds__(kThreads)
lr_vb_finalize_kernel(const LvSlice* __restrict__ slices, const float* __restrict__ Tp,
                      const int* __restrict__ tmap, __bf16* __restrict__ Tpan)
{
    const LvSlice sl = slices[blockIdx.x];
    for (int e = threadIdx.x; e < sl.ncols * kMxRhs; e += kThreads) {
        const int col = e / kMxRhs, c = e % kMxRhs;
        float s = 0.f;
        for (int ch = 0; ch < sl.nchunks; ch++)
            s += Tp[((int64_t)(sl.item0 + ch) * kMxVCols + col) * kMxRhs + c];
        store_split(Tpan, tpan_at(tmap[sl.tcol0 + col], c), s);
    }
}

// U T: P[(q*128 + r)*32 + c] = sum_k U_C[row0 + r][k] T_C[k][c], r < nrows <= 32*NT
// (M = rows, N = right-hand sides, K = the row cluster's rank columns)
template <int NT>
__global__ void __launch_bounds__(kThreads)
lr_ub_kernel(const LrItem* __restrict__ items, int nitems, const __bf16* __restrict__ Up,
             const __bf16* __restrict__ Tpan, float* __restrict__ P, int prio)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = blockIdx.x * kWaves + wave_id();
    if (q >= nitems) return;
    if (prio) __builtin_amdgcn_s_setprio(2);
    const LrItem it = items[q];
    const int c32 = lane & 31, h = lane >> 5;
    const int64_t gstride = 16 * (int64_t)it.ld;
    f32x16 acc[NT];
#pragma unroll
    for (int e = 0; e < NT; e++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[e][r] = 0.f;
    bf_body<NT>(acc, Up + it.uoff + h * gstride + c32 * 8, 2 * gstride, 8 * (int64_t)it.ld,
                Tpan + (it.toff + h) * (2 * kMxRhs * kGroup) + c32 * kGroup, 2 * (2 * kMxRhs * kGroup),
                it.K / 16);
    float* __restrict__ out = P + (int64_t)q * kMxLrRows * kMxRhs + c32;
#pragma unroll
    for (int e = 0; e < NT; e++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int rr = 32 * e + (r >> 2) * 8 + h * 4 + (r & 3);
            if (rr < it.nrows) out[(int64_t)rr * kMxRhs] = acc[e][r];
        }
    }
}

// ---- low-rank products, workgroup form ------------------------------------------
// W wavefronts per workgroup, each with its own WgItem (its own A columns or rows,
// 1, 2 or 4 interleaved 32-wide tiles) against the B panel groups [b0, b0 + nb)
// that all of them share.  The workgroup stages B through LDS kWgGroups groups
// (kWgGroups / 2 reduction steps) at a time, double-buffered, so each B element is
// loaded once per workgroup instead of once per wavefront; A streams from HBM one
// reduction step ahead of the MFMAs, as in the one-wavefront kernels above.
constexpr int kWgGroups = 8;
constexpr int kWgEntries = kWgGroups * 2 * kMxRhs;  // bf16x8 entries of one staged block (8 KB)

template <int NT>
__device__ __forceinline__ void wg_load_a(bf16x8 (&hi)[NT], bf16x8 (&lo)[NT], const __bf16* __restrict__ a,
                                          int64_t a_lo)
{
#pragma unroll
    for (int e = 0; e < NT; e++) {
        hi[e] = *(const bf16x8*)(a + 256 * e);
        lo[e] = *(const bf16x8*)(a + 256 * e + a_lo);
    }
}

template <int NT>
__device__ __forceinline__ void wg_mfma(f32x16 (&acc)[NT], const bf16x8 (&hi)[NT], const bf16x8 (&lo)[NT],
                                        bf16x8 bh, bf16x8 bl)
{
#pragma unroll
    for (int e = 0; e < NT; e++) {
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[e], bh, acc[e], 0, 0, 0);
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[e], bl, acc[e], 0, 0, 0);
        acc[e] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[e], bh, acc[e], 0, 0, 0);
    }
}

// B groups [g, g + kWgGroups) of Bpan (bf16x8 entries, 2 * kMxRhs = one wavefront's
// 64 lanes per group) straight into the LDS block sb (global_load_lds, 16 bytes per
// lane: the data never passes through registers).  Groups at or past bgroups
// re-read the last group (no reduction step uses them).  Complete after a vmcnt(0)
// wait.
template <int W>
__device__ __forceinline__ void wg_stage_b(bf16x8* sb, const bf16x8* __restrict__ Bpan, int64_t g,
                                           int64_t bgroups, int w, int lane)
{
#pragma unroll
    for (int u = 0; u < kWgGroups / W; u++) {
        const int gl = w + u * W;  // group of this wavefront in the block
        const int64_t gg = min(g + gl, bgroups - 1);
        __builtin_amdgcn_global_load_lds((const void*)(Bpan + gg * (2 * kMxRhs) + lane),
                                         (__attribute__((address_space(3))) void*)(sb + gl * (2 * kMxRhs)), 16,
                                         0, 0);
    }
}

// one wavefront of a workgroup item: NT tiles (NT = 0: no item, staging only)
template <int NT, int W, bool VB>
__device__ __forceinline__ void wg_wave(const WgItem& it, int lane, int tid, const __bf16* __restrict__ Apan,
                                        const bf16x8* __restrict__ Bpan, int64_t bgroups,
                                        bf16x8 (*sB)[kWgEntries], const int* __restrict__ tmap,
                                        __bf16* __restrict__ Tpan, float* __restrict__ out)
{
    const int c32 = lane & 31, h = lane >> 5;
    const int w = tid / kWave;
    const int nsteps = it.nb / 2;                      // even (nb is a multiple of 4)
    const int nblk = (it.nb + kWgGroups - 1) / kWgGroups;
    const int64_t gstride = 16 * (int64_t)it.ld;
    const __bf16* __restrict__ a = Apan + it.aoff + h * gstride + c32 * 8;
    const int64_t a_step = 2 * gstride, a_lo = 8 * (int64_t)it.ld;
    f32x16 acc[NT > 0 ? NT : 1];
#pragma unroll
    for (int e = 0; e < (NT > 0 ? NT : 1); e++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[e][r] = 0.f;
    bf16x8 ph[NT > 0 ? NT : 1], pl[NT > 0 ? NT : 1], qh[NT > 0 ? NT : 1], ql[NT > 0 ? NT : 1];
    if constexpr (NT > 0) wg_load_a<NT>(ph, pl, a, a_lo);
    for (int blk = 0; blk < nblk; blk++) {
        const int cur = blk & 1;
        const bool more = blk + 1 < nblk;
        // the next block into the other buffer (last read before the previous barrier)
        if (more) wg_stage_b<W>(sB[cur ^ 1], Bpan, (int64_t)it.b0 + (int64_t)(blk + 1) * kWgGroups, bgroups, w, lane);
        if constexpr (NT > 0) {
            const int s0 = blk * (kWgGroups / 2);
            const int s1 = min(nsteps, s0 + kWgGroups / 2);
            const bf16x8* __restrict__ sb = sB[cur];
            for (int s = s0; s < s1; s += 2) {
                wg_load_a<NT>(qh, ql, a + (s + 1) * a_step, a_lo);
                __builtin_amdgcn_sched_barrier(0);
                {
                    const int gi = 2 * (s - s0) + h;
                    wg_mfma<NT>(acc, ph, pl, sb[gi * 2 * kMxRhs + c32], sb[gi * 2 * kMxRhs + kMxRhs + c32]);
                }
                __builtin_amdgcn_sched_barrier(0);
                wg_load_a<NT>(ph, pl, a + min(s + 2, nsteps - 1) * a_step, a_lo);
                __builtin_amdgcn_sched_barrier(0);
                {
                    const int gi = 2 * (s + 1 - s0) + h;
                    wg_mfma<NT>(acc, qh, ql, sb[gi * 2 * kMxRhs + c32], sb[gi * 2 * kMxRhs + kMxRhs + c32]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (more) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if constexpr (NT > 0) {
#pragma unroll
        for (int e = 0; e < NT; e++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int mm = 32 * e + (r >> 2) * 8 + h * 4 + (r & 3);
                if (mm < it.m) {
                    if (VB) {
                        if (it.tp < 0)
                            store_split(Tpan, tpan_at(tmap[it.o0 + mm], c32), acc[e][r]);
                        else
                            out[((int64_t)it.tp * kMxVCols + mm) * kMxRhs + c32] = acc[e][r];
                    } else {
                        out[((int64_t)it.o0 * kMxLrRows + mm) * kMxRhs + c32] = acc[e][r];
                    }
                }
            }
        }
    }
}

// V^T X (VB) or U T (!VB) over workgroup items: items[W*blockIdx.x + w] is wavefront
// w's share; all W share b0 and nb.  VB: out = the chunk partials Tp (or T panels
// directly); !VB: out = the partial-row slots P.
template <int W, bool VB>
__global__ void __launch_bounds__(W * kWave) __attribute__((amdgpu_waves_per_eu(3)))
lr_wg_kernel(const WgItem* __restrict__ items, const __bf16* __restrict__ Apan, const bf16x8* __restrict__ Bpan,
             int64_t bgroups, const int* __restrict__ tmap, __bf16* __restrict__ Tpan, float* __restrict__ out)
{
    __shared__ bf16x8 sB[2][kWgEntries];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = wave_id();
    const WgItem it = items[(int64_t)blockIdx.x * W + w];
    // the first block of B, staged by the whole workgroup
    wg_stage_b<W>(sB[0], Bpan, it.b0, bgroups, w, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int nt = it.m <= 0 ? 0 : it.m <= 32 ? 1 : it.m <= 64 ? 2 : 4;  // wave-uniform
    if (nt == 4)
        wg_wave<4, W, VB>(it, lane, tid, Apan, Bpan, bgroups, sB, tmap, Tpan, out);
    else if (nt == 2)
        wg_wave<2, W, VB>(it, lane, tid, Apan, Bpan, bgroups, sB, tmap, Tpan, out);
    else if (nt == 1)
        wg_wave<1, W, VB>(it, lane, tid, Apan, Bpan, bgroups, sB, tmap, Tpan, out);
    else
        wg_wave<0, W, VB>(it, lane, tid, Apan, Bpan, bgroups, sB, tmap, Tpan, out);
}

// Ys[i*32 + c] = sum of sorted row i's partials, in list order (deterministic)
__global__ void __launch_bounds__(kThreads)
combine_mx_kernel(const int* __restrict__ ptr, const int* __restrict__ idx, const float* __restrict__ P,
                  float* __restrict__ Ys, int n1)
{
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int i = (int)(t / kMxRhs), c = (int)(t % kMxRhs);
    if (i >= n1) return;
    // four partials loaded at a time, added in list order
    float s = 0.f;
    int q = ptr[i];
    const int qe = ptr[i + 1];
    for (; q + 4 <= qe; q += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = P[(int64_t)idx[q + u] * kMxRhs + c];
#pragma unroll
        for (int u = 0; u < 4; u++) s += v[u];
    }
    for (; q < qe; q++) s += P[(int64_t)idx[q] * kMxRhs + c];
    Ys[t] = s;
}

// Y[p + (c0 + c)*n1] = Ys[inv1[p]*32 + c] for c < nr: 64 caller rows per block,
// read as whole 128-byte sorted rows, transposed through LDS, written as 64
// consecutive doubles per right-hand side
constexpr int kPermRows = 64;
__global__ void __launch_bounds__(kThreads)
permute_out_kernel(const float* __restrict__ Ys, const int* __restrict__ inv1, double* __restrict__ Y,
                   int n1, int c0, int nr)
{
    __shared__ float tile[kMxRhs][kPermRows + 1];
    const int p0 = blockIdx.x * kPermRows;
#pragma unroll
    for (int k = 0; k < kPermRows * kMxRhs / kThreads; k++) {
        const int e = threadIdx.x + kThreads * k;
        const int pl = e / kMxRhs, c = e % kMxRhs;
        const int p = p0 + pl;
        tile[c][pl] = p < n1 ? Ys[(int64_t)inv1[p] * kMxRhs + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPermRows * kMxRhs / kThreads; k++) {
        const int e = threadIdx.x + kThreads * k;
        const int c = e / kPermRows, pl = e % kPermRows;
        const int p = p0 + pl;
        if (c < nr && p < n1) Y[p + (int64_t)(c0 + c) * n1] = (double)tile[c][pl];
    }
}

__global__ void inv32_kernel(const uint64_t* __restrict__ order, int* __restrict__ inv, int n)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) inv[order[i]] = i;
}

template <int KIND, int DIM>
void launch_near_mx(bool reg, unsigned grid, hipStream_t s, const Engine& eng, const KernelParamsF& kp,
                    float* Pn)
{
    const bf16x8* xp = (const bf16x8*)eng.Xp.p;
    // the LDS reservation caps near-field workgroups per CU so the low-rank
    // kernels on the other stream keep register room beside them
    const size_t lds = eng.serialize ? 0 : (size_t)eng.near_lds_pad;
    if (eng.near_wg) {
        const unsigned gw = (unsigned)(eng.h_mx_wg.size() / kNwWaves);
        if (gw == 0) return;
        if (reg)
            near_mx_wg_kernel<KIND, DIM, true><<<gw, kNwWaves * kWave, lds, s>>>(
                eng.cpan[0].p, eng.cpan[1].p, kp, eng.mx_items.p, eng.mx_wg.p, eng.mx_refs.p, xp, Pn);
        else
            near_mx_wg_kernel<KIND, DIM, false><<<gw, kNwWaves * kWave, lds, s>>>(
                eng.cpan[0].p, eng.cpan[1].p, kp, eng.mx_items.p, eng.mx_wg.p, eng.mx_refs.p, xp, Pn);
        return;
    }
    if (reg)
        near_mx_kernel<KIND, DIM, true><<<grid, kThreads, lds, s>>>(
            eng.cpan[0].p, eng.cpan[1].p, kp, eng.mx_items.p, eng.n_mx_items, eng.mx_refs.p, xp, Pn);
    else
        near_mx_kernel<KIND, DIM, false><<<grid, kThreads, lds, s>>>(
            eng.cpan[0].p, eng.cpan[1].p, kp, eng.mx_items.p, eng.n_mx_items, eng.mx_refs.p, xp, Pn);
}

template <int KIND>
void launch_near_mx_dim(bool reg, unsigned grid, hipStream_t s, const Engine& eng, const KernelParamsF& kp,
                        float* Pn)
{
    if (kp.dim == 3)
        launch_near_mx<KIND, 3>(reg, grid, s, eng, kp, Pn);
    else if (kp.dim == 2)
        launch_near_mx<KIND, 2>(reg, grid, s, eng, kp, Pn);
    else
        launch_near_mx<KIND, 0>(reg, grid, s, eng, kp, Pn);
}

// point groups of the X and coordinate panels: 4 zero groups past the last point
// (V panels of the last column cluster reach up to 3 groups past it)
inline int64_t groups_of(int n) { return (n + kGroup - 1) / kGroup + 4; }

}  // namespace

hmg_status hmvp_mixed(Engine& eng, const Coords& c1, const Coords& c2, const KernelParams& kp,
                      const uint64_t* order1, const uint64_t* order2, const double* X, double* Y,
                      int nrhs)
{
    if (eng.layout_prec != HMG_PREC_MIXED || !eng.mx_sched) {
        set_error("mixed-precision product without a mixed-precision layout");
        return HMG_ERR_NOT_READY;
    }
    hipStream_t s = eng.stream, s2 = eng.serialize ? eng.stream : eng.stream2;
    const int n1 = eng.n[0], n2 = eng.n[1];
    const int dim = kp.dim;
    const int64_t g1 = groups_of(n1), g2 = groups_of(n2);
    const KernelParamsF kpf = make_params_f(kp);
    if (!eng.cpan_ready || eng.cpan_scale != kpf.cscale) {
        HMG_TRY(eng.cpan[0].alloc(g1 * dim * kGroup));
        HMG_TRY(eng.cpan[1].alloc(g2 * dim * kGroup));
        coord_panel_kernel<<<grid_for(g1 * kGroup, kThreads), kThreads, 0, s>>>(c1, dim, (int)g1, kpf.cscale,
                                                                                eng.cpan[0].p);
        HMG_HIP(hipGetLastError());
        coord_panel_kernel<<<grid_for(g2 * kGroup, kThreads), kThreads, 0, s>>>(c2, dim, (int)g2, kpf.cscale,
                                                                                eng.cpan[1].p);
        HMG_HIP(hipGetLastError());
        eng.cpan_ready = true;
        eng.cpan_scale = kpf.cscale;
    }
    const int n2p = (int)(g2 * kGroup);
    const int64_t near0 = (int64_t)eng.n_lr_items * kMxLrRows;
    const int64_t slots = near0 + (int64_t)eng.n_mx_items * kMxRowsItem;
    HMG_TRY(eng.Xp.alloc((size_t)g2 * 2 * kMxRhs * kGroup));
    HMG_TRY(eng.Xt32.alloc((size_t)n2 * kMxRhs));
    // T panels: slots no V column maps to (each row cluster's padding) stay zero
    const int64_t tgroups = eng.tsize / kGroup + 2;
    if (!eng.tpan_ready) {
        const size_t n = (size_t)tgroups * 2 * kMxRhs * kGroup;
        HMG_TRY(eng.Tpan.alloc(n));
        HMG_HIP(hipMemsetAsync(eng.Tpan.p, 0, sizeof(__bf16) * n, s));
        eng.tpan_ready = true;
    }
    HMG_TRY(eng.Tp32.alloc((size_t)std::max(eng.n_tp, 1) * kMxVCols * kMxRhs));
    HMG_TRY(eng.P32.alloc((size_t)std::max<int64_t>(slots, 1) * kMxRhs));
    HMG_TRY(eng.Ys32.alloc((size_t)n1 * kMxRhs));
    if (!eng.inv1_ready) {
        HMG_TRY(eng.inv1.alloc(n1));
        inv32_kernel<<<grid_for(n1, kThreads), kThreads, 0, s>>>(order1, eng.inv1.p, n1);
        HMG_HIP(hipGetLastError());
        eng.inv1_ready = true;
    }
    const bool reg = kp.reg != 0.0;
    const __bf16* U = eng.Up.p;
    const __bf16* V = eng.Vp.p;
    for (int c0 = 0; c0 < nrhs; c0 += kMxRhs) {
        const int nr = std::min(kMxRhs, nrhs - c0);
        eng.prof.begin(PROF_GATHER, s);
        x_rows_kernel<<<grid_for(n2, kRowTile), kThreads, 0, s>>>(X, n2, c0, nr, eng.Xt32.p);
        HMG_HIP(hipGetLastError());
        prep_x_kernel<<<grid_for((int64_t)n2p * kMxRhs, kThreads), kThreads, 0, s>>>(
            eng.Xt32.p, order2, n2, n2p, eng.Xp.p);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_GATHER, s);
        HMG_HIP(hipEventRecord(eng.ev_fork, s));
        HMG_HIP(hipStreamWaitEvent(s2, eng.ev_fork, 0));
        if (eng.n_mx_items > 0) {
            const unsigned g = grid_for(eng.n_mx_items, kWaves);
            float* Pn = eng.P32.p + near0 * kMxRhs;
            eng.prof.begin(PROF_NEAR, s2);
            switch (kp.kind) {
            case HMG_KERNEL_GAUSSIAN:
                launch_near_mx_dim<HMG_KERNEL_GAUSSIAN>(reg, g, s2, eng, kpf, Pn);
                break;
            case HMG_KERNEL_EXPONENTIAL:
                launch_near_mx_dim<HMG_KERNEL_EXPONENTIAL>(reg, g, s2, eng, kpf, Pn);
                break;
            case HMG_KERNEL_GAUSSIAN_REF:
                launch_near_mx_dim<HMG_KERNEL_GAUSSIAN_REF>(reg, g, s2, eng, kpf, Pn);
                break;
            default:
                launch_near_mx_dim<HMG_KERNEL_MATERN_REF>(reg, g, s2, eng, kpf, Pn);
                break;
            }
            HMG_HIP(hipGetLastError());
            eng.prof.end(PROF_NEAR, s2);
        }
        HMG_HIP(hipEventRecord(eng.ev_join, s2));
        if (eng.naca > 0 && eng.n_lr_items > 0 && eng.lr_wg) {
            // workgroup form: B staged once per workgroup through LDS
            eng.prof.begin(PROF_LR_VTX, s);
            const unsigned gv = (unsigned)(eng.h_vw_items.size() / kVwWaves);
            if (gv > 0) {
                lr_wg_kernel<kVwWaves, true><<<gv, kVwWaves * kWave, 0, s>>>(
                    eng.vw_items.p, V, (const bf16x8*)eng.Xp.p, g2, eng.tmap.p, eng.Tpan.p, eng.Tp32.p);
                HMG_HIP(hipGetLastError());
            }
            if (eng.n_lv_slices > 0) {
                lr_vb_finalize_kernel<<<eng.n_lv_slices, kThreads, 0, s>>>(eng.lv_slices.p, eng.Tp32.p,
                                                                           eng.tmap.p, eng.Tpan.p);
                HMG_HIP(hipGetLastError());
            }
            eng.prof.end(PROF_LR_VTX, s);
            eng.prof.begin(PROF_LR_U, s);
            const unsigned gu = (unsigned)(eng.h_uw_items.size() / kUwWaves);
            if (gu > 0) {
                lr_wg_kernel<kUwWaves, false><<<gu, kUwWaves * kWave, 0, s>>>(
                    eng.uw_items.p, U, (const bf16x8*)eng.Tpan.p, tgroups, nullptr, nullptr, eng.P32.p);
                HMG_HIP(hipGetLastError());
            }
            eng.prof.end(PROF_LR_U, s);
        } else if (eng.naca > 0 && eng.n_lr_items > 0) {
            const int prio_v = eng.lr_prio & 1, prio_u = (eng.lr_prio >> 1) & 1;
            eng.prof.begin(PROF_LR_VTX, s);
            // items grouped by interleaved tile count: [4 | 2 | 1]
            for (int g = 0; g < 3; g++) {
                const int b = eng.lv_nt_off[g], e = eng.lv_nt_off[g + 1];
                if (e <= b) continue;
                const LvItem* li = eng.lv_items.p + b;
                const unsigned gr = grid_for(e - b, kWaves);
                if (g == 0)
                    lr_vb_kernel<4><<<gr, kThreads, 0, s>>>(li, e - b, V, eng.Xp.p, eng.tmap.p, eng.Tpan.p, eng.Tp32.p,
                                                    prio_v);
                else if (g == 1)
                    lr_vb_kernel<2><<<gr, kThreads, 0, s>>>(li, e - b, V, eng.Xp.p, eng.tmap.p, eng.Tpan.p, eng.Tp32.p,
                                                    prio_v);
                else
                    lr_vb_kernel<1><<<gr, kThreads, 0, s>>>(li, e - b, V, eng.Xp.p, eng.tmap.p, eng.Tpan.p, eng.Tp32.p,
                                                    prio_v);
                HMG_HIP(hipGetLastError());
            }
            if (eng.n_lv_slices > 0) {
                lr_vb_finalize_kernel<<<eng.n_lv_slices, kThreads, 0, s>>>(eng.lv_slices.p, eng.Tp32.p,
                                                                           eng.tmap.p, eng.Tpan.p);
                HMG_HIP(hipGetLastError());
            }
            eng.prof.end(PROF_LR_VTX, s);
            eng.prof.begin(PROF_LR_U, s);
            // items grouped by interleaved tile count: [4 | 2 | 1]
            for (int g = 0; g < 3; g++) {
                const int b = eng.lr_nt_off[g], e = eng.lr_nt_off[g + 1];
                if (e <= b) continue;
                const LrItem* li = eng.lr_items.p + b;
                float* Pg = eng.P32.p + (int64_t)b * kMxLrRows * kMxRhs;
                const unsigned gr = grid_for(e - b, kWaves);
                if (g == 0)
                    lr_ub_kernel<4><<<gr, kThreads, 0, s>>>(li, e - b, U, eng.Tpan.p, Pg, prio_u);
                else if (g == 1)
                    lr_ub_kernel<2><<<gr, kThreads, 0, s>>>(li, e - b, U, eng.Tpan.p, Pg, prio_u);
                else
                    lr_ub_kernel<1><<<gr, kThreads, 0, s>>>(li, e - b, U, eng.Tpan.p, Pg, prio_u);
                HMG_HIP(hipGetLastError());
            }
            eng.prof.end(PROF_LR_U, s);
        }
        HMG_HIP(hipStreamWaitEvent(s, eng.ev_join, 0));
        eng.prof.begin(PROF_COMBINE, s);
        combine_mx_kernel<<<grid_for((int64_t)n1 * kMxRhs, kThreads), kThreads, 0, s>>>(
            eng.comb_mx.ptr.p, eng.comb_mx.idx.p, eng.P32.p, eng.Ys32.p, n1);
        HMG_HIP(hipGetLastError());
        permute_out_kernel<<<grid_for(n1, kPermRows), kThreads, 0, s>>>(eng.Ys32.p, eng.inv1.p, Y, n1, c0, nr);
        HMG_HIP(hipGetLastError());
        eng.prof.end(PROF_COMBINE, s);
    }
    return HMG_OK;
}

}  // namespace hmg
