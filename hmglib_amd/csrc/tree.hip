// tree.hip -- block-cluster-tree traversal on the device, level by level
// (reference: tree.cu:1429-1557, traverse_with_dynamic_arrays_dynamic_output).
//
// Per level and per point set the bounding boxes follow the reference's lookup
// table semantics exactly (tree.cu:852-1147), including its behaviour for size-1
// clusters: unique lower bounds L[t] (rank map via a flag array + scan over the
// N points instead of a sort), keys[L[t]] = t+1 then keys[U[t]] = -(t+1) by the
// same thread, inclusive scan, keys[U[t]] = t+1, and table entry k = min/max over
// the k-th maximal run of equal nonzero keys.  Admissibility uses the diameter
// with an explicit fused multiply-add, as nvcc contracts tree.cu:100 by default;
// this file is compiled with -ffp-contract=off so nothing else is fused.
// Leaves are appended in node order per level (BFS) and finally stably
// partitioned DENSE first (linear_algebra.cu:63-86).
#include "hmg_internal.h"

namespace hmg {

namespace {

constexpr int kT = 256;
inline unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ int clz64(uint64_t x) { return x ? __clzll((long long)x) : 64; }

// tree.cu:40-78
__device__ int find_split(const uint64_t* __restrict__ codes, int first, int last)
{
    const uint64_t fc = codes[first], lc = codes[last];
    if (fc == lc) return (first + last) >> 1;
    const int common = clz64(fc ^ lc);
    int split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        const int ns = split + step;
        if (ns < last) {
            if (clz64(fc ^ codes[ns]) > common) split = ns;
        }
    } while (step > 1);
    return split;
}

__device__ __forceinline__ double diameter(const double* mn, const double* mx, int dim)
{
    double diam = 0.0;
    for (int d = 0; d < dim; d++) {
        const double t = __dsub_rn(mx[d], mn[d]);
        diam = __fma_rn(t, t, diam);
    }
    return sqrt(diam);
}

__device__ __forceinline__ double distance(const work_item& w)
{
    double dist = 0.0;
    for (int d = 0; d < w.dim; d++) {
        const double a = fmax(0.0, __dsub_rn(w.min1[d], w.max2[d]));
        const double b = fmax(0.0, __dsub_rn(w.min2[d], w.max1[d]));
        dist = __dadd_rn(dist, __dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)));
    }
    return sqrt(dist);
}

// tree.cu:121-160
__device__ __forceinline__ bool admissible(const work_item& w, double eta)
{
    const double d1 = diameter(w.min1, w.max1, w.dim);
    const double d2 = diameter(w.min2, w.max2, w.dim);
    const double dist = distance(w);
    return fmax(d1, d2) <= __dmul_rn(eta, dist);
}

// ---- per-level bounding boxes --------------------------------------------------
__global__ void mark_kernel(const work_item* __restrict__ cur, int ncur, int set,
                            int* __restrict__ flags, int* __restrict__ ub)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    const int l = set == 0 ? cur[i].set1_l : cur[i].set2_l;
    const int u = set == 0 ? cur[i].set1_u : cur[i].set2_u;
    if (l > u) return;  // empty cluster (only reachable through the 2-child split)
    flags[l] = 1;
    ub[l] = u;
}

__global__ void compact_kernel(const int* __restrict__ flags, const int* __restrict__ ub,
                               const int* __restrict__ rank, int n, int* __restrict__ Lc,
                               int* __restrict__ Uc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flags[p]) return;
    Lc[rank[p]] = p;
    Uc[rank[p]] = ub[p];
}

// set_bounds_for_keys_using_limits (tree.cu:692-701); flags a level with a singleton
__global__ void keys_set_kernel(const int* __restrict__ Lc, const int* __restrict__ Uc,
                                const int* __restrict__ Tdev, int* __restrict__ keys,
                                int* __restrict__ diag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *Tdev) return;
    keys[Lc[t]] = t + 1;
    keys[Uc[t]] = -(t + 1);
    if (Lc[t] == Uc[t]) diag[0] = 1;
}

// correct_upper_bound_using_limits (tree.cu:723-732)
__global__ void keys_fix_kernel(const int* __restrict__ Uc, const int* __restrict__ Tdev,
                                int* __restrict__ keys)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *Tdev) return;
    keys[Uc[t]] = t + 1;
}

__global__ void run_flags_kernel(const int* __restrict__ keys, int n, int* __restrict__ S)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int k = keys[p];
    S[p] = (k != 0 && (p == 0 || keys[p - 1] != k)) ? 1 : 0;
}

__global__ void run_bounds_kernel(const int* __restrict__ keys, const int* __restrict__ S,
                                  const int* __restrict__ runid, int n, int* __restrict__ rstart,
                                  int* __restrict__ rend)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int k = keys[p];
    if (k == 0) return;
    const int r = runid[p] + S[p] - 1;
    if (S[p]) rstart[r] = p;
    if (p == n - 1 || keys[p + 1] != k) rend[r] = p;
}

// one block per table entry t < T: exact min/max over the t-th run
__global__ void __launch_bounds__(kT)
run_minmax_kernel(Coords c, int dim, const int* __restrict__ rstart, const int* __restrict__ rend,
                  const int* __restrict__ Rdev, const int* __restrict__ Tdev, double* __restrict__ tmin,
                  double* __restrict__ tmax, int* __restrict__ diag)
{
    const int t = blockIdx.x;
    if (t >= *Tdev) return;
    __shared__ double smin[MAX_POINT_DIMENSION][kT];
    __shared__ double smax[MAX_POINT_DIMENSION][kT];
    if (t >= *Rdev) {  // read past the last run: undefined in the reference
        if (threadIdx.x < dim) {
            tmin[(int64_t)t * MAX_POINT_DIMENSION + threadIdx.x] = __builtin_nan("");
            tmax[(int64_t)t * MAX_POINT_DIMENSION + threadIdx.x] = __builtin_nan("");
        }
        if (threadIdx.x == 0) diag[1] = 1;
        return;
    }
    const int a = rstart[t], b = rend[t];
    for (int d = 0; d < dim; d++) {
        double lo = __builtin_inf(), hi = -__builtin_inf();
        for (int p = a + threadIdx.x; p <= b; p += kT) {
            const double v = c.x[d][p];
            lo = v < lo ? v : lo;
            hi = hi < v ? v : hi;
        }
        smin[d][threadIdx.x] = lo;
        smax[d][threadIdx.x] = hi;
    }
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int d = 0; d < dim; d++) {
                const double x = smin[d][threadIdx.x + s], y = smax[d][threadIdx.x + s];
                smin[d][threadIdx.x] = x < smin[d][threadIdx.x] ? x : smin[d][threadIdx.x];
                smax[d][threadIdx.x] = smax[d][threadIdx.x] < y ? y : smax[d][threadIdx.x];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < dim) {
        tmin[(int64_t)t * MAX_POINT_DIMENSION + threadIdx.x] = smin[threadIdx.x][0];
        tmax[(int64_t)t * MAX_POINT_DIMENSION + threadIdx.x] = smax[threadIdx.x][0];
    }
}

// set_bounding_box_minmax_using_lookup_table (tree.cu:1114-1147)
__global__ void assign_box_kernel(work_item* __restrict__ cur, int ncur, int set, int dim,
                                  const int* __restrict__ rank, const double* __restrict__ tmin,
                                  const double* __restrict__ tmax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    work_item& w = cur[i];
    const int l = set == 0 ? w.set1_l : w.set2_l;
    const int u = set == 0 ? w.set1_u : w.set2_u;
    w.dim = dim;
    if (l > u) return;
    const int t = rank[l];
    for (int d = 0; d < dim; d++) {
        const double lo = tmin[(int64_t)t * MAX_POINT_DIMENSION + d];
        const double hi = tmax[(int64_t)t * MAX_POINT_DIMENSION + d];
        if (set == 0) {
            w.min1[d] = lo;
            w.max1[d] = hi;
        } else {
            w.min2[d] = lo;
            w.max2[d] = hi;
        }
    }
    w.dim = dim;
}

// count_for_new_level (tree.cu:215-266)
__global__ void count_kernel(const work_item* __restrict__ cur, int ncur, double eta, int level,
                             int max_level, int c_leaf, int* __restrict__ nch, int* __restrict__ nmv,
                             int* __restrict__ adm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    const work_item w = cur[i];
    const bool a = admissible(w, eta);
    adm[i] = a ? 1 : 0;
    if (w.set1_l > w.set1_u || w.set2_l > w.set2_u) {
        // an empty cluster (2-child split of a size-1 cluster; undefined in the
        // reference): the node is dropped
        nch[i] = 0;
        nmv[i] = 0;
    } else if (w.level_1 == w.level_2) {
        if (a) {
            nch[i] = 0;
            nmv[i] = 1;
        } else if ((w.set1_u - w.set1_l + 1) <= c_leaf || (w.set2_u - w.set2_l + 1) <= c_leaf ||
                   (level + 1) >= max_level) {
            nch[i] = 0;
            nmv[i] = 1;
        } else {
            nch[i] = 4;
            nmv[i] = 0;
        }
    } else {
        nch[i] = 2;
        nmv[i] = 0;
    }
}

__device__ __forceinline__ void make_child(work_item& c, int l1, int u1, int l2, int u2, int lv1,
                                           int lv2)
{
    c = work_item{};
    c.set1_l = l1;
    c.set1_u = u1;
    c.set2_l = l2;
    c.set2_u = u2;
    c.level_1 = (char)lv1;
    c.level_2 = (char)lv2;
}

// generate_new_level (tree.cu:320-447)
__global__ void generate_kernel(const work_item* __restrict__ cur, int ncur,
                                const int* __restrict__ nch, const int* __restrict__ choff,
                                const int* __restrict__ mvoff, const int* __restrict__ adm,
                                const uint64_t* __restrict__ codes1, const uint64_t* __restrict__ codes2,
                                work_item* __restrict__ next, work_item* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    const work_item w = cur[i];
    const int c = nch[i];
    if (c == 4) {
        const int s1 = find_split(codes1, w.set1_l, w.set1_u);
        const int s2 = find_split(codes2, w.set2_l, w.set2_u);
        const int o = choff[i];
        const int a1 = w.level_1 + 1, a2 = w.level_2 + 1;
        make_child(next[o + 0], w.set1_l, s1, w.set2_l, s2, a1, a2);
        make_child(next[o + 1], w.set1_l, s1, s2 + 1, w.set2_u, a1, a2);
        make_child(next[o + 2], s1 + 1, w.set1_u, w.set2_l, s2, a1, a2);
        make_child(next[o + 3], s1 + 1, w.set1_u, s2 + 1, w.set2_u, a1, a2);
    } else if (c == 2) {
        const int o = choff[i];
        if (w.level_1 < w.level_2) {
            const int s1 = find_split(codes1, w.set1_l, w.set1_u);
            make_child(next[o + 0], w.set1_l, s1, w.set2_l, w.set2_u, w.level_1 + 1, w.level_2);
            make_child(next[o + 1], s1 + 1, w.set1_u, w.set2_l, w.set2_u, w.level_1 + 1, w.level_2);
        } else {
            const int s2 = find_split(codes2, w.set2_l, w.set2_u);
            make_child(next[o + 0], w.set1_l, w.set1_u, w.set2_l, s2, w.level_1, w.level_2 + 1);
            make_child(next[o + 1], w.set1_l, w.set1_u, s2 + 1, w.set2_u, w.level_1, w.level_2 + 1);
        }
    } else if (w.set1_l <= w.set1_u && w.set2_l <= w.set2_u) {
        work_item leaf = w;
        leaf.work_type = adm[i] ? WT_ACA : WT_DENSE;
        out[mvoff[i]] = leaf;
    }
}

__global__ void dense_flag_kernel(const work_item* __restrict__ out, int n, int* __restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = out[i].work_type == WT_DENSE ? 1 : 0;
}

__global__ void partition_kernel(const work_item* __restrict__ out, int n, const int* __restrict__ f,
                                 const int* __restrict__ dpos, const int* __restrict__ ndense_dev,
                                 work_item* __restrict__ fin, int4* __restrict__ blk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const work_item w = out[i];
    const int pos = f[i] ? dpos[i] : *ndense_dev + (i - dpos[i]);
    fin[pos] = w;
    blk[pos] = make_int4(w.set1_l, w.set1_u, w.set2_l, w.set2_u);
}

// per level: the clusters' cloud (the point's cloud of the item's first row
// cluster) had work at this level
__global__ void level_mark_kernel(const work_item* __restrict__ cur, int ncur, const int* __restrict__ cid,
                                  int level, int* __restrict__ last)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncur) return;
    const work_item w = cur[i];
    atomicMax(&last[cid[min(w.set1_l, w.set1_u)]], level);
}

// Level-by-level traversal (tree.cu:1429-1557) from the given root items; the
// leaves (node order per level) go to tr.out.  With cid != nullptr the items may
// belong to several point clouds laid side by side (index ranges into the same
// coordinate and code arrays, exact boxes only): last[c] gets the last level at
// which cloud c had items.
struct Traversal {
    DevBuf<work_item> out;
    int nout = 0;
    int levels = 0;
    int quirk_levels = 0;
};

hmg_status traverse(hipStream_t s, int dim, const Coords& c1, const Coords& c2, const uint64_t* codes1,
                    const uint64_t* codes2, const TreeParams& tp, const std::vector<work_item>& roots,
                    const int* cid, int* last, Traversal& tr)
{
    const int nmax = std::max(c1.n, c2.n);
    const Coords* cs[2] = {&c1, &c2};

    DevBuf<int> flags, rank, ub, keys, S, runid, Lc, Uc, rstart, rend, scal;
    DevBuf<double> tmin, tmax;
    HMG_TRY(flags.alloc(nmax));
    HMG_TRY(rank.alloc(nmax));
    HMG_TRY(ub.alloc(nmax));
    HMG_TRY(Lc.alloc(nmax));
    HMG_TRY(Uc.alloc(nmax));
    if (!tp.exact_boxes) {
        HMG_TRY(keys.alloc(nmax));
        HMG_TRY(S.alloc(nmax));
        HMG_TRY(runid.alloc(nmax));
        HMG_TRY(rstart.alloc(nmax));
        HMG_TRY(rend.alloc(nmax));
    }
    HMG_TRY(tmin.alloc((size_t)nmax * MAX_POINT_DIMENSION));
    HMG_TRY(tmax.alloc((size_t)nmax * MAX_POINT_DIMENSION));
    // scal: [0]=T [1]=runs [2]=total children [3]=total leaves [4]=ndense
    //       [8]=singleton flag [9]=undefined-table flag
    HMG_TRY(scal.alloc(16));
    int* dT = scal.p + 0;
    int* dR = scal.p + 1;
    int* dTotCh = scal.p + 2;
    int* dTotMv = scal.p + 3;
    int* dDiag = scal.p + 8;

    DevBuf<work_item> cur, next;
    DevBuf<int> nch, nmv, choff, mvoff, adm;
    HMG_TRY(cur.alloc(std::max<size_t>(roots.size(), 1)));
    HMG_TRY(upload_bytes(cur.p, roots.data(), sizeof(work_item) * roots.size(), s));
    int ncur = (int)roots.size();
    size_t out_cap = 0;
    tr.nout = 0;
    tr.levels = 0;
    tr.quirk_levels = 0;

    for (int level = 0; level < tp.max_level && ncur > 0; level++) {
        tr.levels++;
        if (cid) {
            level_mark_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, cid, level, last);
            HMG_HIP(hipGetLastError());
        }
        HMG_HIP(hipMemsetAsync(dDiag, 0, 2 * sizeof(int), s));
        for (int set = 0; set < 2; set++) {
            const int n = cs[set]->n;
            HMG_HIP(hipMemsetAsync(flags.p, 0, sizeof(int) * n, s));
            mark_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, set, flags.p, ub.p);
            HMG_HIP(hipGetLastError());
            HMG_TRY(scan_i32(flags.p, rank.p, n, false, dT, s));
            compact_kernel<<<grid_for(n, kT), kT, 0, s>>>(flags.p, ub.p, rank.p, n, Lc.p, Uc.p);
            HMG_HIP(hipGetLastError());
            if (tp.exact_boxes) {
                // table entry t = the t-th distinct cluster [Lc[t], Uc[t]] itself
                run_minmax_kernel<<<(unsigned)ncur, kT, 0, s>>>(*cs[set], dim, Lc.p, Uc.p, dT, dT, tmin.p,
                                                                tmax.p, dDiag);
                HMG_HIP(hipGetLastError());
                assign_box_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, set, dim, rank.p, tmin.p,
                                                                    tmax.p);
                HMG_HIP(hipGetLastError());
                continue;
            }
            HMG_HIP(hipMemsetAsync(keys.p, 0, sizeof(int) * n, s));
            keys_set_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(Lc.p, Uc.p, dT, keys.p, dDiag);
            HMG_HIP(hipGetLastError());
            HMG_TRY(scan_i32(keys.p, keys.p, n, true, nullptr, s));
            keys_fix_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(Uc.p, dT, keys.p);
            HMG_HIP(hipGetLastError());
            run_flags_kernel<<<grid_for(n, kT), kT, 0, s>>>(keys.p, n, S.p);
            HMG_HIP(hipGetLastError());
            HMG_TRY(scan_i32(S.p, runid.p, n, false, dR, s));
            run_bounds_kernel<<<grid_for(n, kT), kT, 0, s>>>(keys.p, S.p, runid.p, n, rstart.p, rend.p);
            HMG_HIP(hipGetLastError());
            run_minmax_kernel<<<(unsigned)ncur, kT, 0, s>>>(*cs[set], dim, rstart.p, rend.p, dR, dT,
                                                            tmin.p, tmax.p, dDiag);
            HMG_HIP(hipGetLastError());
            assign_box_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, set, dim, rank.p,
                                                                tmin.p, tmax.p);
            HMG_HIP(hipGetLastError());
        }
        HMG_TRY(nch.alloc(ncur));
        HMG_TRY(nmv.alloc(ncur));
        HMG_TRY(choff.alloc(ncur));
        HMG_TRY(mvoff.alloc(ncur));
        HMG_TRY(adm.alloc(ncur));
        count_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, tp.eta, level, tp.max_level,
                                                       tp.c_leaf, nch.p, nmv.p, adm.p);
        HMG_HIP(hipGetLastError());
        HMG_TRY(scan_i32(nch.p, choff.p, ncur, false, dTotCh, s));
        HMG_TRY(scan_i32(nmv.p, mvoff.p, ncur, false, dTotMv, s));
        int h[10];
        HMG_TRY(download_bytes(h, scal.p, sizeof(int) * 10, s));
        const int tot_ch = h[2], tot_mv = h[3];
        if (h[8]) tr.quirk_levels++;
        if ((size_t)(tr.nout + tot_mv) > out_cap) {
            size_t cap = std::max<size_t>(1024, out_cap);
            while (cap < (size_t)(tr.nout + tot_mv)) cap *= 2;
            DevBuf<work_item> bigger;
            HMG_TRY(bigger.alloc(cap));
            if (tr.nout > 0)
                HMG_HIP(hipMemcpyAsync(bigger.p, tr.out.p, sizeof(work_item) * tr.nout,
                                       hipMemcpyDeviceToDevice, s));
            HMG_HIP(hipStreamSynchronize(s));
            tr.out = std::move(bigger);
            out_cap = cap;
        }
        HMG_TRY(next.alloc(std::max(tot_ch, 1)));
        generate_kernel<<<grid_for(ncur, kT), kT, 0, s>>>(cur.p, ncur, nch.p, choff.p, mvoff.p, adm.p,
                                                          codes1, codes2, next.p, tr.out.p + tr.nout);
        HMG_HIP(hipGetLastError());
        tr.nout += tot_mv;
        std::swap(cur, next);
        ncur = tot_ch;
    }
    HMG_HIP(hipStreamSynchronize(s));
    return HMG_OK;
}

// key of a leaf for the partition of several clouds' block lists: cloud, then
// DENSE before ACA
__global__ void leaf_key_kernel(const work_item* __restrict__ out, int n, const int* __restrict__ cid,
                                uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                                int* __restrict__ counts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const work_item w = out[i];
    const int k = 2 * cid[w.set1_l] + (w.work_type == WT_DENSE ? 0 : 1);
    key[i] = (uint64_t)k;
    idx[i] = (uint32_t)i;
    atomicAdd(&counts[k], 1);
}

// sorted leaf q of cloud c -> its engine's block arrays at position q - first[c],
// row / column ranges back in the cloud's own indices
__global__ void leaf_scatter_kernel(const work_item* __restrict__ out, const uint32_t* __restrict__ idx,
                                    const uint64_t* __restrict__ key, int n, const int* __restrict__ first,
                                    const int* __restrict__ off, work_item* const* __restrict__ blocks,
                                    int4* const* __restrict__ blk, int4* __restrict__ all_blk)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(key[q] >> 1);
    work_item w = out[idx[q]];
    const int o = off[c];
    w.set1_l -= o;
    w.set1_u -= o;
    w.set2_l -= o;
    w.set2_u -= o;
    const int pos = q - first[c];
    const int4 b = make_int4(w.set1_l, w.set1_u, w.set2_l, w.set2_u);
    blocks[c][pos] = w;
    blk[c][pos] = b;
    all_blk[q] = b;
}

}  // namespace

hmg_status build_tree(Engine& eng, const Coords& c1, const Coords& c2, const uint64_t* codes1,
                      const uint64_t* codes2, const TreeParams& tp)
{
    hipStream_t s = eng.stream;
    work_item root{};
    root.set1_l = 0;
    root.set1_u = c1.n - 1;
    root.set2_l = 0;
    root.set2_u = c2.n - 1;
    root.level_1 = (char)tp.root_level_1;
    root.level_2 = (char)tp.root_level_2;
    Traversal tr;
    HMG_TRY(traverse(s, eng.dim, c1, c2, codes1, codes2, tp, {root}, nullptr, nullptr, tr));
    eng.levels = tr.levels;
    eng.quirk_levels = tr.quirk_levels;
    const int nout = tr.nout;

    // stable partition, DENSE first (linear_algebra.cu:4037-4051)
    eng.nblocks = nout;
    HMG_TRY(eng.blocks.alloc(std::max(nout, 1)));
    HMG_TRY(eng.blk.alloc(std::max(nout, 1)));
    int ndense = 0;
    if (nout > 0) {
        DevBuf<int> f, dpos, nd;
        HMG_TRY(f.alloc(nout));
        HMG_TRY(dpos.alloc(nout));
        HMG_TRY(nd.alloc(1));
        dense_flag_kernel<<<grid_for(nout, kT), kT, 0, s>>>(tr.out.p, nout, f.p);
        HMG_HIP(hipGetLastError());
        HMG_TRY(scan_i32(f.p, dpos.p, nout, false, nd.p, s));
        partition_kernel<<<grid_for(nout, kT), kT, 0, s>>>(tr.out.p, nout, f.p, dpos.p, nd.p,
                                                           eng.blocks.p, eng.blk.p);
        HMG_HIP(hipGetLastError());
        HMG_TRY(download_bytes(&ndense, nd.p, sizeof(int), s));
    }
    eng.ndense = ndense;
    eng.naca = nout - ndense;
    eng.h_blk.resize(nout);
    if (nout > 0) HMG_TRY(download_bytes(eng.h_blk.data(), eng.blk.p, sizeof(int4) * nout, s));
    return HMG_OK;
}

hmg_status build_trees_batched(hipStream_t s, int dim, const Coords& c, const uint64_t* codes,
                               const TreeParams& tp, const std::vector<int>& off, const std::vector<int>& n,
                               const int* cid, Engine* const* eng)
{
    const int count = (int)n.size();
    std::vector<work_item> roots((size_t)count);
    for (int q = 0; q < count; q++) {
        work_item& r = roots[q];
        r = work_item{};
        r.set1_l = r.set2_l = off[q];
        r.set1_u = r.set2_u = off[q] + n[q] - 1;
        r.level_1 = (char)tp.root_level_1;
        r.level_2 = (char)tp.root_level_2;
    }
    DevBuf<int> last;
    HMG_TRY(last.alloc(count));
    HMG_HIP(hipMemsetAsync(last.p, 0xff, sizeof(int) * count, s));  // -1: no level yet
    Traversal tr;
    HMG_TRY(traverse(s, dim, c, c, codes, codes, tp, roots, cid, last.p, tr));
    const int nout = tr.nout;
    std::vector<int> lastlv((size_t)count);
    HMG_TRY(download_bytes(lastlv.data(), last.p, sizeof(int) * count, s));
    // partition: a stable sort of the leaves by (cloud, DENSE before ACA)
    DevBuf<uint64_t> key;
    DevBuf<uint32_t> idx;
    DevBuf<int> counts;
    HMG_TRY(key.alloc(std::max(nout, 1)));
    HMG_TRY(idx.alloc(std::max(nout, 1)));
    HMG_TRY(counts.alloc(2 * count));
    HMG_HIP(hipMemsetAsync(counts.p, 0, sizeof(int) * 2 * count, s));
    if (nout > 0) {
        leaf_key_kernel<<<grid_for(nout, kT), kT, 0, s>>>(tr.out.p, nout, cid, key.p, idx.p, counts.p);
        HMG_HIP(hipGetLastError());
        int kbits = 1;
        while ((1ll << kbits) < 2ll * count) kbits++;
        HMG_TRY(radix_sort_u64(key.p, idx.p, nout, kbits, s));
    }
    std::vector<int> cnt((size_t)2 * count);
    HMG_TRY(download_bytes(cnt.data(), counts.p, sizeof(int) * 2 * count, s));
    std::vector<int> first((size_t)count);
    int run = 0;
    for (int q = 0; q < count; q++) {
        first[q] = run;
        run += cnt[2 * q] + cnt[2 * q + 1];
    }
    // each engine's block arrays (on its own stream, complete before the scatter)
    std::vector<work_item*> pb((size_t)count);
    std::vector<int4*> pk((size_t)count);
    for (int q = 0; q < count; q++) {
        Engine& e = *eng[q];
        AllocScope scope(e.stream);
        const int nb = cnt[2 * q] + cnt[2 * q + 1];
        e.nblocks = nb;
        e.ndense = cnt[2 * q];
        e.naca = cnt[2 * q + 1];
        e.levels = lastlv[q] + 1;
        e.quirk_levels = 0;
        HMG_TRY(e.blocks.alloc(std::max(nb, 1)));
        HMG_TRY(e.blk.alloc(std::max(nb, 1)));
        pb[q] = e.blocks.p;
        pk[q] = e.blk.p;
    }
    for (int q = 0; q < count; q++) HMG_HIP(hipStreamSynchronize(eng[q]->stream));
    if (nout > 0) {
        DevBuf<work_item*> dpb;
        DevBuf<int4*> dpk;
        DevBuf<int> dfirst, doff;
        DevBuf<int4> all;
        HMG_TRY(upload(dpb, pb, s));
        HMG_TRY(upload(dpk, pk, s));
        HMG_TRY(upload(dfirst, first, s));
        HMG_TRY(upload(doff, off, s));
        HMG_TRY(all.alloc(nout));
        leaf_scatter_kernel<<<grid_for(nout, kT), kT, 0, s>>>(tr.out.p, idx.p, key.p, nout, dfirst.p, doff.p,
                                                              dpb.p, dpk.p, all.p);
        HMG_HIP(hipGetLastError());
        std::vector<int4> h((size_t)nout);
        HMG_TRY(download_bytes(h.data(), all.p, sizeof(int4) * nout, s));
        for (int q = 0; q < count; q++)
            eng[q]->h_blk.assign(h.begin() + first[q], h.begin() + first[q] + eng[q]->nblocks);
    } else {
        for (int q = 0; q < count; q++) eng[q]->h_blk.clear();
    }
    return HMG_OK;
}

}  // namespace hmg
