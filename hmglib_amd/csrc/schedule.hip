// schedule.hip -- host-side layout of the H-MVP work (runs once per tree / rank).
//
// Blocks are grouped by their row cluster.  All ACA blocks of one row cluster C
// keep their U factors side by side as one column-major m1(C) x K(C) matrix, so
// the low-rank product of C is one contiguous GEMV; all DENSE blocks of C form
// one near-field sweep over the concatenated column ranges.  Work items cover 64
// rows (one wavefront) of one row cluster and write a partial row sum into P;
// every row then adds its partials in a fixed order (near items first, then
// low-rank items, each in item order), so results are deterministic.
#include <unordered_map>

#include "hmg_internal.h"

namespace hmg {

namespace {

constexpr int kSlice = kTileRows;          // rows per low-rank work item (one wavefront)
constexpr int kNearSlice = 2 * kTileRows;  // rows per near-field work item (two per lane)
constexpr int kNearChunkCols = 4096;       // near-field columns per work item
constexpr int kVChunkRows = 1024;          // points per V^T x work item

// Group blocks [b0, b1) by row cluster (l1, u1): groups in order of first
// appearance, blocks in block order inside a group; idx holds b - base.
void group_by_row(const std::vector<int4>& bl, int b0, int b1, int base, std::vector<int>& ptr,
                  std::vector<int>& idx, bool by_column = false)
{
    std::unordered_map<uint64_t, int> gid;
    gid.reserve((size_t)(b1 - b0) / 4 + 16);
    std::vector<int> g_of((size_t)(b1 - b0));
    std::vector<int> cnt;
    for (int b = b0; b < b1; b++) {
        const uint64_t key = by_column ? (((uint64_t)(uint32_t)bl[b].z << 32) | (uint32_t)bl[b].w)
                                       : (((uint64_t)(uint32_t)bl[b].x << 32) | (uint32_t)bl[b].y);
        auto it = gid.find(key);
        int g;
        if (it == gid.end()) {
            g = (int)cnt.size();
            gid.emplace(key, g);
            cnt.push_back(0);
        } else {
            g = it->second;
        }
        cnt[g]++;
        g_of[b - b0] = g;
    }
    ptr.assign(cnt.size() + 1, 0);
    for (size_t g = 0; g < cnt.size(); g++) ptr[g + 1] = ptr[g] + cnt[g];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    idx.assign((size_t)(b1 - b0), 0);
    for (int b = b0; b < b1; b++) idx[fill[g_of[b - b0]]++] = b - base;
}

template <class T>
hmg_status upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s)
{
    HMG_TRY(d.alloc(std::max<size_t>(h.size(), 1)));
    if (!h.empty())
        HMG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return HMG_OK;
}

}  // namespace

hmg_status build_groups(Engine& eng)
{
    group_by_row(eng.h_blk, 0, eng.ndense, 0, eng.h_dgrp_ptr, eng.h_dgrp_blk);
    group_by_row(eng.h_blk, eng.ndense, eng.nblocks, eng.ndense, eng.h_agrp_ptr, eng.h_agrp_blk);
    group_by_row(eng.h_blk, eng.ndense, eng.nblocks, eng.ndense, eng.h_cgrp_ptr, eng.h_cgrp_blk, true);
    // near-field items: (dense row cluster, 64-row slice, chunk of its blocks)
    std::vector<int4> items;
    const int ng = (int)eng.h_dgrp_ptr.size() - 1;
    for (int g = 0; g < ng; g++) {
        const int gb0 = eng.h_dgrp_ptr[g], gb1 = eng.h_dgrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.h_dgrp_blk[gb0]];
        const int l1 = first.x, m1 = first.y - first.x + 1;
        // chunk boundaries over the group's blocks
        std::vector<int> cut{gb0};
        int cols = 0;
        for (int q = gb0; q < gb1; q++) {
            const int4 b = eng.h_blk[eng.h_dgrp_blk[q]];
            const int m2 = b.w - b.z + 1;
            if (cols > 0 && cols + m2 > kNearChunkCols) {
                cut.push_back(q);
                cols = 0;
            }
            cols += m2;
        }
        cut.push_back(gb1);
        for (int r0 = 0; r0 < m1; r0 += kNearSlice)
            for (size_t c = 0; c + 1 < cut.size(); c++)
                items.push_back(make_int4(l1 + r0, std::min(kNearSlice, m1 - r0), cut[c], cut[c + 1]));
    }
    eng.n_near_items = (int)items.size();
    HMG_TRY(upload(eng.near_items, items, eng.stream));
    HMG_TRY(upload(eng.near_blk, eng.h_dgrp_blk, eng.stream));
    HMG_HIP(hipStreamSynchronize(eng.stream));
    return HMG_OK;
}

hmg_status aca_layout(Engine& eng, int k)
{
    const int na = eng.naca;
    eng.h_kb.assign(na, 0);
    eng.kmax = 0;
    for (int a = 0; a < na; a++) {
        const int4 b = eng.h_blk[eng.ndense + a];
        const int m1 = b.y - b.x + 1, m2 = b.w - b.z + 1;
        int kb = k;
        if (kb > m1) kb = m1;
        if (kb > m2) kb = m2;  // set_k_per_item (linear_algebra.cu:769-787)
        if (kb < 0) kb = 0;
        eng.h_kb[a] = kb;
        eng.kmax = std::max(eng.kmax, kb);
    }
    // U and T in row-cluster (U) order
    eng.h_uoff.assign(na + 1, 0);
    eng.h_toff.assign(na + 1, 0);
    std::vector<LrItem> items;
    int64_t urun = 0, trun = 0;
    const int ng = (int)eng.h_agrp_ptr.size() - 1;
    for (int g = 0; g < ng; g++) {
        const int gb0 = eng.h_agrp_ptr[g], gb1 = eng.h_agrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.ndense + eng.h_agrp_blk[gb0]];
        const int l1 = first.x, m1 = first.y - first.x + 1;
        const int64_t uC = urun, tC = trun;
        int K = 0;
        for (int q = gb0; q < gb1; q++) {
            const int a = eng.h_agrp_blk[q];
            eng.h_uoff[a] = urun;
            eng.h_toff[a] = trun;
            urun += (int64_t)m1 * eng.h_kb[a];
            trun += eng.h_kb[a];
            K += eng.h_kb[a];
        }
        if (K == 0) continue;
        for (int r0 = 0; r0 < m1; r0 += kSlice)
            items.push_back(LrItem{l1 + r0, std::min(kSlice, m1 - r0), m1, K, uC + r0, tC});
    }
    eng.h_uoff[na] = urun;
    eng.h_toff[na] = trun;
    eng.tsize = trun;
    eng.n_lr_items = (int)items.size();

    // V in column-cluster order, each cluster's V row-major (m2 x K); V^T x items
    // cover 64 columns x up to kVChunkRows points, slices reduce their row chunks
    eng.h_voff.assign(na + 1, 0);
    eng.h_ldv.assign(std::max(na, 1), 0);
    std::vector<LvItem> vitems;
    std::vector<LvSlice> slices;
    std::vector<int> tmap((size_t)trun);
    int64_t vrun = 0;
    int64_t col = 0;
    const int ncg = (int)eng.h_cgrp_ptr.size() - 1;
    for (int g = 0; g < ncg; g++) {
        const int gb0 = eng.h_cgrp_ptr[g], gb1 = eng.h_cgrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.ndense + eng.h_cgrp_blk[gb0]];
        const int l2 = first.z, m2 = first.w - first.z + 1;
        int K = 0;
        for (int q = gb0; q < gb1; q++) K += eng.h_kb[eng.h_cgrp_blk[q]];
        const int64_t vD = vrun;
        int pos = 0;
        for (int q = gb0; q < gb1; q++) {
            const int a = eng.h_cgrp_blk[q];
            eng.h_voff[a] = vD + pos;
            eng.h_ldv[a] = K;
            for (int l = 0; l < eng.h_kb[a]; l++) tmap[col + pos + l] = (int)(eng.h_toff[a] + l);
            pos += eng.h_kb[a];
        }
        vrun += (int64_t)m2 * K;
        for (int s0 = 0; s0 < K; s0 += kSlice) {
            const int nc = std::min(kSlice, K - s0);
            const int item0 = (int)vitems.size();
            int nch = 0;
            for (int j0 = 0; j0 < m2; j0 += kVChunkRows, nch++)
                vitems.push_back(LvItem{l2 + j0, std::min(kVChunkRows, m2 - j0), K, nc,
                                        vD + s0 + (int64_t)j0 * K});
            slices.push_back(LvSlice{item0, nch, nc, (int)(col + s0)});
        }
        col += K;
    }
    eng.h_voff[na] = vrun;
    eng.n_lv_items = (int)vitems.size();
    eng.n_lv_slices = (int)slices.size();

    // per-row partial lists: near item q owns P slots q*128 .. q*128+127, low-rank
    // item q owns n_near*128 + q*64 .. +63
    const int n1 = eng.n[0];
    std::vector<int> ptr((size_t)n1 + 1, 0);
    std::vector<int4> near_items((size_t)eng.n_near_items);
    if (eng.n_near_items > 0)
        HMG_HIP(hipMemcpy(near_items.data(), eng.near_items.p, sizeof(int4) * eng.n_near_items,
                          hipMemcpyDeviceToHost));
    for (const int4& it : near_items)
        for (int r = 0; r < it.y; r++) ptr[it.x + r + 1]++;
    for (const LrItem& it : items)
        for (int r = 0; r < it.nrows; r++) ptr[it.row0 + r + 1]++;
    for (int i = 0; i < n1; i++) ptr[i + 1] += ptr[i];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    std::vector<int> idx((size_t)ptr[n1]);
    for (int q = 0; q < eng.n_near_items; q++) {
        const int4& it = near_items[q];
        for (int r = 0; r < it.y; r++) idx[fill[it.x + r]++] = q * kNearSlice + r;
    }
    for (int q = 0; q < eng.n_lr_items; q++) {
        const LrItem& it = items[q];
        for (int r = 0; r < it.nrows; r++)
            idx[fill[it.row0 + r]++] = eng.n_near_items * kNearSlice + q * kSlice + r;
    }

    hipStream_t s = eng.stream;
    HMG_TRY(eng.kb.alloc(std::max(na, 1)));
    HMG_TRY(eng.keff.alloc(std::max(na, 1)));
    if (na > 0) {
        HMG_HIP(hipMemcpyAsync(eng.kb.p, eng.h_kb.data(), sizeof(int) * na, hipMemcpyHostToDevice, s));
        HMG_HIP(hipMemsetAsync(eng.keff.p, 0, sizeof(int) * na, s));
    }
    HMG_TRY(upload(eng.uoff, eng.h_uoff, s));
    HMG_TRY(upload(eng.voff, eng.h_voff, s));
    HMG_TRY(upload(eng.toff, eng.h_toff, s));
    HMG_TRY(upload(eng.lr_items, items, s));
    HMG_TRY(upload(eng.ldv, eng.h_ldv, s));
    HMG_TRY(upload(eng.lv_items, vitems, s));
    HMG_TRY(upload(eng.lv_slices, slices, s));
    HMG_TRY(upload(eng.tmap, tmap, s));
    HMG_TRY(eng.Tp.alloc((size_t)std::max(eng.n_lv_items, 1) * kSlice));
    HMG_TRY(upload(eng.row_ptr, ptr, s));
    HMG_TRY(upload(eng.row_idx, idx, s));
    HMG_TRY(eng.P.alloc((size_t)eng.n_near_items * kNearSlice + (size_t)eng.n_lr_items * kSlice));
    HMG_TRY(eng.T.alloc(std::max<int64_t>(eng.tsize, 1)));
    HMG_HIP(hipStreamSynchronize(s));
    eng.aca_k = k;
    return HMG_OK;
}

}  // namespace hmg
