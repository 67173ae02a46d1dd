// schedule.hip -- host-side layout of the H-MVP work (runs once per tree / rank).
//
// Blocks are grouped by their row cluster.  All ACA blocks of one row cluster C
// keep their U factors side by side as one column-major m1(C) x K(C) matrix, so
// the low-rank product of C is one contiguous GEMV; all DENSE blocks of C form
// one near-field sweep over the concatenated column ranges.  Work items write
// partial sums into P; every row then adds its partials in a fixed order (near
// items, then low-rank items, then symmetric column slots), so results are
// deterministic.
//
// When both point sets are the same points, K(C,D) = K(D,C)^T bitwise for
// disjoint clusters C, D (the distance is computed from coordinate differences
// whose squares do not depend on the order).  The matrix-free near field then
// evaluates each transposed pair of DENSE blocks once: the item of (C,D) with
// C before D produces the row sums of C and the column sums of D; (D,C) is not
// evaluated.  Diagonal and overlapping blocks keep the plain row-sum form (the
// regularisation lives there).
#include <algorithm>
#include <map>
#include <unordered_map>
#include <unordered_set>

#include "hmg_internal.h"

namespace hmg {

namespace {

constexpr int kSlice = kTileRows;          // rows per low-rank work item (one wavefront)
constexpr int kNearSlice = 2 * kTileRows;  // rows per near-field work item (two per lane)
constexpr int kNearChunkCols = 4096;       // near-field columns per work item
constexpr int kVChunkRows = 1024;          // points per V^T x work item (FP64)
constexpr int kMxVChunkRows = 4096;        // points per V^T x work item (mixed precision)
constexpr int kMxRows = kMxNearRows;       // rows per mixed-precision near-field item
constexpr int kMxChunkCols = 4096;         // columns per mixed-precision near-field item

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

inline uint64_t range_key(int l, int u) { return ((uint64_t)(uint32_t)l << 32) | (uint32_t)u; }

struct PairKey {
    uint64_t a, b;
    bool operator==(const PairKey& o) const { return a == o.a && b == o.b; }
};
struct PairHash {
    size_t operator()(const PairKey& k) const
    {
        return (size_t)(k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull + (k.a << 6)));
    }
};

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
        const uint64_t key = by_column ? range_key(bl[b].z, bl[b].w) : range_key(bl[b].x, bl[b].y);
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

// cut a run of column counts into chunks of at most kNearChunkCols columns
std::vector<int> column_chunks(const std::vector<int>& m2s)
{
    std::vector<int> cut{0};
    int cols = 0;
    for (int q = 0; q < (int)m2s.size(); q++) {
        if (cols > 0 && cols + m2s[q] > kNearChunkCols) {
            cut.push_back(q);
            cols = 0;
        }
        cols += m2s[q];
    }
    cut.push_back((int)m2s.size());
    return cut;
}

// Per-row partial lists of one schedule.  Near items own P slots
// near0 + q*128 + r, low-rank items q*64 + r: one list entry each.  Symmetric
// column slots come per column cluster as (l2, m2, base, count) (contribution k
// of point l2 + j at base + k*m2 + j): one list entry -(g+1) per group, which
// the combine expands, so neighbouring points read neighbouring slots.
hmg_status build_combine(Combine& cb, int n1, const std::vector<int4>& near_rows, int64_t near0,
                         const std::vector<LrItem>& lr, const std::vector<ColSlots>& cols,
                         hipStream_t s, int near_slice = kNearSlice, int lr_slice = kSlice)
{
    std::vector<int> ptr((size_t)n1 + 1, 0);
    for (const int4& it : near_rows)
        for (int r = 0; r < it.y; r++) ptr[it.x + r + 1]++;
    for (const LrItem& it : lr)
        for (int r = 0; r < it.nrows; r++) ptr[it.row0 + r + 1]++;
    for (const ColSlots& g : cols)
        for (int j = 0; j < g.m2; j++) ptr[g.l2 + j + 1]++;
    for (int i = 0; i < n1; i++) ptr[i + 1] += ptr[i];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    std::vector<int> idx((size_t)ptr[n1]);
    for (size_t q = 0; q < near_rows.size(); q++) {
        const int4& it = near_rows[q];
        for (int r = 0; r < it.y; r++) idx[fill[it.x + r]++] = (int)(near0 + (int64_t)q * near_slice + r);
    }
    for (size_t q = 0; q < lr.size(); q++) {
        const LrItem& it = lr[q];
        for (int r = 0; r < it.nrows; r++) idx[fill[it.row0 + r]++] = (int)(q * lr_slice + r);
    }
    for (size_t g = 0; g < cols.size(); g++)
        for (int j = 0; j < cols[g].m2; j++) idx[fill[cols[g].l2 + j]++] = -(int)g - 1;
    HMG_TRY(upload(cb.ptr, ptr, s));
    HMG_TRY(upload(cb.idx, idx, s));
    HMG_TRY(upload(cb.groups, cols, s));
    cb.total = ptr[n1];
    cb.slots = 0;
    cb.group_entries = 0;
    for (const ColSlots& g : cols) {
        cb.slots += (int64_t)g.m2 * g.count;
        cb.group_entries += g.m2;
    }
    return HMG_OK;
}

// Mixed-precision near field: per dense row cluster, the column ranges of its
// DENSE blocks sorted by start and merged where contiguous, cut into pieces of at
// most kMxChunkCols columns; each piece becomes a run of 8-point groups (an even
// number of groups, masked outside the piece).  Items are 128-row slices of the
// row cluster against consecutive pieces of at most kMxChunkCols columns.
void build_mixed_near(Engine& eng)
{
    eng.h_mx_items.clear();
    eng.h_mx_refs.clear();
    eng.h_mx_wg.clear();
    eng.mx_evals = 0;
    const int ng = (int)eng.h_dgrp_ptr.size() - 1;
    for (int g = 0; g < ng; g++) {
        const int gb0 = eng.h_dgrp_ptr[g], gb1 = eng.h_dgrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.h_dgrp_blk[gb0]];
        const int l1 = first.x, m1 = first.y - first.x + 1;
        std::vector<std::pair<int, int>> rng;  // [lo, hi)
        for (int q = gb0; q < gb1; q++) {
            const int4 b = eng.h_blk[eng.h_dgrp_blk[q]];
            rng.emplace_back(b.z, b.w + 1);
        }
        std::sort(rng.begin(), rng.end());
        std::vector<std::pair<int, int>> merged;
        for (const auto& r : rng) {
            if (!merged.empty() && r.first <= merged.back().second)
                merged.back().second = std::max(merged.back().second, r.second);
            else
                merged.push_back(r);
        }
        std::vector<MxRef> pieces;
        for (const auto& r : merged) {
            for (int lo = r.first; lo < r.second; lo += kMxChunkCols) {
                const int hi = std::min(r.second, lo + kMxChunkCols);
                const int g0 = lo / 8;
                const int gend = (hi + 7) / 8;
                const int g1 = g0 + (int)round_up(gend - g0, 2);
                pieces.push_back(MxRef{g0, g1, lo, hi});
            }
        }
        // items: consecutive pieces up to kMxChunkCols evaluated columns
        size_t p0 = 0;
        while (p0 < pieces.size()) {
            size_t p1 = p0;
            int cols = 0;
            while (p1 < pieces.size()) {
                const int c = (pieces[p1].g1 - pieces[p1].g0) * 8;
                if (p1 > p0 && cols + c > kMxChunkCols) break;
                cols += c;
                p1++;
            }
            const int ref0 = (int)eng.h_mx_refs.size();
            eng.h_mx_refs.insert(eng.h_mx_refs.end(), pieces.begin() + p0, pieces.begin() + p1);
            const int ref1 = (int)eng.h_mx_refs.size();
            const int q0 = (int)eng.h_mx_items.size();
            for (int r0 = 0; r0 < m1; r0 += kMxRows) {
                const int nr = std::min(kMxRows, m1 - r0);
                eng.h_mx_items.push_back(MxItem{l1 + r0, nr, ref0, ref1});
                eng.mx_evals += round_up(nr, 32) * cols;  // whole 32-row tiles are evaluated
            }
            // workgroup form: the row slices of this chunk kNwWaves at a time (they
            // share its column groups); -1 = an idle wavefront
            const int q1 = (int)eng.h_mx_items.size();
            for (int q = q0; q < q1; q += kNwWaves)
                for (int w = 0; w < kNwWaves; w++) eng.h_mx_wg.push_back(q + w < q1 ? q + w : -1);
            p0 = p1;
        }
    }
    eng.n_mx_items = (int)eng.h_mx_items.size();
}

}  // namespace

hmg_status build_groups(Engine& eng)
{
    group_by_row(eng.h_blk, 0, eng.ndense, 0, eng.h_dgrp_ptr, eng.h_dgrp_blk);
    group_by_row(eng.h_blk, eng.ndense, eng.nblocks, eng.ndense, eng.h_agrp_ptr, eng.h_agrp_blk);
    group_by_row(eng.h_blk, eng.ndense, eng.nblocks, eng.ndense, eng.h_cgrp_ptr, eng.h_cgrp_blk, true);
    const int ng = (int)eng.h_dgrp_ptr.size() - 1;

    // stored near field: (dense row cluster, 128-row slice, chunk of its blocks)
    std::vector<int4>& items = eng.h_near_items;
    items.clear();
    for (int g = 0; g < ng; g++) {
        const int gb0 = eng.h_dgrp_ptr[g], gb1 = eng.h_dgrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.h_dgrp_blk[gb0]];
        const int l1 = first.x, m1 = first.y - first.x + 1;
        std::vector<int> m2s;
        for (int q = gb0; q < gb1; q++) {
            const int4 b = eng.h_blk[eng.h_dgrp_blk[q]];
            m2s.push_back(b.w - b.z + 1);
        }
        const std::vector<int> cut = column_chunks(m2s);
        for (int r0 = 0; r0 < m1; r0 += kNearSlice)
            for (size_t c = 0; c + 1 < cut.size(); c++)
                items.push_back(make_int4(l1 + r0, std::min(kNearSlice, m1 - r0), gb0 + cut[c], gb0 + cut[c + 1]));
    }
    eng.n_near_items = (int)items.size();
    HMG_TRY(upload(eng.near_items, items, eng.stream));
    HMG_TRY(upload(eng.near_blk, eng.h_dgrp_blk, eng.stream));

    // matrix-free near field: transposed pairs when both sides are the same points
    std::unordered_set<PairKey, PairHash> dense_set;
    eng.near_sym = eng.same_sets && eng.near_sym_enabled;
    if (eng.near_sym) {
        dense_set.reserve((size_t)eng.ndense * 2 + 16);
        for (int b = 0; b < eng.ndense; b++) {
            const int4 bl = eng.h_blk[b];
            dense_set.insert(PairKey{range_key(bl.x, bl.y), range_key(bl.z, bl.w)});
        }
    }
    eng.h_fly_items.clear();
    eng.h_fly_refs.clear();
    eng.near_evals = 0;
    // kind: 0 = row sums only, 1 = symmetric (row + column sums), 2 = covered by its transpose
    auto kind_of = [&](const int4& b) {
        if (!eng.near_sym) return 0;
        const bool disjoint = b.y < b.z || b.w < b.x;
        if (!disjoint) return 0;
        if (!dense_set.count(PairKey{range_key(b.z, b.w), range_key(b.x, b.y)})) return 0;
        return b.x < b.z ? 1 : 2;
    };
    std::vector<std::pair<int, int>> sym_refs;  // (fly ref index, dense block) awaiting a column slot
    for (int g = 0; g < ng; g++) {
        const int gb0 = eng.h_dgrp_ptr[g], gb1 = eng.h_dgrp_ptr[g + 1];
        const int4 first = eng.h_blk[eng.h_dgrp_blk[gb0]];
        const int l1 = first.x, m1 = first.y - first.x + 1;
        std::vector<int> plain, sym;
        for (int q = gb0; q < gb1; q++) {
            const int b = eng.h_dgrp_blk[q];
            const int k = kind_of(eng.h_blk[b]);
            if (k == 0) plain.push_back(b);
            else if (k == 1) sym.push_back(b);
        }
        std::vector<int> order(plain);
        order.insert(order.end(), sym.begin(), sym.end());
        if (order.empty()) continue;
        std::vector<int> m2s;
        for (int b : order) m2s.push_back(eng.h_blk[b].w - eng.h_blk[b].z + 1);
        const std::vector<int> cut = column_chunks(m2s);
        const int nplain = (int)plain.size();
        for (int r0 = 0; r0 < m1; r0 += kNearSlice) {
            for (size_t c = 0; c + 1 < cut.size(); c++) {
                NearItem it;
                it.row0 = l1 + r0;
                it.nrows = std::min(kNearSlice, m1 - r0);
                it.ref0 = (int)eng.h_fly_refs.size();
                it.ref_sym = -1;
                it.ncols_sym = 0;
                for (int q = cut[c]; q < cut[c + 1]; q++) {
                    if (q >= nplain && it.ref_sym < 0) it.ref_sym = (int)eng.h_fly_refs.size();
                    const int4 b = eng.h_blk[order[q]];
                    if (q >= nplain) {
                        it.ncols_sym += b.w - b.z + 1;
                        sym_refs.emplace_back((int)eng.h_fly_refs.size(), order[q]);
                    }
                    eng.h_fly_refs.push_back(NearRef{b.z, b.w - b.z + 1, -1});
                }
                it.ref1 = (int)eng.h_fly_refs.size();
                if (it.ref_sym < 0) it.ref_sym = it.ref1;
                int cols = 0;
                for (int q = cut[c]; q < cut[c + 1]; q++) cols += m2s[q];
                eng.near_evals += (int64_t)it.nrows * cols;
                eng.h_fly_items.push_back(it);
            }
        }
    }
    eng.n_fly_items = (int)eng.h_fly_items.size();
    // column slots grouped by column cluster; contributions in item order.  Slot
    // bases are relative to the column region of P (resolved by aca_layout).
    std::map<uint64_t, std::vector<int>> by_col;  // column cluster -> refs in item order
    for (const auto& pr : sym_refs) {
        const int4 b = eng.h_blk[pr.second];
        by_col[range_key(b.z, b.w)].push_back(pr.first);
    }
    int64_t run = 0;
    eng.h_col_groups.clear();
    for (const auto& kv : by_col) {
        const int l2 = (int)(kv.first >> 32), u2 = (int)(uint32_t)kv.first;
        const int m2 = u2 - l2 + 1;
        for (size_t k = 0; k < kv.second.size(); k++)
            eng.h_fly_refs[kv.second[k]].cbase = run + (int64_t)k * m2;
        eng.h_col_groups.push_back(ColSlots{l2, m2, run, (int)kv.second.size()});
        run += (int64_t)m2 * (int64_t)kv.second.size();
    }
    eng.n_col_slots = run;
    HMG_HIP(hipStreamSynchronize(eng.stream));
    return HMG_OK;
}

hmg_status aca_layout(Engine& eng, int k, int prec)
{
    const int na = eng.naca;
    const bool mixed = prec == HMG_PREC_MIXED;
    // FP64 leading dims in elements: multiples of 128 bytes
    const int kPad = 16;
    int ntp = 0;
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
    eng.h_ldu.assign(std::max(na, 1), 0);
    eng.h_voff.assign(na + 1, 0);
    eng.h_ldv.assign(std::max(na, 1), 0);
    std::vector<LrItem> items;
    std::vector<LvItem> vitems;
    std::vector<LvSlice> slices;
    std::vector<int> tmap;
    const int lrows = mixed ? kMxLrRows : kSlice;
    // interleaved 32-wide MFMA tiles an item of n rows or columns uses
    auto tiles_of = [](int n) { return n <= 32 ? 1 : n <= 64 ? 2 : 4; };
    for (int g = 0; g < 4; g++) eng.lr_nt_off[g] = eng.lv_nt_off[g] = 0;
    if (!mixed) {
        int64_t urun = 0, trun = 0;
        const int ng = (int)eng.h_agrp_ptr.size() - 1;
        for (int g = 0; g < ng; g++) {
            const int gb0 = eng.h_agrp_ptr[g], gb1 = eng.h_agrp_ptr[g + 1];
            const int4 first = eng.h_blk[eng.ndense + eng.h_agrp_blk[gb0]];
            const int l1 = first.x, m1 = first.y - first.x + 1;
            const int ldu = (int)round_up(m1, kPad);
            const int64_t uC = urun, tC = trun;
            int K = 0;
            for (int q = gb0; q < gb1; q++) {
                const int a = eng.h_agrp_blk[q];
                eng.h_uoff[a] = uC + (int64_t)K * ldu;
                eng.h_toff[a] = trun;
                eng.h_ldu[a] = ldu;
                trun += eng.h_kb[a];
                K += eng.h_kb[a];
            }
            urun += (int64_t)ldu * K;
            if (K == 0) continue;
            for (int r0 = 0; r0 < m1; r0 += lrows)
                items.push_back(LrItem{l1 + r0, std::min(lrows, m1 - r0), ldu, K, uC + r0, tC});
        }
        eng.h_uoff[na] = urun;
        eng.h_toff[na] = trun;
        eng.tsize = trun;

        // V in column-cluster order, each cluster's V row-major (m2 x K, ld = K
        // padded); V^T x items cover 64 columns x up to kVChunkRows points, slices
        // reduce their row chunks
        tmap.assign((size_t)trun, 0);
        int64_t vrun = 0;
        int64_t col = 0;
        const int ncg = (int)eng.h_cgrp_ptr.size() - 1;
        for (int g = 0; g < ncg; g++) {
            const int gb0 = eng.h_cgrp_ptr[g], gb1 = eng.h_cgrp_ptr[g + 1];
            const int4 first = eng.h_blk[eng.ndense + eng.h_cgrp_blk[gb0]];
            const int l2 = first.z, m2 = first.w - first.z + 1;
            int K = 0;
            for (int q = gb0; q < gb1; q++) K += eng.h_kb[eng.h_cgrp_blk[q]];
            const int ld = (int)round_up(K, kPad);
            const int64_t vD = vrun;
            int pos = 0;
            for (int q = gb0; q < gb1; q++) {
                const int a = eng.h_cgrp_blk[q];
                eng.h_voff[a] = vD + pos;
                eng.h_ldv[a] = ld;
                for (int l = 0; l < eng.h_kb[a]; l++) tmap[col + pos + l] = (int)(eng.h_toff[a] + l);
                pos += eng.h_kb[a];
            }
            vrun += (int64_t)m2 * ld;
            const int nch = (m2 + kVChunkRows - 1) / kVChunkRows;
            for (int s0 = 0; s0 < K; s0 += kSlice) {
                const int nc = std::min(kSlice, K - s0);
                const int tp0 = (int)vitems.size();
                for (int c = 0; c < nch; c++) {
                    const int j0 = c * kVChunkRows;
                    vitems.push_back(LvItem{l2 + j0, std::min(kVChunkRows, m2 - j0), ld, nc,
                                            vD + s0 + (int64_t)j0 * ld, (int)(col + s0), tp0 + c});
                }
                slices.push_back(LvSlice{tp0, nch, nc, (int)(col + s0)});
            }
            col += K;
        }
        eng.h_voff[na] = vrun;
    } else {
        // BF16 hi + lo panels (layout described with MxPanel).  U panels and T slots
        // per row cluster; T slots of a cluster padded to 32 (zero slots)
        eng.h_mxp.assign(std::max(na, 1), MxPanel{});
        eng.h_vw_items.clear();
        eng.h_uw_items.clear();
        int64_t urun = 0, trun = 0;
        const int ng = (int)eng.h_agrp_ptr.size() - 1;
        for (int g = 0; g < ng; g++) {
            const int gb0 = eng.h_agrp_ptr[g], gb1 = eng.h_agrp_ptr[g + 1];
            const int4 first = eng.h_blk[eng.ndense + eng.h_agrp_blk[gb0]];
            const int l1 = first.x, m1 = first.y - first.x + 1;
            const int m1p = (int)round_up(m1, 32);
            const int64_t uC = urun, tC = trun;
            int K = 0;
            for (int q = gb0; q < gb1; q++) {
                const int a = eng.h_agrp_blk[q];
                MxPanel& mp = eng.h_mxp[a];
                mp.ubase = uC;
                mp.m1p = m1p;
                mp.koff = K;
                eng.h_uoff[a] = uC;
                eng.h_toff[a] = tC + K;
                K += eng.h_kb[a];
            }
            // whole pairs of 16-column reduction steps
            const int K32 = (int)round_up(K, 32);
            urun += (int64_t)K32 * m1p * 2;
            trun += K32;
            if (K == 0) continue;
            for (int r0 = 0; r0 < m1; r0 += lrows)
                items.push_back(LrItem{l1 + r0, std::min(lrows, m1 - r0), m1p, K32, uC + (int64_t)r0 * 8, tC / 8});
        }
        eng.h_uoff[na] = urun;
        eng.h_toff[na] = trun;
        eng.tsize = trun;
        eng.up_size = urun + kMxPanelSlack;

        // V panels per column cluster; V^T X items: 128 columns x up to 512 point
        // groups (4096 points); a slice with one chunk writes T itself
        tmap.assign((size_t)std::max<int64_t>(trun, 1), 0);
        int64_t vrun = 0;
        int64_t col = 0;
        const int vgroups = kMxVChunkRows / 8;
        const int ncg = (int)eng.h_cgrp_ptr.size() - 1;
        for (int g = 0; g < ncg; g++) {
            const int gb0 = eng.h_cgrp_ptr[g], gb1 = eng.h_cgrp_ptr[g + 1];
            const int4 first = eng.h_blk[eng.ndense + eng.h_cgrp_blk[gb0]];
            const int l2 = first.z, m2 = first.w - first.z + 1;
            int K = 0;
            for (int q = gb0; q < gb1; q++) K += eng.h_kb[eng.h_cgrp_blk[q]];
            const int Kp = (int)round_up(K, 32);
            const int glo = l2 / 8, ghi = (l2 + m2 - 1) / 8;
            // whole pairs of 16-point reduction steps
            const int ngp = (int)round_up(ghi - glo + 1, 4);
            const int64_t vD = vrun;
            int pos = 0;
            for (int q = gb0; q < gb1; q++) {
                const int a = eng.h_cgrp_blk[q];
                MxPanel& mp = eng.h_mxp[a];
                mp.vbase = vD;
                mp.kp = Kp;
                mp.pos = pos;
                mp.glo = glo;
                eng.h_voff[a] = vD;
                for (int l = 0; l < eng.h_kb[a]; l++) tmap[col + pos + l] = (int)(eng.h_toff[a] + l);
                pos += eng.h_kb[a];
            }
            vrun += (int64_t)ngp * 16 * Kp;
            const int nch = (ngp + vgroups - 1) / vgroups;
            const size_t first_item = vitems.size();
            for (int s0 = 0; s0 < K; s0 += kMxVCols) {
                const int nc = std::min(kMxVCols, K - s0);
                const bool direct = nch == 1;
                const int tp0 = direct ? -1 : ntp;
                for (int c = 0; c < nch; c++) {
                    const int gg0 = c * vgroups;
                    vitems.push_back(LvItem{glo + gg0, std::min(vgroups, ngp - gg0), Kp, nc,
                                            vD + (int64_t)gg0 * 16 * Kp + (int64_t)s0 * 8, (int)(col + s0),
                                            direct ? -1 : tp0 + c});
                }
                if (!direct) {
                    slices.push_back(LvSlice{tp0, nch, nc, (int)(col + s0)});
                    ntp += nch;
                }
            }
            // workgroup items: the slices of one point chunk, kVwWaves at a time
            const int nsl = (K + kMxVCols - 1) / kMxVCols;
            for (int c = 0; c < nch; c++) {
                for (int q0 = 0; q0 < nsl; q0 += kVwWaves) {
                    const LvItem& lead = vitems[first_item + (size_t)q0 * nch + c];
                    for (int w = 0; w < kVwWaves; w++) {
                        if (q0 + w < nsl) {
                            const LvItem& it = vitems[first_item + (size_t)(q0 + w) * nch + c];
                            eng.h_vw_items.push_back(WgItem{it.vbase, it.ld, it.ncols, it.j0, it.nrows, it.tcol0, it.tp});
                        } else {
                            eng.h_vw_items.push_back(WgItem{0, lead.ld, 0, lead.j0, lead.nrows, 0, -1});
                        }
                    }
                }
            }
            col += K;
        }
        eng.h_voff[na] = vrun;
        eng.vp_size = vrun + kMxPanelSlack;
        // one kernel instance per tile count: items grouped 4, 2, 1 (each item
        // carries its own rows and output slots)
        std::vector<LrItem> grouped;
        grouped.reserve(items.size());
        int gi = 0;
        for (int nt : {4, 2, 1}) {
            eng.lr_nt_off[gi++] = (int)grouped.size();
            for (const LrItem& it : items)
                if (tiles_of(it.nrows) == nt) grouped.push_back(it);
        }
        eng.lr_nt_off[3] = (int)grouped.size();
        items.swap(grouped);
        // U T workgroup items: the row slices of one row cluster (they share its T
        // slots), kUwWaves at a time; partial-row slots as the grouped items'
        {
            std::unordered_map<int64_t, std::vector<int>> by_t;
            std::vector<int64_t> t_order;
            for (int q = 0; q < (int)items.size(); q++) {
                auto it = by_t.find(items[q].toff);
                if (it == by_t.end()) {
                    t_order.push_back(items[q].toff);
                    by_t[items[q].toff].push_back(q);
                } else {
                    it->second.push_back(q);
                }
            }
            std::sort(t_order.begin(), t_order.end());
            for (int64_t t : t_order) {
                std::vector<int>& qs = by_t[t];
                std::sort(qs.begin(), qs.end(), [&](int a, int b) { return items[a].row0 < items[b].row0; });
                for (size_t q0 = 0; q0 < qs.size(); q0 += kUwWaves) {
                    const LrItem& lead = items[qs[q0]];
                    for (int w = 0; w < kUwWaves; w++) {
                        if (q0 + w < qs.size()) {
                            const LrItem& it = items[qs[q0 + w]];
                            eng.h_uw_items.push_back(
                                WgItem{it.uoff, it.ld, it.nrows, (int)it.toff, it.K / 8, qs[q0 + w], -1});
                        } else {
                            eng.h_uw_items.push_back(WgItem{0, lead.ld, 0, (int)lead.toff, lead.K / 8, 0, -1});
                        }
                    }
                }
            }
        }
        std::vector<LvItem> vgrouped;
        vgrouped.reserve(vitems.size());
        gi = 0;
        for (int nt : {4, 2, 1}) {
            eng.lv_nt_off[gi++] = (int)vgrouped.size();
            for (const LvItem& it : vitems)
                if (tiles_of(it.ncols) == nt) vgrouped.push_back(it);
        }
        eng.lv_nt_off[3] = (int)vgrouped.size();
        vitems.swap(vgrouped);
        eng.tpan_ready = false;
    }
    eng.n_lr_items = (int)items.size();
    eng.n_lv_items = (int)vitems.size();
    eng.n_lv_slices = (int)slices.size();
    eng.n_tp = mixed ? ntp : eng.n_lv_items;

    // FP64: P = [low-rank items x 64][near items x 128][symmetric column slots];
    // mixed precision: P32 = [low-rank items x 128][near items x 64], kMxRhs values
    // per slot (allocated by the product)
    const int64_t near0 = (int64_t)eng.n_lr_items * lrows;
    hipStream_t s = eng.stream;
    const int n1 = eng.n[0];
    int64_t col0 = 0;
    if (!mixed) {
        col0 = near0 + (int64_t)std::max(eng.n_near_items, eng.n_fly_items) * kNearSlice;
        if (col0 + eng.n_col_slots > (int64_t)INT32_MAX) {
            set_error("MVP partial buffer exceeds 2^31 slots");
            return HMG_ERR_INVALID;
        }
        HMG_TRY(build_combine(eng.comb_stored, n1, eng.h_near_items, near0, items, {}, s));
        std::vector<int4> fly_rows((size_t)eng.n_fly_items);
        for (int q = 0; q < eng.n_fly_items; q++)
            fly_rows[q] = make_int4(eng.h_fly_items[q].row0, eng.h_fly_items[q].nrows, 0, 0);
        std::vector<ColSlots> cols(eng.h_col_groups);
        for (ColSlots& c : cols) c.base += col0;
        HMG_TRY(build_combine(eng.comb_fly, n1, fly_rows, near0, items, cols, s));
        std::vector<NearRef> refs(eng.h_fly_refs);
        for (NearRef& r : refs)
            if (r.cbase >= 0) r.cbase += col0;
        HMG_TRY(upload(eng.fly_items, eng.h_fly_items, s));
        HMG_TRY(upload(eng.fly_refs, refs, s));
    } else {
        if (!eng.mx_sched) {
            build_mixed_near(eng);
            HMG_TRY(upload(eng.mx_items, eng.h_mx_items, s));
            HMG_TRY(upload(eng.mx_wg, eng.h_mx_wg, s));
            HMG_TRY(upload(eng.mx_refs, eng.h_mx_refs, s));
            eng.mx_sched = true;
        }
        std::vector<int4> mx_rows((size_t)eng.n_mx_items);
        for (int q = 0; q < eng.n_mx_items; q++)
            mx_rows[q] = make_int4(eng.h_mx_items[q].row0, eng.h_mx_items[q].nrows, 0, 0);
        const int64_t mslots = near0 + (int64_t)eng.n_mx_items * kMxRows;
        if (mslots > (int64_t)INT32_MAX) {
            set_error("mixed-precision partial buffer too large");
            return HMG_ERR_INVALID;
        }
        HMG_TRY(build_combine(eng.comb_mx, n1, mx_rows, near0, items, {}, s, kMxRows, lrows));
    }

    HMG_TRY(eng.kb.alloc(std::max(na, 1)));
    HMG_TRY(eng.keff.alloc(std::max(na, 1)));
    if (na > 0) {
        HMG_TRY(upload_bytes(eng.kb.p, eng.h_kb.data(), sizeof(int) * na, s));
        HMG_HIP(hipMemsetAsync(eng.keff.p, 0, sizeof(int) * na, s));
    }
    HMG_TRY(upload(eng.uoff, eng.h_uoff, s));
    HMG_TRY(upload(eng.voff, eng.h_voff, s));
    HMG_TRY(upload(eng.toff, eng.h_toff, s));
    HMG_TRY(upload(eng.lr_items, items, s));
    HMG_TRY(upload(eng.ldv, eng.h_ldv, s));
    HMG_TRY(upload(eng.ldu, eng.h_ldu, s));
    HMG_TRY(upload(eng.lv_items, vitems, s));
    HMG_TRY(upload(eng.lv_slices, slices, s));
    HMG_TRY(upload(eng.tmap, tmap, s));
    if (mixed) {
        HMG_TRY(upload(eng.mxp, eng.h_mxp, s));
        HMG_TRY(upload(eng.vw_items, eng.h_vw_items, s));
        HMG_TRY(upload(eng.uw_items, eng.h_uw_items, s));
    }
    if (!mixed) {
        HMG_TRY(eng.Tp.alloc((size_t)std::max(eng.n_lv_items, 1) * kSlice));
        eng.p_slots = col0 + eng.n_col_slots;
        HMG_TRY(eng.P.alloc((size_t)eng.p_slots));
        HMG_TRY(eng.T.alloc(std::max<int64_t>(eng.tsize, 1)));
    }
    HMG_HIP(hipStreamSynchronize(s));
    eng.aca_k = k;
    eng.layout_prec = prec;
    return HMG_OK;
}

}  // namespace hmg
