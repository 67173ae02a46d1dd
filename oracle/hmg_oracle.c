/*
 * hmg_oracle.c -- CPU restatement of hmglib's H-matrix build + H-MVP path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See hmg_oracle.h for what pins each
 * function.  Reference citations are file:line in hmglib's src/ directory.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include "hmg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Seeded generator.  SplitMix64 (Steele, Lea, Flood 2014; Vigna's public-domain
 * splitmix64.c) seeds the four words of xoshiro256** (Blackman & Vigna 2018).  */
/* ------------------------------------------------------------------------- */
uint64_t orc_splitmix64_next(uint64_t* state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void orc_rng_seed(orc_rng* r, uint64_t seed)
{
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) r->s[i] = orc_splitmix64_next(&st);
}

uint64_t orc_rng_next(orc_rng* r)
{
    uint64_t* s = r->s;
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

void orc_uniform(uint64_t seed, int64_t n, double* out)
{
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (int64_t i = 0; i < n; i++) out[i] = (double)(orc_rng_next(&r) >> 11) * 0x1.0p-53;
}

void orc_normal(uint64_t seed, int64_t n, double* out)
{
    orc_rng r;
    orc_rng_seed(&r, seed);
    const double two_pi = 6.283185307179586476925286766559;
    /* separate cos() and sin() calls: gcc would otherwise fuse them into sincos(),
     * whose last bit can differ from sin() */
    double (*volatile f_cos)(double) = cos;
    double (*volatile f_sin)(double) = sin;
    for (int64_t i = 0; i < n; i += 2) {
        /* u1 in (0,1]: 1 - [0,1) ; u2 in [0,1) */
        double u1 = 1.0 - (double)(orc_rng_next(&r) >> 11) * 0x1.0p-53;
        double u2 = (double)(orc_rng_next(&r) >> 11) * 0x1.0p-53;
        double rad = sqrt(-2.0 * log(u1));
        out[i] = rad * f_cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = rad * f_sin(two_pi * u2);
    }
}

/* ------------------------------------------------------------------------- */
/* Morton codes: helper.cu:108-138 (exact per-dim min/max), morton.cu:69-155.  */
/* ------------------------------------------------------------------------- */
void orc_minmax(const double* X, int dim, int64_t n, double* mn, double* mx)
{
    for (int d = 0; d < dim; d++) {
        const double* x = X + (int64_t)d * n;
        double lo = x[0], hi = x[0];
        for (int64_t p = 1; p < n; p++) {
            if (x[p] < lo) lo = x[p];
            if (x[p] > hi) hi = x[p];
        }
        mn[d] = lo;
        mx[d] = hi;
    }
}

/* morton.cu:69-115: spread the low `bits` bits of x so that bit b lands at
 * position dim*b; only (3,20), (2,32), (10,4) are defined, else 0. */
uint64_t orc_stretch(uint64_t x, int dim, int bits)
{
    if ((dim == 3 && bits == 20) || (dim == 2 && bits == 32) || (dim == 10 && bits == 4)) {
        uint64_t out = 0;
        for (int b = 0; b < bits; b++)
            if ((x >> b) & 1ULL) out |= 1ULL << (dim * b);
        return out;
    }
    return 0;
}

/* morton.cu:119-123 */
static inline uint64_t coord_to_fp(double coord, double minimum, double mmm, int bits)
{
    double tmp = (coord - minimum) / mmm;
    uint64_t q = (uint64_t)floor(tmp * (double)(1ULL << bits));
    uint64_t cap = (1ULL << bits) - 1;
    return q < cap ? q : cap;
}

/* morton.cu:127-155 */
void orc_morton(const double* X, int dim, int bits, int64_t n,
                const double* mn, const double* mx, uint64_t* codes)
{
    for (int64_t p = 0; p < n; p++) {
        uint64_t c = 0;
        for (int d = 0; d < dim; d++) {
            uint64_t q = coord_to_fp(X[(int64_t)d * n + p], mn[d], mx[d] - mn[d], bits);
            c |= orc_stretch(q, dim, bits) << d;
        }
        codes[p] = c;
    }
}

/* helper.cu:140-167: stable ascending sort by code; order[s] = original index. */
typedef struct { uint64_t code; uint64_t idx; } code_pair;
static int cmp_pair(const void* a, const void* b)
{
    const code_pair* x = (const code_pair*)a;
    const code_pair* y = (const code_pair*)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

void orc_sort(uint64_t* codes, int64_t n, uint64_t* order)
{
    code_pair* pr = (code_pair*)malloc(sizeof(code_pair) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) { pr[i].code = codes[i]; pr[i].idx = (uint64_t)i; }
    qsort(pr, (size_t)n, sizeof(code_pair), cmp_pair);
    for (int64_t i = 0; i < n; i++) { codes[i] = pr[i].code; order[i] = pr[i].idx; }
    free(pr);
}

/* tree.cu:40-78 */
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
int orc_find_split(const uint64_t* codes, int first, int last)
{
    uint64_t fc = codes[first], lc = codes[last];
    if (fc == lc) return (first + last) >> 1;
    int common = clz64(fc ^ lc);
    int split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        int ns = split + step;
        if (ns < last) {
            if (clz64(fc ^ codes[ns]) > common) split = ns;
        }
    } while (step > 1);
    return split;
}

/* ------------------------------------------------------------------------- */
/* H-matrix state                                                             */
/* ------------------------------------------------------------------------- */
struct orc_hm {
    int n[2];
    int dim, bits;
    double eta;
    int max_level, c_leaf;
    double* X[2];          /* sorted SoA coords */
    double mn[2][ORC_MAX_DIM], mx[2][ORC_MAX_DIM];
    uint64_t* codes[2];    /* sorted codes */
    uint64_t* order[2];
    int64_t* inv[2];       /* inv[order[s]] = s */
    orc_work_item* blocks;
    int nblocks, ndense, naca;
    int levels;
    int64_t diag[4];
    int exact_boxes;       /* 1: cluster boxes from the points themselves (no lookup table) */
    /* ACA factors (ACA blocks only, index a = block - ndense) */
    int* kb;
    int* keff;
    int64_t* uoff;
    int64_t* voff;
    int* row_piv;          /* naca * kmax, -1 padded */
    int* col_piv;
    int kmax;
    double* U;
    double* V;
    int64_t usize, vsize;
};

/* tree.cu:94-105 with nvcc's default contraction made explicit: diam = fma(t,t,diam) */
static inline double diameter_fused(const double* mn, const double* mx, int dim)
{
    double diam = 0.0;
    for (int d = 0; d < dim; d++) {
        double t = mx[d] - mn[d];
        diam = fma(t, t, diam);
    }
    return sqrt(diam);
}
static inline double diameter_plain(const double* mn, const double* mx, int dim)
{
    double diam = 0.0;
    for (int d = 0; d < dim; d++) {
        double t = mx[d] - mn[d];
        diam = diam + t * t;
    }
    return sqrt(diam);
}
/* tree.cu:107-118; one of the two squared terms is always zero for boxes, so
 * fused and unfused evaluation agree bit-for-bit. */
static inline double distance(const double* mn1, const double* mx1, const double* mn2,
                              const double* mx2, int dim)
{
    double dist = 0.0;
    for (int d = 0; d < dim; d++) {
        double a = fmax(0.0, mn1[d] - mx2[d]);
        double b = fmax(0.0, mn2[d] - mx1[d]);
        dist = dist + (a * a + b * b);
    }
    return sqrt(dist);
}
/* tree.cu:121-160 */
static inline int admissible(const orc_work_item* w, double eta, int fused)
{
    double d1 = fused ? diameter_fused(w->min1, w->max1, w->dim) : diameter_plain(w->min1, w->max1, w->dim);
    double d2 = fused ? diameter_fused(w->min2, w->max2, w->dim) : diameter_plain(w->min2, w->max2, w->dim);
    double dist = distance(w->min1, w->max1, w->min2, w->max2, w->dim);
    return fmax(d1, d2) <= eta * dist;
}

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* Bounding boxes of one set for all nodes of a level, literally as
 * compute_map_for_lookup_table + compute_lookup_table +
 * set_bounding_box_minmax_using_lookup_table (tree.cu:852-1147):
 *   T = #unique lower bounds; keys[L[t]] = t+1 then keys[U[t]] = -(t+1) (the same
 *   thread, so a singleton keeps -(t+1)); inclusive scan; keys[U[t]] = t+1;
 *   the k-th maximal run of equal nonzero keys gives table entry k. */
static void level_boxes(orc_hm* hm, orc_work_item* nodes, int nn, int set)
{
    const int N = hm->n[set];
    const int dim = hm->dim;
    const double* X = hm->X[set];
    int* L = (int*)malloc(sizeof(int) * nn);
    int* U = (int*)malloc(sizeof(int) * nn);
    int nv = 0;
    for (int i = 0; i < nn; i++) {
        int l = set == 0 ? nodes[i].set1_l : nodes[i].set2_l;
        int u = set == 0 ? nodes[i].set1_u : nodes[i].set2_u;
        if (l > u) continue; /* empty cluster: 2-child split of a size-1 cluster */
        L[nv] = l;
        U[nv] = u;
        nv++;
    }
    qsort(L, nv, sizeof(int), cmp_int);
    qsort(U, nv, sizeof(int), cmp_int);
    int T = 0, Tu = 0;
    for (int i = 0; i < nv; i++) if (i == 0 || L[i] != L[i - 1]) L[T++] = L[i];
    for (int i = 0; i < nv; i++) if (i == 0 || U[i] != U[i - 1]) U[Tu++] = U[i];
    if (Tu != T) hm->diag[3]++;

    int singleton = 0;
    int* keys = (int*)calloc((size_t)N, sizeof(int));
    for (int t = 0; t < T; t++) {
        if (L[t] >= 0 && L[t] < N) keys[L[t]] = t + 1;
        if (t < Tu && U[t] >= 0 && U[t] < N) keys[U[t]] = -(t + 1);
        if (t < Tu && L[t] == U[t]) singleton = 1;
    }
    for (int p = 1; p < N; p++) keys[p] += keys[p - 1];
    for (int t = 0; t < T && t < Tu; t++)
        if (U[t] >= 0 && U[t] < N) keys[U[t]] = t + 1;
    if (singleton) hm->diag[0]++;

    double* tmin = (double*)malloc(sizeof(double) * (size_t)T * dim);
    double* tmax = (double*)malloc(sizeof(double) * (size_t)T * dim);
    for (int64_t q = 0; q < (int64_t)T * dim; q++) { tmin[q] = NAN; tmax[q] = NAN; }
    int run = -1;
    for (int p = 0; p < N; p++) {
        if (keys[p] == 0) continue;
        if (p == 0 || keys[p] != keys[p - 1]) run++;
        if (run >= T) break;
        for (int d = 0; d < dim; d++) {
            double v = X[(int64_t)d * N + p];
            double* lo = &tmin[(int64_t)run * dim + d];
            double* hi = &tmax[(int64_t)run * dim + d];
            /* thrust::minimum / maximum: a<b?a:b, a<b?b:a (no NaN in inputs) */
            if (isnan(*lo) || v < *lo) *lo = v;
            if (isnan(*hi) || *hi < v) *hi = v;
        }
    }
    if (run + 1 < T) hm->diag[3] += T - (run + 1);

    for (int i = 0; i < nn; i++) {
        int l = set == 0 ? nodes[i].set1_l : nodes[i].set2_l;
        int u = set == 0 ? nodes[i].set1_u : nodes[i].set2_u;
        nodes[i].dim = dim;
        if (l > u) continue;
        /* rank of l among unique lower bounds (map, tree.cu:852-918) */
        int lo = 0, hi = T - 1, rank = 0;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            if (L[mid] == l) { rank = mid; break; }
            if (L[mid] < l) lo = mid + 1; else hi = mid - 1;
        }
        double* bmin = set == 0 ? nodes[i].min1 : nodes[i].min2;
        double* bmax = set == 0 ? nodes[i].max1 : nodes[i].max2;
        int differs = 0;
        for (int d = 0; d < dim; d++) {
            bmin[d] = tmin[(int64_t)rank * dim + d];
            bmax[d] = tmax[(int64_t)rank * dim + d];
            /* exact box: the diagnostic, or the box itself in exact mode */
            if (l >= 0 && u < N && l <= u) {
                double elo = X[(int64_t)d * N + l], ehi = elo;
                for (int p = l + 1; p <= u; p++) {
                    double v = X[(int64_t)d * N + p];
                    if (v < elo) elo = v;
                    if (ehi < v) ehi = v;
                }
                if (elo != bmin[d] || ehi != bmax[d]) differs = 1;
                if (hm->exact_boxes) {
                    bmin[d] = elo;
                    bmax[d] = ehi;
                }
            }
        }
        nodes[i].dim = dim;
        if (differs) hm->diag[1]++;
    }
    free(tmin); free(tmax); free(keys); free(L); free(U);
}

/* tree.cu:1429-1557 (traverse_with_dynamic_arrays_dynamic_output), with
 * count_for_new_level (tree.cu:215-266) and generate_new_level (tree.cu:320-447). */
static void traverse(orc_hm* hm, int root_level_1, int root_level_2)
{
    int cap_out = 1024, nout = 0;
    orc_work_item* out = (orc_work_item*)malloc(sizeof(orc_work_item) * cap_out);
    int ncur = 1;
    orc_work_item* cur = (orc_work_item*)calloc(1, sizeof(orc_work_item));
    cur[0].set1_l = 0; cur[0].set1_u = hm->n[0] - 1;
    cur[0].set2_l = 0; cur[0].set2_u = hm->n[1] - 1;
    cur[0].level_1 = (char)root_level_1; cur[0].level_2 = (char)root_level_2;
    hm->levels = 0;
    for (int level = 0; level < hm->max_level; level++) {
        hm->levels++;
        level_boxes(hm, cur, ncur, 0);
        level_boxes(hm, cur, ncur, 1);
        int* nch = (int*)malloc(sizeof(int) * ncur);
        int* nmv = (int*)malloc(sizeof(int) * ncur);
        int* adm = (int*)malloc(sizeof(int) * ncur);
        int tot_ch = 0, tot_mv = 0;
        for (int i = 0; i < ncur; i++) {
            orc_work_item* w = &cur[i];
            adm[i] = admissible(w, hm->eta, 1);
            if (adm[i] != admissible(w, hm->eta, 0)) hm->diag[2]++;
            if (w->set1_l > w->set1_u || w->set2_l > w->set2_u) {
                /* empty cluster (undefined in the reference): the node is dropped */
                nch[i] = 0; nmv[i] = 0;
            } else if (w->level_1 == w->level_2) {
                if (adm[i]) { nch[i] = 0; nmv[i] = 1; }
                else if ((w->set1_u - w->set1_l + 1) <= hm->c_leaf ||
                         (w->set2_u - w->set2_l + 1) <= hm->c_leaf || (level + 1) >= hm->max_level) {
                    nch[i] = 0; nmv[i] = 1;
                } else { nch[i] = 4; nmv[i] = 0; }
            } else { nch[i] = 2; nmv[i] = 0; }
            tot_ch += nch[i];
            tot_mv += nmv[i];
        }
        if (nout + tot_mv > cap_out) {
            while (nout + tot_mv > cap_out) cap_out *= 2;
            out = (orc_work_item*)realloc(out, sizeof(orc_work_item) * cap_out);
        }
        orc_work_item* nxt = (orc_work_item*)calloc(tot_ch > 0 ? tot_ch : 1, sizeof(orc_work_item));
        int co = 0;
        for (int i = 0; i < ncur; i++) {
            orc_work_item w = cur[i];
            if (nch[i] == 4) {
                int s1 = orc_find_split(hm->codes[0], w.set1_l, w.set1_u);
                int s2 = orc_find_split(hm->codes[1], w.set2_l, w.set2_u);
                int l1[4] = {w.set1_l, w.set1_l, s1 + 1, s1 + 1};
                int u1[4] = {s1, s1, w.set1_u, w.set1_u};
                int l2[4] = {w.set2_l, s2 + 1, w.set2_l, s2 + 1};
                int u2[4] = {s2, w.set2_u, s2, w.set2_u};
                for (int c = 0; c < 4; c++) {
                    orc_work_item* ch = &nxt[co++];
                    ch->set1_l = l1[c]; ch->set1_u = u1[c];
                    ch->set2_l = l2[c]; ch->set2_u = u2[c];
                    ch->level_1 = (char)(w.level_1 + 1);
                    ch->level_2 = (char)(w.level_2 + 1);
                }
            } else if (nch[i] == 2) {
                if (w.level_1 < w.level_2) {
                    int s1 = orc_find_split(hm->codes[0], w.set1_l, w.set1_u);
                    orc_work_item* a = &nxt[co++];
                    orc_work_item* b = &nxt[co++];
                    a->set1_l = w.set1_l; a->set1_u = s1; a->set2_l = w.set2_l; a->set2_u = w.set2_u;
                    b->set1_l = s1 + 1; b->set1_u = w.set1_u; b->set2_l = w.set2_l; b->set2_u = w.set2_u;
                    a->level_1 = b->level_1 = (char)(w.level_1 + 1);
                    a->level_2 = b->level_2 = w.level_2;
                } else {
                    int s2 = orc_find_split(hm->codes[1], w.set2_l, w.set2_u);
                    orc_work_item* a = &nxt[co++];
                    orc_work_item* b = &nxt[co++];
                    a->set1_l = w.set1_l; a->set1_u = w.set1_u; a->set2_l = w.set2_l; a->set2_u = s2;
                    b->set1_l = w.set1_l; b->set1_u = w.set1_u; b->set2_l = s2 + 1; b->set2_u = w.set2_u;
                    a->level_1 = b->level_1 = w.level_1;
                    a->level_2 = b->level_2 = (char)(w.level_2 + 1);
                }
            } else if (nmv[i]) {
                w.work_type = adm[i] ? ORC_WT_ACA : ORC_WT_DENSE;
                out[nout++] = w;
            }
        }
        free(nch); free(nmv); free(adm); free(cur);
        cur = nxt;
        ncur = tot_ch;
        if (ncur == 0) break;
    }
    free(cur);
    /* organize_mat_vec_data (linear_algebra.cu:63-86, 4037-4051): stable, DENSE first */
    hm->blocks = (orc_work_item*)malloc(sizeof(orc_work_item) * (nout > 0 ? nout : 1));
    int q = 0;
    for (int i = 0; i < nout; i++) if (out[i].work_type == ORC_WT_DENSE) hm->blocks[q++] = out[i];
    hm->ndense = q;
    for (int i = 0; i < nout; i++) if (out[i].work_type != ORC_WT_DENSE) hm->blocks[q++] = out[i];
    hm->nblocks = nout;
    hm->naca = nout - hm->ndense;
    free(out);
}

orc_hm* orc_build(int n1, int n2, int dim, int bits, const double* X1, const double* X2,
                  double eta, int max_level, int c_leaf, int root_level_1, int root_level_2)
{
    return orc_build_ex(n1, n2, dim, bits, X1, X2, eta, max_level, c_leaf, root_level_1, root_level_2, 0);
}

orc_hm* orc_build_ex(int n1, int n2, int dim, int bits, const double* X1, const double* X2, double eta,
                     int max_level, int c_leaf, int root_level_1, int root_level_2, int exact_boxes)
{
    orc_hm* hm = (orc_hm*)calloc(1, sizeof(orc_hm));
    hm->exact_boxes = exact_boxes;
    hm->n[0] = n1; hm->n[1] = n2;
    hm->dim = dim; hm->bits = bits;
    hm->eta = eta; hm->max_level = max_level; hm->c_leaf = c_leaf;
    const double* Xin[2] = {X1, X2};
    for (int s = 0; s < 2; s++) {
        int64_t n = hm->n[s];
        orc_minmax(Xin[s], dim, n, hm->mn[s], hm->mx[s]);
        hm->codes[s] = (uint64_t*)malloc(sizeof(uint64_t) * n);
        hm->order[s] = (uint64_t*)malloc(sizeof(uint64_t) * n);
        hm->inv[s] = (int64_t*)malloc(sizeof(int64_t) * n);
        orc_morton(Xin[s], dim, bits, n, hm->mn[s], hm->mx[s], hm->codes[s]);
        orc_sort(hm->codes[s], n, hm->order[s]);
        hm->X[s] = (double*)malloc(sizeof(double) * n * dim);
        /* reorder_point_set (helper.cu:293-328) */
        for (int d = 0; d < dim; d++)
            for (int64_t p = 0; p < n; p++)
                hm->X[s][d * n + p] = Xin[s][d * n + (int64_t)hm->order[s][p]];
        for (int64_t p = 0; p < n; p++) hm->inv[s][hm->order[s][p]] = p;
    }
    traverse(hm, root_level_1, root_level_2);
    return hm;
}

void orc_free(orc_hm* hm)
{
    if (!hm) return;
    for (int s = 0; s < 2; s++) {
        free(hm->X[s]); free(hm->codes[s]); free(hm->order[s]); free(hm->inv[s]);
    }
    free(hm->blocks);
    free(hm->kb); free(hm->keff); free(hm->uoff); free(hm->voff);
    free(hm->row_piv); free(hm->col_piv);
    free(hm->U); free(hm->V);
    free(hm);
}

int orc_block_count(const orc_hm* hm) { return hm->nblocks; }
int orc_dense_count(const orc_hm* hm) { return hm->ndense; }
int orc_aca_count(const orc_hm* hm) { return hm->naca; }
int orc_level_count(const orc_hm* hm) { return hm->levels; }
void orc_get_blocks(const orc_hm* hm, orc_work_item* out)
{
    memcpy(out, hm->blocks, sizeof(orc_work_item) * hm->nblocks);
}
void orc_get_codes(const orc_hm* hm, int set, uint64_t* out)
{
    memcpy(out, hm->codes[set], sizeof(uint64_t) * hm->n[set]);
}
void orc_get_order(const orc_hm* hm, int set, uint64_t* out)
{
    memcpy(out, hm->order[set], sizeof(uint64_t) * hm->n[set]);
}
void orc_get_sorted_coords(const orc_hm* hm, int set, double* out)
{
    memcpy(out, hm->X[set], sizeof(double) * hm->n[set] * hm->dim);
}
void orc_get_minmax(const orc_hm* hm, int set, double* mn, double* mx)
{
    memcpy(mn, hm->mn[set], sizeof(double) * hm->dim);
    memcpy(mx, hm->mx[set], sizeof(double) * hm->dim);
}
void orc_get_diagnostics(const orc_hm* hm, int64_t* out4) { memcpy(out4, hm->diag, sizeof(hm->diag)); }

/* ------------------------------------------------------------------------- */
/* Kernel entries on sorted indices (kernel_system_assembler.h:26-104)        */
/* ------------------------------------------------------------------------- */
static inline double kernel_eval(const orc_hm* hm, const orc_kernel* k, int i, int j)
{
    const int dim = hm->dim;
    const int64_t n1 = hm->n[0], n2 = hm->n[1];
    double r2 = 0.0;
    for (int d = 0; d < dim; d++) {
        double t = hm->X[0][d * n1 + i] - hm->X[1][d * n2 + j];
        r2 = r2 + t * t;
    }
    double v;
    switch (k->kind) {
    case ORC_K_GAUSSIAN_REF: {
        double r = sqrt(r2);
        v = exp(-r * r);
        break;
    }
    case ORC_K_MATERN_REF: {
        double r = sqrt(r2);
        v = (yn(1, r + 1.0e-15) * pow(r, 1.0)) /
            (pow(2.0, (double)dim / 2.0) * tgamma(1.0 + (double)dim / 2.0));
        break;
    }
    case ORC_K_GAUSSIAN: {
        double c = 1.0 / (2.0 * k->length_scale * k->length_scale);
        v = exp(-r2 * c);
        break;
    }
    case ORC_K_EXPONENTIAL: {
        double c = 1.0 / k->length_scale;
        v = exp(-sqrt(r2) * c);
        break;
    }
    default:
        v = NAN;
    }
    if (i == j) v += k->regularization;
    return v;
}

double orc_kernel_entry(const orc_hm* hm, const orc_kernel* k, int i, int j)
{
    return kernel_eval(hm, k, i, j);
}

/* ------------------------------------------------------------------------- */
/* ACA, reference mode (linear_algebra.cu:3684-4012; kernels :918-1147,        */
/* :963-1018, :2009-2222).  Residual updates as nvcc contracts them:           */
/* val = fma(-scaling, V, val) over l = 0..r-1.                               */
/* ------------------------------------------------------------------------- */
typedef double (*entry_fn)(const void* ctx, int i, int j);
typedef struct { const orc_hm* hm; const orc_kernel* kern; int l1, l2; } block_ctx;
static double block_entry(const void* c, int i, int j)
{
    const block_ctx* b = (const block_ctx*)c;
    return kernel_eval(b->hm, b->kern, b->l1 + i, b->l2 + j);
}
typedef struct { const double* A; int m1; } dense_ctx;
static double dense_entry(const void* c, int i, int j)
{
    const dense_ctx* d = (const dense_ctx*)c;
    return d->A[(int64_t)j * d->m1 + i];
}

/* Tolerance stop after rank step r (factors u_r, v_r in place), or never when
 * tol->mode == 0.  Mode 1 is the reference's criterion (linear_algebra.cu:498-528,
 * the V'*V of :379 corrected from V'*U): checked when r % 5 == 0, stop when
 * ||u_r|| ||v_r|| <= eps (1 - eta) / (1 + eps) ||U V^T||_F over the first r+1
 * terms (never true for eta >= 1).  Mode 2 is the plain relative criterion,
 * checked at every step: ||u_r|| ||v_r|| <= eps ||U V^T||_F.  ||U V^T||_F^2 =
 * sum_{a,b<=r} (u_a . u_b)(v_a . v_b), summed directly. */
static int tol_stop(const orc_tol* tol, int m1, int m2, int r, const double* U, const double* V)
{
    if (!tol || tol->mode == 0) return 0;
    if (tol->mode == 1 && r % 5 != 0) return 0;
    const double c = tol->mode == 1 ? tol->eps * (1.0 - tol->eta) / (1.0 + tol->eps) : tol->eps;
    double fro2 = 0.0;
    for (int a = 0; a <= r; a++)
        for (int b = 0; b <= r; b++) {
            double uu = 0.0, vv = 0.0;
            for (int i = 0; i < m1; i++) uu += U[(int64_t)a * m1 + i] * U[(int64_t)b * m1 + i];
            for (int j = 0; j < m2; j++) vv += V[(int64_t)a * m2 + j] * V[(int64_t)b * m2 + j];
            fro2 += uu * vv;
        }
    double nu = 0.0, nv = 0.0;
    for (int i = 0; i < m1; i++) nu += U[(int64_t)r * m1 + i] * U[(int64_t)r * m1 + i];
    for (int j = 0; j < m2; j++) nv += V[(int64_t)r * m2 + j] * V[(int64_t)r * m2 + j];
    return sqrt(nu) * sqrt(nv) <= c * sqrt(fro2 > 0.0 ? fro2 : 0.0);
}

static void aca_core(entry_fn entry, const void* ctx, int m1, int m2, int kb, int tie_last,
                     const orc_tol* tol, double* U, double* V, int* keff, int* rp, int* cp)
{
    int i_r = -1;
    int stop = 0;
    *keff = 0;
    for (int r = 0; r < kb; r++) {
        double* v = V + (int64_t)r * m2;
        double* u = U + (int64_t)r * m1;
        /* compute_current_batched_v_r: sequential row search (:2068-2132) */
        for (;;) {
            if (i_r >= m1 - 1) { stop = 1; break; }          /* check_for_maximum_i_r */
            i_r++;
            double nrm = 0.0;
            for (int j = 0; j < m2; j++) {
                double val = entry(ctx, i_r, j);
                for (int l = 0; l < r; l++)
                    val = fma(-U[(int64_t)l * m1 + i_r], V[(int64_t)l * m2 + j], val);
                v[j] = val;
                nrm = nrm + val * val;
            }
            if (sqrt(nrm) >= 1.0e-13) break;                  /* update_ir (:1225) */
        }
        if (stop) break;
        /* tuple_absolute_maximum (:991-1018): last argmax; unbatched path: first */
        int jr = 0;
        double best = fabs(v[0]);
        for (int j = 1; j < m2; j++) {
            double a = fabs(v[j]);
            if (tie_last ? (a >= best) : (a > best)) { best = a; jr = j; }
        }
        double p = v[jr];                                      /* read before scaling */
        for (int j = 0; j < m2; j++) v[j] = v[j] / p;
        for (int i = 0; i < m1; i++) {
            double val = entry(ctx, i, jr);
            for (int l = 0; l < r; l++)
                val = fma(-V[(int64_t)l * m2 + jr], U[(int64_t)l * m1 + i], val);
            u[i] = val;
        }
        rp[r] = i_r;
        cp[r] = jr;
        *keff = r + 1;
        if (tol_stop(tol, m1, m2, r, U, V)) break;
    }
}

static void aca_block(const orc_hm* hm, const orc_kernel* kern, const orc_work_item* w, int kb,
                      int tie_last, const orc_tol* tol, double* U, double* V, int* keff, int* rp, int* cp)
{
    block_ctx c = {hm, kern, w->set1_l, w->set2_l};
    aca_core(block_entry, &c, w->set1_u - w->set1_l + 1, w->set2_u - w->set2_l + 1, kb, tie_last,
             tol, U, V, keff, rp, cp);
}

int orc_aca_matrix_tol(const double* A, int m1, int m2, int k, int tie_last, int tol_mode, double eps,
                       double eta, double* U, double* V, int* rp, int* cp)
{
    dense_ctx c = {A, m1};
    orc_tol tol = {tol_mode, eps, eta};
    int kb = k;
    if (kb > m1) kb = m1;
    if (kb > m2) kb = m2;
    for (int r = 0; r < kb; r++) { rp[r] = -1; cp[r] = -1; }
    int keff = 0;
    aca_core(dense_entry, &c, m1, m2, kb, tie_last, &tol, U, V, &keff, rp, cp);
    return keff;
}

int orc_aca_matrix(const double* A, int m1, int m2, int k, int tie_last, double* U, double* V,
                   int* rp, int* cp)
{
    return orc_aca_matrix_tol(A, m1, m2, k, tie_last, 0, 0.0, 0.0, U, V, rp, cp);
}

static void aca_run(orc_hm* hm, const orc_kernel* kern, int k, int tie_last, int stride, const orc_tol* tol)
{
    free(hm->kb); free(hm->keff); free(hm->uoff); free(hm->voff);
    free(hm->row_piv); free(hm->col_piv); free(hm->U); free(hm->V);
    const int na = hm->naca;
    hm->kb = (int*)calloc(na > 0 ? na : 1, sizeof(int));
    hm->keff = (int*)calloc(na > 0 ? na : 1, sizeof(int));
    hm->uoff = (int64_t*)calloc(na + 1, sizeof(int64_t));
    hm->voff = (int64_t*)calloc(na + 1, sizeof(int64_t));
    int kmax = 0;
    for (int a = 0; a < na; a++) {
        const orc_work_item* w = &hm->blocks[hm->ndense + a];
        int m1 = w->set1_u - w->set1_l + 1, m2 = w->set2_u - w->set2_l + 1;
        int kb = k;
        if (kb > m1) kb = m1;
        if (kb > m2) kb = m2;                                 /* set_k_per_item (:769-787) */
        if (stride > 1 && (a % stride) != 0) kb = 0;
        hm->kb[a] = kb;
        if (kb > kmax) kmax = kb;
        hm->uoff[a + 1] = hm->uoff[a] + (int64_t)m1 * kb;
        hm->voff[a + 1] = hm->voff[a] + (int64_t)m2 * kb;
    }
    hm->kmax = kmax;
    hm->usize = hm->uoff[na];
    hm->vsize = hm->voff[na];
    hm->U = (double*)calloc((size_t)(hm->usize > 0 ? hm->usize : 1), sizeof(double));
    hm->V = (double*)calloc((size_t)(hm->vsize > 0 ? hm->vsize : 1), sizeof(double));
    hm->row_piv = (int*)malloc(sizeof(int) * (size_t)(na > 0 ? na : 1) * (kmax > 0 ? kmax : 1));
    hm->col_piv = (int*)malloc(sizeof(int) * (size_t)(na > 0 ? na : 1) * (kmax > 0 ? kmax : 1));
    for (int64_t q = 0; q < (int64_t)na * (kmax > 0 ? kmax : 1); q++) hm->row_piv[q] = hm->col_piv[q] = -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int a = 0; a < na; a++) {
        if (hm->kb[a] == 0) continue;
        aca_block(hm, kern, &hm->blocks[hm->ndense + a], hm->kb[a], tie_last, tol,
                  hm->U + hm->uoff[a], hm->V + hm->voff[a], &hm->keff[a],
                  hm->row_piv + (int64_t)a * kmax, hm->col_piv + (int64_t)a * kmax);
    }
}

void orc_aca(orc_hm* hm, const orc_kernel* kern, int k, int tie_last)
{
    aca_run(hm, kern, k, tie_last, 1, NULL);
}

void orc_aca_tol(orc_hm* hm, const orc_kernel* kern, int k, int tie_last, int tol_mode, double eps, double eta)
{
    orc_tol tol = {tol_mode, eps, eta};
    aca_run(hm, kern, k, tie_last, 1, &tol);
}

int64_t orc_aca_sample(orc_hm* hm, const orc_kernel* kern, int k, int stride)
{
    aca_run(hm, kern, k, 1, stride, NULL);
    int64_t e = 0;
    for (int a = 0; a < hm->naca; a++) {
        const orc_work_item* w = &hm->blocks[hm->ndense + a];
        e += (int64_t)(w->set1_u - w->set1_l + 1 + w->set2_u - w->set2_l + 1) * hm->kb[a];
    }
    return e;
}

int64_t orc_aca_u_size(const orc_hm* hm) { return hm->usize; }
int64_t orc_aca_v_size(const orc_hm* hm) { return hm->vsize; }
void orc_get_aca(const orc_hm* hm, int* kb, int* keff, int64_t* uoff, int64_t* voff,
                 double* U, double* V, int* row_piv, int* col_piv)
{
    const int na = hm->naca;
    if (kb) memcpy(kb, hm->kb, sizeof(int) * na);
    if (keff) memcpy(keff, hm->keff, sizeof(int) * na);
    if (uoff) memcpy(uoff, hm->uoff, sizeof(int64_t) * (na + 1));
    if (voff) memcpy(voff, hm->voff, sizeof(int64_t) * (na + 1));
    if (U) memcpy(U, hm->U, sizeof(double) * hm->usize);
    if (V) memcpy(V, hm->V, sizeof(double) * hm->vsize);
    if (row_piv) memcpy(row_piv, hm->row_piv, sizeof(int) * (size_t)na * (hm->kmax > 0 ? hm->kmax : 1));
    if (col_piv) memcpy(col_piv, hm->col_piv, sizeof(int) * (size_t)na * (hm->kmax > 0 ? hm->kmax : 1));
}

/* ------------------------------------------------------------------------- */
/* H-MVP (hmglib.cu:271-286, linear_algebra.cu:4292-4379).                    */
/* ------------------------------------------------------------------------- */
static int64_t hmvp_impl(const orc_hm* hm, const orc_kernel* kern, int R, const double* x,
                         double* y, int stride)
{
    const int64_t n1 = hm->n[0], n2 = hm->n[1];
    double* xs = (double*)malloc(sizeof(double) * n2 * R);
    for (int c = 0; c < R; c++)
        for (int64_t p = 0; p < n2; p++) xs[c * n2 + p] = x[c * n2 + (int64_t)hm->order[1][p]];
    /* the near field reads the right-hand sides of one point together */
    double* xr = (double*)malloc(sizeof(double) * n2 * R);
    for (int64_t p = 0; p < n2; p++)
        for (int c = 0; c < R; c++) xr[p * R + c] = xs[c * n2 + p];
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    double* ys = (double*)calloc((size_t)(n1 * R * nth), sizeof(double));
    int64_t touched = 0;
#pragma omp parallel reduction(+:touched)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* yl = ys + (int64_t)tid * n1 * R;
        double* tmp = (double*)malloc(sizeof(double) * (hm->kmax > 0 ? hm->kmax : 1) * R);
        double* acc = (double*)malloc(sizeof(double) * R);
#pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < hm->nblocks; b++) {
            const orc_work_item* w = &hm->blocks[b];
            const int l1 = w->set1_l, l2 = w->set2_l;
            const int m1 = w->set1_u - l1 + 1, m2 = w->set2_u - l2 + 1;
            if (w->work_type == ORC_WT_DENSE) {
                if (stride > 1 && (b % stride) != 0) continue;
                /* each entry evaluated once and applied to every right-hand side;
                 * every (row, rhs) sum still runs over the columns in order */
                for (int i = 0; i < m1; i++) {
                    for (int c = 0; c < R; c++) acc[c] = 0.0;
                    for (int j = 0; j < m2; j++) {
                        const double kv = kernel_eval(hm, kern, l1 + i, l2 + j);
                        const double* xj = xr + (int64_t)(l2 + j) * R;
                        for (int c = 0; c < R; c++) acc[c] = fma(kv, xj[c], acc[c]);
                    }
                    for (int c = 0; c < R; c++) yl[c * n1 + l1 + i] += acc[c];
                }
                touched += (int64_t)m1 * m2;
            } else {
                int a = b - hm->ndense;
                if (!hm->kb || hm->keff[a] == 0) continue;
                const int ke = hm->keff[a];
                const double* U = hm->U + hm->uoff[a];
                const double* V = hm->V + hm->voff[a];
                for (int c = 0; c < R; c++)
                    for (int l = 0; l < ke; l++) {
                        double acc = 0.0;
                        for (int j = 0; j < m2; j++) acc = fma(V[(int64_t)l * m2 + j], xs[c * n2 + l2 + j], acc);
                        tmp[c * ke + l] = acc;
                    }
                for (int c = 0; c < R; c++)
                    for (int i = 0; i < m1; i++) {
                        double acc = 0.0;
                        for (int l = 0; l < ke; l++) acc = fma(U[(int64_t)l * m1 + i], tmp[c * ke + l], acc);
                        yl[c * n1 + l1 + i] += acc;
                    }
                touched += (int64_t)(m1 + m2) * ke;
            }
        }
        free(tmp);
        free(acc);
    }
    for (int c = 0; c < R; c++)
        for (int64_t s = 0; s < n1; s++) {
            double acc = 0.0;
            for (int t = 0; t < nth; t++) acc += ys[(int64_t)t * n1 * R + c * n1 + s];
            y[c * n1 + (int64_t)hm->order[0][s]] = acc;
        }
    free(ys);
    free(xr);
    free(xs);
    return touched;
}

void orc_hmvp(const orc_hm* hm, const orc_kernel* kern, int R, const double* x, double* y)
{
    hmvp_impl(hm, kern, R, x, y, 1);
}

int64_t orc_hmvp_sample(const orc_hm* hm, const orc_kernel* kern, int stride, int R, const double* x, double* y)
{
    return hmvp_impl(hm, kern, R, x, y, stride);
}

/* apply_full_mvp (hmglib.cu:351-401): dense product in sorted space */
void orc_dense_mvp(const orc_hm* hm, const orc_kernel* kern, int R, const double* x, double* y)
{
    const int64_t n1 = hm->n[0], n2 = hm->n[1];
    double* xs = (double*)malloc(sizeof(double) * n2 * R);
    for (int c = 0; c < R; c++)
        for (int64_t p = 0; p < n2; p++) xs[c * n2 + p] = x[c * n2 + (int64_t)hm->order[1][p]];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n1; i++) {
        for (int c = 0; c < R; c++) {
            double acc = 0.0;
            for (int64_t j = 0; j < n2; j++) acc = fma(kernel_eval(hm, kern, (int)i, (int)j), xs[c * n2 + j], acc);
            y[c * n1 + (int64_t)hm->order[0][i]] = acc;
        }
    }
    free(xs);
}

void orc_dense_rows(const orc_hm* hm, const orc_kernel* kern, int nrows, const int64_t* rows,
                    int R, const double* x, double* yrows)
{
    const int64_t n2 = hm->n[1];
    double* xs = (double*)malloc(sizeof(double) * n2 * R);
    for (int c = 0; c < R; c++)
        for (int64_t p = 0; p < n2; p++) xs[c * n2 + p] = x[c * n2 + (int64_t)hm->order[1][p]];
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nrows; q++) {
        int i = (int)hm->inv[0][rows[q]];
        for (int c = 0; c < R; c++) {
            double acc = 0.0;
            for (int64_t j = 0; j < n2; j++) acc = fma(kernel_eval(hm, kern, i, (int)j), xs[c * n2 + j], acc);
            yrows[(int64_t)c * nrows + q] = acc;
        }
    }
    free(xs);
}

/* gen_gaussian_kernel_rhs (hmglib.cu:403-418), evaluated in original order */
void orc_gaussian_rhs(const orc_hm* hm, double* b)
{
    const int64_t n1 = hm->n[0];
    const double sqpi = sqrt(M_PI);
    for (int64_t s = 0; s < n1; s++) {
        double result = 1.0;
        for (int d = 0; d < hm->dim; d++) {
            double xv = hm->X[0][d * n1 + s];
            result = result * sqpi * (erf(1.0 - xv) - erf(0.0 - xv)) / (2.0);
        }
        b[hm->order[0][s]] = result;
    }
}

void orc_work(const orc_hm* hm, int64_t* out2)
{
    int64_t en = 0, el = 0;
    for (int b = 0; b < hm->nblocks; b++) {
        const orc_work_item* w = &hm->blocks[b];
        int64_t m1 = w->set1_u - w->set1_l + 1, m2 = w->set2_u - w->set2_l + 1;
        if (w->work_type == ORC_WT_DENSE) en += m1 * m2;
        else {
            int a = b - hm->ndense;
            int ke = hm->keff ? hm->keff[a] : 0;
            el += (m1 + m2) * ke;
        }
    }
    out2[0] = en;
    out2[1] = el;
}

/* Batch sizing of the reference's setup (hmglib.cu:211-219): the dense batch cap
 * is overridden by 1.05 x the dense entries (predict_precomputing_memory_requirements,
 * linear_algebra.cu:4176-4198), then precompute_work_sizes (:4201-4288) walks the
 * DENSE blocks with compute_current_dense_work_size_magma (:4059-4108, batch size
 * = blocks taken while the running sum of m1*m2 stays <= the cap, the block that
 * crosses it given back) and the ACA blocks with compute_current_aca_work_size
 * (:4158-4174, the same with the running sum of m1); each walk records batch sizes
 * until a batch of size 0, which is recorded too.  Sums are 64-bit here (the
 * reference's int overflows at the scale configuration).  Returns the dense cap;
 * *nd / *na = batch counts, dense_ws / aca_ws receive up to cap_d / cap_a sizes. */
static int batch_walk(const orc_hm* hm, int first, int end, int64_t cap, int by_entries, int* ws, int capn)
{
    int count = 0, at = first;
    for (;;) {
        int taken = 0;
        int64_t sum = 0;
        while (sum <= cap && at + taken < end) {
            const orc_work_item* w = &hm->blocks[at + taken];
            const int64_t m1 = w->set1_u - w->set1_l + 1, m2 = w->set2_u - w->set2_l + 1;
            sum += by_entries ? m1 * m2 : m1;
            taken++;
        }
        if (sum > cap) taken--;
        if (count < capn) ws[count] = taken;
        count++;
        if (taken <= 0) break;
        at += taken;
    }
    return count;
}

int64_t orc_work_sizes(const orc_hm* hm, int max_batched_aca_size, int* dense_ws, int cap_d, int* nd,
                       int* aca_ws, int cap_a, int* na)
{
    int64_t entries = 0;
    for (int b = 0; b < hm->ndense; b++) {
        const orc_work_item* w = &hm->blocks[b];
        entries += (int64_t)(w->set1_u - w->set1_l + 1) * (int64_t)(w->set2_u - w->set2_l + 1);
    }
    const int64_t dense_cap = (int64_t)((double)entries * 1.05);
    *nd = batch_walk(hm, 0, hm->ndense, dense_cap, 1, dense_ws, cap_d);
    *na = batch_walk(hm, hm->ndense, hm->nblocks, (int64_t)max_batched_aca_size, 0, aca_ws, cap_a);
    return dense_cap;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
