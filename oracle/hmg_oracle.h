/*
 * hmg_oracle.h -- CPU restatement of hmglib's H-matrix build + H-MVP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (hmglib_amd/libhmglib.so) never links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the hmglib source tree, src/).  Arithmetic is plain C11, FP64, built with
 * -ffp-contract=off so that every fused multiply-add is explicit.
 *
 * Pinning (what fixes each function independently of this restatement):
 *   orc_rng_*            SplitMix64 / xoshiro256** published algorithms; SplitMix64
 *                        known-answer vector; cross-checked bit-for-bit against the
 *                        product's own generator.
 *   orc_stretch/morton   naive bit interleave (exhaustive over 2^20 for 3D).
 *   orc_sort             Python's stable sorted() on (code, index).
 *   orc_build (tree)     brute-force recursive traversal with exact boxes on inputs
 *                        where the box quirk does not fire; a hand-worked key/scan
 *                        example for the quirk; tiling / admissibility invariants.
 *   orc_aca              exact reproduction U V^T = A when k_b = min(m1,m2); rank-1
 *                        separable blocks; pivot rules checked on hand-built rows.
 *   orc_hmvp/dense       numpy dense products; the Gaussian closed-form right-hand
 *                        side (quadrature identity) at large N.
 */
#ifndef HMG_ORACLE_H
#define HMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded input generator (SplitMix64 seeding xoshiro256**) ---------- */
typedef struct { uint64_t s[4]; } orc_rng;
uint64_t orc_splitmix64_next(uint64_t* state);
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);
/* n doubles (next>>11)*2^-53 */
void orc_uniform(uint64_t seed, int64_t n, double* out);
/* n standard normals, Box-Muller (cos branch first, then sin), u1 in (0,1] */
void orc_normal(uint64_t seed, int64_t n, double* out);

/* ---- 192-byte work item, same field order as tree.h:38-55 ---------------- */
#define ORC_MAX_DIM 5
#define ORC_WT_ACA 1
#define ORC_WT_DENSE 2
typedef struct {
    int set1_l, set1_u, set2_l, set2_u;
    int work_type;
    double max1[ORC_MAX_DIM];
    double min1[ORC_MAX_DIM];
    double max2[ORC_MAX_DIM];
    double min2[ORC_MAX_DIM];
    int dim;
    char level_1, level_2;
} orc_work_item;

/* ---- kernel descriptor ---------------------------------------------------- */
#define ORC_K_GAUSSIAN_REF 1   /* exp(-r*r)            kernel_system_assembler.h:33-36 */
#define ORC_K_MATERN_REF   2   /* yn(1,r+1e-15) r / (2^{d/2} Gamma(1+d/2))  :77-80 */
#define ORC_K_GAUSSIAN     3   /* exp(-r^2/(2 l^2))                                  */
#define ORC_K_EXPONENTIAL  4   /* exp(-r/l)                                          */
typedef struct {
    int kind;
    int dim;
    double regularization;
    double length_scale;
} orc_kernel;

/* ---- H-matrix state -------------------------------------------------------- */
typedef struct orc_hm orc_hm;

/* Morton codes of one point set (morton.cu:69-155, helper.cu:108-138).
 * X is SoA: X[d*n + p].  Writes codes[n], mn[dim], mx[dim]. */
void orc_minmax(const double* X, int dim, int64_t n, double* mn, double* mx);
uint64_t orc_stretch(uint64_t x, int dim, int bits);
void orc_morton(const double* X, int dim, int bits, int64_t n,
                const double* mn, const double* mx, uint64_t* codes);
/* stable sort of codes (in place) with order[s] = original index (helper.cu:140-167) */
void orc_sort(uint64_t* codes, int64_t n, uint64_t* order);
/* LBVH split (tree.cu:40-78) */
int orc_find_split(const uint64_t* codes, int first, int last);

/* Full setup: minmax, Morton, sort, reorder, traversal, DENSE-first stable
 * partition (hmglib.cu:125-225).  X1/X2 SoA in original order. */
orc_hm* orc_build(int n1, int n2, int dim, int bits, const double* X1, const double* X2,
                  double eta, int max_level, int c_leaf, int root_level_1, int root_level_2);
/* orc_build with exact_boxes = 1: every cluster's box is the exact min/max of its
 * points (no lookup table, so a one-point cluster shifts nothing; the engine's
 * HMG_OPT_EXACT_BOXES).  Not the reference's behaviour where a level holds a
 * one-point cluster (tree.cu:1000-1024); identical to it elsewhere. */
orc_hm* orc_build_ex(int n1, int n2, int dim, int bits, const double* X1, const double* X2, double eta,
                     int max_level, int c_leaf, int root_level_1, int root_level_2, int exact_boxes);
void orc_free(orc_hm* hm);

int orc_block_count(const orc_hm* hm);
int orc_dense_count(const orc_hm* hm);
int orc_aca_count(const orc_hm* hm);
void orc_get_blocks(const orc_hm* hm, orc_work_item* out);
void orc_get_codes(const orc_hm* hm, int set, uint64_t* out);     /* sorted codes */
void orc_get_order(const orc_hm* hm, int set, uint64_t* out);
void orc_get_sorted_coords(const orc_hm* hm, int set, double* out); /* SoA dim x n */
void orc_get_minmax(const orc_hm* hm, int set, double* mn, double* mx);
/* diagnostics: levels with a singleton cluster, nodes whose literal box differs
 * from the exact box, nodes whose admissibility differs fused vs unfused diameter,
 * table entries read past the last run (undefined in the reference). */
void orc_get_diagnostics(const orc_hm* hm, int64_t* out4);
int orc_level_count(const orc_hm* hm);

/* kernel entry on sorted indices (kernel_system_assembler.h:38-52,82-96) */
double orc_kernel_entry(const orc_hm* hm, const orc_kernel* k, int i, int j);

/* Batched ACA in reference mode (linear_algebra.cu:3684-4012, kernels :918-1147,
 * :2009-2222): rank k_b = min(k, m1, m2), sequential row pivots skipping rows whose
 * residual 2-norm < 1e-13, column pivot = LAST argmax (tie_last=1) or FIRST argmax
 * (tie_last=0, the unbatched path linear_algebra.cu:473).  Factors are stored per
 * block, column-major: U_b (m1 x k_b) at uoff[b], V_b (m2 x k_b) at voff[b]. */
void orc_aca(orc_hm* hm, const orc_kernel* kern, int k, int tie_last);
/* tolerance stop (not used by the reference's batched path, whose epsilon is dead):
 * mode 1 = linear_algebra.cu:498-528 with :379 corrected (every 5th step,
 * ||u_r|| ||v_r|| <= eps (1-eta)/(1+eps) ||U V^T||_F); mode 2 = every step,
 * ||u_r|| ||v_r|| <= eps ||U V^T||_F */
typedef struct { int mode; double eps; double eta; } orc_tol;
void orc_aca_tol(orc_hm* hm, const orc_kernel* kern, int k, int tie_last, int tol_mode, double eps, double eta);
int orc_aca_matrix_tol(const double* A, int m1, int m2, int k, int tie_last, int tol_mode, double eps,
                       double eta, double* U, double* V, int* rp, int* cp);
/* The same ACA on an explicit column-major m1 x m2 matrix (pivot-rule checks).
 * U (m1 x min(k,m1,m2)), V (m2 x ...), row/col pivots; returns the rank reached. */
int orc_aca_matrix(const double* A, int m1, int m2, int k, int tie_last, double* U, double* V,
                   int* row_piv, int* col_piv);
int64_t orc_aca_u_size(const orc_hm* hm);
int64_t orc_aca_v_size(const orc_hm* hm);
void orc_get_aca(const orc_hm* hm, int* kb, int* keff, int64_t* uoff, int64_t* voff,
                 double* U, double* V, int* row_piv, int* col_piv);

/* H-MVP y = H x, x and y in ORIGINAL order, R right-hand sides column-major
 * (ld n2 / n1).  Near field evaluated on the fly; far field from orc_aca's factors
 * (linear_algebra.cu:4292-4379, hmglib.cu:271-286). */
void orc_hmvp(const orc_hm* hm, const orc_kernel* kern, int R, const double* x, double* y);
/* H-MVP restricted to a subset of ACA/DENSE blocks (stride sampling) -- used only to
 * time a bounded CPU sample. Returns the number of matrix entries touched. */
int64_t orc_hmvp_sample(const orc_hm* hm, const orc_kernel* kern, int stride, int R,
                        const double* x, double* y);
int64_t orc_aca_sample(orc_hm* hm, const orc_kernel* kern, int k, int stride);
/* Dense FP64 product in sorted space, then un-permuted (hmglib.cu:351-401). */
void orc_dense_mvp(const orc_hm* hm, const orc_kernel* kern, int R, const double* x, double* y);
/* Exact dense rows: y_rows[q] = sum_j K(row,j) x_j for ORIGINAL row indices rows[q]. */
void orc_dense_rows(const orc_hm* hm, const orc_kernel* kern, int nrows, const int64_t* rows,
                    int R, const double* x, double* yrows);
/* Closed-form Gaussian right-hand side (hmglib.cu:403-439), original order. */
void orc_gaussian_rhs(const orc_hm* hm, double* b);

/* Algorithmic work of the built block list: E_near = sum_dense m1*m2,
 * E_lr = sum_aca (m1+m2)*keff (if ACA built, else k_b). out[0]=E_near out[1]=E_lr */
void orc_work(const orc_hm* hm, int64_t* out2);

/* Batch sizing of setup_h_matrix (hmglib.cu:211-219, linear_algebra.cu:4059-4288):
 * returns the dense batch cap (1.05 x the dense entries); fills the dense and ACA
 * batch sizes (each walk ends with a recorded 0) and their counts. */
int64_t orc_work_sizes(const orc_hm* hm, int max_batched_aca_size, int* dense_ws, int cap_d, int* nd,
                       int* aca_ws, int cap_a, int* na);

#ifdef __cplusplus
}
#endif
#endif
