/*
 * hmglib.h -- C API of the MI355X-native H-matrix engine (drop-in for hmglib).
 *
 * Keeps hmglib's struct tag, function names and user-visible field names and
 * types (reference: src/hmglib.h:29-120) so that a caller of hmglib recompiles
 * unchanged.  Changes at the type level, all library-owned:
 *   - magma_queue_t magma_queue -> void* (a hipStream_t owned by the library);
 *   - struct system_assembler is a plain kernel descriptor in device memory,
 *     not a C++ object with a virtual get_matrix_entry;
 *   - one trailing field `engine` holds the library's private state.
 * Device pointers are HIP device pointers.  Everything is synchronous.
 *
 * The void entry points behave like the reference: on error they print a
 * message and exit(EXIT_FAILURE) (src/helper.cu:31-37).  Every one has an
 * hmg_* twin returning hmg_status instead.
 */
#ifndef HMGLIB_H
#define HMGLIB_H

#include <stdint.h>
#include "morton.h"
#include "tree.h"
#include "linear_algebra.h"
#include "system_assembler.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: src/hmglib.h:29-102 (same order, same types except as noted) */
struct h_matrix_data
{
	/* point coordinates, per set a HOST array of `dim` DEVICE pointers (SoA,
	 * point_count[i] doubles each).  The caller writes them between
	 * init_h_matrix_data and setup_h_matrix; setup reorders them in place into
	 * Morton order (reference: hmglib.cu:157-158). */
	double** coords_d[2];
	unsigned int* point_ids_d[2];
	double* max_per_dim_d[2];
	double* min_per_dim_d[2];
	double** coords_device[2];
	uint64_t* code_d[2];
	struct point_set* points_d[2];
	struct morton_code* morton_d[2];
	uint64_t* order[2];

	int bits;
	int dim;
	double eta;
	int max_level;
	int c_leaf;
	int mat_vec_data_count;
	int max_elements_in_array;
	int max_elements_in_mat_vec_data_array;
	struct work_item** mat_vec_data;
	int mat_vec_data_array_size;
	int k;
	double epsilon;
	struct mat_vec_data_info mat_vec_info;
	int point_count[2];
	struct system_assembler* assem;
	int max_batched_dense_size;
	int max_batched_aca_size;
	double dense_batching_ratio;

	double* U;
	double* V;
	double* dA;

	void* magma_queue;          /* hipStream_t (reference: magma_queue_t) */

	int* dense_work_size;
	int* aca_work_size;
	int dense_batch_count;
	int aca_batch_count;

	char root_level_set_1;
	char root_level_set_2;

	void* engine;               /* library-private state (not in the reference) */
};

typedef enum {
	HMG_OK = 0,
	HMG_ERR_INVALID = 1,       /* invalid argument or state */
	HMG_ERR_OOM = 2,           /* device memory exhausted or budget exceeded */
	HMG_ERR_HIP = 3,           /* HIP runtime error */
	HMG_ERR_NOT_READY = 4      /* e.g. apply before setup, no kernel set */
} hmg_status;

/* ---- drop-in entry points (reference: src/hmglib.h:104-120) ---------------- */
/* replaces hmglib.h:104 / hmglib.cu:68-123 */
extern void init_h_matrix_data(struct h_matrix_data* data, int point_count[2], int dim, int bits);
/* replaces hmglib.h:106 / hmglib.cu:125-225 */
extern void setup_h_matrix(struct h_matrix_data* data);
/* replaces hmglib.h:108 / hmglib.cu:271-286: y = H x, x and y device vectors in the
 * caller's point order; x is restored exactly, y is overwritten. */
extern void apply_h_matrix_mvp(double* x, double* y, struct h_matrix_data* data);
/* replaces hmglib.h:110 / hmglib.cu:288-301: per-block path, ACA recomputed with the
 * unbatched pivot rule (first argmax, linear_algebra.cu:473), nothing precomputed. */
extern void apply_h_matrix_mvp_without_batching(double* x, double* y, struct h_matrix_data* data);
/* replaces hmglib.h:112 / hmglib.cu:305-348 */
extern void destroy_h_matrix_data(struct h_matrix_data* data);
/* replaces hmglib.h:114 / hmglib.cu:228-236 */
extern void precompute_aca(struct h_matrix_data* data);
/* replaces hmglib.h:116 / hmglib.cu:239-268 */
extern void precompute_dense(struct h_matrix_data* data);
/* replaces hmglib.h:118 / hmglib.cu:351-401: dense product y = K x (never forms K) */
extern void apply_full_mvp(double* x, double* y, struct h_matrix_data* data);
/* replaces hmglib.h:120 / hmglib.cu:420-439 */
extern void set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data);

/* ---- status-returning twins ------------------------------------------------- */
hmg_status hmg_init_h_matrix_data(struct h_matrix_data* data, int point_count[2], int dim, int bits);
hmg_status hmg_setup_h_matrix(struct h_matrix_data* data);
hmg_status hmg_apply_h_matrix_mvp(double* x, double* y, struct h_matrix_data* data);
hmg_status hmg_apply_h_matrix_mvp_without_batching(double* x, double* y, struct h_matrix_data* data);
hmg_status hmg_destroy_h_matrix_data(struct h_matrix_data* data);
hmg_status hmg_precompute_aca(struct h_matrix_data* data);
hmg_status hmg_precompute_dense(struct h_matrix_data* data);
hmg_status hmg_apply_full_mvp(double* x, double* y, struct h_matrix_data* data);
hmg_status hmg_set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data);

/* ---- extensions (new symbols; no counterpart in the reference) -------------- */
/* Multi-RHS apply: X is point_count[1] x nrhs, Y is point_count[0] x nrhs, both
 * column-major device arrays in the caller's point order.  Deterministic: every
 * row of Y is summed in a fixed order, so a result is bitwise the same whatever
 * else runs on the device. */
hmg_status hmg_apply_h_matrix_mvp_multi(const double* X, double* Y, int nrhs, struct h_matrix_data* data);
/* Many H-matrices at once (no counterpart in the reference; for callers with many
 * independent point clouds).  Each entry of data[] is an H-matrix of its own
 * (init_h_matrix_data, coordinates, parameters as usual); threads = host threads
 * that drive them (0: 8).
 * hmg_setup_batch: setup_h_matrix of every entry.  Entries with one point set
 *   against itself (the two sets bitwise equal), exact cluster boxes
 *   (HMG_OPT_EXACT_BOXES), root levels 0 and the same dimension, code width and
 *   tree parameters are set up together, one set of launches per step for all of
 *   them (one sort, every tree in one level-by-level traversal); every other entry
 *   is set up alone, `threads` at a time.  Either way each entry ends exactly as
 *   setup_h_matrix leaves it.
 * hmg_precompute_aca_batch: precompute_aca of every entry (after setup and with
 *   data[i]->assem set), in ONE launch set for all entries that share a kernel,
 *   point dimension, stopping rule and precision: their admissible blocks run side
 *   by side (largest first, 8-wavefront teams for the largest beside the rest).
 *   Ranks and factors are those precompute_aca gives each entry alone.
 * hmg_apply_batch: Y[i] = H_i X[i] for nrhs[i] right-hand sides each (as
 *   hmg_apply_h_matrix_mvp_multi), the entries' products running concurrently.
 * An entry may appear once per call.  The first failure (lowest index) is
 * returned; hmg_last_error names the entry. */
hmg_status hmg_setup_batch(struct h_matrix_data* const* data, int count, int threads);
hmg_status hmg_precompute_aca_batch(struct h_matrix_data* const* data, int count, int threads);
hmg_status hmg_apply_batch(struct h_matrix_data* const* data, const double* const* X, double* const* Y,
                           const int* nrhs, int count, int threads);
/* The products of hmg_apply_batch captured once into one hipGraph and replayed
 * (no counterpart in the reference).  _create runs each product once (so that its
 * buffers exist), then captures all of them for the given operands (device
 * pointers, fixed for the graph's life; their contents may change between
 * launches); _launch replays every product and waits.  A launch refuses
 * (HMG_ERR_NOT_READY) when an H-matrix changed since the capture (setup,
 * precompute, hmg_set_option, another kernel object, a wider product that grew
 * its buffers); capture again then.  The kernel object's parameters are read at
 * capture.  Each H-matrix needs precomputed factors.  No other work may run on
 * the listed H-matrices during a launch. */
typedef struct hmg_batch_graph hmg_batch_graph;
hmg_status hmg_apply_batch_graph_create(struct h_matrix_data* const* data, const double* const* X,
                                        double* const* Y, const int* nrhs, int count, hmg_batch_graph** out);
hmg_status hmg_apply_batch_graph_launch(hmg_batch_graph* graph);
void hmg_apply_batch_graph_destroy(hmg_batch_graph* graph);
/* Y = H X for an H-matrix with precomputed FP64 factors, allocating nothing (no
 * counterpart in the reference; for callers that cannot allocate device memory:
 * code under stream capture, serving loops that forbid allocation).  Same kernels
 * and summation order as hmg_apply_h_matrix_mvp_multi one right-hand side at a
 * time (HMG_OPT_FP64_BLOCK_RHS above nrhs), so the result is bitwise that result
 * (and within rounding of the block path's).  The per-call buffers live in a fixed device array of 32 MiB that
 * the library holds on every device it runs on until the process ends; an
 * H-matrix whose buffers need more is refused with HMG_ERR_INVALID, mixed-precision
 * or missing factors with HMG_ERR_NOT_READY.  Calls on one device share the array:
 * calls that overlap on the same device race. */
hmg_status hmg_apply_h_matrix_mvp_noalloc(const double* X, double* Y, int nrhs, struct h_matrix_data* data);
/* Autotuner of the product's launch shape (no counterpart in the reference): times
 * `reps` products (after one warm-up) for each candidate -- the near field's LDS
 * reservation of 0, 16, 24, 36, 48, 64 or 80 KB per workgroup with the near field
 * and the low-rank kernels on two streams, or everything on one stream; with
 * mixed-precision factors each with the low-rank products and the near field in
 * workgroup or one-wavefront form -- on the caller's device arrays X, Y (Y is
 * overwritten; results do not depend on the choice) and keeps the fastest
 * (HMG_OPT_NEAR_LDS_BYTES, HMG_OPT_SERIALIZE, HMG_OPT_LR_WORKGROUP,
 * HMG_OPT_NEAR_WORKGROUP).  out[0] = LDS bytes, out[1] = 1 for one stream, out[2] =
 * microseconds per product, out[3] = 1 (low-rank workgroup form) | 2 (near-field
 * workgroup form).  Needs precompute_aca. */
hmg_status hmg_autotune_product(struct h_matrix_data* data, const double* X, double* Y, int nrhs, int reps,
                                int64_t out[4]);
/* dense product for nrhs right-hand sides on sampled ORIGINAL rows (checker helper
 * at scale): Yrows[q + c*nrows] = sum_j K(rows[q], j) X[j + c*n2]. rows is a device
 * int64 array. */
hmg_status hmg_dense_rows(const double* X, int nrhs, const int64_t* rows, int nrows, double* Yrows,
                          struct h_matrix_data* data);
/* Algorithmic work of the built H-matrix: out[0] = sum_dense m1*m2 (E_near),
 * out[1] = sum_aca (m1+m2)*rank (E_lr, ranks as built; k_b before precompute_aca),
 * out[2] = #dense blocks, out[3] = #ACA blocks, out[4] = tree levels,
 * out[5] = levels on which a size-1 cluster changed the bounding-box table,
 * out[6] = sum_aca m1*rank, out[7] = sum_aca m2*rank, out[8] = sum_aca rank. */
hmg_status hmg_get_work(const struct h_matrix_data* data, int64_t out[9]);
/* Shape of the H-MVP schedule (out holds 11 values): out[0] = near-field work items, out[1] = low-rank
 * work items, out[2] = rows over low-rank items, out[3] = t values read by the
 * low-rank items (sum of K over items), out[4] = sum_aca m1, out[5] = sum_aca m2,
 * out[6] = partial sums combined per product (entries of the per-row lists),
 * out[7] = 1 if the near field evaluates transposed DENSE pairs once,
 * out[8] = column-sum slots of that symmetric form, out[9] = 1 if both point
 * sets are the same points (sorted coordinates bitwise equal), out[10] = kernel
 * entries the matrix-free near field evaluates per product. */
hmg_status hmg_get_schedule(const struct h_matrix_data* data, int64_t out[11]);
/* Engine options (no counterpart in the reference; defaults in brackets):
 * HMG_OPT_NEAR_SYMMETRIC [1]: when both sets are the same points, evaluate each
 *   transposed pair of DENSE blocks once (row and column sums); takes effect at
 *   the next setup_h_matrix.
 * HMG_OPT_NEAR_LDS_BYTES [65536]: LDS reserved per near-field workgroup, which
 *   caps its occupancy so the low-rank streams keep wave slots beside it.
 * HMG_OPT_SERIALIZE [0]: run the near field on the main stream (no overlap).
 * HMG_OPT_PRECISION [HMG_PREC_FP64]: arithmetic of the next precompute_aca and
 *   of every product.  HMG_PREC_MIXED computes the ACA in FP64 exactly as
 *   HMG_PREC_FP64 does and stores its factors as BF16 hi + lo pairs (the value
 *   rounded to FP32, then split), runs the low-rank products and the near field
 *   (evaluated in FP32) as BF16 hi+lo split MFMA GEMMs accumulated in FP32, for up
 *   to 32 right-hand sides per pass.  X and Y stay FP64 at the boundary.  In this
 *   mode data->U and data->V point at the BF16 factor panels and a stored FP64
 *   near field is not used.
 * HMG_OPT_ACA_ROW_SCAN [1]: when the residual-row search of the ACA rejects a
 *   row, compute the norms of the following rows four at a time (rank steps up
 *   to 64) instead of one at a time, and give rank 0 without evaluating it to a
 *   block whose entries are provably all too small for any row to pass the
 *   search (decreasing kernels, bound from the points' bounding boxes).  The
 *   factors are bitwise the same.
 * HMG_OPT_ACA_BIG_BLOCK [512]: admissible blocks with more rows plus columns
 *   than this run their ACA on a workgroup of 4 wavefronts instead of one
 *   wavefront (row norms summed in another order; same factors up to the
 *   rounding of those norms).
 * HMG_OPT_ACA_HUGE_BLOCK [16384]: blocks with more rows plus columns than this
 *   run on a workgroup of 8 wavefronts, in a launch on the engine's second
 *   stream beside the others (same factors up to the rounding of the norms).
 * HMG_OPT_ACA_TOLERANCE [0]: stopping rule of the ACA.  0: fixed rank k (the
 *   reference's batched path, whose epsilon is unused).  1: the reference's
 *   tolerance criterion (linear_algebra.cu:498-528 with the V'*V of :379
 *   corrected), checked at every fifth rank step: stop after step r when
 *   ||u_r|| ||v_r|| <= epsilon (1 - eta) / (1 + epsilon) ||U V^T||_F (never for
 *   eta >= 1).  2: the plain relative criterion, checked at every step:
 *   ||u_r|| ||v_r|| <= epsilon ||U V^T||_F.  Uses data->epsilon and data->eta at
 *   precompute_aca; ranks stay capped at k and the factor layout keeps k columns.
 * HMG_OPT_LR_WORKGROUP [0]: mixed precision: 1 runs V^T X and U T as workgroups
 *   of wavefronts that share their X (T) panel groups through LDS (each group
 *   loaded once per workgroup); 0: one wavefront per item, each loading its own.
 *   The products are bitwise the same (the autotuner picks the faster).
 * HMG_OPT_EXACT_BOXES [0]: cluster bounding boxes of the tree.  0: the
 *   reference's lookup table (tree.cu:852-1147), bit for bit, including its
 *   behaviour on a level that holds a one-point cluster, where the boxes of the
 *   following clusters are shifted and blocks that are not separated can be taken
 *   as admissible (large approximation errors on clustered clouds).  1: every
 *   cluster's box is the exact min/max of its points; the block list is then the
 *   reference's wherever no level holds a one-point cluster.  Takes effect at the
 *   next setup_h_matrix.
 * HMG_OPT_NEAR_WORKGROUP [0]: mixed precision: 1 runs the near field as
 *   workgroups of two wavefronts (two row slices of one column chunk) that share
 *   each step's X panel groups through LDS; 0: one wavefront per row slice.  The
 *   products are bitwise the same (the autotuner picks the faster).
 * HMG_OPT_FP64_BLOCK_RHS [2]: FP64 products with at least this many right-hand
 *   sides run in passes of 16 or 32 as FP64 MFMA products (each factor element
 *   read once per pass, each near-field entry evaluated once per pass) instead of
 *   one right-hand side at a time (symmetric near field).  A right-hand side's
 *   result is bitwise the same in every pass width; between the two paths it
 *   agrees to rounding.  1 sends every FP64 product through the block path.
 * HMG_OPT_LR_PRIORITY [0]: mixed precision: raises the issue priority of the
 *   low-rank wavefronts over the near field's that run beside them on the same
 *   compute units, bit 0 for V^T X, bit 1 for U T (results unchanged).
 * HMG_OPT_MIXED_ACA_PARTS [16]: mixed precision: the FP64 ACA runs in batches whose
 *   FP64 factors (before the split into BF16 panels) fill at most 1/parts of the
 *   factors' size (at least 1 GiB); results unchanged.
 * HMG_OPT_ACA_LDS_BYTES [0]: unused LDS each workgroup of the main ACA launch
 *   reserves, capping the blocks in flight per compute unit (their factor columns,
 *   re-read at every rank step, then stay in cache); results unchanged. */
#define HMG_OPT_NEAR_SYMMETRIC 1
#define HMG_OPT_NEAR_LDS_BYTES 2
#define HMG_OPT_SERIALIZE 3
#define HMG_OPT_PRECISION 4
#define HMG_OPT_ACA_ROW_SCAN 5
#define HMG_OPT_ACA_BIG_BLOCK 6
#define HMG_OPT_ACA_TOLERANCE 7
#define HMG_OPT_ACA_HUGE_BLOCK 8
#define HMG_OPT_LR_WORKGROUP 9
#define HMG_OPT_EXACT_BOXES 10
#define HMG_OPT_NEAR_WORKGROUP 11
#define HMG_OPT_FP64_BLOCK_RHS 12
#define HMG_OPT_LR_PRIORITY 13
#define HMG_OPT_MIXED_ACA_PARTS 14
#define HMG_OPT_ACA_LDS_BYTES 15
#define HMG_PREC_FP64 0
#define HMG_PREC_MIXED 1
hmg_status hmg_set_option(struct h_matrix_data* data, int option, int64_t value);
/* On-disk / wire format of a built H-matrix (no counterpart in the reference,
 * whose only persistence is debug text dumps, tree.cu:173-183, helper.cu:203-255).
 * hmg_save writes the parameters, the kernel, the points in the caller's order,
 * the effective ranks and the precomputed ACA factors (FP64, or the mixed
 * precision BF16 panels) with a checksum; it needs precompute_aca first.
 * hmg_load takes an uninitialised h_matrix_data, rebuilds the identical block list
 * and layout (setup is deterministic and bit-exact), checks them against the file
 * and installs the factors: products are bitwise those of the saved H-matrix.
 * The loaded H-matrix owns its kernel object (data->assem), freed by
 * destroy_h_matrix_data.  A file that is truncated, of another version or whose
 * checksum fails is refused with HMG_ERR_INVALID before any device memory is
 * touched for the header and points, and with the H-matrix destroyed otherwise. */
hmg_status hmg_save(const struct h_matrix_data* data, const char* path);
hmg_status hmg_load(struct h_matrix_data* data, const char* path);
/* Per-kernel timing with HIP events recorded on the engine's streams around each
 * launch.  enable != 0 starts (and clears) recording.  Kernel classes: 0 = x
 * permutation, 1 = low-rank V^T x, 2 = near field, 3 = ACA, 4 = near-field fill,
 * 5 = low-rank U t, 6 = per-row combine.  Returns the summed milliseconds and the
 * launch count. */
hmg_status hmg_set_profiling(struct h_matrix_data* data, int enable);
hmg_status hmg_get_profile(const struct h_matrix_data* data, int kernel_class, double* total_ms,
                           int64_t* launches);
/* Measurement helper: a plain device copy dst[i] = src[i] of n doubles with the
 * engine's streaming launch shape, used to calibrate HBM byte counters. */
hmg_status hmg_copy_probe(const double* src, double* dst, int64_t n);
/* last error message of this thread */
const char* hmg_last_error(void);
/* Device buffers of 8 GiB and more (factors, panels, scratch) stay mapped after
 * their H-matrix releases them and are reused by later large allocations on the
 * same device (re-mapping device memory costs about 28 ms per GB on MI355X); this
 * returns them to the devices.  An allocation that fails returns them itself. */
void hmg_release_cached_memory(void);
/* sizeof / field offsets of the public structs, for FFI layout checks: which =
 * 0 h_matrix_data, 1 work_item, 2 point_set, 3 morton_code, 4 system_assembler,
 * 5 mat_vec_data_info, 6 hmg_server_config, 7 hmg_build_request, 8
 * hmg_request_info; field < 0 gives sizeof */
int64_t hmg_struct_layout(int which, int field);
/* seeded host generator shared by tests and benchmarks: SplitMix64 seeding
 * xoshiro256**; uniform = (next>>11)*2^-53; normal = Box-Muller. */
void hmg_rng_uniform(uint64_t seed, int64_t n, double* out);
void hmg_rng_normal(uint64_t seed, int64_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif
