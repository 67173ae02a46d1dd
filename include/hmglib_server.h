/*
 * hmglib_server.h -- continuous-batching request loop over H-matrices cached in
 * HBM.  No counterpart in the reference (hmglib is a synchronous library); it
 * serves build and apply requests of different sizes as they arrive.
 *
 *   - A build request (points, tree and kernel parameters) creates an H-matrix
 *     that stays cached on the device; its id is returned at submit.
 *   - An apply request Y = H X names a cached H-matrix.  Apply requests that wait
 *     on the same H-matrix are merged into one product pass of up to
 *     max_rhs_per_pass right-hand sides (each right-hand side is computed
 *     independently, so a result is bitwise the same alone and in the mix).
 *   - Up to `slots` jobs (passes or builds on different H-matrices) run on the
 *     device at once, each on its H-matrix's own streams; requests arriving while
 *     others run are admitted as slots free up.  Builds never hold the last
 *     product_slots of them, so products need not wait behind second-long builds
 *     for a slot.
 *   - A build is admitted when its device memory (measured after the tree is
 *     built, predicted for the factors and product buffers) fits in
 *     hbm_budget_bytes next to the cached H-matrices; least recently used idle
 *     H-matrices without pending applies are evicted to make room.
 * All calls are thread-safe.  Times are seconds on a steady clock since the
 * server was created.
 */
#ifndef HMGLIB_SERVER_H
#define HMGLIB_SERVER_H

#include <stdint.h>
#include "hmglib.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hmg_server hmg_server;

typedef struct {
	int64_t hbm_budget_bytes;   /* admission budget (default 200e9) */
	int slots;                  /* jobs on the device at once (default 4) */
	int max_rhs_per_pass;       /* right-hand sides merged into one pass (default 32) */
	int precision;              /* HMG_PREC_FP64 or HMG_PREC_MIXED (default mixed) */
	int product_slots;          /* slots builds never take, kept for product passes
	                               (0 .. slots - 1; default 0) */
} hmg_server_config;

typedef struct {
	int dim;
	int bits;                   /* Morton bits per dimension (as init_h_matrix_data) */
	int n[2];                   /* target and source point counts */
	/* host arrays, dim x n[s] row-major (coordinate d of point p at [d*n[s] + p]);
	 * coords[1] == coords[0] with n[1] == n[0] builds on one point set.  Copied at
	 * submit. */
	const double* coords[2];
	double eta;
	int max_level;
	int c_leaf;
	int k;                      /* ACA rank */
	int kernel_kind;            /* HMG_KERNEL_* */
	double regularization;
	double length_scale;
	double epsilon;             /* ACA tolerance (used when aca_tolerance != 0) */
	int aca_tolerance;          /* HMG_OPT_ACA_TOLERANCE mode: 0 fixed rank k, 1, 2 */
	int exact_boxes;            /* HMG_OPT_EXACT_BOXES: 0 the reference's cluster boxes, 1 exact */
} hmg_build_request;

typedef struct {
	int status;                 /* HMG_OK or an error; -1 while pending or running */
	double t_submit;
	double t_start;             /* job start on a slot (-1 before) */
	double t_end;               /* results complete on the device (-1 before) */
	int merged;                 /* requests that shared this request's pass */
	int64_t matrix;             /* the H-matrix the request built or applied */
	char message[160];          /* why a failed request failed ("" otherwise) */
} hmg_request_info;

/* cfg may be NULL (defaults) */
hmg_status hmg_server_create(const hmg_server_config* cfg, hmg_server** out);
/* waits for every submitted request, then frees all cached H-matrices */
hmg_status hmg_server_destroy(hmg_server* s);
hmg_status hmg_server_submit_build(hmg_server* s, const hmg_build_request* req, int64_t* request,
                                   int64_t* matrix);
/* X (n[1] x nrhs) and Y (n[0] x nrhs): column-major FP64 DEVICE arrays in the
 * caller's point order, X fully written before submit, both valid until the
 * request completes.  Fails with HMG_ERR_NOT_READY when the H-matrix is unknown,
 * evicted or released; an apply submitted before its build finishes waits for it. */
hmg_status hmg_server_submit_apply(hmg_server* s, int64_t matrix, const double* X, double* Y, int nrhs,
                                   int64_t* request);
/* blocks until the request completes */
hmg_status hmg_server_wait(hmg_server* s, int64_t request, hmg_request_info* info);
/* non-blocking */
hmg_status hmg_server_query(hmg_server* s, int64_t request, hmg_request_info* info);
/* blocks until every submitted request has completed */
hmg_status hmg_server_drain(hmg_server* s);
/* drops a cached H-matrix once its pending applies are done */
hmg_status hmg_server_release(hmg_server* s, int64_t matrix);
/* out[0] = device bytes of cached and building H-matrices, out[1] = cached
 * H-matrices, out[2] = evictions, out[3] = completed requests, out[4] = failed
 * requests, out[5] = product passes, out[6] = apply requests served in passes
 * shared with others, out[7] = builds completed */
hmg_status hmg_server_stats(hmg_server* s, int64_t out[8]);
/* ids of cached, built H-matrices (up to cap); returns the count */
int hmg_server_resident(hmg_server* s, int64_t* ids, int cap);

/* Device memory of one H-matrix (no counterpart in the reference): out[0] =
 * bytes it holds now; out[1] = bytes precompute_aca and the first product in
 * the current precision would add (a prediction from the tree and schedule). */
hmg_status hmg_memory(const struct h_matrix_data* data, int64_t out[2]);

#ifdef __cplusplus
}
#endif

#endif
