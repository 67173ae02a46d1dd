/*
 * hmglib_flow.c -- the reference's own driver sequence (src/hmglib_test.cu:25-142)
 * written against the drop-in C API, for the GPU tests.
 *
 * It calls only what a caller of hmglib calls: init_h_matrix_data, the
 * coordinates written into data.coords_d, setup_h_matrix, the kernel object
 * created through a DEVICE slot and copied back (hmglib_test.cu:84-90),
 * precompute_aca, apply_full_mvp, apply_h_matrix_mvp, plus
 * apply_h_matrix_mvp_without_batching, precompute_dense and
 * set_gaussian_kernel_rhs, then the destroy calls.  The cuRAND inputs of the
 * reference driver (hmglib_test.cu:68-74, 112-115) are replaced by points and a
 * vector read from a file, so that the test can check every output against the
 * CPU oracle on the same inputs.
 *
 * Built as a shared library (tests/ref_flow/Makefile) and called in-process by
 * tests/test_gpu_ref_flow.py: hmglib_flow_run(in_path, out_path) returns 0 on
 * success.  The void entry points exit the process on an engine error, as the
 * reference does.
 *
 * Input (little-endian): int32 n0, n1, dim, bits, k, c_leaf, max_level,
 * max_batched_aca_size, kernel (1 = Gaussian exp(-r^2), 2 = Matern);
 * double eta, epsilon, regularization; double coords of set 0 (dim x n0, row d
 * at d*n0) and of set 1 (dim x n1); double x (n1).
 * Output: int32 dense_count, aca_count, total_count, dense_batch_count,
 * aca_batch_count, max_batched_dense_size, dense_work_size[dense_batch_count],
 * aca_work_size[aca_batch_count]; double rel_err; double y_full, y, y_unbatched,
 * y_stored_dense, b (n0 each); double x after the products (n1).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "hmglib.h"
#include "kernel_system_assembler.h"

static int rd(FILE* f, void* p, size_t n)
{
	return fread(p, 1, n, f) == n ? 0 : -1;
}

static int wr(FILE* f, const void* p, size_t n)
{
	return fwrite(p, 1, n, f) == n ? 0 : -1;
}

static double* dev_vector(size_t n)
{
	double* p = NULL;
	if (hipMalloc((void**)&p, n * sizeof(double)) != hipSuccess) return NULL;
	return p;
}

static int to_host(double* h, const double* d, size_t n)
{
	return hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

int hmglib_flow_run(const char* in_path, const char* out_path)
{
	FILE* in = fopen(in_path, "rb");
	if (!in) return 1;
	int hdr[9];
	double par[3];
	if (rd(in, hdr, sizeof(hdr)) || rd(in, par, sizeof(par))) {
		fclose(in);
		return 2;
	}
	int point_count[2] = {hdr[0], hdr[1]};
	const int dim = hdr[2], bits = hdr[3];
	const size_t n0 = (size_t)point_count[0], n1 = (size_t)point_count[1];
	double* pts[2];
	pts[0] = (double*)malloc(sizeof(double) * dim * n0);
	pts[1] = (double*)malloc(sizeof(double) * dim * n1);
	double* xh = (double*)malloc(sizeof(double) * n1);
	if (rd(in, pts[0], sizeof(double) * dim * n0) || rd(in, pts[1], sizeof(double) * dim * n1) ||
	    rd(in, xh, sizeof(double) * n1)) {
		fclose(in);
		return 3;
	}
	fclose(in);

	/* hmglib_test.cu:43-66 */
	struct h_matrix_data data;
	init_h_matrix_data(&data, point_count, dim, bits);
	data.eta = par[0];
	data.max_level = hdr[6];
	data.c_leaf = hdr[5];
	data.k = hdr[4];
	data.epsilon = par[1];
	data.max_batched_dense_size = 8192;
	data.max_batched_aca_size = hdr[7];

	/* the points, where the reference draws them with cuRAND (hmglib_test.cu:68-74) */
	for (int s = 0; s < 2; s++)
		for (int d = 0; d < dim; d++)
			if (hipMemcpy(data.coords_d[s][d], pts[s] + (size_t)d * (s ? n1 : n0),
			              sizeof(double) * (s ? n1 : n0), hipMemcpyHostToDevice) != hipSuccess)
				return 4;

	/* hmglib_test.cu:78 */
	setup_h_matrix(&data);

	/* kernel object through a device slot (hmglib_test.cu:82-90) */
	const double regularization = par[2];
	struct system_assembler* assem_d = NULL;
	void* slot = NULL;
	if (hipMalloc(&slot, sizeof(void*)) != hipSuccess) return 5;
	if (hdr[8] == 2) {
		create_matern_kernel_system_assembler_object((struct matern_kernel_system_assembler**)slot, regularization);
	} else {
		create_gaussian_kernel_system_assembler_object((struct gaussian_kernel_system_assembler**)slot,
		                                               regularization);
	}
	if (hipMemcpy(&assem_d, slot, sizeof(assem_d), hipMemcpyDeviceToHost) != hipSuccess) return 6;
	data.assem = assem_d;

	/* hmglib_test.cu:93 */
	precompute_aca(&data);

	/* hmglib_test.cu:96-119 */
	double* x = dev_vector(n1);
	double* y = dev_vector(n0);
	double* y_test = dev_vector(n0);
	double* y_nb = dev_vector(n0);
	double* y_st = dev_vector(n0);
	double* b = dev_vector(n0);
	if (!x || !y || !y_test || !y_nb || !y_st || !b) return 7;
	if (hipMemcpy(x, xh, sizeof(double) * n1, hipMemcpyHostToDevice) != hipSuccess) return 8;
	apply_full_mvp(x, y_test, &data);
	apply_h_matrix_mvp(x, y, &data);
	apply_h_matrix_mvp_without_batching(x, y_nb, &data);
	precompute_dense(&data);
	apply_h_matrix_mvp(x, y_st, &data);
	set_gaussian_kernel_rhs(b, &data);

	double* out_v[5];
	for (int q = 0; q < 5; q++) out_v[q] = (double*)malloc(sizeof(double) * n0);
	double* x_after = (double*)malloc(sizeof(double) * n1);
	if (to_host(out_v[0], y_test, n0) || to_host(out_v[1], y, n0) || to_host(out_v[2], y_nb, n0) ||
	    to_host(out_v[3], y_st, n0) || to_host(out_v[4], b, n0) || to_host(x_after, x, n1))
		return 9;

	/* hmglib_test.cu:122-128: relative error of the H-matrix product */
	double num = 0.0, den = 0.0;
	for (size_t i = 0; i < n0; i++) {
		const double e = out_v[0][i] - out_v[1][i];
		num += e * e;
		den += out_v[0][i] * out_v[0][i];
	}
	const double rel_err = sqrt(num) / sqrt(den);
	printf("Relative error in H matrix matrix-vector product: %le\n", rel_err);

	FILE* out = fopen(out_path, "wb");
	if (!out) return 10;
	int info[6] = {data.mat_vec_info.dense_count, data.mat_vec_info.aca_count, data.mat_vec_info.total_count,
	               data.dense_batch_count, data.aca_batch_count, data.max_batched_dense_size};
	int bad = wr(out, info, sizeof(info)) || wr(out, data.dense_work_size, sizeof(int) * data.dense_batch_count) ||
	          wr(out, data.aca_work_size, sizeof(int) * data.aca_batch_count) || wr(out, &rel_err, sizeof(double));
	for (int q = 0; q < 5 && !bad; q++) bad = wr(out, out_v[q], sizeof(double) * n0);
	if (!bad) bad = wr(out, x_after, sizeof(double) * n1);
	fclose(out);

	/* hmglib_test.cu:131-141 */
	hipFree(b);
	hipFree(y_st);
	hipFree(y_nb);
	hipFree(y_test);
	hipFree(x);
	hipFree(y);
	if (hdr[8] == 2)
		destroy_matern_kernel_system_assembler_object((struct matern_kernel_system_assembler**)slot);
	else
		destroy_gaussian_kernel_system_assembler_object((struct gaussian_kernel_system_assembler**)slot);
	hipFree(slot);
	destroy_h_matrix_data(&data);

	for (int q = 0; q < 5; q++) free(out_v[q]);
	free(x_after);
	free(xh);
	free(pts[0]);
	free(pts[1]);
	return bad ? 11 : 0;
}
