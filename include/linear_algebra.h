/*
 * linear_algebra.h -- block-list summary (reference: src/linear_algebra.h:93-98).
 */
#ifndef LINEAR_ALGEBRA_H_
#define LINEAR_ALGEBRA_H_

/* reference: src/linear_algebra.h:93-98 */
struct mat_vec_data_info
{
	int dense_count;
	int aca_count;
	int total_count;
};
#ifndef __cplusplus
typedef struct mat_vec_data_info mat_vec_data_info;
#endif

#endif
