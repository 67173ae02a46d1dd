/*
 * morton.h -- point-set and Morton-code descriptors (reference: src/morton.h:23-46).
 * Both structs live in device memory and are owned by the library; their fields
 * mirror the reference so that existing readers keep working.
 */
#ifndef MORTON_H
#define MORTON_H

#include <stdint.h>

/* reference: src/morton.h:23-31 */
struct point_set
{
	double** coords;
	int dim;
	int size;
	double* max_per_dim;
	double* min_per_dim;
	unsigned int* point_ids;
};

/* reference: src/morton.h:33-41 */
struct morton_code
{
	uint64_t* code;
	int size;
	int dim;
	int bits;
	double* max_values;
	double* min_values;
};

#define MAX_DIM 20

#endif
