/*
 * tree.h -- block-cluster-tree work item (reference: src/tree.h:31-55).
 * Same 192-byte x86-64 layout, so the block list in h_matrix_data.mat_vec_data
 * reads exactly like the reference's.
 */
#ifndef TREE_H
#define TREE_H

#define WT_ACA 1
#define WT_DENSE 2
#define MAX_POINT_DIMENSION 5

/* reference: src/tree.h:38-55 */
struct work_item
{
	int set1_l;
	int set1_u;
	int set2_l;
	int set2_u;
	int work_type;
	double max1[MAX_POINT_DIMENSION];
	double min1[MAX_POINT_DIMENSION];
	double max2[MAX_POINT_DIMENSION];
	double min2[MAX_POINT_DIMENSION];
	int dim;
	char level_1;
	char level_2;
};

#endif
