/*
 * kernel_system_assembler.h -- kernel objects (reference:
 * src/kernel_system_assembler.h:26-102).  As in the reference, `assem` is a
 * DEVICE pointer to a slot that receives the DEVICE address of the new object;
 * the caller copies that address back and stores it in h_matrix_data.assem.
 */
#ifndef kernel_system_assembler_h__
#define kernel_system_assembler_h__

#include "system_assembler.h"

#ifdef __cplusplus
struct gaussian_kernel_system_assembler : system_assembler {};
struct matern_kernel_system_assembler : system_assembler {};
extern "C" {
#else
struct gaussian_kernel_system_assembler { struct system_assembler base; };
struct matern_kernel_system_assembler { struct system_assembler base; };
#endif

/* replaces kernel_system_assembler.h:66 (exp(-r*r) + regularization on i==j) */
extern void create_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem, double regularization);
/* replaces kernel_system_assembler.h:68 */
extern void destroy_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem);
/* replaces kernel_system_assembler.h:100 */
extern void create_matern_kernel_system_assembler_object(struct matern_kernel_system_assembler** assem, double regularization);
/* replaces kernel_system_assembler.h:102 */
extern void destroy_matern_kernel_system_assembler_object(struct matern_kernel_system_assembler** assem);

/* extension: any HMG_KERNEL_* kind with a length scale, same slot protocol */
extern int hmg_create_kernel_object(struct system_assembler** assem, int kind, double regularization, double length_scale);
extern int hmg_destroy_kernel_object(struct system_assembler** assem);
/* extension: returns the device address directly (no slot) or NULL */
extern struct system_assembler* hmg_kernel_new(int kind, double regularization, double length_scale);
extern void hmg_kernel_free(struct system_assembler* assem);

#ifdef __cplusplus
}
#endif

#endif
