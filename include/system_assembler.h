/*
 * system_assembler.h -- kernel descriptor (replaces the virtual device class of
 * src/system_assembler.h:23-32).  A plain struct in device memory: the engine
 * reads it once per call and evaluates the kernel inline in its own HIP kernels.
 */
#ifndef system_assembler_h__
#define system_assembler_h__

#define HMG_KERNEL_GAUSSIAN_REF 1   /* exp(-r*r)                      (kernel_system_assembler.h:33-36) */
#define HMG_KERNEL_MATERN_REF   2   /* yn(1,r+1e-15) r / (2^{d/2} Gamma(1+d/2))        (:77-80) */
#define HMG_KERNEL_GAUSSIAN     3   /* exp(-r^2 / (2 l^2))  -- length-scale extension */
#define HMG_KERNEL_EXPONENTIAL  4   /* exp(-r / l)          -- length-scale extension */

struct system_assembler
{
	int kind;               /* HMG_KERNEL_* */
	int reserved;
	double regularization;  /* added when sorted row index == sorted column index */
	double length_scale;    /* l for HMG_KERNEL_GAUSSIAN / HMG_KERNEL_EXPONENTIAL */
};

#endif
