/* Test infrastructure only: single-precision instantiation of port_core.h. */
#define R float
#define P(x) pkp_s_##x
#define EPS_R 5.9604644775390625e-08f       /* 2^-24, slamch('e') */
#define SFMIN_R 1.17549435e-38f             /* 2^-126, slamch('s') */
#define GEMV_ sgemv_
#define GEMM_ sgemm_
#define BDSQR_ sbdsqr_
#define BDSDC_ sbdsdc_
#define SQRT sqrtf
#define ABS fabsf
#include "port_core.h"
