/* Test infrastructure only: double-precision instantiation of port_core.h. */
#define R double
#define P(x) pkp_d_##x
#define EPS_R 1.1102230246251565e-16        /* 2^-53, dlamch('e') */
#define SFMIN_R 2.2250738585072014e-308     /* 2^-1022, dlamch('s') */
#define GEMV_ dgemv_
#define GEMM_ dgemm_
#define BDSQR_ dbdsqr_
#define BDSDC_ dbdsdc_
#define SQRT sqrt
#define ABS fabs
#include "port_core.h"
