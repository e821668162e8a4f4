// GEMM kernels: FP64 MFMA on FP64 storage (precision mode 0), and the
// register-direct kernel on FP64 storage (modes 0 and 1).
#define CCQ_GEMM_T double
#define CCQ_GEMM_IO double
#define CCQ_GEMM_SUFFIX f64
#define CCQ_GEMM_REG 1
#include "gemm_inst.inc"
