// GEMM kernels: FP32 MFMA on FP32 storage (precision mode 2), and the
// register-direct kernel on FP32 storage.
#define CCQ_GEMM_T float
#define CCQ_GEMM_IO float
#define CCQ_GEMM_SUFFIX f32io
#define CCQ_GEMM_REG 1
#include "gemm_inst.inc"
