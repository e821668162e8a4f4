// GEMM kernels: FP32 MFMA on FP64 storage (precision mode 1).
#define CCQ_GEMM_T float
#define CCQ_GEMM_IO double
#define CCQ_GEMM_SUFFIX f32
#include "gemm_inst.inc"
