// Launch interface of the gfx950 kernels (ccq_amd/csrc/kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccqk {

constexpr int kMaxRank = 8;
constexpr int kMaxDiis = 12;

// Offset tables are int32: every tensor the engine addresses holds < 2^31
// elements (checked when the tables are built).
struct GemmArgs {
  int64_t M, N, K;
  const double* A;
  const int* rowA;
  const int* kA;
  const double* B;
  const int* kB;
  const int* colB;
  double* C;
  const int* rowC;
  const int* colC;
  double alpha;
  int accumulate;
  int a_kfast;
  int b_kfast;
  // launch plan (gemm_plan)
  int bm = 64, bn = 64;
  int nsplit = 1;
  int64_t kchunk = 0;
};

struct UpdateParams {
  int rank;
  int nimg;
  int64_t total;
  int64_t dims[kMaxRank];
  int64_t strides[kMaxRank];
  int group[kMaxRank];   // antisymmetric group id per position, -1 if none
  int isvir[kMaxRank];
  int spin[kMaxRank];
  int64_t occ[kMaxRank]; // occupied count of the position's spin (virtual offset)
  double shift;
};

struct EnergyParams {
  const double* V;       // full-range base array
  int64_t vs0, vs1, vs2, vs3;   // strides of V
  int64_t vo1, vo2;      // virtual offsets of the two virtual indices
  const double* t2;      // [m][n][e][f]
  const double* t1x;     // [m][e]
  const double* t1y;     // [n][f]
  int64_t o1, o2, v1, v2;
  double w2, w11;
};

struct Coeffs {
  double c[kMaxDiis];
};

// Choose tile shape and split-K for a GEMM; ws_cap bounds nsplit*M*N.
void gemm_plan(GemmArgs& g, int64_t ws_cap);
// ws: split-K workspace of at least nsplit*M*N doubles (unused when nsplit == 1).
hipError_t launch_gemm(hipStream_t s, const GemmArgs& g, double* ws);
hipError_t launch_perm_add(hipStream_t s, int64_t M, int64_t N, const double* A,
                           const int64_t* rowA, const int64_t* colA, double* C,
                           const int64_t* rowC, const int64_t* colC, double alpha, int accumulate);
hipError_t launch_fill(hipStream_t s, double* x, int64_t n, double v);
hipError_t launch_update(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                         const int* perms, const double* signs, const double* fa,
                         const double* fb);
hipError_t launch_energy_pairs(hipStream_t s, const EnergyParams& P, double* partial, int blocks);
hipError_t launch_dot_strided(hipStream_t s, int64_t rows, int64_t cols, const double* a,
                              int64_t lda, const double* b, double* partial, int blocks);
hipError_t launch_diis_gram(hipStream_t s, const double* ring, int64_t len, int n, double* partial,
                            int blocks);
hipError_t launch_diff_sq(hipStream_t s, const double* a, const double* b, int64_t len,
                          double* partial, int blocks);
hipError_t launch_lincomb(hipStream_t s, double* out, const double* ring, int64_t len, int n,
                          const Coeffs& c);

}  // namespace ccqk
