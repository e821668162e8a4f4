// Launch interface of the gfx950 kernels (ccq_amd/csrc/kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

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
  // C is contiguous along m (rowC[m] == m) / along n (colC[n] == n)
  int c_mfast = 0, c_nfast = 0;
  // kA[k] == k * kA_aff for every k when kA_aff >= 0 (same for B)
  int64_t kA_aff = -1, kB_aff = -1;
  // launch plan (gemm_plan)
  int kind = 0;            // 0: LDS-tiled kernel, 1: register-direct kernel
  int trans = 0;           // run as C^T = B^T A^T (register-direct kernel)
  int bm = 64, bn = 64, bk = 16;      // LDS-tiled tile
  int rm = 1, rn = 1, wk = 1;         // register-direct wave tile / waves along K
  int nsplit = 1;
  int64_t kchunk = 0;
  int group = 1;           // tile rows per group of the XCD-aware tile order
  // arithmetic of the LDS-tiled kernel: 0 FP64 MFMA; 1 FP32 MFMA (operands
  // rounded to FP32 when staged into LDS, FP32 accumulation, FP64 epilogue)
  int prec = 0;
  // LDS-tiled kernel: 2 = two 4-wave groups share a tile and split its K
  // slabs (tiles up to 64x64), 1 = one group
  int ksplit = 1;
  // split-K workspace of at least gemm_ws_need doubles
  double* ws = nullptr;
  // optional addend folded into the epilogue: C (+)= alpha A B + dcoef D,
  // D read at D[rowD[m] + colD[n]]
  const double* D = nullptr;
  const int* rowD = nullptr;
  const int* colD = nullptr;
  double dcoef = 0.0;
  // mirrored images of an output antisymmetric in a packed pair: the value
  // is also stored, negated, at rowC2[m] + colC[n] and rowC[m] + colC2[n]
  // (and at rowC2[m] + colC2[n] unnegated when both are set)
  const int* rowC2 = nullptr;
  const int* colC2 = nullptr;
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
  // masked update (active-space triples): only tuples with an occupied index
  // >= occ_start, an active virtual (index < vir_end[spin]) of a spin in vneed
  // and no active virtual of a spin in vforbid are updated
  int mask = 0;
  int vneed = 0, vforbid = 0;
  int64_t occ_start = 0;
  int64_t vir_end[2] = {0, 0};
  // canonical tuples (indices ascending inside every antisymmetric group):
  // the occupied halves coc[n_oc][rank/2] and the virtual halves
  // cvc[n_vc][rank/2]; the kernel runs over total = n_oc * n_vc of them
  const int* coc = nullptr;
  const int* cvc = nullptr;
  int64_t n_vc = 1;
  // FP32-storage sweeps: U holds floats
  int u32 = 0;
};

// CC(t;3) triples correction of one spin case (crcc23, src/mm_correct.f90:11-343).
// The block has extents dims[0..5] (occupied, occupied, occupied, virtual x3)
// with spin pattern cs: 0 aaa, 1 aab, 2 abb, 3 bbb.  L3 / M3 are residual
// accumulators with images folded (the value at a tuple is the image average,
// as in jacobi_update).  Only canonical tuples outside the P space count.
// Diagonals, 0-based correlated indices [occ | vir] per spin:
//   fd[s]: Fock; hd[s]: H-bar one-body.
// Two-body H-bar blocks (dense C order) whose diagonals enter D_C / D_D:
//   ovov[so][sv]: spins (so, sv, so, sv); oooo / vvvv [0 aaaa, 1 abab, 2 bbbb].
// d3[0..7]: cal_d3's d3a1, d3a2, d3b1, d3b2, d3c1, d3c2, d3d1, d3d2.
struct Cct3Params {
  int cs;
  int nimg;
  int64_t total;
  int64_t dims[6];
  int64_t strides[6];
  int64_t oa, ob, va, vb;
  int64_t occ_start;
  int64_t vir_end[2];
  const double* fd[2];
  const double* hd[2];
  const double* ovov[2][2];
  const double* oooo[3];
  const double* vvvv[3];
  const double* d3[8];
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

// Choose kernel, tile shape and split-K for a GEMM; ws_cap bounds the
// split-K workspace (doubles).
void gemm_plan(GemmArgs& g, int64_t ws_cap);
// the same, also returning every candidate plan with its modelled cycles,
// best first (the input for the measuring autotuner)
void gemm_plan_ranked(GemmArgs& g, int64_t ws_cap, std::vector<std::pair<double, GemmArgs>>* ranked);
// Output tiles of the planned launch, and split-K workspace (doubles) it
// needs (0 when nsplit == 1).
int64_t gemm_tiles(const GemmArgs& g);
int64_t gemm_ws_need(const GemmArgs& g);
hipError_t launch_gemm(hipStream_t s, const GemmArgs& g);

// Several independent GEMMs in one launch of the LDS-tiled kernel: one grid
// covers the tiles of every member (full K each, no split-K), so products
// whose own tiles would leave CUs idle fill the chip together.  Members must
// share the kernel instantiation (gemm_group_key) and must not write
// anything another member reads or writes.
constexpr int kMaxGroup = 8;
struct GemmGroup {
  GemmArgs g[kMaxGroup];
  int start[kMaxGroup + 1];   // first logical tile of each member
  int n;
};
// kernel instantiation of a plan (after transposition), or -1 when the plan
// cannot join a group (register kernel, split-K)
int64_t gemm_group_key(const GemmArgs& g);
hipError_t launch_gemm_group(hipStream_t s, const GemmArgs* gs, int n);

// A chain of GEMMs accumulating into one output with the same tile mapping
// and the same operand layouts (k-contiguous or not, after transposition),
// run as one product over the concatenated K ranges (LDS-tiled kernel, BK 32,
// split-K over the concatenated K).  Entry i contributes alpha_i A_i B_i
// (its trans flag: run transposed to match the output orientation); the
// output arguments (C, its tables, accumulate, the folded addend) and the
// launch plan (tile, nsplit, kchunk over the concatenated K, group, ws) are
// the first entry's.
constexpr int kMaxChain = 8;
struct GemmChain {
  GemmArgs e[kMaxChain];
  int64_t koff[kMaxChain + 1];   // first concatenated k of each entry
  int n;
};
hipError_t launch_gemm_chain(hipStream_t s, const GemmArgs* es, int n);

// Independent dense block copies / accumulations C_i (+)= alpha_i A_i over
// identical layouts, in one launch (the small adds of one dependency level).
constexpr int kMaxAddGroup = 16;
struct AddGroup {
  const double* A[kMaxAddGroup];
  double* C[kMaxAddGroup];
  double alpha[kMaxAddGroup];
  int accumulate[kMaxAddGroup];
  int64_t start[kMaxAddGroup + 1];   // first element of each member
  int n;
};
// io32: the arrays hold floats (FP32-storage sweeps)
hipError_t launch_axpy_group(hipStream_t s, const AddGroup& ag, int io32 = 0);
// up to kMaxPermGroup permuted adds C_i(perm) (+)= alpha_i A_i(perm) with
// their own offset tables in one launch (the irrep blocks of one add)
constexpr int kMaxPermGroup = 8;
struct PermGroup {
  const double* A[kMaxPermGroup];
  const int* rowA[kMaxPermGroup];
  const int* colA[kMaxPermGroup];
  double* C[kMaxPermGroup];
  const int* rowC[kMaxPermGroup];
  const int* colC[kMaxPermGroup];
  int64_t N[kMaxPermGroup];
  double alpha[kMaxPermGroup];
  int accumulate[kMaxPermGroup];
  int64_t start[kMaxPermGroup + 1];   // first element (row-major M x N) of each member
  int n;
};
hipError_t launch_perm_group(hipStream_t s, const PermGroup& pg, int io32 = 0);
hipError_t launch_perm_add(hipStream_t s, int64_t M, int64_t N, const double* A, const int* rowA,
                           const int* colA, double* C, const int* rowC, const int* colC,
                           double alpha, int accumulate, int io32 = 0);
// Y[rowY[m] + colY[n]] = scale * X[rowX[m] + colX[n]] (int64 tables; view
// set-up, packing of the active-space t_vec)
hipError_t launch_gather64(hipStream_t s, int64_t M, int64_t N, const double* X,
                           const int64_t* rowX, const int64_t* colX, double* Y,
                           const int64_t* rowY, const int64_t* colY, double scale = 1.0);
// C (+)= alpha A for identically laid out dense arrays of n doubles (16-byte aligned)
hipError_t launch_axpy_flat(hipStream_t s, int64_t n, const double* A, double* C, double alpha,
                            int accumulate, int io32 = 0);
// y = (float)x (the FP32 shadows of the FP32-storage sweeps)
hipError_t launch_cast_f32(hipStream_t s, float* y, const double* x, int64_t n);
hipError_t launch_fill(hipStream_t s, double* x, int64_t n, double v);
hipError_t launch_copy(hipStream_t s, double* y, const double* x, int64_t n);
// MFMA issue-rate probe; out holds blocks * 256 doubles; *flops = work of the launch
hipError_t launch_mfma_ceiling(hipStream_t s, int kind, int64_t blocks, int64_t iters, double* out,
                               double* flops);
// sumsq (optional): per-block partial sums of |G| R^2 over the canonical
// tuples updated (the squared norm of the full-storage residual); one entry
// per block of the launch, update_blocks(P) of them.
hipError_t launch_update(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                         const int* perms, const double* signs, const double* fa,
                         const double* fb, double* sumsq = nullptr);
int update_blocks(const UpdateParams& P);
// out[m] = scale * sum_{k < K} X[xr[m] + xk[k]] * Y[yr[m] + yk[k]]
hipError_t launch_hadamard_sum(hipStream_t s, int64_t M, int64_t K, const double* X,
                               const int64_t* xr, const int64_t* xk, const double* Y,
                               const int64_t* yr, const int64_t* yk, double* out, double scale);
// partial[block * 4 + x]: per-block sums of l3 * m3 / D_x, x = A, B, C, D
hipError_t launch_cct3(hipStream_t s, const Cct3Params& P, const double* L3, const double* M3,
                       const int* perms, const double* signs, double* partial, int blocks);
// dst[i] = src[i * (ld + 1)], i < n (diagonal of a square block with row stride ld)
hipError_t launch_diag(hipStream_t s, const double* src, int64_t n, int64_t ld, double* dst);
// ---- learned starting guess (guess.hip) ----
// One T2 spin block t[i][j][a][b] (extents o1 o2 v1 v2) and the integrals it
// is guessed from: V[p][q][r][s] full correlated range, C order, occupied
// first per spin; the virtual a (b) sits at full index va0 + a (vb0 + b).
// f1 / f2: [occ | vir] Fock diagonals of the spins of (i, a) / (j, b).
struct GuessBlock {
  const double* V;
  int64_t nc;
  int64_t o1, o2, v1, v2;
  int64_t va0, vb0;
  const double* f1;
  const double* f2;
  double inv_sm;        // 1 / rms of the MP2 amplitudes of the block
  double* t;            // the block inside t_vec
  int anti = 0;         // same-spin block: evaluate i<j, a<b and write the images
};
constexpr int kGuessFeatures = 5;
constexpr int kGuessHidden = 64;
// packed float parameters: mu[5] sd[5] W1[5][64] b1[64] W2[64][64] b2[64] W3[64] b3
constexpr int kGuessParams = 2 * kGuessFeatures + kGuessFeatures * kGuessHidden + kGuessHidden +
                             kGuessHidden * kGuessHidden + kGuessHidden + kGuessHidden + 1;
// per-block partial sums of m^2 (m = V / D, the MP2 amplitude)
hipError_t launch_guess_msq(hipStream_t s, const GuessBlock& B, double* partial, int blocks);
// t = m (params == nullptr) or t = m * (1 + r(features)) with the BF16 MLP
hipError_t launch_guess_apply(hipStream_t s, const GuessBlock& B, const float* params);
// training data of the block at its current (converged) t: features X[n][5],
// M = m / sm, Y = (t - m) / sm, for every tuple in C order
hipError_t launch_guess_data(hipStream_t s, const GuessBlock& B, float* X, float* M, float* Y);
hipError_t launch_energy_pairs(hipStream_t s, const EnergyParams& P, double* partial, int blocks);
hipError_t launch_dot_strided(hipStream_t s, int64_t rows, int64_t cols, const double* a,
                              int64_t lda, const double* b, double* partial, int blocks);
hipError_t launch_diis_gram(hipStream_t s, const double* ring, int64_t len, int n, double* partial,
                            int blocks);
// dst = src with diff_sq(prev, dst)'s partials in the same pass
hipError_t launch_copy_diff_sq(hipStream_t s, double* dst, const double* src, const double* prev,
                               int64_t len, double* partial, int blocks);
hipError_t launch_diff_sq(hipStream_t s, const double* a, const double* b, int64_t len,
                          double* partial, int blocks);
hipError_t launch_lincomb(hipStream_t s, double* out, const double* ring, int64_t len, int n,
                          const Coeffs& c);

}  // namespace ccqk
