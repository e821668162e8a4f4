// CDNA4 (gfx950) kernels of the coupled-cluster amplitude engine.
//
//  * gemm_f64_tab   FP64 MFMA (v_mfma_f64_16x16x4f64) tile GEMM whose operand
//                   loaders and epilogue address tensors through per-index
//                   offset tables, so any index permutation / sub-block view
//                   of a tensor is read and written in place (no separate
//                   reorder/sum_stripe passes; cf. src/cc/reorder_stripe.f90).
//  * perm_add       C(perm) = beta C + alpha A through the same tables.
//  * jacobi_update  R = A_G(U) over a block's antisymmetry group, t -= R/(D+shift)
//                   on canonical tuples, signed scatter to every image
//                   (src/cc/t2A_update.f90:1885-1901, t1A_update.f90:518-525).
//  * energy_*       FP64 reduction of the CC correlation energy
//                   (src/energy.f90:138-270), deterministic two-level sum.
//  * diis_*         ring Gram matrix, residuum and extrapolation
//                   (src/diis.F90:11-242), deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace ccqk {

constexpr int BM = 64, BN = 64, BK = 16;

__global__ __launch_bounds__(256) void gemm_f64_tab(
    int64_t M, int64_t N, int64_t K,
    const double* __restrict__ A, const int64_t* __restrict__ rowA, const int64_t* __restrict__ kA,
    const double* __restrict__ B, const int64_t* __restrict__ kB, const int64_t* __restrict__ colB,
    double* __restrict__ C, const int64_t* __restrict__ rowC, const int64_t* __restrict__ colC,
    double alpha, int accumulate, int a_kfast, int b_kfast) {
  __shared__ double As[BK][BM];
  __shared__ double Bs[BK][BN];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int wm = (wave & 1) * 32;
  const int wn = (wave >> 1) * 32;

  // per-thread load coordinates (fixed across the k loop)
  int am[4], ak[4], bn[4], bk[4];
  int64_t arow[4], bcol[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = tid + 256 * r;
    if (a_kfast) { ak[r] = e & 15; am[r] = e >> 4; } else { am[r] = e & 63; ak[r] = e >> 6; }
    if (b_kfast) { bk[r] = e & 15; bn[r] = e >> 4; } else { bn[r] = e & 63; bk[r] = e >> 6; }
    const int64_t gm = m0 + am[r];
    const int64_t gn = n0 + bn[r];
    arow[r] = gm < M ? rowA[gm] : -1;
    bcol[r] = gn < N ? colB[gn] : -1;
  }

  d4 acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    double ra[4], rb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t gka = k0 + ak[r];
      ra[r] = (arow[r] >= 0 && gka < K) ? A[arow[r] + kA[gka]] : 0.0;
      const int64_t gkb = k0 + bk[r];
      rb[r] = (bcol[r] >= 0 && gkb < K) ? B[kB[gkb] + bcol[r]] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      As[ak[r]][am[r]] = ra[r];
      Bs[bk[r]][bn[r]] = rb[r];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      const int rr = lane & 15;
      const double a0 = As[kr][wm + rr];
      const double a1 = As[kr][wm + 16 + rr];
      const double b0 = Bs[kr][wn + rr];
      const double b1 = Bs[kr][wn + 16 + rr];
      acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
      acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
      acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
      acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
    }
  }
  // D layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l, register v
  // holds row (l/16) + 4*v, column l%16
  const d4* accs[4] = {&acc00, &acc01, &acc10, &acc11};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = t >> 1, j = t & 1;
    const int64_t gn = n0 + wn + 16 * j + (lane & 15);
    if (gn >= N) continue;
    const int64_t cc = colC[gn];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gm = m0 + wm + 16 * i + (lane >> 4) + 4 * v;
      if (gm < M) {
        double* p = C + rowC[gm] + cc;
        const double val = alpha * (*accs[t])[v];
        if (accumulate) *p += val; else *p = val;
      }
    }
  }
}

__global__ void perm_add(int64_t M, int64_t N, const double* __restrict__ A,
                         const int64_t* __restrict__ rowA, const int64_t* __restrict__ colA,
                         double* __restrict__ C, const int64_t* __restrict__ rowC,
                         const int64_t* __restrict__ colC, double alpha, int accumulate) {
  const int64_t total = M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e / N, n = e - m * N;
    const double v = alpha * A[rowA[m] + colA[n]];
    double* p = C + rowC[m] + colC[n];
    if (accumulate) *p += v; else *p = v;
  }
}

__global__ void fill(double* __restrict__ x, int64_t n, double v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    x[e] = v;
}

// One thread per element of the full block; only canonical tuples (strictly
// increasing inside every antisymmetric group) do work.
__global__ void jacobi_update(UpdateParams P, double* __restrict__ t, const double* __restrict__ U,
                              const int* __restrict__ perms, const double* __restrict__ signs,
                              const double* __restrict__ fdiag_a, const double* __restrict__ fdiag_b) {
  const int rank = P.rank;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P.total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx[8];
    int64_t r = e;
    for (int d = rank - 1; d >= 0; --d) {
      idx[d] = r % P.dims[d];
      r /= P.dims[d];
    }
    bool canon = true;
    for (int d = 0; d + 1 < rank; ++d)
      if (P.group[d] >= 0 && P.group[d] == P.group[d + 1] && idx[d] >= idx[d + 1]) canon = false;
    if (!canon) continue;
    double R = 0.0;
    for (int g = 0; g < P.nimg; ++g) {
      int64_t off = 0;
      for (int d = 0; d < rank; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      R += signs[g] * U[off];
    }
    R /= (double)P.nimg;
    double D = 0.0;
    for (int d = 0; d < rank; ++d)
      if (P.isvir[d]) D += (P.spin[d] ? fdiag_b : fdiag_a)[P.occ[d] + idx[d]];
    for (int d = 0; d < rank; ++d)
      if (!P.isvir[d]) D -= (P.spin[d] ? fdiag_b : fdiag_a)[idx[d]];
    D += P.shift;
    int64_t off0 = 0;
    for (int d = 0; d < rank; ++d) off0 += idx[d] * P.strides[d];
    const double tn = t[off0] - R / D;
    for (int g = 0; g < P.nimg; ++g) {
      int64_t off = 0;
      for (int d = 0; d < rank; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      t[off] = signs[g] * tn;
    }
  }
}

// Energy pieces for one spin case over (m, n, e, f):
//   pair term   w2 * V[e,f,m,n] * t2[m,n,e,f]
//   t1 t1 term  w11 * V[e,f,m,n] * t1x[m,e] * t1y[n,f]
__global__ __launch_bounds__(256) void energy_pairs(EnergyParams P, double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const int64_t total = P.o1 * P.o2 * P.v1 * P.v2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e;
    const int64_t f = r % P.v2; r /= P.v2;
    const int64_t a = r % P.v1; r /= P.v1;
    const int64_t n = r % P.o2; r /= P.o2;
    const int64_t m = r;
    const double v = P.V[(P.vo1 + a) * P.vs0 + (P.vo2 + f) * P.vs1 + m * P.vs2 + n * P.vs3];
    const double t2 = P.t2[((m * P.o2 + n) * P.v1 + a) * P.v2 + f];
    const double t11 = P.t1x[m * P.v1 + a] * P.t1y[n * P.v2 + f];
    acc += v * (P.w2 * t2 + P.w11 * t11);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// sum_x a[x] * b[x] over flat vectors with a generic stride on a (used for f_ov t1)
__global__ __launch_bounds__(256) void dot_strided(int64_t rows, int64_t cols,
                                                   const double* __restrict__ a, int64_t lda,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, j = e - i * cols;
    acc += a[i * lda + j] * b[e];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Gram matrix of ring differences: G[x][y] = <s[x+1]-s[x], s[y+1]-s[y]>,
// x <= y < NS.  partial layout: [block][pair].
template <int NS>
__global__ __launch_bounds__(256) void diis_gram(const double* __restrict__ ring, int64_t len,
                                                 double* __restrict__ partial) {
  constexpr int NP = NS * (NS + 1) / 2;
  __shared__ double red[256];
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (int64_t)gridDim.x * blockDim.x) {
    double d[NS];
#pragma unroll
    for (int x = 0; x < NS; ++x)
      d[x] = ring[(int64_t)(x + 1) * len + e] - ring[(int64_t)x * len + e];
    int q = 0;
#pragma unroll
    for (int x = 0; x < NS; ++x)
#pragma unroll
      for (int y = x; y < NS; ++y) acc[q++] += d[x] * d[y];
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * NP + q] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void diff_sq(const double* __restrict__ a, const double* __restrict__ b,
                                               int64_t len, double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double d = (a ? a[e] : 0.0) - b[e];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void lincomb(double* __restrict__ out, const double* __restrict__ ring, int64_t len,
                        int n, Coeffs c) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += c.c[k] * ring[(int64_t)k * len + e];
    out[e] = s;
  }
}

}  // namespace ccqk

// --------------------------------------------------------------------------
// launchers

namespace ccqk {

static int grid_for(int64_t n, int threads, int cap = 4096) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

hipError_t launch_gemm(hipStream_t s, const GemmArgs& g) {
  dim3 grid((unsigned)((g.M + BM - 1) / BM), (unsigned)((g.N + BN - 1) / BN), 1);
  hipLaunchKernelGGL(gemm_f64_tab, grid, dim3(256), 0, s, g.M, g.N, g.K, g.A, g.rowA, g.kA, g.B,
                     g.kB, g.colB, g.C, g.rowC, g.colC, g.alpha, g.accumulate, g.a_kfast,
                     g.b_kfast);
  return hipGetLastError();
}

hipError_t launch_perm_add(hipStream_t s, int64_t M, int64_t N, const double* A,
                           const int64_t* rowA, const int64_t* colA, double* C,
                           const int64_t* rowC, const int64_t* colC, double alpha,
                           int accumulate) {
  hipLaunchKernelGGL(perm_add, dim3(grid_for(M * N, 256)), dim3(256), 0, s, M, N, A, rowA, colA,
                     C, rowC, colC, alpha, accumulate);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double* x, int64_t n, double v) {
  hipLaunchKernelGGL(fill, dim3(grid_for(n, 256)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}

hipError_t launch_update(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                         const int* perms, const double* signs, const double* fa,
                         const double* fb) {
  hipLaunchKernelGGL(jacobi_update, dim3(grid_for(P.total, 256)), dim3(256), 0, s, P, t, U, perms,
                     signs, fa, fb);
  return hipGetLastError();
}

hipError_t launch_energy_pairs(hipStream_t s, const EnergyParams& P, double* partial, int blocks) {
  hipLaunchKernelGGL(energy_pairs, dim3(blocks), dim3(256), 0, s, P, partial);
  return hipGetLastError();
}

hipError_t launch_dot_strided(hipStream_t s, int64_t rows, int64_t cols, const double* a,
                              int64_t lda, const double* b, double* partial, int blocks) {
  hipLaunchKernelGGL(dot_strided, dim3(blocks), dim3(256), 0, s, rows, cols, a, lda, b, partial);
  return hipGetLastError();
}

hipError_t launch_diis_gram(hipStream_t s, const double* ring, int64_t len, int n, double* partial,
                            int blocks) {
  switch (n) {
#define CCQ_GRAM(NS)                                                                     \
  case NS:                                                                               \
    hipLaunchKernelGGL(diis_gram<NS>, dim3(blocks), dim3(256), 0, s, ring, len, partial); \
    break;
    CCQ_GRAM(1) CCQ_GRAM(2) CCQ_GRAM(3) CCQ_GRAM(4) CCQ_GRAM(5) CCQ_GRAM(6)
    CCQ_GRAM(7) CCQ_GRAM(8) CCQ_GRAM(9) CCQ_GRAM(10) CCQ_GRAM(11) CCQ_GRAM(12)
#undef CCQ_GRAM
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_diff_sq(hipStream_t s, const double* a, const double* b, int64_t len,
                          double* partial, int blocks) {
  hipLaunchKernelGGL(diff_sq, dim3(blocks), dim3(256), 0, s, a, b, len, partial);
  return hipGetLastError();
}

hipError_t launch_lincomb(hipStream_t s, double* out, const double* ring, int64_t len, int n,
                          const Coeffs& c) {
  hipLaunchKernelGGL(lincomb, dim3(grid_for(len, 256)), dim3(256), 0, s, out, ring, len, n, c);
  return hipGetLastError();
}

}  // namespace ccqk
