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

#include <algorithm>
#include <cmath>

#include "kernels.h"

typedef double d4 __attribute__((ext_vector_type(4)));

namespace ccqk {

constexpr int BK = 16;

// LDS tile of an operand: R x BK values.  KF (k-fastest in global memory):
// stored row-major [R][BK+1] so a 16-lane group writing consecutive k and
// the MFMA fragment read (16 consecutive rows, one k) are both bank-conflict
// free.  Otherwise stored k-major [BK][R+2].
template <int R, bool KF>
struct LdsTile {
  static constexpr int size = KF ? R * (BK + 1) : BK * (R + 2);
  __device__ static int at(int row, int k) { return KF ? row * (BK + 1) + k : k * (R + 2) + row; }
};

// Per-thread loader of one operand slab (R rows x BK k-values per slab).
// KF: thread owns k = tid % BK and rows tid / BK + j * (256 / BK);
// else: thread owns row tid % R and k = tid / R + j * (256 / R).
template <int R, bool KF, bool ISB>
struct Loader {
  static constexpr int L = R * BK / 256;
  int row_off[L];        // table offset of each owned row (or -1)
  int kk[L];             // owned k within slab
  int rr[L];             // owned row within tile
  __device__ void init(const int* __restrict__ rowtab, int64_t r0, int64_t R_total, int tid) {
#pragma unroll
    for (int j = 0; j < L; ++j) {
      int row, k;
      if (KF) { k = tid % BK; row = tid / BK + j * (256 / BK); }
      else { row = tid % R; k = tid / R + j * (256 / R); }
      kk[j] = k;
      rr[j] = row;
      const int64_t g = r0 + row;
      row_off[j] = g < R_total ? rowtab[g] : -1;
    }
  }
  __device__ void fetch(double (&reg)[L], const double* __restrict__ X, const int* __restrict__ ktab,
                        int64_t k0, int64_t kend) const {
    if (KF) {
      const int64_t gk = k0 + kk[0];
      const bool kin = gk < kend;
      const int ko = kin ? ktab[gk] : 0;
#pragma unroll
      for (int j = 0; j < L; ++j) reg[j] = (kin && row_off[j] >= 0) ? X[(int64_t)row_off[j] + ko] : 0.0;
    } else {
#pragma unroll
      for (int j = 0; j < L; ++j) {
        const int64_t gk = k0 + kk[j];
        reg[j] = (gk < kend && row_off[j] >= 0) ? X[(int64_t)row_off[j] + ktab[gk]] : 0.0;
      }
    }
  }
  __device__ void stash(double* lds, const double (&reg)[L]) const {
#pragma unroll
    for (int j = 0; j < L; ++j) lds[LdsTile<R, KF>::at(rr[j], kk[j])] = reg[j];
  }
};

// Tile GEMM C(M x N) (+)= alpha * A(M x K) B(K x N) on v_mfma_f64_16x16x4f64.
// 256 threads = 4 waves in a 2 x 2 grid; each wave owns a (BM/2) x (BN/2)
// sub-tile of 16x16 MFMA tiles.  Operands are gathered through int32 offset
// tables (rowA/kA, kB/colB); the next K slab is fetched into registers while
// the current one is multiplied out of LDS.
// Split-K: blockIdx.z owns K range [z*kchunk, (z+1)*kchunk); with nsplit > 1
// the tile goes to ws[z][m][n] and gemm_splitk_reduce finishes it.
template <int BM, int BN, bool AKF, bool BKF>
__global__ __launch_bounds__(256, 2) void gemm_f64_tab(GemmArgs g, int64_t kchunk, double* __restrict__ ws) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  using TA = LdsTile<BM, AKF>;
  using TB = LdsTile<BN, BKF>;
  __shared__ double As[TA::size];
  __shared__ double Bs[TB::size];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int64_t n0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = (kbeg + kchunk < g.K) ? kbeg + kchunk : g.K;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;

  Loader<BM, AKF, false> la;
  Loader<BN, BKF, true> lb;
  la.init(g.rowA, m0, g.M, tid);
  lb.init(g.colB, n0, g.N, tid);
  double ra[Loader<BM, AKF, false>::L], rb[Loader<BN, BKF, true>::L];

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4{0, 0, 0, 0};

  const int64_t nsteps = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nsteps > 0) {
    la.fetch(ra, g.A, g.kA, kbeg, kend);
    lb.fetch(rb, g.B, g.kB, kbeg, kend);
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    __syncthreads();
    la.stash(As, ra);
    lb.stash(Bs, rb);
    __syncthreads();
    if (s + 1 < nsteps) {
      la.fetch(ra, g.A, g.kA, kbeg + (s + 1) * BK, kend);
      lb.fetch(rb, g.B, g.kB, kbeg + (s + 1) * BK, kend);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      const int rr = lane & 15;
      double af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[TA::at(wm + 16 * i + rr, kr)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[TB::at(wn + 16 * j + rr, kr)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }

  // D layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l, register v
  // holds row (l/16) + 4*v, column l%16
  const int64_t M = g.M, N = g.N;
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn + 16 * j + (lane & 15);
    if (gn >= N) continue;
    const int64_t cc = ws ? gn : (int64_t)g.colC[gn];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm + 16 * i + (lane >> 4) + 4 * v;
        if (gm >= M) continue;
        if (ws) {
          ws[((int64_t)blockIdx.z * M + gm) * N + cc] = acc[i][j][v];
        } else {
          double* p = g.C + (int64_t)g.rowC[gm] + cc;
          const double val = g.alpha * acc[i][j][v];
          if (g.accumulate) *p += val; else *p = val;
        }
      }
    }
  }
}

// C = beta C + alpha * sum_z ws[z] in fixed z order (deterministic).
__global__ void gemm_splitk_reduce(GemmArgs g, int nsplit, const double* __restrict__ ws) {
  const int64_t total = g.M * g.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += ws[(int64_t)z * total + e];
    const int64_t m = e / g.N, n = e - m * g.N;
    double* p = g.C + (int64_t)g.rowC[m] + g.colC[n];
    const double val = g.alpha * s;
    if (g.accumulate) *p += val; else *p = val;
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

void gemm_plan(GemmArgs& g, int64_t ws_cap) {
  // Cost model in MFMA cycles of the busiest SIMD.  A 16x16x4 FP64 MFMA
  // occupies a SIMD for 64 cycles (78.6 TF/s over 256 CUs x 4 SIMDs at
  // 2.4 GHz); a WG of 4 waves keeps each SIMD busy for (BM/32)(BN/32)*4 MFMAs
  // per 16-deep K slab.  Resident WGs per CU (from VGPR/LDS use): 2 for
  // 128x128, 3 for 128x64 / 64x128, 4 for 64x64.  Split-K adds the
  // workspace round trip and a reduction pass.
  struct Shape { int bm, bn, occ; };
  const Shape shapes[4] = {{128, 128, 2}, {128, 64, 3}, {64, 128, 3}, {64, 64, 4}};
  const double cus = 256.0;
  const int64_t ksteps = (g.K + BK - 1) / BK;
  double best = 1e300;
  for (const Shape& sh : shapes) {
    const int64_t tiles = ((g.M + sh.bm - 1) / sh.bm) * ((g.N + sh.bn - 1) / sh.bn);
    const double slab = (sh.bm / 32.0) * (sh.bn / 32.0) * 4.0 * 64.0;   // cycles per K slab
    const int64_t maxsplit = std::max<int64_t>(1, std::min<int64_t>(ksteps / 2, 4096));
    for (int64_t split = 1; split <= maxsplit; split *= 2) {
      if (split > 1 && split * g.M * g.N > ws_cap) break;
      const int64_t steps = (ksteps + split - 1) / split;
      // MFMA work is balanced over CUs in whole WGs; a CU holding a single WG
      // cannot hide the global-load latency of a slab (~1500 cycles).
      const double per_cu = std::ceil((double)tiles * split / cus);
      const double resident = std::min<double>(per_cu, sh.occ);
      double cyc = per_cu * (steps + 3) * slab;
      if (resident < 2) cyc += per_cu * steps * 1500.0;
      if (split > 1) cyc += (2.0 * split * g.M * g.N * 8.0) / 8.0e12 * 2.4e9 + 12000.0;
      if (cyc < best) {
        best = cyc;
        g.bm = sh.bm;
        g.bn = sh.bn;
        g.nsplit = (int)split;
        g.kchunk = steps * BK;
      }
    }
  }
}

template <int BM, int BN>
static void launch_tile(hipStream_t s, dim3 grid, const GemmArgs& g, double* w) {
  if (g.a_kfast && g.b_kfast)
    hipLaunchKernelGGL((gemm_f64_tab<BM, BN, true, true>), grid, dim3(256), 0, s, g, g.kchunk, w);
  else if (g.a_kfast)
    hipLaunchKernelGGL((gemm_f64_tab<BM, BN, true, false>), grid, dim3(256), 0, s, g, g.kchunk, w);
  else if (g.b_kfast)
    hipLaunchKernelGGL((gemm_f64_tab<BM, BN, false, true>), grid, dim3(256), 0, s, g, g.kchunk, w);
  else
    hipLaunchKernelGGL((gemm_f64_tab<BM, BN, false, false>), grid, dim3(256), 0, s, g, g.kchunk, w);
}

hipError_t launch_gemm(hipStream_t s, const GemmArgs& g, double* ws) {
  dim3 grid((unsigned)((g.M + g.bm - 1) / g.bm), (unsigned)((g.N + g.bn - 1) / g.bn),
            (unsigned)g.nsplit);
  double* w = g.nsplit > 1 ? ws : nullptr;
  if (g.bm == 128 && g.bn == 128) launch_tile<128, 128>(s, grid, g, w);
  else if (g.bm == 128) launch_tile<128, 64>(s, grid, g, w);
  else if (g.bn == 128) launch_tile<64, 128>(s, grid, g, w);
  else launch_tile<64, 64>(s, grid, g, w);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || g.nsplit == 1) return e;
  hipLaunchKernelGGL(gemm_splitk_reduce, dim3(grid_for(g.M * g.N, 256)), dim3(256), 0, s, g,
                     g.nsplit, (const double*)ws);
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
