// CDNA4 (gfx950) kernels of the coupled-cluster amplitude engine.
//
//  * gemm_tab       MFMA tile GEMM (FP64 v_mfma_f64_16x16x4f64, or FP32
//                   v_mfma_f32_16x16x4f32 for the mixed-precision form) whose
//                   operand loaders and epilogue address tensors through
//                   per-index offset tables, so any index permutation /
//                   sub-block view of a tensor is read and written in place
//                   (no separate reorder/sum_stripe passes; cf.
//                   src/cc/reorder_stripe.f90).
//  * gemm_f64_reg   register-direct FP64 MFMA GEMM for skinny products.
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
#include <vector>
#include <cmath>

#include "kernels.h"

#include "gemm_device.h"

namespace ccqk {

// C (+)= alpha * sum_z ws[z] for split-K GEMMs.  P = 2^pshift threads share an
// element (z = p, p + P, ...), combined through LDS in p order: the summation
// order depends only on nsplit, so results are deterministic.
// Elements are addressed as (m = blockIdx.y, n = blockIdx.x * EPB + e), so no
// integer division is needed.
template <typename IO>
__global__ __launch_bounds__(256) void gemm_splitk_reduce(GemmArgs g, int pshift) {
  __shared__ double red[256];
  const int P = 1 << pshift;
  const int64_t total = g.M * g.N;
  const int p = threadIdx.x & (P - 1);
  const int64_t n = (int64_t)blockIdx.x * (256 >> pshift) + (threadIdx.x >> pshift);
  const bool in = n < g.N;
  const int cnt = (g.nsplit - p + P - 1) / P;
  for (int64_t m = blockIdx.y; m < g.M; m += gridDim.y) {
    double s = 0.0;
    if (in) s = sum_strided(g.ws + (int64_t)p * total + m * g.N + n, (int64_t)P * total, cnt);
    if (P > 1) {
      red[threadIdx.x] = s;
      __syncthreads();
      if (p == 0)
        for (int x = 1; x < P; ++x) s += red[threadIdx.x + x];
      __syncthreads();
    }
    if (in && p == 0) store_c<IO>(g, m, n, s);
  }
}

// C(perm) (+)= alpha A through int32 offset tables: element (m, n) reads
// A[rowA[m] + colA[n]] and updates C[rowC[m] + colC[n]].  A block covers
// 256 >> tn_shift rows x (1 << tn_shift) columns per pass.
template <typename IO>
__global__ __launch_bounds__(256) void perm_add(int64_t M, int N, int tn_shift,
                                                const double* __restrict__ A,
                                                const int* __restrict__ rowA,
                                                const int* __restrict__ colA,
                                                double* __restrict__ C,
                                                const int* __restrict__ rowC,
                                                const int* __restrict__ colC, double alpha,
                                                int accumulate) {
  const int tn = 1 << tn_shift;
  const int tm = 256 >> tn_shift;
  const int n0 = threadIdx.x & (tn - 1);
  for (int64_t m = (int64_t)blockIdx.x * tm + (threadIdx.x >> tn_shift); m < M;
       m += (int64_t)gridDim.x * tm) {
    const int64_t ra = rowA[m], rc = rowC[m];
    for (int n = n0; n < N; n += tn) {
      const double v = alpha * ldg<IO>(A, ra + colA[n]);
      put_c<IO>(C, rc + colC[n], v, accumulate);
    }
  }
}

// Y[rowY[m] + colY[n]] = scale * X[rowX[m] + colX[n]] with int64 tables
// (copies a view out of a base integral array that may exceed 2^31
// elements; packs / unpacks the active-space t_vec)
__global__ __launch_bounds__(256) void gather64(int64_t M, int64_t N, const double* __restrict__ X,
                                                const int64_t* __restrict__ rowX,
                                                const int64_t* __restrict__ colX,
                                                double* __restrict__ Y,
                                                const int64_t* __restrict__ rowY,
                                                const int64_t* __restrict__ colY, double scale) {
  for (int64_t m = blockIdx.x; m < M; m += gridDim.x)
    for (int64_t n = threadIdx.x; n < N; n += blockDim.x)
      Y[rowY[m] + colY[n]] = scale * X[rowX[m] + colX[n]];
}

// C (+)= alpha A over identically laid out dense tensors, two doubles per thread
template <typename IO>
__global__ __launch_bounds__(256) void axpy_flat(int64_t n, const double* __restrict__ A,
                                                 double* __restrict__ C, double alpha,
                                                 int accumulate) {
  if constexpr (sizeof(IO) == 4) {
    // four elements per thread, a grid stride apart, loaded before any store
    const float* __restrict__ Af = reinterpret_cast<const float*>(A);
    float* __restrict__ Cf = reinterpret_cast<float*>(C);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += 4 * stride) {
      float a[4], c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool in = b + j * stride < n;
        a[j] = in ? Af[b + j * stride] : 0.0f;
        c[j] = (in && accumulate) ? Cf[b + j * stride] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (b + j * stride < n) Cf[b + j * stride] = (float)((double)c[j] + alpha * (double)a[j]);
    }
    return;
  }
  const int64_t n2 = n / 2;
  const double2* A2 = reinterpret_cast<const double2*>(A);
  double2* C2 = reinterpret_cast<double2*>(C);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = A2[e];
    double2 c;
    if (accumulate) {
      c = C2[e];
      c.x += alpha * a.x;
      c.y += alpha * a.y;
    } else {
      c.x = alpha * a.x;
      c.y = alpha * a.y;
    }
    C2[e] = c;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
    const double v = alpha * A[n - 1];
    if (accumulate) C[n - 1] += v; else C[n - 1] = v;
  }
}

// Several flat C_i (+)= alpha_i A_i in one launch: element e of the
// concatenation belongs to the member whose range holds it.
template <typename IO>
__global__ __launch_bounds__(256) void axpy_group(AddGroup ag) {
  const int64_t total = ag.start[ag.n];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int m = 0;
#pragma unroll
    for (int i = 1; i < kMaxAddGroup; ++i)
      if (i < ag.n && e >= ag.start[i]) m = i;
    const int64_t k = e - ag.start[m];
    put_c<IO>(ag.C[m], k, ag.alpha[m] * ldg<IO>(ag.A[m], k), ag.accumulate[m]);
  }
}

template <typename IO>
__global__ __launch_bounds__(256) void perm_group(PermGroup pg) {
  const int64_t total = pg.start[pg.n];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int m = 0;
#pragma unroll
    for (int i = 1; i < kMaxPermGroup; ++i)
      if (i < pg.n && e >= pg.start[i]) m = i;
    const int64_t k = e - pg.start[m];
    const int64_t r = k / pg.N[m], c = k - r * pg.N[m];
    const double v = pg.alpha[m] * ldg<IO>(pg.A[m], (int64_t)pg.rowA[m][r] + pg.colA[m][c]);
    put_c<IO>(pg.C[m], (int64_t)pg.rowC[m][r] + pg.colC[m][c], v, pg.accumulate[m]);
  }
}

// y = (float)x; two elements per thread (double2 -> float2) when pairs,
// after a head element when head (x + 1 and y + 1 are the aligned starts)
__global__ void cast_f32(float* __restrict__ y, const double* __restrict__ x, int64_t n, int pairs, int head) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (head) {
    if (t0 == 0) y[0] = (float)x[0];
    ++x;
    ++y;
    --n;
  }
  if (pairs) {
    const double2* x2 = reinterpret_cast<const double2*>(x);
    float2* y2 = reinterpret_cast<float2*>(y);
    for (int64_t e = t0; e < n / 2; e += stride) {
      const double2 v = x2[e];
      y2[e] = make_float2((float)v.x, (float)v.y);
    }
    if (t0 == 0 && (n & 1)) y[n - 1] = (float)x[n - 1];
    return;
  }
  for (int64_t e = t0; e < n; e += stride) y[e] = (float)x[e];
}



__global__ void copy_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = x[e];
}

__global__ void fill(double* __restrict__ x, int64_t n, double v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    x[e] = v;
}

// One thread per element of the full block; only canonical tuples (strictly
// increasing inside every antisymmetric group) do work.
template <bool SUMSQ, typename IO, int RANK>
__global__ __launch_bounds__(256) void jacobi_update(UpdateParams P, double* __restrict__ t,
                                                     const double* __restrict__ U,
                                                     const int* __restrict__ perms,
                                                     const double* __restrict__ signs,
                                                     const double* __restrict__ fdiag_a,
                                                     const double* __restrict__ fdiag_b,
                                                     double* __restrict__ sumsq) {
  // the block's rank (P.rank) as a constant: every index loop unrolls
  constexpr int rank = RANK;
  constexpr int half = RANK / 2;
  double sq = 0.0;
  // one thread per canonical tuple (occupied half oc, virtual half vc; the
  // virtual half runs fastest, so the identity image is read and written
  // along its last index)
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P.total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx[8];
    const int64_t oc = e / P.n_vc, vc = e - oc * P.n_vc;
    for (int d = 0; d < half; ++d) {
      idx[d] = P.coc[oc * half + d];
      idx[half + d] = P.cvc[vc * half + d];
    }
    if (P.mask) {
      bool occ_act = false, vir_a = false, vir_b = false;
      for (int d = 0; d < rank; ++d) {
        if (!P.isvir[d]) {
          occ_act |= idx[d] >= P.occ_start;
        } else if (idx[d] < P.vir_end[P.spin[d]]) {
          if (P.spin[d]) vir_b = true; else vir_a = true;
        }
      }
      const bool need = ((P.vneed & 1) && vir_a) || ((P.vneed & 2) && vir_b);
      const bool forbid = ((P.vforbid & 1) && vir_a) || ((P.vforbid & 2) && vir_b);
      if (!occ_act || !need || forbid) continue;
    }
    double R = 0.0;
    for (int g = 0; g < P.nimg; ++g) {
      int64_t off = 0;
      for (int d = 0; d < rank; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      R += signs[g] * ldg<IO>(U, off);
    }
    R /= (double)P.nimg;
    if (SUMSQ) sq += (double)P.nimg * R * R;
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
  if (SUMSQ) {
    __shared__ double red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) sumsq[blockIdx.x] = red[0];
  }
}

// The same step with one wave per canonical tuple, for blocks with many
// images (a T4 block of one spin has 576): lane l gathers and scatters images
// l, l + 64, ...; the image sum is a fixed butterfly over the wave.
template <bool SUMSQ, typename IO, int RANK>
__global__ __launch_bounds__(256) void jacobi_update_wave(UpdateParams P, double* __restrict__ t,
                                                          const double* __restrict__ U,
                                                          const int* __restrict__ perms,
                                                          const double* __restrict__ signs,
                                                          const double* __restrict__ fdiag_a,
                                                          const double* __restrict__ fdiag_b,
                                                          double* __restrict__ sumsq) {
  // the block's rank (P.rank) as a constant: every index loop unrolls
  constexpr int rank = RANK;
  constexpr int half = RANK / 2;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  double sq = 0.0;
  for (int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; e < P.total; e += nwaves) {
    int64_t idx[8];
    const int64_t oc = e / P.n_vc, vc = e - oc * P.n_vc;
    for (int d = 0; d < half; ++d) {
      idx[d] = P.coc[oc * half + d];
      idx[half + d] = P.cvc[vc * half + d];
    }
    if (P.mask) {
      bool occ_act = false, vir_a = false, vir_b = false;
      for (int d = 0; d < rank; ++d) {
        if (!P.isvir[d]) {
          occ_act |= idx[d] >= P.occ_start;
        } else if (idx[d] < P.vir_end[P.spin[d]]) {
          if (P.spin[d]) vir_b = true; else vir_a = true;
        }
      }
      const bool need = ((P.vneed & 1) && vir_a) || ((P.vneed & 2) && vir_b);
      const bool forbid = ((P.vforbid & 1) && vir_a) || ((P.vforbid & 2) && vir_b);
      if (!occ_act || !need || forbid) continue;
    }
    double R = 0.0;
    for (int g = lane; g < P.nimg; g += 64) {
      int64_t off = 0;
      for (int d = 0; d < rank; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      R += signs[g] * ldg<IO>(U, off);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) R += __shfl_xor(R, m);
    R /= (double)P.nimg;
    if (SUMSQ && lane == 0) sq += (double)P.nimg * R * R;
    double D = 0.0;
    for (int d = 0; d < rank; ++d)
      if (P.isvir[d]) D += (P.spin[d] ? fdiag_b : fdiag_a)[P.occ[d] + idx[d]];
    for (int d = 0; d < rank; ++d)
      if (!P.isvir[d]) D -= (P.spin[d] ? fdiag_b : fdiag_a)[idx[d]];
    D += P.shift;
    int64_t off0 = 0;
    for (int d = 0; d < rank; ++d) off0 += idx[d] * P.strides[d];
    // every lane has read t[off0] before any lane stores an image (the store
    // depends on tn)
    const double tn = t[off0] - R / D;
    for (int g = lane; g < P.nimg; g += 64) {
      int64_t off = 0;
      for (int d = 0; d < rank; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      t[off] = signs[g] * tn;
    }
  }
  if (SUMSQ) {
    __shared__ double red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) sumsq[blockIdx.x] = red[0];
  }
}

// out[m] = scale * sum_k X[xr[m] + xk[k]] * Y[yr[m] + yk[k]]
__global__ __launch_bounds__(256) void hadamard_sum(int64_t M, int64_t K, const double* __restrict__ X,
                                                    const int64_t* __restrict__ xr,
                                                    const int64_t* __restrict__ xk,
                                                    const double* __restrict__ Y,
                                                    const int64_t* __restrict__ yr,
                                                    const int64_t* __restrict__ yk,
                                                    double* __restrict__ out, double scale) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < M;
       m += (int64_t)gridDim.x * blockDim.x) {
    const double* x = X + xr[m];
    const double* y = Y + yr[m];
    double acc = 0.0;
    for (int64_t k = 0; k < K; ++k) acc += x[xk[k]] * y[yk[k]];
    out[m] = scale * acc;
  }
}

__global__ __launch_bounds__(256) void diag_copy(const double* __restrict__ src, int64_t n, int64_t ld,
                                                 double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i * (ld + 1)];
}

// H-bar two-body diagonals of crcc23 (src/mm_correct.f90:116-330)
struct Cct3Diag {
  const Cct3Params& P;
  __device__ int64_t no(int s) const { return s ? P.ob : P.oa; }
  __device__ int64_t nv(int s) const { return s ? P.vb : P.va; }
  // h2(a,i,a,i) = -<ia||ai> = <ia||ia>, block ovov (so, sv, so, sv)
  __device__ double vo(int sv, int64_t a, int so, int64_t i) const {
    const int64_t o = no(so), v = nv(sv);
    return P.ovov[so][sv][((i * v + a) * o + i) * v + a];
  }
  // h2(p,q,p,q) of a same-kind pair; mixed spins live in the abab block with
  // the alpha index first
  __device__ double pair(const double* const* blk, int64_t n_a, int64_t n_b, int s1, int64_t p,
                         int s2, int64_t q) const {
    if (s1 == s2) {
      const int64_t n = s1 ? n_b : n_a;
      return blk[s1 ? 2 : 0][((p * n + q) * n + p) * n + q];
    }
    const int64_t x = s1 ? q : p, y = s1 ? p : q;   // x alpha, y beta
    return blk[1][((x * n_b + y) * n_a + x) * n_b + y];
  }
  __device__ double oo(int s1, int64_t j, int s2, int64_t i) const {
    return pair(P.oooo, P.oa, P.ob, s1, j, s2, i);
  }
  __device__ double vv(int s1, int64_t b, int s2, int64_t a) const {
    return pair(P.vvvv, P.va, P.vb, s1, b, s2, a);
  }
};

// d3 arrays are [x][y][z] with the extents of cal_d3 (src/mm_correct.f90:345-481)
__device__ __forceinline__ double d3at(const double* d, int64_t n1, int64_t n2, int64_t x, int64_t y,
                                       int64_t z) {
  return d[(x * n1 + y) * n2 + z];
}

__global__ __launch_bounds__(256) void cct3_correction(Cct3Params P, const double* __restrict__ L3,
                                                       const double* __restrict__ M3,
                                                       const int* __restrict__ perms,
                                                       const double* __restrict__ signs,
                                                       double* __restrict__ partial) {
  __shared__ double red[4][256];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int cs = P.cs;
  // spins of (i, j, k) and of (a, b, c)
  int sp[3];
  for (int x = 0; x < 3; ++x) sp[x] = x >= 3 - cs ? 1 : 0;
  const Cct3Diag H{P};
  const int64_t oa = P.oa, ob = P.ob, va = P.va, vb = P.vb;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < P.total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx[6];
    int64_t r = e;
    for (int d = 5; d >= 0; --d) {
      idx[d] = r % P.dims[d];
      r /= P.dims[d];
    }
    // canonical: strictly increasing inside same-spin runs
    bool canon = true;
    for (int x = 0; x < 2; ++x)
      if (sp[x] == sp[x + 1] && (idx[x] >= idx[x + 1] || idx[3 + x] >= idx[4 + x])) canon = false;
    if (!canon) continue;
    // outside the P space: no active occupied or no active virtual
    bool occ_act = false, vir_act = false;
    for (int x = 0; x < 3; ++x) {
      occ_act |= idx[x] >= P.occ_start;
      vir_act |= idx[3 + x] < P.vir_end[sp[x]];
    }
    if (occ_act && vir_act) continue;
    double l = 0.0, m = 0.0;
    for (int g = 0; g < P.nimg; ++g) {
      int64_t off = 0;
      for (int d = 0; d < 6; ++d) off += idx[d] * P.strides[perms[g * 8 + d]];
      l += signs[g] * L3[off];
      m += signs[g] * M3[off];
    }
    const double inv = 1.0 / (double)P.nimg;
    const double lm = (l * inv) * (m * inv);
    double da = 0.0, db = 0.0;
    for (int x = 0; x < 3; ++x) {
      const int s = sp[x];
      const double* f = P.fd[s];
      const double* h = P.hd[s];
      const int64_t o = s ? ob : oa;
      da += f[idx[x]] - f[o + idx[3 + x]];
      db += h[idx[x]] - h[o + idx[3 + x]];
    }
    double dc = db;
    for (int x = 0; x < 3; ++x)
      for (int y = 0; y < 3; ++y) dc += H.vo(sp[x], idx[3 + x], sp[y], idx[y]);
    for (int x = 0; x < 3; ++x)
      for (int y = x + 1; y < 3; ++y) {
        dc -= H.oo(sp[y], idx[y], sp[x], idx[x]);
        dc -= H.vv(sp[y], idx[3 + y], sp[x], idx[3 + x]);
      }
    const int64_t i = idx[0], j = idx[1], k = idx[2], a = idx[3], b = idx[4], c = idx[5];
    const double* const* D = P.d3;
    double x3 = 0.0;
    if (cs == 0) {
      // d3a1 (va, oa, oa), d3a2 (va, va, oa)
      auto A1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[0], oa, oa, p, q, s); };
      auto A2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[1], va, oa, p, q, s); };
      x3 = A1(a, j, i) + A1(a, k, i) + A1(a, k, j) + A1(b, j, i) + A1(b, k, i) + A1(b, k, j) +
           A1(c, j, i) + A1(c, k, i) + A1(c, k, j) - A2(b, a, i) - A2(b, a, j) - A2(b, a, k) -
           A2(c, a, i) - A2(c, a, j) - A2(c, a, k) - A2(c, b, i) - A2(c, b, j) - A2(c, b, k);
    } else if (cs == 1) {
      auto A1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[0], oa, oa, p, q, s); };
      auto A2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[1], va, oa, p, q, s); };
      auto B1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[2], ob, oa, p, q, s); };
      auto B2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[3], va, oa, p, q, s); };
      auto C1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[4], ob, oa, p, q, s); };
      auto C2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[5], va, ob, p, q, s); };
      x3 = A1(a, j, i) + B1(a, k, i) + B1(a, k, j) + A1(b, j, i) + B1(b, k, i) + B1(b, k, j) +
           C1(c, k, i) + C1(c, k, j) - A2(b, a, i) - A2(b, a, j) - B2(c, a, i) - B2(c, a, j) -
           C2(c, a, k) - B2(c, b, i) - B2(c, b, j) - C2(c, b, k);
    } else if (cs == 2) {
      auto B1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[2], ob, oa, p, q, s); };
      auto B2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[3], va, oa, p, q, s); };
      auto C1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[4], ob, oa, p, q, s); };
      auto C2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[5], va, ob, p, q, s); };
      auto D1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[6], ob, ob, p, q, s); };
      auto D2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[7], vb, ob, p, q, s); };
      x3 = B1(a, j, i) + B1(a, k, i) + C1(b, j, i) + C1(b, k, i) + D1(b, k, j) + C1(c, j, i) +
           C1(c, k, i) + D1(c, k, j) - B2(b, a, i) - C2(b, a, j) - C2(b, a, k) - B2(c, a, i) -
           C2(c, a, j) - C2(c, a, k) - D2(c, b, j) - D2(c, b, k);
    } else {
      auto D1 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[6], ob, ob, p, q, s); };
      auto D2 = [&](int64_t p, int64_t q, int64_t s) { return d3at(D[7], vb, ob, p, q, s); };
      x3 = D1(a, j, i) + D1(a, k, i) + D1(a, k, j) + D1(b, j, i) + D1(b, k, i) + D1(b, k, j) +
           D1(c, j, i) + D1(c, k, i) + D1(c, k, j) - D2(b, a, i) - D2(b, a, j) - D2(b, a, k) -
           D2(c, a, i) - D2(c, a, j) - D2(c, a, k) - D2(c, b, i) - D2(c, b, j) - D2(c, b, k);
    }
    const double dd = dc + x3;
    acc[0] += lm / da;
    acc[1] += lm / db;
    acc[2] += lm / dc;
    acc[3] += lm / dd;
  }
  for (int x = 0; x < 4; ++x) red[x][threadIdx.x] = acc[x];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int x = 0; x < 4; ++x) red[x][threadIdx.x] += red[x][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int x = 0; x < 4; ++x) partial[blockIdx.x * 4 + x] = red[x][0];
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

// diis_record and the next residuum in one pass: dst = src, and the block
// partials of sum (prev - src)^2 (prev == nullptr: sum src^2) with the
// element-to-thread map and reduction of diff_sq, so the partials equal
// diff_sq(prev, dst)'s
__global__ __launch_bounds__(256) void copy_diff_sq(double* __restrict__ dst, const double* __restrict__ src,
                                                    const double* __restrict__ prev, int64_t len,
                                                    double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double v = src[e];
    dst[e] = v;
    const double d = (prev ? prev[e] : 0.0) - v;
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

// Register-direct kernel variants: (RM, RN, WK).  Output tile
// ((4/WK) * 16 * RM) x (16 * RN).
struct RegShape { int rm, rn, wk; };
static const RegShape kRegShapes[] = {{1, 1, 1}, {2, 1, 1}, {4, 1, 1}, {2, 2, 1}, {1, 1, 4},
                                      {2, 1, 4}, {4, 1, 4}, {6, 1, 4}, {2, 2, 4}};

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

static void tile_dims(const GemmArgs& g, int64_t& tbm, int64_t& tbn) {
  if (g.kind == 1) {
    tbm = (4 / g.wk) * 16 * g.rm;
    tbn = 16 * g.rn;
  } else {
    tbm = g.bm;
    tbn = g.bn;
  }
}

int64_t gemm_tiles(const GemmArgs& g) {
  int64_t tbm, tbn;
  tile_dims(g, tbm, tbn);
  const int64_t M = g.trans ? g.N : g.M, N = g.trans ? g.M : g.N;
  return cdiv(M, tbm) * cdiv(N, tbn);
}

int64_t gemm_ws_need(const GemmArgs& g) { return g.nsplit > 1 ? g.nsplit * g.M * g.N : 0; }

// Split-K: the partials make one HBM round trip and the reduction is one more
// launch (~2 us inside a graph).
static double splitk_cycles(int64_t M, int64_t N, int64_t split) {
  if (split <= 1) return 0.0;
  const double hbm = 8.0e12 / 2.4e9;
  return 2.0 * split * M * N * 8.0 / hbm + 8000.0;
}

void gemm_plan(GemmArgs& g, int64_t ws_cap) { gemm_plan_ranked(g, ws_cap, nullptr); }

// The planner's cycle model over every kernel / tile / split-K choice; the
// best one goes into g, and every candidate with its modelled cycles into
// ranked when given (the autotuner measures the best few).
void gemm_plan_ranked(GemmArgs& g, int64_t ws_cap, std::vector<std::pair<double, GemmArgs>>* ranked) {
  // Cost model in cycles (2.4 GHz).  A 16x16x4 FP64 MFMA occupies a SIMD for
  // 64 cycles (78.6 TF/s over 256 CUs x 4 SIMDs).
  const double cus = 256.0;
  // streaming rates measured for the two kernels on skinny products
  // (bytes per cycle): the LDS-tiled loaders read k-contiguous rows with
  // coalesced 256-byte runs (~4 TB/s), the register-direct fragment loads
  // reach ~2.5 TB/s
  const double hbm_tile = 4.0e12 / 2.4e9;
  const double hbm_reg = 2.5e12 / 2.4e9;
  double best = 1e300;
  auto take = [&](double cyc, int kind, int trans, int bm, int bn, int bk, const RegShape* r,
                  int64_t split, int64_t kchunk, int ksp) {
    if (ranked) {
      GemmArgs c = g;
      c.kind = kind;
      c.trans = trans;
      c.bm = bm;
      c.bn = bn;
      c.bk = bk;
      c.rm = r ? r->rm : 1;
      c.rn = r ? r->rn : 1;
      c.wk = r ? r->wk : 1;
      c.nsplit = (int)split;
      c.kchunk = kchunk;
      c.ksplit = ksp;
      ranked->emplace_back(cyc, c);
    }
    if (cyc >= best) return;
    best = cyc;
    g.kind = kind;
    g.trans = trans;
    g.bm = bm;
    g.bn = bn;
    g.bk = bk;
    g.rm = r ? r->rm : 1;
    g.rn = r ? r->rn : 1;
    g.wk = r ? r->wk : 1;
    g.nsplit = (int)split;
    g.kchunk = kchunk;
  };

  // LDS-tiled kernel.  A WG of 4 waves keeps each SIMD busy for
  // (BM/32)(BN/32)*BK/4 MFMAs per K slab; resident WGs per CU (VGPR/LDS):
  // 2 for 128x128, 3 for 128x64 / 64x128, 4 for 64x64, 6 for 64x32 / 32x64,
  // 8 for 32x32.  A CU holding a single WG cannot hide the global-load
  // latency of a slab (~1500 cycles); every slab costs two barriers and the
  // LDS round trip (~300 cycles).  The small tiles pay off for products with
  // a short side of a few MFMA tiles or a short K (more WGs, no split-K).
  struct Shape { int bm, bn, occ; };
  // resident WGs per CU (128x128 uses more than 256 VGPRs: one)
  const Shape shapes[7] = {{128, 128, 1}, {128, 64, 3}, {64, 128, 3}, {64, 64, 4},
                           {64, 32, 6},   {32, 64, 6},  {32, 32, 8}};
  // the epilogue stores 16 consecutive columns per 16 lanes: when C is
  // contiguous along m only, run the product transposed (C^T = B^T A^T)
  const int ttile = (g.c_mfast && !g.c_nfast) ? 1 : 0;
  const int64_t TM_ = ttile ? g.N : g.M, TN_ = ttile ? g.M : g.N;
  for (int bk : {16, 32}) {
    const int64_t ksteps = cdiv(g.K, bk);
    for (const Shape& sh : shapes) {
      // small tiles only for a short side (measured: no gain, or a loss, on
      // square-ish products such as the 900^3 ring)
      if (sh.bm * sh.bn < 64 * 64 && std::min(g.M, g.N) > 128) continue;
      // no tiles for dot-like products with a tiny side and a long K (the
      // register-direct kernel with split-K measured faster)
      if (std::min(g.M, g.N) < 16 && g.K >= 4096) continue;
      const int64_t tiles = cdiv(TM_, sh.bm) * cdiv(TN_, sh.bn);
      const double slab = (sh.bm / 32.0) * (sh.bn / 32.0) * (bk / 4.0) * 64.0;
      const int64_t maxsplit = std::max<int64_t>(1, std::min<int64_t>(ksteps / 2, 512));
      for (int ksp = 1; ksp <= 2; ++ksp)
      for (int64_t split = 1; split <= maxsplit; split *= 2) {
        if (ksp == 2 && sh.bm * sh.bn > 64 * 64) break;
        // measured inside the sweep: the doubled workgroup pays off only for
        // matrix-vector products (a side of 1 or 2); on the 900^3 ring it
        // beats split-K with hot operands (58 vs 61 us) but loses with the
        // sweep's cold ones (94 vs 82 us)
        if (ksp == 2 && std::min(g.M, g.N) > 2) break;
        if (split > 1 && split * g.M * g.N > ws_cap) break;
        // ksp = 2: a doubled workgroup whose two groups take alternate slabs
        const int64_t steps = cdiv(ksteps, split * ksp);
        const double per_cu = std::ceil((double)tiles * split / cus);
        const double resident = std::min<double>(per_cu, std::max(1, sh.occ / ksp)) * ksp;
        double cyc = per_cu * (steps + 2) * (ksp * slab + 300.0);
        if (resident < 2) cyc += per_cu * steps * 1500.0;
        if (std::min(g.M, g.N) <= 128) {
          // skinny: at least one streaming pass over the operands and C
          const double bytes = 8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N);
          // + the epilogue's table-addressed stores (tile elements per thread)
          cyc = std::max(cyc, bytes / hbm_tile + per_cu * (sh.bm * sh.bn / 256.0) * 8.0);
        }
        cyc += splitk_cycles(g.M, g.N, split);
        const double before = best;
        take(cyc, 0, ttile, sh.bm, sh.bn, bk, nullptr, split, steps * ksp * bk, ksp);
        if (best < before) g.ksplit = ksp;
      }
    }
  }

  // Register-direct kernel for products with a skinny output side (N, or M
  // when run transposed).  A wave walks its 16-deep k steps with one step of
  // prefetch, so a step costs a load round trip (~3000 cycles) unless enough
  // resident waves share the SIMD to keep its MFMA pipe busy instead.
  // A is re-read once per tile column and B once per tile row.
  for (int trans = 0; trans < 2; ++trans) {
    const int64_t M = trans ? g.N : g.M, N = trans ? g.M : g.N;
    for (const RegShape& r : kRegShapes) {
      const int64_t tbm = (4 / r.wk) * 16 * r.rm, tbn = 16 * r.rn;
      if (N > 2 * tbn) continue;
      const int64_t tiles = cdiv(M, tbm) * cdiv(N, tbn);
      const int64_t k16 = cdiv(g.K, 16);
      const double mfma = 4.0 * r.rm * r.rn * 64.0;
      const double vgpr = 40.0 + 24.0 * (r.rm + r.rn) + 8.0 * r.rm * r.rn;
      const double occ = std::min(8.0, std::floor(512.0 / vgpr));
      const double bytes = 8.0 * ((double)M * g.K * cdiv(N, tbn) + (double)g.K * N * cdiv(M, tbm));
      for (int64_t split = 1; split <= std::min<int64_t>(k16, 1024); split *= 2) {
        if (split > 1 && split * g.M * g.N > ws_cap) break;
        const int64_t steps = cdiv(cdiv(k16, split), r.wk);
        const double wps = (double)tiles * split * 4.0 / (cus * 4.0);   // waves per SIMD
        const double rounds = std::ceil(wps / occ);
        const double busy = std::min(wps, occ);
        const double lat = rounds * (steps + 1) * std::max(3000.0, mfma * busy);
        // streaming floor; the latency model still orders the shapes when
        // the floor dominates
        double cyc = std::max(lat, bytes / hbm_reg) + 0.05 * lat;
        cyc += splitk_cycles(g.M, g.N, split);
        const double before = best;
        take(cyc, 1, trans, 64, 64, 16, &r, split, cdiv(k16, split) * 16, 1);
        if (best < before) g.ksplit = 1;
      }
    }
  }
  // tile-row group of the XCD-aware order: about square per-XCD blocks
  auto set_group = [](GemmArgs& x) {
    int64_t tbm, tbn;
    tile_dims(x, tbm, tbn);
    const int64_t tm = cdiv(x.trans ? x.N : x.M, tbm);
    const double per_xcd = (double)gemm_tiles(x) * x.nsplit / 8.0;
    x.group = (int)std::max<int64_t>(1, std::min<int64_t>(tm, (int64_t)std::lround(std::sqrt(per_xcd))));
  };
  set_group(g);
  if (ranked) {
    for (auto& c : *ranked) set_group(c.second);
    std::stable_sort(ranked->begin(), ranked->end(),
                     [](const std::pair<double, GemmArgs>& a, const std::pair<double, GemmArgs>& b) {
                       return a.first < b.first;
                     });
  }
}

// C^T = B^T A^T: the same product with the roles of the operands swapped.
static GemmArgs transposed(const GemmArgs& g) {
  GemmArgs t = g;
  t.M = g.N;
  t.N = g.M;
  t.A = g.B;
  t.rowA = g.colB;
  t.kA = g.kB;
  t.kA_aff = g.kB_aff;
  t.B = g.A;
  t.colB = g.rowA;
  t.kB = g.kA;
  t.kB_aff = g.kA_aff;
  t.rowC = g.colC;
  t.colC = g.rowC;
  t.rowD = g.colD;
  t.colD = g.rowD;
  t.rowC2 = g.colC2;
  t.colC2 = g.rowC2;
  t.a_kfast = g.b_kfast;
  t.b_kfast = g.a_kfast;
  t.trans = 0;
  return t;
}

// the kernel of a plan, by GemmArgs::prec (arithmetic / storage): 0 FP64 /
// FP64, 1 FP32 / FP64, 2 FP32 / FP32 (the register-direct kernel computes in
// FP64 on either storage)
static hipError_t launch_kernels(hipStream_t s, const GemmArgs& g) {
  bool ok;
  if (g.kind == 1) {
    const int64_t TBM = (4 / g.wk) * 16 * g.rm, TBN = 16 * g.rn;
    const dim3 grid((unsigned)cdiv(g.M, TBM), (unsigned)cdiv(g.N, TBN), (unsigned)g.nsplit);
    ok = g.prec == 2 ? launch_reg_f32io(s, grid, g) : launch_reg_f64(s, grid, g);
  } else {
    const dim3 grid((unsigned)cdiv(g.M, g.bm), (unsigned)cdiv(g.N, g.bn), (unsigned)g.nsplit);
    ok = g.prec == 2 ? launch_tile_f32io(s, grid, g)
                     : (g.prec == 1 ? launch_tile_f32(s, grid, g) : launch_tile_f64(s, grid, g));
  }
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

static hipError_t launch_reduce(hipStream_t s, const GemmArgs& g) {
  // threads per element: enough to cover nsplit in a few loads each
  int pshift = 0;
  while ((1 << pshift) < g.nsplit && (1 << pshift) < 32 && g.M * g.N * (2 << pshift) <= 262144)
    ++pshift;
  const int64_t per_block = 256 >> pshift;
  const dim3 grid((unsigned)cdiv(g.N, per_block), (unsigned)std::min<int64_t>(g.M, 65535));
  if (g.prec == 2) hipLaunchKernelGGL(gemm_splitk_reduce<float>, grid, dim3(256), 0, s, g, pshift);
  else hipLaunchKernelGGL(gemm_splitk_reduce<double>, grid, dim3(256), 0, s, g, pshift);
  return hipGetLastError();
}

hipError_t launch_gemm(hipStream_t s, const GemmArgs& g0) {
  const GemmArgs g = g0.trans ? transposed(g0) : g0;
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.nsplit > 1 && !g.ws) return hipErrorInvalidValue;
  hipError_t e = launch_kernels(s, g);
  if (e != hipSuccess || g.nsplit == 1) return e;
  return launch_reduce(s, g);
}

hipError_t launch_gemm_chain(hipStream_t s, const GemmArgs* es, int n) {
  if (n < 1 || n > kMaxChain) return hipErrorInvalidValue;
  GemmChain ch{};
  ch.n = n;
  ch.koff[0] = 0;
  for (int i = 0; i < n; ++i) {
    ch.e[i] = es[i].trans ? transposed(es[i]) : es[i];
    ch.koff[i + 1] = ch.koff[i] + ch.e[i].K;
  }
  const GemmArgs& o = ch.e[0];
  if (o.M <= 0 || o.N <= 0) return hipSuccess;
  if (o.kind != 0 || o.bk != 32 || o.ksplit != 1 || o.nsplit < 1 || (o.nsplit > 1 && !o.ws))
    return hipErrorInvalidValue;
  for (int i = 1; i < n; ++i)
    if (ch.e[i].M != o.M || ch.e[i].N != o.N || ch.e[i].prec != o.prec ||
        ch.e[i].a_kfast != o.a_kfast || ch.e[i].b_kfast != o.b_kfast)
      return hipErrorInvalidValue;
  const dim3 grid((unsigned)cdiv(o.M, o.bm), (unsigned)cdiv(o.N, o.bn), (unsigned)o.nsplit);
  const bool ok = o.prec == 2 ? launch_chain_f32io(s, grid, ch)
                              : (o.prec == 1 ? launch_chain_f32(s, grid, ch) : launch_chain_f64(s, grid, ch));
  if (!ok) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || o.nsplit == 1) return e;
  // the reduction stores the summed, already scaled, partial products
  GemmArgs r = o;
  r.alpha = 1.0;
  return launch_reduce(s, r);
}

// Grouped launches run the KS = 1, BK = 32 tile kernel; members share the
// tile shape (after transposition) and the arithmetic.
int64_t gemm_group_key(const GemmArgs& g0) {
  if (g0.kind != 0 || g0.nsplit != 1 || g0.ksplit != 1 || g0.bk != 32) return -1;
  const GemmArgs g = g0.trans ? transposed(g0) : g0;
  return ((int64_t)g.bm * 1000 + g.bn) * 2 + g.prec;
}

hipError_t launch_gemm_group(hipStream_t s, const GemmArgs* gs, int n) {
  if (n < 1 || n > kMaxGroup) return hipErrorInvalidValue;
  GemmGroup gg{};
  const int64_t key = gemm_group_key(gs[0]);
  if (key < 0) return hipErrorInvalidValue;
  int64_t total = 0;
  gg.n = 0;
  for (int i = 0; i < n; ++i) {
    if (gemm_group_key(gs[i]) != key) return hipErrorInvalidValue;
    const GemmArgs g = gs[i].trans ? transposed(gs[i]) : gs[i];
    if (g.M <= 0 || g.N <= 0) continue;
    gg.g[gg.n] = g;
    gg.start[gg.n] = (int)total;
    total += cdiv(g.M, g.bm) * cdiv(g.N, g.bn);
    if (total > 0x7fffffffLL) return hipErrorInvalidValue;
    ++gg.n;
  }
  if (gg.n == 0) return hipSuccess;
  gg.start[gg.n] = (int)total;
  const GemmArgs& g = gg.g[0];
  const dim3 grid((unsigned)gg.start[gg.n]);
  const bool ok = g.prec == 2 ? launch_group_f32io(s, grid, gg)
                              : (g.prec == 1 ? launch_group_f32(s, grid, gg) : launch_group_f64(s, grid, gg));
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_perm_group(hipStream_t s, const PermGroup& pg, int io32) {
  if (pg.n < 1 || pg.n > kMaxPermGroup) return hipErrorInvalidValue;
  const int64_t total = pg.start[pg.n];
  if (total <= 0) return hipSuccess;
  const dim3 grid(grid_for(total, 256, 8192));
  if (io32) hipLaunchKernelGGL(perm_group<float>, grid, dim3(256), 0, s, pg);
  else hipLaunchKernelGGL(perm_group<double>, grid, dim3(256), 0, s, pg);
  return hipGetLastError();
}

hipError_t launch_perm_add(hipStream_t s, int64_t M, int64_t N, const double* A, const int* rowA,
                           const int* colA, double* C, const int* rowC, const int* colC,
                           double alpha, int accumulate, int io32) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (N > 0x7fffffff) return hipErrorInvalidValue;
  int shift = 0;
  while ((1 << shift) < N && shift < 8) ++shift;
  const int64_t rows_per_block = 256 >> shift;
  const dim3 grid(grid_for(M, (int)rows_per_block, 8192));
  if (io32)
    hipLaunchKernelGGL(perm_add<float>, grid, dim3(256), 0, s, M, (int)N, shift, A, rowA, colA, C, rowC,
                       colC, alpha, accumulate);
  else
    hipLaunchKernelGGL(perm_add<double>, grid, dim3(256), 0, s, M, (int)N, shift, A, rowA, colA, C, rowC,
                       colC, alpha, accumulate);
  return hipGetLastError();
}

hipError_t launch_gather64(hipStream_t s, int64_t M, int64_t N, const double* X,
                           const int64_t* rowX, const int64_t* colX, double* Y,
                           const int64_t* rowY, const int64_t* colY, double scale) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather64, dim3((unsigned)std::min<int64_t>(M, 65536)), dim3(256), 0, s, M, N,
                     X, rowX, colX, Y, rowY, colY, scale);
  return hipGetLastError();
}

hipError_t launch_axpy_flat(hipStream_t s, int64_t n, const double* A, double* C, double alpha,
                            int accumulate, int io32) {
  if (n <= 0) return hipSuccess;
  if (io32) {
    hipLaunchKernelGGL(axpy_flat<float>, dim3(grid_for((n + 3) / 4, 256, 8192)), dim3(256), 0, s, n, A, C,
                       alpha, accumulate);
    return hipGetLastError();
  }
  // double2 access needs 16-byte alignment of both arrays
  if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(C)) & 15) return hipErrorInvalidValue;
  hipLaunchKernelGGL(axpy_flat<double>, dim3(grid_for((n + 1) / 2, 256, 8192)), dim3(256), 0, s, n, A, C,
                     alpha, accumulate);
  return hipGetLastError();
}

hipError_t launch_axpy_group(hipStream_t s, const AddGroup& ag, int io32) {
  if (ag.n < 1 || ag.n > kMaxAddGroup || ag.start[0] != 0) return hipErrorInvalidValue;
  const int64_t total = ag.start[ag.n];
  if (total <= 0) return hipSuccess;
  const dim3 grid(grid_for(total, 256, 8192));
  if (io32) hipLaunchKernelGGL(axpy_group<float>, grid, dim3(256), 0, s, ag);
  else hipLaunchKernelGGL(axpy_group<double>, grid, dim3(256), 0, s, ag);
  return hipGetLastError();
}

hipError_t launch_cast_f32(hipStream_t s, float* y, const double* x, int64_t n) {
  if (n <= 0) return hipSuccess;
  auto aligned = [](const double* a, const float* b) {
    return ((reinterpret_cast<uintptr_t>(a) & 15) | (reinterpret_cast<uintptr_t>(b) & 7)) == 0;
  };
  const int head = !aligned(x, y) && n > 1 && aligned(x + 1, y + 1);
  const int pairs = aligned(x + head, y + head);
  hipLaunchKernelGGL(cast_f32, dim3(grid_for(pairs ? (n + 1) / 2 : n, 256, 8192)), dim3(256), 0, s, y, x,
                     n, pairs, head);
  return hipGetLastError();
}


// MFMA issue-rate probe (the attainable matrix-core rate on this part, the
// ceiling the contraction kernels are compared with): every wave runs `iters`
// steps of 8 independent accumulator chains of one MFMA instruction; the sum
// is stored so nothing is optimised away.  KIND 0: v_mfma_f64_16x16x4f64,
// 1: v_mfma_f32_16x16x4f32, 2: v_mfma_f32_16x16x32_bf16.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
template <int KIND>
__global__ __launch_bounds__(256) void mfma_ceiling(int64_t iters, double* __restrict__ out) {
  const float x = 1.0f + 1e-3f * (float)(threadIdx.x & 63);
  double s = 0.0;
  if constexpr (KIND == 0) {
    d4 acc[8] = {};
    const double a = x, b = 2.0 - x;
    for (int64_t i = 0; i < iters; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  } else if constexpr (KIND == 1) {
    f4 acc[8] = {};
    const float a = x, b = 2.0f - x;
    for (int64_t i = 0; i < iters; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  } else {
    f4 acc[8] = {};
    bf8 a, b;
    for (int q = 0; q < 8; ++q) {
      a[q] = (__bf16)x;
      b[q] = (__bf16)(2.0f - x);
    }
    for (int64_t i = 0; i < iters; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t launch_mfma_ceiling(hipStream_t s, int kind, int64_t blocks, int64_t iters, double* out,
                               double* flops) {
  const dim3 grid((unsigned)blocks), blk(256);
  const double waves = (double)blocks * 4.0 * (double)iters * 8.0;
  if (kind == 0) hipLaunchKernelGGL(mfma_ceiling<0>, grid, blk, 0, s, iters, out);
  else if (kind == 1) hipLaunchKernelGGL(mfma_ceiling<1>, grid, blk, 0, s, iters, out);
  else if (kind == 2) hipLaunchKernelGGL(mfma_ceiling<2>, grid, blk, 0, s, iters, out);
  else return hipErrorInvalidValue;
  // flops per wave instruction: 2*16*16*4 (FP64, FP32), 2*16*16*32 (BF16)
  *flops = waves * (kind == 2 ? 16384.0 : 2048.0);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, double* y, const double* x, int64_t n) {
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, y, x, n);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double* x, int64_t n, double v) {
  hipLaunchKernelGGL(fill, dim3(grid_for(n, 256)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}

// blocks with at least kWaveImages images run one wave per canonical tuple
constexpr int kWaveImages = 128;
int update_blocks(const UpdateParams& P) {
  return P.nimg >= kWaveImages ? grid_for(P.total * 64, 256, 16384) : grid_for(P.total, 256);
}

template <typename IO, int R>
static void launch_update_r(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                            const int* perms, const double* signs, const double* fa,
                            const double* fb, double* sumsq) {
  const dim3 grid(update_blocks(P)), blk(256);
  if (P.nimg >= kWaveImages) {
    if (sumsq)
      hipLaunchKernelGGL((jacobi_update_wave<true, IO, R>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
    else
      hipLaunchKernelGGL((jacobi_update_wave<false, IO, R>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
    return;
  }
  if (sumsq)
    hipLaunchKernelGGL((jacobi_update<true, IO, R>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
  else
    hipLaunchKernelGGL((jacobi_update<false, IO, R>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
}

template <typename IO>
static void launch_update_io(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                             const int* perms, const double* signs, const double* fa,
                             const double* fb, double* sumsq) {
  switch (P.rank) {
    case 2: launch_update_r<IO, 2>(s, P, t, U, perms, signs, fa, fb, sumsq); break;
    case 4: launch_update_r<IO, 4>(s, P, t, U, perms, signs, fa, fb, sumsq); break;
    case 6: launch_update_r<IO, 6>(s, P, t, U, perms, signs, fa, fb, sumsq); break;
    default: launch_update_r<IO, 8>(s, P, t, U, perms, signs, fa, fb, sumsq); break;
  }
}

hipError_t launch_update(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                         const int* perms, const double* signs, const double* fa,
                         const double* fb, double* sumsq) {
  if (!P.coc || !P.cvc || P.n_vc <= 0) return hipErrorInvalidValue;
  if (P.rank != 2 && P.rank != 4 && P.rank != 6 && P.rank != 8) return hipErrorInvalidValue;
  if (P.total <= 0) return hipSuccess;
  if (P.u32) launch_update_io<float>(s, P, t, U, perms, signs, fa, fb, sumsq);
  else launch_update_io<double>(s, P, t, U, perms, signs, fa, fb, sumsq);
  return hipGetLastError();
}

hipError_t launch_hadamard_sum(hipStream_t s, int64_t M, int64_t K, const double* X,
                               const int64_t* xr, const int64_t* xk, const double* Y,
                               const int64_t* yr, const int64_t* yk, double* out, double scale) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(hadamard_sum, dim3(grid_for(M, 256)), dim3(256), 0, s, M, K, X, xr, xk, Y, yr,
                     yk, out, scale);
  return hipGetLastError();
}

hipError_t launch_cct3(hipStream_t s, const Cct3Params& P, const double* L3, const double* M3,
                       const int* perms, const double* signs, double* partial, int blocks) {
  hipLaunchKernelGGL(cct3_correction, dim3(blocks), dim3(256), 0, s, P, L3, M3, perms, signs,
                     partial);
  return hipGetLastError();
}

hipError_t launch_diag(hipStream_t s, const double* src, int64_t n, int64_t ld, double* dst) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n, ld, dst);
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

hipError_t launch_copy_diff_sq(hipStream_t s, double* dst, const double* src, const double* prev,
                               int64_t len, double* partial, int blocks) {
  hipLaunchKernelGGL(copy_diff_sq, dim3(blocks), dim3(256), 0, s, dst, src, prev, len, partial);
  return hipGetLastError();
}

hipError_t launch_lincomb(hipStream_t s, double* out, const double* ring, int64_t len, int n,
                          const Coeffs& c) {
  hipLaunchKernelGGL(lincomb, dim3(grid_for(len, 256)), dim3(256), 0, s, out, ring, len, n, c);
  return hipGetLastError();
}

}  // namespace ccqk
