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

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace ccqk {

// ---------------------------------------------------------------------------
// GEMM building blocks

// Offset of summed index k in an operand: an int32 table, or k * aff when the
// table is affine (aff >= 0), which saves the dependent table load.
__device__ __forceinline__ int64_t koff(const int* __restrict__ tab, int64_t aff, int64_t k) {
  return aff >= 0 ? k * aff : (int64_t)tab[k];
}

// Storage type IO of the tensors a kernel reads and writes: double, or float
// in the FP32-storage sweeps (the argument pointers then hold float arrays;
// arithmetic on loaded values is FP64 in the epilogues)
template <typename IO>
__device__ __forceinline__ double ldg(const double* p, int64_t i) {
  return (double)reinterpret_cast<const IO*>(p)[i];
}
template <typename IO>
__device__ __forceinline__ void stg(double* p, int64_t i, double v) {
  reinterpret_cast<IO*>(p)[i] = (IO)v;
}

template <typename IO>
__device__ __forceinline__ void put_c(double* base, int64_t off, double v, int accumulate) {
  if (accumulate) stg<IO>(base, off, ldg<IO>(base, off) + v); else stg<IO>(base, off, v);
}

template <typename IO>
__device__ __forceinline__ void store_c(const GemmArgs& g, int64_t m, int64_t n, double s, double alpha) {
  const int64_t r = g.rowC[m], c = g.colC[n];
  double v = alpha * s;
  if (g.D) v += g.dcoef * ldg<IO>(g.D, (int64_t)g.rowD[m] + g.colD[n]);
  put_c<IO>(g.C, r + c, v, g.accumulate);
  if (g.rowC2) {
    const int64_t r2 = g.rowC2[m];
    put_c<IO>(g.C, r2 + c, -v, g.accumulate);
    if (g.colC2) put_c<IO>(g.C, r2 + g.colC2[n], v, g.accumulate);
  }
  if (g.colC2) put_c<IO>(g.C, r + g.colC2[n], -v, g.accumulate);
}

template <typename IO>
__device__ __forceinline__ void store_c(const GemmArgs& g, int64_t m, int64_t n, double s) {
  store_c<IO>(g, m, n, s, g.alpha);
}

// sum_{i < cnt} q[i * stride] with four interleaved partial sums (fixed order)
__device__ __forceinline__ double sum_strided(const double* __restrict__ q, int64_t stride, int cnt) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = 0;
  for (; i + 4 <= cnt; i += 4) {
    a0 += q[0];
    a1 += q[stride];
    a2 += q[2 * stride];
    a3 += q[3 * stride];
    q += 4 * stride;
  }
  for (; i < cnt; ++i) {
    a0 += q[0];
    q += stride;
  }
  return (a0 + a1) + (a2 + a3);
}

// Tile order.  Workgroups are dispatched round-robin over the 8 XCDs of
// MI355X, each with its own L2, in linear id order (x fastest, then y, z).
// The linear id is remapped so that every XCD receives a contiguous range of
// logical tiles, and logical tiles are ordered in groups of `group` tile rows
// (all of one split-K slice first), so the A row panels and B column panels a
// workgroup needs are mostly shared with the other workgroups on its XCD.
__device__ __forceinline__ int xcd_logical(int lin, int total) {
  const int q = total >> 3, r = total & 7;
  const int xcd = lin & 7, slot = lin >> 3;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

__device__ __forceinline__ void tile_coords(int group, int& tx, int& ty, int& tz) {
  const int tm = gridDim.x, tn = gridDim.y;
  const int total = tm * tn * (int)gridDim.z;
  const int lin = blockIdx.x + tm * (blockIdx.y + tn * blockIdx.z);
  const int logical = xcd_logical(lin, total);
  const int per_z = tm * tn;
  tz = logical / per_z;
  const int pid = logical - tz * per_z;
  const int gsize = group * tn;
  const int first = (pid / gsize) * group;
  const int gm = tm - first < group ? tm - first : group;
  const int in = pid - (pid / gsize) * gsize;
  tx = first + in % gm;
  ty = in / gm;
}

// Grouped launch (1-D grid over the tiles of every member): the member owning
// the workgroup's logical tile, and the tile inside that member in the order
// above.  Members occupy consecutive logical ranges, so an XCD's range covers
// the tiles of one or two members.
template <int BM, int BN>
__device__ __forceinline__ int group_tile(const GemmGroup& gg, int& tx, int& ty) {
  const int logical = xcd_logical((int)blockIdx.x, gg.start[gg.n]);
  int mem = 0;
#pragma unroll
  for (int i = 1; i < kMaxGroup; ++i)
    if (i < gg.n && logical >= gg.start[i]) mem = i;
  const GemmArgs& g = gg.g[mem];
  const int tm = (int)((g.M + BM - 1) / BM), tn = (int)((g.N + BN - 1) / BN);
  const int pid = logical - gg.start[mem];
  const int group = g.group;
  const int gsize = group * tn;
  const int first = (pid / gsize) * group;
  const int gm = tm - first < group ? tm - first : group;
  const int in = pid - (pid / gsize) * gsize;
  tx = first + in % gm;
  ty = in / gm;
  return mem;
}

// Split-K: workgroup z of a tile stores its partial product of K range
// [z*kchunk, (z+1)*kchunk) at ws[z][m][n] (M x N, row-major); gemm_splitk_reduce
// sums the nsplit partials of every element in a fixed order and writes C.
__device__ __forceinline__ void store_partial(const GemmArgs& g, int z, int64_t m, int64_t n, double v) {
  g.ws[((int64_t)z * g.M + m) * g.N + n] = v;
}

// LDS tile of an operand: R x BK values of type T.  KF (k-fastest in global
// memory): stored row-major [R][BK+P] so a 16-lane group writing consecutive
// k and the MFMA fragment read (16 consecutive rows, one k) are bank-conflict
// free; otherwise stored k-major [BK][R+Q].  FP64: P = 1, Q = 2.  FP32 (four
// k per fragment read by the four 16-lane groups): P = 4 makes the row pitch
// 4 mod 16 words apart and Q = 16 puts the four k rows 16 banks apart, so the
// 64 lanes of one fragment read hit 64 distinct banks.
template <typename T, int R, bool KF, int BK>
struct LdsTile {
  static constexpr int P = sizeof(T) == 8 ? 1 : 4;
  static constexpr int Q = sizeof(T) == 8 ? 2 : 16;
  static constexpr int size = KF ? R * (BK + P) : BK * (R + Q);
  __device__ static int at(int row, int k) { return KF ? row * (BK + P) + k : k * (R + Q) + row; }
};

// Per-thread loader of one operand slab (R rows x BK k-values per slab).
// KF: thread owns k = tid % BK and rows tid / BK + j * (256 / BK);
// else: thread owns row tid % R and k = tid / R + j * (256 / R).
template <int R, bool KF, int BK>
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
  // K offsets: k * kaff when affine, else the workgroup's copy of the table
  // in LDS (lk[k - kbase]) when use_l, else the table in global memory.
  // The registers hold the stored type IO: a conversion right after the
  // load would wait for it there and expose its latency to the MFMA loop.
  template <typename IO>
  __device__ void fetch(IO (&reg)[L], const double* __restrict__ Xd, const int* __restrict__ ktab,
                        int64_t kaff, const int* lk, bool use_l, int64_t kbase, int64_t k0,
                        int64_t kend) const {
    const IO* __restrict__ X = reinterpret_cast<const IO*>(Xd);
    auto ko = [&](int64_t gk) -> int64_t {
      return kaff >= 0 ? gk * kaff : (use_l ? (int64_t)lk[gk - kbase] : (int64_t)ktab[gk]);
    };
    if (KF) {
      const int64_t gk = k0 + kk[0];
      const bool kin = gk < kend;
      const int64_t o = kin ? ko(gk) : 0;
#pragma unroll
      for (int j = 0; j < L; ++j) reg[j] = (kin && row_off[j] >= 0) ? X[(int64_t)row_off[j] + o] : IO(0);
    } else {
#pragma unroll
      for (int j = 0; j < L; ++j) {
        const int64_t gk = k0 + kk[j];
        reg[j] = (gk < kend && row_off[j] >= 0) ? X[(int64_t)row_off[j] + ko(gk)] : IO(0);
      }
    }
  }
  template <typename T, typename IO>
  __device__ void stash(T* lds, const IO (&reg)[L]) const {
#pragma unroll
    for (int j = 0; j < L; ++j) lds[LdsTile<T, R, KF, BK>::at(rr[j], kk[j])] = (T)reg[j];
  }
};

// MFMA step and D-fragment row of one 16x16 tile, by arithmetic type.
// v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 read the same A / B
// fragments (lane l: row l%16, k = l/16); their D layouts differ (measured on
// gfx950): FP64 lane l, register v holds row (l/16) + 4v, FP32 row 4(l/16) + v;
// column l%16 in both.
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef d4 acc_t;
  __device__ static acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mma<float> {
  typedef f4 acc_t;
  __device__ static acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// LDS-tiled GEMM C(M x N) (+)= alpha * A(M x K) B(K x N), for products with
// both output extents large (MFMA bound).  256 threads = 4 waves in a 2 x 2
// grid; each wave owns a (BM/2) x (BN/2) sub-tile of 16x16 MFMA tiles.
// Operands are gathered in their stored type IO (FP64; FP32 in the
// FP32-storage sweeps) through int32 offset tables (rowA/kA, kB/colB) and
// staged into LDS as T; the next K slab is fetched into
// registers while the current one is multiplied out of LDS.  T = double runs
// on v_mfma_f64_16x16x4f64; T = float rounds the operands to FP32 and
// accumulates in FP32 on v_mfma_f32_16x16x4f32 (the mixed-precision form).
// blockIdx.z owns K range [z*kchunk, (z+1)*kchunk) (split-K, completed by
// gemm_splitk_reduce).  KS = 2 doubles the workgroup: two groups of 4 waves
// take alternate K slabs of the same tile (each with its own LDS slab) and
// combine their accumulators through LDS at the end in a fixed order -- the
// K split of a global split-K without its partial products in HBM and its
// reduction launch.
// Epilogue of one tile: C[rowC[m] + colC[n]] (+)= alpha val (+ dcoef D[..]).
// Each lane loads its row offsets once and, per column, all its old C and
// addend values in one batch before storing: the compiler has to assume a
// store through C may change the offset tables, so an element-by-element
// epilogue serialises a table load, a C load and the store for every
// element.  Outputs with mirrored images (packed pairs) take the
// element-by-element path.
template <typename IO, int TM, int TN, typename MM, typename V>
__device__ __forceinline__ void store_tile(const GemmArgs& g, int64_t m0, int64_t n0, int wm, int wn, int lane,
                                           double alpha, const V (&val)[TM][TN]) {
  if (g.rowC2 || g.colC2) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t gn = n0 + wn + 16 * j + (lane & 15);
      if (gn >= g.N) continue;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t gm = m0 + wm + 16 * i + MM::row(lane, v);
          if (gm < g.M) store_c<IO>(g, gm, gn, (double)val[i][j][v], alpha);
        }
    }
    return;
  }
  int rc[TM][4], rd[TM][4];   // int32 offsets as in the tables; -1: row out of range
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gm = m0 + wm + 16 * i + MM::row(lane, v);
      const bool in = gm < g.M;
      rc[i][v] = in ? g.rowC[gm] : -1;
      rd[i][v] = (in && g.D) ? g.rowD[gm] : 0;
    }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn + 16 * j + (lane & 15);
    if (gn >= g.N) continue;
    const int64_t cc = g.colC[gn];
    double out[TM][4];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) out[i][v] = alpha * (double)val[i][j][v];
    if (g.D) {
      const int64_t cd = g.colD[gn];
      double dv[TM][4];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) dv[i][v] = rc[i][v] >= 0 ? ldg<IO>(g.D, (int64_t)rd[i][v] + cd) : 0.0;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) out[i][v] += g.dcoef * dv[i][v];
    }
    if (g.accumulate) {
      double old[TM][4];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) old[i][v] = rc[i][v] >= 0 ? ldg<IO>(g.C, (int64_t)rc[i][v] + cc) : 0.0;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) out[i][v] = old[i][v] + out[i][v];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (rc[i][v] >= 0) stg<IO>(g.C, (int64_t)rc[i][v] + cc, out[i][v]);
  }
}

// LDS bytes of one tile-kernel variant
template <typename T, int BM, int BN, int BK, bool AKF, bool BKF, int KS>
struct TabLds {
  static constexpr int tile = KS * (LdsTile<T, BM, AKF, BK>::size + LdsTile<T, BN, BKF, BK>::size) * (int)sizeof(T);
  static constexpr int red = (KS - 1) * 256 * (BM / 32) * (BN / 32) * 4 * (int)sizeof(T);
  static constexpr int bytes = tile > red ? tile : red;
};

// K-offset tables of a workgroup's K range are copied into LDS when the range
// has at most kKCache entries (the non-k-contiguous loader reads L offsets
// per slab; from global memory each is a dependent load before the data)
constexpr int kKCache = 2048;

template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF, int KS>
__device__ __forceinline__ void gemm_tab_tile(const GemmArgs& g, int tx, int ty, int tz, unsigned char* smem,
                                              int* kts);

template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF, int KS>
__global__ __launch_bounds__(256 * KS, (KS == 1 && !(sizeof(T) == 8 && BM * BN >= 128 * 128)) ? 2 : 1)
void gemm_tab(GemmArgs g) {
  __shared__ __align__(16) unsigned char smem[TabLds<T, BM, BN, BK, AKF, BKF, KS>::bytes];
  __shared__ int kts[2 * kKCache];
  int tx, ty, tz;
  tile_coords(g.group, tx, ty, tz);
  gemm_tab_tile<T, IO, BM, BN, BK, AKF, BKF, KS>(g, tx, ty, tz, smem, kts);
}

// The K-slab loop of the tile kernel: acc += A(m0.., kbeg..kend) B(kbeg..kend, n0..)
// (the caller zeroes acc; group grp of KS takes slabs grp, grp + KS, ...).
template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF, int KS>
__device__ __forceinline__ void tile_mma(const GemmArgs& g, int64_t m0, int64_t n0, int64_t kbeg, int64_t kend,
                                         unsigned char* smem, int* kts,
                                         typename Mma<T>::acc_t (&acc)[BM / 32][BN / 32]) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  using TA = LdsTile<T, BM, AKF, BK>;
  using TB = LdsTile<T, BN, BKF, BK>;
  using MM = Mma<T>;
  const int grp = threadIdx.x >> 8;
  T* As = reinterpret_cast<T*>(smem) + grp * (TA::size + TB::size);
  T* Bs = As + TA::size;
  const int tid = threadIdx.x & 255;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;

  Loader<BM, AKF, BK> la;
  Loader<BN, BKF, BK> lb;
  la.init(g.rowA, m0, g.M, tid);
  lb.init(g.colB, n0, g.N, tid);
  IO ra[Loader<BM, AKF, BK>::L], rb[Loader<BN, BKF, BK>::L];

  // non-affine K tables of this K range into LDS (uniform condition)
  const int64_t klen = kend > kbeg ? kend - kbeg : 0;
  const bool a_l = g.kA_aff < 0 && klen <= kKCache;
  const bool b_l = g.kB_aff < 0 && klen <= kKCache;
  if (a_l || b_l) {
    for (int i = (int)threadIdx.x; i < (int)klen; i += (int)blockDim.x) {
      if (a_l) kts[i] = g.kA[kbeg + i];
      if (b_l) kts[kKCache + i] = g.kB[kbeg + i];
    }
    __syncthreads();
  }

  // (a slab past the end loads zeros, so both groups run the same barriers)
  const int64_t nslabs = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  const int64_t nsteps = (nslabs + KS - 1) / KS;
  if (nsteps > 0) {
    la.template fetch<IO>(ra, g.A, g.kA, g.kA_aff, kts, a_l, kbeg, kbeg + grp * BK, kend);
    lb.template fetch<IO>(rb, g.B, g.kB, g.kB_aff, kts + kKCache, b_l, kbeg, kbeg + grp * BK, kend);
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    __syncthreads();
    la.stash(As, ra);
    lb.stash(Bs, rb);
    __syncthreads();
    if (s + 1 < nsteps) {
      const int64_t k0 = kbeg + ((s + 1) * KS + grp) * BK;
      la.template fetch<IO>(ra, g.A, g.kA, g.kA_aff, kts, a_l, kbeg, k0, kend);
      lb.template fetch<IO>(rb, g.B, g.kB, g.kB_aff, kts + kKCache, b_l, kbeg, k0, kend);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      const int rr = lane & 15;
      T af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[TA::at(wm + 16 * i + rr, kr)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[TB::at(wn + 16 * j + rr, kr)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = MM::step(af[i], bf[j], acc[i][j]);
    }
  }
}

template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF, int KS>
__device__ __forceinline__ void gemm_tab_tile(const GemmArgs& g, int tx, int ty, int tz, unsigned char* smem,
                                              int* kts) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  using MM = Mma<T>;
  const int grp = threadIdx.x >> 8;
  const int tid = threadIdx.x & 255;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t m0 = (int64_t)tx * BM;
  const int64_t n0 = (int64_t)ty * BN;
  const int64_t kbeg = (int64_t)tz * g.kchunk;
  const int64_t kend = (kbeg + g.kchunk < g.K) ? kbeg + g.kchunk : g.K;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;

  typename MM::acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = typename MM::acc_t{0, 0, 0, 0};
  tile_mma<T, IO, BM, BN, BK, AKF, BKF, KS>(g, m0, n0, kbeg, kend, smem, kts, acc);

  if (KS > 1) {
    // group 1 hands its accumulators to group 0 (slab order: group 0's sum
    // plus group 1's sum, the same for every launch)
    T* red = reinterpret_cast<T*>(smem);
    __syncthreads();
    if (grp == 1) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) red[(((i * TN + j) * 4 + v) << 8) + tid] = acc[i][j][v];
    }
    __syncthreads();
    if (grp == 1) return;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[i][j][v] += red[(((i * TN + j) * 4 + v) << 8) + tid];
  }

  if (g.nsplit == 1) {
    store_tile<IO, TM, TN, MM>(g, m0, n0, wm, wn, lane, g.alpha, acc);
    return;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn + 16 * j + (lane & 15);
    if (gn >= g.N) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm + 16 * i + MM::row(lane, v);
        if (gm >= g.M) continue;
        store_partial(g, tz, gm, gn, (double)acc[i][j][v]);
      }
  }
}

// Entries e0 <= e < e1 of a chain accumulating into one output tile (tx, ty):
// the workgroup walks the concatenated K range [kbeg, kend) (kend < 0: all of
// it) entry by entry, each with its own operands and layouts, adds alpha_e
// times the entry's partial product to an FP64 sum, and stores once with the
// first entry's output arguments (split z of a split-K launch: into the
// workspace).
template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF>
__device__ __forceinline__ void tile_entries(const GemmArgs* es, int e0, int e1, int tx, int ty, int tz,
                                             int64_t kbeg, int64_t kend, unsigned char* smem, int* kts) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  using MM = Mma<T>;
  const GemmArgs& o = es[e0];
  const int lane = threadIdx.x & 63;
  const int wave = (threadIdx.x & 255) >> 6;
  const int64_t m0 = (int64_t)tx * BM;
  const int64_t n0 = (int64_t)ty * BN;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;

  d4 sum[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) sum[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  int64_t off = 0;
  for (int e = e0; e < e1; ++e) {
    const GemmArgs& g = es[e];
    const int64_t lo = kend < 0 ? off : (kbeg > off ? kbeg : off);
    const int64_t hi = kend < 0 ? off + g.K : (kend < off + g.K ? kend : off + g.K);
    const int64_t a = lo - off, b = hi - off;
    off += g.K;
    if (a >= b) continue;
    typename MM::acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = typename MM::acc_t{0, 0, 0, 0};
    // the previous entry's waves are past their last LDS reads
    if (e > e0) __syncthreads();
    tile_mma<T, IO, BM, BN, BK, AKF, BKF, 1>(g, m0, n0, a, b, smem, kts, acc);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) sum[i][j][v] += g.alpha * (double)acc[i][j][v];
  }

  if (kend < 0 || o.nsplit == 1) {
    store_tile<IO, TM, TN, MM>(o, m0, n0, wm, wn, lane, 1.0, sum);
    return;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int64_t gn = n0 + wn + 16 * j + (lane & 15);
    if (gn >= o.N) continue;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = m0 + wm + 16 * i + MM::row(lane, v);
        if (gm >= o.M) continue;
        store_partial(o, tz, gm, gn, sum[i][j][v]);
      }
  }
}

// LDS of the kernels that switch between operand layouts per entry
template <typename T, int BM, int BN, int BK>
struct AnyLds {
  static constexpr int b11 = TabLds<T, BM, BN, BK, true, true, 1>::bytes;
  static constexpr int b10 = TabLds<T, BM, BN, BK, true, false, 1>::bytes;
  static constexpr int b01 = TabLds<T, BM, BN, BK, false, true, 1>::bytes;
  static constexpr int b00 = TabLds<T, BM, BN, BK, false, false, 1>::bytes;
  static constexpr int m1 = b11 > b10 ? b11 : b10, m2 = b01 > b00 ? b01 : b00;
  static constexpr int bytes = m1 > m2 ? m1 : m2;
};

// A chain of GEMMs accumulating into one output with the same tile mapping
// and operand layouts (C (+)= sum_e alpha_e A_e B_e (+ dcoef D)), as one
// product over the concatenated K ranges, split along the concatenated K
// (split z of tile (tx, ty) covers [z kchunk, (z + 1) kchunk)).
template <typename T, typename IO, int BM, int BN, int BK, bool AKF, bool BKF>
__global__ __launch_bounds__(256, !(sizeof(T) == 8 && BM * BN >= 128 * 128) ? 2 : 1)
void gemm_tab_chain(GemmChain ch) {
  __shared__ __align__(16) unsigned char smem[TabLds<T, BM, BN, BK, AKF, BKF, 1>::bytes];
  __shared__ int kts[2 * kKCache];
  int tx, ty, tz;
  tile_coords(ch.e[0].group, tx, ty, tz);
  const int64_t ktot = ch.koff[ch.n];
  const int64_t kbeg = (int64_t)tz * ch.e[0].kchunk;
  const int64_t kend = (kbeg + ch.e[0].kchunk < ktot) ? kbeg + ch.e[0].kchunk : ktot;
  tile_entries<T, IO, BM, BN, BK, AKF, BKF>(ch.e, 0, ch.n, tx, ty, tz, kbeg, kend, smem, kts);
}

// The same tile kernel over the tiles of a group of independent GEMMs (full K
// each).  The operand layouts (k-contiguous or not) may differ between
// members: the workgroup branches once, uniformly, to the body of its
// member's layouts.
template <typename T, typename IO, int BM, int BN, int BK>
__global__ __launch_bounds__(256, !(sizeof(T) == 8 && BM * BN >= 128 * 128) ? 2 : 1)
void gemm_tab_group(GemmGroup gg) {
  __shared__ __align__(16) unsigned char smem[AnyLds<T, BM, BN, BK>::bytes];
  __shared__ int kts[2 * kKCache];
  int tx, ty;
  const int mem = group_tile<BM, BN>(gg, tx, ty);
  const GemmArgs& g = gg.g[mem];
  if (g.a_kfast) {
    if (g.b_kfast) gemm_tab_tile<T, IO, BM, BN, BK, true, true, 1>(g, tx, ty, 0, smem, kts);
    else gemm_tab_tile<T, IO, BM, BN, BK, true, false, 1>(g, tx, ty, 0, smem, kts);
  } else {
    if (g.b_kfast) gemm_tab_tile<T, IO, BM, BN, BK, false, true, 1>(g, tx, ty, 0, smem, kts);
    else gemm_tab_tile<T, IO, BM, BN, BK, false, false, 1>(g, tx, ty, 0, smem, kts);
  }
}

// Register-direct GEMM for skinny and dot-shaped products (an output extent
// of at most a few MFMA tiles; HBM / latency bound).  No LDS staging: in a
// 16-deep k step, lane l loads row l%16 of each 16-row block at
// k = kb + 4*(l/16) + s, s = 0..3, and MFMA s contracts exactly those k
// (the A/B fragment layout of v_mfma_f64_16x16x4f64 with the 4 MFMAs of the
// step taking k in the order kb + 4q + s, q = 0..3).  The next step's values
// are loaded while the current step is multiplied.
// Workgroup = 4 waves: WK along K (same output tile, interleaved k steps,
// combined through LDS in wave order), 4/WK along M.  Wave tile (16 RM) x (16 RN).
// The step holds the stored type IO (converted at the MFMA, so a load in
// flight is not waited for before the current step's MFMAs).
template <typename IO, int RM, int RN, int WK>
struct RegStep {
  IO a[RM][4];
  IO b[RN][4];
};

template <typename IO, int RM, int RN, int WK>
__device__ __forceinline__ void reg_load(RegStep<IO, RM, RN, WK>& st, const GemmArgs& g, const int (&ra)[RM],
                                         const int (&cb)[RN], int q, int64_t kb, int64_t kend) {
  const IO* __restrict__ A = reinterpret_cast<const IO*>(g.A);
  const IO* __restrict__ B = reinterpret_cast<const IO*>(g.B);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int64_t gk = kb + 4 * q + s;
    const bool in = gk < kend;
    const int64_t oa = in ? koff(g.kA, g.kA_aff, gk) : 0;
    const int64_t ob = in ? koff(g.kB, g.kB_aff, gk) : 0;
#pragma unroll
    for (int i = 0; i < RM; ++i) st.a[i][s] = (in && ra[i] >= 0) ? A[(int64_t)ra[i] + oa] : IO(0);
#pragma unroll
    for (int j = 0; j < RN; ++j) st.b[j][s] = (in && cb[j] >= 0) ? B[ob + cb[j]] : IO(0);
  }
}

template <typename IO, int RM, int RN, int WK>
__device__ __forceinline__ void reg_mma(d4 (&acc)[RM][RN], const RegStep<IO, RM, RN, WK>& st) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
      for (int j = 0; j < RN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)st.a[i][s], (double)st.b[j][s], acc[i][j], 0, 0, 0);
}

template <typename IO, int RM, int RN, int WK>
__global__ __launch_bounds__(256) void gemm_f64_reg(GemmArgs g) {
  constexpr int WMW = 4 / WK;
  constexpr int TBM = WMW * 16 * RM, TBN = 16 * RN;
  constexpr int WACC = 256 * RM * RN;   // doubles of one wave's accumulators
  constexpr int SMEM = WK > 1 ? (WK - 1) * WMW * WACC : 1;
  __shared__ double red[SMEM];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wk = wave % WK, wmi = wave / WK;
  const int r = lane & 15, q = lane >> 4;
  int tx, ty, tz;
  tile_coords(g.group, tx, ty, tz);
  const int64_t m0 = (int64_t)tx * TBM, n0 = (int64_t)ty * TBN;
  const int64_t wm0 = m0 + wmi * 16 * RM;
  const int64_t kbeg = (int64_t)tz * g.kchunk;
  const int64_t kend = (kbeg + g.kchunk < g.K) ? kbeg + g.kchunk : g.K;
  int ra[RM], cb[RN];
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    const int64_t m = wm0 + 16 * i + r;
    ra[i] = m < g.M ? g.rowA[m] : -1;
  }
#pragma unroll
  for (int j = 0; j < RN; ++j) {
    const int64_t n = n0 + 16 * j + r;
    cb[j] = n < g.N ? g.colB[n] : -1;
  }
  d4 acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) acc[i][j] = d4{0, 0, 0, 0};

  const int64_t nsteps = kend > kbeg ? (kend - kbeg + 15) / 16 : 0;
  RegStep<IO, RM, RN, WK> s0, s1;
  int64_t st = wk;
  if (st < nsteps) reg_load<IO>(s0, g, ra, cb, q, kbeg + 16 * st, kend);
  while (st < nsteps) {
    const int64_t n1 = st + WK;
    if (n1 < nsteps) reg_load<IO>(s1, g, ra, cb, q, kbeg + 16 * n1, kend);
    reg_mma(acc, s0);
    if (n1 >= nsteps) break;
    const int64_t n2 = n1 + WK;
    if (n2 < nsteps) reg_load<IO>(s0, g, ra, cb, q, kbeg + 16 * n2, kend);
    reg_mma(acc, s1);
    st = n2;
  }

  if (WK > 1) {
    // waves wk > 0 hand their accumulators to wave wk == 0 of the same M slot
    if (wk > 0) {
      double* dst = red + ((wk - 1) * WMW + wmi) * WACC;
#pragma unroll
      for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) dst[((i * RN + j) * 4 + v) * 64 + lane] = acc[i][j][v];
    }
    __syncthreads();
    if (wk == 0) {
      for (int w = 1; w < WK; ++w) {
        const double* src = red + ((w - 1) * WMW + wmi) * WACC;
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
          for (int j = 0; j < RN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] += src[((i * RN + j) * 4 + v) * 64 + lane];
      }
    }
  }

  if (wk != 0) return;
  // lane l, register v of 16x16 block (i, j) holds row 16i + l/16 + 4v and
  // column 16j + l%16: the FP64 MFMA output layout of the tile epilogue
  if (g.nsplit == 1) {
    store_tile<IO, RM, RN, Mma<double>>(g, wm0, n0, 0, 0, lane, g.alpha, acc);
    return;
  }
#pragma unroll
  for (int j = 0; j < RN; ++j) {
    const int64_t gn = n0 + 16 * j + r;
    if (gn >= g.N) continue;
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gm = wm0 + 16 * i + q + 4 * v;
        if (gm >= g.M) continue;
        store_partial(g, tz, gm, gn, acc[i][j][v]);
      }
  }
}

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
template <bool SUMSQ, typename IO>
__global__ __launch_bounds__(256) void jacobi_update(UpdateParams P, double* __restrict__ t,
                                                     const double* __restrict__ U,
                                                     const int* __restrict__ perms,
                                                     const double* __restrict__ signs,
                                                     const double* __restrict__ fdiag_a,
                                                     const double* __restrict__ fdiag_b,
                                                     double* __restrict__ sumsq) {
  const int rank = P.rank;
  const int half = rank / 2;
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
template <bool SUMSQ, typename IO>
__global__ __launch_bounds__(256) void jacobi_update_wave(UpdateParams P, double* __restrict__ t,
                                                          const double* __restrict__ U,
                                                          const int* __restrict__ perms,
                                                          const double* __restrict__ signs,
                                                          const double* __restrict__ fdiag_a,
                                                          const double* __restrict__ fdiag_b,
                                                          double* __restrict__ sumsq) {
  const int rank = P.rank;
  const int half = rank / 2;
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

template <typename T, typename IO, int BM, int BN, int BK, int KS>
static void launch_tile_k(hipStream_t s, dim3 grid, const GemmArgs& g) {
  const dim3 blk(256 * KS);
  if (g.a_kfast && g.b_kfast)
    hipLaunchKernelGGL((gemm_tab<T, IO, BM, BN, BK, true, true, KS>), grid, blk, 0, s, g);
  else if (g.a_kfast)
    hipLaunchKernelGGL((gemm_tab<T, IO, BM, BN, BK, true, false, KS>), grid, blk, 0, s, g);
  else if (g.b_kfast)
    hipLaunchKernelGGL((gemm_tab<T, IO, BM, BN, BK, false, true, KS>), grid, blk, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_tab<T, IO, BM, BN, BK, false, false, KS>), grid, blk, 0, s, g);
}

// the doubled workgroup exists for the tiles up to 64x64
template <typename T, typename IO, int BM, int BN, int BK>
static void launch_tile_t(hipStream_t s, dim3 grid, const GemmArgs& g) {
  if constexpr (BM * BN <= 64 * 64) {
    if (g.ksplit == 2) {
      launch_tile_k<T, IO, BM, BN, BK, 2>(s, grid, g);
      return;
    }
  }
  launch_tile_k<T, IO, BM, BN, BK, 1>(s, grid, g);
}

// arithmetic / storage by GemmArgs::prec: 0 FP64 / FP64, 1 FP32 / FP64, 2 FP32 / FP32
template <int BM, int BN, int BK>
static void launch_tile(hipStream_t s, dim3 grid, const GemmArgs& g) {
  if (g.prec == 2) launch_tile_t<float, float, BM, BN, BK>(s, grid, g);
  else if (g.prec == 1) launch_tile_t<float, double, BM, BN, BK>(s, grid, g);
  else launch_tile_t<double, double, BM, BN, BK>(s, grid, g);
}

template <int RM, int RN, int WK>
static void launch_reg(hipStream_t s, const GemmArgs& g) {
  constexpr int64_t TBM = (4 / WK) * 16 * RM, TBN = 16 * RN;
  dim3 grid((unsigned)cdiv(g.M, TBM), (unsigned)cdiv(g.N, TBN), (unsigned)g.nsplit);
  if (g.prec == 2) hipLaunchKernelGGL((gemm_f64_reg<float, RM, RN, WK>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_f64_reg<double, RM, RN, WK>), grid, dim3(256), 0, s, g);
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

static hipError_t launch_kernels(hipStream_t s, const GemmArgs& g) {
  if (g.kind == 1) {
    const int key = g.rm * 100 + g.rn * 10 + g.wk;
    switch (key) {
      case 111: launch_reg<1, 1, 1>(s, g); break;
      case 211: launch_reg<2, 1, 1>(s, g); break;
      case 411: launch_reg<4, 1, 1>(s, g); break;
      case 221: launch_reg<2, 2, 1>(s, g); break;
      case 114: launch_reg<1, 1, 4>(s, g); break;
      case 214: launch_reg<2, 1, 4>(s, g); break;
      case 414: launch_reg<4, 1, 4>(s, g); break;
      case 614: launch_reg<6, 1, 4>(s, g); break;
      case 224: launch_reg<2, 2, 4>(s, g); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  dim3 grid((unsigned)cdiv(g.M, g.bm), (unsigned)cdiv(g.N, g.bn), (unsigned)g.nsplit);
  const int key = g.bm * 1000 + g.bn;
  if (g.bk == 32) {
    switch (key) {
      case 128128: launch_tile<128, 128, 32>(s, grid, g); break;
      case 128064: launch_tile<128, 64, 32>(s, grid, g); break;
      case 64128: launch_tile<64, 128, 32>(s, grid, g); break;
      case 64064: launch_tile<64, 64, 32>(s, grid, g); break;
      case 64032: launch_tile<64, 32, 32>(s, grid, g); break;
      case 32064: launch_tile<32, 64, 32>(s, grid, g); break;
      case 32032: launch_tile<32, 32, 32>(s, grid, g); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (key) {
      case 128128: launch_tile<128, 128, 16>(s, grid, g); break;
      case 128064: launch_tile<128, 64, 16>(s, grid, g); break;
      case 64128: launch_tile<64, 128, 16>(s, grid, g); break;
      case 64064: launch_tile<64, 64, 16>(s, grid, g); break;
      case 64032: launch_tile<64, 32, 16>(s, grid, g); break;
      case 32064: launch_tile<32, 64, 16>(s, grid, g); break;
      case 32032: launch_tile<32, 32, 16>(s, grid, g); break;
      default: return hipErrorInvalidValue;
    }
  }
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

template <typename T, typename IO, int BM, int BN>
static void launch_chain_t(hipStream_t s, dim3 grid, const GemmChain& ch) {
  const GemmArgs& o = ch.e[0];
  if (o.a_kfast && o.b_kfast) hipLaunchKernelGGL((gemm_tab_chain<T, IO, BM, BN, 32, true, true>), grid, dim3(256), 0, s, ch);
  else if (o.a_kfast) hipLaunchKernelGGL((gemm_tab_chain<T, IO, BM, BN, 32, true, false>), grid, dim3(256), 0, s, ch);
  else if (o.b_kfast) hipLaunchKernelGGL((gemm_tab_chain<T, IO, BM, BN, 32, false, true>), grid, dim3(256), 0, s, ch);
  else hipLaunchKernelGGL((gemm_tab_chain<T, IO, BM, BN, 32, false, false>), grid, dim3(256), 0, s, ch);
}

template <int BM, int BN>
static void launch_chain_k(hipStream_t s, const GemmChain& ch) {
  const GemmArgs& o = ch.e[0];
  const dim3 grid((unsigned)cdiv(o.M, BM), (unsigned)cdiv(o.N, BN), (unsigned)o.nsplit);
  if (o.prec == 2) launch_chain_t<float, float, BM, BN>(s, grid, ch);
  else if (o.prec == 1) launch_chain_t<float, double, BM, BN>(s, grid, ch);
  else launch_chain_t<double, double, BM, BN>(s, grid, ch);
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
  switch (o.bm * 1000 + o.bn) {
    case 128128: launch_chain_k<128, 128>(s, ch); break;
    case 128064: launch_chain_k<128, 64>(s, ch); break;
    case 64128: launch_chain_k<64, 128>(s, ch); break;
    case 64064: launch_chain_k<64, 64>(s, ch); break;
    case 64032: launch_chain_k<64, 32>(s, ch); break;
    case 32064: launch_chain_k<32, 64>(s, ch); break;
    case 32032: launch_chain_k<32, 32>(s, ch); break;
    default: return hipErrorInvalidValue;
  }
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

template <int BM, int BN>
static void launch_group(hipStream_t s, const GemmGroup& gg) {
  const dim3 grid((unsigned)gg.start[gg.n]), blk(256);
  if (gg.g[0].prec == 2) hipLaunchKernelGGL((gemm_tab_group<float, float, BM, BN, 32>), grid, blk, 0, s, gg);
  else if (gg.g[0].prec == 1) hipLaunchKernelGGL((gemm_tab_group<float, double, BM, BN, 32>), grid, blk, 0, s, gg);
  else hipLaunchKernelGGL((gemm_tab_group<double, double, BM, BN, 32>), grid, blk, 0, s, gg);
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
  switch (g.bm * 1000 + g.bn) {
    case 128128: launch_group<128, 128>(s, gg); break;
    case 128064: launch_group<128, 64>(s, gg); break;
    case 64128: launch_group<64, 128>(s, gg); break;
    case 64064: launch_group<64, 64>(s, gg); break;
    case 64032: launch_group<64, 32>(s, gg); break;
    case 32064: launch_group<32, 64>(s, gg); break;
    case 32032: launch_group<32, 32>(s, gg); break;
    default: return hipErrorInvalidValue;
  }
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

template <typename IO>
static void launch_update_io(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                             const int* perms, const double* signs, const double* fa,
                             const double* fb, double* sumsq) {
  const dim3 grid(update_blocks(P)), blk(256);
  if (P.nimg >= kWaveImages) {
    if (sumsq)
      hipLaunchKernelGGL((jacobi_update_wave<true, IO>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
    else
      hipLaunchKernelGGL((jacobi_update_wave<false, IO>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
    return;
  }
  if (sumsq)
    hipLaunchKernelGGL((jacobi_update<true, IO>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
  else
    hipLaunchKernelGGL((jacobi_update<false, IO>), grid, blk, 0, s, P, t, U, perms, signs, fa, fb, sumsq);
}

hipError_t launch_update(hipStream_t s, const UpdateParams& P, double* t, const double* U,
                         const int* perms, const double* signs, const double* fa,
                         const double* fb, double* sumsq) {
  if (!P.coc || !P.cvc || P.n_vc <= 0) return hipErrorInvalidValue;
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

hipError_t launch_lincomb(hipStream_t s, double* out, const double* ring, int64_t len, int n,
                          const Coeffs& c) {
  hipLaunchKernelGGL(lincomb, dim3(grid_for(len, 256)), dim3(256), 0, s, out, ring, len, n, c);
  return hipGetLastError();
}

}  // namespace ccqk
