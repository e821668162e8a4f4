// Device code of the GEMM kernels of the coupled-cluster amplitude engine
// (gfx950): operand / epilogue helpers shared with the other kernels, the
// LDS-tiled MFMA tile kernel (single, chained and grouped launches) and the
// register-direct kernel.  Templates: their instantiations live in the units
// that launch them (gemm_f64.hip, gemm_f32.hip, gemm_f32io.hip: one
// arithmetic / storage pair each, compiled in parallel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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


// Launchers of one arithmetic / storage pair (defined in the gemm_*.hip
// units); false when the plan names a shape that has no instantiation.
// f64: FP64 MFMA on FP64 storage; f32: FP32 MFMA on FP64 storage; f32io:
// FP32 MFMA on FP32 storage.  The register-direct kernel depends on the
// storage only (f64: FP64 storage, f32io: FP32 storage).
bool launch_tile_f64(hipStream_t s, dim3 grid, const GemmArgs& g);
bool launch_tile_f32(hipStream_t s, dim3 grid, const GemmArgs& g);
bool launch_tile_f32io(hipStream_t s, dim3 grid, const GemmArgs& g);
bool launch_chain_f64(hipStream_t s, dim3 grid, const GemmChain& ch);
bool launch_chain_f32(hipStream_t s, dim3 grid, const GemmChain& ch);
bool launch_chain_f32io(hipStream_t s, dim3 grid, const GemmChain& ch);
bool launch_group_f64(hipStream_t s, dim3 grid, const GemmGroup& gg);
bool launch_group_f32(hipStream_t s, dim3 grid, const GemmGroup& gg);
bool launch_group_f32io(hipStream_t s, dim3 grid, const GemmGroup& gg);
bool launch_reg_f64(hipStream_t s, dim3 grid, const GemmArgs& g);
bool launch_reg_f32io(hipStream_t s, dim3 grid, const GemmArgs& g);

}  // namespace ccqk
