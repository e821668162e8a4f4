// Learned starting guess for T2: a per-amplitude MLP that maps the MP2
// amplitude's local features to the converged amplitude, trained on the
// engine's own converged amplitudes with the kernels of this file.
//
// Model.  For a T2 tuple (i, j, a, b) of a spin block, m = V / D with
// V = <ij||ab> (or <ij|ab> for the alpha-beta block) and
// D = f_ii + f_jj - f_aa - f_bb.  Features, all even under i<->j and a<->b:
//   x = ( |m| / sm,  f_ii + f_jj,  f_aa + f_bb,  V_ijij,  V_abab )
// (sm = rms of m over the block).  The network predicts r and the guess is
// t = m (1 + r), antisymmetric wherever m is and zero where m is.
// Network: x -> (x - mu) / sd -> 64 tanh -> 64 tanh -> 1, weights and layer
// inputs rounded to BF16 (round to nearest even), products accumulated in
// FP32.  Loss: mean over samples of ((m / sm) r - y)^2 with y = (t_conv - m) / sm.
//
// Training (ccq_hip_mlp_train) runs on the GPU: per step a batch of samples
// drawn by a counter-based hash, one thread per sample for forward and
// backward (activations kept in a scratch buffer), one thread per parameter
// summing the batch in a fixed order (deterministic), then an Adam step.
// ccq has no counterpart (its guess is T = 0, src/calc_driver.f90:210-211).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ccq_hip.h"
#include "kernels.h"

namespace ccqk {

constexpr int F = kGuessFeatures;
constexpr int H = kGuessHidden;
// parameter offsets
constexpr int P_MU = 0, P_SD = F, P_W1 = 2 * F, P_B1 = P_W1 + F * H, P_W2 = P_B1 + H,
              P_B2 = P_W2 + H * H, P_W3 = P_B2 + H, P_B3 = P_W3 + H;
static_assert(P_B3 + 1 == kGuessParams, "parameter layout");

// round to the nearest BF16 (ties to even), returned as float
__host__ __device__ inline float bf16r(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float y;
  memcpy(&y, &u, 4);
  return y;
}

struct Tuple {
  double m;
  float x[F];
};

__device__ inline Tuple tuple_at(const GuessBlock& B, int64_t i, int64_t j, int64_t a, int64_t b) {
  const int64_t nc = B.nc, nc2 = nc * nc, nc3 = nc2 * nc;
  const int64_t A = B.va0 + a, Bv = B.vb0 + b;
  const double V = B.V[i * nc3 + j * nc2 + A * nc + Bv];
  const double so = B.f1[i] + B.f2[j];
  const double sv = B.f1[B.o1 + a] + B.f2[B.o2 + b];
  Tuple t;
  t.m = V / (so - sv);
  t.x[0] = (float)(fabs(t.m) * B.inv_sm);
  t.x[1] = (float)so;
  t.x[2] = (float)sv;
  t.x[3] = (float)B.V[i * nc3 + j * nc2 + i * nc + j];
  t.x[4] = (float)B.V[A * nc3 + Bv * nc2 + A * nc + Bv];
  return t;
}

__device__ inline Tuple guess_tuple(const GuessBlock& B, int64_t e, int64_t& off) {
  int64_t r = e;
  const int64_t b = r % B.v2; r /= B.v2;
  const int64_t a = r % B.v1; r /= B.v1;
  const int64_t j = r % B.o2; r /= B.o2;
  const int64_t i = r;
  off = e;
  return tuple_at(B, i, j, a, b);
}

// p-th pair (x < y) of n in row order (0,1), (0,2), ..., (1,2), ...
__device__ inline void pair_of(int64_t p, int64_t n, int64_t& x, int64_t& y) {
  const int64_t total = n * (n - 1) / 2;
  const int64_t q = total - 1 - p;   // pairs counted from the end
  int64_t k = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((k + 1) * (k + 2) / 2 <= q) ++k;
  while (k * (k + 1) / 2 > q) --k;
  x = n - 2 - k;
  y = n - 1 - (q - k * (k + 1) / 2);
}

__global__ __launch_bounds__(256) void guess_msq(GuessBlock B, double* __restrict__ partial) {
  __shared__ double red[256];
  const int64_t total = B.o1 * B.o2 * B.v1 * B.v2;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t off;
    const Tuple t = guess_tuple(B, e, off);
    acc += t.m * t.m;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// forward pass of one sample; weights (BF16-rounded) in LDS.  Returns r; with
// keep != nullptr stores the rounded input, rounded h1, rounded h2 there.
__device__ inline float mlp_forward(const float* __restrict__ w, const float (&xin)[F], float* keep) {
  float x[F];
#pragma unroll
  for (int k = 0; k < F; ++k) x[k] = bf16r((xin[k] - w[P_MU + k]) / w[P_SD + k]);
  float h1[H];
#pragma unroll 8
  for (int p = 0; p < H; ++p) {
    float acc = w[P_B1 + p];
#pragma unroll
    for (int k = 0; k < F; ++k) acc += w[P_W1 + k * H + p] * x[k];
    h1[p] = bf16r(tanhf(acc));
  }
  float out = w[P_B3];
  for (int q = 0; q < H; ++q) {
    float acc = w[P_B2 + q];
#pragma unroll 16
    for (int p = 0; p < H; ++p) acc += w[P_W2 + p * H + q] * h1[p];
    const float h2 = bf16r(tanhf(acc));
    if (keep) keep[F + H + q] = h2;
    out += w[P_W3 + q] * h2;
  }
  if (keep) {
#pragma unroll
    for (int k = 0; k < F; ++k) keep[k] = x[k];
    for (int p = 0; p < H; ++p) keep[F + p] = h1[p];
  }
  return out;
}

// weights into LDS, BF16-rounded (mu, sd and biases are kept in FP32)
__device__ inline void load_weights(float* w, const float* __restrict__ params) {
  for (int k = threadIdx.x; k < kGuessParams; k += blockDim.x) {
    const bool wt = (k >= P_W1 && k < P_B1) || (k >= P_W2 && k < P_B2) || (k >= P_W3 && k < P_B3);
    w[k] = wt ? bf16r(params[k]) : params[k];
  }
  __syncthreads();
}

// Inference of one sample per thread, the same operations in the same order
// as mlp_forward, with the first hidden layer in LDS (hs: this thread's
// column, stride kApplyThreads) instead of a per-thread array, which the
// run-time indexed loops would otherwise place in scratch memory.
constexpr int kApplyThreads = 128;
__device__ inline float mlp_infer(const float* __restrict__ w, const float (&xin)[F], float* hs) {
  float x[F];
#pragma unroll
  for (int k = 0; k < F; ++k) x[k] = bf16r((xin[k] - w[P_MU + k]) / w[P_SD + k]);
#pragma unroll 8
  for (int p = 0; p < H; ++p) {
    float acc = w[P_B1 + p];
#pragma unroll
    for (int k = 0; k < F; ++k) acc += w[P_W1 + k * H + p] * x[k];
    hs[p * kApplyThreads] = bf16r(tanhf(acc));
  }
  float out = w[P_B3];
  for (int q = 0; q < H; ++q) {
    float acc = w[P_B2 + q];
#pragma unroll 16
    for (int p = 0; p < H; ++p) acc += w[P_W2 + p * H + q] * hs[p * kApplyThreads];
    out += w[P_W3 + q] * bf16r(tanhf(acc));
  }
  return out;
}

__global__ __launch_bounds__(kApplyThreads) void guess_apply(GuessBlock B, const float* __restrict__ params) {
  __shared__ float w[kGuessParams];
  __shared__ float h1s[H * kApplyThreads];
  float* hs = h1s + threadIdx.x;
  if (params) load_weights(w, params);
  if (B.anti) {
    // t[ijab] = -t[jiab] = -t[ijba] = t[jiba]; the network sees features
    // that are even under both swaps, so one evaluation serves all four
    const int64_t po = B.o1 * (B.o1 - 1) / 2, pv = B.v1 * (B.v1 - 1) / 2;
    const int64_t total = po * pv;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
      int64_t i, j, a, b;
      pair_of(e / pv, B.o1, i, j);
      pair_of(e % pv, B.v1, a, b);
      const Tuple t = tuple_at(B, i, j, a, b);
      double v = t.m;
      if (params) v = t.m * (1.0 + (double)mlp_infer(w, t.x, hs));
      const int64_t o = B.o2, vv = B.v2;
      B.t[((i * o + j) * B.v1 + a) * vv + b] = v;
      B.t[((j * o + i) * B.v1 + a) * vv + b] = -v;
      B.t[((i * o + j) * B.v1 + b) * vv + a] = -v;
      B.t[((j * o + i) * B.v1 + b) * vv + a] = v;
    }
    return;
  }
  const int64_t total = B.o1 * B.o2 * B.v1 * B.v2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t off;
    const Tuple t = guess_tuple(B, e, off);
    double v = t.m;
    if (params) v = t.m * (1.0 + (double)mlp_infer(w, t.x, hs));
    B.t[off] = v;
  }
}

__global__ __launch_bounds__(256) void guess_data(GuessBlock B, float* __restrict__ X,
                                                  float* __restrict__ M, float* __restrict__ Y) {
  const int64_t total = B.o1 * B.o2 * B.v1 * B.v2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t off;
    const Tuple t = guess_tuple(B, e, off);
#pragma unroll
    for (int k = 0; k < F; ++k) X[e * F + k] = t.x[k];
    M[e] = (float)(t.m * B.inv_sm);
    Y[e] = (float)((B.t[off] - t.m) * B.inv_sm);
  }
}

static int grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

hipError_t launch_guess_msq(hipStream_t s, const GuessBlock& B, double* partial, int blocks) {
  hipLaunchKernelGGL(guess_msq, dim3(blocks), dim3(256), 0, s, B, partial);
  return hipGetLastError();
}

hipError_t launch_guess_apply(hipStream_t s, const GuessBlock& B, const float* params) {
  if (B.anti && (B.o1 != B.o2 || B.v1 != B.v2 || B.va0 != B.vb0 || B.f1 != B.f2))
    return hipErrorInvalidValue;
  const int64_t total = B.anti ? (B.o1 * (B.o1 - 1) / 2) * (B.v1 * (B.v1 - 1) / 2)
                               : B.o1 * B.o2 * B.v1 * B.v2;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + kApplyThreads - 1) / kApplyThreads;
  hipLaunchKernelGGL(guess_apply, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(kApplyThreads), 0, s, B,
                     params);
  return hipGetLastError();
}

hipError_t launch_guess_data(hipStream_t s, const GuessBlock& B, float* X, float* M, float* Y) {
  const int64_t total = B.o1 * B.o2 * B.v1 * B.v2;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(guess_data, dim3(grid_of(total)), dim3(256), 0, s, B, X, M, Y);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// training

constexpr int KEEP = F + 2 * H;           // rounded x, h1, h2 of a sample
constexpr int SCR = KEEP + 2 * H + 2;     // + gh1, gh2, g_out, e^2

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// one thread per sample of the batch: forward, loss and the back-propagated
// signals gh1, gh2, g_out (activations in scr[s][...])
__global__ __launch_bounds__(256) void mlp_fwd_bwd(const float* __restrict__ params,
                                                   const float* __restrict__ X,
                                                   const float* __restrict__ Mv,
                                                   const float* __restrict__ Yv, int64_t n,
                                                   int64_t batch, uint64_t seed, int64_t step,
                                                   int full, float* __restrict__ scr) {
  __shared__ float w[kGuessParams];
  load_weights(w, params);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= batch) return;
  const int64_t idx = full ? s : (int64_t)(mix64(seed ^ mix64((uint64_t)step * 0x100000001ull + s)) %
                                           (uint64_t)n);
  float x[F];
#pragma unroll
  for (int k = 0; k < F; ++k) x[k] = X[idx * F + k];
  float* q = scr + s * SCR;
  const float r = mlp_forward(w, x, q);
  const float mm = Mv[idx];
  const float e = mm * r - Yv[idx];
  const float g = 2.0f * e * mm / (float)batch;
  q[KEEP + 2 * H] = g;
  q[KEEP + 2 * H + 1] = e * e;
  // gh2 = g W3 (1 - h2^2); gh1 = (W2 gh2) (1 - h1^2)
  float gh2[H];
  for (int qq = 0; qq < H; ++qq) {
    const float h2 = q[F + H + qq];
    gh2[qq] = g * w[P_W3 + qq] * (1.0f - h2 * h2);
    q[KEEP + H + qq] = gh2[qq];
  }
  for (int p = 0; p < H; ++p) {
    float acc = 0.0f;
#pragma unroll 16
    for (int qq = 0; qq < H; ++qq) acc += w[P_W2 + p * H + qq] * gh2[qq];
    const float h1 = q[F + p];
    q[KEEP + p] = acc * (1.0f - h1 * h1);
  }
}

// one thread per parameter (and one for the loss): sum over the batch in
// sample order.  grad[k] for k < kGuessParams (mu / sd get 0), grad[kGuessParams] = loss.
__global__ __launch_bounds__(256) void mlp_grad(const float* __restrict__ scr, int64_t batch,
                                                float* __restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > kGuessParams) return;
  double acc = 0.0;
  if (k == kGuessParams) {
    for (int64_t s = 0; s < batch; ++s) acc += scr[s * SCR + KEEP + 2 * H + 1];
    grad[k] = (float)(acc / (double)batch);
    return;
  }
  int a = -1, b = -1;   // gradient = sum_s act[a] * sig[b] (a = -1: sig alone)
  if (k < P_W1) {
    grad[k] = 0.0f;
    return;
  } else if (k < P_B1) {
    const int kk = (k - P_W1) / H, p = (k - P_W1) % H;
    a = kk;                 // rounded input x[kk]
    b = KEEP + p;           // gh1[p]
  } else if (k < P_W2) {
    b = KEEP + (k - P_B1);
  } else if (k < P_B2) {
    const int p = (k - P_W2) / H, qq = (k - P_W2) % H;
    a = F + p;              // h1[p]
    b = KEEP + H + qq;      // gh2[q]
  } else if (k < P_W3) {
    b = KEEP + H + (k - P_B2);
  } else if (k < P_B3) {
    a = F + H + (k - P_W3); // h2[q]
    b = KEEP + 2 * H;       // g_out
  } else {
    b = KEEP + 2 * H;
  }
  for (int64_t s = 0; s < batch; ++s) {
    const float* q = scr + s * SCR;
    acc += a >= 0 ? (double)q[a] * (double)q[b] : (double)q[b];
  }
  grad[k] = (float)acc;
}

__global__ __launch_bounds__(256) void mlp_adam(float* __restrict__ params, const float* __restrict__ grad,
                                                float* __restrict__ m1, float* __restrict__ m2, float lr,
                                                float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < P_W1 || k >= kGuessParams) return;
  const float g = grad[k];
  m1[k] = 0.9f * m1[k] + 0.1f * g;
  m2[k] = 0.999f * m2[k] + 0.001f * g * g;
  params[k] -= lr * (m1[k] / c1) / (sqrtf(m2[k] / c2) + 1e-8f);
}

}  // namespace ccqk

// ---------------------------------------------------------------------------
// C ABI (standalone: device-resident training data, one GPU)

namespace {

thread_local std::string g_guess_err;

struct DevF {
  float* p = nullptr;
  explicit DevF(int64_t n) {
    if (hipMalloc(&p, sizeof(float) * (n > 0 ? n : 1)) != hipSuccess) p = nullptr;
  }
  ~DevF() {
    if (p) (void)hipFree(p);
  }
};

#define GCHK(x)                                                            \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      g_guess_err = std::string(#x) + ": " + hipGetErrorString(e_);        \
      return 1;                                                            \
    }                                                                      \
  } while (0)

// loss and gradient over all n samples (device buffers; scr holds n * SCR
// floats, grad kGuessParams + 1), on the null stream
hipError_t launch_full_grad(const float* P, const float* X, const float* M, const float* Y,
                            int64_t n, float* scr, float* grad) {
  using namespace ccqk;
  hipLaunchKernelGGL(mlp_fwd_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, P, X, M, Y,
                     n, n, (uint64_t)0, (int64_t)0, 1, scr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp_grad, dim3((kGuessParams + 1 + 255) / 256), dim3(256), 0, 0, scr, n, grad);
  return hipGetLastError();
}

int run_training(const float* X, const float* M, const float* Y, int64_t n, float* params,
                 int64_t nparams, int64_t steps, int64_t batch, double lr, uint64_t seed,
                 double* loss_out, float* grad_out) {
  using namespace ccqk;
  if (nparams != kGuessParams) {
    g_guess_err = "nparams must be " + std::to_string(kGuessParams);
    return 2;
  }
  if (n <= 0 || batch <= 0 || (grad_out && batch != n)) {
    g_guess_err = "bad sample / batch count";
    return 2;
  }
  DevF dX(n * F), dM(n), dY(n), dP(kGuessParams), dG(kGuessParams + 1), dm1(kGuessParams),
      dm2(kGuessParams), dS(batch * SCR);
  if (!dX.p || !dM.p || !dY.p || !dP.p || !dG.p || !dm1.p || !dm2.p || !dS.p) {
    g_guess_err = "out of device memory";
    return 1;
  }
  GCHK(hipMemcpy(dX.p, X, sizeof(float) * n * F, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(dM.p, M, sizeof(float) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(dY.p, Y, sizeof(float) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(dP.p, params, sizeof(float) * kGuessParams, hipMemcpyHostToDevice));
  GCHK(hipMemset(dm1.p, 0, sizeof(float) * kGuessParams));
  GCHK(hipMemset(dm2.p, 0, sizeof(float) * kGuessParams));
  const dim3 gp((kGuessParams + 1 + 255) / 256);
  if (grad_out) {
    GCHK(launch_full_grad(dP.p, dX.p, dM.p, dY.p, n, dS.p, dG.p));
    steps = 0;
  }
  const dim3 gb((unsigned)((batch + 255) / 256));
  double b1 = 1.0, b2 = 1.0;
  for (int64_t st = 0; st < steps; ++st) {
    hipLaunchKernelGGL(mlp_fwd_bwd, gb, dim3(256), 0, 0, dP.p, dX.p, dM.p, dY.p, n, batch, seed, st,
                       0, dS.p);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(mlp_grad, gp, dim3(256), 0, 0, dS.p, batch, dG.p);
    GCHK(hipGetLastError());
    b1 *= 0.9;
    b2 *= 0.999;
    hipLaunchKernelGGL(mlp_adam, gp, dim3(256), 0, 0, dP.p, dG.p, dm1.p, dm2.p, (float)lr,
                       (float)(1.0 - b1), (float)(1.0 - b2));
    GCHK(hipGetLastError());
  }
  float loss = 0.0f;
  GCHK(hipMemcpy(&loss, dG.p + kGuessParams, sizeof(float), hipMemcpyDeviceToHost));
  if (loss_out) *loss_out = loss;
  if (grad_out) GCHK(hipMemcpy(grad_out, dG.p, sizeof(float) * kGuessParams, hipMemcpyDeviceToHost));
  else GCHK(hipMemcpy(params, dP.p, sizeof(float) * kGuessParams, hipMemcpyDeviceToHost));
  GCHK(hipDeviceSynchronize());
  return 0;
}

// ---------------------------------------------------------------------------
// Gradient without device allocation: every device buffer of the gradient
// (the staged samples and parameters, the per-sample activations and
// signals, the gradient and loss) lives in one file-scope array.  Each GPU
// holds its own copy until the process ends; its address is looked up once
// per device and used for every kernel argument and copy.

constexpr int64_t kNaMax = CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES;
constexpr int64_t kNaX = 0;                                   // X[n][F]
constexpr int64_t kNaM = kNaX + kNaMax * ccqk::F;             // M[n]
constexpr int64_t kNaY = kNaM + kNaMax;                       // Y[n]
constexpr int64_t kNaP = kNaY + kNaMax;                       // params
constexpr int64_t kNaG = kNaP + 64 * ((ccqk::kGuessParams + 63) / 64);        // grad + loss
constexpr int64_t kNaS = kNaG + 64 * ((ccqk::kGuessParams + 1 + 63) / 64);    // scr[n][SCR]
constexpr int64_t kNaFloats = kNaS + kNaMax * ccqk::SCR;

constexpr int kNaDevices = 64;
std::mutex g_na_mu;
float* g_na_ptr[kNaDevices] = {};

}  // namespace

static __device__ __align__(256) float scratch[kNaFloats];

namespace {

// the device address of this GPU's scratch (looked up on first use)
float* noalloc_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_guess_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_na_ptr[dev] = static_cast<float*>(p);
  }
  return g_na_ptr[dev];
}

}  // namespace

extern "C" {

int ccq_hip_mlp_grad_noalloc(const float* X, const float* M, const float* Y, int64_t n,
                             const float* params, int64_t nparams, float* grad, double* loss) {
  using namespace ccqk;
  if (nparams != kGuessParams) {
    g_guess_err = "nparams must be " + std::to_string(kGuessParams);
    return 2;
  }
  if (n <= 0 || n > kNaMax) {
    g_guess_err = "sample count must be in 1.." + std::to_string(kNaMax) +
                  " (CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES) for ccq_hip_mlp_grad_noalloc";
    return 2;
  }
  if (!X || !M || !Y || !params || !grad) {
    g_guess_err = "null array";
    return 2;
  }
  int dev = 0;
  GCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaDevices) {
    g_guess_err = "device ordinal beyond the scratch table";
    return 2;
  }
  float* s = noalloc_scratch(dev);
  if (!s) return 1;
  GCHK(hipMemcpy(s + kNaX, X, sizeof(float) * n * F, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(s + kNaM, M, sizeof(float) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(s + kNaY, Y, sizeof(float) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(s + kNaP, params, sizeof(float) * kGuessParams, hipMemcpyHostToDevice));
  GCHK(launch_full_grad(s + kNaP, s + kNaX, s + kNaM, s + kNaY, n, s + kNaS, s + kNaG));
  float l = 0.0f;
  GCHK(hipMemcpy(&l, s + kNaG + kGuessParams, sizeof(float), hipMemcpyDeviceToHost));
  GCHK(hipMemcpy(grad, s + kNaG, sizeof(float) * kGuessParams, hipMemcpyDeviceToHost));
  GCHK(hipDeviceSynchronize());
  if (loss) *loss = l;
  return 0;
}

int64_t ccq_hip_mlp_nparams(void) { return ccqk::kGuessParams; }

int ccq_hip_mlp_train(const float* X, const float* M, const float* Y, int64_t n, float* params,
                      int64_t nparams, int64_t steps, int64_t batch, double lr, uint64_t seed,
                      double* final_loss) {
  return run_training(X, M, Y, n, params, nparams, steps, batch, lr, seed, final_loss, nullptr);
}

int ccq_hip_mlp_grad(const float* X, const float* M, const float* Y, int64_t n, const float* params,
                     int64_t nparams, float* grad, double* loss) {
  std::vector<float> p(params, params + (nparams > 0 ? nparams : 0));
  return run_training(X, M, Y, n, p.data(), nparams, 1, n, 0.0, 0, loss, grad);
}

const char* ccq_hip_mlp_last_error(void) { return g_guess_err.c_str(); }

}  // extern "C"
