// BF16 MLP kernels (training and inference) for the learned (T) screen.
//
// One GEMM form serves all layers: C[m][n] = sum_k A[m][k] * B[n][k] with bf16
// operands that are contiguous in k (v_mfma_f32_16x16x32_bf16, FP32
// accumulation).  Forward Z = X W^T + b, Y = act(Z); backward dX = dZ W (from a
// transposed copy Wt), dW = dZ^T X (from transposed copies), db = sum_b dZ.  The
// elementwise kernels write each bf16 activation both row-major and transposed
// so that every GEMM operand is k-contiguous.  Adam keeps FP32 master weights
// and refreshes the bf16 copies W and Wt.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bf16_split.h"

namespace pamd {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) { return (uint16_t)bf16_bits(f); }

// 8 bf16 along k from a k-contiguous row; zero outside [0, K) (K % 8 == 0)
__device__ __forceinline__ bf16x8 load8(const uint16_t *row, int k, int K, bool rok) {
  u32x4 w = {0u, 0u, 0u, 0u};
  if (rok && k < K) w = *(const u32x4 *)(row + k);
  return __builtin_bit_cast(bf16x8, w);
}

// C (fp32) = A B^T (+ bias[n]) for M x N, one 16x16 tile per wave, 4 waves/block.
// Epilogue modes: 0 store C; 1 store C + bias; C must be M x ldc fp32.
__global__ __launch_bounds__(256) void k_gemm_bf16_tn(int M, int N, int K, const uint16_t *A, int lda,
                                                      const uint16_t *B, int ldb, const float *bias,
                                                      float *C, int ldc, int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_n = (N + 15) / 16;
  const int tile = blockIdx.x * 4 + wave;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  if (tm * 16 >= M) return;
  const int ra = tm * 16 + (lane & 15), rb = tn * 16 + (lane & 15);
  const bool oka = ra < M, okb = rb < N;
  const uint16_t *pa = A + (size_t)(oka ? ra : 0) * lda;
  const uint16_t *pb = B + (size_t)(okb ? rb : 0) * ldb;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    const int k = k0 + 8 * (lane >> 4);
    const bf16x8 a = load8(pa, k, K, oka);
    const bf16x8 b = load8(pb, k, K, okb);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  // D layout: lane l, element r -> row 4*(l/16) + r, column l%16
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = tm * 16 + 4 * (lane >> 4) + r, n = tn * 16 + (lane & 15);
    if (m < M && n < N) {
      float v = acc[r] + (bias ? bias[n] : 0.f);
      float *c = C + (size_t)m * ldc + n;
      *c = accumulate ? *c + v : v;
    }
  }
}

hipError_t launch_gemm_bf16_tn(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B,
                               int ldb, const float *bias, float *C, int ldc, int accumulate,
                               hipStream_t s) {
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL(k_gemm_bf16_tn, dim3((tiles + 3) / 4), dim3(256), 0, s, M, N, K, A, lda, B, ldb,
                     bias, C, ldc, accumulate);
  return hipGetLastError();
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float sg = 1.f / (1.f + __expf(-z));
  return sg * (1.f + z * (1.f - sg));
}

// Y = act(Z) in bf16, row-major [M][N] and transposed [N][M]
__global__ void k_act_out(int M, int N, const float *Z, int act, uint16_t *Y, uint16_t *Yt) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)M * N) return;
  const int m = (int)(t / N), n = (int)(t % N);
  const float z = Z[t];
  const uint16_t y = f32_to_bf16(act ? silu(z) : z);
  if (Y) Y[t] = y;
  if (Yt) Yt[(size_t)n * M + m] = y;
}

// dZ = dY * act'(Z) -> bf16 row-major and transposed; db[n] = sum_m dZ[m][n]
__global__ void k_act_back(int M, int N, const float *Z, const float *dY, int act, uint16_t *dZ,
                           uint16_t *dZt) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)M * N) return;
  const int m = (int)(t / N), n = (int)(t % N);
  const float g = dY[t] * (act ? dsilu(Z[t]) : 1.f);
  const uint16_t b = f32_to_bf16(g);
  dZ[t] = b;
  dZt[(size_t)n * M + m] = b;
}

__global__ void k_colsum(int M, int N, const uint16_t *dZ, float *db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += bf16_to_f32(dZ[(size_t)m * N + n]);
  db[n] = s;
}

// Adam (with decoupled weight decay) on FP32 masters; refresh bf16 W [R][C] and Wt [C][R]
__global__ void k_adam(int R, int Ccols, float *w, const float *g, float *m1, float *m2, uint16_t *wb,
                       uint16_t *wt, float lr, float b1, float b2, float eps, float wd, float c1,
                       float c2) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)R * Ccols;
  if (t >= n) return;
  const float gi = g[t];
  const float a = b1 * m1[t] + (1.f - b1) * gi;
  const float b = b2 * m2[t] + (1.f - b2) * gi * gi;
  m1[t] = a;
  m2[t] = b;
  float x = w[t];
  x -= lr * ((a / c1) / (sqrtf(b / c2) + eps) + wd * x);
  w[t] = x;
  if (wb) {
    const uint16_t xb = f32_to_bf16(x);
    wb[t] = xb;
    if (wt) wt[(t % Ccols) * R + t / Ccols] = xb;
  }
}

// mean squared error over n values, and its gradient 2 (p - y) / n
__global__ void k_mse(int n, const float *pred, const float *y, float *dpred, float *loss) {
  __shared__ float red[256];
  float s = 0.f;
  for (int t = threadIdx.x; t < n; t += 256) {
    const float d = pred[t] - y[t];
    s += d * d;
    if (dpred) dpred[t] = 2.f * d / n;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = red[0] / n;
}

__global__ void k_f32_to_bf16(size_t n, const float *x, uint16_t *y, int M, int N, uint16_t *yt) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint16_t b = f32_to_bf16(x[t]);
  y[t] = b;
  if (yt) yt[(t % N) * M + t / N] = b;
}

static unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_act_out(int M, int N, const float *Z, int act, uint16_t *Y, uint16_t *Yt,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_act_out, dim3(blocks((size_t)M * N)), dim3(256), 0, s, M, N, Z, act, Y, Yt);
  return hipGetLastError();
}

hipError_t launch_act_back(int M, int N, const float *Z, const float *dY, int act, uint16_t *dZ,
                           uint16_t *dZt, float *db, hipStream_t s) {
  hipLaunchKernelGGL(k_act_back, dim3(blocks((size_t)M * N)), dim3(256), 0, s, M, N, Z, dY, act, dZ,
                     dZt);
  hipLaunchKernelGGL(k_colsum, dim3(blocks(N)), dim3(256), 0, s, M, N, dZ, db);
  return hipGetLastError();
}

hipError_t launch_adam(int R, int Ccols, float *w, const float *g, float *m1, float *m2, uint16_t *wb,
                       uint16_t *wt, float lr, float b1, float b2, float eps, float wd, int step,
                       hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks((size_t)R * Ccols)), dim3(256), 0, s, R, Ccols, w, g, m1, m2,
                     wb, wt, lr, b1, b2, eps, wd, c1, c2);
  return hipGetLastError();
}

hipError_t launch_mse(int n, const float *pred, const float *y, float *dpred, float *loss,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, s, n, pred, y, dpred, loss);
  return hipGetLastError();
}

hipError_t launch_f32_to_bf16(size_t n, const float *x, uint16_t *y, int M, int N, uint16_t *yt,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n)), dim3(256), 0, s, n, x, y, M, N, yt);
  return hipGetLastError();
}

}  // namespace pamd
