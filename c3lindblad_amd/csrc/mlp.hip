// Learned warm start: training and inference kernels of the amplitude MLP.
//
// Weights and layer inputs are stored in bfloat16, every product accumulates in
// FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16), and the optimiser keeps
// FP32 master weights.  Shapes are small (batch x 3073 -> 512 -> 512 -> 512 -> 64),
// so one general strided GEMM serves the forward pass and both backward products.
//
// GEMM tiling: a 256-thread workgroup owns a 64x64 output tile, each of its four
// waves a 32x32 quarter (2x2 MFMA blocks of 16x16); the K loop stages a 64x16 A
// tile and a 16x64 B tile in LDS (zero padded at the edges).  MFMA 16x16x16 bf16
// operand layout: lane l holds A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]
// (i = 0..3); the accumulator lane l holds C[4 (l / 16) + i][l % 16].

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/c3l_hip.h"

extern "C" int c3l_internal_fail(int code, const char* message);

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int kTile = 64;
constexpr int kK = 16;

// C[m, n] = alpha * sum_k A(m, k) B(k, n) + beta * C[m, n];
// A(m, k) = A[m * a_rs + k * a_cs], B(k, n) = B[k * b_rs + n * b_cs], C row-major (ldc).
__global__ __launch_bounds__(256) void bf16_gemm_kernel(const uint16_t* __restrict__ A, int64_t a_rs, int64_t a_cs,
                                                        const uint16_t* __restrict__ B, int64_t b_rs, int64_t b_cs,
                                                        float* __restrict__ C, int64_t ldc, int M, int N, int K,
                                                        float alpha, float beta) {
  __shared__ uint16_t As[kTile][kK + 4];  // [m][k]
  __shared__ uint16_t Bs[kK][kTile + 4];  // [k][n]
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  v4f acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kK) {
    // 64x16 A tile and 16x64 B tile: 1024 elements each, 4 per thread
    for (int e = t; e < kTile * kK; e += 256) {
      const int r = e / kK, c = e % kK;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < M && gk < K) ? A[(int64_t)gm * a_rs + (int64_t)gk * a_cs] : (uint16_t)0;
      const int kr = e / kTile, nc = e % kTile;
      const int bk = k0 + kr, bn = n0 + nc;
      Bs[kr][nc] = (bk < K && bn < N) ? B[(int64_t)bk * b_rs + (int64_t)bn * b_cs] : (uint16_t)0;
    }
    __syncthreads();
    const int r = lane & 15, q = (lane >> 4) * 4;
    for (int a = 0; a < 2; ++a) {
      v4s fa;
      for (int i = 0; i < 4; ++i) fa[i] = (short)As[wm + 16 * a + r][q + i];
      for (int b = 0; b < 2; ++b) {
        v4s fb;
        for (int i = 0; i < 4; ++i) fb[i] = (short)Bs[q + i][wn + 16 * b + r];
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa, fb, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int col = lane & 15, row4 = (lane >> 4) * 4;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int i = 0; i < 4; ++i) {
        const int gm = m0 + wm + 16 * a + row4 + i, gn = n0 + wn + 16 * b + col;
        if (gm < M && gn < N) {
          float* dst = C + (int64_t)gm * ldc + gn;
          *dst = alpha * acc[a][b][i] + (beta != 0.f ? beta * *dst : 0.f);
        }
      }
}

__device__ __forceinline__ float gelu(float z) {
  // tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  return 0.5f * z * (1.f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float z) {
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  const float th = tanhf(u);
  const float du = 0.7978845608028654f * (1.f + 3.f * 0.044715f * z * z);
  return 0.5f * (1.f + th) + 0.5f * z * (1.f - th * th) * du;
}

// z[m, n] += bias[n]; act != 0: h = bf16(gelu(z)), else h = bf16(z) (out may be null).
__global__ void bias_act_kernel(float* __restrict__ z, const float* __restrict__ bias, uint16_t* __restrict__ h, int M,
                                int N, int act) {
  const int64_t total = (int64_t)M * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = z[i] + bias[i % N];
    z[i] = v;
    if (h != nullptr) h[i] = f32_to_bf16(act ? gelu(v) : v);
  }
}

// dz = dh * gelu'(z), also as bf16 for the next products
__global__ void gelu_backward_kernel(const float* __restrict__ z, const float* __restrict__ dh, float* __restrict__ dz,
                                     uint16_t* __restrict__ dz16, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float g = dh[i] * gelu_grad(z[i]);
    dz[i] = g;
    dz16[i] = f32_to_bf16(g);
  }
}

// db[n] = sum_m dz[m, n] (one thread per column; M is a batch)
__global__ void column_sum_kernel(const float* __restrict__ dz, float* __restrict__ db, int M, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dz[(int64_t)m * N + n];
  db[n] = s;
}

// loss_parts[block] = sum (y - t)^2 over the block's elements; dy = scale * (y - t)
__global__ void mse_kernel(const float* __restrict__ y, const float* __restrict__ target, float* __restrict__ dy,
                           uint16_t* __restrict__ dy16, float* __restrict__ loss_parts, int64_t total, float scale) {
  __shared__ float part[256];
  float s = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const float d = y[i] - target[i];
    s += d * d;
    const float g = scale * d;
    dy[i] = g;
    dy16[i] = f32_to_bf16(g);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_parts[blockIdx.x] = part[0];
}

// Adam on FP32 master weights; the bf16 copy used by the products is refreshed.
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, uint16_t* __restrict__ w16, int64_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16 != nullptr) w16[i] = f32_to_bf16(wi);
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = f32_to_bf16(a[i]);
}

int64_t blocks_of(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

int mfail(const char* msg) { return c3l_internal_fail(C3L_EINVAL, msg); }

#define M_CHECK()                                                                        \
  do {                                                                                   \
    hipError_t e__ = hipGetLastError();                                                  \
    if (e__ != hipSuccess) return c3l_internal_fail(C3L_EHIP, hipGetErrorString(e__));  \
  } while (0)

}  // namespace

extern "C" {

int c3l_bf16_gemm(const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs,
                  float* C, int64_t ldc, int32_t M, int32_t N, int32_t K, float alpha, float beta, void* stream) {
  if (A == nullptr || B == nullptr || C == nullptr || M <= 0 || N <= 0 || K <= 0 || ldc < N)
    return mfail("c3l_bf16_gemm: invalid argument");
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  hipLaunchKernelGGL(bf16_gemm_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), A, a_rs, a_cs, B, b_rs,
                     b_cs, C, ldc, M, N, K, alpha, beta);
  M_CHECK();
  return C3L_OK;
}

int c3l_mlp_bias_act(float* z, const float* bias, uint16_t* h, int32_t M, int32_t N, int32_t act, void* stream) {
  if (z == nullptr || bias == nullptr || M <= 0 || N <= 0) return mfail("c3l_mlp_bias_act: invalid argument");
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_of((int64_t)M * N)), dim3(256), 0, static_cast<hipStream_t>(stream), z,
                     bias, h, M, N, act);
  M_CHECK();
  return C3L_OK;
}

int c3l_mlp_gelu_backward(const float* z, const float* dh, float* dz, uint16_t* dz16, int64_t total, void* stream) {
  if (z == nullptr || dh == nullptr || dz == nullptr || dz16 == nullptr || total <= 0)
    return mfail("c3l_mlp_gelu_backward: invalid argument");
  hipLaunchKernelGGL(gelu_backward_kernel, dim3(blocks_of(total)), dim3(256), 0, static_cast<hipStream_t>(stream), z, dh,
                     dz, dz16, total);
  M_CHECK();
  return C3L_OK;
}

int c3l_mlp_column_sum(const float* dz, float* db, int32_t M, int32_t N, void* stream) {
  if (dz == nullptr || db == nullptr || M <= 0 || N <= 0) return mfail("c3l_mlp_column_sum: invalid argument");
  hipLaunchKernelGGL(column_sum_kernel, dim3((N + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), dz, db, M,
                     N);
  M_CHECK();
  return C3L_OK;
}

int c3l_mlp_mse(const float* y, const float* target, float* dy, uint16_t* dy16, float* loss_parts, int64_t total,
                float scale, int32_t n_parts, void* stream) {
  if (y == nullptr || target == nullptr || dy == nullptr || dy16 == nullptr || loss_parts == nullptr || total <= 0 ||
      n_parts <= 0)
    return mfail("c3l_mlp_mse: invalid argument");
  hipLaunchKernelGGL(mse_kernel, dim3(n_parts), dim3(256), 0, static_cast<hipStream_t>(stream), y, target, dy, dy16,
                     loss_parts, total, scale);
  M_CHECK();
  return C3L_OK;
}

int c3l_mlp_adam(float* w, float* m, float* v, const float* g, uint16_t* w16, int64_t n, float lr, float b1, float b2,
                 float eps, int32_t step, void* stream) {
  if (w == nullptr || m == nullptr || v == nullptr || g == nullptr || n <= 0 || step <= 0)
    return mfail("c3l_mlp_adam: invalid argument");
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_of(n)), dim3(256), 0, static_cast<hipStream_t>(stream), w, m, v, g, w16, n,
                     lr, b1, b2, eps, c1, c2);
  M_CHECK();
  return C3L_OK;
}

int c3l_to_bf16(const float* a, uint16_t* b, int64_t n, void* stream) {
  if (a == nullptr || b == nullptr || n <= 0) return mfail("c3l_to_bf16: invalid argument");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_of(n)), dim3(256), 0, static_cast<hipStream_t>(stream), a, b, n);
  M_CHECK();
  return C3L_OK;
}

}  // extern "C"
