// Point-set MLP of the learned block proposal on gfx950 matrix cores.
//
//   k_gemm       C = act(alpha * A B^T + bias) (+ C), A and B bf16 with any
//                strides, FP32 accumulation on v_mfma_f32_16x16x16_bf16:
//                64 x 64 output tile per 256-lane workgroup, 32 x 32 per
//                wave (2 x 2 MFMA tiles), K in steps of 16.
//   k_point_pool the shared per-point MLP 3 -> 64 -> 128 -> 256 (ReLU) fused
//                with the max-pool over each scene's points: 64 points per
//                workgroup, layer 1 on the VALU in FP32, layers 2 and 3 on
//                MFMA with bf16 activations staged in LDS, then a per-channel
//                (max, first argmax) reduction merged across workgroups with
//                one 64-bit atomic max per channel.
//   k_adam       Adam step on FP32 master weights plus their bf16 copy.
//
// MFMA 16x16x16 bf16 lane layout (wave64): A holds row (lane % 16), k =
// 4 (lane / 16) + i; B holds column (lane % 16), same k; the four results
// are rows 4 (lane / 16) + j of column (lane % 16).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/set_proposal.h"

namespace sp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  long lda_m, lda_k;
  const uint16_t* B;
  long ldb_n, ldb_k;
  const float* bias;
  float alpha;
  int relu, accumulate;
  float* C;
  long ldc;
  uint16_t* Cb;  // optional bf16 copy of the result (same ldc)
};

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    bf16x4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int m = m0 + i * 16 + r;
      const int n = n0 + i * 16 + r;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = k0 + kq + t;
        a[i][t] = (m < g.M && k < g.K) ? (short)g.A[(long)m * g.lda_m + (long)k * g.lda_k] : (short)0;
        b[i][t] = (n < g.N && k < g.K) ? (short)g.B[(long)n * g.ldb_n + (long)k * g.ldb_k] : (short)0;
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + i * 16 + kq + t;
        const int n = n0 + j * 16 + r;
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][t];
        if (g.bias) v += g.bias[n];
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        float* c = g.C + (long)m * g.ldc + n;
        if (g.accumulate) v += *c;
        *c = v;
        if (g.Cb) g.Cb[(long)m * g.ldc + n] = f2bf(v);
      }
}

// ---------------------------------------------------------------------------
constexpr int kPts = 64;              // points per workgroup
constexpr int kH1 = 64, kH2 = 128, kH3 = 256;
constexpr int kPad = 4;               // bf16 padding per LDS row (bank spread)

struct PoolArgs {
  int B, N;
  const float* pts;      // B x N x 3
  const float* W1;       // 64 x 3
  const float* b1;       // 64
  const uint16_t* W2;    // 128 x 64 bf16
  const float* b2;
  const uint16_t* W3;    // 256 x 128 bf16
  const float* b3;
  unsigned long long* pooled;  // B x 256: (value bits << 32) | (0xffffffff - point)
};

__global__ __launch_bounds__(256) void k_point_pool(PoolArgs a) {
  __shared__ uint16_t s_h1[kPts][kH1 + kPad];
  __shared__ uint16_t s_h2[kPts][kH2 + kPad];
  const int b = blockIdx.y;
  const int p0 = blockIdx.x * kPts;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  // layer 1 (FP32): 64 points x 64 channels, 16 per thread
  for (int e = tid; e < kPts * kH1; e += 256) {
    const int p = e / kH1, c = e % kH1;
    float h = 0.0f;
    if (p0 + p < a.N) {
      const float* x = a.pts + ((long)b * a.N + p0 + p) * 3;
      h = ((a.W1[3 * c] * x[0] + a.W1[3 * c + 1] * x[1]) + a.W1[3 * c + 2] * x[2]) + a.b1[c];
      h = h > 0.0f ? h : 0.0f;
    }
    s_h1[p][c] = f2bf(h);
  }
  __syncthreads();
  // layer 2 (MFMA): wave w owns channels [32 w, 32 w + 32), all 64 points
  {
    f32x4 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k0 = 0; k0 < kH1; k0 += 16) {
      bf16x4 av[4], bv[2];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int t = 0; t < 4; ++t) av[i][t] = (short)s_h1[i * 16 + r][k0 + kq + t];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint16_t* w = a.W2 + (long)(wave * 32 + j * 16 + r) * kH1 + k0 + kq;
#pragma unroll
        for (int t = 0; t < 4; ++t) bv[j][t] = (short)w[t];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma(av[i], bv[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int c = wave * 32 + j * 16 + r;
          float v = acc[i][j][t] + a.b2[c];
          s_h2[i * 16 + kq + t][c] = f2bf(v > 0.0f ? v : 0.0f);
        }
  }
  __syncthreads();
  // layer 3 (MFMA): wave w owns channels [64 w, 64 w + 64), all 64 points
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < kH2; k0 += 16) {
    bf16x4 av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) av[i][t] = (short)s_h2[i * 16 + r][k0 + kq + t];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint16_t* w = a.W3 + (long)(wave * 64 + j * 16 + r) * kH2 + k0 + kq;
#pragma unroll
      for (int t = 0; t < 4; ++t) bv[j][t] = (short)w[t];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(av[i], bv[j], acc[i][j]);
  }
  // bias + ReLU, then (max, first argmax) per channel over the 64 points
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = wave * 64 + j * 16 + r;
    const float bias = a.b3[c];
    float best = -1.0f;
    int bp = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int p = p0 + i * 16 + kq + t;
        float v = acc[i][j][t] + bias;
        v = v > 0.0f ? v : 0.0f;
        if (p < a.N && (v > best || (v == best && p < bp))) {
          best = v;
          bp = p;
        }
      }
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
      const float ov = __shfl_xor(best, off, 64);
      const int op = __shfl_xor(bp, off, 64);
      if (ov > best || (ov == best && op < bp)) {
        best = ov;
        bp = op;
      }
    }
    if (lane < 16 && best >= 0.0f) {
      const unsigned long long key =
          ((unsigned long long)__float_as_uint(best) << 32) | (unsigned long long)(0xffffffffu - (uint32_t)bp);
      atomicMax(a.pooled + (long)b * kH3 + c, key);
    }
  }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_unpack_pool(const unsigned long long* pooled, long n, float* value, int* index) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = pooled[i];
  value[i] = __uint_as_float((uint32_t)(k >> 32));
  index[i] = (int)(0xffffffffu - (uint32_t)(k & 0xffffffffull));
}

__global__ void k_to_bf16(const float* x, uint16_t* y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

__global__ void k_adam(float* p, uint16_t* pb, const float* g, float* m, float* v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

unsigned blocks(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace sp

using namespace sp;

extern "C" {

int sp_gemm_bf16(int M, int N, int K, const uint16_t* A, long lda_m, long lda_k, const uint16_t* B, long ldb_n,
                 long ldb_k, const float* bias, float alpha, int relu, int accumulate, float* C, long ldc,
                 uint16_t* C_bf16, sp_stream_t stream) {
  if (M < 0 || N < 0 || K < 0 || (M > 0 && N > 0 && (!A || !B || !C))) return 1;
  if (M == 0 || N == 0) return 0;
  GemmArgs g{M, N, K, A, lda_m, lda_k, B, ldb_n, ldb_k, bias, alpha, relu, accumulate, C, ldc, C_bf16};
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, g);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int sp_point_pool(int B, int N, const float* pts, const float* W1, const float* b1, const uint16_t* W2,
                  const float* b2, const uint16_t* W3, const float* b3, float* value, int* index,
                  unsigned long long* scratch, sp_stream_t stream) {
  if (B < 0 || N < 0) return 1;
  if (B == 0) return 0;
  if (!pts || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !value || !index || !scratch) return 1;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * kH3;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks(n)), dim3(256), 0, st, scratch, 0ull, n);
  if (N > 0) {
    PoolArgs a{B, N, pts, W1, b1, W2, b2, W3, b3, scratch};
    hipLaunchKernelGGL(k_point_pool, dim3((N + kPts - 1) / kPts, B), dim3(256), 0, st, a);
  }
  hipLaunchKernelGGL(k_unpack_pool, dim3(blocks(n)), dim3(256), 0, st, scratch, n, value, index);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int sp_to_bf16(const float* x, uint16_t* y, long n, sp_stream_t stream) {
  if (n <= 0) return n < 0 ? 1 : 0;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, x, y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int sp_adam(float* param, uint16_t* param_bf16, const float* grad, float* m, float* v, long n, float lr,
            float beta1, float beta2, float eps, int step, sp_stream_t stream) {
  if (n <= 0 || step < 1) return n < 0 || step < 1 ? 1 : 0;
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, param, param_bf16, grad, m, v, n,
                     lr, beta1, beta2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
