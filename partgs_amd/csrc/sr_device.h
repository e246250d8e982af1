// Device-side scalar helpers shared by the surfel-rasteriser kernels.
//
// The library is compiled with -ffp-contract=off: every product and sum is
// rounded separately and evaluated in the order written, so the kernels'
// FP32 results equal those of a host evaluation of the same expressions.
// Fused multiply-adds appear only where fmaf is written explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SR_DEV __device__ __forceinline__

namespace sr {

constexpr float kNear = 0.2f;    // near_n, dsr/cuda_rasterizer/auxiliary.h:37
constexpr float kFar = 100.0f;   // far_n, auxiliary.h:38
constexpr float kFilterSize = 0.707106f;  // auxiliary.h:39 (the literal, not sqrt(1/2))

// SH constants (auxiliary.h:43-60), rounded to float.
constexpr float kSH_C0 = 0.28209479177387814f;
constexpr float kSH_C1 = 0.4886025119029199f;
constexpr float kSH_C2_0 = 1.0925484305920792f;
constexpr float kSH_C2_1 = -1.0925484305920792f;
constexpr float kSH_C2_2 = 0.31539156525252005f;
constexpr float kSH_C2_3 = -1.0925484305920792f;
constexpr float kSH_C2_4 = 0.5462742152960396f;
constexpr float kSH_C3_0 = -0.5900435899266435f;
constexpr float kSH_C3_1 = 2.890611442640554f;
constexpr float kSH_C3_2 = -0.4570457994644658f;
constexpr float kSH_C3_3 = 0.3731763325901154f;
constexpr float kSH_C3_4 = -0.4570457994644658f;
constexpr float kSH_C3_5 = 1.445305721320277f;
constexpr float kSH_C3_6 = -0.5900435899266435f;

// e^x for the blend weights.  Cody-Waite reduction x = k ln2 + r with ln2
// split into a 12-significant-bit head (k*head exact) and a tail, the
// degree-7 Taylor polynomial of e^r (|r| <= ln2/2, truncation < 8e-9
// relative) in Horner form with fmaf, then an exact scaling by 2^k.
SR_DEV float exp_pinned(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return __builtin_inff();
  const float kf = rintf(x * 1.44269502f);
  float r = fmaf(kf, -0.693115234375f, x);
  r = fmaf(kf, -3.19461833e-05f, r);
  float p = 1.98412701e-04f;
  p = fmaf(p, r, 1.38888892e-03f);
  p = fmaf(p, r, 8.33333377e-03f);
  p = fmaf(p, r, 4.16666679e-02f);
  p = fmaf(p, r, 1.66666672e-01f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

// float -> int32 truncation with NaN -> 0 and saturation.
SR_DEV int f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int)x;
}

// float -> uint32 with negative/NaN -> 0 and saturation.
SR_DEV uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967295.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

SR_DEV float fmin_(float a, float b) { return (b < a) ? b : a; }

// The k-th set bit (k < popcount) of a 64-bit mask, by halving.
SR_DEV int select_bit(uint64_t m, uint32_t k) {
  int at = 0;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
    const uint64_t low = m & ((1ull << w) - 1ull);
    const uint32_t n = (uint32_t)__popcll(low);
    if (k >= n) {
      k -= n;
      m >>= w;
      at += w;
    } else {
      m = low;
    }
  }
  return at;
}

// The value of lane 4k in lanes 4k .. 4k + 3 (a DPP quad permutation; all
// four lanes must be active).
SR_DEV float quad_bcast0(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x00, 0xF, 0xF, false));
}

SR_DEV float fmax_(float a, float b) { return (a < b) ? b : a; }

// getRect, dsr/cuda_rasterizer/auxiliary.h:67-77, with tile edge B.
SR_DEV void get_rect(float px, float py, int r, int B, int gx, int gy, int& x0, int& y0, int& x1,
                     int& y1) {
  const float rr = (float)r, BB = (float)B;
  x0 = min(gx, max(0, f2i((px - rr) / BB)));
  y0 = min(gy, max(0, f2i((py - rr) / BB)));
  x1 = min(gx, max(0, f2i((((px + rr) + BB) - 1.0f) / BB)));
  y1 = min(gy, max(0, f2i((((py + rr) + BB) - 1.0f) / BB)));
}

// Wave64 helpers.
SR_DEV float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace sr
