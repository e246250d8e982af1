// Forward kernels of the surfel rasteriser for gfx950 (CDNA4, wave64).
//
//   k_preprocess   per surfel: near test, ray-splat homography T, camera-facing
//                  normal, screen AABB and radius, SH -> RGB, and the per-tile
//                  instance histogram (one atomic per covered tile).
//   k_scan_tiles   one workgroup: exclusive scan of the tile histogram into
//                  tile start offsets, tile ranges and the instance count.
//   k_scatter      per surfel: writes a 64-bit key (depth bits << 32 | surfel
//                  id) for every covered tile into that tile's bucket.
//   k_sort_tiles   one workgroup per tile of up to 2048 instances: sorts its
//                  bucket in LDS (bitonic runs, merge-path levels), emits
//                  point_list, then splits the sorted list into
//                  per-8x8-quadrant sub-lists; larger tiles: one workgroup
//                  per 2048-key chunk, sorted the same way.
//   k_merge_chunks the large tiles' chunks merged: every entry placed in the
//                  tile list and the sub-lists by its ranks in the other
//                  chunks (binary search in LDS).
//   k_blend_fwd    one wave64 per 8 x 8 quadrant, one lane per pixel: front-
//                  to-back alpha blending of colour, depth, alpha, normal,
//                  median depth and depth distortion over the sub-list.
//
// Sorting each tile bucket by (depth bits, surfel id) gives exactly the order
// of the reference's stable radix sort over (tile, depth) keys emitted in
// surfel order (dsr/cuda_rasterizer/rasterizer_impl.cu:70-111, 301-309).
#include "sr_device.h"
#include "sr_internal.h"
#include "sr_pair.h"

namespace sr {

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

static unsigned fill_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b == 0) b = 1;
  return (unsigned)b;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

struct F3 {
  float x, y, z;
};
SR_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
SR_DEV F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
SR_DEV F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
SR_DEV F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }

// SH -> RGB, dsr/cuda_rasterizer/forward.cu:20-71 (same association); S(k)
// is coefficient k.
template <typename Get>
SR_DEV F3 color_from_sh_core(int deg, Get S, F3 dir, uint8_t* clamped) {
  F3 res = kSH_C0 * S(0);
  if (deg > 0) {
    const float x = dir.x, y = dir.y, z = dir.z;
    res = res - (kSH_C1 * y) * S(1);
    res = res + (kSH_C1 * z) * S(2);
    res = res - (kSH_C1 * x) * S(3);
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      res = res + (kSH_C2_0 * xy) * S(4);
      res = res + (kSH_C2_1 * yz) * S(5);
      res = res + (kSH_C2_2 * ((2.0f * zz - xx) - yy)) * S(6);
      res = res + (kSH_C2_3 * xz) * S(7);
      res = res + (kSH_C2_4 * (xx - yy)) * S(8);
      if (deg > 2) {
        res = res + ((kSH_C3_0 * y) * (3.0f * xx - yy)) * S(9);
        res = res + ((kSH_C3_1 * xy) * z) * S(10);
        res = res + ((kSH_C3_2 * y) * ((4.0f * zz - xx) - yy)) * S(11);
        res = res + ((kSH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy)) * S(12);
        res = res + ((kSH_C3_4 * x) * ((4.0f * zz - xx) - yy)) * S(13);
        res = res + ((kSH_C3_5 * z) * (xx - yy)) * S(14);
        res = res + ((kSH_C3_6 * x) * (xx - 3.0f * yy)) * S(15);
      }
    }
  }
  res = f3(res.x + 0.5f, res.y + 0.5f, res.z + 0.5f);
  clamped[0] = res.x < 0.0f;
  clamped[1] = res.y < 0.0f;
  clamped[2] = res.z < 0.0f;
  return f3(fmax_(res.x, 0.0f), fmax_(res.y, 0.0f), fmax_(res.z, 0.0f));
}

// Degree 3 over 16 coefficients in a 16-byte aligned row: the row is read as
// twelve float4 (a quarter of the load instructions); otherwise only the
// coefficients the degree uses are read, one float at a time.
SR_DEV F3 color_from_sh(int deg, int M, const float* sh, F3 dir, uint8_t* clamped) {
  if (deg == 3 && M == 16 && (reinterpret_cast<uintptr_t>(sh) & 15u) == 0u) {
    float v[48];
    const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      const float4 t = s4[i];
      v[4 * i] = t.x;
      v[4 * i + 1] = t.y;
      v[4 * i + 2] = t.z;
      v[4 * i + 3] = t.w;
    }
    return color_from_sh_core(deg, [&](int k) { return f3(v[3 * k], v[3 * k + 1], v[3 * k + 2]); }, dir, clamped);
  }
  return color_from_sh_core(deg, [&](int k) { return f3(sh[3 * k], sh[3 * k + 1], sh[3 * k + 2]); }, dir, clamped);
}

// Screen cull box of a surfel: every pixel where the blend can reach
// alpha >= 1/255 lies inside it.  With rho_max = 2 ln(255 opacity), a pixel
// contributes only if rho3d <= rho_max or rho2d = 2|pix - mean2D|^2 <= rho_max
// (alpha = min(0.99, opacity exp(-rho/2)), dsr/cuda_rasterizer/forward.cu:371-381).
// The box is the union of the disk rho2d <= rho_reject and the bounding box
// of the projected level set rho3d <= rho_reject (the compute_aabb conic with
// cutoff^2 = rho_reject, forward.cu:119-145), where rho_reject exceeds
// rho_max by 0.1% + 0.02, far above FP32 rounding of any pixel's rho; the
// conic's extent (whose FP32 evaluation can cancel) gets a further 0.05 pixel
// plus 1e-3 of its coordinates.  The bounds are on pixel-centre coordinates;
// pixel i has its centre at i + shift (0 here, 0.5 in the _part variant), so
// the pixel-index box is the real box moved by -shift, rounded outwards.  A
// level set reaching the camera plane (d >= 0) projects unbounded: no box
// then.
SR_DEV short4 cull_box(F3 Tu, F3 Tv, F3 Tw, float mx, float my, float opa, int W, int H, float shift,
                       float& rho_rej) {
  const short4 none = make_short4(32767, 32767, -32768, -32768);
  const short4 all = make_short4(-32768, -32768, 32767, 32767);
  if (!(opa >= 1.0f / 255.0f)) {
    rho_rej = -1.0f;
    return none;
  }
  const float rho_max = 2.0f * logf(255.0f * opa);
  rho_rej = rho_max * 1.001f + 0.02f;
  const float r2 = sqrtf(0.5f * rho_rej);
  float x0 = mx - r2, x1 = mx + r2, y0 = my - r2, y1 = my + r2;
  const float c2 = rho_rej;
  const float ww = Tw.x * Tw.x + Tw.y * Tw.y;
  const float d = c2 * ww - Tw.z * Tw.z;
  if (!(d < -1e-3f * (Tw.z * Tw.z))) return all;
  const float fx = c2 / d, fz = -1.0f / d;
  const float cx = fx * (Tu.x * Tw.x + Tu.y * Tw.y) + fz * (Tu.z * Tw.z);
  const float cy = fx * (Tv.x * Tw.x + Tv.y * Tw.y) + fz * (Tv.z * Tw.z);
  const float hx = sqrtf(fmax_(0.0f, cx * cx - (fx * (Tu.x * Tu.x + Tu.y * Tu.y) + fz * (Tu.z * Tu.z))));
  const float hy = sqrtf(fmax_(0.0f, cy * cy - (fx * (Tv.x * Tv.x + Tv.y * Tv.y) + fz * (Tv.z * Tv.z))));
  const float mgx = 0.05f + 1e-3f * (fabsf(cx) + hx), mgy = 0.05f + 1e-3f * (fabsf(cy) + hy);
  x0 = fmin_(x0, cx - hx - mgx);
  x1 = fmax_(x1, cx + hx + mgx);
  y0 = fmin_(y0, cy - hy - mgy);
  y1 = fmax_(y1, cy + hy + mgy);
  x0 = floorf(x0 - shift);
  y0 = floorf(y0 - shift);
  x1 = ceilf(x1 - shift);
  y1 = ceilf(y1 - shift);
  if (!(x0 == x0 && x1 == x1 && y0 == y0 && y1 == y1)) return all;  // NaN: no culling
  x0 = fmax_(x0, -1.0f);
  y0 = fmax_(y0, -1.0f);
  x1 = fmin_(x1, (float)W);
  y1 = fmin_(y1, (float)H);
  if (x1 < x0 || y1 < y0) return none;
  return make_short4((short)x0, (short)y0, (short)x1, (short)y1);
}

// Ray-splat homography and view-space normal of surfel idx seen from view
// v, evaluated in precision R (dsr/cuda_rasterizer/forward.cu:75-115,
// 183-207): T = [L0, 0; L1, 0; p, 1]^T world2ndc ndc2pix, rows Tu, Tv, Tw.
// With R = float the expressions and their order are the reference mode's.
template <typename R> struct Geo {
  R Tu[3], Tv[3], Tw[3], n[3];
};
SR_DEV float sqrt_r(float x) { return sqrtf(x); }
SR_DEV double sqrt_r(double x) { return ::sqrt(x); }
template <typename R>
SR_DEV void surfel_geom(const PreprocessArgs& a, int idx, const float* vm, const float* pm, int W, int H,
                        Geo<R>& g) {
  const R m0 = (R)a.means3D[3 * idx], m1 = (R)a.means3D[3 * idx + 1], m2 = (R)a.means3D[3 * idx + 2];
  if (a.tprec == nullptr) {
    const float* q = a.rots + 4 * idx;
    const R q0 = (R)q[0], q1 = (R)q[1], q2 = (R)q[2], q3 = (R)q[3];
    const R n2 = ((q3 * q3 + q0 * q0) + q1 * q1) + q2 * q2;
    const R s = (R)1 / sqrt_r(n2);
    const R w = q0 * s, x = q1 * s, y = q2 * s, z = q3 * s;
    const R c0[3] = {(R)1 - (R)2 * (y * y + z * z), (R)2 * (x * y + w * z), (R)2 * (x * z - w * y)};
    const R c1[3] = {(R)2 * (x * y - w * z), (R)1 - (R)2 * (x * x + z * z), (R)2 * (y * z + w * x)};
    const R c2[3] = {(R)2 * (x * z + w * y), (R)2 * (y * z - w * x), (R)1 - (R)2 * (x * x + y * y)};
    const R sx = (R)(a.mod * a.scales[2 * idx]), sy = (R)(a.mod * a.scales[2 * idx + 1]);
    const R L0[3] = {sx * c0[0], sx * c0[1], sx * c0[2]};
    const R L1[3] = {sy * c1[0], sy * c1[1], sy * c1[2]};
    auto rowdot = [&](const R* v, int cc) {
      return (v[0] * (R)pm[cc] + v[1] * (R)pm[4 + cc]) + v[2] * (R)pm[8 + cc];
    };
    const R a00 = rowdot(L0, 0), a01 = rowdot(L0, 1), a03 = rowdot(L0, 3);
    const R a10 = rowdot(L1, 0), a11 = rowdot(L1, 1), a13 = rowdot(L1, 3);
    const R pc[3] = {m0, m1, m2};
    const R a20 = rowdot(pc, 0) + (R)pm[12], a21 = rowdot(pc, 1) + (R)pm[13], a23 = rowdot(pc, 3) + (R)pm[15];
    const R hw = (R)W / (R)2, hw1 = (R)(W - 1) / (R)2;
    const R hh = (R)H / (R)2, hh1 = (R)(H - 1) / (R)2;
    g.Tu[0] = a00 * hw + a03 * hw1; g.Tu[1] = a10 * hw + a13 * hw1; g.Tu[2] = a20 * hw + a23 * hw1;
    g.Tv[0] = a01 * hh + a03 * hh1; g.Tv[1] = a11 * hh + a13 * hh1; g.Tv[2] = a21 * hh + a23 * hh1;
    g.Tw[0] = a03; g.Tw[1] = a13; g.Tw[2] = a23;
    for (int i = 0; i < 3; ++i)
      g.n[i] = ((R)vm[i] * c2[0] + (R)vm[4 + i] * c2[1]) + (R)vm[8 + i] * c2[2];
  } else {
    const float* t = a.tprec + 9 * idx;
    for (int i = 0; i < 3; ++i) {
      g.Tu[i] = (R)t[i];
      g.Tv[i] = (R)t[3 + i];
      g.Tw[i] = (R)t[6 + i];
    }
    g.n[0] = (R)0; g.n[1] = (R)0; g.n[2] = (R)1;
  }
}

// compute_aabb with cutoff 3 (dsr/cuda_rasterizer/forward.cu:119-145) up
// to the square roots: the denominator d, the centre p and h0 = p^2 - f T T.
template <typename R> struct Box {
  R d, px, py, h0x, h0y;
};
template <typename R> SR_DEV Box<R> aabb_terms(const Geo<R>& g) {
  Box<R> b;
  const R* Tu = g.Tu;
  const R* Tv = g.Tv;
  const R* Tw = g.Tw;
  b.d = ((R)9 * (Tw[0] * Tw[0]) + (R)9 * (Tw[1] * Tw[1])) + (R)-1 * (Tw[2] * Tw[2]);
  const R inv = (R)1 / b.d;
  const R fx = inv * (R)9, fy = inv * (R)9, fz = inv * (R)-1;
  b.px = (fx * (Tu[0] * Tw[0]) + fy * (Tu[1] * Tw[1])) + fz * (Tu[2] * Tw[2]);
  b.py = (fx * (Tv[0] * Tw[0]) + fy * (Tv[1] * Tw[1])) + fz * (Tv[2] * Tw[2]);
  b.h0x = b.px * b.px - ((fx * (Tu[0] * Tu[0]) + fy * (Tu[1] * Tu[1])) + fz * (Tu[2] * Tu[2]));
  b.h0y = b.py * b.py - ((fx * (Tv[0] * Tv[0]) + fy * (Tv[1] * Tv[1])) + fz * (Tv[2] * Tv[2]));
  return b;
}

// Radius from h0 as the reference forms it, ceil(max(sqrt(max(1e-4, h0x)),
// sqrt(max(1e-4, h0y)), 3 * 0.707106)) (forward.cu:141, 230), in FP32.
SR_DEV int radius_f32(float h0x, float h0y) {
  const float hx = sqrtf(fmax_(1e-4f, h0x));
  const float hy = sqrtf(fmax_(1e-4f, h0y));
  return f2i(ceilf(fmax_(fmax_(hx, hy), 3.0f * kFilterSize)));
}

// The same radius from FP64 h0, exactly: the least integer r >= 3 with
// r^2 >= h0 (the square roots and the ceiling decided on exact squares).
SR_DEV int radius_f64(double h0x, double h0y) {
  const double h0 = fmax(h0x, h0y);
  if (!(h0 < 1.0e18)) return h0 == h0 ? 1000000000 : 3;  // beyond every grid; NaN as the FP32 path's f2i
  int r = (int)ceil(sqrt(fmax(h0, 9.0)));
  while (r > 3 && (double)(r - 1) * (double)(r - 1) >= h0) --r;
  while ((double)r * (double)r < h0) ++r;
  return r;
}

// Near-degenerate surfels, whose FP32 preprocess can take a different
// discrete decision than exact arithmetic would (all tests in FP32 on the
// FP32 values, so the oracle reproduces the selection bit for bit):
//   grazing: |cos| <= 2^-16 (|vx| + |vy| + |vz|), the camera-facing sign
//            test (forward.cu:209-214) within the rounding error of the
//            product and of the unit normal itself, cos == 0 included;
//   conic:   |d| <= 2^-16 (9 Tw.x^2 + 9 Tw.y^2 + Tw.z^2), the AABB
//            denominator (forward.cu:125-126) near its d == 0 exit;
//   tiny:    every tangent entry of Tu, Tv below 2^-7 |Tw.z|, a footprint
//            under 1/128 pixel, where h0 is all cancellation;
//   radius:  h0x or h0y within 2^-20 of its magnitude terms of a square
//            k^2 (k >= 3) where ceil(sqrt(.)) steps (forward.cu:141, 230);
//   edge:    the tile rectangle (getRect, auxiliary.h:67-77) changes when
//            the centre moves by 2^-18 (|p| + 16) pixels.
constexpr float kDegTol = 1.0f / 65536.0f;
SR_DEV bool radius_ambiguous(float h0, float mag) {
  const float e = mag * (1.0f / 1048576.0f);
  const float lo = sqrtf(fmax_(1e-4f, h0 - e)), hi = sqrtf(fmax_(1e-4f, h0 + e));
  return ceilf(fmax_(lo, 3.0f * kFilterSize)) != ceilf(fmax_(hi, 3.0f * kFilterSize));
}
SR_DEV bool rect_ambiguous(float px, float py, int r, int B, int gx, int gy) {
  const float ex = (fabsf(px) + 16.0f) * (1.0f / 262144.0f), ey = (fabsf(py) + 16.0f) * (1.0f / 262144.0f);
  int a0, a1, a2, a3, b0, b1, b2, b3;
  get_rect(px - ex, py - ey, r, B, gx, gy, a0, a1, a2, a3);
  get_rect(px + ex, py + ey, r, B, gx, gy, b0, b1, b2, b3);
  return a0 != b0 || a1 != b1 || a2 != b2 || a3 != b3;
}
SR_DEV bool degenerate(const Geo<float>& g, float vx, float vy, float vz, float cosv, const Box<float>& b, int B,
                       int gx, int gy) {
  const float cabs = (fabsf(vx) + fabsf(vy)) + fabsf(vz);  // |n| = 1; bounds the normal's own error too
  if (!(fabsf(cosv) > kDegTol * cabs)) return true;
  const float dabs = (9.0f * (g.Tw[0] * g.Tw[0]) + 9.0f * (g.Tw[1] * g.Tw[1])) + g.Tw[2] * g.Tw[2];
  if (!(fabsf(b.d) > kDegTol * dabs)) return true;
  const float tmax = fmax_(fmax_(fabsf(g.Tu[0]), fabsf(g.Tu[1])), fmax_(fabsf(g.Tv[0]), fabsf(g.Tv[1])));
  if (tmax < (1.0f / 128.0f) * fabsf(g.Tw[2])) return true;
  const float inv = 1.0f / b.d;
  const float fx = fabsf(inv * 9.0f), fz = fabsf(inv);
  const float mx = b.px * b.px + ((fx * (g.Tu[0] * g.Tu[0]) + fx * (g.Tu[1] * g.Tu[1])) + fz * (g.Tu[2] * g.Tu[2]));
  const float my = b.py * b.py + ((fx * (g.Tv[0] * g.Tv[0]) + fx * (g.Tv[1] * g.Tv[1])) + fz * (g.Tv[2] * g.Tv[2]));
  if (radius_ambiguous(b.h0x, mx) || radius_ambiguous(b.h0y, my)) return true;
  return rect_ambiguous(b.px, b.py, radius_f32(b.h0x, b.h0y), B, gx, gy);
}

// The _part variant's homography, normal and AABB (submodules/
// diff-surfel-rasterization_part, "dsp/": cuda_rasterizer/forward.cu:75-163,
// 235-239): T built in camera space with intrinsics (fx, fy, W/2, H/2), no
// scale modifier, AABB with t = (1, 1, -1), radius ceil(3 max(extent,
// FilterSize)) with the reference's double FilterSize.  Returns false for a
// culled surfel.
constexpr double kPartFS = 0.7071067811865476;
SR_DEV bool part_geom(const PreprocessArgs& a, int idx, const float* vm, const ViewInfo& vi, float vx, float vy,
                      float vz, F3& Tu, F3& Tv, F3& Tw, F3& nrm, float& px, float& py, int& ir) {
  const float fx = (float)vi.W / (2.0f * vi.tanx), fy = (float)vi.H / (2.0f * vi.tany);
  const float cx = (float)((double)(float)vi.W / 2.0), cy = (float)((double)(float)vi.H / 2.0);
  auto Wv = [&](F3 v) {
    return f3((vm[0] * v.x + vm[4] * v.y) + vm[8] * v.z, (vm[1] * v.x + vm[5] * v.y) + vm[9] * v.z,
              (vm[2] * v.x + vm[6] * v.y) + vm[10] * v.z);
  };
  if (a.tprec == nullptr) {
    const float* q = a.rots + 4 * idx;
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const float n2 = ((q3 * q3 + q0 * q0) + q1 * q1) + q2 * q2;
    const float s = 1.0f / sqrtf(n2);
    const float w = q0 * s, x = q1 * s, y = q2 * s, z = q3 * s;
    const F3 c0 = f3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y));
    const F3 c1 = f3(2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x));
    const F3 c2 = f3(2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y));
    const float sx = a.scales[2 * idx], sy = a.scales[2 * idx + 1];
    const F3 M0 = Wv(sx * c0), M1 = Wv(sy * c1);
    F3 tn = Wv(c2);
    const float cosv = ((-tn.x * vx) + (-tn.y * vy)) + (-tn.z * vz);
    if (cosv == 0.0f) return false;  // BACKFACE_CULL
    tn = (cosv > 0.0f ? 1.0f : -1.0f) * tn;
    Tu = f3(fx * M0.x + cx * M0.z, fx * M1.x + cx * M1.z, fx * vx + cx * vz);
    Tv = f3(fy * M0.y + cy * M0.z, fy * M1.y + cy * M1.z, fy * vy + cy * vz);
    Tw = f3(M0.z, M1.z, vz);
    nrm = tn;
  } else {
    const float* t = a.tprec + 9 * idx;
    Tu = f3(t[0], t[1], t[2]);
    Tv = f3(t[3], t[4], t[5]);
    Tw = f3(t[6], t[7], t[8]);
    nrm = f3(0.0f, 0.0f, 1.0f);  // left unset by the reference
  }
  const float d = (1.0f * (Tw.x * Tw.x) + 1.0f * (Tw.y * Tw.y)) + -1.0f * (Tw.z * Tw.z);
  if (d == 0.0f) return false;
  const float inv = 1.0f / d;
  const float f0 = 1.0f * inv, f1 = 1.0f * inv, f2 = -1.0f * inv;
  px = (f0 * (Tu.x * Tw.x) + f1 * (Tu.y * Tw.y)) + f2 * (Tu.z * Tw.z);
  py = (f0 * (Tv.x * Tw.x) + f1 * (Tv.y * Tw.y)) + f2 * (Tv.z * Tw.z);
  const float h0x = px * px - ((f0 * (Tu.x * Tu.x) + f1 * (Tu.y * Tu.y)) + f2 * (Tu.z * Tu.z));
  const float h0y = py * py - ((f0 * (Tv.x * Tv.x) + f1 * (Tv.y * Tv.y)) + f2 * (Tv.z * Tv.z));
  const float hx = sqrtf(fmax_(0.0f, h0x)), hy = sqrtf(fmax_(0.0f, h0y));
  const float radius = (float)ceil(3.0 * fmax((double)fmaxf(hx, hy), kPartFS));
  ir = f2i(radius);
  return true;
}

// Forward preprocess of surfel idx in view v (record r = v * P + idx),
// dsr/cuda_rasterizer/forward.cu:75-145, 148-251.  Returns false for culled
// surfels, else its tile rectangle in the view's tile grid.  Near-degenerate
// surfels (see degenerate()) take the homography, normal, camera-facing
// test and AABB in FP64 and store them rounded to FP32.
SR_DEV bool preprocess_one(const PreprocessArgs& a, int v, int idx, size_t r, int& x0, int& y0, int& x1,
                           int& y1) {
  a.radii[r] = 0;
  a.tiles[r] = 0;
  const float m0 = a.means3D[3 * idx], m1 = a.means3D[3 * idx + 1], m2 = a.means3D[3 * idx + 2];
  const float* vm = a.vm + 16 * v;
  const float* pm = a.pm + 16 * v;
  const float* campos = a.campos + 3 * v;
  const int W = a.vs.v[v].W, H = a.vs.v[v].H, gx = a.vs.v[v].gx, gy = a.vs.v[v].gy;
  const float vx = ((vm[0] * m0 + vm[4] * m1) + vm[8] * m2) + vm[12];
  const float vy = ((vm[1] * m0 + vm[5] * m1) + vm[9] * m2) + vm[13];
  const float vz = ((vm[2] * m0 + vm[6] * m1) + vm[10] * m2) + vm[14];
  if (vz <= 0.2f) {
    if (a.prefiltered) atomicOr(&a.counters[1], kErrPrefiltered);
    return false;
  }
  F3 Tu, Tv, Tw, nrm;
  float px, py, mult;
  int ir;
  if (a.part) {
    if (!part_geom(a, idx, vm, a.vs.v[v], vx, vy, vz, Tu, Tv, Tw, nrm, px, py, ir)) return false;
    mult = 1.0f;
  } else {
  Geo<float> g;
  surfel_geom<float>(a, idx, vm, pm, W, H, g);
  const float cosv = -((vx * g.n[0] + vy * g.n[1]) + vz * g.n[2]);
  Box<float> b = aabb_terms(g);
  if (!degenerate(g, vx, vy, vz, cosv, b, a.vs.B, gx, gy)) {
    mult = cosv > 0.0f ? 1.0f : -1.0f;
    Tu = f3(g.Tu[0], g.Tu[1], g.Tu[2]);
    Tv = f3(g.Tv[0], g.Tv[1], g.Tv[2]);
    Tw = f3(g.Tw[0], g.Tw[1], g.Tw[2]);
    nrm = mult * f3(g.n[0], g.n[1], g.n[2]);
    px = b.px;
    py = b.py;
    ir = radius_f32(b.h0x, b.h0y);
  } else {
    Geo<double> G;
    surfel_geom<double>(a, idx, vm, pm, W, H, G);
    const double dvx = (((double)vm[0] * m0 + (double)vm[4] * m1) + (double)vm[8] * m2) + (double)vm[12];
    const double dvy = (((double)vm[1] * m0 + (double)vm[5] * m1) + (double)vm[9] * m2) + (double)vm[13];
    const double dvz = (((double)vm[2] * m0 + (double)vm[6] * m1) + (double)vm[10] * m2) + (double)vm[14];
    const double dcos = -((dvx * G.n[0] + dvy * G.n[1]) + dvz * G.n[2]);
    if (dcos == 0.0) return false;
    const Box<double> B = aabb_terms(G);
    if (B.d == 0.0) return false;
    mult = dcos > 0.0 ? 1.0f : -1.0f;
    Tu = f3((float)G.Tu[0], (float)G.Tu[1], (float)G.Tu[2]);
    Tv = f3((float)G.Tv[0], (float)G.Tv[1], (float)G.Tv[2]);
    Tw = f3((float)G.Tw[0], (float)G.Tw[1], (float)G.Tw[2]);
    nrm = mult * f3((float)G.n[0], (float)G.n[1], (float)G.n[2]);
    px = (float)B.px;
    py = (float)B.py;
    ir = radius_f64(B.h0x, B.h0y);
    mult *= 2.0f;  // marks the record as evaluated in FP64
  }
  }
  // the rectangle from the stored (FP32) centre, as the scatter recomputes it
  get_rect(px, py, ir, a.vs.B, gx, gy, x0, y0, x1, y1);
  const uint32_t area = (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
  if (area == 0) return false;
  F3 rgb;
  if (a.colors == nullptr) {
    const F3 dir0 = f3(m0 - campos[0], m1 - campos[1], m2 - campos[2]);
    const float len = sqrtf((dir0.x * dir0.x + dir0.y * dir0.y) + dir0.z * dir0.z);
    const F3 dir = f3(dir0.x / len, dir0.y / len, dir0.z / len);
    uint8_t cl[3];
    rgb = color_from_sh(a.D, a.M, a.shs + (size_t)idx * a.M * 3, dir, cl);
    uchar4 c4;
    c4.x = cl[0];
    c4.y = cl[1];
    c4.z = cl[2];
    c4.w = 0;
    reinterpret_cast<uchar4*>(a.clamped)[r] = c4;
  } else {
    rgb = f3(a.colors[3 * idx], a.colors[3 * idx + 1], a.colors[3 * idx + 2]);
  }
  const float opa = a.opac[idx];
  float rho_rej;
  const short4 box = cull_box(Tu, Tv, Tw, px, py, opa, W, H, a.part ? 0.5f : 0.0f, rho_rej);
  a.depths[r] = vz;
  a.radii[r] = ir;
  float4* rc = a.rec + (size_t)kRecQuads * r;
  rc[0] = make_float4(Tu.x, Tu.y, Tu.z, Tv.x);
  rc[1] = make_float4(Tv.y, Tv.z, Tw.x, Tw.y);
  rc[2] = make_float4(Tw.z, px, py, opa);
  rc[3] = make_float4(nrm.x, nrm.y, nrm.z, rgb.x);
  rc[4] = make_float4(rgb.y, rgb.z, rho_rej, mult);
  a.cullbox[r] = box;
  a.tiles[r] = area;
  return true;
}

// Tile histograms with up to kHistTiles tiles are gathered in LDS per
// workgroup and flushed with one global atomic per non-empty bin, which
// keeps hot tiles (thousands of instances) off a single global counter.
constexpr int kHistTiles = 4096;

// Exclusive scan of one value per lane over a 256-lane workgroup (four
// waves); every lane must call it.  Returns the lane's exclusive prefix and
// the workgroup total.
SR_DEV uint32_t block_scan256(uint32_t x, uint32_t* s_w, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, d);
    if (lane >= d) incl += y;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += s_w[w];
  total = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  return off + incl - x;
}

// Grid: (surfel blocks, views); a workgroup's surfels all belong to view
// blockIdx.y, so its tile histogram is that view's.  Each workgroup also
// stores its records' instance total in bsum (view-major), which the tile
// scan turns into the first instance index of every workgroup.
// Partial-gradient slots of a record: one per (instance, quadrant holding
// it), the quadrants found as the sort's sub-lists find them.
SR_DEV uint32_t record_slots(short4 bx, int x0, int y0, int x1, int y1, int B, const ViewInfo& vi) {
  const int QX = B / kQuad;
  return (uint32_t)(quad_span(bx.x, bx.z, x0 * QX, x1 * QX, vi.W) * quad_span(bx.y, bx.w, y0 * QX, y1 * QX, vi.H));
}

__global__ __launch_bounds__(256) void k_preprocess(PreprocessArgs a) {
  __shared__ uint32_t s_hist[kHistTiles];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_ssum[4];
  const int v = blockIdx.y;
  const int gx = a.vs.v[v].gx;
  const int T = gx * a.vs.v[v].gy;
  uint32_t* tile_count = a.tile_count + a.vs.v[v].tile_base;
  const bool lds = T <= kHistTiles;
  if (lds)
    for (int t = threadIdx.x; t < T; t += blockDim.x) s_hist[t] = 0;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
  const bool ok = idx < a.P && preprocess_one(a, v, idx, (size_t)v * a.P + idx, x0, y0, x1, y1);
  {
    uint32_t total = 0;
    (void)block_scan256(ok ? (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0) : 0u, s_wsum, total);
    if (threadIdx.x == 0) a.bsum[(size_t)v * gridDim.x + blockIdx.x] = total;
    // the cull box was stored by preprocess_one (this thread)
    const uint32_t ns = ok ? record_slots(a.cullbox[(size_t)v * a.P + idx], x0, y0, x1, y1, a.vs.B, a.vs.v[v]) : 0u;
    (void)block_scan256(ns, s_ssum, total);
    if (threadIdx.x == 0) a.ssum[(size_t)v * gridDim.x + blockIdx.x] = total;
  }
  if (ok) {
    for (int ty = y0; ty < y1; ++ty)
      for (int tx = x0; tx < x1; ++tx) {
        if (lds) atomicAdd(&s_hist[ty * gx + tx], 1u);
        else atomicAdd(&tile_count[ty * gx + tx], 1u);
      }
  }
  if (lds) {
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
      const uint32_t c = s_hist[t];
      if (c) atomicAdd(&tile_count[t], c);
    }
  }
}

hipError_t launch_preprocess(const PreprocessArgs& a, hipStream_t s) {
  if (a.P == 0 || a.vs.V == 0) return hipSuccess;
  hipLaunchKernelGGL(k_preprocess, dim3((a.P + 255) / 256, a.vs.V), dim3(256), 0, s, a);
  return hipGetLastError();
}


// Exclusive scan of N values per lane over a 1024-lane workgroup: each wave
// scans its 64 lanes with shuffles, one wave scans the 16 wave totals, so
// the workgroup meets at three barriers instead of two per doubling step.
// s: 32 N words of LDS.  Returns the lane's exclusive prefixes in x and the
// workgroup totals in total.
template <int N>
SR_DEV void scan1024_excl(uint32_t* s, uint32_t (&x)[N], uint32_t (&total)[N]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t incl[N];
#pragma unroll
  for (int k = 0; k < N; ++k) incl[k] = x[k];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl[k], off, 64);
      if (lane >= off) incl[k] += y;
    }
  }
  __syncthreads();  // s may still be read by a previous scan
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < N; ++k) s[k * 32 + wave] = incl[k];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      uint32_t t = lane < 16 ? s[k * 32 + lane] : 0u;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)t, off, 64);
        if (lane >= off) t += y;
      }
      if (lane < 16) s[k * 32 + 16 + lane] = t;  // inclusive over the waves
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const uint32_t base = wave == 0 ? 0u : s[k * 32 + 16 + wave - 1];
    total[k] = s[k * 32 + 31];
    x[k] = base + incl[k] - x[k];
  }
}

// Exclusive scan of the tile histogram by one 1024-lane workgroup, with the
// list of tiles of kSmallMax < n <= large_max (count in counters[3]) and the
// number of each one's first kSmallMax-key chunk (lchunk_off, followed by
// the total, which also goes to counters[6]), and the tiles above large_max
// from the end of the same list (count in counters[7]);
// then the same scan for the per-workgroup instance totals of the
// preprocess (in place), which numbers every record's instances in record
// order.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int T, const uint32_t* count,
                                                              uint32_t* start, uint32_t* fill,
                                                              uint32_t* ranges, uint32_t* counters,
                                                              uint32_t capacity, int nb, uint32_t* bsum,
                                                              uint32_t* ssum, uint32_t slot_capacity,
                                                              uint32_t* large, uint32_t* lchunk_off,
                                                              int large_max) {
  __shared__ uint32_t s_part[3 * 32];
  const int tid = threadIdx.x;
  const int per = (T + kScanThreads - 1) / kScanThreads;
  const int b = tid * per;
  const int e = min(T, b + per);
  auto chunks = [](uint32_t c) { return (c + (uint32_t)kSmallMax - 1u) / (uint32_t)kSmallMax; };
  auto is_large = [&](uint32_t c) { return c > (uint32_t)kSmallMax && c <= (uint32_t)large_max; };
  // this lane's tiles: instances, large tiles, their chunks, huge tiles
  uint32_t v1[1] = {0u}, v3[3] = {0u, 0u, 0u};
  for (int i = b; i < e; ++i) {
    const uint32_t c = count[i];
    v1[0] += c;
    if (is_large(c)) {
      v3[0] += 1u;
      v3[1] += chunks(c);
    }
    v3[2] += c > (uint32_t)large_max ? 1u : 0u;
  }
  uint32_t tot1[1], tot3[3];
  scan1024_excl(s_part, v1, tot1);
  scan1024_excl(s_part, v3, tot3);
  uint32_t run = v1[0], at = v3[0], cat = v3[1], hat = v3[2];
  for (int i = b; i < e; ++i) {
    const uint32_t c = count[i];
    start[i] = run;
    fill[i] = run;
    ranges[2 * i] = c ? run : 0u;
    ranges[2 * i + 1] = c ? run + c : 0u;
    run += c;
    if (is_large(c)) {
      large[at] = (uint32_t)i;
      lchunk_off[at] = cat;
      ++at;
      cat += chunks(c);
    }
    if (c > (uint32_t)large_max) large[T - 1 - (int)hat++] = (uint32_t)i;
  }
  if (tid == kScanThreads - 1) {
    start[T] = tot1[0];
    counters[0] = tot1[0];
    if (tot1[0] > capacity) atomicOr(&counters[1], kErrCapacity);
    counters[3] = tot3[0];
    counters[6] = tot3[1];
    counters[7] = tot3[2];
    lchunk_off[tot3[0]] = tot3[1];
  }
  // the per-workgroup instance and slot totals of the preprocess, in place
  const int pb = (nb + kScanThreads - 1) / kScanThreads;
  const int b2 = tid * pb, e2 = min(nb, b2 + pb);
  uint32_t w2[2] = {0u, 0u}, wt[2];
  for (int i = b2; i < e2; ++i) {
    w2[0] += bsum[i];
    w2[1] += ssum[i];
  }
  scan1024_excl(s_part, w2, wt);
  uint32_t run2 = w2[0], run3 = w2[1];
  for (int i = b2; i < e2; ++i) {
    const uint32_t c = bsum[i], d = ssum[i];
    bsum[i] = run2;
    ssum[i] = run3;
    run2 += c;
    run3 += d;
  }
  if (tid == 0) {
    counters[8] = wt[1];
    if (wt[1] > slot_capacity) atomicOr(&counters[1], kErrCapacity | kErrSlots);
  }
}

hipError_t launch_scan_tiles(int T, const uint32_t* count, uint32_t* start, uint32_t* fill,
                             uint32_t* ranges, uint32_t* counters, uint32_t capacity, int nb, uint32_t* bsum,
                             uint32_t* ssum, uint32_t slot_capacity, uint32_t* large, uint32_t* lchunk_off,
                             int large_max, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, T, count, start, fill, ranges,
                     counters, capacity, nb, bsum, ssum, slot_capacity, large, lchunk_off, large_max);
  return hipGetLastError();
}

// duplicateWithKeys, rasterizer_impl.cu:70-111, writing straight into tile
// buckets.  With few enough tiles, each workgroup first counts its instances
// per tile in LDS, reserves one contiguous block per tile with a single
// global atomic, then hands out slots inside the block with LDS atomics.
// Instances are numbered in record order (record r's k-th tile, row-major
// in its rectangle, is instance roff[r] + k) and the key's low word is the
// instance number, which orders equal depths by record exactly as the
// reference's (tile | depth, record) sort does; inst_rs maps it back to the
// record (and to the instance's first partial-gradient slot).
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a) {
  __shared__ uint32_t s_cnt[kHistTiles];
  __shared__ uint32_t s_base[kHistTiles];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_ssum[4];
  if (a.counters[1] & kErrCapacity) return;
  const int v = blockIdx.y;
  const int B = a.vs.B, gx = a.vs.v[v].gx, gy = a.vs.v[v].gy;
  const int T = gx * gy;
  uint32_t* tile_fill = a.tile_fill + a.vs.v[v].tile_base;
  const bool lds = T <= kHistTiles;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t r = (size_t)v * a.P + idx;
  int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
  uint64_t hi = 0;
  const bool ok = idx < a.P && a.radii[r] > 0;
  if (ok) {
    const float4 q2 = a.rec[(size_t)kRecQuads * r + 2];
    get_rect(q2.y, q2.z, a.radii[r], B, gx, gy, x0, y0, x1, y1);
    hi = (uint64_t)__float_as_uint(a.depths[r]) << 32;
  }
  uint32_t total = 0;
  const uint32_t inst0 =
      a.bsum[(size_t)v * gridDim.x + blockIdx.x] +
      block_scan256(ok ? (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0) : 0u, s_wsum, total);
  if (idx < a.P) a.roff[r] = inst0;
  // partial-gradient slots: the record's run, and each instance's first
  // slot (its quadrants' slots follow in quadrant order)
  const ViewInfo& vi = a.vs.v[v];
  const short4 bx = ok ? a.cullbox[r] : make_short4(32767, 32767, -32768, -32768);
  const uint32_t ns = ok ? record_slots(bx, x0, y0, x1, y1, B, vi) : 0u;
  const uint32_t slot0 = a.ssum[(size_t)v * gridDim.x + blockIdx.x] + block_scan256(ns, s_ssum, total);
  if (idx < a.P) a.rslot[r] = make_uint2(slot0, ns);
  if (ok) {
    const int QX = B / kQuad;
    uint32_t i = inst0, sl = slot0;
    for (int ty = y0; ty < y1; ++ty) {
      const uint32_t rows = (uint32_t)quad_span(bx.y, bx.w, ty * QX, ty * QX + QX, vi.H);
      for (int tx = x0; tx < x1; ++tx) {
        a.inst_rs[i++] = make_uint2((uint32_t)r, sl);
        sl += rows * (uint32_t)quad_span(bx.x, bx.z, tx * QX, tx * QX + QX, vi.W);
      }
    }
  }
  uint64_t* keys = a.keys;
  if (!lds) {
    if (ok) {
      uint32_t i = inst0;
      for (int ty = y0; ty < y1; ++ty)
        for (int tx = x0; tx < x1; ++tx) keys[atomicAdd(&tile_fill[ty * gx + tx], 1u)] = hi | i++;
    }
    return;
  }
  for (int t = threadIdx.x; t < T; t += blockDim.x) s_cnt[t] = 0;
  __syncthreads();
  if (ok)
    for (int ty = y0; ty < y1; ++ty)
      for (int tx = x0; tx < x1; ++tx) atomicAdd(&s_cnt[ty * gx + tx], 1u);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += blockDim.x) {
    const uint32_t c = s_cnt[t];
    if (c) s_base[t] = atomicAdd(&tile_fill[t], c);
    s_cnt[t] = 0;
  }
  __syncthreads();
  if (ok) {
    uint32_t i = inst0;
    for (int ty = y0; ty < y1; ++ty)
      for (int tx = x0; tx < x1; ++tx) {
        const int t = ty * gx + tx;
        keys[s_base[t] + atomicAdd(&s_cnt[t], 1u)] = hi | i++;
      }
  }
}

hipError_t launch_scatter(const ScatterArgs& a, hipStream_t s) {
  if (a.P == 0 || a.vs.V == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3((a.P + 255) / 256, a.vs.V), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-tile sort.  Tiles with at most kSmallMax instances are sorted by a
// 256-lane workgroup entirely in LDS (8 keys per lane); larger tiles are cut
// into kSmallMax-key chunks, each sorted the same way by its own workgroup,
// and merged by rank (k_sort_tiles, k_merge_chunks).
// ---------------------------------------------------------------------------

// LDS slot of key i in the merge sort: one pad slot after every K keys, so
// the K-strided accesses of a wave (lane t at t K + k) spread over the banks
// instead of piling onto two of them.
template <int K>
SR_DEV int lds_slot(int i) {
  return i + i / K;
}

// Ascending sort of S keys in LDS by NT lanes, K keys per lane (S a power
// of two, K <= S <= NT K; key i at lds_slot(i)).  Each lane sorts its run of
// K in registers with a bitonic network, then log2(S / K) merge-path levels
// double the sorted runs: at each level a lane finds, by binary search,
// where its K outputs start in the two runs it merges and merges them in
// registers.  A level costs a lane about log2(S) + K dependent LDS reads,
// against the S / NT compare-exchange sweeps per step of an LDS bitonic
// sort.  Keys must be unique apart from the ~0 padding.
template <int NT, int K>
SR_DEV void merge_sort_lds(uint64_t* v, int S) {
  const int tid = threadIdx.x;
  const bool act = tid * K < S;
  auto at = [&](int i) -> uint64_t { return v[lds_slot<K>(i)]; };
  uint64_t r[K];
  if (act) {
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = at(tid * K + k);
#pragma unroll
    for (int kk = 2; kk <= K; kk <<= 1) {
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const int l = i ^ j;
          if (l > i) {
            const uint64_t x = r[i], y = r[l];
            if (((i & kk) == 0) ? (x > y) : (x < y)) {
              r[i] = y;
              r[l] = x;
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) v[lds_slot<K>(tid * K + k)] = r[k];  // the lane's own slots: no hazard
  }
  __syncthreads();
  for (int w = K; w < S; w <<= 1) {
    if (act) {
      const int o = tid * K;
      const int g = o & ~(2 * w - 1);  // first key of the two runs this lane merges
      const int kq = o - g;
      // co-rank: how many of the first kq merged keys come from run A
      // (ties from B, matching the merge below)
      int lo = max(0, kq - w), hi = min(kq, w);
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (at(g + mid) < at(g + w + kq - mid - 1)) lo = mid + 1;
        else hi = mid;
      }
      int i = lo, j = kq - lo;
      uint64_t av = i < w ? at(g + i) : ~0ull, bv = j < w ? at(g + w + j) : ~0ull;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool ta = av < bv;  // ties (padding only) from B
        r[k] = ta ? av : bv;
        if (ta) {
          ++i;
          av = i < w ? at(g + i) : ~0ull;
        } else {
          ++j;
          bv = j < w ? at(g + w + j) : ~0ull;
        }
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[lds_slot<K>(tid * K + k)] = r[k];
    }
    __syncthreads();
  }
}

SR_DEV int next_pow2(int n) {
  int s = 1;
  while (s < n) s <<= 1;
  return s;
}

// Quadrant sub-lists.  Sub-list q of tile t is the tile's sorted list with
// the entries removed whose conservative cull box misses quadrant q (such
// an entry cannot reach alpha >= 1/255 anywhere in the quadrant, so blending
// the sub-list gives the reference's result); it starts at Q * beg + q * n
// and holds (surfel id, position in the tile list) pairs.  Each non-empty
// sub-list appends its 64-entry chunks to the acceptance work list.

// The hit bits of quadrants 4g..4g+3 in the four 16-bit fields of a packed
// counter (at most kSmallMax < 2^16 entries are counted in one scan, so
// sums of these never carry from one field into the next).
SR_DEV uint64_t spread4(uint32_t h, int g) {
  const uint32_t b = h >> (4 * g);
  return (uint64_t)(b & 1u) | ((uint64_t)((b >> 1) & 1u) << 16) | ((uint64_t)((b >> 2) & 1u) << 32) |
         ((uint64_t)((b >> 3) & 1u) << 48);
}
SR_DEV uint32_t field16(uint64_t v, int f) { return (uint32_t)(v >> (16 * f)) & 0xffffu; }

// Exclusive scan of one 64-bit value per lane over a 256-lane workgroup;
// total receives the sum over all lanes.
SR_DEV uint64_t scan256_excl_u64(uint64_t v, uint64_t* s_w, uint64_t& total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)incl, d);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(incl >> 32), d);
    if (lane >= d) incl += ((uint64_t)hi << 32) | lo;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  const uint64_t w0 = s_w[0], w1 = s_w[1], w2 = s_w[2], w3 = s_w[3];
  __syncthreads();
  total = w0 + w1 + w2 + w3;
  const uint64_t off = (wave > 0 ? w0 : 0ull) + (wave > 1 ? w1 : 0ull) + (wave > 2 ? w2 : 0ull);
  return off + incl - v;
}

// After an LDS sort of cn keys (bucket entries c0 .. c0 + cn - 1 of tile
// `tile`, whose bucket starts at beg and holds n), by a 256-lane workgroup:
// lane tid takes the K consecutive sorted entries tid K .. tid K + K - 1,
// looks up their records and cull boxes (all in flight at once), and one
// packed scan per four quadrants numbers every entry's hits.
// kWhole (c0 == beg, cn == n): writes point_list, each position's first
// slot and quadrant mask (pos_slot, pos_quads) and the quadrant sub-lists,
// their lengths and the acceptance work items.  Otherwise (one chunk of a
// larger tile): writes the sorted keys back in place, and per entry
// (record, quadrant mask) to lmeta, its first slot to lslot and, per
// quadrant, the number of the chunk's earlier entries that hit it to lpre,
// for k_merge_chunks.
constexpr int kSortThreads = 256;

// Called by one whole wave; lane q < Q holds the length of the tile's
// sub-list q.  Writes the lengths and clears the overflow flags, and appends
// the sub-lists' 64-entry chunks to the acceptance work list with one atomic
// for the tile (same-address atomics serialise).
SR_DEV void append_work(const SortArgs& a, int tile, int Q, uint32_t cnt) {
  const int lane = threadIdx.x & 63;
  const bool has = lane < Q;
  const uint32_t nch = has ? (cnt + 63u) / 64u : 0u;
  uint32_t incl = nch;
#pragma unroll
  for (int d = 1; d < kMaxQuads; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, d);
    if (lane >= d) incl += y;
  }
  const uint32_t all = (uint32_t)__shfl((int)incl, kMaxQuads - 1);
  uint32_t w0 = 0;
  if (lane == 0 && all > 0) w0 = atomicAdd(a.work_count, all);
  w0 = (uint32_t)__shfl((int)w0, 0);
  if (has) {
    a.sub_count[(size_t)tile * Q + lane] = cnt;
    a.q_flag[(size_t)tile * Q + lane] = 0u;
    const uint32_t first = w0 + incl - nch;
    for (uint32_t ch = 0; ch < nch; ++ch) a.work[first + ch] = make_uint2((uint32_t)(tile * Q + lane), ch);
  }
}

template <bool kWhole>
SR_DEV void place_sorted(const SortArgs& a, int tile, uint32_t beg, int n, uint32_t c0, int cn,
                         const uint64_t* s_keys) {
  constexpr int K = kSmallMax / kSortThreads;
  __shared__ uint64_t s_w[kSortThreads / 64];
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
  const int ltile = tile - (int)vi.tile_base;
  const int tx = ltile % vi.gx, ty = ltile / vi.gx;
  const int tid = threadIdx.x;
  uint32_t inst[K], rid[K], sl[K], hm[K];  // instance number: the key's low half
  // the gathers are unconditional (entries past the run read the run's
  // first entry) so each round has all K loads in flight
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid * K + k;
    inst[k] = (uint32_t)s_keys[lds_slot<K>(i < cn ? i : 0)];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint2 rs = a.inst_rs[inst[k]];
    rid[k] = rs.x;
    sl[k] = rs.y;
  }
  short4 bx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) bx[k] = a.cullbox[rid[k]];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (tid * K + k >= cn) {
      inst[k] = 0u;
      rid[k] = 0u;
      bx[k] = make_short4(32767, 32767, -32768, -32768);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid * K + k;
    if (i < cn) {
      if constexpr (kWhole) {
        a.point_list[c0 + i] = rid[k];
        a.pos_slot[c0 + i] = sl[k];
      } else {
        a.keys[c0 + i] = s_keys[lds_slot<K>(i)];
        a.lslot[c0 + i] = sl[k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid * K + k;
    hm[k] = quad_hits(bx[k], tx, ty, B, QX, Q, vi);
    if (i < cn) {
      // for the gradient slots: the instance's quadrants (a chunk's go to
      // their final positions with the merge)
      if constexpr (kWhole) a.pos_quads[c0 + i] = (uint16_t)hm[k];
      else a.lmeta[c0 + i] = make_uint2(rid[k], hm[k]);
    }
  }
  uint2* base = a.sub + (size_t)Q * beg;
  uint32_t my_cnt = 0;  // kWhole: lane q < Q gets sub-list q's length
  for (int g = 0; g * 4 < Q; ++g) {  // quadrants 4g .. 4g + 3 (uniform loop)
    uint64_t cnt = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) cnt += spread4(hm[k], g);
    uint64_t total;
    uint64_t run = scan256_excl_u64(cnt, s_w, total);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = tid * K + k;
      if (i < cn) {
        if constexpr (kWhole) {
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const int q = 4 * g + f;
            if (q < Q && ((hm[k] >> q) & 1u))
              base[(size_t)q * n + field16(run, f)] = make_uint2(rid[k], (uint32_t)i);
          }
        } else {
          uint16_t* d = a.lpre + (size_t)(c0 + i) * Q + 4 * g;
          if (Q >= 4) *reinterpret_cast<uint64_t*>(d) = run;
          else *d = (uint16_t)run;
        }
      }
      run += spread4(hm[k], g);
    }
    if constexpr (kWhole)
      if (tid >= 4 * g && tid < min(Q, 4 * g + 4)) my_cnt = field16(total, tid - 4 * g);
  }
  if constexpr (kWhole) {
    if (tid < 64) append_work(a, tile, Q, my_cnt);
  }
}

// Loads bucket entries c0 .. c0 + cn - 1 (cn <= kSmallMax) into LDS, padded
// to a power of two, and sorts them.
SR_DEV void sort_run(const SortArgs& a, uint32_t c0, int cn, uint64_t* s_keys) {
  constexpr int K = kSmallMax / kSortThreads;
  const int tid = threadIdx.x;
  const int S = max(next_pow2(cn), K);
  // all K loads in flight at once: unconditional (index clamped into the
  // run), the padding applied afterwards
  uint64_t kv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) kv[k] = a.keys[c0 + (uint32_t)min(tid + k * kSortThreads, cn - 1)];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = tid + k * kSortThreads;
    if (i < S) s_keys[lds_slot<K>(i)] = i < cn ? kv[k] : ~0ull;
  }
  __syncthreads();
  merge_sort_lds<kSortThreads, K>(s_keys, S);
}

// Tiles with more than kSmallMax instances, in two passes over the chunks
// of the scan's list of them: chunk item i (of counters[6]) is chunk
// i - lchunk_off[l] of large tile l, the last with lchunk_off[l] <= i; a
// fixed number of workgroups stride over the items.  The bucket is cut into
// chunks of kSmallMax keys.  k_sort_tiles sorts each chunk in LDS (in place in
// `keys`), looks up its records and their quadrants, and stores per entry
// (record, quadrant mask) in lmeta and, per quadrant, the number of the
// chunk's earlier entries that hit it (lpre).  k_merge_chunks then places
// every entry: its position in the tile list is its rank in its own chunk
// plus, for every other chunk, the number of that chunk's keys below it
// (binary search in LDS); its position
// in sub-list q is likewise its own lpre plus the other chunks' hits of q
// below it.  Keys are unique, so this is the merge of the sorted chunks,
// and the lists equal those of sorting the whole bucket at once.
#ifndef SR_CHUNK_BLOCKS
#define SR_CHUNK_BLOCKS 256
#endif
constexpr uint32_t kChunkBlocks = SR_CHUNK_BLOCKS;  // workgroups striding over the chunk items, per pass

// Large tile and chunk of chunk item `it` (< counters[6]).
SR_DEV void chunk_item(const SortArgs& a, uint32_t it, uint32_t& li, int& c) {
  uint32_t lo = 0, hi = a.counters[3];  // lchunk_off[lo] <= it < lchunk_off[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.lchunk_off[mid] <= it) lo = mid;
    else hi = mid;
  }
  li = lo;
  c = (int)(it - a.lchunk_off[lo]);
}

// Hits of quadrants 4g..4g+3 among the first j entries of a sorted chunk of
// len entries starting at bucket entry c0 (j <= len), packed as by spread4.
SR_DEV uint64_t chunk_hits_below(const SortArgs& a, uint32_t c0, int len, int j, int Q, int g) {
  const int at = min(j, len - 1);
  const uint16_t* p = a.lpre + (size_t)(c0 + (uint32_t)at) * Q + 4 * g;
  uint64_t v = Q >= 4 ? *reinterpret_cast<const uint64_t*>(p) : (uint64_t)*p;
  if (j == len) v += spread4(a.lmeta[c0 + (uint32_t)at].y, g);
  return v;
}

// Places the entries of chunk c of a large tile (bucket at beg, n entries),
// whose chunk-sorted keys are in s_tile.
constexpr int kMergeThreads = 1024;
SR_DEV void merge_chunk(const SortArgs& a, uint32_t beg, int n, int c, const uint64_t* s_tile) {
  constexpr int E = kSmallMax / kMergeThreads;  // entries per lane
  constexpr int kG = 8;                         // chunks searched together
  const int P = (n + kSmallMax - 1) / kSmallMax;
  const int Q = quads_per_tile(a.vs.B);
  const int tid = threadIdx.x;
  auto chunk_len = [&](int cc) { return min(kSmallMax, n - cc * kSmallMax); };
  auto key_at = [&](int idx) -> uint64_t { return s_tile[idx]; };
  const uint32_t c0 = beg + (uint32_t)c * kSmallMax;
  const int cn = chunk_len(c);
  uint64_t x[E];
  uint32_t e[E], rq[E][kMaxQuads];
  bool ok[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int i = tid + r * kMergeThreads;
    ok[r] = i < cn;
    x[r] = key_at(c * kSmallMax + (ok[r] ? i : 0));
    e[r] = (uint32_t)i;
#pragma unroll
    for (int g = 0; g < kMaxQuads / 4; ++g) {
      if (g * 4 < Q) {
        const uint64_t v = chunk_hits_below(a, c0, cn, ok[r] ? i : 0, Q, g);
#pragma unroll
        for (int f = 0; f < 4; ++f) rq[r][4 * g + f] = field16(v, f);
      }
    }
  }
  for (int cg = 0; cg < P; cg += kG) {
    // lower bounds of x in chunks cg .. cg + kG - 1, searched together;
    // every probe stays inside the tile's bucket
    int pos[E][kG];
#pragma unroll
    for (int r = 0; r < E; ++r)
#pragma unroll
      for (int j = 0; j < kG; ++j) pos[r][j] = 0;
    for (int step = kSmallMax; step > 0; step >>= 1) {
      uint64_t probe[E][kG];
#pragma unroll
      for (int r = 0; r < E; ++r)
#pragma unroll
        for (int j = 0; j < kG; ++j) {
          const int cc = cg + j;
          probe[r][j] = 0;
          if (cc < P && cc != c) {  // uniform
            const int p = min(pos[r][j] + step, chunk_len(cc));
            probe[r][j] = key_at(cc * kSmallMax + p - 1);
          }
        }
#pragma unroll
      for (int r = 0; r < E; ++r)
#pragma unroll
        for (int j = 0; j < kG; ++j) {
          const int cc = cg + j;
          if (cc < P && cc != c && pos[r][j] + step <= chunk_len(cc) && probe[r][j] < x[r]) pos[r][j] += step;
        }
    }
#pragma unroll
    for (int j = 0; j < kG; ++j) {
      const int cc = cg + j;
      if (cc >= P || cc == c) continue;
      const uint32_t cc0 = beg + (uint32_t)cc * kSmallMax;
#pragma unroll
      for (int r = 0; r < E; ++r) {
        e[r] += (uint32_t)pos[r][j];
#pragma unroll
        for (int g = 0; g < kMaxQuads / 4; ++g) {
          if (g * 4 < Q) {
            const uint64_t v = chunk_hits_below(a, cc0, chunk_len(cc), pos[r][j], Q, g);
#pragma unroll
            for (int f = 0; f < 4; ++f) rq[r][4 * g + f] += field16(v, f);
          }
        }
      }
    }
  }
  uint2* base = a.sub + (size_t)Q * beg;
#pragma unroll
  for (int r = 0; r < E; ++r) {
    if (!ok[r]) continue;
    const uint2 m = a.lmeta[c0 + tid + r * kMergeThreads];
    a.point_list[beg + e[r]] = m.x;
    a.pos_slot[beg + e[r]] = a.lslot[c0 + tid + r * kMergeThreads];
    a.pos_quads[beg + e[r]] = (uint16_t)m.y;
#pragma unroll
    for (int q = 0; q < kMaxQuads; ++q)
      if (q < Q && ((m.y >> q) & 1u)) base[(size_t)q * n + rq[r][q]] = make_uint2(m.x, e[r]);
  }
}

// Number of elements of A among the first k of merge(A, B) (keys unique).
SR_DEV int co_rank(const uint64_t* A, int na, const uint64_t* Bv, int nb, int k) {
  int lo = max(0, k - nb), hi = min(k, na);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A[mid] < Bv[k - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// A tile above the chunk path's limit (kLargeMax instances: far denser than
// a 16 x 16 tile of a typical view; in batches every tile above kSmallMax),
// by one 1024-lane workgroup: kLargeMax-key chunks sorted in
// LDS (16 keys per lane), merged pairwise in global memory (merge path; the tile's own lmeta
// entries, unused for such a tile, are the second buffer), then split into
// the quadrant sub-lists in rounds of 1024 entries (per-wave ballots and a
// scan over the waves).  O(n log n), where ranking against every chunk
// would be O(n^2 / kSmallMax).
SR_DEV void sort_huge_tile(const SortArgs& a, int tile, uint64_t* s_keys) {
  constexpr int NT = kMergeThreads, K = kLargeMax / NT, kChunk = kLargeMax, kWaves = NT / 64;
  __shared__ uint32_t s_wcnt[kMaxQuads][kWaves];
  __shared__ uint32_t s_woff[kMaxQuads][kWaves];
  __shared__ uint32_t s_total[kMaxQuads];
  const uint32_t beg = a.ranges[2 * tile];
  const int n = (int)(a.ranges[2 * tile + 1] - beg);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t* src = a.keys + beg;
  uint64_t* dst = reinterpret_cast<uint64_t*>(a.lmeta + beg);
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    const int cn = min(kChunk, n - c0);
    const int S = max(next_pow2(cn), K);
    uint64_t kv[K];  // all loads in flight together (index clamped into the chunk)
#pragma unroll
    for (int k = 0; k < K; ++k) kv[k] = src[c0 + min(tid + k * NT, cn - 1)];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = tid + k * NT;
      if (i < S) s_keys[lds_slot<K>(i)] = i < cn ? kv[k] : ~0ull;
    }
    __syncthreads();
    merge_sort_lds<NT, K>(s_keys, S);
    for (int i = tid; i < cn; i += NT) src[c0 + i] = s_keys[lds_slot<K>(i)];
  }
  __syncthreads();
  for (int w = kChunk; w < n; w <<= 1) {
    for (int a0 = 0; a0 < n; a0 += 2 * w) {
      const int na = min(w, n - a0);
      const int nb = max(0, min(w, n - a0 - w));
      const uint64_t* A = src + a0;
      const uint64_t* Bv = src + a0 + na;
      uint64_t* out = dst + a0;
      const int total = na + nb;
      const int per = (total + NT - 1) / NT;
      const int k0 = min(total, tid * per), k1 = min(total, k0 + per);
      if (k0 < k1) {
        int i = co_rank(A, na, Bv, nb, k0);
        int j = k0 - i;
        for (int k = k0; k < k1; ++k) {
          if (j >= nb || (i < na && A[i] < Bv[j])) out[k] = A[i++];
          else out[k] = Bv[j++];
        }
      }
    }
    __syncthreads();
    uint64_t* t = src;
    src = dst;
    dst = t;
  }
  for (int i = tid; i < n; i += NT) {
    const uint2 rs = a.inst_rs[(uint32_t)src[i]];
    a.point_list[beg + i] = rs.x;
    a.pos_slot[beg + i] = rs.y;
  }
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
  const int ltile = tile - (int)vi.tile_base;
  const int tx = ltile % vi.gx, ty = ltile / vi.gx;
  uint2* base = a.sub + (size_t)Q * beg;
  if (tid < kMaxQuads) s_total[tid] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the next round's record and cull box are fetched a round ahead
  const short4 kNone = make_short4(32767, 32767, -32768, -32768);
  uint32_t id_next = tid < n ? a.point_list[beg + tid] : 0u;
  short4 bx_next = tid < n ? a.cullbox[id_next] : kNone;
  for (int c = 0; c < n; c += NT) {  // uniform
    const int e = c + tid;
    const bool valid = e < n;
    const uint32_t id = id_next;
    const short4 bx = bx_next;
    if (e + NT < n) {
      id_next = a.point_list[beg + e + NT];
      bx_next = a.cullbox[id_next];
    }
    const uint32_t hm = valid ? quad_hits(bx, tx, ty, B, QX, Q, vi) : 0u;
#pragma unroll
    for (int q = 0; q < kMaxQuads; ++q) {
      if (q < Q) {
        const uint64_t m = __ballot((hm >> q) & 1u);
        if (lane == 0) s_wcnt[q][wave] = (uint32_t)__popcll(m);
      }
    }
    if (valid) a.pos_quads[beg + e] = (uint16_t)hm;
    __syncthreads();
    if (tid < Q * kWaves) {  // per quadrant, the waves' first slots
      const int q = tid / kWaves, w = tid % kWaves;
      const uint32_t c0 = s_wcnt[q][w];
      uint32_t incl = c0;
#pragma unroll
      for (int d = 1; d < kWaves; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, d, kWaves);
        if (w >= d) incl += y;
      }
      const uint32_t t0 = s_total[q];
      s_woff[q][w] = t0 + incl - c0;
      if (w == kWaves - 1) s_total[q] = t0 + incl;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxQuads; ++q) {  // the ballots again: each lane's rank within its wave
      if (q < Q) {
        const uint64_t m = __ballot((hm >> q) & 1u);
        if ((hm >> q) & 1u) base[(size_t)q * n + s_woff[q][wave] + (uint32_t)__popcll(m & lt)] = make_uint2(id, (uint32_t)e);
      }
    }
  }
  if (tid < Q) {
    const uint32_t cnt = s_total[tid];
    a.sub_count[(size_t)tile * Q + tid] = cnt;
    a.q_flag[(size_t)tile * Q + tid] = 0u;
    const uint32_t nch = (cnt + 63u) / 64u;
    if (nch > 0) {
      const uint32_t w0 = atomicAdd(a.work_count, nch);
      for (uint32_t ch = 0; ch < nch; ++ch) a.work[w0 + ch] = make_uint2((uint32_t)(tile * Q + tid), ch);
    }
  }
}

// The counters[7] huge tiles, one workgroup each (striding).
__global__ __launch_bounds__(kMergeThreads) void k_sort_huge(SortArgs a) {
  __shared__ uint64_t s_keys[kLargeMax + kLargeMax / (kLargeMax / kMergeThreads)];
  if (a.counters[1] & kErrCapacity) return;
  for (uint32_t h = blockIdx.x; h < a.counters[7]; h += gridDim.x) {
    __syncthreads();  // s_keys may still be read for the previous tile
    sort_huge_tile(a, (int)a.large[a.vs.T - 1 - h], s_keys);
  }
}

// The counters[6] chunk items of the large tiles, one workgroup each
// (striding).
__global__ __launch_bounds__(kMergeThreads) void k_merge_chunks(SortArgs a) {
  __shared__ uint64_t s_tile[kLargeMax];
  if (a.counters[1] & kErrCapacity) return;
  const int Q = quads_per_tile(a.vs.B);
  const int tid = threadIdx.x;
  for (uint32_t it = blockIdx.x; it < a.counters[6]; it += gridDim.x) {
    __syncthreads();  // s_tile may still be read for the previous item
    uint32_t li;
    int c;
    chunk_item(a, it, li, c);
    const int tile = (int)a.large[li];
    const uint32_t beg = a.ranges[2 * tile];
    const int n = (int)(a.ranges[2 * tile + 1] - beg);
    const int P = (n + kSmallMax - 1) / kSmallMax;
    for (int i = tid; i < n; i += kMergeThreads) s_tile[i] = a.keys[beg + i];
    __syncthreads();
    merge_chunk(a, beg, n, c, s_tile);
    // the tile's sub-list lengths and acceptance work items, by its chunk 0
    if (c == 0 && tid < 64) {
      uint32_t cnt = 0;
      if (tid < Q)
        for (int cc = 0; cc < P; ++cc) {
          const int len = min(kSmallMax, n - cc * kSmallMax);
          cnt += field16(chunk_hits_below(a, beg + (uint32_t)cc * kSmallMax, len, len, Q, tid / 4), tid % 4);
        }
      append_work(a, tile, Q, cnt);
    }
  }
}

// Every tile's sort in one launch of 256-lane workgroups.  The first
// kChunkBlocks workgroups sort the chunks of the large tiles (workgroup b:
// chunk items b + k kChunkBlocks), so they are dispatched first; workgroup
// kChunkBlocks + t sorts tile t when it holds at most kSmallMax instances
// (and writes empty sub-lists for an empty tile).
__global__ __launch_bounds__(kSortThreads) void k_sort_tiles(SortArgs a) {
  __shared__ uint64_t s_keys[kSmallMax + kSmallMax / (kSmallMax / kSortThreads)];
  if (a.counters[1] & kErrCapacity) return;
  if (blockIdx.x < kChunkBlocks) {
    for (uint32_t it = blockIdx.x; it < a.counters[6]; it += kChunkBlocks) {
      uint32_t li;
      int c;
      chunk_item(a, it, li, c);
      const int tile = (int)a.large[li];
      const uint32_t beg = a.ranges[2 * tile];
      const int n = (int)(a.ranges[2 * tile + 1] - beg);
      const uint32_t c0 = beg + (uint32_t)c * kSmallMax;
      const int cn = min(kSmallMax, n - c * kSmallMax);
      sort_run(a, c0, cn, s_keys);
      place_sorted<false>(a, tile, beg, n, c0, cn, s_keys);
      __syncthreads();  // s_keys is refilled for the next chunk
    }
    return;
  }
  const int tile = (int)(blockIdx.x - kChunkBlocks);
  const uint32_t beg = a.ranges[2 * tile];
  const int n = (int)(a.ranges[2 * tile + 1] - beg);
  if (n > kSmallMax) return;
  if (n <= 0) {
    const int Q = quads_per_tile(a.vs.B);
    if ((int)threadIdx.x < Q) {
      a.sub_count[(size_t)tile * Q + threadIdx.x] = 0;
      a.q_flag[(size_t)tile * Q + threadIdx.x] = 0u;
    }
    return;
  }
  sort_run(a, beg, n, s_keys);
  place_sorted<true>(a, tile, beg, n, beg, n, s_keys);
}

// A second stream per (host thread, device) with fork/join events, for the
// huge-tile sort beside the others.  Default priority; created on first use.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
static hipError_t side_stream(SideStream*& out) {
  thread_local SideStream cache[16];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
  SideStream& c = cache[dev];
  if (c.s == nullptr) {
    e = hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.join, hipEventDisableTiming);
    if (e != hipSuccess) return e;
  }
  out = &c;
  return hipSuccess;
}

hipError_t launch_sort_tiles(const SortArgs& a, hipStream_t s) {
  if (a.vs.T == 0) return hipSuccess;
  hipError_t e = hipSuccess;
  SideStream* side = nullptr;
  if (a.side_huge) {  // fork before the tile sort, so the huge-tile sort does not wait for it
    e = side_stream(side);
    if (e == hipSuccess) e = hipEventRecord(side->fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(side->s, side->fork, 0);
    if (e != hipSuccess) return e;
  }
  // issued first: its many small workgroups are dispatched before the huge
  // tiles' (each of which takes a whole CU's LDS)
  hipLaunchKernelGGL(k_sort_tiles, dim3(kChunkBlocks + a.vs.T), dim3(kSortThreads), 0, s, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const hipStream_t hs = a.side_huge ? side->s : s;  // stream of the huge-tile sort
  // batches: one workgroup per possible tile, so the dispatcher balances
  // thousands of uneven sorts; one view: a few striding workgroups
  const unsigned huge_blocks = a.side_huge ? (unsigned)a.vs.T : kChunkBlocks;
  hipLaunchKernelGGL(k_sort_huge, dim3(huge_blocks), dim3(kMergeThreads), 0, hs, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_merge_chunks, dim3(kChunkBlocks), dim3(kMergeThreads), 0, s, a);
  e = hipGetLastError();
  if (e != hipSuccess || !a.side_huge) return e;
  e = hipEventRecord(side->join, hs);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, side->join, 0);
  return e;
}

// ---------------------------------------------------------------------------
// Acceptance pass.  One wave per (quadrant, chunk of 64 sub-list entries):
// bit j of pixel p's mask is set when entry j of the chunk reaches
// alpha >= 1/255 at p.  Only pixels inside an entry's cull box can (the box
// holds every such pixel), so the wave evaluates exactly the (entry, pixel)
// pairs of the boxes clipped to the quadrant, 64 pairs per step: an LDS
// table names the entry of every pair slot, and the slot's offset in the
// entry's box gives the pixel.  The bits go to per-pixel masks in LDS; each
// accepted pair adds log2(1 - alpha) to its pixel's sum, from which the
// chunk's transmittance follows.  That transmittance is only compared
// against half the saturation threshold (a margin of a factor 2 against its
// rounding, whatever the summation order), so the order in which the lanes
// add is free.  The chunks of all quadrants are independent, so this, the
// bulk of the pair evaluations, runs balanced over the whole GPU; the blend
// kernels then visit only accepted pairs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_accept(AcceptArgs a) {
  __shared__ float4 s_q0[64], s_q1[64], s_q2[64];
  __shared__ float s_rej[64];
  __shared__ uint32_t s_start[64];  // running box area before entry j
  __shared__ uint32_t s_geo[64];    // entry j's clipped box: x0 | y0 << 4 | w << 8
  __shared__ unsigned long long s_mask[64];
  __shared__ float s_logT[64];          // per pixel: sum of log2(1 - alpha) over its accepted pairs
  __shared__ uint8_t s_owner[64 * 64];  // entry of every pair slot
  // r / w for r < 64 and 1 <= w <= 8 as (r * kDiv[w]) >> 9
  constexpr uint32_t kDiv[9] = {0u, 512u, 256u, 171u, 128u, 103u, 86u, 74u, 64u};
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const int lane = threadIdx.x;
  if (a.counters[1] & kErrCapacity) return;
  const uint32_t nwork = *a.work_count;
  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 item = a.work[nwork - 1u - w];  // from the end: see work_capacity
    const int qi = (int)item.x, c = (int)item.y;
    const int tile = qi / Q, q = qi % Q;
    const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
    const int lt = tile - (int)vi.tile_base;
    const int tx = lt % vi.gx, ty = lt / vi.gx;
    const int qx0 = tx * B + (q % QX) * kQuad, qy0 = ty * B + (q / QX) * kQuad;
    const int qw = min(kQuad, vi.W - qx0), qh = min(kQuad, vi.H - qy0);  // pixels inside the image
    const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
    const int cnt = (int)a.sub_count[qi];
    const uint2* list = a.sub + (size_t)Q * rbeg + (size_t)q * ntile;
    const int e0 = c * 64;
    const int m = min(64, cnt - e0);
    __syncthreads();  // the previous item's readers are done
    uint32_t area = 0, geo = 0;
    if (lane < m) {
      const uint32_t rid = list[e0 + lane].x;
      const float4* r = a.rec + (size_t)kRecQuads * rid;
      s_q0[lane] = r[0];
      s_q1[lane] = r[1];
      s_q2[lane] = r[2];
      s_rej[lane] = r[4].z;
      const short4 bx = a.cullbox[rid];
      const int x0 = max((int)bx.x - qx0, 0), x1 = min((int)bx.z - qx0, qw - 1);
      const int y0 = max((int)bx.y - qy0, 0), y1 = min((int)bx.w - qy0, qh - 1);
      if (x1 >= x0 && y1 >= y0) {
        const uint32_t bw = (uint32_t)(x1 - x0 + 1);
        area = bw * (uint32_t)(y1 - y0 + 1);
        geo = (uint32_t)x0 | ((uint32_t)y0 << 4) | (bw << 8);
      }
    }
    uint32_t incl = area;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, off, 64);
      if (lane >= off) incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);  // at most 64 boxes of 64 pixels
    const uint32_t first = incl - area;
    s_start[lane] = first;
    s_geo[lane] = geo;
    for (uint32_t i = 0; i < area; ++i) s_owner[first + i] = (uint8_t)lane;
    s_mask[lane] = 0ull;
    s_logT[lane] = 0.0f;
    __syncthreads();
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t t = base + (uint32_t)lane;
      if (t < total) {
        const int j = s_owner[t];
        const uint32_t g = s_geo[j];
        const uint32_t r = t - s_start[j], bw = (g >> 8) & 15u;
        const uint32_t dy = (r * kDiv[bw]) >> 9, dx = r - dy * bw;
        const int lx = (int)(g & 15u) + (int)dx, ly = (int)((g >> 4) & 15u) + (int)dy;
        Pair e;
        if (eval_pair<true>((float)(qx0 + lx), (float)(qy0 + ly), s_q0[j], s_q1[j], s_q2[j], s_rej[j], e)) {
          const int p = ly * kQuad + lx;
          atomicOr(&s_mask[p], 1ull << j);
          atomicAdd(&s_logT[p], __log2f(1.0f - e.alpha));
        }
      }
    }
    __syncthreads();
    const size_t slot = mask_base(Q, rbeg, ntile, tile, q) + (size_t)e0 + lane;
    a.masks[slot] = s_mask[lane];
    a.chunk_T[slot] = exp2f(s_logT[lane]);
  }
}

hipError_t launch_accept(const AcceptArgs& a, size_t max_items, hipStream_t s) {
  if (max_items == 0) return hipSuccess;
  const unsigned grid = (unsigned)(max_items < 65536 ? max_items : 65536);
  hipLaunchKernelGGL(k_accept, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

// Transmittance below which a pixel is certainly saturated.  The blend stops
// a pixel at its first pair with T (1 - alpha) < 1e-4, T the running product
// of (1 - alpha) over its accepted pairs in order.  The chunk transmittances
// give the same product through sums of log2(1 - alpha): with hardware
// log2 / exp2 (relative error near 2^-22 each) that is off by far less than a
// factor 2 even over thousands of pairs, so below half the threshold the
// pixel stopped in an earlier chunk.
constexpr float kSaturatedT = 0.5e-4f;

// Pair-store offsets: one wave per quadrant, one lane per pixel.  Each
// pixel gets one contiguous run of the store for its accepted pairs (in
// entry order, across every chunk), so the per-pixel blend chains read one
// stream; frag_off[chunk, pixel] records where each chunk's pairs begin.
// Chunks behind a pixel's saturation point (by the acceptance pass's chunk
// transmittances) are cut: their masks are cleared and their pairs never
// stored, as the blend never reaches them.  A quadrant whose pairs do not
// fit is flagged and left to the staged path.
__global__ __launch_bounds__(64) void k_pair_offsets(AcceptArgs a) {
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const int lane = threadIdx.x;
  if (a.counters[1] & kErrCapacity) return;
  const int qi = blockIdx.x;
  const int tile = qi / Q, q = qi % Q;
  const int cnt = (int)a.sub_count[qi];
  const int nch = (cnt + 63) / 64;
  const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
  const size_t base = mask_base(Q, rbeg, ntile, tile, q) + lane;
#ifndef SR_PO_UNROLL
#define SR_PO_UNROLL 8
#endif
  constexpr int kU = SR_PO_UNROLL;  // mask loads in flight
  uint32_t tot = 0;
  int cut = nch;  // first chunk the pixel cannot reach
  float trun = 1.0f;
  for (int c0 = 0; c0 < nch; c0 += kU) {
    uint64_t mk[kU];
    float tk[kU];
    // unconditional loads (chunk index clamped into range) so all 2 kU are
    // in flight together; the values past the end are then dropped
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const size_t cc = (size_t)min(c0 + k, nch - 1) * 64;
      mk[k] = a.masks[base + cc];
      tk[k] = a.chunk_T[base + cc];
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const bool in = c0 + k < nch;
      mk[k] = in ? mk[k] : 0ull;
      tk[k] = in ? tk[k] : 1.0f;
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      if (c0 + k < cut) {
        if (trun < kSaturatedT) {
          cut = c0 + k;
        } else {
          tot += (uint32_t)__popcll(mk[k]);
          trun *= tk[k];
        }
      }
    }
  }
  for (int c = cut; c < nch; ++c) a.masks[base + (size_t)c * 64] = 0ull;
  uint32_t incl = tot;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  unsigned long long qbase = 0;
  if (lane == 0 && total > 0) qbase = atomicAdd(a.frag_fill, (unsigned long long)total);
  qbase = __shfl(qbase, 0, 64);
  const bool fits = qbase + total <= (unsigned long long)a.frag_cap;
  // cost of the quadrant for the blend kernels' longest-first order: its
  // longest pixel chain (the staged path, when it does not fit, its list)
  uint32_t mx = tot;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
  if (lane == 0) a.q_cost[qi] = fits ? mx : (uint32_t)cnt;
  if (!fits) {
    if (lane == 0) a.q_flag[qi] = 1u;
    a.lane_range[(size_t)qi * 64 + lane] = make_uint2(0u, 0u);
    for (int c = 0; c < nch; ++c) a.frag_off[base + (size_t)c * 64] = kNoFrag;
    return;
  }
  uint32_t run = (uint32_t)qbase + (incl - tot);
  a.lane_range[(size_t)qi * 64 + lane] = make_uint2(run, tot);
  for (int c0 = 0; c0 < nch; c0 += kU) {
    uint64_t mk[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) mk[k] = a.masks[base + (size_t)min(c0 + k, nch - 1) * 64];
#pragma unroll
    for (int k = 0; k < kU; ++k) mk[k] = (c0 + k < cut) ? mk[k] : 0ull;
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      if (c0 + k < nch) a.frag_off[base + (size_t)(c0 + k) * 64] = run;
      run += (uint32_t)__popcll(mk[k]);
    }
  }
}

// Orders the quadrants by decreasing cost (64 logarithmic buckets, one
// 1024-lane workgroup), so the blend kernels dispatch the longest per-pixel
// chains first and the tail of the launch is short work.  The order within a
// bucket is arbitrary; each quadrant's result does not depend on it.
__global__ __launch_bounds__(1024) void k_order_quads(int nq, const uint32_t* cost, uint32_t* order) {
  __shared__ uint32_t s_cnt[64], s_base[64];
  const int tid = threadIdx.x;
  if (tid < 64) s_cnt[tid] = 0;
  __syncthreads();
  auto bucket = [](uint32_t c) {  // 63 for the costliest
    const int lg = c == 0u ? 0 : 32 - __clz(c);  // 1 + floor(log2 c)
    int b = lg * 2 + (lg >= 2 ? (int)((c >> (lg - 2)) & 1u) : 0);
    return 63 - min(b, 63);
  };
  // Most quadrants are empty (cost 0, the last bucket): a wave counts and
  // places those with one LDS atomic instead of one per lane.
  const int lane = tid & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i0 = 0; i0 < nq; i0 += 1024) {  // uniform
    const int i = i0 + tid;
    const uint32_t c = i < nq ? cost[i] : 1u;
    const uint64_t zero = __ballot(i < nq && c == 0u);
    if (lane == 0 && zero) atomicAdd(&s_cnt[63], (uint32_t)__popcll(zero));
    if (i < nq && c != 0u) atomicAdd(&s_cnt[bucket(c)], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int b = 0; b < 64; ++b) {
      s_base[b] = run;
      run += s_cnt[b];
    }
  }
  __syncthreads();
  for (int i0 = 0; i0 < nq; i0 += 1024) {
    const int i = i0 + tid;
    const uint32_t c = i < nq ? cost[i] : 1u;
    const uint64_t zero = __ballot(i < nq && c == 0u);
    uint32_t zbase = 0;
    if (lane == 0 && zero) zbase = atomicAdd(&s_base[63], (uint32_t)__popcll(zero));
    zbase = (uint32_t)__shfl((int)zbase, 0);
    if (i < nq) {
      if (c == 0u) order[zbase + (uint32_t)__popcll(zero & lt)] = (uint32_t)i;
      else order[atomicAdd(&s_base[bucket(c)], 1u)] = (uint32_t)i;
    }
  }
}

hipError_t launch_order_quads(int nq, const uint32_t* cost, uint32_t* order, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_order_quads, dim3(1), dim3(1024), 0, s, nq, cost, order);
  return hipGetLastError();
}

__global__ void k_iota_u32(uint32_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

hipError_t launch_iota_u32(uint32_t* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_pair_offsets(const AcceptArgs& a, hipStream_t s) {
  if (a.vs.T == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_offsets, dim3((unsigned)(a.vs.T * quads_per_tile(a.vs.B))), dim3(64), 0, s, a);
  return hipGetLastError();
}

// Pair-store fill: balanced like the acceptance pass, one wave per
// (quadrant, chunk).  The chunk's accepted pairs are dealt out 64 per step
// in pixel order (an LDS table names the pixel of every pair slot; the
// slot's rank among the pixel's pairs picks the entry from the pixel's mask
// bits), so consecutive lanes write consecutive records of a pixel's run.
// Each lane recomputes its pair and stores the entry data the blend chains
// need.
__global__ __launch_bounds__(64) void k_pair_write(AcceptArgs a) {
  __shared__ float4 s_q0[64], s_q1[64], s_q2[64], s_q3[64], s_q4[64];
  __shared__ uint32_t s_pos[64];
  __shared__ unsigned long long s_mask[64];
  __shared__ uint32_t s_start[64], s_foff[64];
  __shared__ uint8_t s_owner[64 * 64];  // pixel of every pair slot
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  const int lane = threadIdx.x;
  if (a.counters[1] & kErrCapacity) return;
  const float mcoef = kFar / (kFar - kNear);
  const float dmd_num = kFar * kNear, dmd_den0 = kFar - kNear;
  const uint32_t nwork = *a.work_count;
  for (uint32_t w = blockIdx.x; w < nwork; w += gridDim.x) {
    const uint2 item = a.work[nwork - 1u - w];  // from the end: see work_capacity
    const int qi = (int)item.x, c = (int)item.y;
    if (a.q_flag[qi] != 0u) continue;
    const int tile = qi / Q, q = qi % Q;
    const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
    const int lt = tile - (int)vi.tile_base;
    const int tx = lt % vi.gx, ty = lt / vi.gx;
    const int qx0 = tx * B + (q % QX) * kQuad, qy0 = ty * B + (q / QX) * kQuad;
    const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
    const int cnt = (int)a.sub_count[qi];
    const uint2* list = a.sub + (size_t)Q * rbeg + (size_t)q * ntile;
    const int e0 = c * 64;
    const int m = min(64, cnt - e0);
    // lane as pixel: its accepted pairs of the chunk and where they go
    const size_t slot = mask_base(Q, rbeg, ntile, tile, q) + (size_t)e0 + lane;
    const uint64_t mm = a.masks[slot];
    const uint32_t foff = mm ? a.frag_off[slot] : 0u;
    const uint32_t n = (uint32_t)__popcll(mm);
    uint32_t incl = n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, off, 64);
      if (lane >= off) incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    if (total == 0u) continue;  // uniform across the wave
    __syncthreads();  // the previous item's readers are done
    const uint32_t first = incl - n;
    s_mask[lane] = mm;
    s_start[lane] = first;
    s_foff[lane] = foff;
    for (uint32_t i = 0; i < n; ++i) s_owner[first + i] = (uint8_t)lane;
    if (lane < m) {
      const uint2 v = list[e0 + lane];
      const float4* rc = a.rec + (size_t)kRecQuads * v.x;
      s_q0[lane] = rc[0];
      s_q1[lane] = rc[1];
      s_q2[lane] = rc[2];
      s_q3[lane] = rc[3];
      s_q4[lane] = rc[4];
      s_pos[lane] = v.y;
    }
    __syncthreads();
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t t = base + (uint32_t)lane;
      if (t >= total) continue;
      const int p = s_owner[t];
      const uint32_t k = t - s_start[p];
      const int j = select_bit(s_mask[p], k);
      Pair e;
      eval_pair<false>((float)(qx0 + (p & 7)), (float)(qy0 + (p >> 3)), s_q0[j], s_q1[j], s_q2[j], 0.0f, e);
      const float4 q3 = s_q3[j], q4 = s_q4[j];
      // the distortion weight m(depth) and its derivative, as the blends form them
      const float m_d = mcoef * (1.0f - kNear / e.depth);
      const float dmd_dd = dmd_num / ((dmd_den0 * e.depth) * e.depth);
      float4* pr = a.ps.rec + kPairQuads * (size_t)(s_foff[p] + k);
      pr[0] = make_float4(e.depth, m_d, m_d * m_d, e.alpha);
      pr[1] = q3;
      pr[2] = make_float4(q4.x, q4.y, __uint_as_float(s_pos[j]), dmd_dd);
    }
  }
}

hipError_t launch_pair_write(const AcceptArgs& a, size_t max_items, hipStream_t s) {
  if (max_items == 0) return hipSuccess;
  const unsigned grid = (unsigned)(max_items < 65536 ? max_items : 65536);
  hipLaunchKernelGGL(k_pair_write, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Forward blend, dsr/cuda_rasterizer/forward.cu:256-441.  One wave64 per
// 8 x 8 quadrant; each lane walks, front to back, only the entries its
// acceptance mask marks, recomputing each pair exactly and applying the
// reference's update.  Entries a lane skips are the ones the reference
// skips, so T, the accumulators and the contributor counts (positions in the
// full tile list) come out bit-identical.
// ---------------------------------------------------------------------------
#ifndef SR_FWD_PREFETCH
#define SR_FWD_PREFETCH 20
#endif
constexpr int kPre = SR_FWD_PREFETCH;  // stored pairs a lane fetches ahead in the blend kernels

// Stored path of the forward blend: every pair of the quadrant is in the
// store and each pixel's pairs form one contiguous run in entry order.  A
// wave takes 16 pixels with 4 lanes each; the 4 lanes of a pixel walk the
// same run in step (each runs the transmittance update, so all agree on
// where the pixel stops) and split the accumulators: role 0 depth, M1, M2
// and distortion, role 1 the normal and red, role 2 green and blue, role 3
// the median and contributor counts.  Every lane executes the same
// instructions (role values are picked by select), so the serial chain of a
// long pixel costs a quarter of the full update per pair; each accumulator
// sees exactly the reference's operations in order.
SR_DEV void blend_fwd_stored(const BlendArgs& a, const ViewInfo& vi, int qi, int tx, int ty, int q, int part) {
  static_assert(kSplit == 4, "4 waves of 16 pixels x 4 roles per quadrant");
  const int B = a.vs.B, QX = B / kQuad;
  const int lane = threadIdx.x, role = lane & 3;
  const int pix = part * 16 + (lane >> 2);
  const uint32_t pxi = (uint32_t)(tx * B + (q % QX) * kQuad + (pix & 7));
  const uint32_t pyi = (uint32_t)(ty * B + (q / QX) * kQuad + (pix >> 3));
  const bool inside = pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
  const int cnt = (int)a.sub_count[qi];
  // lane_range is written by the pair-offset pass, which runs only when the
  // view has instances: read it for non-empty quadrants only
  const uint2 rg = (inside && cnt > 0) ? a.lane_range[(size_t)qi * 64 + pix] : make_uint2(0u, 0u);
  // Each lane loads one record quad per pair: role 0 quad 0 (depth, mm,
  // mm^2, alpha), role 1 quad 1 (normal, red), roles 2 and 3 quad 2 (green,
  // blue, tile position); alpha and depth reach the other lanes of the pixel
  // from role 0 by a quad broadcast.  A lane accumulates x, y and z of its
  // quad in a0..a2 and w in a3, except that role 0's a3 is the distortion.
  const int qsel = role < 2 ? role : 2;
  float T = 1.0f;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, median_depth = 0.0f;
  float median_contributor = -1.0f;
  uint32_t last_contributor = 0;
  int blended = 0;
  bool done = !inside;
  uint32_t f = rg.x;
  int left = (int)rg.y;
  // slots past the run's end keep defined (stale) values; nothing uses them
  float4 fq[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) fq[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  auto fetch = [&](uint32_t at, int n) {
#pragma unroll
    for (int k = 0; k < kPre; ++k)
      if (k < n) fq[k] = a.ps.rec[kPairQuads * (size_t)(at + k) + qsel];
  };
  fetch(f, left);
  while (left > 0 && !done) {  // the four lanes of a pixel share left and done
    float4 cq[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) cq[k] = fq[k];
    const int n = left;
    f += kPre;
    left -= kPre;
    if (left > 0) fetch(f, left);  // next batch in flight
    // Straight-line over the batch: every update is computed and kept only
    // where the pair is live (in the run, the pixel not done) and does not
    // saturate the pixel, so the body has no per-pair branch.
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
      // forward.cu:383-415; role 0's quad: (depth, mm, mm^2, alpha), mm = mcoef (1 - near / depth)
      const float4 v = cq[k];
      const float alpha = quad_bcast0(v.w), depth = quad_bcast0(v.x);
      const float test_T = T * (1.0f - alpha);
      const bool live = k < n && !done;
      const bool sat = test_T < 0.0001f;
      done = done || (live && sat);
      const bool ok = live && !sat;
      const uint32_t contributor = __float_as_uint(v.z) + 1u;  // quad 2's z: position in the tile list
      const float w = alpha * T;
      const float A = 1.0f - T;
      // role 0: the distortion term (mm^2 A + M2 - 2 mm M1), a1 = M1, a2 = M2
      const float x3 = role == 0 ? (v.z * A + a2) - (2.0f * v.y) * a1 : v.w;
      const float n0 = a0 + v.x * w, n1 = a1 + v.y * w, n2 = a2 + v.z * w, n3 = a3 + x3 * w;
      a0 = ok ? n0 : a0;
      a1 = ok ? n1 : a1;
      a2 = ok ? n2 : a2;
      a3 = ok ? n3 : a3;
      const bool med = ok && T > 0.5f;
      median_depth = med ? depth : median_depth;
      median_contributor = med ? (float)contributor : median_contributor;
      T = ok ? test_T : T;
      last_contributor = ok ? contributor : last_contributor;
      blended += ok ? 1 : 0;
    }
  }
  if (!inside) return;
  const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi, pb = vi.pix_base;
  float* final_T = a.final_T + 3 * pb;
  float* out_color = a.out_color + 3 * pb;
  float* out_others = a.out_others + 7 * pb;
  if (role == 0) {
    out_others[pid] = a0;
    final_T[pid + HW] = a1;
    final_T[pid + 2 * HW] = a2;
    out_others[pid + 6 * HW] = a3;  // the distortion
  } else if (role == 1) {
    out_others[pid + 2 * HW] = a0;
    out_others[pid + 3 * HW] = a1;
    out_others[pid + 4 * HW] = a2;
    out_color[pid] = a3 + T * a.bg[0];
  } else if (role == 2) {
    out_color[HW + pid] = a0 + T * a.bg[1];
    out_color[2 * HW + pid] = a1 + T * a.bg[2];
  } else {
    uint32_t* n_contrib = a.n_contrib + 2 * pb;
    final_T[pid] = T;
    n_contrib[pid] = last_contributor;
    n_contrib[pid + HW] = f2u(median_contributor);
    out_others[pid + HW] = 1.0f - T;
    out_others[pid + 5 * HW] = median_depth;
    // the backward passes walk only the blended prefix of the run
    if (cnt > 0) a.lane_range[(size_t)qi * 64 + pix].y = (uint32_t)blended;
  }
}

__global__ __launch_bounds__(64) void k_blend_fwd(BlendArgs a) {
  __shared__ float4 s_q0[64], s_q1[64], s_q2[64], s_q3[64], s_q4[64];
  __shared__ uint32_t s_pos[64];
  const int B = a.vs.B, QX = B / kQuad, Q = QX * QX;
  if (a.counters[1] & kErrCapacity) return;
  const int qi = (int)a.q_order[blockIdx.x / kSplit], part = blockIdx.x % kSplit;
  const int tile = qi / Q, q = qi % Q;
  const ViewInfo& vi = a.vs.v[view_of_tile(a.vs, (uint32_t)tile)];
  const int lt = tile - (int)vi.tile_base;
  const int tx = lt % vi.gx, ty = lt / vi.gx;
  const int lane = threadIdx.x;
  const bool staged = a.q_flag[qi] != 0u;
  if (staged && part != 0) return;  // the staged path uses all 64 lanes of part 0
  if (!staged) {
    blend_fwd_stored(a, vi, qi, tx, ty, q, part);
    return;
  }
  // staged path: lane = pixel of the quadrant
  const int pix = lane;
  const bool lane_ok = true;
  const uint32_t pxi = (uint32_t)(tx * B + (q % QX) * kQuad + (pix & 7));
  const uint32_t pyi = (uint32_t)(ty * B + (q / QX) * kQuad + (pix >> 3));
  const bool inside = lane_ok && pxi < (uint32_t)vi.W && pyi < (uint32_t)vi.H;
  bool done = !inside;
  const uint32_t rbeg = a.ranges[2 * tile], ntile = a.ranges[2 * tile + 1] - rbeg;
  const int cnt = (int)a.sub_count[qi];
  const uint2* list = a.sub + (size_t)Q * rbeg + (size_t)q * ntile;
  const uint64_t* masks = a.masks + mask_base(Q, rbeg, ntile, tile, q);
  const uint32_t* frag_off = a.frag_off + mask_base(Q, rbeg, ntile, tile, q);
  const float pixx = (float)pxi, pixy = (float)pyi;
  float T = 1.0f;
  uint32_t last_contributor = 0;
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, N0 = 0.0f, N1 = 0.0f, N2 = 0.0f;
  float D = 0.0f, M1 = 0.0f, M2 = 0.0f, distortion = 0.0f, median_depth = 0.0f, median_contributor = -1.0f;
  int blended = 0;
  const float mcoef = kFar / (kFar - kNear);
  const int nch = (cnt + 63) / 64;
  // One front-to-back step of the reference's update (forward.cu:383-415);
  // false once T would drop below 1e-4 (the pixel is done).  mm is the
  // distortion weight mcoef (1 - near / depth).
  auto step = [&](float alpha, float depth, float mm, uint32_t pos, const float4& nc, float g, float bl) -> bool {
    const float test_T = T * (1.0f - alpha);
    if (test_T < 0.0001f) return false;
    const uint32_t contributor = pos + 1u;
    const float w = alpha * T;
    const float A = 1.0f - T;
    distortion += (((mm * mm) * A + M2) - (2.0f * mm) * M1) * w;
    D += depth * w;
    M1 += mm * w;
    M2 += (mm * mm) * w;
    if (T > 0.5f) {
      median_depth = depth;
      median_contributor = (float)contributor;
    }
    N0 += nc.x * w;
    N1 += nc.y * w;
    N2 += nc.z * w;
    C0 += nc.w * w;
    C1 += g * w;
    C2 += bl * w;
    T = test_T;
    last_contributor = contributor;
    ++blended;
    return true;
  };
  {
    // Some chunk missed the store: stage the records chunk by chunk and
    // recompute the pairs of chunks without stored pairs.
    ChunkRegs nx{};
    uint2 v2 = make_uint2(0u, 0u);
    if (nch > 0) {
      nx = chunk_load(masks, frag_off, a.rec, lane, 0, chunk_entry(list, cnt, lane, 0, nch));
      v2 = chunk_entry(list, cnt, lane, 1, nch);
    }
    for (int c = 0; c < nch; ++c) {
      const int m = min(64, cnt - c * 64);
      uint64_t mask = done ? 0ull : nx.mask;
      const uint32_t foff = mask ? nx.foff : kNoFrag;
      __syncthreads();
      if (lane < m) {
        s_pos[lane] = nx.v.y;
        s_q0[lane] = nx.q0;
        s_q1[lane] = nx.q1;
        s_q2[lane] = nx.q2;
        s_q3[lane] = nx.q3;
        s_q4[lane] = nx.q4;
      }
      __syncthreads();
      if (c + 1 < nch) {  // in flight while this chunk blends
        const uint2 ve = v2;
        v2 = chunk_entry(list, cnt, lane, c + 2, nch);
        nx = chunk_load(masks, frag_off, a.rec, lane, c + 1, ve);
      }
      if (__ballot(mask != 0ull) == 0ull) continue;
      uint32_t f = foff;
      while (mask != 0ull) {
        const int j = __builtin_ctzll(mask);
        mask &= mask - 1ull;
        float alpha, depth;
        if (foff != kNoFrag) {
          const float4 p0 = a.ps.rec[kPairQuads * (size_t)(f++)];
          alpha = p0.w;
          depth = p0.x;
        } else {
          Pair e;
          if (!eval_pair<false>(pixx, pixy, s_q0[j], s_q1[j], s_q2[j], 0.0f, e)) continue;
          alpha = e.alpha;
          depth = e.depth;
        }
        const float4 q3 = s_q3[j], q4 = s_q4[j];
        if (!step(alpha, depth, mcoef * (1.0f - kNear / depth), s_pos[j], q3, q4.x, q4.y)) {
          done = true;
          break;
        }
      }
      if (__ballot(!done) == 0ull) break;
    }
  }
  if (inside) {
    const size_t HW = (size_t)vi.H * vi.W, pid = (size_t)vi.W * pyi + pxi, pb = vi.pix_base;
    float* final_T = a.final_T + 3 * pb;
    uint32_t* n_contrib = a.n_contrib + 2 * pb;
    float* out_color = a.out_color + 3 * pb;
    float* out_others = a.out_others + 7 * pb;
    final_T[pid] = T;
    n_contrib[pid] = last_contributor;
    out_color[pid] = C0 + T * a.bg[0];
    out_color[HW + pid] = C1 + T * a.bg[1];
    out_color[2 * HW + pid] = C2 + T * a.bg[2];
    n_contrib[pid + HW] = f2u(median_contributor);
    final_T[pid + HW] = M1;
    final_T[pid + 2 * HW] = M2;
    out_others[pid] = D;
    out_others[pid + HW] = 1.0f - T;
    out_others[pid + 2 * HW] = N0;
    out_others[pid + 3 * HW] = N1;
    out_others[pid + 4 * HW] = N2;
    out_others[pid + 5 * HW] = median_depth;
    out_others[pid + 6 * HW] = distortion;
    // the backward cuts the pixel's run into the same segments as a stored run
    a.lane_range[(size_t)qi * 64 + pix].y = (uint32_t)blended;
  }
}

hipError_t launch_blend_fwd(const BlendArgs& a, hipStream_t s) {
  if (a.vs.T == 0) return hipSuccess;
  if (a.vs.B != 8 && a.vs.B != 16 && a.vs.B != 32) return hipErrorInvalidValue;
  const unsigned nblk = (unsigned)(a.vs.T * quads_per_tile(a.vs.B));
  hipLaunchKernelGGL(k_blend_fwd, dim3(nblk * kSplit), dim3(64), 0, s, a);
  return hipGetLastError();
}

// markVisible / checkFrustum, rasterizer_impl.cu:54-66.
__global__ void k_mark_visible(int P, const float* means3D, const float* vm, uint8_t* present) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float m0 = means3D[3 * i], m1 = means3D[3 * i + 1], m2 = means3D[3 * i + 2];
  const float vz = ((vm[2] * m0 + vm[6] * m1) + vm[10] * m2) + vm[14];
  present[i] = vz <= 0.2f ? 0 : 1;
}

hipError_t launch_mark_visible(int P, const float* means3D, const float* vm, uint8_t* present,
                               hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_visible, dim3((P + 255) / 256), dim3(256), 0, s, P, means3D, vm, present);
  return hipGetLastError();
}

// Copies the instance count and the error flags to a caller-visible status
// word, so batched callers can check capacity without reading internal state.
__global__ void k_copy_status(const uint32_t* counters, uint32_t* status) {
  if (threadIdx.x < 9) status[threadIdx.x] = counters[threadIdx.x];
}

hipError_t launch_copy_status(const uint32_t* counters, uint32_t* status, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_status, dim3(1), dim3(64), 0, s, counters, status);
  return hipGetLastError();
}

}  // namespace sr
