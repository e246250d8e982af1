// CPU oracle for the 2DGS surfel rasteriser.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the host CPU, the
// reference's per-surfel and per-pixel algorithm so that the HIP library can
// be checked against it.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product path never does.
//
// Each stage cites the reference file:line it follows (paths relative to the
// reference tree, submodules/diff-surfel-rasterization/ abbreviated "dsr/").
//
// Two precisions from one template:
//   * R = float  : "reference mode".  FP32 with the evaluation order pinned
//     (built with -ffp-contract=off, left-to-right sums, no fused ops except
//     the explicit fmaf of sro_expf).  The HIP kernels evaluate the same
//     expressions in the same order, so integer outputs (radii, tile counts,
//     keys, point_list, ranges) and the forward images agree bit for bit.
//     Backward per-pair terms are FP32; their sums are accumulated in FP64.
//   * R = double : "truth mode" for FP64 central differences.
//
// How each stage is pinned (tests/test_oracle_*.py):
//   expf            vs FP64 libm exp, max error in ulps, whole domain used.
//   preprocess      closed forms: fronto-parallel surfel centre/extent,
//                   geometric ray/plane intersection through T, normal =
//                   R e_z seen from the camera; SH basis orthonormality and
//                   golden vectors of the reference's own eval_sh.
//   binning         brute force: numpy lexsort over (tile, depth bits,
//                   emission order); prefix sums; range scan.
//   render fwd      single-surfel closed form alpha*c + (1-alpha)*bg with
//                   alpha from a geometric intersection; two-surfel
//                   distortion w1 w2 (m1-m2)^2; invariants sum(w) = 1 - T.
//   render bwd,     FP64 central differences of the FP64 forward (loss =
//   preprocess bwd  <g, outputs>) for means3D, scales, rotations (tangent
//                   part, see quat note), opacities, shs, colors, transMat.
//   dL_dmean2D      the densification statistic written at dsr/cuda_rasterizer/
//                   backward.cu:626-629 is not a gradient of anything: parity
//                   unpinned (formula restated only).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Scene description for the oracle: host arrays, float32 or float64 as
// selected by sro_forward's f64 flag (NULL for absent optional inputs).
typedef struct sro_scene_c {
  int P, D, M, W, H, tile, prefiltered;
  const void *means3D, *shs, *colors_precomp, *opacities, *scales, *rotations, *transMat_precomp;
  const void *viewmatrix, *projmatrix, *campos, *bg;
  double scale_modifier, tan_fovx, tan_fovy;
} sro_scene_c;

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// exp: Cody-Waite range reduction x = k ln2 + r, |r| <= ln2/2, then the
// degree-7 Taylor polynomial of e^r in Horner form with fused multiply-adds
// and an exact scaling by 2^k.  ln2 is split as hi (12 significant bits, so
// k*hi is exact for |k| < 2^12) + lo.  Coefficients are 1/n! rounded to
// float.  Truncation error of the series on |r| <= 0.347 is < 8e-9 relative.
// The reference calls exp() on a float (dsr/cuda_rasterizer/forward.cu:379,
// backward.cu:311); this routine fixes the bits on both CPU and GPU.
// ---------------------------------------------------------------------------
inline float sro_expf_impl(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return INFINITY;
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

template <typename R> inline R rexp(R x);
template <> inline float rexp<float>(float x) { return sro_expf_impl(x); }
template <> inline double rexp<double>(double x) { return std::exp(x); }

// Float -> int conversion with the GPU's semantics (NaN -> 0, saturate).
template <typename R> inline int f2i(R x) {
  if (x != x) return 0;
  if (x >= (R)2147483647.0) return INT_MAX;
  if (x <= (R)-2147483648.0) return INT_MIN;
  return (int)x;
}
// Float -> uint32 with saturation (dsr/cuda_rasterizer/forward.cu:430 stores
// a float median_contributor, -1 when unset, into a uint32 buffer).
template <typename R> inline uint32_t f2u(R x) {
  if (!(x > (R)0)) return 0u;
  if (x >= (R)4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

template <typename R> inline R rmin(R a, R b) { return (b < a) ? b : a; }
template <typename R> inline R rmax(R a, R b) { return (a < b) ? b : a; }

// Real spherical-harmonics constants, dsr/cuda_rasterizer/auxiliary.h:43-60.
const double kSH_C0 = 0.28209479177387814;
const double kSH_C1 = 0.4886025119029199;
const double kSH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                          -1.0925484305920792, 0.5462742152960396};
const double kSH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                          0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                          -0.5900435899266435};

template <typename R> struct V3 {
  R x, y, z;
};
template <typename R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Scene {
  int P, D, M, W, H, B, prefiltered;
  bool has_shs, has_colors, has_sr, has_tprec;
  double mod, tanx, tany;
  bool part = false;  // the diff-surfel-rasterization_part variant
  int S = 0;          // its semantic channels (<= 16)
};

template <typename R> struct Ctx {
  Scene sc{};
  int gx = 0, gy = 0, T = 0, N = 0;
  std::vector<R> means3D, shs, colors, opac, scales, rots, tprec, vm, pm, campos, bg;
  std::vector<R> sem, out_sem, g_sem;  // _part: semantics (P x S), image (S x H x W), gradient
  double* acc_sem = nullptr;
  // geometry state (dsr/cuda_rasterizer/rasterizer_impl.h:29-44)
  std::vector<int> radii;
  std::vector<R> depths, means2D, transMat, normal_opacity, rgb;
  std::vector<uint8_t> clamped;
  std::vector<int8_t> flip;  // camera-facing sign of the normal, x2 when the preprocess ran in FP64
  std::vector<uint32_t> tiles_touched, point_offsets;
  int prefilter_violation = 0;
  // binning state
  int L = 0;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> point_list, ranges;
  // image state + outputs
  std::vector<R> out_color, out_others, final_T;
  std::vector<uint32_t> n_contrib;
  // backward: render-stage sums in FP64, then final grads in R
  // backward accumulators in FP64: one copy per OpenMP thread in acc_pool
  // (copy t at t * acc_n), the pointers naming copy 0, into which the other
  // copies are summed in thread order after the parallel pass
  double *acc_transMat = nullptr, *acc_mean2D = nullptr, *acc_normal = nullptr, *acc_opacity = nullptr,
         *acc_colors = nullptr;
  std::vector<double> acc_pool;
  size_t acc_n = 0;
  std::vector<R> g_mean2D, g_colors, g_opacity, g_means3D, g_transMat, g_sh, g_scales, g_rots, g_normal;
  // Frozen decisions (finite-difference support): a recorded forward keeps
  // every discrete choice it made; a replay re-evaluates the same
  // expressions on new inputs with those choices fixed, which makes the
  // outputs a smooth function of the inputs whose gradient is the one the
  // backward computes (it differentiates with the same choices held).
  bool record = false, replay = false;
  std::vector<int8_t> dec_mult;                  // normal flip sign per surfel
  std::vector<std::vector<uint32_t>> dec_pix;    // per pixel: (list position << 3) | flags
  int dec_alpha_clamped = 0;                     // blended pairs whose alpha hit the 0.99 cap
};

template <typename R> void copy_in(std::vector<R>& dst, const void* src, size_t n, int f64) {
  dst.assign(n, (R)0);
  if (!src || n == 0) return;
  if (f64) {
    const double* s = (const double*)src;
    for (size_t i = 0; i < n; ++i) dst[i] = (R)s[i];
  } else {
    const float* s = (const float*)src;
    for (size_t i = 0; i < n; ++i) dst[i] = (R)s[i];
  }
}

// ---------------------------------------------------------------------------
// SH -> RGB, dsr/cuda_rasterizer/forward.cu:20-71.
// ---------------------------------------------------------------------------
template <typename R> V3<R> color_from_sh(const Ctx<R>& c, int idx, uint8_t* clamped) {
  const int deg = c.sc.D, M = c.sc.M;
  const R* m = &c.means3D[3 * idx];
  V3<R> dir = v3(m[0] - c.campos[0], m[1] - c.campos[1], m[2] - c.campos[2]);
  const R len = std::sqrt((dir.x * dir.x + dir.y * dir.y) + dir.z * dir.z);
  dir = v3(dir.x / len, dir.y / len, dir.z / len);
  const R* sh = &c.shs[(size_t)idx * M * 3];
  auto S = [&](int k) { return v3(sh[3 * k], sh[3 * k + 1], sh[3 * k + 2]); };
  const R C0 = (R)(float)kSH_C0, C1 = (R)(float)kSH_C1;
  R C2[5], C3[7];
  for (int i = 0; i < 5; ++i) C2[i] = (R)(float)kSH_C2[i];
  for (int i = 0; i < 7; ++i) C3[i] = (R)(float)kSH_C3[i];
  V3<R> res = C0 * S(0);
  if (deg > 0) {
    const R x = dir.x, y = dir.y, z = dir.z;
    res = res - (C1 * y) * S(1);
    res = res + (C1 * z) * S(2);
    res = res - (C1 * x) * S(3);
    if (deg > 1) {
      const R xx = x * x, yy = y * y, zz = z * z;
      const R xy = x * y, yz = y * z, xz = x * z;
      res = res + (C2[0] * xy) * S(4);
      res = res + (C2[1] * yz) * S(5);
      res = res + (C2[2] * (((R)2 * zz - xx) - yy)) * S(6);
      res = res + (C2[3] * xz) * S(7);
      res = res + (C2[4] * (xx - yy)) * S(8);
      if (deg > 2) {
        res = res + ((C3[0] * y) * ((R)3 * xx - yy)) * S(9);
        res = res + ((C3[1] * xy) * z) * S(10);
        res = res + ((C3[2] * y) * (((R)4 * zz - xx) - yy)) * S(11);
        res = res + ((C3[3] * z) * (((R)2 * zz - (R)3 * xx) - (R)3 * yy)) * S(12);
        res = res + ((C3[4] * x) * (((R)4 * zz - xx) - yy)) * S(13);
        res = res + ((C3[5] * z) * (xx - yy)) * S(14);
        res = res + ((C3[6] * x) * (xx - (R)3 * yy)) * S(15);
      }
    }
  }
  res = v3(res.x + (R)0.5f, res.y + (R)0.5f, res.z + (R)0.5f);
  if (c.replay)  // clamping choices held from the recorded forward
    return v3(clamped[0] ? (R)0 : res.x, clamped[1] ? (R)0 : res.y, clamped[2] ? (R)0 : res.z);
  clamped[0] = res.x < (R)0;
  clamped[1] = res.y < (R)0;
  clamped[2] = res.z < (R)0;
  return v3(rmax(res.x, (R)0), rmax(res.y, (R)0), rmax(res.z, (R)0));
}

// Rotation matrix columns of a (w,x,y,z) quaternion, dsr/cuda_rasterizer/auxiliary.h:213-235.
// The reference normalises with rsqrtf; here 1/sqrt (correctly rounded).
template <typename R> void quat_cols(const R* q, R& w, R& x, R& y, R& z, V3<R> col[3]) {
  const R n2 = ((q[3] * q[3] + q[0] * q[0]) + q[1] * q[1]) + q[2] * q[2];
  const R s = (R)1 / std::sqrt(n2);
  w = q[0] * s;
  x = q[1] * s;
  y = q[2] * s;
  z = q[3] * s;
  col[0] = v3((R)1 - (R)2 * (y * y + z * z), (R)2 * (x * y + w * z), (R)2 * (x * z - w * y));
  col[1] = v3((R)2 * (x * y - w * z), (R)1 - (R)2 * (x * x + z * z), (R)2 * (y * z + w * x));
  col[2] = v3((R)2 * (x * z + w * y), (R)2 * (y * z - w * x), (R)1 - (R)2 * (x * x + y * y));
}

// getRect, dsr/cuda_rasterizer/auxiliary.h:67-77, with tile size B.
template <typename R>
void get_rect(R px, R py, int r, int B, int gx, int gy, int& x0, int& y0, int& x1, int& y1) {
  const R rr = (R)r, BB = (R)B;
  x0 = std::min(gx, std::max(0, f2i((px - rr) / BB)));
  y0 = std::min(gy, std::max(0, f2i((py - rr) / BB)));
  x1 = std::min(gx, std::max(0, f2i((((px + rr) + BB) - (R)1) / BB)));
  y1 = std::min(gy, std::max(0, f2i((((py + rr) + BB) - (R)1) / BB)));
}

// ---------------------------------------------------------------------------
// Forward preprocess, dsr/cuda_rasterizer/forward.cu:75-145 (compute_transmat,
// compute_aabb) and :148-251 (preprocessCUDA), in_frustum auxiliary.h:185-210.
// ---------------------------------------------------------------------------
// Homography rows Tu, Tv, Tw and view-space normal of surfel idx, evaluated
// in precision G from the stored inputs (forward.cu:75-115, 183-207).
template <typename G> struct Geo {
  V3<G> Tu, Tv, Tw, n;
};
template <typename G, typename R> Geo<G> surfel_geom(const Ctx<R>& c, int idx) {
  const Scene& s = c.sc;
  Geo<G> g;
  const R* m = &c.means3D[3 * idx];
  if (!s.has_tprec) {
    const R* qr = &c.rots[4 * idx];
    const G qq[4] = {(G)qr[0], (G)qr[1], (G)qr[2], (G)qr[3]};
    G w, x, y, z;
    V3<G> col[3];
    quat_cols(qq, w, x, y, z, col);
    const G sx = (G)((R)s.mod * c.scales[2 * idx]), sy = (G)((R)s.mod * c.scales[2 * idx + 1]);
    const V3<G> L0 = sx * col[0], L1 = sy * col[1], L2 = col[2];
    const R* pm = c.pm.data();
    const R* vm = c.vm.data();
    // A(i,c) = sum_r M(r,i) F(r,c), F(r,c) = pm[4r+c]; only c = 0, 1, 3 are needed.
    auto rowdot = [&](const V3<G>& a, int cc) {
      return (a.x * (G)pm[cc] + a.y * (G)pm[4 + cc]) + a.z * (G)pm[8 + cc];
    };
    const G a00 = rowdot(L0, 0), a01 = rowdot(L0, 1), a03 = rowdot(L0, 3);
    const G a10 = rowdot(L1, 0), a11 = rowdot(L1, 1), a13 = rowdot(L1, 3);
    const V3<G> pc = v3((G)m[0], (G)m[1], (G)m[2]);
    const G a20 = rowdot(pc, 0) + (G)pm[12], a21 = rowdot(pc, 1) + (G)pm[13], a23 = rowdot(pc, 3) + (G)pm[15];
    const G hw = (G)s.W / (G)2, hw1 = (G)(s.W - 1) / (G)2;
    const G hh = (G)s.H / (G)2, hh1 = (G)(s.H - 1) / (G)2;
    g.Tu = v3(a00 * hw + a03 * hw1, a10 * hw + a13 * hw1, a20 * hw + a23 * hw1);
    g.Tv = v3(a01 * hh + a03 * hh1, a11 * hh + a13 * hh1, a21 * hh + a23 * hh1);
    g.Tw = v3(a03, a13, a23);
    g.n = v3(((G)vm[0] * L2.x + (G)vm[4] * L2.y) + (G)vm[8] * L2.z,
             ((G)vm[1] * L2.x + (G)vm[5] * L2.y) + (G)vm[9] * L2.z,
             ((G)vm[2] * L2.x + (G)vm[6] * L2.y) + (G)vm[10] * L2.z);
  } else {
    const R* t = &c.tprec[9 * idx];
    g.Tu = v3((G)t[0], (G)t[1], (G)t[2]);
    g.Tv = v3((G)t[3], (G)t[4], (G)t[5]);
    g.Tw = v3((G)t[6], (G)t[7], (G)t[8]);
    g.n = v3((G)0, (G)0, (G)1);
  }
  return g;
}

// compute_aabb with cutoff 3 (forward.cu:119-145) up to the square roots.
template <typename G> struct Box {
  G d, px, py, h0x, h0y;
};
template <typename G> Box<G> aabb_terms(const Geo<G>& g) {
  const V3<G> Tu = g.Tu, Tv = g.Tv, Tw = g.Tw;
  Box<G> b;
  b.d = ((G)9 * (Tw.x * Tw.x) + (G)9 * (Tw.y * Tw.y)) + (G)-1 * (Tw.z * Tw.z);
  const G inv = (G)1 / b.d;
  const G fx = inv * (G)9, fy = inv * (G)9, fz = inv * (G)-1;
  b.px = (fx * (Tu.x * Tw.x) + fy * (Tu.y * Tw.y)) + fz * (Tu.z * Tw.z);
  b.py = (fx * (Tv.x * Tw.x) + fy * (Tv.y * Tw.y)) + fz * (Tv.z * Tw.z);
  b.h0x = b.px * b.px - ((fx * (Tu.x * Tu.x) + fy * (Tu.y * Tu.y)) + fz * (Tu.z * Tu.z));
  b.h0y = b.py * b.py - ((fx * (Tv.x * Tv.x) + fy * (Tv.y * Tv.y)) + fz * (Tv.z * Tv.z));
  return b;
}

// ceil(max(sqrt(max(1e-4, h0x)), sqrt(max(1e-4, h0y)), 3 * 0.707106))
// (forward.cu:141, 230) in precision R.
template <typename R> int radius_of(R h0x, R h0y) {
  const R hx = std::sqrt(rmax((R)1e-4f, h0x));
  const R hy = std::sqrt(rmax((R)1e-4f, h0y));
  return f2i(std::ceil(rmax(rmax(hx, hy), (R)(3.0f * 0.707106f))));
}

// The radius from FP64 h0 exactly: the least integer r >= 3 with r^2 >= h0.
int radius_exact(double h0x, double h0y) {
  const double h0 = std::max(h0x, h0y);
  if (!(h0 < 1.0e18)) return h0 == h0 ? 1000000000 : 3;
  int r = (int)std::ceil(std::sqrt(std::max(h0, 9.0)));
  while (r > 3 && (double)(r - 1) * (double)(r - 1) >= h0) --r;
  while ((double)r * (double)r < h0) ++r;
  return r;
}

// Near-degenerate surfels take their homography, normal, camera-facing test
// and AABB in FP64 (selection on the FP32 values, as the HIP kernel makes
// it; thresholds and reasons in partgs_amd/csrc/sr_forward.hip):
// grazing |cos| <= 2^-16 sum|v_i|; conic |d| <= 2^-16 (9 Tw.x^2 +
// 9 Tw.y^2 + Tw.z^2); tiny, every tangent entry of Tu, Tv below 2^-7 |Tw.z|;
// radius, h0 within 2^-20 of its magnitude terms of a ceil step; edge, the
// tile rectangle changes when the centre moves by 2^-18 (|p| + 16) pixels.
bool g_fp64_degenerate = true;
const float kDegTol = 1.0f / 65536.0f;
bool radius_ambiguous(float h0, float mag) {
  const float e = mag * (1.0f / 1048576.0f);
  const float lo = std::sqrt(rmax(1e-4f, h0 - e)), hi = std::sqrt(rmax(1e-4f, h0 + e));
  return std::ceil(rmax(lo, 3.0f * 0.707106f)) != std::ceil(rmax(hi, 3.0f * 0.707106f));
}
bool rect_ambiguous(float px, float py, int r, int B, int gx, int gy) {
  const float ex = (std::fabs(px) + 16.0f) * (1.0f / 262144.0f), ey = (std::fabs(py) + 16.0f) * (1.0f / 262144.0f);
  int a0, a1, a2, a3, b0, b1, b2, b3;
  get_rect(px - ex, py - ey, r, B, gx, gy, a0, a1, a2, a3);
  get_rect(px + ex, py + ey, r, B, gx, gy, b0, b1, b2, b3);
  return a0 != b0 || a1 != b1 || a2 != b2 || a3 != b3;
}
bool degenerate(const Geo<float>& g, float vx, float vy, float vz, float cosv, const Box<float>& b, int B, int gx,
                int gy) {
  const float cabs = (std::fabs(vx) + std::fabs(vy)) + std::fabs(vz);  // |n| = 1
  if (!(std::fabs(cosv) > kDegTol * cabs)) return true;
  const float dabs = (9.0f * (g.Tw.x * g.Tw.x) + 9.0f * (g.Tw.y * g.Tw.y)) + g.Tw.z * g.Tw.z;
  if (!(std::fabs(b.d) > kDegTol * dabs)) return true;
  const float tmax = rmax(rmax(std::fabs(g.Tu.x), std::fabs(g.Tu.y)), rmax(std::fabs(g.Tv.x), std::fabs(g.Tv.y)));
  if (tmax < (1.0f / 128.0f) * std::fabs(g.Tw.z)) return true;
  const float inv = 1.0f / b.d;
  const float fx = std::fabs(inv * 9.0f), fz = std::fabs(inv);
  const float mx = b.px * b.px + ((fx * (g.Tu.x * g.Tu.x) + fx * (g.Tu.y * g.Tu.y)) + fz * (g.Tu.z * g.Tu.z));
  const float my = b.py * b.py + ((fx * (g.Tv.x * g.Tv.x) + fx * (g.Tv.y * g.Tv.y)) + fz * (g.Tv.z * g.Tv.z));
  if (radius_ambiguous(b.h0x, mx) || radius_ambiguous(b.h0y, my)) return true;
  return rect_ambiguous(b.px, b.py, radius_of(b.h0x, b.h0y), B, gx, gy);
}

template <typename R> void preprocess_one(Ctx<R>& c, int idx) {
  const Scene& s = c.sc;
  if (c.replay) {
    if (c.radii[idx] <= 0) return;  // culled in the recorded forward
  } else {
    c.radii[idx] = 0;
    c.tiles_touched[idx] = 0;
  }
  const R* m = &c.means3D[3 * idx];
  const R* vm = c.vm.data();
  const R vx = ((vm[0] * m[0] + vm[4] * m[1]) + vm[8] * m[2]) + vm[12];
  const R vy = ((vm[1] * m[0] + vm[5] * m[1]) + vm[9] * m[2]) + vm[13];
  const R vz = ((vm[2] * m[0] + vm[6] * m[1]) + vm[10] * m[2]) + vm[14];
  if (vz <= (R)0.2f && !c.replay) {
    if (s.prefiltered) c.prefilter_violation = 1;
    return;
  }
  const Geo<R> g = surfel_geom<R>(c, idx);
  // dual-visible normal flip, forward.cu:209-214
  const R cosv = -((vx * g.n.x + vy * g.n.y) + vz * g.n.z);
  const Box<R> b = aabb_terms(g);
  V3<R> Tu = g.Tu, Tv = g.Tv, Tw = g.Tw, nrm = g.n;
  R px = b.px, py = b.py, mult;
  int ir;
  bool fp64 = false;
  if constexpr (std::is_same<R, float>::value)
    fp64 = g_fp64_degenerate && degenerate(g, vx, vy, vz, cosv, b, s.B, c.gx, c.gy);
  if (!fp64) {
    if (cosv == (R)0 && !c.replay) return;
    if (b.d == (R)0 && !c.replay) return;
    mult = cosv > (R)0 ? (R)1 : (R)-1;
    ir = radius_of(b.h0x, b.h0y);
  } else {
    const Geo<double> G = surfel_geom<double>(c, idx);
    const double dvx = (((double)vm[0] * m[0] + (double)vm[4] * m[1]) + (double)vm[8] * m[2]) + (double)vm[12];
    const double dvy = (((double)vm[1] * m[0] + (double)vm[5] * m[1]) + (double)vm[9] * m[2]) + (double)vm[13];
    const double dvz = (((double)vm[2] * m[0] + (double)vm[6] * m[1]) + (double)vm[10] * m[2]) + (double)vm[14];
    const double dcos = -((dvx * G.n.x + dvy * G.n.y) + dvz * G.n.z);
    if (dcos == 0.0 && !c.replay) return;
    const Box<double> B = aabb_terms(G);
    if (B.d == 0.0 && !c.replay) return;
    mult = dcos > 0.0 ? (R)1 : (R)-1;
    Tu = v3((R)G.Tu.x, (R)G.Tu.y, (R)G.Tu.z);
    Tv = v3((R)G.Tv.x, (R)G.Tv.y, (R)G.Tv.z);
    Tw = v3((R)G.Tw.x, (R)G.Tw.y, (R)G.Tw.z);
    nrm = v3((R)G.n.x, (R)G.n.y, (R)G.n.z);
    px = (R)B.px;
    py = (R)B.py;
    ir = radius_exact(B.h0x, B.h0y);
  }
  if (c.replay) mult = (R)c.dec_mult[idx];
  if (c.record) c.dec_mult[idx] = (int8_t)(mult > (R)0 ? 1 : -1);
  c.flip[idx] = (int8_t)((mult > (R)0 ? 1 : -1) * (fp64 ? 2 : 1));
  nrm = mult * nrm;
  if (c.replay) ir = c.radii[idx];
  int x0, y0, x1, y1;
  get_rect(px, py, ir, s.B, c.gx, c.gy, x0, y0, x1, y1);  // from the stored centre, as the binning does
  const uint32_t area = c.replay ? c.tiles_touched[idx] : (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
  if (area == 0) return;
  if (!s.has_colors) {
    const V3<R> rgb = color_from_sh(c, idx, &c.clamped[3 * idx]);
    c.rgb[3 * idx] = rgb.x;
    c.rgb[3 * idx + 1] = rgb.y;
    c.rgb[3 * idx + 2] = rgb.z;
  } else {
    for (int k = 0; k < 3; ++k) c.rgb[3 * idx + k] = c.colors[3 * idx + k];
  }
  c.depths[idx] = vz;
  c.radii[idx] = ir;
  c.means2D[2 * idx] = px;
  c.means2D[2 * idx + 1] = py;
  R* T9 = &c.transMat[9 * idx];
  T9[0] = Tu.x; T9[1] = Tu.y; T9[2] = Tu.z;
  T9[3] = Tv.x; T9[4] = Tv.y; T9[5] = Tv.z;
  T9[6] = Tw.x; T9[7] = Tw.y; T9[8] = Tw.z;
  R* no = &c.normal_opacity[4 * idx];
  no[0] = nrm.x; no[1] = nrm.y; no[2] = nrm.z; no[3] = c.opac[idx];
  c.tiles_touched[idx] = area;
}

// getHigherMsb, dsr/cuda_rasterizer/rasterizer_impl.cu:35-50.
uint32_t higher_msb(uint32_t n) {
  uint32_t msb = sizeof(n) * 4;
  uint32_t step = msb;
  while (step > 1) {
    step /= 2;
    if (n >> msb) msb += step; else msb -= step;
  }
  if (n >> msb) msb++;
  return msb;
}

template <typename R> uint32_t depth_bits(R d) {
  const float f = (float)d;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// ---------------------------------------------------------------------------
// Binning: inclusive scan (rasterizer_impl.cu:278), duplicateWithKeys
// (:70-111), stable radix sort over bits [0, 32+bit) (:301-309), memset +
// identifyTileRanges (:311-318, :116-138).
// ---------------------------------------------------------------------------
template <typename R> void binning(Ctx<R>& c) {
  const int P = c.sc.P;
  c.point_offsets.assign(P, 0);
  uint64_t acc = 0;
  for (int i = 0; i < P; ++i) {
    acc += c.tiles_touched[i];
    c.point_offsets[i] = (uint32_t)acc;
  }
  c.L = P ? (int)c.point_offsets[P - 1] : 0;
  std::vector<uint64_t> ku(c.L);
  std::vector<uint32_t> vu(c.L);
  for (int idx = 0; idx < P; ++idx) {
    if (c.radii[idx] <= 0) continue;
    uint32_t off = idx == 0 ? 0 : c.point_offsets[idx - 1];
    int x0, y0, x1, y1;
    get_rect(c.means2D[2 * idx], c.means2D[2 * idx + 1], c.radii[idx], c.sc.B, c.gx, c.gy, x0, y0, x1, y1);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        uint64_t key = (uint64_t)(uint32_t)(y * c.gx + x);
        key <<= 32;
        key |= depth_bits(c.depths[idx]);
        ku[off] = key;
        vu[off] = (uint32_t)idx;
        ++off;
      }
  }
  const uint32_t bit = higher_msb((uint32_t)(c.gx * c.gy));
  const uint64_t mask = (32 + bit >= 64) ? ~0ull : ((1ull << (32 + bit)) - 1ull);
  // the stable sort by (tile, depth bits): a stable counting pass by tile,
  // then each tile's run stably by depth bits, tiles in parallel
  std::vector<uint32_t> order(c.L), tstart((size_t)c.T + 1, 0u);
  for (int i = 0; i < c.L; ++i) ++tstart[(size_t)((ku[i] & mask) >> 32) + 1];
  for (int t = 0; t < c.T; ++t) tstart[t + 1] += tstart[t];
  {
    std::vector<uint32_t> cur(tstart.begin(), tstart.end() - 1);
    for (int i = 0; i < c.L; ++i) order[cur[(size_t)((ku[i] & mask) >> 32)]++] = (uint32_t)i;
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < c.T; ++t)
    std::stable_sort(order.begin() + tstart[t], order.begin() + tstart[t + 1],
                     [&](uint32_t a, uint32_t b) { return (ku[a] & mask) < (ku[b] & mask); });
  c.keys.resize(c.L);
  c.point_list.resize(c.L);
  for (int i = 0; i < c.L; ++i) {
    c.keys[i] = ku[order[i]];
    c.point_list[i] = vu[order[i]];
  }
  c.ranges.assign(2 * (size_t)c.T, 0u);
  for (int i = 0; i < c.L; ++i) {
    const uint32_t cur = (uint32_t)(c.keys[i] >> 32);
    if (i == 0) c.ranges[2 * cur] = 0;
    else {
      const uint32_t prev = (uint32_t)(c.keys[i - 1] >> 32);
      if (cur != prev) {
        c.ranges[2 * prev + 1] = (uint32_t)i;
        c.ranges[2 * cur] = (uint32_t)i;
      }
    }
    if (i == c.L - 1) c.ranges[2 * cur + 1] = (uint32_t)c.L;
  }
}

// ---------------------------------------------------------------------------
// Per-pixel forward blend, dsr/cuda_rasterizer/forward.cu:256-441.
// ---------------------------------------------------------------------------
template <typename R> void render_pixel(Ctx<R>& c, int tile, int pxi, int pyi) {
  const int W = c.sc.W, H = c.sc.H;
  const uint32_t rs = c.ranges[2 * tile], re = c.ranges[2 * tile + 1];
  const R pixx = (R)pxi, pixy = (R)pyi;
  R T = (R)1;
  uint32_t contributor = 0, last_contributor = 0;
  R C[3] = {0, 0, 0}, N[3] = {0, 0, 0};
  R D = 0, M1 = 0, M2 = 0, distortion = 0, median_depth = 0, median_contributor = (R)-1;
  const R near_n = (R)0.2f, far_n = (R)100.0f;
  const R mcoef = far_n / (far_n - near_n);
  for (uint32_t e = rs; e < re; ++e) {
    contributor++;
    const int id = (int)c.point_list[e];
    const R* T9 = &c.transMat[9 * id];
    const V3<R> Tu = v3(T9[0], T9[1], T9[2]), Tv = v3(T9[3], T9[4], T9[5]), Tw = v3(T9[6], T9[7], T9[8]);
    const V3<R> k = pixx * Tw - Tu;
    const V3<R> l = pixy * Tw - Tv;
    const V3<R> p = cross(k, l);
    if (p.z == (R)0) continue;
    const R sx = p.x / p.z, sy = p.y / p.z;
    const R rho3d = sx * sx + sy * sy;
    const R dx = c.means2D[2 * id] - pixx, dy = c.means2D[2 * id + 1] - pixy;
    const R rho2d = (R)2 * (dx * dx + dy * dy);
    const R rho = rmin(rho3d, rho2d);
    const R depth = (rho3d <= rho2d) ? (sx * Tw.x + sy * Tw.y) + Tw.z : Tw.z;
    if (depth < near_n) continue;
    const R* no = &c.normal_opacity[4 * id];
    const R power = (R)-0.5f * rho;
    if (power > (R)0) continue;
    const R alpha = rmin((R)0.99f, no[3] * rexp<R>(power));
    if (alpha < (R)(1.0f / 255.0f)) continue;
    const R test_T = T * ((R)1 - alpha);
    if (test_T < (R)0.0001f) break;
    const R w = alpha * T;
    const R A = (R)1 - T;
    const R mm = mcoef * ((R)1 - near_n / depth);
    distortion += (((mm * mm) * A + M2) - ((R)2 * mm) * M1) * w;
    D += depth * w;
    M1 += mm * w;
    M2 += (mm * mm) * w;
    if (T > (R)0.5) {
      median_depth = depth;
      median_contributor = (R)contributor;
    }
    for (int ch = 0; ch < 3; ++ch) N[ch] += no[ch] * w;
    for (int ch = 0; ch < 3; ++ch) C[ch] += c.rgb[3 * id + ch] * w;
    if (c.record) {
      const uint32_t fl = (rho3d <= rho2d ? 1u : 0u) | (alpha == (R)0.99f ? 2u : 0u) | (T > (R)0.5 ? 4u : 0u);
      c.dec_pix[(size_t)W * pyi + pxi].push_back((e << 3) | fl);
    }
    T = test_T;
    last_contributor = contributor;
  }
  const size_t HW = (size_t)H * W, pid = (size_t)W * pyi + pxi;
  c.final_T[pid] = T;
  c.n_contrib[pid] = last_contributor;
  for (int ch = 0; ch < 3; ++ch) c.out_color[ch * HW + pid] = C[ch] + T * c.bg[ch];
  c.n_contrib[pid + HW] = f2u(median_contributor);
  c.final_T[pid + HW] = M1;
  c.final_T[pid + 2 * HW] = M2;
  c.out_others[pid] = D;
  c.out_others[pid + HW] = (R)1 - T;
  for (int ch = 0; ch < 3; ++ch) c.out_others[pid + (2 + ch) * HW] = N[ch];
  c.out_others[pid + 5 * HW] = median_depth;
  c.out_others[pid + 6 * HW] = distortion;
}

// Replay of one pixel: the recorded blended entries with their branch
// (3D intersection or screen-space filter), alpha cap and median choices
// held; the expressions are those of render_pixel.
template <typename R> void render_pixel_replay(Ctx<R>& c, int pxi, int pyi) {
  const int W = c.sc.W, H = c.sc.H;
  const R pixx = (R)pxi, pixy = (R)pyi;
  R T = (R)1;
  R C[3] = {0, 0, 0}, N[3] = {0, 0, 0};
  R D = 0, M1 = 0, M2 = 0, distortion = 0, median_depth = 0;
  const R near_n = (R)0.2f, far_n = (R)100.0f;
  const R mcoef = far_n / (far_n - near_n);
  const size_t HW = (size_t)H * W, pid = (size_t)W * pyi + pxi;
  for (const uint32_t word : c.dec_pix[pid]) {
    const uint32_t e = word >> 3, fl = word & 7u;
    const int id = (int)c.point_list[e];
    const R* T9 = &c.transMat[9 * id];
    const V3<R> Tu = v3(T9[0], T9[1], T9[2]), Tv = v3(T9[3], T9[4], T9[5]), Tw = v3(T9[6], T9[7], T9[8]);
    const V3<R> k = pixx * Tw - Tu;
    const V3<R> l = pixy * Tw - Tv;
    const V3<R> p = cross(k, l);
    const R sx = p.x / p.z, sy = p.y / p.z;
    const R rho3d = sx * sx + sy * sy;
    const R dx = c.means2D[2 * id] - pixx, dy = c.means2D[2 * id + 1] - pixy;
    const R rho2d = (R)2 * (dx * dx + dy * dy);
    const R rho = (fl & 1u) ? rho3d : rho2d;
    const R depth = (fl & 1u) ? (sx * Tw.x + sy * Tw.y) + Tw.z : Tw.z;
    const R* no = &c.normal_opacity[4 * id];
    const R alpha = (fl & 2u) ? (R)0.99f : no[3] * rexp<R>((R)-0.5f * rho);
    const R w = alpha * T;
    const R A = (R)1 - T;
    const R mm = mcoef * ((R)1 - near_n / depth);
    distortion += (((mm * mm) * A + M2) - ((R)2 * mm) * M1) * w;
    D += depth * w;
    M1 += mm * w;
    M2 += (mm * mm) * w;
    if (fl & 4u) median_depth = depth;
    for (int ch = 0; ch < 3; ++ch) N[ch] += no[ch] * w;
    for (int ch = 0; ch < 3; ++ch) C[ch] += c.rgb[3 * id + ch] * w;
    T = T * ((R)1 - alpha);
  }
  c.final_T[pid] = T;
  for (int ch = 0; ch < 3; ++ch) c.out_color[ch * HW + pid] = C[ch] + T * c.bg[ch];
  c.final_T[pid + HW] = M1;
  c.final_T[pid + 2 * HW] = M2;
  c.out_others[pid] = D;
  c.out_others[pid + HW] = (R)1 - T;
  for (int ch = 0; ch < 3; ++ch) c.out_others[pid + (2 + ch) * HW] = N[ch];
  c.out_others[pid + 5 * HW] = median_depth;
  c.out_others[pid + 6 * HW] = distortion;
}

// Replays a recorded forward on the inputs now loaded (same surfel count,
// image and tile size): preprocess with the recorded culls, normal flips,
// radii and colour clamps, the recorded tile lists, and the recorded blends.
template <typename R> void run_replay(Ctx<R>& c) {
  const Scene& s = c.sc;
  c.replay = true;
  c.record = false;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.P; ++i) preprocess_one(c, i);
#pragma omp parallel for schedule(dynamic, 16)
  for (int y = 0; y < s.H; ++y)
    for (int x = 0; x < s.W; ++x) render_pixel_replay(c, x, y);
  c.replay = false;
}


// ===========================================================================
// The _part variant (submodules/diff-surfel-rasterization_part, "dsp/"):
// what renderer/gaussian_renderer_2d/__init__.py:86-96 calls.  Its math
// differs from the plain rasteriser's: the homography is built in camera
// space with intrinsics (fx, fy, W/2, H/2) (dsp/cuda_rasterizer/forward.cu:
// 75-128), the AABB uses t = (1, 1, -1) and radius ceil(3 max(extent,
// FilterSize)) (:133-163, 235-239), pixel centres sit at +0.5 (:292), up to
// 16 semantic channels blend beside the colour (:440, 460) and an eighth
// aux channel holds the median weight (:471).  Several constants of the
// reference are double literals (FilterSize, FilterInvSquare, NEAR_PLANE,
// FAR_PLANE; dsp/cuda_rasterizer/auxiliary.h:19-36), so the expressions
// that use them are evaluated in double and rounded, as C++ promotion does.
// ===========================================================================
// Adds v to the calling thread's copy of accumulator element p (p points
// into copy 0; copy t sits t * acc_n doubles further).
template <typename Cx> inline void acc_add(Cx& c, double* p, double v) {
  p[(size_t)omp_get_thread_num() * c.acc_n] += v;
}

const double kP_FS = 0.7071067811865476;  // FilterSize
const double kP_NEAR = 0.2, kP_FAR = 100.0;

// 1/(FilterSize*FilterSize) * x, the macro's evaluation order.
template <typename R> inline R part_fis_mul(R x) { return (R)((1.0 / (kP_FS * kP_FS)) * (double)x); }

template <typename R> void preprocess_one_part(Ctx<R>& c, int idx) {
  const Scene& s = c.sc;
  c.radii[idx] = 0;
  c.tiles_touched[idx] = 0;
  const R* m = &c.means3D[3 * idx];
  const R* vm = c.vm.data();
  const R vx = ((vm[0] * m[0] + vm[4] * m[1]) + vm[8] * m[2]) + vm[12];
  const R vy = ((vm[1] * m[0] + vm[5] * m[1]) + vm[9] * m[2]) + vm[13];
  const R vz = ((vm[2] * m[0] + vm[6] * m[1]) + vm[10] * m[2]) + vm[14];
  if (vz <= (R)0.2f) {  // in_frustum, dsp/cuda_rasterizer/auxiliary.h:180-191
    if (s.prefiltered) c.prefilter_violation = 1;
    return;
  }
  // focal = W / (2 tan(fov/2)) (rasterizer_impl.cu:223-224); intrins.zw =
  // float(W)/2.0, float(H)/2.0 (dsp forward.cu:213)
  const R fx = (R)(float)s.W / ((R)2.0f * (R)(float)s.tanx);
  const R fy = (R)(float)s.H / ((R)2.0f * (R)(float)s.tany);
  const R cx = (R)((double)(float)s.W / 2.0), cy = (R)((double)(float)s.H / 2.0);
  auto Wv = [&](const V3<R>& v) {
    return v3((vm[0] * v.x + vm[4] * v.y) + vm[8] * v.z, (vm[1] * v.x + vm[5] * v.y) + vm[9] * v.z,
              (vm[2] * v.x + vm[6] * v.y) + vm[10] * v.z);
  };
  V3<R> Tu, Tv, Tw, nrm;
  if (!s.has_tprec) {
    // computeTransMat, dsp forward.cu:75-128 (scale_to_mat(.., 1.0f): the
    // scale modifier is not applied)
    R w, x, y, z;
    V3<R> col[3];
    quat_cols(&c.rots[4 * idx], w, x, y, z, col);
    const R sx = c.scales[2 * idx], sy = c.scales[2 * idx + 1];
    const V3<R> M0 = Wv(sx * col[0]), M1 = Wv(sy * col[1]);
    const V3<R> pv = v3(vx, vy, vz);
    V3<R> tn = Wv(col[2]);
    const R cosv = ((-tn.x * pv.x) + (-tn.y * pv.y)) + (-tn.z * pv.z);
    if (cosv == (R)0) return;  // BACKFACE_CULL
    const R mult = cosv > (R)0 ? (R)1 : (R)-1;  // DUAL_VISIABLE
    tn = mult * tn;
    Tu = v3(fx * M0.x + cx * M0.z, fx * M1.x + cx * M1.z, fx * pv.x + cx * pv.z);
    Tv = v3(fy * M0.y + cy * M0.z, fy * M1.y + cy * M1.z, fy * pv.y + cy * pv.z);
    Tw = v3(M0.z, M1.z, pv.z);
    nrm = tn;
  } else {
    // the reference leaves the normal unset here; (0, 0, 1) as the plain
    // rasteriser's precomputed path
    const R* t = &c.tprec[9 * idx];
    Tu = v3(t[0], t[1], t[2]);
    Tv = v3(t[3], t[4], t[5]);
    Tw = v3(t[6], t[7], t[8]);
    nrm = v3((R)0, (R)0, (R)1);
  }
  // computeAABB, dsp forward.cu:133-163
  const R d = ((R)1 * (Tw.x * Tw.x) + (R)1 * (Tw.y * Tw.y)) + (R)-1 * (Tw.z * Tw.z);
  if (d == (R)0) return;
  const R inv = (R)1.0f / d;
  const R f0 = (R)1 * inv, f1 = (R)1 * inv, f2 = (R)-1 * inv;
  const R px = (f0 * (Tu.x * Tw.x) + f1 * (Tu.y * Tw.y)) + f2 * (Tu.z * Tw.z);
  const R py = (f0 * (Tv.x * Tw.x) + f1 * (Tv.y * Tw.y)) + f2 * (Tv.z * Tw.z);
  const R h0x = px * px - ((f0 * (Tu.x * Tu.x) + f1 * (Tu.y * Tu.y)) + f2 * (Tu.z * Tu.z));
  const R h0y = py * py - ((f0 * (Tv.x * Tv.x) + f1 * (Tv.y * Tv.y)) + f2 * (Tv.z * Tv.z));
  const R hx = std::sqrt(rmax((R)0, h0x)), hy = std::sqrt(rmax((R)0, h0y));
  // radius = ceil(3 * max(max(ext.x, ext.y), FilterSize)) with the double
  // FilterSize (forward.cu:235-239)
  const R ext = std::fmax(hx, hy);
  const float radius = (float)std::ceil(3.0 * std::fmax((double)ext, kP_FS));
  const int ir = f2i(radius);
  int x0, y0, x1, y1;
  get_rect(px, py, ir, s.B, c.gx, c.gy, x0, y0, x1, y1);
  const uint32_t area = (uint32_t)(x1 - x0) * (uint32_t)(y1 - y0);
  if (area == 0) return;
  if (!s.has_colors) {
    const V3<R> rgb = color_from_sh(c, idx, &c.clamped[3 * idx]);
    c.rgb[3 * idx] = rgb.x;
    c.rgb[3 * idx + 1] = rgb.y;
    c.rgb[3 * idx + 2] = rgb.z;
  } else {
    for (int k = 0; k < 3; ++k) c.rgb[3 * idx + k] = c.colors[3 * idx + k];
  }
  c.flip[idx] = 1;
  c.depths[idx] = vz;
  c.radii[idx] = ir;
  c.means2D[2 * idx] = px;
  c.means2D[2 * idx + 1] = py;
  R* T9 = &c.transMat[9 * idx];
  T9[0] = Tu.x; T9[1] = Tu.y; T9[2] = Tu.z;
  T9[3] = Tv.x; T9[4] = Tv.y; T9[5] = Tv.z;
  T9[6] = Tw.x; T9[7] = Tw.y; T9[8] = Tw.z;
  R* no = &c.normal_opacity[4 * idx];
  no[0] = nrm.x; no[1] = nrm.y; no[2] = nrm.z; no[3] = c.opac[idx];
  c.tiles_touched[idx] = area;
}

// The per-pair evaluation both passes share (dsp forward.cu:352-385,
// backward.cu:293-318).  Returns false for a skipped pair.
template <typename R> struct PartPair {
  V3<R> k, l, p;
  R sx, sy, rho3d, rho2d, dx, dy, depth, G, alpha;
};
template <typename R> bool part_pair(const Ctx<R>& c, int id, R pixx, R pixy, PartPair<R>& e) {
  const R* T9 = &c.transMat[9 * id];
  const V3<R> Tu = v3(T9[0], T9[1], T9[2]), Tv = v3(T9[3], T9[4], T9[5]), Tw = v3(T9[6], T9[7], T9[8]);
  e.k = v3(-Tu.x + pixx * Tw.x, -Tu.y + pixx * Tw.y, -Tu.z + pixx * Tw.z);
  e.l = v3(-Tv.x + pixy * Tw.x, -Tv.y + pixy * Tw.y, -Tv.z + pixy * Tw.z);
  e.p = cross(e.k, e.l);
  if (e.p.z == (R)0) return false;
  e.sx = e.p.x / e.p.z;
  e.sy = e.p.y / e.p.z;
  e.rho3d = e.sx * e.sx + e.sy * e.sy;
  e.dx = c.means2D[2 * id] - pixx;
  e.dy = c.means2D[2 * id + 1] - pixy;
  e.rho2d = part_fis_mul<R>(e.dx * e.dx + e.dy * e.dy);
  const R rho = rmin(e.rho3d, e.rho2d);
  e.depth = (e.rho3d <= e.rho2d) ? (e.sx * Tw.x + e.sy * Tw.y) + Tw.z : Tw.z;
  if ((double)e.depth < kP_NEAR) return false;
  const R power = (R)-0.5f * rho;
  if (power > (R)0) return false;
  e.G = rexp<R>(power);
  e.alpha = rmin((R)0.99f, c.normal_opacity[4 * id + 3] * e.G);
  return !(e.alpha < (R)(1.0f / 255.0f));
}

// mapped depth (FAR d - FAR NEAR) / ((FAR - NEAR) d) in double (forward.cu:393)
template <typename R> inline R part_mapped(R depth) {
  return (R)((kP_FAR * (double)depth - kP_FAR * kP_NEAR) / ((kP_FAR - kP_NEAR) * (double)depth));
}

// Forward blend of one pixel, dsp/cuda_rasterizer/forward.cu:256-474.
template <typename R> void render_pixel_part(Ctx<R>& c, int tile, int pxi, int pyi) {
  const int W = c.sc.W, H = c.sc.H, S = c.sc.S;
  const uint32_t rs = c.ranges[2 * tile], re = c.ranges[2 * tile + 1];
  const R pixx = (R)((double)(float)pxi + 0.5), pixy = (R)((double)(float)pyi + 0.5);
  R T = (R)1;
  uint32_t contributor = 0, last_contributor = 0;
  R C[3] = {0, 0, 0}, N[3] = {0, 0, 0}, Sem[16] = {0};
  R D = 0, dist1 = 0, dist2 = 0, distortion = 0, median_depth = 0, median_weight = 0, median_contributor = (R)-1;
  for (uint32_t e = rs; e < re; ++e) {
    contributor++;
    const int id = (int)c.point_list[e];
    PartPair<R> q;
    if (!part_pair(c, id, pixx, pixy, q)) continue;
    const R alpha = q.alpha, depth = q.depth;
    const R test_T = T * ((R)1 - alpha);
    if (test_T < (R)0.0001f) break;
    const R A = (R)1 - T;
    const R md = part_mapped(depth);
    const R error = ((md * md) * A + dist2) - ((R)2 * md) * dist1;
    distortion += (error * alpha) * T;
    if (T > (R)0.5) {
      median_depth = depth;
      median_weight = alpha * T;
      median_contributor = (R)contributor;
    }
    const R* no = &c.normal_opacity[4 * id];
    for (int ch = 0; ch < 3; ++ch) N[ch] += (no[ch] * alpha) * T;
    D += (depth * alpha) * T;
    dist1 += (md * alpha) * T;
    dist2 += ((md * md) * alpha) * T;
    for (int ch = 0; ch < 3; ++ch) C[ch] += (c.rgb[3 * id + ch] * alpha) * T;
    for (int ch = 0; ch < S; ++ch) Sem[ch] += (c.sem[(size_t)id * S + ch] * alpha) * T;
    T = test_T;
    last_contributor = contributor;
  }
  const size_t HW = (size_t)H * W, pid = (size_t)W * pyi + pxi;
  c.final_T[pid] = T;
  c.n_contrib[pid] = last_contributor;
  for (int ch = 0; ch < 3; ++ch) c.out_color[ch * HW + pid] = C[ch] + T * c.bg[ch];
  for (int ch = 0; ch < S; ++ch) c.out_sem[ch * HW + pid] = Sem[ch];
  c.n_contrib[pid + HW] = f2u(median_contributor);
  c.final_T[pid + HW] = dist1;
  c.final_T[pid + 2 * HW] = dist2;
  c.out_others[pid] = D;
  c.out_others[pid + HW] = (R)1 - T;
  for (int ch = 0; ch < 3; ++ch) c.out_others[pid + (2 + ch) * HW] = N[ch];
  c.out_others[pid + 5 * HW] = median_depth;
  c.out_others[pid + 6 * HW] = distortion;
  c.out_others[pid + 7 * HW] = median_weight;
}

// Backward blend of one pixel, dsp/cuda_rasterizer/backward.cu:140-470.
template <typename R>
void render_bwd_pixel_part(Ctx<R>& c, int tile, int pxi, int pyi, const R* dL_dpix, const R* dL_dsem,
                           const R* dL_dothers) {
  const int W = c.sc.W, H = c.sc.H, S = c.sc.S;
  const size_t HW = (size_t)H * W, pid = (size_t)W * pyi + pxi;
  const uint32_t rs = c.ranges[2 * tile], re = c.ranges[2 * tile + 1];
  const R pixx = (R)((double)(float)pxi + 0.5), pixy = (R)((double)(float)pyi + 0.5);
  const R T_final = c.final_T[pid];
  R T = T_final;
  uint32_t contributor = re - rs;
  const int last_contributor = (int)c.n_contrib[pid];
  const int median_contributor = (int)c.n_contrib[pid + HW];
  R accum_rec[3] = {0, 0, 0}, dL_dpixel[3], last_color[3] = {0, 0, 0}, dL_dpixel_sem[16] = {0};
  const R dL_ddepth = dL_dothers[pid], dL_daccum = dL_dothers[HW + pid], dL_dreg = dL_dothers[6 * HW + pid];
  R dL_dnormal2D[3];
  for (int i = 0; i < 3; ++i) dL_dnormal2D[i] = dL_dothers[(2 + i) * HW + pid];
  const R dL_dmedian_depth = dL_dothers[5 * HW + pid], dL_dmax_dweight = dL_dothers[7 * HW + pid];
  R last_depth = 0, last_normal[3] = {0, 0, 0}, accum_depth_rec = 0, accum_alpha_rec = 0;
  R accum_normal_rec[3] = {0, 0, 0};
  const R final_D = c.final_T[pid + HW], final_D2 = c.final_T[pid + 2 * HW];
  const R final_A = (R)1 - T_final;
  R last_dL_dT = 0, last_alpha = 0;
  for (int i = 0; i < 3; ++i) dL_dpixel[i] = dL_dpix[i * HW + pid];
  for (int i = 0; i < S; ++i) dL_dpixel_sem[i] = dL_dsem[i * HW + pid];
  for (uint32_t e = re; e > rs; --e) {
    contributor--;
    if (contributor >= (uint32_t)last_contributor) continue;
    const int id = (int)c.point_list[e - 1];
    PartPair<R> q;
    if (!part_pair(c, id, pixx, pixy, q)) continue;
    const R G = q.G, alpha = q.alpha, c_d = q.depth;
    T = T / ((R)1 - alpha);
    const R dchannel_dcolor = alpha * T;
    R dL_dalpha = 0;
    for (int ch = 0; ch < 3; ++ch) {
      const R col = c.rgb[3 * id + ch];
      accum_rec[ch] = last_alpha * last_color[ch] + ((R)1 - last_alpha) * accum_rec[ch];
      last_color[ch] = col;
      dL_dalpha += (col - accum_rec[ch]) * dL_dpixel[ch];
      acc_add(c, &c.acc_colors[3 * id + ch], (double)(dchannel_dcolor * dL_dpixel[ch]));
    }
    // semantics: their gradient only (the alpha term is commented out in
    // the reference, backward.cu:346)
    for (int ch = 0; ch < S; ++ch)
      acc_add(c, &c.acc_sem[(size_t)id * S + ch], (double)(dchannel_dcolor * dL_dpixel_sem[ch]));
    R dL_dz = 0, dL_dweight = 0;
    const R m_d = part_mapped(c_d);
    const R dmd_dd = (R)((kP_FAR * kP_NEAR) / (((kP_FAR - kP_NEAR) * (double)c_d) * (double)c_d));
    if ((int)contributor == median_contributor - 1) {
      dL_dz += dL_dmedian_depth;
      dL_dweight += dL_dmax_dweight;
    }
    dL_dweight += ((final_D2 + (m_d * m_d) * final_A) - ((R)2 * m_d) * final_D) * dL_dreg;
    dL_dalpha += dL_dweight - last_dL_dT;
    last_dL_dT = dL_dweight * alpha + ((R)1 - alpha) * last_dL_dT;
    const R dL_dmd = (((R)2.0f * (T * alpha)) * (m_d * final_A - final_D)) * dL_dreg;
    dL_dz += dL_dmd * dmd_dd;
    accum_depth_rec = last_alpha * last_depth + ((R)1 - last_alpha) * accum_depth_rec;
    last_depth = c_d;
    dL_dalpha += (c_d - accum_depth_rec) * dL_ddepth;
    accum_alpha_rec = (R)((double)last_alpha * 1.0 + (double)(((R)1 - last_alpha) * accum_alpha_rec));
    dL_dalpha += ((R)1 - accum_alpha_rec) * dL_daccum;
    const R* no = &c.normal_opacity[4 * id];
    for (int ch = 0; ch < 3; ++ch) {
      accum_normal_rec[ch] = last_alpha * last_normal[ch] + ((R)1 - last_alpha) * accum_normal_rec[ch];
      last_normal[ch] = no[ch];
      dL_dalpha += (no[ch] - accum_normal_rec[ch]) * dL_dnormal2D[ch];
      acc_add(c, &c.acc_normal[3 * id + ch], (double)((alpha * T) * dL_dnormal2D[ch]));
    }
    dL_dalpha *= T;
    last_alpha = alpha;
    R bg_dot_dpixel = 0;
    for (int i = 0; i < 3; ++i) bg_dot_dpixel += c.bg[i] * dL_dpixel[i];
    dL_dalpha += (-T_final / ((R)1 - alpha)) * bg_dot_dpixel;
    const R dL_dG = no[3] * dL_dalpha;
    dL_dz += (alpha * T) * dL_ddepth;
    const R* T9 = &c.transMat[9 * id];
    if (q.rho3d <= q.rho2d) {
      const R dL_dsx = (dL_dG * -G) * q.sx + dL_dz * T9[6];
      const R dL_dsy = (dL_dG * -G) * q.sy + dL_dz * T9[7];
      const R dsx_pz = dL_dsx / q.p.z, dsy_pz = dL_dsy / q.p.z;
      const V3<R> dL_dp = v3(dsx_pz, dsy_pz, -(dsx_pz * q.sx + dsy_pz * q.sy));
      const V3<R> dL_dk = cross(q.l, dL_dp);
      const V3<R> dL_dl = cross(dL_dp, q.k);
      double* g = &c.acc_transMat[9 * id];
      acc_add(c, &g[0], (double)(-dL_dk.x));
      acc_add(c, &g[1], (double)(-dL_dk.y));
      acc_add(c, &g[2], (double)(-dL_dk.z));
      acc_add(c, &g[3], (double)(-dL_dl.x));
      acc_add(c, &g[4], (double)(-dL_dl.y));
      acc_add(c, &g[5], (double)(-dL_dl.z));
      acc_add(c, &g[6], (double)((pixx * dL_dk.x + pixy * dL_dl.x) + dL_dz * q.sx));
      acc_add(c, &g[7], (double)((pixx * dL_dk.y + pixy * dL_dl.y) + dL_dz * q.sy));
      acc_add(c, &g[8], (double)((pixx * dL_dk.z + pixy * dL_dl.z) + dL_dz * (R)1));
    } else {
      // -G * FilterInvSquare * d.x: ((-G * 1) / (FS * FS)) * d.x in double
      const R dG_ddelx = (R)(((double)(-G) / (kP_FS * kP_FS)) * (double)q.dx);
      const R dG_ddely = (R)(((double)(-G) / (kP_FS * kP_FS)) * (double)q.dy);
      acc_add(c, &c.acc_mean2D[3 * id], (double)(dL_dG * dG_ddelx));
      acc_add(c, &c.acc_mean2D[3 * id + 1], (double)(dL_dG * dG_ddely));
      acc_add(c, &c.acc_transMat[9 * id + 8], (double)dL_dz);
    }
    acc_add(c, &c.acc_opacity[id], (double)(G * dL_dalpha));
  }
}

// Backward preprocess, dsp/cuda_rasterizer/backward.cu:473-700: the AABB
// adjoint added into dL_dtransMat for every visible surfel and the
// densification statistic (computeAABB, :621-671), then the camera-space
// homography's adjoint (computeTransMat, :473-545) and the SH backward.
template <typename R> void preprocess_bwd_one_part(Ctx<R>& c, int idx) {
  const Scene& s = c.sc;
  if (!(c.radii[idx] > 0)) return;
  const R fx = (R)(float)s.W / ((R)2.0f * (R)(float)s.tanx);
  const R fy = (R)(float)s.H / ((R)2.0f * (R)(float)s.tany);
  const R tx = (R)(float)s.tanx, ty = (R)(float)s.tany;
  const R Wh = fx * tx, Hh = fy * ty;  // "W", "H" of BACKWARD::preprocess (:685-686)
  R* gT = &c.g_transMat[9 * idx];
  const R* T9 = &c.transMat[9 * idx];
  const V3<R> T0 = v3(T9[0], T9[1], T9[2]), T1 = v3(T9[3], T9[4], T9[5]), T3 = v3(T9[6], T9[7], T9[8]);
  const R gmx = c.g_mean2D[3 * idx], gmy = c.g_mean2D[3 * idx + 1];
  {
    const V3<R> tv = v3<R>(1, 1, -1);
    const R d = (tv.x * (T3.x * T3.x) + tv.y * (T3.y * T3.y)) + tv.z * (T3.z * T3.z);
    const R invd = (R)1.0f / d;
    const V3<R> f = v3(tv.x * invd, tv.y * invd, tv.z * invd);
    auto had = [](V3<R> a, V3<R> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); };
    const V3<R> dL_dT0 = had(gmx * f, T3);
    const V3<R> dL_dT1 = had(gmy * f, T3);
    V3<R> dL_dT3 = had(gmx * f, T0) + had(gmy * f, T1);
    const V3<R> dL_df = had(gmx * T0, T3) + had(gmy * T1, T3);
    const R dotdf = (dL_df.x * f.x + dL_df.y * f.y) + dL_df.z * f.z;
    const R dL_dd = (R)((double)dotdf * (-1.0 / (double)d));
    const V3<R> dd_dT3 = v3((tv.x * T3.x) * (R)2.0f, (tv.y * T3.y) * (R)2.0f, (tv.z * T3.z) * (R)2.0f);
    dL_dT3 = dL_dT3 + dL_dd * dd_dT3;
    gT[0] += dL_dT0.x; gT[1] += dL_dT0.y; gT[2] += dL_dT0.z;
    gT[3] += dL_dT1.x; gT[4] += dL_dT1.y; gT[5] += dL_dT1.z;
    gT[6] += dL_dT3.x; gT[7] += dL_dT3.y; gT[8] += dL_dT3.z;
    const R z = T9[8];
    c.g_mean2D[3 * idx] = (gT[2] * z) * Wh;
    c.g_mean2D[3 * idx + 1] = (gT[5] * z) * Hh;
  }
  const R* m = &c.means3D[3 * idx];
  if (!s.has_tprec) {
    const R* vm = c.vm.data();
    const R cx = fx * tx, cy = fy * ty;  // intrins.zw of the backward (:596)
    R qw, qx, qy, qz;
    V3<R> Rc[3];
    quat_cols(&c.rots[4 * idx], qw, qx, qy, qz, Rc);
    const R sx = c.scales[2 * idx], sy = c.scales[2 * idx + 1];
    // dL_dM = (P^T dL_dT^T) (:512-518): column j = (fx dTu_j, fy dTv_j, (cx dTu_j + cy dTv_j) + dTw_j)
    V3<R> dM[3];
    for (int j = 0; j < 3; ++j) dM[j] = v3(fx * gT[j], fy * gT[3 + j], (cx * gT[j] + cy * gT[3 + j]) + gT[6 + j]);
    auto Wt = [&](const V3<R>& v) {
      return v3((vm[0] * v.x + vm[1] * v.y) + vm[2] * v.z, (vm[4] * v.x + vm[5] * v.y) + vm[6] * v.z,
                (vm[8] * v.x + vm[9] * v.y) + vm[10] * v.z);
    };
    const V3<R> dRS0 = Wt(dM[0]), dRS1 = Wt(dM[1]), dpw = Wt(dM[2]);
    const R* gn = &c.g_normal[3 * idx];
    V3<R> dL_dtn = Wt(v3(gn[0], gn[1], gn[2]));
    // DUAL_VISIABLE (:528-533): the sign from tn = W R[2] against p_view
    const V3<R> L2 = Rc[2];
    const V3<R> tn = v3((vm[0] * L2.x + vm[4] * L2.y) + vm[8] * L2.z, (vm[1] * L2.x + vm[5] * L2.y) + vm[9] * L2.z,
                        (vm[2] * L2.x + vm[6] * L2.y) + vm[10] * L2.z);
    const R pvx = ((vm[0] * m[0] + vm[4] * m[1]) + vm[8] * m[2]) + vm[12];
    const R pvy = ((vm[1] * m[0] + vm[5] * m[1]) + vm[9] * m[2]) + vm[13];
    const R pvz = ((vm[2] * m[0] + vm[6] * m[1]) + vm[10] * m[2]) + vm[14];
    const R cosv = ((-tn.x * pvx) + (-tn.y * pvy)) + (-tn.z * pvz);
    const R mult = cosv > (R)0 ? (R)1 : (R)-1;
    dL_dtn = mult * dL_dtn;
    const V3<R> dR0 = v3(dRS0.x * sx, dRS0.y * sx, dRS0.z * sx);
    const V3<R> dR1 = v3(dRS1.x * sy, dRS1.y * sy, dRS1.z * sy);
    const R vR[3][3] = {{dR0.x, dR0.y, dR0.z}, {dR1.x, dR1.y, dR1.z}, {dL_dtn.x, dL_dtn.y, dL_dtn.z}};
    const R w = qw, x = qx, y = qy, z = qz;
    R* gq = &c.g_rots[4 * idx];
    gq[0] = (R)2 * ((x * (vR[1][2] - vR[2][1]) + y * (vR[2][0] - vR[0][2])) + z * (vR[0][1] - vR[1][0]));
    gq[1] = (R)2 * ((((R)-2 * x) * (vR[1][1] + vR[2][2]) + y * (vR[0][1] + vR[1][0])) + z * (vR[0][2] + vR[2][0]) +
                    w * (vR[1][2] - vR[2][1]));
    gq[2] = (R)2 * (((x * (vR[0][1] + vR[1][0]) - ((R)2 * y) * (vR[0][0] + vR[2][2])) + z * (vR[1][2] + vR[2][1])) +
                    w * (vR[2][0] - vR[0][2]));
    gq[3] = (R)2 * (((x * (vR[0][2] + vR[2][0]) + y * (vR[1][2] + vR[2][1])) - ((R)2 * z) * (vR[0][0] + vR[1][1])) +
                    w * (vR[0][1] - vR[1][0]));
    c.g_scales[2 * idx] = (dRS0.x * Rc[0].x + dRS0.y * Rc[0].y) + dRS0.z * Rc[0].z;
    c.g_scales[2 * idx + 1] = (dRS1.x * Rc[1].x + dRS1.y * Rc[1].y) + dRS1.z * Rc[1].z;
    c.g_means3D[3 * idx] = dpw.x;
    c.g_means3D[3 * idx + 1] = dpw.y;
    c.g_means3D[3 * idx + 2] = dpw.z;
  }
  if (s.has_shs) sh_backward(c, idx);
}

template <typename R> void run_forward(Ctx<R>& c) {
  const Scene& s = c.sc;
  const int P = s.P;
  c.gx = (s.W + s.B - 1) / s.B;
  c.gy = (s.H + s.B - 1) / s.B;
  c.T = c.gx * c.gy;
  c.N = s.W * s.H;
  c.radii.assign(P, 0);
  c.tiles_touched.assign(P, 0);
  c.depths.assign(P, 0);
  c.means2D.assign(2 * (size_t)P, 0);
  c.transMat.assign(9 * (size_t)P, 0);
  c.normal_opacity.assign(4 * (size_t)P, 0);
  c.rgb.assign(3 * (size_t)P, 0);
  c.clamped.assign(3 * (size_t)P, 0);
  c.flip.assign((size_t)P, 0);
  c.prefilter_violation = 0;
  if (c.record) {
    c.dec_mult.assign(P, 0);
    c.dec_pix.assign((size_t)s.W * s.H, std::vector<uint32_t>());
  }
  if (s.part) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) preprocess_one_part(c, i);
  } else {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) preprocess_one(c, i);
  }
  binning(c);
  c.out_color.assign(3 * (size_t)c.N, 0);
  c.out_others.assign((s.part ? 8 : 7) * (size_t)c.N, 0);
  c.out_sem.assign((size_t)s.S * c.N, 0);
  c.final_T.assign(3 * (size_t)c.N, 0);
  c.n_contrib.assign(2 * (size_t)c.N, 0);
  const int B = s.B;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < c.T; ++t) {
    const int tx = t % c.gx, ty = t / c.gx;
    for (int j = 0; j < B; ++j)
      for (int i = 0; i < B; ++i) {
        const int x = tx * B + i, y = ty * B + j;
        if (x < s.W && y < s.H) {
          if (s.part) render_pixel_part(c, t, x, y);
          else render_pixel(c, t, x, y);
        }
      }
  }
  if (c.record) {
    int n = 0;
    for (const auto& v : c.dec_pix)
      for (const uint32_t w : v) n += (w & 2u) ? 1 : 0;
    c.dec_alpha_clamped = n;
  }
}


// ---------------------------------------------------------------------------
// Per-pixel backward blend, dsr/cuda_rasterizer/backward.cu:143-440.
// ---------------------------------------------------------------------------
template <typename R>
void render_bwd_pixel(Ctx<R>& c, int tile, int pxi, int pyi, const R* dL_dpix, const R* dL_dothers) {
  const int W = c.sc.W, H = c.sc.H;
  const size_t HW = (size_t)H * W, pid = (size_t)W * pyi + pxi;
  const uint32_t rs = c.ranges[2 * tile], re = c.ranges[2 * tile + 1];
  const R pixx = (R)pxi, pixy = (R)pyi;
  const R T_final = c.final_T[pid];
  R T = T_final;
  uint32_t contributor = re - rs;
  const int last_contributor = (int)c.n_contrib[pid];
  const int median_contributor = (int)c.n_contrib[pid + HW];
  R accum_rec[3] = {0, 0, 0}, dL_dpixel[3], last_color[3] = {0, 0, 0};
  const R dL_ddepth = dL_dothers[pid];
  const R dL_daccum = dL_dothers[HW + pid];
  const R dL_dreg = dL_dothers[6 * HW + pid];
  R dL_dnormal2D[3];
  for (int i = 0; i < 3; ++i) dL_dnormal2D[i] = dL_dothers[(2 + i) * HW + pid];
  const R dL_dmedian_depth = dL_dothers[5 * HW + pid];
  R last_depth = 0, last_normal[3] = {0, 0, 0}, accum_depth_rec = 0, accum_alpha_rec = 0;
  R accum_normal_rec[3] = {0, 0, 0};
  const R final_D = c.final_T[pid + HW], final_D2 = c.final_T[pid + 2 * HW];
  const R final_A = (R)1 - T_final;
  R last_dL_dT = 0;
  for (int i = 0; i < 3; ++i) dL_dpixel[i] = dL_dpix[i * HW + pid];
  R last_alpha = 0;
  const R near_n = (R)0.2f, far_n = (R)100.0f;
  const R mcoef = far_n / (far_n - near_n);
  const R dmd_num = far_n * near_n, dmd_den0 = far_n - near_n;
  R bg_dot_dpixel = 0;
  for (int i = 0; i < 3; ++i) bg_dot_dpixel += c.bg[i] * dL_dpixel[i];
  for (uint32_t e = re; e > rs; --e) {
    contributor--;
    if (contributor >= (uint32_t)last_contributor) continue;
    const int id = (int)c.point_list[e - 1];
    const R* T9 = &c.transMat[9 * id];
    const V3<R> Tu = v3(T9[0], T9[1], T9[2]), Tv = v3(T9[3], T9[4], T9[5]), Tw = v3(T9[6], T9[7], T9[8]);
    const V3<R> k = pixx * Tw - Tu;
    const V3<R> l = pixy * Tw - Tv;
    const V3<R> p = cross(k, l);
    if (p.z == (R)0) continue;
    const R sx = p.x / p.z, sy = p.y / p.z;
    const R rho3d = sx * sx + sy * sy;
    const R dx = c.means2D[2 * id] - pixx, dy = c.means2D[2 * id + 1] - pixy;
    const R rho2d = (R)2 * (dx * dx + dy * dy);
    const R rho = rmin(rho3d, rho2d);
    const R c_d = (rho3d <= rho2d) ? (sx * Tw.x + sy * Tw.y) + Tw.z : Tw.z;
    if (c_d < near_n) continue;
    const R* no = &c.normal_opacity[4 * id];
    const R power = (R)-0.5f * rho;
    if (power > (R)0) continue;
    const R G = rexp<R>(power);
    const R alpha = rmin((R)0.99f, no[3] * G);
    if (alpha < (R)(1.0f / 255.0f)) continue;
    T = T / ((R)1 - alpha);
    const R w = alpha * T;
    R dL_dalpha = 0;
    for (int ch = 0; ch < 3; ++ch) {
      const R col = c.rgb[3 * id + ch];
      accum_rec[ch] = last_alpha * last_color[ch] + ((R)1 - last_alpha) * accum_rec[ch];
      last_color[ch] = col;
      dL_dalpha += (col - accum_rec[ch]) * dL_dpixel[ch];
      acc_add(c, &c.acc_colors[3 * id + ch], (double)(w * dL_dpixel[ch]));
    }
    R dL_dz = 0, dL_dweight = 0;
    const R m_d = mcoef * ((R)1 - near_n / c_d);
    const R dmd_dd = dmd_num / ((dmd_den0 * c_d) * c_d);
    if ((int)contributor == median_contributor - 1) dL_dz += dL_dmedian_depth;
    dL_dweight += ((final_D2 + (m_d * m_d) * final_A) - ((R)2 * m_d) * final_D) * dL_dreg;
    dL_dalpha += dL_dweight - last_dL_dT;
    last_dL_dT = dL_dweight * alpha + ((R)1 - alpha) * last_dL_dT;
    const R dL_dmd = (((R)2 * (T * alpha)) * (m_d * final_A - final_D)) * dL_dreg;
    dL_dz += dL_dmd * dmd_dd;
    accum_depth_rec = last_alpha * last_depth + ((R)1 - last_alpha) * accum_depth_rec;
    last_depth = c_d;
    dL_dalpha += (c_d - accum_depth_rec) * dL_ddepth;
    accum_alpha_rec = last_alpha + ((R)1 - last_alpha) * accum_alpha_rec;
    dL_dalpha += ((R)1 - accum_alpha_rec) * dL_daccum;
    for (int ch = 0; ch < 3; ++ch) {
      accum_normal_rec[ch] = last_alpha * last_normal[ch] + ((R)1 - last_alpha) * accum_normal_rec[ch];
      last_normal[ch] = no[ch];
      dL_dalpha += (no[ch] - accum_normal_rec[ch]) * dL_dnormal2D[ch];
      acc_add(c, &c.acc_normal[3 * id + ch], (double)((alpha * T) * dL_dnormal2D[ch]));
    }
    dL_dalpha *= T;
    last_alpha = alpha;
    dL_dalpha += (-T_final / ((R)1 - alpha)) * bg_dot_dpixel;
    const R dL_dG = no[3] * dL_dalpha;
    dL_dz += (alpha * T) * dL_ddepth;
    if (rho3d <= rho2d) {
      const R dL_dsx = (dL_dG * -G) * sx + dL_dz * Tw.x;
      const R dL_dsy = (dL_dG * -G) * sy + dL_dz * Tw.y;
      const R dsx_pz = dL_dsx / p.z, dsy_pz = dL_dsy / p.z;
      const V3<R> dL_dp = v3(dsx_pz, dsy_pz, -(dsx_pz * sx + dsy_pz * sy));
      const V3<R> dL_dk = cross(l, dL_dp);
      const V3<R> dL_dl = cross(dL_dp, k);
      double* g = &c.acc_transMat[9 * id];
      acc_add(c, &g[0], (double)(-dL_dk.x));
      acc_add(c, &g[1], (double)(-dL_dk.y));
      acc_add(c, &g[2], (double)(-dL_dk.z));
      acc_add(c, &g[3], (double)(-dL_dl.x));
      acc_add(c, &g[4], (double)(-dL_dl.y));
      acc_add(c, &g[5], (double)(-dL_dl.z));
      acc_add(c, &g[6], (double)((pixx * dL_dk.x + pixy * dL_dl.x) + dL_dz * sx));
      acc_add(c, &g[7], (double)((pixx * dL_dk.y + pixy * dL_dl.y) + dL_dz * sy));
      acc_add(c, &g[8], (double)((pixx * dL_dk.z + pixy * dL_dl.z) + dL_dz * (R)1));
    } else {
      const R dG_ddelx = (-G * (R)2) * dx;
      const R dG_ddely = (-G * (R)2) * dy;
      acc_add(c, &c.acc_mean2D[3 * id], (double)(dL_dG * dG_ddelx));
      acc_add(c, &c.acc_mean2D[3 * id + 1], (double)(dL_dG * dG_ddely));
      acc_add(c, &c.acc_transMat[9 * id + 8], (double)dL_dz);
    }
    acc_add(c, &c.acc_opacity[id], (double)(G * dL_dalpha));
  }
}

// dnormvdv, dsr/cuda_rasterizer/auxiliary.h:128-138.
template <typename R> V3<R> dnormvdv(V3<R> v, V3<R> dv) {
  const R sum2 = (v.x * v.x + v.y * v.y) + v.z * v.z;
  const R invsum32 = (R)1 / std::sqrt((sum2 * sum2) * sum2);
  V3<R> r;
  r.x = ((((sum2 - v.x * v.x) * dv.x) - (v.y * v.x) * dv.y) - (v.z * v.x) * dv.z) * invsum32;
  r.y = (((-v.x * v.y) * dv.x + (sum2 - v.y * v.y) * dv.y) - (v.z * v.y) * dv.z) * invsum32;
  r.z = (((-v.x * v.z) * dv.x - (v.y * v.z) * dv.y) + (sum2 - v.z * v.z) * dv.z) * invsum32;
  return r;
}

// SH backward, dsr/cuda_rasterizer/backward.cu:20-139.
template <typename R> void sh_backward(Ctx<R>& c, int idx) {
  const int deg = c.sc.D, M = c.sc.M;
  const R* m = &c.means3D[3 * idx];
  const V3<R> dir_orig = v3(m[0] - c.campos[0], m[1] - c.campos[1], m[2] - c.campos[2]);
  const R len = std::sqrt((dir_orig.x * dir_orig.x + dir_orig.y * dir_orig.y) + dir_orig.z * dir_orig.z);
  const V3<R> dir = v3(dir_orig.x / len, dir_orig.y / len, dir_orig.z / len);
  const R* sh = &c.shs[(size_t)idx * M * 3];
  auto S = [&](int k) { return v3(sh[3 * k], sh[3 * k + 1], sh[3 * k + 2]); };
  V3<R> dL_dRGB = v3(c.g_colors[3 * idx], c.g_colors[3 * idx + 1], c.g_colors[3 * idx + 2]);
  if (c.clamped[3 * idx + 0]) dL_dRGB.x = 0;
  if (c.clamped[3 * idx + 1]) dL_dRGB.y = 0;
  if (c.clamped[3 * idx + 2]) dL_dRGB.z = 0;
  const R C0 = (R)(float)kSH_C0, C1 = (R)(float)kSH_C1;
  R C2[5], C3[7];
  for (int i = 0; i < 5; ++i) C2[i] = (R)(float)kSH_C2[i];
  for (int i = 0; i < 7; ++i) C3[i] = (R)(float)kSH_C3[i];
  V3<R> dx = v3<R>(0, 0, 0), dy = v3<R>(0, 0, 0), dz = v3<R>(0, 0, 0);
  const R x = dir.x, y = dir.y, z = dir.z;
  R* g = &c.g_sh[(size_t)idx * M * 3];
  auto put = [&](int k, R coef) {
    g[3 * k] = coef * dL_dRGB.x;
    g[3 * k + 1] = coef * dL_dRGB.y;
    g[3 * k + 2] = coef * dL_dRGB.z;
  };
  put(0, C0);
  if (deg > 0) {
    put(1, -C1 * y);
    put(2, C1 * z);
    put(3, -C1 * x);
    dx = -C1 * S(3);
    dy = -C1 * S(1);
    dz = C1 * S(2);
    if (deg > 1) {
      const R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      put(4, C2[0] * xy);
      put(5, C2[1] * yz);
      put(6, C2[2] * (((R)2 * zz - xx) - yy));
      put(7, C2[3] * xz);
      put(8, C2[4] * (xx - yy));
      dx = dx + ((((C2[0] * y) * S(4) + ((C2[2] * (R)2) * -x) * S(6)) + (C2[3] * z) * S(7)) + ((C2[4] * (R)2) * x) * S(8));
      dy = dy + ((((C2[0] * x) * S(4) + (C2[1] * z) * S(5)) + ((C2[2] * (R)2) * -y) * S(6)) + ((C2[4] * (R)2) * -y) * S(8));
      dz = dz + (((C2[1] * y) * S(5) + (((C2[2] * (R)2) * (R)2) * z) * S(6)) + (C2[3] * x) * S(7));
      if (deg > 2) {
        put(9, (C3[0] * y) * ((R)3 * xx - yy));
        put(10, (C3[1] * xy) * z);
        put(11, (C3[2] * y) * (((R)4 * zz - xx) - yy));
        put(12, (C3[3] * z) * (((R)2 * zz - (R)3 * xx) - (R)3 * yy));
        put(13, (C3[4] * x) * (((R)4 * zz - xx) - yy));
        put(14, (C3[5] * z) * (xx - yy));
        put(15, (C3[6] * x) * (xx - (R)3 * yy));
        // term order and association follow backward.cu:99-122
        auto mul = [](V3<R> v, R f) { return v3(v.x * f, v.y * f, v.z * f); };
        V3<R> ax = mul(mul(mul(C3[0] * S(9), (R)3), (R)2), xy);
        ax = ax + mul(C3[1] * S(10), yz);
        ax = ax + mul(mul(C3[2] * S(11), (R)-2), xy);
        ax = ax + mul(mul(mul(C3[3] * S(12), (R)-3), (R)2), xz);
        ax = ax + mul(C3[4] * S(13), (((R)-3 * xx + (R)4 * zz) - yy));
        ax = ax + mul(mul(C3[5] * S(14), (R)2), xz);
        ax = ax + mul(mul(C3[6] * S(15), (R)3), (xx - yy));
        V3<R> ay = mul(mul(C3[0] * S(9), (R)3), (xx - yy));
        ay = ay + mul(C3[1] * S(10), xz);
        ay = ay + mul(C3[2] * S(11), (((R)-3 * yy + (R)4 * zz) - xx));
        ay = ay + mul(mul(mul(C3[3] * S(12), (R)-3), (R)2), yz);
        ay = ay + mul(mul(C3[4] * S(13), (R)-2), xy);
        ay = ay + mul(mul(C3[5] * S(14), (R)-2), yz);
        ay = ay + mul(mul(mul(C3[6] * S(15), (R)-3), (R)2), xy);
        V3<R> az = mul(C3[1] * S(10), xy);
        az = az + mul(mul(mul(C3[2] * S(11), (R)4), (R)2), yz);
        az = az + mul(mul(C3[3] * S(12), (R)3), (((R)2 * zz - xx) - yy));
        az = az + mul(mul(mul(C3[4] * S(13), (R)4), (R)2), xz);
        az = az + mul(C3[5] * S(14), (xx - yy));
        dx = dx + ax;
        dy = dy + ay;
        dz = dz + az;
      }
    }
  }
  auto dot3 = [](V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; };
  const V3<R> dL_ddir = v3(dot3(dx, dL_dRGB), dot3(dy, dL_dRGB), dot3(dz, dL_dRGB));
  const V3<R> dm = dnormvdv(dir_orig, dL_ddir);
  c.g_means3D[3 * idx] += dm.x;
  c.g_means3D[3 * idx + 1] += dm.y;
  c.g_means3D[3 * idx + 2] += dm.z;
}

// ---------------------------------------------------------------------------
// Backward preprocess, dsr/cuda_rasterizer/backward.cu:443-630 and the
// quaternion VJP auxiliary.h:238-282.
// ---------------------------------------------------------------------------
template <typename R> void preprocess_bwd_one(Ctx<R>& c, int idx) {
  const Scene& s = c.sc;
  if (!(c.radii[idx] > 0)) return;
  const R focal_x = (R)(float)s.W / ((R)2.0f * (R)(float)s.tanx);
  const R focal_y = (R)(float)s.H / ((R)2.0f * (R)(float)s.tany);
  const int W = f2i((focal_x * (R)(float)s.tanx) * (R)2);
  const int H = f2i((focal_y * (R)(float)s.tany) * (R)2);
  const R* pm = c.pm.data();
  const R* vm = c.vm.data();
  R* gT = &c.g_transMat[9 * idx];
  V3<R> T0, T1, T2, L0c, L1c;
  V3<R> Rc[3];
  R qw = 0, qx = 0, qy = 0, qz = 0, scx = 0, scy = 0;
  V3<R> Pc[3][2];  // P columns as (xyz, w) halves: P[k] = (a, b) with a = rows 0..2, b.x = row 3
  R P4[3][4] = {};
  const bool precomp = s.has_tprec;
  const R* m = &c.means3D[3 * idx];
  if (precomp) {
    const R* t = &c.tprec[9 * idx];
    T0 = v3(t[0], t[1], t[2]);
    T1 = v3(t[3], t[4], t[5]);
    T2 = v3(t[6], t[7], t[8]);
  } else {
    quat_cols(&c.rots[4 * idx], qw, qx, qy, qz, Rc);
    scx = c.scales[2 * idx];
    scy = c.scales[2 * idx + 1];
    // scale_to_mat(scale, 1.0f): the scale modifier is not applied here (backward.cu:481)
    L0c = ((R)1 * scx) * Rc[0];
    L1c = ((R)1 * scy) * Rc[1];
    const R hw = (R)W / (R)2, hw1 = (R)(W - 1) / (R)2, hh = (R)H / (R)2, hh1 = (R)(H - 1) / (R)2;
    // P = world2ndc * ndc2pix (backward.cu:503): P[k](i) for rows i = 0..3
    for (int i = 0; i < 4; ++i) {
      P4[0][i] = pm[4 * i] * hw + pm[4 * i + 3] * hw1;
      P4[1][i] = pm[4 * i + 1] * hh + pm[4 * i + 3] * hh1;
      P4[2][i] = pm[4 * i + 3];
    }
    // T = transpose(M) * P, T(i,k) = dot(M_i, P[k]) with M_0 = (L0,0), M_1 = (L1,0), M_2 = (p,1)
    auto dotM = [&](const V3<R>& a, R a4, int k) {
      return ((a.x * P4[k][0] + a.y * P4[k][1]) + a.z * P4[k][2]) + a4 * P4[k][3];
    };
    const V3<R> pc = v3(m[0], m[1], m[2]);
    T0 = v3(dotM(L0c, 0, 0), dotM(L1c, 0, 0), dotM(pc, 1, 0));
    T1 = v3(dotM(L0c, 0, 1), dotM(L1c, 0, 1), dotM(pc, 1, 1));
    T2 = v3(dotM(L0c, 0, 2), dotM(L1c, 0, 2), dotM(pc, 1, 2));
  }
  (void)Pc;
  V3<R> dT0 = v3(gT[0], gT[1], gT[2]), dT1 = v3(gT[3], gT[4], gT[5]), dT2 = v3(gT[6], gT[7], gT[8]);
  const R gmx = c.g_mean2D[3 * idx], gmy = c.g_mean2D[3 * idx + 1];
  if (gmx != (R)0 || gmy != (R)0) {
    // adjoint of the AABB centre (backward.cu:515-530); mixed double steps mirror the reference
    const V3<R> tv = v3<R>(9, 9, -1);
    const R d = (tv.x * (T2.x * T2.x) + tv.y * (T2.y * T2.y)) + tv.z * (T2.z * T2.z);
    const R invd = (R)1.0f / d;
    const V3<R> f = v3(tv.x * invd, tv.y * invd, tv.z * invd);
    auto had = [](V3<R> a, V3<R> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); };
    const V3<R> dL_dT0 = had(gmx * f, T2);
    const V3<R> dL_dT1 = had(gmy * f, T2);
    V3<R> dL_dT3 = had(gmx * f, T0) + had(gmy * f, T1);
    const V3<R> dL_df = had(gmx * T0, T2) + had(gmy * T1, T2);
    const R dotdf = (dL_df.x * f.x + dL_df.y * f.y) + dL_df.z * f.z;
    const R dL_dd = (R)((double)dotdf * (-1.0 / (double)d));
    const V3<R> dd_dT3 = v3((tv.x * T2.x) * (R)2.0f, (tv.y * T2.y) * (R)2.0f, (tv.z * T2.z) * (R)2.0f);
    dL_dT3 = dL_dT3 + dL_dd * dd_dT3;
    dT0 = dT0 + dL_dT0;
    dT1 = dT1 + dL_dT1;
    dT2 = dT2 + dL_dT3;
    if (precomp) {
      gT[0] = dT0.x; gT[1] = dT0.y; gT[2] = dT0.z;
      gT[3] = dT1.x; gT[4] = dT1.y; gT[5] = dT1.z;
      gT[6] = dT2.x; gT[7] = dT2.y; gT[8] = dT2.z;
    }
  }
  if (!precomp) {
    // dL_dM = P * transpose(dL_dT): column cc = P[0]*dT0[cc] + P[1]*dT1[cc] + P[2]*dT2[cc]
    auto comp = [](V3<R> v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); };
    R dM[3][3];
    for (int cc = 0; cc < 3; ++cc)
      for (int i = 0; i < 3; ++i)
        dM[cc][i] = (P4[0][i] * comp(dT0, cc) + P4[1][i] * comp(dT1, cc)) + P4[2][i] * comp(dT2, cc);
    const R* gn = &c.g_normal[3 * idx];
    V3<R> dL_dtn = v3((vm[0] * gn[0] + vm[1] * gn[1]) + vm[2] * gn[2], (vm[4] * gn[0] + vm[5] * gn[1]) + vm[6] * gn[2],
                      (vm[8] * gn[0] + vm[9] * gn[1]) + vm[10] * gn[2]);
    // the camera-facing sign the forward chose (backward.cu:550-555 re-derives
    // it from p_view . normal, which the forward's FP32 test gives too; a
    // surfel whose forward ran in FP64 keeps that decision)
    const R mult = c.flip[idx] > 0 ? (R)1 : (R)-1;
    dL_dtn = mult * dL_dtn;
    const V3<R> dRS0 = v3(dM[0][0], dM[0][1], dM[0][2]);
    const V3<R> dRS1 = v3(dM[1][0], dM[1][1], dM[1][2]);
    const V3<R> dR0 = v3(dRS0.x * scx, dRS0.y * scx, dRS0.z * scx);
    const V3<R> dR1 = v3(dRS1.x * scy, dRS1.y * scy, dRS1.z * scy);
    const V3<R> dR2 = dL_dtn;
    // quat_to_rotmat_vjp with v_R[col][row]
    const R vR[3][3] = {{dR0.x, dR0.y, dR0.z}, {dR1.x, dR1.y, dR1.z}, {dR2.x, dR2.y, dR2.z}};
    const R w = qw, x = qx, y = qy, z = qz;
    R* gq = &c.g_rots[4 * idx];
    gq[0] = (R)2 * ((x * (vR[1][2] - vR[2][1]) + y * (vR[2][0] - vR[0][2])) + z * (vR[0][1] - vR[1][0]));
    gq[1] = (R)2 * ((((R)-2 * x) * (vR[1][1] + vR[2][2]) + y * (vR[0][1] + vR[1][0])) + z * (vR[0][2] + vR[2][0]) +
                    w * (vR[1][2] - vR[2][1]));
    gq[2] = (R)2 * (((x * (vR[0][1] + vR[1][0]) - ((R)2 * y) * (vR[0][0] + vR[2][2])) + z * (vR[1][2] + vR[2][1])) +
                    w * (vR[2][0] - vR[0][2]));
    gq[3] = (R)2 * (((x * (vR[0][2] + vR[2][0]) + y * (vR[1][2] + vR[2][1])) - ((R)2 * z) * (vR[0][0] + vR[1][1])) +
                    w * (vR[0][1] - vR[1][0]));
    c.g_scales[2 * idx] = (dRS0.x * Rc[0].x + dRS0.y * Rc[0].y) + dRS0.z * Rc[0].z;
    c.g_scales[2 * idx + 1] = (dRS1.x * Rc[1].x + dRS1.y * Rc[1].y) + dRS1.z * Rc[1].z;
    c.g_means3D[3 * idx] = dM[2][0];
    c.g_means3D[3 * idx + 1] = dM[2][1];
    c.g_means3D[3 * idx + 2] = dM[2][2];
  }
  if (s.has_shs) sh_backward(c, idx);
  // densification statistic, backward.cu:626-629
  const R depth = c.transMat[9 * idx + 8];
  c.g_mean2D[3 * idx] = (R)(((double)(gT[2] * depth) * 0.5) * (double)(R)(float)W);
  c.g_mean2D[3 * idx + 1] = (R)(((double)(gT[5] * depth) * 0.5) * (double)(R)(float)H);
}

template <typename R>
void run_backward(Ctx<R>& c, const R* dL_dpix, const R* dL_dothers, const R* dL_dsem = nullptr) {
  const Scene& s = c.sc;
  const int P = s.P;
  // per-thread FP64 accumulators, no atomics: transMat 9, mean2D 3, normal 3,
  // opacity 1, colours 3 and the semantic channels per surfel
  const int nt = omp_get_max_threads();
  c.acc_n = (size_t)P * (19 + (size_t)s.S);
  c.acc_pool.assign(c.acc_n * (size_t)nt, 0.0);
  double* base = c.acc_pool.data();
  c.acc_transMat = base;
  c.acc_mean2D = base + 9 * (size_t)P;
  c.acc_normal = base + 12 * (size_t)P;
  c.acc_opacity = base + 15 * (size_t)P;
  c.acc_colors = base + 16 * (size_t)P;
  c.acc_sem = base + 19 * (size_t)P;
  const int B = s.B;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < c.T; ++t) {
    const int tx = t % c.gx, ty = t / c.gx;
    for (int j = 0; j < B; ++j)
      for (int i = 0; i < B; ++i) {
        const int x = tx * B + i, y = ty * B + j;
        if (x < s.W && y < s.H) {
          if (s.part) render_bwd_pixel_part(c, t, x, y, dL_dpix, dL_dsem, dL_dothers);
          else render_bwd_pixel(c, t, x, y, dL_dpix, dL_dothers);
        }
      }
  }
  // the thread copies into copy 0, in thread order
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)c.acc_n; ++i) {
    double v = base[i];
    for (int t = 1; t < nt; ++t) v += base[(size_t)t * c.acc_n + (size_t)i];
    base[i] = v;
  }
  auto round_into = [](std::vector<R>& dst, const double* src, size_t n) {
    dst.resize(n);
    for (size_t i = 0; i < n; ++i) dst[i] = (R)src[i];
  };
  round_into(c.g_transMat, c.acc_transMat, 9 * (size_t)P);
  round_into(c.g_mean2D, c.acc_mean2D, 3 * (size_t)P);
  round_into(c.g_normal, c.acc_normal, 3 * (size_t)P);
  round_into(c.g_opacity, c.acc_opacity, (size_t)P);
  round_into(c.g_colors, c.acc_colors, 3 * (size_t)P);
  c.g_means3D.assign(3 * (size_t)P, (R)0);
  c.g_sh.assign((size_t)P * s.M * 3, (R)0);
  c.g_scales.assign(2 * (size_t)P, (R)0);
  c.g_rots.assign(4 * (size_t)P, (R)0);
  round_into(c.g_sem, c.acc_sem, (size_t)P * s.S);
  if (s.part) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) preprocess_bwd_one_part(c, i);
  } else {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; ++i) preprocess_bwd_one(c, i);
  }
}

struct Handle {
  int f64 = 0;
  Ctx<float> cf;
  Ctx<double> cd;
  std::string err;
};


}  // namespace

namespace {
template <typename R> void load_scene(Ctx<R>& c, const sro_scene_c* s, int f64) {
  Scene& sc = c.sc;
  sc.P = s->P; sc.D = s->D; sc.M = s->M; sc.W = s->W; sc.H = s->H; sc.B = s->tile;
  sc.prefiltered = s->prefiltered;
  sc.has_shs = s->shs != nullptr;
  sc.has_colors = s->colors_precomp != nullptr;
  sc.has_tprec = s->transMat_precomp != nullptr;
  sc.has_sr = s->scales != nullptr;
  // Python floats cross the reference's pybind boundary as C float (rasterize_points.h:18-68).
  sc.mod = f64 ? s->scale_modifier : (double)(float)s->scale_modifier;
  sc.tanx = f64 ? s->tan_fovx : (double)(float)s->tan_fovx;
  sc.tany = f64 ? s->tan_fovy : (double)(float)s->tan_fovy;
  const size_t P = (size_t)s->P;
  copy_in(c.means3D, s->means3D, 3 * P, f64);
  copy_in(c.shs, s->shs, P * (size_t)(s->M > 0 ? s->M : 0) * 3, f64);
  copy_in(c.colors, s->colors_precomp, 3 * P, f64);
  copy_in(c.opac, s->opacities, P, f64);
  copy_in(c.scales, s->scales, 2 * P, f64);
  copy_in(c.rots, s->rotations, 4 * P, f64);
  copy_in(c.tprec, s->transMat_precomp, 9 * P, f64);
  copy_in(c.vm, s->viewmatrix, 16, f64);
  copy_in(c.pm, s->projmatrix, 16, f64);
  copy_in(c.campos, s->campos, 3, f64);
  copy_in(c.bg, s->bg, 3, f64);
}

template <typename T> int copy_out(const std::vector<T>& v, void* dst, size_t nbytes) {
  if (nbytes != v.size() * sizeof(T)) return -(int)(v.size() * sizeof(T)) - 1;
  if (nbytes) std::memcpy(dst, v.data(), nbytes);
  return 0;
}

// An accumulator of n doubles (copy 0 of the pool; absent before a backward).
inline int copy_out_acc(const double* p, size_t n, void* dst, size_t nbytes) {
  if (!p) n = 0;
  if (nbytes != n * sizeof(double)) return -(int)(n * sizeof(double)) - 1;
  if (nbytes) std::memcpy(dst, p, nbytes);
  return 0;
}

template <typename R> int get_named(Ctx<R>& c, const std::string& n, void* dst, size_t nbytes) {
  if (n == "radii") return copy_out(c.radii, dst, nbytes);
  if (n == "depths") return copy_out(c.depths, dst, nbytes);
  if (n == "means2D") return copy_out(c.means2D, dst, nbytes);
  if (n == "transMat") return copy_out(c.transMat, dst, nbytes);
  if (n == "normal_opacity") return copy_out(c.normal_opacity, dst, nbytes);
  if (n == "rgb") return copy_out(c.rgb, dst, nbytes);
  if (n == "clamped") return copy_out(c.clamped, dst, nbytes);
  if (n == "flip") return copy_out(c.flip, dst, nbytes);
  if (n == "tiles_touched") return copy_out(c.tiles_touched, dst, nbytes);
  if (n == "point_offsets") return copy_out(c.point_offsets, dst, nbytes);
  if (n == "keys") return copy_out(c.keys, dst, nbytes);
  if (n == "point_list") return copy_out(c.point_list, dst, nbytes);
  if (n == "ranges") return copy_out(c.ranges, dst, nbytes);
  if (n == "out_color") return copy_out(c.out_color, dst, nbytes);
  if (n == "out_others") return copy_out(c.out_others, dst, nbytes);
  if (n == "out_semantic") return copy_out(c.out_sem, dst, nbytes);
  if (n == "dL_dsemantics") return copy_out(c.g_sem, dst, nbytes);
  if (n == "final_T") return copy_out(c.final_T, dst, nbytes);
  if (n == "n_contrib") return copy_out(c.n_contrib, dst, nbytes);
  if (n == "acc_dL_dtransMat") return copy_out_acc(c.acc_transMat, 9 * (size_t)c.sc.P, dst, nbytes);
  if (n == "acc_dL_dmean2D") return copy_out_acc(c.acc_mean2D, 3 * (size_t)c.sc.P, dst, nbytes);
  if (n == "acc_dL_dnormal") return copy_out_acc(c.acc_normal, 3 * (size_t)c.sc.P, dst, nbytes);
  if (n == "acc_dL_dopacity") return copy_out_acc(c.acc_opacity, (size_t)c.sc.P, dst, nbytes);
  if (n == "acc_dL_dcolors") return copy_out_acc(c.acc_colors, 3 * (size_t)c.sc.P, dst, nbytes);
  if (n == "dL_dmeans2D") return copy_out(c.g_mean2D, dst, nbytes);
  if (n == "dL_dcolors") return copy_out(c.g_colors, dst, nbytes);
  if (n == "dL_dopacity") return copy_out(c.g_opacity, dst, nbytes);
  if (n == "dL_dmeans3D") return copy_out(c.g_means3D, dst, nbytes);
  if (n == "dL_dtransMat") return copy_out(c.g_transMat, dst, nbytes);
  if (n == "dL_dsh") return copy_out(c.g_sh, dst, nbytes);
  if (n == "dL_dscales") return copy_out(c.g_scales, dst, nbytes);
  if (n == "dL_drotations") return copy_out(c.g_rots, dst, nbytes);
  if (n == "dL_dnormal") return copy_out(c.g_normal, dst, nbytes);
  return -1;
}
}  // namespace

extern "C" {

void* sro_create(void) { return new Handle(); }
void sro_destroy(void* h) { delete (Handle*)h; }

void sro_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int sro_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

float sro_expf(float x) { return sro_expf_impl(x); }

void sro_expf_array(const float* x, float* y, int n) {
  for (int i = 0; i < n; ++i) y[i] = sro_expf_impl(x[i]);
}

// SH -> RGB of n points (FP64 mode when f64 != 0), the colour step of the
// forward preprocess on its own: sh is n x M x 3, means n x 3, campos 3.
void sro_sh_rgb(int deg, int M, const void* sh, const void* means, const void* campos, int n, void* rgb,
                uint8_t* clamped, int f64) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    Ctx<R> c;
    c.sc.D = deg;
    c.sc.M = M;
    copy_in(c.shs, sh, (size_t)n * M * 3, f64);
    copy_in(c.means3D, means, (size_t)n * 3, f64);
    copy_in(c.campos, campos, 3, f64);
    for (int i = 0; i < n; ++i) {
      const V3<R> v = color_from_sh(c, i, clamped + 3 * i);
      R* o = (R*)rgb + 3 * i;
      o[0] = v.x;
      o[1] = v.y;
      o[2] = v.z;
    }
  };
  if (f64) run(0.0); else run(0.0f);
}
uint32_t sro_higher_msb(uint32_t n) { return higher_msb(n); }

// on != 0 (the default): near-degenerate surfels take the FP64 preprocess,
// as the HIP kernel does; 0: every surfel stays FP32, which is the
// reference's own arithmetic (used to measure what FP64 changes).
void sro_set_fp64_degenerate(int on) { g_fp64_degenerate = on != 0; }

// Runs preprocess, binning and the forward blend.  Returns 0, 1 when a
// surfel fails the near test although prefiltered is set (the reference
// traps there, auxiliary.h:202-206), or -1 on bad arguments.
int sro_forward(void* hv, const sro_scene_c* s, int f64) {
  Handle* h = (Handle*)hv;
  if (!h || !s || s->P < 0 || s->W <= 0 || s->H <= 0 || s->tile <= 0) return -1;
  if ((s->shs == nullptr) == (s->colors_precomp == nullptr)) return -1;
  if ((s->scales == nullptr || s->rotations == nullptr) == (s->transMat_precomp == nullptr)) return -1;
  h->f64 = f64;
  if (f64) {
    load_scene(h->cd, s, 1);
    run_forward(h->cd);
    return h->cd.prefilter_violation;
  }
  load_scene(h->cf, s, 0);
  run_forward(h->cf);
  return h->cf.prefilter_violation;
}

// Forward that keeps its discrete decisions for sro_replay.  Returns as
// sro_forward, or -2 when the instance count does not fit the record words.
int sro_forward_record(void* hv, const sro_scene_c* s, int f64) {
  Handle* h = (Handle*)hv;
  if (!h) return -1;
  h->cf.record = !f64;
  h->cd.record = f64 != 0;
  const int rc = sro_forward(hv, s, f64);
  h->cf.record = h->cd.record = false;
  const int L = f64 ? h->cd.L : h->cf.L;
  if (L >= (1 << 29)) return -2;
  return rc;
}

// Re-evaluates the forward recorded by sro_forward_record on new inputs
// (same P, image, tile, precision and optional-input pattern) with every
// discrete decision held: culls, normal flips, radii and tile lists, colour
// clamps, and per pixel which entries blend, their branch, alpha cap and
// median.  The outputs replace the recorded run's.
int sro_replay(void* hv, const sro_scene_c* s) {
  Handle* h = (Handle*)hv;
  if (!h || !s) return -1;
  auto go = [&](auto& c, int f64) {
    if (c.dec_pix.empty() || s->P != c.sc.P || s->W != c.sc.W || s->H != c.sc.H || s->tile != c.sc.B) return -1;
    load_scene(c, s, f64);
    run_replay(c);
    return 0;
  };
  return h->f64 ? go(h->cd, 1) : go(h->cf, 0);
}

// Blended pairs of the recorded forward whose alpha was capped at 0.99 (the
// backward differentiates through the cap, so a replay of such a pair is
// not what the gradient describes).
int sro_alpha_capped(void* hv) {
  Handle* h = (Handle*)hv;
  return h->f64 ? h->cd.dec_alpha_clamped : h->cf.dec_alpha_clamped;
}

// The _part variant's forward: as sro_forward, plus S <= 16 semantic
// channels per surfel (semantics P x S, the precision of the run).
int sro_forward_part(void* hv, const sro_scene_c* s, int f64, int S, const void* semantics) {
  Handle* h = (Handle*)hv;
  if (!h || !s || S < 0 || S > 16 || (S > 0 && !semantics)) return -1;
  if (!s || s->P < 0 || s->W <= 0 || s->H <= 0 || s->tile <= 0) return -1;
  if ((s->shs == nullptr) == (s->colors_precomp == nullptr)) return -1;
  if ((s->scales == nullptr || s->rotations == nullptr) == (s->transMat_precomp == nullptr)) return -1;
  h->f64 = f64;
  auto go = [&](auto& c) {
    load_scene(c, s, f64);
    c.sc.part = true;
    c.sc.S = S;
    copy_in(c.sem, semantics, (size_t)s->P * S, f64);
    run_forward(c);
    return c.prefilter_violation;
  };
  return f64 ? go(h->cd) : go(h->cf);
}

// Backward of sro_forward_part: dL_dcolor 3 x H x W, dL_dsemantic S x H x W,
// dL_dothers 8 x H x W.
int sro_backward_part(void* hv, const void* dL_dcolor, const void* dL_dsem, const void* dL_dothers) {
  Handle* h = (Handle*)hv;
  if (h->f64)
    run_backward(h->cd, (const double*)dL_dcolor, (const double*)dL_dothers, (const double*)dL_dsem);
  else
    run_backward(h->cf, (const float*)dL_dcolor, (const float*)dL_dothers, (const float*)dL_dsem);
  return 0;
}

int sro_num_rendered(void* hv) {
  Handle* h = (Handle*)hv;
  return h->f64 ? h->cd.L : h->cf.L;
}

// dL_dcolor (3,H,W) and dL_dothers (7,H,W) in the precision of the forward.
int sro_backward(void* hv, const void* dL_dcolor, const void* dL_dothers) {
  Handle* h = (Handle*)hv;
  if (h->f64) run_backward(h->cd, (const double*)dL_dcolor, (const double*)dL_dothers);
  else run_backward(h->cf, (const float*)dL_dcolor, (const float*)dL_dothers);
  return 0;
}

// Copies a named internal array out; returns 0, or a negative value whose
// magnitude minus one is the required byte count, or -1 for unknown names.
int sro_get(void* hv, const char* name, void* dst, size_t nbytes) {
  Handle* h = (Handle*)hv;
  const std::string n(name);
  return h->f64 ? get_named(h->cd, n, dst, nbytes) : get_named(h->cf, n, dst, nbytes);
}

// markVisible / checkFrustum, rasterizer_impl.cu:54-66,141-153.
void sro_mark_visible(int P, const float* means3D, const float* viewmatrix, uint8_t* present) {
  for (int i = 0; i < P; ++i) {
    const float* m = means3D + 3 * i;
    const float* vm = viewmatrix;
    const float vz = ((vm[2] * m[0] + vm[6] * m[1]) + vm[10] * m[2]) + vm[14];
    present[i] = !(vz <= 0.2f);
  }
}

}  // extern "C"
