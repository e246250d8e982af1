// Per-(pixel, surfel) evaluation shared by the acceptance pass and the two
// blend kernels: the ray-splat intersection, the screen-space low-pass
// filter, depth and alpha of dsr/cuda_rasterizer/forward.cu:349-381 (and the
// identical recomputation in backward.cu:282-314), in the reference's order.
#pragma once

#include "sr_device.h"
#include "sr_internal.h"

namespace sr {

struct Pair {
  float kx, ky, kz, lx, ly, lz;  // the two homogeneous planes k, l
  float px, py, pz;              // p = k x l
  float sx, sy;                  // intersection in splat coordinates
  float rho3d, rho2d, dx, dy;    // exponents; dx, dy = mean2D - pixel
  float depth, G, alpha;
};

// Returns true when the reference blends this pair (alpha >= 1/255 and no
// earlier "continue").  With reject set, pairs whose two exponents both
// provably exceed rho_reject return false before the divisions; their alpha
// would be < 1/255, so the result is the same.
template <bool kReject>
SR_DEV bool eval_pair(float pixx, float pixy, const float4& q0, const float4& q1, const float4& q2, float rho_rej,
                      Pair& e) {
  const float Tux = q0.x, Tuy = q0.y, Tuz = q0.z, Tvx = q0.w, Tvy = q1.x, Tvz = q1.y;
  const float Twx = q1.z, Twy = q1.w, Twz = q2.x;
  e.dx = q2.y - pixx;
  e.dy = q2.z - pixy;
  e.rho2d = 2.0f * (e.dx * e.dx + e.dy * e.dy);
  e.kx = pixx * Twx - Tux;
  e.ky = pixx * Twy - Tuy;
  e.kz = pixx * Twz - Tuz;
  e.lx = pixy * Twx - Tvx;
  e.ly = pixy * Twy - Tvy;
  e.lz = pixy * Twz - Tvz;
  e.pz = e.kx * e.ly - e.ky * e.lx;
  if (e.pz == 0.0f) return false;
  e.px = e.ky * e.lz - e.kz * e.ly;
  e.py = e.kz * e.lx - e.kx * e.lz;
  if (kReject && e.rho2d > rho_rej) {
    const float num = e.px * e.px + e.py * e.py, den = rho_rej * (e.pz * e.pz);
    if (num > den && num < 3.0e38f && den < 3.0e38f) return false;
  }
  e.sx = e.px / e.pz;
  e.sy = e.py / e.pz;
  e.rho3d = e.sx * e.sx + e.sy * e.sy;
  const float rho = fmin_(e.rho3d, e.rho2d);
  e.depth = (e.rho3d <= e.rho2d) ? (e.sx * Twx + e.sy * Twy) + Twz : Twz;
  if (e.depth < kNear) return false;
  const float power = -0.5f * rho;
  if (power > 0.0f) return false;
  e.G = exp_pinned(power);
  e.alpha = fmin_(0.99f, q2.w * e.G);
  return !(e.alpha < 1.0f / 255.0f);
}

// Rebuilds a Pair from its stored (alpha, depth), (G, s.x) and s.y: the
// cheap planes k, l, p.z and the screen offsets are recomputed exactly as in
// eval_pair, so every field equals eval_pair's.
SR_DEV void load_pair(float pixx, float pixy, const float4& q0, const float4& q1, const float4& q2,
                      const float2& fa, const float2& fb, float fc, Pair& e) {
  const float Tux = q0.x, Tuy = q0.y, Tuz = q0.z, Tvx = q0.w, Tvy = q1.x, Tvz = q1.y;
  const float Twx = q1.z, Twy = q1.w, Twz = q2.x;
  e.dx = q2.y - pixx;
  e.dy = q2.z - pixy;
  e.rho2d = 2.0f * (e.dx * e.dx + e.dy * e.dy);
  e.kx = pixx * Twx - Tux;
  e.ky = pixx * Twy - Tuy;
  e.kz = pixx * Twz - Tuz;
  e.lx = pixy * Twx - Tvx;
  e.ly = pixy * Twy - Tvy;
  e.lz = pixy * Twz - Tvz;
  e.pz = e.kx * e.ly - e.ky * e.lx;
  e.alpha = fa.x;
  e.depth = fa.y;
  e.G = fb.x;
  e.sx = fb.y;
  e.sy = fc;
  e.rho3d = e.sx * e.sx + e.sy * e.sy;
}

}  // namespace sr
