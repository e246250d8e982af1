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

// Software pipeline over a quadrant's 64-entry chunks for the blend kernels:
// while one chunk is processed, the per-lane mask and pair-store offset of
// the next chunk, the five record quads of its entries and the sub-list
// entry of the chunk after it are already in flight, so a chunk's staging
// costs one memory latency that overlaps work instead of two in series.
// Lanes past the end get entry (0, 0) and read record 0, so every load is
// unconditional and in bounds; nothing reads those values.
SR_DEV uint2 chunk_entry(const uint2* list, int cnt, int lane, int c, int nch) {
  const int e = c * 64 + lane;
  return (c >= 0 && c < nch && e < cnt) ? list[e] : make_uint2(0u, 0u);
}
struct ChunkRegs {
  uint64_t mask;
  uint32_t foff;
  uint2 v;
  float4 q0, q1, q2, q3, q4;
};
SR_DEV ChunkRegs chunk_load(const uint64_t* masks, const uint32_t* frag_off, const float4* rec, int lane, int c,
                            uint2 ve) {
  ChunkRegs k;
  k.mask = masks[(size_t)c * 64 + lane];
  k.foff = frag_off[(size_t)c * 64 + lane];
  k.v = ve;
  const float4* r = rec + (size_t)kRecQuads * ve.x;
  k.q0 = r[0];
  k.q1 = r[1];
  k.q2 = r[2];
  k.q3 = r[3];
  k.q4 = r[4];
  return k;
}

}  // namespace sr
