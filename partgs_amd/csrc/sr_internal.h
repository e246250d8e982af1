// Internal layouts and kernel-launch entry points of the surfel rasteriser.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sr {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Packed per-surfel record read by both blend kernels (5 x float4 = 80 B):
//   q0 = (Tu.x, Tu.y, Tu.z, Tv.x)   q1 = (Tv.y, Tv.z, Tw.x, Tw.y)
//   q2 = (Tw.z, mean2D.x, mean2D.y, opacity)
//   q3 = (normal.x, normal.y, normal.z, r)   q4 = (g, b, rho_reject, 0)
// Tu/Tv/Tw are the columns of the ray-splat homography T as stored by the
// reference (dsr/cuda_rasterizer/forward.cu:195-198); normal is the
// camera-facing view-space normal (forward.cu:209-214).  rho_reject bounds
// the Gaussian exponent: any pixel with min(rho3d, rho2d) > rho_reject has
// alpha < 1/255 and is skipped by the reference, so the blend kernels may
// skip it early (rho_reject < 0: the surfel never reaches 1/255).
constexpr int kRecQuads = 5;

// The blend kernels run one wave64 per 8 x 8 pixel quadrant of a tile; each
// quadrant walks an order-preserving sub-list of its tile's sorted list,
// holding only the entries whose screen cull box meets the quadrant.
constexpr int kQuad = 8;
__host__ __device__ inline int quads_per_tile(int B) { return (B / kQuad) * (B / kQuad); }

// Geometry state: one entry per surfel.
struct GeomLayout {
  size_t rec, depths, tiles, clamped, dnormal, cullbox, total;
};
inline GeomLayout geom_layout(int P) {
  GeomLayout g{};
  size_t off = 0;
  g.rec = off; off += align_up((size_t)P * kRecQuads * 16);
  g.depths = off; off += align_up((size_t)P * 4);
  g.tiles = off; off += align_up((size_t)P * 4);
  g.clamped = off; off += align_up((size_t)P * 4);  // 3 flags + pad per surfel
  g.dnormal = off; off += align_up((size_t)P * 12);  // backward scratch: dL/dnormal
  g.cullbox = off; off += align_up((size_t)P * 8);   // short4 inclusive pixel box
  g.total = off;
  return g;
}

// Image state: per-tile binning tables and per-pixel blend state.
struct ImageLayout {
  int gx, gy, T, N;
  size_t counters, tile_count, tile_start, tile_fill, ranges, sub_count, final_T, n_contrib, total;
};
inline ImageLayout image_layout(int W, int H, int B) {
  ImageLayout l{};
  l.gx = (W + B - 1) / B;
  l.gy = (H + B - 1) / B;
  l.T = l.gx * l.gy;
  l.N = W * H;
  size_t off = 0;
  l.counters = off; off += kAlign;  // [0] = instance count, [1] = error flags
  l.tile_count = off; off += align_up((size_t)l.T * 4);
  l.tile_start = off; off += align_up(((size_t)l.T + 1) * 4);
  l.tile_fill = off; off += align_up((size_t)l.T * 4);
  l.ranges = off; off += align_up((size_t)l.T * 8);
  l.sub_count = off; off += align_up((size_t)l.T * quads_per_tile(B) * 4);
  l.final_T = off; off += align_up((size_t)l.N * 3 * 4);
  l.n_contrib = off; off += align_up((size_t)l.N * 2 * 4);
  l.total = off;
  return l;
}

// Binning state: one entry per (surfel, tile) instance.
struct BinLayout {
  size_t keys, keys_tmp, point_list, sub, masks, work, frag_off, frag4, frag1, total;
  size_t frag_cap;
};
// Accepted-pair store: per accepted (pixel, entry) pair the acceptance pass
// keeps (alpha, G, depth, s.x) and s.y, where G = exp(-rho / 2) and s the
// intersection in splat coordinates; a lane's pairs of one chunk are
// contiguous from frag_off[mask slot].  Capacity kFragPerInstance pairs per
// tile instance; chunks that do not fit get frag_off = kNoFrag and the blend
// kernels recompute their pairs instead.
constexpr size_t kFragPerInstance = 24;
constexpr uint32_t kNoFrag = 0xFFFFFFFFu;
// Per-pixel acceptance masks: quadrant (t, q) owns mask slots
// [mask_base, mask_base + 64 * chunks), chunk c of its sub-list at
// mask_base + 64 c, one 64-bit word per pixel (bit j: entry 64 c + j reaches
// alpha >= 1/255 at that pixel).
__host__ __device__ inline size_t mask_base(int Q, uint32_t beg, uint32_t n, int tile, int q) {
  return (size_t)Q * beg + (size_t)q * n + (size_t)kQuad * kQuad * ((size_t)tile * Q + q);
}
inline size_t work_capacity(int L, int B, int T) {
  const size_t Q = (size_t)quads_per_tile(B);
  return Q * ((size_t)T + ((size_t)(L > 0 ? L : 1) + 63) / 64);
}
inline BinLayout bin_layout(int L, int B, int T, int pair_store = 0) {
  BinLayout b{};
  size_t off = 0;
  const size_t n = (size_t)(L > 0 ? L : 1);
  const size_t Q = (size_t)quads_per_tile(B);
  b.keys = off; off += align_up(n * 8);
  b.keys_tmp = off; off += align_up(n * 8);
  b.point_list = off; off += align_up(n * 4);
  b.sub = off; off += align_up(n * Q * 8);  // uint2 (surfel id, position in tile)
  b.masks = off; off += align_up((n * Q + (size_t)kQuad * kQuad * Q * (size_t)T) * 8);
  b.work = off; off += align_up(work_capacity(L, B, T) * 8);
  const size_t nslots = n * Q + (size_t)kQuad * kQuad * Q * (size_t)T;
  b.frag_off = off; off += align_up(nslots * 4);
  b.frag_cap = n * (pair_store > 0 ? (size_t)pair_store : kFragPerInstance);
  if (b.frag_cap > 0x7fffffffu) b.frag_cap = 0x7fffffffu;
  b.frag4 = off; off += align_up(b.frag_cap * 16);
  b.frag1 = off; off += align_up(b.frag_cap * 4);
  b.total = off;
  return b;
}

constexpr uint32_t kErrPrefiltered = 1u;

struct PreprocessArgs {
  int P, D, M, W, H, B, gx, gy, prefiltered;
  float mod;
  const float* means3D;
  const float* scales;
  const float* rots;
  const float* opac;
  const float* shs;
  const float* colors;
  const float* tprec;
  const float* vm;
  const float* pm;
  const float* campos;
  float4* rec;
  float* depths;
  int* radii;
  uint32_t* tiles;
  uint8_t* clamped;
  short4* cullbox;
  uint32_t* tile_count;
  uint32_t* counters;
};

struct SortArgs {
  int W, H, B, gx, T;
  const uint32_t* ranges;
  uint64_t* keys;
  uint64_t* keys_tmp;
  uint32_t* point_list;
  const short4* cullbox;
  uint2* sub;
  uint32_t* sub_count;
  uint2* work;          // (quadrant index, chunk) items for the acceptance pass
  uint32_t* work_count;
};

struct AcceptArgs {
  int W, H, B, gx;
  const uint32_t* ranges;
  const uint2* sub;
  const uint32_t* sub_count;
  const float4* rec;
  const uint2* work;
  const uint32_t* work_count;
  uint64_t* masks;
  uint32_t* frag_off;
  float4* frag4;
  float* frag1;
  unsigned long long* frag_fill;
  uint32_t frag_cap;
};

struct BlendArgs {
  int W, H, B, gx, gy, T;
  const uint32_t* ranges;
  const uint2* sub;
  const uint32_t* sub_count;
  const uint64_t* masks;
  const uint32_t* frag_off;
  const float4* frag4;
  const float4* rec;
  const float* bg;
  float* final_T;
  uint32_t* n_contrib;
  float* out_color;
  float* out_others;
};

struct BlendBwdArgs {
  int W, H, B, gx, gy, T;
  const uint32_t* ranges;
  const uint2* sub;
  const uint32_t* sub_count;
  const uint64_t* masks;
  const uint32_t* frag_off;
  const float4* frag4;
  const float* frag1;
  const float4* rec;
  const float* bg;
  const float* final_T;
  const uint32_t* n_contrib;
  const float* dL_dpix;
  const float* dL_dothers;
  float* dL_dtransMat;
  float* dL_dmean2D;
  float* dL_dnormal;
  float* dL_dopacity;
  float* dL_dcolors;
};

struct PreprocessBwdArgs {
  int P, D, M, W, H;
  float tanx, tany;
  const float* means3D;
  const float* scales;
  const float* rots;
  const float* shs;
  const float* tprec;
  const float4* rec;
  const int* radii;
  const uint8_t* clamped;
  const float* vm;
  const float* pm;
  const float* campos;
  float* dL_dtransMat;
  const float* dL_dnormal;
  const float* dL_dcolors;
  float* dL_dsh;
  float* dL_dmean2D;
  float* dL_dmean3D;
  float* dL_dscales;
  float* dL_drots;
};

// Launchers (return hipGetLastError()).
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
hipError_t launch_preprocess(const PreprocessArgs& a, hipStream_t s);
hipError_t launch_scan_tiles(int T, const uint32_t* count, uint32_t* start, uint32_t* fill,
                             uint32_t* ranges, uint32_t* counters, hipStream_t s);
hipError_t launch_scatter(int P, int B, int gx, int gy, const float4* rec, const float* depths,
                          const int* radii, uint32_t* tile_fill, uint64_t* keys, hipStream_t s);
hipError_t launch_sort_tiles(const SortArgs& a, hipStream_t s);
hipError_t launch_accept(const AcceptArgs& a, size_t max_items, hipStream_t s);
hipError_t launch_blend_fwd(const BlendArgs& a, hipStream_t s);
hipError_t launch_blend_bwd(const BlendBwdArgs& a, hipStream_t s);
hipError_t launch_preprocess_bwd(const PreprocessBwdArgs& a, hipStream_t s);
hipError_t launch_mark_visible(int P, const float* means3D, const float* vm, uint8_t* present,
                               hipStream_t s);

}  // namespace sr
