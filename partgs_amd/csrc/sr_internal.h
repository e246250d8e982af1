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
//   q3 = (normal.x, normal.y, normal.z, r)   q4 = (g, b, rho_reject, flip)
// Tu/Tv/Tw are the columns of the ray-splat homography T as stored by the
// reference (dsr/cuda_rasterizer/forward.cu:195-198); normal is the
// camera-facing view-space normal (forward.cu:209-214).  rho_reject bounds
// the Gaussian exponent: any pixel with min(rho3d, rho2d) > rho_reject has
// alpha < 1/255 and is skipped by the reference, so the blend kernels may
// skip it early (rho_reject < 0: the surfel never reaches 1/255).  flip is
// the sign the forward gave the normal to face the camera, times 2 when the
// surfel's preprocess ran in FP64 (near-degenerate); the backward takes the
// sign from it rather than re-deciding it.
constexpr int kRecQuads = 5;

// The blend kernels run one wave64 per 8 x 8 pixel quadrant of a tile; each
// quadrant walks an order-preserving sub-list of its tile's sorted list,
// holding only the entries whose screen cull box meets the quadrant.
constexpr int kQuad = 8;
__host__ __device__ inline int quads_per_tile(int B) { return (B / kQuad) * (B / kQuad); }
// The per-pixel blend chains run 64 / kSplit pixels of a quadrant per wave:
// a lane streams its own pairs, so spreading a heavy quadrant over more
// waves spreads its memory traffic over more in-flight requests.
#ifndef SR_FWD_SPLIT
#define SR_FWD_SPLIT 4
#endif
constexpr int kSplit = SR_FWD_SPLIT;

// A batch of views rasterised together.  Every view has its own camera and
// image size; all share the surfels and the tile edge B.  Surfel i seen from
// view v is record r = v * P + i; tile t of view v is batch tile
// tile_base + t; pixel p of view v is batch pixel pix_base + p, and a view's
// C-channel image planes start at C * pix_base.  The table travels by value
// in the kernel arguments (no device copy, nothing to upload before a launch).
constexpr int kMaxViews = 32;
struct ViewInfo {
  int W, H, gx, gy;
  uint32_t tile_base, pix_base;
  float tanx, tany;
};
struct ViewSet {
  int V, B;
  uint32_t T;  // tiles of all views
  uint32_t N;  // pixels of all views
  ViewInfo v[kMaxViews];
};
inline ViewSet make_viewset(int V, const int* W, const int* H, const float* tanx, const float* tany, int B) {
  ViewSet vs{};
  vs.V = V;
  vs.B = B;
  uint32_t t = 0, n = 0;
  for (int i = 0; i < V; ++i) {
    ViewInfo& vi = vs.v[i];
    vi.W = W[i];
    vi.H = H[i];
    vi.gx = (W[i] + B - 1) / B;
    vi.gy = (H[i] + B - 1) / B;
    vi.tile_base = t;
    vi.pix_base = n;
    vi.tanx = tanx ? tanx[i] : 0.0f;
    vi.tany = tany ? tany[i] : 0.0f;
    t += (uint32_t)(vi.gx * vi.gy);
    n += (uint32_t)(W[i] * H[i]);
  }
  vs.T = t;
  vs.N = n;
  return vs;
}
inline ViewSet single_view(int W, int H, int B, float tanx = 0.0f, float tany = 0.0f) {
  return make_viewset(1, &W, &H, &tanx, &tany, B);
}
// View owning batch tile t (t < vs.T).  The table is uniform across a
// workgroup, so the search runs on scalar registers.
__device__ inline int view_of_tile(const ViewSet& vs, uint32_t t) {
  int lo = 0, hi = vs.V - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vs.v[mid].tile_base <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Geometry state: one entry per record (surfel, view).
struct GeomLayout {
  size_t rec, depths, tiles, clamped, roff, cullbox, bsum, ssum, rslot, total;
};
// Workgroups of 256 records, per view, that preprocess and scatter use:
// at most R / 256 + one partial workgroup per view.
inline size_t record_groups(size_t R) { return R / 256 + kMaxViews + 1; }
inline GeomLayout geom_layout(size_t R) {
  GeomLayout g{};
  size_t off = 0;
  g.rec = off; off += align_up(R * kRecQuads * 16);
  g.depths = off; off += align_up(R * 4);
  g.tiles = off; off += align_up(R * 4);
  g.clamped = off; off += align_up(R * 4);  // 3 flags + pad per record
  g.roff = off; off += align_up(R * 4);     // first instance number of the record
  g.cullbox = off; off += align_up(R * 8);   // short4 inclusive pixel box
  g.bsum = off; off += align_up(record_groups(R) * 4);  // instances per workgroup, then their scan
  g.ssum = off; off += align_up(record_groups(R) * 4);  // partial-gradient slots per workgroup, then their scan
  g.rslot = off; off += align_up(R * 8);  // uint2 per record: first slot, slots
  g.total = off;
  return g;
}

// Image state: per-tile binning tables and per-pixel blend state of all views.
struct ImageLayout {
  int T, N;
  size_t counters, tile_count, tile_start, tile_fill, ranges, large, lchunk_off, sub_count, q_flag, lane_range, q_cost,
      q_order, final_T, n_contrib, total;
};
inline ImageLayout image_layout(const ViewSet& vs) {
  ImageLayout l{};
  l.T = (int)vs.T;
  l.N = (int)vs.N;
  size_t off = 0;
  l.counters = off; off += kAlign;  // [0] instances, [1] error flags, [2] work items,
                                    // [3] large tiles, [4..5] pair-store fill,
                                    // [6] chunks of the large tiles, [7] huge tiles,
                                    // [8] partial-gradient slots
  l.tile_count = off; off += align_up((size_t)l.T * 4);
  l.tile_start = off; off += align_up(((size_t)l.T + 1) * 4);
  l.tile_fill = off; off += align_up((size_t)l.T * 4);
  l.ranges = off; off += align_up((size_t)l.T * 8);
  l.large = off; off += align_up((size_t)l.T * 4);  // tiles for the large-tile sort
  l.lchunk_off = off; off += align_up(((size_t)l.T + 1) * 4);  // their first chunk numbers
  l.sub_count = off; off += align_up((size_t)l.T * quads_per_tile(vs.B) * 4);
  l.q_flag = off; off += align_up((size_t)l.T * quads_per_tile(vs.B) * 4);  // 1: the quadrant missed the pair store
  l.lane_range = off; off += align_up((size_t)l.T * quads_per_tile(vs.B) * 64 * 8);  // (first pair, pairs) per pixel
  l.q_cost = off; off += align_up((size_t)l.T * quads_per_tile(vs.B) * 4);   // longest pixel chain
  l.q_order = off; off += align_up((size_t)l.T * quads_per_tile(vs.B) * 4);  // quadrants, longest chain first
  l.final_T = off; off += align_up((size_t)l.N * 3 * 4);
  l.n_contrib = off; off += align_up((size_t)l.N * 2 * 4);
  l.total = off;
  return l;
}
inline ImageLayout image_layout(int W, int H, int B) { return image_layout(single_view(W, H, B)); }

// Binning state: one entry per (surfel, tile) instance.  Instances are
// numbered in record order: record r's k-th tile (row-major in its tile
// rectangle) is instance roff[r] + k.
//   inst_rs    instance -> (record, its first partial-gradient slot)
//   pos_slot   position in the concatenated sorted tile lists -> the first
//              slot of the instance there
//   pos_quads  position -> the quadrants whose sub-lists hold the instance
//   lslot      large-tile sort: per entry of a sorted chunk, its first slot
//              (the merge moves it to pos_slot)
//   part       one slot per sub-list entry, i.e. per (instance, quadrant
//              holding it): the 18 per-surfel partial gradients the
//              backward blend summed over that quadrant's pixels
//              (kPartStride floats).  Slots are numbered in record order,
//              then instance order, then quadrant order, so a record's
//              slots are one contiguous run (rslot) that the backward
//              preprocess sums in order: no gradient needs a global
//              atomic.  The slot array is the layout's last region, so its
//              size (slot capacity) does not move any other offset.
struct BinLayout {
  size_t keys, lmeta, lpre, point_list, sub, masks, chunk_T, work, frag_off, total;
  size_t inst_rs, pos_slot, pos_quads, lslot, part, part_nz;
  size_t nslot;  // slots the part region holds
  size_t prec, fW;
  size_t frag_cap;
  size_t sem_part;
};
constexpr int kPartStride = 20;  // 18 partial gradients, padded to five float4
// Accepted-pair store: per accepted (pixel, entry) pair the pair-write pass
// keeps one 48-byte record with everything the per-pixel blend chains read,
// so they stream it without staging records:
//   p0 = (depth, m, m^2, alpha)   m = distortion weight of the depth
//   p1 = (normal.xyz, r)
//   p2 = (g, b, position of the entry in the tile list (bits), dm/ddepth)
// (p0, p1 and p2 are what the forward blend's lanes of roles 0, 1 and 2
// accumulate, in their accumulators' order)
// A lane's consecutive pairs are consecutive records, so the fill writes
// whole lines.  (The per-pair gradient pass recomputes s and G.)  fW = (w, dL/dalpha, dL/ddepth, 0) is written by the backward chain
// pass for the per-pair gradient pass, which recomputes the rest of the pair.  All pairs of one pixel of a quadrant are
// contiguous and in entry order: lane_range holds (first, count) per pixel,
// frag_off the first pair of each (chunk, pixel).  The forward blend
// rewrites each count to the pairs the pixel actually blended (a prefix of
// its run: the ones in front of its last contributor), so the backward
// passes never touch the pairs behind the saturation point.  Capacity
// kFragPerInstance pairs per tile instance; a quadrant that does not fit is
// flagged (q_flag, frag_off = kNoFrag) and blended by the staged path that
// recomputes its pairs.
constexpr size_t kFragPerInstance = 24;
constexpr int kPairQuads = 3;  // float4 per stored pair
constexpr uint32_t kNoFrag = 0xFFFFFFFFu;
// Per-pixel acceptance masks: quadrant (t, q) owns mask slots
// [mask_base, mask_base + 64 * chunks), chunk c of its sub-list at
// mask_base + 64 c, one 64-bit word per pixel (bit j: entry 64 c + j reaches
// alpha >= 1/255 at that pixel).
__host__ __device__ inline size_t mask_base(int Q, uint32_t beg, uint32_t n, int tile, int q) {
  return (size_t)Q * beg + (size_t)q * n + (size_t)kQuad * kQuad * ((size_t)tile * Q + q);
}
// (quadrant, chunk) work items, appended by the tile sorts as they finish;
// the kernels that walk the list take it from the end, so the items of the
// large tiles (appended last, by the chunk merge and the huge-tile sort)
// start first (measured: 4 us less per 300k-surfel 800x600 step, in the
// passes that follow).
inline size_t work_capacity(int L, int B, int T) {
  const size_t Q = (size_t)quads_per_tile(B);
  return Q * ((size_t)T + ((size_t)(L > 0 ? L : 1) + 63) / 64);
}
// slot_capacity: slots the part region is sized for; 0 sizes it for every
// (instance, quadrant), which no batch can exceed.
inline BinLayout bin_layout(int L, int B, int T, int pair_store = 0, int pair_capacity = 0, int sem = 0,
                            int slot_capacity = 0) {
  BinLayout b{};
  size_t off = 0;
  const size_t n = (size_t)(L > 0 ? L : 1);
  const size_t Q = (size_t)quads_per_tile(B);
  b.keys = off; off += align_up(n * 8);
  // large-tile sort: per entry of a sorted 2048-key chunk, (record, quadrant
  // mask) and, per quadrant, the chunk's hits before it (16-bit each)
  b.lmeta = off; off += align_up(n * 8);
  b.lpre = off; off += align_up(n * Q * 2);
  b.point_list = off; off += align_up(n * 4);
  b.sub = off; off += align_up(n * Q * 8);  // uint2 (surfel id, position in tile)
  b.masks = off; off += align_up((n * Q + (size_t)kQuad * kQuad * Q * (size_t)T) * 8);
  const size_t nslots = n * Q + (size_t)kQuad * kQuad * Q * (size_t)T;
  b.chunk_T = off; off += align_up(nslots * 4);  // per mask slot: product of (1 - alpha) over its pairs
  b.work = off; off += align_up(work_capacity(L, B, T) * 8);
  b.frag_off = off; off += align_up(nslots * 4);
  b.inst_rs = off; off += align_up(n * 8);
  b.pos_slot = off; off += align_up(n * 4);
  b.pos_quads = off; off += align_up(n * 2);
  b.lslot = off; off += align_up(n * 4);
  b.nslot = slot_capacity > 0 ? (size_t)slot_capacity : n * Q;
  b.part_nz = off; off += align_up(n * Q);  // per slot: 1 when the slot holds sums
  b.frag_cap = pair_capacity > 0 ? (size_t)pair_capacity
                                  : n * (pair_store > 0 ? (size_t)pair_store : kFragPerInstance);
  if (b.frag_cap > 0x7fffffffu) b.frag_cap = 0x7fffffffu;
  b.prec = off; off += align_up(b.frag_cap * kPairQuads * 16);
  b.fW = off; off += align_up(b.frag_cap * 16);
  b.sem_part = off; off += align_up(n * Q * (size_t)sem * 4);  // _part: semantic partials per slot
  b.part = off; off += align_up(b.nslot * kPartStride * 4);         // last: see part
  b.total = off;
  return b;
}

// Device pointers of the accepted-pair store.
struct PairStore {
  float4* rec;  // kPairQuads per pair
  float4* fW;   // 1 per pair
};

constexpr uint32_t kErrPrefiltered = 1u;
constexpr uint32_t kErrCapacity = 2u;  // more tile instances (or slots) than the binning capacity
constexpr uint32_t kErrSlots = 4u;     // with kErrCapacity: the partial-gradient slots did not fit

struct PreprocessArgs {
  ViewSet vs;
  int P, D, M, prefiltered;
  int part;  // the _part variant's homography and AABB
  float mod;
  const float* means3D;
  const float* scales;
  const float* rots;
  const float* opac;
  const float* shs;
  const float* colors;
  const float* tprec;
  const float* vm;      // V x 16
  const float* pm;      // V x 16
  const float* campos;  // V x 3
  float4* rec;
  float* depths;
  int* radii;
  uint32_t* tiles;
  uint8_t* clamped;
  short4* cullbox;
  uint32_t* bsum;
  uint32_t* ssum;  // partial-gradient slots per workgroup
  uint32_t* tile_count;
  uint32_t* counters;
};

struct ScatterArgs {
  ViewSet vs;
  int P;
  const float4* rec;
  const float* depths;
  const int* radii;
  const uint32_t* counters;
  const uint32_t* bsum;
  const uint32_t* ssum;  // first slot of each workgroup
  const short4* cullbox;
  uint32_t* roff;
  uint2* rslot;    // per record: first slot, slots
  uint2* inst_rs;  // per instance: record, first slot
  uint32_t* tile_fill;
  uint64_t* keys;
};

// Tiles with at most kSmallMax instances are sorted by one 256-lane
// workgroup each; up to kLargeMax in kSmallMax-key chunks that are then
// merged by rank; larger ("huge") ones by one 1024-lane workgroup each.
constexpr int kSmallMax = 2048;
constexpr int kLargeMax = 16384;

struct SortArgs {
  ViewSet vs;
  const uint32_t* ranges;
  const uint32_t* counters;
  const uint32_t* large;  // counters[3] tiles of kSmallMax < n <= large_max from the start, then
                          // counters[7] tiles of n > large_max from the end (large[T - 1 - h]);
                          // large_max is kLargeMax, or kSmallMax for batches (no chunk path)
  const uint32_t* lchunk_off;  // per large tile its first kSmallMax-key chunk (counters[6] in all)
  int side_huge;               // host: sort the huge tiles on a second stream beside the rest
  uint64_t* keys;
  uint2* lmeta;
  uint16_t* lpre;
  uint32_t* point_list;
  const short4* cullbox;
  const uint2* inst_rs;
  uint32_t* pos_slot;
  uint16_t* pos_quads;
  uint32_t* lslot;
  uint2* sub;
  uint32_t* sub_count;
  uint2* work;          // (quadrant index, chunk) items for the acceptance pass
  uint32_t* work_count;
  uint32_t* q_flag;
};

struct AcceptArgs {
  ViewSet vs;
  const uint32_t* ranges;
  const uint32_t* counters;
  const uint2* sub;
  const uint32_t* sub_count;
  const float4* rec;
  const short4* cullbox;  // per record: inclusive pixel box holding every pixel with alpha >= 1/255
  const uint2* work;
  const uint32_t* work_count;
  uint64_t* masks;
  float* chunk_T;
  uint32_t* frag_off;
  uint2* lane_range;
  uint32_t* q_cost;
  PairStore ps;
  uint32_t* q_flag;
  unsigned long long* frag_fill;
  uint32_t frag_cap;
};

struct BlendArgs {
  ViewSet vs;
  const uint32_t* ranges;
  const uint32_t* counters;
  const uint2* sub;
  const uint32_t* sub_count;
  const uint64_t* masks;
  const uint32_t* frag_off;
  uint2* lane_range;  // count rewritten to the pairs blended
  const uint32_t* q_order;
  PairStore ps;
  const uint32_t* q_flag;
  const float4* rec;
  const float* bg;
  float* final_T;
  uint32_t* n_contrib;
  float* out_color;
  float* out_others;
};

// Backward blend.  Each (instance, quadrant) entry receives its 18
// partial gradients summed over the quadrant's pixels in part: slots 0..8
// dL_dtransMat, 9..10 dL_dmean2D.xy, 11..13 dL_dnormal, 14 dL_dopacity,
// 15..17 dL_dcolors.
struct BlendBwdArgs {
  ViewSet vs;
  const uint32_t* ranges;
  const uint32_t* counters;
  const uint2* sub;
  const uint32_t* sub_count;
  const uint64_t* masks;
  const uint32_t* frag_off;
  const uint2* lane_range;
  const uint32_t* q_order;
  PairStore ps;
  const uint32_t* q_flag;
  const uint2* work;  // (quadrant, chunk) items of the acceptance pass
  const uint32_t* work_count;
  const float4* rec;
  const float* bg;
  const float* final_T;
  const uint32_t* n_contrib;
  const float* dL_dpix;
  const float* dL_dothers;
  const uint32_t* pos_slot;    // slot_of
  const uint16_t* pos_quads;
  float* part;        // partial gradients per slot (slot_of)
  uint8_t* part_nz;   // 1 where that slot was stored (any non-zero sum)
};

// Backward preprocess over all views of each surfel.  It first reduces each
// visible record's partial gradients (part, over the quadrants of its
// instances) and writes the per-record outputs dL_dtransMat, dL_dmean2D,
// dL_dopacity and dL_dcolors; dL_dopacity_out / dL_dcolors_out receive the
// sums over views (skipped when they alias the per-record arrays, V = 1).
struct PreprocessBwdArgs {
  ViewSet vs;
  int P, D, M;
  const uint32_t* counters;
  const uint32_t* tiles;
  const uint32_t* roff;
  const uint2* rslot;  // per record: first slot, slots
  const float* part;
  const uint8_t* part_nz;
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
  float* dL_dcolors;
  float* dL_dopacity;
  float* dL_dopacity_out;
  float* dL_dcolors_out;
  float* dL_dsh;
  float* dL_dmean2D;
  float* dL_dmean3D;
  float* dL_dscales;
  float* dL_drots;
  // the _part variant (one view): semantic partials in, their sums out
  int S;
  const float* sem_part;
  float* dL_dsem;
};

// The _part variant's blend passes (one view): the plain pipeline's
// records, ranges and quadrant sub-lists, plus semantics.
struct PartArgs {
  ViewSet vs;
  int S;                     // semantic channels, <= 16
  const uint32_t* ranges;
  const uint2* sub;
  const uint32_t* sub_count;
  const float4* rec;
  const float* sem;          // P x S
  const float* bg;
  // forward outputs
  float* final_T;
  uint32_t* n_contrib;
  float* out_color;
  float* out_sem;            // S x H x W
  float* out_others;         // 8 x H x W
  // backward
  const float* final_T_in;
  const uint32_t* n_contrib_in;
  const float* dL_dpix;
  const float* dL_dsem;
  const float* dL_dothers;
  const uint32_t* pos_slot;    // slot_of
  const uint16_t* pos_quads;
  float* part;
  float* sem_part;           // S per slot
  uint8_t* part_nz;
};

constexpr int kMaxQuads = 16;

// Bit q set when cull box bx meets quadrant q (of QX x QX) of tile (tx, ty)
// inside the image; the empty box (32767, 32767, -32768, -32768) meets none.
__device__ inline uint32_t quad_hits(short4 bx, int tx, int ty, int B, int QX, int Q, const ViewInfo& vi) {
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < kMaxQuads; ++q) {
    if (q < Q) {
      const int qx0 = tx * B + (q % QX) * kQuad, qy0 = ty * B + (q / QX) * kQuad;
      const bool hit = qx0 < vi.W && qy0 < vi.H && bx.x <= qx0 + kQuad - 1 && bx.z >= qx0 &&
                       bx.y <= qy0 + kQuad - 1 && bx.w >= qy0;
      m |= (uint32_t)hit << q;
    }
  }
  return m;
}

// quad_hits factorises: quadrant (qx, qy) of tile (tx, ty) is hit when its
// column c = tx * QX + qx meets the box's x span and its row meets the y
// span.  Number of kQuad-wide columns c in [c0, c1) inside the image
// (c * kQuad < W) that meet the pixel span [b0, b1]; with the y span, rows
// and H it counts rows.  popc(quad_hits) over any rectangle of tiles is the
// product of its column and row counts.
__host__ __device__ inline int quad_span(int b0, int b1, int c0, int c1, int W) {
  const int lo0 = b0 > 0 ? b0 / kQuad : 0;
  const int lo = lo0 > c0 ? lo0 : c0;
  const int hi0 = b1 >= 0 ? b1 / kQuad : -1;
  const int hi1 = (W - 1) / kQuad;
  int hi = c1 - 1;
  hi = hi0 < hi ? hi0 : hi;
  hi = hi1 < hi ? hi1 : hi;
  return hi >= lo ? hi - lo + 1 : 0;
}

// Partial-gradient slot of an instance's entry in sub-list q of its tile:
// the instance's first slot plus the quadrants before q that hold it.
__device__ inline size_t slot_of(const uint32_t* pos_slot, const uint16_t* pos_quads, uint32_t pos, int q) {
  return (size_t)pos_slot[pos] + (size_t)__popc((uint32_t)pos_quads[pos] & ((1u << q) - 1u));
}

// Launchers (return hipGetLastError()).
hipError_t launch_part_blend_fwd(const PartArgs& a, hipStream_t s);
hipError_t launch_part_blend_bwd(const PartArgs& a, hipStream_t s);

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
// Sets the message sr_last_error() returns on the calling thread; returns code.
int set_error(int code, const char* msg);
hipError_t launch_preprocess(const PreprocessArgs& a, hipStream_t s);
hipError_t launch_scan_tiles(int T, const uint32_t* count, uint32_t* start, uint32_t* fill,
                             uint32_t* ranges, uint32_t* counters, uint32_t capacity, int nb, uint32_t* bsum,
                             uint32_t* ssum, uint32_t slot_capacity, uint32_t* large, uint32_t* lchunk_off,
                             int large_max, hipStream_t s);
hipError_t launch_scatter(const ScatterArgs& a, hipStream_t s);
hipError_t launch_sort_tiles(const SortArgs& a, hipStream_t s);
hipError_t launch_accept(const AcceptArgs& a, size_t max_items, hipStream_t s);
hipError_t launch_pair_offsets(const AcceptArgs& a, hipStream_t s);
hipError_t launch_pair_write(const AcceptArgs& a, size_t max_items, hipStream_t s);
hipError_t launch_order_quads(int nq, const uint32_t* cost, uint32_t* order, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* p, size_t n, hipStream_t s);
hipError_t launch_blend_fwd(const BlendArgs& a, hipStream_t s);
hipError_t launch_blend_bwd(const BlendBwdArgs& a, hipStream_t s);
hipError_t launch_blend_bwd_grad(const BlendBwdArgs& a, size_t max_items, hipStream_t s);
hipError_t launch_preprocess_bwd(const PreprocessBwdArgs& a, hipStream_t s);
hipError_t launch_part_preprocess_bwd(const PreprocessBwdArgs& a, hipStream_t s);
hipError_t launch_mark_visible(int P, const float* means3D, const float* vm, uint8_t* present,
                               hipStream_t s);
hipError_t launch_copy_status(const uint32_t* counters, uint32_t* status, hipStream_t s);

}  // namespace sr
