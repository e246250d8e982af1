// Host orchestration and the extern "C" entry points declared in
// include/surfel_rast.h.  Mirrors CudaRasterizer::Rasterizer::{forward,
// backward, markVisible} (dsr/cuda_rasterizer/rasterizer_impl.cu:141-448),
// with every launch on the caller's stream.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/surfel_rast.h"
#include "sr_internal.h"

using namespace sr;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

// Sets the calling thread's error message (for entry points in other files).
int sr::set_error(int code, const char* msg) {
  g_err = msg ? msg : "";
  return code;
}

namespace {

#define SR_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---- per-stage timing ------------------------------------------------------
struct Timing {
  std::mutex mu;
  uint32_t mask = 0;  // bit i: stage i is timed
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double ms[SR_NUM_STAGES] = {0};
  int count[SR_NUM_STAGES] = {0};
};
Timing& timing() {
  static Timing t;
  return t;
}

hipEvent_t take_event(Timing& t) {
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct StageScope {
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  StageScope(int st, hipStream_t str) : stage(st), s(str) {
    Timing& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    if (stage < 0 || stage >= SR_NUM_STAGES || !((t.mask >> stage) & 1u)) return;
    a = take_event(t);
    b = take_event(t);
    if (a && b) (void)hipEventRecord(a, s);
  }
  ~StageScope() {
    if (!a || !b) return;
    (void)hipEventRecord(b, s);
    Timing& t = timing();
    std::lock_guard<std::mutex> lk(t.mu);
    t.pending.push_back({stage, a, b});
  }
};

int check_stage(bool debug, hipStream_t s, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SR_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
  if (debug) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  return SR_OK;
}

#define SR_STAGE(st, dbg, stream, name, launch)                 \
  do {                                                          \
    hipError_t le_;                                             \
    {                                                           \
      StageScope sc_(st, stream);                               \
      le_ = (launch);                                           \
    }                                                           \
    if (le_ != hipSuccess)                                      \
      return fail(SR_ERR_HIP, "%s: %s", name, hipGetErrorString(le_)); \
    int rc_ = check_stage(dbg, stream, name);                   \
    if (rc_ != SR_OK) return rc_;                               \
  } while (0)

// An event per (host thread, device), marking the instance count's copy to
// the host (an event is recorded only on streams of the device it was
// created on).
hipEvent_t counts_event() {
  thread_local hipEvent_t cache[16] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return nullptr;
  hipEvent_t& e = cache[dev];
  if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
  return e;
}
// Pinned 64-byte landing zone for the forward counters (instances, error flags, slots).
// Pinned 8-byte landing zone for the instance count and error flags.
uint32_t* pinned_counters() {
  thread_local uint32_t* p = nullptr;
  if (!p) {
    if (hipHostMalloc((void**)&p, 16 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

int validate(const sr_settings* s, const sr_surfels* g) {
  if (!s || !g) return fail(SR_ERR_INVALID, "settings and surfels must not be NULL");
  if (g->P < 0) return fail(SR_ERR_INVALID, "P must be >= 0");
  if (s->image_width <= 0 || s->image_height <= 0 || s->image_width > 32000 || s->image_height > 32000)
    return fail(SR_ERR_INVALID, "image size must be in [1, 32000] (got %d x %d)", s->image_width, s->image_height);
  if (s->pair_store < 0) return fail(SR_ERR_INVALID, "pair_store must be >= 0");
  if (s->pair_capacity < 0) return fail(SR_ERR_INVALID, "pair_capacity must be >= 0");
  if (s->slot_capacity < 0) return fail(SR_ERR_INVALID, "slot_capacity must be >= 0");
  if (s->tile != 8 && s->tile != 16 && s->tile != 32)
    return fail(SR_ERR_INVALID, "tile must be 8, 16 or 32 (got %d)", s->tile);
  if (g->P == 0) return SR_OK;  // the reference glue returns before looking at the inputs
  if ((g->shs == nullptr) == (g->colors_precomp == nullptr))
    return fail(SR_ERR_INVALID, "Please provide excatly one of either SHs or precomputed colors!");
  const bool sr = g->scales != nullptr && g->rotations != nullptr;
  const bool any_sr = g->scales != nullptr || g->rotations != nullptr;
  if ((!sr && g->transMat_precomp == nullptr) || (any_sr && g->transMat_precomp != nullptr))
    return fail(SR_ERR_INVALID,
                "Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!");
  if (g->P > 0 && !g->means3D) return fail(SR_ERR_INVALID, "means3D required");
  if (!s->bg || !s->viewmatrix || !s->projmatrix || !s->campos)
    return fail(SR_ERR_INVALID, "bg, viewmatrix, projmatrix and campos are required");
  if (g->shs && g->M <= 0) return fail(SR_ERR_INVALID, "shs given but M = %d", g->M);
  if (g->shs && s->sh_degree > 3) return fail(SR_ERR_INVALID, "sh_degree must be <= 3");
  if (g->shs && (s->sh_degree + 1) * (s->sh_degree + 1) > g->M)
    return fail(SR_ERR_INVALID, "sh_degree %d needs %d coefficients, M = %d", s->sh_degree,
                (s->sh_degree + 1) * (s->sh_degree + 1), g->M);
  return SR_OK;
}


// ---- pipeline stages shared by the single-view and batched entry points ----
struct Pipeline {
  ViewSet vs;
  int P = 0, s_D = 0, prefiltered = 0, pair_store = 0, pair_capacity = 0;
  float mod = 1.0f;
  const float* vm = nullptr;
  const float* pm = nullptr;
  const float* campos = nullptr;
  const float* bg = nullptr;
  char* geom = nullptr;
  char* img = nullptr;
  char* bin = nullptr;
  GeomLayout gl{};
  ImageLayout il{};
  BinLayout bl{};
  int capacity = 0;
  bool side_huge = false;  // huge-tile sort on a second stream (graph-captured batches)
  bool part = false;       // the _part variant
  int S = 0;               // its semantic channels
  uint32_t* counters() const { return (uint32_t*)(img + il.counters); }
};

PairStore pair_store(char* bin, const BinLayout& bl) {
  PairStore p{};
  p.rec = (float4*)(bin + bl.prec);
  p.fW = (float4*)(bin + bl.fW);
  return p;
}

// Pinned landing zones and events for the accepted-pair count of recent
// forwards, per (host thread, device): a ring of kPairSlots, so enqueuing a
// forward only ever waits for the copy of the forward kPairSlots before it
// (long finished), never for the one still running.
constexpr int kPairSlots = 4;
constexpr int kCountWords = 16;
struct PairCount {
  uint32_t* host = nullptr;  // kPairSlots x kCountWords words: counters[0..8] of a forward
  hipEvent_t ev[kPairSlots] = {};
  unsigned long long seq[kPairSlots] = {};  // 0 = slot never used
  unsigned long long next_seq = 1;
  int next = 0;
};
PairCount& pair_count_slot() {
  thread_local PairCount slots[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
  return slots[dev];
}

// Enqueues the copy of the instance, pair and slot counts (counters[0..8])
// of the forward just enqueued into the next slot of the ring.
int post_pair_count(Pipeline& pl, hipStream_t st) {
  PairCount& pc = pair_count_slot();
  if (!pc.host) {
    if (hipHostMalloc((void**)&pc.host, kPairSlots * kCountWords * sizeof(uint32_t), hipHostMallocDefault) !=
        hipSuccess) {
      pc.host = nullptr;
      return fail(SR_ERR_HIP, "hipHostMalloc failed");
    }
    for (int k = 0; k < kPairSlots; ++k)
      if (hipEventCreateWithFlags(&pc.ev[k], hipEventDisableTiming) != hipSuccess)
        return fail(SR_ERR_HIP, "hipEventCreate failed");
  }
  const int k = pc.next;
  if (pc.seq[k] != 0) SR_HIP(hipEventSynchronize(pc.ev[k]));  // the copy of kPairSlots forwards ago
  SR_HIP(hipMemcpyAsync(pc.host + kCountWords * k, pl.counters(), 9 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SR_HIP(hipEventRecord(pc.ev[k], st));
  pc.seq[k] = pc.next_seq++;
  pc.next = (k + 1) % kPairSlots;
  return SR_OK;
}

// Per-record gradient outputs of the backward preprocess (record r = v P + i):
// dL_dtransMat 9, dL_dmean2D 3, dL_dopacity 1, dL_dcolors 3 floats each.
struct RecordGrads {
  float* transMat;
  float* mean2D;
  float* opacity;
  float* colors;
};

// Workspace of the batched entry points: geometry, image and binning state
// plus the per-record gradient scratch, each section 256-byte aligned.
struct BatchLayout {
  size_t geom, img, bin, g_transMat, g_opacity, g_colors, total;
  GeomLayout gl;
  ImageLayout il;
  BinLayout bl;
};
BatchLayout batch_layout(const ViewSet& vs, int P, int capacity, int pair_store, int pair_capacity,
                         int slot_capacity) {
  BatchLayout b{};
  const size_t R = (size_t)vs.V * (size_t)(P > 0 ? P : 0);
  b.gl = geom_layout(R);
  b.il = image_layout(vs);
  b.bl = bin_layout(capacity, vs.B, (int)vs.T, pair_store, pair_capacity, 0, slot_capacity);
  size_t off = 0;
  b.geom = off; off += align_up(b.gl.total);
  b.img = off; off += align_up(b.il.total);
  b.bin = off; off += align_up(b.bl.total);
  b.g_transMat = off; off += align_up(R * 9 * 4);
  b.g_opacity = off; off += align_up(R * 4);
  b.g_colors = off; off += align_up(R * 3 * 4);
  b.total = off;
  return b;
}

int validate_views(const sr_views* v, const sr_surfels* g, int capacity) {
  if (!v || !g) return fail(SR_ERR_INVALID, "views and surfels must not be NULL");
  if (v->V < 1 || v->V > SR_MAX_VIEWS) return fail(SR_ERR_INVALID, "V must be in [1, %d]", SR_MAX_VIEWS);
  if (!v->widths || !v->heights || !v->tanfovx || !v->tanfovy)
    return fail(SR_ERR_INVALID, "image sizes and fields of view are required");
  if (!v->viewmatrices || !v->projmatrices || !v->campos || !v->bg)
    return fail(SR_ERR_INVALID, "viewmatrices, projmatrices, campos and bg are required");
  if (v->tile != 8 && v->tile != 16 && v->tile != 32)
    return fail(SR_ERR_INVALID, "tile must be 8, 16 or 32 (got %d)", v->tile);
  if (v->pair_store < 0) return fail(SR_ERR_INVALID, "pair_store must be >= 0");
  if (v->pair_capacity < 0) return fail(SR_ERR_INVALID, "pair_capacity must be >= 0");
  if (v->slot_capacity < 0) return fail(SR_ERR_INVALID, "slot_capacity must be >= 0");
  if (capacity < 1) return fail(SR_ERR_INVALID, "capacity must be >= 1");
  size_t pixels = 0;
  for (int i = 0; i < v->V; ++i) {
    if (v->widths[i] <= 0 || v->heights[i] <= 0 || v->widths[i] > 32000 || v->heights[i] > 32000)
      return fail(SR_ERR_INVALID, "view %d: image size must be in [1, 32000]", i);
    pixels += (size_t)v->widths[i] * v->heights[i];
  }
  if (pixels > 0x7fffffffu) return fail(SR_ERR_INVALID, "too many pixels in one batch");
  if (g->P < 1) return fail(SR_ERR_INVALID, "P must be >= 1");
  if ((size_t)g->P * v->V > 0x7fffffffu) return fail(SR_ERR_INVALID, "P x V must stay below 2^31");
  if ((g->shs == nullptr) == (g->colors_precomp == nullptr))
    return fail(SR_ERR_INVALID, "Please provide excatly one of either SHs or precomputed colors!");
  if (!g->scales || !g->rotations || g->transMat_precomp)
    return fail(SR_ERR_INVALID, "batched views need scales and rotations (a precomputed T is per view)");
  if (!g->means3D || !g->opacities) return fail(SR_ERR_INVALID, "means3D and opacities required");
  if (g->shs && (g->M <= 0 || v->sh_degree > 3 || (v->sh_degree + 1) * (v->sh_degree + 1) > g->M))
    return fail(SR_ERR_INVALID, "sh_degree %d does not fit M = %d", v->sh_degree, g->M);
  return SR_OK;
}

Pipeline batch_pipeline(const sr_views* v, const sr_surfels* g, const ViewSet& vs, const BatchLayout& bl,
                        char* ws, int capacity) {
  Pipeline pl{};
  pl.vs = vs;
  pl.P = g->P;
  pl.s_D = v->sh_degree;
  pl.prefiltered = 0;
  pl.mod = v->scale_modifier;
  pl.pair_store = v->pair_store;
  pl.pair_capacity = v->pair_capacity;
  pl.vm = v->viewmatrices;
  pl.pm = v->projmatrices;
  pl.campos = v->campos;
  pl.bg = v->bg;
  pl.geom = ws + bl.geom;
  pl.img = ws + bl.img;
  pl.bin = ws + bl.bin;
  pl.gl = bl.gl;
  pl.il = bl.il;
  pl.bl = bl.bl;
  pl.capacity = capacity;
  return pl;
}

// Clear counters and histogram, preprocess every (surfel, view), scan the
// tile histogram; the instance count lands in counters[0].
int run_front(Pipeline& pl, const sr_surfels* g, int* radii, uint32_t capacity, uint32_t slot_capacity, bool dbg,
              hipStream_t st) {
  uint32_t* counters = pl.counters();
  uint32_t* tile_count = (uint32_t*)(pl.img + pl.il.tile_count);
  // the counters and the tile histogram are adjacent: one clear for both
  static_assert(kAlign % 4 == 0, "counters region in whole words");
  if (tile_count != counters + kAlign / 4) return fail(SR_ERR_INVALID, "image layout: histogram not after counters");
  SR_STAGE(1, dbg, st, "clear", launch_fill_u32(counters, 0u, kAlign / 4 + (size_t)pl.il.T, st));
  PreprocessArgs pa{};
  pa.vs = pl.vs;
  pa.P = pl.P;
  pa.D = pl.s_D;
  pa.M = g->M;
  pa.prefiltered = pl.prefiltered;
  pa.part = pl.part ? 1 : 0;
  pa.mod = pl.mod;
  pa.means3D = g->means3D;
  pa.scales = g->scales;
  pa.rots = g->rotations;
  pa.opac = g->opacities;
  pa.shs = g->shs;
  pa.colors = g->colors_precomp;
  pa.tprec = g->transMat_precomp;
  pa.vm = pl.vm;
  pa.pm = pl.pm;
  pa.campos = pl.campos;
  pa.rec = (float4*)(pl.geom + pl.gl.rec);
  pa.depths = (float*)(pl.geom + pl.gl.depths);
  pa.radii = radii;
  pa.tiles = (uint32_t*)(pl.geom + pl.gl.tiles);
  pa.clamped = (uint8_t*)(pl.geom + pl.gl.clamped);
  pa.cullbox = (short4*)(pl.geom + pl.gl.cullbox);
  pa.bsum = (uint32_t*)(pl.geom + pl.gl.bsum);
  pa.ssum = (uint32_t*)(pl.geom + pl.gl.ssum);
  pa.tile_count = tile_count;
  pa.counters = counters;
  SR_STAGE(0, dbg, st, "preprocess", launch_preprocess(pa, st));
  SR_STAGE(1, dbg, st, "scan",
           launch_scan_tiles(pl.il.T, tile_count, (uint32_t*)(pl.img + pl.il.tile_start),
                             (uint32_t*)(pl.img + pl.il.tile_fill), (uint32_t*)(pl.img + pl.il.ranges), counters,
                             capacity, pl.vs.V * ((pl.P + 255) / 256), (uint32_t*)(pl.geom + pl.gl.bsum),
                             (uint32_t*)(pl.geom + pl.gl.ssum), slot_capacity, (uint32_t*)(pl.img + pl.il.large), (uint32_t*)(pl.img + pl.il.lchunk_off),
                             pl.side_huge ? kSmallMax : kLargeMax, st));
  return SR_OK;
}

// Scatter into tile buckets, sort every tile, split into quadrant sub-lists,
// then the acceptance pass over all (quadrant, chunk) work items.
int run_binning(Pipeline& pl, const int* radii, bool dbg, hipStream_t st, bool pairs = true) {
  uint32_t* counters = pl.counters();
  char* bin = pl.bin;
  const BinLayout& bl = pl.bl;
  ScatterArgs sc{};
  sc.vs = pl.vs;
  sc.P = pl.P;
  sc.rec = (const float4*)(pl.geom + pl.gl.rec);
  sc.depths = (const float*)(pl.geom + pl.gl.depths);
  sc.radii = radii;
  sc.counters = counters;
  sc.bsum = (const uint32_t*)(pl.geom + pl.gl.bsum);
  sc.ssum = (const uint32_t*)(pl.geom + pl.gl.ssum);
  sc.cullbox = (const short4*)(pl.geom + pl.gl.cullbox);
  sc.roff = (uint32_t*)(pl.geom + pl.gl.roff);
  sc.rslot = (uint2*)(pl.geom + pl.gl.rslot);
  sc.inst_rs = (uint2*)(bin + bl.inst_rs);
  sc.tile_fill = (uint32_t*)(pl.img + pl.il.tile_fill);
  sc.keys = (uint64_t*)(bin + bl.keys);
  SR_STAGE(2, dbg, st, "scatter", launch_scatter(sc, st));
  SortArgs sa{};
  sa.vs = pl.vs;
  sa.ranges = (const uint32_t*)(pl.img + pl.il.ranges);
  sa.counters = counters;
  sa.large = (const uint32_t*)(pl.img + pl.il.large);
  sa.keys = (uint64_t*)(bin + bl.keys);
  sa.lmeta = (uint2*)(bin + bl.lmeta);
  sa.lpre = (uint16_t*)(bin + bl.lpre);
  sa.lchunk_off = (const uint32_t*)(pl.img + pl.il.lchunk_off);
  sa.side_huge = pl.side_huge ? 1 : 0;
  sa.point_list = (uint32_t*)(bin + bl.point_list);
  sa.cullbox = (const short4*)(pl.geom + pl.gl.cullbox);
  sa.inst_rs = (const uint2*)(bin + bl.inst_rs);
  sa.pos_slot = (uint32_t*)(bin + bl.pos_slot);
  sa.pos_quads = (uint16_t*)(bin + bl.pos_quads);
  sa.lslot = (uint32_t*)(bin + bl.lslot);
  sa.sub = (uint2*)(bin + bl.sub);
  sa.sub_count = (uint32_t*)(pl.img + pl.il.sub_count);
  sa.work = (uint2*)(bin + bl.work);
  sa.work_count = counters + 2;
  sa.q_flag = (uint32_t*)(pl.img + pl.il.q_flag);
  SR_STAGE(3, dbg, st, "sort", launch_sort_tiles(sa, st));
  if (!pairs) return SR_OK;  // the _part variant's blend walks the sub-lists itself
  AcceptArgs aa{};
  aa.vs = pl.vs;
  aa.ranges = sa.ranges;
  aa.counters = counters;
  aa.sub = sa.sub;
  aa.sub_count = sa.sub_count;
  aa.rec = sc.rec;
  aa.cullbox = (const short4*)(pl.geom + pl.gl.cullbox);
  aa.work = sa.work;
  aa.work_count = counters + 2;
  aa.masks = (uint64_t*)(bin + bl.masks);
  aa.chunk_T = (float*)(bin + bl.chunk_T);
  aa.frag_off = (uint32_t*)(bin + bl.frag_off);
  aa.lane_range = (uint2*)(pl.img + pl.il.lane_range);
  aa.q_cost = (uint32_t*)(pl.img + pl.il.q_cost);
  aa.ps = pair_store(bin, bl);
  aa.q_flag = sa.q_flag;
  aa.frag_fill = (unsigned long long*)(counters + 4);
  aa.frag_cap = (uint32_t)bl.frag_cap;
  const size_t items = work_capacity(pl.capacity, pl.vs.B, pl.il.T);
  SR_STAGE(8, dbg, st, "accept", launch_accept(aa, items, st));
  SR_STAGE(10, dbg, st, "pair_offsets", launch_pair_offsets(aa, st));
  SR_STAGE(10, dbg, st, "order",
           launch_order_quads(pl.il.T * quads_per_tile(pl.vs.B), aa.q_cost, (uint32_t*)(pl.img + pl.il.q_order), st));
  SR_STAGE(10, dbg, st, "pair_write", launch_pair_write(aa, items, st));
  return SR_OK;
}

int run_blend_fwd(Pipeline& pl, float* out_color, float* out_others, bool dbg, hipStream_t st) {
  BlendArgs ba{};
  ba.vs = pl.vs;
  ba.ranges = (const uint32_t*)(pl.img + pl.il.ranges);
  ba.counters = pl.counters();
  ba.sub = (const uint2*)(pl.bin + pl.bl.sub);
  ba.sub_count = (const uint32_t*)(pl.img + pl.il.sub_count);
  ba.masks = (const uint64_t*)(pl.bin + pl.bl.masks);
  ba.frag_off = (const uint32_t*)(pl.bin + pl.bl.frag_off);
  ba.lane_range = (uint2*)(pl.img + pl.il.lane_range);
  ba.q_order = (const uint32_t*)(pl.img + pl.il.q_order);
  ba.ps = pair_store(pl.bin, pl.bl);
  ba.q_flag = (const uint32_t*)(pl.img + pl.il.q_flag);
  ba.rec = (const float4*)(pl.geom + pl.gl.rec);
  ba.bg = pl.bg;
  ba.final_T = (float*)(pl.img + pl.il.final_T);
  ba.n_contrib = (uint32_t*)(pl.img + pl.il.n_contrib);
  ba.out_color = out_color;
  ba.out_others = out_others;
  SR_STAGE(4, dbg, st, "blend_fwd", launch_blend_fwd(ba, st));
  return SR_OK;
}

int run_backward(Pipeline& pl, const sr_surfels* g, const int* radii, const float* dL_dcolor,
                 const float* dL_dothers, const RecordGrads& rg, sr_grads* gr, bool blend, bool dbg,
                 hipStream_t st) {
  if (blend) {
    BlendBwdArgs ba{};
    ba.vs = pl.vs;
    ba.ranges = (const uint32_t*)(pl.img + pl.il.ranges);
    ba.counters = pl.counters();
    ba.sub = (const uint2*)(pl.bin + pl.bl.sub);
    ba.sub_count = (const uint32_t*)(pl.img + pl.il.sub_count);
    ba.masks = (const uint64_t*)(pl.bin + pl.bl.masks);
    ba.frag_off = (const uint32_t*)(pl.bin + pl.bl.frag_off);
    ba.lane_range = (const uint2*)(pl.img + pl.il.lane_range);
    ba.q_order = (const uint32_t*)(pl.img + pl.il.q_order);
    ba.ps = pair_store(pl.bin, pl.bl);
    ba.q_flag = (const uint32_t*)(pl.img + pl.il.q_flag);
    ba.work = (const uint2*)(pl.bin + pl.bl.work);
    ba.work_count = pl.counters() + 2;
    ba.rec = (const float4*)(pl.geom + pl.gl.rec);
    ba.bg = pl.bg;
    ba.final_T = (const float*)(pl.img + pl.il.final_T);
    ba.n_contrib = (const uint32_t*)(pl.img + pl.il.n_contrib);
    ba.dL_dpix = dL_dcolor;
    ba.dL_dothers = dL_dothers;
    ba.pos_slot = (const uint32_t*)(pl.bin + pl.bl.pos_slot);
    ba.pos_quads = (const uint16_t*)(pl.bin + pl.bl.pos_quads);
    ba.part = (float*)(pl.bin + pl.bl.part);
    ba.part_nz = (uint8_t*)(pl.bin + pl.bl.part_nz);
    SR_STAGE(6, dbg, st, "blend_bwd", launch_blend_bwd(ba, st));
    SR_STAGE(9, dbg, st, "blend_bwd_grad",
             launch_blend_bwd_grad(ba, work_capacity(pl.capacity, pl.vs.B, pl.il.T), st));
  }
  PreprocessBwdArgs pb{};
  pb.vs = pl.vs;
  pb.P = pl.P;
  pb.D = pl.s_D;
  pb.M = g->M;
  pb.counters = pl.counters();
  pb.tiles = (const uint32_t*)(pl.geom + pl.gl.tiles);
  pb.roff = (const uint32_t*)(pl.geom + pl.gl.roff);
  pb.rslot = (const uint2*)(pl.geom + pl.gl.rslot);
  pb.part = blend ? (const float*)(pl.bin + pl.bl.part) : nullptr;
  pb.part_nz = blend ? (const uint8_t*)(pl.bin + pl.bl.part_nz) : nullptr;
  pb.means3D = g->means3D;
  pb.scales = g->scales;
  pb.rots = g->rotations;
  pb.shs = g->shs;
  pb.tprec = g->transMat_precomp;
  pb.rec = (const float4*)(pl.geom + pl.gl.rec);
  pb.radii = radii;
  pb.clamped = (const uint8_t*)(pl.geom + pl.gl.clamped);
  pb.vm = pl.vm;
  pb.pm = pl.pm;
  pb.campos = pl.campos;
  pb.dL_dtransMat = rg.transMat;
  pb.dL_dcolors = rg.colors;
  pb.dL_dopacity = rg.opacity;
  pb.dL_dopacity_out = gr->dL_dopacity;
  pb.dL_dcolors_out = gr->dL_dcolors;
  pb.dL_dsh = gr->dL_dsh;
  pb.dL_dmean2D = rg.mean2D;
  pb.dL_dmean3D = gr->dL_dmeans3D;
  pb.dL_dscales = gr->dL_dscales;
  pb.dL_drots = gr->dL_drotations;
  SR_STAGE(7, dbg, st, "preprocess_bwd", launch_preprocess_bwd(pb, st));
  return SR_OK;
}

}  // namespace

extern "C" {

const char* sr_last_error(void) { return g_err.c_str(); }
const char* sr_version(void) { return "surfel_rast gfx950 1.0"; }

size_t sr_geom_bytes(int P) { return geom_layout((size_t)(P < 0 ? 0 : P)).total; }
size_t sr_image_bytes(int W, int H, int tile) { return image_layout(W, H, tile > 0 ? tile : 16).total; }
size_t sr_binning_bytes(int L, int W, int H, int tile, int pair_store, int pair_capacity, int slot_capacity) {
  const int B = tile > 0 ? tile : 16;
  return bin_layout(L, B, image_layout(W, H, B).T, pair_store, pair_capacity, 0, slot_capacity).total;
}

namespace {
// The newest landed count record of the calling thread and device, or NULL.
const uint32_t* latest_counts(int& rc) {
  rc = SR_OK;
  PairCount& pc = pair_count_slot();
  int best = -1;
  for (int k = 0; k < kPairSlots; ++k) {
    if (pc.seq[k] == 0 || (best >= 0 && pc.seq[k] < pc.seq[best])) continue;
    const hipError_t q = hipEventQuery(pc.ev[k]);
    if (q == hipSuccess) best = k;
    else if (q != hipErrorNotReady) {
      rc = fail(SR_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
      return nullptr;
    }
  }
  if (best < 0) {
    rc = fail(SR_ERR_INVALID, "no forward has completed on this thread and device");
    return nullptr;
  }
  return pc.host + kCountWords * best;
}
}  // namespace

int sr_pair_count(unsigned long long* pairs, unsigned int* instances) {
  if (!pairs || !instances) return fail(SR_ERR_INVALID, "NULL argument");
  int rc;
  const uint32_t* c = latest_counts(rc);
  if (!c) return rc;
  *instances = c[0];
  *pairs = (unsigned long long)c[4] | ((unsigned long long)c[5] << 32);
  return SR_OK;
}

int sr_slot_count(unsigned int* slots, unsigned int* instances) {
  if (!slots || !instances) return fail(SR_ERR_INVALID, "NULL argument");
  int rc;
  const uint32_t* c = latest_counts(rc);
  if (!c) return rc;
  *instances = c[0];
  *slots = c[8];
  return SR_OK;
}

size_t sr_state_offset(const char* name, int P, int W, int H, int tile, int L) {
  const GeomLayout g = geom_layout((size_t)(P < 0 ? 0 : P));
  const ImageLayout im = image_layout(W, H, tile > 0 ? tile : 16);
  const BinLayout b = bin_layout(L, tile > 0 ? tile : 16, im.T);
  const std::string n(name ? name : "");
  if (n == "rec") return g.rec;
  if (n == "depths") return g.depths;
  if (n == "tiles_touched") return g.tiles;
  if (n == "clamped") return g.clamped;
  if (n == "counters") return im.counters;
  if (n == "tile_count") return im.tile_count;
  if (n == "tile_start") return im.tile_start;
  if (n == "ranges") return im.ranges;
  if (n == "final_T") return im.final_T;
  if (n == "n_contrib") return im.n_contrib;
  if (n == "keys") return b.keys;
  if (n == "point_list") return b.point_list;
  if (n == "sub") return b.sub;
  if (n == "sub_count") return im.sub_count;
  if (n == "cullbox") return g.cullbox;
  if (n == "masks") return b.masks;
  if (n == "lane_range") return im.lane_range;
  return (size_t)-1;
}

int sr_forward(const sr_settings* s, const sr_surfels* g, sr_alloc_fn alloc, void* ctx, float* out_color,
               float* out_others, int* radii, int* num_rendered, sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate(s, g);
  if (rc != SR_OK) return rc;
  if (!alloc || !out_color || !out_others || !num_rendered) return fail(SR_ERR_INVALID, "NULL output");
  if (g->P > 0 && (!radii || !g->opacities)) return fail(SR_ERR_INVALID, "radii and opacities must not be NULL");
  const int P = g->P, W = s->image_width, H = s->image_height, B = s->tile;
  const bool dbg = s->debug != 0;
  const int cap = s->capacity;
  if (cap < 0) return fail(SR_ERR_INVALID, "capacity must be >= 0 (got %d)", cap);
  *num_rendered = 0;
  if (P == 0) {
    // The reference glue skips the rasteriser for P == 0 and returns its
    // zero-filled outputs (rasterize_points.cu:85-87, 99).
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_color, 0.0f, 3 * (size_t)W * H, st));
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_others, 0.0f, 7 * (size_t)W * H, st));
    return SR_OK;
  }
  const ViewSet vs = single_view(W, H, B, s->tanfovx, s->tanfovy);
  const GeomLayout gl = geom_layout((size_t)P);
  const ImageLayout il = image_layout(vs);
  char* geom = (char*)alloc(ctx, gl.total, 0);
  char* img = (char*)alloc(ctx, il.total, 2);
  if (!geom || !img) return fail(SR_ERR_ALLOC, "allocation callback failed");
  Pipeline pl{};
  pl.vs = vs;
  pl.P = P;
  pl.s_D = s->sh_degree;
  pl.prefiltered = s->prefiltered;
  pl.mod = s->scale_modifier;
  pl.pair_store = s->pair_store;
  pl.pair_capacity = s->pair_capacity;
  pl.vm = s->viewmatrix;
  pl.pm = s->projmatrix;
  pl.campos = s->campos;
  pl.bg = s->bg;
  pl.geom = geom;
  pl.img = img;
  pl.gl = gl;
  pl.il = il;
  uint32_t* host = pinned_counters();
  if (!host) return fail(SR_ERR_HIP, "hipHostMalloc failed");
  if (cap > 0) {
    // Everything enqueued at once on a binning buffer of the caller's
    // capacity; the kernels after the scan do nothing when the instances
    // exceed it.  The instance count is copied back right after the scan
    // and waited for only once the rest is enqueued, by when it has long
    // arrived, so the stream never drains mid-forward.
    pl.bl = bin_layout(cap, B, il.T, s->pair_store, s->pair_capacity, 0, s->slot_capacity);
    pl.bin = (char*)alloc(ctx, pl.bl.total, 1);
    if (!pl.bin) return fail(SR_ERR_ALLOC, "allocation callback failed");
    pl.capacity = cap;
    rc = run_front(pl, g, radii, (uint32_t)cap, (uint32_t)pl.bl.nslot, dbg, st);
    if (rc != SR_OK) return rc;
    SR_HIP(hipMemcpyAsync(host, pl.counters(), 9 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    hipEvent_t counted = counts_event();
    if (!counted) return fail(SR_ERR_HIP, "hipEventCreate failed");
    SR_HIP(hipEventRecord(counted, st));
    rc = run_binning(pl, radii, dbg, st);
    if (rc == SR_OK) rc = run_blend_fwd(pl, out_color, out_others, dbg, st);
    if (rc == SR_OK) rc = post_pair_count(pl, st);
    SR_HIP(hipEventSynchronize(counted));
    if (rc != SR_OK) return rc;
    if (host[1] & kErrPrefiltered)
      return fail(SR_ERR_PREFILTERED, "Point is filtered although prefiltered is set. This shouldn't happen!");
    *num_rendered = (int)host[0];
    if ((host[1] & kErrCapacity) && (uint32_t)cap < host[0])
      return fail(SR_ERR_CAPACITY, "%u tile instances exceed the capacity %d", host[0], cap);
    if (host[1] & kErrCapacity)
      return fail(SR_ERR_CAPACITY, "%u partial-gradient slots exceed the slot capacity %zu", host[8], pl.bl.nslot);
    return SR_OK;
  }
  rc = run_front(pl, g, radii, 0xFFFFFFFFu, 0xFFFFFFFFu, dbg, st);
  if (rc != SR_OK) return rc;
  SR_HIP(hipMemcpyAsync(host, pl.counters(), 9 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SR_HIP(hipStreamSynchronize(st));
  if (host[1] & kErrPrefiltered)
    return fail(SR_ERR_PREFILTERED, "Point is filtered although prefiltered is set. This shouldn't happen!");
  const int L = (int)host[0];
  *num_rendered = L;
  // the slots the forward needs, exactly (the slot region is the layout's
  // last, so the backward locates everything without this count)
  const int S = (int)(host[8] > 0u ? host[8] : 1u);
  pl.bl = bin_layout(L, B, il.T, s->pair_store, s->pair_capacity, 0, S);
  pl.bin = (char*)alloc(ctx, pl.bl.total, 1);
  if (!pl.bin) return fail(SR_ERR_ALLOC, "allocation callback failed");
  pl.capacity = L;
  if (L > 0) {
    rc = run_binning(pl, radii, dbg, st);
    if (rc != SR_OK) return rc;
    rc = run_blend_fwd(pl, out_color, out_others, dbg, st);
    if (rc != SR_OK) return rc;
    return post_pair_count(pl, st);
  } else {
    SR_STAGE(3, dbg, st, "clear",
             launch_fill_u32((uint32_t*)(img + il.sub_count), 0u, (size_t)il.T * quads_per_tile(B), st));
    SR_STAGE(3, dbg, st, "clear",
             launch_fill_u32((uint32_t*)(img + il.q_flag), 0u, (size_t)il.T * quads_per_tile(B), st));
    SR_STAGE(3, dbg, st, "clear",
             launch_fill_u32((uint32_t*)(img + il.lane_range), 0u, (size_t)il.T * quads_per_tile(B) * 128, st));
    SR_STAGE(3, dbg, st, "clear", launch_iota_u32((uint32_t*)(img + il.q_order), (size_t)il.T * quads_per_tile(B), st));
  }
  return run_blend_fwd(pl, out_color, out_others, dbg, st);
}

int sr_backward(const sr_settings* s, const sr_surfels* g, int R, const int* radii, void* geom_buf,
                const void* bin_buf, const void* img_buf, const float* dL_dcolor, const float* dL_dothers,
                sr_grads* gr, sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate(s, g);
  if (rc != SR_OK) return rc;
  if (!gr || !dL_dcolor || !dL_dothers) return fail(SR_ERR_INVALID, "NULL gradient argument");
  const int P = g->P, W = s->image_width, H = s->image_height, B = s->tile;
  if (P == 0) return SR_OK;
  if (!geom_buf || !img_buf || !radii) return fail(SR_ERR_INVALID, "forward state missing");
  if (R < 0) return fail(SR_ERR_INVALID, "num_rendered must be >= 0");
  if (!gr->dL_dmeans2D || !gr->dL_dcolors || !gr->dL_dopacity || !gr->dL_dmeans3D || !gr->dL_dtransMat ||
      !gr->dL_dscales || !gr->dL_drotations || (g->shs && !gr->dL_dsh))
    return fail(SR_ERR_INVALID, "gradient output missing");
  const bool dbg = s->debug != 0;
  Pipeline pl{};
  pl.vs = single_view(W, H, B, s->tanfovx, s->tanfovy);
  pl.P = P;
  pl.s_D = s->sh_degree;
  pl.vm = s->viewmatrix;
  pl.pm = s->projmatrix;
  pl.campos = s->campos;
  pl.bg = s->bg;
  pl.gl = geom_layout((size_t)P);
  pl.il = image_layout(pl.vs);
  // the binning arrays are laid out for the forward's capacity (R when 0)
  const int cap = s->capacity > 0 ? s->capacity : R;
  if (cap < R) return fail(SR_ERR_INVALID, "capacity %d below num_rendered %d", s->capacity, R);
  pl.bl = bin_layout(cap, B, pl.il.T, s->pair_store, s->pair_capacity);
  pl.geom = (char*)geom_buf;
  pl.img = (char*)img_buf;
  pl.bin = (char*)bin_buf;
  pl.capacity = cap;
  // With one view the per-record gradient arrays are the outputs themselves;
  // the backward writes every output once, so nothing is cleared first.
  RecordGrads rg{gr->dL_dtransMat, gr->dL_dmeans2D, gr->dL_dopacity, gr->dL_dcolors};
  return run_backward(pl, g, radii, dL_dcolor, dL_dothers, rg, gr, R > 0 && bin_buf != nullptr, dbg, st);
}

// ---- the _part variant -----------------------------------------------------
namespace {
PartArgs part_args(const Pipeline& pl, const float* sem) {
  PartArgs pa{};
  pa.vs = pl.vs;
  pa.S = pl.S;
  pa.ranges = (const uint32_t*)(pl.img + pl.il.ranges);
  pa.sub = (const uint2*)(pl.bin + pl.bl.sub);
  pa.sub_count = (const uint32_t*)(pl.img + pl.il.sub_count);
  pa.rec = (const float4*)(pl.geom + pl.gl.rec);
  pa.sem = sem;
  pa.bg = pl.bg;
  return pa;
}

int validate_part(const sr_settings* s, const sr_surfels* g, int S, const float* semantics) {
  int rc = validate(s, g);
  if (rc != SR_OK) return rc;
  if (S < 0 || S > 16) return fail(SR_ERR_INVALID, "semantic channels must be in [0, 16] (got %d)", S);
  if (g->P > 0 && S > 0 && !semantics) return fail(SR_ERR_INVALID, "semantics must not be NULL");
  return SR_OK;
}
}  // namespace

int sr_forward_part(const sr_settings* s, const sr_surfels* g, int S, const float* semantics, sr_alloc_fn alloc,
                    void* ctx, float* out_color, float* out_semantic, float* out_others, int* radii,
                    int* num_rendered, sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate_part(s, g, S, semantics);
  if (rc != SR_OK) return rc;
  if (!alloc || !out_color || !out_others || !num_rendered || (S > 0 && !out_semantic))
    return fail(SR_ERR_INVALID, "NULL output");
  if (g->P > 0 && (!radii || !g->opacities)) return fail(SR_ERR_INVALID, "radii and opacities must not be NULL");
  const int P = g->P, W = s->image_width, H = s->image_height, B = s->tile;
  const bool dbg = s->debug != 0;
  *num_rendered = 0;
  if (P == 0) {
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_color, 0.0f, 3 * (size_t)W * H, st));
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_others, 0.0f, 8 * (size_t)W * H, st));
    if (S > 0) SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_semantic, 0.0f, (size_t)S * W * H, st));
    return SR_OK;
  }
  Pipeline pl{};
  pl.vs = single_view(W, H, B, s->tanfovx, s->tanfovy);
  pl.P = P;
  pl.s_D = s->sh_degree;
  pl.prefiltered = s->prefiltered;
  pl.mod = s->scale_modifier;
  pl.part = true;
  pl.S = S;
  pl.vm = s->viewmatrix;
  pl.pm = s->projmatrix;
  pl.campos = s->campos;
  pl.bg = s->bg;
  pl.gl = geom_layout((size_t)P);
  pl.il = image_layout(pl.vs);
  pl.geom = (char*)alloc(ctx, pl.gl.total, 0);
  pl.img = (char*)alloc(ctx, pl.il.total, 2);
  if (!pl.geom || !pl.img) return fail(SR_ERR_ALLOC, "allocation callback failed");
  uint32_t* host = pinned_counters();
  if (!host) return fail(SR_ERR_HIP, "hipHostMalloc failed");
  rc = run_front(pl, g, radii, 0xFFFFFFFFu, 0xFFFFFFFFu, dbg, st);
  if (rc != SR_OK) return rc;
  SR_HIP(hipMemcpyAsync(host, pl.counters(), 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SR_HIP(hipStreamSynchronize(st));
  if (host[1] & kErrPrefiltered)
    return fail(SR_ERR_PREFILTERED, "Point is filtered although prefiltered is set. This shouldn't happen!");
  const int L = (int)host[0];
  *num_rendered = L;
  pl.bl = bin_layout(L, B, pl.il.T, 0, 1, S);
  pl.bin = (char*)alloc(ctx, pl.bl.total, 1);
  if (!pl.bin) return fail(SR_ERR_ALLOC, "allocation callback failed");
  pl.capacity = L;
  if (L > 0) {
    rc = run_binning(pl, radii, dbg, st, false);
    if (rc != SR_OK) return rc;
  } else {
    SR_STAGE(3, dbg, st, "clear",
             launch_fill_u32((uint32_t*)(pl.img + pl.il.sub_count), 0u, (size_t)pl.il.T * quads_per_tile(B), st));
  }
  PartArgs pa = part_args(pl, semantics);
  pa.final_T = (float*)(pl.img + pl.il.final_T);
  pa.n_contrib = (uint32_t*)(pl.img + pl.il.n_contrib);
  pa.out_color = out_color;
  pa.out_sem = out_semantic;
  pa.out_others = out_others;
  SR_STAGE(4, dbg, st, "blend_fwd", launch_part_blend_fwd(pa, st));
  return SR_OK;
}

int sr_backward_part(const sr_settings* s, const sr_surfels* g, int S, const float* semantics, int R,
                     const int* radii, void* geom_buf, const void* bin_buf, const void* img_buf,
                     const float* dL_dcolor, const float* dL_dsemantic, const float* dL_dothers, sr_grads* gr,
                     float* dL_dsemantics, sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate_part(s, g, S, semantics);
  if (rc != SR_OK) return rc;
  if (!gr || !dL_dcolor || !dL_dothers || (S > 0 && (!dL_dsemantic || !dL_dsemantics)))
    return fail(SR_ERR_INVALID, "NULL gradient argument");
  const int P = g->P, W = s->image_width, H = s->image_height, B = s->tile;
  if (P == 0) return SR_OK;
  if (!geom_buf || !img_buf || !radii) return fail(SR_ERR_INVALID, "forward state missing");
  if (R < 0) return fail(SR_ERR_INVALID, "num_rendered must be >= 0");
  if (!gr->dL_dmeans2D || !gr->dL_dcolors || !gr->dL_dopacity || !gr->dL_dmeans3D || !gr->dL_dtransMat ||
      !gr->dL_dscales || !gr->dL_drotations || (g->shs && !gr->dL_dsh))
    return fail(SR_ERR_INVALID, "gradient output missing");
  const bool dbg = s->debug != 0;
  Pipeline pl{};
  pl.vs = single_view(W, H, B, s->tanfovx, s->tanfovy);
  pl.P = P;
  pl.s_D = s->sh_degree;
  pl.part = true;
  pl.S = S;
  pl.vm = s->viewmatrix;
  pl.pm = s->projmatrix;
  pl.campos = s->campos;
  pl.bg = s->bg;
  pl.gl = geom_layout((size_t)P);
  pl.il = image_layout(pl.vs);
  pl.bl = bin_layout(R, B, pl.il.T, 0, 1, S);
  pl.geom = (char*)geom_buf;
  pl.img = (char*)img_buf;
  pl.bin = (char*)bin_buf;
  pl.capacity = R;
  const bool blend = R > 0 && bin_buf != nullptr;
  if (blend) {
    PartArgs pa = part_args(pl, semantics);
    pa.final_T_in = (const float*)(pl.img + pl.il.final_T);
    pa.n_contrib_in = (const uint32_t*)(pl.img + pl.il.n_contrib);
    pa.dL_dpix = dL_dcolor;
    pa.dL_dsem = dL_dsemantic;
    pa.dL_dothers = dL_dothers;
    pa.pos_slot = (const uint32_t*)(pl.bin + pl.bl.pos_slot);
    pa.pos_quads = (const uint16_t*)(pl.bin + pl.bl.pos_quads);
    pa.part = (float*)(pl.bin + pl.bl.part);
    pa.sem_part = (float*)(pl.bin + pl.bl.sem_part);
    pa.part_nz = (uint8_t*)(pl.bin + pl.bl.part_nz);
    SR_STAGE(6, dbg, st, "blend_bwd", launch_part_blend_bwd(pa, st));
  }
  PreprocessBwdArgs pb{};
  pb.vs = pl.vs;
  pb.P = P;
  pb.D = pl.s_D;
  pb.M = g->M;
  pb.counters = pl.counters();
  pb.tiles = (const uint32_t*)(pl.geom + pl.gl.tiles);
  pb.roff = (const uint32_t*)(pl.geom + pl.gl.roff);
  pb.rslot = (const uint2*)(pl.geom + pl.gl.rslot);
  pb.part = blend ? (const float*)(pl.bin + pl.bl.part) : nullptr;
  pb.part_nz = blend ? (const uint8_t*)(pl.bin + pl.bl.part_nz) : nullptr;
  pb.sem_part = blend ? (const float*)(pl.bin + pl.bl.sem_part) : nullptr;
  pb.S = S;
  pb.means3D = g->means3D;
  pb.scales = g->scales;
  pb.rots = g->rotations;
  pb.shs = g->shs;
  pb.tprec = g->transMat_precomp;
  pb.rec = (const float4*)(pl.geom + pl.gl.rec);
  pb.radii = radii;
  pb.clamped = (const uint8_t*)(pl.geom + pl.gl.clamped);
  pb.vm = pl.vm;
  pb.pm = pl.pm;
  pb.campos = pl.campos;
  pb.dL_dtransMat = gr->dL_dtransMat;
  pb.dL_dcolors = gr->dL_dcolors;
  pb.dL_dopacity = gr->dL_dopacity;
  pb.dL_dsh = gr->dL_dsh;
  pb.dL_dmean2D = gr->dL_dmeans2D;
  pb.dL_dmean3D = gr->dL_dmeans3D;
  pb.dL_dscales = gr->dL_dscales;
  pb.dL_drots = gr->dL_drotations;
  pb.dL_dsem = dL_dsemantics;
  SR_STAGE(7, dbg, st, "preprocess_bwd", launch_part_preprocess_bwd(pb, st));
  return SR_OK;
}

size_t sr_batch_workspace_bytes(const sr_views* v, int P, int capacity) {
  if (!v || v->V < 1 || v->V > SR_MAX_VIEWS || P < 0 || capacity < 0 || !v->widths || !v->heights) return 0;
  const int B = v->tile > 0 ? v->tile : 16;
  if (v->slot_capacity < 0) return 0;
  return batch_layout(make_viewset(v->V, v->widths, v->heights, v->tanfovx, v->tanfovy, B), P, capacity,
                      v->pair_store, v->pair_capacity, v->slot_capacity).total;
}

int sr_forward_batched(const sr_views* v, const sr_surfels* g, void* ws, size_t ws_bytes, int capacity,
                       float* out_color, float* out_others, int* radii, uint32_t* status, sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate_views(v, g, capacity);
  if (rc != SR_OK) return rc;
  if (!ws || !out_color || !out_others || !radii || !status) return fail(SR_ERR_INVALID, "NULL argument");
  const ViewSet vs = make_viewset(v->V, v->widths, v->heights, v->tanfovx, v->tanfovy, v->tile);
  const BatchLayout bl = batch_layout(vs, g->P, capacity, v->pair_store, v->pair_capacity, v->slot_capacity);
  if (ws_bytes < bl.total)
    return fail(SR_ERR_INVALID, "workspace holds %zu bytes, needs %zu", ws_bytes, bl.total);
  Pipeline pl = batch_pipeline(v, g, vs, bl, (char*)ws, capacity);
  // Many views hold thousands of tiles above kSmallMax instances: sorting
  // them in chunks and merging by rank does more work in all than one
  // workgroup per tile, and throughput is what counts here (latency of the
  // few largest tiles is what counts for one view).  So every tile above
  // kSmallMax takes the one-workgroup sort, beside the small-tile sort on a
  // second stream (a captured graph makes the fork cheap).
  pl.side_huge = true;
  rc = run_front(pl, g, radii, (uint32_t)capacity, (uint32_t)bl.bl.nslot, false, st);
  if (rc == SR_OK) rc = run_binning(pl, radii, false, st);
  if (rc == SR_OK) rc = run_blend_fwd(pl, out_color, out_others, false, st);
  if (rc != SR_OK) return rc;
  hipError_t e = launch_copy_status(pl.counters(), status, st);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "status: %s", hipGetErrorString(e));
  return SR_OK;
}

int sr_backward_batched(const sr_views* v, const sr_surfels* g, void* ws, size_t ws_bytes, int capacity,
                        const int* radii, const float* dL_dcolor, const float* dL_dothers, sr_grads* gr,
                        sr_stream_t stream_) {
  hipStream_t st = (hipStream_t)stream_;
  int rc = validate_views(v, g, capacity);
  if (rc != SR_OK) return rc;
  if (!ws || !radii || !dL_dcolor || !dL_dothers || !gr) return fail(SR_ERR_INVALID, "NULL argument");
  if (!gr->dL_dmeans2D || !gr->dL_dcolors || !gr->dL_dopacity || !gr->dL_dmeans3D || !gr->dL_dscales ||
      !gr->dL_drotations || (g->shs && !gr->dL_dsh))
    return fail(SR_ERR_INVALID, "gradient output missing");
  const ViewSet vs = make_viewset(v->V, v->widths, v->heights, v->tanfovx, v->tanfovy, v->tile);
  const BatchLayout bl = batch_layout(vs, g->P, capacity, v->pair_store, v->pair_capacity, v->slot_capacity);
  if (ws_bytes < bl.total)
    return fail(SR_ERR_INVALID, "workspace holds %zu bytes, needs %zu", ws_bytes, bl.total);
  Pipeline pl = batch_pipeline(v, g, vs, bl, (char*)ws, capacity);
  char* w = (char*)ws;
  RecordGrads rg{(float*)(w + bl.g_transMat), gr->dL_dmeans2D, (float*)(w + bl.g_opacity),
                 (float*)(w + bl.g_colors)};
  return run_backward(pl, g, radii, dL_dcolor, dL_dothers, rg, gr, true, false, st);
}

int sr_mark_visible(int P, const float* means3D, const float* viewmatrix, const float* projmatrix,
                    uint8_t* present, sr_stream_t stream_) {
  (void)projmatrix;
  if (P < 0) return fail(SR_ERR_INVALID, "P must be >= 0");
  if (P > 0 && (!means3D || !viewmatrix || !present)) return fail(SR_ERR_INVALID, "NULL argument");
  hipError_t e = launch_mark_visible(P, means3D, viewmatrix, present, (hipStream_t)stream_);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "mark_visible: %s", hipGetErrorString(e));
  return SR_OK;
}

int sr_set_timing(int enabled) {
  return sr_set_timing_stages(enabled ? (1u << SR_NUM_STAGES) - 1u : 0u);
}

int sr_set_timing_stages(unsigned int mask) {
  Timing& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  t.mask = mask & ((1u << SR_NUM_STAGES) - 1u);
  return SR_OK;
}

int sr_get_timing(double* ms, int* count, int n) {
  Timing& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  for (auto& p : t.pending) {
    if (hipEventSynchronize(p.b) != hipSuccess) return fail(SR_ERR_HIP, "hipEventSynchronize failed");
    float el = 0.0f;
    if (hipEventElapsedTime(&el, p.a, p.b) != hipSuccess) return fail(SR_ERR_HIP, "hipEventElapsedTime failed");
    t.ms[p.stage] += el;
    t.count[p.stage] += 1;
    t.pool.push_back(p.a);
    t.pool.push_back(p.b);
  }
  t.pending.clear();
  for (int i = 0; i < n && i < SR_NUM_STAGES; ++i) {
    if (ms) ms[i] = t.ms[i];
    if (count) count[i] = t.count[i];
  }
  return SR_OK;
}

void sr_reset_timing(void) {
  double ms[SR_NUM_STAGES];
  int c[SR_NUM_STAGES];
  sr_get_timing(ms, c, SR_NUM_STAGES);
  Timing& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  for (int i = 0; i < SR_NUM_STAGES; ++i) {
    t.ms[i] = 0;
    t.count[i] = 0;
  }
}

}  // extern "C"
