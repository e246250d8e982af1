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

#define SR_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(SR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---- per-stage timing ------------------------------------------------------
struct Timing {
  std::mutex mu;
  bool on = false;
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
    if (!t.on) return;
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

// Pinned 8-byte landing zone for the instance count and error flags.
uint32_t* pinned_counters() {
  thread_local uint32_t* p = nullptr;
  if (!p) {
    if (hipHostMalloc((void**)&p, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

int validate(const sr_settings* s, const sr_surfels* g) {
  if (!s || !g) return fail(SR_ERR_INVALID, "settings and surfels must not be NULL");
  if (g->P < 0) return fail(SR_ERR_INVALID, "P must be >= 0");
  if (s->image_width <= 0 || s->image_height <= 0 || s->image_width > 32000 || s->image_height > 32000)
    return fail(SR_ERR_INVALID, "image size must be in [1, 32000] (got %d x %d)", s->image_width, s->image_height);
  if (s->pair_store < 0) return fail(SR_ERR_INVALID, "pair_store must be >= 0");
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

}  // namespace

extern "C" {

const char* sr_last_error(void) { return g_err.c_str(); }
const char* sr_version(void) { return "surfel_rast gfx950 1.0"; }

size_t sr_geom_bytes(int P) { return geom_layout(P < 0 ? 0 : P).total; }
size_t sr_image_bytes(int W, int H, int tile) { return image_layout(W, H, tile > 0 ? tile : 16).total; }
size_t sr_binning_bytes(int L, int W, int H, int tile, int pair_store) {
  const int B = tile > 0 ? tile : 16;
  return bin_layout(L, B, image_layout(W, H, B).T, pair_store).total;
}

size_t sr_state_offset(const char* name, int P, int W, int H, int tile, int L) {
  const GeomLayout g = geom_layout(P);
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
  *num_rendered = 0;
  if (P == 0) {
    // The reference glue skips the rasteriser for P == 0 and returns its
    // zero-filled outputs (rasterize_points.cu:85-87, 99).
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_color, 0.0f, 3 * (size_t)W * H, st));
    SR_STAGE(4, dbg, st, "clear", launch_fill_f32(out_others, 0.0f, 7 * (size_t)W * H, st));
    return SR_OK;
  }
  const GeomLayout gl = geom_layout(P);
  const ImageLayout il = image_layout(W, H, B);
  char* geom = (char*)alloc(ctx, gl.total, 0);
  char* img = (char*)alloc(ctx, il.total, 2);
  if (!geom || !img) return fail(SR_ERR_ALLOC, "allocation callback failed");
  uint32_t* counters = (uint32_t*)(img + il.counters);
  uint32_t* tile_count = (uint32_t*)(img + il.tile_count);
  uint32_t* tile_start = (uint32_t*)(img + il.tile_start);
  uint32_t* tile_fill = (uint32_t*)(img + il.tile_fill);
  uint32_t* ranges = (uint32_t*)(img + il.ranges);
  // clear counters and the tile histogram (fill kernels, no memset nodes)
  SR_STAGE(1, dbg, st, "clear", launch_fill_u32(counters, 0u, kAlign / 4, st));
  SR_STAGE(1, dbg, st, "clear", launch_fill_u32(tile_count, 0u, (size_t)il.T, st));
  PreprocessArgs pa{};
  pa.P = P;
  pa.D = s->sh_degree;
  pa.M = g->M;
  pa.W = W;
  pa.H = H;
  pa.B = B;
  pa.gx = il.gx;
  pa.gy = il.gy;
  pa.prefiltered = s->prefiltered;
  pa.mod = s->scale_modifier;
  pa.means3D = g->means3D;
  pa.scales = g->scales;
  pa.rots = g->rotations;
  pa.opac = g->opacities;
  pa.shs = g->shs;
  pa.colors = g->colors_precomp;
  pa.tprec = g->transMat_precomp;
  pa.vm = s->viewmatrix;
  pa.pm = s->projmatrix;
  pa.campos = s->campos;
  pa.rec = (float4*)(geom + gl.rec);
  pa.depths = (float*)(geom + gl.depths);
  pa.radii = radii;
  pa.tiles = (uint32_t*)(geom + gl.tiles);
  pa.clamped = (uint8_t*)(geom + gl.clamped);
  pa.cullbox = (short4*)(geom + gl.cullbox);
  pa.tile_count = tile_count;
  pa.counters = counters;
  SR_STAGE(0, dbg, st, "preprocess", launch_preprocess(pa, st));
  SR_STAGE(1, dbg, st, "scan", launch_scan_tiles(il.T, tile_count, tile_start, tile_fill, ranges, counters, st));
  uint32_t* host = pinned_counters();
  if (!host) return fail(SR_ERR_HIP, "hipHostMalloc failed");
  SR_HIP(hipMemcpyAsync(host, counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SR_HIP(hipStreamSynchronize(st));
  if (host[1] & kErrPrefiltered)
    return fail(SR_ERR_PREFILTERED, "Point is filtered although prefiltered is set. This shouldn't happen!");
  const int L = (int)host[0];
  *num_rendered = L;
  const BinLayout bl = bin_layout(L, B, il.T, s->pair_store);
  char* bin = (char*)alloc(ctx, bl.total, 1);
  if (!bin) return fail(SR_ERR_ALLOC, "allocation callback failed");
  uint64_t* keys = (uint64_t*)(bin + bl.keys);
  uint64_t* keys_tmp = (uint64_t*)(bin + bl.keys_tmp);
  uint32_t* point_list = (uint32_t*)(bin + bl.point_list);
  uint2* sub = (uint2*)(bin + bl.sub);
  uint32_t* sub_count = (uint32_t*)(img + il.sub_count);
  if (L > 0) {
    SR_STAGE(2, dbg, st, "scatter",
             launch_scatter(P, B, il.gx, il.gy, (const float4*)(geom + gl.rec), (const float*)(geom + gl.depths),
                            radii, tile_fill, keys, st));
    SortArgs sa{};
    sa.W = W;
    sa.H = H;
    sa.B = B;
    sa.gx = il.gx;
    sa.T = il.T;
    sa.ranges = ranges;
    sa.keys = keys;
    sa.keys_tmp = keys_tmp;
    sa.point_list = point_list;
    sa.cullbox = (const short4*)(geom + gl.cullbox);
    sa.sub = sub;
    sa.sub_count = sub_count;
    sa.work = (uint2*)(bin + bl.work);
    sa.work_count = counters + 2;
    SR_STAGE(3, dbg, st, "sort", launch_sort_tiles(sa, st));
    AcceptArgs aa{};
    aa.W = W;
    aa.H = H;
    aa.B = B;
    aa.gx = il.gx;
    aa.ranges = ranges;
    aa.sub = sub;
    aa.sub_count = sub_count;
    aa.rec = (const float4*)(geom + gl.rec);
    aa.work = (const uint2*)(bin + bl.work);
    aa.work_count = counters + 2;
    aa.masks = (uint64_t*)(bin + bl.masks);
    aa.frag_off = (uint32_t*)(bin + bl.frag_off);
    aa.frag4 = (float4*)(bin + bl.frag4);
    aa.frag1 = (float*)(bin + bl.frag1);
    aa.frag_fill = (unsigned long long*)(counters + 4);
    aa.frag_cap = (uint32_t)bl.frag_cap;
    SR_STAGE(8, dbg, st, "accept", launch_accept(aa, work_capacity(L, B, il.T), st));
  } else {
    SR_STAGE(3, dbg, st, "clear", launch_fill_u32(sub_count, 0u, (size_t)il.T * quads_per_tile(B), st));
  }
  BlendArgs ba{};
  ba.W = W;
  ba.H = H;
  ba.B = B;
  ba.gx = il.gx;
  ba.gy = il.gy;
  ba.T = il.T;
  ba.ranges = ranges;
  ba.sub = sub;
  ba.sub_count = sub_count;
  ba.masks = (const uint64_t*)(bin + bl.masks);
  ba.frag_off = (const uint32_t*)(bin + bl.frag_off);
  ba.frag4 = (const float4*)(bin + bl.frag4);
  ba.rec = (const float4*)(geom + gl.rec);
  ba.bg = s->bg;
  ba.final_T = (float*)(img + il.final_T);
  ba.n_contrib = (uint32_t*)(img + il.n_contrib);
  ba.out_color = out_color;
  ba.out_others = out_others;
  SR_STAGE(4, dbg, st, "blend_fwd", launch_blend_fwd(ba, st));
  return SR_OK;
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
  const GeomLayout gl = geom_layout(P);
  const ImageLayout il = image_layout(W, H, B);
  const BinLayout bl = bin_layout(R, B, il.T, s->pair_store);
  const char* geom = (const char*)geom_buf;
  const char* img = (const char*)img_buf;
  const char* bin = (const char*)bin_buf;
  {
    StageScope sc(5, st);
    hipError_t e = launch_fill_f32(gr->dL_dmeans2D, 0.0f, 3 * (size_t)P, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_dcolors, 0.0f, 3 * (size_t)P, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_dopacity, 0.0f, (size_t)P, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_dmeans3D, 0.0f, 3 * (size_t)P, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_dtransMat, 0.0f, 9 * (size_t)P, st);
    if (e == hipSuccess && g->shs) e = launch_fill_f32(gr->dL_dsh, 0.0f, (size_t)P * g->M * 3, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_dscales, 0.0f, 2 * (size_t)P, st);
    if (e == hipSuccess) e = launch_fill_f32(gr->dL_drotations, 0.0f, 4 * (size_t)P, st);
    if (e != hipSuccess) return fail(SR_ERR_HIP, "clear grads: %s", hipGetErrorString(e));
  }
  // dL/dnormal is an intermediate between the two backward kernels; it lives
  // in the geometry buffer (the reference keeps it in a glue-side tensor,
  // rasterize_points.cu:190).
  float* dL_dnormal = (float*)((char*)geom_buf + gl.dnormal);
  SR_STAGE(5, dbg, st, "clear", launch_fill_f32(dL_dnormal, 0.0f, 3 * (size_t)P, st));
  BlendBwdArgs ba{};
  ba.W = W;
  ba.H = H;
  ba.B = B;
  ba.gx = il.gx;
  ba.gy = il.gy;
  ba.T = il.T;
  ba.ranges = (const uint32_t*)(img + il.ranges);
  ba.sub = bin ? (const uint2*)(bin + bl.sub) : nullptr;
  ba.sub_count = (const uint32_t*)(img + il.sub_count);
  ba.masks = bin ? (const uint64_t*)(bin + bl.masks) : nullptr;
  ba.frag_off = bin ? (const uint32_t*)(bin + bl.frag_off) : nullptr;
  ba.frag4 = bin ? (const float4*)(bin + bl.frag4) : nullptr;
  ba.frag1 = bin ? (const float*)(bin + bl.frag1) : nullptr;
  ba.rec = (const float4*)(geom + gl.rec);
  ba.bg = s->bg;
  ba.final_T = (const float*)(img + il.final_T);
  ba.n_contrib = (const uint32_t*)(img + il.n_contrib);
  ba.dL_dpix = dL_dcolor;
  ba.dL_dothers = dL_dothers;
  ba.dL_dtransMat = gr->dL_dtransMat;
  ba.dL_dmean2D = gr->dL_dmeans2D;
  ba.dL_dnormal = dL_dnormal;
  ba.dL_dopacity = gr->dL_dopacity;
  ba.dL_dcolors = gr->dL_dcolors;
  if (R > 0) SR_STAGE(6, dbg, st, "blend_bwd", launch_blend_bwd(ba, st));
  PreprocessBwdArgs pb{};
  pb.P = P;
  pb.D = s->sh_degree;
  pb.M = g->M;
  pb.W = W;
  pb.H = H;
  pb.tanx = s->tanfovx;
  pb.tany = s->tanfovy;
  pb.means3D = g->means3D;
  pb.scales = g->scales;
  pb.rots = g->rotations;
  pb.shs = g->shs;
  pb.tprec = g->transMat_precomp;
  pb.rec = (const float4*)(geom + gl.rec);
  pb.radii = radii;
  pb.clamped = (const uint8_t*)(geom + gl.clamped);
  pb.vm = s->viewmatrix;
  pb.pm = s->projmatrix;
  pb.campos = s->campos;
  pb.dL_dtransMat = gr->dL_dtransMat;
  pb.dL_dnormal = dL_dnormal;
  pb.dL_dcolors = gr->dL_dcolors;
  pb.dL_dsh = gr->dL_dsh;
  pb.dL_dmean2D = gr->dL_dmeans2D;
  pb.dL_dmean3D = gr->dL_dmeans3D;
  pb.dL_dscales = gr->dL_dscales;
  pb.dL_drots = gr->dL_drotations;
  SR_STAGE(7, dbg, st, "preprocess_bwd", launch_preprocess_bwd(pb, st));
  return SR_OK;
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
  Timing& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  t.on = enabled != 0;
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
