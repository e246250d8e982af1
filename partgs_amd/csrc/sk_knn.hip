// Mean squared distance of every point to its three nearest neighbours:
// distCUDA2 of simple-knn (submodules/simple-knn/spatial.cu:16-27,
// simple_knn.cu:185-220), which sets the initial surfel scales of a point
// cloud (scene/gaussian_model.py:237).
//
// What is computed.  For point q, the squared distance to another point c is
// (dx dx + dy dy) + dz dz in FP32 with d = c - q (simple_knn.cu:129-133),
// each product and sum rounded on its own; the result is
// ((b0 + b1) + b2) / 3 over the three smallest such values b0 <= b1 <= b2
// (:176), FLT_MAX standing in for missing neighbours when there are fewer
// than four points (:150,162).  The reference finds them by exact search
// (its box pruning only skips boxes whose lower bound exceeds a value the
// third neighbour cannot exceed), so the result is a function of the point
// set alone; this search is exact too and returns the same bits whatever
// order it visits the points in.
//
// How.  A coarse Morton bucket per point (5 bits per axis over the bounding
// box, whose extremes include the origin as the reference's reduction with
// init (0, 0, 0) does, :191-202), a counting sort into buckets, leaf boxes of
// 256 consecutive sorted points and groups of 32 boxes with their bounds;
// one 256-lane workgroup per leaf box then searches: its own box first, then
// every group and box whose distance bound does not exceed the lane's current
// third-best (or the bound from its +-3 sorted neighbours, :146-158), with
// each box's points staged once in LDS for all lanes that need it.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../../include/surfel_rast.h"
#include "sr_internal.h"

namespace sk {

constexpr int kThreads = 256;
constexpr int kBox = 256;        // points per leaf box (= kThreads: one workgroup per box)
constexpr int kGroup = 32;       // leaf boxes per group
constexpr int kAxisBits = 5;     // Morton bits per axis of the bucket key
constexpr int kCells = 1 << (3 * kAxisBits);
constexpr int kMaxParts = 1024;  // workgroups of the bounds reduction

struct Layout {
  size_t part, keys, count, start, spts, boxes, groups, total;
  int nparts, nbox, ngrp;
};

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

inline Layout layout(int P) {
  Layout l{};
  const size_t n = (size_t)(P > 0 ? P : 1);
  l.nparts = (int)std::min<long long>(kMaxParts, std::max<long long>(1, ((long long)P + 1023) / 1024));
  l.nbox = (int)((n + kBox - 1) / kBox);
  l.ngrp = (l.nbox + kGroup - 1) / kGroup;
  size_t off = 0;
  l.part = off; off += align_up((size_t)l.nparts * 6 * sizeof(float));
  l.keys = off; off += align_up(n * sizeof(uint32_t));
  l.count = off; off += align_up((size_t)kCells * sizeof(uint32_t));
  l.start = off; off += align_up((size_t)kCells * sizeof(uint32_t));
  l.spts = off; off += align_up(n * sizeof(float4));
  l.boxes = off; off += align_up((size_t)l.nbox * 2 * sizeof(float4));
  l.groups = off; off += align_up((size_t)l.ngrp * 2 * sizeof(float4));
  l.total = off;
  return l;
}

__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }

// Bounding box of the points, per workgroup (grid-stride), as 6 floats.
__global__ __launch_bounds__(kThreads) void k_bounds(int P, const float* __restrict__ pts, float* part) {
  __shared__ float s[6][kThreads / 64];
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < P; i += gridDim.x * kThreads) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float v = pts[3 * (size_t)i + k];
      mn[k] = fmin_(mn[k], v);
      mx[k] = fmax_(mx[k], v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin_(mn[k], __shfl_xor(mn[k], off, 64));
      mx[k] = fmax_(mx[k], __shfl_xor(mx[k], off, 64));
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s[k][w] = mn[k];
      s[3 + k][w] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    float v = s[threadIdx.x][0];
    for (int j = 1; j < kThreads / 64; ++j)
      v = threadIdx.x < 3 ? fmin_(v, s[threadIdx.x][j]) : fmax_(v, s[threadIdx.x][j]);
    part[6 * blockIdx.x + threadIdx.x] = v;
  }
}

__device__ __forceinline__ uint32_t spread5(uint32_t x) {  // 5 bits -> every third bit
  x &= 0x1fu;
  x = (x | (x << 8)) & 0x0100F00Fu;
  x = (x | (x << 4)) & 0x010C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__device__ __forceinline__ uint32_t cell_of(float v, float lo, float hi) {
  const float span = hi - lo;
  if (!(span > 0.0f)) return 0u;
  float t = (v - lo) / span * (float)(1 << kAxisBits);
  t = t < 0.0f ? 0.0f : t;  // also catches NaN
  const uint32_t c = (uint32_t)t;
  return c > (uint32_t)((1 << kAxisBits) - 1) ? (uint32_t)((1 << kAxisBits) - 1) : c;
}

// Bucket key of every point and the bucket histogram.
__global__ __launch_bounds__(kThreads) void k_keys(int P, const float* __restrict__ pts, const float* part,
                                                   int nparts, uint32_t* keys, uint32_t* count) {
  __shared__ float s_b[6];
  if (threadIdx.x < 6) {
    // the reference reduces with init (0, 0, 0): the origin is in the box
    float v = 0.0f;
    for (int j = 0; j < nparts; ++j)
      v = threadIdx.x < 3 ? fmin_(v, part[6 * j + threadIdx.x]) : fmax_(v, part[6 * j + threadIdx.x]);
    s_b[threadIdx.x] = v;
  }
  __syncthreads();
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= P) return;
  const uint32_t cx = cell_of(pts[3 * (size_t)i], s_b[0], s_b[3]);
  const uint32_t cy = cell_of(pts[3 * (size_t)i + 1], s_b[1], s_b[4]);
  const uint32_t cz = cell_of(pts[3 * (size_t)i + 2], s_b[2], s_b[5]);
  const uint32_t key = spread5(cx) | (spread5(cy) << 1) | (spread5(cz) << 2);
  keys[i] = key;
  atomicAdd(&count[key], 1u);
}

// Exclusive scan of the kCells bucket counts by one workgroup of 1024 lanes.
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* count, uint32_t* start) {
  __shared__ uint32_t s[1024];
  constexpr int per = kCells / 1024;
  const int t = threadIdx.x;
  uint32_t loc = 0;
  for (int k = 0; k < per; ++k) loc += count[t * per + k];
  s[t] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  uint32_t run = t == 0 ? 0u : s[t - 1];
  for (int k = 0; k < per; ++k) {
    start[t * per + k] = run;
    run += count[t * per + k];
  }
}

// Points into their buckets (x, y, z, index bits).  The order inside a
// bucket depends on scheduling; the search result does not.
__global__ __launch_bounds__(kThreads) void k_scatter(int P, const float* __restrict__ pts,
                                                      const uint32_t* keys, uint32_t* fill, float4* spts) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= P) return;
  const uint32_t at = atomicAdd(&fill[keys[i]], 1u);
  spts[at] = make_float4(pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2],
                         __uint_as_float((uint32_t)i));
}

// Bounds of every leaf box (one workgroup each).
__global__ __launch_bounds__(kThreads) void k_boxes(int P, const float4* spts, float4* boxes) {
  __shared__ float s[6][kThreads / 64];
  const int i = blockIdx.x * kBox + threadIdx.x;
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  if (i < P) {
    const float4 p = spts[i];
    mn[0] = mx[0] = p.x;
    mn[1] = mx[1] = p.y;
    mn[2] = mx[2] = p.z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin_(mn[k], __shfl_xor(mn[k], off, 64));
      mx[k] = fmax_(mx[k], __shfl_xor(mx[k], off, 64));
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s[k][w] = mn[k];
      s[3 + k][w] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float b[6];
    for (int k = 0; k < 6; ++k) {
      b[k] = s[k][0];
      for (int j = 1; j < kThreads / 64; ++j) b[k] = k < 3 ? fmin_(b[k], s[k][j]) : fmax_(b[k], s[k][j]);
    }
    boxes[2 * blockIdx.x] = make_float4(b[0], b[1], b[2], 0.0f);
    boxes[2 * blockIdx.x + 1] = make_float4(b[3], b[4], b[5], 0.0f);
  }
}

// Bounds of every group of kGroup leaf boxes (one wave each).
__global__ __launch_bounds__(64) void k_groups(int nbox, const float4* boxes, float4* groups) {
  const int b = blockIdx.x * kGroup + (int)threadIdx.x;
  const bool in = threadIdx.x < kGroup && b < nbox;
  float4 lo = in ? boxes[2 * b] : make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.0f);
  float4 hi = in ? boxes[2 * b + 1] : make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo.x = fmin_(lo.x, __shfl_xor(lo.x, off, 64));
    lo.y = fmin_(lo.y, __shfl_xor(lo.y, off, 64));
    lo.z = fmin_(lo.z, __shfl_xor(lo.z, off, 64));
    hi.x = fmax_(hi.x, __shfl_xor(hi.x, off, 64));
    hi.y = fmax_(hi.y, __shfl_xor(hi.y, off, 64));
    hi.z = fmax_(hi.z, __shfl_xor(hi.z, off, 64));
  }
  if (threadIdx.x == 0) {
    groups[2 * blockIdx.x] = lo;
    groups[2 * blockIdx.x + 1] = hi;
  }
}

// distBoxPoint, simple_knn.cu:115-125: a lower bound of the squared
// distance from p to any point inside the box.
__device__ __forceinline__ float box_dist(const float4& lo, const float4& hi, float px, float py, float pz) {
  float dx = 0.0f, dy = 0.0f, dz = 0.0f;
  if (px < lo.x || px > hi.x) dx = fmin_(fabsf(px - lo.x), fabsf(px - hi.x));
  if (py < lo.y || py > hi.y) dy = fmin_(fabsf(py - lo.y), fabsf(py - hi.y));
  if (pz < lo.z || pz > hi.z) dz = fmin_(fabsf(pz - lo.z), fabsf(pz - hi.z));
  return dx * dx + dy * dy + dz * dz;
}

// updateKBest<3>, simple_knn.cu:127-143: insertion into the ascending triple.
__device__ __forceinline__ void update_best(float px, float py, float pz, const float4& c, float* best) {
  const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
  float d = dx * dx + dy * dy + dz * dz;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (best[j] > d) {
      const float t = best[j];
      best[j] = d;
      d = t;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_query(int P, int nbox, int ngrp, const float4* __restrict__ spts,
                                                    const float4* __restrict__ boxes,
                                                    const float4* __restrict__ groups, float* out) {
  __shared__ float4 s_pts[kBox];
  const int own = blockIdx.x;
  const int pos = own * kBox + (int)threadIdx.x;
  const bool valid = pos < P;
  const float4 q = valid ? spts[pos] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float best[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  if (valid) {
    const int lo = pos - 3 < 0 ? 0 : pos - 3, hi = pos + 3 > P - 1 ? P - 1 : pos + 3;
    for (int i = lo; i <= hi; ++i)
      if (i != pos) update_best(q.x, q.y, q.z, spts[i], best);
  }
  const float reject = best[2];
  best[0] = best[1] = best[2] = FLT_MAX;
  // the points of box b against every lane whose bounds admit it
  auto visit = [&](int b, bool need) {
    if (!__syncthreads_or(need ? 1 : 0)) return;
    const int j = b * kBox + (int)threadIdx.x;
    s_pts[threadIdx.x] = j < P ? spts[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    if (need) {
      const int m = P - b * kBox < kBox ? P - b * kBox : kBox;
      for (int k = 0; k < m; ++k)
        if (b * kBox + k != pos) update_best(q.x, q.y, q.z, s_pts[k], best);
    }
  };
  auto admits = [&](float d) { return valid && !(d > reject || d > best[2]); };
  visit(own, valid);
  for (int g = 0; g < ngrp; ++g) {
    const bool gneed = admits(box_dist(groups[2 * g], groups[2 * g + 1], q.x, q.y, q.z));
    if (!__syncthreads_or(gneed ? 1 : 0)) continue;
    const int b1 = (g + 1) * kGroup < nbox ? (g + 1) * kGroup : nbox;
    for (int b = g * kGroup; b < b1; ++b) {
      if (b == own) continue;
      visit(b, gneed && admits(box_dist(boxes[2 * b], boxes[2 * b + 1], q.x, q.y, q.z)));
    }
  }
  if (valid) out[__float_as_uint(q.w)] = (best[0] + best[1] + best[2]) / 3.0f;
}

__global__ void k_zero(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace sk

extern "C" {

size_t sk_dist2_workspace_bytes(int P) {
  if (P < 0) return 0;
  return sk::layout(P).total;
}

int sk_dist2(int P, const float* points, float* mean_dists, void* workspace, size_t workspace_bytes,
             sr_stream_t stream) {
  if (P < 0) return sr::set_error(SR_ERR_INVALID, "sk_dist2: P must be >= 0");
  if (P == 0) return SR_OK;
  if (!points || !mean_dists) return sr::set_error(SR_ERR_INVALID, "sk_dist2: NULL points or output");
  if (P > (1 << 30)) return sr::set_error(SR_ERR_INVALID, "sk_dist2: more than 2^30 points");
  const sk::Layout l = sk::layout(P);
  if (!workspace || workspace_bytes < l.total)
    return sr::set_error(SR_ERR_INVALID, "sk_dist2: workspace smaller than sk_dist2_workspace_bytes(P)");
  char* ws = (char*)workspace;
  float* part = (float*)(ws + l.part);
  uint32_t* keys = (uint32_t*)(ws + l.keys);
  uint32_t* count = (uint32_t*)(ws + l.count);
  uint32_t* start = (uint32_t*)(ws + l.start);
  float4* spts = (float4*)(ws + l.spts);
  float4* boxes = (float4*)(ws + l.boxes);
  float4* groups = (float4*)(ws + l.groups);
  hipStream_t s = (hipStream_t)stream;
  const unsigned pb = (unsigned)((P + sk::kThreads - 1) / sk::kThreads);
  hipLaunchKernelGGL(sk::k_zero, dim3(sk::kCells / 256), dim3(256), 0, s, count, sk::kCells);
  hipLaunchKernelGGL(sk::k_bounds, dim3(l.nparts), dim3(sk::kThreads), 0, s, P, points, part);
  hipLaunchKernelGGL(sk::k_keys, dim3(pb), dim3(sk::kThreads), 0, s, P, points, part, l.nparts, keys, count);
  hipLaunchKernelGGL(sk::k_scan, dim3(1), dim3(1024), 0, s, count, start);
  hipLaunchKernelGGL(sk::k_scatter, dim3(pb), dim3(sk::kThreads), 0, s, P, points, keys, start, spts);
  hipLaunchKernelGGL(sk::k_boxes, dim3(l.nbox), dim3(sk::kThreads), 0, s, P, spts, boxes);
  hipLaunchKernelGGL(sk::k_groups, dim3(l.ngrp), dim3(64), 0, s, l.nbox, boxes, groups);
  hipLaunchKernelGGL(sk::k_query, dim3(l.nbox), dim3(sk::kThreads), 0, s, P, l.nbox, l.ngrp, spts, boxes, groups,
                     mean_dists);
  return hipGetLastError() == hipSuccess ? SR_OK : sr::set_error(SR_ERR_HIP, "sk_dist2: kernel launch failed");
}

}  // extern "C"
