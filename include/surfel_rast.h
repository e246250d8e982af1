/*
 * surfel_rast.h -- C ABI of the MI355X (gfx950) differentiable surfel
 * rasteriser (2D Gaussian splatting) used by PartGS.
 *
 * Every pointer marked "device" is HIP device memory on the current device;
 * "host" pointers are ordinary host memory.  All work is enqueued on the
 * hipStream_t passed in (NULL = the legacy default stream); the only host
 * synchronisation is the 4-byte read of the tile-instance count inside
 * sr_forward, as in the reference.  Functions return SR_OK (0) or an SR_ERR_*
 * code; sr_last_error() gives a message for the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree,
 * dsr = submodules/diff-surfel-rasterization):
 *   sr_forward        <- CudaRasterizer::Rasterizer::forward,
 *                        dsr/cuda_rasterizer/rasterizer.h:31-57 and its torch
 *                        glue RasterizeGaussiansCUDA, dsr/rasterize_points.cu:40-134
 *   sr_backward       <- CudaRasterizer::Rasterizer::backward,
 *                        dsr/cuda_rasterizer/rasterizer.h:59-86 and
 *                        RasterizeGaussiansBackwardCUDA, dsr/rasterize_points.cu:137-233
 *   sr_mark_visible   <- Rasterizer::markVisible, rasterizer.h:23-29 and
 *                        markVisible, dsr/rasterize_points.cu:235-254
 *   sr_*_bytes        <- required<GeometryState/ImageState/BinningState>,
 *                        dsr/cuda_rasterizer/rasterizer_impl.h:67-73
 */
#ifndef SURFEL_RAST_H
#define SURFEL_RAST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* sr_stream_t; /* == hipStream_t */

enum {
  SR_OK = 0,
  SR_ERR_INVALID = 1,     /* bad argument (shape, NULL, unsupported tile) */
  SR_ERR_HIP = 2,         /* a HIP runtime call failed */
  SR_ERR_ALLOC = 3,       /* the allocation callback returned NULL */
  SR_ERR_PREFILTERED = 4, /* prefiltered set but a surfel failed the near test
                             (the reference traps, dsr/cuda_rasterizer/auxiliary.h:202-206) */
  SR_ERR_CAPACITY = 5     /* more tile instances than the caller's capacity */
};

/* GaussianRasterizationSettings, dsr/diff_surfel_rasterization/__init__.py:158-170.
 * tile is the screen tile edge: 16 reproduces the reference (config.h:15-17);
 * 8 and 32 are also accepted. */
typedef struct {
  int image_height;
  int image_width;
  float tanfovx;
  float tanfovy;
  const float* bg;         /* device, 3 */
  float scale_modifier;
  const float* viewmatrix; /* device, 16, as the reference's world_view_transform */
  const float* projmatrix; /* device, 16, full_proj_transform */
  int sh_degree;
  const float* campos;     /* device, 3 */
  int prefiltered;
  int debug;               /* synchronise and check after every stage */
  int tile;
  int pair_store;          /* accepted-pair store per tile instance (0 = 24);
                              chunks beyond it recompute their pairs */
  int capacity;            /* tile instances the binning buffer is sized for.
                              0: sized from the instance count, which the
                              forward reads back before it binns (a stream
                              synchronisation mid-forward, as the reference).
                              > 0: the whole forward is enqueued at once and
                              the count is read back after; more instances
                              than this return SR_ERR_CAPACITY with
                              *num_rendered set (call again with at least
                              that).  The backward must be given the same
                              value, as it locates the binning arrays by it. */
  int pair_capacity;       /* accepted-pair store size in pairs; 0: pair_store
                              pairs per tile instance.  Quadrants whose pairs
                              do not fit are blended by recomputing them (same
                              results, slower).  sr_pair_count reports what a
                              forward asked for, so a caller can size the next
                              one from it.  The backward must be given the
                              same value. */
  int slot_capacity;       /* with capacity > 0: the partial-gradient slots
                              the binning buffer holds (one per tile instance
                              and 8 x 8 quadrant its cull box reaches); 0:
                              every (instance, quadrant) of the capacity,
                              which always fits.  More slots than this return
                              SR_ERR_CAPACITY with *num_rendered within the
                              capacity (call again with a larger value or 0).
                              With capacity 0 the slots are sized from the
                              exact count and this is not used.
                              sr_slot_count reports what a forward needed;
                              the backward does not need the value.  No
                              counterpart in the reference. */
} sr_settings;

/* Per-surfel inputs (device).  Exactly one of shs / colors_precomp and exactly
 * one of (scales and rotations) / transMat_precomp must be non-NULL, as
 * checked by GaussianRasterizer.forward, __init__.py:192-196. */
typedef struct {
  int P;                         /* number of surfels */
  int M;                         /* SH coefficients per surfel (shs is P x M x 3) */
  const float* means3D;          /* P x 3 */
  const float* shs;              /* P x M x 3, or NULL */
  const float* colors_precomp;   /* P x 3, or NULL */
  const float* opacities;        /* P */
  const float* scales;           /* P x 2, or NULL */
  const float* rotations;        /* P x 4 (w, x, y, z), or NULL */
  const float* transMat_precomp; /* P x 9, or NULL */
} sr_surfels;

/* Gradient outputs (device), the tuple of rasterize_points.cu:232.  Shapes:
 * dL_dmeans2D P x 3, dL_dcolors P x 3, dL_dopacity P, dL_dmeans3D P x 3,
 * dL_dtransMat P x 9, dL_dsh P x M x 3 (may be NULL when M == 0),
 * dL_dscales P x 2, dL_drotations P x 4.  sr_backward overwrites them all. */
typedef struct {
  float* dL_dmeans2D;
  float* dL_dcolors;
  float* dL_dopacity;
  float* dL_dmeans3D;
  float* dL_dtransMat;
  float* dL_dsh;
  float* dL_dscales;
  float* dL_drotations;
} sr_grads;

/* Allocation callback: returns device memory of at least nbytes that stays
 * valid until the matching sr_backward has returned, or NULL on failure.
 * which = 0 geometry buffer, 1 binning buffer, 2 image buffer.  Mirrors the
 * std::function<char*(size_t)> resize callbacks, rasterize_points.cu:31-37. */
typedef void* (*sr_alloc_fn)(void* ctx, size_t nbytes, int which);

/* Forward: preprocess, tile binning + per-tile depth sort, blend.
 * out_color 3 x H x W, out_others 7 x H x W (depth, alpha, normal xyz,
 * median depth, distortion; forward.cu:422-440), radii P (int32).
 * *num_rendered receives the number of (surfel, tile) instances.
 * The three buffers obtained from alloc are the state sr_backward needs. */
int sr_forward(const sr_settings* s, const sr_surfels* g, sr_alloc_fn alloc, void* alloc_ctx,
               float* out_color, float* out_others, int* radii, int* num_rendered,
               sr_stream_t stream);

/* Backward.  geom_buf / bin_buf / img_buf are the buffers sr_forward
 * obtained (geom_buf is also used as scratch); R is its num_rendered;
 * radii its radii output.  dL_dcolor is
 * 3 x H x W, dL_dothers 7 x H x W. */
int sr_backward(const sr_settings* s, const sr_surfels* g, int R, const int* radii,
                void* geom_buf, const void* bin_buf, const void* img_buf,
                const float* dL_dcolor, const float* dL_dothers, sr_grads* grads,
                sr_stream_t stream);

/* ---- the _part variant --------------------------------------------------
 * The rasteriser of submodules/diff-surfel-rasterization_part, which
 * renderer/gaussian_renderer_2d/__init__.py:86-96 calls: camera-space
 * homography, pixel centres at +0.5, S <= 16 semantic channels blended
 * beside the colour, an 8-channel out_others (depth, alpha, normal xyz,
 * median depth, distortion, median weight).  The settings, surfels,
 * allocation callback and state buffers are those of sr_forward (its
 * capacity and pair_store fields are not used).
 *   sr_forward_part  <- RasterizeGaussiansCUDA, dsp/rasterize_points.cu:40-146
 *   sr_backward_part <- RasterizeGaussiansBackwardCUDA, dsp/rasterize_points.cu:149-252
 * (dsp = submodules/diff-surfel-rasterization_part).
 * semantics is P x S (device), out_semantic S x H x W, out_others 8 x H x W. */
int sr_forward_part(const sr_settings* s, const sr_surfels* g, int S, const float* semantics,
                    sr_alloc_fn alloc, void* alloc_ctx, float* out_color, float* out_semantic,
                    float* out_others, int* radii, int* num_rendered, sr_stream_t stream);
/* Backward of sr_forward_part: dL_dsemantic S x H x W, dL_dothers 8 x H x W;
 * dL_dsemantics (P x S) receives the semantics' gradient.  As the
 * reference, the semantic images' gradient does not reach the geometry
 * (dsp/cuda_rasterizer/backward.cu:346).  With a precomputed transMat the
 * scale and rotation gradients are zero (the reference reads inputs it was
 * not given there). */
int sr_backward_part(const sr_settings* s, const sr_surfels* g, int S, const float* semantics, int R,
                     const int* radii, void* geom_buf, const void* bin_buf, const void* img_buf,
                     const float* dL_dcolor, const float* dL_dsemantic, const float* dL_dothers,
                     sr_grads* grads, float* dL_dsemantics, sr_stream_t stream);

/* present[i] = 1 when surfel i passes the near test (view z > 0.2). */
int sr_mark_visible(int P, const float* means3D, const float* viewmatrix,
                    const float* projmatrix, uint8_t* present, sr_stream_t stream);

/* Byte sizes of the three state buffers (what sr_forward asks alloc for). */
size_t sr_geom_bytes(int P);
size_t sr_image_bytes(int W, int H, int tile);
size_t sr_binning_bytes(int num_rendered, int W, int H, int tile, int pair_store, int pair_capacity,
                        int slot_capacity);

/* Accepted pairs (*pairs) and tile instances (*instances) of the calling
 * thread's latest completed sr_forward on the current device: the pairs are
 * those the forward asked the pair store for (the accepted pairs in front
 * of each pixel's saturation point).  The counts are copied back
 * asynchronously at the end of each forward and never waited for: returns
 * SR_OK once such a copy has landed, SR_ERR_INVALID before.  A caller sizes
 * the next forward's pair_capacity from them.  No counterpart in the
 * reference. */
int sr_pair_count(unsigned long long* pairs, unsigned int* instances);

/* Partial-gradient slots (*slots) and tile instances (*instances) of the
 * same latest completed forward, under the same rules as sr_pair_count; a
 * caller sizes the next forward's slot_capacity from them.  No counterpart
 * in the reference. */
int sr_slot_count(unsigned int* slots, unsigned int* instances);

/* Byte offsets of the named arrays inside the state buffers, so callers and
 * tests can read internal state (e.g. "depths", "rec", "tiles_touched",
 * "clamped", "ranges", "n_contrib", "final_T", "point_list", "keys").
 * Returns (size_t)-1 for unknown names. */
size_t sr_state_offset(const char* name, int P, int W, int H, int tile, int num_rendered);

/* Per-stage GPU timing with HIP events on the caller's stream.
 * Stages: 0 preprocess, 1 tile scan, 2 instance scatter, 3 tile sort,
 * 4 forward blend, 5 gradient clear, 6 backward blend (per-pixel pass),
 * 7 backward preprocess, 8 acceptance masks, 9 backward blend (per-pair pass),
 * 10 accepted-pair store (offsets and fill). */
#define SR_NUM_STAGES 11
int sr_set_timing(int enabled);
/* Times only the stages whose bit is set in mask (bit i: stage i), so a
 * caller can time one kernel without the other stages' events. */
int sr_set_timing_stages(unsigned int mask);
/* Synchronises pending events; ms[i] = total milliseconds, count[i] = launches. */
int sr_get_timing(double* ms, int* count, int n);
void sr_reset_timing(void);

/* ---- batched views ------------------------------------------------------
 * Many views of one surfel set (uneven image sizes allowed) in one pass:
 * every stage is one launch over all views, nothing is allocated and the
 * host never waits, so a forward + backward can be captured in a hipGraph
 * and replayed.  The caller owns a workspace of sr_batch_workspace_bytes()
 * sized for `capacity` (surfel, tile) instances over all views; when a batch
 * needs more, the forward sets SR_STATUS_CAPACITY in status[1], leaves the
 * outputs undefined, and status[0] holds the instances needed.
 *
 * No counterpart in the reference, which renders one view per call
 * (renderer/gaussian_renderer/__init__.py:9-149 inside the training loop);
 * per view the results equal sr_forward / sr_backward on that view. */
#define SR_MAX_VIEWS 32
#define SR_STATUS_CAPACITY 2u

typedef struct {
  int V;                      /* 1..SR_MAX_VIEWS */
  const int* widths;          /* host, V */
  const int* heights;         /* host, V */
  const float* tanfovx;       /* host, V */
  const float* tanfovy;       /* host, V */
  const float* viewmatrices;  /* device, V x 16 */
  const float* projmatrices;  /* device, V x 16 */
  const float* campos;        /* device, V x 3 */
  const float* bg;            /* device, 3 (shared) */
  float scale_modifier;
  int sh_degree;
  int tile;                   /* 8, 16 or 32 */
  int pair_store;             /* as in sr_settings */
  int pair_capacity;          /* as in sr_settings (pairs over all views) */
  int slot_capacity;          /* partial-gradient slots over all views; 0: every
                                 (instance, quadrant) of the capacity.  More
                                 slots set SR_STATUS_CAPACITY with status[0]
                                 within the capacity; status[8] holds the
                                 slots needed.  Forward and backward must be
                                 given the same value. */
} sr_views;

size_t sr_batch_workspace_bytes(const sr_views* views, int P, int capacity);

/* Outputs: out_color / out_others hold each view's 3 x H x W / 7 x H x W
 * planes back to back in view order (view v starts at 3 / 7 times the pixel
 * count of views 0..v-1); radii is V x P; status (device, 16 words) receives
 * [0] instances, [1] flags (SR_STATUS_CAPACITY), [2] acceptance work items,
 * [4..5] the accepted pairs asked of the pair store (64-bit, low word first;
 * see sr_pair_count), [8] the partial-gradient slots; the other words are
 * internal.
 * The surfels need scales and rotations (no precomputed T, which is per
 * view); prefiltered is not supported. */
int sr_forward_batched(const sr_views* views, const sr_surfels* g, void* workspace, size_t workspace_bytes,
                       int capacity, float* out_color, float* out_others, int* radii, uint32_t* status,
                       sr_stream_t stream);

/* Backward of the last sr_forward_batched on the same workspace.  dL_dcolor /
 * dL_dothers are laid out like the forward outputs.  grads: dL_dmeans2D is
 * V x P x 3 (each view's screen-space statistic, as the reference's
 * per-view calls return it); dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dsh,
 * dL_dscales and dL_drotations are per surfel, summed over the views;
 * dL_dtransMat is not used. */
int sr_backward_batched(const sr_views* views, const sr_surfels* g, void* workspace, size_t workspace_bytes,
                        int capacity, const int* radii, const float* dL_dcolor, const float* dL_dothers,
                        sr_grads* grads, sr_stream_t stream);

/* ---- superquadric block layer -------------------------------------------
 * Blocks -> attached surfels, BlockGaussianModel.prepare_scaling_rot and
 * get_xyz / get_scaling / get_rotation / get_opacity
 * (games/block_mesh_splatting/scene/block_gaussian_model.py:103-121,178-256).
 * The block mesh is a regular (eta, omega) grid: face (i, j) is the triangle
 * (V[i][j], V[i+1][j], V[i][(j+1) % n_omega]) and carries K surfels with
 * barycentric weights bary (K x 3; NULL means K = 1 at the centroid).
 * Surfel index = ((b * n_eta + i) * n_omega + j) * K + k. */
typedef struct {
  int B;                 /* number of blocks */
  int n_eta, n_omega;    /* faces per block = n_eta * n_omega */
  int K;                 /* surfels per face */
  const float* eta;      /* device, n_eta + 1 vertex latitudes */
  const float* omega;    /* device, n_omega vertex longitudes */
  const float* bary;     /* device, K x 3, or NULL */
  float ratio;           /* ratio_block_scene (parametric_sq scale) */
  float scale_min;       /* scale_block_min added to exp(sq_s) */
  float surfel_scale;    /* per-surfel _scale factor on (s1, s2) */
  const float* params;   /* device, B x 12: sq_s(3), sq_r(4, w x y z), sq_t(3), sq_eps(2) */
  const float* occupancy;/* device, B: sq_occ logits */
} sb_blocks;

/* No counterpart in the reference as a native kernel: the reference builds
 * these surfels with ~20 torch ops per iteration (prepare_scaling_rot).
 * Outputs P x 3, P x 2, P x 4, P (P = B * n_eta * n_omega * K). */
int sb_generate_fwd(const sb_blocks* blk, float* means3D, float* scales, float* rotations, float* opacities,
                    sr_stream_t stream);
/* Adjoint of sb_generate_fwd: dL_dparams B x 12 and dL_doccupancy B are
 * overwritten with the gradients of the block parameters, summed in a fixed
 * order (the same bits on every run).  workspace (device) holds at least
 * sb_bwd_workspace_bytes(blk) bytes: 13 floats per surfel.  Replaces the
 * autograd backward of the reference's torch ops in prepare_scaling_rot
 * (block_gaussian_model.py:198-256). */
size_t sb_bwd_workspace_bytes(const sb_blocks* blk);
int sb_generate_bwd(const sb_blocks* blk, const float* dL_dmeans3D, const float* dL_dscales,
                    const float* dL_drotations, const float* dL_dopacities, float* dL_dparams,
                    float* dL_doccupancy, void* workspace, size_t workspace_bytes, sr_stream_t stream);

/* sb_generate_bwd without a workspace and without any allocation: for
 * callers that cannot allocate (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The 13-float-per-surfel contributions go to a fixed scratch
 * array of 13 x 262144 floats (13.6 MB) that the library holds on every GPU
 * it runs on until the process ends; more than 262144 surfels (B x n_eta x
 * n_omega x K) return SR_ERR_INVALID.  Results equal sb_generate_bwd's bit
 * for bit.  Calls share the GPU's scratch: calls that overlap on one GPU
 * (other streams, other threads) race and must be ordered by the caller. */
int sb_generate_bwd_noalloc(const sb_blocks* blk, const float* dL_dmeans3D, const float* dL_dscales,
                            const float* dL_drotations, const float* dL_dopacities, float* dL_dparams,
                            float* dL_doccupancy, sr_stream_t stream);

/* ---- superquadric signed distance --------------------------------------------
 * implicit_sq(points_inv, eps1, eps2, as_sdf=2) with safe powers
 * (games/block_mesh_splatting/utils/superquadric.py:25-46) of N world points
 * in every block's frame, points_inv = ((p - T) @ R^T) / (S ratio)
 * (block_gaussian_model.py:529-531,613-614), S, R, T, eps from the B x 12
 * block parameters as in sb_blocks.  sdf is B x N (block-major).  The
 * backward returns dL/d(block parameters) (B x 12) for an upstream dL/dsdf,
 * summed per block in a fixed order (deterministic); the points get no
 * gradient (the reference samples them without one, :524-528,604-608).
 * workspace: sb_sdf_bwd_workspace_bytes(B, N) bytes of device memory. */
int sb_sdf_fwd(const float* params, int B, float ratio, float scale_min, const float* points, int N, float* sdf,
               sr_stream_t stream);
size_t sb_sdf_bwd_workspace_bytes(int B, int N);
int sb_sdf_bwd(const float* params, int B, float ratio, float scale_min, const float* points, int N,
               const float* dL_dsdf, float* dL_dparams, void* workspace, size_t workspace_bytes, sr_stream_t stream);

/* ---- simple-knn ------------------------------------------------------------
 * distCUDA2 (submodules/simple-knn/spatial.cu:16-27, SimpleKNN::knn at
 * simple_knn.cu:185-220): mean_dists[i] = ((b0 + b1) + b2) / 3 over the three
 * smallest FP32 squared distances (dx dx + dy dy) + dz dz from point i to the
 * other points (FLT_MAX for missing ones when P < 4).  points is P x 3
 * float32; the result depends on the point set only (exact search).
 * workspace: sk_dist2_workspace_bytes(P) bytes of device memory. */
size_t sk_dist2_workspace_bytes(int P);
int sk_dist2(int P, const float* points, float* mean_dists, void* workspace, size_t workspace_bytes,
             sr_stream_t stream);

const char* sr_last_error(void);
const char* sr_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SURFEL_RAST_H */
