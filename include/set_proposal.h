/*
 * set_proposal.h -- C ABI of the learned block proposal's point-set MLP on
 * MI355X matrix cores (bf16 operands, FP32 accumulation).
 *
 * No counterpart in the reference: PartGS initialises its blocks with
 * k-means over the point cloud (games/block_mesh_splatting/scene/
 * dataset_readers.py:487-503, restated in partgs_amd/init.py).  The proposal
 * network predicts the block parameters from the points instead:
 * per point 3 -> 64 -> 128 -> 256 (ReLU), max-pool over the scene, then
 * 256 -> 128 (ReLU) -> 16 x 12 block parameters.
 *
 * All pointers are device memory; bf16 values are raw uint16 bit patterns
 * (round to nearest even).  Work is enqueued on `stream`; functions return 0,
 * 1 for a bad argument, 2 for a failed launch.
 */
#ifndef SET_PROPOSAL_H
#define SET_PROPOSAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* sp_stream_t; /* == hipStream_t */

/* C[m * ldc + n] = act(alpha * sum_k A(m, k) B(n, k) + bias[n]) (+ C[m * ldc + n]
 * when accumulate), with A(m, k) = A[m * lda_m + k * lda_k] and
 * B(n, k) = B[n * ldb_n + k * ldb_k] in bf16; act = ReLU when relu != 0;
 * bias may be NULL.  C_bf16, when not NULL, receives a bf16 copy of C. */
int sp_gemm_bf16(int M, int N, int K, const uint16_t* A, long lda_m, long lda_k, const uint16_t* B, long ldb_n,
                 long ldb_k, const float* bias, float alpha, int relu, int accumulate, float* C, long ldc,
                 uint16_t* C_bf16, sp_stream_t stream);

/* Shared per-point MLP fused with the max-pool over each scene's N points:
 * value[b * 256 + c] = max_p relu(layer3(p))[c] and index[...] the first p
 * attaining it.  pts B x N x 3 (FP32), W1 64 x 3 and b1 (FP32), W2 128 x 64
 * (bf16) and b2, W3 256 x 128 (bf16) and b3.  scratch holds B x 256 words. */
int sp_point_pool(int B, int N, const float* pts, const float* W1, const float* b1, const uint16_t* W2,
                  const float* b2, const uint16_t* W3, const float* b3, float* value, int* index,
                  unsigned long long* scratch, sp_stream_t stream);

/* y = bf16(x) for n values. */
int sp_to_bf16(const float* x, uint16_t* y, long n, sp_stream_t stream);

/* One Adam step (bias-corrected, step >= 1) on FP32 parameters; param_bf16,
 * when not NULL, receives the updated values in bf16. */
int sp_adam(float* param, uint16_t* param_bf16, const float* grad, float* m, float* v, long n, float lr,
            float beta1, float beta2, float eps, int step, sp_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SET_PROPOSAL_H */
