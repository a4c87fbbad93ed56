/* tgp_learn.h -- kernels of the learned move proposer (no counterpart in the
 * reference, whose refinement is gain-driven only, gpu_part.cu:258-397).
 *
 * A 2-layer mean-aggregation GNN over per-vertex features, BF16 weights and
 * activations with FP32 accumulation on the matrix cores:
 *   x_v = [T_v[q]/deg_v (q < k), onehot(part_v) (k), log2(1+deg_v)/16, 1, 0 pad] (F columns)
 *   H   = ReLU(X W1a + mean_N(X W1b) + b1)      hidden 64
 *   Z   = H W2a + mean_N(H) W2b + b2            k logits
 * trained on the plain refinement's own passes (label: the part a vertex holds
 * after the pass).  Proposals are admitted only through the exact integer
 * gain T_v[q] - T_v[part_v] > 0 (tgpl_propose) and the reference's capacity
 * rule in the learned refinement (tgp_partition_learned, tgp.h).
 *
 * All pointers are device pointers; `stream` is a hipStream_t (NULL: the
 * default stream).  Return 0 on success, negative on bad sizes or a launch
 * error.  BF16 values are passed as uint16_t bit patterns.                  */
#ifndef TGP_LEARN_H
#define TGP_LEARN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* C[MxN] = alpha op(A) op(B) + beta C (+ bias[N]) (ReLU); A, B BF16; C FP32
 * and, when Cb is given, a BF16 copy.  op(A)(i,k) = transA ? A[k*lda+i] :
 * A[i*lda+k]; likewise B.                                                    */
int tgpl_gemm(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB,
              float* C, int ldc, uint16_t* Cb, int M, int N, int K, float alpha, float beta,
              const float* bias, int relu, void* stream);
/* The same product with a workspace of work_elems floats: when the output
 * has few 64 x 64 tiles and K is long (the weight gradients, K = vertices),
 * K is split into slices of about 4096 (at most 512 and work_elems / (M N)
 * of them) whose partial products are added in slice order.              */
int tgpl_gemm_ws(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB,
                 float* C, int ldc, uint16_t* Cb, int M, int N, int K, float alpha, float beta,
                 const float* bias, int relu, float* work, int64_t work_elems, void* stream);
/* out = mean over neighbours of X (transpose 0) or its adjoint on a
 * symmetric graph, out[u] = sum_{v in N(u)} X[v]/deg v (transpose 1).        */
int tgpl_mean_agg(int64_t n, const int64_t* xadj, const int32_t* adj, const uint16_t* X, int F,
                  int transpose, uint16_t* out, float* outf, void* stream);
/* X (n x F, BF16) from the partition; tally (optional, n x k) = T_v[q].
 * k <= 256, F >= 2k + 2.                                                     */
int tgpl_features(int64_t n, const int64_t* xadj, const int32_t* adj, const int32_t* w,
                  const int32_t* part, int k, int F, uint16_t* X, int32_t* tally, void* stream);
/* weighted softmax cross-entropy; dZ (optional) = gradient; acc[4] += (loss,
 * weights, correct, rows).                                                   */
int tgpl_softmax_xent(int64_t n, int k, const float* Z, const int32_t* label, const int32_t* cur,
                      float wstay, float wmove, float inv_wsum, float* dZ, uint16_t* dZb,
                      double* acc, void* stream);
int tgpl_relu_bwd(int64_t cnt, const float* H, float* dH, uint16_t* dHb, void* stream);
int tgpl_colsum(int64_t rows, int cols, const float* dY, float* db, void* stream);
int tgpl_adam(int64_t cnt, float* W, const float* G, float* m, float* v, float lr, float b1,
              float b2, float eps, int step, uint16_t* Wb, void* stream);
int tgpl_to_bf16(int64_t cnt, const float* x, uint16_t* y, void* stream);
/* y += a x */
int tgpl_axpy(int64_t cnt, float a, const float* x, float* y, void* stream);
/* gain[v] = exact integer gain of v's best non-own part q when it is
 * positive and logit(q) > logit(part_v) - margin, NaN otherwise; target[v] =
 * q.  tally as written by tgpl_features.                                     */
int tgpl_propose(int64_t n, int k, const float* Z, const int32_t* part, const int32_t* tally,
                 float margin, float* gain, int32_t* target, void* stream);

#ifdef __cplusplus
}
#endif
#endif
