// tgp_learn.hip -- the learned move proposer's kernels (no counterpart in the
// reference): a 2-layer mean-aggregation GNN in BF16 with FP32 accumulation
// on the gfx950 matrix cores, its training steps, and the proposal step the
// learned refinement admits through the exact integer gain and capacity.
//
//   features  x_v = [ T_v[q] / deg_v (q < k), onehot(part_v) (k), log2(1+deg_v)/16, 1, 0... ]
//             (T_v[q] = weight of v's arcs into part q), row stride F (multiple of 32)
//   layer 1   H = ReLU(X W1a + mean_N(X W1b) + b1)          (hidden 64)
//   layer 2   Z = H W2a + mean_N(H) W2b + b2                 (k logits)
//   target    the part the vertex holds after the plain pass (softmax cross-entropy)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tgp_learn.h"

namespace {

typedef int64_t i64;
typedef int32_t i32;
typedef uint16_t bf16;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int WAVE = 64;

__device__ __forceinline__ bf16 f2bf(float f) {  // round to nearest even
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------- BF16 GEMM --
// C[M x N] = alpha * op(A)[M x K] * op(B)[K x N] + beta * C (+ bias[N]) (ReLU)
// A, B in BF16; C in FP32 (and a BF16 copy in Cb when given).  op(A)(i, k) =
// transA ? A[k * lda + i] : A[i * lda + k]; likewise B.  Tiles of 64 x 64 per
// 256-thread block, each wave a 32 x 32 quarter on v_mfma_f32_32x32x8_bf16:
// lane l feeds A rows l % 32 and B columns l % 32 with k = 4 (l / 32) .. +3 of
// each 8-deep step and holds D rows 8 (r / 4) + 4 (l / 32) + r % 4, column
// l % 32, in accumulator r.
constexpr int GT = 64, GK = 32;

// Split K (work != nullptr): block z covers k in [z Kc, (z + 1) Kc) and
// stores its raw partial tile at work[z][M][N]; k_splitk_reduce adds the
// slices in order and applies the epilogue.  Tile loads walk along the
// operand's contiguous dimension.
__global__ __launch_bounds__(256) void k_gemm_bf16(const bf16* A, int lda, int transA, const bf16* B,
                                                   int ldb, int transB, float* C, int ldc, bf16* Cb,
                                                   int M, int N, int K, float alpha, float beta,
                                                   const float* bias, int relu, float* work,
                                                   int Kc) {
  __shared__ bf16 As[GT][GK + 4];
  __shared__ bf16 Bs[GK][GT + 4];
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  const int bm = blockIdx.x * GT, bn = blockIdx.y * GT;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int kbeg = work ? (int)blockIdx.z * Kc : 0;
  const int kend = work ? (kbeg + Kc < K ? kbeg + Kc : K) : K;
  v16f acc = {0};
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
    for (int e = t; e < GT * GK; e += 256) {
      const int i = transA ? e % GT : e / GK, k = transA ? e / GT : e % GK;
      const int gi = bm + i, gk = k0 + k;
      bf16 v = 0;
      if (gi < M && gk < kend) v = transA ? A[(i64)gk * lda + gi] : A[(i64)gi * lda + gk];
      As[i][k] = v;
    }
    for (int e = t; e < GK * GT; e += 256) {
      const int k = transB ? e % GK : e / GT, j = transB ? e / GK : e % GT;
      const int gk = k0 + k, gj = bn + j;
      bf16 v = 0;
      if (gk < kend && gj < N) v = transB ? B[(i64)gj * ldb + gk] : B[(i64)gk * ldb + gj];
      Bs[k][j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < GK; kc += 8) {
      v4s a, b;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int k = kc + 4 * (l >> 5) + q;
        a[q] = (short)As[wr + (l & 31)][k];
        b[q] = (short)Bs[k][wc + (l & 31)];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = bm + wr + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    const int j = bn + wc + (l & 31);
    if (work) {
      if (i < M && j < N) work[((i64)blockIdx.z * M + i) * N + j] = acc[r];
    } else if (i < M && j < N) {
      float v = alpha * acc[r];
      if (beta != 0.0f) v += beta * C[(i64)i * ldc + j];
      if (bias) v += bias[j];
      if (relu && v < 0.0f) v = 0.0f;
      C[(i64)i * ldc + j] = v;
      if (Cb) Cb[(i64)i * ldc + j] = f2bf(v);
    }
  }
}

__global__ void k_splitk_reduce(const float* work, int S, int M, int N, float* C, int ldc,
                                bf16* Cb, float alpha, float beta, const float* bias, int relu) {
  const i64 MN = (i64)M * N;
  for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < MN;
       e += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(e / N), j = (int)(e % N);
    float sum = 0.0f;
    for (int z = 0; z < S; z++) sum += work[z * MN + e];
    float v = alpha * sum;
    if (beta != 0.0f) v += beta * C[(i64)i * ldc + j];
    if (bias) v += bias[j];
    if (relu && v < 0.0f) v = 0.0f;
    C[(i64)i * ldc + j] = v;
    if (Cb) Cb[(i64)i * ldc + j] = f2bf(v);
  }
}

// ------------------------------------------------------------ aggregation --
// out[v] = (1/deg v) sum_{u in N(v)} X[u]   (transpose = 0), or
// out[u] = sum_{v in N(u)} X[v] / deg v     (transpose = 1: the adjoint of the
// mean on a symmetric graph).  BF16 in, FP32 sums, BF16 and optional FP32 out.
// Rows of more than AGG_HEAVY arcs go to a 1024-thread block each
// (k_mean_agg_heavy), the others to a wave each.  Lanes hold features; a
// row's neighbour ids (and, for the adjoint, their degrees) are loaded 64 at a
// time by the lanes and broadcast with readlane, so the feature loads of a
// step are independent of one another.
constexpr i64 AGG_HEAVY = 2048;
constexpr int AGG_U = 16;
constexpr int HB_WAVES = 16;  // waves of a heavy-row block

// Sum over the arcs base, base + stride, ... (64 consecutive arcs per step)
// of row [.., s1): X[u][f], or X[u][f] / deg u for the adjoint.
__device__ __forceinline__ float agg_range(const i64* xadj, const i32* adj, const bf16* X, int F,
                                           int f, bool fok, int transpose, i64 b, i64 s1,
                                           i64 stride, int l) {
  float s = 0.0f;
  for (i64 base = b; base < s1; base += stride) {
    const i64 i = base + l;
    const i32 myu = i < s1 ? adj[i] : 0;
    float mydu = 0.0f;
    if (transpose && i < s1) mydu = (float)(xadj[myu + 1] - xadj[myu]);
    const int cnt = (int)(s1 - base < WAVE ? s1 - base : WAVE);
    for (int j0 = 0; j0 < cnt; j0 += AGG_U) {
      float xs[AGG_U];
#pragma unroll
      for (int jj = 0; jj < AGG_U; jj++) {
        xs[jj] = 0.0f;
        if (j0 + jj < cnt) {
          const i32 u = __builtin_amdgcn_readlane(myu, j0 + jj);
          if (fok) xs[jj] = bf2f(X[(i64)u * F + f]);
        }
      }
#pragma unroll
      for (int jj = 0; jj < AGG_U; jj++) {
        if (j0 + jj < cnt) {
          if (transpose) {
            const float du =
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mydu), j0 + jj));
            s += du > 0.0f ? xs[jj] / du : 0.0f;
          } else {
            s += xs[jj];
          }
        }
      }
    }
  }
  return s;
}

__global__ __launch_bounds__(256) void k_mean_agg(i64 n, const i64* xadj, const i32* adj,
                                                  const bf16* X, int F, int transpose, bf16* out,
                                                  float* outf) {
  const int l = threadIdx.x & 63;
  for (i64 v = (i64)blockIdx.x * 4 + (threadIdx.x >> 6); v < n; v += (i64)gridDim.x * 4) {
    const i64 s0 = xadj[v], s1 = xadj[v + 1];
    if (s1 - s0 > AGG_HEAVY) continue;
    for (int f0 = 0; f0 < F; f0 += WAVE) {
      const int f = f0 + l;
      const bool fok = f < F;
      float s = agg_range(xadj, adj, X, F, f, fok, transpose, s0, s1, WAVE, l);
      if (fok) {
        if (!transpose && s1 > s0) s /= (float)(s1 - s0);
        out[(i64)v * F + f] = f2bf(s);
        if (outf) outf[(i64)v * F + f] = s;
      }
    }
  }
}

// F = 64 (the hidden width): four vertices per wave, 16 lanes each holding 4
// consecutive features (one 8-byte load per neighbour), neighbour ids shared
// within the 16 lanes; four neighbours' loads in flight per step.
constexpr int AGG4_U = 4;
__global__ __launch_bounds__(256) void k_mean_agg64(i64 n, const i64* xadj, const i32* adj,
                                                    const bf16* X, int transpose, bf16* out,
                                                    float* outf) {
  const int gl = threadIdx.x & 15;
  const i64 ngroups = (i64)gridDim.x * (blockDim.x / 16);
  for (i64 v = (i64)blockIdx.x * (blockDim.x / 16) + threadIdx.x / 16; v < n; v += ngroups) {
    const i64 s0 = xadj[v], s1 = xadj[v + 1];
    if (s1 - s0 > AGG_HEAVY) continue;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (i64 base = s0; base < s1; base += 16) {
      const i64 i = base + gl;
      const i32 myu = i < s1 ? adj[i] : 0;
      float mydu = 0.0f;
      if (transpose && i < s1) mydu = (float)(xadj[myu + 1] - xadj[myu]);
      const int cnt = (int)(s1 - base < 16 ? s1 - base : 16);
      for (int j0 = 0; j0 < cnt; j0 += AGG4_U) {
        uint2 xs[AGG4_U];
        float dus[AGG4_U];
#pragma unroll
        for (int jj = 0; jj < AGG4_U; jj++) {
          const int j = j0 + jj < cnt ? j0 + jj : cnt - 1;
          const i32 u = __shfl(myu, j, 16);
          dus[jj] = __shfl(mydu, j, 16);
          xs[jj] = *reinterpret_cast<const uint2*>(X + (i64)u * 64 + 4 * gl);
        }
#pragma unroll
        for (int jj = 0; jj < AGG4_U; jj++) {
          if (j0 + jj >= cnt) break;
          const float x0 = __uint_as_float(xs[jj].x << 16);
          const float x1 = __uint_as_float(xs[jj].x & 0xffff0000u);
          const float x2 = __uint_as_float(xs[jj].y << 16);
          const float x3 = __uint_as_float(xs[jj].y & 0xffff0000u);
          if (transpose) {
            const float du = dus[jj];
            a0 += du > 0.0f ? x0 / du : 0.0f;
            a1 += du > 0.0f ? x1 / du : 0.0f;
            a2 += du > 0.0f ? x2 / du : 0.0f;
            a3 += du > 0.0f ? x3 / du : 0.0f;
          } else {
            a0 += x0;
            a1 += x1;
            a2 += x2;
            a3 += x3;
          }
        }
      }
    }
    if (!transpose && s1 > s0) {
      const float d = (float)(s1 - s0);
      a0 /= d;
      a1 /= d;
      a2 /= d;
      a3 /= d;
    }
    const i64 o = v * 64 + 4 * gl;
    out[o] = f2bf(a0);
    out[o + 1] = f2bf(a1);
    out[o + 2] = f2bf(a2);
    out[o + 3] = f2bf(a3);
    if (outf) *reinterpret_cast<float4*>(outf + o) = make_float4(a0, a1, a2, a3);
  }
}

// Heavy rows: block b takes the vertices b, b + G, b + 2G, ... (G blocks,
// so neighbouring hub ids land in different blocks), 1024 of them per round;
// its 16 waves split each row over AGG_HEAVY arcs and the partial sums are
// added in wave order.
__global__ __launch_bounds__(1024) void k_mean_agg_heavy(i64 n, const i64* xadj, const i32* adj,
                                                         const bf16* X, int F, int transpose,
                                                         bf16* out, float* outf) {
  __shared__ i32 hv[1024];
  __shared__ int nh;
  __shared__ float ps[HB_WAVES][WAVE];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const i64 G = gridDim.x;
  for (i64 r0 = 0; (i64)blockIdx.x + r0 * G < n; r0 += blockDim.x) {
    if (threadIdx.x == 0) nh = 0;
    __syncthreads();
    {
      const i64 v = (i64)blockIdx.x + (r0 + threadIdx.x) * G;
      if (v < n && xadj[v + 1] - xadj[v] > AGG_HEAVY) hv[atomicAdd(&nh, 1)] = (i32)v;
    }
    __syncthreads();
    const int cnt = nh;
    for (int h = 0; h < cnt; h++) {
      const i64 v = hv[h];
      const i64 s0 = xadj[v], s1 = xadj[v + 1];
      for (int f0 = 0; f0 < F; f0 += WAVE) {
        const int f = f0 + l;
        const bool fok = f < F;
        ps[wv][l] = agg_range(xadj, adj, X, F, f, fok, transpose, s0 + (i64)wv * WAVE, s1,
                              (i64)HB_WAVES * WAVE, l);
        __syncthreads();
        if (wv == 0 && fok) {
          float t = 0.0f;
          for (int w2 = 0; w2 < HB_WAVES; w2++) t += ps[w2][l];
          if (!transpose) t /= (float)(s1 - s0);
          out[(i64)v * F + f] = f2bf(t);
          if (outf) outf[(i64)v * F + f] = t;
        }
        __syncthreads();
      }
    }
  }
}

// --------------------------------------------------------------- features --
// Part tallies T_v (integer LDS atomics, exact) and deg_v = sum of v's arc
// weights, then the row of X.  Rows of more than FEAT_HEAVY arcs go to a
// 1024-thread block each (k_features_heavy), the others to a wave each.
constexpr int FEAT_MAXK = 256;
constexpr i64 FEAT_HEAVY = 2048;

__device__ __forceinline__ void feature_row(bf16* row, int F, int k, const i32* tl, i64 deg,
                                            i32 cur, i64 d, i32* tally_row, int l) {
  const float degf = (float)deg;
  for (int f = l; f < F; f += WAVE) {
    float x = 0.0f;
    if (f < k) x = deg > 0 ? (float)tl[f] / degf : 0.0f;
    else if (f < 2 * k) x = (f - k) == cur ? 1.0f : 0.0f;
    else if (f == 2 * k) x = log2f(1.0f + (float)d) / 16.0f;
    else if (f == 2 * k + 1) x = 1.0f;
    row[f] = f2bf(x);
  }
  if (tally_row)
    for (int q = l; q < k; q += WAVE) tally_row[q] = tl[q];
}

__global__ __launch_bounds__(256) void k_features(i64 n, const i64* xadj, const i32* adj,
                                                  const i32* w, const i32* part, int k, int F,
                                                  bf16* X, i32* tally) {
  __shared__ i32 tl[4][FEAT_MAXK];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (i64 v = (i64)blockIdx.x * 4 + wv; v < n; v += (i64)gridDim.x * 4) {
    const i64 s0 = xadj[v], s1 = xadj[v + 1];
    if (s1 - s0 > FEAT_HEAVY) continue;
    for (int q = l; q < k; q += WAVE) tl[wv][q] = 0;
    __builtin_amdgcn_wave_barrier();
    i64 deg = 0;
    for (i64 i = s0 + l; i < s1; i += WAVE) {
      const i32 p = part[adj[i]];
      const i32 wt = w[i];
      deg += wt;
      if (p >= 0 && p < k) atomicAdd(&tl[wv][p], wt);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o);
    __builtin_amdgcn_wave_barrier();
    feature_row(X + v * F, F, k, tl[wv], deg, part[v], s1 - s0, tally ? tally + v * k : nullptr,
                l);
    __builtin_amdgcn_wave_barrier();
  }
}

// Heavy rows, distributed over blocks as in k_mean_agg_heavy.
__global__ __launch_bounds__(1024) void k_features_heavy(i64 n, const i64* xadj, const i32* adj,
                                                         const i32* w, const i32* part, int k,
                                                         int F, bf16* X, i32* tally) {
  __shared__ i32 hv[1024];
  __shared__ int nh;
  __shared__ i32 tl[FEAT_MAXK];
  __shared__ i64 dg[HB_WAVES];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const i64 G = gridDim.x;
  for (i64 r0 = 0; (i64)blockIdx.x + r0 * G < n; r0 += blockDim.x) {
    if (threadIdx.x == 0) nh = 0;
    __syncthreads();
    {
      const i64 v = (i64)blockIdx.x + (r0 + threadIdx.x) * G;
      if (v < n && xadj[v + 1] - xadj[v] > FEAT_HEAVY) hv[atomicAdd(&nh, 1)] = (i32)v;
    }
    __syncthreads();
    const int cnt = nh;
    for (int h = 0; h < cnt; h++) {
      const i64 v = hv[h];
      const i64 s0 = xadj[v], s1 = xadj[v + 1];
      for (int q = threadIdx.x; q < k; q += blockDim.x) tl[q] = 0;
      __syncthreads();
      i64 deg = 0;
      for (i64 i = s0 + threadIdx.x; i < s1; i += blockDim.x) {
        const i32 p = part[adj[i]];
        const i32 wt = w[i];
        deg += wt;
        if (p >= 0 && p < k) atomicAdd(&tl[p], wt);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o);
      if (l == 0) dg[wv] = deg;
      __syncthreads();
      if (wv == 0) {
        i64 t = 0;
        for (int w2 = 0; w2 < HB_WAVES; w2++) t += dg[w2];
        feature_row(X + v * F, F, k, tl, t, part[v], s1 - s0, tally ? tally + v * k : nullptr, l);
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------- loss, grads --
// Softmax cross-entropy over k logits per row with class weights (wstay for
// rows whose label equals their current part, wmove otherwise); dZ = weight *
// (softmax - onehot) / sum of weights.  acc[0] += weighted loss, acc[1] +=
// weights, acc[2] += correct predictions, acc[3] += rows.
__global__ __launch_bounds__(256) void k_softmax_xent(i64 n, int k, const float* Z, const i32* label,
                                                      const i32* cur, float wstay, float wmove,
                                                      float inv_wsum, float* dZ, bf16* dZb,
                                                      double* acc) {
  double lsum = 0, wsum = 0, corr = 0, rows = 0;
  for (i64 v = (i64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (i64)gridDim.x * blockDim.x) {
    const float* z = Z + v * k;
    float mx = z[0];
    int am = 0;
    for (int q = 1; q < k; q++)
      if (z[q] > mx) {
        mx = z[q];
        am = q;
      }
    float se = 0.0f;
    for (int q = 0; q < k; q++) se += __expf(z[q] - mx);
    const i32 y = label[v];
    const float wt = y == cur[v] ? wstay : wmove;
    lsum += (double)wt * (double)(logf(se) + mx - z[y]);
    wsum += wt;
    corr += am == y;
    rows += 1;
    if (dZ) {
      for (int q = 0; q < k; q++) {
        const float g = wt * inv_wsum * (__expf(z[q] - mx) / se - (q == y ? 1.0f : 0.0f));
        dZ[v * k + q] = g;
        if (dZb) dZb[v * k + q] = f2bf(g);
      }
    }
  }
  atomicAdd(&acc[0], lsum);
  atomicAdd(&acc[1], wsum);
  atomicAdd(&acc[2], corr);
  atomicAdd(&acc[3], rows);
}

// dH *= (H > 0), also the BF16 copy
__global__ void k_relu_bwd(i64 cnt, const float* H, float* dH, bf16* dHb) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (i64)gridDim.x * blockDim.x) {
    const float g = H[i] > 0.0f ? dH[i] : 0.0f;
    dH[i] = g;
    if (dHb) dHb[i] = f2bf(g);
  }
}

// db[j] = sum_i dY[i][j]
__global__ void k_colsum(i64 rows, int cols, const float* dY, float* db) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < cols; j += gridDim.x * blockDim.x) {
    double s = 0;
    for (i64 i = 0; i < rows; i++) s += dY[i * cols + j];
    db[j] = (float)s;
  }
}

// Adam on FP32 master weights, BF16 copy refreshed
__global__ void k_adam(i64 cnt, float* W, const float* G, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2, bf16* Wb) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (i64)gridDim.x * blockDim.x) {
    const float g = G[i];
    m[i] = b1 * m[i] + (1.0f - b1) * g;
    v[i] = b2 * v[i] + (1.0f - b2) * g * g;
    W[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    if (Wb) Wb[i] = f2bf(W[i]);
  }
}

__global__ void k_axpy(i64 cnt, float a, const float* x, float* y) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (i64)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void k_to_bf16(i64 cnt, const float* x, bf16* y) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (i64)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// ----------------------------------------------------------------- proposals --
// The learned proposal of each vertex: q* = argmax over parts other than its
// own (lowest q on ties), proposed when its logit is within `margin` of the
// own part's; admitted when the exact integer gain T_v[q*] - T_v[part_v] is
// positive.  gain[v] = that gain (as float) for an admitted
// proposal, NaN otherwise; target[v] = q*.
__global__ void k_propose(i64 n, int k, const float* Z, const i32* part, const i32* tally,
                          float margin, float* gain, i32* target) {
  for (i64 v = (i64)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (i64)gridDim.x * blockDim.x) {
    const i32 cur = part[v];
    const float* z = Z + v * k;
    int best = -1;
    float bz = 0.0f;
    for (int q = 0; q < k; q++) {
      if (q == cur) continue;
      if (best < 0 || z[q] > bz) {
        best = q;
        bz = z[q];
      }
    }
    float g = __int_as_float(0x7fc00000);  // NaN: not a candidate
    if (best >= 0 && cur >= 0 && cur < k && z[best] > z[cur] - margin) {
      const i32 d = tally[v * k + best] - tally[v * k + cur];
      if (d > 0) g = (float)d;
    }
    gain[v] = g;
    target[v] = best < 0 ? cur : best;
  }
}

unsigned grid(i64 work, int per_block, unsigned cap = 65536) {
  i64 b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (unsigned)(b > cap ? cap : b);
}

}  // namespace

extern "C" {

int tgpl_gemm_ws(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB,
                 float* C, int ldc, uint16_t* Cb, int M, int N, int K, float alpha, float beta,
                 const float* bias, int relu, float* work, int64_t work_elems, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0) return -1;
  dim3 g((unsigned)((M + GT - 1) / GT), (unsigned)((N + GT - 1) / GT));
  if (g.y > 65535) return -1;
  // split K when the output has few tiles and K is long: about 4096 per slice,
  // at most 512 slices and what the workspace holds
  const i64 tiles = (i64)g.x * g.y;
  i64 S = 1;
  if (work && tiles < 256 && K > 8192) {
    S = ((i64)K + 4095) / 4096;
    if (S > 512) S = 512;
    if (S > work_elems / ((i64)M * N)) S = work_elems / ((i64)M * N);
  }
  if (S <= 1) {
    k_gemm_bf16<<<g, 256, 0, (hipStream_t)stream>>>(A, lda, transA, B, ldb, transB, C, ldc, Cb,
                                                     M, N, K, alpha, beta, bias, relu, nullptr, 0);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  i64 Kc = ((i64)K + S - 1) / S;
  Kc = (Kc + GK - 1) / GK * GK;
  S = ((i64)K + Kc - 1) / Kc;
  g.z = (unsigned)S;
  k_gemm_bf16<<<g, 256, 0, (hipStream_t)stream>>>(A, lda, transA, B, ldb, transB, C, ldc, Cb, M,
                                                   N, K, alpha, beta, bias, relu, work, (int)Kc);
  k_splitk_reduce<<<grid((i64)M * N, 256, 4096), 256, 0, (hipStream_t)stream>>>(
      work, (int)S, M, N, C, ldc, Cb, alpha, beta, bias, relu);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_gemm(const uint16_t* A, int lda, int transA, const uint16_t* B, int ldb, int transB,
              float* C, int ldc, uint16_t* Cb, int M, int N, int K, float alpha, float beta,
              const float* bias, int relu, void* stream) {
  return tgpl_gemm_ws(A, lda, transA, B, ldb, transB, C, ldc, Cb, M, N, K, alpha, beta, bias,
                      relu, nullptr, 0, stream);
}

int tgpl_mean_agg(int64_t n, const int64_t* xadj, const int32_t* adj, const uint16_t* X, int F,
                  int transpose, uint16_t* out, float* outf, void* stream) {
  if (n <= 0 || F <= 0) return -1;
  if (F == 64)
    k_mean_agg64<<<grid(n, 16), 256, 0, (hipStream_t)stream>>>(n, xadj, adj, X, transpose, out,
                                                                outf);
  else
    k_mean_agg<<<grid(n, 4), 256, 0, (hipStream_t)stream>>>(n, xadj, adj, X, F, transpose, out,
                                                             outf);
  k_mean_agg_heavy<<<grid(n, 1024, 2048), 1024, 0, (hipStream_t)stream>>>(n, xadj, adj, X, F,
                                                                          transpose, out, outf);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_features(int64_t n, const int64_t* xadj, const int32_t* adj, const int32_t* w,
                  const int32_t* part, int k, int F, uint16_t* X, int32_t* tally, void* stream) {
  if (n <= 0 || k < 1 || k > FEAT_MAXK || F < 2 * k + 2) return -1;
  k_features<<<grid(n, 4), 256, 0, (hipStream_t)stream>>>(n, xadj, adj, w, part, k, F, X, tally);
  k_features_heavy<<<grid(n, 1024, 2048), 1024, 0, (hipStream_t)stream>>>(n, xadj, adj, w, part,
                                                                          k, F, X, tally);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_softmax_xent(int64_t n, int k, const float* Z, const int32_t* label, const int32_t* cur,
                      float wstay, float wmove, float inv_wsum, float* dZ, uint16_t* dZb,
                      double* acc, void* stream) {
  if (n <= 0 || k <= 0) return -1;
  k_softmax_xent<<<grid(n, 256, 4096), 256, 0, (hipStream_t)stream>>>(
      n, k, Z, label, cur, wstay, wmove, inv_wsum, dZ, dZb, acc);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_relu_bwd(int64_t cnt, const float* H, float* dH, uint16_t* dHb, void* stream) {
  k_relu_bwd<<<grid(cnt, 256), 256, 0, (hipStream_t)stream>>>(cnt, H, dH, dHb);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_colsum(int64_t rows, int cols, const float* dY, float* db, void* stream) {
  k_colsum<<<grid(cols, 64), 64, 0, (hipStream_t)stream>>>(rows, cols, dY, db);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_adam(int64_t cnt, float* W, const float* G, float* m, float* v, float lr, float b1,
              float b2, float eps, int step, uint16_t* Wb, void* stream) {
  float c1 = 1.0f, c2 = 1.0f;
  for (int i = 0; i < step; i++) {
    c1 *= b1;
    c2 *= b2;
  }
  k_adam<<<grid(cnt, 256), 256, 0, (hipStream_t)stream>>>(cnt, W, G, m, v, lr, b1, b2, eps,
                                                          1.0f - c1, 1.0f - c2, Wb);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_axpy(int64_t cnt, float a, const float* x, float* y, void* stream) {
  k_axpy<<<grid(cnt, 256), 256, 0, (hipStream_t)stream>>>(cnt, a, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_to_bf16(int64_t cnt, const float* x, uint16_t* y, void* stream) {
  k_to_bf16<<<grid(cnt, 256), 256, 0, (hipStream_t)stream>>>(cnt, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tgpl_propose(int64_t n, int k, const float* Z, const int32_t* part, const int32_t* tally,
                 float margin, float* gain, int32_t* target, void* stream) {
  k_propose<<<grid(n, 256), 256, 0, (hipStream_t)stream>>>(n, k, Z, part, tally, margin, gain,
                                                            target);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
