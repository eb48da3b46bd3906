// rfd_mpnn.hip -- kernels of the learned FIRE warm start: a per-particle
// message-passing network over the bond graph, trained and run on the device.
//
// Dense layers are BF16 MFMA tiles (v_mfma_f32_16x16x32_bf16, FP32
// accumulate): operands are FP32 in memory, rounded to BF16 (nearest even) as
// they are loaded, products accumulated in FP32.  Weight gradients are the
// transposed product over all rows, computed as per-workgroup MFMA partials
// reduced in a fixed order, so training is deterministic.  The message step
// works on the half-bond rows of the native tier: one wave per particle, one
// lane per hidden feature.
#include "rfd_internal.h"

using namespace rfd;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMmRows = 64;     // rows per GEMM workgroup (4 waves x 16)
constexpr int kMmMaxN = 128;    // output columns per GEMM (multiple of 16)
constexpr int kMmMaxK = 128;    // reduction length per GEMM (multiple of 32)
constexpr int kTnChunk = 2048;  // rows per workgroup of the transposed product
constexpr int kHid = 64;        // hidden width of the message step

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// Y[M x N] = X[M x K] . B[K x N] + bias (+ Y), B(k, c) = W[k * wsk + c * wsc]
__global__ void __launch_bounds__(256) k_mm(const float *__restrict__ X, uint32_t ldx,
                                            const float *__restrict__ W, uint32_t wsk,
                                            uint32_t wsc, const float *__restrict__ bias,
                                            float *__restrict__ Y, uint32_t ldy, uint32_t M,
                                            uint32_t K, uint32_t Nc, int accumulate) {
  // B staged transposed (column-major, K contiguous, padded against bank
  // conflicts) so that a lane's 8 consecutive k are one 16-byte LDS read
  __shared__ __attribute__((aligned(16))) __bf16 sB[kMmMaxN * (kMmMaxK + 8)];
  const uint32_t Kp = K + 8;
  for (uint32_t q = threadIdx.x; q < K * Nc; q += blockDim.x) {
    const uint32_t k = q / Nc, c = q % Nc;
    sB[c * Kp + k] = to_bf16(W[(size_t)k * wsk + (size_t)c * wsc]);
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t row0 = blockIdx.x * kMmRows + wave * 16;
  const uint32_t ar = row0 + (lane & 15), ak = (lane >> 4) * 8;
  const uint32_t ntile = Nc / 16;
  f32x4 acc[kMmMaxN / 16];
#pragma unroll
  for (int t = 0; t < kMmMaxN / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a;
    if (ar < M) {
      const float *x = X + (size_t)ar * ldx + k0 + ak;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = to_bf16(x[e]);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = to_bf16(0.f);
    }
#pragma unroll
    for (int t = 0; t < kMmMaxN / 16; ++t) {
      if (t < (int)ntile) {
        const bf16x8 b =
            *reinterpret_cast<const bf16x8 *>(&sB[(t * 16 + (lane & 15)) * Kp + k0 + ak]);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // D(i, j): lane j + 16 (i / 4), element i % 4
  const uint32_t dr = row0 + (lane >> 4) * 4, dc = lane & 15;
#pragma unroll
  for (int t = 0; t < kMmMaxN / 16; ++t) {
    if (t >= (int)ntile) continue;
    const uint32_t c = t * 16 + dc;
    const float bc = bias ? bias[c] : 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t r = dr + e;
      if (r >= M) continue;
      float v = acc[t][e] + bc;
      float *y = Y + (size_t)r * ldy + c;
      if (accumulate) v += *y;
      *y = v;
    }
  }
}

// partial G[K x N] = X^T . D over one chunk of rows, per workgroup
__global__ void __launch_bounds__(256) k_mm_tn(const float *__restrict__ X, uint32_t ldx,
                                               const float *__restrict__ D, uint32_t ldd,
                                               float *__restrict__ part, uint32_t M, uint32_t K,
                                               uint32_t Nc) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t m_begin = blockIdx.x * kTnChunk;
  const uint32_t m_end = min(M, m_begin + kTnChunk);
  const uint32_t nkt = K / 16, nct = Nc / 16;
  // wave w owns k-tiles w and w + 4 (K <= 128)
  f32x4 acc[2][kMmMaxN / 16];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int t = 0; t < kMmMaxN / 16; ++t) acc[a][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint32_t mo = (lane >> 4) * 8;
  for (uint32_t m0 = m_begin; m0 < m_end; m0 += 32) {
    bf16x8 bfr[kMmMaxN / 16];
#pragma unroll
    for (int t = 0; t < kMmMaxN / 16; ++t) {
      if (t >= (int)nct) continue;
      const uint32_t c = t * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint32_t m = m0 + mo + e;
        bfr[t][e] = to_bf16(m < m_end ? D[(size_t)m * ldd + c] : 0.f);
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t kt = wave + 4 * a;
      if (kt >= nkt) continue;
      const uint32_t k = kt * 16 + (lane & 15);
      bf16x8 afr;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint32_t m = m0 + mo + e;
        afr[e] = to_bf16(m < m_end ? X[(size_t)m * ldx + k] : 0.f);
      }
#pragma unroll
      for (int t = 0; t < kMmMaxN / 16; ++t)
        if (t < (int)nct) acc[a][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[t], acc[a][t], 0, 0, 0);
    }
  }
  float *out = part + (size_t)blockIdx.x * K * Nc;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const uint32_t kt = wave + 4 * a;
    if (kt >= nkt) continue;
#pragma unroll
    for (int t = 0; t < kMmMaxN / 16; ++t) {
      if (t >= (int)nct) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t k = kt * 16 + (lane >> 4) * 4 + e, c = t * 16 + (lane & 15);
        out[(size_t)k * Nc + c] = acc[a][t][e];
      }
    }
  }
}

// out[i] (+)= sum over blocks b of part[b * n + i], in block order
__global__ void k_reduce_parts(const float *__restrict__ part, uint32_t nblocks, uint32_t n,
                               float *__restrict__ out, int accumulate) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    float s = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) s += part[(size_t)b * n + i];
    out[i] = accumulate ? out[i] + s : s;
  }
}

// column sums of D over a chunk of rows per workgroup (one thread per column)
__global__ void k_colsum(const float *__restrict__ D, uint32_t ldd, uint32_t M, uint32_t nc,
                         float *__restrict__ part) {
  const uint32_t m_begin = blockIdx.x * kTnChunk, m_end = min(M, m_begin + kTnChunk);
  for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
    float s = 0.f;
    for (uint32_t m = m_begin; m < m_end; ++m) s += D[(size_t)m * ldd + c];
    part[(size_t)blockIdx.x * nc + c] = s;
  }
}

// elementwise ops over an M x ncols block with row strides
// (out may alias a or b: each element is read before it is written)
__global__ void k_ewise(int op, uint32_t M, uint32_t nc, float alpha, const float *a,
                        uint32_t lda, const float *b, uint32_t ldb, float *out, uint32_t ldo) {
  const size_t n = (size_t)M * nc;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
       q += (size_t)blockDim.x * gridDim.x) {
    const size_t i = q / nc, c = q % nc;
    const float x = a[i * lda + c];
    float y;
    switch (op) {
      case 0: y = fmaxf(x, 0.f); break;                          // relu
      case 1: y = b[i * ldb + c] + fmaxf(x, 0.f); break;         // residual relu
      case 2: y = b[i * ldb + c] > 0.f ? x : 0.f; break;         // relu backward
      case 3: y = x + b[i * ldb + c]; break;                     // add
      case 5: y = alpha * (x + b[i * ldb + c]); break;           // scaled sum
      case 6: y = alpha * x - b[i * ldb + c]; break;             // scaled difference
      case 7: y = alpha * x; break;                              // scale
      default: y = x; break;                                     // copy
    }
    out[i * ldo + c] = y;
  }
}

// message step: for particle i, a_i = sum over its halves h (partner j) of
// relu(z_j + e_h . We + b); one wave per particle, lane f = hidden feature
__global__ void __launch_bounds__(256) k_edge_fwd(uint32_t N, const uint32_t *__restrict__ hoff,
                                                  const uint32_t *__restrict__ hj, uint32_t hjs,
                                                  const float *__restrict__ E,
                                                  const float *__restrict__ We,
                                                  const float *__restrict__ bm,
                                                  const float *__restrict__ Z,
                                                  float *__restrict__ Q, float *__restrict__ A,
                                                  uint32_t lda) {
  const uint32_t f = threadIdx.x & 63;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  const float w0 = We[f], w1 = We[kHid + f], w2 = We[2 * kHid + f], w3 = We[3 * kHid + f];
  const float bf = bm[f];
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < N; i += waves) {
    float a = 0.f;
    for (uint32_t h = hoff[i]; h < hoff[i + 1]; ++h) {
      const uint32_t j = hj[(size_t)h * hjs];
      const float4 e = *reinterpret_cast<const float4 *>(E + 4 * (size_t)h);
      const float q = Z[(size_t)j * kHid + f] + e.x * w0 + e.y * w1 + e.z * w2 + e.w * w3 + bf;
      if (Q) Q[(size_t)h * kHid + f] = q;
      a += fmaxf(q, 0.f);
    }
    A[(size_t)i * lda + f] = a;
  }
}

// backward of the message step: dq_h = dA_i (q_h > 0) for h in row i;
// dZ_j = sum over halves h' of row j of dq at the mirror half of h'; per
// workgroup partials of dWe (4 x 64) and db (64) in part[block][5][64]
__global__ void __launch_bounds__(256) k_edge_bwd(uint32_t N, const uint32_t *__restrict__ hoff,
                                                  const uint32_t *__restrict__ hj, uint32_t hjs,
                                                  const uint32_t *__restrict__ hm,
                                                  const float *__restrict__ E,
                                                  const float *__restrict__ Q,
                                                  const float *__restrict__ dA, uint32_t ldda,
                                                  float *__restrict__ dZ,
                                                  float *__restrict__ part) {
  __shared__ float sp[4][5][kHid];
  const uint32_t f = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  float gw[4] = {0.f, 0.f, 0.f, 0.f}, gb = 0.f;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < N; i += waves) {
    const float da = dA[(size_t)i * ldda + f];
    float dz = 0.f;
    for (uint32_t h = hoff[i]; h < hoff[i + 1]; ++h) {
      // own half: weight gradients
      const float dq = Q[(size_t)h * kHid + f] > 0.f ? da : 0.f;
      const float4 e = *reinterpret_cast<const float4 *>(E + 4 * (size_t)h);
      gw[0] += e.x * dq;
      gw[1] += e.y * dq;
      gw[2] += e.z * dq;
      gw[3] += e.w * dq;
      gb += dq;
      // the mirror half (j -> i) carries z_i into particle j's sum
      const uint32_t j = hj[(size_t)h * hjs], hmh = hm[h];
      const float dqm = Q[(size_t)hmh * kHid + f] > 0.f ? dA[(size_t)j * ldda + f] : 0.f;
      dz += dqm;
    }
    dZ[(size_t)i * kHid + f] = dz;
  }
  for (int t = 0; t < 4; ++t) sp[wave][t][f] = gw[t];
  sp[wave][4][f] = gb;
  __syncthreads();
  if (wave == 0) {
    for (int t = 0; t < 5; ++t) {
      const float s = ((sp[0][t][f] + sp[1][t][f]) + sp[2][t][f]) + sp[3][t][f];
      part[((size_t)blockIdx.x * 5 + t) * kHid + f] = s;
    }
  }
}

// mean squared error over the first ncols columns, gradient and FP64 partials
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ Y, uint32_t ldy,
                                             const float *__restrict__ T, uint32_t ldt,
                                             uint32_t N, uint32_t nc, float scale,
                                             float *__restrict__ dY, uint32_t lddy,
                                             double *__restrict__ part) {
  __shared__ double s[4];
  double acc = 0;
  const size_t n = (size_t)N * nc;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
       q += (size_t)blockDim.x * gridDim.x) {
    const size_t i = q / nc, c = q % nc;
    const float d = Y[i * ldy + c] - T[i * ldt + c];
    acc += (double)d * (double)d;
    if (dY) dY[i * lddy + c] = 2.f * scale * d;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

// dWe[4 x 64] and db[64] from the per-workgroup partials of k_edge_bwd
__global__ void k_reduce_edge(const float *__restrict__ part, uint32_t nblocks,
                              float *__restrict__ dWe, float *__restrict__ db, int accumulate) {
  for (uint32_t i = threadIdx.x; i < 5 * kHid; i += blockDim.x) {
    float s = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) s += part[(size_t)b * 5 * kHid + i];
    float *o = i < 4 * kHid ? dWe + i : db + (i - 4 * kHid);
    *o = accumulate ? *o + s : s;
  }
}

__global__ void k_sum_parts_f64(const double *__restrict__ part, uint32_t n, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (uint32_t b = 0; b < n; ++b) s += part[b];
    *out = s;
  }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint32_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += grid_stride()) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

uint32_t tn_blocks(uint32_t M) { return (M + kTnChunk - 1) / kTnChunk; }

}  // namespace

extern "C" {

int rfd_mm_bf16(rfd_ctx *ctx, const float *X, uint32_t ldx, const float *W, uint32_t wsk,
                uint32_t wsc, const float *bias, float *Y, uint32_t ldy, uint32_t M, uint32_t K,
                uint32_t Nc, int accumulate) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, X && W && Y, "rfd_mm_bf16: null pointer");
  RFD_CHECK_ARG(ctx, K > 0 && K % 32 == 0 && K <= (uint32_t)kMmMaxK,
                "rfd_mm_bf16: K = %u must be a multiple of 32, at most %d", K, kMmMaxK);
  RFD_CHECK_ARG(ctx, Nc > 0 && Nc % 16 == 0 && Nc <= (uint32_t)kMmMaxN,
                "rfd_mm_bf16: N = %u must be a multiple of 16, at most %d", Nc, kMmMaxN);
  RFD_CHECK_ARG(ctx, ldx >= K && ldy >= Nc, "rfd_mm_bf16: leading dimensions too small");
  if (M == 0) return RFD_OK;
  k_mm<<<(M + kMmRows - 1) / kMmRows, 256, 0, ctx->stream>>>(X, ldx, W, wsk, wsc, bias, Y, ldy,
                                                             M, K, Nc, accumulate);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_mm_tn_bf16(rfd_ctx *ctx, const float *X, uint32_t ldx, const float *D, uint32_t ldd,
                   float *G, uint32_t M, uint32_t K, uint32_t Nc, int accumulate) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, X && D && G, "rfd_mm_tn_bf16: null pointer");
  RFD_CHECK_ARG(ctx, K > 0 && K % 16 == 0 && K <= (uint32_t)kMmMaxK,
                "rfd_mm_tn_bf16: K = %u must be a multiple of 16, at most %d", K, kMmMaxK);
  RFD_CHECK_ARG(ctx, M > 0, "rfd_mm_tn_bf16: no rows");
  RFD_CHECK_ARG(ctx, Nc > 0 && Nc % 16 == 0 && Nc <= (uint32_t)kMmMaxN,
                "rfd_mm_tn_bf16: N = %u must be a multiple of 16, at most %d", Nc, kMmMaxN);
  RFD_CHECK_ARG(ctx, ldx >= K && ldd >= Nc, "rfd_mm_tn_bf16: leading dimensions too small");
  const uint32_t nb = tn_blocks(M);
  void *scr = nullptr;
  int st = rfd_scratch(ctx, sizeof(float) * (size_t)nb * K * Nc, &scr);
  if (st) return st;
  float *part = (float *)scr;
  k_mm_tn<<<nb, 256, 0, ctx->stream>>>(X, ldx, D, ldd, part, M, K, Nc);
  k_reduce_parts<<<rfd_blocks((size_t)K * Nc, 256), 256, 0, ctx->stream>>>(part, nb, K * Nc, G,
                                                                            accumulate);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_colsum(rfd_ctx *ctx, const float *D, uint32_t ldd, float *g, uint32_t M, uint32_t nc,
               int accumulate) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, D && g && nc > 0 && ldd >= nc && M > 0, "rfd_colsum: bad arguments");
  const uint32_t nb = tn_blocks(M);
  void *scr = nullptr;
  int st = rfd_scratch(ctx, sizeof(float) * (size_t)nb * nc, &scr);
  if (st) return st;
  k_colsum<<<nb, 128, 0, ctx->stream>>>(D, ldd, M, nc, (float *)scr);
  k_reduce_parts<<<rfd_blocks(nc, 256), 256, 0, ctx->stream>>>((float *)scr, nb, nc, g,
                                                                accumulate);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_mpnn_ewise(rfd_ctx *ctx, int op, uint32_t M, uint32_t nc, float alpha, const float *a,
                   uint32_t lda, const float *b, uint32_t ldb, float *out, uint32_t ldo) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, op >= 0 && op <= 7 && a && out, "rfd_mpnn_ewise: bad arguments");
  RFD_CHECK_ARG(ctx, op == 0 || op == 4 || op == 7 || b, "rfd_mpnn_ewise: op %d needs b", op);
  if (M == 0 || nc == 0) return RFD_OK;
  k_ewise<<<rfd_blocks((size_t)M * nc, 256), 256, 0, ctx->stream>>>(op, M, nc, alpha, a, lda, b,
                                                                     ldb, out, ldo);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_mpnn_edge_fwd(rfd_ctx *ctx, uint32_t N, const uint32_t *hoff, const uint32_t *hj,
                      uint32_t hj_stride, const float *E, const float *We, const float *bm,
                      const float *Z, float *Q, float *A, uint32_t lda) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, hoff && hj && E && We && bm && Z && A && hj_stride > 0 && lda >= kHid,
                "rfd_mpnn_edge_fwd: bad arguments");
  if (N == 0) return RFD_OK;
  k_edge_fwd<<<rfd_blocks((size_t)N * 64, 256, 16384), 256, 0, ctx->stream>>>(
      N, hoff, hj, hj_stride, E, We, bm, Z, Q, A, lda);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_mpnn_edge_bwd(rfd_ctx *ctx, uint32_t N, const uint32_t *hoff, const uint32_t *hj,
                      uint32_t hj_stride, const uint32_t *hm, const float *E, const float *Q,
                      const float *dA, uint32_t ldda, float *dZ, float *dWe, float *dbm,
                      int accumulate) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, hoff && hj && hm && E && Q && dA && dZ && dWe && dbm && hj_stride > 0 &&
                         ldda >= kHid,
                "rfd_mpnn_edge_bwd: bad arguments");
  if (N == 0) return RFD_OK;
  const uint32_t nb = rfd_blocks((size_t)N * 64, 256, 1024);
  void *scr = nullptr;
  int st = rfd_scratch(ctx, sizeof(float) * (size_t)nb * 5 * kHid + 256, &scr);
  if (st) return st;
  float *part = (float *)scr;
  k_edge_bwd<<<nb, 256, 0, ctx->stream>>>(N, hoff, hj, hj_stride, hm, E, Q, dA, ldda, dZ, part);
  // part[b][t][f]: t < 4 -> dWe[t][f], t = 4 -> db[f]
  k_reduce_edge<<<1, 320, 0, ctx->stream>>>(part, nb, dWe, dbm, accumulate);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

int rfd_mse_loss(rfd_ctx *ctx, const float *Y, uint32_t ldy, const float *T, uint32_t ldt,
                 uint32_t N, uint32_t nc, float *dY, uint32_t lddy, double *loss) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, Y && T && loss && N > 0 && nc > 0, "rfd_mse_loss: bad arguments");
  const uint32_t nb = rfd_blocks((size_t)N * nc, 256, kRedBlocks);
  const float scale = (float)(1.0 / ((double)N * nc));
  k_mse<<<nb, 256, 0, ctx->stream>>>(Y, ldy, T, ldt, N, nc, scale, dY, lddy, ctx->d_red);
  k_sum_parts_f64<<<1, 64, 0, ctx->stream>>>(ctx->d_red, nb, ctx->d_red + kRedBlocks);
  RFD_HIP(ctx, hipMemcpyAsync(ctx->h_red, ctx->d_red + kRedBlocks, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *loss = ctx->h_red[0] * (double)scale;
  return RFD_OK;
}

int rfd_adam(rfd_ctx *ctx, float *p, const float *g, float *m, float *v, uint32_t n, float lr,
             float b1, float b2, float eps, uint32_t t) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, p && g && m && v && n > 0 && t > 0, "rfd_adam: bad arguments");
  const float c1 = (float)(1.0 / (1.0 - pow((double)b1, (double)t)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)b2, (double)t)));
  k_adam<<<rfd_blocks(n, 256), 256, 0, ctx->stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  RFD_HIP(ctx, hipGetLastError());
  return RFD_OK;
}

}  // extern "C"
