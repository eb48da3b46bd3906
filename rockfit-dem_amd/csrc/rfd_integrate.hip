// rfd_integrate.hip -- integration, platens, restriction, FIRE (reference
// slot layout), with deterministic FP64 reductions.
//
// Reductions run in two fixed-order stages: each block reduces its particles
// (wave64 butterfly in double, then the block's waves in index order) into a
// partial, and one block sums the partials in index order.  The result does
// not depend on timing, and it is copied to pinned host memory.
#include "rfd_internal.h"

using namespace rfd;

namespace {

constexpr int kRedThreads = 256;

// block-level sum of NV doubles; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *s_part) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) s_part[wid * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = 0;
      for (int w = 0; w < nw; ++w) a += s_part[w * NV + q];
      v[q] = a;
    }
  }
}

// stage two: sum nblocks partial records of NV doubles (max taken for the
// slots flagged in maxmask) into out[0..NV)
template <int NV>
__global__ void k_finish(const double *__restrict__ part, int nblocks, unsigned maxmask,
                         double *__restrict__ out) {
  __shared__ double s_part[(kRedThreads / 64) * NV];
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  // each thread handles a contiguous chunk (fixed order), then fixed tree
  const int per = (nblocks + kRedThreads - 1) / kRedThreads;
  const int b0 = threadIdx.x * per;
  for (int b = b0; b < b0 + per && b < nblocks; ++b) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x = part[b * NV + q];
      v[q] = ((maxmask >> q) & 1u) ? fmax(v[q], x) : v[q] + x;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = ((maxmask >> q) & 1u) ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) s_part[wid * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = 0;
      for (int w = 0; w < kRedThreads / 64; ++w)
        a = ((maxmask >> q) & 1u) ? fmax(a, s_part[w * NV + q]) : a + s_part[w * NV + q];
      out[q] = a;
    }
  }
}

__global__ void k_zero_fm(float *__restrict__ F, float *__restrict__ M, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)blockDim.x * gridDim.x) {
    F[i] = 0.0f;
    M[i] = 0.0f;
  }
}

// cudaMemset(P.d_V, 0, ...) of the FIRE back-off (md_firerelaxation.cu:292)
__global__ void k_zero_f(float *__restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)blockDim.x * gridDim.x)
    a[i] = 0.0f;
}

// d_CalculateIncrementsDEM (md_calculateincrementsdem.cu:19-49)
__global__ void k_increments(const float *__restrict__ F, float *__restrict__ V,
                             float *__restrict__ R, const float *__restrict__ M,
                             float *__restrict__ W, uint32_t N, float dt_m, float dt,
                             float dt_I) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t a = i + (size_t)d * N;
      const float v = V[a] + F[a] * dt_m;
      V[a] = v;
      R[a] = R[a] + v * dt;
      W[a] = W[a] + M[a] * dt_I;
    }
  }
}

// d_UniaxialCompression3_simple (md_uniaxialcompression3_simple.cu:19-76)
__global__ void __launch_bounds__(kRedThreads)
k_platen_uniaxial(float *__restrict__ R, float *__restrict__ V, float *__restrict__ F,
                  uint32_t N, float cx, float cy, float RR, float Zt, float Zb, float Ztr,
                  float Zbr, float vt, float Zcut, double *__restrict__ part) {
  __shared__ double s_part[(kRedThreads / 64) * 2];
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double v[2] = {0.0, 0.0};  // bottom, top
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const float rx = R[i] - cx, ry = R[i + N1] - cy, rz = R[i + N2];
    if (rx * rx + ry * ry < RR && rz < Zcut) {
      if ((double)(rz - Ztr) > -1e-9) {
        const float fz = F[i + N2];
        if (fz > 0) v[1] += fz;
        if ((double)(rz - Zt) > -1e-9) {
          R[i + N2] = Zt;
          V[i + N2] = vt;
          F[i + N2] = 0.0f;
        }
      }
      if ((double)(rz - Zbr) < 1e-9) {
        const float fz = F[i + N2];
        if (fz < 0) v[0] += fz;
        if ((double)(rz - Zb) < 1e-9) {
          R[i + N2] = Zb;
          V[i + N2] = 0.0f;
          F[i + N2] = 0.0f;
        }
      }
    }
  }
  block_sum<2>(v, s_part);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = v[0];
    part[blockIdx.x * 2 + 1] = v[1];
  }
}

// d_UniaxialCompression_simple (md_uniaxialcompression_simple.cu:29-81), the
// superseded penalty platens: a particle past z = c.z +- Hd2 is pushed back
// by the stiffness C, with Coulomb friction mu against its in-plane velocity
// (dd_FrictionForce :19-28); the reactions are summed in double
__device__ __forceinline__ void penalty_friction(const float *__restrict__ V,
                                                 float *__restrict__ F, size_t N, uint32_t i,
                                                 float fz, float mu) {
  const float vv = V[i] * V[i] + V[i + N] * V[i + N];
  if ((double)vv > 1e-10) {
    const float iv = rsqrt_rn(vv);
    F[i] -= fabsf(fz) * mu * V[i] * iv;
    F[i + N] -= fabsf(fz) * mu * V[i + N] * iv;
  }
}

__global__ void __launch_bounds__(kRedThreads)
k_platen_uniaxial_penalty(const float *__restrict__ R, const float *__restrict__ V,
                          float *__restrict__ F, uint32_t N, float cx, float cy, float cz,
                          float RR, float Hd2, float C, float mu, float Zcut,
                          double *__restrict__ part) {
  __shared__ double s_part[(kRedThreads / 64) * 2];
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double v[2] = {0.0, 0.0};  // bottom, top
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const float rx = R[i] - cx, ry = R[i + N1] - cy, rz = R[i + N2] - cz;
    if (rx * rx + ry * ry < RR && R[i + N2] < Zcut) {
      if (rz - Hd2 > 0) {
        const float fz = (Hd2 - rz) * C;
        penalty_friction(V, F, N, i, fz, mu);
        F[i + N2] += fz;
        v[1] -= fz;
      }
      if (rz + Hd2 < 0) {
        const float fz = (Hd2 + rz) * C;
        penalty_friction(V, F, N, i, fz, mu);
        F[i + N2] -= fz;
        v[0] += fz;
      }
    }
  }
  block_sum<2>(v, s_part);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = v[0];
    part[blockIdx.x * 2 + 1] = v[1];
  }
}

// d_braziltest_simple2 (md_braziltest_simple2.cu:19-92)
__global__ void __launch_bounds__(kRedThreads)
k_platen_brazil(float *__restrict__ R, float *__restrict__ V, float *__restrict__ F,
                uint32_t N, float cx, float cz, float RR, float Yt, float Yb, float Ytr,
                float Ybr, float vt, float Zcut, double *__restrict__ part) {
  __shared__ double s_part[(kRedThreads / 64) * 2];
  const size_t N1 = N, N2 = 2 * (size_t)N;
  double v[2] = {0.0, 0.0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const float rx = R[i] - cx, ry = R[i + N1], rz = R[i + N2] - cz;
    if (rx * rx + rz * rz < RR && rz < Zcut) {
      if ((double)(ry - Ytr) > -1e-9) {
        const float fy = F[i + N1];
        if (fy > 0) {
          v[1] += fy;
          float vv = V[i] * V[i] + V[i + N2] * V[i + N2];
          if ((double)vv > 1e-9) {
            vv = rsqrt_rn(vv);
            F[i] = F[i] - 0.45f * fy * V[i] * vv;
            F[i + N2] = F[i + N2] - 0.45f * fy * V[i + N2] * vv;
          }
        }
        if ((double)(ry - Yt) > -1e-9) {
          R[i + N1] = Yt;
          V[i + N1] = vt;
          F[i + N1] = 0.0f;
        }
      }
      if ((double)(ry - Ybr) < 1e-9) {
        const float fy = F[i + N1];
        if (fy < 0) {
          v[0] += fy;
          float vv = V[i] * V[i] + V[i + N2] * V[i + N2];
          if ((double)vv > 1e-9) {
            vv = rsqrt_rn(vv);
            F[i] = F[i] + 0.45f * fy * V[i] * vv;
            F[i + N2] = F[i + N2] + 0.45f * fy * V[i + N2] * vv;
          }
        }
        if ((double)(ry - Yb) < 1e-9) {
          R[i + N1] = Yb;
          V[i + N1] = 0.0f;
          F[i + N1] = 0.0f;
        }
      }
    }
  }
  block_sum<2>(v, s_part);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2] = v[0];
    part[blockIdx.x * 2 + 1] = v[1];
  }
}

// d_ParallelepipedCutRestriction (md_parallelepipedrestriction.cu:73-158)
__global__ void k_cut_restriction(float *__restrict__ R, float *__restrict__ V,
                                  float *__restrict__ F, uint32_t N, float cx, float cy,
                                  float cz, float hx, float hy, float hz, float px, float py,
                                  float pz) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const size_t N1 = N, N2 = 2 * (size_t)N;
    const float dx = R[i] - cx, dy = R[i + N1] - cy, dz = R[i + N2] - cz;
    const bool out = (dx - hx > 0) || (dx + hx < 0) || (dy - hy > 0) || (dy + hy < 0) ||
                     (dz - hz > 0) || (dz + hz < 0);
    if (out) {
      R[i] = px;
      R[i + N1] = py;
      R[i + N2] = pz;
      V[i] = 0.0f;
      V[i + N1] = 0.0f;
      V[i + N2] = 0.0f;
      F[i] = 0.0f;
      F[i + N1] = 0.0f;
      F[i + N2] = 0.0f;
    }
  }
}

// d_CalculateIncrementsFIRE (md_firerelaxation.cu:18-56)
__global__ void k_fire_increments(const float *__restrict__ F, float *__restrict__ V,
                                  float *__restrict__ R, uint32_t N, float dt_m, float dt,
                                  float alpha) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    const size_t N1 = N, N2 = 2 * (size_t)N;
    const float Fx = F[i], Fy = F[i + N1], Fz = F[i + N2];
    float vx = Fx * dt_m, vy = Fy * dt_m, vz = Fz * dt_m;
    const float vv = vx * vx + vy * vy + vz * vz;
    const float f = Fx * Fx + Fy * Fy + Fz * Fz;
    float vdf;
    if (f > 1e-12f)
      vdf = alpha * sqrtf(vv * (1.0f / f));
    else
      vdf = (float)((double)(alpha * sqrtf(vv)) * 1e6);
    vx = (1.0f - alpha) * vx + vdf * Fx;
    vy = (1.0f - alpha) * vy + vdf * Fy;
    vz = (1.0f - alpha) * vz + vdf * Fz;
    V[i] = vx;
    V[i + N1] = vy;
    V[i + N2] = vz;
    R[i] = R[i] + vx * dt;
    R[i + N1] = R[i + N1] + vy * dt;
    R[i + N2] = R[i + N2] + vz * dt;
  }
}

// d_CalculateDecrementsHalfStepFIRE (md_firerelaxation.cu:58-70)
__global__ void k_fire_halfstep(const float *__restrict__ V, float *__restrict__ R, size_t n,
                                float dt_d2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)blockDim.x * gridDim.x)
    R[i] = R[i] - V[i] * dt_d2;
}

// d_FdotVEntire (md_firerelaxation.cu:372-490), extended to |F|^2, |V|^2,
// max |F_i|^2 and the rotational terms; all in double
__global__ void __launch_bounds__(kRedThreads)
k_fire_power(const float *__restrict__ F, const float *__restrict__ V,
             const float *__restrict__ M, const float *__restrict__ W, uint32_t N,
             double *__restrict__ part) {
  __shared__ double s_part[(kRedThreads / 64) * 4];
  double v[4] = {0, 0, 0, 0};
  double fmx = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += grid_stride()) {
    double f2 = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double f = F[i + (size_t)d * N], u = V[i + (size_t)d * N];
      v[0] += f * u;
      f2 += f * f;
      v[2] += u * u;
    }
    v[1] += f2;
    fmx = fmax(fmx, f2);
    if (M && W) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double m = M[i + (size_t)d * N], w = W[i + (size_t)d * N];
        v[0] += m * w;
        v[1] += m * m;
        v[2] += w * w;
      }
    }
  }
  double s3[3] = {v[0], v[1], v[2]};
  block_sum<3>(s3, s_part);
  __syncthreads();
  fmx = wave_max(fmx);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = fmx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = 0;
    for (int w = 0; w < kRedThreads / 64; ++w) mx = fmax(mx, s_part[w]);
    part[blockIdx.x * 4] = s3[0];
    part[blockIdx.x * 4 + 1] = s3[1];
    part[blockIdx.x * 4 + 2] = s3[2];
    part[blockIdx.x * 4 + 3] = mx;
  }
}

unsigned red_blocks(uint32_t N) { return rfd_blocks(N, kRedThreads, kRedBlocks); }

// copy NV result doubles to the host and wait
int fetch(rfd_ctx *ctx, int nv, double *out) {
  const double *src = ctx->d_red + (size_t)kRedBlocks * 4;
  RFD_HIP(ctx, hipMemcpyAsync(ctx->h_red, src, sizeof(double) * nv, hipMemcpyDeviceToHost,
                              ctx->stream));
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < nv; ++q) out[q] = ctx->h_red[q];
  return RFD_OK;
}

}  // namespace

extern "C" {

int rfd_zero_fm(rfd_ctx *ctx, float *F, float *M, uint32_t N) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, F && M && N > 0, "rfd_zero_fm: bad arguments");
  k_zero_fm<<<rfd_blocks(3 * (size_t)N, 256, 8192), 256, 0, ctx->stream>>>(F, M, 3 * (size_t)N);
  RFD_LAUNCHED(ctx, "rfd_zero_fm");
  return RFD_OK;
}

int rfd_increments(rfd_ctx *ctx, const float *F, float *V, float *R, const float *M, float *W,
                   uint32_t N, float dt_m, float dt, float dt_I) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, F && V && R && M && W && N > 0, "rfd_increments: bad arguments");
  k_increments<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(F, V, R, M, W, N, dt_m, dt, dt_I);
  RFD_LAUNCHED(ctx, "rfd_increments");
  return RFD_OK;
}

int rfd_platen_uniaxial(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N, float cx,
                        float cy, float RR, float Zt, float Zb, float Ztr, float Zbr, float vt,
                        float Zcut, double out_sums[2]) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && F && N > 0 && out_sums, "rfd_platen_uniaxial: bad arguments");
  const unsigned nb = red_blocks(N);
  k_platen_uniaxial<<<nb, kRedThreads, 0, ctx->stream>>>(R, V, F, N, cx, cy, RR, Zt, Zb, Ztr,
                                                          Zbr, vt, Zcut, ctx->d_red);
  k_finish<2><<<1, kRedThreads, 0, ctx->stream>>>(ctx->d_red, nb, 0u,
                                                   ctx->d_red + (size_t)kRedBlocks * 4);
  RFD_LAUNCHED(ctx, "rfd_platen_uniaxial");
  return fetch(ctx, 2, out_sums);
}

int rfd_platen_uniaxial_penalty(rfd_ctx *ctx, const float *R, const float *V, float *F,
                                uint32_t N, float cx, float cy, float cz, float RR, float Hd2,
                                float C, float mu, float Zcut, double out_sums[2]) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && F && N > 0 && out_sums,
                "rfd_platen_uniaxial_penalty: bad arguments");
  const unsigned nb = red_blocks(N);
  k_platen_uniaxial_penalty<<<nb, kRedThreads, 0, ctx->stream>>>(R, V, F, N, cx, cy, cz, RR, Hd2,
                                                                  C, mu, Zcut, ctx->d_red);
  k_finish<2><<<1, kRedThreads, 0, ctx->stream>>>(ctx->d_red, nb, 0u,
                                                   ctx->d_red + (size_t)kRedBlocks * 4);
  RFD_LAUNCHED(ctx, "rfd_platen_uniaxial_penalty");
  return fetch(ctx, 2, out_sums);
}

int rfd_platen_brazil(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N, float cx,
                      float cz, float RR, float Yt, float Yb, float Ytr, float Ybr, float vt,
                      float Zcut, double out_sums[2]) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && F && N > 0 && out_sums, "rfd_platen_brazil: bad arguments");
  const unsigned nb = red_blocks(N);
  k_platen_brazil<<<nb, kRedThreads, 0, ctx->stream>>>(R, V, F, N, cx, cz, RR, Yt, Yb, Ytr, Ybr,
                                                        vt, Zcut, ctx->d_red);
  k_finish<2><<<1, kRedThreads, 0, ctx->stream>>>(ctx->d_red, nb, 0u,
                                                   ctx->d_red + (size_t)kRedBlocks * 4);
  RFD_LAUNCHED(ctx, "rfd_platen_brazil");
  return fetch(ctx, 2, out_sums);
}

int rfd_cut_restriction(rfd_ctx *ctx, float *R, float *V, float *F, uint32_t N,
                        const float center[3], const float sized2[3],
                        const float hidden_point[3]) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, R && V && F && N > 0 && center && sized2 && hidden_point,
                "rfd_cut_restriction: bad arguments");
  k_cut_restriction<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(
      R, V, F, N, center[0], center[1], center[2], sized2[0], sized2[1], sized2[2],
      hidden_point[0], hidden_point[1], hidden_point[2]);
  RFD_LAUNCHED(ctx, "rfd_cut_restriction");
  return RFD_OK;
}

int rfd_fire_increments(rfd_ctx *ctx, const float *F, float *V, float *R, uint32_t N, float dt_m,
                        float dt, float alpha) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, F && V && R && N > 0, "rfd_fire_increments: bad arguments");
  k_fire_increments<<<rfd_blocks(N, 256), 256, 0, ctx->stream>>>(F, V, R, N, dt_m, dt, alpha);
  RFD_LAUNCHED(ctx, "rfd_fire_increments");
  return RFD_OK;
}

int rfd_fire_halfstep(rfd_ctx *ctx, const float *V, float *R, uint32_t N, float dt_d2) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, V && R && N > 0, "rfd_fire_halfstep: bad arguments");
  k_fire_halfstep<<<rfd_blocks(3 * (size_t)N, 256), 256, 0, ctx->stream>>>(V, R, 3 * (size_t)N,
                                                                           dt_d2);
  RFD_LAUNCHED(ctx, "rfd_fire_halfstep");
  return RFD_OK;
}

void rfd_fire_ref_init(rfd_fire_ref *f, float dt0, float alpha0) {
  if (!f) return;
  *f = rfd_fire_ref{};
  // SetFIREData (md_firerelaxation.cu:72-84), with the float fields of
  // firerelax_data (md_data_types.h:296-301)
  f->dt0 = dt0;
  f->dtmax = (float)(100.0 * (double)dt0);
  f->dtmin = (float)(0.001 * (double)dt0);
  f->dtgrow = 1.1f;
  f->dtshrink = 0.5f;
  f->alpha0 = alpha0;
  f->alphashrink = 0.99f;
  f->n_negative_max = 30000;
  f->n_delay = 20;
  f->dt = dt0;
  f->alpha = alpha0;
}

int rfd_fire_relax_reference(rfd_ctx *ctx, const float *F, float *V, float *R, uint32_t N,
                             float dt_m, rfd_fire_ref *f, uint32_t iterations,
                             rfd_callback forces, rfd_callback restriction, void *user) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, F && V && R && N > 0 && f, "rfd_fire_relax_reference: bad arguments");
  RFD_CHECK_ARG(ctx, f->dt > 0.f && f->dtgrow > 0.f && f->dtshrink > 0.f,
                "rfd_fire_relax_reference: state not initialised (rfd_fire_ref_init)");
  for (uint32_t it = 0; it < iterations && !f->stopped; ++it) {
    if (forces) {
      const int st = forces(user);
      if (st) return rfd_set_error(ctx, st, "rfd_fire_relax_reference: force callback returned %d",
                                   st);
    }
    // d_FdotVEntire + the host sum (md_firerelaxation.cu:265-272), in FP64
    double pw[4];
    int st = rfd_fire_power(ctx, F, V, nullptr, nullptr, N, pw);
    if (st) return st;
    f->power = pw[0];
    // md_firerelaxation.cu:274-299
    if (pw[0] > 0.0) {
      ++f->n_positive;
      f->n_negative = 0;
      if (f->n_positive > f->n_delay) {
        f->dt = (f->dt * f->dtgrow < f->dtmax) ? f->dt * f->dtgrow : f->dtmax;
        f->alpha *= f->alphashrink;
      }
    } else {
      f->n_positive = 0;
      ++f->n_negative;
      if (f->n_negative > f->n_negative_max) {
        f->stopped = 1;
        break;
      }
      if (f->steps > f->n_delay) {
        f->dt = (f->dt * f->dtshrink > f->dtmin) ? f->dt * f->dtshrink : f->dtmin;
        f->alpha = f->alpha0;
      }
      st = rfd_fire_halfstep(ctx, V, R, N, (float)(0.5 * (double)f->dt));
      if (st) return st;
      k_zero_f<<<rfd_blocks(3 * (size_t)N, 256), 256, 0, ctx->stream>>>(V, 3 * (size_t)N);
      RFD_LAUNCHED(ctx, "rfd_fire_relax_reference");
    }
    st = rfd_fire_increments(ctx, F, V, R, N, dt_m, f->dt, f->alpha);
    if (st) return st;
    if (restriction) {
      st = restriction(user);
      if (st)
        return rfd_set_error(ctx, st, "rfd_fire_relax_reference: restriction callback returned %d",
                             st);
    }
    ++f->steps;
  }
  return RFD_OK;
}

int rfd_fire_power(rfd_ctx *ctx, const float *F, const float *V, const float *M, const float *W,
                   uint32_t N, double out[4]) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, F && V && N > 0 && out, "rfd_fire_power: bad arguments");
  RFD_CHECK_ARG(ctx, (M == nullptr) == (W == nullptr), "rfd_fire_power: pass both M and W or neither");
  const unsigned nb = red_blocks(N);
  k_fire_power<<<nb, kRedThreads, 0, ctx->stream>>>(F, V, M, W, N, ctx->d_red);
  k_finish<4><<<1, kRedThreads, 0, ctx->stream>>>(ctx->d_red, nb, 0x8u,
                                                   ctx->d_red + (size_t)kRedBlocks * 4);
  RFD_LAUNCHED(ctx, "rfd_fire_power");
  return fetch(ctx, 4, out);
}

}  // extern "C"
