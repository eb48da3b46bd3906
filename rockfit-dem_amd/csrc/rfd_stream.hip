// rfd_stream.hip -- the HBM ceiling the roofline figures are read against:
// a plain streaming copy, 16 bytes per lane per access, with non-temporal
// loads and stores (the data is touched once), unrolled four deep so that
// each lane keeps four loads in flight.
#include "rfd_internal.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kStreamThreads = 256;
constexpr int kStreamUnroll = 4;

__global__ void __launch_bounds__(kStreamThreads)
k_stream_copy(const v4f *__restrict__ src, v4f *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * kStreamThreads;
  size_t i = (size_t)blockIdx.x * kStreamThreads + threadIdx.x;
  for (; i + (kStreamUnroll - 1) * stride < n; i += kStreamUnroll * stride) {
    v4f v[kStreamUnroll];
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kStreamUnroll; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace

extern "C" int rfd_stream_copy(rfd_ctx *ctx, void *dst, const void *src, size_t bytes,
                               uint32_t blocks) {
  if (!ctx) return RFD_E_ARG;
  RFD_CHECK_ARG(ctx, dst && src && bytes > 0 && bytes % 16 == 0 &&
                         (uintptr_t)dst % 16 == 0 && (uintptr_t)src % 16 == 0,
                "rfd_stream_copy: buffers must be 16-byte aligned, bytes a multiple of 16");
  const size_t n = bytes / 16;
  if (blocks == 0) blocks = 256 * 16;  // 16 workgroups per CU of the 256
  k_stream_copy<<<blocks, kStreamThreads, 0, ctx->stream>>>((const v4f *)src, (v4f *)dst, n);
  RFD_LAUNCHED(ctx, "rfd_stream_copy");
  return RFD_OK;
}
