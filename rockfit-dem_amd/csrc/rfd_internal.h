// rfd_internal.h -- shared internals of librfd (context, errors, device math).
//
// All device code is compiled for gfx950 with -ffp-contract=off so that each
// FP32 product and sum rounds separately and in the written order; that is
// what makes the reference-layout kernels reproducible bit for bit on any
// IEEE machine.  Division and sqrt are the correctly rounded IEEE operations
// (-fhip-fp32-correctly-rounded-divide-sqrt).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/rfd.h"

struct rfd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t *d_flags = nullptr;     // device condition flags (RFD_FLAG_*)
  double *d_red = nullptr;         // reduction partials (kRedSlots doubles)
  double *h_red = nullptr;         // pinned host copy of reduction results
  void *scratch = nullptr;         // grow-only workspace
  size_t scratch_bytes = 0;
  void *fixed_scratch = nullptr;   // device address of the fixed workspace
  char err[512] = {0};
};

constexpr int kRedBlocks = 1024;  // max partial-sum blocks of a reduction
constexpr int kRedSlots = kRedBlocks * 4 + 64;

int rfd_set_error(rfd_ctx *ctx, int code, const char *fmt, ...);
// hip error -> positive status with message
int rfd_hip_status(rfd_ctx *ctx, hipError_t e, const char *where);
// grow-only workspace on the context's device
int rfd_scratch(rfd_ctx *ctx, size_t bytes, void **out);
// look up the device address of the fixed (no-allocation) workspace on the
// context's device, the device being current
int rfd_fixed_scratch_attach(rfd_ctx *ctx);

#define RFD_CHECK_ARG(ctx, cond, ...)                  \
  do {                                                 \
    if (!(cond)) return rfd_set_error(ctx, RFD_E_ARG, __VA_ARGS__); \
  } while (0)

#define RFD_HIP(ctx, expr)                                      \
  do {                                                          \
    hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) return rfd_hip_status(ctx, e_, #expr); \
  } while (0)

#define RFD_LAUNCHED(ctx, name) \
  RFD_HIP(ctx, hipGetLastError())

// ------------------------------------------------------------ device math
namespace rfd {

constexpr float kPiF = 3.14159265358979f;      // MCf_pi, md_math_constants.h:87
constexpr float kTwoThirdsF = 0.6666666667f;   // MCf_2d3, md_math_constants.h:63
constexpr uint32_t kSentinel = 0xFFFFFFFFu;

// norm3df replacement: the sum of squares is formed in double (each square is
// exact) and rounded to float once after the square root.
__device__ __forceinline__ float norm3(float x, float y, float z) {
  double dx = x, dy = y, dz = z;
  return (float)__builtin_sqrt(dx * dx + dy * dy + dz * dz);
}

// __frsqrt_rn replacement: 1/sqrt in double, rounded once to float.
__device__ __forceinline__ float rsqrt_rn(float x) {
  return (float)(1.0 / __builtin_sqrt((double)x));
}

__device__ __forceinline__ uint32_t grid_stride() {
  return blockDim.x * gridDim.x;
}

// Workgroup b of a launch of kXcds * per workgroups takes block
// (b % kXcds) * per + b / kXcds. Consecutive workgroups go to the 8 XCDs in
// turn, so each XCD, with its own L2, walks one contiguous run of blocks and
// finds its neighbours' data in that L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t per) {
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// wave64 sum of a double across the wavefront (butterfly over all 64 lanes,
// every lane ends with the same total)
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace rfd

inline unsigned rfd_blocks(size_t n, unsigned tpb, unsigned cap = 65535u * 4) {
  size_t b = (n + tpb - 1) / tpb;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}
