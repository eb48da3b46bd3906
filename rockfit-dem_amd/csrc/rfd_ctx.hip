// rfd_ctx.hip -- context lifetime, errors, flags, workspace.
#include "rfd_internal.h"

int rfd_set_error(rfd_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

int rfd_hip_status(rfd_ctx *ctx, hipError_t e, const char *where) {
  rfd_set_error(ctx, (int)e, "%s: %s", where, hipGetErrorString(e));
  return (int)e;
}

int rfd_scratch(rfd_ctx *ctx, size_t bytes, void **out) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
      RFD_HIP(ctx, hipFree(ctx->scratch));
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    RFD_HIP(ctx, hipMalloc(&ctx->scratch, want));
    ctx->scratch_bytes = want;
  }
  *out = ctx->scratch;
  return RFD_OK;
}

extern "C" {

const char *rfd_version(void) {
  return "rfd 0.1 gfx950 (reference-layout tier + native edge-list tier)";
}

int rfd_ctx_create(rfd_ctx **out, int device, void *stream) {
  if (!out) return RFD_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RFD_E_NODEV;
  if (device < 0 || device >= n) return RFD_E_ARG;
  rfd_ctx *ctx = new rfd_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc(&ctx->d_flags, sizeof(uint32_t) * 4);
  if (e == hipSuccess) e = hipMalloc(&ctx->d_red, sizeof(double) * kRedSlots);
  if (e == hipSuccess) e = hipHostMalloc(&ctx->h_red, sizeof(double) * 64);
  if (e == hipSuccess)
    e = hipMemsetAsync(ctx->d_flags, 0, sizeof(uint32_t) * 4, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    rfd_ctx_destroy(ctx);
    return (int)e;
  }
  const int st = rfd_fixed_scratch_attach(ctx);
  if (st != RFD_OK) {
    rfd_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return RFD_OK;
}

int rfd_ctx_destroy(rfd_ctx *ctx) {
  if (!ctx) return RFD_OK;
  // teardown: errors here have nowhere to go, the context is released anyway
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  if (ctx->d_flags) (void)hipFree(ctx->d_flags);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  delete ctx;
  return RFD_OK;
}

const char *rfd_last_error(const rfd_ctx *ctx) { return ctx ? ctx->err : "null context"; }

int rfd_sync(rfd_ctx *ctx) {
  if (!ctx) return RFD_E_ARG;
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RFD_OK;
}

int rfd_read_flags(rfd_ctx *ctx, uint32_t *flags) {
  if (!ctx || !flags) return RFD_E_ARG;
  uint32_t h[4];
  RFD_HIP(ctx, hipMemcpyAsync(h, ctx->d_flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *flags = h[0];
  RFD_HIP(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(uint32_t) * 4, ctx->stream));
  return RFD_OK;
}

}  // extern "C"
