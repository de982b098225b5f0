// Context management of libpycd_hip.so.
#include "common.hpp"

extern "C" int pycd_abi_version(void) { return 2; }

extern "C" int pycd_ctx_create(int device, pycd_ctx** out) {
  if (!out) return PYCD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PYCD_ERR_HIP;
  if (device < 0 || device >= n) return PYCD_ERR_ARG;
  pycd_ctx* ctx = new (std::nothrow) pycd_ctx();
  if (!ctx) return PYCD_ERR_NOMEM;
  ctx->device = device;
  pycd::DeviceGuard g(device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete ctx;
    return PYCD_ERR_HIP;
  }
  ctx->cu_count = prop.multiProcessorCount;
  if (pycd::kmc_static_scratch(device, &ctx->static_scratch) != PYCD_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PYCD_ERR_HIP;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) ==
          hipSuccess &&
      khz > 0)
    ctx->wall_clock_hz = 1e3 * khz;
  *out = ctx;
  return PYCD_OK;
}

extern "C" void pycd_ctx_destroy(pycd_ctx* ctx) {
  if (!ctx) return;
  {
    pycd::DeviceGuard g(ctx->device);
    if (ctx->stream) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
    }
  }
  delete ctx;
}

extern "C" int pycd_ctx_cu_count(const pycd_ctx* ctx) {
  return ctx ? ctx->cu_count : 0;
}

extern "C" const char* pycd_last_error(const pycd_ctx* ctx) {
  return ctx ? ctx->error.c_str() : "null context";
}

extern "C" int pycd_dev_alloc(pycd_ctx* ctx, int64_t bytes, void** out) {
  if (!ctx || !out || bytes <= 0) return PYCD_ERR_ARG;
  pycd::DeviceGuard g(ctx->device);
  *out = nullptr;
  PYCD_HIP(ctx, hipMalloc(out, size_t(bytes)));
  return PYCD_OK;
}

extern "C" int pycd_dev_free(pycd_ctx* ctx, void* p) {
  if (!ctx) return PYCD_ERR_ARG;
  pycd::DeviceGuard g(ctx->device);
  if (p) PYCD_HIP(ctx, hipFree(p));
  return PYCD_OK;
}

extern "C" int pycd_memcpy(pycd_ctx* ctx, void* dst, const void* src,
                           int64_t bytes) {
  if (!ctx || !dst || !src || bytes < 0) return PYCD_ERR_ARG;
  pycd::DeviceGuard g(ctx->device);
  PYCD_HIP(ctx, hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDefault,
                               ctx->stream));
  PYCD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PYCD_OK;
}

extern "C" int pycd_set_table_variant(pycd_ctx* ctx, int precision,
                                      int variant) {
  if (!ctx) return PYCD_ERR_ARG;
  if (precision < PYCD_PREC_F64 || precision > PYCD_PREC_BF16X3)
    return pycd::fail(ctx, PYCD_ERR_ARG, "unknown precision %d", precision);
  // FP64: 1 = 64 x 64 tiles, 2 = 128 x 128; BF16x3: 1 = K slices of 32,
  // 2 = slices of 64, 3 = slices of 32 in two LDS stages, 4 = 256 x 256
  // tiles in two LDS stages; FP32: 1 only; 0 = default
  const int hi = precision == PYCD_PREC_F32 ? 1
                 : precision == PYCD_PREC_BF16X3 ? 4 : 2;
  if (variant < 0 || variant > hi)
    return pycd::fail(ctx, PYCD_ERR_ARG, "variant %d outside 0..%d", variant,
                      hi);
  ctx->table_variant[precision] = variant;
  return PYCD_OK;
}

extern "C" int pycd_table_variant(const pycd_ctx* ctx, int precision) {
  if (!ctx || precision < PYCD_PREC_F64 || precision > PYCD_PREC_BF16X3)
    return -1;
  return ctx->table_variant[precision];
}
