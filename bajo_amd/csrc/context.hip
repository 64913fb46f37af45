// Context, error reporting and the scratch arena of the bajo HIP library.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "bajo_common.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace bj {

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int32_t hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return e == hipErrorOutOfMemory ? BAJO_ERR_OUT_OF_MEMORY : BAJO_ERR_HIP;
}

// Grows the per-context scratch arena; previous contents are not kept.
int32_t ctx_scratch(bajo_ctx* ctx, size_t bytes, void** out) {
  if (bytes <= ctx->scratch_bytes && ctx->scratch) {
    *out = ctx->scratch;
    return BAJO_OK;
  }
  if (ctx->scratch) {
    BJ_HIP(hipStreamSynchronize(ctx->stream));
    BJ_HIP(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 4096;
  BJ_HIP(hipMalloc(&ctx->scratch, want));
  ctx->scratch_bytes = want;
  *out = ctx->scratch;
  return BAJO_OK;
}

}  // namespace bj

extern "C" {

const char* bajo_last_error(void) { return g_last_error; }

const char* bajo_status_string(int32_t status) {
  switch (status) {
    case BAJO_OK: return "ok";
    case BAJO_ERR_INVALID_ARGUMENT: return "invalid argument";
    case BAJO_ERR_HIP: return "HIP runtime error";
    case BAJO_ERR_OUT_OF_MEMORY: return "out of device memory";
    case BAJO_ERR_UNSUPPORTED: return "unsupported";
    case BAJO_STATUS_STACK_OVERFLOW: return "traversal stack overflow";
    case BAJO_STATUS_BAD_NODE: return "traversal met a node outside the hierarchy";
    case BAJO_STATUS_SORT_LOOKBACK: return "sort look-back gave up waiting (corrupted look-back table)";
    case BAJO_STATUS_REFIT_CORRUPT: return "refit met an inconsistent hierarchy";
    case BAJO_STATUS_HPLOC_FAILED: return "H-PLOC clustering failed (no progress, corrupt cluster list or no single root)";
    case BAJO_STATUS_COLLAPSE_FAILED: return "wide collapse failed (out of nodes, bad root or too deep)";
    case BAJO_STATUS_PATH_OVERFLOW: return "path queue overflow";
    case BAJO_STATUS_SHADOW_OVERFLOW: return "shadow queue overflow";
    default: return "unknown status";
  }
}

int32_t bajo_ctx_create(int32_t device, void* stream, bajo_ctx** out) {
  BJ_REQUIRE(out != nullptr, "bajo_ctx_create: out is NULL");
  *out = nullptr;
  int count = 0;
  BJ_HIP(hipGetDeviceCount(&count));
  BJ_REQUIRE(device >= 0 && device < count, "bajo_ctx_create: device %d outside [0, %d)", device, count);
  BJ_HIP(hipSetDevice(device));
  bajo_ctx* ctx = new bajo_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  ctx->owns_stream = 0;
  ctx->status = nullptr;
  ctx->spin_limit = 0;
  hipError_t e = hipMalloc(&ctx->status, 4 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(ctx->status, 0, 4 * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipFree(ctx->status);
    delete ctx;
    return bj::hip_fail(e, "hipMalloc(context status)", __FILE__, __LINE__);
  }
  *out = ctx;
  return BAJO_OK;
}

int32_t bajo_ctx_destroy(bajo_ctx* ctx) {
  if (!ctx) return BAJO_OK;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  (void)hipFree(ctx->status);
  delete ctx;
  return BAJO_OK;
}

int32_t bajo_ctx_status(bajo_ctx* ctx, uint32_t* status) {
  BJ_REQUIRE(ctx != nullptr && status != nullptr, "bajo_ctx_status: NULL argument");
  uint32_t word = 0;
  BJ_HIP(hipMemcpyAsync(&word, ctx->status, sizeof(word), hipMemcpyDeviceToHost, ctx->stream));
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  if (word) {
    BJ_HIP(hipMemsetAsync(ctx->status, 0, sizeof(word), ctx->stream));
    BJ_HIP(hipStreamSynchronize(ctx->stream));
  }
  *status = word;
  return BAJO_OK;
}

int32_t bajo_ctx_synchronize(bajo_ctx* ctx) {
  BJ_REQUIRE(ctx != nullptr, "bajo_ctx_synchronize: ctx is NULL");
  uint32_t word = 0;
  const int32_t rc = bajo_ctx_status(ctx, &word);
  if (rc) return rc;
  if (word) {
    bj::set_error("device status %u: %s", word, bajo_status_string((int32_t)word));
    return (int32_t)word;
  }
  return BAJO_OK;
}

int32_t bajo_ctx_set_spin_limit(bajo_ctx* ctx, uint32_t spin_limit) {
  BJ_REQUIRE(ctx != nullptr, "bajo_ctx_set_spin_limit: ctx is NULL");
  ctx->spin_limit = spin_limit;
  return BAJO_OK;
}

int32_t bajo_device_name(int32_t device, char* out, int32_t cap) {
  BJ_REQUIRE(out != nullptr && cap > 0, "bajo_device_name: bad output buffer");
  hipDeviceProp_t p;
  BJ_HIP(hipGetDeviceProperties(&p, device));
  snprintf(out, (size_t)cap, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return BAJO_OK;
}

}  // extern "C"
