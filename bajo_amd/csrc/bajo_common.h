// Shared device/host helpers for the bajo MI355X (gfx950) kernels.
//
// The FP32 helpers reproduce bajo's scalar math exactly: select-based
// min/max (core/utils.mojo:60-72), FMA placement of dot/cross
// (core/vec.mojo:435-440, 517-526), the safe reciprocal epsilons
// (core/vec.mojo:373-395, core/ray.mojo:86-112) and the Morton quantization
// (core/morton.mojo:48-89).  Everything here is compiled with
// -ffp-contract=off, so no fused operation appears unless written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bajo_hip.h"

#define BJ_HD __host__ __device__ __forceinline__
#define BJ_D __device__ __forceinline__

namespace bj {

constexpr float kFltMax = 3.40282347e+38f;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFFu;
constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7FFFFFFFu;

struct v3 {
  float x, y, z;
};

BJ_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
BJ_HD float fmax_sel(float a, float b) { return a > b ? a : b; }
BJ_HD float fmin_sel(float a, float b) { return a < b ? a : b; }
BJ_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
BJ_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
BJ_HD v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
BJ_HD v3 scale(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
BJ_HD v3 vmin(v3 a, v3 b) { return mk(fmin_sel(a.x, b.x), fmin_sel(a.y, b.y), fmin_sel(a.z, b.z)); }
BJ_HD v3 vmax(v3 a, v3 b) { return mk(fmax_sel(a.x, b.x), fmax_sel(a.y, b.y), fmax_sel(a.z, b.z)); }
BJ_HD float dot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
BJ_HD v3 cross(v3 a, v3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
BJ_HD v3 normalize(v3 v) {
  float l = sqrtf(dot(v, v));
  float inv = l > 1.0e-20f ? 1.0f / l : 0.0f;
  return scale(v, inv);
}
BJ_HD v3 safe_inv(v3 v) {
  const float e = 1.0e-20f;
  return mk(fabsf(v.x) > e ? 1.0f / v.x : 0.0f, fabsf(v.y) > e ? 1.0f / v.y : 0.0f,
            fabsf(v.z) > e ? 1.0f / v.z : 0.0f);
}
BJ_HD v3 rcp_dir(v3 d) {
  const float e = 1.0e-9f, large = 1.0f / 1.0e-9f;
  return mk(fabsf(d.x) > e ? 1.0f / d.x : (d.x < 0.0f ? -large : large),
            fabsf(d.y) > e ? 1.0f / d.y : (d.y < 0.0f ? -large : large),
            fabsf(d.z) > e ? 1.0f / d.z : (d.z < 0.0f ? -large : large));
}

BJ_HD uint32_t expand_bits_3d(uint32_t n) {
  n = (n ^ (n << 16)) & 0xFF0000FFu;
  n = (n ^ (n << 8)) & 0x0300F00Fu;
  n = (n ^ (n << 4)) & 0x030C30C3u;
  n = (n ^ (n << 2)) & 0x09249249u;
  return n;
}
BJ_HD uint32_t quantize_axis(float v) {
  float s = v * 1024.0f;
  uint32_t u = s > 0.0f ? (s >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)s) : 0u;
  return u > 1023u ? 1023u : u;
}
BJ_HD uint32_t morton3(float x, float y, float z) {
  return (expand_bits_3d(quantize_axis(z)) << 2) | (expand_bits_3d(quantize_axis(y)) << 1) |
         expand_bits_3d(quantize_axis(x));
}

// Monotonic float <-> u32 mapping for exact min/max through integer atomics.
BJ_HD uint32_t f2ord(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
BJ_HD float ord2f(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  return __builtin_bit_cast(float, b);
}

BJ_D uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
// Number of set bits of `mask` in lanes strictly below this lane.
BJ_D uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
BJ_D uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

}  // namespace bj

// Error plumbing shared by every translation unit of the library.
namespace bj {
void set_error(const char* fmt, ...);
int32_t hip_fail(hipError_t e, const char* what, const char* file, int line);
}  // namespace bj

#define BJ_HIP(call)                                                 \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) return bj::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define BJ_REQUIRE(cond, ...)                \
  do {                                       \
    if (!(cond)) {                           \
      bj::set_error(__VA_ARGS__);            \
      return BAJO_ERR_INVALID_ARGUMENT;      \
    }                                        \
  } while (0)

// Library context: one device, one stream, plus a growable scratch arena
// that callers without their own workspace borrow.
// `status` is a device word kernels raise BAJO_STATUS_* codes into
// (atomicMax); it is read and cleared at the context's sync points.
struct bajo_ctx {
  int device;
  hipStream_t stream;
  void* scratch;
  size_t scratch_bytes;
  int owns_stream;
  uint32_t* status;
  uint32_t spin_limit;
};

namespace bj {
int32_t ctx_scratch(bajo_ctx* ctx, size_t bytes, void** out);
inline unsigned div_up(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }
}  // namespace bj
