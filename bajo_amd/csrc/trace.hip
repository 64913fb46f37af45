// Closest-hit / any-hit ray launches over the binary or 8-wide BVH (gfx950).
//
// Replaces GpuTriangleBvh.launch_rays / launch_camera /
// launch_camera_instrumented (bvh/gpu/triangle_bvh.mojo:520-614,730-863) and
// the near-first loop of bvh/gpu/trace.mojo:341-552 with its leaf test
// (bvh/gpu/triangle_bvh.mojo:866-959, determinant-scaled Moller-Trumbore in
// FP32, normal = normalize(e1 x e2)).
//
// MI355X design:
//   - persistent waves: the grid is sized to the resident capacity and each
//     wave fetches 64-ray batches from an atomic counter, so the divergent
//     tail of one batch never idles a whole CU;
//   - nodes are 64 B (both child boxes + children) read with four 16-byte
//     loads; sorted leaf triangles are 48 B (three 16-byte loads);
//   - each lane's traversal stack lives in LDS, laid out depth-major across
//     the block so lanes at any depth hit distinct banks.  Entries deeper
//     than kLdsStack spill to a per-thread global area; the Karras hierarchy
//     over 64-bit (code, index) keys is at most 64 levels deep, so 64 entries
//     always suffice.
#include "traverse.h"
#include "rt_math.h"

namespace bj {
namespace {

// Batched camera launches over a BLAS set: camera view e traces segment e,
// whose root follows from the segment offsets (one-triangle segments are a
// single leaf; an 8-wide segment's root is wide node e).
template <bool kWide>
BJ_D uint32_t segment_root(const uint64_t* __restrict__ off, uint64_t e) {
  if constexpr (kWide) return (uint32_t)e;
  const uint64_t first = off[e];
  return off[e + 1] - first == 1 ? (kLeafFlag | (uint32_t)first) : (uint32_t)first;
}

template <int kMode, bool kStats, bool kCamera, bool kWide, bool kSeg, bool kSph>
__global__ void __launch_bounds__(kTraceBlock) trace_kernel(const AccelView av,
                                                            const float* __restrict__ rays,
                                                            const float* __restrict__ cams, uint32_t cam_w,
                                                            uint32_t cam_h, float inv_height,
                                                            float* __restrict__ hits, uint64_t R,
                                                            uint32_t* __restrict__ stats,
                                                            uint32_t* __restrict__ counter,
                                                            uint32_t* __restrict__ spill,
                                                            const uint64_t* __restrict__ seg_off) {
  __shared__ uint32_t s_stack[kLdsStack * kTraceBlock];
  const uint32_t lane = lane_id();
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0};

  // Batches of 64 rays per atomic while the launch is short; up to eight
  // batches per fetch when every wave would otherwise fetch 8+ times.
  const uint64_t waves = (uint64_t)gridDim.x * (kTraceBlock / 64);
  uint64_t kb = R / (waves * 64u * 8u);
  const uint32_t batch = 64u * (uint32_t)(kb < 1 ? 1 : (kb > 8 ? 8 : kb));
  for (uint32_t base = 0, off = batch;;) {
    if (off >= batch) {
      if (lane == 0) base = atomicAdd(counter, batch);
      base = __shfl(base, 0, 64);
      off = 0;
    }
    if ((uint64_t)base + off >= R) break;
    const uint64_t r = (uint64_t)base + off + lane;
    off += 64;
    if (r >= R) continue;

    Ray ray;
    if constexpr (kCamera) {
      ray = camera_ray(cams, r, cam_w, cam_h, inv_height);
    } else {
      ray.o = mk(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
      ray.tmin = rays[3 * R + r];
      ray.d = mk(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
      ray.tmax = rays[7 * R + r];
    }
    TraceCounters tc;
    AccelView v = av;
    if constexpr (kSeg) {
      const uint32_t root = segment_root<kWide>(seg_off, r / ((uint64_t)cam_w * cam_h));
      if constexpr (kWide) v.wide.root = root;
      else v.bin.root = root;
    }
    Hit hit = trace_accel<kMode, kStats, kWide, kSph>(v, ray, st, &tc);
    store_hit(hits, r, hit);
    if constexpr (kStats) {
      uint32_t* s = stats + 5 * r;
      s[0] = tc.visits; s[1] = tc.inner; s[2] = 0; s[3] = tc.leaves; s[4] = tc.max_sp;
    }
  }
}

// 8-wide launches keep every lane busy: a lane whose ray is finished takes a
// new one as soon as a quarter of the wave is idle, instead of idling until
// the slowest ray of its batch ends (random rays differ in node visits by an
// order of magnitude, so fixed 64-ray batches left most lanes idle).  Each
// wave reserves kRefillChunk ray indices per atomic and hands them to idle
// lanes in lane order; one node visit per loop iteration.
constexpr uint32_t kRefillLanes = 16;
constexpr uint32_t kRefillChunk = 256;

template <int kMode, bool kStats, bool kCamera, bool kSeg, bool kSph>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6))) trace_refill_kernel(const WideView wv, const float* __restrict__ rays,
                                                                   const float* __restrict__ cams, uint32_t cam_w,
                                                                   uint32_t cam_h, float inv_height,
                                                                   float* __restrict__ hits, uint64_t R,
                                                                   uint32_t* __restrict__ stats,
                                                                   uint32_t* __restrict__ counter,
                                                                   uint32_t* __restrict__ spill,
                                                                   const uint64_t* __restrict__ seg_off) {
  __shared__ uint32_t s_stack[kLdsStack * kTraceBlock];
  const uint32_t lane = lane_id();
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0};
  Walk8<kMode, kStats, kSph> w;
  w.cur = kSentinel;
  bool active = false, drained = false;
  uint64_t r = 0, next = 0, stop = 0;  // the wave's reserved, not yet assigned indices [next, stop)
  // short launches are spread one 64-ray batch per wave, long ones take up to
  // kRefillChunk rays per atomic
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t per_wave = (R / waves + 63u) & ~63ull;
  const uint32_t chunk = per_wave < 64u ? 64u : (per_wave > kRefillChunk ? kRefillChunk : (uint32_t)per_wave);
  for (;;) {
    const uint64_t idle = ballot(!active);
    const uint32_t n_idle = (uint32_t)__builtin_popcountll(idle);
    if (!drained && n_idle >= kRefillLanes) {
      if (next >= stop) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(counter, chunk);
        base = __shfl(base, 0, 64);
        next = base;
        stop = (uint64_t)base + chunk < R ? (uint64_t)base + chunk : R;
        if ((uint64_t)base >= R) drained = true;
      }
      if (next < stop) {
        const uint64_t mine = next + mask_rank(idle);
        if (!active && mine < stop) {
          r = mine;
          Ray ray;
          if constexpr (kCamera) {
            ray = camera_ray(cams, r, cam_w, cam_h, inv_height);
          } else {
            ray.o = mk(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
            ray.tmin = rays[3 * R + r];
            ray.d = mk(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
            ray.tmax = rays[7 * R + r];
          }
          if constexpr (kSeg) {
            WideView sv = wv;
            sv.root = segment_root<true>(seg_off, r / ((uint64_t)cam_w * cam_h));
            w.begin(sv, ray, st);
          } else {
            w.begin(wv, ray, st);
          }
          active = true;
        }
        next = next + n_idle < stop ? next + n_idle : stop;
      }
    }
    if (!ballot(active)) {
      if (drained) break;
      continue;
    }
    if (active) {
      if (!w.done()) w.step(wv, st);
      if (w.done()) {
        TraceCounters tc;
        store_hit(hits, r, w.finish(wv, &tc));
        if constexpr (kStats) {
          uint32_t* sp = stats + 5 * r;
          sp[0] = tc.visits; sp[1] = tc.inner; sp[2] = 0; sp[3] = tc.leaves; sp[4] = tc.max_sp;
        }
        active = false;
      }
    }
  }
}

__global__ void trace_init_kernel(uint32_t* counter, uint32_t* status) {
  if (threadIdx.x == 0) { *counter = 0; *status = 0; }
}

struct TraceScratch {
  uint32_t* counter;
  uint32_t* status;
  uint32_t* spill;
  unsigned grid;
};

template <int kMode, bool kStats, bool kCamera, bool kWide, bool kSeg, bool kSph>
int32_t occupancy_grid(unsigned* grid) {
  static unsigned cached = 0;
  if (!cached) {
    int dev = 0, cus = 0, per_cu = 0;
    BJ_HIP(hipGetDevice(&dev));
    BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if constexpr (kWide)
      BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_refill_kernel<kMode, kStats, kCamera, kSeg, kSph>,
                                                          kTraceBlock, 0));
    else
      BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                          trace_kernel<kMode, kStats, kCamera, kWide, kSeg, kSph>,
                                                          kTraceBlock, 0));
    if (per_cu < 1) per_cu = 1;
    cached = (unsigned)(cus * per_cu);
  }
  *grid = cached;
  return BAJO_OK;
}

template <int kMode, bool kStats, bool kCamera, bool kWide, bool kSeg = false, bool kSph = false>
int32_t launch_width(bajo_ctx* ctx, const bajo_bvh* h, const float* rays, const float* cams, uint32_t w, uint32_t hh,
                     float* hits, uint64_t R, uint32_t* stats) {
  unsigned grid = 0;
  int32_t rc = occupancy_grid<kMode, kStats, kCamera, kWide, kSeg, kSph>(&grid);
  if (rc) return rc;
  uint64_t need_blocks = (R + kTraceBlock - 1) / kTraceBlock;
  if ((uint64_t)grid > need_blocks) grid = (unsigned)need_blocks;
  if (grid < 1) grid = 1;
  const uint64_t threads = (uint64_t)grid * kTraceBlock;
  void* scratch = nullptr;
  rc = ctx_scratch(ctx, 256 + threads * kSpill * sizeof(uint32_t), &scratch);
  if (rc) return rc;
  uint32_t* counter = (uint32_t*)scratch;
  uint32_t* status = counter + 1;
  uint32_t* spill = (uint32_t*)((char*)scratch + 256);
  trace_init_kernel<<<1, 64, 0, ctx->stream>>>(counter, status);
  const AccelView av = accel_view(h, status);
  const uint64_t* seg_off = kSeg ? h->seg_off : nullptr;
  if constexpr (kWide)
    trace_refill_kernel<kMode, kStats, kCamera, kSeg, kSph><<<grid, kTraceBlock, 0, ctx->stream>>>(
        av.wide, rays, cams, w, hh, hh ? 1.0f / (float)hh : 0.0f, hits, R, stats, counter, spill, seg_off);
  else
    trace_kernel<kMode, kStats, kCamera, kWide, kSeg, kSph><<<grid, kTraceBlock, 0, ctx->stream>>>(
        av, rays, cams, w, hh, hh ? 1.0f / (float)hh : 0.0f, hits, R, stats, counter, spill, seg_off);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

template <int kMode, bool kStats, bool kCamera>
int32_t launch(bajo_ctx* ctx, const bajo_bvh* h, const float* rays, const float* cams, uint32_t w, uint32_t hh,
               float* hits, uint64_t R, uint32_t* stats) {
  if (h->spheres) {
    if (wide_layout(h->node_width))
      return launch_width<kMode, kStats, kCamera, true, false, true>(ctx, h, rays, cams, w, hh, hits, R, stats);
    return launch_width<kMode, kStats, kCamera, false, false, true>(ctx, h, rays, cams, w, hh, hits, R, stats);
  }
  if (wide_layout(h->node_width))
    return launch_width<kMode, kStats, kCamera, true>(ctx, h, rays, cams, w, hh, hits, R, stats);
  return launch_width<kMode, kStats, kCamera, false>(ctx, h, rays, cams, w, hh, hits, R, stats);
}

// Ambient-occlusion rays of closest hits: bajo's _make_ao_ray
// (rt/gpu/path_shading.mojo:650-665) applied to a ray launch.  AO ray
// j = r * samples + k belongs to primary ray r and draws its direction from
// the Philox stream (seed, path j, stage 1): uniform on the sphere, flipped to
// the side of the normal oriented against the incoming ray
// (rt/geometry.mojo:20-30).  Origin o + t d, t in [0.001, 4.0].  A primary ray
// without a surface hit yields a ray with t_max = -1, which no traversal can
// hit.  Output is field-major like the input rays.
__global__ void __launch_bounds__(256) ao_rays_kernel(const float* __restrict__ rays, const float* __restrict__ hits,
                                                      uint64_t n, uint32_t samples, uint64_t seed,
                                                      float* __restrict__ out) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = n * samples;
  if (j >= m) return;
  const uint64_t r = j / samples;
  const float4* hp = reinterpret_cast<const float4*>(hits + 8 * r);
  const float4 h0 = hp[0], h1 = hp[1];
  const uint32_t prim = __builtin_bit_cast(uint32_t, h0.z);
  const v3 o = mk(rays[r], rays[n + r], rays[2 * n + r]);
  const v3 d = mk(rays[4 * n + r], rays[5 * n + r], rays[6 * n + r]);
  const float tmax = rays[7 * n + r];
  v3 po = mk(0.0f, 0.0f, 0.0f), dir = mk(0.0f, 0.0f, 1.0f);
  float t_hi = -1.0f;
  if (prim != kEmpty && h1.w < tmax) {
    const v3 hn = mk(h1.x, h1.y, h1.z);
    const v3 nn = dot(d, hn) < 0.0f ? hn : mk(-hn.x, -hn.y, -hn.z);
    po = add(o, scale(d, h1.w));
    StageRng rng(seed, (uint32_t)j, 1u);
    float sn, cs;
    sincos_turns(rng.next(), sn, cs);
    const float z = 1.0f - 2.0f * rng.next();
    const float rad = sqrtf(1.0f - z * z);
    dir = mk(rad * cs, rad * sn, z);
    if (dot(dir, nn) < 0.0f) dir = mk(-dir.x, -dir.y, -dir.z);
    t_hi = 4.0f;
  }
  const float f[8] = {po.x, po.y, po.z, 0.001f, dir.x, dir.y, dir.z, t_hi};
#pragma unroll
  for (int k = 0; k < 8; ++k) out[(uint64_t)k * m + j] = f[k];
}

}  // namespace
}  // namespace bj

extern "C" {

int32_t bajo_ao_rays(bajo_ctx* ctx, const float* rays, const float* hits, uint64_t ray_count, uint32_t samples,
                     uint64_t seed, float* ao_rays) {
  BJ_REQUIRE(ctx != nullptr, "bajo_ao_rays: NULL ctx");
  BJ_REQUIRE(rays != nullptr && hits != nullptr && ao_rays != nullptr, "bajo_ao_rays: NULL buffer");
  BJ_REQUIRE(ray_count > 0 && samples > 0, "bajo_ao_rays: ray and sample counts must be positive");
  BJ_REQUIRE(ray_count * samples < (1ull << 31), "bajo_ao_rays: ray_count * samples must be below 2^31");
  const uint64_t m = ray_count * samples;
  bj::ao_rays_kernel<<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(rays, hits, ray_count, samples, seed,
                                                                          ao_rays);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

int32_t bajo_trace_rays(bajo_ctx* ctx, const bajo_bvh* bvh, int32_t mode, const float* rays, float* hits,
                        uint64_t ray_count, uint32_t* stats) {
  BJ_REQUIRE(ctx != nullptr && bvh != nullptr, "bajo_trace_rays: NULL ctx/bvh");
  BJ_REQUIRE(mode == BAJO_TRACE_CLOSEST_HIT || mode == BAJO_TRACE_ANY_HIT, "bajo_trace_rays: bad mode %d", mode);
  BJ_REQUIRE(ray_count > 0, "bajo_trace_rays: ray count must be positive");
  BJ_REQUIRE(ray_count < (1ull << 31), "bajo_trace_rays: ray count must be below 2^31");
  BJ_REQUIRE(rays != nullptr && hits != nullptr, "bajo_trace_rays: NULL rays/hits");
  BJ_REQUIRE(bvh->n > 0, "bajo_trace_rays: empty BVH");
  BJ_REQUIRE(!bvh->boxes || bvh->spheres, "bajo_trace_rays: the hierarchy was built over boxes, not triangles or spheres");
  BJ_REQUIRE(!bvh->seg_count, "bajo_trace_rays: trace a BLAS set through its segment views (bajo_blas_set_get)");
  if (mode == BAJO_TRACE_ANY_HIT) {
    BJ_REQUIRE(stats == nullptr, "bajo_trace_rays: stats are collected for closest-hit queries only");
    return bj::launch<BAJO_TRACE_ANY_HIT, false, false>(ctx, bvh, rays, nullptr, 0, 0, hits, ray_count, nullptr);
  }
  if (stats) return bj::launch<BAJO_TRACE_CLOSEST_HIT, true, false>(ctx, bvh, rays, nullptr, 0, 0, hits, ray_count, stats);
  return bj::launch<BAJO_TRACE_CLOSEST_HIT, false, false>(ctx, bvh, rays, nullptr, 0, 0, hits, ray_count, nullptr);
}

int32_t bajo_trace_camera(bajo_ctx* ctx, const bajo_bvh* bvh, const float* camera, float* hits, uint64_t ray_count,
                          uint32_t width, uint32_t height, uint32_t* stats) {
  BJ_REQUIRE(ctx != nullptr && bvh != nullptr, "bajo_trace_camera: NULL ctx/bvh");
  BJ_REQUIRE(ray_count > 0 && width > 0 && height > 0, "bajo_trace_camera: launch dimensions must be positive");
  BJ_REQUIRE(ray_count < (1ull << 31), "bajo_trace_camera: ray count must be below 2^31");
  BJ_REQUIRE(camera != nullptr && hits != nullptr, "bajo_trace_camera: NULL camera/hits");
  BJ_REQUIRE(bvh->n > 0, "bajo_trace_camera: empty BVH");
  BJ_REQUIRE(!bvh->boxes || bvh->spheres, "bajo_trace_camera: the hierarchy was built over boxes, not triangles or spheres");
  BJ_REQUIRE(!bvh->seg_count, "bajo_trace_camera: trace a BLAS set through its segment views (bajo_blas_set_get)");
  if (stats)
    return bj::launch<BAJO_TRACE_CLOSEST_HIT, true, true>(ctx, bvh, nullptr, camera, width, height, hits, ray_count, stats);
  return bj::launch<BAJO_TRACE_CLOSEST_HIT, false, true>(ctx, bvh, nullptr, camera, width, height, hits, ray_count,
                                                         nullptr);
}

int32_t bajo_blas_set_trace_camera(bajo_ctx* ctx, const bajo_blas_set* set, const float* cameras, float* hits,
                                   uint32_t width, uint32_t height, uint32_t* stats) {
  BJ_REQUIRE(ctx != nullptr && set != nullptr, "bajo_blas_set_trace_camera: NULL ctx/set");
  BJ_REQUIRE(cameras != nullptr && hits != nullptr, "bajo_blas_set_trace_camera: NULL cameras/hits");
  BJ_REQUIRE(width > 0 && height > 0, "bajo_blas_set_trace_camera: launch dimensions must be positive");
  const bajo_bvh* h = bj::blas_set_store(set);
  const uint64_t R = (uint64_t)h->seg_count * width * height;
  BJ_REQUIRE(R < (1ull << 31), "bajo_blas_set_trace_camera: %u segments x %u x %u rays exceed 2^31", h->seg_count,
             width, height);
  constexpr int kC = BAJO_TRACE_CLOSEST_HIT;
  if (wide_layout(h->node_width)) {
    if (stats) return bj::launch_width<kC, true, true, true, true>(ctx, h, nullptr, cameras, width, height, hits, R, stats);
    return bj::launch_width<kC, false, true, true, true>(ctx, h, nullptr, cameras, width, height, hits, R, nullptr);
  }
  if (stats) return bj::launch_width<kC, true, true, false, true>(ctx, h, nullptr, cameras, width, height, hits, R, stats);
  return bj::launch_width<kC, false, true, false, true>(ctx, h, nullptr, cameras, width, height, hits, R, nullptr);
}

int32_t bajo_camera_make(const float* o3, const float* t3, const float* u3, float fov_scale, float focus_dist,
                         float* out20) {
  BJ_REQUIRE(o3 && t3 && u3 && out20, "bajo_camera_make: NULL argument");
  using namespace bj;
  v3 o = mk(o3[0], o3[1], o3[2]);
  v3 fwd = normalize(sub(mk(t3[0], t3[1], t3[2]), o));
  v3 right = normalize(cross(fwd, mk(u3[0], u3[1], u3[2])));
  v3 up = normalize(cross(right, fwd));
  float v[20] = {o.x, o.y, o.z, fwd.x, fwd.y, fwd.z, right.x, right.y, right.z, up.x, up.y, up.z,
                 fov_scale, focus_dist, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 20; ++i) out20[i] = v[i];
  return BAJO_OK;
}

}  // extern "C"
