// BVH2 closest-hit / any-hit traversal for gfx950.
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
#include "bvh_internal.h"

namespace bj {
namespace {

constexpr int kTraceBlock = 256;
constexpr int kLdsStack = 24;
constexpr int kMaxStack = 64;
constexpr int kSpill = kMaxStack - kLdsStack;

struct Ray {
  v3 o, d;
  float tmin, tmax;
};

struct Hit {
  float u, v;
  uint32_t prim, inst;
  v3 n;
  float t;
};

BJ_D Hit miss(float t) { return Hit{0.0f, 0.0f, kEmpty, kEmpty, mk(0, 0, 0), t}; }

BJ_D void store_hit(float* hits, uint64_t r, const Hit& h) {
  float4* p = reinterpret_cast<float4*>(hits + 8 * r);
  p[0] = make_float4(h.u, h.v, __builtin_bit_cast(float, h.prim), __builtin_bit_cast(float, h.inst));
  p[1] = make_float4(h.n.x, h.n.y, h.n.z, h.t);
}

// core/intersect.mojo:193-215 with select min/max.
BJ_D bool slab(v3 o, v3 rcp, float lx, float ly, float lz, float hx, float hy, float hz, float t_max, float& t_out) {
  float t0x = (lx - o.x) * rcp.x, t0y = (ly - o.y) * rcp.y, t0z = (lz - o.z) * rcp.z;
  float t1x = (hx - o.x) * rcp.x, t1y = (hy - o.y) * rcp.y, t1z = (hz - o.z) * rcp.z;
  float tnx = fmin_sel(t0x, t1x), tny = fmin_sel(t0y, t1y), tnz = fmin_sel(t0z, t1z);
  float tfx = fmax_sel(t0x, t1x), tfy = fmax_sel(t0y, t1y), tfz = fmax_sel(t0z, t1z);
  float tmin = fmax_sel(fmax_sel(tnx, tny), fmax_sel(tnz, 0.0f));
  float tmax = fmin_sel(fmin_sel(tfx, tfy), fmin_sel(tfz, t_max));
  t_out = tmin;
  return tmin <= tmax;
}

struct TriData {
  v3 v0, e1, e2;
  uint32_t prim;
};

BJ_D TriData load_tri(const BjTri* __restrict__ tris, uint32_t s) {
  const float4* p = reinterpret_cast<const float4*>(tris + s);
  float4 a = p[0], b = p[1], c = p[2];
  return TriData{mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), __builtin_bit_cast(uint32_t, a.w)};
}

// core/intersect.mojo:462-508.  Returns the mask; fills the scaled values.
BJ_D bool tri_scaled(const Ray& r, const TriData& tr, float t_max, float& abs_det, float& ts, float& us, float& vs) {
  v3 p = cross(r.d, tr.e2);
  float det = dot(tr.e1, p);
  bool pos = det >= 0.0f;
  abs_det = pos ? det : -det;
  v3 tv = sub(r.o, tr.v0);
  float u_num = dot(tv, p);
  us = pos ? u_num : -u_num;
  v3 q = cross(tv, tr.e1);
  float v_num = dot(r.d, q);
  float t_num = dot(tr.e2, q);
  vs = pos ? v_num : -v_num;
  ts = pos ? t_num : -t_num;
  return abs_det > 1e-8f && us >= 0.0f && us <= abs_det && vs >= 0.0f && (us + vs) <= abs_det &&
         ts > r.tmin * abs_det && ts < t_max * abs_det;
}

BJ_D bool leaf_closest(const BjTri* __restrict__ tris, uint32_t s, const Ray& r, Hit& hit) {
  TriData tr = load_tri(tris, s);
  float ad, ts, us, vs;
  if (!tri_scaled(r, tr, hit.t, ad, ts, us, vs)) return false;
  float inv = 1.0f / ad;
  hit.t = ts * inv;
  hit.u = us * inv;
  hit.v = vs * inv;
  hit.prim = tr.prim;
  hit.inst = kEmpty;
  hit.n = normalize(cross(tr.e1, tr.e2));
  return true;
}

BJ_D bool leaf_any(const BjTri* __restrict__ tris, uint32_t s, const Ray& r) {
  TriData tr = load_tri(tris, s);
  float ad, ts, us, vs;
  return tri_scaled(r, tr, r.tmax, ad, ts, us, vs);
}

struct Stack {
  uint32_t* lds;       // this lane's column: entry k at lds[k * kTraceBlock]
  uint32_t* spill;     // this thread's column: entry k at spill[k * spill_stride]
  uint64_t spill_stride;
  int sp;
  BJ_D void push(uint32_t v) {
    if (sp < kLdsStack) lds[sp * kTraceBlock] = v;
    else spill[(uint64_t)(sp - kLdsStack) * spill_stride] = v;
    ++sp;
  }
  BJ_D uint32_t pop() {
    --sp;
    return sp < kLdsStack ? lds[sp * kTraceBlock] : spill[(uint64_t)(sp - kLdsStack) * spill_stride];
  }
};

// bvh/camera.mojo:137-154, camera record bvh/camera.mojo:19-30.
BJ_D Ray camera_ray(const float* __restrict__ cams, uint64_t r, uint32_t w, uint32_t h, float inv_height) {
  uint64_t ppv = (uint64_t)w * h;
  uint64_t view = r / ppv, local = r - view * ppv;
  const float* c = cams + 20 * view;
  uint32_t px = (uint32_t)(local % w), py = (uint32_t)(local / w);
  float sx = (2.0f * ((float)px + 0.5f) - (float)w) * inv_height;
  float sy = 1.0f - 2.0f * ((float)py + 0.5f) * inv_height;
  float fov = c[12];
  v3 dir = add(add(mk(c[3], c[4], c[5]), scale(mk(c[6], c[7], c[8]), sx * fov)), scale(mk(c[9], c[10], c[11]), sy * fov));
  return Ray{mk(c[0], c[1], c[2]), normalize(dir), 0.0f, kFltMax};
}

template <int kMode, bool kStats, bool kCamera>
__global__ void __launch_bounds__(kTraceBlock) trace_kernel(const BjNode* __restrict__ nodes,
                                                            const BjTri* __restrict__ tris, uint32_t root,
                                                            uint32_t node_count, uint32_t prim_count,
                                                            const float* __restrict__ rays,
                                                            const float* __restrict__ cams, uint32_t cam_w,
                                                            uint32_t cam_h, float inv_height,
                                                            float* __restrict__ hits, uint64_t R,
                                                            uint32_t* __restrict__ stats,
                                                            uint32_t* __restrict__ counter,
                                                            uint32_t* __restrict__ spill, uint32_t* __restrict__ status) {
  __shared__ uint32_t s_stack[kLdsStack * kTraceBlock];
  const uint32_t lane = lane_id();
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0};

  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, 64u);
    base = __shfl(base, 0, 64);
    if ((uint64_t)base >= R) break;
    const uint64_t r = (uint64_t)base + lane;
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
    uint32_t n_visit = 0, n_inner = 0, n_leaf = 0, max_sp = 0;
    Hit hit = miss(ray.tmax);
    const v3 rcp = rcp_dir(ray.d);
    st.sp = 0;
    bool occluded = false;

    if ((root & kLeafFlag) && root != kSentinel) {
      n_leaf++;
      if constexpr (kMode == BAJO_TRACE_ANY_HIT) occluded = leaf_any(tris, root & kIndexMask, ray);
      else leaf_closest(tris, root & kIndexMask, ray, hit);
    } else if (root != kSentinel) {
      uint32_t cur = root;
      for (;;) {
        // A valid hierarchy visits each node at most once per ray; the cap only
        // ends a ray early (with a status word) if the hierarchy is corrupted.
        if (cur >= node_count || n_visit > node_count) { atomicOr(status, (uint32_t)BAJO_STATUS_BAD_NODE); break; }
        n_visit++;
        const float4* np = reinterpret_cast<const float4*>(nodes + cur);
        float4 q0 = np[0], q1 = np[1], q2 = np[2];
        uint4 m = reinterpret_cast<const uint4*>(nodes + cur)[3];
        const float node_tmax = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : hit.t;
        float t0, t1;
        bool h0 = slab(ray.o, rcp, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, node_tmax, t0);
        bool h1 = slab(ray.o, rcp, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, node_tmax, t1);
        const uint32_t c0 = m.y, c1 = m.z;
        const bool l0 = (c0 & kLeafFlag) && (c0 & kIndexMask) < prim_count;
        const bool l1 = (c1 & kLeafFlag) && (c1 & kIndexMask) < prim_count;
        if constexpr (kMode == BAJO_TRACE_ANY_HIT) {
          if (h0) {
            if (l0) {
              n_leaf++;
              if (leaf_any(tris, c0 & kIndexMask, ray)) { occluded = true; break; }
            } else if (!(c0 & kLeafFlag)) {
              if (st.sp >= kMaxStack) { atomicOr(status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW); break; }
              st.push(c0);
            }
          }
          if (h1) {
            if (l1) {
              n_leaf++;
              if (leaf_any(tris, c1 & kIndexMask, ray)) { occluded = true; break; }
            } else if (!(c1 & kLeafFlag)) {
              if (st.sp >= kMaxStack) { atomicOr(status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW); break; }
              st.push(c1);
            }
          }
          if (st.sp == 0) break;
          cur = st.pop();
          continue;
        } else {
          bool v0 = false, v1 = false;
          if (h0) {
            if (l0) { n_leaf++; leaf_closest(tris, c0 & kIndexMask, ray, hit); }
            else if (!(c0 & kLeafFlag)) { v0 = true; n_inner++; }
          }
          if (h1) {
            if (l1) { n_leaf++; leaf_closest(tris, c1 & kIndexMask, ray, hit); }
            else if (!(c1 & kLeafFlag)) { v1 = true; n_inner++; }
          }
          // bvh/gpu/trace.mojo:395-426: nearest child first, push the other.
          int near = -1;
          float near_t = kFltMax;
          if (v0) { near = 0; near_t = t0; }
          if (v1 && (near == -1 || t1 < near_t)) { near = 1; near_t = t1; }
          bool overflow = false;
          if (v0 && near != 0 && t0 <= hit.t) {
            if (st.sp >= kMaxStack) overflow = true; else st.push(c0);
          }
          if (v1 && near != 1 && t1 <= hit.t) {
            if (st.sp >= kMaxStack) overflow = true; else st.push(c1);
          }
          if (overflow) { atomicOr(status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW); break; }
          if (kStats && (uint32_t)st.sp > max_sp) max_sp = (uint32_t)st.sp;
          if (near != -1 && near_t <= hit.t) { cur = near == 0 ? c0 : c1; continue; }
          if (st.sp == 0) break;
          cur = st.pop();
        }
      }
    }
    if constexpr (kMode == BAJO_TRACE_ANY_HIT) {
      // bvh/gpu/trace.mojo:44-59: occlusion is reported as Hit.shadow_hit().
      hit = occluded ? Hit{0.0f, 0.0f, kEmpty, kEmpty, mk(0, 0, 0), 0.0f} : miss(ray.tmax);
    }
    store_hit(hits, r, hit);
    if constexpr (kStats) {
      uint32_t* s = stats + 5 * r;
      s[0] = n_visit; s[1] = n_inner; s[2] = 0; s[3] = n_leaf; s[4] = max_sp;
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

template <int kMode, bool kStats, bool kCamera>
int32_t occupancy_grid(unsigned* grid) {
  static unsigned cached = 0;
  if (!cached) {
    int dev = 0, cus = 0, per_cu = 0;
    BJ_HIP(hipGetDevice(&dev));
    BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_kernel<kMode, kStats, kCamera>, kTraceBlock, 0));
    if (per_cu < 1) per_cu = 1;
    cached = (unsigned)(cus * per_cu);
  }
  *grid = cached;
  return BAJO_OK;
}

template <int kMode, bool kStats, bool kCamera>
int32_t launch(bajo_ctx* ctx, const bajo_bvh* h, const float* rays, const float* cams, uint32_t w, uint32_t hh,
               float* hits, uint64_t R, uint32_t* stats) {
  unsigned grid = 0;
  int32_t rc = occupancy_grid<kMode, kStats, kCamera>(&grid);
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
  trace_kernel<kMode, kStats, kCamera><<<grid, kTraceBlock, 0, ctx->stream>>>(
      h->nodes, h->tris, h->root, (uint32_t)(h->n - 1), (uint32_t)h->n, rays, cams, w, hh, hh ? 1.0f / (float)hh : 0.0f, hits, R, stats, counter, spill,
      status);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

}  // namespace
}  // namespace bj

extern "C" {

int32_t bajo_trace_rays(bajo_ctx* ctx, const bajo_bvh* bvh, int32_t mode, const float* rays, float* hits,
                        uint64_t ray_count, uint32_t* stats) {
  BJ_REQUIRE(ctx != nullptr && bvh != nullptr, "bajo_trace_rays: NULL ctx/bvh");
  BJ_REQUIRE(mode == BAJO_TRACE_CLOSEST_HIT || mode == BAJO_TRACE_ANY_HIT, "bajo_trace_rays: bad mode %d", mode);
  BJ_REQUIRE(ray_count > 0, "bajo_trace_rays: ray count must be positive");
  BJ_REQUIRE(ray_count < 0xFFFFFF00ull, "bajo_trace_rays: ray count exceeds 2^32-256");
  BJ_REQUIRE(rays != nullptr && hits != nullptr, "bajo_trace_rays: NULL rays/hits");
  BJ_REQUIRE(bvh->n > 0, "bajo_trace_rays: empty BVH");
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
  BJ_REQUIRE(ray_count < 0xFFFFFF00ull, "bajo_trace_camera: ray count exceeds 2^32-256");
  BJ_REQUIRE(camera != nullptr && hits != nullptr, "bajo_trace_camera: NULL camera/hits");
  BJ_REQUIRE(bvh->n > 0, "bajo_trace_camera: empty BVH");
  if (stats)
    return bj::launch<BAJO_TRACE_CLOSEST_HIT, true, true>(ctx, bvh, nullptr, camera, width, height, hits, ray_count, stats);
  return bj::launch<BAJO_TRACE_CLOSEST_HIT, false, true>(ctx, bvh, nullptr, camera, width, height, hits, ray_count,
                                                         nullptr);
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
