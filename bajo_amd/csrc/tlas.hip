// Instances of triangle or sphere BVHs under a top-level hierarchy (gfx950).
//
// Replaces build_triangle_tlas / build_sphere_tlas and their launch_camera
// (bvh/gpu/tlas.mojo:54-233, 482-715) with bajo's Instance (bvh/types.mojo:
// 153-190), Affine3 (core/transform.mojo: point / vector / inverse /
// normal_from_inverse) and AABB.apply_transform (core/aabb.mojo:101-137).
//
// MI355X design:
//   - the top level is this library's box builder (bajo_build_binary_bounds)
//     over the instances' world boxes, so it is the same LBVH / H-PLOC
//     layout the parity tests already pin;
//   - one persistent grid walks the top level near-first; at an instance
//     leaf the ray is mapped into BLAS space with the inverse transform (t is
//     unchanged) and the BLAS, binary or 8-wide, is traversed with the same
//     code as a plain launch, its stack in LDS; the top-level stack is a
//     per-thread column in global memory (few top-level nodes per ray);
//   - instance transforms, BLAS views and instance-to-BLAS indices live in
//     device arrays uploaded once at creation; the BLAS handles stay owned
//     by the caller.
#include <vector>

#include "traverse.h"

struct TlasBlas {
  bj::BvhView bin;
  bj::WideView wide;
  uint32_t is_wide, is_sphere;
  uint32_t prim_count;
};

struct bajo_tlas {
  int device = 0;
  uint32_t instance_count = 0, blas_count = 0;
  bajo_bvh* top = nullptr;        // boxes hierarchy over the instance world boxes
  float* boxes = nullptr;         // f32[6 I] instance world boxes
  float* inv = nullptr;           // f32[12 I] world -> BLAS transforms
  uint32_t* inst_blas = nullptr;  // u32[I]
  TlasBlas* blas = nullptr;       // device views of the BLASes
  uint32_t* status = nullptr;     // traversal status word
  std::vector<float> host_boxes;
};

namespace bj {
namespace {

constexpr int kTlasStack = 64;

struct TlasView {
  const BjNode* nodes;
  const uint64_t* keys;  // sorted keys of the top level: low word = instance index
  uint32_t root, node_count, instance_count;
  const float* inv;
  const uint32_t* inst_blas;
  const TlasBlas* blas;
  uint32_t* status;
};

// Affine3.point / .vector (core/transform.mojo:155-173) with a row-major 3x4.
BJ_D v3 xf_point(const float* m, v3 p) {
  return mk(fmaf(m[0], p.x, fmaf(m[1], p.y, fmaf(m[2], p.z, m[3]))),
            fmaf(m[4], p.x, fmaf(m[5], p.y, fmaf(m[6], p.z, m[7]))),
            fmaf(m[8], p.x, fmaf(m[9], p.y, fmaf(m[10], p.z, m[11]))));
}
BJ_D v3 xf_vector(const float* m, v3 v) {
  return mk(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[4], v.x, fmaf(m[5], v.y, m[6] * v.z)),
            fmaf(m[8], v.x, fmaf(m[9], v.y, m[10] * v.z)));
}
// Affine3.normal_from_inverse (core/transform.mojo:224-251): the inverse's
// transpose applied to the normal, normalized.
BJ_D v3 normal_from_inverse(const float* m, v3 n) {
  return normalize(mk(fmaf(m[0], n.x, fmaf(m[4], n.y, m[8] * n.z)), fmaf(m[1], n.x, fmaf(m[5], n.y, m[9] * n.z)),
                      fmaf(m[2], n.x, fmaf(m[6], n.y, m[10] * n.z))));
}

// _trace_tlas_ray (bvh/gpu/tlas.mojo:176-233): near-first over the top
// level; each instance leaf traverses its BLAS in BLAS space with t_max =
// the closest hit so far; a closer BLAS hit is promoted with inst = the
// instance (bvh/tlas_common.mojo:8-30); the winning normal is transformed
// once at the end.  Any hit: bajo's shadow hit carrying the instance.
template <int kMode>
BJ_D Hit tlas_trace(const TlasView& tv, const Ray& ray, Stack& st, uint32_t* ts, uint64_t ts_stride) {
  Hit hit = miss(ray.tmax);
  uint32_t occ_inst = kEmpty;
  // returns true once an any-hit query is answered
  auto visit = [&](uint32_t pos) -> bool {
    const uint32_t inst = pos < tv.instance_count ? (uint32_t)tv.keys[pos] : kEmpty;
    if (inst >= tv.instance_count) {
      atomicOr(tv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
      return false;
    }
    const TlasBlas& b = tv.blas[tv.inst_blas[inst]];
    if (b.prim_count == 0) return false;
    const float* m = tv.inv + 12 * (uint64_t)inst;
    Ray lr;
    lr.o = xf_point(m, ray.o);
    lr.d = xf_vector(m, ray.d);
    lr.tmin = ray.tmin;
    lr.tmax = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : hit.t;
    Hit lh;
    if (b.is_sphere)
      lh = b.is_wide ? traverse8<kMode, false, true>(b.wide, lr, st, nullptr)
                     : traverse<kMode, false, true>(b.bin, lr, st, nullptr);
    else
      lh = b.is_wide ? traverse8<kMode, false>(b.wide, lr, st, nullptr) : traverse<kMode, false>(b.bin, lr, st, nullptr);
    if constexpr (kMode == BAJO_TRACE_ANY_HIT) {
      if (lh.prim == kEmpty && lh.t == 0.0f) {
        occ_inst = inst;
        return true;
      }
      return false;
    } else {
      if (lh.prim != kEmpty && lh.t < hit.t) {
        hit = lh;
        hit.inst = inst;
      }
      return false;
    }
  };
  bool done = false;
  if (tv.root & kLeafFlag) {
    done = visit(tv.root & kIndexMask);
  } else if (tv.root != kSentinel) {
    const v3 rcp = rcp_dir(ray.d);
    uint32_t cur = tv.root, n_visit = 0;
    int sp = 0;
    for (;;) {
      if (cur >= tv.node_count || ++n_visit > tv.node_count) {
        atomicOr(tv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
        break;
      }
      const float4* np = reinterpret_cast<const float4*>(tv.nodes + cur);
      const float4 q0 = np[0], q1 = np[1], q2 = np[2];
      const uint4 m = reinterpret_cast<const uint4*>(tv.nodes + cur)[3];
      const float node_tmax = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : hit.t;
      float t0, t1;
      const bool h0 = slab(ray.o, rcp, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, node_tmax, t0);
      const bool h1 = slab(ray.o, rcp, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, node_tmax, t1);
      const uint32_t c0 = m.y, c1 = m.z;
      bool v0 = false, v1 = false;
      if (h0) {
        if (c0 & kLeafFlag) done = visit(c0 & kIndexMask);
        else v0 = true;
      }
      if (!done && h1) {
        if (c1 & kLeafFlag) done = visit(c1 & kIndexMask);
        else v1 = true;
      }
      if (done) break;
      int near = -1;
      float near_t = kFltMax;
      if (v0) { near = 0; near_t = t0; }
      if (v1 && (near == -1 || t1 < near_t)) { near = 1; near_t = t1; }
      const float bound = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : hit.t;
      bool overflow = false;
      if (v0 && near != 0 && t0 <= bound) {
        if (sp >= kTlasStack) overflow = true; else ts[(uint64_t)(sp++) * ts_stride] = c0;
      }
      if (v1 && near != 1 && t1 <= bound) {
        if (sp >= kTlasStack) overflow = true; else ts[(uint64_t)(sp++) * ts_stride] = c1;
      }
      if (overflow) {
        atomicOr(tv.status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW);
        break;
      }
      if (near != -1 && near_t <= bound) {
        cur = near == 0 ? c0 : c1;
        continue;
      }
      if (sp == 0) break;
      cur = ts[(uint64_t)(--sp) * ts_stride];
    }
  }
  if constexpr (kMode == BAJO_TRACE_ANY_HIT) {
    if (done) return Hit{0.0f, 0.0f, kEmpty, occ_inst, mk(0, 0, 0), 0.0f};
    return miss(ray.tmax);
  } else {
    if (hit.prim != kEmpty) hit.n = normal_from_inverse(tv.inv + 12 * (uint64_t)hit.inst, hit.n);
    return hit;
  }
}

template <int kMode, bool kCamera>
__global__ void __launch_bounds__(kTraceBlock) tlas_kernel(const TlasView tv, const float* __restrict__ rays,
                                                           const float* __restrict__ cams, uint32_t cam_w,
                                                           uint32_t cam_h, float inv_height, float* __restrict__ hits,
                                                           uint64_t R, uint32_t* __restrict__ tstack,
                                                           uint32_t* __restrict__ spill) {
  __shared__ uint32_t s_stack[kLdsStack * kTraceBlock];
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t threads = (uint64_t)gridDim.x * blockDim.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, threads, 0};
  for (uint64_t r = gtid; r < R; r += threads) {
    Ray ray;
    if constexpr (kCamera) {
      ray = camera_ray(cams, r, cam_w, cam_h, inv_height);
    } else {
      ray.o = mk(rays[0 * R + r], rays[1 * R + r], rays[2 * R + r]);
      ray.tmin = rays[3 * R + r];
      ray.d = mk(rays[4 * R + r], rays[5 * R + r], rays[6 * R + r]);
      ray.tmax = rays[7 * R + r];
    }
    store_hit(hits, r, tlas_trace<kMode>(tv, ray, st, tstack + gtid, threads));
  }
}

__global__ void tlas_status_kernel(uint32_t* status) {
  if (threadIdx.x == 0) *status = 0;
}

// Affine3.inverse (core/transform.mojo:290-340): adjugate over the
// determinant, translation -inv(M) t; false when |det| <= 1e-8.
bool affine_inverse(const float* a, float* o) {
  const float m00 = a[0], m01 = a[1], m02 = a[2], tx = a[3];
  const float m10 = a[4], m11 = a[5], m12 = a[6], ty = a[7];
  const float m20 = a[8], m21 = a[9], m22 = a[10], tz = a[11];
  float i00 = m11 * m22 - m12 * m21, i01 = m02 * m21 - m01 * m22, i02 = m01 * m12 - m02 * m11;
  float i10 = m12 * m20 - m10 * m22, i11 = m00 * m22 - m02 * m20, i12 = m02 * m10 - m00 * m12;
  float i20 = m10 * m21 - m11 * m20, i21 = m01 * m20 - m00 * m21, i22 = m00 * m11 - m01 * m10;
  const float det = m00 * i00 + m01 * i10 + m02 * i20;
  const bool ok = fabsf(det) > 1.0e-8f;
  const float r = 1.0f / (ok ? det : 1.0f);
  i00 *= r; i01 *= r; i02 *= r;
  i10 *= r; i11 *= r; i12 *= r;
  i20 *= r; i21 *= r; i22 *= r;
  const float v[12] = {i00, i01, i02, -(i00 * tx + i01 * ty + i02 * tz),
                       i10, i11, i12, -(i10 * tx + i11 * ty + i12 * tz),
                       i20, i21, i22, -(i20 * tx + i21 * ty + i22 * tz)};
  for (int k = 0; k < 12; ++k) o[k] = v[k];
  return ok;
}

// AABB.apply_transform (core/aabb.mojo:101-137): the translation plus, per
// column, the smaller / larger of the column scaled by the box's extremes.
void transform_box(const float* m, const float* b, float* out) {
  float lo[3] = {m[3], m[7], m[11]}, hi[3] = {m[3], m[7], m[11]};
  for (int c = 0; c < 3; ++c) {
    for (int r = 0; r < 3; ++r) {
      const float col = m[4 * r + c];
      const float a = col * b[c], z = col * b[3 + c];
      lo[r] = lo[r] + fmin_sel(a, z);
      hi[r] = hi[r] + fmax_sel(a, z);
    }
  }
  for (int k = 0; k < 3; ++k) {
    out[k] = lo[k];
    out[3 + k] = hi[k];
  }
}

void free_tlas(bajo_tlas* t) {
  if (!t) return;
  if (t->top) (void)bajo_bvh_free(t->top);
  (void)hipFree(t->boxes);
  (void)hipFree(t->inv);
  (void)hipFree(t->inst_blas);
  (void)hipFree(t->blas);
  (void)hipFree(t->status);
  delete t;
}

template <int kMode, bool kCamera>
int32_t launch_tlas(bajo_ctx* ctx, const bajo_tlas* t, const float* rays, const float* cams, uint32_t w, uint32_t hh,
                    float* hits, uint64_t R) {
  static unsigned cached = 0;
  if (!cached) {
    int dev = 0, cus = 0, per_cu = 0;
    BJ_HIP(hipGetDevice(&dev));
    BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tlas_kernel<kMode, kCamera>, kTraceBlock, 0));
    cached = (unsigned)(cus * (per_cu < 1 ? 1 : per_cu));
  }
  unsigned grid = cached;
  const uint64_t need = (R + kTraceBlock - 1) / kTraceBlock;
  if ((uint64_t)grid > need) grid = (unsigned)need;
  if (grid < 1) grid = 1;
  const uint64_t threads = (uint64_t)grid * kTraceBlock;
  void* scratch = nullptr;
  int32_t rc = ctx_scratch(ctx, threads * (kSpill + kTlasStack) * sizeof(uint32_t), &scratch);
  if (rc) return rc;
  uint32_t* spill = (uint32_t*)scratch;
  uint32_t* tstack = spill + threads * kSpill;
  const bajo_bvh* top = t->top;
  TlasView tv{top->nodes, top->keys, top->root, (uint32_t)(top->n - 1), t->instance_count, t->inv, t->inst_blas,
              t->blas, t->status};
  tlas_status_kernel<<<1, 64, 0, ctx->stream>>>(t->status);
  tlas_kernel<kMode, kCamera><<<grid, kTraceBlock, 0, ctx->stream>>>(tv, rays, cams, w, hh, hh ? 1.0f / (float)hh : 0.0f,
                                                                    hits, R, tstack, spill);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

}  // namespace
}  // namespace bj

extern "C" {

int32_t bajo_tlas_create(bajo_ctx* ctx, const bajo_bvh* const* blases, uint32_t blas_count, const float* transforms,
                         const uint32_t* blas_indices, uint32_t instance_count, const bajo_build_options* options,
                         bajo_tlas** out) {
  using namespace bj;
  BJ_REQUIRE(ctx && out && blases && transforms && blas_indices, "bajo_tlas_create: NULL argument");
  BJ_REQUIRE(blas_count > 0 && instance_count > 0, "bajo_tlas_create: a TLAS needs BLASes and instances");
  BJ_REQUIRE(instance_count < (1u << 28), "bajo_tlas_create: too many instances");
  bajo_build_options opt = options ? *options : bajo_build_options{BAJO_BUILD_LBVH, BAJO_WIDTH_BINARY, 0, 0};
  BJ_REQUIRE(opt.node_width == BAJO_WIDTH_BINARY, "bajo_tlas_create: the top level is binary (node_width 2)");
  std::vector<TlasBlas> views(blas_count);
  std::vector<float> roots(6 * (size_t)blas_count);
  for (uint32_t b = 0; b < blas_count; ++b) {
    const bajo_bvh* h = blases[b];
    BJ_REQUIRE(h != nullptr && (!h->boxes || h->spheres) && h->n > 0 && !h->seg_count,
               "bajo_tlas_create: BLAS %u is not a triangle or sphere BVH (a BLAS set enters through its segment "
               "views)", b);
    BJ_REQUIRE(h->device == ctx->device, "bajo_tlas_create: BLAS %u belongs to device %d", b, h->device);
    BJ_HIP(hipMemcpyAsync(roots.data() + 6 * (size_t)b, h->scene, 6 * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
  }
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  bajo_tlas* t = new bajo_tlas();
  t->device = ctx->device;
  t->instance_count = instance_count;
  t->blas_count = blas_count;
  auto fail = [&](int32_t rc) {
    free_tlas(t);
    return rc;
  };
  std::vector<float> inv(12 * (size_t)instance_count);
  t->host_boxes.resize(6 * (size_t)instance_count);
  for (uint32_t i = 0; i < instance_count; ++i) {
    if (blas_indices[i] >= blas_count) {
      free_tlas(t);
      BJ_REQUIRE(false, "bajo_tlas_create: instance %u references BLAS %u of %u", i, blas_indices[i], blas_count);
    }
    if (!affine_inverse(transforms + 12 * (size_t)i, inv.data() + 12 * (size_t)i)) {
      free_tlas(t);
      BJ_REQUIRE(false, "bajo_tlas_create: the transform of instance %u is not invertible", i);
    }
    transform_box(transforms + 12 * (size_t)i, roots.data() + 6 * (size_t)blas_indices[i],
                  t->host_boxes.data() + 6 * (size_t)i);
  }
#define TL_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return fail(hip_fail(e_, #call, __FILE__, __LINE__));   \
  } while (0)
  TL_HIP(hipMalloc(&t->status, sizeof(uint32_t)));
  for (uint32_t b = 0; b < blas_count; ++b) {
    const AccelView av = accel_view(blases[b], t->status);
    views[b] = TlasBlas{av.bin, av.wide, wide_layout(blases[b]->node_width) ? 1u : 0u,
                        blases[b]->spheres ? 1u : 0u, (uint32_t)blases[b]->n};
  }
  TL_HIP(hipMalloc(&t->boxes, t->host_boxes.size() * sizeof(float)));
  TL_HIP(hipMalloc(&t->inv, inv.size() * sizeof(float)));
  TL_HIP(hipMalloc(&t->inst_blas, (size_t)instance_count * sizeof(uint32_t)));
  TL_HIP(hipMalloc(&t->blas, views.size() * sizeof(TlasBlas)));
  TL_HIP(hipMemcpy(t->boxes, t->host_boxes.data(), t->host_boxes.size() * sizeof(float), hipMemcpyHostToDevice));
  TL_HIP(hipMemcpy(t->inv, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice));
  TL_HIP(hipMemcpy(t->inst_blas, blas_indices, (size_t)instance_count * sizeof(uint32_t), hipMemcpyHostToDevice));
  TL_HIP(hipMemcpy(t->blas, views.data(), views.size() * sizeof(TlasBlas), hipMemcpyHostToDevice));
#undef TL_HIP
  int32_t rc = bajo_build_binary_bounds(ctx, t->boxes, instance_count, &opt, nullptr, &t->top);
  if (rc) return fail(rc);
  rc = resolve_build(t->top);  // the top-level root is passed by value to traversals
  if (rc) return fail(rc);
  *out = t;
  return BAJO_OK;
}

int32_t bajo_tlas_trace_rays(bajo_ctx* ctx, const bajo_tlas* tlas, int32_t mode, const float* rays, float* hits,
                             uint64_t ray_count) {
  BJ_REQUIRE(ctx && tlas && rays && hits, "bajo_tlas_trace_rays: NULL argument");
  BJ_REQUIRE(mode == BAJO_TRACE_CLOSEST_HIT || mode == BAJO_TRACE_ANY_HIT, "bajo_tlas_trace_rays: bad mode %d", mode);
  BJ_REQUIRE(ray_count > 0 && ray_count < (1ull << 31), "bajo_tlas_trace_rays: ray count must be in [1, 2^31)");
  if (mode == BAJO_TRACE_ANY_HIT)
    return bj::launch_tlas<BAJO_TRACE_ANY_HIT, false>(ctx, tlas, rays, nullptr, 0, 0, hits, ray_count);
  return bj::launch_tlas<BAJO_TRACE_CLOSEST_HIT, false>(ctx, tlas, rays, nullptr, 0, 0, hits, ray_count);
}

int32_t bajo_tlas_trace_camera(bajo_ctx* ctx, const bajo_tlas* tlas, const float* camera, float* hits,
                               uint64_t ray_count, uint32_t width, uint32_t height) {
  BJ_REQUIRE(ctx && tlas && camera && hits, "bajo_tlas_trace_camera: NULL argument");
  BJ_REQUIRE(ray_count > 0 && ray_count < (1ull << 31) && width > 0 && height > 0,
             "bajo_tlas_trace_camera: launch dimensions must be positive");
  return bj::launch_tlas<BAJO_TRACE_CLOSEST_HIT, true>(ctx, tlas, nullptr, camera, width, height, hits, ray_count);
}

int32_t bajo_tlas_get_top(const bajo_tlas* tlas, bajo_bvh** top) {
  BJ_REQUIRE(tlas && top, "bajo_tlas_get_top: NULL argument");
  *top = tlas->top;
  return BAJO_OK;
}

int32_t bajo_tlas_copy_instance_bounds(const bajo_tlas* tlas, float* out6) {
  BJ_REQUIRE(tlas && out6, "bajo_tlas_copy_instance_bounds: NULL argument");
  for (size_t k = 0; k < tlas->host_boxes.size(); ++k) out6[k] = tlas->host_boxes[k];
  return BAJO_OK;
}

int32_t bajo_tlas_status(bajo_ctx* ctx, const bajo_tlas* tlas, uint32_t* status) {
  BJ_REQUIRE(ctx && tlas && status, "bajo_tlas_status: NULL argument");
  BJ_HIP(hipMemcpyAsync(status, tlas->status, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  return BAJO_OK;
}

int32_t bajo_tlas_free(bajo_tlas* tlas) {
  bj::free_tlas(tlas);
  return BAJO_OK;
}

}  // extern "C"
