// Device-side BVH2 traversal shared by the ray-launch kernels and the
// wavefront renderer.
//
// Follows bajo's near-first loop (bvh/gpu/trace.mojo:341-552) and its
// triangle leaf test (bvh/gpu/triangle_bvh.mojo:866-959: determinant-scaled
// Moller-Trumbore in FP32, normal = normalize(e1 x e2)).  Nodes are 64-byte
// records read with four 16-byte loads, leaf triangles 48-byte records read
// with three.  The per-lane stack lives in LDS (depth-major across the
// block, so lanes at any depth hit distinct banks); entries deeper than
// kLdsStack spill to a per-thread global column.  A Karras hierarchy over
// 64-bit (code, index) keys is at most 64 levels deep, so 64 entries always
// suffice.
#pragma once

#include "bvh_internal.h"

namespace bj {

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


struct BvhView {
  const BjNode* nodes;
  const BjTri* tris;
  uint32_t root, node_count, prim_count;
  uint32_t* status;
};

struct TraceCounters {
  uint32_t visits, inner, leaves, max_sp;
};

// Closest hit (kMode == BAJO_TRACE_CLOSEST_HIT) or occlusion (ANY_HIT: the
// returned record is bajo's Hit.shadow_hit() or Hit.miss(t_max)).
template <int kMode, bool kStats>
BJ_D Hit traverse(const BvhView& bv, const Ray& ray, Stack& st, TraceCounters* tc) {
  uint32_t n_visit = 0, n_inner = 0, n_leaf = 0, max_sp = 0;
  const BjNode* __restrict__ nodes = bv.nodes;
  const BjTri* __restrict__ tris = bv.tris;
  const uint32_t root = bv.root, node_count = bv.node_count, prim_count = bv.prim_count;
  uint32_t* status = bv.status;
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
  if constexpr (kStats) {
    tc->visits = n_visit;
    tc->inner = n_inner;
    tc->leaves = n_leaf;
    tc->max_sp = max_sp;
  }
  return hit;
}

}  // namespace bj
