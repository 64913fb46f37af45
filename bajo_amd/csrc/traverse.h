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
// A third traversal mode chosen per lane at run time (closest or any hit), so
// lanes tracing either kind of ray step through one instruction stream: the
// near-first closest-hit walk, ended at the first hit for an any-hit lane.
constexpr int kTraceEither = 2;
constexpr int kTraceShadow = 3;  // 3..6: any hit through one of the shadow walks (8-wide)
constexpr int kLdsStack = 24;
constexpr int kMaxStack = 64;
constexpr int kSpill = kMaxStack - kLdsStack;
// The shallowest LDS stack a kernel uses (the 8-waves-per-SIMD extension
// kernel, whose LDS must fit eight blocks per CU) and the spill it needs.
constexpr int kLdsStackMin = 16;
constexpr int kLdsStackMax = 32;  // the 4-waves-per-SIMD extension kernel
constexpr int kSpillMax = kMaxStack - kLdsStackMin;

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

// The closest-hit candidate held during a traversal: distance, scaled
// barycentrics, reciprocal determinant and triangle position.  The hit record
// (u = us * inv, v = vs * inv, primitive, unit normal of e1 x e2) is formed
// once at the end from the same values, so it equals the record formed at
// each accepted triangle, at a fraction of the per-triangle work.
struct Closest {
  float t, us, vs, inv;
  uint32_t s;
};

BJ_D Closest no_closest(float t) { return Closest{t, 0.0f, 0.0f, 0.0f, kEmpty}; }

BJ_D bool leaf_closest(const BjTri* __restrict__ tris, uint32_t s, const Ray& r, Closest& c) {
  TriData tr = load_tri(tris, s);
  float ad, ts, us, vs;
  if (!tri_scaled(r, tr, c.t, ad, ts, us, vs)) return false;
  const float inv = 1.0f / ad;
  c.t = ts * inv;
  c.us = us;
  c.vs = vs;
  c.inv = inv;
  c.s = s;
  return true;
}

BJ_D Hit closest_hit(const BjTri* __restrict__ tris, const Closest& c) {
  if (c.s == kEmpty) return miss(c.t);
  const TriData tr = load_tri(tris, c.s);
  return Hit{c.us * c.inv, c.vs * c.inv, tr.prim, kEmpty, normalize(cross(tr.e1, tr.e2)), c.t};
}

BJ_D bool leaf_any(const BjTri* __restrict__ tris, uint32_t s, const Ray& r) {
  TriData tr = load_tri(tris, s);
  float ad, ts, us, vs;
  return tri_scaled(r, tr, r.tmax, ad, ts, us, vs);
}

// ---- sphere leaves (bajo's GpuSphereBvh, bvh/gpu/sphere_bvh.mojo:423-478) ----
// A sphere BVH stores one 48-byte record per sorted leaf like a triangle BVH:
// center.xyz + prim id, then the radius (the rest unused).
struct SphData {
  v3 c;
  float r;
  uint32_t prim;
};

BJ_D SphData load_sphere(const BjTri* __restrict__ recs, uint32_t s) {
  const float4* p = reinterpret_cast<const float4*>(recs + s);
  const float4 a = p[0], b = p[1];
  return SphData{mk(a.x, a.y, a.z), b.x, __builtin_bit_cast(uint32_t, a.w)};
}

// intersect_ray_sphere (core/intersect.mojo:309-350): the nearer root inside
// (t_min, t_max), else the farther one.
BJ_D bool sphere_t(const Ray& r, const SphData& sp, float t_max, float& t) {
  const v3 oc = sub(r.o, sp.c);
  const float a = dot(r.d, r.d), half_b = dot(oc, r.d), c = dot(oc, oc) - sp.r * sp.r;
  const float det = half_b * half_b - a * c;
  const bool ok = det >= 0.0f;
  const float sq = sqrtf(fmax_sel(det, 0.0f));
  const float inv_a = 1.0f / a;
  const float t0 = (-half_b - sq) * inv_a, t1 = (-half_b + sq) * inv_a;
  const bool near = ok && t0 > r.tmin && t0 < t_max;
  const bool far = ok && !near && t1 > r.tmin && t1 < t_max;
  t = near ? t0 : t1;
  return near || far;
}

// Leaf tests and the final hit record for either primitive kind.
template <bool kSph>
BJ_D bool leaf_closest_p(const BjTri* __restrict__ recs, uint32_t s, const Ray& r, Closest& c) {
  if constexpr (kSph) {
    float t;
    if (!sphere_t(r, load_sphere(recs, s), c.t, t)) return false;
    c.t = t;
    c.s = s;
    return true;
  } else {
    return leaf_closest(recs, s, r, c);
  }
}

template <bool kSph>
BJ_D bool leaf_any_p(const BjTri* __restrict__ recs, uint32_t s, const Ray& r) {
  if constexpr (kSph) {
    float t;
    return sphere_t(r, load_sphere(recs, s), r.tmax, t);
  } else {
    return leaf_any(recs, s, r);
  }
}

// Sphere hit: u = v = 0, normal = normalize(o + t d - center).
template <bool kSph>
BJ_D Hit closest_hit_p(const BjTri* __restrict__ recs, const Closest& c, const Ray& r) {
  if constexpr (kSph) {
    if (c.s == kEmpty) return miss(c.t);
    const SphData sp = load_sphere(recs, c.s);
    const v3 p = add(r.o, scale(r.d, c.t));
    return Hit{0.0f, 0.0f, sp.prim, kEmpty, normalize(sub(p, sp.c)), c.t};
  } else {
    return closest_hit(recs, c);
  }
}

// A leaf record held in registers (triangle: v0 + prim, e1, e2; sphere:
// center + prim, radius), so a walk can load the next record before it
// tests the current one.
struct LeafRec {
  float4 a, b, c;
};

template <bool kSph>
BJ_D LeafRec load_rec(const BjTri* __restrict__ recs, uint32_t s) {
  const float4* p = reinterpret_cast<const float4*>(recs + s);
  LeafRec r;
  r.a = p[0];
  r.b = p[1];
  if constexpr (kSph) r.c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  else r.c = p[2];
  return r;
}

template <bool kSph>
BJ_D bool rec_closest(const LeafRec& rec, uint32_t s, const Ray& r, Closest& c) {
  if constexpr (kSph) {
    const SphData sp{mk(rec.a.x, rec.a.y, rec.a.z), rec.b.x, __builtin_bit_cast(uint32_t, rec.a.w)};
    float t;
    if (!sphere_t(r, sp, c.t, t)) return false;
    c.t = t;
    c.s = s;
    return true;
  } else {
    const TriData tr{mk(rec.a.x, rec.a.y, rec.a.z), mk(rec.b.x, rec.b.y, rec.b.z), mk(rec.c.x, rec.c.y, rec.c.z),
                     __builtin_bit_cast(uint32_t, rec.a.w)};
    float ad, ts, us, vs;
    if (!tri_scaled(r, tr, c.t, ad, ts, us, vs)) return false;
    const float inv = 1.0f / ad;
    c.t = ts * inv;
    c.us = us;
    c.vs = vs;
    c.inv = inv;
    c.s = s;
    return true;
  }
}

template <bool kSph>
BJ_D bool rec_any(const LeafRec& rec, const Ray& r) {
  if constexpr (kSph) {
    const SphData sp{mk(rec.a.x, rec.a.y, rec.a.z), rec.b.x, __builtin_bit_cast(uint32_t, rec.a.w)};
    float t;
    return sphere_t(r, sp, r.tmax, t);
  } else {
    const TriData tr{mk(rec.a.x, rec.a.y, rec.a.z), mk(rec.b.x, rec.b.y, rec.b.z), mk(rec.c.x, rec.c.y, rec.c.z),
                     __builtin_bit_cast(uint32_t, rec.a.w)};
    float ad, ts, us, vs;
    return tri_scaled(r, tr, r.tmax, ad, ts, us, vs);
  }
}

struct Stack {
  uint32_t* lds;       // this lane's column: entry k at lds[k * kTraceBlock]
  uint32_t* spill;     // this thread's column: entry k at spill[k * spill_stride]
  uint64_t spill_stride;
  int sp;
  int lds_depth = kLdsStack;    // entries held in LDS; deeper ones spill
  int lds_stride = kTraceBlock;  // words between LDS entries (a column per lane of a block, or 1)
  BJ_D void push(uint32_t v) {
    if (sp < lds_depth) lds[sp * lds_stride] = v;
    else spill[(uint64_t)(sp - lds_depth) * spill_stride] = v;
    ++sp;
  }
  BJ_D uint32_t pop() {
    --sp;
    return sp < lds_depth ? lds[sp * lds_stride] : spill[(uint64_t)(sp - lds_depth) * spill_stride];
  }
};


struct BvhView {
  const BjNode* nodes;
  const BjTri* tris;
  uint32_t root, node_count, prim_count;
  uint32_t* status;
  const uint32_t* root_dev;  // when set, the root is read here (H-PLOC leaves it on the device)
};

struct TraceCounters {
  uint32_t visits, inner, leaves, max_sp;
};

// Closest hit (kMode == BAJO_TRACE_CLOSEST_HIT) or occlusion (ANY_HIT: the
// returned record is bajo's Hit.shadow_hit() or Hit.miss(t_max)).
template <int kMode, bool kStats, bool kSph = false>
BJ_D Hit traverse(const BvhView& bv, const Ray& ray, Stack& st, TraceCounters* tc, bool any_rt = false) {
  uint32_t n_visit = 0, n_inner = 0, n_leaf = 0, max_sp = 0;
  const BjNode* __restrict__ nodes = bv.nodes;
  const BjTri* __restrict__ tris = bv.tris;
  const uint32_t root = bv.root_dev ? *bv.root_dev : bv.root;
  const uint32_t node_count = bv.node_count, prim_count = bv.prim_count;
  uint32_t* status = bv.status;
  Closest best = no_closest(ray.tmax);
  const v3 rcp = rcp_dir(ray.d);
  st.sp = 0;
  bool occluded = false;

  if ((root & kLeafFlag) && root != kSentinel) {
    n_leaf++;
    if constexpr (kMode == BAJO_TRACE_ANY_HIT) occluded = leaf_any_p<kSph>(tris, root & kIndexMask, ray);
    else if (leaf_closest_p<kSph>(tris, root & kIndexMask, ray, best) && kMode == kTraceEither && any_rt) occluded = true;
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
      const float node_tmax = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : best.t;
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
            if (leaf_any_p<kSph>(tris, c0 & kIndexMask, ray)) { occluded = true; break; }
          } else if (!(c0 & kLeafFlag)) {
            if (st.sp >= kMaxStack) { atomicOr(status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW); break; }
            st.push(c0);
          }
        }
        if (h1) {
          if (l1) {
            n_leaf++;
            if (leaf_any_p<kSph>(tris, c1 & kIndexMask, ray)) { occluded = true; break; }
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
          if (l0) {
            n_leaf++;
            if (leaf_closest_p<kSph>(tris, c0 & kIndexMask, ray, best) && kMode == kTraceEither && any_rt) {
              occluded = true;
              break;
            }
          } else if (!(c0 & kLeafFlag)) { v0 = true; n_inner++; }
        }
        if (h1) {
          if (l1) {
            n_leaf++;
            if (leaf_closest_p<kSph>(tris, c1 & kIndexMask, ray, best) && kMode == kTraceEither && any_rt) {
              occluded = true;
              break;
            }
          } else if (!(c1 & kLeafFlag)) { v1 = true; n_inner++; }
        }
        // bvh/gpu/trace.mojo:395-426: nearest child first, push the other.
        int near = -1;
        float near_t = kFltMax;
        if (v0) { near = 0; near_t = t0; }
        if (v1 && (near == -1 || t1 < near_t)) { near = 1; near_t = t1; }
        bool overflow = false;
        if (v0 && near != 0 && t0 <= best.t) {
          if (st.sp >= kMaxStack) overflow = true; else st.push(c0);
        }
        if (v1 && near != 1 && t1 <= best.t) {
          if (st.sp >= kMaxStack) overflow = true; else st.push(c1);
        }
        if (overflow) { atomicOr(status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW); break; }
        if (kStats && (uint32_t)st.sp > max_sp) max_sp = (uint32_t)st.sp;
        if (near != -1 && near_t <= best.t) { cur = near == 0 ? c0 : c1; continue; }
        if (st.sp == 0) break;
        cur = st.pop();
      }
    }
  }
  Hit hit;
  if (kMode == BAJO_TRACE_ANY_HIT || (kMode == kTraceEither && any_rt)) {
    // bvh/gpu/trace.mojo:44-59: occlusion is reported as Hit.shadow_hit().
    hit = occluded ? Hit{0.0f, 0.0f, kEmpty, kEmpty, mk(0, 0, 0), 0.0f} : miss(ray.tmax);
  } else {
    hit = closest_hit_p<kSph>(tris, best, ray);
  }
  if constexpr (kStats) {
    tc->visits = n_visit;
    tc->inner = n_inner;
    tc->leaves = n_leaf;
    tc->max_sp = max_sp;
  }
  return hit;
}

// ---- 8-wide traversal -------------------------------------------------------------
// Plays the role of bajo's compressed-BVH8 traversal
// (bvh/gpu/triangle_bvh.mojo:1010-1104, bvh/gpu/compressed_bounds_bvh.mojo:
// 469-576) over the 96-byte quantized nodes built by the collapse.  Per node:
// six 16-byte loads, the ray octant picks the near and far plane bytes, and
// t(q) = (o + q*2^e - ray.o) * rcp is evaluated as fmaf(q, 2^e*rcp,
// (o - ray.o)*rcp).  Leaf children hit by the ray are tested after the box
// loop in one flattened triangle loop (so lanes diverge by the number of
// triangles, not by child slot); inner children continue nearest first with
// the others pushed far-to-near.

struct WideView {
  const BjWide* nodes;
  const BjTri* tris;
  uint32_t node_count, prim_count;
  uint32_t* status;
  uint32_t root;
};

BJ_D float ubyte_f(uint32_t word, int k) { return (float)((word >> (8 * k)) & 0xFFu); }

// Child reference of a one-hot slot mask, as a chain of selects on plain
// register values (an index-based select over an array or struct would be
// folded into a dynamically indexed stack copy, i.e. scratch memory, when the
// index differs between lanes).
BJ_D uint32_t sel8(uint32_t bit, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t c4, uint32_t c5,
                   uint32_t c6, uint32_t c7) {
  uint32_t r = c0;
  r = (bit & 0x02u) ? c1 : r;
  r = (bit & 0x04u) ? c2 : r;
  r = (bit & 0x08u) ? c3 : r;
  r = (bit & 0x10u) ? c4 : r;
  r = (bit & 0x20u) ? c5 : r;
  r = (bit & 0x40u) ? c6 : r;
  r = (bit & 0x80u) ? c7 : r;
  return r;
}

// The 8-wide traversal as a resumable walk: begin() takes a ray, each step()
// visits one node (its eight slab tests, the triangles of the leaf children
// it enters, the pushes of the inner ones) and moves to the next node;
// done() once the stack is empty, the ray is occluded (any hit) or an error
// ended it.  Launches that refill finished lanes with new rays call the steps
// themselves; traverse8 runs one ray to the end.
template <int kMode, bool kStats, bool kSph = false>
struct Walk8 {
  Ray ray;
  v3 rcp;
  bool px, py, pz, occluded, any_rt;
  Closest best;
  uint32_t cur, n_visit, n_inner, n_leaf, max_sp;

  BJ_D void begin(const WideView& wv, const Ray& r, Stack& st, bool any = false) {
    ray = r;
    any_rt = any;
    best = no_closest(r.tmax);
    rcp = rcp_dir(r.d);
    px = rcp.x >= 0.0f;
    py = rcp.y >= 0.0f;
    pz = rcp.z >= 0.0f;
    occluded = false;
    st.sp = 0;
    cur = wv.node_count == 0 ? kSentinel : wv.root;
    n_visit = n_inner = n_leaf = max_sp = 0;
  }

  BJ_D bool done() const { return cur == kSentinel; }

  BJ_D void step(const WideView& wv, Stack& st) {
    if (cur >= wv.node_count || n_visit > wv.node_count) {
      atomicOr(wv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
      cur = kSentinel;
      return;
    }
    n_visit++;
    const uint4* np = reinterpret_cast<const uint4*>(wv.nodes + cur);
    const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4], w5 = np[5];
    const uint32_t meta = w0.w;
    const int n = (int)(meta >> 24);
    const float sx = __builtin_bit_cast(float, (meta & 0xFFu) << 23);
    const float sy = __builtin_bit_cast(float, ((meta >> 8) & 0xFFu) << 23);
    const float sz = __builtin_bit_cast(float, ((meta >> 16) & 0xFFu) << 23);
    const float ax = (__builtin_bit_cast(float, w0.x) - ray.o.x) * rcp.x;
    const float ay = (__builtin_bit_cast(float, w0.y) - ray.o.y) * rcp.y;
    const float az = (__builtin_bit_cast(float, w0.z) - ray.o.z) * rcp.z;
    const float bx = sx * rcp.x, by = sy * rcp.y, bz = sz * rcp.z;
    // near / far plane bytes per axis for this ray octant (words: lo_x w1.xy,
    // lo_y w1.zw, lo_z w2.xy, hi_x w2.zw, hi_y w3.xy, hi_z w3.zw)
    const uint32_t nx0 = px ? w1.x : w2.z, nx1 = px ? w1.y : w2.w, fx0 = px ? w2.z : w1.x, fx1 = px ? w2.w : w1.y;
    const uint32_t ny0 = py ? w1.z : w3.x, ny1 = py ? w1.w : w3.y, fy0 = py ? w3.x : w1.z, fy1 = py ? w3.y : w1.w;
    const uint32_t nz0 = pz ? w2.x : w3.z, nz1 = pz ? w2.y : w3.w, fz0 = pz ? w3.z : w2.x, fz1 = pz ? w3.w : w2.y;
    uint32_t c0 = w4.x, c1 = w4.y, c2 = w4.z, c3 = w4.w, c4 = w5.x, c5 = w5.y, c6 = w5.z, c7 = w5.w;
    asm("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3), "+v"(c4), "+v"(c5), "+v"(c6), "+v"(c7));
    const float node_tmax = kMode == BAJO_TRACE_ANY_HIT ? ray.tmax : best.t;
    float th[8];
    // slot hits as a bit mask, built without branches (per-slot branches
    // cost more than the slab test itself)
    uint32_t hit_mask = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int bt = i & 3;
      const uint32_t qnx = i < 4 ? nx0 : nx1, qfx = i < 4 ? fx0 : fx1;
      const uint32_t qny = i < 4 ? ny0 : ny1, qfy = i < 4 ? fy0 : fy1;
      const uint32_t qnz = i < 4 ? nz0 : nz1, qfz = i < 4 ? fz0 : fz1;
      const float tnx = fmaf(ubyte_f(qnx, bt), bx, ax), tfx = fmaf(ubyte_f(qfx, bt), bx, ax);
      const float tny = fmaf(ubyte_f(qny, bt), by, ay), tfy = fmaf(ubyte_f(qfy, bt), by, ay);
      const float tnz = fmaf(ubyte_f(qnz, bt), bz, az), tfz = fmaf(ubyte_f(qfz, bt), bz, az);
      const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
      const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, node_tmax));
      th[i] = tmin;
      hit_mask |= tmin <= tmax ? (1u << i) : 0u;
    }
    hit_mask &= n >= 8 ? 0xFFu : (1u << n) - 1u;  // slots past the child count are never entered
    const uint32_t leaf_bits = (c0 >> 31) | ((c1 >> 30) & 2u) | ((c2 >> 29) & 4u) | ((c3 >> 28) & 8u) |
                               ((c4 >> 27) & 16u) | ((c5 >> 26) & 32u) | ((c6 >> 25) & 64u) | ((c7 >> 24) & 128u);
    uint32_t leaf_mask = hit_mask & leaf_bits, inner_mask = hit_mask & ~leaf_bits;
    // leaf children: one record per iteration across all hit leaves, in
    // slot order; the next record's load is issued before the current one
    // is tested, so a leaf's records arrive back to back
    uint32_t tri = 0, tri_end = 0;
    auto advance = [&](uint32_t& t, uint32_t& e) -> bool {
      while (t >= e) {
        if (!leaf_mask) return false;
        const uint32_t bit = leaf_mask & (0u - leaf_mask);
        leaf_mask ^= bit;
        const uint32_t ref = sel8(bit, c0, c1, c2, c3, c4, c5, c6, c7);
        t = ref & kWideLeafStartMask;
        e = t + ((ref >> 28) & 7u) + 1u;
        if (e > wv.prim_count) e = wv.prim_count;
      }
      return true;
    };
    bool have = advance(tri, tri_end);
    LeafRec cur_rec;
    if (have) cur_rec = load_rec<kSph>(wv.tris, tri);
    while (have) {
      uint32_t nt = tri + 1, ne = tri_end;
      const bool nhave = advance(nt, ne);
      LeafRec next_rec;
      if (nhave) next_rec = load_rec<kSph>(wv.tris, nt);
      n_leaf++;
      if constexpr (kMode == BAJO_TRACE_ANY_HIT) {
        if (rec_any<kSph>(cur_rec, ray)) {
          occluded = true;
          break;
        }
      } else if constexpr (kMode == kTraceEither) {
        // best.t stays at t_max on an any-hit lane, so the closest-hit test is its occlusion test
        if (rec_closest<kSph>(cur_rec, tri, ray, best) && any_rt) {
          occluded = true;
          break;
        }
      } else {
        rec_closest<kSph>(cur_rec, tri, ray, best);
      }
      tri = nt;
      tri_end = ne;
      have = nhave;
      cur_rec = next_rec;
    }
    if constexpr (kMode == BAJO_TRACE_ANY_HIT || kMode == kTraceEither) {
      if (occluded) {
        cur = kSentinel;
        return;
      }
    }
    if constexpr (kMode != BAJO_TRACE_ANY_HIT) {
      // children entered beyond the (possibly improved) closest hit are dropped
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (th[i] > best.t) inner_mask &= ~(1u << i);
    }
    if (kStats) n_inner += __builtin_popcount(inner_mask);
    if (inner_mask) {
      // push all but the nearest inner child, farthest first
      bool overflow = false;
      while (inner_mask & (inner_mask - 1u)) {
        uint32_t far_bit = 0, far_ref = 0;
        float far_t = -1.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i)
          if (((inner_mask >> i) & 1u) && th[i] >= far_t) {
            far_t = th[i];
            far_bit = 1u << i;
            far_ref = sel8(1u << i, c0, c1, c2, c3, c4, c5, c6, c7);
          }
        inner_mask ^= far_bit;
        if (st.sp >= kMaxStack) {
          overflow = true;
          break;
        }
        st.push(far_ref);
      }
      if (overflow) {
        atomicOr(wv.status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW);
        cur = kSentinel;
        return;
      }
      if (kStats && (uint32_t)st.sp > max_sp) max_sp = (uint32_t)st.sp;
      cur = sel8(inner_mask, c0, c1, c2, c3, c4, c5, c6, c7);
      return;
    }
    cur = st.sp == 0 ? kSentinel : st.pop();

}

  BJ_D Hit finish(const WideView& wv, TraceCounters* tc) const {
    Hit hit;
    if (kMode == BAJO_TRACE_ANY_HIT || (kMode == kTraceEither && any_rt)) {
      hit = occluded ? Hit{0.0f, 0.0f, kEmpty, kEmpty, mk(0, 0, 0), 0.0f} : miss(ray.tmax);
    } else {
      hit = closest_hit_p<kSph>(wv.tris, best, ray);
    }
    if constexpr (kStats) {
      tc->visits = n_visit;
      tc->inner = n_inner;
      tc->leaves = n_leaf;
      tc->max_sp = max_sp;
    }
    return hit;
  }
};

template <int kMode, bool kStats, bool kSph = false>
BJ_D Hit traverse8(const WideView& wv, const Ray& ray, Stack& st, TraceCounters* tc, bool any_rt = false) {
  Walk8<kMode, kStats, kSph> w;
  w.begin(wv, ray, st, any_rt);
  while (!w.done()) w.step(wv, st);
  return w.finish(wv, tc);
}

// Child slab tests of one 8-wide node: which slots the ray enters, which of
// those are leaves, their entry distances and references.
struct WideHits {
  uint32_t c[8];
  float th[8];
  uint32_t leaf, inner;
};

BJ_D void wide_node_hits(const uint4 (&w)[6], const Ray& ray, const v3& rcp, bool px, bool py, bool pz,
                         float node_tmax, WideHits& h) {
  const uint32_t meta = w[0].w;
  const int n = (int)(meta >> 24);
  const float sx = __builtin_bit_cast(float, (meta & 0xFFu) << 23);
  const float sy = __builtin_bit_cast(float, ((meta >> 8) & 0xFFu) << 23);
  const float sz = __builtin_bit_cast(float, ((meta >> 16) & 0xFFu) << 23);
  const float ax = (__builtin_bit_cast(float, w[0].x) - ray.o.x) * rcp.x;
  const float ay = (__builtin_bit_cast(float, w[0].y) - ray.o.y) * rcp.y;
  const float az = (__builtin_bit_cast(float, w[0].z) - ray.o.z) * rcp.z;
  const float bx = sx * rcp.x, by = sy * rcp.y, bz = sz * rcp.z;
  const uint32_t nx0 = px ? w[1].x : w[2].z, nx1 = px ? w[1].y : w[2].w, fx0 = px ? w[2].z : w[1].x,
                 fx1 = px ? w[2].w : w[1].y;
  const uint32_t ny0 = py ? w[1].z : w[3].x, ny1 = py ? w[1].w : w[3].y, fy0 = py ? w[3].x : w[1].z,
                 fy1 = py ? w[3].y : w[1].w;
  const uint32_t nz0 = pz ? w[2].x : w[3].z, nz1 = pz ? w[2].y : w[3].w, fz0 = pz ? w[3].z : w[2].x,
                 fz1 = pz ? w[3].w : w[2].y;
  h.c[0] = w[4].x; h.c[1] = w[4].y; h.c[2] = w[4].z; h.c[3] = w[4].w;
  h.c[4] = w[5].x; h.c[5] = w[5].y; h.c[6] = w[5].z; h.c[7] = w[5].w;
  uint32_t hit_mask = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int bt = i & 3;
    const uint32_t qnx = i < 4 ? nx0 : nx1, qfx = i < 4 ? fx0 : fx1;
    const uint32_t qny = i < 4 ? ny0 : ny1, qfy = i < 4 ? fy0 : fy1;
    const uint32_t qnz = i < 4 ? nz0 : nz1, qfz = i < 4 ? fz0 : fz1;
    const float tnx = fmaf(ubyte_f(qnx, bt), bx, ax), tfx = fmaf(ubyte_f(qfx, bt), bx, ax);
    const float tny = fmaf(ubyte_f(qny, bt), by, ay), tfy = fmaf(ubyte_f(qfy, bt), by, ay);
    const float tnz = fmaf(ubyte_f(qnz, bt), bz, az), tfz = fmaf(ubyte_f(qfz, bt), bz, az);
    const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
    const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, node_tmax));
    h.th[i] = tmin;
    hit_mask |= tmin <= tmax ? (1u << i) : 0u;
  }
  hit_mask &= n >= 8 ? 0xFFu : (1u << n) - 1u;
  uint32_t leaf_bits = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) leaf_bits |= (h.c[i] >> 31) << i;
  h.leaf = hit_mask & leaf_bits;
  h.inner = hit_mask & ~leaf_bits;
}

// Moves (t, e) to the next leaf record of the leaf slots left in lm_a, then
// lm_b (slot order); false once both are empty.
BJ_D bool next_leaf_range2(uint32_t& lm_a, uint32_t& lm_b, const WideHits& ha, const WideHits& hb,
                           uint32_t prim_count, uint32_t& t, uint32_t& e) {
  while (t >= e) {
    uint32_t ref;
    if (lm_a) {
      const uint32_t bit = lm_a & (0u - lm_a);
      lm_a ^= bit;
      ref = sel8(bit, ha.c[0], ha.c[1], ha.c[2], ha.c[3], ha.c[4], ha.c[5], ha.c[6], ha.c[7]);
    } else if (lm_b) {
      const uint32_t bit = lm_b & (0u - lm_b);
      lm_b ^= bit;
      ref = sel8(bit, hb.c[0], hb.c[1], hb.c[2], hb.c[3], hb.c[4], hb.c[5], hb.c[6], hb.c[7]);
    } else {
      return false;
    }
    t = ref & kWideLeafStartMask;
    e = t + ((ref >> 28) & 7u) + 1u;
    if (e > prim_count) e = prim_count;
  }
  return true;
}

// Any hit over the 8-wide layout for shadow rays, whose walks are long serial
// chains (they leave a surface and pass many boxes and leaves near it), in
// four forms picked per launch:
//   kTwo: two nodes per step -- the node on top of the stack is loaded and
//     tested together with the current one, so a step waits for one round of
//     node loads for both;
//   kBatch = 4: the leaf records of the entered leaves are loaded four at a
//     time (a 4-triangle leaf costs one load round instead of four);
//     kBatch = 1 keeps one record in flight ahead of the one being tested.
// Whether a ray is occluded does not depend on the order its nodes and
// records are visited in, so every form returns traverse8's result (an
// occluded ray returns the same constant record); only the visit counters
// can differ.
template <bool kStats, bool kSph, bool kTwo, int kBatch>
BJ_D Hit traverse8_any(const WideView& wv, const Ray& ray, Stack& st, TraceCounters* tc) {
  const v3 rcp = rcp_dir(ray.d);
  const bool px = rcp.x >= 0.0f, py = rcp.y >= 0.0f, pz = rcp.z >= 0.0f;
  uint32_t n_visit = 0, n_inner = 0, n_leaf = 0, max_sp = 0;
  bool occluded = false;
  st.sp = 0;
  uint32_t a = wv.node_count == 0 ? kSentinel : wv.root;
  while (a != kSentinel) {
    const uint32_t b = (kTwo && st.sp > 0) ? st.pop() : kSentinel;
    if (a >= wv.node_count || (b != kSentinel && b >= wv.node_count) || n_visit > wv.node_count) {
      atomicOr(wv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
      break;
    }
    uint4 wa[6], wb[6];
    const uint4* pa = reinterpret_cast<const uint4*>(wv.nodes + a);
#pragma unroll
    for (int k = 0; k < 6; ++k) wa[k] = pa[k];
    if (kTwo && b != kSentinel) {
      const uint4* pb = reinterpret_cast<const uint4*>(wv.nodes + b);
#pragma unroll
      for (int k = 0; k < 6; ++k) wb[k] = pb[k];
    }
    WideHits ha, hb;
    // both nodes' loads are in flight; test a, then b (kept apart so the
    // scheduler does not interleave the two sets of slab temporaries)
    __builtin_amdgcn_sched_barrier(0);
    wide_node_hits(wa, ray, rcp, px, py, pz, ray.tmax, ha);
    n_visit++;
    __builtin_amdgcn_sched_barrier(0);
    if (kTwo && b != kSentinel) {
      wide_node_hits(wb, ray, rcp, px, py, pz, ray.tmax, hb);
      n_visit++;
    } else {
      hb.leaf = hb.inner = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        hb.c[i] = 0;
        hb.th[i] = 0.0f;
      }
    }
    // the leaf records of both nodes (a's leaves first, slot order)
    uint32_t lm_a = ha.leaf, lm_b = hb.leaf;
    uint32_t tri = 0, tri_end = 0;
    if constexpr (kBatch == 4) {
      for (;;) {
        uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        const bool h0 = next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, tri, tri_end);
        if (h0) r0 = tri++;
        const bool h1 = h0 && next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, tri, tri_end);
        if (h1) r1 = tri++;
        const bool h2 = h1 && next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, tri, tri_end);
        if (h2) r2 = tri++;
        const bool h3 = h2 && next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, tri, tri_end);
        if (h3) r3 = tri++;
        if (!h0) break;
        LeafRec q0 = load_rec<kSph>(wv.tris, r0), q1, q2, q3;
        if (h1) q1 = load_rec<kSph>(wv.tris, r1);
        if (h2) q2 = load_rec<kSph>(wv.tris, r2);
        if (h3) q3 = load_rec<kSph>(wv.tris, r3);
        n_leaf += 1u + (h1 ? 1u : 0u) + (h2 ? 1u : 0u) + (h3 ? 1u : 0u);
        if (rec_any<kSph>(q0, ray) || (h1 && rec_any<kSph>(q1, ray)) || (h2 && rec_any<kSph>(q2, ray)) ||
            (h3 && rec_any<kSph>(q3, ray))) {
          occluded = true;
          break;
        }
        if (!h3) break;
      }
    } else {
      bool have = next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, tri, tri_end);
      LeafRec cur_rec;
      if (have) cur_rec = load_rec<kSph>(wv.tris, tri);
      while (have) {
        uint32_t nt = tri + 1, ne = tri_end;
        const bool nhave = next_leaf_range2(lm_a, lm_b, ha, hb, wv.prim_count, nt, ne);
        LeafRec next_rec;
        if (nhave) next_rec = load_rec<kSph>(wv.tris, nt);
        n_leaf++;
        if (rec_any<kSph>(cur_rec, ray)) {
          occluded = true;
          break;
        }
        tri = nt;
        tri_end = ne;
        have = nhave;
        cur_rec = next_rec;
      }
    }
    if (occluded) break;
    uint32_t ia = ha.inner, ib = hb.inner;
    if (kStats) n_inner += __builtin_popcount(ia) + __builtin_popcount(ib);
    // push b's inner children, then a's farthest first; continue at a's
    // nearest (or the top of the stack)
    bool overflow = false;
    while (ib) {
      const uint32_t bit = ib & (0u - ib);
      ib ^= bit;
      if (st.sp >= kMaxStack) { overflow = true; break; }
      st.push(sel8(bit, hb.c[0], hb.c[1], hb.c[2], hb.c[3], hb.c[4], hb.c[5], hb.c[6], hb.c[7]));
    }
    while (!overflow && (ia & (ia - 1u))) {
      uint32_t far_bit = 0;
      float far_t = -1.0f;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (((ia >> i) & 1u) && ha.th[i] >= far_t) {
          far_t = ha.th[i];
          far_bit = 1u << i;
        }
      ia ^= far_bit;
      if (st.sp >= kMaxStack) { overflow = true; break; }
      st.push(sel8(far_bit, ha.c[0], ha.c[1], ha.c[2], ha.c[3], ha.c[4], ha.c[5], ha.c[6], ha.c[7]));
    }
    if (overflow) {
      atomicOr(wv.status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW);
      break;
    }
    if (kStats && (uint32_t)st.sp > max_sp) max_sp = (uint32_t)st.sp;
    if (ia) a = sel8(ia, ha.c[0], ha.c[1], ha.c[2], ha.c[3], ha.c[4], ha.c[5], ha.c[6], ha.c[7]);
    else a = st.sp == 0 ? kSentinel : st.pop();
  }
  if constexpr (kStats) {
    tc->visits = n_visit;
    tc->inner = n_inner;
    tc->leaves = n_leaf;
    tc->max_sp = max_sp;
  }
  return occluded ? Hit{0.0f, 0.0f, kEmpty, kEmpty, mk(0, 0, 0), 0.0f} : miss(ray.tmax);
}

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

// Either layout behind one call: kWide selects the 8-wide traversal.
struct AccelView {
  BvhView bin;
  WideView wide;
};

// Host: the view of a built hierarchy; `status` receives traversal errors.
inline AccelView accel_view(const bajo_bvh* h, uint32_t* status) {
  AccelView av;
  av.bin = BvhView{h->nodes, h->tris, h->root, h->trav_nodes, h->trav_prims, status,
                   h->root_on_device ? h->hp_ctl + kHpCtlRoot : nullptr};
  // a collapse whose node count has not been read back is bounded by its capacity
  const uint32_t wide_nodes = h->wide_pending ? (uint32_t)h->wide_capacity : h->wide_count;
  av.wide = WideView{h->wide, h->tris, wide_layout(h->node_width) ? wide_nodes : 0u, h->trav_prims, status,
                     h->wide_root};
  return av;
}

// Orders one wave's LDS writes before its later LDS reads by other lanes.
BJ_D void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Any hit with G lanes per ray (G = 8, 16, 32 or 64): the walk of rays whose
// chains are long and serial when one lane carries a ray, such as shadow
// rays (they leave a surface and pass many boxes and leaves near it).  A
// step takes up to G/8 nodes off the group's LDS stack, lane 8j + i tests
// child i of node j (the same quantized slab test as Walk8), entered inner
// children are pushed, and the records of all entered leaves are tested one
// per lane.  A step therefore waits for one round of node loads and one of
// record loads whatever the width of the ray's frontier.  The records tested
// are every record whose leaf box the ray enters, as in every other any-hit
// walk, so occlusion comes out the same.
//
// GroupWalk holds one group's state; every lane of the group holds the same
// copy.  step() is called by the group's lanes together (other groups of the
// wave may be stepping their own rays at the same time).
template <int G>
struct GroupWalk {
  static constexpr int kStack = 16 * G;  // a level pushes at most 7G/8 net; 18 levels fit
  static constexpr int kRecs = 8 * G;    // G entered leaves of up to 8 records
  Ray ray;
  v3 rcp;
  bool px, py, pz;
  int sp;
  uint32_t steps;
  // The frontier of the next step, taken off the stack and loaded while the
  // previous step's records were tested: how many nodes it holds, whether
  // this lane's node is one of them, its first four 16-byte words and this
  // lane's child reference.  Every step issues these loads at one place (a
  // dummy node 0 where no node is taken), so they stay in flight across the
  // record tests.
  int nf;
  bool mine;
  uint4 w0, w1, w2, w3;
  uint32_t ref;

  BJ_D uint64_t group_mask() const {
    const uint32_t lane = lane_id();
    return (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << ((lane / G) * G);
  }

  // Starts a ray with the root as its first frontier (loaded here); a root
  // outside the hierarchy ends the ray at once with a status word.
  BJ_D void begin(const WideView& wv, const Ray& r, uint32_t* stack) {
    ray = r;
    rcp = rcp_dir(r.d);
    px = rcp.x >= 0.0f;
    py = rcp.y >= 0.0f;
    pz = rcp.z >= 0.0f;
    steps = 1;
    sp = 0;
    const bool ok = wv.node_count != 0 && wv.root < wv.node_count;
    if (wv.node_count != 0 && !ok && lane_id() % G == 0) atomicOr(wv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
    nf = ok ? 1 : 0;
    mine = ok && (lane_id() % G) < 8;
    if (wv.node_count != 0) load(wv, ok ? wv.root : 0u);
  }

  BJ_D void load(const WideView& wv, uint32_t node) {
    const uint4* np = reinterpret_cast<const uint4*>(wv.nodes + node);
    w0 = np[0];
    w1 = np[1];
    w2 = np[2];
    w3 = np[3];
    ref = reinterpret_cast<const uint32_t*>(wv.nodes + node)[16 + (lane_id() & 7u)];
  }

  // One step; returns true once the ray is finished (then *occluded says how).
  BJ_D bool step(const WideView& wv, uint32_t* stack, uint32_t* recs, uint32_t& visits, uint32_t& tests,
                 bool& occluded) {
    occluded = false;
    if (nf == 0) return true;
    visits += (uint32_t)nf;
    const uint32_t gl = lane_id() % G, j = gl >> 3, i = gl & 7;
    const uint64_t gm = group_mask();
    bool hit = false;
    const uint32_t cref = ref;
    if (mine) {
      const uint32_t meta = w0.w;
      const uint32_t n = meta >> 24;
      const float sx = __builtin_bit_cast(float, (meta & 0xFFu) << 23);
      const float sy = __builtin_bit_cast(float, ((meta >> 8) & 0xFFu) << 23);
      const float sz = __builtin_bit_cast(float, ((meta >> 16) & 0xFFu) << 23);
      const float ax = (__builtin_bit_cast(float, w0.x) - ray.o.x) * rcp.x;
      const float ay = (__builtin_bit_cast(float, w0.y) - ray.o.y) * rcp.y;
      const float az = (__builtin_bit_cast(float, w0.z) - ray.o.z) * rcp.z;
      const float bx = sx * rcp.x, by = sy * rcp.y, bz = sz * rcp.z;
      // byte i of lo_x (w1.xy), lo_y (w1.zw), lo_z (w2.xy), hi_x (w2.zw), hi_y (w3.xy), hi_z (w3.zw)
      const bool hiw = i >= 4;
      const int bt = (int)(i & 3u);
      const uint32_t lox = hiw ? w1.y : w1.x, loy = hiw ? w1.w : w1.z, loz = hiw ? w2.y : w2.x;
      const uint32_t hix = hiw ? w2.w : w2.z, hiy = hiw ? w3.y : w3.x, hiz = hiw ? w3.w : w3.z;
      const uint32_t qnx = px ? lox : hix, qfx = px ? hix : lox;
      const uint32_t qny = py ? loy : hiy, qfy = py ? hiy : loy;
      const uint32_t qnz = pz ? loz : hiz, qfz = pz ? hiz : loz;
      const float tnx = fmaf(ubyte_f(qnx, bt), bx, ax), tfx = fmaf(ubyte_f(qfx, bt), bx, ax);
      const float tny = fmaf(ubyte_f(qny, bt), by, ay), tfy = fmaf(ubyte_f(qfy, bt), by, ay);
      const float tnz = fmaf(ubyte_f(qnz, bt), bz, az), tfz = fmaf(ubyte_f(qfz, bt), bz, az);
      const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
      const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, ray.tmax));
      hit = i < n && tmin <= tmax;
    }
    const bool is_leaf = hit && (cref & kLeafFlag) != 0;
    const bool is_inner = hit && (cref & kLeafFlag) == 0;
    // entered inner children onto the group's stack
    const uint64_t inner = __ballot(is_inner) & gm;
    const int n_inner = __builtin_popcountll(inner);
    if (sp + n_inner > kStack) {
      if (gl == 0) atomicOr(wv.status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW);
      sp = 0;
      nf = 0;
      return true;
    }
    if (is_inner) stack[sp + (int)mask_rank(inner)] = cref;
    sp += n_inner;
    // the entered leaves' records, one per lane
    uint32_t start = 0, k = 0;
    if (is_leaf) {
      start = cref & kWideLeafStartMask;
      const uint32_t e = start + ((cref >> 28) & 7u) + 1u;
      k = (e > wv.prim_count ? wv.prim_count : e) - (start < wv.prim_count ? start : wv.prim_count);
    }
    uint32_t incl = k;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, G);
      if (gl >= (uint32_t)off) incl += y;
    }
    const uint32_t total = __shfl(incl, G - 1, G);
    for (uint32_t m = 0; m < k; ++m) recs[incl - k + m] = start + m;
    wave_lds_sync();
    tests += total;
    // the first round of records is loaded, then the next frontier (up to G/8
    // nodes off the stack), so its node loads are in flight while the
    // records are tested
    LeafRec rec;
    const bool has = gl < total;
    if (has) rec = load_rec<false>(wv.tris, recs[gl]);
    nf = sp < G / 8 ? sp : G / 8;
    mine = (int)j < nf;
    // lanes without a node of their own load their group's first node (the
    // same cache lines as its owners)
    const int at = sp - 1 - (mine ? (int)j : 0);
    const uint32_t top = stack[at > 0 ? at : 0];
    wave_lds_sync();
    sp -= nf;
    const bool bad = (__ballot(mine && top >= wv.node_count) & gm) != 0 || (nf > 0 && ++steps > wv.node_count);
    load(wv, mine && top < wv.node_count ? top : 0u);
    if ((__ballot(has && rec_any<false>(rec, ray)) & gm) != 0) {
      occluded = true;
      nf = 0;
      return true;
    }
    for (uint32_t base = G; base < total; base += G) {
      bool occ = false;
      if (base + gl < total) {
        const LeafRec r2 = load_rec<false>(wv.tris, recs[base + gl]);
        occ = rec_any<false>(r2, ray);
      }
      if ((__ballot(occ) & gm) != 0) {
        occluded = true;
        nf = 0;
        return true;
      }
    }
    wave_lds_sync();
    if (bad) {
      if (gl == 0) atomicOr(wv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
      nf = 0;
      return true;
    }
    return nf == 0;
  }
};

// Closest hit with G lanes per ray (the tail kernel's extension rays), equal
// to the one-lane Walk8 result.  Walk8 accepts a triangle when its scaled
// distance beats the best so far (ts < t_best * det), so which of two
// triangles wins can depend on their test order when their distances lie
// within rounding of each other; the group visits nodes in another order.
// So the group keeps the two smallest candidate distances it has seen (t1
// with its record, t2; candidates pass Walk8's own test against the ray's
// t_max), culls only boxes entered beyond t1 * (1 + margin),
// and reports a near tie when t2 <= t1 * (1 + margin); the caller then runs
// Walk8 for that ray.  Without a near tie every pairwise comparison Walk8
// makes has the exact outcome (the margin, 4e-6, is ten times the rounding of
// the scaled test), so Walk8 ends at the unique nearest candidate, which is
// t1's record; its hit record is formed by testing it again (the same
// arithmetic that formed it in the walk).
template <int G>
struct GroupClosest {
  static constexpr int kStack = 16 * G;  // entries of (node, entry distance); 18 levels fit, as GroupWalk
  static constexpr int kRecs = 8 * G;
  static constexpr float kMargin = 4e-6f;
  Ray ray;
  v3 rcp;
  bool px, py, pz;
  int sp;
  uint32_t steps;
  float t1, t2;
  uint32_t best;  // record of t1 (kEmpty: no candidate)

  BJ_D uint64_t group_mask() const {
    const uint32_t lane = lane_id();
    return (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << ((lane / G) * G);
  }
  BJ_D float cull() const { return fminf(ray.tmax, t1 * (1.0f + kMargin)); }

  BJ_D void begin(const WideView& wv, const Ray& r, uint32_t* stack) {
    ray = r;
    rcp = rcp_dir(r.d);
    px = rcp.x >= 0.0f;
    py = rcp.y >= 0.0f;
    pz = rcp.z >= 0.0f;
    steps = 0;
    t1 = t2 = kFltMax;
    best = kEmpty;
    sp = wv.node_count == 0 ? 0 : 1;
    if (lane_id() % G == 0) {
      stack[0] = wv.root;
      stack[1] = __builtin_bit_cast(uint32_t, 0.0f);
    }
    wave_lds_sync();
  }

  // One step; true once the walk is finished (or ended by an error).
  BJ_D bool step(const WideView& wv, uint32_t* stack, uint32_t* recs, uint32_t& visits, uint32_t& tests) {
    if (sp == 0) return true;
    const uint32_t gl = lane_id() % G, j = gl >> 3, i = gl & 7;
    const uint64_t gm = group_mask();
    const int nf = sp < G / 8 ? sp : G / 8;
    const float limit = cull();
    uint32_t node = 0;
    bool mine = false;
    if ((int)j < nf) {
      node = stack[2 * (sp - 1 - (int)j)];
      mine = __builtin_bit_cast(float, stack[2 * (sp - 1 - (int)j) + 1]) <= limit;
    }
    wave_lds_sync();
    sp -= nf;
    visits += (uint32_t)__builtin_popcountll(__ballot(mine && i == 0) & gm);
    if ((__ballot(mine && node >= wv.node_count) & gm) != 0 || ++steps > wv.node_count) {
      if (gl == 0) atomicOr(wv.status, (uint32_t)BAJO_STATUS_BAD_NODE);
      sp = 0;
      return true;
    }
    bool hit = false;
    uint32_t ref = 0;
    float tin = 0.0f;
    if (mine) {
      const uint4* np = reinterpret_cast<const uint4*>(wv.nodes + node);
      const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
      ref = reinterpret_cast<const uint32_t*>(wv.nodes + node)[16 + i];
      const uint32_t meta = w0.w;
      const uint32_t n = meta >> 24;
      const float sx = __builtin_bit_cast(float, (meta & 0xFFu) << 23);
      const float sy = __builtin_bit_cast(float, ((meta >> 8) & 0xFFu) << 23);
      const float sz = __builtin_bit_cast(float, ((meta >> 16) & 0xFFu) << 23);
      const float ax = (__builtin_bit_cast(float, w0.x) - ray.o.x) * rcp.x;
      const float ay = (__builtin_bit_cast(float, w0.y) - ray.o.y) * rcp.y;
      const float az = (__builtin_bit_cast(float, w0.z) - ray.o.z) * rcp.z;
      const float bx = sx * rcp.x, by = sy * rcp.y, bz = sz * rcp.z;
      const bool hiw = i >= 4;
      const int bt = (int)(i & 3u);
      const uint32_t lox = hiw ? w1.y : w1.x, loy = hiw ? w1.w : w1.z, loz = hiw ? w2.y : w2.x;
      const uint32_t hix = hiw ? w2.w : w2.z, hiy = hiw ? w3.y : w3.x, hiz = hiw ? w3.w : w3.z;
      const uint32_t qnx = px ? lox : hix, qfx = px ? hix : lox;
      const uint32_t qny = py ? loy : hiy, qfy = py ? hiy : loy;
      const uint32_t qnz = pz ? loz : hiz, qfz = pz ? hiz : loz;
      const float tnx = fmaf(ubyte_f(qnx, bt), bx, ax), tfx = fmaf(ubyte_f(qfx, bt), bx, ax);
      const float tny = fmaf(ubyte_f(qny, bt), by, ay), tfy = fmaf(ubyte_f(qfy, bt), by, ay);
      const float tnz = fmaf(ubyte_f(qnz, bt), bz, az), tfz = fmaf(ubyte_f(qfz, bt), bz, az);
      tin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
      const float tout = fminf(fminf(tfx, tfy), fminf(tfz, limit));
      hit = i < n && tin <= tout;
    }
    const bool is_leaf = hit && (ref & kLeafFlag) != 0;
    const bool is_inner = hit && (ref & kLeafFlag) == 0;
    const uint64_t inner = __ballot(is_inner) & gm;
    const int n_inner = __builtin_popcountll(inner);
    if (sp + n_inner > kStack) {
      if (gl == 0) atomicOr(wv.status, (uint32_t)BAJO_STATUS_STACK_OVERFLOW);
      sp = 0;
      return true;
    }
    if (is_inner) {
      const int at = sp + (int)mask_rank(inner);
      stack[2 * at] = ref;
      stack[2 * at + 1] = __builtin_bit_cast(uint32_t, tin);
    }
    sp += n_inner;
    uint32_t start = 0, k = 0;
    if (is_leaf) {
      start = ref & kWideLeafStartMask;
      const uint32_t e = start + ((ref >> 28) & 7u) + 1u;
      k = (e > wv.prim_count ? wv.prim_count : e) - (start < wv.prim_count ? start : wv.prim_count);
    }
    uint32_t incl = k;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, G);
      if (gl >= (uint32_t)off) incl += y;
    }
    const uint32_t total = __shfl(incl, G - 1, G);
    for (uint32_t m = 0; m < k; ++m) recs[incl - k + m] = start + m;
    wave_lds_sync();
    tests += total;
    for (uint32_t base = 0; base < total; base += G) {
      float t = kFltMax;
      uint32_t s = kEmpty;
      if (base + gl < total) {
        s = recs[base + gl];
        const LeafRec rec = load_rec<false>(wv.tris, s);
        const TriData tr{mk(rec.a.x, rec.a.y, rec.a.z), mk(rec.b.x, rec.b.y, rec.b.z),
                         mk(rec.c.x, rec.c.y, rec.c.z), __builtin_bit_cast(uint32_t, rec.a.w)};
        float ad, ts, us, vs;
        if (tri_scaled(ray, tr, ray.tmax, ad, ts, us, vs)) {
          const float inv = 1.0f / ad;
          t = ts * inv;
        } else {
          s = kEmpty;
        }
      }
      // the two smallest distances of this round (the lowest lane wins equal ones)
      float m1 = t;
      uint32_t r1 = s, l1 = gl;
#pragma unroll
      for (int off = 1; off < G; off <<= 1) {
        const float om = __shfl_xor(m1, off, G);
        const uint32_t orr = (uint32_t)__shfl_xor((int)r1, off, G);
        const uint32_t ol = (uint32_t)__shfl_xor((int)l1, off, G);
        if (om < m1 || (om == m1 && ol < l1)) {
          m1 = om;
          r1 = orr;
          l1 = ol;
        }
      }
      float m2 = gl == l1 ? kFltMax : t;
#pragma unroll
      for (int off = 1; off < G; off <<= 1) m2 = fminf(m2, __shfl_xor(m2, off, G));
      if (r1 != kEmpty) {
        if (m1 < t1) {
          t2 = fminf(t1, m2);
          t1 = m1;
          best = r1;
        } else {
          t2 = fminf(t2, m1);
        }
      }
    }
    wave_lds_sync();
    return sp == 0;
  }

  // After the walk: whether Walk8 must decide (a near tie at the nearest
  // candidate), else the hit record of the nearest candidate.
  BJ_D bool near_tie() const { return best != kEmpty && t2 <= t1 * (1.0f + kMargin); }
  BJ_D Hit finish(const WideView& wv) const {
    Closest c = no_closest(ray.tmax);
    if (best != kEmpty) {
      const LeafRec rec = load_rec<false>(wv.tris, best);
      rec_closest<false>(rec, best, ray, c);
    }
    return closest_hit_p<false>(wv.tris, c, ray);
  }
};

// kTraceShadow + w (w = 0..3): any hit through a shadow walk -- w = 0 the
// one-node Walk8, else traverse8_any with bit 0 = two nodes per step and
// bit 1 = leaf records four at a time -- the same result as
// BAJO_TRACE_ANY_HIT; the binary layout keeps its one-node walk.
template <int kMode, bool kStats, bool kWide, bool kSph = false>
BJ_D Hit trace_accel(const AccelView& av, const Ray& ray, Stack& st, TraceCounters* tc, bool any_rt = false) {
  if constexpr (kMode >= kTraceShadow) {
    constexpr int w = kMode - kTraceShadow;
    if constexpr (kWide && w == 0) return traverse8<BAJO_TRACE_ANY_HIT, kStats, kSph>(av.wide, ray, st, tc, false);
    else if constexpr (kWide) return traverse8_any<kStats, kSph, (w & 1) != 0, (w & 2) ? 4 : 1>(av.wide, ray, st, tc);
    else return traverse<BAJO_TRACE_ANY_HIT, kStats, kSph>(av.bin, ray, st, tc, false);
  } else if constexpr (kWide) {
    return traverse8<kMode, kStats, kSph>(av.wide, ray, st, tc, any_rt);
  } else {
    return traverse<kMode, kStats, kSph>(av.bin, ray, st, tc, any_rt);
  }
}

}  // namespace bj
