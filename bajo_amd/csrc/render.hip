// Wavefront path tracer for triangle scenes on gfx950.
//
// Replaces bajo's GPU wavefront (render_gpu / enqueue_render_gpu_triangles,
// rt/gpu/render.mojo:322-420, rt/gpu/triangle_path.mojo:143-235, chunk
// scheduler rt/gpu/resources.mojo:187-285) for Lambertian + emissive
// triangle scenes and RENDER.PATH / NORMALS / AO / NEE / MIS:
//   primary  rt/gpu/common_kernels.mojo:22-94
//   extend   rt/gpu/scene_trace.mojo:41-412 (closest hit, sky, emission,
//            NEE candidate rt/gpu/path_shading.mojo:261-403, inline
//            Lambertian + roulette :463-535, routing :538-648)
//   shadow   rt/gpu/scene_trace.mojo:517-624
//   advance  rt/gpu/wavefront_contract.mojo:441-457 (no host sync)
//   resolve  rt/gpu/common_kernels.mojo:97-136
// Queues keep bajo's field-major planes (rt/wavefront_contract.mojo:41-99).
//
// MI355X design: the extend and shadow kernels are persistent waves that
// fetch 64-path batches; queue appends are wave-aggregated (one atomic per
// wave, slots ranked by ballot) and happen after the per-path work has
// reconverged; HBM is large enough that a whole 1920x1080x8 frame is one
// chunk, so a frame costs max_depth x (extend, shadow, advance) launches.
#include <string.h>

#include <vector>

#include "guide_internal.h"
#include "rt_math.h"
#include "traverse.h"

namespace {
using namespace bj;

constexpr int kRtBlock = kTraceBlock;
// MAT kinds (rt/types.mojo:24-29)
constexpr uint32_t kMatLambertian = 0, kMatMetal = 1, kMatDielectric = 2, kMatEmissive = 3;
// a GGX microfacet conductor (F0.rgb, roughness per material); bajo has no counterpart
constexpr uint32_t kMatGgx = 4;
constexpr uint32_t kSurfaceIndexMask = (1u << 28) - 1;
constexpr uint32_t kPrimSphere = 0, kPrimTriangle = 1;  // bajo's PRIM values
constexpr uint32_t kSphereHit = 1u;  // Hit.inst of a sphere hit in a mixed scene
// scene geometry of the render kernels: triangles, spheres + triangles, or a
// batch of triangle environments (one BVH segment per environment)
constexpr int kGeoTriangles = 0, kGeoMixed = 1, kGeoBatch = 2;
// Shadow walk form used when the settings leave it at 0 (1-based, as in
// bajo_render_settings.shadow_walk): one node per step with leaf records
// loaded four at a time measured fastest on the 4M-triangle NEE frame
// (1.49 ms of shadow launches per render against 1.63 for the one-record
// walk; two nodes per step lost occupancy and time).
constexpr uint32_t kDefaultShadowWalk = 3;
constexpr uint32_t kPathDelta = 0x80000000u, kPathIdMask = 0x7FFFFFFFu;
constexpr int kLightStride = 14;
constexpr int kPathPlanes = 12, kShadowPlanes = 11, kSamplePlanes = 3;
// counter words (bajo's WAVE_COUNTER plus two fetch cursors; the lane-group
// tail claims paths from kFetchExt)
constexpr int kActive = 0, kNext = 1, kShade = 2, kShadow = 3, kStatus = 4, kFetchExt = 5, kFetchShadow = 6,
              kTailDone = 7;
// guided renders: vertices queued for the guide this bounce (word 8 is the traversal status)
constexpr int kVertex = 9;
// blocks of work partials the bounce's shadow kernel wrote (each sets it when it traces the bounce's rays)
constexpr int kSparseShadow = 10;
// Next paths leaving their surface at a grazing angle are appended from the
// back of the next queue, the others from the front; words kNextPair and
// kNextPair + 1 count the two parts (one 64-bit atomic per wave reserves
// both), and kFront holds the front part's length for the next bounce:
// grazing rays walk many boxes along the surface, steep ones leave the
// hierarchy after a few, so each wave of the next bounce gets walks of like
// length.  A path's sample and random numbers do not depend on its queue
// position.
constexpr int kNextPair = 12, kFront = 14;
// (measured on the 4M-triangle frame's bounce 1: 0.7 → 1.69 ms, 0.5 → 1.56,
// 0.35 → 1.50, 0.25 → 1.50, 0.2 → 1.56 for NEE)
constexpr float kGrazingCos = 0.35f;
// shadow rays per bounce at or below which one wave traces each ray
constexpr uint32_t kSparseShadowRays = 8192;
// Lanes per shadow ray above that threshold when the settings leave it at 0:
// 16 up to kGroupShadowRays rays, one lane per ray beyond.  On the
// 4M-triangle NEE frame's first bounce (412 K shadow rays) 16 lanes take
// 0.76 ms against 0.95 with one lane per ray (0.80 with 8, 0.85 with 32);
// on the 4096-environment batch (about 3 M shadow rays of small scenes) one
// lane per ray is fastest.
constexpr uint32_t kDefaultShadowGroup = 16;
constexpr uint32_t kGroupShadowRays = 1u << 20;
// Lanes per path of the tail when the settings leave it at 0: 32 up to
// kGroupTailPaths live paths (the tail takes over at a few thousand), one lane
// per path above.  Four nodes per step shorten the slowest paths' chains: the
// 4M-triangle frame's tail takes 0.29 ms (NEE) and 0.23 ms (PATH) against
// 0.36 and 0.29 with 16 lanes, and 0.36 and 0.31 with 64.
constexpr uint32_t kDefaultTailGroup = 32;
constexpr uint32_t kGroupTailPaths = 1u << 15;
constexpr int kCounterWords = 16;
constexpr uint32_t kStatusPathOverflow = 1, kStatusShadowOverflow = 3;
constexpr int kMaxBounces = 64;
// stats buffer: 2 x kMaxBounces per-bounce ray counts, then traversal work
// totals (extension node visits, extension leaf tests, shadow node visits,
// shadow leaf tests)
constexpr int kWorkBase = 2 * kMaxBounces;
// then the largest node-visit count of one extension ray and of one shadow
// ray, per bounce (kMaxBounces each)
constexpr int kMaxVisitBase = kWorkBase + 4;
constexpr int kWorkStride = 3;  // per-block partials: node visits, leaf tests, max visits of one ray
// then the vertices that sampled through the guide (guided renders)
constexpr int kGuidedWord = kMaxVisitBase + 2 * kMaxBounces;
constexpr int kStatsWords = kGuidedWord + 1;
// Algorithms that compute radiance (the tail and the wavefront's emission
// and sky terms apply), and those whose emission hits take NEE's weights.
constexpr bool radiance_alg(int a) {
  return a == BAJO_RENDER_PATH || a == BAJO_RENDER_NEE || a == BAJO_RENDER_MIS || a == BAJO_RENDER_GUIDED;
}
constexpr bool nee_alg(int a) { return a == BAJO_RENDER_NEE || a == BAJO_RENDER_GUIDED; }

struct SceneView {
  AccelView bvh;
  AccelView spheres;              // sphere BVH of a mixed scene (radii |r|)
  const uint32_t* sphere_surfaces;
  const float* sphere_radii;      // signed: r < 0 turns the normal inward
  const uint32_t* surfaces;
  const float* lambertian;
  const float* metals;       // albedo.xyz, fuzz per metal
  const float* dielectrics;  // refraction index per dielectric
  const float* ggx;          // specular color F0.xyz, roughness per GGX material
  const float* emissive;
  const float* light_fields;
  const uint32_t* light_kinds;
  uint32_t light_count;
  float total_light_weight;
  // A batch of environments (env_paths = 0 otherwise): path ids are
  // environment-major, env_paths per environment; environment e traces BVH
  // segment e (triangles [seg_off[e], seg_off[e+1])), samples the lights
  // [env_light_base[e], env_light_base[e+1]) of total weight env_light_total[e]
  // and draws its random numbers with key seed + e and its local path id, so
  // each environment renders exactly as its scene would alone with that seed.
  const uint64_t* seg_off;
  const uint32_t* env_light_base;
  const float* env_light_total;
  uint32_t env_paths;
};

BJ_D StageRng path_rng(const SceneView& sc, uint64_t seed, uint32_t path_id, uint32_t stage) {
  if (sc.env_paths) {
    const uint32_t e = path_id / sc.env_paths;
    return StageRng(seed + e, path_id - e * sc.env_paths, stage);
  }
  return StageRng(seed, path_id, stage);
}

struct Cam {
  float v[20];
};

struct Queues {
  const uint32_t* src_ids;
  const float* src_fields;
  uint32_t* dst_ids;
  float* dst_fields;
  uint32_t* shadow_ids;
  float* shadow_fields;
  float* sample;
  uint32_t* counters;
  uint64_t capacity;
  uint32_t sample_base;
  // guided renders: the vertex planes of this bounce (kVertexPlanes per
  // entry, 64-entry tiles), and when the frame trains, the radiance added per
  // path and bounce (rgb, max_depth slots per path)
  float* vertices;
  float* contrib;
  uint32_t depth;
};

// Radiance added to a path at `bounce`, kept per bounce for the guide's training.
BJ_D void contrib_add(float* contrib, uint32_t slot, uint32_t depth, uint32_t bounce, v3 c) {
  float* p = contrib + 3ull * ((uint64_t)slot * depth + bounce);
  p[0] += c.x;
  p[1] += c.y;
  p[2] += c.z;
}

// Queue planes live in tiles of 64 entries (one wave): plane p of entry i is
// at ((i / 64) * planes + p) * 64 + i % 64.  A wave touches 256 contiguous
// bytes per plane, and every plane's offset inside a tile is a constant, so
// a kernel needs one base address per queue instead of one per plane.
template <int kPlanes>
BJ_D uint64_t plane_at(uint32_t i, int p) {
  return ((uint64_t)(i >> 6) * kPlanes + (uint64_t)p) * 64u + (i & 63u);
}

BJ_D void accumulate(float* sample, uint32_t slot, v3 c) {
  sample[plane_at<kSamplePlanes>(slot, 0)] += c.x;
  sample[plane_at<kSamplePlanes>(slot, 1)] += c.y;
  sample[plane_at<kSamplePlanes>(slot, 2)] += c.z;
}

// rt/lighting.mojo:7-20.
BJ_D float power_heuristic(float a, float b) {
  float a2 = a * a, b2 = b * b, den = a2 + b2;
  return den > 0.0f ? a2 / den : 0.0f;
}

// ---- primary (rt/gpu/common_kernels.mojo:22-94, bvh/camera.mojo:156-208) ----
struct PrimaryArgs {
  const float* cam;  // 20 floats in device memory (the frame's camera)
  uint32_t width, height, spp;
  uint32_t env_paths;  // paths per environment of a batch (0: one scene)
  uint32_t cam_stride; // 20: one camera per environment at cam + 20 e; 0: one camera
};

// The camera path of `path_id`: origin, t_min, direction, t_max, throughput
// 1 and pdf 0 in the path-plane order of the queues.
BJ_D void primary_path(const PrimaryArgs& pa, uint32_t path_id, uint64_t seed, float (&f)[kPathPlanes]) {
  const uint32_t width = pa.width, height = pa.height, spp = pa.spp;
  uint32_t local = path_id;
  uint64_t key = seed;
  const float* __restrict__ k = pa.cam;
  if (pa.env_paths) {
    const uint32_t e = path_id / pa.env_paths;
    local = path_id - e * pa.env_paths;
    key = seed + e;
    k += (uint64_t)e * pa.cam_stride;
  }
  const uint32_t pixel = local / spp;
  const uint32_t px = pixel % width, py = pixel / width;
  StageRng rng(key, local, 0u);
  // random_in_unit_disk (core/random.mojo:56-61), then the pixel jitter
  float s, c;
  sincos_turns(rng.next(), s, c);
  const float rr = sqrtf(rng.next());
  const float lens_u = rr * c, lens_v = rr * s;
  const float pu = rng.next(), pv = rng.next();
  const float fw = (float)width, fh = (float)height;
  const float aspect = fw / fh;
  const float sx = (((float)px + pu) / fw) * 2.0f - 1.0f;
  const float sy = 1.0f - (((float)py + pv) / fh) * 2.0f;
  const v3 origin = mk(k[0], k[1], k[2]);
  const v3 fwd = mk(k[3], k[4], k[5]), right = mk(k[6], k[7], k[8]), up = mk(k[9], k[10], k[11]);
  const float fov = k[12], focus = k[13];
  const v3 plane = add(add(fwd, scale(right, sx * aspect * fov)), scale(up, sy * fov));
  const v3 focal = add(origin, scale(plane, focus));
  const v3 ro = add(add(origin, scale(mk(k[14], k[15], k[16]), lens_u)), scale(mk(k[17], k[18], k[19]), lens_v));
  const v3 d = normalize(sub(focal, ro));
  const float v[kPathPlanes] = {ro.x, ro.y, ro.z, kT_MinSecondary, d.x, d.y, d.z, kFltMax, 1.0f, 1.0f, 1.0f, 0.0f};
#pragma unroll
  for (int p = 0; p < kPathPlanes; ++p) f[p] = v[p];
}

// Resets the queue counters for a chunk; the status words (kStatus and the
// traversal status at word 8) and the per-bounce ray counts are cleared
// only for the first chunk of a frame.
__global__ void init_counters_kernel(uint32_t* counters, uint32_t active, unsigned long long* stats, int first) {
  const uint32_t t = threadIdx.x;
  if (t < kCounterWords) {
    const bool status_word = t == kStatus || t == 8;
    if (!status_word || first) counters[t] = (t == kActive || t == kFront) ? active : 0u;
  }
  if (first)
    for (uint32_t i = t; i < kStatsWords; i += blockDim.x) stats[i] = 0;
}

// Blocks of the partials reduction: each sums a strided share of the live
// blocks' partials into one row of six sub-totals.
constexpr uint32_t kSumBlocks = 64;

// Sums (a, b, c, d) and maxima (ma, mc) over the 64 lanes of a wave.
BJ_D void wave_reduce_work(unsigned long long (&v)[6]) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const unsigned long long o = __shfl_xor(v[k], off, 64);
      v[k] = k < 4 ? v[k] + o : (o > v[k] ? o : v[k]);
    }
  }
}

// The per-block traversal partials of one bounce's extend and shadow
// launches, reduced by kSumBlocks blocks into `sums` (kSumBlocks rows of 6).
// Blocks past the live entries exited without writing their partials.
__global__ void __launch_bounds__(256) partials_sum_kernel(const uint32_t* __restrict__ counters, uint32_t cap,
                                                           const unsigned long long* __restrict__ ext_work,
                                                           uint32_t ext_blocks,
                                                           const unsigned long long* __restrict__ sh_work,
                                                           uint32_t sh_blocks, unsigned long long* __restrict__ sums) {
  __shared__ unsigned long long s_part[4][6];
  if (counters[kTailDone]) return;  // the tail kernel traced and counted the rest of the chunk
  const uint32_t active = counters[kActive] < cap ? counters[kActive] : cap;
  const uint32_t shadows = counters[kShadow] < cap ? counters[kShadow] : cap;
  const uint32_t ext_live = (active + 255u) / 256u < ext_blocks ? (active + 255u) / 256u : ext_blocks;
  uint32_t sh_live = (shadows + 255u) / 256u < sh_blocks ? (shadows + 255u) / 256u : sh_blocks;
  if (counters[kSparseShadow]) sh_live = counters[kSparseShadow];  // the shadow kernel's partials
  unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ext_live; i += stride) {
    v[0] += ext_work[kWorkStride * i];
    v[1] += ext_work[kWorkStride * i + 1];
    v[4] = ext_work[kWorkStride * i + 2] > v[4] ? ext_work[kWorkStride * i + 2] : v[4];
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sh_live; i += stride) {
    v[2] += sh_work[kWorkStride * i];
    v[3] += sh_work[kWorkStride * i + 1];
    v[5] = sh_work[kWorkStride * i + 2] > v[5] ? sh_work[kWorkStride * i + 2] : v[5];
  }
  wave_reduce_work(v);
  const uint32_t wave = threadIdx.x >> 6;
  if (lane_id() == 0)
    for (int k = 0; k < 6; ++k) s_part[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = (int)threadIdx.x;
    unsigned long long x = 0;
    for (int w = 0; w < 4; ++w) x = k < 4 ? x + s_part[w][k] : (s_part[w][k] > x ? s_part[w][k] : x);
    sums[6 * blockIdx.x + k] = x;
  }
}

// One wave: folds the bounce's partial sums into the render statistics and
// advances the queue counters to the next bounce.
__global__ void __launch_bounds__(64) advance_kernel(uint32_t* counters, unsigned long long* stats, uint32_t bounce,
                                                     uint32_t cap, const unsigned long long* __restrict__ sums) {
  const uint32_t t = threadIdx.x;
  if (counters[kTailDone]) return;  // the tail kernel traced and counted the rest of the chunk
  unsigned long long v[6] = {0, 0, 0, 0, 0, 0};
  if (t < kSumBlocks)
    for (int k = 0; k < 6; ++k) v[k] = sums[6 * t + k];
  wave_reduce_work(v);
  if (t != 0) return;
  for (int k = 0; k < 4; ++k) stats[kWorkBase + k] += v[k];
  if (bounce < kMaxBounces) {
    for (int k = 0; k < 2; ++k) {
      const unsigned long long x = stats[kMaxVisitBase + 2 * bounce + k];
      stats[kMaxVisitBase + 2 * bounce + k] = v[4 + k] > x ? v[4 + k] : x;
    }
    stats[2 * bounce + 0] += counters[kActive];
    stats[2 * bounce + 1] += counters[kShadow];
  }
  stats[kGuidedWord] += counters[kVertex];
  // appends past the arena were dropped (and flagged); never read past it
  const uint32_t nf = counters[kNextPair], nb = counters[kNextPair + 1];
  const uint32_t front = nf < cap ? nf : cap;
  const uint32_t back = nb < cap - front ? nb : cap - front;
  counters[kActive] = front + back;
  counters[kFront] = front;
  counters[kNext] = 0;
  counters[kNextPair] = 0;
  counters[kNextPair + 1] = 0;
  counters[kShade] = 0;
  counters[kShadow] = 0;
  counters[kSparseShadow] = 0;
  counters[kFetchExt] = 0;
  counters[kFetchShadow] = 0;
  counters[kVertex] = 0;
}

struct NextPath {
  uint32_t id;
  float f[kPathPlanes];
};
struct ShadowWork {
  uint32_t id;
  float f[kShadowPlanes];
};

// Block-reduces per-lane traversal work and stores the block's two totals
// (no atomics: the advance kernel sums the per-block partials).  Every
// thread of the block must call it.
BJ_D void store_block_work(unsigned long long* partials, uint64_t a, uint64_t b, uint32_t m) {
  __shared__ unsigned long long s_work[kRtBlock / 64][3];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
    const uint32_t o = (uint32_t)__shfl_xor((int)m, off, 64);
    m = o > m ? o : m;
  }
  const uint32_t wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_work[wave][0] = a;
    s_work[wave][1] = b;
    s_work[wave][2] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0, z = 0;
    for (int w = 0; w < kRtBlock / 64; ++w) {
      x += s_work[w][0];
      y += s_work[w][1];
      z = s_work[w][2] > z ? s_work[w][2] : z;
    }
    partials[kWorkStride * blockIdx.x + 0] = x;
    partials[kWorkStride * blockIdx.x + 1] = y;
    partials[kWorkStride * blockIdx.x + 2] = z;
  }
}

// Paths per fetch: one batch of 64 per atomic while the queue is short, up
// to eight batches when every wave would otherwise fetch more than eight
// times (fewer same-address atomics, still balanced at the tail).
BJ_D uint32_t fetch_batch(uint32_t count) {
  const uint32_t waves = gridDim.x * (kRtBlock / 64);
  uint32_t k = count / (waves * 64u * 8u);
  k = k < 1u ? 1u : (k > 8u ? 8u : k);
  return 64u * k;
}

// Wave-aggregated append: one atomic per wave, slots in lane order.
BJ_D uint32_t wave_reserve(uint32_t* counter, bool want) {
  const uint64_t m = ballot(want);
  if (m == 0) return 0xFFFFFFFFu;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__builtin_popcountll(m));
  base = __shfl(base, (int)leader, 64);
  return want ? base + mask_rank(m) : 0xFFFFFFFFu;
}

// The queue slot of a bounce's entry `idx`: the front part in append order,
// then the back part from the last slot down.
BJ_D uint32_t src_slot(uint32_t idx, uint32_t front, uint64_t cap) {
  return idx < front ? idx : (uint32_t)(cap - 1u - (idx - front));
}

// Appends the next paths of a wave, grazing ones (|cos| to the surface normal
// below kGrazingCos) from the back of the queue, with one atomic on the
// packed pair of counts; returns the slot (the capacity when the queue is
// full).
BJ_D uint32_t reserve_next(const Queues& q, bool want, bool grazing) {
  const uint64_t mf = ballot(want && !grazing), mb = ballot(want && grazing);
  if ((mf | mb) == 0) return 0xFFFFFFFFu;
  const uint32_t leader = (uint32_t)__builtin_ctzll(mf | mb);
  uint32_t lo = 0, hi = 0;
  if (lane_id() == leader) {
    const unsigned long long inc =
        (unsigned long long)__builtin_popcountll(mf) | ((unsigned long long)__builtin_popcountll(mb) << 32);
    const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(q.counters + kNextPair), inc);
    lo = (uint32_t)old;
    hi = (uint32_t)(old >> 32);
  }
  lo = __shfl(lo, (int)leader, 64);
  hi = __shfl(hi, (int)leader, 64);
  if (!want) return 0xFFFFFFFFu;
  const uint32_t cap = (uint32_t)q.capacity;
  if (!grazing) {
    const uint32_t ns = lo + mask_rank(mf);
    return ns < cap ? ns : cap;
  }
  const uint32_t nb = hi + mask_rank(mb);
  return nb < cap ? cap - 1u - nb : cap;
}

// rt/gpu/path_shading.mojo:261-403 (triangle lights only).
template <int kAlg>
BJ_D bool direct_light(const SceneView& sc, uint64_t seed, uint32_t path_id, uint32_t bounce, v3 point, v3 normal,
                       uint32_t kind, uint32_t mi, v3 ray_d, v3 throughput, ShadowWork& sw) {
  uint32_t lbase = 0, lcount = sc.light_count;
  float ltotal = sc.total_light_weight;
  if (sc.env_paths) {
    const uint32_t e = path_id / sc.env_paths;
    lbase = sc.env_light_base[e];
    lcount = sc.env_light_base[e + 1] - lbase;
    ltotal = sc.env_light_total[e];
  }
  if (lcount == 0 || !(ltotal > 0.0f)) return false;
  StageRng rng = path_rng(sc, seed, path_id, rng_light_stage(bounce));
  const float alias_sample = rng.next() * (float)lcount;
  uint32_t sel = (uint32_t)alias_sample;
  if (sel >= lcount) sel = lcount - 1;
  const float frac = alias_sample - (float)sel;
  const uint32_t packed = sc.light_kinds[lbase + sel];
  if (frac > sc.light_fields[(uint64_t)(lbase + sel) * kLightStride + 13]) sel = packed >> 4;
  const float* L = sc.light_fields + (uint64_t)(lbase + sel) * kLightStride;
  const v3 p0 = mk(L[0], L[1], L[2]);
  v3 ln, lp;
  if ((sc.light_kinds[lbase + sel] & 0xFu) == kPrimSphere) {
    // a sphere light: random_unit_vector (core/random.mojo:39-45) as the light normal
    float s, c;
    sincos_turns(rng.next(), s, c);
    const float z = 1.0f - 2.0f * rng.next();
    const float r = sqrtf(1.0f - z * z);
    ln = mk(r * c, r * s, z);
    lp = add(p0, scale(ln, L[9]));
  } else {
    const v3 p1 = mk(L[3], L[4], L[5]), p2 = mk(L[6], L[7], L[8]);
    const v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
    const v3 av = cross(e1, e2);
    const float tas = length2(av);
    if (tas <= 0.0f) return false;
    const float sq = sqrtf(tas);
    ln = mk(av.x / sq, av.y / sq, av.z / sq);
    const float root_u = sqrtf(rng.next());
    const float b1 = root_u * (1.0f - rng.next());
    const float b2 = root_u - b1;
    lp = add(add(p0, scale(e1, b1)), scale(e2, b2));
  }
  const v3 to_light = sub(lp, point);
  const float d2 = length2(to_light);
  if (d2 <= 1.0e-8f) return false;
  const float dist = sqrtf(d2);
  const v3 dir = mk(to_light.x / dist, to_light.y / dist, to_light.z / dist);
  const float scos = fmax_sel(dot(normal, dir), 0.0f);
  const float lcos = fmax_sel(dot(ln, mk(-dir.x, -dir.y, -dir.z)), 0.0f);
  if (scos <= 0.0f || lcos <= 0.0f) return false;
  const float tmax = dist - 0.002f;
  if (tmax <= 0.001f) return false;
  const v3 E = mk(L[10], L[11], L[12]);
  const float light_pdf = d2 * (E.x + E.y + E.z) / (3.0f * lcos * ltotal);
  v3 value;
  float bsdf_pdf;
  if (kind == kMatMetal) {
    // _evaluate_metal (rt/shading.mojo:30-57); a mirror gives pdf 0: no light sample
    const float* m = sc.metals + 4 * (uint64_t)mi;
    const BsdfEval ev = eval_metal(ray_d, normal, mk(m[0], m[1], m[2]), m[3], dir);
    value = ev.value;
    bsdf_pdf = ev.pdf;
  } else if (kind == kMatGgx) {
    const float* g = sc.ggx + 4 * (uint64_t)mi;
    const BsdfEval ev = eval_ggx(ray_d, normal, mk(g[0], g[1], g[2]), g[3], dir);
    value = ev.value;
    bsdf_pdf = ev.pdf;
  } else {
    // _evaluate_lambertian (rt/shading.mojo:13-27)
    const float* a = sc.lambertian + 3 * (uint64_t)mi;
    value = scale(mk(a[0], a[1], a[2]), kInvPi);
    bsdf_pdf = fmax_sel(dot(normal, normalize(dir)), 0.0f) * kInvPi;
  }
  // _direct_light_scale (rt/lighting.mojo:23-38)
  const bool ok = light_pdf > 0.0f && bsdf_pdf > 0.0f;
  float weight = 1.0f;
  if constexpr (kAlg == BAJO_RENDER_MIS) weight = power_heuristic(light_pdf, bsdf_pdf);
  const float sc_ = ok ? scos * weight / light_pdf : 0.0f;
  if (!(sc_ > 0.0f)) return false;
  const v3 contrib = scale(mul(mul(throughput, value), E), sc_);
  const float f[kShadowPlanes] = {point.x, point.y, point.z, 0.001f, dir.x, dir.y, dir.z, tmax,
                                  contrib.x, contrib.y, contrib.z};
  for (int p = 0; p < kShadowPlanes; ++p) sw.f[p] = f[p];
  sw.id = path_id;
  return true;
}

// Wave-aggregated maximum of a per-lane count.
BJ_D void wave_max(unsigned long long* dst, uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o > v ? o : v;
  }
  if (lane_id() == 0 && v) atomicMax(dst, (unsigned long long)v);
}

// Wave-aggregated 64-bit add of a per-lane count.
BJ_D void wave_add(unsigned long long* dst, uint32_t v) {
  uint64_t x = v;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  if (lane_id() == 0 && x) atomicAdd(dst, (unsigned long long)x);
}

// Surface facts of an extension hit (rt/geometry.mojo:20-30: the outward
// normal is oriented against the ray).
struct Surface {
  bool found, front;
  v3 n, point;
  uint32_t kind, mi;
};

BJ_D Surface surface_of(const SceneView& sc, const Ray& ray, const Hit& h) {
  Surface su;
  su.found = h.prim != kEmpty && h.t < ray.tmax;
  su.front = false;
  su.n = mk(0, 0, 0);
  su.point = mk(0, 0, 0);
  su.kind = su.mi = 0;
  if (su.found) {
    // rt/gpu/scene_trace.mojo:137-145: a sphere's outward normal, inward for a negative radius
    const bool sph = h.inst == kSphereHit;
    const v3 hn = sph && sc.sphere_radii[h.prim] < 0.0f ? mk(-h.n.x, -h.n.y, -h.n.z) : h.n;
    su.front = dot(ray.d, hn) < 0.0f;
    su.n = su.front ? hn : mk(-hn.x, -hn.y, -hn.z);
    su.point = add(ray.o, scale(ray.d, h.t));
    const uint32_t surf = sph ? sc.sphere_surfaces[h.prim] : sc.surfaces[h.prim];
    su.kind = surf >> 28;
    su.mi = surf & kSurfaceIndexMask;
  }
  return su;
}

BJ_D v3 albedo_of(const SceneView& sc, const Surface& su) {
  return su.kind == kMatLambertian
             ? mk(sc.lambertian[3 * su.mi + 0], sc.lambertian[3 * su.mi + 1], sc.lambertian[3 * su.mi + 2])
             : mk(0, 0, 0);
}

// The shadow ray an extension hit spawns: the NEE / MIS light sample
// (rt/gpu/path_shading.mojo:261-460) or the AO hemisphere ray (:651-667).
template <int kAlg>
BJ_D bool shade_shadow(const SceneView& sc, uint64_t seed, uint32_t bounce, uint32_t path_id, v3 T, v3 ray_d,
                       const Surface& su, ShadowWork& sw) {
  if (!su.found) return false;
  if constexpr (kAlg == BAJO_RENDER_AO) {
    StageRng rng = path_rng(sc, seed, path_id, 1u);
    float s, c;
    sincos_turns(rng.next(), s, c);
    const float z = 1.0f - 2.0f * rng.next();
    const float r = sqrtf(1.0f - z * z);
    v3 dir = mk(r * c, r * s, z);
    if (dot(dir, su.n) < 0.0f) dir = mk(-dir.x, -dir.y, -dir.z);
    const float fs[kShadowPlanes] = {su.point.x, su.point.y, su.point.z, 0.001f, dir.x, dir.y, dir.z, 4.0f, 0, 0, 0};
    for (int p = 0; p < kShadowPlanes; ++p) sw.f[p] = fs[p];
    sw.id = path_id;
    return true;
  } else if constexpr (nee_alg(kAlg) || kAlg == BAJO_RENDER_MIS) {
    // _sample_direct_light_candidate (rt/gpu/path_shading.mojo:261-403): Lambertian and metal
    // surfaces, and GGX surfaces (their BSDF has a density everywhere above the surface)
    if (su.kind != kMatLambertian && su.kind != kMatMetal && su.kind != kMatGgx) return false;
    return direct_light<kAlg>(sc, seed, path_id, bounce, su.point, su.n, su.kind, su.mi, ray_d, T, sw);
  } else {
    return false;
  }
}

// Radiance an extension hit adds (sky on a miss, emission with the NEE/MIS
// weights, normals) and the next segment (Lambertian bounce with Russian
// roulette), rt/gpu/path_shading.mojo:463-648.  Shared by the wavefront
// extend kernel and the tail kernel so both follow one definition.
// russian_roulette (rt/common.mojo:43-67) at depth bounce + 1: from depth 5
// a path survives with probability max(T) clamped to [0.05, 0.95].
BJ_D bool survives_roulette(const SceneView& sc, uint64_t seed, uint32_t path_id, uint32_t bounce, v3& Tn) {
  if (bounce + 1u < 5u) return true;
  const float mx = fmax_sel(Tn.x, fmax_sel(Tn.y, Tn.z));
  if (mx <= 0.0f) return false;
  const float survival = fmin_sel(fmax_sel(mx, 0.05f), 0.95f);
  StageRng rr = path_rng(sc, seed, path_id, rng_roulette_stage(bounce));
  if (rr.next() >= survival) return false;
  Tn = mk(Tn.x / survival, Tn.y / survival, Tn.z / survival);
  return true;
}

// The next path segment from the hit point (rt/gpu/path_shading.mojo:768-782).
BJ_D void next_path(NextPath& np, uint32_t path_id, bool delta, const Ray& ray, float t_hit, v3 dir, v3 Tn, float pdf) {
  np.id = path_id | (delta ? kPathDelta : 0u);
  const float nf[kPathPlanes] = {ray.o.x + t_hit * ray.d.x, ray.o.y + t_hit * ray.d.y, ray.o.z + t_hit * ray.d.z,
                                 kT_MinSecondary, dir.x, dir.y, dir.z, kFltMax, Tn.x, Tn.y, Tn.z, pdf};
  for (int p = 0; p < kPathPlanes; ++p) np.f[p] = nf[p];
}

// shade_bounce returns kShadeAdd when `add_c` holds radiance for the sample
// and kShadeNext when `np` holds the next path segment (flags in one value:
// bools returned through references were kept in scratch memory).
constexpr uint32_t kShadeAdd = 1u, kShadeNext = 2u;
// guided renders: the hit is a Lambertian / GGX vertex whose continuation the
// guide's sample kernel draws
constexpr uint32_t kShadeVertex = 4u;
template <int kAlg>
BJ_D uint32_t shade_bounce(const SceneView& sc, uint64_t seed, uint32_t bounce, uint32_t path_id, bool delta,
                           const Ray& ray, v3 T, float path_pdf, float t_hit, const Surface& su, v3& add_c,
                           NextPath& np, uint32_t depth = kMaxBounces) {
  uint32_t out = 0;
  if (!su.found) {
    if constexpr (radiance_alg(kAlg)) {
      add_c = mul(T, sky_color(ray.d));
      out = kShadeAdd;
    }
    return out;
  }
  if constexpr (kAlg == BAJO_RENDER_NORMALS) {
    add_c = scale(add(su.n, mk(1.0f, 1.0f, 1.0f)), 0.5f);
    out |= kShadeAdd;
  } else if constexpr (radiance_alg(kAlg)) {
    if (su.kind == kMatEmissive) {
      if (su.front) {
        const v3 E = mk(sc.emissive[3 * su.mi + 0], sc.emissive[3 * su.mi + 1], sc.emissive[3 * su.mi + 2]);
        float w = 1.0f;
        if constexpr (nee_alg(kAlg)) {
          if (bounce > 0 && !delta) w = 0.0f;
        } else if constexpr (kAlg == BAJO_RENDER_MIS) {
          if (bounce > 0 && !delta) {
            const float lc = fmax_sel(dot(su.n, normalize(mk(-ray.d.x, -ray.d.y, -ray.d.z))), 0.0f);
            float lpdf = 0.0f;
            const float ltotal = sc.env_paths ? sc.env_light_total[path_id / sc.env_paths] : sc.total_light_weight;
            if (lc > 0.0f && ltotal > 0.0f) {
              const float d2 = t_hit * t_hit * length2(ray.d);
              lpdf = d2 * (E.x + E.y + E.z) / (3.0f * lc * ltotal);
            }
            w = power_heuristic(path_pdf, lpdf);
          }
        }
        add_c = scale(mul(T, E), w);
        out |= kShadeAdd;
      }
    } else if (kAlg == BAJO_RENDER_GUIDED && (su.kind == kMatLambertian || su.kind == kMatGgx)) {
      // the guide's sample kernel draws the continuation; after the last
      // bounce there is none to draw
      if (bounce + 1 < depth) out |= kShadeVertex;
    } else if (su.kind == kMatLambertian) {
      // _sample_lambertian (rt/shading.mojo:92-118)
      const v3 n = su.n;
      StageRng rng = path_rng(sc, seed, path_id, rng_bsdf_stage(bounce));
      const float ru = rng.next(), rv = rng.next();
      float s, c;
      sincos_turns(ru, s, c);
      const float z = 1.0f - 2.0f * rv;
      const float radius = sqrtf(fmax_sel(1.0f - z * z, 0.0f));
      v3 sd = add(n, mk(radius * c, radius * s, z));
      if (fabsf(sd.x) < 1e-8f && fabsf(sd.y) < 1e-8f && fabsf(sd.z) < 1e-8f) sd = n;
      const v3 dir = normalize(sd);
      const float pdf = fmax_sel(dot(n, normalize(dir)), 0.0f) * kInvPi;
      v3 Tn = mul(T, albedo_of(sc, su));
      if (survives_roulette(sc, seed, path_id, bounce, Tn)) {
        next_path(np, path_id, false, ray, t_hit, dir, Tn, pdf);
        out |= kShadeNext;
      }
    } else if (su.kind == kMatMetal || su.kind == kMatDielectric || su.kind == kMatGgx) {
      // _gpu_rt_shade_one (rt/gpu/path_shading.mojo:669-783): random numbers
      // are drawn only where the sample uses them
      StageRng rng = path_rng(sc, seed, path_id, rng_bsdf_stage(bounce));
      BsdfSample smp;
      if (su.kind == kMatGgx) {
        const float* g = sc.ggx + 4 * (uint64_t)su.mi;
        const float ru = rng.next(), rv = rng.next();
        smp = sample_ggx(ray.d, su.n, mk(g[0], g[1], g[2]), g[3], ru, rv);
      } else if (su.kind == kMatMetal) {
        const float* m = sc.metals + 4 * (uint64_t)su.mi;
        float ru = 0.0f, rv = 0.0f;
        if (m[3] > 1.0e-4f) {
          ru = rng.next();
          rv = rng.next();
        }
        smp = sample_metal(ray.d, su.n, mk(m[0], m[1], m[2]), m[3], ru, rv);
      } else {
        const float ior = sc.dielectrics[su.mi];
        const float ri = su.front ? 1.0f / ior : ior;
        const v3 u = normalize(ray.d);
        const float cos_t = fmin_sel(dot(mk(-u.x, -u.y, -u.z), su.n), 1.0f);
        const float sin_t = sqrtf(fmax_sel(1.0f - cos_t * cos_t, 0.0f));
        const float ru = ri * sin_t <= 1.0f ? rng.next() : 0.0f;
        smp = sample_dielectric(ray.d, su.n, ior, su.front, ru);
      }
      if (smp.ok) {
        v3 Tn = mul(T, smp.weight);
        if (survives_roulette(sc, seed, path_id, bounce, Tn)) {
          next_path(np, path_id, smp.delta, ray, t_hit, smp.direction, Tn, smp.pdf);
          out |= kShadeNext;
        }
      }
    }
  }
  return out;
}

// Radiance a traced shadow ray adds (rt/gpu/scene_trace.mojo:517-624).
template <int kAlg>
BJ_D v3 shadow_radiance(bool occluded, const float* f) {
  if constexpr (kAlg == BAJO_RENDER_AO) {
    const float v = occluded ? 0.08f : 1.0f;
    return mk(v, v, v);
  } else {
    return occluded ? mk(0.0f, 0.0f, 0.0f) : mk(f[8], f[9], f[10]);
  }
}

// kFirst: bounce 0 of a chunk, whose paths are the camera paths of the
// chunk's path ids (generated here instead of read from a queue); it also
// clears the chunk's sample accumulators.
// The scene's closest / any hit (rt/gpu/scene_trace.mojo:113-193): a mixed
// scene traces its spheres first, then its triangles with t_max = the sphere
// hit (a triangle must be strictly closer); any hit: either occludes.  In a
// batch, environment `env` traces its own segment and the hit's triangle
// index (local to the segment) becomes the batch's global index.
template <int kMode, bool kWide, int kGeo>
BJ_D Hit scene_trace(const SceneView& sc, const Ray& ray, Stack& st, TraceCounters* tc, bool any_rt = false,
                     uint32_t env = 0) {
  if constexpr (kGeo == kGeoTriangles) {
    return trace_accel<kMode, true, kWide>(sc.bvh, ray, st, tc, any_rt);
  } else if constexpr (kGeo == kGeoBatch) {
    AccelView av = sc.bvh;
    const uint64_t first = sc.seg_off[env];
    if constexpr (kWide) av.wide.root = env;  // segment roots lead the 8-wide node array
    else av.bin.root = sc.seg_off[env + 1] - first == 1 ? (kLeafFlag | (uint32_t)first) : (uint32_t)first;
    Hit h = trace_accel<kMode, true, kWide>(av, ray, st, tc, any_rt);
    if (h.prim != kEmpty) h.prim += (uint32_t)first;
    return h;
  } else {
    const bool any = kMode == BAJO_TRACE_ANY_HIT || kMode >= kTraceShadow || (kMode == kTraceEither && any_rt);
    TraceCounters ts, tt;
    const Hit hs = trace_accel<kMode, true, kWide, true>(sc.spheres, ray, st, &ts, any_rt);
    const bool sph = any ? (hs.prim == kEmpty && hs.t == 0.0f) : hs.prim != kEmpty;
    Hit h = hs;
    tt.visits = tt.inner = tt.leaves = tt.max_sp = 0;
    if (!(any && sph)) {
      Ray tr = ray;
      if (sph) tr.tmax = hs.t;
      const Hit ht = trace_accel<kMode, true, kWide>(sc.bvh, tr, st, &tt, any_rt);
      if (any) h = ht;
      else if (ht.prim != kEmpty && ht.t < tr.tmax) h = ht;
      else if (sph) h.inst = kSphereHit;
    }
    tc->visits = ts.visits + tt.visits;
    tc->inner = ts.inner + tt.inner;
    tc->leaves = ts.leaves + tt.leaves;
    tc->max_sp = ts.max_sp > tt.max_sp ? ts.max_sp : tt.max_sp;
    return h;
  }
}

template <int kAlg, bool kWide, bool kFirst, int kGeo, int kWaves = 6>
__global__ void __launch_bounds__(kRtBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) extend_kernel(SceneView sc, Queues q, uint64_t seed, uint32_t bounce,
                                                          PrimaryArgs pa, uint32_t* __restrict__ spill,
                                                          unsigned long long* __restrict__ work,
                                                          unsigned long long* __restrict__ stats) {
  // the LDS stack sized to the occupancy: eight waves per SIMD need eight
  // blocks' stacks in a CU's LDS; at four a deeper stack spills less
  constexpr int kDepth = kWaves >= 8 ? kLdsStackMin : (kWaves <= 4 ? kLdsStackMax : kLdsStack);
  __shared__ uint32_t s_stack[kDepth * kRtBlock];
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0, kDepth};
  const uint64_t cap = q.capacity;
  const bool tail_done = q.counters[kTailDone] != 0;
  const uint32_t active = tail_done ? 0u : (q.counters[kActive] < cap ? q.counters[kActive] : (uint32_t)cap);
  uint64_t work_visits = 0, work_leaves = 0;
  // One thread per queue entry; the grid covers the arena, and blocks past
  // the live entries exit at once (a persistent loop here made the compiler
  // keep far more registers live, halving the resident waves).
  if ((uint64_t)blockIdx.x * kRtBlock >= active) return;
  {
    const uint32_t idx = (uint32_t)gtid;
    const uint32_t sidx = kFirst ? idx : src_slot(idx, q.counters[kFront], cap);  // this entry's queue slot
    Surface su;
    su.found = false;
    Ray ray;
    float t_hit = 0.0f;
    uint32_t path_id = 0;
    bool delta = false;
    v3 T = mk(1.0f, 1.0f, 1.0f);
    float path_pdf = 0.0f;
    ShadowWork sw;
    bool want_shadow = false;
    const bool live = idx < active;
    if (live) {
      float f[8];
      if constexpr (kFirst) {
        float pf[kPathPlanes];
        primary_path(pa, q.sample_base + idx, seed, pf);
#pragma unroll
        for (int p = 0; p < 8; ++p) f[p] = pf[p];
#pragma unroll
        for (int p = 0; p < kSamplePlanes; ++p) q.sample[plane_at<kSamplePlanes>(idx, p)] = 0.0f;
      } else {
#pragma unroll
        for (int p = 0; p < 8; ++p) f[p] = q.src_fields[plane_at<kPathPlanes>(sidx, p)];
      }
      ray = Ray{mk(f[0], f[1], f[2]), mk(f[4], f[5], f[6]), f[3], f[7]};
      uint32_t env = 0;
      if constexpr (kGeo == kGeoBatch)
        env = (kFirst ? q.sample_base + idx : (q.src_ids[sidx] & kPathIdMask)) / sc.env_paths;
      TraceCounters tc;
      const Hit h = scene_trace<BAJO_TRACE_CLOSEST_HIT, kWide, kGeo>(sc, ray, st, &tc, false, env);
      work_visits += tc.visits;
      work_leaves += tc.leaves;
      // throughput, pdf and id are read after the traversal rather than
      // held in registers through it
      if constexpr (!kFirst) {
        const uint32_t idw = q.src_ids[sidx];
        path_id = idw & kPathIdMask;
        delta = (idw & kPathDelta) != 0;
        T = mk(q.src_fields[plane_at<kPathPlanes>(sidx, 8)], q.src_fields[plane_at<kPathPlanes>(sidx, 9)], q.src_fields[plane_at<kPathPlanes>(sidx, 10)]);
        path_pdf = q.src_fields[plane_at<kPathPlanes>(sidx, 11)];
      } else {
        path_id = q.sample_base + idx;
        delta = true;
      }
      su = surface_of(sc, ray, h);
      t_hit = h.t;
      want_shadow = shade_shadow<kAlg>(sc, seed, bounce, path_id, T, ray.d, su, sw);
    }
    const uint32_t ss = wave_reserve(q.counters + kShadow, want_shadow);
    if (want_shadow) {
      if (ss < cap) {
        q.shadow_ids[ss] = sw.id;
#pragma unroll
        for (int p = 0; p < kShadowPlanes; ++p) q.shadow_fields[plane_at<kShadowPlanes>(ss, p)] = sw.f[p];
      } else {
        atomicMax(q.counters + kStatus, kStatusShadowOverflow);
      }
    }
    bool want_next = false, want_vertex = false;
    NextPath np;
    if (live) {
      v3 add_c;
      const uint32_t sf =
          shade_bounce<kAlg>(sc, seed, bounce, path_id, delta, ray, T, path_pdf, t_hit, su, add_c, np, q.depth);
      want_next = (sf & kShadeNext) != 0;
      want_vertex = (sf & kShadeVertex) != 0;
      if (sf & kShadeAdd) {
        accumulate(q.sample, path_id - q.sample_base, add_c);
        if constexpr (kAlg == BAJO_RENDER_GUIDED)
          if (q.contrib) contrib_add(q.contrib, path_id - q.sample_base, q.depth, bounce, add_c);
      }
    }
    if constexpr (kAlg == BAJO_RENDER_GUIDED) {
      // the guide's queue: the hit, its oriented normal, the ray, throughput and ids
      const uint32_t vs = wave_reserve(q.counters + kVertex, want_vertex);
      if (want_vertex) {
        if (vs < cap) {
          const float vf[kVertexPlanes] = {su.point.x, su.point.y, su.point.z, su.n.x, su.n.y, su.n.z,
                                           ray.d.x, ray.d.y, ray.d.z, T.x, T.y, T.z,
                                           __builtin_bit_cast(float, (su.kind << 28) | su.mi),
                                           __builtin_bit_cast(float, path_id)};
#pragma unroll
          for (int p = 0; p < kVertexPlanes; ++p) q.vertices[vplane(vs, p)] = vf[p];
        } else {
          atomicMax(q.counters + kStatus, kStatusPathOverflow);
        }
      }
    }
    const bool grazing = want_next && fabsf(dot(mk(np.f[4], np.f[5], np.f[6]), su.n)) < kGrazingCos;
    const uint32_t ns = reserve_next(q, want_next, grazing);
    if (want_next) {
      if (ns < cap) {
        q.dst_ids[ns] = np.id;
#pragma unroll
        for (int p = 0; p < kPathPlanes; ++p) q.dst_fields[plane_at<kPathPlanes>(ns, p)] = np.f[p];
      } else {
        atomicMax(q.counters + kStatus, kStatusPathOverflow);
      }
    }
  }
  store_block_work(work, work_visits, work_leaves, (uint32_t)work_visits);
}

// The remaining bounces of a chunk once few paths are left.  Each thread
// carries one path from `bounce` to the end: extension ray, shading, the
// shadow ray traced at once, the next segment in registers.  A path's
// sample receives the same additions in the same order as in the wavefront
// (extension then shadow, bounce by bounce), so images are identical.  All
// blocks of a launch take the same decision from the queue length; a
// launch that runs marks the chunk done (tag = bounce + 1) so later extend,
// shadow, advance and tail launches of the chunk do nothing.
template <int kAlg, bool kWide, int kGeo>
__global__ void __launch_bounds__(kRtBlock) tail_kernel(SceneView sc, Queues q, uint64_t seed, uint32_t bounce,
                                                        uint32_t depth, uint32_t lo, uint32_t threshold, uint32_t simds,
                                                        unsigned long long* __restrict__ stats,
                                                        uint32_t* __restrict__ spill) {
  __shared__ uint32_t s_stack[kLdsStack * kRtBlock];
  const uint32_t tag = q.counters[kTailDone];
  if (tag != 0 && tag != bounce + 1u) return;
  const uint64_t cap = q.capacity;
  const uint32_t active = q.counters[kActive] < cap ? q.counters[kActive] : (uint32_t)cap;
  if (active <= lo || active > threshold) return;  // (lo: the lane-group tail takes those)
  if (blockIdx.x == 0 && threadIdx.x == 0) q.counters[kTailDone] = bounce + 1u;
  const uint32_t lane = lane_id();
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0};
  uint64_t ext_visits = 0, ext_leaves = 0, sh_visits = 0, sh_leaves = 0;
  const uint32_t nb = depth - bounce;
  // A wave is as slow as its slowest path over all remaining bounces, so the
  // paths are spread over about one wave per SIMD with only `group` lanes
  // each (a power of two, at most 64) instead of packing 64 to a wave.
  // Lanes pair up: the even lane carries the path (its extension rays), the
  // odd lane traces the path's shadow ray of bounce b while the path traces
  // its extension ray of bounce b + 1, which halves the serial chain.  The
  // shadow contribution of bounce b is added to the path's sum before the
  // extension contribution of bounce b + 1, the order of the wavefront.
  const uint32_t want = (2u * active + simds - 1) / (simds ? simds : 1u);
  uint32_t group = 2;
  while (group < want && group < 64) group <<= 1;
  const uint32_t paths = group >> 1;
  const bool is_path = (lane & 1u) == 0;
  const uint32_t mate = lane ^ 1u;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t gw = (uint32_t)(gtid >> 6); (uint64_t)gw * paths < active; gw += waves) {
    const uint32_t idx = gw * paths + (lane >> 1);
    const bool mine = lane < group && idx < active;
    bool alive = mine && is_path;
    uint32_t path_id = 0;
    bool delta = false;
    float f[kPathPlanes];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    uint32_t slot = 0;
    if (alive) {
      // the queue's back part (grazing paths, the long ones) first
      const uint32_t sidx = src_slot(active - 1u - idx, q.counters[kFront], cap);
      const uint32_t idw = q.src_ids[sidx];
      path_id = idw & kPathIdMask;
      delta = (idw & kPathDelta) != 0;
#pragma unroll
      for (int p = 0; p < kPathPlanes; ++p) f[p] = q.src_fields[plane_at<kPathPlanes>(sidx, p)];
      slot = path_id - q.sample_base;
#pragma unroll
      for (int p = 0; p < kSamplePlanes; ++p) acc[p] = q.sample[plane_at<kSamplePlanes>(slot, p)];
    }
    // both lanes of a pair work for the path's environment
    uint32_t env = 0;
    if constexpr (kGeo == kGeoBatch) env = __shfl(path_id, (int)(lane & ~1u), 64) / sc.env_paths;
    bool sh_live = false;  // shadow lane: a shadow ray of the previous bounce is pending
    float sf[kShadowPlanes];
#pragma unroll
    for (int p = 0; p < kShadowPlanes; ++p) sf[p] = 0.0f;
    // iteration k traces the extension rays of bounce + k and the shadow
    // rays of bounce + k - 1; the last iteration only the pending shadows
    for (uint32_t k = 0; k <= nb; ++k) {
      const uint32_t b = bounce + k;
      const bool ext_step = k < nb;
      if (!ext_step) alive = false;  // a path still going after the last bounce ends here
      if (ext_step && b < kMaxBounces) wave_add(stats + 2 * b, alive ? 1u : 0u);
      uint32_t ext_max = 0, sh_max = 0;
      // one traversal call for both lane kinds, so the path's extension ray
      // and its mate's shadow ray step through the same instructions
      Hit h;
      Ray ray;
      if (alive) ray = Ray{mk(f[0], f[1], f[2]), mk(f[4], f[5], f[6]), f[3], f[7]};
      if (sh_live) ray = Ray{mk(sf[0], sf[1], sf[2]), mk(sf[4], sf[5], sf[6]), sf[3], sf[7]};
      if (alive || sh_live) {
        TraceCounters tc;
        h = scene_trace<kTraceEither, kWide, kGeo>(sc, ray, st, &tc, sh_live, env);
        if (alive) {
          ext_visits += tc.visits;
          ext_leaves += tc.leaves;
          ext_max = tc.visits;
        } else {
          sh_visits += tc.visits;
          sh_leaves += tc.leaves;
          sh_max = tc.visits;
        }
      }
      bool add_s = false;
      v3 cs = mk(0, 0, 0);
      if (sh_live) {
        const bool occluded = h.prim == kEmpty && h.t == 0.0f;
        if (kAlg == BAJO_RENDER_AO || !occluded) {
          cs = shadow_radiance<kAlg>(occluded, sf);
          add_s = true;
        }
      }
      // the previous bounce's shadow contribution, from the mate lane
      const bool mate_add = __shfl(add_s, (int)mate, 64);
      const float cx = __shfl(cs.x, (int)mate, 64), cy = __shfl(cs.y, (int)mate, 64), cz = __shfl(cs.z, (int)mate, 64);
      if (mine && is_path && mate_add) {
        acc[0] += cx;
        acc[1] += cy;
        acc[2] += cz;
      }
      bool cast = false;
      float nsf[kShadowPlanes];
#pragma unroll
      for (int p = 0; p < kShadowPlanes; ++p) nsf[p] = 0.0f;
      if (alive) {
        const v3 T = mk(f[8], f[9], f[10]);
        const Surface su = surface_of(sc, ray, h);
        ShadowWork sw;
        if (shade_shadow<kAlg>(sc, seed, b, path_id, T, ray.d, su, sw)) {
          cast = true;
#pragma unroll
          for (int p = 0; p < kShadowPlanes; ++p) nsf[p] = sw.f[p];
        }
        v3 add_c;
        NextPath np;
        const uint32_t sfl = shade_bounce<kAlg>(sc, seed, b, path_id, delta, ray, T, f[11], h.t, su, add_c, np);
        if (sfl & kShadeAdd) {
          acc[0] += add_c.x;
          acc[1] += add_c.y;
          acc[2] += add_c.z;
        }
        alive = (sfl & kShadeNext) != 0;
        if (alive) {
#pragma unroll
          for (int p = 0; p < kPathPlanes; ++p) f[p] = np.f[p];
          delta = (np.id & kPathDelta) != 0;
        }
      }
      // the new shadow ray moves to the mate lane
      const bool mate_cast = __shfl(cast, (int)mate, 64);
#pragma unroll
      for (int p = 0; p < kShadowPlanes; ++p) sf[p] = __shfl(nsf[p], (int)mate, 64);
      sh_live = mine && !is_path && mate_cast;
      if (ext_step && b < kMaxBounces) {
        wave_add(stats + 2 * b + 1, cast ? 1u : 0u);
        wave_max(stats + kMaxVisitBase + 2 * b, ext_max);
      }
      if (k > 0 && b - 1 < kMaxBounces) wave_max(stats + kMaxVisitBase + 2 * (b - 1) + 1, sh_max);
      if (!ballot(alive || sh_live)) break;
    }
    if (mine && is_path) {
#pragma unroll
      for (int p = 0; p < kSamplePlanes; ++p) q.sample[plane_at<kSamplePlanes>(slot, p)] = acc[p];
    }
  }
  wave_add(stats + kWorkBase + 0, (uint32_t)ext_visits);
  wave_add(stats + kWorkBase + 1, (uint32_t)ext_leaves);
  wave_add(stats + kWorkBase + 2, (uint32_t)sh_visits);
  wave_add(stats + kWorkBase + 3, (uint32_t)sh_leaves);
}

// The tail with G lanes per path (triangle and batch scenes, 4- and 8-wide):
// each group carries one path through the remaining bounces -- its extension
// ray by GroupClosest (Walk8 on the group's leader for the rare near ties),
// shading on the leader, then the bounce's shadow ray by GroupWalk -- so the
// sample receives the wavefront's additions in its order (extension, then
// shadow, bounce by bounce) and images equal the per-lane tail's.  The last
// few thousand paths of a chunk walk long serial chains (over 100 node
// visits for the slowest extension and shadow rays); a group shortens each
// step's wait to one round of node loads and one of records.  Runs when the
// chunk's live paths lie in (lo, hi]; the per-lane tail takes the others.
template <int kAlg, int kGeo, int G>
__global__ void __launch_bounds__(kRtBlock) tail_group_kernel(SceneView sc, Queues q, uint64_t seed, uint32_t bounce,
                                                              uint32_t depth, uint32_t lo, uint32_t hi,
                                                              unsigned long long* __restrict__ stats) {
  constexpr int kGroups = kRtBlock / G;
  static_assert(GroupWalk<G>::kStack <= 2 * GroupClosest<G>::kStack, "the walks share one LDS stack area");
  __shared__ uint32_t s_stack[kGroups][2 * GroupClosest<G>::kStack];
  __shared__ uint32_t s_recs[kGroups][GroupWalk<G>::kRecs];
  const uint32_t tag = q.counters[kTailDone];
  if (tag != 0 && tag != bounce + 1u) return;
  const uint64_t cap = q.capacity;
  const uint32_t active = q.counters[kActive] < cap ? q.counters[kActive] : (uint32_t)cap;
  if (active <= lo || active > hi) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) q.counters[kTailDone] = bounce + 1u;
  const uint32_t lane = lane_id(), grp = threadIdx.x / G, gl = lane % G;
  const int lead = (int)((lane / G) * G);  // the group's leader lane in the wave
  const bool leader = gl == 0;
  uint32_t* stack = s_stack[grp];
  uint32_t* recs = s_recs[grp];
  const uint32_t groups = gridDim.x * kGroups;
  uint64_t ext_visits = 0, ext_leaves = 0, sh_visits = 0, sh_leaves = 0;
  // each wave takes 64 / G consecutive paths at a time: the first ones by its
  // place in the grid, then the next unclaimed ones from a cursor (claimed
  // while the current paths run), so waves that drew short paths take more
  for (uint32_t g0 = (blockIdx.x * kRtBlock + (threadIdx.x & ~63u)) / G; g0 < active;) {
    uint32_t claim = 0;
    if (lane == 0) claim = atomicAdd(q.counters + kFetchExt, 64u / G);
    const uint32_t idx = g0 + lane / G;
    bool alive = idx < active;
    uint32_t path_id = 0, slot = 0;
    bool delta = false;
    float f[kPathPlanes];
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int p = 0; p < kPathPlanes; ++p) f[p] = 0.0f;
    if (alive) {
      // the queue's back part (grazing paths, the long ones) first
      const uint32_t sidx = src_slot(active - 1u - idx, q.counters[kFront], cap);
      const uint32_t idw = q.src_ids[sidx];
      path_id = idw & kPathIdMask;
      delta = (idw & kPathDelta) != 0;
#pragma unroll
      for (int p = 0; p < kPathPlanes; ++p) f[p] = q.src_fields[plane_at<kPathPlanes>(sidx, p)];
      slot = path_id - q.sample_base;
      if (leader) {
#pragma unroll
        for (int p = 0; p < kSamplePlanes; ++p) acc[p] = q.sample[plane_at<kSamplePlanes>(slot, p)];
      }
    }
    WideView wv = sc.bvh.wide;
    uint32_t first = 0;
    if constexpr (kGeo == kGeoBatch) {
      const uint32_t env = path_id / sc.env_paths;
      wv.root = env;  // segment roots lead the node array
      first = alive ? (uint32_t)sc.seg_off[env] : 0u;
    }
    for (uint32_t b = bounce; b < depth; ++b) {
      if (b < kMaxBounces) wave_add(stats + 2 * b, (leader && alive) ? 1u : 0u);
      // the extension ray
      const Ray ray{mk(f[0], f[1], f[2]), mk(f[4], f[5], f[6]), f[3], f[7]};
      GroupClosest<G> gc;
      uint32_t ev = 0, et = 0;
      if (alive) gc.begin(wv, ray, stack);
      bool walking = alive;
      while (__ballot(walking) != 0) {
        if (walking && gc.step(wv, stack, recs, ev, et)) walking = false;
      }
      Hit h = miss(ray.tmax);
      if (alive && leader) {
        if (gc.near_tie()) {
          // Walk8 decides; its stack lives in the group's (now free) LDS area
          Stack fst{stack, nullptr, 0, 0, kMaxStack, 1};
          TraceCounters tc;
          h = traverse8<BAJO_TRACE_CLOSEST_HIT, true>(wv, ray, fst, &tc);
          ev += tc.visits;
          et += tc.leaves;
        } else {
          h = gc.finish(wv);
        }
        if constexpr (kGeo == kGeoBatch)
          if (h.prim != kEmpty) h.prim += first;
      }
      if (leader) {
        ext_visits += ev;
        ext_leaves += et;
      }
      if (b < kMaxBounces) wave_max(stats + kMaxVisitBase + 2 * b, leader ? ev : 0u);
      // shading on the leader
      bool cast = false, next = false;
      float sf[kShadowPlanes];
#pragma unroll
      for (int p = 0; p < kShadowPlanes; ++p) sf[p] = 0.0f;
      if (alive && leader) {
        const v3 T = mk(f[8], f[9], f[10]);
        const Surface su = surface_of(sc, ray, h);
        ShadowWork sw;
        if (shade_shadow<kAlg>(sc, seed, b, path_id, T, ray.d, su, sw)) {
          cast = true;
#pragma unroll
          for (int p = 0; p < kShadowPlanes; ++p) sf[p] = sw.f[p];
        }
        v3 add_c;
        NextPath np;
        const uint32_t sfl = shade_bounce<kAlg>(sc, seed, b, path_id, delta, ray, T, f[11], h.t, su, add_c, np);
        if (sfl & kShadeAdd) {
          acc[0] += add_c.x;
          acc[1] += add_c.y;
          acc[2] += add_c.z;
        }
        next = (sfl & kShadeNext) != 0;
        if (next) {
#pragma unroll
          for (int p = 0; p < kPathPlanes; ++p) f[p] = np.f[p];
          delta = (np.id & kPathDelta) != 0;
        }
      }
      if (b < kMaxBounces) wave_add(stats + 2 * b + 1, (leader && cast) ? 1u : 0u);
      // the bounce's shadow ray, traced by the whole group
      cast = __shfl(cast, lead, 64);
#pragma unroll
      for (int p = 0; p < kShadowPlanes; ++p) sf[p] = __shfl(sf[p], lead, 64);
      GroupWalk<G> gw;
      uint32_t sv = 0, stt = 0;
      bool occluded = false;
      if (cast) gw.begin(wv, Ray{mk(sf[0], sf[1], sf[2]), mk(sf[4], sf[5], sf[6]), sf[3], sf[7]}, stack);
      bool shadowing = cast;
      while (__ballot(shadowing) != 0) {
        if (shadowing) {
          bool occ = false;
          if (gw.step(wv, stack, recs, sv, stt, occ)) {
            shadowing = false;
            occluded = occ;
          }
        }
      }
      if (cast && leader && (kAlg == BAJO_RENDER_AO || !occluded)) {
        const v3 cs = shadow_radiance<kAlg>(occluded, sf);
        acc[0] += cs.x;
        acc[1] += cs.y;
        acc[2] += cs.z;
      }
      if (leader) {
        sh_visits += sv;
        sh_leaves += stt;
      }
      if (b < kMaxBounces) wave_max(stats + kMaxVisitBase + 2 * b + 1, leader ? sv : 0u);
      // the next segment, from the leader (a path still going after the last bounce ends)
      alive = __shfl(next, lead, 64) && b + 1 < depth;
#pragma unroll
      for (int p = 0; p < kPathPlanes; ++p) f[p] = __shfl(f[p], lead, 64);
      delta = __shfl(delta, lead, 64);
      if (__ballot(alive) == 0) break;
    }
    if (idx < active && leader) {
#pragma unroll
      for (int p = 0; p < kSamplePlanes; ++p) q.sample[plane_at<kSamplePlanes>(slot, p)] = acc[p];
    }
    g0 = groups + __shfl(claim, 0, 64);
  }
  wave_add(stats + kWorkBase + 0, (uint32_t)ext_visits);
  wave_add(stats + kWorkBase + 1, (uint32_t)ext_leaves);
  wave_add(stats + kWorkBase + 2, (uint32_t)sh_visits);
  wave_add(stats + kWorkBase + 3, (uint32_t)sh_leaves);
}

// kWalk: the shadow walk form (traverse8_any: bit 0 two nodes per step, bit 1
// leaf records four at a time).  One lane per ray, when the bounce's shadow
// count exceeds `sparse_rays`.  A grid of the resident blocks whose waves take
// 64 rays at a time from a cursor: a bounce that leaves the rays to another
// kernel costs only the resident blocks' exits, not one block per arena entry.
template <int kAlg, bool kWide, int kGeo, int kWalk = 0>
__global__ void __launch_bounds__(kRtBlock) __attribute__((amdgpu_waves_per_eu(kWalk == 2 ? 5 : 1)))
shadow_kernel(SceneView sc, Queues q, uint32_t bounce,
                                                          uint32_t* __restrict__ spill,
                                                          unsigned long long* __restrict__ work,
                                                          unsigned long long* __restrict__ stats,
                                                          uint32_t sparse_rays) {
  __shared__ uint32_t s_stack[kLdsStack * kRtBlock];
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Stack st{s_stack + threadIdx.x, spill + gtid, (uint64_t)gridDim.x * blockDim.x, 0};
  const uint64_t cap = q.capacity;
  const uint32_t count =
      q.counters[kTailDone] ? 0u : (q.counters[kShadow] < cap ? q.counters[kShadow] : (uint32_t)cap);
  if (count <= sparse_rays) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) q.counters[kSparseShadow] = gridDim.x;  // blocks of partials
  uint32_t work_visits = 0, work_leaves = 0, ray_max = 0;  // a lane traces a few rays: 32 bits suffice
  const uint32_t threads = gridDim.x * blockDim.x;
  for (uint32_t idx = (uint32_t)gtid; idx < count; idx += threads) {
    {
      // the ray's planes now, its radiance planes (8..10) and id after the walk
      float f[kShadowPlanes];
#pragma unroll
      for (int p = 0; p < 8; ++p) f[p] = q.shadow_fields[plane_at<kShadowPlanes>(idx, p)];
      uint32_t env = 0;
      if constexpr (kGeo == kGeoBatch) env = q.shadow_ids[idx] / sc.env_paths;
      const Ray ray{mk(f[0], f[1], f[2]), mk(f[4], f[5], f[6]), f[3], f[7]};
      TraceCounters tc;
      const Hit h = scene_trace<kTraceShadow + kWalk, kWide, kGeo>(sc, ray, st, &tc, false, env);
      work_visits += tc.visits;
      work_leaves += tc.leaves;
      ray_max = tc.visits > ray_max ? tc.visits : ray_max;
      const bool occluded = h.prim == kEmpty && h.t == 0.0f;
      if (kAlg == BAJO_RENDER_AO || !occluded) {
#pragma unroll
        for (int p = 8; p < kShadowPlanes; ++p) f[p] = q.shadow_fields[plane_at<kShadowPlanes>(idx, p)];
        const uint32_t slot = q.shadow_ids[idx] - q.sample_base;
        const v3 c = shadow_radiance<kAlg>(occluded, f);
        accumulate(q.sample, slot, c);
        if constexpr (kAlg == BAJO_RENDER_GUIDED)
          if (q.contrib) contrib_add(q.contrib, slot, q.depth, bounce, c);
      }
    }
  }
  store_block_work(work, work_visits, work_leaves, ray_max);
}

// Shadow rays with G lanes per ray (GroupWalk) when the bounce's shadow count
// lies in (lo, hi] -- decided on the device from the queue length, so exactly
// one shadow launch of a bounce traces its rays.  After the first bounce only
// a few thousand shadow rays remain, and the per-lane kernel's time is the
// serial chain of its slowest lane (over 100 node visits): there a whole
// wave (G = 64) carries each ray.  Each group takes the next ray as soon as
// it finishes one.  Triangle and batch scenes at widths 4 and 8.
template <int kAlg, int kGeo, int G>
__global__ void __launch_bounds__(kRtBlock) shadow_group_kernel(SceneView sc, Queues q, uint32_t bounce, uint32_t lo,
                                                                uint32_t hi, unsigned long long* __restrict__ work) {
  constexpr int kGroups = kRtBlock / G;
  __shared__ uint32_t s_stack[kGroups][GroupWalk<G>::kStack];
  __shared__ uint32_t s_recs[kGroups][GroupWalk<G>::kRecs];
  const uint64_t cap = q.capacity;
  const uint32_t count =
      q.counters[kTailDone] ? 0u : (q.counters[kShadow] < cap ? q.counters[kShadow] : (uint32_t)cap);
  if (count <= lo || count > hi) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) q.counters[kSparseShadow] = gridDim.x;
  const uint32_t grp = threadIdx.x / G, gl = lane_id() % G;
  uint32_t* stack = s_stack[grp];
  uint32_t* recs = s_recs[grp];
  const uint32_t groups = gridDim.x * kGroups;
  uint32_t idx = blockIdx.x * kGroups + grp;
  uint32_t visits = 0, tests = 0, ray_max = 0, rv = 0;
  GroupWalk<G> walk;
  WideView wv = sc.bvh.wide;
  float f[kShadowPlanes];
  uint32_t slot = 0;
  bool have = false;
  for (;;) {
    if (!have && idx < count) {
#pragma unroll
      for (int p = 0; p < kShadowPlanes; ++p) f[p] = q.shadow_fields[plane_at<kShadowPlanes>(idx, p)];
      const uint32_t sid = q.shadow_ids[idx];
      slot = sid - q.sample_base;
      if constexpr (kGeo == kGeoBatch) wv.root = sid / sc.env_paths;  // segment roots lead the node array
      walk.begin(wv, Ray{mk(f[0], f[1], f[2]), mk(f[4], f[5], f[6]), f[3], f[7]}, stack);
      rv = 0;
      have = true;
    }
    if (__ballot(have) == 0) break;  // every group of the wave is out of rays
    if (have) {
      bool occluded = false;
      if (walk.step(wv, stack, recs, rv, tests, occluded)) {
        if (gl == 0 && (kAlg == BAJO_RENDER_AO || !occluded)) {
          const v3 c = shadow_radiance<kAlg>(occluded, f);
          accumulate(q.sample, slot, c);
          if constexpr (kAlg == BAJO_RENDER_GUIDED)
            if (q.contrib) contrib_add(q.contrib, slot, q.depth, bounce, c);
        }
        visits += rv;
        ray_max = rv > ray_max ? rv : ray_max;
        have = false;
        idx += groups;
      }
    }
  }
  store_block_work(work, gl == 0 ? visits : 0u, gl == 0 ? tests : 0u, ray_max);
}

// The continuation of the guided vertices of one bounce: one-sample MIS of
// the BSDF and the guide's mixture, each chosen with probability 1/2 (the
// guide stage's first number), weighted by f cos / (pdf_bsdf + pdf_guide) / 2.
// The BSDF half draws exactly what the plain renderer draws (same stream,
// same arithmetic); the guide half draws lobe, then direction.  With
// `records` the vertex and the throughput it passes on are kept for the
// frame's training samples.
__global__ void __launch_bounds__(kRtBlock) guide_sample_kernel(SceneView sc, Queues q, uint64_t seed, uint32_t bounce,
                                                                const float* __restrict__ outputs,
                                                                float* __restrict__ records) {
  const uint64_t cap = q.capacity;
  const uint32_t count = q.counters[kVertex] < cap ? q.counters[kVertex] : (uint32_t)cap;
  if ((uint64_t)blockIdx.x * kRtBlock >= count) return;
  const uint32_t idx = blockIdx.x * kRtBlock + threadIdx.x;
  bool want_next = false;
  NextPath np;
  if (idx < count) {
    const float* V = q.vertices;
    const v3 p = mk(V[vplane(idx, 0)], V[vplane(idx, 1)], V[vplane(idx, 2)]);
    const v3 n = mk(V[vplane(idx, 3)], V[vplane(idx, 4)], V[vplane(idx, 5)]);
    const v3 d = mk(V[vplane(idx, 6)], V[vplane(idx, 7)], V[vplane(idx, 8)]);
    const v3 T = mk(V[vplane(idx, 9)], V[vplane(idx, 10)], V[vplane(idx, 11)]);
    const uint32_t surf = __builtin_bit_cast(uint32_t, V[vplane(idx, 12)]);
    const uint32_t path_id = __builtin_bit_cast(uint32_t, V[vplane(idx, 13)]);
    const uint32_t kind = surf >> 28, mi = surf & kSurfaceIndexMask;
    GuideMixture m;
    guide_decode(outputs + (uint64_t)idx * kGuideOut, m);
    StageRng gr = path_rng(sc, seed, path_id, rng_guide_stage(bounce));
    const float u_choose = gr.next(), u_lobe = gr.next(), u1 = gr.next(), u2 = gr.next();
    v3 dir = n, fcos = mk(0.0f, 0.0f, 0.0f);
    float pdf_b = 0.0f;
    bool ok = false;
    const bool ggx = kind == kMatGgx;
    const float* g4 = sc.ggx + 4 * (uint64_t)mi;
    const float* alb = sc.lambertian + 3 * (uint64_t)mi;
    if (u_choose < 1.0f - kGuideMix) {
      StageRng rng = path_rng(sc, seed, path_id, rng_bsdf_stage(bounce));
      const float ru = rng.next(), rv = rng.next();
      if (ggx) {
        const BsdfSample smp = sample_ggx(d, n, mk(g4[0], g4[1], g4[2]), g4[3], ru, rv);
        ok = smp.ok;
        dir = smp.direction;
        pdf_b = smp.pdf;
        fcos = scale(smp.weight, smp.pdf);
      } else {
        // _sample_lambertian (rt/shading.mojo:92-118), as the plain renderer draws it
        float s, c;
        sincos_turns(ru, s, c);
        const float z = 1.0f - 2.0f * rv;
        const float radius = sqrtf(fmax_sel(1.0f - z * z, 0.0f));
        v3 sd = add(n, mk(radius * c, radius * s, z));
        if (fabsf(sd.x) < 1e-8f && fabsf(sd.y) < 1e-8f && fabsf(sd.z) < 1e-8f) sd = n;
        dir = normalize(sd);
        pdf_b = fmax_sel(dot(n, normalize(dir)), 0.0f) * kInvPi;
        fcos = scale(mk(alb[0], alb[1], alb[2]), pdf_b);
        ok = pdf_b > 0.0f;
      }
    } else {
      dir = guide_sample(m, outputs + (uint64_t)idx * kGuideOut, u_lobe, u1, u2);
      const float cos_i = dot(n, dir);
      if (ggx) {
        const BsdfEval ev = eval_ggx(d, n, mk(g4[0], g4[1], g4[2]), g4[3], dir);
        pdf_b = ev.pdf;
        fcos = scale(ev.value, fmax_sel(cos_i, 0.0f));
        ok = ev.pdf > 0.0f && cos_i > 0.0f;
      } else {
        pdf_b = fmax_sel(cos_i, 0.0f) * kInvPi;
        fcos = scale(mk(alb[0], alb[1], alb[2]), pdf_b);
        ok = cos_i > 0.0f;
      }
    }
    const float pmix = guide_mix_pdf(pdf_b, guide_pdf(m, dir));
    v3 Tn = mk(0.0f, 0.0f, 0.0f);
    if (ok && pmix > 0.0f) {
      Tn = mul(T, scale(fcos, 1.0f / pmix));
      want_next = survives_roulette(sc, seed, path_id, bounce, Tn);
    }
    if (want_next) {
      np.id = path_id;
      const float nf[kPathPlanes] = {p.x, p.y, p.z, kT_MinSecondary, dir.x, dir.y, dir.z, kFltMax,
                                     Tn.x, Tn.y, Tn.z, pmix};
#pragma unroll
      for (int k = 0; k < kPathPlanes; ++k) np.f[k] = nf[k];
    }
    if (records) {
      const v3 wo = normalize(mk(-d.x, -d.y, -d.z));
      const float rf[kRecordPlanes] = {p.x, p.y, p.z, n.x, n.y, n.z, wo.x, wo.y, wo.z, dir.x, dir.y, dir.z,
                                       want_next ? pmix : 0.0f, Tn.x, Tn.y, Tn.z};
      float* r = records + ((uint64_t)(path_id - q.sample_base) * q.depth + bounce) * kRecordPlanes;
#pragma unroll
      for (int k = 0; k < kRecordPlanes; ++k) r[k] = rf[k];
    }
  }
  const uint32_t ns = reserve_next(q, want_next, false);
  if (want_next) {
    if (ns < cap) {
      q.dst_ids[ns] = np.id;
#pragma unroll
      for (int k = 0; k < kPathPlanes; ++k) q.dst_fields[plane_at<kPathPlanes>(ns, k)] = np.f[k];
    } else {
      atomicMax(q.counters + kStatus, kStatusPathOverflow);
    }
  }
}

// Per-pixel average of its spp sample slots, summed in slot order.  A block's
// pixels own one contiguous range of slots, which it stages through LDS in
// chunks with coalesced loads; each thread then adds its own slots from LDS.
constexpr uint32_t kResolveChunk = 2048;

__global__ void __launch_bounds__(kRtBlock) resolve_kernel(const float* __restrict__ sample, float* __restrict__ pixels,
                                                           uint64_t cap, uint32_t pixel_base, uint32_t pixel_count,
                                                           uint32_t spp) {
  __shared__ float s_buf[kSamplePlanes][kResolveChunk + kResolveChunk / 32];
  const uint32_t p0 = blockIdx.x * blockDim.x;
  const uint32_t p = p0 + threadIdx.x;
  const uint32_t npix = pixel_count - p0 < blockDim.x ? pixel_count - p0 : blockDim.x;
  const uint64_t base = (uint64_t)p0 * spp, total = (uint64_t)npix * spp;
  const uint64_t mine_lo = (uint64_t)threadIdx.x * spp, mine_hi = mine_lo + spp;
  float acc[kSamplePlanes] = {0.0f, 0.0f, 0.0f};
  for (uint64_t c = 0; c < total; c += kResolveChunk) {
    const uint32_t len = total - c < kResolveChunk ? (uint32_t)(total - c) : kResolveChunk;
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
#pragma unroll
      for (int pl = 0; pl < kSamplePlanes; ++pl)
        s_buf[pl][i + i / 32] = sample[plane_at<kSamplePlanes>((uint32_t)(base + c + i), pl)];
    __syncthreads();
    const uint64_t lo = mine_lo > c ? mine_lo : c, hi = mine_hi < c + len ? mine_hi : c + len;
    for (uint64_t s2 = lo; s2 < hi; ++s2) {
      const uint32_t j = (uint32_t)(s2 - c);
#pragma unroll
      for (int pl = 0; pl < kSamplePlanes; ++pl) acc[pl] += s_buf[pl][j + j / 32];
    }
    __syncthreads();
  }
  if (p >= pixel_count) return;
  const float k = 1.0f / (float)spp;
  const uint64_t o = 3ull * (pixel_base + p);
  pixels[o + 0] = acc[0] * k;
  pixels[o + 1] = acc[1] * k;
  pixels[o + 2] = acc[2] * k;
}

}  // namespace

// ---- scene + arena ---------------------------------------------------------
// Path queues, sample planes, counters and the frame's camera of one render
// (a scene keeps one for bajo_render; a render target owns its own).
struct RenderArena {
  uint64_t cap = 0;
  uint32_t* ids[2] = {nullptr, nullptr};
  float* fields[2] = {nullptr, nullptr};
  uint32_t* shadow_ids = nullptr;
  float* shadow_fields = nullptr;
  float* sample = nullptr;
  uint32_t* counters = nullptr;
  unsigned long long* stats = nullptr;
  uint32_t* spill = nullptr;
  uint64_t spill_threads = 0;
  unsigned long long* work_ext = nullptr;  // per-block traversal work partials
  unsigned long long* work_sh = nullptr;
  unsigned long long* work_sums = nullptr;  // kSumBlocks rows of 6 sub-totals of the partials
  float* cam = nullptr;                    // 20 floats
  float* cams = nullptr;                   // 20 floats per environment (per-environment cameras)
  uint32_t cams_cap = 0;
};

struct bajo_scene {
  int device = 0;
  uint64_t n = 0;
  float* vertices = nullptr;
  bajo_bvh* bvh = nullptr;
  bajo_build_options options{};
  uint32_t* surfaces = nullptr;
  float* lambertian = nullptr;
  float* metals = nullptr;
  float* dielectrics = nullptr;
  float* ggx = nullptr;
  float* emissive = nullptr;
  float* light_fields = nullptr;
  uint32_t* light_kinds = nullptr;
  uint32_t light_count = 0;
  float total_light_weight = 0.0f;
  RenderArena arena;  // reused by bajo_render
  // spheres of a mixed scene: center + |r| (the BVH's input), signed radii,
  // surface ids, and their BVH
  uint64_t n_spheres = 0;
  float* spheres = nullptr;
  float* sphere_radii = nullptr;
  uint32_t* sphere_surfaces = nullptr;
  bajo_bvh* sphere_bvh = nullptr;
  // a batch of environments: the BLAS set (segment e = environment e), the
  // host offsets u64[S + 1], the environments' light ranges u32[S + 1] and
  // total light weights f32[S]
  uint32_t env_count = 0;
  bajo_blas_set* set = nullptr;
  std::vector<uint64_t> env_offsets;
  uint32_t* env_light_base = nullptr;
  float* env_light_total = nullptr;
};

namespace {

void free_arena(RenderArena& a) {
  (void)hipFree(a.ids[0]); (void)hipFree(a.ids[1]); (void)hipFree(a.fields[0]); (void)hipFree(a.fields[1]);
  (void)hipFree(a.shadow_ids); (void)hipFree(a.shadow_fields); (void)hipFree(a.sample); (void)hipFree(a.counters);
  (void)hipFree(a.stats); (void)hipFree(a.spill); (void)hipFree(a.work_ext); (void)hipFree(a.work_sh);
  (void)hipFree(a.work_sums);
  (void)hipFree(a.cam); (void)hipFree(a.cams);
  a = RenderArena();
}

void free_scene(bajo_scene* s) {
  if (!s) return;
  free_arena(s->arena);
  bajo_bvh_free(s->bvh);
  if (s->set) bajo_blas_set_free(s->set);
  (void)hipFree(s->env_light_base); (void)hipFree(s->env_light_total);
  bajo_bvh_free(s->sphere_bvh);
  (void)hipFree(s->spheres); (void)hipFree(s->sphere_radii); (void)hipFree(s->sphere_surfaces);
  (void)hipFree(s->vertices); (void)hipFree(s->surfaces); (void)hipFree(s->lambertian); (void)hipFree(s->emissive);
  (void)hipFree(s->metals); (void)hipFree(s->dielectrics); (void)hipFree(s->ggx);
  (void)hipFree(s->light_fields); (void)hipFree(s->light_kinds);
  delete s;
}

// A view with no hierarchy: every query misses at once.
AccelView empty_view(uint32_t* status) {
  AccelView av;
  av.bin = BvhView{nullptr, nullptr, kSentinel, 0u, 0u, status};
  av.wide = WideView{nullptr, nullptr, 0u, 0u, status, 0u};
  return av;
}

template <typename K>
int32_t grid_of(K kernel, unsigned* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  BJ_HIP(hipGetDevice(&dev));
  BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kRtBlock, 0));
  *grid = (unsigned)(cus * (per_cu < 1 ? 1 : per_cu));
  return BAJO_OK;
}

int32_t ensure_arena(RenderArena& a, uint64_t cap, uint64_t spill_threads) {
  if (cap <= a.cap && spill_threads <= a.spill_threads) return BAJO_OK;
  // keep the frame's camera across a regrowth
  float* cam = a.cam;
  a.cam = nullptr;
  free_arena(a);
  a.cam = cam;
  if (!a.cam) BJ_HIP(hipMalloc(&a.cam, 20 * sizeof(float)));
  const uint64_t tiles = (cap + 63) / 64 * 64;  // plane storage is whole 64-entry tiles
  for (int k = 0; k < 2; ++k) {
    BJ_HIP(hipMalloc(&a.ids[k], cap * sizeof(uint32_t)));
    BJ_HIP(hipMalloc(&a.fields[k], tiles * kPathPlanes * sizeof(float)));
  }
  BJ_HIP(hipMalloc(&a.shadow_ids, cap * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&a.shadow_fields, tiles * kShadowPlanes * sizeof(float)));
  BJ_HIP(hipMalloc(&a.sample, tiles * kSamplePlanes * sizeof(float)));
  BJ_HIP(hipMalloc(&a.counters, kCounterWords * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&a.stats, kStatsWords * sizeof(unsigned long long)));
  BJ_HIP(hipMalloc(&a.spill, spill_threads * kSpillMax * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&a.work_ext, kWorkStride * (spill_threads / kRtBlock) * sizeof(unsigned long long)));
  BJ_HIP(hipMalloc(&a.work_sh, kWorkStride * (spill_threads / kRtBlock) * sizeof(unsigned long long)));
  BJ_HIP(hipMalloc(&a.work_sums, 6 * kSumBlocks * sizeof(unsigned long long)));
  a.cap = cap;
  a.spill_threads = spill_threads;
  return BAJO_OK;
}

// The frame's camera, written by a kernel on the stream (so a captured graph
// replays with whatever camera the buffer holds).
__global__ void set_camera_kernel(float* __restrict__ cam, Cam c) {
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int k = 0; k < 20; ++k) cam[k] = c.v[k];
}

// LightStore.build_alias_table (rt/types.mojo:207-247): Walker-Vose with
// LIFO small/large work lists, in FP32.
void build_alias(const std::vector<float>& w, float total, std::vector<float>& prob, std::vector<uint32_t>& alias) {
  const size_t n = w.size();
  prob.assign(n, 1.0f);
  alias.assign(n, 0u);
  if (n == 0 || !(total > 0.0f)) return;
  std::vector<float> scaled(n);
  std::vector<size_t> small, large;
  for (size_t i = 0; i < n; ++i) {
    alias[i] = (uint32_t)i;
    scaled[i] = w[i] * (float)n / total;
    if (scaled[i] < 1.0f) small.push_back(i);
    else large.push_back(i);
  }
  while (!small.empty() && !large.empty()) {
    size_t si = small.back(); small.pop_back();
    size_t li = large.back(); large.pop_back();
    prob[si] = scaled[si];
    alias[si] = (uint32_t)li;
    scaled[li] = scaled[li] + scaled[si] - 1.0f;
    if (scaled[li] < 1.0f) small.push_back(li);
    else large.push_back(li);
  }
  while (!small.empty()) { size_t i = small.back(); small.pop_back(); prob[i] = 1.0f; alias[i] = (uint32_t)i; }
  while (!large.empty()) { size_t i = large.back(); large.pop_back(); prob[i] = 1.0f; alias[i] = (uint32_t)i; }
}

struct LaunchEvents {
  std::vector<hipEvent_t> ext, sh;  // start/end pairs
  std::vector<hipEvent_t> inf;      // guided renders: network inference start/end pairs
  hipEvent_t begin = nullptr, end = nullptr;
  bool on = false;
  hipEvent_t make() {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  ~LaunchEvents() {
    for (auto e : ext) (void)hipEventDestroy(e);
    for (auto e : sh) (void)hipEventDestroy(e);
    for (auto e : inf) (void)hipEventDestroy(e);
    if (begin) (void)hipEventDestroy(begin);
    if (end) (void)hipEventDestroy(end);
  }
};

// Arena shape of a render: path capacity (whole pixels) and the threads of
// the widest launch (their spill columns and work partials).
struct RenderShape {
  uint64_t cap = 0;
  unsigned g_ext = 0, g_tail = 0, simds = 0;
  uint64_t spill_threads = 0;
};

// Paths of one render: every environment of a batch renders the full frame.
uint64_t render_paths(const bajo_scene* s, const bajo_render_settings* st) {
  return (uint64_t)st->image_width * st->image_height * st->samples_per_pixel * (s->env_count ? s->env_count : 1u);
}

template <int kAlg, bool kWide, int kGeo>
int32_t render_shape(const bajo_scene* s, const bajo_render_settings* st, RenderShape* out) {
  const uint64_t total = render_paths(s, st);
  BJ_REQUIRE(total < (1ull << 31), "bajo_render: %llu paths exceed the 31-bit path id", (unsigned long long)total);
  uint64_t cap = st->path_capacity ? st->path_capacity : total;
  if (cap > total) cap = total;
  cap = (cap / st->samples_per_pixel) * st->samples_per_pixel;
  if (cap == 0) cap = st->samples_per_pixel;
  constexpr bool kTail = kAlg == BAJO_RENDER_PATH || kAlg == BAJO_RENDER_NEE || kAlg == BAJO_RENDER_MIS;
  unsigned g_tail = 0, simds = 0;
  if constexpr (kTail) {
    int32_t rc = grid_of(tail_kernel<kAlg, kWide, kGeo>, &g_tail);
    if (rc) return rc;
    int dev = 0, cus = 0;
    BJ_HIP(hipGetDevice(&dev));
    BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    simds = 4u * (unsigned)cus;
  }
  // extend and shadow launch one thread per arena entry
  const unsigned g_ext = div_up(cap, kRtBlock);
  if (g_tail < 1) g_tail = 1;
  const unsigned g_max = g_ext > g_tail ? g_ext : g_tail;
  out->cap = cap;
  out->g_ext = g_ext;
  out->g_tail = g_tail;
  out->simds = simds;
  out->spill_threads = (uint64_t)g_max * kRtBlock;
  return BAJO_OK;
}

// A guided render's network and whether the frame records training data.
struct GuideRun {
  bajo_guide* g = nullptr;
  bool train = false;
  const float* bounds = nullptr;  // scene box lo.xyz, hi.xyz (device)
};

// Enqueues one render into arena `a` (no host synchronization without `ev`).
// `camera` (host, 20 floats) is written into the arena's camera buffer first;
// NULL renders with the camera the buffer already holds.  Guided renders
// (kAlg = BAJO_RENDER_GUIDED) take the guide in `gr`.
template <int kAlg, bool kWide, int kGeo>
int32_t run_render(bajo_ctx* ctx, bajo_scene* s, RenderArena& a, const bajo_render_settings* st, const float* camera,
                   uint32_t camera_count, float* pixels, LaunchEvents* ev, const GuideRun* gr = nullptr) {
  const uint64_t total = render_paths(s, st);
  RenderShape shape;
  int32_t rc = render_shape<kAlg, kWide, kGeo>(s, st, &shape);
  if (rc) return rc;
  constexpr bool kTail = kAlg == BAJO_RENDER_PATH || kAlg == BAJO_RENDER_NEE || kAlg == BAJO_RENDER_MIS;
  const uint64_t cap = shape.cap;
  const unsigned g_ext = shape.g_ext, g_sh = g_ext, g_tail = shape.g_tail, simds = shape.simds;
  // the tail takes over when the remaining paths fit one pass of its resident
  // threads, or at the schedule's own threshold
  const uint32_t tail_threshold = st->tail_paths ? st->tail_paths : g_tail * kRtBlock;
  BJ_REQUIRE(st->extend_waves == 0 || st->extend_waves == 4 || st->extend_waves == 6 || st->extend_waves == 8,
             "bajo_render: extend_waves must be 0, 4, 6 or 8 (got %u)", st->extend_waves);
  // extension kernels compiled for other occupancies (the hot renders only)
  constexpr bool kTunable = kWide && kGeo != kGeoMixed &&
                            (kAlg == BAJO_RENDER_PATH || kAlg == BAJO_RENDER_NEE || kAlg == BAJO_RENDER_MIS);
  const uint32_t ext_waves = kTunable ? st->extend_waves : 0u;
  BJ_REQUIRE(st->shadow_walk <= 4, "bajo_render: shadow_walk must be 0..4 (got %u)", st->shadow_walk);
  // shadow walk form: 1..4 as given, 0 = the library's default
  const uint32_t shadow_walk = (st->shadow_walk ? st->shadow_walk : kDefaultShadowWalk) - 1u;
  // one ray per wave for bounces with few shadow rays: sparse_rays (0 = the
  // default, UINT32_MAX never), a grid of at most that many waves
  const uint32_t sparse_rays = st->sparse_shadow_rays == 0 ? kSparseShadowRays
                               : (st->sparse_shadow_rays == 0xFFFFFFFFu ? 0u : st->sparse_shadow_rays);
  const unsigned g_wave_cap = div_up(sparse_rays ? sparse_rays : 1u, kRtBlock / 64);
  const unsigned g_wave = g_wave_cap < g_sh ? g_wave_cap : g_sh;
  // lanes per shadow ray above the sparse threshold (1 = the per-lane walk)
  BJ_REQUIRE(st->shadow_group == 0 || st->shadow_group == 1 || st->shadow_group == 8 || st->shadow_group == 16 ||
                 st->shadow_group == 32,
             "bajo_render: shadow_group must be 0, 1, 8, 16 or 32 (got %u)", st->shadow_group);
  // the tail with 16, 32 or 64 lanes per path (triangle and batch scenes, 4- and 8-wide)
  BJ_REQUIRE(st->tail_group == 0 || st->tail_group == 1 || st->tail_group == 16 || st->tail_group == 32 ||
                 st->tail_group == 64,
             "bajo_render: tail_group must be 0, 1, 16, 32 or 64 (got %u)", st->tail_group);
  constexpr bool kGroupTail = kWide && kGeo != kGeoMixed;
  const uint32_t tail_group = st->tail_group ? st->tail_group : kDefaultTailGroup;
  uint32_t group_tail_hi = 0;
  if (kGroupTail && (tail_group == 16 || tail_group == 32 || tail_group == 64))
    group_tail_hi = st->tail_group ? tail_threshold : (kGroupTailPaths < tail_threshold ? kGroupTailPaths : tail_threshold);
  const bool adaptive_group = st->shadow_group == 0;
  const uint32_t shadow_group = adaptive_group ? kDefaultShadowGroup : st->shadow_group;
  // the lane-group shadow kernel loops over its rays with a fixed stride, so
  // its grid is exactly the blocks that are resident at once: a block beyond
  // them would start only when a resident one has finished all of its rays,
  // and run its own share in a second, nearly empty pass (0.86 ms against
  // 0.72 for the 4M-triangle NEE frame's first bounce)
  constexpr bool kSparseGeo = kWide && kGeo != kGeoMixed;
  constexpr int kGroupGeo = kSparseGeo ? kGeo : kGeoTriangles;
  unsigned g_group = 1, g_tgroup = 1;
  if (kSparseGeo && kAlg != BAJO_RENDER_PATH && kAlg != BAJO_RENDER_NORMALS) {
    if (shadow_group == 8) rc = grid_of(shadow_group_kernel<kAlg, kGroupGeo, 8>, &g_group);
    else if (shadow_group == 16) rc = grid_of(shadow_group_kernel<kAlg, kGroupGeo, 16>, &g_group);
    else if (shadow_group == 32) rc = grid_of(shadow_group_kernel<kAlg, kGroupGeo, 32>, &g_group);
    if (rc) return rc;
  }
  if (g_group > g_sh) g_group = g_sh;
  if (g_group < 1) g_group = 1;
  // the per-lane shadow kernel's resident grid, for the walk form this render uses
  unsigned g_lane = 1;
  if (kAlg != BAJO_RENDER_PATH && kAlg != BAJO_RENDER_NORMALS) {
    if (kWide && shadow_walk == 1) rc = grid_of(shadow_kernel<kAlg, kWide, kGeo, kWide ? 1 : 0>, &g_lane);
    else if (kWide && shadow_walk == 2) rc = grid_of(shadow_kernel<kAlg, kWide, kGeo, kWide ? 2 : 0>, &g_lane);
    else if (kWide && shadow_walk == 3) rc = grid_of(shadow_kernel<kAlg, kWide, kGeo, kWide ? 3 : 0>, &g_lane);
    else rc = grid_of(shadow_kernel<kAlg, kWide, kGeo>, &g_lane);
    if (rc) return rc;
  }
  if (g_lane > g_sh) g_lane = g_sh;
  if (g_lane < 1) g_lane = 1;
  // The lane-group tail claims paths from a cursor after its first ones; its
  // grid is four 256-thread blocks per CU, twice the resident count (0.35 ms
  // for the 4M-triangle NEE frame's tail against 0.38 with the resident grid).
  if constexpr (kGroupTail) {
    int dev_id = 0, cu_count = 0;
    BJ_HIP(hipGetDevice(&dev_id));
    BJ_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev_id));
    g_tgroup = 4u * (unsigned)(cu_count < 1 ? 1 : cu_count);
  }
  rc = ensure_arena(a, cap, shape.spill_threads);
  if (rc) return rc;
  hipStream_t stream = ctx->stream;
  PrimaryArgs pa;
  pa.cam = a.cam;
  pa.cam_stride = 0;
  if (camera && camera_count > 1) {
    // one camera per environment of a batch
    BJ_REQUIRE(camera_count == s->env_count, "bajo_render: %u cameras for %u environments", camera_count,
               s->env_count);
    if (camera_count > a.cams_cap) {
      (void)hipFree(a.cams);
      a.cams = nullptr;
      a.cams_cap = 0;
      BJ_HIP(hipMalloc(&a.cams, (size_t)camera_count * 20 * sizeof(float)));
      a.cams_cap = camera_count;
    }
    BJ_HIP(hipMemcpyAsync(a.cams, camera, (size_t)camera_count * 20 * sizeof(float), hipMemcpyHostToDevice, stream));
    pa.cam = a.cams;
    pa.cam_stride = 20;
  } else if (camera) {
    Cam c;
    memcpy(c.v, camera, sizeof(c.v));
    set_camera_kernel<<<1, 64, 0, stream>>>(a.cam, c);
    BJ_HIP(hipGetLastError());
  }
  pa.width = st->image_width;
  pa.height = st->image_height;
  pa.spp = st->samples_per_pixel;
  pa.env_paths = s->env_count ? st->image_width * st->image_height * st->samples_per_pixel : 0u;
  SceneView sv;
  const bajo_bvh* tri = s->set ? blas_set_store(s->set) : s->bvh;
  sv.bvh = tri ? accel_view(tri, a.counters + 8) : empty_view(a.counters + 8);
  sv.spheres = s->sphere_bvh ? accel_view(s->sphere_bvh, a.counters + 8) : empty_view(a.counters + 8);
  sv.sphere_surfaces = s->sphere_surfaces;
  sv.sphere_radii = s->sphere_radii;
  sv.surfaces = s->surfaces;
  sv.lambertian = s->lambertian;
  sv.metals = s->metals;
  sv.dielectrics = s->dielectrics;
  sv.ggx = s->ggx;
  sv.emissive = s->emissive;
  sv.light_fields = s->light_fields;
  sv.light_kinds = s->light_kinds;
  sv.light_count = s->light_count;
  sv.total_light_weight = s->total_light_weight;
  sv.seg_off = s->set ? tri->seg_off : nullptr;
  sv.env_light_base = s->env_light_base;
  sv.env_light_total = s->env_light_total;
  sv.env_paths = pa.env_paths;
  const uint32_t depth = (kAlg == BAJO_RENDER_NORMALS || kAlg == BAJO_RENDER_AO) ? 1u : st->max_depth;
  for (uint64_t begin = 0; begin < total; begin += cap) {
    const uint32_t chunk = (uint32_t)((total - begin) < cap ? (total - begin) : cap);
    init_counters_kernel<<<1, 64, 0, stream>>>(a.counters, chunk, a.stats, begin == 0);
    BJ_HIP(hipGetLastError());
    for (uint32_t b = 0; b < depth; ++b) {
      Queues q{a.ids[b & 1], a.fields[b & 1], a.ids[(b + 1) & 1], a.fields[(b + 1) & 1], a.shadow_ids,
               a.shadow_fields, a.sample, a.counters, cap, (uint32_t)begin, nullptr, nullptr, 0u};
      if constexpr (kAlg == BAJO_RENDER_GUIDED) {
        q.vertices = gr->g->vertices;
        q.contrib = gr->train ? gr->g->contrib : nullptr;
        q.depth = depth;
      }
      if (ev) { ev->ext.push_back(ev->make()); (void)hipEventRecord(ev->ext.back(), stream); }
      if constexpr (kTail) {
        if (b > 0) {
          // lane groups up to group_tail_hi live paths, one lane per path above
          uint32_t lo = 0;
          if constexpr (kGroupTail) {
            constexpr int kTG = kGroupTail ? kGeo : kGeoTriangles;
            if (group_tail_hi && tail_group == 32)
              tail_group_kernel<kAlg, kTG, 32><<<g_tgroup, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, depth, 0u,
                                                                                 group_tail_hi, a.stats);
            else if (group_tail_hi && tail_group == 64)
              tail_group_kernel<kAlg, kTG, 64><<<g_tgroup, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, depth, 0u,
                                                                                 group_tail_hi, a.stats);
            else if (group_tail_hi)
              tail_group_kernel<kAlg, kTG, 16><<<g_tgroup, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, depth, 0u,
                                                                                 group_tail_hi, a.stats);
            if (group_tail_hi) lo = group_tail_hi;
          }
          tail_kernel<kAlg, kWide, kGeo><<<g_tail, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, depth, lo,
                                                                    tail_threshold, simds, a.stats, a.spill);
          BJ_HIP(hipGetLastError());
        }
      }
      if (b == 0)
        extend_kernel<kAlg, kWide, true, kGeo><<<div_up(chunk, kRtBlock), kRtBlock, 0, stream>>>(
            sv, q, st->rng_seed, b, pa, a.spill, a.work_ext, a.stats);
      else if (kTunable && ext_waves == 4)
        extend_kernel<kAlg, kWide, false, kGeo, kTunable ? 4 : 6><<<g_ext, kRtBlock, 0, stream>>>(
            sv, q, st->rng_seed, b, pa, a.spill, a.work_ext, a.stats);
      else if (kTunable && ext_waves == 8)
        extend_kernel<kAlg, kWide, false, kGeo, kTunable ? 8 : 6><<<g_ext, kRtBlock, 0, stream>>>(
            sv, q, st->rng_seed, b, pa, a.spill, a.work_ext, a.stats);
      else
        extend_kernel<kAlg, kWide, false, kGeo><<<g_ext, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, pa, a.spill,
                                                                          a.work_ext, a.stats);
      BJ_HIP(hipGetLastError());
      if (ev) { ev->ext.push_back(ev->make()); (void)hipEventRecord(ev->ext.back(), stream); }
      if constexpr (kAlg == BAJO_RENDER_GUIDED) {
        // the network over this bounce's guided vertices, then their
        // continuations (none after the last bounce)
        if (b + 1 < depth && ev) { ev->inf.push_back(ev->make()); (void)hipEventRecord(ev->inf.back(), stream); }
        if (b + 1 < depth) {
          rc = guide_infer_vertices(stream, gr->g, gr->g->vertices, a.counters + kVertex, cap, gr->bounds,
                                    gr->g->outputs);
          if (rc) return rc;
          if (ev) { ev->inf.push_back(ev->make()); (void)hipEventRecord(ev->inf.back(), stream); }
          guide_sample_kernel<<<g_ext, kRtBlock, 0, stream>>>(sv, q, st->rng_seed, b, gr->g->outputs,
                                                             gr->train ? gr->g->records : nullptr);
          BJ_HIP(hipGetLastError());
        }
      }
      if (kAlg != BAJO_RENDER_PATH && kAlg != BAJO_RENDER_NORMALS) {
        if (ev) { ev->sh.push_back(ev->make()); (void)hipEventRecord(ev->sh.back(), stream); }
        // sparse bounces: one wave per ray; dense ones: lane groups or one
        // lane per ray (triangle and batch scenes, 4- and 8-wide)
        constexpr bool kSparse = kWide && kGeo != kGeoMixed;
        const uint32_t sparse = kSparse ? sparse_rays : 0u;
        const uint32_t dense_group = kSparse ? shadow_group : 1u;
        // counts the lane-group kernel takes: (sparse, group_hi]; the
        // per-lane kernel takes the counts above `per_lane_lo`
        const uint32_t group_hi = dense_group > 1 ? (adaptive_group ? kGroupShadowRays : 0xFFFFFFFFu) : sparse;
        const uint32_t per_lane_lo = group_hi > sparse ? group_hi : sparse;
        if (dense_group > 1 && !adaptive_group) {
          // the lane-group kernel takes every count above the sparse threshold
        } else if (kWide && shadow_walk == 1)
          shadow_kernel<kAlg, kWide, kGeo, kWide ? 1 : 0><<<g_lane, kRtBlock, 0, stream>>>(sv, q, b, a.spill,
                                                                                     a.work_sh, a.stats, per_lane_lo);
        else if (kWide && shadow_walk == 2)
          shadow_kernel<kAlg, kWide, kGeo, kWide ? 2 : 0><<<g_lane, kRtBlock, 0, stream>>>(sv, q, b, a.spill,
                                                                                     a.work_sh, a.stats, per_lane_lo);
        else if (kWide && shadow_walk == 3)
          shadow_kernel<kAlg, kWide, kGeo, kWide ? 3 : 0><<<g_lane, kRtBlock, 0, stream>>>(sv, q, b, a.spill,
                                                                                     a.work_sh, a.stats, per_lane_lo);
        else
          shadow_kernel<kAlg, kWide, kGeo><<<g_lane, kRtBlock, 0, stream>>>(sv, q, b, a.spill, a.work_sh, a.stats,
                                                                          per_lane_lo);
        if constexpr (kSparse) {
          constexpr int kG = kSparse ? kGeo : kGeoTriangles;
          // dense bounces with lane groups (the per-lane kernel then stands aside)
          if (dense_group == 8)
            shadow_group_kernel<kAlg, kG, 8><<<g_group, kRtBlock, 0, stream>>>(sv, q, b, sparse, group_hi,
                                                                             a.work_sh);
          else if (dense_group == 16)
            shadow_group_kernel<kAlg, kG, 16><<<g_group, kRtBlock, 0, stream>>>(sv, q, b, sparse, group_hi,
                                                                              a.work_sh);
          else if (dense_group == 32)
            shadow_group_kernel<kAlg, kG, 32><<<g_group, kRtBlock, 0, stream>>>(sv, q, b, sparse, group_hi,
                                                                              a.work_sh);
          if (sparse)
            shadow_group_kernel<kAlg, kG, 64><<<g_wave, kRtBlock, 0, stream>>>(sv, q, b, 0u, sparse, a.work_sh);
        }
        BJ_HIP(hipGetLastError());
        if (ev) { ev->sh.push_back(ev->make()); (void)hipEventRecord(ev->sh.back(), stream); }
      }
      const bool traced_shadows = kAlg != BAJO_RENDER_PATH && kAlg != BAJO_RENDER_NORMALS;
      partials_sum_kernel<<<kSumBlocks, 256, 0, stream>>>(a.counters, (uint32_t)cap, a.work_ext, g_ext, a.work_sh,
                                                          traced_shadows ? g_sh : 0u, a.work_sums);
      advance_kernel<<<1, 64, 0, stream>>>(a.counters, a.stats, b, (uint32_t)cap, a.work_sums);
    }
    const uint32_t pixel_count = chunk / st->samples_per_pixel;
    resolve_kernel<<<div_up(pixel_count, kRtBlock), kRtBlock, 0, stream>>>(
        a.sample, pixels, cap, (uint32_t)(begin / st->samples_per_pixel), pixel_count, st->samples_per_pixel);
    BJ_HIP(hipGetLastError());
  }
  return BAJO_OK;
}

// Instantiates `fn` for the scene's node width and geometry (triangles only,
// or spheres + triangles).
template <int kAlg, typename F>
int32_t by_scene(const bajo_scene* s, F&& fn) {
  const bool wide = wide_layout(s->options.node_width);
  if (s->env_count)
    return wide ? fn.template operator()<kAlg, true, kGeoBatch>() : fn.template operator()<kAlg, false, kGeoBatch>();
  if (s->n_spheres)
    return wide ? fn.template operator()<kAlg, true, kGeoMixed>() : fn.template operator()<kAlg, false, kGeoMixed>();
  return wide ? fn.template operator()<kAlg, true, kGeoTriangles>()
              : fn.template operator()<kAlg, false, kGeoTriangles>();
}

template <typename F>
int32_t by_algorithm(const bajo_scene* s, int32_t algorithm, F&& fn) {
  switch (algorithm) {
    case BAJO_RENDER_PATH: return by_scene<BAJO_RENDER_PATH>(s, fn);
    case BAJO_RENDER_NORMALS: return by_scene<BAJO_RENDER_NORMALS>(s, fn);
    case BAJO_RENDER_AO: return by_scene<BAJO_RENDER_AO>(s, fn);
    case BAJO_RENDER_NEE: return by_scene<BAJO_RENDER_NEE>(s, fn);
    case BAJO_RENDER_MIS: return by_scene<BAJO_RENDER_MIS>(s, fn);
    default: BJ_REQUIRE(false, "bajo_render: unknown algorithm %d", algorithm);
  }
}

// The instantiation-dispatched calls (by_algorithm picks the template arguments).
struct ShapeCall {
  const bajo_scene* s;
  const bajo_render_settings* st;
  RenderShape* out;
  template <int kAlg, bool kWide, int kGeo>
  int32_t operator()() const { return render_shape<kAlg, kWide, kGeo>(s, st, out); }
};
struct RenderCall {
  bajo_ctx* ctx;
  bajo_scene* s;
  RenderArena* a;
  const bajo_render_settings* st;
  const float* camera;
  uint32_t camera_count;
  float* pixels;
  LaunchEvents* ev;
  template <int kAlg, bool kWide, int kGeo>
  int32_t operator()() const {
    return run_render<kAlg, kWide, kGeo>(ctx, s, *a, st, camera, camera_count, pixels, ev);
  }
};

// Sizes arena `a` for a render of `st` over scene `s` without enqueueing it.
int32_t prepare_arena(const bajo_scene* s, RenderArena& a, const bajo_render_settings* st) {
  RenderShape shape;
  const int32_t rc = by_algorithm(s, st->algorithm, ShapeCall{s, st, &shape});
  if (rc) return rc;
  return ensure_arena(a, shape.cap, shape.spill_threads);
}

// Whether a render of `s` with `st` would reallocate arena `a` (its path
// capacity or the spill columns of its widest launch grow).
int32_t arena_grows(const bajo_scene* s, const RenderArena& a, const bajo_render_settings* st, bool* grows) {
  RenderShape shape;
  const int32_t rc = by_algorithm(s, st->algorithm, ShapeCall{s, st, &shape});
  if (rc) return rc;
  *grows = shape.cap > a.cap || shape.spill_threads > a.spill_threads;
  return BAJO_OK;
}

int32_t dispatch_render(bajo_ctx* ctx, bajo_scene* s, RenderArena& a, const bajo_render_settings* st,
                        const float* camera, uint32_t camera_count, float* pixels, LaunchEvents* ev) {
  return by_algorithm(s, st->algorithm, RenderCall{ctx, s, &a, st, camera, camera_count, pixels, ev});
}

}  // namespace

extern "C" {

int32_t bajo_scene_create(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                          const float* lambertian, uint32_t n_lambertian, const float* emissive, uint32_t n_emissive,
                          bajo_scene** out) {
  return bajo_scene_create_ex(ctx, vertices, n, surfaces, lambertian, n_lambertian, emissive, n_emissive, nullptr,
                              out);
}

int32_t bajo_scene_create_ex(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                             const float* lambertian, uint32_t n_lambertian, const float* emissive,
                             uint32_t n_emissive, const bajo_build_options* options, bajo_scene** out) {
  bajo_materials m{lambertian, n_lambertian, nullptr, 0, nullptr, 0, emissive, n_emissive};
  return bajo_scene_create_materials(ctx, vertices, n, surfaces, &m, options, out);
}

int32_t bajo_scene_create_materials(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                                    const bajo_materials* materials, const bajo_build_options* options,
                                    bajo_scene** out) {
  BJ_REQUIRE(n > 0, "bajo_scene_create: a scene needs triangles and their surfaces");
  return bajo_scene_create_mixed(ctx, vertices, n, surfaces, nullptr, 0, nullptr, materials, options, out);
}

namespace {
int32_t scene_create(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces, const float* spheres,
                     uint64_t n_spheres, const uint32_t* sphere_surfaces, const bajo_materials* materials,
                     const bajo_build_options* options, bool build, bajo_scene** out);
}  // namespace

int32_t bajo_scene_create_mixed(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces,
                                const float* spheres, uint64_t n_spheres, const uint32_t* sphere_surfaces,
                                const bajo_materials* materials, const bajo_build_options* options,
                                bajo_scene** out) {
  return scene_create(ctx, vertices, n, surfaces, spheres, n_spheres, sphere_surfaces, materials, options, true, out);
}

namespace {
// Validates and uploads a scene and its light table; builds its BVHs if `build`.
int32_t scene_create(bajo_ctx* ctx, const float* vertices, uint64_t n, const uint32_t* surfaces, const float* spheres,
                     uint64_t n_spheres, const uint32_t* sphere_surfaces, const bajo_materials* materials,
                     const bajo_build_options* options, bool build, bajo_scene** out) {
  BJ_REQUIRE(ctx && out && materials, "bajo_scene_create: NULL ctx/materials/out");
  BJ_REQUIRE(n + n_spheres > 0, "bajo_scene_create: a scene needs triangles or spheres");
  BJ_REQUIRE(n == 0 || (vertices && surfaces), "bajo_scene_create: triangles need their vertices and surfaces");
  BJ_REQUIRE(n_spheres == 0 || (spheres && sphere_surfaces), "bajo_scene_create: spheres need their surfaces");
  BJ_REQUIRE(n < (1ull << 31) && n_spheres < (1ull << 31), "bajo_scene_create: too many primitives");
  const float *lambertian = materials->lambertian, *metal = materials->metal, *dielectric = materials->dielectric,
              *emissive = materials->emissive, *ggx = materials->ggx;
  const uint32_t n_lambertian = materials->n_lambertian, n_metal = materials->n_metal,
                 n_dielectric = materials->n_dielectric, n_emissive = materials->n_emissive,
                 n_ggx = materials->n_ggx;
  BJ_REQUIRE((n_lambertian == 0 || lambertian) && (n_metal == 0 || metal) && (n_dielectric == 0 || dielectric) &&
                 (n_emissive == 0 || emissive) && (n_ggx == 0 || ggx),
             "bajo_scene_create: NULL material array with a nonzero count");
  for (uint32_t i = 0; i < n_ggx; ++i)
    BJ_REQUIRE(ggx[4 * i + 3] > 0.0f && ggx[4 * i + 3] <= 1.0f,
               "bajo_scene_create: GGX material %u has roughness %g outside (0, 1]", i, (double)ggx[4 * i + 3]);
  auto check_surface = [&](uint32_t sid, const char* what, uint64_t i) -> int32_t {
    const uint32_t kind = sid >> 28, mi = sid & kSurfaceIndexMask;
    BJ_REQUIRE(kind <= kMatGgx, "bajo_scene_create: %s %llu has material kind %u", what, (unsigned long long)i,
               kind);
    BJ_REQUIRE(kind != kMatGgx || mi < n_ggx, "bajo_scene_create: GGX index out of range");
    BJ_REQUIRE(kind != kMatLambertian || mi < n_lambertian, "bajo_scene_create: Lambertian index out of range");
    BJ_REQUIRE(kind != kMatMetal || mi < n_metal, "bajo_scene_create: metal index out of range");
    BJ_REQUIRE(kind != kMatDielectric || mi < n_dielectric, "bajo_scene_create: dielectric index out of range");
    BJ_REQUIRE(kind != kMatEmissive || mi < n_emissive, "bajo_scene_create: emissive index out of range");
    return BAJO_OK;
  };
  for (uint64_t i = 0; i < n; ++i)
    if (int32_t rc = check_surface(surfaces[i], "triangle", i)) return rc;
  for (uint64_t i = 0; i < n_spheres; ++i) {
    if (int32_t rc = check_surface(sphere_surfaces[i], "sphere", i)) return rc;
    // SceneData (rt/types.mojo:448-452): a sphere's radius must be non-zero
    BJ_REQUIRE(spheres[4 * i + 3] != 0.0f, "bajo_scene_create: sphere %llu has radius 0", (unsigned long long)i);
  }
  bajo_scene* s = new bajo_scene();
  s->device = ctx->device;
  s->n = n;
  s->n_spheres = n_spheres;
  s->options = options ? *options : bajo_build_options{BAJO_BUILD_LBVH, BAJO_WIDTH_BINARY, 0, 0};
  auto fail = [&](int32_t rc) { free_scene(s); return rc; };
#define SC_HIP(call)                                              \
  do {                                                            \
    hipError_t e_ = (call);                                       \
    if (e_ != hipSuccess) return fail(hip_fail(e_, #call, __FILE__, __LINE__)); \
  } while (0)
  if (n) {
    SC_HIP(hipMalloc(&s->vertices, n * 9 * sizeof(float)));
    SC_HIP(hipMemcpy(s->vertices, vertices, n * 9 * sizeof(float), hipMemcpyHostToDevice));
    SC_HIP(hipMalloc(&s->surfaces, n * sizeof(uint32_t)));
    SC_HIP(hipMemcpy(s->surfaces, surfaces, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  SC_HIP(hipMalloc(&s->lambertian, (n_lambertian ? n_lambertian : 1) * 3 * sizeof(float)));
  if (n_lambertian) SC_HIP(hipMemcpy(s->lambertian, lambertian, n_lambertian * 3 * sizeof(float), hipMemcpyHostToDevice));
  SC_HIP(hipMalloc(&s->metals, (n_metal ? n_metal : 1) * 4 * sizeof(float)));
  if (n_metal) SC_HIP(hipMemcpy(s->metals, metal, n_metal * 4 * sizeof(float), hipMemcpyHostToDevice));
  SC_HIP(hipMalloc(&s->dielectrics, (n_dielectric ? n_dielectric : 1) * sizeof(float)));
  if (n_dielectric) SC_HIP(hipMemcpy(s->dielectrics, dielectric, n_dielectric * sizeof(float), hipMemcpyHostToDevice));
  SC_HIP(hipMalloc(&s->ggx, (n_ggx ? n_ggx : 1) * 4 * sizeof(float)));
  if (n_ggx) SC_HIP(hipMemcpy(s->ggx, ggx, n_ggx * 4 * sizeof(float), hipMemcpyHostToDevice));
  SC_HIP(hipMalloc(&s->emissive, (n_emissive ? n_emissive : 1) * 3 * sizeof(float)));
  if (n_emissive) SC_HIP(hipMemcpy(s->emissive, emissive, n_emissive * 3 * sizeof(float), hipMemcpyHostToDevice));
  // _build_light_store (rt/types.mojo:490-530): one light per emissive
  // triangle (weight = area * mean radiance), then per emissive sphere
  // (weight = 4 pi r^2 * mean radiance, |r|); one alias table over both.
  std::vector<float> lf, weights;
  std::vector<uint32_t> prim_of_light;
  std::vector<uint8_t> sphere_light;
  float total = 0.0f;
  for (uint64_t i = 0; i < n; ++i) {
    if ((surfaces[i] >> 28) != kMatEmissive) continue;
    const float* t = vertices + 9 * i;
    const v3 p0 = mk(t[0], t[1], t[2]), p1 = mk(t[3], t[4], t[5]), p2 = mk(t[6], t[7], t[8]);
    const float area = 0.5f * sqrtf(length2(cross(sub(p1, p0), sub(p2, p0))));
    const float* e = emissive + 3 * (surfaces[i] & kSurfaceIndexMask);
    const float imp = fmax_sel((e[0] + e[1] + e[2]) / 3.0f, 0.0f);
    const float w = area * imp;
    if (w > 0.0f) {
      weights.push_back(w);
      prim_of_light.push_back((uint32_t)i);
      sphere_light.push_back(0);
      total += w;
    }
  }
  for (uint64_t i = 0; i < n_spheres; ++i) {
    if ((sphere_surfaces[i] >> 28) != kMatEmissive) continue;
    const float r = fabsf(spheres[4 * i + 3]);
    const float* e = emissive + 3 * (sphere_surfaces[i] & kSurfaceIndexMask);
    const float imp = fmax_sel((e[0] + e[1] + e[2]) / 3.0f, 0.0f);
    const float w = 4.0f * 3.14159265358979f * r * r * imp;
    if (w > 0.0f) {
      weights.push_back(w);
      prim_of_light.push_back((uint32_t)i);
      sphere_light.push_back(1);
      total += w;
    }
  }
  std::vector<float> prob;
  std::vector<uint32_t> alias;
  build_alias(weights, total, prob, alias);
  const uint32_t nl = (uint32_t)weights.size();
  std::vector<uint32_t> kinds(nl ? nl : 1, 0u);
  lf.assign((nl ? nl : 1) * kLightStride, 0.0f);
  for (uint32_t l = 0; l < nl; ++l) {
    float* o = lf.data() + (size_t)l * kLightStride;
    const float* e;
    if (sphere_light[l]) {
      const float* sp = spheres + 4 * (uint64_t)prim_of_light[l];
      e = emissive + 3 * (sphere_surfaces[prim_of_light[l]] & kSurfaceIndexMask);
      o[0] = sp[0]; o[1] = sp[1]; o[2] = sp[2];
      o[9] = fabsf(sp[3]);
    } else {
      const float* t = vertices + 9 * (uint64_t)prim_of_light[l];
      e = emissive + 3 * (surfaces[prim_of_light[l]] & kSurfaceIndexMask);
      for (int k = 0; k < 9; ++k) o[k] = t[k];
      o[9] = 0.0f;
    }
    o[10] = e[0]; o[11] = e[1]; o[12] = e[2];
    o[13] = prob[l];
    kinds[l] = (alias[l] << 4) | (sphere_light[l] ? kPrimSphere : kPrimTriangle);
  }
  s->light_count = nl;
  s->total_light_weight = total;
  SC_HIP(hipMalloc(&s->light_fields, lf.size() * sizeof(float)));
  SC_HIP(hipMemcpy(s->light_fields, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice));
  SC_HIP(hipMalloc(&s->light_kinds, kinds.size() * sizeof(uint32_t)));
  SC_HIP(hipMemcpy(s->light_kinds, kinds.data(), kinds.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (n_spheres) {
    // the BVH is built over |r| (sphere_for_acceleration, rt/geometry.mojo:41-46)
    std::vector<float> absr(4 * n_spheres), radii(n_spheres);
    for (uint64_t i = 0; i < n_spheres; ++i) {
      for (int k = 0; k < 3; ++k) absr[4 * i + k] = spheres[4 * i + k];
      absr[4 * i + 3] = fabsf(spheres[4 * i + 3]);
      radii[i] = spheres[4 * i + 3];
    }
    SC_HIP(hipMalloc(&s->spheres, absr.size() * sizeof(float)));
    SC_HIP(hipMemcpy(s->spheres, absr.data(), absr.size() * sizeof(float), hipMemcpyHostToDevice));
    SC_HIP(hipMalloc(&s->sphere_radii, radii.size() * sizeof(float)));
    SC_HIP(hipMemcpy(s->sphere_radii, radii.data(), radii.size() * sizeof(float), hipMemcpyHostToDevice));
    SC_HIP(hipMalloc(&s->sphere_surfaces, n_spheres * sizeof(uint32_t)));
    SC_HIP(hipMemcpy(s->sphere_surfaces, sphere_surfaces, n_spheres * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
#undef SC_HIP
  if (build) {
    const int32_t rc = bajo_scene_rebuild(ctx, s, nullptr);
    if (rc) return fail(rc);
  }
  *out = s;
  return BAJO_OK;
}
}  // namespace

int32_t bajo_scene_create_batch(bajo_ctx* ctx, const float* vertices, const uint64_t* offsets, uint32_t env_count,
                                const uint32_t* surfaces, const bajo_materials* materials,
                                const bajo_build_options* options, bajo_scene** out) {
  BJ_REQUIRE(ctx && vertices && offsets && surfaces && materials && out, "bajo_scene_create_batch: NULL argument");
  BJ_REQUIRE(env_count >= 1, "bajo_scene_create_batch: a batch needs at least one environment");
  BJ_REQUIRE(offsets[0] == 0, "bajo_scene_create_batch: offsets[0] must be 0");
  for (uint32_t e = 0; e < env_count; ++e)
    BJ_REQUIRE(offsets[e + 1] > offsets[e], "bajo_scene_create_batch: environment %u is empty", e);
  const uint64_t n = offsets[env_count];
  BJ_REQUIRE(n < (1ull << 31), "bajo_scene_create_batch: too many triangles");
  const bajo_build_options opt = options ? *options : bajo_build_options{BAJO_BUILD_LBVH, BAJO_WIDTH_BINARY, 0, 0};
  BJ_REQUIRE(opt.method == BAJO_BUILD_LBVH || opt.method == BAJO_BUILD_HPLOC,
             "bajo_scene_create_batch: method must be LBVH (0) or H-PLOC (1)");
  // materials, surfaces and uploads as for one scene; the lights are then
  // rebuilt per environment below
  bajo_scene* s = nullptr;
  int32_t rc = scene_create(ctx, vertices, n, surfaces, nullptr, 0, nullptr, materials, &opt, false, &s);
  if (rc) return rc;
  auto fail = [&](int32_t r) { free_scene(s); return r; };
  (void)hipFree(s->light_fields); (void)hipFree(s->light_kinds);
  s->light_fields = nullptr;
  s->light_kinds = nullptr;
  // _build_light_store (rt/types.mojo:490-530) per environment: its emissive
  // triangles (weight = area * mean radiance) and its own alias table; the
  // alias indices are local to the environment's light range
  const float* emissive = materials->emissive;
  std::vector<float> lf;
  std::vector<uint32_t> kinds, base(env_count + 1, 0u);
  std::vector<float> totals(env_count, 0.0f);
  for (uint32_t e = 0; e < env_count; ++e) {
    std::vector<float> weights;
    std::vector<uint32_t> prims;
    float total = 0.0f;
    for (uint64_t i = offsets[e]; i < offsets[e + 1]; ++i) {
      if ((surfaces[i] >> 28) != kMatEmissive) continue;
      const float* t = vertices + 9 * i;
      const v3 p0 = mk(t[0], t[1], t[2]), p1 = mk(t[3], t[4], t[5]), p2 = mk(t[6], t[7], t[8]);
      const float area = 0.5f * sqrtf(length2(cross(sub(p1, p0), sub(p2, p0))));
      const float* em = emissive + 3 * (surfaces[i] & kSurfaceIndexMask);
      const float w = area * fmax_sel((em[0] + em[1] + em[2]) / 3.0f, 0.0f);
      if (w > 0.0f) {
        weights.push_back(w);
        prims.push_back((uint32_t)i);
        total += w;
      }
    }
    std::vector<float> prob;
    std::vector<uint32_t> alias;
    build_alias(weights, total, prob, alias);
    for (size_t l = 0; l < weights.size(); ++l) {
      const float* t = vertices + 9 * (uint64_t)prims[l];
      const float* em = emissive + 3 * (surfaces[prims[l]] & kSurfaceIndexMask);
      float o[kLightStride] = {};
      for (int k = 0; k < 9; ++k) o[k] = t[k];
      o[10] = em[0]; o[11] = em[1]; o[12] = em[2];
      o[13] = prob[l];
      lf.insert(lf.end(), o, o + kLightStride);
      kinds.push_back((alias[l] << 4) | kPrimTriangle);
    }
    base[e + 1] = base[e] + (uint32_t)weights.size();
    totals[e] = total;
  }
  s->light_count = base[env_count];
  s->total_light_weight = 0.0f;
  if (lf.empty()) {
    lf.assign(kLightStride, 0.0f);
    kinds.assign(1, 0u);
  }
#define SB_HIP(call)                                              \
  do {                                                            \
    hipError_t e_ = (call);                                       \
    if (e_ != hipSuccess) return fail(hip_fail(e_, #call, __FILE__, __LINE__)); \
  } while (0)
  SB_HIP(hipMalloc(&s->light_fields, lf.size() * sizeof(float)));
  SB_HIP(hipMemcpy(s->light_fields, lf.data(), lf.size() * sizeof(float), hipMemcpyHostToDevice));
  SB_HIP(hipMalloc(&s->light_kinds, kinds.size() * sizeof(uint32_t)));
  SB_HIP(hipMemcpy(s->light_kinds, kinds.data(), kinds.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  SB_HIP(hipMalloc(&s->env_light_base, base.size() * sizeof(uint32_t)));
  SB_HIP(hipMemcpy(s->env_light_base, base.data(), base.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  SB_HIP(hipMalloc(&s->env_light_total, totals.size() * sizeof(float)));
  SB_HIP(hipMemcpy(s->env_light_total, totals.data(), totals.size() * sizeof(float), hipMemcpyHostToDevice));
#undef SB_HIP
  s->env_count = env_count;
  s->env_offsets.assign(offsets, offsets + env_count + 1);
  rc = bajo_scene_rebuild(ctx, s, nullptr);
  if (rc) return fail(rc);
  *out = s;
  return BAJO_OK;
}

int32_t bajo_scene_rebuild(bajo_ctx* ctx, bajo_scene* s, bajo_build_timings* timings) {
  BJ_REQUIRE(ctx && s, "bajo_scene_rebuild: NULL argument");
  if (s->env_count) {
    // every environment's BVH in one segmented build (rebuilt in place)
    return bajo_build_triangle_blas_set(ctx, s->vertices, s->env_offsets.data(), s->env_count, &s->options, timings,
                                        &s->set);
  }
  if (s->n) {
    const int32_t rc = bajo_build_triangle_bvh_ex(ctx, s->vertices, s->n, &s->options, timings, &s->bvh);
    if (rc) return rc;
  }
  if (s->n_spheres) {
    // the sphere BVH: the LBVH over the spheres' boxes at the scene's node width and leaf size
    const bajo_build_options so{BAJO_BUILD_LBVH, s->options.node_width, s->options.leaf_size, 0};
    return bajo_build_sphere_bvh(ctx, s->spheres, s->n_spheres, &so, nullptr, &s->sphere_bvh);
  }
  return BAJO_OK;
}

int32_t bajo_rng_stage_floats(uint64_t seed, uint32_t path_id, uint32_t stage, uint32_t count, float* out) {
  BJ_REQUIRE(out || count == 0, "bajo_rng_stage_floats: NULL output");
  StageRng rng(seed, path_id, stage);
  for (uint32_t i = 0; i < count; ++i) out[i] = rng.next();
  return BAJO_OK;
}

int32_t bajo_scene_vertices(const bajo_scene* s, float** vertices, uint64_t* count) {
  BJ_REQUIRE(s && vertices && count, "bajo_scene_vertices: NULL argument");
  *vertices = s->vertices;
  *count = s->n;
  return BAJO_OK;
}

int32_t bajo_scene_set_vertices(bajo_ctx* ctx, bajo_scene* s, const float* vertices, uint64_t n) {
  BJ_REQUIRE(ctx && s && vertices, "bajo_scene_set_vertices: NULL argument");
  BJ_REQUIRE(n == s->n, "bajo_scene_set_vertices: %llu triangles, the scene holds %llu", (unsigned long long)n,
             (unsigned long long)s->n);
  BJ_HIP(hipMemcpyAsync(s->vertices, vertices, n * 9 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  return BAJO_OK;
}

int32_t bajo_scene_free(bajo_scene* s) {
  free_scene(s);
  return BAJO_OK;
}

int32_t bajo_scene_get_bvh(const bajo_scene* s, bajo_bvh** out) {
  BJ_REQUIRE(s && out, "bajo_scene_get_bvh: NULL argument");
  *out = s->bvh;
  return BAJO_OK;
}

int32_t bajo_scene_light_table(const bajo_scene* s, float* fields, uint32_t* kinds, uint32_t* count,
                               float* total_weight) {
  BJ_REQUIRE(s && count && total_weight, "bajo_scene_light_table: NULL argument");
  *count = s->light_count;
  *total_weight = s->total_light_weight;
  if (s->light_count && fields)
    BJ_HIP(hipMemcpy(fields, s->light_fields, (size_t)s->light_count * kLightStride * sizeof(float), hipMemcpyDeviceToHost));
  if (s->light_count && kinds)
    BJ_HIP(hipMemcpy(kinds, s->light_kinds, (size_t)s->light_count * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return BAJO_OK;
}

int32_t bajo_render(bajo_ctx* ctx, bajo_scene* s, const bajo_render_settings* st, const float* camera, float* pixels,
                    bajo_render_stats* stats) {
  return bajo_render_cameras(ctx, s, st, camera, 1, pixels, stats);
}

int32_t bajo_render_cameras(bajo_ctx* ctx, bajo_scene* s, const bajo_render_settings* st, const float* camera,
                            uint32_t camera_count, float* pixels, bajo_render_stats* stats) {
  BJ_REQUIRE(ctx && s && st && camera && pixels, "bajo_render: NULL argument");
  BJ_REQUIRE(camera_count == 1 || camera_count == s->env_count,
             "bajo_render: %u cameras; one, or one per environment (%u)", camera_count, s->env_count);
  BJ_REQUIRE(st->image_width > 0 && st->image_height > 0 && st->samples_per_pixel > 0,
             "bajo_render: image size and samples per pixel must be positive");
  BJ_REQUIRE(st->max_depth <= kMaxBounces, "bajo_render: max_depth above %d", kMaxBounces);
  const uint64_t total = (uint64_t)st->image_width * st->image_height * st->samples_per_pixel;
  BJ_REQUIRE(total < (1ull << 31), "bajo_render: %llu paths exceed the 31-bit path id", (unsigned long long)total);
  int32_t rc;
  LaunchEvents events;
  LaunchEvents* ev = stats ? &events : nullptr;
  if (ev) {
    ev->begin = ev->make();
    ev->end = ev->make();
    BJ_HIP(hipEventRecord(ev->begin, ctx->stream));
  }
  rc = dispatch_render(ctx, s, s->arena, st, camera, camera_count, pixels, ev);
  if (rc) return rc;
  if (stats) {
    BJ_HIP(hipEventRecord(ev->end, ctx->stream));
    memset(stats, 0, sizeof(*stats));
    unsigned long long host[kStatsWords];
    uint32_t counters[kCounterWords];
    BJ_HIP(hipMemcpyAsync(host, s->arena.stats, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
    BJ_HIP(hipMemcpyAsync(counters, s->arena.counters, sizeof(counters), hipMemcpyDeviceToHost, ctx->stream));
    BJ_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    BJ_HIP(hipEventElapsedTime(&ms, ev->begin, ev->end));
    stats->total_ms = ms;
    for (size_t i = 0; i + 1 < ev->ext.size(); i += 2) {
      BJ_HIP(hipEventElapsedTime(&ms, ev->ext[i], ev->ext[i + 1]));
      stats->extend_ms += ms;
      if (i / 2 < 16) stats->extend_bounce_ms[i / 2] = ms;
    }
    for (size_t i = 0; i + 1 < ev->sh.size(); i += 2) {
      BJ_HIP(hipEventElapsedTime(&ms, ev->sh[i], ev->sh[i + 1]));
      stats->shadow_ms += ms;
      if (i / 2 < 16) stats->shadow_bounce_ms[i / 2] = ms;
    }
    stats->extend_launches = (uint32_t)(ev->ext.size() / 2);
    stats->shadow_launches = (uint32_t)(ev->sh.size() / 2);
    stats->extension_node_visits = host[kWorkBase + 0];
    stats->extension_leaf_tests = host[kWorkBase + 1];
    stats->shadow_node_visits = host[kWorkBase + 2];
    stats->shadow_leaf_tests = host[kWorkBase + 3];
    for (int b = 0; b < 16; ++b) {
      stats->max_extension_visits[b] = (uint32_t)host[kMaxVisitBase + 2 * b];
      stats->max_shadow_visits[b] = (uint32_t)host[kMaxVisitBase + 2 * b + 1];
    }
    for (int b = 0; b < kMaxBounces; ++b) {
      stats->extension_rays += host[2 * b];
      stats->shadow_rays += host[2 * b + 1];
      if (b >= 1) stats->incoherent_rays += host[2 * b];
      stats->incoherent_rays += b >= 1 ? host[2 * b + 1] : 0;
      if (b < 16) stats->rays_per_bounce[b] = host[2 * b];
    }
    stats->status = counters[kStatus];
    stats->traversal_status = counters[8];
    if (stats->status) return (int32_t)(BAJO_STATUS_WAVE_BASE + stats->status);
    if (stats->traversal_status) return (int32_t)stats->traversal_status;
  }
  return BAJO_OK;
}

// ---- guided rendering (no bajo counterpart) ------------------------------------

int32_t bajo_render_guided(bajo_ctx* ctx, bajo_scene* s, bajo_guide* guide, const bajo_render_settings* settings,
                           const float* camera, float* pixels, int32_t train, bajo_guide_frame_stats* stats) {
  BJ_REQUIRE(ctx && s && settings && camera && pixels, "bajo_render_guided: NULL argument");
  if (stats) memset(stats, 0, sizeof(*stats));
  bajo_render_settings st = *settings;
  if (!guide) {
    st.algorithm = BAJO_RENDER_NEE;
    return bajo_render(ctx, s, &st, camera, pixels, nullptr);
  }
  BJ_REQUIRE(st.image_width > 0 && st.image_height > 0 && st.samples_per_pixel > 0 && st.max_depth > 0,
             "bajo_render_guided: image size, samples per pixel and depth must be positive");
  BJ_REQUIRE(st.max_depth <= kMaxBounces, "bajo_render_guided: max_depth above %d", kMaxBounces);
  BJ_REQUIRE(s->n_spheres == 0 && s->env_count == 0 && s->bvh != nullptr,
             "bajo_render_guided: guided renders take triangle scenes (no spheres, no batches)");
  BJ_REQUIRE(guide->device == ctx->device, "bajo_render_guided: the guide lives on device %d, the context on %d",
             guide->device, ctx->device);
  const uint64_t total = (uint64_t)st.image_width * st.image_height * st.samples_per_pixel;
  BJ_REQUIRE(total < (1ull << 31), "bajo_render_guided: %llu paths exceed the 31-bit path id",
             (unsigned long long)total);
  BJ_REQUIRE(st.path_capacity == 0 || st.path_capacity >= total,
             "bajo_render_guided: the frame must fit one path chunk (path_capacity 0 or >= %llu)",
             (unsigned long long)total);
  st.algorithm = BAJO_RENDER_GUIDED;
  st.path_capacity = 0;
  hipStream_t stream = ctx->stream;
  int32_t rc = guide_prepare_frame(stream, guide, total, st.max_depth, train != 0);
  if (rc) return rc;
  GuideRun gr;
  gr.g = guide;
  gr.train = train != 0;
  gr.bounds = s->bvh->scene;
  LaunchEvents events;
  LaunchEvents* ev = stats ? &events : nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (ev) {
    ev->begin = ev->make();
    ev->end = ev->make();
    t0 = ev->make();
    t1 = ev->make();
    BJ_HIP(hipEventRecord(ev->begin, stream));
  }
  if (wide_layout(s->options.node_width))
    rc = run_render<BAJO_RENDER_GUIDED, true, kGeoTriangles>(ctx, s, s->arena, &st, camera, 1, pixels, ev, &gr);
  else
    rc = run_render<BAJO_RENDER_GUIDED, false, kGeoTriangles>(ctx, s, s->arena, &st, camera, 1, pixels, ev, &gr);
  if (rc) return rc;
  uint64_t samples = 0;
  uint32_t steps = 0;
  float loss = 0.0f;
  if (ev) BJ_HIP(hipEventRecord(t0, stream));
  if (train) {
    rc = guide_collect(stream, guide, total, st.max_depth, gr.bounds);
    if (rc) return rc;
    rc = guide_train_collected(stream, guide, &samples, &steps, &loss);
    if (rc) return rc;
  }
  if (ev) {
    BJ_HIP(hipEventRecord(t1, stream));
    BJ_HIP(hipEventRecord(ev->end, stream));
    unsigned long long host[kStatsWords];
    uint32_t counters[kCounterWords];
    BJ_HIP(hipMemcpyAsync(host, s->arena.stats, sizeof(host), hipMemcpyDeviceToHost, stream));
    BJ_HIP(hipMemcpyAsync(counters, s->arena.counters, sizeof(counters), hipMemcpyDeviceToHost, stream));
    BJ_HIP(hipStreamSynchronize(stream));
    float ms = 0.0f;
    BJ_HIP(hipEventElapsedTime(&ms, ev->begin, ev->end));
    stats->render_ns = (uint64_t)(ms * 1.0e6f);
    BJ_HIP(hipEventElapsedTime(&ms, t0, t1));
    stats->train_ns = (uint64_t)(ms * 1.0e6f);
    for (size_t i = 0; i + 1 < ev->inf.size(); i += 2) {
      BJ_HIP(hipEventElapsedTime(&ms, ev->inf[i], ev->inf[i + 1]));
      stats->infer_ns += (uint64_t)(ms * 1.0e6f);
    }
    stats->guided_vertices = host[kGuidedWord];
    stats->training_samples = samples;
    stats->adam_steps = steps;
    stats->loss = loss;
    if (counters[kStatus]) return (int32_t)(BAJO_STATUS_WAVE_BASE + counters[kStatus]);
    if (counters[8]) return (int32_t)counters[8];
  }
  return BAJO_OK;
}

// ---- persistent render target (rt/gpu/resources.mojo:38-155) ----------------

struct bajo_render_target {
  int device = 0;
  bajo_render_settings settings{};
  float* pixels = nullptr;  // the caller's f32[3 w h E] device buffer
  uint32_t environments = 1;  // scenes rendered through the target have this many environments
  RenderArena arena;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {
void drop_graph(bajo_render_target* t) {
  if (t->exec) (void)hipGraphExecDestroy(t->exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  t->exec = nullptr;
  t->graph = nullptr;
}

int32_t check_settings(const bajo_render_settings* st, const char* fn) {
  BJ_REQUIRE(st->image_width > 0 && st->image_height > 0 && st->samples_per_pixel > 0,
             "%s: image size and samples per pixel must be positive", fn);
  BJ_REQUIRE(st->max_depth <= kMaxBounces, "%s: max_depth above %d", fn, kMaxBounces);
  BJ_REQUIRE(st->algorithm >= BAJO_RENDER_PATH && st->algorithm <= BAJO_RENDER_MIS, "%s: unknown algorithm %d", fn,
             st->algorithm);
  const uint64_t total = (uint64_t)st->image_width * st->image_height * st->samples_per_pixel;
  BJ_REQUIRE(total < (1ull << 31), "%s: %llu paths exceed the 31-bit path id", fn, (unsigned long long)total);
  return BAJO_OK;
}
}  // namespace

int32_t bajo_render_target_create(bajo_ctx* ctx, const bajo_render_settings* settings, const float* camera,
                                  float* pixels, bajo_render_target** out) {
  return bajo_render_target_create_envs(ctx, settings, 1, camera, pixels, out);
}

int32_t bajo_render_target_create_envs(bajo_ctx* ctx, const bajo_render_settings* settings, uint32_t environments,
                                       const float* camera, float* pixels, bajo_render_target** out) {
  BJ_REQUIRE(ctx && settings && pixels && out, "bajo_render_target_create: NULL argument");
  BJ_REQUIRE(environments >= 1, "bajo_render_target_create: environments must be at least 1");
  int32_t rc = check_settings(settings, "bajo_render_target_create");
  if (rc) return rc;
  const uint64_t paths = (uint64_t)settings->image_width * settings->image_height * settings->samples_per_pixel;
  BJ_REQUIRE(paths * environments < (1ull << 31), "bajo_render_target_create: %llu paths exceed the 31-bit path id",
             (unsigned long long)(paths * environments));
  bajo_render_target* t = new bajo_render_target();
  t->device = ctx->device;
  t->settings = *settings;
  t->pixels = pixels;
  t->environments = environments;
  hipError_t e = hipMalloc(&t->arena.cam, 20 * sizeof(float));
  if (e != hipSuccess) {
    delete t;
    return hip_fail(e, "hipMalloc(camera)", __FILE__, __LINE__);
  }
  if (camera) {
    Cam c;
    memcpy(c.v, camera, sizeof(c.v));
    set_camera_kernel<<<1, 64, 0, ctx->stream>>>(t->arena.cam, c);
    e = hipGetLastError();
    if (e != hipSuccess) {
      free_arena(t->arena);
      delete t;
      return hip_fail(e, "set_camera_kernel", __FILE__, __LINE__);
    }
  }
  *out = t;
  return BAJO_OK;
}

int32_t bajo_render_target_set_camera(bajo_ctx* ctx, bajo_render_target* t, const float* camera) {
  BJ_REQUIRE(ctx && t && camera, "bajo_render_target_set_camera: NULL argument");
  Cam c;
  memcpy(c.v, camera, sizeof(c.v));
  set_camera_kernel<<<1, 64, 0, ctx->stream>>>(t->arena.cam, c);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

int32_t bajo_render_enqueue(bajo_ctx* ctx, bajo_render_target* t, bajo_scene* s) {
  BJ_REQUIRE(ctx && t && s, "bajo_render_enqueue: NULL argument");
  BJ_REQUIRE(t->device == ctx->device && s->device == ctx->device, "bajo_render_enqueue: device mismatch");
  BJ_REQUIRE((s->env_count ? s->env_count : 1u) == t->environments,
             "bajo_render_enqueue: the scene has %u environments, the target's pixels hold %u",
             s->env_count ? s->env_count : 1u, t->environments);
  if (t->exec) {
    // a captured graph holds the arena's buffers: growing the arena frees
    // them, so the graph goes first (a replay then reports nothing captured)
    bool grows = false;
    const int32_t rc = arena_grows(s, t->arena, &t->settings, &grows);
    if (rc) return rc;
    if (grows) {
      BJ_HIP(hipStreamSynchronize(ctx->stream));
      drop_graph(t);
    }
  }
  return dispatch_render(ctx, s, t->arena, &t->settings, nullptr, 0, t->pixels, nullptr);
}

int32_t bajo_render_target_capture(bajo_ctx* ctx, bajo_render_target* t, bajo_scene* s) {
  return bajo_render_target_capture_ex(ctx, t, s, 0);
}

int32_t bajo_render_target_capture_ex(bajo_ctx* ctx, bajo_render_target* t, bajo_scene* s, uint32_t flags) {
  BJ_REQUIRE(ctx && t && s, "bajo_render_target_capture: NULL argument");
  BJ_REQUIRE(t->device == ctx->device && s->device == ctx->device, "bajo_render_target_capture: device mismatch");
  BJ_REQUIRE((s->env_count ? s->env_count : 1u) == t->environments,
             "bajo_render_target_capture: the scene has %u environments, the target's pixels hold %u",
             s->env_count ? s->env_count : 1u, t->environments);
  BJ_REQUIRE((flags & ~(uint32_t)BAJO_CAPTURE_REBUILD) == 0, "bajo_render_target_capture: unknown flags %#x", flags);
  const bool rebuild = (flags & BAJO_CAPTURE_REBUILD) != 0;
  drop_graph(t);
  // every allocation happens before the capture: the arena, and with a
  // rebuild one eager rebuild (it sizes the build buffers and the context
  // scratch, and uploads the segment offsets the captured build reuses)
  int32_t rc = prepare_arena(s, t->arena, &t->settings);
  if (rc) return rc;
  if (rebuild) {
    rc = bajo_scene_rebuild(ctx, s, nullptr);
    if (rc) return rc;
  }
  hipStream_t cs = nullptr;
  BJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  bajo_ctx cap = *ctx;
  cap.stream = cs;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
  }
  rc = rebuild ? bajo_scene_rebuild(&cap, s, nullptr) : BAJO_OK;
  if (!rc) rc = dispatch_render(&cap, s, t->arena, &t->settings, nullptr, 0, t->pixels, nullptr);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  }
  t->graph = g;
  e = hipGraphInstantiate(&t->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(t);
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  return BAJO_OK;
}

int32_t bajo_render_target_replay(bajo_ctx* ctx, bajo_render_target* t) {
  BJ_REQUIRE(ctx && t, "bajo_render_target_replay: NULL argument");
  BJ_REQUIRE(t->exec != nullptr, "bajo_render_target_replay: nothing captured (bajo_render_target_capture)");
  BJ_HIP(hipGraphLaunch(t->exec, ctx->stream));
  return BAJO_OK;
}

int32_t bajo_render_target_free(bajo_render_target* t) {
  if (!t) return BAJO_OK;
  drop_graph(t);
  free_arena(t->arena);
  delete t;
  return BAJO_OK;
}

}  // extern "C"
