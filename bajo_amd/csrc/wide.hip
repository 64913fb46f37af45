// Binary LBVH -> 8-wide BVH with quantized child boxes (gfx950).
//
// Plays the role of bajo's wide collapse and compressed-BVH8 encoding
// (bvh/gpu/builder/wide_collapse.mojo:148-643, bvh/gpu/compressed_bounds_bvh.mojo:
// 84-369): a top-down collapse that opens the child with the largest surface
// area until a node has 8 children -- first among children holding more than
// the leaf size, then among any inner child -- and turns inner children of
// at most the leaf size (1..8 triangles) into leaves.  Child boxes are
// stored as 8-bit offsets on a per-axis power-of-two grid anchored at the
// node's lower corner.
//
// Unlike bajo's persistent work queue with a spin fail-safe, the collapse here
// is level-synchronous: level k processes the wide nodes allocated while
// processing level k-1 (a contiguous index range), so no kernel ever waits on
// another wave.  The host never reads the level sizes back: it enqueues a
// depth-bounded number of full-grid level launches (each followed by a
// one-thread launch that opens the next level; a launch past the last level
// does nothing), then one single-block kernel that finishes any deeper
// levels with block barriers.
// The whole collapse is therefore enqueue-only and can be captured in a HIP
// graph.  Leaf children reference ranges of the sorted triangle records that
// the refit already wrote (an LBVH subtree covers a contiguous sorted range),
// so no triangle data is copied.
#include "bvh_internal.h"

namespace bj {
namespace {

constexpr int kCollapseBlock = 256;
constexpr int kTailBlock = 1024;
// Levels the single-block finishing kernel takes at most (a tree deeper than
// the full-grid launches plus this raises BAJO_STATUS_COLLAPSE_FAILED).
constexpr uint32_t kTailLevelCap = 4096;
// wide_ctl words: node counter, level range, status
constexpr int kCtlCount = 0, kCtlBegin = 1, kCtlEnd = 2, kCtlStatus = 3;
constexpr uint32_t kStatusOutOfNodes = 1, kStatusTooDeep = 2, kStatusBadRoot = 4;

struct Child {
  uint32_t ref, count;
  float b[6];
};

BJ_D float half_area(const float* b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return dx * dy + dx * dz + dy * dz;
}

BJ_D void load_children(const BjNode* __restrict__ bnodes, const uint32_t* __restrict__ leaf_counts, uint32_t node,
                        Child& l, Child& r) {
  const float4* p = reinterpret_cast<const float4*>(bnodes + node);
  const float4 a = p[0], b = p[1], c = p[2];
  const uint4 m = reinterpret_cast<const uint4*>(bnodes + node)[3];
  l.b[0] = a.x; l.b[1] = a.y; l.b[2] = a.z; l.b[3] = a.w; l.b[4] = b.x; l.b[5] = b.y;
  r.b[0] = b.z; r.b[1] = b.w; r.b[2] = c.x; r.b[3] = c.y; r.b[4] = c.z; r.b[5] = c.w;
  l.ref = m.y;
  r.ref = m.z;
  l.count = (l.ref & kLeafFlag) ? 1u : leaf_counts[l.ref];
  r.count = (r.ref & kLeafFlag) ? 1u : leaf_counts[r.ref];
}

// Smallest biased exponent e with o + 254 * 2^(e-127) >= hi (one spare step).
BJ_HD uint32_t grid_exponent(float o, float hi) {
  const float extent = hi - o;
  uint32_t e = 1;
  if (extent > 0.0f) {
    const uint32_t bits = __builtin_bit_cast(uint32_t, extent / 254.0f);
    e = (bits >> 23) & 0xFFu;
    if (bits & 0x7FFFFFu) e += 1;
    if (e < 1) e = 1;
  }
  while (e < 253) {
    const float s = __builtin_bit_cast(float, e << 23);
    if (o + 254.0f * s >= hi) break;
    ++e;
  }
  return e;
}

BJ_HD void quantize(float o, uint32_t e, float lo, float hi, uint8_t& qlo, uint8_t& qhi) {
  const float s = __builtin_bit_cast(float, e << 23);
  const float inv = __builtin_bit_cast(float, (254u - e) << 23);  // exact 2^-(e-127)
  float fl = floorf((lo - o) * inv);
  fl = fl < 0.0f ? 0.0f : (fl > 255.0f ? 255.0f : fl);
  while (fl > 0.0f && o + fl * s > lo) fl -= 1.0f;
  float fh = ceilf((hi - o) * inv);
  fh = fh < 0.0f ? 0.0f : (fh > 255.0f ? 255.0f : fh);
  while (fh < 255.0f && o + fh * s < hi) fh += 1.0f;
  fl = fl > 0.0f ? fl - 1.0f : 0.0f;      // padding: one step below
  fh = fh < 255.0f ? fh + 1.0f : 255.0f;  // and one above
  qlo = (uint8_t)fl;
  qhi = (uint8_t)fh;
}

// `root_dev` (H-PLOC: the root its clustering wrote) overrides `root_binary`;
// a root outside the hierarchy leaves the level range empty and raises
// BAJO_STATUS_COLLAPSE_FAILED.
__global__ void collapse_root_kernel(const float* __restrict__ scene, uint32_t n, uint32_t root_binary,
                                     const uint32_t* __restrict__ root_dev, BjWide* __restrict__ wide,
                                     uint32_t* __restrict__ wide_src, uint32_t* __restrict__ ctl,
                                     uint32_t* __restrict__ status) {
  if (threadIdx.x != 0) return;
  ctl[kCtlStatus] = 0;
  if (n == 1) {
    // single triangle: one wide node with one leaf child spanning the scene box
    BjWide w;
    w.ox = scene[0]; w.oy = scene[1]; w.oz = scene[2];
    w.ex = (uint8_t)grid_exponent(scene[0], scene[3]);
    w.ey = (uint8_t)grid_exponent(scene[1], scene[4]);
    w.ez = (uint8_t)grid_exponent(scene[2], scene[5]);
    w.n = 1;
    for (int i = 0; i < 8; ++i) {
      w.lo_x[i] = w.lo_y[i] = w.lo_z[i] = 255;
      w.hi_x[i] = w.hi_y[i] = w.hi_z[i] = 0;
      w.child[i] = 0;
    }
    quantize(scene[0], w.ex, scene[0], scene[3], w.lo_x[0], w.hi_x[0]);
    quantize(scene[1], w.ey, scene[1], scene[4], w.lo_y[0], w.hi_y[0]);
    quantize(scene[2], w.ez, scene[2], scene[5], w.lo_z[0], w.hi_z[0]);
    w.child[0] = kLeafFlag | 0u;
    wide[0] = w;
    ctl[kCtlCount] = 1;
    ctl[kCtlBegin] = 0;
    ctl[kCtlEnd] = 0;
    return;
  }
  const uint32_t root = root_dev ? *root_dev : root_binary;
  ctl[kCtlCount] = 1;
  ctl[kCtlBegin] = 0;
  if ((root & kLeafFlag) || root >= n - 1) {
    ctl[kCtlStatus] = kStatusBadRoot;
    ctl[kCtlEnd] = 0;
    atomicMax(status, (uint32_t)BAJO_STATUS_COLLAPSE_FAILED);
    return;
  }
  wide_src[0] = root;
  ctl[kCtlEnd] = 1;
}

// Segmented builds: wide node s is the root of segment s (S roots form level
// 0).  A one-triangle segment's root is written here, with a single leaf
// child spanning the segment's box; its wide_src entry marks it done.
__global__ void seg_collapse_root_kernel(const float* __restrict__ seg_scene, const uint64_t* __restrict__ off,
                                         uint32_t S, BjWide* __restrict__ wide, uint32_t* __restrict__ wide_src,
                                         uint32_t* __restrict__ ctl) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gtid == 0) {
    ctl[kCtlStatus] = 0;
    ctl[kCtlCount] = S;
    ctl[kCtlBegin] = 0;
    ctl[kCtlEnd] = S;
  }
  for (uint32_t s = gtid; s < S; s += stride) {
    const uint64_t first = off[s], count = off[s + 1] - first;
    if (count > 1) {
      wide_src[s] = (uint32_t)first;
      continue;
    }
    const float* b = seg_scene + 12 * (uint64_t)s;
    BjWide w;
    w.ox = b[0]; w.oy = b[1]; w.oz = b[2];
    w.ex = (uint8_t)grid_exponent(b[0], b[3]);
    w.ey = (uint8_t)grid_exponent(b[1], b[4]);
    w.ez = (uint8_t)grid_exponent(b[2], b[5]);
    w.n = 1;
    for (int i = 0; i < 8; ++i) {
      w.lo_x[i] = w.lo_y[i] = w.lo_z[i] = 255;
      w.hi_x[i] = w.hi_y[i] = w.hi_z[i] = 0;
      w.child[i] = 0;
    }
    quantize(b[0], w.ex, b[0], b[3], w.lo_x[0], w.hi_x[0]);
    quantize(b[1], w.ey, b[1], b[4], w.lo_y[0], w.hi_y[0]);
    quantize(b[2], w.ez, b[2], b[5], w.lo_z[0], w.hi_z[0]);
    w.child[0] = kLeafFlag | (uint32_t)first;
    wide[s] = w;
    wide_src[s] = kSentinel;
  }
}

// The next level is the nodes reserved by the last one, clipped to the
// capacity; after a capacity failure the level range is left empty.  Run by
// one thread once every wave of the level has reserved its nodes.
BJ_D void level_advance(uint32_t* __restrict__ ctl, uint32_t capacity, uint32_t* __restrict__ status) {
  const uint32_t end = ctl[kCtlEnd];
  uint32_t next = __hip_atomic_load(ctl + kCtlCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (next > capacity) next = capacity;
  if (__hip_atomic_load(ctl + kCtlStatus, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    next = end;
    atomicMax(status, (uint32_t)BAJO_STATUS_COLLAPSE_FAILED);
  }
  ctl[kCtlBegin] = end;
  ctl[kCtlEnd] = next;
}

// LDS of one block's slot reservation: per-wave totals, then offsets.
struct Reserve {
  uint32_t wave_total[kTailBlock / 64];
  uint32_t first;
};

// Collapses the wide nodes [begin, end) of one level.  The whole block steps
// together (block `block` of `blocks`, `stride` threads apart) so that one
// atomic per block and step reserves the nodes its children need: a
// per-node atomic on the shared counter serializes (about 25 ns each).
BJ_D void collapse_level(const BjNode* __restrict__ bnodes, const uint32_t* __restrict__ leaf_counts,
                         BjWide* __restrict__ wide, uint32_t* __restrict__ wide_src, uint32_t* __restrict__ ctl,
                         uint32_t capacity, uint32_t max_leaf, int max_children, uint32_t begin, uint32_t end,
                         uint32_t block_base, uint32_t stride, Reserve* __restrict__ rs) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (uint32_t bbase = begin + block_base; bbase < end; bbase += stride) {
    const uint32_t w = bbase + threadIdx.x;
    const bool active = w < end && wide_src[w] != kSentinel;
    // The children live in registers: every index into ch[] below is a
    // compile-time constant after unrolling (a data-dependent index would
    // place the array in scratch memory).
    Child ch[8];
    int n = 0;
    if (active) {
      load_children(bnodes, leaf_counts, wide_src[w], ch[0], ch[1]);
      n = 2;
      // open largest-area children: first those too big for a leaf, then any inner child
      for (int phase = 0; phase < 2; ++phase) {
        for (int step = 0; step < 6 && n < max_children; ++step) {
          int pick = -1;
          float best = -1.0f;
          uint32_t pick_ref = 0;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            if (i < n) {
              const bool inner = !(ch[i].ref & kLeafFlag);
              const bool eligible = inner && (phase == 0 ? ch[i].count > max_leaf : ch[i].count >= 2);
              const float a = half_area(ch[i].b);
              if (eligible && a > best) {
                best = a;
                pick = i;
                pick_ref = ch[i].ref;
              }
            }
          }
          if (pick < 0) break;
          Child l, r;
          load_children(bnodes, leaf_counts, pick_ref, l, r);
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            if (i == pick) ch[i] = l;
            if (i == n) ch[i] = r;
          }
          ++n;
        }
      }
    }
    // reserve wide nodes for inner children that stay inner (wave-aggregated)
    uint32_t k = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i < n && !(ch[i].ref & kLeafFlag) && ch[i].count > max_leaf) ++k;
    uint32_t incl = k;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, 64);
      if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) rs->wave_total[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t sum = 0;
      for (uint32_t i = 0; i < waves; ++i) {
        const uint32_t t = rs->wave_total[i];
        rs->wave_total[i] = sum;
        sum += t;
      }
      rs->first = sum ? atomicAdd(ctl + kCtlCount, sum) : 0u;
    }
    __syncthreads();
    const uint32_t first = rs->first + rs->wave_total[wave] + (incl - k);
    if (!active) continue;
    if ((uint64_t)first + k > capacity) {
      atomicOr(ctl + kCtlStatus, kStatusOutOfNodes);
      continue;
    }
    // node box and per-axis grids
    float box[6] = {kFltMax, kFltMax, kFltMax, -kFltMax, -kFltMax, -kFltMax};
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i < n)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          box[a] = fmin_sel(box[a], ch[i].b[a]);
          box[a + 3] = fmax_sel(box[a + 3], ch[i].b[a + 3]);
        }
    const uint32_t e[3] = {grid_exponent(box[0], box[3]), grid_exponent(box[1], box[4]),
                           grid_exponent(box[2], box[5])};
    // packed byte planes: lo_x, lo_y, lo_z, hi_x, hi_y, hi_z, two words each
    uint32_t q[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) q[j] = j < 6 ? 0xFFFFFFFFu : 0u;
    uint32_t child[8];
    uint32_t next = first;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      child[i] = 0;
      if (i < n) {
        const int word = i >> 2, shift = 8 * (i & 3);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          uint8_t ql, qh;
          quantize(box[a], e[a], ch[i].b[a], ch[i].b[a + 3], ql, qh);
          q[2 * a + word] = (q[2 * a + word] & ~(0xFFu << shift)) | ((uint32_t)ql << shift);
          q[6 + 2 * a + word] |= (uint32_t)qh << shift;
        }
        const uint32_t ref = ch[i].ref;
        if (ref & kLeafFlag) {
          child[i] = kLeafFlag | (ref & kIndexMask);  // one triangle
        } else if (ch[i].count <= max_leaf) {
          // inner binary node over sorted range [fence - count + 1, fence]
          const uint32_t fence = bnodes[ref].fence;
          child[i] = kLeafFlag | ((ch[i].count - 1u) << 28) | (fence + 1u - ch[i].count);
        } else {
          child[i] = next;
          wide_src[next] = ref;
          ++next;
        }
      }
    }
    const uint32_t meta = e[0] | (e[1] << 8) | (e[2] << 16) | ((uint32_t)n << 24);
    uint4* dst = reinterpret_cast<uint4*>(wide + w);
    dst[0] = make_uint4(__builtin_bit_cast(uint32_t, box[0]), __builtin_bit_cast(uint32_t, box[1]),
                        __builtin_bit_cast(uint32_t, box[2]), meta);
    dst[1] = make_uint4(q[0], q[1], q[2], q[3]);
    dst[2] = make_uint4(q[4], q[5], q[6], q[7]);
    dst[3] = make_uint4(q[8], q[9], q[10], q[11]);
    dst[4] = make_uint4(child[0], child[1], child[2], child[3]);
    dst[5] = make_uint4(child[4], child[5], child[6], child[7]);
  }
}

// One level over the whole grid; a launch whose level range is empty returns
// at once, and blocks past the level's nodes return before any barrier.
__global__ void __launch_bounds__(kCollapseBlock) collapse_level_kernel(const BjNode* __restrict__ bnodes,
                                                                        const uint32_t* __restrict__ leaf_counts,
                                                                        BjWide* __restrict__ wide,
                                                                        uint32_t* __restrict__ wide_src,
                                                                        uint32_t* __restrict__ ctl,
                                                                        uint32_t capacity, uint32_t max_leaf,
                                                                        int max_children) {
  __shared__ Reserve rs;
  const uint32_t begin = ctl[kCtlBegin], end = ctl[kCtlEnd];
  if ((uint64_t)begin + (uint64_t)blockIdx.x * blockDim.x >= end) return;
  collapse_level(bnodes, leaf_counts, wide, wide_src, ctl, capacity, max_leaf, max_children, begin, end,
                 blockIdx.x * blockDim.x, gridDim.x * blockDim.x, &rs);
}

// Opens the next level once every block of the last one has finished.
__global__ void level_advance_kernel(uint32_t* __restrict__ ctl, uint32_t capacity, uint32_t* __restrict__ status) {
  if (threadIdx.x == 0 && ctl[kCtlBegin] != ctl[kCtlEnd]) level_advance(ctl, capacity, status);
}

// Every level left after the full-grid launches, in one block: level by level
// with block barriers (no wait on another block), up to kTailLevelCap levels.
__global__ void __launch_bounds__(kTailBlock) collapse_tail_kernel(const BjNode* __restrict__ bnodes,
                                                                   const uint32_t* __restrict__ leaf_counts,
                                                                   BjWide* __restrict__ wide,
                                                                   uint32_t* __restrict__ wide_src,
                                                                   uint32_t* __restrict__ ctl, uint32_t capacity,
                                                                   uint32_t max_leaf, int max_children,
                                                                   uint32_t* __restrict__ status) {
  __shared__ uint32_t range[2];
  __shared__ Reserve rs;
  for (uint32_t level = 0;; ++level) {
    if (threadIdx.x == 0) {
      range[0] = ctl[kCtlBegin];
      range[1] = ctl[kCtlEnd];
    }
    __syncthreads();
    const uint32_t begin = range[0], end = range[1];
    if (begin == end) break;
    if (level == kTailLevelCap) {
      if (threadIdx.x == 0) {
        atomicOr(ctl + kCtlStatus, kStatusTooDeep);
        atomicMax(status, (uint32_t)BAJO_STATUS_COLLAPSE_FAILED);
      }
      break;
    }
    collapse_level(bnodes, leaf_counts, wide, wide_src, ctl, capacity, max_leaf, max_children, begin, end, 0u,
                   blockDim.x, &rs);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) level_advance(ctl, capacity, status);
    __syncthreads();
  }
}

}  // namespace

// Full-grid level launches enqueued before the finishing kernel: the depth of
// a W-wide tree over `leaves` leaf groups of max_leaf triangles plus the few
// deeper levels an LBVH or H-PLOC tree adds (measured: 2-3 at W = 8, 4-5 at
// W = 4 for 1M-16M triangles).  Launches past the last level return at once.
uint32_t collapse_grid_levels(uint64_t leaves, uint32_t max_leaf, int width) {
  const uint64_t groups = leaves / (max_leaf ? max_leaf : 1) + 1;
  uint32_t depth = 0;
  for (uint64_t span = 1; span < groups; span *= (uint64_t)width) ++depth;
  return depth + (width == 4 ? 5u : 4u);
}

// Builds the W-wide layout of h from its binary hierarchy.  Enqueue-only: the
// host reads nothing back (the node count is read at the first host query,
// H-PLOC's root comes from the device).
int32_t collapse_wide(bajo_ctx* ctx, bajo_bvh* h, uint64_t largest_tree) {
  hipStream_t st = ctx->stream;
  const uint32_t n = (uint32_t)h->n;
  int dev = 0, cus = 0;
  BJ_HIP(hipGetDevice(&dev));
  BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (h->seg_count) {
    const uint32_t S = h->seg_count;
    seg_collapse_root_kernel<<<div_up(S, 256u) < 1024u ? div_up(S, 256u) : 1024u, 256, 0, st>>>(
        h->seg_scene, h->seg_off, S, h->wide, h->wide_src, h->wide_ctl);
  } else {
    const uint32_t* root_dev = (n > 1 && h->root_on_device) ? h->hp_ctl + kHpCtlRoot : nullptr;
    collapse_root_kernel<<<1, 64, 0, st>>>(h->scene, n, n > 1 ? h->root : 0u, root_dev, h->wide, h->wide_src,
                                           h->wide_ctl, ctx->status);
  }
  BJ_HIP(hipGetLastError());
  h->wide_pending = false;
  if (n == 1) {
    h->wide_count = 1;
    return BAJO_OK;
  }
  // grid: enough waves for the widest level (~n/3 nodes), capped at a few per CU
  unsigned grid = div_up(n / 3 + 1, kCollapseBlock);
  const unsigned cap_grid = (unsigned)cus * 8;
  if (grid > cap_grid) grid = cap_grid;
  if (grid < 1) grid = 1;
  const uint32_t levels = collapse_grid_levels(largest_tree, h->wide_leaf, h->node_width);
  for (uint32_t l = 0; l < levels; ++l) {
    collapse_level_kernel<<<grid, kCollapseBlock, 0, st>>>(h->nodes, h->leaf_counts, h->wide, h->wide_src,
                                                           h->wide_ctl, (uint32_t)h->wide_capacity,
                                                           (uint32_t)h->wide_leaf, h->node_width);
    level_advance_kernel<<<1, 64, 0, st>>>(h->wide_ctl, (uint32_t)h->wide_capacity, ctx->status);
  }
  collapse_tail_kernel<<<1, kTailBlock, 0, st>>>(h->nodes, h->leaf_counts, h->wide, h->wide_src, h->wide_ctl,
                                                 (uint32_t)h->wide_capacity, (uint32_t)h->wide_leaf, h->node_width,
                                                 ctx->status);
  BJ_HIP(hipGetLastError());
  h->wide_pending = true;
  return BAJO_OK;
}

// The node count the last collapse left in wide_ctl (clipped to the capacity).
uint32_t collapse_read_count(const uint32_t* host_ctl, uint64_t capacity) {
  const uint32_t c = host_ctl[kCtlCount];
  return (uint64_t)c > capacity ? (uint32_t)capacity : c;
}

}  // namespace bj
