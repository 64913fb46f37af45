// LBVH builder for gfx950.
//
// Replaces build_triangle_bvh[..., LBVH] (bvh/gpu/triangle_bvh.mojo:617-727)
// and build_binary_bvh_with_lbvh (bvh/gpu/builder/lbvh.mojo:332-424).  The
// stages and their bajo counterparts:
//   1. scene bounds: root box and centroid box in one pass over the
//      vertices (bajo writes per-triangle boxes, then reduces them in a
//      host-driven multi-level loop, binary_layout.mojo:128-204,383-415).
//      min/max are folded through order-preserving integer atomics, so the
//      result is independent of block order.
//   2. Morton keys (lbvh.mojo:30-62): key = (morton30 << 32) | prim.  The
//      kernel also fills the digit histogram of the high word for the sort
//      and clears its look-back table, so the sort needs no histogram pass.
//   3. onesweep sort of the high 32 bits only: keys are written in prim
//      order and the sort is stable, which yields the same order as a full
//      64-bit sort (and as bajo's stable (segment, code) pair sort with
//      value = index, lbvh.mojo:94-113) in 4 passes instead of 8.
//   4. Karras topology (lbvh.mojo:164-329), one thread per internal node.
//   5. atomic bottom-up refit (lbvh.mojo:116-161): the second arriver at a
//      node merges both child boxes and climbs.  The same pass packs each
//      sorted leaf's triangle (v0, e1, e2, prim) for traversal.
#include <string.h>

#include <vector>

#include "bvh_internal.h"
#include "sort_internal.h"

namespace bj {
namespace {

constexpr int kBlock = 256;
constexpr int kSortBeginBit = 32;
constexpr int kSortEndBit = 64;
constexpr int kSortPasses = (kSortEndBit - kSortBeginBit) / 8;

BJ_D float ld(const float* p) { return __builtin_nontemporal_load(p); }

BJ_D uint64_t pack2(float a, float b) {
  return (uint64_t)__builtin_bit_cast(uint32_t, a) | ((uint64_t)__builtin_bit_cast(uint32_t, b) << 32);
}
BJ_D void unpack2(uint64_t w, float& a, float& b) {
  a = __builtin_bit_cast(float, (uint32_t)w);
  b = __builtin_bit_cast(float, (uint32_t)(w >> 32));
}

BJ_D void load_tri(const float* __restrict__ v, uint64_t i, v3& a, v3& b, v3& c) {
  const float* t = v + 9 * i;
  a = mk(ld(t + 0), ld(t + 1), ld(t + 2));
  b = mk(ld(t + 3), ld(t + 4), ld(t + 5));
  c = mk(ld(t + 6), ld(t + 7), ld(t + 8));
}

// Leaf box with streaming (non-temporal) loads for the one-pass kernels.
template <bool kBoxes>
BJ_D void leaf_src(const float* __restrict__ src, uint64_t i, v3& lo, v3& hi) {
  if constexpr (kBoxes) {
    const float* b = src + 6 * i;
    lo = mk(ld(b + 0), ld(b + 1), ld(b + 2));
    hi = mk(ld(b + 3), ld(b + 4), ld(b + 5));
  } else {
    v3 a, b, c;
    load_tri(src, i, a, b, c);
    lo = vmin(vmin(a, b), c);
    hi = vmax(vmax(a, b), c);
  }
}

__global__ void __launch_bounds__(256) lbvh_init_kernel(uint32_t* acc, uint32_t* hist, uint32_t* tickets) {
  uint32_t t = threadIdx.x;
  if (t < 12) {
    // (+FLT_MAX, -FLT_MAX) is bajo's empty box (core/aabb.mojo:28-35).
    bool is_min = (t % 6) < 3;
    acc[t] = f2ord(is_min ? kFltMax : -kFltMax);
  }
  for (uint32_t i = t; i < kSortPasses * 256; i += blockDim.x) hist[i] = 0;
  if (t < 8) tickets[t] = 0;
}

// Wave-level min/max folds (select semantics; exact for all non-NaN inputs).
BJ_D float wave_min(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = fmin_sel(x, __shfl_xor(x, off, 64));
  return x;
}
BJ_D float wave_max(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = fmax_sel(x, __shfl_xor(x, off, 64));
  return x;
}

template <bool kBoxes>
__global__ void __launch_bounds__(kBlock) scene_bounds_kernel(const float* __restrict__ vertices, uint64_t n,
                                                              uint32_t* __restrict__ acc) {
  float r[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) r[k] = ((k % 6) < 3) ? kFltMax : -kFltMax;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    v3 lo, hi;
    leaf_src<kBoxes>(vertices, i, lo, hi);
    v3 ce = scale(add(lo, hi), 0.5f);
    r[0] = fmin_sel(r[0], lo.x); r[1] = fmin_sel(r[1], lo.y); r[2] = fmin_sel(r[2], lo.z);
    r[3] = fmax_sel(r[3], hi.x); r[4] = fmax_sel(r[4], hi.y); r[5] = fmax_sel(r[5], hi.z);
    r[6] = fmin_sel(r[6], ce.x); r[7] = fmin_sel(r[7], ce.y); r[8] = fmin_sel(r[8], ce.z);
    r[9] = fmax_sel(r[9], ce.x); r[10] = fmax_sel(r[10], ce.y); r[11] = fmax_sel(r[11], ce.z);
  }
  __shared__ float s[kBlock / 64][12];
  const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    float v = ((k % 6) < 3) ? wave_min(r[k]) : wave_max(r[k]);
    if (lane == 0) s[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    int k = threadIdx.x;
    bool is_min = (k % 6) < 3;
    float v = s[0][k];
    for (int w = 1; w < kBlock / 64; ++w) v = is_min ? fmin_sel(v, s[w][k]) : fmax_sel(v, s[w][k]);
    if (is_min) atomicMin(&acc[k], f2ord(v));
    else atomicMax(&acc[k], f2ord(v));
  }
}

template <bool kBoxes>
__global__ void __launch_bounds__(kBlock) morton_kernel(const float* __restrict__ vertices, uint64_t n,
                                                        const uint32_t* __restrict__ acc, float* __restrict__ scene,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ hist,
                                                        uint32_t* __restrict__ lookback, uint64_t lookback_words) {
  __shared__ uint32_t s_hist[kBlock / 64][kSortPasses * 256];
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < (kBlock / 64) * kSortPasses * 256; i += blockDim.x) (&s_hist[0][0])[i] = 0;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t gstride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gtid; i < lookback_words; i += gstride) lookback[i] = 0;
  if (gtid < 12) scene[gtid] = ord2f(acc[gtid]);
  // bvh/gpu/builder/lbvh.mojo:46-52: cmin and safe_inv(extent) of the centroid box.
  const v3 cmin = mk(ord2f(acc[6]), ord2f(acc[7]), ord2f(acc[8]));
  const v3 cmax = mk(ord2f(acc[9]), ord2f(acc[10]), ord2f(acc[11]));
  const v3 inv = safe_inv(sub(cmax, cmin));
  __syncthreads();
  for (uint64_t i = gtid; i < n; i += gstride) {
    v3 lo, hi;
    leaf_src<kBoxes>(vertices, i, lo, hi);
    v3 q = mul(sub(scale(add(lo, hi), 0.5f), cmin), inv);
    uint32_t code = morton3(q.x, q.y, q.z);
    keys[i] = ((uint64_t)code << 32) | (uint64_t)(uint32_t)i;
#pragma unroll
    for (int p = 0; p < kSortPasses; ++p) atomicAdd(&s_hist[wave][p * 256 + ((code >> (8 * p)) & 0xFFu)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kSortPasses * 256; i += blockDim.x) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) c += s_hist[w][i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// Karras common prefix over (code << 32 | sorted index), -1 outside [0, n).
// The Morton codes (high key words) of a block's neighbourhood staged in LDS:
// the range and split searches of a block's nodes stay mostly within a few
// hundred positions of the block, and a code outside the window is read from
// the keys.  Index j of a hierarchy (a segment's keys start at `base`) is
// global position base + j.
constexpr int64_t kCodeBefore = 512, kCodeAfter = 768;
struct CodeWindow {
  const uint32_t* lds;
  int64_t lo, hi;  // global positions held in lds[0, hi - lo)
  const uint64_t* keys;
  int64_t base;
  BJ_D uint32_t code(int64_t j) const {
    const int64_t g = base + j;
    return (g >= lo && g < hi) ? lds[g - lo] : (uint32_t)(keys[g] >> 32);
  }
};

// Stages the codes of global positions [first - kCodeBefore, first + kCodeAfter)
// of the n keys (clipped) for the block whose first position is `first`.
BJ_D CodeWindow stage_codes(const uint64_t* __restrict__ keys, int64_t n, int64_t first, uint32_t* s_codes) {
  CodeWindow cw;
  cw.lo = first - kCodeBefore < 0 ? 0 : first - kCodeBefore;
  cw.hi = first + kCodeAfter < n ? first + kCodeAfter : n;
  for (int64_t g = cw.lo + threadIdx.x; g < cw.hi; g += blockDim.x) s_codes[g - cw.lo] = (uint32_t)(keys[g] >> 32);
  __syncthreads();
  cw.lds = s_codes;
  cw.keys = keys;
  cw.base = 0;
  return cw;
}

BJ_D int karras_delta(const CodeWindow& keys, int64_t n, int64_t i, uint32_t code_i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t code_j = keys.code(j);
  uint64_t x = ((uint64_t)(code_i ^ code_j) << 32) | (uint64_t)((uint32_t)i ^ (uint32_t)j);
  return x == 0 ? 64 : __builtin_clzll(x);
}

// One Karras node: internal node i of the n sorted keys at `keys`.  Its
// record goes to nodes[node_base + i]; leaf references are leaf_base + the
// sorted position, inner references node_base + the node (both bases are 0
// for a single hierarchy; a segmented build places each segment's nodes and
// leaves at its own bases).
BJ_D void karras_node(const CodeWindow& keys, int64_t n, int64_t i, BjNode* __restrict__ nodes,
                      uint32_t* __restrict__ leaf_parent, uint32_t* __restrict__ leaf_counts,
                      uint32_t* __restrict__ flags, uint32_t* __restrict__ range_first, uint32_t node_base,
                      uint32_t leaf_base) {
  const uint32_t ci = keys.code(i);
  // bvh/gpu/builder/lbvh.mojo:164-191: direction and range.
  int d_next = karras_delta(keys, n, i, ci, i + 1);
  int d_prev = karras_delta(keys, n, i, ci, i - 1);
  int64_t d = d_next < d_prev ? -1 : 1;
  int delta_min = karras_delta(keys, n, i, ci, i - d);
  int64_t lmax = 2;
  while (karras_delta(keys, n, i, ci, i + lmax * d) > delta_min) {
    lmax <<= 1;
    if (lmax > n * 2) break;
  }
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t > 0; t >>= 1)
    if (karras_delta(keys, n, i, ci, i + (l + t) * d) > delta_min) l += t;
  int64_t j = i + l * d;
  int64_t first = i < j ? i : j, last = i < j ? j : i;
  // bvh/gpu/builder/lbvh.mojo:194-218: split by binary search.
  const uint32_t cf = keys.code(first);
  int node_prefix = karras_delta(keys, n, first, cf, last);
  int64_t split = first, step = last - first;
  while (step > 1) {
    step = (step + 1) >> 1;
    int64_t ns = split + step;
    if (ns < last && karras_delta(keys, n, first, cf, ns) > node_prefix) split = ns;
  }
  const uint32_t me = node_base + (uint32_t)i;
  flags[me] = 0;
  leaf_counts[me] = (uint32_t)(last - first + 1);
  if (range_first) range_first[me] = leaf_base + (uint32_t)first;
  if (first == 0 && last == n - 1) nodes[me].parent = kSentinel;
  uint32_t left, right;
  if (split == first) {
    left = kLeafFlag | (leaf_base + (uint32_t)split);
    leaf_parent[leaf_base + split] = me;
  } else {
    left = node_base + (uint32_t)split;
    nodes[node_base + split].parent = me;
  }
  if (split + 1 == last) {
    right = kLeafFlag | (leaf_base + (uint32_t)(split + 1));
    leaf_parent[leaf_base + split + 1] = me;
  } else {
    right = node_base + (uint32_t)(split + 1);
    nodes[node_base + split + 1].parent = me;
  }
  nodes[me].left = left;
  nodes[me].right = right;
  nodes[me].fence = leaf_base + (uint32_t)last;
}

}  // namespace

// Also the guide hierarchy of the H-PLOC builder, which needs each node's
// first sorted position (`range_first`, NULL for the plain LBVH).
__global__ void __launch_bounds__(kBlock) karras_kernel(const uint64_t* __restrict__ keys, uint64_t n_u,
                                                        BjNode* __restrict__ nodes, uint32_t* __restrict__ leaf_parent,
                                                        uint32_t* __restrict__ leaf_counts,
                                                        uint32_t* __restrict__ flags,
                                                        uint32_t* __restrict__ range_first) {
  __shared__ uint32_t s_codes[kCodeBefore + kCodeAfter];
  const int64_t n = (int64_t)n_u;
  const CodeWindow cw = stage_codes(keys, n, (int64_t)blockIdx.x * blockDim.x, s_codes);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  karras_node(cw, n, i, nodes, leaf_parent, leaf_counts, flags, range_first, 0u, 0u);
}

namespace {

// Bottom-up refit, one thread per sorted leaf.  Nodes whose leaf range lies
// inside this block's 256 leaves are merged in LDS (arrival counters and
// child boxes in shared memory; the second arriver writes the node's two
// child boxes with plain stores).  Only nodes whose range crosses a block
// boundary use the cross-XCD hand-off through global memory.
// `gprim` (segmented builds) gives each sorted leaf's input index, whose low
// key word is its index inside the segment; `node_total` bounds the climb.
template <bool kBoxes>
__global__ void __launch_bounds__(kBlock) refit_kernel(const float* __restrict__ vertices,
                                                       const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ gprim, uint64_t n,
                                                       uint64_t node_total, BjNode* __restrict__ nodes,
                                                       uint32_t* __restrict__ leaf_parent,
                                                       const uint32_t* __restrict__ leaf_counts,
                                                       uint32_t* __restrict__ flags, BjTri* __restrict__ tris,
                                                       uint32_t* __restrict__ status) {
  __shared__ uint32_t s_flag[kBlock];
  __shared__ float s_box[kBlock][2][6];
  s_flag[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t prim = (uint32_t)keys[s];
  const uint64_t src = gprim ? (uint64_t)gprim[s] : (uint64_t)prim;
  v3 lo, hi;
  if constexpr (kBoxes) {
    leaf_src<true>(vertices, src, lo, hi);
  } else {
    v3 a, b, c;
    load_tri(vertices, src, a, b, c);
    v3 e1 = sub(b, a), e2 = sub(c, a);
    float4* out = reinterpret_cast<float4*>(tris + s);
    out[0] = make_float4(a.x, a.y, a.z, __builtin_bit_cast(float, prim));
    out[1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    out[2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    // bvh/gpu/triangle_bvh.mojo:711-712: leaf box = vmin(vmin(v0,v1),v2).
    lo = vmin(vmin(a, b), c);
    hi = vmax(vmax(a, b), c);
  }
  if (n < 2) {
    leaf_parent[0] = kSentinel;  // single-leaf segment, lbvh.mojo:246-248
    return;
  }
  const uint64_t block_first = (uint64_t)blockIdx.x * blockDim.x;
  const uint64_t block_last = block_first + blockDim.x - 1 < n - 1 ? block_first + blockDim.x - 1 : n - 1;
  uint32_t cur = kLeafFlag | (uint32_t)s;
  uint32_t parent = leaf_parent[s];
  // A node's topology (parent, left, right, fence: its last 16 bytes) and leaf
  // count were written by the topology launch and do not change here, so the
  // next level's are loaded while this level's boxes are exchanged.
  uint4 topo = make_uint4(kSentinel, kSentinel, kSentinel, 0u);
  uint32_t count = 0;
  if (parent < node_total) {
    topo = reinterpret_cast<const uint4*>(nodes + parent)[3];
    count = leaf_counts[parent];
  }
  // A Karras hierarchy over 64-bit keys is at most 64 levels deep; the index
  // and depth checks only keep a corrupted hierarchy from leaving the arrays,
  // and a climb they stop raises BAJO_STATUS_REFIT_CORRUPT.
  int depth = 0;
  for (; parent != kSentinel && depth < 128; ++depth) {
    if (parent >= node_total) {
      atomicMax(status, (uint32_t)BAJO_STATUS_REFIT_CORRUPT);
      break;
    }
    BjNode* node = nodes + parent;
    const uint32_t left = topo.y, right = topo.z;
    const bool is_left = cur == left;
    if (!is_left && cur != right) {
      atomicMax(status, (uint32_t)BAJO_STATUS_REFIT_CORRUPT);
      break;
    }
    // Karras node i is one end of its leaf range [fence + 1 - count, fence]
    const uint64_t last = topo.w, first = last + 1 - count;
    const uint32_t next = topo.x;
    uint4 next_topo = make_uint4(kSentinel, kSentinel, kSentinel, 0u);
    uint32_t next_count = 0;
    if (next < node_total) {
      next_topo = reinterpret_cast<const uint4*>(nodes + next)[3];
      next_count = leaf_counts[next];
    }
    v3 olo, ohi;
    if (first >= block_first && last <= block_last && parent >= block_first && parent - block_first < kBlock) {
      // every leaf below this node belongs to this block
      float* mine = s_box[parent - block_first][is_left ? 0 : 1];
      mine[0] = lo.x; mine[1] = lo.y; mine[2] = lo.z; mine[3] = hi.x; mine[4] = hi.y; mine[5] = hi.z;
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0xC07F);  // LDS writes done (lgkmcnt 0)
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const uint32_t old = atomicAdd(&s_flag[parent - block_first], 1u);
      if (old == 0) break;
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const float* o = s_box[parent - block_first][is_left ? 1 : 0];
      olo = mk(o[0], o[1], o[2]);
      ohi = mk(o[3], o[4], o[5]);
      // both child boxes of the node
      float* l = node->b;
      const v3 llo = is_left ? lo : olo, lhi = is_left ? hi : ohi;
      const v3 rlo = is_left ? olo : lo, rhi = is_left ? ohi : hi;
      reinterpret_cast<float4*>(l)[0] = make_float4(llo.x, llo.y, llo.z, lhi.x);
      reinterpret_cast<float4*>(l)[1] = make_float4(lhi.y, lhi.z, rlo.x, rlo.y);
      reinterpret_cast<float4*>(l)[2] = make_float4(rlo.z, rhi.x, rhi.y, rhi.z);
    } else {
      // Publish this child's box with write-through (agent-scope) stores and
      // wait until they are performed before the arrival counter is bumped.
      // This is the release half of the hand-off without an L2 writeback:
      // the shared data itself never sits in a non-coherent cache line.
      // Boxes move as three 8-byte words (child halves start at byte 0 and 24).
      uint64_t* mine = reinterpret_cast<uint64_t*>(node->b + (is_left ? 0 : 6));
      const uint64_t w0 = pack2(lo.x, lo.y), w1 = pack2(lo.z, hi.x), w2 = pack2(hi.y, hi.z);
      __hip_atomic_store(mine + 0, w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(mine + 1, w1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(mine + 2, w2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const uint32_t old = __hip_atomic_fetch_add(flags + parent, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old == 0) break;
      // Second arriver: the sibling's stores completed before its increment,
      // which returned before ours; coherent loads read them.
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const uint64_t* other = reinterpret_cast<const uint64_t*>(node->b + (is_left ? 6 : 0));
      float o[6];
      unpack2(__hip_atomic_load(other + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), o[0], o[1]);
      unpack2(__hip_atomic_load(other + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), o[2], o[3]);
      unpack2(__hip_atomic_load(other + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), o[4], o[5]);
      olo = mk(o[0], o[1], o[2]);
      ohi = mk(o[3], o[4], o[5]);
    }
    // _load_and_union_node_bounds (binary_layout.mojo:95-101): merge(left, right).
    if (is_left) {
      lo = vmin(lo, olo);
      hi = vmax(hi, ohi);
    } else {
      lo = vmin(olo, lo);
      hi = vmax(ohi, hi);
    }
    cur = parent;
    parent = next;
    topo = next_topo;
    count = next_count;
  }
  if (depth == 128) atomicMax(status, (uint32_t)BAJO_STATUS_REFIT_CORRUPT);
}

// ---- segmented builds (bajo's BLAS sets) ------------------------------------
// Segment s owns sorted leaves [off_s, off_{s+1}) (every segment nonempty)
// and internal nodes [off_s, off_{s+1} - 1): its hierarchy is the
// single-hierarchy layout of its own triangles, shifted by off_s.  Node slot
// off_{s+1} - 1 stays unused, so a node's index is one end of its leaf range
// as in a single hierarchy, which keeps the refit's block-local merges.

// The segment holding position i: the last s with off[s] <= i
// (off[0] = 0 <= i < off[S]).
BJ_D uint32_t segment_of(const uint64_t* __restrict__ off, uint32_t S, uint64_t i) {
  uint32_t lo = 0, hi = S;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The segment of every 64-position chunk's first position, so that a lookup
// costs one table load and a short forward scan instead of a binary search
// of log2(S) dependent loads (segments of 64 or more triangles need no step).
__global__ void __launch_bounds__(kBlock) seg_chunk_kernel(const uint64_t* __restrict__ off, uint32_t S, uint64_t n,
                                                           uint32_t* __restrict__ chunk) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c * 64 < n) chunk[c] = segment_of(off, S, c * 64);
}

BJ_D uint32_t segment_at(const uint64_t* __restrict__ off, const uint32_t* __restrict__ chunk, uint64_t i) {
  uint32_t s = chunk[i >> 6];
  while (off[s + 1] <= i) ++s;
  return s;
}

__global__ void __launch_bounds__(kBlock) seg_init_kernel(uint32_t* __restrict__ acc, uint32_t S) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < 12ull * S; w += stride)
    acc[w] = f2ord((w % 6) < 3 ? kFltMax : -kFltMax);
}

// Per-segment root and centroid boxes.  Each wave walks a contiguous chunk of
// kSegChunk triangles, 64 at a time (coalesced).  A segmented shuffle
// reduction leaves each run of equal segments' extrema in its first lane;
// the run that reaches the wave's last lane is carried into the next 64 (a
// segment usually spans many of them) and every other run is folded into its
// segment's accumulators by order-preserving integer atomics, so a wave
// issues atomics about once per segment it touches instead of once per 64
// triangles.  Extrema are taken on the order-preserving integer images of
// the coordinates, so the result does not depend on the reduction order.
constexpr uint32_t kSegChunk = 64 * 16;

BJ_D void seg_flush(uint32_t* __restrict__ acc, uint32_t seg, const uint32_t (&r)[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    if ((k % 6) < 3) atomicMin(&acc[12 * (uint64_t)seg + k], r[k]);
    else atomicMax(&acc[12 * (uint64_t)seg + k], r[k]);
  }
}

template <bool kBoxes>
__global__ void __launch_bounds__(kBlock) seg_bounds_kernel(const float* __restrict__ src, uint64_t n,
                                                            const uint64_t* __restrict__ off,
                                                            const uint32_t* __restrict__ chunk,
                                                            uint32_t* __restrict__ acc) {
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const uint32_t lo_init = f2ord(kFltMax), hi_init = f2ord(-kFltMax);
  for (uint64_t c0 = wave * kSegChunk; c0 < n; c0 += waves * kSegChunk) {
    // the carried run: its segment and extrema (the same in every lane)
    uint32_t carry_seg = kSentinel;
    uint32_t carry[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) carry[k] = (k % 6) < 3 ? lo_init : hi_init;
    for (uint32_t it = 0; it < kSegChunk; it += 64) {
      const uint64_t i = c0 + it + lane;
      if (c0 + it >= n) break;
      const bool active = i < n;
      uint32_t r[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) r[k] = (k % 6) < 3 ? lo_init : hi_init;
      uint32_t seg = kSentinel;
      if (active) {
        seg = segment_at(off, chunk, i);
        v3 lo, hi;
        leaf_src<kBoxes>(src, i, lo, hi);
        const v3 ce = scale(add(lo, hi), 0.5f);
        r[0] = f2ord(lo.x); r[1] = f2ord(lo.y); r[2] = f2ord(lo.z);
        r[3] = f2ord(hi.x); r[4] = f2ord(hi.y); r[5] = f2ord(hi.z);
        r[6] = r[9] = f2ord(ce.x);
        r[7] = r[10] = f2ord(ce.y);
        r[8] = r[11] = f2ord(ce.z);
      }
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t os = __shfl_down(seg, d, 64);
        const bool take = lane + d < 64 && os == seg;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const uint32_t o = __shfl_down(r[k], d, 64);
          if (take) r[k] = (k % 6) < 3 ? (o < r[k] ? o : r[k]) : (o > r[k] ? o : r[k]);
        }
      }
      const uint32_t prev = __shfl_up(seg, 1, 64);
      const bool head = active && (lane == 0 || prev != seg);
      const uint64_t heads = __ballot(head);
      const uint32_t last = 63u - (uint32_t)__builtin_clzll(heads);  // lane 0 is always a head here
      // the first run continues the carried one, or the carried run is complete
      if (lane == 0) {
        if (seg == carry_seg) {
#pragma unroll
          for (int k = 0; k < 12; ++k)
            r[k] = (k % 6) < 3 ? (carry[k] < r[k] ? carry[k] : r[k]) : (carry[k] > r[k] ? carry[k] : r[k]);
        } else if (carry_seg != kSentinel) {
          seg_flush(acc, carry_seg, carry);
        }
      }
      // runs that end inside these 64 are complete
      if (head && lane != last) seg_flush(acc, seg, r);
      carry_seg = __shfl(seg, last, 64);
#pragma unroll
      for (int k = 0; k < 12; ++k) carry[k] = __shfl(r[k], last, 64);
    }
    if (lane == 0 && carry_seg != kSentinel) seg_flush(acc, carry_seg, carry);
  }
}

// Sort keys (segment << 30 | morton30) with the input index as value; a
// stable sort of these pairs orders each segment by (code, index), which is
// the order of the single-hierarchy key (code << 32 | index) within it.
template <bool kBoxes>
__global__ void __launch_bounds__(kBlock) seg_morton_kernel(const float* __restrict__ src, uint64_t n,
                                                            const uint64_t* __restrict__ off,
                                                            const uint32_t* __restrict__ chunk, uint32_t S,
                                                            const uint32_t* __restrict__ acc,
                                                            float* __restrict__ scene, uint64_t* __restrict__ keys,
                                                            uint32_t* __restrict__ values) {
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = gtid; w < 12ull * S; w += stride) scene[w] = ord2f(acc[w]);
  for (uint64_t i = gtid; i < n; i += stride) {
    const uint32_t seg = segment_at(off, chunk, i);
    const uint32_t* a = acc + 12 * (uint64_t)seg;
    // bvh/gpu/builder/lbvh.mojo:46-52 with the segment's centroid box
    const v3 cmin = mk(ord2f(a[6]), ord2f(a[7]), ord2f(a[8]));
    const v3 cmax = mk(ord2f(a[9]), ord2f(a[10]), ord2f(a[11]));
    const v3 inv = safe_inv(sub(cmax, cmin));
    v3 lo, hi;
    leaf_src<kBoxes>(src, i, lo, hi);
    const v3 q = mul(sub(scale(add(lo, hi), 0.5f), cmin), inv);
    keys[i] = ((uint64_t)seg << 30) | (uint64_t)morton3(q.x, q.y, q.z);
    values[i] = (uint32_t)i;
  }
}

// Sorted pairs -> the single-hierarchy keys (code << 32 | index inside the
// segment); a one-leaf segment's leaf has no parent.
__global__ void __launch_bounds__(kBlock) seg_keys_kernel(uint64_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ gprim, uint64_t n,
                                                          const uint64_t* __restrict__ off,
                                                          uint32_t* __restrict__ leaf_parent) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t key = keys[k];
  const uint32_t seg = (uint32_t)(key >> 30);
  const uint64_t first = off[seg];
  keys[k] = ((key & 0x3FFFFFFFull) << 32) | (uint64_t)(gprim[k] - (uint32_t)first);
  if (off[seg + 1] - first == 1) leaf_parent[k] = kSentinel;
}

__global__ void __launch_bounds__(kBlock) seg_karras_kernel(const uint64_t* __restrict__ keys, uint64_t n,
                                                            const uint64_t* __restrict__ off,
                                                            const uint32_t* __restrict__ chunk,
                                                            BjNode* __restrict__ nodes,
                                                            uint32_t* __restrict__ leaf_parent,
                                                            uint32_t* __restrict__ leaf_counts,
                                                            uint32_t* __restrict__ flags,
                                                            uint32_t* __restrict__ range_first) {
  __shared__ uint32_t s_codes[kCodeBefore + kCodeAfter];
  CodeWindow cw = stage_codes(keys, (int64_t)n, (int64_t)blockIdx.x * blockDim.x, s_codes);
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j + 1 >= n) return;
  const uint32_t seg = segment_at(off, chunk, j);
  const uint64_t first = off[seg], count = off[seg + 1] - first;
  if (j - first + 1 >= count) return;  // the segment's unused last slot
  cw.base = (int64_t)first;
  karras_node(cw, (int64_t)count, (int64_t)(j - first), nodes, leaf_parent, leaf_counts, flags, range_first,
              (uint32_t)first, (uint32_t)first);
}

// ---- sphere BVHs (bvh/gpu/sphere_bvh.mojo:320-371, 480-530) -------------------
// Sphere.bounds (bvh/types.mojo:122-124): center -/+ radius per axis.
__global__ void __launch_bounds__(kBlock) sphere_boxes_kernel(const float* __restrict__ spheres, uint64_t n,
                                                              float* __restrict__ boxes) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 sp = reinterpret_cast<const float4*>(spheres)[i];
  float* b = boxes + 6 * i;
  b[0] = sp.x - sp.w; b[1] = sp.y - sp.w; b[2] = sp.z - sp.w;
  b[3] = sp.x + sp.w; b[4] = sp.y + sp.w; b[5] = sp.z + sp.w;
}

// The sphere of each sorted leaf into its 48-byte record: center + prim, radius.
__global__ void __launch_bounds__(kBlock) pack_spheres_kernel(const float* __restrict__ spheres,
                                                              const uint64_t* __restrict__ keys, uint64_t n,
                                                              BjTri* __restrict__ recs) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t prim = (uint32_t)keys[k];
  const float4 sp = reinterpret_cast<const float4*>(spheres)[prim];
  float4* out = reinterpret_cast<float4*>(recs + k);
  out[0] = make_float4(sp.x, sp.y, sp.z, __builtin_bit_cast(float, prim));
  out[1] = make_float4(sp.w, 0.0f, 0.0f, 0.0f);
  out[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

int32_t ensure_capacity(bajo_bvh* h, uint64_t n) {
  SortPlan plan = make_sort_plan(n, 8, 0, kSortBeginBit, kSortEndBit);
  if (n <= h->capacity && plan.total <= h->sort_ws_bytes) return BAJO_OK;
  (void)hipFree(h->nodes); (void)hipFree(h->tris); (void)hipFree(h->keys); (void)hipFree(h->leaf_parent); (void)hipFree(h->leaf_counts);
  (void)hipFree(h->flags); (void)hipFree(h->acc); (void)hipFree(h->scene); (void)hipFree(h->sort_ws);
  h->nodes = nullptr; h->tris = nullptr; h->keys = nullptr; h->leaf_parent = nullptr; h->leaf_counts = nullptr;
  h->flags = nullptr; h->acc = nullptr; h->scene = nullptr; h->sort_ws = nullptr;
  h->capacity = 0;
  h->sort_ws_bytes = 0;
  uint64_t internal = n > 1 ? n - 1 : 1;
  BJ_HIP(hipMalloc(&h->nodes, internal * sizeof(BjNode)));
  BJ_HIP(hipMalloc(&h->tris, n * sizeof(BjTri)));
  BJ_HIP(hipMalloc(&h->keys, n * sizeof(uint64_t)));
  BJ_HIP(hipMalloc(&h->leaf_parent, n * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->leaf_counts, internal * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->flags, internal * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->acc, 16 * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->scene, 16 * sizeof(float)));
  BJ_HIP(hipMalloc(&h->sort_ws, plan.total));
  h->capacity = n;
  h->sort_ws_bytes = plan.total;
  return BAJO_OK;
}

void free_wide(bajo_bvh* h) {
  (void)hipFree(h->wide); (void)hipFree(h->wide_src); (void)hipFree(h->wide_ctl); (void)hipHostFree(h->host_ctl);
  h->wide = nullptr; h->wide_src = nullptr; h->wide_ctl = nullptr; h->host_ctl = nullptr;
  h->wide_capacity = 0;
  h->wide_count = 0;
}

// Wide nodes: one per binary node holding more than the leaf size, plus the
// root, so n nodes always suffice.
int32_t ensure_wide_capacity(bajo_bvh* h, uint64_t n) {
  if (n <= h->wide_capacity) return BAJO_OK;
  free_wide(h);
  BJ_HIP(hipMalloc(&h->wide, n * sizeof(BjWide)));
  BJ_HIP(hipMalloc(&h->wide_src, n * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->wide_ctl, 16 * sizeof(uint32_t)));
  BJ_HIP(hipHostMalloc(&h->host_ctl, 16 * sizeof(uint32_t), hipHostMallocDefault));
  h->wide_capacity = n;
  return BAJO_OK;
}

void free_hploc(bajo_bvh* h) {
  (void)hipFree(h->hp_guide); (void)hipFree(h->hp_first); (void)hipFree(h->hp_clusters); (void)hipFree(h->hp_counts); (void)hipFree(h->hp_pos);
  (void)hipFree(h->hp_ctl); (void)hipHostFree(h->host_hp_ctl);
  h->hp_guide = nullptr; h->hp_first = nullptr; h->hp_clusters = nullptr; h->hp_counts = nullptr;
  h->hp_pos = nullptr; h->hp_ctl = nullptr; h->host_hp_ctl = nullptr;
  h->hp_capacity = 0;
}

// H-PLOC workspace: the guide hierarchy, range starts, 32-byte cluster
// entries (later reused for the depth-first keys and parents), list
// lengths and leaf positions.
int32_t ensure_hploc_capacity(bajo_bvh* h, uint64_t n) {
  if (n <= h->hp_capacity) return BAJO_OK;
  free_hploc(h);
  const uint64_t internal = n > 1 ? n - 1 : 1;
  BJ_HIP(hipMalloc(&h->hp_guide, internal * sizeof(BjNode)));
  BJ_HIP(hipMalloc(&h->hp_first, internal * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->hp_clusters, n * 32));
  BJ_HIP(hipMalloc(&h->hp_counts, internal * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->hp_pos, n * sizeof(uint32_t)));
  BJ_HIP(hipMalloc(&h->hp_ctl, 16 * sizeof(uint32_t)));
  BJ_HIP(hipHostMalloc(&h->host_hp_ctl, 16 * sizeof(uint32_t), hipHostMallocDefault));
  h->hp_capacity = n;
  return BAJO_OK;
}

// The staging buffer may still be read by an enqueued upload: wait for it.
void free_segment_staging(bajo_bvh* h) {
  if (h->seg_upload) {
    (void)hipEventSynchronize(h->seg_upload);
    (void)hipEventDestroy(h->seg_upload);
  }
  (void)hipHostFree(h->seg_off_host);
  h->seg_upload = nullptr;
  h->seg_off_host = nullptr;
  h->seg_uploaded = 0;
}

void free_segments(bajo_bvh* h) {
  free_segment_staging(h);
  (void)hipFree(h->seg_off); (void)hipFree(h->seg_acc); (void)hipFree(h->seg_scene); (void)hipFree(h->seg_prim);
  (void)hipFree(h->seg_chunk); (void)hipFree(h->seg_sort_ws);
  h->seg_off = nullptr; h->seg_acc = nullptr; h->seg_scene = nullptr; h->seg_prim = nullptr; h->seg_sort_ws = nullptr;
  h->seg_chunk = nullptr;
  h->seg_chunk_valid = false;
  h->seg_capacity = 0;
  h->seg_leaf_capacity = 0;
  h->seg_sort_ws_bytes = 0;
}

// Segment arrays of a BLAS set: offsets, bounds accumulators and boxes per
// segment; the input index per sorted leaf; the pair-sort workspace.
int32_t ensure_segment_capacity(bajo_bvh* h, uint64_t n, uint32_t S, size_t sort_bytes) {
  if (S > h->seg_capacity) {
    free_segment_staging(h);
    (void)hipFree(h->seg_off); (void)hipFree(h->seg_acc); (void)hipFree(h->seg_scene);
    h->seg_off = nullptr; h->seg_acc = nullptr; h->seg_scene = nullptr;
    h->seg_capacity = 0;
    BJ_HIP(hipHostMalloc(&h->seg_off_host, ((uint64_t)S + 1) * sizeof(uint64_t), hipHostMallocDefault));
    BJ_HIP(hipEventCreateWithFlags(&h->seg_upload, hipEventDisableTiming));
    BJ_HIP(hipMalloc(&h->seg_off, ((uint64_t)S + 1) * sizeof(uint64_t)));
    BJ_HIP(hipMalloc(&h->seg_acc, 12ull * S * sizeof(uint32_t)));
    BJ_HIP(hipMalloc(&h->seg_scene, 12ull * S * sizeof(float)));
    h->seg_capacity = S;
  }
  if (n > h->seg_leaf_capacity) {
    (void)hipFree(h->seg_prim);
    (void)hipFree(h->seg_chunk);
    h->seg_prim = nullptr;
    h->seg_chunk = nullptr;
    h->seg_chunk_valid = false;
    h->seg_leaf_capacity = 0;
    BJ_HIP(hipMalloc(&h->seg_prim, n * sizeof(uint32_t)));
    BJ_HIP(hipMalloc(&h->seg_chunk, (n + 63) / 64 * sizeof(uint32_t)));
    h->seg_leaf_capacity = n;
  }
  if (sort_bytes > h->seg_sort_ws_bytes) {
    (void)hipFree(h->seg_sort_ws);
    h->seg_sort_ws = nullptr;
    h->seg_sort_ws_bytes = 0;
    BJ_HIP(hipMalloc(&h->seg_sort_ws, sort_bytes));
    h->seg_sort_ws_bytes = sort_bytes;
  }
  return BAJO_OK;
}

void free_bvh(bajo_bvh* h) {
  if (!h) return;
  free_segments(h);
  free_hploc(h);
  (void)hipFree(h->nodes); (void)hipFree(h->tris); (void)hipFree(h->keys); (void)hipFree(h->leaf_parent); (void)hipFree(h->leaf_counts);
  (void)hipFree(h->flags); (void)hipFree(h->acc); (void)hipFree(h->scene); (void)hipFree(h->sort_ws);
  free_wide(h);
  delete h;
}

struct StageEvents {
  hipEvent_t ev[7] = {};
  bool on = false;
  int32_t create(bool enable) {
    on = enable;
    if (!on) return BAJO_OK;
    for (auto& e : ev) BJ_HIP(hipEventCreate(&e));
    return BAJO_OK;
  }
  void record(int k, hipStream_t s) {
    if (on) (void)hipEventRecord(ev[k], s);
  }
  ~StageEvents() {
    if (on)
      for (auto& e : ev) (void)hipEventDestroy(e);
  }
};

int32_t build_into(bajo_ctx* ctx, bajo_bvh* h, const float* vertices, uint64_t n, bajo_build_timings* timings) {
  int32_t rc = ensure_capacity(h, n);
  if (rc) return rc;
  if (wide_layout(h->node_width)) {
    rc = ensure_wide_capacity(h, n);
    if (rc) return rc;
  }
  const bool hploc = h->method == BAJO_BUILD_HPLOC && n > 1;
  if (hploc) {
    rc = ensure_hploc_capacity(h, n);
    if (rc) return rc;
  }
  h->n = n;
  h->root = n == 1 ? (kLeafFlag | 0u) : 0u;
  h->root_on_device = false;
  h->root_pending = false;
  h->trav_nodes = (uint32_t)(n - 1);
  h->trav_prims = (uint32_t)n;
  h->wide_root = 0;
  SortPlan plan = make_sort_plan(n, 8, 0, kSortBeginBit, kSortEndBit);
  char* ws = (char*)h->sort_ws;
  uint32_t* hist = (uint32_t*)(ws + plan.off_hist);
  uint32_t* lookback = (uint32_t*)(ws + plan.off_lookback);
  uint32_t* tickets = (uint32_t*)(ws + plan.off_tickets);
  const uint64_t lookback_words = (uint64_t)plan.passes * plan.tiles * 256;
  hipStream_t st = ctx->stream;
  StageEvents ev;
  rc = ev.create(timings != nullptr);
  if (rc) return rc;

  unsigned stream_blocks = div_up(n, (uint64_t)kBlock * 8);
  if (stream_blocks > 4096) stream_blocks = 4096;
  if (stream_blocks < 1) stream_blocks = 1;

  ev.record(0, st);
  lbvh_init_kernel<<<1, 256, 0, st>>>(h->acc, hist, tickets);
  if (h->boxes) scene_bounds_kernel<true><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->acc);
  else scene_bounds_kernel<false><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->acc);
  BJ_HIP(hipGetLastError());
  ev.record(1, st);
  if (h->boxes)
    morton_kernel<true><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->acc, h->scene, h->keys, hist, lookback,
                                                          lookback_words);
  else
    morton_kernel<false><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->acc, h->scene, h->keys, hist, lookback,
                                                           lookback_words);
  BJ_HIP(hipGetLastError());
  ev.record(2, st);
  rc = onesweep_run(ctx, plan, h->keys, nullptr, ws, true);
  if (rc) return rc;
  ev.record(3, st);
  if (hploc) {
    // guide topology, H-PLOC clustering and the depth-first leaf order
    ev.record(4, st);
    rc = build_hploc(ctx, h, vertices);
    if (rc) return rc;
  } else {
    if (n > 1) {
      karras_kernel<<<div_up(n - 1, kBlock), kBlock, 0, st>>>(h->keys, n, h->nodes, h->leaf_parent,
                                                             h->leaf_counts, h->flags, nullptr);
      BJ_HIP(hipGetLastError());
    }
    ev.record(4, st);
    if (h->boxes)
      refit_kernel<true><<<div_up(n, kBlock), kBlock, 0, st>>>(vertices, h->keys, nullptr, n, n - 1, h->nodes,
                                                               h->leaf_parent, h->leaf_counts, h->flags, h->tris,
                                                               ctx->status);
    else
      refit_kernel<false><<<div_up(n, kBlock), kBlock, 0, st>>>(vertices, h->keys, nullptr, n, n - 1, h->nodes,
                                                                h->leaf_parent, h->leaf_counts, h->flags, h->tris,
                                                                ctx->status);
    BJ_HIP(hipGetLastError());
  }
  ev.record(5, st);
  if (wide_layout(h->node_width)) {
    rc = collapse_wide(ctx, h, n);
    if (rc) return rc;
  } else {
    h->wide_count = 0;
    h->wide_pending = false;
  }
  ev.record(6, st);
  if (timings) {
    BJ_HIP(hipEventSynchronize(ev.ev[6]));
    float ms[6];
    for (int k = 0; k < 6; ++k) BJ_HIP(hipEventElapsedTime(&ms[k], ev.ev[k], ev.ev[k + 1]));
    timings->bounds_ns = (uint64_t)(ms[0] * 1e6);
    timings->morton_ns = (uint64_t)(ms[1] * 1e6);
    timings->sort_ns = (uint64_t)(ms[2] * 1e6);
    timings->topology_ns = (uint64_t)(ms[3] * 1e6);
    timings->refit_ns = (uint64_t)(ms[4] * 1e6);
    timings->collapse_ns = (uint64_t)(ms[5] * 1e6);
    float total;
    BJ_HIP(hipEventElapsedTime(&total, ev.ev[0], ev.ev[6]));
    timings->total_ns = (uint64_t)(total * 1e6);
  }
  return BAJO_OK;
}

// Puts a set's segment offsets on the device without a host wait: offsets
// equal to the last upload's are already there; new ones go through the
// pinned staging buffer, after any earlier upload from it has run.  While the
// stream is being captured into a graph, the offsets must be unchanged.
int32_t upload_segment_offsets(hipStream_t st, bajo_bvh* h, const uint64_t* offsets, uint32_t S) {
  const size_t bytes = ((size_t)S + 1) * sizeof(uint64_t);
  if (h->seg_uploaded == S && memcmp(h->seg_off_host, offsets, bytes) == 0) return BAJO_OK;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  BJ_HIP(hipStreamIsCapturing(st, &cap));
  BJ_REQUIRE(cap == hipStreamCaptureStatusNone,
             "set build under graph capture: the segment offsets differ from the last build's");
  if (h->seg_uploaded) BJ_HIP(hipEventSynchronize(h->seg_upload));
  memcpy(h->seg_off_host, offsets, bytes);
  BJ_HIP(hipMemcpyAsync(h->seg_off, h->seg_off_host, bytes, hipMemcpyHostToDevice, st));
  BJ_HIP(hipEventRecord(h->seg_upload, st));
  h->seg_uploaded = S;
  h->seg_chunk_valid = false;
  return BAJO_OK;
}

}  // namespace

int32_t build_entry(bajo_ctx* ctx, const float* vertices, uint64_t n, const bajo_build_options* options,
                    bajo_build_timings* timings, bajo_bvh** out, bool boxes);

int32_t resolve_build(const bajo_bvh* view) {
  bajo_bvh* h = const_cast<bajo_bvh*>(view->owner ? view->owner : view);
  if (!h->wide_pending && !h->root_pending) return BAJO_OK;
  int prev = 0;
  BJ_HIP(hipGetDevice(&prev));
  hipError_t e = hipSetDevice(h->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && h->wide_pending)
    e = hipMemcpy(h->host_ctl, h->wide_ctl, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess && h->root_pending)
    e = hipMemcpy(h->host_hp_ctl, h->hp_ctl, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return hip_fail(e, "reading back the build's sizes", __FILE__, __LINE__);
  if (h->wide_pending) h->wide_count = collapse_read_count(h->host_ctl, h->wide_capacity);
  if (h->root_pending) h->root = h->host_hp_ctl[kHpCtlRoot];
  h->wide_pending = false;
  h->root_pending = false;
  return BAJO_OK;
}

// bajo's segmented BLAS build (bvh/gpu/triangle_bvh.mojo:366-491 over
// builder/segmented_build.mojo:89-162 and builder/lbvh.mojo:30-62, 235-329):
// one pass of every stage over all segments -- per-segment boxes, a stable
// (segment, code) pair sort, Karras per segment, one shared refit and an
// optional 8-wide collapse whose level 0 holds the S segment roots.
int32_t build_segmented(bajo_ctx* ctx, bajo_bvh* h, const float* vertices, const uint64_t* offsets, uint32_t S,
                        bajo_build_timings* timings) {
  const uint64_t n = offsets[S];
  int32_t rc = ensure_capacity(h, n);
  if (rc) return rc;
  if (wide_layout(h->node_width)) {
    rc = ensure_wide_capacity(h, n);
    if (rc) return rc;
  }
  if (h->method == BAJO_BUILD_HPLOC) {
    rc = ensure_hploc_capacity(h, n);
    if (rc) return rc;
  }
  int seg_bits = 0;
  while ((1ull << seg_bits) < S) ++seg_bits;
  const SortPlan plan = make_sort_plan(n, 8, 1, 0, 30 + seg_bits);
  rc = ensure_segment_capacity(h, n, S, plan.total);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  rc = upload_segment_offsets(st, h, offsets, S);
  if (rc) return rc;
  uint64_t largest = 0;
  for (uint32_t s = 0; s < S; ++s)
    if (offsets[s + 1] - offsets[s] > largest) largest = offsets[s + 1] - offsets[s];
  h->n = n;
  h->root_on_device = false;
  h->root_pending = false;
  h->seg_count = S;
  h->root = kSentinel;  // the store itself is traced through its segment views
  h->trav_nodes = (uint32_t)(n - 1);
  h->trav_prims = (uint32_t)n;
  h->wide_root = 0;
  StageEvents ev;
  rc = ev.create(timings != nullptr);
  if (rc) return rc;
  unsigned stream_blocks = div_up(n, (uint64_t)kBlock * 8);
  if (stream_blocks > 4096) stream_blocks = 4096;
  if (stream_blocks < 1) stream_blocks = 1;
  unsigned seg_blocks = div_up(12ull * S, (uint64_t)kBlock);
  if (seg_blocks > 1024) seg_blocks = 1024;
  if (!h->seg_chunk_valid) {
    // once per set of offsets (rebuilds with the same offsets keep the table)
    seg_chunk_kernel<<<div_up((n + 63) / 64, (uint64_t)kBlock), kBlock, 0, st>>>(h->seg_off, S, n, h->seg_chunk);
    BJ_HIP(hipGetLastError());
    h->seg_chunk_valid = true;
  }
  ev.record(0, st);
  seg_init_kernel<<<seg_blocks, kBlock, 0, st>>>(h->seg_acc, S);
  seg_bounds_kernel<false><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->seg_off, h->seg_chunk, h->seg_acc);
  BJ_HIP(hipGetLastError());
  ev.record(1, st);
  seg_morton_kernel<false><<<stream_blocks, kBlock, 0, st>>>(vertices, n, h->seg_off, h->seg_chunk, S, h->seg_acc,
                                                             h->seg_scene, h->keys, h->seg_prim);
  BJ_HIP(hipGetLastError());
  ev.record(2, st);
  rc = onesweep_run(ctx, plan, h->keys, h->seg_prim, h->seg_sort_ws, false);
  if (rc) return rc;
  seg_keys_kernel<<<div_up(n, kBlock), kBlock, 0, st>>>(h->keys, h->seg_prim, n, h->seg_off, h->leaf_parent);
  BJ_HIP(hipGetLastError());
  ev.record(3, st);
  const bool hploc = h->method == BAJO_BUILD_HPLOC && n > S;
  if (n > S) {
    // the LBVH writes its nodes directly; H-PLOC's guide is the same forest
    seg_karras_kernel<<<div_up(n - 1, kBlock), kBlock, 0, st>>>(h->keys, n, h->seg_off, h->seg_chunk,
                                                                hploc ? h->hp_guide : h->nodes, h->leaf_parent,
                                                                h->leaf_counts, h->flags,
                                                                hploc ? h->hp_first : nullptr);
    BJ_HIP(hipGetLastError());
  }
  ev.record(4, st);
  if (hploc) {
    // bvh/gpu/builder/segmented_build.mojo:117-130 with build_method HPLOC;
    // its depth-first pass also packs the one-triangle segments
    rc = build_hploc_segmented(ctx, h, vertices);
    if (rc) return rc;
  } else {
    refit_kernel<false><<<div_up(n, kBlock), kBlock, 0, st>>>(vertices, h->keys, h->seg_prim, n, n - 1, h->nodes,
                                                              h->leaf_parent, h->leaf_counts, h->flags, h->tris,
                                                              ctx->status);
    BJ_HIP(hipGetLastError());
  }
  ev.record(5, st);
  if (wide_layout(h->node_width)) {
    rc = collapse_wide(ctx, h, largest);
    if (rc) return rc;
  } else {
    h->wide_count = 0;
    h->wide_pending = false;
  }
  ev.record(6, st);
  if (timings) {
    BJ_HIP(hipEventSynchronize(ev.ev[6]));
    float ms[6];
    for (int k = 0; k < 6; ++k) BJ_HIP(hipEventElapsedTime(&ms[k], ev.ev[k], ev.ev[k + 1]));
    timings->bounds_ns = (uint64_t)(ms[0] * 1e6);
    timings->morton_ns = (uint64_t)(ms[1] * 1e6);
    timings->sort_ns = (uint64_t)(ms[2] * 1e6);
    timings->topology_ns = (uint64_t)(ms[3] * 1e6);
    timings->refit_ns = (uint64_t)(ms[4] * 1e6);
    timings->collapse_ns = (uint64_t)(ms[5] * 1e6);
    float total;
    BJ_HIP(hipEventElapsedTime(&total, ev.ev[0], ev.ev[6]));
    timings->total_ns = (uint64_t)(total * 1e6);
  }
  return BAJO_OK;
}
}  // namespace bj

extern "C" {

int32_t bajo_build_triangle_bvh(bajo_ctx* ctx, const float* vertices, uint64_t n, int32_t method,
                                bajo_build_timings* timings, bajo_bvh** out) {
  bajo_build_options opt{method, BAJO_WIDTH_BINARY, 0, 0};
  return bajo_build_triangle_bvh_ex(ctx, vertices, n, &opt, timings, out);
}

int32_t bajo_build_triangle_bvh_ex(bajo_ctx* ctx, const float* vertices, uint64_t n,
                                   const bajo_build_options* options, bajo_build_timings* timings,
                                   bajo_bvh** out) {
  return bj::build_entry(ctx, vertices, n, options, timings, out, false);
}

int32_t bajo_build_binary_bounds(bajo_ctx* ctx, const float* leaf_bounds, uint64_t n,
                                 const bajo_build_options* options, bajo_build_timings* timings, bajo_bvh** out) {
  BJ_REQUIRE(out == nullptr || *out == nullptr || !(*out)->spheres,
             "bajo_build_binary_bounds: a sphere BVH is rebuilt by bajo_build_sphere_bvh");
  return bj::build_entry(ctx, leaf_bounds, n, options, timings, out, true);
}

int32_t bajo_build_sphere_bvh(bajo_ctx* ctx, const float* spheres, uint64_t n, const bajo_build_options* options,
                              bajo_build_timings* timings, bajo_bvh** out) {
  using namespace bj;
  BJ_REQUIRE(ctx && spheres && out, "bajo_build_sphere_bvh: NULL argument");
  BJ_REQUIRE(n > 0, "bajo_build_sphere_bvh: a BVH needs nonempty input");
  BJ_REQUIRE(n <= kSortMaxKeys, "bajo_build_sphere_bvh: n=%llu exceeds 2^30-1 spheres", (unsigned long long)n);
  BJ_REQUIRE(*out == nullptr || (*out)->spheres, "bajo_build_sphere_bvh: the handle is not a sphere BVH");
  BJ_REQUIRE(((uintptr_t)spheres & 15u) == 0, "bajo_build_sphere_bvh: spheres must be 16-byte aligned");
  // the leaf boxes live in the context scratch for the duration of the build
  void* boxes = nullptr;
  int32_t rc = ctx_scratch(ctx, 6 * n * sizeof(float), &boxes);
  if (rc) return rc;
  sphere_boxes_kernel<<<div_up(n, (uint64_t)kBlock), kBlock, 0, ctx->stream>>>(spheres, n, (float*)boxes);
  BJ_HIP(hipGetLastError());
  bajo_bvh* h = *out;
  rc = build_entry(ctx, (const float*)boxes, n, options, timings, &h, true);
  if (rc) return rc;
  pack_spheres_kernel<<<div_up(n, (uint64_t)kBlock), kBlock, 0, ctx->stream>>>(spheres, h->keys, n, h->tris);
  BJ_HIP(hipGetLastError());
  h->spheres = true;
  *out = h;
  return BAJO_OK;
}

}  // extern "C"

namespace bj {

int32_t build_entry(bajo_ctx* ctx, const float* vertices, uint64_t n, const bajo_build_options* options,
                    bajo_build_timings* timings, bajo_bvh** out, bool boxes) {
  const char* fn = boxes ? "bajo_build_binary_bounds" : "bajo_build_triangle_bvh";
  const char* items = boxes ? "boxes" : "triangles";
  bajo_build_options opt = options ? *options : bajo_build_options{BAJO_BUILD_LBVH, BAJO_WIDTH_BINARY, 0, 0};
  BJ_REQUIRE(ctx != nullptr && out != nullptr, "%s: ctx/out is NULL", fn);
  BJ_REQUIRE(opt.method == BAJO_BUILD_LBVH || opt.method == BAJO_BUILD_HPLOC,
             "%s: unsupported build method %d", fn, opt.method);
  BJ_REQUIRE(opt.node_width == BAJO_WIDTH_BINARY || wide_layout(opt.node_width),
             "%s: node_width must be 2, 4 or 8, got %d", fn, opt.node_width);
  if (opt.leaf_size == 0) opt.leaf_size = (int32_t)kWideDefaultLeaf;
  BJ_REQUIRE(opt.leaf_size >= 1 && opt.leaf_size <= (int32_t)kWideMaxLeaf,
             "%s: leaf_size must be in [1, 8], got %d", fn, opt.leaf_size);
  BJ_REQUIRE(opt.merging_threshold == 0 || opt.merging_threshold == 16 || opt.merging_threshold == 32,
             "%s: merging_threshold must be 0, 16 or 32, got %d", fn, opt.merging_threshold);
  BJ_REQUIRE(n > 0, "%s: a BVH needs nonempty input", fn);
  BJ_REQUIRE(vertices != nullptr, "%s: input array is NULL", fn);
  BJ_REQUIRE(n <= bj::kSortMaxKeys, "%s: n=%llu exceeds 2^30-1 %s", fn,
             (unsigned long long)n, items);
  BJ_REQUIRE(!wide_layout(opt.node_width) || n < kWideMaxPrims,
             "%s: n=%llu exceeds the 8-wide limit of 2^28-1 %s", fn, (unsigned long long)n, items);
  bajo_bvh* h = *out;
  bool fresh = h == nullptr;
  if (fresh) {
    h = new bajo_bvh();
    memset(h, 0, sizeof(*h));
    h->device = ctx->device;
  } else {
    BJ_REQUIRE(h->device == ctx->device, "%s: handle belongs to device %d", fn, h->device);
  }
  BJ_REQUIRE(fresh || h->boxes == boxes, "bajo_build: a handle keeps its input kind (triangles or boxes)");
  BJ_REQUIRE(fresh || (!h->owner && !h->seg_count), "%s: a BLAS set or its segment view cannot be rebuilt here", fn);
  h->node_width = opt.node_width;
  h->method = opt.method;
  h->hp_threshold = opt.merging_threshold ? (uint32_t)opt.merging_threshold : 16u;
  h->wide_leaf = (uint32_t)opt.leaf_size;
  h->boxes = boxes;
  int32_t rc = bj::build_into(ctx, h, vertices, n, timings);
  if (rc) {
    if (fresh) bj::free_bvh(h);
    return rc;
  }
  *out = h;
  return BAJO_OK;
}

}  // namespace bj

extern "C" {

int32_t bajo_bvh_free(bajo_bvh* bvh) {
  BJ_REQUIRE(bvh == nullptr || bvh->owner == nullptr,
             "bajo_bvh_free: a segment view is released with its BLAS set (bajo_blas_set_free)");
  BJ_REQUIRE(bvh == nullptr || bvh->seg_count == 0, "bajo_bvh_free: a BLAS set is released by bajo_blas_set_free");
  bj::free_bvh(bvh);
  return BAJO_OK;
}

namespace {
// A view's node / leaf references as the single-hierarchy layout of its
// segment would hold them (its own records numbered from 0).
uint32_t rebase_ref(uint32_t r, uint64_t node_base, uint64_t leaf_base) {
  if (r == bj::kSentinel) return r;
  if (r & bj::kLeafFlag) return bj::kLeafFlag | (uint32_t)((r & bj::kIndexMask) - leaf_base);
  return (uint32_t)(r - node_base);
}
}  // namespace

int32_t bajo_bvh_get_info(const bajo_bvh* h, bajo_bvh_info* out) {
  BJ_REQUIRE(h != nullptr && out != nullptr, "bajo_bvh_get_info: NULL argument");
  const int32_t rc = bj::resolve_build(h);
  if (rc) return rc;
  const bajo_bvh* store = h->owner ? h->owner : h;
  out->prim_count = h->n;
  out->node_count = h->n > 0 ? h->n - 1 : 0;
  out->root = h->owner ? rebase_ref(h->root, h->node_base, h->leaf_base) : h->root;
  out->node_width = (uint32_t)h->node_width;
  out->nodes = h->nodes ? h->nodes + h->node_base : nullptr;
  out->triangles = h->tris ? h->tris + h->leaf_base : nullptr;
  out->sorted_keys = h->keys ? h->keys + h->leaf_base : nullptr;
  out->leaf_parent = h->leaf_parent ? h->leaf_parent + h->leaf_base : nullptr;
  out->leaf_counts = h->leaf_counts ? h->leaf_counts + h->node_base : nullptr;
  out->scene_bounds = h->scene;
  out->wide_node_count = wide_layout(h->node_width) ? store->wide_count : 0;
  out->wide_nodes = wide_layout(h->node_width) ? h->wide : nullptr;
  return BAJO_OK;
}

int32_t bajo_bvh_copy_wide(bajo_ctx* ctx, const bajo_bvh* h, void* nodes, uint64_t capacity, uint64_t* count) {
  BJ_REQUIRE(ctx != nullptr && h != nullptr, "bajo_bvh_copy_wide: NULL argument");
  BJ_REQUIRE(wide_layout(h->node_width), "bajo_bvh_copy_wide: the BVH was built with node width %d",
             h->node_width);
  BJ_REQUIRE(h->seg_count == 0, "bajo_bvh_copy_wide: copy a BLAS set's segments through their views");
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  const int32_t rc = bj::resolve_build(h);
  if (rc) return rc;
  if (!h->owner) {
    if (count) *count = h->wide_count;
    if (!nodes) return BAJO_OK;
    BJ_REQUIRE(capacity >= h->wide_count, "bajo_bvh_copy_wide: capacity %llu < %u nodes",
               (unsigned long long)capacity, h->wide_count);
    BJ_HIP(hipStreamSynchronize(ctx->stream));
    BJ_HIP(hipMemcpy(nodes, h->wide, (uint64_t)h->wide_count * sizeof(BjWide), hipMemcpyDeviceToHost));
    return BAJO_OK;
  }
  // a segment view: the nodes reachable from its root, depth-first (children
  // in slot order), with inner references renumbered and leaf triangle
  // ranges relative to the segment's first sorted leaf
  std::vector<BjWide> all(h->owner->wide_count);
  BJ_HIP(hipMemcpy(all.data(), h->wide, all.size() * sizeof(BjWide), hipMemcpyDeviceToHost));
  std::vector<uint32_t> order, stack{h->wide_root}, pos(all.size(), bj::kSentinel);
  while (!stack.empty()) {
    const uint32_t k = stack.back();
    stack.pop_back();
    BJ_REQUIRE(k < all.size() && pos[k] == bj::kSentinel, "bajo_bvh_copy_wide: malformed segment hierarchy");
    pos[k] = (uint32_t)order.size();
    order.push_back(k);
    for (int c = (int)all[k].n - 1; c >= 0; --c)
      if (!(all[k].child[c] & bj::kLeafFlag)) stack.push_back(all[k].child[c]);
  }
  if (count) *count = order.size();
  if (!nodes) return BAJO_OK;
  BJ_REQUIRE(capacity >= order.size(), "bajo_bvh_copy_wide: capacity %llu < %zu nodes",
             (unsigned long long)capacity, order.size());
  BjWide* dst = (BjWide*)nodes;
  for (size_t i = 0; i < order.size(); ++i) {
    BjWide w = all[order[i]];
    for (int c = 0; c < (int)w.n; ++c) {
      const uint32_t r = w.child[c];
      if (r & bj::kLeafFlag) {
        const uint32_t start = (r & kWideLeafStartMask) - (uint32_t)h->leaf_base;
        w.child[c] = (r & ~kWideLeafStartMask) | start;
      } else {
        w.child[c] = pos[r];
      }
    }
    dst[i] = w;
  }
  return BAJO_OK;
}

int32_t bajo_bvh_copy_binary(bajo_ctx* ctx, const bajo_bvh* h, uint64_t* sorted_keys, uint32_t* leaf_ids,
                             uint32_t* node_meta, float* node_bounds, uint32_t* leaf_parent, uint32_t* leaf_counts,
                             float* scene_bounds) {
  BJ_REQUIRE(ctx != nullptr && h != nullptr, "bajo_bvh_copy_binary: NULL argument");
  BJ_REQUIRE(h->seg_count == 0, "bajo_bvh_copy_binary: copy a BLAS set's segments through their views");
  const uint64_t n = h->n, m = n > 0 ? n - 1 : 0;
  const uint64_t nb = h->node_base, lb = h->leaf_base;  // 0 unless a segment view
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  uint64_t* keys = (uint64_t*)malloc(n * sizeof(uint64_t));
  BJ_REQUIRE(keys != nullptr, "bajo_bvh_copy_binary: host allocation failed");
  hipError_t e = hipMemcpy(keys, h->keys + lb, n * sizeof(uint64_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) { free(keys); return bj::hip_fail(e, "hipMemcpy(keys)", __FILE__, __LINE__); }
  if (sorted_keys) memcpy(sorted_keys, keys, n * sizeof(uint64_t));
  if (leaf_ids)
    for (uint64_t i = 0; i < n; ++i) leaf_ids[i] = (uint32_t)keys[i];
  free(keys);
  if ((node_meta || node_bounds) && m > 0) {
    BjNode* nodes = (BjNode*)malloc(m * sizeof(BjNode));
    BJ_REQUIRE(nodes != nullptr, "bajo_bvh_copy_binary: host allocation failed");
    e = hipMemcpy(nodes, h->nodes + nb, m * sizeof(BjNode), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { free(nodes); return bj::hip_fail(e, "hipMemcpy(nodes)", __FILE__, __LINE__); }
    for (uint64_t i = 0; i < m; ++i) {
      if (node_bounds) memcpy(node_bounds + 12 * i, nodes[i].b, 12 * sizeof(float));
      if (node_meta) {
        node_meta[4 * i + 0] = rebase_ref(nodes[i].parent, nb, lb);
        node_meta[4 * i + 1] = rebase_ref(nodes[i].left, nb, lb);
        node_meta[4 * i + 2] = rebase_ref(nodes[i].right, nb, lb);
        node_meta[4 * i + 3] = (uint32_t)(nodes[i].fence - lb);
      }
    }
    free(nodes);
  }
  if (leaf_parent) {
    BJ_HIP(hipMemcpy(leaf_parent, h->leaf_parent + lb, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i) leaf_parent[i] = rebase_ref(leaf_parent[i], nb, lb);
  }
  if (leaf_counts && m > 0)
    BJ_HIP(hipMemcpy(leaf_counts, h->leaf_counts + nb, m * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (scene_bounds) BJ_HIP(hipMemcpy(scene_bounds, h->scene, 12 * sizeof(float), hipMemcpyDeviceToHost));
  return BAJO_OK;
}

// ---- BLAS sets ---------------------------------------------------------------

struct bajo_blas_set {
  bajo_bvh* store = nullptr;
  std::vector<bajo_bvh> views;
};

int32_t bajo_build_triangle_blas_set(bajo_ctx* ctx, const float* vertices, const uint64_t* segment_offsets,
                                     uint32_t segment_count, const bajo_build_options* options,
                                     bajo_build_timings* timings, bajo_blas_set** out) {
  using namespace bj;
  const char* fn = "bajo_build_triangle_blas_set";
  bajo_build_options opt = options ? *options : bajo_build_options{BAJO_BUILD_LBVH, BAJO_WIDTH_BINARY, 0, 0};
  BJ_REQUIRE(ctx && vertices && segment_offsets && out, "%s: NULL argument", fn);
  BJ_REQUIRE(segment_count > 0, "%s: a BLAS set needs at least one segment", fn);
  BJ_REQUIRE(opt.method == BAJO_BUILD_LBVH || opt.method == BAJO_BUILD_HPLOC,
             "%s: method must be LBVH (0) or H-PLOC (1), got %d", fn, opt.method);
  BJ_REQUIRE(opt.node_width == BAJO_WIDTH_BINARY || wide_layout(opt.node_width),
             "%s: node_width must be 2, 4 or 8, got %d", fn, opt.node_width);
  if (opt.leaf_size == 0) opt.leaf_size = (int32_t)kWideDefaultLeaf;
  BJ_REQUIRE(opt.leaf_size >= 1 && opt.leaf_size <= (int32_t)kWideMaxLeaf, "%s: leaf_size must be in [1, 8]", fn);
  BJ_REQUIRE(opt.merging_threshold == 0 || opt.merging_threshold == 16 || opt.merging_threshold == 32,
             "%s: merging_threshold must be 0, 16 or 32, got %d", fn, opt.merging_threshold);
  BJ_REQUIRE(segment_offsets[0] == 0, "%s: segment_offsets[0] must be 0", fn);
  for (uint32_t s = 0; s < segment_count; ++s)
    BJ_REQUIRE(segment_offsets[s + 1] > segment_offsets[s], "%s: segment %u is empty (offsets must increase)", fn, s);
  const uint64_t n = segment_offsets[segment_count];
  BJ_REQUIRE(n <= kSortMaxKeys, "%s: %llu triangles exceed 2^30-1", fn, (unsigned long long)n);
  BJ_REQUIRE(!wide_layout(opt.node_width) || n < kWideMaxPrims, "%s: %llu triangles exceed the 8-wide limit",
             fn, (unsigned long long)n);
  bajo_blas_set* set = *out;
  const bool fresh = set == nullptr;
  if (!fresh) {
    // rebuild in place: same segment count, so the views stay where they are
    BJ_REQUIRE(set->store->device == ctx->device, "%s: the set belongs to device %d", fn, set->store->device);
    BJ_REQUIRE(set->views.size() == segment_count, "%s: a rebuild keeps the segment count (%zu, got %u)", fn,
               set->views.size(), segment_count);
  } else {
    set = new bajo_blas_set();
    set->store = new bajo_bvh();
    memset(set->store, 0, sizeof(*set->store));
    set->store->device = ctx->device;
  }
  bajo_bvh* h = set->store;
  h->node_width = opt.node_width;
  h->method = opt.method;
  h->hp_threshold = opt.merging_threshold ? (uint32_t)opt.merging_threshold : 16u;
  h->wide_leaf = (uint32_t)opt.leaf_size;
  int32_t rc = build_segmented(ctx, h, vertices, segment_offsets, segment_count, timings);
  if (rc) {
    if (fresh) {
      free_bvh(h);
      delete set;
    }
    return rc;
  }
  set->views.resize(segment_count);
  for (uint32_t s = 0; s < segment_count; ++s) {
    bajo_bvh& v = set->views[s];
    v = *h;
    const uint64_t first = segment_offsets[s], count = segment_offsets[s + 1] - first;
    v.seg_count = 0;
    v.owner = h;
    v.n = count;
    v.leaf_base = first;
    v.node_base = first;
    v.root = count == 1 ? (kLeafFlag | (uint32_t)first) : (uint32_t)first;
    v.scene = h->seg_scene + 12 * (uint64_t)s;
    v.wide_root = s;
  }
  *out = set;
  return BAJO_OK;
}

}  // extern "C"

const bajo_bvh* bj::blas_set_store(const bajo_blas_set* set) { return set ? set->store : nullptr; }

extern "C" {

int32_t bajo_blas_set_count(const bajo_blas_set* set, uint32_t* count) {
  BJ_REQUIRE(set && count, "bajo_blas_set_count: NULL argument");
  *count = (uint32_t)set->views.size();
  return BAJO_OK;
}

int32_t bajo_blas_set_get(const bajo_blas_set* set, uint32_t index, bajo_bvh** view) {
  BJ_REQUIRE(set && view, "bajo_blas_set_get: NULL argument");
  BJ_REQUIRE(index < set->views.size(), "bajo_blas_set_get: index %u of %zu segments", index, set->views.size());
  *view = const_cast<bajo_bvh*>(&set->views[index]);
  return BAJO_OK;
}

int32_t bajo_blas_set_free(bajo_blas_set* set) {
  if (!set) return BAJO_OK;
  bj::free_bvh(set->store);
  delete set;
  return BAJO_OK;
}

}  // extern "C"
