// H-PLOC binary BVH builder for gfx950.
//
// Plays the role of bajo's GPU H-PLOC (bvh/gpu/builder/hploc_multi_wave.mojo:
// 187-777, hploc_binary.mojo) and produces the topology of its deterministic
// CPU definition (bvh/cpu/builder/hploc.mojo:450-663): the LBVH over the
// sorted Morton keys guides a bottom-up pass; the second thread to arrive at
// a guide node concatenates the cluster lists of its two halves (left then
// right) and, while the list holds more than the merging threshold (16 as in
// bajo's CPU definition, or 32 as in its GPU build on wave64; 1 at the root),
// runs merge rounds: each cluster picks the neighbour within
// radius 8 whose merged box has the least surface area (right before left,
// nearer first, strict '<'), mutual pairs merge into new nodes in list
// order, and the list is compacted in order.
//
// MI355X design:
//   - one thread per sorted leaf climbs the guide hierarchy; arrival is a
//     relaxed agent-scope counter per guide node, and every datum handed to
//     another wave (cluster lists, their counts, output nodes) is written
//     with write-through stores and read with coherent loads, so no L2
//     writeback or invalidate is needed across the 8 XCDs;
//   - a guide node that needs merging is processed by its whole wave: up to
//     2 x threshold (64) clusters sit one per lane, neighbour boxes move by lane shuffles,
//     mutual pairs are found with one ballot, a new node takes the index
//     (retired leaf position - 1) of the larger lowest leaf position of its
//     two clusters (no shared counter), and the list is compacted with a
//     forward lane permute (ds_permute);
//   - a segmented build (a BLAS set) runs every segment's H-PLOC in the same
//     launch: the guide is the per-segment Karras forest, a guide node whose
//     parent is the sentinel is its segment's root (merged down to one
//     cluster), node numbering by retired leaf positions keeps each
//     segment's nodes in its own slots [first, last), and a fix-up moves each
//     segment's root to slot `first`, where the segmented layout expects it
//     (bajo's segmented H-PLOC, hploc_multi_wave.mojo:126-140, 187-545);
//   - output nodes use the 64-byte binary record of the LBVH; a last pass
//     renumbers the leaves in depth-first order (so every subtree covers a
//     contiguous range of sorted triangles, which the 8-wide collapse and the
//     leaf records rely on), packs the triangles in that order, and writes
//     each node's last leaf position as its fence.
#include "bvh_internal.h"
#include "sort_internal.h"

namespace bj {
namespace {

constexpr int kHpBlock = 256;
constexpr int kHpRadius = 8;
// merging thresholds: bajo's CPU definition (16) and its GPU build on wave64
// (WARP_SIZE / 2 = 32, bvh/gpu/builder/hploc_layout.mojo:6); two lists of at
// most 32 clusters fill the 64 lanes of a wave
constexpr uint32_t kHpThreshold = 16, kHpThresholdMax = 32;
constexpr int kHpMaxLoop = 512;  // climb iterations per wave (depth is far smaller)
// ctl words
constexpr int kHpLeaves = 0, kHpStatus = 1, kHpRoot = kHpCtlRoot;
constexpr uint32_t kHpStatusNoProgress = 1, kHpStatusBadList = 2;

// A cluster: subtree reference (bit 31 = leaf at a sorted position), its
// leaf count and its box; 32 bytes.
struct __attribute__((aligned(16))) HpEntry {
  uint32_t ref, cnt;
  float lo[3], hi[3];
};

BJ_D void store_wt(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
BJ_D void store_wt(float* p, float v) {
  __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __builtin_bit_cast(uint32_t, v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
BJ_D uint32_t load_co(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BJ_D float load_co(const float* p) {
  return __builtin_bit_cast(float, load_co(reinterpret_cast<const uint32_t*>(p)));
}

BJ_D uint64_t pack_u(uint32_t a, uint32_t b) { return (uint64_t)a | ((uint64_t)b << 32); }
BJ_D uint64_t pack_f(float a, float b) {
  return pack_u(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b));
}
BJ_D float lo_f(uint64_t w) { return __builtin_bit_cast(float, (uint32_t)w); }
BJ_D float hi_f(uint64_t w) { return __builtin_bit_cast(float, (uint32_t)(w >> 32)); }

// An entry moves as four 8-byte words: (ref, cnt), (lo.x, lo.y), (lo.z, hi.x), (hi.y, hi.z).
BJ_D void store_entry(HpEntry* p, const HpEntry& e) {
  uint64_t* w = reinterpret_cast<uint64_t*>(p);
  __hip_atomic_store(w + 0, pack_u(e.ref, e.cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 1, pack_f(e.lo[0], e.lo[1]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 2, pack_f(e.lo[2], e.hi[0]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 3, pack_f(e.hi[1], e.hi[2]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
BJ_D HpEntry load_entry(const HpEntry* p) {
  uint64_t* w = reinterpret_cast<uint64_t*>(const_cast<HpEntry*>(p));
  const uint64_t a = __hip_atomic_load(w + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t b = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t c = __hip_atomic_load(w + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t d = __hip_atomic_load(w + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  HpEntry e;
  e.ref = (uint32_t)a;
  e.cnt = (uint32_t)(a >> 32);
  e.lo[0] = lo_f(b); e.lo[1] = hi_f(b);
  e.lo[2] = lo_f(c); e.hi[0] = hi_f(c);
  e.hi[1] = lo_f(d); e.hi[2] = hi_f(d);
  return e;
}

BJ_D void publish_fence() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// bvh/cpu/builder/hploc.mojo:130-142: 2 (dx dy + dx dz + dy dz) of the merged box.
BJ_D float merged_area(const HpEntry& a, const float* o) {
  const float dx = fmax_sel(a.hi[0], o[3]) - fmin_sel(a.lo[0], o[0]);
  const float dy = fmax_sel(a.hi[1], o[4]) - fmin_sel(a.lo[1], o[1]);
  const float dz = fmax_sel(a.hi[2], o[5]) - fmin_sel(a.lo[2], o[2]);
  return 2.0f * (dx * dy + dx * dz + dy * dz);
}

BJ_D uint32_t shfl_u(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }

// Segment of sorted position i (off: u64[S + 1]), by binary search.
BJ_D uint32_t seg_of(const uint64_t* __restrict__ off, uint32_t S, uint64_t i) {
  uint32_t lo = 0, hi = S;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}
BJ_D float shfl_f(float v, int src) { return __shfl(v, src, 64); }
BJ_D uint32_t permute_u(uint32_t dst_lane, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_permute((int)(dst_lane << 2), (int)v);
}
BJ_D float permute_f(uint32_t dst_lane, float v) {
  return __builtin_bit_cast(float, permute_u(dst_lane, __builtin_bit_cast(uint32_t, v)));
}

// Writes output node `idx` = (a, b) with write-through stores; sets the
// children's parent links (inner children; leaves keep theirs in the
// per-position parent array).
BJ_D void write_node(BjNode* __restrict__ nodes, uint32_t* __restrict__ leaf_counts,
                     uint32_t* __restrict__ leaf_parent, uint32_t idx, const HpEntry& a, const HpEntry& b) {
  BjNode* nd = nodes + idx;
  for (int k = 0; k < 3; ++k) {
    store_wt(&nd->b[k], a.lo[k]);
    store_wt(&nd->b[3 + k], a.hi[k]);
    store_wt(&nd->b[6 + k], b.lo[k]);
    store_wt(&nd->b[9 + k], b.hi[k]);
  }
  store_wt(&nd->left, a.ref);
  store_wt(&nd->right, b.ref);
  store_wt(leaf_counts + idx, a.cnt + b.cnt);
  if (a.ref & kLeafFlag) store_wt(leaf_parent + (a.ref & kIndexMask), idx);
  else store_wt(&nodes[a.ref].parent, idx);
  if (b.ref & kLeafFlag) store_wt(leaf_parent + (b.ref & kIndexMask), idx);
  else store_wt(&nodes[b.ref].parent, idx);
}

// The wave processes guide node `g` (all 64 lanes call this together).
BJ_D void process_node(uint32_t g, uint64_t n, const BjNode* __restrict__ guide,
                       const uint32_t* __restrict__ range_first, HpEntry* __restrict__ clusters,
                       uint32_t* __restrict__ ids, uint32_t* __restrict__ counts, BjNode* __restrict__ nodes, uint32_t* __restrict__ leaf_counts,
                       uint32_t* __restrict__ leaf_parent, uint32_t* __restrict__ ctl,
                       const uint64_t* __restrict__ seg_off, uint32_t S, uint32_t* __restrict__ seg_roots,
                       uint32_t threshold) {
  const uint32_t lane = lane_id();
  const uint32_t first = range_first[g], last = guide[g].fence;
  const uint32_t gl = guide[g].left, gr = guide[g].right;
  const uint32_t split = gl & kIndexMask;
  const uint32_t cl = (gl & kLeafFlag) ? 1u : load_co(counts + split);
  const uint32_t cr = (gr & kLeafFlag) ? 1u : load_co(counts + (gr & kIndexMask));
  uint32_t c = cl + cr;
  if (cl > threshold || cr > threshold || cl == 0 || cr == 0 || first > split || split >= last ||
      (uint64_t)last >= n || cl > split - first + 1 || cr > last - split) {
    if (lane == 0) atomicOr(ctl + kHpStatus, kHpStatusBadList);
    return;
  }
  // the hierarchy's (or the segment's) root: no parent in the guide
  const bool root = guide[g].parent == kSentinel;
  const uint32_t thr = root ? 1u : threshold;
  if (c <= thr) {
    // no merging: the list is the left list (already at `first`) followed by
    // the right list, which moves only if the left one is shorter than its range
    if (split + 1 != first + cl && lane < cr) {
      const HpEntry m = load_entry(clusters + split + 1 + lane);
      const uint32_t mid = load_co(ids + split + 1 + lane);
      store_entry(clusters + first + cl + lane, m);
      store_wt(ids + first + cl + lane, mid);
    }
    if (lane == 0) store_wt(counts + g, c);
    publish_fence();
    return;
  }
  HpEntry e{};
  uint32_t id = 0;
  if (lane < cl) {
    e = load_entry(clusters + first + lane);
    id = load_co(ids + first + lane);
  } else if (lane < c) {
    e = load_entry(clusters + split + 1 + (lane - cl));
    id = load_co(ids + split + 1 + (lane - cl));
  }
  for (int round = 0; c > thr && round < 64; ++round) {
    const bool live = lane < c;
    float best_area = kFltMax;
    int best = -1;
    for (int r = 1; r <= kHpRadius; ++r) {
      float o[6];
      const int rs = (int)lane + r;  // right neighbour
      for (int k = 0; k < 3; ++k) {
        o[k] = shfl_f(e.lo[k], rs & 63);
        o[3 + k] = shfl_f(e.hi[k], rs & 63);
      }
      if (live && rs < (int)c) {
        const float a = merged_area(e, o);
        if (a < best_area) {
          best_area = a;
          best = rs;
        }
      }
      const int ls = (int)lane - r;  // left neighbour
      for (int k = 0; k < 3; ++k) {
        o[k] = shfl_f(e.lo[k], ls & 63);
        o[3 + k] = shfl_f(e.hi[k], ls & 63);
      }
      if (live && ls >= 0) {
        const float a = merged_area(e, o);
        if (a < best_area) {
          best_area = a;
          best = ls;
        }
      }
    }
    const int nb_best = (int)shfl_u((uint32_t)best, best < 0 ? (int)lane : best);
    bool mutual = live && best >= 0 && nb_best == (int)lane;
    uint64_t mergers = ballot(mutual && (int)lane < best);
    if (mergers == 0) {
      // no mutual pair (only with degenerate boxes): merge the first two
      // clusters (bvh/cpu/builder/hploc.mojo:470-500)
      best = lane == 0 ? 1 : (lane == 1 ? 0 : -1);
      mutual = lane < 2;
      mergers = 1;
    }
    // partner entry
    const int src = best < 0 ? (int)lane : best;
    const uint32_t pid = shfl_u(id, src);
    HpEntry p;
    p.ref = shfl_u(e.ref, src);
    p.cnt = shfl_u(e.cnt, src);
    for (int q = 0; q < 3; ++q) {
      p.lo[q] = shfl_f(e.lo[q], src);
      p.hi[q] = shfl_f(e.hi[q], src);
    }
    const bool merger = (mergers >> lane) & 1u;
    if (merger) {
      // Node numbering without a shared counter: every cluster carries the
      // lowest sorted leaf position it holds; a merge retires the larger of
      // the two and the new node takes index (retired - 1).  Leaf 0 is never
      // retired, so the n-1 merges use the indices 0..n-2 exactly once.
      const uint32_t idx = (id > pid ? id : pid) - 1u;
      id = id < pid ? id : pid;
      const bool refs_ok = ((e.ref & kLeafFlag) ? (uint64_t)(e.ref & kIndexMask) < n : (uint64_t)e.ref < n - 1) &&
                           ((p.ref & kLeafFlag) ? (uint64_t)(p.ref & kIndexMask) < n : (uint64_t)p.ref < n - 1);
      if ((uint64_t)idx < n - 1 && refs_ok) write_node(nodes, leaf_counts, leaf_parent, idx, e, p);
      else atomicOr(ctl + kHpStatus, kHpStatusBadList);
      e.ref = idx;
      e.cnt += p.cnt;
      for (int q = 0; q < 3; ++q) {
        e.lo[q] = fmin_sel(e.lo[q], p.lo[q]);
        e.hi[q] = fmax_sel(e.hi[q], p.hi[q]);
      }
    }
    // compaction: survivors and new nodes keep their order; the higher
    // member of each merged pair leaves the list
    const bool keep = live && !(mutual && (int)lane > best);
    const uint64_t kept = ballot(keep);
    const uint32_t dst = keep ? mask_rank(kept) : 63u;
    HpEntry m;
    m.ref = permute_u(dst, e.ref);
    m.cnt = permute_u(dst, e.cnt);
    id = permute_u(dst, id);
    for (int q = 0; q < 3; ++q) {
      m.lo[q] = permute_f(dst, e.lo[q]);
      m.hi[q] = permute_f(dst, e.hi[q]);
    }
    e = m;
    c = (uint32_t)__builtin_popcountll(kept);
  }
  if (c > thr) {
    if (lane == 0) atomicOr(ctl + kHpStatus, kHpStatusNoProgress);
    return;
  }
  if (lane < c) {
    store_entry(clusters + first + lane, e);
    store_wt(ids + first + lane, id);
  }
  if (lane == 0) {
    store_wt(counts + g, c);
    if (root) {
      if (seg_roots) store_wt(seg_roots + seg_of(seg_off, S, first), e.ref);
      else store_wt(ctl + kHpRoot, e.ref);
    }
  }
  publish_fence();
}

template <bool kBoxes>
__global__ void __launch_bounds__(kHpBlock) hploc_kernel(const float* __restrict__ vertices,
                                                         const uint64_t* __restrict__ keys, uint64_t n,
                                                         const BjNode* __restrict__ guide,
                                                         const uint32_t* __restrict__ range_first,
                                                         uint32_t* __restrict__ leaf_parent,
                                                         uint32_t* __restrict__ flags, HpEntry* __restrict__ clusters,
                                                         uint32_t* __restrict__ ids,
                                                         uint32_t* __restrict__ counts, BjNode* __restrict__ nodes,
                                                         uint32_t* __restrict__ leaf_counts,
                                                         uint32_t* __restrict__ ctl,
                                                         const uint32_t* __restrict__ gprim,
                                                         const uint64_t* __restrict__ seg_off, uint32_t S,
                                                         uint32_t* __restrict__ seg_roots, uint32_t threshold) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool active = i < n;
  bool pending = false;
  uint32_t cur = kLeafFlag | (uint32_t)i;
  if (active) {
    // a segmented build's keys hold the index inside the segment; gprim the input index
    const uint32_t prim = gprim ? gprim[i] : (uint32_t)keys[i];
    v3 lo, hi;
    leaf_box<kBoxes>(vertices, prim, lo, hi);
    HpEntry e{kLeafFlag | (uint32_t)i, 1u, {lo.x, lo.y, lo.z}, {hi.x, hi.y, hi.z}};
    store_entry(clusters + i, e);
    store_wt(ids + i, (uint32_t)i);
  }
  publish_fence();
  // The current guide node's topology (parent, left, right, fence: the last 16
  // bytes of its record) and first leaf: written by the guide's launch and
  // constant here, so a parent's are loaded together with the arrival atomic.
  uint4 topo = make_uint4(kSentinel, 0u, 0u, 0u);
  uint32_t topo_first = 0;
  for (int it = 0; it < kHpMaxLoop; ++it) {
    if (active && !pending) {
      const uint32_t parent = (cur & kLeafFlag) ? leaf_parent[cur & kIndexMask] : topo.x;
      if (parent == kSentinel || (uint64_t)parent >= n - 1) {
        active = false;
      } else {
        const uint4 ptopo = reinterpret_cast<const uint4*>(guide + parent)[3];
        const uint32_t pfirst = range_first[parent];
        const uint32_t old = __hip_atomic_fetch_add(flags + parent, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) {
          active = false;
        } else {
          cur = parent;
          pending = true;
          topo = ptopo;
          topo_first = pfirst;
        }
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // Nodes whose two lists together stay within the threshold need no
    // merging: each lane finishes its own (appending the right list after
    // the left one) and keeps climbing.  Only merging nodes take the wave.
    if (pending) {
      const uint32_t g = cur;
      const uint32_t first = topo_first, last = topo.w;
      const uint32_t gl = topo.y, gr = topo.z;
      const uint32_t split = gl & kIndexMask;
      const uint32_t cl = (gl & kLeafFlag) ? 1u : load_co(counts + split);
      const uint32_t cr = (gr & kLeafFlag) ? 1u : load_co(counts + (gr & kIndexMask));
      const bool root = topo.x == kSentinel;
      const bool sane = cl >= 1 && cr >= 1 && cl <= threshold && cr <= threshold && first <= split &&
                        split < last && (uint64_t)last < n && cl <= split - first + 1 && cr <= last - split;
      if (sane && !root && cl + cr <= threshold) {
        if (split + 1 != first + cl)
          for (uint32_t k = 0; k < cr; ++k) {
            store_entry(clusters + first + cl + k, load_entry(clusters + split + 1 + k));
            store_wt(ids + first + cl + k, load_co(ids + split + 1 + k));
          }
        store_wt(counts + g, cl + cr);
        pending = false;
      }
    }
    publish_fence();
    uint64_t todo = ballot(pending);
    while (todo) {
      const int leader = __builtin_ctzll(todo);
      todo &= todo - 1;
      const uint32_t g = shfl_u(cur, leader);
      process_node(g, n, guide, range_first, clusters, ids, counts, nodes, leaf_counts, leaf_parent, ctl, seg_off, S,
                   seg_roots, threshold);
      if ((int)lane_id() == leader) {
        pending = false;
        if (topo.x == kSentinel) active = false;  // the leader's own node g
      }
    }
    if (!ballot(active)) return;
  }
  if (active) atomicOr(ctl + kHpStatus, kHpStatusNoProgress);
}

__global__ void hploc_init_kernel(uint32_t* __restrict__ ctl) {
  if (threadIdx.x < 4) ctl[threadIdx.x] = 0;
}

// The root gets no parent from a merge; its leaf count goes to ctl word
// kHpLeaves for the host's completeness check.
__global__ void hploc_root_kernel(BjNode* __restrict__ nodes, uint64_t n, const uint32_t* __restrict__ leaf_counts,
                                  uint32_t* __restrict__ ctl, uint32_t* __restrict__ status) {
  if (threadIdx.x != 0) return;
  const uint32_t root = ctl[kHpRoot];
  const bool valid = !(root & kLeafFlag) && (uint64_t)root < n - 1;
  if (valid) {
    nodes[root].parent = kSentinel;
    ctl[kHpLeaves] = leaf_counts[root];
  }
  // the build's own checks, raised into the context status word
  if (ctl[kHpStatus] != 0 || !valid || (uint64_t)ctl[kHpLeaves] != n) {
    if (!valid) ctl[kHpRoot] = kSentinel;  // traversals then see an empty hierarchy
    atomicMax(status, (uint32_t)BAJO_STATUS_HPLOC_FAILED);
  }
}

// Raises a segmented build's clustering status into the context status word.
__global__ void hploc_status_kernel(const uint32_t* __restrict__ ctl, uint32_t* __restrict__ status) {
  if (threadIdx.x == 0 && ctl[kHpStatus] != 0) atomicMax(status, (uint32_t)BAJO_STATUS_HPLOC_FAILED);
}

__global__ void seg_roots_init_kernel(uint32_t* __restrict__ seg_roots, uint32_t S) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) seg_roots[s] = kSentinel;
}

// Segmented builds: moves segment s's root (any slot of the segment) to slot
// off[s] by swapping it with the node there, relinking both nodes'
// children and the moved node's parent.  One thread per segment; segments
// own disjoint slots.  Checks each root's leaf count.
__global__ void seg_root_fix_kernel(BjNode* __restrict__ nodes, uint32_t* __restrict__ leaf_counts,
                                    uint32_t* __restrict__ leaf_parent, const uint64_t* __restrict__ off, uint32_t S,
                                    const uint32_t* __restrict__ seg_roots, uint32_t* __restrict__ ctl) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const uint64_t first = off[s], cnt = off[s + 1] - first;
  if (cnt < 2) return;  // a one-triangle segment is its own root
  const uint32_t r = seg_roots[s], x = (uint32_t)first;
  if ((r & kLeafFlag) || r < x || (uint64_t)r >= first + cnt - 1 || leaf_counts[r] != cnt) {
    atomicOr(ctl + kHpStatus, kHpStatusBadList);
    return;
  }
  // every reference the fix-up follows must stay inside the segment
  const uint32_t end_node = (uint32_t)(first + cnt - 1), end_leaf = (uint32_t)(first + cnt);
  auto ok = [&](uint32_t ref) {
    return (ref & kLeafFlag) ? ((ref & kIndexMask) >= x && (ref & kIndexMask) < end_leaf) : (ref >= x && ref < end_node);
  };
  if (!ok(nodes[r].left) || !ok(nodes[r].right) || !ok(nodes[x].left) || !ok(nodes[x].right) ||
      (r != x && !ok(nodes[x].parent))) {
    atomicOr(ctl + kHpStatus, kHpStatusBadList);
    return;
  }
  if (r != x) {
    BjNode R = nodes[r], X = nodes[x];
    const uint32_t cR = leaf_counts[r], cX = leaf_counts[x];
    auto remap = [&](uint32_t ref) { return (ref & kLeafFlag) ? ref : ref == r ? x : ref == x ? r : ref; };
    R.left = remap(R.left);
    R.right = remap(R.right);
    X.left = remap(X.left);
    X.right = remap(X.right);
    X.parent = remap(X.parent);
    nodes[x] = R;
    nodes[r] = X;
    leaf_counts[x] = cR;
    leaf_counts[r] = cX;
    const uint32_t moved[2] = {x, r};
    for (int k = 0; k < 2; ++k) {
      const uint32_t y = moved[k];
      const uint32_t kids[2] = {nodes[y].left, nodes[y].right};
      for (int c = 0; c < 2; ++c) {
        const uint32_t ref = kids[c];
        if (ref & kLeafFlag) leaf_parent[ref & kIndexMask] = y;
        else nodes[ref].parent = y;
      }
    }
    // the parent of the node moved to r (unless that parent is the root, whose
    // references were remapped above) still points at x
    const uint32_t p = nodes[r].parent;
    if (p != x) {
      if (nodes[p].left == x) nodes[p].left = r;
      else if (nodes[p].right == x) nodes[p].right = r;
    }
  }
  nodes[x].parent = kSentinel;
}

// Depth-first position of each sorted leaf: the leaves of the left subtrees
// of every ancestor reached from the right come first.  Also writes the
// fence (last depth-first position) of every node whose rightmost leaf this
// is, and which side of its parent the leaf hangs on (bit 31 of `pos`).
__global__ void __launch_bounds__(kHpBlock) dfs_position_kernel(BjNode* nodes, uint64_t n,
                                                                const uint32_t* __restrict__ leaf_parent,
                                                                const uint32_t* __restrict__ leaf_counts,
                                                                uint32_t* __restrict__ pos_out,
                                                                const uint64_t* __restrict__ seg_off, uint32_t S) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t ref = kLeafFlag | (uint32_t)j;
  uint32_t parent = leaf_parent[j];
  // a segment's leaves start at its first sorted position
  uint32_t pos = seg_off ? (uint32_t)seg_off[seg_of(seg_off, S, j)] : 0u;
  bool leaf_is_right = false;
  for (int depth = 0; parent != kSentinel && (uint64_t)parent < n - 1 && depth < 4096; ++depth) {
    const uint32_t l = nodes[parent].left;
    const bool from_right = l != ref;
    if (depth == 0) leaf_is_right = from_right;
    if (from_right) pos += (l & kLeafFlag) ? 1u : ((uint64_t)l < n - 1 ? leaf_counts[l] : 0u);
    ref = parent;
    parent = nodes[parent].parent;
  }
  pos_out[j] = pos | (leaf_is_right ? 0x80000000u : 0u);
  // fences: ancestors of which this leaf is the rightmost
  ref = kLeafFlag | (uint32_t)j;
  parent = leaf_parent[j];
  for (int depth = 0; parent != kSentinel && (uint64_t)parent < n - 1 && depth < 4096; ++depth) {
    if (nodes[parent].right != ref) break;
    nodes[parent].fence = pos;
    ref = parent;
    parent = nodes[parent].parent;
  }
}

// Moves each leaf to its depth-first position: triangle record, key, parent
// link, and the parent's child reference.
template <bool kBoxes>
__global__ void __launch_bounds__(kHpBlock) dfs_emit_kernel(const float* __restrict__ vertices,
                                                            const uint64_t* __restrict__ keys, uint64_t n,
                                                            const uint32_t* __restrict__ pos_in,
                                                            const uint32_t* __restrict__ leaf_parent,
                                                            BjNode* __restrict__ nodes, BjTri* __restrict__ tris,
                                                            uint64_t* __restrict__ keys_out,
                                                            uint32_t* __restrict__ leaf_parent_out,
                                                            const uint32_t* __restrict__ gprim) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t pw = pos_in[j];
  const uint32_t pos = pw & 0x7FFFFFFFu;
  if ((uint64_t)pos >= n) return;
  const uint64_t key = keys[j];
  const uint32_t prim = (uint32_t)key;
  if constexpr (!kBoxes) {
    v3 a, b, c;
    load_tri_vertices(vertices, gprim ? gprim[j] : prim, a, b, c);
    const v3 e1 = sub(b, a), e2 = sub(c, a);
    float4* out = reinterpret_cast<float4*>(tris + pos);
    out[0] = make_float4(a.x, a.y, a.z, __builtin_bit_cast(float, prim));
    out[1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    out[2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
  }
  keys_out[pos] = key;
  const uint32_t parent = leaf_parent[j];
  leaf_parent_out[pos] = parent;
  if (parent != kSentinel && (uint64_t)parent < n - 1) {
    if (pw & 0x80000000u) nodes[parent].right = kLeafFlag | pos;
    else nodes[parent].left = kLeafFlag | pos;
  }
}

}  // namespace

int32_t build_hploc(bajo_ctx* ctx, bajo_bvh* h, const float* vertices) {
  hipStream_t st = ctx->stream;
  const uint64_t n = h->n;
  if (n < 2) return BAJO_OK;  // a single leaf is its own root; the LBVH path packed it
  const uint32_t threshold = h->hp_threshold ? h->hp_threshold : kHpThreshold;
  HpEntry* clusters = reinterpret_cast<HpEntry*>(h->hp_clusters);
  hploc_init_kernel<<<1, 64, 0, st>>>(h->hp_ctl);
  karras_kernel<<<div_up(n - 1, kHpBlock), kHpBlock, 0, st>>>(h->keys, n, h->hp_guide, h->leaf_parent,
                                                               h->leaf_counts, h->flags, h->hp_first);
  // hp_pos holds the clusters' lowest leaf positions until the depth-first pass
  if (h->boxes)
    hploc_kernel<true><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_guide, h->hp_first,
                                                                 h->leaf_parent, h->flags, clusters, h->hp_pos,
                                                                 h->hp_counts, h->nodes, h->leaf_counts, h->hp_ctl,
                                                                 nullptr, nullptr, 0u, nullptr, threshold);
  else
    hploc_kernel<false><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_guide, h->hp_first,
                                                                  h->leaf_parent, h->flags, clusters, h->hp_pos,
                                                                  h->hp_counts, h->nodes, h->leaf_counts, h->hp_ctl,
                                                                  nullptr, nullptr, 0u, nullptr, threshold);
  hploc_root_kernel<<<1, 64, 0, st>>>(h->nodes, n, h->leaf_counts, h->hp_ctl, ctx->status);
  BJ_HIP(hipGetLastError());
  // the cluster area is free now: it holds the reordered keys and leaf parents
  uint64_t* keys_dfs = reinterpret_cast<uint64_t*>(h->hp_clusters);
  uint32_t* parent_dfs = reinterpret_cast<uint32_t*>(keys_dfs + n);
  dfs_position_kernel<<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(h->nodes, n, h->leaf_parent,
                                                                h->leaf_counts, h->hp_pos, nullptr, 0u);
  if (h->boxes)
    dfs_emit_kernel<true><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_pos, h->leaf_parent,
                                                                    h->nodes, h->tris, keys_dfs, parent_dfs, nullptr);
  else
    dfs_emit_kernel<false><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_pos,
                                                                     h->leaf_parent, h->nodes, h->tris, keys_dfs,
                                                                     parent_dfs, nullptr);
  BJ_HIP(hipGetLastError());
  BJ_HIP(hipMemcpyAsync(h->keys, keys_dfs, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
  BJ_HIP(hipMemcpyAsync(h->leaf_parent, parent_dfs, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  // no read-back: traversals and the collapse take the root from hp_ctl, a
  // host query reads it (bajo_bvh_get_info), and a failed build has raised
  // BAJO_STATUS_HPLOC_FAILED for the next sync point
  h->root = kSentinel;
  h->root_on_device = true;
  h->root_pending = true;
  return BAJO_OK;
}

// H-PLOC over the segments of a segmented build: `h->hp_guide` / `h->hp_first`
// already hold the per-segment Karras forest and `h->keys` the sorted keys
// with in-segment indices (`h->seg_prim`: input indices).  Writes every
// segment's nodes (root at its first slot), fences and packed triangles.
int32_t build_hploc_segmented(bajo_ctx* ctx, bajo_bvh* h, const float* vertices) {
  hipStream_t st = ctx->stream;
  const uint64_t n = h->n;
  const uint32_t S = h->seg_count;
  if (n < 2 || n == S) return BAJO_OK;  // only one-triangle segments: packed by the caller
  const uint32_t threshold = h->hp_threshold ? h->hp_threshold : kHpThreshold;
  HpEntry* clusters = reinterpret_cast<HpEntry*>(h->hp_clusters);
  uint32_t* seg_roots = h->seg_acc;  // free after the Morton codes: u32[S] of the 12 S words
  hploc_init_kernel<<<1, 64, 0, st>>>(h->hp_ctl);
  seg_roots_init_kernel<<<div_up(S, kHpBlock), kHpBlock, 0, st>>>(seg_roots, S);
  hploc_kernel<false><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_guide, h->hp_first,
                                                                h->leaf_parent, h->flags, clusters, h->hp_pos,
                                                                h->hp_counts, h->nodes, h->leaf_counts, h->hp_ctl,
                                                                h->seg_prim, h->seg_off, S, seg_roots, threshold);
  seg_root_fix_kernel<<<div_up(S, kHpBlock), kHpBlock, 0, st>>>(h->nodes, h->leaf_counts, h->leaf_parent, h->seg_off,
                                                               S, seg_roots, h->hp_ctl);
  BJ_HIP(hipGetLastError());
  uint64_t* keys_dfs = reinterpret_cast<uint64_t*>(h->hp_clusters);
  uint32_t* parent_dfs = reinterpret_cast<uint32_t*>(keys_dfs + n);
  dfs_position_kernel<<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(h->nodes, n, h->leaf_parent, h->leaf_counts,
                                                                h->hp_pos, h->seg_off, S);
  dfs_emit_kernel<false><<<div_up(n, kHpBlock), kHpBlock, 0, st>>>(vertices, h->keys, n, h->hp_pos, h->leaf_parent,
                                                                   h->nodes, h->tris, keys_dfs, parent_dfs,
                                                                   h->seg_prim);
  BJ_HIP(hipGetLastError());
  BJ_HIP(hipMemcpyAsync(h->keys, keys_dfs, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
  BJ_HIP(hipMemcpyAsync(h->leaf_parent, parent_dfs, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  hploc_status_kernel<<<1, 64, 0, st>>>(h->hp_ctl, ctx->status);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

}  // namespace bj
