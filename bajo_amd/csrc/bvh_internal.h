// Device-side layout of a built triangle BVH (shared by builder and tracer).
#pragma once

#include "bajo_common.h"

// 64-byte binary node: bounds[0..5] = left child box, bounds[6..11] = right
// child box (bajo's node_bounds stride 12, bvh/constants.mojo:38-45), then
// meta = parent, left, right, fence (bajo's node_meta stride 4).
struct __attribute__((aligned(64))) BjNode {
  float b[12];
  uint32_t parent, left, right, fence;
};
static_assert(sizeof(BjNode) == 64, "BjNode must be 64 bytes");

// 48-byte sorted leaf triangle: v0.xyz + prim id, e1.xyz, e2.xyz (bajo's
// TRI_LEAF_PACKED_STRIDE record, bvh/gpu/triangle_bvh.mojo:81-149).
struct __attribute__((aligned(16))) BjTri {
  float v0x, v0y, v0z;
  uint32_t prim;
  float e1x, e1y, e1z, pad0;
  float e2x, e2y, e2z, pad1;
};
static_assert(sizeof(BjTri) == 48, "BjTri must be 48 bytes");

// 96-byte 8-wide node with quantized child boxes.  Child i spans
// [o + lo_i * 2^e, o + hi_i * 2^e] per axis (one spare quantization step of
// padding on each side, so the boxes stay conservative under the traversal's
// fused dequantization).  child[i]: bit 31 = leaf; a leaf holds sorted
// triangles [start, start + count) with start in bits 0..27 and count - 1 in
// bits 28..30; an inner child holds a wide node index.
struct __attribute__((aligned(16))) BjWide {
  float ox, oy, oz;
  uint8_t ex, ey, ez, n;  // biased FP32 exponents of the scales, child count
  uint8_t lo_x[8], lo_y[8], lo_z[8], hi_x[8], hi_y[8], hi_z[8];
  uint32_t child[8];
};
static_assert(sizeof(BjWide) == 96, "BjWide must be 96 bytes");

constexpr uint32_t kWideLeafStartMask = 0x0FFFFFFFu;
// Node widths with the quantized wide layout (4 or 8 live child slots).
BJ_HD bool wide_layout(int node_width) { return node_width == BAJO_WIDTH_WIDE4 || node_width == BAJO_WIDTH_WIDE8; }
constexpr uint32_t kWideMaxLeaf = 8;      // largest leaf child the encoding holds
constexpr uint32_t kWideDefaultLeaf = 4;  // triangles per leaf child ("fat leaf")
constexpr uint64_t kWideMaxPrims = 1ull << 28;

struct bajo_bvh {
  int device;
  uint64_t n;         // triangles
  uint64_t capacity;  // allocated triangle capacity
  uint32_t root;
  BjNode* nodes;       // n-1
  BjTri* tris;         // n
  uint64_t* keys;      // n
  uint32_t* leaf_parent;  // n
  uint32_t* leaf_counts;  // n-1
  uint32_t* flags;        // n-1
  uint32_t* acc;          // 12 ordered-int bounds accumulators
  float* scene;           // 12 floats
  void* sort_ws;
  size_t sort_ws_bytes;
  // 8-wide layout (node_width == 8): collapsed from the binary LBVH
  int node_width;
  BjWide* wide;
  uint32_t* wide_src;   // binary node each wide node was opened from
  uint32_t* wide_ctl;   // wide counter, level begin/end, status
  uint32_t* host_ctl;   // pinned host mirror of wide_ctl
  uint64_t wide_capacity;
  uint32_t wide_count;
  uint32_t wide_leaf;  // largest leaf child, 1..kWideMaxLeaf
  // Builds enqueue without reading anything back: the wide node count stays
  // in wide_ctl and H-PLOC's root in hp_ctl until a host query reads them.
  bool wide_pending;    // wide_count not read back yet (traversals bound by wide_capacity)
  bool root_on_device;  // H-PLOC: traversals read the root from hp_ctl[kHpCtlRoot]
  bool root_pending;    // root not read back yet
  // H-PLOC (method BAJO_BUILD_HPLOC): guide LBVH and cluster workspace
  int method;
  uint32_t hp_threshold;  // merging threshold (16 or 32)
  BjNode* hp_guide;       // n-1 guide nodes (Karras over the sorted keys)
  uint32_t* hp_first;     // n-1 first sorted position of each guide node
  void* hp_clusters;      // n cluster entries of 32 bytes
  uint32_t* hp_counts;    // n-1 cluster list lengths
  uint32_t* hp_pos;       // n depth-first leaf positions
  uint32_t* hp_ctl;       // node counter, status, root
  uint32_t* host_hp_ctl;  // pinned host mirror of hp_ctl
  uint64_t hp_capacity;
  // input kind: triangles f32[9N] (traceable) or leaf boxes f32[6N]; a
  // sphere BVH is a boxes build whose sorted leaf records hold the spheres
  bool boxes;
  bool spheres;
  // segmented build (a BLAS set): seg_count > 0 segments share this storage;
  // segment s owns sorted leaves [off_s, off_{s+1}) and nodes [off_s, off_{s+1} - 1)
  uint32_t seg_count;
  uint64_t seg_capacity;
  uint64_t* seg_off;      // device u64[S + 1]
  uint64_t* seg_off_host; // pinned u64[S + 1]: the offsets of the last upload (its staging buffer)
  uint32_t seg_uploaded;  // segment count of that upload (0 = none)
  hipEvent_t seg_upload;  // recorded after that upload
  uint32_t* seg_acc;      // 12 S ordered-int bounds accumulators
  float* seg_scene;       // 12 S: root box and centroid box of each segment
  uint32_t* seg_prim;     // n: input index of each sorted leaf
  uint32_t* seg_chunk;    // ceil(n / 64): the segment holding position 64 c
  bool seg_chunk_valid;   // seg_chunk describes the offsets last uploaded
  uint64_t seg_leaf_capacity;
  void* seg_sort_ws;
  size_t seg_sort_ws_bytes;
  // traversal: index bounds of the node / triangle arrays and the 8-wide
  // root.  A segment view of a BLAS set (owner != NULL) shares its owner's
  // arrays; its records start at node_base / leaf_base.
  uint32_t trav_nodes, trav_prims, wide_root;
  const bajo_bvh* owner;
  uint64_t node_base, leaf_base;
};

namespace bj {
// hp_ctl word holding H-PLOC's root reference.
constexpr int kHpCtlRoot = 2;
// Collapses h's binary hierarchy into h->wide (W-wide quantized nodes);
// `largest_tree` bounds the leaves of one tree (a segment of a set); wide.hip.
int32_t collapse_wide(bajo_ctx* ctx, bajo_bvh* h, uint64_t largest_tree);
// The node count in a host copy of wide_ctl, clipped to the capacity.
uint32_t collapse_read_count(const uint32_t* host_ctl, uint64_t capacity);
// Reads back what a build left on the device (wide node count, H-PLOC root)
// of h or its owning set; synchronizes the device if anything is pending.
int32_t resolve_build(const bajo_bvh* h);
// Builds h's binary hierarchy with H-PLOC over the sorted keys; hploc.hip.
int32_t build_hploc(bajo_ctx* ctx, bajo_bvh* h, const float* src);
int32_t build_hploc_segmented(bajo_ctx* ctx, bajo_bvh* h, const float* vertices);
// Builds the segments of a BLAS set into h (lbvh.hip); offsets are host u64[S + 1].
int32_t build_segmented(bajo_ctx* ctx, bajo_bvh* h, const float* vertices, const uint64_t* offsets, uint32_t S,
                        bajo_build_timings* timings);
// The shared storage of a BLAS set (lbvh.hip).
const bajo_bvh* blas_set_store(const bajo_blas_set* set);
// Karras topology over sorted 64-bit keys (lbvh.hip); `range_first` may be NULL.
__global__ void karras_kernel(const uint64_t* __restrict__ keys, uint64_t n_u, BjNode* __restrict__ nodes,
                              uint32_t* __restrict__ leaf_parent, uint32_t* __restrict__ leaf_counts,
                              uint32_t* __restrict__ flags, uint32_t* __restrict__ range_first);
// Triangle i's vertices from the f32[9N] input.
BJ_D void load_tri_vertices(const float* __restrict__ v, uint64_t i, v3& a, v3& b, v3& c) {
  const float* t = v + 9 * i;
  a = mk(t[0], t[1], t[2]);
  b = mk(t[3], t[4], t[5]);
  c = mk(t[6], t[7], t[8]);
}
// Leaf i's box: the box of triangle i's vertices (bvh/gpu/triangle_bvh.mojo:
// 711-712, vmin(vmin(v0, v1), v2)) or, for a box input, the six floats
// lo.xyz, hi.xyz of bajo's leaf_bounds (bvh/gpu/builder/binary_layout.mojo).
template <bool kBoxes>
BJ_D void leaf_box(const float* __restrict__ src, uint64_t i, v3& lo, v3& hi) {
  if constexpr (kBoxes) {
    const float* b = src + 6 * i;
    lo = mk(b[0], b[1], b[2]);
    hi = mk(b[3], b[4], b[5]);
  } else {
    v3 a, b, c;
    load_tri_vertices(src, i, a, b, c);
    lo = vmin(vmin(a, b), c);
    hi = vmax(vmax(a, b), c);
  }
}
}  // namespace bj
