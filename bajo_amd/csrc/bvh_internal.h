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
};
