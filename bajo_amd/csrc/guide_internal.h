// Internal interface of the path-guiding network shared with the renderer.
#pragma once

#include "guide_math.h"

// The network, its Adam state and its training buffers (device memory).
struct bajo_guide {
  int device = 0;
  bajo_guide_options opt{};
  uint32_t step = 0;             // Adam steps taken
  float* params = nullptr;       // f32[kGuideParams] master weights
  float* adam_m = nullptr;       // first moments
  float* adam_v = nullptr;       // second moments
  uint8_t* pack = nullptr;       // bf16 weight images (forward + transposed) and f32 biases
  float* ema = nullptr;          // exponential moving average of the weights (what renders sample from)
  uint8_t* infer_pack = nullptr; // the same images of `ema`
  // training batch buffers (bf16, batch-contiguous rows), for `batch_cap` samples
  uint32_t batch_cap = 0;
  uint16_t* x0t = nullptr;       // [224][batch_cap] encoded inputs, row 216 = 1
  uint16_t* h1t = nullptr;       // [80][batch_cap] hidden activations, row 64 = 1
  uint16_t* h2t = nullptr;
  uint16_t* h3t = nullptr;
  uint16_t* d1t = nullptr;       // [64][batch_cap] loss gradients of the hidden layers
  uint16_t* d2t = nullptr;
  uint16_t* d3t = nullptr;
  uint16_t* dot = nullptr;       // [48][batch_cap] loss gradients of the outputs
  float* tile_part = nullptr;    // per 16-sample tile: sum of weights, weighted loss
  float* grad_part = nullptr;    // per batch chunk and gradient tile: 16 x 16 partial sums
  // rendering: training records of one frame and the samples made from them
  uint64_t record_cap = 0;       // path-bounce slots
  float* records = nullptr;      // kRecordPlanes floats per slot
  float* contrib = nullptr;      // radiance added per path and bounce (rgb)
  uint64_t sample_cap = 0;
  float* samples = nullptr;      // kGuideSampleFloats per training sample
  uint32_t* sample_count = nullptr;
  uint32_t* block_counts = nullptr;  // per collection block: samples, then their offset
  uint64_t block_cap = 0;
  // rendering: the guide queue of one bounce and the network's outputs for it
  uint64_t vertex_cap = 0;
  float* vertices = nullptr;     // kVertexPlanes per entry, 64-entry tiles
  float* outputs = nullptr;      // kGuideOut per entry
  uint64_t sample_cursor = 0;    // next training sample of the collected set
};

namespace bj {

// A training sample: the 9 network inputs, the direction taken and its weight
// (luminance of the radiance that arrived along it over the density it was
// drawn with).
constexpr int kGuideSampleFloats = 13;
// A training record of one path vertex: position, normal, outgoing direction,
// direction taken, its density (0: no continuation) and the path throughput
// after the vertex.
constexpr int kRecordPlanes = 16;

// Vertex planes of the renderer's guide queue, in tiles of 64 entries:
// point 0-2, oriented normal 3-5, ray direction 6-8, throughput 9-11,
// surface id 12 and path id 13 (bit patterns).
constexpr int kVertexPlanes = 14;
BJ_HD uint64_t vplane(uint64_t i, int p) { return ((i >> 6) * kVertexPlanes + (uint64_t)p) * 64u + (i & 63u); }

// Network forward over `count_dev[0]` (capped at `cap`) render vertices read
// from the extension kernel's vertex planes; writes 40 outputs per vertex.
int32_t guide_infer_vertices(hipStream_t stream, const bajo_guide* g, const float* vertex_fields,
                             const uint32_t* count_dev, uint64_t cap, const float* scene_bounds, float* outputs);
// Makes room for a frame of `paths` paths of `depth` bounces: the guide
// queue and outputs, and with `train` the training records, per-bounce
// radiance and samples (records and radiance cleared on the stream).
int32_t guide_prepare_frame(hipStream_t stream, bajo_guide* g, uint64_t paths, uint32_t depth, bool train);
// Turns a frame's records into training samples (g->samples, count in
// g->sample_count[0]).
int32_t guide_collect(hipStream_t stream, bajo_guide* g, uint64_t paths, uint32_t depth, const float* scene_bounds);
// Synchronizes, reads the collected sample count and takes up to
// steps_per_frame Adam steps of batch_size samples, strided over the set.
int32_t guide_train_collected(hipStream_t stream, bajo_guide* g, uint64_t* samples_out, uint32_t* steps_out,
                              float* loss_out);

}  // namespace bj
