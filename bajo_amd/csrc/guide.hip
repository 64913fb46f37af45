// Learned path guiding on gfx950: the network's forward pass, its fused
// forward + backward pass, the weight gradient and Adam, all hand-written
// for the BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
// bajo has no counterpart; the network and its output are defined in
// guide_math.h.
//
// MI355X design:
// - Weights live in one device image: BF16 rows padded by 8 elements (so
//   16 lanes reading one 16-byte column slice of 16 rows hit different LDS
//   banks), transposed copies of W2..W4 for the backward pass, FP32 biases.
//   A block copies the image into LDS once (55 KB forward, 83 KB training)
//   and its waves loop over 16-sample tiles.
// - MFMA operand layout (16x16x32): A[i][k] at lane i + 16 (k / 8), element
//   k % 8; B[k][j] at lane j + 16 (k / 8); D[i][j] at lane j + 16 (i / 4),
//   element i % 4.  A row of weights W[j][k..k+7] is one 16-byte LDS load,
//   and so is a row of activations, so every operand is a ds_read_b128.
// - The input encoding is computed straight into A fragments: lane (i, g)
//   of k-step s needs features 32 s + 8 g .. + 7, i.e. four frequencies of
//   one input, so layer 1 reads nothing but the sample's 9 inputs.
// - Training writes activations and deltas transposed (batch-contiguous
//   rows, BF16) and the weight gradient is a batch-reduction GEMM
//   dW = D^T X with the batch as the MFMA K dimension: 111 output tiles x
//   batch chunks of 2048, partial tiles summed in fixed order by the Adam
//   kernel (deterministic, no atomics).
#include <math.h>
#include <string.h>

#include <vector>

#include "guide_internal.h"

namespace bj {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// ---- weight image --------------------------------------------------------------
constexpr int kW1Stride = kGuideK0 + 8;      // 232
constexpr int kRowStride = kGuideHidden + 8;  // 72: rows of 64-wide operands
constexpr int kPackW1 = 0;                                       // [64][232]
constexpr int kPackW2 = kPackW1 + kGuideHidden * kW1Stride;      // [64][72]
constexpr int kPackW3 = kPackW2 + kGuideHidden * kRowStride;     // [64][72]
constexpr int kPackW4 = kPackW3 + kGuideHidden * kRowStride;     // [48][72]
constexpr int kPackFwd = kPackW4 + kGuideOutPad * kRowStride;    // forward weights end
constexpr int kPackW2t = kPackFwd;                               // [64 in][72]: W2^T
constexpr int kPackW3t = kPackW2t + kGuideHidden * kRowStride;   // W3^T
constexpr int kPackW4t = kPackW3t + kGuideHidden * kRowStride;   // [64 hidden][72]: W4^T (outputs < 40)
constexpr int kPackAll = kPackW4t + kGuideHidden * kRowStride;
constexpr int kBiasFloats = 3 * kGuideHidden + kGuideOutPad;     // b1 b2 b3 b4(48)
constexpr size_t kPackBytes = (size_t)kPackAll * 2 + kBiasFloats * 4;
static_assert((kPackAll * 2) % 16 == 0 && (kPackFwd * 2) % 16 == 0, "pack images are 16-byte aligned");

// ---- training layout -----------------------------------------------------------
constexpr int kX0Rows = kGuideK0;            // 224: features, row 216 = 1
constexpr int kHRows = kGuideHidden + 16;    // 80: hidden units, row 64 = 1
constexpr int kDRows = kGuideHidden;         // 64
constexpr int kDoRows = kGuideOutPad;        // 48 (rows 40..47 zero)
constexpr int kChunk = 2048;                 // batch samples per weight-gradient partial
// gradient tiles (16 x 16 of dW[out][in], the bias in the column of the constant input)
constexpr int kT1 = 0, kT2 = 4 * 14, kT3 = kT2 + 4 * 5, kT4 = kT3 + 4 * 5, kTiles = kT4 + 3 * 5;  // 111

BJ_HD uint16_t f2bf(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
BJ_HD float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

BJ_D bf16x8 frag16(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
BJ_D f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// Where parameter p lives: its gradient tile / element and its positions in
// the weight image (forward and, for W2..W4, transposed; -1 = none; biases
// go to the FP32 bias block).
struct ParamSlot {
  int tile, elem, fwd, tr, bias;
};
BJ_HD ParamSlot param_slot(int p) {
  ParamSlot s{0, 0, -1, -1, -1};
  int j, k, base_t, ntiles, fwd_base, fwd_stride, tr_base, bias_base;
  if (p < kGuideOffB1) {
    j = p / kGuideFeatures; k = p - j * kGuideFeatures;
    s.tile = kT1 + (j >> 4) * 14 + (k >> 4); s.elem = (j & 15) * 16 + (k & 15);
    s.fwd = kPackW1 + j * kW1Stride + k;
    return s;
  }
  if (p < kGuideOffW2) {
    j = p - kGuideOffB1; k = kGuideFeatures;
    s.tile = kT1 + (j >> 4) * 14 + (k >> 4); s.elem = (j & 15) * 16 + (k & 15);
    s.bias = j;
    return s;
  }
  if (p < kGuideOffW3) { base_t = kT2; ntiles = 5; fwd_base = kPackW2; tr_base = kPackW2t; bias_base = 64; p -= kGuideOffW2; }
  else if (p < kGuideOffW4) { base_t = kT3; ntiles = 5; fwd_base = kPackW3; tr_base = kPackW3t; bias_base = 128; p -= kGuideOffW3; }
  else { base_t = kT4; ntiles = 5; fwd_base = kPackW4; tr_base = kPackW4t; bias_base = 192; p -= kGuideOffW4; }
  fwd_stride = kRowStride;
  const int rows = base_t == kT4 ? kGuideOut : kGuideHidden;
  if (p < rows * kGuideHidden) {
    j = p / kGuideHidden; k = p - j * kGuideHidden;
    s.fwd = fwd_base + j * fwd_stride + k;
    s.tr = tr_base + k * kRowStride + j;
  } else {
    j = p - rows * kGuideHidden; k = kGuideHidden;
    s.bias = bias_base + j;
  }
  s.tile = base_t + (j >> 4) * ntiles + (k >> 4);
  s.elem = (j & 15) * 16 + (k & 15);
  return s;
}

BJ_D void pack_param(uint8_t* pack, const ParamSlot& s, float v) {
  uint16_t* w = (uint16_t*)pack;
  float* b = (float*)(pack + (size_t)kPackAll * 2);
  if (s.bias >= 0) {
    b[s.bias] = v;
    return;
  }
  const uint16_t h = f2bf(v);
  w[s.fwd] = h;
  if (s.tr >= 0) w[s.tr] = h;
}

__global__ void __launch_bounds__(256) guide_pack_kernel(const float* __restrict__ params, uint8_t* __restrict__ pack) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kGuideParams) pack_param(pack, param_slot(p), params[p]);
}

// ---- layer-1 A fragments ---------------------------------------------------------
// Chunk q (0..27) of a sample's features: input d = q / 3, frequencies
// 4 (q % 3) .. + 3, (sin, cos) pairs; chunk 27 is padding (zeros).
BJ_D bf16x8 encode_chunk(const float (&x)[kGuideInputs], int q) {
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f}, c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (q < 27) {
    const int d = q / 3, fb = q - 3 * d;
    float xd = x[0];
#pragma unroll
    for (int a = 1; a < kGuideInputs; ++a) xd = d == a ? x[a] : xd;
    guide_encode_group(xd, fb, s, c);
  }
  // round to nearest even (the hardware conversion equals f2bf for finite values)
  return bf16x8{(__bf16)s[0], (__bf16)c[0], (__bf16)s[1], (__bf16)c[1],
                (__bf16)s[2], (__bf16)c[2], (__bf16)s[3], (__bf16)c[3]};
}

// Bias + ReLU of a 16 x 64 layer result into a row-major BF16 tile.
BJ_D void store_hidden(const f32x4 (&acc)[4], const float* bias, uint16_t* sh, uint32_t lane) {
  const uint32_t col = lane & 15u, g = lane >> 4;
#pragma unroll
  for (int jt = 0; jt < 4; ++jt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[jt][r] + bias[16 * jt + col];
      sh[(4 * g + r) * kRowStride + 16 * jt + col] = f2bf(v > 0.0f ? v : 0.0f);
    }
}

// One 64-wide layer: acc = H (16 x 64, row-major BF16 tile) x W^T rows.
template <int kTiles>
BJ_D void layer64(const uint16_t* sh, const uint16_t* sw, f32x4 (&acc)[kTiles], uint32_t lane) {
  const uint32_t i = lane & 15u, g = lane >> 4;
#pragma unroll
  for (int jt = 0; jt < kTiles; ++jt) acc[jt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const bf16x8 a = frag16(sh + i * kRowStride + 32 * s + 8 * g);
#pragma unroll
    for (int jt = 0; jt < kTiles; ++jt) acc[jt] = mfma(a, frag16(sw + (16 * jt + i) * kRowStride + 32 * s + 8 * g), acc[jt]);
  }
}

// Layer 1 from the encoded inputs: lane (i, g) holds sample i's inputs.
BJ_D void layer1(const float (&x)[kGuideInputs], const uint16_t* sw, f32x4 (&acc)[4], uint32_t lane,
                 uint16_t* x0t = nullptr, uint64_t ldb = 0, uint64_t col = 0) {
  const uint32_t i = lane & 15u, g = lane >> 4;
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) acc[jt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    const bf16x8 a = encode_chunk(x, 4 * s + (int)g);
    if (x0t) {
      // training: the transposed features, with the constant input 1 at row 216
      const uint4 w = __builtin_bit_cast(uint4, a);
      const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
      const uint32_t k0 = 32u * s + 8u * g;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        uint16_t h = (uint16_t)(ww[e >> 1] >> (16 * (e & 1)));
        if (k0 + e == (uint32_t)kGuideFeatures) h = 0x3F80u;
        x0t[(uint64_t)(k0 + e) * ldb + col + i] = h;
      }
    }
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) acc[jt] = mfma(a, frag16(sw + (16 * jt + i) * kW1Stride + 32 * s + 8 * g), acc[jt]);
  }
}

// Block-wide copy of the first `elems` BF16 weights and the biases into LDS.
BJ_D void load_pack(const uint8_t* __restrict__ pack, uint16_t* sw, int elems, float* sb) {
  const uint4* src = (const uint4*)pack;
  uint4* dst = (uint4*)sw;
  for (int i = threadIdx.x; i < elems / 8; i += blockDim.x) dst[i] = src[i];
  const float* bsrc = (const float*)(pack + (size_t)kPackAll * 2);
  for (int i = threadIdx.x; i < kBiasFloats; i += blockDim.x) sb[i] = bsrc[i];
  __syncthreads();
}

// ---- forward ---------------------------------------------------------------------
// Inputs: explicit f32[9 n], or render vertices (their position, normal and
// ray direction planes) normalized against the scene box.
struct InputRows {
  const float* x;
  BJ_D void load(uint64_t r, float (&v)[kGuideInputs]) const {
#pragma unroll
    for (int a = 0; a < kGuideInputs; ++a) v[a] = x[r * kGuideInputs + a];
  }
};

struct VertexRows {
  const float* f;
  const float* bounds;  // scene box lo.xyz, hi.xyz
  BJ_D void load(uint64_t r, float (&v)[kGuideInputs]) const {
    const v3 lo = mk(bounds[0], bounds[1], bounds[2]);
    const v3 hi = mk(bounds[3], bounds[4], bounds[5]);
    const v3 ext = sub(hi, lo);
    const v3 inv = mk(ext.x > 0.0f ? 1.0f / ext.x : 0.0f, ext.y > 0.0f ? 1.0f / ext.y : 0.0f,
                      ext.z > 0.0f ? 1.0f / ext.z : 0.0f);
    const v3 p = mk(f[vplane(r, 0)], f[vplane(r, 1)], f[vplane(r, 2)]);
    const v3 n = mk(f[vplane(r, 3)], f[vplane(r, 4)], f[vplane(r, 5)]);
    const v3 d = mk(f[vplane(r, 6)], f[vplane(r, 7)], f[vplane(r, 8)]);
    const GuideInput g = guide_input(p, n, normalize(mk(-d.x, -d.y, -d.z)), lo, inv);
#pragma unroll
    for (int a = 0; a < kGuideInputs; ++a) v[a] = g.v[a];
  }
};

template <typename Rows>
__global__ void __launch_bounds__(256) guide_forward_kernel(const uint8_t* __restrict__ pack, Rows rows,
                                                            const uint32_t* __restrict__ count_dev, uint64_t count,
                                                            float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t s_w[kPackFwd];
  __shared__ float s_b[kBiasFloats];
  __shared__ __attribute__((aligned(16))) uint16_t s_h[4][16 * kRowStride];
  uint64_t n = count;
  if (count_dev) n = *count_dev < count ? *count_dev : count;
  const uint64_t tiles = (n + 15) / 16;
  if ((uint64_t)blockIdx.x * 4 >= tiles) return;  // uniform per block: before the barrier
  load_pack(pack, s_w, kPackFwd, s_b);
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t i = lane & 15u, g = lane >> 4;
  uint16_t* sh = s_h[wave];
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + wave; tile < tiles; tile += (uint64_t)gridDim.x * 4) {
    const uint64_t r = tile * 16 + i;
    float x[kGuideInputs];
    if (r < n) rows.load(r, x);
    else
#pragma unroll
      for (int a = 0; a < kGuideInputs; ++a) x[a] = 0.0f;
    f32x4 acc[4];
    layer1(x, s_w + kPackW1, acc, lane);
    store_hidden(acc, s_b, sh, lane);
    __builtin_amdgcn_wave_barrier();
    layer64<4>(sh, s_w + kPackW2, acc, lane);
    __builtin_amdgcn_wave_barrier();
    store_hidden(acc, s_b + 64, sh, lane);
    __builtin_amdgcn_wave_barrier();
    layer64<4>(sh, s_w + kPackW3, acc, lane);
    __builtin_amdgcn_wave_barrier();
    store_hidden(acc, s_b + 128, sh, lane);
    __builtin_amdgcn_wave_barrier();
    f32x4 o[3];
    layer64<3>(sh, s_w + kPackW4, o, lane);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int jt = 0; jt < 3; ++jt) {
      const uint32_t col = 16u * jt + i;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const uint64_t row = tile * 16 + 4 * g + rr;
        if (col < (uint32_t)kGuideOut && row < n) out[row * kGuideOut + col] = o[jt][rr] + s_b[192 + col];
      }
    }
  }
}

// ---- loss gradient of one sample (outputs -> d loss / d outputs) ------------------
// Loss -w ln q(dir), q = sum_j pi_j vMF_j(dir) in log space; returns the
// weighted loss.
BJ_D float output_gradient(const float* o, v3 dir, float w, float* d) {
  float lp[kGuideLobes], lc[kGuideLobes], kap[kGuideLobes], mdw[kGuideLobes], len[kGuideLobes], e2v[kGuideLobes];
  v3 mu[kGuideLobes];
  bool inside[kGuideLobes];
  float amax = o[4];
#pragma unroll
  for (int j = 1; j < kGuideLobes; ++j) amax = fmax_sel(amax, o[5 * j + 4]);
  float se = 0.0f;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) se += exp_f(o[5 * j + 4] - amax);
  const float lse_a = amax + ln_f(se);
  float cmax = -kFltMax;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) {
    const float* oj = o + 5 * j;
    const float l2 = oj[0] * oj[0] + oj[1] * oj[1] + oj[2] * oj[2];
    len[j] = sqrtf(l2);
    mu[j] = l2 > 1.0e-24f ? mk(oj[0] / len[j], oj[1] / len[j], oj[2] / len[j]) : mk(0.0f, 0.0f, 1.0f);
    inside[j] = oj[3] > kGuideLogKappaMin && oj[3] < kGuideLogKappaMax;
    const float s = fmin_sel(fmax_sel(oj[3], kGuideLogKappaMin), kGuideLogKappaMax);
    kap[j] = exp_f(s);
    e2v[j] = exp_f(-2.0f * kap[j]);
    mdw[j] = dot(mu[j], dir);
    lp[j] = oj[4] - lse_a;
    lc[j] = lp[j] + (s - kLn2Pi - ln_f(1.0f - e2v[j])) + kap[j] * (mdw[j] - 1.0f);
    cmax = fmax_sel(cmax, lc[j]);
  }
  float sc = 0.0f;
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) sc += exp_f(lc[j] - cmax);
  const float lq = cmax + ln_f(sc);
#pragma unroll
  for (int j = 0; j < kGuideLobes; ++j) {
    const float r = exp_f(lc[j] - lq);
    const float pi = exp_f(lp[j]);
    d[5 * j + 4] = w * (pi - r);
    const float dlk = 1.0f / kap[j] - 2.0f * e2v[j] / (1.0f - e2v[j]) + (mdw[j] - 1.0f);
    d[5 * j + 3] = inside[j] ? -w * r * dlk * kap[j] : 0.0f;
    const float k = len[j] > 1.0e-12f ? -w * r * kap[j] / len[j] : 0.0f;
    d[5 * j + 0] = k * (dir.x - mdw[j] * mu[j].x);
    d[5 * j + 1] = k * (dir.y - mdw[j] * mu[j].y);
    d[5 * j + 2] = k * (dir.z - mdw[j] * mu[j].z);
  }
  return -w * lq;
}

// Masked backward step: acc (16 x 64 D fragments) times the ReLU mask of
// the forward activations `sh`, to BF16 in `sd` (row-major) .
BJ_D void store_delta(const f32x4 (&acc)[4], const uint16_t* sh, uint16_t* sd, uint32_t lane) {
  const uint32_t col = lane & 15u, g = lane >> 4;
#pragma unroll
  for (int jt = 0; jt < 4; ++jt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int at = (4 * g + r) * kRowStride + 16 * jt + col;
      sd[at] = sh[at] != 0 ? f2bf(acc[jt][r]) : (uint16_t)0;
    }
}

// Writes rows [0, rows) of a row-major 16 x 72 tile transposed: lane l
// writes row l's 16 samples (32 contiguous bytes) at dst[l * ldb + col].
BJ_D void store_transposed(const uint16_t* st, int rows, uint16_t* dst, uint64_t ldb, uint64_t col, uint32_t lane) {
  if ((int)lane < rows) {
    uint32_t w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
      w[e] = (uint32_t)st[(2 * e) * kRowStride + lane] | ((uint32_t)st[(2 * e + 1) * kRowStride + lane] << 16);
    uint4* p = (uint4*)(dst + (uint64_t)lane * ldb + col);
    p[0] = uint4{w[0], w[1], w[2], w[3]};
    p[1] = uint4{w[4], w[5], w[6], w[7]};
  }
}

// Constant rows 64..79 of a hidden-activation buffer: 1 at row 64, 0 below.
BJ_D void store_bias_rows(uint16_t* dst, uint64_t ldb, uint64_t col, uint32_t lane) {
  if (lane < 16) {
    const uint32_t v = lane == 0 ? 0x3F803F80u : 0u;
    uint4* p = (uint4*)(dst + (uint64_t)(kGuideHidden + lane) * ldb + col);
    p[0] = uint4{v, v, v, v};
    p[1] = uint4{v, v, v, v};
  }
}

struct TrainArgs {
  const float* samples;  // f32[13 * src_count]
  // batch row r is sample (src_offset + r * src_stride) % src_count
  uint64_t src_count, src_offset, src_stride;
  uint64_t n;            // samples
  uint64_t n_pad;        // n rounded up to a whole chunk (padded rows carry weight 0)
  uint64_t ldb;          // row stride of the transposed buffers
  uint16_t *x0t, *h1t, *h2t, *h3t, *d1t, *d2t, *d3t, *dot;
  float* tile_part;      // per tile: sum w, weighted loss
};

// Forward + backward of 16-sample tiles: activations and deltas of every
// layer written transposed for the weight gradient.
__global__ void __launch_bounds__(256) guide_train_kernel(const uint8_t* __restrict__ pack, TrainArgs ta) {
  __shared__ __attribute__((aligned(16))) uint16_t s_w[kPackAll];
  __shared__ float s_b[kBiasFloats];
  __shared__ __attribute__((aligned(16))) uint16_t s_h[4][3][16 * kRowStride];
  __shared__ __attribute__((aligned(16))) uint16_t s_d[4][16 * kRowStride];
  __shared__ float s_o[4][16 * kGuideOutPad];
  const uint64_t tiles = ta.n_pad / 16;
  if ((uint64_t)blockIdx.x * 4 >= tiles) return;
  load_pack(pack, s_w, kPackAll, s_b);
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t i = lane & 15u, g = lane >> 4;
  uint16_t* h1 = s_h[wave][0];
  uint16_t* h2 = s_h[wave][1];
  uint16_t* h3 = s_h[wave][2];
  uint16_t* sd = s_d[wave];
  float* so = s_o[wave];
  for (uint64_t tile = (uint64_t)blockIdx.x * 4 + wave; tile < tiles; tile += (uint64_t)gridDim.x * 4) {
    const uint64_t col = tile * 16;
    const uint64_t r = col + i;
    float x[kGuideInputs];
    const float* sp = ta.samples + ((ta.src_offset + r * ta.src_stride) % ta.src_count) * kGuideSampleFloats;
#pragma unroll
    for (int a = 0; a < kGuideInputs; ++a) x[a] = r < ta.n ? sp[a] : 0.0f;
    f32x4 acc[4];
    layer1(x, s_w + kPackW1, acc, lane, ta.x0t, ta.ldb, col);
    store_hidden(acc, s_b, h1, lane);
    __builtin_amdgcn_wave_barrier();
    layer64<4>(h1, s_w + kPackW2, acc, lane);
    store_hidden(acc, s_b + 64, h2, lane);
    __builtin_amdgcn_wave_barrier();
    layer64<4>(h2, s_w + kPackW3, acc, lane);
    store_hidden(acc, s_b + 128, h3, lane);
    __builtin_amdgcn_wave_barrier();
    f32x4 o[3];
    layer64<3>(h3, s_w + kPackW4, o, lane);
#pragma unroll
    for (int jt = 0; jt < 3; ++jt)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) so[(4 * g + rr) * kGuideOutPad + 16 * jt + i] = o[jt][rr] + s_b[192 + 16 * jt + i];
    store_transposed(h1, kGuideHidden, ta.h1t, ta.ldb, col, lane);
    store_transposed(h2, kGuideHidden, ta.h2t, ta.ldb, col, lane);
    store_transposed(h3, kGuideHidden, ta.h3t, ta.ldb, col, lane);
    store_bias_rows(ta.h1t, ta.ldb, col, lane);
    store_bias_rows(ta.h2t, ta.ldb, col, lane);
    store_bias_rows(ta.h3t, ta.ldb, col, lane);
    __builtin_amdgcn_wave_barrier();
    // loss gradient of the outputs: lane i < 16 takes sample i
    float wsum = 0.0f, lsum = 0.0f;
    if (lane < 16) {
      float d[kGuideOut];
      float w = 0.0f;
      v3 dir = mk(0.0f, 0.0f, 1.0f);
      if (r < ta.n) {
        dir = mk(sp[9], sp[10], sp[11]);
        w = sp[12];
      }
      const float loss = output_gradient(so + lane * kGuideOutPad, dir, w, d);
      wsum = w;
      lsum = w > 0.0f ? loss : 0.0f;
#pragma unroll
      for (int c = 0; c < kGuideOut; ++c) sd[lane * kRowStride + c] = f2bf(d[c]);
#pragma unroll
      for (int c = kGuideOut; c < kGuideHidden; ++c) sd[lane * kRowStride + c] = 0;
    }
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) {
      wsum += __shfl_xor(wsum, off, 64);
      lsum += __shfl_xor(lsum, off, 64);
    }
    if (lane == 0) {
      ta.tile_part[2 * tile] = wsum;
      ta.tile_part[2 * tile + 1] = lsum;
    }
    __builtin_amdgcn_wave_barrier();
    store_transposed(sd, kDoRows, ta.dot, ta.ldb, col, lane);
    // backward: d3 = dO W4 (B rows from W4^T), masked by h3; then d2, d1
    layer64<4>(sd, s_w + kPackW4t, acc, lane);
    __builtin_amdgcn_wave_barrier();
    store_delta(acc, h3, sd, lane);
    __builtin_amdgcn_wave_barrier();
    store_transposed(sd, kDRows, ta.d3t, ta.ldb, col, lane);
    layer64<4>(sd, s_w + kPackW3t, acc, lane);
    __builtin_amdgcn_wave_barrier();
    store_delta(acc, h2, sd, lane);
    __builtin_amdgcn_wave_barrier();
    store_transposed(sd, kDRows, ta.d2t, ta.ldb, col, lane);
    layer64<4>(sd, s_w + kPackW2t, acc, lane);
    __builtin_amdgcn_wave_barrier();
    store_delta(acc, h1, sd, lane);
    __builtin_amdgcn_wave_barrier();
    store_transposed(sd, kDRows, ta.d1t, ta.ldb, col, lane);
    __builtin_amdgcn_wave_barrier();
  }
}

// ---- weight gradient: dW[out][in] = sum over the batch of D[out] X[in] ----------
struct WgradArgs {
  const uint16_t *x0t, *h1t, *h2t, *h3t, *d1t, *d2t, *d3t, *dot;
  uint64_t ldb;
  float* part;  // [chunk][kTiles][256]
};

__global__ void __launch_bounds__(64) guide_wgrad_kernel(WgradArgs wa) {
  const uint32_t t = blockIdx.x, chunk = blockIdx.y;
  const uint32_t lane = lane_id();
  const uint32_t i = lane & 15u, g = lane >> 4;
  const uint16_t *dsrc, *xsrc;
  uint32_t mt, nt;
  if (t < (uint32_t)kT2) { mt = t / 14; nt = t % 14; dsrc = wa.d1t; xsrc = wa.x0t; }
  else if (t < (uint32_t)kT3) { mt = (t - kT2) / 5; nt = (t - kT2) % 5; dsrc = wa.d2t; xsrc = wa.h1t; }
  else if (t < (uint32_t)kT4) { mt = (t - kT3) / 5; nt = (t - kT3) % 5; dsrc = wa.d3t; xsrc = wa.h2t; }
  else { mt = (t - kT4) / 5; nt = (t - kT4) % 5; dsrc = wa.dot; xsrc = wa.h3t; }
  const uint64_t b0 = (uint64_t)chunk * kChunk;
  const uint16_t* a_row = dsrc + (uint64_t)(16 * mt + i) * wa.ldb + b0 + 8 * g;
  const uint16_t* b_row = xsrc + (uint64_t)(16 * nt + i) * wa.ldb + b0 + 8 * g;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
  for (int s = 0; s < kChunk / 32; ++s) acc = mfma(frag16(a_row + 32 * s), frag16(b_row + 32 * s), acc);
  float* out = wa.part + ((uint64_t)chunk * kTiles + t) * 256;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(4 * g + r) * 16 + i] = acc[r];
}

// ---- reduce + Adam ----------------------------------------------------------------
struct AdamArgs {
  float* params;
  float* m;
  float* v;
  uint8_t* pack;
  float* ema;            // moving average of the weights and its weight images
  uint8_t* infer_pack;
  float ema_decay;
  const float* part;
  uint32_t chunks;
  const float* tile_part;
  uint32_t tiles;
  float lr, beta1, beta2, eps, bias1, bias2;  // bias1/2 = 1 - beta^t
  float* grad_out;  // gradient mode: write the gradient, no update
};

// Sum over the batch's tiles of per-tile values (stride 2), fixed order.
BJ_D float block_sum(const float* v, uint32_t count, float* s_tmp) {
  float acc = 0.0f;
  for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) acc += v[2 * k];
  s_tmp[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s_tmp[threadIdx.x] += s_tmp[threadIdx.x + w];
    __syncthreads();
  }
  const float total = s_tmp[0];
  __syncthreads();
  return total;
}

// Decay of the weights' moving average that renders sample from: the last
// steps of a frame are noisy (lr 1e-2 on one frame's samples), and the
// average over about 100 steps is the steadier guide.
constexpr float kEmaDecay = 0.99f;

__global__ void __launch_bounds__(256) guide_adam_kernel(AdamArgs aa) {
  __shared__ float s_tmp[256];
  const float wsum = block_sum(aa.tile_part, aa.tiles, s_tmp);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kGuideParams) return;
  const ParamSlot sl = param_slot(p);
  float gsum = 0.0f;
  for (uint32_t c = 0; c < aa.chunks; ++c) gsum += aa.part[((uint64_t)c * kTiles + sl.tile) * 256 + sl.elem];
  const float grad = wsum > 0.0f ? gsum / wsum : 0.0f;
  if (aa.grad_out) {
    aa.grad_out[p] = grad;
    return;
  }
  const float m = aa.beta1 * aa.m[p] + (1.0f - aa.beta1) * grad;
  const float v = aa.beta2 * aa.v[p] + (1.0f - aa.beta2) * grad * grad;
  aa.m[p] = m;
  aa.v[p] = v;
  const float mh = m / aa.bias1, vh = v / aa.bias2;
  const float w = aa.params[p] - aa.lr * mh / (sqrtf(vh) + aa.eps);
  aa.params[p] = w;
  pack_param(aa.pack, sl, w);
  const float e = aa.ema_decay * aa.ema[p] + (1.0f - aa.ema_decay) * w;
  aa.ema[p] = e;
  pack_param(aa.infer_pack, sl, e);
}

// ---- mixture sampling (tests and tools) ---------------------------------------------
__global__ void __launch_bounds__(256) guide_mixture_kernel(const float* __restrict__ outputs, uint64_t n,
                                                            const float* __restrict__ u, float* __restrict__ dirs,
                                                            float* __restrict__ pdfs) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  GuideMixture m;
  guide_decode(outputs + r * kGuideOut, m);
  const v3 d = guide_sample(m, outputs + r * kGuideOut, u[3 * r], u[3 * r + 1], u[3 * r + 2]);
  dirs[3 * r] = d.x;
  dirs[3 * r + 1] = d.y;
  dirs[3 * r + 2] = d.z;
  pdfs[r] = guide_pdf(m, d);
}

// ---- host -------------------------------------------------------------------------
struct Pcg32 {
  uint64_t state = 0, inc = 0;
  Pcg32(uint64_t initstate, uint64_t initseq) {
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t next() {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// He-uniform weights (bound sqrt(6 / fan_in)), output layer scaled by 0.1,
// zero biases, drawn in parameter order from PCG32(seed, 54).
void init_params(uint64_t seed, std::vector<float>& p) {
  p.assign(kGuideParams, 0.0f);
  Pcg32 rng(seed, 54);
  auto fill = [&](int off, int rows, int fan_in, float gain) {
    const float bound = gain * sqrtf(6.0f / (float)fan_in);
    for (int k = 0; k < rows * fan_in; ++k) p[off + k] = (2.0f * rng.uniform() - 1.0f) * bound;
  };
  fill(kGuideOffW1, kGuideHidden, kGuideFeatures, 1.0f);
  fill(kGuideOffW2, kGuideHidden, kGuideHidden, 1.0f);
  fill(kGuideOffW3, kGuideHidden, kGuideHidden, 1.0f);
  fill(kGuideOffW4, kGuideOut, kGuideHidden, 0.1f);
}

int32_t upload_params(bajo_ctx* ctx, bajo_guide* g, const float* host) {
  BJ_HIP(hipMemcpyAsync(g->params, host, kGuideParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  BJ_HIP(hipMemsetAsync(g->adam_m, 0, kGuideParams * sizeof(float), ctx->stream));
  BJ_HIP(hipMemsetAsync(g->adam_v, 0, kGuideParams * sizeof(float), ctx->stream));
  BJ_HIP(hipMemcpyAsync(g->ema, g->params, kGuideParams * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  guide_pack_kernel<<<div_up(kGuideParams, 256), 256, 0, ctx->stream>>>(g->params, g->pack);
  guide_pack_kernel<<<div_up(kGuideParams, 256), 256, 0, ctx->stream>>>(g->params, g->infer_pack);
  BJ_HIP(hipGetLastError());
  g->step = 0;
  return BAJO_OK;
}

void free_guide(bajo_guide* g) {
  if (!g) return;
  (void)hipFree(g->params); (void)hipFree(g->adam_m); (void)hipFree(g->adam_v); (void)hipFree(g->pack);
  (void)hipFree(g->ema); (void)hipFree(g->infer_pack);
  (void)hipFree(g->x0t); (void)hipFree(g->h1t); (void)hipFree(g->h2t); (void)hipFree(g->h3t);
  (void)hipFree(g->d1t); (void)hipFree(g->d2t); (void)hipFree(g->d3t); (void)hipFree(g->dot);
  (void)hipFree(g->tile_part); (void)hipFree(g->grad_part);
  (void)hipFree(g->records); (void)hipFree(g->contrib); (void)hipFree(g->samples); (void)hipFree(g->sample_count);
  (void)hipFree(g->block_counts);
  (void)hipFree(g->vertices); (void)hipFree(g->outputs);
  delete g;
}

template <typename K>
int32_t blocks_per_cu(K kernel, int* out) {
  int dev = 0, cus = 0, per_cu = 0;
  BJ_HIP(hipGetDevice(&dev));
  BJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  BJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
  *out = cus * (per_cu < 1 ? 1 : per_cu);
  return BAJO_OK;
}

// Forward + backward + weight-gradient partials of one batch (n <= batch_cap).
int32_t run_backward(hipStream_t stream, bajo_guide* g, const float* samples, uint64_t src_count, uint64_t src_offset,
                     uint64_t src_stride, uint64_t n, uint64_t* ldb_out, uint32_t* chunks_out) {
  const uint64_t ldb = (n + kChunk - 1) / kChunk * kChunk;  // padded batch
  const uint32_t chunks = (uint32_t)(ldb / kChunk);
  TrainArgs ta{samples, src_count, src_offset, src_stride, n, ldb, g->batch_cap, g->x0t, g->h1t, g->h2t, g->h3t,
               g->d1t, g->d2t, g->d3t, g->dot, g->tile_part};
  const uint64_t tiles = ldb / 16;
  int grid = 0;
  int32_t rc = blocks_per_cu(guide_train_kernel, &grid);
  if (rc) return rc;
  const uint64_t need = (tiles + 3) / 4;
  if ((uint64_t)grid > need) grid = (int)need;
  guide_train_kernel<<<grid, 256, 0, stream>>>(g->pack, ta);
  BJ_HIP(hipGetLastError());
  WgradArgs wa{g->x0t, g->h1t, g->h2t, g->h3t, g->d1t, g->d2t, g->d3t, g->dot, g->batch_cap, g->grad_part};
  guide_wgrad_kernel<<<dim3(kTiles, chunks), 64, 0, stream>>>(wa);
  BJ_HIP(hipGetLastError());
  *ldb_out = ldb;
  *chunks_out = chunks;
  return BAJO_OK;
}

int32_t run_adam(hipStream_t stream, bajo_guide* g, uint64_t ldb, uint32_t chunks, float* grad_out) {
  AdamArgs aa;
  aa.params = g->params;
  aa.m = g->adam_m;
  aa.v = g->adam_v;
  aa.pack = g->pack;
  aa.ema = g->ema;
  aa.infer_pack = g->infer_pack;
  aa.part = g->grad_part;
  aa.chunks = chunks;
  aa.tile_part = g->tile_part;
  aa.tiles = (uint32_t)(ldb / 16);
  aa.lr = g->opt.learning_rate;
  aa.beta1 = g->opt.beta1;
  aa.beta2 = g->opt.beta2;
  aa.eps = g->opt.epsilon;
  aa.grad_out = grad_out;
  aa.ema_decay = 0.0f;
  if (!grad_out) {
    g->step += 1;
    aa.bias1 = 1.0f - powf(g->opt.beta1, (float)g->step);
    aa.bias2 = 1.0f - powf(g->opt.beta2, (float)g->step);
    // the average follows the first steps closely, then settles at kEmaDecay
    const float warm = (1.0f + (float)g->step) / (10.0f + (float)g->step);
    aa.ema_decay = warm < kEmaDecay ? warm : kEmaDecay;
  } else {
    aa.bias1 = aa.bias2 = 1.0f;
  }
  guide_adam_kernel<<<div_up(kGuideParams, 256), 256, 0, stream>>>(aa);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

}  // namespace

int32_t guide_train_samples(hipStream_t stream, bajo_guide* g, const float* samples, uint64_t n) {
  for (uint64_t b = 0; b < n; b += g->batch_cap) {
    const uint64_t m = n - b < g->batch_cap ? n - b : g->batch_cap;
    uint64_t ldb = 0;
    uint32_t chunks = 0;
    int32_t rc = run_backward(stream, g, samples, n, b, 1, m, &ldb, &chunks);
    if (rc) return rc;
    rc = run_adam(stream, g, ldb, chunks, nullptr);
    if (rc) return rc;
  }
  return BAJO_OK;
}

namespace {

// The training sample of path p at bounce k, if it has one: the vertex's
// inputs, the direction taken, and as weight the radiance added at bounces
// after it over the throughput the path carried out of it (per channel), as
// luminance over the density its direction was drawn with.
BJ_D bool collect_sample(const float* __restrict__ records, const float* __restrict__ contrib, uint64_t p,
                         uint32_t k, uint32_t depth, v3 lo, v3 inv, float* out) {
  const float* r = records + (p * depth + k) * kRecordPlanes;
  const float pdf = r[12];
  if (!(pdf > 0.0f)) return false;
  v3 L = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t j = k + 1; j < depth; ++j) {
    const float* c = contrib + 3 * (p * depth + j);
    L = add(L, mk(c[0], c[1], c[2]));
  }
  const v3 T = mk(r[13], r[14], r[15]);
  L = mk(T.x > 0.0f ? L.x / T.x : 0.0f, T.y > 0.0f ? L.y / T.y : 0.0f, T.z > 0.0f ? L.z / T.z : 0.0f);
  const float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
  if (!(lum > 0.0f)) return false;
  const GuideInput gi = guide_input(mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), mk(r[6], r[7], r[8]), lo, inv);
#pragma unroll
  for (int a = 0; a < kGuideInputs; ++a) out[a] = gi.v[a];
  out[9] = r[9];
  out[10] = r[10];
  out[11] = r[11];
  out[12] = lum / pdf;
  return true;
}

BJ_D void bounds_frame(const float* __restrict__ bounds, v3& lo, v3& inv) {
  lo = mk(bounds[0], bounds[1], bounds[2]);
  const v3 ext = sub(mk(bounds[3], bounds[4], bounds[5]), lo);
  inv = mk(ext.x > 0.0f ? 1.0f / ext.x : 0.0f, ext.y > 0.0f ? 1.0f / ext.y : 0.0f, ext.z > 0.0f ? 1.0f / ext.z : 0.0f);
}

// The frame's samples are stored in a fixed order -- block by block (256
// paths each), bounce by bounce within a block, then path order -- so the
// training batches, and the trained network, do not depend on the order in
// which waves run.  Pass 1 counts each block's samples.
__global__ void __launch_bounds__(256) guide_count_kernel(const float* __restrict__ records,
                                                          const float* __restrict__ contrib, uint64_t paths,
                                                          uint32_t depth, const float* __restrict__ bounds,
                                                          uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t s_wave[4];
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  v3 lo, inv;
  bounds_frame(bounds, lo, inv);
  uint32_t mine = 0;
  float out[kGuideSampleFloats];
  if (p < paths)
    for (uint32_t k = 0; k < depth; ++k) mine += collect_sample(records, contrib, p, k, depth, lo, inv, out) ? 1u : 0u;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mine += (uint32_t)__shfl_xor((int)mine, off, 64);
  if (lane_id() == 0) s_wave[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
}

// Pass 2 (one block): exclusive prefix of the block counts, in place; the
// total becomes the frame's sample count.
__global__ void __launch_bounds__(1024) guide_offsets_kernel(uint32_t* __restrict__ block_counts, uint32_t blocks,
                                                             uint32_t* __restrict__ count) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < blocks; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < blocks ? block_counts[i] : 0u;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
      if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = s_carry;
    for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
    if (i < blocks) block_counts[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += s_wave[w];
      s_carry += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) count[0] = s_carry;
}

// Pass 3: each block writes its samples from its offset, bounce by bounce,
// in path order.
__global__ void __launch_bounds__(256) guide_collect_kernel(const float* __restrict__ records,
                                                            const float* __restrict__ contrib, uint64_t paths,
                                                            uint32_t depth, const float* __restrict__ bounds,
                                                            const uint32_t* __restrict__ block_offsets,
                                                            float* __restrict__ samples, uint64_t cap) {
  __shared__ uint32_t s_wave[4];
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  v3 lo, inv;
  bounds_frame(bounds, lo, inv);
  uint64_t next = block_offsets[blockIdx.x];
  for (uint32_t k = 0; k < depth; ++k) {
    float out[kGuideSampleFloats];
    const bool want = p < paths && collect_sample(records, contrib, p, k, depth, lo, inv, out);
    const uint64_t m = ballot(want);
    if (lane == 0) s_wave[wave] = (uint32_t)__builtin_popcountll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < 4; ++w) {
      if (w < wave) before += s_wave[w];
      total += s_wave[w];
    }
    const uint64_t slot = next + before + mask_rank(m);
    if (want && slot < cap)
#pragma unroll
      for (int a = 0; a < kGuideSampleFloats; ++a) samples[slot * kGuideSampleFloats + a] = out[a];
    next += total;
    __syncthreads();
  }
}

uint64_t gcd_u64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

int32_t guide_infer_vertices(hipStream_t stream, const bajo_guide* g, const float* vertex_fields,
                             const uint32_t* count_dev, uint64_t cap, const float* scene_bounds, float* outputs) {
  int grid = 0;
  int32_t rc = blocks_per_cu(guide_forward_kernel<VertexRows>, &grid);
  if (rc) return rc;
  const uint64_t need = (cap + 63) / 64;
  if ((uint64_t)grid > need) grid = (int)need;
  guide_forward_kernel<VertexRows><<<grid, 256, 0, stream>>>(g->infer_pack, VertexRows{vertex_fields, scene_bounds},
                                                             count_dev, cap, outputs);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

int32_t guide_prepare_frame(hipStream_t stream, bajo_guide* g, uint64_t paths, uint32_t depth, bool train) {
  if (paths > g->vertex_cap) {
    BJ_HIP(hipStreamSynchronize(stream));
    (void)hipFree(g->vertices);
    (void)hipFree(g->outputs);
    g->vertices = g->outputs = nullptr;
    g->vertex_cap = 0;
    const uint64_t tiles = (paths + 63) / 64;
    BJ_HIP(hipMalloc(&g->vertices, tiles * 64 * kVertexPlanes * sizeof(float)));
    BJ_HIP(hipMalloc(&g->outputs, paths * kGuideOut * sizeof(float)));
    g->vertex_cap = paths;
  }
  if (!train) return BAJO_OK;
  const uint64_t slots = paths * depth;
  if (slots > g->record_cap) {
    BJ_HIP(hipStreamSynchronize(stream));
    (void)hipFree(g->records);
    (void)hipFree(g->contrib);
    (void)hipFree(g->samples);
    g->records = g->contrib = g->samples = nullptr;
    g->record_cap = g->sample_cap = 0;
    BJ_HIP(hipMalloc(&g->records, slots * kRecordPlanes * sizeof(float)));
    BJ_HIP(hipMalloc(&g->contrib, slots * 3 * sizeof(float)));
    BJ_HIP(hipMalloc(&g->samples, slots * kGuideSampleFloats * sizeof(float)));
    g->record_cap = g->sample_cap = slots;
  }
  const uint64_t blocks = (paths + 255) / 256;
  if (blocks > g->block_cap) {
    BJ_HIP(hipStreamSynchronize(stream));
    (void)hipFree(g->block_counts);
    g->block_counts = nullptr;
    g->block_cap = 0;
    BJ_HIP(hipMalloc(&g->block_counts, blocks * sizeof(uint32_t)));
    g->block_cap = blocks;
  }
  BJ_HIP(hipMemsetAsync(g->records, 0, slots * kRecordPlanes * sizeof(float), stream));
  BJ_HIP(hipMemsetAsync(g->contrib, 0, slots * 3 * sizeof(float), stream));
  BJ_HIP(hipMemsetAsync(g->sample_count, 0, sizeof(uint32_t), stream));
  return BAJO_OK;
}

int32_t guide_collect(hipStream_t stream, bajo_guide* g, uint64_t paths, uint32_t depth, const float* scene_bounds) {
  const unsigned blocks = div_up(paths, 256);
  BJ_REQUIRE(blocks <= g->block_cap, "guide_collect: %u blocks, room for %llu", blocks,
             (unsigned long long)g->block_cap);
  guide_count_kernel<<<blocks, 256, 0, stream>>>(g->records, g->contrib, paths, depth, scene_bounds, g->block_counts);
  guide_offsets_kernel<<<1, 1024, 0, stream>>>(g->block_counts, blocks, g->sample_count);
  guide_collect_kernel<<<blocks, 256, 0, stream>>>(g->records, g->contrib, paths, depth, scene_bounds,
                                                  g->block_counts, g->samples, g->sample_cap);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

int32_t guide_train_collected(hipStream_t stream, bajo_guide* g, uint64_t* samples_out, uint32_t* steps_out,
                              float* loss_out) {
  uint32_t count = 0;
  BJ_HIP(hipMemcpyAsync(&count, g->sample_count, sizeof(count), hipMemcpyDeviceToHost, stream));
  BJ_HIP(hipStreamSynchronize(stream));
  const uint64_t total = count < g->sample_cap ? count : g->sample_cap;
  *samples_out = total;
  *steps_out = 0;
  *loss_out = 0.0f;
  if (total == 0) return BAJO_OK;
  const uint64_t n = total < g->batch_cap ? total : g->batch_cap;
  // a stride coprime with the set's size spreads each batch over the whole frame
  uint64_t stride = total / n;
  if (stride < 1) stride = 1;
  while (gcd_u64(stride, total) != 1) ++stride;
  uint64_t ldb = 0;
  for (uint32_t k = 0; k < g->opt.steps_per_frame; ++k) {
    uint32_t chunks = 0;
    const uint64_t offset = g->sample_cursor++ % total;
    int32_t rc = run_backward(stream, g, g->samples, total, offset, stride, n, &ldb, &chunks);
    if (rc) return rc;
    rc = run_adam(stream, g, ldb, chunks, nullptr);
    if (rc) return rc;
    *steps_out += 1;
  }
  std::vector<float> part(ldb / 16 * 2);
  BJ_HIP(hipMemcpyAsync(part.data(), g->tile_part, part.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
  BJ_HIP(hipStreamSynchronize(stream));
  double ws = 0.0, ls = 0.0;
  for (size_t i = 0; i < part.size(); i += 2) {
    ws += part[i];
    ls += part[i + 1];
  }
  *loss_out = ws > 0.0 ? (float)(ls / ws) : 0.0f;
  return BAJO_OK;
}

}  // namespace bj

extern "C" {

int32_t bajo_guide_create(bajo_ctx* ctx, const bajo_guide_options* options, bajo_guide** out) {
  BJ_REQUIRE(ctx != nullptr && out != nullptr, "bajo_guide_create: NULL argument");
  *out = nullptr;
  bajo_guide_options o{};
  if (options) o = *options;
  if (!(o.learning_rate > 0.0f)) o.learning_rate = 1.0e-2f;
  if (!(o.beta1 > 0.0f)) o.beta1 = 0.9f;
  if (!(o.beta2 > 0.0f)) o.beta2 = 0.999f;
  if (!(o.epsilon > 0.0f)) o.epsilon = 1.0e-8f;
  if (o.batch_size == 0) o.batch_size = 65536;
  if (o.steps_per_frame == 0) o.steps_per_frame = 4;
  BJ_REQUIRE(o.batch_size % 32 == 0 && o.batch_size <= (1u << 22),
             "bajo_guide_create: batch_size %u must be a multiple of 32 and at most 2^22", o.batch_size);
  BJ_REQUIRE(o.beta1 < 1.0f && o.beta2 < 1.0f, "bajo_guide_create: Adam decays must be below 1");
  bajo_guide* g = new bajo_guide();
  g->device = ctx->device;
  g->opt = o;
  g->batch_cap = (o.batch_size + bj::kChunk - 1) / bj::kChunk * bj::kChunk;
  const uint64_t cap = g->batch_cap;
  auto fail = [&](hipError_t e, const char* what) {
    bj::free_guide(g);
    return bj::hip_fail(e, what, __FILE__, __LINE__);
  };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&g->params, bj::kGuideParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->adam_m, bj::kGuideParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->adam_v, bj::kGuideParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->pack, bj::kPackBytes);
  if (e == hipSuccess) e = hipMalloc(&g->ema, bj::kGuideParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->infer_pack, bj::kPackBytes);
  if (e == hipSuccess) e = hipMalloc(&g->x0t, (size_t)bj::kX0Rows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->h1t, (size_t)bj::kHRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->h2t, (size_t)bj::kHRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->h3t, (size_t)bj::kHRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->d1t, (size_t)bj::kDRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->d2t, (size_t)bj::kDRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->d3t, (size_t)bj::kDRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->dot, (size_t)bj::kDoRows * cap * 2);
  if (e == hipSuccess) e = hipMalloc(&g->tile_part, (size_t)(cap / 16) * 2 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->grad_part, (size_t)(cap / bj::kChunk) * bj::kTiles * 256 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&g->sample_count, 4 * sizeof(uint32_t));
  if (e != hipSuccess) return fail(e, "hipMalloc(guide)");
  // padding entries of the weight image stay zero
  e = hipMemsetAsync(g->pack, 0, bj::kPackBytes, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g->infer_pack, 0, bj::kPackBytes, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g->sample_count, 0, 4 * sizeof(uint32_t), ctx->stream);
  if (e != hipSuccess) return fail(e, "hipMemsetAsync(guide)");
  std::vector<float> p;
  bj::init_params(o.seed, p);
  int32_t rc = bj::upload_params(ctx, g, p.data());
  if (rc == BAJO_OK) {
    e = hipStreamSynchronize(ctx->stream);  // the host vector dies here
    if (e != hipSuccess) return fail(e, "hipStreamSynchronize(guide)");
  }
  if (rc) {
    bj::free_guide(g);
    return rc;
  }
  *out = g;
  return BAJO_OK;
}

int32_t bajo_guide_free(bajo_guide* guide) {
  if (guide) {
    (void)hipDeviceSynchronize();
    bj::free_guide(guide);
  }
  return BAJO_OK;
}

int32_t bajo_guide_get_params(bajo_ctx* ctx, const bajo_guide* guide, float* params) {
  BJ_REQUIRE(ctx && guide && params, "bajo_guide_get_params: NULL argument");
  BJ_HIP(hipMemcpyAsync(params, guide->params, bj::kGuideParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  return BAJO_OK;
}

int32_t bajo_guide_set_params(bajo_ctx* ctx, bajo_guide* guide, const float* params) {
  BJ_REQUIRE(ctx && guide && params, "bajo_guide_set_params: NULL argument");
  int32_t rc = bj::upload_params(ctx, guide, params);
  if (rc) return rc;
  BJ_HIP(hipStreamSynchronize(ctx->stream));
  return BAJO_OK;
}

int32_t bajo_guide_forward(bajo_ctx* ctx, const bajo_guide* guide, const float* inputs, uint64_t n, float* outputs) {
  BJ_REQUIRE(ctx && guide, "bajo_guide_forward: NULL argument");
  BJ_REQUIRE(n == 0 || (inputs && outputs), "bajo_guide_forward: NULL buffer");
  if (n == 0) return BAJO_OK;
  int grid = 0;
  int32_t rc = bj::blocks_per_cu(bj::guide_forward_kernel<bj::InputRows>, &grid);
  if (rc) return rc;
  const uint64_t need = (n + 63) / 64;
  if ((uint64_t)grid > need) grid = (int)need;
  bj::guide_forward_kernel<bj::InputRows><<<grid, 256, 0, ctx->stream>>>(guide->pack, bj::InputRows{inputs}, nullptr,
                                                                         n, outputs);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

int32_t bajo_guide_gradient(bajo_ctx* ctx, bajo_guide* guide, const float* samples, uint64_t n, float* gradient,
                            float* loss) {
  BJ_REQUIRE(ctx && guide && gradient, "bajo_guide_gradient: NULL argument");
  BJ_REQUIRE(n > 0 && samples, "bajo_guide_gradient: empty batch");
  BJ_REQUIRE(n <= guide->batch_cap, "bajo_guide_gradient: %llu samples exceed the batch size %u",
             (unsigned long long)n, guide->batch_cap);
  uint64_t ldb = 0;
  uint32_t chunks = 0;
  int32_t rc = bj::run_backward(ctx->stream, guide, samples, n, 0, 1, n, &ldb, &chunks);
  if (rc) return rc;
  rc = bj::run_adam(ctx->stream, guide, ldb, chunks, gradient);
  if (rc) return rc;
  if (loss) {
    std::vector<float> part(ldb / 16 * 2);
    BJ_HIP(hipMemcpyAsync(part.data(), guide->tile_part, part.size() * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
    BJ_HIP(hipStreamSynchronize(ctx->stream));
    double ws = 0.0, ls = 0.0;
    for (size_t k = 0; k < part.size(); k += 2) {
      ws += part[k];
      ls += part[k + 1];
    }
    *loss = ws > 0.0 ? (float)(ls / ws) : 0.0f;
  }
  return BAJO_OK;
}

int32_t bajo_guide_train(bajo_ctx* ctx, bajo_guide* guide, const float* samples, uint64_t n) {
  BJ_REQUIRE(ctx && guide, "bajo_guide_train: NULL argument");
  BJ_REQUIRE(n == 0 || samples, "bajo_guide_train: samples is NULL");
  return bj::guide_train_samples(ctx->stream, guide, samples, n);
}

int32_t bajo_guide_training_samples(bajo_ctx* ctx, const bajo_guide* guide, float* samples, uint64_t cap,
                                    uint64_t* count) {
  BJ_REQUIRE(ctx && guide && count, "bajo_guide_training_samples: NULL argument");
  BJ_REQUIRE(cap == 0 || samples, "bajo_guide_training_samples: samples is NULL");
  uint32_t c = 0;
  if (guide->sample_count) {
    BJ_HIP(hipMemcpyAsync(&c, guide->sample_count, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    BJ_HIP(hipStreamSynchronize(ctx->stream));
  }
  uint64_t have = c < guide->sample_cap ? c : guide->sample_cap;
  *count = have;
  if (have > cap) have = cap;
  if (have) {
    BJ_HIP(hipMemcpyAsync(samples, guide->samples, have * bj::kGuideSampleFloats * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->stream));
    BJ_HIP(hipStreamSynchronize(ctx->stream));
  }
  return BAJO_OK;
}

int32_t bajo_guide_mixture_sample(bajo_ctx* ctx, const float* outputs, uint64_t n, const float* uniforms, float* dirs,
                                  float* pdfs) {
  BJ_REQUIRE(ctx != nullptr, "bajo_guide_mixture_sample: ctx is NULL");
  BJ_REQUIRE(n == 0 || (outputs && uniforms && dirs && pdfs), "bajo_guide_mixture_sample: NULL buffer");
  if (n == 0) return BAJO_OK;
  bj::guide_mixture_kernel<<<bj::div_up(n, 256), 256, 0, ctx->stream>>>(outputs, n, uniforms, dirs, pdfs);
  BJ_HIP(hipGetLastError());
  return BAJO_OK;
}

}  // extern "C"
