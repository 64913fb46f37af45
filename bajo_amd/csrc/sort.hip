// Onesweep LSD radix sort for gfx950 (wave64).
//
// Replaces bajo/sort/gpu/onesweep.mojo (global_histogram :31-99, scan_global
// :102-118, digit_binning :121-327) and the reduce-then-scan sort of
// bajo/sort/gpu/radix_sort.mojo:510-609 with one engine:
//   1. one histogram pass computes the 256-bin digit histogram of every
//      pass at once (per-wave LDS histograms, one global atomic per bin);
//   2. one binning launch per 8-bit digit: a tile of 4096 keys is taken by
//      atomic ticket (tickets are handed out in dispatch order, so every tile
//      a block waits on is already resident), ranked stably with wave64
//      ballot match masks, published to a decoupled look-back table, staged
//      in LDS as full-width keys (64-bit keys stay 64-bit; bajo's
//      onesweep.mojo:244,311 staged u32) and written out in digit-contiguous
//      runs.
// The global exclusive digit offsets are recomputed by every tile from the
// 256-bin histogram, so no separate scan launch is needed.
#include <mutex>

#include "sort_internal.h"

namespace bj {

namespace {

template <typename K>
BJ_D uint32_t digit_of(K key, uint32_t shift) {
  return (uint32_t)(key >> shift) & 0xFFu;
}

// Exclusive scan of one value per thread over a 256-thread block.
BJ_D uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) {
    uint32_t t = s_tmp[w];
    before += (w < (int)wave) ? t : 0u;
    all += t;
  }
  __syncthreads();
  if (total) *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(256) sort_init_kernel(uint32_t* hist, uint32_t hist_words, uint32_t* tickets,
                                                        uint32_t ticket_words) {
  for (uint32_t i = threadIdx.x; i < hist_words; i += blockDim.x) hist[i] = 0;
  for (uint32_t i = threadIdx.x; i < ticket_words; i += blockDim.x) tickets[i] = 0;
}

template <typename K>
__global__ void __launch_bounds__(kHistBlock) sort_hist_kernel(const K* __restrict__ keys, uint64_t n,
                                                               uint32_t begin_bit, uint32_t passes,
                                                               uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ lookback,
                                                               uint64_t lookback_words) {
  __shared__ uint32_t s_hist[kHistBlock / 64][8 * 256];
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < (kHistBlock / 64) * 8 * 256; i += blockDim.x) (&s_hist[0][0])[i] = 0;
  // The look-back table of every pass starts NOT_READY.
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t gstride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gtid; i < lookback_words; i += gstride) lookback[i] = 0;
  __syncthreads();
  for (uint64_t i = gtid; i < n; i += gstride) {
    K k = keys[i];
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_hist[wave][p * 256 + digit_of(k, begin_bit + 8 * p)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < passes * 256; i += blockDim.x) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kHistBlock / 64; ++w) c += s_hist[w][i];
    if (c) atomicAdd(&hist[i], c);
  }
}

template <typename K, bool kValues>
__global__ void __launch_bounds__(kSortBlock) sort_bin_kernel(const K* __restrict__ in_keys, K* __restrict__ out_keys,
                                                              const uint32_t* __restrict__ in_vals,
                                                              uint32_t* __restrict__ out_vals, uint64_t n,
                                                              uint32_t shift, const uint32_t* __restrict__ pass_hist,
                                                              uint32_t* __restrict__ lookback,
                                                              uint32_t* __restrict__ ticket,
                                                              uint32_t* __restrict__ status, uint32_t spin_limit) {
  __shared__ K s_keys[kSortTile];
  __shared__ uint32_t s_vals[kValues ? kSortTile : 1];
  __shared__ uint32_t s_wave_hist[kSortWaves][256];
  __shared__ uint32_t s_digit_start[256];
  __shared__ uint32_t s_adjust[256];
  __shared__ uint32_t s_tmp[kSortWaves];
  __shared__ uint32_t s_tile;

  const uint32_t t = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = t >> 6;

  if (t == 0) s_tile = atomicAdd(ticket, 1u);
  for (uint32_t i = t; i < kSortWaves * 256; i += kSortBlock) (&s_wave_hist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t tile_base = (uint64_t)tile * kSortTile;

  // ---- load (wave-striped: each load instruction covers 64 consecutive keys)
  K key[kSortKpt];
  uint32_t val[kValues ? kSortKpt : 1];
  uint32_t rank[kSortKpt];
  const uint64_t wave_base = tile_base + (uint64_t)wave * 64 * kSortKpt;
#pragma unroll
  for (int k = 0; k < kSortKpt; ++k) {
    uint64_t idx = wave_base + (uint64_t)k * 64 + lane;
    bool valid = idx < n;
    key[k] = valid ? in_keys[idx] : (K)0;
    if constexpr (kValues) val[k] = valid ? in_vals[idx] : 0u;
  }

  // ---- stable wave-level ranking with ballot match masks
#pragma unroll
  for (int k = 0; k < kSortKpt; ++k) {
    uint64_t idx = wave_base + (uint64_t)k * 64 + lane;
    bool valid = idx < n;
    uint32_t d = digit_of(key[k], shift);
    uint64_t match = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t m = ballot(bit);
      match &= bit ? m : ~m;
    }
    uint32_t below = mask_rank(match);
    uint32_t base = valid ? s_wave_hist[wave][d] : 0u;
    rank[k] = base + below;
    bool last = valid && ((match >> lane) == 1ull);
    __builtin_amdgcn_wave_barrier();
    if (last) s_wave_hist[wave][d] = base + (uint32_t)__builtin_popcountll(match);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  // ---- per-digit block counts and per-wave exclusive offsets (thread t = digit t)
  uint32_t run = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) {
    uint32_t c = s_wave_hist[w][t];
    s_wave_hist[w][t] = run;
    run += c;
  }
  const uint32_t count = run;

  // ---- decoupled look-back: publish this tile's count, then sum predecessors
  uint32_t* my_entry = lookback + (uint64_t)tile * 256 + t;
  uint32_t exclusive = 0;
  if (tile == 0) {
    __hip_atomic_store(my_entry, kLookbackInc | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(my_entry, kLookbackAgg | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Predecessors' entries are read kLookAhead at a time (their loads in
    // flight together) and consumed in order: aggregates are added until an
    // inclusive prefix, or the first entry not yet published, which is polled
    // again.
    constexpr int kLookAhead = 4;
    int64_t j = (int64_t)tile - 1;
    uint32_t spins = 0;
    bool done = false;
    while (!done && j >= 0) {
      uint32_t v[kLookAhead];
#pragma unroll
      for (int k = 0; k < kLookAhead; ++k)
        v[k] = j - k >= 0 ? __hip_atomic_load(lookback + (uint64_t)(j - k) * 256 + t, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
      int k = 0;
      bool waiting = false;
      for (; k < kLookAhead && j - k >= 0; ++k) {
        const uint32_t flag = v[k] & ~kLookbackMask;
        if (flag == 0) {
          waiting = true;
          break;
        }
        exclusive += v[k] & kLookbackMask;
        if (flag == kLookbackInc) {
          done = true;
          break;
        }
      }
      if (done) break;
      j -= k;
      if (waiting) {
        // Every predecessor holds an earlier ticket and is resident, so it
        // publishes; the cap only bounds the wait if the table was corrupted,
        // and giving up is reported through the status word.
        if (++spins > spin_limit) {
          atomicMax(status, (uint32_t)BAJO_STATUS_SORT_LOOKBACK);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(my_entry, kLookbackInc | (exclusive + count), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // ---- block-local digit starts and global destinations
  uint32_t digit_start = block_exclusive_scan_256(count, s_tmp, nullptr);
  uint32_t global_excl = block_exclusive_scan_256(pass_hist[t], s_tmp, nullptr);
  s_digit_start[t] = digit_start;
  s_adjust[t] = global_excl + exclusive - digit_start;
  __syncthreads();

  // ---- scatter into LDS in sorted order
#pragma unroll
  for (int k = 0; k < kSortKpt; ++k) {
    uint64_t idx = wave_base + (uint64_t)k * 64 + lane;
    if (idx < n) {
      uint32_t d = digit_of(key[k], shift);
      uint32_t pos = s_digit_start[d] + s_wave_hist[wave][d] + rank[k];
      s_keys[pos] = key[k];
      if constexpr (kValues) s_vals[pos] = val[k];
    }
  }
  __syncthreads();

  // ---- write digit-contiguous runs
  const uint64_t rem = n - tile_base;
  const uint32_t tile_count = rem < (uint64_t)kSortTile ? (uint32_t)rem : (uint32_t)kSortTile;
  for (uint32_t i = t; i < tile_count; i += kSortBlock) {
    K k = s_keys[i];
    uint32_t pos = s_adjust[digit_of(k, shift)] + i;
    if (pos < n) {
      out_keys[pos] = k;
      if constexpr (kValues) out_vals[pos] = s_vals[i];
    }
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename K>
int32_t run_typed(bajo_ctx* ctx, const SortPlan& p, K* keys, uint32_t* values, char* ws, bool hist_ready) {
  K* alt_keys = (K*)(ws + p.off_alt_keys);
  uint32_t* alt_vals = p.has_values ? (uint32_t*)(ws + p.off_alt_vals) : nullptr;
  uint32_t* hist = (uint32_t*)(ws + p.off_hist);
  uint32_t* lookback = (uint32_t*)(ws + p.off_lookback);
  uint32_t* tickets = (uint32_t*)(ws + p.off_tickets);
  const uint64_t lb_pass_words = (uint64_t)p.tiles * 256;
  if (!hist_ready) {
    sort_init_kernel<<<1, 256, 0, ctx->stream>>>(hist, (uint32_t)p.passes * 256, tickets, (uint32_t)p.passes);
    BJ_HIP(hipGetLastError());
    unsigned blocks = div_up(p.n, (uint64_t)kHistBlock * 16);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    sort_hist_kernel<K><<<blocks, kHistBlock, 0, ctx->stream>>>(keys, p.n, (uint32_t)p.begin_bit, (uint32_t)p.passes,
                                                                hist, lookback, lb_pass_words * p.passes);
    BJ_HIP(hipGetLastError());
  }
  K* src = keys;
  K* dst = alt_keys;
  uint32_t* vsrc = values;
  uint32_t* vdst = alt_vals;
  const uint32_t spin_limit = ctx->spin_limit ? ctx->spin_limit : kLookbackSpinLimit;
  for (int pass = 0; pass < p.passes; ++pass) {
    uint32_t shift = (uint32_t)(p.begin_bit + 8 * pass);
    if (p.has_values)
      sort_bin_kernel<K, true><<<p.tiles, kSortBlock, 0, ctx->stream>>>(
          src, dst, vsrc, vdst, p.n, shift, hist + pass * 256, lookback + lb_pass_words * pass, tickets + pass,
          ctx->status, spin_limit);
    else
      sort_bin_kernel<K, false><<<p.tiles, kSortBlock, 0, ctx->stream>>>(
          src, dst, nullptr, nullptr, p.n, shift, hist + pass * 256, lookback + lb_pass_words * pass, tickets + pass,
          ctx->status, spin_limit);
    BJ_HIP(hipGetLastError());
    K* tk = src; src = dst; dst = tk;
    uint32_t* tv = vsrc; vsrc = vdst; vdst = tv;
  }
  if (src != keys) {
    BJ_HIP(hipMemcpyAsync(keys, src, p.n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
    if (p.has_values) BJ_HIP(hipMemcpyAsync(values, vsrc, p.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
  }
  return BAJO_OK;
}

}  // namespace

SortPlan make_sort_plan(uint64_t n, int key_bytes, int has_values, int begin_bit, int end_bit) {
  SortPlan p;
  p.n = n;
  p.key_bytes = key_bytes;
  p.has_values = has_values;
  p.begin_bit = begin_bit;
  p.end_bit = end_bit;
  p.passes = (end_bit - begin_bit + 7) / 8;
  p.tiles = div_up(n ? n : 1, kSortTile);
  size_t off = 0;
  p.off_alt_keys = off;
  off = align256(off + n * (size_t)key_bytes);
  p.off_alt_vals = off;
  off = align256(off + (has_values ? n * sizeof(uint32_t) : 0));
  p.off_hist = off;
  off = align256(off + (size_t)(p.passes ? p.passes : 1) * 256 * sizeof(uint32_t));
  p.off_lookback = off;
  off = align256(off + (size_t)(p.passes ? p.passes : 1) * p.tiles * 256 * sizeof(uint32_t));
  p.off_tickets = off;
  off = align256(off + 8 * sizeof(uint32_t));
  p.total = off;
  return p;
}

int32_t onesweep_run(bajo_ctx* ctx, const SortPlan& plan, void* keys, uint32_t* values, void* ws, bool hist_ready) {
  if (plan.n <= 1 || plan.passes == 0) return BAJO_OK;
  if (plan.key_bytes == 8) return run_typed<uint64_t>(ctx, plan, (uint64_t*)keys, values, (char*)ws, hist_ready);
  return run_typed<uint32_t>(ctx, plan, (uint32_t*)keys, values, (char*)ws, hist_ready);
}

}  // namespace bj

namespace {

int32_t sort_entry(bajo_ctx* ctx, void* keys, uint32_t* values, uint64_t n, int key_bytes, int begin_bit,
                   int end_bit, void* workspace, uint64_t workspace_bytes, const char* name) {
  BJ_REQUIRE(ctx != nullptr, "%s: ctx is NULL", name);
  BJ_REQUIRE(n == 0 || keys != nullptr, "%s: keys is NULL", name);
  BJ_REQUIRE(n <= bj::kSortMaxKeys, "%s: n=%llu exceeds the 2^30-1 key limit", name, (unsigned long long)n);
  BJ_REQUIRE(begin_bit >= 0 && end_bit <= key_bytes * 8 && begin_bit <= end_bit && begin_bit % 8 == 0 &&
                 end_bit % 8 == 0,
             "%s: bit range [%d, %d) must be byte aligned inside the key", name, begin_bit, end_bit);
  if (n <= 1 || begin_bit == end_bit) return BAJO_OK;
  bj::SortPlan plan = bj::make_sort_plan(n, key_bytes, values != nullptr, begin_bit, end_bit);
  void* ws = workspace;
  if (ws) {
    BJ_REQUIRE(workspace_bytes >= plan.total, "%s: workspace of %llu bytes is below the %llu required", name,
               (unsigned long long)workspace_bytes, (unsigned long long)plan.total);
  } else {
    int32_t rc = bj::ctx_scratch(ctx, plan.total, &ws);
    if (rc) return rc;
  }
  return bj::onesweep_run(ctx, plan, keys, values, ws, false);
}

// ---- fixed-scratch sort (no allocation) --------------------------------------
// The workspace of bajo_onesweep_radix_sort_pairs_u64_u32_noalloc lives in a
// file-scope device array: alternate keys and values, digit histograms,
// look-back table and tile tickets of one sort of at most kNoAllocMaxKeys
// pairs.  Each device holds its own copy for the life of the process; host
// code uses the address hipGetSymbolAddress returns for the current device,
// never the array's name.
constexpr uint64_t kNoAllocMaxKeys = 1ull << 20;
constexpr size_t kNoAllocScratchBytes = 16ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// with that device current.
int32_t noalloc_scratch(int device, void** out) {
  BJ_REQUIRE(device >= 0 && device < kNoAllocMaxDevices, "noalloc sort: device %d outside [0, %d)", device,
             kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    int prev = 0;
    BJ_HIP(hipGetDevice(&prev));
    BJ_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return bj::hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    BJ_REQUIRE(ptr != nullptr, "noalloc sort: hipGetSymbolAddress returned NULL");
    g_noalloc_ptr[device] = ptr;
  }
  *out = g_noalloc_ptr[device];
  return BAJO_OK;
}

}  // namespace

extern "C" {

uint64_t bajo_sort_workspace_bytes(uint64_t n, int32_t key_bytes, int32_t has_values) {
  if (key_bytes != 4 && key_bytes != 8) return 0;
  return bj::make_sort_plan(n, key_bytes, has_values != 0, 0, key_bytes * 8).total;
}

int32_t bajo_device_radix_sort_pairs_u64_u32(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                             void* workspace, uint64_t workspace_bytes) {
  BJ_REQUIRE(n == 0 || values != nullptr, "bajo_device_radix_sort_pairs_u64_u32: values is NULL");
  return sort_entry(ctx, keys, values, n, 8, 0, 64, workspace, workspace_bytes, "bajo_device_radix_sort_pairs_u64_u32");
}

int32_t bajo_device_radix_sort_pairs_u32_u32(bajo_ctx* ctx, uint32_t* keys, uint32_t* values, uint64_t n,
                                             void* workspace, uint64_t workspace_bytes) {
  BJ_REQUIRE(n == 0 || values != nullptr, "bajo_device_radix_sort_pairs_u32_u32: values is NULL");
  return sort_entry(ctx, keys, values, n, 4, 0, 32, workspace, workspace_bytes, "bajo_device_radix_sort_pairs_u32_u32");
}

int32_t bajo_device_radix_sort_keys_u64(bajo_ctx* ctx, uint64_t* keys, uint64_t n, void* workspace,
                                        uint64_t workspace_bytes) {
  return sort_entry(ctx, keys, nullptr, n, 8, 0, 64, workspace, workspace_bytes, "bajo_device_radix_sort_keys_u64");
}

int32_t bajo_device_radix_sort_keys_u32(bajo_ctx* ctx, uint32_t* keys, uint64_t n, void* workspace,
                                        uint64_t workspace_bytes) {
  return sort_entry(ctx, keys, nullptr, n, 4, 0, 32, workspace, workspace_bytes, "bajo_device_radix_sort_keys_u32");
}

int32_t bajo_onesweep_radix_sort_keys_u64(bajo_ctx* ctx, uint64_t* keys, uint64_t n, int32_t begin_bit,
                                          int32_t end_bit, void* workspace, uint64_t workspace_bytes) {
  return sort_entry(ctx, keys, nullptr, n, 8, begin_bit, end_bit, workspace, workspace_bytes,
                    "bajo_onesweep_radix_sort_keys_u64");
}

int32_t bajo_onesweep_radix_sort_pairs_u64_u32(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                               int32_t begin_bit, int32_t end_bit, void* workspace,
                                               uint64_t workspace_bytes) {
  BJ_REQUIRE(n == 0 || values != nullptr, "bajo_onesweep_radix_sort_pairs_u64_u32: values is NULL");
  return sort_entry(ctx, keys, values, n, 8, begin_bit, end_bit, workspace, workspace_bytes,
                    "bajo_onesweep_radix_sort_pairs_u64_u32");
}
int32_t bajo_onesweep_radix_sort_pairs_u64_u32_noalloc(bajo_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t n,
                                                       int32_t begin_bit, int32_t end_bit) {
  const char* name = "bajo_onesweep_radix_sort_pairs_u64_u32_noalloc";
  BJ_REQUIRE(ctx != nullptr, "%s: ctx is NULL", name);
  BJ_REQUIRE(n == 0 || (keys != nullptr && values != nullptr), "%s: keys or values is NULL", name);
  BJ_REQUIRE(n <= kNoAllocMaxKeys, "%s: n=%llu exceeds the fixed scratch's limit of %llu pairs", name,
             (unsigned long long)n, (unsigned long long)kNoAllocMaxKeys);
  BJ_REQUIRE(begin_bit >= 0 && end_bit <= 64 && begin_bit <= end_bit && begin_bit % 8 == 0 && end_bit % 8 == 0,
             "%s: bit range [%d, %d) must be byte aligned inside the key", name, begin_bit, end_bit);
  if (n <= 1 || begin_bit == end_bit) return BAJO_OK;
  const bj::SortPlan plan = bj::make_sort_plan(n, 8, 1, begin_bit, end_bit);
  BJ_REQUIRE(plan.total <= kNoAllocScratchBytes, "%s: workspace of %llu bytes exceeds the fixed scratch", name,
             (unsigned long long)plan.total);
  void* ws = nullptr;
  const int32_t rc = noalloc_scratch(ctx->device, &ws);
  if (rc) return rc;
  return bj::onesweep_run(ctx, plan, keys, values, ws, false);
}

}  // extern "C"
