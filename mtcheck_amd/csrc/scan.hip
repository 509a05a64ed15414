// scan.hip -- device-wide exclusive prefix sum over uint32 counts (tiles of 2048).
#include "internal.hpp"

namespace mtc {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread across a 256-thread block.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* ws) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < kScanThreads / 64; i++) {
    if (i < w) pre += ws[i];
    tot += ws[i];
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(const uint32_t* in, uint64_t count, uint32_t* sums) {
  __shared__ uint32_t ws[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; k++)
    if (base + k < count) s += in[base + k];
  uint32_t tot;
  block_excl_scan(s, &tot, ws);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out may alias in.  tile_base: scanned tile sums (or null for one tile).
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* in, uint32_t* out, uint64_t count,
                                                             const uint32_t* tile_base, uint32_t* total) {
  __shared__ uint32_t ws[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; k++) {
    v[k] = (base + k < count) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, &tot, ws) + (tile_base ? tile_base[blockIdx.x] : 0u);
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; k++) {
    if (base + k < count) out[base + k] = pre;
    pre += v[k];
  }
  if (total && !tile_base && threadIdx.x == 0) *total = tot;
}

// One level of the scan.  Tile sums go to sums[depth]: the workspace's buffers
// (grown on demand) or, when fixed is given, caller-provided buffers of at least
// scan_sums_words(count, depth) words.
static void scan_level(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total, Workspace* ws,
                       uint32_t* const* fixed, int depth, hipStream_t st) {
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    k_scan_apply<<<1, kScanThreads, 0, st>>>(in, out, count, nullptr, total);
    MTC_HIP(hipGetLastError());
    return;
  }
  if (depth > 2) throw HipError{hipErrorInvalidValue, "scan depth"};
  uint32_t* sums = nullptr;
  if (fixed) {
    sums = fixed[depth];
  } else {
    DevBuf<uint32_t>& buf = depth == 0 ? ws->u32a : (depth == 1 ? ws->u32b : ws->u32c);
    buf.reserve(tiles);
    sums = buf.p;
  }
  k_scan_sums<<<(uint32_t)tiles, kScanThreads, 0, st>>>(in, count, sums);
  MTC_HIP(hipGetLastError());
  scan_level(sums, sums, tiles, total, ws, fixed, depth + 1, st);
  k_scan_apply<<<(uint32_t)tiles, kScanThreads, 0, st>>>(in, out, count, sums, nullptr);
  MTC_HIP(hipGetLastError());
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total, Workspace& ws,
                        hipStream_t st) {
  scan_level(in, out, count, total, &ws, nullptr, 0, st);  // count == 0 still writes *total = 0
}

uint64_t scan_sums_words(uint64_t count, int depth) {
  uint64_t t = count;
  for (int d = 0; d <= depth; d++) t = (t + kScanTile - 1) / kScanTile;
  return t;
}

void exclusive_scan_u32_fixed(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total,
                              uint32_t* const sums[3], hipStream_t st) {
  scan_level(in, out, count, total, nullptr, sums, 0, st);
}

}  // namespace mtc
