// workload.hip -- device side of the synthetic checkpoint-series generator
// (include/mtcheck_workload.h).
#include "../../include/mtcheck_workload.h"
#include "internal.hpp"

namespace mtc {

constexpr uint32_t kWThreads = 256;

__device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t stream_word(uint64_t key, uint64_t w) {
  return sm_mix(key + (w + 1) * 0x9E3779B97F4A7C15ull);
}

// Bytes [begin, end): whole 8-byte words are stored at once, edges byte by byte.
__device__ inline void fill_range(uint8_t* buf, uint64_t begin, uint64_t end, uint64_t key, bool zero, uint32_t t,
                                  uint32_t nt) {
  if (begin >= end) return;
  const uint64_t w0 = begin >> 3, w1 = (end + 7) >> 3;
  for (uint64_t w = w0 + t; w < w1; w += nt) {
    const uint64_t v = zero ? 0 : stream_word(key, w);
    const uint64_t b0 = w << 3;
    if (b0 >= begin && b0 + 8 <= end) {
      *reinterpret_cast<uint64_t*>(buf + b0) = v;
    } else {
      for (uint32_t b = 0; b < 8; b++)
        if (b0 + b >= begin && b0 + b < end) buf[b0 + b] = (uint8_t)(v >> (8 * b));
    }
  }
}

__global__ __launch_bounds__(kWThreads) void k_wl_stream(uint8_t* buf, uint64_t begin, uint64_t end, uint64_t key) {
  const uint64_t words = ((end + 7) >> 3) - (begin >> 3);
  const uint64_t per = (words + gridDim.x - 1) / gridDim.x;
  const uint64_t wb = (begin >> 3) + per * blockIdx.x;
  const uint64_t lo = std::max<uint64_t>(begin, wb << 3);
  const uint64_t hi = std::min<uint64_t>(end, (wb + per) << 3);
  fill_range(buf, lo, hi, key, false, threadIdx.x, kWThreads);
}

__global__ __launch_bounds__(kWThreads) void k_wl_runs(uint8_t* buf, uint64_t len, uint32_t C, const uint32_t* runs,
                                                       uint32_t nruns, uint64_t key, int zero) {
  for (uint32_t r = blockIdx.x; r < nruns; r += gridDim.x) {
    const uint64_t b = (uint64_t)runs[2 * r] * C;
    uint64_t e = b + (uint64_t)runs[2 * r + 1] * C;
    if (e > len) e = len;
    if (b < e) fill_range(buf, b, e, key, zero != 0, threadIdx.x, kWThreads);
  }
}

__global__ __launch_bounds__(kWThreads) void k_wl_gather(const uint8_t* buf, uint64_t len, uint32_t C,
                                                         const uint32_t* pairs, uint32_t m, uint8_t* tmp) {
  for (uint32_t i = blockIdx.x; i < m; i += gridDim.x) {
    const uint64_t s = (uint64_t)pairs[2 * i + 1] * C;
    for (uint32_t b = threadIdx.x; b < C; b += kWThreads) tmp[(uint64_t)i * C + b] = (s + b < len) ? buf[s + b] : 0;
  }
}

__global__ __launch_bounds__(kWThreads) void k_wl_scatter(uint8_t* buf, uint64_t len, uint32_t C,
                                                          const uint32_t* pairs, uint32_t m, const uint8_t* tmp) {
  for (uint32_t i = blockIdx.x; i < m; i += gridDim.x) {
    const uint64_t d = (uint64_t)pairs[2 * i] * C;
    for (uint32_t b = threadIdx.x; b < C; b += kWThreads)
      if (d + b < len) buf[d + b] = tmp[(uint64_t)i * C + b];
  }
}

}  // namespace mtc

using mtc::HipError;

template <class F>
static int wl_guard(F&& f) {
  try {
    f();
    return MTC_OK;
  } catch (const HipError& e) {
    (void)hipGetLastError();
    return e.err == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
  }
}

extern "C" {

int mtc_wl_fill_stream(void* dev_buf, uint64_t begin, uint64_t end, uint64_t key, void* stream) {
  if (!dev_buf || begin > end) return MTC_ERR_INVALID_ARG;
  if (begin == end) return MTC_OK;
  return wl_guard([&] {
    const uint64_t words = ((end + 7) >> 3) - (begin >> 3);
    const uint64_t grid = std::min<uint64_t>(65536, std::max<uint64_t>(1, words / 4096));
    mtc::k_wl_stream<<<(uint32_t)grid, mtc::kWThreads, 0, (hipStream_t)stream>>>((uint8_t*)dev_buf, begin, end, key);
    MTC_HIP(hipGetLastError());
  });
}

int mtc_wl_fill_runs(void* dev_buf, uint64_t len, uint32_t chunk_size, const uint32_t* dev_runs, uint32_t nruns,
                     uint64_t key, int zero, void* stream) {
  if (!dev_buf || !chunk_size || (nruns && !dev_runs)) return MTC_ERR_INVALID_ARG;
  if (!nruns) return MTC_OK;
  return wl_guard([&] {
    mtc::k_wl_runs<<<std::min<uint32_t>(nruns, 65536), mtc::kWThreads, 0, (hipStream_t)stream>>>(
        (uint8_t*)dev_buf, len, chunk_size, dev_runs, nruns, key, zero);
    MTC_HIP(hipGetLastError());
  });
}

int mtc_wl_copy_chunks(void* dev_buf, uint64_t len, uint32_t chunk_size, const uint32_t* dev_pairs, uint32_t npairs,
                       void* dev_tmp, void* stream) {
  if (!dev_buf || !chunk_size || (npairs && (!dev_pairs || !dev_tmp))) return MTC_ERR_INVALID_ARG;
  if (!npairs) return MTC_OK;
  return wl_guard([&] {
    const uint32_t grid = std::min<uint32_t>(npairs, 65536);
    mtc::k_wl_gather<<<grid, mtc::kWThreads, 0, (hipStream_t)stream>>>((const uint8_t*)dev_buf, len, chunk_size,
                                                                      dev_pairs, npairs, (uint8_t*)dev_tmp);
    mtc::k_wl_scatter<<<grid, mtc::kWThreads, 0, (hipStream_t)stream>>>((uint8_t*)dev_buf, len, chunk_size, dev_pairs,
                                                                       npairs, (const uint8_t*)dev_tmp);
    MTC_HIP(hipGetLastError());
  });
}

}  // extern "C"
