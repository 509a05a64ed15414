// dedup.hip -- the checkpoint path: leaf hashing and labelling, the two fused
// Merkle-forest sweeps, root compaction in node-id order, and diff assembly.
//
// Serial semantics followed: include/reference_impl.hpp:15-132 (labels and region
// roots) and src/tree_approach.cpp:511-710 (diff layout), in the canonical order
// of the serial reference (roots ascending; shifted entries grouped by source
// checkpoint, ascending node id inside a group).
//
// Every kernel of the pipeline serves a batch of deduplicators ("members") in one
// launch: a Launch<A> carries one argument block per member and the block offsets
// of each member's share of the grid, so one checkpoint of one buffer is a batch
// of one and a step over hundreds of buffers costs the same number of launches.
// Grids depend only on the members' geometry and every data-dependent count (roots,
// shifted entries, chunks to insert) is read on the device, so a whole step has no
// host synchronisation until its end and can be captured as one hipGraph.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <tuple>
#include <unordered_set>

#include "args.hpp"

namespace mtc {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kFuse = 8;                  // most tree levels fused per sweep stage (one wavefront tile)
constexpr uint64_t kFirstTag = 0xFFFFFFFFull;  // staging tag of a first-occurrence root
// Staging tag of a shifted root that is a leaf: its table slot is leaf_slot of its
// chunk (leaf n - 1 + c holds chunk c), read where the entry is consumed.  Table
// slots are at most 2^31, so neither tag is a slot.
constexpr uint64_t kLeafTag = 0xFFFFFFFEull;
constexpr uint32_t kScanTile = 8 * kThreads;   // elements per block-tile of the batched scan
constexpr uint32_t kScanOneItems = 16;          // elements per thread of a one-launch scan
constexpr uint32_t kScanOneMax = kScanOneItems * kThreads;
constexpr uint32_t kSortBlocks = 1024;         // most blocks per member of the stable sort
#ifndef MTC_TPW_A
#define MTC_TPW_A 8
#endif
constexpr uint32_t kTilesPerWaveA = MTC_TPW_A;  // stage-0 tiles per wavefront, sweep A (<= 64)
#ifndef MTC_TPW_B
#define MTC_TPW_B 8
#endif
constexpr uint32_t kTilesPerWaveB = MTC_TPW_B;  // stage-0 tiles per wavefront, sweep B (<= 64)
// Wavefronts per block of the sweeps.  Wavefronts of one block draw very
// different numbers of busy tiles, and a block holds its slot (and sweep A's LDS
// records) until its slowest wavefront ends: with four wavefronts per block only
// about 2500 of the chip's 8192 wavefront slots were busy during sweep A's first
// stage.  One wavefront per block: sweep A 178 -> 153 us, sweep B 210 -> 184 us
// (8 GiB, 128-byte chunks, 2% change).
#ifndef MTC_SWEEP_WPB_A
#define MTC_SWEEP_WPB_A 1
#endif
#ifndef MTC_SWEEP_WPB_B
#define MTC_SWEEP_WPB_B 1
#endif
#ifndef MTC_SWEEP_B_WAVES
#define MTC_SWEEP_B_WAVES 5  // sweep B's wavefronts per SIMD (register budget)
#endif
constexpr uint32_t kSweepWpbA = MTC_SWEEP_WPB_A;
constexpr uint32_t kSweepWpbB = MTC_SWEEP_WPB_B;

// ---------------------------------------------------------------------------
// Batched launches
// ---------------------------------------------------------------------------

// The member of this block: a table entry per 64 blocks, then a short forward walk
// (members with no blocks are skipped by the walk).  All values are wave-uniform.
template <class A>
__device__ __forceinline__ Where where(const Launch<A>& L) {
  const uint32_t b = blockIdx.x;
  uint32_t m = 0;
  if (L.count > 1) {
    m = L.group[b >> 6];
    while (L.prefix[m + 1] <= b) m++;
  }
  return Where{m, b - L.prefix[m], L.prefix[m + 1] - L.prefix[m]};
}

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void k_fill(Launch<FillSeg> L) {
  const Where w = where(L);
  const FillSeg s = L.args[w.m].global();
  for (uint64_t i = (uint64_t)w.blk * kThreads + threadIdx.x; i < s.count; i += (uint64_t)w.nblk * kThreads)
    s.p[i] = s.v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Per-checkpoint counters are striped over kStripes addresses (by block index) so
// that a million wavefronts do not serialise on one atomic; k_reduce_ctr folds them.
constexpr uint32_t kStripes = 256;
enum : uint32_t { CTR_CLAIMS = 0, CTR_CHANGED = 1, CTR_KINDS = 2 };

__device__ __forceinline__ unsigned long long* stripe(unsigned long long* ctr, uint32_t kind) {
  return ctr + kind * kStripes + (blockIdx.x & (kStripes - 1));
}

// Fold the striped counters, then copy the member's control block into the step's
// contiguous result array (one device-to-host copy for the whole batch).  Run by
// one whole block of kStripes threads, after every counter and control field of
// the checkpoint is final.
// s: CTR_KINDS words of LDS.
__device__ void finish_member(const unsigned long long* ctr, const uint32_t* changed, Ctl* ctl, Ctl* out,
                              unsigned long long* s) {
  if (threadIdx.x < CTR_KINDS) s[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t k = 0; k < CTR_KINDS; k++) {
    const unsigned long long v = ctr[k * kStripes + threadIdx.x];
    if (v) atomicAdd(&s[k], v);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->claims = s[CTR_CLAIMS];
    ctl->changed = *changed;
    ctl->t_end = wall_clock64();
  }
  __syncthreads();
  const uint32_t* src = reinterpret_cast<const uint32_t*>(ctl);
  uint32_t* dst = reinterpret_cast<uint32_t*>(out);
  if (threadIdx.x < sizeof(Ctl) / 4) dst[threadIdx.x] = src[threadIdx.x];
}

static uint32_t grid_for(uint64_t items, uint32_t per_block = kThreads, uint32_t cap = 1u << 20) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ inline uint32_t block_scan_excl(uint32_t v, uint32_t* total, uint32_t* ws) {
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
  for (uint32_t i = 0; i < kThreads / 64; i++) {
    if (i < w) pre += ws[i];
    tot += ws[i];
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

// ---------------------------------------------------------------------------
// Batched exclusive scan of u32 arrays in two launches: per-tile sums, then the
// tiles with their offsets (each block adds up the sums of the tiles before its
// own).  The element count may live on the device (count_dev), bounded by
// count_max.
// ---------------------------------------------------------------------------

// Element e covers [o, o + size) of the running total: it is the first element of
// every 256-position tile that starts inside that range.
__device__ __forceinline__ void mark_tiles(uint32_t* tile_lo, uint32_t o, uint32_t size, uint32_t e) {
  for (uint32_t b = (o + 255) / 256; (uint64_t)b * 256 < (uint64_t)o + size; b++) tile_lo[b] = e;
}

__device__ __forceinline__ uint32_t scan_count(const ScanSeg& s) {
  return s.count_dev ? min(*s.count_dev, s.count_max) : s.count_max;
}

__device__ __forceinline__ void scan_stamp(const ScanSeg& s, const Where& w) {
  if (s.stamp && w.blk == 0 && threadIdx.x == 0) *s.stamp = wall_clock64();
}

__device__ __forceinline__ void ph_scan_reduce(const Launch<ScanSeg>& L, const Where& w, uint8_t* lds) {
  uint32_t* ws = reinterpret_cast<uint32_t*>(lds);
  const ScanSeg s = L.args[w.m].global();
  scan_stamp(s, w);
  const uint32_t cnt = scan_count(s), ntiles = (cnt + kScanTile - 1) / kScanTile;
  for (uint32_t tile = w.blk; tile < ntiles; tile += w.nblk) {
    const uint64_t base = (uint64_t)tile * kScanTile;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      const uint64_t i = base + j * kThreads + threadIdx.x;
      if (i < cnt) v += s.in[i];
    }
    uint32_t tot;
    block_scan_excl(v, &tot, ws);
    if (threadIdx.x == 0) s.sums[tile] = tot;
  }
}

__global__ __launch_bounds__(kThreads) void k_scan_reduce(Launch<ScanSeg> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * (kThreads / 64)];
  ph_scan_reduce(L, where(L), lds);
}

// The second launch of the scan: each block sums the tile sums before its tiles
// itself (reading each sum once, as its tiles advance), so the top-level scan
// needs no launch of its own; the block of the last tile writes the total.
__device__ __forceinline__ void ph_scan_down2(const Launch<ScanSeg>& L, const Where& w, uint8_t* lds) {
  uint32_t* ws = reinterpret_cast<uint32_t*>(lds);
  const ScanSeg s = L.args[w.m].global();
  const uint32_t cnt = scan_count(s), ntiles = (cnt + kScanTile - 1) / kScanTile;
  if (ntiles == 0 && w.blk == 0 && threadIdx.x == 0 && s.total) *s.total = 0;
  uint32_t pre = 0, done = 0;  // pre = sums[0] + ... + sums[done - 1]
  for (uint32_t tile = w.blk; tile < ntiles; tile += w.nblk) {
    uint32_t acc = 0;
    for (uint32_t j = done + threadIdx.x; j < tile; j += kThreads) acc += s.sums[j];
    uint32_t more;
    block_scan_excl(acc, &more, ws);
    pre += more;
    const uint64_t base = (uint64_t)tile * kScanTile + 8ull * threadIdx.x;
    uint32_t v[8], run = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) {
      v[j] = (base + j < cnt) ? s.in[base + j] : 0u;
      const uint32_t x = v[j];
      v[j] = run;
      run += x;
    }
    uint32_t tot;
    const uint32_t off = pre + block_scan_excl(run, &tot, ws);
#pragma unroll
    for (uint32_t j = 0; j < 8; j++)
      if (base + j < cnt) {
        if (s.tile_lo) mark_tiles(s.tile_lo, off + v[j], (j + 1 < 8 ? v[j + 1] : run) - v[j], (uint32_t)(base + j));
        s.out[base + j] = off + v[j];
      }
    if (tile + 1 == ntiles && threadIdx.x == 0 && s.total) *s.total = pre + tot;
    pre += tot;  // this tile's sum
    done = tile + 1;
  }
}

__global__ __launch_bounds__(kThreads) void k_scan_down2(Launch<ScanSeg> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * (kThreads / 64)];
  ph_scan_down2(L, where(L), lds);
}

// A scan whose every segment has at most kScanOneMax elements: one block per
// segment scans it and writes its total, in one launch instead of two.
__device__ __forceinline__ void ph_scan_one(const Launch<ScanSeg>& L, const Where& w, uint8_t* lds) {
  uint32_t* ws = reinterpret_cast<uint32_t*>(lds);
  const ScanSeg s = L.args[w.m].global();
  scan_stamp(s, w);
  const uint32_t cnt = scan_count(s);
  const uint32_t base = kScanOneItems * threadIdx.x;
  uint32_t v[kScanOneItems], run = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanOneItems; j++) {
    v[j] = (base + j < cnt) ? s.in[base + j] : 0u;
    const uint32_t x = v[j];
    v[j] = run;
    run += x;
  }
  uint32_t tot;
  const uint32_t off = block_scan_excl(run, &tot, ws);
#pragma unroll
  for (uint32_t j = 0; j < kScanOneItems; j++)
    if (base + j < cnt) {
      if (s.tile_lo) mark_tiles(s.tile_lo, off + v[j], (j + 1 < kScanOneItems ? v[j + 1] : run) - v[j], base + j);
      s.out[base + j] = off + v[j];
    }
  if (threadIdx.x == 0 && s.total) *s.total = tot;
}

__global__ __launch_bounds__(kThreads) void k_scan_one(Launch<ScanSeg> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * (kThreads / 64)];
  ph_scan_one(L, where(L), lds);
}

// ---------------------------------------------------------------------------
// LDS-staged tile hashing.  A 256-thread block hashes 256 consecutive chunks, one
// per thread.  The chunks are brought in as segments of SEG bytes per chunk with
// fully coalesced 16-byte loads (each block-wide load instruction covers 4 KiB),
// written to LDS with a 16-byte pad per chunk (so the per-thread 16-byte reads
// of the hashing phase hit distinct banks), then each thread hashes its own chunk
// from LDS.  Murmur3 keeps (h1, h2) across segments; MD5 keeps its state and
// compresses every 64 bytes.
// ---------------------------------------------------------------------------
#ifndef MTC_SEG_MAX
#define MTC_SEG_MAX 128
#endif
constexpr int kSegMax = MTC_SEG_MAX;  // most bytes of a chunk staged per pass (LDS per block: 256 * (SEG + 16))

template <int CB>
struct TileShape {
  static constexpr int SEG = CB < kSegMax ? CB : kSegMax;  // bytes of each chunk per pass
  static constexpr int PITCH = SEG + 16;           // LDS bytes per chunk
  static constexpr int LOADS = SEG / 16;           // 16-byte loads per thread per pass
  static constexpr int PASSES = CB / SEG;
  static constexpr int LDS = kThreads * PITCH;
};

// Returns the digest of chunk (first + threadIdx.x) when it is a full chunk; the
// caller hashes a short last chunk separately.  `full` = number of full chunks.
template <int HK, int CB>
__device__ __forceinline__ Digest staged_hash(const uint8_t* data, uint64_t first, uint64_t full, uint8_t* lds) {
  using S = TileShape<CB>;
  const uint32_t t = threadIdx.x;
  uint64_t h1 = 0, h2 = 0;
  uint32_t st[4];
  md5_init(st);
  uint32_t X[16];
#pragma unroll 1
  for (int pass = 0; pass < S::PASSES; pass++) {
    uint4 v[S::LOADS];
#pragma unroll
    for (int i = 0; i < S::LOADS; i++) {
      const uint32_t u = i * (kThreads * 16) + 16 * t;  // byte within this pass's tile
      const uint32_t j = u / S::SEG, w = u % S::SEG;
      v[i] = uint4{0, 0, 0, 0};
      if (first + j < full)
        v[i] = *reinterpret_cast<const uint4*>(data + (first + j) * (uint64_t)CB + (uint64_t)pass * S::SEG + w);
    }
    __syncthreads();  // the previous pass's reads are done
#pragma unroll
    for (int i = 0; i < S::LOADS; i++) {
      const uint32_t u = i * (kThreads * 16) + 16 * t;
      const uint32_t j = u / S::SEG, w = u % S::SEG;
      *reinterpret_cast<uint4*>(lds + j * S::PITCH + w) = v[i];
    }
    __syncthreads();
    const uint8_t* mine = lds + t * S::PITCH;
    if constexpr (HK == HASH_MD5) {
#pragma unroll
      for (int q = 0; q < S::LOADS; q++) {
        const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * q);
        const int wq = q & 3;  // LOADS is a multiple of 4 whenever PASSES > 1
        X[4 * wq] = b.x;
        X[4 * wq + 1] = b.y;
        X[4 * wq + 2] = b.z;
        X[4 * wq + 3] = b.w;
        if (wq == 3) md5_compress(st, X);
      }
    } else {
#pragma unroll
      for (int q = 0; q < S::LOADS; q++) {
        const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * q);
        m3_block(h1, h2, (uint64_t)b.x | ((uint64_t)b.y << 32), (uint64_t)b.z | ((uint64_t)b.w << 32));
      }
    }
  }
  if constexpr (HK == HASH_MD5) {
    constexpr int RW = (CB % 64) / 16;  // words of a final partial 64-byte block already in X
    uint32_t T[16];
#pragma unroll
    for (int w = 0; w < 16; w++) T[w] = (w < 4 * RW) ? X[w] : 0u;
    md5_finish_words(st, T, (uint32_t)(CB % 64), (uint64_t)CB);
    return md5_digest(st);
  } else {
    return m3_final(h1, h2, CB);
  }
}

// ---------------------------------------------------------------------------
// Leaf kernel: hash, fixed test, table insert (src/tree_approach.cpp:305-351 and
// the serial loop of include/reference_impl.hpp:44-71).  One lane per chunk.
// ---------------------------------------------------------------------------

// Digest of chunk tile*256 + threadIdx.x (zero beyond n).  CB > 0 selects the
// LDS-staged path for 16-byte-aligned data and chunk sizes that are multiples of
// 16; a short last chunk and CB == 0 use byte-granular hashing.
template <int HK, int CB>
__device__ __forceinline__ Digest tile_digest(const uint8_t* data, uint64_t len, uint32_t n, uint32_t C,
                                              uint64_t tile) {
  const uint64_t first = tile * kThreads;
  const uint64_t c = first + threadIdx.x;
  Digest d{0, 0};
  if constexpr (CB > 0) {
    __shared__ __align__(16) uint8_t lds[TileShape<CB>::LDS];
    const uint64_t full = len / CB;
    d = staged_hash<HK, CB>(data, first, full, lds);
    if (c < n && c >= full) d = hash_bytes<HK>(data + c * CB, len - c * CB);
  } else {
    if (c < n) {
      const uint64_t off = c * C;
      d = hash_bytes<HK>(data + off, (off + C <= len) ? C : len - off);
    }
  }
  return d;
}

// Fixed test, label and packing of one tile's chunks (chunk t*256 + threadIdx.x
// has digest d).  Chunks to insert go to this block's region, ascending.
// The previous checkpoint's digest of this lane's chunk in tile t (zero after a
// reset).  Callers issue it with the tile's data loads so that its latency hides
// behind theirs.
__device__ __forceinline__ Digest leaf_prev(const LeafArgs& a, uint64_t t) {
  const uint64_t c = t * kThreads + threadIdx.x;
  if (c >= a.n || a.baseline) return Digest{0, 0};
  const uint4 v = ld_stream16(a.leaf_dg + c);
  return Digest{(uint64_t)v.x | ((uint64_t)v.y << 32), (uint64_t)v.z | ((uint64_t)v.w << 32)};
}

__device__ __forceinline__ void leaf_epilogue(const LeafArgs& a, uint64_t t, const Digest& d, const Digest& prev,
                                              uint32_t* region, uint32_t& run, uint32_t* s_wcnt) {
  const uint64_t c = t * kThreads + threadIdx.x;
  const bool valid = c < a.n;
  const bool fixed = valid && a.baseline != 2 && d == prev;
  // A fixed leaf's digest was inserted by an earlier checkpoint of this baseline,
  // so its insert is a no-op except right after a reset, where every leaf inserts.
  const bool ins = valid && (a.baseline || !fixed);
  if (valid && a.all_labels) a.labels[c] = fixed ? L_FIXED : L_PENDING;
  if (ins && !a.list_dg) a.leaf_dg[c] = d;
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  const uint64_t m = __ballot(ins);
  if (lane == 0) s_wcnt[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t base = run, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < kThreads / 64; i++) {
    base += (i < w) ? s_wcnt[i] : 0u;
    tot += s_wcnt[i];
  }
  if (ins) {
    const uint32_t k = base + __popcll(m & ((1ull << lane) - 1));
    region[k] = (uint32_t)c;
    if (a.list_dg && !a.hash_later) a.list_dg[(region - a.list) + k] = d;
  }
  run += tot;
  __syncthreads();  // s_wcnt is reused by the next tile
}

// One block streams a contiguous range of tiles (256 chunks each), software
// pipelined: the 16-byte loads of the next (tile, segment) are in flight while the
// current one is hashed out of LDS.  Chunks to insert are packed into the block's
// region of the list; the table work runs densely afterwards.
template <int HK, int CB>
__global__ __launch_bounds__(kThreads) void k_leaf(Launch<LeafArgs> L) {
  __shared__ uint32_t s_wcnt[kThreads / 64];
  const Where wh = where(L);
  const LeafArgs a = L.args[wh.m].global();
  if (a.stamp && wh.blk == 0 && threadIdx.x == 0) *a.stamp = wall_clock64();
  const uint64_t ntiles = (a.n + kThreads - 1) / kThreads;
  const uint64_t t0 = (uint64_t)wh.blk * a.tiles_per_block;
  const uint64_t t1 = std::min<uint64_t>(t0 + a.tiles_per_block, ntiles);
  uint32_t* region = a.list + t0 * kThreads;
  uint32_t run = 0;
  if constexpr (CB == 0) {
    for (uint64_t t = t0; t < t1; t++) {
      const Digest pv = leaf_prev(a, t);
      leaf_epilogue(a, t, tile_digest<HK, 0>(a.data, a.len, a.n, a.C, t), pv, region, run, s_wcnt);
    }
  } else {
    using S = TileShape<CB>;
    __shared__ __align__(16) uint8_t lds[S::LDS];
    const uint64_t full = a.len / CB;
    const uint64_t steps = (t1 > t0 ? t1 - t0 : 0) * S::PASSES;
    const uint32_t tid = threadIdx.x;
    auto load = [&](uint64_t step, uint4* v) {
      const uint64_t t = t0 + step / S::PASSES;
      const uint32_t pass = (uint32_t)(step % S::PASSES);
#pragma unroll
      for (int i = 0; i < S::LOADS; i++) {
        const uint32_t u = i * (kThreads * 16) + 16 * tid;
        const uint32_t j = u / S::SEG, w = u % S::SEG;
        const uint64_t chunk = t * kThreads + j;
        v[i] = uint4{0, 0, 0, 0};
        if (chunk < full) v[i] = ld_stream16(a.data + chunk * CB + pass * S::SEG + w);
      }
    };
    uint4 v[S::LOADS];
    Digest pv{0, 0};  // previous digest of this lane's chunk in the current tile
    if (steps) {
      load(0, v);
      pv = leaf_prev(a, t0);
    }
    uint64_t h1 = 0, h2 = 0;
    uint32_t st[4], X[16];
    for (uint64_t step = 0; step < steps; step++) {
      uint4 nv[S::LOADS];
      Digest npv = pv;
      if (step + 1 < steps) {
        load(step + 1, nv);
        if ((step + 1) % S::PASSES == 0) npv = leaf_prev(a, t0 + (step + 1) / S::PASSES);
      }
      __syncthreads();  // the previous segment's LDS reads are done
#pragma unroll
      for (int i = 0; i < S::LOADS; i++) {
        const uint32_t u = i * (kThreads * 16) + 16 * tid;
        *reinterpret_cast<uint4*>(lds + (u / S::SEG) * S::PITCH + u % S::SEG) = v[i];
      }
      __syncthreads();
      const uint32_t pass = (uint32_t)(step % S::PASSES);
      if (pass == 0) {
        h1 = h2 = 0;
        md5_init(st);
      }
      const uint8_t* mine = lds + tid * S::PITCH;
      if constexpr (HK == HASH_MD5) {
#pragma unroll
        for (int q = 0; q < S::LOADS; q++) {
          const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * q);
          const int wq = q & 3;  // LOADS is a multiple of 4 whenever PASSES > 1
          X[4 * wq] = b.x;
          X[4 * wq + 1] = b.y;
          X[4 * wq + 2] = b.z;
          X[4 * wq + 3] = b.w;
          if (wq == 3) md5_compress(st, X);
        }
      } else {
#pragma unroll
        for (int q = 0; q < S::LOADS; q++) {
          const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * q);
          m3_block(h1, h2, (uint64_t)b.x | ((uint64_t)b.y << 32), (uint64_t)b.z | ((uint64_t)b.w << 32));
        }
      }
      if (pass == S::PASSES - 1) {
        const uint64_t t = t0 + step / S::PASSES;
        const uint64_t c = t * kThreads + tid;
        Digest d;
        if constexpr (HK == HASH_MD5) {
          constexpr int RW = (CB % 64) / 16;
          uint32_t T[16];
#pragma unroll
          for (int w = 0; w < 16; w++) T[w] = (w < 4 * RW) ? X[w] : 0u;
          md5_finish_words(st, T, (uint32_t)(CB % 64), (uint64_t)CB);
          d = md5_digest(st);
        } else {
          d = m3_final(h1, h2, CB);
        }
        if (c >= full) d = (c < a.n) ? hash_bytes<HK>(a.data + c * CB, a.len - c * CB) : Digest{0, 0};
        leaf_epilogue(a, t, d, pv, region, run, s_wcnt);
      }
#pragma unroll
      for (int i = 0; i < S::LOADS; i++) v[i] = nv[i];
      pv = npv;
    }
  }
  if (threadIdx.x == 0) a.blk_cnt[wh.blk] = run;
}

// One 256-chunk tile per block, chunks of 16..128 bytes staged whole: the kernel
// used when a block takes a single tile.  LDS holds exactly the tile (32 KiB at
// 128-byte chunks: five blocks per CU), its 16-byte columns swizzled instead of
// padded: column ^ ((row / (128 / SEG) ^ row / (256 / SEG)) & (SEG / 16 - 1))
// gives the rows whose 16-byte accesses are served together distinct columns,
// whether the banks span 128 bytes (8 lanes per pass) or 256 bytes (16 lanes per
// pass), so both the row-wise stores and the per-lane reads are free of bank
// conflicts.  (A plain column ^ row swizzle leaves 2-way conflicts on the reads
// with 256-byte banks: SQ_LDS_BANK_CONFLICT was a quarter of SQ_LDS_IDX_ACTIVE.)
// Each wavefront reserves its run of the block's region with one atomic add on
// the block's count (zero when the kernel starts: filled by the step, or cleared
// by the previous step's k_compact), so no block-wide count is needed; within a
// region the runs are in arrival order.
template <int SEG>
__device__ __forceinline__ uint32_t lean_col(uint32_t row, uint32_t col) {
  return col ^ (((row / (128 / SEG)) ^ (row / (256 / SEG))) & (SEG / 16 - 1));
}

template <int HK, int CB>
__device__ __forceinline__ void leaf1_tile(const LeafArgs& a, uint32_t blk, uint8_t* lds) {
  static_assert(CB >= 16 && (CB & (CB - 1)) == 0, "power-of-two chunk size");
  constexpr int SEG = CB < 128 ? CB : 128;  // bytes of each chunk per pass
  constexpr int PASSES = CB / SEG, LOADS = SEG / 16;
  if (a.stamp && blk == 0 && threadIdx.x == 0) *a.stamp = wall_clock64();
  const uint64_t t = blk;
  const uint64_t full = a.len / CB;
  const uint32_t tid = threadIdx.x;
  const Digest pv = leaf_prev(a, t);
  const uint8_t* mine = lds + tid * SEG;
  uint64_t h1 = 0, h2 = 0;
  uint32_t st[4], X[16];
  md5_init(st);
#pragma unroll 1
  for (int pass = 0; pass < PASSES; pass++) {
    uint4 v[LOADS];
#pragma unroll
    for (int i = 0; i < LOADS; i++) {
      const uint32_t u = i * (kThreads * 16) + 16 * tid;
      const uint64_t chunk = t * kThreads + u / SEG;
      v[i] = uint4{0, 0, 0, 0};
      if (chunk < full) v[i] = ld_stream16(a.data + chunk * CB + (uint64_t)pass * SEG + u % SEG);
    }
    if (PASSES > 1) __syncthreads();  // the previous pass's reads are done
#pragma unroll
    for (int i = 0; i < LOADS; i++) {
      const uint32_t u = i * (kThreads * 16) + 16 * tid;
      const uint32_t j = u / SEG, col = (u % SEG) / 16;
      *reinterpret_cast<uint4*>(lds + j * SEG + 16 * lean_col<SEG>(j, col)) = v[i];
    }
    __syncthreads();
    if constexpr (HK == HASH_MD5) {
#pragma unroll
      for (int q = 0; q < LOADS; q++) {
        const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * lean_col<SEG>(tid, q));
        const int wq = q & 3;  // LOADS is a multiple of 4 whenever PASSES > 1
        X[4 * wq] = b.x;
        X[4 * wq + 1] = b.y;
        X[4 * wq + 2] = b.z;
        X[4 * wq + 3] = b.w;
        if (wq == 3) md5_compress(st, X);
      }
    } else {
#pragma unroll
      for (int q = 0; q < LOADS; q++) {
        const uint4 b = *reinterpret_cast<const uint4*>(mine + 16 * lean_col<SEG>(tid, q));
        m3_block(h1, h2, (uint64_t)b.x | ((uint64_t)b.y << 32), (uint64_t)b.z | ((uint64_t)b.w << 32));
      }
    }
  }
  Digest d;
  if constexpr (HK == HASH_MD5) {
    constexpr int RW = (CB % 64) / 16;  // words of a final partial 64-byte block already in X
    uint32_t T[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = (k < 4 * RW) ? X[k] : 0u;
    md5_finish_words(st, T, (uint32_t)(CB % 64), (uint64_t)CB);
    d = md5_digest(st);
  } else {
    d = m3_final(h1, h2, CB);
  }
  const uint64_t c = t * kThreads + tid;
  if (c >= full) d = (c < a.n) ? hash_bytes<HK>(a.data + c * CB, a.len - c * CB) : Digest{0, 0};
  const bool valid = c < a.n;
  const bool fixed = valid && a.baseline != 2 && d == pv;
  const bool ins = valid && (a.baseline || !fixed);  // as in leaf_epilogue
  // Writes beside the read stream are costly (a label byte per chunk cost the
  // kernel 10% of its rate at 128-byte chunks, scattered digests of changed chunks
  // as much again): once the labels read L_FIXED, none is written here (k_insert
  // labels the listed chunks), and each listed digest goes beside its list entry,
  // packed, for k_insert to move to leaf_dg.
  if (valid && a.all_labels) a.labels[c] = fixed ? L_FIXED : L_PENDING;
  if (ins && !a.list_dg) a.leaf_dg[c] = d;
  const uint32_t lane = lane_id();
  const uint64_t m = __ballot(ins);
  uint32_t base = 0;
  if (lane == 0 && m) base = atomicAdd(a.blk_cnt + blk, (uint32_t)__popcll(m));
  base = __shfl(base, 0);
  if (ins) {
    const uint64_t k = t * kThreads + base + __popcll(m & ((1ull << lane) - 1));
    a.list[k] = (uint32_t)c;
    if (a.list_dg && !a.hash_later) a.list_dg[k] = d;
  }
}

template <int HK, int CB>
__global__ __launch_bounds__(kThreads, 5) void k_leaf1(Launch<LeafArgs> L) {
  __shared__ __align__(16) uint8_t lds[kThreads * (CB < 128 ? CB : 128)];
  const Where wh = where(L);
  leaf1_tile<HK, CB>(L.args[wh.m].global(), wh.blk, lds);
}

// The same kernel for a step of one member, started before anything else of the
// step: its arguments come by value, and the threads of its first blocks copy the
// step's argument arena from pinned host memory to the device (what the later
// kernels read) and clear the step's counters before their tiles.  That saves the
// upload copy, the clearing kernel and the idle time in front of each.  The block
// counts k_leaf1 adds to are zero already: k_compact clears them after reading.
// (Spread over every block, the same work made the kernel 3% slower: each block
// read the fill list before its loads.)
constexpr uint32_t kPrologueBlocks = 64;
template <int HK, int CB>
__global__ __launch_bounds__(kThreads, 5) void k_leaf1_direct(LeafArgs a, Prologue p) {
  __shared__ __align__(16) uint8_t lds[kThreads * (CB < 128 ? CB : 128)];
  if (blockIdx.x < kPrologueBlocks) {
    const uint64_t gt = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint64_t gs = (uint64_t)min(gridDim.x, kPrologueBlocks) * kThreads;
    for (uint64_t i = gt; i < p.n16; i += gs) p.dst[i] = p.src[i];
    for (uint32_t f = 0; f < p.nfill; f++) {
      const FillSeg s = p.fill[f].global();
      for (uint64_t i = gt; i < s.count; i += gs) s.p[i] = s.v;
    }
  }
  leaf1_tile<HK, CB>(a.global(), blockIdx.x, lds);
}

// Concatenate the per-block regions into the dense list of chunks to insert.

__device__ __forceinline__ Digest chunk_m3_bytes(const uint8_t* p, uint64_t len) { return m3_bytes(p, len); }

// Murmur3 digest of chunk c read straight from memory with 16-byte loads, eight in
// flight (data 16-byte aligned, C a multiple of 16); a short last chunk byte by byte.
__device__ __forceinline__ Digest chunk_m3(const uint8_t* data, uint64_t len, uint32_t C, uint32_t c) {
  const uint64_t off = (uint64_t)c * C;
  if (off + C > len) return chunk_m3_bytes(data + off, len - off);
  const uint4* p = reinterpret_cast<const uint4*>(data + off);
  uint64_t h1 = 0, h2 = 0;
  const uint32_t nb = C / 16;
  uint32_t q = 0;
  for (; q + 8 <= nb; q += 8) {
    uint4 w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = p[q + k];
#pragma unroll
    for (int k = 0; k < 8; k++) m3_block(h1, h2, (uint64_t)w[k].x | ((uint64_t)w[k].y << 32), (uint64_t)w[k].z | ((uint64_t)w[k].w << 32));
  }
  for (; q < nb; q++) {
    const uint4 w = p[q];
    m3_block(h1, h2, (uint64_t)w.x | ((uint64_t)w.y << 32), (uint64_t)w.z | ((uint64_t)w.w << 32));
  }
  return m3_final(h1, h2, C);
}

constexpr uint32_t kCompactPer = 32;  // leaf blocks per compaction block

// The regions of kCompactPer consecutive leaf blocks land contiguously in the dense
// list (their offsets are a prefix sum), so each thread copies entries by their
// position in the group and finds the source region by a search in LDS.
__device__ __forceinline__ void ph_compact(const Launch<CompactArgs>& L, const Where& w, uint8_t* lds) {
  uint32_t* s_pre = reinterpret_cast<uint32_t*>(lds);
  const CompactArgs a = L.args[w.m].global();
  const uint32_t b0 = w.blk * kCompactPer, nb = min(kCompactPer, a.nblk - b0);
  // One wavefront loads the counts at once and scans them (a walk by one thread
  // waits for each load in turn).
  if (threadIdx.x < 64) {
    const uint32_t lane = threadIdx.x;
    const uint32_t v = lane < nb ? a.blk_cnt[b0 + lane] : 0u;
    if (lane == 0) s_pre[kCompactPer + 1] = a.blk_off[b0];
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane < nb) s_pre[lane] = x - v;
    if (lane + 1 == nb) s_pre[nb] = x;
  }
  __syncthreads();
  // The last reader of the counts clears them (the leaf scan read them before).
  if (threadIdx.x < nb) a.blk_cnt[b0 + threadIdx.x] = 0;
  const uint32_t total = s_pre[nb];
  const uint32_t off = s_pre[kCompactPer + 1];
  uint32_t* dst = a.dense + off;
  for (uint32_t e = threadIdx.x; e < total; e += kThreads) {
    uint32_t lo = 0, hi = nb;  // last region whose start <= e
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_pre[mid] <= e) lo = mid;
      else hi = mid;
    }
    const uint64_t src = (uint64_t)(b0 + lo) * a.tiles_per_block * kThreads + (e - s_pre[lo]);
    const uint32_t c = a.list[src];
    dst[e] = c;
    if (a.dense_dg) a.dense_dg[off + e] = a.hash_chunks ? chunk_m3(a.data, a.len, a.C, c) : a.list_dg[src];
  }
}

__global__ __launch_bounds__(kThreads, 8) void k_compact(Launch<CompactArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * (kCompactPer + 2)];
  ph_compact(L, where(L), lds);
}

// Stage-0 sweep tile of a chunk: bottom-level chunks by position, depth D-1 leaves
// by their position one level up.
__device__ __forceinline__ uint64_t chunk_tile(const Geo& g, uint32_t R0, uint64_t c) {
  const uint64_t lead = (1ull << g.D) - g.n;  // chunks stored at depth D-1
  if (c >= lead) return (c - lead) >> R0;
  return (c + g.half) >> (R0 - 1);
}

// Table inserts (src/tree_approach.cpp:320-349; include/reference_impl.hpp:51): the
// packed value (cur << 32 | leaf) with atomicMin leaves the lowest leaf of a new
// digest in the table, and leaves an earlier checkpoint's entry as it is (k_resolve
// then labels the leaf a shifted duplicate).  Each probe reads the slot's sector
// before any claim: a digest already in the table costs loads only, which matters
// for digests that many changed chunks share (a copied zero chunk), where claim-
// first atomics would serialise on one address; after a reset, the heads of all
// zero-filled runs meet there.  Consecutive entries with the same digest (one
// zero-filled run) share the first one's table access.
__device__ __forceinline__ void ph_insert(const Launch<DenseArgs>& L, const Where& w, uint8_t* lds) {
  (void)lds;
  const DenseArgs a = L.args[w.m].global();
  const uint64_t total = *a.total;
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)w.nblk * kThreads;
  uint32_t claims = 0;
  for (uint64_t wbase = (uint64_t)w.blk * kThreads + (threadIdx.x & ~63u); wbase < total; wbase += stride) {
    const uint64_t i = wbase + lane;
    bool v = i < total;
    uint32_t c = 0;
    Digest d{0, 0};
    if (v) {
      c = a.dense[i];
      d = a.dense_dg ? a.dense_dg[i] : a.leaf_dg[c];
    }
    const uint64_t plo = __shfl_up(d.lo, 1), phi = __shfl_up(d.hi, 1);
    const int pv = __shfl_up((int)v, 1);
    const uint32_t pc = __shfl_up(c, 1);
    // Same digest as the entry before, and a higher chunk: a block's runs of the
    // dense list may come in any order.
    const bool follower = v && lane > 0 && pv && plo == d.lo && phi == d.hi && pc < c;
    const bool head = v && !follower;
    uint32_t slot = ~0u;
    bool claimed = false;
    uint64_t seen = TEMPTY;
    if (head) slot = table_insert(a.tab, a.mask, d, ((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c), true,
                                  &a.ctl->err, &claimed, false, &seen, true);
    const uint64_t mine = __ballot(head) & ((lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1));
    const uint32_t hslot = __shfl(slot, mine ? 63 - __clzll((long long)mine) : (int)lane);
    if (follower) slot = hslot;
    if (v) {
      a.leaf_slot[c] = slot;
      if (a.dense_dg) a.leaf_dg[c] = d;
      // Every leaf that is not a fixed duplicate is on this list, so sweep B walks
      // only the stage-0 tiles flagged here.
      if (a.R0) a.tile_busy[chunk_tile(a.g, a.R0, c)] = 1;
    }
    // Decided here, so k_resolve skips them: a follower's leaf is above its head's,
    // so it never holds the minimum, and an entry of an earlier checkpoint keeps
    // its value.  A zero digest is left to k_resolve: after a reset it may belong
    // to a fixed leaf, whose label stays.  When the leaf kernel labelled nothing,
    // every listed chunk is a changed one (pending) until labelled here.
    const bool nz = (d.lo | d.hi) != 0;
    const bool older = head && !claimed && seen != TEMPTY && (uint32_t)(seen >> 32) < a.cur;
    int nl = (v && !a.leaf_labels) ? (int)L_PENDING : -1;
    if ((follower || older) && nz) nl = L_SHIFT;
    // A claimer is the first occurrence unless a marked match below it takes the
    // entry in k_fixmin, which then relabels it.
    if (claimed && nz) {
      nl = L_FIRST;
      if (a.R0) a.tile_first[chunk_tile(a.g, a.R0, c)] = 1;
    }
    // A head that matched an entry of this checkpoint: k_fixmin applies its min.
    if (head && !claimed && !older && slot <= a.mask + 1) {
      const uint8_t l = a.leaf_labels ? a.labels[c] : (uint8_t)L_PENDING;
      nl = l == L_FIXED ? L_MIN_FIXED : L_MIN_PENDING;
    }
    if (nl >= 0) a.labels[c] = (uint8_t)nl;
    // Only entries marked for a min or still pending (a zero digest, a failed
    // probe, or a leaf kernel's label kept) need k_fixmin and k_resolve.  A wave
    // lists them at the front of its own 64 words and counts them per group, so
    // no counter is shared between wavefronts (one returning atomic add per
    // wavefront on a single counter made k_insert 1.96 ms instead of 1.11 ms at 20%
    // change on 8 GiB of 128-byte chunks).
    const bool later = v && (nl < 0 || nl == L_PENDING || nl == L_MIN_PENDING || nl == L_MIN_FIXED);
    const uint64_t pm = __ballot(later);
    if (lane == 0) a.pcnt[wbase >> 6] = (uint8_t)__popcll(pm);
    if (later) a.pend[wbase + __popcll(pm & ((1ull << lane) - 1))] = c;
    claims += claimed;
  }
  const uint32_t wc = wave_sum(claims);
  if (lane == 0 && wc) atomicAdd(stripe(a.ctr, CTR_CLAIMS), (unsigned long long)wc);
}

__global__ __launch_bounds__(kThreads) void k_insert(Launch<DenseArgs> L) {
  ph_insert(L, where(L), nullptr);
}

// The chunks k_insert left for k_fixmin and k_resolve.  A wavefront takes 64
// groups of the dense list at a time (one count per lane), ranks their entries with
// a wave scan and gives each lane one entry per round: lane e of a round finds its
// group by bisection over the exclusive prefix, so every lane has work whether the
// entries sit in few groups or many.
constexpr uint32_t kPendPerBlock = (kThreads / 64) * 64 * 64;  // dense entries per block and pass
template <class F>
__device__ __forceinline__ void for_pending(const DenseArgs& a, const Where& w, F&& f) {
  const uint32_t lane = lane_id();
  const uint64_t groups = ((uint64_t)*a.total + 63) / 64;
  const uint64_t wave = ((uint64_t)w.blk * kThreads + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)w.nblk * kThreads) >> 6;
  for (uint64_t g0 = wave * 64; g0 < groups; g0 += waves * 64) {
    const uint32_t cnt = g0 + lane < groups ? a.pcnt[g0 + lane] : 0u;
    uint32_t x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    const uint32_t total = __shfl(x, 63), excl = x - cnt;
    for (uint32_t b = 0; b < total; b += 64) {
      const uint32_t e = b + lane;
      // the last group whose entries start at or before e
      uint32_t j = 0;
#pragma unroll
      for (uint32_t step = 32; step; step >>= 1)
        if (__shfl(excl, (int)(j + step)) <= e) j += step;
      const uint32_t k = e - __shfl(excl, (int)j);
      if (e < total) f(a.pend[(g0 + j) * 64 + k]);
    }
  }
}

// Once every claimer has stored its value: the mins of the leaves whose insert
// matched an entry of this checkpoint (digests that several changed chunks share).
// The first match below the entry's holder sees the holder's value and relabels it
// a shifted duplicate (a claimer is labelled first by k_insert); later, lower
// matches do the same to the match they displace, whose own label k_resolve then
// decides from the final value either way.  Fixed leaves keep their label.
__device__ __forceinline__ void ph_fixmin(const Launch<DenseArgs>& L, const Where& w, uint8_t* lds) {
  (void)lds;
  const DenseArgs a = L.args[w.m].global();
  for_pending(a, w, [&](uint32_t c) {
    const uint8_t l = a.labels[c];
    if (l != L_MIN_PENDING && l != L_MIN_FIXED) return;
    const uint64_t v = ((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c);
    const uint64_t old = fetch_min_relaxed(&a.tab[a.leaf_slot[c]].val, v);
    if (v < old && (uint32_t)(old >> 32) == a.cur) {  // displaced a leaf of this checkpoint
      const uint32_t node = (uint32_t)old;
      if (node >= a.n - 1 && node < 2 * a.n - 1) {
        const uint32_t cd = node - (a.n - 1);
        const uint8_t ld = a.labels[cd];
        if (ld != L_FIXED && ld != L_MIN_FIXED) a.labels[cd] = L_SHIFT;
      }
    }
    a.labels[c] = l == L_MIN_FIXED ? L_FIXED : L_PENDING;
  });
}

__global__ __launch_bounds__(kThreads) void k_fixmin(Launch<DenseArgs> L) {
  ph_fixmin(L, where(L), nullptr);
}

// Once every insert has landed: a pending leaf is FIRST when its digest's entry is
// (cur, itself), SHIFT otherwise (include/reference_impl.hpp:52-69).
__device__ __forceinline__ void ph_resolve(const Launch<DenseArgs>& L, const Where& w, uint8_t* lds) {
  (void)lds;
  const DenseArgs a = L.args[w.m].global();
  for_pending(a, w, [&](uint32_t c) {
    if (a.labels[c] != L_PENDING) return;  // a fixed leaf right after a reset
    const uint32_t s = a.leaf_slot[c];
    uint8_t l = L_SHIFT;
    if (s <= a.mask + 1) l = (a.tab[s].val == (((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c))) ? L_FIRST : L_SHIFT;
    a.labels[c] = l;
    if (a.R0 && l == L_FIRST) a.tile_first[chunk_tile(a.g, a.R0, c)] = 1;
  });
}

__global__ __launch_bounds__(kThreads) void k_resolve(Launch<DenseArgs> L) {
  ph_resolve(L, where(L), nullptr);
}

// Standalone chunk hashing (mtc_hash_chunks).
template <int HK, int CB>
__global__ __launch_bounds__(kThreads) void k_hash_chunks(const uint8_t* data, uint64_t len, uint32_t n, uint32_t C,
                                                          Digest* out) {
  const uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const Digest d = tile_digest<HK, CB>(data, len, n, C, blockIdx.x);
  if (c < n) out[c] = d;
}

// ---------------------------------------------------------------------------
// Forest sweeps.  A stage covers parent depths [d_out, d_in-1] (at most kFuse
// levels).  One wavefront owns a tile: positions [t*T, (t+1)*T) at depth d_in
// (T = 2^(d_in-d_out) <= 256) and the subtree above them.  Lane l holds positions
// 4l..4l+3; levels 1 and 2 are combined in registers, levels 3.. by
// __shfl_down, so no block barrier is needed.  Stage 0 reads the leaves (bottom
// level and, when n is not a power of two, the depth D-1 leaves); later stages
// read the previous stage's frontier (one entry per tile root).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t bottom_chunk(const Geo& g, uint64_t p) { return (1ull << g.D) - g.n + p; }

__device__ __forceinline__ bool is_region(uint8_t l) { return l == L_FIRST || l == L_SHIFT; }

__device__ __forceinline__ uint64_t heap_id(uint32_t depth, uint64_t pos) { return ((1ull << depth) - 1) + pos; }

// Wave-wide exclusive prefix and total of small per-lane counts (0..7).
__device__ __forceinline__ uint32_t wave_rank(uint32_t cnt, uint32_t* total) {
  const uint32_t lane = lane_id();
  const uint64_t lt = (1ull << lane) - 1;
  uint32_t r = 0, t = 0;
#pragma unroll
  for (int b = 0; b < 3; b++) {
    const uint64_t m = __ballot((cnt >> b) & 1);
    r += (uint32_t)__popcll(m & lt) << b;
    t += (uint32_t)__popcll(m) << b;
  }
  *total = t;
  return r;
}

// Wave-wide total of small per-lane counts (0..7): three ballots, no per-lane prefix.
__device__ __forceinline__ uint32_t wave_total3(uint32_t cnt) {
  return (uint32_t)__popcll(__ballot(cnt & 1)) + 2u * (uint32_t)__popcll(__ballot(cnt & 2)) +
         4u * (uint32_t)__popcll(__ballot(cnt & 4));
}

__device__ __forceinline__ Digest shfl_down_dg(const Digest& d, int s) {
  return Digest{__shfl_down(d.lo, s), __shfl_down(d.hi, s)};
}

// Sweep A: first-occurrence forest (include/reference_impl.hpp:74-82).  Resolves
// pending leaves, hashes every node whose children are both first occurrences and
// inserts it (insert-if-absent).
template <int HK>
struct SweepA {
  const SweepArgs& a;
  Digest* rec_dg;      // this wavefront's LDS record of nodes to insert (256 entries)
  uint32_t* rec_node;
  uint32_t count = 0;  // entries recorded so far (wave-uniform)
  Digest d0, d1;       // this lane's records of the current level
  uint32_t n0 = 0, n1 = 0, nrec = 0;

  // Both children first occurrences: hash the pair and record the node; the
  // wavefront inserts all its records together once the tile is done (they only
  // have to land before sweep B looks anything up).
  __device__ void combine(uint32_t depth, uint64_t pos, uint8_t la, const Digest& da, uint8_t lb, const Digest& db,
                          uint8_t& lo, Digest& dout) {
    lo = L_NONE;
    if (la == L_FIRST && lb == L_FIRST) {
      dout = hash_pair<HK>(da, db);
      const uint32_t node = (uint32_t)heap_id(depth, pos);
      if (nrec == 0) {
        d0 = dout;
        n0 = node;
      } else {
        d1 = dout;
        n1 = node;
      }
      nrec++;
      lo = L_FIRST;
      if (a.dbg_tree) a.dbg_tree[node] = dout;
    }
  }

  __device__ void flush() {
    uint32_t tot;
    const uint32_t r = wave_rank(nrec, &tot);
    if (nrec > 0) {
      rec_dg[count + r] = d0;
      rec_node[count + r] = n0;
    }
    if (nrec > 1) {
      rec_dg[count + r + 1] = d1;
      rec_node[count + r + 1] = n1;
    }
    count += tot;
    nrec = 0;
  }
};

// One tile of sweep A (the whole wavefront).
template <int HK>
__device__ __forceinline__ void sweep_a_tile(const SweepArgs& a, uint64_t tile, uint32_t lane, Digest* rec_dg,
                                             uint32_t* rec_node) {
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
  uint8_t l0[4];
  bool any = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t lp = 4 * lane + i;
    l0[i] = L_NONE;
    if (lp >= T) continue;
    const uint64_t p = base + lp;
    if (a.stage0) {
      if (p >= a.g.m) continue;
      l0[i] = a.labels[bottom_chunk(a.g, p)];
    } else {
      l0[i] = a.in_lab[p];
    }
    any |= (l0[i] == L_FIRST);
  }
  // Digests of positions 4 lane + i, loaded where a pair of first occurrences is
  // hashed (holding all four per lane cost wavefront slots in registers).
  auto dg_at = [&](int i) -> Digest {
    const uint64_t p = base + 4 * lane + i;
    return a.stage0 ? a.leaf_dg[bottom_chunk(a.g, p)] : a.in_dg[p];
  };
  // Depth D-1 leaves among this tile's level-1 positions (stage 0 only).
  uint8_t lf[2] = {L_NONE, L_NONE};
  Digest gf[2];
  if (a.stage0) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t j = 2 * lane + k;
      const uint64_t q = (base >> 1) + j;
      if (j >= T / 2 || q < a.g.half) continue;
      const uint64_t c = q - a.g.half;
      lf[k] = a.labels[c];
      if (lf[k] == L_FIRST) gf[k] = a.leaf_dg[c];
      any |= (lf[k] == L_FIRST);
    }
  }
  if (!__ballot(any)) {
    if (lane == 0) a.out_lab[tile] = L_NONE;
    return;
  }
  SweepA<HK> s{a, rec_dg, rec_node};
  // level 1 (two parents per lane)
  uint8_t l1[2] = {L_NONE, L_NONE};
  Digest g1[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const uint32_t j = 2 * lane + k;
    if (j >= T / 2) continue;
    const uint64_t q = (base >> 1) + j;
    if (a.stage0 && q >= a.g.half) {
      l1[k] = lf[k];
      g1[k] = gf[k];
    } else {
      const bool both = l0[2 * k] == L_FIRST && l0[2 * k + 1] == L_FIRST;
      const Digest da = both ? dg_at(2 * k) : Digest{0, 0};
      const Digest db = both ? dg_at(2 * k + 1) : Digest{0, 0};
      s.combine(a.d_in - 1, q, l0[2 * k], da, l0[2 * k + 1], db, l1[k], g1[k]);
    }
  }
  s.flush();
  uint8_t lc = l1[0];
  Digest gc = g1[0];
  if (R >= 2) {  // level 2 (one parent per lane)
    lc = L_NONE;
    if (lane < T / 4) s.combine(a.d_in - 2, (base >> 2) + lane, l1[0], g1[0], l1[1], g1[1], lc, gc);
    s.flush();
  }
  for (uint32_t r = 3; r <= R; r++) {  // level r: lanes that are multiples of 2^(r-2)
    const int sh = 1 << (r - 3);
    const uint8_t lp = (uint8_t)__shfl_down((int)lc, sh);
    const Digest gp = shfl_down_dg(gc, sh);
    const uint32_t j = lane >> (r - 2);
    if ((lane & ((2u << (r - 3)) - 1)) == 0 && j < (T >> r)) {
      uint8_t lo;
      Digest go;
      s.combine(a.d_in - r, (base >> r) + j, lc, gc, lp, gp, lo, go);
      lc = lo;
      gc = go;
    }
    s.flush();
  }
  if (lane == 0) {
    a.out_lab[tile] = lc;
    if (lc == L_FIRST) a.out_dg[tile] = gc;
  }
  // Insert-if-absent of this tile's first-occurrence nodes (reference_impl.hpp:80).
  // LDS accesses of one wavefront execute in order, so the records are visible.
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  uint32_t claims = 0;
  for (uint32_t i = lane; i < s.count; i += 64) {
    bool cl = false;
    table_insert(a.tab, a.mask, rec_dg[i], ((uint64_t)a.cur << 32) | rec_node[i], false, &a.ctl->err, &cl,
                 true);
    claims += cl;
  }
  const uint32_t wc = wave_sum(claims);
  if (lane == 0 && wc) atomicAdd(stripe(a.ctr, CTR_CLAIMS), (unsigned long long)wc);
}

// Stage 0: a wavefront owns tiles_per_wave consecutive tiles, reads their
// first-occurrence flags with one load per lane and walks only the flagged ones.
template <int HK, int WPB>
__device__ __forceinline__ void sweep_a_stage(const SweepArgs& a, uint32_t blk, uint8_t* lds) {
  auto& rec_dg = *reinterpret_cast<Digest(*)[WPB][256]>(lds);
  auto& rec_node = *reinterpret_cast<uint32_t(*)[WPB][256]>(lds + sizeof(Digest) * WPB * 256);
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  const uint64_t ntiles = 1ull << a.d_out;
  const uint32_t tpw = a.stage0 ? a.tiles_per_wave : 1u;
  const uint64_t t0 = ((uint64_t)blk * WPB + w) * tpw;
  if (t0 >= ntiles) return;  // whole wavefront
  const uint64_t t1 = min(t0 + tpw, ntiles);
  if (!a.stage0) {
    for (uint64_t t = t0; t < t1; t++) sweep_a_tile<HK>(a, t, lane, rec_dg[w], rec_node[w]);
    return;
  }
  const uint64_t mine = t0 + lane;
  const bool has = mine < t1 && a.tile_first[mine];
  if (mine < t1 && !has) a.out_lab[mine] = L_NONE;  // no first-occurrence leaf below this tile
  uint64_t todo = __ballot(has);
  while (todo) {
    const int i = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    sweep_a_tile<HK>(a, t0 + i, lane, rec_dg[w], rec_node[w]);
  }
}

template <int HK, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_sweep_a(Launch<SweepArgs> L) {
  __shared__ __align__(16) uint8_t lds[(sizeof(Digest) + sizeof(uint32_t)) * WPB * 256];
  const Where wh = where(L);
  sweep_a_stage<HK, WPB>(L.args[wh.m].global(), wh.blk, lds);
}

// A member's last stages in one block: each stage's frontier and table inserts are
// visible to the next stage's wavefronts of the same block after the barrier.  Each
// launch costs about 5 us of dispatch on its own, more than these stages' work.
template <int HK>
__global__ __launch_bounds__(64 * kTailWaves) void k_sweep_a_tail(Launch<SweepTail> L) {
  __shared__ __align__(16) uint8_t lds[(sizeof(Digest) + sizeof(uint32_t)) * kTailWaves * 256];
  const Where wh = where(L);
  const uint32_t nst = L.args[wh.m].nst;
  for (uint32_t i = 0; i < nst; i++) {
    sweep_a_stage<HK, kTailWaves>(L.args[wh.m].st[i].global(), 0, lds);
    __syncthreads();
  }
}

// Sweep B: shifted-duplicate forest and region roots (include/reference_impl.hpp
// :83-106).  Read-only on the table.  Roots are staged per (depth, group) in
// ascending node id, a group being the consecutive tiles one wavefront walks (one
// tile after the first stage); a shifted root carries the table slot of its digest.
// Where a level's `tot` roots (`ftot` of them first occurrences) go in the staging
// array: after the roots the group has staged at depth c so far.  The per-level
// tallies live in the wavefront's LDS (volatile, so every lane reads lane 0's last
// update; one wavefront's LDS accesses execute in order), and the group's root
// counts are written from them once its tiles are done.
__device__ __forceinline__ uint64_t group_stage(const SweepArgs& a, uint32_t c, uint64_t gtile,
                                                volatile uint32_t* run, uint32_t tot, uint32_t ftot) {
  const uint32_t i = c - a.d_out - 1;
  const uint32_t r0 = __builtin_amdgcn_readfirstlane(run[2 * i]);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  if (lane_id() == 0) {
    run[2 * i] = r0 + tot;
    run[2 * i + 1] = run[2 * i + 1] + ftot;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  return heap_id(c, gtile << (c - a.d_out)) + r0;
}

template <int HK>
struct SweepB {
  const SweepArgs& a;
  uint64_t* ent;  // this wavefront's LDS: up to four roots per lane and level, entry k of lane l at [64 k + l]
  volatile uint32_t* run;  // this wavefront's LDS: roots staged so far per level (all, first), for the group
  uint64_t gtile;          // the group's first tile
  uint32_t ne = 0, nf = 0;

  __device__ void emit(uint8_t l, uint32_t slot, uint64_t node) {
    if (!is_region(l)) return;
    ent[64 * ne + lane_id()] = ((l == L_FIRST ? kFirstTag : (uint64_t)slot) << 32) | (uint32_t)node;
    ne++;
    nf += (l == L_FIRST);
  }

  __device__ void combine(uint32_t depth, uint64_t pos, uint8_t la, const Digest& da, uint32_t sa, uint8_t lb,
                          const Digest& db, uint32_t sb, uint8_t& lo, Digest& dout, uint32_t& sout) {
    const uint64_t node = heap_id(depth, pos), cl = 2 * node + 1;
    lo = L_DONE;
    if (la != lb) {
      emit(la, sa, cl);
      emit(lb, sb, cl + 1);
    } else if (la == L_FIXED) {
      lo = L_FIXED;
    } else if (la == L_SHIFT) {
      const Digest d = hash_pair<HK>(da, db);
      if (a.dbg_tree) a.dbg_tree[node] = d;
      const uint32_t s = table_find(a.tab, a.mask, d);
      if (s != ~0u) {
        lo = L_SHIFT;
        dout = d;
        sout = s;
      } else {
        emit(la, sa, cl);
        emit(lb, sb, cl + 1);
      }
    } else if (la == L_FIRST) {
      lo = L_FIRST;
    }
    if (a.dbg_labels) a.dbg_labels[node] = lo;
  }

  // Stage this level's roots (children at depth c) after the group's earlier ones.
  __device__ void flush(uint32_t c) {
    if (!__ballot(ne != 0)) return;  // no root at this level (nf is zero too)
    uint32_t tot;
    const uint32_t rank = wave_rank(ne, &tot);
    const uint32_t ftot = wave_total3(nf);
    const uint32_t lane = lane_id();
    uint64_t* dst = a.staging + group_stage(a, c, gtile, run, tot, ftot) + rank;
    for (uint32_t k = 0; k < ne; k++) dst[k] = ent[64 * k + lane];
    ne = nf = 0;
  }
};

// Every leaf below the tile is a fixed duplicate: the tile is fixed and emits no roots.
__device__ void fixed_tile(const SweepArgs& a, uint64_t tile, uint32_t R, uint32_t T, uint64_t base, uint32_t lane) {
  if (lane == 0) a.out_lab[tile] = L_FIXED;  // root counts stay zero
  if (a.dbg_labels)
    for (uint32_t r = 1; r <= R; r++)
      for (uint32_t j = lane; j < (T >> r); j += 64) {
        const uint64_t q = (base >> r) + j;
        if (a.stage0 && r == 1 && q >= a.g.half) continue;
        a.dbg_labels[heap_id(a.d_in - r, q)] = L_FIXED;
      }
}

// Labels of a tile's positions, four per lane (packed in one word, L_NONE where
// the position does not exist): one 32-bit load when whole and aligned.
struct Lab4 {
  uint32_t w;     // label of position 4 lane + i in bits 8i..8i+7
  uint32_t nval;  // valid positions of this lane
  uint64_t c0;    // index of the first one in labels (stage 0) or in_lab
  __device__ uint8_t l(int i) const { return (uint8_t)(w >> (8 * i)); }
};

__device__ __forceinline__ Lab4 load_lab4(const SweepArgs& a, uint64_t base, uint32_t T, uint32_t lane) {
  Lab4 r;
  const uint32_t lp0 = 4 * lane;
  r.c0 = a.stage0 ? bottom_chunk(a.g, base + lp0) : base + lp0;
  const uint8_t* lsrc = a.stage0 ? a.labels + r.c0 : a.in_lab + r.c0;
  r.nval = lp0 < T ? min(4u, T - lp0) : 0u;
  if (a.stage0) r.nval = (base + lp0 >= a.g.m) ? 0u : (uint32_t)min<uint64_t>(r.nval, a.g.m - (base + lp0));
  if (r.nval == 4 && ((uintptr_t)lsrc & 3) == 0) {
    r.w = *reinterpret_cast<const uint32_t*>(lsrc);
  } else {
    r.w = 0x01010101u * L_NONE;  // L_NONE in every byte
#pragma unroll
    for (int i = 0; i < 4; i++)
      if ((uint32_t)i < r.nval) r.w = (r.w & ~(0xFFu << (8 * i))) | ((uint32_t)lsrc[i] << (8 * i));
  }
  return r;
}

// One tile of sweep B (the whole wavefront), its labels already loaded.
template <int HK>
__device__ __forceinline__ void sweep_b_tile(const SweepArgs& a, uint64_t tile, const Lab4& lb, uint32_t lane,
                                             uint64_t* ent, volatile uint32_t* run, uint64_t gtile) {
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
  const uint32_t lp0 = 4 * lane;
  const uint64_t c0 = lb.c0;
  const uint32_t nval = lb.nval;
  uint8_t l0[4];
#pragma unroll
  for (int i = 0; i < 4; i++) l0[i] = lb.l(i);
  bool busy = false;
#pragma unroll
  for (int i = 0; i < 4; i++) busy |= ((uint32_t)i < nval && l0[i] != L_FIXED);
  uint8_t lf[2] = {L_NONE, L_NONE};  // depth D-1 leaves among the level-1 positions
  if (a.stage0) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t j = 2 * lane + k;
      const uint64_t q = (base >> 1) + j;
      if (j >= T / 2 || q < a.g.half) continue;
      lf[k] = a.labels[q - a.g.half];
      busy |= (lf[k] != L_FIXED);
    }
  }
  if (!__ballot(busy)) {
    fixed_tile(a, tile, R, T, base, lane);
    return;
  }
  // Sparse tile: no two siblings carry the same region label, so every region leaf
  // is a root and every ancestor of a non-fixed leaf is done (the levels below
  // would derive exactly this); emit the roots without walking the levels.
  if (a.stage0 && !a.dbg_labels && base + T <= a.g.m) {
    bool pair = false;
#pragma unroll
    for (int k = 0; k < 2; k++) pair |= (l0[2 * k] == l0[2 * k + 1] && is_region(l0[2 * k]));
    if (!__ballot(pair)) {
      // The region leaves are the tile's roots, ascending (what SweepB::emit and
      // flush would stage, written directly).
      uint32_t rm = 0, fm = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        rm |= (uint32_t)is_region(l0[i]) << i;
        fm |= (uint32_t)(l0[i] == L_FIRST) << i;
      }
      if (__ballot(rm != 0)) {
        uint32_t tot;
        const uint32_t rank = wave_rank((uint32_t)__popc(rm), &tot);
        const uint32_t ftot = wave_total3((uint32_t)__popc(fm));
        uint64_t* dst = a.staging + group_stage(a, a.d_in, gtile, run, tot, ftot) + rank;
        const uint32_t node0 = (uint32_t)heap_id(a.d_in, base + lp0);
#pragma unroll
        for (int i = 0; i < 4; i++)
          if ((rm >> i) & 1) {
            const uint64_t tag = ((fm >> i) & 1) ? kFirstTag : kLeafTag;
            dst[__popc(rm & ((1u << i) - 1))] = (tag << 32) | (node0 + (uint32_t)i);
          }
      }
      if (lane == 0) {
        a.out_lab[tile] = L_DONE;
        if (a.d_out == 0) {
          a.cnt_f[a.seg_base[0]] = 0;
          a.cnt_s[a.seg_base[0]] = 0;
        }
      }
      return;
    }
  }
  // Level 1.  The digests and table slots of shifted positions are loaded only
  // where a pair needs them (a shifted pair is hashed; a shifted root carries its
  // slot; a leaf's slot is read when its root entry is consumed): holding all four
  // per lane cost the kernel half its wavefront slots in registers.
  SweepB<HK> s{a, ent, run, gtile};
  uint8_t l1[2] = {L_NONE, L_NONE};
  Digest g1[2];
  uint32_t s1[2] = {0, 0};
  auto slot_at = [&](int i) -> uint32_t { return a.stage0 ? (uint32_t)kLeafTag : a.in_slot[c0 + i]; };
  auto dg_at = [&](int i) -> Digest { return a.stage0 ? a.leaf_dg[c0 + i] : a.in_dg[c0 + i]; };
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const uint32_t j = 2 * lane + k;
    if (j >= T / 2) continue;
    const uint64_t q = (base >> 1) + j;
    if (a.stage0 && q >= a.g.half) {
      l1[k] = lf[k];
      if (lf[k] == L_SHIFT) {
        g1[k] = a.leaf_dg[q - a.g.half];
        s1[k] = (uint32_t)kLeafTag;
      }
      continue;
    }
    const uint8_t la = l0[2 * k], lb2 = l0[2 * k + 1];
    const bool both_shift = la == L_SHIFT && lb2 == L_SHIFT;
    const Digest da = both_shift ? dg_at(2 * k) : Digest{0, 0};
    const Digest db = both_shift ? dg_at(2 * k + 1) : Digest{0, 0};
    const uint32_t sa = (la == L_SHIFT) ? slot_at(2 * k) : 0u;
    const uint32_t sb = (lb2 == L_SHIFT) ? slot_at(2 * k + 1) : 0u;
    s.combine(a.d_in - 1, q, la, da, sa, lb2, db, sb, l1[k], g1[k], s1[k]);
  }
  s.flush(a.d_in);
  uint8_t lc = l1[0];
  Digest gc = g1[0];
  uint32_t sc = s1[0];
  if (R >= 2) {
    lc = L_NONE;
    if (lane < T / 4)
      s.combine(a.d_in - 2, (base >> 2) + lane, l1[0], g1[0], s1[0], l1[1], g1[1], s1[1], lc, gc, sc);
    s.flush(a.d_in - 1);
  }
  for (uint32_t r = 3; r <= R; r++) {
    const int sh = 1 << (r - 3);
    const uint8_t lp = (uint8_t)__shfl_down((int)lc, sh);
    // Digests and slots matter only for shifted pairs (a fresh run has none).
    Digest gp{0, 0};
    uint32_t sp = 0;
    if (__ballot(lc == L_SHIFT)) {
      gp = shfl_down_dg(gc, sh);
      sp = __shfl_down(sc, sh);
    }
    const uint32_t j = lane >> (r - 2);
    if ((lane & ((2u << (r - 3)) - 1)) == 0 && j < (T >> r)) {
      uint8_t lo;
      Digest go;
      uint32_t so = 0;
      s.combine(a.d_in - r, (base >> r) + j, lc, gc, sc, lp, gp, sp, lo, go, so);
      lc = lo;
      gc = go;
      sc = so;
    }
    s.flush(a.d_in - r + 1);
  }
  if (lane == 0) {
    a.out_lab[tile] = lc;
    if (lc == L_SHIFT) {
      a.out_dg[tile] = gc;
      a.out_slot[tile] = sc;
    }
    if (a.d_out == 0) {
      // Node 0: a first-occurrence root (:105-106); a shifted root too, which the
      // reference omits and so cannot restart (documented departure), unless the
      // reference's bytes are asked for.
      const bool root = a.compat ? lc == L_FIRST : is_region(lc);
      if (root) a.staging[0] = ((lc == L_FIRST ? kFirstTag : (uint64_t)sc) << 32) | 0u;
      a.cnt_f[a.seg_base[0]] = (root && lc == L_FIRST) ? 1u : 0u;
      a.cnt_s[a.seg_base[0]] = (root && lc == L_SHIFT) ? 1u : 0u;
    }
  }
}

template <int HK>
__device__ __forceinline__ void ph_sweep_b_tiles(const SweepArgs& a, uint64_t t0, uint32_t tpw, uint64_t ntiles,
                                                 uint32_t R, uint32_t T, uint32_t lane, uint64_t* ent,
                                                 volatile uint32_t* run) {
  if (tpw == 1) {
    sweep_b_tile<HK>(a, t0, load_lab4(a, t0 << R, T, lane), lane, ent, run, t0);
    return;
  }
  const uint64_t t1 = min(t0 + tpw, ntiles);
  const uint64_t mine = t0 + lane;
  const bool has = mine < t1 && a.tile_busy[mine];
  if (mine < t1 && !has) a.out_lab[mine] = L_FIXED;
  uint64_t todo = __ballot(has);
  if (!todo) return;
  // The next flagged tile's labels are loaded before this one is walked.
  int i = __ffsll((long long)todo) - 1;
  Lab4 lb = load_lab4(a, (t0 + i) << R, T, lane);
  while (todo) {
    const uint64_t t = t0 + i;
    const Lab4 cur = lb;
    todo &= todo - 1;
    if (todo) {
      i = __ffsll((long long)todo) - 1;
      lb = load_lab4(a, (t0 + i) << R, T, lane);
    }
    sweep_b_tile<HK>(a, t, cur, lane, ent, run, t0);
  }
}

// One wavefront per tile, or at stage 0 tiles_per_wave consecutive tiles (one
// root segment per depth): their busy flags are read with one load per lane and
// only the flagged tiles are walked; every leaf below an unflagged tile is a fixed
// duplicate, so the tile is fixed and adds no roots.
template <int HK, int WPB>
__device__ __forceinline__ void sweep_b_stage(const SweepArgs& a, uint32_t blk, uint64_t* ent_all,
                                              uint32_t* run_all) {
  const uint32_t w = threadIdx.x >> 6;
  uint64_t* ent = ent_all + 256 * w;  // this wavefront's root entries
  volatile uint32_t* run = run_all + 2 * kFuse * w;
  const uint32_t lane = lane_id();
  const uint64_t ntiles = 1ull << a.d_out;
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint32_t tpw = a.stage0 ? a.tiles_per_wave : 1u;
  const uint64_t grp = (uint64_t)blk * WPB + w, t0 = grp * tpw;
  if (t0 >= ntiles) return;  // whole wavefront
  if (lane < 2 * R) run[lane] = 0;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  ph_sweep_b_tiles<HK>(a, t0, tpw, ntiles, R, T, lane, ent, run);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // The group's root counts per depth (zero-filled before the sweep).
  if (lane < R) {
    const uint32_t tot = run[2 * lane], ftot = run[2 * lane + 1];
    if (tot) {
      const uint32_t seg = a.seg_base[a.d_out + 1 + lane] + (uint32_t)grp;
      a.cnt_f[seg] = ftot;
      a.cnt_s[seg] = tot - ftot;
    }
  }
  if (a.d_out == 0 && lane == 0) a.ctl->t_forest = wall_clock64();  // the member's last sweep stage
}

// Five wavefronts per SIMD (96 VGPRs, two spill): sweep B took 150 us per
// checkpoint against 169 us at six (80 VGPRs, eighteen spill) at 8 GiB, 128-byte
// chunks, 2% change.
template <int HK, int WPB>
__global__ __launch_bounds__(64 * WPB, MTC_SWEEP_B_WAVES) void k_sweep_b(Launch<SweepArgs> L) {
  __shared__ __align__(16) uint64_t ent[WPB * 256];
  __shared__ uint32_t run[WPB * 2 * kFuse];
  const Where wh = where(L);
  sweep_b_stage<HK, WPB>(L.args[wh.m].global(), wh.blk, ent, run);
}

// A single chunk: the leaf is the root.
__device__ __forceinline__ void ph_single_leaf(const Launch<SweepArgs>& L, const Where& w) {
  if (threadIdx.x != 0) return;
  const SweepArgs a = L.args[w.m].global();
  const uint8_t l = a.labels[0];
  const bool root = a.compat ? l == L_FIRST : is_region(l);
  if (root) a.staging[0] = (l == L_FIRST ? (kFirstTag << 32) : ((uint64_t)a.leaf_slot[0] << 32)) | 0u;
  a.cnt_f[0] = (root && l == L_FIRST) ? 1u : 0u;
  a.cnt_s[0] = (root && l == L_SHIFT) ? 1u : 0u;
  a.ctl->t_forest = wall_clock64();
}

__global__ void k_single_leaf(Launch<SweepArgs> L) { ph_single_leaf(L, where(L)); }

// ---------------------------------------------------------------------------
// Root compaction, in segment order (depth, then group).  A wavefront takes 64
// segments, one per lane, and copies their roots 64 at a time: lane e of a round
// finds its segment by bisection over the exclusive prefix of the root counts, and
// its rank among the segment's first-occurrence (or shifted) roots from a ballot.
// ---------------------------------------------------------------------------

// Lanes lo..hi-1 (0 <= lo, hi <= 64).
__device__ __forceinline__ uint64_t lanes_between(uint32_t lo, uint32_t hi) {
  if (hi <= lo) return 0;
  const uint64_t upto = hi >= 64 ? ~0ull : ((1ull << hi) - 1);
  return upto & ~((1ull << lo) - 1);
}

// The per-depth segment bases and geometry (at most 2 x 33 + 1 words) are staged in
// LDS: each lane finds its depth with a walk over them.
__device__ __forceinline__ void ph_scatter_roots(const Launch<RootArgs>& L, const Where& w, uint8_t* lds) {
  uint32_t* s_base = reinterpret_cast<uint32_t*>(lds);
  const RootArgs a = L.args[w.m].global();
  const uint32_t* s_geo = s_base + a.ndepth + 1;
  if (threadIdx.x <= 2 * a.ndepth) s_base[threadIdx.x] = a.seg_base[threadIdx.x];
  __syncthreads();
  const uint32_t lane = lane_id();
  const uint32_t seg = w.blk * kThreads + threadIdx.x;
  uint32_t nf = 0, ns = 0;
  if (seg < a.nseg) {
    nf = a.cnt_f[seg];
    ns = a.cnt_s[seg];
  }
  const uint32_t nr = nf + ns;
  if (!__ballot(nr != 0)) return;
  uint64_t sb = 0;
  uint32_t of = 0, os = 0;
  if (nr) {
    uint32_t c = 0;
    while (c + 1 < a.ndepth && s_base[c + 1] <= seg) c++;
    const uint32_t geo = s_geo[c];
    sb = ((1ull << c) - 1) + (((uint64_t)(seg - s_base[c]) * (geo & 0xFFFFu)) << (geo >> 16));
    of = a.off_f[seg];
    os = a.off_s[seg];
  }
  uint32_t x = nr;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  const uint32_t total = __shfl(x, 63), excl = x - nr;
  const uint32_t sb_lo = (uint32_t)sb, sb_hi = (uint32_t)(sb >> 32);
  for (uint32_t b = 0; b < total; b += 64) {
    const uint32_t e = b + lane;
    uint32_t j = 0;  // the last segment whose roots start at or before e
#pragma unroll
    for (uint32_t step = 32; step; step >>= 1)
      if (__shfl(excl, (int)(j + step)) <= e) j += step;
    const uint32_t ej = __shfl(excl, (int)j);
    const uint64_t sbj = ((uint64_t)(uint32_t)__shfl(sb_hi, (int)j) << 32) | (uint32_t)__shfl(sb_lo, (int)j);
    const uint32_t ofj = __shfl(of, (int)j), osj = __shfl(os, (int)j);
    const bool valid = e < total;
    const uint64_t ent = valid ? a.staging[sbj + (e - ej)] : 0ull;
    const bool isf = valid && (ent >> 32) == kFirstTag;
    const uint64_t bf = __ballot(isf), bv = __ballot(valid);
    const uint64_t before = lanes_between(ej > b ? ej - b : 0u, lane);  // this segment's lanes below
    if (valid) {
      if (isf) a.first_roots[ofj + __popcll(bf & before)] = (uint32_t)ent;
      else a.shift_tmp[osj + __popcll(bv & ~bf & before)] = ent;
    }
    // Each segment's lane advances its offsets past the roots this round placed.
    const uint32_t lo = excl > b ? min(excl - b, 64u) : 0u;
    const uint32_t hi = excl + nr > b ? min(excl + nr - b, 64u) : 0u;
    const uint64_t mine = lanes_between(lo, hi);
    of += __popcll(bf & mine);
    os += __popcll(bv & ~bf & mine);
  }
}

__global__ __launch_bounds__(kThreads) void k_scatter_roots(Launch<RootArgs> L) {
  __shared__ uint32_t s_base[128];
  ph_scatter_roots(L, where(L), reinterpret_cast<uint8_t*>(s_base));
}

// ---------------------------------------------------------------------------
// Shifted entries: source checkpoint and previous node from the table
// (src/tree_approach.cpp:574-580,670-696), then a stable sort by source.
// ---------------------------------------------------------------------------
constexpr uint32_t kLdsBins = 4096;  // source-checkpoint bins counted in LDS

// Block g of G's contiguous range of the S entries (whole 256-entry rows).
__device__ __forceinline__ void sort_range(uint32_t S, uint32_t g, uint32_t G, uint32_t& lo, uint32_t& hi) {
  const uint32_t per = ((S + G - 1) / G + kThreads - 1) / kThreads * kThreads;
  lo = min(S, g * per);
  hi = min(S, lo + per);
}

__device__ __forceinline__ void ph_shift_src(const Launch<ShiftArgs>& L, const Where& w, uint8_t* lds) {
  uint32_t* h = reinterpret_cast<uint32_t*>(lds);
  const ShiftArgs a = L.args[w.m].global();
  const bool in_lds = a.nbins <= kLdsBins;
  if (in_lds)
    for (uint32_t b = threadIdx.x; b < a.nbins; b += kThreads) h[b] = 0;
  __syncthreads();
  const uint32_t S = *a.S;
  uint32_t lo = 0, hi = S, i0 = w.blk * kThreads, stride = w.nblk * kThreads;
  if (a.hist0) {  // the first sort pass's block ranges
    sort_range(S, w.blk, w.nblk, lo, hi);
    i0 = lo;
    stride = kThreads;
  }
  for (uint32_t i = i0 + threadIdx.x; i < hi; i += stride) {
    const uint64_t e = a.shift_tmp[i];
    const uint32_t node = (uint32_t)e;
    uint32_t slot = (uint32_t)(e >> 32);
    if (slot == (uint32_t)kLeafTag) slot = (node >= a.n - 1 && node < 2 * a.n - 1) ? a.leaf_slot[node - (a.n - 1)] : ~0u;
    uint32_t k = 0, prev = 0;
    if (slot > a.mask + 1) {
      atomicOr(&a.ctl->err, ERR_CORRUPT);
    } else {
      const uint64_t v = a.tab[slot].val;
      const uint32_t src = (uint32_t)(v >> 32);
      prev = (uint32_t)v;
      k = src - a.ref;
      if (src < a.ref || src > a.cur) {
        atomicOr(&a.ctl->err, ERR_CORRUPT);
        k = 0;
      }
    }
    a.key[i] = k;
    a.val[i] = ((uint64_t)prev << 32) | node;
    if (in_lds) atomicAdd(&h[k], 1u);
    else atomicAdd(&a.src_hist[k], 1u);
  }
  __syncthreads();
  if (in_lds)
    for (uint32_t b = threadIdx.x; b < a.nbins; b += kThreads)
      if (h[b]) atomicAdd(&a.src_hist[b], h[b]);
  if (a.hist0)  // this block's count of each first-pass digit (key mod nb0)
    for (uint32_t d = threadIdx.x; d < a.nb0; d += kThreads) {
      uint32_t c = 0;
      for (uint32_t b = d; b < a.nbins; b += a.nb0) c += h[b];
      a.hist0[d * w.nblk + w.blk] = c;
    }
}

__global__ __launch_bounds__(kThreads) void k_shift_src(Launch<ShiftArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * kLdsBins];
  ph_shift_src(L, where(L), lds);
}

// Stable counting sort by one 8-bit digit with a fixed grid: block g of G owns a
// contiguous range of the S entries (S read on the device), counts its digits,
// and after a scan of the digit-major (digit, block) counts moves its entries in
// order.


__device__ __forceinline__ void ph_sort_count(const Launch<SortArgs>& L, const Where& w, uint8_t* lds) {
  uint32_t* h = reinterpret_cast<uint32_t*>(lds);
  const SortArgs a = L.args[w.m].global();
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t lo, hi;
  sort_range(*a.S, w.blk, w.nblk, lo, hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kThreads) atomicAdd(&h[(a.key_in[i] >> a.shift) & (a.nb - 1)], 1u);
  __syncthreads();
  if (threadIdx.x < a.nb) a.hist[threadIdx.x * w.nblk + w.blk] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_sort_count(Launch<SortArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * 256];
  ph_sort_count(L, where(L), lds);
}

__device__ __forceinline__ void ph_sort_move(const Launch<SortArgs>& L, const Where& wh, uint8_t* lds) {
  auto& cnt = *reinterpret_cast<uint32_t(*)[kThreads / 64][256]>(lds);
  uint32_t* run = reinterpret_cast<uint32_t*>(lds + sizeof(uint32_t) * (kThreads / 64) * 256);
  const SortArgs a = L.args[wh.m].global();
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t lt = (1ull << lane) - 1;
  run[tid] = 0;
  uint32_t lo, hi;
  sort_range(*a.S, wh.blk, wh.nblk, lo, hi);
  for (uint32_t r0 = lo; r0 < hi; r0 += kThreads) {
    for (uint32_t j = tid; j < (kThreads / 64) * 256; j += kThreads) (&cnt[0][0])[j] = 0;
    __syncthreads();
    const uint32_t i = r0 + tid;
    const bool valid = i < hi;
    const uint32_t k = valid ? a.key_in[i] : 0u;
    const uint32_t dg = (k >> a.shift) & (a.nb - 1);
    // rank among earlier lanes of this wave with the same digit
    bool done = !valid;
    uint32_t rank = 0;
    while (true) {
      const uint64_t pending = __ballot(!done);
      if (pending == 0) break;
      const int leader = __ffsll((long long)pending) - 1;
      const uint32_t ld = __shfl(dg, leader);
      const bool same = !done && dg == ld;
      const uint64_t sm = __ballot(same);
      if (same) {
        rank = __popcll(sm & lt);
        done = true;
      }
      if ((int)lane == leader) cnt[w][ld] = __popcll(sm);
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = a.hist[dg * wh.nblk + wh.blk] + run[dg] + rank;
      for (uint32_t ww = 0; ww < w; ww++) pos += cnt[ww][dg];
      a.key_out[pos] = k;
      a.val_out[pos] = a.val_in[i];
    }
    __syncthreads();
    if (tid < a.nb) {
      uint32_t add = 0;
      for (uint32_t ww = 0; ww < kThreads / 64; ww++) add += cnt[ww][tid];
      run[tid] += add;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_sort_move(Launch<SortArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * ((kThreads / 64) * 256 + 256)];
  ph_sort_move(L, where(L), lds);
}

// (chkpt, count) table in ascending checkpoint id (src/tree_approach.cpp:657-665).

// The (chkpt, count) table (src/tree_approach.cpp:657-665): the source checkpoints
// with at least one shifted entry, ascending.  One whole block.
__device__ void build_ptable(const PtableArgs& a, uint32_t* s_ws) {
  uint32_t base = 0;
  for (uint32_t b0 = 0; b0 < a.nbins; b0 += kThreads) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t cnt = b < a.nbins ? a.src_hist[b] : 0u;
    uint32_t tot;
    const uint32_t r = block_scan_excl(cnt > 0, &tot, s_ws);
    if (cnt > 0) {
      a.ptable[2 * (base + r)] = a.ref + b;
      a.ptable[2 * (base + r) + 1] = cnt;
    }
    base += tot;
  }
  if (threadIdx.x == 0) a.ctl->P = base;
}

__device__ __forceinline__ void ph_region_sizes(const Launch<RegionArgs>& L, const Where& w, uint8_t* lds) {
  uint32_t* s_ws = reinterpret_cast<uint32_t*>(lds);
  const RegionArgs a = L.args[w.m].global();
  if (w.blk == 0) build_ptable(a.pt, s_ws);
  const uint32_t F = *a.F;
  for (uint32_t i = w.blk * kThreads + threadIdx.x; i < F; i += w.nblk * kThreads) {
    uint32_t st[2], ct[2];
    const int r = region_runs(a.g, a.first_roots[i], a.compat, st, ct);
    a.sizes[i] = r == 0 ? 0 : (r == 1 ? ct[0] : ct[0] + ct[1]);
  }
}

__global__ __launch_bounds__(kThreads) void k_region_sizes(Launch<RegionArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * (kThreads / 64)];
  ph_region_sizes(L, where(L), lds);
}

// ---------------------------------------------------------------------------
// Diff assembly (src/tree_approach.cpp:593,622-706): header, F roots, P table,
// S pairs, then the gathered chunks.
// ---------------------------------------------------------------------------

__device__ void diff_meta(const DiffArgs& a, const Where& w) {
  const uint32_t F = a.ctl->F, P = a.ctl->P, S = a.ctl->S;
  uint32_t* out = reinterpret_cast<uint32_t*>(a.diff);  // every metadata field is 4-byte aligned
  const uint64_t oF = 8, oP = oF + F, oS = oP + 2ull * P;
  const uint64_t tid = (uint64_t)w.blk * kThreads + threadIdx.x, stride = (uint64_t)w.nblk * kThreads;
  for (uint64_t i = tid; i < F; i += stride) out[oF + i] = a.first_roots[i];
  for (uint64_t i = tid; i < 2ull * P; i += stride) out[oP + i] = a.ptable[i];
  for (uint64_t i = tid; i < S; i += stride) {
    const uint64_t v = a.shift_sorted[i];
    out[oS + 2 * i] = (uint32_t)v;
    out[oS + 2 * i + 1] = (uint32_t)(v >> 32);
  }
  if (tid == 0) {
    const uint64_t chunks = *a.total_chunks;
    const uint64_t data_off = 32 + 4ull * F + 8ull * P + 8ull * S;
    mtc_header h;
    h.ref_id = a.ref_id;
    h.chkpt_id = a.chkpt_id;
    h.datalen = a.len;
    h.chunk_size = a.C;
    h.num_first_ocur = F;
    h.num_prior_chkpts = P;
    h.num_shift_dupl = S;
    const uint32_t* hw = reinterpret_cast<const uint32_t*>(&h);
    for (int k = 0; k < 8; k++) out[k] = hw[k];
    a.ctl->data_chunks = chunks;
    a.ctl->diff_size = data_off + chunks * a.C;
  }
}

constexpr uint32_t kAsmWindow = 512;  // region offsets staged per tile of 256 diff chunks

// The diff in one launch: metadata (every block, grid-stride), then block 0 folds
// the counters and copies the control block out (its diff size is final once
// block 0's own metadata thread has written it), then the chunks of every
// first-occurrence region are gathered, regions in root order and leaves in
// ascending chunk order; the short last chunk is zero padded (:636-650).
__device__ __forceinline__ void ph_assemble(const Launch<DiffArgs>& L, const Where& wh, uint8_t* lds) {
  uint32_t* s_src = reinterpret_cast<uint32_t*>(lds);
  unsigned long long* s_fold = reinterpret_cast<unsigned long long*>(lds + sizeof(uint32_t) * kThreads);
  uint32_t* s_off = reinterpret_cast<uint32_t*>(lds + sizeof(uint32_t) * kThreads + sizeof(unsigned long long) * CTR_KINDS);
  static_assert(kThreads == kStripes, "one block folds the striped counters");
  const DiffArgs a = L.args[wh.m].global();
  diff_meta(a, wh);
  // Nothing reads the leaf labels after sweep B: the listed chunks go back to
  // L_FIXED, so the next step's leaf kernel labels only the chunks it lists.
  if (a.reset_labels) {
    const uint64_t listed = *a.changed;
    for (uint64_t i = (uint64_t)wh.blk * kThreads + threadIdx.x; i < listed; i += (uint64_t)wh.nblk * kThreads)
      a.reset_labels[a.dense[i]] = L_FIXED;
  }
  if (wh.blk == 0) {
    __syncthreads();  // the header thread's control fields
    finish_member(a.ctr, a.changed, a.ctl, a.out, s_fold);
  }
  const uint32_t F = a.ctl->F;
  if (F == 0) return;
  const uint64_t total = *a.total_chunks;
  const uint64_t data_off = 32 + 4ull * F + 8ull * a.ctl->P + 8ull * a.ctl->S;
  uint8_t* dst0 = a.diff + data_off;
  const uint32_t C = a.C;
  const uint32_t U = a.aligned ? C / 16 : C;  // copy units per chunk
  for (uint64_t tile = wh.blk; tile * kThreads < total; tile += wh.nblk) {
    const uint64_t j0 = tile * kThreads;
    const uint64_t j = j0 + threadIdx.x;
    // The tile's regions start at tile_lo[tile]: their offsets are staged in LDS
    // (a search over every region's offset in global memory waited for about 18
    // dependent loads per chunk).
    const uint32_t lo0 = a.tile_lo[tile];
    const uint32_t win = min(F - lo0, kAsmWindow);
    for (uint32_t t = threadIdx.x; t < win; t += kThreads) s_off[t] = a.region_off[lo0 + t];
    __syncthreads();
    if (j < total) {
      uint32_t lo, hi;  // last region whose offset <= j
      if (win == kAsmWindow && j >= s_off[kAsmWindow - 1]) {  // many empty regions: the rest in global memory
        lo = lo0 + kAsmWindow - 1;
        hi = F;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (a.region_off[mid] <= j) lo = mid;
          else hi = mid;
        }
      } else {
        uint32_t l = 0, h = win;
        while (h - l > 1) {
          const uint32_t mid = (l + h) >> 1;
          if (s_off[mid] <= j) l = mid;
          else h = mid;
        }
        lo = lo0 + l;
      }
      uint32_t st[2], ct[2];
      const int r = region_runs(a.g, a.first_roots[lo], a.compat, st, ct);
      const uint32_t off = lo - lo0 < win ? s_off[lo - lo0] : a.region_off[lo];
      s_src[threadIdx.x] = run_at(r, st, ct, (uint32_t)(j - off));
    }
    __syncthreads();
    const uint64_t cnt = (total - j0 < kThreads) ? total - j0 : kThreads;
    const uint64_t units = cnt * U;
    for (uint64_t u = threadIdx.x; u < units; u += kThreads) {
      const uint32_t jj = (uint32_t)(u / U), w = (uint32_t)(u % U);
      const uint64_t chunk = s_src[jj];
      const uint64_t so = chunk * C;
      uint8_t* dst = dst0 + (j0 + jj) * C;
      if (a.aligned) {
        const uint64_t b0 = so + 16ull * w;
        if (b0 + 16 <= a.len) {
          const uint4 v = *reinterpret_cast<const uint4*>(a.data + b0);
          *reinterpret_cast<u32x4_nt_a4*>(dst + 16ull * w) = u32x4_nt_a4{v.x, v.y, v.z, v.w};
        } else {
          for (int b = 0; b < 16; b++) dst[16ull * w + b] = (b0 + b < a.len) ? a.data[b0 + b] : 0;
        }
      } else {
        dst[w] = (so + w < a.len) ? a.data[so + w] : 0;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_assemble(Launch<DiffArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(uint32_t) * kThreads + sizeof(unsigned long long) * CTR_KINDS +
                                       sizeof(uint32_t) * kAsmWindow];
  ph_assemble(L, where(L), lds);
}

// Diff of the basic approach (src/basic_approach.cpp:54-142): the header (F = the
// changed chunks, P = S = 0), their chunk indices, then their bytes, C each, the
// short last chunk zero padded.  The reference orders the indices as its atomics
// land; here they are ascending.  Block 0 also folds the counters and copies the
// control block out, as k_assemble does for tree diffs.

__device__ __forceinline__ void ph_basic_diff(const Launch<BasicDiffArgs>& L, const Where& wh, uint8_t* lds) {
  unsigned long long* s_fold = reinterpret_cast<unsigned long long*>(lds);
  const BasicDiffArgs a = L.args[wh.m].global();
  const uint32_t F = *a.total;
  uint32_t* out = reinterpret_cast<uint32_t*>(a.diff);
  const uint64_t tid = (uint64_t)wh.blk * kThreads + threadIdx.x, stride = (uint64_t)wh.nblk * kThreads;
  for (uint64_t i = tid; i < F; i += stride) out[8 + i] = a.dense[i];
  if (wh.blk == 0) {
    if (threadIdx.x == 0) {
      mtc_header h;
      h.ref_id = a.ref_id;
      h.chkpt_id = a.chkpt_id;
      h.datalen = a.len;
      h.chunk_size = a.C;
      h.num_first_ocur = F;
      h.num_prior_chkpts = 0;
      h.num_shift_dupl = 0;
      const uint32_t* hw = reinterpret_cast<const uint32_t*>(&h);
      for (int k = 0; k < 8; k++) out[k] = hw[k];
      a.ctl->F = F;
      a.ctl->P = 0;
      a.ctl->S = 0;
      a.ctl->data_chunks = F;
      a.ctl->diff_size = 32 + 4ull * F + (uint64_t)F * a.C;
    }
    __syncthreads();
    finish_member(a.ctr, a.total, a.ctl, a.out, s_fold);
  }
  uint8_t* dst0 = a.diff + 32 + 4ull * F;
  const uint32_t C = a.C;
  const uint32_t U = a.aligned ? C / 16 : C;  // copy units per chunk
  for (uint64_t u = tid; u < (uint64_t)F * U; u += stride) {
    const uint64_t j = u / U, w = u % U;
    const uint64_t so = (uint64_t)a.dense[j] * C;
    uint8_t* dst = dst0 + j * C;
    if (a.aligned) {
      const uint64_t b0 = so + 16 * w;
      if (b0 + 16 <= a.len) {
        const uint4 v = *reinterpret_cast<const uint4*>(a.data + b0);
        *reinterpret_cast<u32x4_nt_a4*>(dst + 16 * w) = u32x4_nt_a4{v.x, v.y, v.z, v.w};
      } else {
        for (int b = 0; b < 16; b++) dst[16 * w + b] = (b0 + b < a.len) ? a.data[b0 + b] : 0;
      }
    } else {
      dst[w] = (so + w < a.len) ? a.data[so + w] : 0;
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_basic_diff(Launch<BasicDiffArgs> L) {
  __shared__ __align__(16) uint8_t lds[sizeof(unsigned long long) * CTR_KINDS];
  ph_basic_diff(L, where(L), lds);
}

// Reinsert every entry of the old table into a larger one (a rehash).
__global__ __launch_bounds__(kThreads) void k_rehash(const Slot* old, uint64_t old_slots, Slot* tab, uint64_t mask,
                                                     Ctl* ctl) {
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < old_slots; i += (uint64_t)gridDim.x * kThreads) {
    const Slot s = old[i];
    const bool used = (i + 1 == old_slots) ? (s.k1 != TEMPTY && s.val != TEMPTY) : (s.k0 != TEMPTY);
    if (!used) continue;
    bool cl;
    table_insert(tab, mask, Digest{(i + 1 == old_slots) ? TEMPTY : s.k0, s.k1}, s.val, true, &ctl->err, &cl);
  }
}

// Windowed rehash into a table 2^g times larger (g = 1, 2).  An entry's new home
// keeps its old home as low bits (both are the same hash masked), so the entries
// whose old home lies in window [b*W, (b+1)*W) are exactly the ones whose new home
// lies in one of the 2^g windows [j*old_cap + b*W, ...+W).  Block b collects them
// (from its own old slots and from the cluster running past its window end),
// places them by linear probing inside its new windows in LDS, and writes those
// windows out whole, empty slots included, so no fill of the new table is needed.
// An entry that probes past its window's end is left for k_rehash_left, which
// inserts it with the ordinary atomic probe once every window is written.  Any
// insertion order gives a valid linear-probing table, and entries are distinct,
// so lookups find exactly what the old table held.
constexpr uint32_t kRehashW = 1024;
constexpr uint32_t kRehashMaxParts = 4;
constexpr uint32_t kOwnEmpty = ~0u;

__global__ __launch_bounds__(kThreads) void k_rehash_windows(const Slot* old, uint64_t old_cap, Slot* tab,
                                                             uint64_t new_mask, uint32_t cap_log, uint32_t* left,
                                                             uint32_t probe_limit) {
  __shared__ uint32_t own[kRehashMaxParts * kRehashW];
  const uint32_t parts = (uint32_t)((new_mask + 1) >> cap_log);
  const uint64_t old_mask = old_cap - 1;
  const uint64_t base = (uint64_t)blockIdx.x * kRehashW;
  for (uint32_t i = threadIdx.x; i < parts * kRehashW; i += kThreads) own[i] = kOwnEmpty;
  if (blockIdx.x == 0 && threadIdx.x == 0) tab[new_mask + 1] = old[old_cap];  // the overflow slot
  __syncthreads();
  auto place = [&](uint64_t p, uint64_t k0, uint64_t k1) {
    const uint64_t h = slot_home(Digest{k0, k1}, new_mask);
    const uint64_t off0 = (h & old_mask) - base;  // wraps above W when the home is elsewhere
    if (off0 >= kRehashW) return;
    uint32_t* w = own + (uint32_t)(h >> cap_log) * kRehashW;
    const uint32_t end = (uint32_t)std::min<uint64_t>(kRehashW, off0 + probe_limit);
    for (uint32_t off = (uint32_t)off0; off < end; off++)
      if (atomicCAS(&w[off], kOwnEmpty, (uint32_t)p) == kOwnEmpty) return;
    left[1 + atomicAdd(left, 1u)] = (uint32_t)p;  // left[0] counts; at most one entry per old slot
  };
  for (uint32_t i = threadIdx.x; i < kRehashW; i += kThreads) {
    const ulonglong2 k = *reinterpret_cast<const ulonglong2*>(&old[base + i].k0);
    if (k.x != TEMPTY) place(base + i, k.x, k.y);
  }
  // The cluster that runs past the window end, up to its first empty slot.
  if (threadIdx.x < 64) {
    const uint32_t lane = threadIdx.x;
    for (uint64_t q = base + kRehashW, seen = 0; seen + kRehashW < old_cap; q += 64, seen += 64) {  // never its own window
      const uint64_t p = q & old_mask;  // old_cap is a multiple of 64
      const ulonglong2 k = *reinterpret_cast<const ulonglong2*>(&old[p + lane].k0);
      const uint64_t empty = __ballot(k.x == TEMPTY);
      const uint32_t first = empty ? (uint32_t)__ffsll((long long)empty) - 1 : 64u;
      if (lane < first) place(p + lane, k.x, k.y);
      if (empty) break;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < parts * kRehashW; i += kThreads) {
    const uint32_t j = i / kRehashW, r = i % kRehashW;
    Slot* dst = tab + ((uint64_t)j << cap_log) + base + r;
    const uint32_t o = own[i];
    ulonglong2 a{TEMPTY, TEMPTY}, b{TEMPTY, TEMPTY};
    if (o != kOwnEmpty) {
      a = *reinterpret_cast<const ulonglong2*>(&old[o].k0);
      b = *reinterpret_cast<const ulonglong2*>(&old[o].val);
    }
    *reinterpret_cast<ulonglong2*>(&dst->k0) = a;
    *reinterpret_cast<ulonglong2*>(&dst->val) = b;
  }
}

// The entries that k_rehash_windows could not place inside their windows.
__global__ __launch_bounds__(kThreads) void k_rehash_left(const Slot* old, const uint32_t* left, uint32_t count,
                                                          Slot* tab, uint64_t mask, Ctl* ctl) {
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < count; i += gridDim.x * kThreads) {
    const Slot s = old[left[1 + i]];
    bool cl;
    table_insert(tab, mask, Digest{s.k0, s.k1}, s.val, true, &ctl->err, &cl);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

template <int HK>
static void launch_hash_chunks(const uint8_t* data, uint64_t len, uint32_t n, uint32_t C, Digest* out, bool fast,
                               hipStream_t st) {
  const uint32_t grid = (uint32_t)((n + kThreads - 1) / kThreads);
  if (fast) {
    switch (C) {
      case 16: k_hash_chunks<HK, 16><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 32: k_hash_chunks<HK, 32><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 64: k_hash_chunks<HK, 64><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 128: k_hash_chunks<HK, 128><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 256: k_hash_chunks<HK, 256><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 512: k_hash_chunks<HK, 512><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 1024: k_hash_chunks<HK, 1024><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 2048: k_hash_chunks<HK, 2048><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      case 4096: k_hash_chunks<HK, 4096><<<grid, kThreads, 0, st>>>(data, len, n, C, out); return;
      default: break;
    }
  }
  k_hash_chunks<HK, 0><<<grid, kThreads, 0, st>>>(data, len, n, C, out);
}

void hash_chunks(const uint8_t* data, uint64_t len, uint32_t C, int hk, Digest* out, hipStream_t st) {
  const uint32_t n = (uint32_t)((len + C - 1) / C);
  const bool fast = (C % 16 == 0) && ((uintptr_t)data % 16 == 0);
  if (hk == HASH_MD5) launch_hash_chunks<HASH_MD5>(data, len, n, C, out, fast, st);
  else launch_hash_chunks<HASH_MURMUR3>(data, len, n, C, out, fast, st);
  MTC_HIP(hipGetLastError());
}

// Chunk sizes with an LDS-staged leaf kernel; other sizes hash byte by byte.
static int staged_cb(uint32_t C, bool aligned) {
  if (!aligned) return 0;
  switch (C) {
    case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048: case 4096: return (int)C;
    default: return 0;
  }
}

template <int HK, int CB>
static void launch_leaf_cb(Launch<LeafArgs> L, uint32_t grid, hipStream_t st) {
  k_leaf<HK, CB><<<grid, kThreads, 0, st>>>(L);
}

template <int HK>
static void launch_leaf(int cb, Launch<LeafArgs> L, uint32_t grid, hipStream_t st) {
  switch (cb) {
    case 16: launch_leaf_cb<HK, 16>(L, grid, st); break;
    case 32: launch_leaf_cb<HK, 32>(L, grid, st); break;
    case 64: launch_leaf_cb<HK, 64>(L, grid, st); break;
    case 128: launch_leaf_cb<HK, 128>(L, grid, st); break;
    case 256: launch_leaf_cb<HK, 256>(L, grid, st); break;
    case 512: launch_leaf_cb<HK, 512>(L, grid, st); break;
    case 1024: launch_leaf_cb<HK, 1024>(L, grid, st); break;
    case 2048: launch_leaf_cb<HK, 2048>(L, grid, st); break;
    case 4096: launch_leaf_cb<HK, 4096>(L, grid, st); break;
    default: launch_leaf_cb<HK, 0>(L, grid, st); break;
  }
}

template <int HK>
static void launch_leaf1(int cb, Launch<LeafArgs> L, uint32_t grid, hipStream_t st) {
  switch (cb) {
    case 16: k_leaf1<HK, 16><<<grid, kThreads, 0, st>>>(L); break;
    case 32: k_leaf1<HK, 32><<<grid, kThreads, 0, st>>>(L); break;
    case 64: k_leaf1<HK, 64><<<grid, kThreads, 0, st>>>(L); break;
    case 128: k_leaf1<HK, 128><<<grid, kThreads, 0, st>>>(L); break;
    case 256: k_leaf1<HK, 256><<<grid, kThreads, 0, st>>>(L); break;
    case 512: k_leaf1<HK, 512><<<grid, kThreads, 0, st>>>(L); break;
    case 1024: k_leaf1<HK, 1024><<<grid, kThreads, 0, st>>>(L); break;
    case 2048: k_leaf1<HK, 2048><<<grid, kThreads, 0, st>>>(L); break;
    default: k_leaf1<HK, 4096><<<grid, kThreads, 0, st>>>(L); break;
  }
}

template <int HK>
static void launch_leaf1_direct(int cb, const LeafArgs& a, const Prologue& p, uint32_t grid, hipStream_t st) {
  switch (cb) {
    case 16: k_leaf1_direct<HK, 16><<<grid, kThreads, 0, st>>>(a, p); break;
    case 32: k_leaf1_direct<HK, 32><<<grid, kThreads, 0, st>>>(a, p); break;
    case 64: k_leaf1_direct<HK, 64><<<grid, kThreads, 0, st>>>(a, p); break;
    case 128: k_leaf1_direct<HK, 128><<<grid, kThreads, 0, st>>>(a, p); break;
    case 256: k_leaf1_direct<HK, 256><<<grid, kThreads, 0, st>>>(a, p); break;
    case 512: k_leaf1_direct<HK, 512><<<grid, kThreads, 0, st>>>(a, p); break;
    case 1024: k_leaf1_direct<HK, 1024><<<grid, kThreads, 0, st>>>(a, p); break;
    case 2048: k_leaf1_direct<HK, 2048><<<grid, kThreads, 0, st>>>(a, p); break;
    default: k_leaf1_direct<HK, 4096><<<grid, kThreads, 0, st>>>(a, p); break;
  }
}

// Plan sweep stages and root segments for n chunks (fuse: levels per stage).  The
// roots of depth c are counted per segment: a group of `group` consecutive tiles of
// the first stage (the tiles one wavefront of sweep B walks), one tile in later
// stages.  Per depth: the first segment (seg_base_h) and the segment geometry
// (seg_geo_h: tiles per segment | log2 positions per tile at depth c << 16).
static void plan_tree(mtc_dedup* h, uint32_t fuse_req, uint32_t group) {
  const Geo g = make_geo(h->n);
  const uint32_t fuse = std::max<uint32_t>(1, std::min<uint32_t>(kFuse, fuse_req));
  h->plan_fuse = fuse;
  h->plan_group = group;
  h->stages.clear();
  std::vector<uint32_t> nb(g.D + 1, 1);  // segments of depth c
  h->seg_geo_h.assign(g.D + 1, 1u);
  if (h->n > 1) {
    uint32_t d = g.D;
    while (d > 0) {
      const uint32_t dout = d > fuse ? d - fuse : 0;
      const uint32_t gs = h->stages.empty() ? group : 1u;
      h->stages.push_back(Stage{d, dout});
      for (uint32_t c = dout + 1; c <= d; c++) {
        nb[c] = (uint32_t)(((1ull << dout) + gs - 1) / gs);
        h->seg_geo_h[c] = gs | ((c - dout) << 16);
      }
      d = dout;
    }
  }
  h->seg_base_h.assign(g.D + 2, 0);
  for (uint32_t c = 0; c <= g.D; c++) h->seg_base_h[c + 1] = h->seg_base_h[c] + nb[c];
  h->plan_dirty = true;
}

// Tiles per wavefront of sweep B's first stage, which is also the segment group:
// every tile, one per wavefront, when the debug labels are recorded.
static uint32_t sweep_b_group(const mtc_dedup* h) {
  if (h->debug) return 1;
  return h->tune.sweep_tiles_b ? h->tune.sweep_tiles_b : kTilesPerWaveB;
}

// Digit values of the sort's first pass for nbins source checkpoints.
static uint32_t digits0(uint32_t nbins) {
  uint32_t bits = 0;
  while ((1ull << bits) < nbins) bits++;
  return 1u << std::min<uint32_t>(8, bits);
}

static uint32_t sort_blocks(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(kSortBlocks, (n + 16 * kThreads - 1) / (16 * kThreads))); }

static uint64_t scan_sums(uint64_t count) { return (count + kScanTile - 1) / kScanTile + 1; }

static void ensure_buffers(mtc_dedup* h, hipStream_t st) {
  const uint64_t n = h->n, nodes = 2 * n - 1;
  h->leaf_dg.ensure(n);
  h->labels.ensure(n);
  h->leaf_slot.ensure(n);
  const Geo g = make_geo(h->n);
  const uint64_t half_cap = h->stages.size() > 1 ? (1ull << h->stages[0].d_out) : 1;
  h->fr_lab.ensure(2 * half_cap);
  h->fr_dg.ensure(2 * half_cap);
  h->fr_slot.ensure(2 * half_cap);
  h->staging.ensure(nodes);
  const uint32_t nseg = h->seg_base_h.back();
  h->seg_cnt_f.ensure(nseg);
  h->seg_cnt_s.ensure(nseg);
  h->seg_off_f.ensure(nseg + 1);
  h->seg_off_s.ensure(nseg + 1);
  h->seg_base.ensure(h->seg_base_h.size() + h->seg_geo_h.size());
  if (h->plan_dirty) {
    std::vector<uint32_t> up(h->seg_base_h);
    up.insert(up.end(), h->seg_geo_h.begin(), h->seg_geo_h.end());
    MTC_HIP(hipMemcpyAsync(h->seg_base.p, up.data(), up.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    MTC_HIP(hipStreamSynchronize(st));  // the host vector is pageable
    h->plan_dirty = false;
  }
  h->list.ensure(n);
  h->dense.ensure(n);
  if (h->approach != MTC_APPROACH_BASIC) {  // listed digests beside the list, then dense
    h->list_dg.ensure(n);
    h->dense_dg.ensure(n);
    h->pend.ensure(n);
    h->pcnt.ensure((n + 63) / 64);
  }
  const uint64_t max_blk = (n + kThreads - 1) / kThreads;  // at most one block per tile
  h->blk_cnt.ensure(max_blk);
  h->blk_off.ensure(max_blk + 1);
  const uint64_t ntile0 = h->stages.empty() ? 1 : (1ull << h->stages[0].d_out);
  h->tile_flags.ensure(2 * ntile0 + 4);
  h->first_roots.ensure(n);
  h->shift_tmp.ensure(n);
  for (int b = 0; b < 2; b++) {
    h->sort_key[b].ensure(n);
    h->sort_val[b].ensure(n);
  }
  h->sort_hist.ensure(256ull * sort_blocks(h->n));
  h->region_off.ensure(n + 1);
  h->tile_lo.ensure(n / 256 + 2);
  // scan partial sums: leaf blocks, two root-segment arrays, the sort, the regions
  h->scan_ws.ensure(scan_sums(max_blk) + 2 * scan_sums(nseg) + scan_sums(256ull * sort_blocks(h->n)) + scan_sums(n));
  // P <= checkpoints since the baseline; reserve in steps of 4096 so the arena is
  // not reallocated every checkpoint.
  const uint64_t pcap = ((uint64_t)h->current_id + 2 + 4095) / 4096 * 4096;
  h->src_hist.ensure(pcap);
  h->ptable.ensure(2 * pcap);
  h->diff.ensure(32 + 12 * n + 8 * pcap + n * (uint64_t)h->chunk_size);
  if (h->async_copies) h->diff_alt.ensure(32 + 12 * n + 8 * pcap + n * (uint64_t)h->chunk_size);
  h->ctr.ensure(CTR_KINDS * kStripes);
  (void)g;
}

// MTC_REHASH=atomic sends every entry through the atomic probe, and
// MTC_REHASH_PROBE_LIMIT=k ends the in-window probe after k slots; both exist to
// test the less used paths (results never depend on them).
static void rehash_to(mtc_dedup* h, uint64_t cap, hipStream_t st) {
  if (cap > kMaxTableSlots) throw Status{MTC_ERR_TABLE_FULL, "table capacity above 2^31 slots"};
  const uint64_t old_cap = h->table_cap;
  h->table_tmp.ensure(cap + 1);
  uint32_t cap_log = 0;
  while ((1ull << cap_log) < old_cap) cap_log++;
  const char* mode = std::getenv("MTC_REHASH");
  const char* lim = std::getenv("MTC_REHASH_PROBE_LIMIT");
  const uint32_t probe_limit = lim ? (uint32_t)std::max(1L, std::atol(lim)) : kRehashW;
  if (old_cap % kRehashW == 0 && cap / old_cap <= kRehashMaxParts && !(mode && std::strcmp(mode, "atomic") == 0)) {
    // Windowed rehash; the entries that overflow a window go through k_rehash_left.
    DevBuf<uint32_t> left_buf;
    left_buf.ensure(old_cap + 1);
    k_fill_u64<<<1, 64, 0, st>>>(reinterpret_cast<uint64_t*>(left_buf.p), 1, 0);  // the count (and left[1])
    MTC_HIP(hipGetLastError());
    k_rehash_windows<<<(uint32_t)(old_cap / kRehashW), kThreads, 0, st>>>(h->table.p, old_cap, h->table_tmp.p, cap - 1,
                                                                         cap_log, left_buf.p, probe_limit);
    MTC_HIP(hipGetLastError());
    uint32_t left = 0;
    MTC_HIP(hipMemcpyAsync(&left, left_buf.p, sizeof(left), hipMemcpyDeviceToHost, st));
    MTC_HIP(hipStreamSynchronize(st));
    if (left) {
      k_rehash_left<<<grid_for(left, kThreads, 65536), kThreads, 0, st>>>(h->table.p, left_buf.p, left, h->table_tmp.p,
                                                                          cap - 1, h->ctl_d);
      MTC_HIP(hipGetLastError());
      MTC_HIP(hipStreamSynchronize(st));  // before left_buf is freed
    }
  } else {  // every entry through the atomic probe
    k_fill_u64<<<grid_for(4 * (cap + 1), kThreads, 65536), kThreads, 0, st>>>(
        reinterpret_cast<uint64_t*>(h->table_tmp.p), 4 * (cap + 1), TEMPTY);
    MTC_HIP(hipGetLastError());
    k_rehash<<<grid_for(old_cap + 1, kThreads, 65536), kThreads, 0, st>>>(h->table.p, old_cap + 1, h->table_tmp.p,
                                                                          cap - 1, h->ctl_d);
    MTC_HIP(hipGetLastError());
  }
  std::swap(h->table, h->table_tmp);
  h->table_tmp.release();
  h->table_cap = cap;
}

// The table is reset at a baseline (filled by the step's first launch).  Probing
// always ends while one slot stays empty, so a checkpoint may start only if the
// table holds all its possible inserts (the reference rehashes to size +
// num_nodes, src/tree_approach.cpp:1622-1623).  Returns false when the worst case
// (2n-1 new entries) fits and the table is at most 80% full: the step runs in one
// piece.  Otherwise the step pauses after the leaf kernel, when an exact bound is
// known (twice the chunks to insert: every new internal node has two new leaves
// below it), and grow_exact rehashes only if that bound does not fit.
// The next checkpoint runs in one piece in a table of `cap` slots: its worst case
// (2n - 1 new entries) fits and the table is at most 80% full.
static bool table_fits_step(const mtc_dedup* h, uint64_t cap) {
  const uint64_t nodes = 2ull * h->n - 1;
  return h->table_size + nodes < cap && h->table_size <= cap / 10 * 8;
}

// Slots of a baseline's digest table (without the overflow slot): a power of two
// holding n entries at the load factor, at least 1024.
uint64_t table_slots(uint64_t n, double load_factor) {
  const double lf = load_factor > 0.05 ? load_factor : 0.125;
  return next_pow2(std::max<uint64_t>((uint64_t)((double)n / lf), 1024));
}

// Device bytes a deduplicator of n chunks holds (the buffers of ensure_buffers,
// the table at the given load factor and the diff arena).
uint64_t state_bytes(uint64_t n, uint32_t chunk_size, double load_factor) {
  const uint64_t table = 32 * (table_slots(n, load_factor) + 1);
  const uint64_t per_chunk = 16 + 1 + 4 + 4 + 4 + 8 * 2 + 4 + 8 + 2 * (4 + 8) + 4 + 12 + 1 + 2 * 16 + 4 + 1;
  return table + per_chunk * n + (uint64_t)chunk_size * n + (8u << 20);
}

static bool prepare_table(mtc_dedup* h, bool reset, hipStream_t st) {
  (void)st;
  if (reset) {
    const uint64_t cap = table_slots(h->n, h->tune.load_factor);
    if (cap > kMaxTableSlots) throw Status{MTC_ERR_TABLE_FULL, "digest table above 2^31 slots"};
    h->table.ensure(cap + 1);
    h->table_cap = cap;
    h->table_size = 0;
  }
  return !table_fits_step(h, h->table_cap);
}

// Grows, between two checkpoints of a handle, a table that the next checkpoint
// could not run in without pausing after the leaf kernel (and perhaps rehashing
// there): the serving loop calls it once a client's result is posted, so the
// rehash falls outside its requests.  A rehash that reports a failed insert
// leaves the handle to start a new baseline.
void pregrow(mtc_dedup* h, hipStream_t st) {
  if (h->approach == MTC_APPROACH_BASIC || !h->has_state || h->table_cap == 0) return;
  if (table_fits_step(h, h->table_cap)) return;
  uint64_t cap = h->table_cap;
  while (!table_fits_step(h, cap)) cap *= 2;
  rehash_to(h, cap, st);
  uint32_t err = 0;
  MTC_HIP(hipMemcpyAsync(&err, &h->ctl_d->err, sizeof(err), hipMemcpyDeviceToHost, st));
  MTC_HIP(hipStreamSynchronize(st));
  if (err) h->has_state = false;
}

static void grow_exact(mtc_dedup* h, uint64_t to_insert, hipStream_t st) {
  const uint64_t need = h->table_size + 2 * to_insert + 1;
  uint64_t cap = h->table_cap;
  while (need > cap / 10 * 9 || h->table_size > cap / 10 * 8) cap *= 2;
  if (cap != h->table_cap) rehash_to(h, cap, st);
}

// Host-side phase timer, printed per step when MTC_HOST_TRACE is set.
struct HostTrace {
  bool on;
  std::chrono::steady_clock::time_point t0, last;
  std::string line;
  HostTrace() : on(std::getenv("MTC_HOST_TRACE") != nullptr) { t0 = last = std::chrono::steady_clock::now(); }
  void mark(const char* what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    line += std::string(what) + "=" + std::to_string(std::chrono::duration<double, std::micro>(now - last).count()) + "us ";
    last = now;
  }
  void print(size_t members) {
    if (on) std::fprintf(stderr, "[mtc host] members=%zu %stotal=%.1fus\n", members, line.c_str(),
                         std::chrono::duration<double, std::micro>(last - t0).count());
  }
};

// ---------------------------------------------------------------------------
// One step: a checkpoint of every member, as one sequence of batched launches.
// ---------------------------------------------------------------------------
namespace {

struct LaunchRec {
  uint64_t args = 0, prefix = 0, group = 0;
  uint32_t count = 0, grid = 0;
};

// Host image of the step's argument arena (uploaded with one copy).
struct ArenaImage {
  std::vector<uint8_t> bytes;
  uint64_t put(const void* p, uint64_t size) {
    const uint64_t off = (bytes.size() + 255) & ~255ull;
    bytes.resize(off + size);
    if (size) std::memcpy(bytes.data() + off, p, size);
    return off;
  }
  // One batched launch: an argument block and a block count per launch member.
  template <class A>
  LaunchRec add(const std::vector<A>& args, const std::vector<uint32_t>& blocks) {
    LaunchRec r;
    r.count = (uint32_t)args.size();
    std::vector<uint32_t> prefix(r.count + 1, 0);
    for (uint32_t i = 0; i < r.count; i++) prefix[i + 1] = prefix[i] + blocks[i];
    r.grid = prefix[r.count];
    r.args = put(args.data(), sizeof(A) * args.size());
    r.prefix = put(prefix.data(), sizeof(uint32_t) * prefix.size());
    if (r.count > 1) {
      std::vector<uint32_t> group((r.grid + 63) / 64);
      uint32_t m = 0;
      for (uint32_t gi = 0; gi < group.size(); gi++) {
        while (prefix[m + 1] <= 64 * gi) m++;
        group[gi] = m;
      }
      r.group = put(group.data(), sizeof(uint32_t) * group.size());
    }
    return r;
  }
};

template <class A>
Launch<A> bind(const LaunchRec& r, uint8_t* base) {
  return Launch<A>{reinterpret_cast<const A*>(base + r.args), reinterpret_cast<const uint32_t*>(base + r.prefix),
                   r.count > 1 ? reinterpret_cast<const uint32_t*>(base + r.group) : nullptr, r.count};
}

struct Member {
  mtc_dedup* h;
  const uint8_t* data;
  uint64_t len;
  bool baseline;
  Ctl* ctl_out = nullptr;  // this member's entry of the step's result array
  int arena = 0;           // diff arena this step writes
  // derived
  uint32_t n = 0, cur = 0, nblk_leaf = 0, nseg = 0, nbins = 0, passes = 0, Gs = 0, R0 = 0;
  uint64_t tpb = 1, ntile0 = 1;
  int cb = 0;
  bool split = false;  // the table bound is checked after the leaf kernel
  bool all_labels = true;  // the leaf kernel labels every chunk, not only the listed ones
  bool hash_later = false;  // k_compact hashes the listed chunks; the leaf kernel stores no digest
  uint64_t blk_zero = 0;   // leading block counts known to be zero when the step starts
  Geo g{};
  uint32_t* ws_blk = nullptr;  // scan partial sums of this member
  uint32_t* ws_f = nullptr;
  uint32_t* ws_s = nullptr;
  uint32_t* ws_sort = nullptr;
  uint32_t* ws_reg = nullptr;
};

// A step hashes its listed chunks again in k_compact when the previous checkpoint
// of the handle changed fewer than n / kHashLaterBelow chunks, for 128-byte chunks
// and buffers of at least kHashLaterMinChunks chunks.  Measured per checkpoint at
// 2% change: 8 GiB of 128-byte chunks 20 us faster (the leaf kernel 45 us faster,
// k_compact 25 us slower); 64 MiB 4 us slower; 64-byte chunks no faster; 4096-byte
// chunks 70 us slower (one lane reads a whole chunk).
constexpr uint64_t kHashLaterBelow = 16;
constexpr uint32_t kHashLaterMinChunks = 1u << 21;
constexpr uint32_t kHashLaterChunk = 128;
// MTC_HASH_LATER_MIN_CHUNKS=k replaces both bounds (k chunks, any staged chunk
// size), so that tests can run the mode on any buffer (results never depend on it).
static bool hash_later_fits(uint64_t n, uint32_t C) {
  const char* e = std::getenv("MTC_HASH_LATER_MIN_CHUNKS");
  if (e) return n >= std::strtoull(e, nullptr, 10);
  return C == kHashLaterChunk && n >= kHashLaterMinChunks;
}

// Recorded launches of a step: run eagerly or captured into a graph.  A tag names
// each launch's kernel (sweep B: 44 and 45, sweep A's tails: 48 and 49, one chunk: 30, leaf
// kernels: 100 up).
struct Ops {
  std::vector<std::function<void(uint8_t*, hipStream_t)>> fn;
  std::vector<uint64_t> sig;
  std::vector<uint32_t> tag;
  template <class A, class K>
  void add(uint32_t t, const LaunchRec& r, K launch) {
    if (r.grid == 0) return;
    sig.push_back(((uint64_t)t << 32) | r.grid);
    sig.push_back((r.args << 32) ^ (r.prefix << 8) ^ r.group ^ r.count);
    fn.push_back([r, launch](uint8_t* base, hipStream_t st) { launch(bind<A>(r, base), r.grid, st); });
    tag.push_back(t);
  }
};

}  // namespace

// Checks a member against its handle before anything of the step changes: a step
// that is refused leaves every handle as it was.
static void validate_member(Member& M) {
  const mtc_dedup* h = M.h;
  if (M.len == 0) throw Status{MTC_ERR_INVALID_ARG, "empty checkpoint buffer"};
  const uint64_t C = h->chunk_size;
  const uint64_t n64 = (M.len + C - 1) / C;
  if (n64 > (1ull << 31)) throw Status{MTC_ERR_INVALID_ARG, "more than 2^31 chunks"};
  M.baseline = M.baseline || h->current_id == 0 || !h->has_state;
  if (!M.baseline && M.len != h->data_len)
    throw Status{MTC_ERR_STATE, "incremental checkpoint with a different length than the previous one"};
  if (M.baseline && h->approach != MTC_APPROACH_BASIC && table_slots(n64, h->tune.load_factor) > kMaxTableSlots)
    throw Status{MTC_ERR_TABLE_FULL, "a digest table for " + std::to_string(n64) +
                                         " chunks needs more than 2^31 slots at this load factor"};
}

// Sets a validated member up: geometry, buffers, table and launch shape.  The
// handle holds no state until the step has completed (run_step sets has_state).
static void prepare_member(Member& M, hipStream_t st) {
  mtc_dedup* h = M.h;
  const uint64_t C = h->chunk_size;
  const uint64_t n64 = (M.len + C - 1) / C;
  if (M.baseline) {
    h->has_state = false;
    h->n = (uint32_t)n64;
    h->data_len = M.len;
    plan_tree(h, h->tune.fuse, sweep_b_group(h));
  } else if (h->plan_group != sweep_b_group(h)) {
    plan_tree(h, h->plan_fuse, sweep_b_group(h));  // same stages, new root segments
  }
  ensure_buffers(h, st);
  // Until this step completes, the labels are not known to be reset.  A baseline
  // lists every chunk, fixed ones included, so its leaf kernel labels them all.
  M.all_labels = M.baseline || !h->labels_clean || h->debug || h->approach == MTC_APPROACH_BASIC;
  h->labels_clean = false;
  M.split = h->approach == MTC_APPROACH_BASIC ? false : prepare_table(h, M.baseline, st);
  if (M.baseline) h->baseline_id = h->current_id;
  M.n = h->n;
  M.cur = h->current_id;
  M.g = make_geo(M.n);
  M.cb = staged_cb(h->chunk_size, (uintptr_t)M.data % 16 == 0);
  // With asynchronous copies, checkpoints alternate diff arenas; this step waits
  // (on the device) for the copy still reading the arena it is about to write.
  // The pending flag is cleared once the step has completed.
  M.arena = h->async_copies ? 1 - h->diff_cur : 0;
  if (h->copy_pending[M.arena]) MTC_HIP(hipStreamWaitEvent(st, h->copy_done[M.arena], 0));
  const uint64_t ntiles = (M.n + kThreads - 1) / kThreads;
  M.tpb = std::max<uint64_t>(1, h->tune.leaf_tiles_per_block);
  M.nblk_leaf = (uint32_t)((ntiles + M.tpb - 1) / M.tpb);
  // The leaf kernel's digest stores cost it far more than their bytes (stores
  // beside its read stream); when the previous checkpoint changed few chunks of a
  // large buffer, k_compact hashes the listed chunks again instead (Murmur3,
  // single-tile leaf).
  M.hash_later = !M.baseline && h->approach != MTC_APPROACH_BASIC && h->hash_kind == HASH_MURMUR3 && M.cb >= 16 &&
                 M.tpb == 1 && hash_later_fits(M.n, h->chunk_size) && h->ctl_h &&
                 h->ctl_h->changed * kHashLaterBelow <= M.n;
  // Block counts that k_compact cleared at the end of the last completed step (the
  // same buffer); unknown again until this step completes.
  M.blk_zero = h->blk_zero_p == h->blk_cnt.p ? h->blk_zero : 0;
  h->blk_zero = 0;
  M.nseg = h->seg_base_h.back();
  M.nbins = M.cur - h->baseline_id + 1;
  uint32_t bits = 0;
  while ((1ull << bits) < M.nbins) bits++;
  M.passes = (bits + 7) / 8;
  M.Gs = sort_blocks(M.n);
  M.R0 = h->stages.empty() ? 0 : h->stages[0].d_in - h->stages[0].d_out;
  M.ntile0 = h->stages.empty() ? 1 : (1ull << h->stages[0].d_out);
  uint32_t* w = h->scan_ws.p;
  M.ws_blk = w;
  w += scan_sums((M.n + kThreads - 1) / kThreads);
  M.ws_f = w;
  w += scan_sums(M.nseg);
  M.ws_s = w;
  w += scan_sums(M.nseg);
  M.ws_sort = w;
  w += scan_sums(256ull * M.Gs);
  M.ws_reg = w;
  if (h->debug) {
    h->dbg_tree.ensure(2ull * M.n - 1);
    h->dbg_labels.ensure(2ull * M.n + 3);  // whole words for the zero fill
  }
}

static void add_scan(Ops& ops, ArenaImage& img, const std::vector<ScanSeg>& segs) {
  if (segs.empty()) return;
  std::vector<uint32_t> tiles, ones(segs.size(), 1);
  if (std::all_of(segs.begin(), segs.end(), [](const ScanSeg& s) { return s.count_max <= kScanOneMax; })) {
    ops.add<ScanSeg>(13, img.add(segs, ones), [](Launch<ScanSeg> L, uint32_t g, hipStream_t st) { k_scan_one<<<g, kThreads, 0, st>>>(L); });
    return;
  }
  for (const auto& s : segs) tiles.push_back(std::max<uint32_t>(1, std::min<uint32_t>(512, (s.count_max + kScanTile - 1) / kScanTile)));
  const LaunchRec rt = img.add(segs, tiles);
  ops.add<ScanSeg>(10, rt, [](Launch<ScanSeg> L, uint32_t g, hipStream_t st) { k_scan_reduce<<<g, kThreads, 0, st>>>(L); });
  ops.add<ScanSeg>(14, rt, [](Launch<ScanSeg> L, uint32_t g, hipStream_t st) { k_scan_down2<<<g, kThreads, 0, st>>>(L); });
}

// The scan of the leaf blocks' counts (the offsets of each block's list of
// chunks to insert, and their total).
static void add_leaf_scan(std::vector<Member>& ms, Ops& ops, ArenaImage& img) {
  std::vector<ScanSeg> ss;
  for (auto& M : ms)  // its first block's start closes the leaf phase
    ss.push_back(ScanSeg{M.h->blk_cnt.p, M.h->blk_off.p, nullptr, M.nblk_leaf, M.h->blk_off.p + M.nblk_leaf, M.ws_blk,
                         &M.h->ctl_d->t_leaf});
  add_scan(ops, img, ss);
}

// After the leaf kernel, for tree members: table passes, both forests, roots and
// the diff.
static void build_tree_ops(std::vector<Member>& ms, Ops& ops, ArenaImage& img) {
  const size_t B = ms.size();
  // 3. dense table passes over the changed chunks
  {
    std::vector<CompactArgs> ca;
    std::vector<uint32_t> cb;
    std::vector<DenseArgs> da;
    std::vector<uint32_t> db, pb;
    for (auto& M : ms) {
      mtc_dedup* h = M.h;
      ca.push_back(CompactArgs{h->list.p, h->blk_cnt.p, h->blk_off.p, M.tpb, h->dense.p, M.nblk_leaf, h->list_dg.p,
                               h->dense_dg.p, M.data, M.len, h->chunk_size, M.hash_later ? 1 : 0});
      cb.push_back((M.nblk_leaf + kCompactPer - 1) / kCompactPer);
      da.push_back(DenseArgs{h->dense.p, h->blk_off.p + M.nblk_leaf, M.n, h->leaf_dg.p, h->leaf_slot.p, h->labels.p,
                             h->table.p, h->table_cap - 1, M.cur, h->ctl_d, h->ctr.p, M.g, M.R0, h->tile_flags.p,
                             h->tile_flags.p + M.ntile0, h->dense_dg.p, M.all_labels ? 1 : 0, h->pend.p,
                             h->pcnt.p});
      db.push_back(grid_for(M.n, 8 * kThreads, 16384));
      pb.push_back(grid_for(M.n, kPendPerBlock, 4096));
    }
    ops.add<CompactArgs>(20, img.add(ca, cb), [](Launch<CompactArgs> L, uint32_t g, hipStream_t st) { k_compact<<<g, kThreads, 0, st>>>(L); });
    const LaunchRec rd = img.add(da, db), rp = img.add(da, pb);
    ops.add<DenseArgs>(22, rd, [](Launch<DenseArgs> L, uint32_t g, hipStream_t st) { k_insert<<<g, kThreads, 0, st>>>(L); });
    ops.add<DenseArgs>(24, rp, [](Launch<DenseArgs> L, uint32_t g, hipStream_t st) { k_fixmin<<<g, kThreads, 0, st>>>(L); });
    ops.add<DenseArgs>(23, rp, [](Launch<DenseArgs> L, uint32_t g, hipStream_t st) { k_resolve<<<g, kThreads, 0, st>>>(L); });
  }
  // 4. sweeps: every stage of sweep A (and its inserts) before any stage of sweep B
  {
    std::vector<SweepArgs> base(B);
    size_t max_stages = 0;
    for (size_t i = 0; i < B; i++) {
      Member& M = ms[i];
      mtc_dedup* h = M.h;
      SweepArgs& sa = base[i];
      sa = SweepArgs{};
      sa.g = M.g;
      sa.cur = M.cur;
      sa.leaf_dg = h->leaf_dg.p;
      sa.labels = h->labels.p;
      sa.leaf_slot = h->leaf_slot.p;
      sa.tab = h->table.p;
      sa.mask = h->table_cap - 1;
      sa.ctl = h->ctl_d;
      sa.ctr = h->ctr.p;
      sa.staging = h->staging.p;
      sa.cnt_f = h->seg_cnt_f.p;
      sa.cnt_s = h->seg_cnt_s.p;
      sa.seg_base = h->seg_base.p;
      sa.dbg_tree = h->debug ? h->dbg_tree.p : nullptr;
      sa.dbg_labels = h->debug ? h->dbg_labels.p : nullptr;
      sa.tile_first = h->tile_flags.p;
      sa.tile_busy = h->tile_flags.p + M.ntile0;
      sa.compat = h->compat;
      max_stages = std::max(max_stages, h->stages.size());
    }
    // single-chunk members
    {
      std::vector<uint32_t> sb(B, 0);
      bool any = false;
      for (size_t i = 0; i < B; i++)
        if (ms[i].n == 1) sb[i] = 1, any = true;
      if (any)
        ops.add<SweepArgs>(30, img.add(base, sb), [](Launch<SweepArgs> L, uint32_t g, hipStream_t st) { k_single_leaf<<<g, 64, 0, st>>>(L); });
    }
    // Stage s of member i in this pass (its frontiers alternate between two halves).
    auto stage_args = [&](size_t i, size_t s, int pass) {
      mtc_dedup* h = ms[i].h;
      SweepArgs a = base[i];
      a.d_in = h->stages[s].d_in;
      a.d_out = h->stages[s].d_out;
      a.stage0 = (s == 0);
      const uint64_t half_cap = h->stages.size() > 1 ? (1ull << h->stages[0].d_out) : 1;
      const uint64_t in_off = (s % 2 == 1) ? 0 : half_cap, out_off = (s % 2 == 0) ? 0 : half_cap;
      a.in_lab = h->fr_lab.p + in_off;
      a.in_dg = h->fr_dg.p + in_off;
      a.in_slot = h->fr_slot.p + in_off;
      a.out_lab = h->fr_lab.p + out_off;
      a.out_dg = h->fr_dg.p + out_off;
      a.out_slot = h->fr_slot.p + out_off;
      // stage 0: several tiles per wavefront, their flags read at once; sweep B's
      // tiles per wavefront are the root segment group the plan was made for
      const uint32_t ta = h->tune.sweep_tiles_a ? h->tune.sweep_tiles_a : kTilesPerWaveA;
      a.tiles_per_wave = !a.stage0 ? 1u : (pass == 0 ? ta : h->plan_group);
      return a;
    };
    auto stage_waves = [](const SweepArgs& a) {
      return ((1ull << a.d_out) + a.tiles_per_wave - 1) / a.tiles_per_wave;
    };
    for (int pass = 0; pass < 2; pass++) {
      // Sweep A: a member's stages from tail[i] on each fit one block of kTailWaves
      // wavefronts, so they run in one launch after its other stages.  (Sweep B's
      // last stages in one kernel of their own took 13.0 us against 10.7 us for two
      // launches of the kernel that had just run stage 0 at 8 GiB: its code has to
      // be fetched again, and the walk of one of its tiles is mostly instructions.)
      std::vector<size_t> tail(B);
      for (size_t i = 0; i < B; i++) {
        const size_t ns = ms[i].n > 1 ? ms[i].h->stages.size() : 0;
        tail[i] = ns;
        while (pass == 0 && tail[i] > 0 && ns - (tail[i] - 1) <= kTailStages &&
               stage_waves(stage_args(i, tail[i] - 1, pass)) <= kTailWaves)
          tail[i]--;
        if (ns - tail[i] < 2) tail[i] = ns;  // a single stage keeps its own launch
      }
      for (size_t s = 0; s < max_stages; s++) {
        struct Group {
          std::vector<SweepArgs> args;
          std::vector<uint32_t> blocks;
        };
        std::map<int, Group> byhk;
        for (size_t i = 0; i < B; i++) {
          mtc_dedup* h = ms[i].h;
          Group& slot = byhk[h->hash_kind];
          SweepArgs a = base[i];
          uint32_t blocks = 0;
          if (s < tail[i]) {
            a = stage_args(i, s, pass);
            const uint32_t wpb = pass == 0 ? kSweepWpbA : kSweepWpbB;
            blocks = (uint32_t)((stage_waves(a) + wpb - 1) / wpb);
          }
          slot.args.push_back(a);
          slot.blocks.push_back(blocks);
        }
        for (auto& kv : byhk) {
          const int hk = kv.first;
          const uint32_t tag = 40 + 4 * (uint32_t)pass + (uint32_t)hk;
          const LaunchRec r = img.add(kv.second.args, kv.second.blocks);
          if (pass == 0)
            ops.add<SweepArgs>(tag, r, [hk](Launch<SweepArgs> L, uint32_t g, hipStream_t st) {
              if (hk == HASH_MD5) k_sweep_a<HASH_MD5, kSweepWpbA><<<g, 64 * kSweepWpbA, 0, st>>>(L);
              else k_sweep_a<HASH_MURMUR3, kSweepWpbA><<<g, 64 * kSweepWpbA, 0, st>>>(L);
            });
          else
            ops.add<SweepArgs>(tag, r, [hk](Launch<SweepArgs> L, uint32_t g, hipStream_t st) {
              if (hk == HASH_MD5) k_sweep_b<HASH_MD5, kSweepWpbB><<<g, 64 * kSweepWpbB, 0, st>>>(L);
              else k_sweep_b<HASH_MURMUR3, kSweepWpbB><<<g, 64 * kSweepWpbB, 0, st>>>(L);
            });
        }
      }
      if (pass == 1) continue;
      std::map<int, std::pair<std::vector<SweepTail>, std::vector<uint32_t>>> tails;
      for (size_t i = 0; i < B; i++) {
        auto& t = tails[ms[i].h->hash_kind];
        SweepTail st{};
        const size_t ns = ms[i].n > 1 ? ms[i].h->stages.size() : 0;
        for (size_t s = tail[i]; s < ns; s++) st.st[st.nst++] = stage_args(i, s, pass);
        t.first.push_back(st);
        t.second.push_back(st.nst ? 1u : 0u);
      }
      for (auto& kv : tails) {
        const int hk = kv.first;
        const LaunchRec r = img.add(kv.second.first, kv.second.second);
        ops.add<SweepTail>(48 + (uint32_t)hk, r, [hk](Launch<SweepTail> L, uint32_t g, hipStream_t st) {
          if (hk == HASH_MD5) k_sweep_a_tail<HASH_MD5><<<g, 64 * kTailWaves, 0, st>>>(L);
          else k_sweep_a_tail<HASH_MURMUR3><<<g, 64 * kTailWaves, 0, st>>>(L);
        });
      }
    }
  }
  // 5. root totals and compaction (segments are (depth, group) in node-id order)
  {
    std::vector<ScanSeg> ss;
    for (auto& M : ms) {
      mtc_dedup* h = M.h;
      ss.push_back(ScanSeg{h->seg_cnt_f.p, h->seg_off_f.p, nullptr, M.nseg, &h->ctl_d->F, M.ws_f});
      ss.push_back(ScanSeg{h->seg_cnt_s.p, h->seg_off_s.p, nullptr, M.nseg, &h->ctl_d->S, M.ws_s});
    }
    add_scan(ops, img, ss);
    std::vector<RootArgs> ra;
    std::vector<uint32_t> rb;
    for (auto& M : ms) {
      mtc_dedup* h = M.h;
      ra.push_back(RootArgs{h->staging.p, h->seg_cnt_f.p, h->seg_cnt_s.p, h->seg_off_f.p, h->seg_off_s.p, h->seg_base.p,
                            M.g.D + 1, M.nseg, h->first_roots.p, h->shift_tmp.p});
      rb.push_back(grid_for(M.nseg));
    }
    ops.add<RootArgs>(50, img.add(ra, rb), [](Launch<RootArgs> L, uint32_t g, hipStream_t st) { k_scatter_roots<<<g, kThreads, 0, st>>>(L); });
  }
  // 6. shifted entries: source checkpoint, stable sort by source, (chkpt, count) table
  std::vector<const uint64_t*> sorted(B);
  {
    std::vector<ShiftArgs> sa;
    std::vector<uint32_t> sb;
    uint32_t max_passes = 0;
    for (size_t i = 0; i < B; i++) {
      Member& M = ms[i];
      mtc_dedup* h = M.h;
      // k_shift_src counts the first pass's digits per sort block itself when its
      // histogram of sources is in LDS (fewer than kLdsBins checkpoints since the
      // baseline): one launch fewer.
      const bool fold = M.passes > 0 && M.nbins <= kLdsBins;
      sa.push_back(ShiftArgs{h->shift_tmp.p, &h->ctl_d->S, h->leaf_slot.p, M.n, h->table.p, h->table_cap - 1, h->baseline_id, M.cur,
                             h->sort_key[0].p, h->sort_val[0].p, h->src_hist.p, M.nbins, h->ctl_d,
                             fold ? h->sort_hist.p : nullptr, fold ? digits0(M.nbins) : 0u});
      sb.push_back(M.Gs);
      max_passes = std::max(max_passes, M.passes);
      sorted[i] = h->sort_val[M.passes % 2].p;
    }
    ops.add<ShiftArgs>(60, img.add(sa, sb), [](Launch<ShiftArgs> L, uint32_t g, hipStream_t st) { k_shift_src<<<g, kThreads, 0, st>>>(L); });
    for (uint32_t p = 0; p < max_passes; p++) {
      std::vector<SortArgs> so;
      std::vector<uint32_t> sblk, cblk;
      std::vector<ScanSeg> hs;
      for (auto& M : ms) {
        mtc_dedup* h = M.h;
        uint32_t bits = 0;
        while ((1ull << bits) < M.nbins) bits++;
        const uint32_t shift = 8 * p;
        const uint32_t nb = (p < M.passes) ? (1u << std::min<uint32_t>(8, bits - shift)) : 1u;
        so.push_back(SortArgs{h->sort_key[p % 2].p, h->sort_val[p % 2].p, h->sort_key[1 - p % 2].p,
                              h->sort_val[1 - p % 2].p, &h->ctl_d->S, shift, nb, h->sort_hist.p});
        sblk.push_back(p < M.passes ? M.Gs : 0u);
        cblk.push_back(p < M.passes && !(p == 0 && M.nbins <= kLdsBins) ? M.Gs : 0u);  // else counted by k_shift_src
        if (p < M.passes) hs.push_back(ScanSeg{h->sort_hist.p, h->sort_hist.p, nullptr, nb * M.Gs, nullptr, M.ws_sort});
      }
      const LaunchRec r = img.add(so, sblk), rc = img.add(so, cblk);
      ops.add<SortArgs>(61, rc, [](Launch<SortArgs> L, uint32_t g, hipStream_t st) { k_sort_count<<<g, kThreads, 0, st>>>(L); });
      add_scan(ops, img, hs);
      ops.add<SortArgs>(62, r, [](Launch<SortArgs> L, uint32_t g, hipStream_t st) { k_sort_move<<<g, kThreads, 0, st>>>(L); });
    }
  }
  // 7. region sizes and offsets, metadata, gathered data, counters
  {
    std::vector<RegionArgs> ra;
    std::vector<uint32_t> rb;
    std::vector<ScanSeg> ss;
    for (auto& M : ms) {
      mtc_dedup* h = M.h;
      ra.push_back(RegionArgs{h->first_roots.p, &h->ctl_d->F, M.g, h->region_off.p, h->compat,
                              PtableArgs{h->src_hist.p, M.nbins, h->baseline_id, h->ptable.p, h->ctl_d}});
      rb.push_back(grid_for(M.n, 16 * kThreads, 256));
      ss.push_back(ScanSeg{h->region_off.p, h->region_off.p, &h->ctl_d->F, M.n, h->region_off.p + M.n, M.ws_reg,
                           nullptr, h->tile_lo.p});
    }
    ops.add<RegionArgs>(70, img.add(ra, rb), [](Launch<RegionArgs> L, uint32_t g, hipStream_t st) { k_region_sizes<<<g, kThreads, 0, st>>>(L); });
    add_scan(ops, img, ss);
    std::vector<DiffArgs> da;
    std::vector<uint32_t> gb;
    for (size_t i = 0; i < B; i++) {
      Member& M = ms[i];
      mtc_dedup* h = M.h;
      DiffArgs d{};
      d.diff = M.arena ? h->diff_alt.p : h->diff.p;
      d.first_roots = h->first_roots.p;
      d.ptable = h->ptable.p;
      d.shift_sorted = sorted[i];
      d.region_off = h->region_off.p;
      d.tile_lo = h->tile_lo.p;
      d.total_chunks = h->region_off.p + M.n;
      d.data = M.data;
      d.len = M.len;
      d.n = M.n;
      d.C = h->chunk_size;
      d.ref_id = h->baseline_id;
      d.chkpt_id = M.cur;
      d.g = M.g;
      d.ctl = h->ctl_d;
      d.aligned = ((h->chunk_size % 16) == 0 && (uintptr_t)M.data % 16 == 0) ? 1 : 0;
      d.compat = h->compat;
      d.ctr = h->ctr.p;
      d.changed = h->blk_off.p + M.nblk_leaf;
      d.out = M.ctl_out;
      d.dense = h->dense.p;
      d.reset_labels = h->debug ? nullptr : h->labels.p;  // the debug view copies them after the step
      da.push_back(d);
      gb.push_back(grid_for(M.n, 8 * kThreads, 8192));
    }
    ops.add<DiffArgs>(72, img.add(da, gb), [](Launch<DiffArgs> L, uint32_t g, hipStream_t st) { k_assemble<<<g, kThreads, 0, st>>>(L); });
  }
}

// After the leaf kernel, for basic members (src/basic_approach.cpp:14-142): the
// diff of the changed chunks, whose ascending list the leaf pass and compaction
// leave.
static void build_basic_ops(std::vector<Member>& ms, Ops& ops, ArenaImage& img) {
  // the leaf blocks' list regions, concatenated in block order (ascending)
  {
    std::vector<CompactArgs> ca;
    std::vector<uint32_t> cb;
    for (auto& M : ms) {
      mtc_dedup* h = M.h;
      ca.push_back(CompactArgs{h->list.p, h->blk_cnt.p, h->blk_off.p, M.tpb, h->dense.p, M.nblk_leaf, nullptr, nullptr});
      cb.push_back((M.nblk_leaf + kCompactPer - 1) / kCompactPer);
    }
    ops.add<CompactArgs>(21, img.add(ca, cb), [](Launch<CompactArgs> L, uint32_t g, hipStream_t st) { k_compact<<<g, kThreads, 0, st>>>(L); });
  }
  std::vector<BasicDiffArgs> da;
  std::vector<uint32_t> gb;
  for (auto& M : ms) {
    mtc_dedup* h = M.h;
    BasicDiffArgs d{};
    d.diff = M.arena ? h->diff_alt.p : h->diff.p;
    d.dense = h->dense.p;
    d.total = h->blk_off.p + M.nblk_leaf;
    d.data = M.data;
    d.len = M.len;
    d.n = M.n;
    d.C = h->chunk_size;
    d.ref_id = h->baseline_id;
    d.chkpt_id = M.cur;
    d.ctl = h->ctl_d;
    d.aligned = ((h->chunk_size % 16) == 0 && (uintptr_t)M.data % 16 == 0) ? 1 : 0;
    d.ctr = h->ctr.p;
    d.out = M.ctl_out;
    da.push_back(d);
    gb.push_back(grid_for(M.n, 8 * kThreads, 8192));
  }
  ops.add<BasicDiffArgs>(80, img.add(da, gb), [](Launch<BasicDiffArgs> L, uint32_t g, hipStream_t st) { k_basic_diff<<<g, kThreads, 0, st>>>(L); });
}

// The step's launch sequence.  Part 1: the fills (control blocks, counters, tile
// flags, histograms, debug views) and the leaf kernels, plus the scan of the leaf
// blocks' counts when `leaf_scan` (steps that read the count of chunks to insert
// before part 2).  Part 2: the rest.  Part 0: both.
// Launch of a step's single member's leaf kernel with its arguments by value (see
// k_leaf1_direct), when build_ops leaves it out of the recorded launches.
struct DirectLeaf {
  bool ok = false;
  int hk = 0, cb = 0;
  uint32_t grid = 0;
  LeafArgs la{};
  Prologue pro{};
};

static bool lean_leaf(const Member& M) { return M.cb >= 16 && M.tpb == 1 && M.h->approach != MTC_APPROACH_BASIC; }

static void build_ops(std::vector<Member>& ms, Ops& ops, ArenaImage& img, int phase, bool leaf_scan,
                      DirectLeaf* direct = nullptr) {
  const size_t B = ms.size();
  // One member whose leaf kernel adds to block counts known to be zero: its leaf
  // kernel goes first, by value, and does the fills itself.
  const bool dl = direct && phase == 0 && B == 1 && lean_leaf(ms[0]) && ms[0].blk_zero >= ms[0].nblk_leaf;
  if (phase != 2) {
    {
      std::vector<FillSeg> fs;
      std::vector<uint32_t> fb;
      auto fill = [&](void* p, uint64_t words, uint32_t v) {
        if (!words) return;
        fs.push_back(FillSeg{reinterpret_cast<uint32_t*>(p), words, v});
        fb.push_back(grid_for(words, kThreads * 4, 4096));
      };
      for (auto& M : ms) {
        mtc_dedup* h = M.h;
        if (dl) {  // the leaf kernel's first block stamps t_begin meanwhile
          constexpr uint64_t tb = offsetof(Ctl, t_begin) / 4, words = sizeof(Ctl) / 4;
          fill(h->ctl_d, tb, 0);
          fill(reinterpret_cast<uint32_t*>(h->ctl_d) + tb + 2, words - tb - 2, 0);
        } else {
          fill(h->ctl_d, sizeof(Ctl) / 4, 0);
        }
        fill(h->ctr.p, 2ull * CTR_KINDS * kStripes, 0);
        fill(h->tile_flags.p, (2 * M.ntile0 + 3) / 4, 0);
        fill(h->seg_cnt_f.p, M.nseg, 0);
        fill(h->seg_cnt_s.p, M.nseg, 0);
        fill(h->src_hist.p, M.nbins, 0);
        if (!dl) fill(h->blk_cnt.p, M.nblk_leaf, 0);  // counts that k_leaf1 accumulates
        if (h->debug) {
          fill(h->dbg_tree.p, 4 * (2ull * M.n - 1), 0);
          fill(h->dbg_labels.p, (2ull * M.n - 1 + 3) / 4, 0);
        }
      }
      if (dl) {
        if (fs.size() > kPrologueFills) throw HipError{hipErrorInvalidValue, "too many fills for a direct leaf launch"};
        for (size_t i = 0; i < fs.size(); i++) direct->pro.fill[i] = fs[i];
        direct->pro.nfill = (uint32_t)fs.size();
      } else {
        ops.add<FillSeg>(1, img.add(fs, fb), [](Launch<FillSeg> L, uint32_t g, hipStream_t st) { k_fill<<<g, kThreads, 0, st>>>(L); });
      }
    }
    // leaf kernel, one launch per (hash, staged chunk size, single tile) group
    std::map<std::tuple<int, int, bool>, std::vector<size_t>> groups;
    for (size_t i = 0; i < B; i++) groups[{ms[i].h->hash_kind, ms[i].cb, lean_leaf(ms[i])}].push_back(i);
    for (auto& kv : groups) {
      std::vector<LeafArgs> la;
      std::vector<uint32_t> lb;
      for (size_t i : kv.second) {
        Member& M = ms[i];
        mtc_dedup* h = M.h;
        // basic approach: every chunk of the first checkpoint is changed (src/basic_approach.cpp:36-45)
        const int reset = !M.baseline ? 0 : (h->approach == MTC_APPROACH_BASIC && M.cur == 0) ? 2 : 1;
        la.push_back(LeafArgs{M.data, M.len, M.n, h->chunk_size, h->leaf_dg.p, h->labels.p, reset,
                              h->list.p, h->blk_cnt.p, M.tpb, M.all_labels ? 1 : 0,
                              h->approach == MTC_APPROACH_BASIC ? nullptr : h->list_dg.p, &h->ctl_d->t_begin,
                              M.hash_later ? 1 : 0});
        lb.push_back(M.nblk_leaf);
      }
      const int hk = std::get<0>(kv.first), cb = std::get<1>(kv.first);
      const bool lean = std::get<2>(kv.first);
      if (dl) {
        direct->ok = true;
        direct->hk = hk;
        direct->cb = cb;
        direct->grid = lb[0];
        direct->la = la[0];
        continue;
      }
      const uint32_t tag = 100 + (uint32_t)hk * 32 + (lean ? 16u : 0u) + (uint32_t)(cb ? __builtin_ctz((unsigned)cb) : 0);
      ops.add<LeafArgs>(tag, img.add(la, lb), [hk, cb, lean](Launch<LeafArgs> L, uint32_t g, hipStream_t st) {
        if (lean) {
          if (hk == HASH_MD5) launch_leaf1<HASH_MD5>(cb, L, g, st);
          else launch_leaf1<HASH_MURMUR3>(cb, L, g, st);
        } else if (hk == HASH_MD5) {
          launch_leaf<HASH_MD5>(cb, L, g, st);
        } else {
          launch_leaf<HASH_MURMUR3>(cb, L, g, st);
        }
      });
    }
    if (leaf_scan || phase == 0) add_leaf_scan(ms, ops, img);
    if (phase == 1) return;
  }
  if (phase == 2 && !leaf_scan) add_leaf_scan(ms, ops, img);
  std::vector<Member> tree, basic;
  for (auto& M : ms) (M.h->approach == MTC_APPROACH_BASIC ? basic : tree).push_back(M);
  if (!tree.empty()) build_tree_ops(tree, ops, img);
  if (!basic.empty()) build_basic_ops(basic, ops, img);
}

void StepCtx::reset_graph() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  sig.clear();
}

StepCtx::~StepCtx() {
  reset_graph();
  if (arena_h) (void)hipHostFree(arena_h);
  if (lead_h) (void)hipHostFree(lead_h);
  if (ctl_h) (void)hipHostFree(ctl_h);
  if (cnt_h) (void)hipHostFree(cnt_h);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
}

static void run_step_body(StepCtx& ctx, std::vector<Member>& ms, hipStream_t st, bool use_graph);

// Runs one checkpoint of every member.  Every member is checked before any handle
// changes.  Once work is under way, a failure of any kind (a HIP error, or a
// member whose kernels report one) leaves every member of the step without state,
// so that each handle's next checkpoint is a new baseline: the step may already
// have overwritten leaf digests and added table entries of every member.
void run_step(StepCtx& ctx, std::vector<Member>& ms, hipStream_t st, bool use_graph) {
  for (auto& M : ms) validate_member(M);
  try {
    run_step_body(ctx, ms, st, use_graph);
  } catch (...) {
    for (auto& M : ms) M.h->has_state = false;
    throw;
  }
}

// A step in two parts when it is replayed from a graph: part 1 (fills and leaf
// kernels) runs at once from its own arena while part 2 is built, then part 2 is
// replayed from a graph when its launch shape repeats.  Otherwise every launch
// goes out from one arena.  A step whose table might overflow pauses between the
// parts for the exact bound (the scan of the leaf blocks' counts then runs in
// part 1).
static void run_step_body(StepCtx& ctx, std::vector<Member>& ms, hipStream_t st, bool use_graph) {
  HostTrace ht;
  const size_t B = ms.size();
  if (!ctx.ev[0])
    for (auto& e : ctx.ev) MTC_HIP(hipEventCreate(&e));
  MTC_HIP(hipEventRecord(ctx.ev[0], st));
  for (auto& M : ms) prepare_member(M, st);
  // Reset tables of baseline members (basic members keep no table).
  for (auto& M : ms) {
    if (!M.baseline || M.h->approach == MTC_APPROACH_BASIC) continue;
    if (!M.h->table.p || M.h->table_cap == 0) throw Status{MTC_ERR_STATE, "digest table not allocated"};
    k_fill_u64<<<grid_for(4 * (M.h->table_cap + 1), kThreads, 65536), kThreads, 0, st>>>(
        reinterpret_cast<uint64_t*>(M.h->table.p), 4 * (M.h->table_cap + 1), TEMPTY);
  }
  MTC_HIP(hipGetLastError());
  MTC_HIP(hipEventRecord(ctx.ev[1], st));
  ht.mark("prepare");

  const bool debug = std::any_of(ms.begin(), ms.end(), [](const Member& M) { return M.h->debug != 0; });
  const bool split = std::any_of(ms.begin(), ms.end(), [](const Member& M) { return M.split; });
  const bool graph = use_graph && !debug;
  // Each member's last kernel writes its control block straight into pinned host
  // memory: no copy (and no idle time before one) ends the step.
  if (ctx.ctl_h_cap < B) {
    if (ctx.ctl_h) (void)hipHostFree(ctx.ctl_h);
    ctx.ctl_h = nullptr;
    MTC_HIP(hipHostMalloc((void**)&ctx.ctl_h, B * sizeof(Ctl)));
    ctx.ctl_h_cap = (uint32_t)B;
  }
  for (size_t i = 0; i < B; i++) ms[i].ctl_out = ctx.ctl_h + i;
  // Uploads an arena image and runs its launches, replayed from a graph or eagerly.
  // Phase boundaries are wall-clock stamps the kernels write into each member's
  // control block: an event recorded between two launches stalled the queue for
  // about 6 us each.
  auto run_ops = [&](Ops& ops, ArenaImage& img, bool as_graph, DevBuf<uint8_t>& arena_d, uint8_t*& arena_h,
                     uint64_t& arena_h_cap, DirectLeaf* direct = nullptr) {
    const uint64_t bytes = (img.bytes.size() + 15) & ~15ull;  // whole 16-byte words
    img.bytes.resize(bytes);
    const uint8_t* old_arena = arena_d.p;
    arena_d.ensure(bytes);
    if (arena_h_cap < bytes) {
      if (arena_h) (void)hipHostFree(arena_h);
      arena_h = nullptr;
      MTC_HIP(hipHostMalloc((void**)&arena_h, bytes));
      arena_h_cap = bytes;
    }
    std::memcpy(arena_h, img.bytes.data(), bytes);
    if (direct && direct->ok) {  // the leaf kernel copies the arena
      direct->pro.src = reinterpret_cast<const uint4*>(arena_h);
      direct->pro.dst = reinterpret_cast<uint4*>(arena_d.p);
      direct->pro.n16 = bytes / 16;
      if (direct->hk == HASH_MD5) launch_leaf1_direct<HASH_MD5>(direct->cb, direct->la, direct->pro, direct->grid, st);
      else launch_leaf1_direct<HASH_MURMUR3>(direct->cb, direct->la, direct->pro, direct->grid, st);
      MTC_HIP(hipGetLastError());
    } else {
      MTC_HIP(hipMemcpyAsync(arena_d.p, arena_h, bytes, hipMemcpyHostToDevice, st));
    }
    uint8_t* base = arena_d.p;
    if (as_graph) {
      ops.sig.push_back(reinterpret_cast<uint64_t>(base));
      if (!ctx.exec || ctx.sig != ops.sig || old_arena != base) {
        ctx.reset_graph();
        hipGraph_t g = nullptr;
        MTC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (auto& f : ops.fn) f(base, st);
        MTC_HIP(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&ctx.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        MTC_HIP(e);
        ctx.sig = ops.sig;
      }
      MTC_HIP(hipGraphLaunch(ctx.exec, st));
      return;
    }
    for (size_t i = 0; i < ops.fn.size(); i++) {
      ops.fn[i](base, st);
      MTC_HIP(hipGetLastError());
    }
  };
  if (split || graph) {
    Ops ops1;
    ArenaImage img1;
    build_ops(ms, ops1, img1, 1, split);
    run_ops(ops1, img1, false, ctx.lead_d, ctx.lead_h, ctx.lead_h_cap);
    if (split) {  // the exact table bound of every member
      if (ctx.cnt_h_cap < B) {
        if (ctx.cnt_h) (void)hipHostFree(ctx.cnt_h);
        ctx.cnt_h = nullptr;
        MTC_HIP(hipHostMalloc((void**)&ctx.cnt_h, B * sizeof(uint32_t)));
        ctx.cnt_h_cap = (uint32_t)B;
      }
      for (size_t i = 0; i < B; i++)
        MTC_HIP(hipMemcpyAsync(ctx.cnt_h + i, ms[i].h->blk_off.p + ms[i].nblk_leaf, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, st));
      MTC_HIP(hipStreamSynchronize(st));
      ht.mark("phase 1");
      for (size_t i = 0; i < B; i++)
        if (ms[i].split) grow_exact(ms[i].h, ctx.cnt_h[i], st);
    }
    Ops ops2;
    ArenaImage img2;
    build_ops(ms, ops2, img2, 2, split);
    ht.mark("build");
    run_ops(ops2, img2, graph && !split, ctx.arena_d, ctx.arena_h, ctx.arena_h_cap);
  } else {
    Ops ops;
    ArenaImage img;
    DirectLeaf direct;
    build_ops(ms, ops, img, 0, false, &direct);
    ht.mark("build");
    run_ops(ops, img, false, ctx.arena_d, ctx.arena_h, ctx.arena_h_cap, &direct);
  }
  if (debug)
    for (auto& M : ms)
      if (M.h->debug) {  // leaf digests and labels into the debug views
        MTC_HIP(hipMemcpyAsync(M.h->dbg_tree.p + (M.n - 1), M.h->leaf_dg.p, 16ull * M.n, hipMemcpyDeviceToDevice, st));
        MTC_HIP(hipMemcpyAsync(M.h->dbg_labels.p + (M.n - 1), M.h->labels.p, M.n, hipMemcpyDeviceToDevice, st));
      }
  ht.mark("launch");
  MTC_HIP(hipStreamSynchronize(st));
  ht.mark("sync");
  float ms_setup = 0;
  MTC_HIP(hipEventElapsedTime(&ms_setup, ctx.ev[0], ctx.ev[1]));
  if (ctx.clock_khz == 0) {
    int dev = 0, khz = 0;
    MTC_HIP(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    ctx.clock_khz = khz;
  }
  // MTC_TEST_FAIL_MEMBER=i makes member i report an inconsistent state (tests of
  // the failure path; nothing else reads it).
  const char* inject = std::getenv("MTC_TEST_FAIL_MEMBER");
  const long fail_member = inject ? std::atol(inject) : -1;
  for (size_t i = 0; i < B; i++) {
    const uint32_t e = ctx.ctl_h[i].err | ((long)i == fail_member ? (uint32_t)ERR_CORRUPT : 0u);
    if (e & (ERR_TABLE_FULL | ERR_TABLE_SPIN | ERR_OVERFLOW_SLOT))
      throw Status{MTC_ERR_TABLE_FULL, "member " + std::to_string(i) + ": digest table insert failed (err bits " +
                                           std::to_string(e) + ")"};
    if (e)
      throw Status{MTC_ERR_STATE, "member " + std::to_string(i) + ": inconsistent table state (err bits " +
                                      std::to_string(e) + ")"};
  }
  for (size_t i = 0; i < B; i++) {
    mtc_dedup* h = ms[i].h;
    const Ctl& c = ctx.ctl_h[i];
    h->has_state = true;
    // k_assemble reset the listed chunks' labels unless the debug view needed them
    h->labels_clean = h->approach != MTC_APPROACH_BASIC && !h->debug;
    h->blk_zero = std::max<uint64_t>(ms[i].blk_zero, ms[i].nblk_leaf);
    h->blk_zero_p = h->blk_cnt.p;
    h->copy_pending[ms[i].arena] = false;  // the step waited for that copy and has completed
    *h->ctl_h = c;
    h->table_size += c.claims;
    h->last_diff_size = c.diff_size;
    h->diff_cur = ms[i].arena;
    h->last_header = mtc_header{h->baseline_id, ms[i].cur, ms[i].len, h->chunk_size, c.F, c.P, c.S};
    // Timers (src/tree_approach.cpp:1640,1657,1672): set-up (buffers, table reset);
    // labels and forests (leaf kernel, table passes, both sweeps); the diff (root
    // compaction to assembly).  The basic approach has no forests: its whole step
    // counts as the second.  The step runs from the leaf kernel's first block to
    // its last kernel writing the member's record; the leaf time, to the leaf
    // scan's first block.
    const double tick_ms = 1.0 / ctx.clock_khz;
    const float ms_total = c.t_end > c.t_begin ? (float)((c.t_end - c.t_begin) * tick_ms) : 0.0f;
    const float ms_leaf = c.t_leaf > c.t_begin ? (float)((c.t_leaf - c.t_begin) * tick_ms) : 0.0f;
    const float ms_forest = c.t_forest > c.t_begin ? std::min(ms_total, (float)((c.t_forest - c.t_begin) * tick_ms)) : ms_total;
    h->last_ms[0] = ms_setup;
    h->last_ms[1] = ms_forest;
    h->last_ms[2] = ms_total - ms_forest;
    h->last_ms[3] = 0;
    h->last_ms[4] = ms_leaf;
    h->last_leaf_labels = ms[i].all_labels ? 1 : 0;
    h->last_leaf_digests = ms[i].hash_later ? 0 : 1;
  }
  ht.print(B);
}

void checkpoint_impl(mtc_dedup* h, const uint8_t* data, uint64_t len, bool make_baseline, hipStream_t st) {
  std::vector<Member> ms{Member{h, data, len, make_baseline}};
  run_step(h->step, ms, st, false);
}

void checkpoint_batch_impl(mtc_dedup* const* hs, const uint8_t* const* datas, const uint64_t* lens,
                           const int* baselines, uint32_t count, StepCtx& ctx, hipStream_t st, bool use_graph) {
  if (count > kMaxBatch) throw Status{MTC_ERR_INVALID_ARG, "more than 65536 members in one batch"};
  std::vector<Member> ms;
  ms.reserve(count);
  std::unordered_set<const mtc_dedup*> seen;
  seen.reserve(2 * (size_t)count);
  for (uint32_t i = 0; i < count; i++) {
    if (!seen.insert(hs[i]).second) throw Status{MTC_ERR_INVALID_ARG, "a handle appears twice in one batch"};
    ms.push_back(Member{hs[i], datas[i], lens[i], baselines && baselines[i] != 0});
  }
  run_step(ctx, ms, st, use_graph);
}

}  // namespace mtc
