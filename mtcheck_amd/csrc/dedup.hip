// dedup.hip -- the checkpoint path: leaf hashing and labelling, the two fused
// Merkle-forest sweeps, root compaction in node-id order, and diff assembly.
//
// Serial semantics followed: include/reference_impl.hpp:15-132 (labels and region
// roots) and src/tree_approach.cpp:511-710 (diff layout), in the canonical order
// of the serial reference (roots ascending; shifted entries grouped by source
// checkpoint, ascending node id inside a group).
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "internal.hpp"

namespace mtc {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kFuse = 8;              // tree levels fused per sweep stage (one wavefront tile)
constexpr uint64_t kLeafBlocks = 1ull << 40; // leaf-kernel blocks (default: one 256-chunk tile each)
constexpr uint64_t kFirstTag = 0xFFFFFFFFull;  // staging tag of a first-occurrence root

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
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

__global__ void k_reduce_ctr(const unsigned long long* ctr, const uint32_t* changed, Ctl* c) {
  __shared__ unsigned long long s[CTR_KINDS];
  if (threadIdx.x < CTR_KINDS) s[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t k = 0; k < CTR_KINDS; k++) {
    const unsigned long long v = ctr[k * kStripes + threadIdx.x];
    if (v) atomicAdd(&s[k], v);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    c->claims = s[CTR_CLAIMS];
    c->changed = *changed;
  }
}

__global__ void k_zero_ctl(Ctl* c, unsigned long long* ctr) {
  for (uint32_t i = threadIdx.x; i < CTR_KINDS * kStripes; i += blockDim.x) ctr[i] = 0;
  if (threadIdx.x == 0) {
    c->err = 0;
    c->F = c->S = c->P = 0;
    c->claims = 0;
    c->data_chunks = 0;
    c->diff_size = 0;
    c->changed = 0;
  }
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

// ---------------------------------------------------------------------------
// LDS-staged tile hashing.  A 256-thread block hashes 256 consecutive chunks, one
// per thread.  The chunks are brought in as segments of SEG bytes per chunk with
// fully coalesced 16-byte loads (each block-wide load instruction covers 4 KiB),
// written to LDS with a 16-byte pad per chunk (so the per-thread 16-byte reads
// of the hashing phase hit distinct banks), then each thread hashes its own chunk
// from LDS.  Murmur3 keeps (h1, h2) across segments; MD5 keeps its state and
// compresses every 64 bytes.
// ---------------------------------------------------------------------------
template <int CB>
struct TileShape {
  static constexpr int SEG = CB < 128 ? CB : 128;  // bytes of each chunk per pass
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
struct LeafArgs {
  const uint8_t* data;
  uint64_t len;
  uint32_t n;
  uint32_t C;
  Digest* leaf_dg;
  uint8_t* labels;
  int baseline;
  uint32_t* list;     // chunks to insert, packed per block
  uint32_t* blk_cnt;  // their count per block
  uint64_t tiles_per_block;
};

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
__device__ __forceinline__ void leaf_epilogue(const LeafArgs& a, uint64_t t, const Digest& d, uint32_t* region,
                                              uint32_t& run, uint32_t* s_wcnt) {
  const uint64_t c = t * kThreads + threadIdx.x;
  const bool valid = c < a.n;
  Digest prev{0, 0};
  if (valid && !a.baseline) prev = a.leaf_dg[c];
  const bool fixed = valid && d == prev;
  // A fixed leaf's digest was inserted by an earlier checkpoint of this baseline,
  // so its insert is a no-op except right after a reset, where every leaf inserts.
  const bool ins = valid && (a.baseline || !fixed);
  if (valid) {
    a.labels[c] = fixed ? L_FIXED : L_PENDING;
    if (ins) a.leaf_dg[c] = d;
  }
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
  if (ins) region[base + __popcll(m & ((1ull << lane) - 1))] = (uint32_t)c;
  run += tot;
  __syncthreads();  // s_wcnt is reused by the next tile
}

// One block streams a contiguous range of tiles (256 chunks each), software
// pipelined: the 16-byte loads of the next (tile, segment) are in flight while the
// current one is hashed out of LDS.  Chunks to insert are packed into the block's
// region of the list; the table work runs densely afterwards.
template <int HK, int CB>
__global__ __launch_bounds__(kThreads) void k_leaf(LeafArgs a) {
  __shared__ uint32_t s_wcnt[kThreads / 64];
  const uint64_t ntiles = (a.n + kThreads - 1) / kThreads;
  const uint64_t t0 = (uint64_t)blockIdx.x * a.tiles_per_block;
  const uint64_t t1 = std::min<uint64_t>(t0 + a.tiles_per_block, ntiles);
  uint32_t* region = a.list + t0 * kThreads;
  uint32_t run = 0;
  if constexpr (CB == 0) {
    for (uint64_t t = t0; t < t1; t++)
      leaf_epilogue(a, t, tile_digest<HK, 0>(a.data, a.len, a.n, a.C, t), region, run, s_wcnt);
  } else {
    using S = TileShape<CB>;
    __shared__ __align__(16) uint8_t lds[S::LDS];
    const uint64_t full = a.len / CB;
    const uint64_t steps = (t1 - t0) * S::PASSES;
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
        if (chunk < full) v[i] = *reinterpret_cast<const uint4*>(a.data + chunk * CB + pass * S::SEG + w);
      }
    };
    uint4 v[S::LOADS];
    if (steps) load(0, v);
    uint64_t h1 = 0, h2 = 0;
    uint32_t st[4], X[16];
    for (uint64_t step = 0; step < steps; step++) {
      uint4 nv[S::LOADS];
      if (step + 1 < steps) load(step + 1, nv);
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
        leaf_epilogue(a, t, d, region, run, s_wcnt);
      }
#pragma unroll
      for (int i = 0; i < S::LOADS; i++) v[i] = nv[i];
    }
  }
  if (threadIdx.x == 0) a.blk_cnt[blockIdx.x] = run;
}

// Concatenate the per-block regions into the dense list of chunks to insert.
__global__ __launch_bounds__(kThreads) void k_compact(const uint32_t* list, const uint32_t* blk_cnt,
                                                      const uint32_t* blk_off, uint64_t tiles_per_block,
                                                      uint32_t* dense) {
  const uint32_t b = blockIdx.x, cnt = blk_cnt[b];
  const uint32_t* src = list + (uint64_t)b * tiles_per_block * kThreads;
  uint32_t* dst = dense + blk_off[b];
  for (uint32_t i = threadIdx.x; i < cnt; i += kThreads) dst[i] = src[i];
}

struct DenseArgs {
  const uint32_t* dense;  // chunks to insert, ascending
  const uint32_t* total;  // their number
  uint32_t n;
  const Digest* leaf_dg;
  uint32_t* leaf_slot;
  uint8_t* labels;
  Slot* tab;
  uint64_t mask;
  uint32_t cur;
  Ctl* ctl;
  unsigned long long* ctr;
  Geo g;
  uint32_t R0;          // levels of the first sweep stage
  uint8_t* tile_first;  // per stage-0 tile: holds a first-occurrence leaf
  uint8_t* tile_busy;   // per stage-0 tile: holds a leaf that is not a fixed duplicate
};

// Stage-0 sweep tile of a chunk: bottom-level chunks by position, depth D-1 leaves
// by their position one level up.
__device__ __forceinline__ uint64_t chunk_tile(const Geo& g, uint32_t R0, uint64_t c) {
  const uint64_t lead = (1ull << g.D) - g.n;  // chunks stored at depth D-1
  if (c >= lead) return (c - lead) >> R0;
  return (c + g.half) >> (R0 - 1);
}

// Read-only pass before any insert of this checkpoint: a digest already in the
// table belongs to an earlier checkpoint, so its leaf is a shifted duplicate and
// needs no insert (include/reference_impl.hpp:51,66-67).  Plain loads, so popular
// digests (zero chunks) are served from cache.
__global__ __launch_bounds__(kThreads) void k_lookup(DenseArgs a) {
  const uint64_t total = *a.total;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kThreads) {
    const uint32_t c = a.dense[i];
    const uint32_t s = table_find(a.tab, a.mask, a.leaf_dg[c]);
    if (s == ~0u) continue;
    a.leaf_slot[c] = s;
    if (a.labels[c] == L_PENDING) {
      a.labels[c] = L_SHIFT;
      if (a.R0) a.tile_busy[chunk_tile(a.g, a.R0, c)] = 1;
    }
  }
}

// Table inserts (src/tree_approach.cpp:320-349; include/reference_impl.hpp:51): the
// packed value (cur << 32 | leaf) with atomicMin leaves the lowest leaf of a new
// digest in the table.  Consecutive entries with the same digest (zero-filled
// runs) share the first one's table access.
__global__ __launch_bounds__(kThreads) void k_insert(DenseArgs a) {
  const uint64_t total = *a.total;
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  uint32_t claims = 0;
  for (uint64_t wbase = (uint64_t)blockIdx.x * kThreads + (threadIdx.x & ~63u); wbase < total; wbase += stride) {
    const uint64_t i = wbase + lane;
    bool v = i < total;
    uint32_t c = 0;
    Digest d{0, 0};
    if (v) {
      c = a.dense[i];
      v = a.labels[c] != L_SHIFT;  // found by k_lookup: nothing to insert
      if (v) d = a.leaf_dg[c];
    }
    const uint64_t plo = __shfl_up(d.lo, 1), phi = __shfl_up(d.hi, 1);
    const int pv = __shfl_up((int)v, 1);
    const bool follower = v && lane > 0 && pv && plo == d.lo && phi == d.hi;
    const bool head = v && !follower;
    uint32_t slot = ~0u;
    bool claimed = false;
    if (head) slot = table_insert(a.tab, a.mask, d, ((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c), true,
                                  &a.ctl->err, &claimed, true);
    const uint64_t mine = __ballot(head) & ((lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1));
    const uint32_t hslot = __shfl(slot, mine ? 63 - __clzll((long long)mine) : (int)lane);
    if (follower) slot = hslot;
    if (v) a.leaf_slot[c] = slot;
    claims += claimed;
  }
  const uint32_t wc = wave_sum(claims);
  if (lane == 0 && wc) atomicAdd(stripe(a.ctr, CTR_CLAIMS), (unsigned long long)wc);
}

// Once every insert has landed: a pending leaf is FIRST when its digest's entry is
// (cur, itself), SHIFT otherwise (include/reference_impl.hpp:52-69).
__global__ __launch_bounds__(kThreads) void k_resolve(DenseArgs a) {
  const uint64_t total = *a.total;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (uint64_t)gridDim.x * kThreads) {
    const uint32_t c = a.dense[i];
    if (a.labels[c] != L_PENDING) continue;  // a fixed leaf right after a reset
    const uint32_t s = a.leaf_slot[c];
    uint8_t l = L_SHIFT;
    if (s <= a.mask + 1) l = (a.tab[s].val == (((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c))) ? L_FIRST : L_SHIFT;
    a.labels[c] = l;
    if (a.R0) {
      const uint64_t t = chunk_tile(a.g, a.R0, c);
      a.tile_busy[t] = 1;
      if (l == L_FIRST) a.tile_first[t] = 1;
    }
  }
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
struct SweepArgs {
  Geo g;
  uint32_t cur;
  uint32_t d_in, d_out;
  int stage0;
  const Digest* leaf_dg;
  uint8_t* labels;
  const uint32_t* leaf_slot;
  const uint8_t* in_lab;
  const Digest* in_dg;
  const uint32_t* in_slot;
  uint8_t* out_lab;
  Digest* out_dg;
  uint32_t* out_slot;
  Slot* tab;
  uint64_t mask;
  Ctl* ctl;
  unsigned long long* ctr;
  uint64_t* staging;
  uint32_t* cnt_f;
  uint32_t* cnt_s;
  const uint32_t* seg_base;
  Digest* dbg_tree;
  uint8_t* dbg_labels;
  const uint8_t* tile_first;  // stage 0: tile holds a first-occurrence leaf
  const uint8_t* tile_busy;   // stage 0: tile holds a leaf that is not a fixed duplicate
};

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

template <int HK>
__global__ __launch_bounds__(kThreads) void k_sweep_a(SweepArgs a) {
  const uint32_t lane = lane_id();
  const uint64_t tile = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (tile >= (1ull << a.d_out)) return;  // whole wavefront
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
  if (a.stage0 && !a.tile_first[tile]) {  // no first-occurrence leaf below this tile
    if (lane == 0) a.out_lab[tile] = L_NONE;
    return;
  }
  uint8_t l0[4];
  Digest g0[4];
  bool any = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t lp = 4 * lane + i;
    l0[i] = L_NONE;
    if (lp >= T) continue;
    const uint64_t p = base + lp;
    if (a.stage0) {
      if (p >= a.g.m) continue;
      const uint64_t c = bottom_chunk(a.g, p);
      l0[i] = a.labels[c];
      if (l0[i] == L_FIRST) g0[i] = a.leaf_dg[c];
    } else {
      l0[i] = a.in_lab[p];
      if (l0[i] == L_FIRST) g0[i] = a.in_dg[p];
    }
    any |= (l0[i] == L_FIRST);
  }
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
  __shared__ Digest rec_dg[kThreads / 64][256];
  __shared__ uint32_t rec_node[kThreads / 64][256];
  const uint32_t w = threadIdx.x >> 6;
  SweepA<HK> s{a, rec_dg[w], rec_node[w]};
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
      s.combine(a.d_in - 1, q, l0[2 * k], g0[2 * k], l0[2 * k + 1], g0[2 * k + 1], l1[k], g1[k]);
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
    table_insert(a.tab, a.mask, rec_dg[w][i], ((uint64_t)a.cur << 32) | rec_node[w][i], false, &a.ctl->err, &cl,
                 true);
    claims += cl;
  }
  const uint32_t wc = wave_sum(claims);
  if (lane == 0 && wc) atomicAdd(stripe(a.ctr, CTR_CLAIMS), (unsigned long long)wc);
}

// Sweep B: shifted-duplicate forest and region roots (include/reference_impl.hpp
// :83-106).  Read-only on the table.  Roots are staged per (depth, tile) in
// ascending node id; a shifted root carries the table slot of its digest.
template <int HK>
struct SweepB {
  const SweepArgs& a;
  uint64_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;  // up to four roots per lane and level, in order
  uint32_t ne = 0, nf = 0;

  __device__ void emit(uint8_t l, uint32_t slot, uint64_t node) {
    if (!is_region(l)) return;
    const uint64_t v = ((l == L_FIRST ? kFirstTag : (uint64_t)slot) << 32) | (uint32_t)node;
    e0 = ne == 0 ? v : e0;  // selects instead of an indexed array keep the entries in registers
    e1 = ne == 1 ? v : e1;
    e2 = ne == 2 ? v : e2;
    e3 = ne == 3 ? v : e3;
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

  // Stage this level's roots (children at depth c) for tile `tile`.
  __device__ void flush(uint32_t c, uint64_t tile) {
    uint32_t tot, ftot;
    const uint32_t rank = wave_rank(ne, &tot);
    wave_rank(nf, &ftot);
    uint64_t* dst = a.staging + heap_id(c, tile << (c - a.d_out)) + rank;
    if (ne > 0) dst[0] = e0;
    if (ne > 1) dst[1] = e1;
    if (ne > 2) dst[2] = e2;
    if (ne > 3) dst[3] = e3;
    if (lane_id() == 0 && tot) {  // counts are zero-filled before the sweep
      a.cnt_f[a.seg_base[c] + tile] = ftot;
      a.cnt_s[a.seg_base[c] + tile] = tot - ftot;
    }
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

template <int HK>
__global__ __launch_bounds__(kThreads) void k_sweep_b(SweepArgs a) {
  const uint32_t lane = lane_id();
  const uint64_t tile = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (tile >= (1ull << a.d_out)) return;  // whole wavefront
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
  if (a.stage0 && !a.tile_busy[tile]) {
    fixed_tile(a, tile, R, T, base, lane);
    return;
  }
  uint8_t l0[4];
  Digest g0[4];
  uint32_t s0[4];
  bool busy = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t lp = 4 * lane + i;
    l0[i] = L_NONE;
    if (lp >= T) continue;
    const uint64_t p = base + lp;
    if (a.stage0) {
      if (p >= a.g.m) continue;
      const uint64_t c = bottom_chunk(a.g, p);
      l0[i] = a.labels[c];
      if (l0[i] == L_SHIFT) {
        g0[i] = a.leaf_dg[c];
        s0[i] = a.leaf_slot[c];
      }
    } else {
      l0[i] = a.in_lab[p];
      if (l0[i] == L_SHIFT) {
        g0[i] = a.in_dg[p];
        s0[i] = a.in_slot[p];
      }
    }
    busy |= (l0[i] != L_FIXED);
  }
  uint8_t lf[2] = {L_NONE, L_NONE};
  Digest gf[2];
  uint32_t sf[2];
  if (a.stage0) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t j = 2 * lane + k;
      const uint64_t q = (base >> 1) + j;
      if (j >= T / 2 || q < a.g.half) continue;
      const uint64_t c = q - a.g.half;
      lf[k] = a.labels[c];
      if (lf[k] == L_SHIFT) {
        gf[k] = a.leaf_dg[c];
        sf[k] = a.leaf_slot[c];
      }
      busy |= (lf[k] != L_FIXED);
    }
  }
  if (!__ballot(busy)) {
    fixed_tile(a, tile, R, T, base, lane);
    return;
  }
  SweepB<HK> s{a};
  uint8_t l1[2] = {L_NONE, L_NONE};
  Digest g1[2];
  uint32_t s1[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const uint32_t j = 2 * lane + k;
    if (j >= T / 2) continue;
    const uint64_t q = (base >> 1) + j;
    if (a.stage0 && q >= a.g.half) {
      l1[k] = lf[k];
      g1[k] = gf[k];
      s1[k] = sf[k];
    } else {
      s.combine(a.d_in - 1, q, l0[2 * k], g0[2 * k], s0[2 * k], l0[2 * k + 1], g0[2 * k + 1], s0[2 * k + 1], l1[k],
                g1[k], s1[k]);
    }
  }
  s.flush(a.d_in, tile);
  uint8_t lc = l1[0];
  Digest gc = g1[0];
  uint32_t sc = s1[0];
  if (R >= 2) {
    lc = L_NONE;
    if (lane < T / 4)
      s.combine(a.d_in - 2, (base >> 2) + lane, l1[0], g1[0], s1[0], l1[1], g1[1], s1[1], lc, gc, sc);
    s.flush(a.d_in - 1, tile);
  }
  for (uint32_t r = 3; r <= R; r++) {
    const int sh = 1 << (r - 3);
    const uint8_t lp = (uint8_t)__shfl_down((int)lc, sh);
    const Digest gp = shfl_down_dg(gc, sh);
    const uint32_t sp = __shfl_down(sc, sh);
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
    s.flush(a.d_in - r + 1, tile);
  }
  if (lane == 0) {
    a.out_lab[tile] = lc;
    if (lc == L_SHIFT) {
      a.out_dg[tile] = gc;
      a.out_slot[tile] = sc;
    }
    if (a.d_out == 0) {
      // Node 0: a first-occurrence root (:105-106); a shifted root too, which the
      // reference omits and so cannot restart (documented departure).
      const bool root = is_region(lc);
      if (root) a.staging[0] = ((lc == L_FIRST ? kFirstTag : (uint64_t)sc) << 32) | 0u;
      a.cnt_f[a.seg_base[0]] = (root && lc == L_FIRST) ? 1u : 0u;
      a.cnt_s[a.seg_base[0]] = (root && lc == L_SHIFT) ? 1u : 0u;
    }
  }
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

// A single chunk: the leaf is the root.
__global__ void k_single_leaf(SweepArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint8_t l = a.labels[0];
  const bool root = is_region(l);
  if (root) a.staging[0] = (l == L_FIRST ? (kFirstTag << 32) : ((uint64_t)a.leaf_slot[0] << 32)) | 0u;
  a.cnt_f[0] = (root && l == L_FIRST) ? 1u : 0u;
  a.cnt_s[0] = (root && l == L_SHIFT) ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// Root compaction: one thread per (depth, block) segment, in segment order.
// ---------------------------------------------------------------------------
__global__ void k_scatter_roots(const uint64_t* staging, const uint32_t* cnt_f, const uint32_t* cnt_s,
                                const uint32_t* off_f, const uint32_t* off_s, const uint32_t* seg_base,
                                uint32_t ndepth, uint32_t nseg, uint32_t* first_roots, uint64_t* shift_tmp) {
  const uint32_t seg = blockIdx.x * blockDim.x + threadIdx.x;
  if (seg >= nseg) return;
  const uint32_t nf = cnt_f[seg], ns = cnt_s[seg];
  if (nf + ns == 0) return;
  uint32_t c = 0;
  while (c + 1 < ndepth && seg_base[c + 1] <= seg) c++;
  const uint32_t nb = seg_base[c + 1] - seg_base[c];  // blocks of the emitting stage (a power of two)
  const uint32_t span = c - (31u - (uint32_t)__clz((int)nb));
  const uint64_t sb = ((1ull << c) - 1) + ((uint64_t)(seg - seg_base[c]) << span);
  uint32_t of = off_f[seg], os = off_s[seg];
  for (uint32_t j = 0; j < nf + ns; j++) {
    const uint64_t e = staging[sb + j];
    if ((e >> 32) == kFirstTag) first_roots[of++] = (uint32_t)e;
    else shift_tmp[os++] = e;
  }
}

// ---------------------------------------------------------------------------
// Shifted entries: source checkpoint and previous node from the table
// (src/tree_approach.cpp:574-580,670-696), then a stable sort by source.
// ---------------------------------------------------------------------------
constexpr uint32_t kLdsBins = 4096;  // source-checkpoint bins counted in LDS

__global__ __launch_bounds__(kThreads) void k_shift_src(const uint64_t* shift_tmp, uint32_t S, const Slot* tab,
                                                        uint64_t mask, uint32_t ref, uint32_t cur, uint32_t* key,
                                                        uint64_t* val, uint32_t* src_hist, uint32_t nbins, Ctl* ctl) {
  __shared__ uint32_t h[kLdsBins];
  const bool lds = nbins <= kLdsBins;
  if (lds)
    for (uint32_t b = threadIdx.x; b < nbins; b += kThreads) h[b] = 0;
  __syncthreads();
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < S) {
    const uint64_t e = shift_tmp[i];
    const uint32_t node = (uint32_t)e, slot = (uint32_t)(e >> 32);
    uint32_t k = 0, prev = 0;
    if (slot > mask + 1) {
      atomicOr(&ctl->err, ERR_CORRUPT);
    } else {
      const uint64_t v = tab[slot].val;
      const uint32_t src = (uint32_t)(v >> 32);
      prev = (uint32_t)v;
      k = src - ref;
      if (src < ref || src > cur) {
        atomicOr(&ctl->err, ERR_CORRUPT);
        k = 0;
      }
    }
    key[i] = k;
    val[i] = ((uint64_t)prev << 32) | node;
    if (lds) atomicAdd(&h[k], 1u);
    else atomicAdd(&src_hist[k], 1u);
  }
  __syncthreads();
  if (lds)
    for (uint32_t b = threadIdx.x; b < nbins; b += kThreads)
      if (h[b]) atomicAdd(&src_hist[b], h[b]);
}

// Histogram of one 8-bit digit over a 256-entry segment.
__global__ __launch_bounds__(kThreads) void k_sort_hist(const uint32_t* key, uint32_t S, uint32_t shift,
                                                        uint32_t nbins, uint32_t nseg, uint32_t* hist) {
  __shared__ uint32_t h[256];
  const uint32_t seg = blockIdx.x, tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const uint32_t i = seg * kThreads + tid;
  if (i < S) atomicAdd(&h[(key[i] >> shift) & (nbins - 1)], 1u);
  __syncthreads();
  if (tid < nbins) hist[tid * nseg + seg] = h[tid];
}

// Stable scatter by digit: rank within the segment = earlier lanes of this wave
// with the same digit + the same digit in earlier waves.
__global__ __launch_bounds__(kThreads) void k_sort_scatter(const uint32_t* key, const uint64_t* val, uint32_t S,
                                                           uint32_t shift, uint32_t nbins, uint32_t nseg,
                                                           const uint32_t* hist_off, uint32_t* key_out,
                                                           uint64_t* val_out) {
  __shared__ uint32_t cnt[kThreads / 64][256];
  const uint32_t seg = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (uint32_t j = tid; j < (kThreads / 64) * 256; j += kThreads) (&cnt[0][0])[j] = 0;
  __syncthreads();
  const uint32_t i = seg * kThreads + tid;
  const bool valid = i < S;
  const uint32_t k = valid ? key[i] : 0u;
  const uint32_t dg = (k >> shift) & (nbins - 1);
  bool done = !valid;
  uint32_t rank = 0;
  const uint64_t lt = (1ull << lane) - 1;
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
  if (!valid) return;
  uint32_t pos = hist_off[dg * nseg + seg] + rank;
  for (uint32_t ww = 0; ww < w; ww++) pos += cnt[ww][dg];
  key_out[pos] = k;
  val_out[pos] = val[i];
}

// (chkpt, count) table in ascending checkpoint id (src/tree_approach.cpp:657-665).
__global__ __launch_bounds__(kThreads) void k_ptable(const uint32_t* src_hist, uint32_t nbins, uint32_t ref,
                                                     uint32_t* ptable, Ctl* ctl) {
  __shared__ uint32_t s_ws[kThreads / 64];
  uint32_t base = 0;
  for (uint32_t b0 = 0; b0 < nbins; b0 += kThreads) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t cnt = b < nbins ? src_hist[b] : 0u;
    uint32_t tot;
    const uint32_t r = block_scan_excl(cnt > 0, &tot, s_ws);
    if (cnt > 0) {
      ptable[2 * (base + r)] = ref + b;
      ptable[2 * (base + r) + 1] = cnt;
    }
    base += tot;
  }
  if (threadIdx.x == 0) ctl->P = base;
}

__global__ __launch_bounds__(kThreads) void k_region_sizes(const uint32_t* first_roots, uint32_t F, Geo g,
                                                           uint32_t* sizes) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < F) sizes[i] = leaf_count(g, first_roots[i]);
}

// ---------------------------------------------------------------------------
// Diff assembly (src/tree_approach.cpp:593,622-706): header, F roots, P table,
// S pairs, then the gathered chunks.
// ---------------------------------------------------------------------------
struct DiffArgs {
  uint8_t* diff;
  const uint32_t* first_roots;
  const uint32_t* ptable;
  const uint64_t* shift_sorted;  // (prev << 32 | node)
  const uint32_t* region_off;
  const uint32_t* total_chunks;
  const uint8_t* data;
  uint64_t len;
  uint32_t n, C;
  uint32_t ref_id, chkpt_id;
  Geo g;
  Ctl* ctl;
  int aligned;
};

__device__ __forceinline__ void put32(uint8_t* p, uint32_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
  p[2] = (uint8_t)(v >> 16);
  p[3] = (uint8_t)(v >> 24);
}

__global__ __launch_bounds__(kThreads) void k_diff_meta(DiffArgs a) {
  const uint32_t F = a.ctl->F, P = a.ctl->P, S = a.ctl->S;
  uint32_t* out = reinterpret_cast<uint32_t*>(a.diff);  // every metadata field is 4-byte aligned
  const uint64_t oF = 8, oP = oF + F, oS = oP + 2ull * P;
  const uint64_t tid = (uint64_t)blockIdx.x * kThreads + threadIdx.x, stride = (uint64_t)gridDim.x * kThreads;
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
    for (int w = 0; w < 8; w++) out[w] = hw[w];
    a.ctl->data_chunks = chunks;
    a.ctl->diff_size = data_off + chunks * a.C;
  }
}

typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

// Gather the chunks of every first-occurrence region, regions in root order and
// leaves in ascending chunk order; the short last chunk is zero padded (:636-650).
__global__ __launch_bounds__(kThreads) void k_gather(DiffArgs a) {
  __shared__ uint32_t s_src[kThreads];
  const uint32_t F = a.ctl->F;
  const uint64_t total = *a.total_chunks;
  const uint64_t data_off = 32 + 4ull * F + 8ull * a.ctl->P + 8ull * a.ctl->S;
  uint8_t* dst0 = a.diff + data_off;
  const uint32_t C = a.C;
  const uint32_t U = a.aligned ? C / 16 : C;  // copy units per chunk
  for (uint64_t tile = blockIdx.x; tile * kThreads < total; tile += gridDim.x) {
    const uint64_t j0 = tile * kThreads;
    const uint64_t j = j0 + threadIdx.x;
    if (j < total) {
      uint32_t lo = 0, hi = F;  // last region whose offset <= j
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.region_off[mid] <= j) lo = mid;
        else hi = mid;
      }
      uint32_t st[2], ct[2];
      const int r = leaf_runs(a.g, a.first_roots[lo], st, ct);
      s_src[threadIdx.x] = run_at(r, st, ct, (uint32_t)(j - a.region_off[lo]));
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
          *reinterpret_cast<u32x4_a4*>(dst + 16ull * w) = u32x4_a4{v.x, v.y, v.z, v.w};
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

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int HK>
static void launch_leaf(const LeafArgs& a, bool fast, hipStream_t st) {
  const uint64_t ntiles = (a.n + kThreads - 1) / kThreads;
  const uint32_t grid = (uint32_t)((ntiles + a.tiles_per_block - 1) / a.tiles_per_block);
  if (fast) {
    switch (a.C) {
      case 16: k_leaf<HK, 16><<<grid, kThreads, 0, st>>>(a); return;
      case 32: k_leaf<HK, 32><<<grid, kThreads, 0, st>>>(a); return;
      case 64: k_leaf<HK, 64><<<grid, kThreads, 0, st>>>(a); return;
      case 128: k_leaf<HK, 128><<<grid, kThreads, 0, st>>>(a); return;
      case 256: k_leaf<HK, 256><<<grid, kThreads, 0, st>>>(a); return;
      case 512: k_leaf<HK, 512><<<grid, kThreads, 0, st>>>(a); return;
      case 1024: k_leaf<HK, 1024><<<grid, kThreads, 0, st>>>(a); return;
      case 2048: k_leaf<HK, 2048><<<grid, kThreads, 0, st>>>(a); return;
      case 4096: k_leaf<HK, 4096><<<grid, kThreads, 0, st>>>(a); return;
      default: break;
    }
  }
  k_leaf<HK, 0><<<grid, kThreads, 0, st>>>(a);
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

static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

static void fill_table(Slot* t, uint64_t slots, hipStream_t st) {
  const uint64_t words = slots * 4;
  k_fill_u64<<<grid_for(words, kThreads, 65536), kThreads, 0, st>>>(reinterpret_cast<uint64_t*>(t), words, TEMPTY);
  MTC_HIP(hipGetLastError());
}

// Plan sweep stages and root segments for n chunks.
static void plan_tree(mtc_dedup* h) {
  const Geo g = make_geo(h->n);
  h->stages.clear();
  std::vector<uint32_t> nb(g.D + 1, 1);  // blocks of the stage emitting roots at depth c
  if (h->n > 1) {
    uint32_t d = g.D;
    while (d > 0) {
      const uint32_t dout = d > kFuse ? d - kFuse : 0;
      h->stages.push_back(Stage{d, dout});
      for (uint32_t c = dout + 1; c <= d; c++) nb[c] = 1u << dout;
      d = dout;
    }
  }
  h->seg_base_h.assign(g.D + 2, 0);
  for (uint32_t c = 0; c <= g.D; c++) h->seg_base_h[c + 1] = h->seg_base_h[c] + nb[c];
  h->plan_dirty = true;
}

static void ensure_buffers(mtc_dedup* h, hipStream_t st) {
  const uint64_t n = h->n, nodes = 2 * n - 1;
  h->leaf_dg.ensure(n);
  h->labels.ensure(n);
  h->leaf_slot.ensure(n);
  const Geo g = make_geo(h->n);
  const uint64_t half_cap = (g.D > kFuse) ? (1ull << (g.D - kFuse)) : 1;
  h->fr_lab.ensure(2 * half_cap);
  h->fr_dg.ensure(2 * half_cap);
  h->fr_slot.ensure(2 * half_cap);
  h->staging.ensure(nodes);
  const uint32_t nseg = h->seg_base_h.back();
  h->seg_cnt_f.ensure(nseg);
  h->seg_cnt_s.ensure(nseg);
  h->seg_off_f.ensure(nseg + 1);
  h->seg_off_s.ensure(nseg + 1);
  h->seg_base.ensure(h->seg_base_h.size());
  if (h->plan_dirty) {
    MTC_HIP(hipMemcpyAsync(h->seg_base.p, h->seg_base_h.data(), h->seg_base_h.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st));
    MTC_HIP(hipStreamSynchronize(st));  // the host vector is pageable
    h->plan_dirty = false;
  }
  h->list.ensure(n);
  h->dense.ensure(n);
  const uint64_t max_blk = (n + kThreads - 1) / kThreads;  // at most one block per tile
  h->blk_cnt.ensure(max_blk);
  h->blk_off.ensure(max_blk + 1);
  h->tile_flags.ensure(2 * ((g.D > kFuse) ? (1ull << (g.D - kFuse)) : 1) + 4);
  h->first_roots.ensure(n);
  h->shift_tmp.ensure(n);
  h->region_off.ensure(n + 1);
  // P <= checkpoints since the baseline; reserve in steps of 4096 so the arena is
  // not reallocated every checkpoint.
  const uint64_t pcap = ((uint64_t)h->current_id + 2 + 4095) / 4096 * 4096;
  h->diff.ensure(32 + 12 * n + 8 * pcap + n * (uint64_t)h->chunk_size);
}

static void ensure_table(mtc_dedup* h, bool reset, hipStream_t st) {
  const uint64_t nodes = 2ull * h->n - 1;
  if (reset) {
    const uint64_t cap = next_pow2(std::max<uint64_t>(4 * (uint64_t)h->n, 1024));
    h->table.ensure(cap + 1);
    h->table_cap = cap;
    h->table_size = 0;
    fill_table(h->table.p, cap + 1, st);
    return;
  }
  // Every checkpoint can add at most 2n-1 entries; keep room for all of them, so
  // probing always ends (the reference rehashes to size + num_nodes,
  // src/tree_approach.cpp:1622-1623), and grow early once the table is 80% full.
  auto fits = [&](uint64_t cap) { return h->table_size + nodes < cap && h->table_size <= cap / 10 * 8; };
  if (fits(h->table_cap)) return;
  uint64_t cap = h->table_cap;
  while (!fits(cap)) cap *= 2;
  if (cap > (1ull << 31)) throw Status{MTC_ERR_TABLE_FULL, "table capacity above 2^31 slots"};
  h->table_tmp.ensure(cap + 1);
  fill_table(h->table_tmp.p, cap + 1, st);
  k_rehash<<<grid_for(h->table_cap + 1, kThreads, 65536), kThreads, 0, st>>>(h->table.p, h->table_cap + 1,
                                                                             h->table_tmp.p, cap - 1, h->ctl_d);
  MTC_HIP(hipGetLastError());
  std::swap(h->table, h->table_tmp);
  h->table_tmp.release();
  h->table_cap = cap;
}

static void sync_ctl(mtc_dedup* h, hipStream_t st) {
  MTC_HIP(hipMemcpyAsync(h->ctl_h, h->ctl_d, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  MTC_HIP(hipStreamSynchronize(st));
  if (h->ctl_h->err & (ERR_TABLE_FULL | ERR_TABLE_SPIN | ERR_OVERFLOW_SLOT))
    throw Status{MTC_ERR_TABLE_FULL, "digest table insert failed (err bits " + std::to_string(h->ctl_h->err) + ")"};
  if (h->ctl_h->err) throw Status{MTC_ERR_STATE, "inconsistent table state (err bits " + std::to_string(h->ctl_h->err) + ")"};
}

template <int HK>
static void run_sweeps(mtc_dedup* h, SweepArgs sa, hipStream_t st) {
  const Geo g = make_geo(h->n);
  const uint64_t half_cap = (g.D > kFuse) ? (1ull << (g.D - kFuse)) : 1;
  if (h->n == 1) {
    k_single_leaf<<<1, 64, 0, st>>>(sa);
    MTC_HIP(hipGetLastError());
    return;
  }
  // Sweep A over all stages, its inserts, then sweep B: every first-occurrence
  // insert must land before any shifted-duplicate lookup (reference_impl.hpp:74-104).
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      const uint32_t nseg = h->seg_base_h.back();
      k_fill_u32<<<grid_for(nseg, kThreads, 4096), kThreads, 0, st>>>(h->seg_cnt_f.p, nseg, 0);
      k_fill_u32<<<grid_for(nseg, kThreads, 4096), kThreads, 0, st>>>(h->seg_cnt_s.p, nseg, 0);
    }
    for (size_t s = 0; s < h->stages.size(); s++) {
      SweepArgs a = sa;
      a.d_in = h->stages[s].d_in;
      a.d_out = h->stages[s].d_out;
      a.stage0 = (s == 0);

      const uint64_t in_off = (s % 2 == 1) ? 0 : half_cap, out_off = (s % 2 == 0) ? 0 : half_cap;
      a.in_lab = h->fr_lab.p + in_off;
      a.in_dg = h->fr_dg.p + in_off;
      a.in_slot = h->fr_slot.p + in_off;
      a.out_lab = h->fr_lab.p + out_off;
      a.out_dg = h->fr_dg.p + out_off;
      a.out_slot = h->fr_slot.p + out_off;
      const uint32_t blocks = (uint32_t)(((1ull << a.d_out) + 3) / 4);  // four wavefront tiles per block
      if (pass == 0) k_sweep_a<HK><<<blocks, kThreads, 0, st>>>(a);
      else k_sweep_b<HK><<<blocks, kThreads, 0, st>>>(a);
      MTC_HIP(hipGetLastError());
    }
  }
}

// Host-side phase timer, printed per checkpoint when MTC_HOST_TRACE is set.
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
  void print() {
    if (on) std::fprintf(stderr, "[mtc host] %stotal=%.1fus\n", line.c_str(),
                         std::chrono::duration<double, std::micro>(last - t0).count());
  }
};

void checkpoint_impl(mtc_dedup* h, const uint8_t* data, uint64_t len, bool make_baseline, hipStream_t st) {
  HostTrace ht;
  if (!h->ev[0])
    for (auto& e : h->ev) MTC_HIP(hipEventCreate(&e));
  hipEvent_t* ev = h->ev;
  MTC_HIP(hipEventRecord(ev[0], st));
  const uint64_t C = h->chunk_size;
  const uint64_t n64 = (len + C - 1) / C;
  if (n64 > (1ull << 31)) throw Status{MTC_ERR_INVALID_ARG, "more than 2^31 chunks"};
  const bool baseline = make_baseline || h->current_id == 0 || !h->has_state;
  if (!baseline && len != h->data_len)
    throw Status{MTC_ERR_STATE, "incremental checkpoint with a different length than the previous one"};
  if (baseline) {
    h->n = (uint32_t)n64;
    h->data_len = len;
    plan_tree(h);
  }
  ensure_buffers(h, st);
  h->ctr.ensure(CTR_KINDS * kStripes);
  k_zero_ctl<<<1, kThreads, 0, st>>>(h->ctl_d, h->ctr.p);
  ensure_table(h, baseline, st);
  if (baseline) h->baseline_id = h->current_id;
  h->has_state = true;
  const uint32_t n = h->n, cur = h->current_id;
  const Geo g = make_geo(n);
  if (h->debug) {
    h->dbg_tree.ensure(2ull * n - 1);
    h->dbg_labels.ensure(2ull * n - 1);
    k_fill_u64<<<grid_for(4ull * n, kThreads, 65536), kThreads, 0, st>>>(reinterpret_cast<uint64_t*>(h->dbg_tree.p),
                                                                        2 * (2ull * n - 1), 0);
  }
  MTC_HIP(hipEventRecord(ev[1], st));
  ht.mark("setup");

  // The leaf kernel runs one contiguous range of tiles per block (kLeafBlocks blocks).
  const uint64_t ntiles_leaf = (n + kThreads - 1) / kThreads;
  static const uint64_t leaf_blocks = [] {
    const char* e = std::getenv("MTC_LEAF_BLOCKS");  // tuning knob for the persistent grid
    return (e && std::atoll(e) > 0) ? (uint64_t)std::atoll(e) : kLeafBlocks;
  }();
  const uint64_t tpb = (ntiles_leaf + leaf_blocks - 1) / leaf_blocks;
  const uint32_t nblk = (uint32_t)((ntiles_leaf + tpb - 1) / tpb);
  LeafArgs la{data, len, n, h->chunk_size, h->leaf_dg.p, h->labels.p, baseline ? 1 : 0, h->list.p, h->blk_cnt.p, tpb};
  const bool fast = (C % 16 == 0) && ((uintptr_t)data % 16 == 0);
  if (h->hash_kind == HASH_MD5) launch_leaf<HASH_MD5>(la, fast, st);
  else launch_leaf<HASH_MURMUR3>(la, fast, st);
  MTC_HIP(hipGetLastError());
  MTC_HIP(hipEventRecord(ev[4], st));

  // Dense table inserts over the changed chunks, then FIRST/SHIFT resolution.
  exclusive_scan_u32(h->blk_cnt.p, h->blk_off.p, nblk, h->blk_off.p + nblk, h->ws, st);
  k_compact<<<nblk, kThreads, 0, st>>>(h->list.p, h->blk_cnt.p, h->blk_off.p, tpb, h->dense.p);
  MTC_HIP(hipGetLastError());
  const uint32_t R0 = h->stages.empty() ? 0 : h->stages[0].d_in - h->stages[0].d_out;
  const uint64_t ntile0 = h->stages.empty() ? 1 : (1ull << h->stages[0].d_out);
  k_fill_u32<<<grid_for((2 * ntile0 + 3) / 4), kThreads, 0, st>>>(reinterpret_cast<uint32_t*>(h->tile_flags.p),
                                                                 (2 * ntile0 + 3) / 4, 0);
  DenseArgs ia{h->dense.p, h->blk_off.p + nblk, n, h->leaf_dg.p, h->leaf_slot.p, h->labels.p, h->table.p,
               h->table_cap - 1, cur, h->ctl_d, h->ctr.p, g, R0, h->tile_flags.p, h->tile_flags.p + ntile0};
  const uint32_t dense_grid = grid_for(n, kThreads, 16384);
  if (!baseline) {  // a reset table holds nothing to find
    k_lookup<<<dense_grid, kThreads, 0, st>>>(ia);
    MTC_HIP(hipGetLastError());
  }
  k_insert<<<dense_grid, kThreads, 0, st>>>(ia);
  MTC_HIP(hipGetLastError());
  k_resolve<<<dense_grid, kThreads, 0, st>>>(ia);
  MTC_HIP(hipGetLastError());

  SweepArgs sa{};
  sa.g = g;
  sa.cur = cur;
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
  sa.tile_busy = h->tile_flags.p + ntile0;
  if (h->hash_kind == HASH_MD5) run_sweeps<HASH_MD5>(h, sa, st);
  else run_sweeps<HASH_MURMUR3>(h, sa, st);
  ht.mark("sweeps launched");
  if (h->debug) {  // leaf digests and labels into the debug views
    MTC_HIP(hipMemcpyAsync(h->dbg_tree.p + (n - 1), h->leaf_dg.p, 16ull * n, hipMemcpyDeviceToDevice, st));
    MTC_HIP(hipMemcpyAsync(h->dbg_labels.p + (n - 1), h->labels.p, n, hipMemcpyDeviceToDevice, st));
  }

  // Root totals, then order them: segments are (depth, block) in node-id order.
  const uint32_t nseg = h->seg_base_h.back();
  Ctl* cd = h->ctl_d;
  exclusive_scan_u32(h->seg_cnt_f.p, h->seg_off_f.p, nseg, &cd->F, h->ws, st);
  exclusive_scan_u32(h->seg_cnt_s.p, h->seg_off_s.p, nseg, &cd->S, h->ws, st);
  MTC_HIP(hipEventRecord(ev[2], st));
  sync_ctl(h, st);
  ht.mark("sync 1");
  const uint32_t F = h->ctl_h->F, S = h->ctl_h->S;
  k_scatter_roots<<<grid_for(nseg), kThreads, 0, st>>>(h->staging.p, h->seg_cnt_f.p, h->seg_cnt_s.p, h->seg_off_f.p,
                                                       h->seg_off_s.p, h->seg_base.p, g.D + 1, nseg,
                                                       h->first_roots.p, h->shift_tmp.p);
  MTC_HIP(hipGetLastError());

  // Shifted entries grouped by source checkpoint (stable, so ascending node id inside).
  const uint32_t nbins = cur - h->baseline_id + 1;
  h->src_hist.ensure(nbins);
  h->ptable.ensure(2ull * nbins);
  k_fill_u32<<<grid_for(nbins), kThreads, 0, st>>>(h->src_hist.p, nbins, 0);
  uint64_t* sorted = nullptr;
  if (S > 0) {
    for (int b = 0; b < 2; b++) {
      h->sort_key[b].ensure(S);
      h->sort_val[b].ensure(S);
    }
    k_shift_src<<<grid_for(S), kThreads, 0, st>>>(h->shift_tmp.p, S, h->table.p, h->table_cap - 1, h->baseline_id, cur,
                                                 h->sort_key[0].p, h->sort_val[0].p, h->src_hist.p, nbins, cd);
    MTC_HIP(hipGetLastError());
    uint32_t bits = 0;
    while ((1ull << bits) < nbins) bits++;
    const uint32_t nsg = (S + kThreads - 1) / kThreads;
    h->sort_hist.ensure(256ull * nsg);
    int cur_buf = 0;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
      const uint32_t db = std::min<uint32_t>(8, bits - shift);
      const uint32_t nb = 1u << db;
      k_sort_hist<<<nsg, kThreads, 0, st>>>(h->sort_key[cur_buf].p, S, shift, nb, nsg, h->sort_hist.p);
      MTC_HIP(hipGetLastError());
      exclusive_scan_u32(h->sort_hist.p, h->sort_hist.p, (uint64_t)nb * nsg, nullptr, h->ws, st);
      k_sort_scatter<<<nsg, kThreads, 0, st>>>(h->sort_key[cur_buf].p, h->sort_val[cur_buf].p, S, shift, nb, nsg,
                                              h->sort_hist.p, h->sort_key[1 - cur_buf].p, h->sort_val[1 - cur_buf].p);
      MTC_HIP(hipGetLastError());
      cur_buf = 1 - cur_buf;
    }
    sorted = h->sort_val[cur_buf].p;
  }
  k_ptable<<<1, kThreads, 0, st>>>(h->src_hist.p, nbins, h->baseline_id, h->ptable.p, cd);
  MTC_HIP(hipGetLastError());

  // Region sizes and data offsets.
  if (F > 0) {
    k_region_sizes<<<grid_for(F), kThreads, 0, st>>>(h->first_roots.p, F, g, h->region_off.p);
    MTC_HIP(hipGetLastError());
  }
  exclusive_scan_u32(h->region_off.p, h->region_off.p, F, h->region_off.p + n, h->ws, st);

  DiffArgs da{};
  da.diff = h->diff.p;
  da.first_roots = h->first_roots.p;
  da.ptable = h->ptable.p;
  da.shift_sorted = sorted;
  da.region_off = h->region_off.p;
  da.total_chunks = h->region_off.p + n;
  da.data = data;
  da.len = len;
  da.n = n;
  da.C = h->chunk_size;
  da.ref_id = h->baseline_id;
  da.chkpt_id = cur;
  da.g = g;
  da.ctl = cd;
  da.aligned = fast ? 1 : 0;
  k_diff_meta<<<grid_for(std::max<uint64_t>(std::max(F, S), 2ull * nbins), kThreads, 4096), kThreads, 0, st>>>(da);
  MTC_HIP(hipGetLastError());
  if (F > 0) {
    k_gather<<<grid_for(n, kThreads, 8192), kThreads, 0, st>>>(da);
    MTC_HIP(hipGetLastError());
  }
  k_reduce_ctr<<<1, kStripes, 0, st>>>(h->ctr.p, h->blk_off.p + nblk, cd);
  MTC_HIP(hipEventRecord(ev[3], st));
  sync_ctl(h, st);
  ht.mark("final sync");
  h->table_size += h->ctl_h->claims;
  h->last_diff_size = h->ctl_h->diff_size;
  h->last_header = mtc_header{h->baseline_id, cur, len, h->chunk_size, h->ctl_h->F, h->ctl_h->P, h->ctl_h->S};
  MTC_HIP(hipEventElapsedTime(&h->last_ms[0], ev[0], ev[1]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[1], ev[1], ev[2]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[2], ev[2], ev[3]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[4], ev[1], ev[4]));
  h->last_ms[3] = 0;
  ht.mark("epilogue");
  ht.print();
}

}  // namespace mtc
