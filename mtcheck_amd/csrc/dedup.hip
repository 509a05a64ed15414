// dedup.hip -- the checkpoint path: leaf hashing and labelling, the two fused
// Merkle-forest sweeps, root compaction in node-id order, and diff assembly.
//
// Serial semantics followed: include/reference_impl.hpp:15-132 (labels and region
// roots) and src/tree_approach.cpp:511-710 (diff layout), in the canonical order
// of the serial reference (roots ascending; shifted entries grouped by source
// checkpoint, ascending node id inside a group).
#include "internal.hpp"

namespace mtc {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kFuse = 8;              // tree levels fused per sweep stage (one wavefront tile)
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

__global__ void k_reduce_ctr(const unsigned long long* ctr, Ctl* c) {
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
    c->changed = s[CTR_CHANGED];
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
  uint32_t* leaf_slot;
  Slot* tab;
  uint64_t mask;
  uint32_t cur;
  int baseline;
  Ctl* ctl;
  unsigned long long* ctr;
};

// Digest of chunk blockIdx.x*256 + threadIdx.x (zero beyond n).  CB > 0 selects
// the LDS-staged path for 16-byte-aligned data and chunk sizes that are multiples
// of 16; a short last chunk and CB == 0 use byte-granular hashing.
template <int HK, int CB>
__device__ __forceinline__ Digest tile_digest(const uint8_t* data, uint64_t len, uint32_t n, uint32_t C) {
  const uint64_t first = (uint64_t)blockIdx.x * kThreads;
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

template <int HK, int CB>
__global__ __launch_bounds__(kThreads) void k_leaf(LeafArgs a) {
  const uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const bool valid = c < a.n;
  const Digest d = tile_digest<HK, CB>(a.data, a.len, a.n, a.C);
  Digest prev{0, 0};
  if (valid && !a.baseline) prev = a.leaf_dg[c];
  const bool fixed = valid && d == prev;
  // A fixed leaf's digest was inserted by an earlier checkpoint of this baseline,
  // so its insert is a no-op except right after a reset, where every leaf inserts.
  const bool ins = valid && (a.baseline || !fixed);
  // Lanes holding the same digest as the lane before them (zero-filled runs) reuse
  // that lane's table slot: the lowest leaf of a run carries the minimum anyway.
  const uint32_t lane = lane_id();
  uint64_t plo = __shfl_up(d.lo, 1), phi = __shfl_up(d.hi, 1);
  int pins = __shfl_up((int)ins, 1);
  const bool follower = ins && lane > 0 && pins && plo == d.lo && phi == d.hi;
  const bool head = ins && !follower;
  uint32_t slot = ~0u;
  bool claimed = false;
  if (head) {
    const uint64_t v = ((uint64_t)a.cur << 32) | (uint32_t)(a.n - 1 + c);
    slot = table_insert(a.tab, a.mask, d, v, true, &a.ctl->err, &claimed);
  }
  const uint64_t heads = __ballot(head);
  const uint64_t le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1);
  const uint64_t mine = heads & le;
  const int src_lane = mine ? 63 - __clzll((long long)mine) : (int)lane;
  const uint32_t hslot = __shfl(slot, src_lane);
  if (follower) slot = hslot;
  wave_count_add(stripe(a.ctr, CTR_CLAIMS), claimed);
  wave_count_add(stripe(a.ctr, CTR_CHANGED), valid && !fixed);
  if (valid) {
    a.labels[c] = fixed ? L_FIXED : L_PENDING;
    if (ins) {
      a.leaf_dg[c] = d;
      a.leaf_slot[c] = slot;
    }
  }
}

// Standalone chunk hashing (mtc_hash_chunks).
template <int HK, int CB>
__global__ __launch_bounds__(kThreads) void k_hash_chunks(const uint8_t* data, uint64_t len, uint32_t n, uint32_t C,
                                                          Digest* out) {
  const uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const Digest d = tile_digest<HK, CB>(data, len, n, C);
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
};

// FIRST if this leaf holds the table entry of its digest (lowest leaf of a digest
// new to this checkpoint), SHIFT otherwise.
__device__ __forceinline__ uint8_t resolve_leaf(const SweepArgs& a, uint64_t c) {
  const uint32_t s = a.leaf_slot[c];
  if (s > a.mask + 1) return L_SHIFT;  // insert failed; error already raised
  const uint64_t v = a.tab[s].val;
  const uint64_t mine = ((uint64_t)a.cur << 32) | (uint32_t)(a.g.n - 1 + c);
  return v == mine ? L_FIRST : L_SHIFT;
}

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
  uint32_t claims = 0;

  __device__ void combine(uint32_t depth, uint64_t pos, uint8_t la, const Digest& da, uint8_t lb, const Digest& db,
                          uint8_t& lo, Digest& dout) {
    lo = L_NONE;
    if (la == L_FIRST && lb == L_FIRST) {
      dout = hash_pair<HK>(da, db);
      const uint32_t node = (uint32_t)heap_id(depth, pos);
      bool cl = false;
      table_insert(a.tab, a.mask, dout, ((uint64_t)a.cur << 32) | node, false, &a.ctl->err, &cl);
      claims += cl;
      lo = L_FIRST;
      if (a.dbg_tree) a.dbg_tree[node] = dout;
    }
  }
};

template <int HK>
__global__ __launch_bounds__(kThreads) void k_sweep_a(SweepArgs a) {
  const uint32_t lane = lane_id();
  const uint64_t tile = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (tile >= (1ull << a.d_out)) return;  // whole wavefront
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
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
      uint8_t l = a.labels[c];
      if (l == L_PENDING) {
        l = resolve_leaf(a, c);
        a.labels[c] = l;
      }
      l0[i] = l;
      if (l == L_FIRST) g0[i] = a.leaf_dg[c];
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
      uint8_t l = a.labels[c];
      if (l == L_PENDING) {
        l = resolve_leaf(a, c);
        a.labels[c] = l;
      }
      lf[k] = l;
      if (l == L_FIRST) gf[k] = a.leaf_dg[c];
      any |= (l == L_FIRST);
    }
  }
  if (!__ballot(any)) {
    if (lane == 0) a.out_lab[tile] = L_NONE;
    return;
  }
  SweepA<HK> s{a};
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
  uint8_t lc = l1[0];
  Digest gc = g1[0];
  if (R >= 2) {  // level 2 (one parent per lane)
    lc = L_NONE;
    if (lane < T / 4) s.combine(a.d_in - 2, (base >> 2) + lane, l1[0], g1[0], l1[1], g1[1], lc, gc);
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
  }
  const uint32_t wc = wave_sum(s.claims);
  if (lane == 0) {
    if (wc) atomicAdd(stripe(a.ctr, CTR_CLAIMS), (unsigned long long)wc);
    a.out_lab[tile] = lc;
    if (lc == L_FIRST) a.out_dg[tile] = gc;
  }
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
    if (lane_id() == 0) {
      a.cnt_f[a.seg_base[c] + tile] = ftot;
      a.cnt_s[a.seg_base[c] + tile] = tot - ftot;
    }
    ne = nf = 0;
  }
};

template <int HK>
__global__ __launch_bounds__(kThreads) void k_sweep_b(SweepArgs a) {
  const uint32_t lane = lane_id();
  const uint64_t tile = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (tile >= (1ull << a.d_out)) return;  // whole wavefront
  const uint32_t R = a.d_in - a.d_out, T = 1u << R;
  const uint64_t base = tile << R;
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
    // every leaf below is a fixed duplicate: the whole tile is fixed, no roots
    if (lane < R) {
      const uint32_t c = a.d_out + 1 + lane;
      a.cnt_f[a.seg_base[c] + tile] = 0;
      a.cnt_s[a.seg_base[c] + tile] = 0;
    }
    if (lane == 0) {
      a.out_lab[tile] = L_FIXED;
      if (a.d_out == 0) {
        a.cnt_f[a.seg_base[0]] = 0;
        a.cnt_s[a.seg_base[0]] = 0;
      }
    }
    if (a.dbg_labels)
      for (uint32_t r = 1; r <= R; r++)
        for (uint32_t j = lane; j < (T >> r); j += 64) {
          const uint64_t q = (base >> r) + j;
          if (a.stage0 && r == 1 && q >= a.g.half) continue;
          a.dbg_labels[heap_id(a.d_in - r, q)] = L_FIXED;
        }
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
  uint8_t l = a.labels[0];
  if (l == L_PENDING) {
    l = resolve_leaf(a, 0);
    a.labels[0] = l;
  }
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
  const uint32_t grid = (uint32_t)((a.n + kThreads - 1) / kThreads);
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
  MTC_HIP(hipMemcpyAsync(h->seg_base.p, h->seg_base_h.data(), h->seg_base_h.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice, st));
  h->first_roots.ensure(n);
  h->shift_tmp.ensure(n);
  h->region_off.ensure(n + 1);
  h->diff.ensure(32 + 12 * n + 8ull * (h->current_id + 2) + n * (uint64_t)h->chunk_size);
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
  // Sweep A over all stages, then sweep B: every first-occurrence insert must land
  // before any shifted-duplicate lookup (reference_impl.hpp:74-104).
  for (int pass = 0; pass < 2; pass++) {
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

void checkpoint_impl(mtc_dedup* h, const uint8_t* data, uint64_t len, bool make_baseline, hipStream_t st) {
  hipEvent_t ev[5];
  for (auto& e : ev) MTC_HIP(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < 5; i++) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
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

  LeafArgs la{data, len, n, h->chunk_size, h->leaf_dg.p, h->labels.p, h->leaf_slot.p, h->table.p,
              h->table_cap - 1, cur, baseline ? 1 : 0, h->ctl_d, h->ctr.p};
  const bool fast = (C % 16 == 0) && ((uintptr_t)data % 16 == 0);
  if (h->hash_kind == HASH_MD5) launch_leaf<HASH_MD5>(la, fast, st);
  else launch_leaf<HASH_MURMUR3>(la, fast, st);
  MTC_HIP(hipGetLastError());
  MTC_HIP(hipEventRecord(ev[4], st));

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
  if (h->hash_kind == HASH_MD5) run_sweeps<HASH_MD5>(h, sa, st);
  else run_sweeps<HASH_MURMUR3>(h, sa, st);
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
  k_reduce_ctr<<<1, kStripes, 0, st>>>(h->ctr.p, cd);
  MTC_HIP(hipEventRecord(ev[3], st));
  sync_ctl(h, st);
  h->table_size += h->ctl_h->claims;
  h->last_diff_size = h->ctl_h->diff_size;
  MTC_HIP(hipMemcpy(&h->last_header, h->diff.p, sizeof(mtc_header), hipMemcpyDeviceToHost));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[0], ev[0], ev[1]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[1], ev[1], ev[2]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[2], ev[2], ev[3]));
  MTC_HIP(hipEventElapsedTime(&h->last_ms[4], ev[1], ev[4]));
  h->last_ms[3] = 0;
}

}  // namespace mtc
