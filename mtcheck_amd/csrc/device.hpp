// device.hpp -- device primitives for the MI355X tree deduplicator:
// chunk hashes (MTCheck's Murmur3 variant and MD5), the digest -> (tree, node)
// open-addressing table, heap-tree geometry and wave helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtc {

// Node labels: include/utils.hpp:51-57 (FIRST_OCUR..DONE), plus kernel-internal
// states.  PENDING = non-fixed leaf whose FIRST/SHIFT status is
// decided once every leaf of the checkpoint has been inserted.
enum : uint8_t {
  L_FIRST = 0,
  L_FIXED = 1,
  L_SHIFT = 2,
  L_FIRST_DUPL = 3,
  L_DONE = 4,
  L_PENDING = 5,
  L_NONE = 6,
  L_MIN_PENDING = 7,  // leaf insert matched an entry of this checkpoint; its min is still to apply
  L_MIN_FIXED = 8,    // the same for a fixed leaf right after a reset (it stays fixed)
};

enum : int { HASH_MURMUR3 = 0, HASH_MD5 = 1 };

struct Digest {
  uint64_t lo, hi;
};

__host__ __device__ __forceinline__ bool operator==(const Digest& a, const Digest& b) {
  return a.lo == b.lo && a.hi == b.hi;
}

// --------------------------------------------------------------------------
// Murmur3 as MTCheck computes it: MurmurHash3_x64_128, seed 0, with rotl64
// declared to return uint32_t (include/kokkos_murmur3.hpp:31,214-292), so each
// rotate keeps only its low 32 bits.
// --------------------------------------------------------------------------
template <int R>
__host__ __device__ __forceinline__ uint64_t rotl_lo32(uint64_t x) {
  return (uint32_t)((x << R) | (x >> (64 - R)));
}

__host__ __device__ __forceinline__ void m3_block(uint64_t& h1, uint64_t& h2, uint64_t k1, uint64_t k2) {
  constexpr uint64_t c1 = 0x87c37b91114253d5ULL, c2 = 0x4cf5ad432745937fULL;
  k1 *= c1;
  k1 = rotl_lo32<31>(k1);
  k1 *= c2;
  h1 ^= k1;
  h1 = rotl_lo32<27>(h1);
  h1 += h2;
  h1 = h1 * 5 + 0x52dce729;
  k2 *= c2;
  k2 = rotl_lo32<33>(k2);
  k2 *= c1;
  h2 ^= k2;
  h2 = rotl_lo32<31>(h2);
  h2 += h1;
  h2 = h2 * 5 + 0x38495ab5;
}

__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

__host__ __device__ __forceinline__ Digest m3_final(uint64_t h1, uint64_t h2, uint64_t len) {
  h1 ^= len;
  h2 ^= len;
  h1 += h2;
  h2 += h1;
  h1 = fmix64(h1);
  h2 = fmix64(h2);
  h1 += h2;
  h2 += h1;
  return Digest{h1, h2};
}

// Tail words: the last len%16 bytes, little-endian, as the switch at :251-271.
__host__ __device__ __forceinline__ void m3_tail(uint64_t& h1, uint64_t& h2, uint64_t t1, uint64_t t2, uint32_t rem) {
  constexpr uint64_t c1 = 0x87c37b91114253d5ULL, c2 = 0x4cf5ad432745937fULL;
  if (rem > 8) {
    t2 *= c2;
    t2 = rotl_lo32<33>(t2);
    t2 *= c1;
    h2 ^= t2;
  }
  if (rem > 0) {
    t1 *= c1;
    t1 = rotl_lo32<31>(t1);
    t1 *= c2;
    h1 ^= t1;
  }
}

// Byte-granular Murmur3 for any length and alignment (short last chunk, odd chunk sizes).
__host__ __device__ inline Digest m3_bytes(const uint8_t* p, uint64_t len) {
  uint64_t h1 = 0, h2 = 0;
  const uint64_t nb = len / 16;
  for (uint64_t i = 0; i < nb; i++) {
    uint64_t k1 = 0, k2 = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      k1 |= (uint64_t)p[16 * i + b] << (8 * b);
      k2 |= (uint64_t)p[16 * i + 8 + b] << (8 * b);
    }
    m3_block(h1, h2, k1, k2);
  }
  const uint32_t rem = (uint32_t)(len & 15);
  uint64_t t1 = 0, t2 = 0;
  for (uint32_t b = 0; b < rem; b++) {
    uint64_t v = (uint64_t)p[16 * nb + b];
    if (b < 8) t1 |= v << (8 * b);
    else t2 |= v << (8 * (b - 8));
  }
  m3_tail(h1, h2, t1, t2, rem);
  return m3_final(h1, h2, len);
}

// Hash of two adjacent child digests (32 bytes), reference_impl.hpp:79,95.
__device__ __forceinline__ Digest m3_pair(const Digest& l, const Digest& r) {
  uint64_t h1 = 0, h2 = 0;
  m3_block(h1, h2, l.lo, l.hi);
  m3_block(h1, h2, r.lo, r.hi);
  return m3_final(h1, h2, 32);
}

// --------------------------------------------------------------------------
// MD5 (RFC 1321), include/kokkos_md5.hpp:61-182 (compression), :233-265 (finish).
// --------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

#define MTC_MD5_F(x, y, z) (((x) & (y)) | (~(x) & (z)))
#define MTC_MD5_G(x, y, z) (((x) & (z)) | ((y) & ~(z)))
#define MTC_MD5_H(x, y, z) ((x) ^ (y) ^ (z))
#define MTC_MD5_I(x, y, z) ((y) ^ ((x) | ~(z)))
#define MTC_MD5_STEP(f, a, b, c, d, x, t, s) \
  a += f(b, c, d) + (x) + (t);               \
  a = rotl32(a, s) + b;

__host__ __device__ __forceinline__ void md5_compress(uint32_t st[4], const uint32_t X[16]) {
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
  MTC_MD5_STEP(MTC_MD5_F, a, b, c, d, X[0], 0xd76aa478, 7)
  MTC_MD5_STEP(MTC_MD5_F, d, a, b, c, X[1], 0xe8c7b756, 12)
  MTC_MD5_STEP(MTC_MD5_F, c, d, a, b, X[2], 0x242070db, 17)
  MTC_MD5_STEP(MTC_MD5_F, b, c, d, a, X[3], 0xc1bdceee, 22)
  MTC_MD5_STEP(MTC_MD5_F, a, b, c, d, X[4], 0xf57c0faf, 7)
  MTC_MD5_STEP(MTC_MD5_F, d, a, b, c, X[5], 0x4787c62a, 12)
  MTC_MD5_STEP(MTC_MD5_F, c, d, a, b, X[6], 0xa8304613, 17)
  MTC_MD5_STEP(MTC_MD5_F, b, c, d, a, X[7], 0xfd469501, 22)
  MTC_MD5_STEP(MTC_MD5_F, a, b, c, d, X[8], 0x698098d8, 7)
  MTC_MD5_STEP(MTC_MD5_F, d, a, b, c, X[9], 0x8b44f7af, 12)
  MTC_MD5_STEP(MTC_MD5_F, c, d, a, b, X[10], 0xffff5bb1, 17)
  MTC_MD5_STEP(MTC_MD5_F, b, c, d, a, X[11], 0x895cd7be, 22)
  MTC_MD5_STEP(MTC_MD5_F, a, b, c, d, X[12], 0x6b901122, 7)
  MTC_MD5_STEP(MTC_MD5_F, d, a, b, c, X[13], 0xfd987193, 12)
  MTC_MD5_STEP(MTC_MD5_F, c, d, a, b, X[14], 0xa679438e, 17)
  MTC_MD5_STEP(MTC_MD5_F, b, c, d, a, X[15], 0x49b40821, 22)
  MTC_MD5_STEP(MTC_MD5_G, a, b, c, d, X[1], 0xf61e2562, 5)
  MTC_MD5_STEP(MTC_MD5_G, d, a, b, c, X[6], 0xc040b340, 9)
  MTC_MD5_STEP(MTC_MD5_G, c, d, a, b, X[11], 0x265e5a51, 14)
  MTC_MD5_STEP(MTC_MD5_G, b, c, d, a, X[0], 0xe9b6c7aa, 20)
  MTC_MD5_STEP(MTC_MD5_G, a, b, c, d, X[5], 0xd62f105d, 5)
  MTC_MD5_STEP(MTC_MD5_G, d, a, b, c, X[10], 0x02441453, 9)
  MTC_MD5_STEP(MTC_MD5_G, c, d, a, b, X[15], 0xd8a1e681, 14)
  MTC_MD5_STEP(MTC_MD5_G, b, c, d, a, X[4], 0xe7d3fbc8, 20)
  MTC_MD5_STEP(MTC_MD5_G, a, b, c, d, X[9], 0x21e1cde6, 5)
  MTC_MD5_STEP(MTC_MD5_G, d, a, b, c, X[14], 0xc33707d6, 9)
  MTC_MD5_STEP(MTC_MD5_G, c, d, a, b, X[3], 0xf4d50d87, 14)
  MTC_MD5_STEP(MTC_MD5_G, b, c, d, a, X[8], 0x455a14ed, 20)
  MTC_MD5_STEP(MTC_MD5_G, a, b, c, d, X[13], 0xa9e3e905, 5)
  MTC_MD5_STEP(MTC_MD5_G, d, a, b, c, X[2], 0xfcefa3f8, 9)
  MTC_MD5_STEP(MTC_MD5_G, c, d, a, b, X[7], 0x676f02d9, 14)
  MTC_MD5_STEP(MTC_MD5_G, b, c, d, a, X[12], 0x8d2a4c8a, 20)
  MTC_MD5_STEP(MTC_MD5_H, a, b, c, d, X[5], 0xfffa3942, 4)
  MTC_MD5_STEP(MTC_MD5_H, d, a, b, c, X[8], 0x8771f681, 11)
  MTC_MD5_STEP(MTC_MD5_H, c, d, a, b, X[11], 0x6d9d6122, 16)
  MTC_MD5_STEP(MTC_MD5_H, b, c, d, a, X[14], 0xfde5380c, 23)
  MTC_MD5_STEP(MTC_MD5_H, a, b, c, d, X[1], 0xa4beea44, 4)
  MTC_MD5_STEP(MTC_MD5_H, d, a, b, c, X[4], 0x4bdecfa9, 11)
  MTC_MD5_STEP(MTC_MD5_H, c, d, a, b, X[7], 0xf6bb4b60, 16)
  MTC_MD5_STEP(MTC_MD5_H, b, c, d, a, X[10], 0xbebfbc70, 23)
  MTC_MD5_STEP(MTC_MD5_H, a, b, c, d, X[13], 0x289b7ec6, 4)
  MTC_MD5_STEP(MTC_MD5_H, d, a, b, c, X[0], 0xeaa127fa, 11)
  MTC_MD5_STEP(MTC_MD5_H, c, d, a, b, X[3], 0xd4ef3085, 16)
  MTC_MD5_STEP(MTC_MD5_H, b, c, d, a, X[6], 0x04881d05, 23)
  MTC_MD5_STEP(MTC_MD5_H, a, b, c, d, X[9], 0xd9d4d039, 4)
  MTC_MD5_STEP(MTC_MD5_H, d, a, b, c, X[12], 0xe6db99e5, 11)
  MTC_MD5_STEP(MTC_MD5_H, c, d, a, b, X[15], 0x1fa27cf8, 16)
  MTC_MD5_STEP(MTC_MD5_H, b, c, d, a, X[2], 0xc4ac5665, 23)
  MTC_MD5_STEP(MTC_MD5_I, a, b, c, d, X[0], 0xf4292244, 6)
  MTC_MD5_STEP(MTC_MD5_I, d, a, b, c, X[7], 0x432aff97, 10)
  MTC_MD5_STEP(MTC_MD5_I, c, d, a, b, X[14], 0xab9423a7, 15)
  MTC_MD5_STEP(MTC_MD5_I, b, c, d, a, X[5], 0xfc93a039, 21)
  MTC_MD5_STEP(MTC_MD5_I, a, b, c, d, X[12], 0x655b59c3, 6)
  MTC_MD5_STEP(MTC_MD5_I, d, a, b, c, X[3], 0x8f0ccc92, 10)
  MTC_MD5_STEP(MTC_MD5_I, c, d, a, b, X[10], 0xffeff47d, 15)
  MTC_MD5_STEP(MTC_MD5_I, b, c, d, a, X[1], 0x85845dd1, 21)
  MTC_MD5_STEP(MTC_MD5_I, a, b, c, d, X[8], 0x6fa87e4f, 6)
  MTC_MD5_STEP(MTC_MD5_I, d, a, b, c, X[15], 0xfe2ce6e0, 10)
  MTC_MD5_STEP(MTC_MD5_I, c, d, a, b, X[6], 0xa3014314, 15)
  MTC_MD5_STEP(MTC_MD5_I, b, c, d, a, X[13], 0x4e0811a1, 21)
  MTC_MD5_STEP(MTC_MD5_I, a, b, c, d, X[4], 0xf7537e82, 6)
  MTC_MD5_STEP(MTC_MD5_I, d, a, b, c, X[11], 0xbd3af235, 10)
  MTC_MD5_STEP(MTC_MD5_I, c, d, a, b, X[2], 0x2ad7d2bb, 15)
  MTC_MD5_STEP(MTC_MD5_I, b, c, d, a, X[9], 0xeb86d391, 21)
  st[0] += a;
  st[1] += b;
  st[2] += c;
  st[3] += d;
}

__host__ __device__ __forceinline__ void md5_init(uint32_t st[4]) {
  st[0] = 0x67452301u;
  st[1] = 0xefcdab89u;
  st[2] = 0x98badcfeu;
  st[3] = 0x10325476u;
}

__host__ __device__ __forceinline__ Digest md5_digest(const uint32_t st[4]) {
  return Digest{(uint64_t)st[0] | ((uint64_t)st[1] << 32), (uint64_t)st[2] | ((uint64_t)st[3] << 32)};
}

// Final block(s) for a message whose last `rem` (< 64) bytes are in tailw
// (little-endian words, zero beyond rem), total length len.
__host__ __device__ __forceinline__ void md5_finish_words(uint32_t st[4], uint32_t tailw[16], uint32_t rem, uint64_t len) {
  tailw[rem >> 2] |= 0x80u << (8 * (rem & 3));
  if (rem >= 56) {
    md5_compress(st, tailw);
#pragma unroll
    for (int i = 0; i < 16; i++) tailw[i] = 0;
  }
  tailw[14] = (uint32_t)(len << 3);
  tailw[15] = (uint32_t)(len >> 29);
  md5_compress(st, tailw);
}

__host__ __device__ inline Digest md5_bytes(const uint8_t* p, uint64_t len) {
  uint32_t st[4];
  md5_init(st);
  uint32_t X[16];
  const uint64_t nb = len / 64;
  for (uint64_t i = 0; i < nb; i++) {
#pragma unroll
    for (int w = 0; w < 16; w++) {
      const uint8_t* q = p + 64 * i + 4 * w;
      X[w] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
    }
    md5_compress(st, X);
  }
  const uint32_t rem = (uint32_t)(len & 63);
#pragma unroll
  for (int w = 0; w < 16; w++) X[w] = 0;
  for (uint32_t b = 0; b < rem; b++) X[b >> 2] |= (uint32_t)p[64 * nb + b] << (8 * (b & 3));
  md5_finish_words(st, X, rem, len);
  return md5_digest(st);
}

__device__ __forceinline__ Digest md5_pair(const Digest& l, const Digest& r) {
  uint32_t st[4];
  md5_init(st);
  uint32_t X[16] = {(uint32_t)l.lo, (uint32_t)(l.lo >> 32), (uint32_t)l.hi, (uint32_t)(l.hi >> 32),
                    (uint32_t)r.lo, (uint32_t)(r.lo >> 32), (uint32_t)r.hi, (uint32_t)(r.hi >> 32),
                    0x80u, 0, 0, 0, 0, 0, 256u, 0};
  md5_compress(st, X);
  return md5_digest(st);
}

template <int HK>
__device__ __forceinline__ Digest hash_pair(const Digest& l, const Digest& r) {
  if constexpr (HK == HASH_MD5) return md5_pair(l, r);
  else return m3_pair(l, r);
}

template <int HK>
__device__ __forceinline__ Digest hash_bytes(const uint8_t* p, uint64_t len) {
  if constexpr (HK == HASH_MD5) return md5_bytes(p, len);
  else return m3_bytes(p, len);
}

// --------------------------------------------------------------------------
// Digest table: open addressing, 32-byte slots {k0, k1, val, pad}.  val packs the
// reference's NodeID as (tree << 32) | node, so a 64-bit atomicMin keeps
// entries of older checkpoints and, within the current checkpoint, the lowest
// node (the lowest-offset rule of reference_impl.hpp:51-69).  Empty words hold
// TEMPTY.  Every access to a slot word inside a kernel that inserts is an
// agent-scope atomic, so claims are coherent across the XCDs' L2 caches.
// A digest whose k0 equals TEMPTY lives in the one overflow slot at index cap.
// --------------------------------------------------------------------------
constexpr uint64_t TEMPTY = ~0ull;

struct alignas(32) Slot {
  uint64_t k0, k1, val, pad;
};

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t cas_relaxed(uint64_t* p, uint64_t expect, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint64_t fetch_min_relaxed(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void min_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Streaming 16-byte accesses for data that is touched once (chunk bytes, the
// previous leaf digests, restart copies): non-temporal, so they do not displace
// what later kernels reuse (the leaf kernel runs 7-9% faster with them).  The
// 4-byte-aligned forms serve diff data sections, which start 4-byte aligned.
typedef uint32_t u32x4_nt __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_nt_a4 __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ uint4 ld_stream16(const void* p) {
  const u32x4_nt x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt*>(p));
  return uint4{x.x, x.y, x.z, x.w};
}
__device__ __forceinline__ uint4 ld_stream16_a4(const void* p) {
  const u32x4_nt_a4 x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt_a4*>(p));
  return uint4{x.x, x.y, x.z, x.w};
}
__device__ __forceinline__ void st_stream16(void* p, const uint4& v) {
  __builtin_nontemporal_store(u32x4_nt{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4_nt*>(p));
}
__device__ __forceinline__ void st_stream16_a4(void* p, const uint4& v) {
  __builtin_nontemporal_store(u32x4_nt_a4{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4_nt_a4*>(p));
}

__device__ __forceinline__ uint64_t slot_home(const Digest& d, uint64_t mask) {
  return (d.hi ^ (d.lo >> 32) ^ (d.lo << 17)) & mask;
}

// Error bits reported through the status word.
enum : uint32_t { ERR_TABLE_FULL = 1u, ERR_TABLE_SPIN = 2u, ERR_CORRUPT = 4u, ERR_OVERFLOW_SLOT = 8u };

constexpr uint32_t kMaxSpin = 1u << 20;

// Probing runs in sector order: slots are 32 bytes and HBM is read in 64-byte
// sectors, so slots 2k and 2k+1 arrive together.  Both routines read the keys of
// the probe position's whole sector at once (one access per sector instead of one
// per slot; reading a whole 128-byte line costs two accesses and mostly goes
// unused at low load).  The probe sequence itself is plain linear probing.
constexpr uint32_t kLineSlots = 2;

// The value of a matched entry only ever decreases (use_min) or is set once (insert-
// if-absent).  Skipping the atomic when it cannot change keeps lanes that repeat a
// popular digest (zero chunks) from serialising on one address.
// Returns the value it read.
__device__ __forceinline__ uint64_t table_update_val(Slot* s, uint64_t v, bool use_min) {
  const uint64_t now = ld_relaxed(&s->val);
  if (use_min) {
    if (now > v) min_relaxed(&s->val, v);
  } else if (now == TEMPTY) {
    cas_relaxed(&s->val, TEMPTY, v);
  }
  return now;
}

// Insert d with value v.  use_min: val = min(val, v) (leaf inserts); otherwise
// val is set only if still empty (insert-if-absent, first-occurrence nodes).
// cas_first: the digest is expected to be new, so the home slot is claimed without
// reading it first (a digest already present then costs a failed claim, an atomic).  Returns the slot index, or ~0u on error (bit set in *err).
// *claimed = 1 when this call created the entry.  *seen (if given) receives the
// value a matched entry held when it was read (TEMPTY when this call claimed it, or
// when its claimer had not stored the value yet).  defer (with use_min): a claimer
// stores v and a matcher leaves the value alone; the caller applies the matchers'
// mins after the whole insert pass, when every claimer's store has landed.
__device__ inline uint32_t table_insert(Slot* tab, uint64_t mask, const Digest& d, uint64_t v, bool use_min,
                                        uint32_t* err, bool* claimed, bool cas_first = false,
                                        uint64_t* seen = nullptr, bool defer = false) {
  *claimed = false;
  if (seen) *seen = TEMPTY;
  if (d.lo == TEMPTY) {  // overflow slot, keyed by k1 alone
    Slot* s = tab + mask + 1;
    uint64_t old = cas_relaxed(&s->k1, TEMPTY, d.hi);
    if (old != TEMPTY && old != d.hi) {
      atomicOr(err, ERR_OVERFLOW_SLOT);
      return ~0u;
    }
    *claimed = (old == TEMPTY);
    if (use_min) min_relaxed(&s->val, v);
    else cas_relaxed(&s->val, TEMPTY, v);
    return (uint32_t)(mask + 1);
  }
  // One slot per iteration.  A claim's stores sit in a plain if-block that every
  // lane leaves before any lane can exit or loop again, and a lane that waits for
  // k1 does so by repeating the loop: a claimer in the same wavefront has then
  // always published k1 (an inner wait loop could run while the claimer is masked
  // off before its store).
  uint64_t i = slot_home(d, mask);
  uint64_t probes = 0;
  uint32_t spins = 0;
  uint64_t kl0 = 0, kl1 = 0;  // k0 of the current sector's slots
  bool loaded = false;
  bool assume_empty = cas_first;
  while (true) {
    Slot* s = tab + i;
    uint64_t k0 = TEMPTY;
    if (!assume_empty) {
      if (!loaded) {
        const Slot* line = tab + (i & ~(uint64_t)(kLineSlots - 1));
        kl0 = ld_relaxed(&line[0].k0);
        kl1 = ld_relaxed(&line[1].k0);
        loaded = true;
      }
      k0 = (i & 1) ? kl1 : kl0;
    }
    assume_empty = false;
    bool mine = false;
    if (k0 == TEMPTY) {
      const uint64_t old = cas_relaxed(&s->k0, TEMPTY, d.lo);
      if (old == TEMPTY) {
        st_relaxed(&s->k1, d.hi);
        // Insert-if-absent: the claimer's value stands.  A matcher only claims an
        // empty value, so its CAS either fails or is overwritten by this store.
        if (use_min && !defer) min_relaxed(&s->val, v);
        else st_relaxed(&s->val, v);
        mine = true;
      } else {
        k0 = old;
      }
    }
    if (mine) {
      *claimed = true;
      return (uint32_t)i;
    }
    if (k0 == d.lo) {
      const uint64_t k1 = ld_relaxed(&s->k1);
      if (k1 == d.hi) {
        const uint64_t now = defer ? ld_relaxed(&s->val) : table_update_val(s, v, use_min);
        if (seen) *seen = now;
        return (uint32_t)i;
      }
      if (k1 == TEMPTY) {  // claimer has not published k1 yet: look again
        if (++spins > kMaxSpin) {
          atomicOr(err, ERR_TABLE_SPIN);
          return ~0u;
        }
        continue;
      }
    }
    if (++probes > mask) {
      atomicOr(err, ERR_TABLE_FULL);
      return ~0u;
    }
    i = (i + 1) & mask;
    if ((i & (kLineSlots - 1)) == 0) loaded = false;
  }
}

// Read-only lookup, valid once the inserting kernel has completed.  Reads the
// (k0, k1) pairs of the probe position's whole sector with 16-byte loads.
__device__ inline uint32_t table_find(const Slot* tab, uint64_t mask, const Digest& d) {
  if (d.lo == TEMPTY) {
    const Slot* s = tab + mask + 1;
    return (s->k1 == d.hi && s->val != TEMPTY) ? (uint32_t)(mask + 1) : ~0u;
  }
  uint64_t i = slot_home(d, mask);
  for (uint64_t probes = 0; probes <= mask;) {
    const Slot* line = tab + (i & ~(uint64_t)(kLineSlots - 1));
    ulonglong2 k[kLineSlots];
#pragma unroll
    for (uint32_t q = 0; q < kLineSlots; q++) k[q] = *reinterpret_cast<const ulonglong2*>(&line[q].k0);
    const uint32_t j0 = (uint32_t)(i & (kLineSlots - 1));
#pragma unroll
    for (uint32_t q = 0; q < kLineSlots; q++) {
      if (q < j0) continue;
      if (k[q].x == TEMPTY) return ~0u;
      if (k[q].x == d.lo && k[q].y == d.hi) return (uint32_t)(line + q - tab);
      probes++;
    }
    i = ((i | (kLineSlots - 1)) + 1) & mask;
  }
  return ~0u;
}

// --------------------------------------------------------------------------
// Heap-tree geometry (include/kokkos_merkle_tree.hpp:16-33): 2n-1 nodes,
// children 2v+1 and 2v+2, leaf n-1+c holds chunk c.  The bottom level (depth D)
// holds m = 2n - 2^D leaves (chunks 2^D-n ...); when n is not a power of two the
// remaining chunks 0 .. 2^D-n-1 are leaves at depth D-1, positions m/2 ...
// --------------------------------------------------------------------------
struct Geo {
  uint32_t n;     // chunks
  uint32_t D;     // depth of the bottom level
  uint64_t m;     // leaves on the bottom level
  uint64_t half;  // m/2 = first depth-(D-1) position that is a leaf
};

__host__ __device__ inline uint32_t ilog2_64(uint64_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  return 63u - (uint32_t)__clzll((long long)x);
#else
  return 63u - (uint32_t)__builtin_clzll(x);
#endif
}

__host__ __device__ inline Geo make_geo(uint32_t n) {
  Geo g;
  g.n = n;
  g.D = ilog2_64(2ull * n - 1);
  g.m = 2ull * n - (1ull << g.D);
  g.half = g.m / 2;
  return g;
}

// Leaves of `node` as up to two ascending runs of chunk indices.
__host__ __device__ inline int leaf_runs(const Geo& g, uint32_t node, uint32_t st[2], uint32_t ct[2]) {
  const uint64_t nn = g.n, nodes = 2 * nn - 1;
  const uint32_t dv = ilog2_64((uint64_t)node + 1);
  const uint32_t h = g.D - dv;
  int r = 0;
  if (h >= 1) {  // depth D-1 leaves (lowest chunk indices)
    uint64_t lo = (((uint64_t)node + 1) << (h - 1)) - 1, hi = (((uint64_t)node + 2) << (h - 1)) - 2;
    if (lo < nn - 1) lo = nn - 1;
    if (lo <= hi) {
      st[r] = (uint32_t)(lo - (nn - 1));
      ct[r] = (uint32_t)(hi - lo + 1);
      r++;
    }
  }
  {
    uint64_t lo = (((uint64_t)node + 1) << h) - 1, hi = (((uint64_t)node + 2) << h) - 2;
    if (hi > nodes - 1) hi = nodes - 1;
    if (lo <= hi) {
      st[r] = (uint32_t)(lo - (nn - 1));
      ct[r] = (uint32_t)(hi - lo + 1);
      r++;
    }
  }
  return r;
}

// The reference's leaf range of a node (include/kokkos_merkle_tree.hpp:102-128):
// one run from the leftmost bottom-level leaf under it, as long as the walk to the
// rightmost one says, clamped to the last node.  Where a node's leaves lie on two
// depths (a chunk count that is not a power of two) this is not the node's leaf
// set; it is kept only to reproduce the reference's bytes on request.
__host__ __device__ inline int ref_leaf_run(const Geo& g, uint32_t node, uint32_t st[2], uint32_t ct[2]) {
  const uint64_t num_nodes = 2ull * g.n - 1;
  uint64_t l = 2ull * node + 1, r = 2ull * node + 2;
  while (l < num_nodes) {
    l = 2 * l + 1;
    r = 2 * r + 2;
  }
  l = (l - 1) / 2;
  r = (r - 2) / 2;
  if (r > num_nodes) r = num_nodes - 1;
  st[0] = (uint32_t)(l - (g.n - 1));
  ct[0] = (uint32_t)(r - l + 1);
  return 1;
}

// Leaves of a region root as stored in the data section: its true leaf set, or
// with `compat` the reference's single run.
__host__ __device__ inline int region_runs(const Geo& g, uint32_t node, int compat, uint32_t st[2], uint32_t ct[2]) {
  return compat ? ref_leaf_run(g, node, st, ct) : leaf_runs(g, node, st, ct);
}

__host__ __device__ inline uint32_t leaf_count(const Geo& g, uint32_t node) {
  uint32_t st[2], ct[2];
  int r = leaf_runs(g, node, st, ct);
  return r == 0 ? 0 : (r == 1 ? ct[0] : ct[0] + ct[1]);
}

// j-th leaf (ascending chunk order) of a node given its runs.
__host__ __device__ inline uint32_t run_at(int r, const uint32_t st[2], const uint32_t ct[2], uint32_t j) {
  return (r == 2 && j >= ct[0]) ? st[1] + (j - ct[0]) : st[0] + j;
}

// Rank of `chunk` inside a node's leaf set (ascending chunk order).
__host__ __device__ inline uint32_t run_rank(int r, const uint32_t st[2], const uint32_t ct[2], uint32_t chunk) {
  return (r == 2 && chunk >= st[1]) ? ct[0] + (chunk - st[1]) : chunk - st[0];
}

// --------------------------------------------------------------------------
// Wave helpers (64-lane wavefronts).
// --------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Adds popcount(ballot(flag)) to *ctr with one atomic per wave.
__device__ __forceinline__ void wave_count_add(unsigned long long* ctr, bool flag) {
  uint64_t m = __ballot(flag);
  uint64_t active = __ballot(1);
  if (m && lane_id() == (uint32_t)(__ffsll((long long)active) - 1)) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

}  // namespace mtc
