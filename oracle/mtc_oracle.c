/*
 * mtc_oracle.c -- CPU restatement of MTCheck's tree (lowest-offset) deduplication
 * path.  TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the HIP
 * library never links or calls it.
 *
 * What it restates (reference paths relative to the TauferLab/MTCheck tree):
 *   hashes        include/kokkos_murmur3.hpp:14-53,214-292,432-437
 *                 include/kokkos_md5.hpp:49-323, include/hash_functions.hpp:16-20
 *   digest table  include/map_helpers.hpp:7-113 (insert-if-absent UnorderedMap)
 *   leaf ranges   include/kokkos_merkle_tree.hpp:102-142
 *   labelling     include/reference_impl.hpp:15-132 (dedup_low_offset_ref)
 *   diff          src/tree_approach.cpp:511-710 (collect_diff), serial order
 *   lifecycle     src/tree_approach.cpp:1577-1689,1740-1748
 *   restart       src/tree_approach.cpp:712-1128
 *
 * Two deliberate departures, both confined to inputs on which the reference
 * cannot restart its own diff (compat_ref=1 reproduces the reference bytes):
 *   D1  regions whose leaves sit on two tree levels (chunk count not a power of
 *       two) use their true leaf set, ascending chunk order, instead of the
 *       truncated range of kokkos_merkle_tree.hpp:102-128;
 *   D2  a shifted-duplicate node 0 is emitted as a region root (reference_impl.hpp
 *       :105-106 emits node 0 only when it is a first occurrence).
 *
 * Pinning (see tests/): MD5 against RFC 1321 and hashlib; the Murmur3 body against
 * smhasher's published verification value with a full rotate; labels against the
 * reference's test_case_01..10 counts; diffs and restarts against invariants
 * (region disjointness and cover, content equality of shifted pairs, round trips).
 */
#include "mtc_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ========================================================================== */
/* Murmur3: include/kokkos_murmur3.hpp:214-292                                 */
/* ========================================================================== */

static inline uint64_t rd64(const uint8_t* p) {
  uint64_t v;
  memcpy(&v, p, 8); /* x86-64 is little-endian, like getblock64 on the GPU */
  return v;
}

/* kokkos_murmur3.hpp:31 declares rotl64 with a uint32_t return type, so every
 * rotate result is cut to its low 32 bits.  `full` selects the canonical rotate. */
static inline uint64_t rot(uint64_t x, int r, int full) {
  uint64_t v = (x << r) | (x >> (64 - r));
  return full ? v : (uint64_t)(uint32_t)v;
}

static inline uint64_t fmix64(uint64_t k) { /* kokkos_murmur3.hpp:45-53 */
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

static void murmur3_x64_128(const uint8_t* data, uint64_t len, uint32_t seed, uint8_t out[16], int full) {
  const uint64_t nblocks = len / 16;
  uint64_t h1 = seed, h2 = seed;
  const uint64_t c1 = 0x87c37b91114253d5ULL;
  const uint64_t c2 = 0x4cf5ad432745937fULL;
  for (uint64_t i = 0; i < nblocks; i++) { /* body, :229-241 */
    uint64_t k1 = rd64(data + 16 * i);
    uint64_t k2 = rd64(data + 16 * i + 8);
    k1 *= c1; k1 = rot(k1, 31, full); k1 *= c2; h1 ^= k1;
    h1 = rot(h1, 27, full); h1 += h2; h1 = h1 * 5 + 0x52dce729;
    k2 *= c2; k2 = rot(k2, 33, full); k2 *= c1; h2 ^= k2;
    h2 = rot(h2, 31, full); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
  }
  const uint8_t* tail = data + nblocks * 16; /* tail, :246-271 */
  uint64_t k1 = 0, k2 = 0;
  unsigned rem = (unsigned)(len & 15);
  if (rem > 8) {
    for (unsigned b = rem; b > 8; b--) k2 ^= (uint64_t)tail[b - 1] << (8 * (b - 9));
    k2 *= c2; k2 = rot(k2, 33, full); k2 *= c1; h2 ^= k2;
  }
  if (rem > 0) {
    unsigned top = rem > 8 ? 8 : rem;
    for (unsigned b = top; b > 0; b--) k1 ^= (uint64_t)tail[b - 1] << (8 * (b - 1));
    k1 *= c1; k1 = rot(k1, 31, full); k1 *= c2; h1 ^= k1;
  }
  h1 ^= len; h2 ^= len; /* finalization, :277-289 */
  h1 += h2; h2 += h1;
  h1 = fmix64(h1); h2 = fmix64(h2);
  h1 += h2; h2 += h1;
  memcpy(out, &h1, 8);
  memcpy(out + 8, &h2, 8);
}

void orc_murmur3_mtcheck(const void* data, uint64_t len, uint8_t out[16]) {
  murmur3_x64_128((const uint8_t*)data, len, 0, out, 0); /* hash(), :432-437 */
}

void orc_murmur3_canonical(const void* data, uint64_t len, uint32_t seed, uint8_t out[16]) {
  murmur3_x64_128((const uint8_t*)data, len, seed, out, 1);
}

/* ========================================================================== */
/* MD5 (RFC 1321): include/kokkos_md5.hpp:49-323                              */
/* ========================================================================== */

static const uint32_t md5_T[64] = {
  0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
  0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
  0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
  0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
  0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
  0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
  0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
  0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};
static const unsigned md5_S[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                   5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                   4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                   6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};

static void md5_block(uint32_t st[4], const uint8_t blk[64]) {
  uint32_t X[16];
  for (int i = 0; i < 16; i++)
    X[i] = (uint32_t)blk[4 * i] | ((uint32_t)blk[4 * i + 1] << 8) | ((uint32_t)blk[4 * i + 2] << 16) |
           ((uint32_t)blk[4 * i + 3] << 24);
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
  for (int i = 0; i < 64; i++) {
    uint32_t f;
    int g;
    if (i < 16) { f = (b & c) | (~b & d); g = i; }
    else if (i < 32) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
    else if (i < 48) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
    else { f = c ^ (b | ~d); g = (7 * i) & 15; }
    uint32_t t = a + f + md5_T[i] + X[g];
    a = d; d = c; c = b;
    b = b + ((t << md5_S[i]) | (t >> (32 - md5_S[i])));
  }
  st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

void orc_md5(const void* data, uint64_t len, uint8_t out[16]) {
  const uint8_t* p = (const uint8_t*)data;
  uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  uint64_t full = len / 64;
  for (uint64_t i = 0; i < full; i++) md5_block(st, p + 64 * i);
  uint8_t last[128];
  uint64_t rem = len - 64 * full;
  memset(last, 0, sizeof last);
  memcpy(last, p + 64 * full, rem);
  last[rem] = 0x80;
  uint64_t nb = (rem < 56) ? 1 : 2;
  uint64_t bits = len * 8;
  for (int i = 0; i < 8; i++) last[64 * nb - 8 + i] = (uint8_t)(bits >> (8 * i));
  for (uint64_t i = 0; i < nb; i++) md5_block(st, last + 64 * i);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(st[i] >> (8 * j));
}

void orc_hash(int kind, const void* data, uint64_t len, uint8_t out[16]) {
  if (kind == ORC_HASH_MD5) orc_md5(data, len, out);
  else orc_murmur3_mtcheck(data, len, out);
}

/* ========================================================================== */
/* Leaf sets: include/kokkos_merkle_tree.hpp:102-142                          */
/* ========================================================================== */

static inline uint32_t ilog2_u64(uint64_t x) { return 63u - (uint32_t)__builtin_clzll(x); }

int orc_leaf_runs(uint32_t node, uint32_t n, uint32_t start[2], uint32_t count[2]) {
  /* Heap layout: children 2v+1, 2v+2; leaves are nodes n-1 .. 2n-2, chunk = leaf-(n-1).
   * All leaves sit on depth D (the bottom level) or D-1.  Leaves of depth D-1 hold
   * the lowest chunk indices, so a node spanning both levels owns a run at the start
   * of the buffer and a run at its end. */
  uint64_t nn = n, num_nodes = 2 * nn - 1;
  uint32_t D = ilog2_u64(num_nodes);
  uint32_t dv = ilog2_u64((uint64_t)node + 1);
  uint32_t h = D - dv;
  int runs = 0;
  /* depth D-1 part (only leaves, i.e. ids >= n-1) */
  if (h >= 1) {
    uint64_t lo = (((uint64_t)node + 1) << (h - 1)) - 1, hi = (((uint64_t)node + 2) << (h - 1)) - 2;
    if (lo < nn - 1) lo = nn - 1;
    if (lo <= hi) { start[runs] = (uint32_t)(lo - (nn - 1)); count[runs] = (uint32_t)(hi - lo + 1); runs++; }
  }
  /* depth D part */
  {
    uint64_t lo = (((uint64_t)node + 1) << h) - 1, hi = (((uint64_t)node + 2) << h) - 2;
    if (hi > num_nodes - 1) hi = num_nodes - 1;
    if (lo <= hi) { start[runs] = (uint32_t)(lo - (nn - 1)); count[runs] = (uint32_t)(hi - lo + 1); runs++; }
  }
  return runs;
}

uint32_t orc_ref_num_leaf_descendents(uint32_t node, uint32_t num_nodes) { /* :102-114 */
  uint64_t l = 2 * (uint64_t)node + 1, r = 2 * (uint64_t)node + 2;
  while (l < num_nodes) { l = 2 * l + 1; r = 2 * r + 2; }
  l = (l - 1) / 2;
  r = (r - 2) / 2;
  if (r > num_nodes) r = num_nodes - 1;
  return (uint32_t)(r - l + 1);
}

uint32_t orc_ref_leftmost_leaf(uint32_t node, uint32_t num_nodes) { /* :116-128 */
  uint64_t l = 2 * (uint64_t)node + 1;
  while (l < num_nodes) l = 2 * l + 1;
  return (uint32_t)((l - 1) / 2);
}

/* ========================================================================== */
/* Digest -> NodeID table (insert-if-absent), include/map_helpers.hpp:38-108  */
/* ========================================================================== */

typedef struct {
  uint8_t key[16];
  uint32_t node, tree; /* NodeID; node == UINT32_MAX marks an empty slot */
} orc_slot;

typedef struct {
  orc_slot* slots;
  uint64_t cap, size;
} orc_map;

static uint64_t map_home(const uint8_t* key, uint64_t cap) {
  uint64_t a, b;
  memcpy(&a, key, 8);
  memcpy(&b, key + 8, 8);
  uint64_t h = fmix64(a ^ (b * 0x9e3779b97f4a7c15ULL));
  return h & (cap - 1);
}

static int map_init(orc_map* m, uint64_t cap) {
  uint64_t c = 16;
  while (c < cap) c <<= 1;
  m->slots = (orc_slot*)malloc(c * sizeof(orc_slot));
  if (!m->slots) return ORC_ERR_NOMEM;
  for (uint64_t i = 0; i < c; i++) m->slots[i].node = UINT32_MAX;
  m->cap = c;
  m->size = 0;
  return ORC_OK;
}

static void map_free(orc_map* m) {
  free(m->slots);
  m->slots = NULL;
  m->cap = m->size = 0;
}

static orc_slot* map_find(const orc_map* m, const uint8_t* key) {
  if (!m->slots) return NULL;
  uint64_t i = map_home(key, m->cap);
  for (;;) {
    orc_slot* s = &m->slots[i];
    if (s->node == UINT32_MAX) return NULL;
    if (memcmp(s->key, key, 16) == 0) return s;
    i = (i + 1) & (m->cap - 1);
  }
}

static int map_grow(orc_map* m) {
  orc_map nm;
  if (map_init(&nm, m->cap * 2) != ORC_OK) return ORC_ERR_NOMEM;
  for (uint64_t i = 0; i < m->cap; i++) {
    orc_slot* s = &m->slots[i];
    if (s->node == UINT32_MAX) continue;
    uint64_t j = map_home(s->key, nm.cap);
    while (nm.slots[j].node != UINT32_MAX) j = (j + 1) & (nm.cap - 1);
    nm.slots[j] = *s;
    nm.size++;
  }
  map_free(m);
  *m = nm;
  return ORC_OK;
}

/* Kokkos::UnorderedMap::insert: returns 1 on success (key was absent), 0 when
 * the key exists (the stored value is left as is).  *slot receives the entry. */
static int map_insert(orc_map* m, const uint8_t* key, uint32_t node, uint32_t tree, orc_slot** slot) {
  if (2 * (m->size + 1) > m->cap) map_grow(m);
  uint64_t i = map_home(key, m->cap);
  for (;;) {
    orc_slot* s = &m->slots[i];
    if (s->node == UINT32_MAX) {
      memcpy(s->key, key, 16);
      s->node = node;
      s->tree = tree;
      m->size++;
      if (slot) *slot = s;
      return 1;
    }
    if (memcmp(s->key, key, 16) == 0) {
      if (slot) *slot = s;
      return 0;
    }
    i = (i + 1) & (m->cap - 1);
  }
}

/* ========================================================================== */
/* Deduplicator                                                                */
/* ========================================================================== */

struct orc_dedup {
  uint32_t chunk_size;
  int hash_kind, compat;
  uint32_t current_id, baseline_id;
  uint32_t n, num_nodes;
  uint8_t* tree;   /* 16 * num_nodes */
  uint8_t* labels; /* num_nodes */
  orc_map map;
  uint32_t *first, *shift;
  uint32_t nfirst, nshift;
  uint8_t* diff;
  uint64_t diff_size;
};

orc_dedup* orc_dedup_create(uint32_t chunk_size, int hash_kind, int compat_ref) {
  if (chunk_size == 0) return NULL;
  orc_dedup* d = (orc_dedup*)calloc(1, sizeof(orc_dedup));
  if (!d) return NULL;
  d->chunk_size = chunk_size;
  d->hash_kind = hash_kind;
  d->compat = compat_ref;
  return d;
}

void orc_dedup_destroy(orc_dedup* d) {
  if (!d) return;
  free(d->tree);
  free(d->labels);
  map_free(&d->map);
  free(d->first);
  free(d->shift);
  free(d->diff);
  free(d);
}

#define DG(t, i) ((t) + 16 * (uint64_t)(i))

/* dedup_low_offset_ref, include/reference_impl.hpp:15-132 */
static int label_low_offset(orc_dedup* d, const uint8_t* data, uint64_t len, uint8_t* cur_tree) {
  const uint32_t n = d->n, num_nodes = d->num_nodes, cur = d->current_id;
  const uint64_t C = d->chunk_size;
  uint8_t* labels = d->labels;
  memset(labels, ORC_DONE, num_nodes); /* :39 */
  /* Leaves in ascending order, :44-71 */
  for (uint64_t leaf = n - 1; leaf < num_nodes; leaf++) {
    uint64_t off = (leaf - (n - 1)) * C;
    uint64_t nb = (leaf == num_nodes - 1) ? len - off : C;
    orc_hash(d->hash_kind, data + off, nb, DG(cur_tree, leaf));
    orc_slot* s = NULL;
    int ok = map_insert(&d->map, DG(cur_tree, leaf), (uint32_t)leaf, cur, &s);
    if (memcmp(DG(d->tree, leaf), DG(cur_tree, leaf), 16) == 0) {
      labels[leaf] = ORC_FIXED_DUPL;
    } else if (ok) {
      labels[leaf] = ORC_FIRST_OCUR;
    } else if (s->tree == cur) {
      if (s->node > leaf) { /* unreachable with ascending leaves; kept as written */
        labels[leaf] = ORC_FIRST_OCUR;
        labels[s->node] = ORC_SHIFT_DUPL;
        s->node = (uint32_t)leaf;
      } else {
        labels[leaf] = ORC_SHIFT_DUPL;
      }
    } else {
      labels[leaf] = ORC_SHIFT_DUPL;
    }
  }
  if (n < 2) return ORC_OK;
  /* First-occurrence forest, :74-82 */
  for (int64_t node = (int64_t)n - 2; node >= 0; node--) {
    uint64_t l = 2 * (uint64_t)node + 1, r = l + 1;
    if (labels[l] == ORC_FIRST_OCUR && labels[r] == ORC_FIRST_OCUR) {
      labels[node] = ORC_FIRST_OCUR;
      orc_hash(d->hash_kind, DG(cur_tree, l), 32, DG(cur_tree, node));
      map_insert(&d->map, DG(cur_tree, node), (uint32_t)node, cur, NULL);
    }
  }
  return ORC_OK;
}

/* Shifted-duplicate forest and root set, reference_impl.hpp:83-119 */
static void forest_roots(orc_dedup* d, uint8_t* cur_tree, uint8_t* is_root) {
  const uint32_t n = d->n;
  uint8_t* labels = d->labels;
  for (int64_t node = (int64_t)n - 2; node >= 0; node--) {
    uint64_t l = 2 * (uint64_t)node + 1, r = l + 1;
    uint8_t a = labels[l], b = labels[r];
    if (a != b) {
      labels[node] = ORC_DONE;
      if (a != ORC_FIXED_DUPL && a != ORC_DONE) is_root[l] = 1;
      if (b != ORC_FIXED_DUPL && b != ORC_DONE) is_root[r] = 1;
    } else if (a == ORC_FIXED_DUPL) {
      labels[node] = ORC_FIXED_DUPL;
    } else if (a == ORC_SHIFT_DUPL) {
      orc_hash(d->hash_kind, DG(cur_tree, l), 32, DG(cur_tree, node));
      if (map_find(&d->map, DG(cur_tree, node))) {
        labels[node] = ORC_SHIFT_DUPL;
      } else {
        labels[node] = ORC_DONE;
        is_root[l] = 1;
        is_root[r] = 1;
      }
    }
  }
  if (labels[0] == ORC_FIRST_OCUR) is_root[0] = 1; /* :105-106 */
  if (!d->compat && labels[0] == ORC_SHIFT_DUPL) is_root[0] = 1; /* D2 */
}

/* Leaves of a region root in data-section order (D1 unless compat). */
static int region_runs(const orc_dedup* d, uint32_t node, uint32_t start[2], uint32_t count[2]) {
  if (d->compat) {
    start[0] = orc_ref_leftmost_leaf(node, d->num_nodes) - (d->n - 1);
    count[0] = orc_ref_num_leaf_descendents(node, d->num_nodes);
    return 1;
  }
  return orc_leaf_runs(node, d->n, start, count);
}

static void put32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static uint32_t get32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }

/* collect_diff, src/tree_approach.cpp:511-710, in serial (canonical) order */
static int collect_diff(orc_dedup* d, const uint8_t* data, uint64_t len) {
  const uint32_t F = d->nfirst, S = d->nshift, cur = d->current_id;
  const uint64_t C = d->chunk_size;
  /* Region sizes; every first root must map to itself in the table (:546-558). */
  uint64_t total_chunks = 0;
  for (uint32_t i = 0; i < F; i++) {
    uint32_t node = d->first[i];
    orc_slot* s = map_find(&d->map, DG(d->tree, node));
    if (!s || s->node != node || s->tree != cur) return ORC_ERR_STATE;
    uint32_t st[2], ct[2];
    int nr = region_runs(d, node, st, ct);
    for (int k = 0; k < nr; k++) total_chunks += ct[k];
  }
  /* Source checkpoint of each shifted root (:574-580) and the (chkpt, count) table. */
  uint32_t* src = (uint32_t*)malloc(sizeof(uint32_t) * (S ? S : 1));
  uint32_t* prev = (uint32_t*)malloc(sizeof(uint32_t) * (S ? S : 1));
  uint64_t* cnt = (uint64_t*)calloc((size_t)cur + 1, sizeof(uint64_t));
  if (!src || !prev || !cnt) { free(src); free(prev); free(cnt); return ORC_ERR_NOMEM; }
  for (uint32_t i = 0; i < S; i++) {
    orc_slot* s = map_find(&d->map, DG(d->tree, d->shift[i]));
    if (!s || s->tree > cur) { free(src); free(prev); free(cnt); return ORC_ERR_STATE; }
    src[i] = s->tree;
    prev[i] = s->node;
    cnt[s->tree]++;
  }
  uint32_t P = 0;
  for (uint32_t c = 0; c <= cur; c++) P += cnt[c] > 0;
  uint64_t size = 32 + 4ull * F + 8ull * P + 8ull * S + C * total_chunks; /* :622 */
  free(d->diff);
  d->diff = (uint8_t*)calloc(size, 1);
  if (!d->diff) { free(src); free(prev); free(cnt); return ORC_ERR_NOMEM; }
  d->diff_size = size;
  uint8_t* p = d->diff;
  orc_header h;
  h.ref_id = d->baseline_id; /* :700-706 */
  h.chkpt_id = cur;
  h.datalen = len;
  h.chunk_size = d->chunk_size;
  h.num_first_ocur = F;
  h.num_prior_chkpts = P;
  h.num_shift_dupl = S;
  memcpy(p, &h, 32);
  p += 32;
  for (uint32_t i = 0; i < F; i++, p += 4) put32(p, d->first[i]); /* :630-633 */
  for (uint32_t c = 0; c <= cur; c++) /* :657-665, ascending checkpoint id */
    if (cnt[c]) { put32(p, c); put32(p + 4, (uint32_t)cnt[c]); p += 8; }
  for (uint32_t c = 0; c <= cur; c++) { /* :670-696, grouped by source, ascending node */
    if (!cnt[c]) continue;
    for (uint32_t i = 0; i < S; i++)
      if (src[i] == c) { put32(p, d->shift[i]); put32(p + 4, prev[i]); p += 8; }
  }
  for (uint32_t i = 0; i < F; i++) { /* :636-650 */
    uint32_t st[2], ct[2];
    int nr = region_runs(d, d->first[i], st, ct);
    for (int k = 0; k < nr; k++)
      for (uint32_t j = 0; j < ct[k]; j++, p += C) {
        uint64_t off = (uint64_t)(st[k] + j) * C;
        uint64_t nb = (off + C > len) ? len - off : C;
        memcpy(p, data + off, nb);
      }
  }
  free(src);
  free(prev);
  free(cnt);
  return ORC_OK;
}

/* TreeDeduplicator::checkpoint, src/tree_approach.cpp:1577-1689 (+ :1747) */
int orc_dedup_checkpoint(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline) {
  if (!d || !data || len == 0) return ORC_ERR_ARG;
  uint64_t n64 = (len + d->chunk_size - 1) / d->chunk_size; /* :1594-1598 */
  if (n64 > (1ull << 31)) return ORC_ERR_ARG;
  uint32_t n = (uint32_t)n64, num_nodes = 2 * n - 1;
  int baseline = make_baseline || d->current_id == 0;
  if (!baseline && n != d->n) return ORC_ERR_STATE;
  if (baseline) { /* fresh tree and table, :1601-1606; baseline_id, :1644-1650 */
    free(d->tree);
    free(d->labels);
    map_free(&d->map);
    d->tree = (uint8_t*)calloc(num_nodes, 16);
    d->labels = (uint8_t*)malloc(num_nodes);
    if (!d->tree || !d->labels || map_init(&d->map, 2ull * num_nodes) != ORC_OK) return ORC_ERR_NOMEM;
    d->n = n;
    d->num_nodes = num_nodes;
    d->baseline_id = d->current_id;
  }
  uint8_t* cur_tree = (uint8_t*)calloc(num_nodes, 16); /* fresh host view, reference_impl.hpp:27 */
  uint8_t* is_root = (uint8_t*)calloc(num_nodes, 1);
  if (!cur_tree || !is_root) { free(cur_tree); free(is_root); return ORC_ERR_NOMEM; }
  label_low_offset(d, data, len, cur_tree);
  forest_roots(d, cur_tree, is_root);
  free(d->tree);
  d->tree = cur_tree; /* reference_impl.hpp:121 */
  free(d->first);
  free(d->shift);
  uint32_t F = 0, S = 0;
  for (uint32_t v = 0; v < num_nodes; v++) { /* ascending std::set walk, :109-119 */
    if (!is_root[v]) continue;
    if (d->labels[v] == ORC_FIRST_OCUR) F++;
    else if (d->labels[v] == ORC_SHIFT_DUPL) S++;
  }
  d->first = (uint32_t*)malloc(sizeof(uint32_t) * (F ? F : 1));
  d->shift = (uint32_t*)malloc(sizeof(uint32_t) * (S ? S : 1));
  d->nfirst = d->nshift = 0;
  for (uint32_t v = 0; v < num_nodes; v++) {
    if (!is_root[v]) continue;
    if (d->labels[v] == ORC_FIRST_OCUR) d->first[d->nfirst++] = v;
    else if (d->labels[v] == ORC_SHIFT_DUPL) d->shift[d->nshift++] = v;
  }
  free(is_root);
  int rc = collect_diff(d, data, len);
  d->current_id += 1;
  return rc;
}

/* ========================================================================== */
/* Multi-threaded checkpoint (OpenMP): the CPU baseline on all host cores.     */
/* Same results as orc_dedup_checkpoint, byte for byte:                        */
/*  - leaf hashes and the fixed test run in parallel; table inserts of the    */
/*    leaves stay serial, in ascending leaf order.  A fixed leaf's insert is   */
/*    skipped after the baseline's own checkpoint: its digest is the previous */
/*    checkpoint's leaf, which that checkpoint inserted, so the serial insert */
/*    finds it and changes nothing;                                            */
/*  - both forests run one depth at a time, deepest first, nodes of a depth   */
/*    in parallel (a node reads only its children, one depth below).  The    */
/*    first-occurrence inserts are made serially afterwards in descending node */
/*    order, the serial loop's order; no forest label depends on them, and    */
/*    the shifted-duplicate forest (which reads the table) runs after them, as */
/*    in reference_impl.hpp:74-119.                                            */
/* ========================================================================== */

/* Internal nodes at heap depth k: [2^k - 1, min(2^(k+1) - 2, n - 2)]. */
static void depth_range(uint32_t n, uint32_t k, int64_t* lo, int64_t* hi) {
  *lo = (int64_t)((1ull << k) - 1);
  int64_t h = (int64_t)((1ull << (k + 1)) - 2);
  *hi = h < (int64_t)n - 2 ? h : (int64_t)n - 2;
}

static int label_low_offset_par(orc_dedup* d, const uint8_t* data, uint64_t len, uint8_t* cur_tree, int threads) {
  const uint32_t n = d->n, num_nodes = d->num_nodes, cur = d->current_id;
  const uint64_t C = d->chunk_size;
  const int first_after_reset = (cur == d->baseline_id);
  uint8_t* labels = d->labels;
  memset(labels, ORC_DONE, num_nodes);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t leaf = (int64_t)n - 1; leaf < (int64_t)num_nodes; leaf++) {
    uint64_t off = (uint64_t)(leaf - ((int64_t)n - 1)) * C;
    uint64_t nb = (leaf == (int64_t)num_nodes - 1) ? len - off : C;
    orc_hash(d->hash_kind, data + off, nb, DG(cur_tree, leaf));
    labels[leaf] = memcmp(DG(d->tree, leaf), DG(cur_tree, leaf), 16) == 0 ? ORC_FIXED_DUPL : ORC_DONE;
  }
  for (uint64_t leaf = n - 1; leaf < num_nodes; leaf++) { /* :44-71, ascending */
    if (labels[leaf] == ORC_FIXED_DUPL) {
      if (first_after_reset) map_insert(&d->map, DG(cur_tree, leaf), (uint32_t)leaf, cur, NULL);
      continue;
    }
    orc_slot* s = NULL;
    int ok = map_insert(&d->map, DG(cur_tree, leaf), (uint32_t)leaf, cur, &s);
    if (ok) {
      labels[leaf] = ORC_FIRST_OCUR;
    } else if (s->tree == cur && s->node > leaf) { /* unreachable with ascending leaves */
      labels[leaf] = ORC_FIRST_OCUR;
      labels[s->node] = ORC_SHIFT_DUPL;
      s->node = (uint32_t)leaf;
    } else {
      labels[leaf] = ORC_SHIFT_DUPL;
    }
  }
  if (n < 2) return ORC_OK;
  const uint32_t depth_int = ilog2_u64((uint64_t)n - 1); /* depth of node n - 2 */
  for (int64_t k = depth_int; k >= 0; k--) { /* first-occurrence forest, :74-82 */
    int64_t lo, hi;
    depth_range(n, (uint32_t)k, &lo, &hi);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t node = lo; node <= hi; node++) {
      uint64_t l = 2 * (uint64_t)node + 1, r = l + 1;
      if (labels[l] == ORC_FIRST_OCUR && labels[r] == ORC_FIRST_OCUR) {
        labels[node] = ORC_FIRST_OCUR;
        orc_hash(d->hash_kind, DG(cur_tree, l), 32, DG(cur_tree, node));
      }
    }
  }
  for (int64_t node = (int64_t)n - 2; node >= 0; node--)
    if (labels[node] == ORC_FIRST_OCUR) map_insert(&d->map, DG(cur_tree, node), (uint32_t)node, cur, NULL);
  return ORC_OK;
}

static void forest_roots_par(orc_dedup* d, uint8_t* cur_tree, uint8_t* is_root, int threads) {
  const uint32_t n = d->n;
  uint8_t* labels = d->labels;
  if (n >= 2) {
    const uint32_t depth_int = ilog2_u64((uint64_t)n - 1);
    for (int64_t k = depth_int; k >= 0; k--) { /* :83-104, a depth at a time */
      int64_t lo, hi;
      depth_range(n, (uint32_t)k, &lo, &hi);
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int64_t node = lo; node <= hi; node++) {
        uint64_t l = 2 * (uint64_t)node + 1, r = l + 1;
        uint8_t a = labels[l], b = labels[r];
        if (a != b) {
          labels[node] = ORC_DONE;
          if (a != ORC_FIXED_DUPL && a != ORC_DONE) is_root[l] = 1;
          if (b != ORC_FIXED_DUPL && b != ORC_DONE) is_root[r] = 1;
        } else if (a == ORC_FIXED_DUPL) {
          labels[node] = ORC_FIXED_DUPL;
        } else if (a == ORC_SHIFT_DUPL) {
          orc_hash(d->hash_kind, DG(cur_tree, l), 32, DG(cur_tree, node));
          if (map_find(&d->map, DG(cur_tree, node))) {
            labels[node] = ORC_SHIFT_DUPL;
          } else {
            labels[node] = ORC_DONE;
            is_root[l] = 1;
            is_root[r] = 1;
          }
        }
      }
    }
  }
  if (labels[0] == ORC_FIRST_OCUR) is_root[0] = 1;
  if (!d->compat && labels[0] == ORC_SHIFT_DUPL) is_root[0] = 1; /* D2 */
}

int orc_dedup_checkpoint_par(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline, int threads) {
  if (!d || !data || len == 0 || threads < 1) return ORC_ERR_ARG;
  uint64_t n64 = (len + d->chunk_size - 1) / d->chunk_size;
  if (n64 > (1ull << 31)) return ORC_ERR_ARG;
  uint32_t n = (uint32_t)n64, num_nodes = 2 * n - 1;
  int baseline = make_baseline || d->current_id == 0;
  if (!baseline && n != d->n) return ORC_ERR_STATE;
  if (baseline) {
    free(d->tree);
    free(d->labels);
    map_free(&d->map);
    d->tree = (uint8_t*)calloc(num_nodes, 16);
    d->labels = (uint8_t*)malloc(num_nodes);
    if (!d->tree || !d->labels || map_init(&d->map, 2ull * num_nodes) != ORC_OK) return ORC_ERR_NOMEM;
    d->n = n;
    d->num_nodes = num_nodes;
    d->baseline_id = d->current_id;
  }
  uint8_t* cur_tree = (uint8_t*)calloc(num_nodes, 16);
  uint8_t* is_root = (uint8_t*)calloc(num_nodes, 1);
  if (!cur_tree || !is_root) { free(cur_tree); free(is_root); return ORC_ERR_NOMEM; }
  label_low_offset_par(d, data, len, cur_tree, threads);
  forest_roots_par(d, cur_tree, is_root, threads);
  free(d->tree);
  d->tree = cur_tree;
  free(d->first);
  free(d->shift);
  uint32_t F = 0, S = 0;
  for (uint32_t v = 0; v < num_nodes; v++) {
    if (!is_root[v]) continue;
    if (d->labels[v] == ORC_FIRST_OCUR) F++;
    else if (d->labels[v] == ORC_SHIFT_DUPL) S++;
  }
  d->first = (uint32_t*)malloc(sizeof(uint32_t) * (F ? F : 1));
  d->shift = (uint32_t*)malloc(sizeof(uint32_t) * (S ? S : 1));
  d->nfirst = d->nshift = 0;
  for (uint32_t v = 0; v < num_nodes; v++) {
    if (!is_root[v]) continue;
    if (d->labels[v] == ORC_FIRST_OCUR) d->first[d->nfirst++] = v;
    else if (d->labels[v] == ORC_SHIFT_DUPL) d->shift[d->nshift++] = v;
  }
  free(is_root);
  int rc = collect_diff(d, data, len);
  d->current_id += 1;
  return rc;
}

/* ========================================================================== */
/* Basic approach: src/basic_approach.cpp:14-142 (dedup_data, collect_diff),   */
/* :468-550 (checkpoint), :153-330 (restart).  The reference writes the        */
/* changed chunks' indices in the order its atomics land; here, ascending.     */
/* The handle's `tree` holds the chunk digests (the reference's HashList).     */
/* ========================================================================== */

int orc_basic_checkpoint(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline) {
  if (!d || !data || len == 0) return ORC_ERR_ARG;
  const uint64_t C = d->chunk_size;
  uint64_t n64 = (len + C - 1) / C;
  if (n64 > (1ull << 31)) return ORC_ERR_ARG;
  const uint32_t n = (uint32_t)n64;
  if (make_baseline || !d->tree) { /* a fresh HashList, :490-493 */
    free(d->tree);
    d->tree = (uint8_t*)calloc(n, 16);
    if (!d->tree) return ORC_ERR_NOMEM;
    d->n = n;
  }
  if (n != d->n) return ORC_ERR_STATE;
  if (d->current_id == 0 || make_baseline) d->baseline_id = d->current_id; /* :510-512 */
  uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * n);
  if (!ids) return ORC_ERR_NOMEM;
  uint32_t F = 0;
  for (uint32_t c = 0; c < n; c++) { /* :26-45 */
    const uint64_t off = (uint64_t)c * C;
    const uint64_t nb = (c == n - 1) ? len - off : C;
    uint8_t h[16];
    orc_hash(d->hash_kind, data + off, nb, h);
    if (d->current_id > 0 && memcmp(DG(d->tree, c), h, 16) == 0) continue;
    memcpy(DG(d->tree, c), h, 16);
    ids[F++] = c;
  }
  const uint64_t size = 32 + (uint64_t)F * (4 + C); /* :73-75 */
  free(d->diff);
  d->diff = (uint8_t*)calloc(size, 1);
  if (!d->diff) { free(ids); return ORC_ERR_NOMEM; }
  d->diff_size = size;
  orc_header hd;
  hd.ref_id = d->baseline_id;
  hd.chkpt_id = d->current_id;
  hd.datalen = len;
  hd.chunk_size = d->chunk_size;
  hd.num_first_ocur = F;
  hd.num_prior_chkpts = 0;
  hd.num_shift_dupl = 0;
  memcpy(d->diff, &hd, 32);
  for (uint32_t i = 0; i < F; i++) put32(d->diff + 32 + 4ull * i, ids[i]);
  uint8_t* dat = d->diff + 32 + 4ull * F;
  for (uint32_t i = 0; i < F; i++) {
    const uint64_t off = (uint64_t)ids[i] * C;
    const uint64_t nb = (off + C > len) ? len - off : C;
    memcpy(dat + (uint64_t)i * C, data + off, nb);
  }
  d->nfirst = F;
  d->nshift = 0;
  free(ids);
  d->current_id += 1;
  return ORC_OK;
}

/* restart_chkpt of the basic approach: every chunk comes from the newest diff of
 * the chain ref_id..chkpt_idx that holds it. */
int orc_restart_basic(const uint8_t* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_idx,
                      uint8_t* out, uint64_t out_len) {
  if (!diffs || !sizes || chkpt_idx >= count || !diffs[chkpt_idx] || sizes[chkpt_idx] < 32) return ORC_ERR_ARG;
  orc_header hk;
  memcpy(&hk, diffs[chkpt_idx], 32);
  if (hk.chkpt_id != chkpt_idx || hk.ref_id > chkpt_idx || hk.chunk_size == 0 || out_len < hk.datalen)
    return ORC_ERR_CORRUPT;
  const uint64_t C = hk.chunk_size, len = hk.datalen;
  const uint32_t n = (uint32_t)((len + C - 1) / C);
  uint8_t* done = (uint8_t*)calloc(n, 1);
  if (!done) return ORC_ERR_NOMEM;
  for (int64_t idx = chkpt_idx; idx >= (int64_t)hk.ref_id; idx--) {
    if (!diffs[idx] || sizes[idx] < 32) { free(done); return ORC_ERR_CORRUPT; }
    orc_header h;
    memcpy(&h, diffs[idx], 32);
    if (h.chkpt_id != (uint32_t)idx || h.ref_id != hk.ref_id || h.datalen != len || h.chunk_size != C ||
        h.num_prior_chkpts || h.num_shift_dupl || 32 + (uint64_t)h.num_first_ocur * (4 + C) > sizes[idx]) {
      free(done);
      return ORC_ERR_CORRUPT;
    }
    const uint8_t* dat = diffs[idx] + 32 + 4ull * h.num_first_ocur;
    for (uint32_t i = 0; i < h.num_first_ocur; i++) {
      const uint32_t c = get32(diffs[idx] + 32 + 4ull * i);
      if (c >= n) { free(done); return ORC_ERR_CORRUPT; }
      if (done[c]) continue;
      done[c] = 1;
      const uint64_t off = (uint64_t)c * C;
      const uint64_t nb = (off + C > len) ? len - off : C;
      memcpy(out + off, dat + (uint64_t)i * C, nb);
    }
  }
  for (uint32_t c = 0; c < n; c++)
    if (!done[c]) { free(done); return ORC_ERR_CORRUPT; }
  free(done);
  return ORC_OK;
}

uint64_t orc_dedup_diff_size(const orc_dedup* d) { return d->diff_size; }
void orc_dedup_copy_diff(const orc_dedup* d, uint8_t* dst) { memcpy(dst, d->diff, d->diff_size); }
uint32_t orc_dedup_num_first_ocur(const orc_dedup* d) { return d->nfirst; }
uint32_t orc_dedup_num_shift_dupl(const orc_dedup* d) { return d->nshift; }
uint32_t orc_dedup_num_nodes(const orc_dedup* d) { return d->num_nodes; }
uint32_t orc_dedup_current_id(const orc_dedup* d) { return d->current_id; }
void orc_dedup_copy_labels(const orc_dedup* d, uint8_t* dst) { memcpy(dst, d->labels, d->num_nodes); }
void orc_dedup_copy_tree(const orc_dedup* d, uint8_t* dst) { memcpy(dst, d->tree, 16ull * d->num_nodes); }
void orc_dedup_copy_first_roots(const orc_dedup* d, uint32_t* dst) { memcpy(dst, d->first, 4ull * d->nfirst); }
void orc_dedup_copy_shift_roots(const orc_dedup* d, uint32_t* dst) { memcpy(dst, d->shift, 4ull * d->nshift); }
uint64_t orc_dedup_table_size(const orc_dedup* d) { return d->map.size; }

int orc_dedup_table_find(const orc_dedup* d, const uint8_t digest[16], uint32_t* node, uint32_t* tree) {
  orc_slot* s = map_find(&d->map, digest);
  if (!s) return 0;
  *node = s->node;
  *tree = s->tree;
  return 1;
}

/* ========================================================================== */
/* Restart: src/tree_approach.cpp:712-1128                                    */
/* ========================================================================== */

typedef struct {
  orc_header h;
  const uint32_t* first;  /* F node ids */
  const uint8_t* table;   /* P (chkpt, count) pairs */
  const uint8_t* shift;   /* S (node, prev) pairs */
  const uint8_t* data;
  uint64_t data_len;
} orc_view;


static int parse_diff(const uint8_t* buf, uint64_t size, orc_view* v) {
  if (!buf || size < 32) return ORC_ERR_CORRUPT;
  memcpy(&v->h, buf, 32);
  uint64_t meta = 4ull * v->h.num_first_ocur + 8ull * v->h.num_prior_chkpts + 8ull * v->h.num_shift_dupl;
  if (v->h.chunk_size == 0 || v->h.datalen == 0 || 32 + meta > size) return ORC_ERR_CORRUPT;
  v->first = (const uint32_t*)(buf + 32); /* :777-784 */
  v->table = buf + 32 + 4ull * v->h.num_first_ocur;
  v->shift = v->table + 8ull * v->h.num_prior_chkpts;
  v->data = buf + 32 + meta;
  v->data_len = size - 32 - meta;
  return ORC_OK;
}

/* Per-chunk data offset of every leaf inside a first region of one diff
 * (the reference's distinct_map restricted to leaves, :825-865 / :1017-1041). */
static int first_offsets(const orc_view* v, uint32_t n, uint64_t* off) {
  const uint64_t C = v->h.chunk_size;
  for (uint32_t c = 0; c < n; c++) off[c] = UINT64_MAX;
  uint64_t pos = 0;
  for (uint32_t i = 0; i < v->h.num_first_ocur; i++) {
    uint32_t node = get32((const uint8_t*)(v->first + i));
    if ((uint64_t)node >= 2ull * n - 1) return ORC_ERR_CORRUPT;
    uint32_t st[2], ct[2];
    int nr = orc_leaf_runs(node, n, st, ct);
    for (int k = 0; k < nr; k++)
      for (uint32_t j = 0; j < ct[k]; j++, pos++) off[st[k] + j] = pos * C;
  }
  if (pos * C > v->data_len) return ORC_ERR_CORRUPT;
  return ORC_OK;
}

/* Expand a node's leaf set into ascending chunk indices. */
static uint32_t leaf_list(uint32_t node, uint32_t n, uint32_t* out) {
  uint32_t st[2], ct[2], m = 0;
  int nr = orc_leaf_runs(node, n, st, ct);
  for (int k = 0; k < nr; k++)
    for (uint32_t j = 0; j < ct[k]; j++) out[m++] = st[k] + j;
  return m;
}

/* Source checkpoint of each shifted entry from the (chkpt, count) table:
 * inclusive scan over checkpoint ids, smallest id whose running sum exceeds the
 * entry index (:872-898, :1043-1067). */
static int entry_sources(const orc_view* v, uint32_t* src) {
  uint32_t top = v->h.chkpt_id;
  uint64_t* cum = (uint64_t*)calloc((size_t)top + 1, sizeof(uint64_t));
  if (!cum) return ORC_ERR_NOMEM;
  for (uint32_t i = 0; i < v->h.num_prior_chkpts; i++) {
    uint32_t c = get32(v->table + 8ull * i), k = get32(v->table + 8ull * i + 4);
    if (c > top) { free(cum); return ORC_ERR_CORRUPT; }
    cum[c] += k;
  }
  for (uint32_t c = 1; c <= top; c++) cum[c] += cum[c - 1];
  if (cum[top] != v->h.num_shift_dupl) { free(cum); return ORC_ERR_CORRUPT; }
  uint32_t c = 0;
  for (uint32_t i = 0; i < v->h.num_shift_dupl; i++) {
    while (i >= cum[c]) c++;
    src[i] = c;
  }
  free(cum);
  return ORC_OK;
}

static void copy_chunk(uint8_t* out, uint64_t datalen, uint64_t C, uint32_t chunk, const uint8_t* src) {
  uint64_t off = (uint64_t)chunk * C;
  uint64_t nb = (off + C > datalen) ? datalen - off : C;
  memcpy(out + off, src, nb);
}

int orc_restart(const uint8_t* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_idx,
                uint8_t* out, uint64_t out_len, int compat_ref) {
  (void)compat_ref;
  if (!diffs || !sizes || chkpt_idx >= count) return ORC_ERR_ARG;
  orc_view vk;
  int rc = parse_diff(diffs[chkpt_idx], sizes[chkpt_idx], &vk);
  if (rc) return rc;
  const uint64_t C = vk.h.chunk_size, datalen = vk.h.datalen;
  if (vk.h.chkpt_id != chkpt_idx || vk.h.ref_id > chkpt_idx) return ORC_ERR_CORRUPT;
  if (out_len < datalen) return ORC_ERR_ARG;
  uint64_t n64 = (datalen + C - 1) / C;
  if (n64 > (1ull << 31)) return ORC_ERR_CORRUPT;
  const uint32_t n = (uint32_t)n64;
  uint32_t* own_node = (uint32_t*)malloc(4ull * n); /* node_list, :759-760 */
  uint32_t* own_tree = (uint32_t*)malloc(4ull * n);
  uint8_t* done = (uint8_t*)calloc(n, 1);
  uint64_t* off = (uint64_t*)malloc(8ull * n);
  uint32_t* rep_leaf = (uint32_t*)malloc(4ull * n); /* repeat_map, leaf -> (leaf, src) */
  uint32_t* rep_src = (uint32_t*)malloc(4ull * n);
  uint32_t* la = (uint32_t*)malloc(4ull * n);
  uint32_t* lb = (uint32_t*)malloc(4ull * n);
  uint32_t* src = NULL;
  rc = ORC_ERR_NOMEM;
  if (!own_node || !own_tree || !done || !off || !rep_leaf || !rep_src || !la || !lb) goto out;
  for (uint32_t c = 0; c < n; c++) { own_node[c] = UINT32_MAX; own_tree[c] = UINT32_MAX; }

  /* ---- the requested diff (:803-937) ---- */
  if ((rc = first_offsets(&vk, n, off))) goto out;
  for (uint32_t c = 0; c < n; c++)
    if (off[c] != UINT64_MAX) {
      copy_chunk(out, datalen, C, c, vk.data + off[c]);
      own_node[c] = c + n - 1;
      own_tree[c] = chkpt_idx;
      done[c] = 1;
    }
  src = (uint32_t*)malloc(4ull * (vk.h.num_shift_dupl ? vk.h.num_shift_dupl : 1));
  if (!src) { rc = ORC_ERR_NOMEM; goto out; }
  if ((rc = entry_sources(&vk, src))) goto out;
  for (uint32_t i = 0; i < vk.h.num_shift_dupl; i++) {
    uint32_t node = get32(vk.shift + 8ull * i), prev = get32(vk.shift + 8ull * i + 4);
    rc = ORC_ERR_CORRUPT;
    if ((uint64_t)node >= 2ull * n - 1 || (uint64_t)prev >= 2ull * n - 1) goto out;
    uint32_t na = leaf_list(node, n, la), nb = leaf_list(prev, n, lb);
    if (na != nb) goto out;
    for (uint32_t j = 0; j < na; j++) {
      own_node[la[j]] = lb[j] + n - 1;
      own_tree[la[j]] = src[i];
      if (src[i] == chkpt_idx) {
        if (off[lb[j]] == UINT64_MAX) goto out;
        copy_chunk(out, datalen, C, la[j], vk.data + off[lb[j]]);
        done[la[j]] = 1;
      }
    }
  }
  for (uint32_t c = 0; c < n; c++) /* :932-937 */
    if (own_node[c] == UINT32_MAX) { own_node[c] = c + n - 1; own_tree[c] = chkpt_idx - 1; }
  free(src);
  src = NULL;

  /* ---- older diffs down to the baseline (:943-1121) ---- */
  for (int64_t idx = (int64_t)chkpt_idx - 1; idx >= (int64_t)vk.h.ref_id; idx--) {
    orc_view v;
    if ((rc = parse_diff(diffs[idx], sizes[idx], &v))) goto out;
    rc = ORC_ERR_CORRUPT;
    if (v.h.chkpt_id != (uint32_t)idx || v.h.datalen != datalen || v.h.chunk_size != C) goto out;
    if ((rc = first_offsets(&v, n, off))) goto out;
    src = (uint32_t*)malloc(4ull * (v.h.num_shift_dupl ? v.h.num_shift_dupl : 1));
    if (!src) { rc = ORC_ERR_NOMEM; goto out; }
    if ((rc = entry_sources(&v, src))) goto out;
    for (uint32_t c = 0; c < n; c++) rep_leaf[c] = UINT32_MAX;
    for (uint32_t i = 0; i < v.h.num_shift_dupl; i++) {
      uint32_t node = get32(v.shift + 8ull * i), prev = get32(v.shift + 8ull * i + 4);
      rc = ORC_ERR_CORRUPT;
      if ((uint64_t)node >= 2ull * n - 1 || (uint64_t)prev >= 2ull * n - 1) goto out;
      uint32_t na = leaf_list(node, n, la), nb = leaf_list(prev, n, lb);
      if (na != nb) goto out;
      for (uint32_t j = 0; j < na; j++) { rep_leaf[la[j]] = lb[j]; rep_src[la[j]] = src[i]; }
    }
    free(src);
    src = NULL;
    for (uint32_t c = 0; c < n; c++) { /* :1083-1118 */
      if (done[c] || own_tree[c] != (uint32_t)idx) continue;
      uint32_t lc = own_node[c] - (n - 1);
      if (off[lc] != UINT64_MAX) {
        copy_chunk(out, datalen, C, c, v.data + off[lc]);
        done[c] = 1;
      } else if (rep_leaf[lc] != UINT32_MAX) {
        if (rep_src[lc] == (uint32_t)idx) {
          rc = ORC_ERR_CORRUPT;
          if (off[rep_leaf[lc]] == UINT64_MAX) goto out;
          copy_chunk(out, datalen, C, c, v.data + off[rep_leaf[lc]]);
          done[c] = 1;
        } else {
          own_node[c] = rep_leaf[lc] + n - 1;
          own_tree[c] = rep_src[lc];
        }
      } else {
        own_tree[c] = (uint32_t)idx - 1;
      }
    }
  }
  rc = ORC_OK;
  for (uint32_t c = 0; c < n; c++)
    if (!done[c]) { rc = ORC_ERR_CORRUPT; break; }
out:
  free(own_node); free(own_tree); free(done); free(off); free(rep_leaf); free(rep_src);
  free(la); free(lb); free(src);
  return rc;
}
