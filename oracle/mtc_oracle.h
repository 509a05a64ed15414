/*
 * mtc_oracle.h -- CPU restatement of MTCheck's tree (lowest-offset) deduplication
 * path.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library, the
 * CPU baseline timed by bench.py, and nothing else.  No product code links it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the TauferLab/MTCheck source tree.
 */
#ifndef MTC_ORACLE_H
#define MTC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node labels, include/utils.hpp:51-57 */
enum { ORC_FIRST_OCUR = 0, ORC_FIXED_DUPL = 1, ORC_SHIFT_DUPL = 2, ORC_FIRST_DUPL = 3, ORC_DONE = 4 };

/* Hash selector.  include/hash_functions.hpp:16-20 picks one at compile time. */
enum { ORC_HASH_MURMUR3 = 0, ORC_HASH_MD5 = 1 };

/* Status codes */
enum { ORC_OK = 0, ORC_ERR_ARG = 1, ORC_ERR_CORRUPT = 2, ORC_ERR_STATE = 3, ORC_ERR_NOMEM = 4 };

/* header_t, include/utils.hpp:29-37 (32 bytes, little-endian, no padding) */
typedef struct orc_header {
  uint32_t ref_id;
  uint32_t chkpt_id;
  uint64_t datalen;
  uint32_t chunk_size;
  uint32_t num_first_ocur;
  uint32_t num_prior_chkpts;
  uint32_t num_shift_dupl;
} orc_header;

/* ---- hashes ---------------------------------------------------------------- */
/* MurmurHash3_x64_128 exactly as include/kokkos_murmur3.hpp:214-292 computes it
 * (rotl64 returns uint32_t, :31), seed 0 as in hash() :432-437. */
void orc_murmur3_mtcheck(const void* data, uint64_t len, uint8_t out[16]);
/* Same body with a full 64-bit rotate: canonical smhasher MurmurHash3_x64_128.
 * Exists only so tests can pin the shared body against published vectors. */
void orc_murmur3_canonical(const void* data, uint64_t len, uint32_t seed, uint8_t out[16]);
/* RFC 1321 MD5, include/kokkos_md5.hpp:49-323 */
void orc_md5(const void* data, uint64_t len, uint8_t out[16]);
void orc_hash(int kind, const void* data, uint64_t len, uint8_t out[16]);

/* ---- heap-tree leaf sets ------------------------------------------------------ */
/* True leaf set of `node` in a heap tree over n chunks, as up to two runs of
 * consecutive chunk indices in ascending order.  Returns the number of runs. */
int orc_leaf_runs(uint32_t node, uint32_t n, uint32_t start[2], uint32_t count[2]);
/* include/kokkos_merkle_tree.hpp:102-114 and :116-128, restated (correct only
 * when all leaves of `node` share one depth). */
uint32_t orc_ref_num_leaf_descendents(uint32_t node, uint32_t num_nodes);
uint32_t orc_ref_leftmost_leaf(uint32_t node, uint32_t num_nodes);

/* ---- deduplicator ------------------------------------------------------------- */
typedef struct orc_dedup orc_dedup;

/* compat_ref != 0 reproduces two reference behaviours the library deliberately
 * departs from: region leaf ranges from the helpers above, and a shifted-duplicate
 * root 0 that is never emitted. */
orc_dedup* orc_dedup_create(uint32_t chunk_size, int hash_kind, int compat_ref);
void orc_dedup_destroy(orc_dedup* d);
/* One checkpoint; advances current_id like src/tree_approach.cpp:1740-1748. */
int orc_dedup_checkpoint(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline);
/* The basic approach (src/basic_approach.cpp): changed chunks in ascending order. */
int orc_basic_checkpoint(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline);
int orc_restart_basic(const uint8_t* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_idx,
                      uint8_t* out, uint64_t out_len);

/* The same checkpoint with OpenMP on `threads` threads (identical results). */
int orc_dedup_checkpoint_par(orc_dedup* d, const uint8_t* data, uint64_t len, int make_baseline, int threads);
uint64_t orc_dedup_diff_size(const orc_dedup* d);
void orc_dedup_copy_diff(const orc_dedup* d, uint8_t* dst);
uint32_t orc_dedup_num_first_ocur(const orc_dedup* d);
uint32_t orc_dedup_num_shift_dupl(const orc_dedup* d);
uint32_t orc_dedup_num_nodes(const orc_dedup* d);
uint32_t orc_dedup_current_id(const orc_dedup* d);
/* Labels (num_nodes bytes) and the tree (16*num_nodes bytes) of the last checkpoint. */
void orc_dedup_copy_labels(const orc_dedup* d, uint8_t* dst);
void orc_dedup_copy_tree(const orc_dedup* d, uint8_t* dst);
void orc_dedup_copy_first_roots(const orc_dedup* d, uint32_t* dst);
void orc_dedup_copy_shift_roots(const orc_dedup* d, uint32_t* dst);
/* Table lookup: returns 1 and the (node, tree) pair when the digest is present. */
int orc_dedup_table_find(const orc_dedup* d, const uint8_t digest[16], uint32_t* node, uint32_t* tree);
uint64_t orc_dedup_table_size(const orc_dedup* d);

/* ---- restart ------------------------------------------------------------------ */
/* Rebuild checkpoint `chkpt_idx` from diffs[0..count) (indexed by checkpoint id)
 * into out (>= datalen bytes).  src/tree_approach.cpp:712-1128. */
int orc_restart(const uint8_t* const* diffs, const uint64_t* sizes, uint32_t count,
                uint32_t chkpt_idx, uint8_t* out, uint64_t out_len, int compat_ref);

#ifdef __cplusplus
}
#endif
#endif
