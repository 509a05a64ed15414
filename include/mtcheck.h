/*
 * mtcheck.h -- C ABI of the MI355X tree deduplicator (libmtcheck_amd.so).
 *
 * Drop-in boundary for MTCheck's tree approach: each entry point names the
 * reference interface it replaces (paths relative to the TauferLab/MTCheck tree).
 * Plain pointers and sizes only; device pointers are HIP device memory on the
 * handle's device; `stream` is a hipStream_t (NULL = the null stream).
 * Every function returns an mtc_status and never aborts.
 */
#ifndef MTCHECK_H
#define MTCHECK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference returns void and prints on impossible states
 * (src/tree_approach.cpp:556,843); this ABI reports them instead. */
typedef enum mtc_status {
  MTC_OK = 0,
  MTC_ERR_INVALID_ARG = 1,
  MTC_ERR_UNSUPPORTED_APPROACH = 2,
  MTC_ERR_OUT_OF_MEMORY = 3,
  MTC_ERR_TABLE_FULL = 4,
  MTC_ERR_CORRUPT_DIFF = 5,
  MTC_ERR_HIP = 6,
  MTC_ERR_STATE = 7,
  MTC_ERR_BUFFER_TOO_SMALL = 8,
  MTC_PENDING = 9 /* a served request has not finished yet */
} mtc_status;

/* Approaches: the DedupMode enum, include/utils.hpp:39-49.  Every tree mode is
 * served by the tree deduplicator, as src/dedup_files.cpp:54-63 dispatches them;
 * Basic is the basic deduplicator; Full and List return
 * MTC_ERR_UNSUPPORTED_APPROACH. */
typedef enum mtc_approach {
  MTC_APPROACH_UNKNOWN = 0,
  MTC_APPROACH_FULL = 1,
  MTC_APPROACH_BASIC = 2,
  MTC_APPROACH_LIST = 3,
  MTC_APPROACH_TREE = 4,
  MTC_APPROACH_TREE_LOW_OFFSET_REF = 5,
  MTC_APPROACH_TREE_LOW_OFFSET = 6,
  MTC_APPROACH_TREE_LOW_ROOT_REF = 7,
  MTC_APPROACH_TREE_LOW_ROOT = 8
} mtc_approach;

/* Chunk hash: include/hash_functions.hpp:16-20 (chosen at run time here). */
typedef enum mtc_hash_kind { MTC_HASH_MURMUR3 = 0, MTC_HASH_MD5 = 1 } mtc_hash_kind;

/* Diff header, include/utils.hpp:29-37: 32 bytes, little-endian. */
typedef struct mtc_header {
  uint32_t ref_id;
  uint32_t chkpt_id;
  uint64_t datalen;
  uint32_t chunk_size;
  uint32_t num_first_ocur;
  uint32_t num_prior_chkpts;
  uint32_t num_shift_dupl;
} mtc_header;

typedef struct mtc_dedup mtc_dedup;

/* Per-checkpoint statistics: include/tree_approach.hpp:53-59 (num_first_ocur,
 * num_shift_dupl) and the timers of src/tree_approach.cpp:1640,1657,1672,1688. */
typedef struct mtc_stats {
  uint32_t current_id;
  uint32_t baseline_id;
  uint32_t num_first_ocur;
  uint32_t num_shift_dupl;
  uint32_t num_prior_chkpts;
  uint32_t num_chunks;
  uint64_t data_len;
  uint64_t diff_size;
  uint64_t data_chunks;     /* chunks stored in the data section */
  uint64_t table_entries;
  uint64_t table_capacity;
  float ms_setup, ms_dedup, ms_collect, ms_copy; /* timers[0..3] */
  uint64_t changed_chunks; /* leaves that are not fixed duplicates */
  float ms_leaf;           /* the leaf (hash + label + insert) kernel alone */
  /* What the last leaf kernel wrote besides its list of changed chunks (4 bytes
   * each): a label byte for every chunk (a baseline, or after a debug or failed
   * step), and the changed chunks' digests (16 bytes each; otherwise the
   * compaction hashes those chunks again). */
  uint32_t leaf_labels_written;
  uint32_t leaf_digests_stored;
} mtc_stats;

/* TreeDeduplicator(uint32_t bytes_per_chunk), src/tree_approach.cpp:5-9, plus the
 * approach selection of src/utils.cpp:15-36.  `device` is a HIP device ordinal.
 * MTC_APPROACH_BASIC gives BasicDeduplicator(bytes_per_chunk)
 * (src/basic_approach.cpp:468-549): its diff lists every chunk whose digest changed
 * since the previous checkpoint, ids ascending, then their bytes. */
int mtc_create(mtc_dedup** out, uint32_t chunk_size, int approach, int hash_kind, int device);

/* ~TreeDeduplicator, src/tree_approach.cpp (destructor). */
int mtc_destroy(mtc_dedup* h);

/* TreeDeduplicator::checkpoint(header_t&, uint8_t*, size_t, HostMirror&, bool),
 * src/tree_approach.cpp:1577-1689.  Deduplicates `len` bytes at device pointer
 * `dev_data`; the diff stays in library-owned device memory (mtc_diff_device,
 * mtc_copy_diff).  make_baseline (or the first checkpoint) resets the tree and the
 * table and sets ref_id to the current id (:1601-1606,1644-1650).  advance_id != 0
 * increments the checkpoint id afterwards, as the other three overloads do
 * (:1723,1747,1775).  An incremental checkpoint must keep the previous length.
 * Returns once the diff is complete on the device (the stream is synchronised). */
int mtc_checkpoint(mtc_dedup* h, const void* dev_data, uint64_t len, int make_baseline, int advance_id,
                   void* stream, mtc_header* header_out, uint64_t* diff_size_out);

/* The device diff of the last checkpoint (valid until the next checkpoint). */
int mtc_diff_device(mtc_dedup* h, const void** dev_diff, uint64_t* size);

/* Copy of the last diff, src/tree_approach.cpp:1678-1684: to host memory (the
 * reference's D2H copy) or to device memory (keeping a chain resident in HBM).
 * `capacity` must be at least the diff size. */
int mtc_copy_diff(mtc_dedup* h, void* host_dst, uint64_t capacity, void* stream);

/* The same copy, queued on the handle's own copy stream; returns at once
 * (src/tree_approach.cpp:1678-1684 copies synchronously; the overlap is this
 * library's).  From then on the handle alternates two device diff arenas, so the
 * next checkpoint runs while this copy is in flight; the checkpoint after it waits,
 * on the device, until the copy out of its arena has finished.  host_dst should
 * be pinned (hipHostMalloc): a pageable destination makes the copy synchronous.
 * host_dst must stay valid, and is complete, once mtc_copy_wait returns. */
int mtc_copy_diff_async(mtc_dedup* h, void* host_dst, uint64_t capacity);

/* Waits for every copy queued by mtc_copy_diff_async on this handle. */
int mtc_copy_wait(mtc_dedup* h);

/* num_first_ocur() / num_shift_dupl() (include/tree_approach.hpp:53-59) and timers. */
int mtc_get_stats(const mtc_dedup* h, mtc_stats* out);

/* TreeDeduplicator::restart(View, vector<HostMirror>&, logname, chkpt_id),
 * src/tree_approach.cpp:1786-1797 (-> restart_chkpt :712-1128), and the raw-pointer
 * overload :1809-1817.  diffs[i] is the diff of checkpoint i (entries below the
 * chain's ref_id may be NULL); diffs_on_device selects device or host pointers.
 * The restored buffer (header.datalen bytes) is written to device pointer dev_out. */
int mtc_restart(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                int diffs_on_device, void* dev_out, uint64_t out_len, int device, void* stream);

/* BasicDeduplicator::restart(View, vector<HostMirror>&, logname, chkpt_id),
 * src/basic_approach.cpp:593-604 (-> restart_chkpt :155-322), and the raw-pointer
 * overload :606-614: the same call as mtc_restart for basic diffs (header with
 * P = S = 0, the changed chunks' indices, then their bytes, C each). */
int mtc_restart_basic(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                      int diffs_on_device, void* dev_out, uint64_t out_len, int device, void* stream);

/* No counterpart in the reference: mtc_restart for callers that cannot allocate
 * device memory (code that runs where hipMalloc is not allowed, tight serving
 * loops, harnesses that forbid allocation).  It calls no hipMalloc, hipFree,
 * hipHostMalloc or other allocating device routine: its scratch is one static
 * 64 MiB device array per GPU.  Device-resident tree chains only (dev_diffs[i]
 * device pointers), of at most 256 diffs from the chain's ref_id to chkpt_id;
 * a restart whose scratch (about 8 bytes per chunk plus 16 per first-occurrence
 * root) exceeds 64 MiB -- roughly 7.5 million chunks -- returns
 * MTC_ERR_INVALID_ARG.  Calls on one GPU are serialised by the library; the
 * stream is synchronised before the call returns. */
int mtc_restart_noalloc(const void* const* dev_diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                        void* dev_out, uint64_t out_len, int device, void* stream);

/* The calling thread's last mtc_restart / mtc_restart_basic: the upload of a host
 * chain and the whole restart, in milliseconds (restart_timers[0..1],
 * src/tree_approach.cpp:755-758,1124-1127,1792-1793; 0 upload for device chains). */
int mtc_restart_times(float* upload_ms, float* total_ms);

/* calculate_digest_host, include/hash_functions.hpp:46-82: digest of a whole host
 * buffer with the selected hash (restart verification). */
int mtc_digest_host(const void* host_data, uint64_t len, int hash_kind, uint8_t out[16]);

/* Leaf digests only: hash each chunk of a device buffer into dev_digests
 * (16 bytes per chunk), the hashing half of the leaf kernel
 * (src/tree_approach.cpp:309-323). */
int mtc_hash_chunks(const void* dev_data, uint64_t len, uint32_t chunk_size, int hash_kind, void* dev_digests,
                    void* stream);

/* Reference-identical bytes where this library departs from them on purpose
 * (from the next checkpoint on; off by default).  With on != 0 a first-occurrence
 * region stores the reference's leaf range (num_leaf_descendents/leftmost_leaf,
 * include/kokkos_merkle_tree.hpp:102-128) instead of the root's true leaf set,
 * and a buffer equal to an earlier checkpoint drops shifted root 0
 * (include/reference_impl.hpp:105-106).  The two differ only when the chunk count
 * is not a power of two, or for that shifted root; there the reference's diff
 * does not restore the buffer, and neither does mtc_restart. */
int mtc_set_compat(mtc_dedup* h, int on);

/* Debug views, as the serial reference leaves them (include/reference_impl.hpp
 * :121): the full tree (16 * (2n-1) bytes, zero where no digest was computed) and
 * every node's label.  Enable before a checkpoint with mtc_set_debug(h, 1). */
int mtc_set_debug(mtc_dedup* h, int on);
int mtc_copy_tree(const mtc_dedup* h, void* host_dst, uint64_t capacity);
int mtc_copy_labels(const mtc_dedup* h, void* host_dst, uint64_t capacity);

/* Schedule parameters of a handle (no counterpart in the reference, whose Kokkos
 * policies are fixed): 256-chunk tiles streamed per leaf-kernel block (1..4096),
 * table load factor at a baseline (capacity = chunks / load_factor, in [0.05, 0.7]; default 0.125),
 * and tree levels fused per sweep stage (1..8).  The load factor and fuse depth
 * take effect at the next baseline.  Results do not depend on them. */
int mtc_set_tuning(mtc_dedup* h, uint32_t leaf_tiles_per_block, float load_factor, uint32_t fuse_levels);

/* First-stage tiles of 256 positions each wavefront of sweep A (first occurrences)
 * and sweep B (shifted duplicates and roots) walks, 1..64, or 0 for the library's
 * default (8 and 8); no counterpart in the reference.  Applies from the next
 * checkpoint.  Results do not depend on it. */
int mtc_set_sweep_tiles(mtc_dedup* h, uint32_t tiles_a, uint32_t tiles_b);

/* Batched checkpoints.  The reference checkpoints one buffer per deduplicator
 * object per call (src/tree_approach.cpp:1577-1776; one object per MPI rank,
 * src/dedup_files_mpi.cpp:72-81).  A batch runs one checkpoint of each of `count`
 * handles (each with its own state, chunk size, hash and history) as one sequence
 * of launches on the batch's stream, after the work already queued on `stream`.
 * With use_graph, steps whose launch shape repeats are replayed from one captured
 * hipGraph.  Each handle's result is identical to mtc_checkpoint on it alone. */
typedef struct mtc_batch mtc_batch;
int mtc_batch_create(mtc_batch** out, int device, int use_graph);
int mtc_batch_destroy(mtc_batch* b);
int mtc_checkpoint_batch(mtc_batch* b, mtc_dedup* const* handles, const void* const* dev_data, const uint64_t* lens,
                         const int* make_baseline, uint32_t count, int advance_id, void* stream,
                         mtc_header* headers_out, uint64_t* diff_sizes_out);
const char* mtc_batch_last_error(const mtc_batch* b);

/* Served checkpoints (no counterpart in the reference, which checkpoints inside
 * the application's own loop).  A server keeps each client's tree and table
 * resident on `device`, admits a client's first checkpoint only while the
 * estimated resident state of all admitted clients stays within memory_budget
 * bytes, and runs `lanes` worker threads, each with its own stream: a free lane
 * takes every queued request it may run (at most one per client, in arrival
 * order, at most max_batch) as one batched step, so requests arriving while others
 * run join the next step and the steps of different lanes overlap on the device.
 * A client's checkpoints run in submission order (the result equals its serial
 * sequence); the diff of a request stays in the client's handle
 * (mtc_server_client) until the client's next checkpoint runs.  mtc_server_submit
 * records an event on `stream`, so the buffer may still be in production there.
 * After a result is posted, the lane may grow that client's digest table before
 * taking its next checkpoint; reading the client's diff and statistics meanwhile
 * is safe, and the grown bytes count against the memory budget. */
typedef struct mtc_server mtc_server;
int mtc_server_create(mtc_server** out, int device, uint64_t memory_budget, uint32_t lanes, uint32_t max_batch);
int mtc_server_destroy(mtc_server* s);
int mtc_server_add_client(mtc_server* s, uint32_t chunk_size, int hash_kind, uint64_t max_len, uint32_t* client);
int mtc_server_client(mtc_server* s, uint32_t client, mtc_dedup** handle);
int mtc_server_remove_client(mtc_server* s, uint32_t client);
int mtc_server_submit(mtc_server* s, uint32_t client, const void* dev_data, uint64_t len, int make_baseline,
                      void* stream, uint64_t* ticket);
/* MTC_PENDING while unfinished (wait == 0); otherwise the request's status, and the
 * record is released.  Times are CLOCK_MONOTONIC seconds. */
int mtc_server_result(mtc_server* s, uint64_t ticket, int wait, mtc_header* header, uint64_t* diff_size,
                      double* t_submit, double* t_start, double* t_done);
int mtc_server_stats(mtc_server* s, uint64_t* steps, uint64_t* served, uint64_t* bytes, uint64_t* resident_bytes,
                     uint32_t* queued);
const char* mtc_server_last_error(mtc_server* s);

/* Message for the last error on a handle (or the thread's last handle-less call). */
const char* mtc_last_error(const mtc_dedup* h);
const char* mtc_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
